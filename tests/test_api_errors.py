"""Error behaviour of the learned-shader, batch and framebuffer-extension
C ABI (include/srp/srp_neural.h, include/srp/srp_hip.h), on the CPU: invalid arguments are rejected before any device work and
reported through the context's message callback (the reference's error
channel, utils/message_callback.c:18-35), naming the API function."""
import ctypes
import os

import numpy as np
import pytest

from scene_runner import REPO

LIB_SO = os.path.join(REPO, "srp_amd", "libsrp_amd.so")

CALLBACK = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_char_p, ctypes.c_char_p, ctypes.c_void_p)


class _Callback(ctypes.Structure):
    _fields_ = [("func", CALLBACK), ("userParameter", ctypes.c_void_p)]


@pytest.fixture(scope="module")
def api():
    if not os.path.exists(LIB_SO):
        pytest.skip("libsrp_amd.so is not built")
    lib = ctypes.CDLL(LIB_SO, mode=ctypes.RTLD_LOCAL)
    got = []

    def on_message(kind, severity, fn, text, user):
        got.append((fn.decode(), text.decode()))

    cb = CALLBACK(on_message)
    lib.srpSetMessageCallback.argtypes = [_Callback]
    # pointer arguments are passed whole whichever test runs first
    lib.srpNeuralNetSetEncoding.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    lib.srpSetMessageCallback(_Callback(cb, None))
    yield lib, got
    lib.srpSetMessageCallback(_Callback(CALLBACK(), None))


def test_dataset_capacity_must_be_positive(api):
    lib, got = api
    got.clear()
    lib.srpNewNeuralDataset.restype = ctypes.c_void_p
    lib.srpNewNeuralDataset.argtypes = [ctypes.c_size_t]
    assert lib.srpNewNeuralDataset(0) is None
    assert got == [("srpNewNeuralDataset", "Capacity must be positive")]


@pytest.mark.parametrize("bad", [0.0, -1.0, 1024.0, float("nan"), float("inf")])
def test_encoding_step_out_of_range(api, bad):
    lib, got = api
    got.clear()
    lib.srpNeuralNetSetEncoding.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    steps = np.full(8, 0.5, np.float32)
    steps[5] = bad
    lib.srpNeuralNetSetEncoding(None, steps.ctypes.data)   # rejected before the net is touched
    assert len(got) == 1 and got[0][0] == "srpNeuralNetSetEncoding"
    assert "feature 5 is not in (0, 1024)" in got[0][1]


_HALF_STEPS = np.full(8, 0.5, np.float32)   # alive for the whole module


@pytest.mark.parametrize("call, fn, text", [
    (lambda L: L.srpNeuralDatasetStage(None, None, 0), "srpNeuralDatasetStage", "Invalid dataset, framebuffer or slot (0)"),
    (lambda L: L.srpNeuralDatasetAppend(None, None, ctypes.c_float(1.0), ctypes.c_uint64(1)),
     "srpNeuralDatasetAppend", "Nothing staged for a framebuffer of this size"),
    (lambda L: L.srpNeuralDatasetRead(None, ctypes.c_size_t(0), ctypes.c_size_t(1), None, None),
     "srpNeuralDatasetRead", "Range out of bounds"),
    (lambda L: L.srpNeuralDatasetWrite(None, None, None, ctypes.c_size_t(1)),
     "srpNeuralDatasetWrite", "1 samples exceed the capacity"),
    (lambda L: L.srpNeuralNetTrain(None, None, 10, 8192, ctypes.c_float(1e-3), ctypes.c_float(1e-4),
                                   ctypes.c_uint64(1), None), "srpNeuralNetTrain", "Nothing to train on"),
    (lambda L: L.srpNeuralNetGradients(None, None, ctypes.c_size_t(256), None, None),
     "srpNeuralNetGradients", "Count exceeds the dataset"),
    (lambda L: L.srpNeuralNetInfer(None, None, ctypes.c_size_t(1), None, 0),
     "srpNeuralNetInfer", "Count exceeds the dataset"),
    (lambda L: L.srpNeuralNetGetParams(None, None), "srpNeuralNetGetParams", "Invalid network or parameter array"),
    (lambda L: L.srpNeuralNetSetParams(None, None), "srpNeuralNetSetParams", "Invalid network or parameter array"),
    (lambda L: L.srpNeuralNetGetEncoding(None, None), "srpNeuralNetGetEncoding", "Invalid network or step array"),
    (lambda L: L.srpNeuralNetSetEncoding(None, None), "srpNeuralNetSetEncoding", "Invalid step array"),
    (lambda L: L.srpNeuralNetSetEncoding(None, _HALF_STEPS.ctypes.data), "srpNeuralNetSetEncoding", "Invalid network"),
])
def test_missing_objects_are_reported(api, call, fn, text):
    """Every dataset and training call given no object reports it and
    returns before any device work (nothing here has a GPU)."""
    lib, got = api
    got.clear()
    call(lib)
    assert got == [(fn, text)]


@pytest.mark.parametrize("run", [-1, 65])
def test_tile_order_out_of_range(api, run):
    lib, got = api
    got.clear()
    lib.srpHipSetTileOrder(run)
    assert got == [("srpHipSetTileOrder", f"Tile run {run} is not in [0, 64]")]


@pytest.mark.parametrize("fn, args", [
    ("srpFramebufferSetDeferred", (None, True)),
    ("srpFramebufferDownload", (None,)),
    ("srpFramebufferUpload", (None,)),
    ("srpFramebufferFillDepth", (None, ctypes.c_float(1.0))),
    ("srpFramebufferFillStencil", (None, ctypes.c_uint8(0))),
])
def test_framebuffer_extensions_need_a_framebuffer(api, fn, args):
    lib, got = api
    got.clear()
    getattr(lib, fn)(*args)
    assert got == [(fn, "Invalid framebuffer")]


def test_batch_needs_draws(api):
    lib, got = api
    got.clear()
    lib.srpNewBatch.restype = ctypes.c_void_p
    lib.srpNewBatch.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    assert lib.srpNewBatch(None, 0) is None
    assert got == [("srpNewBatch", "No draws")]


def test_compare_needs_framebuffers_of_one_size(api):
    lib, got = api
    got.clear()
    lib.srpNeuralCompare.argtypes = [ctypes.c_void_p] * 6
    ps, mx, n = ctypes.c_double(-1), ctypes.c_double(-1), ctypes.c_uint64(7)
    lib.srpNeuralCompare(None, None, None, ctypes.addressof(ps), ctypes.addressof(mx), ctypes.addressof(n))
    assert got == [("srpNeuralCompare", "Framebuffers must exist and have one size")]
    assert ps.value == -1 and mx.value == -1 and n.value == 7   # outputs untouched
