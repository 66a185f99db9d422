"""ctypes driver of the neural shader's C ABI (include/srp/srp_neural.h) and
of the terrain workload's frame driver in the GPU scene module
(tests/scenes/bench.c)."""
import ctypes
import os

import numpy as np

from scene_runner import GPU_SO, REPO

LIB_SO = os.path.join(REPO, "srp_amd", "libsrp_amd.so")
N_PARAMS = 37187

_c = ctypes
_vp = _c.c_void_p
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


class Neural:
    """The library's neural API plus the scene module's terrain driver."""

    def __init__(self):
        self.lib = _c.CDLL(LIB_SO, mode=_c.RTLD_LOCAL)
        self.scn = _c.CDLL(GPU_SO, mode=_c.RTLD_LOCAL)
        L, S = self.lib, self.scn
        self.new_ds = _sig(L, "srpNewNeuralDataset", _vp, [_c.c_size_t])
        self.free_ds = _sig(L, "srpFreeNeuralDataset", None, [_vp])
        self.ds_size = _sig(L, "srpNeuralDatasetSize", _c.c_size_t, [_vp])
        self._ds_read = _sig(L, "srpNeuralDatasetRead", None, [_vp, _c.c_size_t, _c.c_size_t, _f32p, _f32p])
        self._ds_write = _sig(L, "srpNeuralDatasetWrite", None, [_vp, _f32p, _f32p, _c.c_size_t])
        self.new_net = _sig(L, "srpNewNeuralNet", _vp, [_c.c_uint64])
        self.free_net = _sig(L, "srpFreeNeuralNet", None, [_vp])
        self._get = _sig(L, "srpNeuralNetGetParams", None, [_vp, _f32p])
        self._set = _sig(L, "srpNeuralNetSetParams", None, [_vp, _f32p])
        self._train = _sig(L, "srpNeuralNetTrain", None,
                           [_vp, _vp, _c.c_int, _c.c_int, _c.c_float, _c.c_float, _c.c_uint64, _vp])
        self._grad = _sig(L, "srpNeuralNetGradients", None, [_vp, _vp, _c.c_size_t, _f32p, _vp])
        self._infer = _sig(L, "srpNeuralNetInfer", None, [_vp, _vp, _c.c_size_t, _f32p, _c.c_int])
        self.net_device = _sig(L, "srpNeuralNetDevice", _vp, [_vp])
        self.time_inference = _sig(L, "srpNeuralNetTimeInference", _c.c_double, [_vp, _vp, _c.c_size_t, _c.c_int])
        self._dense = _sig(S, "srpBenchTerrainDenseMs", _c.c_double, [_f32p, _c.c_size_t, _f32p, _c.c_int])
        self._bounds = _sig(S, "srpBenchTerrainBounds", None, [_f32p])
        self._set_enc = _sig(L, "srpNeuralNetSetEncoding", None, [_vp, _f32p])
        self._get_enc = _sig(L, "srpNeuralNetGetEncoding", None, [_vp, _f32p])
        self._compare = _sig(L, "srpNeuralCompare", None, [_vp, _vp, _vp, _vp, _vp, _vp])
        self.sync = _sig(L, "srpHipSynchronize", None, [])
        self.set_deferred = _sig(L, "srpHipSetDeferredShading", None, [_c.c_int])
        self.get_deferred = _sig(L, "srpHipGetDeferredShading", _c.c_int, [])
        self.profile_enable = _sig(L, "srpHipProfileEnable", None, [_c.c_bool])
        self.profile_slots = _sig(L, "srpHipProfileSlots", None, [_c.c_uint])
        self._profile_read = _sig(L, "srpHipProfileRead", None, [_vp, _vp])
        self.reset_stats = _sig(L, "srpHipResetStats", None, [])
        self._stats = _sig(L, "srpHipGetStats", None, [_vp])
        self.terrain_setup = _sig(S, "srpBenchTerrainSetup", _c.c_int, [_c.c_int] * 6)
        self.terrain_pose = _sig(S, "srpBenchTerrainPose", None, [_c.c_int])
        self.terrain_shader = _sig(S, "srpBenchTerrainShader", None, [_c.c_int, _vp, _vp])
        self.terrain_collect = _sig(S, "srpBenchTerrainCollect", _c.c_size_t,
                                    [_vp, _c.c_int, _c.c_int, _c.c_float, _c.c_uint64])
        self.frame = _sig(S, "srpBenchFrame", None, [])
        self.fb = _sig(S, "srpBenchFramebuffer", _vp, [])
        self.use_fb = _sig(S, "srpBenchUseFramebuffer", None, [_c.c_int])
        self._read = _sig(S, "srpBenchRead", None, [_vp, _vp, _vp])
        self.finish = _sig(S, "srpBenchFinish", None, [])
        self.neural_fs = _c.cast(S.sc_fsTerrainNeural, _vp).value

    def terrain_dense_ms(self, features, count, reps):
        """Device ms per dense pass of the procedural material over features."""
        b = np.zeros(6, np.float32)
        self._bounds(b)
        return self._dense(np.ascontiguousarray(features[:count], np.float32), count, b, reps)

    # -- data
    def ds_read(self, ds, first, count):
        F = np.empty((count, 8), np.float32)
        T = np.empty((count, 3), np.float32)
        self._ds_read(ds, first, count, F, T)
        return F, T

    def ds_write(self, ds, F, T):
        F = np.ascontiguousarray(F, np.float32)
        T = np.ascontiguousarray(T, np.float32)
        self._ds_write(ds, F, T, F.shape[0])

    # -- network
    def get_params(self, net):
        p = np.empty(N_PARAMS, np.float32)
        self._get(net, p)
        return p

    def set_params(self, net, p):
        self._set(net, np.ascontiguousarray(p, np.float32))

    def set_encoding(self, net, steps):
        self._set_enc(net, np.ascontiguousarray(steps, np.float32))

    def get_encoding(self, net):
        st = np.empty(8, np.float32)
        self._get_enc(net, st)
        return st

    def train(self, net, ds, steps, batch, lr0, lr1, seed):
        loss = np.empty(steps, np.float32)
        self._train(net, ds, steps, batch, lr0, lr1, seed, loss.ctypes.data)
        return loss

    def gradients(self, net, ds, count):
        g = np.empty(N_PARAMS, np.float32)
        loss = np.zeros(1, np.float32)
        self._grad(net, ds, count, g, loss.ctypes.data)
        return g, float(loss[0])

    def infer(self, net, ds, count, lanes=False):
        y = np.empty((count, 3), np.float32)
        self._infer(net, ds, count, y, 1 if lanes else 0)
        return y

    def compare(self, fa, fb, mask):
        """PSNR, largest difference and pixel count over mask's covered pixels."""
        ps = _c.c_double(0)
        mx = _c.c_double(0)
        n = _c.c_uint64(0)
        self._compare(fa, fb, mask, _c.addressof(ps), _c.addressof(mx), _c.addressof(n))
        return ps.value, mx.value, n.value

    # -- frames
    def read_frame(self, w, h):
        c = np.empty(w * h, np.uint32)
        d = np.empty(w * h, np.float32)
        s = np.empty(w * h, np.uint8)
        self._read(c.ctypes.data, d.ctypes.data, s.ctypes.data)
        return c.reshape(h, w), d.reshape(h, w), s.reshape(h, w)

    def profile_read(self):
        ms = (_c.c_double * 4)()
        n = (_c.c_uint64 * 4)()
        self._profile_read(ms, n)
        return list(ms), list(n)

    def stats(self):
        a = (_c.c_uint64 * 4)()
        self._stats(a)
        return list(a)


def rgb_of(color):
    """uint32 R<<24|G<<16|B<<8|A -> [..., 3] uint8."""
    return np.stack([(color >> 24) & 255, (color >> 16) & 255, (color >> 8) & 255], -1).astype(np.uint8)
