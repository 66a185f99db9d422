"""INTEGRATION.md's sample program, built verbatim from the document by
srp_amd/build.py:build_integration_sample (tests/integration/build/): the
GPU program (a HIP shader module and a C host linked to libsrp_amd.so) must
write the same colour, depth and stencil buffers as the same host program
linked to the CPU oracle, which in turn equals the reference renderer
compiled in place (oracle/_ref) on this program."""
import os
import shutil
import subprocess
import sys

import numpy as np
import pytest

from scene_runner import REPO, RES_DIR

sys.path.insert(0, REPO)
from srp_amd import build as srp_build  # noqa: E402

BUILD = srp_build.INTEGRATION_BUILD
REF_ROOT = os.environ.get("SRP_REFERENCE", "/root/reference")   # the reference's sources (oracle/Makefile's REF)
TEX = os.path.join(RES_DIR, "stoneWall.png")
N = 320 * 240


def _run(exe, prefix):
    r = subprocess.run([exe, TEX, prefix], stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    b = np.fromfile(prefix + ".bin", np.uint8)
    assert b.size == 9 * N
    return b[:4 * N].view(np.uint32), b[4 * N:8 * N].view(np.uint32), b[8 * N:]


def test_built_sample_is_the_document_verbatim():
    """The built sources equal the blocks INTEGRATION.md shows (rebuild
    after editing the document)."""
    for name, body in srp_build.integration_sample_sources().items():
        with open(os.path.join(BUILD, "src", name)) as f:
            assert f.read() == body, name
    assert os.path.exists(os.path.join(BUILD, "sample_gpu")) and os.path.exists(os.path.join(BUILD, "sample_oracle"))


def test_oracle_build_equals_compiled_reference(tmp_path):
    ref_obj = os.path.join(REPO, "oracle", "_ref", "obj")
    cc = shutil.which("gcc")
    if not (cc and os.path.isdir(ref_obj) and os.path.isdir(os.path.join(REF_ROOT, "include"))):
        pytest.skip("compiled reference (oracle/_ref) not built")
    # the reference has no PNG writer of its own in the library (its tests
    # link stb_image_write); the sample's call is served by the library's
    shim = tmp_path / "save.c"
    shim.write_text('#include <srp/srp.h>\n#include "png_encode.h"\n'
                    "int srpFramebufferSaveImage(const SRPFramebuffer* fb, const char* path)\n"
                    "{ return srpPngSaveColorBuffer(path, (int) fb->width, (int) fb->height, fb->color); }\n")
    exe = str(tmp_path / "sample_ref")
    host = os.path.join(REPO, "srp_amd", "csrc", "host")
    objs = [os.path.join(ref_obj, f) for f in sorted(os.listdir(ref_obj)) if f.endswith(".o")]
    # the reference's own headers first; srp_hip.h (the sample's PNG call)
    # only exists in the library's include directory
    subprocess.run([cc, "-std=c11", "-O2", "-ffp-contract=off", "-I", os.path.join(REF_ROOT, "include"),
                    "-I", os.path.join(REPO, "include"),
                    "-I", os.path.join(BUILD, "src"), "-I", host, os.path.join(BUILD, "src", "main.c"),
                    os.path.join(REPO, "tests", "integration", "host_shaders.c"), str(shim),
                    os.path.join(host, "png_encode.c")] + objs + ["-lz", "-lm", "-o", exe], check=True)
    ref = _run(exe, str(tmp_path / "ref"))
    got = _run(os.path.join(BUILD, "sample_oracle"), str(tmp_path / "oracle"))
    for a, b in zip(ref, got):
        assert np.array_equal(a, b)
    assert int(np.count_nonzero(ref[2])) > 5000 and len(np.unique(ref[0])) > 1000


@pytest.mark.gpu
def test_gpu_sample_equals_oracle(tmp_path):
    print("[integration] sample_gpu vs sample_oracle", flush=True)
    gpu = _run(os.path.join(BUILD, "sample_gpu"), str(tmp_path / "gpu"))
    cpu = _run(os.path.join(BUILD, "sample_oracle"), str(tmp_path / "oracle"))
    for a, b in zip(gpu, cpu):
        assert np.array_equal(a, b)
    with open(tmp_path / "gpu.png", "rb") as f1, open(tmp_path / "oracle.png", "rb") as f2:
        assert f1.read() == f2.read()
