"""The C-ABI library loads without a GPU and exports every function the
public headers declare (include/srp/*.h), with the reference's names."""
import ctypes
import os
import re
import subprocess

import pytest

from scene_runner import REPO

LIB = os.path.join(REPO, "srp_amd", "libsrp_amd.so")
INCLUDE = os.path.join(REPO, "include", "srp")


def declared_functions():
    names = set()
    for fn in os.listdir(INCLUDE):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(INCLUDE, fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b((?:srp|vec[234]|mat4)\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_headers_declare_reference_api():
    names = declared_functions()
    # 19 context + 8 buffer + 3 framebuffer + 5 texture + 24 vec + 10 mat functions
    reference_api = [n for n in names if not n.startswith("srpHip") and "Neural" not in n and "Batch" not in n and n not in (
        "srpFramebufferSetDeferred", "srpFramebufferDownload", "srpFramebufferUpload", "srpFramebufferFillDepth",
        "srpFramebufferFillStencil", "srpFramebufferSaveImage", "srpNewTextureFromMemory", "srpTextureGetFilteredColorHost", "srpMakeVec2",
        "srpMakeVec3", "srpMakeVec4", "srpMakeVec4From3", "srpMat4Diag_")]
    assert len(reference_api) == 69, reference_api


def test_library_exports_every_declared_function():
    lib = ctypes.CDLL(LIB, mode=ctypes.RTLD_LOCAL)
    missing = [n for n in declared_functions()
               if n not in ("srpMakeVec2", "srpMakeVec3", "srpMakeVec4", "srpMakeVec4From3", "srpMat4Diag_")
               and not hasattr(lib, n)]
    assert not missing, missing


def test_library_is_gfx950_code_object():
    out = subprocess.run(["strings", LIB], stdout=subprocess.PIPE, text=True).stdout
    assert "gfx950" in out
