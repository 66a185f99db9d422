"""Build recipe for libsrp_amd and HIP shader modules (gfx950).

Every translation unit, host or device, is compiled with FP contraction off
and without fast-math: bit-exact parity with the reference renderer depends
on it.  HIP code also keeps correctly rounded float division/sqrt and
denormals, and disables packed FP32 instructions.

    python -m srp_amd.build            # library + test scene module + oracle
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "srp_amd")
INCLUDE = os.path.join(REPO, "include")
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build", "obj")
LIB = os.path.join(PKG, "libsrp_amd.so")
SCENE_DIR = os.path.join(REPO, "tests", "scenes")
SCENE_BUILD = os.path.join(SCENE_DIR, "build")
SCENE_LIB = os.path.join(SCENE_BUILD, "libscenes_amd.so")

ARCH = os.environ.get("SRP_OFFLOAD_ARCH", "gfx950")
HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CC = os.environ.get("CC", "gcc")

NUMERIC_FLAGS = ["-ffp-contract=off", "-fno-fast-math"]
HIP_FLAGS = [
    "-x", "hip", f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-pass-failed",
    # host code: automatic variables start zeroed (launch-argument structs
    # handed to kernels can never carry stack garbage)
    "-Xarch_host", "-ftrivial-auto-var-init=zero",
] + NUMERIC_FLAGS
C_FLAGS = ["-O2", "-fPIC", "-std=c11", "-Wall", "-Wno-unused-function"] + NUMERIC_FLAGS


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}")
    return r.stdout


def _hip_obj(src, obj, includes):
    inc = sum((["-I", d] for d in includes), [])
    _run([HIPCC] + HIP_FLAGS + inc + ["-c", src, "-o", obj])
    return obj


def _c_obj(src, obj, includes, defines=()):
    inc = sum((["-I", d] for d in includes), [])
    _run([CC] + C_FLAGS + inc + [f"-D{d}" for d in defines] + ["-c", src, "-o", obj])
    return obj


def build_library(jobs=8):
    """libsrp_amd.so: host API + shader-independent kernels."""
    os.makedirs(OBJ, exist_ok=True)
    inc = [INCLUDE, os.path.join(CSRC, "lib"), os.path.join(CSRC, "host")]
    hip_srcs = ["lib/runtime.cpp", "lib/assemble.hip", "lib/util.hip", "neural/mlp.hip"]
    c_srcs = ["host/math.c", "host/png_decode.c", "host/image_decode.c", "host/png_encode.c", "host/obj_load.c"]
    tasks = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for s in hip_srcs:
            o = os.path.join(OBJ, os.path.basename(s) + ".o")
            tasks.append(ex.submit(_hip_obj, os.path.join(CSRC, s), o, inc))
        for s in c_srcs:
            o = os.path.join(OBJ, os.path.basename(s) + ".o")
            tasks.append(ex.submit(_c_obj, os.path.join(CSRC, s), o, inc))
        objs = [t.result() for t in tasks]
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz"])
    os.replace(tmp, LIB)
    return LIB


def build_shader_module(hip_sources, c_sources, out, includes=(), jobs=8, defines=()):
    """A shader module: HIP files with SRP_HIP_*_SHADER registrations plus
    optional host C files, linked against libsrp_amd.so."""
    objdir = os.path.join(os.path.dirname(out), "obj")
    os.makedirs(objdir, exist_ok=True)
    inc = [INCLUDE] + list(includes)
    tasks = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for s in hip_sources:
            o = os.path.join(objdir, os.path.basename(s) + ".o")
            tasks.append(ex.submit(_hip_obj, s, o, inc))
        for s in c_sources:
            o = os.path.join(objdir, os.path.basename(s) + ".o")
            tasks.append(ex.submit(_c_obj, s, o, inc, defines))
        objs = [t.result() for t in tasks]
    rpath = os.path.relpath(PKG, os.path.dirname(out))
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-L", PKG, "-lsrp_amd", f"-Wl,-rpath,$ORIGIN/{rpath}", "-lm"])
    os.replace(tmp, out)
    return out


def build_scenes(jobs=8):
    """The GPU build of the test scene programs (tests/scenes)."""
    host = os.path.join(CSRC, "host")
    return build_shader_module(
        [os.path.join(SCENE_DIR, f) for f in ("device_shaders.hip", "device_shaders_1.hip", "device_shaders_2.hip",
                                              "device_shaders_3.hip", "device_shaders_4.hip", "copy_probe.hip")],
        [os.path.join(SCENE_DIR, f) for f in ("scenes.c", "fuzz.c", "bench.c", "messages.c", "terrain.c", "batch.c")] + [os.path.join(host, "obj_load.c")],
        SCENE_LIB, includes=[SCENE_DIR, host], jobs=jobs, defines=["SRP_GPU_BUILD"])


INTEGRATION_DIR = os.path.join(REPO, "tests", "integration")
INTEGRATION_BUILD = os.path.join(INTEGRATION_DIR, "build")


def integration_sample_sources(doc=os.path.join(REPO, "INTEGRATION.md")):
    """The sample program's files as INTEGRATION.md shows them: every fenced
    block preceded by a `<!-- sample: NAME -->` line, indentation removed."""
    import re
    import textwrap
    text = open(doc).read()
    files = {}
    for m in re.finditer(r"<!-- sample: (\S+) -->\s*\n(\s*)```\w*\n(.*?)\n\s*```", text, flags=re.S):
        files[m.group(1)] = textwrap.dedent(m.group(3)) + "\n"
    return files


def build_integration_sample():
    """INTEGRATION.md's sample, verbatim: the GPU program (shaders.hip as a
    shader module, main.c linked to libsrp_amd.so) and the same main.c with
    tests/integration/host_shaders.c against the CPU oracle."""
    src = os.path.join(INTEGRATION_BUILD, "src")
    os.makedirs(src, exist_ok=True)
    for name, body in integration_sample_sources().items():
        with open(os.path.join(src, name), "w") as f:
            f.write(body)
    inc = ["-I", INCLUDE, "-I", src]
    shaders = _hip_obj(os.path.join(src, "shaders.hip"), os.path.join(INTEGRATION_BUILD, "shaders.o"), [INCLUDE, src])
    main_o = os.path.join(INTEGRATION_BUILD, "main.o")
    _run([CC, "-std=c11", "-O2", "-Wall"] + NUMERIC_FLAGS + inc + ["-c", os.path.join(src, "main.c"), "-o", main_o])
    rpath = os.path.relpath(PKG, INTEGRATION_BUILD)
    _run([HIPCC, f"--offload-arch={ARCH}", main_o, shaders, "-L", PKG, "-lsrp_amd", f"-Wl,-rpath,$ORIGIN/{rpath}",
          "-o", os.path.join(INTEGRATION_BUILD, "sample_gpu")])
    host = os.path.join(CSRC, "host")
    _run([CC, "-std=c11", "-O2", "-Wall", "-Wno-unused-function"] + NUMERIC_FLAGS + inc + ["-I", host] +
         [os.path.join(src, "main.c"), os.path.join(INTEGRATION_DIR, "host_shaders.c"),
          os.path.join(REPO, "oracle", "srp_oracle.c"), os.path.join(host, "png_decode.c"),
          os.path.join(host, "image_decode.c"),
          os.path.join(host, "png_encode.c"), "-lz", "-lm", "-o", os.path.join(INTEGRATION_BUILD, "sample_oracle")])
    return os.path.join(INTEGRATION_BUILD, "sample_gpu")


def build_oracle():
    """The CPU checkers (test infrastructure): the oracle and, when the
    reference sources are present, the in-place reference builds (a shared
    object for the parity tests, executables built as the reference ships for
    the CPU timing)."""
    _run(["make", "-s", "-C", os.path.join(REPO, "oracle")])
    if os.path.isdir("/root/reference/src"):
        _run(["make", "-s", "-j8", "-C", os.path.join(REPO, "oracle"), "ref", "refbench"])


def build_all(jobs=8):
    build_library(jobs)
    build_scenes(jobs)
    build_oracle()
    build_integration_sample()


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    {"all": build_all, "lib": build_library, "scenes": build_scenes, "oracle": build_oracle,
     "sample": build_integration_sample}[what]()
    print("built", what)
