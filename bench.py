"""Benchmark of the draw pipeline on one MI355X.

Workload (default): the indexed textured mesh — 1920x1080 RGBA8 colour,
FP32 depth, 8-bit stencil; a 256x512-segment UV sphere (131,841 vertices,
786,432 uint32 indices, 262,144 triangles); perspective MVP (fov 60 deg,
near 0.1, far 100); 1024x1024 seeded-noise texture, nearest sampling,
perspective-correct UV, flat per-face tag; scissor to the central 75%;
stencil INCR on depth pass; depth LESS against a +1 depth buffer.

A step is one frame: clear, depth/stencil init, the draw, until the
colour/depth/stencil buffers are final in device memory (inputs resident
in HBM; no host copies in the timed region).  The metric is shaded
fragments (fragment shader invocations, identical to the reference's
count) per second.  The CPU baseline is the oracle (a single-threaded C
restatement of the reference's draw path) on the same frame, one core.

    python bench.py [--steps K] [--warmup W] [--workload mesh|mix]
"""
import argparse
import ctypes
import json
import os
import sys
import tempfile
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "tests"))

from scene_runner import GPU_SO, ORACLE_SO, REF_SO, write_noise_png  # noqa: E402

HBM_PEAK_GBS = 8000.0


def load(path):
    lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
    lib.srpBenchMeshSetup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_char_p]
    lib.srpBenchMixSetup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    lib.srpBenchTrisSetup.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    lib.srpBenchTerrainSetup.argtypes = [ctypes.c_int] * 6
    lib.srpBenchRead.argtypes = [ctypes.c_void_p] * 3
    return lib


def stats(lib, fn):
    a = (ctypes.c_uint64 * 4)()
    getattr(lib, fn)(a)
    return list(a)


# The workloads (the benchmark configurations), each a dict of its sizes:
#   mesh     indexed textured UV sphere (the headline)
#   tris     random clip-space triangles, 512x512 (the plain-method workload)
#   mix      clipped triangles + lines + points, 1080p
#   terrain  1,000,000-triangle heightfield with the procedural material, 1080p
def workload(args, kind, scale=1.0):
    if kind == "mesh":
        return dict(kind=kind, w=args.width, h=args.height, rings=args.rings, segs=args.segs)
    if kind == "tris":
        return dict(kind=kind, w=512, h=512, tris=max(1, int(round(2000 * scale))), seed=0x5EED0001)
    if kind == "mix":
        return dict(kind=kind, w=args.width, h=args.height, tris=max(1, int(round(args.tris * scale))),
                    lines=max(1, int(round(args.lines * scale))), points=max(4, int(round(args.points * scale))),
                    seed=0x5EED0003)
    return dict(kind=kind, w=args.width, h=args.height, gx=args.grid_x, gz=args.grid_z, pose=256)


def describe(wl):
    k = wl["kind"]
    if k == "mesh":
        return (f"indexed textured UV sphere {wl['rings']}x{wl['segs']} segs, {2 * wl['rings'] * wl['segs']:,} tris, "
                f"{wl['w']}x{wl['h']}, scissor 75%, stencil INCR, depth LESS")
    if k == "tris":
        return f"{wl['tris']:,} random clip-space triangles, {wl['w']}x{wl['h']}, depth LESS, no culling"
    if k == "mix":
        return (f"{wl['tris']:,} tris + {wl['lines']:,} lines + {wl['points']:,} points, {wl['w']}x{wl['h']}, "
                f"LEQUAL, cull back")
    return (f"{wl['gx']}x{wl['gz']}-cell terrain ({2 * wl['gx'] * wl['gz']:,} tris), procedural fBm + Blinn-Phong "
            f"material, pose {wl['pose']}, {wl['w']}x{wl['h']}, depth LESS, cull back")


def setup(lib, wl, tex):
    k = wl["kind"]
    if k == "mesh":
        rc = lib.srpBenchMeshSetup(wl["w"], wl["h"], wl["rings"], wl["segs"], tex.encode())
    elif k == "tris":
        rc = lib.srpBenchTrisSetup(wl["w"], wl["h"], wl["tris"], wl["seed"])
    elif k == "mix":
        rc = lib.srpBenchMixSetup(wl["w"], wl["h"], wl["tris"], wl["lines"], wl["points"], wl["seed"])
    else:
        rc = lib.srpBenchTerrainSetup(wl["w"], wl["h"], wl["gx"], wl["gz"], wl["pose"], 0)
    if rc != 0:
        raise RuntimeError(f"bench setup failed: {rc}")


def input_triangles(wl):
    k = wl["kind"]
    return {"mesh": lambda: 2 * wl["rings"] * wl["segs"], "tris": lambda: wl["tris"], "mix": lambda: wl["tris"],
            "terrain": lambda: 2 * wl["gx"] * wl["gz"]}[k]()


def geometry_bytes(wl):
    """Vertex + index bytes one frame reads (each unique vertex once)."""
    k = wl["kind"]
    if k == "mesh":
        nv = (wl["rings"] + 1) * (wl["segs"] + 1)
        return nv * 20 + wl["rings"] * wl["segs"] * 6 * 4
    if k == "tris":
        return 3 * wl["tris"] * 24                      # {pos vec3, colour vec3}
    if k == "mix":
        # non-indexed {clip vec4, st vec2, level, tag} = 32 B per vertex
        return (3 * wl["tris"] + 2 * wl["lines"] + wl["points"]) * 32
    nv = (wl["gx"] + 1) * (wl["gz"] + 1)
    return nv * 32 + 2 * wl["gx"] * wl["gz"] * 3 * 4    # {pos, normal, uv} + u32 indices


def raster_bytes(wl, shaded_per_frame):
    """Algorithmic bytes of the triangle rasteriser per frame: colour, depth
    and stencil read and written once (W*H*9*2), plus one RGBA8 texel per
    shaded fragment for the textured mesh."""
    return wl["w"] * wl["h"] * 9 * 2 + shaded_per_frame * (4 if wl["kind"] == "mesh" else 0)


def gpu_frames(wl, tex, steps, warmup):
    """GPU frames of a workload (inputs resident, deferred framebuffer):
    wall time per frame, per-kernel HIP-event times, counters, last frame."""
    gpu = load(GPU_SO)
    setup(gpu, wl, tex)
    for _ in range(warmup):
        gpu.srpBenchFrame()
    gpu.srpHipSynchronize()
    gpu.srpHipResetStats()
    # the timed frames record events around the rasteriser only (each
    # interval costs the stream a barrier); the other stages are timed on
    # a few more frames afterwards
    gpu.srpHipProfileSlots(1)
    gpu.srpHipProfileEnable(True)
    gpu.srpHipSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        gpu.srpBenchFrame()
    gpu.srpHipSynchronize()
    dt = time.perf_counter() - t0
    ms = (ctypes.c_double * 4)()
    launches = (ctypes.c_uint64 * 4)()
    gpu.srpHipProfileRead(ms, launches)
    st = stats(gpu, "srpHipGetStats")
    gpu.srpHipProfileSlots(0xF)
    gpu.srpHipProfileEnable(True)
    staged = min(steps, 20)
    for _ in range(staged):
        gpu.srpBenchFrame()
    ms2 = (ctypes.c_double * 4)()
    launches2 = (ctypes.c_uint64 * 4)()
    gpu.srpHipProfileRead(ms2, launches2)
    gpu.srpHipProfileEnable(False)
    frame = read_frame(gpu, wl["w"] * wl["h"])
    gpu.srpBenchFinish()
    return {
        "ms_per_frame": dt / steps * 1e3,
        "shaded_per_frame": st[1] // steps,
        "covered_tested_per_frame": st[0] // steps,
        "primitives_per_frame": st[2] // steps,
        "draws_per_frame": st[3] // steps,
        "raster_ms": ms[0] / max(launches[0], 1),
        "draw_ms": ms2[1] / max(launches2[1], 1),
        "frag_ms": ms2[2] / max(launches2[2], 1) if launches2[2] else None,
        "front_ms": ms2[3] / max(launches2[3], 1),
        "raster_launches_per_frame": launches[0] / steps,
    }, frame


REF_EXE_DIR = os.path.join(REPO, "oracle", "_ref")


def ref_exe():
    """The reference built as it ships (oracle/Makefile target refbench: -O3
    -DNDEBUG, one executable) for the highest x86-64 level this host supports
    (its -march=native stands for the host that builds it), or None."""
    flags = set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        pass
    levels = []
    if {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl", "avx2", "fma", "bmi2"} <= flags:
        levels.append("v4")
    if {"avx2", "fma", "bmi1", "bmi2", "movbe"} <= flags:
        levels.append("v3")
    for lv in levels:
        p = os.path.join(REF_EXE_DIR, f"ref_bench_{lv}")
        if os.path.exists(p):
            return p
    return None


def run_ref_exe(exe, wl, tex, budget_s, min_frames, max_frames):
    """Frames of the as-shipped reference on one host core (a child process
    pinned there): per-frame seconds and the last frame."""
    import subprocess
    k = wl["kind"]
    keys = {"mesh": ("rings", "segs"), "tris": ("tris", "seed"), "mix": ("tris", "lines", "points", "seed"),
            "terrain": ("gx", "gz", "pose")}[k]
    args = [wl["w"], wl["h"]] + [wl[key] for key in keys] + ([tex or ""] if k == "mesh" else [])
    dump = os.path.join(tempfile.mkdtemp(prefix="srp_ref_"), "frame.bin")
    core = min(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None
    pre = (lambda: os.sched_setaffinity(0, {core})) if core is not None else None
    r = subprocess.run([exe, k] + [str(a) for a in args] + [str(budget_s), str(min_frames), str(max_frames), dump],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, preexec_fn=pre, timeout=900)
    if r.returncode != 0:
        raise RuntimeError(f"reference timing driver failed: {r.stderr}")
    times = json.loads(r.stdout.strip().splitlines()[-1])["frame_s"]
    n = wl["w"] * wl["h"]
    b = np.fromfile(dump, np.uint8)
    frame = (b[:4 * n].view(np.uint32).copy(), b[4 * n:8 * n].view(np.float32).copy(), b[8 * n:9 * n].copy())
    os.remove(dump)
    os.rmdir(os.path.dirname(dump))
    return times, frame


def reference_frames(wl, tex, budget_s, min_frames, max_frames):
    """The reference renderer timed on one host core: built as it ships
    (ref_exe), else compiled in place as a shared object (oracle/_ref), else
    the oracle; frames until budget_s (at least min_frames, at most
    max_frames) after one warm-up frame."""
    exe = ref_exe()
    if exe:
        times, frame = run_ref_exe(exe, wl, tex, budget_s, min_frames, max_frames)
        return "reference", times, frame
    path = REF_SO if os.path.exists(REF_SO) else ORACLE_SO
    lib = load(path)
    setup(lib, wl, tex)

    def run():
        lib.srpBenchFrame()                      # warm-up
        t0 = time.perf_counter()
        times = []
        while len(times) < min_frames or (time.perf_counter() - t0 < budget_s and len(times) < max_frames):
            t = time.perf_counter()
            lib.srpBenchFrame()
            times.append(time.perf_counter() - t)
        return times

    times = on_one_core(run)
    frame = read_frame(lib, wl["w"] * wl["h"])
    lib.srpBenchFinish()
    return ("reference" if path == REF_SO else "port"), times, frame


def frames_equal(a, b):
    return all(np.array_equal(x.view(np.uint32), y.view(np.uint32)) for x, y in zip(a, b))


def workload_leg(args, tex, kind, steps, warmup, cpu_scale, cpu_budget, cpu_min, cpu_max):
    """One workload at full size on the GPU, beside the CPU reference on a
    bounded sample of it (the same generator at cpu_scale of its primitive
    counts, or the full frame when cpu_scale is 1).  The CPU's shaded count
    is the GPU's on the same sample, whose frame must equal the reference's
    (checked)."""
    wl = workload(args, kind)
    g, gframe = gpu_frames(wl, tex, steps, warmup)
    out = {
        "workload": describe(wl),
        "frame_ms": round(g["ms_per_frame"], 4),
        "shaded_per_frame": int(g["shaded_per_frame"]),
        "shaded_mfrags_per_s": round(g["shaded_per_frame"] / g["ms_per_frame"] * 1e-3, 3),
        "triangles_per_s": round(input_triangles(wl) / (g["ms_per_frame"] * 1e-3), 1),
        "primitives_per_frame": int(g["primitives_per_frame"]),
        "kernel_ms": {"raster_tris": round(g["raster_ms"], 4), "fragment_lists": round(g["frag_ms"], 4)
                      if g["frag_ms"] is not None else None, "vertex_and_count_pass": round(g["front_ms"], 4)},
    }
    rb = raster_bytes(wl, g["shaded_per_frame"])
    if g["raster_ms"] > 0:
        ach = rb / (g["raster_ms"] * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "kernel": "srp_raster_batched", "achieved": round(ach, 2),
                           "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 6),
                           "algorithmic_bytes": int(rb),
                           "frame_bytes": int(rb + geometry_bytes(wl)),
                           "frame_frac": round((rb + geometry_bytes(wl)) / (g["ms_per_frame"] * 1e-3) / 1e9
                                               / HBM_PEAK_GBS, 6)}
    if cpu_budget <= 0:
        return out
    swl = workload(args, kind, cpu_scale)
    kind_cpu, times, cframe = reference_frames(swl, tex, cpu_budget, cpu_min, cpu_max)
    if cpu_scale == 1.0:
        sframe, sshaded = gframe, g["shaded_per_frame"]
    else:
        s, sframe = gpu_frames(swl, tex, 1, 0)
        sshaded = s["shaded_per_frame"]
    med = float(np.median(times))
    model, ncpu = host_cpu()
    out["cpu_baseline"] = {
        "value": round(sshaded / med / 1e6, 4), "unit": "shaded Mfrags/s", "cores": 1, "kind": kind_cpu,
        "sample": (f"{len(times)} frame(s) of " + ("the same frame" if cpu_scale == 1.0 else
                   f"the same generator at {cpu_scale:g} of its primitive counts ({describe(swl)})") +
                   f" after one warm-up frame; median {med * 1e3:.1f} ms/frame; shaded count of that frame "
                   f"{sshaded:,} (the GPU's, whose frame equals the reference's)"),
        "ms_per_frame_median": round(med * 1e3, 3),
        "host_cpu": model, "host_cpus_visible": ncpu,
    }
    out["parity_sample_vs_" + kind_cpu] = frames_equal(sframe, cframe)
    out["speedup_vs_cpu"] = round(out["shaded_mfrags_per_s"] / out["cpu_baseline"]["value"], 1) \
        if out["cpu_baseline"]["value"] else None
    return out


def read_frame(lib, n):
    c = np.zeros(n, np.uint32)
    d = np.zeros(n, np.float32)
    s = np.zeros(n, np.uint8)
    lib.srpBenchRead(c.ctypes.data, d.ctypes.data, s.ctypes.data)
    return c, d, s


def host_cpu():
    """Model name of the host CPU (from /proc/cpuinfo), and the cores visible."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count()


def on_one_core(fn):
    """Run fn() pinned to one host core (restoring the affinity after)."""
    cores = None
    try:
        cores = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(cores)})
    except (AttributeError, OSError):
        pass
    try:
        return fn()
    finally:
        if cores:
            os.sched_setaffinity(0, cores)


def reference_baseline(args, wl, tex, budget_s, port):
    """The reference renderer timed on one host core for ~budget_s: built
    as it ships (one -O3 executable, ref_exe) where that was built, else
    compiled in place as a shared object (oracle/_ref); its shaded count per
    frame is the oracle's (the frames are equal, which is checked).  Returns
    (baseline dict, frame)."""
    exe = ref_exe()
    if exe:
        times, frame = run_ref_exe(exe, wl, tex, budget_s, 5, 200)
        median = float(np.median(times))
        model, ncpu = host_cpu()
        shaded = port["shaded_per_frame"]
        return {
            "value": round(shaded / median / 1e6, 4),
            "unit": "shaded Mfrags/s",
            "cores": 1,
            "kind": "reference",
            "sample": f"{len(times)} full frames of the same workload by the reference built as it ships "
                      f"({os.path.basename(exe)}: -O3 -DNDEBUG -march=x86-64-{os.path.basename(exe)[-2:]}) after "
                      f"one warm-up frame, {sum(times):.1f} s, median {median * 1e3:.1f} ms/frame (value is the "
                      f"median frame; shaded count per frame from the oracle, whose frame it equals)",
            "ms_per_frame": round(sum(times) / len(times) * 1e3, 3),
            "ms_per_frame_median": round(median * 1e3, 3),
            "port_ms_per_frame": port["ms_per_frame"],
            "host_cpu": model,
            "host_cpus_visible": ncpu,
        }, frame
    lib = load(REF_SO)
    setup(lib, wl, tex)

    def run():
        lib.srpBenchFrame()                      # warm-up
        t0 = time.perf_counter()
        times = []
        while len(times) < 5 or (time.perf_counter() - t0 < budget_s and len(times) < 200):
            t = time.perf_counter()
            lib.srpBenchFrame()
            times.append(time.perf_counter() - t)
        return len(times), time.perf_counter() - t0, float(np.median(times))

    frames, dt, median = on_one_core(run)
    model, ncpu = host_cpu()
    frame = read_frame(lib, wl["w"] * wl["h"])
    lib.srpBenchFinish()
    shaded = port["shaded_per_frame"]
    return {
        "value": round(shaded / median / 1e6, 4),
        "unit": "shaded Mfrags/s",
        "cores": 1,
        "kind": "reference",
        "sample": f"{frames} full frames of the same workload by the compiled reference after one warm-up frame, "
                  f"{dt:.1f} s, median {median * 1e3:.1f} ms/frame (value is the median frame; shaded count per "
                  f"frame from the oracle, whose frame it equals)",
        "ms_per_frame": round(dt / frames * 1e3, 3),
        "ms_per_frame_median": round(median * 1e3, 3),
        "port_ms_per_frame": port["ms_per_frame"],
        "host_cpu": model,
        "host_cpus_visible": ncpu,
    }, frame


def cpu_baseline(args, wl, tex, budget_s):
    """Oracle frames on one host core until ~budget_s of work (at least 2)."""
    lib = load(ORACLE_SO)
    setup(lib, wl, tex)
    cores = None
    try:
        cores = os.sched_getaffinity(0)
        os.sched_setaffinity(0, {min(cores)})
    except (AttributeError, OSError):
        pass
    try:
        lib.srpBenchFrame()                      # warm-up
        lib.srpOracleResetStats()
        t0 = time.perf_counter()
        frames = 0
        while frames < 2 or (time.perf_counter() - t0 < budget_s and frames < 200):
            lib.srpBenchFrame()
            frames += 1
        dt = time.perf_counter() - t0
    finally:
        if cores:
            os.sched_setaffinity(0, cores)
    st = stats(lib, "srpOracleGetStats")
    lib.srpOracleBboxPixels.restype = ctypes.c_uint64
    bbox = int(lib.srpOracleBboxPixels())
    frame = read_frame(lib, wl["w"] * wl["h"])
    lib.srpBenchFinish()
    return {
        "value": round(st[1] / dt / 1e6, 4),
        "unit": "shaded Mfrags/s",
        "cores": 1,
        "kind": "port",
        "sample": f"{frames} full frames of the same workload, {dt:.1f} s, {dt / frames * 1e3:.1f} ms/frame",
        "ms_per_frame": round(dt / frames * 1e3, 3),
        "shaded_per_frame": st[1] // frames,
        "covered_per_frame": st[0] // frames,
        "bbox_pixels_per_frame": bbox // frames,
    }, frame, st[1] // frames


VALU_FP32_PEAK_TFLOPS = 157.3
# Reference operation count of the textured mesh frame, per pixel or fragment:
# 3 barycentric steps per bounding-box pixel (triangle.c:73-111 walks the box),
# per covered fragment 6 for the depth and 6 for the reciprocal w plus 10 per
# perspective-interpolated float (two: the uv), and per shaded fragment the
# fragment shader's 23 (nearest texel: wrap 4, scale 2, convert 2, index 2,
# unpack 6; shade factor 4; three products).
VALU_OPS_BBOX, VALU_OPS_COVERED, VALU_OPS_SHADED = 3, 6 + 6 + 2 * 10, 23


def valu_figure(bbox, covered, shaded, raster_ms, frame_ms):
    """Secondary bound: the reference's arithmetic for the frame against the
    FP32 vector peak (the rasteriser does all of it)."""
    ops = VALU_OPS_BBOX * bbox + VALU_OPS_COVERED * covered + VALU_OPS_SHADED * shaded
    achieved = ops / (raster_ms * 1e-3) / 1e12 if raster_ms > 0 else 0.0
    return {
        "bound": "valu",
        "ops_per_frame": int(ops),
        "bbox_pixels_per_frame": int(bbox),
        "achieved": round(achieved, 4),
        "peak": VALU_FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / VALU_FP32_PEAK_TFLOPS, 6),
        "frame_achieved": round(ops / (frame_ms * 1e-3) / 1e12, 4),
    }


MFMA_BF16_PEAK_TFLOPS = 2500.0
# algorithmic FLOPs of the network per fragment: 2*(512*64 + 64*64 + 64*3)
NN_FWD_FLOPS = 74112
# forward + input gradients of layers 3, 2 + weight gradients of all layers
NN_TRAIN_FLOPS = NN_FWD_FLOPS + 2 * (64 * 3 + 64 * 64 + 512 * 64 + 64 * 64 + 3 * 64)


def neural_leg(args):
    """The terrain's neural fragment shader: capture training fragments from
    256 poses with the ordinary pipeline, train the network on the GPU,
    evaluate it against the procedural shader on 64 held-out poses, and time
    frames with each shader."""
    from neural_runner import Neural
    nn = Neural()
    W, H = args.width, args.height
    if nn.terrain_setup(W, H, args.grid_x, args.grid_z, 0, 0) != 0:
        raise RuntimeError("terrain setup failed")
    ds = nn.new_ds(8_000_000)
    net = nn.new_net(0x5EED0005)
    if args.nn_freq:
        st = [float(v) for v in args.nn_freq.split(",")]
        nn.set_encoding(net, st * 8 if len(st) == 1 else st)
    out = {"freq_steps": [round(float(v), 4) for v in nn.get_encoding(net)]}
    try:
        t0 = time.perf_counter()
        got = nn.terrain_collect(ds, 0, 256, args.nn_keep, 0x5EED0005)
        nn.sync()
        out["train_fragments"] = int(got)
        out["capture_s"] = round(time.perf_counter() - t0, 3)
        dev = nn.net_device(net)

        def evaluate(poses):
            ps, mx = [], []
            for pose in poses:
                nn.terrain_pose(pose)
                nn.use_fb(0)
                nn.terrain_shader(0, None, None)
                nn.frame()
                fa = nn.fb()
                nn.use_fb(1)
                nn.terrain_shader(0, nn.neural_fs, dev)
                nn.frame()
                p, m, _ = nn.compare(fa, nn.fb(), fa)
                ps.append(p)
                mx.append(m)
            nn.use_fb(0)
            return ps, mx

        # poses 0-255 give the training fragments, 320-351 are the validation
        # poses the encoding was chosen on (--nn-select), and 256-319 are the
        # test poses, used for nothing but this report
        test_poses = range(256, 256 + 64)
        val_poses = range(320, 320 + 32)
        ps0, _ = evaluate(test_poses[:8])
        nn.sync()
        t0 = time.perf_counter()
        loss = nn.train(net, ds, args.nn_steps, args.nn_batch, args.nn_lr, args.nn_lr * 0.02, 0x5EED0005)
        t_train = time.perf_counter() - t0
        samples = args.nn_steps * args.nn_batch
        out["train"] = {"steps": args.nn_steps, "batch": args.nn_batch, "s": round(t_train, 3),
                        "ms_per_step": round(t_train / args.nn_steps * 1e3, 4),
                        "loss_first": float(loss[0]), "loss_last": float(np.mean(loss[-50:])),
                        "mfma_tflops": round(samples * NN_TRAIN_FLOPS / t_train / 1e12, 2),
                        "mfma_frac": round(samples * NN_TRAIN_FLOPS / t_train / 1e12 / MFMA_BF16_PEAK_TFLOPS, 5)}
        pv, mv = evaluate(val_poses)
        ps, mx = evaluate(test_poses)
        out["psnr_db_val_mean"] = round(float(np.mean(pv)), 3)
        out["psnr_db_val_min"] = round(float(np.min(pv)), 3)
        out["psnr_db_mean"] = round(float(np.mean(ps)), 3)
        out["psnr_db_min"] = round(float(np.min(ps)), 3)
        out["max_abs_err"] = float(np.max(mx))
        out["psnr_db_untrained"] = round(float(np.mean(ps0)), 3)
        out["eval_poses"] = {"test": [test_poses.start, test_poses.stop - 1],
                             "validation": [val_poses.start, val_poses.stop - 1], "training": [0, 255]}

        # the network and the procedural material alone, dense over the same
        # captured fragments (no rasteriser around them)
        nd = min(int(got), 1 << 20)
        ms_net = nn.time_inference(net, ds, nd, 10)
        feats, _ = nn.ds_read(ds, 0, nd)
        ms_plain = nn.terrain_dense_ms(feats, nd, 10)
        ps_net, ps_plain = ms_net * 1e9 / nd, ms_plain * 1e9 / nd
        ps_ideal = NN_FWD_FLOPS / (MFMA_BF16_PEAK_TFLOPS * 1e12) * 1e12
        out["dense"] = {"fragments": nd, "network_ps_per_fragment": round(ps_net, 2),
                        "procedural_ps_per_fragment": round(ps_plain, 2),
                        "network_speedup": round(ps_plain / ps_net, 3) if ps_net > 0 else None,
                        "network_mfma_tflops": round(NN_FWD_FLOPS / ps_net, 2) if ps_net > 0 else None,
                        "network_mfma_frac": round(NN_FWD_FLOPS / ps_net / MFMA_BF16_PEAK_TFLOPS, 5) if ps_net > 0 else None,
                        "network_ps_at_full_mfma": round(ps_ideal, 2),
                        "speedup_bound_at_full_mfma": round(ps_plain / ps_ideal, 3)}

        def time_frames(capture, fs, netdev, pose=256):
            nn.use_fb(0)
            nn.terrain_pose(pose)
            nn.terrain_shader(capture, fs, netdev)
            for _ in range(3):
                nn.frame()
            nn.sync()
            nn.reset_stats()
            nn.profile_slots(1)   # events around the rasteriser (and deferred shading) only
            nn.profile_enable(True)
            nn.sync()
            t0 = time.perf_counter()
            for _ in range(args.nn_frames):
                nn.frame()
            nn.sync()
            dt = time.perf_counter() - t0
            ms, n = nn.profile_read()
            nn.profile_enable(False)
            nn.profile_slots(0xF)
            st = nn.stats()
            return dt / args.nn_frames * 1e3, ms[0] / max(n[0], 1), st[1] // args.nn_frames

        plain_ms, plain_raster, shaded = time_frames(0, None, None)
        # the network deferred (once per visible pixel after the rasteriser),
        # then shaded inside the rasteriser for every fragment that passes
        # the depth test (the library's default)
        nn.set_deferred(1)
        dfr_ms, dfr_raster, _ = time_frames(0, nn.neural_fs, dev)
        nn.set_deferred(0)
        neural_ms, neural_raster, shaded_n = time_frames(0, nn.neural_fs, dev)
        out["frame_ms_neural_deferred"] = round(dfr_ms, 4)
        out["raster_ms_neural_deferred"] = round(dfr_raster, 4)
        # the same frame with a trivial fragment shader (capture pass 0): what
        # the rasteriser costs without shading work
        base_ms, base_raster, _ = time_frames(1, None, None)
        out["frame_ms_trivial_fs"] = round(base_ms, 4)
        out["raster_ms_trivial_fs"] = round(base_raster, 4)
        out["fs_ms_plain"] = round(plain_raster - base_raster, 4)
        out["fs_ms_neural"] = round(neural_raster - base_raster, 4)
        out["neural_speedup_fs"] = round((plain_raster - base_raster) / max(neural_raster - base_raster, 1e-6), 3)
        out["frame_ms_plain"] = round(plain_ms, 4)
        out["frame_ms_neural"] = round(neural_ms, 4)
        out["raster_ms_plain"] = round(plain_raster, 4)
        out["raster_ms_neural"] = round(neural_raster, 4)
        out["shaded_per_frame"] = int(shaded)
        out["neural_speedup_frame"] = round(plain_ms / neural_ms, 3)
        out["neural_speedup_raster"] = round(plain_raster / neural_raster, 3)
        # the visible pixels (one draw after a clear: depth below 1); the
        # network's useful work is one evaluation per visible pixel
        _, depth, _ = nn.read_frame(args.width, args.height)
        resolved = int(np.count_nonzero(depth < 1.0))
        out["resolved_pixels_per_frame"] = resolved
        out["neural_raster_mfma_tflops"] = round(resolved * NN_FWD_FLOPS / (neural_raster * 1e-3) / 1e12, 2)
        out["neural_raster_mfma_frac"] = round(resolved * NN_FWD_FLOPS / (neural_raster * 1e-3) / 1e12
                                               / MFMA_BF16_PEAK_TFLOPS, 5)
    finally:
        nn.free_net(net)
        nn.free_ds(ds)
        nn.finish()
    return out


NN_SELECT_CANDIDATES = [
    [0.5] * 8,
    [0.75, 0.5, 0.5, 0.5, 0.5, 0.5, 1.25, 0.75],
    [1.0, 0.5, 0.5, 0.5, 0.5, 0.5, 2.0, 1.0],
    [1.0, 0.5, 0.75, 0.5, 0.5, 0.5, 2.0, 1.0],
    [1.5, 0.5, 0.75, 0.5, 0.5, 0.5, 3.0, 1.5],
    [1.0, 0.75, 0.5, 0.25, 0.25, 0.25, 2.0, 1.0],
]


def neural_select(args):
    """Choose the encoding's per-feature frequency steps on the validation
    poses 320-351 (never on the test poses 256-319): each candidate network
    is trained as the neural leg trains (same data, schedule and seed) and
    scored by its mean PSNR over the validation poses."""
    from neural_runner import Neural
    nn = Neural()
    if nn.terrain_setup(args.width, args.height, args.grid_x, args.grid_z, 0, 0) != 0:
        raise RuntimeError("terrain setup failed")
    ds = nn.new_ds(8_000_000)
    table = []
    try:
        got = nn.terrain_collect(ds, 0, 256, args.nn_keep, 0x5EED0005)
        nn.sync()
        for steps in NN_SELECT_CANDIDATES:
            net = nn.new_net(0x5EED0005)
            try:
                nn.set_encoding(net, steps)
                nn.train(net, ds, args.nn_steps, args.nn_batch, args.nn_lr, args.nn_lr * 0.02, 0x5EED0005)
                dev = nn.net_device(net)
                ps = []
                for pose in range(320, 352):
                    nn.terrain_pose(pose)
                    nn.use_fb(0)
                    nn.terrain_shader(0, None, None)
                    nn.frame()
                    fa = nn.fb()
                    nn.use_fb(1)
                    nn.terrain_shader(0, nn.neural_fs, dev)
                    nn.frame()
                    ps.append(nn.compare(fa, nn.fb(), fa)[0])
                nn.use_fb(0)
                table.append({"freq_steps": steps, "psnr_db_val_mean": round(float(np.mean(ps)), 3),
                              "psnr_db_val_min": round(float(np.min(ps)), 3)})
                print(f"bench.py: encoding {steps}: validation {np.mean(ps):.3f} dB", file=sys.stderr, flush=True)
            finally:
                nn.free_net(net)
    finally:
        nn.free_ds(ds)
        nn.finish()
    best = max(table, key=lambda r: r["psnr_db_val_mean"])
    return {"train_fragments": int(got), "steps": args.nn_steps, "validation_poses": [320, 351],
            "candidates": table, "best": best["freq_steps"]}


def batch_leg(args):
    """The batched-scenes workload: 64 independent scenes rendered as one
    batch (one persistent rasteriser launch for all their tiles), tuned by
    the autotuner, captured into a HIP graph and replayed; compared with
    drawing the scenes one at a time on the same GPU, and with the CPU
    oracle (one frame)."""
    from batch_runner import C4_SEED, BatchScenes, frames_equal
    from neural_runner import Neural
    nn = Neural()
    gb = BatchScenes(GPU_SO, gpu=True)
    gb.setup(args.batch_scenes, 1.0, C4_SEED, batched=True)
    out = {}
    try:
        info = gb.info()
        out["scenes"] = args.batch_scenes
        out["triangles"] = int(sum(t for _, t in gb.scenes))
        out["primitives_after_culling"] = int(info[1])
        out["tile_pairs"] = int(info[2])
        out["tiles"] = int(info[3])

        def timed(fn, frames):
            fn()
            nn.sync()
            nn.reset_stats()
            t0 = time.perf_counter()
            for _ in range(frames):
                fn()
            nn.sync()
            dt = (time.perf_counter() - t0) / frames
            return dt * 1e3, nn.stats()[1] // frames

        seq_ms, seq_shaded = timed(lambda: gb.frame(False), 3)
        seq = gb.read_all()
        eager_default_ms, _ = timed(lambda: gb.frame(True), 5)
        t0 = time.perf_counter()
        tuned_ms, tuning = gb.autotune(3)
        out["autotune"] = {"s": round(time.perf_counter() - t0, 3), "best_ms": round(tuned_ms, 4),
                           "waves_per_cu": tuning[0], "mask_box": tuning[1], "xcd_queues": tuning[2],
                           "pixel_runs": tuning[3], "tile_size": tuning[4],
                           "waves_per_tile": 1 if tuning[4] == 8 else 4, "candidates": 56}
        # the best of each tile size (8x8 a wave a tile; 16x16 four waves a tile)
        by_tile = {}
        for tile in (8, 16):
            best = None
            for w in (4, 8, 16, 28):
                gb.set_tuning(w, tuning[1] if tile == 8 else 24, tuning[2], tuning[3] if tile == 8 else 1, tile)
                ms_t, _ = timed(lambda: gb.frame(True), 3)
                best = ms_t if best is None else min(best, ms_t)
            by_tile[str(tile)] = round(best, 4)
        out["autotune"]["best_ms_by_tile_size"] = by_tile
        gb.set_tuning(*tuning)
        eager_ms, shaded = timed(lambda: gb.frame(True), 10)
        assert gb.capture() == 0
        graph_ms, g_shaded = timed(lambda: gb.frame(True), args.batch_frames)
        got = gb.read_all()
        out["frame_ms_graph"] = round(graph_ms, 4)
        out["frame_ms_eager"] = round(eager_ms, 4)
        out["frame_ms_eager_default_tuning"] = round(eager_default_ms, 4)
        out["frame_ms_draw_by_draw"] = round(seq_ms, 4)
        out["shaded_per_frame"] = int(g_shaded)
        out["shaded_mfrags_per_s"] = round(g_shaded / (graph_ms * 1e-3) / 1e6, 2)
        out["triangles_per_s"] = round(out["triangles"] / (graph_ms * 1e-3), 1)
        out["parity_vs_draw_by_draw"] = frames_equal(got, seq) == [] and g_shaded == seq_shaded == shaded
        if not args.no_cpu_baseline:
            ob = BatchScenes(ORACLE_SO, gpu=False)
            ob.setup(args.batch_scenes, 1.0, C4_SEED, batched=False)
            try:
                t0 = time.perf_counter()
                ob.frame(False)
                cpu_s = time.perf_counter() - t0
                ref = ob.read_all()
            finally:
                ob.finish()
            out["parity_vs_oracle"] = frames_equal(got, ref) == []
            out["cpu_frame_s"] = round(cpu_s, 3)
            out["speedup_vs_cpu"] = round(cpu_s * 1e3 / graph_ms, 1)
    finally:
        gb.finish()
    return out


def copy_probe(mib, reps=5):
    """Plain device copy of a known size: its FETCH_SIZE / WRITE_SIZE per
    dispatch calibrate the counters, its rate is the achievable copy BW."""
    lib = ctypes.CDLL(GPU_SO, mode=ctypes.RTLD_LOCAL)
    lib.srpBenchCopyProbe.argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    nbytes = mib << 20
    ms = ctypes.c_double(0.0)
    rc = lib.srpBenchCopyProbe(nbytes, reps, ctypes.byref(ms))
    if rc != 0:
        raise RuntimeError(f"copy probe failed: {rc}")
    print(json.dumps({"copy_probe_bytes_read": nbytes, "copy_probe_bytes_written": nbytes, "dispatches": reps + 1,
                      "ms_per_copy": round(ms.value, 4),
                      "GB_per_s": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--workload", choices=["mesh", "mix", "tris", "terrain"], default="mesh")
    ap.add_argument("--no-legs", action="store_true", help="skip the primitive-mix, random-triangle and terrain legs")
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--rings", type=int, default=256)
    ap.add_argument("--segs", type=int, default=512)
    ap.add_argument("--tris", type=int, default=500000)
    ap.add_argument("--lines", type=int, default=100000)
    ap.add_argument("--points", type=int, default=100000)
    ap.add_argument("--tile", type=int, default=0, help="raster tile side (8 or 16); 0 = library default")
    ap.add_argument("--tile-order", type=int, default=-1,
                    help="tiles dealt to the XCDs in runs of this many (0: one range per XCD); -1 = library default")
    ap.add_argument("--mask-box", type=int, default=-1,
                    help="largest triangle box side given exact binning masks (0..24); -1 = library default")
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of oracle work for the CPU baseline")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--grid-x", type=int, default=1000, help="terrain cells along x (neural leg)")
    ap.add_argument("--grid-z", type=int, default=500, help="terrain cells along z (neural leg)")
    ap.add_argument("--nn-keep", type=float, default=0.0135, help="fraction of covered pixels kept per training pose")
    ap.add_argument("--nn-steps", type=int, default=150000)
    ap.add_argument("--nn-batch", type=int, default=8192)
    ap.add_argument("--nn-lr", type=float, default=5e-3)
    ap.add_argument("--nn-frames", type=int, default=50)
    ap.add_argument("--nn-freq", default="1.0,0.5,0.75,0.5,0.5,0.5,2.0,1.0",
                    help="encoding frequency step: one value, or 8 comma-separated (per feature: position xyz, "
                         "normal xyz, uv); the default is the best of NN_SELECT_CANDIDATES on the validation "
                         "poses 320-351 (--nn-select), never on the test poses 256-319 that the leg reports")
    ap.add_argument("--neural-only", action="store_true", help="run only the neural-shader leg and print its JSON")
    ap.add_argument("--nn-select", action="store_true",
                    help="only choose the encoding's frequency steps on the validation poses and print the table")
    ap.add_argument("--no-neural", action="store_true", help="skip the neural-shader leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-frames", type=int, default=20)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-scenes leg")
    ap.add_argument("--batch-only", action="store_true", help="run only the batched-scenes leg and print its JSON")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="only run the plain copy kernel over this many MiB (calibrates byte counters) and exit")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: the draw pipeline runs on one GPU; --gpus {args.gpus} is not supported "
              "(a frame would split by screen tiles, a batch by scene; neither is implemented)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe_mb:
        copy_probe(args.copy_probe_mb)
        return
    if args.nn_select:
        print(json.dumps({"neural_select": neural_select(args)}))
        return
    if args.neural_only:
        print(json.dumps({"neural": neural_leg(args)}))
        return
    if args.batch_only:
        print(json.dumps({"batch": batch_leg(args)}))
        return

    tmp = tempfile.mkdtemp(prefix="srp_bench_")
    tex = write_noise_png(os.path.join(tmp, "noise1024.png"), 1024, 0x5EED0002)
    wl = workload(args, args.workload)

    gpu = load(GPU_SO)
    if args.tile:
        gpu.srpHipSetTileSize(args.tile)
    if args.mask_box >= 0:
        gpu.srpHipSetRasterOptions.argtypes = [ctypes.c_int] * 3
        gpu.srpHipSetRasterOptions(-1, -1, args.mask_box)
    if args.tile_order >= 0:
        gpu.srpHipSetTileOrder(args.tile_order)
    tile = int(gpu.srpHipGetTileSize())
    g, gframe = gpu_frames(wl, tex, args.steps, args.warmup)
    steps = args.steps
    shaded_per_frame = g["shaded_per_frame"]
    ms_per_step = g["ms_per_frame"]
    value = shaded_per_frame / (ms_per_step * 1e-3) / 1e6
    raster_ms = g["raster_ms"]

    # Dominant kernel: the triangle rasteriser.  Algorithmic bytes per launch
    # (the per-frame figure of the workload definition): colour + depth +
    # stencil of the frame read and written once (W*H*9*2) plus one RGBA8
    # texel per shaded fragment (mesh workload).  The whole frame adds the
    # vertex and index data: B_frame = VB + IB + W*H*9*2 + shaded*T_tex.
    rbytes = raster_bytes(wl, shaded_per_frame)
    achieved = rbytes / (raster_ms * 1e-3) / 1e9 if raster_ms > 0 else 0.0
    frame_bytes = rbytes + geometry_bytes(wl)
    traffic = None
    tpath = os.path.join(REPO, "profiles", f"raster_{args.workload}_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("bytes_per_launch")
    names = {"mesh": "indexed textured mesh", "mix": "primitive mix", "tris": "random triangles, 512x512",
             "terrain": "terrain, procedural material"}

    result = {
        "metric": f"shaded Mfrags/s @1080p ({names[args.workload]})",
        "value": round(value, 3),
        "unit": "shaded Mfrags/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": {"mesh": "synthetic (seeded UV sphere + noise texture)", "mix": "synthetic (seeded clip-space primitives)",
                 "tris": "synthetic (seeded clip-space triangles)",
                 "terrain": "synthetic (seeded fBm heightfield)"}[args.workload],
        "config": {"workload": describe(wl), "width": wl["w"], "height": wl["h"], "tile": tile,
                   "tile_order": int(gpu.srpHipGetTileOrder())},
        "frame_ms": round(ms_per_step, 4),
        "triangles_per_s": round(input_triangles(wl) / (ms_per_step * 1e-3), 1),
        "shaded_per_frame": int(shaded_per_frame),
        "primitives_per_frame": int(g["primitives_per_frame"]),
        "draws_per_frame": int(g["draws_per_frame"]),
        "kernel_ms": {
            "raster_tris": round(raster_ms, 4),
            "draw_total": round(g["draw_ms"], 4),
            "fragment_lists": round(g["frag_ms"], 4) if g["frag_ms"] is not None else None,
            "vertex_and_count_pass": round(g["front_ms"], 4),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "srp_raster_batched",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 5),
            "traffic": traffic,
            "algorithmic_bytes": int(rbytes),
            "frame_bytes": int(frame_bytes),
            "frame_achieved": round(frame_bytes / (ms_per_step * 1e-3) / 1e9, 2),
        },
        "cpu_baseline": None,
    }

    if not args.no_cpu_baseline:
        # the oracle (a C restatement of the reference) checks the frame and
        # counts covered fragments and box pixels; the compiled reference,
        # where it was built, is the timed baseline
        have_ref = os.path.exists(REF_SO) or ref_exe() is not None
        port_budget = min(args.cpu_budget, 3.0) if have_ref else args.cpu_budget
        cpu, cframe, cpu_shaded = cpu_baseline(args, wl, tex, port_budget)
        bbox_per_frame = cpu["bbox_pixels_per_frame"]
        covered_per_frame = cpu["covered_per_frame"]
        result["covered_per_frame"] = int(covered_per_frame)
        result["parity_vs_oracle"] = bool(frames_equal(gframe, cframe) and cpu_shaded == shaded_per_frame)
        if have_ref:
            cpu, rframe = reference_baseline(args, wl, tex, args.cpu_budget, cpu)
            result["parity_vs_reference"] = frames_equal(gframe, rframe)
        if args.workload == "mesh" and cpu_shaded == shaded_per_frame:
            result["valu"] = valu_figure(bbox_per_frame, covered_per_frame, shaded_per_frame,
                                         raster_ms, ms_per_step)
        for k in ("shaded_per_frame", "covered_per_frame", "bbox_pixels_per_frame"):
            cpu.pop(k, None)
        result["cpu_baseline"] = cpu
        result["speedup_vs_cpu"] = round(value / cpu["value"], 1) if cpu["value"] else None

    if args.workload == "mesh" and not args.no_legs:
        cpu_on = not args.no_cpu_baseline
        print("bench.py: primitive-mix leg (1080p, 500k tris + 100k lines + 100k points)", file=sys.stderr, flush=True)
        result["mix"] = workload_leg(args, tex, "mix", 10, 2, 1.0 / 25.0, 5.0 if cpu_on else 0.0, 1, 3)
        print("bench.py: random-triangle leg (512x512, 2,000 tris)", file=sys.stderr, flush=True)
        result["tris512"] = workload_leg(args, tex, "tris", 50, 3, 1.0, 4.0 if cpu_on else 0.0, 1, 5)
        print("bench.py: terrain leg (1080p, 1M tris, procedural material)", file=sys.stderr, flush=True)
        result["terrain"] = workload_leg(args, tex, "terrain", 50, 3, 1.0, 3.0 if cpu_on else 0.0, 2, 8)
    if not args.no_batch and args.workload == "mesh":
        print("bench.py: batched-scenes leg (64 scenes, autotune, graph)", file=sys.stderr, flush=True)
        result["batch"] = batch_leg(args)
    if not args.no_neural and args.workload == "mesh":
        print("bench.py: neural shader leg (capture, training, evaluation)", file=sys.stderr, flush=True)
        result["neural"] = neural_leg(args)

    print(json.dumps(result))


if __name__ == "__main__":
    main()
