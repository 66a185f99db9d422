"""ctypes driver of the batched-scenes workload (tests/scenes/batch.c) in a
scene module: the GPU build renders it as one batch (srpBatchRender, a
captured HIP graph once captured) or draw by draw; the oracle build draw by
draw."""
import ctypes

import numpy as np

C4_SEED = 0x5EED0004


class BatchScenes:
    def __init__(self, path, gpu):
        self.lib = ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)
        L = self.lib
        self.gpu = gpu
        L.srpBenchBatchSetup.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_uint64, ctypes.c_int]
        L.srpBenchBatchSetup.restype = ctypes.c_int
        L.srpBenchBatchFrame.argtypes = [ctypes.c_int]
        L.srpBenchBatchScene.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_uint64)]
        L.srpBenchBatchRead.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        if gpu:
            L.srpBenchBatchCapture.restype = ctypes.c_int
            L.srpBenchBatchAutotune.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int * 5)]
            L.srpBenchBatchAutotune.restype = ctypes.c_double
            L.srpBenchBatchSetTuning.argtypes = [ctypes.c_int] * 5
            L.srpBenchBatchInfo.argtypes = [ctypes.POINTER(ctypes.c_uint64 * 4)]
        self.n = 0

    def setup(self, n, scale=1.0, seed=C4_SEED, batched=True):
        rc = self.lib.srpBenchBatchSetup(n, scale, seed, 1 if batched else 0)
        if rc != 0:
            raise RuntimeError(f"batch setup failed: {rc}")
        self.n = n
        self.scenes = []
        for k in range(n):
            s = ctypes.c_int()
            t = ctypes.c_uint64()
            self.lib.srpBenchBatchScene(k, ctypes.byref(s), ctypes.byref(t))
            self.scenes.append((s.value, t.value))

    def frame(self, batched=True):
        self.lib.srpBenchBatchFrame(1 if batched else 0)

    def read(self, k):
        side = self.scenes[k][0]
        c = np.empty(side * side, np.uint32)
        d = np.empty(side * side, np.float32)
        s = np.empty(side * side, np.uint8)
        self.lib.srpBenchBatchRead(k, c.ctypes.data, d.ctypes.data, s.ctypes.data)
        return c, d, s

    def read_all(self):
        return [self.read(k) for k in range(self.n)]

    def capture(self):
        return self.lib.srpBenchBatchCapture()

    def autotune(self, trials):
        """Mean frame ms of the best candidate and it: (waves/CU, mask box,
        xcd queues, pixel runs, tile size)."""
        out = (ctypes.c_int * 5)()
        ms = self.lib.srpBenchBatchAutotune(trials, ctypes.byref(out))
        return ms, tuple(out)

    def set_tuning(self, waves, box, xcd, runs=1, tile=8):
        self.lib.srpBenchBatchSetTuning(waves, box, xcd, runs, tile)

    def regenerate(self, k, seed):
        self.lib.srpBenchBatchRegenerate.argtypes = [ctypes.c_int, ctypes.c_uint64]
        self.lib.srpBenchBatchRegenerate(k, seed)

    def drop_scene(self, k):
        self.lib.srpBenchBatchDropScene.argtypes = [ctypes.c_int]
        self.lib.srpBenchBatchDropScene(k)

    def info(self):
        a = (ctypes.c_uint64 * 4)()
        self.lib.srpBenchBatchInfo(ctypes.byref(a))
        return list(a)

    def finish(self):
        self.lib.srpBenchBatchFinish()
        self.n = 0


def frames_equal(a, b):
    """Number of scenes whose colour, depth or stencil differ."""
    bad = []
    for k, (x, y) in enumerate(zip(a, b)):
        if not (np.array_equal(x[0], y[0]) and np.array_equal(x[1].view(np.uint32), y[1].view(np.uint32))
                and np.array_equal(x[2], y[2])):
            bad.append(k)
    return bad
