"""ba_eval_cost_noalloc (the allocation-free cost entry point) against ba_eval_cost on the same
problems: two calls in a row, the second on a smaller problem, so the shared per-GPU work space
is reused with fewer cameras and observations.  The two entry points run the same kernels over the
same block partition, so the costs agree bit for bit.

The test allocates the device copies of the problems itself through the HIP runtime the library is
linked against (the entry point's callers own their memory); the entry point allocates nothing."""
import ctypes

import numpy as np
import pytest

import slamplay_amd as S
from slamplay_amd import _capi as C

H2D, D2H = 1, 2


def _hip():
    C.lib()  # the library's own HIP runtime is loaded first; libamdhip64 below resolves to it
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipStreamCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
    h.hipStreamSynchronize.argtypes = [ctypes.c_void_p]
    h.hipStreamDestroy.argtypes = [ctypes.c_void_p]
    h.hipSetDevice.argtypes = [ctypes.c_int]
    return h


class _DeviceProblem:
    """Device copies of a problem's arrays (owned by the caller, freed on close)."""

    def __init__(self, hip, p):
        self.hip = hip
        self.bufs = []
        arrays = [np.ascontiguousarray(a) for a in (p.camera_index(), p.point_index(), p.observations(),
                                                    p.mutable_cameras(), p.mutable_points())]
        ptrs = [self._upload(a) for a in arrays]
        v = C.ProblemView()
        v.num_cameras, v.num_points, v.num_observations = p.num_cameras(), p.num_points(), p.num_observations()
        v.camera_block_size = 9
        v.camera_index = ctypes.cast(ptrs[0], C.i32p)
        v.point_index = ctypes.cast(ptrs[1], C.i32p)
        v.observations = ctypes.cast(ptrs[2], C.f64p)
        v.cameras = ctypes.cast(ptrs[3], C.f64p)
        v.points = ctypes.cast(ptrs[4], C.f64p)
        self.view = v
        self.cost = self._alloc(8)

    def _alloc(self, n):
        ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(ptr), max(n, 8)) == 0
        self.bufs.append(ptr)
        return ptr

    def _upload(self, a):
        ptr = self._alloc(a.nbytes)
        assert self.hip.hipMemcpy(ptr, a.ctypes.data, a.nbytes, H2D) == 0
        return ptr

    def read_cost(self):
        out = np.zeros(1)
        assert self.hip.hipMemcpy(out.ctypes.data, self.cost, 8, D2H) == 0
        return float(out[0])

    def close(self):
        for b in self.bufs:
            self.hip.hipFree(b)
        self.bufs = []


@pytest.mark.gpu
def test_noalloc_cost_matches_eval_cost():
    hip = _hip()
    assert hip.hipSetDevice(0) == 0
    stream = ctypes.c_void_p()
    assert hip.hipStreamCreate(ctypes.byref(stream)) == 0
    # configuration 2 (250 cameras, 473,296 observations), then configuration 1 (16 cameras, 14,853)
    problems = [S.synthetic(2, seed=1), S.synthetic(1, seed=1)]
    assert problems[1].num_observations() < problems[0].num_observations()
    for p in problems:
        print("noalloc cost: %d cameras, %d observations" % (p.num_cameras(), p.num_observations()), flush=True)
        for loss, scale in ((C.LOSS_HUBER, 1.0), (C.LOSS_NONE, 1.0)):
            ref = S.eval_cost(p, loss=loss, loss_scale=scale)
            dp = _DeviceProblem(hip, p)
            try:
                C.check(C.lib().ba_eval_cost_noalloc(ctypes.byref(dp.view), loss, scale, dp.cost, stream))
                assert hip.hipStreamSynchronize(stream) == 0
                got = dp.read_cost()
            finally:
                dp.close()
            assert got == ref, (got, ref)
    assert hip.hipStreamDestroy(stream) == 0


def test_noalloc_rejects_sizes_beyond_its_buffer():
    """Sizes above the fixed work space are refused before anything is enqueued (no device use)."""
    v = C.ProblemView()
    v.num_cameras, v.num_points, v.num_observations, v.camera_block_size = 5000, 10, 10, 9
    with pytest.raises(C.BAError) as e:
        C.check(C.lib().ba_eval_cost_noalloc(ctypes.byref(v), C.LOSS_HUBER, 1.0, ctypes.c_void_p(8), None))
    assert e.value.status == 6  # BA_ERR_CAPACITY
    v.num_cameras, v.num_observations = 10, (1 << 24) + 1
    with pytest.raises(C.BAError) as e:
        C.check(C.lib().ba_eval_cost_noalloc(ctypes.byref(v), C.LOSS_HUBER, 1.0, ctypes.c_void_p(8), None))
    assert e.value.status == 6  # BA_ERR_CAPACITY
