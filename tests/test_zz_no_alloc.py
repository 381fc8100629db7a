"""rrlbm_diag_flow_noalloc: the diag_flow sums of caller-owned device fields
with the partial sums in the library's fixed device array (no device memory
allocated by the call), against rrlbm_diag_flow of a context holding the same
fields: the same kernels, so the same bits.  Two calls in a row, the second
on fewer cells (the array keeps the first call's partial sums, which the second
must not read), then the size limits."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class HipBuffers:
    """Device buffers from the HIP runtime librrlbm.so links (already loaded
    in this process): the test allocates the caller's fields, the entry point
    under test allocates nothing."""

    def __init__(self):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.hip.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(p), a.nbytes) == 0
        self.ptrs.append(p.value)
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # hipMemcpyHostToDevice
        return p.value

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(ctypes.c_void_p(p))
        self.ptrs = []


def fields(n, seed):
    rng = np.random.default_rng(seed)
    return (1.0 + 0.01 * rng.standard_normal(n), 0.05 * rng.standard_normal(n), 0.05 * rng.standard_normal(n),
            0.05 * rng.standard_normal(n))


def diag_by_context(rr, dims, f):
    with rr.Solver(lx=dims[0], ly=dims[1], lz=dims[2], storage=1) as s:
        for fld, a in zip((rr.Field.RHO, rr.Field.UX, rr.Field.UY, rr.Field.UZ), f):
            s.upload(fld, a)
        return s.diag_flow()


def test_diag_flow_noalloc_matches_diag_flow(rr):
    bufs = HipBuffers()
    try:
        # 640 x 96 x 80 = 4 915 200 cells (19 200 blocks), then 40 x 30 x 7 = 8400 cells
        for dims, seed in (((640, 96, 80), 1), ((40, 30, 7), 2)):
            n = dims[0] * dims[1] * dims[2]
            f = fields(n, seed)
            ref = diag_by_context(rr, dims, f)
            ptrs = [bufs.upload(a) for a in f]
            got = rr.diag_flow_noalloc(*ptrs, n)
            assert got == ref, (dims, got, ref)
            assert got["ncells"] == n
            bufs.free()
    finally:
        bufs.free()


def test_diag_flow_noalloc_limits(rr):
    bufs = HipBuffers()
    try:
        f = fields(64, 3)
        ptrs = [bufs.upload(a) for a in f]
        for n in (0, -5, (1 << 24) + 1):
            with pytest.raises(rr.RRLBMError) as ei:
                rr.diag_flow_noalloc(*ptrs, n)
            assert ei.value.status == 1
        with pytest.raises(rr.RRLBMError) as ei:
            rr.diag_flow_noalloc(0, ptrs[1], ptrs[2], ptrs[3], 64)
        assert ei.value.status == 1
        # a non-finite cell is reported, as rrlbm_diag_flow does
        bad = list(f)
        bad[1] = f[1].copy()
        bad[1][7] = np.nan
        ptrs = [bufs.upload(a) for a in bad]
        with pytest.raises(rr.RRLBMError) as ei:
            rr.diag_flow_noalloc(*ptrs, 64)
        assert ei.value.status == 4
    finally:
        bufs.free()
