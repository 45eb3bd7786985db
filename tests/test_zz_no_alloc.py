"""mfs_step_noalloc, the variant of mfs_step that allocates nothing, against
mfs_step itself, through the C ABI.

The variant keeps its reduction partials and sweep ticket counters in the
library's static device array and writes the history records into a device
buffer the caller owns.  Its results must be mfs_step's bit for bit: all 11
arrays and every history record.  Two calls in a row, the second on a
smaller mesh, so the second starts from what the first left in the array.
"""
import ctypes as C

import numpy as np
import pytest

import minifs_amd as mfs
from minifs_amd._lib import lib

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz")
HIP_H2D, HIP_D2H = 1, 2


class DeviceBuffer:
    """A device buffer of doubles from the HIP runtime (the caller's memory)."""

    def __init__(self, count: int):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.count = count
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), 8 * count) == 0
        assert self.hip.hipMemset(self.ptr, 0xFF, 8 * count) == 0  # NaN until written

    def read(self) -> np.ndarray:
        out = np.empty(self.count)
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, 8 * self.count, HIP_D2H) == 0
        return out

    def free(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _step_noalloc(s: mfs.Solver, iters: int) -> np.ndarray:
    buf = DeviceBuffer(7 * iters)
    rc = lib().mfs_step_noalloc(s._h, iters, buf.ptr)
    assert rc == 0, lib().mfs_last_error()
    s.sync()
    h = buf.read().reshape(iters, 7)
    buf.free()
    return h


def _same_history(h, hr):
    assert h.shape == hr.shape
    np.testing.assert_array_equal(h, hr)


def _compare(a: mfs.Solver, b: mfs.Solver, what: str):
    for nm in FIELDS:
        x, y = a.field(nm), b.field(nm)
        bad = np.flatnonzero(x != y)
        assert bad.size == 0, f"{what}: {nm} differs in {bad.size} entries, first {bad[:4]}"


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED], ids=["fp64", "mixed"])
def test_noalloc_equals_step_two_calls(precision):
    """First call: 16^3 x 6 = 24,576 cells, 6 iterations; second call: a
    6^3 x 6 = 1,296-cell mesh, 4 iterations, after the first left its
    partials and counters in the static array."""
    for n, mach, alpha, iters in ((16, 0.8, 1.25, 6), (6, 0.5, 0.0, 4)):
        print(f"[noalloc] precision={precision} N={n} iters={iters}", flush=True)
        p = mfs.make_params(mach=mach, alpha=alpha, cfl=10.0)
        g = mfs.Grid.box_bump(n, seed=11 + n).decouple().preprocess(p)
        ref = mfs.Solver(g, p, precision=precision)
        ref.step(iters)
        ref.sync()
        s = mfs.Solver(g, p, precision=precision)
        h = _step_noalloc(s, iters)
        _same_history(h, ref.history())
        _compare(ref, s, f"N={n}")
        assert s.history().shape == (0, 7)  # the records went to the caller's buffer only
        ref.close()
        s.close()
        g.close()


@pytest.mark.parametrize("opts", [{}, {"gradient_stream": 1}, {"gradient_stream": 4}], ids=["default", "fused", "side"])
def test_noalloc_continues_a_solve(opts):
    """mfs_step and mfs_step_noalloc interleaved on one context give the
    same solve as mfs_step alone (the running dt extrema, the CFL ramp's
    iteration count and the state carry over), with the gradient fused with
    the next residual pass or on the context's second stream."""
    print(f"[noalloc] interleaved with mfs_step {opts}", flush=True)
    p = mfs.make_params(mach=0.7, alpha=2.0, cfl=10.0)
    g = mfs.Grid.box_bump(10, seed=5).decouple().preprocess(p)
    ref = mfs.Solver(g, p, **opts)
    ref.step(9)
    ref.sync()
    s = mfs.Solver(g, p, **opts)
    s.step(3)
    h_mid = _step_noalloc(s, 4)
    s.step(2)
    s.sync()
    h = s.history()
    hr = ref.history()
    _same_history(h[:3], hr[:3])
    _same_history(h_mid, hr[3:7])
    _same_history(h[3:], hr[7:])
    _compare(ref, s, "interleaved")
    ref.close()
    s.close()
    g.close()


def test_noalloc_rejects_bad_calls():
    p = mfs.make_params(mach=0.5, alpha=0.0, cfl=10.0)
    g = mfs.Grid.box_bump(4).decouple().preprocess(p)
    s = mfs.Solver(g, p)
    assert lib().mfs_step_noalloc(s._h, 2, None) == -1  # MFS_ERR_ARG: no history buffer
    assert lib().mfs_step_noalloc(s._h, -1, None) == -1
    assert lib().mfs_step_noalloc(s._h, 0, None) == 0
    s.close()
    b = mfs.Batch(mfs.MFS_FP64)
    m = b.add(g, p)
    buf = DeviceBuffer(7)
    assert lib().mfs_step_noalloc(m._h, 1, buf.ptr) == -5  # MFS_ERR_STATE: batch member
    assert b"batch" in lib().mfs_last_error()
    buf.free()
    b.close()
    g.close()
