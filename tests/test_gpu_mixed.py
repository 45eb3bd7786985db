"""GPU tests of mixed-precision contexts (MFS_MIXED): FP32 implicit operator
(state snapshot, geometry, Diag, DQ, FluxLUSGS3D, limiter) and FP32
gradients, FP64 residual, residual norm and state update.  The iteration is
a defect correction, so it converges to the FP64 solution: its first
residual equals the FP64 one exactly, and the converged states agree to the
convergence level."""
import numpy as np
import pytest

import minifs_amd as mfs

pytestmark = pytest.mark.gpu


def _run(g, p, precision, drop, max_iters, every=50):
    s = mfs.Solver(g, p, precision=precision)
    while True:
        s.step(every)
        h = s.history()
        if h[-1, 0] <= drop * h[0, 0] or len(h) >= max_iters:
            break
    st = {nm: s.field(nm) for nm in ("rho", "u", "v", "w", "p")}
    s.close()
    return h, st


def test_mixed_first_residual_is_the_fp64_residual():
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(10).decouple().preprocess(p)
    a, b = mfs.Solver(g, p), mfs.Solver(g, p, precision=mfs.MFS_MIXED)
    a.step(1)
    b.step(1)
    np.testing.assert_array_equal(a.field("res"), b.field("res"))
    np.testing.assert_array_equal(a.history(), b.history())
    # one FP32 operator application: the increments agree to FP32 accuracy
    da, db = a.field("dq"), b.field("dq")
    assert np.linalg.norm(da - db) <= 1e-5 * np.linalg.norm(da)
    a.close()
    b.close()


@pytest.mark.parametrize("mach,alpha", [(0.5, 0.0), (0.8, 1.25)])
def test_mixed_converges_to_the_fp64_solution(mach, alpha):
    print(f"[mixed] convergence M={mach} a={alpha}", flush=True)
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=10.0)
    g = mfs.Grid.box_bump(8).decouple().preprocess(p)
    h64, s64 = _run(g, p, mfs.MFS_FP64, 1e-10, 20000)
    h32, s32 = _run(g, p, mfs.MFS_MIXED, 1e-10, 20000)
    assert h64[-1, 0] <= 1e-10 * h64[0, 0]
    assert h32[-1, 0] <= 1e-10 * h32[0, 0]
    # same fixed point: state error at the level the convergence leaves
    for nm in ("rho", "u", "v", "w"):
        a, b = s64[nm], s32[nm]
        assert np.max(np.abs(a - b)) <= 1e-8 * np.max(np.abs(a)), nm
    assert np.max(np.abs(s64["p"] - s32["p"])) <= 1e-8 * p.p_bar
    F64, L64 = mfs.wall_force(g, s64["p"], p)
    F32, L32 = mfs.wall_force(g, s32["p"], p)
    assert abs(L32 - L64) <= 1e-9 * abs(L64)
    # the FP32 operator does not slow the convergence much
    assert len(h32) <= 1.5 * len(h64) + 100


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
@pytest.mark.parametrize("waves", [4, 8, 16])
def test_full_sweep_grid_is_resident(precision, waves):
    """A mesh with more sweep chunks than the sweep grid has waves, so every
    workgroup of the sweep grid is launched and the sweeps need them all
    resident at once: at most 16 waves per CU (a measured bound, not an
    occupancy query), for every workgroup size and both operator
    precisions."""
    print(f"[mixed] full sweep grid precision={precision} waves={waves}", flush=True)
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(25).decouple().preprocess(p)
    s = mfs.Solver(g, p, precision=precision, sweep_waves=waves)
    info = s.info()
    assert info["compute_units"] > 0
    assert info["sweep_workgroups"] * waves <= 16 * info["compute_units"]
    assert g.info["nTCell"] // 16 > info["sweep_workgroups"] * waves  # more chunks than waves
    s.step(3)
    h = s.history()
    assert np.all(np.isfinite(h)) and h[-1, 0] < h[0, 0] * 10
    s.close()
    g.close()
