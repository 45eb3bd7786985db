"""GPU parity: the HIP path against the reference's own serial build.

The reference build (oracle/_ref/ref_driver) runs the reference's
InviscidFlux / ComputeTimeStep / ForwardLUSGS / CalculateGradient on the
same TDTree-renumbered mesh file; the HIP path must reproduce its state,
residual, time step, DQ and gradients bit for bit, and its per-iteration
residual history to 1e-12 relative (the norm is a parallel tree sum on the
GPU and a sequential sum in the reference).
"""
import os

import numpy as np
import pytest

import minifs_amd as mfs
from tests.refrun import STATE, load_state, run_reference

pytestmark = pytest.mark.gpu

HIST_RTOL = 1e-12


def _case(tmp_path, ref_driver, n, mach, alpha, cfl, iters, dumps, cells="tet", tree=(4, 512)):
    print(f"[parity] box N={n} {cells} mach={mach} alpha={alpha} cfl={cfl} iters={iters} tree={tree}", flush=True)
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g = mfs.Grid.box_bump(n, cells=cells).decouple(*tree)
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    out = str(tmp_path / "ref")
    ref = run_reference(ref_driver, mesh, bc, out, mach, alpha, cfl, iters, dumps=dumps)
    g.preprocess(p)
    return p, g, ref, out


def _assert_state_equal(s, out, it, nC):
    for j, nm in enumerate(STATE):
        a = s.field(nm)
        b = load_state(out, it, nm)
        assert a.shape == b.shape
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{nm} differs at it{it} in {bad.size} cells, first {bad[:5]}: {a[bad[:3]]} vs {b[bad[:3]]}"
    for nm in ("res", "dt", "dq", "dqdx", "dqdy", "dqdz"):
        a = s.field(nm)
        b = load_state(out, it, nm)
        assert a.shape == b.shape, nm
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{nm} differs at it{it} in {bad.size} entries, first {bad[:5]}: {a[bad[:3]]} vs {b[bad[:3]]}"


@pytest.mark.parametrize("n,iters", [(6, 30), (20, 100)])
def test_step_bitwise_vs_reference(tmp_path, ref_driver, n, iters):
    p, g, ref, out = _case(tmp_path, ref_driver, n, 0.5, 0.0, 10.0, iters, dumps=(1, iters))
    s = mfs.Solver(g, p)
    s.step(1)
    _assert_state_equal(s, out, 1, g.info["nTCell"])
    s.step(iters - 1)
    _assert_state_equal(s, out, iters, g.info["nTCell"])
    h = s.history()
    hr = ref["history"]
    assert h.shape == hr.shape
    np.testing.assert_array_equal(h[:, 5:], hr[:, 5:])  # running dt extrema are exact
    np.testing.assert_allclose(h[:, :5], hr[:, :5], rtol=HIST_RTOL, atol=0)
    s.close()


@pytest.mark.parametrize("tree", [(2, 128), (3, 64)])
def test_other_tree_orders_bitwise_vs_reference(tmp_path, ref_driver, tree):
    """Any cell order is a valid input: a mesh decoupled with other tree
    parameters (a shallower sweep DAG) still gives the reference's bits."""
    p, g, ref, out = _case(tmp_path, ref_driver, 12, 0.8, 1.25, 10.0, 25, dumps=(25,), tree=tree)
    s = mfs.Solver(g, p)
    s.step(25)
    _assert_state_equal(s, out, 25, g.info["nTCell"])
    h = s.history()
    np.testing.assert_array_equal(h[:, 5:], ref["history"][:, 5:])
    np.testing.assert_allclose(h[:, :5], ref["history"][:, :5], rtol=HIST_RTOL, atol=0)
    s.close()


@pytest.mark.parametrize("bc_text", ["2\n1 3\n2 4\n", "2\n1 3\n2 61\n", "2\n1 4\n2 6\n", "2\n1 4\n2 10\n"],
                         ids=["symmetry", "far-field-mod", "symmetric-bump", "interface"])
def test_other_boundary_types_bitwise_vs_reference(tmp_path, ref_driver, bc_text):
    """The box's patches remapped in bc.para to SYMM (4), FAR_FIELD_MOD (61)
    and INTERFACE (10): the symmetry-node values, the other-type node
    weights, the node-mean face values and the ghost rules those types
    take, against the reference build."""
    print(f"[parity] bc {bc_text!r}", flush=True)
    mach, alpha, cfl, iters = 0.6, 2.0, 10.0, 20
    g0 = mfs.Grid.box_bump(8, seed=5).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g0.save(mesh, bc)
    g0.close()
    open(bc, "w").write(bc_text)
    out = str(tmp_path / "ref")
    ref = run_reference(ref_driver, mesh, bc, out, mach, alpha, cfl, iters, dumps=(1, iters))
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g = mfs.Grid.load(mesh, bc).preprocess(p)
    s = mfs.Solver(g, p)
    s.step(1)
    _assert_state_equal(s, out, 1, g.info["nTCell"])
    s.step(iters - 1)
    _assert_state_equal(s, out, iters, g.info["nTCell"])
    h = s.history()
    np.testing.assert_array_equal(h[:, 5:], ref["history"][:, 5:])
    np.testing.assert_allclose(h[:, :5], ref["history"][:, :5], rtol=HIST_RTOL, atol=0)
    s.close()


def test_cfl_ramp_bitwise_vs_reference(tmp_path, ref_driver):
    """The CFL ramp of LimitTimeStep (compute_timestep.cpp:868-871): cfl
    from cfl_start to cfl_end over cfl_nstep iterations, then cfl_end; the
    run crosses the end of the ramp."""
    print("[parity] cfl ramp 1 -> 10 over 12 iterations", flush=True)
    mach, alpha, iters = 0.7, 1.0, 18
    g = mfs.Grid.box_bump(7, seed=21).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    out = str(tmp_path / "ref")
    ref = run_reference(ref_driver, mesh, bc, out, mach, alpha, 10.0, iters, dumps=(5, iters), ramp=(1.0, 12))
    p = mfs.make_params(mach=mach, alpha=alpha, cfl_start=1.0, cfl_end=10.0, cfl_nstep=12)
    g.preprocess(p)
    s = mfs.Solver(g, p)
    s.step(5)
    _assert_state_equal(s, out, 5, g.info["nTCell"])
    s.step(iters - 5)
    _assert_state_equal(s, out, iters, g.info["nTCell"])
    h = s.history()
    np.testing.assert_array_equal(h[:, 5:], ref["history"][:, 5:])
    np.testing.assert_allclose(h[:, :5], ref["history"][:, :5], rtol=HIST_RTOL, atol=0)
    s.close()


@pytest.mark.parametrize("dims", [(1, 1, 1), (1, 2, 7), (9, 2, 3)])
def test_tiny_and_ragged_boxes_bitwise_vs_reference(tmp_path, ref_driver, dims):
    """Meshes of 6, 84 and 324 cells: a single chunk, chunks shorter than a
    wave, one-cell-thick boxes; the whole path against the reference build."""
    print(f"[parity] box {dims}", flush=True)
    mach, alpha, cfl, iters = 0.5, 2.0, 10.0, 12
    g = mfs.Grid.box_bump(*dims, seed=8).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    out = str(tmp_path / "ref")
    ref = run_reference(ref_driver, mesh, bc, out, mach, alpha, cfl, iters, dumps=(iters,))
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g.preprocess(p)
    for prec in (mfs.MFS_FP64, mfs.MFS_MIXED):
        s = mfs.Solver(g, p, precision=prec)
        s.step(iters)
        if prec == mfs.MFS_FP64:
            _assert_state_equal(s, out, iters, g.info["nTCell"])
            np.testing.assert_allclose(s.history()[:, :5], ref["history"][:, :5], rtol=HIST_RTOL, atol=0)
        else:  # the mixed path runs and stays finite on the same meshes
            assert np.isfinite(s.history()).all()
        s.close()


@pytest.mark.parametrize("cells", ["hex", "prism"])
def test_polyhedral_cells_bitwise_vs_reference(tmp_path, ref_driver, cells):
    """Hexahedra (6 quadrilateral faces) and prisms (5 faces, mixed
    triangles and quadrilaterals) run the kernels' 6-face variants and the
    8-lanes-per-cell sweep."""
    p, g, ref, out = _case(tmp_path, ref_driver, 7, 0.8, 1.25, 10.0, 20, dumps=(20,), cells=cells)
    s = mfs.Solver(g, p)
    s.step(20)
    _assert_state_equal(s, out, 20, g.info["nTCell"])
    h = s.history()
    np.testing.assert_array_equal(h[:, 5:], ref["history"][:, 5:])
    np.testing.assert_allclose(h[:, :5], ref["history"][:, :5], rtol=HIST_RTOL, atol=0)
    s.close()


def test_entry_points_match_step(tmp_path, ref_driver):
    """The per-kernel entry points (reference call order) and the fused
    mfs_step produce the same bits in every field.  The residual norms are
    tree sums over different workgroup partitions (the fused residual kernel
    groups cells by mesh id), so they agree to rounding; the dt extrema are
    exact."""
    p, g, ref, out = _case(tmp_path, ref_driver, 8, 0.8, 1.25, 10.0, 5, dumps=(5,))
    a = mfs.Solver(g, p)
    b = mfs.Solver(g, p)
    a.iterate_reference_order(5)
    b.step(5)
    for nm in STATE + ("res", "dt", "dq", "dqdx", "dqdy", "dqdz"):
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    _assert_state_equal(a, out, 5, g.info["nTCell"])
    ha, hb = a.history(), b.history()
    np.testing.assert_array_equal(ha[:, 5:], hb[:, 5:])
    np.testing.assert_allclose(ha[:, :5], hb[:, :5], rtol=HIST_RTOL, atol=0)
    a.close()
    b.close()


def test_ghost_refresh_vs_reference(tmp_path, ref_driver):
    """mfs_step with SetGhostVariables after every update (the call the
    reference's main.cpp:283 leaves commented out) against the reference
    build running its own SetGhostVariables each iteration.  The far-field
    ghosts of a subsonic flow go through pow, whose GPU result may differ
    from the host's in the last bit, so the state is compared to 1e-12 of
    each field's scale and the history to 1e-10 (the plain path is bitwise)."""
    n, iters, mach, alpha, cfl = 8, 20, 0.5, 2.0, 10.0
    print(f"[parity] ghost refresh box N={n} iters={iters}", flush=True)
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g = mfs.Grid.box_bump(n).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    out = str(tmp_path / "ref")
    ref = run_reference(ref_driver, mesh, bc, out, mach, alpha, cfl, iters, dumps=(iters,), ghost_refresh=True)
    g.preprocess(p)
    s = mfs.Solver(g, p)
    s.ghost_refresh(True)
    s.step(iters)
    exact = []
    for nm in STATE:
        a, b = s.field(nm), load_state(out, iters, nm)
        scale = np.max(np.abs(b)) if nm != "p" else p.p_bar
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * scale, err_msg=nm)
        exact.append(np.mean(a == b))
    print(f"[parity] ghost refresh: fraction of bitwise-equal state entries {min(exact):.6f}", flush=True)
    np.testing.assert_allclose(s.history(), ref["history"], rtol=1e-10, atol=0)
    s.close()
    g.close()


@pytest.mark.parametrize("ghost_refresh", [False, True], ids=["plain", "ghost-refresh"])
def test_guard_branches_vs_reference(tmp_path, ref_driver, ghost_refresh):
    """The robustness branches (all three DQ limiter stages in both sweeps,
    CFL3D relaxation, the update's rho/p clamps, LimitTimeStep below p_break
    and p_min, CellIsMG, the ratio_dtmax clamp, and with the ghost refresh
    the ghost clamps) on a perturbed state with tightened limits
    (tests/guards.py), uploaded with mfs_upload, against the reference build
    started from the same state.  Plain: bitwise after each of three
    iterations.  With the ghost refresh the far-field ghosts go through pow,
    whose GPU result may differ in the last bit, so the state is compared to
    1e-12 of each field's scale.  The oracle is pinned on the same workload
    (tests/test_oracle.py); here each branch's fingerprint is counted in the
    GPU's own output."""
    from tests import guards

    print(f"[parity] guard branches ghost_refresh={ghost_refresh}", flush=True)
    mesh, bc, info, st, init = guards.make_case(tmp_path)
    nC = info["nTCell"]
    out = str(tmp_path / "ref")
    hist_r = guards.run_reference(ref_driver, mesh, bc, init, out, ghost_refresh)
    p = guards.params()
    g = mfs.Grid.load(mesh, bc).preprocess(p)
    s = mfs.Solver(g, p)
    for k, a in st.items():
        s.upload(k, a)
    s.ghost_refresh(ghost_refresh)
    gpu_dir = tmp_path / "gpu"
    os.makedirs(gpu_dir)
    for it in range(1, guards.ITERS + 1):
        s.step(1)
        if not ghost_refresh:
            _assert_state_equal(s, out, it, nC)
        else:
            for nm in STATE:
                a, b = s.field(nm), load_state(out, it, nm)
                scale = np.max(np.abs(b)) if nm != "p" else p.p_bar
                np.testing.assert_allclose(a, b, rtol=0, atol=1e-12 * scale, err_msg=f"it{it} {nm}")
        for nm in ("rho", "p", "dt", "dq", "u", "v", "w"):
            s.field(nm).tofile(str(gpu_dir / f"it{it}_{nm}.f64"))
    h = s.history()
    if not ghost_refresh:
        np.testing.assert_array_equal(h[:, 5:], hist_r[:, 5:])
        np.testing.assert_allclose(h[:, :5], hist_r[:, :5], rtol=HIST_RTOL, atol=0)
    else:
        np.testing.assert_allclose(h, hist_r, rtol=1e-10, atol=0)
    ev = guards.branch_predicates(str(gpu_dir), st, h, nC, ghost_refresh)
    print(f"[parity] guard fingerprints in the GPU output: {ev}", flush=True)
    assert all(v > 0 for v in ev.values()), ev
    s.close()
    if not ghost_refresh:  # the FP32 operator takes the same branches and stays finite
        m = mfs.Solver(g, p, precision=mfs.MFS_MIXED)
        for k, a in st.items():
            m.upload(k, a)
        m.step(guards.ITERS)
        rho, pr = m.field("rho"), m.field("p")
        assert np.isfinite(m.history()).all() and np.isfinite(rho).all() and np.isfinite(pr).all()
        for a, lo, hi in ((rho, p.rho_min, p.rho_max), (pr, p.p_min, p.p_max)):
            assert np.all((a[:nC] >= lo) & (a[:nC] <= hi))
            assert np.any(a[:nC] == lo) and np.any(a[:nC] == hi)
        m.close()
    g.close()
