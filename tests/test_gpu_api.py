"""GPU tests of the C ABI beyond the reference build: the HIP path against
the oracle (C restatement) on a larger mesh and another flow condition,
state upload/download, error reporting, and independent contexts."""
import os

import numpy as np
import pytest

import minifs_amd as mfs
from oracle import Oracle

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz")


def _mesh(tmp_path, n, seed=20260101):
    g = mfs.Grid.box_bump(n, seed=seed).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    return g, mesh, bc


@pytest.mark.parametrize("n,mach,alpha,cfl,iters", [(40, 0.8, 1.25, 10.0, 4), (12, 0.3, -3.0, 5.0, 25)])
def test_step_bitwise_vs_oracle(tmp_path, n, mach, alpha, cfl, iters):
    print(f"[api] oracle parity N={n} M={mach} a={alpha} cfl={cfl} iters={iters}", flush=True)
    g, mesh, bc = _mesh(tmp_path, n)
    o = Oracle(mesh, bc, mach, alpha, cfl)
    hist_o, _ = o.run(iters, min(16, os.cpu_count() or 1))
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g.preprocess(p)
    s = mfs.Solver(g, p)
    s.step(iters)
    for nm in FIELDS:
        a, b = s.field(nm), o.get(nm)
        bad = np.flatnonzero(a != b)
        assert bad.size == 0, f"{nm}: {bad.size} entries differ, first {bad[:4]}"
    h = s.history()
    np.testing.assert_array_equal(h[:, 5:], hist_o[:, 5:])
    np.testing.assert_allclose(h[:, :5], hist_o[:, :5], rtol=1e-12, atol=0)
    s.close()
    o.close()


def test_upload_download_round_trip_and_restart(tmp_path):
    """A state uploaded into a fresh context continues the solve exactly
    where the first context was (the solve has no hidden state besides the
    running dt extrema and the iteration counter)."""
    p = mfs.make_params(mach=0.6, alpha=2.0, cfl=10.0)
    g = mfs.Grid.box_bump(10).decouple().preprocess(p)
    a = mfs.Solver(g, p)
    a.step(6)
    snap = {nm: a.field(nm) for nm in ("rho", "u", "v", "w", "p")}
    a.step(1)
    b = mfs.Solver(g, p)
    for nm, v in snap.items():
        b.upload(nm, v)
        np.testing.assert_array_equal(b.field(nm), v)
    b.step(1)
    for nm in ("rho", "u", "v", "w", "p", "res", "dq"):
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    a.close()
    b.close()


def test_errors_are_reported():
    p = mfs.make_params(mach=0.5, alpha=0.0, cfl=10.0)
    g = mfs.Grid.box_bump(3).decouple().preprocess(p)
    s = mfs.Solver(g, p)
    n = g.info["nTCell"] + g.info["nBFace"]
    L = mfs.lib()
    buf = np.zeros(n)
    dp = buf.ctypes.data_as(mfs._lib._DP)
    assert L.mfs_download(s._h, 99, dp, n) < 0
    assert L.mfs_download(s._h, mfs._lib.F_RHO, dp, n - 1) < 0
    grad = np.zeros(5 * n)
    assert L.mfs_upload(s._h, mfs._lib.F_DQDX, grad.ctypes.data_as(mfs._lib._DP), 5 * n) < 0  # read-only
    assert L.mfs_step(s._h, -1) < 0
    assert "negative" in L.mfs_last_error().decode()
    # the context stays usable after argument errors
    s.step(2)
    assert s.history().shape == (2, 7)
    sc = s.scalars()
    assert sc["iter_done"] == 2
    assert sc["dt_max"] == s.history()[-1, 5] and sc["dt_min"] == s.history()[-1, 6]
    s.close()


def test_contexts_are_independent(tmp_path):
    p = mfs.make_params(mach=0.7, alpha=1.0, cfl=10.0)
    g = mfs.Grid.box_bump(8).decouple().preprocess(p)
    a, b = mfs.Solver(g, p), mfs.Solver(g, p)
    a.step(3)
    b.step(1)
    b.step(2)
    for nm in FIELDS:
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    np.testing.assert_array_equal(a.history(), b.history())
    a.close()
    b.close()


def test_threads_stepping_contexts_on_one_device():
    """Two host threads step a context each, and a third steps a batch, at
    the same time on one device.  Each sweep grid needs all of its waves
    resident, so the library orders the sweeps of different streams on a
    device one after the other (the other kernels may overlap).  Every
    result equals the same solve stepped alone."""
    import threading

    print("[api] three threads stepping on one device", flush=True)
    cases = [(mfs.make_params(mach=m, alpha=a, cfl=10.0), n) for m, a, n in ((0.7, 1.0, 14), (0.5, -2.0, 12), (0.8, 3.0, 10))]
    grids = [mfs.Grid.box_bump(n, seed=31 + n).decouple().preprocess(p) for p, n in cases]
    iters = 40
    ref = []
    for g, (p, _) in zip(grids, cases):
        s = mfs.Solver(g, p)
        s.step(iters)
        ref.append(({nm: s.field(nm) for nm in FIELDS}, s.history()))
        s.close()
    solvers = [mfs.Solver(grids[0], cases[0][0]), mfs.Solver(grids[1], cases[1][0])]
    batch = mfs.Batch(mfs.MFS_FP64)
    member = batch.add(grids[2], cases[2][0])
    errors = []

    def run(stepper):
        try:
            for _ in range(iters // 5):
                stepper.step(5)
            stepper.sync()
        except Exception as e:  # reported by the main thread
            errors.append(repr(e))

    threads = [threading.Thread(target=run, args=(x,)) for x in (solvers[0], solvers[1], batch)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for s, (fields, hist) in zip(solvers + [member], ref):
        for nm in FIELDS:
            np.testing.assert_array_equal(s.field(nm), fields[nm], err_msg=nm)
        np.testing.assert_array_equal(s.history(), hist)
    for s in solvers:
        s.close()
    batch.close()


@pytest.mark.parametrize("tile_cells,sweep_waves,by_tiles", [(64, 4, 0), (1000, 16, 1), (8192, 8, 0), (64, 8, 1)])
def test_layout_options_do_not_change_results(tile_cells, sweep_waves, by_tiles):
    """Tile size, sweep workgroup size and sweep schedule change the device
    order and the schedule, never the per-cell arithmetic: the fields are the
    same bits as with the defaults.  The residual norms are tree sums over
    blocks of cells in device order, so the history agrees to rounding."""
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(9).decouple().preprocess(p)
    a = mfs.Solver(g, p)
    b = mfs.Solver(g, p, tile_cells=tile_cells, sweep_waves=sweep_waves, sweep_by_tiles=by_tiles)
    info = b.info()
    assert info["tile_cells"] == tile_cells and info["sweep_waves"] == sweep_waves
    assert info["sweep_by_tiles"] == by_tiles
    a.step(6)
    b.step(6)
    for nm in FIELDS:
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    np.testing.assert_allclose(a.history(), b.history(), rtol=1e-12, atol=0)
    a.close()
    b.close()


@pytest.mark.parametrize("cells,precision,tail_grid,gradient_stream,tail_chunks",
                         [("tet", mfs.MFS_FP64, 0, 1, 694), ("tet", mfs.MFS_FP64, 1, 3, 694),
                          ("tet", mfs.MFS_FP64, 7, 5, 694), ("tet", mfs.MFS_MIXED, 0, 4, 694),
                          ("tet", mfs.MFS_MIXED, 100000, 5, 694), ("hex", mfs.MFS_FP64, 0, 5, 472),
                          ("prism", mfs.MFS_MIXED, 0, 5, 791)])
def test_sweep_tail_launch_does_not_change_results(cells, precision, tail_grid, gradient_stream, tail_chunks):
    """A sweep whose order ends in a run of thin levels runs them as a launch
    of their own (the forward sweep's last launch, the backward sweep's
    first).  The chunks, their order and their arithmetic are the same, so
    every field and history record is the same bits as with one launch per
    sweep, for tail grids of one workgroup, a few, the default and more than
    the tail has chunks (clamped), on 20^3-hex boxes of tetrahedra (4 lanes
    per cell), hexahedra and prisms (8 lanes per cell)."""
    print(f"[api] sweep tail grid {tail_grid} {cells} precision={precision} gradient_stream={gradient_stream}",
          flush=True)
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(20, cells=cells).decouple().preprocess(p)
    a = mfs.Solver(g, p, precision=precision, gradient_stream=gradient_stream, sweep_tail_grid=-1)
    b = mfs.Solver(g, p, precision=precision, gradient_stream=gradient_stream, sweep_tail_grid=tail_grid)
    ia, ib = a.info(), b.info()
    assert ia["sweep_tail_chunks"] == 0 and ia["sweep_tail_workgroups"] == 0
    assert ib["sweep_tail_chunks"] == tail_chunks
    want = {0: 64, 1: 1, 7: 7}.get(tail_grid, (tail_chunks + 3) // 4)
    assert ib["sweep_tail_workgroups"] == want
    for s in (a, b):
        s.step(5)
        # one iteration through the entry points: forward_lusgs takes the
        # same sweep launches as mfs_step
        s.zero_residuals()
        s.inviscid_flux()
        s.compute_timestep()
        s.forward_lusgs()
        s.calculate_gradient()
        s.iteration_done()
        s.step(3)
        s.sync()
    for nm in FIELDS:
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    np.testing.assert_array_equal(a.history(), b.history())
    a.close()
    b.close()
    g.close()


def test_kernel_timing_levels():
    """Timing level 1 brackets every kernel group with HIP events, level 2
    the two sweeps only; neither changes a result bit.  The 20^3 x 6 mesh's
    sweeps run as two launches each: each launch is timed, and a sweep's
    span holds both."""
    print("[api] kernel timing levels", flush=True)
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(20).decouple().preprocess(p)
    ref = mfs.Solver(g, p)
    ref.step(4)
    for level in (1, 2):
        s = mfs.Solver(g, p)
        s.enable_timing(level)
        s.step(4)
        s.sync()
        for sweep, first, second in (("fwd_sweep", "fwd_sweep_wide", "fwd_sweep_tail"),
                                     ("bwd_sweep", "bwd_sweep_tail", "bwd_sweep_wide")):
            span = s.timing(sweep)
            a, b = s.timing(first), s.timing(second)
            assert span[1] == a[1] == b[1] == 4 and a[0] > 0 and b[0] > 0, (level, sweep, span, a, b)
            assert span[0] >= a[0] + b[0], (level, sweep, span, a, b)
        for nm in ("flux_ts", "reduce", "node", "gradient"):
            ms, n = s.timing(nm)
            assert (n > 0) == (level == 1), (level, nm, n)
        for nm in FIELDS:
            np.testing.assert_array_equal(s.field(nm), ref.field(nm), err_msg=nm)
        np.testing.assert_array_equal(s.history(), ref.history())
        s.close()
    ref.close()
    g.close()


@pytest.mark.parametrize("cells,precision,ghosts,mode,n", [("tet", mfs.MFS_FP64, False, 2, 10),
                                                            ("tet", mfs.MFS_MIXED, False, 3, 10),
                                                            ("tet", mfs.MFS_FP64, True, 4, 10),
                                                            ("hex", mfs.MFS_FP64, False, 4, 10),
                                                            ("prism", mfs.MFS_MIXED, True, 2, 10),
                                                            ("tet", mfs.MFS_MIXED, False, 4, 10),
                                                            ("tet", mfs.MFS_FP64, False, 5, 20),
                                                            ("tet", mfs.MFS_MIXED, True, 5, 20),
                                                            ("tet", mfs.MFS_FP64, False, 5, 10)])
def test_gradient_on_second_stream_same_results(cells, precision, ghosts, mode, n):
    """mfs_step with the gradient on a second stream (started after the
    update, after the next residual pass, after the next forward sweep or
    after that sweep's wide levels) gives the same bits as with the gradient
    on the solver stream: every field and every history record, across
    several mfs_step calls and with kernel entry points in between.  The
    20^3 x 6 mesh has a thin tail of 694 chunks, so its sweeps run as two
    launches each; the 10^3 x 6 mesh has none (mode 5 then starts the
    gradient after the forward sweep)."""
    print(f"[api] gradient stream {mode} vs 1: {cells} N={n} precision={precision} ghosts={ghosts}", flush=True)
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(n, cells=cells).decouple().preprocess(p)
    a = mfs.Solver(g, p, precision=precision, gradient_stream=1)
    b = mfs.Solver(g, p, precision=precision, gradient_stream=mode)
    assert (b.info()["sweep_tail_chunks"] > 0) == (n == 20)
    for s in (a, b):
        s.ghost_refresh(ghosts)
        s.step(7)
        s.step(1)
        s.zero_residuals()
        s.inviscid_flux()
        s.compute_timestep()
        s.forward_lusgs()
        s.calculate_gradient()
        s.iteration_done()
        s.step(5)
        s.sync()
    for nm in FIELDS:
        np.testing.assert_array_equal(a.field(nm), b.field(nm), err_msg=nm)
    np.testing.assert_array_equal(a.history(), b.history())
    a.close()
    b.close()
    g.close()


@pytest.mark.parametrize("entry_points,setup", [(False, ""), (True, ""), (False, "tree+device")])
def test_native_driver_matches_python_api(entry_points, setup):
    """mfs_solve (the native driver of the reference's main loop) gives the
    same residual history record as the library driven from Python, through
    mfs_step or through the kernel entry points; also with another tree
    order and the set-up metrics on the GPU."""
    import json
    import os
    import subprocess

    exe = os.path.join(os.path.dirname(mfs.__file__), "mfs_solve")
    cmd = [exe, "--box", "8", "--steps", "4", "--warmup", "1", "--mach", "0.8", "--alpha", "1.25", "--cfl", "10",
           "--json"] + (["--entry-points"] if entry_points else []) + \
        (["--tree", "2,64", "--device-metrics"] if setup else [])
    print("[native]", " ".join(cmd), flush=True)
    out = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rec = json.loads(out.stdout.strip().splitlines()[-1])
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(8).decouple(*((2, 64) if setup else (4, 512))).preprocess(p)
    s = mfs.Solver(g, p)
    s.step(5)
    h = s.history()
    s.close()
    assert rec["cells"] == g.info["nTCell"]
    got = np.array(rec["last_residual"] + [rec["dt_max"], rec["dt_min"]])
    if entry_points:  # the separate flux kernel sums the norm in other blocks
        np.testing.assert_allclose(got, h[-1], rtol=1e-12, atol=0)
    else:
        np.testing.assert_array_equal(got, h[-1])


def test_set_ghost_variables_matches_oracle(tmp_path):
    """mfs_set_ghost_variables on a state the iterations produced: interior
    cells unchanged, ghosts as the oracle's SetGhostVariables sets them
    (to 1e-14 of each field's scale: the pow of the subsonic far field may
    differ in the last bit)."""
    mach, alpha, cfl = 0.5, 2.0, 10.0
    p = mfs.make_params(mach=mach, alpha=alpha, cfl=cfl)
    g = mfs.Grid.box_bump(6).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    g.preprocess(p)
    nC = g.info["nTCell"]
    s = mfs.Solver(g, p)
    s.step(4)
    before = {nm: s.field(nm) for nm in ("rho", "u", "v", "w", "p")}
    s.set_ghost_variables()
    after = {nm: s.field(nm) for nm in ("rho", "u", "v", "w", "p")}
    o = Oracle(mesh, bc, mach, alpha, cfl)
    for nm, v in before.items():
        o.set(nm, v)
    o.set_ghosts()
    changed = False
    for nm in before:
        np.testing.assert_array_equal(after[nm][:nC], before[nm][:nC], err_msg=nm)
        want = o.get(nm)
        scale = np.max(np.abs(want)) if nm != "p" else p.p_bar
        np.testing.assert_allclose(after[nm][nC:], want[nC:], rtol=0, atol=1e-14 * scale, err_msg=nm)
        changed |= not np.array_equal(after[nm][nC:], before[nm][nC:])
    assert changed  # the ghosts did follow the interior
    o.close()
    s.close()
    g.close()
