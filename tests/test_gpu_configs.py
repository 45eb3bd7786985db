"""The benchmark workloads at full size on the GPU.

- 6.0M-cell box+bump mesh, Mach 0.5, 500 iterations, plain FP64: history
  and final arrays against the reference's serial build (committed fixture
  made by tests/golden/make_config2.py).
- The same mesh at Mach 0.8, AoA 1.25: both precisions to a 1e-10 residual
  drop; the mixed-precision state and wall force against the FP64 ones.
- 64 seeded solves of 15-60 hexes per axis (x6 tetrahedra) stepped together
  in one batch, each against the same solve alone.
- The first 16 requests of the seeded Poisson request stream (20k-3M cells)
  through the serving loop to a 1e-8 drop, each against the same request
  solved alone.
"""
import concurrent.futures as cf
import hashlib
import json
import os
import time

import numpy as np
import pytest

import minifs_amd as mfs
from minifs_amd.serve import Server, make_batch_requests, make_requests, run_sequential

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz")
STATE = ("rho", "u", "v", "w", "p")


def _sha_file(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for blk in iter(lambda: f.read(1 << 24), b""):
            h.update(blk)
    return h.hexdigest()


def _digest(s, names=FIELDS):
    return {nm: hashlib.sha256(s.field(nm).tobytes()).hexdigest() for nm in names}


def test_500_iterations_6M_cells_vs_reference_fixture(tmp_path):
    """100^3 x 6 = 6.0M cells, Mach 0.5, AoA 0, CFL 10, 500 iterations of
    mfs_step (device set-up metrics): the per-iteration history against the
    reference build's (norms to 1e-12 relative, running dt extrema exact) and
    all 11 final arrays bitwise (SHA-256 of each)."""
    man = json.load(open(os.path.join(GOLD, "box100.json")))
    hist_r = np.fromfile(os.path.join(GOLD, "box100_history.f64")).reshape(-1, 7)
    print("[config] 6.0M cells, 500 iterations against the reference fixture", flush=True)
    g = mfs.Grid.box_bump(100, seed=20260101).decouple(4, 512)
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    assert _sha_file(mesh) == man["mesh_sha256"], "the generator no longer writes the mesh the reference ran on"
    assert open(bc).read() == man["bc_text"]
    os.remove(mesh)
    p = mfs.make_params(mach=man["mach"], alpha=man["alpha"], cfl=man["cfl"])
    g.preprocess(p, device=True)
    s = mfs.Solver(g, p)
    t0 = time.perf_counter()
    s.step(man["iters"])
    s.sync()
    print(f"[config] 500 iterations in {time.perf_counter() - t0:.2f} s", flush=True)
    h = s.history()
    assert h.shape == hist_r.shape
    np.testing.assert_array_equal(h[:, 5:], hist_r[:, 5:])
    np.testing.assert_allclose(h[:, :5], hist_r[:, :5], rtol=1e-12, atol=0)
    got = _digest(s)
    bad = [nm for nm in FIELDS if got[nm] != man["final_sha256"][nm]]
    assert not bad, f"final arrays differ from the reference's: {bad}"
    s.close()
    g.close()


def _converge(g, p, precision, drop, max_iters, every=100):
    s = mfs.Solver(g, p, precision=precision)
    t0 = time.perf_counter()
    while True:
        s.step(every)
        h = s.history()
        if h[-1, 0] <= drop * h[0, 0] or len(h) >= max_iters:
            break
    el = time.perf_counter() - t0
    st = {nm: s.field(nm) for nm in STATE}
    s.close()
    return h, st, el


def test_mixed_precision_converges_to_fp64_on_6M_cells():
    """100^3 x 6 = 6.0M cells, Mach 0.8, AoA 1.25, CFL 10: FP64 and mixed
    precision each to a 1e-10 drop of the density residual.  The converged
    mixed state agrees with the FP64 one to 1e-8 (L2 and Linf, relative to
    each field's scale; p to p_bar), and so does the wall-force lift."""
    print("[config] 6.0M cells, Mach 0.8: FP64 and mixed to a 1e-10 drop", flush=True)
    p = mfs.make_params(mach=0.8, alpha=1.25, cfl=10.0)
    g = mfs.Grid.box_bump(100, seed=20260101).decouple(4, 512).preprocess(p, device=True)
    h64, s64, t64 = _converge(g, p, mfs.MFS_FP64, 1e-10, 20000)
    print(f"[config] FP64: {len(h64)} iterations, {t64:.1f} s", flush=True)
    h32, s32, t32 = _converge(g, p, mfs.MFS_MIXED, 1e-10, 20000)
    print(f"[config] mixed: {len(h32)} iterations, {t32:.1f} s ({t64 / t32:.3f}x)", flush=True)
    assert h64[-1, 0] <= 1e-10 * h64[0, 0]
    assert h32[-1, 0] <= 1e-10 * h32[0, 0]
    err = {}
    for nm in STATE:
        a, b = s64[nm], s32[nm]
        scale = p.p_bar if nm == "p" else np.max(np.abs(a))
        err[nm] = (np.linalg.norm(a - b) / np.linalg.norm(a) if nm != "p" else
                   np.linalg.norm(a - b) / (np.sqrt(a.size) * scale), np.max(np.abs(a - b)) / scale)
    print(f"[config] mixed vs FP64 state error (L2, Linf): {err}", flush=True)
    assert all(l2 <= 1e-8 and linf <= 1e-8 for l2, linf in err.values()), err
    _, L64 = mfs.wall_force(g, s64["p"], p)
    _, L32 = mfs.wall_force(g, s32["p"], p)
    print(f"[config] lift FP64 {L64!r} mixed {L32!r}", flush=True)
    assert abs(L32 - L64) <= 1e-8 * abs(L64)
    g.close()


def test_64_batched_solves_equal_alone():
    """64 seeded solves (seed 4001): boxes of U{15..60}^3 hexes x 6
    tetrahedra (about 20k to 1.3M cells), bump height U[0.02, 0.08], Mach
    U[0.3, 0.85], AoA U[-3, 5], mixed precision, stepped together for 50
    iterations in one batch.  Every member's 11 arrays and history equal the
    same solve's run alone."""
    print("[config] 64 batched solves, 50 iterations, mixed precision", flush=True)
    reqs = make_batch_requests(64, 15, 60, 4001)
    iters = 50

    def build(r):
        p = mfs.make_params(mach=r.mach, alpha=r.alpha, cfl=10.0)
        g = mfs.Grid.box_bump(r.nx, r.ny, r.nz, bump_height=r.bump_height, seed=r.seed).decouple()
        return g.preprocess(p, device=True), p

    b = mfs.Batch(mfs.MFS_MIXED)
    alone, members, cells = [], [], 0
    with cf.ThreadPoolExecutor(16) as pool:
        for k, (g, p) in enumerate(pool.map(build, reqs)):
            s = mfs.Solver(g, p, precision=mfs.MFS_MIXED)
            s.step(iters)
            alone.append((_digest(s), s.history()))
            s.close()
            members.append(b.add(g, p))
            cells += g.info["nTCell"]
            g.close()
            if k % 16 == 15:
                print(f"[config] {k + 1} solves run alone and added", flush=True)
    t0 = time.perf_counter()
    b.step(iters)
    b.sync()
    el = time.perf_counter() - t0
    print(f"[config] batch of {len(b)} solves, {cells} cells: {iters} iterations in {el:.2f} s "
          f"({cells * iters / el / 1e6:.0f} M cell-iters/s)", flush=True)
    for k, (m, (dig, h)) in enumerate(zip(members, alone)):
        np.testing.assert_array_equal(m.history(), h, err_msg=f"solve {k}")
        got = _digest(m)
        bad = [nm for nm in FIELDS if got[nm] != dig[nm]]
        assert not bad, f"solve {k}: {bad} differ between the batch and the solve alone"
    b.close()


def test_served_requests_equal_alone():
    """The first 16 requests of the seeded request stream (seed 5001, Poisson
    arrivals at 4 per second, 20k-3M cells log-uniform, Mach U[0.3, 0.85],
    AoA U[-3, 5]) served by the continuous-batching loop in mixed precision
    to a 1e-8 drop, under a 200 GB device budget; every result (iterations
    and the SHA-256 of rho, u, v, w, p) equals the same request solved alone."""
    print("[config] 16 served requests, mixed precision, to a 1e-8 drop", flush=True)
    served = make_requests(16, 4.0, 20000, 3e6, 5001)
    srv = Server(mfs.MFS_MIXED, hbm_budget=200e9, drop=1e-8, max_iters=20000, check_every=20, workers=16)
    stats = srv.run(served)
    print(f"[config] served: {stats}", flush=True)
    assert stats["solves"] == 16 and stats["converged"] == 16
    assert stats["peak_concurrent"] >= 2
    alone = make_requests(16, 4.0, 20000, 3e6, 5001)
    seq = run_sequential(alone, drop=1e-8, max_iters=20000, check_every=20, precision=mfs.MFS_MIXED)
    print(f"[config] alone: {seq}", flush=True)
    for a, b in zip(sorted(served, key=lambda r: r.rid), sorted(alone, key=lambda r: r.rid)):
        assert a.iters == b.iters, (a.rid, a.iters, b.iters)
        assert a.state_sha256 == b.state_sha256, (a.rid, a.n_cells)
