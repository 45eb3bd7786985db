"""Serving loop for steady solves of uneven mesh size and flow condition.

Requests arrive over time.  Each one is a box+bump mesh (size, bump height,
seed, cell kind) and a flow condition, solved to a residual drop.  The
server overlaps three things:

* set-up (mesh generation and TDTree ordering on the host, ComputeMetrics on
  the GPU; the C++ code releases the GIL through ctypes) on a pool of worker
  threads;
* admission of set-up solves while the device memory they need fits the
  budget: their device contexts (host layout, upload) are built on worker
  threads and then attached to the one device batch (mfs_batch_attach);
* batched device iterations of every admitted solve (one launch per kernel
  for all of them), with convergence checked every `check_every` steps;
  converged solves leave the batch and free their memory for the next.

Each solve's arithmetic is the single-solve path's, so a result is the same
alone and in the mix.  `run_sequential` is the baseline: the same requests
one at a time, each set up and solved alone on the plain FP64 path.

Both build every context with the layout options the autotuner measured for
the nearest mesh size (`tune.tuned_options`, tile size and sweep workgroup
size); the options change the device order and the schedule, never a result.
"""
from __future__ import annotations

import concurrent.futures as cf
import dataclasses
import hashlib
import math
import time

import numpy as np

from . import MFS_FP64, Batch, Grid, Solver, make_params
from .tune import tuned_options


def tuned_solver(g: Grid, p, precision: int) -> Solver:
    """A context with the tuned layout options for this mesh size."""
    return Solver(g, p, precision=precision, **tuned_options(g.info["nTCell"], precision))


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the run
    nx: int
    ny: int
    nz: int
    mach: float
    alpha: float
    bump_height: float = 0.05
    seed: int = 20260101
    cells: str = "tet"
    # filled by the server
    t_ready: float = math.nan     # host set-up finished
    t_admit: float = math.nan     # joined the device batch
    t_done: float = math.nan      # converged (or gave up)
    iters: int = 0
    drop: float = math.nan        # final rho-residual norm / first
    converged: bool = False
    n_cells: int = 0
    rho_sum: float = math.nan     # checksum of the final density
    state_sha256: str = ""        # SHA-256 of the final rho, u, v, w, p (reference layout, ghosts included)

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def make_requests(n: int, rate: float, cells_lo: float, cells_hi: float, seed: int) -> list[Request]:
    """Poisson arrivals at `rate` per second; cell counts log-uniform in
    [cells_lo, cells_hi] as 6 n^3 tetrahedra of a near-cubic box; Mach ~
    U[0.3, 0.85], AoA ~ U[-3, 5] degrees, bump height ~ U[0.02, 0.08]."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        cells = math.exp(rng.uniform(math.log(cells_lo), math.log(cells_hi)))
        side = (cells / 6.0) ** (1.0 / 3.0)
        nx = max(2, int(round(side * rng.uniform(0.85, 1.15))))
        ny = max(2, int(round(side * rng.uniform(0.85, 1.15))))
        nz = max(2, int(round(cells / 6.0 / (nx * ny))))
        out.append(Request(i, t, nx, ny, nz, float(rng.uniform(0.3, 0.85)), float(rng.uniform(-3.0, 5.0)),
                           float(rng.uniform(0.02, 0.08)), int(rng.integers(1, 2**31))))
    return out


def make_batch_requests(n: int, lo: int, hi: int, seed: int) -> list[Request]:
    """n solves present at t = 0: boxes of nx, ny, nz ~ U{lo..hi} hexes x 6
    tetrahedra, bump height ~ U[0.02, 0.08], Mach ~ U[0.3, 0.85], AoA ~
    U[-3, 5] degrees, mesh seed drawn from the same generator."""
    r = np.random.default_rng(seed)
    out = []
    for i in range(n):
        nx, ny, nz = (int(r.integers(lo, hi + 1)) for _ in range(3))
        bh, mach, alpha = float(r.uniform(0.02, 0.08)), float(r.uniform(0.3, 0.85)), float(r.uniform(-3.0, 5.0))
        out.append(Request(i, 0.0, nx, ny, nz, mach, alpha, bh, int(r.integers(1, 2**31))))
    return out


def state_digest(solver) -> str:
    """SHA-256 of a solve's state arrays rho, u, v, w, p (reference layout)."""
    h = hashlib.sha256()
    for nm in ("rho", "u", "v", "w", "p"):
        h.update(solver.field(nm).tobytes())
    return h.hexdigest()


def _finish(r: Request, solver, h: np.ndarray, drop: float, t: float) -> None:
    r.t_done = t
    r.iters = len(h)
    r.drop = float(h[-1, 0] / h[0, 0])
    r.converged = _converged(h, drop)
    r.rho_sum = float(np.sum(solver.field("rho")))
    r.state_sha256 = state_digest(solver)


def _setup(r: Request, cfl: float, tree: tuple[int, int] = (4, 512), device_metrics: bool = True):
    p = make_params(mach=r.mach, alpha=r.alpha, cfl=cfl)
    g = Grid.box_bump(r.nx, r.ny, r.nz, bump_height=r.bump_height, seed=r.seed, cells=r.cells)
    g.decouple(*tree).preprocess(p, device=device_metrics)
    return g, p


def _converged(h: np.ndarray, drop: float) -> bool:
    return len(h) > 0 and h[-1, 0] <= drop * h[0, 0]


class Server:
    def __init__(self, precision: int, hbm_budget: float, drop: float = 1e-8, max_iters: int = 20000,
                 check_every: int = 20, cfl: float = 10.0, workers: int = 8, bytes_per_cell: float | None = None,
                 tree: tuple[int, int] = (4, 512), device_metrics: bool = True):
        self.precision = precision
        self.budget = hbm_budget
        self.drop = drop
        self.max_iters = max_iters
        self.check_every = check_every
        self.cfl = cfl
        self.workers = workers
        self.tree = tuple(tree)     # TDTree parameters of the meshes' cell order
        self.device_metrics = device_metrics  # ComputeMetrics on the GPU (bitwise the host's)
        # device bytes per cell, refined from the contexts actually created
        self.bytes_per_cell = bytes_per_cell or (620.0 if precision == MFS_FP64 else 900.0)

    def run(self, requests: list[Request], progress=None) -> dict:
        reqs = sorted(requests, key=lambda r: r.arrival)
        batch = Batch(self.precision)
        # two pools, so that building a device context (host layout and
        # upload, once memory is reserved) never queues behind mesh set-ups
        setup_pool = cf.ThreadPoolExecutor(self.workers)
        ctx_pool = cf.ThreadPoolExecutor(max(1, self.workers // 2))
        start = time.perf_counter()
        now = lambda: time.perf_counter() - start  # noqa: E731
        nxt = 0                     # next request not yet submitted for set-up
        building: dict[cf.Future, Request] = {}
        ready: list[tuple[Request, Grid, object]] = []
        creating: dict[cf.Future, tuple[Request, Grid, float]] = {}
        live: dict[int, tuple[Request, object, Grid]] = {}   # slot -> (request, member, grid)
        used = 0.0                  # device bytes reserved or held
        done = []
        peak_live = 0
        try:
            while len(done) < len(reqs):
                t = now()
                while nxt < len(reqs) and reqs[nxt].arrival <= t:
                    building[setup_pool.submit(_setup, reqs[nxt], self.cfl, self.tree, self.device_metrics)] = reqs[nxt]
                    nxt += 1
                for fut in [f for f in building if f.done()]:
                    r = building.pop(fut)
                    g, p = fut.result()
                    r.t_ready = now()
                    ready.append((r, g, p))
                ready.sort(key=lambda e: e[0].arrival)
                # reserve memory in arrival order while the budget allows; the
                # context is built on a worker while the batch steps
                while ready:
                    r, g, p = ready[0]
                    need = g.info["nTCell"] * self.bytes_per_cell
                    if (live or creating) and used + need > self.budget:
                        break
                    ready.pop(0)
                    used += need
                    creating[ctx_pool.submit(tuned_solver, g, p, self.precision)] = (r, g, need)
                # built contexts join the batch (cheap: no host layout work here)
                for fut in [f for f in creating if f.done()]:
                    r, g, need = creating.pop(fut)
                    s = fut.result()
                    nc = g.info["nTCell"]
                    dev_bytes = s.info()["device_bytes"]
                    self.bytes_per_cell = max(self.bytes_per_cell, dev_bytes / nc)
                    used += dev_bytes - need
                    m = batch.attach(s)
                    r.t_admit = now()
                    r.n_cells = nc
                    live[m.slot] = (r, m, g)
                peak_live = max(peak_live, len(live))
                if live:
                    batch.step(self.check_every)
                    batch.sync()
                    for slot, (r, m, g) in list(live.items()):
                        h = m.history()
                        if _converged(h, self.drop) or len(h) >= self.max_iters:
                            _finish(r, m, h, self.drop, now())
                            used -= m.info()["device_bytes"]
                            batch.remove(m)
                            g.close()
                            del live[slot]
                            done.append(r)
                            if progress:
                                progress(r)
                elif building or creating:
                    # nothing on the device: wait for the next set-up or context
                    cf.wait(list(building) + list(creating), timeout=0.05, return_when=cf.FIRST_COMPLETED)
                elif nxt < len(reqs):
                    time.sleep(max(0.0, min(0.05, reqs[nxt].arrival - now())))
            wall = now()
        finally:
            # also when a set-up or a context fails: no worker outlives the
            # run, and the batch frees the device memory of its solves
            setup_pool.shutdown(wait=True, cancel_futures=True)
            ctx_pool.shutdown(wait=True, cancel_futures=True)
            batch.close()
        return summarize(done, wall, peak_live)


def run_sequential(requests: list[Request], drop: float = 1e-8, max_iters: int = 20000, check_every: int = 20,
                   cfl: float = 10.0, precision: int = MFS_FP64, progress=None,
                   tree: tuple[int, int] = (4, 512), device_metrics: bool = True) -> dict:
    """The baseline: requests served one at a time in arrival order, each set
    up and solved alone."""
    reqs = sorted(requests, key=lambda r: r.arrival)
    start = time.perf_counter()
    now = lambda: time.perf_counter() - start  # noqa: E731
    done = []
    for r in reqs:
        wait = r.arrival - now()
        if wait > 0:
            time.sleep(wait)
        g, p = _setup(r, cfl, tree, device_metrics)
        r.t_ready = r.t_admit = now()
        s = tuned_solver(g, p, precision)
        r.n_cells = g.info["nTCell"]
        while True:
            s.step(check_every)
            h = s.history()
            if _converged(h, drop) or len(h) >= max_iters:
                break
        _finish(r, s, h, drop, now())
        s.close()
        g.close()
        done.append(r)
        if progress:
            progress(r)
    return summarize(done, now(), 1)


def summarize(done: list[Request], wall: float, peak_live: int) -> dict:
    lat = np.array([r.latency for r in done])
    return {
        "solves": len(done),
        "converged": int(sum(r.converged for r in done)),
        "wall_s": round(wall, 3),
        "solves_per_hour": round(len(done) / wall * 3600.0, 1) if wall > 0 else None,
        "p50_s": round(float(np.percentile(lat, 50)), 3) if len(lat) else None,
        "p99_s": round(float(np.percentile(lat, 99)), 3) if len(lat) else None,
        "cell_iters": int(sum(r.n_cells * r.iters for r in done)),
        "peak_concurrent": peak_live,
    }
