"""GPU test of the serving loop: requests that arrive while others are on
the device are admitted into the running batch, and every result is the
same as the same request solved alone."""
import numpy as np
import pytest

import minifs_amd as mfs
from minifs_amd.serve import Server, make_requests, run_sequential

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [mfs.MFS_FP64, mfs.MFS_MIXED])
def test_served_results_equal_alone(precision):
    print(f"[serve] precision={precision}", flush=True)
    reqs = make_requests(6, rate=40.0, cells_lo=2000, cells_hi=20000, seed=5001)
    srv = Server(precision, hbm_budget=2e9, drop=1e-6, max_iters=4000, check_every=20, workers=4)
    stats = srv.run(reqs)
    assert stats["solves"] == 6 and stats["converged"] == 6
    assert stats["peak_concurrent"] >= 2  # requests overlapped on the device
    alone = make_requests(6, rate=40.0, cells_lo=2000, cells_hi=20000, seed=5001)
    run_sequential(alone, drop=1e-6, max_iters=4000, check_every=20, precision=precision)
    for a, b in zip(sorted(reqs, key=lambda r: r.rid), sorted(alone, key=lambda r: r.rid)):
        assert a.iters == b.iters, (a.rid, a.iters, b.iters)
        assert a.state_sha256 == b.state_sha256, (a.rid, a.state_sha256, b.state_sha256)
        assert a.latency >= 0 and np.isfinite(a.latency)


def test_budget_limits_admission():
    """A budget below two solves keeps them one at a time, and still serves all."""
    reqs = make_requests(3, rate=100.0, cells_lo=5000, cells_hi=6000, seed=7)
    srv = Server(mfs.MFS_FP64, hbm_budget=1.5 * 6000 * 700, drop=1e-4, max_iters=2000, check_every=20, workers=2)
    stats = srv.run(reqs)
    assert stats["solves"] == 3
    assert stats["peak_concurrent"] == 1
