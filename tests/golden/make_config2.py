"""Writes the 6.0M-cell fixtures of tests/golden/ by running the reference's
own serial build (oracle/_ref/ref_driver, compiled from the reference sources
by oracle/Makefile) for 500 iterations on the 100x100x100x6 box+bump mesh.

The workload: box+bump generator, seed 20260101, TDTree order with 4 parts
and 512-cell leaves; FP64 Euler, Mach 0.5, AoA 0, CFL 10, 500 LU-SGS
iterations (the driver loop of src/main.cpp:198-291).

Fixtures:
  box100_history.f64 : the reference's 500 x 7 history (five residual norms,
                       running dt_max, dt_min after every iteration)
  box100.json        : SHA-256 of the mesh file written by Grid.save, and
                       SHA-256 of the reference's 11 arrays after iteration
                       500 (rho, u, v, w, p, res, dt, dq, dqdx, dqdy, dqdz,
                       little-endian, the reference layout)

The arrays themselves (about 1.5 GB) are not committed.  The reference takes
about half an hour on one core.

Run from the repository root:  python tests/golden/make_config2.py [WORKDIR]
"""
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
HERE = os.path.dirname(os.path.abspath(__file__))

import numpy as np  # noqa: E402

import minifs_amd as mfs  # noqa: E402

REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
FIELDS = ["rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz"]
N, SEED, TREE = 100, 20260101, (4, 512)
MACH, ALPHA, CFL, ITERS = 0.5, 0.0, 10.0, 500


def sha(path):
    h = hashlib.sha256()
    with open(path, "rb") as f:
        for blk in iter(lambda: f.read(1 << 24), b""):
            h.update(blk)
    return h.hexdigest()


def main():
    if not os.path.exists(REF):
        sys.exit("build oracle/_ref/ref_driver first (make -C oracle ref)")
    work = sys.argv[1] if len(sys.argv) > 1 else tempfile.mkdtemp(prefix="mfs_c2_")
    os.makedirs(work, exist_ok=True)
    mesh, bc = os.path.join(work, "box100.mfl"), os.path.join(work, "box100.bc")
    g = mfs.Grid.box_bump(N, seed=SEED).decouple(*TREE)
    info = g.info
    g.save(mesh, bc)
    g.close()
    out = os.path.join(work, "ref")
    os.makedirs(out, exist_ok=True)
    t0 = time.time()
    subprocess.run([REF, mesh, bc, out, repr(MACH), repr(ALPHA), repr(CFL), str(ITERS), str(ITERS), "nogeo"],
                   check=True, stdout=subprocess.DEVNULL)
    wall = time.time() - t0
    hist = np.fromfile(os.path.join(out, "history.f64")).reshape(-1, 7)
    assert hist.shape == (ITERS, 7)
    hist.tofile(os.path.join(HERE, "box100_history.f64"))
    man = {
        "generator": f"minifs_amd.Grid.box_bump({N}, seed={SEED}).decouple{TREE}",
        "nTCell": info["nTCell"], "nTFace": info["nTFace"], "nBFace": info["nBFace"], "nTNode": info["nTNode"],
        "mach": MACH, "alpha": ALPHA, "cfl": CFL, "iters": ITERS,
        "mesh_sha256": sha(mesh), "bc_text": open(bc).read(),
        "final_sha256": {nm: sha(os.path.join(out, f"it{ITERS}_{nm}.f64")) for nm in FIELDS},
        "reference_summary": open(os.path.join(out, "summary.txt")).read().splitlines()[0],
        "reference_wall_s": round(wall, 1),
    }
    with open(os.path.join(HERE, "box100.json"), "w") as f:
        json.dump(man, f, indent=1)
        f.write("\n")
    print(json.dumps(man, indent=1))


if __name__ == "__main__":
    main()
