"""The reference-side binding, compiled and run.

oracle/_ref/shim_driver is the reference's own set-up (ReadMFlowGrid_binary,
Decoupling, PreprocessGrid: src/Base/*.cpp and src/Kernels/load_grid.cpp,
compiled from the reference sources) and the driver loop of
src/main.cpp:198-291, linked with minifs_amd/shim/mfs_gpu_shim.cpp in place
of cal_flux.cpp, compute_timestep.cpp, forward_lusgs.cpp and
cal_gradient.cpp.  The device context is built from the arrays of the
reference's own PreprocessGrid through mfs_mesh_view.  Every dump (state,
res, dt, DQ, gradients) and the history must equal the pure reference
build's bit for bit: both drivers sum the same res on the host.
"""
import os
import subprocess

import numpy as np
import pytest

import minifs_amd as mfs
from tests.refrun import run_reference

pytestmark = pytest.mark.gpu

FIELDS = ("rho", "u", "v", "w", "p", "res", "dt", "dq", "dqdx", "dqdy", "dqdz")


@pytest.mark.parametrize("n,cells,mach,alpha,iters", [(20, "tet", 0.5, 0.0, 100), (8, "tet", 0.8, 1.25, 40),
                                                      (6, "hex", 0.5, 0.0, 30), (6, "prism", 0.8, 1.25, 30)])
def test_shim_equals_reference(tmp_path, ref_driver, shim_driver, n, cells, mach, alpha, iters):
    print(f"[shim] box N={n} {cells} mach={mach} alpha={alpha} iters={iters}", flush=True)
    g = mfs.Grid.box_bump(n, cells=cells).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    g.close()
    dumps = (1, iters)
    ref = run_reference(ref_driver, mesh, bc, str(tmp_path / "ref"), mach, alpha, 10.0, iters, dumps=dumps,
                        geometry=False)
    shim = run_reference(shim_driver, mesh, bc, str(tmp_path / "shim"), mach, alpha, 10.0, iters, dumps=dumps,
                         geometry=False)
    np.testing.assert_array_equal(shim["history"], ref["history"])
    for it in dumps:
        for nm in FIELDS:
            a = np.fromfile(os.path.join(shim["dir"], f"it{it}_{nm}.f64"))
            b = np.fromfile(os.path.join(ref["dir"], f"it{it}_{nm}.f64"))
            assert a.shape == b.shape, (it, nm)
            bad = np.flatnonzero(a != b)
            assert bad.size == 0, f"it{it} {nm}: {bad.size} entries differ, first {bad[:4]}"


def test_shim_refuses_host_ghost_refresh(tmp_path, shim_driver):
    """The per-iteration SetGhostVariables of the reference stays a host
    function that the device state never sees; the shim driver says so."""
    g = mfs.Grid.box_bump(4).decouple()
    mesh, bc = str(tmp_path / "m.mfl"), str(tmp_path / "m.bc")
    g.save(mesh, bc)
    g.close()
    os.makedirs(tmp_path / "o")
    r = subprocess.run([shim_driver, mesh, bc, str(tmp_path / "o"), "0.5", "0", "10", "2", "", "ghosts"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2 and "not supported" in r.stderr
