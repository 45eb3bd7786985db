"""MI355X-native MiniFS solver path.

Host-side mirror of the reference's set-up and kernel interface
(src/Kernels/load_grid.h, cal_flux.h, compute_timestep.h, forward_lusgs.h,
cal_gradient.h; driver loop src/main.cpp:198-291) over the C ABI of
libmfs_hip.so.  Grid is the PolyGrid after LoadGrid/Decoupling/
PreprocessGrid; Solver is the device context on which the kernel entry
points run.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib
from ._lib import FIELD_IDS, MFS_FP64, MFS_MIXED, MfsError, check, lib

__all__ = ["Grid", "Solver", "Batch", "make_params", "wall_force", "MFS_FP64", "MFS_MIXED", "MfsError", "lib"]

BC_WALL, BC_SYMM, BC_FAR_FIELD, BC_FAR_FIELD_MOD, BC_INTERFACE = 3, 4, 6, 61, 10


def make_params(mach: float = 0.8395, alpha: float = 3.06, cfl: float | None = None, **over) -> _lib.MfsParams:
    """Reference defaults (para_field_global.cpp:37-146) with the flow
    condition overridden, then the derived values recomputed
    (para_field_global.cpp:150-209)."""
    p = _lib.MfsParams()
    lib().mfs_params_default(C.byref(p))
    p.mach = mach
    p.alpha = alpha
    if cfl is not None:
        p.cfl_start = cfl
        p.cfl_end = cfl
    for k, v in over.items():
        setattr(p, k, v)
    check(lib().mfs_params_derive(C.byref(p)), "mfs_params_derive")
    return p


_INT_ARRAYS = {"nNPF", "f2n", "f2c", "bctype", "cellPerm", "facePerm", "nodePerm", "tdtree_nodes"}


class Grid:
    """Host grid: FlowStar mesh + boundary conditions, optionally TDTree
    renumbered (Decoupling) and preprocessed (metrics, freestream, ghosts)."""

    def __init__(self, handle: C.c_void_p):
        self._h = handle

    @classmethod
    def box_bump(cls, nx: int, ny: int | None = None, nz: int | None = None, jitter: float = 0.1,
                 bump_height: float = 0.05, bump_sigma: float = 0.1, seed: int = 20260101,
                 cells: str = "tet") -> "Grid":
        """Box + Gaussian-bump mesh; cells = "tet" (6 per hex), "hex" or "prism" (2 per hex)."""
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        kind = {"tet": 0, "hex": 1, "prism": 2}[cells]
        h = C.c_void_p()
        check(lib().mfs_grid_generate_box(nx, ny, nz, kind, jitter, bump_height, bump_sigma, seed, C.byref(h)),
              "mfs_grid_generate_box")
        return cls(h)

    @classmethod
    def load(cls, mesh_path: str, bc_path: str) -> "Grid":
        h = C.c_void_p()
        check(lib().mfs_grid_load(mesh_path.encode(), bc_path.encode(), C.byref(h)), "mfs_grid_load")
        return cls(h)

    def save(self, mesh_path: str, bc_path: str) -> None:
        check(lib().mfs_grid_save(self._h, mesh_path.encode(), bc_path.encode()), "mfs_grid_save")

    def decouple(self, nb_parts: int = 4, part_size: int = 512) -> "Grid":
        """TDTree decomposition and renumbering (Decoupling); nb_parts and
        part_size are the tree's parameters (the reference uses 4 and 512)."""
        check(lib().mfs_grid_decouple_ex(self._h, nb_parts, part_size), "mfs_grid_decouple_ex")
        return self

    def preprocess(self, params: _lib.MfsParams, device: bool = False) -> "Grid":
        """PreprocessGrid; with device=True its ComputeMetrics runs on the GPU
        (bitwise the same results)."""
        fn = lib().mfs_grid_preprocess_device if device else lib().mfs_grid_preprocess
        check(fn(self._h, C.byref(params)), "mfs_grid_preprocess")
        return self

    @property
    def info(self) -> dict:
        i = _lib.MfsGridInfo()
        check(lib().mfs_grid_get_info(self._h, C.byref(i)), "mfs_grid_get_info")
        return {f: getattr(i, f) for f, _ in i._fields_}

    def array(self, name: str) -> np.ndarray:
        nbytes = lib().mfs_grid_array_bytes(self._h, name.encode())
        dt = np.int32 if name in _INT_ARRAYS else np.float64
        out = np.empty(nbytes // np.dtype(dt).itemsize, dtype=dt)
        if nbytes:
            check(lib().mfs_grid_get_array(self._h, name.encode(), out.ctypes.data_as(C.c_void_p), nbytes),
                  "mfs_grid_get_array")
        return out

    def view(self) -> _lib.MfsMeshView:
        v = _lib.MfsMeshView()
        check(lib().mfs_grid_view(self._h, C.byref(v)), "mfs_grid_view")
        return v

    def close(self) -> None:
        if self._h:
            lib().mfs_grid_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def wall_force(grid: Grid, p_field: np.ndarray, params: _lib.MfsParams) -> tuple[np.ndarray, float]:
    """Pressure force on the WALL patch of a preprocessed grid,
    F = sum over wall faces of (p_c1 + p_bar) n A (p is the gauge pressure
    of the reference's state, reference layout), and the lift-like component
    L = -F_x sin(alpha) + F_z cos(alpha).  Not part of the reference's loop:
    the figure the mixed-precision results are compared on."""
    f2c = grid.array("f2c").reshape(-1, 2)
    wall = np.flatnonzero(grid.array("bctype") == 3)  # WALL (src/Base/constant.h)
    pp = np.asarray(p_field)[f2c[wall, 0]] + params.p_bar
    area = grid.array("area")[wall]
    F = np.array([np.sum(pp * grid.array(nm)[wall] * area) for nm in ("xfn", "yfn", "zfn")])
    a = np.deg2rad(params.alpha)
    return F, float(-F[0] * np.sin(a) + F[2] * np.cos(a))


class Solver:
    """Device context with the reference's kernel entry points."""

    def __init__(self, grid: Grid, params: _lib.MfsParams, precision: int = MFS_FP64, tile_cells: int = 0,
                 sweep_waves: int = 0, sweep_by_tiles: int = 0, gradient_stream: int = 0, sweep_tail_grid: int = 0):
        self._h = C.c_void_p()
        view = grid.view()
        self.params = params
        opt = _lib.MfsCtxOptions(tile_cells, sweep_waves, sweep_by_tiles, gradient_stream, sweep_tail_grid)
        check(lib().mfs_ctx_create_ex(C.byref(view), C.byref(params), precision, C.byref(opt), C.byref(self._h)),
              "mfs_ctx_create_ex")
        self.grid = grid

    # -- the reference's per-iteration kernels (main.cpp:201-291)
    def zero_residuals(self):
        check(lib().mfs_zero_residuals(self._h), "mfs_zero_residuals")

    def inviscid_flux(self):
        check(lib().mfs_inviscid_flux(self._h), "mfs_inviscid_flux")

    def compute_timestep(self):
        check(lib().mfs_compute_timestep(self._h), "mfs_compute_timestep")

    def forward_lusgs(self):
        check(lib().mfs_forward_lusgs(self._h), "mfs_forward_lusgs")

    def pass_interface_data(self):
        check(lib().mfs_pass_interface_data(self._h), "mfs_pass_interface_data")

    def set_ghost_variables(self):
        """SetGhostVariables on the device state (load_grid.cpp:953-1212)."""
        check(lib().mfs_set_ghost_variables(self._h), "mfs_set_ghost_variables")

    def ghost_refresh(self, on: bool = True):
        """Make step() call SetGhostVariables after each update (the call
        main.cpp:283 leaves commented out)."""
        check(lib().mfs_ctx_set_ghost_refresh(self._h, 1 if on else 0), "mfs_ctx_set_ghost_refresh")

    def calculate_gradient(self):
        check(lib().mfs_calculate_gradient(self._h), "mfs_calculate_gradient")

    def iteration_done(self):
        check(lib().mfs_iteration_done(self._h), "mfs_iteration_done")

    def iterate_reference_order(self, n: int = 1):
        """One driver-loop iteration through the individual entry points."""
        for _ in range(n):
            self.zero_residuals()
            self.inviscid_flux()
            self.compute_timestep()
            self.forward_lusgs()
            self.pass_interface_data()
            self.calculate_gradient()
            self.iteration_done()

    def step(self, n: int = 1):
        check(lib().mfs_step(self._h, n), "mfs_step")

    def sync(self):
        check(lib().mfs_sync(self._h), "mfs_sync")

    def history(self) -> np.ndarray:
        n = lib().mfs_history_count(self._h)
        out = np.zeros((n, 7), dtype=np.float64)
        if n:
            got = check(lib().mfs_get_history(self._h, out.ctypes.data_as(_lib._DP), n), "mfs_get_history")
            out = out[:got]
        return out

    def field(self, name: str) -> np.ndarray:
        fid = FIELD_IDS[name]
        n = lib().mfs_field_count(self._h, fid)
        out = np.empty(n, dtype=np.float64)
        check(lib().mfs_download(self._h, fid, out.ctypes.data_as(_lib._DP), n), "mfs_download")
        return out

    def upload(self, name: str, values: np.ndarray):
        fid = FIELD_IDS[name]
        a = np.ascontiguousarray(values, dtype=np.float64)
        check(lib().mfs_upload(self._h, fid, a.ctypes.data_as(_lib._DP), a.size), "mfs_upload")

    def scalars(self) -> dict:
        s = _lib.MfsScalars()
        check(lib().mfs_get_scalars(self._h, C.byref(s)), "mfs_get_scalars")
        return {f: getattr(s, f) for f, _ in s._fields_}

    def info(self) -> dict:
        i = _lib.MfsCtxInfo()
        check(lib().mfs_ctx_get_info(self._h, C.byref(i)), "mfs_ctx_get_info")
        return {f: getattr(i, f) for f, _ in i._fields_}

    def enable_timing(self, on: bool | int = True):
        """True/1: time every kernel group; 2: the two sweeps only; False/0: off."""
        check(lib().mfs_ctx_enable_timing(self._h, int(on)), "mfs_ctx_enable_timing")

    def timing(self, name: str) -> tuple[float, int]:
        ms = C.c_double()
        cnt = C.c_int()
        check(lib().mfs_ctx_get_timing(self._h, name.encode(), C.byref(ms), C.byref(cnt)), "mfs_ctx_get_timing")
        return ms.value, cnt.value

    def reset_timing(self):
        check(lib().mfs_ctx_reset_timing(self._h), "mfs_ctx_reset_timing")

    def close(self):
        if self._h:
            lib().mfs_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _BatchMember(Solver):
    """A solve owned by a Batch: the Solver queries on its context; the batch
    frees it."""

    def __init__(self, batch: "Batch", slot: int, grid: Grid, params: _lib.MfsParams):
        self._h = C.c_void_p(lib().mfs_batch_ctx(batch._h, slot))
        self.params = params
        self.grid = grid
        self.slot = slot
        self._batch = batch

    def close(self):
        self._h = None  # owned by the batch

    def solve(self, *a, **k):
        raise RuntimeError("step a batch member through its Batch")


class Batch:
    """Several independent solves stepped together (mfs_batch_*)."""

    def __init__(self, precision: int = MFS_FP64, gradient_stream: int = 0):
        self._h = C.c_void_p()
        opt = _lib.MfsBatchOptions(gradient_stream)
        check(lib().mfs_batch_create_ex(precision, C.byref(opt), C.byref(self._h)), "mfs_batch_create_ex")
        self.members: dict[int, _BatchMember] = {}

    def add(self, grid: Grid, params: _lib.MfsParams) -> _BatchMember:
        view = grid.view()
        slot = C.c_int()
        check(lib().mfs_batch_add(self._h, C.byref(view), C.byref(params), C.byref(slot)), "mfs_batch_add")
        m = _BatchMember(self, slot.value, grid, params)
        self.members[slot.value] = m
        return m

    def attach(self, solver: Solver) -> _BatchMember:
        """Move a Solver's context (same precision, built on any thread,
        possibly stepped alone already) into the batch; the Solver object is
        emptied and the returned member stands for the solve."""
        slot = C.c_int()
        check(lib().mfs_batch_attach(self._h, solver._h, C.byref(slot)), "mfs_batch_attach")
        solver._h = None
        m = _BatchMember(self, slot.value, solver.grid, solver.params)
        self.members[slot.value] = m
        return m

    def remove(self, member: _BatchMember):
        check(lib().mfs_batch_remove(self._h, member.slot), "mfs_batch_remove")
        self.members.pop(member.slot, None)
        member._h = None

    def step(self, n: int = 1):
        check(lib().mfs_batch_step(self._h, n), "mfs_batch_step")

    def sync(self):
        check(lib().mfs_batch_sync(self._h), "mfs_batch_sync")

    def __len__(self):
        return lib().mfs_batch_count(self._h)

    def close(self):
        if self._h:
            for m in self.members.values():
                m._h = None
            self.members.clear()
            lib().mfs_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
