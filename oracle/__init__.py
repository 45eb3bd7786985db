"""Test infrastructure: ctypes binding of the C restatement (liboracle.so)
and the path of the reference build (oracle/_ref/ref_driver).

Only tests/, __graft_entry__.smoke() and bench.py's CPU leg use this
package, and only as the checker or the CPU baseline.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REF_DRIVER = os.path.join(HERE, "_ref", "ref_driver")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run make -C oracle port")
        L = C.CDLL(LIB)
        L.or_load.argtypes = [C.c_char_p, C.c_char_p, C.POINTER(C.c_void_p)]
        L.or_preprocess.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double]
        L.or_run.argtypes = [C.c_void_p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.or_get.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]
        L.or_sizes.argtypes = [C.c_void_p] + [C.POINTER(C.c_int)] * 4
        L.or_set_ghost_refresh.argtypes = [C.c_void_p, C.c_int]
        L.or_set.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_double), C.c_size_t]
        L.or_set_ghosts.argtypes = [C.c_void_p]
        L.or_set_limits.argtypes = [C.c_void_p] + [C.c_double] * 4
        L.or_counts.argtypes = [C.c_void_p, C.POINTER(C.c_longlong), C.c_int, C.c_int]
        L.or_free.argtypes = [C.c_void_p]
        L.or_free.restype = None
        L.or_error.restype = C.c_char_p
        _lib = L
    return _lib


# Guard-branch counter names, in the order of the or_counts enum (mfs_oracle.h).
COUNTS = ("fwd_limit1", "fwd_limit2", "fwd_limit3", "bwd_limit1", "bwd_limit2", "bwd_limit3",
          "relax_rho", "relax_p", "clamp_rho_lo", "clamp_rho_hi", "clamp_p_lo", "clamp_p_hi",
          "cfl_ramp", "cfl_pmin", "cfl_det", "dtmax", "ghost_rho_lo", "ghost_rho_hi", "ghost_p_lo", "ghost_p_hi")


class Oracle:
    """CPU restatement run on one FlowStar mesh file.  limits = (ratio_rhop_min,
    ratio_rhop_max, ratio_p_break, ratio_dtmax) replaces the reference's
    defaults (1e-4, 10, 0.01, 1e80) before the set-up."""

    def __init__(self, mesh: str, bc: str, mach: float, alpha: float, cfl: float, limits=None):
        self._h = C.c_void_p()
        self._check(lib().or_load(mesh.encode(), bc.encode(), C.byref(self._h)))
        if limits is not None:
            self._check(lib().or_set_limits(self._h, *[float(x) for x in limits]))
        self._check(lib().or_preprocess(self._h, mach, alpha, cfl))
        s = [C.c_int() for _ in range(4)]
        lib().or_sizes(self._h, *[C.byref(x) for x in s])
        self.nC, self.nF, self.nB, self.nN = (x.value for x in s)

    @staticmethod
    def _check(rc):
        if rc != 0:
            raise RuntimeError("oracle: " + lib().or_error().decode())

    def run(self, iters: int, threads: int = 0):
        hist = np.zeros((iters, 7))
        t = C.c_double()
        self._check(lib().or_run(self._h, iters, threads, hist.ctypes.data_as(C.POINTER(C.c_double)), C.byref(t)))
        return hist, t.value

    def get(self, name: str) -> np.ndarray:
        n = self.nC + self.nB
        sizes = {"res": 5 * self.nC, "dt": self.nC, "dq": 5 * n, "dqdx": 5 * n, "dqdy": 5 * n, "dqdz": 5 * n}
        for k in ("rho", "u", "v", "w", "p", "xcc", "ycc", "zcc", "vol"):
            sizes[k] = n
        for k in ("xfn", "yfn", "zfn", "area", "xfc", "yfc", "zfc"):
            sizes[k] = self.nF
        out = np.empty(sizes[name])
        self._check(lib().or_get(self._h, name.encode(), out.ctypes.data_as(C.POINTER(C.c_double)), out.size))
        return out

    def set_ghost_refresh(self, on: bool):
        self._check(lib().or_set_ghost_refresh(self._h, 1 if on else 0))

    def set(self, name: str, values: np.ndarray):
        a = np.ascontiguousarray(values, dtype=np.float64)
        self._check(lib().or_set(self._h, name.encode(), a.ctypes.data_as(C.POINTER(C.c_double)), a.size))

    def set_ghosts(self):
        self._check(lib().or_set_ghosts(self._h))

    def counts(self, reset: bool = False) -> dict:
        """How often each guard branch was taken since the set-up or the last reset."""
        out = (C.c_longlong * len(COUNTS))()
        self._check(lib().or_counts(self._h, out, len(COUNTS), 1 if reset else 0))
        return dict(zip(COUNTS, out))

    def close(self):
        if self._h:
            lib().or_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def run_reference(mesh: str, bc: str, out_dir: str, mach: float, alpha: float, cfl: float, iters: int,
                  dumps=(0, 1)) -> np.ndarray:
    """Runs the reference build; returns its per-iteration history."""
    if not os.path.exists(REF_DRIVER):
        raise RuntimeError("oracle/_ref/ref_driver not built")
    os.makedirs(out_dir, exist_ok=True)
    subprocess.run([REF_DRIVER, mesh, bc, out_dir, repr(mach), repr(alpha), repr(cfl), str(iters),
                    ",".join(str(d) for d in dumps)], check=True, stdout=subprocess.DEVNULL, timeout=1200)
    return np.fromfile(os.path.join(out_dir, "history.f64")).reshape(-1, 7)
