"""ctypes front end of the scalar C oracle (oracle/tf_oracle.c) -- TEST
INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU-baseline
leg of the benchmark, never by the product package.

Same algorithm and evaluation order as oracle/np_tinyfluids.py (bit-equal in
float64), but scalar C with OpenMP, so it reaches 128^3-512^3 in minutes.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libtforacle.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        for sfx, R in (("f64", ctypes.c_double), ("f32", ctypes.c_float)):
            getattr(L, f"tfo_{sfx}_sweep").argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, R, R, R, ctypes.c_int]
            getattr(L, f"tfo_{sfx}_open_boundaries").argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"tfo_{sfx}_cfl_time_step")
            f.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, R, R, R]
            f.restype = R
            getattr(L, f"tfo_{sfx}_evolve_state").argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, R, R, R, ctypes.c_int]
            f = getattr(L, f"tfo_{sfx}_time_integration")
            f.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, R, R, R, ctypes.c_int, ctypes.c_long, R, P, P]
            f.restype = ctypes.c_long
            getattr(L, f"tfo_{sfx}_prim_to_cons").argtypes = [P, P, ctypes.c_size_t, R]
            getattr(L, f"tfo_{sfx}_cons_to_prim").argtypes = [P, P, ctypes.c_size_t, R]
            getattr(L, f"tfo_{sfx}_conserved_totals").argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, P]
        L.tfo_num_threads.restype = ctypes.c_int
        _lib = L
    return _lib


def _sfx(a):
    if a.dtype == np.float64:
        return "f64"
    if a.dtype == np.float32:
        return "f32"
    raise TypeError(f"unsupported dtype {a.dtype}")


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def num_threads() -> int:
    return int(lib().tfo_num_threads())


def cfl_time_step(state, grid_spacing, gamma, C_CFL=0.4):
    s = np.ascontiguousarray(state)
    _, nx, ny, nz = s.shape
    return getattr(lib(), f"tfo_{_sfx(s)}_cfl_time_step")(_ptr(s), nx, ny, nz, grid_spacing, gamma, C_CFL)


def evolve_state(state, grid_spacing, dt, gamma, handle_boundaries=False):
    s = np.array(state, copy=True, order="C")
    tmp = np.empty_like(s)
    _, nx, ny, nz = s.shape
    getattr(lib(), f"tfo_{_sfx(s)}_evolve_state")(_ptr(s), _ptr(tmp), nx, ny, nz, grid_spacing, dt, gamma,
                                                   int(handle_boundaries))
    return s


def sweep(state, grid_spacing, dt, gamma, axis):
    s = np.ascontiguousarray(state)
    out = np.empty_like(s)
    _, nx, ny, nz = s.shape
    getattr(lib(), f"tfo_{_sfx(s)}_sweep")(_ptr(s), _ptr(out), nx, ny, nz, grid_spacing, dt, gamma, axis)
    return out


def open_boundaries(state):
    s = np.array(state, copy=True, order="C")
    _, nx, ny, nz = s.shape
    getattr(lib(), f"tfo_{_sfx(s)}_open_boundaries")(_ptr(s), nx, ny, nz)
    return s


def time_integration(state, grid_spacing, t_final, gamma, handle_boundaries=False, max_steps=-1, C_CFL=0.4,
                     in_place=False, record_dt=False):
    """Returns (state, num_iterations, t, dts or None)."""
    s = state if in_place else np.array(state, copy=True, order="C")
    assert s.flags.c_contiguous
    tmp = np.empty_like(s)
    _, nx, ny, nz = s.shape
    R = np.float64 if s.dtype == np.float64 else np.float32
    t_out = np.zeros(1, dtype=R)
    cap = max_steps if max_steps >= 0 else 100000
    dts = np.zeros(cap, dtype=R) if record_dt else None
    n = getattr(lib(), f"tfo_{_sfx(s)}_time_integration")(
        _ptr(s), _ptr(tmp), nx, ny, nz, grid_spacing, t_final, gamma, int(handle_boundaries), max_steps, C_CFL,
        _ptr(t_out), _ptr(dts) if dts is not None else None)
    return s, int(n), float(t_out[0]), (dts[:n] if dts is not None else None)


def run_steps(state, grid_spacing, gamma, num_steps, handle_boundaries=False, C_CFL=0.4, in_place=False,
              record_dt=False):
    """Exactly ``num_steps`` steps. Returns (state, t, dts or None)."""
    s, n, t, dts = time_integration(state, grid_spacing, float("inf"), gamma, handle_boundaries, num_steps, C_CFL,
                                    in_place, record_dt)
    assert n == num_steps
    return s, t, dts


def prim_to_cons(state, gamma):
    s = np.ascontiguousarray(state)
    out = np.empty_like(s)
    getattr(lib(), f"tfo_{_sfx(s)}_prim_to_cons")(_ptr(s), _ptr(out), s[0].size, gamma)
    return out


def cons_to_prim(state, gamma):
    s = np.ascontiguousarray(state)
    out = np.empty_like(s)
    getattr(lib(), f"tfo_{_sfx(s)}_cons_to_prim")(_ptr(s), _ptr(out), s[0].size, gamma)
    return out


def conserved_totals(state, gamma):
    s = np.ascontiguousarray(state)
    _, nx, ny, nz = s.shape
    out = np.zeros(5, dtype=np.float64)
    getattr(lib(), f"tfo_{_sfx(s)}_conserved_totals")(_ptr(s), nx, ny, nz, gamma, _ptr(out))
    return out
