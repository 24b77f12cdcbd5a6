"""ctypes binding of libtfhip.so (the C ABI declared in include/tinyfluids.h).

The library is built in-tree by ``build_library()`` (hipcc for gfx950) and
loaded from this directory. There is no CPU fallback: every entry point of the
package needs the library and a ROCm device, and raises if either is missing.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
# TINYFLUIDS_LIB selects another build of the same library (A/B timing).
LIB_PATH = os.environ.get("TINYFLUIDS_LIB") or os.path.join(_HERE, "libtfhip.so")
_CSRC = os.path.join(_HERE, "csrc")

TF_OK = 0
TF_ERR_ARG = -1
TF_ERR_HIP = -2
TF_ERR_NONFINITE = -3
TF_ERR_UNSUPPORTED = -4
TF_CTRL_BYTES = 128

_lock = threading.Lock()
_lib = None

P = ctypes.c_void_p
I = ctypes.c_int
I64 = ctypes.c_int64
D = ctypes.c_double
SZ = ctypes.c_size_t


class SimDesc(ctypes.Structure):
    """tf_sim_desc"""
    _fields_ = [
        ("buf0", P), ("buf1", P), ("ctrl", P),
        ("nx", I), ("ny", I), ("nz", I), ("handle_boundaries", I),
        ("grid_spacing", D), ("gamma", D), ("c_cfl", D), ("t_final", D),
        ("max_steps", I64),
    ]


class SimStatus(ctypes.Structure):
    """tf_sim_status"""
    _fields_ = [
        ("t", D), ("dt", D), ("n_iter", I64), ("active", ctypes.c_int32), ("nonfinite", ctypes.c_int32),
        ("max_speed", ctypes.c_float), ("pad", ctypes.c_int32 * 3),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "tf_abi_version": (I, []),
    "tf_last_error": (ctypes.c_char_p, []),
    "tf_workspace_bytes": (SZ, [I, I, I]),
    "tf_conserved_state_from_primitive": (I, [P, P, I64, D, P]),
    "tf_primitive_state_from_conserved": (I, [P, P, I64, D, P]),
    "tf_cfl_time_step": (I, [P, I, I, I, D, D, D, P, P, P, P]),
    "tf_open_boundaries": (I, [P, P, I, I, I, I, P]),
    "tf_speed_of_sound": (I, [P, P, P, I64, D, P]),
    "tf_get_wave_speeds": (I, [P, P, I64, D, I, P, P, P]),
    "tf_get_max_wave_speeds": (I, [P, P, I64, D, I, P, P, P]),
    "tf_euler_flux": (I, [P, P, I64, D, I, P]),
    "tf_hll_solver": (I, [P, P, P, I64, D, I, P]),
    "tf_set_along_axis": (I, [P, P, I, I, I, I, I, I, P]),
    "tf_evolve_state_along_axis": (I, [P, P, I, I, I, D, P, D, I, P]),
    "tf_evolve_state": (I, [P, P, I, I, I, D, P, D, I, P, P, P]),
    "tf_time_integration": (I, [P, P, I, I, I, D, D, D, I, I64, D, P,
                                ctypes.POINTER(I64), ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I), P]),
    "tf_time_integration_noalloc": (I, [P, P, I, I, I, D, D, D, I, I64, D,
                                        ctypes.POINTER(I64), ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I),
                                        P]),
    "tf_conserved_totals": (I, [P, I, I, I, D, P, P, P]),
    "tf_sim_reset": (I, [ctypes.POINTER(SimDesc), P]),
    "tf_batch_workspace_bytes": (SZ, [I]),
    "tf_batch_create": (I, [ctypes.POINTER(P), ctypes.POINTER(SimDesc), I, P, SZ, I, I, P]),
    "tf_batch_create_async": (I, [ctypes.POINTER(P), ctypes.POINTER(SimDesc), I, P, SZ, I, I, P]),
    "tf_batch_destroy": (I, [P]),
    "tf_batch_sim_tiling": (I, [P, I, ctypes.POINTER(I), ctypes.POINTER(I), ctypes.POINTER(I)]),
    "tf_batch_step": (I, [P, I, P]),
    "tf_batch_reset": (I, [P, P]),
    "tf_batch_step_graph": (I, [P, I, P]),
    "tf_batch_run_graph": (I, [P, I, P]),
    "tf_batch_status": (I, [P, ctypes.POINTER(SimStatus), P]),
    "tf_batch_status_bytes": (SZ, [P]),
    "tf_batch_status_enqueue": (I, [P, P, P]),
    "tf_batch_status_unpack": (I, [P, P, ctypes.POINTER(SimStatus)]),
    "tf_batch_step_timed": (I, [P, I, ctypes.POINTER(ctypes.c_float), P]),
    "tf_set_line_mode": (I, [I]),
    "tf_hbm_copy": (I, [P, P, SZ, P]),
    "tf_get_tiling": (I, [I, I, I, ctypes.POINTER(I), ctypes.POINTER(I)]),
    "tf_set_tiling": (I, [I, I, I, I, I]),
    "tf_autotune_evolve": (I, [P, P, I, I, I, D, D, P, ctypes.POINTER(I), ctypes.POINTER(I),
                               ctypes.POINTER(ctypes.c_float), P]),
    "tf_autotune_loop": (I, [P, P, I, I, I, D, D, I, P, ctypes.POINTER(I), ctypes.POINTER(I),
                             ctypes.POINTER(ctypes.c_float), P]),
}


class TinyfluidsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def build_library(force: bool = True) -> str:
    """Compile libtfhip.so for gfx950 with hipcc (works without a GPU)."""
    if force and os.path.exists(LIB_PATH):
        os.remove(LIB_PATH)
    subprocess.run(["make", "-s", "-C", _CSRC], check=True)
    return LIB_PATH


def load():
    """Load libtfhip.so and declare its signatures. Raises if it is absent."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; run tinyfluids_amd.build() (hipcc, gfx950) first. "
                "There is no CPU fallback.")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
        return _lib


def check(rc: int):
    if rc != TF_OK:
        msg = load().tf_last_error().decode(errors="replace")
        if rc == TF_ERR_ARG:
            raise ValueError(msg)
        raise TinyfluidsError(rc, msg)
