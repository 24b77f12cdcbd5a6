"""Drop-in for the reference's run loop
(tinyfluids/jax_tinyfluids/time_integration.py:94-128, loop body :63-92).

``time_integration`` keeps the reference signature and returns
``(state, num_iterations)``. The loop runs on the device: each step is one
fused stencil launch plus a one-thread update of t, dt and the loop flag in a
device-resident control block; the host checks the flag once per chunk of
steps, so steps past the end are empty launches.
"""

from __future__ import annotations

import ctypes
import math
import os

import torch

from . import _lib
from .fluid import _ptr, _scalar, _stream, check_state, workspace

# Grid shapes whose loop tiling has been measured in this process. The first
# run on a new shape times the candidate tilings on its own state (as JAX
# compiles on the first call of a shape); TINYFLUIDS_AUTOTUNE=0 keeps the
# untuned default. Results do not depend on the tiling.
_TUNED = set()


def autotune(primitive_state, grid_spacing, gamma, handle_boundaries=False, scratch=None):
    """Measures the loop tilings of this state's shape (tf_autotune_loop) and
    keeps the fastest for later runs. Returns (tile_x, tile_y, ms per step)."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    scratch = torch.empty_like(s) if scratch is None else scratch
    ws = workspace(nx, ny, nz, s.device)
    tx, ty, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0.0)
    _lib.check(_lib.load().tf_autotune_loop(
        _ptr(s), _ptr(scratch), nx, ny, nz, _scalar(grid_spacing, "grid_spacing"), _scalar(gamma, "gamma"),
        int(bool(handle_boundaries)), _ptr(ws), ctypes.byref(tx), ctypes.byref(ty), ctypes.byref(ms),
        _stream(s.device)))
    _TUNED.add((nx, ny, nz))
    return tx.value, ty.value, ms.value


def _run(primitive_state, grid_spacing, t_final, gamma, handle_boundaries, max_steps, C_CFL):
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    state = s.clone()
    scratch = torch.empty_like(state)
    if (nx, ny, nz) not in _TUNED and os.environ.get("TINYFLUIDS_AUTOTUNE", "1") != "0":
        autotune(state, grid_spacing, gamma, handle_boundaries, scratch)
    ws = workspace(nx, ny, nz, state.device)
    n = ctypes.c_int64(0)
    t = ctypes.c_double(0.0)
    in_scratch = ctypes.c_int(0)
    nonfinite = ctypes.c_int(0)
    _lib.check(_lib.load().tf_time_integration(
        _ptr(state), _ptr(scratch), nx, ny, nz, _scalar(grid_spacing, "grid_spacing"), float(t_final),
        _scalar(gamma, "gamma"), int(bool(handle_boundaries)), int(max_steps), float(C_CFL), _ptr(ws),
        ctypes.byref(n), ctypes.byref(t), ctypes.byref(in_scratch), ctypes.byref(nonfinite), _stream(state.device)))
    out = scratch if in_scratch.value else state
    return out, int(n.value), float(t.value), bool(nonfinite.value)


def time_integration(primitive_state, grid_spacing, t_final, gamma, handle_boundaries, shard_mapped, padding=None,
                     split=None, *, check_finite=False):
    """time_integration.py:94-128. ``shard_mapped``, ``padding`` and ``split``
    belong to the reference's multi-GPU branch, which this one-GPU build does
    not run (the grid wraps periodically inside the kernel).

    As in the reference, a non-finite dt ends the loop after the step that
    used it and the partial state is returned; ``check_finite=True`` raises
    FloatingPointError instead."""
    if shard_mapped or padding is not None or split is not None:
        raise NotImplementedError("shard_mapped/padding/split: multi-GPU domain decomposition is not part of this "
                                  "single-GPU build")
    state, n, _, nonfinite = _run(primitive_state, grid_spacing, t_final, gamma, handle_boundaries, -1, 0.4)
    if check_finite and nonfinite:
        raise FloatingPointError(f"non-finite signal speed after {n} iterations")
    return state, n


def run_steps(primitive_state, grid_spacing, gamma, num_steps, handle_boundaries=False, C_CFL=0.4):
    """Exactly ``num_steps`` steps (the benchmark workloads quote step counts).
    Returns (state, t, nonfinite)."""
    state, n, t, nonfinite = _run(primitive_state, grid_spacing, math.inf, gamma, handle_boundaries,
                                  int(num_steps), C_CFL)
    return state, t, nonfinite
