"""tinyfluids_amd: the jax_tinyfluids time step (first-order Godunov, HLL
fluxes, periodic split X -> Y -> Z sweeps, CFL dt) on MI355X / gfx950.

Python entry points keep the reference's names and the primitive
(5, Nx, Ny, Nz) state layout; they call hand-written HIP kernels through the
ctypes C ABI of libtfhip.so (include/tinyfluids.h). There is no CPU fallback.
"""

from ._lib import build_library, load as load_library, TinyfluidsError  # noqa: F401
from .fluid import (  # noqa: F401
    DENSITY_INDEX, VELOCITY_X_INDEX, VELOCITY_Y_INDEX, VELOCITY_Z_INDEX, PRESSURE_INDEX,
    VAR_AXIS, X_AXIS, Y_AXIS, Z_AXIS,
    conserved_state_from_primitive, primitive_state_from_conserved, _cfl_time_step, _open_boundaries,
    _evolve_state_along_axis, _evolve_state, _evolve_state_split, conserved_totals, max_signal_speed,
    speed_of_sound, get_wave_speeds, get_max_wave_speeds, _euler_flux, _hll_solver, _set_along_axis,
    _open_left_boundary, _open_right_boundary,
)
from .time_integration import time_integration, run_steps  # noqa: F401


def build():
    """Compile libtfhip.so for gfx950 (hipcc; no GPU needed)."""
    return build_library(force=True)
