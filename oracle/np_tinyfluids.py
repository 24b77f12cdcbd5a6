"""NumPy restatement of the jax_tinyfluids time step -- TEST INFRASTRUCTURE ONLY.

This module is the oracle: a CPU restatement of the reference algorithm, used by
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` as the checker. The product path (``tinyfluids_amd``) never imports
it and has no CPU fallback.

Every function follows the reference op for op, in the same evaluation order,
parametrised by dtype: float64 is the oracle proper, float32 reproduces the
reference's own precision (JAX with x64 disabled, so Python scalars such as
``gamma``, ``grid_spacing``, ``C_CFL`` and the ``t`` accumulator are rounded to
float32 before they meet the state, as JAX's weak typing does inside ``jit``).

Reference files (paths relative to the reference repository root):
  fluid.py            = tinyfluids/jax_tinyfluids/fluid.py
  time_integration.py = tinyfluids/jax_tinyfluids/time_integration.py

Pinning (every function has at least one independent check in tests/):
  conserved_state_from_primitive / primitive_state_from_conserved
      hand-computed scalar cases, f64 round trip, textbook identities.
  speed_of_sound, get_wave_speeds, _euler_flux, _hll_solver
      hand-computed fluxes, HLL consistency F(W,W)=F(W), supersonic upwind
      limits (S_L>=0 gives F_L, S_R<=0 gives F_R).
  _evolve_state_along_axis / _evolve_state
      periodic conservation, uniform-velocity advection keeps v and p exact,
      mirror symmetry, convergence to the exact (Toro) Sod solution on every
      axis, scalar C restatement bit-equal in float64.
  _open_boundaries      equals an index clamp for one ghost cell and a
                        separable index map for 0-3 ghost cells down to
                        n = g + 1 (independent formulations, brute force).
  _cfl_time_step        closed form for uniform states; per-cell identity.
  time_integration      iteration counts 24/51/79 to t=0.2 held in the
                        reference's results/scaling_data*.npz; the
                        reference's own check figures (figures/check_{N}.png)
                        rendered from its runs.
"""

from __future__ import annotations

import numpy as np

# constants, fluid.py:14-22
DENSITY_INDEX = 0
VELOCITY_X_INDEX = 1
VELOCITY_Y_INDEX = 2
VELOCITY_Z_INDEX = 3
PRESSURE_INDEX = 4
VAR_AXIS = 0
X_AXIS = 1
Y_AXIS = 2
Z_AXIS = 3


def _s(x, dtype):
    """A Python scalar as a scalar of the state dtype (JAX weak typing in jit)."""
    return np.asarray(x, dtype=np.float64).astype(dtype)[()]


# ------------------------------------------------------------------ fluid.py:28-67
def conserved_state_from_primitive(primitive_state, gamma):
    dt = primitive_state.dtype
    g = _s(gamma, dt)
    rho = primitive_state[DENSITY_INDEX]
    p = primitive_state[PRESSURE_INDEX]
    utotal = np.sqrt(
        primitive_state[VELOCITY_X_INDEX] ** 2
        + primitive_state[VELOCITY_Y_INDEX] ** 2
        + primitive_state[VELOCITY_Z_INDEX] ** 2
    )
    E = p / (g - _s(1, dt)) + _s(0.5, dt) * rho * utotal ** 2
    out = primitive_state.copy()
    out[PRESSURE_INDEX] = E
    out[VELOCITY_X_INDEX] = rho * primitive_state[VELOCITY_X_INDEX]
    out[VELOCITY_Y_INDEX] = rho * primitive_state[VELOCITY_Y_INDEX]
    out[VELOCITY_Z_INDEX] = rho * primitive_state[VELOCITY_Z_INDEX]
    return out


# ------------------------------------------------------------------ fluid.py:69-103
def primitive_state_from_conserved(conserved_state, gamma):
    dt = conserved_state.dtype
    g = _s(gamma, dt)
    rho = conserved_state[DENSITY_INDEX]
    E = conserved_state[PRESSURE_INDEX]
    ux = conserved_state[VELOCITY_X_INDEX] / rho
    uy = conserved_state[VELOCITY_Y_INDEX] / rho
    uz = conserved_state[VELOCITY_Z_INDEX] / rho
    u = np.sqrt(ux ** 2 + uy ** 2 + uz ** 2)
    p = (g - _s(1, dt)) * rho * (E / rho - _s(0.5, dt) * u ** 2)
    out = conserved_state.copy()
    out[PRESSURE_INDEX] = p
    out[VELOCITY_X_INDEX] = ux
    out[VELOCITY_Y_INDEX] = uy
    out[VELOCITY_Z_INDEX] = uz
    return out


# ------------------------------------------------------------------ fluid.py:105-117
def speed_of_sound(rho, p, gamma):
    g = _s(gamma, rho.dtype)
    return np.sqrt(g * p / rho)


# ------------------------------------------------------------------ fluid.py:119-160
def get_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index):
    rho_L = primitives_left[DENSITY_INDEX]
    u_L = primitives_left[flux_direction_index]
    p_L = primitives_left[PRESSURE_INDEX]
    rho_R = primitives_right[DENSITY_INDEX]
    u_R = primitives_right[flux_direction_index]
    p_R = primitives_right[PRESSURE_INDEX]
    c_L = speed_of_sound(rho_L, p_L, gamma)
    c_R = speed_of_sound(rho_R, p_R, gamma)
    zero = _s(0, primitives_left.dtype)
    wave_speeds_right_plus = np.maximum(np.maximum(u_L + c_L, u_R + c_R), zero)
    wave_speeds_left_minus = np.minimum(np.minimum(u_L - c_L, u_R - c_R), zero)
    return wave_speeds_left_minus, wave_speeds_right_plus


# ------------------------------------------------------------------ fluid.py:162-195
def get_max_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index):
    sl, sr = get_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index)
    return np.maximum(np.max(np.abs(sr)), np.max(np.abs(sl)))


# ------------------------------------------------------------------ fluid.py:206-259
def _euler_flux(primitive_state, gamma, flux_direction_index):
    dt = primitive_state.dtype
    g = _s(gamma, dt)
    rho = primitive_state[DENSITY_INDEX]
    p = primitive_state[PRESSURE_INDEX]
    flux_vector = primitive_state.copy()
    utotal = np.sqrt(
        primitive_state[VELOCITY_X_INDEX] ** 2
        + primitive_state[VELOCITY_Y_INDEX] ** 2
        + primitive_state[VELOCITY_Z_INDEX] ** 2
    )
    E = p / (g - _s(1, dt)) + _s(0.5, dt) * rho * utotal ** 2
    flux_vector[PRESSURE_INDEX] = flux_vector[PRESSURE_INDEX] + E
    flux_vector[VELOCITY_X_INDEX] = primitive_state[VELOCITY_X_INDEX] * rho
    flux_vector[VELOCITY_Y_INDEX] = primitive_state[VELOCITY_Y_INDEX] * rho
    flux_vector[VELOCITY_Z_INDEX] = primitive_state[VELOCITY_Z_INDEX] * rho
    flux_vector = primitive_state[flux_direction_index] * flux_vector
    flux_vector[flux_direction_index] = flux_vector[flux_direction_index] + p
    return flux_vector


# ------------------------------------------------------------------ fluid.py:261-305
def _hll_solver(primitives_left, primitives_right, gamma, flux_direction_index):
    fluxes_left = _euler_flux(primitives_left, gamma, flux_direction_index)
    fluxes_right = _euler_flux(primitives_right, gamma, flux_direction_index)
    sl, sr = get_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index)
    conserved_left = conserved_state_from_primitive(primitives_left, gamma)
    conserved_right = conserved_state_from_primitive(primitives_right, gamma)
    return (sr * fluxes_left - sl * fluxes_right + sl * sr * (conserved_right - conserved_left)) / (sr - sl)


# ------------------------------------------------------------------ fluid.py:317-413
def _set_along_axis(primitive_state, axis, set_index, get_index):
    s_set = (slice(None),) * axis + (set_index,) + (slice(None),) * (primitive_state.ndim - axis - 1)
    s_get = (slice(None),) * axis + (get_index,) + (slice(None),) * (primitive_state.ndim - axis - 1)
    out = primitive_state.copy()
    out[s_set] = primitive_state[s_get]
    return out


def _open_right_boundary(primitive_state, num_ghost_cells, axis):
    get_index = -num_ghost_cells - 1
    for set_index in range(-num_ghost_cells, 0):
        primitive_state = _set_along_axis(primitive_state, axis, set_index, get_index)
    return primitive_state


def _open_left_boundary(primitive_state, num_ghost_cells, axis):
    get_index = num_ghost_cells
    for set_index in range(num_ghost_cells):
        primitive_state = _set_along_axis(primitive_state, axis, set_index, get_index)
    return primitive_state


def _open_boundaries(primitive_state, num_ghost_cells):
    for axis in (X_AXIS, Y_AXIS, Z_AXIS):
        primitive_state = _open_left_boundary(primitive_state, num_ghost_cells, axis)
        primitive_state = _open_right_boundary(primitive_state, num_ghost_cells, axis)
    return primitive_state


# ------------------------------------------------------------------ fluid.py:425-504
def _cfl_time_step(primitive_state, grid_spacing, gamma, C_CFL=0.4):
    dt = primitive_state.dtype
    speeds = []
    for axis, vel in ((X_AXIS, VELOCITY_X_INDEX), (Y_AXIS, VELOCITY_Y_INDEX), (Z_AXIS, VELOCITY_Z_INDEX)):
        left = np.roll(primitive_state, shift=1, axis=axis)
        speeds.append(get_max_wave_speeds(left, primitive_state, gamma, vel))
    max_wave_speed = np.maximum(np.maximum(speeds[0], speeds[1]), speeds[2])
    return _s(C_CFL, dt) * _s(grid_spacing, dt) / max_wave_speed


# ------------------------------------------------------------------ fluid.py:506-557
def _evolve_state_along_axis(primitive_state, grid_spacing, dt, gamma, axis):
    d = primitive_state.dtype
    conservative_states = conserved_state_from_primitive(primitive_state, gamma)
    primitive_state_left = np.roll(primitive_state, shift=1, axis=axis)
    fluxes = _hll_solver(primitive_state_left, primitive_state, gamma, axis)
    conserved_change = _s(1, d) / _s(grid_spacing, d) * (fluxes - np.roll(fluxes, shift=-1, axis=axis)) * _s(dt, d)
    conservative_states = conservative_states + conserved_change
    return primitive_state_from_conserved(conservative_states, gamma)


# ------------------------------------------------------------------ fluid.py:560-578
def _evolve_state(primitive_state, grid_spacing, dt, gamma, handle_boundaries=False):
    if handle_boundaries:
        primitive_state = _open_boundaries(primitive_state, 1)
    for axis in (X_AXIS, Y_AXIS, Z_AXIS):
        primitive_state = _evolve_state_along_axis(primitive_state, grid_spacing, dt, gamma, axis)
    return primitive_state


# ------------------------------------------------------------------ time_integration.py:35-92, 127-128
def time_integration(primitive_state, grid_spacing, t_final, gamma, handle_boundaries=False, max_steps=None):
    """``while t < t_final: dt = cfl; state = evolve(state, dt); t += dt; n += 1``.

    No dt clipping (the last step overshoots t_final), and the loop leaves as
    soon as ``t < t_final`` is False, which includes a NaN ``t``
    (time_integration.py:85-90). ``t`` has the state's dtype.
    """
    d = primitive_state.dtype
    t = _s(0.0, d)
    tf = _s(t_final, d)
    n = 0
    while t < tf and (max_steps is None or n < max_steps):
        dt = _cfl_time_step(primitive_state, grid_spacing, gamma)
        primitive_state = _evolve_state(primitive_state, grid_spacing, dt, gamma, handle_boundaries)
        t = t + dt
        n += 1
    return primitive_state, n


def run_steps(primitive_state, grid_spacing, gamma, num_steps, handle_boundaries=False, C_CFL=0.4):
    """Fixed step count variant of the loop above (the benchmark configurations
    quote step counts, not t_final). Returns (state, t, list of dt)."""
    d = primitive_state.dtype
    t = _s(0.0, d)
    dts = []
    for _ in range(num_steps):
        dt = _cfl_time_step(primitive_state, grid_spacing, gamma, C_CFL)
        primitive_state = _evolve_state(primitive_state, grid_spacing, dt, gamma, handle_boundaries)
        t = t + dt
        dts.append(dt)
    return primitive_state, t, dts


# ------------------------------------------------------------------ diagnostics (not in the reference)
def conserved_totals(primitive_state, gamma):
    """Totals of mass, momentum (3) and energy, summed in float64."""
    cons = conserved_state_from_primitive(primitive_state.astype(np.float64), gamma)
    return cons.reshape(5, -1).sum(axis=1)


def relative_l1(test_prim, oracle_prim, gamma):
    """Per conserved field: sum|q_test - q_oracle| / sum|q_oracle|, in float64.

    Fields that are identically zero in the oracle report the absolute
    sum|q_test| instead (relative error is undefined there)."""
    ct = conserved_state_from_primitive(np.asarray(test_prim, dtype=np.float64), gamma).reshape(5, -1)
    co = conserved_state_from_primitive(np.asarray(oracle_prim, dtype=np.float64), gamma).reshape(5, -1)
    num = np.abs(ct - co).sum(axis=1)
    den = np.abs(co).sum(axis=1)
    return np.where(den > 0, num / np.where(den > 0, den, 1.0), num)
