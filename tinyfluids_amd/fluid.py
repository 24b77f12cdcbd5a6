"""Drop-in for the reference's physics module (tinyfluids/jax_tinyfluids/fluid.py).

Same names, argument order and defaults; states are torch float32 tensors of
shape (5, Nx, Ny, Nz) on a ROCm device, holding primitive variables
(rho, vx, vy, vz, p) as in fluid.py:10-22. Every function runs a gfx950 HIP
kernel through the C ABI on the current torch stream and returns a new tensor
(the reference's arrays are immutable, so callers never see an input change).
Like the reference's jitted functions, calls are asynchronous; synchronise
as torch does.

The reference's per-interface helpers (speed_of_sound, get_wave_speeds,
get_max_wave_speeds, _euler_flux, _hll_solver, fluid.py:105-305) are fused
into the step kernels; they are also here as stand-alone elementwise kernels,
as are its boundary helpers (_set_along_axis, _open_left_boundary,
_open_right_boundary, fluid.py:317-379).
"""

from __future__ import annotations

import ctypes

import torch

from . import _lib

# fluid.py:14-22
DENSITY_INDEX = 0
VELOCITY_X_INDEX = 1
VELOCITY_Y_INDEX = 2
VELOCITY_Z_INDEX = 3
PRESSURE_INDEX = 4
VAR_AXIS = 0
X_AXIS = 1
Y_AXIS = 2
Z_AXIS = 3


def _stream(device=None) -> ctypes.c_void_p:
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def _ptr(t: torch.Tensor) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr())


def check_state(state, name: str = "primitive_state") -> torch.Tensor:
    """The STATE_TYPE contract (fluid.py:10): a float32 (5, Nx, Ny, Nz) device
    tensor. Raises TypeError like the reference's jaxtyping/beartype checks."""
    if not isinstance(state, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor, got {type(state).__name__}")
    if state.dim() != 4 or state.shape[0] != 5:
        raise TypeError(f"{name} must have shape (5, Nx, Ny, Nz), got {tuple(state.shape)}")
    if state.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {state.dtype}")
    if state.device.type != "cuda":
        raise TypeError(f"{name} must be on a ROCm device (no CPU path), got {state.device}")
    if not state.is_contiguous():
        state = state.contiguous()
    return state


def _scalar(x, name) -> float:
    if isinstance(x, torch.Tensor):
        if x.numel() != 1:
            raise TypeError(f"{name} must be a scalar")
        return float(x.item())
    return float(x)


def _dt_tensor(dt, device) -> torch.Tensor:
    """dt as a float64 scalar on the device (a traced scalar in the reference)."""
    if isinstance(dt, torch.Tensor):
        if dt.numel() != 1:
            raise TypeError("dt must be a scalar")
        return dt.reshape(()).to(device=device, dtype=torch.float64)
    return torch.tensor(float(dt), dtype=torch.float64, device=device)


def workspace(nx: int, ny: int, nz: int, device) -> torch.Tensor:
    n = _lib.load().tf_workspace_bytes(nx, ny, nz)
    return torch.empty(n, dtype=torch.uint8, device=device)


def conserved_state_from_primitive(primitive_state, gamma):
    """fluid.py:28-67."""
    s = check_state(primitive_state)
    out = torch.empty_like(s)
    _lib.check(_lib.load().tf_conserved_state_from_primitive(
        _ptr(s), _ptr(out), s[0].numel(), _scalar(gamma, "gamma"), _stream(s.device)))
    return out


def primitive_state_from_conserved(conserved_state, gamma):
    """fluid.py:69-103."""
    s = check_state(conserved_state, "conserved_state")
    out = torch.empty_like(s)
    _lib.check(_lib.load().tf_primitive_state_from_conserved(
        _ptr(s), _ptr(out), s[0].numel(), _scalar(gamma, "gamma"), _stream(s.device)))
    return out


def _cfl_time_step(primitive_state, grid_spacing, gamma, C_CFL=0.4):
    """fluid.py:425-504. Returns dt as a 0-d float64 device tensor."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    dt = torch.empty((), dtype=torch.float64, device=s.device)
    ws = workspace(nx, ny, nz, s.device)
    _lib.check(_lib.load().tf_cfl_time_step(
        _ptr(s), nx, ny, nz, _scalar(grid_spacing, "grid_spacing"), _scalar(gamma, "gamma"),
        _scalar(C_CFL, "C_CFL"), _ptr(dt), None, _ptr(ws), _stream(s.device)))
    return dt


def max_signal_speed(primitive_state, gamma):
    """max over cells and axes of |v_d| + c (the denominator of the CFL step)."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    m = torch.empty((), dtype=torch.float32, device=s.device)
    ws = workspace(nx, ny, nz, s.device)
    _lib.check(_lib.load().tf_cfl_time_step(
        _ptr(s), nx, ny, nz, 1.0, _scalar(gamma, "gamma"), 0.4, None, _ptr(m), _ptr(ws), _stream(s.device)))
    return m


def _field(x, name):
    if not isinstance(x, torch.Tensor) or x.dtype != torch.float32 or x.device.type != "cuda":
        raise TypeError(f"{name} must be a float32 tensor on a ROCm device")
    return x.contiguous()


def _direction(flux_direction_index):
    if flux_direction_index not in (VELOCITY_X_INDEX, VELOCITY_Y_INDEX, VELOCITY_Z_INDEX):
        raise ValueError(f"flux_direction_index must be 1, 2 or 3, got {flux_direction_index}")
    return int(flux_direction_index)


def _pair(primitives_left, primitives_right):
    left = check_state(primitives_left, "primitives_left")
    right = check_state(primitives_right, "primitives_right")
    if left.shape != right.shape:
        raise TypeError(f"left and right states differ in shape: {tuple(left.shape)} vs {tuple(right.shape)}")
    return left, right


def speed_of_sound(rho, p, gamma):
    """fluid.py:105-117: sqrt(gamma p / rho), elementwise."""
    rho, p = _field(rho, "rho"), _field(p, "p")
    if rho.shape != p.shape:
        raise TypeError(f"rho and p differ in shape: {tuple(rho.shape)} vs {tuple(p.shape)}")
    c = torch.empty_like(rho)
    _lib.check(_lib.load().tf_speed_of_sound(_ptr(rho), _ptr(p), _ptr(c), rho.numel(), _scalar(gamma, "gamma"),
                                             _stream(rho.device)))
    return c


def get_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index):
    """fluid.py:119-160: (wave_speeds_left_minus, wave_speeds_right_plus)."""
    left, right = _pair(primitives_left, primitives_right)
    sl, sr = torch.empty_like(left[0]), torch.empty_like(left[0])
    _lib.check(_lib.load().tf_get_wave_speeds(_ptr(left), _ptr(right), left[0].numel(), _scalar(gamma, "gamma"),
                                              _direction(flux_direction_index), _ptr(sl), _ptr(sr),
                                              _stream(left.device)))
    return sl, sr


def get_max_wave_speeds(primitives_left, primitives_right, gamma, flux_direction_index):
    """fluid.py:162-195: max(max |S_R|, max |S_L|) as a 0-d float32 device tensor."""
    left, right = _pair(primitives_left, primitives_right)
    m = torch.empty((), dtype=torch.float32, device=left.device)
    ws = workspace(*left.shape[1:], left.device)
    _lib.check(_lib.load().tf_get_max_wave_speeds(_ptr(left), _ptr(right), left[0].numel(), _scalar(gamma, "gamma"),
                                                  _direction(flux_direction_index), _ptr(m), _ptr(ws),
                                                  _stream(left.device)))
    return m


def _euler_flux(primitive_state, gamma, flux_direction_index):
    """fluid.py:206-259."""
    s = check_state(primitive_state)
    out = torch.empty_like(s)
    _lib.check(_lib.load().tf_euler_flux(_ptr(s), _ptr(out), s[0].numel(), _scalar(gamma, "gamma"),
                                         _direction(flux_direction_index), _stream(s.device)))
    return out


def _hll_solver(primitives_left, primitives_right, gamma, flux_direction_index):
    """fluid.py:261-305."""
    left, right = _pair(primitives_left, primitives_right)
    out = torch.empty_like(left)
    _lib.check(_lib.load().tf_hll_solver(_ptr(left), _ptr(right), _ptr(out), left[0].numel(), _scalar(gamma, "gamma"),
                                         _direction(flux_direction_index), _stream(left.device)))
    return out


def _set_along_axis(primitive_state, axis, set_index, get_index):
    """fluid.py:317-331 (axis 0..3 of the state; negative indices count from the end)."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    out = torch.empty_like(s)
    _lib.check(_lib.load().tf_set_along_axis(_ptr(s), _ptr(out), nx, ny, nz, int(axis), int(set_index),
                                             int(get_index), _stream(s.device)))
    return out


def _open_right_boundary(primitive_state, num_ghost_cells, axis):
    """fluid.py:333-354: the last num_ghost_cells slices along axis take the
    value of the slice before them."""
    get_index = -num_ghost_cells - 1
    for set_index in range(-num_ghost_cells, 0):
        primitive_state = _set_along_axis(primitive_state, axis, set_index, get_index)
    return primitive_state


def _open_left_boundary(primitive_state, num_ghost_cells, axis):
    """fluid.py:356-379: the first num_ghost_cells slices along axis take the
    value of the slice after them."""
    for set_index in range(num_ghost_cells):
        primitive_state = _set_along_axis(primitive_state, axis, set_index, num_ghost_cells)
    return primitive_state


def _open_boundaries(primitive_state, num_ghost_cells=1):
    """fluid.py:317-413: open boundaries with num_ghost_cells ghost cells per
    side (the reference step uses one, :572)."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    out = torch.empty_like(s)
    _lib.check(_lib.load().tf_open_boundaries(_ptr(s), _ptr(out), nx, ny, nz, int(num_ghost_cells),
                                              _stream(s.device)))
    return out


def _evolve_state_along_axis(primitive_state, grid_spacing, dt, gamma, axis):
    """fluid.py:506-557: one split sweep along axis 1 (x), 2 (y) or 3 (z)."""
    s = check_state(primitive_state)
    if axis not in (X_AXIS, Y_AXIS, Z_AXIS):
        raise ValueError(f"axis must be 1, 2 or 3, got {axis}")
    _, nx, ny, nz = s.shape
    out = torch.empty_like(s)
    d = _dt_tensor(dt, s.device)
    _lib.check(_lib.load().tf_evolve_state_along_axis(
        _ptr(s), _ptr(out), nx, ny, nz, _scalar(grid_spacing, "grid_spacing"), _ptr(d), _scalar(gamma, "gamma"),
        int(axis), _stream(s.device)))
    return out


def _evolve_state(primitive_state, grid_spacing, dt, gamma, handle_boundaries=False):
    """fluid.py:560-578: [open boundaries] then X, Y, Z split sweeps, fused
    into one gfx950 kernel pass."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    out = torch.empty_like(s)
    d = _dt_tensor(dt, s.device)
    ws = workspace(nx, ny, nz, s.device)
    _lib.check(_lib.load().tf_evolve_state(
        _ptr(s), _ptr(out), nx, ny, nz, _scalar(grid_spacing, "grid_spacing"), _ptr(d), _scalar(gamma, "gamma"),
        int(bool(handle_boundaries)), None, _ptr(ws), _stream(s.device)))
    return out


def _evolve_state_split(primitive_state, grid_spacing, dt, gamma, handle_boundaries=False):
    """The same step as three separate per-axis sweep launches (the
    reference's structure, fluid.py:570-576); a cross-check of the fused kernel."""
    s = _open_boundaries(primitive_state) if handle_boundaries else check_state(primitive_state)
    for axis in (X_AXIS, Y_AXIS, Z_AXIS):
        s = _evolve_state_along_axis(s, grid_spacing, dt, gamma, axis)
    return s


def conserved_totals(primitive_state, gamma):
    """Totals of mass, momentum and energy, summed in float64 in a fixed
    order on the device. Returns a float64 tensor of 5 on the device."""
    s = check_state(primitive_state)
    _, nx, ny, nz = s.shape
    out = torch.empty(5, dtype=torch.float64, device=s.device)
    ws = workspace(nx, ny, nz, s.device)
    _lib.check(_lib.load().tf_conserved_totals(
        _ptr(s), nx, ny, nz, _scalar(gamma, "gamma"), _ptr(out), _ptr(ws), _stream(s.device)))
    return out
