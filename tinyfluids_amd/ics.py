"""Initial conditions for the benchmark workloads, built in float64 NumPy and
cast once, so the GPU path and the CPU oracle start from identical bits.

``setup_ics`` follows the reference harness scaling_jax.py:23-41 (blast cube,
dx = 1/(N-1)). The other generators describe the periodic workloads used to
measure this implementation: spherical blast, blast with seeded density
noise, Kelvin-Helmholtz shear layer, and the seeded uneven batch/serving mixes.
All grids have cubic cells of width ``dx`` and cell centres at (i + 1/2) dx.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

GAMMA_BLAST = 5.0 / 3.0
GAMMA_KH = 1.4


def _centres(n: int, dx: float) -> np.ndarray:
    return (np.arange(n, dtype=np.float64) + 0.5) * dx


def setup_ics(num_cells: int, num_injection_cells: int = 2, dtype=np.float32):
    """Blast cube of scaling_jax.py:23-41: rho=0.125, p=0.1 background;
    rho=1, p=1 on [c-k, c+k)^3 with c = N//2; dx = 1/(N-1)."""
    n = num_cells
    grid_spacing = 1 / (n - 1)
    state = np.zeros((5, n, n, n), dtype=np.float64)
    state[0] = 0.125
    state[4] = 0.1
    c = n // 2
    s = slice(c - num_injection_cells, c + num_injection_cells)
    state[0, s, s, s] = 1.0
    state[4, s, s, s] = 1.0
    return state.astype(dtype), grid_spacing


def blast_front_radius(state, grid_spacing: float, p_background: float = 0.1) -> float:
    """Radius of the pressure front of the setup_ics blast along +x through
    the centre of the injection cube: the outermost cell whose pressure
    exceeds the midpoint between the background and the peak on that ray.
    The cube spans cells [c - k, c + k) with c = N // 2, so its centre sits
    half a cell below cell c. At t = 0.2 the reference's figures show the
    front at r ~ 0.38 (scaling_jax.py:75-82, figures/check_{N}.png)."""
    c = state.shape[1] // 2
    return front_radius_of_ray(state[4, c:, c, c], grid_spacing, p_background)


def front_radius_of_ray(p_ray, grid_spacing: float, p_background: float = 0.1) -> float:
    """blast_front_radius from the pressure ray p[4, c:, c, c] alone."""
    p = np.asarray(p_ray, dtype=np.float64)
    threshold = p_background + 0.5 * (p.max() - p_background)
    i = int(np.nonzero(p > threshold)[0].max())
    return (i + 0.5) * grid_spacing


def blast(shape, dx=None, centre=(0.5, 0.5, 0.5), radius=0.1, p_in=10.0, p_out=1.0, rho=1.0,
          dtype=np.float32):
    """Spherical blast: rho, v=0, p_out background; p_in where |r - centre| < radius.
    Returns (state, dx)."""
    nx, ny, nz = shape
    if dx is None:
        dx = 1.0 / nx
    x = _centres(nx, dx)[:, None, None]
    y = _centres(ny, dx)[None, :, None]
    z = _centres(nz, dx)[None, None, :]
    # the distance test in float64; the state is written in the output dtype
    # directly (rho, p_in, p_out rounded once, as a float64 state cast would)
    inside = ((x - centre[0]) ** 2 + (y - centre[1]) ** 2) + (z - centre[2]) ** 2 < radius * radius
    state = np.zeros((5, nx, ny, nz), dtype=dtype)
    state[0] = np.asarray(rho, dtype=np.float64).astype(dtype)
    state[4] = np.where(inside, np.asarray(p_in, dtype=np.float64).astype(dtype),
                        np.asarray(p_out, dtype=np.float64).astype(dtype))
    return state, dx


def blast_noise(n: int, seed: int = 1, dtype=np.float32):
    """Blast of ``blast`` on an n^3 box of length 1 with
    rho = 1 + 0.01*U(-1,1) per cell, drawn with default_rng(seed) in float64."""
    state, dx = blast((n, n, n), dtype=np.float64)
    rng = np.random.default_rng(seed)
    state[0] = 1.0 + 0.01 * rng.uniform(-1.0, 1.0, (n, n, n))
    return state.astype(dtype), dx


def kelvin_helmholtz_plane(nx: int, ny: int, dx: float, dtype=np.float32):
    """The z-independent Kelvin-Helmholtz shear layer as an (5, nx, ny) plane:
    rho=2, vx=+0.5 for |y-0.5|<0.25, else rho=1, vx=-0.5; p=2.5;
    vy=0.01*sin(4*pi*x); vz=0."""
    x = _centres(nx, dx)[:, None]
    y = _centres(ny, dx)[None, :]
    inner = np.abs(y - 0.5) < 0.25
    plane = np.zeros((5, nx, ny), dtype=np.float64)
    plane[0] = np.where(inner, 2.0, 1.0) * np.ones_like(x)
    plane[1] = np.where(inner, 0.5, -0.5) * np.ones_like(x)
    plane[2] = 0.01 * np.sin(4.0 * math.pi * x) * np.ones_like(y)
    plane[4] = 2.5
    return plane.astype(dtype)


def kelvin_helmholtz(shape, dx=None, dtype=np.float32):
    """Full 3-D KH state (the plane broadcast along z). Returns (state, dx)."""
    nx, ny, nz = shape
    if dx is None:
        dx = 1.0 / ny
    plane = kelvin_helmholtz_plane(nx, ny, dx, dtype)
    return np.ascontiguousarray(np.broadcast_to(plane[:, :, :, None], (5, nx, ny, nz))), dx


@dataclass
class SimSpec:
    """One simulation of an uneven batch: grid, physics and its own step count."""
    shape: tuple
    gamma: float
    dx: float
    steps: int
    kind: str = "blast"
    centre: tuple = (0.5, 0.5, 0.5)
    radius: float = 0.1
    p_ratio: float = 10.0
    arrival: float = 0.0
    meta: dict = field(default_factory=dict)

    def initial_state(self, dtype=np.float32):
        if self.kind == "blast":
            return blast(self.shape, self.dx, self.centre, self.radius, self.p_ratio, 1.0, dtype=dtype)[0]
        if self.kind == "kh":
            return kelvin_helmholtz(self.shape, self.dx, dtype=dtype)[0]
        raise ValueError(f"unknown initial condition {self.kind!r}")

    @property
    def cells(self) -> int:
        return int(self.shape[0]) * int(self.shape[1]) * int(self.shape[2])


BATCH_SIDES = (32, 48, 64, 96, 128)


def uneven_batch(num_sims: int = 256, steps: int = 100, seed: int = 3, sides=BATCH_SIDES):
    """Seeded uneven batch of cubic blasts: side uniform over ``sides``,
    centre U[0.2,0.8]^3, radius U[0.05,0.15], p_in/p_out log-uniform in [2,100],
    gamma in {1.4, 5/3}; box length 1, so dx = 1/side."""
    rng = np.random.default_rng(seed)
    sims = []
    for _ in range(num_sims):
        n = int(rng.choice(sides))
        centre = tuple(float(c) for c in rng.uniform(0.2, 0.8, 3))
        radius = float(rng.uniform(0.05, 0.15))
        ratio = float(math.exp(rng.uniform(math.log(2.0), math.log(100.0))))
        gamma = float(rng.choice([GAMMA_KH, GAMMA_BLAST]))
        sims.append(SimSpec((n, n, n), gamma, 1.0 / n, steps, "blast", centre, radius, ratio))
    return sims


SERVE_SIDES = (32, 48, 64, 96, 128, 192, 256)


def serving_stream(num_jobs: int, rate: float, seed: int = 4, sides=SERVE_SIDES, max_aspect: int = 4,
                   steps_range=(50, 2000)):
    """Seeded Poisson job stream: inter-arrival Exp(rate) seconds; grid sides
    drawn from ``sides`` with aspect ratio up to ``max_aspect``:1 (the long side
    is capped at 256); steps log-uniform in ``steps_range``; half the jobs are
    blasts (as ``uneven_batch``), half Kelvin-Helmholtz shear layers."""
    rng = np.random.default_rng(seed)
    t = 0.0
    jobs = []
    for _ in range(num_jobs):
        t += float(rng.exponential(1.0 / rate))
        base = int(rng.choice(sides))
        aspect = int(rng.integers(1, max_aspect + 1))
        axis = int(rng.integers(0, 3))
        shape = [base, base, base]
        shape[axis] = min(base * aspect, 256)
        shape = tuple(shape)
        steps = int(round(math.exp(rng.uniform(math.log(steps_range[0]), math.log(steps_range[1])))))
        dx = 1.0 / min(shape)
        if rng.uniform() < 0.5:
            centre = tuple(float(c) for c in rng.uniform(0.2, 0.8, 3))
            radius = float(rng.uniform(0.05, 0.15))
            ratio = float(math.exp(rng.uniform(math.log(2.0), math.log(100.0))))
            gamma = float(rng.choice([GAMMA_KH, GAMMA_BLAST]))
            jobs.append(SimSpec(shape, gamma, dx, steps, "blast", centre, radius, ratio, arrival=t))
        else:
            jobs.append(SimSpec(shape, GAMMA_KH, dx, steps, "kh", arrival=t))
    return jobs
