"""Pins the CPU oracle (oracle/np_tinyfluids.py and oracle/tf_oracle.c).

Each oracle function is checked against something other than its own
formula: hand-computed values, textbook identities (HLL consistency and
upwind limits), invariants (conservation, mirror symmetry, uniform-flow
advection), the exact Riemann solution of the Sod problem (Toro, "Riemann
Solvers and Numerical Methods for Fluid Dynamics", ch. 4) on every axis, an
independent index-clamp formulation of the open boundaries, bit-equality of
the vectorised NumPy and scalar C restatements, and the iteration counts the
reference's own runs recorded.
"""

import json
import math
import os

import numpy as np
import pytest

from oracle import c_oracle, np_tinyfluids as npo
from tinyfluids_amd import ics

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def rand_state(shape, seed, amp=0.5, dtype=np.float64):
    rng = np.random.default_rng(seed)
    s = np.empty((5,) + tuple(shape))
    s[0] = 1.0 + amp * rng.uniform(0, 1, shape)
    s[1:4] = amp * rng.uniform(-1, 1, (3,) + tuple(shape))
    s[4] = 1.0 + amp * rng.uniform(0, 1, shape)
    return s.astype(dtype)


def cell(rho, vx, vy, vz, p):
    return np.array([rho, vx, vy, vz, p], dtype=np.float64).reshape(5, 1, 1, 1)


# ---------------------------------------------------------------- closed forms
def test_conversions_hand_computed():
    w = cell(1.0, 1.0, 2.0, 2.0, 1.0)
    u = npo.conserved_state_from_primitive(w, 5 / 3)
    # E = p/(gamma-1) + rho |v|^2 / 2 = 1.5 + 4.5
    np.testing.assert_allclose(u.ravel(), [1.0, 1.0, 2.0, 2.0, 6.0], rtol=1e-15)
    w2 = cell(2.0, -0.5, 0.0, 3.0, 0.7)
    u2 = npo.conserved_state_from_primitive(w2, 1.4)
    np.testing.assert_allclose(u2.ravel(), [2.0, -1.0, 0.0, 6.0, 0.7 / 0.4 + 0.5 * 2.0 * 9.25], rtol=1e-15)
    back = npo.primitive_state_from_conserved(u2, 1.4)
    np.testing.assert_allclose(back.ravel(), w2.ravel(), rtol=1e-14)


def test_conversion_round_trip_f64():
    s = rand_state((6, 5, 7), 0)
    back = npo.primitive_state_from_conserved(npo.conserved_state_from_primitive(s, 5 / 3), 5 / 3)
    np.testing.assert_allclose(back, s, rtol=1e-14, atol=1e-15)


def test_speed_of_sound():
    assert npo.speed_of_sound(np.array(2.0), np.array(3.0), 1.4) == pytest.approx(math.sqrt(2.1), rel=1e-15)


def test_euler_flux_hand_computed():
    w = cell(1.0, 1.0, 2.0, 2.0, 1.0)  # E = 6
    fx = npo._euler_flux(w, 5 / 3, 1).ravel()
    np.testing.assert_allclose(fx, [1.0, 1.0 + 1.0, 2.0, 2.0, 1.0 * (6.0 + 1.0)], rtol=1e-15)
    fy = npo._euler_flux(w, 5 / 3, 2).ravel()
    np.testing.assert_allclose(fy, [2.0, 2.0, 4.0 + 1.0, 4.0, 2.0 * 7.0], rtol=1e-15)
    fz = npo._euler_flux(w, 5 / 3, 3).ravel()
    np.testing.assert_allclose(fz, [2.0, 2.0, 4.0, 4.0 + 1.0, 14.0], rtol=1e-15)


def test_davis_wave_speeds_hand_computed():
    wl = cell(1.0, 0.5, 0.0, 0.0, 1.0)
    wr = cell(0.125, 0.0, 0.0, 0.0, 0.1)
    sl, sr = npo.get_wave_speeds(wl, wr, 1.4, 1)
    cl, cr = math.sqrt(1.4), math.sqrt(1.4 * 0.1 / 0.125)
    assert sr.item() == pytest.approx(max(0.5 + cl, cr), rel=1e-15)
    assert sl.item() == pytest.approx(min(0.5 - cl, -cr), rel=1e-15)
    # the 0 bound: both states supersonic to the right
    sl2, sr2 = npo.get_wave_speeds(cell(1, 5, 0, 0, 1), cell(1, 6, 0, 0, 1), 1.4, 1)
    assert sl2.item() == 0.0 and sr2.item() == pytest.approx(6 + math.sqrt(1.4))


def test_hll_consistency_and_upwind_limits():
    rng = np.random.default_rng(1)
    for d in (1, 2, 3):
        w = rand_state((4, 3, 2), 2)
        f = npo._hll_solver(w, w, 5 / 3, d)
        np.testing.assert_allclose(f, npo._euler_flux(w, 5 / 3, d), rtol=1e-13, atol=1e-14)
        # supersonic to the right: u - c > 0 on both sides -> F = F_L
        wl = rand_state((3, 3, 3), 3)
        wr = rand_state((3, 3, 3), 4)
        wl[d] += 5.0
        wr[d] += 5.0
        np.testing.assert_allclose(npo._hll_solver(wl, wr, 1.4, d), npo._euler_flux(wl, 1.4, d), rtol=1e-13)
        # supersonic to the left -> F = F_R
        wl[d] -= 10.0
        wr[d] -= 10.0
        np.testing.assert_allclose(npo._hll_solver(wl, wr, 1.4, d), npo._euler_flux(wr, 1.4, d), rtol=1e-13)
    assert rng is not None


def test_cfl_closed_form_uniform():
    shape = (4, 5, 6)
    s = np.empty((5,) + shape)
    s[0], s[1], s[2], s[3], s[4] = 2.0, 0.3, -0.7, 0.1, 3.0
    dt = npo._cfl_time_step(s, 0.01, 1.4)
    assert dt == pytest.approx(0.4 * 0.01 / (0.7 + math.sqrt(1.4 * 3.0 / 2.0)), rel=1e-15)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_cfl_per_cell_identity_bitwise(dtype):
    """max over interfaces of max(|S_R|,|S_L|) == max over cells of max_d |v_d| + c."""
    s = rand_state((7, 6, 9), 5, amp=2.0, dtype=dtype)
    ref = npo._cfl_time_step(s, 0.02, 5 / 3)
    g = np.asarray(5 / 3).astype(dtype)[()]
    c = np.sqrt(g * s[4] / s[0])
    m = np.max(np.maximum(np.maximum(np.abs(s[1]) + c, np.abs(s[2]) + c), np.abs(s[3]) + c))
    dt = np.asarray(0.4).astype(dtype)[()] * np.asarray(0.02).astype(dtype)[()] / m
    assert dt == ref


def test_open_boundaries_equal_index_clamp():
    for shape in [(5, 6, 7), (2, 4, 3), (3, 2, 2)]:
        s = rand_state(shape, 6)
        out = npo._open_boundaries(s, 1)
        gx, gy, gz = (np.array([1 if n == 2 else min(max(i, 1), n - 2) for i in range(n)]) for n in shape)
        ref = s[:, gx][:, :, gy][:, :, :, gz]
        assert np.array_equal(out, ref)
        assert np.array_equal(c_oracle.open_boundaries(s), ref)


def test_open_boundaries_any_ghost_count_equal_index_map():
    """For g ghost cells the reference's sequential copies (left then right,
    x then y then z) equal the separable index map s(i) = max(j, g) with
    j = n-g-1 for i >= n-g, else i: brute force on small grids, down to
    n = g + 1 where the two sides overlap."""
    for g in range(0, 4):
        for shape in [(9, 8, 7), (g + 1, g + 2, 2 * g + 1), (2 * g + 2, g + 1, 6)]:
            s = rand_state(shape, 7 + g)
            out = npo._open_boundaries(s.copy(), g)
            idx = [np.array([max(n - g - 1 if i >= n - g else i, g) for i in range(n)]) for n in shape]
            ref = s[:, idx[0]][:, :, idx[1]][:, :, :, idx[2]]
            assert np.array_equal(out, ref), (g, shape)


# ---------------------------------------------------------------- invariants
def test_uniform_state_is_steady():
    s = np.empty((5, 4, 5, 6))
    s[0], s[1], s[2], s[3], s[4] = 1.3, 0.2, -0.1, 0.4, 0.9
    out = npo._evolve_state(s, 0.1, 0.01, 1.4)
    np.testing.assert_allclose(out, s, rtol=1e-14, atol=1e-15)


def test_periodic_conservation():
    s = rand_state((6, 7, 8), 7)
    tot0 = npo.conserved_totals(s, 5 / 3)
    out, _, _ = npo.run_steps(s, 0.1, 5 / 3, 4)
    np.testing.assert_allclose(npo.conserved_totals(out, 5 / 3), tot0, rtol=1e-13, atol=1e-12)


def test_uniform_velocity_advection_keeps_v_and_p():
    """Uniform v and p with varying density: HLL's mass, momentum and energy
    fluxes stay proportional, so v and p stay uniform (exactly, up to round-off)."""
    rng = np.random.default_rng(8)
    shape = (8, 6, 5)
    s = np.empty((5,) + shape)
    s[0] = 1.0 + rng.uniform(0, 1, shape)
    s[1], s[2], s[3], s[4] = 0.3, -0.2, 0.1, 1.0
    out, _, _ = npo.run_steps(s, 0.1, 1.4, 5)
    np.testing.assert_allclose(out[1:], s[1:], rtol=1e-12, atol=1e-12)
    assert np.abs(out[0] - s[0]).max() > 1e-3  # the density did move


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_mirror_symmetry(axis):
    s = rand_state((6, 5, 7), 9)
    m = np.flip(s, axis=axis).copy()
    m[axis] = -m[axis]
    out = npo._evolve_state(s, 0.1, 0.01, 5 / 3)
    outm = npo._evolve_state(m, 0.1, 0.01, 5 / 3)
    back = np.flip(outm, axis=axis).copy()
    back[axis] = -back[axis]
    np.testing.assert_allclose(back, out, rtol=1e-12, atol=1e-13)


def test_one_dimensional_reduction_is_axis_independent():
    """A profile along one axis evolves the same whichever axis carries it."""
    prof = rand_state((9, 1, 1), 10)
    res = []
    for axis in (1, 2, 3):
        shape = [1, 1, 1]
        shape[axis - 1] = 9
        s = np.empty((5,) + tuple(shape))
        for k in range(5):
            s[k] = prof[k].reshape(shape)
        s[1:4] = 0.0
        s[axis] = prof[1].reshape(shape)
        out, _, _ = npo.run_steps(s, 0.1, 1.4, 3)
        res.append(np.stack([out[0].ravel(), out[axis].ravel(), out[4].ravel()]))
    np.testing.assert_allclose(res[1], res[0], rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(res[2], res[0], rtol=1e-13, atol=1e-14)


# ---------------------------------------------------------------- exact Sod (Toro ch. 4)
def exact_riemann(wl, wr, gamma, xi):
    """Exact solution (rho, u, p) of the 1-D Riemann problem at x/t = xi."""
    rl, ul, pl = wl
    rr, ur, pr = wr
    cl, cr = math.sqrt(gamma * pl / rl), math.sqrt(gamma * pr / rr)
    g1, g2 = (gamma - 1) / (2 * gamma), (gamma + 1) / (2 * gamma)

    def f(p, rk, pk, ck):
        if p > pk:
            a, b = 2 / ((gamma + 1) * rk), (gamma - 1) / (gamma + 1) * pk
            q = math.sqrt(a / (p + b))
            return (p - pk) * q, q * (1 - 0.5 * (p - pk) / (p + b))
        return 2 * ck / (gamma - 1) * ((p / pk) ** g1 - 1), (p / pk) ** (-g2) / (rk * ck)

    p = max(1e-8, 0.5 * (pl + pr))
    for _ in range(100):
        fl, dl = f(p, rl, pl, cl)
        fr, dr = f(p, rr, pr, cr)
        pn = max(1e-10, p - (fl + fr + ur - ul) / (dl + dr))
        if abs(pn - p) < 1e-14 * p:
            p = pn
            break
        p = pn
    u = 0.5 * (ul + ur) + 0.5 * (f(p, rr, pr, cr)[0] - f(p, rl, pl, cl)[0])
    out = []
    for s in xi:
        if s <= u:  # left of the contact
            if p > pl:
                sh = ul - cl * math.sqrt(g2 * p / pl + g1)
                out.append(wl if s < sh else ((rl * (p / pl + (gamma - 1) / (gamma + 1))
                                               / ((gamma - 1) / (gamma + 1) * p / pl + 1)), u, p))
            else:
                csl = cl * (p / pl) ** g1
                if s < ul - cl:
                    out.append(wl)
                elif s > u - csl:
                    out.append((rl * (p / pl) ** (1 / gamma), u, p))
                else:
                    c = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * (ul - s))
                    uu = 2 / (gamma + 1) * (cl + (gamma - 1) / 2 * ul + s)
                    out.append((rl * (c / cl) ** (2 / (gamma - 1)), uu, pl * (c / cl) ** (1 / g1)))
        else:
            if p > pr:
                sh = ur + cr * math.sqrt(g2 * p / pr + g1)
                out.append(wr if s > sh else ((rr * (p / pr + (gamma - 1) / (gamma + 1))
                                               / ((gamma - 1) / (gamma + 1) * p / pr + 1)), u, p))
            else:
                csr = cr * (p / pr) ** g1
                if s > ur + cr:
                    out.append(wr)
                elif s < u + csr:
                    out.append((rr * (p / pr) ** (1 / gamma), u, p))
                else:
                    c = 2 / (gamma + 1) * (cr - (gamma - 1) / 2 * (ur - s))
                    uu = 2 / (gamma + 1) * (-cr + (gamma - 1) / 2 * ur + s)
                    out.append((rr * (c / cr) ** (2 / (gamma - 1)), uu, pr * (c / cr) ** (1 / g1)))
    return np.array(out)


def test_exact_riemann_solver_known_star_state():
    # Toro table 4.2, test 1 (Sod): p* = 0.30313, u* = 0.92745
    sol = exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, [0.92745])
    assert sol[0, 2] == pytest.approx(0.30313, rel=1e-4)


def sod_error(n, axis, t_end=0.1):
    """Periodic double Sod tube along `axis`; L1 density error against the
    exact solution around the discontinuity at x = 0.75."""
    shape = [1, 1, 1]
    shape[axis - 1] = n
    dx = 1.0 / n
    x = (np.arange(n) + 0.5) * dx
    inner = (x >= 0.25) & (x < 0.75)
    s = np.zeros((5,) + tuple(shape))
    s[0] = np.where(inner, 1.0, 0.125).reshape(shape)
    s[4] = np.where(inner, 1.0, 0.1).reshape(shape)
    t = 0.0
    while t < t_end:
        dt = npo._cfl_time_step(s, dx, 1.4)
        s = npo._evolve_state(s, dx, dt, 1.4)
        t += dt
    sel = (x >= 0.5) & (x < 1.0)
    ex = exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, (x[sel] - 0.75) / t)
    rho = s[0].ravel()[sel]
    u = s[axis].ravel()[sel]
    assert np.all(np.abs(np.delete(s[1:4], axis - 1, axis=0)) < 1e-12)
    return float(np.sum(np.abs(rho - ex[:, 0])) * dx), float(np.sum(np.abs(u - ex[:, 1])) * dx)


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_sod_converges_to_exact_solution(axis):
    e100 = sod_error(100, axis)
    e400 = sod_error(400, axis)
    assert e400[0] < 0.008 and e400[1] < 0.012, e400
    assert e400[0] < 0.6 * e100[0] and e400[1] < 0.6 * e100[1], (e100, e400)


# ---------------------------------------------------------------- two restatements agree
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("hb", [False, True])
def test_numpy_and_c_restatements_bit_equal(dtype, hb):
    s = rand_state((5, 6, 7), 11, dtype=dtype)
    a, na = npo.time_integration(s, 0.1, 0.05, 5 / 3, hb)
    b, nb, _, _ = c_oracle.time_integration(s, 0.1, 0.05, 5 / 3, hb)
    assert na == nb and na >= 2
    assert np.array_equal(a, b)
    for axis in (1, 2, 3):
        assert np.array_equal(npo._evolve_state_along_axis(s, 0.1, 0.01, 1.4, axis),
                              c_oracle.sweep(s, 0.1, 0.01, 1.4, axis))
    assert npo._cfl_time_step(s, 0.1, 1.4) == c_oracle.cfl_time_step(s, 0.1, 1.4)


def test_fp32_reference_precision_meets_tolerance_at_small_size():
    s0, dx = ics.blast((24, 24, 24), radius=0.25)
    a64, _, _ = c_oracle.run_steps(s0.astype(np.float64), dx, 5 / 3, 40)
    a32, _, _ = c_oracle.run_steps(s0, dx, 5 / 3, 40)
    assert np.all(npo.relative_l1(a32, a64, 5 / 3) < 1e-5)


# ---------------------------------------------------------------- reference-held known answers
def load_counts():
    with open(os.path.join(GOLDEN, "iteration_counts.json")) as f:
        return json.load(f)["counts"]


def test_iteration_count_fixture_matches_reference_results():
    path = "/root/reference/results/scaling_data.npz"
    if not os.path.exists(path):
        pytest.skip("reference checkout not present")
    d = np.load(path, allow_pickle=False)
    ratio = d["execution_times"].astype(np.float64) / d["execution_times_per_iteration"].astype(np.float64)
    counts = load_counts()["scaling_data.npz"]
    assert {int(n): int(round(r)) for n, r in zip(d["num_cells_list"], ratio)} == {int(k): v for k, v in
                                                                                   counts.items()}


@pytest.mark.parametrize("n", [32, 64])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
@pytest.mark.parametrize("hb", [True, False])
def test_iteration_counts_known_answer(n, dtype, hb):
    counts = load_counts()
    expect = counts["scaling_data.npz"][str(n)]
    assert all(c[str(n)] == expect for c in counts.values())
    s, dx = ics.setup_ics(n, n // 16, dtype=dtype)
    _, num, _, _ = c_oracle.time_integration(s, dx, 0.2, 5 / 3, hb)
    assert num == expect


def test_iteration_count_known_answer_96():
    s, dx = ics.setup_ics(96, 6, dtype=np.float32)
    _, num, _, _ = c_oracle.time_integration(s, dx, 0.2, 5 / 3, True)
    assert num == load_counts()["scaling_data.npz"]["96"]


def test_golden_fixtures_reproduce():
    g = np.load(os.path.join(GOLDEN, "golden_steps.npz"), allow_pickle=False)
    names = sorted({k.split("__")[0] for k in g.files})
    assert len(names) >= 5
    for name in names:
        dx, gamma, steps, hb = g[f"{name}__params"]
        out, _, dts = c_oracle.run_steps(g[f"{name}__input"].astype(np.float64), dx, gamma, int(steps), bool(hb),
                                         record_dt=True)
        assert np.array_equal(out, g[f"{name}__output"]), name
        assert np.array_equal(dts, g[f"{name}__dts"]), name
