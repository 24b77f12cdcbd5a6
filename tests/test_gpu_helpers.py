"""The reference's per-interface and boundary helpers as stand-alone kernels
(speed_of_sound, get_wave_speeds, get_max_wave_speeds, _euler_flux,
_hll_solver, _set_along_axis, _open_left_boundary, _open_right_boundary;
fluid.py:105-379) through the C ABI, against the float64 oracle, which
restates them op for op (oracle/np_tinyfluids.py).

Tolerances: float32 device arithmetic (one v_rcp_f32 per interface, a
Newton-refined 1/rho, v_sqrt_f32) against float64: 2e-6 relative for speeds,
and for fluxes 4e-6 of the field's largest magnitude (a flux component can
cancel to near zero); index work is bit-exact.
"""

import numpy as np
import pytest
import torch

from oracle import np_tinyfluids as npo

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
SHAPE = (7, 9, 70)


def tf():
    import tinyfluids_amd
    return tinyfluids_amd


def state(seed, shape=SHAPE, amp=0.5):
    rng = np.random.default_rng(seed)
    s = np.empty((5,) + shape)
    s[0] = 1.0 + amp * rng.uniform(0, 1, shape)
    s[1:4] = amp * rng.uniform(-1, 1, (3,) + shape)
    s[4] = 1.0 + amp * rng.uniform(0, 1, shape)
    return s.astype(np.float32)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


@pytest.mark.parametrize("gamma", [5 / 3, 1.4])
def test_speed_of_sound(gamma):
    s = state(1)
    c = tf().speed_of_sound(dev(s[0]), dev(s[4]), gamma).cpu().numpy()
    ref = npo.speed_of_sound(s[0].astype(np.float64), s[4].astype(np.float64), gamma)
    np.testing.assert_allclose(c, ref, rtol=2e-6, atol=0)


@pytest.mark.parametrize("d", [1, 2, 3])
def test_wave_speeds_and_their_maximum(d):
    L, R = state(2), state(3)
    sl, sr = tf().get_wave_speeds(dev(L), dev(R), 5 / 3, d)
    rl, rr = npo.get_wave_speeds(L.astype(np.float64), R.astype(np.float64), 5 / 3, d)
    np.testing.assert_allclose(sl.cpu().numpy(), rl, rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(sr.cpu().numpy(), rr, rtol=2e-6, atol=1e-7)
    assert torch.all(sr >= 0) and torch.all(sl <= 0)
    m = float(tf().get_max_wave_speeds(dev(L), dev(R), 5 / 3, d))
    ref = npo.get_max_wave_speeds(L.astype(np.float64), R.astype(np.float64), 5 / 3, d)
    assert abs(m - ref) <= 2e-6 * ref
    # the maximum is the exact maximum of the speeds the kernel gives
    assert m == max(float(sr.max()), float((-sl).max()))


def test_wave_speeds_propagate_nan():
    """jnp.maximum / jnp.minimum propagate NaN (fluid.py:157-158, 192-195)."""
    L, R = state(4), state(5)
    L[4, 1, 2, 3] = np.nan
    sl, sr = tf().get_wave_speeds(dev(L), dev(R), 1.4, 1)
    assert torch.isnan(sl[1, 2, 3]) and torch.isnan(sr[1, 2, 3])
    assert int(torch.isnan(sr).sum()) == 1
    assert np.isnan(float(tf().get_max_wave_speeds(dev(L), dev(R), 1.4, 1)))


@pytest.mark.parametrize("d", [1, 2, 3])
def test_euler_flux(d):
    s = state(6)
    got = tf()._euler_flux(dev(s), 5 / 3, d).cpu().numpy()
    ref = npo._euler_flux(s.astype(np.float64), 5 / 3, d)
    for f in range(5):
        assert np.abs(got[f] - ref[f]).max() <= 4e-6 * np.abs(ref[f]).max(), f


@pytest.mark.parametrize("d", [1, 2, 3])
def test_hll_solver(d):
    L, R = state(7), state(8)
    got = tf()._hll_solver(dev(L), dev(R), 1.4, d).cpu().numpy()
    ref = npo._hll_solver(L.astype(np.float64), R.astype(np.float64), 1.4, d)
    for f in range(5):
        assert np.abs(got[f] - ref[f]).max() <= 4e-6 * np.abs(ref[f]).max(), f
    # consistency: F(W, W) = F(W) (fluid.py:299-303 with U_R = U_L)
    same = tf()._hll_solver(dev(L), dev(L), 1.4, d)
    flux = tf()._euler_flux(dev(L), 1.4, d)
    assert torch.allclose(same, flux, rtol=4e-6, atol=1e-6)


@pytest.mark.parametrize("axis,set_index,get_index", [(0, 2, 4), (1, 0, 1), (1, -1, -2), (2, 3, 0), (3, -1, 5),
                                                      (3, 0, -1)])
def test_set_along_axis(axis, set_index, get_index):
    s = state(9)
    got = tf()._set_along_axis(dev(s), axis, set_index, get_index).cpu().numpy()
    assert np.array_equal(got, npo._set_along_axis(s, axis, set_index, get_index))


def test_set_along_axis_rejects_an_index_out_of_range():
    with pytest.raises(ValueError):
        tf()._set_along_axis(dev(state(9)), 1, 7, 0)
    with pytest.raises(ValueError):
        tf()._set_along_axis(dev(state(9)), 2, 0, -10)


@pytest.mark.parametrize("g", [1, 2, 3])
@pytest.mark.parametrize("axis", [1, 2, 3])
def test_open_left_and_right_boundaries(g, axis):
    s = state(10)
    left = tf()._open_left_boundary(dev(s), g, axis)
    right = tf()._open_right_boundary(dev(s), g, axis)
    assert np.array_equal(left.cpu().numpy(), npo._open_left_boundary(s, g, axis))
    assert np.array_equal(right.cpu().numpy(), npo._open_right_boundary(s, g, axis))


@pytest.mark.parametrize("g", [1, 2])
def test_open_boundaries_equal_their_left_right_sequence(g):
    """_open_boundaries (one kernel) equals the reference's sequence of left
    and right boundaries along x, y, z (fluid.py:382-413)."""
    s = dev(state(11))
    seq = s
    for axis in (1, 2, 3):
        seq = tf()._open_right_boundary(tf()._open_left_boundary(seq, g, axis), g, axis)
    assert torch.equal(tf()._open_boundaries(s, g), seq)
