"""GPU parity of the HIP path against the float64 oracle (through the C ABI).

Tolerances: float32 device arithmetic against the float64 oracle, measured as
relative L1 per conserved field, sum|q32 - q64| / sum|q64| (absolute sum for
fields that are zero in the oracle):
  - one step / one sweep / conversions: 2e-6
  - full runs (100-200 steps through shocks): 1e-5
"""


import numpy as np
import pytest
import torch

from oracle import c_oracle, np_tinyfluids as npo
from tinyfluids_amd import ics

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
ONE_STEP_TOL = 2e-6
RUN_TOL = 1e-5


def tf():
    import tinyfluids_amd
    return tinyfluids_amd


def random_state(shape, seed, amp=0.5):
    rng = np.random.default_rng(seed)
    s = np.empty((5,) + tuple(shape))
    s[0] = 1.0 + amp * rng.uniform(0, 1, shape)
    s[1:4] = amp * rng.uniform(-1, 1, (3,) + tuple(shape))
    s[4] = 1.0 + amp * rng.uniform(0, 1, shape)
    return s.astype(np.float32)


def to_dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to(DEV)


def rel_l1(gpu_tensor, ref64, gamma):
    return npo.relative_l1(gpu_tensor.cpu().numpy(), ref64, gamma)


def test_library_is_native():
    import tinyfluids_amd._lib as L
    lib = L.load()
    assert lib.tf_abi_version() == 1
    assert torch.cuda.is_available()


@pytest.mark.parametrize("nfloats", [4, 1 << 20, 5 * 33 * 17 * 4 + 4])
def test_hbm_copy_copies_every_byte(nfloats):
    import ctypes
    import tinyfluids_amd._lib as L
    src = torch.randn(nfloats, device=DEV)
    dst = torch.full((nfloats + 4,), -7.0, device=DEV)
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(L.load().tf_hbm_copy(src.data_ptr(), dst.data_ptr(), nfloats * 4, stream))
    torch.cuda.synchronize()
    assert torch.equal(dst[:nfloats], src)
    assert torch.all(dst[nfloats:] == -7.0)


@pytest.mark.parametrize("gamma", [5 / 3, 1.4])
def test_conversions(gamma):
    s = random_state((6, 7, 9), 1)
    cons = tf().conserved_state_from_primitive(to_dev(s), gamma)
    ref = npo.conserved_state_from_primitive(s.astype(np.float64), gamma)
    np.testing.assert_allclose(cons.cpu().numpy(), ref, rtol=2e-6, atol=1e-7)
    back = tf().primitive_state_from_conserved(cons, gamma)
    np.testing.assert_allclose(back.cpu().numpy(), s, rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("gamma", [5 / 3, 1.4])
def test_primitive_from_conserved_against_oracle(gamma):
    """primitive_state_from_conserved (fluid.py:69-103) on one conserved
    input, against the float64 oracle on the same float32 values."""
    rng = np.random.default_rng(11)
    prim = random_state((7, 6, 65), 9)
    prim[1:4] *= 2.0  # kinetic energy up to about 60% of E
    cons = npo.conserved_state_from_primitive(prim.astype(np.float64), gamma).astype(np.float32)
    cons[1:4, 0, 0, :3] = 0.0  # zero momentum: v must be exactly 0
    cons[0] *= rng.uniform(1.0, 2.0, cons[0].shape).astype(np.float32)  # other rho, p stays > 0
    got = tf().primitive_state_from_conserved(to_dev(cons), gamma).cpu().numpy()
    ref = npo.primitive_state_from_conserved(cons.astype(np.float64), gamma)
    np.testing.assert_allclose(got[[0, 1, 2, 3]], ref[[0, 1, 2, 3]], rtol=2e-6, atol=0)
    assert np.all(got[1:4, 0, 0, :3] == 0.0)
    # p = (gamma-1)(E - m.v/2) loses bits to the subtraction: bound by E's scale
    np.testing.assert_allclose(got[4], ref[4], rtol=0, atol=4e-7 * np.abs(cons[4]).max())


@pytest.mark.parametrize("shape", [(8, 8, 8), (5, 9, 70), (33, 3, 130)])
def test_cfl_time_step(shape):
    s = random_state(shape, 2)
    dt = tf()._cfl_time_step(to_dev(s), 0.01, 5 / 3)
    ref = npo._cfl_time_step(s.astype(np.float64), 0.01, 5 / 3)
    assert abs(float(dt) - ref) <= 1e-6 * ref


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_sweep_along_axis(axis):
    s = random_state((9, 10, 11), 3)
    out = tf()._evolve_state_along_axis(to_dev(s), 0.05, 0.004, 5 / 3, axis)
    ref = npo._evolve_state_along_axis(s.astype(np.float64), 0.05, 0.004, 5 / 3, axis)
    assert np.all(rel_l1(out, ref, 5 / 3) < ONE_STEP_TOL)


@pytest.mark.parametrize("shape", [(4, 4, 4), (7, 9, 70), (16, 12, 62), (3, 17, 125), (9, 2, 5), (1, 6, 8)])
@pytest.mark.parametrize("hb", [False, True])
def test_fused_step(shape, hb):
    s = random_state(shape, 4)
    if hb and min(shape) < 2:
        # open boundaries read cell 1 of every axis (the get index
        # num_ghost_cells of _open_left_boundary, fluid.py:356-379), which an
        # axis of one cell does not have: rejected rather than made up
        with pytest.raises(ValueError, match="open boundaries"):
            tf()._evolve_state(to_dev(s), 0.05, 0.004, 5 / 3, hb)
        return
    print(f"fused step shape={shape} handle_boundaries={hb}", flush=True)
    out = tf()._evolve_state(to_dev(s), 0.05, 0.004, 5 / 3, hb)
    ref = npo._evolve_state(s.astype(np.float64), 0.05, 0.004, 5 / 3, hb)
    assert np.all(rel_l1(out, ref, 5 / 3) < ONE_STEP_TOL)


@pytest.mark.parametrize("shape", [(16, 16, 1), (16, 1, 16), (1, 16, 16), (1, 1, 64), (33, 1, 1), (1, 1, 1),
                                   (2, 3, 1), (70, 1, 3)])
def test_degenerate_axes(shape):
    """Grids with extent 1 (2-D and 1-D problems, a single cell): a periodic
    axis of one cell has zero flux difference, so the step equals the lower-
    dimensional one; single steps and a 20-step run follow the oracle."""
    s = random_state(shape, 8)
    print(f"degenerate shape={shape}", flush=True)
    out = tf()._evolve_state(to_dev(s), 0.05, 0.004, 5 / 3)
    ref = npo._evolve_state(s.astype(np.float64), 0.05, 0.004, 5 / 3)
    assert np.all(rel_l1(out, ref, 5 / 3) < ONE_STEP_TOL)
    run, t, _ = tf().run_steps(to_dev(s), 0.05, 5 / 3, 20)
    ref_run, t_ref, _ = c_oracle.run_steps(s.astype(np.float64), 0.05, 5 / 3, 20)
    assert np.all(rel_l1(run, ref_run, 5 / 3) < RUN_TOL)
    assert abs(t - t_ref) < 1e-5 * t_ref


@pytest.mark.parametrize("shape", [(12, 10, 70), (9, 8, 5), (6, 7, 200)])
@pytest.mark.parametrize("with_workspace", [True, False])
def test_single_step_reports_signal_speed_of_new_state(shape, with_workspace):
    """tf_evolve_state's max_speed_out: the float bits of max |v_d| + c of the
    state it wrote, equal to the CFL maximum of that state (whole-line path
    with a workspace, 62-cell path without); the output does not depend on
    whether the maximum is asked for."""
    import ctypes
    import tinyfluids_amd._lib as L
    from tinyfluids_amd.fluid import workspace
    lib = L.load()
    s = to_dev(random_state(shape, 9))
    dt = torch.tensor(0.003, dtype=torch.float64, device=DEV)
    ws = workspace(*shape, DEV) if with_workspace else None
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    bits = torch.zeros(1, dtype=torch.int32, device=DEV)
    out, out2 = torch.empty_like(s), torch.empty_like(s)
    L.check(lib.tf_evolve_state(s.data_ptr(), out.data_ptr(), *shape, 0.05, dt.data_ptr(), 1.4, 0, bits.data_ptr(),
                                ws.data_ptr() if ws is not None else None, st))
    L.check(lib.tf_evolve_state(s.data_ptr(), out2.data_ptr(), *shape, 0.05, dt.data_ptr(), 1.4, 0, None,
                                ws.data_ptr() if ws is not None else None, st))
    torch.cuda.synchronize()
    assert torch.equal(out, out2)
    got = bits.view(torch.float32).item()
    assert got == tf().max_signal_speed(out, 1.4).item()


@pytest.mark.parametrize("tx", [2, 4, 8])
def test_fused_tilings_agree_bitwise(tx):
    """Per-cell arithmetic does not depend on the tiling."""
    import tinyfluids_amd._lib as L
    s = to_dev(random_state((12, 40, 130), 5))
    lib = L.load()
    base = tf()._evolve_state(s, 0.05, 0.003, 1.4)
    for ly in (1, 7, 40):
        L.check(lib.tf_set_tiling(12, 40, 130, tx, ly))
        out = tf()._evolve_state(s, 0.05, 0.003, 1.4)
        assert torch.equal(out, base), (tx, ly)
    L.check(lib.tf_set_tiling(12, 40, 130, 4, 40))


def test_fused_matches_split_path():
    s = to_dev(random_state((10, 11, 12), 6))
    a = tf()._evolve_state(s, 0.05, 0.004, 5 / 3, True)
    b = tf()._evolve_state_split(s, 0.05, 0.004, 5 / 3, True)
    assert torch.allclose(a, b, rtol=2e-6, atol=1e-7)


def test_open_boundaries():
    s = random_state((5, 6, 7), 7)
    out = tf()._open_boundaries(to_dev(s))
    ref = c_oracle.open_boundaries(s)
    assert np.array_equal(out.cpu().numpy(), ref)


@pytest.mark.parametrize("g", [0, 1, 2, 3])
@pytest.mark.parametrize("shape", [(9, 8, 7), (4, 5, 6), (7, 4, 12)])
def test_open_boundaries_ghost_cells(g, shape):
    """Any ghost-cell count equals the reference's sequential left/right copies
    (restated in the NumPy oracle), down to dimensions of g + 1."""
    if min(shape) < g + 1:
        pytest.skip("dimension below g + 1")
    s = random_state(shape, 10 + g)
    out = tf()._open_boundaries(to_dev(s), g)
    ref = npo._open_boundaries(s.copy(), g)
    assert np.array_equal(out.cpu().numpy(), ref)


def test_blast_64_100_steps():
    """64^3 periodic spherical blast, gamma 5/3, CFL 0.4, 100 steps."""
    s0, dx = ics.blast((64, 64, 64))
    g = ics.GAMMA_BLAST
    out, t, nonfinite = tf().run_steps(to_dev(s0), dx, g, 100)
    ref, t_ref, _ = c_oracle.run_steps(s0.astype(np.float64), dx, g, 100)
    err = rel_l1(out, ref, g)
    print("blast64 rel L1", err, "t", t, t_ref, flush=True)
    assert not nonfinite
    assert np.all(err < RUN_TOL)
    assert abs(t - t_ref) < 1e-5 * t_ref
    tot0 = c_oracle.conserved_totals(s0, g)
    tot = tf().conserved_totals(out, g).cpu().numpy()
    # mass and energy: float32 round-off over 100 steps; momenta stay ~0
    assert abs(tot[0] - tot0[0]) < 1e-6 * tot0[0]
    assert abs(tot[4] - tot0[4]) < 1e-5 * tot0[4]
    assert np.all(np.abs(tot[1:4]) < 1e-6 * tot0[0])


@pytest.mark.parametrize("hb", [True, False])
@pytest.mark.parametrize("n,iters", [(32, 24), (64, 51)])
def test_iteration_counts(n, iters, hb):
    s, dx = ics.setup_ics(n, n // 16)
    _, num = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, hb, False)
    assert num == iters


@pytest.mark.parametrize("hb", [False, True])
@pytest.mark.parametrize("n", [96, 128, 196, 256])
def test_iteration_counts_reference_sizes(n, hb):
    """The reference's own step counts to t = 0.2 for setup_ics(N, N//16) at
    every size its results/*.npz hold beyond 64 (tests/golden/
    iteration_counts.json; the three files agree)."""
    import json
    import os
    with open(os.path.join(os.path.dirname(__file__), "golden", "iteration_counts.json")) as f:
        expected = json.load(f)["counts"]["scaling_data.npz"][str(n)]
    s, dx = ics.setup_ics(n, n // 16)
    _, num = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, hb, False)
    assert num == expected


def test_harness_problem_with_open_boundaries_meets_oracle():
    """setup_ics(96, 6) to t = 0.2 with open boundaries (79 steps): relative L1
    per conserved field within 1e-5 of the float64 oracle, same step count
    and final time."""
    s, dx = ics.setup_ics(96, 6)
    out, num = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, True, False)
    ref, n_ref, t_ref = c_oracle.time_integration(s.astype(np.float64), dx, 0.2, 5 / 3, True)[:3]
    assert num == n_ref == 79
    err = rel_l1(out, ref, 5 / 3)
    print("setup_ics(96) open boundaries rel L1", err, flush=True)
    assert np.all(err < RUN_TOL)


def test_harness_blast_front_at_reference_radius():
    """The reference harness problem at N = 128 (setup_ics(128, 8), gamma
    5/3, t = 0.2, 108 steps) on the GPU: the pressure front sits at
    r ~ 0.38, as in the reference's figures/check_{N}.png."""
    s, dx = ics.setup_ics(128, 8)
    out, num = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, False, False)
    assert num == 108
    assert 0.37 <= ics.blast_front_radius(out.cpu().numpy(), dx) <= 0.40


@pytest.mark.parametrize("n", [32, 64, 128, 256, 512])
def test_harness_run_renders_the_reference_check_figure(n):
    """The reference's own outputs: figures/check_{N}.png of its GH200 runs
    (scaling_jax.py:75-82; tests/golden/reference_check_figures.npz). The
    HIP run of the same problem, rendered the same way, matches them to
    within colormap quantisation (max 16/255 per channel, mean 0.3/255)."""
    import os
    from tinyfluids_amd.harness import check_image
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_check_figures.npz"))[f"check_{n}"]
    s, dx = ics.setup_ics(n, n // 16)
    out, _ = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, False, False)
    img = check_image(out[:, :, :, n // 2].cpu().numpy())[::2, ::2].astype(np.int16)
    d = np.abs(img - ref.astype(np.int16))
    assert d.max() <= 16 and d.mean() <= 0.3, (n, d.max(), d.mean())


def test_concurrent_callers_on_their_own_streams():
    """Threads calling the drop-in loop at once, each on its own torch stream
    and with shapes tuned concurrently on first use, get bit for bit what the
    same calls give one after another (the calls share no device state; the
    tiling registry is locked)."""
    import threading

    cases = [((40, 36, 44), 7, None), ((52, 48, 20), 8, None), ((40, 36, 44), 9, None), ((48, 48, 48), 0, 3)]

    def run(shape, seed, n):
        if n is None:
            out, _, _ = tf().run_steps(to_dev(random_state(shape, seed)), 1.0 / shape[1], 1.4, 25, seed % 2 == 1)
            return out, 25
        s0, dx = ics.setup_ics(shape[0], n)
        return tf().time_integration(to_dev(s0), dx, 0.2, 5 / 3, True, False)

    results = [None] * len(cases)
    errors = []

    def worker(k):
        try:
            st = torch.cuda.Stream(DEV)
            with torch.cuda.stream(st):
                out, num = run(*cases[k])
                results[k] = (out.cpu(), num)
        except Exception as e:  # reported below
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    for k, case in enumerate(cases):
        out, num = run(*case)
        assert num == results[k][1], case
        assert torch.equal(out.cpu(), results[k][0]), case


def test_nan_ends_loop_like_reference():
    s, dx = ics.blast((16, 16, 16))
    s[4, 3, 4, 5] = -1.0  # negative pressure: sound speed NaN, dt NaN
    _, num = tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, False, False)
    assert num == 1
    with pytest.raises(FloatingPointError):
        tf().time_integration(to_dev(s), dx, 0.2, 5 / 3, False, False, check_finite=True)


@pytest.mark.parametrize("shape,hb", [((13, 36, 40), False), ((9, 21, 70), True)])
def test_autotune_loop_is_read_only_and_tiling_free(shape, hb):
    """The loop tuner reads the state and writes only scratch; runs under the
    tuned tiling, the untuned default and forced tilings agree bit for bit,
    and a time_integration run ends on the oracle's step count."""
    import tinyfluids_amd._lib as L
    from tinyfluids_amd.time_integration import autotune, run_steps
    s = to_dev(random_state(shape, 31))
    keep = s.clone()
    tx, ty, ms = autotune(s, 0.05, 1.4, hb)
    assert torch.equal(s, keep)
    assert tx in (2, 4, 8) and 1 <= ty <= shape[1] and ms > 0
    ref, t_ref, _ = run_steps(s, 0.05, 1.4, 7, handle_boundaries=hb)
    for ftx, fty in ((2, 1), (4, 3), (8, shape[1])):
        L.check(L.load().tf_set_tiling(*shape, ftx, fty))
        out, t, _ = run_steps(s, 0.05, 1.4, 7, handle_boundaries=hb)
        assert torch.equal(out, ref), (ftx, fty)
        assert t == t_ref
    s64 = keep.cpu().numpy().astype(np.float64)
    _, n_ref = npo.time_integration(s64, 0.05, 0.02, 1.4, hb)
    _, n = tf().time_integration(s, 0.05, 0.02, 1.4, hb, False)
    assert n == n_ref
