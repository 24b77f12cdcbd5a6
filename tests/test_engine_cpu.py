"""Host-side logic of the workloads (no GPU): workload generators and the
server's admission checks."""


import numpy as np
import pytest

from tinyfluids_amd import ics


def test_uneven_batch_is_seeded_and_in_range():
    a = ics.uneven_batch(64, seed=3)
    b = ics.uneven_batch(64, seed=3)
    assert [s.shape for s in a] == [s.shape for s in b]
    for s in a:
        assert s.shape[0] in ics.BATCH_SIDES and s.shape[0] == s.shape[1] == s.shape[2]
        assert 0.2 <= min(s.centre) and max(s.centre) <= 0.8
        assert 0.05 <= s.radius <= 0.15 and 2.0 <= s.p_ratio <= 100.0
        assert s.gamma in (ics.GAMMA_KH, ics.GAMMA_BLAST) and s.steps == 100
    assert len({s.shape for s in a}) > 1


def test_serving_stream_shapes_and_arrivals():
    jobs = ics.serving_stream(200, rate=10.0, seed=4)
    arr = np.array([j.arrival for j in jobs])
    assert np.all(np.diff(arr) > 0)
    assert 10.0 < 200 / arr[-1] * 2 and 200 / arr[-1] < 20.0  # mean rate near 10 jobs/s
    for j in jobs:
        assert 50 <= j.steps <= 2000
        assert max(j.shape) <= 256 and min(j.shape) >= 32
        assert max(j.shape) / min(j.shape) <= 4
        assert j.kind in ("blast", "kh")


def test_initial_states_are_valid():
    for s in ics.uneven_batch(8, seed=5)[:3] + ics.serving_stream(4, 1.0, seed=6):
        x = s.initial_state()
        assert x.shape == (5,) + s.shape and x.dtype == np.float32
        assert np.all(x[0] > 0) and np.all(x[4] > 0)


def test_server_rejects_jobs_larger_than_budget():
    torch = pytest.importorskip("torch")
    from tinyfluids_amd import engine
    big = ics.SimSpec((64, 64, 64), 1.4, 1 / 64, 10, "kh")
    srv = engine.Server(hbm_budget_bytes=engine.state_bytes(big) - 1, device=torch.device("cpu"))
    with pytest.raises(ValueError):
        srv.serve([big])


def test_serving_rejects_a_wrong_number_of_initial_states():
    torch = pytest.importorskip("torch")
    from tinyfluids_amd import engine
    specs = [ics.SimSpec((8, 8, 8), 1.4, 1 / 8, 5, "kh") for _ in range(3)]
    initial = [s.initial_state() for s in specs[:2]]
    with pytest.raises(ValueError, match="2 initial states for 3 jobs"):
        engine.Server(hbm_budget_bytes=1 << 30, device=torch.device("cpu")).serve(specs, initial=initial)
    with pytest.raises(ValueError, match="2 initial states for 3 jobs"):
        engine.serve_plain(specs, device=torch.device("cpu"), initial=initial)


def test_setup_ics_matches_reference_harness():
    s, dx = ics.setup_ics(32, 2)
    assert dx == 1 / 31
    assert s[0, 16, 16, 16] == 1.0 and s[0, 0, 0, 0] == np.float32(0.125)
    assert s[4, 14, 14, 14] == 1.0 and s[4, 13, 14, 14] == np.float32(0.1)
    assert int((s[0] == 1.0).sum()) == 4 ** 3


def test_blast_front_of_reference_harness_problem():
    """The oracle's run of the reference harness problem (setup_ics(64, 4),
    gamma 5/3, t = 0.2) puts the pressure front at r ~ 0.38-0.39, where the
    reference's own figures/check_{N}.png show it; the front helper reads a
    hand-built profile exactly."""
    from oracle import c_oracle
    s, dx = ics.setup_ics(64, 4)
    out, n = c_oracle.time_integration(s, dx, 0.2, 5 / 3, False)[:2]
    assert n == 51
    assert 0.37 <= ics.blast_front_radius(out, dx) <= 0.40
    probe = np.zeros((5, 10, 10, 10), dtype=np.float32)
    probe[4] = 0.1
    probe[4, 5:8, 5, 5] = [2.0, 1.0, 0.9]  # midpoint 0.1 + (2.0 - 0.1) / 2 = 1.05: only cell c is above
    assert ics.blast_front_radius(probe, 0.1) == 0.5 * 0.1
    probe[4, 5:8, 5, 5] = [2.0, 1.5, 1.2]  # cells c .. c+2 are above
    assert abs(ics.blast_front_radius(probe, 0.1) - 2.5 * 0.1) < 1e-12


def test_oracle_run_renders_the_reference_check_figures():
    """Golden outputs of the reference itself: its harness figures
    figures/check_{32,64}.png (scaling_jax.py:75-82; tests/golden/
    reference_check_figures.npz). The oracle's run of the same problem,
    rendered the same way, matches them to within colormap quantisation:
    max 16/255 per channel, mean 0.3/255 (measured 10 and 0.11 at N = 32,
    3 and 0.005 at N = 64). A wrong sign, a dropped sweep or one step more
    or less moves the blast and fails this."""
    import os
    from oracle import c_oracle
    from tinyfluids_amd.harness import check_image
    ref = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_check_figures.npz"))
    for n in (32, 64):
        s, dx = ics.setup_ics(n, n // 16)
        out = c_oracle.time_integration(s, dx, 0.2, 5 / 3, False)[0]
        img = check_image(out[:, :, :, n // 2])[::2, ::2].astype(np.int16)
        d = np.abs(img - ref[f"check_{n}"].astype(np.int16))
        assert d.max() <= 16 and d.mean() <= 0.3, (n, d.max(), d.mean())


def _jobs(spec_list):
    from tinyfluids_amd.engine import JobRecord
    return [JobRecord(i, s, s.arrival) for i, s in enumerate(spec_list)]


def _spec(shape, steps, arrival):
    return ics.SimSpec(tuple(shape), ics.GAMMA_BLAST, 1.0 / shape[0], steps, arrival=arrival)


def test_server_groups_by_size():
    from tinyfluids_amd.engine import Server
    assert Server.group_of(_spec((64, 64, 64), 10, 0.0)) == 0           # 2^18 cells
    assert Server.group_of(_spec((128, 128, 64), 10, 0.0)) == 1         # 2^20
    assert Server.group_of(_spec((256, 256, 256), 10, 0.0)) == 2        # 2^24
    assert Server.group_of(_spec((256, 128, 128), 10, 0.0)) == 1        # 2^22


def test_server_batch_choice_oldest_then_shortest_work():
    """The oldest job is always in the batch; the rest come shortest remaining
    work first (policy "shortest") or in arrival order ("fcfs"), until the
    batch holds fill_cells cells."""
    from tinyfluids_amd.engine import Server
    specs = [_spec((128, 128, 128), 1000, 0.0),   # oldest, 2M cells
             _spec((64, 64, 64), 2000, 0.1),      # 262k cells, long
             _spec((64, 64, 64), 10, 0.2),        # 262k cells, short
             _spec((128, 128, 128), 20, 0.3)]     # 2M cells, short-ish
    jobs = _jobs(specs)
    srv = Server(1 << 40, fill_cells=(1 << 21) + (1 << 18) + 1)
    chosen = srv._select([0, 1, 2, 3], jobs)
    # remaining work: job2 2.6M < job3 42M < job1 524M -> fill: 0 (2M), 2 (262k), then 3 overflows the fill
    assert chosen == [0, 2, 3]
    fcfs = Server(1 << 40, fill_cells=(1 << 21) + (1 << 18) + 1, policy="fcfs")
    assert fcfs._select([0, 1, 2, 3], jobs) == [0, 1, 2]
    jobs[0].steps_done = 999  # the oldest stays first whatever its remaining work
    assert srv._select([3, 2, 0], jobs)[0] == 0
    with pytest.raises(ValueError):
        Server(1 << 30, policy="random")


def test_server_tick_length_follows_step_cost():
    from tinyfluids_amd.engine import Server
    small = _jobs([_spec((32, 32, 32), 5000, 0.0)])
    big = _jobs([_spec((256, 256, 256), 5000, 0.0)])
    srv = Server(1 << 40, tick_ms=4.0)
    n_small, n_big = srv._tick_steps([0], small), srv._tick_steps([0], big)
    assert n_small > 4 * n_big >= 4          # small grids take many more steps per tick
    assert n_small <= 1024
    nearly_done = _jobs([_spec((32, 32, 32), 5000, 0.0)])
    nearly_done[0].steps_done = 4997
    assert srv._tick_steps([0], nearly_done) == 3  # no more steps than are left
    assert Server(1 << 40, steps_per_tick=7)._tick_steps([0], big) == 7


def test_server_next_tick_counts_the_queued_ones():
    """A group keeps ticks queued ahead of its status reads: the next tick's
    steps count those queued for each job (a job new to the plan is not held
    back by ticks it is not in), and a group whose queued ticks cover every
    job's steps queues no more."""
    from tinyfluids_amd.engine import Server
    srv = Server(1 << 40, tick_ms=4.0)
    jobs = _jobs([_spec((32, 32, 32), 100, 0.0), _spec((32, 32, 32), 40, 0.0)])
    assert srv._tick_steps([0], jobs, [(None, [0], 60)]) == 40
    assert srv._tick_steps([0], jobs, [(None, [0], 60), (None, [0], 40)]) == 0
    # job 1 joined the plan after the queued ticks of job 0 alone
    assert srv._tick_steps([0, 1], jobs, [(None, [0], 100)]) == 40
    assert Server(1 << 40, steps_per_tick=7)._tick_steps([0], jobs, [(None, [0], 100)]) == 0
