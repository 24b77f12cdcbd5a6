"""GPU tests of the batched engine and the continuous-batching server:
a simulation's result is bit-identical alone (plain single-grid loop),
batched, graph-replayed, or served in a mix; batched results meet the
float64 oracle; the server honours its HBM budget."""

import math

import numpy as np
import pytest
import torch

from oracle import c_oracle, np_tinyfluids as npo
from tinyfluids_amd import ics

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def specs_mixed():
    return [
        ics.SimSpec((16, 16, 16), ics.GAMMA_BLAST, 1 / 16, 7, "blast", (0.4, 0.5, 0.6), 0.2, 30.0),
        ics.SimSpec((24, 20, 70), ics.GAMMA_KH, 1 / 20, 9, "kh"),
        ics.SimSpec((32, 32, 32), ics.GAMMA_KH, 1 / 32, 5, "blast", (0.3, 0.3, 0.7), 0.12, 5.0),
        ics.SimSpec((8, 40, 12), ics.GAMMA_BLAST, 1 / 40, 11, "blast", (0.5, 0.5, 0.5), 0.3, 50.0),
        ics.SimSpec((10, 12, 130), ics.GAMMA_BLAST, 1 / 130, 6, "blast", (0.5, 0.5, 0.5), 0.2, 3.0),
        ics.SimSpec((3, 5, 64), ics.GAMMA_KH, 1 / 64, 8, "kh"),
    ]


def test_batched_equals_solo_bitwise():
    from tinyfluids_amd import engine
    specs = specs_mixed()
    init = engine.upload(specs, DEV)
    b = engine.run_batched(specs, init, graph=True)
    p = engine.run_plain(specs, init)
    for i, (x, y, st) in enumerate(zip(b.states, p.states, b.statuses)):
        assert st.n_iter == specs[i].steps and not st.active
        assert torch.equal(x, y), f"sim {i} {specs[i].shape}"


@pytest.mark.parametrize("tile_x,tile_y", [(2, 0), (4, 3), (8, 16), (6, 0)])
def test_batched_tilings_bitwise(tile_x, tile_y):
    from tinyfluids_amd import engine
    specs = specs_mixed()
    init = engine.upload(specs, DEV)
    ref = engine.run_batched(specs, init, graph=False)
    b = engine.run_batched(specs, init, graph=True, tile_x=tile_x, tile_y=tile_y)
    for i, (x, y) in enumerate(zip(b.states, ref.states)):
        assert torch.equal(x, y), (i, tile_x, tile_y)


def test_graph_replays_match_eager():
    """The captured graphs are replayed several times, with eager work and
    synchronisation in between; every replay matches eager stepping. Step
    counts cover one short graph, 16-step chunks plus a remainder graph, and
    re-capture when the counts change."""
    from tinyfluids_amd.batch import Batch, Simulation
    specs = specs_mixed()[:3]
    a = [Simulation(s.initial_state(), s.dx, s.gamma, device=DEV) for s in specs]
    b = [Simulation(s.initial_state(), s.dx, s.gamma, device=DEV) for s in specs]
    ba, bb = Batch(a), Batch(b)
    for s in a + b:
        s.reset()
    done = 0
    for rep, steps in enumerate((4, 4, 37, 16, 5, 37)):
        ba.step_graph(steps)
        bb.step(steps)
        done += steps
        torch.cuda.synchronize()
        _ = torch.ones(1000, device=DEV).sum().item()  # eager work between replays
        sa, sb = ba.status(), bb.status()
        for i in range(len(specs)):
            assert sa[i].n_iter == sb[i].n_iter == done
            assert sa[i].t == sb[i].t
            assert torch.equal(a[i].state(sa[i]), b[i].state(sb[i])), (rep, i)


def test_async_plans_and_status_copies_match_synchronous_ones():
    """Plans built without waiting (tf_batch_create_async), while the stream
    still runs another plan's steps, and statuses read in two halves
    (tf_batch_status_enqueue / _unpack, with further steps queued behind the
    copy) equal the synchronous ones: the simulations are switched between
    plans of different members mid-run, as a server does."""
    from tinyfluids_amd.batch import Batch, Simulation
    specs = specs_mixed()
    a = [Simulation(s.initial_state(), s.dx, s.gamma, max_steps=s.steps, device=DEV) for s in specs]
    b = [Simulation(s.initial_state(), s.dx, s.gamma, max_steps=s.steps, device=DEV) for s in specs]
    for s in a + b:
        s.reset()
    ref = Batch(a)
    stream = torch.cuda.Stream(DEV)  # default priority
    with torch.cuda.stream(stream):
        first = Batch(b[:4], asynchronous=True)
        first.step(3)
        copy = first.status_async()
        first.step(2)  # queued behind the copy: the copy sees 3 steps
        second = Batch(b[2:], asynchronous=True)  # built while `first` may still run
        second.step(4)
        copies = [second.status_async()]
        full = Batch(b, asynchronous=True)
        full.step(20)
        copies.append(full.status_async())
    st3 = copy.result()
    assert [x.n_iter for x in st3] == [min(3, s.steps) for s in specs[:4]]
    out = [c.result() for c in copies][-1]
    stream.synchronize()
    ref.step(25)
    want = ref.status()
    for i, s in enumerate(specs):
        assert out[i].n_iter == want[i].n_iter == s.steps and not out[i].active
        assert out[i].t == want[i].t
        assert torch.equal(b[i].state(out[i]), a[i].state(want[i])), i
    for plan in (first, second, full, ref):
        plan.close()


def test_whole_run_graph_matches_eager():
    """A whole run as one captured graph (tf_batch_run_graph), cold and
    replayed after the states are restored, equals eager stepping."""
    from tinyfluids_amd import engine
    specs = specs_mixed()
    init = engine.upload(specs, DEV)
    ref = engine.run_batched(specs, init, graph=False)
    for warm in (False, True):
        b = engine.run_batched(specs, init, graph="whole", warm=warm)
        for i, (x, y, st) in enumerate(zip(b.states, ref.states, b.statuses)):
            assert st.n_iter == specs[i].steps and not st.active, (warm, i)
            assert torch.equal(x, y), (warm, i)


def test_batch_reset_equals_per_simulation_reset():
    """tf_batch_reset leaves every control block as tf_sim_reset does (t, dt,
    step count, flag, signal speed), also after steps were taken."""
    from tinyfluids_amd.batch import Batch, Simulation
    specs = specs_mixed()
    a = [Simulation(s.initial_state(), s.dx, s.gamma, max_steps=s.steps, device=DEV) for s in specs]
    b = [Simulation(s.initial_state(), s.dx, s.gamma, max_steps=s.steps, device=DEV) for s in specs]
    ba, bb = Batch(a), Batch(b)
    for s in a:
        s.reset()
    ba.step(3)
    for s in a:
        s.reset()
    bb.reset()
    bb.step(3)
    bb.reset()
    sa, sb = ba.status(), bb.status()
    for i in range(len(specs)):
        for f in ("t", "dt", "n_iter", "active", "nonfinite", "max_speed"):
            assert getattr(sa[i], f) == getattr(sb[i], f), (i, f)
        assert sa[i].n_iter == 0 and sa[i].active and sa[i].dt > 0
    ba.step(5)
    bb.step(5)
    sa, sb = ba.status(), bb.status()
    for i in range(len(specs)):
        assert torch.equal(a[i].state(sa[i]), b[i].state(sb[i])), i


def test_batched_meets_oracle():
    from tinyfluids_amd import engine
    spec = ics.SimSpec((40, 36, 44), ics.GAMMA_BLAST, 1 / 40, 20, "blast", (0.45, 0.5, 0.55), 0.15, 20.0)
    other = specs_mixed()[:2]
    specs = [spec] + other
    init = engine.upload(specs, DEV)
    res = engine.run_batched(specs, init)
    ref, t_ref, _ = c_oracle.run_steps(spec.initial_state().astype(np.float64), spec.dx, spec.gamma, spec.steps)
    err = npo.relative_l1(res.states[0].cpu().numpy(), ref, spec.gamma)
    assert np.all(err < 1e-5), err
    assert abs(res.statuses[0].t - t_ref) < 1e-6 * t_ref


def test_server_results_same_alone_and_in_mix():
    """Served jobs (three at once, then a trickle) equal the same jobs run
    alone, bit for bit, and the float64 oracle within 1e-5."""
    from tinyfluids_amd import engine
    specs = specs_mixed()
    for k, s in enumerate(specs):
        s.arrival = 0.0 if k < 3 else 0.002 * k  # three at once, then a trickle
    srv = engine.Server(hbm_budget_bytes=1 << 30, steps_per_tick=3, keep_states=True)
    rep = srv.serve(specs)
    alone = engine.serve_plain(specs, keep_states=True)
    assert rep.max_resident >= 2  # jobs really overlapped
    for a, b in zip(rep.jobs, alone.jobs):
        assert a.steps_done == a.spec.steps
        assert torch.equal(a.state, b.state), a.index
        assert a.finished >= a.arrival
        # and each served result meets the float64 oracle (5-11 steps)
        ref, t_ref, _ = c_oracle.run_steps(a.spec.initial_state().astype(np.float64), a.spec.dx, a.spec.gamma,
                                           a.spec.steps)
        err = npo.relative_l1(a.state.cpu().numpy(), ref, a.spec.gamma)
        assert np.all(err < 1e-5), (a.index, err)
        assert abs(a.t - t_ref) < 1e-6 * t_ref, (a.index, a.t, t_ref)


@pytest.mark.parametrize("steps_per_tick,depth", [(1, 3), (2, 2), (0, 1)])
def test_server_pipeline_churn_bitwise(steps_per_tick, depth):
    """The pipelined server under churn: ticks of one or two steps with two
    or three ticks queued per group, so plans are rebuilt while the old
    plans' ticks are still queued, finished jobs are released a tick later,
    and kept states are cloned behind queued ticks; also one tick queued at a
    time. Jobs of all three size groups arrive in bursts. Every result
    equals the job run alone, bit for bit, with the same step count and t,
    and the HBM budget holds."""
    from tinyfluids_amd import engine
    specs = specs_mixed() + [
        ics.SimSpec((128, 96, 96), ics.GAMMA_KH, 1 / 96, 5, "kh"),
        ics.SimSpec((256, 256, 128), ics.GAMMA_BLAST, 1 / 256, 4, "blast", (0.5, 0.4, 0.5), 0.15, 20.0),
        ics.SimSpec((20, 24, 28), ics.GAMMA_BLAST, 1 / 28, 13, "blast", (0.5, 0.5, 0.5), 0.25, 8.0),
        ics.SimSpec((64, 64, 64), ics.GAMMA_KH, 1 / 64, 3, "kh"),
    ]
    for k, s in enumerate(specs):
        s.arrival = 0.001 * (k // 3)  # bursts of three
    budget = 1 << 30
    srv = engine.Server(hbm_budget_bytes=budget, steps_per_tick=steps_per_tick, depth=depth, keep_states=True,
                        tick_ms=0.05)
    rep = srv.serve(specs)
    alone = engine.serve_plain(specs, keep_states=True)
    assert rep.peak_bytes <= budget
    assert {engine.Server.group_of(s) for s in specs} == {0, 1, 2}
    for a, b in zip(rep.jobs, alone.jobs):
        assert a.steps_done == a.spec.steps and not a.nonfinite, a.index
        assert a.t == b.t, a.index
        assert torch.equal(a.state, b.state), a.index


def test_server_respects_hbm_budget():
    from tinyfluids_amd import engine
    specs = [ics.SimSpec((32, 32, 32), ics.GAMMA_KH, 1 / 32, 6, "kh", arrival=0.0) for _ in range(5)]
    budget = 2 * engine.state_bytes(specs[0]) + 100
    rep = engine.Server(hbm_budget_bytes=budget, steps_per_tick=2).serve(specs)
    assert rep.max_resident <= 2
    assert rep.peak_bytes <= budget
    assert all(j.steps_done == 6 for j in rep.jobs)
    s = rep.summary()
    assert s["p99_latency_s"] >= s["p50_latency_s"] > 0


LINE_SHAPES = [((5, 6, 8), False), ((7, 5, 16), True), ((6, 4, 24), False), ((4, 9, 32), False),
               ((3, 5, 48), True), ((6, 3, 64), False), ((3, 4, 100), False), ((4, 3, 128), True),
               ((3, 3, 192), False), ((2, 3, 256), False), ((9, 2, 20), False),
               # several whole lines per workgroup across waves
               ((9, 4, 48), False), ((13, 3, 96), True), ((11, 3, 80), False), ((7, 3, 33), True),
               ((17, 2, 5), False),
               # 5..16 waves per z-line
               ((3, 4, 300), False), ((4, 3, 512), True), ((2, 5, 520), False), ((2, 4, 700), True),
               ((3, 3, 1024), False), ((2, 3, 1000), True)]


def halo_path_steps(x, dx, gamma, n, hb):
    """n steps through the single-step entry points: the CFL dt, then the
    fused step with the line tiling switched off (62-cell z-tiles)."""
    import tinyfluids_amd as tf
    import tinyfluids_amd._lib as L
    lib = L.load()
    lib.tf_set_line_mode(0)
    try:
        for _ in range(n):
            dt = tf._cfl_time_step(x, dx, gamma)
            x = tf._evolve_state(x, dx, dt, gamma, hb)
    finally:
        lib.tf_set_line_mode(1)
    return x


@pytest.mark.parametrize("line_mode", [1, 0])
@pytest.mark.parametrize("tile_x", [2, 4, 8])
def test_line_tiling_bitwise(line_mode, tile_x):
    """Batched plans with whole-z-line workgroups (W = 1..16 waves, several
    lines per workgroup where that leaves fewer lanes idle) give the plain
    loop's results bit for bit, and both equal the 62-cell z-tiling of the
    single-step path."""
    import tinyfluids_amd._lib as L
    from tinyfluids_amd.batch import Batch, Simulation
    from tinyfluids_amd.time_integration import run_steps
    rng = np.random.default_rng(21)
    sims, init = [], []
    for (shape, hb) in LINE_SHAPES:
        s = np.empty((5,) + shape)
        s[0] = 1.0 + 0.5 * rng.uniform(0, 1, shape)
        s[1:4] = 0.5 * rng.uniform(-1, 1, (3,) + shape)
        s[4] = 1.0 + 0.5 * rng.uniform(0, 1, shape)
        x = torch.from_numpy(s.astype(np.float32)).to(DEV)
        init.append((x, hb))
        sims.append(Simulation(x, 0.05, 1.4, max_steps=4, handle_boundaries=hb, device=DEV))
    lib = L.load()
    try:
        lib.tf_set_line_mode(line_mode)
        b = Batch(sims, tile_x=tile_x)
    finally:
        lib.tf_set_line_mode(1)
    for s in sims:
        s.reset()
    b.step_graph(4)
    st = b.status()
    for i, ((x, hb), sim) in enumerate(zip(init, sims)):
        ref, _, _ = run_steps(x, 0.05, 1.4, 4, handle_boundaries=hb)
        assert st[i].n_iter == 4
        assert torch.equal(sim.state(st[i]), ref), (LINE_SHAPES[i], line_mode, tile_x)
        if line_mode == 1 and tile_x == 4:
            assert torch.equal(halo_path_steps(x, 0.05, 1.4, 4, hb), ref), LINE_SHAPES[i]


def test_tune_batch_picks_a_measured_candidate():
    """The batch autotuner times every candidate tile width and y-chunk (the
    plan's default chunks and 8, 16, 32 rows) on the real batch and returns
    the fastest; the run with it matches the default plan."""
    from tinyfluids_amd import engine
    specs = specs_mixed()
    init = engine.upload(specs, DEV)
    tuned = engine.tune_batch(specs, init, candidates=(2, 4), steps=2)
    assert tuned["tile_x"] in (2, 4) and tuned["tile_y"] in (0, 8, 16, 32)
    assert set(tuned["candidates"]) == {f"{tx},{ty}" for tx in (2, 4) for ty in (0, 8, 16, 32)}
    assert all(v > 0 and math.isfinite(v) for v in tuned["candidates"].values())
    assert tuned["ms_per_step"] == min(tuned["candidates"].values())
    ref = engine.run_batched(specs, init, graph=False)
    b = engine.run_batched(specs, init, graph=True, tile_x=tuned["tile_x"], tile_y=tuned["tile_y"])
    for x, y in zip(b.states, ref.states):
        assert torch.equal(x, y)
