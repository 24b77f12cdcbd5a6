"""Benchmark of the fused gfx950 time step.

Workload (one step = CFL dt + the X -> Y -> Z split HLL/Godunov update, the
body of time_integration.py:63-83): Kelvin-Helmholtz shear layer on a
1024^3 periodic float32 grid, gamma = 1.4, dx = 1/1024, 50 timed steps.
The state stays resident in HBM; each step is one fused stencil launch (dt
reduction fused into its epilogue) plus a one-thread dt update. The tiling
is the one the loop tuner measures fastest (for 1024^3: whole z-lines, 16
waves per workgroup).

Prints one JSON line: cell-updates/s, the HBM roofline of the stencil kernel
(40 algorithmic bytes per cell-update: 5 float32 read + 5 written) beside the
measured flat-copy ceiling of the box and the committed counter profiles,
and the CPU restatement of the reference algorithm (C, float32, OpenMP)
timed on a bounded z-slab of the same workload; further legs: the non-cubic
grid, the reference's own harness against its published GH200 times (with
its mid-plane check figures written to --figures), the uneven batch and the
served job stream against the plain path.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_CELL = 40


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--shape", type=str, default="1024,1024,1024")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="bound of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--tile", type=str, default="", help="force tile_x,tile_y (skips autotuning)")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--serve-rates", type=str, default="16,32,64,128", help="Poisson arrival rates (jobs/s) to sweep")
    ap.add_argument("--serve-fcfs", action="store_true", help="also serve with the arrival-order batch policy")
    ap.add_argument("--no-harness", action="store_true")
    ap.add_argument("--harness-sizes", type=str, default="32,64,96,128,196,256,384,512,768")
    ap.add_argument("--figures", type=str, default="figures",
                    help="directory for the harness's mid-plane density/pressure images ('' for none)")
    ap.add_argument("--no-shapes", action="store_true", help="skip the non-cubic 1536x512x512 roofline line")
    return ap.parse_args()


def cpu_baseline(nx, ny, budget_s):
    """C restatement of the reference algorithm (float32, OpenMP, all host
    threads) on a z-slab nx x ny x nz_s of the same Kelvin-Helmholtz workload;
    steps run until the time budget is used (at least one)."""
    import numpy as np

    from oracle import c_oracle
    from tinyfluids_amd import ics

    nz_s = 16
    state, dx = ics.kelvin_helmholtz((nx, ny, nz_s), dx=1.0 / ny)
    state = np.ascontiguousarray(state)
    c_oracle.lib()
    steps = 0
    t0 = time.perf_counter()
    while True:
        c_oracle.run_steps(state, dx, ics.GAMMA_KH, 1, in_place=True)
        steps += 1
        if time.perf_counter() - t0 >= budget_s:
            break
    el = time.perf_counter() - t0
    cells = nx * ny * nz_s
    return {
        "value": cells * steps / el,
        "unit": "cell-updates/s",
        "cores": c_oracle.num_threads(),
        "kind": "port",
        "sample": f"C float32 restatement of the reference step (oracle/tf_oracle.c), KH {nx}x{ny}x{nz_s} "
                  f"z-slab, {steps} steps in {el:.1f} s, OpenMP over all host threads; "
                  f"host {os.cpu_count()} logical CPUs",
    }


def cpu_numpy_64():
    """The NumPy restatement of the reference step (oracle/np_tinyfluids.py)
    on the 64^3 spherical blast (gamma 5/3, CFL 0.4), float32 (the reference's
    precision) and float64 (the oracle): 1 warm-up step, median of 3 timed
    steps. NumPy's elementwise work runs on one core in effect."""
    import numpy as np

    from oracle import np_tinyfluids as npo
    from tinyfluids_amd import ics

    out = {}
    s0, dx = ics.blast((64, 64, 64))
    for name, dt in (("f32", np.float32), ("f64", np.float64)):
        s = s0.astype(dt)
        s = npo.run_steps(s, dx, ics.GAMMA_BLAST, 1)[0]
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            s = npo.run_steps(s, dx, ics.GAMMA_BLAST, 1)[0]
            times.append(time.perf_counter() - t0)
        out[name] = 64 ** 3 / sorted(times)[1]
    return {"value_f32": out["f32"], "value_f64": out["f64"], "unit": "cell-updates/s", "cores": 1,
            "kind": "port", "sample": "NumPy restatement of the reference step, 64^3 blast, median of 3 steps"}


# Wall time of time_integration to t = 0.2 on setup_ics(N, N // 16) on one
# NVIDIA GH200, from the reference's own log (scaling_jax_57440.out:6-14).
GH200_SECONDS = {32: 0.001980, 64: 0.006428, 96: 0.024395, 128: 0.082908, 196: 0.559784, 256: 1.464719,
                 384: 7.514550, 512: 24.537891, 768: 128.716188}


def harness_leg(dev, sizes, figures=""):
    """The reference's own benchmark (scaling_jax.py:42-73, 86-93): the blast
    cube setup_ics(N, N // 16), gamma 5/3, dx 1/(N-1), time_integration to
    t = 0.2 with open boundaries, as the harness sets them (scaling_jax.py:
    234-236); one warm-up call (whose result gives the check figure and the
    blast-front radius), then the minimum of 3 timed
    calls, each returning after the run (block_until_ready)."""
    import torch

    import tinyfluids_amd as tf
    from tinyfluids_amd import ics
    from tinyfluids_amd.harness import plot_check

    out = []
    for n in sizes:
        s, dx = ics.setup_ics(n, n // 16)
        x = torch.from_numpy(s).to(dev)
        del s
        y, iters = tf.time_integration(x, dx, 0.2, 5 / 3, True, False)
        c = n // 2
        front = ics.front_radius_of_ray(y[4, c:, c, c].cpu().numpy(), dx)
        if figures:
            os.makedirs(figures, exist_ok=True)
            plot_check(y[:, :, :, c].cpu().numpy(), os.path.join(figures, f"check_{n:d}.png"))
        del y
        times = []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            y, it = tf.time_integration(x, dx, 0.2, 5 / 3, True, False)
            times.append(time.perf_counter() - t0)
            assert it == iters
            del y
        t = min(times)
        row = {"N": n, "iterations": iters, "seconds": t, "cell_updates_per_s": n ** 3 * iters / t,
               "blast_front_radius": front}
        if n in GH200_SECONDS:
            row["gh200_seconds"] = GH200_SECONDS[n]
            row["speedup_vs_gh200"] = GH200_SECONDS[n] / t
        out.append(row)
        del x
        torch.cuda.empty_cache()
        print(f"bench: harness {row}", file=sys.stderr, flush=True)
    return out


# The sources the stencil kernels are built from: a committed profile
# summary records their hash (source_hash:), and only summaries taken on the
# current sources are reported.
KERNEL_SOURCES = ("tinyfluids_amd/csrc/tf_kernels.hip", "tinyfluids_amd/csrc/tf_device.h",
                  "tinyfluids_amd/csrc/tf_layout.h", "tinyfluids_amd/csrc/tf_internal.h")


def kernel_source_hash(root=ROOT):
    """First 16 hex digits of the SHA-256 of the kernel sources, in order."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        with open(os.path.join(root, rel), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def profile_fields(text):
    """The `key: value` lines of a profile summary."""
    out = {}
    for line in text.splitlines():
        if ":" in line and not line.startswith((" ", "|", "#", "`")):
            k, v = line.split(":", 1)
            out[k.strip()] = v.strip()
    return out


def profile_paths(stem, root=ROOT):
    """Committed summaries of one kind: profiles/<stem>.md (the default
    tiling) and profiles/<stem>_tile*.md (other tilings the tuner picks)."""
    import glob
    return sorted(glob.glob(os.path.join(root, "profiles", stem + ".md"))) + \
        sorted(glob.glob(os.path.join(root, "profiles", stem + "_tile*.md")))


def current_profiles(stem, kernel, tile, root=ROOT):
    """(path, text) of the summaries of this kernel and tiling taken on the
    current kernel sources."""
    want = kernel_source_hash(root)
    out = []
    for path in profile_paths(stem, root):
        text = open(path).read()
        f = profile_fields(text)
        if f.get("kernel") == kernel and f.get("tile") == f"{tile[0]},{tile[1]}" and f.get("source_hash") == want:
            out.append((path, text))
    return out


def stale_profiles(root=ROOT):
    """Committed profile summaries taken on other kernel sources (or that
    record none): their figures describe another build."""
    import glob
    want = kernel_source_hash(root)
    return sorted(os.path.relpath(p, root) for p in glob.glob(os.path.join(root, "profiles", "*.md"))
                  if profile_fields(open(p).read()).get("source_hash") != want)


def measured_traffic(nx, ny, nz, kernel, tile, root=ROOT):
    """HBM bytes per launch of the fused step from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary (scaled by a plain copy's counter reading),
    when it was taken on this workload, kernel, tiling and kernel sources;
    else (None, None)."""
    if (nx, ny, nz) != (1024, 1024, 1024):
        return None, None
    for path, text in current_profiles("fused_step_kh1024_hbm_traffic", kernel, tile, root):
        return int(profile_fields(text)["traffic_bytes_per_launch"]), os.path.relpath(path, root)
    return None, None


def measured_trace(nx, ny, nz, kernel, tile, root=ROOT):
    """Average launch duration of the fused step in the committed
    rocprofv3 --kernel-trace --stats summary of this workload, kernel,
    tiling and kernel sources (to compare with kernel_ms, measured here with
    HIP events); else None."""
    import re
    if (nx, ny, nz) != (1024, 1024, 1024):
        return None
    for path, text in current_profiles("fused_step_kh1024_kernel_trace", kernel, tile, root):
        m = re.search(r"Average fused-step launch: \*\*([0-9.]+) ms\*\*", text)
        if m:
            return {"source": os.path.relpath(path, root), "avg_ms": float(m.group(1))}
    return None


def measured_valu(nx, ny, nz, kernel, tile, root=ROOT):
    """VALU instructions per cell-update and the VALU-busy fraction of the
    fused step from the committed rocprofv3 SQ-counter summary of this
    workload, kernel, tiling and kernel sources; else None."""
    if (nx, ny, nz) != (1024, 1024, 1024):
        return None
    for path, text in current_profiles("fused_step_kh1024_valu", kernel, tile, root):
        out = {"source": os.path.relpath(path, root)}
        for line in text.splitlines():
            cells = [c.strip() for c in line.strip().strip("|").split("|")]
            if len(cells) != 2:
                continue
            if cells[0].startswith("VALU instructions per cell-update"):
                out["insts_per_cell"] = float(cells[1])
            elif cells[0].startswith("SQ_ACTIVE_INST_VALU"):
                out["busy"] = float(cells[1])
        return out if len(out) == 3 else None
    return None


def batched_leg(dev):
    """256 uneven blasts, sides {32,48,64,96,128}^3, 100 steps each: all
    advanced together, against the plain path that runs them one after
    another through run_steps (tilings of every shape measured beforehand).
    Three launch forms of the batch: replays of a captured 16-step graph
    (capture inside the timed region), the whole run as one captured graph
    (capture inside), and that graph replayed once captured (warm)."""
    import torch

    from tinyfluids_amd import engine, ics

    specs = ics.uneven_batch(256, 100, seed=3)
    init = engine.upload(specs, dev)
    tuned = engine.tune_batch(specs, init)
    tx, ty = tuned["tile_x"], tuned["tile_y"]
    engine.run_batched(specs[:16], init[:16], keep_states=False, tile_x=tx)  # warm-up
    engine.run_plain(specs[:4], init[:4], keep_states=False)
    b = engine.run_batched(specs, init, graph=True, tile_x=tx, tile_y=ty)
    w = engine.run_batched(specs, init, graph="whole", tile_x=tx, tile_y=ty, keep_states=False)
    ww = engine.run_batched(specs, init, graph="whole", tile_x=tx, tile_y=ty, warm=True)
    p = engine.run_plain(specs, init)
    identical = all(torch.equal(x, y) and torch.equal(z, y) for x, y, z in zip(b.states, p.states, ww.states))
    cells = sum(s.cells * s.steps for s in specs)
    del init
    best = min(b.seconds, ww.seconds)
    return {
        "workload": "256 seeded blasts, sides {32,48,64,96,128}^3, 100 steps each, gamma in {1.4,5/3}",
        "sims": len(specs), "batched_s": best, "plain_s": p.seconds,
        "sims_per_s": len(specs) / best, "plain_sims_per_s": len(specs) / p.seconds,
        "speedup": p.seconds / best, "cell_updates_per_s": cells / best,
        "plain_cell_updates_per_s": cells / p.seconds, "bitwise_identical_to_plain": identical,
        "launch_forms": {
            "graph_16_step_replays_capture_timed": {"seconds": b.seconds, "sims_per_s": len(specs) / b.seconds},
            "one_graph_of_the_run_capture_timed": {"seconds": w.seconds, "sims_per_s": len(specs) / w.seconds},
            "one_graph_of_the_run_replayed": {"seconds": ww.seconds, "sims_per_s": len(specs) / ww.seconds},
        },
        "tile_x": tx, "autotune_ms_per_step": tuned["candidates"],
    }


def serve_leg(dev, njobs, rates, fcfs=False):
    """Poisson job stream (grids 32..256 per side, aspect up to 4:1, 50-2000
    steps, blast or shear layer) at each arrival rate: continuous batching
    against first come, first served one job at a time. The rates bracket
    the plain path's saturation, so the served throughput is measured where
    arrivals do not bound it. Each request carries its initial state (built
    on the host before the clock starts); both paths upload it when the job
    starts."""
    import numpy as np

    from tinyfluids_amd import engine, ics

    out = []
    # every request carries its initial state: built before the clock starts
    # (the same stream at every rate; only the arrival times scale)
    base = ics.serving_stream(njobs, 1.0, seed=4)
    initial = [s.initial_state() for s in base]
    for rate in rates:
        specs = ics.serving_stream(njobs, rate, seed=4)
        r = engine.Server(hbm_budget_bytes=64 << 30).serve(specs, initial=initial).summary()
        p = engine.serve_plain(specs, initial=initial).summary()
        row = {"rate_jobs_per_s": rate, "jobs": njobs,
               "arrival_span_s": specs[-1].arrival,
               "sims_per_s": r["sims_per_s"], "p50_latency_s": r["p50_latency_s"],
               "p99_latency_s": r["p99_latency_s"], "mean_latency_s": r["mean_latency_s"],
               "cell_updates_per_s": r["cell_updates_per_s"], "max_resident": r["max_resident"],
               "ticks": r["ticks"], "makespan_s": r["makespan_s"], "status_wait_s": r.get("status_wait_s"),
               "plan_rebuild_s": r.get("plan_rebuild_s"),
               "plain_sims_per_s": p["sims_per_s"], "plain_p50_latency_s": p["p50_latency_s"],
               "plain_p99_latency_s": p["p99_latency_s"], "plain_mean_latency_s": p["mean_latency_s"],
               "plain_cell_updates_per_s": p["cell_updates_per_s"],
               "speedup_sims_per_s": r["sims_per_s"] / p["sims_per_s"]}
        if fcfs:
            f = engine.Server(hbm_budget_bytes=64 << 30, policy="fcfs").serve(specs, initial=initial).summary()
            row["fcfs_batched"] = {"sims_per_s": f["sims_per_s"], "p50_latency_s": f["p50_latency_s"],
                                   "p99_latency_s": f["p99_latency_s"]}
        out.append(row)
        print(f"bench: serve rate {rate}: {row}", file=sys.stderr, flush=True)
    return out


def plan_kernel(tx, waves_per_line):
    """Name of the stencil kernel a plan runs for a simulation's tiling."""
    if waves_per_line == 0:
        return f"tf::k_step_sims<{tx}>"
    if waves_per_line <= 4:
        return f"tf::k_line_sims<{tx}, {waves_per_line}>"
    if waves_per_line <= 8:
        return f"tf::k_line_wide<{tx}, 8>"
    return "tf::k_line_wide<2, 16>"


def measure_grid(dev, nx, ny, nz, steps, warmup, tile, no_tune, copy_ceiling=False):
    """KH on nx x ny x nz resident in HBM: tiling (forced, measured by the
    loop tuner, or the default), `warmup` steps, then `steps` timed steps
    (wall clock and HIP events around the stencil launches). With
    copy_ceiling, also the rate of a flat copy of the state buffer
    (tf_hbm_copy: every byte read once and written once), the practical
    ceiling of a read+write kernel on this device; its four dispatches of
    known size also calibrate the HBM byte counters of a profiled run."""
    import ctypes

    import torch

    from tinyfluids_amd import _lib, ics
    from tinyfluids_amd.batch import Batch, Simulation
    from tinyfluids_amd.time_integration import autotune

    dx = 1.0 / ny
    gamma = ics.GAMMA_KH
    plane = torch.from_numpy(ics.kelvin_helmholtz_plane(nx, ny, dx)).to(dev)
    buf0 = plane[:, :, :, None].expand(5, nx, ny, nz).contiguous()
    del plane
    buf1 = torch.empty_like(buf0)
    print(f"bench: KH {nx}x{ny}x{nz} fp32, state {buf0.numel() * 4 / 1e9:.2f} GB", file=sys.stderr, flush=True)
    lib = _lib.load()
    tx = ty = 0
    tuned_ms = None
    if tile:
        tx, ty = (int(v) for v in tile.split(","))
        _lib.check(lib.tf_set_tiling(nx, ny, nz, tx, ty))
    elif not no_tune:
        _, _, tuned_ms = autotune(buf0, dx, gamma, False, buf1)
    copy_gbs = None
    if copy_ceiling:
        nbytes = buf0.numel() * 4
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        _lib.check(lib.tf_hbm_copy(buf0.data_ptr(), buf1.data_ptr(), nbytes, stream))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(3):
            _lib.check(lib.tf_hbm_copy(buf0.data_ptr(), buf1.data_ptr(), nbytes, stream))
        e1.record()
        torch.cuda.synchronize()
        copy_gbs = 2 * nbytes / (e0.elapsed_time(e1) / 3 / 1e3) / 1e9
        print(f"bench: flat copy of {nbytes} bytes: {copy_gbs:.0f} GB/s (read + write)", file=sys.stderr, flush=True)
    sim = Simulation(buf0, dx, gamma, buffers=(buf0, buf1))
    batch = Batch([sim], tile_x=tx, tile_y=ty)
    ptx, pty, pw = batch.tiling(0)
    print(f"bench: plan tile_x={ptx} tile_y={pty} waves/z-line={pw} (tuned {tuned_ms} ms)", file=sys.stderr,
          flush=True)
    sim.reset()
    batch.step(warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stencil_ms = batch.step_timed(steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = batch.status()[0]
    assert st.n_iter == warmup + steps and not st.nonfinite, (st.n_iter, st.nonfinite)
    batch.close()
    del buf0, buf1, sim, batch
    torch.cuda.empty_cache()
    return {"tile_x": ptx, "tile_y": pty, "waves_per_line": pw, "kernel": plan_kernel(ptx, pw), "status": st,
            "seconds": el, "stencil_ms": stencil_ms, "copy_gbs": copy_gbs}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU domain decomposition is not built "
                                   "(bench.py --gpus 1)"}))
        sys.exit(2)
    import torch

    nx, ny, nz = (int(v) for v in args.shape.split(","))
    dev = torch.device("cuda:0")
    m = measure_grid(dev, nx, ny, nz, args.steps, args.warmup, args.tile, args.no_tune, copy_ceiling=True)
    tx, ly, st, el, stencil_ms = m["tile_x"], m["tile_y"], m["status"], m["seconds"], m["stencil_ms"]
    kernel = m["kernel"]
    cells = nx * ny * nz
    value = cells * args.steps / el
    traffic, traffic_src = measured_traffic(nx, ny, nz, kernel, (tx, ly))
    kernel_s = stencil_ms / 1e3 / args.steps
    achieved = BYTES_PER_CELL * cells / kernel_s / 1e9
    rec = {
        "metric": "cell-updates/s (frac of HBM roofline)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Kelvin-Helmholtz shear layer {nx}x{ny}x{nz} periodic, gamma 1.4, dx 1/{ny}, "
                               f"CFL 0.4, fused split HLL/Godunov step with dt",
                   "grid": [nx, ny, nz], "tile_x": tx, "tile_y": ly, "waves_per_z_line": m["waves_per_line"],
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": kernel, "kernel_ms": kernel_s * 1e3,
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL * cells,
                     "copy_ceiling": m["copy_gbs"],
                     "valu": measured_valu(nx, ny, nz, kernel, (tx, ly)),
                     "kernel_trace": measured_trace(nx, ny, nz, kernel, (tx, ly)),
                     "kernel_source_hash": kernel_source_hash(),
                     "stale_profiles": stale_profiles(),
                     "frac_of_copy_ceiling": achieved / m["copy_gbs"] if m["copy_gbs"] else None},
        "sim_time": st.t,
    }
    if not args.no_shapes and (nx, ny, nz) == (1024, 1024, 1024):
        m2 = measure_grid(dev, 1536, 512, 512, 20, 2, "", args.no_tune)
        c2 = 1536 * 512 * 512
        k2 = m2["stencil_ms"] / 1e3 / 20
        rec["shapes"] = [{"grid": [1536, 512, 512], "workload": "Kelvin-Helmholtz, gamma 1.4, dx 1/512, 20 steps",
                          "value": c2 * 20 / m2["seconds"], "ms_per_step": m2["seconds"] * 1e3 / 20,
                          "kernel_ms": k2 * 1e3, "frac": BYTES_PER_CELL * c2 / k2 / 1e9 / HBM_PEAK_GBS,
                          "tile_x": m2["tile_x"], "tile_y": m2["tile_y"], "kernel": m2["kernel"]}]
        print(f"bench: shapes {rec['shapes']}", file=sys.stderr, flush=True)
    torch.cuda.empty_cache()
    if not args.no_harness:
        rec["reference_harness"] = harness_leg(dev, [int(v) for v in args.harness_sizes.split(",")], args.figures)
    if not args.no_batched:
        rec["batched"] = batched_leg(dev)
        print(f"bench: batched {rec['batched']}", file=sys.stderr, flush=True)
        torch.cuda.empty_cache()
    if not args.no_serve:
        rec["serving"] = serve_leg(dev, args.serve_jobs, [float(r) for r in args.serve_rates.split(",")],
                                   args.serve_fcfs)
        torch.cuda.empty_cache()
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(nx, ny, args.cpu_seconds)
        rec["cpu_numpy_64"] = cpu_numpy_64()
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
