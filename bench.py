"""Benchmark of the fused gfx950 time step.

Workload (one step = CFL dt + the X -> Y -> Z split HLL/Godunov update, the
body of time_integration.py:63-83): Kelvin-Helmholtz shear layer on a
1024^3 periodic float32 grid, gamma = 1.4, dx = 1/1024, 50 timed steps.
The state stays resident in HBM; each step is one fused stencil launch (dt
reduction fused into its epilogue) plus a one-thread dt update.

Prints one JSON line: cell-updates/s, the HBM roofline of the stencil kernel
(40 algorithmic bytes per cell-update: 5 float32 read + 5 written), and the
CPU restatement of the reference algorithm (C, float32, OpenMP) timed on a
bounded z-slab of the same workload.
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
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also copy the state buffer once (a plain copy kernel of known size, for HBM counters)")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU domain decomposition is not built "
                                   "(bench.py --gpus 1)"}))
        sys.exit(2)
    import torch

    import tinyfluids_amd as tf
    from tinyfluids_amd import ics
    from tinyfluids_amd.batch import Batch, Simulation
    from tinyfluids_amd import _lib

    nx, ny, nz = (int(v) for v in args.shape.split(","))
    dev = torch.device("cuda:0")
    dx = 1.0 / ny
    gamma = ics.GAMMA_KH
    plane = torch.from_numpy(ics.kelvin_helmholtz_plane(nx, ny, dx)).to(dev)
    buf0 = plane[:, :, :, None].expand(5, nx, ny, nz).contiguous()
    del plane
    buf1 = torch.empty_like(buf0)
    print(f"bench: KH {nx}x{ny}x{nz} fp32, state {buf0.numel() * 4 / 1e9:.2f} GB", file=sys.stderr, flush=True)

    lib = _lib.load()
    import ctypes
    tx, ly = ctypes.c_int(0), ctypes.c_int(0)
    tuned_ms = None
    if args.tile:
        tx.value, ly.value = (int(v) for v in args.tile.split(","))
        _lib.check(lib.tf_set_tiling(nx, ny, nz, tx.value, ly.value))
    elif not args.no_tune:
        ws = tf.fluid.workspace(nx, ny, nz, dev)
        ms = ctypes.c_float(0)
        _lib.check(lib.tf_autotune_evolve(buf0.data_ptr(), buf1.data_ptr(), nx, ny, nz, dx, gamma, ws.data_ptr(),
                                          ctypes.byref(tx), ctypes.byref(ly), ctypes.byref(ms),
                                          tf.fluid._stream(dev)))
        tuned_ms = ms.value
    else:
        _lib.check(lib.tf_get_tiling(nx, ny, nz, ctypes.byref(tx), ctypes.byref(ly)))
    print(f"bench: tiling tile_x={tx.value} tile_y={ly.value} (tuned {tuned_ms} ms)", file=sys.stderr, flush=True)

    if args.calibrate_copy:
        buf1.copy_(buf0)
        torch.cuda.synchronize()
        print(f"bench: calibration copy of {buf0.numel() * 4} bytes", file=sys.stderr, flush=True)
    sim = Simulation(buf0, dx, gamma, buffers=(buf0, buf1))
    batch = Batch([sim], tile_x=tx.value, tile_y=ly.value)
    sim.reset()
    batch.step(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    stencil_ms = batch.step_timed(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    st = batch.status()[0]
    assert st.n_iter == args.warmup + args.steps and not st.nonfinite, (st.n_iter, st.nonfinite)

    cells = nx * ny * nz
    value = cells * args.steps / el
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
                   "grid": [nx, ny, nz], "tile_x": tx.value, "tile_y": ly.value, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": f"tf::k_step_sims<{tx.value}>", "kernel_ms": kernel_s * 1e3,
                     "algorithmic_bytes_per_launch": BYTES_PER_CELL * cells},
        "sim_time": st.t,
    }
    del buf0, buf1, sim, batch
    torch.cuda.empty_cache()
    if not args.no_cpu:
        rec["cpu_baseline"] = cpu_baseline(nx, ny, args.cpu_seconds)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
