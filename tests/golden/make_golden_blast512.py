"""Writes tests/golden/blast512_noise.npz: the float64 oracle on the 512^3
periodic blast with seeded density noise (rho = 1 + 0.01*U(-1,1) from
default_rng(1); p = 10 inside r < 0.1 of the box centre, 1 outside; v = 0;
gamma 5/3; CFL 0.4; dx = 1/512), 200 steps.

The full state is 5.4 GB per float64 copy, so the fixture keeps
- every 32nd cell along each axis (16^3 cells x 5 primitive fields) at steps
  0, 50, 100, 150 and 200 (float64),
- the float64 totals of mass, momentum and energy every 10 steps,
- the float64 dt of every step;
and tests/golden/blast512_planes.npz the mid-planes through the blast centre,
one per axis (x = 256, y = 256, z = 256) at full resolution (512 x 512 cells x
5 primitive fields each), at steps 100 and 200. The planes are stored as
float32 roundings of the float64 oracle (relative rounding <= 6e-8 per value,
far below the 1e-5 relative L1 bound they are compared at), 22 MB.

Runs the scalar C oracle (OpenMP) for about half an hour on 8 cores.
Run from the repository root: python tests/golden/make_golden_blast512.py

With --float32-errors (after the above): runs the reference algorithm in
float32, the reference's own precision, from the same float32 initial state
and writes tests/golden/blast512_f32_errors.npz: its relative L1 per
conserved field against the float64 fixtures on the same subsample and
mid-planes (the scale of float32 round-off on those measures).
"""

from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import c_oracle  # noqa: E402
from tinyfluids_amd import ics  # noqa: E402

N = 512
STEPS = 200
SUB = 32
CHECK = (0, 50, 100, 150, 200)
PLANE_STEPS = (100, 200)
MID = N // 2
PLANE_STRIDE = 1


def mid_planes(state):
    """The three mid-planes of a (5, N, N, N) state, every PLANE_STRIDE-th cell in-plane."""
    k = PLANE_STRIDE
    return {"x": state[:, MID, ::k, ::k], "y": state[:, ::k, MID, ::k], "z": state[:, ::k, ::k, MID]}


def main():
    s32, dx = ics.blast_noise(N, seed=1)
    state = s32.astype(np.float64)
    del s32
    g = ics.GAMMA_BLAST
    sub = {}
    planes = {}
    totals = []
    dts = []
    t0 = time.time()
    for step in range(0, STEPS + 1, 10):
        if step % 10 == 0:
            totals.append(c_oracle.conserved_totals(state, g))
        if step in CHECK:
            sub[f"sub_{step}"] = state[:, ::SUB, ::SUB, ::SUB].copy()
        if step in PLANE_STEPS:
            for axis, plane in mid_planes(state).items():
                planes[f"{axis}_{step}"] = plane.astype(np.float32)
        if step == STEPS:
            break
        _, n, _, d = c_oracle.time_integration(state, dx, float("inf"), g, False, 10, 0.4, in_place=True,
                                               record_dt=True)
        assert n == 10
        dts.extend(d.tolist())
        print(f"step {step + 10}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "blast512_noise.npz"), totals=np.array(totals), dts=np.array(dts),
                        stride=np.array(SUB), checkpoints=np.array(CHECK), dx=np.array(dx), gamma=np.array(g),
                        **sub)
    np.savez_compressed(os.path.join(HERE, "blast512_planes.npz"), index=np.array(MID),
                        plane_stride=np.array(PLANE_STRIDE), steps=np.array(PLANE_STEPS), **planes)
    print("done", time.time() - t0)


def float32_errors():
    from oracle import np_tinyfluids as npo
    g = ics.GAMMA_BLAST
    fix = np.load(os.path.join(HERE, "blast512_noise.npz"), allow_pickle=False)
    planes = np.load(os.path.join(HERE, "blast512_planes.npz"), allow_pickle=False)
    state, dx = ics.blast_noise(N, seed=1)
    out = {}
    t0 = time.time()
    for step in range(0, STEPS + 1, 10):
        if step in CHECK:
            out[f"sub_{step}"] = npo.relative_l1(state[:, ::SUB, ::SUB, ::SUB], fix[f"sub_{step}"], g)
        if step in PLANE_STEPS:
            for axis, got in mid_planes(state).items():
                out[f"{axis}_{step}"] = npo.relative_l1(got, planes[f"{axis}_{step}"], g)
        if step == STEPS:
            break
        c_oracle.time_integration(state, dx, float("inf"), g, False, 10, 0.4, in_place=True)
        print(f"float32 step {step + 10}: {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(os.path.join(HERE, "blast512_f32_errors.npz"), **out)
    print("done", {k: v.tolist() for k, v in out.items()})


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "--float32-errors":
        float32_errors()
    else:
        main()
