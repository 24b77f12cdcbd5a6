"""Writes the committed golden fixtures of the test suite.

- iteration_counts.json: the number of steps the reference's own runs took to
  reach t = 0.2 for setup_ics(N, N//16) (scaling_jax.py:23-41, 84-93), which
  is the exact ratio execution_times / execution_times_per_iteration stored
  in the reference's results/scaling_data*.npz. Read from the reference
  checkout when it is present (only when this script runs).
- golden_steps.npz: small inputs and the float64 oracle's outputs after a few
  steps (inputs are float32, as the device path takes them).

Run from the repository root: python tests/golden/make_golden.py
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import np_tinyfluids as npo  # noqa: E402
from tinyfluids_amd import ics  # noqa: E402

REFERENCE_RESULTS = "/root/reference/results"


def iteration_counts():
    out = {}
    for name in ("scaling_data.npz", "scaling_data_sharding.npz", "scaling_data_shard_mapped.npz"):
        d = np.load(os.path.join(REFERENCE_RESULTS, name), allow_pickle=False)
        ratio = d["execution_times"].astype(np.float64) / d["execution_times_per_iteration"].astype(np.float64)
        counts = {int(n): int(round(r)) for n, r in zip(d["num_cells_list"], ratio)}
        assert all(abs(r - round(r)) < 1e-3 for r in ratio), (name, ratio)
        out[name] = counts
    return out


def random_state(shape, seed):
    rng = np.random.default_rng(seed)
    s = np.empty((5,) + tuple(shape))
    s[0] = 1.0 + 0.5 * rng.uniform(0, 1, shape)
    s[1:4] = 0.5 * rng.uniform(-1, 1, (3,) + tuple(shape))
    s[4] = 1.0 + 0.5 * rng.uniform(0, 1, shape)
    return s.astype(np.float32)


def cases():
    """name -> (input float32, dx, gamma, steps, handle_boundaries)"""
    c = {}
    c["random_6x7x8_periodic"] = (random_state((6, 7, 8), 11), 0.1, 5 / 3, 3, False)
    c["random_6x7x8_open"] = (random_state((6, 7, 8), 12), 0.1, 5 / 3, 3, True)
    s, dx = ics.blast((12, 12, 12), radius=0.25)
    c["blast_12_periodic"] = (s, dx, ics.GAMMA_BLAST, 8, False)
    s, dx = ics.kelvin_helmholtz((12, 12, 4))
    c["kh_12x12x4_periodic"] = (s, dx, ics.GAMMA_KH, 8, False)
    s, dx = ics.setup_ics(12, 2)
    c["setup_ics_12_open"] = (s, dx, 5 / 3, 5, True)
    c["random_3x70x5_periodic"] = (random_state((3, 70, 5), 13), 0.05, 1.4, 2, False)
    return c


def main():
    counts = iteration_counts()
    with open(os.path.join(HERE, "iteration_counts.json"), "w") as f:
        json.dump({"source": "reference results/*.npz: execution_times / execution_times_per_iteration",
                   "t_final": 0.2, "gamma": 5 / 3, "ics": "setup_ics(N, N//16), grid_spacing 1/(N-1)",
                   "counts": counts}, f, indent=1, sort_keys=True)
    arrays = {}
    for name, (s, dx, gamma, steps, hb) in cases().items():
        out, t, dts = npo.run_steps(s.astype(np.float64), dx, gamma, steps, hb)
        arrays[f"{name}__input"] = s
        arrays[f"{name}__output"] = out
        arrays[f"{name}__dts"] = np.array(dts, dtype=np.float64)
        arrays[f"{name}__params"] = np.array([dx, gamma, steps, float(hb)], dtype=np.float64)
    np.savez_compressed(os.path.join(HERE, "golden_steps.npz"), **arrays)
    print("wrote", sorted(os.listdir(HERE)))


if __name__ == "__main__":
    main()
