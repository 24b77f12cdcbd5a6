"""Writes reference_check_figures.npz: the reference's own check figures
figures/check_{32,64,128,256,512}.png (written by scaling_jax.py:75-82 after its
GH200 runs of setup_ics(N, N // 16) to t = 0.2, gamma 5/3), as RGB uint8
arrays at half resolution (every second row and column). Run here, where
the reference checkout exists; the output is committed."""

import os
import sys

import matplotlib.image as mpimg
import numpy as np

REF = "/root/reference/figures"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_check_figures.npz")


def main():
    arrays = {}
    for n in (32, 64, 128, 256, 512):
        img = mpimg.imread(os.path.join(REF, f"check_{n}.png"))[..., :3]
        arrays[f"check_{n}"] = (img * 255.0).round().astype(np.uint8)[::2, ::2]
    np.savez_compressed(OUT, **arrays)
    print("wrote", OUT, {k: v.shape for k, v in arrays.items()}, file=sys.stderr)


if __name__ == "__main__":
    main()
