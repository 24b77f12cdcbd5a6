"""The reference harness's check figure (scaling_jax.py:75-82): density and
pressure on the mid-plane z = N // 2 of the final state, side by side, in
matplotlib's default style. Rendered the same way, the reference's own
figures/check_{N}.png serve as golden outputs for the whole run."""

from __future__ import annotations

import io

import numpy as np


def plot_check(mid, path) -> None:
    """Writes the figure for mid = state[:, :, :, N // 2] ((5, N, N), primitive
    variables) to `path` (a file name or a binary file object), as PNG."""
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    fig, axs = plt.subplots(1, 2, figsize=(12, 6))
    axs[0].imshow(mid[0], extent=[0, 1, 0, 1])
    axs[0].set_title("Density")
    axs[1].imshow(mid[4], extent=[0, 1, 0, 1])
    axs[1].set_title("Pressure")
    fig.savefig(path, format="png")
    plt.close(fig)


def check_image(mid) -> np.ndarray:
    """The rendered figure as an (H, W, 3) uint8 RGB array."""
    import matplotlib.image as mpimg
    buf = io.BytesIO()
    plot_check(np.asarray(mid, dtype=np.float32), buf)
    buf.seek(0)
    return (mpimg.imread(buf)[..., :3] * 255.0).round().astype(np.uint8)
