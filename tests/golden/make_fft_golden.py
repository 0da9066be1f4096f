"""Writes tests/golden/fft_golden.npz: seeded complex inputs and numpy.fft.fftn outputs for
small 3D shapes (radices 2, 3, 4, 5, 8 and mixed).  Run in the build container:
    python tests/golden/make_fft_golden.py
The fixture is plain numeric data (no pickles)."""
import os

import numpy as np

SHAPES = [(2, 2, 2), (4, 4, 4), (3, 5, 6), (8, 6, 4), (6, 12, 10), (24, 4, 8), (5, 9, 32), (48, 3, 2)]


def main():
    rng = np.random.default_rng(20260101)
    out = {}
    for k, (mz, my, mx) in enumerate(SHAPES):
        a = rng.standard_normal((mz, my, mx)) + 1j * rng.standard_normal((mz, my, mx))
        out[f"in_{k}"] = a
        out[f"fwd_{k}"] = np.fft.fftn(a)
        out[f"inv_{k}"] = np.fft.ifftn(a) * a.size  # unnormalised inverse
    out["shapes"] = np.array(SHAPES)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fft_golden.npz")
    np.savez_compressed(path, **out)
    print("wrote", path)


if __name__ == "__main__":
    main()
