"""Writes tests/golden/fft_golden.npz: inputs and numpy.fft.rfftn outputs (FP64) on small grids,
including radix-3 and radix-5 lengths.  Generated in this container with numpy; the committed
file is data only."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(20261202)
    out = {}
    for shape in [(8, 8, 8), (16, 8, 12), (12, 8, 16), (20, 10, 8), (6, 12, 18)]:
        key = "x".join(map(str, shape))
        f = rng.standard_normal(shape)
        out["in_" + key] = f
        out["out_" + key] = np.fft.rfftn(f)
    np.savez_compressed(os.path.join(HERE, "fft_golden.npz"), **out)


if __name__ == "__main__":
    main()
