"""Writes fft_vectors.npz: unnormalised forward and inverse 3-D DFTs computed
with numpy.fft (a library routine) for two small seeded complex grids.  The
oracle's own FFT is checked against these vectors."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(2024)
    out = {}
    for name, shape in (("a", (6, 5, 4)), ("b", (9, 10, 8))):
        x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
        out["in_" + name] = x
        out["fwd_" + name] = np.fft.fftn(x)
        out["inv_" + name] = np.fft.ifftn(x) * x.size
    np.savez_compressed(os.path.join(HERE, "fft_vectors.npz"), **out)


if __name__ == "__main__":
    main()
