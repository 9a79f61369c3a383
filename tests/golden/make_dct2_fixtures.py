"""Write tests/golden/dct2.npz: known-answer vectors of the transform pair of
a cell-centred Neumann axis, computed by scipy (the reference has no such
axis; its Neumann axes are node-centred, FFTW_REDFT00).

  dct2   scipy.fft.dct(x, type=2): y_k = 2 sum_j x_j cos(pi k (2j + 1) / (2n))
  dct3   scipy.fft.dct(x, type=3): y_j = x_0 + 2 sum_{k>=1} x_k cos(pi k (2j + 1) / (2n))

Inputs are N(0,1) from numpy's PCG64, seed 20261203 + n.
"""
import os

import numpy as np
import scipy.fft as sf

LENGTHS = [2, 3, 4, 5, 8, 16, 17, 31, 64, 128, 256]


def main():
    data = {}
    for n in LENGTHS:
        x = np.random.default_rng(20261203 + n).standard_normal(n)
        data[f"x_{n}"] = x
        data[f"dct2_{n}"] = sf.dct(x, type=2)
        data[f"dct3_{n}"] = sf.dct(x, type=3)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dct2.npz")
    np.savez(out, lengths=np.array(LENGTHS), **data)
    print("wrote", out)


if __name__ == "__main__":
    main()
