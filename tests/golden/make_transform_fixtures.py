"""Write tests/golden/transforms.npz: known-answer vectors of the three 1-D
real transforms the Poisson solve is built on, computed by scipy/numpy (not
by the reference, which calls FFTW: src/PressureSolverStructures.cpp:23-41).

  redft00  FFTW_REDFT00 (DCT-I):  scipy.fft.dct(x, type=1)  (same definition,
           unnormalised: Y_k = x_0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi j k/(n-1)))
  r2hc     FFTW_R2HC: numpy.fft.rfft repacked into halfcomplex order
           r_0..r_{n/2}, i_{(n+1)/2-1}..i_1
  hc2r     FFTW_HC2R (unnormalised inverse): n * irfft of the unpacked spectrum

Inputs are N(0,1) from numpy's PCG64, seed 20261202 + n.
"""
import os

import numpy as np
import scipy.fft as sf

LENGTHS = [2, 3, 4, 5, 7, 8, 9, 16, 17, 23, 31, 33, 64, 65, 126, 127, 129, 256, 257]


def r2hc(x):
    n = len(x)
    r = np.fft.rfft(x)
    out = np.zeros(n)
    out[: n // 2 + 1] = r.real
    for k in range(1, (n + 1) // 2):
        out[n - k] = r[k].imag
    return out


def hc2r(h):
    n = len(h)
    spec = np.zeros(n // 2 + 1, dtype=complex)
    spec[0] = h[0]
    for k in range(1, (n + 1) // 2):
        spec[k] = h[k] + 1j * h[n - k]
    if n % 2 == 0:
        spec[n // 2] = h[n // 2]
    return np.fft.irfft(spec, n) * n


def main():
    data = {}
    for n in LENGTHS:
        x = np.random.default_rng(20261202 + n).standard_normal(n)
        data[f"x_{n}"] = x
        data[f"redft00_{n}"] = sf.dct(x, type=1)
        data[f"r2hc_{n}"] = r2hc(x)
        data[f"hc2r_{n}"] = hc2r(x)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "transforms.npz")
    np.savez(out, lengths=np.array(LENGTHS), **data)
    print("wrote", out)


if __name__ == "__main__":
    main()
