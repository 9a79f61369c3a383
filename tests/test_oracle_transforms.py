"""Pins the oracle's 1-D transforms (FFTW_REDFT00, FFTW_R2HC, FFTW_HC2R as the
reference uses them, src/PressureSolverStructures.cpp:23-41) against
library fixtures, O(N^2) direct sums and the inverse identities."""
import os

import numpy as np
import pytest

import mif_oracle as mo

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "transforms.npz"))
LENGTHS = [int(n) for n in FIX["lengths"]]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("name", ["redft00", "r2hc", "hc2r"])
def test_transform_matches_fixture(name, n):
    x = FIX[f"x_{n}"]
    assert rel(mo.transform(name, x), FIX[f"{name}_{n}"]) < 1e-13


@pytest.mark.parametrize("n", LENGTHS)
@pytest.mark.parametrize("name", ["redft00", "r2hc", "hc2r"])
def test_fast_equals_direct(name, n):
    x = FIX[f"x_{n}"]
    assert rel(mo.transform(name, x), mo.transform(name + "_direct", x)) < 1e-12


@pytest.mark.parametrize("n", LENGTHS)
def test_inverse_identities(n):
    x = FIX[f"x_{n}"]
    # HC2R(R2HC(x)) = n x ; REDFT00 is its own inverse up to 2(n-1)
    assert rel(mo.transform("hc2r", mo.transform("r2hc", x)), n * x) < 1e-13
    if n >= 2:
        assert rel(mo.transform("redft00", mo.transform("redft00", x)), 2 * (n - 1) * x) < 1e-13


def test_dct1_diagonalises_neumann_operator():
    # the DCT-I basis vectors cos(pi k j/(n-1)) are eigenvectors of the node
    # Neumann operator with rows (-2, 2), (1,-2,1), (2,-2) and eigenvalues
    # 2(cos(pi k/(n-1)) - 1)  (src/PressureSolverStructures.cpp:5-7)
    n = 17
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = -2
        if i > 0:
            A[i, i - 1] = 1
        if i < n - 1:
            A[i, i + 1] = 1
    A[0, 1] = 2
    A[n - 1, n - 2] = 2
    j = np.arange(n)
    for k in range(n):
        v = np.cos(np.pi * k * j / (n - 1))
        lam = 2 * (np.cos(np.pi * k / (n - 1)) - 1)
        assert np.max(np.abs(A @ v - lam * v)) < 1e-12
        # and the oracle's DCT-I of that vector is a single spike at k
        y = mo.transform("redft00", v)
        spike = np.zeros(n)
        spike[k] = (n - 1) * (2.0 if k in (0, n - 1) else 1.0)
        assert np.max(np.abs(y - spike)) < 1e-11
