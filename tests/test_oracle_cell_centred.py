"""Pins the oracle's cell-centred Neumann Poisson solve (no reference
counterpart: the reference's Neumann axes are node-centred DCT-I; this is the
256^3 cell-centred DCT-II axis of the stand-alone Poisson workload).

  * the 1-D pair against scipy fixtures (tests/golden/dct2.npz);
  * cos(pi k (j + 1/2) / N) are eigenvectors of the cell-centred operator
    with rows (-1, 1), (1, -2, 1), (1, -1) and eigenvalues 2 (cos(pi k/N) - 1);
  * the 3-D solve leaves a residual of the exact 7-point operator at
    rounding level and equals the dense pseudo-inverse solution on tiny
    grids (zero mean, the null space of the singular operator)."""
import os

import numpy as np
import pytest

import mif_oracle as mo

FIX = np.load(os.path.join(os.path.dirname(__file__), "golden", "dct2.npz"))
LENGTHS = [int(n) for n in FIX["lengths"]]


def rel(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("n", LENGTHS)
def test_pair_matches_scipy(n):
    x = FIX[f"x_{n}"]
    assert rel(mo.transform("dct2_direct", x), FIX[f"dct2_{n}"] / 2) < 1e-13
    assert rel(mo.transform("dct3_direct", x), FIX[f"dct3_{n}"]) < 1e-13
    # the DCT-III inverts the DCT-II up to n
    assert rel(mo.transform("dct3_direct", mo.transform("dct2_direct", x)), n * x) < 1e-13


def cell_operator(n, h=1.0):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, i] = -2
        if i > 0:
            A[i, i - 1] = 1
        if i < n - 1:
            A[i, i + 1] = 1
    A[0, 0] = -1
    A[n - 1, n - 1] = -1
    return A / (h * h)


def test_dct2_basis_diagonalises_cell_operator():
    n = 16
    A = cell_operator(n)
    j = np.arange(n)
    for k in range(n):
        v = np.cos(np.pi * k * (j + 0.5) / n)
        lam = 2 * (np.cos(np.pi * k / n) - 1)
        assert np.max(np.abs(A @ v - lam * v)) < 1e-12
        spike = np.zeros(n)
        spike[k] = n if k == 0 else n / 2
        assert np.max(np.abs(mo.transform("dct2_direct", v) - spike)) < 1e-12


def grid(Ng, per, size):
    return mo.Oracle(mo.make_grid(Ng, size, (0.0, -1.0, 0.0), 1.0, 1.0, 1, per))


def laplacian(b, o, cell, n):
    """The exact 7-point operator on a packed (nz, ny, nx) array: periodic
    axes wrap, cell-centred axes mirror the first/last cell, node Neumann
    axes reflect across the end node."""
    out = np.zeros_like(b)
    for a in range(3):
        ax = 2 - a
        h = o.g.h[a]
        if o.g.periodic[a]:
            lo, hi = np.roll(b, 1, axis=ax), np.roll(b, -1, axis=ax)
        else:
            idx = np.arange(n[a])
            if cell[a]:
                im, ip = np.maximum(idx - 1, 0), np.minimum(idx + 1, n[a] - 1)
            else:
                im = np.where(idx == 0, 1, idx - 1)
                ip = np.where(idx == n[a] - 1, n[a] - 2, idx + 1)
            lo, hi = np.take(b, im, axis=ax), np.take(b, ip, axis=ax)
        out += (lo - 2 * b + hi) / (h * h)
    return out


@pytest.mark.parametrize("Ng", [(9, 17, 9), (17, 33, 13)])
def test_solve_residual(Ng):
    per, cell = (1, 0, 1), (0, 1, 0)
    o = grid(Ng, per, (2 * np.pi, 2.0, np.pi))
    n = [Ng[0] - 1, Ng[1] - 1, Ng[2] - 1]
    rng = np.random.default_rng(5)
    b = rng.standard_normal((n[2], n[1], n[0]))
    b -= b.mean()  # uniform weights: the left null vector of the operator
    phi = b.copy()
    o.poisson_packed_cell(phi, cell)
    r = laplacian(phi, o, cell, n) - b
    assert np.linalg.norm(r) <= 1e-13 * np.linalg.norm(b) * 10
    assert abs(phi.mean()) < 1e-13 * np.max(np.abs(phi))


def test_solve_equals_dense_pseudo_inverse():
    Ng, per, cell = (5, 5, 4), (1, 0, 0), (0, 1, 0)
    o = grid(Ng, per, (1.0, 2.0, 1.5))
    n = [Ng[0] - 1, Ng[1] - 1, Ng[2]]  # periodic x, cell-centred y, node Neumann z
    size = n[0] * n[1] * n[2]
    A = np.zeros((size, size))
    for q in range(size):
        e = np.zeros(size)
        e[q] = 1
        A[:, q] = laplacian(e.reshape(n[2], n[1], n[0]), o, cell, n).ravel()
    rng = np.random.default_rng(9)
    b = rng.standard_normal(size)
    # weights of the range: uniform on x and y, trapezoid on the node Neumann z
    wz = np.ones(n[2])
    wz[[0, -1]] = 0.5
    w = np.kron(wz, np.ones(n[0] * n[1]))
    b -= (w @ b) / w.sum()
    phi = b.reshape(n[2], n[1], n[0]).copy()
    o.poisson_packed_cell(phi, cell)
    # A phi = b, and phi has no component along the null vector (constant)
    assert np.max(np.abs(A @ phi.ravel() - b)) < 1e-12 * np.max(np.abs(b)) * 10
    ref = np.linalg.lstsq(A, b, rcond=None)[0]
    ref -= ref.mean()
    got = phi.ravel() - phi.mean()
    assert np.max(np.abs(got - ref)) < 1e-11 * np.max(np.abs(ref))
