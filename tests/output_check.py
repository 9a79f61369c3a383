"""Test-side restatement of the output writers at one rank
(src/VTKDatExport.cpp:115-313 writeVTK, :342-584 writeDat, :587-669
writeVTKFullMesh of the reference) over numpy fields in the reference layout
(arrays indexed [k, j, i]), and parsers of the files they produce.  Used to
check the C++ writers' files; the arithmetic of the interpolations is
pinned by the closed-form checks in test_gpu_io.py."""
from __future__ import annotations

import math
import struct

import numpy as np


class Geometry:
    """Node coordinates of one domain (src/Constants.cpp:58-120 at one rank)."""

    def __init__(self, Ng, size, origin, periodic):
        self.Ng = tuple(int(n) for n in Ng)
        self.periodic = tuple(bool(p) for p in periodic)
        self.min = tuple(float(o) for o in origin)
        self.h = tuple(float(s) / (n - 1) for s, n in zip(size, self.Ng))
        self.base = tuple(-1 if p else 0 for p in self.periodic)
        S = [n + (1 if p else 0) for n, p in zip(self.Ng, self.periodic)]
        self.ext = {"p": tuple(S),
                    "u": (S[0] + (0 if self.periodic[0] else 1), S[1], S[2]),
                    "v": (S[0], S[1] + 1, S[2]),
                    "w": (S[0], S[1], S[2] + 1)}

    def coord(self, a, i, stag=None):
        x = self.min[a] + (self.base[a] + i) * self.h[a]
        return x - self.h[a] / 2 if stag == a else x

    def field_from(self, f, comp):
        """Values of f(x, y, z) at the points of field comp ('u', 'v', 'w' or 'p')."""
        stag = {"u": 0, "v": 1, "w": 2, "p": None}[comp]
        e = self.ext[comp]
        x = np.array([self.coord(0, i, stag) for i in range(e[0])])
        y = np.array([self.coord(1, j, stag) for j in range(e[1])])
        z = np.array([self.coord(2, k, stag) for k in range(e[2])])
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return f(X, Y, Z)

    def ranges(self):
        N = self.Ng
        start = (1 if self.periodic[0] else 0, 0, 0)
        sg = tuple(b + s for b, s in zip(self.base, start))
        return N, start, sg


def pos_to_index(pos, mn, d):
    fi = (pos - mn) / d
    i1 = math.floor(fi)
    return int(i1), float(np.float32(1.0 - (fi - i1)))


def vtk_points(G: Geometry, u, v, w, p):
    N, sl, sg = G.ranges()
    el = [s + n for s, n in zip(sl, N)]
    eg = [s + n for s, n in zip(sg, N)]
    pts, vals = [], []

    def put(i, j, k):
        pts.append((G.coord(0, i), G.coord(1, j), G.coord(2, k)))
        vals.append(((u[k, j, i] + u[k, j, i + 1]) / 2, (v[k, j, i] + v[k, j + 1, i]) / 2,
                     (w[k, j, i] + w[k + 1, j, i]) / 2, p[k, j, i]))

    kg = pos_to_index(0.0, G.min[2], G.h[2])[0]
    if sg[2] <= kg < eg[2]:
        k = kg - sg[2] + sl[2]
        for i in range(sl[0], el[0]):
            for j in range(sl[1], el[1]):
                put(i, j, k)
    i = pos_to_index(0.0, G.min[0], G.h[0])[0] - sg[0] + sl[0]
    for j in range(sl[1], el[1]):
        for k in range(sl[2], el[2]):
            put(i, j, k)
    jg = pos_to_index(0.0, G.min[1], G.h[1])[0]
    if sg[1] <= jg < eg[1]:
        j = jg - sg[1] + sl[1]
        for i in range(sl[0], el[0]):
            for k in range(sl[2], el[2]):
                put(i, j, k)
    return np.array(pts), np.array(vals)


def dat_rows(G: Geometry, u, v, w, p, direction, x, y, z):
    N, sl, sg = G.ranges()
    el = [s + n for s, n in zip(sl, N)]
    eg = [s + n for s, n in zip(sg, N)]
    ig, wi = pos_to_index(x, G.min[0], G.h[0])
    jg, wj = pos_to_index(y, G.min[1], G.h[1])
    kg, wk = pos_to_index(z, G.min[2], G.h[2])
    ax, ay, az = (abs(q - 1.0) < 1e-6 for q in (wi, wj, wk))
    U = lambda i, j, k: u[k, j, i]  # noqa: E731
    V = lambda i, j, k: v[k, j, i]  # noqa: E731
    W = lambda i, j, k: w[k, j, i]  # noqa: E731
    P = lambda i, j, k: p[k, j, i]  # noqa: E731
    rows = []
    if direction == 0 and sg[1] <= jg < eg[1] and sg[2] <= kg < eg[2]:
        j, k = jg - sg[1] + sl[1], kg - sg[2] + sl[2]
        for i in range(sl[0], el[0]):
            uu = (wj * wk * (U(i, j, k) + U(i + 1, j, k)) / 2 + (1 - wj) * wk * (U(i, j + 1, k) + U(i + 1, j + 1, k)) / 2
                  + wj * (1 - wk) * (U(i, j, k + 1) + U(i + 1, j, k + 1)) / 2
                  + (1 - wj) * (1 - wk) * (U(i, j + 1, k + 1) + U(i + 1, j + 1, k + 1)) / 2)
            vv = (wk * (V(i, j, k) + V(i, j + 1, k)) / 2 + (1 - wk) * (V(i, j, k + 1) + V(i, j + 1, k + 1)) / 2
                  if ay else wk * V(i, j + 1, k) + (1 - wk) * V(i, j + 1, k + 1))
            ww = (wj * (W(i, j, k) + W(i, j, k + 1)) / 2 + (1 - wj) * (W(i, j + 1, k) + W(i, j + 1, k + 1)) / 2
                  if az else wj * W(i, j, k + 1) + (1 - wj) * W(i, j + 1, k + 1))
            pp = (wj * wk * P(i, j, k) + (1 - wj) * wk * P(i, j + 1, k) + wj * (1 - wk) * P(i, j, k + 1)
                  + (1 - wj) * (1 - wk) * P(i, j + 1, k + 1))
            rows.append((G.coord(0, i), y, z, uu, vv, ww, pp))
    elif direction == 1 and sg[0] <= ig < eg[0] and sg[2] <= kg < eg[2]:
        i, k = ig - sg[0] + sl[0], kg - sg[2] + sl[2]
        for j in range(sl[1], el[1]):
            # the reference reads v where u is meant in the unaligned-x branch
            uu = (wk * (U(i, j, k) + U(i + 1, j, k)) / 2 + (1 - wk) * (U(i, j, k + 1) + U(i + 1, j, k + 1)) / 2
                  if ax else wk * U(i + 1, j, k) + (1 - wk) * V(i + 1, j, k + 1))
            vv = (wi * wk * (V(i, j, k) + V(i, j + 1, k)) / 2 + (1 - wi) * wk * (V(i + 1, j, k) + V(i + 1, j + 1, k)) / 2
                  + wi * (1 - wk) * (V(i, j, k + 1) + V(i, j + 1, k + 1)) / 2
                  + (1 - wi) * (1 - wk) * (V(i + 1, j, k + 1) + V(i + 1, j + 1, k + 1)) / 2)
            ww = (wi * (W(i, j, k) + W(i, j, k + 1)) / 2 + (1 - wi) * (W(i + 1, j, k) + W(i + 1, j, k + 1)) / 2
                  if az else wi * W(i, j, k + 1) + (1 - wi) * W(i + 1, j, k + 1))
            pp = (wi * wk * P(i, j, k) + (1 - wi) * wk * P(i + 1, j, k) + wi * (1 - wk) * P(i, j, k + 1)
                  + (1 - wi) * (1 - wk) * P(i + 1, j, k + 1))
            rows.append((x, G.coord(1, j), z, uu, vv, ww, pp))
    elif direction == 2 and sg[0] <= ig < eg[0] and sg[1] <= jg < eg[1]:
        i, j = ig - sg[0] + sl[0], jg - sg[1] + sl[1]
        for k in range(sl[2], el[2]):
            uu = (wj * (U(i, j, k) + U(i + 1, j, k)) / 2 + (1 - wj) * (U(i, j + 1, k) + U(i + 1, j + 1, k)) / 2
                  if ax else wj * U(i + 1, j, k) + (1 - wj) * V(i + 1, j + 1, k))
            vv = (wi * (V(i, j, k) + V(i, j + 1, k)) / 2 + (1 - wi) * (V(i + 1, j, k) + V(i + 1, j + 1, k)) / 2
                  if ay else wi * V(i, j + 1, k) + (1 - wi) * V(i + 1, j + 1, k))
            ww = (wi * wj * (W(i, j, k) + W(i, j, k + 1)) / 2 + (1 - wi) * wj * (W(i + 1, j, k) + W(i + 1, j, k + 1)) / 2
                  + wi * (1 - wj) * (W(i, j + 1, k) + W(i, j + 1, k + 1)) / 2
                  + (1 - wi) * (1 - wj) * (W(i + 1, j + 1, k) + W(i + 1, j + 1, k + 1)) / 2)
            pp = (wi * wj * P(i, j, k) + (1 - wi) * wj * P(i + 1, j, k) + wi * (1 - wj) * P(i, j + 1, k)
                  + (1 - wi) * (1 - wj) * P(i + 1, j + 1, k))
            rows.append((x, y, G.coord(2, k), uu, vv, ww, pp))
    rows.sort(key=lambda r: r[direction])
    return np.array(rows)


def read_vtk_binary(path):
    """(points (n, 3), {name: values}) of a legacy binary VTK written by writeVTK."""
    data = open(path, "rb").read()
    head = b"# vtk DataFile Version 2.0\nvtk output\nBINARY\nDATASET UNSTRUCTURED_GRID \nPOINTS "
    assert data.startswith(head), data[:80]
    pos = len(head)
    eol = data.index(b"\n", pos)
    n_str, typ = data[pos:eol].split()
    n = int(n_str)
    assert typ == b"double"
    pos = eol + 1
    pts = np.frombuffer(data, dtype=">f8", count=3 * n, offset=pos).reshape(n, 3)
    pos += 24 * n
    out = {}
    blocks = [(b"\nPOINT_DATA %d\nSCALARS u double 1\nLOOKUP_TABLE default\n" % n, "u"),
              (b"\nSCALARS v double 1\nLOOKUP_TABLE default\n", "v"),
              (b"\nSCALARS w double 1\nLOOKUP_TABLE default\n", "w"),
              (b"\nSCALARS p double 1\nLOOKUP_TABLE default\n", "p")]
    for tag, name in blocks:
        assert data[pos:pos + len(tag)] == tag, (name, data[pos:pos + len(tag)])
        pos += len(tag)
        out[name] = np.frombuffer(data, dtype=">f8", count=n, offset=pos).astype(np.float64)
        pos += 8 * n
    assert pos == len(data)
    return pts.astype(np.float64), out


def read_vtk_ascii(path):
    """(dimensions, origin, spacing, {name: values}) of writeVTKFullMesh's file."""
    text = open(path).read()
    lines = text.split("\n")
    assert lines[0] == "# vtk DataFile Version 3.0" and lines[3] == "DATASET STRUCTURED_POINTS"
    dims = tuple(int(t) for t in lines[4].split()[1:])
    origin = tuple(float(t) for t in lines[5].split()[1:])
    spacing = tuple(float(t) for t in lines[6].split()[1:])
    n = int(lines[7].split()[1])
    rest = "\n".join(lines[8:])
    out = {}
    for name in ("u", "v", "w", "|u|", "p"):
        tag = f"SCALARS {name} double 1\nLOOKUP_TABLE default\n"
        start = rest.index(tag) + len(tag)
        vals = rest[start:].split("SCALARS")[0].split()
        out[name] = np.array([float(t) for t in vals[:n]])
        assert len(vals) >= n
    return dims, origin, spacing, out


def read_dat(path):
    return np.loadtxt(path, ndmin=2)


def struct_check():
    """Big-endian doubles round-trip (the byte order the writer uses)."""
    return struct.unpack(">d", struct.pack(">d", 1.5))[0] == 1.5
