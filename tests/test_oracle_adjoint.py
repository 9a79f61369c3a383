"""Pins the oracle's adjoint of timestep() (a builder extension with no
reference counterpart) by central finite differences of the forward map
(src/Timestep.cpp:97-156 restated in the oracle), on every boundary type."""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, grids

def forward(o, u0, p0, f, nsteps, record=None):
    o.set_body_force(f)
    vel = u0.copy()
    p = p0.copy()
    buf, buf2 = o.velocity(), o.velocity()
    pb = o.zeros(3)
    ev = mo.make_fn(mo.FN_ZERO)
    for s in range(nsteps):
        un = vel.copy()
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
        if record is not None:
            record.append((un, buf.copy(), buf2.copy()))
    return vel, p


def functional(vel, p, wu, wp):
    return sum(float((vel[c] * wu[c]).sum()) for c in range(3)) + float((p * wp).sum())


CASES = [
    ((9, 7, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),
    ((7, 6, 8), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0)),
    ((7, 8, 6), (2 * PI, 2 * PI, 1.0), (0.0, 0.0, 0.0), (1, 1, 0)),
    ((6, 7, 8), (2 * PI, 2 * PI, 2 * PI), (0.0, 0.0, 0.0), (1, 1, 1)),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "p" + "".join(map(str, c[3])))
def test_adjoint_matches_central_differences(case):
    Ng, size, origin, per = case
    RNG = np.random.default_rng(2026 + CASES.index(case))
    nsteps = 2
    og, _ = grids(Ng, size, origin, 50.0, 0.02, nsteps, per)
    o = mo.Oracle(og)
    u0 = o.velocity()
    for c in range(3):
        u0[c][...] = RNG.standard_normal(u0[c].shape)
    p0 = RNG.standard_normal(o.shape(3))
    f = RNG.standard_normal(3)
    wu = [RNG.standard_normal(u0[c].shape) for c in range(3)]
    wp = RNG.standard_normal(o.shape(3))
    rec = []
    forward(o, u0, p0, f, nsteps, rec)
    ubar = mo.Fields(*[w.copy() for w in wu])
    pbar = wp.copy()
    fbar = np.zeros(3)
    for un, y2, y3 in reversed(rec):
        o.timestep_adjoint(un, y2, y3, ubar, pbar, fbar)
    for trial in range(3):
        du = [RNG.standard_normal(u0[c].shape) for c in range(3)]
        dp = RNG.standard_normal(o.shape(3))
        df = RNG.standard_normal(3)
        eps = 1e-6
        up = mo.Fields(*[u0[c] + eps * du[c] for c in range(3)])
        um = mo.Fields(*[u0[c] - eps * du[c] for c in range(3)])
        Jp = functional(*forward(o, up, p0 + eps * dp, f + eps * df, nsteps), wu, wp)
        Jm = functional(*forward(o, um, p0 - eps * dp, f - eps * df, nsteps), wu, wp)
        fd = (Jp - Jm) / (2 * eps)
        ad = sum(float((ubar[c] * du[c]).sum()) for c in range(3)) + float((pbar * dp).sum()) + float(fbar @ df)
        assert abs(ad - fd) <= 1e-7 * max(1.0, abs(fd)), (trial, ad, fd)
