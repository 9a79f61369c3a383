"""Derive the closed forms the oracle and the HIP library evaluate.

Restates, from the equations alone, the manufactured fields that the reference
generates at build time with sympy (generators/manufsol.py:31-72,
generators/manufsol_velocity.py:32-57, generators/manufsol_pressure.py:36-39):
the forcing of the velocity-only test is f = u_t + (u.grad)u - lap(u)/Re and
the Ethier-Steinman pressure gradient is grad p.  Running this script prints
the expressions that are written out by hand in oracle/mif_oracle.c and in the
device function table; it is a derivation aid, not part of any build.
"""
import sympy as sp

t, x, y, z, Re = sp.symbols("t x y z Re")


def lap(f):
    return sp.diff(f, x, 2) + sp.diff(f, y, 2) + sp.diff(f, z, 2)


def forcing(u, v, w):
    out = []
    for c in (u, v, w):
        f = sp.diff(c, t) + u * sp.diff(c, x) + v * sp.diff(c, y) + w * sp.diff(c, z) - lap(c) / Re
        out.append(sp.simplify(f))
    return out


if __name__ == "__main__":
    # velocity-only manufactured field
    u = sp.sin(x) * sp.cos(y) * sp.sin(z) * sp.sin(t)
    v = sp.cos(x) * sp.sin(y) * sp.sin(z) * sp.sin(t)
    w = 2 * sp.cos(x) * sp.cos(y) * sp.cos(z) * sp.sin(t)
    assert sp.simplify(sp.diff(u, x) + sp.diff(v, y) + sp.diff(w, z)) == 0
    for name, f in zip(("fx", "fy", "fz"), forcing(u, v, w)):
        print(name, "=", sp.expand(f))

    # Ethier-Steinman
    a, d = sp.pi / 4, sp.pi / 2
    E = sp.exp(-d * d * t / Re)
    ue = -a * (sp.exp(a * x) * sp.sin(a * y + d * z) + sp.exp(a * z) * sp.cos(a * x + d * y)) * E
    ve = -a * (sp.exp(a * y) * sp.sin(a * z + d * x) + sp.exp(a * x) * sp.cos(a * y + d * z)) * E
    we = -a * (sp.exp(a * z) * sp.sin(a * x + d * y) + sp.exp(a * y) * sp.cos(a * z + d * x)) * E
    pe = -a * a / 2 * (sp.exp(2 * a * x) + sp.exp(2 * a * y) + sp.exp(2 * a * z)
                       + 2 * sp.sin(a * x + d * y) * sp.cos(a * z + d * x) * sp.exp(a * (y + z))
                       + 2 * sp.sin(a * y + d * z) * sp.cos(a * x + d * y) * sp.exp(a * (z + x))
                       + 2 * sp.sin(a * z + d * x) * sp.cos(a * y + d * z) * sp.exp(a * (x + y))) * E * E
    assert sp.simplify(sp.diff(ue, x) + sp.diff(ve, y) + sp.diff(we, z)) == 0
    fes = forcing(ue, ve, we)
    # exact NS solution: forcing + grad p == 0
    for c, f in zip((x, y, z), fes):
        assert sp.simplify(f + sp.diff(pe, c)) == 0
    for c in (x, y, z):
        print("dp/d" + str(c), "=", sp.expand(sp.diff(pe, c)))
