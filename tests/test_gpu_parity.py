"""Parity of the HIP library (through the C ABI) with the CPU oracle.

Every case runs the same seeded inputs through both and requires the FP64
results to agree to 1e-11 relative (max-abs difference over max-abs value of
the oracle field), the tolerance the timestep path is specified to.
"""
import numpy as np
import pytest

import mif_oracle as mo
from helpers import PI, channel_initial, grids, rel_err, seeded_normal

pytestmark = pytest.mark.gpu

TOL = 1e-11

mif = pytest.importorskip("mif_amd")


def _print_case(name):
    print(f"[case] {name}", flush=True)


POISSON_CASES = [
    # (Ng, size, origin, periodic)
    ((16, 16, 16), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0)),       # all walls, DCT-I L=30
    ((8, 12, 11), (2 * PI, 2 * PI, 2 * PI), (0.0, 0.0, 0.0), (0, 0, 1)),  # walls x/y, periodic z
    ((33, 17, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),      # channel
    ((23, 9, 14), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 1, 1)),          # generic radices (L = 44, 8, 13)
    ((65, 65, 65), (2 * PI, 2.0, 2 * PI), (0.0, -1.0, 0.0), (1, 0, 1)),  # power-of-two path L = 64, 128
    ((257, 129, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),    # L = 256 (x, y)
    ((129, 257, 129), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),   # L = 128, 512, 128
    ((65, 33, 129), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),        # Neumann L = 128, 64, 256
    # radix-3 kernels (L = 768, 1536) and L = 1024 on every axis and mode
    ((769, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),      # x periodic L = 768
    ((9, 769, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),      # y Neumann L = 1536
    ((17, 513, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),      # y Neumann L = 1024
    ((9, 9, 769), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1)),          # z periodic L = 768 (fused z)
    ((9, 9, 513), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # z Neumann L = 1024 (fused z)
    ((5, 5, 769), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # z Neumann L = 1536 (fused z)
    ((513, 9, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # x Neumann L = 1024
    ((769, 5, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # x Neumann L = 1536
    ((9, 769, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 1, 0)),          # y periodic L = 768
    # radix-3 kernels of lengths 96, 192, 384
    ((193, 97, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),     # x periodic L = 192, y Neumann L = 192
    ((97, 193, 33), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),     # x periodic L = 96, y Neumann L = 384
    ((385, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),      # x periodic L = 384
    ((9, 9, 193), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1)),          # z periodic L = 192 (fused z)
    ((9, 9, 97), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 1)),           # z periodic L = 96 (fused z)
    ((9, 9, 193), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # z Neumann L = 384 (fused z)
    ((193, 5, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),          # x Neumann L = 384
    ((49, 5, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 0, 0)),           # x Neumann L = 96
    ((9, 193, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (0, 1, 0)),          # y periodic L = 192
]


@pytest.mark.parametrize("case", POISSON_CASES, ids=lambda c: "x".join(map(str, c[0])) + "-p" + "".join(map(str, c[3])))
def test_poisson_solve_matches_oracle(case):
    Ng, size, origin, per = case
    _print_case(f"poisson {Ng} periodic={per}")
    og, gs = grids(Ng, size, origin, 1.0, 1.0, 1, per)
    o = mo.Oracle(og)
    b = seeded_normal(tuple(reversed(o.npts)), 20261202)
    # oracle on the packed array
    ref = b.copy()
    o.poisson_packed(ref)
    # GPU on a pressure field holding b at its owner points
    ctx = mif.Context(gs)
    p = ctx.pressure()
    host = np.zeros(p.shape)
    lo = [1 if q else 0 for q in per]
    host[lo[2]:lo[2] + b.shape[0], lo[1]:lo[1] + b.shape[1], lo[0]:lo[0] + b.shape[2]] = b
    p.upload(host)
    mif.poisson_solve(p)
    out = p.download()
    got = out[lo[2]:lo[2] + b.shape[0], lo[1]:lo[1] + b.shape[1], lo[0]:lo[0] + b.shape[2]]
    assert rel_err(got, ref) < TOL
    # ghosts hold the periodic images
    full = np.zeros(p.shape)
    full[lo[2]:lo[2] + b.shape[0], lo[1]:lo[1] + b.shape[1], lo[0]:lo[0] + b.shape[2]] = ref
    o.apply_periodic(full, 3)
    assert rel_err(out, full) < TOL


BC_CASES = [
    ((12, 10, 9), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0)),
    ((12, 10, 9), (2 * PI, 2 * PI, 1.0), (0.0, 0.0, 0.0), (1, 1, 0)),
    ((12, 10, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), (1, 0, 1)),
    ((7, 8, 9), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), (1, 1, 1)),
]


@pytest.mark.parametrize("case", BC_CASES, ids=lambda c: "p" + "".join(map(str, c[3])))
def test_apply_bc_matches_oracle(case):
    Ng, size, origin, per = case
    _print_case(f"apply_bc {Ng} periodic={per}")
    og, gs = grids(Ng, size, origin, 1e3, 1e-4, 1, per)
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_ES_VELOCITY, Re=1e3)
    vel = o.velocity()
    for c in range(3):
        vel[c][...] = seeded_normal(vel[c].shape, 7 + c)
    ctx = mif.Context(gs)
    gv = ctx.velocity()
    gv.upload(vel)
    o.apply_bc(vel, ev, 0.37)
    gv.apply_bc(mif.function(mif.FN_ES_VELOCITY, Re=1e3), 0.37)
    for c, got in enumerate(gv.download()):
        assert rel_err(got, vel[c]) < 1e-14, c


def _run_full(Ng, size, origin, Re, T, nsteps, per, init, fn_kind, seed=0):
    og, gs = grids(Ng, size, origin, Re, T, nsteps, per)
    o = mo.Oracle(og)
    ev = mo.make_fn(fn_kind, Re=Re)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    if init == "es":
        o.set_velocity(vel, ev, 0.0)
        o.set(p, 3, mo.make_fn(mo.FN_ES_PRESSURE, Re=Re), 0, 0.0)
    else:
        v0 = channel_initial(o, seed)
        for c in range(3):
            vel[c][...] = v0[c]
    ctx = mif.Context(gs)
    gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gp, gpb = ctx.pressure(), ctx.pressure()
    gvel.upload(vel)
    gp.upload(p)
    fn = mif.function(fn_kind, Re=Re)
    for s in range(nsteps):
        t = s * o.dt
        o.timestep(vel, buf, buf2, ev, t, p, pb)
        mif.timestep(gvel, gbuf, gbuf2, fn, t, gp, gpb)
    return o, vel, p, gvel.download(), gp.download()


def test_full_timestep_ethier_steinman_walls():
    _print_case("timestep ES 16^3 walls")
    o, vel, p, gvel, gp = _run_full((16, 16, 16), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), 1e3, 1e-4, 3, (0, 0, 0), "es",
                                    mo.FN_ES_VELOCITY)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


def test_full_timestep_channel_periodic():
    _print_case("timestep channel 17x9x17")
    o, vel, p, gvel, gp = _run_full((17, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-3, 4, (1, 0, 1),
                                    "channel", mo.FN_ZERO, seed=3)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


def test_timestep_velocity_matches_oracle():
    _print_case("timestep_velocity 12^3")
    Re = 1e4
    og, gs = grids((12, 12, 12), (1.0, 1.0, 1.0), (0.0, 0.0, 0.0), Re, 1e-3, 3, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_VT_VELOCITY, Re=Re)
    fo = mo.make_fn(mo.FN_VT_FORCING, Re=Re)
    vel, buf, rhs = o.velocity(), o.velocity(), o.velocity()
    o.set_velocity(vel, ev, 0.0)
    ctx = mif.Context(gs)
    gvel, gbuf, grhs = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gvel.upload(vel)
    gev, gfo = mif.function(mif.FN_VT_VELOCITY, Re=Re), mif.function(mif.FN_VT_FORCING, Re=Re)
    for s in range(3):
        o.timestep_velocity(vel, buf, rhs, ev, fo, s * o.dt)
        mif.timestep_velocity(gvel, gbuf, grhs, gev, gfo, s * o.dt)
    for c, got in enumerate(gvel.download()):
        assert rel_err(got, vel[c]) < TOL, c
    # norms through the library equal the oracle's on the same field
    T = 3 * o.dt
    n_ref = o.error_norms_velocity(vel, ev, T)
    n_gpu = mif.error_norms_velocity(gvel, gev, T)
    assert np.allclose(n_gpu, n_ref, rtol=1e-9, atol=0)


def test_full_timestep_channel_pow2_fused_path():
    # power-of-two x length: divergence fused into the x forward pass and the
    # x inverse pass writing the pressure field with its periodic images
    _print_case("timestep channel 65x33x65")
    o, vel, p, gvel, gp = _run_full((65, 33, 65), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-3, 3, (1, 0, 1),
                                    "channel", mo.FN_ZERO, seed=5)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


def test_full_timestep_channel_radix3_fused_path():
    # x length 768 = 3 * 2^8: the fused x passes of the radix-3 kernels
    _print_case("timestep channel 769x9x17")
    o, vel, p, gvel, gp = _run_full((769, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 1e-3, 2, (1, 0, 1),
                                    "channel", mo.FN_ZERO, seed=7)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


@pytest.mark.parametrize("nx", [97, 193, 385])
def test_full_timestep_channel_radix3_fused_path_small(nx):
    # x lengths 96, 192, 384 = 3 * 2^k: the fused x passes of the radix-3
    # kernels below 768
    _print_case(f"timestep channel {nx}x9x17")
    o, vel, p, gvel, gp = _run_full((nx, 9, 17), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 1e-3, 2, (1, 0, 1),
                                    "channel", mo.FN_ZERO, seed=11)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


@pytest.mark.parametrize("ny", [129, 513])
def test_full_timestep_channel_long_wall_normal_axis(ny):
    # Neumann y of 129 nodes (the 2N-point DCT-I kernel) and of 513 nodes
    # (the DCT-I through a 512-point FFT of two lines).  The pressure
    # accumulates the solves' rounding times the Laplacian's condition number
    # (~1.6e4 at 129 nodes, 16x that of the 33-node channels above; ~2.8e5 at
    # 513): any two FP64 solve orderings differ by ~4e-11 relative in p at
    # 129 nodes (both transform algorithms measured), so pressure is held to
    # 1e-11 times kappa / 1e3 and the velocity, which sees only gradients, to
    # the 1e-11 bar.
    _print_case(f"timestep channel 9x{ny}x9")
    o, vel, p, gvel, gp = _run_full((9, ny, 9), (2 * PI, 2.0, PI), (0.0, -1.0, 0.0), 1e3, 5e-3, 3, (1, 0, 1),
                                    "channel", mo.FN_ZERO, seed=9)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    kappa = (4.0 / (2.0 / (ny - 1)) ** 2) / (2 * (1 - np.cos(2 * PI / 8)) / (2 * PI / 8) ** 2)
    assert rel_err(gp, p) < TOL * kappa / 1e3, (rel_err(gp, p), kappa)


def test_full_timestep_walls_pow2_fused_path():
    _print_case("timestep ES 33^3 walls")
    o, vel, p, gvel, gp = _run_full((33, 33, 33), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), 1e3, 1e-4, 2, (0, 0, 0), "es",
                                    mo.FN_ES_VELOCITY)
    for c in range(3):
        assert rel_err(gvel[c], vel[c]) < TOL, c
    assert rel_err(gp, p) < TOL


def test_timestep_nhn_matches_oracle():
    _print_case("timestep_nhn ES 16^3 walls")
    Re, T, nsteps = 1e3, 1e-4, 2
    Ng, size, origin = (16, 16, 16), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0)
    og, gs = grids(Ng, size, origin, Re, T, nsteps, (0, 0, 0))
    o = mo.Oracle(og)
    ev = mo.make_fn(mo.FN_ES_VELOCITY, Re=Re)
    gpf = mo.make_fn(mo.FN_ES_GRAD_P, Re=Re)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    p, pb = o.zeros(3), o.zeros(3)
    o.set_velocity(vel, ev, 0.0)
    o.set(p, 3, mo.make_fn(mo.FN_ES_PRESSURE, Re=Re), 0, 0.0)
    ctx = mif.Context(gs)
    gvel, gbuf, gbuf2 = ctx.velocity(), ctx.velocity(), ctx.velocity()
    gp, gpb = ctx.pressure(), ctx.pressure()
    gvel.upload(vel)
    gp.upload(p)
    gev, ggp = mif.function(mif.FN_ES_VELOCITY, Re=Re), mif.function(mif.FN_ES_GRAD_P, Re=Re)
    for s in range(nsteps):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb, grad_p=gpf)
        mif.timestep_nhn(gvel, gbuf, gbuf2, gev, ggp, s * o.dt, gp, gpb)
    for c, got in enumerate(gvel.download()):
        assert rel_err(got, vel[c]) < TOL, c
    assert rel_err(gp.download(), p) < TOL


@pytest.mark.parametrize("Ng,per", [((33, 33, 33), (0, 0, 0)), ((17, 13, 11), (0, 0, 1))])
def test_solve_pressure_and_adjust_and_norms(Ng, per):
    # stand-alone pressure test of the reference (test/pressure_test_hn.cpp,
    # test/pressure_test_mixed.cpp): dt = 1, [0, 2pi]^3
    _print_case(f"solve_pressure {Ng} periodic={per}")
    og, gs = grids(Ng, (2 * PI,) * 3, (0.0, 0.0, 0.0), 1.0, 1.0, 1, per)
    o = mo.Oracle(og)
    vel = o.velocity()
    o.set_velocity(vel, mo.make_fn(mo.FN_PT_VELOCITY), 1.0)
    p = o.zeros(3)
    o.solve_pressure(p, vel, o.dt)
    pe = mo.make_fn(mo.FN_PT_PRESSURE)
    ctx = mif.Context(gs)
    gvel = ctx.velocity()
    gvel.set(mif.function(mif.FN_PT_VELOCITY), 1.0)
    gpres = ctx.pressure()
    mif.solve_pressure_equation_homogeneous_periodic(gpres, gvel, ctx.dt)
    assert rel_err(gpres.download(), p) < TOL
    o.adjust_pressure(p, pe, 1.0)
    mif.adjust_pressure(gpres, mif.function(mif.FN_PT_PRESSURE), 1.0)
    assert rel_err(gpres.download(), p) < TOL
    n_ref = o.error_norms_pressure(p, pe, 1.0)
    n_gpu = mif.error_norms_pressure(gpres, mif.function(mif.FN_PT_PRESSURE), 1.0)
    assert np.allclose(n_gpu, n_ref, rtol=1e-9)


def test_solve_pressure_nhn_matches_oracle():
    # test/pressure_test_nhn.cpp: [-pi/2, 0]^3, Neumann data from grad p
    _print_case("solve_pressure_nhn 17x13x11")
    og, gs = grids((17, 13, 11), (PI / 2,) * 3, (-PI / 2,) * 3, 1.0, 1.0, 1, (0, 0, 0))
    o = mo.Oracle(og)
    vel = o.velocity()
    o.set_velocity(vel, mo.make_fn(mo.FN_PT_VELOCITY), 1.0)
    p = o.zeros(3)
    # the oracle takes the gradient as a time-difference descriptor frozen at t = 1
    gfix = mo.make_fn(mo.FN_TIME_DIFFERENCE, base=mo.FN_PT_GRAD_P, t0=-1.0, t1=1.0)
    # base(1) - base(-1) = 2 * grad p(1) for this field (linear in t); halve it
    o.solve_pressure(p, vel, o.dt, grad_p=gfix)
    ctx = mif.Context(gs)
    gvel = ctx.velocity()
    gvel.set(mif.function(mif.FN_PT_VELOCITY), 1.0)
    gpres = ctx.pressure()
    mif.solve_pressure_equation_non_homogeneous_neumann(gpres, gvel, mif.function(mif.FN_PT_GRAD_P), 2.0, ctx.dt)
    assert rel_err(gpres.download(), p) < TOL
