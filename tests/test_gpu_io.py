"""The output writers, the restart files and the `mif` driver on the GPU.

1. io_test (tests/cpp/io_test.cpp) sets linear fields from host callbacks,
   so every interpolation the writers make is exact: the VTK planes and the
   full mesh hold the closed forms at the pressure nodes, the dat lines the
   closed forms at the line points (the one branch where the reference reads
   v for u is checked against the restatement in output_check.py); the text
   dump/load keeps 6 significant digits, the binary one is exact.
2. The `mif` driver runs both lid-driven cases; its solution.vtk and
   profile*.dat equal the writers' formulas applied to the oracle's fields
   after the same steps (1e-11 relative in the binary VTK, the 9 printed
   digits in the dat files)."""
import os
import subprocess

import numpy as np
import pytest

import mif_oracle as mo
from output_check import Geometry, dat_rows, read_dat, read_vtk_ascii, read_vtk_binary, vtk_points

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "mpi-incompressible-fluid_amd", "_build")

fu = lambda x, y, z: 1 + 2 * x + 3 * y + 5 * z  # noqa: E731
fv = lambda x, y, z: 2 - x + 4 * y + z  # noqa: E731
fw = lambda x, y, z: -1 + x + y - 2 * z  # noqa: E731
fp = lambda x, y, z: 3 + x - y + 2 * z  # noqa: E731


def _geometry(N, pz):
    if pz:
        return Geometry((N, N, N), (1.0, 1.0, 1.0), (-0.5, -0.5, -0.5), (0, 0, 1))
    return Geometry((N, N, N), (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0))


@pytest.mark.parametrize("pz", [0, 1])
def test_writers_and_restart_files(tmp_path, pz):
    N = 9
    print(f"[case] io_test N={N} periodic_z={pz}", flush=True)
    r = subprocess.run([os.path.join(BUILD, "io_test"), str(tmp_path), str(N), str(pz)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    G = _geometry(N, pz)
    u, v, w, p = (G.field_from(f, c) for f, c in ((fu, "u"), (fv, "v"), (fw, "w"), (fp, "p")))

    # binary VTK: the three planes, values = closed forms at the nodes
    pts, vals = read_vtk_binary(tmp_path / "solution.vtk")
    ref_pts, ref_vals = vtk_points(G, u, v, w, p)
    assert pts.shape == ref_pts.shape
    assert np.allclose(pts, ref_pts, rtol=0, atol=1e-14)
    X, Y, Z = pts[:, 0], pts[:, 1], pts[:, 2]
    for name, f in (("u", fu), ("v", fv), ("w", fw), ("p", fp)):
        assert np.allclose(vals[name], f(X, Y, Z), rtol=0, atol=1e-12), name
    assert np.allclose(np.stack([vals[n] for n in "uvwp"], 1), ref_vals, rtol=0, atol=1e-12)

    # ASCII full mesh (default stream formatting: 6 significant digits)
    dims, origin, spacing, full = read_vtk_ascii(tmp_path / "full.vtk")
    Nw, sl, _ = G.ranges()
    assert dims == tuple(Nw)
    xs = [G.coord(0, i) for i in range(sl[0], sl[0] + Nw[0])]
    ys = [G.coord(1, j) for j in range(sl[1], sl[1] + Nw[1])]
    zs = [G.coord(2, k) for k in range(sl[2], sl[2] + Nw[2])]
    Zg, Yg, Xg = (a.ravel() for a in np.meshgrid(zs, ys, xs, indexing="ij"))
    for name, f in (("u", fu), ("v", fv), ("w", fw), ("p", fp)):
        ref = f(Xg, Yg, Zg)
        assert np.allclose(full[name], ref, rtol=1e-5, atol=1e-5), name
    mag = np.sqrt(fu(Xg, Yg, Zg) ** 2 + fv(Xg, Yg, Zg) ** 2 + fw(Xg, Yg, Zg) ** 2)
    assert np.allclose(full["|u|"], mag, rtol=1e-5, atol=1e-5)

    # dat lines: on node planes and halfway between them the interpolation of
    # a linear field is exact, except where the reference reads v for u
    c = [G.min[a] + (G.Ng[a] - 1) * G.h[a] / 2 for a in range(3)]
    h = G.h
    lines = {"line_x": (0, c[0], c[1], c[2]), "line_y": (1, c[0], c[1], c[2]), "line_z": (2, c[0], c[1], c[2]),
             "line_x_half": (0, c[0], c[1] + h[1] / 2, c[2] + h[2] / 2),
             "line_y_half": (1, c[0] + h[0] / 2, c[1], c[2] + h[2] / 2),
             "line_z_half": (2, c[0] + h[0] / 2, c[1] + h[1] / 2, c[2])}
    for name, (d, x, y, z) in lines.items():
        got = read_dat(tmp_path / f"{name}.dat")
        ref = dat_rows(G, u, v, w, p, d, x, y, z)
        assert got.shape == ref.shape, name
        assert np.allclose(got, ref, rtol=1e-8, atol=1e-8), name
        quirk = name in ("line_y_half", "line_z_half")
        for col, f in ((3, fu), (4, fv), (5, fw), (6, fp)):
            if quirk and col == 3:
                continue
            exact = f(got[:, 0], got[:, 1], got[:, 2])
            assert np.allclose(got[:, col], exact, rtol=1e-8, atol=1e-8), (name, col)

    # restart files
    out = dict(line.split() for line in r.stdout.splitlines())
    assert float(out["text_round_trip"]) < 1e-5 and float(out["binary_round_trip"]) == 0.0
    assert out["binary_wrong_staggering_refused"] == "1" and out["list_load"] == "1"
    text = (tmp_path / "u.txt").read_text().split("\n")
    assert text[0] == "Tensor (row major). Dimensions: [ %d %d %d ]" % G.ext["u"]

    # the Python binding reads and writes the same files
    import mif_amd as mif

    gs = mif.GridSpec(Nx_global=N, Ny_global=N, Nz_global=N, x_size=1.0, y_size=1.0, z_size=1.0 if pz else 2.0,
                      min_x=-0.5 if pz else 0.0, min_y=-0.5 if pz else 0.0, min_z=-0.5 if pz else -1.0, Re=1e3,
                      final_time=1e-3, num_time_steps=1, periodic=(0, 0, pz))
    ctx = mif.Context(gs)
    pf, uf = ctx.pressure(), ctx.velocity().components[0]
    pf.load_binary(str(tmp_path / "p.bin"))
    assert np.array_equal(pf.download(), p)
    with pytest.raises(ValueError):
        uf.load_binary(str(tmp_path / "p.bin"))
    uf.load(str(tmp_path / "u.txt"))
    assert np.allclose(uf.download(), u, rtol=1e-5, atol=1e-5)
    uf.dump(str(tmp_path / "u_py.txt"))
    assert (tmp_path / "u_py.txt").read_text() == (tmp_path / "u.txt").read_text()
    pf.dump_binary(str(tmp_path / "p_py.bin"))
    assert (tmp_path / "p_py.bin").read_bytes() == (tmp_path / "p.bin").read_bytes()
    ctx.close()


def _oracle_lid(N, Nt, dt, case2):
    Re = 1e3
    if case2:
        size, origin, per = (1.0, 1.0, 1.0), (-0.5, -0.5, -0.5), (0, 0, 1)
    else:
        size, origin, per = (1.0, 1.0, 2.0), (0.0, 0.0, -1.0), (0, 0, 0)
    o = mo.Oracle(mo.make_grid((N, N, N), size, origin, Re, dt * Nt, Nt, per))
    ev = mo.make_fn(mo.FN_LID_2 if case2 else mo.FN_LID_1, Re=Re)
    vel, buf, buf2 = o.velocity(), o.velocity(), o.velocity()
    o.set_velocity(vel, ev, 0.0)
    p, pb = o.zeros(3), o.zeros(3)
    for s in range(Nt):
        o.timestep(vel, buf, buf2, ev, s * o.dt, p, pb)
    return Geometry((N, N, N), size, origin, per), vel, p


@pytest.mark.parametrize("case2", [False, True])
def test_mif_driver_lid_cavity(tmp_path, case2):
    N, Nt, dt = 17, 4, 1e-3
    print(f"[case] mif lid case_2={case2} N={N} Nt={Nt}", flush=True)
    (tmp_path / "input.txt").write_text(f"Nt : {Nt}\ndt : {dt}\nNx : {N}\nNy : {N}\nNz : {N}\nPy : 1\nPz : 1\n"
                                        f"test_case_2 : {'true' if case2 else 'false'}\n")
    # stale outputs of an earlier run are removed first
    (tmp_path / "profile3.dat").write_text("stale")
    r = subprocess.run([os.path.join(BUILD, "mif"), "input.txt"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    G, vel, p = _oracle_lid(N, Nt, dt, case2)
    u, v, w = vel
    pts, vals = read_vtk_binary(tmp_path / "solution.vtk")
    ref_pts, ref_vals = vtk_points(G, u, v, w, p)
    assert np.allclose(pts, ref_pts, rtol=0, atol=1e-14)
    got = np.stack([vals[n] for n in "uvwp"], 1)
    # per column, relative to its largest value (w stays at rounding level in
    # these planes, so the floor is a millionth of the largest column)
    scale = np.max(np.abs(ref_vals), axis=0)
    scale = np.maximum(scale, 1e-6 * scale.max())
    assert np.all(np.max(np.abs(got - ref_vals), axis=0) <= 1e-11 * scale), (np.max(np.abs(got - ref_vals), 0), scale)
    profiles = ([(1, 0.5, 0.5, 0), (0, 0.5, 0.5, 0)] if not case2 else [(1, 0, 0, 0), (0, 0, 0, 0), (2, 0, 0, 0)])
    for q, (d, x, y, z) in enumerate(profiles, start=1):
        got = read_dat(tmp_path / f"profile{q}.dat")
        ref = dat_rows(G, u, v, w, p, d, x, y, z)
        assert got.shape == ref.shape
        assert np.allclose(got, ref, rtol=2e-8, atol=1e-9 * np.max(np.abs(ref[:, 3:]))), q
    assert case2 or not (tmp_path / "profile3.dat").exists()
    # the lid moves the fluid: v on the profile next to the lid is not zero
    assert np.max(np.abs(read_dat(tmp_path / "profile2.dat")[:, 4])) > 1e-3
