"""The Fortran host layer (reference module names, iso_c_binding) drives the
GPU projection and matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from channel3d_parallel_amd import fields
from oracle import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "projection_driver")


CHANNEL_FD = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "channel_fd")


def test_fortran_rk3_time_loop(tmp_path):
    """The Fortran time loop (compute_dt + compute_time_step_RK3 with
    constant mass flow, as app/main.f90 runs it) on the 32x64x32-cell channel for 5 steps."""
    if not os.path.exists(CHANNEL_FD):
        pytest.fail("Fortran channel_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz, nsteps = 34, 65, 34, 5
    nu, cfl = 1.0 / 2800.0, 0.5
    dpdx0 = (180 * nu) ** 2
    print(f"\n[case] fortran channel_fd {nx}x{ny}x{nz} steps={nsteps}", flush=True)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=1)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    r = subprocess.run([CHANNEL_FD, str(nx), str(ny), str(nz), "1", "2.0", repr(nu), repr(cfl),
                        repr(dpdx0), str(nsteps), "1", str(fin), str(fout)],
                       check=True, timeout=120, capture_output=True, text=True)
    print(r.stdout)
    raw = np.fromfile(fout)
    g = orc.Grid(nx, ny, nz)
    outs, off = [], 0
    for s in (g.shape_u, g.shape_v, g.shape_w, g.shape_p):
        n = int(np.prod(s))
        outs.append(raw[off:off + n].reshape(s, order="F"))
        off += n
    dpdx_gpu, t_gpu, div_gpu, utau_gpu = raw[off:off + 4]
    g.apply_bc(U, V, W)
    q0 = g.mean_mass_flow_u(U)
    P = g.zeros("p")
    dpdx, t = dpdx0, 0.0
    for _ in range(nsteps):
        dt = g.compute_dt(U, V, W, cfl, nu)
        dpdx, _ = g.time_step_rk3(U, V, W, P, dt, nu, dpdx, q0)
        t += dt
    for name, a, b in (("U", outs[0], U), ("V", outs[1], V), ("W", outs[2], W), ("P", outs[3], P)):
        rr = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"  {name}: rel L2 = {rr:.3e}")
        assert rr <= 1e-11, name
    assert abs(dpdx_gpu - dpdx) <= 1e-9 * abs(dpdx)
    assert abs(t_gpu - t) <= 1e-14 * t
    assert div_gpu <= 1e-10
    utau, _, _, _ = g.wall_shear(U, W, nu)  # statistics::compute_wall_shear
    print(f"  utau gpu {utau_gpu:.15e} oracle {utau:.15e}")
    assert abs(utau_gpu - utau) <= 1e-10 * utau


@pytest.mark.parametrize("nx,ny,nz,gt,rk,prec,nref", [(34, 65, 34, 1, 1, 0, -1), (66, 33, 34, 0, 2, 0, -1),
                                                     (34, 65, 34, 1, 2, 0, 1), (34, 65, 34, 1, 1, 1, -1)])
def test_fortran_projection_driver(tmp_path, nx, ny, nz, gt, rk, prec, nref):
    """The Fortran host's projection: FP64 (the reference's arithmetic), FP64
    with one defect-correction sweep, and mixed precision with its default
    sweep (c3d_initialize's prec argument, c3d_set_refinement)."""
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (make -C channel3d_parallel_amd/fortran)")
    print(f"\n[case] fortran driver {nx}x{ny}x{nz} gt={gt} rk={rk} prec={prec} nref={nref}", flush=True)
    U, V, W = fields.white_faces(nx, ny, nz, seed=21)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    dt = 3e-3
    subprocess.run([DRIVER, str(nx), str(ny), str(nz), str(gt), "2.0", repr(dt), str(rk),
                    str(fin), str(fout), str(prec), str(nref)], check=True, timeout=120)
    raw = np.fromfile(fout)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    shapes = [g.shape_u, g.shape_v, g.shape_w, g.shape_p, g.shape_rhs]
    outs, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        outs.append(raw[off:off + n].reshape(s, order="F"))
        off += n
    max_div = raw[off]
    g.apply_bc(U, V, W)
    P = np.zeros(g.shape_p, order="F")
    ref = g.projection(U, V, W, rk_step=rk, dt=dt, P=P)
    g.apply_bc(U, V, W)  # the driver reads back before the trailing BC; compare interiors
    tol = 1e-11 if prec == 1 else 1e-12  # the mixed path's bar
    for name, a, b in (("U", outs[0], U), ("V", outs[1], V), ("W", outs[2], W),
                       ("P_interim", outs[4], ref["P_interim"])):
        sl = (slice(1, -2), slice(1, -1), slice(1, -2))
        r = np.linalg.norm(a[sl] - b[sl]) / np.linalg.norm(b[sl])
        assert r <= tol, (name, r)
    if rk == 1:
        assert np.linalg.norm(outs[3] - P) <= tol * np.linalg.norm(P)
    # one mixed sweep leaves ~2e-7 of the first solve's divergence (1e-10 bar)
    assert max_div <= (1e-10 if prec == 1 else max(1e-12, 4 * g.max_abs_div(U, V, W)))


CYL_FD = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "cylinders_fd")


def test_fortran_cylinders_matches_c_abi(tmp_path):
    """The Fortran IB time loop (setup_IB_operators, compute_time_step_RK3 with
    compute_IB_projection) gives bitwise the fields of the same steps driven
    through the C ABI's compute_time_step_RK3: same kernels, same order."""
    from channel3d_parallel_amd import Channel, ib_geometry
    if not os.path.exists(CYL_FD):
        pytest.fail("Fortran cylinders_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz, nsteps = 34, 33, 34, 3
    nu, dt, dpdx = 1.0 / 2800.0, 0.004, 3.0 / 2800.0
    print(f"\n[case] fortran cylinders_fd {nx}x{ny}x{nz} steps={nsteps}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=0)
    m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=16, nth=24)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.0, grid_type=0)
    fmk, fin, fout = tmp_path / "mk.bin", tmp_path / "in.bin", tmp_path / "out.bin"
    with open(fmk, "wb") as f:
        np.array([m.nb, m.nzb, m.nxb, m.nxb1], dtype=np.int32).tofile(f)
        np.concatenate([m.xb, m.yb, m.zb, m.normals, m.sb, m.ub]).tofile(f)
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    r = subprocess.run([CYL_FD, str(nx), str(ny), str(nz), repr(nu), repr(dt), repr(dpdx),
                        str(nsteps), "1", "0", str(fmk), str(fin), str(fout)],
                       check=True, timeout=120, capture_output=True, text=True)
    print(r.stdout)
    raw = np.fromfile(fout)
    outs, off = [], 0
    for s in (g.shape_u, g.shape_v, g.shape_w):
        n = int(np.prod(s))
        outs.append(raw[off:off + n].reshape(s, order="F"))
        off += n
    fb_f = raw[off:off + 4 * m.nb]
    div_f, slip_f, iters_f = raw[off + 4 * m.nb:off + 4 * m.nb + 3]
    g.apply_bc(U, V, W)
    with Channel(nx, ny, nz, grid_type=0, nu=nu) as ch:
        ch.set_fields(U, V, W)
        ch.setup_IB(m)
        ch.set_pressure_gradient(dpdx, 0.0, 0.0)
        ch.set_IB_solver(tol=1e-10, maxit=1000, guess_mode=0)
        ch.set_rhs_exact(True)  # the Fortran loop calls compute_rhs and the update separately
        for _ in range(nsteps):
            ch.compute_time_step_RK3(dt, x_mass_cte=0)
        ref = ch.get_fields()
        fb = ch.get_IB_force()
        _, accum, _ = ch.ib_stats()
    for name, a, b in (("U", outs[0], ref["U"]), ("V", outs[1], ref["V"]), ("W", outs[2], ref["W"])):
        assert np.array_equal(a, b), name
    assert np.array_equal(fb_f, fb)
    assert int(iters_f) == accum
    assert div_f < 1e-12 and slip_f < 1e-9


WAVES_FD = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "waves_fd")


@pytest.mark.parametrize("body,p1,p2,p3,ramp", [("standing_wave", 0.04, 2.0, 1.0, 0.0),
                                                 ("traveling_wave_x", 0.03, 1.0, 1.0, 0.5)])
def test_fortran_waves_matches_c_abi(tmp_path, body, p1, p2, p3, ramp):
    """The Fortran time loop with moving walls (setup_IB_geometry, then the
    per-stage move inside compute_time_step_RK3) gives bitwise the fields,
    force, marker heights and time of the same steps driven through the
    C ABI's compute_time_step_RK3."""
    from channel3d_parallel_amd import Channel
    if not os.path.exists(WAVES_FD):
        pytest.fail("Fortran waves_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz, alpha, mbw, nsteps = 34, 65, 18, 2.6, 0.1, 3
    nxb, nzb = 32, 8
    nu, dt, dpdx, t0 = 1.0 / 2800.0, 0.01, 3.0 / 2800.0, 0.2
    print(f"\n[case] fortran waves_fd {body} {nx}x{ny}x{nz} steps={nsteps}", flush=True)
    kw = dict(alpha=alpha, grid_type=2, min_buffer_width=mbw)
    g = orc.Grid(nx, ny, nz, **kw)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.05, seed=3, **kw)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    r = subprocess.run([WAVES_FD, str(nx), str(ny), str(nz), repr(alpha), repr(mbw), repr(nu),
                        repr(dt), repr(dpdx), str(nsteps), body, str(nxb), str(nzb), repr(p1),
                        repr(p2), repr(p3), repr(ramp), repr(t0), str(fin), str(fout)],
                       check=True, timeout=120, capture_output=True, text=True)
    print(r.stdout)
    nb = 2 * nxb * nzb
    raw = np.fromfile(fout)
    outs, off = [], 0
    for s in (g.shape_u, g.shape_v, g.shape_w):
        n = int(np.prod(s))
        outs.append(raw[off:off + n].reshape(s, order="F"))
        off += n
    fb_f, yb_f = raw[off:off + 4 * nb], raw[off + 4 * nb:off + 5 * nb]
    t_f, div_f, slip_f, iters_f = raw[off + 5 * nb:off + 5 * nb + 4]
    g.apply_bc(U, V, W)
    with Channel(nx, ny, nz, nu=nu, **kw) as ch:
        ch.set_fields(U, V, W)
        ch.time = t0
        ch.setup_IB_wave(body, nxb, nzb, p1, p2, p3, ramp)
        ch.set_pressure_gradient(dpdx, 0.0, 0.0)
        ch.set_IB_solver(tol=1e-10, maxit=1000, guess_mode=0)
        ch.set_rhs_exact(True)  # the Fortran loop calls compute_rhs and the update separately
        for _ in range(nsteps):
            ch.compute_time_step_RK3(dt, x_mass_cte=0)
        ref = ch.get_fields()
        fb = ch.get_IB_force()
        yb = ch.get_IB_markers()["yb"]
        _, accum, _ = ch.ib_stats()
        t = ch.time
    for name, a, b in (("U", outs[0], ref["U"]), ("V", outs[1], ref["V"]), ("W", outs[2], ref["W"])):
        assert np.array_equal(a, b), name
    assert np.array_equal(fb_f, fb) and np.array_equal(yb_f, yb)
    assert t_f == t and int(iters_f) == accum
    assert div_f < 1e-12 and slip_f < 1e-9


def test_fortran_output_data_snapshot(tmp_path):
    """input_output::output_data from the Fortran time loop writes the
    reference's snapshot format: the oracle's reader of that format finds
    the final fields, t, dt, dPdx and nu of the run, and no body."""
    from oracle import snapshot
    if not os.path.exists(CHANNEL_FD):
        pytest.fail("Fortran channel_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz, nsteps = 34, 65, 34, 3
    nu, cfl = 1.0 / 2800.0, 0.5
    dpdx0 = (180 * nu) ** 2
    print(f"\n[case] fortran output_data {nx}x{ny}x{nz} steps={nsteps}", flush=True)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=2)
    fin, fout, fsnap = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "test.3"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    subprocess.run([CHANNEL_FD, str(nx), str(ny), str(nz), "1", "2.0", repr(nu), repr(cfl),
                    repr(dpdx0), str(nsteps), "0", str(fin), str(fout), str(fsnap)],
                   check=True, timeout=120, capture_output=True, text=True)
    g = orc.Grid(nx, ny, nz)
    raw = np.fromfile(fout)
    d = snapshot.read(fsnap)
    off = 0
    for name, s in (("U", g.shape_u), ("V", g.shape_v), ("W", g.shape_w), ("P", g.shape_p)):
        n = int(np.prod(s))
        assert np.array_equal(d[name], raw[off:off + n].reshape(s, order="F")), name
        off += n
    dpdx, t = raw[off:off + 2]
    assert d["t"] == t and d["dpdx"] == dpdx and d["nu"] == nu and d["dt"] > 0
    assert np.array_equal(d["y"], g.get(1)) and d["xb"].size == 0 and d["nxb"] == 0
    assert not d["Hu_interior"].any()


def test_fortran_restart_from_snapshot(tmp_path):
    """read_input_data in the Fortran host: 3 steps + snapshot, then a new
    run restarted from it for 2 steps, against 5 uninterrupted steps.  The
    restarted run recomputes the flow rate to hold from the restored field
    (app/main.f90:90, as the reference does), so it agrees to round-off;
    the time, restored from the file, is exact."""
    if not os.path.exists(CHANNEL_FD):
        pytest.fail("Fortran channel_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz = 34, 65, 34
    nu, cfl = 1.0 / 2800.0, 0.5
    dpdx0 = (180 * nu) ** 2
    print(f"\n[case] fortran restart {nx}x{ny}x{nz}", flush=True)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=6)
    fin = tmp_path / "in.bin"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)

    def run(src, nsteps, out, snap=None):
        args = [CHANNEL_FD, str(nx), str(ny), str(nz), "1", "2.0", repr(nu), repr(cfl),
                repr(dpdx0), str(nsteps), "0", src, str(out)]
        subprocess.run(args + ([str(snap)] if snap else []), check=True, timeout=120,
                       capture_output=True, text=True)
        return np.fromfile(out)

    full = run(str(fin), 5, tmp_path / "a.bin")
    run(str(fin), 3, tmp_path / "b.bin", tmp_path / "snap.3")
    rest = run("snapshot:" + str(tmp_path / "snap.3"), 2, tmp_path / "c.bin")
    g = orc.Grid(nx, ny, nz)
    n = sum(int(np.prod(s)) for s in (g.shape_u, g.shape_v, g.shape_w, g.shape_p))
    a, c = full[:n], rest[:n]
    assert np.linalg.norm(a - c) <= 1e-13 * np.linalg.norm(a)
    dpdx_a, t_a = full[n:n + 2]
    dpdx_c, t_c = rest[n:n + 2]
    assert t_c == t_a and abs(dpdx_c - dpdx_a) <= 1e-9 * abs(dpdx_a)


DROPIN_FD = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "dropin_fd")


def test_fortran_host_array_procedures_leave_resident_state(tmp_path):
    """The reference's host-array procedures (apply_boundary_conditions,
    divergence, compute_rhs_u/v/w, compute_mean_mass_flow_U,
    compute_dPx_for_constant_mass_flow, regT, regT_1n, regTc_1n, regu/v/w,
    remove_mean_per_body, compute_heaviside), called by the Fortran host on its
    own scratch arrays between two resident RK3 steps: the resident velocity is
    bitwise unchanged by them, the two steps end bitwise where they end
    without the calls, and each result matches the oracle on the same arrays."""
    from channel3d_parallel_amd import ib_geometry
    if not os.path.exists(DROPIN_FD):
        pytest.fail("Fortran dropin_fd not built (make -C channel3d_parallel_amd/fortran)")
    nx, ny, nz = 34, 33, 34
    nu, dt, dpdx = 1.0 / 2800.0, 0.004, 3.0 / 2800.0
    print(f"\n[case] fortran dropin_fd {nx}x{ny}x{nz}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=0)
    m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=16, nth=24)
    nb = m.nb
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.0, grid_type=0)
    Ua, Va, Wa = fields.white_faces(nx, ny, nz, seed=31)
    rng = np.random.default_rng(32)
    Pa = np.asfortranarray(rng.standard_normal(g.shape_rhs))
    f = rng.standard_normal(3 * nb)
    fmk, fin, fsc = tmp_path / "mk.bin", tmp_path / "in.bin", tmp_path / "sc.bin"
    with open(fmk, "wb") as fh:
        np.array([m.nb, m.nzb, m.nxb, m.nxb1], dtype=np.int32).tofile(fh)
        np.concatenate([m.xb, m.yb, m.zb, m.normals, m.sb, m.ub]).tofile(fh)
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    np.concatenate([Ua.ravel("F"), Va.ravel("F"), Wa.ravel("F"), Pa.ravel("F"), f]).tofile(fsc)
    raws = []
    for mode in (0, 1):
        fout = tmp_path / f"out{mode}.bin"
        subprocess.run([DROPIN_FD, str(nx), str(ny), str(nz), repr(nu), repr(dt), repr(dpdx),
                        str(mode), str(fmk), str(fin), str(fsc), str(fout)],
                       check=True, timeout=120, capture_output=True, text=True)
        raws.append(np.fromfile(fout))
    nstate = int(np.prod(g.shape_u) + np.prod(g.shape_v) + np.prod(g.shape_w)) + 4 * nb
    assert np.array_equal(raws[0][:nstate], raws[1][:nstate]), "the calls changed the time steps"
    assert raws[0][nstate] == 1.0 and raws[1][nstate] == 1.0, "resident velocity changed"
    raw, off = raws[1], nstate + 1

    def take(shape):
        nonlocal off
        n = int(np.prod(shape))
        a = raw[off:off + n].reshape(shape, order="F")
        off += n
        return a

    bc = [take(s) for s in (g.shape_u, g.shape_v, g.shape_w)]
    dv = take(g.shape_rhs)
    F = [take(s) for s in (g.shape_fu, g.shape_fv, g.shape_fw)]
    qflow, dpx = take((2,))
    Eu, E1nu, E1np = take((3 * nb,)), take((3 * nb,)), take((nb,))
    R = [take(s) for s in (g.shape_u, g.shape_v, g.shape_w)]
    fm = take((nb,))
    assert off == raw.size
    g.apply_bc(Ua, Va, Wa)
    for a, b, n in zip(bc, (Ua, Va, Wa), "UVW"):
        assert np.array_equal(a, b), "apply_boundary_conditions " + n
    assert np.array_equal(dv, g.divergence(Ua, Va, Wa)), "divergence"
    for a, b, n in zip(F, g.rhs(Ua, Va, Wa, nu, dpdx, 0.0, 0.0), "uvw"):
        assert np.array_equal(a, b), "compute_rhs_" + n
    ib = orc.IB(g, m)
    ib.heaviside()
    Hu = ib.heaviside_coeffs()["Hu_int"]
    q = g.mean_mass_flow_u(Ua, Hu)
    assert abs(qflow - q) <= 1e-13 * abs(q), (qflow, q)
    assert abs(dpx - (1.0 - q)) <= 1e-13 * abs(q)
    for a, b, n in ((Eu, ib.E(Ua, Va, Wa), "regT"), (E1nu, ib.E(Ua, Va, Wa, with_dxn=True), "regT_1n"),
                    (E1np, ib.E1n_c(Pa), "regTc_1n")):
        err = np.max(np.abs(a - b)) / np.max(np.abs(b))
        assert err <= 1e-13, (n, err)
    for comp, n in enumerate(("regu", "regv", "regw")):
        ref = ib.R(comp, f[comp * nb:(comp + 1) * nb])
        assert np.max(np.abs(R[comp] - ref)) <= 1e-13 * np.max(np.abs(ref)), n
        assert np.array_equal(R[comp] != 0, ref != 0), n + " support"
    ref = ib.remove_mean(f[:nb])
    assert np.max(np.abs(fm - ref)) <= 1e-13 * np.max(np.abs(f[:nb])), "remove_mean_per_body"
