"""GPU parity of the RK3 row (stencils, stage update, compute_dt, mass flow,
full time step) against the oracle.

Tolerances: right-hand sides and stage updates are bitwise (same operations
in the same order, FMA contraction off); the bulk-flow reduction sums in a
different order (relative 1e-14); full time steps include the projection
(relative L2 1e-12 after one step, 1e-10 after 100 steps of the 32x64x32-cell
channel; at 192x128x192 and 384x192x384 one step to 1e-12 once the
constant-flow correction's uniform shift, the oracle's serial-summation error,
is taken out).
"""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields
from oracle import orc

pytestmark = pytest.mark.gpu

NU = 1.0 / 2800.0
DPDX = (180.0 * NU) ** 2  # u_tau^2 for Re_tau = 180


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def state(nx, ny, nz, gt=1, seed=5, kind="random"):
    print(f"\n[case] {nx}x{ny}x{nz} gt={gt} seed={seed} kind={kind}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    if kind == "random":
        U, V, W = fields.white_faces(nx, ny, nz, seed)
    else:
        U, V, W = fields.channel_ic(nx, ny, nz, seed=seed, grid_type=gt)
    g.apply_bc(U, V, W)
    return g, U, V, W


@pytest.mark.parametrize("nx,ny,nz,gt", [(34, 33, 18, 1), (18, 17, 34, 0)])
def test_rhs_bitwise(nx, ny, nz, gt):
    g, U, V, W = state(nx, ny, nz, gt)
    with Channel(nx, ny, nz, grid_type=gt, nu=NU) as ch:
        ch.set_fields(U, V, W)
        ch.set_pressure_gradient(DPDX, 0.0, 0.0)
        ch.compute_rhs(2)
        gu, gv, gw = ch.get_rhs(2)
    ru, rv, rw = g.rhs(U, V, W, NU, DPDX, 0.0, 0.0)
    assert np.array_equal(gu, ru) and np.array_equal(gv, rv) and np.array_equal(gw, rw)


def test_rk3_update_bitwise():
    nx, ny, nz = 34, 33, 18
    g, U, V, W = state(nx, ny, nz)
    dt = 0.0123
    with Channel(nx, ny, nz, nu=NU) as ch:
        ch.set_fields(U, V, W)
        ch.save_state()
        ch.compute_rhs(1)
        ch.rk3_update(1, dt)
        ch.compute_rhs(2)
        ch.rk3_update(2, dt)
        ch.compute_rhs(3)
        ch.rk3_update(3, dt)
        out = ch.get_fields()
    F = [g.rhs(U, V, W, NU)]
    Us = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F")]
    cur = [a.copy(order="F") for a in Us]
    for st in (1, 2, 3):
        if st > 1:
            F.append(g.rhs(*cur, NU))
        for q in range(3):
            g.rk_update(q, st, dt, F[0][q], F[1][q] if st > 1 else None,
                        F[2][q] if st > 2 else None, Us[q], cur[q])
    assert np.array_equal(out["U"], cur[0]) and np.array_equal(out["V"], cur[1])
    assert np.array_equal(out["W"], cur[2])


def test_compute_dt_and_mass_flow():
    nx, ny, nz = 34, 65, 34
    g, U, V, W = state(nx, ny, nz, kind="ic", seed=1)
    with Channel(nx, ny, nz, nu=NU) as ch:
        ch.set_fields(U, V, W)
        for cfl in (0.5, -0.002):
            assert ch.compute_dt(cfl) == g.compute_dt(U, V, W, cfl, NU)
        q = ch.mean_mass_flow_u()
        # ~1e5-term sum in a different order than the oracle's sequential loop
        assert abs(q - g.mean_mass_flow_u(U)) <= 1e-13 * abs(q)
        ch.add_to_u(0.25)
        U2 = ch.get_fields()["U"]
    ref = U.copy(order="F")
    ref[1:-1, 1:-1, 1:-1] += 0.25
    assert np.array_equal(U2, ref)


@pytest.mark.parametrize("nsteps,tol,exact", [(1, 1e-12, False), (100, 1e-10, False),
                                               (1, 1e-12, True), (100, 1e-10, True)])
def test_channel_32x64x32_time_steps(nsteps, tol, exact):
    """32x64x32 cells (34x65x34 faces), tanh alpha 2, Re_b 2800,
    laminar parabola + 10% divergence-free perturbation, constant flow rate,
    dt from compute_dt at CFL 0.5; the fused right-hand sides (default) and
    the bitwise ones (set_rhs_exact) against the oracle."""
    nx, ny, nz = 34, 65, 34
    g, U, V, W = state(nx, ny, nz, kind="ic", seed=1)
    q0 = g.mean_mass_flow_u(U)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Pr = g.zeros("p")
    dpdx_r = DPDX
    with Channel(nx, ny, nz, nu=NU) as ch:
        ch.set_rhs_exact(exact)
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(DPDX, 0.0, 0.0)
        for n in range(nsteps):
            dt = ch.compute_dt(0.5)
            dt_r = g.compute_dt(Ur, Vr, Wr, 0.5, NU)
            assert abs(dt - dt_r) <= 1e-15 * dt_r
            ch.compute_time_step_RK3(dt, x_mass_cte=1, q0=q0)
            dpdx_r, _ = g.time_step_rk3(Ur, Vr, Wr, Pr, dt_r, NU, dpdx_r, q0)
        out = ch.get_fields()
        dpdx = ch.get_pressure_gradient()[0]
        div = ch.check_divergence()
        q = ch.mean_mass_flow_u()
    for name, a, b in (("U", out["U"], Ur), ("V", out["V"], Vr), ("W", out["W"], Wr),
                       ("P", out["P"], Pr)):
        r = rel(a, b)
        print(f"  after {nsteps} steps {name}: rel L2 = {r:.3e}")
        assert r <= tol, name
    print(f"  dPdx gpu={dpdx:.15e} oracle={dpdx_r:.15e}; max|div|={div:.2e}; Q={q:.15f} Q0={q0:.15f}")
    # dPdx = (dPdx_prev + (Q0 - <U>)/dt)/2: the bulk-flow round-off of the
    # reduction (~1e-14) is amplified by 1/dt, so the bound is absolute
    assert abs(dpdx - dpdx_r) <= 1e-11
    assert div <= 1e-10
    assert abs(q - q0) <= 1e-12


def _bulk_flow_exact(g, U):
    """mean_mass_flow_U (src/mass_flow.f90:24-53) with both sums exactly
    rounded (math.fsum), the terms formed as the oracle forms them."""
    import math
    nx, nyg, nzg = g.nx, g.nyg, g.nzg
    yg = g.get(4)
    dy = yg[2:nyg - 1] - yg[1:nyg - 2]  # yg(j) - yg(j-1), j = 3..nyg-1
    Ui = U[1:nx - 1, :, 1:nzg - 1]
    terms = (Ui[:, 2:nyg - 1, :] + Ui[:, 1:nyg - 2, :]) * 0.5 * dy[None, :, None]
    return math.fsum(terms.ravel()) / math.fsum(np.broadcast_to(dy[None, :, None], terms.shape).ravel())


@pytest.mark.parametrize("nx,ny,nz,exact,mixed", [(194, 129, 194, False, False),
                                                   (194, 129, 194, True, False),
                                                   (386, 193, 386, False, False),
                                                   (386, 193, 386, False, True)])
def test_channel_time_step_large(nx, ny, nz, exact, mixed):
    """One RK3 step at 192x128x192 cells (194x129x194 faces; both
    right-hand-side modes) and at 384x192x384 (386x193x386 faces, the
    benchmarked size; the default fused mode, in FP64 and in mixed precision
    with its default refinement sweep), tanh alpha 2: laminar parabola plus a
    divergence-free perturbation, constant flow rate, dt from compute_dt at
    CFL 0.5, against the (FP64) oracle.

    The constant-flow correction adds d = Q0 - Q(U) to every interior U face
    (src/mass_flow.f90).  The oracle sums Q in the reference's serial order,
    whose rounding grows with the cell count: against an exactly rounded sum
    it is off by 3e-14 at 32x64x32, 5e-12 here and 5e-11 at 384x192x384,
    while the library's row sums and fixed-order tree match the exact sum.
    So U is compared after removing the uniform shift the two corrections
    differ by, and that shift must equal the oracle's own summation error on
    the field it corrected (to 1e-13); V, W and P to 1e-12 as at 32x64x32.
    Mixed precision: each of the step's three projections leaves ~3e-13 of
    the velocity and P_interim against FP64, so its bars are 1e-11 (shift to
    1e-12: the mixed fields' own bulk flow moves by their difference)."""
    from channel3d_parallel_amd._lib import PREC_FP64, PREC_MIXED
    tol, stol = (1e-11, 1e-12) if mixed else (1e-12, 1e-13)
    g, U, V, W = state(nx, ny, nz, kind="ic", seed=1)
    q0 = g.mean_mass_flow_u(U)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Pr = g.zeros("p")
    with Channel(nx, ny, nz, nu=NU, precision=PREC_MIXED if mixed else PREC_FP64) as ch:
        ch.set_rhs_exact(exact)
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(DPDX, 0.0, 0.0)
        dt = ch.compute_dt(0.5)
        dt_r = g.compute_dt(Ur, Vr, Wr, 0.5, NU)
        assert abs(dt - dt_r) <= 1e-15 * dt_r
        ch.compute_time_step_RK3(dt, x_mass_cte=1, q0=q0)
        dpdx_r, _ = g.time_step_rk3(Ur, Vr, Wr, Pr, dt_r, NU, DPDX, q0)
        out = ch.get_fields()
        dpdx = ch.get_pressure_gradient()[0]
        div = ch.check_divergence()
        q = ch.mean_mass_flow_u()
    for name, a, b in (("V", out["V"], Vr), ("W", out["W"], Wr), ("P", out["P"], Pr)):
        r = rel(a, b)
        print(f"  one step {name}: rel L2 = {r:.3e}")
        assert r <= tol, name
    # the oracle's correction d_r (dPdx = (DPDX + d/dt)/2) and the field it corrected
    d_r = (2.0 * dpdx_r - DPDX) * dt_r
    Ustar = Ur.copy(order="F")
    Ustar[1:nx - 1, 1:ny, 1:nz] -= d_r
    expected = g.mean_mass_flow_u(Ustar) - _bulk_flow_exact(g, Ustar)
    diff = (out["U"] - Ur)[1:nx - 1, 1:ny, 1:nz]
    shift = float(diff.mean())
    r = np.linalg.norm(diff - shift) / np.linalg.norm(Ur[1:nx - 1, 1:ny, 1:nz])
    print(f"  one step U: rel L2 = {rel(out['U'], Ur):.3e}; uniform shift {shift:.3e} "
          f"(oracle summation error {expected:.3e}); without it rel L2 = {r:.3e}")
    assert abs(shift - expected) <= stol
    assert r <= tol
    div_r = g.max_abs_div(Ur, Vr, Wr)
    print(f"  dPdx gpu={dpdx:.15e} oracle={dpdx_r:.15e}; max|div| gpu={div:.2e} oracle={div_r:.2e}")
    # the dPdx update carries the same difference, divided by 2 dt
    assert abs((dpdx - dpdx_r) - shift / (2.0 * dt)) <= 1e-11
    # mixed precision: its divergence bar (one sweep: ~1e-10 at this size)
    assert div <= (1e-10 if mixed else max(1e-12, 4 * div_r))
    assert abs(q - q0) <= 1e-12


@pytest.mark.parametrize("nx,ny,nz,gt", [(34, 65, 34, 1), (194, 129, 194, 1), (130, 129, 130, 0)])
def test_wall_shear_vs_oracle(nx, ny, nz, gt):
    """Wall shear of compute_statistics (src/statistics.f90:41-131) on the
    resident fields: utau and the plane-mean profiles equal the oracle's up
    to summation order (1e-13 relative); wtau likewise or NaN in both (a
    negative argument of the square root, as in the reference)."""
    print(f"\n[case] wall shear {nx}x{ny}x{nz} gt={gt}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.1, seed=4, grid_type=gt)
    g.apply_bc(U, V, W)
    with Channel(nx, ny, nz, grid_type=gt, nu=NU) as ch:
        ch.set_fields(U, V, W)
        ut, wt, um, wm = ch.wall_shear()
    uo, wo, umo, wmo = g.wall_shear(U, W, NU)
    print(f"  utau gpu {ut:.15e} oracle {uo:.15e}; wtau gpu {wt:.6e} oracle {wo:.6e}")
    assert np.isfinite(uo) and abs(ut - uo) <= 1e-13 * uo
    scale = np.max(np.abs(umo))
    assert np.max(np.abs(um - umo)) <= 1e-13 * scale
    assert np.max(np.abs(wm - wmo)) <= 1e-13 * scale
    assert (np.isnan(wt) and np.isnan(wo)) or abs(wt - wo) <= 1e-12 * abs(wo)


@pytest.mark.parametrize("nx,ny,nz,gt", [(34, 65, 34, 1), (194, 129, 194, 1), (130, 129, 130, 0)])
def test_plane_statistics_vs_oracle(nx, ny, nz, gt):
    """Plane-mean profiles of compute_statistics (src/statistics.f90:41-111)
    on the resident fields equal the oracle's up to summation order: 1e-13
    of the component's magnitude (the largest entry of its mean square, or
    its square root for first moments; the plane mean of V of a
    divergence-free field is zero up to round-off, so its own size is no
    scale)."""
    print(f"\n[case] plane statistics {nx}x{ny}x{nz} gt={gt}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.1, seed=6, grid_type=gt)
    g.apply_bc(U, V, W)
    with Channel(nx, ny, nz, grid_type=gt, nu=NU) as ch:
        ch.set_fields(U, V, W)
        st = ch.plane_statistics()
    ref = g.plane_statistics(U, V, W)
    sq = {c: np.max(ref[f"{c}2mean"]) for c in "UVW"}
    scale = {"Umean": np.sqrt(sq["U"]), "Vmean": np.sqrt(sq["V"]), "Wmean": np.sqrt(sq["W"]),
             "U2mean": sq["U"], "V2mean": sq["V"], "W2mean": sq["W"],
             "UVmean": np.sqrt(sq["U"] * sq["V"])}
    for k, r in ref.items():
        err = np.max(np.abs(st[k] - r)) / scale[k]
        print(f"  {k}: max err / scale {err:.2e}")
        assert err <= 1e-13, k


@pytest.mark.parametrize("nx,ny,nz,gt", [(34, 65, 34, 1), (194, 129, 194, 1)])
def test_monitor_vs_oracle(nx, ny, nz, gt):
    """output_monitor's field summaries (src/monitor.f90:27-39, 57-59) of the
    resident fields: maxima exactly, means to 1e-13 relative (summation
    order), and the non-finite count, also after planting NaN/Inf values in
    the interiors (and one in a ghost cell, which is not counted)."""
    print(f"\n[case] monitor {nx}x{ny}x{nz} gt={gt}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.1, seed=8, grid_type=gt)
    g.apply_bc(U, V, W)
    with Channel(nx, ny, nz, grid_type=gt, nu=NU) as ch:
        ch.set_fields(U, V, W)
        m = ch.monitor()
        U2, W2 = U.copy(order="F"), W.copy(order="F")
        U2[4, 3, 2] = np.nan
        W2[2, 2, 3] = -np.inf
        W2[0, 2, 3] = np.nan      # ghost column i = 1
        ch.set_fields(U2, V, W2)
        m_bad = ch.monitor()
    r = g.monitor(U, V, W)
    print(f"  gpu {m}\n  oracle {r}")
    for k in ("maxU", "maxV", "maxW", "nonfinite"):
        assert m[k] == r[k], k
    for k in ("meanU", "meanV", "meanW"):
        assert abs(m[k] - r[k]) <= 1e-13 * max(abs(r["meanU"]), 1e-300), k
    assert m_bad["nonfinite"] == 2.0 == g.monitor(U2, V, W2)["nonfinite"]


@pytest.mark.parametrize("where", ["interior", "ib"])
def test_time_step_stops_on_nan(where):
    """A non-finite value in the state makes compute_time_step_RK3 return
    C3D_ENAN (the NaN stop of compute_statistics, src/statistics.f90:142-144,
    checked at every step); the steps before it return normally."""
    from channel3d_parallel_amd import ib_geometry
    from channel3d_parallel_amd.solver import C3DError
    nx, ny, nz = 34, 33, 34
    gt = 0 if where == "ib" else 1
    g, U, V, W = state(nx, ny, nz, gt=gt, kind="ic", seed=4)
    with Channel(nx, ny, nz, nu=NU, grid_type=gt) as ch:
        ch.set_fields(U, V, W)
        ch.set_pressure_gradient(DPDX, 0.0, 0.0)
        if where == "ib":
            m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=16, nth=24)
            ch.setup_IB(m)
        ch.compute_time_step_RK3(1e-3)  # finite: no error
        out = ch.get_fields()
        Ub = out["U"]
        Ub[nx // 2, ny // 2, nz // 2] = np.nan
        ch.set_fields(Ub, out["V"], out["W"])
        with pytest.raises(C3DError) as e:
            ch.compute_time_step_RK3(1e-3)
        assert e.value.code == -5, e.value
        assert "non-finite" in str(e.value)
        print(f"  {where}: {e.value}")
