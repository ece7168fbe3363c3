"""Refined projections against the FP64 oracle: the mixed-precision path
(FP32 forward transforms, FP64 tridiagonal solve and inverse transforms,
FP64 residual refinement) and FP64 defect correction.

The refined potential is kept as the first solve's phi1 plus a correction
phic, each differenced on its own in the velocity correction.  Bars: without
refinement the FP32 spectra limit the potential to relative 1e-5; each sweep
cuts the FP64 residual by >= 1e3 until it meets the FP64 floor; one sweep
(the mixed default) brings max|div| to 1e-10 or below and the fields to the
FP64 oracle within relative 1e-11; two sweeps reach the floor (no worse than
the plain FP64 path's max|div|).  The all-FP32 variant (C3D_MIXED_INV64=0)
needs one sweep more.
"""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields, ib_geometry
from channel3d_parallel_amd._lib import PREC_FP64, PREC_MIXED
from oracle import orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nx,ny,nz,gt,rk", [(66, 65, 66, 1, 2), (130, 129, 66, 1, 1), (66, 129, 34, 0, 2)])
def test_mixed_projection(nx, ny, nz, gt, rk):
    print(f"\n[case] mixed {nx}x{ny}x{nz} gt={gt} rk={rk}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    U, V, W = fields.white_faces(nx, ny, nz, seed=2)
    g.apply_bc(U, V, W)
    dt = 0.01
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Pr = g.zeros("p")
    ref = g.projection(Ur, Vr, Wr, rk_step=rk, dt=dt, P=Pr)
    g.apply_bc(Ur, Vr, Wr)
    with Channel(nx, ny, nz, grid_type=gt) as c64:
        c64.set_fields(U, V, W, P=g.zeros("p"))
        c64.compute_non_IB_projection(dt, rk)
        c64.apply_boundary_conditions()
        div64 = c64.check_divergence()
    res = {}
    with Channel(nx, ny, nz, grid_type=gt, precision=PREC_MIXED) as cm:
        for nref in (0, 1, 2, 3):
            cm.set_refinement(nref)
            cm.set_fields(U, V, W, P=g.zeros("p"))
            cm.compute_non_IB_projection(dt, rk)
            cm.apply_boundary_conditions()
            out = cm.get_fields()
            res[nref] = dict(div=cm.check_divergence(), rr=cm.refinement_residuals(),
                             eU=max(rel(out["U"], Ur), rel(out["V"], Vr), rel(out["W"], Wr)),
                             eP=rel(out["P_interim"], ref["P_interim"]),
                             eP1=rel(out["P"], Pr) if rk == 1 else 0.0)
            r = res[nref]
            print(f"  nref={nref}: max|div|={r['div']:.2e} (fp64 {div64:.2e}) "
                  f"relU={r['eU']:.2e} relPint={r['eP']:.2e} relP={r['eP1']:.2e} "
                  f"residuals={np.array2string(r['rr'], precision=2, floatmode='unique')}")
    assert res[0]["eP"] < 1e-5 and res[0]["eU"] < 1e-5
    for nref in (1, 2, 3):
        rr = res[nref]["rr"]
        assert np.all(rr[1:] <= np.maximum(1e-3 * rr[:-1], 4.0 * div64)), rr
        assert res[nref]["eU"] < 1e-11 and res[nref]["eP"] < 1e-11 and res[nref]["eP1"] < 1e-11
    assert res[1]["div"] <= max(1e-10, div64)
    for nref in (2, 3):
        assert res[nref]["div"] <= div64


def test_mixed_projection_384():
    """The 384x192x384-cell workload (386x193x386 faces), tanh alpha 2, the
    seed-3 k^-5/3 random-phase field at unit RMS.  One sweep (the default)
    reaches the FP64 oracle to relative 1e-11 (U, V, W, P_interim) and
    max|div| <= 1e-10, which the plain FP64 path (7.2e-10) does not; every
    sweep cuts the FP64 residual by >= 1e3 until the FP64 floor; the wall
    stress is the FP64 path's; 32 back-to-back solves from the same input
    give the same bits."""
    nx, ny, nz, rk, dt = 386, 193, 386, 2, 1e-3
    print(f"\n[case] mixed {nx}x{ny}x{nz} k^-5/3 seed 3", flush=True)
    g = orc.Grid(nx, ny, nz)
    U, V, W = fields.spectral_faces(nx, ny, nz, 3)
    g.apply_bc(U, V, W)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    ref = g.projection(Ur, Vr, Wr, rk_step=rk, dt=dt)
    g.apply_bc(Ur, Vr, Wr)
    with Channel(nx, ny, nz) as c64:
        c64.set_fields(U, V, W)
        c64.compute_non_IB_projection(dt, rk)
        c64.apply_boundary_conditions()
        div64 = c64.check_divergence()
        tau64 = c64.wall_shear()[0]
    res = {}
    with Channel(nx, ny, nz, precision=PREC_MIXED) as cm:
        for nref in (0, 1, 2):
            cm.set_refinement(nref)
            cm.set_fields(U, V, W)
            cm.compute_non_IB_projection(dt, rk)
            cm.apply_boundary_conditions()
            out = cm.get_fields()
            res[nref] = dict(div=cm.check_divergence(), rr=cm.refinement_residuals(),
                             tau=cm.wall_shear()[0],
                             eU=max(rel(out["U"], Ur), rel(out["V"], Vr), rel(out["W"], Wr)),
                             eP=rel(out["P_interim"], ref["P_interim"]))
            r = res[nref]
            print(f"  nref={nref}: max|div|={r['div']:.2e} (fp64 {div64:.2e}) relU={r['eU']:.2e} "
                  f"relPint={r['eP']:.2e} utau={r['tau']:.15e} (fp64 {tau64:.15e}) "
                  f"residuals={np.array2string(r['rr'], precision=2, floatmode='unique')}", flush=True)
        first = out
        # 32 back-to-back solves of the same input (the 32-solve batch) repeat the bits
        for _ in range(32):
            cm.set_fields(U, V, W)
            cm.compute_non_IB_projection(dt, rk)
        cm.apply_boundary_conditions()
        again = cm.get_fields()
    assert np.array_equal(again["U"], first["U"]) and np.array_equal(again["P_interim"],
                                                                      first["P_interim"])
    # without refinement the FP32 forward transforms leave the potential's
    # error in P_interim and the velocity (amplified by the correction's
    # gradient, 1/dy_min)
    assert res[0]["eP"] < 1e-5 and res[0]["eU"] < 1e-4
    rr = res[2]["rr"]
    assert rr.size == 2 and np.all(rr[1:] <= np.maximum(1e-3 * rr[:-1], 4.0 * div64)), rr
    for nref in (1, 2):
        assert res[nref]["eU"] < 1e-11 and res[nref]["eP"] < 1e-11
        assert res[nref]["div"] <= min(1e-10, div64)
        assert abs(res[nref]["tau"] - tau64) <= 1e-12 * abs(tau64)


def test_fp64_defect_correction_192():
    """FP64 defect correction on the 192x128x192-cell workload (194x129x194
    faces, i.i.d. N(0,1) faces, no-slip walls): one sweep takes max|div| from
    the single solve's 3.5e-11 to the FP64 floor.  On the cells the solve owns
    that floor is the rounding of the residual b - M phi, a few ulps of
    max|b| (2643 here: 4 eps max|b| = 2.3e-12); the reference's divergence
    loop (src/differential_operators.f90:64-72) also evaluates the periodic
    duplicate cells i = nx, k = nz, whose spacing x(nx) - x(nx-1) differs from
    x(2) - x(1) by rounding (relative 5.8e-15 here), so their divergence
    differs from their twins' by that fraction of the x- and z-derivative
    terms, ~2e-12 whatever the velocity.  The fields stay within relative
    1e-12 of the FP64 oracle."""
    nx, ny, nz, rk, dt = 194, 129, 194, 2, 0.01
    print(f"\n[case] fp64 defect correction {nx}x{ny}x{nz} white faces seed 2", flush=True)
    g = orc.Grid(nx, ny, nz)
    U, V, W = fields.white_faces(nx, ny, nz, seed=2)
    g.apply_bc(U, V, W)
    bmax = g.max_abs_div(U, V, W)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    ref = g.projection(Ur, Vr, Wr, rk_step=rk, dt=dt)
    g.apply_bc(Ur, Vr, Wr)
    x, z = g.get(0), g.get(2)
    seam = max(abs((x[nx - 1] - x[nx - 2]) / (x[1] - x[0]) - 1),
               abs((z[nz - 1] - z[nz - 2]) / (z[1] - z[0]) - 1))
    nxp, nzp = nx - 2, nz - 2
    res = {}
    with Channel(nx, ny, nz) as c:
        for nref in (0, 1, 2):
            c.set_refinement(nref)
            c.set_fields(U, V, W)
            c.compute_non_IB_projection(dt, rk)
            c.apply_boundary_conditions()
            out = c.get_fields()
            d = np.abs(c.divergence())
            res[nref] = dict(div=c.check_divergence(), inner=float(d[:nxp, :, :nzp].max()),
                             rr=c.refinement_residuals(),
                             eU=max(rel(out["U"], Ur), rel(out["V"], Vr), rel(out["W"], Wr)),
                             eP=rel(out["P_interim"], ref["P_interim"]))
            r = res[nref]
            print(f"  nref={nref}: max|div|={r['div']:.3e} owned cells {r['inner']:.3e} "
                  f"relU={r['eU']:.2e} relPint={r['eP']:.2e} "
                  f"residuals={np.array2string(r['rr'], precision=2, floatmode='unique')}", flush=True)
    print(f"  periodic seam spacing mismatch {seam:.2e}, input max|div| {bmax:.4e}")
    assert seam > 0.0  # the duplicate cells' spacing is not their twins'
    for nref in (0, 1, 2):
        assert res[nref]["eU"] < 1e-12 and res[nref]["eP"] < 1e-12
    assert res[1]["div"] <= 0.1 * res[0]["div"]
    for nref in (1, 2):
        assert res[nref]["inner"] <= 4.0 * np.finfo(float).eps * bmax
        assert res[nref]["div"] <= 4e-12


def test_mixed_fp32_inverse_variant(monkeypatch):
    """C3D_MIXED_INV64=0 (read at context creation): the inverse half in FP32
    too.  Its first solve is less accurate, so its default is two sweeps,
    which reach the FP64 oracle to relative 1e-11 and the plain FP64 path's
    max|div|."""
    nx, ny, nz, rk, dt = 66, 65, 66, 2, 0.01
    print(f"\n[case] mixed FP32 inverse half {nx}x{ny}x{nz}", flush=True)
    g = orc.Grid(nx, ny, nz)
    U, V, W = fields.white_faces(nx, ny, nz, seed=2)
    g.apply_bc(U, V, W)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    ref = g.projection(Ur, Vr, Wr, rk_step=rk, dt=dt)
    g.apply_bc(Ur, Vr, Wr)
    with Channel(nx, ny, nz) as c64:
        c64.set_fields(U, V, W)
        c64.compute_non_IB_projection(dt, rk)
        c64.apply_boundary_conditions()
        div64 = c64.check_divergence()
    monkeypatch.setenv("C3D_MIXED_INV64", "0")
    with Channel(nx, ny, nz, precision=PREC_MIXED) as cm:
        cm.set_fields(U, V, W)
        cm.compute_non_IB_projection(dt, rk)
        rr = cm.refinement_residuals()
        cm.apply_boundary_conditions()
        out = cm.get_fields()
        div = cm.check_divergence()
    eU = max(rel(out["U"], Ur), rel(out["V"], Vr), rel(out["W"], Wr))
    eP = rel(out["P_interim"], ref["P_interim"])
    print(f"  default sweeps {rr.size}: max|div|={div:.2e} (fp64 {div64:.2e}) relU={eU:.2e} "
          f"relPint={eP:.2e} residuals={np.array2string(rr, precision=2, floatmode='unique')}")
    assert rr.size == 2 and rr[1] <= 1e-3 * rr[0]
    assert eU < 1e-11 and eP < 1e-11
    assert div <= div64


def test_mixed_time_steps_32x64x32():
    """100 RK3 steps of the 32x64x32-cell channel (34x65x34 faces, tanh alpha
    2, Re_b 2800, laminar parabola + 10% divergence-free perturbation,
    constant flow rate, dt from compute_dt at CFL 0.5) with the mixed-precision
    projection (default single sweep) at every stage, against the FP64 oracle:
    the same bar as the FP64 path (relative L2 1e-10 after 100 steps)."""
    nu = 1.0 / 2800.0
    dpdx0 = (180.0 * nu) ** 2
    nx, ny, nz = 34, 65, 34
    print(f"\n[case] mixed RK3 {nx}x{ny}x{nz} 100 steps", flush=True)
    g = orc.Grid(nx, ny, nz)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=1, grid_type=1)
    g.apply_bc(U, V, W)
    q0 = g.mean_mass_flow_u(U)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Pr = g.zeros("p")
    dpdx_r = dpdx0
    with Channel(nx, ny, nz, nu=nu, precision=PREC_MIXED) as ch:
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(dpdx0, 0.0, 0.0)
        for _ in range(100):
            dt = ch.compute_dt(0.5)
            dt_r = g.compute_dt(Ur, Vr, Wr, 0.5, nu)
            ch.compute_time_step_RK3(dt, x_mass_cte=1, q0=q0)
            dpdx_r, _ = g.time_step_rk3(Ur, Vr, Wr, Pr, dt_r, nu, dpdx_r, q0)
        out = ch.get_fields()
        div = ch.check_divergence()
    errs = {n: rel(a, b) for n, a, b in (("U", out["U"], Ur), ("V", out["V"], Vr),
                                         ("W", out["W"], Wr), ("P", out["P"], Pr))}
    print(f"  rel L2 after 100 steps: {errs}; max|div|={div:.2e}")
    for n, e in errs.items():
        assert e <= 1e-10, n
    assert div <= 1e-10


def test_mixed_with_body():
    """The IB projection after a mixed-precision non-IB projection (the Schur
    solve's own Poisson solves stay FP64): two cylinders (24 x 16 markers
    each) in a 34x65x34-face uniform channel, laminar flow, against the same
    projections on an FP64 context.  Both Schur solves stop at the absolute
    tolerance 1e-10, so forces agree to 1e-8 relative and fields to 1e-10;
    slip and divergence meet the FP64 path's bars."""
    nu = 1.0 / 2800.0
    nx, ny, nz, dt, tol = 34, 65, 34, 0.005, 1e-10
    print(f"\n[case] mixed + IB {nx}x{ny}x{nz}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=0)
    m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=16, nth=24, r=0.2)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.0, grid_type=0)
    g.apply_bc(U, V, W)
    res = {}
    for name, prec in (("fp64", PREC_FP64), ("mixed", PREC_MIXED)):
        with Channel(nx, ny, nz, grid_type=0, nu=nu, precision=prec) as ch:
            ch.setup_IB(m)
            ch.set_fields(U, V, W)
            ch.set_IB_solver(tol=tol, maxit=1000, guess_mode=1)
            ch.compute_non_IB_projection(dt, 2)
            ch.apply_boundary_conditions()
            it, err = ch.compute_IB_projection(dt, 2)
            res[name] = dict(f=ch.get_fields(), fb=ch.get_IB_force(), it=it, err=err,
                             slip=ch.check_slip(), div=ch.check_divergence())
            print(f"  {name}: iterations {it}, slip {res[name]['slip']:.2e}, "
                  f"max|div| {res[name]['div']:.2e}")
    a, b = res["mixed"], res["fp64"]
    assert a["err"] < tol ** 2 and b["err"] < tol ** 2
    assert rel(a["fb"], b["fb"]) < 1e-8
    for n in ("U", "V", "W"):
        assert rel(a["f"][n], b["f"][n]) < 1e-10, n
    assert a["slip"] < 1e-9 and a["div"] < 1e-12
