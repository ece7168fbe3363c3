"""Pins of the immersed-boundary oracle (oracle/c3d_oracle_ib.c).

The reference ships no IB golden vectors, so each function is pinned by what
the mathematics fixes:
  deltafnc          closed-form values at r = 0, 1, 2 from both branches,
                    symmetry, discrete moment sums;
  stencils          brute-force scan of every grid node with periodic images
                    (weights and canonical indices), index range invariants;
  E / R             adjointness <E u, sb*f> = dx*dymin*dz <u, R f>, exactness
                    on constants, E1n on linear fields;
  mean removal      per-body sums vanish;
  Heaviside         Hc_interior solves the Poisson problem it is defined by,
                    indicator values inside and outside the bodies;
  Schur             linearity and a numpy recomposition of its steps from the
                    pinned operators;
  BiCGSTAB          against numpy.linalg.solve on the densely assembled Schur
                    matrix of a tiny case;
  IB projection     divergence-free result, slip at the solver tolerance,
                    recomposition from the pinned operators.
"""
import math

import numpy as np
import pytest

from channel3d_parallel_amd import fields, ib_geometry
from oracle import orc

SQ3 = math.sqrt(3.0)


def test_delta_closed_forms():
    d = orc.deltafnc
    assert d(0.0, 0.0, 1.0) == pytest.approx(5.0 / 12.0 + SQ3 * math.pi / 27.0, abs=1e-15)
    phi1 = 7.0 / 24.0 - SQ3 * math.pi / 54.0
    # both branches meet at r = 1 and vanish at r = 2
    assert d(1.0, 0.0, 1.0) == pytest.approx(phi1, abs=1e-15)
    assert d(1.0 + 1e-13, 0.0, 1.0) == pytest.approx(phi1, abs=1e-12)
    assert abs(d(2.0, 0.0, 1.0)) < 1e-15
    assert d(2.0 + 1e-9, 0.0, 1.0) == 0.0
    # symmetric, scaled by 1/dr
    for r in (0.3, 0.77, 1.5, 1.99):
        assert d(r, 0.0, 1.0) == d(-r, 0.0, 1.0)
        assert d(0.5 * r, 0.0, 0.5) == pytest.approx(2.0 * d(r, 0.0, 1.0), rel=1e-14)


def test_delta_discrete_moments():
    """Sum_i phi(i - s) = 1 and Sum_i (i - s) phi(i - s) = 0 for any shift s
    (the smoothed 3-point kernel keeps the zeroth and first discrete moments
    of the 3-point kernel to within its smoothing error)."""
    d = orc.deltafnc
    m0, m1 = [], []
    for s in np.linspace(0.0, 1.0, 41):
        pts = np.arange(-3, 4) - s
        w = np.array([d(p, 0.0, 1.0) for p in pts])
        m0.append(w.sum())
        m1.append((pts * w).sum())
    m0, m1 = np.array(m0), np.array(m1)
    # observed: |m0 - 1| <= 4e-3 and |m1| <= 2e-3, exact at s = 0, 1/2, 1
    assert np.max(np.abs(m0 - 1.0)) < 5e-3
    assert np.max(np.abs(m1)) < 3e-3
    for i in (0, 20, 40):
        assert abs(m0[i] - 1.0) < 1e-14 and abs(m1[i]) < 1e-14


def _case(nx=18, ny=17, nz=18, nth=8, nzb=4, c1=(2.0, 0.7), c2=(4.0, 1.3), r=0.3):
    g = orc.Grid(nx, ny, nz, grid_type=0)
    m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=nzb, nth=nth, r=r, c1=c1, c2=c2)
    return g, m, orc.IB(g, m)


def _node_positions(g, s):
    """Positions of the canonical nodes of stencil s (Fortran indices)."""
    x, y, z = g.get(0), g.get(1), g.get(2)
    xm, ym, zm = 0.5 * (x[:-1] + x[1:]), 0.5 * (y[:-1] + y[1:]), 0.5 * (z[:-1] + z[1:])
    nx, ny, nz = g.nx, g.ny, g.nz
    ii = np.arange(2, nx)  # canonical x: faces 2..nx-1 (u) / centres 2..nxg-2
    xs = x[ii - 1] if s == 0 else xm[ii - 2]
    if s == 1:
        jj = np.arange(1, ny + 1)
        ys = y[jj - 1]
    else:
        jj = np.arange(2, ny + 1)
        ys = ym[jj - 2]
    kk = np.arange(2, nz)  # canonical z: faces 2..nz-1 (w) / centres 2..nzg-2
    zs = z[kk - 1] if s == 2 else zm[kk - 2]
    return ii, xs, jj, ys, kk, zs


@pytest.mark.parametrize("s", [0, 1, 3])
def test_stencil_brute_force(s):
    """Every node of the canonical (periodic) grid within the kernel support of
    a marker appears in its 125-point stencil with the product weight of its
    nearest periodic image; nothing else has a nonzero weight."""
    g, m, ib = _case()
    sc = g.scalars
    dx, dy, dz, Lx, Lz = sc["dx"], sc["dymin"], sc["dz"], sc["Lxp"], sc["Lzp"]
    xi, yi, zi, w, _ = ib.stencil(s)
    ii, xs, jj, ys, kk, zs = _node_positions(g, s)
    for l in range(0, m.nb, 7):
        got = {}
        for q in range(125):
            if w[l, q] != 0.0:
                key = (xi[l, q], yi[l, q], zi[l, q])
                assert key not in got
                got[key] = w[l, q]
        ddx = xs - m.xb[l]
        ddx -= Lx * np.round(ddx / Lx)
        ddz = zs - m.zb[l]
        ddz -= Lz * np.round(ddz / Lz)
        fx = np.array([orc.deltafnc(v, 0.0, dx) for v in ddx])
        fy = np.array([orc.deltafnc(v, m.yb[l], dy) for v in ys])
        fz = np.array([orc.deltafnc(v, 0.0, dz) for v in ddz])
        want = {}
        for a in np.nonzero(fx)[0]:
            for b in np.nonzero(fy)[0]:
                for c in np.nonzero(fz)[0]:
                    want[(ii[a], jj[b], kk[c])] = dx * dy * dz * fx[a] * fy[b] * fz[c]
        assert set(got) == set(want)
        for k in want:
            assert got[k] == pytest.approx(want[k], rel=1e-12, abs=1e-16)


def test_w_stencil_window_shift():
    """The w stencil's x window starts one cell left of the centre stencil's
    (src/immersed_boundary_operators.f90:156); its x indices are the centre
    indices shifted by -1 with periodic wrap."""
    g, m, ib = _case()
    cx = ib.stencil(3)[0]
    wx = ib.stencil(2)[0]
    per = g.nx - 2
    assert np.array_equal((cx - 2 - 1) % per + 2, wx)


def test_stencil_index_ranges():
    g, m, ib = _case(nx=34, ny=33, nz=34, nth=24, nzb=16)
    for s in range(4):
        xi, yi, zi, w, dxn = ib.stencil(s)
        assert xi.min() >= 2 and xi.max() <= g.nx - 1
        assert zi.min() >= 2 and zi.max() <= g.nz - 1
        assert yi.min() >= 1 and yi.max() <= (g.ny if s == 1 else g.nyg)
        # the kernel formula cancels to ~1e-17 (either sign) as r -> 2
        assert np.all(w >= -1e-15 * w.max())
        # dxn = n . (x_node - xb) with |n| = 1 and the node within 3.5 cells
        # of the marker per direction (the w window is shifted one cell)
        assert np.max(np.abs(dxn)) < 3.5 * math.sqrt(g.scalars["dx"] ** 2 + g.scalars["dymin"] ** 2
                                                      + g.scalars["dz"] ** 2)


def test_E_R_adjoint_and_constants():
    g, m, ib = _case(nx=34, ny=33, nz=34, nth=24, nzb=16)
    sc = g.scalars
    vol = sc["dx"] * sc["dymin"] * sc["dz"]
    rng = np.random.default_rng(3)
    U, V, W = (rng.standard_normal(g.zeros(c).shape).copy(order="F") for c in "uvw")
    f = rng.standard_normal(m.nb)
    Eu = ib.E(U, V, W)
    for c, F in enumerate((U, V, W)):
        lhs = np.dot(Eu[c * m.nb:(c + 1) * m.nb], m.sb * f)
        rhs = vol * np.sum(F * ib.R(c, f))
        assert lhs == pytest.approx(rhs, rel=1e-12)
    # constants: E of a constant field is the stencil weight sum
    one = [np.ones(g.zeros(c).shape, order="F") for c in "uvw"]
    E1 = ib.E(*one)
    for c in range(3):
        wsum = ib.stencil(c)[3].sum(axis=1)
        assert np.allclose(E1[c * m.nb:(c + 1) * m.nb], wsum, rtol=1e-14, atol=0)
    # u and v stencils carry the full kernel: weight sums = product of
    # 1-D moment sums, within the kernel's moment error (< 1.3e-2)
    assert np.max(np.abs(E1[:2 * m.nb] - 1.0)) < 1.3e-2


def test_E1n_linear_field():
    """E1n of the linear centre field p = n . (x - xb) for one marker equals
    Sum w * dxn^2; the oracle's regTc_1n sums w * P * dxn."""
    g, m, ib = _case()
    xi, yi, zi, w, dxn = ib.stencil(3)
    l = 5
    P = g.zeros("rhs")
    # place dxn at the stencil nodes of marker l (canonical centre indices)
    for q in range(125):
        P[xi[l, q] - 2, yi[l, q] - 2, zi[l, q] - 2] = dxn[l, q]
    out = ib.E1n_c(P)
    assert out[l] == pytest.approx(np.sum(w[l] * dxn[l] ** 2), rel=1e-13)


def test_remove_mean_per_body():
    g, m, ib = _case()
    f = np.random.default_rng(1).standard_normal(m.nb) + 3.0
    r = ib.remove_mean(f)
    idx = np.arange(m.nb) % m.nxb
    b1, b2 = idx < m.nxb1, idx >= m.nxb1
    assert abs(r[b1].sum()) < 1e-12 and abs(r[b2].sum()) < 1e-12
    assert np.allclose(f[b1] - r[b1], f[b1].mean(), rtol=1e-14)
    assert np.allclose(f[b2] - r[b2], f[b2].mean(), rtol=1e-14)


def test_heaviside_indicator_and_definition():
    g, m, ib = _case(nx=66, ny=65, nz=34, nth=48, nzb=16, r=0.4)
    Hc, Hu, Hv, Hw = ib.heaviside()
    # Hc_interior = L^-1(-D R n): re-derive with the pinned operators
    Rn = [ib.R(c, m.normals[c * m.nb:(c + 1) * m.nb]) for c in range(3)]
    g.apply_bc(*Rn)
    d = -g.divergence(*Rn)
    g.poisson(d, rk_step=0)
    assert np.allclose(1.0 - Hc, d, rtol=0, atol=1e-13)
    # exterior indicator: ~0 at the cylinder centres, ~1 far away
    xg, yg, zg = g.get(3), g.get(4), g.get(5)
    for (xc, yc) in ((2.0, 0.7), (4.0, 1.3)):
        i = int(np.argmin(np.abs(xg - xc))) + 1
        j = int(np.argmin(np.abs(yg - yc))) + 1
        assert abs(Hc[i - 2, j - 2, 3]) < 0.05
    i = int(np.argmin(np.abs(xg - 0.5))) + 1
    j = int(np.argmin(np.abs(yg - 1.0))) + 1
    assert abs(Hc[i - 2, j - 2, 3] - 1.0) < 0.02
    co = ib.heaviside_coeffs()
    assert np.array_equal(co["Hu_int"][1:-1, 1:-1, 1:-1], 1.0 - Hu[1:-1, 1:-1, 1:-1])
    # E1nH / E1nHc are E1n of the exterior fields
    assert np.allclose(co["E1nHc"], ib.E1n_c(Hc), rtol=0, atol=1e-15)
    assert np.allclose(co["E1nH"], ib.E(Hu, Hv, Hw, with_dxn=True), rtol=0, atol=1e-15)


def _schur_numpy(g, m, ib, f, dt, nu):
    nb = m.nb
    co = ib.heaviside_coeffs()
    dudn, pj = f[:3 * nb], f[3 * nb:]
    Fr = [ib.R(c, dt * nu * dudn[c * nb:(c + 1) * nb] - dt * pj * m.normals[c * nb:(c + 1) * nb])
          for c in range(3)]
    g.apply_bc(*Fr)
    phi = g.divergence(*Fr)
    g.poisson(phi, rk_step=2)
    phi = phi / dt
    Sp = ib.remove_mean(-ib.E1n_c(phi)) - co["E1nHc"] * pj
    G = list(g.gradient(dt * phi))
    g.apply_bc(*G)
    G = [G[c] - Fr[c] for c in range(3)]
    Su = ib.E(*G) - co["E1nH"] * dudn
    return np.concatenate([Su, Sp])


def test_schur_linearity_and_composition():
    g, m, ib = _case()
    ib.heaviside()
    rng = np.random.default_rng(7)
    dt, nu = 0.01, 1.0 / 2800.0
    f1, f2 = rng.standard_normal(4 * m.nb), rng.standard_normal(4 * m.nb)
    S1, S2 = ib.schur(f1, dt, nu), ib.schur(f2, dt, nu)
    S12 = ib.schur(2.0 * f1 - 0.5 * f2, dt, nu)
    assert np.allclose(S12, 2.0 * S1 - 0.5 * S2, rtol=0, atol=1e-12 * np.abs(S12).max())
    assert np.allclose(S1, _schur_numpy(g, m, ib, f1, dt, nu), rtol=0, atol=1e-13 * np.abs(S1).max())


def test_bicgstab_vs_dense_solve():
    g, m, ib = _case(nth=6, nzb=4)
    ib.heaviside()
    dt, nu = 0.01, 1.0 / 2800.0
    n = 4 * m.nb
    S = np.zeros((n, n))
    for q in range(n):
        e = np.zeros(n)
        e[q] = 1.0
        S[:, q] = ib.schur(e, dt, nu)
    b = S @ np.random.default_rng(2).standard_normal(n)
    x, it, err = ib.bicgstab(b, tol=1e-11, maxit=2000, dt=dt, nu=nu)
    xd = np.linalg.solve(S, b)
    assert err < 1e-22
    assert np.linalg.norm(S @ x - b) < 1e-10
    cond = np.linalg.cond(S)
    assert np.linalg.norm(x - xd) <= 10 * cond * 1e-11 + 1e-12 * np.linalg.norm(xd)
    print(f"nb={m.nb} iters={it} cond={cond:.3e}")


def test_ib_projection_invariants():
    nx, ny, nz = 34, 33, 34
    g, m, ib = _case(nx=nx, ny=ny, nz=nz, nth=24, nzb=16, r=0.2)
    ib.heaviside()
    U, V, W = fields.channel_ic(nx, ny, nz, seed=1, grid_type=0)
    g.apply_bc(U, V, W)
    dt, nu = 0.005, 1.0 / 2800.0
    out = g.projection(U, V, W, rk_step=2, dt=dt)
    fb = np.zeros(4 * m.nb)
    U0, V0, W0 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    it, err = ib.projection(U, V, W, out["U_interim"], out["V_interim"], out["W_interim"],
                            out["P_interim"], fb, dt, nu, tol=1e-10, maxit=500)
    assert err < 1e-20 and 1 <= it < 500
    assert g.max_abs_div(U, V, W) < 1e-12
    assert ib.max_slip(U, V, W, fb) < 1e-9
    # recomposition: U = U_int - R f - G(dt * (P_int - L^-1 D R f / dt))
    nb = m.nb
    dudn, pj = fb[:3 * nb], fb[3 * nb:]
    Fr = [ib.R(c, dt * nu * dudn[c * nb:(c + 1) * nb] - dt * pj * m.normals[c * nb:(c + 1) * nb])
          for c in range(3)]
    g.apply_bc(*Fr)
    phi = g.divergence(*Fr)
    g.poisson(phi, rk_step=2)
    G = g.gradient(dt * (out["P_interim"] - phi / dt))
    ref = [a - b - c for a, b, c in zip((out["U_interim"], out["V_interim"], out["W_interim"]), Fr, G)]
    g.apply_bc(*ref)
    for a, b in zip((U, V, W), ref):
        assert np.allclose(a, b, rtol=0, atol=1e-13)
    # the unprojected field did not satisfy the slip condition
    assert ib.max_slip(U0, V0, W0, np.zeros(4 * nb)) > 1e-3
