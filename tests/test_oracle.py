"""Pins the C oracle (CPU, no GPU) against closed forms, invariants, numpy's
FFT and LAPACK's zgtsv, so that the GPU parity tests check against a
trusted restatement of the reference."""
import numpy as np
import pytest
from scipy.linalg import lapack

from channel3d_parallel_amd import fields
from oracle import orc
from tests import refimpl

RNG = np.random.default_rng(20261202)


def rand_fields(g, rng=RNG):
    U = np.asfortranarray(rng.standard_normal(g.shape_u))
    V = np.asfortranarray(rng.standard_normal(g.shape_v))
    W = np.asfortranarray(rng.standard_normal(g.shape_w))
    g.apply_bc(U, V, W)
    return U, V, W


# ---------------------------------------------------------------- FFT ----
@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 9, 12, 15])
def test_fft_brute_force(n):
    a = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    k = np.arange(n)
    for sign in (-1, 1):
        dft = np.exp(sign * 2j * np.pi * np.outer(k, k) / n) @ a
        assert np.allclose(orc.fft(a, sign), dft, rtol=0, atol=1e-13 * max(1, np.abs(a).sum()))


@pytest.mark.parametrize("n", [16, 32, 48, 64, 96, 128, 192, 256, 384, 512])
def test_fft_vs_numpy(n):
    a = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    ref = np.fft.fft(a)
    assert np.abs(orc.fft(a, -1) - ref).max() <= 1e-14 * np.abs(ref).max()
    refi = np.fft.ifft(a) * n
    assert np.abs(orc.fft(a, +1) - refi).max() <= 1e-14 * np.abs(refi).max()


def test_fft_rejects_unsupported_length():
    with pytest.raises(ValueError):
        orc.fft(np.ones(67 * 2), -1)


# -------------------------------------------------------------- zgtsv ----
@pytest.mark.parametrize("n,scale,cplx", [(2, 1, False), (3, 0.3, False),
                                          (50, 5, False), (50, 0.5, False), (40, 0.5, True),
                                          (64, 0.1, True)])
def test_zgtsv_bitwise_vs_lapack(n, scale, cplx):
    """Same elimination, pivot test and back substitution as LAPACK: bitwise."""
    def r(m):
        v = RNG.standard_normal(m)
        return v + 1j * RNG.standard_normal(m) if cplx else v.astype(complex)
    dl, d, du = r(n - 1), r(n) * scale, r(n - 1)
    b = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    x, info = orc.zgtsv(dl, d, du, b)
    ref = lapack.zgtsv(dl, d, du, b.copy())
    assert info == ref[4] == 0
    if cplx:
        # complex division algorithm differs from the Fortran runtime's
        assert np.abs(x - ref[3]).max() <= 1e-12 * np.abs(ref[3]).max()
    else:
        assert np.array_equal(x, ref[3])


def test_zgtsv_pivoting_happens_and_matches():
    n = 30
    dl = np.full(n - 1, 3.0 + 0j)
    d = np.full(n, 1.0 + 0j)
    du = np.full(n - 1, 0.5 + 0j)
    b = RNG.standard_normal(n) + 0j
    x, info = orc.zgtsv(dl, d, du, b)
    ref = lapack.zgtsv(dl, d, du, b.copy())
    assert info == 0 and np.array_equal(x, ref[3])
    A = np.diag(d) + np.diag(dl, -1) + np.diag(du, 1)
    assert np.linalg.norm(A @ x - b) <= 1e-13 * np.linalg.norm(A) * np.linalg.norm(x)


def test_zgtsv_single_row():
    x, info = orc.zgtsv([], [4.0], [], [2.0 + 6.0j])
    assert info == 0 and x[0] == 0.5 + 1.5j


def test_zgtsv_singular_info():
    x, info = orc.zgtsv(np.zeros(2), np.array([0.0, 1.0, 1.0]), np.ones(2), np.ones(3))
    assert info == 1


# --------------------------------------------------------------- grid ----
@pytest.mark.parametrize("nx,ny,nz,gt", [(34, 65, 34, 1), (194, 129, 194, 1), (130, 129, 130, 0)])
def test_grid_closed_form(nx, ny, nz, gt):
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    x, y, z = refimpl.grid_closed_form(nx, ny, nz, grid_type=gt)
    assert np.allclose(g.get(0), x, rtol=0, atol=1e-14)
    assert np.allclose(g.get(1), y, rtol=0, atol=1e-14)
    assert np.allclose(g.get(2), z, rtol=0, atol=1e-14)
    xg = g.get(3)
    assert np.allclose(xg[1:-1], 0.5 * (x[1:] + x[:-1]), atol=1e-15)
    assert abs(xg[0] + 0.5 * (x[1] - x[0])) < 1e-14  # mirror ghost
    yg = g.get(4)
    assert abs(yg[0] - (2 * y[0] - yg[1])) < 1e-14 and abs(yg[-1] - (2 * y[-1] - yg[-2])) < 1e-14


@pytest.mark.parametrize("ny,mbw,alpha", [(65, 0.1, 2.0), (129, 0.0, 2.6), (97, 0.25, 1.5),
                                           (49, 0.05, 0.0)])
def test_grid_buffered_closed_form(ny, mbw, alpha):
    """grid_type 2 (src/input_output.f90:180-247) against its closed form:
    tanh points 1 + (h/h1) tanh(alpha eta)/tanh(alpha) on the middle
    m = ny - 2(n_u - 1) points, h1 the first spacing of the tanh points on
    [-1, 1], h = Ly h1 / (2 + (n_u - 1) h1), straight lines of spacing h
    beyond; the channel walls y = 0 and y = 2 sit in the middle of the
    buffers; n_u is the smallest count with n_u h >= mbw + 6 h."""
    g = orc.Grid(34, ny, 34, alpha=alpha, grid_type=2, min_buffer_width=mbw)
    nu = g.n_uniform
    y = g.get(1)

    def spacing(n):
        m = ny - 2 * (n - 1)
        t = (lambda e: np.tanh(alpha * e) / np.tanh(alpha)) if alpha > 0 else (lambda e: e)
        h1 = t(2.0 / (m - 1) - 1.0) - t(-1.0)
        return m, t, h1, 2.0 * h1 / (2.0 + (n - 1) * h1)

    m, t, h1, h = spacing(nu)
    eta = 2.0 * np.arange(m) / (m - 1) - 1.0
    ref = np.empty(ny)
    ref[nu - 1:nu - 1 + m] = 1.0 + (h / h1) * t(eta)
    b = np.arange(1, nu)
    ref[nu - 1 - b] = ref[nu - 1] - b * h
    ref[ny - nu + b] = ref[ny - nu] + b * h
    assert np.allclose(y, ref, rtol=0, atol=2e-14)
    assert abs(y[0] + 0.5 * (nu - 1) * h) < 2e-14 and abs(y[-1] - 2.0 - 0.5 * (nu - 1) * h) < 2e-14
    assert nu * h >= mbw + 6 * h
    if nu > 1:
        _, _, _, hp = spacing(nu - 1)
        assert (nu - 1) * hp < mbw + 6 * hp
    # the Python field generator builds the same faces
    _, yf, _ = fields.grid_faces(34, ny, 34, alpha=alpha, grid_type=2, min_buffer_width=mbw)
    assert np.allclose(yf, y, rtol=0, atol=1e-14)


def test_grid_buffered_too_wide():
    """Buffers that cannot fit stop the set-up (src/input_output.f90:199)."""
    with pytest.raises(ValueError):
        orc.Grid(34, 17, 34, grid_type=2, min_buffer_width=1.0)


@pytest.mark.parametrize("ny,dymin", [(65, 4.8674e-3), (129, 2.3606e-3), (193, 1.5579e-3)])
def test_grid_dymin(ny, dymin):
    g = orc.Grid(34, ny, 34)
    assert abs(g.scalars["dymin"] - dymin) < 5e-8


def test_modified_wavenumbers():
    g = orc.Grid(194, 33, 98)
    s = g.scalars
    for arr, n, d in ((g.get(6), 192, s["dx"]), (g.get(7), 96, s["dz"])):
        i = np.arange(n)
        assert np.allclose(arr, -4 * np.sin(np.pi * i / n) ** 2 / d ** 2, rtol=1e-12, atol=1e-9)
        assert np.array_equal(arr[1:], arr[1:][::-1])  # even in the index
        assert arr[0] == 0.0


def test_dyy_is_div_of_grad():
    """Dyy must equal the y-divergence of the y-gradient (wall flux zero)."""
    g = orc.Grid(10, 33, 10)
    phi_y = RNG.standard_normal(g.nym)
    p = g.zeros("rhs")
    p[:, :, :] = phi_y[None, :, None]
    gx, gy, gz = g.gradient(p)
    V = gy
    U = np.zeros(g.shape_u, order="F")
    W = np.zeros(g.shape_w, order="F")
    d = g.divergence(U, V, W)[2, :, 2]
    D, L, Uu = g.get(8), g.get(9), g.get(10)
    ref = D * phi_y
    ref[1:] += L * phi_y[:-1]
    ref[:-1] += Uu * phi_y[1:]
    assert np.abs(d - ref).max() <= 1e-12 * np.abs(ref).max()


# ----------------------------------------------------- div / grad / BC ----
def test_bc_closed_form():
    g = orc.Grid(10, 9, 12)
    U, V, W = (np.asfortranarray(RNG.standard_normal(s)) for s in (g.shape_u, g.shape_v, g.shape_w))
    U0, V0, W0 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    g.apply_bc(U, V, W)
    nx, ny, nz, nxg, nyg, nzg = g.nx, g.ny, g.nz, g.nxg, g.nyg, g.nzg
    # U: x faces periodic, z centres periodic, y centres odd
    assert np.array_equal(U[0, 1:-1, 1:-2], U0[nx - 2, 1:-1, 1:-2])
    assert np.array_equal(U[nx - 1, 1:-1, 1:-2], U0[1, 1:-1, 1:-2])
    assert np.array_equal(U[:, 0, :], -U[:, 1, :]) and np.array_equal(U[:, -1, :], -U[:, -2, :])
    assert np.array_equal(U[:, :, nzg - 2], U[:, :, 1]) and np.array_equal(U[:, :, nzg - 1], U[:, :, 2])
    assert np.array_equal(U[:, :, 0], U[:, :, nzg - 3])
    # V: walls zero, x centres periodic
    assert not V[:, 0, :].any() and not V[:, ny - 1, :].any()
    assert np.array_equal(V[0], V[nxg - 3]) and np.array_equal(V[nxg - 2], V[1]) and np.array_equal(V[nxg - 1], V[2])
    # W: z faces periodic
    assert np.array_equal(W[:, :, 0], W[:, :, nz - 2]) and np.array_equal(W[:, :, nz - 1], W[:, :, 1])
    assert np.array_equal(W[1:-2, 1:-1, 1:-1], W0[1:-2, 1:-1, 1:-1])


def test_divergence_brute_force():
    g = orc.Grid(8, 7, 10)
    U, V, W = rand_fields(g)
    x, y, z = g.get(0), g.get(1), g.get(2)
    d = g.divergence(U, V, W)
    ref = np.zeros(g.shape_rhs, order="F")
    for k in range(2, g.nzg):
        for j in range(2, g.nyg):
            for i in range(2, g.nxg):
                ref[i - 2, j - 2, k - 2] = ((U[i - 1, j - 1, k - 1] - U[i - 2, j - 1, k - 1]) / (x[i - 1] - x[i - 2])
                                            + (V[i - 1, j - 1, k - 1] - V[i - 1, j - 2, k - 1]) / (y[j - 1] - y[j - 2])
                                            + (W[i - 1, j - 1, k - 1] - W[i - 1, j - 1, k - 2]) / (z[k - 1] - z[k - 2]))
    assert np.array_equal(d, ref)
    assert not d[:, :, -1].any()


def test_div_grad_adjoint():
    """sum_cells vol*(D u)*p = -sum_faces vol_f*u*(G p) on unique cells and faces."""
    g = orc.Grid(12, 11, 14)
    U, V, W = rand_fields(g)
    p = np.asfortranarray(RNG.standard_normal(g.shape_rhs))
    nxp, nzp = g.nxp, g.nzp
    p[nxp, :, :] = p[0, :, :]   # periodic duplicates of the centre field
    p[:, :, nzp] = p[:, :, 0]
    x, y, z, xg, yg, zg = (g.get(w) for w in range(6))
    hx, hy, hz = np.diff(x)[1:nxp + 1], np.diff(y), np.diff(z)[1:nzp + 1]
    hgx, hgy, hgz = np.diff(xg)[1:nxp + 1], np.diff(yg)[1:-1], np.diff(zg)[1:nzp + 1]
    d = g.divergence(U, V, W)[:nxp, :, :nzp]
    lhs = np.einsum("ijk,i,j,k->", d * p[:nxp, :, :nzp], hx, hy, hz)
    gx, gy, gz = g.gradient(p)
    rx = np.einsum("ijk,i,j,k->", U[1:nxp + 1, 1:-1, 1:nzp + 1] * gx[1:nxp + 1, 1:-1, 1:nzp + 1], hgx, hy, hz)
    ry = np.einsum("ijk,i,j,k->", V[1:nxp + 1, 1:-1, 1:nzp + 1] * gy[1:nxp + 1, 1:-1, 1:nzp + 1], hx, hgy, hz)
    rz = np.einsum("ijk,i,j,k->", W[1:nxp + 1, 1:-1, 1:nzp + 1] * gz[1:nxp + 1, 1:-1, 1:nzp + 1], hx, hy, hgz)
    assert abs(lhs + (rx + ry + rz)) <= 1e-12 * (abs(rx) + abs(ry) + abs(rz))


def test_total_divergence_vanishes():
    g = orc.Grid(16, 13, 18)
    U, V, W = rand_fields(g)
    d = g.divergence(U, V, W)[: g.nxp, :, : g.nzp]
    vol = np.einsum("i,j,k->ijk", np.diff(g.get(0))[1:g.nxp + 1], np.diff(g.get(1)), np.diff(g.get(2))[1:g.nzp + 1])
    assert abs((d * vol).sum()) <= 1e-13 * np.abs(d * vol).sum()


# ------------------------------------------------------------ Poisson ----
@pytest.mark.parametrize("nx,ny,nz,gt", [(10, 9, 8, 1), (18, 17, 14, 1), (34, 33, 26, 0)])
def test_poisson_vs_numpy_lapack(nx, ny, nz, gt):
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    rhs = np.asfortranarray(RNG.standard_normal(g.shape_rhs))
    ref = refimpl.poisson_numpy(rhs, g.get(6), g.get(7), g.get(8), g.get(9), g.get(10))
    out = g.poisson(rhs.copy(order="F"))
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    # plane nzg is left as given, except the x-periodic duplicate column
    assert np.array_equal(out[:-1, :, -1], rhs[:-1, :, -1])
    assert np.array_equal(out[-1, :, -1], rhs[0, :, -1])


def test_poisson_residual_and_gauge():
    """For a compatible rhs (a divergence), D G phi = rhs and the (0,0) mode
    of phi vanishes at j = 2 (src/poisson.f90:69)."""
    g = orc.Grid(34, 33, 18)
    U, V, W = rand_fields(g)
    rhs = g.divergence(U, V, W)
    phi = g.poisson(rhs.copy(order="F"))
    gx, gy, gz = g.gradient(phi)
    Z = np.zeros_like
    # x/z periodic faces of G phi, V walls are zero in the gradient already
    gx[0], gx[-1] = gx[-2], gx[1]
    gz[:, :, 0], gz[:, :, -1] = gz[:, :, -2], gz[:, :, 1]
    res = g.divergence(gx, gy, gz)[: g.nxp, :, : g.nzp] - rhs[: g.nxp, :, : g.nzp]
    assert np.abs(res).max() <= 1e-12 * np.abs(rhs).max()
    assert abs(phi[: g.nxp, 0, : g.nzp].mean()) <= 1e-12 * np.abs(phi).max()
    del Z


def test_poisson_pressure_side_effect():
    g = orc.Grid(10, 9, 8)
    rhs = np.asfortranarray(RNG.standard_normal(g.shape_rhs))
    P = np.asfortranarray(RNG.standard_normal(g.shape_p))
    dt = 0.0123
    sol = g.poisson(rhs, rk_step=1, dt=dt, P=P)
    inner = sol[:, :, :-1] / (dt * (8.0 / 15.0))
    assert np.array_equal(P[1:-1, 1:-1, 1:-1], inner)
    assert np.array_equal(P[:, :, 0], P[:, :, 1]) and np.array_equal(P[:, :, -1], P[:, :, -2])
    assert np.array_equal(P[:, 0, :], P[:, 1, :]) and np.array_equal(P[0], P[1])
    P2 = np.asfortranarray(RNG.standard_normal(g.shape_p))
    P2c = P2.copy(order="F")
    g.poisson(np.asfortranarray(RNG.standard_normal(g.shape_rhs)), rk_step=2, dt=dt, P=P2)
    assert np.array_equal(P2, P2c)


def test_thomas_agrees_with_zgtsv():
    g = orc.Grid(34, 65, 34)
    U, V, W = rand_fields(g)
    rhs = g.divergence(U, V, W)
    a = g.poisson(rhs.copy(order="F"), tri_mode=0)
    b = g.poisson(rhs.copy(order="F"), tri_mode=1)
    assert np.linalg.norm(a - b) <= 1e-12 * np.linalg.norm(a)


# --------------------------------------------------------- projection ----
def test_projection_divergence_free_and_idempotent():
    g = orc.Grid(34, 65, 34)
    U, V, W = rand_fields(g)
    d0 = g.max_abs_div(U, V, W)
    U0, V0, W0 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    dt = 0.01
    out = g.projection(U, V, W, dt=dt)
    # U_interim is the velocity the projection started from (src/projection.f90:26-28)
    assert np.array_equal(out["U_interim"], U0)
    assert np.array_equal(out["V_interim"], V0)
    assert np.array_equal(out["W_interim"], W0)
    g.apply_bc(U, V, W)
    d1 = g.max_abs_div(U, V, W)
    assert d1 <= 1e-13 * d0
    U1, V1, W1 = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    g.projection(U, V, W, dt=dt)
    g.apply_bc(U, V, W)
    assert np.linalg.norm(U - U1) <= 1e-13 * np.linalg.norm(U1)


def test_projection_matches_composition():
    """P_interim = phi/dt and U = U* - G(dt*phi/dt) exactly as composed."""
    g = orc.Grid(10, 9, 12)
    U, V, W = rand_fields(g)
    U0 = U.copy(order="F")
    dt = 0.37
    rhs = g.divergence(U, V, W)
    phi = g.poisson(rhs.copy(order="F"))
    out = g.projection(U, V, W, dt=dt)
    assert np.array_equal(out["P_interim"], phi / dt)
    gx, _, _ = g.gradient(dt * (phi / dt))
    assert np.array_equal(U, U0 - gx)
