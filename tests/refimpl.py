"""Independent numpy/scipy statements of the reference formulas, used to pin
the C oracle (tests only).

The Poisson pipeline here calls numpy.fft.fft2/ifft2 (the DFT the
reference gets from FFTW, src/poisson.f90:43-47,78-82) and scipy's LAPACK
``zgtsv`` (the routine the reference calls, src/poisson.f90:72), so it
shares no code with the oracle's own FFT and tridiagonal restatements.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack


def grid_closed_form(nx, ny, nz, Lxp=2 * np.pi, Lzp=np.pi, Ly=2.0, alpha=2.0, grid_type=1):
    """Face coordinates in closed form (src/input_output.f90:146-178)."""
    x = Lxp * np.arange(nx) / (nx - 2)
    z = Lzp * np.arange(nz) / (nz - 2)
    eta = 2.0 * np.arange(ny) / (ny - 1) - 1.0
    if grid_type == 0:
        y = Ly * np.arange(ny) / (ny - 1)
    else:
        y = 0.5 * Ly * (1.0 + np.tanh(alpha * eta) / np.tanh(alpha))
    return x, y, z


def poisson_numpy(rhs, kxx, kzz, dyy_d, dyy_l, dyy_u):
    """solve_poisson_equation with numpy.fft + LAPACK zgtsv (src/poisson.f90:33-88).

    rhs: rhs_p-shaped array (nxp+1, nym, nzp+2) in the reference layout.
    Returns the solution with the periodic copies applied and the last
    z plane left as given."""
    nxp, nzp, nym = kxx.size, kzz.size, dyy_d.size
    out = np.array(rhs, copy=True, order="F")
    hat = np.empty((nxp, nym, nzp), dtype=np.complex128)
    for j in range(nym):
        hat[:, j, :] = np.fft.fft2(out[:nxp, j, :nzp].astype(np.complex128))
    for k in range(nzp):
        for i in range(nxp):
            d = (dyy_d + kxx[i]) + kzz[k]
            d = d.astype(np.complex128)
            if i == 0 and k == 0:
                d[0] = 1.5 * d[0]
            r = lapack.zgtsv(dyy_l.astype(np.complex128), d, dyy_u.astype(np.complex128),
                             hat[i, :, k].copy())
            hat[i, :, k] = r[3]
    for j in range(nym):
        plane = np.fft.ifft2(hat[:, j, :]) * (nxp * nzp)
        out[:nxp, j, :nzp] = plane.real / float(nxp * nzp)
    out[nxp, :, :] = out[0, :, :]
    out[:, :, nzp] = out[:, :, 0]
    return out


# ---------------------------------------------------------------------------
# RK3 right-hand sides in the reference's array-pass form (src/equations.f90):
# term arrays of shape (nxg, nyg, nzg), interpolation into array sections,
# whole-array squares and products.  1-based Fortran index i maps to i-1.
# ---------------------------------------------------------------------------
def _interp_x(u):
    n1 = u.shape[0]
    out = np.zeros_like(u)
    out[: n1 - 1] = 0.5 * (u[: n1 - 1] + u[1:n1])
    return out


def _interp_y(u):
    n2 = u.shape[1]
    out = np.zeros_like(u)
    out[:, : n2 - 1] = 0.5 * (u[:, : n2 - 1] + u[:, 1:n2])
    return out


def _interp_z(u):
    n3 = u.shape[2]
    out = np.zeros_like(u)
    out[:, :, : n3 - 1] = 0.5 * (u[:, :, : n3 - 1] + u[:, :, 1:n3])
    return out


def rhs_arraypass(which, U, V, W, x, y, z, xg, yg, zg, nu, dpd):
    """compute_rhs_u/v/w following the reference's whole-array passes."""
    nx, nyg, nzg = U.shape
    nxg, ny, nz = nx + 1, nyg - 1, nzg - 1
    t1 = np.zeros((nxg, nyg, nzg))
    t2 = np.zeros((nxg, nyg, nzg))
    X = lambda a, i: a[i - 1]  # noqa: E731
    if which == "u":
        # 1) u interpolated to x-centres into term_1(2:nxg-1, 1:nyg, 1:nzg)
        t1[1:nxg - 1, :, :] = _interp_x(U)[: nx - 1]
        t1 = t1 * t1
        I, J, K = np.meshgrid(np.arange(2, nx), np.arange(2, nyg), np.arange(2, nzg), indexing="ij")
        A = -(t1[I, J - 1, K - 1] - t1[I - 1, J - 1, K - 1]) / (xg[I] - xg[I - 1])
        # 2) uv at y faces: term_1(1:nx,1:ny,1:nzg), term_2(1:nx,1:ny,1:nzg)
        t1[:nx, :ny, :] = _interp_y(U)[:, :ny, :]
        t2[:nx, :ny, :] = _interp_x(V)[:nx, :, :]
        t1 = t1 * t2
        B = -(t1[I - 1, J - 1, K - 1] - t1[I - 1, J - 2, K - 1]) / (y[J - 1] - y[J - 2])
        # 3) uw at z faces
        t1[:nx, :, :nz] = _interp_z(U)[:, :, :nz]
        t2[:nx, :, :nz] = _interp_x(W)[:nx, :, :]
        t1 = t1 * t2
        C = -(t1[I - 1, J - 1, K - 1] - t1[I - 1, J - 1, K - 2]) / (z[K - 1] - z[K - 2])
        # 4) viscous
        t1[:nx, : nyg - 1, :] = (U[:, 1:nyg, :] - U[:, : nyg - 1, :]) / (yg[1:nyg] - yg[: nyg - 1])[None, :, None]
        dx_1, dx_2, dx_3 = x[I - 1] - x[I - 2], x[I] - x[I - 1], xg[I] - xg[I - 1]
        dy_3 = y[J - 1] - y[J - 2]
        dz_1, dz_2, dz_3 = zg[K - 1] - zg[K - 2], zg[K] - zg[K - 1], z[K - 1] - z[K - 2]
        Uc = U[I - 1, J - 1, K - 1]
        Vis = (1.0 / dx_3 * (nu * 1.0 / dx_2 * (U[I, J - 1, K - 1] - Uc) - nu * 1.0 / dx_1 * (Uc - U[I - 2, J - 1, K - 1]))
               + 1.0 / dy_3 * (nu * t1[I - 1, J - 1, K - 1] - nu * t1[I - 1, J - 2, K - 1])
               + 1.0 / dz_3 * (nu * 1.0 / dz_2 * (U[I - 1, J - 1, K] - Uc) - nu * 1.0 / dz_1 * (Uc - U[I - 1, J - 1, K - 2])))
    elif which == "v":
        t1[:, 1:nyg - 1, :] = _interp_y(V)[:, : ny - 1, :]
        t1 = t1 * t1
        I, J, K = np.meshgrid(np.arange(2, nxg), np.arange(2, ny), np.arange(2, nzg), indexing="ij")
        A = -(t1[I - 1, J, K - 1] - t1[I - 1, J - 1, K - 1]) / (yg[J] - yg[J - 1])
        t1[:nx, :ny, :] = _interp_y(U)[:, :ny, :]
        t2[:nx, :ny, :] = _interp_x(V)[:nx, :, :]
        t1 = t1 * t2
        B = -(t1[I - 1, J - 1, K - 1] - t1[I - 2, J - 1, K - 1]) / (x[I - 1] - x[I - 2])
        t1[:, :ny, :nz] = _interp_z(V)[:, :, :nz]
        t2[:, :ny, :nz] = _interp_y(W)[:, :ny, :]
        t1 = t1 * t2
        C = -(t1[I - 1, J - 1, K - 1] - t1[I - 1, J - 1, K - 2]) / (z[K - 1] - z[K - 2])
        dx_1, dx_2, dx_3 = xg[I - 1] - xg[I - 2], xg[I] - xg[I - 1], x[I - 1] - x[I - 2]
        dy_1, dy_2, dy_3 = y[J - 1] - y[J - 2], y[J] - y[J - 1], yg[J] - yg[J - 1]
        dz_1, dz_2, dz_3 = zg[K - 1] - zg[K - 2], zg[K] - zg[K - 1], z[K - 1] - z[K - 2]
        Vc = V[I - 1, J - 1, K - 1]
        Vis = (1.0 / dx_3 * (nu * 1.0 / dx_2 * (V[I, J - 1, K - 1] - Vc) - nu * 1.0 / dx_1 * (Vc - V[I - 2, J - 1, K - 1]))
               + 1.0 / dy_3 * (nu * 1.0 / dy_2 * (V[I - 1, J, K - 1] - Vc) - nu * 1.0 / dy_1 * (Vc - V[I - 1, J - 2, K - 1]))
               + 1.0 / dz_3 * (nu * 1.0 / dz_2 * (V[I - 1, J - 1, K] - Vc) - nu * 1.0 / dz_1 * (Vc - V[I - 1, J - 1, K - 2])))
    else:
        t1[:, :, 1:nz] = _interp_z(W)[:, :, : nz - 1]
        t1 = t1 * t1
        I, J, K = np.meshgrid(np.arange(2, nxg), np.arange(2, nyg), np.arange(2, nz), indexing="ij")
        A = -(t1[I - 1, J - 1, K] - t1[I - 1, J - 1, K - 1]) / (zg[K] - zg[K - 1])
        t1[:nx, :, :nz] = _interp_z(U)[:, :, :nz]
        t2[:nx, :, :nz] = _interp_x(W)[:nx, :, :]
        t1 = t1 * t2
        B = -(t1[I - 1, J - 1, K - 1] - t1[I - 2, J - 1, K - 1]) / (x[I - 1] - x[I - 2])
        t1[:, :ny, :nz] = _interp_z(V)[:, :, :nz]
        t2[:, :ny, :nz] = _interp_y(W)[:, :ny, :]
        t1 = t1 * t2
        C = -(t1[I - 1, J - 1, K - 1] - t1[I - 1, J - 2, K - 1]) / (y[J - 1] - y[J - 2])
        t1[:, : nyg - 1, :nz] = (W[:, 1:nyg, :] - W[:, : nyg - 1, :]) / (yg[1:nyg] - yg[: nyg - 1])[None, :, None]
        dx_1, dx_2, dx_3 = xg[I - 1] - xg[I - 2], xg[I] - xg[I - 1], x[I - 1] - x[I - 2]
        dy_3 = y[J - 1] - y[J - 2]
        dz_1, dz_2, dz_3 = z[K - 1] - z[K - 2], z[K] - z[K - 1], zg[K] - zg[K - 1]
        Wc = W[I - 1, J - 1, K - 1]
        Vis = (1.0 / dx_3 * (nu * 1.0 / dx_2 * (W[I, J - 1, K - 1] - Wc) - nu * 1.0 / dx_1 * (Wc - W[I - 2, J - 1, K - 1]))
               + 1.0 / dy_3 * (nu * t1[I - 1, J - 1, K - 1] - nu * t1[I - 1, J - 2, K - 1])
               + 1.0 / dz_3 * (nu * 1.0 / dz_2 * (W[I - 1, J - 1, K] - Wc) - nu * 1.0 / dz_1 * (Wc - W[I - 1, J - 1, K - 2])))
    del X
    r = A
    r = r + B
    r = r + C
    r = r + Vis
    r = r + dpd * 1.0
    return np.asfortranarray(r)
