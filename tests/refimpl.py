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
