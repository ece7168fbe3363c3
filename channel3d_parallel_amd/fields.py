"""Seeded synthetic input fields (host side, numpy).

Both the GPU path and the oracle consume the same bytes produced here.

* SplitMix64 (Steele, Lea, Flood 2014) gives the uniform stream:
  z_n = seed + n*0x9E3779B97F4A7C15, then the two xor-shift-multiply mixes;
  u = (z >> 11) * 2^-53 in [0, 1).
* Box-Muller turns consecutive uniform pairs into normals.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n raw 64-bit outputs starting at stream position `offset`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform(seed: int, n: int, offset: int = 0) -> np.ndarray:
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n standard normals (Box-Muller on consecutive uniform pairs)."""
    m = (n + 1) // 2
    u = uniform(seed, 2 * m, 2 * offset)
    u1 = 1.0 - u[0::2]  # (0, 1]
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2.0 * np.pi * u2)
    z[1::2] = r * np.sin(2.0 * np.pi * u2)
    return z[:n]


def shapes(nx, ny, nz):
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    return (nx, nyg, nzg), (nxg, ny, nzg), (nxg, nyg, nz)


def white_faces(nx, ny, nz, seed=2):
    """I.i.d. N(0,1) per face, filled in column-major order over the
    ghost-inclusive U, V, W arrays (one stream: U, then V, then W).
    Boundary conditions are not applied."""
    su, sv, sw = shapes(nx, ny, nz)
    nu, nv, nw = np.prod(su), np.prod(sv), np.prod(sw)
    z = normal(seed, nu + nv + nw)
    U = np.asfortranarray(z[:nu].reshape(su, order="F"))
    V = np.asfortranarray(z[nu:nu + nv].reshape(sv, order="F"))
    W = np.asfortranarray(z[nu + nv:].reshape(sw, order="F"))
    return U, V, W


def _buffered_y(ny, Ly, alpha, min_buffer_width):
    """grid_type 2 (src/input_output.f90:180-247): tanh-clustered interior
    with n_uniform - 1 extra points of the smallest spacing beyond each end,
    n_uniform the first count whose buffer holds min_buffer_width + 6 dymin,
    shifted by +1."""
    nbuf = 1
    while True:
        m = ny - 2 * (nbuf - 1)
        s = np.arange(m, dtype=np.float64)
        s = 2.0 * s / s[-1] - 1.0
        if alpha > 0:
            s = np.tanh(alpha * s) / np.tanh(alpha)
        h_unit = s[1] - s[0]
        h = Ly * h_unit / (2.0 + (nbuf - 1) * h_unit)
        y = np.zeros(ny)
        y[nbuf - 1:nbuf - 1 + m] = s * h / h_unit
        for b in range(1, nbuf):
            y[ny - nbuf + b] = y[ny - nbuf + b - 1] + h
            y[nbuf - 1 - b] = y[nbuf - b] - h
        h1 = y[1] - y[0]
        if nbuf * h1 >= min_buffer_width + 6 * h1:
            return y + 1.0
        nbuf += 1
        if 2 * nbuf >= ny:
            raise ValueError("uniform buffers exceed the grid")


def grid_faces(nx, ny, nz, Lxp=2 * np.pi, Lzp=np.pi, Ly=2.0, alpha=2.0, grid_type=1,
               min_buffer_width=0.0):
    """Face coordinates of create_grid (src/input_output.f90:140-247)."""
    x = Lxp * np.arange(nx, dtype=np.float64) / (nx - 2)
    z = Lzp * np.arange(nz, dtype=np.float64) / (nz - 2)
    y = Ly * np.arange(ny, dtype=np.float64) / (ny - 1)
    if grid_type == 1:
        y = y - 1.0
        if alpha > 0:
            y = np.tanh(alpha * y) / np.tanh(alpha)
        y = y - y.min()
        y = y * Ly / y.max()
    elif grid_type == 2:
        y = _buffered_y(ny, Ly, alpha, min_buffer_width)
    return x, y, z


def channel_ic(nx, ny, nz, Ub=1.0, pert=0.1, seed=1, Lxp=2 * np.pi, Lzp=np.pi, Ly=2.0, alpha=2.0,
               grid_type=1, min_buffer_width=0.0):
    """Laminar parabola with bulk velocity Ub plus a divergence-free
    perturbation of RMS pert*Ub.  On the buffered grid (grid_type 2) the
    parabola spans the channel 0 < y < Ly and the buffers beyond the walls
    start at rest, as init_flow sets them (src/input_output.f90:274-280).

    The perturbation is the discrete curl of a vector potential A whose
    components live on cell edges (Ax on (x-centre, y-face, z-face), Ay on
    (x-face, y-centre, z-face), Az on (x-face, y-face, z-centre)), i.i.d.
    N(0,1) from the SplitMix64 stream `seed`, periodic in x and z and zero on
    the walls.  The curl uses the face spacings that the divergence uses, so
    the discrete divergence of the perturbation telescopes to round-off.
    Boundary conditions are not applied."""
    x, y, z = grid_faces(nx, ny, nz, Lxp, Lzp, Ly, alpha, grid_type, min_buffer_width)
    nxg, nyg, nzg = nx + 1, ny + 1, nz + 1
    nxp, nzp = nx - 2, nz - 2
    hx = np.zeros(nxg + 1)
    hx[2:nx + 1] = x[1:] - x[:-1]  # hx[i] = x(i) - x(i-1), Fortran index
    hy = np.zeros(nyg + 1)
    hy[2:ny + 1] = y[1:] - y[:-1]
    hz = np.zeros(nzg + 1)
    hz[2:nz + 1] = z[1:] - z[:-1]
    # periodic potentials on the unique index box; A[i,j,k] for i in 0..nxp-1 (x index 2..nxp+1)
    n = nxp * ny * nzp
    a = normal(seed, 3 * n)
    Ax = a[:n].reshape((nxp, ny, nzp), order="F").copy()
    Ay = a[n:2 * n].reshape((nxp, ny, nzp), order="F").copy()
    Az = a[2 * n:].reshape((nxp, ny, nzp), order="F").copy()
    Ax[:, 0, :] = Ax[:, -1, :] = 0.0  # y-face components vanish on the walls
    Az[:, 0, :] = Az[:, -1, :] = 0.0

    def px(A, s):  # periodic shift in x: value at index i-s
        return np.roll(A, s, axis=0)

    def pz(A, s):
        return np.roll(A, s, axis=2)

    # unique faces/centres: x index i = 2..nxp+1 -> 0..nxp-1; z likewise;
    # y: faces j = 1..ny -> 0..ny-1, centres j = 2..ny -> use face pairs
    hxu = hx[2:nxp + 2][:, None, None]
    hzu = hz[2:nzp + 2][None, None, :]
    hyc = hy[2:ny + 1][None, :, None]  # spacing of y centre j = 2..ny
    # U at (x-face i, y-centre j, z-centre k): (Az(j) - Az(j-1))/hy_j - (Ay(k) - Ay(k-1))/hz_k
    Up = (Az[:, 1:, :] - Az[:, :-1, :]) / hyc - (Ay[:, 1:, :] - pz(Ay, 1)[:, 1:, :]) / hzu
    # V at (x-centre i, y-face j, z-centre k): (Ax(k) - Ax(k-1))/hz_k - (Az(i) - Az(i-1))/hx_i
    Vp = (Ax - pz(Ax, 1)) / hzu - (Az - px(Az, 1)) / hxu
    # W at (x-centre i, y-centre j, z-face k): (Ay(i) - Ay(i-1))/hx_i - (Ax(j) - Ax(j-1))/hy_j
    Wp = (Ay[:, 1:, :] - px(Ay, 1)[:, 1:, :]) / hxu - (Ax[:, 1:, :] - Ax[:, :-1, :]) / hyc
    rms = np.sqrt((np.mean(Up ** 2) + np.mean(Vp[:, 1:-1, :] ** 2) + np.mean(Wp ** 2)) / 3.0)
    s = pert * Ub / rms
    U = np.zeros((nx, nyg, nzg), order="F")
    V = np.zeros((nxg, ny, nzg), order="F")
    W = np.zeros((nxg, nyg, nz), order="F")
    ym = 0.5 * (y[:-1] + y[1:])  # centres yg(2..ny)
    lam = 1.5 * Ub * ym * (Ly - ym) / (0.25 * Ly * Ly)  # parabola with bulk Ub
    lam = np.where((ym > 0.0) & (ym < Ly), lam, 0.0)
    U[1:nxp + 1, 1:ny, 1:nzp + 1] = lam[None, :, None] + s * Up
    V[1:nxp + 1, :, 1:nzp + 1] = s * Vp
    W[1:nxp + 1, 1:ny, 1:nzp + 1] = s * Wp
    return U, V, W


def spectral_faces(nx, ny, nz, seed=3, slope=-11.0 / 6.0):
    """Random-phase fields with |u_hat(k)| ~ |k|^slope on the index box
    (nxp, nym, nzp) (slope -11/6 gives an E(k) ~ k^-5/3 shell spectrum),
    each component rescaled to unit RMS and placed on the unique interior
    faces.  Boundary conditions are not applied."""
    nxp, nym, nzp = nx - 2, ny - 1, nz - 2
    su, sv, sw = shapes(nx, ny, nz)
    kx = np.fft.fftfreq(nxp) * nxp
    ky = np.fft.fftfreq(nym) * nym
    kz = np.fft.rfftfreq(nzp) * nzp
    kk = np.sqrt(kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2)
    amp = np.where(kk > 0, kk, 1.0) ** slope
    amp[0, 0, 0] = 0.0
    out = []
    ncell = nxp * nym * nzp
    for c in range(3):
        white = normal(seed, ncell, offset=c * ncell).reshape((nxp, nym, nzp), order="F")
        f = np.fft.irfftn(np.fft.rfftn(white) * amp, s=(nxp, nym, nzp), axes=(0, 1, 2))
        f /= np.sqrt(np.mean(f * f))
        out.append(f)
        del white
    U = np.zeros(su, order="F")
    V = np.zeros(sv, order="F")
    W = np.zeros(sw, order="F")
    U[1:nx - 1, 1:ny, 1:nzp + 1] = out[0]
    V[1:nxp + 1, 1:ny - 1, 1:nzp + 1] = out[1][:, : ny - 2, :]
    W[1:nxp + 1, 1:ny, 1:nz - 1] = out[2]
    return U, V, W
