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
        f = np.fft.irfftn(np.fft.rfftn(white) * amp, s=(nxp, nym, nzp))
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
