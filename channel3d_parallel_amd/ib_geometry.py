"""Lagrangian marker sets for z-spanning cylinders and the centre wall.

Arrays follow the reference's conventions (src/immersed_boundary_geometry.f90):
``xb, yb, zb, sb`` have ``nb`` entries; ``normals`` and ``ub`` are
component-major ``[x(nb), y(nb), z(nb)]``.  Markers are ordered slab by slab
in z: marker ``i`` of slab ``j`` (both 1-based) is ``l = i + (j-1)*nxb``, the
first ``nxb1`` of a slab belonging to body 1 and the rest to body 2.  That
grouping is what the per-body mean removal of the pressure jump uses.

``2 * 3.14159`` is evaluated as the reference does: an integer times a
default-real (binary32) literal, promoted to FP64 when it meets FP64 data
(src/immersed_boundary_geometry.f90:37-38,129-133).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

# 2 * 3.14159 in binary32, promoted exactly to binary64
TWO_PI_F32 = float(np.float32(2) * np.float32(3.14159))


@dataclass
class Markers:
    xb: np.ndarray
    yb: np.ndarray
    zb: np.ndarray
    normals: np.ndarray
    sb: np.ndarray
    ub: np.ndarray
    nzb: int
    nxb: int
    nxb1: int
    yref: np.ndarray | None = None  # y_ref_index per marker (None: 1)

    @property
    def nb(self) -> int:
        return int(self.xb.size)


def center_wall(Lxp, Lzp, Ly, ny, nxb, nzb):
    """``center_wall`` (src/immersed_boundary_geometry.f90:80-115): a static
    planar wall at y = Ly/2 across the whole channel, nxb x nzb markers at
    the cell-centred positions ((i - 1/2) dxb, Ly/2, (j - 1/2) dzb), area
    dxb dzb, normal (0, -1, 0), at rest, y pivots counted from face ny / 2
    (integer division, :87).  One body: nxb1 = nxb, so the mean removal
    subtracts the plain mean (:410-411).  The reference allows it on the
    uniform grid only (:81)."""
    nb = nxb * nzb
    dxb, dzb = Lxp / nxb, Lzp / nzb
    i = np.tile(np.arange(1, nxb + 1, dtype=np.float64), nzb)
    j = np.repeat(np.arange(1, nzb + 1, dtype=np.float64), nxb)
    normals = np.zeros(3 * nb)
    normals[nb:2 * nb] = -1.0
    return Markers(xb=(i - 0.5) * dxb, yb=np.full(nb, 0.5 * Ly), zb=(j - 0.5) * dzb,
                   normals=normals, sb=np.full(nb, dxb * dzb), ub=np.zeros(3 * nb),
                   nzb=nzb, nxb=nxb, nxb1=nxb, yref=np.full(nb, ny // 2, dtype=np.int32))


def _ring(r, xc, yc, n, dzb, nzb, j, speed, normal_sign):
    dsb = TWO_PI_F32 * r / n
    th = np.arange(n, dtype=np.float64) * dsb / r
    c, s = np.cos(th), np.sin(th)
    return dict(x=r * c + xc, y=r * s + yc, z=np.full(n, (j - 0.5) * dzb),
                ux=-s * speed, uy=c * speed, nx=normal_sign * c, ny=normal_sign * s,
                sb=np.full(n, dsb * dzb))


def _assemble(rings_per_slab, nzb, nxb, nxb1):
    nb = nxb * nzb
    xb, yb, zb, sb = (np.zeros(nb) for _ in range(4))
    normals, ub = np.zeros(3 * nb), np.zeros(3 * nb)
    for j, rings in enumerate(rings_per_slab):
        o = j * nxb
        for rg in rings:
            n = rg["x"].size
            sl = slice(o, o + n)
            xb[sl], yb[sl], zb[sl], sb[sl] = rg["x"], rg["y"], rg["z"], rg["sb"]
            ub[sl], ub[nb + o:nb + o + n] = rg["ux"], rg["uy"]
            normals[sl], normals[nb + o:nb + o + n] = rg["nx"], rg["ny"]
            o += n
    return Markers(xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1)


def double_cylinders_z(dx, dz, Lxp, Ly, Lzp, nzb, r1, r2, speed):
    """Concentric cylinders about (Lxp/2, Ly/2): the inner one (radius r1)
    rotates at tangential speed ``speed`` with normals -r_hat; the outer
    one (radius r2) is static with normals +r_hat
    (src/immersed_boundary_geometry.f90:31-38 marker count, :117-174 set-up)."""
    dzb = Lzp / nzb
    dxb = dzb * dx / dz
    nxb1 = int(TWO_PI_F32 * r1 / dxb)
    nxb2 = int(TWO_PI_F32 * r2 / dxb)
    xc, yc = 0.5 * Lxp, 0.5 * Ly
    slabs = [[_ring(r1, xc, yc, nxb1, dzb, nzb, j, speed, -1.0),
              _ring(r2, xc, yc, nxb2, dzb, nzb, j, 0.0, +1.0)] for j in range(1, nzb + 1)]
    return _assemble(slabs, nzb, nxb1 + nxb2, nxb1)


def two_cylinders(Lzp, nzb=128, nth=96, r=0.2, c1=(2.0, 0.7), c2=(4.0, 1.3),
                  speed1=0.5, speed2=-0.5, normal_sign=1.0):
    """Two separate z-spanning cylinders of radius ``r`` centred at ``c1`` and
    ``c2`` (x, y with the walls at y = 0 and y = Ly), ``nth`` azimuthal by
    ``nzb`` spanwise markers each, rotating at tangential speeds ``speed1``
    and ``speed2`` (counter-rotating by default).  Ordering and marker areas
    follow ``double_cylinders_z``; this body type has no reference counterpart.

    Normals point out of the bodies (``normal_sign=+1``): the Heaviside set-up
    takes the region whose outward normal is ``n`` as the interior, and its
    Poisson solve pins the plane mean next to the bottom wall to zero, so the
    interior must be a region that does not touch the walls.  Then
    ``H_interior`` is the body indicator and ``H_exterior`` the fluid one."""
    dzb = Lzp / nzb
    slabs = [[_ring(r, c1[0], c1[1], nth, dzb, nzb, j, speed1, normal_sign),
              _ring(r, c2[0], c2[1], nth, dzb, nzb, j, speed2, normal_sign)] for j in range(1, nzb + 1)]
    return _assemble(slabs, nzb, 2 * nth, nth)
