"""Marker sets of the static bodies (channel3d_parallel_amd/ib_geometry.py)
pinned by closed forms: the marker count of compute_nb
(src/immersed_boundary_geometry.f90:31-40), radii and centres, unit normals
(inward on the inner cylinder of double_cylinders_z, outward elsewhere),
tangential wall speeds, slab heights, equal angular spacing, and the total
marker area sum(sb) = (2 * 3.14159 in binary32) * r * Lz per cylinder
(:129-141).  2 * 3.14159 is formed here by doubling the binary32 literal,
an exact operation, independently of the module's product of two binary32
numbers."""
import math

import numpy as np
import pytest

from channel3d_parallel_amd import ib_geometry
from oracle import orc

TWO_PI_32 = 2.0 * float(np.float32(3.14159))


def ring(m, lo, n, nxb):
    """Indices of markers lo .. lo+n-1 of every slab."""
    return np.concatenate([np.arange(j * nxb + lo, j * nxb + lo + n) for j in range(m.nzb)])


def check_ring(m, idx, xc, yc, r, speed, outward):
    nb = m.nb
    dx, dy = m.xb[idx] - xc, m.yb[idx] - yc
    rad = np.hypot(dx, dy)
    assert np.max(np.abs(rad - r)) <= 1e-14 * max(r, 1.0)
    nxv, nyv, nzv = m.normals[idx], m.normals[nb + idx], m.normals[2 * nb + idx]
    assert np.max(np.abs(np.hypot(nxv, nyv) - 1.0)) <= 1e-15 and np.all(nzv == 0.0)
    sgn = 1.0 if outward else -1.0
    assert np.max(np.abs(nxv - sgn * dx / r)) <= 1e-13 and np.max(np.abs(nyv - sgn * dy / r)) <= 1e-13
    ux, uy, uz = m.ub[idx], m.ub[nb + idx], m.ub[2 * nb + idx]
    assert np.all(uz == 0.0)
    # rigid rotation about the axis: u = speed * (-dy, dx) / r, tangent to the ring
    assert np.max(np.abs(ux - speed * (-dy / r))) <= 1e-13 and np.max(np.abs(uy - speed * dx / r)) <= 1e-13


def test_double_cylinders_z_closed_form():
    g = orc.Grid(66, 33, 34, grid_type=0)
    s = g.scalars
    nzb, r1, r2, speed = 16, 0.3, 0.6, 0.7
    m = ib_geometry.double_cylinders_z(s["dx"], s["dz"], s["Lxp"], 2.0, s["Lzp"], nzb, r1, r2, speed)
    dzb = s["Lzp"] / nzb
    dxb = dzb * s["dx"] / s["dz"]
    n1, n2 = math.floor(TWO_PI_32 * r1 / dxb), math.floor(TWO_PI_32 * r2 / dxb)
    assert (m.nxb1, m.nxb, m.nzb) == (n1, n1 + n2, nzb)
    assert m.nb == (n1 + n2) * nzb
    xc, yc = 0.5 * s["Lxp"], 1.0
    inner, outer = ring(m, 0, n1, m.nxb), ring(m, n1, n2, m.nxb)
    check_ring(m, inner, xc, yc, r1, speed, outward=False)
    check_ring(m, outer, xc, yc, r2, 0.0, outward=True)
    # slab heights and the first marker of each ring at theta = 0
    j = np.repeat(np.arange(1, nzb + 1), m.nxb)
    assert np.max(np.abs(m.zb - (j - 0.5) * dzb)) <= 1e-15
    assert np.allclose(m.xb[inner[::n1]], xc + r1, rtol=0, atol=1e-15)
    # equal angular spacing dsb/r = 2 pi_32 / n
    th = np.unwrap(np.arctan2(m.yb[inner[:n1]] - yc, m.xb[inner[:n1]] - xc))
    assert np.max(np.abs(np.diff(th) - TWO_PI_32 / n1)) <= 1e-13
    # marker areas: each ring carries (2 pi_32) r dzb per slab
    assert abs(np.sum(m.sb[inner]) - TWO_PI_32 * r1 * s["Lzp"]) <= 1e-12
    assert abs(np.sum(m.sb[outer]) - TWO_PI_32 * r2 * s["Lzp"]) <= 1e-12


@pytest.mark.parametrize("nth,nzb,r,speed1,speed2", [(96, 128, 0.2, 0.5, -0.5), (24, 16, 0.25, 0.3, -0.6)])
def test_two_cylinders_closed_form(nth, nzb, r, speed1, speed2):
    Lzp = math.pi
    c1, c2 = (2.0, 0.7), (4.0, 1.3)
    m = ib_geometry.two_cylinders(Lzp, nzb=nzb, nth=nth, r=r, c1=c1, c2=c2, speed1=speed1,
                                  speed2=speed2)
    assert (m.nxb1, m.nxb, m.nzb, m.nb) == (nth, 2 * nth, nzb, 2 * nth * nzb)
    a, b = ring(m, 0, nth, m.nxb), ring(m, nth, nth, m.nxb)
    check_ring(m, a, c1[0], c1[1], r, speed1, outward=True)
    check_ring(m, b, c2[0], c2[1], r, speed2, outward=True)
    for idx in (a, b):
        assert abs(np.sum(m.sb[idx]) - TWO_PI_32 * r * Lzp) <= 1e-12
    dzb = Lzp / nzb
    j = np.repeat(np.arange(1, nzb + 1), m.nxb)
    assert np.max(np.abs(m.zb - (j - 0.5) * dzb)) <= 1e-15
