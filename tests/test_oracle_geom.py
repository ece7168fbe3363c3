"""The oracle's moving-wall geometry (setup_IB_geometry for standing_wave and
traveling_wave_x, src/immersed_boundary_geometry.f90:176-390) and the
buffered grid it lives on, against closed forms: markers on the wall curve,
marker velocities equal to the time derivative of the positions, unit
normals orthogonal to the wall, areas that add up to the polygon length,
the traveling wave's ramp, and delta stencils that sum to one on the
uniform buffer (the y_ref_index pivots)."""
import numpy as np
import pytest

from oracle import orc

NX, NY, NZ, MBW = 34, 65, 18, 0.1


@pytest.fixture(scope="module")
def grid():
    return orc.Grid(NX, NY, NZ, alpha=2.6, grid_type=2, min_buffer_width=MBW)


def curve(body, x, t, p1, p2, p3, Lxp, ramp=0.0, top=False):
    """Wall position and its x derivative in closed form."""
    if body == "standing_wave":
        k = 2 * np.pi * p3 / Lxp
        y = p1 * np.sin(k * x) * np.cos(p2 * t)
        return y, p1 * k * np.cos(k * x) * np.cos(p2 * t)
    amp = p1 * 0.5 * (1 - np.cos(np.pi * t / ramp)) if (ramp > 0 and t < ramp) else p1
    ph = p3 * (x - p2 * t) + (np.pi if top else 0.0)
    return amp / (p2 * p3) * np.sin(ph), amp / p2 * np.cos(ph)


CASES = [("standing_wave", 0.04, 0.7, 1.0, 0.0, 0.3), ("standing_wave", 0.03, 2.0, 2.0, 0.0, 1.1),
         ("traveling_wave_x", 0.03, 1.0, 1.0, 0.0, 0.4),
         ("traveling_wave_x", 0.02, 1.5, 1.0, 2.0, 0.5)]


@pytest.mark.parametrize("body,p1,p2,p3,ramp,t", CASES)
def test_wave_markers_closed_form(grid, body, p1, p2, p3, ramp, t):
    nxb, nzb = 24, 4
    m = orc.WaveMarkers(grid, body, nxb, nzb, p1, p2, p3, t, ramp)
    assert m.stop == 0
    Lxp, Lzp = grid.scalars["Lxp"], grid.scalars["Lzp"]
    nb = m.nb
    xw = (np.arange(1, nxb + 1) - 0.5) * Lxp / nxb
    for j in range(nzb):
        for wall in (0, 1):
            sl = slice(2 * nxb * j + wall * nxb, 2 * nxb * j + (wall + 1) * nxb)
            y0, dydx = curve(body, xw, t, p1, p2, p3, Lxp, ramp, top=bool(wall))
            assert np.allclose(m.xb[sl], xw, rtol=0, atol=1e-15)
            assert np.allclose(m.zb[sl], (j + 0.5) * Lzp / nzb, rtol=0, atol=1e-15)
            assert np.allclose(m.yb[sl], y0 + 2.0 * wall, rtol=0, atol=1e-15)
            assert np.all(m.yref[sl] == (NY if wall else 1))
            n = np.stack([m.normals[sl], m.normals[nb:][sl], m.normals[2 * nb:][sl]])
            assert np.allclose(np.linalg.norm(n, axis=0), 1.0, atol=1e-15)
            assert np.allclose(n[0] + n[1] * dydx, 0.0, atol=1e-14)  # orthogonal to (1, y')
            assert np.all(n[1] < 0) if wall == 0 else np.all(n[1] > 0)  # away from the fluid
            assert np.all(n[2] == 0.0) and np.all(m.ub[sl] == 0.0) and np.all(m.ub[2 * nb:][sl] == 0)
            # areas: each chord of the closed periodic polygon is shared by its two ends
            xs, ys = m.xb[sl], m.yb[sl]
            chords = np.hypot(np.diff(np.r_[xs, xs[0] + Lxp]), np.diff(np.r_[ys, ys[0]]))
            assert abs(m.sb[sl].sum() - Lzp / nzb * chords.sum()) < 1e-14
    # velocities: d yb / dt by central differences
    h = 1e-5
    a = orc.WaveMarkers(grid, body, nxb, nzb, p1, p2, p3, t + h, ramp)
    b = orc.WaveMarkers(grid, body, nxb, nzb, p1, p2, p3, t - h, ramp)
    fd = (a.yb - b.yb) / (2 * h)
    assert np.allclose(m.ub[nb:2 * nb], fd, rtol=0, atol=1e-9 * max(1.0, np.abs(fd).max()))


def test_standing_wave_flat_walls(grid):
    """Amplitude 0: flat walls at y = 0 and y = 2 with areas dxb dzb (the
    flat-wall channel of examples/input_parameters_ib_mfu)."""
    nxb, nzb = 16, 5
    m = orc.WaveMarkers(grid, "standing_wave", nxb, nzb, 0.0, 0.0, 0.0, 0.0)
    s = grid.scalars
    assert np.all(m.yb[m.yref == 1] == 0.0) and np.all(m.yb[m.yref == NY] == 2.0)
    assert np.allclose(m.sb, s["Lxp"] / nxb * s["Lzp"] / nzb, rtol=1e-14, atol=0)
    assert np.all(m.normals[:m.nb] == 0.0)


def test_traveling_wave_ramp(grid):
    p1, p2, p3, ramp = 0.02, 1.5, 1.0, 2.0
    m0 = orc.WaveMarkers(grid, "traveling_wave_x", 8, 2, p1, p2, p3, 0.0, ramp)
    assert np.all(m0.yb[m0.yref == 1] == 0.0) and np.all(m0.ub == 0.0)
    half = orc.WaveMarkers(grid, "traveling_wave_x", 8, 2, p1, p2, p3, 1.0, ramp)
    full = orc.WaveMarkers(grid, "traveling_wave_x", 8, 2, p1, p2, p3, 1.0, 0.0)
    bot = half.yref == 1
    assert np.allclose(half.yb[bot], 0.5 * full.yb[bot], rtol=0, atol=1e-16)


def test_wave_outside_buffer_stops(grid):
    """Markers below y(3) stop the set-up (src/immersed_boundary_geometry.f90:199-207)."""
    y = grid.get(1)
    p1 = 1.1 * abs(y[2])  # the lowest of 16 markers reaches 0.98 p1 below y = 0
    m = orc.WaveMarkers(grid, "standing_wave", 16, 2, p1, 1.0, 1.0, 0.0)
    assert m.stop > 0 and m.yb[m.stop - 1] < y[2]


@pytest.mark.parametrize("body,p1,p2,p3,ramp,t", CASES[:3])
def test_wave_stencils_partition_of_unity(grid, body, p1, p2, p3, ramp, t):
    """On the uniform buffer each marker's 125 weights add up to one and
    its nodes surround it; a pivot taken from the wrong reference face
    (y(1) for the top wall) would shift the window.  The w stencil is left
    out of the sum: its x window starts one cell early in the reference
    (src/immersed_boundary_operators.f90:156, ii = xm_pivot_index + i - 1),
    so its weights do not add up to one."""
    m = orc.WaveMarkers(grid, body, 16, 3, p1, p2, p3, t, ramp)
    ib = orc.IB(grid, m)
    yg, yf = grid.get(4), grid.get(1)
    for s in range(4):
        xi, yi, zi, w, dxn = ib.stencil(s)
        w = w.reshape(m.nb, 125)
        if s != 2:
            assert np.allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-12), s
        ycoord = (yf if s == 1 else yg)[yi.reshape(m.nb, 125) - 1]
        assert np.all(np.abs(ycoord - m.yb[:, None]) < 3.01 * grid.scalars["dymin"])


def test_center_wall_markers_and_single_body_mean():
    """center_wall (src/immersed_boundary_geometry.f90:80-115): markers on
    the plane y = 1 at the cell-centred x, z positions, pivots from face
    ny/2, weights that add up to one on the uniform grid, and the one-body
    mean removal f - mean(f) (:410-411)."""
    from channel3d_parallel_amd import ib_geometry
    g = orc.Grid(34, 33, 34, grid_type=0)
    s = g.scalars
    m = ib_geometry.center_wall(s["Lxp"], s["Lzp"], 2.0, 33, 32, 16)
    assert m.nb == 512 and m.nxb1 == m.nxb and np.all(m.yref == 16)
    assert np.all(m.yb == 1.0) and np.allclose(m.sb, s["Lxp"] / 32 * s["Lzp"] / 16, rtol=1e-15)
    assert np.all(m.normals[m.nb:2 * m.nb] == -1.0) and np.all(m.ub == 0.0)
    ib = orc.IB(g, m)
    for st in range(4):
        w = ib.stencil(st)[3].reshape(m.nb, 125)
        assert np.allclose(w.sum(axis=1), 1.0, rtol=0, atol=1e-12), st
    f = np.random.default_rng(2).standard_normal(m.nb)
    r = ib.remove_mean(f.copy())
    assert np.allclose(r, f - f.mean(), rtol=0, atol=1e-15)
