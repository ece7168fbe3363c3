"""GPU parity of the moving-wall path: setup_IB_geometry for standing_wave
and traveling_wave_x, with setup_IB_operators and compute_heaviside redone
on the device at every Runge-Kutta stage (src/time_integration.f90:110-115),
on the buffered grid (grid_type 2).

Bars:
  marker geometry vs oracle        x, z, y_ref_index bit-exact; y, normals,
                                   velocities, areas to 2e-15 absolute
                                   (device sin/cos vs libm)
  device-built operators vs the    bit-exact: stencils, Heaviside fields and
  host-built ones (same markers)   coefficients, spreading, Schur products
  operators vs oracle (same        as the static body: indices and dxn
  markers)                         bit-exact, weights 1e-14, Heaviside 1e-12,
                                   Schur 1e-12 of the largest entry
  a body moved to t vs one set up  bit-exact (nothing of the old position
  at t                             survives the move)
  RK3 steps vs oracle composition  both BiCGSTAB solves to 1e-12 absolute;
                                   fields to 1e-9 relative (two solutions
                                   converged to a residual tol differ by
                                   ~tol times the Schur conditioning: at
                                   1e-10 the fields part by ~2e-8)
"""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields
from oracle import orc

pytestmark = pytest.mark.gpu

NU = 1.0 / 2800.0
NX, NY, NZ, MBW, ALPHA = 34, 65, 18, 0.1, 2.6
BODIES = {"standing_wave": (0.04, 2.0, 1.0, 0.0), "traveling_wave_x": (0.03, 1.0, 1.0, 0.5)}
NXB, NZB = 32, 8


def grid():
    return orc.Grid(NX, NY, NZ, alpha=ALPHA, grid_type=2, min_buffer_width=MBW)


def channel():
    return Channel(NX, NY, NZ, alpha=ALPHA, grid_type=2, nu=NU, min_buffer_width=MBW)


class Markers:
    """Marker set in the layout Channel.setup_IB / orc.IB read."""

    def __init__(self, d, nzb, nxb_wall):
        for k, v in d.items():
            setattr(self, k, v)
        self.nb, self.nzb, self.nxb, self.nxb1 = d["xb"].size, nzb, 2 * nxb_wall, nxb_wall


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("body", list(BODIES))
@pytest.mark.parametrize("t", [0.0, 0.37])
def test_wave_markers_vs_oracle(body, t):
    print(f"\n[case] {body} markers t={t}", flush=True)
    p1, p2, p3, ramp = BODIES[body]
    g = grid()
    o = orc.WaveMarkers(g, body, NXB, NZB, p1, p2, p3, t, ramp)
    with channel() as ch:
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp, t)
        d = ch.get_IB_markers()
    assert np.array_equal(d["xb"], o.xb) and np.array_equal(d["zb"], o.zb)
    assert np.array_equal(d["yref"], o.yref)
    for k in ("yb", "normals", "ub", "sb"):
        assert np.max(np.abs(d[k] - getattr(o, k))) <= 2e-15, k


def operators(ch, nb, f):
    out = {}
    for s in range(4):
        out[f"stencil{s}"] = ch.ib_stencil(s)
    out.update(ch.ib_heaviside())
    out["schur2"] = ch.schur(f, 0.004, 2)
    out["schur1"] = ch.schur(f, 0.004, 1)
    for c in range(3):
        out[f"reg{c}"] = ch.reg(c, f[:nb])
    return out


def assert_same(a, b):
    for k in a:
        va, vb = a[k], b[k]
        if isinstance(va, tuple):
            for x, y in zip(va, vb):
                assert np.array_equal(x, y), k
        else:
            assert np.array_equal(va, vb), k


@pytest.mark.parametrize("body", list(BODIES))
def test_device_operators_match_host_setup(body):
    """The wave body's device-built stencils and spreading maps against the
    host-built ones of c3d_ib_setup_ref on the same (device-generated)
    markers: every operator output bit for bit."""
    print(f"\n[case] {body} device vs host set-up", flush=True)
    p1, p2, p3, ramp = BODIES[body]
    nb = 2 * NXB * NZB
    f = np.random.default_rng(3).standard_normal(4 * nb)
    with channel() as ch:
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp, 0.21)
        m = Markers(ch.get_IB_markers(), NZB, NXB)
        a = operators(ch, nb, f)
    with channel() as ch:
        ch.setup_IB(m)
        b = operators(ch, nb, f)
    assert_same(a, b)


@pytest.mark.parametrize("body", list(BODIES))
def test_move_equals_fresh_setup(body):
    print(f"\n[case] {body} move vs fresh set-up", flush=True)
    p1, p2, p3, ramp = BODIES[body]
    nb = 2 * NXB * NZB
    f = np.random.default_rng(4).standard_normal(4 * nb)
    with channel() as ch:
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp, 0.0)
        operators(ch, nb, f)
        ch.move_IB(0.9)
        a = operators(ch, nb, f)
    with channel() as ch:
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp, 0.9)
        b = operators(ch, nb, f)
    assert_same(a, b)


@pytest.mark.parametrize("body", list(BODIES))
def test_wave_operators_vs_oracle(body):
    print(f"\n[case] {body} operators vs oracle", flush=True)
    p1, p2, p3, ramp = BODIES[body]
    g = grid()
    nb = 2 * NXB * NZB
    f = np.random.default_rng(5).standard_normal(4 * nb)
    with channel() as ch:
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp, 0.6)
        m = Markers(ch.get_IB_markers(), NZB, NXB)
        ib = orc.IB(g, m)
        for s in range(4):
            xi, yi, zi, w, dxn = ch.ib_stencil(s)
            oxi, oyi, ozi, ow, odxn = ib.stencil(s)
            assert np.array_equal(xi, oxi) and np.array_equal(yi, oyi) and np.array_equal(zi, ozi), s
            assert np.array_equal(dxn, odxn), s
            assert np.max(np.abs(w - ow)) <= 1e-14 * np.max(np.abs(ow)), s
        Hc, Hu, Hv, Hw = ib.heaviside()
        co = ib.heaviside_coeffs()
        h = ch.ib_heaviside()
        assert np.max(np.abs(h["Hc_ext"] - Hc)) < 1e-12
        for k in ("Hu_int", "Hv_int", "Hw_int"):
            assert np.max(np.abs(h[k] - co[k])) < 1e-12, k
        assert np.max(np.abs(h["E1nH"] - co["E1nH"])) < 1e-12 * np.max(np.abs(co["E1nH"])) + 1e-16
        Sg = ch.schur(f, 0.004, 2)
    So = ib.schur(f, 0.004, NU)
    assert np.max(np.abs(Sg - So)) <= 1e-12 * np.max(np.abs(So))


@pytest.mark.parametrize("body", list(BODIES))
def test_time_steps_with_moving_walls(body):
    """Two RK3 steps: at each stage the walls move to t0 + rk_t dt and the
    operators are rebuilt, against the oracle composed stage by stage."""
    print(f"\n[case] {body} RK3 with moving walls", flush=True)
    p1, p2, p3, ramp = BODIES[body]
    g = grid()
    U, V, W = fields.channel_ic(NX, NY, NZ, pert=0.05, seed=2, alpha=ALPHA, grid_type=2,
                                min_buffer_width=MBW)
    g.apply_bc(U, V, W)
    dPdx = 3.0 * NU
    dt, tol, t0 = 0.01, 1e-12, 0.2
    rk_t = (8.0 / 15.0, 2.0 / 3.0, 1.0)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    nb = 2 * NXB * NZB
    fb = np.zeros(4 * nb)
    t = t0
    its = []
    for _ in range(2):
        Us = [Ur.copy(order="F"), Vr.copy(order="F"), Wr.copy(order="F")]
        F = []
        for st in (1, 2, 3):
            F.append(g.rhs(Ur, Vr, Wr, NU, dPdx))
            for q, X in enumerate((Ur, Vr, Wr)):
                g.rk_update(q, st, dt, F[0][q], F[1][q] if st > 1 else None,
                            F[2][q] if st > 2 else None, Us[q], X)
            m = orc.WaveMarkers(g, body, NXB, NZB, p1, p2, p3, t + rk_t[st - 1] * dt, ramp)
            assert m.stop == 0
            ib = orc.IB(g, m)
            ib.heaviside()
            g.apply_bc(Ur, Vr, Wr)
            out = g.projection(Ur, Vr, Wr, rk_step=st, dt=dt, P=g.zeros("p"))
            g.apply_bc(Ur, Vr, Wr)
            fb[:] = 0.0
            it, _ = ib.projection(Ur, Vr, Wr, out["U_interim"], out["V_interim"], out["W_interim"],
                                  out["P_interim"], fb, dt, NU, tol=tol, maxit=2000, rk_step=st,
                                  P=g.zeros("p"))
            its.append(it)
            g.apply_bc(Ur, Vr, Wr)
        t = t + dt
    with channel() as ch:
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(dPdx, 0.0, 0.0)
        ch.time = t0
        ch.setup_IB_wave(body, NXB, NZB, p1, p2, p3, ramp)
        ch.set_IB_solver(tol=tol, maxit=2000, guess_mode=0)
        for _ in range(2):
            ch.compute_time_step_RK3(dt, x_mass_cte=0)
        r = ch.get_fields()
        last, accum, err = ch.ib_stats()
        div, slip = ch.check_divergence(), ch.check_slip()
        t_end = ch.time
        d = ch.get_IB_markers()
    print(f"  oracle iterations {its}; gpu total {accum}; max|div| {div:.2e} slip {slip:.2e}")
    assert t_end == t
    mo = orc.WaveMarkers(g, body, NXB, NZB, p1, p2, p3, t0 + dt + dt, ramp)
    assert np.max(np.abs(d["yb"] - mo.yb)) <= 2e-15
    assert err < tol ** 2 and div < 1e-12 and slip < 1e-11
    for a, b in ((r["U"], Ur), (r["V"], Vr), (r["W"], Wr)):
        assert rel(a, b) < 1e-9


def test_flat_walls_stay_flat():
    """Amplitude 0 (the flat-wall channel of examples/input_parameters_ib_mfu):
    a move leaves every operator unchanged."""
    print("\n[case] flat walls", flush=True)
    nb = 2 * 16 * 4
    f = np.random.default_rng(6).standard_normal(4 * nb)
    with channel() as ch:
        ch.setup_IB_wave("standing_wave", 16, 4, 0.0, 0.0, 0.0, 0.0, 0.0)
        a = operators(ch, nb, f)
        ch.move_IB(5.0)
        b = operators(ch, nb, f)
    assert_same(a, b)


def test_wave_setup_errors():
    """The reference's stop conditions: amplitude above the buffer width
    (src/immersed_boundary_geometry.f90:177) and markers below y(3)."""
    from channel3d_parallel_amd.solver import C3DError
    with channel() as ch:
        with pytest.raises(C3DError):
            ch.setup_IB_wave("standing_wave", 16, 4, 0.2, 1.0, 1.0)
    with channel() as ch:
        with pytest.raises(C3DError):
            ch.setup_IB_wave("traveling_wave_x", 16, 4, 0.1, 1.0, 1.0)  # 0.1/(1*1) > |y(3)|


def test_device_map_dense_markers():
    """Markers four per cell in x and two per cell in z (64 x 32 per wall on
    a 16 x 16 cell plane): about 200 entries per support node, a wide count
    range for the device sort; operators bitwise equal to the host-built
    set-up of the same markers."""
    nx, ny, nz = 18, 65, 18
    print(f"\n[case] dense wall markers {nx}x{ny}x{nz}", flush=True)
    kw = dict(alpha=2.6, grid_type=2, min_buffer_width=0.1, nu=NU)
    nxb, nzb = 64, 32
    nb = 2 * nxb * nzb
    f = np.random.default_rng(9).standard_normal(4 * nb)
    with Channel(nx, ny, nz, **kw) as ch:
        ch.setup_IB_wave("standing_wave", nxb, nzb, 0.04, 2.0, 1.0, 0.0, 0.3)
        m = Markers(ch.get_IB_markers(), nzb, nxb)
        a = operators(ch, nb, f)
    with Channel(nx, ny, nz, **kw) as ch:
        ch.setup_IB(m)
        b = operators(ch, nb, f)
    assert_same(a, b)


def test_time_step_moving_walls_benchmark_size():
    """The benchmarked moving-wall workload: standing wave (amplitude 0.03,
    omega 2, one wavelength in x), 128 x 64 markers per wall (nb = 16,384),
    130x129x66 faces on the buffered grid (alpha 2.6, min_buffer_width 0.1),
    constant pressure gradient 3 nu.  One RK3 step with the device
    re-set-up at every stage against the oracle composed stage by stage, at
    the bars of test_time_steps_with_moving_walls."""
    nx, ny, nz, nxb, nzb, mbw, alpha = 130, 129, 66, 128, 64, 0.1, 2.6
    p1, p2, p3, ramp = 0.03, 2.0, 1.0, 0.0
    print(f"\n[case] standing wave {nx}x{ny}x{nz}, {nxb}x{nzb} markers per wall: one RK3 step", flush=True)
    g = orc.Grid(nx, ny, nz, alpha=alpha, grid_type=2, min_buffer_width=mbw)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.0, alpha=alpha, grid_type=2, min_buffer_width=mbw)
    g.apply_bc(U, V, W)
    dPdx, dt, tol, t0 = 3.0 * NU, 0.005, 1e-12, 0.0
    rk_t = (8.0 / 15.0, 2.0 / 3.0, 1.0)
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Us = [U.copy(order="F"), V.copy(order="F"), W.copy(order="F")]
    fb = np.zeros(4 * 2 * nxb * nzb)
    F, its = [], []
    for st in (1, 2, 3):
        F.append(g.rhs(Ur, Vr, Wr, NU, dPdx))
        for q, X in enumerate((Ur, Vr, Wr)):
            g.rk_update(q, st, dt, F[0][q], F[1][q] if st > 1 else None,
                        F[2][q] if st > 2 else None, Us[q], X)
        m = orc.WaveMarkers(g, "standing_wave", nxb, nzb, p1, p2, p3, t0 + rk_t[st - 1] * dt, ramp)
        assert m.stop == 0
        ib = orc.IB(g, m)
        ib.heaviside()
        g.apply_bc(Ur, Vr, Wr)
        out = g.projection(Ur, Vr, Wr, rk_step=st, dt=dt, P=g.zeros("p"))
        g.apply_bc(Ur, Vr, Wr)
        fb[:] = 0.0
        it, _ = ib.projection(Ur, Vr, Wr, out["U_interim"], out["V_interim"], out["W_interim"],
                              out["P_interim"], fb, dt, NU, tol=tol, maxit=2000, rk_step=st,
                              P=g.zeros("p"))
        its.append(it)
        g.apply_bc(Ur, Vr, Wr)
    with Channel(nx, ny, nz, alpha=alpha, grid_type=2, nu=NU, min_buffer_width=mbw) as ch:
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(dPdx, 0.0, 0.0)
        ch.time = t0
        ch.setup_IB_wave("standing_wave", nxb, nzb, p1, p2, p3, ramp)
        ch.set_IB_solver(tol=tol, maxit=2000, guess_mode=0)
        ch.compute_time_step_RK3(dt, x_mass_cte=0)
        r = ch.get_fields()
        last, accum, err = ch.ib_stats()
        div, slip = ch.check_divergence(), ch.check_slip()
    scale = np.linalg.norm(Ur)
    d = [np.linalg.norm(a - b) / scale for a, b in ((r["U"], Ur), (r["V"], Vr), (r["W"], Wr))]
    print(f"  oracle iterations {its}; gpu total {accum}; max|div| {div:.2e} slip {slip:.2e}; "
          f"|difference| / |U|: U {d[0]:.2e} V {d[1]:.2e} W {d[2]:.2e}")
    assert err < tol ** 2 and div < 1e-12 and slip < 1e-11
    assert max(d) < 1e-9
