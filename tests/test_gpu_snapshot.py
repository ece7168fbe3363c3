"""Snapshots and restart through the C ABI: the library's writer against the
oracle's reader of the reference format, the oracle's writer against the
library's reader, the reader's checks, and restarts that continue bit for
bit once P, t, dPdx and fb are restored (the reference restores U, V, W
only, src/input_output.f90:335-489).

Bars: every value bit-exact (the file stores FP64 values as they are)."""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields, ib_geometry
from channel3d_parallel_amd._lib import RESTORE_ALL
from channel3d_parallel_amd.solver import C3DError
from oracle import orc, snapshot

pytestmark = pytest.mark.gpu

NU = 1.0 / 2800.0


def cylinders_case(ch, g):
    m = ib_geometry.two_cylinders(g.scalars["Lzp"], nzb=16, nth=24)
    ch.setup_IB(m)
    return m


def test_library_snapshot_reads_as_reference_format(tmp_path):
    nx, ny, nz = 34, 33, 34
    print(f"\n[case] snapshot write {nx}x{ny}x{nz} with cylinders", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=0)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=5, grid_type=0)
    P = np.asfortranarray(np.random.default_rng(2).standard_normal(g.shape_p))
    p = tmp_path / "snap.000010"
    with Channel(nx, ny, nz, grid_type=0, nu=NU) as ch:
        m = cylinders_case(ch, g)
        ch.set_fields(U, V, W, P=P)
        ch.set_pressure_gradient(2.5e-3, 0.0, 0.0)
        ch.time = 3.5
        fb = np.random.default_rng(3).standard_normal(4 * m.nb)
        ch.set_IB_force(fb)
        t1 = np.random.default_rng(4).standard_normal(3 * m.nb)
        ch.write_snapshot(p, 0.004, tangents_1=t1)
        h = ch.ib_heaviside()
        mk = ch.get_IB_markers()
        fbr = ch.get_IB_force()
    assert p.stat().st_size == snapshot.expected_size(nx, ny, nz, m.nb)
    d = snapshot.read(p)
    for k, w in (("x", 0), ("y", 1), ("z", 2)):
        assert np.array_equal(d[k], g.get(w)), k
    for k, w in (("xm", 3), ("ym", 4), ("zm", 5)):
        assert np.array_equal(d[k], g.get(w)[1:-1]), k
    for k, a in (("U", U), ("V", V), ("W", W), ("P", P)):
        assert np.array_equal(d[k], a), k
    for k in ("Hu_int", "Hv_int", "Hw_int"):
        assert np.array_equal(d[k.replace("_int", "_interior")], h[k]), k
    assert (d["t"], d["dt"], d["dpdx"], d["nu"]) == (3.5, 0.004, 2.5e-3, NU)
    for k in ("xb", "yb", "zb", "ub", "sb", "normals"):
        assert np.array_equal(d[k], mk[k]), k
    assert np.array_equal(d["fb"], fbr) and np.array_equal(d["fb"], fb)
    assert np.array_equal(d["tangents_1"], t1) and not d["tangents_2"].any()
    assert (d["nxb"], d["nzb"]) == (m.nxb, m.nzb)


def test_reference_format_file_restores(tmp_path):
    """A file written by the format's restatement (as tools/file_io.py
    writes it) restores U, V, W, P, t, dPdx and returns dt."""
    nx, ny, nz = 34, 65, 18
    print(f"\n[case] snapshot read {nx}x{ny}x{nz} body-free", flush=True)
    g = orc.Grid(nx, ny, nz)
    rng = np.random.default_rng(7)
    d = {"x": g.get(0), "y": g.get(1), "z": g.get(2), "xm": g.get(3)[1:-1],
         "ym": g.get(4)[1:-1], "zm": g.get(5)[1:-1]}
    for k, s in (("U", g.shape_u), ("V", g.shape_v), ("W", g.shape_w), ("P", g.shape_p),
                 ("Hu_interior", g.shape_u), ("Hv_interior", g.shape_v),
                 ("Hw_interior", g.shape_w)):
        d[k] = np.asfortranarray(rng.standard_normal(s))
    d.update(t=7.0, dt=0.002, dpdx=1e-3, nu=NU, nxb=0, nzb=0)
    for k in ("xb", "yb", "zb", "fb", "ub", "sb", "normals", "tangents_1", "tangents_2"):
        d[k] = np.zeros(0)
    p = tmp_path / "in.snap"
    snapshot.write(p, d)
    with Channel(nx, ny, nz, nu=NU) as ch:
        dt = ch.read_snapshot(p)
        f = ch.get_fields()
        t = ch.time
        dpdx = ch.get_pressure_gradient()[0]
    assert dt == 0.002 and t == 7.0 and dpdx == 1e-3
    for k in ("U", "V", "W", "P"):
        assert np.array_equal(f[k], d[k]), k
    # fields-only files restore U, V, W (and P) but have no time section
    q = tmp_path / "orig.snap"
    snapshot.write(q, d, fields_only=True)
    with Channel(nx, ny, nz, nu=NU) as ch:
        ch.read_snapshot(q, what=1)
        assert np.array_equal(ch.get_fields()["U"], d["U"])
        with pytest.raises(C3DError):
            ch.read_snapshot(q, what=RESTORE_ALL)


def test_reader_rejects_other_grids(tmp_path):
    nx, ny, nz = 34, 33, 34
    print(f"\n[case] snapshot grid checks {nx}x{ny}x{nz}", flush=True)
    p = tmp_path / "snap"
    with Channel(nx, ny, nz, grid_type=1, alpha=2.0) as ch:
        ch.write_snapshot(p, 0.01)
    with Channel(nx, ny, nz, grid_type=1, alpha=2.1) as ch:  # other stretching
        with pytest.raises(C3DError):
            ch.read_snapshot(p)
    with Channel(nx, ny + 2, nz, grid_type=1, alpha=2.0) as ch:  # other size
        with pytest.raises(C3DError):
            ch.read_snapshot(p)
    with Channel(nx, ny, nz, grid_type=1, alpha=2.0) as ch:
        ch.read_snapshot(p, what=0)


def test_restart_continues_bitwise_constant_flow(tmp_path):
    """3 steps in one run against 2 steps, a snapshot, and 1 step in a new
    context: identical with P, t and dPdx restored; the reference's restart
    (U, V, W only) loses dPdx and the step differs."""
    nx, ny, nz = 34, 65, 34
    print(f"\n[case] restart constant flow {nx}x{ny}x{nz}", flush=True)
    U, V, W = fields.channel_ic(nx, ny, nz, seed=9)
    dpdx0 = (180 * NU) ** 2
    g = orc.Grid(nx, ny, nz)
    g.apply_bc(U, V, W)
    q0 = g.mean_mass_flow_u(U)
    p = tmp_path / "snap"
    with Channel(nx, ny, nz, nu=NU) as ch:
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        ch.set_pressure_gradient(dpdx0, 0.0, 0.0)
        dt = ch.compute_dt(0.5)
        for n in range(3):
            ch.compute_time_step_RK3(dt, x_mass_cte=1, q0=q0)
            if n == 1:
                ch.write_snapshot(p, dt)
        ref = ch.get_fields()
    for what, same in ((RESTORE_ALL, True), (0, False)):
        with Channel(nx, ny, nz, nu=NU) as ch:
            ch.set_pressure_gradient(dpdx0, 0.0, 0.0)
            ch.read_snapshot(p, what=what)
            ch.compute_time_step_RK3(dt, x_mass_cte=1, q0=q0)
            out = ch.get_fields()
        eq = all(np.array_equal(out[k], ref[k]) for k in ("U", "V", "W", "P"))
        assert eq == same, what


@pytest.mark.parametrize("body", ["cylinders", "standing_wave"])
def test_restart_continues_bitwise_with_body(tmp_path, body):
    """IB runs: fb (the warm start of the next solve) and t (the walls'
    position) restored make the continued run identical, iteration counts
    included."""
    if body == "cylinders":
        nx, ny, nz, kw = 34, 33, 34, dict(grid_type=0)
    else:
        nx, ny, nz, kw = 34, 65, 18, dict(alpha=2.6, grid_type=2, min_buffer_width=0.1)
    print(f"\n[case] restart with {body} {nx}x{ny}x{nz}", flush=True)
    g = orc.Grid(nx, ny, nz, **kw)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.02, seed=4, **kw)
    g.apply_bc(U, V, W)
    dt, p = 0.004, tmp_path / "snap"

    def setup(ch):
        if body == "cylinders":
            cylinders_case(ch, g)
        else:
            ch.setup_IB_wave("standing_wave", 32, 8, 0.04, 2.0, 1.0, 0.0, 0.0)
        ch.set_pressure_gradient(3.0 * NU, 0.0, 0.0)
        ch.set_IB_solver(tol=1e-10, maxit=1000, guess_mode=0)

    with Channel(nx, ny, nz, nu=NU, **kw) as ch:
        ch.set_fields(U, V, W, P=np.zeros(g.shape_p, order="F"))
        setup(ch)
        for n in range(3):
            ch.compute_time_step_RK3(dt, x_mass_cte=0)
            if n == 1:
                ch.write_snapshot(p, dt)
                _, acc2, _ = ch.ib_stats()
        ref = ch.get_fields()
        fb_ref = ch.get_IB_force()
        _, acc3, _ = ch.ib_stats()
        t_ref = ch.time
    with Channel(nx, ny, nz, nu=NU, **kw) as ch:
        setup(ch)
        ch.read_snapshot(p)
        ch.compute_time_step_RK3(dt, x_mass_cte=0)
        out = ch.get_fields()
        fb = ch.get_IB_force()
        _, acc, _ = ch.ib_stats()
        t = ch.time
    assert t == t_ref and acc == acc3 - acc2
    for k in ("U", "V", "W", "P"):
        assert np.array_equal(out[k], ref[k]), k
    assert np.array_equal(fb, fb_ref)


def test_reader_rejects_truncated_and_corrupt_files(tmp_path):
    """Truncated fields and corrupt marker counts are errors, not reads past
    the data."""
    import struct
    nx, ny, nz = 34, 33, 34
    print(f"\n[case] snapshot corrupt files {nx}x{ny}x{nz}", flush=True)
    p = tmp_path / "snap"
    with Channel(nx, ny, nz) as ch:
        ch.write_snapshot(p, 0.01)
    raw = p.read_bytes()
    g = orc.Grid(nx, ny, nz)
    mesh = sum(4 + 8 * n for n in (nx, ny, nz, nx - 1, ny - 1, nz - 1))
    cut = tmp_path / "cut"
    cut.write_bytes(raw[:mesh + 12 + 1000])  # inside U
    fields_end = mesh + sum(12 + 8 * int(np.prod(s)) for s in
                            (g.shape_u, g.shape_v, g.shape_w, g.shape_p, g.shape_u, g.shape_v,
                             g.shape_w))
    bad = bytearray(raw)
    bad[fields_end + 32:fields_end + 36] = struct.pack(">i", -5)  # xb length
    corrupt = tmp_path / "corrupt"
    corrupt.write_bytes(bytes(bad))
    with Channel(nx, ny, nz) as ch:
        for f in (cut, corrupt):
            with pytest.raises(C3DError):
                ch.read_snapshot(f)
        ch.read_snapshot(p)  # the intact file still reads
