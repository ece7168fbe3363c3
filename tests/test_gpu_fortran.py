"""The Fortran host layer (reference module names, iso_c_binding) drives the
GPU projection and matches the oracle."""
import os
import subprocess

import numpy as np
import pytest

from channel3d_parallel_amd import fields
from oracle import orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "channel3d_parallel_amd", "fortran", "bin", "projection_driver")


@pytest.mark.parametrize("nx,ny,nz,gt,rk", [(34, 65, 34, 1, 1), (66, 33, 34, 0, 2)])
def test_fortran_projection_driver(tmp_path, nx, ny, nz, gt, rk):
    if not os.path.exists(DRIVER):
        pytest.fail("Fortran driver not built (make -C channel3d_parallel_amd/fortran)")
    print(f"\n[case] fortran driver {nx}x{ny}x{nz} gt={gt} rk={rk}", flush=True)
    U, V, W = fields.white_faces(nx, ny, nz, seed=21)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    np.concatenate([U.ravel("F"), V.ravel("F"), W.ravel("F")]).tofile(fin)
    dt = 3e-3
    subprocess.run([DRIVER, str(nx), str(ny), str(nz), str(gt), "2.0", repr(dt), str(rk),
                    str(fin), str(fout)], check=True, timeout=120)
    raw = np.fromfile(fout)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    shapes = [g.shape_u, g.shape_v, g.shape_w, g.shape_p, g.shape_rhs]
    outs, off = [], 0
    for s in shapes:
        n = int(np.prod(s))
        outs.append(raw[off:off + n].reshape(s, order="F"))
        off += n
    max_div = raw[off]
    g.apply_bc(U, V, W)
    P = np.zeros(g.shape_p, order="F")
    ref = g.projection(U, V, W, rk_step=rk, dt=dt, P=P)
    g.apply_bc(U, V, W)  # the driver reads back before the trailing BC; compare interiors
    for name, a, b in (("U", outs[0], U), ("V", outs[1], V), ("W", outs[2], W),
                       ("P_interim", outs[4], ref["P_interim"])):
        sl = (slice(1, -2), slice(1, -1), slice(1, -2))
        r = np.linalg.norm(a[sl] - b[sl]) / np.linalg.norm(b[sl])
        assert r <= 1e-12, (name, r)
    if rk == 1:
        assert np.linalg.norm(outs[3] - P) <= 1e-12 * np.linalg.norm(P)
    assert max_div <= max(1e-12, 4 * g.max_abs_div(U, V, W))
