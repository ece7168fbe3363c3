"""Mixed-precision projection (FP32 spectra, FP64 tridiagonal arithmetic and
FP64 residual refinement) against the FP64 oracle.

Bars: without refinement the FP32 spectral storage limits the potential to
relative 1e-5; each refinement sweep must cut the FP64 residual by >= 1e3;
with two sweeps the fields agree with the FP64 oracle to relative 1e-11 and
max|div| is no worse than twice the plain FP64 GPU path's on the same input.
"""
import numpy as np
import pytest

from channel3d_parallel_amd import Channel, fields
from channel3d_parallel_amd._lib import PREC_MIXED
from oracle import orc

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("nx,ny,nz,gt,rk", [(66, 65, 66, 1, 2), (130, 129, 66, 1, 1), (66, 129, 34, 0, 2)])
def test_mixed_projection(nx, ny, nz, gt, rk):
    print(f"\n[case] mixed {nx}x{ny}x{nz} gt={gt} rk={rk}", flush=True)
    g = orc.Grid(nx, ny, nz, grid_type=gt)
    U, V, W = fields.white_faces(nx, ny, nz, seed=2)
    g.apply_bc(U, V, W)
    dt = 0.01
    Ur, Vr, Wr = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    Pr = g.zeros("p")
    ref = g.projection(Ur, Vr, Wr, rk_step=rk, dt=dt, P=Pr)
    g.apply_bc(Ur, Vr, Wr)
    with Channel(nx, ny, nz, grid_type=gt) as c64:
        c64.set_fields(U, V, W, P=g.zeros("p"))
        c64.compute_non_IB_projection(dt, rk)
        c64.apply_boundary_conditions()
        div64 = c64.check_divergence()
    res = {}
    with Channel(nx, ny, nz, grid_type=gt, precision=PREC_MIXED) as cm:
        for nref in (0, 1, 2, 3):
            cm.set_refinement(nref)
            cm.set_fields(U, V, W, P=g.zeros("p"))
            cm.compute_non_IB_projection(dt, rk)
            cm.apply_boundary_conditions()
            out = cm.get_fields()
            res[nref] = dict(div=cm.check_divergence(), rr=cm.refinement_residuals(),
                             eU=max(rel(out["U"], Ur), rel(out["V"], Vr), rel(out["W"], Wr)),
                             eP=rel(out["P_interim"], ref["P_interim"]),
                             eP1=rel(out["P"], Pr) if rk == 1 else 0.0)
            r = res[nref]
            print(f"  nref={nref}: max|div|={r['div']:.2e} (fp64 {div64:.2e}) "
                  f"relU={r['eU']:.2e} relPint={r['eP']:.2e} relP={r['eP1']:.2e} "
                  f"residuals={np.array2string(r['rr'], precision=2)}")
    assert res[0]["eP"] < 1e-5 and res[0]["eU"] < 1e-5
    for nref in (2, 3):
        rr = res[nref]["rr"]
        assert np.all(rr[1:] <= 1e-3 * rr[:-1]), rr
        assert res[nref]["eU"] < 1e-11 and res[nref]["eP"] < 1e-11 and res[nref]["eP1"] < 1e-11
        assert res[nref]["div"] <= 2.0 * div64
