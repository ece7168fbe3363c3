"""Writes ib_twocyl130_projection.json: summary of one IB projection of the
two-cylinder body on the 130x129x130 uniform grid, computed by the oracle.

Case: laminar parabola (U_b = 1, no perturbation), boundary conditions,
non-IB projection at rk_step 2 with dt = 0.01, boundary conditions, then the
IB projection from a zero initial guess at absolute tolerance 1e-10.  The
summary holds the BiCGSTAB iteration count and residual, the slip and
divergence metrics, block norms of the force and norms and weighted sums of
the projected velocity (weights: SplitMix64 uniforms, seed 77), enough to
check a GPU run without storing fields.

Run from the repository root: python tests/golden/make_ib_twocyl130.py
"""
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from channel3d_parallel_amd import fields, ib_geometry  # noqa: E402
from oracle import orc  # noqa: E402

NX, NY, NZ = 130, 129, 130
NU, DT, TOL = 1.0 / 2800.0, 0.01, 1e-10


def summarize(U, V, W, fb, nb):
    out = {}
    for name, a in (("U", U), ("V", V), ("W", W)):
        w = fields.uniform(77, a.size).reshape(a.shape, order="F")
        out[name + "_l2"] = float(np.linalg.norm(a))
        out[name + "_wsum"] = float(np.sum(w * a))
    for k, name in enumerate(("dudn_x", "dudn_y", "dudn_z", "p_jump")):
        out["fb_" + name + "_l2"] = float(np.linalg.norm(fb[k * nb:(k + 1) * nb]))
    return out


def main():
    g = orc.Grid(NX, NY, NZ, grid_type=0)
    m = ib_geometry.two_cylinders(g.scalars["Lzp"])
    ib = orc.IB(g, m)
    ib.heaviside()
    U, V, W = fields.channel_ic(NX, NY, NZ, pert=0.0, grid_type=0)
    g.apply_bc(U, V, W)
    out = g.projection(U, V, W, rk_step=2, dt=DT)
    g.apply_bc(U, V, W)
    fb = np.zeros(4 * m.nb)
    t = time.time()
    it, err = ib.projection(U, V, W, out["U_interim"], out["V_interim"], out["W_interim"],
                            out["P_interim"], fb, DT, NU, tol=TOL, maxit=2000)
    res = dict(nx=NX, ny=NY, nz=NZ, nu=NU, dt=DT, tol=TOL, nb=m.nb, iterations=it, err=err,
               max_slip=ib.max_slip(U, V, W, fb), max_div=g.max_abs_div(U, V, W),
               oracle_seconds=time.time() - t)
    res.update(summarize(U, V, W, fb, m.nb))
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "ib_twocyl130_projection.json"),
              "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res, indent=1))


if __name__ == "__main__":
    main()
