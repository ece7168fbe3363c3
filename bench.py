"""Projection-solve throughput on one MI355X.

A step is one fractional-step projection solve (divergence, FFT-x/z,
tridiagonal-y, inverse FFTs, gradient correction: compute_non_IB_projection,
src/projection.f90:22-46) on the device-resident 384x193x384 channel grid
(386x193x386 faces, tanh-stretched y, alpha 2) with a seeded k^-5/3 random
field, inputs resident in HBM before the timed region.

Prints one JSON line (rank 0).  --gpus above 1 is not supported: this build
runs on one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, published

WORKLOADS = {
    # name: (nx, ny, nz) face counts, field kind, seed
    "proj384": (386, 193, 386, "spectral", 3),
    "proj192": (194, 129, 194, "white", 2),
}


def stage_bytes(nx, ny, nz, word=8):
    """Algorithmic HBM bytes per launch of each stage (each datum read or
    written once)."""
    nxp, nzp, nym = nx - 2, nz - 2, ny - 1
    ncell = nxp * nym * nzp                     # pressure unknowns
    nspec = nym * (nxp // 2 + 1) * nzp          # R2C half spectrum (complex)
    nrhs = (nx - 1) * (ny - 1) * nz             # rhs_p / P_interim array
    ngrad = (nx - 2) * (ny - 1) * (nz - 1)      # cells the gradient updates
    c = 2 * word
    return {
        "fwd_x_div": 3 * word * ncell + c * nspec,   # U,V,W read; spectrum written
        "fft_z_fwd": 2 * c * nspec,
        "tri_y": 2 * c * nspec,
        "fft_z_inv": 2 * c * nspec,
        "inv_x": c * nspec + word * nrhs,            # spectrum read; P_interim written
        "grad_correct": word * ngrad + 6 * word * ngrad,  # P_interim + U,V,W read; U,V,W written
    }


def read_traffic(kernel, workload):
    """HBM bytes per dispatch from the committed rocprofv3 --pmc summary,
    if one exists for this kernel and workload."""
    path = os.path.join(ROOT, "profiles", f"pmc_{kernel}_{workload}.txt")
    if not os.path.exists(path):
        return None
    m = re.search(r"hbm_bytes_per_dispatch\s*=\s*([0-9.eE+]+)", open(path).read())
    return float(m.group(1)) if m else None


def cpu_baseline(nx, ny, nz, U, V, W, dt, nsolves, threads):
    """The oracle (C FP64 restatement, OpenMP) timed on the host cores."""
    from oracle import orc

    g = orc.Grid(nx, ny, nz)
    u, v, w = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
    g.projection(u, v, w, dt=dt, nthreads=threads)  # warm-up
    ts = []
    for _ in range(nsolves):
        u, v, w = U.copy(order="F"), V.copy(order="F"), W.copy(order="F")
        t0 = time.perf_counter()
        g.projection(u, v, w, dt=dt, nthreads=threads)
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="proj384", choices=sorted(WORKLOADS))
    ap.add_argument("--cpu-solves", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: --gpus {a.gpus} is not supported; this build measures one GPU "
              "(the multi-GPU design is documented, not built).", file=sys.stderr)
        sys.exit(2)

    from channel3d_parallel_amd import Channel, fields

    nx, ny, nz, kind, seed = WORKLOADS[a.workload]
    print(f"[bench] workload={a.workload} grid={nx}x{ny}x{nz} faces, generating fields", file=sys.stderr,
          flush=True)
    if kind == "spectral":
        U, V, W = fields.spectral_faces(nx, ny, nz, seed)
    else:
        U, V, W = fields.white_faces(nx, ny, nz, seed)
    dt = 1e-3
    ncell = (nx - 2) * (ny - 1) * (nz - 2)

    with Channel(nx, ny, nz, grid_type=1, alpha=2.0) as ch:
        ch.set_fields(U, V, W)
        ch.apply_boundary_conditions()
        f = ch.get_fields()
        U, V, W = f["U"], f["V"], f["W"]  # BC-applied input for the CPU leg
        del f
        div_in = ch.check_divergence()
        # one clean projection (BC, project, BC) for the divergence figure
        ch.compute_non_IB_projection(dt)
        ch.apply_boundary_conditions()
        max_div = ch.check_divergence()
        copy_bytes = 1 << 30
        copy_ms = ch.copy_bandwidth(copy_bytes, iters=10)
        print("[bench] warm-up", file=sys.stderr, flush=True)
        for _ in range(a.warmup):
            ch.compute_non_IB_projection(dt)
        print(f"[bench] timing {a.steps} projection solves", file=sys.stderr, flush=True)
        t0 = time.perf_counter()
        tot_ms, stages = ch.time_projection(a.steps, dt)
        wall_ms = (time.perf_counter() - t0) * 1e3 / a.steps

    sb = stage_bytes(nx, ny, nz)
    dom = max(stages, key=stages.get)
    achieved = sb[dom] / (stages[dom] * 1e-3) / 1e9
    total_bytes = sum(sb.values())
    out = {
        "metric": "projection solves/s (FP64, 384x193x384)",
        "value": 1e3 / tot_ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": tot_ms,
        "wall_ms_per_step": wall_ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded k^-5/3 random-phase field, unit RMS)",
        "config": {"workload": f"{a.workload}: projection solve, {nx - 2}x{ny}x{nz - 2} cells "
                               f"({nx}x{ny}x{nz} faces), tanh y alpha=2, Lx=2pi Lz=pi Ly=2",
                   "cells": ncell, "precision": "fp64"},
        "ns_per_point": tot_ms * 1e6 / ncell,
        "input_max_div": div_in,
        "post_projection_max_div": max_div,
        "copy_GBps": 2 * copy_bytes / (copy_ms * 1e-3) / 1e9,
        "whole_solve_GBps": total_bytes / (tot_ms * 1e-3) / 1e9,
        "whole_solve_hbm_frac": total_bytes / (tot_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "stages_ms": stages,
        "stage_alg_bytes": sb,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": read_traffic(dom, a.workload)},
    }
    if not a.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        print(f"[bench] CPU baseline: oracle, {threads} threads, {a.cpu_solves} solves",
              file=sys.stderr, flush=True)
        t = cpu_baseline(nx, ny, nz, U, V, W, dt, a.cpu_solves, threads)
        out["cpu_baseline"] = {"value": 1.0 / t, "unit": "solves/s", "cores": threads,
                               "kind": "port",
                               "sample": f"{a.cpu_solves} full {nx - 2}x{ny}x{nz - 2} projection "
                                         "solves (median), C FP64 restatement with OpenMP; "
                                         "not the Fortran/FFTW binary"}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
