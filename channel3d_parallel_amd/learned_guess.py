"""Learned initial guess for the IB Schur solve: data generation, training
and the iteration A/B comparison.

Training pairs come from the IB solver itself: for every IB projection of a
two-cylinder run with seeded rotation speeds, the features the network sees
(per marker and 3x3 neighbourhood: the Schur residual of the previous force,
b - S f_prev, and that force) and the converged force increment.  Each
solve's pairs are normalised by that solve's own scales, as the library
does at inference: the residual channels and the increment by the rms of
the residual, the force channels by the rms of the force.  The increment
solves S d = b - S f_prev, linear in the residual, and the start takes the
step along the predicted direction that minimises the residual, so only
the direction matters.
The network is trained on the GPU by the library's own kernels
(c3d_mlp_train).  Plain BiCGSTAB from a zero guess stays the reference:
every variant is solved to the same absolute tolerance.
"""
from __future__ import annotations

import numpy as np

from . import fields, ib_geometry
from .solver import Channel

NU = 1.0 / 2800.0


def speeds(seed, run):
    """Rotation speeds of the two cylinders, uniform in [0.3, 0.7] U_b."""
    u = fields.uniform(seed, 2, offset=2 * run)
    return 0.3 + 0.4 * float(u[0]), -(0.3 + 0.4 * float(u[1]))


def make_channel(nx, ny, nz, nth, nzb, s1, s2, dt_guess_mode=0, tol=1e-10, maxit=2000):
    ch = Channel(nx, ny, nz, grid_type=0, nu=NU)
    m = ib_geometry.two_cylinders(ch.grid(11)[9], nzb=nzb, nth=nth, speed1=s1, speed2=s2)
    ch.setup_IB(m)
    U, V, W = fields.channel_ic(nx, ny, nz, pert=0.0, grid_type=0)
    ch.set_fields(U, V, W, P=np.zeros((nx + 1, ny + 1, nz + 1), order="F"))
    ch.apply_boundary_conditions()
    ch.set_pressure_gradient(3.0 * NU, 0.0, 0.0)
    ch.set_IB_solver(tol=tol, maxit=maxit, guess_mode=dt_guess_mode)
    return ch, m


def stage(ch, st, dt, before_ib=None):
    """One RK3 stage of compute_time_step_RK3 (src/time_integration.f90:98-123)
    through the per-procedure entry points; before_ib(ch, st) runs right
    before the IB projection and may run it itself (returning True)."""
    if st == 1:
        ch.save_state()
    ch.compute_rhs(st)
    ch.rk3_update(st, dt)
    ch.apply_boundary_conditions()
    ch.compute_non_IB_projection(dt, st)
    ch.apply_boundary_conditions()
    done = before_ib(ch, st) if before_ib else False
    if not done:
        ch.compute_IB_projection(dt, st)
    ch.apply_boundary_conditions()


def normalise(X, T):
    """Per-solve normalisation of one solve's pairs (X [nb, 54] as
    ib_last_pair returns them, T [nb, 4]): residual channels and T divided by
    the rms of the residual (the centre marker's channels 0-3 over all
    markers), force channels by the rms of dudn_x, dudn_y."""
    X = X.astype(np.float64).reshape(-1, 9, 6)
    rr = np.sqrt(np.mean(X[:, 4, 0:4] ** 2))
    rf = np.sqrt(np.mean(X[:, 4, 4:6] ** 2))
    ir = 1.0 / rr if rr > 0 else 0.0
    jf = 1.0 / rf if rf > 0 else 0.0
    X[:, :, 0:4] *= ir
    X[:, :, 4:6] *= jf
    return X.reshape(-1, 54), T * ir


def generate_pairs(nx, ny, nz, nth, nzb, runs, steps, dt, seed=11, max_pairs=None, rng_seed=5,
                   warm_steps=0, base_mode=0):
    """Normalised features X [n, 54] and force increments T [n, 4] from
    `runs` runs: `warm_steps` RK3 steps past the start-up transient, then
    `steps` steps whose every IB solve gives nb pairs.  The solves start from
    guess mode `base_mode` (0: the previous force, the base the learned start
    builds on)."""
    Xs, Ts = [], []
    for run in range(runs):
        s1, s2 = speeds(seed, run)
        ch, m = make_channel(nx, ny, nz, nth, nzb, s1, s2, dt_guess_mode=base_mode)
        with ch:
            for _ in range(warm_steps):
                ch.compute_time_step_RK3(dt)
            for _ in range(steps):
                for st in (1, 2, 3):
                    stage(ch, st, dt)
                    X, dF = ch.ib_last_pair()
                    Xn, Tn = normalise(X, dF.reshape(4, m.nb).T.astype(np.float64))
                    Xs.append(Xn)
                    Ts.append(Tn)
    X = np.concatenate(Xs)
    T = np.concatenate(Ts)
    if max_pairs and X.shape[0] > max_pairs:
        idx = np.random.default_rng(rng_seed).choice(X.shape[0], max_pairs, replace=False)
        X, T = X[idx], T[idx]
    return X, T


def scales(X, T):
    """Per-channel input scales (6 channels shared by the 9 neighbours) and
    output scales; a channel without spread (dudn_z of z-invariant bodies)
    gets a scale 1e-6 of the largest so its targets stay small."""
    xc = X.reshape(-1, 9, 6)
    isc = np.sqrt(np.mean(xc.astype(np.float64) ** 2, axis=(0, 1)))
    isc = np.maximum(isc, 1e-6 * isc.max())
    osc = np.sqrt(np.mean(T ** 2, axis=0))
    osc = np.maximum(osc, 1e-6 * osc.max())
    return isc.astype(np.float32), osc.astype(np.float32)


def train(ch, X, T, epochs=40, batch=256, lr=1e-3, seed=1):
    """Initialises and trains the network of channel `ch`; returns the
    per-epoch loss history and the scales."""
    isc, osc = scales(X, T)
    ch.mlp_init(seed)
    ch.mlp_set_scales(isc, osc)
    Xs = (X.reshape(-1, 9, 6) / isc).reshape(-1, 54).astype(np.float32)
    Ts = (T / osc).astype(np.float32)
    hist = ch.mlp_train(Xs, Ts, epochs, batch=batch, lr=lr, seed=seed + 100)
    return hist, isc, osc


def per_stage(rec, name):
    """Iterations of `name` per RK3 stage (solves in stage order 1, 2, 3)."""
    it = rec[name]
    return {f"stage{st}": it[st - 1::3] for st in (1, 2, 3)}


def ab_compare(ch, dt, steps, tol=1e-10, maxit=2000, learned=None, stage_ch=None):
    """At every IB projection of `steps` RK3 steps, solve from the zero guess,
    from the previous force (the reference's warm start) and from the
    least-squares history start (guess_mode 4, fed by its own solves of the
    same states) on `ch`, all from the same state, and continue with the
    zero-guess solution.  The same-stage start (guess_mode 5 on `stage_ch`,
    which keeps each stage's force of the previous step) and the learned
    guess (guess_mode 3 on `learned`, a channel with the trained network)
    run on channels of their own, which repeat the stage's projection from
    the same interim velocity, so that the stage forces each keeps are its
    own.  Returns per-solve iteration counts and force differences against
    the zero-guess force."""
    names = [("previous", 0), ("history", 4), ("zero", 1)]
    sides = [(n, c, m) for n, c, m in (("stage", stage_ch, 5), ("learned", learned, 3)) if c is not None]
    rec = {n: [] for n, _ in names}
    rec.update({n: [] for n, _, _ in sides})
    for n in list(rec):
        if n != "zero":
            rec[f"max_force_diff_{n}_vs_zero"] = []
    for _ in range(steps):
        for st in (1, 2, 3):
            if st == 1:
                ch.save_state()
            ch.compute_rhs(st)
            ch.rk3_update(st, dt)
            ch.apply_boundary_conditions()
            pre = ch.get_fields()  # the interim velocity entering the projections
            ch.compute_non_IB_projection(dt, st)
            ch.apply_boundary_conditions()
            snap = ch.get_fields()
            fb0 = ch.get_IB_force()
            out = {}
            for n, c, m in sides:
                c.set_fields(pre["U"], pre["V"], pre["W"], P=pre["P"])
                c.compute_non_IB_projection(dt, st)
                c.apply_boundary_conditions()
                c.set_IB_force(fb0)
                c.set_IB_solver(tol=tol, maxit=maxit, guess_mode=m)
                it, _ = c.compute_IB_projection(dt, st)
                rec[n].append(it)
                out[n] = c.get_IB_force()
            for n, m in names:  # the zero-guess solve last: the state continues from it
                ch.set_fields(snap["U"], snap["V"], snap["W"], P=snap["P"])
                ch.set_IB_force(fb0)
                ch.set_IB_solver(tol=tol, maxit=maxit, guess_mode=m)
                it, _ = ch.compute_IB_projection(dt, st)
                rec[n].append(it)
                out[n] = ch.get_IB_force()
            scale = max(np.max(np.abs(out["zero"])), 1e-300)
            for n in out:
                if n != "zero":
                    rec[f"max_force_diff_{n}_vs_zero"].append(
                        float(np.max(np.abs(out[n] - out["zero"])) / scale))
            ch.apply_boundary_conditions()
    return rec


def side_channels(ch, nx, ny, nz, nth, nzb, s1, s2, isc=None, osc=None):
    """Channels for ab_compare's per-stage starts, on the body of `ch`: one
    for guess_mode 5 and, when the scales are given, one with a copy of
    `ch`'s trained network for guess_mode 3."""
    stage_ch, _ = make_channel(nx, ny, nz, nth, nzb, s1, s2)
    learned = None
    if isc is not None:
        learned, _ = make_channel(nx, ny, nz, nth, nzb, s1, s2)
        learned.mlp_init(1)
        learned.mlp_set_scales(isc, osc)
        learned.mlp_set_params(ch.mlp_params())
    return stage_ch, learned
