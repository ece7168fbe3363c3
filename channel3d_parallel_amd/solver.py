"""Python host mirror of the reference's projection-path module procedures.

Each method calls one C-ABI entry point of libc3d.so and keeps the name and
argument meaning of the reference routine it replaces:

========================================  ======================================
reference (file:line)                     here
========================================  ======================================
``solve_poisson_equation(rhs)``           :meth:`Channel.solve_poisson_equation`
src/poisson.f90:33-101
``compute_non_IB_projection``             :meth:`Channel.compute_non_IB_projection`
src/projection.f90:22-46
``apply_boundary_conditions(U,V,W)``      :meth:`Channel.apply_boundary_conditions`
src/boundary_conditions.f90:20-55
``divergence(div,U,V,W)``                 :meth:`Channel.divergence`
src/differential_operators.f90:54-74
``gradient(Gpx,Gpy,Gpz,p)``               :meth:`Channel.gradient`
src/differential_operators.f90:16-52
``check_divergence(max_div)``             :meth:`Channel.check_divergence`
src/monitor.f90:210-221
========================================  ======================================

Arrays are numpy float64 in the reference's Fortran layout.  The reference
stops the program on fatal conditions and ignores solver status; here every
failure raises :class:`C3DError` with the library's message.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import _lib
from ._lib import fptr, iptr, ptr


class C3DError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_lib.ERRORS.get(code, code)}: {msg}")
        self.code = code


class Channel:
    """One GPU-resident channel state (the reference's module-global state)."""

    def __init__(self, nx, ny, nz, Lxp=2 * np.pi, Lzp=np.pi, Ly=2.0, alpha=2.0, grid_type=1,
                 nu=1.0, precision=_lib.PREC_FP64, min_buffer_width=0.0):
        self.L = _lib.load()
        d = _lib.c3d_desc(nx, ny, nz, grid_type, Lxp, Lzp, Ly, alpha, nu, precision,
                          min_buffer_width)
        h = C.c_void_p()
        rc = self.L.c3d_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise C3DError(rc, f"c3d_create({nx},{ny},{nz}) failed")
        self.h = h
        self.nu = nu
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxg, self.nyg, self.nzg = nx + 1, ny + 1, nz + 1
        self.nxp, self.nzp, self.nym = nx - 2, nz - 2, ny - 1

    def close(self):
        if getattr(self, "h", None):
            self.L.c3d_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, rc):
        if rc != 0:
            raise C3DError(rc, self.L.c3d_last_error(self.h).decode())

    # shapes ----------------------------------------------------------------
    shape_u = property(lambda s: (s.nx, s.nyg, s.nzg))
    shape_v = property(lambda s: (s.nxg, s.ny, s.nzg))
    shape_w = property(lambda s: (s.nxg, s.nyg, s.nz))
    shape_p = property(lambda s: (s.nxg, s.nyg, s.nzg))
    shape_rhs = property(lambda s: (s.nxg - 2, s.nyg - 2, s.nzg - 1))

    def zeros(self, which):
        return np.zeros(getattr(self, "shape_" + which), order="F")

    def grid(self, what: int) -> np.ndarray:
        out = np.zeros(max(self.nxg, self.nyg, self.nzg) + 16)
        n = self.L.c3d_grid_get(self.h, what, ptr(out), out.size)
        if n < 0:
            self._chk(n)
        return out[:n].copy()

    # state -------------------------------------------------------------------
    def set_fields(self, U=None, V=None, W=None, P=None):
        for a, s in ((U, self.shape_u), (V, self.shape_v), (W, self.shape_w), (P, self.shape_p)):
            if a is not None and a.shape != s:
                raise ValueError(f"shape {a.shape} != {s}")
        self._chk(self.L.c3d_set_fields(self.h, ptr(U), ptr(V), ptr(W), ptr(P)))

    def get_fields(self):
        U, V, W, P, Pint = (self.zeros("u"), self.zeros("v"), self.zeros("w"), self.zeros("p"),
                            self.zeros("rhs"))
        self._chk(self.L.c3d_get_fields(self.h, ptr(U), ptr(V), ptr(W), ptr(P), ptr(Pint)))
        return dict(U=U, V=V, W=W, P=P, P_interim=Pint)

    # reference procedures ----------------------------------------------------
    def _uvw(self, U, V, W):
        out = []
        for a, s in ((U, self.shape_u), (V, self.shape_v), (W, self.shape_w)):
            if a.shape != s or a.dtype != np.float64 or not a.flags["F_CONTIGUOUS"]:
                raise ValueError(f"expected a Fortran-ordered float64 array of shape {s}")
            out.append(a)
        return out

    def apply_boundary_conditions(self, U=None, V=None, W=None):
        """apply_boundary_conditions(U_, V_, W_): on the resident velocity
        without arguments, else in place on the given host arrays (the
        resident velocity is then left alone)."""
        if U is None and V is None and W is None:
            self._chk(self.L.c3d_apply_boundary_conditions(self.h))
            return None
        U, V, W = self._uvw(U, V, W)
        self._chk(self.L.c3d_apply_boundary_conditions_arrays(self.h, ptr(U), ptr(V), ptr(W)))
        return U, V, W

    def divergence(self, U=None, V=None, W=None):
        """divergence(div_, U_, V_, W_) of the resident velocity, or of the
        given host arrays."""
        d = self.zeros("rhs")
        if U is None and V is None and W is None:
            self._chk(self.L.c3d_divergence(self.h, ptr(d)))
        else:
            U, V, W = self._uvw(U, V, W)
            self._chk(self.L.c3d_divergence_arrays(self.h, ptr(U), ptr(V), ptr(W), ptr(d)))
        return d

    def gradient(self, p):
        gx, gy, gz = self.zeros("u"), self.zeros("v"), self.zeros("w")
        p = np.asfortranarray(p, dtype=np.float64)
        self._chk(self.L.c3d_gradient(self.h, ptr(p), ptr(gx), ptr(gy), ptr(gz)))
        return gx, gy, gz

    def solve_poisson_equation(self, rhs, rk_step=2, dt=1.0):
        """In place on a Fortran-ordered rhs_p-shaped array."""
        if rhs.shape != self.shape_rhs or not rhs.flags["F_CONTIGUOUS"]:
            raise ValueError("rhs must be Fortran-ordered with shape %s" % (self.shape_rhs,))
        self._chk(self.L.c3d_solve_poisson_equation(self.h, ptr(rhs), rk_step, dt))
        return rhs

    def compute_non_IB_projection(self, dt, rk_step=2):
        self._chk(self.L.c3d_compute_non_ib_projection(self.h, dt, rk_step))

    def check_divergence(self) -> float:
        v = C.c_double()
        self._chk(self.L.c3d_check_divergence(self.h, C.byref(v)))
        return v.value

    MONITOR_NAMES = ("maxU", "maxV", "maxW", "meanU", "meanV", "meanW", "nonfinite")

    def monitor(self):
        """output_monitor's field summaries (src/monitor.f90:27-39, 57-59) of the
        resident fields, plus the count of non-finite entries."""
        out = np.zeros(7)
        self._chk(self.L.c3d_monitor(self.h, ptr(out)))
        return dict(zip(self.MONITOR_NAMES, out.tolist()))

    STAT_NAMES = ("Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean")

    def plane_statistics(self):
        """Plane-mean profiles of compute_statistics (src/statistics.f90:41-111)
        from the resident fields, as a dict of [nyg] arrays."""
        n = self.ny + 1
        out = np.zeros(7 * n)
        self._chk(self.L.c3d_plane_statistics(self.h, ptr(out)))
        return {k: out[q * n:(q + 1) * n].copy() for q, k in enumerate(self.STAT_NAMES)}

    def wall_shear(self):
        """(utau, wtau, Umean[nyg], Wmean[nyg]) of compute_statistics
        (src/statistics.f90:41-131) from the resident fields."""
        ut, wt = C.c_double(), C.c_double()
        um, wm = np.zeros(self.ny + 1), np.zeros(self.ny + 1)
        self._chk(self.L.c3d_wall_shear(self.h, C.byref(ut), C.byref(wt), ptr(um), ptr(wm)))
        return ut.value, wt.value, um, wm

    # RK3 time integration (src/time_integration.f90, src/equations.f90) -------
    def set_pressure_gradient(self, dPdx=0.0, dPdy=0.0, dPdz=0.0):
        self._chk(self.L.c3d_set_pressure_gradient(self.h, dPdx, dPdy, dPdz))

    def get_pressure_gradient(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self._chk(self.L.c3d_get_pressure_gradient(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def save_state(self):
        self._chk(self.L.c3d_save_state(self.h))

    def compute_rhs(self, slot=1):
        """compute_rhs_u/v/w of the resident velocity into stage slot 1..3."""
        self._chk(self.L.c3d_compute_rhs(self.h, slot))

    def get_rhs(self, slot=1):
        fu = np.zeros((self.nx - 2, self.nyg - 2, self.nzg - 2), order="F")
        fv = np.zeros((self.nxg - 2, self.ny - 2, self.nzg - 2), order="F")
        fw = np.zeros((self.nxg - 2, self.nyg - 2, self.nz - 2), order="F")
        self._chk(self.L.c3d_get_rhs(self.h, slot, ptr(fu), ptr(fv), ptr(fw)))
        return fu, fv, fw

    def compute_rhs_arrays(self, U, V, W):
        """compute_rhs_u/v/w(U_, V_, W_, rhs) of host arrays -> (Fu, Fv, Fw)
        with the interior shapes of get_rhs."""
        U, V, W = self._uvw(U, V, W)
        fu = np.zeros((self.nx - 2, self.nyg - 2, self.nzg - 2), order="F")
        fv = np.zeros((self.nxg - 2, self.ny - 2, self.nzg - 2), order="F")
        fw = np.zeros((self.nxg - 2, self.nyg - 2, self.nz - 2), order="F")
        self._chk(self.L.c3d_compute_rhs_arrays(self.h, ptr(U), ptr(V), ptr(W), ptr(fu), ptr(fv),
                                                ptr(fw)))
        return fu, fv, fw

    def compute_rhs_u(self, U, V, W):
        return self.compute_rhs_arrays(U, V, W)[0]

    def compute_rhs_v(self, U, V, W):
        return self.compute_rhs_arrays(U, V, W)[1]

    def compute_rhs_w(self, U, V, W):
        return self.compute_rhs_arrays(U, V, W)[2]

    def rk3_update(self, rk_step, dt):
        self._chk(self.L.c3d_rk3_update(self.h, rk_step, dt))

    def compute_dt(self, cfl):
        v = C.c_double()
        self._chk(self.L.c3d_compute_dt(self.h, cfl, C.byref(v)))
        return v.value

    def mean_mass_flow_u(self, U=None):
        """compute_mean_mass_flow_U of the resident U, or of a host array."""
        v = C.c_double()
        if U is None:
            self._chk(self.L.c3d_mean_mass_flow_u(self.h, C.byref(v)))
        else:
            if U.shape != self.shape_u or not U.flags["F_CONTIGUOUS"] or U.dtype != np.float64:
                raise ValueError(f"expected a Fortran-ordered float64 array of shape {self.shape_u}")
            self._chk(self.L.c3d_mean_mass_flow_u_array(self.h, ptr(U), C.byref(v)))
        return v.value

    def add_to_u(self, d):
        self._chk(self.L.c3d_add_to_u(self.h, d))

    def set_rhs_exact(self, on=True):
        """Bitwise (reference divisions, no FMA) right-hand sides in
        compute_time_step_RK3 instead of the fused round-off-level pass."""
        self._chk(self.L.c3d_set_rhs_exact(self.h, int(bool(on))))

    def compute_time_step_RK3(self, dt, x_mass_cte=0, q0=0.0):
        """One compute_time_step_RK3 (src/time_integration.f90:86-193)."""
        self._chk(self.L.c3d_time_step_rk3(self.h, dt, x_mass_cte, q0))

    # immersed boundary (src/immersed_boundary_operators.f90, src/heaviside.f90,
    # src/projection.f90:48-186) ------------------------------------------------
    def setup_IB(self, m):
        """setup_IB_operators + compute_heaviside for a marker set ``m``
        (attributes xb, yb, zb, normals, sb, ub, nzb, nxb, nxb1 and, if
        present, yref = y_ref_index per marker)."""
        arrs = [np.ascontiguousarray(a, dtype=np.float64)
                for a in (m.xb, m.yb, m.zb, m.normals, m.sb, m.ub)]
        self.nb = int(arrs[0].size)
        yref = getattr(m, "yref", None)
        yr = None if yref is None else np.ascontiguousarray(yref, dtype=np.int32)
        self._chk(self.L.c3d_ib_setup_ref(self.h, self.nb, *[ptr(a) for a in arrs],
                                          None if yr is None else iptr(yr),
                                          int(m.nzb), int(m.nxb), int(m.nxb1)))

    def write_snapshot(self, path, dt, tangents_1=None, tangents_2=None):
        """output_data: the reference's big-endian snapshot of the resident
        state (fields, Heaviside masks, t, dt, dpdx, nu, body, fb)."""
        t1 = None if tangents_1 is None else np.ascontiguousarray(tangents_1, dtype=np.float64)
        t2 = None if tangents_2 is None else np.ascontiguousarray(tangents_2, dtype=np.float64)
        self._chk(self.L.c3d_write_snapshot(self.h, os.fsencode(path), dt,
                                            None if t1 is None else ptr(t1),
                                            None if t2 is None else ptr(t2)))

    def read_snapshot(self, path, what=_lib.RESTORE_ALL):
        """read_input_data: U, V, W from a snapshot (the reference's restart)
        plus, per ``what``, P, t and dPdx, and fb.  Returns the recorded dt."""
        dt = np.zeros(1)
        self._chk(self.L.c3d_read_snapshot(self.h, os.fsencode(path), what, ptr(dt)))
        return float(dt[0])

    WAVE_BODIES = {"standing_wave": 1, "traveling_wave_x": 2}

    def setup_IB_wave(self, body, nxb, nzb, p1, p2, p3, ramp=0.0, t=None):
        """Moving wall body (standing_wave or traveling_wave_x) built on the
        device at time t (default: the context's time)."""
        t = self.time if t is None else t
        self._chk(self.L.c3d_ib_setup_wave(self.h, self.WAVE_BODIES[body], nxb, nzb, p1, p2, p3,
                                           ramp, t))
        self.nb = 2 * nxb * nzb

    def move_IB(self, t):
        self._chk(self.L.c3d_ib_move(self.h, t))

    def get_IB_markers(self):
        """Marker arrays of the current body: xb, yb, zb, sb, yref [nb];
        normals, ub [3nb]."""
        nb = self.nb
        xb, yb, zb, sb = (np.zeros(nb) for _ in range(4))
        nrm, ub = np.zeros(3 * nb), np.zeros(3 * nb)
        yref = np.zeros(nb, dtype=np.int32)
        self._chk(self.L.c3d_ib_get_markers(self.h, ptr(xb), ptr(yb), ptr(zb), ptr(nrm), ptr(sb),
                                            ptr(ub), iptr(yref)))
        return dict(xb=xb, yb=yb, zb=zb, normals=nrm, sb=sb, ub=ub, yref=yref)

    @property
    def time(self):
        t = np.zeros(1)
        self._chk(self.L.c3d_get_time(self.h, ptr(t)))
        return float(t[0])

    @time.setter
    def time(self, t):
        self._chk(self.L.c3d_set_time(self.h, float(t)))

    def ib_stencil(self, s):
        n = self.nb * 125
        xi, yi, zi = (np.zeros(n, dtype=np.int32) for _ in range(3))
        w, dxn = np.zeros(n), np.zeros(n)
        self._chk(self.L.c3d_ib_get_stencil(self.h, s, iptr(xi), iptr(yi), iptr(zi), ptr(w),
                                            ptr(dxn)))
        sh = (self.nb, 125)
        return xi.reshape(sh), yi.reshape(sh), zi.reshape(sh), w.reshape(sh), dxn.reshape(sh)

    def ib_heaviside(self):
        a, b = np.zeros(self.nb), np.zeros(3 * self.nb)
        Hu, Hv, Hw, Hc = self.zeros("u"), self.zeros("v"), self.zeros("w"), self.zeros("rhs")
        self._chk(self.L.c3d_ib_get_heaviside(self.h, ptr(a), ptr(b), ptr(Hu), ptr(Hv), ptr(Hw),
                                              ptr(Hc)))
        return dict(E1nHc=a, E1nH=b, Hu_int=Hu, Hv_int=Hv, Hw_int=Hw, Hc_ext=Hc)

    def regT(self, with_dxn=False, U=None, V=None, W=None):
        """E (regT) or E1n (regT_1n) of the resident velocity, or of the
        given host arrays -> [3nb]."""
        out = np.zeros(3 * self.nb)
        if U is None and V is None and W is None:
            self._chk(self.L.c3d_ib_interpolate(self.h, int(with_dxn), ptr(out)))
        else:
            U, V, W = self._uvw(U, V, W)
            self._chk(self.L.c3d_ib_interpolate_arrays(self.h, int(with_dxn), ptr(U), ptr(V),
                                                       ptr(W), ptr(out)))
        return out

    def compute_heaviside(self):
        """compute_heaviside (src/heaviside.f90:20-54) for the current markers."""
        self._chk(self.L.c3d_ib_compute_heaviside(self.h))

    def regTc_1n(self, p):
        p = np.asfortranarray(p, dtype=np.float64)
        out = np.zeros(self.nb)
        self._chk(self.L.c3d_ib_interpolate_c(self.h, ptr(p), ptr(out)))
        return out

    def reg(self, comp, f):
        """regu / regv / regw (comp 0, 1, 2) of f[nb] -> field."""
        F = self.zeros("uvw"[comp])
        self._chk(self.L.c3d_ib_spread(self.h, comp, ptr(np.ascontiguousarray(f, dtype=np.float64)),
                                       ptr(F)))
        return F

    def remove_mean_per_body(self, f):
        f = np.array(f, dtype=np.float64)
        self._chk(self.L.c3d_ib_remove_mean(self.h, ptr(f)))
        return f

    def schur(self, f, dt, rk_step=2):
        Sf = np.zeros(4 * self.nb)
        self._chk(self.L.c3d_ib_schur(self.h, ptr(np.ascontiguousarray(f, dtype=np.float64)),
                                      ptr(Sf), dt, rk_step))
        return Sf

    def bicgstab(self, b, x0=None, tol=1e-10, maxit=1000, dt=1.0, rk_step=2):
        """Returns (x, iterations, ||r||^2); raises C3DError on no convergence."""
        x = np.zeros(4 * self.nb) if x0 is None else np.array(x0, dtype=np.float64)
        it, err = C.c_int(), C.c_double()
        self._chk(self.L.c3d_ib_bicgstab(self.h, ptr(x), ptr(np.ascontiguousarray(b, dtype=np.float64)),
                                         tol, maxit, dt, rk_step, C.byref(it), C.byref(err)))
        return x, it.value, err.value

    def set_IB_solver(self, tol=1e-10, maxit=1000, guess_mode=0):
        self._chk(self.L.c3d_ib_set_solver(self.h, tol, maxit, guess_mode))

    def set_IB_graphs(self, on=True):
        """Replay each BiCGSTAB iteration from a captured HIP graph (default)
        or launch it kernel by kernel; the results are identical."""
        self._chk(self.L.c3d_ib_set_graphs(self.h, 1 if on else 0))

    def set_IB_force(self, fb):
        self._chk(self.L.c3d_ib_set_force(self.h, ptr(np.ascontiguousarray(fb, dtype=np.float64))))

    def get_IB_force(self):
        fb = np.zeros(4 * self.nb)
        self._chk(self.L.c3d_ib_get_force(self.h, ptr(fb)))
        return fb

    def compute_IB_projection(self, dt, rk_step=2):
        """Returns (iterations, ||r||^2)."""
        it, err = C.c_int(), C.c_double()
        self._chk(self.L.c3d_compute_ib_projection(self.h, dt, rk_step, C.byref(it), C.byref(err)))
        return it.value, err.value

    def check_slip(self) -> float:
        v = C.c_double()
        self._chk(self.L.c3d_check_slip(self.h, C.byref(v)))
        return v.value

    def ib_stats(self):
        a, b, e = C.c_int(), C.c_longlong(), C.c_double()
        self._chk(self.L.c3d_ib_stats(self.h, C.byref(a), C.byref(b), C.byref(e)))
        return a.value, b.value, e.value

    def set_exterior_pressure_gradient(self, flag):
        self._chk(self.L.c3d_ib_set_exterior_pressure_gradient(self.h, int(flag)))

    # learned initial guess for the Schur solve -------------------------------------
    def mlp_init(self, seed=1):
        self._chk(self.L.c3d_mlp_init(self.h, seed))

    def mlp_set_scales(self, iscale, oscale):
        self._chk(self.L.c3d_mlp_set_scales(self.h, fptr(np.asarray(iscale, dtype=np.float32)),
                                            fptr(np.asarray(oscale, dtype=np.float32))))

    def mlp_params(self):
        p = np.zeros(self.L.c3d_mlp_param_count(), dtype=np.float32)
        self._chk(self.L.c3d_mlp_get_params(self.h, fptr(p)))
        return p

    def mlp_set_params(self, p):
        self._chk(self.L.c3d_mlp_set_params(self.h, fptr(np.ascontiguousarray(p, dtype=np.float32))))

    def mlp_forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 4), dtype=np.float32)
        self._chk(self.L.c3d_mlp_forward(self.h, fptr(X), X.shape[0], fptr(Y)))
        return Y

    def mlp_gradients(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        G = np.zeros(self.L.c3d_mlp_param_count(), dtype=np.float32)
        loss = C.c_float()
        self._chk(self.L.c3d_mlp_gradients(self.h, fptr(X), fptr(T), X.shape[0], fptr(G),
                                           C.byref(loss)))
        return G, loss.value

    def mlp_train(self, X, T, epochs, batch=256, lr=1e-3, seed=7):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        hist = np.zeros(epochs, dtype=np.float32)
        self._chk(self.L.c3d_mlp_train(self.h, fptr(X), fptr(T), X.shape[0], epochs, batch, lr,
                                       seed, fptr(hist)))
        return hist

    def mlp_gemm(self, A, B, ta=False, tb=False):
        A = np.ascontiguousarray(A, dtype=np.float32)
        B = np.ascontiguousarray(B, dtype=np.float32)
        M, K = (A.shape[1], A.shape[0]) if ta else A.shape
        N = B.shape[0] if tb else B.shape[1]
        Cm = np.zeros((M, N), dtype=np.float32)
        self._chk(self.L.c3d_mlp_gemm_test(self.h, M, N, K, int(ta), int(tb), fptr(A), fptr(B),
                                           fptr(Cm)))
        return Cm

    def ib_last_pair(self):
        """(features [nb, 54] unscaled, force increment [4nb]) of the last IB solve."""
        X = np.zeros((self.nb, 54), dtype=np.float32)
        dF = np.zeros(4 * self.nb)
        self._chk(self.L.c3d_ib_last_pair(self.h, fptr(X), ptr(dF)))
        return X, dF

    # mixed precision -----------------------------------------------------------
    def set_refinement(self, nref):
        """FP64 refinement sweeps per projection of a precision='mixed' channel."""
        self._chk(self.L.c3d_set_refinement(self.h, nref))

    def refinement_residuals(self):
        out = np.zeros(8)
        n = self.L.c3d_refinement_residuals(self.h, ptr(out), out.size)
        if n < 0:
            self._chk(n)
        return out[:n].copy()

    # measurement ---------------------------------------------------------------
    def copy_bandwidth(self, nbytes, iters=20):
        """Average ms of a device streaming copy of nbytes (calibration)."""
        ms = C.c_double()
        self._chk(self.L.c3d_copy_bandwidth(self.h, nbytes, iters, C.byref(ms)))
        return ms.value

    def time_projection(self, iters, dt=1.0):
        ns = self.L.c3d_stage_count()
        tot = C.c_double()
        ks = np.zeros(ns)
        self._chk(self.L.c3d_time_projection(self.h, iters, dt, C.byref(tot), ptr(ks), ns))
        names = [self.L.c3d_stage_name(i).decode() for i in range(ns)]
        return tot.value, dict(zip(names, ks))
