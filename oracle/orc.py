"""ctypes wrapper of the C oracle (test infrastructure, see package docstring).

Arrays are numpy float64 arrays in Fortran order with the reference's shapes:
``U(nx,nyg,nzg)``, ``V(nxg,ny,nzg)``, ``W(nxg,nyg,nz)``, ``P(nxg,nyg,nzg)``
and ``rhs_p(nxg-2, nyg-2, nzg-1)`` (reference bounds ``2:nxg-1, 2:nyg-1,
2:nzg``), see src/initialization.f90:116-155.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "liborc.so")
_dp = C.POINTER(C.c_double)


def build(force: bool = False) -> str:
    """Compile liborc.so with the committed Makefile."""
    if force or not os.path.exists(_SO):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _SO


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_SO):
            build()
        L = C.CDLL(_SO)
        L.orc_grid_new.restype = C.c_void_p
        L.orc_grid_new.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                   C.c_double, C.c_double, C.c_int, C.c_double]
        L.orc_grid_free.argtypes = [C.c_void_p]
        L.orc_grid_get.argtypes = [C.c_void_p, C.c_int, _dp]
        L.orc_grid_get.restype = C.c_int
        L.orc_apply_bc.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_divergence.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_gradient.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_fft.argtypes = [C.c_int, C.c_int, _dp]
        L.orc_fft.restype = C.c_int
        L.orc_zgtsv.argtypes = [C.c_int, _dp, _dp, _dp, _dp]
        L.orc_zgtsv.restype = C.c_int
        L.orc_poisson.argtypes = [C.c_void_p, _dp, C.c_int, C.c_double, _dp, C.c_int, C.c_int]
        L.orc_projection.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                     C.c_int, C.c_double, C.c_int, C.c_int]
        L.orc_max_abs_div.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_max_abs_div.restype = C.c_double
        for f in (L.orc_rhs_u, L.orc_rhs_v, L.orc_rhs_w):
            f.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, C.c_double, C.c_double, _dp]
        L.orc_rk_update.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_compute_dt.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_double, C.c_double]
        L.orc_compute_dt.restype = C.c_double
        L.orc_mean_mass_flow_u.argtypes = [C.c_void_p, _dp, _dp]
        L.orc_mean_mass_flow_u.restype = C.c_double
        L.orc_wall_shear.argtypes = [C.c_void_p, _dp, _dp, C.c_double, _dp, _dp, _dp, _dp]
        L.orc_wall_shear.restype = None
        L.orc_plane_statistics.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_plane_statistics.restype = None
        L.orc_monitor.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
        L.orc_monitor.restype = None
        L.orc_time_step_rk3.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, C.c_double,
                                        C.c_double, _dp, C.c_double, C.c_int]
        _lib = L
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 or a.dtype == np.complex128
    assert a.flags["F_CONTIGUOUS"] or a.ndim == 1, "oracle arrays must be Fortran-contiguous"
    return a.ctypes.data_as(_dp)


def fft(a: np.ndarray, sign: int = -1) -> np.ndarray:
    """Unnormalised 1-D complex DFT (sign -1 forward, +1 backward)."""
    b = np.array(a, dtype=np.complex128, copy=True)
    if lib().orc_fft(b.size, sign, _p(b)) != 0:
        raise ValueError("unsupported FFT length %d" % b.size)
    return b


def zgtsv(dl, d, du, b):
    """Restated LAPACK zgtsv on copies; returns (x, info)."""
    dl = np.array(dl, dtype=np.complex128)
    d = np.array(d, dtype=np.complex128)
    du = np.array(du, dtype=np.complex128)
    x = np.array(b, dtype=np.complex128)
    info = lib().orc_zgtsv(d.size, _p(dl), _p(d), _p(du), _p(x))
    return x, info


class Grid:
    """Single-domain grid (src/input_output.f90:140-206, src/initialization.f90:166-390).

    ``nx, ny, nz`` are face counts as in the reference namelist
    (``nx_global`` ...).  ``grid_type`` 0 uniform y, 1 tanh-stretched y,
    2 tanh-stretched y with uniform buffers of at least
    ``min_buffer_width`` + 6 dymin at each wall (src/input_output.f90:180-247).
    """

    def __init__(self, nx, ny, nz, Lxp=2 * np.pi, Lzp=np.pi, Ly=2.0, alpha=2.0, grid_type=1,
                 min_buffer_width=0.0):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.nxg, self.nyg, self.nzg = nx + 1, ny + 1, nz + 1
        self.nxp, self.nzp, self.nym = nx - 2, nz - 2, ny - 1
        self.h = lib().orc_grid_new(nx, ny, nz, Lxp, Lzp, Ly, alpha, grid_type, min_buffer_width)
        if not self.h:
            raise ValueError("invalid grid")
        self.Lxp_in, self.Lzp_in, self.Ly_in, self.alpha, self.grid_type = Lxp, Lzp, Ly, alpha, grid_type
        self.min_buffer_width = min_buffer_width

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_grid_free(self.h)
            self.h = None

    def get(self, what: int) -> np.ndarray:
        sizes = {0: self.nx, 1: self.ny, 2: self.nz, 3: self.nxg, 4: self.nyg, 5: self.nzg,
                 6: self.nxp, 7: self.nzp, 8: self.nym, 9: self.nym - 1, 10: self.nym - 1, 11: 10,
                 12: 1}
        out = np.zeros(sizes[what])
        lib().orc_grid_get(self.h, what, _p(out))
        return out

    @property
    def n_uniform(self):
        """Uniform buffer points per wall (grid_type 2; 1 otherwise)."""
        return int(self.get(12)[0])

    @property
    def scalars(self):
        s = self.get(11)
        return dict(zip(["dx", "dz", "dxmin", "dymin", "dzmin", "Lx", "Ly", "Lz", "Lxp", "Lzp"], s))

    # shapes -------------------------------------------------------------
    @property
    def shape_u(self):
        return (self.nx, self.nyg, self.nzg)

    @property
    def shape_v(self):
        return (self.nxg, self.ny, self.nzg)

    @property
    def shape_w(self):
        return (self.nxg, self.nyg, self.nz)

    @property
    def shape_p(self):
        return (self.nxg, self.nyg, self.nzg)

    @property
    def shape_rhs(self):
        return (self.nxg - 2, self.nyg - 2, self.nzg - 1)

    def zeros(self, which):
        return np.zeros(getattr(self, "shape_" + which), order="F")

    # operators ----------------------------------------------------------
    def apply_bc(self, U, V, W):
        lib().orc_apply_bc(self.h, _p(U), _p(V), _p(W))

    def divergence(self, U, V, W):
        d = self.zeros("rhs")
        lib().orc_divergence(self.h, _p(d), _p(U), _p(V), _p(W))
        return d

    def gradient(self, p):
        gx, gy, gz = self.zeros("u"), self.zeros("v"), self.zeros("w")
        lib().orc_gradient(self.h, _p(gx), _p(gy), _p(gz), _p(p))
        return gx, gy, gz

    def poisson(self, rhs, rk_step=2, dt=1.0, P=None, tri_mode=0, nthreads=0):
        """In-place solve_poisson_equation (src/poisson.f90:33-101)."""
        lib().orc_poisson(self.h, _p(rhs), rk_step, dt, _p(P), tri_mode, nthreads)
        return rhs

    def projection(self, U, V, W, rk_step=2, dt=1.0, P=None, tri_mode=0, nthreads=0):
        """compute_non_IB_projection in place on U,V,W; returns dict of outputs."""
        Ui, Vi, Wi = self.zeros("u"), self.zeros("v"), self.zeros("w")
        Pint, rhs = self.zeros("rhs"), self.zeros("rhs")
        lib().orc_projection(self.h, _p(U), _p(V), _p(W), _p(Ui), _p(Vi), _p(Wi), _p(Pint),
                             _p(rhs), _p(P), rk_step, dt, tri_mode, nthreads)
        return dict(U_interim=Ui, V_interim=Vi, W_interim=Wi, P_interim=Pint, rhs_p=rhs)

    def max_abs_div(self, U, V, W) -> float:
        return lib().orc_max_abs_div(self.h, _p(U), _p(V), _p(W))

    # RK3 (c3d_oracle_rk.c) ---------------------------------------------------
    @property
    def shape_fu(self):
        return (self.nx - 2, self.nyg - 2, self.nzg - 2)

    @property
    def shape_fv(self):
        return (self.nxg - 2, self.ny - 2, self.nzg - 2)

    @property
    def shape_fw(self):
        return (self.nxg - 2, self.nyg - 2, self.nz - 2)

    def rhs(self, U, V, W, nu, dPdx=0.0, dPdy=0.0, dPdz=0.0, H=(None, None, None)):
        """compute_rhs_u/v/w (src/equations.f90) -> (Fu, Fv, Fw)."""
        fu, fv, fw = (np.zeros(s, order="F") for s in (self.shape_fu, self.shape_fv, self.shape_fw))
        lib().orc_rhs_u(self.h, _p(U), _p(V), _p(W), _p(H[0]), nu, dPdx, _p(fu))
        lib().orc_rhs_v(self.h, _p(U), _p(V), _p(W), _p(H[1]), nu, dPdy, _p(fv))
        lib().orc_rhs_w(self.h, _p(U), _p(V), _p(W), _p(H[2]), nu, dPdz, _p(fw))
        return fu, fv, fw

    def rk_update(self, which, stage, dt, F1, F2, F3, Xo, X):
        z = F1
        lib().orc_rk_update(self.h, which, stage, dt, _p(F1), _p(F2 if F2 is not None else z),
                            _p(F3 if F3 is not None else z), _p(Xo), _p(X))

    def compute_dt(self, U, V, W, cfl, nu) -> float:
        return lib().orc_compute_dt(self.h, _p(U), _p(V), _p(W), cfl, nu)

    def mean_mass_flow_u(self, U, Hu=None) -> float:
        return lib().orc_mean_mass_flow_u(self.h, _p(U), _p(Hu))

    MONITOR_NAMES = ("maxU", "maxV", "maxW", "meanU", "meanV", "meanW", "nonfinite")

    def monitor(self, U, V, W):
        """output_monitor's field summaries (src/monitor.f90:27-39, 57-59)."""
        out = np.zeros(7)
        lib().orc_monitor(self.h, _p(U), _p(V), _p(W), _p(out))
        return dict(zip(self.MONITOR_NAMES, out.tolist()))

    STAT_NAMES = ("Umean", "Vmean", "Wmean", "U2mean", "V2mean", "W2mean", "UVmean")

    def plane_statistics(self, U, V, W):
        """Plane-mean profiles of compute_statistics (src/statistics.f90:41-111)
        as a dict of [nyg] arrays."""
        out = np.zeros(7 * self.nyg)
        lib().orc_plane_statistics(self.h, _p(U), _p(V), _p(W), _p(out))
        return {k: out[q * self.nyg:(q + 1) * self.nyg].copy() for q, k in enumerate(self.STAT_NAMES)}

    def wall_shear(self, U, W, nu):
        """(utau, wtau, Umean[nyg], Wmean[nyg]) of compute_statistics
        (src/statistics.f90:41-131)."""
        ut, wt = C.c_double(), C.c_double()
        n = self.nyg
        um, wm = np.zeros(n), np.zeros(n)
        lib().orc_wall_shear(self.h, _p(U), _p(W), nu, C.cast(C.pointer(ut), _dp),
                             C.cast(C.pointer(wt), _dp), _p(um), _p(wm))
        return ut.value, wt.value, um, wm

    def time_step_rk3(self, U, V, W, P, dt, nu, dPdx, q0=float("nan"), nthreads=0):
        """One compute_time_step_RK3 in place; returns (new dPdx, P_interim)."""
        d = np.array([dPdx], dtype=np.float64)
        Pint = self.zeros("rhs")
        lib().orc_time_step_rk3(self.h, _p(U), _p(V), _p(W), _p(P), _p(Pint), dt, nu, _p(d), q0,
                                nthreads)
        return float(d[0]), Pint


_ip = C.POINTER(C.c_int)


def _ib_sigs(L):
    if getattr(L, "_ib_ready", False):
        return
    L.orc_deltafnc.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_deltafnc.restype = C.c_double
    L.orc_ib_new.restype = C.c_void_p
    L.orc_ib_new.argtypes = [C.c_void_p, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _ip,
                             C.c_int, C.c_int, C.c_int]
    L.orc_wave_geometry.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_double,
                                    C.c_double, C.c_double, C.c_double, C.c_double,
                                    C.c_double, _dp, _dp, _dp, _dp, _dp, _dp, _ip]
    L.orc_wave_geometry.restype = C.c_int
    L.orc_ib_free.argtypes = [C.c_void_p]
    L.orc_ib_get_stencil.argtypes = [C.c_void_p, C.c_int, _ip, _ip, _ip, _dp, _dp]
    L.orc_ib_regT.argtypes = [C.c_void_p, _dp, _dp, _dp, C.c_int, _dp]
    L.orc_ib_regTc_1n.argtypes = [C.c_void_p, _dp, _dp]
    L.orc_ib_reg.argtypes = [C.c_void_p, C.c_int, _dp, _dp]
    L.orc_ib_remove_mean.argtypes = [C.c_void_p, _dp]
    L.orc_ib_heaviside.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.orc_ib_get_heaviside.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.orc_ib_schur.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int]
    L.orc_ib_bicgstab.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.c_double,
                                  C.c_double, C.c_int, _dp, _dp, C.c_int]
    L.orc_ib_bicgstab.restype = C.c_int
    L.orc_ib_projection.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp,
                                    C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, _dp,
                                    C.c_int]
    L.orc_ib_projection.restype = C.c_int
    L.orc_ib_max_slip.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp]
    L.orc_ib_max_slip.restype = C.c_double
    L._ib_ready = True


def deltafnc(x1, x2, dr) -> float:
    L = lib()
    _ib_sigs(L)
    return L.orc_deltafnc(x1, x2, dr)


WAVE_BODIES = {"standing_wave": 1, "traveling_wave_x": 2}


class WaveMarkers:
    """Markers of a moving wall body at one time (orc_wave_geometry):
    ``xb, yb, zb, sb, yref`` [nb], ``normals, ub`` [3nb]; slab j holds nxb
    bottom-wall then nxb top-wall markers (``nxb`` here is 2 x the wall's
    marker count, ``nxb1`` the bottom wall's), so that the two walls are
    the two bodies of remove_mean_per_body.  ``stop`` is the 1-based index
    of the first marker outside [y(3), y(ny-2)] (0 if none)."""

    def __init__(self, grid: Grid, body: str, nxb: int, nzb: int, p1: float, p2: float,
                 p3: float, t: float, ramp: float = 0.0, Lyc: float = 2.0):
        L = lib()
        _ib_sigs(L)
        nb = 2 * nxb * nzb
        self.xb, self.yb, self.zb, self.sb = (np.zeros(nb) for _ in range(4))
        self.normals, self.ub = np.zeros(3 * nb), np.zeros(3 * nb)
        self.yref = np.zeros(nb, dtype=np.int32)
        self.stop = L.orc_wave_geometry(grid.h, WAVE_BODIES[body], nxb, nzb, p1, p2, p3, ramp, t,
                                        Lyc, _p(self.xb), _p(self.yb), _p(self.zb),
                                        _p(self.normals), _p(self.sb), _p(self.ub),
                                        self.yref.ctypes.data_as(_ip))
        if self.stop < 0:
            raise ValueError("invalid wave body")
        self.nb, self.nzb, self.nxb, self.nxb1 = nb, nzb, 2 * nxb, nxb


class IB:
    """Immersed-boundary operators on a uniform-y oracle grid.

    ``m`` is a marker set with attributes ``xb, yb, zb, normals, sb, ub,
    nzb, nxb, nxb1`` (see channel3d_parallel_amd.ib_geometry).
    """

    def __init__(self, grid: Grid, m):
        L = lib()
        _ib_sigs(L)
        self.g, self.m = grid, m
        self.nb = int(m.xb.size)
        self._keep = [np.ascontiguousarray(a, dtype=np.float64)
                      for a in (m.xb, m.yb, m.zb, m.normals, m.sb, m.ub)]
        yref = getattr(m, "yref", None)
        self._yref = None if yref is None else np.ascontiguousarray(yref, dtype=np.int32)
        self.h = L.orc_ib_new(grid.h, self.nb, *[_p(a) for a in self._keep],
                              None if self._yref is None else self._yref.ctypes.data_as(_ip),
                              m.nzb, m.nxb, m.nxb1)
        self.heaviside_done = False

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_ib_free(self.h)
            self.h = None

    def stencil(self, s):
        n = self.nb * 125
        xi, yi, zi = (np.zeros(n, dtype=np.int32) for _ in range(3))
        w, dxn = np.zeros(n), np.zeros(n)
        lib().orc_ib_get_stencil(self.h, s, xi.ctypes.data_as(_ip), yi.ctypes.data_as(_ip),
                                 zi.ctypes.data_as(_ip), _p(w), _p(dxn))
        sh = (self.nb, 125)
        return xi.reshape(sh), yi.reshape(sh), zi.reshape(sh), w.reshape(sh), dxn.reshape(sh)

    def E(self, U, V, W, with_dxn=False):
        out = np.zeros(3 * self.nb)
        lib().orc_ib_regT(self.h, _p(U), _p(V), _p(W), int(with_dxn), _p(out))
        return out

    def E1n_c(self, P):
        out = np.zeros(self.nb)
        lib().orc_ib_regTc_1n(self.h, _p(P), _p(out))
        return out

    def R(self, comp, f):
        F = self.g.zeros("uvw"[comp])
        lib().orc_ib_reg(self.h, comp, _p(np.ascontiguousarray(f, dtype=np.float64)), _p(F))
        return F

    def remove_mean(self, f):
        f = np.array(f, dtype=np.float64)
        lib().orc_ib_remove_mean(self.h, _p(f))
        return f

    def heaviside(self):
        g = self.g
        Hc, Hu, Hv, Hw = g.zeros("rhs"), g.zeros("u"), g.zeros("v"), g.zeros("w")
        lib().orc_ib_heaviside(self.h, _p(Hc), _p(Hu), _p(Hv), _p(Hw))
        self.heaviside_done = True
        return Hc, Hu, Hv, Hw

    def heaviside_coeffs(self):
        g = self.g
        a, b = np.zeros(self.nb), np.zeros(3 * self.nb)
        Hu, Hv, Hw = g.zeros("u"), g.zeros("v"), g.zeros("w")
        lib().orc_ib_get_heaviside(self.h, _p(a), _p(b), _p(Hu), _p(Hv), _p(Hw))
        return dict(E1nHc=a, E1nH=b, Hu_int=Hu, Hv_int=Hv, Hw_int=Hw)

    def schur(self, f, dt, nu, nthreads=0):
        assert self.heaviside_done
        Sf = np.zeros(4 * self.nb)
        lib().orc_ib_schur(self.h, _p(np.ascontiguousarray(f, dtype=np.float64)), _p(Sf),
                           dt, nu, nthreads)
        return Sf

    def bicgstab(self, b, x0=None, tol=1e-10, maxit=1000, dt=1.0, nu=1.0, rk_step=2, P=None,
                 nthreads=0):
        assert self.heaviside_done
        x = np.zeros(4 * self.nb) if x0 is None else np.array(x0, dtype=np.float64)
        err = np.zeros(1)
        it = lib().orc_ib_bicgstab(self.h, _p(x), _p(np.ascontiguousarray(b, dtype=np.float64)),
                                   tol, maxit, dt, nu, rk_step, _p(P), _p(err), nthreads)
        return x, it, float(err[0])

    def projection(self, U, V, W, Uint, Vint, Wint, Pint, fb, dt, nu, tol=1e-10, maxit=1000,
                   rk_step=2, P=None, nthreads=0):
        """compute_IB_projection in place on U, V, W and fb; returns (iters, ||r||^2)."""
        assert self.heaviside_done
        err = np.zeros(1)
        it = lib().orc_ib_projection(self.h, _p(U), _p(V), _p(W), _p(Uint), _p(Vint), _p(Wint),
                                     _p(Pint), _p(P), _p(fb), dt, nu, tol, maxit, rk_step,
                                     _p(err), nthreads)
        return it, float(err[0])

    def max_slip(self, U, V, W, fb) -> float:
        return lib().orc_ib_max_slip(self.h, _p(U), _p(V), _p(W), _p(fb))
