"""CPU oracle for the SWM hot path -- TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg,
as the checker / baseline.  The product (swm_amd) never imports it.

``Oracle(setup)`` wraps liboracle.so (swm_oracle.c, a pass-for-pass C
restatement of the reference Fortran).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
MAX_STREAMS = 4
MAX_TRACERS = 4

_dp = C.POINTER(C.c_double)
_i8p = C.POINTER(C.c_int8)


class _Grid(C.Structure):
    _fields_ = [("H", _dp), ("ocean", _i8p), ("cos_lat", _dp), ("tan_lat", _dp), ("f", _dp)]


class _Desc(C.Structure):
    _fields_ = [
        ("Nx", C.c_int32), ("Ny", C.c_int32), ("variant", C.c_int32),
        ("lateral_mixing", C.c_int32), ("quadratic_friction", C.c_int32), ("barotropic", C.c_int32),
        ("has_tau_scale", C.c_int32), ("n_wind", C.c_int32), ("n_osc", C.c_int32), ("n_tracer", C.c_int32),
        ("tau_scale", C.c_double), ("G", C.c_double), ("Ah", C.c_double), ("dt", C.c_double),
        ("A", C.c_double), ("dLambda", C.c_double), ("dTheta", C.c_double), ("RHO0", C.c_double),
        ("minD", C.c_double), ("lbc", C.c_double), ("ab_chi", C.c_double),
        ("grid", _Grid * 4),
        ("impl_u", _dp), ("impl_v", _dp), ("impl_eta", _dp), ("gamma_sq_u", _dp), ("gamma_sq_v", _dp),
        ("F_x_const", _dp), ("F_y_const", _dp), ("F_eta_const", _dp),
        ("wind_comp", C.c_int32 * MAX_STREAMS), ("wind_tau", _dp * MAX_STREAMS),
        ("osc_comp", C.c_int32 * MAX_STREAMS), ("osc_re", _dp * MAX_STREAMS), ("osc_im", _dp * MAX_STREAMS),
        ("osc_omega", C.c_double * MAX_STREAMS),
        ("u_bs", _dp), ("v_bs", _dp), ("zeta_bs", _dp),
        ("u0", _dp), ("v0", _dp), ("eta0", _dp),
        ("kappa_h", C.c_double * MAX_TRACERS),
        ("trc_C", _dp * MAX_TRACERS), ("trc_C0", _dp * MAX_TRACERS),
        ("trc_gamma_C", _dp * MAX_TRACERS), ("trc_cons", _dp * MAX_TRACERS),
        ("n_tdf", C.c_int32), ("tdf_kind", C.c_int32 * MAX_STREAMS), ("tdf_comp", C.c_int32 * MAX_STREAMS),
        ("tdf_nrec", C.c_int32 * MAX_STREAMS), ("tdf_chunk", C.c_int32 * MAX_STREAMS),
        ("tdf_dt", C.c_double * MAX_STREAMS), ("tdf_rec", _dp * MAX_STREAMS),
        ("n_dff", C.c_int32), ("dff_field", C.c_int32 * 3), ("dff_nrec", C.c_int32 * 3),
        ("dff_chunk", C.c_int32 * 3), ("dff_dt", C.c_double * 3), ("dff_rec", _dp * 3),
    ]


def build(force: bool = False) -> str:
    src = os.path.join(_HERE, "swm_oracle.c")
    if force or not os.path.exists(_LIB) or os.path.getmtime(_LIB) < os.path.getmtime(src):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)
    return _LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(_Desc)]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_step.argtypes = [C.c_void_p, C.c_int64]
        L.orc_itt.restype = C.c_int64
        L.orc_itt.argtypes = [C.c_void_p]
        L.orc_get.argtypes = [C.c_void_p, C.c_char_p, C.c_int, _dp]
        L.orc_get.restype = C.c_int
        L.orc_sor.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_int, C.POINTER(C.c_int64), _dp]
        L.orc_sor.restype = C.c_int
        L.orc_sor_fixed.argtypes = [C.c_void_p, _dp, _dp, C.c_double, C.c_double, C.c_int64,
                                    C.POINTER(C.c_int64), _dp]
        L.orc_sor_fixed.restype = C.c_int
        L.orc_divergence.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_psi.argtypes = [C.c_void_p, _dp, _dp, _dp]
        L.orc_nondivergent.argtypes = [C.c_void_p, _dp, _dp, _dp, _dp, _dp]
        L.orc_sor_apply.argtypes = [C.c_void_p, _dp, _dp]
        _lib = L
    return _lib


def _ptr(a, keep):
    if a is None:
        return None
    a = np.ascontiguousarray(a)
    keep.append(a)
    if a.dtype == np.int8:
        return a.ctypes.data_as(_i8p)
    assert a.dtype == np.float64, a.dtype
    return a.ctypes.data_as(_dp)


class Oracle:
    """Reference-order CPU model built from a swm_amd.host.Setup."""

    def __init__(self, s):
        self.s = s
        self._keep = []
        k = self._keep
        d = _Desc()
        d.Nx, d.Ny, d.variant = s.Nx, s.Ny, s.variant
        d.lateral_mixing, d.quadratic_friction, d.barotropic = int(s.lateral_mixing), int(s.quadratic_friction), int(s.barotropic)
        d.has_tau_scale = int(s.tau_scale is not None)
        d.tau_scale = s.tau_scale if s.tau_scale is not None else 0.0
        for name in ("G", "Ah", "dt", "A", "dLambda", "dTheta", "RHO0", "minD", "lbc", "ab_chi"):
            setattr(d, name, float(getattr(s, name)))
        for gi, gname in enumerate(("H", "u", "v", "eta")):
            g = s.grids[gname]
            d.grid[gi].H = _ptr(np.asarray(g.H, np.float64), k)
            d.grid[gi].ocean = _ptr(np.asarray(g.ocean, np.int8), k)
            d.grid[gi].cos_lat = _ptr(g.cos_lat, k)
            d.grid[gi].tan_lat = _ptr(g.tan_lat, k)
            d.grid[gi].f = _ptr(g.f, k)
        for name in ("impl_u", "impl_v", "impl_eta", "gamma_sq_u", "gamma_sq_v", "F_x_const", "F_y_const",
                     "F_eta_const", "u_bs", "v_bs", "zeta_bs", "u0", "v0", "eta0"):
            v = getattr(s, name)
            setattr(d, name, _ptr(None if v is None else np.asarray(v, np.float64), k))
        assert len(s.wind) <= MAX_STREAMS and len(s.osc) <= MAX_STREAMS and len(s.tracers) <= MAX_TRACERS
        d.n_wind = len(s.wind)
        for n, (comp, tau) in enumerate(s.wind):
            d.wind_comp[n] = comp
            d.wind_tau[n] = _ptr(tau, k)
        d.n_osc = len(s.osc)
        for n, (comp, re, im, om) in enumerate(s.osc):
            d.osc_comp[n] = comp
            d.osc_re[n] = _ptr(re, k)
            d.osc_im[n] = _ptr(im, k)
            d.osc_omega[n] = om
        streams = getattr(s, "streams", [])
        assert len(streams) <= MAX_STREAMS
        d.n_tdf = len(streams)
        for n, st in enumerate(streams):
            d.tdf_kind[n], d.tdf_comp[n] = st.kind, st.comp
            d.tdf_nrec[n], d.tdf_chunk[n], d.tdf_dt[n] = st.records.shape[0], st.chunk_size, st.rec_dt
            d.tdf_rec[n] = _ptr(np.ascontiguousarray(st.records, np.float64), k)
        dyn = getattr(s, "dyn", [])
        assert len(dyn) <= 3
        d.n_dff = len(dyn)
        for n, st in enumerate(dyn):
            d.dff_field[n], d.dff_nrec[n], d.dff_chunk[n], d.dff_dt[n] = (st.comp, st.records.shape[0], st.chunk_size,
                                                                         st.rec_dt)
            d.dff_rec[n] = _ptr(np.ascontiguousarray(st.records, np.float64), k)
        d.n_tracer = len(s.tracers)
        for n, t in enumerate(s.tracers):
            d.kappa_h[n] = t["kappa_h"]
            d.trc_C[n] = _ptr(t["C"], k)
            d.trc_C0[n] = _ptr(t["C0"], k)
            d.trc_gamma_C[n] = _ptr(t["gamma_C"], k)
            d.trc_cons[n] = _ptr(t["cons"], k)
        self._desc = d
        self._h = lib().orc_create(C.byref(d))
        if not self._h:
            raise RuntimeError("orc_create failed")

    def close(self):
        if self._h:
            lib().orc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        lib().orc_step(self._h, int(n))

    @property
    def itt(self) -> int:
        return lib().orc_itt(self._h)

    def get(self, name: str, tracer: int = -1) -> np.ndarray:
        out = np.empty(self.s.shape, dtype=np.float64)
        if lib().orc_get(self._h, name.encode(), tracer, out.ctypes.data_as(_dp)) != 0:
            raise KeyError(name)
        return out

    def divergence(self, u, v):
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        out = np.empty(self.s.shape)
        lib().orc_divergence(self._h, u.ctypes.data_as(_dp), v.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def nondivergent(self, u, v, chi):
        """computeNonDivergentFlowField of (u, v) given the velocity potential
        chi (the elliptic-solver build): (u - dchi/dx, v - dchi/dy)."""
        u = np.ascontiguousarray(u, np.float64)
        v = np.ascontiguousarray(v, np.float64)
        chi = np.ascontiguousarray(chi, np.float64)
        un, vn = np.empty(self.s.shape), np.empty(self.s.shape)
        lib().orc_nondivergent(self._h, u.ctypes.data_as(_dp), v.ctypes.data_as(_dp), chi.ctypes.data_as(_dp),
                               un.ctypes.data_as(_dp), vn.ctypes.data_as(_dp))
        return un, vn

    def psi(self, u=None, v=None):
        """computeStreamfunction of (u, v); default the master u, v at N0."""
        u = self.get("u") if u is None else np.ascontiguousarray(u, np.float64)
        v = self.get("v") if v is None else np.ascontiguousarray(v, np.float64)
        out = np.empty(self.s.shape)
        lib().orc_psi(self._h, u.ctypes.data_as(_dp), v.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def sor_apply(self, chi):
        chi = np.ascontiguousarray(chi, np.float64)
        out = np.empty(self.s.shape)
        lib().orc_sor_apply(self._h, chi.ctypes.data_as(_dp), out.ctypes.data_as(_dp))
        return out

    def sor(self, B, chi, epsilon=1e-4, first_guess=True):
        B = np.ascontiguousarray(B, np.float64)
        chi = np.array(chi, dtype=np.float64, copy=True)
        it = C.c_int64()
        an = C.c_double()
        rc = lib().orc_sor(self._h, B.ctypes.data_as(_dp), chi.ctypes.data_as(_dp), float(epsilon),
                           int(bool(first_guess)), C.byref(it), C.byref(an))
        return chi, it.value, an.value, rc

    def sor_fixed(self, B, chi, omega=1.9, rel_tol=1e-10, max_sweeps=5000):
        B = np.ascontiguousarray(B, np.float64)
        chi = np.array(chi, dtype=np.float64, copy=True)
        it = C.c_int64()
        rr = C.c_double()
        rc = lib().orc_sor_fixed(self._h, B.ctypes.data_as(_dp), chi.ctypes.data_as(_dp), float(omega),
                                 float(rel_tol), int(max_sweeps), C.byref(it), C.byref(rr))
        return chi, it.value, rr.value, rc
