"""Python host binding of the MI355X SWM path (ctypes over libswmgpu.so).

``SwmGpu`` plays the role the reference's Fortran module procedures play for
``src/model.f90``: it is built from the host-side initialisation
(:class:`swm_amd.host.Setup`), takes over the state produced by
``SWM_initialConditions`` / ``TRC_initTracer`` and then advances it with
``step`` (SWM_timestep + TRC_timestep + model_advance).  Fields stay on the GPU
and are copied back by ``get`` (the register read a due diag task performs).

There is no CPU fallback: if the HIP library cannot be loaded or no GPU is
present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import host as _host

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libswmgpu.so")
ABI_VERSION = 2

FULLY_NONLINEAR, LINEARISED_MEAN_STATE, LINEARISED_STATE_OF_REST = 1, 2, 3
LATERAL_MIXING, QUADRATIC_BOTTOM_FRICTION, BAROTROPIC, TAU_SCALE, ELLIPTIC_SOLVER = 1, 2, 4, 8, 16
FP32_LOW_PART = 32              # SWM_FP32_LOW_PART
FP64, FP32 = 0, 1

EXPORTED = [
    "swm_gpu_create", "swm_gpu_upload_state", "swm_gpu_add_tracer", "swm_gpu_step", "swm_gpu_itt",
    "swm_gpu_sync", "swm_gpu_sor", "swm_gpu_sor_fixed", "swm_gpu_velocity_potential", "swm_gpu_divergence",
    "swm_gpu_sums",
    "swm_gpu_last_timing", "swm_gpu_last_error", "swm_gpu_destroy",
    "swm_gpu_batch_create", "swm_gpu_batch_attach", "swm_gpu_batch_detach", "swm_gpu_batch_set_steps",
    "swm_gpu_batch_remaining", "swm_gpu_batch_step", "swm_gpu_batch_last_timing", "swm_gpu_batch_last_error",
    "swm_gpu_batch_destroy", "swm_gpu_kernel_shape_count", "swm_gpu_set_kernel_shape", "swm_gpu_tune",
    "swm_gpu_set_sor_check_every", "swm_gpu_batch_set_shape", "swm_gpu_batch_tune", "swm_gpu_avg_add",
    "swm_gpu_avg_read", "swm_gpu_add_forcing_stream", "swm_gpu_sums_noalloc", "swm_gpu_elliptic_stats",
    "swm_gpu_add_dyn_stream", "swm_gpu_batch_shape_count", "swm_gpu_set_launch_mode", "swm_gpu_launch_mode",
    "swm_gpu_set_sor_halo", "swm_gpu_set_sor_tile_rows",
]

_dp = C.POINTER(C.c_double)
_i8p = C.POINTER(C.c_int8)


class GridDesc(C.Structure):
    _fields_ = [("ocean", _i8p), ("H", _dp), ("cos_lat", _dp), ("tan_lat", _dp), ("f", _dp)]


class Desc(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("variant", C.c_int32),
        ("flags", C.c_int32), ("precision", C.c_int32),
        ("G", C.c_double), ("Ah", C.c_double), ("dt", C.c_double), ("A", C.c_double), ("dLambda", C.c_double),
        ("dTheta", C.c_double), ("RHO0", C.c_double), ("minD", C.c_double), ("lbc", C.c_double),
        ("ab_chi", C.c_double), ("tau_scale", C.c_double),
        ("grid", GridDesc * 4),
        ("impl_u", _dp), ("impl_v", _dp), ("impl_eta", _dp), ("gamma_sq_u", _dp), ("gamma_sq_v", _dp),
        ("F_x_const", _dp), ("F_y_const", _dp), ("F_eta_const", _dp), ("tau_x", _dp), ("tau_y", _dp),
        ("osc_re", _dp * 3), ("osc_im", _dp * 3), ("osc_omega", C.c_double * 3),
        ("u_bs", _dp), ("v_bs", _dp), ("zeta_bs", _dp),
    ]


class TracerDesc(C.Structure):
    _fields_ = [("kappa_h", C.c_double), ("CH", _dp), ("C", _dp), ("G_CH", _dp), ("C0", _dp),
                ("gamma_C", _dp), ("impl", _dp)]


class SwmGpuError(RuntimeError):
    pass


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libswmgpu.so (built by ``__graft_entry__.build``).  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise SwmGpuError(f"{path} is missing: build the HIP extension first (__graft_entry__.build())")
    L = C.CDLL(path)
    p = C.c_void_p
    L.swm_gpu_create.argtypes = [C.POINTER(Desc), C.POINTER(p)]
    L.swm_gpu_create.restype = C.c_int
    L.swm_gpu_upload_state.argtypes = [p, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int64]
    L.swm_gpu_upload_state.restype = C.c_int
    L.swm_gpu_add_tracer.argtypes = [p, C.POINTER(TracerDesc), C.POINTER(C.c_int32)]
    L.swm_gpu_add_tracer.restype = C.c_int
    L.swm_gpu_add_forcing_stream.argtypes = [p, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int32, _dp,
                                             C.POINTER(C.c_int32)]
    L.swm_gpu_add_forcing_stream.restype = C.c_int
    L.swm_gpu_add_dyn_stream.argtypes = [p, C.c_int32, C.c_int32, C.c_double, C.c_int32, _dp]
    L.swm_gpu_add_dyn_stream.restype = C.c_int
    L.swm_gpu_step.argtypes = [p, C.c_int64]
    L.swm_gpu_step.restype = C.c_int
    L.swm_gpu_itt.argtypes = [p]
    L.swm_gpu_itt.restype = C.c_int64
    L.swm_gpu_sync.argtypes = [p, C.c_char_p, C.c_int32, _dp]
    L.swm_gpu_sync.restype = C.c_int
    L.swm_gpu_sor.argtypes = [p, _dp, _dp, C.c_double, C.c_int32, C.POINTER(C.c_int64), _dp]
    L.swm_gpu_sor.restype = C.c_int
    L.swm_gpu_sor_fixed.argtypes = [p, _dp, _dp, C.c_double, C.c_double, C.c_int64, C.c_int32,
                                    C.POINTER(C.c_int64), _dp]
    L.swm_gpu_sor_fixed.restype = C.c_int
    L.swm_gpu_velocity_potential.argtypes = [p, C.c_double, C.c_double, C.c_int64, C.c_int32, C.c_int32,
                                             C.POINTER(C.c_int64), _dp]
    L.swm_gpu_velocity_potential.restype = C.c_int
    L.swm_gpu_divergence.argtypes = [p, _dp]
    L.swm_gpu_divergence.restype = C.c_int
    L.swm_gpu_sums.argtypes = [p, C.c_int32, _dp]
    L.swm_gpu_sums.restype = C.c_int
    L.swm_gpu_sums_noalloc.argtypes = [p, C.c_int32, _dp]
    L.swm_gpu_sums_noalloc.restype = C.c_int
    L.swm_gpu_elliptic_stats.argtypes = [p, C.POINTER(C.c_int64), _dp]
    L.swm_gpu_elliptic_stats.restype = C.c_int
    L.swm_gpu_last_timing.argtypes = [p, _dp, _dp]
    L.swm_gpu_last_timing.restype = C.c_int
    L.swm_gpu_last_error.argtypes = [p]
    L.swm_gpu_last_error.restype = C.c_char_p
    L.swm_gpu_batch_create.argtypes = [C.POINTER(p)]
    L.swm_gpu_batch_create.restype = C.c_int
    L.swm_gpu_batch_attach.argtypes = [p, p, C.c_int64]
    L.swm_gpu_batch_attach.restype = C.c_int
    L.swm_gpu_batch_detach.argtypes = [p, p]
    L.swm_gpu_batch_detach.restype = C.c_int
    L.swm_gpu_batch_set_steps.argtypes = [p, p, C.c_int64]
    L.swm_gpu_batch_set_steps.restype = C.c_int
    L.swm_gpu_batch_remaining.argtypes = [p, p, C.POINTER(C.c_int64)]
    L.swm_gpu_batch_remaining.restype = C.c_int
    L.swm_gpu_batch_step.argtypes = [p, C.c_int64, C.POINTER(C.c_int32)]
    L.swm_gpu_batch_step.restype = C.c_int
    L.swm_gpu_batch_last_timing.argtypes = [p, _dp]
    L.swm_gpu_batch_last_timing.restype = C.c_int
    L.swm_gpu_batch_last_error.argtypes = [p]
    L.swm_gpu_batch_last_error.restype = C.c_char_p
    L.swm_gpu_batch_destroy.argtypes = [p]
    L.swm_gpu_batch_destroy.restype = None
    L.swm_gpu_kernel_shape_count.argtypes = []
    L.swm_gpu_kernel_shape_count.restype = C.c_int
    L.swm_gpu_batch_shape_count.argtypes = []
    L.swm_gpu_batch_shape_count.restype = C.c_int32
    L.swm_gpu_set_launch_mode.argtypes = [p, C.c_int32]
    L.swm_gpu_set_launch_mode.restype = C.c_int
    L.swm_gpu_launch_mode.argtypes = [p, C.POINTER(C.c_int32), _dp, _dp]
    L.swm_gpu_launch_mode.restype = C.c_int
    L.swm_gpu_set_sor_halo.argtypes = [p, C.c_int32]
    L.swm_gpu_set_sor_halo.restype = C.c_int
    L.swm_gpu_set_sor_tile_rows.argtypes = [p, C.c_int32]
    L.swm_gpu_set_sor_tile_rows.restype = C.c_int
    L.swm_gpu_set_kernel_shape.argtypes = [p, C.c_int32]
    L.swm_gpu_set_kernel_shape.restype = C.c_int
    L.swm_gpu_tune.argtypes = [p, C.c_int32, C.POINTER(C.c_int32), _dp]
    L.swm_gpu_tune.restype = C.c_int
    L.swm_gpu_set_sor_check_every.argtypes = [p, C.c_int32]
    L.swm_gpu_set_sor_check_every.restype = C.c_int
    L.swm_gpu_batch_set_shape.argtypes = [p, C.c_int32]
    L.swm_gpu_batch_set_shape.restype = C.c_int
    L.swm_gpu_batch_tune.argtypes = [p, C.c_int32, C.POINTER(C.c_int32), _dp]
    L.swm_gpu_batch_tune.restype = C.c_int
    L.swm_gpu_avg_add.argtypes = [p, C.c_char_p, C.c_int32, C.c_int32, C.POINTER(C.c_int32)]
    L.swm_gpu_avg_add.restype = C.c_int
    L.swm_gpu_avg_read.argtypes = [p, C.c_int32, _dp, _dp, C.c_int32]
    L.swm_gpu_avg_read.restype = C.c_int
    L.swm_gpu_destroy.argtypes = [p]
    L.swm_gpu_destroy.restype = None
    _lib = L
    return L


def _arr(a, keep, dtype=np.float64):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dtype)
    keep.append(a)
    return a.ctypes.data_as(_i8p if dtype == np.int8 else _dp)


def _flags(s) -> int:
    f = 0
    if s.lateral_mixing:
        f |= LATERAL_MIXING
    if s.quadratic_friction:
        f |= QUADRATIC_BOTTOM_FRICTION
    if s.barotropic:
        f |= BAROTROPIC
    if s.tau_scale is not None:
        f |= TAU_SCALE
    if getattr(s, "elliptic_solver", False):
        f |= ELLIPTIC_SOLVER
    return f


def initial_state(s):
    """State after SWM_initialConditions (src/swm_module.f90:170-179): masked ICs,
    zero AB history, itt = 0."""
    u = s.grids["u"].ocean.astype(np.float64) * s.u0
    v = s.grids["v"].ocean.astype(np.float64) * s.v0
    eta = s.grids["eta"].ocean.astype(np.float64) * s.eta0
    return u, v, eta


def initial_D(s, eta):
    """D after computeD at initialisation (src/swm_timestep_module.f90:182-200)."""
    He = s.grids["eta"].H
    if s.variant == _host.FULLY_NONLINEAR:
        return np.where(s.grids["eta"].ocean == 1, eta + He, 1.0)
    return He.copy()


class SwmGpu:
    """Device-resident SWM model (C-ABI context) built from a host Setup."""

    def __init__(self, s, precision: int = FP64, low_part: bool = False):
        """low_part: an FP32 context carries a bfloat16 low part of its state
        (SWM_FP32_LOW_PART)."""
        self.L = load_library()
        self.s = s
        self.precision = precision
        keep = self._keep = []
        d = Desc()
        d.abi_version = ABI_VERSION
        d.Nx, d.Ny, d.variant, d.precision = s.Nx, s.Ny, s.variant, precision
        d.flags = _flags(s) | (FP32_LOW_PART if low_part else 0)
        for name in ("G", "Ah", "dt", "A", "dLambda", "dTheta", "RHO0", "minD", "lbc", "ab_chi"):
            setattr(d, name, float(getattr(s, name)))
        d.tau_scale = float(s.tau_scale) if s.tau_scale is not None else 0.0
        for gi, gname in enumerate(("H", "u", "v", "eta")):
            g = s.grids[gname]
            d.grid[gi].ocean = _arr(g.ocean, keep, np.int8)
            d.grid[gi].H = _arr(g.H, keep)
            d.grid[gi].cos_lat = _arr(g.cos_lat, keep)
            d.grid[gi].tan_lat = _arr(g.tan_lat, keep)
            d.grid[gi].f = _arr(g.f, keep)
        d.impl_u, d.impl_v, d.impl_eta = _arr(s.impl_u, keep), _arr(s.impl_v, keep), _arr(s.impl_eta, keep)
        d.gamma_sq_u, d.gamma_sq_v = _arr(s.gamma_sq_u, keep), _arr(s.gamma_sq_v, keep)
        d.F_x_const, d.F_y_const, d.F_eta_const = (_arr(s.F_x_const, keep), _arr(s.F_y_const, keep),
                                                   _arr(s.F_eta_const, keep))
        for comp, tau in s.wind:
            if comp == _host.COMP_X:
                if d.tau_x:
                    raise SwmGpuError("one zonal windstress stream is supported")
                d.tau_x = _arr(tau, keep)
            else:
                if d.tau_y:
                    raise SwmGpuError("one meridional windstress stream is supported")
                d.tau_y = _arr(tau, keep)
        for comp, re, im, om in s.osc:
            if d.osc_re[comp]:
                raise SwmGpuError("one oscillating stream per component is supported")
            d.osc_re[comp] = _arr(re, keep)
            d.osc_im[comp] = _arr(im, keep)
            d.osc_omega[comp] = om
        if s.variant == _host.LINEARISED_MEAN_STATE:
            d.u_bs, d.v_bs, d.zeta_bs = _arr(s.u_bs, keep), _arr(s.v_bs, keep), _arr(s.zeta_bs, keep)
        self._desc = d
        h = C.c_void_p()
        rc = self.L.swm_gpu_create(C.byref(d), C.byref(h))
        if rc != 0:
            raise SwmGpuError(f"swm_gpu_create failed ({rc}): {self.L.swm_gpu_last_error(None).decode()}")
        self._h = h
        self._keep = []   # inputs are copied to the device during create
        for st in getattr(s, "streams", []):
            self.add_forcing_stream(st.kind, st.comp, st.records, st.rec_dt, st.chunk_size)
        for st in getattr(s, "dyn", []):
            self.add_dyn_stream(st.comp, st.records, st.rec_dt, st.chunk_size)

    # -- error handling -------------------------------------------------------
    def _check(self, rc: int, what: str):
        if rc != 0:
            raise SwmGpuError(f"{what} failed ({rc}): {self.L.swm_gpu_last_error(self._h).decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self.L.swm_gpu_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- hand-over ------------------------------------------------------------
    def upload_state(self, u, v, eta, G_u=None, G_v=None, G_eta=None, itt: int = 0):
        keep = []
        rc = self.L.swm_gpu_upload_state(self._h, _arr(u, keep), _arr(v, keep), _arr(eta, keep),
                                         _arr(G_u, keep), _arr(G_v, keep), _arr(G_eta, keep), int(itt))
        self._check(rc, "swm_gpu_upload_state")

    def add_tracer(self, kappa_h, CH, C_, G_CH=None, C0=None, gamma_C=None, impl=None) -> int:
        keep = []
        t = TracerDesc(float(kappa_h), _arr(CH, keep), _arr(C_, keep), _arr(G_CH, keep), _arr(C0, keep),
                       _arr(gamma_C, keep), _arr(impl, keep))
        tid = C.c_int32()
        self._check(self.L.swm_gpu_add_tracer(self._h, C.byref(t), C.byref(tid)), "swm_gpu_add_tracer")
        return tid.value

    def add_forcing_stream(self, kind: int, comp: int, records, rec_dt: float, chunk_size: int = 100) -> int:
        """A time-interpolated forcing stream (swm_gpu_add_forcing_stream):
        records (nrec, Ny, Nx) spaced rec_dt model seconds."""
        rec = np.ascontiguousarray(records, dtype=np.float64)
        if rec.ndim == 2:
            rec = rec[None]
        if rec.shape[1:] != (self.s.Ny, self.s.Nx):
            raise SwmGpuError("records must be (nrec, Ny, Nx)")
        sid = C.c_int32()
        self._check(self.L.swm_gpu_add_forcing_stream(self._h, int(kind), int(comp), rec.shape[0], float(rec_dt),
                                                      int(chunk_size), rec.ctypes.data_as(_dp), C.byref(sid)),
                    "swm_gpu_add_forcing_stream")
        return sid.value

    def add_dyn_stream(self, comp: int, records, rec_dt: float, chunk_size: int = 100):
        """A DYNFROMFILE dataset (swm_gpu_add_dyn_stream): records (nrec, Ny, Nx)
        of u, v or eta spaced rec_dt model seconds."""
        rec = np.ascontiguousarray(records, dtype=np.float64)
        if rec.ndim == 2:
            rec = rec[None]
        if rec.shape[1:] != (self.s.Ny, self.s.Nx):
            raise SwmGpuError("records must be (nrec, Ny, Nx)")
        self._check(self.L.swm_gpu_add_dyn_stream(self._h, int(comp), rec.shape[0], float(rec_dt), int(chunk_size),
                                                  rec.ctypes.data_as(_dp)), "swm_gpu_add_dyn_stream")

    def init_from_setup(self):
        """SWM_initialConditions + TRC_initTracer hand-over for a fresh run."""
        s = self.s
        u, v, eta = initial_state(s)
        self.upload_state(u, v, eta, itt=0)
        D = initial_D(s, eta)
        for t in s.tracers:
            CH = D * t["C"]
            impl = 1.0 + s.dt * t["cons"]
            self.add_tracer(t["kappa_h"], CH, t["C"], C0=t["C0"], gamma_C=t["gamma_C"], impl=impl)
        return self

    # -- stepping -------------------------------------------------------------
    def step(self, nsteps: int = 1):
        self._check(self.L.swm_gpu_step(self._h, int(nsteps)), "swm_gpu_step")

    @property
    def itt(self) -> int:
        return int(self.L.swm_gpu_itt(self._h))

    def last_timing(self):
        tot, ker = C.c_double(), C.c_double()
        self._check(self.L.swm_gpu_last_timing(self._h, C.byref(tot), C.byref(ker)), "swm_gpu_last_timing")
        return tot.value, ker.value

    # -- register reads -------------------------------------------------------
    def get(self, name: str, tracer: int = -1) -> np.ndarray:
        out = np.empty((self.s.Ny, self.s.Nx), dtype=np.float64)
        self._check(self.L.swm_gpu_sync(self._h, name.encode(), int(tracer), out.ctypes.data_as(_dp)),
                    f"swm_gpu_sync({name})")
        return out

    def elliptic_stats(self):
        """(sweeps, anorm) of the last computeVelocityPotential solve of the
        elliptic-solver build (swm_gpu_elliptic_stats)."""
        it, an = C.c_int64(), C.c_double()
        self._check(self.L.swm_gpu_elliptic_stats(self._h, C.byref(it), C.byref(an)), "swm_gpu_elliptic_stats")
        return it.value, an.value

    def divergence(self) -> np.ndarray:
        out = np.empty((self.s.Ny, self.s.Nx), dtype=np.float64)
        self._check(self.L.swm_gpu_divergence(self._h, out.ctypes.data_as(_dp)), "swm_gpu_divergence")
        return out

    def velocity_potential(self, omega=1.9, rel_tol=1e-10, max_sweeps=5000, fp32_sweeps=False, reset=False):
        """Device-resident computeVelocityPotential: returns (iters, rel_res, converged);
        the solution stays on the device (get("CHI"))."""
        it, rr = C.c_int64(), C.c_double()
        rc = self.L.swm_gpu_velocity_potential(self._h, float(omega), float(rel_tol), int(max_sweeps),
                                               int(bool(fp32_sweeps)), int(bool(reset)), C.byref(it), C.byref(rr))
        if rc not in (0, 1):
            self._check(rc, "swm_gpu_velocity_potential")
        return it.value, rr.value, rc == 0

    def avg_add(self, name: str, tracer: int = -1, square: bool = False) -> int:
        """Register an averaging diag buffer (dt * field**p added after every step)."""
        aid = C.c_int32()
        self._check(self.L.swm_gpu_avg_add(self._h, name.encode(), int(tracer), int(bool(square)), C.byref(aid)),
                    "swm_gpu_avg_add")
        return aid.value

    def avg_read(self, avg_id: int, reset: bool = True):
        """(buffer / bufferCount, bufferCount); resets the buffer when asked."""
        out = np.empty((self.s.Ny, self.s.Nx), dtype=np.float64)
        w = C.c_double()
        self._check(self.L.swm_gpu_avg_read(self._h, int(avg_id), out.ctypes.data_as(_dp), C.byref(w),
                                            int(bool(reset))), "swm_gpu_avg_read")
        return out, w.value

    def set_kernel_shape(self, shape: int):
        self._check(self.L.swm_gpu_set_kernel_shape(self._h, int(shape)), "swm_gpu_set_kernel_shape")

    def tune(self, steps: int = 20):
        """Time every kernel shape from a snapshot (state restored exactly);
        selects the fastest.  Returns (best_shape, per-step ms per shape)."""
        n = self.L.swm_gpu_kernel_shape_count()
        times = (C.c_double * n)()
        best = C.c_int32()
        self._check(self.L.swm_gpu_tune(self._h, int(steps), C.byref(best), C.cast(times, _dp)), "swm_gpu_tune")
        return best.value, list(times)

    def set_sor_check_every(self, every: int):
        self._check(self.L.swm_gpu_set_sor_check_every(self._h, int(every)), "swm_gpu_set_sor_check_every")

    def set_sor_halo(self, halo: int):
        """LDS halo of the blocked fixed-omega SOR sweeps in cells (2 x sweeps
        per launch: 2, 4, 8; 0 = one launch per colour; -1 = from the cadence)."""
        self._check(self.L.swm_gpu_set_sor_halo(self._h, int(halo)), "swm_gpu_set_sor_halo")

    def set_sor_tile_rows(self, rows: int):
        """Rows of the blocked SOR sweeps' 64-column tile (16 or 32)."""
        self._check(self.L.swm_gpu_set_sor_tile_rows(self._h, int(rows)), "swm_gpu_set_sor_tile_rows")

    def set_launch_mode(self, graphs: bool):
        """Replay captured six-step graphs (True) or launch each step (False)."""
        self._check(self.L.swm_gpu_set_launch_mode(self._h, int(bool(graphs))), "swm_gpu_set_launch_mode")

    def launch_mode(self):
        """(graphs, ms per step direct, ms per step graphs) after a tuning run."""
        g, a, b = C.c_int32(), C.c_double(), C.c_double()
        self._check(self.L.swm_gpu_launch_mode(self._h, C.byref(g), C.byref(a), C.byref(b)), "swm_gpu_launch_mode")
        return bool(g.value), a.value, b.value

    def sums(self, tracer: int = -1):
        """FP64 compensated (volume, energy, tracer inventory) sums, cos-weighted
        over ocean eta points (swm_gpu_sums)."""
        out = (C.c_double * 3)()
        self._check(self.L.swm_gpu_sums(self._h, int(tracer), C.cast(out, _dp)), "swm_gpu_sums")
        return float(out[0]), float(out[1]), float(out[2])

    # -- elliptic solver ------------------------------------------------------
    def sor(self, B, chi, epsilon=1e-4, first_guess=True):
        """main_ElSolv_SOR contract; returns (chi, iters, anorm, converged)."""
        B = np.ascontiguousarray(B, np.float64)
        chi = np.array(chi, dtype=np.float64, copy=True)
        it, an = C.c_int64(), C.c_double()
        rc = self.L.swm_gpu_sor(self._h, B.ctypes.data_as(_dp), chi.ctypes.data_as(_dp), float(epsilon),
                                int(bool(first_guess)), C.byref(it), C.byref(an))
        if rc not in (0, 1):
            self._check(rc, "swm_gpu_sor")
        return chi, it.value, an.value, rc == 0

    def sor_fixed(self, B, chi, omega=1.9, rel_tol=1e-10, max_sweeps=5000, fp32_sweeps=False):
        B = np.ascontiguousarray(B, np.float64)
        chi = np.array(chi, dtype=np.float64, copy=True)
        it, rr = C.c_int64(), C.c_double()
        rc = self.L.swm_gpu_sor_fixed(self._h, B.ctypes.data_as(_dp), chi.ctypes.data_as(_dp), float(omega),
                                      float(rel_tol), int(max_sweeps), int(bool(fp32_sweeps)), C.byref(it),
                                      C.byref(rr))
        if rc not in (0, 1):
            self._check(rc, "swm_gpu_sor_fixed")
        return chi, it.value, rr.value, rc == 0


class SwmBatch:
    """Batched ensemble (swm_gpu_batch_*): attached SwmGpu members advance
    together, one launch per step covering every member's tiles."""

    def __init__(self):
        self.L = load_library()
        h = C.c_void_p()
        rc = self.L.swm_gpu_batch_create(C.byref(h))
        if rc != 0:
            raise SwmGpuError(f"swm_gpu_batch_create failed ({rc}): {self.L.swm_gpu_batch_last_error(None).decode()}")
        self._h = h
        self.members = []      # keeps attached contexts alive

    def _check(self, rc: int, what: str):
        if rc != 0:
            raise SwmGpuError(f"{what} failed ({rc}): {self.L.swm_gpu_batch_last_error(self._h).decode()}")

    def attach(self, m: "SwmGpu", nsteps: int):
        self._check(self.L.swm_gpu_batch_attach(self._h, m._h, int(nsteps)), "swm_gpu_batch_attach")
        self.members.append(m)

    def detach(self, m: "SwmGpu"):
        self._check(self.L.swm_gpu_batch_detach(self._h, m._h), "swm_gpu_batch_detach")
        self.members.remove(m)

    def set_steps(self, m: "SwmGpu", nsteps: int):
        self._check(self.L.swm_gpu_batch_set_steps(self._h, m._h, int(nsteps)), "swm_gpu_batch_set_steps")

    def remaining(self, m: "SwmGpu") -> int:
        r = C.c_int64()
        self._check(self.L.swm_gpu_batch_remaining(self._h, m._h, C.byref(r)), "swm_gpu_batch_remaining")
        return r.value

    def step(self, nsteps: int) -> int:
        """Advance every member with steps left by up to nsteps; returns the
        number of members with steps left afterwards."""
        a = C.c_int32()
        self._check(self.L.swm_gpu_batch_step(self._h, int(nsteps), C.byref(a)), "swm_gpu_batch_step")
        return a.value

    def set_shape(self, shape: int):
        self._check(self.L.swm_gpu_batch_set_shape(self._h, int(shape)), "swm_gpu_batch_set_shape")

    def tune(self, steps: int = 10):
        """Time each batch tile and launch form from a snapshot of every member
        (restored exactly); selects the fastest.  Returns (best_shape, per-step
        ms per form)."""
        times = (C.c_double * self.L.swm_gpu_batch_shape_count())()
        best = C.c_int32()
        self._check(self.L.swm_gpu_batch_tune(self._h, int(steps), C.byref(best), C.cast(times, _dp)),
                    "swm_gpu_batch_tune")
        return best.value, list(times)

    def last_ms(self) -> float:
        t = C.c_double()
        self._check(self.L.swm_gpu_batch_last_timing(self._h, C.byref(t)), "swm_gpu_batch_last_timing")
        return t.value

    def close(self):
        if getattr(self, "_h", None):
            self.L.swm_gpu_batch_destroy(self._h)
            self._h = None
            self.members = []

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
