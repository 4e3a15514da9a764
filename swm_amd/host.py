"""Host-side initialisation feeding the SWM time-stepping path.

This restates, on the host and in NumPy, the reference's initialisation code
whose products the hot path consumes every step: the domain and the four
staggered grids (src/domain_module.f90:49-207, src/grid_module.f90:55-195),
the damping coefficients (src/swm_damping_module.f90:44-359), the forcing set-up
(src/swm_forcing_module.f90:78-127,592-646) and the initial conditions
(src/swm_module.f90:156-181, src/tracer_vars.f90:124-185).  In the drop-in
setting the reference's own Fortran does this work and hands the arrays to the
C-ABI; here it lets tests and the benchmark run the same configurations
without a Fortran toolchain.

Precision quirks of the reference are reproduced on purpose:

* ``PI`` and ``D2R`` are built from a *single-precision* literal
  (grid_module.f90:6-7, swm_damping_module.f90:282-283);
* default ``G`` and ``OMEGA`` are single-precision literals widened to double
  (vars_module.f90:27, domain_module.f90:13), ``TAU_SCALE`` likewise
  (a CPP literal such as ``1e-2``);
* trigonometric and exponential functions use the C library (``math``) as a
  gfortran build on Linux does, element by element on 1-D row vectors.

Fields are NumPy arrays of shape ``(Ny, Nx)`` -- the C view of the Fortran
``(Nx, Ny)`` arrays (i fastest).  Masks are ``int8`` ocean masks.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from . import namelist as _nml

__all__ = [
    "FULLY_NONLINEAR", "LINEARISED_MEAN_STATE", "LINEARISED_STATE_OF_REST",
    "Grid", "Setup", "f32", "init_domain", "setup_from_selfcheck",
]

FULLY_NONLINEAR = 1
LINEARISED_MEAN_STATE = 2
LINEARISED_STATE_OF_REST = 3
VARIANT_NAMES = {
    "FULLY_NONLINEAR": FULLY_NONLINEAR,
    "LINEARISED_MEAN_STATE": LINEARISED_MEAN_STATE,
    "LINEARISED_STATE_OF_REST": LINEARISED_STATE_OF_REST,
}

COMP_X, COMP_Y, COMP_ETA = 0, 1, 2


def f32(x: float) -> float:
    """A Fortran default-real (single precision) literal widened to double."""
    return float(np.float32(x))


PI = f32(3.14159265358979323846)     # grid_module.f90:6
D2R = PI / 180.0                      # grid_module.f90:7


@dataclass
class Grid:
    """grid_t (grid_module.f90:9-20) for one of the H, u, v, eta grids."""
    lon: np.ndarray
    lat: np.ndarray
    H: np.ndarray
    ocean: np.ndarray
    sin_lat: np.ndarray = None
    cos_lat: np.ndarray = None
    tan_lat: np.ndarray = None
    f: np.ndarray = None

    @property
    def land(self) -> np.ndarray:
        return (1 - self.ocean).astype(np.int8)

    def bc(self, lbc: float) -> np.ndarray:
        """setGridEuler boundary factor, grid_module.f90:188-194."""
        return np.where(self.ocean == 1, 1.0, lbc)


def _libm(fn, x: np.ndarray) -> np.ndarray:
    return np.array([fn(float(v)) for v in np.asarray(x, dtype=np.float64)], dtype=np.float64)


def _set_lat(g: Grid) -> None:
    """setLat, grid_module.f90:55-74 (C-library sin/cos/tan)."""
    arg = g.lat * D2R
    g.sin_lat = _libm(math.sin, arg)
    g.cos_lat = _libm(math.cos, arg)
    g.tan_lat = _libm(math.tan, arg)


def _set_f(g: Grid, coriolis_approx: int, theta0: float, OMEGA: float) -> None:
    """setf, grid_module.f90:116-145."""
    rtheta0 = theta0 * D2R
    if coriolis_approx == 1:
        g.f = np.full_like(g.lat, 2 * OMEGA * math.sin(rtheta0))
    elif coriolis_approx == 2:
        rtheta = D2R * g.lat
        g.f = 2 * OMEGA * math.sin(rtheta0) + (2 * OMEGA * math.cos(rtheta0)) * rtheta
    elif coriolis_approx == 3:
        g.f = (2 * OMEGA) * g.sin_lat
    elif coriolis_approx == 0:
        g.f = np.zeros_like(g.lat)
    else:
        raise ValueError("Invalid value of coriolis_approx in domain_nl.")


def _cs(a: np.ndarray, axis: int) -> np.ndarray:
    """cshift(a, 1, dim): element (i+1) along the Fortran dimension (periodic)."""
    return np.roll(a, -1, axis=axis)


def init_domain(Nx: int, Ny: int, lon_s: float, lon_e: float, lat_s: float, lat_e: float,
                H_in: Optional[np.ndarray] = None, H_missing: Optional[np.ndarray] = None,
                H_overwrite: float = 0.0, lbc: float = 2.0, coriolis_approx: int = 3,
                theta0: float = 0.0, OMEGA: float = f32(7.272205e-5), A: float = 6371000.0):
    """initDomain, domain_module.f90:49-207.  Returns (grids, dLambda, dTheta)."""
    dLambda = D2R * (lon_e - lon_s) / (Nx - 1)
    dTheta = D2R * (lat_e - lat_s) / (Ny - 1)
    j = np.arange(1, Ny + 1, dtype=np.float64)
    i = np.arange(1, Nx + 1, dtype=np.float64)
    lat_v = (j - 1) * (lat_e - lat_s) / (Ny - 1) + lat_s
    lat_H = lat_v.copy()
    lat_u = lat_v + (lat_e - lat_s) / f32(2.0 * (Ny - 1))
    lat_eta = lat_u.copy()
    lon_H = (i - 1) * (lon_e - lon_s) / (Nx - 1) + lon_s
    lon_u = lon_H.copy()
    lon_v = lon_H + (lon_e - lon_s) / f32(2.0 * (Nx - 1))
    lon_eta = lon_v.copy()

    if H_in is not None:
        H = np.array(H_in, dtype=np.float64)
        if H_missing is not None:
            H[H_missing.astype(bool)] = 0.0
        H[H <= 0.0] = 0.0
    else:
        H = np.ones((Ny, Nx))
        H[:, 0] = 0.0
        H[:, Nx - 1] = 0.0
    H[0, :] = 0.0
    H[Ny - 1, :] = 0.0

    land_H = (H == 0.0).astype(np.int8)
    land_u = ((land_H + _cs(land_H, 0)) == 2).astype(np.int8)
    land_v = ((land_H + _cs(land_H, 1)) == 2).astype(np.int8)
    land_eta = ((land_H + _cs(land_H, 0) + _cs(land_H, 1) + _cs(_cs(land_H, 1), 0)) == 4).astype(np.int8)
    ocean_H, ocean_u, ocean_v, ocean_eta = (1 - land_H, 1 - land_u, 1 - land_v, 1 - land_eta)
    ocean_H, ocean_u, ocean_v, ocean_eta = (a.astype(np.int8) for a in (ocean_H, ocean_u, ocean_v, ocean_eta))

    H_u = np.zeros_like(H)
    H_v = np.zeros_like(H)
    H_eta = np.zeros_like(H)
    H_jp = _cs(H, 0)
    H_ip = _cs(H, 1)
    H_ipjp = _cs(H_ip, 0)
    oH = ocean_H.astype(np.int64)
    with np.errstate(divide="ignore", invalid="ignore"):
        m = ocean_u != 0
        H_u[m] = ((H + H_jp) / (oH + _cs(oH, 0)))[m]
        m = ocean_v != 0
        H_v[m] = ((H + H_ip) / (oH + _cs(oH, 1)))[m]
        m = ocean_eta != 0
        H_eta[m] = ((((H + H_ip) + H_jp) + H_ipjp) / (((oH + _cs(oH, 1)) + _cs(oH, 0)) + _cs(_cs(oH, 1), 0)))[m]
    if H_overwrite != 0.0:
        H = ocean_H * H_overwrite
        H_u = ocean_u * H_overwrite
        H_v = ocean_v * H_overwrite
        H_eta = ocean_eta * H_overwrite

    grids = {
        "H": Grid(lon_H, lat_H, H, ocean_H),
        "u": Grid(lon_u, lat_u, H_u, ocean_u),
        "v": Grid(lon_v, lat_v, H_v, ocean_v),
        "eta": Grid(lon_eta, lat_eta, H_eta, ocean_eta),
    }
    for g in grids.values():
        _set_lat(g)
        _set_f(g, coriolis_approx, theta0, OMEGA)
    return grids, dLambda, dTheta


def noland_interp(D: np.ndarray, ocean_eta: np.ndarray, target: str) -> np.ndarray:
    """interp2/interp4 of an eta-grid field with the *_noland interpolators
    (calc_lib.f90:160-171, include/interp.inc) onto the H, u or v grid."""
    oe = ocean_eta.astype(np.int64)
    sh = lambda a, di, dj: np.roll(np.roll(a, di, axis=1), dj, axis=0)   # a(i-di, j-dj)
    if target == "u":       # eta2u_noland: points (i, j), (i-1, j)
        pts = [(0, 0), (1, 0)]
    elif target == "v":     # eta2v_noland: points (i, j), (i, j-1)
        pts = [(0, 0), (0, 1)]
    else:                   # eta2H_noland: (i,j), (i-1,j-1), (i-1,j), (i,j-1)
        pts = [(0, 0), (1, 1), (1, 0), (0, 1)]
    masks = [sh(oe, di, dj) for di, dj in pts]
    s = sum(masks).astype(np.float32).astype(np.float64)
    weight = 1.0 / np.maximum(s, 1.0)
    out = None
    for (di, dj), mk in zip(pts, masks):
        term = mk.astype(np.float32).astype(np.float64) * weight * sh(D, di, dj)
        out = term if out is None else out + term
    return out


STREAM_WIND, STREAM_CUSTOM, STREAM_HEATING = 0, 1, 2
DEF_FORCING_CHUNKSIZE = 100     # SWM_DEF_FORCING_CHUNKSIZE, include/swm_module.h:30


@dataclass
class ForcingStream:
    """A time-dependent forcing stream read through a memoryChunk
    (src/memchunk_module.f90): nrec (Ny, Nx) records spaced rec_dt model
    seconds (getDt), record 1 at model time 0, cycling after the last."""
    kind: int                  # STREAM_WIND ("W"), STREAM_CUSTOM ("C"), STREAM_HEATING ("H")
    comp: int                  # COMP_X, COMP_Y, COMP_ETA
    records: np.ndarray        # (nrec, Ny, Nx)
    rec_dt: float
    chunk_size: int = DEF_FORCING_CHUNKSIZE


@dataclass
class Setup:
    """Everything the hot path reads, as the reference holds it after init."""
    Nx: int
    Ny: int
    variant: int
    grids: dict
    dLambda: float
    dTheta: float
    lateral_mixing: bool = False
    quadratic_friction: bool = False
    barotropic: bool = True
    elliptic_solver: bool = False     # built with CALC_LIB_ELLIPTIC_SOLVER = ElSolv_SOR (include/calc_lib.h)
    tau_scale: Optional[float] = None
    G: float = f32(9.80665)
    Ah: float = 0.0
    dt: float = 10.0
    A: float = 6371000.0
    RHO0: float = 1024.0
    minD: float = 1.0
    lbc: float = 2.0
    ab_chi: float = 0.1
    Nt: int = 0
    impl_u: np.ndarray = None
    impl_v: np.ndarray = None
    impl_eta: np.ndarray = None
    gamma_lin_u: np.ndarray = None
    gamma_lin_v: np.ndarray = None
    gamma_lin_eta: np.ndarray = None
    gamma_sq_u: np.ndarray = None
    gamma_sq_v: np.ndarray = None
    F_x_const: np.ndarray = None
    F_y_const: np.ndarray = None
    F_eta_const: np.ndarray = None
    wind: list = field(default_factory=list)        # [(comp, tau)] re-applied every step
    osc: list = field(default_factory=list)         # [(comp, re, im, omega)]
    streams: list = field(default_factory=list)     # [ForcingStream] time-interpolated, after wind and osc
    dyn: list = field(default_factory=list)         # [ForcingStream] DYNFROMFILE datasets (comp: u, v, eta)
    u_bs: np.ndarray = None
    v_bs: np.ndarray = None
    zeta_bs: np.ndarray = None
    u0: np.ndarray = None
    v0: np.ndarray = None
    eta0: np.ndarray = None
    tracers: list = field(default_factory=list)     # [dict(kappa_h, C, C0, gamma_C, cons)]

    @property
    def shape(self):
        return (self.Ny, self.Nx)

    def masks(self):
        return {k: g.ocean for k, g in self.grids.items()}


# ---------------------------------------------------------------------------
# swm_vars / computeD aliasing for the linear variants
# ---------------------------------------------------------------------------

def set_mean_state(s: Setup, psi_bs: np.ndarray, mask_psi: bool = False) -> None:
    """LINEARISED_MEAN_STATE basic state from its streamfunction on the H grid
    (SWM_timestep_init, swm_timestep_module.f90:76-87): u_bs = evSF_zonal,
    v_bs = evSF_meridional (calc_lib.f90:840-880), zeta_bs = laplacian2D
    (calc_lib.f90:1170-1206, the `vorticity` of a scalar).  Evaluated with the
    grid thicknesses as initDomain leaves them, i.e. before
    apply_linear_D_alias, as the reference computes it before computeD.

    mask_psi multiplies the streamfunction by the H-grid ocean mask first, the
    o_H Psi of evSF_zonal's documented formula (calc_lib.f90:832); the code
    itself (calc_lib.f90:840-850, 1176-1205) uses Psi as read, which is the
    default."""
    psi = np.asarray(psi_bs, dtype=np.float64)
    if mask_psi:
        psi = psi * s.grids["H"].ocean
    A, dL, dT = s.A, s.dLambda, s.dTheta
    gH, gu, gv = s.grids["H"], s.grids["u"], s.grids["v"]
    oH = gH.ocean.astype(np.int64)
    ip = lambda a: np.roll(a, -1, axis=1)     # a(ip1(i), j)
    im = lambda a: np.roll(a, 1, axis=1)      # a(im1(i), j)
    jp = lambda a: np.roll(a, -1, axis=0)     # a(i, jp1(j))
    jm = lambda a: np.roll(a, 1, axis=0)      # a(i, jm1(j))
    bc_u, bc_v = gu.bc(s.lbc), gv.bc(s.lbc)
    # pder_meridional3D(psi, H_grid): out grid u, ind0 = jp1, indm1 = jp0
    dth = np.where(jp(oH) + oH == 0, 0.0, bc_u * (jp(psi) - psi) / (A * dT))
    u_bs = -dth
    # pder_zonal3D(psi, H_grid): out grid v, ind0 = ip1, indm1 = ip0
    cvr = gv.cos_lat[:, None]
    dla = np.where(ip(oH) + oH == 0, 0.0, bc_v / (A * cvr * dL) * (ip(psi) - psi))
    v_bs = dla
    if s.barotropic:
        with np.errstate(divide="ignore", invalid="ignore"):
            u_bs = np.where(gu.ocean == 1, u_bs / gu.H, u_bs)
            v_bs = np.where(gv.ocean == 1, v_bs / gv.H, v_bs)
    # laplacian2D(psi, H_grid): lambda2 -> v grid (ip0, im1), theta2 -> u grid (jp0, jm1)
    cH = gH.cos_lat[:, None]
    cHp, cHm = np.roll(gH.cos_lat, -1)[:, None], np.roll(gH.cos_lat, 1)[:, None]
    t1 = 1.0 / (A * cH * dL) ** 2 * (bc_v * (ip(psi) - psi) - im(bc_v) * (psi - im(psi)))
    t2 = 1.0 / (A ** 2 * cH * dT ** 2) * (bc_u * (cHp * jp(psi) - cH * psi) - jm(bc_u) * (cH * psi - cHm * jm(psi)))
    zeta = np.where(gH.ocean == 1, t1 + t2, 0.0)
    s.u_bs, s.v_bs, s.zeta_bs = u_bs, v_bs, zeta


def apply_linear_D_alias(s: Setup) -> None:
    """For the linearised variants D, Dh, Du, Dv are pointers to the grid H
    arrays (swm_vars.f90:93-105) and computeD at itt=0 overwrites H_grid%H,
    u_grid%H and v_grid%H with interpolations of H_eta
    (swm_timestep_module.f90:201-217).  Everything initialised afterwards
    (damping, forcing) sees the overwritten values."""
    if s.variant == FULLY_NONLINEAR:
        return
    He = s.grids["eta"].H
    oe = s.grids["eta"].ocean
    s.grids["H"].H = noland_interp(He, oe, "H")
    s.grids["u"].H = noland_interp(He, oe, "u")
    s.grids["v"].H = noland_interp(He, oe, "v")


# ---------------------------------------------------------------------------
# damping, swm_damping_module.f90
# ---------------------------------------------------------------------------

def sponge_layer(s: Setup, gamma: np.ndarray, gstring: str, pos: str, max_val: float,
                 length: float, cutoff: float) -> np.ndarray:
    """getSpongeLayer, swm_damping_module.f90:257-359 (SPONGE_SCALE_UNIT = degree)."""
    Nx, Ny = s.Nx, s.Ny
    isp = {"U": (1, Ny - 1, 2, Nx - 1), "V": (2, Ny - 1, 1, Nx - 1), "E": (1, Ny - 1, 1, Nx - 1)}
    g = {"U": s.grids["u"], "V": s.grids["v"], "E": s.grids["eta"]}[gstring[0].upper()]
    ib = isp[gstring[0].upper()]
    lat, lon = g.lat, g.lon
    coef = D2R * s.A / length / (D2R * s.A)
    ecut = math.exp(-cutoff)
    gamma = gamma.copy()
    P = pos.upper()
    if "N" in P:
        prof = max_val * (_libm(math.exp, -np.abs(lat - lat[ib[1] - 1]) * coef) - ecut)
        gamma = np.maximum(prof[:, None], gamma)
    if "S" in P:
        prof = max_val * (_libm(math.exp, -np.abs(lat - lat[ib[0] - 1]) * coef) - ecut)
        gamma = np.maximum(prof[:, None], gamma)
    if "W" in P:
        prof = max_val * (_libm(math.exp, -np.abs(lon - lon[ib[2] - 1]) * coef) - ecut)
        gamma = np.maximum(prof[None, :], gamma)
    if "E" in P:
        prof = max_val * (_libm(math.exp, -np.abs(lon - lon[ib[3] - 1]) * coef) - ecut)
        gamma = np.maximum(prof[None, :], gamma)
    return gamma


def _damping_coef(s: Setup, name: str, terms: list, loader) -> np.ndarray:
    """getDampingCoefficient, swm_damping_module.f90:137-203."""
    coef = np.zeros(s.shape)
    gid = {"GAMMA_LIN_U": "U", "GAMMA_LIN_V": "V", "GAMMA_SQ_U": "U", "GAMMA_SQ_V": "V",
           "GAMMA_LIN_ETA": "E"}[name]
    for t in terms:
        if str(t.get("term", "")).upper() != name:
            continue
        typ = t.get("type", "")
        if typ == "sponge":
            coef = sponge_layer(s, coef, gid, t.get("sl_location", ""), float(t.get("value", 0.0)),
                                float(t.get("sl_length", 1.0)), float(t.get("sl_cutoff", 15.0)))
        elif typ == "file":
            coef = np.maximum(loader(t["filename"], t["varname"]), coef)
        elif typ == "uniform":
            coef = np.maximum(coef, float(t.get("value", 0.0)))
        else:
            raise ValueError("Unkown type identifyer in swm_damping_nl. Check your namelists!")
    return coef


def init_damping(s: Setup, terms: list, loader=None) -> None:
    """SWM_damping_init, swm_damping_module.f90:44-117."""
    ou, ov = s.grids["u"].ocean, s.grids["v"].ocean
    H_u, H_v = s.grids["u"].H, s.grids["v"].H
    gu = _damping_coef(s, "GAMMA_LIN_U", terms, loader)
    gv = _damping_coef(s, "GAMMA_LIN_V", terms, loader)
    if s.barotropic:
        with np.errstate(divide="ignore", invalid="ignore"):
            gu = np.where(ou == 1, gu / H_u, gu)
            gv = np.where(ov == 1, gv / H_v, gv)
    if s.quadratic_friction:
        su = _damping_coef(s, "GAMMA_SQ_U", terms, loader)
        sv = _damping_coef(s, "GAMMA_SQ_V", terms, loader)
        if s.barotropic:
            with np.errstate(divide="ignore", invalid="ignore"):
                su = np.where(ou == 1, su / H_u, su)
                sv = np.where(ov == 1, sv / H_v, sv)
        s.gamma_sq_u, s.gamma_sq_v = su, sv
    ge = _damping_coef(s, "GAMMA_LIN_ETA", terms, loader)
    s.gamma_lin_u, s.gamma_lin_v, s.gamma_lin_eta = gu, gv, ge
    s.impl_u = 1.0 + s.dt * gu
    s.impl_v = 1.0 + s.dt * gv
    s.impl_eta = 1.0 + s.dt * ge


# ---------------------------------------------------------------------------
# forcing, swm_forcing_module.f90
# ---------------------------------------------------------------------------

def add_windstress(s: Setup, comp: int, tau: np.ndarray) -> None:
    """A windstress stream of one record (processWindstress,
    swm_forcing_module.f90:274-324).

    FULLY_NONLINEAR makes windstress time dependent (initStream 631-632): it is
    re-applied every step with the current Du/Dv, so it is kept as a stream.
    Otherwise a constant stream is folded once into F_*_const with the
    (overwritten) H_u/H_v, as SWM_forcing_getForcing(isTDF=.FALSE.) does.
    Datasets of several records go through add_forcing_stream.
    """
    for name in ("F_x_const", "F_y_const", "F_eta_const"):
        if getattr(s, name) is None:
            setattr(s, name, np.zeros(s.shape))
    tau = np.asarray(tau, dtype=np.float64)
    if s.variant == FULLY_NONLINEAR:
        if s.osc or s.streams:
            raise NotImplementedError("windstress streams are applied before oscillating and time-interpolated "
                                      "streams; list them first")
        s.wind.append((comp, tau))
        return
    F = s.F_x_const if comp == COMP_X else s.F_y_const
    g = s.grids["u"] if comp == COMP_X else s.grids["v"]
    num = s.tau_scale * tau if s.tau_scale is not None else tau
    den = s.RHO0 * g.H if s.barotropic else s.RHO0
    with np.errstate(divide="ignore", invalid="ignore"):
        F[:] = np.where(g.ocean == 1, F + num / den, F)


def add_oscillation(s: Setup, comp: int, re: np.ndarray, im: np.ndarray, omega: float) -> None:
    """An oscillating forcing stream (processOscillation, 326-374)."""
    for name in ("F_x_const", "F_y_const", "F_eta_const"):
        if getattr(s, name) is None:
            setattr(s, name, np.zeros(s.shape))
    if s.streams:
        raise NotImplementedError("oscillating streams are applied before time-interpolated streams; list them first")
    s.osc.append((comp, np.asarray(re, dtype=np.float64), np.asarray(im, dtype=np.float64), float(omega)))


DEF_NT_CHUNKSIZE = 100          # DFF_Nt_chunksize default, include/io.h:12
STREAM_DYN = 3                  # kind of a DYNFROMFILE dataset


def add_dyn_from_file(s: Setup, comp: int, records: np.ndarray, rec_dt: float,
                      chunk_size: int = DEF_NT_CHUNKSIZE) -> None:
    """A DYNFROMFILE dataset (DFF_initDynFromFile, dynFromFile_module.f90:56-95):
    nrec (Ny, Nx) records of u (COMP_X), v (COMP_Y) or eta (COMP_ETA), spaced
    rec_dt model seconds, added times the grid's ocean mask to the master field
    at every model_advance (DFF_advance, 137-160), i.e. to the flow the tracer
    is advected by and to the U, V, ETA register fields."""
    rec = np.asarray(records, dtype=np.float64)
    if rec.ndim == 2:
        rec = rec[None]
    if rec.shape[1:] != s.shape:
        raise ValueError(f"records must be (nrec, {s.Ny}, {s.Nx})")
    if any(d.comp == comp for d in s.dyn):
        raise ValueError("one dataset per field")
    s.dyn.append(ForcingStream(kind=STREAM_DYN, comp=comp, records=rec, rec_dt=float(rec_dt),
                               chunk_size=int(chunk_size)))


def add_forcing_stream(s: Setup, kind: int, comp: int, records: np.ndarray, rec_dt: float,
                       chunk_size: int = DEF_FORCING_CHUNKSIZE) -> None:
    """A forcing stream on a dataset of several records, interpolated in time
    every step (SWM_forcing_initStream, swm_forcing_module.f90:593-640;
    getChunkData, memchunk_module.f90:214-243).  A dataset of one record is
    constant in time and is folded into F_*_const as the reference does
    (except windstress under FULLY_NONLINEAR, which add_windstress keeps as a
    per-step stream).  Time-dependent streams are processed in the order they
    are added, after the windstress and oscillating streams."""
    for name in ("F_x_const", "F_y_const", "F_eta_const"):
        if getattr(s, name) is None:
            setattr(s, name, np.zeros(s.shape))
    rec = np.asarray(records, dtype=np.float64)
    if rec.ndim == 2:
        rec = rec[None]
    if rec.shape[1:] != s.shape:
        raise ValueError(f"records must be (nrec, {s.Ny}, {s.Nx})")
    if kind == STREAM_WIND and comp == COMP_ETA:
        raise ValueError("windstress has components X (zonal) and Y (meridional) only")
    if kind == STREAM_HEATING and comp != COMP_ETA:
        raise ValueError("heating acts on the continuity equation (COMP_ETA)")
    if rec.shape[0] <= 1:
        if kind == STREAM_WIND:
            add_windstress(s, comp, rec[0])
            return
        F = {COMP_X: s.F_x_const, COMP_Y: s.F_y_const, COMP_ETA: s.F_eta_const}[comp]
        g = s.grids[{COMP_X: "u", COMP_Y: "v", COMP_ETA: "eta"}[comp]]
        F[:] = np.where(g.ocean == 1, F + rec[0], F)
        return
    if chunk_size < 2:
        raise ValueError("a time-dependent dataset needs chunk_size >= 2 (getChunkData reads time(chunkCounter+1))")
    if len(s.streams) >= 4:
        raise ValueError("at most 4 time-interpolated forcing streams")
    s.streams.append(ForcingStream(kind=kind, comp=comp, records=np.ascontiguousarray(rec), rec_dt=float(rec_dt),
                                   chunk_size=int(chunk_size)))


def finalize(s: Setup) -> Setup:
    """Fill defaults for anything the configuration did not set."""
    z = lambda: np.zeros(s.shape)
    for name in ("F_x_const", "F_y_const", "F_eta_const"):
        if getattr(s, name) is None:
            setattr(s, name, z())
    if s.impl_u is None:
        init_damping(s, [])
    if s.quadratic_friction and s.gamma_sq_u is None:
        s.gamma_sq_u, s.gamma_sq_v = z(), z()
    if s.variant == LINEARISED_MEAN_STATE:
        if s.Ny > s.Nx:
            raise ValueError("LINEARISED_MEAN_STATE needs Ny <= Nx: the reference indexes ip1(j) "
                             "(swm_timestep_module.f90:360) and ip1 has Nx entries")
        for name in ("u_bs", "v_bs", "zeta_bs"):
            if getattr(s, name) is None:
                setattr(s, name, z())
    for a in ("u0", "v0", "eta0"):
        if getattr(s, a) is None:
            setattr(s, a, z())
    return s


def add_tracer(s: Setup, C: np.ndarray, kappa_h: float = 0.0, C0=None, gamma_C=None, cons=None) -> None:
    """TRC_add_to_list, tracer_vars.f90:124-185 (fields given directly)."""
    full = lambda v: np.full(s.shape, float(v)) if np.isscalar(v) or v is None else np.asarray(v, np.float64)
    s.tracers.append(dict(kappa_h=float(kappa_h), C=np.asarray(C, np.float64),
                          C0=full(0.0 if C0 is None else C0),
                          gamma_C=full(0.0 if gamma_C is None else gamma_C),
                          cons=full(0.0 if cons is None else cons)))


# ---------------------------------------------------------------------------
# selfcheck experiments (netCDF inputs)
# ---------------------------------------------------------------------------

def _read_nc(path: str, varname: str):
    """readInitialCondition (io_module.f90:198-205): last record, plus the
    missing-value mask of getVar2Dhandle (596-617)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables[varname]
        data = np.array(v.data)
        if data.ndim == 3:
            data = data[-1]
        miss = np.zeros(data.shape, dtype=np.int8)
        for att in ("missing_value", "_FillValue"):
            if hasattr(v, att):
                mv = np.asarray(getattr(v, att)).astype(np.float64)
                miss[data.astype(np.float64) == mv] = 1
        return data.astype(np.float64), miss


_SECONDS_PER = {"second": 1.0, "seconds": 1.0, "s": 1.0, "minute": 60.0, "minutes": 60.0, "hour": 3600.0,
                "hours": 3600.0, "h": 3600.0, "day": 86400.0, "days": 86400.0, "d": 86400.0}


def _read_records(path: str, varname: str):
    """Every record of a forcing dataset and its record spacing in model
    seconds, dt = (t_last - t_first)/(nrec - 1) (getDt,
    memchunk_module.f90:252-273; the calendar offset cancels in the
    difference, only the unit's length matters)."""
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables[varname]
        data = np.array(v.data, dtype=np.float64)
        if data.ndim == 2:
            return data[None], 0.0
        if data.shape[0] <= 1:
            return data, 0.0
        tname = v.dimensions[0]
        t = np.array(nc.variables[tname].data, dtype=np.float64)
        units = nc.variables[tname].units
        units = units.decode() if isinstance(units, bytes) else str(units)
        scale = _SECONDS_PER[units.split()[0].lower()]
    return data, (t[-1] * scale - t[0] * scale) / float(np.float32(data.shape[0] - 1.0))


def setup_from_selfcheck(exp_dir: str, defines: dict) -> Setup:
    """Build a Setup from a reference selfcheck experiment directory
    (model.namelist + input/*.nc) and its model.h defines."""
    import os
    nml = _nml.read_namelist_file(os.path.join(exp_dir, "model.namelist"))
    g1 = lambda name: (_nml.groups(nml, name) or [{}])[0]
    model_nl, domain_nl = g1("model_nl"), g1("domain_nl")
    path = lambda p: os.path.join(exp_dir, p)
    Nx, Ny = int(domain_nl.get("nx", 100)), int(domain_nl.get("ny", 100))
    H_in = H_miss = None
    if domain_nl.get("in_file_h"):
        H_in, H_miss = _read_nc(path(domain_nl["in_file_h"]), domain_nl.get("in_varname_h", "H"))
    variant = [VARIANT_NAMES[k] for k in VARIANT_NAMES if k in defines][0]
    if "lbc" not in domain_nl:
        # domain_module.f90:53 shadows the module default with an uninitialised
        # local; its value only reaches the bc factors, which a linear model
        # without lateral mixing never reads.
        if "LATERAL_MIXING" in defines or variant != LINEARISED_STATE_OF_REST:
            raise ValueError("domain_nl has no lbc: the reference reads an uninitialised local "
                             "(domain_module.f90:53); give lbc explicitly")
        lbc = 0.0
    else:
        lbc = float(domain_nl["lbc"])
    OMEGA = float(domain_nl.get("omega", f32(7.272205e-5)))
    A = float(domain_nl.get("a", 6371000.0))
    grids, dL, dT = init_domain(
        Nx, Ny, float(domain_nl.get("lon_s", -20.0)), float(domain_nl.get("lon_e", 20.0)),
        float(domain_nl.get("lat_s", -20.0)), float(domain_nl.get("lat_e", 20.0)),
        H_in, H_miss, float(domain_nl.get("h_overwrite", 0.0)), lbc,
        int(domain_nl.get("coriolis_approx", 3)), float(domain_nl.get("theta0", 0.0)), OMEGA, A)
    ts = defines.get("TAU_SCALE")
    s = Setup(Nx=Nx, Ny=Ny, variant=variant, grids=grids, dLambda=dL, dTheta=dT,
              lateral_mixing="LATERAL_MIXING" in defines,
              quadratic_friction="QUADRATIC_BOTTOM_FRICTION" in defines,
              barotropic="BAROTROPIC" in defines,
              tau_scale=None if ts is None else f32(float(ts)),
              G=float(model_nl.get("g", f32(9.80665))), Ah=float(model_nl.get("ah", 0.0)),
              dt=float(model_nl.get("dt", 10.0)), A=A, RHO0=float(domain_nl.get("rho0", 1024.0)),
              lbc=lbc, ab_chi=float(g1("tint_nl").get("ab_chi", 0.1)))
    s.Nt = int(float(model_nl.get("run_length", 100000.0)) / s.dt)
    if variant == LINEARISED_MEAN_STATE:
        ts_nl = g1("swm_timestep_nl")
        psi_bs, _ = _read_records(path(ts_nl["filename"]), ts_nl["varname"])
        set_mean_state(s, psi_bs[0])      # getChunkData(SWM_MC_bs_psi, 0): the first record
    apply_linear_D_alias(s)
    loader = lambda fn, vn: _read_nc(path(fn), vn)[0]
    init_damping(s, _nml.groups(nml, "swm_damping_nl"), loader)
    for st in _nml.groups(nml, "swm_forcing_nl"):
        if not st.get("filename"):
            continue
        comp = {"Z": COMP_X, "M": COMP_Y, "C": COMP_ETA}[str(st.get("component", "Z"))[0].upper()]
        ftype = str(st.get("forcingtype", ""))[0].upper()
        if ftype == "O":
            re, _ = _read_nc(path(st["filename"]), st["varname"])
            im, _ = _read_nc(path(st["filename2"]), st["varname2"])
            add_oscillation(s, comp, re, im, float(st.get("omega", 0.0)))
        elif ftype in ("W", "C", "H"):
            rec, rec_dt = _read_records(path(st["filename"]), st["varname"])
            kind = {"W": STREAM_WIND, "C": STREAM_CUSTOM, "H": STREAM_HEATING}[ftype]
            if ftype == "H":
                comp = COMP_ETA
            add_forcing_stream(s, kind, comp, rec, rec_dt, int(st.get("chunksize", DEF_FORCING_CHUNKSIZE)))
        else:
            raise NotImplementedError(f"forcing type {ftype}")
    if model_nl.get("file_eta_init"):
        s.eta0, _ = _read_nc(path(model_nl["file_eta_init"]), model_nl.get("varname_eta_init", "ETA"))
    return finalize(s)
