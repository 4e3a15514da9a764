"""Synthetic workloads for the SWM time-stepping path.

Cartesian basins are realised on the reference's spherical metric
(src/domain_module.f90:97-127, src/grid_module.f90:116-145) with a very large
radius, A = 1e15 m, so that every |theta| < 1e-8 rad and cos(theta) rounds to
exactly 1.0 in FP64.  A beta plane f = f0 + beta*y uses coriolis_approx = 2 with
OMEGA = beta*A/(2 cos theta0) and tan(theta0) = f0/(beta*A).  Linear bottom drag
r is given, as swm_damping_nl "uniform" with value r*H (BAROTROPIC divides by
H_u, src/swm_damping_module.f90:78-81).

All randomness is drawn from numpy's PCG64 with the given seed.
"""
from __future__ import annotations

import math

import numpy as np

from . import host as H

A_FLAT = 1.0e15


def _extent_deg(n: int, d: float, A: float) -> float:
    """Degrees spanned by n grid points spaced d metres (dLambda = D2R*range/(n-1))."""
    return (n - 1) * d / (A * H.D2R)


def _smooth_noise(rng, Ny, Nx, sigma):
    from scipy.ndimage import gaussian_filter
    return gaussian_filter(rng.standard_normal((Ny, Nx)), sigma=sigma, mode="wrap")


def basin(Nx, Ny, dx, dy, *, variant, H0, dt, G=9.81, Ah=0.0, lbc=2.0, f0=0.0, beta=0.0,
          drag=0.0, lateral_mixing=False, land_fraction=0.0, seed=0, A=A_FLAT,
          quadratic_friction=False, gamma_sq=0.0, psi_bs=None, side_walls=True):
    """A closed flat-bottom basin of depth H0 with optional seeded islands.
    LINEARISED_MEAN_STATE takes its basic state from the streamfunction psi_bs
    (zero if None).  side_walls=False leaves the zonal boundary open: a channel
    periodic in x between the land rows that initDomain puts at j = 1, Ny."""
    rng = np.random.default_rng(seed)
    lon_e = _extent_deg(Nx, dx, A)
    lat_e = _extent_deg(Ny, dy, A)
    Hin = np.full((Ny, Nx), float(H0))
    if side_walls:
        Hin[:, 0] = 0.0
        Hin[:, -1] = 0.0
    if land_fraction > 0.0:
        noise = _smooth_noise(rng, Ny, Nx, sigma=max(2.0, min(Nx, Ny) / 128.0))
        thr = np.quantile(noise, 1.0 - land_fraction)
        Hin[noise > thr] = 0.0
    if beta != 0.0:
        coriolis, theta0 = 2, math.atan(f0 / (beta * A)) / H.D2R
        OMEGA = beta * A / (2.0 * math.cos(theta0 * H.D2R))
    elif f0 != 0.0:
        coriolis, theta0, OMEGA = 1, 90.0, f0 / 2.0
    else:
        coriolis, theta0, OMEGA = 0, 0.0, 0.0
    grids, dL, dT = H.init_domain(Nx, Ny, 0.0, lon_e, 0.0, lat_e, H_in=Hin, lbc=lbc,
                                  coriolis_approx=coriolis, theta0=theta0, OMEGA=OMEGA, A=A)
    s = H.Setup(Nx=Nx, Ny=Ny, variant=variant, grids=grids, dLambda=dL, dTheta=dT,
                lateral_mixing=lateral_mixing, quadratic_friction=quadratic_friction, barotropic=True,
                G=G, Ah=Ah, dt=dt, A=A, RHO0=1024.0, lbc=lbc)
    if variant == H.LINEARISED_MEAN_STATE:
        H.set_mean_state(s, np.zeros((Ny, Nx)) if psi_bs is None else psi_bs)
    H.apply_linear_D_alias(s)
    terms = []
    if drag:
        terms += [dict(term="GAMMA_LIN_U", type="uniform", value=drag * H0),
                  dict(term="GAMMA_LIN_V", type="uniform", value=drag * H0)]
    if quadratic_friction and gamma_sq:
        terms += [dict(term="GAMMA_SQ_U", type="uniform", value=gamma_sq),
                  dict(term="GAMMA_SQ_V", type="uniform", value=gamma_sq)]
    H.init_damping(s, terms)
    s._rng = rng
    return s


def double_gyre_tau(s, tau0):
    """tau_x = tau0 cos(2 pi y / Ly) on the u grid (row-uniform)."""
    lat = s.grids["u"].lat
    Ly = s.A * (s.grids["v"].lat[-1] - s.grids["v"].lat[0]) * H.D2R
    y = s.A * (lat - s.grids["v"].lat[0]) * H.D2R
    row = tau0 * np.cos(2.0 * math.pi * y / Ly)
    return np.repeat(row[:, None], s.Nx, axis=1)


def fourier_modes(s, rng, n_modes=16, amp=1.0, kmax=8):
    """Sum of n seeded Fourier modes amp*cos(2 pi (kx x + ky y) + phase) on the
    eta grid, evaluated separably as cos(a)cos(b) - sin(a)sin(b)."""
    x = np.arange(s.Nx) / s.Nx
    y = np.arange(s.Ny) / s.Ny
    C = np.zeros((s.Ny, s.Nx))
    for _ in range(n_modes):
        kx, ky = rng.integers(1, kmax + 1, size=2)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        ax = 2.0 * math.pi * kx * x + ph
        by = 2.0 * math.pi * ky * y
        C += amp * (np.outer(np.cos(by), np.cos(ax)) - np.outer(np.sin(by), np.sin(ax)))
    return C


def c1_linear(seed=0, N=200, dx=10e3, dt=60.0, nsteps=2000):
    """Linear non-rotating basin (LINEARISED_STATE_OF_REST): H=1000 m, g=9.81,
    f=0, Gaussian bump 0.1 m (sigma 100 km, centre jittered +-200 km) plus
    N(0, 1e-4 m) noise, u=v=0, no forcing or mixing."""
    s = basin(N, N, dx, dx, variant=H.LINEARISED_STATE_OF_REST, H0=1000.0, dt=dt, G=9.81, seed=seed)
    rng = s._rng
    x = (np.arange(N) + 0.5) * dx
    y = (np.arange(N) + 0.5) * dx
    cx = x.mean() + rng.uniform(-200e3, 200e3)
    cy = y.mean() + rng.uniform(-200e3, 200e3)
    r2 = (x[None, :] - cx) ** 2 + (y[:, None] - cy) ** 2
    s.eta0 = 0.1 * np.exp(-r2 / (2.0 * 100e3 ** 2)) + rng.normal(0.0, 1e-4, size=(N, N))
    s.Nt = nsteps
    return H.finalize(s)


def c2_nonlinear(seed=0, N=1024, dx=5e3, dt=30.0, nsteps=10000, tracer=True, Ah=100.0, tau0=0.1):
    """Nonlinear beta-plane basin (FULLY_NONLINEAR + LATERAL_MIXING): f0=1e-4,
    beta=2e-11, H=500 m, Ah, linear drag 1e-6 1/s, double-gyre windstress,
    ~5% seeded islands, one passive tracer (16 Fourier modes, kappa=50)."""
    s = basin(N, N, dx, dx, variant=H.FULLY_NONLINEAR, H0=500.0, dt=dt, G=9.81, Ah=Ah, lbc=2.0,
              f0=1e-4, beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.05, seed=seed)
    H.add_windstress(s, H.COMP_X, double_gyre_tau(s, tau0))
    if tracer:
        H.add_tracer(s, fourier_modes(s, s._rng), kappa_h=50.0)
    s.Nt = nsteps
    return H.finalize(s)


def c3_elliptic(seed=0, Nx=2048, Ny=1024, dx=10e3, dt=30.0, nsteps=2000, Ah=200.0, tau0=0.1):
    """Elliptic-path basin: nonlinear beta-plane double gyre on 2048 x 1024 whose
    velocity potential (computeVelocityPotential) is solved every step."""
    s = basin(Nx, Ny, dx, dx, variant=H.FULLY_NONLINEAR, H0=500.0, dt=dt, G=9.81, Ah=Ah, lbc=2.0,
              f0=1e-4, beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.0, seed=seed)
    H.add_windstress(s, H.COMP_X, double_gyre_tau(s, tau0))
    s.Nt = nsteps
    return H.finalize(s)


ENSEMBLE_SIZES = (128, 256, 384, 512, 768, 1024)


def stable_dt(H0: float, dx: float, G: float = 9.81, cfl: float = 0.2) -> float:
    """Time step with gravity-wave Courant number c*dt/dx <= cfl (c = sqrt(G*H0)),
    rounded down to whole seconds: AB3 on the C-grid stays stable there."""
    return float(max(1.0, math.floor(cfl * dx / math.sqrt(G * H0))))


def ensemble_member(seed: int, Nx: int, Ny: int, H0: float, Ah: float, tau0: float, tracer: bool,
                    dx: float = 5e3, nsteps: int = 5000):
    """One nonlinear beta-plane basin of the batched ensemble: the C2 physics with
    the member's depth, viscosity and wind amplitude, dt from stable_dt."""
    s = basin(Nx, Ny, dx, dx, variant=H.FULLY_NONLINEAR, H0=H0, dt=stable_dt(H0, dx), G=9.81, Ah=Ah, lbc=2.0,
              f0=1e-4, beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.05, seed=seed)
    H.add_windstress(s, H.COMP_X, double_gyre_tau(s, tau0))
    if tracer:
        H.add_tracer(s, fourier_modes(s, s._rng), kappa_h=50.0)
    s.Nt = nsteps
    return H.finalize(s)


def ensemble_params(seed: int = 0, n: int = 64):
    """Parameters of the batched ensemble: nx, ny each from ENSEMBLE_SIZES,
    H ~ U(200, 4000) m, Ah ~ U(20, 500) m^2/s, wind amplitude ~ U(0.02, 0.2) N/m^2,
    a tracer with probability 1/2."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        out.append(dict(seed=seed * 100003 + k, Nx=int(rng.choice(ENSEMBLE_SIZES)), Ny=int(rng.choice(ENSEMBLE_SIZES)),
                        H0=float(rng.uniform(200.0, 4000.0)), Ah=float(rng.uniform(20.0, 500.0)),
                        tau0=float(rng.uniform(0.02, 0.2)), tracer=bool(rng.random() < 0.5)))
    return out


def c4_ensemble(seed: int = 0, n: int = 64, nsteps: int = 5000):
    """The batched ensemble: n independent members of uneven size."""
    return [ensemble_member(nsteps=nsteps, **p) for p in ensemble_params(seed, n)]


def serve_member(seed: int, Nx: int, Ny: int, H0: float, Ah: float, tau0: float, tracer: bool,
                 dx: float = 5e3, nsteps: int = 1000):
    """One request of the serving stream: a closed nonlinear beta-plane basin
    (walls, no islands) with the ensemble's physics and parameters."""
    s = basin(Nx, Ny, dx, dx, variant=H.FULLY_NONLINEAR, H0=H0, dt=stable_dt(H0, dx), G=9.81, Ah=Ah, lbc=2.0,
              f0=1e-4, beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.0, seed=seed)
    H.add_windstress(s, H.COMP_X, double_gyre_tau(s, tau0))
    if tracer:
        H.add_tracer(s, fourier_modes(s, s._rng), kappa_h=50.0)
    s.Nt = nsteps
    return H.finalize(s)


def mean_state_basin(seed=0, Nx=96, Ny=64, dx=10e3, dt=20.0, H0=100.0, U=0.1, lateral_mixing=True, Ah=200.0,
                     land_fraction=0.05):
    """LINEARISED_MEAN_STATE basin: a perturbation on a steady basic flow of
    speed ~U from a smooth seeded streamfunction (zero on land), f-plane plus
    beta, linear drag, double-gyre windstress, a Gaussian eta bump."""
    rng = np.random.default_rng(seed)
    stub = basin(Nx, Ny, dx, dx, variant=H.LINEARISED_STATE_OF_REST, H0=H0, dt=dt, seed=seed,
                 land_fraction=land_fraction)
    psi = fourier_modes(stub, rng, n_modes=6, kmax=3)
    psi = psi * (U * H0 * dx * 4.0 / np.abs(psi).max()) * stub.grids["H"].ocean
    s = basin(Nx, Ny, dx, dx, variant=H.LINEARISED_MEAN_STATE, H0=H0, dt=dt, f0=1e-4, beta=2e-11, drag=1e-6,
              lateral_mixing=lateral_mixing, Ah=Ah, lbc=0.0, land_fraction=land_fraction, seed=seed, psi_bs=psi)
    H.add_windstress(s, H.COMP_X, double_gyre_tau(s, 0.05))
    x = (np.arange(Nx) - Nx / 2) * dx
    y = (np.arange(Ny) - Ny / 2) * dx
    s.eta0 = 0.1 * np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (2 * (8 * dx) ** 2))
    return H.finalize(s)
