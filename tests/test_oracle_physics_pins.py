"""Pins of the oracle's FULLY_NONLINEAR, lateral-mixing and tracer terms that
follow from the mathematics, chosen so that a plausible slip (a sign, a
transposed coefficient, a shifted index, a swapped face velocity) fails one
of them:

* lateral mixing (swm_lateralmixing_module.f90:162-240, coefficients 66-146)
  on a flat, all-ocean channel with constant D: the tension/shear form reduces
  to Ah times the 5-point Laplacian of each velocity component, whose discrete
  symbol is exact: for a mode sin(k i) along a row the result is
  -Ah 4 sin^2(k/2) / dx^2 times the mode, for either component and direction;
  a rigid translation has no mixing at all;
* on the sphere, a solid-body rotation u = U cos(theta) has no strain, so its
  mixing tendency vanishes as the grid is refined, at second order: the tan
  terms of the coefficients cancel the curvature of cos(theta);
* on a spherical band symmetric about the equator with mirrored coasts, a
  flow with u even and v odd in latitude has a mixing trend of the same
  parities: the metric terms and their coastal mask weights mirror;
* tracer (tracer_module.f90:154-209, 291-342): diffusion of a mode with no
  flow and advection of a mode by a uniform flow have exact discrete symbols
  (-kappa h 4 sin^2(k/2)/dx^2 and -U h sin(k)/dx times the shifted mode); a
  tracer of constant concentration stays constant under any flow (its fluxes
  are the continuity equation's mass fluxes times C); the area-weighted
  inventory sum(CH cos) is conserved with advection and diffusion;
* FULLY_NONLINEAR against LINEARISED_STATE_OF_REST (itself pinned by the
  recorded selfcheck runs): started from an eta bump of amplitude eps the two
  differ by O(eps^2), so halving eps divides the difference by 4; any slip in
  the nonlinear model's linear part (Coriolis through Pot*MV, the pressure
  gradient through EDens, the mass flux, the mixing) leaves an O(eps)
  difference instead.
"""
import math

import numpy as np
import pytest

import oracle
from swm_amd import configs, host


def channel(Nx=64, Ny=48, dx=5e3, H0=100.0, dt=5.0, **kw):
    """Flat all-ocean channel, periodic in x, land rows at j = 1, Ny."""
    kw.setdefault("variant", host.FULLY_NONLINEAR)
    s = configs.basin(Nx, Ny, dx, dx, H0=H0, dt=dt, side_walls=False, **kw)
    return s


def interior_rows(s, margin):
    m = np.zeros(s.shape, dtype=bool)
    m[margin:s.Ny - margin, :] = True
    return m


def latmix_of(s, u0=None, v0=None):
    s.u0 = np.zeros(s.shape) if u0 is None else u0
    s.v0 = np.zeros(s.shape) if v0 is None else v0
    s.eta0 = np.zeros(s.shape)
    o = oracle.Oracle(host.finalize(s))
    o.step(1)                     # latmix is computed from the state before the step
    out = o.get("latmix_u"), o.get("latmix_v")
    o.close()
    return out


@pytest.mark.parametrize("comp,direction,m", [("u", "x", 3), ("u", "y", 2), ("v", "x", 5), ("v", "y", 3)])
def test_mixing_symbol_of_a_mode(comp, direction, m):
    Ah, dx = 300.0, 5e3
    s = channel(Nx=64, Ny=48, dx=dx, Ah=Ah, lateral_mixing=True, lbc=0.0)
    i = np.arange(s.Nx)[None, :]
    j = np.arange(s.Ny)[:, None]
    if direction == "x":
        k = 2.0 * math.pi * m / s.Nx
        mode = np.sin(k * i) + 0.0 * j
    else:
        k = 2.0 * math.pi * m / 40.0
        mode = np.cos(k * (j - 17)) + 0.0 * i
    ux = mode * 0.7
    lu, lv = latmix_of(s, u0=ux if comp == "u" else None, v0=ux if comp == "v" else None)
    got = lu if comp == "u" else lv
    other = lv if comp == "u" else lu
    dlen = s.A * s.dLambda if direction == "x" else s.A * s.dTheta
    want = -Ah * 4.0 * math.sin(k / 2.0) ** 2 / dlen ** 2 * ux
    ok = interior_rows(s, 3)
    assert np.abs(got[ok] - want[ok]).max() <= 1e-9 * np.abs(want[ok]).max()
    assert np.abs(other[ok]).max() <= 1e-9 * np.abs(want[ok]).max()


def test_mixing_of_a_rigid_translation_vanishes():
    s = channel(Ah=300.0, lateral_mixing=True, lbc=0.0)
    lu, lv = latmix_of(s, u0=np.full(s.shape, 0.4), v0=np.full(s.shape, -0.3))
    ok = interior_rows(s, 3)
    scale = 300.0 * 0.4 / (s.A * s.dLambda) ** 2
    assert np.abs(lu[ok]).max() <= 1e-12 * scale and np.abs(lv[ok]).max() <= 1e-12 * scale


def test_mixing_of_solid_body_rotation_converges_at_second_order():
    """u = U cos(theta) on a 40-degree spherical band: |latmix_u| in the
    interior falls by ~4 when the grid spacing halves (metric terms consistent
    to second order)."""
    err = []
    for n in (32, 64, 128):
        grids, dL, dT = host.init_domain(n, n, 0.0, 40.0, -20.0, 20.0, H_in=np.full((n, n), 500.0), lbc=0.0,
                                         coriolis_approx=0)
        s = host.Setup(Nx=n, Ny=n, variant=host.FULLY_NONLINEAR, grids=grids, dLambda=dL, dTheta=dT,
                       lateral_mixing=True, barotropic=True, Ah=1e4, dt=10.0, lbc=0.0)
        host.init_damping(s, [])
        cu = grids["u"].cos_lat[:, None]
        lu, lv = latmix_of(s, u0=np.broadcast_to(2.0 * cu, s.shape).copy())
        lat_u = grids["u"].lat
        band = (np.abs(lat_u) < 10.0)[:, None] & np.ones((1, n), dtype=bool)
        err.append(np.abs(lu[band]).max())
    # each halving of the spacing divides the residual by about four
    assert 3.5 < err[0] / err[1] < 4.5 and 3.5 < err[1] / err[2] < 4.5, err


def tracer_channel(C, u0=None, v0=None, eta0=None, kappa=0.0, **kw):
    s = channel(**kw)
    host.add_tracer(s, C, kappa_h=kappa)
    s.u0 = np.zeros(s.shape) if u0 is None else u0
    s.v0 = np.zeros(s.shape) if v0 is None else v0
    s.eta0 = np.zeros(s.shape) if eta0 is None else eta0
    return host.finalize(s)


def test_tracer_diffusion_symbol():
    kappa, H0 = 50.0, 100.0
    s0 = channel()
    k = 2.0 * math.pi * 4 / s0.Nx
    C = np.sin(k * np.arange(s0.Nx))[None, :] * np.ones((s0.Ny, 1))
    s = tracer_channel(C, kappa=kappa, H0=H0)
    o = oracle.Oracle(s)
    o.step(1)
    G = o.get("G_CH", 0)
    o.close()
    want = -kappa * H0 * 4.0 * math.sin(k / 2.0) ** 2 / (s.A * s.dLambda) ** 2 * C
    ok = interior_rows(s, 2)
    assert np.abs(G[ok] - want[ok]).max() <= 1e-10 * np.abs(want[ok]).max()


@pytest.mark.parametrize("direction", ["x", "y"])
def test_tracer_advection_symbol(direction):
    H0, U = 100.0, 0.3
    s0 = channel()
    i = np.arange(s0.Nx)[None, :] + 0.0 * np.arange(s0.Ny)[:, None]
    j = np.arange(s0.Ny)[:, None] + 0.0 * np.arange(s0.Nx)[None, :]
    if direction == "x":
        k = 2.0 * math.pi * 5 / s0.Nx
        C = np.sin(k * i)
        s = tracer_channel(C, u0=np.full(s0.shape, U), H0=H0)
        want = -U * H0 * math.sin(k) / (s.A * s.dLambda) * np.cos(k * i)
    else:
        k = 2.0 * math.pi * 3 / 40.0
        C = np.sin(k * j)
        s = tracer_channel(C, v0=np.full(s0.shape, U), H0=H0)
        want = -U * H0 * math.sin(k) / (s.A * s.dTheta) * np.cos(k * j)
    o = oracle.Oracle(s)
    o.step(1)
    G = o.get("G_CH", 0)
    o.close()
    ok = interior_rows(s, 2)
    assert np.abs(G[ok] - want[ok]).max() <= 1e-10 * np.abs(want[ok]).max()


def flow_channel(C, kappa, seed=4, Nx=64, Ny=128):
    """A channel with a compact eta bump and a compact tracer patch in its
    middle rows, beta-plane, mixing, so every field is exactly zero within
    some rows of the land rows for the first steps."""
    s0 = channel(Nx=Nx, Ny=Ny)
    rng = np.random.default_rng(seed)
    eta = np.zeros(s0.shape)
    eta[56:72, 10:40] = 0.2 * rng.random((16, 30))
    return tracer_channel(C, eta0=eta, kappa=kappa, Nx=Nx, Ny=Ny, Ah=200.0, lateral_mixing=True, lbc=0.0,
                          f0=1e-4, beta=2e-11, H0=100.0, dt=5.0)


def test_constant_concentration_stays_constant():
    """C = 1 under the flow of an adjusting eta bump: CH/h stays 1 wherever it
    is finite (the land rows' 0/0 coefficients spread NaN one row per step)."""
    s = flow_channel(np.ones((128, 64)), kappa=40.0)
    o = oracle.Oracle(s)
    o.step(30)
    C = o.get("C", 0)
    u = o.get("u")
    o.close()
    fin = np.isfinite(C) & (s.grids["eta"].ocean == 1)
    assert fin.sum() > 0.4 * C.size
    assert np.abs(u).max() > 1e-3                  # the flow is not trivial
    assert np.abs(C[fin] - 1.0).max() <= 1e-12


def test_tracer_inventory_conserved():
    """sum(CH cos) over the middle rows, where every flux through their edges
    is still exactly zero, is constant with advection and diffusion."""
    rng = np.random.default_rng(9)
    C = np.zeros((128, 64))
    C[58:70, 5:50] = rng.random((12, 45))
    s = flow_channel(C, kappa=40.0)
    o = oracle.Oracle(s)
    rows = slice(30, 98)
    ce = s.grids["eta"].cos_lat[rows, None]
    inv0 = (o.get("CH", 0)[rows] * ce).sum()
    o.step(20)
    CH = o.get("CH", 0)[rows]
    o.close()
    assert np.isfinite(CH).all()
    assert np.abs(CH[:4]).max() == 0.0 and np.abs(CH[-4:]).max() == 0.0   # edges untouched
    assert abs((CH * ce).sum() - inv0) <= 1e-13 * np.abs(CH * ce).sum()


def fn_minus_linear(eps):
    out = {}
    for var in (host.FULLY_NONLINEAR, host.LINEARISED_STATE_OF_REST):
        s = configs.basin(48, 40, 10e3, 10e3, variant=var, H0=100.0, dt=20.0, f0=1e-4, beta=2e-11, Ah=200.0,
                          lateral_mixing=True, lbc=2.0, land_fraction=0.06, seed=3)
        x = (np.arange(48) - 22.0) * 10e3
        y = (np.arange(40) - 19.0) * 10e3
        s.eta0 = eps * np.exp(-(x[None, :] ** 2 + y[:, None] ** 2) / (2 * (60e3) ** 2))
        o = oracle.Oracle(host.finalize(s))
        o.step(60)
        # ocean points only: the FULLY_NONLINEAR clamp sets land eta to minD - 0
        oc = [s.grids[g].ocean == 1 for g in ("eta", "u", "v")]
        out[var] = [o.get(n)[m] for n, m in zip(("eta", "u", "v"), oc)]
        o.close()
    return [np.abs(a - b).max() for a, b in zip(out[host.FULLY_NONLINEAR], out[host.LINEARISED_STATE_OF_REST])], \
        [np.abs(b).max() for b in out[host.LINEARISED_STATE_OF_REST]]


def test_nonlinear_minus_linear_is_second_order_in_amplitude():
    d1, a1 = fn_minus_linear(0.02)
    d2, a2 = fn_minus_linear(0.01)
    for k in range(3):
        assert 3.7 < d1[k] / d2[k] < 4.3, (k, d1, d2)
        assert 1.9 < a1[k] / a2[k] < 2.1               # the linear run is linear
        assert d1[k] < 0.05 * a1[k]                     # and the nonlinear part is small


@pytest.mark.parametrize("direction", ["zonal", "meridional"])
def test_nonrotating_shear_flow_is_steady(direction):
    """f = 0, no mixing or forcing: a zonal flow u(y) (or a meridional flow
    v(x)) on a flat layer with eta = 0 is an exact steady state of the
    nonlinear shallow-water equations, because the vortex-force term
    zeta/D * (D u) balances the gradient of the kinetic energy u^2/2.  In the
    discrete vector-invariant form the two cancel term by term
    (interp2(Pot, H2v) * interp4(MU, u2v) against the difference of EDens),
    so G_u, G_v, G_eta vanish to rounding: this fixes the sign of the Pot*MU
    and Pot*MV terms, the u^2 and v^2 in EDens and their factor 1/2."""
    s = channel(Nx=64, Ny=48, dx=5e3, H0=100.0, dt=5.0)
    assert all(np.all(g.f == 0.0) for g in s.grids.values())
    j = np.arange(s.Ny)[:, None] + 0.0 * np.arange(s.Nx)[None, :]
    i = np.arange(s.Nx)[None, :] + 0.0 * np.arange(s.Ny)[:, None]
    if direction == "zonal":
        s.u0 = 0.8 * np.sin(math.pi * (j - 1) / (s.Ny - 3)) ** 2
        s.v0 = np.zeros(s.shape)
        scale = 0.8 ** 2 / (s.A * s.dTheta)
    else:
        s.u0 = np.zeros(s.shape)
        s.v0 = 0.8 * np.sin(2 * math.pi * 3 * i / s.Nx) + 0.3 * np.cos(2 * math.pi * i / s.Nx)
        scale = 0.8 ** 2 / (s.A * s.dLambda)
    s.eta0 = np.zeros(s.shape)
    o = oracle.Oracle(host.finalize(s))
    o.step(1)
    ok = interior_rows(s, 3)
    for name in ("G_u", "G_v", "G_eta"):
        assert np.abs(o.get(name)[ok]).max() <= 1e-12 * scale, name
    o.close()


@pytest.mark.parametrize("lbc", [0.0, 2.0])
def test_mixing_operator_symmetric_and_dissipative(lbc):
    """On the flat metric, with islands and free- or no-slip coasts, the
    mixing trend is the divergence of a stress built from the strain of the
    same flow, so as an operator L on (u, v) it is symmetric and negative in
    the energy inner product <a, b> = sum(D cos (a_u b_u + a_v b_v)) over
    ocean u and v points: <a, L b> = <L a, b> and <a, L a> < 0."""
    rng = np.random.default_rng(11)

    def apply(u, v):
        s = configs.basin(60, 50, 5e3, 5e3, variant=host.FULLY_NONLINEAR, H0=500.0, dt=10.0, Ah=300.0,
                          lateral_mixing=True, lbc=lbc, land_fraction=0.08, seed=5)
        lu, lv = latmix_of(s, u0=u, v0=v)
        return lu, lv, s

    a = (rng.standard_normal((50, 60)), rng.standard_normal((50, 60)))
    b = (rng.standard_normal((50, 60)), rng.standard_normal((50, 60)))
    La, Lb = apply(*a), apply(*b)
    s = La[2]
    gu, gv = s.grids["u"], s.grids["v"]
    wu = gu.ocean * gu.H * gu.cos_lat[:, None]
    wv = gv.ocean * gv.H * gv.cos_lat[:, None]
    dot = lambda x, Ly: (wu * x[0] * gu.ocean * Ly[0]).sum() + (wv * x[1] * gv.ocean * Ly[1]).sum()
    ab, ba, aa = dot(a, Lb), dot(b, La), dot(a, La)
    assert abs(ab - ba) <= 1e-12 * abs(aa)
    assert aa < 0.0 and dot(b, Lb) < 0.0


@pytest.mark.parametrize("lbc", [0.0, 2.0])
def test_mixing_mirror_symmetric_about_the_equator(lbc):
    """On a spherical band symmetric about the equator, with land (coasts,
    islands) mirrored onto itself, a flow with u even and v odd in latitude
    has a mixing trend with the same parities: the metric (tan) terms and the
    ocean masks that weight them at a southern coast must mirror those at the
    northern one.  H and v rows j mirror onto Ny-1-j, eta and u rows onto
    Ny-2-j (the last eta row is land)."""
    n, ny = 40, 36
    rng = np.random.default_rng(3)
    Hin = np.full((ny, n), 500.0)
    Hin[:, 0] = 0.0
    Hin[:, -1] = 0.0
    for _ in range(8):                                  # islands of 3x3 land corners
        i, j = rng.integers(2, n - 5), rng.integers(2, ny - 5)
        Hin[j:j + 3, i:i + 3] = 0.0
    Hin = np.where(np.arange(ny)[:, None] < ny // 2, Hin, Hin[::-1])
    grids, dL, dT = host.init_domain(n, ny, 0.0, 30.0, -30.0, 30.0, H_in=Hin, lbc=lbc, coriolis_approx=0)
    s = host.Setup(Nx=n, Ny=ny, variant=host.FULLY_NONLINEAR, grids=grids, dLambda=dL, dTheta=dT,
                   lateral_mixing=True, barotropic=True, Ah=1e4, dt=10.0, lbc=lbc)
    host.init_damping(s, [])
    e, v_ = slice(0, ny - 1), slice(1, ny - 1)
    assert np.abs(grids["eta"].lat[e] + grids["eta"].lat[e][::-1]).max() < 1e-9
    assert np.array_equal(grids["v"].ocean[v_], grids["v"].ocean[v_][::-1])
    # the band's first and last v rows are land: coasts along both edges, and islands inside
    assert not grids["v"].ocean[0].any() and not grids["v"].ocean[ny - 1].any()
    assert (grids["eta"].ocean[1:ny - 2, 1:n - 1] == 0).any()
    u = rng.standard_normal((ny, n))
    u[e] = 0.5 * (u[e] + u[e][::-1])
    v = rng.standard_normal((ny, n))
    v[v_] = 0.5 * (v[v_] - v[v_][::-1])
    v[0] = v[ny - 1] = 0.0
    lu, lv = latmix_of(s, u0=u, v0=v)
    scale = max(np.abs(lu).max(), np.abs(lv).max())
    assert np.abs(lu[e] - lu[e][::-1]).max() <= 1e-13 * scale
    assert np.abs(lv[v_] + lv[v_][::-1]).max() <= 1e-13 * scale


@pytest.mark.parametrize("lbc", [0.0, 2.0])
def test_nondivergent_flow_divergence_is_the_solver_residual(lbc):
    """computeNonDivergentFlowField: div(u - grad chi) = div(u) - L chi, with L
    the SOR operator (ElSolv_SOR.f90:54-62), because the divergence of the
    gradient is that 5-point operator; on every ocean cell with free slip, and
    on the cells with ocean faces with no slip (whose land faces carry the bc
    factor 2 into the gradient).  A converged chi leaves a small divergence."""
    s = configs.basin(40, 30, 10e3, 10e3, variant=host.FULLY_NONLINEAR, H0=500.0, dt=10.0, seed=3,
                      land_fraction=0.08, lbc=lbc)
    o = oracle.Oracle(host.finalize(s))
    rng = np.random.default_rng(0)
    u = rng.standard_normal(s.shape) * s.grids["u"].ocean
    v = rng.standard_normal(s.shape) * s.grids["v"].ocean
    B = o.divergence(u, v)
    chi = rng.standard_normal(s.shape)
    un, vn = o.nondivergent(u, v, chi)
    d = o.divergence(un, vn)
    Lc = o.sor_apply(chi)
    ou, ov, oe = s.grids["u"].ocean == 1, s.grids["v"].ocean == 1, s.grids["eta"].ocean == 1
    cells = oe if lbc == 0.0 else oe & ou & np.roll(ou, -1, axis=1) & ov & np.roll(ov, -1, axis=0)
    assert cells.sum() > 800
    assert np.abs(d - (B - Lc))[cells].max() <= 1e-12 * np.abs(Lc).max()
    chi_c, it, rr, rc = o.sor_fixed(B, np.zeros(s.shape), 1.8, 1e-12, 20000)
    assert rc == 0
    dn = o.divergence(*o.nondivergent(u, v, chi_c))
    assert np.abs(dn[cells]).max() <= 1e-9 * np.abs(B[cells]).max()
