"""Pins of the oracle that follow from the mathematics rather than from a
recorded run: the AB3 start-up ramp, flux-form mass conservation, properties of
the SOR operator and its solution, and the propagation law of the tracer's
0/0 coefficients."""
import numpy as np

import oracle
from swm_amd import configs, host


def rest_basin(**kw):
    s = configs.basin(24, 18, 10e3, 10e3, variant=host.LINEARISED_STATE_OF_REST, H0=100.0, dt=30.0, seed=1, **kw)
    return s


def test_ab3_ramp_closed_form():
    """With a spatially uniform constant F_eta and no flow, eta evolves with the
    constant tendency only.  integrate_AB (time_integration_module.f90:93-102)
    uses slots 1..min(3, itt), oldest first, with the new tendency in slot 3:
    step 1 adds nothing, step 2 adds (1.5 + ab_chi) dt F, later steps dt F."""
    s = rest_basin()
    F = 3.0e-6
    s.F_eta_const = np.full(s.shape, F)
    host.finalize(s)
    o = oracle.Oracle(s)
    oe = s.grids["eta"].ocean == 1
    expect = {1: 0.0, 2: (1.5 + s.ab_chi) * s.dt * F}
    for n in range(3, 12):
        expect[n] = expect[n - 1] + s.dt * F
    for n in range(1, 12):
        o.step(1)
        eta = o.get("eta")
        assert np.allclose(eta[oe], expect[n], rtol=1e-13, atol=0.0), n
        assert np.all(eta[~oe] == 0.0)
        assert np.all(o.get("u") == 0.0)


def test_mass_conservation_linear_basin():
    """Flux-form continuity with MU = MV = 0 on land: sum(eta * cos) over ocean
    points is constant (cos = 1 on this flat metric) up to rounding."""
    s = configs.c1_linear(seed=4, N=60, dt=20.0)
    o = oracle.Oracle(s)
    oe = s.grids["eta"].ocean == 1
    m0 = o.get("eta")[oe].sum()
    scale = np.abs(o.get("eta")[oe]).sum()
    o.step(300)
    eta = o.get("eta")
    assert np.isfinite(eta).all()
    assert abs(eta[oe].sum() - m0) <= 1e-12 * scale


def test_mass_conservation_nonlinear_until_clamp():
    s = configs.c2_nonlinear(seed=2, N=64, dt=10.0, tracer=False)
    o = oracle.Oracle(s)
    oe = s.grids["eta"].ocean == 1
    o.step(200)
    eta = o.get("eta")
    assert np.all(eta[oe] > (1.0 - s.grids["eta"].H[oe]))   # clamp never active
    assert abs(eta[oe].sum()) <= 1e-12 * np.abs(eta[oe]).sum() * oe.sum()


def sor_setup(land=0.05, Nx=20, Ny=14, seed=3):
    s = configs.basin(Nx, Ny, 10e3, 10e3, variant=host.FULLY_NONLINEAR, H0=500.0, dt=10.0, seed=seed,
                      land_fraction=land)
    return host.finalize(s)


def test_sor_operator_null_space_and_symmetry():
    """The elliptic operator of ElSolv_SOR.f90:54-61 annihilates constants on
    every ocean cell (c5 = -(c1+c2+c3+c4)) and is self-adjoint with the cos(lat)
    area weights.  A transposed or shifted coefficient breaks one of these."""
    s = sor_setup()
    o = oracle.Oracle(s)
    oe = s.grids["eta"].ocean == 1
    ce = s.grids["eta"].cos_lat[:, None]
    one = o.sor_apply(np.ones(s.shape))
    assert np.abs(one[oe]).max() <= 1e-12 * np.abs(o.sor_apply(np.random.default_rng(0).random(s.shape))).max()
    rng = np.random.default_rng(7)
    x = rng.standard_normal(s.shape) * oe
    y = rng.standard_normal(s.shape) * oe
    a = (ce * x * o.sor_apply(y))[oe].sum()
    b = (ce * y * o.sor_apply(x))[oe].sum()
    assert abs(a - b) <= 1e-12 * (abs(a) + abs(b))


def test_sor_interior_stencil_is_the_textbook_laplacian():
    """On interior ocean cells of the flat metric the stencil is
    (x(i+1)-2x+x(i-1))/dx^2 + (x(j+1)-2x+x(j-1))/dy^2."""
    s = sor_setup(land=0.0)
    o = oracle.Oracle(s)
    x = np.random.default_rng(5).standard_normal(s.shape)
    Lx = o.sor_apply(x)
    dx = s.A * s.dLambda
    dy = s.A * s.dTheta
    lap = ((np.roll(x, -1, 1) - 2 * x + np.roll(x, 1, 1)) / dx ** 2 +
           (np.roll(x, -1, 0) - 2 * x + np.roll(x, 1, 0)) / dy ** 2)
    inner = np.zeros(s.shape, bool)
    inner[2:-2, 2:-2] = True
    assert np.allclose(Lx[inner], lap[inner], rtol=1e-12, atol=1e-12 * np.abs(lap).max())


def test_sor_converges_to_dense_solution():
    """Fixed-omega SOR reaches the least-squares solution of L chi = B obtained
    by a dense solve of the probed operator (up to the constant null mode)."""
    s = sor_setup()
    o = oracle.Oracle(s)
    oe = s.grids["eta"].ocean == 1
    idx = np.flatnonzero(oe.ravel())
    n = idx.size
    L = np.zeros((n, n))
    for k, p in enumerate(idx):
        e = np.zeros(s.Nx * s.Ny)
        e[p] = 1.0
        L[:, k] = o.sor_apply(e.reshape(s.shape)).ravel()[idx]
    rng = np.random.default_rng(2)
    B = o.divergence(rng.standard_normal(s.shape) * s.grids["u"].ocean,
                     rng.standard_normal(s.shape) * s.grids["v"].ocean)
    chi, it, rel, rc = o.sor_fixed(B, np.zeros(s.shape), 1.7, 1e-13, 20000)
    assert rc == 0
    ref = np.linalg.lstsq(L, B.ravel()[idx], rcond=None)[0]
    got = chi.ravel()[idx]
    got, ref = got - got.mean(), ref - ref.mean()
    assert np.abs(got - ref).max() <= 1e-9 * np.abs(ref).max()


def test_sor_reference_omega_schedule():
    """Chebyshev acceleration of ElSolv_SOR.f90:206-210: the converged solution
    does not depend on the schedule, so the reference mode (rho_J from a single
    precision literal) and the fixed-omega mode agree once both converge."""
    s = sor_setup(land=0.0, Nx=12, Ny=10)
    o = oracle.Oracle(s)
    rng = np.random.default_rng(3)
    B = o.divergence(rng.standard_normal(s.shape) * s.grids["u"].ocean,
                     rng.standard_normal(s.shape) * s.grids["v"].ocean)
    chi_r, it_r, an, rc = o.sor(B, np.zeros(s.shape), epsilon=1e-16, first_guess=True)
    chi_f, it_f, rel, rc2 = o.sor_fixed(B, np.zeros(s.shape), 1.5, 1e-14, 50000)
    oe = s.grids["eta"].ocean == 1
    a, b = chi_r[oe] - chi_r[oe].mean(), chi_f[oe] - chi_f[oe].mean()
    assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_tracer_nan_spread_law():
    """TRC_initCoeffs divides by ocean_eta(nb) + ocean_eta(i,j), which is 0 at a
    land eta cell with a land neighbour (tracer_module.f90:318-333).  The NaN
    reaches CH at step 2 (the first AB step that uses a tendency) and then
    spreads by one cell (4-neighbourhood, periodic) per step through the
    advection and diffusion sums, even with the flow at rest."""
    s = configs.basin(30, 22, 5e3, 5e3, variant=host.FULLY_NONLINEAR, H0=100.0, dt=5.0, seed=6,
                      land_fraction=0.08)
    host.add_tracer(s, np.ones(s.shape), kappa_h=10.0)
    host.finalize(s)
    o = oracle.Oracle(s)
    land = s.grids["eta"].ocean == 0
    nb = lambda m: m | np.roll(m, 1, 0) | np.roll(m, -1, 0) | np.roll(m, 1, 1) | np.roll(m, -1, 1)
    seed = land & (np.roll(land, 1, 0) | np.roll(land, -1, 0) | np.roll(land, 1, 1) | np.roll(land, -1, 1))
    o.step(1)
    assert not np.isnan(o.get("CH", 0)).any()
    expect = seed.copy()
    for n in range(2, 9):
        o.step(1)
        assert np.array_equal(np.isnan(o.get("CH", 0)), expect), n
        expect = nb(expect)
