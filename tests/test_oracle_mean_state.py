"""LINEARISED_MEAN_STATE in the oracle (swm_timestep_module.f90:230-310,
349-443), pinned by special cases that reduce its terms to closed forms: a
zero basic state is the state-of-rest model bit for bit; with u = v = 0 and a
uniform zonal basic flow the only tendency is the basic-state eta flux, whose
first interpolation sits at column j+1 of row j as the reference indexes it
(ip1(j)); the vorticity-basic-flow products in G_u and G_v; mass
conservation under a meridional basic flow (a flux divergence); and
Pot * Dh - f = zeta_bs."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, host

STATE = ("u", "v", "eta", "G_u", "G_v", "G_eta")


def plain_basin(variant, seed=2, nx=48, ny=40, land=0.0, mixing=False):
    s = configs.basin(nx, ny, 10e3, 10e3, variant=variant, H0=100.0, dt=20.0, f0=1e-4, beta=2e-11, seed=seed,
                      land_fraction=land, lateral_mixing=mixing, Ah=300.0, lbc=0.0)
    return s


def test_zero_basic_state_is_state_of_rest():
    runs = {}
    for var in (host.LINEARISED_STATE_OF_REST, host.LINEARISED_MEAN_STATE):
        s = plain_basin(var, land=0.05, mixing=True)
        host.add_windstress(s, host.COMP_X, configs.double_gyre_tau(s, 0.05))
        s.eta0 = 0.05 * configs.fourier_modes(s, np.random.default_rng(1), n_modes=3)
        o = oracle.Oracle(host.finalize(s))
        o.step(60)
        runs[var] = {n: o.get(n) for n in STATE}
        o.close()
    for n in STATE:
        assert same(runs[host.LINEARISED_STATE_OF_REST][n], runs[host.LINEARISED_MEAN_STATE][n]), n


def all_ocean_interior(s):
    """Ocean eta points whose stencil neighbours are ocean too."""
    oe = s.grids["eta"].ocean == 1
    ok = oe.copy()
    for ax in (0, 1):
        for sh in (1, -1):
            ok &= np.roll(oe, sh, axis=ax)
    return ok


def test_eta_flux_of_a_uniform_zonal_basic_flow():
    """u = v = 0, u_bs = U0, v_bs = zeta_bs = 0: G_eta = -(E1(j) - E0(i,j)) U0 / dLambda / (A cos),
    E0 = eta on the u point (i,j), E1(j) = eta on the u point of column j+1 in row j."""
    s = plain_basin(host.LINEARISED_MEAN_STATE, nx=48, ny=40)
    U0 = 0.5
    s.u_bs = np.full(s.shape, U0)
    s.v_bs = np.zeros(s.shape)
    s.zeta_bs = np.zeros(s.shape)
    rng = np.random.default_rng(5)
    s.eta0 = 0.1 * rng.standard_normal(s.shape)
    s = host.finalize(s)
    o = oracle.Oracle(s)
    eta = o.get("eta")
    o.step(1)
    Ge = o.get("G_eta")
    o.close()
    Ny, Nx = s.shape
    E0 = 0.5 * eta + 0.5 * np.roll(eta, 1, axis=1)             # (eta(i,j) + eta(i-1,j)) / 2
    col = (np.arange(Ny) + 1) % Nx
    E1 = 0.5 * eta[np.arange(Ny), col] + 0.5 * eta[np.arange(Ny), col - 1]
    want = -(E1[:, None] * U0 - E0 * U0) / s.dLambda / (s.A * s.grids["eta"].cos_lat[:, None])
    ok = all_ocean_interior(s)
    ok &= (np.roll(ok, -1, axis=1)) & np.array([[True]] * Ny)
    rows_ok = np.array([bool(ok[j, col[j]] and ok[j, col[j] - 1]) for j in range(Ny)])
    m = ok & rows_ok[:, None]
    assert m.sum() > 100
    assert np.abs(Ge[m] - want[m]).max() <= 1e-12 * np.abs(want[m]).max()
    # the same flux with ip1(i) in place of ip1(j) differs: the index is the reference's
    E1i = 0.5 * np.roll(eta, -1, axis=1) + 0.5 * eta
    wrong = -(E1i * U0 - E0 * U0) / s.dLambda / (s.A * s.grids["eta"].cos_lat[:, None])
    assert np.abs(Ge[m] - wrong[m]).max() > 1e-6 * np.abs(want[m]).max()


@pytest.mark.parametrize("component", ["u", "v"])
def test_vorticity_times_basic_flow(component):
    """G_u = interp2(zeta, H2u) * V0 for u random, v = eta = 0, v_bs = V0;
    G_v = -interp2(zeta, H2v) * U0 for v random, u = eta = 0, u_bs = U0
    (no mixing, no forcing, so nothing else contributes)."""
    s = plain_basin(host.LINEARISED_MEAN_STATE, nx=40, ny=36)
    U0 = 0.5
    rng = np.random.default_rng(7)
    s.u_bs, s.v_bs, s.zeta_bs = np.zeros(s.shape), np.zeros(s.shape), np.zeros(s.shape)
    if component == "u":
        s.v_bs = np.full(s.shape, U0)
        s.u0 = rng.standard_normal(s.shape)
    else:
        s.u_bs = np.full(s.shape, U0)
        s.v0 = rng.standard_normal(s.shape)
    s = host.finalize(s)
    o = oracle.Oracle(s)
    zeta = o.get("zeta")
    o.step(1)
    ok = all_ocean_interior(s)
    if component == "u":
        G = o.get("G_u")
        want = (0.5 * np.roll(zeta, -1, axis=0) + 0.5 * zeta) * U0
        m = ok & (s.grids["u"].ocean == 1)
    else:
        G = o.get("G_v")
        want = -(0.5 * np.roll(zeta, -1, axis=1) + 0.5 * zeta) * U0
        m = ok & (s.grids["v"].ocean == 1)
    o.close()
    assert np.abs(want[m]).max() > 0
    assert np.abs(G[m] - want[m]).max() <= 1e-14 * np.abs(want[m]).max()


def test_meridional_basic_flow_conserves_mass():
    """u_bs = 0: the basic-state eta flux is a meridional flux divergence, so
    sum(eta * cos) stays constant (closed basin, 300 steps)."""
    s = plain_basin(host.LINEARISED_MEAN_STATE, nx=40, ny=36, land=0.05)
    rng = np.random.default_rng(9)
    s.u_bs = np.zeros(s.shape)
    s.v_bs = 0.3 * rng.standard_normal(s.shape) * s.grids["v"].ocean
    s.zeta_bs = 1e-5 * rng.standard_normal(s.shape)
    s.eta0 = 0.1 * configs.fourier_modes(s, rng, n_modes=4)
    s = host.finalize(s)
    o = oracle.Oracle(s)
    ce = s.grids["eta"].cos_lat[:, None]
    m0 = np.sum(o.get("eta") * ce)
    scale = np.sum(np.abs(o.get("eta")) * ce)
    o.step(300)
    e = o.get("eta")
    o.close()
    assert np.isfinite(e).all() and np.abs(e).max() > 1e-3
    assert abs(np.sum(e * ce) - m0) <= 1e-12 * scale


def test_potential_vorticity_uses_basic_state_vorticity():
    s = configs.mean_state_basin(seed=3, Nx=48, Ny=40)
    o = oracle.Oracle(s)
    pot, Dh = o.get("Pot"), o.get("Dh")
    o.close()
    f = s.grids["H"].f[:, None]
    m = Dh > 1.0
    assert np.abs(pot[m] * Dh[m] - f[np.nonzero(m)[0], 0] - s.zeta_bs[m]).max() <= 1e-12 * np.abs(s.zeta_bs).max()


def test_rows_beyond_nx_are_rejected():
    s = plain_basin(host.LINEARISED_MEAN_STATE, nx=30, ny=40)
    with pytest.raises(ValueError):
        host.finalize(s)
