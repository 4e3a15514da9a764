"""Blocked fixed-omega SOR (several red-black sweeps per launch on LDS tiles,
used when the convergence test runs every 2 or 4 sweeps or a multiple): the
same chi, bit for bit, as the oracle's sequential sweeps in FP64 and as the
one-colour-per-launch sweeps in FP32, for sweep counts that leave a remainder,
grids that no tile divides, a grid narrower than the tile's halo, and odd
periodic extents (301 columns, 75 rows), where blocks clear of the wrap visit
only their colour's cells and blocks across it test each cell's colour."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu


def basin(nx, ny, land, seed):
    s = configs.basin(nx, ny, 10e3, 10e3, variant=host.FULLY_NONLINEAR, H0=500.0, dt=10.0, seed=seed,
                      land_fraction=land)
    return host.finalize(s)


def rhs(o, s, seed):
    rng = np.random.default_rng(seed)
    return o.divergence(rng.standard_normal(s.shape) * s.grids["u"].ocean,
                        rng.standard_normal(s.shape) * s.grids["v"].ocean)


@pytest.mark.parametrize("nx,ny,every,sweeps", [(96, 64, 4, 40), (130, 70, 4, 42), (130, 70, 2, 41),
                                                (7, 5, 4, 12), (200, 33, 16, 50), (301, 40, 2, 31),
                                                (96, 75, 4, 40)])
def test_blocked_sweeps_equal_sequential_sweeps(gpu_lib, nx, ny, every, sweeps):
    s = basin(nx, ny, 0.08 if nx > 20 else 0.0, seed=nx + ny)
    o = oracle.Oracle(s)
    B = rhs(o, s, 3)
    chi0 = 1e3 * np.random.default_rng(5).standard_normal(s.shape)
    chi_o, it_o, _, rc_o = o.sor_fixed(B, chi0, 1.7, 0.0, sweeps)     # tol 0: exactly `sweeps` sweeps
    g = gpu.SwmGpu(s)
    g.set_sor_check_every(every)
    chi_g, it_g, _, ok = g.sor_fixed(B, chi0, 1.7, 0.0, sweeps)
    assert it_o == it_g == sweeps and not ok and rc_o == 1
    assert same(chi_g, chi_o)
    g.close()
    o.close()


def test_blocked_fp32_sweeps_equal_unblocked(gpu_lib, monkeypatch):
    s = basin(150, 90, 0.05, seed=9)
    o = oracle.Oracle(s)
    B = rhs(o, s, 4)
    o.close()
    out = {}
    for blocked in ("1", "0"):
        monkeypatch.setenv("SWM_GPU_SOR_BLOCK", blocked)
        g = gpu.SwmGpu(s)
        g.set_sor_check_every(4)
        out[blocked] = g.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, 36, fp32_sweeps=True)
        g.close()
    assert out["1"][1] == out["0"][1] == 36
    assert same(out["1"][0], out["0"][0])


def test_blocked_converges_like_unblocked(gpu_lib, monkeypatch):
    """With a tolerance the blocked run stops at a test sweep like the
    unblocked run with the same cadence (the residual sums differ only in
    summation order)."""
    s = basin(96, 64, 0.05, seed=2)
    o = oracle.Oracle(s)
    B = rhs(o, s, 6)
    o.close()
    res = {}
    for blocked in ("1", "0"):
        monkeypatch.setenv("SWM_GPU_SOR_BLOCK", blocked)
        g = gpu.SwmGpu(s)
        g.set_sor_check_every(8)
        res[blocked] = g.sor_fixed(B, np.zeros(s.shape), 1.9, 1e-9, 5000)
        g.close()
    chi_b, it_b, rr_b, ok_b = res["1"]
    chi_u, it_u, rr_u, ok_u = res["0"]
    assert ok_b and ok_u and it_b % 8 == 0 and abs(it_b - it_u) <= 8 and rr_b < 1e-9


@pytest.mark.parametrize("every,stop", [(2, 20), (2, 22), (4, 24), (4, 28), (2, 32), (4, 96)])
def test_blocked_stops_at_chosen_test_sweep(gpu_lib, every, stop):
    """Converging blocked runs with K = 2 and 4 sweeps per launch whose
    tolerance is passed first at sweep `stop`: launch n does sweeps
    nK+1..(n+1)K, so stop/K - 1 is odd for 20, 28, 32, 96 (the result is in
    the caller's buffer) and even for 22, 24 (the result is copied back from
    the second buffer); 32 and 96 end a host batch, whose test runs in the
    one-block check instead of the next launch.  The tolerance is the geometric
    mean of the oracle's relative residuals at sweeps stop - every and stop,
    and every earlier test sweep's residual lies above it, so the run must stop
    at `stop` with the oracle's chi of that many sweeps, bit for bit."""
    s = basin(96, 64, 0.05, seed=2)
    o = oracle.Oracle(s)
    B = rhs(o, s, 6)
    rr = {k: o.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, k)[2] for k in range(every, stop + 1, every)}
    tol = float(np.sqrt(rr[stop] * rr[stop - every]))
    assert rr[stop] < tol and all(rr[k] > tol for k in rr if k < stop)
    chi_o = o.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, stop)[0]
    o.close()
    g = gpu.SwmGpu(s)
    g.set_sor_check_every(every)
    chi_g, it_g, rr_g, ok = g.sor_fixed(B, np.zeros(s.shape), 1.9, tol, 5000)
    assert ok and it_g == stop
    assert same(chi_g, chi_o)
    assert abs(rr_g - rr[stop]) <= 1e-12 * rr[stop]
    g.close()


@pytest.mark.parametrize("blocked", ["1", "0"])
def test_reference_mode_one_sweep_blocks(gpu_lib, monkeypatch, blocked):
    """main_ElSolv_SOR (Chebyshev omega per half-sweep, test after every sweep)
    with blocked launches on LDS tiles (two sweeps per launch by default), and
    with one colour per launch: the oracle's chi bit for bit and its sweep
    count, on a grid of several tiles with an odd row count."""
    monkeypatch.setenv("SWM_GPU_SOR_BLOCK", blocked)
    s = basin(130, 75, 0.06, seed=13)
    o = oracle.Oracle(s)
    B = rhs(o, s, 8)
    chi_o, it_o, an_o, rc_o = o.sor(B, np.zeros(s.shape), epsilon=1e-4, first_guess=True)
    g = gpu.SwmGpu(s)
    chi_g, it_g, an_g, ok = g.sor(B, np.zeros(s.shape), epsilon=1e-4, first_guess=True)
    assert ok and rc_o == 0 and it_g == it_o
    assert same(chi_g, chi_o)
    assert abs(an_g - an_o) <= 1e-12 * an_o
    g.close()
    o.close()


# Tolerances whose first passed test in the oracle's reference-mode solve of
# the problem below is the given sweep; each lies inside the interval of
# tolerances that stop there with a margin of at least 0.3 % on either side,
# far above the rounding of a differently ordered sum of res^2.
REF_STOPS = {2270: 0.014168205745143653, 2271: 0.01402977285267049, 2272: 0.013896860975592949,
             2300: 0.01105465311498026, 2301: 0.010981280786951706, 2302: 0.010906213987353636}


@pytest.mark.parametrize("halo", [-1, 8, 4, 2, 0])
def test_reference_mode_stops_at_any_sweep_of_a_launch(gpu_lib, halo):
    """main_ElSolv_SOR with four sweeps per launch (halo -1 or 8) or two (halo
    4) — each launch also keeps the chi and residual partials of its earlier
    sweeps, and the next launch's block 0 tests the sweeps in order — with one
    sweep per launch (halo 2) and one colour per launch (halo 0).  Host
    batches end at sweep 2272, so 2270..2272 lie in the last launch of a batch
    (tested by the one-block checks that close it) and 2300..2302 inside one
    (tested by the next launch); with four sweeps per launch they are the 2nd,
    3rd, 4th and 4th, 1st, 2nd sweep of their launch.  For each, the oracle's
    sweep count and chi, bit for bit, and its norm to rounding."""
    s = basin(130, 75, 0.06, seed=13)
    o = oracle.Oracle(s)
    B = rhs(o, s, 8)
    g = gpu.SwmGpu(s)
    g.set_sor_check_every(16)   # fixed-omega cadence that the 8-cell halo's four sweeps divide
    g.set_sor_halo(halo)
    for k, eps in REF_STOPS.items():
        chi_o, it_o, an_o, rc_o = o.sor(B, np.zeros(s.shape), epsilon=eps, first_guess=True)
        assert rc_o == 0 and it_o == k
        chi_g, it_g, an_g, ok = g.sor(B, np.zeros(s.shape), epsilon=eps, first_guess=True)
        assert ok and it_g == k, (halo, k, it_g)
        assert same(chi_g, chi_o), (halo, k)
        assert abs(an_g - an_o) <= 1e-12 * an_o
    g.close()
    o.close()
