"""The benchmark configurations at their stated shapes, through the C-ABI.

C1: the linear non-rotating 200 x 200 basin, dt = 60 s, 2000 steps, against
the oracle (AB3 is unstable at this dt for the basin's fastest gravity wave,
so the run overflows: the Inf/NaN patterns must agree exactly), and the same
basin at dt = 25 s, inside AB3's stability region, finite and bit identical.
C3: red-black SOR on the 2048 x 1024 elliptic basin (omega = 1.9): FP64 sweeps
bit identical to the oracle's at a fixed sweep count, FP32 sweeps within a
stated bar, and the reference mode (Chebyshev omega, absolute RMS stopping
rule) with its sweep count.  The C4 and C5 shapes are covered in
test_gpu_batch.py and test_gpu_serve.py.
"""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu

pytestmark = pytest.mark.gpu

STATE = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_U", "G_u"), ("G_V", "G_v"), ("G_ETA", "G_eta")]


@pytest.mark.parametrize("dt,finite", [(60.0, False), (25.0, True)])
def test_c1_linear_200_full_run(gpu_lib, dt, finite):
    s = configs.c1_linear(seed=0, N=200, dt=dt, nsteps=2000)
    assert s.Nx == s.Ny == 200
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    for n in (1, 2, 3, 250, 2000):
        g.step(n - g.itt)
        o.step(n - o.itt)
        for gn, on in STATE:
            assert same(g.get(gn), o.get(on)), (dt, n, gn)
    eta = g.get("SWM_ETA")
    assert bool(np.isfinite(eta).all()) == finite
    g.close()
    o.close()


@pytest.fixture(scope="module")
def c3():
    s = configs.c3_elliptic(seed=0, nsteps=4)
    assert (s.Nx, s.Ny) == (2048, 1024)
    o = oracle.Oracle(s)
    g = gpu.SwmGpu(s).init_from_setup()
    g.step(5)
    o.step(5)
    assert same(g.get("SWM_U"), o.get("u"))
    B = o.divergence(o.get("u"), o.get("v"))
    yield s, o, g, B
    g.close()
    o.close()


@pytest.mark.parametrize("every", [1, 2, 4])
def test_c3_sor_fixed_sweeps_bit_exact(gpu_lib, c3, every):
    """omega = 1.9, 48 sweeps, tolerance 0 (so both run every sweep): FP64 chi
    bit identical to the oracle with one-, two- and four-sweep blocks."""
    s, o, g, B = c3
    chi_o, it_o, rr_o, _ = o.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, 48)
    g.set_sor_check_every(every)
    chi_g, it_g, rr_g, ok = g.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, 48)
    g.set_sor_check_every(1)
    assert it_o == it_g == 48 and not ok
    assert same(chi_g, chi_o)
    assert abs(rr_g - rr_o) <= 1e-12 * rr_o


def test_c3_sor_fp32_sweeps_error_bar(gpu_lib, c3):
    """FP32 sweeps (FP32 chi and B, FP64 residual norm) after 48 sweeps from
    zero stay within 1e-4 relative max-norm of the FP64 oracle's chi (the
    FP32 contract's bar on eta), and the
    FP64 residual they report is within 1e-3 of the oracle's."""
    s, o, g, B = c3
    chi_o, _, rr_o, _ = o.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, 48)
    g.set_sor_check_every(2)
    chi_f, it_f, rr_f, _ = g.sor_fixed(B, np.zeros(s.shape), 1.9, 0.0, 48, fp32_sweeps=True)
    g.set_sor_check_every(1)
    assert it_f == 48
    oc = s.grids["eta"].ocean == 1
    err = np.abs(chi_f[oc] - chi_o[oc]).max() / np.abs(chi_o[oc]).max()
    assert err <= 1e-4, err
    assert abs(rr_f - rr_o) <= 1e-3 * rr_o


def test_c3_sor_reference_mode(gpu_lib, c3):
    """main_ElSolv_SOR on the 2048 x 1024 grid (Chebyshev omega, stop when
    sqrt(sum res^2) < EPS = 1e-4): the right-hand side is scaled so that the
    first sweep's correction norm lies just above EPS and the solve stops
    after a few hundred sweeps (432 with this state); same sweep count and chi
    as the oracle."""
    s, o, g, B = c3
    Bs = B * (4e-14 / np.abs(B).max())
    chi_o, it_o, an_o, rc_o = o.sor(Bs, np.zeros(s.shape), epsilon=1e-4, first_guess=True)
    chi_g, it_g, an_g, ok = g.sor(Bs, np.zeros(s.shape), epsilon=1e-4, first_guess=True)
    assert rc_o == 0 and ok and 2 <= it_o <= 1000
    assert it_g == it_o
    assert same(chi_g, chi_o)
    assert abs(an_g - an_o) <= 1e-12 * an_o
