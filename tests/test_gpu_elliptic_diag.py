"""Diagnostic variables of the elliptic-solver build on the device
(diagVar CHI, U_ND, V_ND, PSI; src/diagVar.f90:115-134): computeVelocityPotential
(main_ElSolv_SOR of the divergence with EPS and first_guess = itt == 1, chi kept
between steps), computeNonDivergentFlowField and the streamfunction of the
non-divergent flow, bit for bit against the oracle in FP64, at the first step
and later steps (warm start), once per step however often they are asked."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu

pytestmark = pytest.mark.gpu

EPS = float(np.float32(1e-4))   # calc_lib.h's EPS, a default-kind (single precision) literal


def setup(N=96, lbc=2.0):
    s = configs.c2_nonlinear(seed=8, N=N, dt=10.0, tracer=False)
    s.elliptic_solver = True
    s.lbc = lbc
    return s


@pytest.mark.parametrize("lbc", [2.0, 0.0])
def test_velocity_potential_and_nondivergent_flow(gpu_lib, lbc):
    s = setup(lbc=lbc)
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    chi_o = np.zeros(s.shape)
    for itt in (1, 2, 7):
        g.step(itt - g.itt)
        o.step(itt - o.itt)
        u, v = o.get("u"), o.get("v")
        B = o.divergence(u, v)
        chi_o, it_o, an_o, rc_o = o.sor(B, chi_o, epsilon=EPS, first_guess=(itt == 1))
        un, vn = o.nondivergent(u, v, chi_o)
        psi_o = o.psi(un, vn)
        assert same(g.get("PSI"), psi_o), itt          # the first request of the step solves
        it_g, an_g = g.elliptic_stats()
        assert it_g == it_o and abs(an_g - an_o) <= 1e-12 * an_o, (itt, it_g, it_o)
        assert same(g.get("CHI"), chi_o), itt
        assert same(g.get("U_ND"), un) and same(g.get("V_ND"), vn), itt
        assert g.elliptic_stats()[0] == it_g           # no second solve within the step
    g.close()
    o.close()


def test_without_the_solver_u_nd_is_u(gpu_lib):
    s = setup()
    s.elliptic_solver = False
    g = gpu.SwmGpu(s).init_from_setup()
    g.step(3)
    assert same(g.get("U_ND"), g.get("SWM_U")) and same(g.get("V_ND"), g.get("SWM_V"))
    g.close()


def test_fp32_nondivergent_flow_close_to_fp64(gpu_lib):
    s = setup()
    a = gpu.SwmGpu(s, gpu.FP64).init_from_setup()
    b = gpu.SwmGpu(s, gpu.FP32).init_from_setup()
    a.step(5)
    b.step(5)
    for name in ("U_ND", "V_ND", "PSI"):
        x, y = a.get(name), b.get(name)
        assert np.abs(x - y).max() <= 1e-3 * np.abs(x).max(), name
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_solve_is_the_same_at_every_sweeps_per_launch(gpu_lib, precision):
    """The elliptic build's per-step solve (reference mode, test after every
    sweep) with four, two and one sweep(s) per launch (halo 8, 4, 2: launches
    of several sweeps keep and test each earlier sweep's chi and residual) and
    one colour per launch (halo 0): the same sweep count and the same CHI bit
    for bit at every step, in the context's precision (FP32 sweeps for an FP32
    context)."""
    s = setup()
    runs = {}
    for halo in (8, 4, 2, 0):
        g = gpu.SwmGpu(s, precision).init_from_setup()
        g.set_sor_check_every(16)        # fixed-omega cadence that the 8-cell halo's four sweeps divide
        g.set_sor_halo(halo)
        runs[halo] = []
        for itt in (1, 2, 7):
            g.step(itt - g.itt)
            chi = g.get("CHI")
            runs[halo].append((g.elliptic_stats()[0], chi))
        g.close()
    for halo in (4, 2, 0):
        for (it_a, chi_a), (it_b, chi_b) in zip(runs[8], runs[halo]):
            assert it_a == it_b, (halo, it_a, it_b)
            assert same(chi_a, chi_b), halo
