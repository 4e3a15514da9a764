"""Checks at the workloads' full sizes.  The 1024 x 1024 basin runs its full
10 000 steps (at dt = 18 s, the largest stable whole-second step; the stated
30 s overflows) and keeps the flux-form volume sum(eta * cos) to rounding in
FP64 and within the FP32 bar; the 2048 x 1024 elliptic-workload basin and an
8192 x 4096 basin (33.5 M cells) are bit-identical to the
oracle for a few steps."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

STATE = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_U", "G_u"), ("G_V", "G_v"), ("G_ETA", "G_eta")]


def volume_scale(g, s):
    eta = g.get("SWM_ETA")
    ce = s.grids["eta"].cos_lat[:, None]
    return float(np.sum(np.abs(eta) * ce * (s.grids["eta"].ocean == 1)))


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_c2_full_run_conserves_volume(gpu_lib, precision):
    s = configs.c2_nonlinear(seed=0, N=1024, dt=18.0)
    g = gpu.SwmGpu(s, precision).init_from_setup()
    v0 = g.sums()[0]
    g.step(10000)
    assert g.itt == 10000
    eta = g.get("SWM_ETA")
    assert np.isfinite(eta).all() and np.abs(eta).max() > 1e-3
    v1 = g.sums()[0]
    scale = volume_scale(g, s)
    bar = 1e-11 if precision == gpu.FP64 else 1e-5
    assert abs(v1 - v0) <= bar * scale, (v0, v1, scale)
    g.close()


def test_elliptic_workload_grid_bit_identical(gpu_lib):
    s = configs.c3_elliptic(seed=0)
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    for n in (1, 3, 5):
        g.step(n - g.itt)
        o.step(n - o.itt)
        for gn, on in STATE:
            assert same(g.get(gn), o.get(on)), f"itt={n} {gn}"
    g.close()
    o.close()


def test_large_grid_bit_identical(gpu_lib):
    """8192 x 4096 FP64 (the default 32 x 8 late-history tiles), 3 steps."""
    big = configs.basin(8192, 4096, 5e3, 5e3, variant=host.FULLY_NONLINEAR, H0=500.0, dt=10.0, Ah=100.0, f0=1e-4,
                        beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.03, seed=5)
    host.add_windstress(big, host.COMP_X, configs.double_gyre_tau(big, 0.1))
    big = host.finalize(big)
    g = gpu.SwmGpu(big).init_from_setup()
    o = oracle.Oracle(big)
    g.step(3)
    o.step(3)
    for gn, on in STATE:
        assert same(g.get(gn), o.get(on)), gn
    g.close()
    o.close()
