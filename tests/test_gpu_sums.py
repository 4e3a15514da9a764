"""FP64 compensated conservation sums (swm_gpu_sums) against exactly rounded
sums (math.fsum) of the same per-cell terms, and the FP32 storage path's
volume drift against the FP64 oracle."""
import math

import numpy as np
import pytest

import oracle
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu


def cell_terms(s, u, v, eta, ch=None):
    """Per-cell terms in the kernel's operation order (FP64)."""
    oc = s.grids["eta"].ocean == 1
    w = np.broadcast_to(s.grids["eta"].cos_lat[:, None], s.shape)
    H = s.grids["eta"].H
    D = eta + H if s.variant == host.FULLY_NONLINEAR else H
    ue, vn = np.roll(u, -1, axis=1), np.roll(v, -1, axis=0)
    ke = 0.5 * D * (0.5 * (u * u + ue * ue) + 0.5 * (v * v + vn * vn))
    pe = 0.5 * s.G * (eta * eta)
    out = [(eta * w)[oc], ((ke + pe) * w)[oc]]
    if ch is not None:
        out.append((ch * w)[oc])
    return out


def test_sums_fp64_match_exact_rounding(gpu_lib):
    s = configs.c2_nonlinear(seed=4, N=160, dt=10.0)
    g = gpu.SwmGpu(s).init_from_setup()
    # step 1: CH is still finite (the land 0/0 coefficients reach CH at step 2
    # and then spread into the ocean, as in the reference)
    g.step(1)
    ch = g.get("CH", 0)
    assert np.isfinite(ch).all()
    got_t = g.sums(0)[2]
    t = cell_terms(s, g.get("SWM_U"), g.get("SWM_V"), g.get("SWM_ETA"), ch)[2]
    assert abs(got_t - math.fsum(t)) <= 1e-15 * abs(math.fsum(t)) + 1e-20 * np.abs(t).sum() * t.size
    g.step(24)
    terms = cell_terms(s, g.get("SWM_U"), g.get("SWM_V"), g.get("SWM_ETA"))
    got = g.sums(0)
    for k, t in enumerate(terms):
        exact = math.fsum(t)
        # Neumaier summation: |error| <= 2u|S| + O(n u^2) sum|x|
        assert abs(got[k] - exact) <= 1e-15 * abs(exact) + 1e-20 * np.abs(t).sum() * t.size, (k, got[k], exact)
    assert g.sums(-1)[2] == 0.0
    # reproducible run to run
    assert g.sums(-1)[:2] == got[:2]
    g.close()


def test_sums_fp32_volume_drift_vs_fp64_oracle(gpu_lib):
    """FP32 storage keeps the cos-weighted volume within 1e-6 of sum|eta*w| of
    the FP64 oracle after 400 steps (flux-form continuity conserves it)."""
    s = configs.c2_nonlinear(seed=3, N=128, dt=10.0, tracer=False)
    g = gpu.SwmGpu(s, precision=gpu.FP32).init_from_setup()
    o = oracle.Oracle(s)
    g.step(400)
    o.step(400)
    ref = cell_terms(s, o.get("u"), o.get("v"), o.get("eta"))
    vol32, en32, _ = g.sums()
    scale = np.abs(ref[0]).sum()
    assert abs(vol32 - math.fsum(ref[0])) <= 1e-6 * scale
    assert abs(en32 - math.fsum(ref[1])) <= 1e-4 * abs(math.fsum(ref[1]))
    g.close()
