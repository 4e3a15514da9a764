"""Serving (continuous batching): requests that arrive while others are on the
GPU join the running batch, and each result is bit-identical to the same
request run alone."""
import numpy as np
import pytest

from conftest import same
from swm_amd import gpu, serve

pytestmark = pytest.mark.gpu


def small_stream(seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(7):
        reqs.append(serve.Request(rid=k, arrival=0.02 * k, Nx=int(rng.choice([48, 80, 112, 160])),
                                  Ny=int(rng.choice([32, 64, 96])), nsteps=int(rng.integers(20, 90)),
                                  H0=float(rng.uniform(200, 4000)), Ah=float(rng.uniform(20, 500)),
                                  tau0=float(rng.uniform(0.02, 0.2)), tracer=bool(k % 2), seed=1000 + k, keep=True))
    return reqs


def alone(r, precision):
    s = serve.build_setup(r)
    m = gpu.SwmGpu(s, precision).init_from_setup()
    m.step(r.nsteps)
    out = {"SWM_ETA": m.get("SWM_ETA"), "SWM_U": m.get("SWM_U"), "SWM_V": m.get("SWM_V"), "itt": m.itt}
    if r.tracer:
        out["CH"] = m.get("CH", 0)
    m.close()
    return out


@pytest.mark.parametrize("precision,budget", [(gpu.FP32, 200e9), (gpu.FP64, 200e9), (gpu.FP32, 3.5e6)])
def test_served_results_equal_alone(gpu_lib, precision, budget):
    reqs = small_stream(7)
    srv = serve.Server(precision=precision, hbm_budget=budget, chunk=16)
    realtime = budget > 1e9
    stats = srv.run(reqs, realtime=realtime)
    assert stats["requests"] == len(reqs)
    assert stats["chunks"] >= max(r.nsteps for r in reqs) // 16
    for r in reqs:
        assert r.result["itt"] == r.nsteps and r.t_done >= r.t_admit
        if realtime:   # otherwise every request is available at once
            assert r.t_admit >= r.arrival - 1e-3
        ref = alone(r, precision)
        for k in r.result:
            if k != "itt":
                assert same(r.result[k], ref[k]), (r.rid, k)
    if budget < 1e9:
        # the budget holds about one request: admission waited for retirements
        assert any(r.t_admit > min(q.t_done for q in reqs) for r in reqs)
