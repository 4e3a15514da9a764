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


def test_c5_stream_slice_with_largest_grid(gpu_lib):
    """A slice of the serving stream's own request sequence (seed 0, 20 req/s,
    realtime arrivals, FP32 fields) whose first request is given the stream's
    largest grid, 2048 x 1024, and 3000 steps; the later requests keep their
    grids with run lengths cut to 40..80 steps, so they arrive while the large
    one is on the GPU.  Every result equals the same request run alone, and
    the large request shares its launches with others."""
    reqs = serve.request_stream(seed=0, rate=20.0, duration=0.5)[:6]
    assert len(reqs) >= 4
    reqs[0].Nx, reqs[0].Ny, reqs[0].nsteps = 2048, 1024, 3000
    for k, r in enumerate(reqs[1:]):
        r.nsteps = 40 + 10 * k
    for r in reqs:
        r.keep = True
    srv = serve.Server(precision=gpu.FP32, hbm_budget=200e9, chunk=16)
    stats = srv.run(reqs, realtime=True)
    assert stats["requests"] == len(reqs)
    assert stats["max_members_per_launch"] >= 2
    big = reqs[0]
    assert big.max_batch >= 2, "the 2048 x 1024 request never shared a launch"
    for r in reqs:
        assert r.result["itt"] == r.nsteps
        ref = alone(r, gpu.FP32)
        for k in r.result:
            if k != "itt":
                assert same(r.result[k], ref[k]), (r.rid, r.Nx, r.Ny, k)
