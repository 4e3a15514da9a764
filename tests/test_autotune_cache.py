"""The autotuner's host logic without a GPU: workload signatures, the SOR
(halo, cadence) candidate set, the JSON tuning cache, and that a cache hit
applies the stored choice to the context without timing anything."""
import json

from swm_amd import autotune, configs, gpu


class FakeCtx:
    """Duck-typed stand-in for SwmGpu: records what the tuner applies."""

    def __init__(self, s, precision=gpu.FP64):
        self.s, self.precision = s, precision
        self.applied = []

    def set_kernel_shape(self, k):
        self.applied.append(("shape", k))

    def set_launch_mode(self, g):
        self.applied.append(("graphs", g))

    def set_sor_halo(self, h):
        self.applied.append(("halo", h))

    def set_sor_check_every(self, c):
        self.applied.append(("every", c))

    def set_sor_tile_rows(self, r):
        self.applied.append(("rows", r))

    def tune(self, steps):
        raise AssertionError("a cache hit must not time anything")


def test_signature_separates_what_the_cost_depends_on():
    a = FakeCtx(configs.c2_nonlinear(seed=1, N=64, dt=10.0))
    b = FakeCtx(configs.c2_nonlinear(seed=2, N=64, dt=10.0))          # other seed: same workload
    c = FakeCtx(configs.c2_nonlinear(seed=1, N=64, dt=10.0, tracer=False))
    d = FakeCtx(configs.c2_nonlinear(seed=1, N=64, dt=10.0), gpu.FP32)
    assert autotune.signature(a) == autotune.signature(b)
    assert len({autotune.signature(x) for x in (a, c, d)}) == 3
    assert autotune.batch_signature([a, c]) == autotune.batch_signature([c, a])


def test_sor_candidates_respect_the_cadence_and_lds():
    cand = autotune.sor_candidates()
    assert (0, 1, 16) in cand and (2, 1, 32) in cand and (8, 4, 16) in cand and (4, 16, 32) in cand
    assert (4, 1, 16) not in cand and (8, 2, 16) not in cand
    assert (8, 8, 32) not in cand and (0, 4, 32) not in cand     # LDS limit; no tile without blocking
    assert all(h <= 2 or c % (h // 2) == 0 for h, c, r in cand)


def test_cache_round_trip_and_hits_apply(tmp_path):
    path = str(tmp_path / "tune.json")
    m = FakeCtx(configs.c2_nonlinear(seed=1, N=64, dt=10.0))
    cache = autotune.TuneCache(path)
    cache.put("context", autotune.signature(m), {"best_shape": 3, "ms_per_step": [1.0], "graphs": True,
                                                  "ms_direct": 2.0, "ms_graph": 1.5})
    cache.put("sor", f"{autotune.signature(m)}/w1.9/f0", {"halo": 4, "check_every": 8, "rows": 32,
                                                         "sweeps_per_launch": 2, "rounds": []})
    on_disk = json.load(open(path))
    assert cache.device in on_disk
    again = autotune.TuneCache(path)
    res = autotune.tune_context(m, cache=again)
    assert res["cached"] and m.applied == [("shape", 3), ("graphs", True)]
    m.applied.clear()
    res = autotune.tune_sor(m, omega=1.9, cache=again)
    assert res["cached"] and m.applied == [("halo", -1), ("rows", 32), ("every", 8), ("halo", 4)]
