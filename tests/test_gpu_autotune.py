"""Autotuner: tuning runs every candidate from a device snapshot and leaves
the model state exactly as it was; the selected shapes keep FP64 parity; the
SOR check cadence changes only when the residual is tested."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import autotune, configs, gpu

pytestmark = pytest.mark.gpu

FIELDS = ["SWM_U", "SWM_V", "SWM_ETA", "G_U", "G_V", "G_ETA", "G_U_M1", "G_V_M1", "G_ETA_M1"]


def snapshot(m, ntrc):
    out = {f: m.get(f) for f in FIELDS}
    for t in range(ntrc):
        out[f"CH{t}"] = m.get("CH", t)
    out["itt"] = m.itt
    return out


def same_snap(a, b):
    return all((a[k] == b[k]) if k == "itt" else same(a[k], b[k]) for k in a)


def test_context_tune_restores_state_and_keeps_parity(gpu_lib):
    s = configs.c2_nonlinear(seed=8, N=130, dt=10.0)
    m = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    m.step(5)
    o.step(5)
    before = snapshot(m, 1)
    res = autotune.tune_context(m, steps=3)
    assert 0 <= res["best_shape"] < len(res["ms_per_step"])
    assert all(t > 0 for t in res["ms_per_step"])
    assert same_snap(before, snapshot(m, 1))
    m.step(7)
    o.step(7)
    for f, of in (("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_ETA", "G_eta")):
        assert same(m.get(f), o.get(of)), f
    assert same(m.get("CH", 0), o.get("CH", 0))
    m.close()


def test_batch_tune_restores_members(gpu_lib):
    setups = [configs.c2_nonlinear(seed=20 + k, N=n, dt=10.0, tracer=bool(k % 2)) for k, n in enumerate((64, 96, 150))]
    mixed = [gpu.SwmGpu(s).init_from_setup() for s in setups]
    alone = [gpu.SwmGpu(s).init_from_setup() for s in setups]
    with gpu.SwmBatch() as b:
        for m in mixed:
            b.attach(m, 30)
        b.step(4)
        before = [snapshot(m, len(s.tracers)) for m, s in zip(mixed, setups)]
        res = autotune.tune_batch(b, steps=3)
        assert res["best_shape"] in range(4)
        assert all(same_snap(x, snapshot(m, len(s.tracers))) for x, m, s in zip(before, mixed, setups))
        assert [b.remaining(m) for m in mixed] == [26, 26, 26]
        b.step(26)
    for m, a, s in zip(mixed, alone, setups):
        a.step(30)
        assert same_snap(snapshot(a, len(s.tracers)), snapshot(m, len(s.tracers)))
        m.close()
        a.close()


def test_sor_check_cadence(gpu_lib):
    s = configs.c2_nonlinear(seed=2, N=96, dt=10.0, tracer=False)
    m = gpu.SwmGpu(s).init_from_setup()
    m.step(15)
    it1, rr1, ok1 = m.velocity_potential(1.9, 1e-10, 5000, reset=True)
    m.set_sor_check_every(8)
    it8, rr8, ok8 = m.velocity_potential(1.9, 1e-10, 5000, reset=True)
    assert ok1 and ok8 and it8 % 8 == 0 and it1 <= it8 < it1 + 8 and rr8 < 1e-10
    chi8 = m.get("CHI")
    m.set_sor_check_every(1)
    m.velocity_potential(1.9, 0.0, it8, reset=True)     # exactly it8 sweeps, tested every sweep
    assert same(m.get("CHI"), chi8)
    res = autotune.tune_sor(m, sweeps=64, cadences=(1, 4))
    assert res["best_check_every"] in (1, 4)
    m.close()
