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
    res = autotune.tune_sor(m, sweeps=32, cadences=(1, 4))
    assert (res["halo"], res["check_every"], res["rows"]) in autotune.sor_candidates(cadences=(1, 4))
    assert len(res["rounds"]) >= 2 and res["rounds"][-1]["sweeps"] > res["rounds"][0]["sweeps"]
    m.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_sor_halo_widths_give_the_same_chi(gpu_lib, fp32):
    """One launch per colour (halo 0) and blocked launches of 1, 2, 4 sweeps
    (halo 2, 4, 8 cells) give chi bit for bit, for a sweep count that is not a
    multiple of the sweeps per launch; a halo whose sweeps per launch do not
    divide the cadence is refused."""
    s = configs.c2_nonlinear(seed=5, N=130, dt=10.0, tracer=False)
    m = gpu.SwmGpu(s).init_from_setup()
    m.step(9)
    m.set_sor_check_every(8)
    chis = {}
    for halo, rows in ((0, 16), (2, 16), (4, 16), (8, 16), (2, 32), (4, 32)):
        m.set_sor_halo(-1)
        m.set_sor_tile_rows(rows)
        m.set_sor_halo(halo)
        it, rr, ok = m.velocity_potential(1.9, 0.0, 45, fp32_sweeps=fp32, reset=True)
        assert it == 45 and not ok
        chis[(halo, rows)] = m.get("CHI")
    for key in chis:
        assert same(chis[key], chis[(0, 16)]), key
    with pytest.raises(gpu.SwmGpuError):
        m.set_sor_halo(8)                # 8-cell halo with 32-row tiles
    m.set_sor_tile_rows(16)
    m.set_sor_halo(8)
    with pytest.raises(gpu.SwmGpuError):
        m.set_sor_check_every(2)         # halo 8 = 4 sweeps per launch, tested every 2
    m.set_sor_halo(-1)
    m.set_sor_check_every(2)
    with pytest.raises(gpu.SwmGpuError):
        m.set_sor_halo(8)
    m.close()


def test_tune_chooses_launch_form_and_cache_replays_it(gpu_lib, tmp_path):
    """swm_gpu_tune also times graph replay against direct launches for the
    best shape; a second context of the same workload takes both from the
    tuning cache without timing anything, and both keep FP64 parity."""
    s = configs.c2_nonlinear(seed=6, N=150, dt=10.0)
    cache = autotune.TuneCache(str(tmp_path / "tune.json"))
    a = gpu.SwmGpu(s).init_from_setup()
    a.step(4)
    res = autotune.tune_context(a, steps=12, cache=cache)
    graphs, ms_direct, ms_graph = a.launch_mode()
    assert not res["cached"] and ms_direct > 0 and ms_graph > 0 and graphs == (ms_graph < ms_direct)
    b = gpu.SwmGpu(s).init_from_setup()
    res2 = autotune.tune_context(b, cache=autotune.TuneCache(str(tmp_path / "tune.json")))
    assert res2["cached"] and res2["best_shape"] == res["best_shape"] and b.launch_mode()[0] == graphs
    o = oracle.Oracle(s)
    a.step(20)
    b.step(24)
    o.step(24)
    for f, of in (("SWM_U", "u"), ("SWM_ETA", "eta")):
        assert same(a.get(f), o.get(of)) and same(b.get(f), o.get(of)), f
    for m in (a, b, o):
        m.close()
