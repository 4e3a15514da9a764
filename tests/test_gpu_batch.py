"""Batched ensembles (swm_gpu_batch_*): every member of a batch gets the same
result as the same context stepped alone, bit for bit, in FP64 and FP32;
members stop at their own step counts; detach / re-attach keeps the state."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

FIELDS = ["SWM_U", "SWM_V", "SWM_ETA", "G_U", "G_V", "G_ETA", "G_U_M1", "G_V_M1", "G_ETA_M1"]


def members():
    """Uneven domains: different sizes, with and without a tracer, islands or not."""
    return [
        configs.c2_nonlinear(seed=11, N=96, dt=10.0),
        configs.c2_nonlinear(seed=12, N=130, dt=8.0, tracer=False),
        host.finalize(configs.basin(70, 45, 5e3, 5e3, variant=host.FULLY_NONLINEAR, H0=800.0, dt=6.0, Ah=50.0,
                                    f0=1e-4, beta=2e-11, seed=13, land_fraction=0.0)),
        configs.c2_nonlinear(seed=14, N=64, dt=10.0),
    ]


def check_same(a, b, ntrc, tag):
    for f in FIELDS:
        assert same(a.get(f), b.get(f)), f"{tag} {f}"
    for t in range(ntrc):
        for f in ("CH", "G_CH", "G_CH_M1", "C"):
            assert same(a.get(f, t), b.get(f, t)), f"{tag} tracer {f}"


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_batch_members_equal_alone(gpu_lib, precision):
    setups = members()
    steps = [37, 25, 40, 9]
    alone = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
    mixed = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
    for m, n in zip(alone, steps):
        m.step(n)
    with gpu.SwmBatch() as b:
        for m, n in zip(mixed, steps):
            b.attach(m, n)
        active = b.step(15)
        assert active == 3                       # the 9-step member is done
        assert [b.remaining(m) for m in mixed] == [22, 10, 25, 0]
        assert mixed[3].itt == 9                 # stopped at its own count
        active = b.step(100)
        assert active == 0
        for m, a, s, n in zip(mixed, alone, setups, steps):
            assert m.itt == a.itt == n
            check_same(m, a, len(s.tracers), f"member {s.Nx}x{s.Ny}")
    for m in alone + mixed:
        m.close()


@pytest.mark.parametrize("shape", [0, 1, 2, 3])
def test_every_batch_form_equals_alone(gpu_lib, shape):
    """64x4 / 32x8 tiles, one block per tile (0, 1) or the persistent grid
    draining per-XCD tile queues (2, 3): bit-identical members, across graph
    replays (the queue counters are cleared inside the captured step)."""
    setups = members()
    steps = [37, 25, 40, 9]
    for precision in (gpu.FP64, gpu.FP32):
        alone = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
        mixed = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
        for m, n in zip(alone, steps):
            m.step(n)
        with gpu.SwmBatch() as b:
            b.set_shape(shape)
            for m, n in zip(mixed, steps):
                b.attach(m, n)
            b.step(6)
            b.step(100)
            for m, a, s, n in zip(mixed, alone, setups, steps):
                assert m.itt == a.itt == n
                check_same(m, a, len(s.tracers), f"shape {shape} member {s.Nx}x{s.Ny}")
        for m in alone + mixed:
            m.close()


@pytest.mark.parametrize("shape", [1, 3])
def test_batch_of_more_members_than_one_block(gpu_lib, shape):
    """300 small members (the step state of more than one 256-thread block of
    the member advance), one block per tile and the persistent queue: every
    member advances exactly its own steps, equal to itself alone."""
    setups = [host.finalize(configs.basin(32 + 16 * (k % 3), 16 + 16 * (k % 2), 5e3, 5e3,
                                          variant=host.FULLY_NONLINEAR, H0=300.0 + k, dt=5.0, Ah=50.0, f0=1e-4,
                                          beta=2e-11, lateral_mixing=True, seed=100 + k)) for k in range(300)]
    mixed = [gpu.SwmGpu(s).init_from_setup() for s in setups]
    with gpu.SwmBatch() as b:
        b.set_shape(shape)
        for k, m in enumerate(mixed):
            b.attach(m, 5 + k % 7)
        b.step(3)
        b.step(20)
        assert all(m.itt == 5 + k % 7 for k, m in enumerate(mixed))
    for k in (0, 1, 255, 256, 257, 299):
        a = gpu.SwmGpu(setups[k]).init_from_setup()
        a.step(5 + k % 7)
        check_same(mixed[k], a, 0, f"member {k}")
        a.close()
    for m in mixed:
        m.close()


def test_batch_matches_oracle_and_resumes_alone(gpu_lib):
    """A member's batched run is bit-identical to the oracle; after detach it
    continues alone, and re-attaching with a new count continues the batch."""
    s = configs.c2_nonlinear(seed=21, N=80, dt=10.0)
    o = oracle.Oracle(s)
    m = gpu.SwmGpu(s).init_from_setup()
    other = gpu.SwmGpu(configs.c2_nonlinear(seed=22, N=120, dt=10.0)).init_from_setup()
    with gpu.SwmBatch() as b:
        b.attach(m, 12)
        b.attach(other, 30)
        b.step(12)
        o.step(12)
        assert same(m.get("SWM_ETA"), o.get("eta")) and same(m.get("CH", 0), o.get("CH", 0))
        with pytest.raises(gpu.SwmGpuError):
            m.step(1)                            # attached contexts advance through the batch
        b.detach(m)
        m.step(5)
        o.step(5)
        assert same(m.get("SWM_U"), o.get("u"))
        b.attach(m, 7)
        b.set_steps(other, 7)
        assert b.step(50) == 0
        o.step(7)
        assert m.itt == 24 and other.itt == 19
        assert same(m.get("SWM_ETA"), o.get("eta")) and same(m.get("G_ETA"), o.get("G_eta"))
    m.close()
    other.close()


def test_batch_rejects_mixed_precision(gpu_lib):
    a = gpu.SwmGpu(configs.c2_nonlinear(seed=1, N=48, dt=10.0), gpu.FP64).init_from_setup()
    f = gpu.SwmGpu(configs.c2_nonlinear(seed=1, N=48, dt=10.0), gpu.FP32).init_from_setup()
    with gpu.SwmBatch() as b:
        b.attach(a, 3)
        with pytest.raises(gpu.SwmGpuError):
            b.attach(f, 3)
        b.step(3)
        assert a.itt == 3
    a.close()
    f.close()


@pytest.mark.parametrize("precision", [gpu.FP32, gpu.FP64])
def test_batch_tune_warmup_close_equals_alone(gpu_lib, precision):
    """The ensemble benchmark's sequence: attach, tune (snapshot, timed runs of
    each tile shape, restore), warm-up, timed steps, close.  After the batch is
    closed every member equals the same member stepped alone, bit for bit, and
    its FP64 sums equal the alone member's (volume and energy finite; the
    tracer inventory is NaN once the land cells' 0/0 coefficients have made CH
    NaN there, as in the reference)."""
    setups = [configs.ensemble_member(seed=31 + k, Nx=nx, Ny=ny, H0=h, Ah=ah, tau0=t, tracer=tr, nsteps=40)
              for k, (nx, ny, h, ah, t, tr) in enumerate([(128, 256, 900.0, 60.0, 0.1, True),
                                                          (256, 128, 3000.0, 300.0, 0.05, False),
                                                          (384, 128, 450.0, 150.0, 0.15, True),
                                                          (128, 128, 2000.0, 25.0, 0.02, False)])]
    warm, steps = 3, 17
    mixed = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
    b = gpu.SwmBatch()
    for m in mixed:
        b.attach(m, warm + steps)
    best, times = b.tune(steps=4)
    assert best in range(len(times)) and len(times) == 4 and all(t > 0 for t in times)
    assert all(m.itt == 0 for m in mixed)          # the tuning runs were undone
    b.step(warm)
    b.step(steps)
    b.close()
    for m, s in zip(mixed, setups):
        a = gpu.SwmGpu(s, precision).init_from_setup()
        a.step(warm + steps)
        assert m.itt == a.itt == warm + steps
        check_same(m, a, len(s.tracers), f"member {s.Nx}x{s.Ny}")
        sm, sa = m.sums(0 if s.tracers else -1), a.sums(0 if s.tracers else -1)
        assert np.isfinite(sm[0]) and np.isfinite(sm[1]), sm
        assert same(np.array(sm), np.array(sa)), (sm, sa)
        a.close()
        m.close()


def test_c4_ensemble_all_members(gpu_lib):
    """The 64-member ensemble at its stated shapes (nx, ny from 128..1024, seed
    0), attached to one batch: after 8 steps (the graph-replayed part
    included) every member equals the same member stepped alone, bit for bit,
    in FP64 and FP32; three members of different sizes also equal the
    oracle."""
    setups = configs.c4_ensemble(seed=0, n=64, nsteps=8)
    assert len(setups) == 64 and len({(s.Nx, s.Ny) for s in setups}) > 10
    for precision in (gpu.FP64, gpu.FP32):
        mixed = [gpu.SwmGpu(s, precision).init_from_setup() for s in setups]
        with gpu.SwmBatch() as b:
            for m in mixed:
                b.attach(m, 8)
            assert b.step(3) == 64
            assert b.step(10) == 0
        for m, s in zip(mixed, setups):
            a = gpu.SwmGpu(s, precision).init_from_setup()
            a.step(8)
            for f in ("SWM_U", "SWM_V", "SWM_ETA", "G_ETA"):
                assert same(m.get(f), a.get(f)), (precision, s.Nx, s.Ny, f)
            if s.tracers:
                assert same(m.get("CH", 0), a.get("CH", 0)), (precision, s.Nx, s.Ny, "CH")
            a.close()
        if precision == gpu.FP64:
            picks = sorted(range(64), key=lambda k: setups[k].Nx * setups[k].Ny)
            for k in (picks[0], picks[32], picks[-1]):
                o = oracle.Oracle(setups[k])
                o.step(8)
                assert same(mixed[k].get("SWM_ETA"), o.get("eta")) and same(mixed[k].get("SWM_U"), o.get("u"))
                o.close()
        for m in mixed:
            m.close()
