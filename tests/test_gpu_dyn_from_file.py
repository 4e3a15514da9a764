"""DYNFROMFILE datasets on the GPU (swm_gpu_add_dyn_stream): the tracer
advected by SWM's flow plus the time-interpolated dataset flow, and the master
U, V, ETA fields, bit-identical to the oracle's restatement of
dynFromFile_module.f90 in FP64, chunked and fully resident, on every kernel
family that takes datasets, and after a tuning run that rewinds the clock."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

NAMES = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("U", "U"), ("V", "V"), ("ETA", "ETA")]


def compare(g, o, tag):
    for gn, on in NAMES:
        a, b = g.get(gn), o.get(on)
        assert same(a, b), f"{tag} {gn}: max |diff| {np.nanmax(np.abs(a - b))}"
    for t in range(len(g.s.tracers)):
        for gn in ("CH", "C"):
            assert same(g.get(gn, t), o.get(gn, t)), f"{tag} tracer {gn}"


def with_datasets(chunk, seed=3, N=96, comps=(host.COMP_X, host.COMP_Y, host.COMP_ETA), nrec=4):
    """C2 physics (nonlinear basin, one tracer) with u, v and eta datasets of
    nrec records spaced 2.5 steps apart: model times fall on records, between
    them and, after 4 * 2.5 steps, wrap to the first record."""
    s = configs.c2_nonlinear(seed=seed, N=N, dt=10.0)
    rng = np.random.default_rng(seed + 100)
    amp = {host.COMP_X: 0.2, host.COMP_Y: 0.15, host.COMP_ETA: 0.3}
    for comp in comps:
        host.add_dyn_from_file(s, comp, amp[comp] * rng.standard_normal((nrec,) + s.shape), rec_dt=2.5 * s.dt,
                               chunk_size=chunk)
    return host.finalize(s)


def run_pair(s, checkpoints, shape=None):
    g = gpu.SwmGpu(s).init_from_setup()
    if shape is not None:
        g.set_kernel_shape(shape)
    o = oracle.Oracle(s)
    compare(g, o, "init")
    for n in checkpoints:
        g.step(n - g.itt)
        o.step(n - o.itt)
        compare(g, o, f"itt={n}")
    g.close()
    o.close()


@pytest.mark.parametrize("chunk", [100, 3])
def test_dataset_flow_advects_the_tracer(gpu_lib, chunk):
    run_pair(with_datasets(chunk), [1, 2, 3, 5, 11, 26, 40])


@pytest.mark.parametrize("shape", [0, 1, 9, 11])
def test_datasets_every_kernel_family(gpu_lib, shape):
    """LDS tiles (64x4, 32x8), persistent tiles and the 32x8 tile of the
    marching build; the marching wavefronts do not read datasets."""
    run_pair(with_datasets(4, seed=5, N=70), [1, 3, 17], shape=shape)


def test_eta_dataset_only(gpu_lib):
    """An eta dataset changes only the master ETA field: the tracer and SWM
    stay as without it, and the marching kernels may still run."""
    s = with_datasets(100, seed=7, N=64, comps=(host.COMP_ETA,), nrec=1)
    run_pair(s, [1, 7, 20], shape=5)


def test_tune_rewinds_the_dataset_clock(gpu_lib):
    s = with_datasets(3, seed=8, N=64)
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    g.step(6)
    o.step(6)
    best, times = g.tune(steps=4)
    assert g.itt == 6
    assert times[5] < 0                       # the marching wavefronts were skipped
    g.step(14)
    o.step(14)
    compare(g, o, "after tune")
    g.close()
    o.close()


def test_dataset_contexts_step_alone(gpu_lib):
    s = with_datasets(100, seed=1, N=48, comps=(host.COMP_X,))
    g = gpu.SwmGpu(s).init_from_setup()
    with gpu.SwmBatch() as b:
        with pytest.raises(gpu.SwmGpuError):
            b.attach(g, 5)
    with pytest.raises(gpu.SwmGpuError):
        g.set_kernel_shape(5)
    with pytest.raises(gpu.SwmGpuError):
        g.add_dyn_stream(host.COMP_X, np.zeros((2, s.Ny, s.Nx)), 10.0)
    g.close()


def test_dataset_fp32_error_bar(gpu_lib):
    s = with_datasets(100, seed=9, N=80)
    g = gpu.SwmGpu(s, gpu.FP32).init_from_setup()
    o = oracle.Oracle(s)
    for n in (3, 60):
        g.step(n - g.itt)
        o.step(n - o.itt)
        for gn, on in (("U", "U"), ("V", "V"), ("ETA", "ETA")):
            a, b = g.get(gn), o.get(on)
            assert np.abs(a - b).max() <= 1e-4 * np.abs(b).max(), (n, gn)
        if n == 3:
            # the tracer before the land cells' 0/0 coefficients have spread
            # NaN over the basin
            c32, c64 = g.get("CH", 0), o.get("CH", 0)
            fin = np.isfinite(c64)
            assert fin.sum() > 0.5 * fin.size and np.array_equal(fin, np.isfinite(c32))
            assert np.abs(c32[fin] - c64[fin]).max() <= 1e-4 * np.abs(c64[fin]).max()
    g.close()
    o.close()
