"""Time-interpolated forcing streams on the GPU (swm_gpu_add_forcing_stream):
bit-identical to the oracle's memchunk restatement in FP64, for windstress
under FULLY_NONLINEAR (divided by the step's Du/Dv) and under the linear
variants, custom and heating streams, chunked and fully resident chunks,
every kernel shape, and after a tuning run that rewinds the model time."""
import numpy as np
import pytest

import oracle
from conftest import linear_osc_setup, load_golden, same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

STATE = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_U", "G_u"), ("G_V", "G_v"), ("G_ETA", "G_eta")]
FORCING = [("F_X", "F_x"), ("F_Y", "F_y"), ("F_ETA", "F_eta")]


def compare(g, o, names, tag):
    for gn, on in names:
        a, b = g.get(gn), o.get(on)
        assert same(a, b), f"{tag} {gn}: max |diff| {np.nanmax(np.abs(a - b))}"
    for t in range(len(g.s.tracers)):
        for gn, on in (("CH", "CH"), ("C", "C")):
            assert same(g.get(gn, t), o.get(on, t)), f"{tag} tracer {gn}"


def nonlinear_with_streams(chunk, seed=3, N=96):
    """C2 physics with the steady windstress replaced by two windstress
    streams of 5 records, spaced 2.5 steps apart (times fall on and between
    records), and a custom stream on the continuity equation."""
    s = configs.c2_nonlinear(seed=seed, N=N, dt=10.0)
    s.wind.clear()
    rng = np.random.default_rng(seed)
    taux = configs.double_gyre_tau(s, 0.1)[None] * (1.0 + 0.5 * rng.standard_normal((5,) + s.shape))
    tauy = 0.05 * rng.standard_normal((5,) + s.shape)
    host.add_forcing_stream(s, host.STREAM_WIND, host.COMP_X, taux, rec_dt=2.5 * s.dt, chunk_size=chunk)
    host.add_forcing_stream(s, host.STREAM_WIND, host.COMP_Y, tauy, rec_dt=2.5 * s.dt, chunk_size=chunk)
    host.add_forcing_stream(s, host.STREAM_CUSTOM, host.COMP_ETA, 1e-7 * rng.standard_normal((3,) + s.shape),
                            rec_dt=4.0 * s.dt, chunk_size=chunk)
    return host.finalize(s)


def run_pair(s, checkpoints, names, shape=None):
    g = gpu.SwmGpu(s).init_from_setup()
    if shape is not None:
        g.set_kernel_shape(shape)
    o = oracle.Oracle(s)
    compare(g, o, names, "init")
    for n in checkpoints:
        g.step(n - g.itt)
        o.step(n - o.itt)
        compare(g, o, names, f"itt={n}")
    g.close()
    o.close()


@pytest.mark.parametrize("chunk", [100, 3])
def test_nonlinear_windstress_streams(gpu_lib, chunk):
    s = nonlinear_with_streams(chunk)
    run_pair(s, [1, 2, 3, 5, 13, 26, 40], STATE + FORCING)


@pytest.mark.parametrize("shape", [1, 0, 5, 9])
def test_streams_every_kernel_family(gpu_lib, shape):
    """LDS tiles (32x8, 64x4), marching wavefronts and persistent tiles."""
    s = nonlinear_with_streams(4, seed=5, N=70)
    run_pair(s, [1, 3, 17], STATE + FORCING, shape=shape)


def test_linear_variant_streams_after_oscillation(gpu_lib):
    """LINEARISED_STATE_OF_REST + BAROTROPIC (test_fixed_length_output set-up)
    with its oscillating zonal forcing, then a meridional windstress stream
    divided by RHO0*H_v, a custom zonal stream and a heating stream."""
    npz = load_golden("fixed_length_output.npz")
    s = linear_osc_setup(npz)
    s.tau_scale = host.f32(1e-2)
    rng = np.random.default_rng(11)
    host.add_forcing_stream(s, host.STREAM_WIND, host.COMP_Y, rng.standard_normal((7,) + s.shape),
                            rec_dt=3.0 * s.dt, chunk_size=5)
    host.add_forcing_stream(s, host.STREAM_CUSTOM, host.COMP_X, 1e-6 * rng.standard_normal((4,) + s.shape),
                            rec_dt=1.5 * s.dt)
    host.add_forcing_stream(s, host.STREAM_HEATING, host.COMP_ETA, 1e-6 * rng.standard_normal((2,) + s.shape),
                            rec_dt=5.0 * s.dt)
    run_pair(s, [1, 2, 3, 9, 22, 45], STATE + FORCING)


def test_not_barotropic_divides_by_rho0_only(gpu_lib):
    npz = load_golden("fixed_length_output.npz")
    s = linear_osc_setup(npz)
    s.barotropic = False
    host.add_forcing_stream(s, host.STREAM_WIND, host.COMP_X, np.random.default_rng(2).standard_normal((3,) + s.shape),
                            rec_dt=2.0 * s.dt)
    run_pair(s, [1, 4, 11], STATE + FORCING)


def test_tune_rewinds_the_stream_clock(gpu_lib):
    """swm_gpu_tune runs every shape from a snapshot and restores itt; the
    streams restart their chunk at the earlier time and the run stays
    bit-identical to the oracle."""
    s = nonlinear_with_streams(3, seed=7, N=64)
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    g.step(6)
    o.step(6)
    g.tune(steps=4)
    assert g.itt == 6
    g.step(14)
    o.step(14)
    compare(g, o, STATE + FORCING, "after tune")
    g.close()
    o.close()


def test_streams_fp32_error_bar(gpu_lib):
    s = nonlinear_with_streams(100, seed=9, N=80)
    g = gpu.SwmGpu(s, gpu.FP32).init_from_setup()
    o = oracle.Oracle(s)
    g.step(60)
    o.step(60)
    e32, e64 = g.get("SWM_ETA"), o.get("eta")
    assert np.abs(e32 - e64).max() <= 1e-4 * np.abs(e64).max()
    g.close()
    o.close()


def test_stream_contexts_step_alone(gpu_lib):
    s = nonlinear_with_streams(100, seed=1, N=48)
    g = gpu.SwmGpu(s).init_from_setup()
    with gpu.SwmBatch() as b:
        with pytest.raises(gpu.SwmGpuError):
            b.attach(g, 5)
    with pytest.raises(gpu.SwmGpuError):
        g.add_forcing_stream(host.STREAM_HEATING, host.COMP_X, np.zeros((2, s.Ny, s.Nx)), 10.0)
    g.close()
