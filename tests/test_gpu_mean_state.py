"""LINEARISED_MEAN_STATE on the GPU: bit-identical to the oracle in FP64 on a
basic flow with islands, lateral mixing (free slip), windstress, sponges and
a time-interpolated stream, for every LDS tile shape and the diagnostic
fields; FP32 within its error bar; the shapes and batches it does not run on
are refused."""
import numpy as np
import pytest

import oracle
from conftest import diaglist_grid_setup, load_golden, same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

STATE = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_U", "G_u"), ("G_V", "G_v"), ("G_ETA", "G_eta"),
         ("G_U_M1", "G_u_m1"), ("G_V_M1", "G_v_m1"), ("G_ETA_M1", "G_eta_m1")]
DERIVED = [("SWM_RELVORT", "zeta"), ("SWM_POT", "Pot"), ("SWM_EDENS", "EDens"), ("SWM_MU", "MU"), ("SWM_MV", "MV"),
           ("SWM_DU", "Du"), ("F_X", "F_x"), ("F_Y", "F_y"), ("F_ETA", "F_eta"), ("SWM_LATMIX_U", "latmix_u"),
           ("SWM_LATMIX_V", "latmix_v")]


def compare(g, o, names, tag):
    for gn, on in names:
        a, b = g.get(gn), o.get(on)
        assert same(a, b), f"{tag} {gn}: max |diff| {np.nanmax(np.abs(a - b))}"


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


@pytest.mark.parametrize("shape", [11, 1, 0, 2, 3, 4, 12])
def test_mean_state_every_tile_shape(gpu_lib, shape):
    s = configs.mean_state_basin(seed=4, Nx=100, Ny=70)
    run_pair(s, [1, 2, 3, 9, 31], STATE + DERIVED, shape=shape)


def test_mean_state_sponges_and_stream(gpu_lib):
    """Spatially varying implicit damping (full arrays) and a meridional
    windstress stream interpolated in time, on a grid with Ny = Nx."""
    s = configs.mean_state_basin(seed=6, Nx=72, Ny=72, land_fraction=0.08)
    rng = np.random.default_rng(2)
    s.impl_u = 1.0 + 1e-3 * rng.random(s.shape)
    s.impl_eta = 1.0 + 1e-4 * rng.random(s.shape)
    host.add_forcing_stream(s, host.STREAM_WIND, host.COMP_Y, 0.02 * rng.standard_normal((6,) + s.shape),
                            rec_dt=3.5 * s.dt, chunk_size=4)
    run_pair(s, [1, 5, 18, 40], STATE + DERIVED)


def test_mean_state_on_the_diaglist_domain(gpu_lib):
    """The test_diaglist domain (file bathymetry, H_overwrite, spherical f,
    61 x 41) with a basic flow from a smooth streamfunction."""
    npz = load_golden("diaglist_psi.npz")
    base = diaglist_grid_setup(npz)
    s = host.Setup(Nx=base.Nx, Ny=base.Ny, variant=host.LINEARISED_MEAN_STATE, grids=base.grids,
                   dLambda=base.dLambda, dTheta=base.dTheta, dt=base.dt, A=base.A, lbc=0.0, barotropic=True,
                   lateral_mixing=True, Ah=5e3, tau_scale=host.f32(1e-2))
    rng = np.random.default_rng(8)
    psi = configs.fourier_modes(s, rng, n_modes=4, kmax=2) * 1e4 * s.grids["H"].ocean
    host.set_mean_state(s, psi)
    host.init_damping(s, [])
    s.eta0 = 0.5 * configs.fourier_modes(s, rng, n_modes=2, kmax=2)
    s = host.finalize(s)
    run_pair(s, [1, 3, 25, 60], STATE + DERIVED)


def test_mean_state_fp32_error_bar(gpu_lib):
    s = configs.mean_state_basin(seed=5, Nx=96, Ny=64)
    g = gpu.SwmGpu(s, gpu.FP32).init_from_setup()
    o = oracle.Oracle(s)
    g.step(200)
    o.step(200)
    e32, e64 = g.get("SWM_ETA"), o.get("eta")
    assert np.abs(e32 - e64).max() <= 1e-4 * np.abs(e64).max()
    g.close()
    o.close()


def test_mean_state_refusals(gpu_lib):
    s = configs.mean_state_basin(seed=1, Nx=48, Ny=40)
    g = gpu.SwmGpu(s).init_from_setup()
    with pytest.raises(gpu.SwmGpuError):
        g.set_kernel_shape(5)          # marching wavefronts
    with pytest.raises(gpu.SwmGpuError):
        g.set_kernel_shape(9)          # persistent tiles
    with gpu.SwmBatch() as b:
        with pytest.raises(gpu.SwmGpuError):
            b.attach(g, 3)
    best, times = g.tune(steps=3)
    assert best in (0, 1, 2, 3, 4, 11, 12) and all(t < 0 for t in times[5:11]) and min(times[11:]) > 0
    g.close()
