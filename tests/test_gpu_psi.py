"""Streamfunction diagnostic on the GPU (swm_gpu_sync "PSI",
computeStreamfunction, calc_lib.f90:727-780): bit-identical to the reference's
recorded psi for the recorded flow, and to the oracle on model states; the
row sums are staged through LDS in chunks, so a grid wider than one chunk is
covered too."""
import numpy as np
import pytest

import oracle
from conftest import diaglist_grid_setup, linear_osc_setup, load_golden, same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu


def test_psi_of_the_reference_record(gpu_lib):
    """Upload the recorded u, v of test_diaglist and read PSI back: equal, bit
    for bit, to the reference's recorded psi at three snapshots."""
    npz = load_golden("diaglist_psi.npz")
    s = diaglist_grid_setup(npz)
    g = gpu.SwmGpu(s)
    for k in range(npz["psi"].shape[0]):
        u = np.where(npz["u"][k] == npz["u_missing"], 0.0, npz["u"][k])
        v = np.where(npz["v"][k] == npz["v_missing"], 0.0, npz["v"][k])
        g.upload_state(u, v, np.zeros(s.shape), itt=0)
        got = g.get("PSI")
        ref = npz["psi"][k]
        m = ref != npz["psi_missing"]
        assert np.array_equal(got[m], ref[m]), k
    g.close()


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_psi_of_a_nonlinear_run(gpu_lib, precision):
    s = configs.c2_nonlinear(seed=4, N=96, dt=10.0)
    g = gpu.SwmGpu(s, precision).init_from_setup()
    o = oracle.Oracle(s)
    g.step(30)
    o.step(30)
    if precision == gpu.FP64:
        assert same(g.get("PSI"), o.psi())
    else:
        # the oracle's psi of the FP32 state: only the FP32 thicknesses and
        # metric factors differ
        want = o.psi(g.get("SWM_U"), g.get("SWM_V"))
        assert np.abs(g.get("PSI") - want).max() <= 1e-6 * np.abs(want).max()
    g.close()
    o.close()


def test_psi_not_barotropic(gpu_lib):
    """Without BAROTROPIC the paths integrate the velocities themselves."""
    npz = load_golden("fixed_length_output.npz")
    s = linear_osc_setup(npz)
    s.barotropic = False
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    g.step(40)
    o.step(40)
    assert same(g.get("PSI"), o.psi())
    g.close()
    o.close()


def test_psi_wide_grid(gpu_lib):
    """Nx = 2309: two LDS chunks and ten 256-cell segments per row."""
    s = host.finalize(configs.basin(2309, 21, 5e3, 5e3, variant=host.FULLY_NONLINEAR, H0=300.0, dt=10.0,
                                    f0=1e-4, beta=2e-11, seed=8, land_fraction=0.05))
    rng = np.random.default_rng(3)
    u = rng.standard_normal(s.shape) * s.grids["u"].ocean
    v = rng.standard_normal(s.shape) * s.grids["v"].ocean
    g = gpu.SwmGpu(s)
    g.upload_state(u, v, np.zeros(s.shape), itt=0)
    o = oracle.Oracle(s)
    assert same(g.get("PSI"), o.psi(u, v))
    g.close()
    o.close()
