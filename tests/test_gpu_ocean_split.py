"""FP32 tiles with no land in their cells or halo take a mask-free
instantiation of the step's phases 2 and 3.  The persistent tile kernel
(shape 10) runs every tile through the general path, the default 32x8 kernel
(shape 11) sends all-ocean tiles to the mask-free one: on an island grid with
tiles of both kinds the two give the same FP32 state bit for bit."""
import numpy as np
import pytest

from conftest import same
from swm_amd import configs, gpu

pytestmark = pytest.mark.gpu

FIELDS = ["SWM_U", "SWM_V", "SWM_ETA", "G_U", "G_V", "G_ETA", "G_U_M1", "G_V_M1", "G_ETA_M1"]


def tile_kinds(s, tx=32, ty=8):
    """(all-ocean tiles, tiles with land) of tx x ty tiles with a one-cell halo."""
    allo = np.ones(s.shape, bool)
    for k in ("H", "u", "v", "eta"):
        allo &= np.asarray(s.grids[k].ocean).astype(bool).reshape(s.shape)
    ny, nx = s.shape
    ocean = land = 0
    for j0 in range(0, ny, ty):
        rows = np.arange(j0 - 1, j0 + ty + 1) % ny
        for i0 in range(0, nx, tx):
            cols = np.arange(i0 - 1, i0 + tx + 1) % nx
            if allo[np.ix_(rows, cols)].all():
                ocean += 1
            else:
                land += 1
    return ocean, land


@pytest.mark.parametrize("low_part", [False, True])
def test_fp32_mask_free_tiles_equal_general_path(gpu_lib, low_part):
    s = configs.c2_nonlinear(seed=4, N=384, dt=10.0)
    ocean, land = tile_kinds(s)
    assert ocean > 100 and land > 100, (ocean, land)   # 158 and 418
    runs = []
    for shape in (11, 10):
        g = gpu.SwmGpu(s, gpu.FP32, low_part=low_part).init_from_setup()
        g.set_kernel_shape(shape)
        g.step(27)
        runs.append({f: g.get(f) for f in FIELDS} | {"CH": g.get("CH", 0), "G_CH": g.get("G_CH", 0)})
        g.close()
    for f in runs[0]:
        assert same(runs[0][f], runs[1][f]), f
