"""Static fields that vary along a row are kept as full arrays instead of two
values per row; the tile kernels then read them from the array and not from
their per-row table.  Every such field at once — windstress in x and y,
constant forcing on all three equations, implicit damping of u, v and eta,
tracer relaxation, target and consumption — bit-identical to the oracle for
the shapes that stage the per-row table in LDS (32x8, 32x8 late, 64x8) and
those that read it with scalar loads (64x4, 64x4 late)."""
import numpy as np
import pytest

import oracle
from conftest import same
from swm_amd import configs, gpu, host

pytestmark = pytest.mark.gpu

STATE = [("SWM_U", "u"), ("SWM_V", "v"), ("SWM_ETA", "eta"), ("G_U", "G_u"), ("G_V", "G_v"), ("G_ETA", "G_eta")]


def full_field_basin():
    s = configs.basin(90, 60, 8e3, 8e3, variant=host.FULLY_NONLINEAR, H0=300.0, dt=15.0, Ah=150.0, f0=1e-4,
                      beta=2e-11, drag=1e-6, lateral_mixing=True, land_fraction=0.06, seed=12)
    rng = np.random.default_rng(21)
    ou, ov, oe = (s.grids[g].ocean for g in ("u", "v", "eta"))
    host.add_windstress(s, host.COMP_X, 0.05 * rng.standard_normal(s.shape) * ou)
    host.add_windstress(s, host.COMP_Y, 0.05 * rng.standard_normal(s.shape) * ov)
    s.F_x_const = 1e-7 * rng.standard_normal(s.shape) * ou
    s.F_y_const = 1e-7 * rng.standard_normal(s.shape) * ov
    s.F_eta_const = 1e-8 * rng.standard_normal(s.shape) * oe
    s.impl_u = s.impl_u * (1.0 + 1e-3 * rng.random(s.shape))
    s.impl_v = s.impl_v * (1.0 + 1e-3 * rng.random(s.shape))
    s.impl_eta = 1.0 + 1e-4 * rng.random(s.shape)
    host.add_tracer(s, configs.fourier_modes(s, rng), kappa_h=40.0, C0=0.5 + 0.1 * rng.random(s.shape),
                    gamma_C=1e-6 * rng.random(s.shape), cons=1e-7 * rng.random(s.shape))
    return host.finalize(s)


@pytest.mark.parametrize("shape", [11, 1, 2, 0, 12])
def test_full_array_fields_every_row_table_path(gpu_lib, shape):
    s = full_field_basin()
    g = gpu.SwmGpu(s).init_from_setup()
    g.set_kernel_shape(shape)
    o = oracle.Oracle(s)
    for n in (1, 4, 23):
        g.step(n - g.itt)
        o.step(n - o.itt)
        for gn, on in STATE:
            assert same(g.get(gn), o.get(on)), f"shape {shape} itt={n} {gn}"
        a, b = g.get("CH", 0), o.get("CH", 0)
        assert same(a, b), f"shape {shape} itt={n} CH"
    g.close()
    o.close()
