import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def load_golden(name):
    return np.load(os.path.join(GOLDEN, name), allow_pickle=False)


def linear_osc_setup(npz):
    """Setup of test_fixed_length_output / test_oscfrc (LINEARISED_STATE_OF_REST,
    BAROTROPIC, oscillating zonal forcing) from a golden fixture: the recorded
    input fields and the experiment's numeric namelist values."""
    from swm_amd import host
    p = json.loads(str(npz["params"]))
    Hin = npz["H"].astype(np.float64)
    miss = (npz["H"] == npz["H_missing"]).astype(np.int8)
    grids, dL, dT = host.init_domain(
        int(p["domain_nl.nx"]), int(p["domain_nl.ny"]), float(p["domain_nl.lon_s"]), float(p["domain_nl.lon_e"]),
        float(p["domain_nl.lat_s"]), float(p["domain_nl.lat_e"]), Hin, miss, float(p["domain_nl.h_overwrite"]),
        0.0, 3, float(p["domain_nl.theta0"]), float(p["domain_nl.omega"]), float(p["domain_nl.a"]))
    s = host.Setup(Nx=int(p["domain_nl.nx"]), Ny=int(p["domain_nl.ny"]), variant=host.LINEARISED_STATE_OF_REST,
                   grids=grids, dLambda=dL, dTheta=dT, barotropic=True, Ah=float(p["model_nl.ah"]),
                   dt=float(p["model_nl.dt"]), A=float(p["domain_nl.a"]), RHO0=float(p["domain_nl.rho0"]), lbc=0.0)
    s.Nt = int(float(p["model_nl.run_length"]) / s.dt)
    host.apply_linear_D_alias(s)
    host.init_damping(s, [])
    host.add_oscillation(s, host.COMP_X, npz["taux_re"].astype(np.float64), npz["taux_im"].astype(np.float64),
                         float(p["swm_forcing_nl0.omega"]))
    return host.finalize(s)


def diaglist_grid_setup(npz, variant=None):
    """The test_diaglist domain (file bathymetry, H_overwrite = 3, BAROTROPIC)
    with the linear D alias applied to the grid thicknesses."""
    from swm_amd import host
    variant = host.LINEARISED_STATE_OF_REST if variant is None else variant
    p = json.loads(str(npz["params"]))
    Hin = npz["H"].astype(np.float64)
    miss = (npz["H"] == npz["H_missing"]).astype(np.int8)
    grids, dL, dT = host.init_domain(int(p["domain_nl.nx"]), int(p["domain_nl.ny"]), p["domain_nl.lon_s"],
                                     p["domain_nl.lon_e"], p["domain_nl.lat_s"], p["domain_nl.lat_e"], Hin, miss,
                                     p["domain_nl.h_overwrite"], 0.0, 3, p["domain_nl.theta0"], p["domain_nl.omega"],
                                     p["domain_nl.a"])
    s = host.Setup(Nx=int(p["domain_nl.nx"]), Ny=int(p["domain_nl.ny"]), variant=variant, grids=grids,
                   dLambda=dL, dTheta=dT, dt=float(p["model_nl.dt"]), A=float(p["domain_nl.a"]), lbc=0.0,
                   barotropic=True)
    host.apply_linear_D_alias(s)
    return host.finalize(s)


def diaglist_recorded_setup(snapshot, mask_psi=True):
    """The test_diaglist experiment (LINEARISED_MEAN_STATE, BAROTROPIC, free
    slip, file bathymetry with H_overwrite = 3, spherical f) initialised from
    one of the recorded snapshots of u, v, eta (0, 1, 2 = output records 1, 5
    and 10), with the basic state from the recorded input streamfunction
    (diaglist_derived.npz).  mask_psi: o_H * Psi, as the recording was made."""
    from swm_amd import host
    npz = load_golden("diaglist_psi.npz")
    der = load_golden("diaglist_derived.npz")
    p = json.loads(str(npz["params"]))
    Hin = npz["H"].astype(np.float64)
    miss = (npz["H"] == npz["H_missing"]).astype(np.int8)
    grids, dL, dT = host.init_domain(int(p["domain_nl.nx"]), int(p["domain_nl.ny"]), p["domain_nl.lon_s"],
                                     p["domain_nl.lon_e"], p["domain_nl.lat_s"], p["domain_nl.lat_e"], Hin, miss,
                                     p["domain_nl.h_overwrite"], 0.0, 3, p["domain_nl.theta0"], p["domain_nl.omega"],
                                     p["domain_nl.a"])
    s = host.Setup(Nx=int(p["domain_nl.nx"]), Ny=int(p["domain_nl.ny"]), variant=host.LINEARISED_MEAN_STATE,
                   grids=grids, dLambda=dL, dTheta=dT, dt=float(p["model_nl.dt"]), A=float(p["domain_nl.a"]),
                   lbc=0.0, barotropic=True, lateral_mixing=True, Ah=float(p["model_nl.ah"]))
    host.set_mean_state(s, der["psi_bs"], mask_psi=mask_psi)
    host.apply_linear_D_alias(s)
    host.init_damping(s, [])
    mv = float(der["missing"])
    for name, key in (("u0", "u"), ("v0", "v"), ("eta0", "eta")):
        a = der[key][snapshot].copy()
        a[a == mv] = 0.0
        setattr(s, name, a)
    return host.finalize(s), der


def coastal_eta(s):
    """Ocean eta points with a land H corner or a land u/v face."""
    ip = lambda a: np.roll(a, -1, axis=1)
    jp = lambda a: np.roll(a, -1, axis=0)
    oH, ou, ov = s.grids["H"].ocean, s.grids["u"].ocean, s.grids["v"].ocean
    c = (oH == 0) | (ip(oH) == 0) | (jp(oH) == 0) | (ip(jp(oH)) == 0)
    c |= (ou == 0) | (ip(ou) == 0) | (ov == 0) | (jp(ov) == 0)
    return c & (s.grids["eta"].ocean == 1)


def same(a, b):
    """Bitwise-equal values (NaN positions must agree; +0 and -0 compare equal)."""
    na, nb = np.isnan(a), np.isnan(b)
    return bool(np.array_equal(na, nb) and np.array_equal(a[~na], b[~nb]))


def gpu_available():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu_lib():
    if not gpu_available():
        pytest.skip("no GPU")
    from swm_amd import gpu
    return gpu.load_library()
