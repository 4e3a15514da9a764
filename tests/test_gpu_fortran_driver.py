"""The Fortran drop-in run end to end: swm_amd/fortran/swm_gpu_driver (the
ISO_C_BINDING shim plus a time loop in the order of src/model.f90:153-197,
built by __graft_entry__.build with amdflang and linked to libswmgpu.so) runs
create, upload_state, add_tracer, step and sync on a 96 x 96 FULLY_NONLINEAR
basin with mixing, windstress and a tracer.  Its output fields equal the
Python binding's bit for bit at every output step, in FP64 and FP32."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, same
from swm_amd import configs, gpu

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(ROOT, "swm_amd", "fortran", "swm_gpu_driver")


def write_input(path, s, precision, Nt, out_every):
    u, v, eta = gpu.initial_state(s)
    D = gpu.initial_D(s, eta)
    taux = None
    for comp, tau in s.wind:
        assert comp == 0
        taux = tau
    with open(path, "wb") as fh:
        np.array([s.Nx, s.Ny, s.variant, gpu._flags(s), precision, Nt, out_every, len(s.tracers),
                  int(taux is not None)], dtype=np.int32).tofile(fh)
        np.array([s.G, s.Ah, s.dt, s.A, s.dLambda, s.dTheta, s.RHO0, s.minD, s.lbc, s.ab_chi,
                  s.tau_scale if s.tau_scale is not None else 0.0], dtype=np.float64).tofile(fh)
        for name in ("H", "u", "v", "eta"):
            g = s.grids[name]
            np.ascontiguousarray(g.ocean, np.int8).tofile(fh)
            for a in (g.H, g.cos_lat, g.tan_lat, g.f):
                np.ascontiguousarray(a, np.float64).tofile(fh)
        for a in (s.impl_u, s.impl_v, s.impl_eta, s.F_x_const, s.F_y_const, s.F_eta_const):
            np.ascontiguousarray(a, np.float64).tofile(fh)
        if taux is not None:
            np.ascontiguousarray(taux, np.float64).tofile(fh)
        for a in (u, v, eta):
            np.ascontiguousarray(a, np.float64).tofile(fh)
        for t in s.tracers:
            np.array([t["kappa_h"]], np.float64).tofile(fh)
            for a in (D * t["C"], t["C"], t["C0"], t["gamma_C"], 1.0 + s.dt * t["cons"]):
                np.ascontiguousarray(a, np.float64).tofile(fh)


@pytest.mark.parametrize("precision", [gpu.FP64, gpu.FP32])
def test_fortran_driver_equals_python_binding(gpu_lib, tmp_path, precision):
    if not os.path.exists(DRIVER):
        pytest.skip("the Fortran driver was not built (no amdflang where build() ran)")
    s = configs.c2_nonlinear(seed=3, N=96, dt=10.0)
    Nt, every = 40, 8
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    write_input(fin, s, precision, Nt, every)
    r = subprocess.run([DRIVER, fin, fout], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(fout, "rb").read()
    n = s.Nx * s.Ny
    per = 8 + 8 * n * (3 + len(s.tracers))
    assert len(raw) == per * (Nt // every)
    g = gpu.SwmGpu(s, precision).init_from_setup()
    for k in range(Nt // every):
        rec = raw[k * per:(k + 1) * per]
        itt = int(np.frombuffer(rec[:8], np.int64)[0])
        fields = np.frombuffer(rec[8:], np.float64).reshape(3 + len(s.tracers), s.Ny, s.Nx)
        g.step(itt - g.itt)
        assert itt == (k + 1) * every == g.itt
        for f, name in zip(fields, ("SWM_U", "SWM_V", "SWM_ETA")):
            assert same(f, g.get(name)), (itt, name)
        for t in range(len(s.tracers)):
            assert same(fields[3 + t], g.get("CH", t)), (itt, "CH")
    g.close()
