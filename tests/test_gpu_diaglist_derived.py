"""The device's derived register fields (swm_gpu_sync) against the reference's
recorded test_diaglist run: the GPU twin of test_oracle_diaglist_derived.py.
A LINEARISED_MEAN_STATE context is handed the recorded (u, v, eta) and the
basic state of the recording (o_H * Psi); SWM_RELVORT, SWM_MU, SWM_MV and
SWM_D equal the recording bit for bit, SWM_POT off the f = 0 row and SWM_EDENS
off the coast too, and every field equals the oracle's bit for bit."""
import numpy as np
import pytest

import oracle
from conftest import coastal_eta, diaglist_recorded_setup, same
from swm_amd import gpu

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("snap", [0, 1, 2])
def test_gpu_derived_fields_match_recording(gpu_lib, snap):
    s, der = diaglist_recorded_setup(snap, mask_psi=True)
    g = gpu.SwmGpu(s).init_from_setup()
    o = oracle.Oracle(s)
    mv = float(der["missing"])
    f0 = np.broadcast_to(s.grids["H"].f[:, None] == 0.0, s.shape)
    inner = ~coastal_eta(s)
    for gn, on, rn, where in (("SWM_RELVORT", "zeta", "swm_relvort", None), ("SWM_MU", "MU", "swm_mu", None),
                              ("SWM_MV", "MV", "swm_mv", None), ("SWM_D", "D", "swm_d", None),
                              ("SWM_POT", "Pot", "swm_pot", ~f0), ("SWM_EDENS", "EDens", "swm_edens", inner)):
        a = g.get(gn)
        assert same(a, o.get(on)), gn
        rec = der[rn][snap]
        m = rec != mv
        if where is not None:
            m = m & where
        assert same(a[m], rec[m]), gn
    g.close()
    o.close()
