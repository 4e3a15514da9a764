"""Extract golden vectors from the reference's selfcheck experiments.

Reads netCDF data files under /root/reference/selfcheck (inputs and the
reference model's recorded outputs) and the numeric run parameters of each
experiment's namelist, and writes small .npz fixtures next to this script.
Only data is stored: input fields, recorded outputs, output times and the
scalar run parameters.  Run from the repository root:

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SELF = "/root/reference/selfcheck"
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from swm_amd import namelist as nl  # noqa: E402


def read_var(path, name, all_records=False):
    from scipy.io import netcdf_file
    with netcdf_file(path, "r", mmap=False) as nc:
        v = nc.variables[name]
        data = np.array(v.data)
        attrs = {a: np.asarray(getattr(v, a)).astype(np.float64).item() for a in ("missing_value", "_FillValue")
                 if hasattr(v, a)}
        t = None
        for tn in ("TIME", "time"):
            if tn in nc.variables:
                t = np.array(nc.variables[tn].data, dtype=np.float64)
        return data, attrs, t


def params(exp):
    groups = nl.read_namelist_file(os.path.join(SELF, exp, "model.namelist"))
    keep = {}
    for g in ("model_nl", "domain_nl"):
        for k, v in (nl.groups(groups, g) or [{}])[0].items():
            if isinstance(v, (int, float)):
                keep[f"{g}.{k}"] = v
    for n, st in enumerate(nl.groups(groups, "swm_forcing_nl")):
        for k in ("omega",):
            if k in st:
                keep[f"swm_forcing_nl{n}.{k}"] = st[k]
    return json.dumps(keep, sort_keys=True)


def main():
    # test_fixed_length_output: LINEARISED_STATE_OF_REST, oscillating forcing, 346 steps
    exp = os.path.join(SELF, "test_fixed_length_output")
    H, Ha, _ = read_var(os.path.join(exp, "input/H_sc.nc"), "H")
    re, _, _ = read_var(os.path.join(exp, "input/taux_sc.nc"), "TAUX_REAL")
    im, _, _ = read_var(os.path.join(exp, "input/taux_sc.nc"), "TAUX_IMAG")
    u, ua, t = read_var(os.path.join(exp, "output/reference_000000000001_u_snapshot.nc"), "u")
    np.savez_compressed(os.path.join(HERE, "fixed_length_output.npz"),
                        H=H, H_missing=np.float32(Ha["missing_value"]), taux_re=re, taux_im=im,
                        u=u, u_missing=ua["missing_value"], time=t, params=params("test_fixed_length_output"))
    # test_oscfrc: F_x of the same forcing at 11 times over 34680 steps
    exp = os.path.join(SELF, "test_oscfrc")
    fx, fa, t = read_var(os.path.join(exp, "output/reference_000000000001_taux.nc"), "f_x")
    re2, _, _ = read_var(os.path.join(exp, "input/taux_sc.nc"), "TAUX_REAL")
    im2, _, _ = read_var(os.path.join(exp, "input/taux_sc.nc"), "TAUX_IMAG")
    H2, Ha2, _ = read_var(os.path.join(exp, "input/H_sc.nc"), "H")
    np.savez_compressed(os.path.join(HERE, "oscfrc.npz"), H=H2, H_missing=np.float32(Ha2["missing_value"]),
                        taux_re=re2, taux_im=im2, f_x=fx, f_x_missing=fa["missing_value"], time=t,
                        params=params("test_oscfrc"))
    # test_diaglist: initial-time static outputs (domain init, H overwrite through the D alias)
    exp = os.path.join(SELF, "test_diaglist")
    H3, Ha3, _ = read_var(os.path.join(exp, "input/H_sc.nc"), "H")
    out = {}
    for fname, var in (("H_out", "H"), ("impl_u", "impl_u"), ("impl_v", "impl_v"), ("impl_n", "impl_eta")):
        d, a, _ = read_var(os.path.join(exp, f"output/reference_000000000001_{fname}.nc"), var)
        out[fname] = d[0]
        out[fname + "_missing"] = a["missing_value"]
    np.savez_compressed(os.path.join(HERE, "diaglist_static.npz"), H=H3, H_missing=np.float32(Ha3["missing_value"]),
                        params=params("test_diaglist"), **out)
    # test_diaglist: recorded u, v and psi snapshots (psi is computeStreamfunction
    # of the recorded u, v; three of the eleven snapshots)
    snap = {}
    for fname, var in (("u_out", "u"), ("v_out", "v"), ("psi_out", "psi")):
        d, a, t = read_var(os.path.join(exp, f"output/reference_000000000001_{fname}.nc"), var)
        snap[var] = d[[1, 5, 10]]
        snap[var + "_missing"] = a["missing_value"]
        snap["time"] = t[[1, 5, 10]]
    np.savez_compressed(os.path.join(HERE, "diaglist_psi.npz"), H=H3, H_missing=np.float32(Ha3["missing_value"]),
                        params=params("test_diaglist"), **snap)
    # test_diaglist: recorded state and derived register fields at three of the
    # eleven snapshots, and the basic-state streamfunction input
    der = {}
    for fname, var in (("u_out", "u"), ("v_out", "v"), ("eta_out", "eta"), ("relvort", "swm_relvort"),
                       ("potvort", "swm_pot"), ("edens", "swm_edens"), ("mu", "swm_mu"), ("mv", "swm_mv"),
                       ("d", "swm_d")):
        d, a, t = read_var(os.path.join(exp, f"output/reference_000000000001_{fname}.nc"), var)
        der[var] = d[[1, 5, 10]]
        der["missing"] = a["missing_value"]
        der["time"] = t[[1, 5, 10]]
    bs, _, _ = read_var(os.path.join(exp, "input/bs_sc.nc"), "PSI")
    np.savez_compressed(os.path.join(HERE, "diaglist_derived.npz"), psi_bs=bs.astype(np.float64), **der)
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
