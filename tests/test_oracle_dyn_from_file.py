"""Pins of the oracle's DYNFROMFILE restatement (dynFromFile_module.f90:56-160,
model.f90:108-113, 238-259) that follow from the mathematics:

* the master fields u, v, eta after model_advance are the SWM fields plus the
  grid's ocean mask times the dataset linearly interpolated at itt*dt, with
  the dataset repeating after its last record (memchunk_module.f90:181-207):
  closed form checked at record times, between records, across chunk reloads
  and past the end of the dataset;
* the tracer is advected by the master flow, not by SWM's: with SWM at rest
  and a uniform dataset velocity U the advective tendency of a mode has the
  exact discrete symbol -U h sin(k)/dx of the shifted mode, in x and in y.
"""
import math

import numpy as np
import pytest

import oracle
from swm_amd import configs, host

from test_oracle_physics_pins import channel, interior_rows, tracer_channel


def interp_cyclic(rec, rec_dt, t):
    nrec = rec.shape[0]
    if nrec == 1:
        return rec[0]
    n = int(math.floor(t / rec_dt + 1e-9))
    a = t - n * rec_dt
    r0, r1 = rec[n % nrec], rec[(n + 1) % nrec]
    return r0 + a * (r1 - r0) / rec_dt


def flow_with_datasets(chunk, nrec=3, seed=5):
    s = configs.c2_nonlinear(seed=seed, N=48, dt=10.0)
    rng = np.random.default_rng(seed)
    recs = {}
    for comp, amp in ((host.COMP_X, 0.05), (host.COMP_Y, 0.04), (host.COMP_ETA, 0.1)):
        recs[comp] = amp * rng.standard_normal((nrec,) + s.shape)
        host.add_dyn_from_file(s, comp, recs[comp], rec_dt=2.5 * s.dt, chunk_size=chunk)
    return host.finalize(s), recs


@pytest.mark.parametrize("chunk", [100, 2])
def test_master_fields_are_swm_plus_interpolated_dataset(chunk):
    s, recs = flow_with_datasets(chunk)
    o = oracle.Oracle(s)
    grid = {host.COMP_X: ("U", "u", "u"), host.COMP_Y: ("V", "v", "v"), host.COMP_ETA: ("ETA", "eta", "eta")}
    nrec, rec_dt = 3, 2.5 * s.dt
    # 0 and 5 fall on records, 2, 6 and 11 between, 8 and 19 past the end of the
    # dataset (period 3 records = 7.5 steps)
    for itt in (0, 2, 5, 6, 8, 11, 19):
        o.step(itt - o.itt)
        t = itt * s.dt
        for comp, (master, swm, g) in grid.items():
            want = s.grids[g].ocean * interp_cyclic(recs[comp], rec_dt, t) + o.get(swm)
            got = o.get(master)
            assert np.abs(got - want).max() <= 1e-14 * np.abs(want).max(), (itt, master)
            assert np.all(got[s.grids[g].ocean == 0] == o.get(swm)[s.grids[g].ocean == 0])
    o.close()


@pytest.mark.parametrize("direction", ["x", "y"])
def test_tracer_advected_by_the_dataset_flow(direction):
    H0, U = 100.0, 0.3
    s0 = channel()
    i = np.arange(s0.Nx)[None, :] + 0.0 * np.arange(s0.Ny)[:, None]
    j = np.arange(s0.Ny)[:, None] + 0.0 * np.arange(s0.Nx)[None, :]
    if direction == "x":
        k = 2.0 * math.pi * 5 / s0.Nx
        C = np.sin(k * i)
        s = tracer_channel(C, H0=H0)
        host.add_dyn_from_file(s, host.COMP_X, np.full(s.shape, U), rec_dt=100.0)
        want = -U * H0 * math.sin(k) / (s.A * s.dLambda) * np.cos(k * i)
    else:
        k = 2.0 * math.pi * 3 / 40.0
        C = np.sin(k * j)
        s = tracer_channel(C, H0=H0)
        host.add_dyn_from_file(s, host.COMP_Y, np.full(s.shape, U), rec_dt=100.0)
        want = -U * H0 * math.sin(k) / (s.A * s.dTheta) * np.cos(k * j)
    o = oracle.Oracle(s)
    o.step(1)
    G = o.get("G_CH", 0)
    assert np.abs(o.get("u")).max() == 0.0 and np.abs(o.get("v")).max() == 0.0   # SWM stays at rest
    o.close()
    ok = interior_rows(s, 2)
    assert np.abs(G[ok] - want[ok]).max() <= 1e-10 * np.abs(want[ok]).max()


def test_no_dataset_leaves_the_tracer_at_rest():
    """The same channel without the dataset: no flow, no tendency (guards the
    symbol test against a tendency that does not come from the dataset)."""
    s0 = channel()
    C = np.sin(2.0 * math.pi * 5 / s0.Nx * np.arange(s0.Nx))[None, :] * np.ones((s0.Ny, 1))
    s = tracer_channel(C, H0=100.0)
    o = oracle.Oracle(s)
    o.step(1)
    G = o.get("G_CH", 0)
    o.close()
    assert np.abs(G[interior_rows(s, 2)]).max() == 0.0
