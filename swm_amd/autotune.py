"""Autotuner: picks launch parameters for the workload at hand by timing them
on the real data.

* step-kernel shape of a context (LDS tiles 64x4 / 32x8 / 64x8 / 128x4 /
  32x16, marching wavefronts, persistent tiles): swm_gpu_tune, which runs each
  candidate from a device snapshot and restores the state exactly;
* tile shape of a batch (64x4 / 32x8): swm_gpu_batch_tune, likewise;
* SOR convergence-test cadence of the fixed-omega velocity potential: each
  candidate cadence solves the same right-hand side from the same start for a
  fixed number of sweeps; the cadence with the least time per sweep wins.

All candidates give the same FP64 step results bit for bit; the SOR cadence
only changes how often the residual is reduced, so iterations-to-tolerance are
reported to a multiple of the cadence."""
from __future__ import annotations

import time
from typing import Sequence

from . import gpu


def tune_context(m: "gpu.SwmGpu", steps: int = 20) -> dict:
    best, times = m.tune(steps)
    return {"best_shape": best, "ms_per_step": times}


def tune_batch(b: "gpu.SwmBatch", steps: int = 10) -> dict:
    best, times = b.tune(steps)
    return {"best_shape": best, "ms_per_step": times}


def tune_sor(m: "gpu.SwmGpu", omega: float = 1.9, fp32_sweeps: bool = False,
             cadences: Sequence[int] = (1, 2, 4, 8, 16), sweeps: int = 256) -> dict:
    """Time `sweeps` fixed-omega sweeps (tolerance 0, so none stops early) per
    cadence from a zero start; selects the cadence with the least time."""
    res = {}
    for k in cadences:
        m.set_sor_check_every(k)
        m.velocity_potential(omega, 0.0, sweeps, fp32_sweeps=fp32_sweeps, reset=True)   # warm caches
        t0 = time.perf_counter()
        m.velocity_potential(omega, 0.0, sweeps, fp32_sweeps=fp32_sweeps, reset=True)
        res[k] = (time.perf_counter() - t0) / sweeps * 1e3
    best = min(res, key=res.get)
    m.set_sor_check_every(best)
    return {"best_check_every": best, "ms_per_sweep": res}
