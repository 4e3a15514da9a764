"""Autotuner: picks launch parameters for the workload at hand by timing them
on the real data, and remembers them per device and workload.

Search spaces (every candidate gives the same results bit for bit):

* a context's step: kernel shape (LDS tiles 64x4 / 32x8 / 64x8 / 128x4 /
  32x16, marching wavefronts, persistent tiles, late-history tiles) and launch
  form (replay of captured six-step graphs, or direct launches) -
  swm_gpu_tune runs each candidate from a device snapshot and restores the
  state exactly;
* a batch's step: tile (64x4 / 32x8) x launch form (one block per tile, or a
  persistent grid draining per-XCD tile queues) - swm_gpu_batch_tune, likewise;
* the fixed-omega SOR: halo width of the blocked sweeps (0 = one launch per
  colour; 2, 4, 8 cells = 1, 2, 4 sweeps per launch, each sweep consuming two
  halo cells) x convergence-test cadence (1..16 sweeps) x rows of the 64-wide
  LDS tile (16 or 32: fewer halo cells per output, more LDS), where the sweeps
  per launch must divide the cadence and the 8-cell halo takes 16-row tiles.
  The candidates are raced by successive halving:
  every survivor solves the same right-hand side from the same zero start for
  a fixed number of sweeps (tolerance 0, so none stops early); the slower half
  is dropped and the sweep count doubled until one pair is left.  The cadence
  changes only where the residual is tested, so iterations-to-tolerance are
  reported to a multiple of it.

Results go into a TuneCache (JSON, keyed by device name and a workload
signature: grid, precision, variant, flags, tracers), so a served request
stream or a restarted job tunes each distinct workload once.  The cache file
is the one given, or $SWM_TUNE_CACHE, or none (in memory only)."""
from __future__ import annotations

import json
import os
import time
from typing import Iterable, Optional, Sequence

from . import gpu


def device_name() -> str:
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.get_device_name(0)
    except Exception:   # no torch or no device: one shared key
        pass
    return "device"


def signature(m: "gpu.SwmGpu") -> str:
    """Workload key of a context: what the kernels' cost depends on."""
    s = m.s
    return (f"{s.Nx}x{s.Ny}/p{m.precision}/v{s.variant}/f{gpu._flags(s)}/t{len(s.tracers)}"
            f"/s{len(getattr(s, 'streams', []))}/d{len(getattr(s, 'dyn', []))}")


def batch_signature(members: Sequence["gpu.SwmGpu"]) -> str:
    """Workload key of a batch: the sorted member signatures."""
    return "batch[" + ",".join(sorted(signature(m) for m in members)) + "]"


class TuneCache:
    """Tuning results per (device, workload signature), optionally kept in a
    JSON file (read at construction, written after each new entry)."""

    def __init__(self, path: Optional[str] = None):
        self.path = path if path is not None else os.environ.get("SWM_TUNE_CACHE")
        self.data = {}
        if self.path and os.path.exists(self.path):
            with open(self.path) as fh:
                self.data = json.load(fh)
        self.device = device_name()

    def get(self, kind: str, key: str):
        return self.data.get(self.device, {}).get(kind, {}).get(key)

    def put(self, kind: str, key: str, value: dict):
        self.data.setdefault(self.device, {}).setdefault(kind, {})[key] = value
        if self.path:
            tmp = self.path + ".tmp"
            with open(tmp, "w") as fh:
                json.dump(self.data, fh, indent=1, sort_keys=True)
            os.replace(tmp, self.path)


def tune_context(m: "gpu.SwmGpu", steps: int = 24, cache: Optional[TuneCache] = None) -> dict:
    """Kernel shape and launch form of a context's step (state restored)."""
    key = signature(m)
    hit = cache.get("context", key) if cache else None
    if hit is not None:
        m.set_kernel_shape(hit["best_shape"])
        m.set_launch_mode(hit["graphs"])
        return dict(hit, cached=True)
    best, times = m.tune(steps)
    graphs, ms_direct, ms_graph = m.launch_mode()
    res = {"best_shape": best, "ms_per_step": times, "graphs": graphs, "ms_direct": ms_direct,
           "ms_graph": ms_graph}
    if cache:
        cache.put("context", key, res)
    return dict(res, cached=False)


def tune_batch(b: "gpu.SwmBatch", steps: int = 10, cache: Optional[TuneCache] = None,
               members: Optional[Sequence["gpu.SwmGpu"]] = None) -> dict:
    """Tile and launch form of a batch's step (members restored).  With a
    cache, `members` (the attached contexts) give the workload key."""
    key = batch_signature(members) if (cache and members) else None
    hit = cache.get("batch", key) if key else None
    if hit is not None:
        b.set_shape(hit["best_shape"])
        return dict(hit, cached=True)
    best, times = b.tune(steps)
    res = {"best_shape": best, "ms_per_step": times}
    if key:
        cache.put("batch", key, res)
    return dict(res, cached=False)


def sor_candidates(halos: Iterable[int] = (0, 2, 4, 8), cadences: Iterable[int] = (1, 2, 4, 8, 16),
                   rows: Iterable[int] = (16, 32)):
    """(halo, cadence, rows) whose sweeps per launch (halo / 2) divide the
    cadence; 32-row tiles only for blocked sweeps of at most 2 per launch (the
    unblocked colour launches have no tile)."""
    return [(h, c, r) for h in halos for c in cadences for r in rows
            if (h <= 2 or c % (h // 2) == 0) and (r == 16 or 0 < h <= 4)]


def apply_sor(m: "gpu.SwmGpu", halo: int, cadence: int, rows: int = 16):
    m.set_sor_halo(-1)            # no setting may contradict another in between
    m.set_sor_tile_rows(rows)
    m.set_sor_check_every(cadence)
    m.set_sor_halo(halo)


def tune_sor(m: "gpu.SwmGpu", omega: float = 1.9, fp32_sweeps: bool = False,
             halos: Sequence[int] = (0, 2, 4, 8), cadences: Sequence[int] = (1, 2, 4, 8, 16),
             rows: Sequence[int] = (16, 32), sweeps: int = 64, cache: Optional[TuneCache] = None) -> dict:
    """Successive halving over (halo, cadence, rows): time `sweeps`
    fixed-omega sweeps of every candidate, keep the faster half, double the
    sweeps, until one is left; applies and returns it."""
    key = f"{signature(m)}/w{omega}/f{int(bool(fp32_sweeps))}"
    hit = cache.get("sor", key) if cache else None
    if hit is not None:
        apply_sor(m, hit["halo"], hit["check_every"], hit.get("rows", 16))
        return dict(hit, cached=True)
    cand = sor_candidates(halos, cadences, rows)
    rounds = []
    n = sweeps
    while len(cand) > 1:
        times = {}
        for h, c, r in cand:
            apply_sor(m, h, c, r)
            m.velocity_potential(omega, 0.0, n, fp32_sweeps=fp32_sweeps, reset=True)   # warm caches
            t0 = time.perf_counter()
            m.velocity_potential(omega, 0.0, n, fp32_sweeps=fp32_sweeps, reset=True)
            times[(h, c, r)] = (time.perf_counter() - t0) / n * 1e3
        rounds.append({"sweeps": n, "ms_per_sweep": {f"halo{h}/every{c}/rows{r}": t for (h, c, r), t in times.items()}})
        cand = sorted(cand, key=times.get)[:max(1, len(cand) // 2)]
        n *= 2
    (h, c, r), = cand
    apply_sor(m, h, c, r)
    res = {"halo": h, "check_every": c, "rows": r, "sweeps_per_launch": h // 2, "rounds": rounds}
    if cache:
        cache.put("sor", key, res)
    return dict(res, cached=False)
