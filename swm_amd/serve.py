"""Serving: continuous batching of shallow-water runs on one GPU.

Requests (a basin, its forcing and parameters, a step count) arrive over time.
The server keeps one batch (swm_gpu_batch_*) resident on the device and
advances it in chunks of steps.  Between chunks it
  * retires members whose own step count has run out (their results are read
    back and their device memory freed), and
  * admits waiting requests in arrival order while their device footprint fits
    the HBM budget; an admitted request joins the running batch, so requests
    that arrive while others are on the GPU share its launches.
A member advances exactly its own number of steps (the batch kernel skips
members with no steps left), and the batch kernel runs the same arithmetic
as a context stepped alone, so each result is the same alone and in the mix.

Host-side setup of arriving requests (grids, masks, forcing), their admission
and their device upload run on a producer thread while the GPU works on the
current chunk.
"""
from __future__ import annotations

import dataclasses
import math
import queue
import threading
import time
from typing import Callable, List, Optional

import numpy as np

from . import configs, gpu


@dataclasses.dataclass
class Request:
    rid: int
    arrival: float            # seconds after the start of the stream
    Nx: int
    Ny: int
    nsteps: int
    H0: float
    Ah: float
    tau0: float
    tracer: bool
    seed: int
    keep: bool = False        # keep the final state for inspection
    # filled in by the server
    t_ready: float = math.nan
    t_admit: float = math.nan
    t_done: float = math.nan
    result: Optional[dict] = None
    max_batch: int = 0        # most members that shared one launch with this request

    @property
    def cells(self) -> int:
        return self.Nx * self.Ny

    @property
    def latency(self) -> float:
        return self.t_done - self.arrival


def request_stream(seed: int = 0, rate: float = 20.0, duration: float = 600.0, min_cells=(64, 64),
                   max_cells=(2048, 1024), steps=(500, 20000), tracer_p: float = 0.5,
                   keep_every: int = 0) -> List[Request]:
    """Poisson arrivals at `rate` per second over `duration` seconds.  Grid
    sizes: Nx and Ny log-uniform in [64, 2048] and [64, 1024], rounded to
    multiples of 16; length log-uniform in [500, 20000] steps; H, Ah and wind
    amplitude as in the batched ensemble; a tracer with probability tracer_p."""
    rng = np.random.default_rng(seed)
    out, t, k = [], 0.0, 0
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        lg = lambda lo, hi: math.exp(rng.uniform(math.log(lo), math.log(hi)))
        Nx = int(round(lg(min_cells[0], max_cells[0]) / 16.0)) * 16
        Ny = int(round(lg(min_cells[1], max_cells[1]) / 16.0)) * 16
        out.append(Request(rid=k, arrival=t, Nx=max(16, Nx), Ny=max(16, Ny), nsteps=int(round(lg(*steps))),
                           H0=float(rng.uniform(200.0, 4000.0)), Ah=float(rng.uniform(20.0, 500.0)),
                           tau0=float(rng.uniform(0.02, 0.2)), tracer=bool(rng.random() < tracer_p),
                           seed=seed * 1000003 + k, keep=bool(keep_every and k % keep_every == 0)))
        k += 1
    return out


def build_setup(r: Request):
    return configs.serve_member(seed=r.seed, Nx=r.Nx, Ny=r.Ny, H0=r.H0, Ah=r.Ah, tau0=r.tau0, tracer=r.tracer,
                                nsteps=r.nsteps)


def footprint_bytes(Nx: int, Ny: int, ntracers: int, precision: int) -> int:
    """Device memory of one context: the state planes (u, v, eta x2, 9 AB slots,
    3 low-part planes, 7 per tracer), one scratch plane, the mask and the
    per-field static arrays (an upper bound: every static field stored in
    full)."""
    esz = 4 if precision == gpu.FP32 else 8
    align = 128 // esz
    P = (Nx + align - 1) // align * align
    plane = P * Ny * esz
    statics = 19 + 3 * ntracers
    return plane * (18 + 7 * ntracers + 1 + statics) + P * Ny + (1 << 20)


class Server:
    """Continuous-batching request loop over one SwmBatch.

    A producer thread takes requests in arrival order, builds their host-side
    setup, reserves their device footprint against the HBM budget (waiting
    for retirements when it does not fit, first come first served), creates
    and uploads their context, and hands it over.  The main thread attaches
    handed-over contexts to the running batch between chunks, advances the
    batch, and retires finished members.  Context creation and upload thus
    overlap the GPU's work on the current chunk."""

    def __init__(self, precision: int = gpu.FP32, hbm_budget: float = 200e9, chunk: int = 64,
                 builder: Callable = build_setup, max_ready: int = 64):
        self.precision = precision
        self.budget = float(hbm_budget)
        self.chunk = int(chunk)
        self.builder = builder
        self.max_ready = max_ready
        self.used = 0.0
        self.cv = threading.Condition()
        self.chunks = 0
        self.max_batch = 0        # most members advanced by one launch
        self.gpu_ms = 0.0

    def _producer(self, reqs, ready: "queue.Queue", t0: float, realtime: bool, stop: threading.Event):
        try:
            for r in reqs:
                if stop.is_set():
                    return
                if realtime:
                    wait = t0 + r.arrival - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                s = self.builder(r)
                r.t_ready = time.perf_counter() - t0
                fp = footprint_bytes(s.Nx, s.Ny, len(s.tracers), self.precision)
                with self.cv:   # admission: first come first served within the budget
                    while self.used > 0 and self.used + fp > self.budget and not stop.is_set():
                        self.cv.wait(timeout=0.5)
                    self.used += fp
                m = gpu.SwmGpu(s, self.precision).init_from_setup()
                ready.put((r, m, fp))
        except BaseException as e:   # surfaced by the main loop
            ready.put(e)
            return
        ready.put(None)

    def _release(self, fp: float):
        with self.cv:
            self.used -= fp
            self.cv.notify_all()

    def run(self, requests: List[Request], realtime: bool = True, progress: Optional[Callable[[dict], None]] = None,
            progress_every: float = 30.0) -> dict:
        """Serve every request; returns summary statistics.  With realtime
        False all requests are available at once (arrival order kept).
        progress, if given, receives a status dict every progress_every
        seconds of wall time (requests done and active, chunks, GPU time)."""
        reqs = sorted(requests, key=lambda r: r.arrival)
        ready: "queue.Queue" = queue.Queue(maxsize=self.max_ready)
        stop = threading.Event()
        t0 = time.perf_counter()
        prod = threading.Thread(target=self._producer, args=(reqs, ready, t0, realtime, stop), daemon=True)
        prod.start()
        batch = gpu.SwmBatch()
        active = {}             # id(ctx) -> (request, ctx, footprint)
        produced_all = False
        cell_steps = 0
        done = 0
        t_report = progress_every
        try:
            while True:
                # attach every handed-over context (block only when the GPU is idle)
                block = not active and not produced_all
                while True:
                    try:
                        item = ready.get(timeout=0.5) if block else ready.get_nowait()
                    except queue.Empty:
                        break
                    block = False
                    if item is None:
                        produced_all = True
                        break
                    if isinstance(item, BaseException):
                        raise item
                    r, m, fp = item
                    batch.attach(m, r.nsteps)
                    r.t_admit = time.perf_counter() - t0
                    active[id(m)] = (r, m, fp)
                if not active:
                    if produced_all:
                        break
                    continue
                batch.step(self.chunk)
                self.chunks += 1
                self.max_batch = max(self.max_batch, len(active))
                for r, _, _ in active.values():
                    r.max_batch = max(r.max_batch, len(active))
                self.gpu_ms += batch.last_ms()
                now = time.perf_counter() - t0
                for key in list(active):
                    r, m, fp = active[key]
                    if batch.remaining(m) == 0:
                        if r.keep:
                            r.result = {"SWM_ETA": m.get("SWM_ETA"), "SWM_U": m.get("SWM_U"),
                                        "SWM_V": m.get("SWM_V"), "itt": m.itt}
                            if r.tracer:
                                r.result["CH"] = m.get("CH", 0)
                        batch.detach(m)
                        m.close()
                        self._release(fp)
                        r.t_done = now
                        cell_steps += r.cells * r.nsteps
                        done += 1
                        del active[key]
                if progress and now >= t_report:
                    t_report = now + progress_every
                    progress({"t_s": now, "done": done, "active": len(active), "requests": len(reqs),
                              "chunks": self.chunks, "gpu_busy_s": self.gpu_ms * 1e-3})
        finally:
            stop.set()
            with self.cv:
                self.cv.notify_all()
            prod.join(timeout=5.0)
            for r, m, fp in active.values():
                m.close()
            while True:
                try:
                    item = ready.get_nowait()
                except queue.Empty:
                    break
                if isinstance(item, tuple):
                    item[1].close()
            batch.close()
        wall = time.perf_counter() - t0
        lat = np.array([r.latency for r in reqs])
        return {"requests": len(reqs), "wall_s": wall, "cell_steps": cell_steps,
                "cell_updates_per_s": cell_steps / wall if wall > 0 else 0.0,
                "gpu_busy_s": self.gpu_ms * 1e-3, "chunks": self.chunks, "max_members_per_launch": self.max_batch,
                "latency_p50_s": float(np.percentile(lat, 50)), "latency_p95_s": float(np.percentile(lat, 95)),
                "latency_p99_s": float(np.percentile(lat, 99)), "latency_max_s": float(lat.max())}
