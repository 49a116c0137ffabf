"""Batched and served migration workloads over BatchEngine.

Two workloads, both seeded (numpy PCG64) so a run can be repeated exactly:

  crop_batch(...)     aperture crops of the 2301 x 751 faulted model: widths
                      drawn from U[400, 1200] cells, nz = 751, nt from
                      U[2000, 8000] at dt = 0.35 ms, orders 8 and 12 mixed;
                      one shot per crop, source at the crop centre, receivers
                      every 12 m over the crop.
  served_stream(...)  a Poisson stream of shot-migration requests, each drawn
                      uniformly from three resident models:
                        201 x 151 @ 10 m three-layer model, nt 1400, dt 1 ms,
                          12 Hz, 201 receivers, store 2;
                        1201 x 401 @ 8 m faulted model, nt 4000, dt 0.6 ms,
                          15 Hz, 401 receivers, store 4;
                        2301 x 751 @ 4 m faulted model, nt 8000, dt 0.35 ms,
                          15 Hz, 767 receivers, store 4;
                      all order 8, nb 60, FP32 fields with FP16 snapshots.

serve() drives a BatchEngine in real time: requests are admitted after their
arrival time passes (first come first served, within a fill limit of cells
per tick), the engine runs short chunks of ticks in between, and each
request's latency runs from its arrival to its image being on the host.
serial_plain() times the same requests on the plain per-shot engine, one
after the other, and derives the FIFO latency a single plain engine would
give for the same arrival times.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import BatchEngine, HipEngine, SNAP_F16, SNAP_F32, models
from .types import RTMParams, ShotRecord, VelocityModel


@dataclass
class Request:
    key: str                   # model name
    shot: ShotRecord
    arrival: float = 0.0       # seconds after the stream starts
    latency: float = -1.0
    image: np.ndarray | None = None
    illum: np.ndarray | None = None


@dataclass
class ModelSpec:
    model: VelocityModel
    params: RTMParams
    nt: int
    dt: float
    snapshot_format: int = SNAP_F16
    shots: list = field(default_factory=list)   # ShotRecords with modelled traces

    @property
    def cells_per_tick(self) -> float:
        """Extended cells one step of a job on this model updates."""
        return float((self.model.nx + 2 * self.params.nb) * (self.model.nz + 2 * self.params.nb))

    @property
    def cell_steps(self) -> float:
        return 2.0 * self.nt * self.cells_per_tick


# Admission limits of serve().  FILL_CELLS extended cells per tick: one
# 2301 x 751 job (2421 x 871 cells, already enough to spread over every CU)
# with a 1201 x 401 and a 201 x 151 job beside it, but never two 2301 x 751
# jobs, which run slower side by side than one after the other (their fields
# no longer fit the caches).  MAX_CONCURRENT jobs at once: the batch engine
# runs up to three jobs of which one fills the chip each on its own stream,
# where the smaller ones fill the large one's launch gaps.
FILL_CELLS = 3.0e6
MAX_CONCURRENT = 3


def served_models(nsources: int = 4, snapshot_format: int = SNAP_F16) -> dict[str, ModelSpec]:
    """The three resident models of the served stream, with `nsources`
    source positions each (traces modelled later by model_traces)."""
    specs = {}
    m = models.layered_3()
    specs["layered_201x151"] = ModelSpec(m, RTMParams(nb=60, order=8, store_interval=2, f0=12.0),
                                         1400, 0.001, snapshot_format)
    m = models.faulted_layers(1201, 401, 8.0, 8.0, seed=1201401)
    specs["faulted_1201x401"] = ModelSpec(m, RTMParams(nb=60, order=8, store_interval=4, f0=15.0),
                                          4000, 0.0006, snapshot_format)
    m = models.faulted_layers(2301, 751, 4.0, 4.0)
    specs["faulted_2301x751"] = ModelSpec(m, RTMParams(nb=60, order=8, store_interval=4, f0=15.0),
                                          8000, 0.00035, snapshot_format)
    rec_dx = {"layered_201x151": 10.0, "faulted_1201x401": 24.0, "faulted_2301x751": 12.0}
    for key, sp in specs.items():
        xmax = (sp.model.nx - 1) * sp.model.dx
        sxs = xmax * (np.arange(nsources) + 1) / (nsources + 1)
        sp.shots = models.surface_shots(sp.model, sp.nt, sp.dt, sxs, rec_dx=rec_dx[key])
    return specs


def model_traces(specs: dict[str, ModelSpec], device: int = 0) -> None:
    """Model every source's traces with the plain engine (forward modelling)."""
    for sp in specs.values():
        with HipEngine(device) as eng:
            eng.setup(sp.model, sp.params, sp.nt, sp.dt)
            for s in sp.shots:
                s.traces = eng.model_shot(s)


def served_stream(specs: dict[str, ModelSpec], rate: float, njobs: int, seed: int = 5) -> list[Request]:
    """Poisson arrivals at `rate` per second; model and source uniform."""
    rng = np.random.default_rng(seed)
    keys = sorted(specs)
    t = 0.0
    out = []
    for _ in range(njobs):
        t += rng.exponential(1.0 / rate)
        key = keys[rng.integers(len(keys))]
        shot = specs[key].shots[rng.integers(len(specs[key].shots))]
        out.append(Request(key, shot, t))
    return out


def serve(specs: dict[str, ModelSpec], requests: list[Request], device: int = 0, chunk_s: float = 0.004,
          use_graph: bool = True, memory_budget: int = 0, keep_images: bool = False, reserve: int = 0,
          fill_cells: float = FILL_CELLS) -> dict:
    """Runs the stream through one BatchEngine in real time; fills each
    request's latency (and image when keep_images).  reserve > 0 places all
    job memory in one block of that many bytes, reserved before the stream
    starts (-1: the whole budget).

    Admission is first come first served with a fill limit: an arrived
    request joins the running batch while the batch's extended cells per
    tick, its own included, stay within fill_cells and fewer than
    MAX_CONCURRENT requests run, or when nothing runs; otherwise it waits,
    and those behind it too.  One 2301 x 751 job already fills the chip, so
    large jobs run one after the other (processor sharing among them would
    delay every one of them to the end of the busy period), while smaller
    ones run beside them.  fill_cells <= 0 admits every request on arrival.
    Returns a summary."""
    with BatchEngine(device, memory_budget=memory_budget, use_graph=use_graph) as b:
        mids = {k: b.add_model(sp.model, sp.params, sp.nt, sp.dt, sp.snapshot_format) for k, sp in specs.items()}
        if reserve:
            b.reserve(max(reserve, 0))     # the server's job memory, placed before requests arrive
        pending = sorted(range(len(requests)), key=lambda i: requests[i].arrival)
        out: dict[int, tuple[int, float, float]] = {}   # job id -> (request index, submit lag, cells per tick)
        waiting: list[int] = []                          # arrived, not yet admitted, in arrival order
        nxt = 0
        chunk = 16
        t_ret, prev_ran = None, 0          # return time and ticks of the previous run() call
        t0 = time.perf_counter()
        while nxt < len(pending) or out or waiting:
            now = time.perf_counter() - t0
            while nxt < len(pending) and requests[pending[nxt]].arrival <= now:
                waiting.append(pending[nxt])
                nxt += 1
            while waiting:
                r = requests[waiting[0]]
                cells = specs[r.key].cells_per_tick
                busy = sum(c for _, _, c in out.values())
                if out and fill_cells > 0 and (busy + cells > fill_cells or len(out) >= MAX_CONCURRENT):
                    break
                jid = b.submit_shot(mids[r.key], r.shot)
                out[jid] = (waiting.pop(0), (time.perf_counter() - t0) - r.arrival, cells)
            if not out:
                t_ret = None
                time.sleep(max(0.0, requests[pending[nxt]].arrival - (time.perf_counter() - t0)))
                continue
            ticks0 = b.stats()["ticks"]
            b.run(max_ticks=chunk, until_done=False)
            c1 = time.perf_counter()
            ran = b.stats()["ticks"] - ticks0
            # run() returns once the previous call's ticks are done and leaves
            # its own queued, so the time between two returns is the device
            # time of the previous call's ticks: keep a chunk near chunk_s
            if t_ret is not None and prev_ran > 0 and c1 > t_ret:
                chunk = int(np.clip(chunk_s * prev_ran / (c1 - t_ret), 4, 4096))
            t_ret, prev_ran = c1, ran
            for jid in list(out):
                state, lat = b.state(jid)
                if state == 2:
                    i, lag, _ = out.pop(jid)
                    r = requests[i]
                    r.latency = lag + lat
                    img, ill = b.result(jid)
                    if keep_images:
                        r.image, r.illum = img, ill
        makespan = time.perf_counter() - t0
        st = b.stats()
    extra = {k: st[k] for k in ("ticks", "rebuilds", "tunes", "groups")}
    extra.update({k: round(st[k], 4) for k in ("run_seconds", "admit_seconds", "tune_seconds", "retire_seconds",
                                                  "rebuild_seconds", "capture_seconds")})
    extra["fill_cells"] = fill_cells
    return summary(requests, makespan, specs, extra=extra)


def serial_plain(specs: dict[str, ModelSpec], requests: list[Request], device: int = 0,
                 keep_images: bool = False) -> dict:
    """Service time of every request on the plain engine (models set up once,
    one shot after the other), then the FIFO latencies of a single plain
    engine under the same arrival times."""
    engines = {}
    try:
        for k, sp in specs.items():
            e = HipEngine(device)
            e.setup(sp.model, sp.params, sp.nt, sp.dt, snapshot_format=sp.snapshot_format)
            engines[k] = e
            e.migrate([sp.shots[0]])          # warm-up (and step-kernel autotune)
        service = []
        for r in requests:
            t = time.perf_counter()
            img, ill = engines[r.key].migrate([r.shot])
            service.append(time.perf_counter() - t)
            if keep_images:
                r.image, r.illum = img, ill
    finally:
        for e in engines.values():
            e.close()
    done = 0.0
    for r, s in sorted(zip(requests, service), key=lambda p: p[0].arrival):
        done = max(done, r.arrival) + s
        r.latency = done - r.arrival
    makespan = max(done, 0.0) - min(r.arrival for r in requests) if requests else 0.0
    out = summary(requests, makespan, specs)
    out["busy_s"] = float(sum(service))
    out["shots_per_s_busy"] = len(requests) / float(sum(service)) if service else 0.0
    return out


def summary(requests: list[Request], makespan: float, specs: dict[str, ModelSpec], extra=None) -> dict:
    lat = np.array([r.latency for r in requests])
    cells = sum(specs[r.key].cell_steps for r in requests)
    out = {"jobs": len(requests), "makespan_s": makespan,
           "shots_per_s": len(requests) / makespan if makespan > 0 else 0.0,
           "gcell_steps_per_s": cells / makespan * 1e-9 if makespan > 0 else 0.0,
           "latency_p50_s": float(np.percentile(lat, 50)), "latency_p99_s": float(np.percentile(lat, 99)),
           "per_model": {k: int(sum(r.key == k for r in requests)) for k in sorted(specs)}}
    if extra:
        out.update(extra)
    return out


# ---- aperture-cropped batch ------------------------------------------------

def crop_batch(njobs: int = 64, seed: int = 64, snapshot_format: int = SNAP_F32, full=None):
    """Aperture crops of the 2301 x 751 model (one shot each) with their
    specs; `full` is the 2301 x 751 model when the caller already has it."""
    rng = np.random.default_rng(seed)
    full = full if full is not None else models.faulted_layers(2301, 751, 4.0, 4.0)
    specs = {}
    reqs = []
    for k in range(njobs):
        width = int(rng.integers(400, 1201))
        ix0 = int(rng.integers(0, full.nx - width + 1))
        nt = int(rng.integers(2000, 8001))
        order = int(rng.choice([8, 12]))
        m = models.crop(full, ix0, width)
        p = RTMParams(nb=60, order=order, store_interval=4, f0=15.0)
        sp = ModelSpec(m, p, nt, 0.00035, snapshot_format)
        xmid = m.ox + 0.5 * (m.nx - 1) * m.dx
        sp.shots = models.surface_shots(m, nt, 0.00035, [xmid], rec_dx=12.0, rec_x0=m.ox)
        key = f"crop{k:02d}"
        specs[key] = sp
        reqs.append(Request(key, sp.shots[0], 0.0))
    return specs, reqs
