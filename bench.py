#!/usr/bin/env python3
"""Throughput of the MI355X RTM engine on the 2301 x 751 workload.

Workload (one GPU): 2301 x 751 cells at 4 m, seeded model of 20 dipping layers
cut by 3 normal faults, 1500 -> 5500 m/s with a smooth +-5 % perturbation;
nt = 8000, dt = 0.35 ms, 15 Hz Ricker, order 8, nb = 60, store interval 4,
767 receivers every 12 m, shots every 48 m.

A step is one migrated shot: forward propagation with the snapshot store,
then backward propagation with the fused imaging condition, on inputs already
resident on the device (traces are modelled on the GPU before the timed
region; the timed region includes only the per-shot trace upload and the
final image download that the engine does anyway).

Reported (one JSON line):
  value        Gcell-steps/s, cell-steps = 2 nt (nx+2nb)(nz+2nb) per shot
               (the reference's accounting, src/benchmark/benchmark.cpp:22-25)
  roofline     the backward step kernel (the dominant one): algorithmic bytes
               per launch / its mean duration, from HIP events recorded on the
               engine's stream around each shot's 8000 back-to-back backward
               step launches in the timed region (divided by 8000); the
               algorithmic bytes are the reference algorithm's (each stored
               step imaged on its own), and executed_bytes_per_launch those of
               the paired imaging the engine runs (fewer)
  cpu_baseline the reference's cpu-opt engine, compiled from its own sources
               (oracle/_ref), timed on this host on a bounded sample of the
               same workload (one shot, nt capped to a ~20 s budget)
  traffic      HBM bytes per backward-step launch from the rocprofv3 counter
               runs summarised in profiles/step_kernel_traffic_2301x751.md
               (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading
               for the copy probe of known size), when that file is present
  batched      the batch engine on 8 shots of the same workload at once
  crops        64 aperture crops of the model (widths U[400,1200], nt
               U[2000,8000], orders 8 and 12) submitted at once, longest
               first, to the batch engine, against the plain engine migrating
               them one by one in the same order; timed without device
               allocation on either side, and (with_setup) with engine
               creation, allocation and setup on both
  served       seeded Poisson streams (8/s and 64/s, 16 requests each) over
               three resident models served in real time by the batch engine,
               against one plain engine serving them first come first served:
               throughput, p50 and p99 latency and their ratios
  whole_shot_hbm_frac  algorithmic bytes of whole shots / elapsed time / 8 TB/s,
               for every line (fields, tables, snapshot store and load, image)

--nt shortens the shot for counter-collection runs only; the reported line
always uses the full nt = 8000 unless --nt is given.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "rtm-acceleration_amd"))

HBM_PEAK_GBS = 8000.0           # MI355X HBM3E, published
METRIC = "Gcell-steps/s & % HBM roofline; migrated shots/s + p99 latency vs plain GPU"


def workload(nshots_needed: int, nt: int = 8000):
    from rtm_amd import RTMParams, models
    model = models.faulted_layers(2301, 751, 4.0, 4.0)
    params = RTMParams(nb=60, order=8, store_interval=4, f0=15.0)
    dt = 0.00035
    rec_x = 12.0
    shot_x = 48.0 * np.arange(192)
    # take shots spread over the line rather than the first few
    pick = np.linspace(20, 170, nshots_needed).astype(int)
    shots = models.surface_shots(model, nt, dt, shot_x[pick], rec_dx=rec_x)
    return model, params, nt, dt, shots


def algorithmic_bytes(model, params, nt, snap_bytes=4):
    """Bytes of the reference's algorithm: 16 per extended cell and step,
    plus per stored step and interior cell the snapshot write (forward) and
    read (backward) and the image/illumination read-modify-write (16)."""
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    n_ext, n_int = nxe * nze, model.nx * model.nz
    nsnap = (nt - 1) // params.store_interval + 1
    per_shot = 32.0 * nt * n_ext + nsnap * n_int * (2 * snap_bytes + 16)
    fwd_launch = 16.0 * n_ext + nsnap * n_int * snap_bytes / nt
    bwd_launch = 16.0 * n_ext + nsnap * n_int * (snap_bytes + 16) / nt
    return n_ext, per_shot, fwd_launch, bwd_launch


def executed_backward_bytes(model, params, nt, snap_bytes=4):
    """Mean bytes per backward launch as the engine images: stored steps in
    pairs, each pair reading two snapshots and the held field (4) and
    updating image and illumination once (16); an unpaired one as above."""
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    n_ext, n_int = nxe * nze, model.nx * model.nz
    nsnap = (nt - 1) // params.store_interval + 1
    per_int = (nsnap // 2) * (2 * snap_bytes + 20) + (nsnap % 2) * (snap_bytes + 16)
    return 16.0 * n_ext + per_int * n_int / nt


def traffic_per_launch():
    """HBM bytes per backward-step launch from the committed counter summary
    (a line 'traffic_bytes_per_launch: N'), or None."""
    path = os.path.join(REPO, "profiles", "step_kernel_traffic_2301x751.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.strip().startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(model, params, dt, shot, budget_s: float = 20.0):
    """The reference's cpu-opt engine on one shot of the same grid with nt
    reduced so the run stays within a few tens of seconds."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as O
    kind = "reference" if O.ref_available() else "port"
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    nt_probe = 20
    t0 = time.perf_counter()
    if kind == "reference":
        O.ref_migrate("cpu-opt", model, params, nt_probe, dt, [shot])
        threads = O.ref().ref_num_threads(b"cpu-opt")
    else:
        orc = O.Oracle(model, params, nt_probe, dt, variant="cpu-opt")
        orc.migrate([shot])
        threads = O.lib().orc_num_threads(orc.h)
    probe = time.perf_counter() - t0
    nt_sample = int(max(40, min(shot.nt, nt_probe * budget_s / max(probe, 1e-3))))
    t0 = time.perf_counter()
    if kind == "reference":
        _, _, times = O.ref_migrate("cpu-opt", model, params, nt_sample, dt, [shot])
        prop = times[0] + times[1]
    else:
        orc = O.Oracle(model, params, nt_sample, dt, variant="cpu-opt")
        orc.migrate([shot])
        f, b, _ = orc.times()
        prop = f + b
    wall = time.perf_counter() - t0
    cells = 2.0 * nt_sample * nxe * nze
    out = {"value": round(cells / wall * 1e-9, 4), "unit": "Gcell-steps/s", "cores": int(threads),
           "kind": kind, "cpu_model": cpu_model(),
           "sample": f"cpu-opt engine, 1 shot of the 2301x751 workload, nt={nt_sample} of {shot.nt} "
                     f"({cells * 1e-9:.2f} G cell-steps, {wall:.1f} s wall, {prop:.1f} s propagation), "
                     f"IEEE subnormals kept, {os.cpu_count()} host cpus visible"}
    if kind == "reference":
        out["cpu_engine"] = cpu_single_thread(model, params, dt, shot, budget_s / 2)
    return out


def cpu_model() -> str:
    """The host CPU's model name (/proc/cpuinfo), as the reference's CSV
    records it (src/benchmark/benchmark.cpp cpu_model_name)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_single_thread(model, params, dt, shot, budget_s: float):
    """The reference's plain cpu engine (one thread, -O2) on the same shot,
    nt reduced to a budget of about budget_s seconds."""
    import oracle_lib as O
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    nt_probe = 4
    t0 = time.perf_counter()
    O.ref_migrate("cpu", model, params, nt_probe, dt, [shot])
    probe = time.perf_counter() - t0
    nt_sample = int(max(8, min(shot.nt, nt_probe * budget_s / max(probe, 1e-3))))
    t0 = time.perf_counter()
    O.ref_migrate("cpu", model, params, nt_sample, dt, [shot])
    wall = time.perf_counter() - t0
    cells = 2.0 * nt_sample * nxe * nze
    return {"value": round(cells / wall * 1e-9, 4), "unit": "Gcell-steps/s",
            "cores": int(O.ref().ref_num_threads(b"cpu")), "kind": "reference",
            "sample": f"cpu engine, 1 shot of the 2301x751 workload, nt={nt_sample} of {shot.nt} "
                      f"({cells * 1e-9:.2f} G cell-steps, {wall:.1f} s wall), IEEE subnormals kept"}


def batched_same_shots(model, params, nt, dt, shots, warm_shot, snapshot_format=0, use_graph=True,
                       kernel_family=0, tile_x=0):
    """The batch engine on the same shots: one resident model, every shot
    submitted at once and advanced together, one launch per time step
    (direct launches at this tick size; hipGraph replay only for ticks small
    enough to be launch-bound).  Returns throughput and per-shot latency."""
    import rtm_amd
    with rtm_amd.BatchEngine(0, use_graph=use_graph, kernel_family=kernel_family, tile_x=tile_x) as b:
        # job memory placed once, outside the timed region: what the warm-up
        # job and the timed shots need, not the whole budget (a later engine
        # in this process would wait for the device to clear all of it)
        jb = b.job_bytes(model, params, nt, dt, warm_shot, snapshot_format)
        b.reserve(int(jb * (len(shots) + 1) * 1.02) + (256 << 20))
        mid = b.add_model(model, params, nt, dt, snapshot_format)
        b.submit_shot(mid, warm_shot)             # warm-up: step-kernel autotune, graph capture
        b.run()
        t0 = time.perf_counter()
        ids = [b.submit_shot(mid, s) for s in shots]
        b.run()
        elapsed = time.perf_counter() - t0
        lat = [b.state(j)[1] for j in ids]
        st = b.stats()
        img = sum(b.result(j)[0].astype(np.float64) for j in ids)
    n_ext = (model.nx + 2 * params.nb) * (model.nz + 2 * params.nb)
    bytes_shot = algorithmic_bytes(model, params, nt, 4 if snapshot_format == 0 else 2)[1]
    return {"shots": len(shots), "elapsed_s": round(elapsed, 4),
            "whole_shot_hbm_frac": round(bytes_shot * len(shots) / elapsed * 1e-9 / HBM_PEAK_GBS, 4),
            "value": round(2.0 * nt * n_ext * len(shots) / elapsed * 1e-9, 3), "unit": "Gcell-steps/s",
            "shots_per_s": round(len(shots) / elapsed, 4),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                           "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
            "snapshots": ["f32", "f16", "bf16"][snapshot_format], "groups": st["groups"],
            "lean_launch_fraction": round(st["lean_launches"] / max(st["step_launches"], 1), 4),
            "image_checksum": float(np.abs(img).sum())}


# concurrent jobs of the crops batch: 8-24 give the same tick throughput
# (12-16 M extended cells per tick); fewer need less device memory
CROPS_ACTIVE = 12


def crops_run(ncrops: int, use_graph: bool = True, merge_groups: bool = False) -> dict:
    """ncrops aperture crops (widths U[400,1200], nt U[2000,8000], orders
    8/12) migrated by the batch engine, all submitted at once, and by the
    plain engine one after the other, both taking the crops longest first
    (most cell-steps first: the batch's last jobs then finish together
    instead of one long job running alone at the end; the per-shot loop's
    total does not depend on the order).  Both timings include everything a
    caller pays: engine creation, device allocation, model upload and setup,
    step-kernel choice, the migration and the image download."""
    import rtm_amd
    from rtm_amd import serve
    specs, reqs = serve.crop_batch(ncrops)
    print(f"bench: modelling {len(reqs)} crop shots", file=sys.stderr, flush=True)
    for sp in specs.values():
        with rtm_amd.HipEngine(0) as e:
            e.setup(sp.model, sp.params, sp.nt, sp.dt)
            sp.shots[0].traces = e.model_shot(sp.shots[0])
    cells = sum(sp.cell_steps for sp in specs.values())
    hbm_bytes = sum(algorithmic_bytes(sp.model, sp.params, sp.nt)[1] for sp in specs.values())
    order = sorted(specs, key=lambda k: (-specs[k].cell_steps, k))
    print("bench: crops one by one on the plain engine", file=sys.stderr, flush=True)
    plain_t, migrate_t = [], []
    t_all = time.perf_counter()
    for k in order:
        sp = specs[k]
        t = time.perf_counter()
        with rtm_amd.HipEngine(0) as e:
            e.setup(sp.model, sp.params, sp.nt, sp.dt)
            e.set_step_kernel(rtm_amd.STEP_MARCH_DPP, 4)    # no per-crop autotuning
            t_m = time.perf_counter()
            e.migrate([sp.shots[0]])
            migrate_t.append(time.perf_counter() - t_m)
        plain_t.append(time.perf_counter() - t)
    plain_elapsed = time.perf_counter() - t_all
    print("bench: crops batched", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    with rtm_amd.BatchEngine(0, use_graph=use_graph, merge_groups=merge_groups, max_active=CROPS_ACTIVE) as b:
        # at most CROPS_ACTIVE jobs run at once (about 12 M cells per tick,
        # which fills the chip); one block holds the largest CROPS_ACTIVE of
        # them, and a job admitted later reuses what a finished one freed
        # (the device clears what it allocates, so a larger block costs more)
        jb = sorted((b.job_bytes(sp.model, sp.params, sp.nt, sp.dt, sp.shots[0], sp.snapshot_format)
                     for sp in specs.values()), reverse=True)
        b.reserve(int(sum(jb[:CROPS_ACTIVE]) * 1.02) + (256 << 20))
        ids = [b.submit(sp.model, sp.params, sp.nt, sp.dt, sp.shots[0], sp.snapshot_format)
               for sp in (specs[k] for k in order)]
        b.run()
        lat = [b.state(j)[1] for j in ids]
        for j in ids:
            b.result(j)                           # the download retires the job
        elapsed = time.perf_counter() - t0
        st = b.stats()
    run_s = st["run_seconds"]
    plain_s = sum(migrate_t)
    done = np.cumsum(migrate_t)                   # all submitted at t=0, served in order
    return {
        "crops": len(reqs), "order": "longest first (most cell-steps first), on both sides",
        "value": round(cells / run_s * 1e-9, 3), "unit": "Gcell-steps/s",
        "elapsed_s": round(run_s, 3), "shots_per_s": round(len(reqs) / run_s, 4),
        "whole_shot_hbm_frac": round(hbm_bytes / run_s * 1e-9 / HBM_PEAK_GBS, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
        "timing": ("no device allocation on either side: the plain engine's migrate calls (each with its "
                   "trace upload, field resets and image download) against the batch engine's run (which also "
                   "admits every job: model tables, traces, field clears, image downloads); the with_setup "
                   "figures add engine creation, allocation and setup on both sides, and depend on how much "
                   "device memory earlier work in the process freed (the driver clears reused memory when it "
                   "is allocated again)"),
        "plain_per_shot_loop": {"step_kernel": "march-dpp tile 4",
                                "value": round(cells / plain_s * 1e-9, 3),
                                "elapsed_s": round(plain_s, 3),
                                "shots_per_s": round(len(reqs) / plain_s, 4),
                                "whole_shot_hbm_frac": round(hbm_bytes / plain_s * 1e-9 / HBM_PEAK_GBS, 4),
                                "latency_ms": {"p50": round(float(np.percentile(done, 50)) * 1e3, 2),
                                               "p99": round(float(np.percentile(done, 99)) * 1e3, 2)}},
        "speedup_vs_plain": round(plain_s / run_s, 3),
        "with_setup": {"elapsed_s": round(elapsed, 3), "plain_elapsed_s": round(plain_elapsed, 3),
                       "speedup_vs_plain": round(plain_elapsed / elapsed, 3)},
        "batch": {"max_active": CROPS_ACTIVE, "ticks": st["ticks"], "rebuilds": st["rebuilds"], "groups": st["groups"],
                  "merged_launch": bool(st["merged_launch"]),
                  "seconds": {k: round(st[k], 4) for k in ("run_seconds", "admit_seconds", "tune_seconds",
                                                           "retire_seconds")}},
    }


def crops_workload(args):
    """The crops run as its own bench line."""
    r = crops_run(args.crops, use_graph=not args.batch_no_graph, merge_groups=args.batch_merge)
    return {
        "metric": METRIC, "value": r["value"], "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": 1, "warmup": 0, "ms_per_step": round(r["elapsed_s"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded faulted-layer model, aperture crops; traces modelled on the GPU)",
        "config": {"workload": f"{r['crops']} aperture crops of the 2301x751 @4 m model, widths U[400,1200], "
                               "nz 751, nt U[2000,8000], dt 0.35 ms, orders 8/12, nb 60, store 4, FP32 "
                               "snapshots, all submitted at once (longest first) to the batch engine",
                   "engine": "hip-batched"},
        "crops": r,
    }


def served_sweep(rates, jobs: int, reserve_gib: int, chunk_ms: float, use_graph: bool = True,
                 plain_serial: bool = True) -> list:
    """Seeded Poisson streams over three resident models at each rate, served
    in real time by the batch engine, against one plain engine serving the
    same requests first come first served (its service times measured, the
    queueing replayed on the same arrival times)."""
    from rtm_amd import serve
    specs = serve.served_models()
    print("bench: modelling traces of the served models", file=sys.stderr, flush=True)
    serve.model_traces(specs)
    sweep = []
    for rate in rates:
        reqs = serve.served_stream(specs, rate, jobs, seed=int(rate * 1000) + 7)
        print(f"bench: serving {len(reqs)} requests at {rate}/s", file=sys.stderr, flush=True)
        batched = serve.serve(specs, reqs, reserve=reserve_gib << 30, use_graph=use_graph, chunk_s=chunk_ms * 1e-3)
        hbm = sum(algorithmic_bytes(specs[r.key].model, specs[r.key].params, specs[r.key].nt, 2)[1] for r in reqs)
        batched["whole_shot_hbm_frac"] = hbm / batched["makespan_s"] * 1e-9 / HBM_PEAK_GBS
        row = {"rate_per_s": rate, "batched": _round(batched)}
        if plain_serial:
            plain = serve.serial_plain(specs, reqs)
            plain["whole_shot_hbm_frac"] = hbm / plain["makespan_s"] * 1e-9 / HBM_PEAK_GBS
            row.update({"plain_serial": _round(plain),
                        "shots_per_s_ratio": round(batched["shots_per_s"] / plain["shots_per_s"], 3),
                        "p50_ratio": round(batched["latency_p50_s"] / plain["latency_p50_s"], 3),
                        "p99_ratio": round(batched["latency_p99_s"] / plain["latency_p99_s"], 3)})
        sweep.append(row)
    return sweep


SERVED_WORKLOAD = ("served stream: requests uniform over 201x151 @10 m (nt 1400), 1201x401 @8 m (nt 4000), "
                   "2301x751 @4 m (nt 8000) resident models, order 8, FP32 fields, FP16 snapshots")


def served_workload(args):
    """Poisson streams over three resident models, swept in rate, against the
    plain engine serving the same requests serially (FIFO)."""
    rates = [float(r) for r in args.rates.split(",")]
    sweep = served_sweep(rates, args.jobs, args.serve_reserve_gib, args.serve_chunk_ms,
                         use_graph=not args.batch_no_graph, plain_serial=not args.no_plain_serial)
    top = sweep[-1]
    return {
        "metric": METRIC, "value": round(top["batched"]["gcell_steps_per_s"], 3), "unit": "Gcell-steps/s",
        "n_gpus": 1, "steps": args.jobs, "warmup": 0,
        "ms_per_step": round(top["batched"]["makespan_s"] / args.jobs * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded models; traces modelled on the GPU; seeded Poisson arrivals)",
        "config": {"workload": SERVED_WORKLOAD + "; value at the highest rate",
                   "engine": "hip-batched", "rates": rates, "jobs_per_rate": args.jobs},
        "shots_per_s": round(top["batched"]["shots_per_s"], 4),
        "latency_ms": {"p50": round(top["batched"]["latency_p50_s"] * 1e3, 2),
                       "p99": round(top["batched"]["latency_p99_s"] * 1e3, 2)},
        "sweep": sweep,
    }


def salt_workload(args):
    """Reduced-precision snapshots on the 4001x1201 salt model: one shot with
    FP32 snapshots and FP32 stack, one with FP16 snapshots (per-snapshot
    scale) and the FP64 stack, on the plain engine."""
    import rtm_amd
    from rtm_amd import models
    model = models.salt_body(4001, 1201, 5.0, 5.0)
    params = rtm_amd.RTMParams(nb=80, order=12, store_interval=4, f0=12.0)
    nt, dt = args.salt_nt, 0.0004
    shot = models.surface_shots(model, nt, dt, [0.5 * (model.nx - 1) * model.dx], rec_dx=10.0)[0]
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    cells = 2.0 * nt * nxe * nze
    out = {}
    for name, fmt, stack in (("f32", rtm_amd.SNAP_F32, rtm_amd.STACK_F32),
                             ("f16", rtm_amd.SNAP_F16, rtm_amd.STACK_F64)):
        with rtm_amd.HipEngine(0) as eng:
            eng.setup(model, params, nt, dt, snapshot_format=fmt, image_stack=stack)
            eng.set_step_kernel(args.family, args.tile)
            print(f"bench: salt {name}: modelling", file=sys.stderr, flush=True)
            if shot.traces is None:
                shot.traces = eng.model_shot(shot)
            eng.migrate([shot])                       # warm-up
            eng.set_kernel_timing(-1)
            eng.reset_times()
            t0 = time.perf_counter()
            eng.migrate([shot])
            el = time.perf_counter() - t0
            ks = eng.kernel_stats()
            times = eng.times()
            info = eng.info()
        sb = 4 if fmt == rtm_amd.SNAP_F32 else 2
        _, per_shot, fwd_b, bwd_b = algorithmic_bytes(model, params, nt, sb)
        bwd_us = ks["sampled_ms"][1] / max(ks["sampled"][1], 1) * 1e3
        fwd_us = ks["sampled_ms"][0] / max(ks["sampled"][0], 1) * 1e3
        out[name] = {"value": round(cells / el * 1e-9, 3), "shot_s": round(el, 4),
                     "forward_kernel_us": round(fwd_us, 3),
                     "engine_times_s": {k: round(v, 4) for k, v in times.items() if v},
                     "snapshot_gb": round(info["snapshot_bytes"] / 1e9, 2),
                     "backward_kernel_us": round(bwd_us, 3),
                     "backward_bytes_per_launch": round(bwd_b),
                     "backward_frac_of_peak": round(bwd_b / (bwd_us * 1e-6) * 1e-9 / HBM_PEAK_GBS, 4),
                     "whole_shot_hbm_frac": round(per_shot / el * 1e-9 / HBM_PEAK_GBS, 4)}
    return {
        "metric": METRIC, "value": out["f16"]["value"], "unit": "Gcell-steps/s", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": round(out["f16"]["shot_s"] * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded salt-body model; traces modelled on the GPU)",
        "config": {"workload": f"4001x1201 @5 m salt model, nt {nt}, dt 0.4 ms, 12 Hz, order 12, nb 80, "
                               "store 4, one surface shot; FP16 snapshots + FP64 stack (value) vs FP32",
                   "engine": "hip"},
        "snapshots": out, "speedup_f16_vs_f32": round(out["f32"]["shot_s"] / out["f16"]["shot_s"], 3),
    }


def _round(d):
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in d.items()}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--family", type=int, default=-1,
                    help="step-kernel family: -1 = the engine's autotuner (default), 0 LDS tile, 1 march, "
                         "2 DPP march, 3 stream")
    ap.add_argument("--tile", type=int, default=0, help="step-kernel tile width (0 = the autotuner's)")
    ap.add_argument("--batch-family", type=int, default=0,
                    help="batch step kernel (0 = autotuned, 1 march, 2 DPP march, 3 stream)")
    ap.add_argument("--batch-tile", type=int, default=0, help="batch step kernel tile width (0 = default)")
    ap.add_argument("--snapshots", choices=("f32", "f16", "bf16"), default="f32",
                    help="snapshot format of the plain engine's line (the workload's is f32)")
    ap.add_argument("--batch-snapshots", choices=("f32", "f16", "bf16"), default="f32",
                    help="snapshot format of the batched comparison's jobs")
    ap.add_argument("--batch-merge", action="store_true",
                    help="batch engine: one launch per tick for jobs of every stencil order (crops run)")
    ap.add_argument("--batch-no-graph", action="store_true",
                    help="batched comparison with direct launches instead of hipGraph replay")
    ap.add_argument("--sample-every", type=int, default=-1,
                    help="kernel timing: -1 = events around each shot's whole step loop (default), "
                         "N > 0 = events around every N-th step launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--nt", type=int, default=8000, help="time steps per shot (8000 = the workload)")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30, help="copy-probe buffer size")
    ap.add_argument("--salt-nt", type=int, default=10000, help="time steps of the salt workload")
    ap.add_argument("--workload", choices=["plain", "crops", "served", "salt"], default="plain",
                    help="plain: the 2301x751 per-shot line (default), with the batch engine on the same "
                         "shots; crops: 64 aperture crops batched vs one by one; served: Poisson streams; "
                         "salt: 4001x1201 order 12 with FP16 vs FP32 snapshots")
    ap.add_argument("--no-batched", action="store_true", help="plain workload without the batched comparison")
    ap.add_argument("--batched-shots", type=int, default=8,
                    help="shots the batch engine migrates at once in the plain workload's comparison")
    ap.add_argument("--crops", type=int, default=64, help="crops in the crops run")
    ap.add_argument("--no-crops", action="store_true", help="plain workload without the crops run")
    ap.add_argument("--no-served", action="store_true", help="plain workload without the served sweep")
    ap.add_argument("--line-rates", default="8,64",
                    help="arrival rates (1/s) of the served sweep in the plain workload's line")
    ap.add_argument("--line-jobs", type=int, default=16, help="requests per rate in that sweep")
    ap.add_argument("--rates", default="1,4,16,64", help="arrival rates (1/s) of the served workload")
    ap.add_argument("--jobs", type=int, default=32, help="requests per rate in the served workload")
    ap.add_argument("--no-plain-serial", action="store_true", help="served workload without the plain baseline")
    ap.add_argument("--serve-chunk-ms", type=float, default=10.0,
                    help="device time the served loop runs between looking at arrivals")
    ap.add_argument("--serve-reserve-gib", type=int, default=160,
                    help="device memory the served batch engine reserves for jobs before the stream starts")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not produced "
                                   "(shots shard across GPUs with one final image reduce)"}))
        sys.exit(2)

    import rtm_amd
    if rtm_amd.device_count() < 1:
        sys.exit("bench: no HIP device")
    if args.workload != "plain":
        line = {"crops": crops_workload, "served": served_workload, "salt": salt_workload}[args.workload](args)
        print(json.dumps(line), flush=True)
        return
    K, W = args.steps, args.warmup
    NB = 0 if args.no_batched else max(K, args.batched_shots)
    model, params, nt, dt, shots = workload(max(K, NB) + W, args.nt)
    copy_peak = rtm_amd.bandwidth_probe(0, args.probe_bytes, 5)
    fmt = ("f32", "f16", "bf16").index(args.snapshots)
    snap_b = 4 if fmt == 0 else 2
    n_ext, bytes_shot, fwd_b, bwd_b = algorithmic_bytes(model, params, nt, snap_b)

    eng = rtm_amd.HipEngine(0)
    eng.setup(model, params, nt, dt, snapshot_format=fmt)
    eng.set_step_kernel(args.family, args.tile)
    print(f"bench: device {eng.info()['device_name']}, modelling {len(shots)} shots", file=sys.stderr, flush=True)
    for s in shots:
        s.traces = eng.model_shot(s)

    if W:
        eng.migrate(shots[:W])
    print("bench: warmup done, timing", file=sys.stderr, flush=True)
    eng.reset_times()
    eng.set_kernel_timing(args.sample_every)
    lat = []
    eng.image_zero()
    t_start = time.perf_counter()
    for s in shots[W:W + K]:
        t0 = time.perf_counter()
        eng.forward(s)
        eng.backward(s)
        lat.append(time.perf_counter() - t0)
    img, _ = eng.image()
    elapsed = time.perf_counter() - t_start
    eng.set_kernel_timing(0)
    ks = eng.kernel_stats()
    info = eng.info()
    # post-processing of the stacked image on the device (outside the timed
    # region: the reference applies it once per run, after the last shot)
    eng.postprocess(illum=True, laplacian=True)        # first call sizes its buffer
    t0 = time.perf_counter()
    for _ in range(5):
        eng.postprocess(illum=True, laplacian=True)
    post_ms = (time.perf_counter() - t0) / 5 * 1e3
    eng.close()

    cells = 2.0 * nt * n_ext * K
    value = cells / elapsed * 1e-9
    fwd_ms = ks["sampled_ms"][0] / max(ks["sampled"][0], 1)
    bwd_ms = ks["sampled_ms"][1] / max(ks["sampled"][1], 1)
    achieved = bwd_b / (bwd_ms * 1e-3) * 1e-9
    bwd_exec = executed_backward_bytes(model, params, nt, snap_b)
    shot_bytes_rate = bytes_shot * K / elapsed * 1e-9
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(elapsed / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded faulted-layer model; traces modelled on the GPU)",
        "config": {"workload": "2301x751 @4 m, nt 8000, dt 0.35 ms, order 8, nb 60, store 4, "
                               f"767 receivers, plain per-shot engine, {args.snapshots.upper()} snapshots",
                   "engine": "hip",
                   "step_kernel": {d: {"family": ["lds-tile", "march", "march-dpp", "stream"][info["step_family"][k]],
                                       "tile_x": info["tile_x"][k], "autotuned": bool(info["autotuned"][k]),
                                       "waves_per_cu": info["waves_per_cu"][k] or "no cap"}
                                   for k, d in enumerate(("forward", "backward"))},
                   "ldz": info["ldz"]},
        "shots_per_s": round(K / elapsed, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
        "whole_shot_hbm_frac": round(shot_bytes_rate / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_per_launch() if fmt == 0 else None,   # measured with FP32 snapshots
                     "copy_probe_gbs": round(copy_peak, 1),
                     "frac_of_copy_probe": round(achieved / copy_peak, 4) if copy_peak else None,
                     "kernel": "backward fused step (stencil + sponge + injection + imaging), order 8; "
                               "mean over the phase's launches: imaging build on stored steps, lean build "
                               "on the others",
                     "bytes_per_launch": round(bwd_b), "mean_launch_us": round(bwd_ms * 1e3, 3),
                     "executed_bytes_per_launch": round(bwd_exec),
                     "frac_executed_bytes": round(bwd_exec / (bwd_ms * 1e-3) * 1e-9 / HBM_PEAK_GBS, 4),
                     "forward_kernel_us": round(fwd_ms * 1e3, 3),
                     "sampled_launches": ks["sampled"]},
        "image_checksum": float(np.abs(img).sum()),
        "postprocess": {"ms": round(post_ms, 3),
                        "what": "illumination compensation + Laplacian filter of the stacked image on the "
                                "device, including the download of the result"},
    }
    if not args.no_batched:
        print("bench: batch engine on the same shots", file=sys.stderr, flush=True)
        bt = batched_same_shots(model, params, nt, dt, shots[W:W + NB], shots[0],
                                snapshot_format=("f32", "f16", "bf16").index(args.batch_snapshots),
                                use_graph=not args.batch_no_graph, kernel_family=args.batch_family,
                                tile_x=args.batch_tile)
        bt["speedup_vs_plain_shots_per_s"] = round(bt["shots_per_s"] / line["shots_per_s"], 3)
        line["batched"] = bt
    if not args.no_crops:
        line["crops"] = crops_run(args.crops, use_graph=not args.batch_no_graph, merge_groups=args.batch_merge)
    if not args.no_served:
        rates = [float(r) for r in args.line_rates.split(",")]
        line["served"] = {"workload": SERVED_WORKLOAD, "jobs_per_rate": args.line_jobs,
                          "sweep": served_sweep(rates, args.line_jobs, args.serve_reserve_gib, args.serve_chunk_ms,
                                                use_graph=not args.batch_no_graph)}
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(model, params, dt, shots[0])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
