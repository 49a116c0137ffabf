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
               per launch / its mean duration from HIP events recorded on the
               engine's stream around sampled launches in the timed region
  cpu_baseline the reference's cpu-opt engine, compiled from its own sources
               (oracle/_ref), timed on this host on a bounded sample of the
               same workload (one shot, nt capped to a ~20 s budget)
  traffic      HBM bytes per backward-step launch from the rocprofv3 counter
               runs summarised in profiles/step_kernel_traffic_2301x751.md
               (FETCH_SIZE + WRITE_SIZE, each scaled by the counter's reading
               for the copy probe of known size), when that file is present

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
    nxe, nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb
    n_ext, n_int = nxe * nze, model.nx * model.nz
    nsnap = (nt - 1) // params.store_interval + 1
    per_shot = 32.0 * nt * n_ext + nsnap * n_int * (2 * snap_bytes + 16)
    fwd_launch = 16.0 * n_ext + nsnap * n_int * snap_bytes / nt
    bwd_launch = 16.0 * n_ext + nsnap * n_int * (snap_bytes + 16) / nt
    return n_ext, per_shot, fwd_launch, bwd_launch


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
    return {"value": round(cells / wall * 1e-9, 4), "unit": "Gcell-steps/s", "cores": int(threads),
            "kind": kind,
            "sample": f"cpu-opt engine, 1 shot of the 2301x751 workload, nt={nt_sample} of {shot.nt} "
                      f"({cells * 1e-9:.2f} G cell-steps, {wall:.1f} s wall, {prop:.1f} s propagation), "
                      f"IEEE subnormals kept, {os.cpu_count()} host cpus visible"}


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--family", type=int, default=-1, help="step-kernel family (-1 built-in, 0 LDS tile, 1 march)")
    ap.add_argument("--tile", type=int, default=0, help="step-kernel tile width (0 = built-in)")
    ap.add_argument("--sample-every", type=int, default=32, help="event-time every N-th step launch")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--nt", type=int, default=8000, help="time steps per shot (8000 = the workload)")
    ap.add_argument("--probe-bytes", type=int, default=1 << 30, help="copy-probe buffer size")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; multi-GPU figures are not produced "
                                   "(shots shard across GPUs with one final image reduce)"}))
        sys.exit(2)

    import rtm_amd
    if rtm_amd.device_count() < 1:
        sys.exit("bench: no HIP device")
    K, W = args.steps, args.warmup
    model, params, nt, dt, shots = workload(K + W, args.nt)
    copy_peak = rtm_amd.bandwidth_probe(0, args.probe_bytes, 5)
    n_ext, bytes_shot, fwd_b, bwd_b = algorithmic_bytes(model, params, nt)

    eng = rtm_amd.HipEngine(0)
    eng.setup(model, params, nt, dt)
    eng.set_step_kernel(args.family, args.tile)
    print(f"bench: device {eng.info()['device_name']}, modelling {K + W} shots", file=sys.stderr, flush=True)
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
    eng.close()

    cells = 2.0 * nt * n_ext * K
    value = cells / elapsed * 1e-9
    fwd_ms = ks["sampled_ms"][0] / max(ks["sampled"][0], 1)
    bwd_ms = ks["sampled_ms"][1] / max(ks["sampled"][1], 1)
    achieved = bwd_b / (bwd_ms * 1e-3) * 1e-9
    shot_bytes_rate = bytes_shot * K / elapsed * 1e-9
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "Gcell-steps/s", "n_gpus": 1,
        "steps": K, "warmup": W, "ms_per_step": round(elapsed / K * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded faulted-layer model; traces modelled on the GPU)",
        "config": {"workload": "2301x751 @4 m, nt 8000, dt 0.35 ms, order 8, nb 60, store 4, "
                               "767 receivers, plain per-shot engine, FP32 snapshots",
                   "engine": "hip",
                   "step_kernel": {d: {"family": ["lds-tile", "march", "march-dpp"][info["step_family"][k]],
                                       "tile_x": info["tile_x"][k], "autotuned": bool(info["autotuned"][k])}
                                   for k, d in enumerate(("forward", "backward"))},
                   "ldz": info["ldz"]},
        "shots_per_s": round(K / elapsed, 4),
        "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                       "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
        "whole_shot_hbm_frac": round(shot_bytes_rate / HBM_PEAK_GBS, 4),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic_per_launch(),
                     "copy_probe_gbs": round(copy_peak, 1),
                     "frac_of_copy_probe": round(achieved / copy_peak, 4) if copy_peak else None,
                     "kernel": "backward fused step (stencil + sponge + injection + imaging), order 8",
                     "bytes_per_launch": round(bwd_b), "mean_launch_us": round(bwd_ms * 1e3, 3),
                     "forward_kernel_us": round(fwd_ms * 1e3, 3),
                     "sampled_launches": ks["sampled"]},
        "image_checksum": float(np.abs(img).sum()),
    }
    if not args.no_cpu:
        print("bench: CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(model, params, dt, shots[0])
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
