"""Served stream: requests arriving in real time (seeded Poisson) on three
resident models are admitted while others run, and every result equals the
plain engine's migration of that request alone ("alone == in the mix",
bit-exact: both use the same snapshot format and the same FP32 arithmetic).
Each case prints its parameters first so a failure names it."""
import copy

import numpy as np
import pytest

from conftest import bits_equal, line_shots, small_model

pytestmark = pytest.mark.gpu


def _specs():
    import rtm_amd
    from rtm_amd import serve
    out = {}
    for key, (nx, nz, order, nt, store, fmt) in {
        "a": (61, 41, 8, 220, 2, rtm_amd.SNAP_F16),
        "b": (90, 35, 12, 300, 3, rtm_amd.SNAP_F16),
        "c": (45, 52, 4, 160, 1, rtm_amd.SNAP_F32),
    }.items():
        m = small_model(nx, nz)
        p = rtm_amd.RTMParams(nb=16, order=order, store_interval=store, f0=15.0)
        sp = serve.ModelSpec(m, p, nt, 0.001, fmt)
        xmax = (nx - 1) * m.dx
        sp.shots = line_shots(m, nt, 0.001, [0.25 * xmax, 0.6 * xmax], duplicate=True, clamp=False)
        out[key] = sp
    serve.model_traces(out)
    return out


@pytest.mark.parametrize("rate", [200.0, 2000.0])
def test_served_equals_alone(rate):
    from rtm_amd import serve
    print(f"case: served stream rate={rate}/s", flush=True)
    specs = _specs()
    reqs = serve.served_stream(specs, rate, 24, seed=3)
    alone = copy.deepcopy(reqs)
    s = serve.serve(specs, reqs, keep_images=True, chunk_s=0.001)
    assert s["jobs"] == 24 and all(r.latency > 0 for r in reqs)
    assert s["latency_p99_s"] >= s["latency_p50_s"] > 0
    p = serve.serial_plain(specs, alone, keep_images=True)
    assert p["jobs"] == 24
    for k, (r, a) in enumerate(zip(reqs, alone)):
        assert bits_equal(a.image, r.image), f"request {k} ({r.key}) image differs"
        assert bits_equal(a.illum, r.illum), f"request {k} ({r.key}) illumination differs"


def test_admission_under_small_budget():
    """A budget that holds only a few jobs at a time still serves everything."""
    import rtm_amd
    from rtm_amd import serve
    print("case: served stream under a small budget", flush=True)
    specs = _specs()
    reqs = serve.served_stream(specs, 500.0, 10, seed=9)
    with rtm_amd.BatchEngine(0) as probe:
        sizes = [probe.job_bytes(sp.model, sp.params, sp.nt, sp.dt, sp.shots[0], sp.snapshot_format)
                 for sp in specs.values()]
    alone = copy.deepcopy(reqs)
    serve.serve(specs, reqs, keep_images=True, memory_budget=int(max(sizes) * 1.5) + sum(sizes))
    serve.serial_plain(specs, alone, keep_images=True)
    for r, a in zip(reqs, alone):
        assert bits_equal(a.image, r.image)
