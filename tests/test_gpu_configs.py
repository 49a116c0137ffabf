"""The batched, served and reduced-precision workloads at their stated sizes.

* 64 aperture crops of the 2301 x 751 model (widths U[400,1200], nz 751, nt
  U[2000,8000], orders 8 and 12 mixed, seeded) all submitted to one batch:
  every image equals the plain engine's migration of that crop alone, bit for
  bit; the largest order-8 and order-12 crops also equal the CPU oracle's
  cpu-opt at full nt (the per-shot body both reproduce is src/rtm/rtm.cpp:88-101).
* A served stream over the three resident models (201 x 151 @ 10 m nt 1400,
  1201 x 401 @ 8 m nt 4000, 2301 x 751 @ 4 m nt 8000) with FP16 snapshots,
  requests arriving while others run: every result equals the plain engine's
  migration of that shot alone; FP32 requests of the 201 x 151 workload with
  the reference's settings (the golden one's shots, direct wave muted) equal
  the oracle's image of that shot.
* The 4001 x 1201 salt model (order 12, nb 80, store 4, nt 10 000), whose
  snapshot store holds 1.2e10 elements, past 2^31 (slots are size_t in the
  reference, src/rtm/forward.cpp:121-123): two FP32 runs are bit-identical,
  the FP32 image equals the oracle's cpu-opt bit for bit, and the FP16-snapshot
  image is within 1e-3 relative L2 of the FP32 one.
Each case prints its parameters first so a failure names it."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import bits_equal

pytestmark = pytest.mark.gpu


def _rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# ---- 64 aperture crops ------------------------------------------------------

@pytest.fixture(scope="module")
def crops(hip_engine_cls):
    """The 64 crops with GPU-modelled traces and each crop migrated alone."""
    from rtm_amd import serve
    specs, _ = serve.crop_batch(64, seed=64)
    alone = {}
    for key in sorted(specs):
        sp = specs[key]
        with hip_engine_cls() as eng:
            eng.setup(sp.model, sp.params, sp.nt, sp.dt)
            sp.shots[0].traces = eng.model_shot(sp.shots[0])
            alone[key] = eng.migrate([sp.shots[0]])
    return specs, alone


def test_crops_64_batched_equal_alone(crops):
    import rtm_amd
    specs, alone = crops
    orders = sorted({sp.params.order for sp in specs.values()})
    nts = [sp.nt for sp in specs.values()]
    widths = [sp.model.nx for sp in specs.values()]
    print(f"case: 64 crops in one batch, orders {orders}, nt {min(nts)}..{max(nts)}, "
          f"widths {min(widths)}..{max(widths)}", flush=True)
    assert orders == [8, 12] and len(specs) == 64
    with rtm_amd.BatchEngine(0) as b:
        ids = {key: b.submit(sp.model, sp.params, sp.nt, sp.dt, sp.shots[0], sp.snapshot_format)
               for key, sp in specs.items()}
        b.run()
        for key, jid in ids.items():
            I, S = b.result(jid)
            assert bits_equal(alone[key][0], I), key
            assert bits_equal(alone[key][1], S), key
        st = b.stats()
    print(f"  batch ticks {st['ticks']}, groups {st['groups']}", flush=True)


@pytest.mark.parametrize("order", [8, 12])
def test_largest_crop_matches_oracle(crops, order):
    specs, alone = crops
    key = max((k for k, sp in specs.items() if sp.params.order == order),
              key=lambda k: specs[k].cell_steps)
    sp = specs[key]
    print(f"case: crop {key} {sp.model.nx}x{sp.model.nz} order {order} nt {sp.nt} vs oracle cpu-opt", flush=True)
    I_ref, S_ref = O.Oracle(sp.model, sp.params, sp.nt, sp.dt, variant="cpu-opt").migrate([sp.shots[0]])
    assert bits_equal(alone[key][0], I_ref), f"rel L2 {_rel(alone[key][0], I_ref):.3e}"
    assert bits_equal(alone[key][1], S_ref)


# ---- served stream over the three models -----------------------------------

def test_served_three_models_equal_alone(hip_engine_cls):
    import rtm_amd
    import rtm_amd.models as M
    from rtm_amd import serve
    import golden_data as G
    specs = serve.served_models(nsources=2, snapshot_format=rtm_amd.SNAP_F16)
    serve.model_traces(specs)
    # FP32 requests of the 201x151 workload with the reference's settings
    man, v, _, _ = G.layered3()
    s = man["settings"]
    m = rtm_amd.VelocityModel(201, 151, 10.0, 10.0, v)
    gold_shots = M.synth_line(m, s["nshots"], s["nt"], s["dt"], s["rec_dx"], s["src_z"], s["rec_z"])
    gold_params = rtm_amd.RTMParams(nb=s["nb"], order=s["order"], store_interval=s["store_interval"], f0=s["f0"])
    with hip_engine_cls() as eng:
        eng.setup(m, rtm_amd.RTMParams(nb=s["model_nb"], order=s["model_order"], store_interval=1, f0=s["f0"]),
                  s["nt"], s["dt"])
        t = float(np.float32(1) / np.float32(s["f0"]))
        for sh in gold_shots:
            sh.traces = eng.model_shot(sh)
            rtm_amd.mute_direct_wave(sh, s["mute_direct"], t, t)
    specs["layered_golden_f32"] = serve.ModelSpec(m, gold_params, s["nt"], s["dt"], rtm_amd.SNAP_F32,
                                                  gold_shots[:2])
    requests = serve.served_stream(specs, rate=24.0, njobs=16, seed=11)
    for key in sorted(specs):         # every model at least once
        if not any(r.key == key for r in requests):
            requests.append(serve.Request(key, specs[key].shots[0], requests[-1].arrival + 0.01))
    print(f"case: served stream of {len(requests)} requests over {sorted(specs)}", flush=True)
    summary = serve.serve(specs, requests, keep_images=True, fill_cells=0)     # every request admitted on arrival
    print(f"  p50 {summary['latency_p50_s']:.3f} s, p99 {summary['latency_p99_s']:.3f} s", flush=True)
    alone = {}
    for r in requests:
        k = (r.key, id(r.shot))
        if k not in alone:
            sp = specs[r.key]
            with hip_engine_cls() as eng:
                eng.setup(sp.model, sp.params, sp.nt, sp.dt, snapshot_format=sp.snapshot_format)
                alone[k] = eng.migrate([r.shot])
        assert bits_equal(alone[k][0], r.image), r.key
        assert bits_equal(alone[k][1], r.illum), r.key
    for sh in gold_shots[:2]:
        k = ("layered_golden_f32", id(sh))
        if k in alone:
            I_ref, S_ref = O.Oracle(m, gold_params, s["nt"], s["dt"]).migrate([sh])
            assert bits_equal(alone[k][0], I_ref) and bits_equal(alone[k][1], S_ref)


# ---- 4001 x 1201 salt model, snapshots past 2^31 elements -------------------

@pytest.fixture(scope="module")
def salt(hip_engine_cls):
    import rtm_amd
    from rtm_amd import models
    model = models.salt_body(4001, 1201, 5.0, 5.0)
    params = rtm_amd.RTMParams(nb=80, order=12, store_interval=4, f0=12.0)
    nt, dt = 10000, 0.0004
    shot = models.surface_shots(model, nt, dt, [0.5 * (model.nx - 1) * model.dx], rec_dx=10.0)[0]
    nsnap = (nt - 1) // params.store_interval + 1
    assert nsnap * model.nx * model.nz > 2 ** 31
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, dt)
        shot.traces = eng.model_shot(shot)
        I1, S1 = eng.migrate([shot])
        I2, S2 = eng.migrate([shot])
    return model, params, nt, dt, shot, (I1, S1), (I2, S2)


def test_salt_fp32_deterministic_and_fp16_close(hip_engine_cls, salt):
    import rtm_amd
    model, params, nt, dt, shot, (I1, S1), (I2, S2) = salt
    print("case: salt 4001x1201 order 12 nt 10000, FP32 twice and FP16", flush=True)
    assert bits_equal(I1, I2) and bits_equal(S1, S2)
    assert np.abs(I1).max() > 0
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, dt, snapshot_format=rtm_amd.SNAP_F16)
        I16, S16 = eng.migrate([shot])
    rel = _rel(I16, I1)
    print(f"  FP16 vs FP32 image: relative L2 {rel:.3e}", flush=True)
    assert rel <= 1e-3
    assert _rel(S16, S1) <= 1e-3


def test_salt_fp32_matches_oracle(salt):
    model, params, nt, dt, shot, (I1, S1), _ = salt
    print("case: salt 4001x1201 order 12 nt 10000 vs oracle cpu-opt", flush=True)
    I_ref, S_ref = O.Oracle(model, params, nt, dt, variant="cpu-opt").migrate([shot])
    assert bits_equal(I1, I_ref), f"rel L2 {_rel(I1, I_ref):.3e}"
    assert bits_equal(S1, S_ref)
