"""The HIP engine against the golden vectors written by the reference build.

* tiny grids, every order: traces and images bit-identical;
* the 4-shot 201x151 workload: GPU-modelled shots reproduce the reference's
  RTMS checksum, and the GPU migration (after the direct-wave mute) reproduces
  the reference image and illumination map bit for bit, and the device
  post-processing reproduces the reference's compensated and filtered images;
* Marmousi: the 12 GPU-modelled shots reproduce the reference-held sha256 of
  the dataset (data/real/marmousi_dataset.sha256:1).
"""
import numpy as np
import pytest

import golden_data as G
from conftest import bits_equal

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", G.tiny_cases(), ids=lambda c: c["name"])
def test_tiny_cases(hip_engine_cls, case):
    import rtm_amd
    print(f"case: tiny {case['name']}", flush=True)
    m = rtm_amd.VelocityModel(case["nx"], case["nz"], case["dx"], case["dz"], case["v"])
    p = rtm_amd.RTMParams(nb=case["nb"], order=case["order"], store_interval=case["store"], f0=case["f0"])
    shots = [rtm_amd.ShotRecord(float(sx), case["sz"], case["rx"], case["rz"], case["nt"], case["dt"])
             for sx in case["sx"]]
    with hip_engine_cls() as eng:
        eng.setup(m, p, case["nt"], case["dt"])
        for s, ref in zip(shots, case["traces"]):
            assert bits_equal(eng.model_shot(s), ref)
            s.traces = ref
        eng.setup(m, p, case["nt_migrate"], case["dt"])
        image, illum = eng.migrate(shots)
    assert bits_equal(image, case["image"])
    assert bits_equal(illum, case["illum"])


def test_layered3_pipeline(hip_engine_cls):
    import rtm_amd
    import rtm_amd.models as M
    print("case: layered3 4 shots", flush=True)
    man, v, ref_image, ref_illum = G.layered3()
    s = man["settings"]
    m = rtm_amd.VelocityModel(201, 151, 10.0, 10.0, v)
    shots = M.synth_line(m, s["nshots"], s["nt"], s["dt"], s["rec_dx"], s["src_z"], s["rec_z"])
    with hip_engine_cls() as eng:
        eng.setup(m, rtm_amd.RTMParams(nb=s["model_nb"], order=s["model_order"], store_interval=1,
                                       f0=s["f0"]), s["nt"], s["dt"])
        traces = [eng.model_shot(sh) for sh in shots]
        assert G.sha256(G.rtms_bytes(shots, traces)) == man["sha256"]["shots.rtms"]
        t = float(np.float32(1) / np.float32(s["f0"]))
        for sh, tr in zip(shots, traces):
            sh.traces = tr
            rtm_amd.mute_direct_wave(sh, s["mute_direct"], t, t)
        eng.setup(m, rtm_amd.RTMParams(nb=s["nb"], order=s["order"], store_interval=s["store_interval"],
                                       f0=s["f0"]), s["nt"], s["dt"])
        image, illum = eng.migrate(shots)
        # run twice: no state leaks between migrations (scripts/run_synthetic.sh:26-40)
        image2, illum2 = eng.migrate(shots)
        # the reference's post-processed outputs (src/main.cpp:211-222), on the device
        post = {"illum.f32": eng.postprocess(illum=True),
                "filtered.f32": eng.postprocess(illum=False, laplacian=True),
                "filtered_illum.f32": eng.postprocess(illum=True, laplacian=True)}
    assert bits_equal(image, ref_image)
    assert bits_equal(illum, ref_illum)
    assert bits_equal(image2, image) and bits_equal(illum2, illum)
    for name, arr in post.items():
        assert G.sha256(arr.tobytes()) == man["sha256"][name], name


def test_marmousi_dataset_checksum(hip_engine_cls):
    import rtm_amd
    import rtm_amd.models as M
    print("case: marmousi 12 shots", flush=True)
    man, v = G.marmousi()
    s = man["settings"]
    m = rtm_amd.VelocityModel(s["nx"], s["nz"], s["dx"], s["dx"], v)
    shots = M.synth_line(m, s["nshots"], s["nt"], s["dt"], s["rec_dx"], s["src_z"], s["rec_z"])
    with hip_engine_cls() as eng:
        eng.setup(m, rtm_amd.RTMParams(nb=s["nb"], order=s["order"], store_interval=1, f0=s["f0"]),
                  s["nt"], s["dt"])
        traces = [eng.model_shot(sh) for sh in shots]
    for k, tr in enumerate(traces):
        assert G.sha256(tr.tobytes()) == man["per_shot_traces_sha256"][k], f"shot {k}"
    assert G.sha256(G.rtms_bytes(shots, traces)) == man["recorded_sha256"]["marmousi_shots.bin"]
