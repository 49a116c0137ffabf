"""The oracle against the golden vectors the reference build wrote.

These pin the CPU restatement (oracle/) to the reference's own outputs:
bit-identical traces and images on small grids for every order, the 4-shot
201x151 workload through the reference CLI pipeline, and the reference-held
Marmousi checksum (data/real/marmousi_dataset.sha256:1) for one shot.
"""
import numpy as np
import pytest

import golden_data as G
import oracle_lib as O
from conftest import bits_equal


def _tiny_setup(c):
    m = O.Model(c["nx"], c["nz"], c["dx"], c["dz"], c["v"])
    p = O.Params(nb=c["nb"], order=c["order"], store_interval=c["store"], f0=c["f0"])
    shots = [O.Shot(float(sx), c["sz"], c["rx"], c["rz"], c["nt"], c["dt"]) for sx in c["sx"]]
    return m, p, shots


@pytest.mark.parametrize("variant", ["cpu", "cpu-opt"])
@pytest.mark.parametrize("case", G.tiny_cases(), ids=lambda c: c["name"])
def test_tiny_cases_bit_identical(case, variant):
    m, p, shots = _tiny_setup(case)
    orc = O.Oracle(m, p, case["nt"], case["dt"], variant=variant)
    for s, ref_tr in zip(shots, case["traces"]):
        assert bits_equal(orc.model_shot(s), ref_tr)
        s.traces = ref_tr
    orc_m = O.Oracle(m, p, case["nt_migrate"], case["dt"], variant=variant)
    image, illum = orc_m.migrate(shots)
    assert bits_equal(image, case["image"])
    assert bits_equal(illum, case["illum"])


def layered3_shots(model_cls=O.Model, shot_cls=O.Shot):
    """The 4-shot acquisition of the layered workload, as rtm_synth builds it."""
    import rtm_amd.models as M
    man, v, _, _ = G.layered3()
    s = man["settings"]
    m = model_cls(201, 151, 10.0, 10.0, v)
    shots = M.synth_line(m, s["nshots"], s["nt"], s["dt"], s["rec_dx"], s["src_z"], s["rec_z"])
    return man, m, shots


def test_layered3_pipeline():
    """rtm_synth (modelling, store 1) then mute + rtm (order 8, nb 60, store 2)."""
    man, m, shots = layered3_shots()
    s = man["settings"]
    mp = O.Params(nb=s["model_nb"], order=s["model_order"], store_interval=1, f0=s["f0"])
    orc = O.Oracle(m, mp, s["nt"], s["dt"], variant="cpu-opt")
    traces = [orc.model_shot(sh) for sh in shots]
    assert G.sha256(G.rtms_bytes(shots, traces)) == man["sha256"]["shots.rtms"]
    for sh, tr in zip(shots, traces):
        sh.traces = tr
        sh.traces = O.mute_direct_wave(sh, s["mute_direct"], np.float32(1) / np.float32(s["f0"]),
                                       np.float32(1) / np.float32(s["f0"]))
    p = O.Params(nb=s["nb"], order=s["order"], store_interval=s["store_interval"], f0=s["f0"])
    image, illum = O.Oracle(m, p, s["nt"], s["dt"], variant="cpu-opt").migrate(shots)
    _, _, ref_image, ref_illum = G.layered3()
    assert bits_equal(image, ref_image)
    assert bits_equal(illum, ref_illum)


def test_marmousi_first_shot_checksum():
    """One of the 12 Marmousi shots (scripts/dataset_marmousi.env:21-31): the
    reference build's traces, whose 12-shot RTMS file reproduces the
    reference-held sha256."""
    import rtm_amd.models as M
    man, v = G.marmousi()
    assert man["reproduced_matches_recorded"]
    s = man["settings"]
    m = O.Model(s["nx"], s["nz"], s["dx"], s["dx"], v)
    shots = M.synth_line(m, s["nshots"], s["nt"], s["dt"], s["rec_dx"], s["src_z"], s["rec_z"])
    p = O.Params(nb=s["nb"], order=s["order"], store_interval=1, f0=s["f0"])
    tr = O.Oracle(m, p, s["nt"], s["dt"], variant="cpu-opt").model_shot(shots[0])
    assert G.sha256(tr.astype(np.float32).tobytes()) == man["per_shot_traces_sha256"][0]


def test_layered3_postprocessing():
    """The oracle's illumination compensation and Laplacian filter, applied to
    the reference image and illumination map, reproduce the post-processed
    images the reference CLI wrote (--illumination, --filter laplacian, both;
    src/main.cpp:211-222)."""
    man, _, ref_image, ref_illum = G.layered3()
    comp = O.illumination_compensation(ref_image, ref_illum, 1e-4)
    filt = O.laplacian_filter(ref_image, 201, 151, 10.0, 10.0)
    both = O.laplacian_filter(comp, 201, 151, 10.0, 10.0)
    assert G.sha256(comp.tobytes()) == man["sha256"]["illum.f32"]
    assert G.sha256(filt.tobytes()) == man["sha256"]["filtered.f32"]
    assert G.sha256(both.tobytes()) == man["sha256"]["filtered_illum.f32"]
