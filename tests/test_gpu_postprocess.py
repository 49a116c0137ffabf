"""Post-processing of the resident image on the device (illumination
compensation and the Laplacian filter, src/rtm/imaging.cpp:48-75) against the
oracle's restatement of the same functions, which the reference's own
post-processed outputs pin (test_oracle_golden.py):

* every flag combination, bit for bit, on a grid large enough that the
  maximum of S is reduced over several grid-stride passes;
* an engine with the FP64 stack: the source is the stack rounded to float,
  exactly what rtm_hip_image_get returns;
* the resident image is left unchanged, and a second call after a further
  migration uses the new image;
* bad flags and an engine that is not set up are errors.
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import bits_equal, line_shots

pytestmark = pytest.mark.gpu


def _model(nx, nz, dx, dz):
    import rtm_amd
    v = np.full((nx, nz), 1600.0, np.float32)
    v[:, nz // 3:] = 2400.0
    v[:, 2 * nz // 3:] = 3100.0
    return rtm_amd.VelocityModel(nx, nz, dx, dz, v.reshape(-1))


def _expected(image, illum, nx, nz, dx, dz, eps):
    comp = O.illumination_compensation(image, illum, eps)
    return {(True, False): comp,
            (False, True): O.laplacian_filter(image, nx, nz, dx, dz),
            (True, True): O.laplacian_filter(comp, nx, nz, dx, dz)}


@pytest.mark.parametrize("stack", ["f32", "f64"])
def test_postprocess_matches_oracle(hip_engine_cls, stack):
    import rtm_amd
    nx, nz, dx, dz, nt, dt = 601, 401, 8.0, 6.0, 160, 0.0008
    print(f"case: postprocess {nx}x{nz} stack {stack}", flush=True)
    m = _model(nx, nz, dx, dz)
    p = rtm_amd.RTMParams(nb=20, order=8, store_interval=2, f0=20.0)
    shots = line_shots(m, nt, dt, [1200.0, 3000.0], rec_dx=24.0)
    st = rtm_amd.STACK_F64 if stack == "f64" else rtm_amd.STACK_F32
    with hip_engine_cls() as eng:
        eng.setup(m, p, nt, dt, image_stack=st)
        for s in shots:
            s.traces = eng.model_shot(s)
        image, illum = eng.migrate(shots[:1])
        assert np.abs(image).max() > 0 and illum.max() > 0
        for eps in (1e-4, 0.3):
            for (ill, lap), want in _expected(image, illum, nx, nz, dx, dz, eps).items():
                got = eng.postprocess(illum=ill, laplacian=lap, eps_relative=eps)
                assert bits_equal(got, want), (ill, lap, eps)
        # the resident image is unchanged by post-processing
        again, again_s = eng.image()
        assert bits_equal(again, image) and bits_equal(again_s, illum)
        # a further migration replaces the image the next call works on
        image2, illum2 = eng.migrate(shots)
        got = eng.postprocess(illum=True, laplacian=True)
    assert not bits_equal(image2, image)
    assert bits_equal(got, _expected(image2, illum2, nx, nz, dx, dz, 1e-4)[(True, True)])


def test_postprocess_errors(hip_engine_cls):
    import rtm_amd
    from rtm_amd import _capi
    m = _model(41, 31, 10.0, 10.0)
    with hip_engine_cls() as eng:
        with pytest.raises(RuntimeError, match="not set up"):
            eng.postprocess()
        eng.setup(m, rtm_amd.RTMParams(nb=20, order=8), 10, 0.001)
        with pytest.raises(RuntimeError, match="post-processing flags"):
            eng.postprocess(illum=False, laplacian=False)
        with pytest.raises(RuntimeError, match="post-processing flags"):
            out = np.empty(41 * 31, np.float32)
            eng._check(eng.lib.rtm_hip_image_postprocess(eng.h, 4, 1e-4, _capi.fptr(out)))
        # an image of zeros: compensation gives 0 / 1e-30 = 0, the filter 0
        assert not eng.postprocess(illum=True, laplacian=True).any()
