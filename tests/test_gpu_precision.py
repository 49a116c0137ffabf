"""Reduced-precision snapshots and the FP64 image stack (no reference
counterpart; the reference stores FP32 snapshots and stacks in FP32).

* 16-bit snapshot codes are exact: each decoded snapshot equals
  decode(encode(s)) of the oracle's FP32 snapshot s, where encode is
  round-to-nearest-even of s * 2^k to FP16/BF16 and 2^k is the power of two
  that keeps max|s| * 2^k below 2^15.  numpy's float16 cast and a bitwise
  BF16 rounding give the expected codes independently of the device.
* Images from 16-bit snapshots stay within the stated tolerance of the FP64
  oracle: relative L2 <= 1e-3 for FP16.  BF16 (8-bit mantissa) is reported
  and bounded at 1e-2.
* The FP64 stack of one shot equals the FP32 image; over several shots it
  agrees with the FP32 running sum to rounding.
* The 601 x 401 crop of the salt model at order 12 (nt 10000): the GPU's FP32
  image is bit-identical to the FP32 oracle (the reference's arithmetic), whose
  own distance to the FP64 oracle is 1.6e-4 here: FP32 leapfrog stepping at
  ~200 steps per period cancels about three of FP32's seven digits in
  2p^n - p^{n-1} every step, a phase drift that grows linearly in time.  FP16
  snapshots add < 1e-4 on top of that and stay <= 1e-3 from FP64.
"""
import numpy as np
import pytest

import oracle_lib as O
from conftest import bits_equal, line_shots, small_model

pytestmark = pytest.mark.gpu


def expected_codes(snap32, fmt):
    """Decode(encode(s)) of one FP32 snapshot, computed in numpy."""
    m = np.float32(np.abs(snap32).max())
    k = 0 if m == 0 else 15 - int(np.frexp(m)[1])
    s = np.ldexp(snap32.astype(np.float32), k).astype(np.float32)
    if fmt == "f16":
        dec = s.astype(np.float16).astype(np.float32)
    else:
        b = s.view(np.uint32).astype(np.uint64)
        b = (b + 0x7FFF + ((b >> 16) & 1)) >> 16
        dec = (b.astype(np.uint32) << 16).view(np.float32)
    return np.ldexp(dec, -k).astype(np.float32)


def rel_l2(a, ref):
    a = np.asarray(a, np.float64)
    ref = np.asarray(ref, np.float64)
    return float(np.linalg.norm(a - ref) / np.linalg.norm(ref))


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
@pytest.mark.parametrize("nt,store", [(202, 3), (200, 3), (120, 1)])
def test_snapshot_codes_exact(hip_engine_cls, fmt, nt, store):
    import rtm_amd
    print(f"case: snapshot codes fmt={fmt} nt={nt} store={store}", flush=True)
    model = small_model()
    params = rtm_amd.RTMParams(nb=20, order=8, store_interval=store, f0=15.0)
    shot = line_shots(model, nt, 0.001, [200.0])[0]
    orc = O.Oracle(model, params, nt, 0.001)
    ref_snaps, _ = orc.forward(shot, snapshots=True)
    ref_snaps = ref_snaps.reshape(orc.nsnap, -1)
    code = rtm_amd.SNAP_F16 if fmt == "f16" else rtm_amd.SNAP_BF16
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, 0.001, snapshot_format=code)
        got, _ = eng.forward(shot, snapshots=True)
    got = got.reshape(orc.nsnap, -1)
    for s in range(orc.nsnap):
        assert bits_equal(got[s], expected_codes(ref_snaps[s], fmt)), f"snapshot {s}"


def _pair(fmt_code, stack, nt=300, order=8, store=2, nshots=2):
    import rtm_amd
    model = small_model(61, 45)
    params = rtm_amd.RTMParams(nb=20, order=order, store_interval=store, f0=15.0)
    shots = line_shots(model, nt, 0.001, [150.0, 420.0][:nshots])
    orc = O.Oracle(model, params, nt, 0.001)
    for s in shots:
        s.traces = orc.model_shot(s)
    return model, params, shots, orc


@pytest.mark.parametrize("fmt", ["f16", "bf16"])
def test_16bit_image_accuracy(hip_engine_cls, fmt):
    import rtm_amd
    print(f"case: 16-bit image accuracy {fmt}", flush=True)
    model, params, shots, orc = _pair(None, None)
    I32, _ = orc.migrate(shots)
    I64, _ = O.Oracle(model, params, shots[0].nt, 0.001, precision=64).migrate(shots)
    code = rtm_amd.SNAP_F16 if fmt == "f16" else rtm_amd.SNAP_BF16
    with hip_engine_cls() as eng:
        eng.setup(model, params, shots[0].nt, 0.001, snapshot_format=code, image_stack=rtm_amd.STACK_F64)
        I16, _ = eng.migrate(shots)
    e64, e32 = rel_l2(I16, I64), rel_l2(I16, I32)
    print(f"  {fmt}: rel L2 vs FP64 oracle {e64:.3e}, vs FP32 image {e32:.3e}; "
          f"FP32 image vs FP64 {rel_l2(I32, I64):.3e}")
    assert e64 <= (1e-3 if fmt == "f16" else 1e-2)


def test_fp64_stack(hip_engine_cls):
    import rtm_amd
    print("case: FP64 stack", flush=True)
    model, params, shots, orc = _pair(None, None, nshots=2)
    I1, S1 = orc.migrate(shots[:1])
    I2, S2 = orc.migrate(shots)
    with hip_engine_cls() as eng:
        eng.setup(model, params, shots[0].nt, 0.001, image_stack=rtm_amd.STACK_F64)
        J1, T1 = eng.migrate(shots[:1])
        D1, _ = eng.image_f64()
        J2, T2 = eng.migrate(shots)
        D2, E2 = eng.image_f64()
    # one shot: the stack is 0 + (double)I_shot, which rounds back to I_shot
    assert bits_equal(J1, I1) and bits_equal(T1, S1)
    assert np.array_equal(D1, I1.astype(np.float64))
    # two shots: FP64 sum of two FP32 shot images vs the FP32 running sum
    assert rel_l2(J2, I2) < 1e-6 and rel_l2(T2, S2) < 1e-6
    assert rel_l2(D2, I2) < 1e-6


def test_salt_crop_against_fp64_oracle(hip_engine_cls):
    """Order 12, nb 80, store 4, nt 10000, dt 0.4 ms, 12 Hz on a 601 x 401
    crop (5 m) of the seeded salt model; one surface shot, 601 receivers."""
    import rtm_amd
    from rtm_amd import models
    print("case: salt crop 601x401 order 12 vs FP64 oracle", flush=True)
    full = models.salt_body()
    v = full.v.reshape(full.nx, full.nz)[1700:2301, :401]
    crop = rtm_amd.VelocityModel(601, 401, 5.0, 5.0, np.ascontiguousarray(v).reshape(-1))
    params = rtm_amd.RTMParams(nb=80, order=12, store_interval=4, f0=12.0)
    nt, dt = 10000, 0.0004
    shot = models.surface_shots(crop, nt, dt, [1500.0], rec_dx=5.0, src_z=10.0, rec_z=10.0)[0]
    with hip_engine_cls() as eng:
        eng.setup(crop, params, nt, dt)
        shot.traces = eng.model_shot(shot)
        I32, _ = eng.migrate([shot])
        eng.setup(crop, params, nt, dt, snapshot_format=rtm_amd.SNAP_F16, image_stack=rtm_amd.STACK_F64)
        I16, _ = eng.migrate([shot])
    Iref, _ = O.Oracle(crop, params, nt, dt, variant="cpu-opt").migrate([shot])
    I64, _ = O.Oracle(crop, params, nt, dt, precision=64).migrate([shot])
    e32, e16, q16 = rel_l2(I32, I64), rel_l2(I16, I64), rel_l2(I16, I32)
    print(f"  rel L2 vs FP64 oracle: FP32 snapshots {e32:.3e}, FP16 snapshots {e16:.3e}; "
          f"FP16 vs FP32 snapshots {q16:.3e}")
    assert bits_equal(I32, Iref)          # the reference's FP32 arithmetic, exactly
    assert e32 <= 3e-4                    # that arithmetic's own distance to FP64
    assert q16 <= 1e-4                    # what 16-bit snapshots add
    assert e16 <= 1e-3
