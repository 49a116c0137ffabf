"""HIP engine vs the CPU oracle on small grids: bit-identical traces and images.

The oracle (oracle/liboracle.so) is itself bit-identical to the reference's
-O2 build (tests/test_oracle_vs_reference.py), so equality here is equality
with the reference's CPU path.  Each case prints its parameters first so a
failure names it.
"""
import numpy as np
import pytest

from conftest import bits_equal, line_shots, small_model
import oracle_lib as O

pytestmark = pytest.mark.gpu


def _case(order, store, nt=160, dt=0.001):
    model = small_model()
    import rtm_amd
    params = rtm_amd.RTMParams(nb=max(order // 2 + 1, 12), order=order, store_interval=store, f0=15.0)
    shots = line_shots(model, nt, dt, [100.0, 250.0])
    return model, params, shots


@pytest.mark.parametrize("order", [2, 4, 6, 8, 10, 12])
def test_model_shot_bit_identical(hip_engine_cls, order):
    print(f"case: model_shot order={order}", flush=True)
    model, params, shots = _case(order, 1)
    orc = O.Oracle(model, params, shots[0].nt, shots[0].dt)
    with hip_engine_cls() as eng:
        eng.setup(model, params, shots[0].nt, shots[0].dt)
        for s in shots:
            ref = orc.model_shot(s)
            got = eng.model_shot(s)
            assert bits_equal(ref, got), f"max diff {np.abs(ref - got).max()}"


KERNELS = [(-1, 0), (0, 4), (0, 8), (0, 16), (1, 4), (1, 8), (1, 16), (2, 2), (2, 3), (2, 4), (2, 5), (2, 6), (2, 8), (2, 16),
           (3, 8), (3, 16), (3, 32)]  # (family, tile_x)


@pytest.mark.parametrize("family,tile", KERNELS)
@pytest.mark.parametrize("order,store", [(2, 1), (4, 3), (6, 2), (8, 2), (10, 5), (12, 1), (8, 7)])
def test_migrate_bit_identical(hip_engine_cls, order, store, family, tile):
    print(f"case: migrate order={order} store={store} family={family} tile={tile}", flush=True)
    model, params, shots = _case(order, store)
    orc = O.Oracle(model, params, shots[0].nt, shots[0].dt)
    for s in shots:
        s.traces = orc.model_shot(s)
    I_ref, S_ref = orc.migrate(shots)
    with hip_engine_cls() as eng:
        eng.setup(model, params, shots[0].nt, shots[0].dt)
        eng.set_step_kernel(family, tile)
        I, S = eng.migrate(shots)
    assert bits_equal(I_ref, I), f"image max diff {np.abs(I_ref - I).max()} of {np.abs(I_ref).max()}"
    assert bits_equal(S_ref, S)


@pytest.mark.parametrize("family,tile", [(0, 16), (1, 16), (2, 4), (2, 16), (3, 32)])
def test_dense_receivers_bit_identical(hip_engine_cls, family, tile):
    """More than 64 injection entries in one tile (a 12 x 12 patch of
    receivers, some sharing cells), so entries past the first 64 of a tile
    take the in-epilogue path; modelling and migration stay bit-identical."""
    import rtm_amd
    print(f"case: dense receivers family={family} tile={tile}", flush=True)
    model = small_model(61, 47)
    params = rtm_amd.RTMParams(nb=12, order=8, store_interval=2, f0=15.0)
    nt, dt = 140, 0.001
    gx, gz = np.meshgrid(np.arange(12) * 10.0 + 150.0, np.arange(12) * 10.0 + 60.0, indexing="ij")
    rx = np.concatenate([gx.ravel(), [150.0, 151.0, 260.0]]).astype(np.float32)   # duplicate cells
    rz = np.concatenate([gz.ravel(), [60.0, 61.0, 170.0]]).astype(np.float32)
    shot = rtm_amd.ShotRecord(300.0, 20.0, rx, rz, nt, dt)
    orc = O.Oracle(model, params, nt, dt)
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, dt)
        eng.set_step_kernel(family, tile)
        tr = eng.model_shot(shot)
        ref = orc.model_shot(shot)
        assert bits_equal(ref, tr)
        shot.traces = ref
        I, S = eng.migrate([shot])
    I_ref, S_ref = orc.migrate([shot])
    assert bits_equal(I_ref, I) and bits_equal(S_ref, S)


@pytest.mark.parametrize("family,tile", KERNELS)
def test_source_on_receiver_cell_bit_identical(hip_engine_cls, family, tile):
    """The source on the cell of two receivers: modelling injects, then records
    both traces at that cell, in entry order; a tile whose entries hit several
    of its columns and one cell more than once.  Traces and image are
    bit-identical to the oracle for every step kernel."""
    import rtm_amd
    print(f"case: source on a receiver cell family={family} tile={tile}", flush=True)
    model = small_model(41, 31)
    params = rtm_amd.RTMParams(nb=12, order=8, store_interval=3, f0=15.0)
    nt, dt = 120, 0.001
    rx = np.array([200.0, 201.0, 50.0, 120.0, 330.0, 210.0, 190.0], np.float32)
    rz = np.array([40.0, 41.0, 40.0, 20.0, 60.0, 40.0, 40.0], np.float32)
    shot = rtm_amd.ShotRecord(200.0, 40.0, rx, rz, nt, dt)
    orc = O.Oracle(model, params, nt, dt)
    ref = orc.model_shot(shot)
    assert np.abs(ref[0]).max() > 0 and bits_equal(ref[0], ref[1])
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, dt)
        eng.set_step_kernel(family, tile)
        tr = eng.model_shot(shot)
        assert bits_equal(ref, tr), f"trace max diff {np.abs(ref - tr).max()}"
        shot.traces = ref
        I, S = eng.migrate([shot])
    I_ref, S_ref = orc.migrate([shot])
    assert bits_equal(I_ref, I) and bits_equal(S_ref, S)


@pytest.mark.parametrize("fmt", [0, 1])
def test_model_shot_leaves_snapshot_store(hip_engine_cls, fmt):
    """Modelling stores no snapshots (the reference's model_shot passes no
    snapshot store, src/rtm/rtm.cpp:105-111): a backward pass after
    forward(A), model_shot(B) images A exactly as migrate([A])."""
    print(f"case: model_shot between forward and backward, snapshot format {fmt}", flush=True)
    model, params, shots = _case(8, 2)
    with hip_engine_cls() as eng:
        eng.setup(model, params, shots[0].nt, shots[0].dt, snapshot_format=fmt)
        for s in shots:
            s.traces = eng.model_shot(s)
        I_ref, S_ref = eng.migrate([shots[0]])
        eng.image_zero()
        eng.forward(shots[0])
        tr = eng.model_shot(shots[1])
        eng.backward(shots[0])
        I, S = eng.image()
    assert bits_equal(tr, shots[1].traces)
    assert bits_equal(I_ref, I) and bits_equal(S_ref, S)
