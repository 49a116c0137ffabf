"""The oracle against the reference itself, compiled from its own sources
(oracle/_ref, built by oracle/Makefile where the reference sources exist):
bit-identical images, traces and snapshots on seeded random cases, for both
the "cpu" (src/rtm/*.cpp) and "cpu-opt" (src/cpu_opt/cpu_opt.cpp) engines."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import bits_equal

pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (no reference sources)")


def random_case(seed):
    rng = np.random.default_rng(seed)
    order = int(rng.choice([2, 4, 6, 8, 10, 12]))
    nx, nz = int(rng.integers(25, 48)), int(rng.integers(20, 40))
    nb = int(rng.integers(order // 2 + 1, order // 2 + 12))
    dx, dz = float(rng.choice([5.0, 8.0, 10.0])), float(rng.choice([5.0, 8.0, 10.0]))
    v = rng.uniform(1500.0, 2500.0, (nx, nz)).astype(np.float32)
    m = O.Model(nx, nz, dx, dz, v.reshape(-1), ox=float(rng.choice([0.0, 13.0])))
    store = int(rng.integers(1, 6))
    p = O.Params(nb=nb, order=order, store_interval=store, f0=float(rng.uniform(10, 20)),
                 sponge_alpha=float(rng.uniform(0.003, 0.01)))
    dt = float(0.8 * O.lib().orc_max_stable_dt(2500.0, dx, dz, order))
    nt = int(rng.integers(60, 140))
    nrec = int(rng.integers(1, 40))
    rx = rng.uniform(-30.0, nx * dx + 30.0, nrec).astype(np.float32)     # some clamp
    rz = rng.uniform(0.0, 4 * dz, nrec).astype(np.float32)
    if nrec > 3:
        rx[1], rz[1] = rx[0], rz[0]                                     # duplicate cell
    shots = [O.Shot(float(rng.uniform(0, nx * dx)), float(rng.uniform(0, 3 * dz)), rx, rz, nt, dt)
             for _ in range(int(rng.integers(1, 3)))]
    return m, p, nt, dt, shots


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("variant", ["cpu", "cpu-opt"])
def test_random_cases_bit_identical(seed, variant):
    m, p, nt, dt, shots = random_case(seed)
    ref_tr = O.ref_model_shots(variant, m, p, nt, dt, shots)
    orc = O.Oracle(m, p, nt, dt, variant=variant)
    for s, t in zip(shots, ref_tr):
        assert bits_equal(orc.model_shot(s), t)
        s.traces = t
    I_ref, S_ref, _ = O.ref_migrate(variant, m, p, nt, dt, shots)
    I, S = orc.migrate(shots)
    assert bits_equal(I, I_ref) and bits_equal(S, S_ref)
    snap_ref = O.ref_forward_snapshots(variant, m, p, nt, dt, shots[0])
    snap, _ = orc.forward(shots[0], snapshots=True)
    assert bits_equal(snap, snap_ref)


def test_mute_matches_reference():
    rng = np.random.default_rng(7)
    nt, dt = 300, 0.0008
    rx = rng.uniform(0, 2000, 50).astype(np.float32)
    rz = rng.uniform(0, 40, 50).astype(np.float32)
    shot = O.Shot(700.0, 10.0, rx, rz, nt, dt, rng.standard_normal((50, nt)).astype(np.float32))
    for v, tp in [(1500.0, 1 / 12.0), (2000.0, 0.01), (1800.0, 0.0002)]:
        assert bits_equal(O.mute_direct_wave(shot, v, tp, tp), O.ref_mute(shot, v, tp, tp))


def source_on_receiver_case():
    """The source on the cell of two receivers (injection, then the two
    recordings, in entry order) and receivers on cells of their own."""
    rng = np.random.default_rng(21)
    nx, nz, dx, dz = 41, 31, 10.0, 10.0
    v = rng.uniform(1500.0, 2500.0, (nx, nz)).astype(np.float32)
    m = O.Model(nx, nz, dx, dz, v.reshape(-1))
    p = O.Params(nb=12, order=8, store_interval=3, f0=15.0)
    dt = float(0.8 * O.lib().orc_max_stable_dt(2500.0, dx, dz, 8))
    nt = 120
    rx = np.array([200.0, 201.0, 50.0, 120.0, 330.0], np.float32)
    rz = np.array([40.0, 41.0, 40.0, 20.0, 60.0], np.float32)
    return m, p, nt, dt, [O.Shot(200.0, 40.0, rx, rz, nt, dt)]


@pytest.mark.parametrize("variant", ["cpu", "cpu-opt"])
def test_source_on_receiver_cell_bit_identical(variant):
    m, p, nt, dt, shots = source_on_receiver_case()
    ref_tr = O.ref_model_shots(variant, m, p, nt, dt, shots)
    orc = O.Oracle(m, p, nt, dt, variant=variant)
    tr = orc.model_shot(shots[0])
    assert bits_equal(tr, ref_tr[0])
    assert np.abs(tr[0]).max() > 0 and bits_equal(tr[0], tr[1])     # the source cell is recorded twice
    shots[0].traces = ref_tr[0]
    I_ref, S_ref, _ = O.ref_migrate(variant, m, p, nt, dt, shots)
    I, S = orc.migrate(shots)
    assert bits_equal(I, I_ref) and bits_equal(S, S_ref)
