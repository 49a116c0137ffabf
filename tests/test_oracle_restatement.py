"""An independent numpy float32 restatement of the reference step on tiny
grids, and invariants the oracle must keep.

numpy float32 array arithmetic rounds every operation to float32 without
contraction, so a restatement with the reference's association
(src/rtm/forward.cpp:38-45, 56-63) must match the oracle bit for bit.  It
uses the oracle's tables (whose closed forms are pinned separately)."""
import numpy as np
import pytest

import oracle_lib as O
from conftest import bits_equal


def numpy_forward(o, shot, nt, store):
    """Reference-order forward propagation in numpy float32; returns
    (snapshots, recorded traces, final three buffers)."""
    nxe, nze, nb, h = o.nxe, o.nze, o.nb, o.params.order // 2
    f32 = np.float32
    vdt2 = o.table("vdt2").reshape(nxe, nze)
    w = o.table("sponge").reshape(nxe, nze)
    wav = o.table("wavelet")
    c = o.table("coeffs")
    c0, cx, cz = c[0], c[1:h + 1], c[h + 1:]
    src, rec = o.map_geometry(shot)
    sx, sz = divmod(src, nze)
    pp = np.zeros((nxe, nze), f32)
    pc = np.zeros_like(pp)
    pn = np.zeros_like(pp)
    snaps, recs = [], np.zeros((len(rec), nt), f32)
    for it in range(nt):
        core = pc[h:nxe - h, h:nze - h]
        lap = c0 * core
        for k in range(1, h + 1):
            tx = cx[k - 1] * (pc[h + k:nxe - h + k, h:nze - h] + pc[h - k:nxe - h - k, h:nze - h])
            tz = cz[k - 1] * (pc[h:nxe - h, h + k:nze - h + k] + pc[h:nxe - h, h - k:nze - h - k])
            lap = lap + (tx + tz)
        pn[h:nxe - h, h:nze - h] = (f32(2.0) * core - pp[h:nxe - h, h:nze - h]) + \
            vdt2[h:nxe - h, h:nze - h] * lap
        pn[sx, sz] = pn[sx, sz] + vdt2[sx, sz] * wav[it]
        pc *= w
        pn *= w
        pp, pc, pn = pc, pn, pp
        for ir, idx in enumerate(rec):
            recs[ir, it] = pc.reshape(-1)[idx]
        if it % store == 0:
            snaps.append(pc[nb:nxe - nb, nb:nze - nb].copy())
    return np.stack(snaps), recs, (pp, pc, pn)


@pytest.mark.parametrize("order,store", [(2, 1), (4, 2), (8, 3), (12, 1)])
def test_numpy_restatement_matches_oracle(order, store):
    nx, nz, nt = 19, 15, 40
    rng = np.random.default_rng(order)
    v = rng.uniform(1500, 2500, (nx, nz)).astype(np.float32)
    m = O.Model(nx, nz, 10.0, 10.0, v.reshape(-1))
    p = O.Params(nb=order // 2 + 3, order=order, store_interval=store, f0=25.0)
    o = O.Oracle(m, p, nt, 0.001)
    rx = np.array([0.0, 50.0, 50.0, 120.0, 180.0], np.float32)
    shot = O.Shot(90.0, 30.0, rx, np.full_like(rx, 10.0), nt, 0.001)
    snaps_np, rec_np, (_, pc, _) = numpy_forward(o, shot, nt, store)
    snaps, rec = o.forward(shot, snapshots=True, record=True)
    assert bits_equal(snaps, snaps_np.reshape(-1))
    assert bits_equal(rec, rec_np)
    # the outer `half` ring is never written (src/rtm/forward.cpp:48-49)
    h = order // 2
    assert np.all(pc[:h] == 0) and np.all(pc[-h:] == 0) and np.all(pc[:, :h] == 0) and np.all(pc[:, -h:] == 0)


def _small(order=8, store=2, nt=120):
    nx, nz = 31, 25
    v = np.full((nx, nz), 1800.0, np.float32)
    v[:, 12:] = 2600.0
    m = O.Model(nx, nz, 10.0, 10.0, v.reshape(-1))
    p = O.Params(nb=12, order=order, store_interval=store, f0=18.0)
    rx = np.arange(0, 300, 10, dtype=np.float32)
    shots = [O.Shot(150.0, 20.0, rx, np.full_like(rx, 20.0), nt, 0.001)]
    o = O.Oracle(m, p, nt, 0.001)
    shots[0].traces = o.model_shot(shots[0])
    return m, p, o, shots


def test_linearity_in_the_receiver_data():
    """Scaling the traces by 2 scales the receiver field, hence the image, by
    exactly 2 (a power-of-two scale commutes with rounding away from
    underflow); the illumination depends only on the source field."""
    m, p, o, shots = _small()
    I1, S1 = o.migrate(shots)
    s2 = O.Shot(shots[0].sx, shots[0].sz, shots[0].rx, shots[0].rz, shots[0].nt, shots[0].dt,
                shots[0].traces * np.float32(2))
    I2, S2 = o.migrate([s2])
    assert bits_equal(S1, S2)
    big = np.abs(I1) > 1e-30
    assert np.array_equal(I2[big], 2 * I1[big])


def test_zero_data_gives_zero_image_and_shots_add():
    m, p, o, shots = _small()
    z = O.Shot(shots[0].sx, shots[0].sz, shots[0].rx, shots[0].rz, shots[0].nt, shots[0].dt,
               np.zeros_like(shots[0].traces))
    I0, S0 = o.migrate([z])
    assert np.all(I0 == 0)
    assert np.all(S0 >= 0) and S0.max() > 0
    # migrate runs a running FP32 sum over shots in order (src/rtm/rtm.cpp:92-101)
    I1, S1 = o.migrate(shots)
    Ia, Sa = o.migrate([shots[0], z])
    assert bits_equal(Ia, I1)
    assert np.allclose(Sa, 2 * S1, rtol=1e-6)


def test_fp64_agrees_with_fp32():
    m, p, o, shots = _small(nt=200)
    I32, _ = o.migrate(shots)
    I64, _ = O.Oracle(m, p, 200, 0.001, precision=64).migrate(shots)
    assert np.linalg.norm(I32 - I64) / np.linalg.norm(I64) < 1e-5
