"""Host-side workload logic, checked without a GPU: the seeded model
generators are deterministic and have the documented ranges, the crop and
Poisson request generators follow their stated distributions, the FIFO
latency of the serial baseline is computed as a single server would give it,
and the bench's algorithmic byte counts match the closed forms."""
import importlib.util
import os

import numpy as np
import pytest

from conftest import REPO


def test_models_deterministic_and_in_range():
    from rtm_amd import models
    a = models.faulted_layers(301, 151, 4.0, 4.0, seed=11)
    b = models.faulted_layers(301, 151, 4.0, 4.0, seed=11)
    c = models.faulted_layers(301, 151, 4.0, 4.0, seed=12)
    assert np.array_equal(a.v, b.v) and not np.array_equal(a.v, c.v)
    # 1500 -> 5500 m/s layers scaled by at most +-5 %
    assert a.v.min() >= 1500 * 0.95 - 1e-3 and a.v.max() <= 5500 * 1.05 + 1e-3
    # velocity rises with depth on average
    col = a.v.reshape(301, 151).mean(axis=0)
    assert col[-10:].mean() > col[:10].mean() + 2000
    s = models.salt_body(401, 121, 5.0, 5.0, seed=3)
    assert np.isclose(s.v.max(), 4500.0) and s.v.min() >= 1500.0 - 1e-3
    l3 = models.layered_3()
    v = l3.v.reshape(201, 151)
    assert set(np.unique(v)) == {1500.0, 2200.0, 3000.0}
    assert v[0, 49] == 1500.0 and v[0, 50] == 2200.0 and v[0, 100] == 3000.0


def test_crop_batch_distribution():
    from rtm_amd import serve
    full = serve.models.faulted_layers(2301, 75, 4.0, 4.0)       # shallow stand-in: same widths, cheap
    specs, reqs = serve.crop_batch(64, seed=64, full=full)
    assert len(specs) == len(reqs) == 64
    widths = np.array([sp.model.nx for sp in specs.values()])
    nts = np.array([sp.nt for sp in specs.values()])
    orders = [sp.params.order for sp in specs.values()]
    assert widths.min() >= 400 and widths.max() <= 1200
    assert nts.min() >= 2000 and nts.max() <= 8000
    assert set(orders) == {8, 12}
    for sp in specs.values():
        m = sp.model
        assert m.ox >= 0 and m.ox + (m.nx - 1) * m.dx <= (full.nx - 1) * full.dx + 1e-3
        s = sp.shots[0]
        assert m.ox <= s.sx <= m.ox + (m.nx - 1) * m.dx
        assert s.rx.min() >= m.ox - 1e-3 and s.rx.max() <= m.ox + (m.nx - 1) * m.dx + 1e-3
    again, _ = serve.crop_batch(64, seed=64, full=full)
    assert [sp.nt for sp in again.values()] == list(nts)


def test_served_stream_poisson():
    from rtm_amd import serve
    specs = {k: serve.ModelSpec(None, None, 100, 0.001) for k in ("a", "b", "c")}
    for sp in specs.values():
        sp.shots = [object(), object()]
    reqs = serve.served_stream(specs, 50.0, 4000, seed=1)
    t = np.array([r.arrival for r in reqs])
    gaps = np.diff(np.concatenate([[0.0], t]))
    assert np.all(gaps > 0)
    assert abs(gaps.mean() - 1 / 50.0) < 0.1 / 50.0           # mean inter-arrival 1/rate
    assert abs(gaps.std() / gaps.mean() - 1.0) < 0.1           # exponential: CV 1
    keys = [r.key for r in reqs]
    for k in specs:
        assert abs(keys.count(k) / len(keys) - 1 / 3) < 0.05
    assert [r.arrival for r in serve.served_stream(specs, 50.0, 10, seed=1)] == list(t[:10])


def test_fifo_latency_of_serial_baseline():
    """serial_plain's queueing: done_k = max(arrival_k, done_{k-1}) + service_k."""
    from rtm_amd import serve
    reqs = [serve.Request("a", None, a) for a in (0.0, 0.1, 0.15, 2.0)]
    service = [0.5, 0.2, 0.1, 0.3]
    done = 0.0
    for r, s in zip(reqs, service):
        done = max(done, r.arrival) + s
        r.latency = done - r.arrival
    assert [round(r.latency, 6) for r in reqs] == [0.5, 0.6, 0.65, 0.3]
    summ = serve.summary(reqs, 2.3, {"a": serve.ModelSpec(serve.models.layered_3(41, 31),
                                                           serve.RTMParams(nb=10, order=4), 50, 0.001)})
    assert summ["jobs"] == 4 and summ["latency_p50_s"] == pytest.approx(np.percentile([0.5, 0.6, 0.65, 0.3], 50))


def test_bench_algorithmic_bytes():
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(REPO, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    from rtm_amd import RTMParams, VelocityModel
    m = VelocityModel(2301, 751, 4.0, 4.0, np.zeros(2301 * 751, np.float32))
    p = RTMParams(nb=60, order=8, store_interval=4)
    n_ext, per_shot, fwd, bwd = bench.algorithmic_bytes(m, p, 8000)
    assert n_ext == 2421 * 871
    nint, nsnap = 2301 * 751, 2000
    assert fwd == pytest.approx(16 * n_ext + nsnap * nint * 4 / 8000)
    assert bwd == pytest.approx(16 * n_ext + nsnap * nint * 20 / 8000)
    assert round(bwd) == 42379311
    assert per_shot == pytest.approx(8000 * (fwd + bwd))
    # 16-bit snapshots: 2 bytes per stored value
    _, per16, fwd16, bwd16 = bench.algorithmic_bytes(m, p, 8000, 2)
    assert fwd - fwd16 == pytest.approx(nsnap * nint * 2 / 8000)
    assert bwd - bwd16 == pytest.approx(nsnap * nint * 2 / 8000)
