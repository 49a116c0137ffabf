"""Batched engine: every job's image in a heterogeneous batch is bit-identical
to the same shot migrated alone by the plain engine ("alone == in the mix").

Jobs differ in grid, stencil order, nt, store interval, snapshot format and
trace stride; max_active and the memory budget force jobs to wait and join
while others are mid-way (forward and backward phases in one launch).  The
plain engine's FP32 image is itself bit-identical to the oracle (and so to
the reference's CPU path), which the first test re-checks for F32 jobs.
Each case prints its parameters first so a failure names it.
"""
import numpy as np
import pytest

from conftest import bits_equal, line_shots, small_model
import oracle_lib as O

pytestmark = pytest.mark.gpu

# (nx, nz, order, nb, nt, store, snapshot format, extra trace samples)
JOBS = [
    (61, 41, 8, 20, 180, 2, 0, 0),
    (45, 33, 12, 14, 150, 1, 0, 7),
    (80, 29, 4, 12, 210, 3, 1, 0),
    (52, 47, 8, 16, 121, 4, 2, 0),
    (37, 25, 2, 8, 97, 1, 1, 3),
    (70, 38, 12, 18, 165, 5, 0, 0),
    (33, 64, 6, 10, 133, 2, 2, 0),
    (90, 31, 10, 16, 190, 7, 1, 0),
]


def _job(k, nx, nz, order, nb, nt, store, fmt, extra, seed=11):
    import rtm_amd
    model = small_model(nx, nz)
    rng = np.random.default_rng(seed + k)
    model.v = (model.v.reshape(nx, nz) * rng.uniform(0.95, 1.05, (nx, nz))).astype(np.float32).reshape(-1)
    params = rtm_amd.RTMParams(nb=nb, order=order, store_interval=store, f0=15.0)
    dt = 0.001
    sx = float(rng.uniform(0, (nx - 1) * model.dx))
    shot = line_shots(model, nt + extra, dt, [sx])[0]
    shot.traces = (rng.standard_normal((shot.nrec, nt + extra)) * 1e-2).astype(np.float32)
    return model, params, nt, dt, shot, fmt


def _alone(hip_engine_cls, model, params, nt, dt, shot, fmt):
    with hip_engine_cls() as eng:
        eng.setup(model, params, nt, dt, snapshot_format=fmt)
        return eng.migrate([shot])


@pytest.fixture(scope="module")
def jobs():
    return [_job(k, *j) for k, j in enumerate(JOBS)]


@pytest.fixture(scope="module")
def alone(hip_engine_cls, jobs):
    return [_alone(hip_engine_cls, *j) for j in jobs]


def test_alone_matches_oracle_for_f32(jobs, alone):
    for k, (model, params, nt, dt, shot, fmt) in enumerate(jobs):
        if fmt != 0:
            continue
        print(f"case: job {k} alone vs oracle", flush=True)
        I_ref, S_ref = O.Oracle(model, params, nt, dt).migrate([shot])
        assert bits_equal(I_ref, alone[k][0])
        assert bits_equal(S_ref, alone[k][1])


@pytest.mark.parametrize("merge", [False, True])
@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("max_active", [0, 3])
def test_batch_equals_alone(jobs, alone, use_graph, max_active, merge):
    """Per-group launches and the one merged launch per tick (every order
    2-12 and snapshot width in one kernel) give the same bits."""
    import rtm_amd
    print(f"case: batch of {len(jobs)} jobs graph={use_graph} max_active={max_active} merged={merge}", flush=True)
    with rtm_amd.BatchEngine(0, max_active=max_active, use_graph=use_graph, merge_groups=merge) as b:
        ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
        b.run()
        st = b.stats()
        assert st["jobs_done"] == len(jobs) and st["jobs_active"] == 0 and st["jobs_queued"] == 0
        # 2 nt steps, plus one bookkeeping tick between the phases of a job
        # with 16-bit snapshots
        longest = max(2 * j[2] + (j[5] != 0) for j in jobs)
        if max_active == 0:
            assert st["ticks"] == longest          # every job ran from the first tick
        else:
            assert st["ticks"] > longest
        if max_active:
            assert st["rebuilds"] > 2      # jobs joined mid-run
            # the jobs that waited were admitted ahead, while the ticks up to
            # the finish that let them in ran
            assert st["staged_admissions"] == len(jobs) - max_active
        if merge:
            assert st["merged_launch"] in (0, 1) and all(g["tile_x"] in (2, 3, 4) for g in st["groups"])
        for k, jid in enumerate(ids):
            state, lat = b.state(jid)
            assert state == rtm_amd.JOB_DONE and lat > 0
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}: image max diff {np.abs(alone[k][0] - I).max()}"
            assert bits_equal(alone[k][1], S), f"job {k}: illumination differs"


@pytest.mark.parametrize("family,tile", [(1, 4), (2, 5), (2, 16), (3, 8), (3, 32)])
def test_batch_kernel_families(jobs, alone, family, tile):
    """Every batched kernel family and width gives the plain engine's bits."""
    import rtm_amd
    print(f"case: batch with kernel family {family} tile {tile}", flush=True)
    with rtm_amd.BatchEngine(0, kernel_family=family, tile_x=tile) as b:
        ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
        b.run()
        assert all(g["tile_x"] == tile for g in b.stats()["groups"])
        for k, jid in enumerate(ids):
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}"
            assert bits_equal(alone[k][1], S), f"job {k}"


@pytest.mark.parametrize("max_ticks", [0, 37])
def test_concurrent_streams_equal_alone(hip_engine_cls, max_ticks):
    """Two or three jobs, one of which fills the chip (>= 1 M extended cells),
    run each on the per-shot kernels on a stream of its own: the results are
    the plain engine's bits, for FP32, FP16 (codes of step nt-2 in the
    bookkeeping tick) and BF16 (slices of step nt-1) jobs; with max_ticks the
    runs stop and resume mid-way (one call's ticks left queued)."""
    import rtm_amd
    print(f"case: concurrent per-job streams, max_ticks={max_ticks}", flush=True)
    specs = [(1601, 601, 8, 20, 300, 4, 0, 0),      # 1641 x 641 extended cells
             (61, 41, 8, 20, 100, 2, 1, 0),        # (nt - 2) % store == 0
             (52, 47, 12, 16, 121, 4, 2, 3)]       # (nt - 1) % store == 0
    jobs = [_job(20 + k, *sp) for k, sp in enumerate(specs)]
    alone = [_alone(hip_engine_cls, *j) for j in jobs]
    with rtm_amd.BatchEngine(0) as b:
        ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
        if max_ticks:
            while b.stats()["jobs_done"] < len(jobs):
                b.run(max_ticks=max_ticks, until_done=False)
        else:
            b.run()
        st = b.stats()
        # the small jobs finish while the large one runs: concurrent ticks
        # until the second job finishes, then the large job alone
        assert st["concurrent_ticks"] >= 2 * 121 + 1
        for k, jid in enumerate(ids):
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}: image max diff {np.abs(alone[k][0] - I).max()}"
            assert bits_equal(alone[k][1], S), f"job {k}: illumination differs"


def test_interleaved_submission_and_budget(jobs, alone):
    """Jobs submitted while others run, under a budget that holds two jobs."""
    import rtm_amd
    print("case: interleaved submissions under a two-job budget", flush=True)
    with rtm_amd.BatchEngine(0) as probe:
        sizes = [probe.job_bytes(m, p, nt, dt, s, f) for (m, p, nt, dt, s, f) in jobs]
    budget = sorted(sizes)[-1] + sorted(sizes)[-2]
    with rtm_amd.BatchEngine(0, memory_budget=budget) as b:
        ids = {}
        for k in range(3):
            ids[k] = b.submit(*jobs[k][:5], snapshot_format=jobs[k][5])
        b.run(max_ticks=60, until_done=False)
        for k in range(3, len(jobs)):
            ids[k] = b.submit(*jobs[k][:5], snapshot_format=jobs[k][5])
            b.run(max_ticks=25, until_done=False)
            assert b.stats()["bytes_in_use"] <= budget
        b.run()
        for k, jid in ids.items():
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}"
            assert bits_equal(alone[k][1], S), f"job {k}"


def test_fp64_stack_at_offsets(jobs, alone):
    import rtm_amd
    print("case: FP64 stack of the batch at per-job offsets", flush=True)
    NX, NZ = 160, 90
    want_i = np.zeros((NX, NZ))
    want_s = np.zeros((NX, NZ))
    with rtm_amd.BatchEngine(0) as b:
        b.set_stack(NX, NZ)
        for k, (m, p, nt, dt, s, f) in enumerate(jobs):
            ix0, iz0 = (7 * k) % (NX - m.nx + 1), (5 * k) % (NZ - m.nz + 1)
            b.submit(m, p, nt, dt, s, snapshot_format=f, stack_at=(ix0, iz0))
            want_i[ix0:ix0 + m.nx, iz0:iz0 + m.nz] += alone[k][0].reshape(m.nx, m.nz).astype(np.float64)
            want_s[ix0:ix0 + m.nx, iz0:iz0 + m.nz] += alone[k][1].reshape(m.nx, m.nz).astype(np.float64)
        b.run()
        gi, gs = b.stack()
    # FP64 sums of the same FP32 addends; only the order of addition may differ
    np.testing.assert_allclose(gi.reshape(NX, NZ), want_i, rtol=1e-12, atol=1e-12 * np.abs(want_i).max())
    np.testing.assert_allclose(gs.reshape(NX, NZ), want_s, rtol=1e-12, atol=1e-12 * np.abs(want_s).max())


def test_batch_errors(jobs):
    import rtm_amd
    print("case: batch argument errors", flush=True)
    m, p, nt, dt, s, f = jobs[0]
    with rtm_amd.BatchEngine(0, memory_budget=1 << 20) as b:
        with pytest.raises(RuntimeError, match="budget"):
            b.submit(m, p, nt, dt, s)
    with rtm_amd.BatchEngine(0) as b:
        bad = rtm_amd.RTMParams(nb=p.nb, order=7, store_interval=1)
        with pytest.raises(RuntimeError, match="unsupported spatial order"):
            b.submit(m, bad, nt, dt, s)
        with pytest.raises(RuntimeError, match="fewer than nt"):
            b.submit(m, p, s.nt + 1, dt, s)
        with pytest.raises(RuntimeError, match="no stack is set"):
            b.submit(m, p, nt, dt, s, stack_at=(0, 0))
        b.set_stack(m.nx, m.nz)
        with pytest.raises(RuntimeError, match="does not fit"):
            b.submit(m, p, nt, dt, s, stack_at=(1, 0))
        jid = b.submit(m, p, nt, dt, s)
        with pytest.raises(RuntimeError, match="not finished"):
            b.result(jid)
        b.run()
        b.result(jid)
        with pytest.raises(RuntimeError, match="unknown job"):
            b.state(jid)


def test_resident_models_serve_many_shots(hip_engine_cls, jobs):
    """Shots on registered (resident) models: each equals the plain engine's
    migration of that one shot; models survive their jobs until removed."""
    import rtm_amd
    print("case: resident models, several shots each", flush=True)
    picks = [jobs[0], jobs[2]]
    with rtm_amd.BatchEngine(0) as b:
        mids, expect = [], {}
        for (m, p, nt, dt, s, f) in picks:
            mids.append(b.add_model(m, p, nt, dt, snapshot_format=f))
        resident_bytes = b.stats()["bytes_in_use"]
        assert resident_bytes > 0 and b.stats()["models_resident"] == 2
        for round_ in range(2):
            ids = []
            for k, (m, p, nt, dt, s, f) in enumerate(picks):
                for sx in (0.2, 0.7):
                    shot = rtm_amd.ShotRecord(sx * (m.nx - 1) * m.dx, s.sz, s.rx, s.rz, s.nt, s.dt,
                                              s.traces * (1.0 + round_))
                    jid = b.submit_shot(mids[k], shot)
                    ids.append(jid)
                    if jid not in expect:
                        with hip_engine_cls() as eng:
                            eng.setup(m, p, nt, dt, snapshot_format=f)
                            expect[jid] = eng.migrate([shot])
            b.run()
            assert b.stats()["bytes_in_use"] == resident_bytes     # jobs freed, models kept
            for jid in ids:
                I, S = b.result(jid)
                assert bits_equal(expect[jid][0], I), f"job {jid}"
                assert bits_equal(expect[jid][1], S), f"job {jid}"
        # removal with a shot still queued: the model lives until that shot is done
        m, p, nt, dt, s, f = picks[0]
        jid = b.submit_shot(mids[0], s)
        b.remove_model(mids[0])
        b.run()
        I, _ = b.result(jid)
        with hip_engine_cls() as eng:
            eng.setup(m, p, nt, dt, snapshot_format=f)
            assert bits_equal(eng.migrate([s])[0], I)
        b.remove_model(mids[1])
        st = b.stats()
        assert st["models_resident"] == 0 and st["bytes_in_use"] == 0
        with pytest.raises(RuntimeError, match="unknown model"):
            b.submit_shot(mids[0], s)


@pytest.mark.parametrize("use_graph", [True, False])
def test_single_short_job_after_autotune(jobs, alone, use_graph):
    """A lone job whose group is tuned after its first steps (the tuning
    rebuild follows ticks still in flight) keeps the plain engine's bits."""
    import rtm_amd
    print(f"case: lone job 4 with autotuning, graph={use_graph}", flush=True)
    for _ in range(3):
        with rtm_amd.BatchEngine(0, use_graph=use_graph) as b:
            jid = b.submit(*jobs[4][:5], snapshot_format=jobs[4][5])
            b.run()
            I, S = b.result(jid)
        assert bits_equal(alone[4][0], I)
        assert bits_equal(alone[4][1], S)


def test_retune_on_load_class_change(jobs, alone):
    """A small job joining a lone order-8 job moves the lone job's group from
    the per-shot kernels to the batch kernel (a new load class): the group is
    re-tuned mid-stream on dry steps and the joiner's own group (16-bit
    snapshots) is tuned on first sight; when the joiner retires, the first
    class's cached choice comes back without a search; both images keep the
    plain engine's bits."""
    import rtm_amd
    print("case: jobs 0 and 3 (order 8), job 3 joining at tick 40", flush=True)
    with rtm_amd.BatchEngine(0) as b:
        j0 = b.submit(*jobs[0][:5], snapshot_format=jobs[0][5])
        b.run(max_ticks=40, until_done=False)
        assert b.stats()["tunes"] == 1
        j3 = b.submit(*jobs[3][:5], snapshot_format=jobs[3][5])
        b.run()
        assert b.stats()["tunes"] == 3
        for k, jid in ((0, j0), (3, j3)):
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}"
            assert bits_equal(alone[k][1], S), f"job {k}"


@pytest.mark.parametrize("fmt", [0, 1])
def test_dense_receivers_in_batch(hip_engine_cls, jobs, fmt):
    """More than 64 receiver entries in one tile (a 12 x 12 patch, some
    sharing cells) beside an ordinary job: entries past the first 64 of a
    tile, and their amplitudes in entry order, keep the plain engine's bits."""
    import rtm_amd
    print(f"case: dense receivers in a batch, snapshot format {fmt}", flush=True)
    model = small_model(61, 47)
    params = rtm_amd.RTMParams(nb=12, order=8, store_interval=2, f0=15.0)
    nt, dt = 140, 0.001
    gx, gz = np.meshgrid(np.arange(12) * 10.0 + 150.0, np.arange(12) * 10.0 + 60.0, indexing="ij")
    rx = np.concatenate([gx.ravel(), [150.0, 151.0, 260.0]]).astype(np.float32)
    rz = np.concatenate([gz.ravel(), [60.0, 61.0, 170.0]]).astype(np.float32)
    shot = rtm_amd.ShotRecord(300.0, 20.0, rx, rz, nt + 5, dt)
    rng = np.random.default_rng(7)
    shot.traces = (rng.standard_normal((len(rx), nt + 5)) * 1e-2).astype(np.float32)
    I0, S0 = _alone(hip_engine_cls, model, params, nt, dt, shot, fmt)
    with rtm_amd.BatchEngine(0) as b:
        other = b.submit(*jobs[0][:5], snapshot_format=jobs[0][5])
        jid = b.submit(model, params, nt, dt, shot, snapshot_format=fmt)
        b.run()
        I, S = b.result(jid)
        b.result(other)
    assert bits_equal(I0, I) and bits_equal(S0, S)


@pytest.mark.parametrize("nt,store", [(1, 1), (2, 1), (2, 2), (3, 1), (3, 2), (4, 3)])
def test_short_16bit_jobs(hip_engine_cls, nt, store):
    """16-bit jobs whose last stored steps fall on the forward phase's last
    two steps, including nt = 1 and 2 (the bookkeeping tick settles every
    scale), equal the plain engine alone, with and without a longer job."""
    import rtm_amd
    print(f"case: short 16-bit job nt={nt} store={store}", flush=True)
    model, params, _, dt, shot, _ = _job(3, 40, 30, 8, 10, nt, store, 1, 2)
    I0, S0 = _alone(hip_engine_cls, model, params, nt, dt, shot, 1)
    for with_other in (False, True):
        with rtm_amd.BatchEngine(0) as b:
            if with_other:
                b.submit(*_other_job_args(), snapshot_format=1)
            jid = b.submit(model, params, nt, dt, shot, snapshot_format=1)
            b.run()
            I, S = b.result(jid)
        assert bits_equal(I0, I) and bits_equal(S0, S), f"with_other={with_other}"


def _other_job_args():
    return _job(9, 52, 47, 8, 16, 121, 4, 1, 0)[:5]


def test_reserved_arena(jobs, alone):
    """Jobs placed in a reserved block: a block that holds only a few jobs at
    a time makes later jobs wait for memory (and for a free run of it), and
    every image still equals the plain engine's."""
    import rtm_amd
    print("case: batch in a reserved arena", flush=True)
    with rtm_amd.BatchEngine(0) as probe:
        sizes = sorted(probe.job_bytes(m, p, nt, dt, s, f) for (m, p, nt, dt, s, f) in jobs)
    for reserve in (int(sum(sizes[-3:]) * 1.05), 1 << 30):
        with rtm_amd.BatchEngine(0) as b:
            b.reserve(reserve)
            assert b.stats()["bytes_budget"] <= reserve
            ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
            with pytest.raises(RuntimeError, match="before submitting"):
                b.reserve(1 << 20)
            b.run()
            assert b.stats()["bytes_in_use"] == 0
            for k, jid in enumerate(ids):
                I, S = b.result(jid)
                assert bits_equal(alone[k][0], I), f"job {k}, reserve {reserve}"
                assert bits_equal(alone[k][1], S), f"job {k}, reserve {reserve}"


@pytest.mark.parametrize("use_graph", [True, False])
@pytest.mark.parametrize("fmt", [0, 1])
def test_lean_tick_patterns(hip_engine_cls, fmt, use_graph):
    """Ticks in which no job of a group stores or images run the lean build.
    Three jobs of one group with store intervals 3, 5 and 7 and different nt
    make the lean/full pattern of almost every 32-tick run a new one, so the
    per-pattern graph cache fills and is emptied mid-batch; every image equals
    the job's migration alone."""
    import rtm_amd
    print(f"case: lean tick patterns, snapshot format {fmt}, graph={use_graph}", flush=True)
    specs = [(48, 36, 8, 12, 300, 3), (52, 30, 8, 12, 331, 5), (44, 40, 8, 12, 362, 7)]
    jobs = [_job(20 + k, nx, nz, order, nb, nt, store, fmt, 0) for k, (nx, nz, order, nb, nt, store) in
            enumerate(specs)]
    alone = [_alone(hip_engine_cls, *j) for j in jobs]
    with rtm_amd.BatchEngine(0, use_graph=use_graph) as b:
        ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
        b.run()
        st = b.stats()
        assert st["jobs_done"] == len(jobs)
        assert 0 < st["lean_launches"] < st["step_launches"]
        if use_graph and fmt == 0:
            assert st["graphs_captured"] > 8           # the pattern cache was emptied mid-batch
        for k, jid in enumerate(ids):
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}: image max diff {np.abs(alone[k][0] - I).max()}"
            assert bits_equal(alone[k][1], S), f"job {k}: illumination differs"


@pytest.mark.parametrize("tile", [2, 3, 4])
def test_merged_launch_widths(jobs, alone, tile):
    """The merged launch at each of its tile widths, pinned (DPP march)."""
    import rtm_amd
    print(f"case: merged launch, tile {tile}", flush=True)
    with rtm_amd.BatchEngine(0, kernel_family=rtm_amd.STEP_MARCH_DPP, tile_x=tile, merge_groups=True) as b:
        ids = [b.submit(m, p, nt, dt, s, snapshot_format=f) for (m, p, nt, dt, s, f) in jobs]
        b.run(max_ticks=50, until_done=False)
        assert b.stats()["merged_launch"] == 1          # jobs of several groups on the device
        b.run()
        for k, jid in enumerate(ids):
            I, S = b.result(jid)
            assert bits_equal(alone[k][0], I), f"job {k}"
            assert bits_equal(alone[k][1], S), f"job {k}"
