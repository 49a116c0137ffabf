"""Error behaviour of the C ABI (the reference throws std::runtime_error for
the same conditions, src/rtm/rtm_cpu.cpp:15-22, src/rtm/rtm.cpp:33-34)."""
import numpy as np
import pytest

from conftest import small_model

pytestmark = pytest.mark.gpu


def test_setup_validation(hip_engine_cls):
    import rtm_amd
    m = small_model()
    with hip_engine_cls() as eng:
        with pytest.raises(RuntimeError, match="unsupported spatial order"):
            eng.setup(m, rtm_amd.RTMParams(nb=20, order=7), 10, 0.001)
        with pytest.raises(RuntimeError, match="nb must be at least"):
            eng.setup(m, rtm_amd.RTMParams(nb=4, order=8), 10, 0.001)
        with pytest.raises(RuntimeError, match="store-interval"):
            eng.setup(m, rtm_amd.RTMParams(nb=20, order=8, store_interval=0), 10, 0.001)
        with pytest.raises(RuntimeError, match="not set up"):
            eng.image()
        with pytest.raises(RuntimeError, match="snapshot format"):
            eng.setup(m, rtm_amd.RTMParams(nb=20, order=8), 10, 0.001, snapshot_format=9)
        with pytest.raises(RuntimeError, match="tile width"):
            eng.set_step_kernel(1, 5)
        with pytest.raises(RuntimeError, match="family"):
            eng.set_step_kernel(7, 8)
        with pytest.raises(RuntimeError, match="not available"):
            eng.set_step_kernel(3, 4)
        with pytest.raises(RuntimeError, match="not available"):
            eng.set_step_kernel(0, 5)
    with pytest.raises(RuntimeError, match="batch kernel family"):
        rtm_amd.BatchEngine(0, kernel_family=0, tile_x=8)
    with pytest.raises(RuntimeError, match="batch kernel family"):
        rtm_amd.BatchEngine(0, kernel_family=3, tile_x=4)


def test_shot_validation(hip_engine_cls):
    import rtm_amd
    m = small_model()
    rx = np.arange(0, 300, 10, dtype=np.float32)
    with hip_engine_cls() as eng:
        eng.setup(m, rtm_amd.RTMParams(nb=12, order=8), 100, 0.001)
        short = rtm_amd.ShotRecord(100.0, 20.0, rx, np.full_like(rx, 20.0), 50, 0.001,
                                   np.zeros((rx.size, 50), np.float32))
        with pytest.raises(RuntimeError, match="fewer than the engine's nt"):
            eng.backward(short)
        no_data = rtm_amd.ShotRecord(100.0, 20.0, rx, np.full_like(rx, 20.0), 100, 0.001)
        with pytest.raises(RuntimeError, match="needs traces"):
            eng.backward(no_data)
        # the engine keeps working after refused calls
        ok = rtm_amd.ShotRecord(100.0, 20.0, rx, np.full_like(rx, 20.0), 100, 0.001)
        tr = eng.model_shot(ok)
        assert np.abs(tr).max() > 0


def test_kernel_timing_modes(hip_engine_cls):
    """Sampled timing counts one launch per sampled step; loop timing counts
    nt launches per forward/backward loop; 0 records nothing."""
    import rtm_amd
    from conftest import line_shots
    m = small_model()
    nt = 64
    shot = line_shots(m, nt, 0.001, [150.0])[0]
    shot.traces = np.zeros((shot.nrec, nt), np.float32)
    with hip_engine_cls() as eng:
        eng.setup(m, rtm_amd.RTMParams(nb=12, order=8), nt, 0.001)
        for every, want in [(0, 0), (8, nt // 8), (-1, nt)]:
            eng.reset_times()
            eng.set_kernel_timing(every)
            eng.forward(shot)
            eng.backward(shot)
            ks = eng.kernel_stats()
            assert ks["launches"] == [nt, nt]
            assert ks["sampled"] == [want, want], (every, ks)
            assert all(ms > 0 for ms in ks["sampled_ms"]) == (want > 0)
