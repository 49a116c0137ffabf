"""Shared test setup: import paths, the `gpu` marker, small model builders."""
import os
import sys

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "rtm-acceleration_amd")
for p in (PKG, os.path.join(REPO, "tests")):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")
    config.addinivalue_line("markers", "slow: long CPU-only check (set RTM_SLOW=1 to run)")


def pytest_collection_modifyitems(config, items):
    if os.environ.get("RTM_SLOW") == "1":
        return
    skip = pytest.mark.skip(reason="slow check; set RTM_SLOW=1")
    for it in items:
        if "slow" in it.keywords:
            it.add_marker(skip)


@pytest.fixture(scope="session")
def hip_engine_cls():
    """The product engine class; fails loudly when the library or device is missing."""
    import rtm_amd
    if rtm_amd.device_count() < 1:
        pytest.fail("no HIP device visible to a test marked gpu")
    return rtm_amd.HipEngine


def small_model(nx=41, nz=31, dx=10.0, dz=10.0):
    """Two layers and a fast box: enough contrast for reflections."""
    import rtm_amd
    v = np.full((nx, nz), 1500.0, np.float32)
    v[:, nz // 2:] = 2200.0
    v[nx // 2: nx // 2 + 5, nz - 10: nz - 6] = 2800.0
    return rtm_amd.VelocityModel(nx, nz, dx, dz, v.reshape(-1))


def line_shots(model, nt, dt, sxs, sz=20.0, rec_dx=10.0, rz=20.0, duplicate=True, clamp=True):
    """Surface line of receivers; optionally a duplicate cell and a clamped receiver."""
    import rtm_amd
    xmax = (model.nx - 1) * model.dx
    rx = np.arange(0.0, xmax + 1e-3, rec_dx, dtype=np.float32)
    if duplicate and rx.size > 6:
        rx[5] = rx[4] + 0.3 * model.dx          # rounds onto receiver 4's cell
    if clamp:
        rx = np.append(rx, np.float32(xmax + 7.0 * model.dx))   # off the model end
    rzz = np.full_like(rx, rz)
    return [rtm_amd.ShotRecord(sx, sz, rx, rzz, nt, dt) for sx in sxs]


def bits_equal(a, b):
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))
