"""The C ABI library, checked without a GPU: it loads, exports every entry
point include/rtm_hip.h declares, fails cleanly when no device is present
(as the reference's CUDA engine does, src/cuda/rtm_cuda.cu:33), and its
host-only mute matches the oracle bit for bit."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle_lib as O
from conftest import REPO, bits_equal

HEADER = os.path.join(REPO, "include", "rtm_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rtm_hip_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import rtm_amd
    lib = C.CDLL(rtm_amd.library_path())
    names = declared_functions()
    assert len(names) >= 18
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding covers exactly the declared set
    from rtm_amd import _capi
    assert sorted(_capi.SIGNATURES) == names


@pytest.mark.skipif(os.path.exists("/dev/kfd"), reason="a GPU is present")
def test_no_device_fails_cleanly():
    import rtm_amd
    assert rtm_amd.device_count() == 0
    with pytest.raises(RuntimeError, match="no HIP device"):
        rtm_amd.HipEngine(0)


def test_mute_matches_oracle():
    import rtm_amd
    rng = np.random.default_rng(3)
    nt, dt = 500, 0.001
    rx = rng.uniform(0, 2000, 40).astype(np.float32)
    rz = rng.uniform(0, 40, 40).astype(np.float32)
    tr = rng.standard_normal((40, nt)).astype(np.float32)
    for v, tp in [(1500.0, 1 / 12.0), (2500.0, 0.03)]:
        s = rtm_amd.ShotRecord(650.0, 20.0, rx, rz, nt, dt, tr.copy())
        rtm_amd.mute_direct_wave(s, v, tp, tp)
        ref = O.mute_direct_wave(O.Shot(650.0, 20.0, rx, rz, nt, dt, tr), v, tp, tp)
        assert bits_equal(s.traces, ref)


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes structure has the size and field offsets the C compiler
    gives the header's struct (a mismatch would pass garbage through the ABI)."""
    from rtm_amd import _capi
    pairs = {"rtm_hip_grid": _capi.Grid, "rtm_hip_params": _capi.Params, "rtm_hip_shot": _capi.Shot,
             "rtm_hip_times": _capi.Times, "rtm_hip_info": _capi.Info,
             "rtm_hip_kernel_stats": _capi.KernelStats, "rtm_hip_batch_config": _capi.BatchConfig,
             "rtm_hip_job": _capi.Job, "rtm_hip_batch_stats": _capi.BatchStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "rtm_hip.h"', 'int main(void) {']
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in cls._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {f[0]}));')
    lines += ["return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    import subprocess
    subprocess.run(["gcc", "-I", os.path.join(REPO, "include"), str(src), "-o", str(exe)], check=True)
    got = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    want = []
    for cname, cls in pairs.items():
        want.append(f"{cname} size {C.sizeof(cls)}")
        for f in cls._fields_:
            want.append(f"{cname} {f[0]} {getattr(cls, f[0]).offset}")
    assert [g for g in got if g] == want
