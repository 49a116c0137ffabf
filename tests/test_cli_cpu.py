"""The command-line drop-ins, checked without a GPU: flags, exit codes and
the parity gate of rtm_compare (src/tools/compare_main.cpp:52-77: exit 0
pass, 2 fail, 1 error)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import REPO

BIN = os.path.join(REPO, "rtm-acceleration_amd", "bin")
GOLD = os.path.join(REPO, "tests", "golden", "layered3")


def run(*args, **kw):
    return subprocess.run([os.path.join(BIN, args[0]), *args[1:]], capture_output=True, text=True, **kw)


def test_rtm_lists_engines_and_requires_inputs():
    r = run("rtm", "--list-engines")
    assert r.returncode == 0 and r.stdout.split() == ["hip", "hip-fp16", "hip-bf16"]
    r = run("rtm")
    assert r.returncode == 1 and "--velocity, --shots and --output are required" in r.stderr
    r = run("rtm", "--bogus")
    assert r.returncode == 1 and "unknown option: --bogus" in r.stderr


@pytest.mark.skipif(not os.path.exists(os.path.join(REPO, "oracle", "_ref", "rtm_hip")),
                    reason="oracle/_ref/rtm_hip is built only where the reference sources are present")
def test_reference_driver_lists_hip():
    """The reference's own rtm (src/main.cpp), built with the hip engines
    registered in its factory (oracle/Makefile), lists them after its own."""
    r = subprocess.run([os.path.join(REPO, "oracle", "_ref", "rtm_hip"), "--list-engines"], capture_output=True,
                       text=True)
    assert r.returncode == 0 and r.stdout.split() == ["cpu", "cpu-opt", "hip", "hip-fp16", "hip-bf16"]
    r = subprocess.run([os.path.join(REPO, "oracle", "_ref", "rtm_hip"), "--engine", "hip2"], capture_output=True,
                       text=True)
    assert r.returncode == 1 and "--velocity, --shots and --output are required" in r.stderr


def test_rtm_compare_gate(tmp_path):
    ref = os.path.join(GOLD, "image.f32")
    assert run("rtm_compare", ref, ref, "--nx", "201", "--nz", "151").returncode == 0
    img = np.fromfile(ref, np.float32)
    small = tmp_path / "small.f32"
    (img * np.float32(1 + 1e-7)).astype(np.float32).tofile(small)        # within the gate
    assert run("rtm_compare", ref, str(small)).returncode == 0
    big = tmp_path / "big.f32"
    (img + np.float32(1e-2) * np.abs(img).max()).astype(np.float32).tofile(big)
    csv = tmp_path / "cmp.csv"
    r = run("rtm_compare", ref, str(big), "--csv", str(csv), "--engine", "x", "--dataset", "y")
    assert r.returncode == 2 and "FAIL" in r.stdout
    lines = csv.read_text().splitlines()
    assert lines[0].startswith("date,host,engine,dataset,max_abs") and lines[1].endswith(",0")
    assert run("rtm_compare", ref, str(tmp_path / "missing.f32")).returncode == 1


def test_rtm_without_device_fails_cleanly(tmp_path):
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is present")
    r = run("rtm", "--velocity", os.path.join(GOLD, "velocity.bin"), "--shots", "x.bin", "--output",
            str(tmp_path / "o.bin"))
    assert r.returncode == 1 and "no HIP device" in r.stderr


def test_bandwidth_probe_csv(tmp_path):
    """rtm_bandwidth_probe's CSV is the reference's (src/tools/bandwidth_probe.cpp:
    date,host,device_kind,device_name,threads,gbps, header once, one row per
    device); the CPU triad row comes first; a bad option exits 1."""
    csv = tmp_path / "bw.csv"
    for _ in range(2):
        r = run("rtm_bandwidth_probe", "--mb", "16", "--repeats", "1", "--csv", str(csv), timeout=120)
        assert r.returncode == 0, r.stderr
        assert "triad" in r.stdout
    lines = csv.read_text().splitlines()
    assert lines[0] == "date,host,device_kind,device_name,threads,gbps"
    assert lines.count(lines[0]) == 1
    rows = [ln.split(",") for ln in lines[1:]]
    assert rows[0][2] == "cpu" and int(rows[0][4]) >= 1 and float(rows[0][5]) > 0.1
    assert all(len(x) == 6 and x[2] in ("cpu", "gpu") for x in rows)
    r = run("rtm_bandwidth_probe", "--bogus")
    assert r.returncode == 1 and "unknown option: --bogus" in r.stderr
