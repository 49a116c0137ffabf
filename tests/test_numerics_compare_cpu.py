"""Host numerics of the product, checked without a GPU.

* The CFL bound and the points per wavelength that the engine library
  computes (rtm::max_stable_dt, rtm::points_per_wavelength, the reference's
  src/rtm/rtm.cpp:57-70), which the rtm CLI uses to refuse an unstable time
  step (src/main.cpp:188-190), equal the oracle's bit for bit for every order
  and the spacings and velocities of the five workloads and of random cases
  (the oracle's functions are pinned by closed forms in
  test_oracle_numerics.py).
* rtm_compare on images that differ: its metrics are field for field those of
  the reference's own rtm_compare, built from its sources (oracle/_ref), in
  the CSV row both append; a failing image exits 2 from both, a passing one 0.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle_lib as O
from conftest import PKG, REPO

LIB = os.path.join(PKG, "lib")
BIN = os.path.join(PKG, "bin")
GOLD = os.path.join(REPO, "tests", "golden", "layered3")

# (vmax, vmin, f0, dx, dz) of the 201x151, 2301x751, 4001x1201, Marmousi and
# 1201x401 workloads
WORKLOADS = [(3000.0, 1500.0, 12.0, 10.0, 10.0), (5775.0, 1500.0, 15.0, 4.0, 4.0), (4500.0, 1500.0, 12.0, 5.0, 5.0),
             (4700.0, 1500.0, 10.0, 12.5, 12.5), (4800.0, 1500.0, 12.0, 8.0, 8.0)]


@pytest.fixture(scope="module")
def numerics_exe(tmp_path_factory):
    exe = tmp_path_factory.mktemp("num") / "numerics_check"
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", os.path.join(REPO, "include"),
                    os.path.join(REPO, "tests", "native", "numerics_check.cpp"), "-o", str(exe),
                    "-L", LIB, "-lrtm_hip", f"-Wl,-rpath,{LIB}"], check=True)
    return str(exe)


def test_cfl_and_dispersion_match_oracle(numerics_exe):
    rng = np.random.default_rng(7)
    cases = [w + (order,) for w in WORKLOADS for order in (2, 4, 6, 8, 10, 12)]
    for _ in range(60):
        vmin = float(np.float32(rng.uniform(1000, 2500)))
        cases.append((float(np.float32(vmin * rng.uniform(1.0, 4.0))), vmin, float(np.float32(rng.uniform(5, 40))),
                      float(np.float32(rng.uniform(1, 25))), float(np.float32(rng.uniform(1, 25))),
                      int(rng.choice([2, 4, 6, 8, 10, 12]))))
    args = [f"{vmax!r}:{vmin!r}:{f0!r}:{dx!r}:{dz!r}:{order}" for vmax, vmin, f0, dx, dz, order in cases]
    out = subprocess.run([numerics_exe] + args, capture_output=True, text=True, check=True).stdout.split("\n")
    L = O.lib()
    for (vmax, vmin, f0, dx, dz, order), line in zip(cases, out):
        _, dt_hex, ppw_hex = line.split()
        dt_ref = np.float32(L.orc_max_stable_dt(vmax, dx, dz, order))
        ppw_ref = np.float32(L.orc_points_per_wavelength(vmin, f0, dx, dz))
        assert int(dt_hex, 16) == int(dt_ref.view(np.uint32)), (vmax, dx, dz, order)
        assert int(ppw_hex, 16) == int(ppw_ref.view(np.uint32)), (vmin, f0, dx, dz)


@pytest.mark.skipif(not O.ref_available(), reason="the reference build (oracle/_ref) needs the reference sources")
@pytest.mark.parametrize("noise,code", [(1e-7, 0), (3e-3, 2), (0.5, 2)])
def test_compare_metrics_match_reference(tmp_path, noise, code):
    ref_img = os.path.join(GOLD, "image.f32")
    img = np.fromfile(ref_img, np.float32)
    rng = np.random.default_rng(3)
    test = tmp_path / "test.f32"
    (img + (noise * np.abs(img).max() * rng.standard_normal(img.size)).astype(np.float32)).astype(
        np.float32).tofile(test)
    rows = []
    for exe, name in ((os.path.join(BIN, "rtm_compare"), "ours.csv"),
                      (os.path.join(REPO, "oracle", "_ref", "rtm_compare"), "ref.csv")):
        csv = tmp_path / name
        r = subprocess.run([exe, ref_img, str(test), "--nx", "201", "--nz", "151", "--csv", str(csv),
                            "--engine", "e", "--dataset", "d"], capture_output=True, text=True)
        assert r.returncode == code, (exe, r.stdout[-500:], r.stderr[-500:])
        header, row = csv.read_text().splitlines()
        rows.append((header, row.split(",")))
    (h0, r0), (h1, r1) = rows
    assert h0 == h1
    # date and host may differ by a second; every metric and the verdict match
    assert r0[2:] == r1[2:]
