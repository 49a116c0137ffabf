"""SEG-Y shot input (rtm::io::read_shots_segy, csrc/io.cpp) against the
reference's reader (src/io/seismic_io.cpp:126-249) compiled from its own
sources (oracle/_ref/segy_dump), on SEG-Y files written here:

* every sample format (1 IBM float over random 32-bit words, including
  values that overflow or underflow float; 2 int32; 3 int16; 5 IEEE float),
  coordinate scalars > 0, < 0 and 0, shots grouped by field record number and,
  with record number 0, by changes of source x, uneven receiver counts, a
  trailing partial trace, extension dispatch;
* the errors, message for message;
* IBM words with known values (independent of the reference).
Both readers run through the same driver, tests/native/segy_dump.cpp, and
their decoded shots must be byte-identical.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from conftest import PKG, REPO

REF_DUMP = os.path.join(REPO, "oracle", "_ref", "segy_dump")


@pytest.fixture(scope="module")
def ours(tmp_path_factory):
    exe = tmp_path_factory.mktemp("segy") / "segy_dump"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(REPO, "include", "rtm"),
                    "-I", os.path.join(REPO, "include"), os.path.join(REPO, "tests", "native", "segy_dump.cpp"),
                    os.path.join(PKG, "csrc", "io.cpp"), "-o", str(exe)], check=True)
    return str(exe)


def write_segy(path, fmt, dt_us, ns, traces, tail=b""):
    """traces: list of (fldr, scalco, sx, gx, samples as raw big-endian bytes)."""
    with open(path, "wb") as f:
        f.write(b" " * 3200)
        bh = bytearray(400)
        struct.pack_into(">h", bh, 16, dt_us)
        struct.pack_into(">h", bh, 20, ns)
        struct.pack_into(">h", bh, 24, fmt)
        f.write(bh)
        for fldr, scalco, sx, gx, data in traces:
            th = bytearray(240)
            struct.pack_into(">i", th, 8, fldr)
            struct.pack_into(">h", th, 70, scalco)
            struct.pack_into(">i", th, 72, sx)
            struct.pack_into(">i", th, 80, gx)
            f.write(th)
            f.write(data)
        f.write(tail)


def samples(rng, fmt, ns):
    if fmt == 1:        # any 32-bit word is an IBM float
        w = rng.integers(0, 2**32, ns, dtype=np.uint64).astype(">u4")
        w[:4] = [0, 0x80000000, 0x7FFFFFFF, 0x00100000]     # 0, -0, overflow, underflow
        return w.tobytes()
    if fmt == 5:
        return rng.standard_normal(ns).astype(">f4").tobytes()
    if fmt == 2:
        return rng.integers(-2**31, 2**31, ns).astype(">i4").tobytes()
    return rng.integers(-2**15, 2**15, ns).astype(">i2").tobytes()


def run(exe, path, out, sz=12.5, rz=7.0):
    r = subprocess.run([exe, str(path), str(sz), str(rz), str(out)], capture_output=True, text=True)
    return r.returncode, r.stdout, (open(out, "rb").read() if r.returncode == 0 else b"")


def both(ours, path, tmp_path, **kw):
    a = run(ours, path, tmp_path / "ours.dump", **kw)
    if not os.path.exists(REF_DUMP):
        pytest.skip("oracle/_ref is not built (reference sources absent)")
    b = run(REF_DUMP, path, tmp_path / "ref.dump", **kw)
    assert a[0] == b[0] and a[1] == b[1], (a[:2], b[:2])
    assert a[2] == b[2]
    return a


@pytest.mark.parametrize("fmt", [1, 2, 3, 5])
@pytest.mark.parametrize("scalco", [0, 10, -100])
def test_formats_and_grouping(ours, tmp_path, fmt, scalco):
    rng = np.random.default_rng(fmt * 1000 + scalco + 7)
    ns, bps = 37, (2 if fmt == 3 else 4)
    trs = []
    for fldr, nrec in ((11, 5), (12, 3), (12, 2), (40, 7), (11, 1)):   # 4 shots: 11, 12 (5 traces), 40, 11
        for r in range(nrec):
            trs.append((fldr, scalco, 1000 * fldr + 3, 25 * r - 13, samples(rng, fmt, ns)))
    p = tmp_path / f"line_{fmt}.sgy"
    write_segy(p, fmt, 1500, ns, trs, tail=b"\x01" * (240 + ns * bps - 1))
    rc, _, dump = both(ours, p, tmp_path)
    assert rc == 0
    assert struct.unpack_from("<i", dump, 0)[0] == 4


def test_record_zero_groups_by_source_x(ours, tmp_path):
    rng = np.random.default_rng(3)
    trs = [(0, -10, sx, 10 * r, samples(rng, 5, 16)) for sx in (100, 100, 250, 100) for r in range(2)]
    p = tmp_path / "zero.segy"
    write_segy(p, 5, 2000, 16, trs)
    rc, _, dump = both(ours, p, tmp_path)
    assert rc == 0 and struct.unpack_from("<i", dump, 0)[0] == 3     # 100, 250, 100


def test_extension_dispatch(ours, tmp_path):
    rng = np.random.default_rng(5)
    p = tmp_path / "UPPER.SGY"
    write_segy(p, 3, 1000, 8, [(1, 1, 0, r, samples(rng, 3, 8)) for r in range(3)])
    assert both(ours, p, tmp_path)[0] == 0
    q = tmp_path / "not_rtms.bin"       # any other extension is RTMS
    q.write_bytes(p.read_bytes())
    rc, msg, _ = both(ours, q, tmp_path)
    assert rc == 3 and "RTMS" in msg


@pytest.mark.parametrize("case", ["short", "format4", "ns0", "ns_negative", "dt0", "no_traces", "missing"])
def test_errors(ours, tmp_path, case):
    p = tmp_path / "bad.sgy"
    fmt, dt, ns, trs = 5, 1000, 8, [(1, 0, 0, 0, b"\0" * 32)]
    if case == "format4":
        fmt = 4
    elif case == "ns0":
        ns = 0
    elif case == "ns_negative":
        ns = -5
    elif case == "dt0":
        dt = 0
    elif case == "no_traces":
        trs = []
    if case == "short":
        p.write_bytes(b"\0" * 3599)
    elif case != "missing":
        write_segy(p, fmt, dt, ns, trs)
    rc, msg, _ = both(ours, p, tmp_path)
    assert rc == 3 and msg.startswith("error: "), msg


def test_ibm_known_values(ours, tmp_path):
    """IBM hexadecimal floats: 0x42640000 = 0.390625*16^2 = 100, 0xC276A000 =
    -(0x76A000/2^24)*16^2 = -118.625, 0x41100000 = 1/16*16 = 1, 0x3F200000 =
    0.125/16 = 2^-7, 0x00000000 = 0; and 16^-65 (0x00100000) underflows float
    to 0, 0x7FFFFFFF overflows to inf."""
    words = [0x42640000, 0xC276A000, 0x41100000, 0x3F200000, 0, 0x00100000, 0x7FFFFFFF]
    want = np.array([100.0, -118.625, 1.0, 2.0**-7, 0.0, 0.0, np.inf], np.float32)
    p = tmp_path / "ibm.sgy"
    write_segy(p, 1, 4000, len(words), [(1, 0, 0, 0, np.array(words, ">u4").tobytes())])
    rc, _, dump = run(ours, p, tmp_path / "ibm.dump", sz=3.0, rz=4.0)
    assert rc == 0
    nshots, nt, dt, sx, sz, nrec = struct.unpack_from("<iifffi", dump, 0)
    assert (nshots, nt, nrec, sx, sz) == (1, len(words), 1, 0.0, 3.0)
    assert dt == np.float32(4000) * np.float32(1e-6)
    tr = np.frombuffer(dump, "<f4", len(words), 24 + 8)
    assert np.array_equal(tr, want)
