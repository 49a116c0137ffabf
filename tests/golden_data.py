"""Loaders for the committed golden fixtures (tests/golden/, written by
tests/golden/make_golden.py with the reference compiled from source)."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def layered3():
    d = os.path.join(HERE, "layered3")
    with open(os.path.join(d, "manifest.json")) as f:
        man = json.load(f)
    v = np.fromfile(os.path.join(d, "velocity.bin"), np.float32)
    image = np.fromfile(os.path.join(d, "image.f32"), np.float32)
    illum = np.fromfile(os.path.join(d, "illummap.f32"), np.float32)
    return man, v, image, illum


def tiny_cases():
    with open(os.path.join(HERE, "tiny_cases.json")) as f:
        meta = json.load(f)["cases"]
    z = np.load(os.path.join(HERE, "tiny_cases.npz"))
    out = []
    for m in meta:
        n = m["name"]
        out.append(dict(m, v=z[f"{n}/v"], rx=z[f"{n}/rx"], rz=z[f"{n}/rz"], sx=z[f"{n}/sx"],
                        traces=z[f"{n}/traces"], image=z[f"{n}/image"], illum=z[f"{n}/illum"]))
    return out


def marmousi():
    d = os.path.join(HERE, "marmousi")
    with open(os.path.join(d, "manifest.json")) as f:
        man = json.load(f)
    v = np.fromfile(os.path.join(d, "marmousi_vp_12.5m.bin"), np.float32)
    return man, v


def sha256(b: bytes) -> str:
    return hashlib.sha256(b).hexdigest()


def rtms_bytes(shots, traces) -> bytes:
    """RTMS v1 byte stream (src/io/seismic_io.cpp:39-88) of shots + traces."""
    nrec, nt = shots[0].nrec, shots[0].nt
    head = b"RTMS" + np.array([1, len(shots), nrec, nt], np.int32).tobytes() + \
        np.array([shots[0].dt], np.float32).tobytes() + np.array([0, 0], np.int32).tobytes()
    body = b"".join(np.array([s.sx, s.sz], np.float32).tobytes() + np.asarray(s.rx, np.float32).tobytes() +
                    np.asarray(s.rz, np.float32).tobytes() + np.asarray(t, np.float32).tobytes()
                    for s, t in zip(shots, traces))
    return head + body
