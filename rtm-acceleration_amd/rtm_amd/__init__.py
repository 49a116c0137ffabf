"""rtm_amd -- Python access to the MI355X RTM engine.

The engine is C++/HIP (rtm-acceleration_amd/csrc) behind the C ABI of
include/rtm_hip.h; this package only binds that ABI with ctypes, for tests,
benchmarks and scripting.  It mirrors the reference's engine interface
(include/rtm_engine.hpp:20-83): setup, forward_propagation,
backward_propagation, migrate, model_shot, times.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _capi
from ._capi import (SNAP_BF16, SNAP_F16, SNAP_F32, STACK_F32, STACK_F64, STEP_LDS_TILE, STEP_MARCH,
                    STEP_MARCH_DPP)
from .types import RTMParams, ShotRecord, TimeAxis, VelocityModel

__all__ = ["HipEngine", "RTMParams", "ShotRecord", "TimeAxis", "VelocityModel",
           "SNAP_F32", "SNAP_F16", "SNAP_BF16", "STACK_F32", "STACK_F64", "STEP_LDS_TILE", "STEP_MARCH", "STEP_MARCH_DPP", "device_count", "mute_direct_wave", "bandwidth_probe",
           "library_path"]


def library_path() -> str:
    return _capi.LIB_PATH


def device_count() -> int:
    return int(_capi.load().rtm_hip_device_count())


def bandwidth_probe(device: int = 0, nbytes: int = 1 << 30, repeats: int = 5) -> float:
    """Best device copy rate, GB/s (read + write bytes)."""
    lib = _capi.load()
    out = C.c_double()
    rc = lib.rtm_hip_bandwidth_probe(device, nbytes, repeats, C.byref(out))
    if rc != _capi.OK:
        raise RuntimeError(f"bandwidth probe failed ({rc})")
    return out.value


def mute_direct_wave(shot: ShotRecord, v_surface: float, t_pad: float, t_taper: float) -> None:
    """In-place direct-arrival mute of shot.traces (src/io/seismic_io.cpp:18-35)."""
    lib = _capi.load()
    keep: list = []
    s = _capi.shot_struct(shot, keep)
    tr = np.ascontiguousarray(shot.traces, dtype=np.float32)
    lib.rtm_hip_mute_direct_wave(C.byref(s), _capi.fptr(tr.reshape(-1)), v_surface, t_pad, t_taper)
    shot.traces = tr.reshape(shot.nrec, shot.nt)


class HipEngine:
    """One engine handle: one device, one stream, device-resident state."""

    def __init__(self, device: int = 0):
        self.lib = _capi.load()
        h = C.c_void_p()
        rc = self.lib.rtm_hip_create(device, C.byref(h))
        if rc != _capi.OK:
            raise RuntimeError(self.lib.rtm_hip_last_error(None).decode())
        self.h = h
        self.model = None
        self.params = None
        self.nt = 0
        self.dt = 0.0

    def close(self):
        if getattr(self, "h", None):
            self.lib.rtm_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != _capi.OK:
            raise RuntimeError(self.lib.rtm_hip_last_error(self.h).decode())

    # ---- RTMEngine::setup -------------------------------------------------
    def setup(self, model: VelocityModel, params: RTMParams, nt: int, dt: float,
              snapshot_format: int = SNAP_F32, image_stack: int = STACK_F32):
        g = _capi.Grid(model.nx, model.nz, model.dx, model.dz, model.ox, model.oz)
        p = _capi.Params(params.nb, params.order, params.sponge_alpha, params.store_interval,
                         params.f0, int(params.verbose), snapshot_format, image_stack)
        v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
        self._check(self.lib.rtm_hip_setup(self.h, C.byref(g), _capi.fptr(v), C.byref(p), nt, dt))
        self.model, self.params, self.nt, self.dt = model, params, nt, dt
        return self

    def info(self) -> dict:
        i = _capi.Info()
        self._check(self.lib.rtm_hip_get_info(self.h, C.byref(i)))
        out = {n: getattr(i, n) for n, _ in _capi.Info._fields_}
        for n in ("step_family", "tile_x", "autotuned"):
            out[n] = list(out[n])
        out["device_name"] = i.device_name.decode()
        return out

    @property
    def n_int(self) -> int:
        """Interior cells of the set-up grid (0 before setup)."""
        i = _capi.Info()
        self._check(self.lib.rtm_hip_get_info(self.h, C.byref(i)))
        return int(i.nx) * int(i.nz)

    # ---- propagation ------------------------------------------------------
    def forward(self, shot: ShotRecord, snapshots: bool = False, record: bool = False):
        """forward_propagation; returns (snapshots or None, traces or None)."""
        keep: list = []
        s = _capi.shot_struct(shot, keep)
        snap = np.empty(self.info()["num_snapshots"] * self.n_int, np.float32) if snapshots else None
        rec = np.zeros(shot.nrec * self.nt, np.float32) if (record and shot.nrec) else None
        self._check(self.lib.rtm_hip_forward(self.h, C.byref(s), _capi.fptr(snap), _capi.fptr(rec)))
        return snap, (None if rec is None else rec.reshape(shot.nrec, self.nt))

    def backward(self, shot: ShotRecord):
        keep: list = []
        s = _capi.shot_struct(shot, keep)
        self._check(self.lib.rtm_hip_backward(self.h, C.byref(s)))

    def image_zero(self):
        self._check(self.lib.rtm_hip_image_zero(self.h))

    def image(self):
        n = self.n_int
        if n == 0:    # not set up: let the engine report it
            self._check(self.lib.rtm_hip_image_get(self.h, None, None))
        img = np.empty(n, np.float32)
        ill = np.empty(n, np.float32)
        self._check(self.lib.rtm_hip_image_get(self.h, _capi.fptr(img), _capi.fptr(ill)))
        return img, ill

    def image_f64(self):
        """FP64 stacks (engines set up with image_stack=STACK_F64)."""
        n = self.n_int
        img = np.empty(max(n, 1), np.float64)
        ill = np.empty(max(n, 1), np.float64)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.rtm_hip_image_get_f64(self.h, img.ctypes.data_as(dp), ill.ctypes.data_as(dp)))
        return img[:n], ill[:n]

    def migrate(self, shots):
        keep: list = []
        arr = (_capi.Shot * max(len(shots), 1))()
        for i, sh in enumerate(shots):
            arr[i] = _capi.shot_struct(sh, keep)
        n = self.n_int
        img = np.empty(max(n, 1), np.float32)
        ill = np.empty(max(n, 1), np.float32)
        self._check(self.lib.rtm_hip_migrate(self.h, len(shots), arr, _capi.fptr(img), _capi.fptr(ill)))
        return img, ill

    def model_shot(self, shot: ShotRecord) -> np.ndarray:
        keep: list = []
        s = _capi.shot_struct(shot, keep)
        out = np.zeros(shot.nrec * self.nt, np.float32)
        self._check(self.lib.rtm_hip_model_shot(self.h, C.byref(s), _capi.fptr(out)))
        return out.reshape(shot.nrec, self.nt)

    # ---- accounting -------------------------------------------------------
    def times(self) -> dict:
        t = _capi.Times()
        self._check(self.lib.rtm_hip_get_times(self.h, C.byref(t)))
        return {n: getattr(t, n) for n, _ in _capi.Times._fields_}

    def reset_times(self):
        self._check(self.lib.rtm_hip_reset_times(self.h))

    def set_kernel_timing(self, every: int):
        self._check(self.lib.rtm_hip_set_kernel_timing(self.h, every))

    def kernel_stats(self) -> dict:
        k = _capi.KernelStats()
        self._check(self.lib.rtm_hip_get_kernel_stats(self.h, C.byref(k)))
        return {"launches": list(k.launches), "sampled": list(k.sampled),
                "sampled_ms": list(k.sampled_ms)}

    def set_step_kernel(self, family: int = -1, tile_x: int = 0):
        """Step-kernel mapping (STEP_MARCH / STEP_LDS_TILE, tile width 4/8/16);
        -1 / 0 select the built-in choice.  Results do not depend on it."""
        self._check(self.lib.rtm_hip_set_step_kernel(self.h, family, tile_x))
