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
from ._capi import (JOB_DONE, JOB_FAILED, JOB_QUEUED, JOB_RUNNING, SNAP_BF16, SNAP_F16, SNAP_F32,
                    STACK_F32, STACK_F64, STEP_LDS_TILE, STEP_MARCH, STEP_MARCH_DPP, STEP_STREAM)
from .types import RTMParams, ShotRecord, TimeAxis, VelocityModel

__all__ = ["HipEngine", "BatchEngine", "JOB_QUEUED", "JOB_RUNNING", "JOB_DONE", "JOB_FAILED", "RTMParams", "ShotRecord", "TimeAxis", "VelocityModel",
           "SNAP_F32", "SNAP_F16", "SNAP_BF16", "STACK_F32", "STACK_F64", "STEP_LDS_TILE", "STEP_MARCH", "STEP_MARCH_DPP", "STEP_STREAM", "device_count", "mute_direct_wave", "bandwidth_probe",
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
        for n in ("step_family", "tile_x", "autotuned", "waves_per_cu"):
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

    def postprocess(self, illum: bool = True, laplacian: bool = False, eps_relative: float = 1e-4):
        """Illumination compensation and/or the Laplacian filter of the
        resident image, on the device (src/rtm/imaging.cpp:48-75)."""
        flags = (_capi.POST_ILLUM if illum else 0) | (_capi.POST_LAPLACIAN if laplacian else 0)
        out = np.empty(max(self.n_int, 1), np.float32)
        self._check(self.lib.rtm_hip_image_postprocess(self.h, flags, eps_relative, _capi.fptr(out)))
        return out[:self.n_int]

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


class BatchEngine:
    """Batched / served migration (rtm_hip_batch_*): independent shot jobs,
    each on its own grid, advanced together one time step per tick."""

    def __init__(self, device: int = 0, memory_budget: int = 0, max_active: int = 0,
                 use_graph: bool = True, kernel_family: int = 0, tile_x: int = 0, merge_groups: bool = False):
        self.lib = _capi.load()
        cfg = _capi.BatchConfig(int(memory_budget), int(max_active), int(bool(use_graph)), int(kernel_family),
                                int(tile_x), int(bool(merge_groups)))
        h = C.c_void_p()
        rc = self.lib.rtm_hip_batch_create(device, C.byref(cfg), C.byref(h))
        if rc != _capi.OK:
            raise RuntimeError(self.lib.rtm_hip_batch_last_error(None).decode())
        self.h = h
        self.sizes: dict[int, int] = {}
        self.model_sizes: dict[int, int] = {}

    def close(self):
        if getattr(self, "h", None):
            self.lib.rtm_hip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _check(self, rc: int):
        if rc != _capi.OK:
            raise RuntimeError(self.lib.rtm_hip_batch_last_error(self.h).decode())

    @staticmethod
    def _job(model: VelocityModel, params: RTMParams, nt: int, dt: float, shot: ShotRecord,
             snapshot_format: int, stack_at, keep: list) -> _capi.Job:
        j = _capi.Job()
        j.grid = _capi.Grid(model.nx, model.nz, model.dx, model.dz, model.ox, model.oz)
        v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
        keep.append(v)
        j.v = _capi.fptr(v)
        j.params = _capi.Params(params.nb, params.order, params.sponge_alpha, params.store_interval,
                                params.f0, int(params.verbose), snapshot_format, STACK_F32)
        j.nt, j.dt = int(nt), float(dt)
        j.shot = _capi.shot_struct(shot, keep)
        j.stack_ix0, j.stack_iz0 = (-1, -1) if stack_at is None else (int(stack_at[0]), int(stack_at[1]))
        return j

    def reserve(self, nbytes: int = 0):
        """Reserve device memory for all later jobs (0: the whole budget)."""
        self._check(self.lib.rtm_hip_batch_reserve(self.h, int(nbytes)))

    def add_model(self, model: VelocityModel, params: RTMParams, nt: int, dt: float,
                  snapshot_format: int = SNAP_F32) -> int:
        """Resident model (the arguments of HipEngine.setup); returns its id."""
        g = _capi.Grid(model.nx, model.nz, model.dx, model.dz, model.ox, model.oz)
        p = _capi.Params(params.nb, params.order, params.sponge_alpha, params.store_interval,
                         params.f0, int(params.verbose), snapshot_format, STACK_F32)
        v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
        mid = C.c_int()
        self._check(self.lib.rtm_hip_batch_add_model(self.h, C.byref(g), _capi.fptr(v), C.byref(p), int(nt),
                                                     float(dt), C.byref(mid)))
        self.model_sizes[mid.value] = model.nx * model.nz
        return int(mid.value)

    def remove_model(self, model_id: int):
        self._check(self.lib.rtm_hip_batch_remove_model(self.h, int(model_id)))
        self.model_sizes.pop(model_id, None)

    def submit_shot(self, model_id: int, shot: ShotRecord, stack_at=None) -> int:
        """Queue one shot on a resident model."""
        keep: list = []
        s = _capi.shot_struct(shot, keep)
        ix0, iz0 = (-1, -1) if stack_at is None else (int(stack_at[0]), int(stack_at[1]))
        jid = C.c_longlong()
        self._check(self.lib.rtm_hip_batch_submit_shot(self.h, int(model_id), C.byref(s), ix0, iz0, C.byref(jid)))
        self.sizes[jid.value] = self.model_sizes[model_id]
        return int(jid.value)

    def job_bytes(self, model, params, nt, dt, shot, snapshot_format: int = SNAP_F32) -> int:
        keep: list = []
        j = self._job(model, params, nt, dt, shot, snapshot_format, None, keep)
        out = C.c_size_t()
        self._check(self.lib.rtm_hip_batch_job_bytes(self.h, C.byref(j), C.byref(out)))
        return int(out.value)

    def submit(self, model: VelocityModel, params: RTMParams, nt: int, dt: float, shot: ShotRecord,
               snapshot_format: int = SNAP_F32, stack_at=None) -> int:
        """Queue one shot migration; stack_at = (ix0, iz0) adds its image into
        the FP64 stack at that offset when it finishes."""
        keep: list = []
        j = self._job(model, params, nt, dt, shot, snapshot_format, stack_at, keep)
        jid = C.c_longlong()
        self._check(self.lib.rtm_hip_batch_submit(self.h, C.byref(j), C.byref(jid)))
        self.sizes[jid.value] = model.nx * model.nz
        return int(jid.value)

    def run(self, max_ticks: int = 0, until_done: bool = True):
        self._check(self.lib.rtm_hip_batch_run(self.h, int(max_ticks), int(bool(until_done))))

    def state(self, jid: int):
        st, lat = C.c_int(), C.c_double()
        self._check(self.lib.rtm_hip_batch_job_state(self.h, jid, C.byref(st), C.byref(lat)))
        return int(st.value), float(lat.value)

    def result(self, jid: int):
        n = self.sizes[jid]
        img = np.empty(n, np.float32)
        ill = np.empty(n, np.float32)
        self._check(self.lib.rtm_hip_batch_result(self.h, jid, _capi.fptr(img), _capi.fptr(ill)))
        del self.sizes[jid]
        return img, ill

    def set_stack(self, nx: int, nz: int):
        self._check(self.lib.rtm_hip_batch_set_stack(self.h, nx, nz))
        self.stack_shape = (nx, nz)

    def stack(self):
        n = self.stack_shape[0] * self.stack_shape[1]
        img = np.empty(n, np.float64)
        ill = np.empty(n, np.float64)
        dp = C.POINTER(C.c_double)
        self._check(self.lib.rtm_hip_batch_get_stack(self.h, img.ctypes.data_as(dp), ill.ctypes.data_as(dp)))
        return img, ill

    def stats(self) -> dict:
        s = _capi.BatchStats()
        self._check(self.lib.rtm_hip_batch_get_stats(self.h, C.byref(s)))
        out = {n: getattr(s, n) for n, _ in _capi.BatchStats._fields_}
        k = s.ngroups
        names = ["lds-tile", "march", "march-dpp", "stream"]
        out["groups"] = [{"half": s.group_half[i], "snap16": bool(s.group_snap16[i]), "tile_x": s.group_tile_x[i],
                          "family": names[s.group_family[i]], "waves_per_cu": s.group_waves_per_cu[i]}
                         for i in range(k)]
        for n in ("group_half", "group_tile_x", "group_family", "group_waves_per_cu", "group_snap16"):
            del out[n]
        return out
