"""ctypes binding of the engine's C ABI (include/rtm_hip.h).

The shared library is rtm-acceleration_amd/lib/librtm_hip.so, built in-tree by
csrc/Makefile.  Loading fails loudly when it is missing: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_ROOT, "lib", "librtm_hip.so")

OK, ERR_ARG, ERR_DEVICE, ERR_MEMORY = 0, -1, -2, -3
STEP_LDS_TILE, STEP_MARCH, STEP_MARCH_DPP, STEP_STREAM = 0, 1, 2, 3
SNAP_F32, SNAP_F16, SNAP_BF16 = 0, 1, 2
POST_ILLUM, POST_LAPLACIAN = 1, 2
STACK_F32, STACK_F64 = 0, 1
JOB_QUEUED, JOB_RUNNING, JOB_DONE, JOB_FAILED = 0, 1, 2, 3


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oz", C.c_float)]


class Params(C.Structure):
    _fields_ = [("nb", C.c_int), ("order", C.c_int), ("sponge_alpha", C.c_float),
                ("store_interval", C.c_int), ("f0", C.c_float), ("verbose", C.c_int),
                ("snapshot_format", C.c_int), ("image_stack", C.c_int)]


class Shot(C.Structure):
    _fields_ = [("sx", C.c_float), ("sz", C.c_float), ("nrec", C.c_int),
                ("rx", C.POINTER(C.c_float)), ("rz", C.POINTER(C.c_float)),
                ("nt", C.c_int), ("dt", C.c_float), ("traces", C.POINTER(C.c_float))]


class Times(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("io", "forward", "backward", "imaging", "total", "h2d", "kernel", "d2h")]


class Info(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("nb", C.c_int), ("half", C.c_int),
                ("nxe", C.c_int), ("nze", C.c_int), ("ldz", C.c_int), ("nt", C.c_int),
                ("store_interval", C.c_int), ("num_snapshots", C.c_int),
                ("snapshot_bytes", C.c_size_t), ("device_bytes", C.c_size_t),
                ("snapshot_format", C.c_int), ("image_stack", C.c_int), ("step_family", C.c_int * 2), ("tile_x", C.c_int * 2),
                ("autotuned", C.c_int * 2), ("waves_per_cu", C.c_int * 2),
                ("device_name", C.c_char * 128)]


class KernelStats(C.Structure):
    _fields_ = [("launches", C.c_longlong * 2), ("sampled", C.c_longlong * 2),
                ("sampled_ms", C.c_double * 2)]


class BatchConfig(C.Structure):
    _fields_ = [("memory_budget", C.c_size_t), ("max_active", C.c_int), ("use_graph", C.c_int),
                ("kernel_family", C.c_int), ("tile_x", C.c_int), ("merge_groups", C.c_int)]


class Job(C.Structure):
    _fields_ = [("grid", Grid), ("v", C.POINTER(C.c_float)), ("params", Params), ("nt", C.c_int),
                ("dt", C.c_float), ("shot", Shot), ("stack_ix0", C.c_int), ("stack_iz0", C.c_int)]


class BatchStats(C.Structure):
    _fields_ = [("jobs_submitted", C.c_longlong), ("jobs_done", C.c_longlong),
                ("jobs_active", C.c_int), ("jobs_queued", C.c_int), ("ticks", C.c_longlong),
                ("rebuilds", C.c_longlong), ("cell_steps", C.c_double), ("run_seconds", C.c_double),
                ("admit_seconds", C.c_double), ("tune_seconds", C.c_double), ("retire_seconds", C.c_double),
                ("busy_latency_s", C.c_double), ("models_resident", C.c_int), ("bytes_in_use", C.c_size_t),
                ("bytes_budget", C.c_size_t), ("ngroups", C.c_int), ("group_half", C.c_int * 12),
                ("group_tile_x", C.c_int * 12), ("group_family", C.c_int * 12),
                ("group_waves_per_cu", C.c_int * 12), ("group_snap16", C.c_int * 12), ("tunes", C.c_int),
                ("rebuild_seconds", C.c_double), ("capture_seconds", C.c_double),
                ("graphs_captured", C.c_longlong), ("step_launches", C.c_longlong),
                ("lean_launches", C.c_longlong), ("merged_launch", C.c_int),
                ("concurrent_ticks", C.c_longlong), ("staged_admissions", C.c_longlong)]


# Every entry point declared in include/rtm_hip.h, with its ctypes signature.
_f32p = C.POINTER(C.c_float)
_H = C.c_void_p
SIGNATURES = {
    "rtm_hip_device_count": (C.c_int, []),
    "rtm_hip_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "rtm_hip_destroy": (None, [_H]),
    "rtm_hip_last_error": (C.c_char_p, [_H]),
    "rtm_hip_setup": (C.c_int, [_H, C.POINTER(Grid), _f32p, C.POINTER(Params), C.c_int, C.c_float]),
    "rtm_hip_get_info": (C.c_int, [_H, C.POINTER(Info)]),
    "rtm_hip_forward": (C.c_int, [_H, C.POINTER(Shot), _f32p, _f32p]),
    "rtm_hip_backward": (C.c_int, [_H, C.POINTER(Shot)]),
    "rtm_hip_image_zero": (C.c_int, [_H]),
    "rtm_hip_image_get": (C.c_int, [_H, _f32p, _f32p]),
    "rtm_hip_image_get_f64": (C.c_int, [_H, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "rtm_hip_image_postprocess": (C.c_int, [_H, C.c_int, C.c_float, _f32p]),
    "rtm_hip_image_postprocess_noalloc": (C.c_int, [_H, C.c_int, C.c_float, C.c_void_p]),
    "rtm_hip_migrate": (C.c_int, [_H, C.c_int, C.POINTER(Shot), _f32p, _f32p]),
    "rtm_hip_model_shot": (C.c_int, [_H, C.POINTER(Shot), _f32p]),
    "rtm_hip_get_times": (C.c_int, [_H, C.POINTER(Times)]),
    "rtm_hip_reset_times": (C.c_int, [_H]),
    "rtm_hip_set_kernel_timing": (C.c_int, [_H, C.c_int]),
    "rtm_hip_get_kernel_stats": (C.c_int, [_H, C.POINTER(KernelStats)]),
    "rtm_hip_set_step_kernel": (C.c_int, [_H, C.c_int, C.c_int]),
    "rtm_hip_mute_direct_wave": (None, [C.POINTER(Shot), _f32p, C.c_float, C.c_float, C.c_float]),
    "rtm_hip_bandwidth_probe": (C.c_int, [C.c_int, C.c_size_t, C.c_int, C.POINTER(C.c_double)]),
    "rtm_hip_batch_create": (C.c_int, [C.c_int, C.POINTER(BatchConfig), C.POINTER(C.c_void_p)]),
    "rtm_hip_batch_destroy": (None, [_H]),
    "rtm_hip_batch_last_error": (C.c_char_p, [_H]),
    "rtm_hip_batch_add_model": (C.c_int, [_H, C.POINTER(Grid), _f32p, C.POINTER(Params), C.c_int, C.c_float,
                                          C.POINTER(C.c_int)]),
    "rtm_hip_batch_remove_model": (C.c_int, [_H, C.c_int]),
    "rtm_hip_batch_submit_shot": (C.c_int, [_H, C.c_int, C.POINTER(Shot), C.c_int, C.c_int, C.POINTER(C.c_longlong)]),
    "rtm_hip_batch_reserve": (C.c_int, [_H, C.c_size_t]),
    "rtm_hip_batch_job_bytes": (C.c_int, [_H, C.POINTER(Job), C.POINTER(C.c_size_t)]),
    "rtm_hip_batch_submit": (C.c_int, [_H, C.POINTER(Job), C.POINTER(C.c_longlong)]),
    "rtm_hip_batch_run": (C.c_int, [_H, C.c_int, C.c_int]),
    "rtm_hip_batch_job_state": (C.c_int, [_H, C.c_longlong, C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "rtm_hip_batch_result": (C.c_int, [_H, C.c_longlong, _f32p, _f32p]),
    "rtm_hip_batch_set_stack": (C.c_int, [_H, C.c_int, C.c_int]),
    "rtm_hip_batch_get_stack": (C.c_int, [_H, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "rtm_hip_batch_get_stats": (C.c_int, [_H, C.POINTER(BatchStats)]),
}

_lib = None


def load():
    """Load librtm_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP engine library not built: {LIB_PATH} "
                               "(run __graft_entry__.build() or make -C rtm-acceleration_amd/csrc)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_f32p)


def shot_struct(shot, keep: list) -> Shot:
    """C view of a shot-like object (sx, sz, rx, rz, nt, dt, traces)."""
    rx = np.ascontiguousarray(shot.rx, dtype=np.float32).reshape(-1)
    rz = np.ascontiguousarray(shot.rz, dtype=np.float32).reshape(-1)
    s = Shot()
    s.sx, s.sz, s.nrec, s.nt, s.dt = float(shot.sx), float(shot.sz), int(rx.size), int(shot.nt), float(shot.dt)
    s.rx, s.rz = fptr(rx), fptr(rz)
    keep += [rx, rz]
    if shot.traces is not None:
        tr = np.ascontiguousarray(shot.traces, dtype=np.float32).reshape(-1)
        if tr.size != rx.size * int(shot.nt):
            raise ValueError("traces must hold nrec*nt samples")
        s.traces = fptr(tr)
        keep.append(tr)
    else:
        s.traces = None
    return s
