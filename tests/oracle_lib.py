"""ctypes access to the CPU checker (test infrastructure only).

Two libraries, both built by ``oracle/Makefile``:

* ``oracle/liboracle.so`` -- the in-repo restatement of the reference's CPU
  engines (``cpu`` = src/rtm/*.cpp, ``cpu-opt`` = src/cpu_opt/cpu_opt.cpp) in
  FP32 and FP64;
* ``oracle/_ref/librtmref.so`` -- the reference itself compiled from its own
  sources plus a thin C shim; present only where the reference sources were
  available at build time.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module; the product never does.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(REPO, "oracle", "liboracle.so")
REF_SO = os.path.join(REPO, "oracle", "_ref", "librtmref.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrcGrid(C.Structure):
    _fields_ = [("nx", C.c_int), ("nz", C.c_int), ("dx", C.c_float), ("dz", C.c_float),
                ("ox", C.c_float), ("oz", C.c_float)]


class OrcParams(C.Structure):
    _fields_ = [("nb", C.c_int), ("order", C.c_int), ("sponge_alpha", C.c_float),
                ("store_interval", C.c_int), ("f0", C.c_float), ("verbose", C.c_int)]


class OrcShot(C.Structure):
    _fields_ = [("sx", C.c_float), ("sz", C.c_float), ("nrec", C.c_int),
                ("rx", C.POINTER(C.c_float)), ("rz", C.POINTER(C.c_float)),
                ("nt", C.c_int), ("dt", C.c_float), ("traces", C.POINTER(C.c_float))]


@dataclass
class Model:
    """Velocity model, interior, depth-fastest v[ix*nz+iz] (include/rtm_types.hpp:42-47)."""
    nx: int
    nz: int
    dx: float
    dz: float
    v: np.ndarray
    ox: float = 0.0
    oz: float = 0.0


@dataclass
class Params:
    """rtm::RTMParams defaults (include/rtm_types.hpp:33-40)."""
    nb: int = 60
    order: int = 8
    sponge_alpha: float = 0.0053
    store_interval: int = 1
    f0: float = 12.0
    verbose: int = 0


@dataclass
class Shot:
    """rtm::ShotRecord (include/rtm_types.hpp:49-57); traces trace-major (nrec, nt)."""
    sx: float
    sz: float
    rx: np.ndarray
    rz: np.ndarray
    nt: int
    dt: float
    traces: np.ndarray | None = None
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nrec(self) -> int:
        return int(len(self.rx))


def shot_struct(shot, keep: list) -> OrcShot:
    """C view of any shot-like object (sx, sz, rx, rz, nt, dt, traces)."""
    rx = np.ascontiguousarray(shot.rx, dtype=np.float32).reshape(-1)
    rz = np.ascontiguousarray(shot.rz, dtype=np.float32).reshape(-1)
    s = OrcShot()
    s.sx, s.sz, s.nrec, s.nt, s.dt = float(shot.sx), float(shot.sz), len(rx), int(shot.nt), float(shot.dt)
    s.rx = rx.ctypes.data_as(C.POINTER(C.c_float))
    s.rz = rz.ctypes.data_as(C.POINTER(C.c_float))
    keep += [rx, rz]
    if shot.traces is not None:
        tr = np.ascontiguousarray(shot.traces, dtype=np.float32).reshape(-1)
        assert tr.size == len(rx) * int(shot.nt)
        s.traces = tr.ctypes.data_as(C.POINTER(C.c_float))
        keep.append(tr)
    else:
        s.traces = None
    return s


def _grid(m: Model) -> OrcGrid:
    return OrcGrid(m.nx, m.nz, m.dx, m.dz, m.ox, m.oz)


def _params(p: Params) -> OrcParams:
    return OrcParams(p.nb, p.order, p.sponge_alpha, p.store_interval, p.f0, p.verbose)


def _shots_array(shots, keep: list):
    arr = (OrcShot * max(len(shots), 1))()
    for i, s in enumerate(shots):
        arr[i] = shot_struct(s, keep)
    return arr


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle library not built: {ORACLE_SO} (run oracle/Makefile)")
        L = C.CDLL(ORACLE_SO)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_int, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_last_error.restype = C.c_char_p
        L.orc_setup.argtypes = [C.c_void_p, C.POINTER(OrcGrid), _f32p, C.POINTER(OrcParams),
                                C.c_int, C.c_float]
        L.orc_num_snapshots.argtypes = [C.c_void_p]
        L.orc_num_threads.argtypes = [C.c_void_p]
        L.orc_get_table_f32.argtypes = [C.c_void_p, C.c_int, _f32p]
        L.orc_get_table_f64.argtypes = [C.c_void_p, C.c_int, _f64p]
        L.orc_map_geometry.argtypes = [C.c_void_p, C.POINTER(OrcShot),
                                       C.POINTER(C.c_longlong), np.ctypeslib.ndpointer(np.int64)]
        for suf, ptr in (("f32", _f32p), ("f64", _f64p)):
            getattr(L, f"orc_forward_{suf}").argtypes = [C.c_void_p, C.POINTER(OrcShot), C.c_void_p, C.c_void_p]
            getattr(L, f"orc_backward_{suf}").argtypes = [C.c_void_p, C.POINTER(OrcShot), ptr, ptr, ptr]
            getattr(L, f"orc_migrate_{suf}").argtypes = [C.c_void_p, C.c_int, C.POINTER(OrcShot), ptr, ptr]
        L.orc_model_shot_f32.argtypes = [C.c_void_p, C.POINTER(OrcShot), _f32p]
        L.orc_times.argtypes = [C.c_void_p] + [C.POINTER(C.c_double)] * 3
        L.orc_stencil_half.argtypes = [C.c_int]
        L.orc_stencil_coefficients_f32.argtypes = [C.c_int, _f32p]
        L.orc_stencil_coefficients_f64.argtypes = [C.c_int, _f64p]
        L.orc_stencil_abs_sum.restype = C.c_float
        L.orc_stencil_abs_sum.argtypes = [C.c_int]
        L.orc_ricker.restype = C.c_float
        L.orc_ricker.argtypes = [C.c_float, C.c_float]
        L.orc_max_stable_dt.restype = C.c_float
        L.orc_max_stable_dt.argtypes = [C.c_float, C.c_float, C.c_float, C.c_int]
        L.orc_points_per_wavelength.restype = C.c_float
        L.orc_points_per_wavelength.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_mute_direct_wave.argtypes = [C.POINTER(OrcShot), _f32p, C.c_float, C.c_float, C.c_float]
        L.orc_illumination_compensation.argtypes = [_f32p, _f32p, C.c_size_t, C.c_float]
        L.orc_laplacian_filter.argtypes = [_f32p, C.c_int, C.c_int, C.c_float, C.c_float]
        _lib = L
    return _lib


class Oracle:
    """One oracle engine: precision 32/64, variant 'cpu' or 'cpu-opt'."""

    def __init__(self, model: Model, params: Params, nt: int, dt: float,
                 precision: int = 32, variant: str = "cpu"):
        L = lib()
        self.L = L
        self.prec = precision
        self.dtype = np.float32 if precision == 32 else np.float64
        self.h = L.orc_create(precision, {"cpu": 0, "cpu-opt": 1}[variant])
        if not self.h:
            raise RuntimeError(L.orc_last_error().decode())
        self.model, self.params, self.nt, self.dt = model, params, nt, dt
        v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
        g, p = _grid(model), _params(params)
        if L.orc_setup(self.h, C.byref(g), v, C.byref(p), nt, dt) != 0:
            raise RuntimeError(L.orc_last_error().decode())
        self.nsnap = L.orc_num_snapshots(self.h)
        self.nb = params.nb
        self.nxe, self.nze = model.nx + 2 * params.nb, model.nz + 2 * params.nb

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def _err(self):
        raise RuntimeError(self.L.orc_last_error().decode())

    @property
    def n_int(self):
        return self.model.nx * self.model.nz

    def table(self, which: str) -> np.ndarray:
        idx = {"vdt2": 0, "sponge": 1, "wavelet": 2, "coeffs": 3}[which]
        h = self.params.order // 2
        n = {0: self.nxe * self.nze, 1: self.nxe * self.nze, 2: self.nt, 3: 2 * h + 1}[idx]
        out = np.zeros(n, dtype=self.dtype)
        fn = self.L.orc_get_table_f32 if self.prec == 32 else self.L.orc_get_table_f64
        if fn(self.h, idx, out) != 0:
            self._err()
        return out

    def map_geometry(self, shot: Shot):
        keep: list = []
        s = shot_struct(shot, keep)
        src = C.c_longlong()
        rec = np.zeros(max(shot.nrec, 1), dtype=np.int64)
        if self.L.orc_map_geometry(self.h, C.byref(s), C.byref(src), rec) != 0:
            self._err()
        return int(src.value), rec[: shot.nrec]

    def forward(self, shot: Shot, snapshots: bool = True, record: bool = False):
        snap = np.zeros(self.nsnap * self.n_int, dtype=self.dtype) if snapshots else None
        rec = np.zeros(shot.nrec * self.nt, dtype=self.dtype) if record else None
        keep: list = []
        s = shot_struct(shot, keep)
        fn = self.L.orc_forward_f32 if self.prec == 32 else self.L.orc_forward_f64
        if fn(self.h, C.byref(s), None if snap is None else snap.ctypes.data,
              None if rec is None else rec.ctypes.data) != 0:
            self._err()
        return snap, (None if rec is None else rec.reshape(shot.nrec, self.nt))

    def backward(self, shot: Shot, snapshots: np.ndarray, image: np.ndarray, illum: np.ndarray):
        keep: list = []
        s = shot_struct(shot, keep)
        fn = self.L.orc_backward_f32 if self.prec == 32 else self.L.orc_backward_f64
        if fn(self.h, C.byref(s), snapshots, image, illum) != 0:
            self._err()

    def migrate(self, shots):
        image = np.zeros(self.n_int, dtype=self.dtype)
        illum = np.zeros(self.n_int, dtype=self.dtype)
        keep: list = []
        arr = _shots_array(shots, keep)
        fn = self.L.orc_migrate_f32 if self.prec == 32 else self.L.orc_migrate_f64
        if fn(self.h, len(shots), arr, image, illum) != 0:
            self._err()
        return image, illum

    def model_shot(self, shot: Shot) -> np.ndarray:
        out = np.zeros(shot.nrec * self.nt, dtype=np.float32)
        keep: list = []
        s = shot_struct(shot, keep)
        if self.L.orc_model_shot_f32(self.h, C.byref(s), out) != 0:
            self._err()
        return out.reshape(shot.nrec, self.nt)

    def times(self):
        f, b, i = C.c_double(), C.c_double(), C.c_double()
        self.L.orc_times(self.h, C.byref(f), C.byref(b), C.byref(i))
        return f.value, b.value, i.value


def coefficients(order: int, precision: int = 32) -> np.ndarray:
    L = lib()
    h = L.orc_stencil_half(order)
    if h < 0:
        raise ValueError("unsupported order")
    if precision == 32:
        out = np.zeros(h + 1, np.float32)
        L.orc_stencil_coefficients_f32(order, out)
    else:
        out = np.zeros(h + 1, np.float64)
        L.orc_stencil_coefficients_f64(order, out)
    return out


def mute_direct_wave(shot: Shot, v_surface: float, t_pad: float, t_taper: float) -> np.ndarray:
    tr = np.ascontiguousarray(shot.traces, dtype=np.float32).reshape(-1).copy()
    keep: list = []
    s = shot_struct(shot, keep)
    lib().orc_mute_direct_wave(C.byref(s), tr, v_surface, t_pad, t_taper)
    return tr.reshape(shot.nrec, shot.nt)


# ---- the reference itself (oracle/_ref) ------------------------------------

_ref = None


def illumination_compensation(image: np.ndarray, illum: np.ndarray, eps_relative: float = 1e-4) -> np.ndarray:
    """I / (S + eps_relative*max S + 1e-30) on a copy (src/rtm/imaging.cpp:48-56)."""
    out = np.array(image, np.float32, copy=True).reshape(-1)
    s = np.ascontiguousarray(illum, np.float32).reshape(-1)
    lib().orc_illumination_compensation(out, s, out.size, eps_relative)
    return out


def laplacian_filter(image: np.ndarray, nx: int, nz: int, dx: float, dz: float) -> np.ndarray:
    """Laplacian high-pass of a copy, 0 on the outer ring (src/rtm/imaging.cpp:63-75)."""
    out = np.array(image, np.float32, copy=True).reshape(-1)
    lib().orc_laplacian_filter(out, nx, nz, dx, dz)
    return out


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("oracle/_ref/librtmref.so not built (reference sources absent)")
        L = C.CDLL(REF_SO)
        L.ref_last_error.restype = C.c_char_p
        L.ref_num_threads.argtypes = [C.c_char_p]
        L.ref_migrate.argtypes = [C.c_char_p, C.POINTER(OrcGrid), _f32p, C.POINTER(OrcParams),
                                  C.c_int, C.c_float, C.c_int, C.POINTER(OrcShot), _f32p, _f32p,
                                  _f64p]
        L.ref_model_shots.argtypes = [C.c_char_p, C.POINTER(OrcGrid), _f32p, C.POINTER(OrcParams),
                                      C.c_int, C.c_float, C.c_int, C.POINTER(OrcShot), _f32p]
        L.ref_forward_snapshots.argtypes = [C.c_char_p, C.POINTER(OrcGrid), _f32p,
                                            C.POINTER(OrcParams), C.c_int, C.c_float,
                                            C.POINTER(OrcShot), _f32p]
        L.ref_mute_direct_wave.argtypes = [C.POINTER(OrcShot), _f32p, C.c_float, C.c_float, C.c_float]
        _ref = L
    return _ref


def ref_migrate(engine: str, model: Model, params: Params, nt: int, dt: float, shots):
    L = ref()
    v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
    image = np.zeros(model.nx * model.nz, np.float32)
    illum = np.zeros_like(image)
    times = np.zeros(4, np.float64)
    g, p = _grid(model), _params(params)
    keep: list = []
    arr = _shots_array(shots, keep)
    if L.ref_migrate(engine.encode(), C.byref(g), v, C.byref(p), nt, dt, len(shots), arr,
                     image, illum, times) != 0:
        raise RuntimeError(L.ref_last_error().decode())
    return image, illum, times


def ref_model_shots(engine: str, model: Model, params: Params, nt: int, dt: float, shots):
    L = ref()
    v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
    out = np.zeros(sum(s.nrec for s in shots) * nt, np.float32)
    g, p = _grid(model), _params(params)
    keep: list = []
    arr = _shots_array(shots, keep)
    if L.ref_model_shots(engine.encode(), C.byref(g), v, C.byref(p), nt, dt, len(shots), arr, out) != 0:
        raise RuntimeError(L.ref_last_error().decode())
    res, off = [], 0
    for s in shots:
        n = s.nrec * nt
        res.append(out[off:off + n].reshape(s.nrec, nt))
        off += n
    return res


def ref_forward_snapshots(engine: str, model: Model, params: Params, nt: int, dt: float, shot: Shot):
    L = ref()
    v = np.ascontiguousarray(model.v, dtype=np.float32).reshape(-1)
    nsnap = (nt - 1) // params.store_interval + 1
    out = np.zeros(nsnap * model.nx * model.nz, np.float32)
    g, p = _grid(model), _params(params)
    keep: list = []
    s = shot_struct(shot, keep)
    if L.ref_forward_snapshots(engine.encode(), C.byref(g), v, C.byref(p), nt, dt, C.byref(s), out) != 0:
        raise RuntimeError(L.ref_last_error().decode())
    return out


def ref_mute(shot: Shot, v: float, t_pad: float, t_taper: float) -> np.ndarray:
    tr = np.ascontiguousarray(shot.traces, dtype=np.float32).reshape(-1).copy()
    keep: list = []
    s = shot_struct(shot, keep)
    ref().ref_mute_direct_wave(C.byref(s), tr, v, t_pad, t_taper)
    return tr.reshape(shot.nrec, shot.nt)
