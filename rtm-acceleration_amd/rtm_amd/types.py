"""Plain data types of the engine, mirroring the reference's
include/rtm_types.hpp:16-57 (Grid is folded into VelocityModel here)."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class VelocityModel:
    """Interior velocity, m/s, depth-fastest: v[ix*nz + iz]."""
    nx: int
    nz: int
    dx: float
    dz: float
    v: np.ndarray
    ox: float = 0.0
    oz: float = 0.0

    def __post_init__(self):
        self.v = np.ascontiguousarray(self.v, dtype=np.float32).reshape(-1)
        if self.v.size != self.nx * self.nz:
            raise ValueError("velocity model size does not match nx*nz")

    def vmin(self) -> float:
        return float(self.v.min())

    def vmax(self) -> float:
        return float(self.v.max())


@dataclass
class RTMParams:
    """Defaults as in the reference (include/rtm_types.hpp:33-40)."""
    nb: int = 60
    order: int = 8
    sponge_alpha: float = 0.0053
    store_interval: int = 1
    f0: float = 12.0
    verbose: int = 0


@dataclass
class TimeAxis:
    nt: int
    dt: float


@dataclass
class ShotRecord:
    """Source/receivers in metres; traces trace-major, shape (nrec, nt)."""
    sx: float
    sz: float
    rx: np.ndarray
    rz: np.ndarray
    nt: int
    dt: float
    traces: np.ndarray | None = None

    def __post_init__(self):
        self.rx = np.ascontiguousarray(self.rx, dtype=np.float32).reshape(-1)
        self.rz = np.ascontiguousarray(self.rz, dtype=np.float32).reshape(-1)
        if self.rx.size != self.rz.size:
            raise ValueError("rx and rz differ in length")
        if self.traces is not None:
            self.traces = np.ascontiguousarray(self.traces, dtype=np.float32).reshape(self.nrec, self.nt)

    @property
    def nrec(self) -> int:
        return int(self.rx.size)

    def allocate(self):
        self.traces = np.zeros((self.nrec, self.nt), np.float32)
