"""Deterministic synthetic velocity models and acquisition geometries.

These build the workloads the engine is measured on.  The reference ships
only a built-in three-layer model with a box diffractor
(src/tools/make_synthetic.cpp:90-118) and the Marmousi file; the larger models
below have no reference counterpart and are defined here, seeded, so that the
same call returns the same array on any machine (numpy PCG64 + float64 maths,
rounded once to float32).

    layered_3()          201 x 151 @ 10 m, 1500/2200/3000 m/s, interfaces at
                         z = 500 m and 1000 m (the 4-shot parity workload)
    faulted_layers(...)  dipping layers offset by normal faults, velocity
                         rising with depth, smooth +-5 % perturbation
                         (2301 x 751 @ 4 m and 1201 x 401 @ 8 m workloads)
    salt_body(...)       1500-4000 m/s gradient with a ~4500 m/s salt body
                         made of overlapping ellipses (4001 x 1201 @ 5 m)
"""
from __future__ import annotations

import numpy as np

from .types import ShotRecord, VelocityModel


def layered_3(nx: int = 201, nz: int = 151, dx: float = 10.0, dz: float = 10.0,
              velocities=(1500.0, 2200.0, 3000.0), interfaces=(500.0, 1000.0)) -> VelocityModel:
    """Flat layers; a cell at depth z = iz*dz takes the velocity of the layer
    whose top is the deepest interface <= z."""
    z = np.arange(nz, dtype=np.float64) * dz
    col = np.full(nz, velocities[0], np.float64)
    for vel, zi in zip(velocities[1:], interfaces):
        col[z >= zi] = vel
    v = np.tile(col.astype(np.float32), (nx, 1))
    return VelocityModel(nx, nz, dx, dz, v.reshape(-1))


def _smooth_noise(rng, nx: int, nz: int, corr_x: float, corr_z: float) -> np.ndarray:
    """Zero-mean, unit-peak smooth field: white noise filtered by a Gaussian
    in the wavenumber domain (correlation lengths in cells)."""
    w = rng.standard_normal((nx, nz))
    kx = np.fft.fftfreq(nx)[:, None]
    kz = np.fft.rfftfreq(nz)[None, :]
    filt = np.exp(-2.0 * (np.pi ** 2) * ((kx * corr_x) ** 2 + (kz * corr_z) ** 2))
    f = np.fft.irfft2(np.fft.rfft2(w) * filt, s=(nx, nz))
    f -= f.mean()
    m = np.abs(f).max()
    return f / m if m > 0 else f


def faulted_layers(nx: int = 2301, nz: int = 751, dx: float = 4.0, dz: float = 4.0,
                   nlayers: int = 20, nfaults: int = 3, vtop: float = 1500.0,
                   vbottom: float = 5500.0, perturbation: float = 0.05,
                   seed: int = 20261203) -> VelocityModel:
    """Gently dipping layers cut by normal faults, v rising with depth.

    Layer boundaries are straight lines with dips drawn in [-3, 3] degrees;
    each fault is a line dipping 60-70 degrees whose hanging wall (the side
    x > fault trace) is dropped by 3-8 % of the model depth.  Layer velocities
    increase linearly from vtop to vbottom, then a smooth random field scales
    the model by up to +-perturbation."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx, dtype=np.float64) * dx)[:, None]
    z = (np.arange(nz, dtype=np.float64) * dz)[None, :]
    depth = (nz - 1) * dz
    width = (nx - 1) * dx

    # Fault throw: at each (x, z) the depth at which to look up the layering.
    zsrc = np.broadcast_to(z, (nx, nz)).copy()
    for _ in range(nfaults):
        x0 = rng.uniform(0.15, 0.85) * width
        dip = np.deg2rad(rng.uniform(60.0, 70.0))
        throw = rng.uniform(0.03, 0.08) * depth
        trace_x = x0 + z / np.tan(dip)            # fault trace x at each depth
        hanging = x > trace_x
        zsrc = np.where(hanging, zsrc - throw, zsrc)

    tops = np.sort(rng.uniform(0.02, 0.98, nlayers - 1)) * depth
    dips = np.deg2rad(rng.uniform(-3.0, 3.0, nlayers - 1))
    layer = np.zeros((nx, nz), np.int32)
    for k in range(nlayers - 1):
        boundary = tops[k] + np.tan(dips[k]) * (x - 0.5 * width)
        layer += (zsrc >= boundary).astype(np.int32)
    vel_of_layer = np.linspace(vtop, vbottom, nlayers)
    v = vel_of_layer[layer]

    noise = _smooth_noise(rng, nx, nz, corr_x=40.0, corr_z=15.0)
    v = v * (1.0 + perturbation * noise)
    return VelocityModel(nx, nz, dx, dz, v.astype(np.float32).reshape(-1))


def salt_body(nx: int = 4001, nz: int = 1201, dx: float = 5.0, dz: float = 5.0,
              vtop: float = 1500.0, vbottom: float = 4000.0, vsalt: float = 4500.0,
              nellipses: int = 6, seed: int = 4500) -> VelocityModel:
    """Linear gradient background with a salt body = union of ellipses centred
    in the middle third of the model."""
    rng = np.random.default_rng(seed)
    x = (np.arange(nx, dtype=np.float64) * dx)[:, None]
    z = (np.arange(nz, dtype=np.float64) * dz)[None, :]
    depth = (nz - 1) * dz
    width = (nx - 1) * dx
    v = vtop + (vbottom - vtop) * (z / depth) + 0.0 * x
    salt = np.zeros((nx, nz), bool)
    for _ in range(nellipses):
        cx = rng.uniform(0.35, 0.65) * width
        cz = rng.uniform(0.35, 0.6) * depth
        ax = rng.uniform(0.06, 0.15) * width
        az = rng.uniform(0.06, 0.14) * depth
        salt |= ((x - cx) / ax) ** 2 + ((z - cz) / az) ** 2 <= 1.0
    v = np.where(salt, vsalt, v)
    return VelocityModel(nx, nz, dx, dz, v.astype(np.float32).reshape(-1))


def crop(model: VelocityModel, ix0: int, width: int) -> VelocityModel:
    """Columns [ix0, ix0+width) of a model (an aperture crop); origin moved."""
    v = model.v.reshape(model.nx, model.nz)[ix0:ix0 + width]
    return VelocityModel(width, model.nz, model.dx, model.dz, np.ascontiguousarray(v).reshape(-1),
                         ox=model.ox + ix0 * model.dx, oz=model.oz)


def surface_shots(model: VelocityModel, nt: int, dt: float, shot_x, rec_dx: float,
                  rec_x0: float = 0.0, src_z: float = 20.0, rec_z: float = 20.0):
    """One ShotRecord per source x with a fixed receiver line from rec_x0 to the
    model end every rec_dx metres (traces unset)."""
    xmax = model.ox + (model.nx - 1) * model.dx
    nrec = int((xmax - rec_x0) / rec_dx) + 1
    rx = (rec_x0 + rec_dx * np.arange(nrec, dtype=np.float64)).astype(np.float32)
    rz = np.full(nrec, rec_z, np.float32)
    return [ShotRecord(float(sx), src_z, rx, rz, nt, dt) for sx in shot_x]


def synth_line(model: VelocityModel, nshots: int, nt: int, dt: float, rec_dx: float,
               src_z: float = 20.0, rec_z: float = 20.0, shot_x0: float = -1.0,
               shot_dx: float = -1.0):
    """The reference's rtm_synth acquisition (src/tools/make_synthetic.cpp:124-131):
    nrec = (int)((nx-1)*dx/rec_dx)+1 receivers at ir*rec_dx; shots evenly
    spread at xmax*(is+1)/(nshots+1), or shot_x0 + is*shot_dx when shot_dx > 0.
    Positions are computed in float32 as the reference does."""
    f32 = np.float32
    nrec = int(f32(model.nx - 1) * f32(model.dx) / f32(rec_dx)) + 1
    rx = np.array([f32(ir) * f32(rec_dx) for ir in range(nrec)], np.float32)
    rz = np.full(nrec, rec_z, np.float32)
    xmax = f32(model.nx - 1) * f32(model.dx)
    shots = []
    for i in range(nshots):
        if shot_dx > 0:
            sx = f32(shot_x0 if shot_x0 >= 0 else 0.0) + f32(shot_dx) * f32(i)
        else:
            sx = xmax * f32(i + 1) / f32(nshots + 1)
        shots.append(ShotRecord(float(sx), src_z, rx, rz, nt, dt))
    return shots
