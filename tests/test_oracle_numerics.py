"""Pins of the oracle's numerics against what the mathematics fixes.

Each check is chosen so that a plausible slip (a wrong coefficient, a dropped
factor, a transposed table) fails it.  References: src/rtm/rtm.cpp:14-70
(coefficients, Ricker, CFL, dispersion), src/rtm/rtm_cpu.cpp:33-79 (tables),
src/rtm/rtm_cpu.cpp:99-123 (geometry), src/io/seismic_io.cpp:18-35 (mute).
"""
from fractions import Fraction
import math

import numpy as np
import pytest

import oracle_lib as O

ORDERS = [2, 4, 6, 8, 10, 12]


def taylor_coefficients(order):
    """Exact central second-derivative weights from the Taylor system:
    sum_k c_k * 2 k^(2m) / (2m)! ... solved in rationals (c_0 + 2 sum c_k = 0,
    2 sum c_k k^2 = 2, 2 sum c_k k^(2m) = 0 for m = 2..h)."""
    h = order // 2
    # unknowns c_1..c_h: sum_k c_k k^(2m) = [m == 1] for m = 1..h
    A = [[Fraction(k) ** (2 * m) for k in range(1, h + 1)] for m in range(1, h + 1)]
    b = [Fraction(1 if m == 1 else 0) for m in range(1, h + 1)]
    n = h
    for i in range(n):                       # Gauss-Jordan in exact arithmetic
        p = next(r for r in range(i, n) if A[r][i] != 0)
        A[i], A[p] = A[p], A[i]
        b[i], b[p] = b[p], b[i]
        for r in range(n):
            if r != i and A[r][i] != 0:
                f = A[r][i] / A[i][i]
                A[r] = [a - f * c for a, c in zip(A[r], A[i])]
                b[r] -= f * b[i]
    ck = [b[i] / A[i][i] for i in range(n)]
    c0 = -2 * sum(ck)
    return [c0] + ck


@pytest.mark.parametrize("order", ORDERS)
def test_coefficients_are_the_taylor_weights(order):
    exact = taylor_coefficients(order)
    c32 = O.coefficients(order, 32)
    c64 = O.coefficients(order, 64)
    assert len(c32) == order // 2 + 1
    for k, e in enumerate(exact):
        assert c64[k] == pytest.approx(float(e), rel=1e-15, abs=1e-300)
        # float set: the float quotient of the exact fraction (one rounding)
        assert abs(float(c32[k]) - float(e)) <= abs(float(e)) * 2.0 ** -23


@pytest.mark.parametrize("order", ORDERS)
def test_coefficients_differentiate_polynomials_exactly(order):
    """sum c_k f(x+k) is f'' for every polynomial of degree <= order+1."""
    c = O.coefficients(order, 64)
    h = order // 2
    x = 0.37
    for deg in range(order + 2):
        f = lambda t: t ** deg
        approx = c[0] * f(x) + sum(c[k] * (f(x + k) + f(x - k)) for k in range(1, h + 1))
        exact = deg * (deg - 1) * x ** (deg - 2) if deg >= 2 else 0.0
        assert approx == pytest.approx(exact, abs=1e-9 * max(1.0, (x + h) ** deg))


def test_stencil_abs_sum_values():
    # the values quoted for the CFL bound: 4, 16/3, 6.0444, 6.5016, 6.8267, 7.0729
    expect = {2: 4.0, 4: 5.3333, 6: 6.0444, 8: 6.5016, 10: 6.8267, 12: 7.0729}
    for order, e in expect.items():
        exact = sum(abs(c) for c in taylor_coefficients(order)[1:]) * 2 + abs(taylor_coefficients(order)[0])
        assert O.lib().orc_stencil_abs_sum(order) == pytest.approx(float(exact), rel=2e-6)
        assert O.lib().orc_stencil_abs_sum(order) == pytest.approx(e, abs=6e-5)
    assert O.lib().orc_stencil_half(7) == -1


@pytest.mark.parametrize("t,f0", [(0.0, 12.0), (0.01, 12.0), (-0.05, 15.0), (0.1, 10.0), (0.02, 30.0)])
def test_ricker_closed_form(t, f0):
    a = math.pi ** 2 * f0 ** 2 * t ** 2
    assert O.lib().orc_ricker(t, f0) == pytest.approx((1 - 2 * a) * math.exp(-a), rel=3e-6, abs=1e-7)
    assert O.lib().orc_ricker(0.0, f0) == 1.0


def test_cfl_and_dispersion_closed_forms():
    for order in ORDERS:
        S = float(sum(abs(c) for c in taylor_coefficients(order)[1:]) * 2 + abs(taylor_coefficients(order)[0]))
        for vmax, dx, dz in [(3000.0, 10.0, 10.0), (5775.0, 4.0, 4.0), (4500.0, 5.0, 7.5)]:
            exact = 2.0 / (vmax * math.sqrt(S * (1 / dx ** 2 + 1 / dz ** 2)))
            assert O.lib().orc_max_stable_dt(vmax, dx, dz, order) == pytest.approx(exact, rel=1e-6)
    assert O.lib().orc_points_per_wavelength(1500.0, 12.0, 10.0, 8.0) == pytest.approx(1500 / 30 / 10, rel=1e-6)


def _setup(nx=23, nz=17, nb=7, order=4, dx=10.0, dz=12.0, nt=50, dt=0.001, f0=15.0, alpha=0.0053):
    rng = np.random.default_rng(1)
    v = rng.uniform(1500, 3000, (nx, nz)).astype(np.float32)
    m = O.Model(nx, nz, dx, dz, v.reshape(-1))
    p = O.Params(nb=nb, order=order, sponge_alpha=alpha, store_interval=1, f0=f0)
    return m, p, O.Oracle(m, p, nt, dt), O.Oracle(m, p, nt, dt, precision=64)


def test_tables_closed_forms():
    m, p, o32, o64 = _setup()
    nxe, nze, nb = m.nx + 2 * p.nb, m.nz + 2 * p.nb, p.nb
    v = m.v.reshape(m.nx, m.nz).astype(np.float64)
    # vdt2: edge replication of (v dt)^2 -- a transposed or shifted pad fails here
    vdt2 = o64.table("vdt2").reshape(nxe, nze)
    ix = np.clip(np.arange(nxe) - nb, 0, m.nx - 1)
    iz = np.clip(np.arange(nze) - nb, 0, m.nz - 1)
    dt = float(np.float32(0.001))
    assert np.allclose(vdt2, (v[ix][:, iz] * dt) ** 2, rtol=1e-14)
    assert np.allclose(o32.table("vdt2").reshape(nxe, nze), vdt2, rtol=3e-7)
    # sponge: separable Cerjan taper, exactly 1 at distance >= nb from every edge
    w = o32.table("sponge").reshape(nxe, nze)
    alpha = float(np.float32(p.sponge_alpha))

    def taper(n):
        d = np.minimum(np.arange(n), n - 1 - np.arange(n))
        return np.where(d >= nb, 1.0, np.exp(-(alpha * (nb - d)) ** 2))
    assert np.allclose(w, np.outer(taper(nxe), taper(nze)), rtol=1e-6)
    assert np.all(w[nb:nxe - nb, nb:nze - nb] == 1.0)
    assert np.all(w[:nb, :] < 1.0) and np.all(w[:, :nb] < 1.0)
    # wavelet: Ricker delayed by 1.2/f0
    t = np.arange(50) * dt - 1.2 / 15.0
    a = np.pi ** 2 * 15.0 ** 2 * t ** 2
    assert np.allclose(o64.table("wavelet"), (1 - 2 * a) * np.exp(-a), rtol=1e-6, atol=1e-12)
    # coefficients pre-divided by the spacing
    c = o64.table("coeffs")
    cc = O.coefficients(4, 64)
    assert c[0] == pytest.approx(cc[0] * (1 / 100.0 + 1 / 144.0))
    assert np.allclose(c[1:3], cc[1:] / 100.0) and np.allclose(c[3:5], cc[1:] / 144.0)


def test_geometry_brute_force():
    m, p, o32, _ = _setup(nx=30, nz=20, nb=9, dx=10.0, dz=10.0)
    rx = np.array([0.0, 4.9, 5.0, 5.1, 15.0, 289.9, 290.0, 300.0, -20.0, 1000.0], np.float32)
    rz = np.array([0.0, 14.99, 15.0, 25.0, -3.0, 190.0, 195.0, 400.0, 0.0, 50.0], np.float32)
    shot = O.Shot(55.0, 45.0, rx, rz, 10, 0.001)
    src, rec = o32.map_geometry(shot)
    nze = m.nz + 2 * p.nb

    def lround(q):                          # half away from zero
        return int(math.copysign(math.floor(abs(q) + 0.5), q))

    def brute(x, z):
        ix = min(max(lround(float(np.float32(x) / np.float32(10.0))), 0), m.nx - 1)
        iz = min(max(lround(float(np.float32(z) / np.float32(10.0))), 0), m.nz - 1)
        return (ix + p.nb) * nze + iz + p.nb
    assert src == brute(55.0, 45.0)
    assert list(rec) == [brute(x, z) for x, z in zip(rx, rz)]


def test_mute_closed_form():
    nt, dt = 400, 0.001
    rx = np.array([0.0, 100.0, 500.0, 1500.0], np.float32)
    shot = O.Shot(100.0, 0.0, rx, np.zeros_like(rx), nt, dt, np.ones((4, nt), np.float32))
    v, tp = 1500.0, 1.0 / 12.0
    out = O.mute_direct_wave(shot, v, tp, tp)
    for ir, x in enumerate(rx):
        i0 = int(np.float32(np.float32(abs(x - 100.0)) / np.float32(v) + np.float32(tp)) / np.float32(dt))
        ntap = max(1, int(np.float32(tp) / np.float32(dt)))
        assert np.all(out[ir, :min(i0 + 1, nt)] == 0.0)
        for it in range(i0 + 1, min(i0 + ntap, nt)):
            u = (it - i0) / ntap
            assert out[ir, it] == pytest.approx(0.5 * (1 - math.cos(math.pi * u)), abs=2e-6)
        assert np.all(out[ir, min(i0 + ntap, nt):] == 1.0)


def test_laplacian_filter_closed_form():
    """On I = a*ix^2 + b*iz^2 + c*ix*iz + linear terms the second differences
    are exact: the filter gives 2a/dx^2 + 2b/dz^2 inside and 0 on the outer
    ring; a transposed axis or a wrong spacing changes the value."""
    nx, nz, dx, dz = 9, 7, 2.0, 4.0
    ix, iz = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij")
    img = (3.0 * ix**2 + 5.0 * iz**2 + 1.0 * ix * iz + 7.0 * ix - 2.0 * iz + 1.0).astype(np.float32)
    out = O.laplacian_filter(img.reshape(-1), nx, nz, dx, dz).reshape(nx, nz)
    want = 2 * 3.0 / dx**2 + 2 * 5.0 / dz**2
    assert np.all(out[1:-1, 1:-1] == np.float32(want))
    assert not out[0].any() and not out[-1].any() and not out[:, 0].any() and not out[:, -1].any()


def test_illumination_compensation_closed_form():
    """I / (S + eps*max S + 1e-30): with S = 0 everywhere the divisor is 1e-30;
    with max S = 4 and eps 0.5 it is S + 2."""
    img = np.array([1.0, -2.0, 3.0, 0.0], np.float32)
    S = np.array([0.0, 4.0, 2.0, 1.0], np.float32)
    out = O.illumination_compensation(img, S, 0.5)
    assert np.array_equal(out, (img / (S + np.float32(2.0) + np.float32(1e-30))).astype(np.float32))
    assert out[1] == np.float32(-2.0 / 6.0)
    z = O.illumination_compensation(img, np.zeros(4, np.float32), 1e-4)
    assert np.array_equal(z, img / np.float32(1e-30))
