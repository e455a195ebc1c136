"""Pins the oracle's horizontal spectral operators (src/fft.F90).

The reference delegates its real transforms to FFTW 2.1.5 (absent here), so
the oracle's transform is checked against an O(n^2) long-double DFT written
from the definition, and the derivative / dealias operators against the
closed forms that src/test_dbydxy.F90:118-335 prints (with thresholds).
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, hc2r, r2hc
from tests.conftest import reference_params


def brute_r2hc(x):
    """FFTW-2 halfcomplex of x from the DFT definition in long double."""
    n = len(x)
    xl = np.asarray(x, dtype=np.longdouble)
    t = np.arange(n, dtype=np.longdouble)
    out = np.zeros(n, dtype=np.longdouble)
    pi = np.longdouble("3.14159265358979323846264338327950288")
    for m in range(n // 2 + 1):
        ang = -2 * pi * m * t / n
        out[m] = np.sum(xl * np.cos(ang))
        if 0 < m < (n + 1) // 2:
            out[n - m] = np.sum(xl * np.sin(ang))
    return out


@pytest.mark.parametrize("n", [8, 30, 64, 96])
def test_r2hc_matches_long_double_dft(n):
    x = np.random.RandomState(n).randn(n)
    ref = brute_r2hc(x).astype(np.float64)
    got = r2hc(x)
    assert np.abs(got - ref).max() <= 4e-15 * np.abs(ref).max()


@pytest.mark.parametrize("n", [8, 30, 64, 96])
def test_hc2r_is_unnormalised_inverse(n):
    x = np.random.RandomState(100 + n).randn(n)
    back = hc2r(r2hc(x)) / n
    assert np.abs(back - x).max() < 1e-14


def test_halfcomplex_storage_order():
    # x = cos(2 pi t/8): r_1 = 4 at index 1; sin(2 pi t/8): i_1 = -4 at index n-1
    t = np.arange(8)
    h = r2hc(np.cos(2 * np.pi * t / 8))
    assert abs(h[1] - 4) < 1e-14 and np.abs(np.delete(h, 1)).max() < 1e-14
    h = r2hc(np.sin(2 * np.pi * t / 8))
    assert abs(h[7] + 4) < 1e-14 and np.abs(np.delete(h, 7)).max() < 1e-14


def _box(nx=16, ny=12, nz=8, xmax=4.0, ymax=3.0):
    return Oracle(nx, ny, nz, reference_params(xmax=xmax, ymax=ymax))


def test_derivatives_closed_form_x_and_y():
    o = _box()
    nx, ny, nz = o.nx, o.ny, o.nz
    x = np.arange(nx) * 4.0 / nx
    y = np.arange(ny) * 3.0 / ny
    X = np.broadcast_to(x[None, None, :], (nz, ny, nx))
    Y = np.broadcast_to(y[None, :, None], (nz, ny, nx))
    for iw in range(nx // 2 + 1):
        om = iw * 2 * np.pi / 4.0
        for f, d1, d2 in ((np.cos(om * X), -om * np.sin(om * X), -om * om * np.cos(om * X)),
                          (np.sin(om * X), om * np.cos(om * X), -om * om * np.sin(om * X))):
            if iw == nx // 2:
                d1 = 0 * d1                       # d1 zeroes the Nyquist mode
            a = o.d1bydx(f)
            b1, b2 = o.d2bydx(f)
            assert np.abs(a - d1).max() < 1e-12 * max(1, om)
            assert np.abs(b1 - d1).max() < 1e-12 * max(1, om)
            assert np.abs(b2 - d2).max() < 1e-12 * max(1, om * om)   # d2 keeps Nyquist
    for iw in range(ny // 2 + 1):
        om = iw * 2 * np.pi / 3.0
        f, d1, d2 = np.cos(om * Y), -om * np.sin(om * Y), -om * om * np.cos(om * Y)
        if iw == ny // 2:
            d1 = 0 * d1
        a = o.d1bydy(f)
        b1, b2 = o.d2bydy(f)
        assert np.abs(a - d1).max() < 1e-12 * max(1, om)
        assert np.abs(b1 - d1).max() < 1e-12 * max(1, om)
        assert np.abs(b2 - d2).max() < 1e-12 * max(1, om * om)


def test_dealias_zeroes_top_third_and_is_idempotent():
    o = _box(nx=24, ny=18)
    nx, ny, nz = o.nx, o.ny, o.nz
    x = np.arange(nx) * 4.0 / nx
    X = np.broadcast_to(x[None, None, :], (nz, ny, nx))
    icut = (nx // 2) // 3
    for iw in range(nx // 2 + 1):
        f = np.cos(iw * 2 * np.pi / 4.0 * X)
        g = o.dealias(f)
        killed = iw >= nx // 2 - icut
        assert np.abs(g - (0 * f if killed else f)).max() < 1e-13
    r = np.random.RandomState(3).randn(nz, ny, nx)
    g = o.dealias(r)
    assert np.abs(o.dealias(g) - g).max() < 1e-13


def test_forward_inverse_2d_round_trip_and_layout():
    o = _box()
    nx, ny, nz = o.nx, o.ny, o.nz
    r = np.random.RandomState(5).randn(nz, ny, nx)
    F = o.forwardfft(r)
    assert np.abs(o.inversefft(F) - r).max() < 1e-13
    # double-halfcomplex: x-halfcomplex rows, then y-halfcomplex of each column
    ref = np.apply_along_axis(r2hc, 2, r)
    ref = np.apply_along_axis(r2hc, 1, ref)
    assert np.abs(F - ref).max() < 1e-12
    # DC equals the plane sum
    assert np.abs(F[:, 0, 0] - r.sum(axis=(1, 2))).max() < 1e-12
