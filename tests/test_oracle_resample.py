"""Pins of the oracle's spectral resampling (expand, interpolate/fft.F90:
832-927) and of the block box filter used for coarse-grained data.

expand: a field whose modes lie below the coarse Nyquist frequency is
band-limited, so zero-padding its spectrum must give the continuous
function at the fine points exactly (up to round-off).  The coarse Nyquist
coefficient is placed as an ordinary cosine of the fine grid (the
reference's lines 879-887), so (−1)^i expands to 2 cos(π i'/2): the closed
form pins that placement.  Box filter: constants and the mean are kept, and
a single Fourier mode is multiplied by the Dirichlet kernel of the block.
"""
import numpy as np
import pytest

from oracle.pyoracle import OracleError, box_filter, expand


def grid(nz, ny, nx):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return k, j, i


@pytest.mark.parametrize("shape", [(3, 8, 12), (2, 16, 16), (1, 30, 20)])
def test_expand_band_limited_is_exact(shape):
    nz, ny, nx = shape
    rng = np.random.default_rng(1)
    k, j, i = grid(nz, ny, nx)
    K, J, I = grid(nz, 2 * ny, 2 * nx)

    def field(jj, ii, n_y, n_x):
        f = 0.3 + 0.0 * ii
        for mx in range(0, nx // 2):
            for my in range(0, ny // 2):
                a, ph = rng.standard_normal(), rng.uniform(0, 2 * np.pi)
                f = f + a * np.cos(2 * np.pi * (mx * ii / n_x + my * jj / n_y) + ph)
        return f
    state = rng.bit_generator.state
    coarse = field(j, i, ny, nx) * (1 + k)
    rng.bit_generator.state = state
    fine = field(J, I, 2 * ny, 2 * nx) * (1 + K)
    out = expand(coarse)
    assert out.shape == (nz, 2 * ny, 2 * nx)
    assert np.abs(out - fine).max() < 1e-12 * np.abs(fine).max()
    # the coarse points are kept
    assert np.abs(out[:, ::2, ::2] - coarse).max() < 1e-12 * np.abs(coarse).max()


def test_expand_places_nyquist_as_a_cosine():
    nz, ny, nx = 1, 8, 8
    k, j, i = grid(nz, ny, nx)
    K, J, I = grid(nz, 2 * ny, 2 * nx)
    out = expand((-1.0) ** i)
    np.testing.assert_allclose(out, 2 * np.cos(np.pi * I / 2), atol=1e-13)
    out = expand((-1.0) ** j)
    np.testing.assert_allclose(out, 2 * np.cos(np.pi * J / 2), atol=1e-13)


def test_expand_rejects_odd_sizes():
    with pytest.raises(OracleError):
        expand(np.zeros((2, 8, 7)))


@pytest.mark.parametrize("f", [1, 2, 4])
def test_box_filter_constants_mean_and_modes(f):
    nz, ny, nx = 8, 16, 32
    rng = np.random.default_rng(2)
    a = rng.standard_normal((nz, ny, nx))
    out = box_filter(a, f)
    assert out.shape == (nz // f, ny // f, nx // f)
    assert abs(out.mean() - a.mean()) < 1e-14
    np.testing.assert_allclose(box_filter(np.full((nz, ny, nx), 2.5), f), 2.5, rtol=0, atol=1e-15)
    # Dirichlet kernel of the block on a single x mode
    k, j, i = grid(nz, ny, nx)
    m = 3
    mode = np.cos(2 * np.pi * m * i / nx)
    got = box_filter(mode, f)
    K, J, I = grid(nz // f, ny // f, nx // f)
    g = np.sin(np.pi * m * f / nx) / (f * np.sin(np.pi * m / nx))
    want = g * np.cos(2 * np.pi * m * (f * I + (f - 1) / 2) / nx)
    assert np.abs(got - want).max() < 1e-14
    # brute force on a small block
    assert abs(out[0, 0, 0] - a[:f, :f, :f].mean()) < 1e-15


def test_box_filter_rejects_non_dividing_factor():
    with pytest.raises(OracleError):
        box_filter(np.zeros((8, 16, 30)), 4)
