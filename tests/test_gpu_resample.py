"""GPU spectral resampling (the reference's expand) and box filter against
the oracle, through the C ABI."""
import numpy as np
import pytest

from oracle import pyoracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(3, 8, 12), (4, 64, 64), (2, 30, 20), (5, 96, 128), (2, 14, 10)])
def test_expand_matches_oracle(shape):
    import fug30v0kxc_amd as fa
    print(f"[resample] expand {shape}", flush=True)
    rng = np.random.default_rng(7)
    a = rng.standard_normal(shape) + 3.0        # white noise: Nyquist content included
    got = fa.expand(a)
    want = pyoracle.expand(a)
    assert got.shape == want.shape
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_expand_keeps_coarse_points_of_band_limited_fields():
    import fug30v0kxc_amd as fa
    nz, ny, nx = 2, 32, 48
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a = np.cos(2 * np.pi * (5 * i / nx + 3 * j / ny)) + 0.5 * np.sin(2 * np.pi * 7 * j / ny) * (1 + k)
    out = fa.expand(a)
    assert np.abs(out[:, ::2, ::2] - a).max() < 1e-13


@pytest.mark.parametrize("f", [1, 2, 4])
def test_box_filter_matches_oracle(f):
    import fug30v0kxc_amd as fa
    print(f"[resample] box filter {f}", flush=True)
    rng = np.random.default_rng(8)
    a = rng.standard_normal((16, 32, 64))
    got = fa.box_filter(a, f)
    want = pyoracle.box_filter(a, f)
    assert np.abs(got - want).max() <= 1e-15 * max(1.0, np.abs(want).max())


def test_errors():
    import fug30v0kxc_amd as fa
    with pytest.raises(fa.MhdError) as e:
        fa.expand(np.zeros((2, 8, 7)))
    assert e.value.code == -1
    with pytest.raises(fa.MhdError) as e:
        fa.expand(np.zeros((2, 1026, 8)))          # 2 ny = 2052 is past the longest transform
    assert e.value.code == -20
    with pytest.raises(fa.MhdError):
        fa.box_filter(np.zeros((8, 16, 30)), 4)
