"""The host-side measures of closure runs (fug30v0kxc_amd/closure_eval.py):
the centred explained variance, the energy spectrum and the fallback rule.
No GPU is used."""
import numpy as np

from fug30v0kxc_amd import closure_eval as ce


def test_explained_variance_is_centred():
    rng = np.random.default_rng(0)
    t = rng.standard_normal((3, 500)) + np.array([[5.0], [-2.0], [0.0]])
    assert np.allclose(ce.explained_variance(t, t), 1.0)
    # the mean alone explains nothing, however large the mean is
    mean_only = np.broadcast_to(t.mean(axis=1, keepdims=True), t.shape)
    assert np.allclose(ce.explained_variance(mean_only, t), 0.0, atol=1e-12)
    # half the fluctuation: 1 - 1/4
    half = t.mean(axis=1, keepdims=True) + 0.5 * (t - t.mean(axis=1, keepdims=True))
    assert np.allclose(ce.explained_variance(half, t), 0.75)


def test_energy_spectrum_of_a_single_mode():
    nz, ny, nx = 3, 16, 16
    x = np.arange(nx) * 2 * np.pi / nx
    u = np.broadcast_to(np.cos(3 * x)[None, None, :], (nz, ny, nx))
    st = {"u": u, "v": np.zeros_like(u), "w": np.zeros_like(u)}
    e = ce.energy_spectrum(st)
    # 0.5 <u^2> = 0.25 in shell 3 (two Fourier coefficients of 1/4 each)
    assert abs(e[3] - 0.25) < 1e-14 and abs(e.sum() - 0.25) < 1e-14
    assert ce.spectral_error(st, st) == 0.0


def test_fallback_rule_takes_nusselt_into_account():
    assert ce.choose(0.05) == "closure"
    assert ce.choose(0.15) == "plain"
    assert ce.choose(0.05, nu_error=0.2) == "plain"
    assert ce.choose(0.05, nu_error=0.02) == "closure"
    assert abs(ce.nusselt_error(1.1, 1.0) - 0.1) < 1e-15
