"""Judging a closure-coupled coarse run against box-filtered fine states
(host-side analysis on numpy arrays; no counterpart in the reference).

The measures are the horizontal kinetic-energy spectrum (shells of integer
|k| on the coarse grid, averaged over z), its relative L1 distance up to the
2/3 cutoff, the a-priori fraction of subgrid-flux variance a prediction
explains (the centred coefficient of determination), and the rule that
falls back to the plain coarse run when the coupled run's spectral error or
its Nusselt number strays beyond a threshold from the filtered fine run.
"""
import numpy as np

FALLBACK_SPECTRAL_ERROR = 0.10
FALLBACK_NUSSELT_ERROR = 0.10


def energy_spectrum(state, kmax=None):
    """E(k), k = 0..kmax, of 0.5 (u^2 + v^2 + w^2) per plane, averaged over z."""
    nz, ny, nx = state["u"].shape
    if kmax is None:
        kmax = int(min(nx, ny) // 2 * 2 // 3)
    kx = np.fft.fftfreq(nx) * nx
    ky = np.fft.fftfreq(ny) * ny
    shell = np.rint(np.sqrt(kx[None, :] ** 2 + ky[:, None] ** 2)).astype(int)
    e = np.zeros((ny, nx))
    for n in ("u", "v", "w"):
        h = np.fft.fft2(state[n], axes=(1, 2)) / (nx * ny)
        e += 0.5 * (np.abs(h) ** 2).mean(axis=0)
    out = np.zeros(kmax + 1)
    for k in range(kmax + 1):
        out[k] = e[shell == k].sum()
    return out


def spectral_error(state, reference, kmax=None):
    """sum_k |E(k) - E_ref(k)| / sum_k E_ref(k) over 1 <= k <= kmax."""
    a = energy_spectrum(state, kmax)[1:]
    b = energy_spectrum(reference, kmax)[1:]
    return float(np.abs(a - b).sum() / b.sum())


def explained_variance(prediction, target):
    """Per-output centred R^2 = 1 - sum (p - t)^2 / sum (t - mean t)^2: the
    fraction of the subgrid flux's variance the prediction accounts for.
    Predicting the mean scores 0, as does the plain coarse method."""
    p = np.asarray(prediction).reshape(len(prediction), -1)
    t = np.asarray(target).reshape(len(target), -1)
    tc = t - t.mean(axis=1, keepdims=True)
    return 1.0 - ((p - t) ** 2).sum(axis=1) / (tc ** 2).sum(axis=1)


def nusselt_error(nu, nu_ref):
    """|Nu - Nu_ref| / Nu_ref."""
    return abs(nu - nu_ref) / abs(nu_ref)


def choose(closure_error, threshold=FALLBACK_SPECTRAL_ERROR, nu_error=None,
           nu_threshold=FALLBACK_NUSSELT_ERROR):
    """'closure' while the coupled run stays within the threshold of the
    filtered fine spectra and, when given, its Nusselt number within
    nu_threshold of the filtered fine run's; else 'plain' (the fallback)."""
    if closure_error > threshold:
        return "plain"
    if nu_error is not None and nu_error > nu_threshold:
        return "plain"
    return "closure"
