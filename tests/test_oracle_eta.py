"""eta(z) extension of the induction equation in the oracle (the reference
has a constant eta): eta(z) = zeta C_k [1 + (eta_c - 1)(1 + tanh((z -
eta_z0)/eta_w))/2], applied as E = u x B - (eta(z) - zeta C_k) J.

Pins:
* a z-independent field Bz = b cos(kx) (pol ∝ cos(kx), tor = 0) carries the
  current J = (0, -dBz/dx, 0) and decays at the local rate eta(z) k^2, so
  poldot = -eta(z) k^2 pol exactly, with tordot = 0 and no mean-field
  change;
* Ohmic heating scales with the local diffusivity:
  tdot = F zeta C_k (gamma - 1) (eta(z)/eta_0) J^2 / rho;
* eta_c = 1 reproduces the constant-eta right-hand sides exactly.
"""
import numpy as np

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

NX, NY, NZ = 16, 12, 10


def eta_ratio(p, z):
    return 1.0 + (p["eta_c"] - 1.0) * 0.5 * (1.0 + np.tanh((z - p["eta_z0"]) / p["eta_w"]))


def box(**kw):
    p = dict(reference_params(), eta_c=10.0, eta_z0=0.5, eta_w=0.05)
    p.update(kw)
    o = Oracle(NX, NY, NZ, p)
    o.setup_static()
    z = np.arange(NZ) / (NZ - 1)
    return o, p, z


def test_profile():
    o, p, z = box()
    np.testing.assert_allclose(o["etar"], eta_ratio(p, z), rtol=1e-15)
    assert o["etar"][0] < 1.001 and o["etar"][-1] > 9.99


def test_local_decay_rate_of_a_vertical_field_mode():
    o, p, z = box()
    for n in ("u", "v", "w"):
        o[n][:] = 0.0
    x = np.arange(NX) * p["xmax"] / NX
    kx = 2 * np.pi * 3 / p["xmax"]
    P = np.cos(kx * x)[None, None, :] * np.ones((NZ, NY, 1))
    o["pol"][:] = o.forwardfft(P)
    o["tor"][:] = 0.0
    o["bxbar"][:] = 0.0
    o["bybar"][:] = 0.0
    o.bandj()
    o.mag_deriv()
    eta0 = p["zeta"] * p["ck"]
    want = -eta0 * eta_ratio(p, z)[:, None, None] * kx ** 2 * P
    got = o.inversefft(o["poldot"][2])
    assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    assert np.abs(o.inversefft(o["tordot"][2])).max() < 1e-12 * np.abs(want).max()
    assert np.abs(o["bxbardot"][2]).max() < 1e-14 and np.abs(o["bybardot"][2]).max() < 1e-14


def test_ohmic_heating_scales_with_eta():
    o, p, z = box()
    for n in ("u", "v", "w", "bx", "by"):
        o[n][:] = 0.0
    o["bz"][:] = 1.0
    o["r"][:] = 2.0
    o["t"][:] = 1.0
    j2 = 0.1 + 0.05 * np.cos(np.arange(NX) * 2 * np.pi / NX)[None, None, :] * np.ones((NZ, NY, 1))
    o["jsq"][:] = j2
    o.hydro_deriv()
    want = p["f"] * p["zeta"] * p["ck"] * (p["gamma"] - 1) * eta_ratio(p, z)[:, None, None] * j2 / 2.0
    assert np.abs(o["tdot"][2] - want).max() < 1e-11 * np.abs(want).max()


def test_unit_contrast_is_the_reference():
    o1, p1, _ = box(eta_c=1.0)
    o2 = Oracle(NX, NY, NZ, reference_params())
    o2.setup_static()
    for o in (o1, o2):
        o["u"][:] = 1e-2 * np.random.default_rng(4).standard_normal((NZ, NY, NX))
    d1, d2 = o1.step(5, 1e-4, 5), o2.step(5, 1e-4, 5)
    for n in ("u", "t", "bx", "bz"):
        assert np.array_equal(o1[n], o2[n]), n
