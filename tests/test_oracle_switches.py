"""The reference's physics switches (src/control.f90:23-34) in the oracle,
pinned by what each one imposes:

radbc      the top row satisfies the one-sided third-order form of
           dT/dz = theta T^4 (src/dbydz.F90:4037-4057);
constflux  the bottom row has the one-sided third-order dT/dz = theta
           (4133-4143);
perfect    pol = 0 and zero one-sided d/dz of tor, Bx-bar, By-bar on both
           walls (4314-4400); the static field is B = x (initial.F90:63-71);
dynamo     F = 1 and a seed field (magnetic.f90:29-57) that bandj returns;
kinematic  F = 0: the flow feels no Lorentz force, B still evolves.
"""
import numpy as np

from oracle.pyoracle import (OPT_CONSTFLUX, OPT_DYNAMO, OPT_KINEMATIC, OPT_PERFECT, OPT_RADBC,
                             Oracle)
from tests.conftest import reference_params

NX, NY, NZ = 16, 12, 10


def make(opts, **kw):
    p = dict(reference_params(**kw), opts=opts)
    o = Oracle(NX, NY, NZ, p)
    o.setup_static()
    return o, p


def one_sided(f0, f1, f2, f3):
    """11 f0 - 18 f1 + 9 f2 - 2 f3 = 6 dz * (one-sided third-order df/dz)."""
    return 11 * f0 - 18 * f1 + 9 * f2 - 2 * f3


def test_radbc_top_row():
    o, p = make(OPT_RADBC, t_amp=0.1)
    rng = np.random.default_rng(0)
    o["t"][1:4] += 0.05 * rng.standard_normal((3, NY, NX))
    o.boundary()
    t = o["t"]
    dz = 1.0 / (NZ - 1)
    resid = 6 * p["theta"] * dz * t[0] ** 4 + one_sided(t[0], t[1], t[2], t[3])
    assert np.abs(resid).max() < 1e-12
    assert np.all(t[-1] == 1 + p["theta"])            # bottom stays fixed


def test_constflux_bottom_row():
    o, p = make(OPT_CONSTFLUX, t_amp=0.1)
    o.boundary()
    t = o["t"]
    dz = 1.0 / (NZ - 1)
    slope = one_sided(t[-1], t[-2], t[-3], t[-4]) / (6 * dz)   # backward d/dz at the bottom
    assert np.abs(slope - p["theta"]).max() < 1e-11
    assert np.all(t[0] == 1.0)


def test_perfect_conductor_static_and_walls():
    o, p = make(OPT_PERFECT, b_amp=0.0)
    assert np.all(o["bx"] == 1.0) and np.all(o["by"] == 0.0) and np.all(o["bz"] == 0.0)
    assert np.allclose(o["bxbar"], 1.0) and abs(o.scalar("bzbar")) < 1e-15
    rng = np.random.default_rng(1)
    o["tor"][:] = rng.standard_normal(o["tor"].shape)
    o["pol"][:] = rng.standard_normal(o["pol"].shape)
    o["bxbar"][:] = rng.standard_normal(NZ)
    o["bybar"][:] = rng.standard_normal(NZ)
    o.mag_boundary()
    assert np.all(o["pol"][0] == 0.0) and np.all(o["pol"][-1] == 0.0)
    for f in (o["tor"], o["bxbar"], o["bybar"]):
        assert np.abs(one_sided(f[0], f[1], f[2], f[3])).max() < 1e-12
        assert np.abs(one_sided(f[-1], f[-2], f[-3], f[-4])).max() < 1e-12


def test_dynamo_seed_field():
    o, p = make(OPT_DYNAMO)
    o.bandj()
    i = np.arange(NX)[None, None, :]
    j = np.arange(NY)[None, :, None]
    seed = 0.003 * np.cos(2 * np.pi * (i + 1) / NX) * np.cos(2 * np.pi * (j + 1) / NY)
    assert np.abs(o["bz"] - seed).max() < 1e-15
    assert np.abs(o["bx"]).max() < 1e-15 and np.abs(o["by"]).max() < 1e-15
    # F = 1: the magnetic energy (a volume integral over xmax x ymax x 1) is reported as F * ME
    d = o.step(2, 1e-4, 1)
    assert abs(d[3] - 0.5 * (0.003 ** 2) / 4 * p["xmax"] * p["ymax"]) < 1e-3 * d[3]


def test_dynamo_perfect_seed_field():
    o, p = make(OPT_DYNAMO | OPT_PERFECT)
    o.bandj()
    i = np.arange(NX)[None, None, :]
    j = np.arange(NY)[None, :, None]
    assert np.abs(o["bx"] - 0.003 * np.cos(2 * np.pi * (j + 1) / NY)).max() < 1e-15
    assert np.abs(o["by"] - 0.003 * np.cos(2 * np.pi * (i + 1) / NX)).max() < 1e-15


def test_kinematic_flow_ignores_the_field():
    o1, _ = make(OPT_KINEMATIC, b_amp=1e-2)
    o2, _ = make(OPT_KINEMATIC, b_amp=0.0)
    rng = np.random.default_rng(3)
    for o in (o1, o2):
        o["u"][:] = 1e-2 * rng.standard_normal((NZ, NY, NX))
        rng = np.random.default_rng(3)
    for o in (o1, o2):
        o.hydro_deriv()
    for n in ("udot", "vdot", "wdot", "tdot", "rdot"):
        assert np.array_equal(o1[n][2], o2[n][2]), n
    # ... while the field is advected and diffused
    o1.mag_deriv()
    assert np.abs(o1["poldot"][2]).max() > 0
    d = o1.step(2, 1e-4, 1)
    assert d[3] > 0                                # ME reported without the F factor
