"""GPU parity of the reference's per-step routines (mag_init, hydro_deriv,
mag_deriv, bandj, boundary, diagnostics) and of whole time steps, through the
C ABI, against the CPU oracle on identical seeded inputs (FP64)."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def make_pair(nx, ny, nz, seed=11, **kw):
    from fug30v0kxc_amd import Solver
    p = reference_params(**kw)
    o = Oracle(nx, ny, nz, p)
    o.setup_static()
    rng = np.random.RandomState(seed)
    # a non-trivial flow on top of the seeded static state
    for n in ("u", "v", "w"):
        o[n][:] = 1e-2 * rng.randn(nz, ny, nx)
    o["w"][[0, -1]] = 0.0
    s = Solver(nx, ny, nz, p)
    s.set_state(*[o[n] for n in NAMES])
    return s, o


@pytest.fixture(scope="module")
def pair():
    print("[physics] box 32x24x16", flush=True)
    s, o = make_pair(32, 24, 16, b_amp=1e-3)
    yield s, o
    s.close()


def test_mag_init(pair):
    s, o = pair
    for n in ("pol", "tor", "jsq"):
        assert rel(s.field(n), o[n]) < 1e-11, n
    for n in ("bxbar", "bybar"):
        assert np.abs(s.field(n) - o[n]).max() < 1e-14, n


def test_hydro_deriv(pair):
    s, o = pair
    s.hydro_deriv()
    o.hydro_deriv()
    for n in ("rdot", "tdot", "udot", "vdot", "wdot"):
        assert rel(s.field(n, 3), o[n][2]) < 1e-11, n


def test_mag_deriv(pair):
    s, o = pair
    s.mag_deriv()
    o.mag_deriv()
    for n in ("poldot", "tordot"):
        assert rel(s.field(n, 3), o[n][2]) < 1e-11, n
    for n in ("bxbardot", "bybardot"):
        got, ref = s.field(n, 3), o[n][2]
        assert np.abs(got - ref).max() <= 1e-11 * max(np.abs(ref).max(), 1e-12), n


def test_bandj(pair):
    s, o = pair
    s.bandj()
    o.bandj()
    for n in ("bx", "by", "bz", "jsq"):
        assert rel(s.field(n), o[n]) < 1e-11, n


def test_boundaries(pair):
    s, o = pair
    for n in ("u", "v", "w", "t", "pol", "tor"):
        s.set_field(n, o[n])         # identical inputs: the BC rows are pure algebra
    s.set_field("bxbar", o["bxbar"])
    s.set_field("bybar", o["bybar"])
    s.boundary()
    s.mag_boundary()
    o.boundary()
    o.mag_boundary()
    for n in ("u", "v", "w", "t", "pol", "tor"):
        assert rel(s.field(n), o[n]) < 1e-13, n


def test_diagnostics(pair):
    s, o = pair
    assert abs(s.nusselt() - o.nusselt()) < 1e-12
    db_s, db_o = s.checkdivb(), o.checkdivb()
    assert abs(db_s[0] - db_o[0]) < 1e-10 and abs(db_s[1] - db_o[1]) < 1e-10
    for n in ("t", "u", "bz"):
        mm_s, mm_o = s.sminmax(n), o.sminmax(o[n])
        assert np.allclose(mm_s, mm_o, rtol=1e-14, atol=0)
    assert abs(s.calcdt() - o.calcdt()) < 1e-15 * o.calcdt() * 10
    ga_s, ga_o = np.array(s.GlobalAverages()), o.global_averages()
    assert np.allclose(ga_s, ga_o, rtol=1e-11, atol=1e-18)
    assert rel(s.field("u"), o["u"]) < 1e-12


def test_diagnostics_tall_box():
    # nz = 520: the reductions' per-plane partials (8 values a plane) exceed the
    # finishing workgroup's LDS and are combined in device memory instead
    print("[physics] diagnostics 8x8x520", flush=True)
    s, o = make_pair(8, 8, 520, seed=5, b_amp=1e-3)
    try:
        assert abs(s.nusselt() - o.nusselt()) < 1e-12
        db_s, db_o = s.checkdivb(), o.checkdivb()
        assert abs(db_s[0] - db_o[0]) < 1e-10 and abs(db_s[1] - db_o[1]) < 1e-10
        for n in ("t", "u"):
            assert np.allclose(s.sminmax(n), o.sminmax(o[n]), rtol=1e-14, atol=0)
        assert abs(s.calcdt() - o.calcdt()) < 1e-15 * o.calcdt() * 10
        ga_s, ga_o = np.array(s.GlobalAverages()), o.global_averages()
        assert np.allclose(ga_s, ga_o, rtol=1e-11, atol=1e-18)
    finally:
        s.close()


@pytest.mark.parametrize("box,steps", [((16, 16, 12), 30), ((64, 64, 48), 12)],
                         ids=["16x16x12-30", "64x64x48-12"])
def test_steps_fixed_dt(box, steps):
    print(f"[physics] steps {box} x {steps}", flush=True)
    s, o = make_pair(*box, seed=3, b_amp=1e-3)
    ds = s.step(steps, 1e-4, 10)
    do = o.step(steps, 1e-4, 10)
    st = s.get_state()
    for n in NAMES:
        assert rel(st[n], o[n]) < 1e-10, n
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= 1e-10 * abs(do[i]), key
    sc = s.scalars()
    assert sc["nit"] == steps and abs(sc["ttime"] - o.scalar("ttime")) < 1e-15
    s.close()


def test_steps_variable_dt():
    print("[physics] variable dt 16x16x12", flush=True)
    s, o = make_pair(16, 16, 12, seed=4)
    s.step(25, 0.0, 5)
    o.step(25, 0.0, 5)
    assert abs(s.scalars()["dt"] - o.scalar("dt")) <= 1e-12 * o.scalar("dt")
    st = s.get_state()
    for n in NAMES:
        assert rel(st[n], o[n]) < 1e-10, n
    s.close()


def test_static_matches_oracle_seeding():
    from fug30v0kxc_amd import Solver
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = Solver(16, 16, 8, p)
    s.static()
    o = Oracle(16, 16, 8, p)
    o.setup_static()
    st = s.get_state()
    for n in NAMES:
        assert np.array_equal(st[n], o[n]), n
    assert rel(s.field("pol"), o["pol"]) < 1e-12
    s.close()
