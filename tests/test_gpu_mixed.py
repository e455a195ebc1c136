"""The mixed-precision path (MHD_PREC_MIXED) against the FP64 oracle.

The mixed path does its transforms and stencils on FP32 fluctuation arrays:
T − T₀(z), ρ − ρ₀(z), u, v, w, and B without its means. It keeps FP64 for
the state, the means, the backgrounds, the point-wise arithmetic and the
reductions. The bar is ≤ 1e-5 relative error against the FP64 oracle for
Nu, kinetic and magnetic energy, and for every dumped field in the
max-norm relative to the field's largest value.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
TOL = 1e-5


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def mixed(nx, ny, nz, p):
    from fug30v0kxc_amd import MHD_PREC_MIXED, Solver
    return Solver(nx, ny, nz, p, precision=MHD_PREC_MIXED)


def test_mixed_static_and_state_round_trip():
    print("[mixed] static 16x16x12", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = mixed(16, 16, 12, p)
    s.static()
    o = Oracle(16, 16, 12, p)
    o.setup_static()
    st = s.get_state()
    for n in ("r", "t", "u", "v", "w"):
        assert np.array_equal(st[n], o[n]), n        # FP64 state
    for n in ("bx", "by", "bz"):
        # FP32 fluctuations about FP64 means
        assert np.abs(st[n] - o[n]).max() < 1e-9, n
    for n in ("bxbar", "bybar"):
        assert np.abs(s.field(n) - o[n]).max() < 1e-12, n
    assert abs(s.scalars()["bzbar"] - o.scalar("bzbar")) < 1e-12
    assert rel(s.field("pol"), o["pol"]) < 1e-5
    s.close()


@pytest.fixture(scope="module")
def pair():
    print("[mixed] RHS 32x24x16", flush=True)
    p = reference_params(b_amp=1e-3)
    o = Oracle(32, 24, 16, p)
    o.setup_static()
    rng = np.random.RandomState(11)
    for n in ("u", "v", "w"):
        o[n][:] = 1e-2 * rng.randn(16, 24, 32)
    o["w"][[0, -1]] = 0.0
    s = mixed(32, 24, 16, p)
    s.set_state(*[o[n] for n in NAMES])
    yield s, o
    s.close()


def test_mixed_hydro_rhs(pair):
    s, o = pair
    s.hydro_deriv()
    o.hydro_deriv()
    for n in ("rdot", "tdot", "udot", "vdot", "wdot"):
        e = rel(s.field(n, 3), o[n][2])
        print(f"  {n} rel err {e:.2e}")
        assert e < TOL, n


def test_mixed_mag_rhs(pair):
    s, o = pair
    s.mag_deriv()
    o.mag_deriv()
    for n in ("poldot", "tordot"):
        e = rel(s.field(n, 3), o[n][2])
        print(f"  {n} rel err {e:.2e}")
        assert e < TOL, n


@pytest.mark.parametrize("box,steps", [((16, 16, 12), 40), ((64, 64, 48), 30), ((512, 512, 8), 10)],
                         ids=["16x16x12-40", "64x64x48-30", "512x512x8-10"])
def test_mixed_steps_against_fp64_oracle(box, steps):
    print(f"[mixed] steps {box} x {steps}", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = mixed(*box, p)
    s.static()
    o = Oracle(*box, p)
    o.setup_static()
    ds = s.step(steps, 1e-4, 10)
    do = o.step(steps, 1e-4, 10)
    st = s.get_state()
    for n in NAMES:
        e = rel(st[n], o[n])
        print(f"  {n} rel err {e:.2e}")
        assert e < TOL, n
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        e = abs(ds[key] - do[i]) / abs(do[i])
        print(f"  {key} rel err {e:.2e}")
        assert e < TOL, key
    s.close()


def test_mixed_variable_dt():
    print("[mixed] variable dt 16x16x12", flush=True)
    p = reference_params(t_amp=1e-3)
    s = mixed(16, 16, 12, p)
    s.static()
    o = Oracle(16, 16, 12, p)
    o.setup_static()
    s.step(25, 0.0, 5)
    o.step(25, 0.0, 5)
    assert abs(s.scalars()["dt"] - o.scalar("dt")) <= 1e-6 * o.scalar("dt")
    st = s.get_state()
    for n in NAMES:
        assert rel(st[n], o[n]) < TOL, n
    s.close()


def test_mixed_rejects_setting_b_by_field():
    from fug30v0kxc_amd import MhdError
    p = reference_params()
    s = mixed(16, 16, 8, p)
    s.static()
    with pytest.raises(MhdError):
        s.set_field("bx", np.zeros((8, 16, 16)))
    s.close()


def test_mixed_large_box_keeps_fp64_backgrounds():
    """A large box (a 3 GB pool) in the mixed mode: its FP64 backgrounds T0(z),
    rho0(z) must survive the pool's initial fill, or T is carried as full FP32
    values and drifts from the FP64 run by ~1e-7 within a few steps instead of
    ~1e-10."""
    import fug30v0kxc_amd as fa
    print("[mixed] 256x256x160 backgrounds", flush=True)
    p = reference_params(ra=1e6, q=1000.0, t_amp=1e-3, b_amp=1e-3)
    a = fa.Solver(256, 256, 160, p, precision=fa.MHD_PREC_FP64)
    a.static()
    b = fa.Solver(256, 256, 160, p, precision=fa.MHD_PREC_MIXED)
    b.static()
    a.step(10, 5e-5, 0)
    b.step(10, 5e-5, 0)
    sa, sb = a.get_state(), b.get_state()
    for n in ("t", "r"):
        e = np.abs(sa[n] - sb[n]).max() / np.abs(sa[n]).max()
        assert e < 1e-9, (n, e)
