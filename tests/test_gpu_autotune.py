"""Autotuner: every launch shape it can choose must give the oracle's
results; the alternative radix plans are forced for both directions and the
operators and whole steps compared with the oracle."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu
NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.mark.parametrize("shape", [(64, 128, 16), (256, 160, 10), (128, 64, 12)])
def test_alternative_plans_match_oracle(shape):
    from fug30v0kxc_amd import Solver
    nx, ny, nz = shape
    print(f"[autotune] alternative plans {shape}", flush=True)
    p = reference_params()
    s = Solver(nx, ny, nz, p)
    s.set_tune(variant_x=1, variant_y=1, mult_x=2, mult_y=2, zc=8)
    o = Oracle(nx, ny, nz, p)
    rng = np.random.default_rng(1)
    f = rng.standard_normal((nz, ny, nx))
    for op in ("d1bydx", "d1bydy", "forwardfft", "inversefft", "dealias"):
        got = getattr(s, op)(f)
        want = getattr(o, op)(f)
        assert rel(got, want) < 1e-12, op
    for op in ("d2bydx", "d2bydy"):
        g1, g2 = getattr(s, op)(f)
        w1, w2 = getattr(o, op)(f)
        assert rel(g1, w1) < 1e-12 and rel(g2, w2) < 1e-12, op
    s.close()


def test_autotune_then_steps_match_oracle():
    from fug30v0kxc_amd import Solver
    print("[autotune] tune + steps 64x64x48", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = Solver(64, 64, 48, p)
    s.static()
    t = s.autotune()
    assert all(v in (0, 1, 2) for v in t["variant_x"] + t["variant_y"])
    assert all(m in (1, 2) for m in t["mult_x"] + t["mult_y"]) and t["zc"] >= 8
    assert t["ms_x"] > 0 and t["ms_col"] > 0
    o = Oracle(64, 64, 48, p)
    o.setup_static()
    s.step(12, 1e-4, 0)
    o.step(12, 1e-4, 0)
    st = s.get_state()
    for n in NAMES:
        assert rel(st[n], o[n]) < 1e-10, n
    s.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 1e-5)])
def test_autotune_between_steps_matches_oracle(prec, tol):
    """The tuner runs the fused hydro kernel into the buffers the next step
    writes; tuning between steps leaves the run unchanged."""
    from fug30v0kxc_amd import Solver
    print(f"[autotune] tune between steps 64x64x48 precision {prec}", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = Solver(64, 64, 48, p, precision=prec)
    s.static()
    o = Oracle(64, 64, 48, p)
    o.setup_static()
    s.step(5, 1e-4, 0)
    s.autotune()
    s.step(7, 1e-4, 0)
    o.step(12, 1e-4, 0)
    st = s.get_state()
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    for n in NAMES:
        scale = bscale if n.startswith("b") else np.abs(o[n]).max()
        assert np.abs(st[n] - o[n]).max() / scale < tol, n
    s.close()


def test_forced_alternative_steps_match_oracle():
    from fug30v0kxc_amd import MHD_PREC_MIXED, Solver
    print("[autotune] forced alternatives, steps 64x64x48, both precisions", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    o = Oracle(64, 64, 48, p)
    o.setup_static()
    o.step(12, 1e-4, 0)
    for prec, tol in ((0, 1e-10), (MHD_PREC_MIXED, 1e-5)):
        s = Solver(64, 64, 48, p, precision=prec)
        # a grid multiple per transform kind, as the tuner may choose
        s.set_tune(variant_x=1, variant_y=1, mult_x=[1, 2, 2, 1], mult_y=[2, 1, 1, 2], zc=12)
        s.static()
        s.step(12, 1e-4, 0)
        st = s.get_state()
        bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
        for n in NAMES:
            scale = bscale if n.startswith("b") else np.abs(o[n]).max()
            assert np.abs(st[n] - o[n]).max() / scale < tol, (prec, n)
        s.close()


@pytest.mark.parametrize("shape", [(256, 64, 10), (64, 256, 10)])
def test_second_alternative_plan_matches_oracle(shape):
    # 256 points have a second alternative plan (two radix-16 passes)
    from fug30v0kxc_amd import Solver
    nx, ny, nz = shape
    print(f"[autotune] second alternative plan {shape}", flush=True)
    p = reference_params()
    s = Solver(nx, ny, nz, p)
    s.set_tune(variant_x=2 if nx == 256 else 0, variant_y=2 if ny == 256 else 0, mult_x=1, mult_y=2, zc=8)
    o = Oracle(nx, ny, nz, p)
    f = np.random.default_rng(2).standard_normal((nz, ny, nx))
    for op in ("d1bydx", "d1bydy", "forwardfft", "inversefft", "dealias"):
        assert rel(getattr(s, op)(f), getattr(o, op)(f)) < 1e-12, op
    for op in ("d2bydx", "d2bydy"):
        g1, g2 = getattr(s, op)(f)
        w1, w2 = getattr(o, op)(f)
        assert rel(g1, w1) < 1e-12 and rel(g2, w2) < 1e-12, op
    s.close()


@pytest.mark.parametrize("prec,tol,vy", [(0, 1e-10, 2), (1, 1e-5, 2), (1, 1e-5, [1, 2, 0, 2])])
def test_second_alternative_plan_steps_match_oracle(prec, tol, vy):
    # vy: one plan for every transform kind along y, or one per kind (d1, d1+d2, r2hc, hc2r)
    from fug30v0kxc_amd import Solver
    print(f"[autotune] plans {vy}, steps 256x256x12 precision {prec}", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    o = Oracle(256, 256, 12, p)
    o.setup_static()
    o.step(4, 1e-4, 0)
    s = Solver(256, 256, 12, p, precision=prec)
    s.set_tune(variant_x=2, variant_y=vy, mult_x=1, mult_y=1, zc=8)
    s.static()
    s.step(4, 1e-4, 0)
    st = s.get_state()
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    for n in NAMES:
        scale = bscale if n.startswith("b") else np.abs(o[n]).max()
        assert np.abs(st[n] - o[n]).max() / scale < tol, n
    s.close()


def test_set_tune_rejects_invalid_choices():
    from fug30v0kxc_amd import MhdError, Solver
    s = Solver(48, 48, 8, reference_params())         # 48 has no alternative plan
    with pytest.raises(MhdError):
        s.set_tune(variant_x=1)
    with pytest.raises(MhdError):
        s.set_tune(mult_x=0)
    with pytest.raises(MhdError):
        s.set_tune(mult_y=[1, 2, 9, 1])
    s.close()
    s = Solver(128, 256, 8, reference_params())       # 128: one alternative; 256: two
    s.set_tune(variant_x=1, variant_y=2)
    with pytest.raises(MhdError):
        s.set_tune(variant_x=2)
    with pytest.raises(MhdError):
        s.set_tune(variant_y=3)
    s.close()
