"""The CPU-runnable configuration end to end: 64x64x48, aspect 4, gamma 5/3,
m = 1, theta = 10, Ra = 1e5, Pr = 1, Q = 100, zeta_0 = 0.2, uniform vertical
B0, seeded T perturbation 1e-3, 200 AB3 steps at fixed dt = 1e-4.

* plain method (the reference's kappa(z), constant eta), FP64: every field,
  Nu, KE and ME within 1e-10 of the oracle;
* with the eta(z) step of contrast 10 at z = 0.5 (width 0.05): FP64 within
  1e-10, mixed precision within 1e-5.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

BOX = (64, 64, 48)
STEPS, DT = 200, 1e-4
NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
ETA = dict(eta_c=10.0, eta_z0=0.5, eta_w=0.05)


def errors(st, o):
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    return {n: np.abs(st[n] - o[n]).max() /
            (bscale if n in ("bx", "by", "bz") else max(np.abs(o[n]).max(), 1e-300)) for n in NAMES}


def oracle_run(p):
    o = Oracle(*BOX, p)
    o.setup_static()
    d = o.step(STEPS, DT, 100)
    return o, d


def gpu_run(p, precision):
    from fug30v0kxc_amd import Solver
    s = Solver(*BOX, p, precision=precision)
    s.static()
    d = s.step(STEPS, DT, 100)
    st = s.get_state()
    s.close()
    return st, d


def check(st, ds, o, do, tol):
    for n, e in errors(st, o).items():
        print(f"  {n} {e:.2e}")
        assert e < tol, n
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        e = abs(ds[key] - do[i]) / abs(do[i])
        print(f"  {key} {e:.2e}")
        assert e < tol, key


def test_c1_plain_fp64():
    from fug30v0kxc_amd import MHD_PREC_FP64
    print("[configs] 64x64x48 x 200, plain, fp64", flush=True)
    p = reference_params(ra=1e5, q=100.0, t_amp=1e-3)
    o, do = oracle_run(p)
    st, ds = gpu_run(p, MHD_PREC_FP64)
    check(st, ds, o, do, 1e-10)


@pytest.fixture(scope="module")
def eta_oracle():
    p = dict(reference_params(ra=1e5, q=100.0, t_amp=1e-3), **ETA)
    o, do = oracle_run(p)
    return p, o, do


def test_c1_eta_fp64(eta_oracle):
    from fug30v0kxc_amd import MHD_PREC_FP64
    print("[configs] 64x64x48 x 200, eta(z) x10, fp64", flush=True)
    p, o, do = eta_oracle
    st, ds = gpu_run(p, MHD_PREC_FP64)
    check(st, ds, o, do, 1e-10)


def test_c1_eta_mixed(eta_oracle):
    from fug30v0kxc_amd import MHD_PREC_MIXED
    print("[configs] 64x64x48 x 200, eta(z) x10, mixed", flush=True)
    p, o, do = eta_oracle
    st, ds = gpu_run(p, MHD_PREC_MIXED)
    check(st, ds, o, do, 1e-5)
