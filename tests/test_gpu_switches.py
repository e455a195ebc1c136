"""The reference's physics switches on the GPU against the oracle: static
start, then whole time steps with each option (and combinations), FP64 to
1e-10 relative in every field; one mixed-precision case to 1e-5."""
import numpy as np
import pytest

from oracle.pyoracle import (OPT_CONSTFLUX, OPT_DYNAMO, OPT_KINEMATIC, OPT_PERFECT, OPT_RADBC,
                             Oracle)
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
CASES = {
    "radbc": OPT_RADBC,
    "constflux": OPT_CONSTFLUX,
    "perfect": OPT_PERFECT,
    "kinematic": OPT_KINEMATIC,
    "dynamo": OPT_DYNAMO,
    "dynamo+perfect": OPT_DYNAMO | OPT_PERFECT,
    "radbc+constflux+perfect": OPT_RADBC | OPT_CONSTFLUX | OPT_PERFECT,
}


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def field_errors(st, o):
    """Max-norm error per field relative to the field's largest value; the
    components of B relative to the largest |B| component (a component can
    stay at round-off level, e.g. bx of the dynamo seed)."""
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    out = {}
    for n in NAMES:
        scale = bscale if n in ("bx", "by", "bz") else max(np.abs(o[n]).max(), 1e-300)
        out[n] = np.abs(st[n] - o[n]).max() / scale
    return out


def run_pair(opts, precision, steps=20, box=(16, 16, 12)):
    from fug30v0kxc_amd import Solver
    p = dict(reference_params(t_amp=1e-2, b_amp=1e-3), opts=opts)
    s = Solver(*box, p, precision=precision)
    s.static()
    o = Oracle(*box, p)
    o.setup_static()
    ds = s.step(steps, 1e-4, 10)
    do = o.step(steps, 1e-4, 10)
    return s, o, ds, do


@pytest.mark.parametrize("name", list(CASES))
def test_switch_steps_fp64(name):
    from fug30v0kxc_amd import MHD_PREC_FP64
    print(f"[switches] {name}", flush=True)
    s, o, ds, do = run_pair(CASES[name], MHD_PREC_FP64)
    for n, e in field_errors(s.get_state(), o).items():
        assert e < 1e-10, (name, n, e)
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= 1e-10 * max(abs(do[i]), 1e-300), (name, key)
    s.close()


def test_switch_steps_mixed():
    from fug30v0kxc_amd import MHD_PREC_MIXED
    print("[switches] mixed radbc+constflux+perfect", flush=True)
    s, o, ds, do = run_pair(OPT_RADBC | OPT_CONSTFLUX | OPT_PERFECT, MHD_PREC_MIXED)
    for n, e in field_errors(s.get_state(), o).items():
        assert e < 1e-5, (n, e)
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= 1e-5 * abs(do[i]), key
    s.close()


ETA_CASES = {
    "eta10@0.5": dict(eta_c=10.0, eta_z0=0.5, eta_w=0.05),
    "eta100@0.3": dict(eta_c=100.0, eta_z0=0.3, eta_w=0.05),
}


@pytest.mark.parametrize("name", list(ETA_CASES))
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_eta_profile_steps(name, prec):
    """eta(z) extension (E = u x B - (eta - eta_0) J, Ohmic heating eta J^2)."""
    from fug30v0kxc_amd import MHD_PREC_FP64, MHD_PREC_MIXED, Solver
    print(f"[switches] {name} {prec}", flush=True)
    p = dict(reference_params(t_amp=1e-2, b_amp=1e-3), **ETA_CASES[name])
    s = Solver(16, 16, 12, p, precision=MHD_PREC_FP64 if prec == "fp64" else MHD_PREC_MIXED)
    s.static()
    o = Oracle(16, 16, 12, p)
    o.setup_static()
    ds = s.step(20, 1e-4, 10)
    do = o.step(20, 1e-4, 10)
    tol = 1e-10 if prec == "fp64" else 1e-5
    for n, e in field_errors(s.get_state(), o).items():
        assert e < tol, (name, n, e)
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= tol * abs(do[i]), (name, key)
    s.close()
