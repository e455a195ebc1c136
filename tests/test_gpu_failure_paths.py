"""GPU runs of the two failure paths the reference defines inside its time
loop -- the density floor of the AB3 update (src/main.F90:438-443) and the
div B abort (src/main.F90:578-590, abort_comms(12)) -- and GPU hydro_deriv on
the closed-form states that pin its wall rows, through the C ABI."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, OracleError
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
FLOOR_CELLS = [(4, 5, 6), (1, 0, 0), (0, 7, 3), (9, 2, 9), (5, 11, 15)]


def _floor_pair(precision, nx=16, ny=12, nz=10):
    import fug30v0kxc_amd as fa
    p = reference_params(b_amp=1e-3)
    o = Oracle(nx, ny, nz, p)
    o.setup_static()
    rng = np.random.RandomState(3)
    for n in ("u", "v", "w"):
        o[n][:] = 1e-2 * rng.randn(nz, ny, nx)
    o["w"][[0, -1]] = 0.0
    # cells below the floor, in the interior, next to and on both walls
    for c in FLOOR_CELLS:
        o["r"][c] = 0.03
    s = fa.Solver(nx, ny, nz, p, precision=precision)
    s.set_state(*[o[n] for n in NAMES])
    return s, o


def _constant_closure():
    """A closure whose fluxes are constants (W3 = b3 = 0): no effect on the
    run, but its coupling switches the AB3 update to the unfused kernel."""
    import fug30v0kxc_amd as fa
    rng = np.random.RandomState(41)
    params = (0.3 * rng.randn(18828)).astype(np.float32)
    params[18048:] = 0.0
    cl = fa.Closure(seed=0)
    cl.params = params
    cl.norm = np.concatenate([np.zeros(8), np.ones(8), np.linspace(-0.2, 0.3, 12), np.ones(12)])
    return cl


@pytest.mark.parametrize("mode", ["fp64", "mixed", "fp64-unfused"])
def test_density_floor_matches_oracle(mode):
    import fug30v0kxc_amd as fa
    print(f"[floor] {mode} 16x12x10", flush=True)
    prec = fa.MHD_PREC_MIXED if mode == "mixed" else fa.MHD_PREC_FP64
    s, o = _floor_pair(prec)
    cl = None
    if mode == "fp64-unfused":
        cl = _constant_closure()
        s.attach_closure(cl)
    try:
        dt = 1e-5
        for nstep in (1, 3):
            s.step(nstep, dt, 0)
            o.step(nstep, dt, 0)
            if nstep == 1:
                # every planted cell reached the floor in the first (Euler) step
                assert all(o["r"][c] == 0.05 for c in FLOOR_CELLS)
            r = s.get_state()["r"]
            for c in FLOOR_CELLS:
                if o["r"][c] == 0.05:
                    assert r[c] == 0.05, c
                    for lev in (1, 2):
                        assert s.field("rdot", lev)[c] == 0.0, (c, lev)
            st = s.get_state()
            tol = 1e-5 if mode == "mixed" else 1e-10
            # B components against the largest |B| component (the field is
            # mostly vertical), as everywhere in the mixed-precision checks
            bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
            for n in NAMES:
                err = np.abs(st[n] - o[n]).max() / (bscale if n[0] == "b" else np.abs(o[n]).max())
                assert err < tol, (n, err)
            for lev in (1, 2):
                g, ref = s.field("rdot", lev), o["rdot"][lev - 1]
                assert np.abs(g - ref).max() <= tol * np.abs(ref).max(), lev
    finally:
        if cl is not None:
            s.attach_closure(None)
        s.close()


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_divb_abort_returns_12(mode):
    # F = 0 (Q = 0, not kinematic): the reference never calls mag_init
    # (src/main.F90:326-333), so B keeps the non-solenoidal state it was given
    # and checkdivb at the ndu06 step aborts with code 12
    import fug30v0kxc_amd as fa
    print(f"[divb] abort {mode}", flush=True)
    nx, ny, nz = 16, 16, 12
    p = reference_params(q=0.0)
    o = Oracle(nx, ny, nz, p)
    o.setup_static()
    x = (np.arange(nx) * 4.0 / nx)[None, None, :]
    o["bx"][:] = 0.1 * np.sin(2 * np.pi * x / 4.0) + np.zeros((nz, ny, nx))
    prec = fa.MHD_PREC_MIXED if mode == "mixed" else fa.MHD_PREC_FP64
    s = fa.Solver(nx, ny, nz, p, precision=prec)
    s.set_state(*[o[n] for n in NAMES])
    try:
        db = s.checkdivb()
        assert abs(db[0] - 0.1 * 2 * np.pi / 4.0) < 1e-6
        # no diagnostics step: no check
        s.step(2, 1e-5, 0)
        o.step(2, 1e-5, 0)
        with pytest.raises(fa.MhdError) as ei:
            s.step(3, 1e-5, 3)
        assert ei.value.code == 12
        with pytest.raises(OracleError) as eo:
            o.step(3, 1e-5, 3)
        assert eo.value.code == 12
        # the abort comes at the diagnostics step, after its update, as in the
        # reference; the state up to there agrees
        st = s.get_state()
        tol = 1e-5 if mode == "mixed" else 1e-10
        for n in ("r", "t", "u"):
            assert np.abs(st[n] - o[n]).max() <= tol * np.abs(o[n]).max(), n
    finally:
        s.close()


def _closed_form_state(kind, nx=16, ny=12, nz=14):
    z = np.arange(nz) / (nz - 1.0)
    p = reference_params(t_amp=0.0)
    o = Oracle(nx, ny, nz, p)
    o.setup_static()
    ones = np.ones((nz, ny, nx))
    if kind == "rdot":
        a = 0.03
        o["r"][:] = (1.0 + 2.0 * z)[:, None, None] * ones
        o["w"][:] = (a * z * (1.0 - z))[:, None, None] * ones
        ref = -(2.0 * a * z * (1.0 - z) + (1.0 + 2.0 * z) * a * (1.0 - 2.0 * z))
    elif kind == "tdot":
        b = 2.0
        o["r"][:] = 1.0
        o["t"][:] = (1.0 + 10.0 * z + b * z ** 3)[:, None, None] * ones
        tz, tzz = 10.0 + 3 * b * z ** 2, 6 * b * z
        tzz[[0, -1]] = 0.0
        ref = o.scalar("gck") * (o["kappa"] * tzz + o["dkappa"] * tz)
    else:
        c = np.array([0.1, -0.4, 0.7, 0.5, -0.9])
        o["r"][:] = 1.0
        o["t"][:] = 1.0
        o["u"][:] = sum(c[i] * z ** i for i in range(5))[:, None, None] * ones
        ref = o.scalar("sck") * sum(i * (i - 1) * c[i] * z ** (i - 2) for i in range(2, 5))
    return o, p, ref


@pytest.mark.parametrize("kind", ["rdot", "tdot", "udot"])
def test_hydro_deriv_wall_rows_closed_forms(kind):
    """The oracle's pinned closed forms (flag 1 of div u on the walls, flag 0
    of T_zz, flag 3 of u_zz, one-sided upwind rows) on the GPU."""
    import fug30v0kxc_amd as fa
    print(f"[walls] {kind}", flush=True)
    o, p, ref = _closed_form_state(kind)
    s = fa.Solver(o.nx, o.ny, o.nz, p)
    try:
        s.set_state(*[o[n] for n in NAMES])
        s.hydro_deriv()
        got = s.field(kind, 3)
        assert np.abs(got - ref[:, None, None]).max() < 1e-10 * np.abs(ref).max()
        wd = s.field("wdot", 3)
        assert np.abs(wd[[0, -1]]).max() <= 1e-14 * max(np.abs(wd).max(), 1.0)
    finally:
        s.close()
