"""Pins of the oracle's calcdt, nusselt, checkdivb, the wall rows of
hydro_deriv and the density floor of the AB3 update, each against a closed
form, a scaling law or an invariant that a plausible slip (a wrong limit,
constant, flag or history level) breaks."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params


def _zgrid(nz):
    return np.arange(nz) / (nz - 1.0)          # src/initial.F90:41, z = (k-1) dz


def _quiet(nx, ny, nz, override=None, **kw):
    """Static polytrope without noise, u = v = w = 0, B = z-hat; `override`
    replaces converted parameters (e.g. sigma at a fixed C_k)."""
    p = reference_params(t_amp=0.0, **kw)
    p.update(override or {})
    o = Oracle(nx, ny, nz, p)
    o.static()
    return o


# ---------------------------------------------------------------------------
# calcdt, src/utils.F90:144-244 with the coefficients of src/constants.f90:104-107
# ---------------------------------------------------------------------------
def _dt(o, **fields):
    for n, v in fields.items():
        o[n][:] = v
    return o.calcdt()


def test_calcdt_scales_with_the_safety_factor():
    a = _quiet(16, 16, 12, sf=0.2)
    b = _quiet(16, 16, 12, sf=0.4)
    for o in (a, b):
        o["u"][:] = 3.0
    assert b.calcdt() == 2.0 * a.calcdt()


@pytest.mark.parametrize("xmax,ymax,nx,ny", [(4.0, 4.0, 16, 16), (4.0, 2.0, 16, 16), (2.0, 4.0, 8, 16)])
def test_calcdt_horizontal_cfl(xmax, ymax, nx, ny):
    # |u_H| dominates: 1/dt = (|u_H| + c)/(0.16 sf dd) is affine in |u_H| with
    # slope 1/(0.16 sf dd), dd = min(dx, dy); only the horizontal speed counts
    def inv(u, v, nz=12):
        o = _quiet(nx, ny, nz, xmax=xmax, ymax=ymax)
        return 1.0 / _dt(o, u=u, v=v)
    dd = min(xmax / nx, ymax / ny)
    slope = (inv(300.0, 400.0) - inv(150.0, 200.0)) / (500.0 - 250.0)
    assert abs(slope * 0.16 * 0.2 * dd - 1.0) < 1e-10
    # the sign and split of (u, v) do not matter, nor nz
    assert abs(inv(-400.0, 300.0) / inv(300.0, -400.0) - 1.0) < 1e-14
    assert abs(inv(500.0, 0.0) / inv(300.0, 400.0) - 1.0) < 1e-14
    assert abs(inv(300.0, 400.0, nz=23) / inv(300.0, 400.0) - 1.0) < 1e-14


@pytest.mark.parametrize("nz", [12, 23])
def test_calcdt_vertical_cfl(nz):
    # |w| dominates: slope 1/(0.3 sf dz), independent of the horizontal grid
    def inv(w, nx=16):
        o = _quiet(nx, 16, nz)
        return 1.0 / _dt(o, w=w)
    dz = 1.0 / (nz - 1)
    slope = (inv(4000.0) - inv(2000.0)) / 2000.0
    assert abs(slope * 0.3 * 0.2 * dz - 1.0) < 1e-10
    assert abs(inv(-4000.0) / inv(4000.0) - 1.0) < 1e-14
    assert abs(inv(4000.0, nx=32) / inv(4000.0) - 1.0) < 1e-14


def _diffusive(nx, nz, sigma=1.0, rho_min=1e-4, cell=(3, 2, 5)):
    # F = 0 (Q = 0) so the Alfven speed drops out; one low-density cell; dd is
    # dx = 4/nx for nx >= 16 (ny = 16); sigma changes at a fixed C_k
    o = _quiet(nx, 16, nz, override={"sigma": sigma}, q=0.0)
    o["r"][cell] = rho_min
    return o.calcdt()


def test_calcdt_diffusive_limits_scale_with_density_and_grid():
    # dt = sf dtdiff0 rho_min: proportional to the density minimum
    a, b = _diffusive(16, 12, rho_min=1e-4), _diffusive(16, 12, rho_min=2e-4)
    assert abs(b / a - 2.0) < 1e-12
    # vertical diffusion binds at dd = 1/4, dz = 1/11: dt proportional to dz^2
    c = _diffusive(16, 23)
    assert abs(c / a - (11.0 / 22.0) ** 2) < 1e-12
    # a fine horizontal grid binds the horizontal limit: dt proportional to dd^2
    d, e = _diffusive(128, 12), _diffusive(256, 12)
    assert abs(e / d - 0.25) < 1e-12
    # thermal (1/gamma) against viscous (0.75/sigma): viscous binds for
    # sigma > 0.75 gamma, and then dt is proportional to 1/sigma
    f2, f4 = _diffusive(256, 12, sigma=2.0), _diffusive(256, 12, sigma=4.0)
    assert abs(f2 / f4 - 2.0) < 1e-12
    assert f2 < d


def test_calcdt_takes_the_global_extrema():
    # one fast cell anywhere sets the CFL limit; one slow cell does not
    # (uniform T and rho, so the sound and Alfven speeds are the same everywhere)
    a, b = _quiet(16, 16, 12), _quiet(16, 16, 12)
    for o in (a, b):
        o["t"][:] = 2.0
        o["r"][:] = 1.5
    a["u"][7, 3, 11] = 800.0
    b["u"][:] = 800.0
    b["u"][7, 3, 11] = 0.0
    assert a.calcdt() == b.calcdt()


# ---------------------------------------------------------------------------
# nusselt, src/dbydz.F90:4874-4925 (nusstmp, src/constants.f90:94)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("gamma,cm", [(5.0 / 3.0, 1.0), (1.4, 1.5), (5.0 / 3.0, 0.5)])
def test_nusselt_cubic_profile_pins_nusstmp(gamma, cm):
    # T = 1 + theta z + a z^3 (+ x-dependent parts with zero plane mean): the
    # one-sided 4-point bottom row is exact for cubics, so <T_z>(1) = theta + 3a
    # and Nu - 1 = 3a / (theta - nusstmp); nusstmp is the adiabatic gradient
    # theta (m + 1)(1 - 1/gamma) of the polytrope
    nx, ny, nz, theta = 16, 8, 13, 10.0
    z = _zgrid(nz)[:, None, None]
    x = np.arange(nx)[None, None, :] * 4.0 / nx
    nus = []
    for a in (0.5, -1.25):
        o = _quiet(nx, ny, nz, gamma=gamma, cm=cm)
        o["t"][:] = 1.0 + theta * z + (a + 0.3 * np.cos(2 * np.pi * x / 4.0)) * z ** 3
        nus.append((a, o.nusselt()))
    for a, nu in nus:
        nusstmp = theta - 3.0 * a / (nu - 1.0)
        assert abs(nusstmp - theta * (cm + 1.0) * (1.0 - 1.0 / gamma)) < 1e-9


# ---------------------------------------------------------------------------
# checkdivb, src/utils.F90:249-318
# ---------------------------------------------------------------------------
def test_checkdivb_known_divergence():
    # B = (A sin kx, C sin ly, c1 z + c2 z^2): div B = A k cos kx + C l cos ly
    # + c1 + 2 c2 z; spectral d/dx, d/dy are exact for these modes and the
    # flag-5 z rows for quadratics, so max and min are closed forms
    nx, ny, nz = 16, 12, 10
    o = _quiet(nx, ny, nz)
    L = 4.0
    x = (np.arange(nx) * L / nx)[None, None, :]
    y = (np.arange(ny) * L / ny)[None, :, None]
    z = _zgrid(nz)[:, None, None]
    A, C, c1, c2 = 0.7, -0.2, 0.05, 0.3
    k, l_ = 2 * np.pi * 2 / L, 2 * np.pi * 1 / L
    o["bx"][:] = A * np.sin(k * x) + 0 * y + 0 * z
    o["by"][:] = C * np.sin(l_ * y) + 0 * x + 0 * z
    o["bz"][:] = c1 * z + c2 * z * z + 0 * x + 0 * y
    mx, mn = o.checkdivb()
    amp = abs(A) * k + abs(C) * l_
    zterm = c1 + 2 * c2 * _zgrid(nz)
    assert abs(mx - (amp + zterm.max())) < 1e-12
    assert abs(mn - (-amp + zterm.min())) < 1e-12


# ---------------------------------------------------------------------------
# hydro_deriv wall rows, src/hydro_deriv.f90:443-482 (and the rows of the
# derivatives that enter there, src/dbydz.F90:927-998 and its twins)
# ---------------------------------------------------------------------------
def test_wdot_vanishes_on_the_walls():
    # d(rho)/dz on the walls is replaced by wdot/T, so the wall wdot is zero
    # for any state
    nx, ny, nz = 16, 12, 10
    o = Oracle(nx, ny, nz, reference_params(b_amp=1e-3))
    o.setup_static()
    rng = np.random.RandomState(7)
    for n in ("u", "v", "w"):
        o[n][:] = 1e-2 * rng.randn(nz, ny, nx)
    o.hydro_deriv()
    wd = o["wdot"][2]
    assert np.abs(wd[1:-1]).max() > 0.05
    assert np.abs(wd[[0, -1]]).max() < 1e-14 * np.abs(wd[1:-1]).max()


def test_rdot_closed_form_including_walls():
    # u = v = 0, rho = 1 + 2z, w = a z (1 - z) >= 0, horizontally uniform:
    # rdot = -(rho w)' at every level.  On the walls w = 0, so the wdot/T
    # substitution drops out and rdot = -rho w', which needs the flag-1 rows
    # of divu's dw/dz (flag 0 or 2 would give 0)
    nx, ny, nz, a = 8, 8, 14, 0.03
    o = _quiet(nx, ny, nz)
    z = _zgrid(nz)
    rho = 1.0 + 2.0 * z
    w = a * z * (1.0 - z)
    o["r"][:] = rho[:, None, None]
    o["w"][:] = w[:, None, None]
    o.hydro_deriv()
    ref = -(2.0 * w + rho * a * (1.0 - 2.0 * z))
    got = o["rdot"][2]
    assert np.abs(got - ref[:, None, None]).max() < 1e-13
    assert abs(got[0, 0, 0] + a) < 1e-14 and abs(got[-1, 0, 0] - 3.0 * a) < 1e-13


def test_tdot_closed_form_wall_rows_drop_t_zz():
    # u = w = 0, rho = 1, B = z-hat (J = 0), T = 1 + theta z + b z^3:
    # tdot = gck (kappa T_zz + kappa' T_z); T_zz takes flag 0 (zero on the
    # walls), T_z the upwind flag-5 rows (one-sided, exact for cubics)
    nx, ny, nz, b = 8, 8, 13, 2.0
    o = _quiet(nx, ny, nz)
    o.ini_trans_params()
    z = _zgrid(nz)
    o["r"][:] = 1.0
    o["t"][:] = (1.0 + 10.0 * z + b * z ** 3)[:, None, None]
    o.hydro_deriv()
    tz, tzz = 10.0 + 3 * b * z ** 2, 6 * b * z
    tzz[[0, -1]] = 0.0
    ref = o.scalar("gck") * (o["kappa"] * tzz + o["dkappa"] * tz)
    got = o["tdot"][2]
    assert np.abs(got - ref[:, None, None]).max() < 1e-11 * np.abs(ref).max()
    # and the walls really differ from the interior formula
    full = o.scalar("gck") * o["kappa"][-1] * 6 * b
    assert abs(full) > 1e3 * 1e-11 * np.abs(ref).max()


def test_udot_closed_form_flag3_wall_rows():
    # v = w = 0, rho = T = 1, B = z-hat, u = u(z) quartic: udot = sck u''
    # at every level; the walls use the one-sided six-point rows of flag 3
    nx, ny, nz = 8, 8, 12
    o = _quiet(nx, ny, nz)
    z = _zgrid(nz)
    c = np.array([0.1, -0.4, 0.7, 0.5, -0.9])
    u = sum(c[i] * z ** i for i in range(5))
    uzz = sum(i * (i - 1) * c[i] * z ** (i - 2) for i in range(2, 5))
    o["r"][:] = 1.0
    o["t"][:] = 1.0
    o["u"][:] = u[:, None, None]
    o.hydro_deriv()
    ref = o.scalar("sck") * uzz
    got = o["udot"][2]
    assert np.abs(got - ref[:, None, None]).max() < 1e-10 * np.abs(ref).max()
    assert np.abs(o["vdot"][2]).max() < 1e-14


# ---------------------------------------------------------------------------
# density floor of the AB3 update, src/main.F90:438-443
# ---------------------------------------------------------------------------
def test_density_floor_clamps_and_clears_history():
    nx, ny, nz = 16, 12, 10
    o = Oracle(nx, ny, nz, reference_params(b_amp=1e-3))
    o.setup_static()
    rng = np.random.RandomState(3)
    for n in ("u", "v", "w"):
        o[n][:] = 1e-2 * rng.randn(nz, ny, nx)
    o["w"][[0, -1]] = 0.0
    cells = [(4, 5, 6), (1, 0, 0), (0, 7, 3), (nz - 1, 2, 9)]
    for c in cells:
        o["r"][c] = 0.03
    # the next right-hand side of rho, computed on its own
    ref = Oracle(nx, ny, nz, reference_params(b_amp=1e-3))
    ref.setup_static()
    for n in ("r", "t", "u", "v", "w", "bx", "by", "bz", "jsq"):
        ref[n][:] = o[n]
    ref.hydro_deriv()
    rdot_new = ref["rdot"][2].copy()
    dt = 1e-5
    o.step(1, dt, 0)
    r, rd = o["r"], o["rdot"]
    floored = np.zeros((nz, ny, nx), bool)
    for c in cells:
        floored[c] = True
        assert r[c] == 0.05
        # after the shift 2->1, 3->2: both kept levels are zero at the cell
        assert rd[0][c] == 0.0 and rd[1][c] == 0.0
    # everywhere else: Euler update, untouched history (level 2 = new RHS)
    assert np.all(r[~floored] > 0.05)
    assert np.array_equal(rd[1][~floored], rdot_new[~floored])
    assert np.abs(rdot_new[floored]).max() > 0.0
