"""Physics pins of the oracle's hydro_deriv and mag_deriv
(src/hydro_deriv.f90:5-505, src/magnetic.f90:285-487).

Each case is a field for which the compressible MHD equations give the time
derivative in closed form. The closed forms come from the physics, not from
the code's formulas:

* shear-wave viscous decay ν k²;
* the Stokes bulk term (4/3)ν k² of a compressional mode, with its
  compressive and viscous heating;
* the pressure gradient and thermal diffusion with κ(z);
* the magnetic tension B_z ∂_z B_x and the magnetic pressure of a
  force-free-violating field;
* Ohmic heating;
* unbalanced gravity;
* continuity for uniform advection;
* magnetic diffusion ζC_k∇² and uniform-flow advection −U∂_x of the
  poloidal/toroidal potentials;
* mean-field generation B_z ∂_z U by shear;
* a Coriolis force that does no work.

Fields are single Fourier modes in x/y times polynomials in z of a degree
the stencils differentiate exactly, so the oracle must agree to round-off.
"""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

NX, NY, NZ = 16, 12, 10
TOL = 1e-11


def box(**kw):
    p = reference_params(**kw)
    o = Oracle(NX, NY, NZ, p)
    x = np.arange(NX) * (p["xmax"] / NX)
    y = np.arange(NY) * (p["ymax"] / NY)
    z = np.arange(NZ) * (1.0 / (NZ - 1))
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    kx = 2 * np.pi * 2 / p["xmax"]        # mode 2 in x
    ky = 2 * np.pi * 3 / p["ymax"]        # mode 3 in y
    return o, p, X, Y, Z, kx, ky


def quiet(o, r=1.0, t=1.0):
    """ρ, T uniform, u = 0, B = ẑ, J² = 0."""
    for n in ("u", "v", "w", "bx", "by", "jsq"):
        o[n][:] = 0.0
    o["bz"][:] = 1.0
    o["r"][:] = r
    o["t"][:] = t


def rhs(o):
    o.hydro_deriv()
    return {n: o[n][2].copy() for n in ("rdot", "tdot", "udot", "vdot", "wdot")}


def close(a, b, scale=1.0):
    assert np.abs(a - b).max() <= TOL * max(scale, np.abs(b).max())


def grav(p):
    return p["theta"] * (p["cm"] + 1.0)


def test_uniform_state_feels_only_gravity():
    o, p, *_ = box()
    quiet(o)
    d = rhs(o)
    for n in ("rdot", "tdot", "udot", "vdot"):
        close(d[n], 0.0 * d[n], 1.0)
    close(d["wdot"][1:-1], np.full_like(d["wdot"][1:-1], grav(p)))
    # dw/dt = 0 is imposed on the top and bottom rows
    assert np.abs(d["wdot"][[0, -1]]).max() < 1e-13


def test_shear_wave_decays_at_viscous_rate_and_heats():
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    eps = 1e-2
    o["u"][:] = eps * np.sin(ky * Y)
    d = rhs(o)
    nu = p["sigma"] * p["ck"]
    close(d["udot"], -nu * ky**2 * o["u"], eps * nu * ky**2)
    close(d["vdot"], 0.0 * d["vdot"], eps)
    g1 = p["gamma"] - 1.0
    close(d["tdot"], nu * g1 * (eps * ky * np.cos(ky * Y)) ** 2, nu * (eps * ky) ** 2)


def test_compressional_mode_bulk_viscosity_and_heating():
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    eps = 1e-2
    u = eps * np.sin(kx * X)
    o["u"][:] = u
    d = rhs(o)
    nu = p["sigma"] * p["ck"]
    ux = eps * kx * np.cos(kx * X)
    want_u = -u * ux - (4.0 / 3.0) * nu * kx**2 * u
    close(d["udot"], want_u, eps * kx**2)
    g1 = p["gamma"] - 1.0
    want_t = -g1 * ux + (4.0 / 3.0) * nu * g1 * ux**2
    close(d["tdot"], want_t, eps * kx)
    close(d["rdot"], -ux, eps * kx)


def test_pressure_gradient_and_thermal_diffusion():
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    a = 1e-2
    o["t"][:] = 1.0 + a * np.cos(kx * X)
    d = rhs(o)
    close(d["udot"], a * kx * np.sin(kx * X), a * kx)
    kappa = o["kappa"][:, None, None]
    want = p["gamma"] * p["ck"] * kappa * (-a * kx**2 * np.cos(kx * X))
    close(d["tdot"], want, a * kx**2)


def test_vertical_conduction_with_variable_kappa():
    """ρT_t = γC_k (κ T_zz + κ' T_z) for T = 1 + a z² (interior rows)."""
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    a = 0.3
    o["t"][:] = 1.0 + a * Z**2
    d = rhs(o)
    kap = o["kappa"][:, None, None]
    dkap = o["dkappa"][:, None, None]
    want = p["gamma"] * p["ck"] * (kap * 2 * a + dkap * 2 * a * Z) / 1.0
    close(d["tdot"][1:-1], want[1:-1].repeat(1, 0) * np.ones_like(d["tdot"][1:-1]), 1.0)


def test_continuity_under_uniform_advection():
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    a, U = 1e-2, 0.3
    rho = 1.0 + a * np.cos(kx * X)
    o["r"][:] = rho
    o["u"][:] = U
    d = rhs(o)
    close(d["rdot"], U * a * kx * np.sin(kx * X), a * kx)
    close(d["udot"], a * kx * np.sin(kx * X) / rho, a * kx)


def test_magnetic_tension_and_pressure():
    """B = (b z², 0, 1): u_t = F B_z ∂_z B_x = 2 F b z; w_t adds
    -∂_z(F B_x²/2) = -2 F b² z³ to gravity (interior rows)."""
    o, p, X, Y, Z, kx, ky = box(b_amp=0.0)
    quiet(o)
    b = 0.05
    o["bx"][:] = b * Z**2
    d = rhs(o)
    F = p["f"]
    close(d["udot"], 2 * F * b * Z, F * b)
    want_w = grav(p) - 2 * F * b**2 * Z**3
    close(d["wdot"][2:-2], want_w[2:-2], grav(p))


def test_lorentz_force_of_sheared_field():
    """B = (0, b sin(k x), 1): J = ∂_x B_y ẑ, (J×B)_x = -F b² k sin cos."""
    o, p, X, Y, Z, kx, ky = box()
    quiet(o)
    b = 0.05
    o["by"][:] = b * np.sin(kx * X)
    d = rhs(o)
    F = p["f"]
    close(d["udot"], -F * b**2 * kx * np.sin(kx * X) * np.cos(kx * X), F * b**2 * kx)
    # (J×B)_y = J_z B_x - J_x B_z = 0 and (J×B)_z = J_x B_y - J_y B_x = 0
    close(d["vdot"], 0.0 * d["vdot"], F * b * kx)
    close(d["wdot"][1:-1], np.full_like(d["wdot"][1:-1], grav(p)), grav(p))


def test_ohmic_heating():
    o, p, X, Y, Z, kx, ky = box()
    quiet(o, r=2.0)
    j2 = 0.1 + 0.05 * np.cos(kx * X) ** 2
    o["jsq"][:] = j2
    d = rhs(o)
    want = p["f"] * p["zeta"] * p["ck"] * (p["gamma"] - 1.0) * j2 / 2.0
    close(d["tdot"], want, want.max())


def test_coriolis_force_does_no_work():
    o, p, X, Y, Z, kx, ky = box(tayl=1e4, psi=0.7)
    quiet(o)
    U = (0.3, -0.2, 0.1)
    for n, val in zip(("u", "v", "w"), U):
        o[n][:] = val
    d = rhs(o)
    a = np.array([d["udot"][3, 0, 0], d["vdot"][3, 0, 0], d["wdot"][3, 0, 0] - grav(p)])
    assert np.abs(a).max() > 1e-3             # the force is there ...
    assert abs(np.dot(a, U)) < 1e-13 * np.abs(a).max()   # ... and does no work


# ---- induction (src/magnetic.f90) ------------------------------------------

def mag_box(**kw):
    o, p, X, Y, Z, kx, ky = box(**kw)
    quiet(o)
    o["bxbar"][:] = 0.0
    o["bybar"][:] = 0.0
    o.set_scalar("bzbar", 1.0)
    return o, p, X, Y, Z, kx, ky


def test_uniform_flow_advects_and_diffuses_potentials():
    """pol = cos(kx x + ky y) z², tor = sin(kx x) z² advected by u = U x̂:
    pol_t = ζC_k ∇² pol − U ∂_x pol (and the same for tor)."""
    o, p, X, Y, Z, kx, ky = mag_box()
    U = 0.4
    ph = kx * X + ky * Y
    P = np.cos(ph) * Z**2
    T = np.sin(kx * X) * Z**2
    o["pol"][:] = o.forwardfft(P)
    o["tor"][:] = o.forwardfft(T)
    o.bandj()
    o["u"][:] = U
    o.mag_deriv()
    eta = p["zeta"] * p["ck"]
    want_p = eta * (-(kx**2 + ky**2) * P + 2 * np.cos(ph)) + U * kx * np.sin(ph) * Z**2
    want_t = eta * (-kx**2 * T + 2 * np.sin(kx * X)) - U * kx * np.cos(kx * X) * Z**2
    got_p = o.inversefft(o["poldot"][2])
    got_t = o.inversefft(o["tordot"][2])
    close(got_p, want_p, 1.0)
    close(got_t, want_t, 1.0)
    assert np.abs(o["bxbardot"][2]).max() < 1e-12
    assert np.abs(o["bybardot"][2]).max() < 1e-12


def test_shear_generates_mean_field_from_vertical_field():
    """u = U(z) x̂ with B = ẑ: ∂_t Bx̄ = B_z U'(z) (+ ζC_k Bx̄_zz = 0)."""
    o, p, X, Y, Z, kx, ky = mag_box()
    o["pol"][:] = 0.0
    o["tor"][:] = 0.0
    o.bandj()
    z = np.arange(NZ) * (1.0 / (NZ - 1))
    o["u"][:] = (z**2)[:, None, None]
    o["v"][:] = (0.5 * z**3)[:, None, None]
    o.mag_deriv()
    close(o["bxbardot"][2], 2 * z, 1.0)
    close(o["bybardot"][2], 1.5 * z**2, 1.0)


def test_mean_field_diffusion():
    o, p, X, Y, Z, kx, ky = mag_box()
    z = np.arange(NZ) * (1.0 / (NZ - 1))
    o["pol"][:] = 0.0
    o["tor"][:] = 0.0
    o["bxbar"][:] = 0.2 * z**2 + 0.1 * z**5
    o["bybar"][:] = -0.3 * z**4
    o.bandj()
    o.mag_deriv()
    eta = p["zeta"] * p["ck"]
    close(o["bxbardot"][2], eta * (0.4 + 2.0 * z**3), eta)
    close(o["bybardot"][2], eta * (-3.6 * z**2), eta)
