"""GPU operator parity: every horizontal and vertical operator of the HIP
library (through the C ABI) against the CPU oracle on the same inputs.
FP64: the transforms round differently from the oracle's, so agreement is
to a few ulp of the field's scale."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

SIZES = [(16, 12, 8), (64, 64, 48), (30, 30, 12), (96, 64, 10), (256, 256, 6), (512, 512, 6),
         # lengths without a register-blocked plan (run-time radices, prime
         # factors up to 101), including the reference's 4 x 4 x 8 smoke box
         (4, 4, 8), (2, 8, 6), (10, 14, 6), (36, 22, 7), (100, 50, 6), (202, 34, 6), (2048, 6, 6)]


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module", params=SIZES, ids=lambda s: "x".join(map(str, s)))
def pair(request):
    from fug30v0kxc_amd import Solver
    nx, ny, nz = request.param
    p = reference_params()
    print(f"[ops] box {nx}x{ny}x{nz}", flush=True)
    s = Solver(nx, ny, nz, p)
    o = Oracle(nx, ny, nz, p)
    yield s, o
    s.close()


def field(shape, seed):
    rng = np.random.RandomState(seed)
    return rng.randn(*shape) + 3.0


def test_horizontal_derivatives(pair):
    s, o = pair
    f = field(s.shape, 1)
    assert rel(s.d1bydx(f), o.d1bydx(f)) < 1e-12
    assert rel(s.d1bydy(f), o.d1bydy(f)) < 1e-12
    a1, a2 = s.d2bydx(f)
    b1, b2 = o.d2bydx(f)
    assert rel(a1, b1) < 1e-12 and rel(a2, b2) < 1e-12
    a1, a2 = s.d2bydy(f)
    b1, b2 = o.d2bydy(f)
    assert rel(a1, b1) < 1e-12 and rel(a2, b2) < 1e-12


def test_dealias_and_2d_transforms(pair):
    s, o = pair
    f = field(s.shape, 2)
    assert rel(s.dealias(f), o.dealias(f)) < 1e-13
    F = o.forwardfft(f)
    assert rel(s.forwardfft(f), F) < 1e-13
    assert rel(s.inversefft(F), o.inversefft(F)) < 1e-13


@pytest.mark.parametrize("flag", [0, 1, 2, 5])
def test_d1bydz(pair, flag):
    s, o = pair
    f = field(s.shape, 3)
    assert rel(s.d1bydz(f, flag), o.d1bydz(f, flag)) < 1e-13


@pytest.mark.parametrize("flag", [0, 3])
def test_d2bydz(pair, flag):
    s, o = pair
    f = field(s.shape, 4)
    assert rel(s.d2bydz(f, flag), o.d2bydz(f, flag)) < 1e-13


@pytest.mark.parametrize("flag", [0, 1, 2, 5])
def test_d1upbydz_uses_solver_w(pair, flag):
    s, o = pair
    w = np.random.RandomState(5).randn(*s.shape)
    w[0, 0, :4] = 0.0                       # w = 0 takes the w >= 0 branch
    s.set_field("w", w)
    f = field(s.shape, 6)
    assert rel(s.d1upbydz(f, flag), o.d1upbydz(f, flag, w)) < 1e-13


def test_1d_z_operators(pair):
    s, o = pair
    v = np.random.RandomState(7).randn(s.nz)
    for flag in (0, 1, 2, 5):
        assert rel(s.d1bydz1D(v, flag), o.d1bydz1D(v, flag)) < 1e-13
    for flag in (0, 3):
        assert rel(s.d2bydz1D(v, flag), o.d2bydz1D(v, flag)) < 1e-13


def test_unknown_flag_is_code_7(pair):
    from fug30v0kxc_amd import MhdError
    s, _ = pair
    with pytest.raises(MhdError) as e:
        s.d1bydz(np.zeros(s.shape), 3)
    assert e.value.code == 7
    with pytest.raises(MhdError) as e:
        s.d2bydz(np.zeros(s.shape), 5)
    assert e.value.code == 7


def test_create_error_codes():
    from fug30v0kxc_amd import MhdError, Solver
    p = reference_params()
    for dims, code in (((15, 16, 8), -1), ((16, 15, 8), -2), ((16, 16, 5), 1), ((2050, 16, 8), -20)):
        with pytest.raises(MhdError) as e:
            Solver(*dims, p)
        assert e.value.code == code


@pytest.mark.parametrize("variant", [0, 1])
def test_512_transforms_both_radix_plans(variant):
    """512-point lines: column groups narrower than a cache line (paired onto
    one XCD) and both radix plans (8x8x8 is the alternative)."""
    from fug30v0kxc_amd import Solver
    nx, ny, nz = 512, 512, 6
    print(f"[ops] 512 variant {variant}", flush=True)
    p = reference_params()
    s = Solver(nx, ny, nz, p)
    s.set_tune(variant_x=variant, variant_y=variant, mult_x=1, mult_y=2, zc=8)
    o = Oracle(nx, ny, nz, p)
    f = field(s.shape, 7)
    assert rel(s.d1bydy(f), o.d1bydy(f)) < 1e-12
    a1, a2 = s.d2bydy(f)
    b1, b2 = o.d2bydy(f)
    assert rel(a1, b1) < 1e-12 and rel(a2, b2) < 1e-12
    a1, a2 = s.d2bydx(f)
    b1, b2 = o.d2bydx(f)
    assert rel(a1, b1) < 1e-12 and rel(a2, b2) < 1e-12
    F = o.forwardfft(f)
    assert rel(s.forwardfft(f), F) < 1e-13
    assert rel(s.inversefft(F), o.inversefft(F)) < 1e-13
    s.close()
