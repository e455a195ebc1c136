"""Pins the oracle's z finite differences (src/dbydz.F90) by polynomial
exactness of every row, for every flag, and the upwind branch choice."""
import numpy as np
import pytest

from oracle.pyoracle import Oracle, OracleError
from tests.conftest import reference_params

NZ = 11


@pytest.fixture(scope="module")
def o():
    return Oracle(4, 4, NZ, reference_params())


def poly(deg, shift=0.3):
    z = np.arange(NZ) / (NZ - 1.0)
    f = (z + shift) ** deg
    d1 = deg * (z + shift) ** (deg - 1) if deg >= 1 else 0 * z
    d2 = deg * (deg - 1) * (z + shift) ** (deg - 2) if deg >= 2 else 0 * z
    return f, d1, d2


def as3d(p):
    return np.ascontiguousarray(np.broadcast_to(p[:, None, None], (NZ, 4, 4)))


def col(a):
    return a[:, 1, 2]


@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_d1_flag5_exact_cubic(o, deg):
    f, d1, _ = poly(deg)
    assert np.abs(col(o.d1bydz(as3d(f), 5)) - d1).max() < 1e-10
    assert np.abs(o.d1bydz1D(f, 5) - d1).max() < 1e-10


def test_d1_interior_exact_quartic(o):
    f, d1, _ = poly(4)
    got = col(o.d1bydz(as3d(f), 5))
    assert np.abs(got[2:-2] - d1[2:-2]).max() < 1e-10
    assert np.abs(got[[0, 1, -2, -1]] - d1[[0, 1, -2, -1]]).max() > 1e-6  # rows are 3rd order


@pytest.mark.parametrize("flag", [0, 2])
def test_d1_zero_flags_zero_the_end_rows(o, flag):
    f, d1, _ = poly(3)
    got = col(o.d1bydz(as3d(f), flag))
    assert got[0] == 0.0 and got[-1] == 0.0
    assert np.abs(got[1:-1] - d1[1:-1]).max() < 1e-10


def test_d1_flag1_exact_for_cubics_vanishing_at_walls(o):
    z = np.arange(NZ) / (NZ - 1.0)
    for f, d in ((z * (1 - z), 1 - 2 * z), (z * z * (1 - z), 2 * z - 3 * z * z)):
        got = col(o.d1bydz(as3d(f), 1))
        assert np.abs(got - d).max() < 1e-10


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4, 5])
def test_d2_flag3_exact_quintic(o, deg):
    f, _, d2 = poly(deg)
    assert np.abs(col(o.d2bydz(as3d(f), 3)) - d2).max() < 1e-8
    assert np.abs(o.d2bydz1D(f, 3) - d2).max() < 1e-8


def test_d2_flag0(o):
    f, _, d2 = poly(5)
    got = col(o.d2bydz(as3d(f), 0))
    assert got[0] == 0.0 and got[-1] == 0.0
    assert np.abs(got[1:-1] - d2[1:-1]).max() < 1e-8


@pytest.mark.parametrize("sign", [1.0, -1.0, 0.0])
def test_upwind_branches_exact(o, sign):
    f, d1, _ = poly(3)
    w = as3d(np.full(NZ, sign))
    got = col(o.d1upbydz(as3d(f), 5, w))
    assert np.abs(got - d1).max() < 1e-10
    f4, d14, _ = poly(4)
    got4 = col(o.d1upbydz(as3d(f4), 5, w))
    assert np.abs(got4[3:-3] - d14[3:-3]).max() < 1e-9   # 4th-order interior


def test_upwind_branch_is_selected_by_sign_of_w(o):
    # f = z^6: the two one-sided interior rows have equal error on odd powers,
    # so an even power beyond their order separates them
    f, _, _ = poly(6)
    wp = col(o.d1upbydz(as3d(f), 5, as3d(np.ones(NZ))))
    wm = col(o.d1upbydz(as3d(f), 5, as3d(-np.ones(NZ))))
    w0 = col(o.d1upbydz(as3d(f), 5, as3d(np.zeros(NZ))))
    assert np.abs(wp[3:-3] - wm[3:-3]).min() > 1e-8
    assert np.array_equal(w0, wp)           # w = 0 takes the w >= 0 branch


@pytest.mark.parametrize("flag", [3, 4, -1])
def test_unknown_d1_flag_is_error_7(o, flag):
    with pytest.raises(OracleError) as e:
        o.d1bydz(np.zeros((NZ, 4, 4)), flag)
    assert e.value.code == 7


def test_unknown_d2_flag_is_error_7(o):
    with pytest.raises(OracleError) as e:
        o.d2bydz(np.zeros((NZ, 4, 4)), 5)
    assert e.value.code == 7


def test_create_errors():
    p = reference_params()
    for dims, code in (((15, 16, 8), -1), ((16, 15, 8), -2), ((16, 16, 5), 1)):
        with pytest.raises(OracleError) as e:
            Oracle(*dims, p)
        assert e.value.code == code
