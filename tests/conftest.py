import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP C-ABI)")


def reference_params(ra=1e5, q=100.0, t_amp=1e-3, b_amp=0.0, gamma=5.0 / 3.0, theta=10.0,
                     cm=1.0, zeta=0.2, sigma=1.0, xmax=4.0, ymax=4.0, sf=0.2, tayl=0.0,
                     psi=1.570796327):
    """Parameter block of the BASELINE physics (INPUT flag 2: Ra, Q given)."""
    from oracle.pyoracle import ra_q_to_ck_f
    ck, f = ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q)
    return dict(gamma=gamma, sigma=sigma, theta=theta, ck=ck, cm=cm, f=f, tayl=tayl, psi=psi,
                zeta=zeta, xmax=xmax, ymax=ymax, sf=sf, t_amp=t_amp, b_amp=b_amp)


@pytest.fixture
def params():
    return reference_params()
