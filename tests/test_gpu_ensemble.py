"""Ensembles: boxes of different sizes, parameters and precisions stepped by
replays of one captured hipGraph must reproduce, bit for bit, the same boxes
stepped one call at a time.  The graph is replayed several times with eager
work (state downloads, reductions) and synchronisations between replays, and
every replay is checked."""
import numpy as np
import pytest

from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
DT = 1e-4


def boxes():
    from fug30v0kxc_amd import MHD_PREC_FP64, MHD_PREC_MIXED, Solver
    specs = [((16, 16, 12), reference_params(ra=1e5, q=100.0, t_amp=1e-3, b_amp=1e-3), MHD_PREC_FP64),
             ((32, 24, 16), dict(reference_params(ra=3e5, q=30.0, t_amp=1e-3, b_amp=1e-3),
                                 eta_c=10.0, eta_z0=0.5, eta_w=0.05), MHD_PREC_FP64),
             ((24, 20, 12), reference_params(ra=1e4, q=500.0, t_amp=1e-3, b_amp=1e-3), MHD_PREC_MIXED)]
    out = []
    for shape, p, prec in specs:
        s = Solver(*shape, p, precision=prec)
        s.static()
        out.append(s)
    return out


def test_ensemble_replays_match_single_box_steps():
    from fug30v0kxc_amd import Ensemble
    print("[ensemble] 3 boxes", flush=True)
    ens_boxes, ref_boxes = boxes(), boxes()
    with Ensemble(ens_boxes, DT) as ens:
        assert ens.block == 30 and ens.nodes > 3 * 30 * 10
        for b in ref_boxes:
            b.step(2, DT, 0)                       # the ensemble's start-up steps
        for replay in range(3):
            ens.step(30)
            for b in ref_boxes:
                b.step(30, DT, 0)
            for i, (e, r) in enumerate(zip(ens_boxes, ref_boxes)):
                se, sr = e.get_state(), r.get_state()
                for n in NAMES:
                    assert np.array_equal(se[n], sr[n]), (replay, i, n)
                assert e.scalars()["nit"] == r.scalars()["nit"]
                assert e.scalars()["ttime"] == r.scalars()["ttime"]
                e.sminmax("t")                     # eager work between replays
    for b in ens_boxes + ref_boxes:
        b.close()


def test_ensemble_rejects_bad_requests():
    from fug30v0kxc_amd import Ensemble, MhdError
    bs = boxes()[:2]
    with Ensemble(bs, DT) as ens:
        with pytest.raises(MhdError):
            ens.step(31)
    with pytest.raises(MhdError):
        Ensemble([bs[0], bs[0]], DT)
    for b in bs:
        b.close()
