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


def test_ensemble_batches_each_class_into_one_launch_per_kernel():
    """Boxes with the same horizontal lengths and precision form one class
    whatever their nz, Ra, Q or eta(z): their launches are batched, so the
    graph holds one node per kernel of a class, not per box."""
    from fug30v0kxc_amd import MHD_PREC_MIXED, Ensemble, Solver
    print("[ensemble] classes", flush=True)
    specs = [((32, 32, 12), reference_params(ra=1e5, q=100.0, t_amp=1e-3, b_amp=1e-3)),
             ((32, 32, 20), dict(reference_params(ra=3e5, q=30.0, t_amp=1e-3, b_amp=1e-3),
                                 eta_c=20.0, eta_z0=0.5, eta_w=0.05)),
             ((32, 32, 16), reference_params(ra=1e4, q=500.0, t_amp=1e-3, b_amp=1e-3)),
             ((48, 32, 12), reference_params(ra=1e5, q=100.0, t_amp=1e-3, b_amp=1e-3))]

    def make():
        out = []
        for shape, p in specs:
            s = Solver(*shape, p, precision=MHD_PREC_MIXED)
            s.static()
            out.append(s)
        return out
    ens_boxes, ref_boxes = make(), make()
    with Ensemble(ens_boxes, DT) as ens:
        assert ens.classes == 2
        one = Ensemble([ref_boxes[0]], DT)
        per_class = one.nodes
        one.close()
        # two classes: about twice one box's nodes, not four times
        assert ens.nodes < 2 * per_class + 16
        ens.step(60)
    for b in ref_boxes:
        b.step(60 if b is ref_boxes[0] else 62, DT, 0)
    for i, (e, r) in enumerate(zip(ens_boxes, ref_boxes)):
        se, sr = e.get_state(), r.get_state()
        for n in NAMES:
            assert np.array_equal(se[n], sr[n]), (i, n)
    for b in ens_boxes + ref_boxes:
        b.close()
