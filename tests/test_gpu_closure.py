"""GPU parity of the learned closure against the oracle (closure_oracle.c):
subgrid targets, the bf16 matrix-core forward pass, the gradient of the
training kernel, training itself, and the closure's coupling to the step."""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

F = po.CLOSURE_FIELDS


def _fine(shape, seed=0):
    rng = np.random.default_rng(seed)
    return {n: rng.standard_normal(shape) * (0.3 if n not in ("r", "t") else 0.1) + (1.5 if n in ("r", "t") else 0.0)
            for n in F}


def _f32(d):
    """values representable in float32 (the data set stores float32)"""
    return {k: np.asarray(v, np.float32).astype(np.float64) for k, v in d.items()}


def _params(seed, scale=0.3):
    rng = np.random.default_rng(seed)
    return (rng.standard_normal(po.N_PARAM) * scale).astype(np.float32)


# f a power of two: the lanes of a coarse cell split its fine columns (rows of
# 8 to 256 fine points, so a wave spans one or several coarse rows); f = 3: one
# thread per coarse cell
@pytest.mark.parametrize("f,shape", [(2, (8, 12, 16)), (4, (8, 16, 8)), (8, (16, 16, 64)),
                                     (4, (8, 8, 256)), (3, (9, 12, 24))])
def test_targets_match_oracle(f, shape):
    import fug30v0kxc_amd as fa
    print(f"[closure] targets f={f} {shape}", flush=True)
    fine = _fine(shape, seed=f)
    cg, tg = fa.closure_targets(fine, f, 0.7)
    co, to = po.closure_targets(fine, f, 0.7)
    for n in F:
        assert np.abs(cg[n] - co[n]).max() <= 1e-14 * max(1.0, np.abs(co[n]).max()), n
    for o in range(12):
        assert np.abs(tg[o] - to[o]).max() <= 1e-13 * max(1e-3, np.abs(to[o]).max()), o


def _closure_with_data(nsnap=2, shape=(8, 8, 12), seed=1):
    import fug30v0kxc_amd as fa
    cl = fa.Closure(seed=seed)
    snaps = []
    for s in range(nsnap):
        fine = _fine((2 * shape[0], 2 * shape[1], 2 * shape[2]), seed=10 + s)
        c, t = po.closure_targets(fine, 2, 0.5)
        c = _f32(c)
        t = np.asarray(t, np.float32).astype(np.float64)
        cl.add(c, t)
        snaps.append((c, t))
    return cl, snaps


# (6, 10, 12): neighbourhoods gathered per cell; (3, 5, 64) and (4, 3, 32):
# rows a multiple of 32 cells long, neighbourhoods staged per x-row segment
# (x wrap at both row ends, a partly empty last tile in the first)
@pytest.mark.parametrize("shape", [(6, 10, 12), (3, 5, 64), (4, 3, 32)])
def test_forward_matches_oracle(shape):
    print(f"[closure] forward {shape}", flush=True)
    cl, snaps = _closure_with_data()
    p = _params(3)
    cl.params = p
    norm = cl.norm
    coarse = _fine(shape, seed=4)
    yg = cl.predict(coarse)
    yo = po.mlp_forward(coarse, p.astype(np.float64), norm, round_bf16=True)
    tsd = norm[28:40, None, None, None]
    err = np.abs(yg - yo) / tsd                       # in scaled units
    scale = np.abs((yo - norm[16:28, None, None, None]) / tsd).max()
    # bf16 activations: an FP32-accumulated pre-activation can land on the
    # other side of a bf16 rounding boundary than the FP64 one, rarely; such a
    # flip moves all 12 outputs of its cell, so the rare cases are counted per cell
    cell = err.max(axis=0) / scale
    assert np.mean(cell > 1e-4) <= 0.01, np.argwhere(cell > 1e-4)[:10]
    assert cell.max() <= 2e-2
    yn = po.mlp_forward(coarse, p.astype(np.float64), norm, round_bf16=False)
    # the rounding points matter: the unrounded network differs far more
    assert np.abs(yn - yo).max() > 10 * np.quantile(np.abs(yg - yo), 0.5)


@pytest.mark.parametrize("ncell", [32, 77])
def test_gradient_matches_oracle(ncell):
    print(f"[closure] gradient ncell={ncell}", flush=True)
    cl, snaps = _closure_with_data()
    p = _params(5, 0.2)
    cl.params = p
    norm = cl.norm
    n3 = snaps[0][0]["r"].size
    rng = np.random.default_rng(ncell)
    cells = rng.integers(0, 2 * n3, size=ncell)
    Lg, gg = cl.grad(cells)
    Lo, go = po.mlp_grad(snaps, p.astype(np.float64), norm, cells, round_bf16=True)
    assert abs(Lg - Lo) <= 1e-4 * Lo
    offs = [0, 13824, 13888, 17984, 18048, 18816, 18828]
    for a, b in zip(offs[:-1], offs[1:]):
        d = np.linalg.norm(gg[a:b] - go[a:b]) / np.linalg.norm(go[a:b])
        assert d < 1e-2, (a, d)
    # and the gradient is a descent direction for the oracle's loss
    L2, _ = po.mlp_grad(snaps, p.astype(np.float64) - 1e-3 * gg.astype(np.float64) / np.abs(gg).max(), norm, cells,
                        round_bf16=False)
    L1, _ = po.mlp_grad(snaps, p.astype(np.float64), norm, cells, round_bf16=False)
    assert L2 < L1


def test_training_fits_a_teacher_network():
    """Targets made by a fixed random network are learnable exactly; Adam on
    the GPU must bring the data-set loss well below its starting value."""
    import fug30v0kxc_amd as fa
    print("[closure] training", flush=True)
    shape = (12, 16, 16)
    teacher = _params(21, 0.15).astype(np.float64)
    cl = fa.Closure(seed=2)
    rng = np.random.default_rng(22)
    norm = np.concatenate([np.zeros(8), np.ones(8), np.zeros(12), np.ones(12)])
    for s in range(3):
        coarse = _f32({n: rng.standard_normal(shape) for n in F})
        t = po.mlp_forward(coarse, teacher, norm, round_bf16=False)
        cl.add(coarse, np.asarray(t, np.float32).astype(np.float64))
    cl.norm = norm
    L0 = cl.eval()
    cl.train(400, batch=2048, lr=2e-3)
    L1 = cl.eval()
    print(f"[closure] loss {L0:.4g} -> {L1:.4g}", flush=True)
    assert L1 < 0.2 * L0


def _run_pair(precision, params, norm, steps=4, box=(16, 16, 12)):
    import fug30v0kxc_amd as fa
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = fa.Solver(*box, p, precision=precision)
    s.static()
    o = po.Oracle(*box, p)
    o.setup_static()
    cl = fa.Closure(seed=0)
    cl.params = params
    cl.norm = norm
    s.attach_closure(cl)
    o.attach_closure(params.astype(np.float64), norm)
    return s, o, cl


def _flux_norm():
    # scaled so that the fluxes are a visible but small correction
    return np.concatenate([np.zeros(8), np.ones(8), np.zeros(12), np.full(12, 1e-3)])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-10), ("mixed", 2e-5)])
def test_closure_rhs_matches_oracle(prec, tol):
    """The coupling on its own: the oracle is given the fluxes the GPU network
    produced for the same state and adds their divergences and eps itself."""
    import fug30v0kxc_amd as fa
    print(f"[closure] coupled RHS {prec}", flush=True)
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    s = fa.Solver(16, 16, 12, p, precision=fa.MHD_PREC_FP64 if prec == "fp64" else fa.MHD_PREC_MIXED)
    s.static()
    o = po.Oracle(16, 16, 12, p)
    o.setup_static()
    s.step(3, 1e-4, 0)
    o.step(3, 1e-4, 0)
    cl = fa.Closure(seed=0)
    cl.params = _params(31, 0.3)
    cl.norm = _flux_norm()
    s.attach_closure(cl)
    flux = cl.predict(s.get_state())
    o.closure_fix_fluxes(flux)
    s.hydro_deriv()
    o.hydro_deriv()
    s.mag_deriv()
    o.mag_deriv()
    for n in ("udot", "vdot", "wdot", "tdot", "rdot", "poldot", "tordot"):
        g, r = s.field(n, 3), o[n][2]
        e = np.abs(g - r).max() / np.abs(r).max()
        assert e < tol, (n, e)
    # the fluxes are not negligible: without them the RHS differ by far more
    o.closure_fix_fluxes(None)
    o.hydro_deriv()
    assert np.abs(s.field("udot", 3) - o["udot"][2]).max() > 1e3 * tol * np.abs(o["udot"][2]).max()


@pytest.mark.parametrize("prec,tol", [("fp64", 5e-4), ("mixed", 1e-3)])
def test_closure_steps_match_oracle(prec, tol):
    import fug30v0kxc_amd as fa
    print(f"[closure] coupled steps {prec}", flush=True)
    params = _params(31, 0.3)
    s, o, cl = _run_pair(fa.MHD_PREC_FP64 if prec == "fp64" else fa.MHD_PREC_MIXED, params, _flux_norm())
    s.step(12, 1e-4, 0)
    o.step(12, 1e-4, 0)
    st = s.get_state()
    # end to end the network is evaluated on each side: FP32 against FP64
    # pre-activations round a few bf16 activations differently per step
    # (B components against the largest |B| component, as the field is mostly vertical)
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    errs = {n: float(np.abs(st[n] - o[n]).max() / (bscale if n[0] == "b" else np.abs(o[n]).max()))
            for n in ("r", "t", "u", "v", "w", "bx", "by", "bz")}
    print(f"[closure] {prec} errors after 12 steps {errs}", flush=True)
    for n, e in errs.items():
        assert e < tol, (n, e)
    # the closure changed the run: compare with the plain oracle
    o2 = po.Oracle(16, 16, 12, reference_params(t_amp=1e-3, b_amp=1e-3))
    o2.setup_static()
    o2.step(12, 1e-4, 0)
    assert np.abs(o["u"] - o2["u"]).max() > 100 * np.abs(st["u"] - o["u"]).max()


def test_constant_fluxes_leave_the_run_unchanged():
    """W3 = b3 = 0: every flux is its constant mean, whose divergence and curl
    vanish, so the coupled run equals the plain one to rounding."""
    import fug30v0kxc_amd as fa
    print("[closure] constant fluxes", flush=True)
    params = _params(41, 0.3)
    params[18048:] = 0.0
    norm = np.concatenate([np.zeros(8), np.ones(8), np.linspace(-0.2, 0.3, 12), np.ones(12)])
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    a = fa.Solver(16, 16, 12, p)
    a.static()
    b = fa.Solver(16, 16, 12, p)
    b.static()
    cl = fa.Closure(seed=0)
    cl.params = params
    cl.norm = norm
    b.attach_closure(cl)
    a.step(6, 1e-4, 0)
    b.step(6, 1e-4, 0)
    sa, sb = a.get_state(), b.get_state()
    for n in sa:
        assert np.abs(sa[n] - sb[n]).max() <= 1e-10 * max(np.abs(sa[n]).max(), 1e-30), n
    b.attach_closure(None)
    b.step(1, 1e-4, 0)


def test_closure_gain_on_fine_run_data():
    """Trained on box-filtered snapshots of a fine run of this solver, the
    network explains much of the subgrid fluxes of a later, held-out snapshot,
    which the plain coarse method (no subgrid term) leaves entirely out; the
    closure-coupled coarse run from the filtered state stays finite and near
    the filtered fine spectra."""
    import fug30v0kxc_amd as fa
    from fug30v0kxc_amd import closure_eval as ce
    print("[closure] fine-run data", flush=True)
    p = reference_params(ra=1e6, q=500.0, t_amp=0.1, b_amp=0.01)
    fine = fa.Solver(64, 64, 48, p)
    fine.static()
    fine.step(1000, 0.0, 0)
    dtf = fine.scalars()["dt"]
    snaps = []
    for _ in range(5):
        fine.step(200, dtf, 0)
        snaps.append(fa.closure_targets(fine.get_state(), 2, p["f"]))
    cl = fa.Closure(seed=3)
    for c, t in snaps[:4]:
        cl.add(c, t)
    L0 = cl.eval()
    cl.train(3000, batch=2048, lr=1e-3)
    L1 = cl.eval()
    r2 = ce.explained_variance(cl.predict(snaps[4][0]), snaps[4][1])
    print(f"[closure] loss {L0:.3f} -> {L1:.3f}; held-out explained variance {np.round(r2, 3)}", flush=True)
    assert L1 < 0.7 * L0
    assert r2.mean() > 0.1 and (r2 > 0).sum() >= 8
    fine.step(100, dtf, 0)
    ref = fa.closure_targets(fine.get_state(), 2, p["f"])[0]
    errs = {}
    for mode in ("plain", "closure"):
        co = fa.Solver(32, 32, 24, p)
        co.static()
        co.set_state(*[snaps[4][0][n] for n in fa.CLOSURE_FIELDS])
        if mode == "closure":
            co.attach_closure(cl)
        co.step(50, 2 * dtf, 0)
        st = co.get_state()
        assert all(np.isfinite(v).all() for v in st.values())
        errs[mode] = ce.spectral_error(st, ref)
    print(f"[closure] spectral error against the filtered fine run {errs}", flush=True)
    assert errs["closure"] < 0.25 and abs(errs["closure"] - errs["plain"]) < 0.1
    assert ce.choose(errs["closure"]) == ("closure" if errs["closure"] <= 0.1 else "plain")


def test_closure_refusals():
    """The closure needs the unfused update and the EMF's extra term: it is
    refused with eta(z) on and inside an ensemble; bad arguments are errors."""
    import fug30v0kxc_amd as fa
    print("[closure] refusals", flush=True)
    cl = fa.Closure(seed=0)
    p = dict(reference_params(t_amp=1e-3), eta_c=10.0, eta_z0=0.5, eta_w=0.05)
    s = fa.Solver(16, 16, 12, p)
    s.static()
    with pytest.raises(fa.MhdError) as e:
        s.attach_closure(cl)
    assert e.value.code == -10
    q = fa.Solver(16, 16, 12, reference_params(t_amp=1e-3))
    q.static()
    q.attach_closure(cl)
    with pytest.raises(fa.MhdError):
        fa.Ensemble([q], 1e-4)
    q.attach_closure(None)
    with fa.Ensemble([q], 1e-4) as ens:
        ens.step(ens.block)
    with pytest.raises(fa.MhdError):
        cl.train(10, batch=100)          # batch not a multiple of 32
    with pytest.raises(fa.MhdError):
        cl.train(10, batch=64)           # no data yet
    with pytest.raises(fa.MhdError):
        cl.params = np.zeros(10, np.float32)
    bad = np.concatenate([np.zeros(8), np.zeros(8), np.zeros(12), np.ones(12)])
    with pytest.raises(fa.MhdError):
        cl.norm = bad                    # zero deviations
