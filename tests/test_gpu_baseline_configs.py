"""Every benchmark configuration, and the reference's own evolved state,
through the GPU step against the FP64 oracle (C1 is in test_gpu_configs.py):

* C2, 256x256x160, Ra 1e6, Q 1000, FP64: 3 steps within 1e-10;
* C3's physics (eta contrast 100 at z = 0.3, Ra 1e6, Q 1e3) in the mixed
  mode at 512x512 planes (nz = 16): within 1e-5;
* C4, the seeded 32-box ensemble, mixed: graph replays bit-identical to
  stepping each box on its own, and one box of every (n, nz) class within
  1e-5 of the oracle;
* bin/restart.nc (30x30x60, gamma 1.4, Ra 1e5, Q 2500, iteration 105000, an
  O(1) developed state; tests/golden/restart_fields.npz): 20 steps with the
  reference's calcdt, FP64 within 1e-10 and mixed within 1e-5.
"""
import os

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

NAMES = ("r", "t", "u", "v", "w", "bx", "by", "bz")
HERE = os.path.dirname(os.path.abspath(__file__))


def errors(st, o):
    bscale = max(np.abs(o[n]).max() for n in ("bx", "by", "bz"))
    return {n: float(np.abs(st[n] - o[n]).max() /
                     (bscale if n in ("bx", "by", "bz") else max(np.abs(o[n]).max(), 1e-300))) for n in NAMES}


def check(st, o, tol, what):
    errs = errors(st, o)
    print(f"[baseline] {what}: " + " ".join(f"{n} {e:.1e}" for n, e in errs.items()), flush=True)
    for n, e in errs.items():
        assert e < tol, (what, n, e)


def test_c2_fp64_full_box():
    import fug30v0kxc_amd as fa
    print("[baseline] C2 256x256x160 fp64, 3 steps", flush=True)
    p = reference_params(ra=1e6, q=1000.0, t_amp=1e-3, b_amp=1e-3)
    s = fa.Solver(256, 256, 160, p)
    s.static()
    ds = s.step(3, 5e-5, 3)
    st = s.get_state()
    s.close()
    o = Oracle(256, 256, 160, p)
    o.setup_static()
    do = o.step(3, 5e-5, 3)
    check(st, o, 1e-10, "C2 fp64")
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= 1e-10 * abs(do[i]), key


def test_c3_physics_mixed_512_planes():
    import fug30v0kxc_amd as fa
    print("[baseline] C3 physics 512x512x16 mixed, eta x100 at z=0.3", flush=True)
    p = reference_params(ra=1e6, q=1000.0, t_amp=1e-3, b_amp=1e-3)
    p.update(eta_c=100.0, eta_z0=0.3, eta_w=0.05)
    nx, ny, nz, steps, dt = 512, 512, 16, 6, 5e-6
    s = fa.Solver(nx, ny, nz, p, precision=fa.MHD_PREC_MIXED)
    s.static()
    ds = s.step(steps, dt, steps)
    st = s.get_state()
    s.close()
    o = Oracle(nx, ny, nz, p)
    o.setup_static()
    do = o.step(steps, dt, steps)
    check(st, o, 1e-5, "C3 mixed")
    for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
        assert abs(ds[key] - do[i]) <= 1e-5 * abs(do[i]), key


def test_c4_ensemble_graph_and_oracle():
    import bench
    import fug30v0kxc_amd as fa
    specs = bench.c4_specs()
    dt, steps = bench.C4_DT, 60
    print(f"[baseline] C4 ensemble of {len(specs)} boxes, {steps} steps", flush=True)
    graph = []
    for n, nz, p in specs:
        s = fa.Solver(n, n, nz, p, precision=fa.MHD_PREC_MIXED)
        s.static()
        graph.append(s)
    ens = fa.Ensemble(graph, dt)          # up to two start-up steps per box
    ens.step(steps)
    ens.close()
    total = 2 + steps
    reps = {}
    try:
        for i, (s, (n, nz, p)) in enumerate(zip(graph, specs)):
            assert s.scalars()["nit"] == total
            one = fa.Solver(n, n, nz, p, precision=fa.MHD_PREC_MIXED)
            one.static()
            one.step(total, dt, 0)
            a, b = s.get_state(), one.get_state()
            one.close()
            for k in NAMES:
                assert np.array_equal(a[k], b[k]), (i, k)
            reps.setdefault((n, nz), (i, a))
        print(f"[baseline] C4 graph == per-box stepping for all boxes; classes {sorted(reps)}", flush=True)
        for (n, nz), (i, st) in sorted(reps.items()):
            o = Oracle(n, n, nz, specs[i][2])
            o.setup_static()
            o.step(total, dt, 0)
            check(st, o, 1e-5, f"C4 box {i} ({n}^2 x {nz})")
    finally:
        for s in graph:
            s.close()


def _restart_state():
    d = np.load(os.path.join(HERE, "golden", "restart_fields.npz"), allow_pickle=False)
    names = [str(x) for x in d["attr_names"]]
    p = dict(zip(names, (float(v) for v in d["attrs"])))
    # INPUT's SF (bin/INPUT line 17); noise amplitudes are unused on a restart
    p.update(sf=0.2, t_amp=0.0, b_amp=0.0)
    fields = {k: d[K] for k, K in zip(NAMES, ("R", "T", "U", "V", "W", "Bx", "By", "Bz"))}
    return p, fields


@pytest.fixture(scope="module")
def restart_oracle():
    p, fields = _restart_state()
    nz, ny, nx = fields["r"].shape
    o = Oracle(nx, ny, nz, p)
    for k in NAMES:
        o[k][:] = fields[k]
    o.ini_trans_params()
    o.mag_init()
    do = o.step(20, 0.0, 10)        # the reference's calcdt every 10 steps
    return p, fields, o, do


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_restart_fixture_steps(restart_oracle, prec):
    import fug30v0kxc_amd as fa
    p, fields, o, do = restart_oracle
    print(f"[baseline] restart.nc 30x30x60, 20 steps, {prec}", flush=True)
    nz, ny, nx = fields["r"].shape
    s = fa.Solver(nx, ny, nz, p, precision=fa.MHD_PREC_FP64 if prec == "fp64" else fa.MHD_PREC_MIXED)
    s.set_state(*[fields[k] for k in NAMES])
    try:
        ds = s.step(20, 0.0, 10)
        st = s.get_state()
        tol = 1e-10 if prec == "fp64" else 1e-5
        check(st, o, tol, f"restart {prec}")
        assert abs(s.scalars()["dt"] - o.scalar("dt")) <= tol * o.scalar("dt")
        for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
            assert abs(ds[key] - do[i]) <= tol * abs(do[i]), key
        # a developed state: the flow is O(1), not a small perturbation
        assert np.abs(o["w"]).max() > 0.1
    finally:
        s.close()
