#!/usr/bin/env python3
"""Throughput of the per-timestep compressible-MHD core on one MI355X.

Workload: the reference physics at 256x256x160 (aspect 4, gamma 5/3, theta 10,
m 1, Ra 1e6, Q 1000, zeta 0.2, kappa(z) of the reference), seeded T and Bz
perturbations of amplitude 1e-3, fixed dt, diagnostics every 100 steps.  A
step is one full pass of the time-loop body (hydro RHS, magnetic RHS, AB3
update, dealias cadence, B and J reconstruction, boundary conditions).
Inputs are resident in HBM before the timed region.

The headline value is the FP64 plain method.  The same workload then runs in
the mixed-precision mode (FP32 transforms, stencils and working arrays; FP64
state, means, backgrounds and reductions); its rate, its speed-up and its
relative error against the FP64 run (fields, Nu, kinetic and magnetic
energy) are reported under "mixed".

Prints one JSON line.  The roofline object covers the dominant kernel family,
hydro_col (the whole hydrodynamic right-hand side after the x/y transform
passes, with the fused AB3 update of u, v, w, rho, T: 55 arrays per cell in
FP64, 300 bytes per cell in the mixed mode), timed with HIP events on the
solver's stream over the timed steps; roofline_xderiv2 does the same for the
x-direction first+second derivative pass of T, u, v, w (with d/dx of rho in
the same launch) over a few further steps.  step_roofline puts
the whole step against the same peak at its algorithmic bytes per cell-step
(139 working-precision array passes, the FP64 state read and written, the RHS
history: 1448 B in FP64, 780 B mixed).  cpu_baseline times
the CPU oracle (a C restatement of the reference) on a bounded sample of the
same workload: a few steps of the full box on all host threads, and one
step on one thread.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NX, NY, NZ = 256, 256, 160
DT = 5e-5
NDU06 = 100
# the mixed-precision error is measured after this many steps of both modes
ERR_STEPS = 2000
AUTOTUNE = True

# (nx, ny, nz, dt, extra parameters, description)
WORKLOADS = {
    "c2": (256, 256, 160, 5e-5, {},
           "256x256x160 compressible MHD, Ra=1e6, Q=1000, T/Bz noise 1e-3, fixed dt 5e-5"),
    "c3": (512, 512, 256, 5e-6, {"eta_c": 100.0, "eta_z0": 0.3, "eta_w": 0.05},
           "512x512x256 compressible MHD, Ra=1e6, Q=1000, eta(z) contrast 100 at z=0.3 (width 0.05), "
           "T/Bz noise 1e-3, fixed dt 5e-6"),
}
HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-steps/s at 256²×160: FP64 vs mixed-FP32; rel err (Nu, E) vs FP64 oracle"


def workload_params(extra=None):
    from fug30v0kxc_amd import reference_params
    p = reference_params(ra=1e6, q=1000.0, t_amp=1e-3, b_amp=1e-3)
    p.update(extra or {})
    return p


def cpu_baseline(params, steps=3, steps1=1):
    """Time the CPU oracle (serial C restatement of the reference, OpenMP over
    planes and points, results independent of the thread count) on the full
    256x256x160 box: `steps` steps on all of this host's threads, then
    `steps1` step on one thread, each after one untimed step."""
    from oracle import pyoracle
    o = pyoracle.Oracle(NX, NY, NZ, params)
    o.setup_static()
    nthreads = pyoracle.threads()
    o.step(1, DT, 0)          # the Euler start-up step, untimed
    t0 = time.perf_counter()
    o.step(steps, DT, 0)
    dt = time.perf_counter() - t0
    pyoracle.set_threads(1)
    t1 = time.perf_counter()
    o.step(steps1, DT, 0)
    dt1 = time.perf_counter() - t1
    pyoracle.set_threads(nthreads)
    cells = NX * NY * NZ
    return {"value": round(cells * steps / dt / 1e6, 3), "unit": "Mcell-steps/s", "cores": nthreads,
            "kind": "port", "host_logical_cpus": os.cpu_count(),
            "single_thread": round(cells * steps1 / dt1 / 1e6, 3),
            "sample": f"CPU oracle (C restatement of the reference, FP64), full {NX}x{NY}x{NZ} box: "
                      f"{steps} AB3 steps on {nthreads} threads ({dt:.1f} s) and {steps1} on 1 thread "
                      f"({dt1:.1f} s), each after 1 untimed step"}


SECOND_KERNEL_STEPS = 5


def run(precision, steps, warmup, kernel, wl, second="xderiv2"):
    """Static start, `warmup` untimed steps, `steps` timed steps (kernel family
    `kernel` timed with events), then SECOND_KERNEL_STEPS untimed steps with
    the family `second` timed; returns the live solver with the timings."""
    from fug30v0kxc_amd import MHD_PREC_FP64, MHD_PREC_MIXED, Solver
    nx, ny, nz, dt, extra, _ = wl
    prec = MHD_PREC_FP64 if precision == "fp64" else MHD_PREC_MIXED
    s = Solver(nx, ny, nz, workload_params(extra), precision=prec)
    s.static()
    tune = s.autotune() if AUTOTUNE else None
    s.step(warmup, dt, NDU06)
    s.sync()
    s.kernel_timing(kernel, True)
    s.sync()
    t0 = time.perf_counter()
    diag = s.step(steps, dt, NDU06)
    s.sync()
    t1 = time.perf_counter()
    kms, klaunch = s.kernel_time()
    s.kernel_timing(kernel, False)
    out = {"solver": s, "elapsed": t1 - t0, "diag": diag, "kms": kms, "klaunch": klaunch, "tune": tune,
           "cells": nx * ny * nz, "device_bytes": s.device_bytes(), "done": warmup + steps}
    if second and second != kernel:
        s.kernel_timing(second, True)
        s.step(SECOND_KERNEL_STEPS, dt, NDU06)
        s.sync()
        out["kms2"], out["klaunch2"] = s.kernel_time()
        s.kernel_timing(second, False)
        out["done"] += SECOND_KERNEL_STEPS
    return out


C4_SEED = 20261202
C4_DT = 1e-6


def c4_specs(seed=C4_SEED, nbox=32):
    """The ensemble workload's box table: (n, nz, params) of `nbox` seeded
    independent boxes, nx = ny in {64, 96, 128}, nz in {48, 64, 96}, Ra and Q
    log-uniform in [1e4, 1e6] and [10, 1e3], eta(z) contrast log-uniform in
    [1, 100] at z = 0.5 (width 0.05), T and Bz noise 1e-3."""
    import numpy as np
    from fug30v0kxc_amd import reference_params
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(nbox):
        n = int(rng.choice([64, 96, 128]))
        nz = int(rng.choice([48, 64, 96]))
        ra = 10 ** rng.uniform(4, 6)
        q = 10 ** rng.uniform(1, 3)
        ec = 10 ** rng.uniform(0, 2)
        p = reference_params(ra=ra, q=q, t_amp=1e-3, b_amp=1e-3)
        p.update(eta_c=ec, eta_z0=0.5, eta_w=0.05)
        specs.append((n, nz, p))
    return specs


def ensemble_bench(steps, seed=C4_SEED):
    """32 seeded independent boxes (c4_specs), mixed precision, autotuned once
    per horizontal size (on the first box of that size; boxes of one size then
    launch the same kernels, which the ensemble batches); `steps` steps of all
    boxes by hipGraph replays against the same boxes stepped one after
    another."""
    import numpy as np
    from fug30v0kxc_amd import MHD_PREC_MIXED, Ensemble, Solver
    specs = c4_specs(seed)
    dt = C4_DT
    cells = sum(n * n * nz for n, nz, _ in specs)

    def shape(t):
        return {k: v for k, v in t.items() if not k.startswith("ms_")}

    def make(tunes=None):
        out, per_size = [], {}
        for i, (n, nz, p) in enumerate(specs):
            s = Solver(n, n, nz, p, precision=MHD_PREC_MIXED)
            s.static()
            if tunes is None:
                if n not in per_size:
                    per_size[n] = s.autotune()
                else:
                    s.set_tune(**shape(per_size[n]))
                s.tune = per_size[n]
            else:
                s.set_tune(**shape(tunes[i]))
            out.append(s)
        return out
    ens_boxes = make()
    t0 = time.perf_counter()
    ens = Ensemble(ens_boxes, dt)
    t_capture = time.perf_counter() - t0
    ens.step(ens.block)                      # first replay (untimed)
    t0 = time.perf_counter()
    ens.step(steps)
    t_graph = time.perf_counter() - t0
    nodes = ens.nodes
    ens.close()
    seq = make([b.tune for b in ens_boxes])     # the same launch shapes, so bit-identical
    for s in seq:
        s.step(2 + ens.block, dt, 0)
        s.sync()
    t0 = time.perf_counter()
    for s in seq:
        s.step(steps, dt, 0)
    for s in seq:
        s.sync()
    t_seq = time.perf_counter() - t0
    same = all(np.array_equal(a.get_state()["t"], b.get_state()["t"]) for a, b in zip(ens_boxes, seq))
    for s in ens_boxes + seq:
        s.close()
    return {"boxes": len(specs), "cells": cells, "steps": steps, "dt": dt,
            "graph_mcell_steps_per_s": round(cells * steps / t_graph / 1e6, 2),
            "sequential_mcell_steps_per_s": round(cells * steps / t_seq / 1e6, 2),
            "graph_speedup": round(t_seq / t_graph, 3), "graph_nodes": nodes,
            "capture_s": round(t_capture, 2), "identical_to_sequential": bool(same)}


def closure_bench(nfine, f=4, ra=1e6, nblocks=4, seed=20261203):
    """Learned closure: a fine run (nfine^2 x 3 nfine / 4, Ra `ra`, Q 500, T
    and Bz noise 0.1 and 0.01, mixed precision) is coarse-grained (factor f
    horizontally, onto the coarse vertices in z) into snapshots straight from
    its device arrays; the stencil MLP is trained on the GPU on the earlier
    ones.  Reported: the centred fraction of subgrid-flux variance it explains
    on a later, held-out snapshot (the plain coarse method has no subgrid term
    and scores 0); coarse runs with and without the closure from the
    filtered held-out state against the filtered fine run (kinetic-energy
    spectra, and Nu of the filtered fine state on the coarse grid) and the
    fallback decision; cost per unit of simulated time; the data path's and
    the forward and training kernels' rates."""
    import numpy as np
    from fug30v0kxc_amd import MHD_PREC_MIXED, Closure, Solver, closure_targets, reference_params
    from fug30v0kxc_amd import closure_eval as ce
    fine_box = (nfine, nfine, 3 * nfine // 4)
    coarse_box = tuple(n // f for n in fine_box)
    p = reference_params(ra=ra, q=500.0, t_amp=0.1, b_amp=0.01)
    fine = Solver(*fine_box, p, precision=MHD_PREC_MIXED)
    fine.static()
    blk = 500

    def note(msg):
        print(f"[c5 {nfine}/{f}] {msg}", file=sys.stderr, flush=True)

    # spin-up with the reference's calcdt until t = 0.25
    nspin = 0
    while fine.scalars()["ttime"] < 0.25:
        fine.step(blk, 0.0, blk)
        nspin += blk
        sc = fine.scalars()
        note(f"spin-up: {nspin} steps, t = {sc['ttime']:.4f}, dt = {sc['dt']:.3e}")
    dtf = fine.scalars()["dt"]
    cl = Closure(seed=1)
    add_ms = []
    for i in range(5):
        fine.step(blk, dtf, blk)
        fine.sync()
        t0 = time.perf_counter()
        cl.add_solver(fine, f)          # device-resident coarse-graining into the data set
        add_ms.append((time.perf_counter() - t0) * 1e3)
        note(f"snapshot {i + 1} of 5")
    fine.step(blk, dtf, blk)
    held = closure_targets(fine.get_state(), f, p["f"])
    t1 = time.perf_counter()
    loss = cl.train(6000, batch=4096, lr=1e-3)
    train_s = time.perf_counter() - t1
    note(f"trained: 6000 steps in {train_s:.2f} s, last batch loss {loss:.4f}")
    r2 = ce.explained_variance(cl.predict(held[0]), held[1])
    fwd_ms, train_ms, nfwd = cl.time(batch=8192, reps=20)
    flops_cell = 2 * (216 * 64 + 64 * 64 + 64 * 12)

    # Nu of a filtered fine state on the coarse grid
    probe = Solver(*coarse_box, p, precision=MHD_PREC_MIXED)

    def nu_filtered(coarse_state):
        probe.set_state(*[coarse_state[n] for n in ("r", "t", "u", "v", "w", "bx", "by", "bz")])
        return probe.nusselt()

    # a posteriori: nblocks blocks from the held-out state
    ref = []
    t2 = time.perf_counter()
    for _ in range(nblocks):
        d = fine.step(blk, dtf, blk)
        fine.sync()
        ref.append((closure_targets(fine.get_state(), f, p["f"])[0], d["nusselt"]))
        note(f"reference block {len(ref)} of {nblocks}")
    fine_ms_per_step = (time.perf_counter() - t2) / (nblocks * blk) * 1e3
    fine.close()
    nu_ref = [nu_filtered(st) for st, _ in ref]
    probe.close()
    runs = {}
    for mode in ("plain", "closure"):
        co = Solver(*coarse_box, p, precision=MHD_PREC_MIXED)
        co.static()
        c0 = held[0]
        co.set_state(*[c0[n] for n in ("r", "t", "u", "v", "w", "bx", "by", "bz")])
        if mode == "closure":
            co.attach_closure(cl)
        errs, nus = [], []
        co.step(2, f * dtf, 0)          # start-up steps, untimed
        co.sync()
        tc = time.perf_counter()
        nsteps = 2
        for st_ref, _ in ref:
            n = blk // f - (nsteps if not errs else 0)
            d = co.step(n, f * dtf, n)
            nsteps += n
            errs.append(ce.spectral_error(co.get_state(), st_ref))
            nus.append(d["nusselt"])
        co.sync()
        ms = (time.perf_counter() - tc) / (nsteps - 2) * 1e3
        nu_err = float(np.mean([ce.nusselt_error(a, b) for a, b in zip(nus, nu_ref)]))
        runs[mode] = {"spectral_error": [round(e, 4) for e in errs], "mean_spectral_error": round(float(np.mean(errs)), 4),
                      "nusselt": [round(x, 4) for x in nus], "mean_nusselt_error": round(nu_err, 4),
                      "ms_per_step": round(ms, 4)}
        note(f"coarse {mode}: mean spectral error {runs[mode]['mean_spectral_error']}, "
             f"Nu error {runs[mode]['mean_nusselt_error']}, {ms:.4f} ms/step")
        co.close()
    cost_fine = fine_ms_per_step * f                 # one coarse step's simulated time
    return {
        "fine_box": list(fine_box), "coarse_box": list(coarse_box), "filter": f, "ra": ra, "dt_fine": dtf,
        "data_path": {"add_snapshot_ms": [round(x, 2) for x in add_ms], "fine_cells": int(np.prod(fine_box)),
                      "coarse_cells": int(np.prod(coarse_box))},
        "training": {"snapshots": 5, "samples": cl.samples, "steps": 6000, "batch": 4096, "seconds": round(train_s, 3),
                     "last_batch_loss": round(loss, 4)},
        "a_priori_explained_variance": {"closure": [round(float(x), 3) for x in r2], "closure_mean": round(float(np.mean(r2)), 3),
                                        "plain": 0.0, "definition": "centred R^2 per flux, held-out snapshot"},
        "a_posteriori": {**runs, "fine_nusselt": [round(x[1], 4) for x in ref],
                         "filtered_fine_nusselt_on_coarse_grid": [round(x, 4) for x in nu_ref],
                         "decision": ce.choose(runs["closure"]["mean_spectral_error"],
                                               nu_error=runs["closure"]["mean_nusselt_error"]),
                         "fallback_threshold": ce.FALLBACK_SPECTRAL_ERROR,
                         "fallback_nusselt_threshold": ce.FALLBACK_NUSSELT_ERROR},
        "cost_ratio_fine_over_closure_coarse": round(cost_fine / runs["closure"]["ms_per_step"], 1),
        "mlp_forward": {"cells": nfwd, "ms": round(fwd_ms, 4),
                        "tflops": round(nfwd * flops_cell / (fwd_ms * 1e-3) / 1e12, 1), "peak_tflops_bf16": 2500.0},
        "mlp_train_step_ms": round(train_ms, 4),
    }


def _best(fn, reps=3):
    """Best wall time (s) of reps calls after one untimed call, and the last result."""
    out = fn()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        out = fn()
        best = min(best, time.perf_counter() - t0)
    return best, out


def aux_bench(cpu=True):
    """The rows either side of the step: netCDF dumps of a 256x256x160 FP64
    state (FP64 and FP32 variables, read back), expand to 512x512x160, the
    factor-4 box filter of a 512x512x384 field, the closure targets of a
    256x256x192 field set, and the diagnostic and calcdt reductions at
    256x256x160.  These entry points take host arrays, so their times include
    the transfers; the kernels alone are in the kernel trace of this command.
    With cpu, the CPU oracle's expand and box filter are timed on bounded
    samples beside them."""
    import tempfile
    import numpy as np
    import fug30v0kxc_amd as fa
    from fug30v0kxc_amd import MHD_PREC_FP64, Solver, closure_targets
    out = {}
    rng = np.random.default_rng(7)

    def note(msg):
        print(f"[aux] {msg}", file=sys.stderr, flush=True)

    # diagnostics and calcdt on a live FP64 box
    s = Solver(NX, NY, NZ, workload_params(), precision=MHD_PREC_FP64)
    s.static()
    s.step(3, DT, 0)
    s.sync()
    diag = {}
    for name, fn in (("GlobalAverages", s.GlobalAverages), ("nusselt", s.nusselt),
                     ("checkdivb", s.checkdivb), ("calcdt", s.calcdt)):
        t, _ = _best(lambda: [fn() for _ in range(20)], reps=3)
        diag[name] = round(t / 20 * 1e6, 1)
    out["diagnostics_us_per_call"] = diag
    st = s.get_state()
    s.close()
    note("diagnostics done")

    # dumps
    n3 = NX * NY * NZ
    with tempfile.TemporaryDirectory() as td:
        path = os.path.join(td, "T16_001.nc")
        dumps = {}
        for iprec in (8, 4):
            tw, _ = _best(lambda: fa.write_dump(path, st, iteration=3, iprec=iprec))
            size = os.path.getsize(path)
            tr, (back, _) = _best(lambda: fa.read_dump(path))
            if iprec == 8:
                exact = all(np.array_equal(back[k], st[k]) for k in fa.DUMP_FIELDS)
            else:
                exact = all(np.array_equal(back[k], st[k].astype(np.float32).astype(np.float64))
                            for k in fa.DUMP_FIELDS)
            dumps[f"fp{8 * iprec}"] = {"bytes": size, "write_ms": round(tw * 1e3, 1),
                                       "write_gbs": round(size / tw / 1e9, 2), "read_ms": round(tr * 1e3, 1),
                                       "read_gbs": round(size / tr / 1e9, 2), "round_trip_exact": bool(exact)}
        out["dump_256x256x160"] = dumps
    note("dumps done")

    # expand 256x256x160 -> 512x512x160 (one field)
    f = st["t"]
    te, e = _best(lambda: fa.expand(f))
    # without Nyquist content (which expand keeps at full amplitude as a fine-grid
    # cosine, interpolate/fft.F90:888, 913) the coarse points are fine points:
    # check that on a band-limited field of the same shape
    spec = np.fft.rfft2(rng.standard_normal(f.shape))
    spec[:, NY // 2, :] = 0.0
    spec[:, :, NX // 2] = 0.0
    bl = np.fft.irfft2(spec, s=(NY, NX))
    rt = float(np.abs(fa.expand(bl)[:, ::2, ::2] - bl).max() / np.abs(bl).max())
    out["expand_256x256x160"] = {"ms": round(te * 1e3, 2), "out_cells": int(e.size),
                                 "host_bytes": int(8 * (f.size + e.size)),
                                 "host_inclusive_gbs": round(8 * (f.size + e.size) / te / 1e9, 2),
                                 "coarse_points_rel_err": float(f"{rt:.2e}")}
    del e
    note("expand done")

    # box filter 512x512x384 -> 128x128x96
    g = rng.standard_normal((384, 512, 512))
    tb, b = _best(lambda: fa.box_filter(g, 4))
    out["box_filter_512x512x384_f4"] = {"ms": round(tb * 1e3, 2), "in_cells": int(g.size),
                                        "mean_rel_err": float(f"{abs(b.mean() - g.mean()) / g.std():.2e}")}
    note("box filter done")

    # closure targets of 8 fields at 256x256x192, factor 4
    fine = {k: rng.standard_normal((192, 256, 256)) for k in fa.CLOSURE_FIELDS}
    tt, _ = _best(lambda: closure_targets(fine, 4, 0.5))
    out["closure_targets_256x256x192_f4"] = {"ms": round(tt * 1e3, 2), "in_cells": 192 * 256 * 256}
    note("targets done")

    if cpu:
        from oracle import pyoracle
        nth = pyoracle.threads()
        fs = np.ascontiguousarray(f[:16])
        tc, _ = _best(lambda: pyoracle.expand(fs), reps=1)
        tcb, _ = _best(lambda: pyoracle.box_filter(g, 4), reps=1)
        out["cpu_baseline"] = {"kind": "port", "cores": nth,
                               "expand_ms_per_plane": round(tc / 16 * 1e3, 3),
                               "expand_ms_full_field_equiv": round(tc / 16 * NZ * 1e3, 1),
                               "box_filter_512x512x384_ms": round(tcb * 1e3, 1),
                               "sample": f"CPU oracle expand of 16 of the {NZ} planes (scaled to the field) "
                                         f"and box filter of the whole 512x512x384 field, {nth} threads"}
        note("cpu baseline done")
    return out


def advance(r, total, dt):
    """Step a run on to `total` steps; returns the last diagnostics."""
    left = total - r["done"]
    d = r["diag"]
    if left > 0:
        d = r["solver"].step(left, dt, NDU06)
        r["done"] = total
    return d


def measured_traffic(precision, kernel):
    """HBM bytes per launch of a kernel family from the committed FETCH_SIZE /
    WRITE_SIZE summary (profiles/hbm_bytes_256x256x160.md, lines
    `traffic_bytes_per_launch <kernel> <precision> <bytes>`), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_bytes_256x256x160.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        f = line.split()
        if len(f) == 4 and f[0] == "traffic_bytes_per_launch" and f[1] == kernel and f[2] == precision:
            return int(f[3])
    return None


# algorithmic bytes per cell of one launch of each timed kernel family
#  xderiv2: 4 fields, 1 read + 2 writes each, and rho, 1 read + 1 write
#           (its first derivative only), working precision: 14 arrays
#  yderiv2: the same along y with three single-output jobs (rho, d/dx u,
#           d/dx v): 18 arrays
#  hydro:   hydro_col with the fused AB3 update (interior and wall-row launches):
#           reads the state (5), B and J^2 (4), the 26 x/y derivative arrays and
#           history levels 1, 2 of 5 fields; writes 5 RHS and the 5 new state
#           fields.  FP64: 55 x 8 = 440 B.  Mixed: FP32 state shadows, B, J^2,
#           derivatives and history, plus the FP64 state read and written and
#           the new FP32 shadows: 4 (5 + 4 + 26 + 10 + 5 + 5) + 8 (5 + 5) = 300 B.
KERNEL_BYTES_PER_CELL = {"xderiv2": {8: 112, 4: 56}, "yderiv2": {8: 144, 4: 72}, "hydro": {8: 440, 4: 300}}
KERNEL_DESC = {"xderiv2": "SpecK<T, plan, 0, 2, 1, 0> (xderiv2: d/dx and d2/dx2 of T, u, v, w; d/dx of rho)",
               "yderiv2": "SpecK<T, plan, 1, B, 1, 0> (yderiv2: d/dy and d2/dy2 of T, u, v, w; d/dy of rho, "
                          "d/dx u, d/dx v)",
               "hydro": "HydroColKernel<T, mixed, false> + <T, mixed, true> (hydro_col, interior and wall "
                        "rows): hydro_deriv's right-hand side after the x/y passes with the fused AB3 "
                        "update of u, v, w, rho, T"}


# algorithmic bytes per cell-step of the whole step: 139 working-precision array
# passes, the FP64 state (7 in, 7 out) and the RHS history (21 in, 7 out)
def step_bytes_per_cell(elem):
    return 139 * elem + 14 * 8 + 28 * elem


def step_roofline(cells, ms_per_step, elem):
    """The whole step against the HBM peak, at the algorithmic bytes per cell-step."""
    b = step_bytes_per_cell(elem)
    achieved = b * cells / (ms_per_step * 1e-3) / 1e9
    return {"bound": "hbm", "bytes_per_cell_step": b, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "ceiling_mcell_steps_per_s": round(HBM_PEAK_GBS * 1e9 / b / 1e6, 1)}


def roofline_of(r, kernel, elem, workload="c2", second=False):
    cells = r["cells"]
    bytes_per_launch = KERNEL_BYTES_PER_CELL[kernel][elem] * cells
    kms, nl = (r["kms2"], r["klaunch2"]) if second else (r["kms"], r["klaunch"])
    avg_ms = kms / max(nl, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = measured_traffic("fp64" if elem == 8 else "mixed", kernel) if workload == "c2" else None
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic, "kernel": KERNEL_DESC[kernel],
            "avg_launch_ms": round(avg_ms, 4), "launches": nl,
            "bytes_per_launch": bytes_per_launch}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--precision", default="both", choices=("both", "fp64", "mixed"),
                    help="both: FP64 headline plus the mixed run; fp64 / mixed: that run alone "
                         "(e.g. for a kernel profile of one mode)")
    ap.add_argument("--kernel", default="hydro", choices=("hydro", "xderiv2", "yderiv2"),
                    help="kernel family for the roofline object (xderiv2 is also reported beside it)")
    ap.add_argument("--c5-fine", type=int, default=256,
                    help="c5 (closure) workload: fine nx = ny (nz = 3/4 of it); 512 is the 512^2 x 384 case")
    ap.add_argument("--c5-factor", type=int, default=4, help="c5: coarse-graining factor")
    ap.add_argument("--c5-ra", type=float, default=1e6, help="c5: Rayleigh number")
    ap.add_argument("--c5-blocks", type=int, default=4, help="c5: a-posteriori blocks of 500 fine steps")
    ap.add_argument("--workload", default="c2", choices=sorted(WORKLOADS) + ["c4", "c5", "aux"],
                    help="c2: 256x256x160 (default); c3: 512x512x256 with the eta(z) step; "
                         "c4: ensemble of 32 boxes in hipGraph replays; c5: learned closure; "
                         "aux: dumps, expand, box filter, closure targets, diagnostics")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--err-steps", type=int, default=ERR_STEPS,
                    help="steps after which the mixed run is compared with the FP64 run")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a plain 1 GiB device copy (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = 1 << 27                      # 2^27 doubles = 1 GiB read + 1 GiB written
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 8 * n, "copy_bytes_written": 8 * n, "launches": 3}))
        return 0
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the whole box sits on one MI355X, see the sharding design notes)", file=sys.stderr)
        return 2

    import numpy as np
    global AUTOTUNE
    AUTOTUNE = not args.no_autotune
    if args.workload == "c4":
        steps = max(30, (args.steps // 30) * 30)
        r = ensemble_bench(steps)
        out = {"metric": METRIC, "value": r["graph_mcell_steps_per_s"], "unit": "Mcell-steps/s",
               "n_gpus": 1, "steps": steps, "warmup": 32, "ms_per_step": None,
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "f32 (FP64 state and reductions)", "data": "synthetic",
               "config": {"workload": "ensemble of 32 seeded boxes, mixed precision, hipGraph replays",
                          **{k: v for k, v in r.items() if k not in ("graph_mcell_steps_per_s",)}},
               "roofline": None, "cpu_baseline": None}
        print(json.dumps(out))
        return 0
    if args.workload == "aux":
        r = aux_bench(cpu=not args.no_cpu_baseline)
        ex = r["expand_256x256x160"]
        out = {"metric": "expand 256x256x160 -> 512x512x160 output cells per second, host arrays in and out",
               "value": round(ex["out_cells"] / (ex["ms"] * 1e-3) / 1e6, 1), "unit": "Mcell/s", "n_gpus": 1,
               "steps": None, "warmup": None, "ms_per_step": None, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
               "config": {"workload": "rows either side of the step (dumps, expand, box filter, "
                                      "closure targets, diagnostics)", **{k: v for k, v in r.items()
                                                                          if k != "cpu_baseline"}},
               "roofline": None, "cpu_baseline": r.get("cpu_baseline")}
        print(json.dumps(out))
        return 0
    if args.workload == "c5":
        r = closure_bench(args.c5_fine, args.c5_factor, args.c5_ra, args.c5_blocks)
        out = {"metric": "closure: explained subgrid-flux variance (held out) and coarse/fine spectral error",
               "value": r["a_priori_explained_variance"]["closure_mean"], "unit": "fraction", "n_gpus": 1,
               "steps": None, "warmup": None, "ms_per_step": r["a_posteriori"]["closure"]["ms_per_step"],
               "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
               "dtype": "bf16 (FP32 accumulate); mixed-precision solver", "data": "synthetic",
               "config": {"workload": f"learned closure, fine run coarse-grained by {args.c5_factor}, "
                                      f"Ra {args.c5_ra:g}, Q 500", **r},
               "roofline": {"bound": "mfma", "achieved": r["mlp_forward"]["tflops"], "peak": 2500.0,
                            "unit": "TFLOP/s", "frac": round(r["mlp_forward"]["tflops"] / 2500.0, 4),
                            "traffic": None},
               "cpu_baseline": None}
        print(json.dumps(out))
        return 0
    wl = WORKLOADS[args.workload]
    nx, ny, nz, dt, extra, desc = wl
    cells = nx * ny * nz
    if args.precision == "mixed":
        rm = run("mixed", args.steps, args.warmup, args.kernel, wl)
        vm = cells * args.steps / rm["elapsed"] / 1e6
        out = {"metric": METRIC, "value": round(vm, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
               "steps": args.steps, "warmup": args.warmup,
               "ms_per_step": round(rm["elapsed"] / args.steps * 1e3, 4), "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32 (FP64 state and reductions)",
               "data": "synthetic",
               "config": {"workload": desc + ", mixed precision, diagnostics every 100 steps",
                          "nx": nx, "ny": ny, "nz": nz, "dt": dt, "ndu06": NDU06, "precision": "mixed"},
               "roofline": roofline_of(rm, args.kernel, 4, args.workload), "cpu_baseline": None,
               "step_roofline": step_roofline(cells, rm["elapsed"] / args.steps * 1e3, 4),
               "device_bytes": rm["device_bytes"]}
        if "kms2" in rm:
            out["roofline_xderiv2"] = roofline_of(rm, "xderiv2", 4, args.workload, second=True)
        rm["solver"].close()
        print(json.dumps(out))
        return 0
    r64 = run("fp64", args.steps, args.warmup, args.kernel, wl)
    value = cells * args.steps / r64["elapsed"] / 1e6
    d64 = r64["diag"]

    mixed = None
    if args.precision == "both":
        rm = run("mixed", args.steps, args.warmup, args.kernel, wl)
        vm = cells * args.steps / rm["elapsed"] / 1e6
        # error of the mixed run against the FP64 run after err_steps steps of
        # both (the FP64 path matches the FP64 oracle to 1e-10 in the parity tests)
        nerr = max(args.err_steps, r64["done"])
        e64 = advance(r64, nerr, dt)
        em = advance(rm, nerr, dt)
        s64, sm = r64["solver"].get_state(), rm["solver"].get_state()
        bscale = max(float(np.abs(s64[n]).max()) for n in ("bx", "by", "bz"))
        fe = {n: float(np.abs(sm[n] - s64[n]).max() /
                       (bscale if n in ("bx", "by", "bz") else max(float(np.abs(s64[n]).max()), 1e-300)))
              for n in s64}
        qe = {k: (float(abs(em[k] - e64[k]) / abs(e64[k])) if e64.get(k) else None)
              for k in ("nusselt", "kinenergy", "magenergy")}
        mixed = {"value": round(vm, 2), "ms_per_step": round(rm["elapsed"] / args.steps * 1e3, 4),
                 "speedup_vs_fp64": round(vm / value, 3),
                 "rel_err_vs_fp64": {"after_steps": nerr,
                                     **{k: (float(f"{v:.3e}") if v is not None else None)
                                        for k, v in qe.items()},
                                     "fields": {k: float(f"{v:.3e}") for k, v in fe.items()}},
                 "roofline": roofline_of(rm, args.kernel, 4, args.workload),
                 "roofline_xderiv2": (roofline_of(rm, "xderiv2", 4, args.workload, second=True)
                                      if "kms2" in rm else None),
                 "step_roofline": step_roofline(cells, rm["elapsed"] / args.steps * 1e3, 4),
                 "tune": rm.get("tune"),
                 "device_bytes": rm["device_bytes"]}
        rm["solver"].close()
    st = r64["solver"].get_state()
    finite = all(float(np.abs(v).max()) < 1e30 for v in st.values())
    r64["solver"].close()

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(workload_params()) if args.workload == "c2" else None

    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(r64["elapsed"] / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": desc + ", FP64 plain method, diagnostics every 100 steps",
                      "nx": nx, "ny": ny, "nz": nz, "dt": dt, "ndu06": NDU06,
                      "precision": "fp64"},
           "roofline": roofline_of(r64, args.kernel, 8, args.workload), "cpu_baseline": cpu,
           "roofline_xderiv2": (roofline_of(r64, "xderiv2", 8, args.workload, second=True)
                                if "kms2" in r64 else None),
           "step_roofline": step_roofline(cells, r64["elapsed"] / args.steps * 1e3, 8),
           "mixed": mixed,
           "tune": r64.get("tune"),
           "device_bytes": r64["device_bytes"],
           # the diagnostics of the last ndu06 step inside the timed steps (null when
           # none fell there), and whether the final state is finite
           "diag": ({"nit": d64.get("nit"), "nusselt": d64.get("nusselt"), "kinenergy": d64.get("kinenergy"),
                     "magenergy": d64.get("magenergy")} if d64.get("nit", -1) >= 0 else None),
           "state_finite": finite}
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
