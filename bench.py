#!/usr/bin/env python3
"""Throughput of the per-timestep compressible-MHD core on one MI355X.

Workload: the reference physics at 256x256x160 (aspect 4, gamma 5/3, theta 10,
m 1, Ra 1e6, Q 1000, zeta 0.2, kappa(z) of the reference), seeded T and Bz
perturbations of amplitude 1e-3, fixed dt, diagnostics every 100 steps.  A
step is one full pass of the time-loop body (hydro RHS, magnetic RHS, AB3
update, dealias cadence, B and J reconstruction, boundary conditions).
Inputs are resident in HBM before the timed region.

Prints one JSON line.  The roofline object covers the dominant kernel family
(the x-direction spectral first+second derivative of T, u, v, w), timed with
HIP events on the solver's stream over the timed steps.  cpu_baseline times
the CPU oracle (a serial C restatement of the reference) on a bounded sample
of the same workload: a 256x256x16 slab (same plane size, so the same
per-cell transform cost), a few steps, one thread.
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
HBM_PEAK_GBS = 8000.0
METRIC = "Mcell-steps/s at 256²×160: FP64 vs mixed-FP32; rel err (Nu, E) vs FP64 oracle"


def workload_params():
    from fug30v0kxc_amd import reference_params
    return reference_params(ra=1e6, q=1000.0, t_amp=1e-3, b_amp=1e-3)


def cpu_baseline(params, nz=16, steps=6):
    """Time the CPU oracle on a 256x256x`nz` slab (1 thread)."""
    from oracle.pyoracle import Oracle
    o = Oracle(NX, NY, nz, params)
    o.setup_static()
    o.step(1, DT, 0)          # first step includes the Euler start-up
    t0 = time.perf_counter()
    o.step(steps, DT, 0)
    dt = time.perf_counter() - t0
    return {"value": NX * NY * nz * steps / dt / 1e6, "unit": "Mcell-steps/s", "cores": 1,
            "kind": "port",
            "sample": f"CPU oracle (serial C restatement of the reference), {NX}x{NY}x{nz} slab, "
                      f"{steps} FP64 AB3 steps after 1 untimed, 1 thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--kernel", default="xderiv2", help="kernel family for the roofline object")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              "(the whole box sits on one MI355X, see the sharding design notes)", file=sys.stderr)
        return 2

    from fug30v0kxc_amd import Solver
    p = workload_params()
    s = Solver(NX, NY, NZ, p)
    s.static()
    s.step(args.warmup, DT, NDU06)
    s.sync()

    s.kernel_timing(args.kernel, True)
    s.sync()
    t0 = time.perf_counter()
    diag = s.step(args.steps, DT, NDU06)
    s.sync()
    t1 = time.perf_counter()
    kms, klaunch = s.kernel_time()
    s.kernel_timing(args.kernel, False)

    elapsed = t1 - t0
    cells = NX * NY * NZ
    value = cells * args.steps / elapsed / 1e6
    st = s.get_state()
    finite = all(float(abs(v).max()) < 1e30 for v in st.values())

    # algorithmic bytes of one x-derivative launch: 4 fields, 1 read + 2 writes each
    bytes_per_launch = {"xderiv2": 4 * 3 * 8 * cells, "yderiv2": 4 * 3 * 8 * cells}[args.kernel]
    avg_ms = kms / max(klaunch, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                "kernel": f"spec_kernel {args.kernel} (4 fields: d/dx and d2/dx2 of T,u,v,w)",
                "avg_launch_ms": round(avg_ms, 4), "launches": klaunch,
                "bytes_per_launch": bytes_per_launch}

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(p)

    out = {"metric": METRIC, "value": round(value, 2), "unit": "Mcell-steps/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(elapsed / args.steps * 1e3, 4), "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": "256x256x160 compressible MHD, FP64 plain, Ra=1e6, Q=1000, "
                                  "T/Bz noise 1e-3, fixed dt 5e-5, diagnostics every 100 steps",
                      "nx": NX, "ny": NY, "nz": NZ, "dt": DT, "ndu06": NDU06,
                      "precision": "fp64"},
           "roofline": roofline, "cpu_baseline": cpu,
           "diag": {"nusselt": diag.get("nusselt"), "kinenergy": diag.get("kinenergy"),
                    "magenergy": diag.get("magenergy"), "finite": finite}}
    if cpu:
        out["gpu_over_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(out))
    s.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
