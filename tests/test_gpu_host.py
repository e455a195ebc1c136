"""The Fortran host program (fug30v0kxc_amd/host/p3d_gpu, iso_c_binding over
the C ABI) end to end: the reference's INPUT file, a static start, the time
loop with diagnostics and T16_XXX.nc dumps (FP32 between restarts, FP64 at
the end: src/main.F90:695-740), and a restart from restart.nc (NFRAME 1,
src/main.F90:275-287, first step Euler).  Dumped fields are compared with
the CPU oracle run on the same parameters: FP64 dumps to 1e-10 relative,
FP32 dumps to float32 rounding."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle.pyoracle import Oracle
from tests.conftest import reference_params

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "fug30v0kxc_amd", "host", "p3d_gpu")
INTERP = os.path.join(ROOT, "fug30v0kxc_amd", "host", "interp_gpu")
NAMES = {"r": "R", "t": "T", "u": "U", "v": "V", "w": "W", "bx": "Bx", "by": "By", "bz": "Bz"}
NX, NY, NZ = 16, 16, 12
DT = 1e-4


def input_text(nframe, ntotal, ndu06, ndu16, ndurestart, ra=1e5, q=100.0):
    p = reference_params(ra=ra, q=q)
    rows = [
        "Compressible magnetoconvection", "----", "2", "----", "----",
        repr(p["gamma"]), repr(p["sigma"]), repr(p["theta"]), repr(ra), repr(p["cm"]), repr(q),
        repr(p["tayl"]), repr(p["psi"]), repr(p["zeta"]), repr(p["xmax"]), repr(p["ymax"]),
        repr(p["sf"]), str(nframe), str(ntotal), str(ndu06), str(ndu16), str(ndurestart),
        str(NX), str(NY), str(NZ)]
    return "\n".join(rows) + "\n"


def run_host(cwd, text, ext):
    (cwd / "INPUT").write_text(text)
    (cwd / "INPUT.ext").write_text(ext)
    print(f"[host] running p3d_gpu in {cwd}", flush=True)
    r = subprocess.run([EXE], cwd=cwd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def dump(path):
    from scipy.io import netcdf_file
    f = netcdf_file(str(path), "r", mmap=False)
    out = {n: np.array(f.variables[nc].data) for n, nc in NAMES.items()}
    prec = {n: a.dtype.itemsize for n, a in out.items()}
    out = {n: a.astype(np.float64) for n, a in out.items()}
    meta = {k: np.array(f.variables[k].data)[0] for k in ("time", "iteration", "timestep")}
    f.close()
    meta["itemsize"] = prec["t"]
    return out, meta


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


@pytest.fixture(scope="module")
def host_run(tmp_path_factory):
    if not os.path.exists(EXE):
        pytest.skip("p3d_gpu not built (no Fortran compiler where build() ran)")
    d = tmp_path_factory.mktemp("host")
    out = run_host(d, input_text(0, 10, 5, 5, 100), f"dt_fixed {DT}\nt_amp 1e-3\nb_amp 1e-3\n")
    return d, out


def test_static_start_and_dumps(host_run):
    d, out = host_run
    assert "Nusselt" in out and "T16_002.nc" in out
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    o = Oracle(NX, NY, NZ, p)
    o.setup_static()
    o.step(5, DT, 5)
    f1, m1 = dump(d / "T16_001.nc")          # FP32 dump at step 5
    assert m1["itemsize"] == 4
    assert int(m1["iteration"]) == 5
    for n in NAMES:
        ref = o[n]
        assert np.abs(f1[n] - ref).max() <= 2e-7 * max(np.abs(ref).max(), 1e-30), n
    o.step(5, DT, 5)
    f2, m2 = dump(d / "T16_002.nc")          # FP64 dump at the last step
    assert m2["itemsize"] == 8
    assert int(m2["iteration"]) == 10
    assert abs(m2["time"] - o.scalar("ttime")) < 1e-15
    for n in NAMES:
        assert rel(f2[n], o[n]) < 1e-10, n


def test_restart(host_run):
    d, _ = host_run
    r = d / "restart"
    r.mkdir()
    shutil.copy(d / "T16_002.nc", r / "restart.nc")
    run_host(r, input_text(1, 4, 2, 4, 1), f"dt_fixed {DT}\n")
    got, meta = dump(r / "T16_003.nc")       # iteration 14 -> 14/4 = 3
    assert int(meta["iteration"]) == 14
    start, m0 = dump(r / "restart.nc")
    p = reference_params(t_amp=1e-3, b_amp=1e-3)
    o = Oracle(NX, NY, NZ, p)
    o.setup_static()
    for n in NAMES:
        o[n][:] = start[n]
    o.mag_init()
    o.set_scalar("iter", int(m0["iteration"]))
    o.set_scalar("ttime", float(m0["time"]))
    o.set_scalar("dt1", float(m0["timestep"]))
    o.set_scalar("first", 1)
    o.set_scalar("second", 0)
    o.step(4, DT, 2)
    for n in NAMES:
        assert rel(got[n], o[n]) < 1e-10, n


def test_bad_input_fails_loudly(tmp_path):
    if not os.path.exists(EXE):
        pytest.skip("p3d_gpu not built")
    (tmp_path / "INPUT").write_text(input_text(0, 1, 1, 1, 1).replace(f"\n{NX}\n", "\n15\n"))
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert "ERROR" in r.stdout


def test_interpolation_tool(tmp_path):
    """interp_gpu: restart.nc (nx, ny, nz) -> output.nc (2nx, 2ny, nz), each
    field expanded as the reference's interpolate/main.F90:299-306."""
    if not os.path.exists(INTERP):
        pytest.skip("interp_gpu not built")
    import fug30v0kxc_amd as fa
    from oracle import pyoracle
    rng = np.random.default_rng(5)
    shape = (6, 12, 16)
    fields = {n: rng.standard_normal(shape) + (2.0 if n in ("r", "t") else 0.0) for n in fa.DUMP_FIELDS}
    fa.write_dump(tmp_path / "restart.nc", fields, iteration=42, time=1.25, timestep=1e-4,
                  params=reference_params(), iprec=8)
    print(f"[host] running interp_gpu in {tmp_path}", flush=True)
    r = subprocess.run([INTERP], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got, meta = dump(tmp_path / "output.nc")
    assert int(meta["iteration"]) == 42 and meta["itemsize"] == 8
    for n in fa.DUMP_FIELDS:
        want = pyoracle.expand(fields[n])
        assert got[n].shape == (6, 24, 32)
        assert np.abs(got[n] - want).max() <= 1e-12 * np.abs(want).max(), n


# ---- the reference's own smoke inputs ----------------------------------------
# bin/originalINPUT (4 x 4 x 8, gamma 5/3) and bin/test1/INPUT (30 x 30 x 60,
# gamma 1.4), copied unchanged to tests/golden/reference_inputs: Ra 1e5,
# Q 2500, a static start with the reference's T noise (0.1), calcdt every 10
# steps, 10 steps with diagnostics every 2.  No INPUT.ext: every setting is
# the reference's.  They write no dump (NDU16 210 > TOTAL), so the printed
# diagnostics are compared with the oracle's at each NDU06 step.
REF_INPUTS = os.path.join(ROOT, "tests", "golden", "reference_inputs")


def _diag_lines(out):
    """Per NDU06 step: {key: value} from the host's printout."""
    steps, cur = [], None
    for ln in out.splitlines():
        f = ln.replace("=", " ").split()
        if not f:
            continue
        if f[0] == "NIT":
            cur = {"nit": int(f[1]), "ttime": float(f[3])}
            steps.append(cur)
        elif cur is not None and f[0] in ("Nusselt", "kinenergy", "magenergy", "dt"):
            cur[f[0].lower()] = float(f[1])
        elif cur is not None and f[0] in ("rmin", "tmin", "wmin", "bzmin"):
            cur[f[0]] = float(f[1])
            cur[f[2]] = float(f[3])
    return steps


@pytest.mark.parametrize("name,box", [("originalINPUT", (4, 4, 8)), ("test1_INPUT", (30, 30, 60))])
def test_reference_smoke_inputs(tmp_path, name, box):
    if not os.path.exists(EXE):
        pytest.skip("p3d_gpu not built")
    text = open(os.path.join(REF_INPUTS, name)).read()
    (tmp_path / "INPUT").write_text(text)
    print(f"[host] p3d_gpu on the reference's {name}", flush=True)
    r = subprocess.run([EXE], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    got = _diag_lines(r.stdout)
    assert [d["nit"] for d in got] == [2, 4, 6, 8, 10]
    vals = text.split("\n")
    gamma = float(vals[5].split()[0])
    p = reference_params(ra=1e5, q=2500.0, gamma=gamma, t_amp=0.1)
    o = Oracle(*box, p)
    o.setup_static()
    prev_dt = None
    for d in got:
        od = o.step(2, 0.0, 2)
        assert abs(d["ttime"] - o.scalar("ttime")) <= 1e-12 * o.scalar("ttime")
        # the printed dt is the step's; calcdt at every 10th step comes after it
        dt_used = o.scalar("dt") if d["nit"] % 10 else prev_dt
        prev_dt = o.scalar("dt")
        assert abs(d["dt"] - dt_used) <= 1e-10 * dt_used
        for key, i in (("kinenergy", 2), ("magenergy", 3), ("nusselt", 4)):
            assert abs(d[key] - od[i]) <= 1e-10 * abs(od[i]), (d["nit"], key, d[key], od[i])
        for n, lo, hi in (("r", "rmin", "rmax"), ("t", "tmin", "tmax"), ("w", "wmin", "wmax")):
            scale = np.abs(o[n]).max()
            assert abs(d[lo] - o[n].min()) <= 1e-10 * scale and abs(d[hi] - o[n].max()) <= 1e-10 * scale
