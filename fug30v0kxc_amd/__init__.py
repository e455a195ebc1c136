"""fug30v0kxc_amd -- MI355X-native per-timestep core of the p3d compressible
magnetoconvection solver (reference "VarElecCond", Fortran 90).

The package is the host-side mirror of the reference's time-step interface.
The reference exposes argument-less module subroutines acting on module
state (src/main.F90:340-747): hydro_deriv, mag_deriv, bandj, magdealias,
boundary, mag_boundary, GlobalAverages, nusselt, checkdivb, sminmax, calcdt,
plus the operators d1bydx, d1bydy, d2bydx, d2bydy, dealias, forwardfft,
inversefft (src/fft.F90) and d1bydz, d1upbydz, d2bydz (src/dbydz.F90).
`Solver` keeps those names and argument meanings; every call goes through
the C ABI (include/mhd.h) to hand-written gfx950 kernels.  Errors carry the
reference's abort / iret codes (MhdError.code).

Arrays are numpy float64 of shape (nz, ny, nx) in C order, which is the
reference's Fortran (nx, ny, nz) column-major layout.
"""
import ctypes

import numpy as np

from . import _abi
from ._abi import (FIELD_IDS, HISTORY_FIELDS, KERNELS, MHD_PREC_FP64, MHD_PREC_MIXED, OPS,
                   OPT_CONSTFLUX, OPT_DYNAMO, OPT_KINEMATIC, OPT_PERFECT, OPT_RADBC,
                   VECTOR_FIELDS, MhdDiag, MhdParams, MhdScalars)

__all__ = ["Solver", "MhdError", "ra_q_to_ck_f", "reference_params", "write_dump", "read_dump", "expand", "box_filter", "Ensemble", "MHD_PREC_FP64",
           "MHD_PREC_MIXED"]


class MhdError(RuntimeError):
    """A non-zero return code of the C ABI (the reference's abort codes)."""

    def __init__(self, code, what):
        msg = _abi.load().mhd_strerror(code).decode()
        super().__init__(f"{what}: {msg} (code {code})")
        self.code = code


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        raise ValueError(f"expected shape {shape}, got {a.shape}")
    return a


def ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q):
    """INPUT flag 2: Rayleigh and Chandrasekhar numbers -> CK, F
    (src/initial.F90:183-195)."""
    ck = ctypes.c_double()
    f = ctypes.c_double()
    _abi.load().mhd_ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q, ctypes.byref(ck),
                                 ctypes.byref(f))
    return ck.value, f.value


def reference_params(ra=1e5, q=100.0, t_amp=1e-3, b_amp=0.0, gamma=5.0 / 3.0, sigma=1.0,
                     theta=10.0, cm=1.0, zeta=0.2, xmax=4.0, ymax=4.0, sf=0.2, tayl=0.0,
                     psi=1.570796327):
    """Parameter dict of the reference's INPUT (flag 2) with Ra, Q converted."""
    ck, f = ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q)
    return dict(gamma=gamma, sigma=sigma, theta=theta, ck=ck, cm=cm, f=f, tayl=tayl, psi=psi,
                zeta=zeta, xmax=xmax, ymax=ymax, sf=sf, t_amp=t_amp, b_amp=b_amp)


DUMP_FIELDS = ("r", "t", "u", "v", "w", "bx", "by", "bz")   # R T U V W Bx By Bz


def write_dump(path, fields, iteration=0, time=0.0, timestep=0.0, params=None, iprec=8,
               history="", version=3):
    """Write a T16-style dump (WriteNetCDF layout, netCDF classic).  `fields`
    maps r, t, u, v, w, bx, by, bz to (nz, ny, nx) arrays."""
    arrs = [_f64(fields[n]) for n in DUMP_FIELDS]
    nz, ny, nx = arrs[0].shape
    m = _abi.MhdDump()
    m.nx, m.ny, m.nz, m.version, m.iteration = nx, ny, nz, version, iteration
    m.time, m.timestep = time, timestep
    for k, v in (params or {}).items():
        if hasattr(m, k) and k not in ("nx", "ny", "nz"):
            setattr(m, k, v)
    m.history = history.encode()[:255]
    ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dp(a) for a in arrs])
    code = _abi.load().mhd_write_dump(str(path).encode(), ctypes.byref(m), ptrs, iprec)
    if code:
        raise MhdError(code, f"write_dump({path})")


def read_dump(path):
    """(fields, meta) of a dump/restart file; fields as (nz, ny, nx) float64."""
    L = _abi.load()
    m = _abi.MhdDump()
    code = L.mhd_read_dump(str(path).encode(), ctypes.byref(m), None)
    if code:
        raise MhdError(code, f"read_dump({path})")
    arrs = [np.empty((m.nz, m.ny, m.nx)) for _ in DUMP_FIELDS]
    ptrs = (ctypes.POINTER(ctypes.c_double) * 8)(*[_dp(a) for a in arrs])
    code = L.mhd_read_dump(str(path).encode(), ctypes.byref(m), ptrs)
    if code:
        raise MhdError(code, f"read_dump({path})")
    meta = {n: getattr(m, n) for n, _ in _abi.MhdDump._fields_ if n != "history"}
    meta["history"] = m.history.decode(errors="replace")
    return dict(zip(DUMP_FIELDS, arrs)), meta


def expand(field):
    """The reference's expand (interpolate/fft.F90:832-927) on the GPU: a
    (nz, ny, nx) field -> (nz, 2ny, 2nx) by spectral zero-padding."""
    a = _f64(field)
    nz, ny, nx = a.shape
    out = np.empty((nz, 2 * ny, 2 * nx))
    code = _abi.load().mhd_expand(nx, ny, nz, _dp(a), _dp(out))
    if code:
        raise MhdError(code, "expand")
    return out


def box_filter(field, factor):
    """Mean over factor^3 blocks: (nz, ny, nx) -> (nz, ny, nx) // factor."""
    a = _f64(field)
    nz, ny, nx = a.shape
    f = int(factor)
    if f < 1:
        raise ValueError("factor must be >= 1")
    out = np.empty((max(nz // f, 0), max(ny // f, 0), max(nx // f, 0)))
    code = _abi.load().mhd_box_filter(nx, ny, nz, f, _dp(a), _dp(out))
    if code:
        raise MhdError(code, "box_filter")
    return out


class Solver:
    """One box of the solver on one GPU (one C-ABI context)."""

    def __init__(self, nx, ny, nz, params, precision=MHD_PREC_FP64):
        self._lib = _abi.load()
        self._ctx = None
        p = MhdParams(**params)
        ctx = ctypes.c_void_p()
        self._chk(self._lib.mhd_create(ctypes.byref(p), nx, ny, nz, precision,
                                       ctypes.byref(ctx)), "mhd_create")
        self._ctx = ctx
        self.nx, self.ny, self.nz = nx, ny, nz
        self.shape = (nz, ny, nx)
        self.params = dict(params)
        self.precision = precision

    def close(self):
        if self._ctx is not None:
            self._lib.mhd_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @staticmethod
    def _chk(code, what):
        if code != 0:
            raise MhdError(code, what)

    # -- state ------------------------------------------------------------
    def static(self):
        """static() + ini_trans_params() + mag_init() (src/main.F90:275-333)."""
        self._chk(self._lib.mhd_static(self._ctx), "static")

    def set_state(self, r, t, u, v, w, bx, by, bz):
        arrs = [_f64(a, self.shape) for a in (r, t, u, v, w, bx, by, bz)]
        self._chk(self._lib.mhd_set_state(self._ctx, *[_dp(a) for a in arrs]), "set_state")

    def get_state(self):
        """dict of the 8 dumped fields (R, T, U, V, W, Bx, By, Bz)."""
        names = ("r", "t", "u", "v", "w", "bx", "by", "bz")
        out = {n: np.empty(self.shape) for n in names}
        self._chk(self._lib.mhd_get_state(self._ctx, *[_dp(out[n]) for n in names]),
                  "get_state")
        return out

    def field(self, name, level=3):
        n = self.nz if name in VECTOR_FIELDS else None
        out = np.empty(n) if n else np.empty(self.shape)
        self._chk(self._lib.mhd_get_field(self._ctx, FIELD_IDS[name], level, _dp(out)),
                  f"get_field({name})")
        return out

    def set_field(self, name, value, level=3):
        shape = (self.nz,) if name in VECTOR_FIELDS else self.shape
        a = _f64(value, shape)
        self._chk(self._lib.mhd_set_field(self._ctx, FIELD_IDS[name], level, _dp(a)),
                  f"set_field({name})")

    def scalars(self):
        s = MhdScalars()
        self._chk(self._lib.mhd_get_scalars(self._ctx, ctypes.byref(s)), "get_scalars")
        return {n: getattr(s, n) for n, _ in MhdScalars._fields_}

    def set_scalars(self, **kw):
        s = MhdScalars()
        self._chk(self._lib.mhd_get_scalars(self._ctx, ctypes.byref(s)), "get_scalars")
        for k, v in kw.items():
            setattr(s, k, v)
        self._chk(self._lib.mhd_set_scalars(self._ctx, ctypes.byref(s)), "set_scalars")

    # -- the reference's per-step routines ----------------------------------
    def hydro_deriv(self):
        self._chk(self._lib.mhd_hydro_deriv(self._ctx), "hydro_deriv")

    def mag_deriv(self):
        self._chk(self._lib.mhd_mag_deriv(self._ctx), "mag_deriv")

    def bandj(self):
        self._chk(self._lib.mhd_bandj(self._ctx), "bandj")

    def magdealias(self):
        self._chk(self._lib.mhd_magdealias(self._ctx), "magdealias")

    def boundary(self):
        self._chk(self._lib.mhd_boundary(self._ctx), "boundary")

    def mag_boundary(self):
        self._chk(self._lib.mhd_mag_boundary(self._ctx), "mag_boundary")

    def mag_init(self):
        self._chk(self._lib.mhd_mag_init(self._ctx), "mag_init")

    def step(self, nsteps, dt_fixed=0.0, ndu06=0):
        """nsteps passes of the time-loop body (src/main.F90:340-683).
        Returns the last diagnostics as a dict (NaN-free only if computed)."""
        d = MhdDiag()
        d.nit = -1
        self._chk(self._lib.mhd_step(self._ctx, nsteps, dt_fixed, ndu06, ctypes.byref(d)),
                  "step")
        return {n: getattr(d, n) for n, _ in MhdDiag._fields_}

    # -- diagnostics (src/utils.F90) ----------------------------------------
    def GlobalAverages(self):
        """(ubar, vbar, kenergy, menergy); subtracts ubar, vbar from u, v."""
        out = np.zeros(4)
        self._chk(self._lib.mhd_global_averages(self._ctx, _dp(out)), "GlobalAverages")
        return tuple(out)

    def nusselt(self):
        v = ctypes.c_double()
        self._chk(self._lib.mhd_nusselt(self._ctx, ctypes.byref(v)), "nusselt")
        return v.value

    def checkdivb(self):
        out = np.zeros(2)
        self._chk(self._lib.mhd_checkdivb(self._ctx, _dp(out)), "checkdivb")
        return tuple(out)

    def sminmax(self, name):
        out = np.zeros(2)
        self._chk(self._lib.mhd_sminmax(self._ctx, FIELD_IDS[name], _dp(out)), "sminmax")
        return tuple(out)

    def calcdt(self):
        v = ctypes.c_double()
        self._chk(self._lib.mhd_calcdt(self._ctx, ctypes.byref(v)), "calcdt")
        return v.value

    # -- operators (src/fft.F90, src/dbydz.F90) -----------------------------
    def _op(self, name, fin, flag=0, two=False, vec=False):
        shape = (self.nz,) if vec else self.shape
        a = _f64(fin, shape)
        o1 = np.empty(shape)
        o2 = np.empty(shape) if two else None
        self._chk(self._lib.mhd_op(self._ctx, OPS[name], _dp(a), _dp(o1),
                                   _dp(o2) if two else None, flag), name)
        return (o1, o2) if two else o1

    def d1bydx(self, fin):
        return self._op("d1bydx", fin)

    def d1bydy(self, fin):
        return self._op("d1bydy", fin)

    def d2bydx(self, fin):
        """(first, second) x-derivative, like d2bydx(fin, f1out, f2out)."""
        return self._op("d2bydx", fin, two=True)

    def d2bydy(self, fin):
        return self._op("d2bydy", fin, two=True)

    def dealias(self, z):
        return self._op("dealias", z)

    def forwardfft(self, fin):
        return self._op("forwardfft", fin)

    def inversefft(self, fin):
        return self._op("inversefft", fin)

    def d1bydz(self, fin, dflag):
        return self._op("d1bydz", fin, dflag)

    def d1upbydz(self, fin, dflag):
        """Upwind d/dz selected by the solver's own w (as the reference)."""
        return self._op("d1upbydz", fin, dflag)

    def d2bydz(self, fin, dflag):
        return self._op("d2bydz", fin, dflag)

    def d1bydz1D(self, fin, dflag):
        return self._op("d1bydz1D", fin, dflag, vec=True)

    def d2bydz1D(self, fin, dflag):
        return self._op("d2bydz1D", fin, dflag, vec=True)

    # -- measurement ------------------------------------------------------
    def kernel_timing(self, kernel, enable=True):
        self._chk(self._lib.mhd_kernel_timing(self._ctx, KERNELS[kernel], int(enable)),
                  "kernel_timing")

    def kernel_time(self):
        ms = ctypes.c_double()
        n = ctypes.c_long()
        self._chk(self._lib.mhd_kernel_time(self._ctx, ctypes.byref(ms), ctypes.byref(n)),
                  "kernel_time")
        return ms.value, n.value

    def autotune(self):
        """Time this box's launch shapes (per transform direction and kind the
        radix plan and the persistent-grid multiple; the planes per thread of
        the hydro column kernel) and keep the fastest; returns the choices and
        their times (variant_x, variant_y, mult_x, mult_y: one per kind, d1,
        d1+d2, r2hc, hc2r)."""
        t = _abi.MhdTune()
        self._chk(self._lib.mhd_autotune(self._ctx, ctypes.byref(t)), "autotune")
        out = {n: getattr(t, n) for n, _ in _abi.MhdTune._fields_}
        for n in ("variant_x", "variant_y", "mult_x", "mult_y"):
            out[n] = list(getattr(t, n))
        return out

    def set_tune(self, variant_x=0, variant_y=0, mult_x=1, mult_y=1, zc=8):
        """Apply launch-shape choices (as returned by autotune); a single
        variant or multiple applies to every transform kind of its direction."""
        def kinds(v):
            v = [v] * _abi.TUNE_KINDS if isinstance(v, int) else list(v)
            if len(v) != _abi.TUNE_KINDS:
                raise ValueError(f"a value per transform kind ({_abi.TUNE_KINDS}) or one for all")
            return (ctypes.c_int * _abi.TUNE_KINDS)(*v)
        t = _abi.MhdTune(variant_x=kinds(variant_x), variant_y=kinds(variant_y), mult_x=kinds(mult_x),
                         mult_y=kinds(mult_y), zc=zc)
        self._chk(self._lib.mhd_set_tune(self._ctx, ctypes.byref(t)), "set_tune")

    def sync(self):
        self._chk(self._lib.mhd_sync(self._ctx), "sync")

    def device_bytes(self):
        return self._lib.mhd_device_bytes(self._ctx)


class Ensemble:
    """Independent boxes (Solver objects on this GPU) stepped together: the
    kernel launches of `block` steps of every box are recorded, the k-th
    launches of the boxes of one class (same lengths, precision, options)
    become one launch over a table of boxes, and those batched launches are
    captured once into one hipGraph and replayed.  Results equal
    Solver.step(n, dt) per box, bit for bit.  A box must not be stepped on
    its own while it belongs to an ensemble."""

    def __init__(self, solvers, dt):
        self._lib = _abi.load()
        self._e = None
        self.solvers = list(solvers)
        arr = (ctypes.c_void_p * len(self.solvers))(*[sv._ctx.value for sv in self.solvers])
        e = ctypes.c_void_p()
        code = self._lib.mhd_ensemble_create(arr, len(self.solvers), float(dt), ctypes.byref(e))
        if code:
            raise MhdError(code, "mhd_ensemble_create")
        self._e = e
        self.dt = dt
        self.block = self._lib.mhd_ensemble_block()

    @property
    def nodes(self):
        return self._lib.mhd_ensemble_nodes(self._e)

    @property
    def classes(self):
        return self._lib.mhd_ensemble_classes(self._e)

    def step(self, nsteps):
        code = self._lib.mhd_ensemble_step(self._e, int(nsteps))
        if code:
            raise MhdError(code, "mhd_ensemble_step")

    def close(self):
        if self._e is not None:
            self._lib.mhd_ensemble_destroy(self._e)
            self._e = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()



CLOSURE_FIELDS = ("r", "t", "u", "v", "w", "bx", "by", "bz")
CLOSURE_OUTPUTS = ("tau_xx", "tau_yy", "tau_zz", "tau_xy", "tau_xz", "tau_yz",
                   "eps_x", "eps_y", "eps_z", "q_x", "q_y", "q_z")


def _ptr_array(arrs):
    P = ctypes.POINTER(ctypes.c_double)
    return (P * len(arrs))(*[_dp(a) for a in arrs])


def closure_targets(fine, factor, fmag):
    """Box-filtered fields and the 12 subgrid targets of a fine state (dict of
    (nz, ny, nx) arrays r t u v w bx by bz), computed on the GPU.
    Returns (coarse dict, (12, nz/f, ny/f, nx/f) array)."""
    arrs = [_f64(fine[n]) for n in CLOSURE_FIELDS]
    nz, ny, nx = arrs[0].shape
    f = int(factor)
    shp = (nz // f, ny // f, nx // f)
    coarse = [np.zeros(shp) for _ in range(8)]
    tgt = np.zeros((12,) + shp)
    tl = [tgt[o] for o in range(12)]
    code = _abi.load().mhd_closure_targets(nx, ny, nz, f, float(fmag), _ptr_array(arrs), _ptr_array(coarse),
                                           _ptr_array(tl))
    if code:
        raise MhdError(code, "mhd_closure_targets")
    return dict(zip(CLOSURE_FIELDS, coarse)), tgt


class Closure:
    """The learned subgrid closure: a 216-64-64-12 stencil MLP with bf16
    matrix-core kernels, trained on the GPU on box-filtered fine states."""

    def __init__(self, seed=0):
        self._lib = _abi.load()
        self._c = None
        c = ctypes.c_void_p()
        code = self._lib.mhd_closure_create(int(seed), ctypes.byref(c))
        if code:
            raise MhdError(code, "mhd_closure_create")
        self._c = c
        self.nparam = self._lib.mhd_closure_nparam()

    @staticmethod
    def _chk(code, what):
        if code:
            raise MhdError(code, what)

    def close(self):
        if self._c is not None:
            self._lib.mhd_closure_destroy(self._c)
            self._c = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def params(self):
        p = np.zeros(self.nparam, np.float32)
        self._chk(self._lib.mhd_closure_get_params(
            self._c, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), self.nparam), "get_params")
        return p

    @params.setter
    def params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        self._chk(self._lib.mhd_closure_set_params(
            self._c, p.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), len(p)), "set_params")

    @property
    def norm(self):
        n = np.zeros(40)
        self._chk(self._lib.mhd_closure_get_norm(self._c, _dp(n)), "get_norm")
        return n

    @norm.setter
    def norm(self, n):
        n = _f64(n)
        self._chk(self._lib.mhd_closure_set_norm(self._c, _dp(n)), "set_norm")

    def add(self, coarse, target):
        """Append one coarse snapshot: dict of fields and (12, ...) targets."""
        arrs = [_f64(coarse[k]) for k in CLOSURE_FIELDS]
        tl = [_f64(target[o]) for o in range(12)]
        cz, cy, cx = arrs[0].shape
        self._chk(self._lib.mhd_closure_add(self._c, cx, cy, cz, _ptr_array(arrs), _ptr_array(tl)), "add")

    def add_fine(self, fine, factor, fmag):
        """Box-filter a fine state and append it; returns (coarse, targets)."""
        coarse, tgt = closure_targets(fine, factor, fmag)
        self.add(coarse, tgt)
        return coarse, tgt

    def add_solver(self, solver, factor):
        """Filter a fine Solver's current state on the device and append it
        (no host round trip of the fields)."""
        self._chk(self._lib.mhd_closure_add_state(self._c, solver._ctx, int(factor)), "add_solver")

    @property
    def samples(self):
        n = ctypes.c_long(0)
        self._chk(self._lib.mhd_closure_samples(self._c, ctypes.byref(n)), "samples")
        return n.value

    def train(self, steps, batch=4096, lr=1e-3):
        loss = ctypes.c_double(0.0)
        self._chk(self._lib.mhd_closure_train(self._c, int(steps), int(batch), float(lr), ctypes.byref(loss)),
                  "train")
        return loss.value

    def grad(self, cells):
        c = np.ascontiguousarray(cells, dtype=np.int64)
        g = np.zeros(self.nparam, np.float32)
        loss = ctypes.c_double(0.0)
        self._chk(self._lib.mhd_closure_grad(self._c, len(c), c.ctypes.data_as(ctypes.POINTER(ctypes.c_long)),
                                             ctypes.byref(loss), g.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
                  "grad")
        return loss.value, g

    def eval(self):
        loss = ctypes.c_double(0.0)
        self._chk(self._lib.mhd_closure_eval(self._c, ctypes.byref(loss)), "eval")
        return loss.value

    def predict(self, coarse):
        arrs = [_f64(coarse[k]) for k in CLOSURE_FIELDS]
        nz, ny, nx = arrs[0].shape
        out = np.zeros((12, nz, ny, nx))
        ol = [out[o] for o in range(12)]
        self._chk(self._lib.mhd_closure_predict(self._c, nx, ny, nz, _ptr_array(arrs), _ptr_array(ol)), "predict")
        return out

    def time(self, batch=8192, reps=20):
        """(ms per forward pass over the data set, ms per training step, cells per forward pass)"""
        f, t = ctypes.c_double(0.0), ctypes.c_double(0.0)
        n = ctypes.c_long(0)
        self._chk(self._lib.mhd_closure_time(self._c, int(batch), int(reps), ctypes.byref(f), ctypes.byref(t),
                                             ctypes.byref(n)), "time")
        return f.value, t.value, n.value


def _attach_closure(self, closure):
    """Couple a trained Closure to this (coarse) run; None detaches."""
    self._closure = closure
    code = self._lib.mhd_closure_attach(self._ctx, closure._c if closure is not None else None)
    if code:
        raise MhdError(code, "mhd_closure_attach")


Solver.attach_closure = _attach_closure
__all__ += ["Closure", "closure_targets", "CLOSURE_FIELDS", "CLOSURE_OUTPUTS"]
