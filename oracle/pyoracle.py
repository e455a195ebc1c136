"""ctypes wrapper around the CPU oracle (oracle/mhd_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg.  The product package never imports this module.

Fields are exposed as numpy views of shape (nz, ny, nx) in C order, which is
the reference's Fortran (nx, ny, nz) column-major layout.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

FIELDS = ["u", "v", "w", "t", "r", "bx", "by", "bz", "pol", "tor", "jsq", "ex", "ey", "ez",
          "wk1", "wk2", "wk3", "wk4", "udot", "vdot", "wdot", "tdot", "rdot", "poldot",
          "tordot", "bxbar", "bybar", "kappa", "dkappa", "bxbardot", "bybardot",
          "akx1", "akx2", "aky1", "aky2", "jx", "jy", "jz", "etar"]
FID = {n: i for i, n in enumerate(FIELDS)}
SCALARS = ["dt", "dt1", "dt2", "ttime", "bzbar", "nit", "iter", "first", "second",
           "gck", "sck", "zck", "dz", "xn", "xzn", "nusstmp"]
SID = {n: i for i, n in enumerate(SCALARS)}


# switches of orc_params.opts (src/control.f90:23-34)
OPT_RADBC, OPT_CONSTFLUX, OPT_PERFECT, OPT_KINEMATIC, OPT_DYNAMO = 1, 2, 4, 8, 16


class OrcParams(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in
                ("gamma", "sigma", "theta", "ck", "cm", "f", "tayl", "psi", "zeta",
                 "xmax", "ymax", "sf", "t_amp", "b_amp", "eta_c", "eta_z0", "eta_w")]
                + [("opts", ctypes.c_int)])


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        dp = ctypes.POINTER(ctypes.c_double)
        vp = ctypes.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int, ctypes.c_int,
                                 ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        L.orc_destroy.argtypes = [vp]
        L.orc_threads.restype = ctypes.c_int
        L.orc_threads.argtypes = []
        L.orc_set_threads.argtypes = [ctypes.c_int]
        ci = ctypes.c_int
        L.orc_expand.argtypes = [ci, ci, ci, dp, dp]
        L.orc_box_filter.argtypes = [ci, ci, ci, ci, dp, dp]
        pp = ctypes.POINTER(dp)
        L.orc_closure_targets.argtypes = [ci, ci, ci, ci, ctypes.c_double, pp, pp, pp]
        L.orc_zweight.argtypes = [ci, ci, ci, ci]
        L.orc_zweight.restype = ctypes.c_double
        L.orc_mlp_forward.argtypes = [ci, ci, ci, pp, dp, dp, ci, pp]
        L.orc_mlp_grad.argtypes = [ci, ci, ci, pp, pp, dp, dp, ci, ctypes.POINTER(ctypes.c_long), ci,
                                   dp, dp]
        L.orc_closure_attach.argtypes = [vp, dp, dp]
        L.orc_closure_fix_fluxes.argtypes = [vp, dp]
        L.orc_bf16.restype = ctypes.c_float
        L.orc_bf16.argtypes = [ctypes.c_float]
        L.orc_field.restype = dp
        L.orc_field.argtypes = [vp, ctypes.c_int]
        L.orc_scalar.restype = ctypes.c_double
        L.orc_scalar.argtypes = [vp, ctypes.c_int]
        L.orc_set_scalar.argtypes = [vp, ctypes.c_int, ctypes.c_double]
        L.orc_ra_q_to_ck_f.argtypes = [ctypes.c_double] * 7 + [dp, dp]
        for name in ("orc_r2hc", "orc_hc2r"):
            getattr(L, name).argtypes = [ctypes.c_int, dp, dp]
        for name in ("orc_d1bydx", "orc_d1bydy", "orc_forwardfft", "orc_inversefft"):
            getattr(L, name).argtypes = [vp, dp, dp]
        for name in ("orc_d2bydx", "orc_d2bydy"):
            getattr(L, name).argtypes = [vp, dp, dp, dp]
        L.orc_dealias.argtypes = [vp, dp]
        for name in ("orc_d1bydz", "orc_d2bydz", "orc_d1bydz1D", "orc_d2bydz1D"):
            getattr(L, name).argtypes = [vp, dp, dp, ctypes.c_int]
            getattr(L, name).restype = ctypes.c_int
        L.orc_d1upbydz.argtypes = [vp, dp, dp, ctypes.c_int, dp]
        L.orc_d1upbydz.restype = ctypes.c_int
        L.orc_ran0_fill.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_double]
        L.orc_ran0_seq.argtypes = [ctypes.c_int, ctypes.c_int, dp]
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_ab3_coeffs.argtypes = [ip, ip, ctypes.c_double, dp, dp, dp]
        for name in ("orc_static", "orc_ini_trans_params", "orc_magdealias", "orc_boundary",
                     "orc_mag_boundary"):
            getattr(L, name).argtypes = [vp]
        for name in ("orc_mag_init", "orc_hydro_deriv", "orc_mag_deriv", "orc_bandj"):
            getattr(L, name).argtypes = [vp]
            getattr(L, name).restype = ctypes.c_int
        L.orc_global_averages.argtypes = [vp, dp]
        L.orc_nusselt.argtypes = [vp]
        L.orc_nusselt.restype = ctypes.c_double
        L.orc_checkdivb.argtypes = [vp, dp]
        L.orc_checkdivb.restype = ctypes.c_int
        L.orc_sminmax.argtypes = [vp, dp, dp]
        L.orc_calcdt.argtypes = [vp]
        L.orc_calcdt.restype = ctypes.c_double
        L.orc_step.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, dp]
        L.orc_step.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q):
    ck = ctypes.c_double()
    f = ctypes.c_double()
    lib().orc_ra_q_to_ck_f(gamma, sigma, theta, cm, zeta, ra, q, ctypes.byref(ck),
                           ctypes.byref(f))
    return ck.value, f.value


def r2hc(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_r2hc(len(x), _p(x), _p(out))
    return out


def hc2r(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    lib().orc_hc2r(len(x), _p(x), _p(out))
    return out


def ran0_layer(layer_id, n, amp):
    out = np.empty(n)
    lib().orc_ran0_fill(layer_id, _p(out), n, amp)
    return out


def ran0_seq(idum, n):
    out = np.empty(n)
    lib().orc_ran0_seq(idum, n, _p(out))
    return out


def ab3_coeffs(first, second, dt, dt1, dt2):
    """Returns (c0, c1, c2, first, second, dt1, dt2) after the call."""
    f = ctypes.c_int(int(first))
    s = ctypes.c_int(int(second))
    a = ctypes.c_double(dt1)
    b = ctypes.c_double(dt2)
    c = np.zeros(3)
    lib().orc_ab3_coeffs(ctypes.byref(f), ctypes.byref(s), dt, ctypes.byref(a), ctypes.byref(b),
                         _p(c))
    return c[0], c[1], c[2], f.value, s.value, a.value, b.value


def threads():
    """OpenMP threads of the oracle's plane / point loops."""
    return lib().orc_threads()


def set_threads(n):
    lib().orc_set_threads(int(n))


def expand(f):
    """expand of a (nz, ny, nx) field -> (nz, 2ny, 2nx)."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, ny, nx = f.shape
    out = np.zeros((nz, 2 * ny, 2 * nx))
    e = lib().orc_expand(nx, ny, nz, _p(f), _p(out))
    if e:
        raise OracleError(e, "orc_expand")
    return out


def box_filter(f, factor):
    """factor-fold block average of a (nz, ny, nx) field."""
    f = np.ascontiguousarray(f, dtype=np.float64)
    nz, ny, nx = f.shape
    out = np.zeros((nz // factor, ny // factor, nx // factor))
    e = lib().orc_box_filter(nx, ny, nz, int(factor), _p(f), _p(out))
    if e:
        raise OracleError(e, "orc_box_filter")
    return out


class OracleError(RuntimeError):
    def __init__(self, code, what):
        super().__init__(f"{what} returned {code}")
        self.code = code


class Oracle:
    """One oracle context = one box of the reference solver (serial)."""

    def __init__(self, nx, ny, nz, params):
        L = lib()
        p = OrcParams(**params)
        ier = ctypes.c_int(0)
        self._c = L.orc_create(ctypes.byref(p), nx, ny, nz, ctypes.byref(ier))
        if not self._c:
            raise OracleError(ier.value, "orc_create")
        self.nx, self.ny, self.nz = nx, ny, nz
        self.params = dict(params)

    def __del__(self):
        if getattr(self, "_c", None):
            lib().orc_destroy(self._c)
            self._c = None

    def attach_closure(self, params, norm):
        """Couple a closure network (packed parameters, scaling) to the step; None detaches."""
        if params is None:
            lib().orc_closure_attach(self._c, None, None)
            return
        self._cl = (np.ascontiguousarray(params, dtype=np.float64), np.ascontiguousarray(norm, dtype=np.float64))
        e = lib().orc_closure_attach(self._c, _p(self._cl[0]), _p(self._cl[1]))
        if e:
            raise OracleError(e, "orc_closure_attach")

    def closure_fix_fluxes(self, flux):
        """Couple fixed closure fluxes (12, nz, ny, nx) to the step; None detaches."""
        if flux is None:
            lib().orc_closure_fix_fluxes(self._c, None)
            return
        f = np.ascontiguousarray(flux, dtype=np.float64)
        e = lib().orc_closure_fix_fluxes(self._c, _p(f))
        if e:
            raise OracleError(e, "orc_closure_fix_fluxes")

    # -- state access ----------------------------------------------------
    def field(self, name):
        L = lib()
        ptr = L.orc_field(self._c, FID[name])
        nx, ny, nz = self.nx, self.ny, self.nz
        if name in ("udot", "vdot", "wdot", "tdot", "rdot", "poldot", "tordot"):
            shape = (3, nz, ny, nx)
        elif name in ("bxbardot", "bybardot"):
            shape = (3, nz)
        elif name in ("bxbar", "bybar", "kappa", "dkappa", "etar"):
            shape = (nz,)
        elif name in ("akx1", "akx2"):
            shape = (nx,)
        elif name in ("aky1", "aky2"):
            shape = (ny,)
        else:
            shape = (nz, ny, nx)
        return np.ctypeslib.as_array(ptr, shape=shape)

    def __getitem__(self, name):
        return self.field(name)

    def scalar(self, name):
        return lib().orc_scalar(self._c, SID[name])

    def set_scalar(self, name, v):
        lib().orc_set_scalar(self._c, SID[name], float(v))

    def _new(self):
        return np.zeros((self.nz, self.ny, self.nx))

    # -- operators (src/fft.F90, src/dbydz.F90) ---------------------------
    def d1bydx(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = self._new()
        lib().orc_d1bydx(self._c, _p(f), _p(o))
        return o

    def d1bydy(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = self._new()
        lib().orc_d1bydy(self._c, _p(f), _p(o))
        return o

    def d2bydx(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        a, b = self._new(), self._new()
        lib().orc_d2bydx(self._c, _p(f), _p(a), _p(b))
        return a, b

    def d2bydy(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        a, b = self._new(), self._new()
        lib().orc_d2bydy(self._c, _p(f), _p(a), _p(b))
        return a, b

    def dealias(self, z):
        z = np.array(z, dtype=np.float64, order="C")
        lib().orc_dealias(self._c, _p(z))
        return z

    def forwardfft(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = self._new()
        lib().orc_forwardfft(self._c, _p(f), _p(o))
        return o

    def inversefft(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = self._new()
        lib().orc_inversefft(self._c, _p(f), _p(o))
        return o

    def _zop(self, name, f, flag, shape=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = np.zeros_like(f)
        e = getattr(lib(), name)(self._c, _p(f), _p(o), flag)
        if e:
            raise OracleError(e, name)
        return o

    def d1bydz(self, f, flag):
        return self._zop("orc_d1bydz", f, flag)

    def d2bydz(self, f, flag):
        return self._zop("orc_d2bydz", f, flag)

    def d1bydz1D(self, f, flag):
        return self._zop("orc_d1bydz1D", f, flag)

    def d2bydz1D(self, f, flag):
        return self._zop("orc_d2bydz1D", f, flag)

    def d1upbydz(self, f, flag, w=None):
        f = np.ascontiguousarray(f, dtype=np.float64)
        o = np.zeros_like(f)
        wp = _p(np.ascontiguousarray(w, dtype=np.float64)) if w is not None else None
        if w is not None:
            wa = np.ascontiguousarray(w, dtype=np.float64)
            wp = _p(wa)
        e = lib().orc_d1upbydz(self._c, _p(f), _p(o), flag, wp)
        if e:
            raise OracleError(e, "orc_d1upbydz")
        return o

    # -- physics ---------------------------------------------------------
    def static(self):
        lib().orc_static(self._c)

    def ini_trans_params(self):
        lib().orc_ini_trans_params(self._c)

    def _chk(self, name):
        e = getattr(lib(), name)(self._c)
        if e:
            raise OracleError(e, name)

    def mag_init(self):
        self._chk("orc_mag_init")

    def hydro_deriv(self):
        self._chk("orc_hydro_deriv")

    def mag_deriv(self):
        self._chk("orc_mag_deriv")

    def bandj(self):
        self._chk("orc_bandj")

    def magdealias(self):
        lib().orc_magdealias(self._c)

    def boundary(self):
        lib().orc_boundary(self._c)

    def mag_boundary(self):
        lib().orc_mag_boundary(self._c)

    def global_averages(self):
        out = np.zeros(4)
        lib().orc_global_averages(self._c, _p(out))
        return out

    def nusselt(self):
        return lib().orc_nusselt(self._c)

    def checkdivb(self):
        out = np.zeros(2)
        e = lib().orc_checkdivb(self._c, _p(out))
        if e:
            raise OracleError(e, "orc_checkdivb")
        return out

    def sminmax(self, z):
        z = np.ascontiguousarray(z, dtype=np.float64)
        out = np.zeros(2)
        lib().orc_sminmax(self._c, _p(z), _p(out))
        return out

    def calcdt(self):
        return lib().orc_calcdt(self._c)

    def step(self, nsteps, dt_fixed=0.0, ndu06=0):
        diag = np.full(7, np.nan)
        e = lib().orc_step(self._c, nsteps, dt_fixed, ndu06, _p(diag))
        if e:
            raise OracleError(e, "orc_step")
        return diag

    def setup_static(self):
        """static() + ini_trans_params() + mag_init(), src/main.F90:275-333."""
        self.static()
        self.ini_trans_params()
        # src/main.F90:320-333 (F as rescaled for the dynamo / kinematic problems)
        opts = self.params.get("opts", 0)
        f = 0.0 if opts & OPT_KINEMATIC else (1.0 if opts & OPT_DYNAMO else self.params["f"])
        if f != 0.0 or opts & OPT_KINEMATIC:
            self.mag_init()


# ---- learned subgrid closure (closure_oracle.c) ------------------------------
CLOSURE_FIELDS = ("r", "t", "u", "v", "w", "bx", "by", "bz")
CLOSURE_TARGETS = ("tau_xx", "tau_yy", "tau_zz", "tau_xy", "tau_xz", "tau_yz",
                   "eps_x", "eps_y", "eps_z", "q_x", "q_y", "q_z")
N_PARAM = 18828


def _ptrs(arrs):
    P = ctypes.POINTER(ctypes.c_double)
    return (P * len(arrs))(*[_p(a) for a in arrs])


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def closure_targets(fine, f, fmag):
    """Box-filtered fields and the 12 subgrid targets of a fine state (dict of
    (nz, ny, nx) arrays) -> (coarse dict, (12, cz, cy, cx) array)."""
    arrs = [np.ascontiguousarray(fine[n], dtype=np.float64) for n in CLOSURE_FIELDS]
    nz, ny, nx = arrs[0].shape
    shp = (nz // f, ny // f, nx // f)
    coarse = [np.zeros(shp) for _ in range(8)]
    tgt = np.zeros((12,) + shp)
    tl = [tgt[o] for o in range(12)]
    e = lib().orc_closure_targets(nx, ny, nz, int(f), ctypes.c_double(fmag), _ptrs(arrs), _ptrs(coarse),
                                  _ptrs(tl))
    if e:
        raise OracleError(e, "orc_closure_targets")
    return dict(zip(CLOSURE_FIELDS, coarse)), tgt


def mlp_forward(coarse, params, norm, round_bf16=True):
    """Closure prediction (12, nz, ny, nx) from coarse fields."""
    arrs = [np.ascontiguousarray(coarse[n], dtype=np.float64) for n in CLOSURE_FIELDS]
    nz, ny, nx = arrs[0].shape
    out = np.zeros((12, nz, ny, nx))
    ol = [out[o] for o in range(12)]
    p = np.ascontiguousarray(params, dtype=np.float64)
    nm = np.ascontiguousarray(norm, dtype=np.float64)
    lib().orc_mlp_forward(nx, ny, nz, _ptrs(arrs), _p(p), _p(nm), int(bool(round_bf16)), _ptrs(ol))
    return out


def mlp_grad(snaps, params, norm, cells, round_bf16=True):
    """Loss and gradient over cells (snapshot * n3 + cell) of snapshots given
    as a list of (coarse dict, targets (12, ...)) pairs."""
    arrs, tl = [], []
    for coarse, tgt in snaps:
        arrs += [np.ascontiguousarray(coarse[n], dtype=np.float64) for n in CLOSURE_FIELDS]
        tl += [np.ascontiguousarray(tgt[o], dtype=np.float64) for o in range(12)]
    nz, ny, nx = arrs[0].shape
    p = np.ascontiguousarray(params, dtype=np.float64)
    nm = np.ascontiguousarray(norm, dtype=np.float64)
    c = np.ascontiguousarray(cells, dtype=np.int64)
    loss = ctypes.c_double(0.0)
    g = np.zeros(N_PARAM)
    lib().orc_mlp_grad(nx, ny, nz, _ptrs(arrs), _ptrs(tl), _p(nm), _p(p), len(c),
                       c.ctypes.data_as(ctypes.POINTER(ctypes.c_long)), int(bool(round_bf16)),
                       ctypes.byref(loss), _p(g))
    return loss.value, g
