"""ctypes binding of the C ABI in include/mhd.h (libmhd_hip.so, gfx950).

Loading fails loudly when the HIP library has not been built: there is no
CPU fallback in the product path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmhd_hip.so")

# return codes (include/mhd.h)
MHD_OK = 0
MHD_E_ODD_NX = -1
MHD_E_ODD_NY = -2
MHD_E_ALLOC = -5
MHD_E_NZ = 1
MHD_E_FLAG = 7
MHD_E_DIVB = 12
MHD_E_ARG = -10
MHD_E_SIZE = -20
MHD_E_DEVICE = -30
MHD_E_STATE = -40
MHD_E_SCRATCH = -50
MHD_NOALLOC_SCRATCH_BYTES = 64 << 20

# MHD_OPT_* switches of mhd_params_t.opts
OPT_RADBC, OPT_CONSTFLUX, OPT_PERFECT, OPT_KINEMATIC, OPT_DYNAMO = 1, 2, 4, 8, 16

MHD_PREC_FP64 = 0
MHD_PREC_MIXED = 1

FIELD_IDS = {n: i for i, n in enumerate(
    ["u", "v", "w", "t", "r", "bx", "by", "bz", "pol", "tor", "jsq",
     "udot", "vdot", "wdot", "rdot", "tdot", "poldot", "tordot",
     "bxbar", "bybar", "bxbardot", "bybardot", "kappa", "dkappa"])}
VECTOR_FIELDS = {"bxbar", "bybar", "bxbardot", "bybardot", "kappa", "dkappa"}
HISTORY_FIELDS = {"udot", "vdot", "wdot", "rdot", "tdot", "poldot", "tordot",
                  "bxbardot", "bybardot"}

OPS = {"d1bydx": 1, "d1bydy": 2, "d2bydx": 3, "d2bydy": 4, "d1bydz": 5, "d1upbydz": 6,
       "d2bydz": 7, "dealias": 8, "forwardfft": 9, "inversefft": 10, "d1bydz1D": 11,
       "d2bydz1D": 12}

KERNELS = {"none": 0, "xderiv2": 1, "yderiv2": 2, "zderiv": 3, "hydro": 4, "ab3": 5}


class MhdParams(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_double) for n in
                ("gamma", "sigma", "theta", "ck", "cm", "f", "tayl", "psi", "zeta",
                 "xmax", "ymax", "sf", "t_amp", "b_amp", "eta_c", "eta_z0", "eta_w")]
                + [("opts", ctypes.c_int)])


class MhdScalars(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("dt1", ctypes.c_double), ("dt2", ctypes.c_double),
                ("ttime", ctypes.c_double), ("bzbar", ctypes.c_double),
                ("nit", ctypes.c_int), ("iter", ctypes.c_int), ("first", ctypes.c_int),
                ("second", ctypes.c_int)]


class MhdDiag(ctypes.Structure):
    _fields_ = [("ubar", ctypes.c_double), ("vbar", ctypes.c_double),
                ("kinenergy", ctypes.c_double), ("magenergy", ctypes.c_double),
                ("nusselt", ctypes.c_double), ("divbmax", ctypes.c_double),
                ("divbmin", ctypes.c_double), ("dt", ctypes.c_double), ("nit", ctypes.c_int)]


class MhdDump(ctypes.Structure):
    _fields_ = ([("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                 ("version", ctypes.c_int), ("iteration", ctypes.c_int),
                 ("time", ctypes.c_double), ("timestep", ctypes.c_double)]
                + [(n, ctypes.c_double) for n in ("gamma", "sigma", "theta", "ck", "cm", "f", "tayl",
                                                   "psi", "zeta", "xmax", "ymax")]
                + [("history", ctypes.c_char * 256)])


TUNE_KINDS = 4   # radix-plan choices per direction: d1, d1+d2, r2hc, hc2r (MHD_TK_*)


class MhdTune(ctypes.Structure):
    _fields_ = ([(n, ctypes.c_int * TUNE_KINDS) for n in ("variant_x", "variant_y", "mult_x", "mult_y")]
                + [("zc", ctypes.c_int)]
                + [(n, ctypes.c_double) for n in ("ms_x", "ms_y", "ms_col")])


_lib = None


def load():
    """Load libmhd_hip.so and declare every entry point of include/mhd.h."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build() "
                           "(make -C fug30v0kxc_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "mhd_ra_q_to_ck_f": (None, [d, d, d, d, d, d, d, dp, dp]),
        "mhd_strerror": (ctypes.c_char_p, [i]),
        "mhd_create": (i, [ctypes.POINTER(MhdParams), i, i, i, i, ctypes.POINTER(vp)]),
        "mhd_destroy": (None, [vp]),
        "mhd_static": (i, [vp]),
        "mhd_set_state": (i, [vp] + [dp] * 8),
        "mhd_get_state": (i, [vp] + [dp] * 8),
        "mhd_get_field": (i, [vp, i, i, dp]),
        "mhd_set_field": (i, [vp, i, i, dp]),
        "mhd_get_scalars": (i, [vp, ctypes.POINTER(MhdScalars)]),
        "mhd_set_scalars": (i, [vp, ctypes.POINTER(MhdScalars)]),
        "mhd_hydro_deriv": (i, [vp]),
        "mhd_mag_deriv": (i, [vp]),
        "mhd_bandj": (i, [vp]),
        "mhd_magdealias": (i, [vp]),
        "mhd_boundary": (i, [vp]),
        "mhd_mag_boundary": (i, [vp]),
        "mhd_mag_init": (i, [vp]),
        "mhd_step": (i, [vp, i, d, i, ctypes.POINTER(MhdDiag)]),
        "mhd_global_averages": (i, [vp, dp]),
        "mhd_nusselt": (i, [vp, dp]),
        "mhd_checkdivb": (i, [vp, dp]),
        "mhd_sminmax": (i, [vp, i, dp]),
        "mhd_calcdt": (i, [vp, dp]),
        "mhd_op": (i, [vp, i, dp, dp, dp, i]),
        "mhd_kernel_timing": (i, [vp, i, i]),
        "mhd_kernel_time": (i, [vp, dp, ctypes.POINTER(ctypes.c_long)]),
        "mhd_sync": (i, [vp]),
        "mhd_device_bytes": (ctypes.c_long, [vp]),
        "mhd_write_dump": (i, [ctypes.c_char_p, ctypes.POINTER(MhdDump), ctypes.POINTER(dp), i]),
        "mhd_read_dump": (i, [ctypes.c_char_p, ctypes.POINTER(MhdDump), ctypes.POINTER(dp)]),
        "mhd_autotune": (i, [vp, ctypes.POINTER(MhdTune)]),
        "mhd_set_tune": (i, [vp, ctypes.POINTER(MhdTune)]),
        "mhd_ensemble_create": (i, [ctypes.POINTER(vp), i, d, ctypes.POINTER(vp)]),
        "mhd_ensemble_step": (i, [vp, i]),
        "mhd_ensemble_block": (i, []),
        "mhd_ensemble_nodes": (ctypes.c_long, [vp]),
        "mhd_ensemble_classes": (i, [vp]),
        "mhd_ensemble_destroy": (None, [vp]),
        "mhd_expand": (i, [i, i, i, dp, dp]),
        "mhd_expand_noalloc": (i, [i, i, i, vp, vp, vp]),
        "mhd_dump_header": (ctypes.c_long, [ctypes.POINTER(MhdDump), i, ctypes.c_char_p, ctypes.c_long,
                                            ctypes.POINTER(i)]),
        "mhd_box_filter": (i, [i, i, i, i, dp, dp]),
        "mhd_closure_nparam": (i, []),
        "mhd_closure_create": (i, [ctypes.c_uint, ctypes.POINTER(vp)]),
        "mhd_closure_destroy": (None, [vp]),
        "mhd_closure_set_params": (i, [vp, fp, i]),
        "mhd_closure_get_params": (i, [vp, fp, i]),
        "mhd_closure_targets": (i, [i, i, i, i, d, ctypes.POINTER(dp), ctypes.POINTER(dp), ctypes.POINTER(dp)]),
        "mhd_closure_add": (i, [vp, i, i, i, ctypes.POINTER(dp), ctypes.POINTER(dp)]),
        "mhd_closure_add_state": (i, [vp, vp, i]),
        "mhd_closure_samples": (i, [vp, ctypes.POINTER(ctypes.c_long)]),
        "mhd_closure_get_norm": (i, [vp, dp]),
        "mhd_closure_set_norm": (i, [vp, dp]),
        "mhd_closure_train": (i, [vp, i, i, d, dp]),
        "mhd_closure_grad": (i, [vp, i, ctypes.POINTER(ctypes.c_long), dp, fp]),
        "mhd_closure_eval": (i, [vp, dp]),
        "mhd_closure_predict": (i, [vp, i, i, i, ctypes.POINTER(dp), ctypes.POINTER(dp)]),
        "mhd_closure_time": (i, [vp, i, i, dp, dp, ctypes.POINTER(ctypes.c_long)]),
        "mhd_closure_attach": (i, [vp, vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
