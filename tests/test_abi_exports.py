"""The C-ABI library loads and exports every function include/mhd.h declares
(no device calls: runs without a GPU)."""
import ctypes
import os
import re
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mhd.h")
LIB = os.path.join(ROOT, "fug30v0kxc_amd", "libmhd_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(mhd_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_the_reference_routines():
    names = declared_functions()
    for n in ("mhd_create", "mhd_step", "mhd_hydro_deriv", "mhd_mag_deriv", "mhd_bandj",
              "mhd_boundary", "mhd_mag_boundary", "mhd_op", "mhd_destroy"):
        assert n in names


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build the library first (__graft_entry__.build())"
    ctypes.CDLL(LIB)   # loads without a GPU
    out = subprocess.check_output(["nm", "-D", "--defined-only", LIB], text=True)
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


def test_python_binding_declares_every_symbol():
    from fug30v0kxc_amd import _abi
    L = _abi.load()
    for n in declared_functions():
        assert getattr(L, n).argtypes is not None or n == "mhd_strerror", n


def test_error_strings_without_device():
    from fug30v0kxc_amd import _abi
    L = _abi.load()
    assert L.mhd_strerror(12) == b"div B is not zero"
    assert L.mhd_strerror(7) == b"unknown derivative flag"
