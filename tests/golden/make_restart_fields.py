"""Writes tests/golden/restart_fields.npz: the eight fields, the three scalar
variables and the global attributes of the reference's evolved-state
restart fixture (bin/restart.nc: 30x30x60, gamma 1.4, Ra 1e5, Q 2500, at
iteration 105000), read with scipy's netCDF-3 reader.  Run where the
reference tree is mounted; the .npz (plain arrays, no pickles) is the
committed fixture that drives the GPU step from the reference's own state."""
import os

import numpy as np
from scipy.io import netcdf_file

SRC = "/root/reference/VarElecCond/bin/restart.nc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "restart_fields.npz")
FIELDS = ("R", "T", "U", "V", "W", "Bx", "By", "Bz")
ATTRS = ("gamma", "sigma", "theta", "ck", "cm", "f", "tayl", "psi", "zeta", "xmax", "ymax")


def main():
    f = netcdf_file(SRC, "r", mmap=False)
    out = {n: np.array(f.variables[n].data, dtype=np.float64) for n in FIELDS}
    for n in ("time", "timestep"):
        out[n] = np.array(f.variables[n].data, dtype=np.float64)
    out["iteration"] = np.array(f.variables["iteration"].data, dtype=np.int64)
    out["attrs"] = np.array([float(np.atleast_1d(f._attributes[a])[0]) for a in ATTRS])
    out["attr_names"] = np.array(ATTRS)
    f.close()
    np.savez_compressed(OUT, **out)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
