"""Writes tests/golden/restart_layout.json: the structure of the reference's
NetCDF restart fixture (bin/restart.nc, read with scipy's own netCDF-3
reader): dimensions, global attributes (names, types, values), variables
(names, dimensions, types) and per-field statistics.  Run in a container
where the reference tree is mounted; the JSON is the committed fixture."""
import json
import os
import sys

import numpy as np
from scipy.io import netcdf_file

SRC = "/root/reference/VarElecCond/bin/restart.nc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "restart_layout.json")


def main():
    f = netcdf_file(SRC, "r", mmap=False)
    attrs = {}
    for k, v in f._attributes.items():
        if isinstance(v, bytes):
            attrs[k] = {"type": "char", "value": v.decode(errors="replace")}
        else:
            a = np.atleast_1d(v)
            attrs[k] = {"type": str(a.dtype), "value": a.tolist()}
    variables = []
    for name, var in f.variables.items():
        data = np.array(var.data, dtype=np.float64)
        entry = {"name": name, "dims": list(var.dimensions), "typecode": var.typecode(),
                 "attrs": {k: (v.decode() if isinstance(v, bytes) else v)
                           for k, v in var._attributes.items()}}
        if data.ndim == 3:
            entry["stats"] = {"min": float(data.min()), "max": float(data.max()),
                              "mean": float(data.mean()),
                              "plane_means_first3": data.mean(axis=(1, 2))[:3].tolist()}
        else:
            entry["value"] = data.tolist()
        variables.append(entry)
    out = {"source": "VarElecCond/bin/restart.nc", "version_byte": int(f.version_byte),
           "dimensions": dict(f.dimensions), "global_attributes": attrs,
           "attribute_order": list(f._attributes.keys()), "variables": variables}
    json.dump(out, open(OUT, "w"), indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    sys.exit(main())
