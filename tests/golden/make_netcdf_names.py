"""Write the reference's NetCDF output names (dimensions and variables) as a
JSON fixture.

Reads the default names the reference's output module assigns
(netcdf_io_module.f90, the `<key>_dimname` / `<key>_varname` components) and
writes tests/golden/netcdf_names.json.  Run in a container that has the
reference tree at /root/reference; the JSON is committed and the tests read
only it.
"""
import json
import os
import re

SRC = "/root/reference/src/netcdf_io/netcdf_io_module.f90"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    text = open(SRC).read()
    pat = re.compile(r"::\s*(\w+)_(dimname|varname)\s*=\s*\"([^\"]+)\"")
    out = {"dimensions": {}, "variables": {}}
    for key, kind, name in pat.findall(text):
        out["dimensions" if kind == "dimname" else "variables"][key] = name
    with open(os.path.join(HERE, "netcdf_names.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(len(out["dimensions"]), "dimensions,", len(out["variables"]), "variables")


if __name__ == "__main__":
    main()
