"""Write the reference's example initial conditions as JSON fixtures.

Reads the ASCII inputs of the reference's Swifter/Swiftest comparison
examples (central body, massive bodies with Gmass, rhill, radius, rh, vh, test
particles with rh, vh; the format of swiftest_io.f90:3279-3380) and the run
parameters the tests use, and writes tests/golden/<example>.json.  Run in a
container that has the reference tree at /root/reference; the JSON files are
committed and the tests read only them.
"""
import json
import os

REF = "/root/reference/examples/Swifter_Swiftest"
HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS = ("DT", "TSTOP", "ISTEP_OUT", "CHK_CLOSE", "RHILL_PRESENT", "GR", "GMTINY", "TU2S", "ENCOUNTER_CHECK")


def _lines(path):
    with open(path) as f:
        return [ln.split() for ln in f if ln.strip() and not ln.lstrip().startswith("!")]


def read_example(name, cb_file):
    d = os.path.join(REF, name)
    cb = _lines(os.path.join(d, cb_file))
    out = {"example": name, "cb": {"name": cb[0][0], "GM": float(cb[1][0]), "radius": float(cb[2][0])}}
    pl = _lines(os.path.join(d, "pl.swiftest.in"))
    n = int(pl[0][0])
    bodies = []
    for k in range(n):
        a, r, x, v = pl[1 + 4 * k: 5 + 4 * k]
        bodies.append({"name": a[0], "Gmass": float(a[1]), "rhill": float(a[2]), "radius": float(r[0]),
                       "rh": [float(t) for t in x], "vh": [float(t) for t in v]})
    out["pl"] = bodies
    tp = _lines(os.path.join(d, "tp.swiftest.in"))
    m = int(tp[0][0])
    parts = []
    for k in range(m):
        a, x, v = tp[1 + 3 * k: 4 + 3 * k]
        parts.append({"name": a[0], "rh": [float(t) for t in x], "vh": [float(t) for t in v]})
    out["tp"] = parts
    prm = {}
    for ln in _lines(os.path.join(d, "param.swiftest.in")):
        if ln[0] in PARAMS:
            prm[ln[0]] = ln[1]
    out["param"] = prm
    return out


def main():
    for name, cb in (("8pl_0tp", "cb.swiftest.in"), ("108pl_50tp", "cb.in")):
        data = read_example(name, cb)
        path = os.path.join(HERE, f"{name}.json")
        with open(path, "w") as f:
            json.dump(data, f, indent=1)
        print(path, len(data["pl"]), "pl", len(data["tp"]), "tp", data["param"])


if __name__ == "__main__":
    main()
