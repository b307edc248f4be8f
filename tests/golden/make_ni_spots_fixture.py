"""Extract the reference's own C-output fixture rows into a small JSON vector.

The reference test fixtures ``tests/fixtures/sample.solutions.txt`` and
``tests/fixtures/sample.spots.txt`` were written by the reference C indexer
(``src/LaueMatchingHeaders.h:888-901`` for spots, ``:1130-1149`` for
solutions) with the geometry of ``tests/fixtures/params_NiIndent.txt``.  Each
spot row holds the hkl, the truncated pixel (X, Y) and the unit scattering
vector printed with ``%lf``; each solution row holds the 3x3 matrix
``Recip1..9`` (= orientation x reciprocal basis, row-major) used to compute
them.  This script reads those files as data (run it where the reference tree
is mounted) and writes ``ni_spots_golden.json`` beside itself.
"""
import json
import os
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
FIX = os.path.join(REF, "tests", "fixtures")


def main():
    params = {}
    for line in open(os.path.join(FIX, "params_NiIndent.txt")):
        parts = line.split()
        if len(parts) >= 2:
            params[parts[0]] = parts[1:]
    geom = {
        "P_Array": [float(x) for x in params["P_Array"][:3]],
        "R_Array": [float(x) for x in params["R_Array"][:3]],
        "PxX": float(params["PxX"][0]), "PxY": float(params["PxY"][0]),
        "NrPxX": int(params["NrPxX"][0]), "NrPxY": int(params["NrPxY"][0]),
        "Elo": float(params["Elo"][0]), "Ehi": float(params["Ehi"][0]),
        "LatticeParameter": [float(x) for x in params["LatticeParameter"][:6]],
    }
    grains = {}
    for line in open(os.path.join(FIX, "sample.solutions.txt")):
        if line.startswith("%"):
            continue
        f = line.split()
        grains[int(f[0])] = {"recip_rowmajor": [float(x) for x in f[7:16]],
                             "orient_rowmajor": [float(x) for x in f[22:31]],
                             "spots": []}
    for line in open(os.path.join(FIX, "sample.spots.txt")):
        if line.startswith("%"):
            continue
        f = line.split()
        g = int(f[0])
        grains[g]["spots"].append({
            "spot_nr": int(f[1]), "hkl": [int(f[2]), int(f[3]), int(f[4])],
            "X": int(f[5]), "Y": int(f[6]), "qhat": [float(x) for x in f[7:10]]})
    out = {"source": "reference tests/fixtures/{sample.solutions.txt,sample.spots.txt,"
                     "params_NiIndent.txt}",
           "geometry": geom, "grains": [grains[k] for k in sorted(grains)]}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ni_spots_golden.json")
    with open(dst, "w") as fh:
        json.dump(out, fh, indent=1)
    print("wrote", dst, sum(len(g["spots"]) for g in out["grains"]), "spots")


if __name__ == "__main__":
    main()
