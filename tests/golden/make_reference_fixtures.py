"""Writes tests/golden/reference_observables.json: the exact-diagonalisation observables that the
reference's regression tests compare their DMRG runs against, read as data (h5dump text output)
from the reference's HDF5 files:

* regression/applications/dmrg/ref/test_fermiladder.diag.h5: fermion Hubbard on the open ladder,
  L=6 rungs (12 sites), t = t' = 1, U = 8, 10 up + 10 down electrons (the model of
  test_coded_fermiladder_ts.py:38-92): energy, local densities up / down and the one-body density
  matrices up / down (12 x 12, site = 2 * x + y for the label "( x,y )").
* regression/applications/dmrg/ref/test_ff.diag.h5: spinless free fermions on the open 5 x 5 square
  lattice, t = 1, N = 10 (test_ff_ts.py): energy.

Usage: python tests/golden/make_reference_fixtures.py  (needs /root/reference and h5dump).
"""
import json
import os
import re
import subprocess

REF = "/root/reference/dmrg/regression/applications/dmrg/ref"
H5DUMP = "/opt/conda/bin/h5dump"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_observables.json")


def dump(path, dataset):
    txt = subprocess.run([H5DUMP, "-m", "%.17g", "-w", "0", "-d", dataset, path], capture_output=True, text=True,
                         check=True).stdout
    body = txt[txt.index("DATA {") + 6:txt.rindex("}")]
    body = body[:body.rindex("}")] if body.count("}") else body
    return body


def numbers(path, dataset):
    body = dump(path, dataset)
    body = re.sub(r"\(\d+(,\d+)*\):", " ", body)
    return [float(x) for x in re.findall(r"[-+]?\d[\d.]*(?:[eE][-+]?\d+)?", body)]


def strings(path, dataset):
    return re.findall(r'"([^"]*)"', dump(path, dataset))


def site_of(label):
    x, y = (int(v) for v in re.findall(r"-?\d+", label))
    return 2 * x + y


def matrix(path, name):
    vals = numbers(path, "/spectrum/sectors/0/results/%s/mean/value" % name)
    labels = strings(path, "/spectrum/sectors/0/results/%s/labels" % name)
    assert len(vals) == len(labels) == 144, (len(vals), len(labels))
    M = [[None] * 12 for _ in range(12)]
    for v, lab in zip(vals, labels):
        a, b = lab.split("--")
        M[site_of(a)][site_of(b)] = v
    return M


def main():
    lad = os.path.join(REF, "test_fermiladder.diag.h5")
    ff = os.path.join(REF, "test_ff.diag.h5")
    out = {
        "fermiladder": {
            "model": "fermion Hubbard, open ladder, L=6 rungs (site = 2 * rung + leg), t=1, t'=1, U=8, "
                     "Nup=10, Ndown=10",
            "source": "regression/applications/dmrg/ref/test_fermiladder.diag.h5 (exact diagonalisation)",
            "energy": numbers(lad, "/spectrum/sectors/0/energies")[0],
            "local_density_up": numbers(lad, "/spectrum/sectors/0/results/Local density up/mean/value"),
            "local_density_down": numbers(lad, "/spectrum/sectors/0/results/Local density down/mean/value"),
            "onebody_up": matrix(lad, "Onebody density matrix up"),
            "onebody_down": matrix(lad, "Onebody density matrix down"),
            "dmrg_test_energy": 55.08693016528693,
        },
        "free_fermions": {
            "model": "spinless fermions, open 5 x 5 square lattice, t=1, N=10",
            "source": "regression/applications/dmrg/ref/test_ff.diag.h5 (exact diagonalisation)",
            "energy": numbers(ff, "/spectrum/sectors/0/energies")[0],
        },
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", OUT)


if __name__ == "__main__":
    main()
