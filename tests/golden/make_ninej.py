"""Write ninej_sympy.json: exact Wigner 9j symbols (sympy) over every argument set with doubled
spins a..i in the ranges the SU2U1 couplings of a two-site QC sweep reach (bond spins up to 2S=8,
operator / site spins up to 2).  Inputs and outputs only; rerun with `python make_ninej.py`."""
import itertools
import json
import os

from sympy import Rational
from sympy.physics.wigner import wigner_9j


def tri(a, b, c):
    return abs(a - b) <= c <= a + b and (a + b + c) % 2 == 0


def main():
    rows = []
    big = range(0, 9)      # bond spins (2S)
    small = range(0, 3)    # site / operator spins (2S)
    for a, c, g, i in itertools.product(big, big, big, big):
        for b, h in itertools.product(small, small):
            for d, e, f in itertools.product(small, small, small):
                if not (tri(a, b, c) and tri(g, h, i) and tri(d, e, f) and tri(a, d, g) and tri(b, e, h) and tri(c, f, i)):
                    continue
                v = wigner_9j(*[Rational(x, 2) for x in (a, b, c, d, e, f, g, h, i)], prec=None)
                rows.append([a, b, c, d, e, f, g, h, i, float(v)])
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ninej_sympy.json")
    with open(out, "w") as fh:
        json.dump({"source": "sympy.physics.wigner.wigner_9j (exact rational arithmetic)", "args": "doubled spins a..i",
                   "rows": rows}, fh)
    print(len(rows), "rows ->", out)


if __name__ == "__main__":
    main()
