"""Extract the 6-31G** Cartesian basis for H, C, N and O into a JSON data table.

The source is LIO's plain-text basis file (`dat/basis/6-31G**`, an EMSL Basis
Set Exchange export: Hehre, Ditchfield & Pople 1972; Hariharan & Pople 1973).
Its layout (parsed by `lioamber/basis_data.f90:790-822`) is, per element:

    gaussian
     Z  n_primitives  n_shells
     <primitives per shell ...>
     <angular momentum per shell ...>
     exponent coefficient          (n_primitives lines)

Only the published exponents and raw contraction coefficients are kept; the
primitive normalisation and function ordering are applied by
`lio_amd.systems`, which is the caller-side builder the tests and the
benchmark use.  Run once in a container that holds the LIO sources:

    python tests/golden/make_basis_table.py /root/reference/dat/basis/6-31G**
"""
import json
import os
import sys

KEEP = (1, 6, 7, 8)


def parse(path):
    lines = open(path).read().splitlines()
    out = {}
    i = 0
    while i < len(lines):
        if lines[i].strip() != "gaussian":
            i += 1
            continue
        z, nraw, ncon = map(int, lines[i + 1].split())
        nprim = list(map(int, lines[i + 2].split()))[:ncon]
        lval = list(map(int, lines[i + 3].split()))[:ncon]
        prims = [tuple(map(float, lines[i + 4 + k].split()[:2])) for k in range(nraw)]
        i += 4 + nraw
        if z not in KEEP:
            continue
        shells, pos = [], 0
        for n, l in zip(nprim, lval):
            shells.append({"l": l, "exp": [p[0] for p in prims[pos:pos + n]],
                           "coef": [p[1] for p in prims[pos:pos + n]]})
            pos += n
        out[str(z)] = shells
    return out


if __name__ == "__main__":
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/dat/basis/6-31G**"
    table = {"name": "6-31G**", "cartesian": True, "source": "EMSL BSE via LIO dat/basis/6-31G**",
             "elements": parse(src)}
    dst = os.path.join(os.path.dirname(__file__), "..", "..", "lio_amd", "data", "basis_6-31Gss.json")
    with open(dst, "w") as f:
        json.dump(table, f, indent=1)
    print("wrote", os.path.normpath(dst))
