"""Extract the reference's SOL 105 CQUAD4 golden data into a small JSON fixture.

Reads (as text/JSON data only) the validation deck
validation/cases/TG_plate_uniaxial_buckling/model.bdf (a simply supported square
plate of 5 x 5 CQUAD4 under uniaxial edge compression: static preload subcase 1,
buckling subcase 2 with EIGB 1 INV 0.0 1.0E9 4) and its recorded results
results/timoshenko_plate_buckling.BUCKLING.JSON: the eigenvalues of the buckling
subcase, the mode shapes (NORM = MAX) and the static preload displacements, in
grid order.  Output: tests/golden/reference_buckling.json.
Run from the repository root: ``python tests/golden/make_buckling_fixture.py``.
"""
import json
import os

REF = "/root/reference/validation/cases/TG_plate_uniaxial_buckling"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_buckling.json")
COMP = ("t1", "t2", "t3", "r1", "r2", "r3")


def rows(entries):
    return {"grid_id": [int(e["grid_id"]) for e in entries],
            "values": [[float(e[c]) for c in COMP] for e in entries]}


def main():
    deck = open(os.path.join(REF, "model.bdf")).read()
    res = json.load(open(os.path.join(REF, "results", "timoshenko_plate_buckling.BUCKLING.JSON")))
    diag = res["solver_diagnostics"][0]
    data = {
        "case": "TG_plate_uniaxial_buckling",
        "source": "validation/cases/TG_plate_uniaxial_buckling (model.bdf, results/timoshenko_plate_buckling.BUCKLING.JSON)",
        "deck": deck,
        "eigenvalues": [float(v) for v in res["eigenvalues"]],
        "modes": [rows(m["mode_shape"]) for m in res["modes"]],
        "static_displacements": rows(res["static_displacements"]),
        "extraction": {k: diag.get(k) for k in ("eigenvalue_extraction_type", "eigenvalue_extraction_method",
                                                "buckling_subcase")},
        "free_dofs": int(diag["details"]["free_dofs"]),
        "fixed_dofs": int(diag["details"]["fixed_dofs"]),
    }
    json.dump(data, open(OUT, "w"), indent=1)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")


if __name__ == "__main__":
    main()
