"""Extract the reference's own SOL 101 golden data into small JSON fixtures.

Reads (as text/JSON data only) the flat-shell validation decks of the reference
and their recorded displacement vectors ``results/*.JU.JSON``:

* validation/cases/MYSTRAN_sol101_simple  (1 CQUAD4, FORCE and PLOAD4 subcases)
* validation/cases/MH_curved_beam_in_plane (6 tapered CQUAD4, in-plane FORCE)
* validation/cases/MH_curved_beam_refined  (96 CQUAD4, in-plane FORCE)
* validation/cases/MH_twisted_beam_{in_plane,out_of_plane} (warped CQUAD4, FORCE)
* validation/cases/MH_pinched_cylinder{,_refined}, MH_hemispherical_shell
  (curved shells of warped CQUAD4, FORCE, symmetry SPC1)
* validation/cases/MH_hemispherical_shell/convergence/{N08,N16} (the same
  hemisphere on 8x8 and 16x16 CQUAD4 meshes)
* validation/cases/MH_scordelis_lo{,/convergence/N08,/convergence/N16}
  (the Scordelis-Lo roof under self-weight: GRAV on shell mass)

Where a result file also records CQUAD4 shell recovery (the MYSTRAN case:
forces, forces_bilin, stresses, strains of its element), each subcase carries
those rows under "shell_recovery", and the case its solution sequence ("sol",
the executive SOL number, which selects the recovery branch).

The TACS-backend uCRM-9 vector is not on this path and is not extracted.  Output: tests/golden/reference_ju.json.
Run from the repository root: ``python tests/golden/make_fixtures.py``.
"""
import json
import os
import sys

REF = "/root/reference/validation/cases"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_ju.json")

CASES = [
    ("MYSTRAN_sol101_simple", "results/sol101_simple.JU.JSON", {1: 1992, 3: 1994}, 101),
    ("MH_curved_beam_in_plane", "results/curved_beam.JU.JSON", {1: 1}, 1),
    ("MH_curved_beam_refined", "results/curved_beam_refined.JU.JSON", {1: 1}, 1),
    ("MH_twisted_beam_in_plane", "results/twisted_beam.JU.JSON", {1: 1}, 1),
    ("MH_twisted_beam_out_of_plane", "results/twisted_beam_out_of_plane.JU.JSON", {1: 1}, 1),
    ("MH_pinched_cylinder", "results/pinched_cylinder.JU.JSON", {1: 1}, 1),
    ("MH_pinched_cylinder_refined", "results/pinched_cylinder_refined.JU.JSON", {1: 1}, 1),
    ("MH_hemispherical_shell", "results/hemispherical_shell.JU.JSON", {1: 1}, 1),
    ("MH_hemispherical_shell_N08", "results/MH_hemispherical_shell_N08.JU.JSON", {1: 1}, 1),
    ("MH_hemispherical_shell_N16", "results/MH_hemispherical_shell_N16.JU.JSON", {1: 1}, 1),
    ("MH_scordelis_lo", "results/scordelis_lo.JU.JSON", {1: 1}, 1),
    ("MH_scordelis_lo_N08", "results/MH_scordelis_lo_N08.JU.JSON", {1: 1}, 1),
    ("MH_scordelis_lo_N16", "results/MH_scordelis_lo_N16.JU.JSON", {1: 1}, 1),
]
# deck directory under validation/cases when it is not the case name
CASE_DIR = {"MH_hemispherical_shell_N08": "MH_hemispherical_shell/convergence/N08",
            "MH_hemispherical_shell_N16": "MH_hemispherical_shell/convergence/N16",
            "MH_scordelis_lo_N08": "MH_scordelis_lo/convergence/N08",
            "MH_scordelis_lo_N16": "MH_scordelis_lo/convergence/N16"}


def num(s):
    s = s.strip()
    if not s:
        return None
    try:
        return float(s)
    except ValueError:
        pass
    # Nastran exponent shorthand: 1.0+7, 10.-3
    body = s
    for i in range(len(body) - 1, 0, -1):
        if body[i] in "+-" and body[i - 1] not in "eE":
            return float(body[:i] + "e" + body[i:])
    raise ValueError(s)


def cards(path):
    """Bulk cards as lists of string fields (free or small fixed field, '+' continuations)."""
    lines = open(path).read().splitlines()
    out = []
    in_bulk = False
    for raw in lines:
        if raw.strip().upper().startswith("BEGIN BULK"):
            in_bulk = True
            continue
        if not in_bulk or raw.startswith("$") or not raw.strip():
            continue
        if raw.strip().upper().startswith("ENDDATA"):
            break
        line = raw.rstrip()
        if "," in line:
            f = [x.strip() for x in line.split(",")]
        else:
            f = [line[i:i + 8].strip() for i in range(0, min(len(line), 72), 8)]
        if f[0].startswith("+") and out:
            out[-1].extend(f[1:])
        else:
            out.append(f)
    return out


def parse(case, loads_by_subcase, spc_sid):
    cdir = CASE_DIR.get(case, case)
    deck = os.path.join(REF, cdir, "model.bdf")
    grids, quads, pshell, mat1, spc, forces, pload4, grav = {}, [], {}, {}, [], {}, {}, {}
    k6rot = 100.0
    for f in cards(deck):
        name = f[0].upper()
        g = lambda i: f[i] if i < len(f) else ""
        if name == "PARAM" and g(1).upper() == "K6ROT":
            k6rot = num(g(2))
        elif name == "GRID":
            grids[int(g(1))] = [num(g(3)) or 0.0, num(g(4)) or 0.0, num(g(5)) or 0.0]
        elif name == "CQUAD4":
            quads.append({"eid": int(g(1)), "pid": int(g(2)), "nodes": [int(g(k)) for k in range(3, 7)]})
        elif name == "PSHELL":
            pshell[int(g(1))] = {"mid1": int(g(2)), "t": num(g(3)),
                                 "mid2": int(g(4)) if g(4) else 0,
                                 "bend_ratio": num(g(5)) if g(5) else 1.0,
                                 "mid3": int(g(6)) if g(6) else 0,
                                 "ts_t": num(g(7)) if g(7) else 5.0 / 6.0,
                                 "nsm": num(g(8)) or 0.0}
        elif name == "MAT1":
            mat1[int(g(1))] = {"E": num(g(2)), "G": num(g(3)) or 0.0, "nu": num(g(4)) or 0.0,
                               "rho": num(g(5)) or 0.0}
        elif name == "GRAV":
            if int(g(2) or 0) != 0:
                raise ValueError("GRAV with a coordinate system")
            grav.setdefault(int(g(1)), []).append({"a": num(g(3)) or 0.0,
                                                   "n": [num(g(4)) or 0.0, num(g(5)) or 0.0, num(g(6)) or 0.0]})
        elif name == "SPC1":
            sid = int(g(1))
            comps = [int(c) for c in g(2)]
            ids = [x for x in f[3:] if x]
            k = 0
            while k < len(ids):
                if k + 2 < len(ids) and ids[k + 1].upper() == "THRU":
                    for n in range(int(ids[k]), int(ids[k + 2]) + 1):
                        spc.append((sid, n, comps))
                    k += 3
                else:
                    spc.append((sid, int(ids[k]), comps))
                    k += 1
        elif name == "FORCE":
            sid = int(g(1))
            forces.setdefault(sid, []).append({"grid": int(g(2)), "mag": num(g(4)),
                                               "dir": [num(g(5)) or 0.0, num(g(6)) or 0.0, num(g(7)) or 0.0]})
        elif name == "PLOAD4":
            pload4.setdefault(int(g(1)), []).append({"eid": int(g(2)), "p": num(g(3))})
    ju = json.load(open(os.path.join(REF, cdir, CASE_JU[case])))
    disp = ju["displacements"]
    n_sub = len(ju["solver_diagnostics"])
    per = len(disp) // n_sub
    subcases = []
    for k, diag in enumerate(ju["solver_diagnostics"]):
        sid = int(diag.get("sid", k + 1))
        if sid not in loads_by_subcase:
            continue
        lsid = loads_by_subcase[sid]
        rows = disp[k * per:(k + 1) * per]
        sub = {
            "subcase": sid, "load_sid": lsid,
            "force": forces.get(lsid, []), "pload4": pload4.get(lsid, []), "grav": grav.get(lsid, []),
            "expected": {str(r["grid_id"]): [r["t1"], r["t2"], r["t3"], r["r1"], r["r2"], r["r3"]] for r in rows},
            "reference_free_dofs": diag["details"]["bc_partition"]["free_dofs"],
            "reference_relative_residual": diag["details"]["linear_solver"]["relative_residual"],
        }
        nq = len(quads)
        if nq and len(ju.get("stresses", {}).get("quad4", [])) == n_sub * nq:
            sub["shell_recovery"] = {
                "forces": ju["forces"]["quad4"][k * nq:(k + 1) * nq],
                "forces_bilin": ju["forces_bilin"]["quad4"][5 * k * nq:5 * (k + 1) * nq],
                "stresses": ju["stresses"]["quad4"][k * nq:(k + 1) * nq],
                "strains": ju["strains"]["quad4"][k * nq:(k + 1) * nq],
            }
        subcases.append(sub)
    sol = 101
    for line in open(os.path.join(REF, cdir, "model.bdf")).read().splitlines():
        t = line.split("$")[0].strip().upper()
        if t.startswith("CEND") or t.startswith("BEGIN"):
            break
        if t.startswith("SOL ") and t.split()[1].isdigit():
            sol = int(t.split()[1])
    return {"case": case, "dir": cdir, "sol": sol, "k6rot": k6rot, "grids": {str(k): v for k, v in grids.items()},
            "cquad4": quads, "pshell": {str(k): v for k, v in pshell.items()},
            "mat1": {str(k): v for k, v in mat1.items()},
            "spc1": [{"grid": n, "comps": c} for (s, n, c) in spc if s == spc_sid],
            "subcases": subcases,
            "provenance": f"validation/cases/{cdir}/model.bdf + {CASE_JU[case]}"}


CASE_JU = {c: j for c, j, _, _ in CASES}

if __name__ == "__main__":
    if not os.path.isdir(REF):
        sys.exit("reference validation cases not present")
    data = [parse(c, l, s) for c, _, l, s in CASES]
    json.dump(data, open(OUT, "w"), indent=1)
    print("wrote", OUT, os.path.getsize(OUT), "bytes")
