"""Shared test helpers: golden-deck models and tolerances."""
import json
import os

import numpy as np

from jf_2026_05_25_amd import meshes

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_ju.json")


def golden_cases():
    with open(GOLDEN) as f:
        return json.load(f)


def golden_model(case: dict, sub: dict, membrane_modes: int = 15) -> meshes.Model:
    """Model for one subcase of a golden deck: GRIDs in ascending ID order
    (src/solver/assembly.jl:2653-2657), PSHELL/MAT1 sections, SPC1, FORCE/PLOAD4
    and GRAV (lumped shell self-weight)."""
    ids = sorted(int(k) for k in case["grids"])
    idx = {g: i for i, g in enumerate(ids)}
    xyz = np.array([case["grids"][str(g)] for g in ids], dtype=float)
    quads = np.array([[idx[n] for n in q["nodes"]] for q in case["cquad4"]], dtype=np.int32)
    props = np.zeros((len(quads), meshes.SP_NCOL))
    for e, q in enumerate(case["cquad4"]):
        ps = case["pshell"][str(q["pid"])]
        mat = case["mat1"][str(ps["mid1"])]
        br = ps["bend_ratio"] if ps["mid2"] else 0.0
        props[e] = meshes.pshell_props(1, ps["t"], mat["E"], mat["nu"], mat["G"], br, ps["ts_t"],
                                       mid3_blank=(ps["mid3"] == 0), membrane_modes=membrane_modes)[0]
    ndof = 6 * len(ids)
    spc = np.zeros(ndof, dtype=np.uint8)
    for s in case["spc1"]:
        for c in s["comps"]:
            spc[6 * idx[s["grid"]] + c - 1] = 1
    F = np.zeros(ndof)
    for f in sub["force"]:
        for c in range(3):
            F[6 * idx[f["grid"]] + c] += f["dir"][c] * f["mag"]
    eid_row = {q["eid"]: e for e, q in enumerate(case["cquad4"])}
    for p in sub["pload4"]:
        e = eid_row[p["eid"]]
        F += meshes.pload4_forces(xyz, quads[e:e + 1], p["p"])
    for gcard in sub.get("grav", []):
        # lumped shell self-weight: (rho T + NSM) * area / 4 per corner, area half the
        # norm of the diagonal cross product (src/solver/loads.jl:758-772)
        g = gcard["a"] * np.asarray(gcard["n"], float)
        for e, q in enumerate(case["cquad4"]):
            ps = case["pshell"][str(q["pid"])]
            rho = case["mat1"][str(ps["mid1"])].get("rho", 0.0)
            nsm = ps.get("nsm", 0.0)
            if rho <= 0 and nsm <= 0:
                continue
            X = xyz[quads[e]]
            area = 0.5 * np.linalg.norm(np.cross(X[2] - X[0], X[3] - X[1]))
            share = (rho * ps["t"] + nsm) * area / 4
            for n in quads[e]:
                F[6 * n:6 * n + 3] += share * g
    expected = np.zeros(ndof)
    for g, v in sub["expected"].items():
        expected[6 * idx[int(g)]:6 * idx[int(g)] + 6] = v
    m = meshes.Model(case["case"], xyz, quads, props, np.zeros((0, 3), np.int32),
                     np.zeros((0, 8)), np.zeros((0, 10), np.int32), np.zeros((0, 2)),
                     spc, F[:, None], case["k6rot"], {"expected": expected})
    return m


def golden_grav(case: dict, sub: dict):
    """The subcase's GRAV cards as LoadSet / oracle.load_vector arguments:
    {"grav": (A, N), "shell_areal": rho T + NSM per CQUAD4} (empty without GRAV)."""
    gr = sub.get("grav", [])
    if not gr:
        return {}
    areal = []
    for q in case["cquad4"]:
        ps = case["pshell"][str(q["pid"])]
        rho = case["mat1"][str(ps["mid1"])].get("rho", 0.0)
        nsm = ps.get("nsm", 0.0)
        areal.append(0.0 if (rho <= 0 and nsm <= 0) else rho * ps["t"] + nsm)
    return {"grav": (np.array([g["a"] for g in gr]), np.array([g["n"] for g in gr], float)),
            "shell_areal": np.array(areal)}


def flap_model(n: int = 4, k6rot: float = 0.01) -> meshes.Model:
    """A near-mechanism for the post-factor pivot check: plate A on [0,1]^2
    (n x n CQUAD4, edge x = 0 clamped) and plate B on [1,2] x [1,2] joined to A at
    the single grid point (1,1).  B's in-plane rotation about that point is
    resisted only by drilling stiffness (proportional to K6ROT), so with a small
    K6ROT one pivot ratio is tiny (about 8e-8 K6ROT under some orders) and the next
    smallest is about 2e-4: the check flags one DOF, which one depending on the
    elimination order."""
    xs = np.linspace(0.0, 1.0, n + 1)
    pts = [(x, y, 0.0) for y in xs for x in xs]
    ia = {(i, j): j * (n + 1) + i for j in range(n + 1) for i in range(n + 1)}
    ib = {}
    for j in range(n + 1):
        for i in range(n + 1):
            if i == 0 and j == 0:
                ib[(i, j)] = ia[(n, n)]
            else:
                ib[(i, j)] = len(pts)
                pts.append((1.0 + xs[i], 1.0 + xs[j], 0.0))
    quads = [[idx[(i, j)], idx[(i + 1, j)], idx[(i + 1, j + 1)], idx[(i, j + 1)]]
             for idx in (ia, ib) for j in range(n) for i in range(n)]
    xyz = np.array(pts)
    quads = np.array(quads, np.int32)
    props = meshes.pshell_props(len(quads), 0.002, 70e9, 0.33)
    ndof = 6 * len(xyz)
    spc = np.zeros(ndof, np.uint8)
    for j in range(n + 1):
        spc[6 * ia[(0, j)]:6 * ia[(0, j)] + 6] = 1
    F = np.zeros((ndof, 2))
    F[2::6, 0] = 1.0
    F[0::6, 1] = np.linspace(-1.0, 1.0, len(xyz))
    return meshes.Model("flap", xyz, quads, props, np.zeros((0, 3), np.int32), np.zeros((0, 8)),
                        np.zeros((0, 10), np.int32), np.zeros((0, 2)), spc, F, k6rot)
