"""Shell result rows from recovered element fields, in the reference's schema.

The device computes per-element resultants and fibre fields
(``DeviceMesh.shell_recover``); this module turns them into the rows that
recover_shell_stresses! pushes into the results (src/solver/stress_recovery.jl:
1230-1264): ``forces`` (centre N, M, Q), ``forces_bilin`` (CQUAD4 only: the
centre row "CEN/4" and one row per corner grid), ``stresses`` (z1 = -T/2,
z2 = +T/2 with von Mises and principal values) and ``strains`` (z1: membrane
strain, fibre tag 0; z2: curvature in the printed sign, fibre tag -1).
"""
from __future__ import annotations

import math

import numpy as np

Q4_FIELDS = 44
T3_FIELDS = 20


def principal_2d(xx: float, yy: float, xy: float):
    """(major, minor) of a symmetric 2 x 2 tensor, the larger-magnitude root
    formed without cancellation and the other from the determinant identity
    (_recovery_principal_2d, src/solver/stress_recovery.jl:17-34)."""
    if xy == 0.0:
        return max(xx, yy), min(xx, yy)
    scale = max(abs(xx), abs(yy), abs(xy))
    a, b, c = xx / scale, yy / scale, xy / scale
    mean = (a + b) / 2
    radius = math.hypot((a - b) / 2, c)
    large = (mean + math.copysign(radius, mean)) * scale
    if not math.isfinite(large):
        return (mean + radius) * scale, (mean - radius) * scale
    hi, lo = (xx, yy) if abs(xx) >= abs(yy) else (yy, xx)
    small = (hi / large) * lo - (xy / large) * xy
    return max(large, small), min(large, small)


def von_mises(s) -> float:
    """sqrt(sx^2 - sx sy + sy^2 + 3 txy^2) (stress_recovery.jl:1244)."""
    return math.sqrt(s[0] ** 2 - s[0] * s[1] + s[1] ** 2 + 3 * s[2] ** 2)


def _stress_entry(s, t):
    major, minor = principal_2d(s[0], s[1], s[2])
    return {"fiber_dist": t, "normal_x": s[0], "normal_y": s[1], "shear_xy": s[2],
            "von_mises": von_mises(s), "major": major, "minor": minor}


def _strain_entry(e, t):
    # engineering shear in the row, tensor shear (half) in the principal values
    major, minor = principal_2d(e[0], e[1], e[2] / 2)
    return {"fiber_dist": t, "normal_x": e[0], "normal_y": e[1], "shear_xy": e[2], "major": major, "minor": minor}


def _force_row(eid, n, m, q, grid=None):
    row = {"eid": eid, "fx": n[0], "fy": n[1], "fxy": n[2], "mx": m[0], "my": m[1], "mxy": m[2],
           "qx": q[0], "qy": q[1]}
    if grid is not None:
        row["grid_id"] = grid
    return row


def shell_rows(raw: np.ndarray, eids, thickness, kind: str, corner_grids=None) -> dict:
    """Rows for the elements of one kind.  raw: (n, 44) for "quad4" or (n, 20)
    for "tria3" as returned by DeviceMesh.shell_recover; eids: element ids;
    thickness: T per element (the fibre distances and the curvature divisor);
    corner_grids: (n, 4) grid ids of each CQUAD4 (forces_bilin rows)."""
    out = {"forces": [], "stresses": [], "strains": [], "forces_bilin": []}
    raw = np.asarray(raw, dtype=float)
    for k, eid in enumerate(eids):
        r = [float(v) for v in raw[k]]
        T = float(thickness[k])
        N, M, Q = r[0:3], r[3:6], r[6:8]
        s1, s2, e1, e2 = r[8:11], r[11:14], r[14:17], r[17:20]
        out["forces"].append(_force_row(eid, N, M, Q))
        if kind == "quad4":
            out["forces_bilin"].append(_force_row(eid, N, M, Q, "CEN/4"))
            for c in range(4):
                out["forces_bilin"].append(_force_row(eid, r[20 + 3 * c:23 + 3 * c], r[32 + 3 * c:35 + 3 * c], Q,
                                                      int(corner_grids[k][c])))
        out["stresses"].append({"eid": eid, "z1": _stress_entry(s1, -T / 2), "z2": _stress_entry(s2, T / 2)})
        eps = [(a + b) / 2.0 for a, b in zip(e1, e2)]
        kap = [(a - b) / T for a, b in zip(e1, e2)]
        out["strains"].append({"eid": eid, "z1": _strain_entry(eps, 0.0), "z2": _strain_entry(kap, -1.0)})
    return out
