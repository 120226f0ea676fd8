"""Deck text to JU.JSON results on the GPU: the golden decks (written in
small-field form from the fixtures) read by jf_2026_05_25_amd.deck, every
subcase through the device path (element kernels, assembly, BC partition,
factor, refined solve, device load vector), the displacements and free-DOF
counts compared with the reference's recorded results."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_golden_decks_end_to_end(tmp_path):
    from jf_2026_05_25_amd import deck as D
    from jf_2026_05_25_amd import solver
    from _decks import fixture_deck
    from _helpers import golden_cases, golden_model
    from test_oracle_system import GOLDEN_TOL
    ctx = solver.Context(0)
    bad = []
    try:
        for case in golden_cases():
            dk = D.read_deck(fixture_deck(case, "small"))
            for precision in (64, 32):
                print(f"deck {case['case']} precision {precision}", flush=True)
                res = D.solve_deck(dk, precision=precision, membrane_modes=6, ctx=ctx)
                path = tmp_path / f"{case['case']}_{precision}.JU.JSON"
                D.write_ju(str(path), dk, res)
                ju = json.load(open(path))
                rows = ju["displacements"]
                per = len(dk.grid_ids)
                for k, s in enumerate(case["subcases"]):
                    u = np.array([[r[c] for c in ("t1", "t2", "t3", "r1", "r2", "r3")]
                                  for r in rows[k * per:(k + 1) * per]]).ravel()
                    exp = golden_model(case, s).meta["expected"]
                    err = np.linalg.norm(u - exp) / np.linalg.norm(exp)
                    if not err < GOLDEN_TOL[(case["case"], s["subcase"])]:
                        bad.append((case["case"], s["subcase"], precision, err))
                    d = ju["solver_diagnostics"][k]
                    if d["sid"] != s["subcase"] or d["details"]["bc_partition"]["free_dofs"] != s["reference_free_dofs"]:
                        bad.append((case["case"], s["subcase"], precision, "diagnostics"))
                    if not d["details"]["linear_solver"]["relative_residual"] <= 1e-10:
                        bad.append((case["case"], s["subcase"], precision, "residual"))
    finally:
        ctx.close()
    assert not bad, bad
