"""GPU shell force / stress / strain recovery against the oracle and the
reference's recorded rows (recover_shell_stresses!, src/solver/stress_recovery.jl:505).

* Raw per-element fields (44 per CQUAD4, 20 per CTRIA3) from DeviceMesh.shell_recover
  equal the oracle's to 1e-11 of each field's largest magnitude over the mesh, in
  both branches, for a plate of CQUAD4 and CTRIA3, a wing box (shells in three
  planes) and a golden deck of warped CQUAD4 (finite-warp displacement map).
* The MYSTRAN deck's recorded rows (forces, corner forces, stresses, strains) from
  the recorded displacements, through the device kernel and the row formatter.
* The full wing box (65,536 CQUAD4) recovered from its FP64 solution: every
  field finite, a sample of 2,000 elements against the oracle.
"""
import time

import numpy as np
import pytest

from jf_2026_05_25_amd import meshes, recovery

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jf_2026_05_25_amd import solver
    c = solver.Context(0)
    yield c
    c.close()


def _models():
    from _helpers import golden_cases, golden_model
    c = [g for g in golden_cases() if g["case"] == "MH_twisted_beam_out_of_plane"][0]
    return {
        "plate_mixed": meshes.flat_plate(16, 12, tria_fraction=0.2),
        "box": meshes.wing_box(24, 6, nrhs=1),
        "twisted": golden_model(c, c["subcases"][0], membrane_modes=6),
    }


def _field_close(got, exp, rtol):
    """Per column: |got - exp| <= rtol * (largest |exp| of that column, floored
    by 1e-30 so that all-zero columns compare exactly)."""
    if exp.size == 0:
        return True, 0.0
    scale = np.maximum(np.max(np.abs(exp), axis=0), 1e-30)
    err = np.max(np.abs(got - exp) / scale)
    return err <= rtol, err


@pytest.mark.parametrize("mean_shear", [True, False])
@pytest.mark.parametrize("name", ["plate_mixed", "box", "twisted"])
def test_recovery_matches_oracle(ctx, oracle, name, mean_shear):
    from jf_2026_05_25_amd import solver
    m = _models()[name]
    rng = np.random.default_rng(7)
    u = rng.normal(size=m.ndof) * 1e-3
    mesh = solver.DeviceMesh(m, ctx)
    try:
        q, t = mesh.shell_recover(u, mean_shear)
    finally:
        mesh.free()
    qo, to = oracle.shell_recover(m, u, mean_shear)
    for got, exp, kind in ((q, qo, "quad4"), (t, to, "tria3")):
        ok, err = _field_close(got, exp, 1e-11)
        assert ok, (name, kind, err)
    assert len(q) == len(m.quads) and len(t) == len(m.trias)


def test_mystran_rows_on_gpu(ctx):
    """The recorded MYSTRAN rows (SOL 1, the branch of other sequences) from the
    recorded displacements, through the device kernel."""
    from _helpers import golden_cases, golden_model
    from jf_2026_05_25_amd import solver
    c = [g for g in golden_cases() if g["case"] == "MYSTRAN_sol101_simple"][0]
    for s in c["subcases"]:
        m = golden_model(c, s, membrane_modes=6)
        mesh = solver.DeviceMesh(m, ctx)
        try:
            q, _ = mesh.shell_recover(m.meta["expected"], mean_shear=(c["sol"] in (101, 105, 106)))
        finally:
            mesh.free()
        grids = [[int(g) for g in e["nodes"]] for e in c["cquad4"]]
        rows = recovery.shell_rows(q, [e["eid"] for e in c["cquad4"]], m.quad_props[:, 0], "quad4", grids)
        ref = s["shell_recovery"]
        for key in ("forces", "forces_bilin", "stresses", "strains"):
            for a, b in zip(rows[key], ref[key]):
                fa, fb = _flat(a), _flat(b)
                assert fa.keys() == fb.keys()
                scale = max(abs(v) for v in fb.values() if isinstance(v, float))
                for k in fb:
                    if isinstance(fb[k], float):
                        assert abs(fa[k] - fb[k]) <= 1e-9 * max(abs(fb[k]), scale), (key, k, fa[k], fb[k])
                    else:
                        assert fa[k] == fb[k]


def _flat(d, pre=""):
    out = {}
    for k, v in d.items():
        if isinstance(v, dict):
            out.update(_flat(v, pre + k + "."))
        elif isinstance(v, (int, float)) and not isinstance(v, bool) and k not in ("eid", "grid_id"):
            out[pre + k] = float(v)
        else:
            out[pre + k] = v
    return out


@pytest.mark.slow
def test_recovery_full_wing_box(ctx, oracle):
    from jf_2026_05_25_amd import solver
    m = meshes.wing_box()                       # 256 x 64 CQUAD4 per face
    asm = solver.assemble_stiffness(m, ctx)
    u, fixed, spc, diag = solver.apply_bc_and_solve(asm, m, precision=64)
    u = u.reshape(m.ndof, -1)[:, 0]
    q, t = asm.mesh.shell_recover(u, True)        # warm-up (module load)
    t0 = time.perf_counter()
    q, t = asm.mesh.shell_recover(u, True)
    dt = time.perf_counter() - t0
    print(f"\nshell recovery, {len(m.quads)} CQUAD4: {dt * 1e3:.2f} ms per call (host buffers)")
    assert q.shape == (len(m.quads), 44) and np.all(np.isfinite(q))
    idx = np.random.default_rng(3).choice(len(m.quads), 2000, replace=False)
    exp = np.array([oracle.q4_recover(m.xyz[m.quads[e]], m.quad_props[e],
                                      np.concatenate([u[6 * n:6 * n + 6] for n in m.quads[e]]), True) for e in idx])
    ok, err = _field_close(q[idx], exp, 1e-10)
    assert ok, err
