"""GPU parity of the device load-vector builder (FORCE / MOMENT / shell PLOAD4 /
GRAV on shell mass of one load set) with the oracle: bit-exact (the sums run in
the reference's order on both sides), on random load sets over the wing box and a mixed plate, with
repeated cards, explicit directions, a zero-area element, an empty set, the
device-output entry point, and end to end on the golden decks (device-built
loads -> solve -> the reference's recorded displacements)."""
import ctypes as C

import numpy as np
import pytest

from jf_2026_05_25_amd import meshes

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from jf_2026_05_25_amd import solver
    c = solver.Context(0)
    yield c
    c.close()


def random_cards(m, rng, npl, nf, nm):
    ne = len(m.quads) + len(m.trias)
    pl_e = rng.integers(0, ne, npl)
    pl_p = rng.normal(0.0, 1e3, npl)
    pl_d = rng.normal(size=(npl, 3))
    pl_d[rng.random(npl) < 0.5] = 0.0                 # half the cards: element normal
    f_n = rng.integers(0, m.n_nodes, nf)
    f_n[: nf // 3] = f_n[0]                            # repeated grid
    f_m = rng.normal(0.0, 50.0, nf)
    f_d = rng.normal(size=(nf, 3))
    m_n = rng.integers(0, m.n_nodes, nm)
    m_m = rng.normal(0.0, 5.0, nm)
    m_d = rng.normal(size=(nm, 3))
    return (pl_e, pl_p, pl_d), (f_n, f_m, f_d), (m_n, m_m, m_d)


@pytest.mark.parametrize("name,scale", [("box", 1.0), ("box", 0.7), ("plate_mixed", 1.0), ("plate_mixed", -2.5)])
def test_load_vector_bitwise_equals_oracle(ctx, oracle, name, scale):
    from jf_2026_05_25_amd import solver
    m = meshes.wing_box(24, 4, tria_fraction=0.1) if name == "box" else meshes.flat_plate(12, 10, tria_fraction=0.3)
    rng = np.random.default_rng(17)
    pl, fc, mc = random_cards(m, rng, 3 * (len(m.quads) + len(m.trias)), 40, 25)
    print(f"load set: {name}, scale {scale}, {len(pl[0])} PLOAD4, {len(fc[0])} FORCE, {len(mc[0])} MOMENT", flush=True)
    # two GRAV cards on shell mass, a fifth of the shells massless
    ne = len(m.quads) + len(m.trias)
    areal = rng.uniform(1.0, 20.0, ne)
    areal[rng.random(ne) < 0.2] = 0.0
    gr = {"grav": (rng.normal(0.0, 10.0, 2), rng.normal(size=(2, 3))), "shell_areal": areal}
    mesh = solver.DeviceMesh(m, ctx)
    ls = solver.LoadSet(mesh, pload4=pl, force=fc, moment=mc, **gr)
    F = ls.vector(scale)
    Fo = oracle.load_vector(m, pload4=pl, force=fc, moment=mc, scale=scale, **gr)
    assert np.array_equal(F, Fo), np.abs(F - Fo).max()
    # without the GRAV cards
    ls.set_gravity(([], np.zeros((0, 3))), areal)
    assert np.array_equal(ls.vector(scale), oracle.load_vector(m, pload4=pl, force=fc, moment=mc, scale=scale))
    ls.set_gravity(gr["grav"], areal)
    # device output on the context's stream
    nb = 8 * m.ndof
    dptr = C.c_void_p()
    L = ctx.L
    assert L.jfem_dev_alloc(ctx.h, nb, C.byref(dptr)) == 0
    try:
        ls.vector_dev(dptr.value, scale)
        Fd = np.zeros(m.ndof)
        assert L.jfem_memcpy_d2h(ctx.h, Fd.ctypes.data, dptr, nb) == 0
        assert np.array_equal(Fd, Fo)
    finally:
        L.jfem_dev_free(ctx.h, dptr)
    ls.free()


def test_load_vector_edge_cases(ctx, oracle):
    from jf_2026_05_25_amd import solver
    from jf_2026_05_25_amd._lib import JfemError
    # a zero-area quad next to a regular one; an empty set; cards out of range
    X = np.array([[0, 0, 0], [1.0, 0, 0], [1.0, 1.0, 0], [0, 1.0, 0], [2.0, 0, 0], [2.0, 0, 0]])
    quads = np.array([[0, 1, 2, 3], [1, 4, 5, 1]], np.int32)
    m = meshes.Model("edge", X, quads, np.zeros((2, 8)), np.zeros((0, 3), np.int32), np.zeros((0, 8)),
                     np.zeros((0, 10), np.int32), np.zeros((0, 2)), np.zeros(36, np.uint8), np.zeros((36, 1)),
                     100.0, {})
    mesh = solver.DeviceMesh(m, ctx)
    pl = (np.array([1, 0, 1]), np.array([5.0, 4.0, 7.0]))
    F = solver.LoadSet(mesh, pload4=pl).vector()
    Fo = oracle.load_vector(m, pload4=pl)
    assert np.array_equal(F, Fo)
    assert F[2] == 1.0 and np.count_nonzero(F) == 4      # only the unit square, p = 4
    assert np.all(solver.LoadSet(mesh).vector(3.0) == 0.0)
    with pytest.raises(JfemError):
        solver.LoadSet(mesh, pload4=(np.array([2]), np.array([1.0])))
    with pytest.raises(JfemError):
        solver.LoadSet(mesh, force=(np.array([6]), np.array([1.0]), np.array([[1.0, 0, 0]])))


def test_golden_decks_with_device_loads(ctx, oracle):
    """Device-built loads of the golden decks' subcases, solved on the GPU (FP64),
    reproduce the reference's recorded displacements as the oracle does."""
    from _helpers import golden_cases, golden_grav, golden_model
    from test_oracle_system import GOLDEN_TOL
    from jf_2026_05_25_amd import solver
    bad = []
    for c in golden_cases():
        ids = sorted(int(k) for k in c["grids"])
        idx = {g: i for i, g in enumerate(ids)}
        eid_row = {q["eid"]: e for e, q in enumerate(c["cquad4"])}
        for s in c["subcases"]:
            m = golden_model(c, s, membrane_modes=6)
            fc, pl = s["force"], s["pload4"]
            force = ([idx[f["grid"]] for f in fc], solver.single_precision([f["mag"] for f in fc]),
                     solver.single_precision([f["dir"] for f in fc])) if fc else None
            pload4 = ([eid_row[p["eid"]] for p in pl], [p["p"] for p in pl]) if pl else None
            print(f"golden deck {c['case']} subcase {s['subcase']}", flush=True)
            asm = solver.assemble_stiffness(m, ctx)
            gr = golden_grav(c, s)
            F = solver.LoadSet(asm.mesh, pload4=pload4, force=force, **gr).vector()
            if not np.array_equal(F, oracle.load_vector(m, pload4=pload4, force=force, **gr)):
                bad.append((c["case"], s["subcase"], "load vector"))
                continue
            u, fixed, spc, diag = solver.apply_bc_and_solve(asm, m, F_applied=F[:, None], precision=64)
            exp = m.meta["expected"]
            err = np.linalg.norm(u.reshape(m.ndof, -1)[:, 0] - exp) / np.linalg.norm(exp)
            if not err < GOLDEN_TOL[(c["case"], s["subcase"])]:
                bad.append((c["case"], s["subcase"], err))
    assert not bad, bad
