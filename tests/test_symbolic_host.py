"""Host-side symbolic analysis (ordering, supernodes, front maps, task lists),
checked on the CPU: a sequential harness runs the same task decomposition the
device kernels run and must reproduce a dense solve of K_ff."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from jf_2026_05_25_amd import meshes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _build_harness(tmp_path_factory, chunk=None):
    # chunk None: the library's own split of the solve sums; else partial sums of
    # `chunk` blocks for sums over more than 2 chunk blocks
    out = str(tmp_path_factory.mktemp("harness") / f"libharness{chunk or ''}.so")
    defs = [] if chunk is None else [f"-DJFEM_SOLVE_CHUNK={chunk}", f"-DJFEM_SOLVE_SPLIT={2 * chunk}",
                                     f"-DJFEM_SOLVE_SPLIT_ROWS={2 * chunk}"]
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", *defs, "-o", out,
                    os.path.join(HERE, "symbolic_harness.cpp"),
                    os.path.join(ROOT, "jf_2026_05_25_amd", "csrc", "symbolic.cpp")], check=True)
    L = ctypes.CDLL(out)
    L.harness_solve.restype = ctypes.c_int
    vp = ctypes.c_void_p
    L.harness_solve.argtypes = [ctypes.c_int64, vp, vp, vp, vp, vp, ctypes.c_int, vp, vp, vp]
    return L


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    return _build_harness(tmp_path_factory)


@pytest.fixture(scope="module")
def harness_chunk1(tmp_path_factory):
    # block sums over more than two blocks split into one-block partial tasks, so
    # that the small models run every kind of split task
    return _build_harness(tmp_path_factory, 1)


def run(L, K, xyz, fixed, B):
    K = K.tocsr(); K.sort_indices()
    n = K.shape[0]
    B = np.asarray(B, float).reshape(n, -1)
    nrhs = B.shape[1]
    Bc = np.ascontiguousarray(B.T).ravel(); X = np.zeros_like(Bc)
    info = np.zeros(10, np.int64)
    rp = K.indptr.astype(np.int64); ci = K.indices.astype(np.int32); v = np.ascontiguousarray(K.data)
    xyzc = np.ascontiguousarray(xyz, float).ravel(); fx = np.ascontiguousarray(fixed, np.uint8)
    st = L.harness_solve(n, rp.ctypes.data, ci.ctypes.data, v.ctypes.data, xyzc.ctypes.data,
                         fx.ctypes.data, nrhs, Bc.ctypes.data, X.ctypes.data, info.ctypes.data)
    assert st == 0, st
    return X.reshape(nrhs, n).T, info


MODELS = {
    "plate": lambda: meshes.flat_plate(24, 20),
    "plate_tria": lambda: meshes.flat_plate(16, 16, tria_fraction=0.3),
    "box": lambda: meshes.wing_box(24, 4, tria_fraction=0.1, nrhs=3),
    "block": lambda: meshes.solid_block(3, nrhs=2),
    "stiffened": lambda: meshes.stiffened_plate(12, n_stiff=2, blade_cells=2, nrhs=2),
    # a six-panel separator: inversion by doubling at block sizes 1, 2 and 4 (uneven last pair)
    "block5": lambda: meshes.solid_block(5, nrhs=2),
}


@pytest.mark.parametrize("chunk", [4, 1])
@pytest.mark.parametrize("name", list(MODELS))
def test_task_decomposition_solves(harness, harness_chunk1, oracle, name, chunk):
    m = MODELS[name]()
    K = oracle.assemble(m)
    fixed = oracle.bc_partition(K, m.spc)
    x, info = run(harness if chunk == 4 else harness_chunk1, K, m.xyz, fixed, m.loads)
    free = np.where(fixed == 0)[0]
    Kd = K.toarray()[np.ix_(free, free)]
    Ku = np.triu(Kd); Ks = Ku + np.triu(Ku, 1).T
    xe = np.linalg.solve(Ks, m.loads[free])
    np.testing.assert_allclose(x[free], xe, rtol=1e-8, atol=1e-9 * np.abs(xe).max())
    assert np.all(x[fixed == 1] == 0)
    ns, nlev, nL, nU, npt, nut, maxf, nleaf = info[:8]
    assert ns > 1 and nlev > 1 and nL >= np.count_nonzero(np.tril(Ks))
    assert 0 < nleaf < ns            # bottom subtrees (one wavefront per task) and the rest
    if chunk == 1 and name in ("box", "block5"):
        assert info[8] > 0 and info[9] > 0   # partial-sum tasks in both lists


def test_fill_is_nested_dissection_quality(harness, oracle):
    """Geometric nested dissection on a 40x40 plate: nnz(L) within a small factor
    of n log n (a banded ordering would give ~n * bandwidth)."""
    m = meshes.flat_plate(40, 40)
    K = oracle.assemble(m)
    fixed = oracle.bc_partition(K, m.spc)
    x, info = run(harness, K, m.xyz, fixed, m.loads)
    n = int((fixed == 0).sum())
    assert info[2] < 40 * n * np.log2(n)
