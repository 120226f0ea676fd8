"""Synthetic SOL 101 models: the five workloads of this path, built in memory.

A :class:`Model` holds exactly what the reference's ``assemble_stiffness`` and
``apply_bc_and_solve`` consume for this path (src/solver/assembly.jl:2630,
src/solver/boundary_conditions.jl:1324), already resolved to arrays:

* grid points in ascending-ID order (index ``i`` owns global DOFs ``6i..6i+5``,
  src/solver/assembly.jl:2653-2657),
* CQUAD4 / CTRIA3 connectivity and per-element PSHELL+MAT1 section records,
* CTETRA10 connectivity (corners, then edges (1,2),(2,3),(1,3),(1,4),(2,4),(3,4),
  src/tetra10.jl:6) and per-element MAT1 (E, nu),
* the SPC1 mask and the load vectors (one column per right-hand side).

Shell property columns (``SP_*``) match ``include/jfem_hip.h``.
All generators are deterministic for a given seed.
"""
from __future__ import annotations

from dataclasses import dataclass, field
import math

import numpy as np

SP_H, SP_E, SP_NU, SP_G, SP_BR, SP_TST, SP_RIGID, SP_MODES = range(8)
SP_NCOL = 8


@dataclass
class Model:
    name: str
    xyz: np.ndarray                      # (n_nodes, 3) float64
    quads: np.ndarray                    # (nq, 4) int32
    quad_props: np.ndarray               # (nq, 8) float64
    trias: np.ndarray                    # (nt, 3) int32
    tria_props: np.ndarray               # (nt, 8) float64
    tets: np.ndarray                     # (ntet, 10) int32
    tet_props: np.ndarray                # (ntet, 2) float64: E, nu
    spc: np.ndarray                      # (ndof,) uint8, 1 = SPC1-fixed
    loads: np.ndarray                    # (ndof, nrhs) float64
    k6rot: float = 100.0
    meta: dict = field(default_factory=dict)

    @property
    def n_nodes(self) -> int:
        return int(self.xyz.shape[0])

    @property
    def ndof(self) -> int:
        return 6 * self.n_nodes

    @property
    def nrhs(self) -> int:
        return int(self.loads.shape[1])

    def summary(self) -> dict:
        return {"name": self.name, "nodes": self.n_nodes, "dofs": self.ndof,
                "cquad4": int(len(self.quads)), "ctria3": int(len(self.trias)),
                "ctetra10": int(len(self.tets)), "spc_dofs": int(self.spc.sum()),
                "nrhs": self.nrhs}

    def with_thickness(self, quad_h: np.ndarray | None = None,
                       tria_h: np.ndarray | None = None) -> "Model":
        """Same topology, new per-element thickness (design variants)."""
        qp = self.quad_props.copy()
        tp = self.tria_props.copy()
        if quad_h is not None:
            qp[:, SP_H] = quad_h
        if tria_h is not None:
            tp[:, SP_H] = tria_h
        return Model(self.name, self.xyz, self.quads, qp, self.trias, tp, self.tets,
                     self.tet_props, self.spc, self.loads, self.k6rot, dict(self.meta))


def pshell_props(n: int, h, E: float, nu: float, G: float = 0.0, bend_ratio: float = 1.0,
                 ts_t: float = 5.0 / 6.0, mid3_blank: bool = False,
                 membrane_modes: int = 15) -> np.ndarray:
    """PSHELL,MID1=MID2=(MID3)=MAT1 section rows (src/solver/assembly.jl:3380-3404).

    ``bend_ratio`` is 12I/T^3 (0 for a blank MID2); ``mid3_blank`` selects the
    rigid transverse-shear limit of a blank MID3 (src/solver/assembly.jl:4182-4187).
    ``membrane_modes`` is the CQUAD4 membrane internal-mode mask (bit i = mode i+1):
    15 (all four) is the default; 6 keeps modes 2 and 3 only, the reference's
    JFEM_SOL101_Q4_CROSS_MEMBRANE_WEIGHTS_ENABLED option (src/solver/assembly.jl:2806-2809,
    4575-4583).
    """
    p = np.zeros((n, SP_NCOL))
    p[:, SP_H] = h
    p[:, SP_E] = E
    p[:, SP_NU] = nu
    p[:, SP_G] = G
    p[:, SP_BR] = bend_ratio
    p[:, SP_TST] = ts_t
    p[:, SP_RIGID] = 1.0 if (mid3_blank and bend_ratio > 1e-12) else 0.0
    p[:, SP_MODES] = membrane_modes
    return p


def _empty_i(n):
    return np.zeros((0, n), dtype=np.int32)


def _empty_f(n):
    return np.zeros((0, n))


def pload4_forces(xyz: np.ndarray, quads: np.ndarray, pressure) -> np.ndarray:
    """Uniform PLOAD4 on CQUAD4: each corner gets (area*P/4) along normalize(d13 x d24)
    (src/solver/loads.jl:541-586).  Returns a (6n,) vector."""
    ndof = 6 * xyz.shape[0]
    F = np.zeros(ndof)
    if len(quads) == 0:
        return F
    X = xyz[quads]                               # (nq,4,3)
    n = np.cross(X[:, 2] - X[:, 0], X[:, 3] - X[:, 1])
    nrm = np.linalg.norm(n, axis=1)
    area = 0.5 * nrm
    d = n / nrm[:, None]
    p = np.broadcast_to(np.asarray(pressure, dtype=float), (len(quads),))
    fnode = d * (area * p / 4.0)[:, None]
    for k in range(4):
        for c in range(3):
            np.add.at(F, 6 * quads[:, k] + c, fnode[:, c])
    return F


# ---------------------------------------------------------------------------
# C1: flat square plate, one edge clamped, uniform pressure
# ---------------------------------------------------------------------------
def flat_plate(nx: int = 40, ny: int = 40, Lx: float = 0.5, Ly: float = 0.5,
               E: float = 70e9, nu: float = 0.33, t: float = 0.002,
               pressure: float = 1000.0, tria_fraction: float = 0.0,
               thickness=None, seed: int = 20260525, name: str = "flat_plate") -> Model:
    """Square plate in z=0 meshed nx x ny CQUAD4, edge x=0 SPC1 123456,
    uniform PLOAD4.  ``tria_fraction`` splits that share of cells into two CTRIA3."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, Lx, nx + 1)
    ys = np.linspace(0.0, Ly, ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    xyz = np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)
    nid = lambda i, j: j * (nx + 1) + i
    cells = [(nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
             for j in range(ny) for i in range(nx)]
    cells = np.array(cells, dtype=np.int32)
    split = rng.random(len(cells)) < tria_fraction if tria_fraction > 0 else np.zeros(len(cells), bool)
    quads = cells[~split]
    tri_cells = cells[split]
    trias = np.concatenate([tri_cells[:, [0, 1, 2]], tri_cells[:, [0, 2, 3]]]) if len(tri_cells) else _empty_i(3)
    if thickness is None:
        hq = np.full(len(quads), t)
        ht = np.full(len(trias), t)
    else:
        hq, ht = thickness(rng, len(quads)), thickness(rng, len(trias))
    spc = np.zeros(6 * len(xyz), dtype=np.uint8)
    for j in range(ny + 1):
        spc[6 * nid(0, j):6 * nid(0, j) + 6] = 1
    F = pload4_forces(xyz, quads, pressure)
    if len(trias):
        Xt = xyz[trias]
        n = np.cross(Xt[:, 1] - Xt[:, 0], Xt[:, 2] - Xt[:, 0])
        nrm = np.linalg.norm(n, axis=1)
        fn = (n / nrm[:, None]) * (0.5 * nrm * pressure / 3.0)[:, None]
        for k in range(3):
            for c in range(3):
                np.add.at(F, 6 * trias[:, k] + c, fn[:, c])
    return Model(name, xyz, quads, pshell_props(len(quads), hq, E, nu), trias,
                 pshell_props(len(trias), ht, E, nu), _empty_i(10), _empty_f(2), spc,
                 F[:, None].copy(), 100.0,
                 {"config": f"{nx}x{ny} CQUAD4 plate {Lx}x{Ly} m, t={t}, p={pressure} Pa"})


def ss_plate_compression(n: int, size: float = 1.0, t: float = 0.01, E: float = 210e9, nu: float = 0.3,
                         rigid_shear: bool = False, tria_fraction: float = 0.0) -> Model:
    """Simply supported square plate (size x size x t, n x n cells) under a uniform
    uniaxial compression sigma_x = -1 Pa: w = 0 on the four edges, u = 0 on x = 0,
    v = 0 at the origin, consistent edge forces on x = size (the SOL 105 preload
    of a Timoshenko-Gere buckling plate; first root 4 pi^2 D / (size^2 t)).
    ``tria_fraction`` splits that share of cells into two CTRIA3."""
    m = flat_plate(n, n, Lx=size, Ly=size, E=E, nu=nu, t=t, tria_fraction=tria_fraction,
                   name="ss_plate_compression")
    xy = m.xyz
    spc = np.zeros(m.ndof, np.uint8)
    x0, x1 = np.isclose(xy[:, 0], 0.0), np.isclose(xy[:, 0], size)
    y0, y1 = np.isclose(xy[:, 1], 0.0), np.isclose(xy[:, 1], size)
    spc[6 * np.where(x0 | x1 | y0 | y1)[0] + 2] = 1
    spc[6 * np.where(x0)[0]] = 1
    spc[6 * np.where(x0 & y0)[0] + 1] = 1
    F = np.zeros(m.ndof)
    right = np.where(x1)[0]
    w = np.where(y0[right] | y1[right], 0.5, 1.0) * (size / n)
    F[6 * right] = -t * w
    m.spc = spc
    m.loads = F[:, None]
    if rigid_shear:
        m.quad_props[:, SP_RIGID] = 1.0
    kind = "CQUAD4" if tria_fraction <= 0 else f"shell ({tria_fraction:g} of cells as CTRIA3 pairs)"
    m.meta["config"] = f"simply supported {n}x{n} {kind} plate {size}x{size} m, t={t}, sigma_x = -1 Pa"
    return m


# ---------------------------------------------------------------------------
# C2: wing box (closed rectangular tube of four CQUAD4 faces)
# ---------------------------------------------------------------------------
def wing_box(nx: int = 256, nw: int = 64, L: float = 4.0, W: float = 1.0, H: float = 0.2,
             E: float = 70e9, nu: float = 0.33, t_range=(1e-3, 4e-3), nrhs: int = 8,
             force_sigma: float = 1000.0, tria_fraction: float = 0.0,
             seed: int = 20260525, name: str = "wing_box") -> Model:
    """Box 4 x 1 x 0.2 m: top/bottom skins and two spar webs, each nx x nw CQUAD4,
    sharing corner-line grids; per-element thickness ~U(t_range); root x=0 SPC1
    123456; ``nrhs`` random nodal force sets ~N(0, force_sigma) on translations."""
    rng = np.random.default_rng(seed)
    # perimeter of the cross-section: 4*nw points around (y,z), closed loop
    per = []
    for k in range(nw):                       # bottom skin y: 0 -> W at z=0
        per.append((W * k / nw, 0.0))
    for k in range(nw):                       # web y=W, z: 0 -> H
        per.append((W, H * k / nw))
    for k in range(nw):                       # top skin y: W -> 0 at z=H
        per.append((W - W * k / nw, H))
    for k in range(nw):                       # web y=0, z: H -> 0
        per.append((0.0, H - H * k / nw))
    per = np.array(per)
    npr = len(per)
    xs = np.linspace(0.0, L, nx + 1)
    xyz = np.zeros(((nx + 1) * npr, 3))
    for i, x in enumerate(xs):
        xyz[i * npr:(i + 1) * npr, 0] = x
        xyz[i * npr:(i + 1) * npr, 1] = per[:, 0]
        xyz[i * npr:(i + 1) * npr, 2] = per[:, 1]
    nid = lambda i, k: i * npr + (k % npr)
    cells = np.array([(nid(i, k), nid(i + 1, k), nid(i + 1, k + 1), nid(i, k + 1))
                      for i in range(nx) for k in range(npr)], dtype=np.int32)
    split = rng.random(len(cells)) < tria_fraction if tria_fraction > 0 else np.zeros(len(cells), bool)
    quads = cells[~split]
    tc = cells[split]
    trias = np.concatenate([tc[:, [0, 1, 2]], tc[:, [0, 2, 3]]]) if len(tc) else _empty_i(3)
    hq = rng.uniform(t_range[0], t_range[1], len(quads))
    ht = rng.uniform(t_range[0], t_range[1], len(trias))
    ndof = 6 * len(xyz)
    spc = np.zeros(ndof, dtype=np.uint8)
    for k in range(npr):
        spc[6 * nid(0, k):6 * nid(0, k) + 6] = 1
    F = np.zeros((ndof, nrhs))
    trans = np.zeros(ndof, bool)
    for c in range(3):
        trans[c::6] = True
    free_trans = trans & (spc == 0)
    F[free_trans, :] = rng.normal(0.0, force_sigma, (int(free_trans.sum()), nrhs))
    return Model(name, xyz, quads, pshell_props(len(quads), hq, E, nu), trias,
                 pshell_props(len(trias), ht, E, nu), _empty_i(10), _empty_f(2), spc, F, 100.0,
                 {"config": f"wing box {L}x{W}x{H} m, {nx}x{nw} CQUAD4 per face, {nrhs} RHS"})


# ---------------------------------------------------------------------------
# C3: solid block of 5-tet-split hexes, quadratic CTETRA10
# ---------------------------------------------------------------------------
_HEX_CORNERS = np.array([(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0),
                         (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)])
# five-tet split of a hex with corners 0..7 (and its mirror for alternate cells);
# each tet listed with positive orientation in the reference frame
_TETS_A = [(0, 1, 3, 4), (1, 2, 3, 6), (1, 4, 5, 6), (3, 6, 7, 4), (1, 3, 4, 6)]
_TETS_B = [(0, 1, 2, 5), (0, 2, 3, 7), (0, 5, 4, 7), (2, 5, 7, 6), (0, 2, 5, 7)]
_TET_EDGES = ((0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3))


def solid_block(n: int = 24, size: float = 1.0, E: float = 210e9, nu: float = 0.3,
                traction: float = 1e6, nrhs: int = 1, seed: int = 20260525,
                name: str = "solid_block") -> Model:
    """Cube meshed n^3 hexes, each split into five CTETRA10 (alternating split for a
    conforming mesh); base z=0 SPC1 123; random top-face traction ~U(-1,1)*traction
    per face triangle, applied as consistent quadratic-face loads (midside A/3)."""
    rng = np.random.default_rng(seed)
    m = 2 * n + 1                                   # half-grid points per axis
    h = size / (2 * n)
    gid = -np.ones(m ** 3, dtype=np.int64)

    def hid(a, b, c):
        return (c * m + b) * m + a

    tets = []
    for k in range(n):
        for j in range(n):
            for i in range(n):
                pat = _TETS_A if (i + j + k) % 2 == 0 else _TETS_B
                corner = [(2 * (i + dx), 2 * (j + dy), 2 * (k + dz)) for dx, dy, dz in _HEX_CORNERS]
                for tet in pat:
                    cs = [corner[q] for q in tet]
                    nodes = [hid(*c) for c in cs]
                    for a, b in _TET_EDGES:
                        mid = tuple((np.array(cs[a]) + np.array(cs[b])) // 2)
                        nodes.append(hid(*mid))
                    tets.append(nodes)
    tets = np.array(tets, dtype=np.int64)
    used = np.unique(tets)
    gid[used] = np.arange(len(used))
    tets = gid[tets].astype(np.int32)
    a = used % m
    b = (used // m) % m
    c = used // (m * m)
    xyz = np.stack([a * h, b * h, c * h], axis=1).astype(float)
    # orientation fix: make every tet positive
    P = xyz[tets[:, :4]]
    vol = np.einsum("ij,ij->i", np.cross(P[:, 1] - P[:, 0], P[:, 2] - P[:, 0]), P[:, 3] - P[:, 0])
    neg = vol < 0
    if neg.any():
        t = tets[neg]
        # swap corners 2 and 3 and their edge nodes: edges (1,3)<->(1,4), (2,3)<->(2,4)
        t2 = t.copy()
        t2[:, 2], t2[:, 3] = t[:, 3], t[:, 2]
        t2[:, 6], t2[:, 7] = t[:, 7], t[:, 6]      # (1,3) <-> (1,4)
        t2[:, 5], t2[:, 8] = t[:, 8], t[:, 5]      # (2,3) <-> (2,4)
        t2[:, 9] = t[:, 9]                          # (3,4) stays (4,3)
        tets[neg] = t2
    ndof = 6 * len(xyz)
    spc = np.zeros(ndof, dtype=np.uint8)
    base = np.where(np.abs(xyz[:, 2]) < 1e-12)[0]
    for c_ in range(3):
        spc[6 * base + c_] = 1
    # top-face triangles: tet faces with all three corners on z=size
    F = np.zeros((ndof, nrhs))
    top = np.abs(xyz[:, 2] - size) < 1e-12
    faces = ((0, 1, 2, 4, 5, 6), (0, 1, 3, 4, 8, 7), (1, 2, 3, 5, 9, 8), (0, 2, 3, 6, 9, 7))
    # edge index lookup within the tet node list: (0,1)->4,(1,2)->5,(0,2)->6,(0,3)->7,(1,3)->8,(2,3)->9
    tri = []
    for f in faces:
        on = top[tets[:, f[0]]] & top[tets[:, f[1]]] & top[tets[:, f[2]]]
        for row in tets[on]:
            tri.append([row[q] for q in f])
    tri = np.array(tri, dtype=np.int64).reshape(-1, 6)
    for r in range(nrhs):
        tr = rng.uniform(-1.0, 1.0, (len(tri), 3)) * traction
        Pc = xyz[tri[:, :3]]
        area = 0.5 * np.linalg.norm(np.cross(Pc[:, 1] - Pc[:, 0], Pc[:, 2] - Pc[:, 0]), axis=1)
        fm = tr * (area / 3.0)[:, None]
        for q in range(3, 6):
            for c_ in range(3):
                np.add.at(F[:, r], 6 * tri[:, q] + c_, fm[:, c_])
    props = np.zeros((len(tets), 2))
    props[:, 0] = E
    props[:, 1] = nu
    return Model(name, xyz, _empty_i(4), _empty_f(8), _empty_i(3), _empty_f(8), tets, props,
                 spc, F, 100.0,
                 {"config": f"{n}^3 hexes x 5 CTETRA10, cube {size} m"})


# ---------------------------------------------------------------------------
# C5: stiffened plate (design-variant base)
# ---------------------------------------------------------------------------
def stiffened_plate(n: int = 128, L: float = 1.0, n_stiff: int = 4, blade_h: float = 0.04,
                    blade_cells: int = 4, E: float = 70e9, nu: float = 0.33, t: float = 0.003,
                    nrhs: int = 4, force_sigma: float = 1000.0, seed: int = 20260525,
                    name: str = "stiffened_plate") -> Model:
    """n x n CQUAD4 plate in z=0 with ``n_stiff`` blade stiffeners (vertical CQUAD4
    strips along x, ``blade_cells`` high) sharing the plate grid line; edge x=0
    clamped; ``nrhs`` random nodal force sets on free translations."""
    rng = np.random.default_rng(seed)
    xs = np.linspace(0.0, L, n + 1)
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    xyz = [np.stack([X.ravel(), Y.ravel(), np.zeros(X.size)], axis=1)]
    nid = lambda i, j: j * (n + 1) + i
    quads = [(nid(i, j), nid(i + 1, j), nid(i + 1, j + 1), nid(i, j + 1))
             for j in range(n) for i in range(n)]
    base = (n + 1) * (n + 1)
    rows = [int(round((s + 1) * n / (n_stiff + 1))) for s in range(n_stiff)]
    for s, jrow in enumerate(rows):
        start = base + sum(len(a) for a in xyz[1:])
        pts = []
        for k in range(1, blade_cells + 1):
            for i in range(n + 1):
                pts.append((xs[i], xs[jrow], blade_h * k / blade_cells))
        xyz.append(np.array(pts))
        bid = lambda i, k: nid(i, jrow) if k == 0 else start + (k - 1) * (n + 1) + i
        for k in range(blade_cells):
            for i in range(n):
                quads.append((bid(i, k), bid(i + 1, k), bid(i + 1, k + 1), bid(i, k + 1)))
    xyz = np.concatenate(xyz)
    quads = np.array(quads, dtype=np.int32)
    ndof = 6 * len(xyz)
    spc = np.zeros(ndof, dtype=np.uint8)
    clamp = np.where(np.abs(xyz[:, 0]) < 1e-12)[0]
    for c in range(6):
        spc[6 * clamp + c] = 1
    F = np.zeros((ndof, nrhs))
    trans = np.zeros(ndof, bool)
    for c in range(3):
        trans[c::6] = True
    ft = trans & (spc == 0)
    F[ft, :] = rng.normal(0.0, force_sigma, (int(ft.sum()), nrhs))
    return Model(name, xyz, quads, pshell_props(len(quads), t, E, nu), _empty_i(3), _empty_f(8),
                 _empty_i(10), _empty_f(2), spc, F, 100.0,
                 {"config": f"{n}x{n} CQUAD4 plate + {n_stiff} blades"})


def design_variants(base: Model, count: int = 64, spread: float = 0.2, seed: int = 20260525):
    """``count`` copies of ``base`` with per-element thickness x U(1-spread, 1+spread)."""
    rng = np.random.default_rng(seed + 1)
    out = []
    for v in range(count):
        fq = rng.uniform(1 - spread, 1 + spread, len(base.quads))
        ft = rng.uniform(1 - spread, 1 + spread, len(base.trias))
        m = base.with_thickness(base.quad_props[:, SP_H] * fq, base.tria_props[:, SP_H] * ft)
        m.name = f"{base.name}_v{v}"
        out.append(m)
    return out


# ---------------------------------------------------------------------------
# C4: serving stream request generator
# ---------------------------------------------------------------------------
def shell_request(target_dofs: float, rng: np.random.Generator, nrhs: int,
                  tria_fraction: float = 0.1, idx: int = 0) -> Model:
    """A flat plate or a box shell of roughly ``target_dofs`` DOFs, 10% CTRIA3,
    random thickness field, ``nrhs`` random load sets."""
    seed = int(rng.integers(0, 2 ** 31 - 1))
    if rng.random() < 0.5:
        side = max(2, int(round(math.sqrt(target_dofs / 6.0))) - 1)
        m = flat_plate(side, side, 0.5, 0.5, tria_fraction=tria_fraction,
                       thickness=lambda r, k: r.uniform(1e-3, 4e-3, k), seed=seed,
                       name=f"req{idx}_plate")
        F = np.zeros((m.ndof, nrhs))
        r2 = np.random.default_rng(seed + 7)
        trans = np.zeros(m.ndof, bool)
        for c in range(3):
            trans[c::6] = True
        ft = trans & (m.spc == 0)
        F[ft, :] = r2.normal(0.0, 1000.0, (int(ft.sum()), nrhs))
        m.loads = F
        return m
    nodes = target_dofs / 6.0
    nw = max(1, int(round(math.sqrt(nodes / 16.0))))
    nx = max(2, int(round(nodes / (4 * nw))) - 1)
    return wing_box(nx, nw, nrhs=nrhs, tria_fraction=tria_fraction, seed=seed, name=f"req{idx}_box")
