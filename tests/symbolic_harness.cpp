// CPU harness for the symbolic analysis (test infrastructure).
//
// Runs jfem::build_symbolic and then executes, sequentially on the CPU, the same
// task decomposition the device kernels execute (scatter, panel tasks, update
// tiles, the update matrices pushed into the parents launch by launch, the
// inversion tiles of X = L_11^{-1}, and the dataflow solve lists in their order),
// so the structures and task lists can be checked against a dense solve without
// a GPU.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../jf_2026_05_25_amd/csrc/symbolic.h"

using namespace jfem;

extern "C" int harness_solve(int64_t ndof, const int64_t *rowptr, const int32_t *col, const double *val,
                             const double *xyz, const uint8_t *fixed, int nrhs, const double *b, double *x,
                             int64_t *info) {
    SymbolicHost S;
    std::string err;
    if (build_symbolic(ndof, rowptr, col, xyz, fixed, S, err)) return 1;
    std::vector<double> L(S.L_entries + 1, 0.0), U(S.U_entries + 1, 0.0);
    std::vector<int64_t> Lstamp(S.L_entries + 1, 0), Ustamp(S.U_entries + 1, 0);   // last push launch per entry
    for (size_t i = 0; i < S.amap_src.size(); ++i) L[S.amap_dst[i]] = val[S.amap_src[i]];
    const int64_t nlev = (int64_t)S.level_ptr.size() - 1;
    std::vector<int64_t> lev_of(S.ns, -1);
    for (int64_t l = 0; l < nlev; ++l)
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) lev_of[S.level_sn[q]] = l;
    for (int64_t l = 0; l < nlev; ++l) {
        // factor the b x b diagonal block at column k of supernode s in place
        auto potrf = [&](int s, int64_t k) -> int {
            int64_t ns = S.ncol[s], ms = S.m[s];
            int b = (int)std::min<int64_t>(kPanel, ns - k);
            double *Ls = L.data() + S.lofs[s];
            for (int j = 0; j < b; ++j) {
                double d = Ls[(k + j) + ms * (k + j)];
                if (!(d > 0)) return 3;
                d = std::sqrt(d);
                Ls[(k + j) + ms * (k + j)] = d;
                for (int i = j + 1; i < b; ++i) Ls[(k + i) + ms * (k + j)] /= d;
                for (int c2 = j + 1; c2 < b; ++c2)
                    for (int i = c2; i < b; ++i)
                        Ls[(k + i) + ms * (k + c2)] -= Ls[(k + i) + ms * (k + j)] * Ls[(k + c2) + ms * (k + j)];
            }
            return 0;
        };
        // step 0 of the level: diagonal blocks of the first panels
        if (S.step_ptr[l + 1] > S.step_ptr[l]) {
            int64_t stp = S.step_ptr[l];
            for (int64_t q = S.pstep_off[stp]; q < S.pstep_mid[stp]; ++q) {
                PanelTask t = S.ptasks[q];
                if (t.r0 != t.k || t.k != 0) return 2;
                if (int e = potrf(t.s, t.k)) return e;
            }
        }
        for (int64_t stp = S.step_ptr[l]; stp < S.step_ptr[l + 1]; ++stp) {
            // tile solves: rows [max(r0, k+b), r0+kTile) of every panel task, X <- X L_kk^{-T}
            for (int64_t q = S.pstep_off[stp]; q < S.pstep_off[stp + 1]; ++q) {
                PanelTask t = S.ptasks[q];
                int s = t.s;
                int64_t ns = S.ncol[s], ms = S.m[s], k = t.k;
                int b = (int)std::min<int64_t>(kPanel, ns - k);
                double *Ls = L.data() + S.lofs[s];
                int64_t r1 = std::min<int64_t>(t.r0 + kTile, ms);
                for (int64_t i = std::max<int64_t>(t.r0, k + b); i < r1; ++i)
                    for (int j = 0; j < b; ++j) {
                        double v = Ls[i + ms * (k + j)];
                        for (int p = 0; p < j; ++p) v -= Ls[i + ms * (k + p)] * Ls[(k + j) + ms * (k + p)];
                        Ls[i + ms * (k + j)] = v / Ls[(k + j) + ms * (k + j)];
                    }
            }
            // trailing update tiles; the tile holding the next panel's diagonal block
            // (r0 == c0 == k + b < ns) then factors that block
            for (int64_t q = S.ustep_off[stp]; q < S.ustep_off[stp + 1]; ++q) {
                UpdateTask t = S.utasks[q];
                int s = t.s;
                int64_t ns = S.ncol[s], ms = S.m[s], k = t.k;
                int b = (int)std::min<int64_t>(kPanel, ns - k);
                int64_t rb = t.r0 < ns ? ns : ms, cb = t.c0 < ns ? ns : ms;
                int64_t r1 = std::min<int64_t>(t.r0 + kTile, rb), c1 = std::min<int64_t>(t.c0 + kTile, cb);
                const double *Ls = L.data() + S.lofs[s];
                for (int64_t c = t.c0; c < c1; ++c)
                    for (int64_t r = std::max<int64_t>(t.r0, c); r < r1; ++r) {
                        double acc = 0;
                        for (int p = 0; p < b; ++p) acc += Ls[r + ms * (k + p)] * Ls[c + ms * (k + p)];
                        if (c < ns) L[S.lofs[s] + r + ms * c] -= acc;
                        else U[S.uofs[s] + (r - ns) + (ms - ns) * (c - ns)] -= acc;
                    }
                if (t.r0 == t.c0 && t.c0 == k + b && t.c0 < ns)
                    if (int e = potrf(s, t.c0)) return e;
            }
        }
        // after the level's last step: the update matrices, full rank per front,
        // pushed into the parents' fronts, one launch per sibling rank: a task's
        // fields match the symbolic arrays (10), its parent lies on a later level
        // (14), and no parent entry is written twice within one launch (13)
        if (S.slev_ptr.size() != (size_t)nlev + 1 || S.slev_ptr[l] > S.slev_ptr[l + 1]) return 10;
        for (int64_t q = S.slev_ptr[l]; q < S.slev_ptr[l + 1]; ++q) {
            const int64_t launch = q + 1;
            for (int64_t e = S.spush_off[q]; e < S.spush_off[q + 1]; ++e) {
                const PushTask t = S.stasks[e];
                const int s = t.s, p = S.parent[s];
                const int64_t ns = S.ncol[s], ms = S.m[s];
                if (p < 0 || t.ns != ns || t.ms != ms || t.nsp != S.ncol[p] || t.msp != S.m[p] || t.lofs != S.lofs[s] ||
                    t.uofs != S.uofs[s] || t.relofs != S.relofs[s] || t.lofsp != S.lofs[p] || t.uofsp != S.uofs[p])
                    return 10;
                if (t.r0 < ns || t.c0 < ns || t.r0 < t.c0 || t.r0 >= ms) return 12;
                if (lev_of[p] <= l || lev_of[s] != l) return 14;
                const int64_t r1 = std::min<int64_t>(t.r0 + kTile, ms), c1 = std::min<int64_t>(t.c0 + kTile, ms);
                const double *Ls = L.data() + S.lofs[s];
                const int32_t *rel = S.relmap.data() + S.relofs[s] - ns;
                const int64_t nsp = S.ncol[p], msp = S.m[p];
                for (int64_t c = t.c0; c < c1; ++c)
                    for (int64_t r = std::max<int64_t>(t.r0, c); r < r1; ++r) {
                        double acc = 0;
                        for (int64_t k = 0; k < ns; ++k) acc += Ls[r + ms * k] * Ls[c + ms * k];
                        const double tv = U[S.uofs[s] + (r - ns) + (ms - ns) * (c - ns)] - acc;
                        const int64_t pr = rel[r], pc = rel[c];
                        if (pr < pc || pr >= msp || pc < 0) return 10;
                        const bool inL = pc < nsp;
                        const int64_t at = inL ? S.lofs[p] + pr + msp * pc : S.uofs[p] + (pr - nsp) + (msp - nsp) * (pc - nsp);
                        int64_t &stamp = inL ? Lstamp[at] : Ustamp[at];
                        if (stamp == launch) return 13;
                        stamp = launch;
                        (inL ? L : U)[at] += tv;
                    }
            }
        }
    }
    // solves, interleaved right-hand sides (v[i * nrhs + r])
    const int64_t n = S.n;
    const int64_t nw = std::max<int64_t>(S.upd_rows, 1);
    std::vector<double> X;
    // Dataflow solves: the tasks of each list run sequentially in list order, as
    // the kernels compute them (four partial sums per row in wave order, inverse
    // diagonal blocks).  Each wait must already be satisfied (a task may wait only
    // on tasks listed before it: returns 9), every index a kernel forms -- clamped
    // loads included -- must lie inside its array or block (10), every vector entry
    // is written once and read only after it is written (11), and children's rows
    // land inside the task's rows (8).
    {
        std::vector<double> Linv(std::max<int64_t>(S.inv_entries, 1), 0.0);
        // X = L_11^{-1} regions of the multi-panel supernodes: in range, disjoint
        // from the diagonal inverses; per block the launch that wrote it (lower
        // blocks: X, upper blocks: the transposed T of the inversion)
        std::vector<std::vector<int>> xw(S.ns);
        for (const FlowTask &t : S.flow_b) {
            if ((t.xofs >= 0) != (t.np >= 2)) return 10;
            if (t.xofs >= 0 && (t.xofs < S.iofs[S.ns] || t.xofs + t.ns * t.ns > S.inv_entries)) return 10;
            if (t.xofs >= 0 && xw[t.s].empty()) xw[t.s].assign((size_t)t.np * t.np, -1);
        }
        for (const FlowTask &t : S.flow_b) {
            if (t.kind != 0 && t.kind != 3) continue;      // one task per panel forms t_p
            const int b = t.nr;
            const int64_t k = t.r0, ms = t.ms;
            if (b < 1 || b > kPanel || t.iofs < 0 || t.iofs + (int64_t)b * b > S.inv_entries) return 10;
            if (k + b > t.ns || t.ns > ms) return 10;
            for (int c = 0; c < b; ++c) {
                std::vector<double> xc(b, 0.0);
                for (int i = 0; i < b; ++i) {
                    double v = i == c ? 1.0 : 0.0;
                    for (int j = 0; j < i; ++j) v -= L[t.lofs + (k + i) + ms * (k + j)] * xc[j];
                    xc[i] = v / L[t.lofs + (k + i) + ms * (k + i)];
                }
                for (int i = 0; i < b; ++i) {
                    Linv[t.iofs + i + (int64_t)b * c] = xc[i];
                    if (t.xofs >= 0) Linv[t.xofs + (k + i) + t.ns * (k + c)] = xc[i];
                }
            }
            if (t.xofs >= 0) xw[t.s][(size_t)t.p * t.np + t.p] = 0;
        }
        // the inversion launches (InvTask), sequentially, each tile as the kernel
        // forms it: loads clamped into the matrix, entries beyond ns masked; a
        // tile may read only blocks written by an earlier launch (9)
        {
            if (S.istep_off.empty() || (S.istep_off.size() - 1) % 2 != 0) return 10;
            std::vector<int64_t> snode_of_lofs;
            for (size_t q = 0; q + 1 < S.istep_off.size(); ++q) {
                const int ph = (int)(q % 2), launch = (int)q + 1;
                for (int64_t e = S.istep_off[q]; e < S.istep_off[q + 1]; ++e) {
                    const InvTask t = S.itasks[e];
                    int64_t s = -1;
                    for (int64_t c = 0; c < S.ns; ++c)
                        if (S.lofs[c] == t.lofs && (int64_t)t.ns == S.ncol[c]) { s = c; break; }
                    if (s < 0 || xw[s].empty()) return 10;
                    const int np = (int)((t.ns + kPanel - 1) / kPanel);
                    if (t.i <= t.j || t.i >= np || t.j < 0 || t.k0 >= t.k1 || t.k0 < 0 || t.k1 > np) return 10;
                    const int64_t ns = t.ns, ms = t.ms, r0 = (int64_t)kPanel * t.i, c0 = (int64_t)kPanel * t.j;
                    const int nr = (int)std::min<int64_t>(kPanel, ns - r0);
                    if (c0 + kPanel > ns) return 10;                 // panel j is never the partial one
                    std::vector<double> C((size_t)kTile * kTile, 0.0);
                    for (int k = t.k0; k < t.k1; ++k) {
                        const int64_t kc = (int64_t)kPanel * k;
                        const int kb = (int)std::min<int64_t>(kPanel, ns - kc);
                        if (ph == 0) {
                            if (k < t.j || xw[s][(size_t)k * np + t.j] < 0 || xw[s][(size_t)k * np + t.j] >= launch) return 9;
                        } else {
                            if (k > t.i || xw[s][(size_t)t.i * np + k] < 0 || xw[s][(size_t)t.i * np + k] >= launch) return 9;
                            if (xw[s][(size_t)t.j * np + k] < 0 || xw[s][(size_t)t.j * np + k] >= launch) return 9;
                        }
                        for (int y = 0; y < kTile; ++y)         // every load the kernel forms
                            for (int x = 0; x < kTile; ++x) {
                                const int64_t ra = r0 + std::min(x, nr - 1), ka = kc + std::min(y, kb - 1);
                                const int64_t ia = ph == 0 ? t.lofs + ra + ms * ka : t.xofs + ra + ns * ka;
                                const int64_t ib = ph == 0 ? t.xofs + (kc + std::min(x, kb - 1)) + ns * (c0 + y)
                                                           : t.xofs + (c0 + x) + ns * ka;
                                if (ph == 0 && (ia < t.lofs || ia >= t.lofs + ms * ns || ia >= S.L_entries)) return 10;
                                if (ph == 1 && (ia < t.xofs || ia >= t.xofs + ns * ns)) return 10;
                                if (ib < t.xofs || ib >= t.xofs + ns * ns) return 10;
                            }
                        for (int c = 0; c < kTile; ++c)
                            for (int r = 0; r < nr; ++r) {
                                double a = 0;
                                for (int kk = 0; kk < kb; ++kk) {
                                    const double av = ph == 0 ? L[t.lofs + (r0 + r) + ms * (kc + kk)]
                                                              : Linv[t.xofs + (r0 + r) + ns * (kc + kk)];
                                    const double bv = ph == 0 ? Linv[t.xofs + (kc + kk) + ns * (c0 + c)]
                                                              : Linv[t.xofs + (c0 + c) + ns * (kc + kk)];
                                    a += av * bv;
                                }
                                C[r + (size_t)kTile * c] += a;
                            }
                    }
                    for (int c = 0; c < kTile; ++c)
                        for (int r = 0; r < nr; ++r) {
                            if (ph == 0) Linv[t.xofs + (c0 + c) + ns * (r0 + r)] = C[r + (size_t)kTile * c];
                            else Linv[t.xofs + (r0 + r) + ns * (c0 + c)] = -C[r + (size_t)kTile * c];
                        }
                    int &f = ph == 0 ? xw[s][(size_t)t.j * np + t.i] : xw[s][(size_t)t.i * np + t.j];
                    if (f >= 0 && ph == 1) return 11;                // each X block written once
                    f = launch;
                }
            }
            for (int64_t s = 0; s < S.ns; ++s)                   // every lower block of every X written
                if (!xw[s].empty()) {
                    const int np = (int)((S.ncol[s] + kPanel - 1) / kPanel);
                    for (int i = 0; i < np; ++i)
                        for (int j = 0; j <= i; ++j)
                            if (xw[s][(size_t)i * np + j] < 0) return 11;
                }
        }
        const int64_t nsn = S.ns;
        std::vector<int32_t> ready(nsn, 0), zdone(nsn, 0), vdone(nsn, 0);
        std::vector<double> Y2(n * nrhs), W2(nw * nrhs, 0.0), Z2(n * nrhs, 0.0), X2(n * nrhs, 0.0);
        std::vector<double> V2(n * nrhs, 0.0), T2(n * nrhs, 0.0);
        std::vector<char> wW2(nw, 0), wZ2(n, 0), wX2(n, 0), wV2(n, 0), wT2(n, 0);
        // partial sums of the chunked block sums: value, written flag, (s, r0, k0, k1)
        const int64_t npf = std::max<int64_t>(S.flow_f_parts, 1), npb = std::max<int64_t>(S.flow_b_parts, 1);
        std::vector<double> PF2(npf * kTile * nrhs, 0.0), PB2(npb * kTile * nrhs, 0.0);
        std::vector<char> wPF(npf, 0), wPB(npb, 0);
        std::vector<int64_t> idPF(4 * npf, -1), idPB(4 * npb, -1);
        // a partial task's result into its slot (written once)
        auto put_part = [&](const FlowTask &t, int64_t nparts, std::vector<double> &P2, std::vector<char> &wP,
                            std::vector<int64_t> &id, const std::vector<double> &acc) -> int {
            if (t.pofs < 0 || t.pofs >= nparts || wP[t.pofs]) return 11;
            wP[t.pofs] = 1;
            const bool rows = t.u1 > t.u0;                // backward: over update rows, else over blocks
            id[4 * t.pofs] = t.s; id[4 * t.pofs + 1] = t.r0;
            id[4 * t.pofs + 2] = rows ? t.u0 : t.k0; id[4 * t.pofs + 3] = rows ? t.u1 : t.k1;
            const int64_t st = (int64_t)kTile * nrhs;
            for (int i = 0; i < t.nr; ++i)
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[i * nrhs + r];
                    P2[(t.pofs * kTile + i) * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                }
            return 0;
        };
        // a consumer's partials: slots [pofs, pofs + npart), written, of its rows,
        // covering the blocks [lo, hi) in order; their sums added to out (nr x nrhs)
        auto add_parts = [&](const FlowTask &t, int64_t nparts, const std::vector<double> &P2, const std::vector<char> &wP,
                             const std::vector<int64_t> &id, int32_t lo, int32_t hi, std::vector<double> &out) -> int {
            if (t.npart < 0 || (t.npart > 0 && (t.pofs < 0 || t.pofs + t.npart > nparts))) return 10;
            int64_t at = lo;
            for (int u = 0; u < t.npart; ++u) {
                const int64_t q = t.pofs + u;
                if (!wP[q]) return 9;
                if (id[4 * q] != t.s || id[4 * q + 1] != t.r0 || id[4 * q + 2] != at) return 10;
                at = id[4 * q + 3];
                for (int i = 0; i < t.nr; ++i)
                    for (int r = 0; r < nrhs; ++r) out[i * nrhs + r] += P2[(q * kTile + i) * nrhs + r];
            }
            return at == hi ? 0 : 10;
        };
        for (int64_t i = 0; i < n; ++i)
            for (int r = 0; r < nrhs; ++r) Y2[i * nrhs + r] = b[S.perm[i] + ndof * r];
        // an L index formed for block (s, rows [r0, r0+nr), cols [c, c+nc)), clamped as load_cols16 does
        auto Lblock_ok = [&](const FlowTask &t, int64_t r0, int nr, int64_t c, int nc) {
            if (nr < 1 || nc < 1) return false;
            const int64_t lo = t.lofs + r0 + t.ms * c, hi = lo + (nr - 1) + t.ms * (nc - 1);
            return r0 >= 0 && r0 + nr <= t.ms && c >= 0 && c + nc <= t.ns && lo >= t.lofs && hi < t.lofs + t.ms * t.ns &&
                   hi < S.L_entries;
        };
        // bottom subtrees, run one or two wavefronts per task: the head of the
        // forward list and the tail of the backward list, one task per supernode
        // (one panel, a whole front of at most kWholeRows rows), whose children
        // are all bottom supernodes (12)
        {
            if (S.flow_f_leaf < 0 || S.flow_f_leaf > (int64_t)S.flow_f.size() || S.flow_b_leaf != S.flow_f_leaf) return 12;
            std::vector<char> leaf(nsn, 0);
            for (int64_t e = 0; e < S.flow_f_leaf; ++e) {
                const FlowTask &t = S.flow_f[e];
                if (t.kind != 1 || t.p != 0 || t.np != 1 || t.ms > kWholeRows || leaf[t.s]) return 12;
                leaf[t.s] = 1;
            }
            for (int64_t e = S.flow_f_leaf; e < (int64_t)S.flow_f.size(); ++e)
                if (leaf[S.flow_f[e].s]) return 12;
            const int64_t b0 = (int64_t)S.flow_b.size() - S.flow_b_leaf;
            for (int64_t e = 0; e < (int64_t)S.flow_b.size(); ++e) {
                const FlowTask &t = S.flow_b[e];
                if ((e >= b0) != (leaf[t.s] != 0)) return 12;
                if (e >= b0 && (t.np != 1 || t.p != 0 || t.xofs >= 0 || t.u0 != t.ns || t.u1 != t.ms)) return 12;
            }
            for (int64_t s = 0; s < nsn; ++s)
                if (leaf[s])
                    for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q)
                        if (!leaf[S.child[q]]) return 12;
        }
        for (size_t e = 0; e < S.flow_f.size(); ++e) {
            const FlowTask &t = S.flow_f[e];
            const bool pivot = t.p >= 0;
            const bool multi = pivot && t.xofs >= 0;     // z_p through X and the v of the earlier panels
            if (multi != (pivot && t.kind != 1 && t.np >= 2)) return 10;
            const int kend = pivot ? t.p : t.np;
            const bool part = t.kind == 2;
            if (t.kind < 0 || t.kind > 2) return 10;
            if (part && !(t.k0 >= 0 && t.k0 < t.k1 && t.k1 - t.k0 <= kSolveChunk && t.k1 <= kend - kSolveChunk))
                return 10;
            if (t.kind == 0 && (t.k1 != kend || t.k0 != solve_own_first(kend, kSolveSplit))) return 10;
            if (part && (t.need != 0 || t.nch != 0)) return 10;
            const int nr = t.nr;
            if (t.s < 0 || t.s >= nsn || nr < 1 || nr > kTile) return 10;
            if (pivot && (t.r0 != kPanel * t.p || t.r0 + nr > t.ns)) return 10;
            if (!pivot && (t.r0 < t.ns || t.r0 + nr > t.ms || t.parent < 0)) return 10;
            const bool whole = t.kind == 1;
            if (whole && (t.p != 0 || t.np != 1 || t.ms > kWholeRows || nr != t.ns)) return 10;
            std::vector<double> vs(kWholeRows * nrhs, 0.0);
            if (pivot)
                for (int i = 0; i < nr; ++i)
                    for (int r = 0; r < nrhs; ++r) vs[i * nrhs + r] = Y2[(t.c0 + t.r0 + i) * nrhs + r];
            if (t.need > 0) {
                if (ready[t.s] < t.need) return 9;
                if (t.gb < 0 || t.gb + t.nch > (int64_t)S.flow_gb.size()) return 10;
                for (int j = 0; j < t.nch; ++j) {
                    const FlowChild ch = S.flow_gb[t.gb + j];
                    for (int64_t i = ch.i0; i < ch.i1; ++i) {
                        if (ch.relofs + i < 0 || ch.relofs + i >= (int64_t)S.relmap.size() || ch.relofs + i >= nw) return 10;
                        const int64_t row = S.relmap[ch.relofs + i] - t.r0;
                        if (row < 0 || row >= (whole ? t.ms : nr)) return 8;
                        if (!wW2[ch.relofs + i]) return 11;
                        for (int r = 0; r < nrhs; ++r) vs[row * nrhs + r] += W2[(ch.relofs + i) * nrhs + r];
                    }
                }
            }
            if (whole) {
                // z = L11^{-1} v[0:ns] (four partial sums), then the update rows
                if (t.iofs < 0 || t.iofs + (int64_t)nr * nr > S.inv_entries) return 10;
                std::vector<double> acc(4 * kTile * nrhs, 0.0), z(kTile * nrhs, 0.0);
                for (int w = 0; w < 4; ++w)
                    for (int i = 0; i < nr; ++i)
                        for (int c = 16 * w; c < std::min(16 * w + 16, nr); ++c)
                            for (int r = 0; r < nrhs; ++r)
                                acc[(w * kTile + i) * nrhs + r] += Linv[t.iofs + i + (int64_t)nr * c] * vs[c * nrhs + r];
                for (int i = 0; i < nr; ++i) {
                    const int64_t zi = t.c0 + i;
                    if (wZ2[zi]) return 11;
                    wZ2[zi] = 1;
                    for (int r = 0; r < nrhs; ++r) {
                        const double *a = &acc[i * nrhs + r];
                        const int64_t st = (int64_t)kTile * nrhs;
                        z[i * nrhs + r] = Z2[zi * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                    }
                }
                for (int64_t i = t.ns; i < t.ms; ++i) {
                    const int64_t wi = t.relofs + i - t.ns;
                    if (wi < 0 || wi >= nw || wW2[wi]) return 11;
                    wW2[wi] = 1;
                    for (int r = 0; r < nrhs; ++r) {
                        double a = 0;
                        for (int c = 0; c < nr; ++c) a += L[t.lofs + i + t.ms * c] * z[c * nrhs + r];
                        W2[wi * nrhs + r] = vs[i * nrhs + r] - a;
                    }
                }
                zdone[t.s] = 1;
                if (t.parent >= 0) ready[t.parent] += 1;
                continue;
            }
            if (multi && !part) {                        // v_p published for the later panels of s
                if (vdone[t.s] != t.p) return 9;
                for (int i = 0; i < nr; ++i) {
                    const int64_t vi = t.c0 + t.r0 + i;
                    if (wV2[vi]) return 11;
                    wV2[vi] = 1;
                    for (int r = 0; r < nrhs; ++r) V2[vi * nrhs + r] = vs[i * nrhs + r];
                }
                vdone[t.s] = t.p + 1;
            }
            std::vector<double> acc(4 * kTile * nrhs, 0.0);
            for (int k = t.k0; k < t.k1; ++k) {
                const int64_t col = (int64_t)kPanel * k;
                const int bk = (int)std::min<int64_t>(kPanel, t.ns - col);
                if (multi) {
                    if (vdone[t.s] < k + 1) return 9;
                    const int64_t lo = t.xofs + t.r0 + t.ns * col, hi = lo + (nr - 1) + t.ns * (bk - 1);
                    if (lo < t.xofs || hi >= t.xofs + t.ns * t.ns || k >= t.p) return 10;
                } else {
                    if (zdone[t.s] < k + 1) return 9;
                    if (!Lblock_ok(t, t.r0, nr, col, bk)) return 10;
                }
                const std::vector<double> &src = multi ? V2 : Z2;
                for (int c = 0; c < bk; ++c)
                    if (!(multi ? wV2 : wZ2)[t.c0 + col + c]) return 11;
                for (int w = 0; w < 4; ++w)
                    for (int i = 0; i < nr; ++i)
                        for (int c = 16 * w; c < std::min(16 * w + 16, bk); ++c) {
                            const double m = multi ? Linv[t.xofs + (t.r0 + i) + t.ns * (col + c)]
                                                   : L[t.lofs + (t.r0 + i) + t.ms * (col + c)];
                            for (int r = 0; r < nrhs; ++r)
                                acc[(w * kTile + i) * nrhs + r] += m * src[(t.c0 + col + c) * nrhs + r];
                        }
            }
            if (part) {
                if (const int rc = put_part(t, S.flow_f_parts, PF2, wPF, idPF, acc)) return rc;
                continue;
            }
            if (pivot && (t.iofs < 0 || t.iofs + (int64_t)nr * nr > S.inv_entries)) return 10;
            std::vector<double> sum(kTile * nrhs, 0.0);   // the block sum, then the partials of [0, k0)
            if (multi)                                   // + X_pp v_p, one reduction for the whole sum
                for (int w = 0; w < 4; ++w)
                    for (int i = 0; i < nr; ++i)
                        for (int c = 16 * w; c < std::min(16 * w + 16, nr); ++c)
                            for (int r = 0; r < nrhs; ++r)
                                acc[(w * kTile + i) * nrhs + r] += Linv[t.iofs + i + (int64_t)nr * c] * vs[c * nrhs + r];
            for (int i = 0; i < nr; ++i)
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[i * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    sum[i * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                }
            if (const int rc = add_parts(t, S.flow_f_parts, PF2, wPF, idPF, 0, t.k0, sum)) return rc;
            if (multi) {
                for (int i = 0; i < nr; ++i) {
                    const int64_t zi = t.c0 + t.r0 + i;
                    if (wZ2[zi]) return 11;
                    wZ2[zi] = 1;
                    for (int r = 0; r < nrhs; ++r) Z2[zi * nrhs + r] = sum[i * nrhs + r];
                }
                zdone[t.s] += 1;
                continue;
            }
            std::vector<double> v(kTile * nrhs, 0.0);
            for (int i = 0; i < nr; ++i)
                for (int r = 0; r < nrhs; ++r) v[i * nrhs + r] = vs[i * nrhs + r] - sum[i * nrhs + r];
            if (!pivot) {
                for (int i = 0; i < nr; ++i) {
                    const int64_t wi = t.relofs + t.r0 - t.ns + i;
                    if (wi < 0 || wi >= nw || wW2[wi]) return 11;
                    wW2[wi] = 1;
                    for (int r = 0; r < nrhs; ++r) W2[wi * nrhs + r] = v[i * nrhs + r];
                }
                ready[t.parent] += 1;
                continue;
            }
            // one-panel supernode: z = L_11^{-1} v
            if (t.np != 1 || t.p != 0) return 10;
            std::fill(acc.begin(), acc.end(), 0.0);
            for (int w = 0; w < 4; ++w)
                for (int i = 0; i < nr; ++i)
                    for (int c = 16 * w; c < std::min(16 * w + 16, nr); ++c)
                        for (int r = 0; r < nrhs; ++r)
                            acc[(w * kTile + i) * nrhs + r] += Linv[t.iofs + i + (int64_t)nr * c] * v[c * nrhs + r];
            for (int i = 0; i < nr; ++i) {
                const int64_t zi = t.c0 + t.r0 + i;
                if (wZ2[zi]) return 11;
                wZ2[zi] = 1;
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[i * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    Z2[zi * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                }
            }
            zdone[t.s] = 1;
        }
        for (int64_t i = 0; i < n; ++i)
            if (!wZ2[i]) return 11;
        std::vector<int32_t> tdone(nsn, 0);
        for (size_t e = 0; e < S.flow_b.size(); ++e) {
            const FlowTask &t = S.flow_b[e];
            const int nc = t.nr;
            if (t.parent >= 0 && t.pnp != (int32_t)((S.ncol[t.parent] + kPanel - 1) / kPanel)) return 10;
            const bool multi = t.xofs >= 0;              // x_p through X and the t of the later panels
            if (multi != (t.np >= 2)) return 10;
            const bool tphase = t.kind == 0 || t.kind == 3;
            const bool tpart = t.kind == 2 && t.u1 > t.u0;   // partial sum over far update rows
            const int later = t.np - 1 - t.p;            // panels after p
            const int nut = (int)((t.ms - t.ns + kTile - 1) / kTile);
            const bool whole = t.ns <= kPanel && t.ms <= kWholeRows;   // never split
            const int64_t urow1 = !whole && solve_own_first(nut, kSolveSplitRows) > 0 ? t.ns + (int64_t)kTile * kSolveChunk : t.ms;
            if (tphase && (t.u0 != t.ns || t.u1 != urow1)) return 10;
            if (!tphase && !tpart && t.u0 != t.u1) return 10;
            if (tpart && !(t.u0 >= urow1 && t.u1 <= t.ms && t.u1 - t.u0 <= kTile * kSolveChunk && t.k0 == t.k1 &&
                           (t.u0 - t.ns) % kTile == 0))
                return 10;
            if (t.kind == 0 && (solve_own_first(later, kSolveSplit) > 0 || t.k0 != t.p + 1 || t.k1 != t.np)) return 10;
            if (t.kind == 3 && (solve_own_first(later, kSolveSplit) == 0 || !multi || t.k0 != t.k1)) return 10;
            if (t.kind == 2 && !tpart &&
                !(multi && t.k0 >= t.p + 1 + kSolveChunk && t.k0 < t.k1 && t.k1 <= t.np && t.k1 - t.k0 <= kSolveChunk))
                return 10;
            if (t.kind == 4 && (solve_own_first(later, kSolveSplit) == 0 || t.k0 != t.p + 1 || t.k1 != t.p + 1 + kSolveChunk)) return 10;
            if (t.kind == 1 || t.kind < 0 || t.kind > 4) return 10;
            std::vector<double> acc(4 * kTile * nrhs, 0.0);
            std::vector<double> v(kTile * nrhs, 0.0);     // t_p
            if (tphase || tpart) {
                if (tphase) {
                    if (tdone[t.s] != later) return 9;   // the panels' t tasks from the last panel
                    tdone[t.s] += 1;
                }
                const int nchunk = (t.u1 - t.u0 + kTile - 1) / kTile;
                for (int e2 = 0; e2 < nchunk; ++e2) {  // update rows (ancestors' values), from the last tile
                    const int64_t row0 = t.u0 + (int64_t)kTile * (nchunk - 1 - e2);
                    const int nre = (int)std::min<int64_t>(kTile, (int64_t)t.u1 - row0);
                    const int64_t lo = t.lofs + row0 + t.ms * t.r0, hi = lo + (nre - 1) + t.ms * (nc - 1);
                    if (!(row0 >= 0 && row0 + nre <= t.ms && hi < t.lofs + t.ms * t.ns)) return 10;
                    for (int i = 0; i < nre; ++i) {
                        if (t.rofs + row0 + i >= (int64_t)S.rows.size()) return 10;
                        const int64_t g = S.rows[t.rofs + row0 + i];
                        if (g < 0 || g >= n || !wX2[g]) return 11;
                        for (int c = 0; c < nc; ++c) {
                            const int w = i / 16;
                            for (int r = 0; r < nrhs; ++r)
                                acc[(w * kTile + c) * nrhs + r] += L[t.lofs + (row0 + i) + t.ms * (t.r0 + c)] * X2[g * nrhs + r];
                        }
                    }
                }
                if (tpart) {
                    if (const int rc = put_part(t, S.flow_b_parts, PB2, wPB, idPB, acc)) return rc;
                    continue;
                }
                std::vector<double> us(kTile * nrhs, 0.0);   // update-row sum, then the far rows' partials
                for (int c = 0; c < nc; ++c)
                    for (int r = 0; r < nrhs; ++r) {
                        const double *a = &acc[c * nrhs + r];
                        const int64_t st = (int64_t)kTile * nrhs;
                        us[c * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                    }
                if (const int rc = add_parts(t, S.flow_b_parts, PB2, wPB, idPB, t.u1, (int32_t)t.ms, us)) return rc;
                for (int c = 0; c < nc; ++c)
                    for (int r = 0; r < nrhs; ++r) v[c * nrhs + r] = Z2[(t.c0 + t.r0 + c) * nrhs + r] - us[c * nrhs + r];
                if (multi)
                    for (int c = 0; c < nc; ++c) {
                        const int64_t ti = t.c0 + t.r0 + c;
                        if (wT2[ti]) return 11;
                        wT2[ti] = 1;
                        for (int r = 0; r < nrhs; ++r) T2[ti * nrhs + r] = v[c * nrhs + r];
                    }
                if (t.kind == 3) continue;
                std::fill(acc.begin(), acc.end(), 0.0);
            } else if (t.kind == 4) {                    // t_p as the panel's t task published it
                for (int c = 0; c < nc; ++c) {
                    const int64_t ti = t.c0 + t.r0 + c;
                    if (!wT2[ti]) return 11;
                    for (int r = 0; r < nrhs; ++r) v[c * nrhs + r] = T2[ti * nrhs + r];
                }
            }
            if (multi)
                for (int q = t.k1 - 1; q >= t.k0; --q) { // later panels of X, from the last
                    const int64_t row0 = (int64_t)kPanel * q;
                    const int nre = (int)std::min<int64_t>(kTile, t.ns - row0);
                    const int64_t lo = t.xofs + row0 + t.ns * t.r0, hi = lo + (nre - 1) + t.ns * (nc - 1);
                    if (lo < t.xofs || hi >= t.xofs + t.ns * t.ns) return 10;
                    for (int i = 0; i < nre; ++i) {
                        const int64_t g = t.c0 + row0 + i;
                        if (!wT2[g]) return 11;
                        for (int c = 0; c < nc; ++c)
                            for (int r = 0; r < nrhs; ++r)
                                acc[((i / 16) * kTile + c) * nrhs + r] +=
                                    Linv[t.xofs + (row0 + i) + t.ns * (t.r0 + c)] * T2[g * nrhs + r];
                    }
                }
            if (t.kind == 2) {
                if (const int rc = put_part(t, S.flow_b_parts, PB2, wPB, idPB, acc)) return rc;
                continue;
            }
            for (int i = 0; i < nc; ++i)
                for (int c = 0; c < nc; ++c)
                    for (int r = 0; r < nrhs; ++r)
                        acc[((i / 16) * kTile + c) * nrhs + r] += Linv[t.iofs + i + (int64_t)nc * c] * v[i * nrhs + r];
            std::vector<double> sum(kTile * nrhs, 0.0);
            for (int c = 0; c < nc; ++c)
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[c * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    sum[c * nrhs + r] = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                }
            if (t.kind == 4)
                if (const int rc = add_parts(t, S.flow_b_parts, PB2, wPB, idPB, t.k1, t.np, sum)) return rc;
            for (int c = 0; c < nc; ++c) {
                const int64_t xi = t.c0 + t.r0 + c;
                if (wX2[xi]) return 11;
                wX2[xi] = 1;
                for (int r = 0; r < nrhs; ++r) X2[xi * nrhs + r] = sum[c * nrhs + r];
            }
        }
        for (int64_t i = 0; i < n; ++i)
            if (!wX2[i]) return 11;
        X.swap(X2);
    }
    for (int64_t i = 0; i < ndof * nrhs; ++i) x[i] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int r = 0; r < nrhs; ++r) x[S.perm[i] + ndof * r] = X[i * nrhs + r];
    info[0] = S.ns;
    info[1] = nlev;
    info[2] = S.L_entries;
    info[3] = S.U_entries;
    info[4] = (int64_t)S.ptasks.size();
    info[5] = (int64_t)S.utasks.size();
    info[6] = S.max_front;
    info[7] = S.flow_f_leaf;
    info[8] = S.flow_f_parts;
    info[9] = S.flow_b_parts;
    return 0;
}
