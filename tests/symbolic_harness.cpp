// CPU harness for the symbolic analysis (test infrastructure).
//
// Runs jfem::build_symbolic and then executes, sequentially on the CPU, the same
// task decomposition the device kernels execute (scatter, extend-add tasks, panel
// tasks in two phases per step, update tiles, and the dataflow solve lists in
// their order), so the structures and task lists can be checked against a dense
// solve without a GPU.
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
    for (size_t i = 0; i < S.amap_src.size(); ++i) L[S.amap_dst[i]] = val[S.amap_src[i]];
    const int64_t nlev = (int64_t)S.level_ptr.size() - 1;
    for (int64_t l = 0; l < nlev; ++l) {
        for (int64_t q = S.ea_off[l]; q < S.ea_off[l + 1]; ++q) {
            EaTask t = S.etasks[q];
            int s = t.s;
            int64_t ns = S.ncol[s], ms = S.m[s], mu = ms - ns;
            for (int32_t qc = S.child_ptr[s]; qc < S.child_ptr[s + 1]; ++qc) {
                int c = S.child[qc];
                const int32_t *rel = S.relmap.data() + S.relofs[c];
                const int32_t *bd = S.ea_bounds.data() + 4 * (S.ea_boff[q] + (qc - S.child_ptr[s]));
                for (int64_t j = bd[0]; j < bd[1]; ++j) {
                    int64_t pc = rel[j];
                    if (pc < t.c0 || pc >= t.c1) return 6;
                    for (int64_t i = std::max<int64_t>(j, bd[2]); i < bd[3]; ++i) {
                        int64_t pr = rel[i];
                        if (pr < t.r0 || pr >= t.r0 + kEaRows) return 7;
                        int64_t nu = S.m[c] - S.ncol[c];
                        double v = U[S.uofs[c] + i + nu * j];
                        if (pc < ns) L[S.lofs[s] + pr + ms * pc] += v;
                        else U[S.uofs[s] + (pr - ns) + mu * (pc - ns)] += v;
                    }
                }
            }
        }
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
        // after the level's last step: the update matrices, full rank per front
        for (int64_t q = S.sstep_off[l]; q < S.sstep_off[l + 1]; ++q) {
            UpdateTask t = S.stasks[q];
            int s = t.s;
            int64_t ns = S.ncol[s], ms = S.m[s];
            if (t.r0 < ns || t.c0 < ns || t.r0 < t.c0 || t.r0 >= ms) return 12;
            int64_t r1 = std::min<int64_t>(t.r0 + kTile, ms), c1 = std::min<int64_t>(t.c0 + kTile, ms);
            const double *Ls = L.data() + S.lofs[s];
            for (int64_t c = t.c0; c < c1; ++c)
                for (int64_t r = std::max<int64_t>(t.r0, c); r < r1; ++r) {
                    double acc = 0;
                    for (int64_t p = 0; p < ns; ++p) acc += Ls[r + ms * p] * Ls[c + ms * p];
                    U[S.uofs[s] + (r - ns) + (ms - ns) * (c - ns)] -= acc;
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
        for (const FlowTask &t : S.flow_b) {
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
                for (int i = 0; i < b; ++i) Linv[t.iofs + i + (int64_t)b * c] = xc[i];
            }
        }
        // chain products: boundary q of s (backward task (s, q), q < np-1):
        // M_{q+1} = L_{q+1,q+1}^{-1} A at bofs - 4096, G_q = (A L_qq^{-1})^T at bofs,
        // A = L[q+1 rows, q cols], both 64 x 64 column-major
        const int64_t PP = (int64_t)kPanel * kPanel;
        for (const FlowTask &t : S.flow_b) {
            if ((t.bofs >= 0) != (t.p + 1 < t.np)) return 10;
            if (t.bofs < 0) continue;
            if (t.bofs - PP < 0 || t.bofs + PP > S.inv_entries || t.nr != kPanel) return 10;
            const int b1 = (int)std::min<int64_t>(kPanel, t.ns - (int64_t)kPanel * (t.p + 1));
            if (t.iofs + PP + (int64_t)b1 * b1 > S.inv_entries) return 10;
            auto A = [&](int k, int m) { return k < b1 ? L[t.lofs + (kPanel * (t.p + 1) + k) + t.ms * (kPanel * t.p + m)] : 0.0; };
            for (int c = 0; c < kPanel; ++c)
                for (int i = 0; i < kPanel; ++i) {
                    double sm = 0, sg = 0;
                    if (i < b1)
                        for (int m = 0; m <= i; ++m) sm += Linv[t.iofs + PP + i + (int64_t)b1 * m] * A(m, c);
                    if (c < b1)
                        for (int m = i; m < kPanel; ++m) sg += A(c, m) * Linv[t.iofs + m + (int64_t)kPanel * i];
                    Linv[t.bofs - PP + i + kPanel * c] = sm;
                    Linv[t.bofs + i + kPanel * c] = sg;
                }
        }
        const int64_t nsn = S.ns;
        std::vector<int32_t> ready(nsn, 0), zdone(nsn, 0), xdone(nsn, 0);
        std::vector<double> Y2(n * nrhs), W2(nw * nrhs, 0.0), Z2(n * nrhs, 0.0), X2(n * nrhs, 0.0);
        std::vector<char> wW2(nw, 0), wZ2(n, 0), wX2(n, 0);
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
                if (e >= b0 && (t.np != 1 || t.p != 0 || t.bofs >= 0 || t.u0 != t.ns || t.u1 != t.ms)) return 12;
            }
            for (int64_t s = 0; s < nsn; ++s)
                if (leaf[s])
                    for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1]; ++q)
                        if (!leaf[S.child[q]]) return 12;
        }
        for (size_t e = 0; e < S.flow_f.size(); ++e) {
            const FlowTask &t = S.flow_f[e];
            const bool pivot = t.p >= 0;
            const bool chain = pivot && t.bofs >= 0;
            if (chain != (pivot && t.kind != 1 && t.p >= 1)) return 10;
            if (chain && (t.bofs < 0 || t.bofs + PP > S.inv_entries)) return 10;
            const int kend = pivot ? (chain ? t.p - 1 : t.p) : t.np;
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
            std::vector<double> acc(4 * kTile * nrhs, 0.0);
            for (int k = 0; k < kend; ++k) {
                if (zdone[t.s] < k + 1) return 9;
                const int64_t col = (int64_t)kPanel * k;
                const int bk = (int)std::min<int64_t>(kPanel, t.ns - col);
                if (!Lblock_ok(t, t.r0, nr, col, bk)) return 10;
                for (int c = 0; c < bk; ++c)
                    if (!wZ2[t.c0 + col + c]) return 11;
                for (int w = 0; w < 4; ++w)
                    for (int i = 0; i < nr; ++i)
                        for (int c = 16 * w; c < std::min(16 * w + 16, bk); ++c)
                            for (int r = 0; r < nrhs; ++r)
                                acc[(w * kTile + i) * nrhs + r] +=
                                    L[t.lofs + (t.r0 + i) + t.ms * (col + c)] * Z2[(t.c0 + col + c) * nrhs + r];
            }
            std::vector<double> v(kTile * nrhs, 0.0);
            for (int i = 0; i < nr; ++i)
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[i * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    v[i * nrhs + r] = vs[i * nrhs + r] - (((a[0] + a[st]) + a[2 * st]) + a[3 * st]);
                }
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
            if (t.iofs < 0 || t.iofs + (int64_t)nr * nr > S.inv_entries) return 10;
            if (zdone[t.s] != t.p) return 9;
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
                    double z = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                    if (chain)      // z_p = u - M_p z_{p-1}
                        for (int c = 0; c < kPanel; ++c) {
                            const int64_t zc = t.c0 + t.r0 - kPanel + c;
                            if (!wZ2[zc]) return 11;
                            z -= Linv[t.bofs + i + kPanel * c] * Z2[zc * nrhs + r];
                        }
                    Z2[zi * nrhs + r] = z;
                }
            }
            zdone[t.s] = t.p + 1;
        }
        for (int64_t i = 0; i < n; ++i)
            if (!wZ2[i]) return 11;
        for (size_t e = 0; e < S.flow_b.size(); ++e) {
            const FlowTask &t = S.flow_b[e];
            const int nc = t.nr;
            if (t.parent >= 0 && t.pnp != (int32_t)((S.ncol[t.parent] + kPanel - 1) / kPanel)) return 10;
            if (t.u0 != t.ns || t.u1 != t.ms) return 10;
            std::vector<double> acc(4 * kTile * nrhs, 0.0);
            const int nchunk = (t.u1 - t.u0 + kTile - 1) / kTile;
            const int ntile = nchunk + (t.np - 1 - t.p);
            const bool chain = t.bofs >= 0;              // panel p+1 (the last tile) through G_p
            for (int e2 = 0; e2 < (chain ? ntile - 1 : ntile); ++e2) {
                const bool upd = e2 < nchunk;
                const int q = upd ? -1 : t.np - 1 - (e2 - nchunk);
                const int64_t row0 = upd ? t.u0 + (int64_t)kTile * (nchunk - 1 - e2) : (int64_t)kPanel * q;
                const int nre = (int)std::min<int64_t>(kTile, (upd ? (int64_t)t.u1 : t.ns) - row0);
                if (!upd && xdone[t.s] < t.np - q) return 9;
                if (!Lblock_ok(t, row0, nre, t.r0, nc)) {
                    // update rows lie below the pivot columns: check against the whole front
                    const int64_t lo = t.lofs + row0 + t.ms * t.r0, hi = lo + (nre - 1) + t.ms * (nc - 1);
                    if (!(row0 >= 0 && row0 + nre <= t.ms && hi < t.lofs + t.ms * t.ns)) return 10;
                }
                for (int i = 0; i < nre; ++i) {
                    if (upd && (t.rofs + row0 + i >= (int64_t)S.rows.size())) return 10;
                    const int64_t g = upd ? S.rows[t.rofs + row0 + i] : t.c0 + row0 + i;
                    if (g < 0 || g >= n || !wX2[g]) return 11;
                    for (int c = 0; c < nc; ++c) {
                        const int w = i / 16;
                        for (int r = 0; r < nrhs; ++r)
                            acc[(w * kTile + c) * nrhs + r] += L[t.lofs + (row0 + i) + t.ms * (t.r0 + c)] * X2[g * nrhs + r];
                    }
                }
            }
            std::vector<double> v(kTile * nrhs, 0.0);
            for (int c = 0; c < nc; ++c)
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[c * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    v[c * nrhs + r] = Z2[(t.c0 + t.r0 + c) * nrhs + r] - (((a[0] + a[st]) + a[2 * st]) + a[3 * st]);
                }
            std::fill(acc.begin(), acc.end(), 0.0);
            for (int i = 0; i < nc; ++i)
                for (int c = 0; c < nc; ++c)
                    for (int r = 0; r < nrhs; ++r)
                        acc[((i / 16) * kTile + c) * nrhs + r] += Linv[t.iofs + i + (int64_t)nc * c] * v[i * nrhs + r];
            if (xdone[t.s] != t.np - 1 - t.p) return 9;
            const int b1 = chain ? (int)std::min<int64_t>(kPanel, t.ns - t.r0 - kPanel) : 0;
            for (int c = 0; c < nc; ++c) {
                const int64_t xi = t.c0 + t.r0 + c;
                if (wX2[xi]) return 11;
                wX2[xi] = 1;
                for (int r = 0; r < nrhs; ++r) {
                    const double *a = &acc[c * nrhs + r];
                    const int64_t st = (int64_t)kTile * nrhs;
                    double x = ((a[0] + a[st]) + a[2 * st]) + a[3 * st];
                    for (int k = 0; k < b1; ++k) {     // x_p = u - G_p x_{p+1}
                        const int64_t xk = t.c0 + t.r0 + kPanel + k;
                        if (!wX2[xk]) return 11;
                        x -= Linv[t.bofs + c + kPanel * k] * X2[xk * nrhs + r];
                    }
                    X2[xi * nrhs + r] = x;
                }
            }
            xdone[t.s] += 1;
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
    return 0;
}
