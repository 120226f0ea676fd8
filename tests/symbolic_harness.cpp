// CPU harness for the symbolic analysis (test infrastructure).
//
// Runs jfem::build_symbolic and then executes, sequentially on the CPU, the same
// task decomposition the device kernels execute (scatter, extend-add tasks, panel
// tasks in two phases per step, update tiles, and the level-scheduled solves), so
// the structures and task lists can be checked against a dense solve without a GPU.
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
                int64_t nu = S.m[c] - S.ncol[c];
                const int32_t *rel = S.relmap.data() + S.relofs[c];
                for (int64_t j = 0; j < nu; ++j) {
                    int64_t pc = rel[j];
                    if (pc < t.c0 || pc >= t.c1) continue;
                    for (int64_t i = j; i < nu; ++i) {
                        int64_t pr = rel[i];
                        double v = U[S.uofs[c] + i + nu * j];
                        if (pc < ns) L[S.lofs[s] + pr + ms * pc] += v;
                        else U[S.uofs[s] + (pr - ns) + mu * (pc - ns)] += v;
                    }
                }
            }
        }
        for (int64_t stp = S.step_ptr[l]; stp < S.step_ptr[l + 1]; ++stp) {
            for (int phase = 0; phase < 2; ++phase) {
                int64_t a = phase == 0 ? S.pstep_off[stp] : S.pstep_mid[stp];
                int64_t e = phase == 0 ? S.pstep_mid[stp] : S.pstep_off[stp + 1];
                for (int64_t q = a; q < e; ++q) {
                    PanelTask t = S.ptasks[q];
                    int s = t.s;
                    int64_t ns = S.ncol[s], ms = S.m[s], k = t.k;
                    int b = (int)std::min<int64_t>(kPanel, ns - k);
                    double *Ls = L.data() + S.lofs[s];
                    if (phase == 0) {
                        if (t.r0 != k) return 2;
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
                    } else if (t.r0 == k) {
                        return 4;
                    }
                    int64_t r1 = std::min<int64_t>(t.r0 + kTile, ms);
                    for (int64_t i = std::max<int64_t>(t.r0, k + b); i < r1; ++i)
                        for (int j = 0; j < b; ++j) {
                            double v = Ls[i + ms * (k + j)];
                            for (int p = 0; p < j; ++p) v -= Ls[i + ms * (k + p)] * Ls[(k + j) + ms * (k + p)];
                            Ls[i + ms * (k + j)] = v / Ls[(k + j) + ms * (k + j)];
                        }
                }
            }
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
            }
        }
    }
    // solves, interleaved right-hand sides
    const int64_t n = S.n;
    std::vector<double> Y(n * nrhs), W(std::max<int64_t>(S.upd_rows, 1) * nrhs);
    for (int64_t i = 0; i < n; ++i)
        for (int r = 0; r < nrhs; ++r) Y[i * nrhs + r] = b[S.perm[i] + ndof * r];
    for (int64_t l = 0; l < nlev; ++l) {
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
            int s = S.level_sn[q];
            int64_t ns = S.ncol[s], ms = S.m[s];
            double *Ws = W.data() + S.relofs[s] * nrhs, *Ys = Y.data() + S.col0[s] * nrhs;
            const double *Ls = L.data() + S.lofs[s];
            for (int64_t i = 0; i < (ms - ns) * nrhs; ++i) Ws[i] = 0;
            for (int32_t qc = S.child_ptr[s]; qc < S.child_ptr[s + 1]; ++qc) {
                int c = S.child[qc];
                int64_t nu = S.m[c] - S.ncol[c];
                const int32_t *rel = S.relmap.data() + S.relofs[c];
                const double *Wc = W.data() + S.relofs[c] * nrhs;
                for (int64_t i = 0; i < nu; ++i)
                    for (int r = 0; r < nrhs; ++r) {
                        int64_t pr = rel[i];
                        if (pr < ns) Ys[pr * nrhs + r] += Wc[i * nrhs + r];
                        else Ws[(pr - ns) * nrhs + r] += Wc[i * nrhs + r];
                    }
            }
            for (int64_t j = 0; j < ns; ++j)
                for (int r = 0; r < nrhs; ++r) {
                    Ys[j * nrhs + r] /= Ls[j + ms * j];
                    for (int64_t i = j + 1; i < ns; ++i) Ys[i * nrhs + r] -= Ls[i + ms * j] * Ys[j * nrhs + r];
                }
        }
        for (int64_t q = S.fwd_off[l]; q < S.fwd_off[l + 1]; ++q) {
            PanelTask t = S.fwd_tiles[q];
            int s = t.s;
            int64_t ns = S.ncol[s], ms = S.m[s];
            const double *Ls = L.data() + S.lofs[s];
            double *Ws = W.data() + S.relofs[s] * nrhs;
            const double *Ys = Y.data() + S.col0[s] * nrhs;
            for (int64_t i = t.r0; i < std::min<int64_t>(t.r0 + kTile, ms); ++i)
                for (int r = 0; r < nrhs; ++r) {
                    double v = 0;
                    for (int64_t j = 0; j < ns; ++j) v += Ls[i + ms * j] * Ys[j * nrhs + r];
                    Ws[(i - ns) * nrhs + r] -= v;
                }
        }
    }
    for (int64_t l = nlev - 1; l >= 0; --l) {
        for (int64_t q = S.bwd_off[l]; q < S.bwd_off[l + 1]; ++q) {
            PanelTask t = S.bwd_tiles[q];
            int s = t.s;
            int64_t ns = S.ncol[s], ms = S.m[s];
            const double *Ls = L.data() + S.lofs[s];
            const int32_t *rw = S.rows.data() + S.rofs[s];
            double *Ys = Y.data() + S.col0[s] * nrhs;
            for (int64_t j = t.r0; j < std::min<int64_t>(t.r0 + kTile, ns); ++j)
                for (int r = 0; r < nrhs; ++r) {
                    double v = 0;
                    for (int64_t i = ns; i < ms; ++i) v += Ls[i + ms * j] * Y[(int64_t)rw[i] * nrhs + r];
                    Ys[j * nrhs + r] -= v;
                }
        }
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
            int s = S.level_sn[q];
            int64_t ns = S.ncol[s], ms = S.m[s];
            const double *Ls = L.data() + S.lofs[s];
            double *Ys = Y.data() + S.col0[s] * nrhs;
            for (int64_t j = ns - 1; j >= 0; --j)
                for (int r = 0; r < nrhs; ++r) {
                    double v = Ys[j * nrhs + r];
                    for (int64_t i = j + 1; i < ns; ++i) v -= Ls[i + ms * j] * Ys[i * nrhs + r];
                    Ys[j * nrhs + r] = v / Ls[j + ms * j];
                }
        }
    }
    for (int64_t i = 0; i < ndof * nrhs; ++i) x[i] = 0;
    for (int64_t i = 0; i < n; ++i)
        for (int r = 0; r < nrhs; ++r) x[S.perm[i] + ndof * r] = Y[i * nrhs + r];
    info[0] = S.ns;
    info[1] = nlev;
    info[2] = S.L_entries;
    info[3] = S.U_entries;
    info[4] = (int64_t)S.ptasks.size();
    info[5] = (int64_t)S.utasks.size();
    info[6] = S.max_front;
    return 0;
}
