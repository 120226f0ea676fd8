// Symbolic analysis of K_ff for the supernodal multifrontal Cholesky.
//
//  1. grid points with free DOFs are the supervariables (their DOFs share one
//     sparsity pattern); the grid-point graph comes from the CSR pattern of K;
//  2. fill-reducing ordering: geometric nested dissection on the grid coordinates
//     (median plane along the axis that gives the smallest one-sided vertex
//     separator);
//  3. elimination tree, grid-point column structures, fundamental supernodes and
//     relaxed amalgamation (merge a child into its parent when the merged front
//     adds few explicit zeros);
//  4. DOF-level fronts: column ranges, row lists, storage offsets, the K -> front
//     scatter map (upper triangle of K in original numbering, as
//     cholesky(Symmetric(K_ff)) reads it, src/solver/boundary_conditions.jl:1660),
//     child -> parent relative maps, level schedule and the task lists of the
//     batched panel/update/update-matrix/inversion/solve kernels.
#include <algorithm>
#include <cstdint>
#include <cstdint>
#include <cstring>
#include <functional>
#include <climits>
#include <numeric>
#include <string>
#include <vector>

#include "symbolic.h"

namespace jfem {

namespace {

struct Graph {
    std::vector<int64_t> ptr;
    std::vector<int32_t> adj;
};

// geometric nested dissection over the active grid points
struct ND {
    const double *xyz;
    const Graph &g;
    std::vector<int32_t> mark;   // 0 = outside, 1 = side A, 2 = side B
    std::vector<int32_t> out;
    int leaf;
    ND(const double *x, const Graph &gr, int64_t nn, int lf) : xyz(x), g(gr), mark(nn, 0), leaf(lf) {}

    // median split of `set` along axis ax: returns the split index h (0 = none)
    size_t split(std::vector<int32_t> &set, int ax) {
        std::stable_sort(set.begin(), set.end(), [&](int32_t a, int32_t b) { return xyz[3 * a + ax] < xyz[3 * b + ax]; });
        size_t h = set.size() / 2;
        // keep points with equal coordinate on one side
        while (h < set.size() && h > 0 && xyz[3 * set[h] + ax] == xyz[3 * set[h - 1] + ax]) ++h;
        if (h == set.size()) {
            h = set.size() / 2;
            while (h > 0 && xyz[3 * set[h] + ax] == xyz[3 * set[h - 1] + ax]) --h;
        }
        return (h == set.size()) ? 0 : h;
    }

    // one-sided separators of a split: SA = A-side points adjacent to B, SB likewise
    void separators(const std::vector<int32_t> &set, size_t h, std::vector<int32_t> &A, std::vector<int32_t> &B,
                    std::vector<int32_t> &SA, std::vector<int32_t> &SB) {
        A.clear(); B.clear(); SA.clear(); SB.clear();
        for (size_t k = 0; k < set.size(); ++k) mark[set[k]] = k < h ? 1 : 2;
        for (size_t k = 0; k < set.size(); ++k) {
            int32_t v = set[k];
            int other = k < h ? 2 : 1;
            bool sep = false;
            for (int64_t p = g.ptr[v]; p < g.ptr[v + 1]; ++p)
                if (mark[g.adj[p]] == other) { sep = true; break; }
            if (k < h) (sep ? SA : A).push_back(v);
            else (sep ? SB : B).push_back(v);
        }
        for (int32_t v : set) mark[v] = 0;
    }

    void run(std::vector<int32_t> set) {
        if ((int64_t)set.size() <= leaf) {
            out.insert(out.end(), set.begin(), set.end());
            return;
        }
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int32_t v : set)
            for (int a = 0; a < 3; ++a) {
                lo[a] = std::min(lo[a], xyz[3 * v + a]);
                hi[a] = std::max(hi[a], xyz[3 * v + a]);
            }
        // choose the axis whose median plane gives the smallest separator
        // (ties: larger extent); meshes with stretched cells otherwise cut along
        // the coarse direction and produce large separators
        int best_ax = -1;
        size_t best_sep = SIZE_MAX;
        double best_ext = -1;
        std::vector<int32_t> A, B, SA, SB;
        for (int ax = 0; ax < 3; ++ax) {
            if (!(hi[ax] - lo[ax] > 0)) continue;
            size_t h = split(set, ax);
            if (h == 0) continue;
            separators(set, h, A, B, SA, SB);
            size_t sz = std::min(SA.size(), SB.size());
            if (A.empty() && SA.size() <= SB.size()) continue;
            if (B.empty() && SB.size() < SA.size()) continue;
            double ext = hi[ax] - lo[ax];
            if (sz < best_sep || (sz == best_sep && ext > best_ext)) { best_sep = sz; best_ax = ax; best_ext = ext; }
        }
        if (best_ax < 0) {
            out.insert(out.end(), set.begin(), set.end());
            return;
        }
        size_t h = split(set, best_ax);
        separators(set, h, A, B, SA, SB);
        std::vector<int32_t> S;
        if (SA.size() <= SB.size()) {
            S = SA;
            B.insert(B.end(), SB.begin(), SB.end());
        } else {
            S = SB;
            A.insert(A.end(), SA.begin(), SA.end());
        }
        if (A.empty() || B.empty()) {
            out.insert(out.end(), set.begin(), set.end());
            return;
        }
        const int ax = best_ax;
        std::stable_sort(B.begin(), B.end(), [&](int32_t a, int32_t b) { return xyz[3 * a + ax] < xyz[3 * b + ax]; });
        run(std::move(A));
        run(std::move(B));
        out.insert(out.end(), S.begin(), S.end());
    }
};

}  // namespace

// Order the dataflow lists by estimated start time.  A workgroup that waits holds
// its slot on the GPU, so a task should be dispatched about when its inputs are
// ready.  List scheduling over the task graph with unlimited workers (a fixed
// latency per dependency, a duration growing with the tiles a task streams)
// gives each supernode's panel chain a start time.  The panel tasks of a supernode are
// keyed at the start of its chain (so the whole chain is resident when it runs),
// in chain order, and forward update tasks right after them; sorting by key
// (ties by the construction order, itself topological) keeps every task after
// the tasks it waits on (a backward chain is keyed after its parent's panels,
// whose keys lie below the parent's first finish).
// Move the tasks of the bottom subtrees to the front of the forward list and to
// the back of the backward list (stable, so both lists stay in dependency order:
// a bottom supernode's children are bottom supernodes, its ancestors come first
// backward).
static void split_flow_leaves(SymbolicHost &S) {
    const int64_t nsn = S.ns;
    std::vector<char> leaf(nsn, 0);
    for (int64_t s = 0; s < nsn; ++s) {          // children precede parents
        bool ok = S.ncol[s] <= kPanel && S.m[s] <= kWholeRows;
        for (int32_t q = S.child_ptr[s]; q < S.child_ptr[s + 1] && ok; ++q) ok = leaf[S.child[q]] != 0;
        leaf[s] = ok;
    }
    auto part = [&](std::vector<FlowTask> &v, bool leaves_first) {
        std::stable_partition(v.begin(), v.end(), [&](const FlowTask &t) { return (leaf[t.s] != 0) == leaves_first; });
        int64_t n = 0;
        for (const FlowTask &t : v) n += leaf[t.s] ? 1 : 0;
        return n;
    };
    S.flow_f_leaf = part(S.flow_f, true);
    S.flow_b_leaf = part(S.flow_b, false);
}

static void order_flow_lists(SymbolicHost &S, const std::vector<int32_t> &snode_of_pos) {
    const double lat = 1.0, tile = 1.5, base = 3.0, eps = 1e-6;
    const int64_t nsn = S.ns;
    auto stable_by = [](std::vector<FlowTask> &v, const std::vector<double> &key) {
        std::vector<size_t> idx(v.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return key[a] < key[b]; });
        std::vector<FlowTask> out(v.size());
        for (size_t i = 0; i < idx.size(); ++i) out[i] = v[idx[i]];
        v.swap(out);
    };
    // forward: children's outputs (whole-front or update tasks) -> the panel chain;
    // update tasks trail the chain of their supernode
    {
        std::vector<double> out_fin(nsn, 0.0), chain(nsn, 0.0), last_fin(nsn, 0.0), key(S.flow_f.size());
        for (size_t i = 0; i < S.flow_f.size(); ++i) {
            const FlowTask &t = S.flow_f[i];
            if (t.kind == 2) continue;                 // keyed with the task it precedes, below
            double ready = 0.0;
            for (int32_t q = S.child_ptr[t.s]; q < S.child_ptr[t.s + 1]; ++q) ready = std::max(ready, out_fin[S.child[q]]);
            if (t.nch) ready += lat;
            if (t.kind == 1) {
                key[i] = ready;
                out_fin[t.s] = ready + base + tile * (double)(t.ms - t.ns) / kTile;
            } else if (t.p >= 0) {
                // the panels of s run side by side: each publishes its v, then
                // forms z_p from the v of the panels before it (one exchange)
                if (t.p == 0) chain[t.s] = ready;
                key[i] = chain[t.s] + eps * t.p;
                last_fin[t.s] = ready + (t.np > 1 ? lat : 0.0) + base;
            } else {
                key[i] = chain[t.s] + eps * (t.np + 1);
                const double stream = tile * (t.k1 - t.k0 + (t.npart > 0 ? kSolveChunk : 0));
                out_fin[t.s] = std::max({out_fin[t.s], ready + base + stream, last_fin[t.s] + lat + tile});
            }
        }
        for (size_t i = S.flow_f.size(); i-- > 0;)
            if (S.flow_f[i].kind == 2) key[i] = key[i + 1];   // (a partial task never ends the list)
        stable_by(S.flow_f, key);
    }
    // backward: a supernode's panel tasks read its update rows' values, which its
    // ancestors publish from their last panel to their first; they are keyed at
    // the time the parent's first-finishing panel is done, and then trail the
    // parent's chain (their tiles go through the update rows from the last)
    {
        std::vector<double> first_fin(nsn, 0.0), key(S.flow_b.size());
        std::vector<double> chain_key(nsn, 0.0);
        std::vector<int32_t> xtask(nsn, 0);
        auto tpart = [](const FlowTask &t) { return t.kind == 2 && t.u1 > t.u0; };
        for (size_t i = 0; i < S.flow_b.size(); ++i) {
            const FlowTask &t = S.flow_b[i];
            if (tpart(t)) continue;                    // keyed with the t task it precedes, below
            if (t.kind == 2 || t.kind == 4) {          // after the panels' t tasks, in list order
                key[i] = chain_key[t.s] + eps * (t.np + xtask[t.s]++);
                continue;
            }
            if (t.p == t.np - 1) {
                chain_key[t.s] = t.parent >= 0 ? first_fin[t.parent] : 0.0;
                first_fin[t.s] = chain_key[t.s] + base + tile * (double)std::min<int64_t>(t.ms - t.ns, 2 * kTile * kSolveChunk) / kTile;
            }
            key[i] = chain_key[t.s] + eps * (t.np - 1 - t.p);
        }
        for (size_t i = S.flow_b.size(); i-- > 0;)
            if (tpart(S.flow_b[i])) key[i] = key[i + 1];   // (a partial task never ends the list)
        stable_by(S.flow_b, key);
    }
}

int build_symbolic(int64_t ndof, const int64_t *rowptr, const int32_t *col, const double *xyz, const uint8_t *fixed,
                   SymbolicHost &S, std::string &err) {
    const int64_t nn = ndof / 6;
    S = SymbolicHost();
    S.ndof = ndof;
    // ---- supervariables: grid points with free DOFs
    std::vector<int32_t> wdof(nn, 0);
    for (int64_t g = 0; g < ndof; ++g)
        if (!fixed[g]) wdof[g / 6]++;
    std::vector<int32_t> act;  // active grid points
    for (int64_t a = 0; a < nn; ++a)
        if (wdof[a]) act.push_back((int32_t)a);
    const int64_t na = (int64_t)act.size();
    S.n = 0;
    for (int64_t a = 0; a < nn; ++a) S.n += wdof[a];
    if (na == 0) {
        S.level_ptr.assign(1, 0);
        S.step_ptr.assign(1, 0);
        S.pstep_off.assign(1, 0);
        S.ustep_off.assign(1, 0);
        S.slev_ptr.assign(1, 0);
        S.spush_off.assign(1, 0);
        S.iofs.assign(1, 0);
        S.iperm.assign(ndof, -1);
        return 0;
    }
    std::vector<int32_t> aidx(nn, -1);
    for (int64_t i = 0; i < na; ++i) aidx[act[i]] = (int32_t)i;
    // ---- grid-point graph among active points (free-free couplings)
    Graph G;
    G.ptr.assign(na + 1, 0);
    {
        std::vector<int32_t> stamp(na, -1);
        std::vector<int32_t> tmp;
        for (int64_t i = 0; i < na; ++i) {
            int64_t a = act[i];
            for (int d = 0; d < 6; ++d) {
                int64_t gr = 6 * a + d;
                if (fixed[gr]) continue;
                for (int64_t p = rowptr[gr]; p < rowptr[gr + 1]; ++p) {
                    int32_t c = col[p];
                    if (fixed[c]) continue;
                    int32_t b = aidx[c / 6];
                    if (b == i || stamp[b] == i) continue;
                    stamp[b] = (int32_t)i;
                    G.adj.push_back(b);
                }
            }
            G.ptr[i + 1] = (int64_t)G.adj.size();
        }
    }
    // ---- nested dissection ordering
    std::vector<double> axyz(3 * na);
    for (int64_t i = 0; i < na; ++i)
        for (int c = 0; c < 3; ++c) axyz[3 * i + c] = xyz[3 * act[i] + c];
    std::vector<int32_t> order;
    {
        ND nd(axyz.data(), G, na, 24);
        std::vector<int32_t> all(na);
        std::iota(all.begin(), all.end(), 0);
        nd.run(std::move(all));
        order = std::move(nd.out);
    }
    if ((int64_t)order.size() != na) {
        err = "ordering lost grid points";
        return 1;
    }
    // ---- elimination tree (Liu) on the permuted grid-point graph
    std::vector<int32_t> pos(na);
    for (int64_t k = 0; k < na; ++k) pos[order[k]] = (int32_t)k;
    std::vector<int32_t> par(na, -1), anc(na, -1);
    for (int64_t k = 0; k < na; ++k) {
        int32_t v = order[k];
        for (int64_t p = G.ptr[v]; p < G.ptr[v + 1]; ++p) {
            int32_t i = pos[G.adj[p]];
            while (i != -1 && i < k) {
                int32_t nx = anc[i];
                anc[i] = (int32_t)k;
                if (nx == -1) { par[i] = (int32_t)k; break; }
                i = nx;
            }
        }
    }
    // postorder
    {
        std::vector<int32_t> head(na, -1), next(na, -1), post;
        post.reserve(na);
        for (int64_t j = na - 1; j >= 0; --j)
            if (par[j] != -1) { next[j] = head[par[j]]; head[par[j]] = (int32_t)j; }
        std::vector<int32_t> stack;
        for (int64_t r = 0; r < na; ++r) {
            if (par[r] != -1) continue;
            stack.push_back((int32_t)r);
            while (!stack.empty()) {
                int32_t p = stack.back();
                int32_t c = head[p];
                if (c == -1) { post.push_back(p); stack.pop_back(); }
                else { head[p] = next[c]; stack.push_back(c); }
            }
        }
        std::vector<int32_t> norder(na), npar(na);
        std::vector<int32_t> newpos(na);
        for (int64_t k = 0; k < na; ++k) newpos[post[k]] = (int32_t)k;
        for (int64_t k = 0; k < na; ++k) {
            norder[k] = order[post[k]];
            npar[k] = par[post[k]] == -1 ? -1 : newpos[par[post[k]]];
        }
        order.swap(norder);
        par.swap(npar);
        for (int64_t k = 0; k < na; ++k) pos[order[k]] = (int32_t)k;
    }
    // ---- grid-point column structures (strictly below), bottom-up
    std::vector<int32_t> nchild(na, 0);
    for (int64_t j = 0; j < na; ++j)
        if (par[j] != -1) nchild[par[j]]++;
    std::vector<std::vector<int32_t>> st(na);
    {
        std::vector<int32_t> stamp(na, -1);
        std::vector<std::vector<int32_t>> kids(na);
        for (int64_t j = 0; j < na; ++j)
            if (par[j] != -1) kids[par[j]].push_back((int32_t)j);
        for (int64_t j = 0; j < na; ++j) {
            std::vector<int32_t> s;
            int32_t v = order[j];
            for (int64_t p = G.ptr[v]; p < G.ptr[v + 1]; ++p) {
                int32_t i = pos[G.adj[p]];
                if (i > j && stamp[i] != j) { stamp[i] = (int32_t)j; s.push_back(i); }
            }
            for (int32_t c : kids[j]) {
                for (int32_t i : st[c])
                    if (i > j && stamp[i] != j) { stamp[i] = (int32_t)j; s.push_back(i); }
                // a child's structure is no longer needed unless it heads a supernode
            }
            std::sort(s.begin(), s.end());
            st[j] = std::move(s);
        }
    }
    // ---- fundamental supernodes (grid-point level)
    std::vector<int32_t> snode_of(na);
    std::vector<int32_t> sfirst, slast;
    for (int64_t j = 0; j < na; ++j) {
        bool join = j > 0 && par[j - 1] == j && nchild[j] == 1 && st[j - 1].size() == st[j].size() + 1;
        if (!join) { sfirst.push_back((int32_t)j); slast.push_back((int32_t)j); }
        else slast.back() = (int32_t)j;
        snode_of[j] = (int32_t)sfirst.size() - 1;
    }
    const int64_t nf = (int64_t)sfirst.size();
    std::vector<int32_t> fpar(nf, -1);
    for (int64_t s = 0; s < nf; ++s) {
        int32_t p = par[slast[s]];
        fpar[s] = p == -1 ? -1 : snode_of[p];
    }
    // DOF weights
    std::vector<int64_t> fcol(nf, 0), frow(nf, 0);
    for (int64_t s = 0; s < nf; ++s) {
        for (int32_t j = sfirst[s]; j <= slast[s]; ++j) fcol[s] += wdof[act[order[j]]];
        for (int32_t i : st[slast[s]]) frow[s] += wdof[act[order[i]]];
    }
    // ---- relaxed amalgamation
    std::vector<std::vector<int32_t>> members(nf);   // fundamental supernodes merged into s (in column order)
    std::vector<double> actual(nf);
    std::vector<int64_t> mcol(fcol);
    std::vector<int32_t> merged_into(nf, -1);
    std::vector<std::vector<int32_t>> fkids(nf);
    for (int64_t s = 0; s < nf; ++s) {
        members[s].push_back((int32_t)s);
        double c = (double)fcol[s];
        actual[s] = c * (c + 1) / 2 + c * (double)frow[s];
        if (fpar[s] != -1) fkids[fpar[s]].push_back((int32_t)s);
    }
    auto dense = [](double c, double r) { return c * (c + 1) / 2 + c * r; };
    for (int64_t s = 0; s < nf; ++s) {
        // children in reverse column order: the last child is adjacent to s
        std::vector<int32_t> kids = fkids[s];
        std::vector<int32_t> keep;
        for (auto it = kids.rbegin(); it != kids.rend(); ++it) {
            int32_t c = *it;
            double nc = (double)(mcol[c] + mcol[s]);
            double dz = dense(nc, (double)frow[s]);
            double z = 1.0 - (actual[c] + actual[s]) / dz;
            // merged width nc in DOF columns, z the share of explicit zeros in the
            // merged front.  Looser than the usual CPU settings (4/16/48 columns):
            // on the GPU every supernode level costs launches in the factor and a
            // dependency step in each triangular solve, while the extra
            // arithmetic on small fronts is nearly free.
            bool ok = nc <= 12 || (nc <= 24 && z < 0.8) || (nc <= 72 && z < 0.2) || (nc <= 256 && z < 0.05);
            if (ok) {
                merged_into[c] = (int32_t)s;
                std::vector<int32_t> mem = members[c];
                mem.insert(mem.end(), members[s].begin(), members[s].end());
                members[s].swap(mem);
                actual[s] += actual[c];
                mcol[s] += mcol[c];
            }
        }
    }
    // final supernodes = fundamental supernodes not merged; tree over them
    auto top = [&](int32_t s) {
        while (merged_into[s] != -1) s = merged_into[s];
        return s;
    };
    std::vector<int32_t> fin;                 // final supernode -> representative fundamental id
    std::vector<int32_t> fin_of(nf, -1);
    for (int64_t s = 0; s < nf; ++s)
        if (merged_into[s] == -1) { fin_of[s] = (int32_t)fin.size(); fin.push_back((int32_t)s); }
    const int64_t nsn = (int64_t)fin.size();
    std::vector<int32_t> spar(nsn, -1);
    for (int64_t k = 0; k < nsn; ++k) {
        int32_t p = fpar[fin[k]];
        spar[k] = p == -1 ? -1 : fin_of[top(p)];
    }
    // children of final supernodes (every fundamental child of any member that is not itself a member)
    std::vector<std::vector<int32_t>> skids(nsn);
    for (int64_t k = 0; k < nsn; ++k)
        if (spar[k] != -1) skids[spar[k]].push_back((int32_t)k);
    // postorder of the final tree; emit grid points of each supernode (members in order)
    std::vector<int32_t> spost;
    spost.reserve(nsn);
    {
        std::vector<int32_t> stack, it(nsn, 0);
        for (int64_t r = 0; r < nsn; ++r) {
            if (spar[r] != -1) continue;
            stack.push_back((int32_t)r);
            while (!stack.empty()) {
                int32_t p = stack.back();
                if (it[p] < (int32_t)skids[p].size()) {
                    stack.push_back(skids[p][it[p]++]);
                } else {
                    spost.push_back(p);
                    stack.pop_back();
                }
            }
        }
    }
    std::vector<int32_t> snew(nsn);
    for (int64_t k = 0; k < nsn; ++k) snew[spost[k]] = (int32_t)k;
    std::vector<int32_t> final_order;         // new grid-point order (old position indices j)
    final_order.reserve(na);
    std::vector<int32_t> sn_gfirst(nsn + 1);
    for (int64_t k = 0; k < nsn; ++k) {
        int32_t rep = fin[spost[k]];
        sn_gfirst[k] = (int32_t)final_order.size();
        for (int32_t f : members[rep])
            for (int32_t j = sfirst[f]; j <= slast[f]; ++j) final_order.push_back(j);
    }
    sn_gfirst[nsn] = (int32_t)final_order.size();
    std::vector<int32_t> gpos(na);            // old position j -> new position
    for (int64_t k = 0; k < na; ++k) gpos[final_order[k]] = (int32_t)k;
    // DOF positions
    std::vector<int64_t> dofpos(na + 1, 0);   // by new grid position
    for (int64_t k = 0; k < na; ++k) dofpos[k + 1] = dofpos[k] + wdof[act[order[final_order[k]]]];
    S.perm.assign(S.n, 0);
    S.iperm.assign(ndof, -1);
    for (int64_t k = 0; k < na; ++k) {
        int32_t a = act[order[final_order[k]]];
        int64_t p = dofpos[k];
        for (int d = 0; d < 6; ++d)
            if (!fixed[6 * a + d]) { S.perm[p] = (int32_t)(6 * a + d); S.iperm[6 * a + d] = (int32_t)p; ++p; }
    }
    // ---- DOF-level fronts
    S.ns = nsn;
    S.col0.resize(nsn); S.ncol.resize(nsn); S.m.resize(nsn); S.lofs.resize(nsn + 1); S.uofs.resize(nsn + 1);
    S.rofs.resize(nsn + 1); S.parent.resize(nsn);
    int64_t lo = 0, uo = 0, ro = 0;
    std::vector<int32_t> snode_of_pos(S.n);
    for (int64_t k = 0; k < nsn; ++k) {
        int32_t rep = fin[spost[k]];
        int32_t topf = members[rep].back();       // highest fundamental supernode of the merged set
        S.col0[k] = dofpos[sn_gfirst[k]];
        S.ncol[k] = dofpos[sn_gfirst[k + 1]] - dofpos[sn_gfirst[k]];
        std::vector<int32_t> sg;                  // structure grid points in new positions
        for (int32_t i : st[slast[topf]]) sg.push_back(gpos[i]);
        std::sort(sg.begin(), sg.end());
        S.rofs[k] = ro;
        for (int64_t c = 0; c < S.ncol[k]; ++c) S.rows.push_back((int32_t)(S.col0[k] + c));
        for (int32_t gp : sg)
            for (int64_t p = dofpos[gp]; p < dofpos[gp + 1]; ++p) S.rows.push_back((int32_t)p);
        S.m[k] = (int64_t)S.rows.size() - ro;
        ro = (int64_t)S.rows.size();
        S.lofs[k] = lo;
        lo += S.m[k] * S.ncol[k];
        S.uofs[k] = uo;
        uo += (S.m[k] - S.ncol[k]) * (S.m[k] - S.ncol[k]);
        S.parent[k] = spar[spost[k]] == -1 ? -1 : snew[spar[spost[k]]];
        for (int64_t c = 0; c < S.ncol[k]; ++c) snode_of_pos[S.col0[k] + c] = (int32_t)k;
        S.max_front = std::max<int64_t>(S.max_front, S.m[k]);
        for (int64_t j = 0; j < S.ncol[k]; ++j) S.flops += (double)(S.m[k] - j) * (double)(S.m[k] - j);
        S.upd_rows += S.m[k] - S.ncol[k];
    }
    S.lofs[nsn] = lo; S.uofs[nsn] = uo; S.rofs[nsn] = ro;
    S.L_entries = lo; S.U_entries = uo;
    // children lists and relative maps
    S.child_ptr.assign(nsn + 1, 0);
    for (int64_t k = 0; k < nsn; ++k)
        if (S.parent[k] != -1) S.child_ptr[S.parent[k] + 1]++;
    for (int64_t k = 0; k < nsn; ++k) S.child_ptr[k + 1] += S.child_ptr[k];
    S.child.resize(S.child_ptr[nsn]);
    {
        std::vector<int32_t> fill(S.child_ptr.begin(), S.child_ptr.end() - 1);
        for (int64_t k = 0; k < nsn; ++k)
            if (S.parent[k] != -1) S.child[fill[S.parent[k]]++] = (int32_t)k;
    }
    S.relofs.resize(nsn + 1);
    int64_t rel = 0;
    for (int64_t k = 0; k < nsn; ++k) {
        S.relofs[k] = rel;
        int32_t p = S.parent[k];
        int64_t nu = S.m[k] - S.ncol[k];
        if (p == -1) {
            if (nu != 0) { err = "root supernode with a non-empty update"; return 1; }
            continue;
        }
        const int32_t *pr = S.rows.data() + S.rofs[p];
        for (int64_t i = 0; i < nu; ++i) {
            int32_t r = S.rows[S.rofs[k] + S.ncol[k] + i];
            const int32_t *it = std::lower_bound(pr, pr + S.m[p], r);
            if (it == pr + S.m[p] || *it != r) { err = "child row missing from parent front"; return 1; }
            S.relmap.push_back((int32_t)(it - pr));
        }
        rel += nu;
    }
    S.relofs[nsn] = rel;
    // K -> front scatter map: entries (g, c) with c >= g (upper triangle, original
    // numbering), in K order.  The entries are grouped by the supernode owning
    // their column (a stable counting sort); per supernode a position map
    // row -> local row (stamped with the supernode, never cleared) gives each
    // entry's front row in O(1).
    {
        // pass 1: kept entries per supernode (and in total)
        std::vector<int64_t> sptr(nsn + 1, 0);
        int64_t ne = 0;
        for (int64_t g = 0; g < ndof; ++g) {
            if (fixed[g]) continue;
            const int32_t a = S.iperm[g];
            for (int64_t p = rowptr[g]; p < rowptr[g + 1]; ++p) {
                const int64_t c = col[p];
                if (c < g || fixed[c]) continue;
                sptr[snode_of_pos[std::min(a, S.iperm[c])] + 1]++;
                ++ne;
            }
        }
        for (int64_t q = 0; q < nsn; ++q) sptr[q + 1] += sptr[q];
        // pass 2: output slot k (K order) of each entry, grouped by supernode
        S.amap_src.resize(ne);
        S.amap_dst.resize(ne);
        std::vector<int64_t> order(ne);
        std::vector<int32_t> e_lo(ne), e_hi(ne);
        {
            std::vector<int64_t> fill(sptr.begin(), sptr.end() - 1);
            int64_t k = 0;
            for (int64_t g = 0; g < ndof; ++g) {
                if (fixed[g]) continue;
                const int32_t a = S.iperm[g];
                for (int64_t p = rowptr[g]; p < rowptr[g + 1]; ++p) {
                    const int64_t c = col[p];
                    if (c < g || fixed[c]) continue;
                    const int32_t b = S.iperm[c], lo = std::min(a, b);
                    S.amap_src[k] = (int32_t)p;
                    e_lo[k] = lo;
                    e_hi[k] = std::max(a, b);
                    order[fill[snode_of_pos[lo]]++] = k;
                    ++k;
                }
            }
        }
        std::vector<int32_t> stamp(S.n, -1), local(S.n, 0);
        for (int64_t q = 0; q < nsn; ++q) {
            const int32_t *rr = S.rows.data() + S.rofs[q];
            for (int64_t i = 0; i < S.m[q]; ++i) { stamp[rr[i]] = (int32_t)q; local[rr[i]] = (int32_t)i; }
            for (int64_t t = sptr[q]; t < sptr[q + 1]; ++t) {
                const int64_t k = order[t];
                const int32_t hi = e_hi[k];
                if (stamp[hi] != q) { err = "matrix entry outside the symbolic structure"; return 1; }
                S.amap_dst[k] = S.lofs[q] + local[hi] + S.m[q] * (int64_t)(e_lo[k] - S.col0[q]);
            }
        }
    }
    // levels (height in the supernode tree)
    std::vector<int32_t> lev(nsn, 0);
    int32_t nlev = 0;
    for (int64_t k = 0; k < nsn; ++k) {
        int32_t l = 0;
        for (int32_t q = S.child_ptr[k]; q < S.child_ptr[k + 1]; ++q) l = std::max(l, lev[S.child[q]] + 1);
        lev[k] = l;
        nlev = std::max(nlev, l + 1);
    }
    S.level_ptr.assign(nlev + 1, 0);
    for (int64_t k = 0; k < nsn; ++k) S.level_ptr[lev[k] + 1]++;
    for (int32_t l = 0; l < nlev; ++l) S.level_ptr[l + 1] += S.level_ptr[l];
    S.level_sn.resize(nsn);
    {
        std::vector<int32_t> fill(S.level_ptr.begin(), S.level_ptr.end() - 1);
        for (int64_t k = 0; k < nsn; ++k) S.level_sn[fill[lev[k]]++] = (int32_t)k;
    }
    // task lists
    S.step_ptr.assign(nlev + 1, 0);
    S.pstep_off.clear(); S.ustep_off.clear(); S.pstep_mid.clear(); S.pstep_bw.clear();
    for (int32_t l = 0; l < nlev; ++l) {
        int64_t maxsteps = 0;
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
            int32_t s = S.level_sn[q];
            maxsteps = std::max<int64_t>(maxsteps, (S.ncol[s] + kPanel - 1) / kPanel);
        }
        S.step_ptr[l] = (int64_t)S.pstep_off.size();
        for (int64_t t = 0; t < maxsteps; ++t) {
            S.pstep_off.push_back((int64_t)S.ptasks.size());
            S.ustep_off.push_back((int64_t)S.utasks.size());
            // diagonal tiles first (they factor and store the diagonal block), then the rest
            int32_t bw = 1;
            for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
                int32_t s = S.level_sn[q];
                int64_t k = t * kPanel;
                if (k < S.ncol[s]) {
                    S.ptasks.push_back({s, (int32_t)k, (int32_t)k, 0});
                    bw = std::max<int32_t>(bw, (int32_t)std::min<int64_t>(kPanel, S.ncol[s] - k));
                }
            }
            S.pstep_bw.push_back(bw);
            S.pstep_mid.push_back((int64_t)S.ptasks.size());
            for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
                int32_t s = S.level_sn[q];
                int64_t k = t * kPanel;
                if (k >= S.ncol[s]) continue;
                int64_t b = std::min<int64_t>(kPanel, S.ncol[s] - k);
                for (int64_t r0 = k + kTile; r0 < S.m[s]; r0 += kTile) S.ptasks.push_back({s, (int32_t)k, (int32_t)r0, 0});
                // tile origins of the trailing columns of the front (pivot columns, then
                // update rows); per step only the tiles of pivot columns: the update
                // matrix takes its whole rank-ncol product once, after the last step
                std::vector<int32_t> bnd;
                for (int64_t c = k + b; c < S.ncol[s]; c += kTile) bnd.push_back((int32_t)c);
                const size_t npiv = bnd.size();
                for (int64_t c = std::max<int64_t>(k + b, S.ncol[s]); c < S.m[s]; c += kTile) bnd.push_back((int32_t)c);
                for (size_t ci = 0; ci < npiv; ++ci)
                    for (size_t ri = ci; ri < bnd.size(); ++ri) {
                        S.utasks.push_back({s, (int32_t)k, bnd[ri], bnd[ci]});
                        int64_t r0 = bnd[ri], c0 = bnd[ci];
                        int64_t r1 = std::min<int64_t>(r0 + kTile, r0 < S.ncol[s] ? S.ncol[s] : S.m[s]);
                        int64_t c1 = std::min<int64_t>(c0 + kTile, c0 < S.ncol[s] ? S.ncol[s] : S.m[s]);
                        double cnt = 0;
                        for (int64_t c = c0; c < c1; ++c) cnt += (double)std::max<int64_t>(0, r1 - std::max(r0, c));
                        S.update_flops += 2.0 * (double)b * cnt;
                    }
            }
        }
    }
    // per level: the update-matrix tiles of its fronts (lower triangles), taken
    // after the level's last panel step with the full rank ncol of each front and
    // added into the parents' fronts; one launch per sibling rank
    S.stasks.clear();
    S.spush_off.assign(1, 0);
    S.slev_ptr.assign(nlev + 1, 0);
    S.schur_flops = 0;
    {
        std::vector<int32_t> rank(nsn, 0), seen(nsn, -1), cnt_at(nsn, 0);
        for (int32_t l = 0; l < nlev; ++l) {
            S.slev_ptr[l] = (int64_t)S.spush_off.size() - 1;
            int32_t maxrank = -1;
            for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
                const int32_t s = S.level_sn[q], p = S.parent[s];
                if (p < 0 || S.m[s] <= S.ncol[s]) continue;
                if (seen[p] != l) { seen[p] = l; cnt_at[p] = 0; }
                rank[s] = cnt_at[p]++;
                maxrank = std::max(maxrank, rank[s]);
            }
            for (int32_t r = 0; r <= maxrank; ++r) {
                for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
                    const int32_t s = S.level_sn[q], p = S.parent[s];
                    if (p < 0 || S.m[s] <= S.ncol[s] || rank[s] != r) continue;
                    const int64_t ns = S.ncol[s], ms = S.m[s];
                    for (int64_t c0 = ns; c0 < ms; c0 += kTile)
                        for (int64_t r0 = c0; r0 < ms; r0 += kTile) {
                            PushTask t{};
                            t.s = s; t.r0 = (int32_t)r0; t.c0 = (int32_t)c0;
                            t.ns = (int32_t)ns; t.ms = (int32_t)ms;
                            t.nsp = (int32_t)S.ncol[p]; t.msp = (int32_t)S.m[p];
                            t.lofs = S.lofs[s]; t.uofs = S.uofs[s]; t.relofs = S.relofs[s];
                            t.lofsp = S.lofs[p]; t.uofsp = S.uofs[p];
                            S.stasks.push_back(t);
                            const int64_t r1 = std::min<int64_t>(r0 + kTile, ms), c1 = std::min<int64_t>(c0 + kTile, ms);
                            double cnt = 0;
                            for (int64_t c = c0; c < c1; ++c) cnt += (double)std::max<int64_t>(0, r1 - std::max(r0, c));
                            S.schur_flops += 2.0 * (double)ns * cnt;
                        }
                }
                S.spush_off.push_back((int64_t)S.stasks.size());
            }
        }
        S.slev_ptr[nlev] = (int64_t)S.spush_off.size() - 1;
    }
    S.step_ptr[nlev] = (int64_t)S.pstep_off.size();
    S.pstep_off.push_back((int64_t)S.ptasks.size());
    S.ustep_off.push_back((int64_t)S.utasks.size());
    // dataflow solve lists
    auto npan = [&](int64_t s) { return (int32_t)((S.ncol[s] + kPanel - 1) / kPanel); };
    auto nupd = [&](int64_t s) { return (int32_t)((S.m[s] - S.ncol[s] + kTile - 1) / kTile); };
    S.iofs.assign(nsn + 1, 0);
    for (int64_t k = 0; k < nsn; ++k) {
        int64_t e = 0;
        for (int64_t c = 0; c < S.ncol[k]; c += kPanel) {
            const int64_t b = std::min<int64_t>(kPanel, S.ncol[k] - c);
            e += b * b;
        }
        S.iofs[k + 1] = S.iofs[k] + e;
    }
    // X = L_11^{-1} of every supernode with two or more panels, after the
    // diagonal-block inverses (ns x ns each, 64-entry aligned)
    std::vector<int64_t> xofs(nsn, -1);
    int64_t nx = (S.iofs[nsn] + 63) & ~(int64_t)63;
    int32_t max_np = 0;
    for (int64_t k = 0; k < nsn; ++k) {
        const int32_t np = npan(k);
        max_np = std::max(max_np, np);
        if (np < 2) continue;
        xofs[k] = nx;
        nx += (S.ncol[k] * S.ncol[k] + 63) & ~(int64_t)63;
    }
    S.inv_entries = nx;
    // inversion tasks by doubling (InvTask); within a launch the longest k ranges first
    S.itasks.clear();
    S.istep_off.assign(1, 0);
    S.inv_flops = 0;
    for (int32_t B = 1; B < max_np; B *= 2)
        for (int ph = 0; ph < 2; ++ph) {
            const size_t first = S.itasks.size();
            for (int64_t s = 0; s < nsn; ++s) {
                const int32_t np = npan(s);
                if (np <= B) continue;
                for (int32_t a = 0; a + B < np; a += 2 * B) {
                    const int32_t c1 = std::min(a + 2 * B, np);
                    for (int32_t i = a + B; i < c1; ++i)
                        for (int32_t j = a; j < a + B; ++j) {
                            InvTask t;
                            t.lofs = S.lofs[s]; t.xofs = xofs[s];
                            t.ns = (int32_t)S.ncol[s]; t.ms = (int32_t)S.m[s];
                            t.i = i; t.j = j;
                            t.k0 = ph == 0 ? j : a + B;
                            t.k1 = ph == 0 ? a + B : i + 1;
                            S.itasks.push_back(t);
                            S.inv_flops += 2.0 * kPanel * kPanel * kPanel * (t.k1 - t.k0);
                        }
                }
            }
            std::stable_sort(S.itasks.begin() + first, S.itasks.end(),
                             [](const InvTask &x, const InvTask &y) { return x.k1 - x.k0 > y.k1 - y.k0; });
            S.istep_off.push_back((int64_t)S.itasks.size());
        }
    auto whole = [&](int64_t s) { return S.ncol[s] <= kPanel && S.m[s] <= kWholeRows; };
    std::vector<int32_t> need(nsn, 0);
    for (int64_t k = 0; k < nsn; ++k)
        for (int32_t q = S.child_ptr[k]; q < S.child_ptr[k + 1]; ++q) {
            const int32_t c = S.child[q];
            need[k] += whole(c) ? 1 : nupd(c);
        }
    auto base = [&](int32_t s) {
        FlowTask t{};
        t.s = s;
        t.np = npan(s);
        t.need = need[s];
        t.parent = S.parent[s];
        t.pnp = t.parent >= 0 ? npan(t.parent) : 0;
        t.nch = S.child_ptr[s + 1] - S.child_ptr[s];
        t.ns = S.ncol[s]; t.ms = S.m[s]; t.lofs = S.lofs[s]; t.c0 = S.col0[s];
        t.relofs = S.relofs[s]; t.rofs = S.rofs[s];
        return t;
    };
    S.flow_f.clear(); S.flow_b.clear(); S.flow_gb.clear();
    auto add_fwd = [&](FlowTask t, int64_t rows_end) {
        t.gb = (int64_t)S.flow_gb.size();
        for (int32_t q = S.child_ptr[t.s]; q < S.child_ptr[t.s + 1]; ++q) {
            const int32_t c = S.child[q];
            const int32_t *rel = S.relmap.data() + S.relofs[c];
            const int64_t nu = S.m[c] - S.ncol[c];
            FlowChild f;
            f.relofs = S.relofs[c];
            f.i0 = (int32_t)(std::lower_bound(rel, rel + nu, (int64_t)t.r0) - rel);
            f.i1 = (int32_t)(std::lower_bound(rel, rel + nu, rows_end) - rel);
            S.flow_gb.push_back(f);
        }
        S.flow_f.push_back(t);
    };
    // a forward task summing blocks k in [0, kend): its own sum over the last
    // kSolveChunk, the partial tasks over the rest listed just before it
    S.flow_f_parts = S.flow_b_parts = 0;
    auto add_chunked = [&](FlowTask t, int32_t kend, int64_t rows_end) {
        const int32_t own0 = solve_own_first(kend, kSolveSplit);
        t.npart = 0;
        t.pofs = (int32_t)S.flow_f_parts;
        for (int32_t k = 0; k < own0; k += kSolveChunk) {
            FlowTask q = t;
            q.kind = 2;
            q.k0 = k;
            q.k1 = std::min(k + kSolveChunk, own0);
            q.need = 0;
            q.nch = 0;
            q.gb = 0;
            q.npart = 0;
            q.pofs = (int32_t)S.flow_f_parts++;
            S.flow_f.push_back(q);
            ++t.npart;
        }
        t.k0 = own0;
        t.k1 = kend;
        add_fwd(t, rows_end);
    };
    for (int32_t l = 0; l < nlev; ++l)
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
            const int32_t s = S.level_sn[q];
            FlowTask t = base(s);
            t.xofs = -1;
            if (whole(s)) {
                t.kind = 1;
                t.p = 0; t.r0 = 0; t.nr = (int32_t)S.ncol[s];
                t.iofs = S.iofs[s];
                t.k0 = t.k1 = t.npart = t.pofs = 0;
                add_fwd(t, S.m[s]);
                continue;
            }
            for (int32_t p = 0; p < t.np; ++p) {
                t.p = p;
                t.r0 = p * kPanel;
                t.nr = (int32_t)std::min<int64_t>(kPanel, S.ncol[s] - t.r0);
                t.iofs = S.iofs[s] + (int64_t)kPanel * kPanel * p;
                t.xofs = xofs[s];
                add_chunked(t, p, (int64_t)t.r0 + t.nr);
            }
            t.iofs = 0;
            t.xofs = -1;
            for (int32_t u = 0; u < nupd(s); ++u) {
                t.p = -1 - u;
                t.r0 = (int32_t)(S.ncol[s] + (int64_t)u * kTile);
                t.nr = (int32_t)std::min<int64_t>(kTile, S.m[s] - t.r0);
                add_chunked(t, t.np, (int64_t)t.r0 + t.nr);
            }
        }
    for (int32_t l = nlev - 1; l >= 0; --l)
        for (int32_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
            const int32_t s = S.level_sn[q];
            FlowTask t = base(s);
            t.u0 = (int32_t)S.ncol[s];
            t.u1 = (int32_t)S.m[s];
            auto panel = [&](int32_t p) {
                t.p = p;
                t.r0 = p * kPanel;
                t.nr = (int32_t)std::min<int64_t>(kPanel, S.ncol[s] - t.r0);
                t.iofs = S.iofs[s] + (int64_t)kPanel * kPanel * p;
                t.xofs = xofs[s];
                t.npart = 0;
                t.pofs = 0;
            };
            // per panel the partial sums of its far update rows and its t (and x,
            // up to 2 kSolveChunk later panels); then the split panels' partial
            // sums over X and x
            const int32_t nut = (int32_t)((S.m[s] - S.ncol[s] + kTile - 1) / kTile);   // update-row tiles
            const int32_t urow1 = (int32_t)(S.ncol[s] + (int64_t)kTile * kSolveChunk);
            const bool usplit = !whole(s) && solve_own_first(nut, kSolveSplitRows) > 0;   // (bottom-subtree tasks never split)
            for (int32_t p = t.np - 1; p >= 0; --p) {
                panel(p);
                const int32_t first = (int32_t)S.flow_b_parts;
                if (usplit)
                    for (int64_t r0 = urow1; r0 < S.m[s]; r0 += (int64_t)kTile * kSolveChunk) {
                        t.kind = 2;
                        t.k0 = t.k1 = 0;
                        t.u0 = (int32_t)r0;
                        t.u1 = (int32_t)std::min<int64_t>(r0 + (int64_t)kTile * kSolveChunk, S.m[s]);
                        t.pofs = (int32_t)S.flow_b_parts++;
                        S.flow_b.push_back(t);
                    }
                const bool split = solve_own_first(t.np - 1 - p, kSolveSplit) > 0;
                t.kind = split ? 3 : 0;
                t.k0 = split ? 0 : p + 1;
                t.k1 = split ? 0 : t.np;
                t.u0 = (int32_t)S.ncol[s];
                t.u1 = usplit ? urow1 : (int32_t)S.m[s];
                t.npart = (int32_t)S.flow_b_parts - first;
                t.pofs = first;
                S.flow_b.push_back(t);
            }
            t.u0 = t.u1 = (int32_t)S.ncol[s];                 // (the X tasks read no update rows)
            for (int32_t p = t.np - 1; p >= 0; --p) {
                if (solve_own_first(t.np - 1 - p, kSolveSplit) == 0) continue;
                panel(p);
                const int32_t own1 = p + 1 + kSolveChunk;
                const int32_t first = (int32_t)S.flow_b_parts;
                for (int32_t q0 = own1; q0 < t.np; q0 += kSolveChunk) {
                    t.kind = 2;
                    t.k0 = q0;
                    t.k1 = std::min(q0 + kSolveChunk, t.np);
                    t.pofs = (int32_t)S.flow_b_parts++;
                    S.flow_b.push_back(t);
                }
                t.kind = 4;
                t.k0 = p + 1;
                t.k1 = own1;
                t.npart = (int32_t)S.flow_b_parts - first;
                t.pofs = first;
                S.flow_b.push_back(t);
            }
        }
    order_flow_lists(S, snode_of_pos);
    split_flow_leaves(S);
    return 0;
}

}  // namespace jfem
