// Host-side symbolic work of libpf, done once per network:
//  * the Ybus CSR pattern and the per-slot contribution lists that make the
//    GPU assembly a deterministic segmented sum (no float atomics);
//  * ordering + symbolic factorisation of the 2x2-block Jacobian on the bus
//    graph (minus the slack), elimination-tree levels, and the "factor
//    program" (for every node, every destination block and the list of
//    source-block pairs that update it);
//  * the host numeric factor/solve over that program (metodo hibridoC).
//
// Why a fixed 2x2-block pattern: the reference rebuilds the Jacobian pattern
// and redoes the symbolic analysis whenever a PV bus switches to PQ
// (CPU/esparso/spCalcJac.cpp:402-568; PARDISO phase 11 each iteration,
// CPU/esparso/spSolver.cpp:141-153).  Here every non-slack bus keeps two
// unknowns (dtheta, dV) and two equations (dP, dQ); for a PV bus the dQ row
// is replaced by dV = 0 and its dV column is zeroed, which leaves the
// reference's reduced system embedded unchanged, so one symbolic analysis
// serves every switching state.

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <queue>

#include "pf_internal.hpp"

namespace pf {

YPattern build_ypattern(const pf_network &net) {
    struct T {
        int row, col, code;
    };
    std::vector<T> t;
    t.reserve(net.nB + 4 * (size_t)net.nL);
    for (int l = 0; l < net.nL; l++) {
        int f = net.from[l], to = net.to[l];
        t.push_back({f, f, 4 * l + 0});
        t.push_back({f, to, 4 * l + 1});
        t.push_back({to, f, 4 * l + 2});
        t.push_back({to, to, 4 * l + 3});
    }
    for (int i = 0; i < net.nB; i++) t.push_back({i, i, INT32_MAX});   // shunt, last in its slot
    std::sort(t.begin(), t.end(), [](const T &a, const T &b) {
        if (a.row != b.row) return a.row < b.row;
        if (a.col != b.col) return a.col < b.col;
        return a.code < b.code;
    });
    YPattern y;
    y.nB = net.nB;
    y.rowptr.assign(net.nB + 1, 0);
    y.diag.assign(net.nB, -1);
    y.br_ff.assign(net.nL, -1);
    y.br_ft.assign(net.nL, -1);
    y.br_tf.assign(net.nL, -1);
    y.br_tt.assign(net.nL, -1);
    y.contrib_ptr.push_back(0);
    for (size_t i = 0; i < t.size();) {
        size_t j = i;
        int slot = (int)y.colind.size();
        while (j < t.size() && t[j].row == t[i].row && t[j].col == t[i].col) {
            if (t[j].code != INT32_MAX) {
                y.contrib.push_back(t[j].code);
                int l = t[j].code >> 2;
                switch (t[j].code & 3) {
                    case 0: y.br_ff[l] = slot; break;
                    case 1: y.br_ft[l] = slot; break;
                    case 2: y.br_tf[l] = slot; break;
                    default: y.br_tt[l] = slot; break;
                }
            }
            j++;
        }
        y.colind.push_back(t[i].col);
        y.coorow.push_back(t[i].row);
        if (t[i].row == t[i].col) y.diag[t[i].row] = slot;
        y.rowptr[t[i].row + 1] = (int)y.colind.size();
        y.contrib_ptr.push_back((int)y.contrib.size());
        i = j;
    }
    for (int r = 0; r < net.nB; r++) y.rowptr[r + 1] = std::max(y.rowptr[r + 1], y.rowptr[r]);
    y.nnz = (int)y.colind.size();
    return y;
}

// Minimum degree on the explicit elimination graph; ties broken by the
// lowest index, so the ordering is a pure function of the pattern.
std::vector<int> min_degree(int n, const std::vector<int> &rowptr, const std::vector<int> &colind) {
    std::vector<std::vector<int>> adj(n);
    for (int i = 0; i < n; i++)
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
            int j = colind[p];
            if (j == i) continue;
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    using E = std::pair<int, int>;   // (degree, node)
    std::priority_queue<E, std::vector<E>, std::greater<E>> pq;
    for (int i = 0; i < n; i++) pq.push({(int)adj[i].size(), i});
    std::vector<char> dead(n, 0);
    std::vector<int> perm;
    perm.reserve(n);
    std::vector<int> tmp;
    while ((int)perm.size() < n) {
        auto [d, v] = pq.top();
        pq.pop();
        if (dead[v] || d != (int)adj[v].size()) continue;
        dead[v] = 1;
        perm.push_back(v);
        const std::vector<int> nv = adj[v];
        for (int u : nv) {
            tmp.clear();
            std::set_union(adj[u].begin(), adj[u].end(), nv.begin(), nv.end(), std::back_inserter(tmp));
            tmp.erase(std::remove_if(tmp.begin(), tmp.end(), [&](int x) { return x == u || x == v; }),
                      tmp.end());
            adj[u].swap(tmp);
            pq.push({(int)adj[u].size(), u});
        }
        adj[v].clear();
        adj[v].shrink_to_fit();
    }
    return perm;
}

namespace {

// Symbolic analysis for a given elimination order perm (node j eliminates
// block perm[j]); gp/gi is the bus graph without the slack in block numbering.
Symbolic analyze_order(const pf_network &net, const YPattern &yp, int tail_width,
                       const std::vector<int> &perm, const std::vector<int> &gp, const std::vector<int> &gi) {
    Symbolic s;
    const int nB = net.nB, sl = net.slack;
    const int n = nB - 1;
    s.n = n;
    auto blk = [&](int k) { return k - (k > sl ? 1 : 0); };
    s.perm = perm;
    s.iperm.assign(n, 0);
    for (int j = 0; j < n; j++) s.iperm[s.perm[j]] = j;
    // upper adjacency in elimination order
    std::vector<std::vector<int>> up(n);
    for (int b = 0; b < n; b++)
        for (int p = gp[b]; p < gp[b + 1]; p++) {
            int a = s.iperm[b], c = s.iperm[gi[p]];
            if (a < c) up[a].push_back(c);
            else if (c < a) up[c].push_back(a);
        }
    // column structures by merging children (elimination tree):
    // struct(j) = (upper adjacency of j  U  struct of every child) \ {<= j}
    std::vector<std::vector<int>> st(n), children(n);
    s.parent.assign(n, -1);
    std::vector<int> tmp;
    for (int j = 0; j < n; j++) {
        std::vector<int> cur = up[j];
        std::sort(cur.begin(), cur.end());
        cur.erase(std::unique(cur.begin(), cur.end()), cur.end());
        for (int c : children[j]) {
            tmp.clear();
            std::set_union(cur.begin(), cur.end(), st[c].begin(), st[c].end(), std::back_inserter(tmp));
            cur.swap(tmp);
        }
        cur.erase(std::remove_if(cur.begin(), cur.end(), [&](int x) { return x <= j; }), cur.end());
        if (!cur.empty()) {
            s.parent[j] = cur[0];
            children[cur[0]].push_back(j);
        }
        st[j].swap(cur);
    }
    s.colptr.assign(n + 1, 0);
    for (int j = 0; j < n; j++) s.colptr[j + 1] = s.colptr[j] + (int)st[j].size();
    s.nnzL = s.colptr[n];
    s.rowind.resize(s.nnzL);
    for (int j = 0; j < n; j++) std::copy(st[j].begin(), st[j].end(), s.rowind.begin() + s.colptr[j]);
    auto pos = [&](int i, int j) -> int {   // position of row i in column j, -1 if absent
        auto b = s.rowind.begin() + s.colptr[j], e = s.rowind.begin() + s.colptr[j + 1];
        auto it = std::lower_bound(b, e, i);
        return (it != e && *it == i) ? (int)(it - s.rowind.begin()) : -1;
    };
    // levels: leaves 0, parent above every child
    s.level.assign(n, 0);
    for (int j = 0; j < n; j++)
        if (s.parent[j] >= 0) s.level[s.parent[j]] = std::max(s.level[s.parent[j]], s.level[j] + 1);
    s.nlev = 0;
    for (int j = 0; j < n; j++) s.nlev = std::max(s.nlev, s.level[j] + 1);
    s.lvl_ptr.assign(s.nlev + 1, 0);
    for (int j = 0; j < n; j++) s.lvl_ptr[s.level[j] + 1]++;
    for (int l = 0; l < s.nlev; l++) s.lvl_ptr[l + 1] += s.lvl_ptr[l];
    s.lvl_nodes.resize(n);
    {
        std::vector<int> fill(s.lvl_ptr.begin(), s.lvl_ptr.end() - 1);
        for (int j = 0; j < n; j++) s.lvl_nodes[fill[s.level[j]]++] = j;
    }
    s.tail_start = s.nlev;
    while (s.tail_start > 0 &&
           s.lvl_ptr[s.tail_start] - s.lvl_ptr[s.tail_start - 1] <= tail_width)
        s.tail_start--;
    // row structures (forward solve): for row i, every k < i with i in struct(col k)
    s.rs_ptr.assign(n + 1, 0);
    for (int p = 0; p < s.nnzL; p++) s.rs_ptr[s.rowind[p] + 1]++;
    for (int i = 0; i < n; i++) s.rs_ptr[i + 1] += s.rs_ptr[i];
    s.rs_k.resize(s.nnzL);
    s.rs_p.resize(s.nnzL);
    {
        std::vector<int> fill(s.rs_ptr.begin(), s.rs_ptr.end() - 1);
        for (int k = 0; k < n; k++)
            for (int p = s.colptr[k]; p < s.colptr[k + 1]; p++) {
                int i = s.rowind[p];
                s.rs_k[fill[i]] = k;
                s.rs_p[fill[i]] = p;
                fill[i]++;
            }
    }
    s.lr_pos.resize(s.nnzL);
    for (int q = 0; q < s.nnzL; q++) s.lr_pos[s.rs_p[q]] = q;
    // factor program
    s.fd_ptr.assign(n + 1, 0);
    s.fc_ptr.assign(1, 0);
    std::vector<int> local(n, -1);
    std::vector<std::vector<std::pair<int, int>>> dl;
    for (int j = 0; j < n; j++) {
        int nd = 1 + (s.colptr[j + 1] - s.colptr[j]);
        s.fd_ptr[j + 1] = s.fd_ptr[j] + nd;
        dl.assign(nd, {});
        for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) local[s.rowind[p]] = 1 + (p - s.colptr[j]);
        for (int q = s.rs_ptr[j]; q < s.rs_ptr[j + 1]; q++) {
            int k = s.rs_k[q], pjk = s.rs_p[q];
            dl[0].push_back({pjk, pjk});
            for (int p = pjk + 1; p < s.colptr[k + 1]; p++) {   // rows of col k below j
                int i = s.rowind[p];
                dl[local[i]].push_back({p, pjk});
            }
        }
        for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) local[s.rowind[p]] = -1;
        for (int d = 0; d < nd; d++) {
            for (auto &pr : dl[d]) {
                s.fc_a.push_back(pr.first);
                s.fc_b.push_back(pr.second);
            }
            s.fc_ptr.push_back((int)s.fc_a.size());
        }
    }
    // flattened destinations per level
    s.ld_ptr.assign(s.nlev + 1, 0);
    for (int l = 0; l < s.nlev; l++) {
        s.ld_ptr[l + 1] = s.ld_ptr[l];
        for (int q = s.lvl_ptr[l]; q < s.lvl_ptr[l + 1]; q++) {
            int j = s.lvl_nodes[q];
            for (int d = s.fd_ptr[j]; d < s.fd_ptr[j + 1]; d++) {
                s.ld_node.push_back(j);
                s.ld_d.push_back(d);
            }
            s.ld_ptr[l + 1] += s.fd_ptr[j + 1] - s.fd_ptr[j];
        }
    }
    // Jacobian slot map
    s.slot_e.assign(s.nslot(), -1);
    for (int k = 0; k < nB; k++) {
        if (k == sl) continue;
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (m == sl) continue;
            int a = s.iperm[blk(k)], c = s.iperm[blk(m)];
            int slot;
            if (a == c) slot = a;
            else if (a > c) slot = n + pos(a, c);
            else slot = n + s.nnzL + pos(c, a);
            s.slot_e[slot] = e;
        }
    }
    return s;
}

}  // namespace

// Minimum degree, then the same order renumbered level by level (stable in
// the minimum-degree order inside a level).  Children sit on lower levels
// than their parents, so the renumbering is a topological order of the same
// elimination tree: the fill is unchanged, while every level becomes a
// contiguous range of nodes (lvl_nodes is the identity) and the factor
// program, the row structures and the flattened destinations of a level are
// contiguous too.
Symbolic analyze(const pf_network &net, const YPattern &yp, int tail_width, int chain_cap, int top_width) {
    const int nB = net.nB, sl = net.slack;
    const int n = nB - 1;
    auto blk = [&](int k) { return k - (k > sl ? 1 : 0); };
    // bus graph without the slack, in block numbering
    std::vector<int> gp(n + 1, 0), gi;
    for (int k = 0; k < nB; k++) {
        if (k == sl) continue;
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (m == sl || m == k) continue;
            gi.push_back(blk(m));
        }
        gp[blk(k) + 1] = (int)gi.size();
    }
    const std::vector<int> md = min_degree(n, gp, gi);
    const Symbolic s0 = analyze_order(net, yp, tail_width, md, gp, gi);
    std::vector<int> by_level(s0.lvl_nodes.begin(), s0.lvl_nodes.end());   // nodes grouped by level
    std::vector<int> perm(n);
    for (int t = 0; t < n; t++) perm[t] = md[by_level[t]];
    Symbolic s = analyze_order(net, yp, tail_width, perm, gp, gi);
    // dense top: the last elimination-tree levels, each at most top_width
    // nodes wide, at most chain_cap nodes in all.  They are factored as one
    // dense block (its Schur complement, inverted), so the level walks end
    // below them; entries of the block outside the pattern of L are zero.
    int m = 0, taken = 0;
    while (taken < s.nlev) {
        const int l = s.nlev - 1 - taken, w = s.lvl_ptr[l + 1] - s.lvl_ptr[l];
        if (w > top_width || m + w > chain_cap) break;
        m += w;
        taken++;
    }
    if (m < 4) m = taken = 0;
    s.ch_m = m;
    s.ch_c0 = n - m;
    s.nlev_s = s.nlev - taken;
    if (m > 0) {
        std::vector<int> pcol(s.nnzL);
        for (int k = 0; k < n; k++)
            for (int p = s.colptr[k]; p < s.colptr[k + 1]; p++) pcol[p] = k;
        const int d0 = s.fd_ptr[s.ch_c0], d1 = s.fd_ptr[n];
        s.ch_split.resize(d1 - d0);
        for (int d = d0; d < d1; d++) {
            int q = s.fc_ptr[d];
            while (q < s.fc_ptr[d + 1] && pcol[s.fc_b[q]] < s.ch_c0) q++;   // pairs ascend in k
            s.ch_split[d - d0] = q;
        }
    }
    return s;
}

void partition_subtrees(Symbolic &s, int max_parts) {
    const int n = s.n, T = s.tail_start;
    s.n_parts = 0;
    s.part_ptr.clear();
    s.part_nodes.clear();
    if (T == 0 || max_parts <= 0) return;
    const int nw = s.lvl_ptr[T];   // nodes [0, nw) are below the tail (level order)
    // work of a node: its destinations and update pairs; summed over subtrees
    std::vector<long long> work(nw);
    for (int j = 0; j < nw; j++)
        work[j] += (s.fd_ptr[j + 1] - s.fd_ptr[j]) + (s.fc_ptr[s.fd_ptr[j + 1]] - s.fc_ptr[s.fd_ptr[j]]);
    std::vector<int> roots;
    for (int j = 0; j < nw; j++) {   // children precede parents
        const int p = s.parent[j];
        if (p >= 0 && p < nw) work[p] += work[j];
        else roots.push_back(j);
    }
    std::sort(roots.begin(), roots.end(), [&](int a, int b) { return work[a] != work[b] ? work[a] > work[b] : a < b; });
    const int np = std::min<int>(max_parts, (int)roots.size());
    std::vector<long long> load(np, 0);
    std::vector<int> part(n, -1);
    for (int r : roots) {
        int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        part[r] = g;
        load[g] += work[r];
    }
    for (int j = nw - 1; j >= 0; j--)   // parents precede children going down
        if (part[j] < 0) part[j] = part[s.parent[j]];
    s.n_parts = np;
    s.part_ptr.assign((size_t)np * T + 1, 0);
    for (int j = 0; j < nw; j++) s.part_ptr[(size_t)part[j] * T + s.level[j] + 1]++;
    for (size_t i = 0; i + 1 < s.part_ptr.size(); i++) s.part_ptr[i + 1] += s.part_ptr[i];
    s.part_nodes.resize(nw);
    std::vector<int> fill(s.part_ptr.begin(), s.part_ptr.end() - 1);
    for (int j = 0; j < nw; j++) s.part_nodes[fill[(size_t)part[j] * T + s.level[j]]++] = j;
}

void chunk_tail(Symbolic &s, int chunk_pairs) {
    s.ck_lvl.assign(1, 0);
    s.ck_q0.clear();
    s.ck_q1.clear();
    s.te_node.clear();
    s.te_p.clear();
    s.te_ck.clear();
    s.ck_max = 0;
    for (int l = s.tail_start; l < s.nlev_s; l++) {
        const int before = (int)s.ck_q0.size();
        for (int e = s.ld_ptr[l]; e < s.ld_ptr[l + 1]; e++) {
            const int j = s.ld_node[e], d = s.ld_d[e], t = d - s.fd_ptr[j];
            s.te_node.push_back(j);
            s.te_p.push_back(t == 0 ? -1 : s.colptr[j] + t - 1);
            s.te_ck.push_back((int)s.ck_q0.size());
            for (int q = s.fc_ptr[d]; q < s.fc_ptr[d + 1]; q += chunk_pairs) {
                s.ck_q0.push_back(q);
                s.ck_q1.push_back(std::min(q + chunk_pairs, s.fc_ptr[d + 1]) | (t == 0 ? kChunkDiag : 0));
            }
        }
        s.ck_max = std::max(s.ck_max, (int)s.ck_q0.size() - before);
        s.ck_lvl.push_back((int)s.ck_q0.size());
    }
    s.te_ck.push_back((int)s.ck_q0.size());
}

TailSweeps tail_sweeps(const Symbolic &s) {
    TailSweeps t;
    t.t0 = s.lvl_ptr[s.tail_start];
    t.t1 = s.lvl_ptr[s.nlev_s];
    t.fx_ptr.assign(1, 0);
    t.fi_ptr.assign(1, 0);
    t.bx_ptr.assign(1, 0);
    t.bi_ptr.assign(1, 0);
    for (int j = t.t0; j < t.t1; j++) {
        for (int q = s.rs_ptr[j]; q < s.rs_ptr[j + 1]; q++) {
            const int k = s.rs_k[q];
            if (k < t.t0) {
                t.fx_k.push_back(k);
                t.fx_q.push_back(q);
            } else {
                t.fi_k.push_back(k - t.t0);
                t.fi_q.push_back(q);
            }
        }
        for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) {
            const int i = s.rowind[p];
            if (i >= t.t1) {
                t.bx_i.push_back(i);
                t.bx_p.push_back(p);
            } else {
                t.bi_i.push_back(i - t.t0);
                t.bi_p.push_back(p);
            }
        }
        t.fx_ptr.push_back((int)t.fx_k.size());
        t.fi_ptr.push_back((int)t.fi_k.size());
        t.bx_ptr.push_back((int)t.bx_i.size());
        t.bi_ptr.push_back((int)t.bi_i.size());
    }
    return t;
}

// Narrow steps with long product lists take 16 or 4 lanes per destination:
// a sum of up to ~40 products then costs a few LDS loads per lane and a
// short tree instead of a chain of dependent additions in one lane.  The
// width depends on the program only, so every launch shape sums alike.
int flat_group_width(int nd, int maxpd) {
    if (nd <= 64 && maxpd >= 16) return 16;
    if (nd <= 256 && maxpd >= 6) return 4;
    return 1;
}

void flat_group_sum(const double *v, int lo, int hi, int G, double *out) {
    double part[16][4] = {};
    for (int l = 0; l < G; l++)
        for (int q = lo + l; q < hi; q += G)
            for (int c = 0; c < 4; c++) part[l][c] += v[4 * (size_t)q + c];
    for (int w = 1; w < G; w *= 2)   // (p0 + p1) + (p2 + p3), then pairs of those, ...
        for (int l = 0; l < G; l += 2 * w)
            for (int c = 0; c < 4; c++) part[l][c] += part[l + w][c];
    for (int c = 0; c < 4; c++) out[c] = part[0][c];
}

namespace {

// one level step of a flat program: the destinations of the given nodes,
// the pivots' pairs first (pairs from sources k < t0 left out)
void flat_step(const Symbolic &s, FlatTail &f, const int *nodes, int nn, int t0, const std::vector<int> &pcol,
               std::vector<int> &dlo, std::vector<int> &dhi, std::vector<int> &pos_de) {
    const int base = f.fl_ptr.back();
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) f.fl_split.push_back((int)f.fl_a.size() - base);
        for (int a = 0; a < nn; a++) {
            const int j = nodes[a];
            for (int d = s.fd_ptr[j]; d < s.fd_ptr[j + 1]; d++) {
                const int t = d - s.fd_ptr[j];
                if ((t == 0) != (pass == 0)) continue;
                int q = s.fc_ptr[d];
                while (q < s.fc_ptr[d + 1] && pcol[s.fc_b[q]] < t0) q++;   // pairs ascend in k
                if (t > 0) pos_de[s.colptr[j] + t - 1] = (int)f.de_node.size();
                f.de_node.push_back(j);
                f.de_p.push_back(t == 0 ? -1 : s.colptr[j] + t - 1);
                f.de_lr.push_back(t == 0 ? 0 : s.lr_pos[s.colptr[j] + t - 1]);
                f.de_q0.push_back(s.fc_ptr[d]);
                f.de_ext.push_back(q);
                f.de_lo.push_back((int)f.fl_a.size() - base);
                for (; q < s.fc_ptr[d + 1]; q++) {
                    f.fl_a.push_back(s.fc_a[q]);
                    f.fl_b.push_back(s.fc_b[q]);
                }
                f.de_hi.push_back((int)f.fl_a.size() - base);
                if (t == 0) {
                    dlo[j] = f.de_lo.back();
                    dhi[j] = f.de_hi.back();
                }
                f.de_dlo.push_back(dlo[j]);
                f.de_dhi.push_back(dhi[j]);
            }
        }
    }
    const int np = (int)f.fl_a.size() - base;
    const int E0 = f.de_ptr.back(), E1 = (int)f.de_node.size();
    // pairs whose source is a node of the previous step: the slots of their
    // two positions among that step's destinations
    const int P0 = f.de_ptr.size() >= 2 ? f.de_ptr[f.de_ptr.size() - 2] : E0;
    for (int q = base; q < base + np; q++) {
        const int ea = pos_de[f.fl_a[q]], eb = pos_de[f.fl_b[q]];
        const bool fresh = ea >= P0 && ea < E0;
        if (fresh != (eb >= P0 && eb < E0)) f.stage_ok = false;   // both lie in column k: cannot happen
        f.fl_s.push_back(fresh && E0 - P0 <= kFlatStageMax ? (ea - P0) | (eb - P0) << 16 : -1);
    }
    int maxpd = 0;
    for (int E = E0; E < E1; E++) maxpd = std::max(maxpd, f.de_hi[E] - f.de_lo[E]);
    f.st_g.push_back(flat_group_width(E1 - E0, maxpd));
    f.max_pairs = std::max(f.max_pairs, np);
    f.max_col = std::max(f.max_col, np - f.fl_split.back());
    f.fl_ptr.push_back((int)f.fl_a.size());
    f.de_ptr.push_back((int)f.de_node.size());
}

void flat_seal(FlatTail &f) {
    f.fl_a.push_back(0);
    f.fl_b.push_back(0);
    f.fl_s.push_back(-1);
    f.de_node.push_back(0);
    f.de_p.push_back(-1);
    f.de_q0.push_back(0);
    f.de_ext.push_back(0);
    f.de_lo.push_back(0);
    f.de_hi.push_back(0);
    f.de_dlo.push_back(0);
    f.de_dhi.push_back(0);
    f.de_lr.push_back(0);
}

std::vector<int> position_columns(const Symbolic &s) {
    std::vector<int> pcol(s.nnzL);
    for (int k = 0; k < s.n; k++)
        for (int p = s.colptr[k]; p < s.colptr[k + 1]; p++) pcol[p] = k;
    return pcol;
}

}  // namespace

FlatTail flat_tail(const Symbolic &s, int lv0, bool split_below) {
    FlatTail f;
    f.lv0 = lv0;
    const int t0 = split_below ? s.lvl_ptr[lv0] : 0;
    const std::vector<int> pcol = position_columns(s);
    std::vector<int> dlo(s.n), dhi(s.n), pos_de(s.nnzL, -1);
    f.fl_ptr.assign(1, 0);
    f.de_ptr.assign(1, 0);
    for (int l = lv0; l < s.nlev_s; l++) {   // nodes are numbered level by level
        std::vector<int> nodes;
        for (int j = s.lvl_ptr[l]; j < s.lvl_ptr[l + 1]; j++) nodes.push_back(j);
        flat_step(s, f, nodes.data(), (int)nodes.size(), t0, pcol, dlo, dhi, pos_de);
    }
    flat_seal(f);
    return f;
}

FlatTail flat_parts(const Symbolic &s) {
    FlatTail f;
    const int T = s.tail_start;
    const std::vector<int> pcol = position_columns(s);
    std::vector<int> dlo(s.n), dhi(s.n), pos_de(s.nnzL, -1);
    f.fl_ptr.assign(1, 0);
    f.de_ptr.assign(1, 0);
    for (int g = 0; g < s.n_parts; g++)
        for (int l = 0; l < T; l++) {
            const int a = s.part_ptr[(size_t)g * T + l], b = s.part_ptr[(size_t)g * T + l + 1];
            flat_step(s, f, s.part_nodes.data() + a, b - a, 0, pcol, dlo, dhi, pos_de);
        }
    flat_seal(f);
    return f;
}

PartSweeps part_sweeps(const Symbolic &s) {
    PartSweeps ps;
    const int T = s.tail_start;
    std::vector<int> loc(s.n, -1), grp(s.n, -1);
    for (int g = 0; g < s.n_parts; g++) {
        const int u0 = s.part_ptr[(size_t)g * T], u1 = s.part_ptr[(size_t)(g + 1) * T];
        for (int u = u0; u < u1; u++) {
            loc[s.part_nodes[u]] = u - u0;
            grp[s.part_nodes[u]] = g;
        }
        ps.max_nodes = std::max(ps.max_nodes, u1 - u0);
    }
    ps.fi_ptr.assign(1, 0);
    ps.bi_ptr.assign(1, 0);
    ps.bx_ptr.assign(1, 0);
    for (int g = 0; g < s.n_parts; g++) {
        const int u0 = s.part_ptr[(size_t)g * T], u1 = s.part_ptr[(size_t)(g + 1) * T];
        const int f0 = (int)ps.fi_k.size(), b0 = (int)ps.bi_i.size();
        for (int u = u0; u < u1; u++) {
            const int j = s.part_nodes[u];
            for (int q = s.rs_ptr[j]; q < s.rs_ptr[j + 1]; q++) {   // sources: the group's own nodes
                ps.fi_k.push_back(loc[s.rs_k[q]]);
                ps.fi_q.push_back(q);
            }
            for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) {
                const int i = s.rowind[p];
                if (grp[i] == g) {
                    ps.bi_i.push_back(loc[i]);
                    ps.bi_p.push_back(p);
                } else {
                    ps.bx_i.push_back(i);
                    ps.bx_p.push_back(p);
                }
            }
            ps.fi_ptr.push_back((int)ps.fi_k.size());
            ps.bi_ptr.push_back((int)ps.bi_i.size());
            ps.bx_ptr.push_back((int)ps.bx_i.size());
        }
        ps.max_fi = std::max(ps.max_fi, (int)ps.fi_k.size() - f0);
        ps.max_bi = std::max(ps.max_bi, (int)ps.bi_i.size() - b0);
    }
    return ps;
}

// ------------------------------------------------------------- host numeric

namespace {

inline void mms(double *c, const double *a, const double *b) {   // c -= a*b (2x2)
    double c0 = c[0] - (a[0] * b[0] + a[1] * b[2]);
    double c1 = c[1] - (a[0] * b[1] + a[1] * b[3]);
    double c2 = c[2] - (a[2] * b[0] + a[3] * b[2]);
    double c3 = c[3] - (a[2] * b[1] + a[3] * b[3]);
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

inline void mul2(double *c, const double *a, const double *b) {   // c = a*b (2x2)
    double c0 = a[0] * b[0] + a[1] * b[2], c1 = a[0] * b[1] + a[1] * b[3];
    double c2 = a[2] * b[0] + a[3] * b[2], c3 = a[2] * b[1] + a[3] * b[3];
    c[0] = c0; c[1] = c1; c[2] = c2; c[3] = c3;
}

// Trailing dense chain: Schur complement of the chain block and its
// Gauss-Jordan inverse, stored row-major at block n + 3 nnzL (the
// arithmetic of the device chain factorisation).
bool host_chain_factor(const Symbolic &s, double *blk) {
    const int n = s.n, m = s.ch_m, c0 = s.ch_c0, L0 = n, U0 = n + s.nnzL, SB = n + 3 * s.nnzL;
    double *S = blk + 4 * (size_t)SB;
    std::fill(S, S + 4 * (size_t)m * m, 0.0);   // entries outside the pattern of L
    auto at = [&](int i, int j) { return S + 4 * ((size_t)i * m + j); };
    const int dbase = s.fd_ptr[c0];
    for (int j = c0; j < n; j++)
        for (int d = s.fd_ptr[j]; d < s.fd_ptr[j + 1]; d++) {
            const int t = d - s.fd_ptr[j], a = j - c0, qend = s.ch_split[d - dbase];
            if (t == 0) {
                double D[4];
                std::copy(blk + 4 * (size_t)j, blk + 4 * (size_t)j + 4, D);
                for (int q = s.fc_ptr[d]; q < qend; q++)
                    mms(D, blk + 4 * (size_t)(L0 + s.fc_a[q]), blk + 4 * (size_t)(U0 + s.fc_b[q]));
                std::copy(D, D + 4, at(a, a));
            } else {
                const int p = s.colptr[j] + t - 1, i = s.rowind[p] - c0;
                double Lb[4], Ub[4];
                std::copy(blk + 4 * (size_t)(L0 + p), blk + 4 * (size_t)(L0 + p) + 4, Lb);
                std::copy(blk + 4 * (size_t)(U0 + p), blk + 4 * (size_t)(U0 + p) + 4, Ub);
                for (int q = s.fc_ptr[d]; q < qend; q++) {
                    const int pik = s.fc_a[q], pjk = s.fc_b[q];
                    mms(Lb, blk + 4 * (size_t)(L0 + pik), blk + 4 * (size_t)(U0 + pjk));
                    mms(Ub, blk + 4 * (size_t)(L0 + pjk), blk + 4 * (size_t)(U0 + pik));
                }
                std::copy(Lb, Lb + 4, at(i, a));
                std::copy(Ub, Ub + 4, at(a, i));
            }
        }
    std::vector<double> row(4 * (size_t)m), col(4 * (size_t)m);
    for (int k = 0; k < m; k++) {
        for (int t = 0; t < m; t++) {
            std::copy(at(k, t), at(k, t) + 4, &row[4 * t]);
            std::copy(at(t, k), at(t, k) + 4, &col[4 * t]);
        }
        const double *pv = &row[4 * k];
        const double det = pv[0] * pv[3] - pv[1] * pv[2];
        if (!(det != 0.0) || !std::isfinite(det)) return false;
        const double P[4] = {pv[3] / det, -pv[1] / det, -pv[2] / det, pv[0] / det};
        for (int i = 0; i < m; i++)
            for (int j = 0; j < m; j++) {
                double *sij = at(i, j);
                if (i == k) {
                    if (j == k) std::copy(P, P + 4, sij);
                    else mul2(sij, P, &row[4 * j]);
                } else if (j == k) {
                    mul2(sij, &col[4 * i], P);
                    for (int c = 0; c < 4; c++) sij[c] = -sij[c];
                } else {
                    double u[4];
                    mul2(u, P, &row[4 * j]);
                    mms(sij, &col[4 * i], u);
                }
            }
    }
    return true;
}

}  // namespace

bool host_block_factor(const Symbolic &s, double *blk) {
    const int n = s.n, L0 = n, U0 = n + s.nnzL;
    const int nf = s.ch_m > 0 ? s.ch_c0 : n;   // nodes factored one by one
    for (int j = 0; j < nf; j++) {
        double dinv[4];
        for (int d = s.fd_ptr[j]; d < s.fd_ptr[j + 1]; d++) {
            int t = d - s.fd_ptr[j];
            if (t == 0) {
                double *D = blk + 4 * (size_t)j;
                for (int q = s.fc_ptr[d]; q < s.fc_ptr[d + 1]; q++)
                    mms(D, blk + 4 * (size_t)(L0 + s.fc_a[q]), blk + 4 * (size_t)(U0 + s.fc_b[q]));
                double det = D[0] * D[3] - D[1] * D[2];
                if (!(det != 0.0) || !std::isfinite(det)) return false;
                dinv[0] = D[3] / det; dinv[1] = -D[1] / det;
                dinv[2] = -D[2] / det; dinv[3] = D[0] / det;
                for (int k = 0; k < 4; k++) D[k] = dinv[k];
            } else {
                int p = s.colptr[j] + t - 1;
                double *Lb = blk + 4 * (size_t)(L0 + p), *Ub = blk + 4 * (size_t)(U0 + p);
                for (int q = s.fc_ptr[d]; q < s.fc_ptr[d + 1]; q++) {
                    int pik = s.fc_a[q], pjk = s.fc_b[q];
                    mms(Lb, blk + 4 * (size_t)(L0 + pik), blk + 4 * (size_t)(U0 + pjk));
                    mms(Ub, blk + 4 * (size_t)(L0 + pjk), blk + 4 * (size_t)(U0 + pik));
                }
                double l0 = Lb[0] * dinv[0] + Lb[1] * dinv[2];
                double l1 = Lb[0] * dinv[1] + Lb[1] * dinv[3];
                double l2 = Lb[2] * dinv[0] + Lb[3] * dinv[2];
                double l3 = Lb[2] * dinv[1] + Lb[3] * dinv[3];
                Lb[0] = l0; Lb[1] = l1; Lb[2] = l2; Lb[3] = l3;
            }
        }
    }
    if (s.ch_m > 0) return host_chain_factor(s, blk);
    return true;
}

bool host_flat_factor(const Symbolic &s, const FlatTail &parts, const FlatTail &top, double *blk) {
    const int n = s.n, L0 = n, U0 = n + s.nnzL;
    const int nf = s.ch_m > 0 ? s.ch_c0 : n;
    std::vector<double> inv(4 * (size_t)n, 0.0);   // inverse pivots; the pivot slots keep A_jj
    auto L = [&](int p) { return blk + 4 * (size_t)(L0 + p); };
    auto U = [&](int p) { return blk + 4 * (size_t)(U0 + p); };
    auto run = [&](const FlatTail &f) -> bool {
        const int nsteps = (int)f.fl_ptr.size() - 1;
        // the previous step's destinations' L and U blocks by their index in
        // its list: the pairs with a source there (fl_s >= 0) read them here
        std::vector<double> stage, next;
        for (int t = 0; t < nsteps; t++) {
            const int base = f.fl_ptr[t], np = f.fl_ptr[t + 1] - base, ns = f.fl_split[t], G = f.st_g[t];
            auto sum = [G](const std::vector<double> &v, int lo, int hi, double *a) {
                flat_group_sum(v.data(), lo, hi, G, a);
            };
            std::vector<double> pa(4 * (size_t)np), pu(4 * (size_t)std::max(np - ns, 0));
            for (int e = 0; e < np; e++) {   // every product from earlier steps' blocks
                const int fs = f.fl_s[base + e];
                const double *la = L(f.fl_a[base + e]), *ua = U(f.fl_a[base + e]);
                const double *lb = L(f.fl_b[base + e]), *ub = U(f.fl_b[base + e]);
                if (fs >= 0 && f.stage_ok) {
                    const size_t sa = (size_t)(fs & 0xffff), sb = (size_t)(fs >> 16);
                    if (8 * std::max(sa, sb) + 8 > stage.size()) return false;
                    la = &stage[8 * sa];
                    ua = &stage[8 * sa + 4];
                    lb = &stage[8 * sb];
                    ub = &stage[8 * sb + 4];
                }
                mul2(&pa[4 * (size_t)e], la, ub);
                if (e >= ns) mul2(&pu[4 * (size_t)(e - ns)], lb, ua);
            }
            next.assign(8 * (size_t)(f.de_ptr[t + 1] - f.de_ptr[t]), std::nan(""));
            for (int E = f.de_ptr[t]; E < f.de_ptr[t + 1]; E++) {
                const int j = f.de_node[E], p = f.de_p[E];
                double a[4], P[4];
                sum(pa, f.de_lo[E], f.de_hi[E], a);
                double ad[4];
                if (p < 0) std::copy(a, a + 4, ad);
                else sum(pa, f.de_dlo[E], f.de_dhi[E], ad);
                const double *D = blk + 4 * (size_t)j;
                for (int c = 0; c < 4; c++) P[c] = D[c] - ad[c];
                double det = P[0] * P[3] - P[1] * P[2];
                if (!(det != 0.0) || !std::isfinite(det)) return false;
                const double Pi[4] = {P[3] / det, -P[1] / det, -P[2] / det, P[0] / det};
                if (p < 0) {
                    std::copy(Pi, Pi + 4, &inv[4 * (size_t)j]);
                } else {
                    double u[4], Lb[4], lb[4];
                    sum(pu, f.de_lo[E] - ns, f.de_hi[E] - ns, u);
                    for (int c = 0; c < 4; c++) {
                        Lb[c] = L(p)[c] - a[c];
                        U(p)[c] -= u[c];
                    }
                    mul2(lb, Lb, Pi);
                    std::copy(lb, lb + 4, L(p));
                    std::copy(lb, lb + 4, &next[8 * (size_t)(E - f.de_ptr[t])]);
                    std::copy(U(p), U(p) + 4, &next[8 * (size_t)(E - f.de_ptr[t]) + 4]);
                }
            }
            stage.swap(next);
        }
        return true;
    };
    if (!parts.fl_ptr.empty() && !run(parts)) return false;
    // the narrow top's pairs from below it, applied first
    for (int E = 0; E < (top.de_ptr.empty() ? 0 : top.de_ptr.back()); E++) {
        const int j = top.de_node[E], p = top.de_p[E];
        double a[4] = {0, 0, 0, 0}, u[4] = {0, 0, 0, 0}, t4[4];
        for (int q = top.de_q0[E]; q < top.de_ext[E]; q++) {
            mul2(t4, L(s.fc_a[q]), U(s.fc_b[q]));
            for (int c = 0; c < 4; c++) a[c] += t4[c];
            if (p >= 0) {
                mul2(t4, L(s.fc_b[q]), U(s.fc_a[q]));
                for (int c = 0; c < 4; c++) u[c] += t4[c];
            }
        }
        double *v0 = p < 0 ? blk + 4 * (size_t)j : L(p);
        for (int c = 0; c < 4; c++) v0[c] -= a[c];
        if (p >= 0)
            for (int c = 0; c < 4; c++) U(p)[c] -= u[c];
    }
    if (!top.fl_ptr.empty() && !run(top)) return false;
    for (int j = 0; j < nf; j++) std::copy(&inv[4 * (size_t)j], &inv[4 * (size_t)j] + 4, blk + 4 * (size_t)j);
    if (s.ch_m > 0) return host_chain_factor(s, blk);
    return true;
}

void host_split_solve(const Symbolic &s, const TailSweeps &ts, const PartSweeps &ps, const double *blk,
                      double *x) {
    const int n = s.n, L0 = n, U0 = n + s.nnzL, t0 = ts.t0, t1 = ts.t1, T = s.tail_start;
    auto Lr = [&](int q) { return blk + 4 * (size_t)(L0 + s.rs_p[q]); };   // L by row position q
    auto sub = [&](double *y, const double *B, int k) {
        y[0] -= B[0] * x[2 * k] + B[1] * x[2 * k + 1];
        y[1] -= B[2] * x[2 * k] + B[3] * x[2 * k + 1];
    };
    const bool groups = s.n_parts > 0 && !ps.fi_ptr.empty();
    if (groups) {   // below the top: each subtree group level by level, entries by local index
        for (int g = 0; g < s.n_parts; g++) {
            const int u0 = s.part_ptr[(size_t)g * T];
            for (int u = u0; u < s.part_ptr[(size_t)(g + 1) * T]; u++) {   // level-major within the group
                const int j = s.part_nodes[u];
                double y[2] = {x[2 * j], x[2 * j + 1]};
                for (int e = ps.fi_ptr[u]; e < ps.fi_ptr[u + 1]; e++) sub(y, Lr(ps.fi_q[e]), s.part_nodes[u0 + ps.fi_k[e]]);
                x[2 * j] = y[0];
                x[2 * j + 1] = y[1];
            }
        }
    } else {
        for (int j = 0; j < t0; j++) {   // below the top: every row entry
            double y[2] = {x[2 * j], x[2 * j + 1]};
            for (int q = s.rs_ptr[j]; q < s.rs_ptr[j + 1]; q++) sub(y, Lr(q), s.rs_k[q]);
            x[2 * j] = y[0];
            x[2 * j + 1] = y[1];
        }
    }
    for (int a = 0; a < t1 - t0; a++) {   // the top: entries below it first
        double y[2] = {x[2 * (t0 + a)], x[2 * (t0 + a) + 1]};
        for (int e = ts.fx_ptr[a]; e < ts.fx_ptr[a + 1]; e++) sub(y, Lr(ts.fx_q[e]), ts.fx_k[e]);
        x[2 * (t0 + a)] = y[0];
        x[2 * (t0 + a) + 1] = y[1];
    }
    for (int a = 0; a < t1 - t0; a++) {   // then its own entries, in node (level) order
        double y[2] = {x[2 * (t0 + a)], x[2 * (t0 + a) + 1]};
        for (int e = ts.fi_ptr[a]; e < ts.fi_ptr[a + 1]; e++) sub(y, Lr(ts.fi_q[e]), t0 + ts.fi_k[e]);
        x[2 * (t0 + a)] = y[0];
        x[2 * (t0 + a) + 1] = y[1];
    }
    if (s.ch_m > 0) {   // x_C = S^-1 (x_C - L_CN y_N)
        const int m = s.ch_m, c0 = s.ch_c0;
        std::vector<double> r(2 * (size_t)m);
        for (int a = 0; a < m; a++) {
            const int c = c0 + a;
            double y[2] = {x[2 * c], x[2 * c + 1]};
            for (int q = s.rs_ptr[c]; q < s.rs_ptr[c + 1] && s.rs_k[q] < c0; q++) sub(y, Lr(q), s.rs_k[q]);
            r[2 * a] = y[0];
            r[2 * a + 1] = y[1];
        }
        const double *S = blk + 4 * (size_t)(n + 3 * s.nnzL);
        for (int a = 0; a < m; a++) {
            double a0 = 0, a1 = 0;
            for (int b = 0; b < m; b++) {
                const double *Sab = S + 4 * ((size_t)a * m + b);
                a0 += Sab[0] * r[2 * b] + Sab[1] * r[2 * b + 1];
                a1 += Sab[2] * r[2 * b] + Sab[3] * r[2 * b + 1];
            }
            x[2 * (c0 + a)] = a0;
            x[2 * (c0 + a) + 1] = a1;
        }
    }
    std::vector<double> y(2 * (size_t)(t1 - t0));
    for (int a = 0; a < t1 - t0; a++) {   // the top backward: entries on the dense chain first
        y[2 * a] = x[2 * (t0 + a)];
        y[2 * a + 1] = x[2 * (t0 + a) + 1];
        for (int e = ts.bx_ptr[a]; e < ts.bx_ptr[a + 1]; e++) {
            double v[2] = {y[2 * a], y[2 * a + 1]};
            sub(v, blk + 4 * (size_t)(U0 + ts.bx_p[e]), ts.bx_i[e]);
            y[2 * a] = v[0];
            y[2 * a + 1] = v[1];
        }
    }
    for (int a = t1 - t0 - 1; a >= 0; a--) {   // then its own, in reverse node order
        const int j = t0 + a;
        double v[2] = {y[2 * a], y[2 * a + 1]};
        for (int e = ts.bi_ptr[a]; e < ts.bi_ptr[a + 1]; e++) sub(v, blk + 4 * (size_t)(U0 + ts.bi_p[e]), t0 + ts.bi_i[e]);
        const double *Di = blk + 4 * (size_t)j;
        x[2 * j] = Di[0] * v[0] + Di[1] * v[1];
        x[2 * j + 1] = Di[2] * v[0] + Di[3] * v[1];
    }
    if (groups) {   // each group: the entries above it first, then its own in reverse
        for (int g = 0; g < s.n_parts; g++) {
            const int u0 = s.part_ptr[(size_t)g * T], u1 = s.part_ptr[(size_t)(g + 1) * T];
            std::vector<double> yg(2 * (size_t)(u1 - u0));
            for (int u = u0; u < u1; u++) {
                const int j = s.part_nodes[u];
                double v[2] = {x[2 * j], x[2 * j + 1]};
                for (int e = ps.bx_ptr[u]; e < ps.bx_ptr[u + 1]; e++) sub(v, blk + 4 * (size_t)(U0 + ps.bx_p[e]), ps.bx_i[e]);
                yg[2 * (u - u0)] = v[0];
                yg[2 * (u - u0) + 1] = v[1];
            }
            for (int u = u1 - 1; u >= u0; u--) {
                const int j = s.part_nodes[u];
                double v[2] = {yg[2 * (u - u0)], yg[2 * (u - u0) + 1]};
                for (int e = ps.bi_ptr[u]; e < ps.bi_ptr[u + 1]; e++)
                    sub(v, blk + 4 * (size_t)(U0 + ps.bi_p[e]), s.part_nodes[u0 + ps.bi_i[e]]);
                const double *Di = blk + 4 * (size_t)j;
                x[2 * j] = Di[0] * v[0] + Di[1] * v[1];
                x[2 * j + 1] = Di[2] * v[0] + Di[3] * v[1];
            }
        }
        return;
    }
    for (int j = t0 - 1; j >= 0; j--) {
        double v[2] = {x[2 * j], x[2 * j + 1]};
        for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) sub(v, blk + 4 * (size_t)(U0 + p), s.rowind[p]);
        const double *Di = blk + 4 * (size_t)j;
        x[2 * j] = Di[0] * v[0] + Di[1] * v[1];
        x[2 * j + 1] = Di[2] * v[0] + Di[3] * v[1];
    }
}

void host_block_solve(const Symbolic &s, const double *blk, double *x) {
    const int n = s.n, L0 = n, U0 = n + s.nnzL;
    const int nf = s.ch_m > 0 ? s.ch_c0 : n;   // chain rows take only x_C - L_CN y_N here
    for (int j = 0; j < n; j++) {
        double y0 = x[2 * j], y1 = x[2 * j + 1];
        for (int q = s.rs_ptr[j]; q < s.rs_ptr[j + 1]; q++) {
            const int k = s.rs_k[q];
            if (k >= nf) break;                  // row entries ascend in k
            const double *L = blk + 4 * (size_t)(L0 + s.rs_p[q]);
            y0 -= L[0] * x[2 * k] + L[1] * x[2 * k + 1];
            y1 -= L[2] * x[2 * k] + L[3] * x[2 * k + 1];
        }
        x[2 * j] = y0;
        x[2 * j + 1] = y1;
    }
    if (s.ch_m > 0) {   // x_C = S^-1 (x_C - L_CN y_N)
        const int m = s.ch_m, c0 = s.ch_c0;
        const double *S = blk + 4 * (size_t)(n + 3 * s.nnzL);
        std::vector<double> r(x + 2 * (size_t)c0, x + 2 * (size_t)n);
        for (int a = 0; a < m; a++) {
            double a0 = 0, a1 = 0;
            for (int b = 0; b < m; b++) {
                const double *Sab = S + 4 * ((size_t)a * m + b);
                a0 += Sab[0] * r[2 * b] + Sab[1] * r[2 * b + 1];
                a1 += Sab[2] * r[2 * b] + Sab[3] * r[2 * b + 1];
            }
            x[2 * (c0 + a)] = a0;
            x[2 * (c0 + a) + 1] = a1;
        }
    }
    for (int j = nf - 1; j >= 0; j--) {
        double y0 = x[2 * j], y1 = x[2 * j + 1];
        for (int p = s.colptr[j]; p < s.colptr[j + 1]; p++) {
            const double *U = blk + 4 * (size_t)(U0 + p);
            int i = s.rowind[p];
            y0 -= U[0] * x[2 * i] + U[1] * x[2 * i + 1];
            y1 -= U[2] * x[2 * i] + U[3] * x[2 * i + 1];
        }
        const double *Di = blk + 4 * (size_t)j;
        x[2 * j] = Di[0] * y0 + Di[1] * y1;
        x[2 * j + 1] = Di[2] * y0 + Di[3] * y1;
    }
}

}  // namespace pf
