// Batched NR: many scenarios (perturbed injections, optional warm start) of
// one or more topologies solved in one persistent launch per topology.
//
// One workgroup owns one scenario for its whole NR solve and pulls the next
// scenario from a per-topology work counter when done, so uneven scenario
// sizes and iteration counts balance themselves.  Small topologies keep the
// whole working set (state vectors, mismatch, step, 2x2-block LU) in LDS;
// larger ones use a per-workgroup slice of a global workspace (L2 resident).
// The topology (Ybus, symbolic factor program) is shared read-only.
//
// Per scenario the pass is the same as the single-network solver
// (CPU/newtonRaphson.cpp:180-305): P,Q -> limits trigger -> mismatch ->
// convergence / switch at convergence -> Jacobian -> FP32 block LU ->
// solves -> FP64 refinement -> update.  A start that diverges (non-finite or
// huge mismatch, singular pivot, or no convergence) is retried from the flat
// start inside the same workgroup.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "kernels_decl.hpp"
#include "kernels_impl.hpp"
#include "pf_internal.hpp"

using namespace pf;

namespace pf {

struct BatchTopo {
    DevNet N;
    DevSym Y;
    const double *V0flat, *th0flat;     // flat start of this topology
    int nslot;                          // Jacobian slots
    int nblk;                           // factor blocks: slots, row-ordered L, chain inverse
    DevFlat F;                          // flat pair program of the narrow top (F.fl_ptr null: none)
};

struct BatchJob {
    int nscen;                          // scenarios of this topology
    const int *nscen_dev;               // when set: the count, read on the device
    const int *scen;                    // global scenario index
    const long long *off;               // bus offset of scenario s in the packed arrays
    const double *Pliq, *Qliq, *Qload;  // packed [sum nB]
    const double *V0, *th0;             // packed warm start, may be null
    double *Vout, *thout;               // packed results
    int *iters, *conv;
    double *maxmis;
    const long long *pv_off;            // PV offset of scenario s in the packed limQ
    unsigned char *limQout;             // packed final limQ [sum nPV]
    int *flips;                         // PV<->PQ flips of the final solve, per scenario
    int *counter;                       // work queue head
    char *gws;                          // global workspace (per workgroup slice)
    long long ws_bytes;                 // bytes per workgroup slice
    int max_iter, lim, refine;
    double tol, diverge;
};

constexpr int GL = 16;   // lanes per node in the in-workgroup LU and solves
// dense top block: at most kBatchChain nodes when the working set sits in
// LDS (MODE 2), kBatchChainWide otherwise; S is inverted in register tiles
// of up to kBatchTile<T> x kBatchTile<T> blocks per thread, or where the
// working set lives when the workgroup cannot cover it
constexpr int kBatchChain = 32;
constexpr int kBatchChainWide = 64;
template <typename T> constexpr int kBatchTile = 2;   // wider tiles spill in the persistent kernel
template <int MODE> constexpr int batch_chain_cap() { return MODE == 2 ? kBatchChain : kBatchChainWide; }
// levels of at least kWideLevel nodes: a lane group per node; narrower ones:
// the whole workgroup on the level's destinations.  A constant, not the
// number of lane groups, so that every workgroup size sums in the same order.
constexpr int kWideLevel = 32;
// the narrow levels at the top of the tree (below the dense block), when
// their update pairs are chunked: chunks of kBatchChunkPairs pairs, one lane
// group of kBatchChunkLanes per chunk (one pair per lane), at most
// kBatchMaxChunks chunks on a level (LDS partial sums; kernels whose working
// set lives in global memory only)
constexpr int kBatchChunkPairs = 8, kBatchChunkLanes = 8, kBatchMaxChunks = 512;
static_assert(kWideLevel <= 4 * kBatchChain, "the narrow-level pivots pass through the dense-top LDS buffer");
// update pairs loaded per round in the factorisation (registers are tight
// in the persistent kernel)
constexpr int kBatchUnroll = 1;

// per-scenario working set, in LDS or in a global slice
template <typename T>
struct Work {
    double *V, *th, *P, *Q, *ql, *g, *x, *r;
    double2 *E, *I;
    unsigned char *isq, *limQ;
    T *blk;
};

// nblk: factor blocks (Jacobian slots and the row-ordered copy of L)
template <typename T>
__host__ __device__ inline long long work_bytes(int nB, int n, int nPV, int nblk) {
    long long b = 0;
    b += 8LL * nB * 5;                 // V th P Q ql
    b += 8LL * 2 * n * 3;              // g x r
    b = (b + 15) & ~15LL;              // E and I hold double2: 16-byte aligned
    b += 16LL * nB * 2;                // E I
    b += (long long)sizeof(T) * 4 * nblk;
    b += nB + nPV + 16;
    return (b + 255) & ~255LL;
}

template <typename T>
__device__ __forceinline__ Work<T> carve(char *base, int nB, int n, int nPV, int nblk) {
    Work<T> w;
    char *p = base;
    auto take = [&](size_t bytes) { char *q = p; p += bytes; return q; };
    w.V = (double *)take(8 * (size_t)nB);
    w.th = (double *)take(8 * (size_t)nB);
    w.P = (double *)take(8 * (size_t)nB);
    w.Q = (double *)take(8 * (size_t)nB);
    w.ql = (double *)take(8 * (size_t)nB);
    w.g = (double *)take(16 * (size_t)n);
    w.x = (double *)take(16 * (size_t)n);
    w.r = (double *)take(16 * (size_t)n);
    p = base + ((p - base + 15) & ~(size_t)15);   // 8 mod 16 when nB is odd
    w.E = (double2 *)take(16 * (size_t)nB);
    w.I = (double2 *)take(16 * (size_t)nB);
    w.blk = (T *)take(sizeof(T) * 4 * (size_t)nblk);
    w.isq = (unsigned char *)take(nB);
    w.limQ = (unsigned char *)take(nPV);
    return w;
}

template <typename T>
__device__ __forceinline__ void wg_factor_node(const DevSym &Y, T *blk, int j, int lane, int *bad) {
    const int d0 = Y.fd_ptr[j], nd = Y.fd_ptr[j + 1] - d0;
    const int L0 = Y.n, U0 = Y.n + Y.nnzL;
    B4<T> dinv{0, 0, 0, 0};
    for (int base = 0; base < nd; base += GL) {
        const int t = base + lane;
        B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
        int p = -1;
        if (t < nd) {
            const int d = d0 + t;
            B4<T> sa{0, 0, 0, 0}, su{0, 0, 0, 0};
            if (t == 0) {
                pairs_acc<T, false, kBatchUnroll>(Y, blk, Y.fc_ptr[d], Y.fc_ptr[d + 1], 1, sa, su);
                a = load_b4(blk, j);
            } else {
                pairs_acc<T, true, kBatchUnroll>(Y, blk, Y.fc_ptr[d], Y.fc_ptr[d + 1], 1, sa, su);
                p = Y.colptr[j] + t - 1;
                a = load_b4(blk, L0 + p);
                u = load_b4(blk, U0 + p);
            }
            a = B4<T>{a.a - sa.a, a.b - sa.b, a.c - sa.c, a.d - sa.d};
            u = B4<T>{u.a - su.a, u.b - su.b, u.c - su.c, u.d - su.d};
        }
        if (base == 0) {
            B4<T> di{0, 0, 0, 0};
            if (lane == 0) {
                T det = a.a * a.d - a.b * a.c;
                if (!(det != (T)0) || !isfinite((double)det)) {
                    *bad = 1;
                    det = (T)1;
                }
                di = B4<T>{a.d / det, -a.b / det, -a.c / det, a.a / det};
                store_b4(blk, Y.di0 + j, di);
            }
            dinv.a = __shfl(di.a, 0, GL);
            dinv.b = __shfl(di.b, 0, GL);
            dinv.c = __shfl(di.c, 0, GL);
            dinv.d = __shfl(di.d, 0, GL);
        }
        if (t < nd && t > 0) {
            const B4<T> l = mul(a, dinv);
            store_b4(blk, L0 + p, l);
            store_b4(blk, U0 + Y.nnzL + Y.lr_pos[p], l);
            store_b4(blk, U0 + p, u);
        }
    }
}

// Narrow elimination-tree levels [l0, l1) (fewer than kWideLevel nodes
// each) by the whole workgroup, two barriers per level, with the same
// arithmetic as coop_factor_range: groups of G lanes take the level's
// destination blocks, the lanes of a group split the destination's update
// pairs (pair q + G*i for lane q) and reduce them in a fixed order; after a
// barrier the pivot's inverse is stored and each column block is scaled by
// it.  What is loaded is arranged for latency: the indices of a group's
// first destination on the next level and of each lane's first pair are
// loaded one level ahead, a destination's current blocks are loaded with its
// first pair's blocks, the updated pivots pass through LDS (s_piv, one per
// node of the level) and the first updated column block stays in registers,
// so the second half of a level waits on no global load.
template <typename T, int G>
__device__ __forceinline__ void wg_factor_narrow(const DevSym &Y, T *blk, int l0, int l1, int *bad, B4<T> *s_piv) {
    const int L0 = Y.n, U0 = Y.n + Y.nnzL, LR0 = Y.n + 2 * Y.nnzL, DI = Y.di0;
    const int grp = (int)threadIdx.x / G, lane = (int)threadIdx.x % G, ng = (int)blockDim.x / G;
    struct Dest {
        int j, t, q, qe, pa, pb, p;   // p: the column block's position (t > 0)
    };
    auto dest = [&](int e0, int e) {
        Dest r;
        r.j = Y.ld_node[e0 + e];
        const int d = Y.ld_d[e0 + e];
        r.t = d - Y.fd_ptr[r.j];
        r.q = Y.fc_ptr[d] + lane;
        r.qe = Y.fc_ptr[d + 1];
        r.pa = r.pb = 0;
        if (r.q < r.qe) {
            r.pa = Y.fc_a[r.q];
            r.pb = Y.fc_b[r.q];
        }
        r.p = r.t ? Y.colptr[r.j] + r.t - 1 : 0;
        return r;
    };
    Dest cur{};
    cur.j = -1;
    if (l0 < l1 && grp < Y.ld_ptr[l0 + 1] - Y.ld_ptr[l0]) cur = dest(Y.ld_ptr[l0], grp);
    for (int l = l0; l < l1; l++) {
        const int e0 = Y.ld_ptr[l], ne = Y.ld_ptr[l + 1] - e0, s0 = Y.lvl_ptr[l];
        Dest nxt{};
        nxt.j = -1;
        if (l + 1 < l1 && grp < Y.ld_ptr[l + 2] - Y.ld_ptr[l + 1]) nxt = dest(Y.ld_ptr[l + 1], grp);
        B4<T> keep{0, 0, 0, 0};   // the first destination's updated column block
        int keep_lrp = 0;
        for (int e = grp; e < ne; e += ng) {
            const Dest D = (e == grp) ? cur : dest(e0, e);
            B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
            // the destination's own blocks, loaded with its first pair's
            B4<T> v0{0, 0, 0, 0}, v1{0, 0, 0, 0};
            if (lane == 0) {
                v0 = load_b4(blk, D.t ? L0 + D.p : D.j);
                if (D.t) v1 = load_b4(blk, U0 + D.p);
            }
            const int lrp = (lane == 0 && D.t) ? Y.lr_pos[D.p] : 0;
            if (D.q < D.qe) {
                mma2(a, load_b4(blk, L0 + D.pa), load_b4(blk, U0 + D.pb));
                if (D.t != 0) mma2(u, load_b4(blk, L0 + D.pb), load_b4(blk, U0 + D.pa));
            }
            if (D.t == 0) pairs_acc<T, false, 1>(Y, blk, D.q + G, D.qe, G, a, u);
            else pairs_acc<T, true, 1>(Y, blk, D.q + G, D.qe, G, a, u);
            a = group_sum<T, G>(a);
            u = group_sum<T, G>(u);
            if (lane == 0) {
                if (D.t == 0) {
                    const B4<T> P{v0.a - a.a, v0.b - a.b, v0.c - a.c, v0.d - a.d};
                    store_b4(blk, D.j, P);
                    s_piv[D.j - s0] = P;
                } else {
                    const B4<T> Lb{v0.a - a.a, v0.b - a.b, v0.c - a.c, v0.d - a.d};
                    store_b4(blk, L0 + D.p, Lb);
                    store_b4(blk, U0 + D.p, B4<T>{v1.a - u.a, v1.b - u.b, v1.c - u.c, v1.d - u.d});
                    if (e == grp) {
                        keep = Lb;
                        keep_lrp = lrp;
                    }
                }
            }
        }
        __syncthreads();
        if (lane == 0) {
            for (int e = grp; e < ne; e += ng) {
                int j, t, p, lrp;
                B4<T> Lb;
                if (e == grp) {
                    j = cur.j; t = cur.t; p = cur.p; lrp = keep_lrp; Lb = keep;
                } else {
                    j = Y.ld_node[e0 + e];
                    t = Y.ld_d[e0 + e] - Y.fd_ptr[j];
                    p = t ? Y.colptr[j] + t - 1 : 0;
                    lrp = t ? Y.lr_pos[p] : 0;
                    Lb = t ? load_b4(blk, L0 + p) : B4<T>{0, 0, 0, 0};
                }
                const B4<T> P = s_piv[j - s0];
                T det = P.a * P.d - P.b * P.c;
                const bool sing = !(det != (T)0) || !isfinite((double)det);
                if (sing) det = (T)1;
                const B4<T> Pi{P.d / det, -P.b / det, -P.c / det, P.a / det};
                if (t == 0) {
                    if (sing) atomicOr(bad, 1);
                    store_b4(blk, DI + j, Pi);
                } else {
                    const B4<T> lb = mul(Lb, Pi);
                    store_b4(blk, L0 + p, lb);
                    store_b4(blk, LR0 + lrp, lb);
                }
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// Forward and backward sweeps, GL lanes per node.  The first entry of the
// first node a lane group takes on a level (index and factor block, which do
// not depend on x) is loaded one level ahead, so its latency overlaps the
// previous level's arithmetic and barrier; the summation order is the same
// as without it.
template <typename T>
__device__ __forceinline__ void wg_solve(const DevSym &Y, const T *blk, double *x, double *s_r) {
    const int grp = (int)threadIdx.x / GL, lane = (int)threadIdx.x % GL, ng = (int)blockDim.x / GL;
    const int LR0 = Y.n + 2 * Y.nnzL, U0 = Y.n + Y.nnzL;
    const int nl = Y.nlev_s;
    auto fetch_fwd = [&](int l, int &k, B4<T> &L) {
        k = -1;
        const int s = Y.lvl_ptr[l];
        if (grp < Y.lvl_ptr[l + 1] - s) {
            const int j = s + grp, q = Y.rs_ptr[j] + lane;
            if (q < Y.rs_ptr[j + 1]) {
                k = Y.rs_k[q];
                L = load_b4(blk, LR0 + q);
            }
        }
    };
    int pk = -1;
    B4<T> pL{0, 0, 0, 0};
    if (nl > 0) fetch_fwd(0, pk, pL);
    for (int l = 0; l < nl; l++) {
        int nk = -1;
        B4<T> nL{0, 0, 0, 0};
        if (l + 1 < nl) fetch_fwd(l + 1, nk, nL);
        const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
        for (int i = grp; i < c; i += ng) {
            const int j = (s + i);
            double a0 = 0, a1 = 0;
            int q = Y.rs_ptr[j] + lane;
            if (i == grp) {
                if (pk >= 0) {
                    a0 += (double)pL.a * x[2 * pk] + (double)pL.b * x[2 * pk + 1];
                    a1 += (double)pL.c * x[2 * pk] + (double)pL.d * x[2 * pk + 1];
                }
                q += GL;
            }
            for (; q < Y.rs_ptr[j + 1]; q += GL) {
                B4<T> L = load_b4(blk, LR0 + q);
                const int k = Y.rs_k[q];
                a0 += (double)L.a * x[2 * k] + (double)L.b * x[2 * k + 1];
                a1 += (double)L.c * x[2 * k] + (double)L.d * x[2 * k + 1];
            }
            a0 = xor_sum<GL>(a0);
            a1 = xor_sum<GL>(a1);
            if (lane == 0) {
                x[2 * j] -= a0;
                x[2 * j + 1] -= a1;
            }
        }
        __syncthreads();
        pk = nk;
        pL = nL;
    }
    auto fetch_bwd = [&](int l, int &r, B4<T> &U, B4<T> &D) {
        r = -1;
        const int s = Y.lvl_ptr[l];
        if (grp < Y.lvl_ptr[l + 1] - s) {
            const int j = s + grp, p = Y.colptr[j] + lane;
            if (lane == 0) D = load_b4(blk, Y.di0 + j);
            if (p < Y.colptr[j + 1]) {
                r = Y.rowind[p];
                U = load_b4(blk, U0 + p);
            }
        }
    };
    int pr = -1;
    B4<T> pU{0, 0, 0, 0}, pD{0, 0, 0, 0};
    if (nl > 0) fetch_bwd(nl - 1, pr, pU, pD);
    if (Y.ch_m > 0) chain_solve<T, GL>(Y, blk, x, s_r);
    for (int l = nl - 1; l >= 0; l--) {
        int nr = -1;
        B4<T> nU{0, 0, 0, 0}, nD{0, 0, 0, 0};
        if (l > 0) fetch_bwd(l - 1, nr, nU, nD);
        const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
        for (int i = grp; i < c; i += ng) {
            const int j = (s + i);
            double a0 = 0, a1 = 0;
            int p = Y.colptr[j] + lane;
            if (i == grp) {
                if (pr >= 0) {
                    a0 += (double)pU.a * x[2 * pr] + (double)pU.b * x[2 * pr + 1];
                    a1 += (double)pU.c * x[2 * pr] + (double)pU.d * x[2 * pr + 1];
                }
                p += GL;
            }
            for (; p < Y.colptr[j + 1]; p += GL) {
                B4<T> U = load_b4(blk, U0 + p);
                const int r = Y.rowind[p];
                a0 += (double)U.a * x[2 * r] + (double)U.b * x[2 * r + 1];
                a1 += (double)U.c * x[2 * r] + (double)U.d * x[2 * r + 1];
            }
            a0 = xor_sum<GL>(a0);
            a1 = xor_sum<GL>(a1);
            if (lane == 0) {
                const double y0 = x[2 * j] - a0, y1 = x[2 * j + 1] - a1;
                const B4<T> D = (i == grp) ? pD : load_b4(blk, Y.di0 + j);
                x[2 * j] = (double)D.a * y0 + (double)D.b * y1;
                x[2 * j + 1] = (double)D.c * y0 + (double)D.d * y1;
            }
        }
        __syncthreads();
        pr = nr;
        pU = nU;
        pD = nD;
    }
}

// Sum of the products [lo, hi) of an LDS array in index order (four
// loads in flight before their additions, which keep the order).
__device__ __forceinline__ double2 prod_sum_in_order(const double2 *v, int lo, int hi) {
    double a0 = 0, a1 = 0;
    int q = lo;
    for (; q + 4 <= hi; q += 4) {
        const double2 p0 = v[q], p1 = v[q + 1], p2 = v[q + 2], p3 = v[q + 3];
        a0 += p0.x; a1 += p0.y;
        a0 += p1.x; a1 += p1.y;
        a0 += p2.x; a1 += p2.y;
        a0 += p3.x; a1 += p3.y;
    }
    for (; q < hi; q++) {
        a0 += v[q].x;
        a1 += v[q].y;
    }
    return make_double2(a0, a1);
}

// Forward and backward sweeps whose factor entries are loaded a level
// ahead (they do not depend on x).  The entries of a level are contiguous
// (rs_ptr / colptr in node order, nodes numbered level by level).
//  * Narrow levels (at most kSweepGroupLevel nodes, a constant, so every
//    launch shape sums alike): 16 lanes per node, lane i taking entries
//    i, i + 16, ...; its first kSweepAhead entries come from registers
//    filled during the level before.  One barrier per level.
//  * Wide levels: one entry per lane, the product L_jk x_k (backward
//    U_ji x_i) into LDS (prod, `page` at a time), then one thread per node
//    subtracts its products in entry order (backward: and applies the
//    inverse pivot after the level's last page).  Two barriers per page.
constexpr int kSweepGroupLevel = 16, kSweepAhead = 2;

template <typename T>
__device__ __forceinline__ void wg_solve_flat(const DevSym &Y, const T *blk, double *x, double *s_r, double2 *prod,
                                              int page) {
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / GL, lane = tid % GL;
    const int LR0 = Y.n + 2 * Y.nnzL, U0 = Y.n + Y.nnzL, nl = Y.nlev_s;
    constexpr int RA = kSweepAhead;
    int pk[RA];
    B4<T> pB[RA], pD{0, 0, 0, 0};
    int ns = 0, nc = 0, ne0 = 0, ne1 = 0;   // the coming level's first node and width, this lane's node's entries
    // the entries this lane takes first on level l: a narrow level's node
    // grp (entries lane + GL r), a wide level's entry tid; ptr / idx / B0 are
    // the row (forward) or column (backward) structure and its blocks.  The
    // level's bounds and the node's entry range stay in registers for the
    // level itself, which then starts with no metadata read.
    auto ahead = [&](int l, const int *ptr, const int *idx, int B0, bool diag) __attribute__((always_inline)) {
        const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
        ns = s;
        nc = c;
#pragma unroll
        for (int r = 0; r < RA; r++) pk[r] = -1;
        if (c <= kSweepGroupLevel) {
            if (grp < c) {
                const int j = s + grp, e1 = ptr[j + 1];
                ne0 = ptr[j];
                ne1 = e1;
#pragma unroll
                for (int r = 0; r < RA; r++) {
                    const int e = ptr[j] + lane + GL * r;
                    if (e < e1) {
                        pk[r] = idx[e];
                        pB[r] = load_b4(blk, B0 + e);
                    }
                }
                if (diag && lane == 0) pD = load_b4(blk, Y.di0 + j);
            }
        } else {
            const int e = ptr[s] + tid;
            if (e < ptr[s + c]) {
                pk[0] = idx[e];
                pB[0] = load_b4(blk, B0 + e);
            }
            if (diag && tid < c) pD = load_b4(blk, Y.di0 + s + tid);
        }
    };
    // ---- forward
    if (nl > 0) ahead(0, Y.rs_ptr, Y.rs_k, LR0, false);
    for (int l = 0; l < nl; l++) {
        const int s = ns, c = nc, je0 = ne0, je1 = ne1;
        int ck[RA];
        B4<T> cB[RA];
#pragma unroll
        for (int r = 0; r < RA; r++) {
            ck[r] = pk[r];
            cB[r] = pB[r];
        }
        if (c <= kSweepGroupLevel) {
            double xk[RA][2];   // x of the first entries, read before the next level's requests
#pragma unroll
            for (int r = 0; r < RA; r++) {
                const int k = ck[r] >= 0 ? ck[r] : 0;
                xk[r][0] = x[2 * k];
                xk[r][1] = x[2 * k + 1];
            }
            // the node's own x: no earlier level of this sweep writes it
            const int jo = s + (grp < c ? grp : 0);
            const double xj0 = x[2 * jo], xj1 = x[2 * jo + 1];
            if (l + 1 < nl) ahead(l + 1, Y.rs_ptr, Y.rs_k, LR0, false);
            if (grp < c) {
                const int j = s + grp;
                double a0 = 0, a1 = 0;
#pragma unroll
                for (int r = 0; r < RA; r++)
                    if (ck[r] >= 0) {
                        const double x0 = xk[r][0], x1 = xk[r][1];
                        a0 += (double)cB[r].a * x0 + (double)cB[r].b * x1;
                        a1 += (double)cB[r].c * x0 + (double)cB[r].d * x1;
                    }
                for (int e = je0 + lane + GL * RA; e < je1; e += GL) {
                    const B4<T> L = load_b4(blk, LR0 + e);
                    const int k = Y.rs_k[e];
                    const double x0 = x[2 * k], x1 = x[2 * k + 1];
                    a0 += (double)L.a * x0 + (double)L.b * x1;
                    a1 += (double)L.c * x0 + (double)L.d * x1;
                }
                a0 = xor_sum<GL>(a0);
                a1 = xor_sum<GL>(a1);
                if (lane == 0) {
                    x[2 * j] = xj0 - a0;
                    x[2 * j + 1] = xj1 - a1;
                }
            }
            __syncthreads();
            continue;
        }
        const int e_lo = Y.rs_ptr[s], e_hi = Y.rs_ptr[s + c];
        if (e_lo == e_hi && l + 1 < nl) ahead(l + 1, Y.rs_ptr, Y.rs_k, LR0, false);
        for (int pg = e_lo; pg < e_hi; pg += page) {
            const int pe = min(pg + page, e_hi);
            for (int e = pg + tid; e < pe; e += bd) {
                int k;
                B4<T> L;
                if (e == e_lo + tid) {
                    k = ck[0];
                    L = cB[0];
                } else {
                    k = Y.rs_k[e];
                    L = load_b4(blk, LR0 + e);
                }
                const double x0 = x[2 * k], x1 = x[2 * k + 1];
                prod[e - pg] = make_double2((double)L.a * x0 + (double)L.b * x1, (double)L.c * x0 + (double)L.d * x1);
            }
            if (pg == e_lo && l + 1 < nl) ahead(l + 1, Y.rs_ptr, Y.rs_k, LR0, false);
            __syncthreads();
            for (int i = tid; i < c; i += bd) {
                const int j = s + i, lo = max(Y.rs_ptr[j], pg), hi = min(Y.rs_ptr[j + 1], pe);
                if (lo < hi) {
                    const double2 a = prod_sum_in_order(prod, lo - pg, hi - pg);
                    x[2 * j] -= a.x;
                    x[2 * j + 1] -= a.y;
                }
            }
            __syncthreads();
        }
    }
    if (Y.ch_m > 0) chain_solve<T, GL>(Y, blk, x, s_r);
    // ---- backward
    if (nl > 0) ahead(nl - 1, Y.colptr, Y.rowind, U0, true);
    for (int l = nl - 1; l >= 0; l--) {
        const int s = ns, c = nc, je0 = ne0, je1 = ne1;
        int ck[RA];
        B4<T> cB[RA];
#pragma unroll
        for (int r = 0; r < RA; r++) {
            ck[r] = pk[r];
            cB[r] = pB[r];
        }
        const B4<T> myD = pD;   // narrow: the inverse pivot of node s + grp (lane 0); wide: of node s + tid
        if (c <= kSweepGroupLevel) {
            double xk[RA][2];   // x of the first entries, read before the next level's requests
#pragma unroll
            for (int r = 0; r < RA; r++) {
                const int k = ck[r] >= 0 ? ck[r] : 0;
                xk[r][0] = x[2 * k];
                xk[r][1] = x[2 * k + 1];
            }
            // the node's own x: no earlier level of this sweep writes it
            const int jo = s + (grp < c ? grp : 0);
            const double xj0 = x[2 * jo], xj1 = x[2 * jo + 1];
            if (l > 0) ahead(l - 1, Y.colptr, Y.rowind, U0, true);
            if (grp < c) {
                const int j = s + grp;
                double a0 = 0, a1 = 0;
#pragma unroll
                for (int r = 0; r < RA; r++)
                    if (ck[r] >= 0) {
                        const double x0 = xk[r][0], x1 = xk[r][1];
                        a0 += (double)cB[r].a * x0 + (double)cB[r].b * x1;
                        a1 += (double)cB[r].c * x0 + (double)cB[r].d * x1;
                    }
                for (int e = je0 + lane + GL * RA; e < je1; e += GL) {
                    const B4<T> U = load_b4(blk, U0 + e);
                    const int i = Y.rowind[e];
                    const double x0 = x[2 * i], x1 = x[2 * i + 1];
                    a0 += (double)U.a * x0 + (double)U.b * x1;
                    a1 += (double)U.c * x0 + (double)U.d * x1;
                }
                a0 = xor_sum<GL>(a0);
                a1 = xor_sum<GL>(a1);
                if (lane == 0) {
                    const double y0 = xj0 - a0, y1 = xj1 - a1;
                    x[2 * j] = (double)myD.a * y0 + (double)myD.b * y1;
                    x[2 * j + 1] = (double)myD.c * y0 + (double)myD.d * y1;
                }
            }
            __syncthreads();
            continue;
        }
        const int e_lo = Y.colptr[s], e_hi = Y.colptr[s + c];
        if (e_lo == e_hi) {     // no entries: only the inverse pivots
            if (l > 0) ahead(l - 1, Y.colptr, Y.rowind, U0, true);
            for (int i = tid; i < c; i += bd) {
                const int j = s + i;
                const B4<T> D = (i == tid) ? myD : load_b4(blk, Y.di0 + j);
                const double y0 = x[2 * j], y1 = x[2 * j + 1];
                x[2 * j] = (double)D.a * y0 + (double)D.b * y1;
                x[2 * j + 1] = (double)D.c * y0 + (double)D.d * y1;
            }
            __syncthreads();
            continue;
        }
        for (int pg = e_lo; pg < e_hi; pg += page) {
            const int pe = min(pg + page, e_hi);
            for (int e = pg + tid; e < pe; e += bd) {
                int r;
                B4<T> U;
                if (e == e_lo + tid) {
                    r = ck[0];
                    U = cB[0];
                } else {
                    r = Y.rowind[e];
                    U = load_b4(blk, U0 + e);
                }
                const double x0 = x[2 * r], x1 = x[2 * r + 1];
                prod[e - pg] = make_double2((double)U.a * x0 + (double)U.b * x1, (double)U.c * x0 + (double)U.d * x1);
            }
            if (pg == e_lo && l > 0) ahead(l - 1, Y.colptr, Y.rowind, U0, true);
            __syncthreads();
            const bool last = pe == e_hi;
            for (int i = tid; i < c; i += bd) {
                const int j = s + i, lo = max(Y.colptr[j], pg), hi = min(Y.colptr[j + 1], pe);
                double y0 = x[2 * j], y1 = x[2 * j + 1];
                if (lo < hi) {
                    const double2 a = prod_sum_in_order(prod, lo - pg, hi - pg);
                    y0 -= a.x;
                    y1 -= a.y;
                }
                if (last) {
                    const B4<T> D = (i == tid) ? myD : load_b4(blk, Y.di0 + j);
                    x[2 * j] = (double)D.a * y0 + (double)D.b * y1;
                    x[2 * j + 1] = (double)D.c * y0 + (double)D.d * y1;
                } else if (lo < hi) {
                    x[2 * j] = y0;
                    x[2 * j + 1] = y1;
                }
            }
            __syncthreads();
        }
    }
}

// I = Y E and S = E conj(I): 4 lanes per row
template <typename T>
__device__ __forceinline__ void wg_pq(const DevNet &N, Work<T> &w) {
    // two rows a lane group at a time (r and r + rows_per), their entries
    // walked together so that their loads are in flight together; each row
    // sums in the same order as alone
    const int rows_per = (int)blockDim.x / 4;
    const int lane = (int)threadIdx.x & 3;
    for (int r0 = 0; r0 < N.nB; r0 += 2 * rows_per) {
        int row[2], e[2], e1[2];
        double ar[2] = {0, 0}, ai[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            row[u] = r0 + u * rows_per + (int)threadIdx.x / 4;
            const bool ok = row[u] < N.nB;
            e[u] = ok ? N.rowptr[row[u]] + lane : 0;
            e1[u] = ok ? N.rowptr[row[u] + 1] : 0;
        }
        while (e[0] < e1[0] || e[1] < e1[1]) {
#pragma unroll
            for (int u = 0; u < 2; u++)
                if (e[u] < e1[u]) {
                    const double2 Em = w.E[N.colind[e[u]]];
                    const double2 y = N.y[e[u]];
                    ar[u] += y.x * Em.x - y.y * Em.y;
                    ai[u] += y.x * Em.y + y.y * Em.x;
                    e[u] += 4;
                }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            ar[u] = quad_sum(ar[u]);
            ai[u] = quad_sum(ai[u]);
            if (row[u] < N.nB && lane == 0) {
                const double2 Ek = w.E[row[u]];
                w.I[row[u]] = make_double2(ar[u], ai[u]);
                w.P[row[u]] = Ek.x * ar[u] + Ek.y * ai[u];
                w.Q[row[u]] = Ek.y * ar[u] - Ek.x * ai[u];
            }
        }
    }
}

// r = g - J x (FP64, complex voltages)
template <typename T>
__device__ __forceinline__ void wg_residual(const DevNet &N, Work<T> &w) {
    // two rows a lane group at a time, as in wg_pq
    const int rows_per = (int)blockDim.x / 4;
    const int lane = (int)threadIdx.x & 3;
    for (int r0 = 0; r0 < N.nB; r0 += 2 * rows_per) {
        int row[2], e[2], e1[2];
        bool active[2];
        double ar[2] = {0, 0}, ai[2] = {0, 0};
#pragma unroll
        for (int u = 0; u < 2; u++) {
            row[u] = r0 + u * rows_per + (int)threadIdx.x / 4;
            active[u] = row[u] < N.nB && row[u] != N.slack;
            e[u] = active[u] ? N.rowptr[row[u]] + lane : 0;
            e1[u] = active[u] ? N.rowptr[row[u] + 1] : 0;
        }
        while (e[0] < e1[0] || e[1] < e1[1]) {
#pragma unroll
            for (int u = 0; u < 2; u++)
                if (e[u] < e1[u]) {
                    const int m = N.colind[e[u]];
                    if (m != N.slack) {
                        const int jm = N.node_of_bus[m];
                        const double dth = w.x[2 * jm];
                        const double dv = w.isq[m] ? w.x[2 * jm + 1] : 0.0;
                        const double2 Em = w.E[m];
                        const double f = dv / w.V[m];
                        const double dr = Em.x * f - Em.y * dth, di = Em.y * f + Em.x * dth;
                        const double2 y = N.y[e[u]];
                        ar[u] += y.x * dr - y.y * di;
                        ai[u] += y.x * di + y.y * dr;
                    }
                    e[u] += 4;
                }
        }
#pragma unroll
        for (int u = 0; u < 2; u++) {
            ar[u] = quad_sum(ar[u]);
            ai[u] = quad_sum(ai[u]);
            if (active[u] && lane == 0) {
                const int k = row[u], j = N.node_of_bus[k];
                const double dth = w.x[2 * j];
                const bool q = w.isq[k];
                const double dv = q ? w.x[2 * j + 1] : 0.0;
                const double2 Ek = w.E[k], Ik = w.I[k];
                const double f = dv / w.V[k];
                const double dr = Ek.x * f - Ek.y * dth, di = Ek.y * f + Ek.x * dth;
                const double sr = dr * Ik.x + di * Ik.y + Ek.x * ar[u] + Ek.y * ai[u];
                const double si = di * Ik.x - dr * Ik.y + Ek.y * ar[u] - Ek.x * ai[u];
                w.r[2 * j] = w.g[2 * j] - sr;
                w.r[2 * j + 1] = q ? w.g[2 * j + 1] - si : -w.x[2 * j + 1];
            }
        }
    }
}

// limits pass: mode 0 chkLimQ (or undo when !chk), flips accumulated in *flips
template <typename T>
__device__ __forceinline__ void wg_limits(const DevNet &N, Work<T> &w, const double *Qliq, const double *Qload,
                          bool chk, int *flips) {
    int f = 0;
    for (int i = threadIdx.x; i < N.nPV; i += blockDim.x) {
        const int b = N.pv_bus[i];
        const double qload = Qload[b], qmin = N.qmin[i], qmax = N.qmax[i];
        unsigned char lq = w.limQ[i];
        double qll = w.ql[b];
        if (!lq) {
            if (chk) {
                const double qg = w.Q[b] + qload;
                if (qg < qmin - kFolga) { qll = qmin - qload; lq = 1; f++; }
                else if (qg > qmax + kFolga) { qll = qmax - qload; lq = 1; f++; }
            }
        } else {
            if (qll == qmax - qload) {
                if (w.V[b] > N.vset[i] + kFolga) { lq = 0; qll = Qliq[b]; f++; }
            } else {
                if (w.V[b] < N.vset[i] - kFolga) { lq = 0; qll = Qliq[b]; f++; }
            }
        }
        w.limQ[i] = lq;
        w.ql[b] = qll;
        w.isq[b] = lq;
    }
    if (f) atomicAdd(flips, f);
}

template <typename T>
__device__ __forceinline__ void wg_set_E(const DevNet &N, Work<T> &w) {
    for (int k = threadIdx.x; k < N.nB; k += blockDim.x) {
        double s, c;
        sincos(w.th[k], &s, &c);
        w.E[k] = make_double2(w.V[k] * c, w.V[k] * s);
    }
}

// One NR solve of one scenario by the whole workgroup.  Returns (via shared
// result) iterations, converged flag and final max mismatch.
template <typename T>
__device__ __forceinline__ void wg_nr(const BatchTopo &tp, const BatchJob &J, Work<T> &w, const double *Pliq,
                      const double *Qliq, const double *Qload, const double *V0, const double *th0,
                      int *s_it, int *s_conv, double *s_maxg, int *s_fl, B4<T> *s_chain, B4<T> *s_ck) {
    const DevNet &N = tp.N;
    const DevSym &Y = tp.Y;
    __shared__ unsigned long long s_maxG_bits, s_maxDx_bits;
    __shared__ int s_bad, s_flips, s_sing;
    const int tid = threadIdx.x;
    // a warm start sets |V| of the PQ buses and the angles of the non-slack
    // buses; PV buses start at their setpoint and the slack at its data
    for (int k = tid; k < N.nB; k += blockDim.x) {
        const bool pq = N.pv_of_bus[k] < 0 && k != N.slack;
        w.V[k] = pq ? V0[k] : tp.V0flat[k];
        w.th[k] = (k == N.slack) ? tp.th0flat[k] : th0[k];
        w.ql[k] = Qliq[k];
        w.isq[k] = pq ? 1 : 0;                                   // initLim: PQ buses only
    }
    for (int i = tid; i < N.nPV; i += blockDim.x) w.limQ[i] = 0;
    if (tid == 0) {
        s_maxDx_bits = 0;
        s_flips = 0;
        s_sing = 0;
    }
    __syncthreads();
    wg_set_E(N, w);
    __syncthreads();
    int it = 0, conv = 0;
    double maxG = 0;
    while (it < J.max_iter) {
        wg_pq(N, w);
        if (tid == 0) { s_maxG_bits = 0; s_bad = 0; }
        __syncthreads();
        if (J.lim && it >= 1) {
            const bool chk = u64_as_double(s_maxDx_bits) < 0.1;
            int any = 0;
            if (!chk)
                for (int i = tid; i < N.nPV; i += blockDim.x) any |= w.limQ[i];
            any = __syncthreads_or(any);
            if (chk || any) wg_limits(N, w, Qliq, Qload, chk, &s_flips);
            __syncthreads();
        }
        // mismatch
        {
            double m = 0.0;
            int bad = 0;
            for (int k = tid; k < N.nB; k += blockDim.x) {
                if (k == N.slack) continue;
                const int j = N.node_of_bus[k];
                const double gp = Pliq[k] - w.P[k];
                const double gq = w.isq[k] ? w.ql[k] - w.Q[k] : 0.0;
                w.g[2 * j] = gp;
                w.g[2 * j + 1] = gq;
                m = dmax_nan(m, dmax_nan(fabs(gp), fabs(gq)));
                const int i = N.pv_of_bus[k];
                if (i >= 0) {
                    const double qg = w.Q[k] + Qload[k];
                    bad |= (qg < N.qmin[i] - kFolga) || (qg > N.qmax[i] + kFolga);
                }
            }
            m = wave_max(m);
            bad = __any(bad);
            if ((tid & 63) == 0) {
                atomicMax(&s_maxG_bits, double_as_u64(m));
                if (bad) atomicOr(&s_bad, 1);
            }
        }
        __syncthreads();
        maxG = u64_as_double(s_maxG_bits);
        const int simple_bad = s_bad;
        __syncthreads();
        if (maxG <= J.tol) {
            if (J.lim && simple_bad) {                    // switch taken at convergence
                wg_limits(N, w, Qliq, Qload, true, &s_flips);
                if (tid == 0) s_maxDx_bits = 0;
                __syncthreads();
                it++;
                continue;
            }
            conv = 1;
            break;
        }
        if (!(maxG < J.diverge)) break;                   // NaN or diverging: give up this start
        // Jacobian blocks, four slots a thread at a time so that their index
        // loads are in flight together
        constexpr int JF = 4;
        for (int s0 = tid; s0 < tp.nslot; s0 += JF * blockDim.x) {
            int es[JF];
#pragma unroll
            for (int r = 0; r < JF; r++) {
                const int s = s0 + r * (int)blockDim.x;
                es[r] = s < tp.nslot ? Y.slot_e[s] : -1;
            }
            int ks[JF], ms[JF];
            double2 ys[JF];
#pragma unroll
            for (int r = 0; r < JF; r++) {
                const int e = es[r] >= 0 ? es[r] : 0;
                ks[r] = N.coorow[e];
                ms[r] = N.colind[e];
                ys[r] = N.y[e];
            }
#pragma unroll
            for (int r = 0; r < JF; r++) {
            const int s = s0 + r * (int)blockDim.x, e = es[r];
            if (s >= tp.nslot) break;
            B4<T> out{0, 0, 0, 0};
            if (e >= 0) {
                const int k = ks[r], m = ms[r];
                const double2 y = ys[r];
                double H, Nn, M, L;
                if (k != m) {
                    const double2 Em = w.E[m], Ek = w.E[k];
                    const double tr = y.x * Em.x - y.y * Em.y, ti = y.x * Em.y + y.y * Em.x;
                    const double cr = Ek.x * tr + Ek.y * ti, ci = Ek.y * tr - Ek.x * ti;
                    const double vm = w.V[m];
                    H = ci; Nn = cr / vm; M = -cr; L = ci / vm;
                } else {
                    const double v = w.V[k], v2 = v * v, p = w.P[k], q = w.Q[k];
                    H = -q - v2 * y.y; Nn = (p + v2 * y.x) / v; M = p - v2 * y.x; L = (q - v2 * y.y) / v;
                }
                if (!w.isq[m]) { Nn = 0; L = 0; }
                if (!w.isq[k]) { M = 0; L = (k == m) ? 1.0 : 0.0; }
                out = B4<T>{(T)H, (T)Nn, (T)M, (T)L};
            }
            store_b4(w.blk, s, out);
            }
        }
        if (tid == 0) s_maxDx_bits = 0;
        __syncthreads();
        // factor, level by level
        // wide levels: a lane group per node; narrow levels: the whole
        // workgroup on the level's destinations (coop_factor_range)
        {
            const int grp = tid / GL, lane = tid % GL, ng = (int)blockDim.x / GL;
            for (int l = 0; l < Y.nlev_s;) {
                if (s_ck && tp.F.fl_ptr && l == tp.F.lv0 && ((size_t)w.x & 15) == 0) {   // flat pair lists
                    T *pa = reinterpret_cast<T *>(w.x);   // x and r: free until the solve
                    flat_factor_levels<T, 1>(Y, tp.F, w.blk, 0, Y.nlev_s - tp.F.lv0, &s_sing, pa,
                                                pa + 4 * (size_t)tp.F.max_pairs);
                    break;
                }
                if (s_ck && Y.ck_lvl && l >= Y.tail_start) {   // the chunked narrow top
                    chunk_factor_levels<T, kBatchChunkLanes, 1>(Y, w.blk, l, Y.nlev_s, &s_sing, s_ck,
                                                                  s_ck + kBatchMaxChunks, s_chain);
                    break;
                }
                const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
                if (c >= kWideLevel) {
                    for (int i = grp; i < c; i += ng) wg_factor_node(Y, w.blk, (s + i), lane, &s_sing);
                    __syncthreads();
                    l++;
                } else {   // a run of narrow levels
                    int l1 = l + 1;
                    while (l1 < Y.nlev_s && Y.lvl_ptr[l1 + 1] - Y.lvl_ptr[l1] < kWideLevel) l1++;
                    wg_factor_narrow<T, kCoopLanes>(Y, w.blk, l, l1, &s_sing, s_chain);
                    l = l1;
                }
            }
            if (Y.ch_m > 0) {
                chain_factor_mem<T, GL, kBatchUnroll, kBatchTile<T>>(Y, w.blk, &s_sing, s_chain);
                __syncthreads();
            }
        }
        if (s_sing) break;
        for (int i = tid; i < 2 * Y.n; i += blockDim.x) w.x[i] = w.g[i];
        __syncthreads();
        // the flat sweeps take their products in s_ck (free after the factorisation)
        const int page = (int)(sizeof(B4<T>) * 2 * kBatchMaxChunks / sizeof(double2));
        if (s_ck) wg_solve_flat(Y, (const T *)w.blk, w.x, (double *)s_chain, reinterpret_cast<double2 *>(s_ck), page);
        else wg_solve(Y, (const T *)w.blk, w.x, (double *)s_chain);
        for (int rs = 0; rs < J.refine; rs++) {
            wg_residual(N, w);
            __syncthreads();
            if (s_ck) wg_solve_flat(Y, (const T *)w.blk, w.r, (double *)s_chain, reinterpret_cast<double2 *>(s_ck), page);
            else wg_solve(Y, (const T *)w.blk, w.r, (double *)s_chain);
            for (int i = tid; i < 2 * Y.n; i += blockDim.x) w.x[i] += w.r[i];
            __syncthreads();
        }
        // update
        {
            double m = 0.0;
            for (int k = tid; k < N.nB; k += blockDim.x) {
                if (k == N.slack) continue;
                const int j = N.node_of_bus[k];
                const double dth = w.x[2 * j];
                double v = w.V[k];
                const double th = w.th[k] + dth;
                m = dmax_nan(m, fabs(dth));
                if (w.isq[k]) {
                    const double dv = w.x[2 * j + 1];
                    v += dv;
                    m = dmax_nan(m, fabs(dv));
                }
                w.V[k] = v;
                w.th[k] = th;
                double sn, cs;
                sincos(th, &sn, &cs);
                w.E[k] = make_double2(v * cs, v * sn);
            }
            m = wave_max(m);
            if ((tid & 63) == 0) atomicMax(&s_maxDx_bits, double_as_u64(m));
        }
        __syncthreads();
        it++;
    }
    if (tid == 0) {
        *s_it = it;
        *s_conv = conv;
        *s_maxg = maxG;
        *s_fl = s_flips;
    }
    __syncthreads();
}

// bytes of LDS that MODE 1 uses: the step vectors g, x, r and the level /
// row-structure metadata the triangular solves and the factor walk
template <typename T>
__host__ __device__ inline long long partial_lds_bytes(int n, int nlev) {
    long long b = 16LL * n * 3;                              // g x r
    b += 4LL * ((nlev + 1) * 2 + 3 * (n + 1));               // lvl_ptr ld_ptr rs_ptr colptr fd_ptr
    return (b + 255) & ~255LL;
}

// MODE 2: the whole per-scenario working set in LDS.
// MODE 1: working set in a global slice, except the step vectors and the
//         level / row-structure metadata, staged in LDS (latency of the
//         level-by-level walk); MODE 0: everything in the global slice.
// The float LDS-resident kernel at 256 threads keeps three waves per SIMD
// (three workgroups per CU, within 168 VGPRs): small networks are latency
// bound, so more resident scenarios per CU beat the few registers it costs.
template <typename T, int MODE, int MAXT>
constexpr int batch_min_waves() { return (MODE == 2 && MAXT == 256 && sizeof(T) == 4) ? 3 : 1; }

template <typename T, int MODE, int MAXT>   // MAXT: largest workgroup (register budget)
__global__ void __launch_bounds__(MAXT, (batch_min_waves<T, MODE, MAXT>())) k_batch(BatchTopo tp0, BatchJob J) {
    extern __shared__ __align__(16) char smem[];
    __shared__ int s_item, s_it, s_conv, s_fl;
    __shared__ B4<T> s_chain[4 * batch_chain_cap<MODE>() + 2];   // dense top: row / column / pivot staging
    __shared__ B4<T> s_ck[MODE == 2 ? 1 : 2 * kBatchMaxChunks];  // chunk partial sums (a, u)
    __shared__ double s_maxg;
    BatchTopo tp = tp0;
    const DevNet &N = tp.N;
    char *base = (MODE == 2) ? smem : J.gws + (long long)blockIdx.x * J.ws_bytes;
    Work<T> w = carve<T>(base, N.nB, tp.Y.n, N.nPV, tp.nblk);
    if (MODE == 1) {
        const int n = tp.Y.n, nlev = tp.Y.nlev;
        double *v = reinterpret_cast<double *>(smem);
        w.g = v;
        w.x = v + 2 * n;
        w.r = v + 4 * n;
        int *ip = reinterpret_cast<int *>(v + 6 * n);
        auto stage = [&](const int *src, int cnt) {
            int *dst = ip;
            for (int i = threadIdx.x; i < cnt; i += blockDim.x) dst[i] = src[i];
            ip += cnt;
            return (const int *)dst;
        };
        tp.Y.lvl_ptr = stage(tp0.Y.lvl_ptr, nlev + 1);
        tp.Y.ld_ptr = stage(tp0.Y.ld_ptr, nlev + 1);
        tp.Y.rs_ptr = stage(tp0.Y.rs_ptr, n + 1);
        tp.Y.colptr = stage(tp0.Y.colptr, n + 1);
        tp.Y.fd_ptr = stage(tp0.Y.fd_ptr, n + 1);
        __syncthreads();
    }
    const int nscen = J.nscen_dev ? *J.nscen_dev : J.nscen;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(J.counter, 1);
        __syncthreads();
        const int item = s_item;
        __syncthreads();
        if (item >= nscen) break;                           // every wave reaches this exit
        const int sc = J.scen[item];
        const long long off = J.off[sc];
        const double *Pl = J.Pliq + off, *Ql = J.Qliq + off, *Qd = J.Qload + off;
        const bool warm = J.V0 != nullptr;
        int total_it = 0;
        // one call site (the solve is inlined once): the warm start, then,
        // when it does not converge, the flat-start fallback
        for (int attempt = warm ? 0 : 1; attempt < 2; ++attempt) {
            const bool from_warm = attempt == 0;
            wg_nr(tp, J, w, Pl, Ql, Qd, from_warm ? J.V0 + off : tp.V0flat,
                  from_warm ? J.th0 + off : tp.th0flat, &s_it, &s_conv, &s_maxg, &s_fl, s_chain,
                  MODE == 2 ? nullptr : s_ck);
            total_it += s_it;
            if (s_conv) break;                              // shared: uniform over the workgroup
        }
        for (int k = threadIdx.x; k < N.nB; k += blockDim.x) {
            J.Vout[off + k] = w.V[k];
            J.thout[off + k] = w.th[k];
        }
        const long long pvo = J.pv_off[sc];
        for (int i = threadIdx.x; i < N.nPV; i += blockDim.x) J.limQout[pvo + i] = w.limQ[i];
        if (threadIdx.x == 0) {
            J.iters[sc] = total_it;
            J.conv[sc] = s_conv;
            J.maxmis[sc] = s_maxg;
            J.flips[sc] = s_fl;
        }
        __syncthreads();
    }
}


// ------------------------------------------------------------ fixed scratch
// The allocation-free batch entry point keeps its per-call tables (packed
// offsets, per-network scenario lists, work counters) and the per-workgroup
// working-set slices in this fixed array, one copy per GPU, held until the
// process ends.  Host code must use the address hipGetSymbolAddress returns
// for the current device (noalloc_scratch), never the array's name.
constexpr long long kScratchBytes = 64LL << 20;
constexpr int kScratchHeader = 1024;        // counts [64], counters [64], invalid flag
static __device__ __align__(256) char scratch[kScratchBytes];

struct NoallocNets {
    int n_nets;
    int nB[64], nPV[64];
};

// exclusive prefix of v over the workgroup (1024 threads at most); *total:
// the sum over all threads
__device__ long long block_scan_excl(long long v, long long *total) {
    __shared__ long long buf[1024];
    const int t = threadIdx.x;
    buf[t] = v;
    __syncthreads();
    for (int d = 1; d < (int)blockDim.x; d <<= 1) {
        const long long a = t >= d ? buf[t - d] : 0;
        __syncthreads();
        buf[t] += a;
        __syncthreads();
    }
    const long long incl = buf[t];
    *total = buf[blockDim.x - 1];
    __syncthreads();
    return incl - v;
}

// One workgroup: the packed bus / PV offsets of every scenario, the list of
// scenarios of network i (in order) at scen + i * n_scen with its length in
// count[i], the work counters cleared, and a flag for net_of entries outside
// [0, n_nets) (those scenarios are left unsolved).
__global__ void __launch_bounds__(1024) k_noalloc_tables(const int *net_of, int n_scen, NoallocNets nn,
                                                         long long *off, long long *pvoff, int *scen,
                                                         int *count, int *counter, int *invalid) {
    const int T = blockDim.x, t = threadIdx.x;
    const int per = (n_scen + T - 1) / T;
    const int s0 = min(n_scen, t * per), s1 = min(n_scen, s0 + per);
    long long a = 0, b = 0;
    int bad = 0;
    for (int s = s0; s < s1; s++) {
        const int k = net_of[s];
        if (k < 0 || k >= nn.n_nets) { bad = 1; continue; }
        a += nn.nB[k];
        b += nn.nPV[k];
    }
    long long ta, tb;
    long long ea = block_scan_excl(a, &ta), eb = block_scan_excl(b, &tb);
    for (int s = s0; s < s1; s++) {
        const int k = net_of[s];
        off[s] = ea;
        pvoff[s] = eb;
        if (k < 0 || k >= nn.n_nets) continue;
        ea += nn.nB[k];
        eb += nn.nPV[k];
    }
    bad = __syncthreads_or(bad);
    for (int i = 0; i < nn.n_nets; i++) {
        long long c = 0, tot;
        for (int s = s0; s < s1; s++) c += net_of[s] == i;
        long long e = block_scan_excl(c, &tot);
        for (int s = s0; s < s1; s++)
            if (net_of[s] == i) scen[(long long)i * n_scen + e++] = s;
        if (t == 0) {
            count[i] = (int)tot;
            counter[i] = 0;
        }
    }
    if (t == 0) *invalid = bad;
}

__global__ void k_zero_int(int *p, int n) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

// The instantiation for a precision, working-set placement and workgroup
// size: compiled for at most 256, 512 or 1024 threads, so that smaller
// workgroups get the larger register budget.
using BatchKernel = void (*)(BatchTopo, BatchJob);
template <typename T, int MODE>
BatchKernel batch_kernel_t(int block) {
    if (block <= 256) return k_batch<T, MODE, 256>;
    if (block <= 512) return k_batch<T, MODE, 512>;
    return k_batch<T, MODE, 1024>;
}
inline BatchKernel batch_kernel(bool mixed, int mode, int block) {
    if (mixed)
        return mode == 2 ? batch_kernel_t<float, 2>(block)
                         : mode == 1 ? batch_kernel_t<float, 1>(block) : batch_kernel_t<float, 0>(block);
    return mode == 2 ? batch_kernel_t<double, 2>(block)
                     : mode == 1 ? batch_kernel_t<double, 1>(block) : batch_kernel_t<double, 0>(block);
}

// Starting state of scenario s (one workgroup each) from row s of a device
// array of network outputs, de-normalised as the inference's outputs are.
__global__ void k_start_from_rows(const float *Y, long long ld, int v_col, int th_col, const float *mu,
                                  const float *sd, const long long *off, int n_scen, long long total, double *V0,
                                  double *T0) {
    const int s = (int)blockIdx.x;
    const long long o = off[s], nb = (s + 1 < n_scen ? off[s + 1] : total) - o;
    const float *row = Y + (long long)s * ld;
    for (long long i = threadIdx.x; i < nb; i += blockDim.x) {
        V0[o + i] = (double)(row[v_col + i] * sd[v_col + i] + mu[v_col + i]);
        T0[o + i] = (double)(row[th_col + i] * sd[th_col + i] + mu[th_col + i]);
    }
}

}  // namespace pf

// ======================================================================== host

namespace {
#define BHIP(expr)                                                                 \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            pf::set_last_error(std::string(#expr) + ": " + hipGetErrorString(e_));     \
            return PF_DEVICE_ERROR;                                                \
        }                                                                          \
    } while (0)
#define BCHK(expr)                  \
    do {                            \
        pf_status s_ = (expr);      \
        if (s_ != PF_OK) return s_; \
    } while (0)
constexpr long long kLdsLimit = 64 * 1024;
}  // namespace

namespace {
std::mutex g_scratch_mu;
char *g_scratch_ptr[64] = {};   // address of `scratch` on each device

// once per device, with that device current: the device address of the
// fixed scratch array
pf_status noalloc_setup(int device) {
    if (device < 0 || device >= 64) return PF_INVALID_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PF_OK;
    void *p = nullptr;
    BHIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = (char *)p;
    return PF_OK;
}
}  // namespace

struct BatchNet {
    pf_network net;
    YPattern yp;
    Symbolic sym;
    BatchTopo tp{};
    bool lds = false;          // the whole working set fits in LDS (MODE 2)
    long long partial = 0;     // MODE 1 LDS bytes, 0 if it does not fit
    int mode = 0;              // k_batch MODE
    long long shm = 0;         // dynamic LDS bytes
    int block = 256;
    int grid = 0;
    int per_cu = 1;            // workgroups per CU the grid is sized for
    long long ws = 0;
    std::vector<int> scen;
    // device
    int *d_scen = nullptr;
    char *d_ws = nullptr;
    size_t ws_cap = 0;         // bytes allocated at d_ws
};

struct pf_batch {
    int device = 0;
    pf_options opt;
    std::vector<std::unique_ptr<BatchNet>> nets;
    std::vector<void *> allocs;
    hipStream_t st[2] = {nullptr, nullptr};
    hipEvent_t ev_fork = nullptr, ev_mid = nullptr, ev_join = nullptr, t0 = nullptr, t1 = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    int n_scen = 0;
    long long total = 0;
    std::vector<long long> off, pv_off;
    std::vector<int> net_of;
    long long total_pv = 0;
    double *d_P = nullptr, *d_Q = nullptr, *d_L = nullptr, *d_V0 = nullptr, *d_T0 = nullptr;
    double *warm_V0 = nullptr, *warm_T0 = nullptr;
    double *d_V = nullptr, *d_T = nullptr, *d_mm = nullptr;
    int *d_it = nullptr, *d_cv = nullptr, *d_counter = nullptr, *d_fl = nullptr;
    long long *d_off = nullptr, *d_pvoff = nullptr;
    unsigned char *d_lq = nullptr;
    int cus = 0;

    template <typename T>
    pf_status alloc(T **p, size_t n) {
        void *q = nullptr;
        BHIP(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)));
        allocs.push_back(q);
        *p = (T *)q;
        return PF_OK;
    }
    template <typename T>
    pf_status upload(T **p, const std::vector<T> &v) {
        BCHK(alloc(p, v.size()));
        if (!v.empty()) BHIP(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return PF_OK;
    }
    ~pf_batch() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : {ev_fork, ev_mid, ev_join, t0, t1})
            if (e) (void)hipEventDestroy(e);
        for (void *p : allocs) (void)hipFree(p);
        for (auto s : st)
            if (s) (void)hipStreamDestroy(s);
    }
    pf_status add_net(const pf_network &net);
    pf_status launch_net(size_t i, hipStream_t s);
    pf_status launch_all();
    pf_status configure(BatchNet &nb, int mode, int block, int per_cu);
    pf_status occupancy(const BatchNet &nb, int mode, int block, int *per_cu);
    void drop_graph() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        gexec = nullptr;
        graph = nullptr;
    }
};

pf_status pf_batch::add_net(const pf_network &src) {
    auto b = std::make_unique<BatchNet>();
    int max_lds_i = 0;
    BHIP(hipDeviceGetAttribute(&max_lds_i, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const long long max_lds = max_lds_i - 8192;   // room for the kernel's static LDS
    b->net = src;
    b->yp = build_ypattern(src);
    // tail_start: the first level of the narrow top (fewer than kWideLevel
    // nodes per level up to the dense block)
    b->sym = analyze(src, b->yp, kWideLevel - 1, opt.lu_blocked ? kBatchChain : 0, kTopWidth);
    {   // a working set beyond LDS runs in MODE 1 / 0, whose kernels take a wider dense top
        const bool mx = opt.precision != PF_FP64;
        const long long w0 = mx ? work_bytes<float>(src.nB, b->sym.n, src.nPV, b->sym.nblk())
                                : work_bytes<double>(src.nB, b->sym.n, src.nPV, b->sym.nblk());
        if (w0 > kLdsLimit && opt.lu_blocked)
            b->sym = analyze(src, b->yp, kWideLevel - 1, kBatchChainWide, kTopWidth);
    }
    chunk_tail(b->sym, kBatchChunkPairs);
    if (!b->sym.level_ordered()) {
        pf::set_last_error("symbolic analysis: nodes not numbered level by level");
        return PF_INVALID_ARG;
    }
    const pf_network &net = b->net;
    const YPattern &yp = b->yp;
    const Symbolic &sym = b->sym;
    const int nB = net.nB, n = sym.n, sl = net.slack;
    std::vector<int> node_of_bus(nB, -1), pv_of_bus(nB, -1);
    for (int k = 0; k < nB; k++)
        if (k != sl) node_of_bus[k] = sym.iperm[k - (k > sl)];
    for (int i = 0; i < net.nPV; i++) pv_of_bus[net.pv[i]] = i;
    DevNet &N = b->tp.N;
    DevSym &Y = b->tp.Y;
    N.nB = nB; N.nL = net.nL; N.slack = sl; N.nPV = net.nPV; N.n = n;
    BCHK(upload((int **)&N.rowptr, yp.rowptr));
    BCHK(upload((int **)&N.colind, yp.colind));
    BCHK(upload((int **)&N.coorow, yp.coorow));
    BCHK(upload((int **)&N.ydiag, yp.diag));
    BCHK(upload((int **)&N.yc_ptr, yp.contrib_ptr));
    BCHK(upload((int **)&N.yc, yp.contrib));
    BCHK(upload((double **)&N.br_r, net.r));
    BCHK(upload((double **)&N.br_x, net.x));
    BCHK(upload((double **)&N.br_b, net.bch));
    BCHK(upload((double **)&N.br_tr, net.tap_re));
    BCHK(upload((double **)&N.br_ti, net.tap_im));
    BCHK(upload((double **)&N.gsh, net.gsh));
    BCHK(upload((double **)&N.bsh, net.bsh));
    BCHK(alloc(&N.y, yp.nnz));
    BCHK(upload((int **)&N.node_of_bus, node_of_bus));
    N.bus_of_node = nullptr;
    BCHK(upload((int **)&N.pv_bus, net.pv));
    BCHK(upload((int **)&N.pv_of_bus, pv_of_bus));
    BCHK(upload((double **)&N.qmin, net.Qmin));
    BCHK(upload((double **)&N.qmax, net.Qmax));
    BCHK(upload((double **)&N.vset, net.Vset));
    Y.n = n; Y.nnzL = sym.nnzL; Y.nslot = sym.nslot(); Y.nlev = sym.nlev;
    Y.ch_m = sym.ch_m; Y.ch_c0 = sym.ch_c0; Y.nlev_s = sym.nlev_s; Y.di0 = sym.di0();
    BCHK(upload((int **)&Y.ch_split, sym.ch_split));
    BCHK(upload((int **)&Y.colptr, sym.colptr));
    BCHK(upload((int **)&Y.rowind, sym.rowind));
    Y.pcol = nullptr;
    BCHK(upload((int **)&Y.lvl_ptr, sym.lvl_ptr));
    BCHK(upload((int **)&Y.lvl_nodes, sym.lvl_nodes));
    BCHK(upload((int **)&Y.fd_ptr, sym.fd_ptr));
    BCHK(upload((int **)&Y.fc_ptr, sym.fc_ptr));
    BCHK(upload((int **)&Y.fc_a, sym.fc_a));
    BCHK(upload((int **)&Y.fc_b, sym.fc_b));
    BCHK(upload((int **)&Y.rs_ptr, sym.rs_ptr));
    BCHK(upload((int **)&Y.rs_k, sym.rs_k));
    BCHK(upload((int **)&Y.rs_p, sym.rs_p));
    BCHK(upload((int **)&Y.lr_pos, sym.lr_pos));
    BCHK(upload((int **)&Y.slot_e, sym.slot_e));
    BCHK(upload((int **)&Y.ld_ptr, sym.ld_ptr));
    BCHK(upload((int **)&Y.ld_node, sym.ld_node));
    BCHK(upload((int **)&Y.ld_d, sym.ld_d));
    Y.tail_start = sym.tail_start;
    Y.te0 = sym.ld_ptr[sym.tail_start];
    Y.ck_lvl = Y.ck_q0 = Y.ck_q1 = Y.te_node = Y.te_p = Y.te_ck = nullptr;
    if (opt.lu_blocked && sym.ck_max <= kBatchMaxChunks && !sym.ck_q0.empty()) {
        BCHK(upload((int **)&Y.ck_lvl, sym.ck_lvl));
        BCHK(upload((int **)&Y.ck_q0, sym.ck_q0));
        BCHK(upload((int **)&Y.ck_q1, sym.ck_q1));
        BCHK(upload((int **)&Y.te_node, sym.te_node));
        BCHK(upload((int **)&Y.te_p, sym.te_p));
        BCHK(upload((int **)&Y.te_ck, sym.te_ck));
    }
    // the factorisation as flat pair lists, its products in the step vectors
    // x and r (4 n doubles, free during the factorisation): from the first
    // level when one level's products fit there, else from the narrow top
    b->tp.F = DevFlat{};
    if (opt.lu_blocked && sym.tail_start < sym.nlev_s) {
        const bool mx = opt.precision != PF_FP64;
        auto need = [&](const FlatTail &f) {
            return mx ? flat_tail_lds<float>(f.max_pairs, f.max_col) : flat_tail_lds<double>(f.max_pairs, f.max_col);
        };
        FlatTail ft = flat_tail(sym, 0, false);
        if (need(ft) > 32LL * sym.n) ft = flat_tail(sym, sym.tail_start, false);
        if (need(ft) <= 32LL * sym.n) {
            DevFlat &F = b->tp.F;
            F.lv0 = ft.lv0;
            F.max_pairs = ft.max_pairs;
            F.max_col = ft.max_col;
            F.de_sentinel = ft.de_ptr.back();
            BCHK(upload((int **)&F.fl_ptr, ft.fl_ptr));
            BCHK(upload((int **)&F.fl_split, ft.fl_split));
            BCHK(upload((int **)&F.fl_a, ft.fl_a));
            BCHK(upload((int **)&F.fl_b, ft.fl_b));
            BCHK(upload((int **)&F.fl_s, ft.fl_s));
            BCHK(upload((int **)&F.de_ptr, ft.de_ptr));
            BCHK(upload((int **)&F.de_node, ft.de_node));
            BCHK(upload((int **)&F.de_p, ft.de_p));
            BCHK(upload((int **)&F.de_q0, ft.de_q0));
            BCHK(upload((int **)&F.de_ext, ft.de_ext));
            BCHK(upload((int **)&F.de_lo, ft.de_lo));
            BCHK(upload((int **)&F.de_hi, ft.de_hi));
            BCHK(upload((int **)&F.de_dlo, ft.de_dlo));
            BCHK(upload((int **)&F.de_dhi, ft.de_dhi));
            BCHK(upload((int **)&F.de_lr, ft.de_lr));
            BCHK(upload((int **)&F.st_g, ft.st_g));
        }
    }
    BCHK(upload((double **)&b->tp.V0flat, net.V));
    BCHK(upload((double **)&b->tp.th0flat, net.theta));
    b->tp.nslot = sym.nslot();
    b->tp.nblk = sym.nblk();
    k_ybus<<<(yp.nnz + 255) / 256, 256, 0, st[0]>>>(N, yp.nnz);
    BHIP(hipGetLastError());
    BHIP(hipStreamSynchronize(st[0]));
    const bool mixed = opt.precision != PF_FP64;
    b->ws = mixed ? work_bytes<float>(nB, n, net.nPV, sym.nblk())
                  : work_bytes<double>(nB, n, net.nPV, sym.nblk());
    b->lds = b->ws <= kLdsLimit;
    b->block = b->lds ? 256 : 1024;
    if (b->lds) {
        b->mode = 2;
        b->shm = b->ws;
    } else {
        long long pb = mixed ? partial_lds_bytes<float>(n, sym.nlev) : partial_lds_bytes<double>(n, sym.nlev);
        hipError_t e = hipSuccess;
        if (pb <= max_lds)
            for (int blk : {256, 512, 1024})
                if (e == hipSuccess)
                    e = hipFuncSetAttribute((const void *)batch_kernel(mixed, 1, blk),
                                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)pb);
        if (pb <= max_lds && e == hipSuccess) {
            b->partial = pb;
            b->mode = 1;
            b->shm = pb;
        } else {
            (void)hipGetLastError();
            b->partial = 0;
            b->mode = 0;
            b->shm = 0;
        }
    }
    nets.push_back(std::move(b));
    return PF_OK;
}

// resident workgroups per CU of k_batch<T, mode> at this block size
pf_status pf_batch::occupancy(const BatchNet &nb, int mode, int block, int *per_cu) {
    const bool mixed = opt.precision != PF_FP64;
    const size_t shm = mode == 2 ? (size_t)nb.ws : mode == 1 ? (size_t)nb.partial : 0;
    int pc = 0;
    BHIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&pc, (const void *)batch_kernel(mixed, mode, block), block,
                                                      shm));
    *per_cu = pc;
    return PF_OK;
}

// launch shape of one topology: mode, workgroup size and workgroups per CU
// (capped by occupancy); grows the global workspace when needed.
pf_status pf_batch::configure(BatchNet &nb, int mode, int block, int per_cu) {
    auto refuse = [&](const char *why) {
        pf::set_last_error(std::string("batch launch shape (mode ") + std::to_string(mode) + ", block " +
                           std::to_string(block) + "): " + why);
        return PF_INVALID_ARG;
    };
    if (mode == 2 && !nb.lds) return refuse("the working set does not fit in LDS");
    if (mode == 1 && (nb.lds || nb.partial == 0)) return refuse("mode 1 is for working sets beyond LDS whose step vectors fit");
    if (mode == 0 && nb.lds) return refuse("the working set fits in LDS: use mode 2");
    if (block < 64 || block > 1024 || block % 64) return refuse("block must be a multiple of 64 up to 1024");
    // the sweeps of modes 1 and 0 give each node of a narrow level its own 16 lanes
    if (mode != 2 && block < GL * kSweepGroupLevel) return refuse("modes 1 and 0 take at least 256 threads");
    int occ = 0;
    BCHK(occupancy(nb, mode, block, &occ));
    if (occ < 1) return refuse("no workgroup fits on a CU (registers or LDS)");
    per_cu = per_cu <= 0 ? occ : std::min(per_cu, occ);
    nb.mode = mode;
    nb.block = block;
    nb.per_cu = per_cu;
    nb.shm = mode == 2 ? nb.ws : mode == 1 ? nb.partial : 0;
    nb.grid = std::max(1, std::min<int>((int)nb.scen.size(), per_cu * cus));
    if (mode != 2) {
        const size_t need = (size_t)nb.grid * (size_t)nb.ws;
        if (need > nb.ws_cap) {
            char *p = nullptr;
            BCHK(alloc(&p, need));
            nb.d_ws = p;
            nb.ws_cap = need;
        }
    }
    return PF_OK;
}

pf_status pf_batch::launch_net(size_t i, hipStream_t s) {
    const bool mixed = opt.precision != PF_FP64;
    {
        BatchNet &b = *nets[i];
        if (b.scen.empty()) return PF_OK;
        BatchJob J{};
        J.nscen = (int)b.scen.size();
        J.scen = b.d_scen;
        J.off = d_off;
        J.Pliq = d_P; J.Qliq = d_Q; J.Qload = d_L;
        J.V0 = d_V0; J.th0 = d_T0;
        J.Vout = d_V; J.thout = d_T;
        J.iters = d_it; J.conv = d_cv; J.maxmis = d_mm;
        J.pv_off = d_pvoff; J.limQout = d_lq; J.flips = d_fl;
        J.counter = d_counter + i;
        J.gws = b.d_ws;
        J.ws_bytes = b.ws;
        J.max_iter = opt.no_max_iter;
        J.lim = opt.lim_inj_reat ? 1 : 0;
        J.refine = mixed ? opt.refine_steps : 0;
        J.tol = opt.tolerancia;
        J.diverge = 1e6;
        const size_t shm = (size_t)b.shm;
        BatchTopo tp = b.tp;
        void *args[] = {&tp, &J};
        BHIP(hipLaunchKernel((const void *)batch_kernel(mixed, b.mode, b.block), dim3(b.grid), dim3(b.block), args,
                             b.mode == 0 ? 0 : shm, s));
    }
    return PF_OK;
}

pf_status pf_batch::launch_all() {
    BHIP(hipEventRecord(ev_fork, st[0]));
    BHIP(hipStreamWaitEvent(st[1], ev_fork, 0));
    k_zero_int<<<1, 64, 0, st[0]>>>(d_counter, (int)nets.size());
    BHIP(hipEventRecord(ev_mid, st[0]));
    BHIP(hipStreamWaitEvent(st[1], ev_mid, 0));
    // the largest topology first: its few long-running workgroups take their
    // CUs before the small topology's many short ones fill the rest, so the
    // two launches overlap instead of the long one waiting for a full GPU
    std::vector<size_t> order(nets.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t c) { return nets[a]->net.nB > nets[c]->net.nB; });
    for (size_t r = 0; r < order.size(); r++) BCHK(launch_net(order[r], st[r & 1]));
    BHIP(hipEventRecord(ev_join, st[1]));
    BHIP(hipStreamWaitEvent(st[0], ev_join, 0));
    return PF_OK;
}

pf_status pf::batch_start_from_rows(pf_batch *b, int device, int rows, const float *Y, long long ld, int v_col,
                                    int th_col, const float *mu, const float *sd, void *stream) {
    if (!b || !b->n_scen || !Y || !mu || !sd || v_col < 0 || th_col < 0) return PF_INVALID_ARG;
    if (device != b->device) {
        pf::set_last_error("warm start: the network and the batch are on different GPUs");
        return PF_INVALID_ARG;
    }
    if (rows != b->n_scen) {
        pf::set_last_error("warm start: one row per batch scenario");
        return PF_INVALID_ARG;
    }
    int nbmax = 0;
    for (int s : b->net_of) nbmax = std::max(nbmax, b->nets[s]->net.nB);
    if ((long long)v_col + nbmax > ld || (long long)th_col + nbmax > ld) {
        pf::set_last_error("warm start: a row holds fewer outputs than a scenario's buses");
        return PF_INVALID_ARG;
    }
    BHIP(hipSetDevice(b->device));
    const size_t tot = (size_t)b->total;
    if (!b->warm_V0) {
        BCHK(b->alloc(&b->warm_V0, tot));
        BCHK(b->alloc(&b->warm_T0, tot));
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    k_start_from_rows<<<b->n_scen, 256, 0, st>>>(Y, ld, v_col, th_col, mu, sd, b->d_off, b->n_scen, b->total,
                                                b->warm_V0, b->warm_T0);
    BHIP(hipGetLastError());
    BHIP(hipStreamSynchronize(st));
    const bool had_warm = b->d_V0 != nullptr;
    b->d_V0 = b->warm_V0;
    b->d_T0 = b->warm_T0;
    if (!had_warm) b->drop_graph();   // kernel arguments change: capture again
    return PF_OK;
}

extern "C" {

pf_status pf_batch_create(const pf_network *const *nets, int n_nets, const pf_options *opt, int device,
                          pf_batch **out) {
    if (!nets || n_nets <= 0 || n_nets > 64 || !opt || !out) return PF_INVALID_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return PF_DEVICE_ERROR;
    auto b = std::make_unique<pf_batch>();
    b->device = device;
    b->opt = *opt;
    BHIP(hipSetDevice(device));
    for (auto &s : b->st) BHIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    BHIP(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming));
    BHIP(hipEventCreateWithFlags(&b->ev_mid, hipEventDisableTiming));
    BHIP(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming));
    BHIP(hipEventCreate(&b->t0));
    BHIP(hipEventCreate(&b->t1));
    hipDeviceProp_t prop;
    BHIP(hipGetDeviceProperties(&prop, device));
    b->cus = prop.multiProcessorCount;
    BCHK(noalloc_setup(device));
    for (int i = 0; i < n_nets; i++) {
        if (!nets[i] || nets[i]->nB < 2 || nets[i]->slack < 0) return PF_INVALID_ARG;
        BCHK(b->add_net(*nets[i]));
    }
    BCHK(b->alloc(&b->d_counter, 64));
    *out = b.release();
    return PF_OK;
}

void pf_batch_destroy(pf_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    delete b;
}

pf_status pf_batch_set_scenarios(pf_batch *b, int n_scen, const int *net_of, const double *Pliq,
                                 const double *Qliq, const double *Qload, const double *V0,
                                 const double *theta0) {
    if (!b || n_scen <= 0 || !net_of || !Pliq || !Qliq || !Qload || ((V0 == nullptr) != (theta0 == nullptr)))
        return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    const bool first = b->n_scen == 0;
    if (!first) {
        // a later call refreshes the data of the same scenario layout
        if (n_scen != b->n_scen || !std::equal(net_of, net_of + n_scen, b->net_of.begin())) {
            pf::set_last_error("pf_batch_set_scenarios: a later call must keep the first call's layout");
            return PF_INVALID_ARG;
        }
    } else {
        b->n_scen = n_scen;
        b->net_of.assign(net_of, net_of + n_scen);
        b->off.resize(n_scen);
        b->pv_off.resize(n_scen);
        long long tot = 0, tot_pv = 0;
        for (auto &nb : b->nets) nb->scen.clear();
        for (int s = 0; s < n_scen; s++) {
            if (net_of[s] < 0 || net_of[s] >= (int)b->nets.size()) {
                b->n_scen = 0;
                pf::set_last_error("pf_batch_set_scenarios: net_of holds an index outside the batch's networks");
                return PF_INVALID_ARG;
            }
            b->off[s] = tot;
            b->pv_off[s] = tot_pv;
            tot += b->nets[net_of[s]]->net.nB;
            tot_pv += b->nets[net_of[s]]->net.nPV;
            b->nets[net_of[s]]->scen.push_back(s);
        }
        b->total = tot;
        b->total_pv = tot_pv;
        BCHK(b->alloc(&b->d_P, tot));
        BCHK(b->alloc(&b->d_Q, tot));
        BCHK(b->alloc(&b->d_L, tot));
        BCHK(b->alloc(&b->d_V, tot));
        BCHK(b->alloc(&b->d_T, tot));
        BCHK(b->alloc(&b->d_mm, n_scen));
        BCHK(b->alloc(&b->d_it, n_scen));
        BCHK(b->alloc(&b->d_cv, n_scen));
        BCHK(b->alloc(&b->d_fl, n_scen));
        BCHK(b->alloc(&b->d_lq, tot_pv));
        BCHK(b->upload(&b->d_off, b->off));
        BCHK(b->upload(&b->d_pvoff, b->pv_off));
        for (auto &nbp : b->nets) {
            BatchNet &nb = *nbp;
            if (nb.scen.empty()) continue;
            BCHK(b->upload(&nb.d_scen, nb.scen));
            // default shape: as many resident workgroups as fit for LDS
            // topologies, one per CU otherwise (pf_batch_autotune refines it)
            BCHK(b->configure(nb, nb.mode, nb.block, nb.lds ? 0 : 1));
        }
    }
    const size_t tot = (size_t)b->total;
    BHIP(hipMemcpy(b->d_P, Pliq, sizeof(double) * tot, hipMemcpyHostToDevice));
    BHIP(hipMemcpy(b->d_Q, Qliq, sizeof(double) * tot, hipMemcpyHostToDevice));
    BHIP(hipMemcpy(b->d_L, Qload, sizeof(double) * tot, hipMemcpyHostToDevice));
    const bool had_warm = b->d_V0 != nullptr;
    if (V0) {
        if (!b->warm_V0) {
            BCHK(b->alloc(&b->warm_V0, tot));
            BCHK(b->alloc(&b->warm_T0, tot));
        }
        BHIP(hipMemcpy(b->warm_V0, V0, sizeof(double) * tot, hipMemcpyHostToDevice));
        BHIP(hipMemcpy(b->warm_T0, theta0, sizeof(double) * tot, hipMemcpyHostToDevice));
        b->d_V0 = b->warm_V0;
        b->d_T0 = b->warm_T0;
    } else {
        b->d_V0 = b->d_T0 = nullptr;
    }
    if (had_warm != (b->d_V0 != nullptr)) b->drop_graph();   // kernel arguments change: capture again
    return PF_OK;
}

pf_status pf_batch_solve(pf_batch *b, double *elapsed_ms) {
    if (!b || !b->n_scen) return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    if (!b->gexec) {
        BHIP(hipStreamBeginCapture(b->st[0], hipStreamCaptureModeThreadLocal));
        pf_status s = b->launch_all();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(b->st[0], &g);
        if (s != PF_OK) return s;
        BHIP(e);
        b->graph = g;
        BHIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
    }
    BHIP(hipEventRecord(b->t0, b->st[0]));
    BHIP(hipGraphLaunch(b->gexec, b->st[0]));
    BHIP(hipEventRecord(b->t1, b->st[0]));
    BHIP(hipEventSynchronize(b->t1));
    float ms = 0;
    BHIP(hipEventElapsedTime(&ms, b->t0, b->t1));
    if (elapsed_ms) *elapsed_ms = ms;
    return PF_OK;
}

// ------------------------------------------------------------- autotuner
// For every topology of the current scenario set: time one launch of its
// persistent kernel (alone, HIP events on one stream) for each candidate
// launch shape -- kernel MODE (working set in LDS / partly in LDS / global),
// workgroup size, and for LDS topologies the resident workgroups per CU --
// keep the fastest, and remember it in a text cache keyed by the device, the
// topology's sizes, the numeric options and the scenario count, so later
// runs apply it without measuring.
}  // extern "C"

namespace {

std::string tune_key(const pf_batch &b, const BatchNet &nb, const char *arch) {
    char k[512];
    std::snprintf(k, sizeof k, "%s cus=%d nB=%d n=%d nnzL=%d nlev=%d nPV=%d prec=%d refine=%d lim=%d maxit=%d nscen=%zu",
                  arch, b.cus, nb.net.nB, nb.sym.n, nb.sym.nnzL, nb.sym.nlev, nb.net.nPV, (int)b.opt.precision,
                  b.opt.refine_steps, b.opt.lim_inj_reat ? 1 : 0, b.opt.no_max_iter, nb.scen.size());
    return k;
}

bool cache_lookup(const char *path, const std::string &key, int *mode, int *block, int *per_cu) {
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char line[1024];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        char *tab = std::strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        int m, bl, pc;
        if (key == line && std::sscanf(tab + 1, "%d %d %d", &m, &bl, &pc) == 3) {
            *mode = m; *block = bl; *per_cu = pc;
            found = true;   // the last entry for a key wins
        }
    }
    std::fclose(f);
    return found;
}

void cache_store(const char *path, const std::string &key, int mode, int block, int per_cu, float ms) {
    if (!path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s\t%d %d %d %.4f\n", key.c_str(), mode, block, per_cu, ms);
    std::fclose(f);
}

}  // namespace

extern "C" {

pf_status pf_batch_autotune(pf_batch *b, const char *cache_path, int reps, int *n_measured) {
    if (!b || !b->n_scen || reps < 0) return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st[0]));
    BHIP(hipStreamSynchronize(b->st[1]));
    hipDeviceProp_t prop;
    BHIP(hipGetDeviceProperties(&prop, b->device));
    const int nrep = std::max(1, reps);
    int measured = 0;
    b->drop_graph();
    for (size_t i = 0; i < b->nets.size(); i++) {
        BatchNet &nb = *b->nets[i];
        if (nb.scen.empty()) continue;
        const std::string key = tune_key(*b, nb, prop.gcnArchName);
        int m = 0, bl = 0, pc = 0;
        if (cache_lookup(cache_path, key, &m, &bl, &pc) && b->configure(nb, m, bl, pc) == PF_OK) continue;
        struct Cand { int mode, block, per_cu; };
        std::vector<Cand> cands;
        if (nb.lds) {
            for (int block : {64, 128, 256, 512, 1024}) {
                int occ = 0;
                if (b->occupancy(nb, 2, block, &occ) != PF_OK || occ < 1) continue;
                cands.push_back({2, block, occ});
                if (occ > 2) cands.push_back({2, block, occ / 2});
            }
        } else {
            for (int mode : {1, 0}) {
                if (mode == 1 && nb.partial == 0) continue;
                for (int block : {256, 512, 1024}) {
                    int occ = 0;
                    if (b->occupancy(nb, mode, block, &occ) != PF_OK || occ < 1) continue;
                    cands.push_back({mode, block, 1});
                    if (occ >= 2) cands.push_back({mode, block, 2});
                }
            }
        }
        float best = 1e30f;
        Cand pick{nb.mode, nb.block, nb.per_cu};
        for (const Cand &c : cands) {
            BCHK(b->configure(nb, c.mode, c.block, c.per_cu));
            float t = 1e30f;
            for (int r = 0; r <= nrep; r++) {   // the first launch is a warm-up
                k_zero_int<<<1, 64, 0, b->st[0]>>>(b->d_counter, (int)b->nets.size());
                BHIP(hipEventRecord(b->t0, b->st[0]));
                BCHK(b->launch_net(i, b->st[0]));
                BHIP(hipEventRecord(b->t1, b->st[0]));
                BHIP(hipEventSynchronize(b->t1));
                float ms = 0;
                BHIP(hipEventElapsedTime(&ms, b->t0, b->t1));
                if (r > 0) t = std::min(t, ms);
            }
            measured++;
            if (t < best) {
                best = t;
                pick = c;
            }
        }
        BCHK(b->configure(nb, pick.mode, pick.block, pick.per_cu));
        cache_store(cache_path, key, pick.mode, pick.block, pick.per_cu, best);
    }
    b->drop_graph();   // launch shapes changed: the next solve captures again
    if (n_measured) *n_measured = measured;
    return PF_OK;
}

pf_status pf_batch_set_shape(pf_batch *b, int net, int mode, int block, int per_cu) {
    if (!b || !b->n_scen || net < 0 || net >= (int)b->nets.size()) return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    BatchNet &nb = *b->nets[net];
    if (nb.scen.empty()) return PF_INVALID_ARG;
    BHIP(hipStreamSynchronize(b->st[0]));
    BHIP(hipStreamSynchronize(b->st[1]));
    BCHK(b->configure(nb, mode, block, per_cu));
    b->drop_graph();
    return PF_OK;
}

pf_status pf_batch_get_shape(pf_batch *b, int net, int *mode, int *block, int *grid, long long *lds_bytes) {
    if (!b || net < 0 || net >= (int)b->nets.size()) return PF_INVALID_ARG;
    const BatchNet &nb = *b->nets[net];
    if (mode) *mode = nb.mode;
    if (block) *block = nb.block;
    if (grid) *grid = nb.grid;
    if (lds_bytes) *lds_bytes = nb.shm;
    return PF_OK;
}

pf_status pf_batch_get(pf_batch *b, double *V, double *theta, int *iterations, int *converged,
                       double *max_mismatch) {
    if (!b || !b->n_scen) return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st[0]));
    if (V) BHIP(hipMemcpy(V, b->d_V, sizeof(double) * b->total, hipMemcpyDeviceToHost));
    if (theta) BHIP(hipMemcpy(theta, b->d_T, sizeof(double) * b->total, hipMemcpyDeviceToHost));
    if (iterations) BHIP(hipMemcpy(iterations, b->d_it, sizeof(int) * b->n_scen, hipMemcpyDeviceToHost));
    if (converged) BHIP(hipMemcpy(converged, b->d_cv, sizeof(int) * b->n_scen, hipMemcpyDeviceToHost));
    if (max_mismatch) BHIP(hipMemcpy(max_mismatch, b->d_mm, sizeof(double) * b->n_scen, hipMemcpyDeviceToHost));
    return PF_OK;
}

pf_status pf_batch_get_limq(pf_batch *b, unsigned char *limQ, int *flips) {
    if (!b || !b->n_scen) return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    BHIP(hipStreamSynchronize(b->st[0]));
    if (limQ && b->total_pv > 0)
        BHIP(hipMemcpy(limQ, b->d_lq, (size_t)b->total_pv, hipMemcpyDeviceToHost));
    if (flips) BHIP(hipMemcpy(flips, b->d_fl, sizeof(int) * b->n_scen, hipMemcpyDeviceToHost));
    return PF_OK;
}

}  // extern "C"

// ------------------------------------------------- allocation-free entry

extern "C" pf_status pf_batch_solve_noalloc(pf_batch *b, int n_scen, const int *net_of, const double *Pliq,
                                            const double *Qliq, const double *Qload, const double *V0,
                                            const double *theta0, double *V, double *theta, int *iterations,
                                            int *converged, double *max_mismatch, unsigned char *limQ,
                                            int *flips, void *stream) {
    if (!b || n_scen <= 0 || !net_of || !Pliq || !Qliq || !Qload || !V || !theta || !iterations ||
        !converged || !max_mismatch || !limQ || !flips || ((V0 == nullptr) != (theta0 == nullptr)))
        return PF_INVALID_ARG;
    BHIP(hipSetDevice(b->device));
    char *base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        base = g_scratch_ptr[b->device];
    }
    if (!base) {
        pf::set_last_error("scratch address not set up on this device");
        return PF_DEVICE_ERROR;
    }
    const int nn = (int)b->nets.size();
    // layout: header, per-scenario offsets, per-network lists, workspaces
    auto up = [](long long x) { return (x + 255) & ~255LL; };
    const long long o_off = kScratchHeader;
    const long long o_pvoff = up(o_off + 8LL * n_scen);
    const long long o_scen = up(o_pvoff + 8LL * n_scen);
    const long long o_ws = up(o_scen + 4LL * n_scen * nn);
    if (o_ws > kScratchBytes) {
        pf::set_last_error("pf_batch_solve_noalloc: too many scenarios for the fixed scratch");
        return PF_INVALID_ARG;
    }
    int n_big = 0;
    for (auto &nb : b->nets) n_big += nb->mode != 2;
    const long long share = n_big ? (kScratchBytes - o_ws) / n_big : 0;
    struct Plan { int grid; long long ws_off; };
    Plan plan[64];            // at most 64 networks per batch (pf_batch_create)
    long long o = o_ws;
    for (int i = 0; i < nn; i++) {
        BatchNet &nb = *b->nets[i];
        int per_cu = nb.per_cu;
        if (nb.grid == 0) BCHK(b->occupancy(nb, nb.mode, nb.block, &per_cu));
        int grid = std::max(1, per_cu * b->cus);
        if (nb.mode != 2) {
            grid = (int)std::min<long long>(grid, share / nb.ws);
            if (grid < 1) {
                pf::set_last_error("pf_batch_solve_noalloc: a working-set slice does not fit the fixed scratch");
                return PF_INVALID_ARG;
            }
        }
        plan[i] = {grid, o};
        if (nb.mode != 2) o += up((long long)grid * nb.ws);
    }
    if (o > kScratchBytes) {
        pf::set_last_error("pf_batch_solve_noalloc: working sets exceed the fixed scratch");
        return PF_INVALID_ARG;
    }
    hipStream_t s = stream ? (hipStream_t)stream : b->st[0];
    int *count = (int *)base, *counter = (int *)(base + 256), *invalid = (int *)(base + 512);
    long long *off = (long long *)(base + o_off), *pvoff = (long long *)(base + o_pvoff);
    int *scen = (int *)(base + o_scen);
    NoallocNets info{};
    info.n_nets = nn;
    for (int i = 0; i < nn; i++) {
        info.nB[i] = b->nets[i]->net.nB;
        info.nPV[i] = b->nets[i]->net.nPV;
    }
    k_noalloc_tables<<<1, 1024, 0, s>>>(net_of, n_scen, info, off, pvoff, scen, count, counter, invalid);
    BHIP(hipGetLastError());
    const bool mixed = b->opt.precision != PF_FP64;
    for (int i = 0; i < nn; i++) {
        BatchNet &nb = *b->nets[i];
        BatchJob J{};
        J.nscen = 0;
        J.nscen_dev = count + i;
        J.scen = scen + (long long)i * n_scen;
        J.off = off;
        J.Pliq = Pliq; J.Qliq = Qliq; J.Qload = Qload;
        J.V0 = V0; J.th0 = theta0;
        J.Vout = V; J.thout = theta;
        J.iters = iterations; J.conv = converged; J.maxmis = max_mismatch;
        J.pv_off = pvoff; J.limQout = limQ; J.flips = flips;
        J.counter = counter + i;
        J.gws = nb.mode != 2 ? base + plan[i].ws_off : nullptr;
        J.ws_bytes = nb.ws;
        J.max_iter = b->opt.no_max_iter;
        J.lim = b->opt.lim_inj_reat ? 1 : 0;
        J.refine = mixed ? b->opt.refine_steps : 0;
        J.tol = b->opt.tolerancia;
        J.diverge = 1e6;
        BatchTopo tp = nb.tp;
        void *args[] = {&tp, &J};
        BHIP(hipLaunchKernel((const void *)batch_kernel(mixed, nb.mode, nb.block), dim3(plan[i].grid),
                             dim3(nb.block), args, nb.mode == 0 ? 0 : (size_t)nb.shm, s));
    }
    if (!stream) {
        BHIP(hipStreamSynchronize(s));
        int bad = 0;
        BHIP(hipMemcpy(&bad, invalid, sizeof(int), hipMemcpyDeviceToHost));
        if (bad) {
            pf::set_last_error("pf_batch_solve_noalloc: net_of holds an index outside the batch's networks");
            return PF_INVALID_ARG;
        }
    }
    return PF_OK;
}
