// Device helpers shared by the kernel translation units.
#pragma once

#include "kernels.hpp"

namespace pf {

// 2x2 block [[a, b], [c, d]] (row-major), stored as 4 consecutive T.
template <typename T>
struct B4 {
    T a, b, c, d;
};

__device__ __forceinline__ B4<double> load_b4(const double *p, int s) {
    const double2 *q = reinterpret_cast<const double2 *>(p) + 2 * (size_t)s;
    double2 u = q[0], v = q[1];
    return B4<double>{u.x, u.y, v.x, v.y};
}
__device__ __forceinline__ B4<float> load_b4(const float *p, int s) {
    float4 u = reinterpret_cast<const float4 *>(p)[s];
    return B4<float>{u.x, u.y, u.z, u.w};
}
__device__ __forceinline__ void store_b4(double *p, int s, B4<double> v) {
    double2 *q = reinterpret_cast<double2 *>(p) + 2 * (size_t)s;
    q[0] = make_double2(v.a, v.b);
    q[1] = make_double2(v.c, v.d);
}
__device__ __forceinline__ void store_b4(float *p, int s, B4<float> v) {
    reinterpret_cast<float4 *>(p)[s] = make_float4(v.a, v.b, v.c, v.d);
}

// c -= a * b
template <typename T>
__device__ __forceinline__ void mms(B4<T> &c, const B4<T> &a, const B4<T> &b) {
    c.a -= a.a * b.a + a.b * b.c;
    c.b -= a.a * b.b + a.b * b.d;
    c.c -= a.c * b.a + a.d * b.c;
    c.d -= a.c * b.b + a.d * b.d;
}

// c -= a * b as two fused multiply-adds per entry, in the order of the
// terms (two roundings per entry instead of three, two instructions instead
// of three)
template <typename T>
__device__ __forceinline__ void mms_fused(B4<T> &c, const B4<T> &a, const B4<T> &b) {
    c.a = fma(-a.b, b.c, fma(-a.a, b.a, c.a));
    c.b = fma(-a.b, b.d, fma(-a.a, b.b, c.b));
    c.c = fma(-a.d, b.c, fma(-a.c, b.a, c.c));
    c.d = fma(-a.d, b.d, fma(-a.c, b.b, c.d));
}

template <typename T>
__device__ __forceinline__ B4<T> mul(const B4<T> &a, const B4<T> &b) {
    return B4<T>{a.a * b.a + a.b * b.c, a.a * b.b + a.b * b.d, a.c * b.a + a.d * b.c,
                 a.c * b.b + a.d * b.d};
}

// c += a * b
template <typename T>
__device__ __forceinline__ void mma2(B4<T> &c, const B4<T> &a, const B4<T> &b) {
    c.a += a.a * b.a + a.b * b.c;
    c.b += a.a * b.b + a.b * b.d;
    c.c += a.c * b.a + a.d * b.c;
    c.d += a.c * b.b + a.d * b.d;
}

// v + v of the lane that the DPP control selects within the row of 16 lanes
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_add(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xF, 0xF, false);
    return v + __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The xor butterfly over aligned groups of W lanes,
//     for (off = W / 2; off > 0; off /= 2) v += v of lane (lane ^ off),
// every lane ending with the group's sum.  For W = 2, 4 and 16 the partner
// values come from DPP moves (no LDS round trip): offsets 2 and 1 are quad
// permutations, and 8 and 4 are row rotations, which read a lane holding
// the same bits as lane ^ off because each step leaves the values periodic
// in the offsets already taken.  Other widths use ds_bpermute.
template <int W, typename T>
__device__ __forceinline__ T xor_sum(T v) {
    if constexpr (W == 2 || W == 4 || W == 16) {
        if constexpr (W == 16) {
            v = dpp_add<0x128>(v);   // row_ror:8  = lane ^ 8
            v = dpp_add<0x124>(v);   // row_ror:4 ~ lane ^ 4 (period 8 now)
        }
        if constexpr (W >= 4) v = dpp_add<0x4E>(v);   // quad_perm [2,3,0,1] = lane ^ 2
        v = dpp_add<0xB1>(v);                          // quad_perm [1,0,3,2] = lane ^ 1
    } else {
#pragma unroll
        for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
    }
    return v;
}

// The sum of an aligned quad of lanes as (p0 + p1) + (p2 + p3), in every
// lane of the quad.
template <typename T>
__device__ __forceinline__ T quad_sum(T v) {
    return dpp_add<0x4E>(dpp_add<0xB1>(v));
}

template <typename T, int G>
__device__ __forceinline__ B4<T> group_sum(B4<T> v) {
    return B4<T>{xor_sum<G>(v.a), xor_sum<G>(v.b), xor_sum<G>(v.c), xor_sum<G>(v.d)};
}

// Sums the update pairs q, q + step, ... < qe of one destination block:
// a += L(p_ik) U(p_jk) and, for a column destination, u += L(p_jk) U(p_ik).
// UNR pairs at a time: their indices are loaded first, then all their
// blocks, so a lane waits for one round of dependent loads per UNR pairs
// instead of one per pair; the sum is taken in pair order either way.
template <typename T, bool WITH_U, int UNR>
__device__ __forceinline__ void pairs_acc(const DevSym &Y, const T *blk, int q, int qe, int step, B4<T> &a,
                                          B4<T> &u) {
    const int L0 = Y.n, U0 = Y.n + Y.nnzL;
    if constexpr (UNR == 1) {
        // one pair at a time, its successor's indices loaded while its
        // blocks are: one dependent load round per pair instead of two
        if (q >= qe) return;
        int pa = Y.fc_a[q], pb = Y.fc_b[q];
        for (;;) {
            const int qn = q + step;
            int na = 0, nb = 0;
            if (qn < qe) {
                na = Y.fc_a[qn];
                nb = Y.fc_b[qn];
            }
            mma2(a, load_b4(blk, L0 + pa), load_b4(blk, U0 + pb));
            if (WITH_U) mma2(u, load_b4(blk, L0 + pb), load_b4(blk, U0 + pa));
            if (qn >= qe) return;
            q = qn;
            pa = na;
            pb = nb;
        }
    }
    for (; q + (UNR - 1) * step < qe; q += UNR * step) {
        int pa[UNR], pb[UNR];
#pragma unroll
        for (int r = 0; r < UNR; r++) {
            pa[r] = Y.fc_a[q + r * step];
            pb[r] = Y.fc_b[q + r * step];
        }
        B4<T> la[UNR], ub[UNR], lb[UNR], ua[UNR];
#pragma unroll
        for (int r = 0; r < UNR; r++) {
            la[r] = load_b4(blk, L0 + pa[r]);
            ub[r] = load_b4(blk, U0 + pb[r]);
            if (WITH_U) {
                lb[r] = load_b4(blk, L0 + pb[r]);
                ua[r] = load_b4(blk, U0 + pa[r]);
            }
        }
#pragma unroll
        for (int r = 0; r < UNR; r++) {
            mma2(a, la[r], ub[r]);
            if (WITH_U) mma2(u, lb[r], ua[r]);
        }
    }
    for (; q < qe; q += step) {
        const int pik = Y.fc_a[q], pjk = Y.fc_b[q];
        mma2(a, load_b4(blk, L0 + pik), load_b4(blk, U0 + pjk));
        if (WITH_U) mma2(u, load_b4(blk, L0 + pjk), load_b4(blk, U0 + pik));
    }
}

// Narrow elimination-tree levels [l0, l1) factored by a whole workgroup:
// groups of G lanes take a level's destination blocks (every node's pivot and
// column entries), the lanes of a group split that destination's update pairs
// and reduce them in a fixed order.  After a barrier the same group finishes
// its destination: the pivot's inverse goes to the inverse-pivot region (the
// updated pivot itself stays in place, so the column groups can invert it
// too) and a column block is scaled by it, in both L copies.  Two barriers
// per level whatever its width, so the narrow top of the tree (one node per
// level) keeps every lane busy.  The indices of a group's first destination
// and of each lane's first pair on the next level do not depend on values and
// are loaded one level ahead.
template <typename T, int G, int UNR>
__device__ __forceinline__ void coop_factor_range(const DevSym &Y, T *blk, int l0, int l1, int *bad) {
    const int L0 = Y.n, U0 = Y.n + Y.nnzL, LR0 = Y.n + 2 * Y.nnzL, DI = Y.di0;
    const int grp = (int)threadIdx.x / G, lane = (int)threadIdx.x % G, ng = (int)blockDim.x / G;
    struct Dest {
        int j, d, t, q, qe, pa, pb;
    };
    auto dest = [&](int e0, int e) {
        Dest r;
        r.j = Y.ld_node[e0 + e];
        r.d = Y.ld_d[e0 + e];
        r.t = r.d - Y.fd_ptr[r.j];
        r.q = Y.fc_ptr[r.d] + lane;
        r.qe = Y.fc_ptr[r.d + 1];
        r.pa = r.pb = 0;
        if (r.q < r.qe) {
            r.pa = Y.fc_a[r.q];
            r.pb = Y.fc_b[r.q];
        }
        return r;
    };
    Dest cur{-1, -1, 0, 0, 0, 0, 0};
    if (l0 < l1 && grp < Y.ld_ptr[l0 + 1] - Y.ld_ptr[l0]) cur = dest(Y.ld_ptr[l0], grp);
    for (int l = l0; l < l1; l++) {
        const int e0 = Y.ld_ptr[l], ne = Y.ld_ptr[l + 1] - e0;
        Dest nxt{-1, -1, 0, 0, 0, 0, 0};
        if (l + 1 < l1 && grp < Y.ld_ptr[l + 2] - Y.ld_ptr[l + 1]) nxt = dest(Y.ld_ptr[l + 1], grp);
        for (int e = grp; e < ne; e += ng) {
            const Dest D = (e == grp) ? cur : dest(e0, e);
            B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
            if (D.q < D.qe) {
                mma2(a, load_b4(blk, L0 + D.pa), load_b4(blk, U0 + D.pb));
                if (D.t != 0) mma2(u, load_b4(blk, L0 + D.pb), load_b4(blk, U0 + D.pa));
            }
            if (D.t == 0) pairs_acc<T, false, UNR>(Y, blk, D.q + G, D.qe, G, a, u);
            else pairs_acc<T, true, UNR>(Y, blk, D.q + G, D.qe, G, a, u);
            a = group_sum<T, G>(a);
            u = group_sum<T, G>(u);
            if (lane == 0) {
                if (D.t == 0) {
                    const B4<T> P = load_b4(blk, D.j);
                    store_b4(blk, D.j, B4<T>{P.a - a.a, P.b - a.b, P.c - a.c, P.d - a.d});
                } else {
                    const int p = Y.colptr[D.j] + D.t - 1;
                    const B4<T> Lb = load_b4(blk, L0 + p), Ub = load_b4(blk, U0 + p);
                    store_b4(blk, L0 + p, B4<T>{Lb.a - a.a, Lb.b - a.b, Lb.c - a.c, Lb.d - a.d});
                    store_b4(blk, U0 + p, B4<T>{Ub.a - u.a, Ub.b - u.b, Ub.c - u.c, Ub.d - u.d});
                }
            }
        }
        __syncthreads();
        for (int e = grp; e < ne; e += ng) {
            if (lane != 0) continue;
            const Dest D = (e == grp) ? cur : dest(e0, e);
            const B4<T> P = load_b4(blk, D.j);
            T det = P.a * P.d - P.b * P.c;
            const bool sing = !(det != (T)0) || !isfinite((double)det);
            if (sing) det = (T)1;
            const B4<T> Pi{P.d / det, -P.b / det, -P.c / det, P.a / det};
            if (D.t == 0) {
                if (sing) atomicOr(bad, 1);
                store_b4(blk, DI + D.j, Pi);
            } else {
                const int p = Y.colptr[D.j] + D.t - 1;
                const B4<T> lb = mul(load_b4(blk, L0 + p), Pi);
                store_b4(blk, L0 + p, lb);
                store_b4(blk, LR0 + Y.lr_pos[p], lb);
            }
        }
        __syncthreads();
        cur = nxt;
    }
}

// Levels [lv0, lv1) of the chunked tail (Symbolic::ck_*): the update pairs
// of each level cut into chunks of a few pairs, each inside one
// destination, so that one round of loads covers a level whatever its
// longest destination:
//   A. lane groups of G sum one chunk each (pair q0 + lane + G*i, tree over
//      the group) into LDS (s_pa, s_pu: one per chunk of the level);
//   B. one thread per destination adds its chunks in order and updates the
//      block; an updated pivot also goes to LDS (s_piv, one per node of the
//      level);
//   C. one thread per destination inverts the pivot into the inverse-pivot
//      region or scales its column block (both L copies).
// Three barriers per level; the sums are in a fixed order, whatever the
// workgroup size.  What B and C need that does not depend on values (the
// destination's indices, chunk range, current blocks, row-copy position) is
// loaded before A, and C takes the pivot from LDS and the updated column
// block from B's registers, so only A waits on global loads.
template <typename T, int G, int UNR>
__device__ __forceinline__ void chunk_factor_levels(const DevSym &Y, T *blk, int lv0, int lv1, int *bad, B4<T> *s_pa,
                                    B4<T> *s_pu, B4<T> *s_piv) {
    const int L0 = Y.n, U0 = Y.n + Y.nnzL, LR0 = Y.n + 2 * Y.nnzL, DI = Y.di0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / G, lane = tid % G, ng = bd / G;
    for (int l = lv0; l < lv1; l++) {
        const int c0 = Y.ck_lvl[l - Y.tail_start], nc = Y.ck_lvl[l - Y.tail_start + 1] - c0;
        const int e0 = Y.ld_ptr[l] - Y.te0, ne = Y.ld_ptr[l + 1] - Y.ld_ptr[l], s0 = Y.lvl_ptr[l];
        // this thread's first destination, consumed after the chunk pass: its
        // node, column position (-1: the pivot), chunk range, current blocks
        // and row-copy position
        int mj = -1, mp = -1, mcb = 0, mce = 0, mlrp = 0;
        B4<T> mv0{0, 0, 0, 0}, mv1{0, 0, 0, 0};
        auto fetch = [&](int E, int &j, int &p, int &cb, int &ce, int &lrp, B4<T> &v0,
                         B4<T> &v1) __attribute__((always_inline)) {
            j = Y.te_node[E];
            p = Y.te_p[E];
            cb = Y.te_ck[E];
            ce = Y.te_ck[E + 1];
            if (p < 0) {
                v0 = load_b4(blk, j);
            } else {
                v0 = load_b4(blk, L0 + p);
                v1 = load_b4(blk, U0 + p);
                lrp = Y.lr_pos[p];
            }
        };
        if (tid < ne) fetch(e0 + tid, mj, mp, mcb, mce, mlrp, mv0, mv1);
        for (int c = grp; c < nc; c += ng) {
            const int q0 = Y.ck_q0[c0 + c], q1f = Y.ck_q1[c0 + c], q1 = q1f & kChunkMask;
            B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
            if (q1f & kChunkDiag) pairs_acc<T, false, UNR>(Y, blk, q0 + lane, q1, G, a, u);
            else pairs_acc<T, true, UNR>(Y, blk, q0 + lane, q1, G, a, u);
            a = group_sum<T, G>(a);
            u = group_sum<T, G>(u);
            if (lane == 0) {
                s_pa[c] = a;
                s_pu[c] = u;
            }
        }
        __syncthreads();
        B4<T> keep{0, 0, 0, 0};
        for (int e = tid; e < ne; e += bd) {
            int j = mj, p = mp, cb = mcb, ce = mce, lrp = mlrp;
            B4<T> v0 = mv0, v1 = mv1;
            if (e != tid) fetch(e0 + e, j, p, cb, ce, lrp, v0, v1);
            B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
            for (int c = cb; c < ce; c++) {
                const B4<T> x = s_pa[c - c0], y = s_pu[c - c0];
                a = B4<T>{a.a + x.a, a.b + x.b, a.c + x.c, a.d + x.d};
                u = B4<T>{u.a + y.a, u.b + y.b, u.c + y.c, u.d + y.d};
            }
            if (p < 0) {
                const B4<T> P{v0.a - a.a, v0.b - a.b, v0.c - a.c, v0.d - a.d};
                store_b4(blk, j, P);
                s_piv[j - s0] = P;
            } else {
                const B4<T> Lb{v0.a - a.a, v0.b - a.b, v0.c - a.c, v0.d - a.d};
                store_b4(blk, L0 + p, Lb);
                store_b4(blk, U0 + p, B4<T>{v1.a - u.a, v1.b - u.b, v1.c - u.c, v1.d - u.d});
                if (e == tid) keep = Lb;
            }
        }
        __syncthreads();
        for (int e = tid; e < ne; e += bd) {
            int j = mj, p = mp, lrp = mlrp;
            B4<T> Lb = keep;
            if (e != tid) {
                const int E = e0 + e;
                j = Y.te_node[E];
                p = Y.te_p[E];
                lrp = p >= 0 ? Y.lr_pos[p] : 0;
                Lb = p >= 0 ? load_b4(blk, L0 + p) : B4<T>{0, 0, 0, 0};
            }
            const B4<T> P = s_piv[j - s0];
            T det = P.a * P.d - P.b * P.c;
            const bool sing = !(det != (T)0) || !isfinite((double)det);
            if (sing) det = (T)1;
            const B4<T> Pi{P.d / det, -P.b / det, -P.c / det, P.a / det};
            if (p < 0) {
                if (sing) atomicOr(bad, 1);
                store_b4(blk, DI + j, Pi);
            } else {
                const B4<T> lb = mul(Lb, Pi);
                store_b4(blk, L0 + p, lb);
                store_b4(blk, LR0 + lrp, lb);
            }
        }
        __syncthreads();
    }
}

// Sum of the blocks [lo, hi) of an LDS array in index order; W loads are
// in flight before their additions, which keep the sequential order.
template <typename T, int W>
__device__ __forceinline__ B4<T> lds_sum_in_order(const T *v, int lo, int hi) {
    B4<T> a{0, 0, 0, 0};
    int q = lo;
    for (; q + W <= hi; q += W) {
        B4<T> x[W];
#pragma unroll
        for (int r = 0; r < W; r++) x[r] = load_b4(v, q + r);
#pragma unroll
        for (int r = 0; r < W; r++) a = B4<T>{a.a + x[r].a, a.b + x[r].b, a.c + x[r].c, a.d + x[r].d};
    }
    for (; q < hi; q++) {
        const B4<T> x = load_b4(v, q);
        a = B4<T>{a.a + x.a, a.b + x.b, a.c + x.c, a.d + x.d};
    }
    return a;
}

// Sum of an aligned group of g = 4 or 16 lanes' values p_0 .. p_{g-1} as the
// balanced tree in lane order: (p0 + p1) + (p2 + p3) in each quad, then (g =
// 16) the quads in pairs (row_half_mirror pairs lane i with 7 - i) and the
// halves (row_mirror: i with 15 - i).  Every step adds two values that are
// the same bits in both lanes of a pair, so every lane ends with the same
// sum.  g is uniform over the wave.
template <typename T>
__device__ __forceinline__ B4<T> group_tree_sum(B4<T> v, int g) {
    v = B4<T>{dpp_add<0xB1>(v.a), dpp_add<0xB1>(v.b), dpp_add<0xB1>(v.c), dpp_add<0xB1>(v.d)};   // lane ^ 1
    v = B4<T>{dpp_add<0x4E>(v.a), dpp_add<0x4E>(v.b), dpp_add<0x4E>(v.c), dpp_add<0x4E>(v.d)};   // lane ^ 2
    if (g == 16) {
        v = B4<T>{dpp_add<0x141>(v.a), dpp_add<0x141>(v.b), dpp_add<0x141>(v.c), dpp_add<0x141>(v.d)};
        v = B4<T>{dpp_add<0x140>(v.a), dpp_add<0x140>(v.b), dpp_add<0x140>(v.c), dpp_add<0x140>(v.d)};
    }
    return v;
}

// a load from an array that stays constant while the kernel runs: through
// the constant address space, so that a uniform index is a scalar load
__device__ __forceinline__ int ld_const(const int *p, int i) {
    return ((const __attribute__((address_space(4))) int *)p)[i];
}

template <typename T>
__device__ __forceinline__ B4<T> add4(const B4<T> &a, const B4<T> &x) {
    return B4<T>{a.a + x.a, a.b + x.b, a.c + x.c, a.d + x.d};
}
template <typename T>
__device__ __forceinline__ B4<T> sub4(const B4<T> &a, const B4<T> &x) {
    return B4<T>{a.a - x.a, a.b - x.b, a.c - x.c, a.d - x.d};
}

// Level steps [t0, t1) of a flat pair program (DevFlat) in one workgroup,
// two barriers per step:
//   A. one update pair per lane (strided), its block product(s) into LDS
//      (s_pa; s_pu for the column destinations' second product);
//   B. st_g[t] lanes per destination add its products (lane i: products
//      i, i + G, ... in order, then the lanes' sums as a balanced tree); a
//      pivot stores the inverse of its updated pivot, a column destination
//      forms the same updated pivot from the pivot's products, inverts it
//      and stores its L block scaled by it (both L copies) and its U block.
//      The pivot slots are left as they were (nothing after the
//      factorisation reads them; the inverse pivots carry the result).
// Only A waits on global loads, one round of them.  Everything whose
// address does not depend on this step's results is requested a step
// ahead: the next step's pair indices and its first destination's indices
// (this step's are in registers), the metadata of the step after that.  A
// step's own loads (its pairs' blocks and its destination's blocks, which
// nothing changes before this step) are then issued together and only they
// are waited for.  Every such load is unconditional: a lane without a pair
// reads pair 0's slot, a lane without a destination reads the program's
// sentinel destination (a pivot of node 0 without pairs), and a step past
// t1 is empty.  s_pa and s_pu hold F.max_pairs and F.max_col blocks.
// With staging (s_st: 2 cap blocks, cap <= kFlatStageMax; cap = 0: none),
// a step that follows one of at most cap destinations does not wait for
// its pairs' blocks in A: B of the previous step leaves each destination's
// L and U blocks in s_st as well, the pairs whose source is a node of that
// step (FlatTail::fl_s) read them there, and every other pair's blocks
// (final two steps before) are requested at the start of the previous
// step's B, whose work covers their latency.  (The step's destinations'
// own blocks are still requested in A: B needs them last.)  The arithmetic
// is the same either way.
template <typename T, int R>
__device__ __forceinline__ void flat_factor_levels(const DevSym &Y, const DevFlat &F, T *blk, int t0, int t1, int *bad, T *s_pa,
                                   T *s_pu, T *s_st = nullptr, int cap = 0) {
    if (t0 >= t1) return;
    const int L0 = Y.n, U0 = Y.n + Y.nnzL, LR0 = Y.n + 2 * Y.nnzL, DI = Y.di0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x;
    struct Step {
        int base, np, ns, e0, ne, g;
        bool f;   // the pair blocks come from staging and from requests made during the previous B
    };
    // the program's arrays are constant while it runs: their uniform loads
    // go through the scalar cache
    auto step_at = [&](int t) __attribute__((always_inline)) {
        const bool in = t < t1;
        const int tc = in ? t : t1 - 1;
        Step S;
        S.base = ld_const(F.fl_ptr, tc);
        S.np = in ? ld_const(F.fl_ptr, tc + 1) - S.base : 0;
        S.ns = ld_const(F.fl_split, tc);
        S.e0 = ld_const(F.de_ptr, tc);
        S.ne = in ? ld_const(F.de_ptr, tc + 1) - S.e0 : 0;
        S.g = ld_const(F.st_g, tc);
        S.f = cap > 0 && in && t > t0 && S.e0 - ld_const(F.de_ptr, tc > 0 ? tc - 1 : 0) <= cap;
        return S;
    };
    auto lane_group = [](int g, int x) { return g == 1 ? x : g == 4 ? x >> 2 : x >> 4; };
    int pa[R], pb[R], ps[R];
    auto pairs_of = [&](const Step &S) __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = tid + r * bd;
            const bool in = e < S.np;
            const int q = in ? S.base + e : S.base;   // the sentinel pair when the step is empty
            const int x = F.fl_a[q], y = F.fl_b[q], z = F.fl_s[q];
            pa[r] = in ? x : 0;
            pb[r] = in ? y : 0;
            ps[r] = in ? z : -1;
        }
    };
    struct Idx {
        int j, p, lo, hi, dlo, dhi, lr;
    };
    struct Val {
        B4<T> v0, v1, vd;   // own block (L, or the pivot), U block (a pivot: its pivot again), node's pivot
    };
    auto idx_of = [&](int E) __attribute__((always_inline)) {
        return Idx{F.de_node[E], F.de_p[E], F.de_lo[E], F.de_hi[E], F.de_dlo[E], F.de_dhi[E], F.de_lr[E]};
    };
    auto first_dest = [&](const Step &S) __attribute__((always_inline)) {
        const int me = lane_group(S.g, tid);
        return me < S.ne ? S.e0 + me : F.de_sentinel;
    };
    auto val_of = [&](const Idx &D) __attribute__((always_inline)) {
        const bool col = D.p >= 0;
        return Val{load_b4(blk, col ? L0 + D.p : D.j), load_b4(blk, col ? U0 + D.p : D.j), load_b4(blk, D.j)};
    };
    // destination D of a step with g lanes per destination: lane i adds
    // the products i, i + g, ... of its three sums (a: own, ad: the node's
    // pivot's, u: the U block's) in order, two at a time with their loads
    // in flight together; the group then adds its lanes' sums as a tree and
    // its first lane stores the pivot inverse or the scaled L and the U block
    auto dest_b = [&](int g, int ns, const Idx &D, const Val &V, int slot, bool stage) __attribute__((always_inline)) {
        const int lane = tid & (g - 1);
        const bool col = D.p >= 0;
        const int na = D.hi - D.lo, nd = col ? D.dhi - D.dlo : 0;
        const int nr = na > nd ? na : nd;
        B4<T> a{0, 0, 0, 0}, ad{0, 0, 0, 0}, u{0, 0, 0, 0};
        for (int q = lane; q < nr; q += 2 * g) {
            const int q2 = q + g;
            const bool a1 = q < na, a2 = q2 < na, d1 = q < nd, d2 = q2 < nd;
            B4<T> xa1{0, 0, 0, 0}, xa2{0, 0, 0, 0}, xu1{0, 0, 0, 0}, xu2{0, 0, 0, 0}, xd1{0, 0, 0, 0},
                xd2{0, 0, 0, 0};
            if (a1) xa1 = load_b4(s_pa, D.lo + q);
            if (a2) xa2 = load_b4(s_pa, D.lo + q2);
            if (col && a1) xu1 = load_b4(s_pu, D.lo - ns + q);
            if (col && a2) xu2 = load_b4(s_pu, D.lo - ns + q2);
            if (d1) xd1 = load_b4(s_pa, D.dlo + q);
            if (d2) xd2 = load_b4(s_pa, D.dlo + q2);
            if (a1) {
                a = add4(a, xa1);
                u = add4(u, xu1);
            }
            if (a2) {
                a = add4(a, xa2);
                u = add4(u, xu2);
            }
            if (d1) ad = add4(ad, xd1);
            if (d2) ad = add4(ad, xd2);
        }
        if (g > 1) {
            a = group_tree_sum(a, g);
            if (col) {
                ad = group_tree_sum(ad, g);
                u = group_tree_sum(u, g);
            }
        }
        if (lane != 0) return;
        const B4<T> P = col ? sub4(V.vd, ad) : sub4(V.v0, a);
        T det = P.a * P.d - P.b * P.c;
        const bool sing = !(det != (T)0) || !isfinite((double)det);
        if (sing) det = (T)1;
        const B4<T> Pi{P.d / det, -P.b / det, -P.c / det, P.a / det};
        if (!col) {   // the pivot slot keeps its value: B of a later round reads it
            if (sing) atomicOr(bad, 1);
            store_b4(blk, DI + D.j, Pi);
        } else {
            const B4<T> lb = mul(sub4(V.v0, a), Pi), ub = sub4(V.v1, u);
            store_b4(blk, L0 + D.p, lb);
            store_b4(blk, LR0 + D.lr, lb);
            store_b4(blk, U0 + D.p, ub);
            if (stage) {
                store_b4(s_st, 2 * slot, lb);
                store_b4(s_st, 2 * slot + 1, ub);
            }
        }
    };
    Step S = step_at(t0), S1 = step_at(t0 + 1);
    pairs_of(S);
    Idx nx = idx_of(first_dest(S));   // this thread's first destination of the coming step
    B4<T> la[R], ub[R], lb[R], ua[R];   // this step's pair blocks
    auto request = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            la[r] = load_b4(blk, L0 + pa[r]);
            ub[r] = load_b4(blk, U0 + pb[r]);
            lb[r] = load_b4(blk, L0 + pb[r]);
            ua[r] = load_b4(blk, U0 + pa[r]);
        }
    };
    for (int t = t0; t < t1; t++) {
        const Idx mi = nx;
        const Val mv = val_of(mi);
        if (S.f) {   // requested during the last B; the pairs from its nodes from staging
#pragma unroll
            for (int r = 0; r < R; r++)
                if (ps[r] >= 0) {
                    const int sa = ps[r] & 0xffff, sb = ps[r] >> 16;
                    la[r] = load_b4(s_st, 2 * sa);
                    ua[r] = load_b4(s_st, 2 * sa + 1);
                    lb[r] = load_b4(s_st, 2 * sb);
                    ub[r] = load_b4(s_st, 2 * sb + 1);
                }
        } else {
            request();
        }
        pairs_of(S1);
        nx = idx_of(first_dest(S1));
#pragma unroll
        for (int r = 0; r < R; r++) {
            const int e = tid + r * bd;
            if (e < S.np) {
                store_b4(s_pa, e, mul(la[r], ub[r]));
                if (e >= S.ns) store_b4(s_pu, e - S.ns, mul(lb[r], ua[r]));
            }
        }
        for (int e = tid + R * bd; e < S.np; e += bd) {
            const int qa = F.fl_a[S.base + e], qb = F.fl_b[S.base + e];
            store_b4(s_pa, e, mul(load_b4(blk, L0 + qa), load_b4(blk, U0 + qb)));
            if (e >= S.ns) store_b4(s_pu, e - S.ns, mul(load_b4(blk, L0 + qb), load_b4(blk, U0 + qa)));
        }
        __syncthreads();
        {   // B: the thread's first destination from registers, further ones (wide steps) loaded here
            if (S1.f) request();   // the next step's pair blocks that are final already
            const int grp = lane_group(S.g, tid), ng = lane_group(S.g, bd);
            if (grp < S.ne) dest_b(S.g, S.ns, mi, mv, grp, S1.f);
            for (int e = grp + ng; e < S.ne; e += ng) {
                const Idx D = idx_of(S.e0 + e);
                dest_b(S.g, S.ns, D, val_of(D), e, S1.f);
            }
        }
        __syncthreads();
        S = S1;
        S1 = step_at(t + 2);
    }
}

// ------------------------------------------------- trailing dense chain
// The top of a power-network elimination tree ends in a chain of single
// nodes whose columns are full (the separator between the network's areas).
// Instead of one level per node, the chain is one dense block: its Schur
// complement S = A_CC - L_CN U_NC (the updates from every node before the
// chain, all of them independent) and then S^-1 by Gauss-Jordan on 2x2
// blocks, without pivoting like the rest of the factorisation.  The forward
// and backward solves of the chain collapse into
//     x_C = S^-1 (b_C - L_CN y_N)
// so a solve costs two barriers there instead of two levels per node.

// S = chain block minus the updates of the nodes before the chain, written
// row-major (ch_m x ch_m blocks) at block n + 3 nnzL over a zeroed S (the
// entries outside the pattern of L stay zero).  Groups of G lanes
// per destination block (group group0 + threadIdx.x / G of ngroups, so that
// one or many workgroups can share the pass); every destination is
// independent.  The caller orders it after every node before the chain.
template <typename T, int G, int UNR>
__device__ __forceinline__ void chain_schur(const DevSym &Y, T *blk, int group0, int ngroups) {
    const int m = Y.ch_m, c0 = Y.ch_c0;
    const int L0 = Y.n, U0 = Y.n + Y.nnzL, SB = Y.n + 3 * Y.nnzL;
    const int tid = (int)threadIdx.x;
    {
        const int e0 = Y.ld_ptr[Y.nlev_s], ne = Y.ld_ptr[Y.nlev] - e0, dbase = Y.fd_ptr[c0];
        const int grp = group0 + tid / G, lane = tid % G, ng = ngroups;
        for (int e = grp; e < ne; e += ng) {
            const int j = Y.ld_node[e0 + e], d = Y.ld_d[e0 + e];
            const int t = d - Y.fd_ptr[j], a = j - c0, qend = Y.ch_split[d - dbase];
            B4<T> acc{0, 0, 0, 0}, uacc{0, 0, 0, 0};
            if (t == 0) pairs_acc<T, false, UNR>(Y, blk, Y.fc_ptr[d] + lane, qend, G, acc, uacc);
            else pairs_acc<T, true, UNR>(Y, blk, Y.fc_ptr[d] + lane, qend, G, acc, uacc);
            acc = group_sum<T, G>(acc);
            uacc = group_sum<T, G>(uacc);
            if (lane == 0) {
                if (t == 0) {
                    const B4<T> D = load_b4(blk, j);
                    store_b4(blk, SB + a * m + a, B4<T>{D.a - acc.a, D.b - acc.b, D.c - acc.c, D.d - acc.d});
                } else {
                    const int p = Y.colptr[j] + t - 1, i = Y.rowind[p] - c0;
                    const B4<T> Lb = load_b4(blk, L0 + p), Ub = load_b4(blk, U0 + p);
                    store_b4(blk, SB + i * m + a, B4<T>{Lb.a - acc.a, Lb.b - acc.b, Lb.c - acc.c, Lb.d - acc.d});
                    store_b4(blk, SB + a * m + i,
                             B4<T>{Ub.a - uacc.a, Ub.b - uacc.b, Ub.c - uacc.c, Ub.d - uacc.d});
                }
            }
        }
    }
}

// 2x2 pivot inverse; a zero or non-finite determinant raises *bad (thread 0)
template <typename T>
__device__ __forceinline__ B4<T> pivot_inverse(const B4<T> &pv, int *bad) {
    T det = pv.a * pv.d - pv.b * pv.c;
    if (!(det != (T)0) || !isfinite((double)det)) {
        if (threadIdx.x == 0) atomicOr(bad, 1);
        det = (T)1;
    }
    return B4<T>{pv.d / det, -pv.b / det, -pv.c / det, pv.a / det};
}

// 2x2 pivot inverse formed by one thread, which raises *bad itself
template <typename T>
__device__ __forceinline__ B4<T> pivot_inverse_own(const B4<T> &pv, int *bad) {
    T det = pv.a * pv.d - pv.b * pv.c;
    if (!(det != (T)0) || !isfinite((double)det)) {
        atomicOr(bad, 1);
        det = (T)1;
    }
    return B4<T>{pv.d / det, -pv.b / det, -pv.c / det, pv.a / det};
}

// One Gauss-Jordan update of block (i, j) at step k, from the step's pivot
// inverse P and the old row k (rk = S_kj) and column k (ci = S_ik):
//   S_kk <- P; S_kj <- P S_kj; S_ik <- -S_ik P; S_ij <- S_ij - S_ik (P S_kj).
template <typename T>
__device__ __forceinline__ void gj_update(B4<T> &s, int i, int j, int k, const B4<T> &P, const B4<T> &rk,
                                          const B4<T> &ci) {
    if (i == k) {
        s = (j == k) ? P : mul(P, rk);
    } else if (j == k) {
        const B4<T> c = mul(ci, P);
        s = B4<T>{-c.a, -c.b, -c.c, -c.d};
    } else {
        mms(s, ci, mul(P, rk));
    }
}

// Overwrites S (formed by chain_schur) with S^-1.  S lives in registers
// during the elimination, as TR x TC tiles of blocks: thread t owns rows
// (t / ntc) * TR + a and columns (t % ntc) * TC + c, with ntc = ceil(m / TC)
// (requires ceil(m / TR) * ceil(m / TC) <= blockDim).  Row k and column k go
// through LDS (lds: 4 ch_m + 2 blocks, double buffered: one barrier per
// step); a thread reads TR + TC blocks of them per step, not two per block it
// owns.  The thread holding the pivot block inverts it before the barrier
// and leaves the inverse in LDS beside the row, so the other threads do not
// each invert it (the update is bound by arithmetic, not by the barrier).
// In FP32 every block of a tile takes the general update
// S_ij -= S_ik (P S_kj) without a test; the few threads holding row k or
// column k then overwrite those blocks with P S_kj, -S_ik P and P, so the
// update loop carries no per-block branches (blocks past m are padding,
// updated but never stored).  Both forms give the same values.
// FUSED (FP32): the general update as mms_fused, which rounds differently
// from mms; the memory fallback and the batch kernel keep mms.
// The caller orders it after chain_schur and synchronises after.
template <typename T, int TR, int TC, bool FUSED = false>
__device__ __forceinline__ void chain_invert(const DevSym &Y, T *blk, int *bad, B4<T> *lds) {
    const int m = Y.ch_m, SB = Y.n + 3 * Y.nnzL;
    const int tid = (int)threadIdx.x;
    const int ntc = (m + TC - 1) / TC, ntr = (m + TR - 1) / TR;
    const bool own = tid < ntr * ntc;
    const int i0 = own ? (tid / ntc) * TR : m, j0 = own ? (tid % ntc) * TC : m;
    B4<T> reg[TR][TC];
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int c = 0; c < TC; c++)
            reg[a][c] = (i0 + a < m && j0 + c < m) ? load_b4(blk, SB + (i0 + a) * m + j0 + c) : B4<T>{0, 0, 0, 0};
    // LDS index of tile column c / row a, clamped into [0, m)
    auto jc = [&](int c) { return min(j0 + c, m - 1); };
    auto ic = [&](int a) { return min(i0 + a, m - 1); };
    for (int k = 0; k < m; k++) {
        B4<T> *s_row = lds + (k & 1) * (2 * m + 1), *s_col = s_row + m, *s_p = s_col + m;
        const int ar = k - i0, ac = k - j0;   // pivot row / column inside this tile when in [0, TR) / [0, TC)
        if (k >= i0 && k < i0 + TR) {   // this thread holds part of row k
#pragma unroll
            for (int a = 0; a < TR; a++)
                if (i0 + a == k) {
#pragma unroll
                    for (int c = 0; c < TC; c++) {
                        if (j0 + c < m) s_row[j0 + c] = reg[a][c];
                        if (j0 + c == k) *s_p = pivot_inverse_own(reg[a][c], bad);
                    }
                }
        }
        if (k >= j0 && k < j0 + TC) {   // ... of column k
#pragma unroll
            for (int c = 0; c < TC; c++)
                if (j0 + c == k) {
#pragma unroll
                    for (int a = 0; a < TR; a++)
                        if (i0 + a < m) s_col[i0 + a] = reg[a][c];
                }
        }
        __syncthreads();
        if (!own) continue;
        const B4<T> P = *s_p;
        if constexpr (sizeof(T) == 4) {   // registers to spare: read row and column once
            B4<T> pr[TC], ci[TR];
#pragma unroll
            for (int c = 0; c < TC; c++) pr[c] = mul(P, s_row[jc(c)]);
#pragma unroll
            for (int a = 0; a < TR; a++) ci[a] = s_col[ic(a)];
#pragma unroll
            for (int a = 0; a < TR; a++)
#pragma unroll
                for (int c = 0; c < TC; c++) {
                    if constexpr (FUSED) mms_fused(reg[a][c], ci[a], pr[c]);
                    else mms(reg[a][c], ci[a], pr[c]);
                }
            if ((unsigned)ar < (unsigned)TR) {   // row k: P S_kj
#pragma unroll
                for (int a = 0; a < TR; a++)
                    if (a == ar) {
#pragma unroll
                        for (int c = 0; c < TC; c++) reg[a][c] = pr[c];
                    }
            }
            if ((unsigned)ac < (unsigned)TC) {   // column k: -S_ik P, and P on the pivot
#pragma unroll
                for (int c = 0; c < TC; c++)
                    if (c == ac) {
#pragma unroll
                        for (int a = 0; a < TR; a++) {
                            const B4<T> v = mul(ci[a], P);
                            reg[a][c] = (a == ar) ? P : B4<T>{-v.a, -v.b, -v.c, -v.d};
                        }
                    }
            }
        } else {
            // column by column: one P S_kj live at a time, S_ik re-read from
            // LDS; the register budget leaves no room for the unconditional
            // form, so the pivot row and column are tested per block
#pragma unroll
            for (int c = 0; c < TC; c++) {
                const int j = j0 + c;
                if (j >= m) continue;
                const B4<T> prc = (j != k) ? mul(P, s_row[j]) : P;
#pragma unroll
                for (int a = 0; a < TR; a++) {
                    const int i = i0 + a;
                    if (i >= m) continue;
                    if (i == k) {
                        reg[a][c] = prc;
                    } else if (j == k) {
                        const B4<T> v = mul(s_col[i], P);
                        reg[a][c] = B4<T>{-v.a, -v.b, -v.c, -v.d};
                    } else {
                        mms(reg[a][c], s_col[i], prc);
                    }
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < TR; a++)
#pragma unroll
        for (int c = 0; c < TC; c++)
            if (i0 + a < m && j0 + c < m) store_b4(blk, SB + (i0 + a) * m + j0 + c, reg[a][c]);
}

// Schur complement and inverse of the dense top block in one workgroup (lds:
// 4 ch_m + 2 blocks).  S sits in registers in the smallest square tiles of at
// most TMAX x TMAX blocks that the workgroup covers; when none does, S is
// updated where it is stored (LDS or a global slice), two barriers per
// step.  Every variant does the same arithmetic.
template <typename T, int G, int UNR, int TMAX>
__device__ __forceinline__ void chain_factor_mem(const DevSym &Y, T *blk, int *bad, B4<T> *lds) {
    const int m = Y.ch_m, mm = m * m, SB = Y.n + 3 * Y.nnzL;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x;
    for (int q = tid; q < mm; q += bd) store_b4(blk, SB + q, B4<T>{0, 0, 0, 0});   // outside L: zero
    __syncthreads();
    chain_schur<T, G, UNR>(Y, blk, 0, bd / G);
    __syncthreads();
    auto fits = [&](int t) { return ((m + t - 1) / t) * ((m + t - 1) / t) <= bd; };
    if (fits(1)) { chain_invert<T, 1, 1>(Y, blk, bad, lds); return; }
    if (TMAX >= 2 && fits(2)) { chain_invert<T, 2, 2>(Y, blk, bad, lds); return; }
    if (TMAX >= 3 && fits(3)) { chain_invert<T, (TMAX >= 3 ? 3 : 1), (TMAX >= 3 ? 3 : 1)>(Y, blk, bad, lds); return; }
    if (TMAX >= 4 && fits(4)) { chain_invert<T, (TMAX >= 4 ? 4 : 1), (TMAX >= 4 ? 4 : 1)>(Y, blk, bad, lds); return; }
    B4<T> *s_row = lds, *s_col = lds + m;
    for (int k = 0; k < m; k++) {
        for (int t = tid; t < m; t += bd) {
            s_row[t] = load_b4(blk, SB + k * m + t);
            s_col[t] = load_b4(blk, SB + t * m + k);
        }
        __syncthreads();
        const B4<T> P = pivot_inverse(s_row[k], bad);
        for (int q = tid; q < mm; q += bd) {
            const int i = q / m, j = q - i * m;
            B4<T> s = load_b4(blk, SB + q);
            gj_update(s, i, j, k, P, s_row[j], s_col[i]);
            store_b4(blk, SB + q, s);
        }
        __syncthreads();
    }
}

// Chain part of a solve, after the forward sweep over the levels before the
// chain: x_C <- S^-1 (x_C - L_CN x_N), where x_C still holds the right-hand
// side (s_r: 2 ch_m doubles of LDS).  The backward sweep then starts below
// the chain.  Two barriers.
template <typename T, int G>
__device__ __forceinline__ void chain_solve(const DevSym &Y, const T *blk, double *x, double *s_r) {
    const int m = Y.ch_m, c0 = Y.ch_c0;
    const int LR0 = Y.n + 2 * Y.nnzL, SB = Y.n + 3 * Y.nnzL;
    const int tid = (int)threadIdx.x, grp = tid / G, lane = tid % G, ng = (int)blockDim.x / G;
    for (int a = grp; a < m; a += ng) {
        const int c = c0 + a;
        double a0 = 0, a1 = 0;
        for (int q = Y.rs_ptr[c] + lane; q < Y.rs_ptr[c + 1]; q += G) {
            const int k = Y.rs_k[q];
            if (k >= c0) break;                 // row entries ascend in k
            const B4<T> L = load_b4(blk, LR0 + q);
            const double x0 = x[2 * k], x1 = x[2 * k + 1];
            a0 += (double)L.a * x0 + (double)L.b * x1;
            a1 += (double)L.c * x0 + (double)L.d * x1;
        }
        a0 = xor_sum<G>(a0);
        a1 = xor_sum<G>(a1);
        if (lane == 0) {
            s_r[2 * a] = x[2 * c] - a0;
            s_r[2 * a + 1] = x[2 * c + 1] - a1;
        }
    }
    __syncthreads();
    for (int a = grp; a < m; a += ng) {
        double a0 = 0, a1 = 0;
        for (int b = lane; b < m; b += G) {
            const B4<T> S = load_b4(blk, SB + a * m + b);
            const double r0 = s_r[2 * b], r1 = s_r[2 * b + 1];
            a0 += (double)S.a * r0 + (double)S.b * r1;
            a1 += (double)S.c * r0 + (double)S.d * r1;
        }
        a0 = xor_sum<G>(a0);
        a1 = xor_sum<G>(a1);
        if (lane == 0) {
            x[2 * (c0 + a)] = a0;
            x[2 * (c0 + a) + 1] = a1;
        }
    }
    __syncthreads();
}

// Smith's complex division (CACM Algorithm 116, 1962): (ar + j ai) / (br + j bi)
__device__ __forceinline__ void smith_div(double ar, double ai, double br, double bi, double &cr,
                                          double &ci) {
#pragma clang fp contract(off)
    if (fabs(br) >= fabs(bi)) {
        double r = bi / br, d = br + bi * r;
        cr = (ar + ai * r) / d;
        ci = (ai - ar * r) / d;
    } else {
        double r = br / bi, d = bi + br * r;
        cr = (ar * r + ai) / d;
        ci = (ai * r - ar) / d;
    }
}

__device__ __forceinline__ unsigned long long double_as_u64(double v) {
    return (unsigned long long)__double_as_longlong(v);
}
__device__ __forceinline__ double u64_as_double(unsigned long long v) {
    return __longlong_as_double((long long)v);
}

// max that keeps a NaN (a non-finite mismatch must never look converged)
__device__ __forceinline__ double dmax_nan(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ double wave_sum(double v) {
    v += __shfl_xor(v, 32);
    v += __shfl_xor(v, 16);
    return xor_sum<16>(v);   // the butterfly's offsets 8 .. 1 (period 16 from here)
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = dmax_nan(v, __shfl_xor(v, off));
    return v;
}

}  // namespace pf
