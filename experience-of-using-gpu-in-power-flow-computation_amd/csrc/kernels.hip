// Device kernels of the single-network NR solver (gfx950, wave64).
//
// Stage -> reference kernel it replaces:
//   k_ybus          host calcYbusSp_Matpower, CPU/esparso/spCalcYbus.cpp:4-90
//   k_pq            calcPeficiente_0based / calcQeficiente_0based, GPU/PQcalc.cuh:177-219
//   k_mismatch      calcResPLim / calcResQLim, GPU/newtonRaphson.cuh:35-58
//   k_limits        host chkLimQ / undo_chkLimQ, CPU/limInjReat.cpp:38-131
//   k_jfill         SpCalcH/L/M/N, GPU/Jacobiano.cuh:252-344
//   k_factor_*,
//   k_fwd_*/k_bwd_* cusolverSpDcsrlsvqr, GPU/newtonRaphson.cuh:466-468
//   k_update        attOlim / attVlim, GPU/newtonRaphson.cuh:76-119
//   k_dense         cusolverDnDgetrf/getrs, GPU/newtonRaphson.cuh:507-530
#include "kernels_impl.hpp"

namespace pf {

// ------------------------------------------------------------------- Ybus
// One thread per CSR slot; the slot's branch terms are summed in branch
// order and the bus shunt last, with no contraction into FMA: the result is
// bitwise the same as a sequential host assembly in branch order.
__global__ void k_ybus(DevNet N, int nnz) {
#pragma clang fp contract(off)
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nnz) return;
    double re = 0.0, im = 0.0;
    for (int q = N.yc_ptr[e]; q < N.yc_ptr[e + 1]; q++) {
        int code = N.yc[q];
        int l = code >> 2, t = code & 3;
        double r = N.br_r[l], x = N.br_x[l], b = N.br_b[l], tr = N.br_tr[l], ti = N.br_ti[l];
        double vr, vi;
        if (t == 0 || t == 3) {
            double yr, yi;
            smith_div(1.0, 0.0, r, x, yr, yi);                    // y = 1/z
            if (t == 0) {
                double s = 1.0 / (tr * tr + ti * ti);              // 1/|t|^2
                vr = yr * s;
                vi = (yi + b / 2.) * s;                            // (y + jb/2)/|t|^2
            } else {
                vr = yr;
                vi = yi + b / 2.;                                   // y + jb/2
            }
        } else {
            double ar, ai;
            if (t == 1) smith_div(-1.0, 0.0, tr, -ti, ar, ai);    // -1/conj(t)
            else smith_div(-1.0, 0.0, tr, ti, ar, ai);            // -1/t
            smith_div(ar, ai, r, x, vr, vi);                       // .../z
        }
        re += vr;
        im += vi;
    }
    int k = N.coorow[e];
    if (k == N.colind[e]) {
        re += N.gsh[k];
        im += N.bsh[k];
    }
    N.y[e] = make_double2(re, im);
}

// --------------------------------------------------------------- voltages
__global__ void k_init_E(DevNet N, DevState S) {
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= N.nB) return;
    double s, c;
    sincos(S.th[k], &s, &c);
    S.E[k] = make_double2(S.V[k] * c, S.V[k] * s);
}

// --------------------------------------------------------------------- PQ
// S_k = E_k conj(sum_m Y_km E_m): LPR lanes per bus row, E staged in LDS for
// a window of buses around the block's rows (bus numbering keeps most
// neighbours close), gathered from global memory outside the window.
template <int LPR>
__global__ void __launch_bounds__(256) k_pq(DevNet N, DevState S, int gate) {
    constexpr int ROWS = 256 / LPR;
    constexpr int WIN = 192;
    __shared__ double2 win[ROWS + 2 * WIN];
    if (gate && S.ctl->done) return;
    const int r0 = blockIdx.x * ROWS;
    const int r1 = min(N.nB, r0 + ROWS);
    const int lo = max(0, r0 - WIN), hi = min(N.nB, r1 + WIN);
    for (int i = threadIdx.x; i < hi - lo; i += blockDim.x) win[i] = S.E[lo + i];
    __syncthreads();
    const int row = r0 + (int)threadIdx.x / LPR;
    const int lane = (int)threadIdx.x % LPR;
    double ar = 0.0, ai = 0.0;
    if (row < N.nB) {
        for (int e = N.rowptr[row] + lane; e < N.rowptr[row + 1]; e += LPR) {
            int m = N.colind[e];
            double2 Em = (m >= lo && m < hi) ? win[m - lo] : S.E[m];
            double2 y = N.y[e];
            ar += y.x * Em.x - y.y * Em.y;
            ai += y.x * Em.y + y.y * Em.x;
        }
    }
    ar = xor_sum<LPR>(ar);
    ai = xor_sum<LPR>(ai);
    if (row < N.nB && lane == 0) {
        double2 Ek = win[row - lo];
        S.I[row] = make_double2(ar, ai);
        S.P[row] = Ek.x * ar + Ek.y * ai;
        S.Q[row] = Ek.y * ar - Ek.x * ai;
    }
}

// ---------------------------------------------------------------- limits
// One workgroup.  mode 0: start of a pass (reference trigger at iteration
// >= 1: chkLimQ when max|previous dx| < 0.1, else undo_chkLimQ if any bus is
// limited, CPU/newtonRaphson.cpp:202-211); mode 1: switch taken at
// convergence (:252-256); mode 2 / 3: unconditional chkLimQ / undo_chkLimQ.
__global__ void __launch_bounds__(1024) k_limits(DevNet N, DevState S, int mode, int lim) {
    __shared__ int s_flips;
    Ctl *ctl = S.ctl;
    if (mode == 0 && (ctl->done || !lim || ctl->it < 1)) return;
    if (mode == 1 && !ctl->post) return;
    bool do_chk;
    if (mode == 0) do_chk = u64_as_double(ctl->maxDx_bits) < 0.1;
    else do_chk = (mode == 1 || mode == 2);
    if (threadIdx.x == 0) s_flips = 0;
    int any = 0;
    if (!do_chk) {
        for (int i = threadIdx.x; i < N.nPV; i += blockDim.x) any |= S.limQ[i];
    }
    any = __syncthreads_or(any);
    if (!do_chk && !any) return;
    int flips = 0;
    for (int i = threadIdx.x; i < N.nPV; i += blockDim.x) {
        const int b = N.pv_bus[i];
        const double qload = S.Qload[b];
        const double qmin = N.qmin[i], qmax = N.qmax[i];
        unsigned char lq = S.limQ[i];
        double qll = S.QliqLim[b];
        if (!lq) {
            if (do_chk) {
                double qg = S.Q[b] + qload;
                if (qg < qmin - kFolga) {
                    qll = qmin - qload;
                    lq = 1;
                    flips++;
                } else if (qg > qmax + kFolga) {
                    qll = qmax - qload;
                    lq = 1;
                    flips++;
                }
            }
        } else {
            if (qll == qmax - qload) {
                if (S.V[b] > N.vset[i] + kFolga) {
                    lq = 0;
                    qll = S.Qliq[b];
                    flips++;
                }
            } else {
                if (S.V[b] < N.vset[i] - kFolga) {
                    lq = 0;
                    qll = S.Qliq[b];
                    flips++;
                }
            }
        }
        S.limQ[i] = lq;
        S.QliqLim[b] = qll;
        S.isq[b] = lq;
    }
    if (flips) atomicAdd(&s_flips, flips);
    __syncthreads();
    if (threadIdx.x == 0 && s_flips) atomicAdd(&ctl->flips, s_flips);
}

// --------------------------------------------------------------- mismatch
__global__ void __launch_bounds__(256) k_mismatch(DevNet N, DevState S, int gate) {
    if (gate && S.ctl->done) return;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    unsigned bad = 0;
    if (k < N.nB && k != N.slack) {
        int j = N.node_of_bus[k];
        double gp = S.Pliq[k] - S.P[k];
        double gq = S.isq[k] ? S.QliqLim[k] - S.Q[k] : 0.0;
        S.g[2 * j] = gp;
        S.g[2 * j + 1] = gq;
        m = dmax_nan(fabs(gp), fabs(gq));
        int i = N.pv_of_bus[k];
        if (i >= 0) {
            double qg = S.Q[k] + S.Qload[k];
            bad = (qg < N.qmin[i] - kFolga) || (qg > N.qmax[i] + kFolga);
        }
    }
    m = wave_max(m);
    bad = __any(bad);
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&S.ctl->maxG_bits, double_as_u64(m));
        if (bad) atomicOr(&S.ctl->simple_bad, 1u);
    }
}

// Decision of a pass (CPU/newtonRaphson.cpp:247-264): converged, switch at
// convergence (zero step, as the reference's step there is O(tol)), or step.
__global__ void k_control(DevState S, int lim) {
    if (threadIdx.x != 0) return;
    Ctl *c = S.ctl;
    if (c->done) return;
    double maxG = u64_as_double(c->maxG_bits);
    c->maxG = maxG;
    if (maxG <= c->tol) {
        if (lim && c->simple_bad) {
            c->post = 1;
            c->step = 0;
            c->it++;
        } else {
            c->done = 1;
            c->step = 0;
            c->post = 0;
        }
    } else {
        c->step = 1;
        c->post = 0;
        c->it++;
    }
    c->maxG_bits = 0;
    c->simple_bad = 0;
    c->maxDx_bits = 0;
}

// ----------------------------------------------------------- Jacobian fill
// One thread per factor slot.  With c = E_k conj(Y_km E_m), k != m:
//   H = Im c, N = Re c / V_m, M = -Re c, L = Im c / V_m
// and on the diagonal (CPU/esparso/spCalcJac.cpp:426-547):
//   H = -Q - V^2 B, N = (P + V^2 G)/V, M = P - V^2 G, L = (Q - V^2 B)/V.
// Rows of buses without a Q equation become dV = 0; their dV columns are 0.
template <typename T>
__global__ void __launch_bounds__(256) k_jfill(DevNet N, DevState S, DevSym Y, T *blk, int gate) {
    if (gate && !S.ctl->step) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Y.nslot) {   // threads past the slots clear the dense top block S
        const int q = s - Y.nslot;
        if (q < Y.ch_m * Y.ch_m) store_b4(blk, Y.n + 3 * Y.nnzL + q, B4<T>{0, 0, 0, 0});
        return;
    }
    int e = Y.slot_e[s];
    B4<T> out;
    if (e < 0) {
        out = B4<T>{0, 0, 0, 0};
    } else {
        int k = N.coorow[e], m = N.colind[e];
        double2 y = N.y[e];
        double H, Nn, M, L;
        if (k != m) {
            double2 Em = S.E[m], Ek = S.E[k];
            double tr = y.x * Em.x - y.y * Em.y, ti = y.x * Em.y + y.y * Em.x;   // Y_km E_m
            double cr = Ek.x * tr + Ek.y * ti, ci = Ek.y * tr - Ek.x * ti;       // E_k conj(.)
            double vm = S.V[m];
            H = ci;
            Nn = cr / vm;
            M = -cr;
            L = ci / vm;
        } else {
            double v = S.V[k], v2 = v * v, p = S.P[k], q = S.Q[k];
            H = -q - v2 * y.y;
            Nn = (p + v2 * y.x) / v;
            M = p - v2 * y.x;
            L = (q - v2 * y.y) / v;
        }
        if (!S.isq[m]) { Nn = 0; L = 0; }
        if (!S.isq[k]) { M = 0; L = (k == m) ? 1.0 : 0.0; }
        out = B4<T>{(T)H, (T)Nn, (T)M, (T)L};
    }
    store_b4(blk, s, out);
}

// ------------------------------------------------------------ block LU
// Left-looking 2x2-block LU without pivoting over the symbolic program:
// node j gathers every update from its descendants, so the nodes of one
// elimination-tree level are independent.  One wavefront per node, one lane
// per destination block (diagonal first).
template <typename T>
__device__ __forceinline__ void factor_node(const DevSym &Y, T *blk, int j, int lane, int *err) {
    const int d0 = Y.fd_ptr[j], nd = Y.fd_ptr[j + 1] - d0;
    const int L0 = Y.n, U0 = Y.n + Y.nnzL;
    B4<T> dinv{0, 0, 0, 0};
    for (int base = 0; base < nd; base += 64) {
        const int t = base + lane;
        B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
        int p = -1;
        if (t < nd) {
            const int d = d0 + t;
            B4<T> sa{0, 0, 0, 0}, su{0, 0, 0, 0};
            if (t == 0) {
                pairs_acc<T, false, 4>(Y, blk, Y.fc_ptr[d], Y.fc_ptr[d + 1], 1, sa, su);
                a = load_b4(blk, j);
            } else {
                pairs_acc<T, true, 4>(Y, blk, Y.fc_ptr[d], Y.fc_ptr[d + 1], 1, sa, su);
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
                    atomicOr(err, 1);
                    det = (T)1;
                }
                di = B4<T>{a.d / det, -a.b / det, -a.c / det, a.a / det};
                store_b4(blk, Y.di0 + j, di);
            }
            dinv.a = __shfl(di.a, 0);
            dinv.b = __shfl(di.b, 0);
            dinv.c = __shfl(di.c, 0);
            dinv.d = __shfl(di.d, 0);
        }
        if (t < nd && t > 0) {
            const B4<T> l = mul(a, dinv);
            store_b4(blk, L0 + p, l);
            store_b4(blk, U0 + Y.nnzL + Y.lr_pos[p], l);
            store_b4(blk, U0 + p, u);
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_factor_level(DevSym Y, T *blk, Ctl *ctl, int l0, int count, int gate) {
    if (gate && !ctl->step) return;
    const int idx = blockIdx.x * 4 + (int)threadIdx.x / 64;
    if (idx >= count) return;
    factor_node(Y, blk, (l0 + idx), (int)threadIdx.x & 63, &ctl->singular);
}

template <typename T>
__global__ void __launch_bounds__(1024) k_factor_tail(DevSym Y, T *blk, Ctl *ctl, int lv0, int lv1, int gate) {
    if (gate && !ctl->step) return;
    if (Y.ck_lvl) {
        __shared__ B4<T> s_pa[kMaxChunks], s_pu[kMaxChunks], s_piv[kTailWidth];
        chunk_factor_levels<T, 4, sizeof(T) == 4 ? 4 : 2>(Y, blk, lv0, lv1, &ctl->singular, s_pa, s_pu, s_piv);
    } else {
        coop_factor_range<T, kCoopLanes, 4>(Y, blk, lv0, lv1, &ctl->singular);
    }
}

// Flat tail factorisation, first part: every tail destination minus its
// update pairs from sources below the tail (final after the subtree
// groups), G lanes per destination across the GPU.
template <typename T>
__global__ void __launch_bounds__(256) k_tail_factor_ext(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int ne, int gate) {
    if (gate && !ctl->step) return;
    constexpr int G = 8;
    const int E = (int)blockIdx.x * (256 / G) + (int)threadIdx.x / G, lane = (int)threadIdx.x % G;
    if (E >= ne) return;                                 // whole groups leave together
    const int L0 = Y.n, U0 = Y.n + Y.nnzL;
    const int p = F.de_p[E], q0 = F.de_q0[E], q1 = F.de_ext[E];
    B4<T> a{0, 0, 0, 0}, u{0, 0, 0, 0};
    if (p < 0) pairs_acc<T, false, 2>(Y, blk, q0 + lane, q1, G, a, u);
    else pairs_acc<T, true, 2>(Y, blk, q0 + lane, q1, G, a, u);
    a = group_sum<T, G>(a);
    u = group_sum<T, G>(u);
    if (lane == 0 && q1 > q0) {
        if (p < 0) {
            const int j = F.de_node[E];
            const B4<T> v = load_b4(blk, j);
            store_b4(blk, j, B4<T>{v.a - a.a, v.b - a.b, v.c - a.c, v.d - a.d});
        } else {
            const B4<T> lb = load_b4(blk, L0 + p), ub = load_b4(blk, U0 + p);
            store_b4(blk, L0 + p, B4<T>{lb.a - a.a, lb.b - a.b, lb.c - a.c, lb.d - a.d});
            store_b4(blk, U0 + p, B4<T>{ub.a - u.a, ub.b - u.b, ub.c - u.c, ub.d - u.d});
        }
    }
}

// Flat tail factorisation, second part: the tail levels in one workgroup
// (flat_factor_levels), the level's products in dynamic LDS, in FP32 also
// the staged blocks of steps of at most cap destinations after them (the
// FP64 blocks leave no registers for the requests made a step early; in
// FP32 two pairs per lane do, three would not).
template <typename T>
__global__ void __launch_bounds__(1024) k_factor_tail_flat(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int nsteps, int gate,
                                                           int cap) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    T *s_pa = reinterpret_cast<T *>(smem), *s_pu = s_pa + 4 * (size_t)F.max_pairs;
    if constexpr (sizeof(T) == 4)
        flat_factor_levels<T, 2>(Y, F, blk, 0, nsteps, &ctl->singular, s_pa, s_pu, s_pu + 4 * (size_t)F.max_col, cap);
    else
        flat_factor_levels<T, 1>(Y, F, blk, 0, nsteps, &ctl->singular, s_pa, s_pu);
}

// Levels below the tail as flat pair programs: workgroup g walks the level
// steps of its group of subtrees (g * part_levels .. + part_levels).
template <typename T>
__global__ void __launch_bounds__(256) k_factor_parts_flat(DevSym Y, DevFlat F, T *blk, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    T *s_pa = reinterpret_cast<T *>(smem);
    const int t0 = (int)blockIdx.x * Y.part_levels;
    flat_factor_levels<T, 2>(Y, F, blk, t0, t0 + Y.part_levels, &ctl->singular, s_pa,
                                s_pa + 4 * (size_t)F.max_pairs);
}

// Schur complement of the trailing dense chain: every destination is
// independent, so the pass is spread over many workgroups (16 lanes each).
template <typename T>
__global__ void __launch_bounds__(256) k_chain_schur(DevSym Y, T *blk, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    chain_schur<T, 16, 4>(Y, blk, (int)blockIdx.x * 16, (int)gridDim.x * 16);
}

// Gauss-Jordan inverse of the chain's Schur complement, in one workgroup (in FP32
// with the fused update: this solver's FP32 factor is refined in FP64).
template <typename T>
__global__ void __launch_bounds__(ChainTiles<T>::threads) k_chain_invert(DevSym Y, T *blk, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    __shared__ B4<T> lds[4 * kChainMax + 2];
    chain_invert<T, ChainTiles<T>::tr, ChainTiles<T>::tc, sizeof(T) == 4>(Y, blk, &ctl->singular, lds);
}

// ----------------------------------------------------- triangular solves
// 16 lanes per node; x holds the right-hand side and is overwritten.
template <typename T>
__device__ __forceinline__ void fwd_node(const DevSym &Y, const T *blk, double *x, int j, int sl) {
    double a0 = 0, a1 = 0;
    for (int q = Y.rs_ptr[j] + sl; q < Y.rs_ptr[j + 1]; q += 16) {
        B4<T> L = load_b4(blk, Y.n + 2 * Y.nnzL + q);
        const int k = Y.rs_k[q];
        const double x0 = x[2 * k], x1 = x[2 * k + 1];
        a0 += (double)L.a * x0 + (double)L.b * x1;
        a1 += (double)L.c * x0 + (double)L.d * x1;
    }
    a0 = xor_sum<16>(a0);
    a1 = xor_sum<16>(a1);
    if (sl == 0) {
        x[2 * j] -= a0;
        x[2 * j + 1] -= a1;
    }
}

template <typename T>
__device__ __forceinline__ void bwd_node(const DevSym &Y, const T *blk, double *x, int j, int sl) {
    double a0 = 0, a1 = 0;
    const int U0 = Y.n + Y.nnzL;
    for (int p = Y.colptr[j] + sl; p < Y.colptr[j + 1]; p += 16) {
        B4<T> U = load_b4(blk, U0 + p);
        const int i = Y.rowind[p];
        const double x0 = x[2 * i], x1 = x[2 * i + 1];
        a0 += (double)U.a * x0 + (double)U.b * x1;
        a1 += (double)U.c * x0 + (double)U.d * x1;
    }
    a0 = xor_sum<16>(a0);
    a1 = xor_sum<16>(a1);
    if (sl == 0) {
        const double y0 = x[2 * j] - a0, y1 = x[2 * j + 1] - a1;
        B4<T> D = load_b4(blk, Y.di0 + j);
        x[2 * j] = (double)D.a * y0 + (double)D.b * y1;
        x[2 * j + 1] = (double)D.c * y0 + (double)D.d * y1;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_fwd_level(DevSym Y, const T *blk, double *x, Ctl *ctl, int l0, int count, int gate) {
    if (gate && !ctl->step) return;
    const int idx = blockIdx.x * 16 + (int)threadIdx.x / 16;
    if (idx >= count) return;
    fwd_node(Y, blk, x, (l0 + idx), (int)threadIdx.x & 15);
}

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_level(DevSym Y, const T *blk, double *x, Ctl *ctl, int l0, int count, int gate) {
    if (gate && !ctl->step) return;
    const int idx = blockIdx.x * 16 + (int)threadIdx.x / 16;
    if (idx >= count) return;
    bwd_node(Y, blk, x, (l0 + idx), (int)threadIdx.x & 15);
}

// Levels below the tail in one launch: workgroup g walks the levels of its
// own group of subtrees (every update of a node comes from its descendants,
// so the groups never wait on each other).
// 1024 threads: 16 wavefronts (a node each) in the factorisation, 64 lane
// groups of 16 (a node each) in the sweeps.
template <typename T>
__global__ void __launch_bounds__(1024) k_factor_parts(DevSym Y, T *blk, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    const int w = (int)threadIdx.x / 64, lane = (int)threadIdx.x & 63, nl = Y.part_levels;
    const int nw = (int)blockDim.x / 64;
    const int *pp = Y.part_ptr + (size_t)blockIdx.x * nl;
    for (int l = 0; l < nl; l++) {
        const int s = pp[l], e = pp[l + 1];
        if (s == e) continue;
        for (int i = s + w; i < e; i += nw) factor_node(Y, blk, Y.part_nodes[i], lane, &ctl->singular);
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_fwd_parts(DevSym Y, const T *blk, double *x, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    const int g = (int)threadIdx.x / 16, lane = (int)threadIdx.x & 15, nl = Y.part_levels;
    const int ng = (int)blockDim.x / 16;
    const int *pp = Y.part_ptr + (size_t)blockIdx.x * nl;
    for (int l = 0; l < nl; l++) {
        const int s = pp[l], e = pp[l + 1];
        if (s == e) continue;
        for (int i = s + g; i < e; i += ng) fwd_node(Y, blk, x, Y.part_nodes[i], lane);
        __syncthreads();
    }
}

template <typename T>
__global__ void __launch_bounds__(1024) k_bwd_parts(DevSym Y, const T *blk, double *x, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    const int g = (int)threadIdx.x / 16, lane = (int)threadIdx.x & 15, nl = Y.part_levels;
    const int ng = (int)blockDim.x / 16;
    const int *pp = Y.part_ptr + (size_t)blockIdx.x * nl;
    for (int l = nl - 1; l >= 0; l--) {
        const int s = pp[l], e = pp[l + 1];
        if (s == e) continue;
        for (int i = s + g; i < e; i += ng) bwd_node(Y, blk, x, Y.part_nodes[i], lane);
        __syncthreads();
    }
}

// narrow levels: one wavefront per node.  The first entry of a wave's first
// node on the next level (index and factor block, independent of x) is
// loaded before the current level's arithmetic and barrier.
template <typename T>
__global__ void __launch_bounds__(1024) k_fwd_tail(DevSym Y, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate) {
    if (gate && !ctl->step) return;
    const int w = (int)threadIdx.x / 64, lane = (int)threadIdx.x & 63;
    const int LR0 = Y.n + 2 * Y.nnzL;
    auto fetch = [&](int l, int &k, B4<T> &L) {
        k = -1;
        const int s = Y.lvl_ptr[l];
        if (w < Y.lvl_ptr[l + 1] - s) {
            const int j = s + w, q = Y.rs_ptr[j] + lane;
            if (q < Y.rs_ptr[j + 1]) {
                k = Y.rs_k[q];
                L = load_b4(blk, LR0 + q);
            }
        }
    };
    int pk = -1;
    B4<T> pL{0, 0, 0, 0};
    if (lv0 < lv1) fetch(lv0, pk, pL);
    for (int l = lv0; l < lv1; l++) {
        int nk = -1;
        B4<T> nL{0, 0, 0, 0};
        if (l + 1 < lv1) fetch(l + 1, nk, nL);
        const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
        for (int idx = w; idx < c; idx += 16) {
            const int j = (s + idx);
            double a0 = 0, a1 = 0;
            int q = Y.rs_ptr[j] + lane;
            if (idx == w) {
                if (pk >= 0) {
                    a0 += (double)pL.a * x[2 * pk] + (double)pL.b * x[2 * pk + 1];
                    a1 += (double)pL.c * x[2 * pk] + (double)pL.d * x[2 * pk + 1];
                }
                q += 64;
            }
            for (; q < Y.rs_ptr[j + 1]; q += 64) {
                B4<T> L = load_b4(blk, LR0 + q);
                const int k = Y.rs_k[q];
                const double x0 = x[2 * k], x1 = x[2 * k + 1];
                a0 += (double)L.a * x0 + (double)L.b * x1;
                a1 += (double)L.c * x0 + (double)L.d * x1;
            }
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            if (lane == 0) {
                x[2 * j] -= a0;
                x[2 * j + 1] -= a1;
            }
        }
        __syncthreads();
        pk = nk;
        pL = nL;
    }
    __shared__ double s_r[2 * kChainMax];
    if (Y.ch_m > 0) chain_solve<T, 16>(Y, blk, x, s_r);
}

template <typename T>
__global__ void __launch_bounds__(1024) k_bwd_tail(DevSym Y, const T *blk, double *x, Ctl *ctl, int lv0, int lv1, int gate) {
    if (gate && !ctl->step) return;
    const int w = (int)threadIdx.x / 64, lane = (int)threadIdx.x & 63;
    const int U0 = Y.n + Y.nnzL;
    auto fetch = [&](int l, int &r, B4<T> &U, B4<T> &D) {
        r = -1;
        const int s = Y.lvl_ptr[l];
        if (w < Y.lvl_ptr[l + 1] - s) {
            const int j = s + w, p = Y.colptr[j] + lane;
            if (lane == 0) D = load_b4(blk, Y.di0 + j);
            if (p < Y.colptr[j + 1]) {
                r = Y.rowind[p];
                U = load_b4(blk, U0 + p);
            }
        }
    };
    int pr = -1;
    B4<T> pU{0, 0, 0, 0}, pD{0, 0, 0, 0};
    if (lv0 < lv1) fetch(lv1 - 1, pr, pU, pD);
    for (int l = lv1 - 1; l >= lv0; l--) {
        int nr = -1;
        B4<T> nU{0, 0, 0, 0}, nD{0, 0, 0, 0};
        if (l > lv0) fetch(l - 1, nr, nU, nD);
        const int s = Y.lvl_ptr[l], c = Y.lvl_ptr[l + 1] - s;
        for (int idx = w; idx < c; idx += 16) {
            const int j = (s + idx);
            double a0 = 0, a1 = 0;
            int p = Y.colptr[j] + lane;
            if (idx == w) {
                if (pr >= 0) {
                    a0 += (double)pU.a * x[2 * pr] + (double)pU.b * x[2 * pr + 1];
                    a1 += (double)pU.c * x[2 * pr] + (double)pU.d * x[2 * pr + 1];
                }
                p += 64;
            }
            for (; p < Y.colptr[j + 1]; p += 64) {
                B4<T> U = load_b4(blk, U0 + p);
                const int i = Y.rowind[p];
                const double x0 = x[2 * i], x1 = x[2 * i + 1];
                a0 += (double)U.a * x0 + (double)U.b * x1;
                a1 += (double)U.c * x0 + (double)U.d * x1;
            }
            a0 = wave_sum(a0);
            a1 = wave_sum(a1);
            if (lane == 0) {
                const double y0 = x[2 * j] - a0, y1 = x[2 * j + 1] - a1;
                const B4<T> D = (idx == w) ? pD : load_b4(blk, Y.di0 + j);
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

// ------------------------------------------ tail sweeps with x in LDS
// The tail's unknowns live in LDS for the whole level walk, so a level costs
// a barrier and LDS round trips instead of dependent reads through the
// cache hierarchy.  Entries on nodes outside the tail (final before the
// sweep) are summed first by a grid-wide pass, 16 lanes per tail node.

__device__ __forceinline__ void sum16(double &a0, double &a1) {
    a0 = xor_sum<16>(a0);
    a1 = xor_sum<16>(a1);
}

// forward rows of the tail nodes: x_j -= sum over entries on nodes < t0
template <typename T>
__global__ void __launch_bounds__(256) k_tail_fwd_ext(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    const int a = (int)blockIdx.x * 16 + (int)threadIdx.x / 16, lane = (int)threadIdx.x & 15;
    if (a >= D.t1 - D.t0) return;                        // whole 16-lane groups leave together
    const int LR0 = Y.n + 2 * Y.nnzL;
    double a0 = 0, a1 = 0;
    for (int e = D.fx_ptr[a] + lane; e < D.fx_ptr[a + 1]; e += 16) {
        const B4<T> L = load_b4(blk, LR0 + D.fx_q[e]);
        const int k = D.fx_k[e];
        const double x0 = x[2 * k], x1 = x[2 * k + 1];
        a0 += (double)L.a * x0 + (double)L.b * x1;
        a1 += (double)L.c * x0 + (double)L.d * x1;
    }
    sum16(a0, a1);
    if (lane == 0) {
        const int j = D.t0 + a;
        x[2 * j] -= a0;
        x[2 * j + 1] -= a1;
    }
}

// backward columns of the tail nodes: x_j -= sum over entries on the dense
// top (nodes >= t1, solved by the forward sweep's end)
template <typename T>
__global__ void __launch_bounds__(256) k_tail_bwd_ext(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    const int a = (int)blockIdx.x * 16 + (int)threadIdx.x / 16, lane = (int)threadIdx.x & 15;
    if (a >= D.t1 - D.t0) return;
    const int U0 = Y.n + Y.nnzL;
    double a0 = 0, a1 = 0;
    for (int e = D.bx_ptr[a] + lane; e < D.bx_ptr[a + 1]; e += 16) {
        const B4<T> U = load_b4(blk, U0 + D.bx_p[e]);
        const int i = D.bx_i[e];
        const double x0 = x[2 * i], x1 = x[2 * i + 1];
        a0 += (double)U.a * x0 + (double)U.b * x1;
        a1 += (double)U.c * x0 + (double)U.d * x1;
    }
    sum16(a0, a1);
    if (lane == 0) {
        const int j = D.t0 + a;
        x[2 * j] -= a0;
        x[2 * j + 1] -= a1;
    }
}

// LDS layout of one sweep (tail_sweep_lds gives its size): x of the tail
// nodes, [inverse pivots], [entry blocks], level bounds, entry pointers,
// entry local indices, [entry slots]
template <typename T, bool STAGED>
struct TailLds {
    double2 *x;
    B4<T> *dg, *eb;
    int *lp, *ptr, *ek, *es;
    __device__ TailLds(char *smem, int nt, int nlt, int ne, bool backward) {
        x = reinterpret_cast<double2 *>(smem);
        B4<T> *b = reinterpret_cast<B4<T> *>(x + nt);
        dg = b;
        if (backward) b += nt;
        eb = b;
        if (STAGED) b += ne;
        lp = reinterpret_cast<int *>(b);
        ptr = lp + nlt + 1;
        ek = ptr + nt + 1;
        es = ek + ne;
    }
};

// levels [lv0, lv1) of the tail, forward: 16 lanes per node, then the dense
// top (chain_solve) from the finished x
template <typename T, bool STAGED>
__global__ void __launch_bounds__(1024) k_fwd_tail_lds(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int lv0,
                                                       int lv1, int gate) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    const int nt = D.t1 - D.t0, nlt = lv1 - lv0, ne = D.nfi, t0 = D.t0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / 16, lane = tid & 15, ng = bd / 16;
    const int LR0 = Y.n + 2 * Y.nnzL;
    TailLds<T, STAGED> S(smem, nt, nlt, ne, false);
    for (int a = tid; a < nt; a += bd) S.x[a] = make_double2(x[2 * (t0 + a)], x[2 * (t0 + a) + 1]);
    for (int l = tid; l <= nlt; l += bd) S.lp[l] = Y.lvl_ptr[lv0 + l] - t0;
    for (int a = tid; a <= nt; a += bd) S.ptr[a] = D.fi_ptr[a];
    for (int e = tid; e < ne; e += bd) {
        S.ek[e] = D.fi_k[e];
        if constexpr (STAGED) store_b4(reinterpret_cast<T *>(S.eb), e, load_b4(blk, LR0 + D.fi_q[e]));
        else S.es[e] = LR0 + D.fi_q[e];
    }
    __syncthreads();
    for (int l = 0; l < nlt; l++) {
        const int s = S.lp[l], c = S.lp[l + 1] - s;
        for (int idx = grp; idx < c; idx += ng) {
            const int a = s + idx;
            double a0 = 0, a1 = 0;
            for (int e = S.ptr[a] + lane; e < S.ptr[a + 1]; e += 16) {
                const B4<T> L = load_b4(STAGED ? reinterpret_cast<const T *>(S.eb) : blk, STAGED ? e : S.es[e]);
                const double2 xk = S.x[S.ek[e]];
                a0 += (double)L.a * xk.x + (double)L.b * xk.y;
                a1 += (double)L.c * xk.x + (double)L.d * xk.y;
            }
            sum16(a0, a1);
            if (lane == 0) S.x[a] = make_double2(S.x[a].x - a0, S.x[a].y - a1);
        }
        __syncthreads();
    }
    for (int a = tid; a < nt; a += bd) {
        x[2 * (t0 + a)] = S.x[a].x;
        x[2 * (t0 + a) + 1] = S.x[a].y;
    }
    __syncthreads();
    __shared__ double s_r[2 * kChainMax];
    if (Y.ch_m > 0) chain_solve<T, 16>(Y, blk, x, s_r);
}

// levels [lv0, lv1) of the tail, backward, after k_tail_bwd_ext
template <typename T, bool STAGED>
__global__ void __launch_bounds__(1024) k_bwd_tail_lds(DevSym Y, DevTail D, const T *blk, double *x, Ctl *ctl, int lv0,
                                                       int lv1, int gate) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    const int nt = D.t1 - D.t0, nlt = lv1 - lv0, ne = D.nbi, t0 = D.t0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / 16, lane = tid & 15, ng = bd / 16;
    const int U0 = Y.n + Y.nnzL;
    TailLds<T, STAGED> S(smem, nt, nlt, ne, true);
    for (int a = tid; a < nt; a += bd) {
        S.x[a] = make_double2(x[2 * (t0 + a)], x[2 * (t0 + a) + 1]);
        store_b4(reinterpret_cast<T *>(S.dg), a, load_b4(blk, Y.di0 + t0 + a));
    }
    for (int l = tid; l <= nlt; l += bd) S.lp[l] = Y.lvl_ptr[lv0 + l] - t0;
    for (int a = tid; a <= nt; a += bd) S.ptr[a] = D.bi_ptr[a];
    for (int e = tid; e < ne; e += bd) {
        S.ek[e] = D.bi_i[e];
        if constexpr (STAGED) store_b4(reinterpret_cast<T *>(S.eb), e, load_b4(blk, U0 + D.bi_p[e]));
        else S.es[e] = U0 + D.bi_p[e];
    }
    __syncthreads();
    for (int l = nlt - 1; l >= 0; l--) {
        const int s = S.lp[l], c = S.lp[l + 1] - s;
        for (int idx = grp; idx < c; idx += ng) {
            const int a = s + idx;
            double a0 = 0, a1 = 0;
            for (int e = S.ptr[a] + lane; e < S.ptr[a + 1]; e += 16) {
                const B4<T> U = load_b4(STAGED ? reinterpret_cast<const T *>(S.eb) : blk, STAGED ? e : S.es[e]);
                const double2 xi = S.x[S.ek[e]];
                a0 += (double)U.a * xi.x + (double)U.b * xi.y;
                a1 += (double)U.c * xi.x + (double)U.d * xi.y;
            }
            sum16(a0, a1);
            if (lane == 0) {
                const double y0 = S.x[a].x - a0, y1 = S.x[a].y - a1;
                const B4<T> P = load_b4(reinterpret_cast<const T *>(S.dg), a);
                S.x[a] = make_double2((double)P.a * y0 + (double)P.b * y1, (double)P.c * y0 + (double)P.d * y1);
            }
        }
        __syncthreads();
    }
    for (int a = tid; a < nt; a += bd) {
        x[2 * (t0 + a)] = S.x[a].x;
        x[2 * (t0 + a) + 1] = S.x[a].y;
    }
}

// ----------------------------------- subtree groups' sweeps with x in LDS
// Workgroup g holds its group's unknowns, its entries' local indices and
// factor blocks in LDS for the whole sweep (a group's forward rows only
// reach its own nodes; the backward columns' entries above the group are
// final before its sweep and are subtracted first).  16 lanes per node, one
// barrier per level.  LDS: x [max_nodes], then (backward: the inverse
// pivots [max_nodes]), the entry blocks and their local indices.
template <typename T>
__global__ void __launch_bounds__(256) k_fwd_parts_lds(DevSym Y, DevParts P, const T *blk, double *x, Ctl *ctl,
                                                       int gate) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    const int T_ = Y.part_levels, g = (int)blockIdx.x;
    const int u0 = Y.part_ptr[(size_t)g * T_], u1 = Y.part_ptr[(size_t)(g + 1) * T_], nn = u1 - u0;
    const int e0 = P.fi_ptr[u0], ne = P.fi_ptr[u1] - e0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / 16, lane = tid & 15, ng = bd / 16;
    const int LR0 = Y.n + 2 * Y.nnzL;
    double2 *xl = reinterpret_cast<double2 *>(smem);
    T *eb = reinterpret_cast<T *>(xl + P.max_nodes);
    int *ek = reinterpret_cast<int *>(eb + 4 * (size_t)P.max_fi);
    for (int a = tid; a < nn; a += bd) {
        const int j = Y.part_nodes[u0 + a];
        xl[a] = make_double2(x[2 * j], x[2 * j + 1]);
    }
    for (int e = tid; e < ne; e += bd) {
        ek[e] = P.fi_k[e0 + e];
        store_b4(eb, e, load_b4(blk, LR0 + P.fi_q[e0 + e]));
    }
    __syncthreads();
    for (int l = 0; l < T_; l++) {
        const int a0 = Y.part_ptr[(size_t)g * T_ + l] - u0, a1 = Y.part_ptr[(size_t)g * T_ + l + 1] - u0;
        for (int a = a0 + grp; a < a1; a += ng) {
            double s0 = 0, s1 = 0;
            for (int e = P.fi_ptr[u0 + a] - e0 + lane; e < P.fi_ptr[u0 + a + 1] - e0; e += 16) {
                const B4<T> L = load_b4(eb, e);
                const double2 xk = xl[ek[e]];
                s0 += (double)L.a * xk.x + (double)L.b * xk.y;
                s1 += (double)L.c * xk.x + (double)L.d * xk.y;
            }
            sum16(s0, s1);
            if (lane == 0) xl[a] = make_double2(xl[a].x - s0, xl[a].y - s1);
        }
        __syncthreads();
    }
    for (int a = tid; a < nn; a += bd) {
        const int j = Y.part_nodes[u0 + a];
        x[2 * j] = xl[a].x;
        x[2 * j + 1] = xl[a].y;
    }
}

template <typename T>
__global__ void __launch_bounds__(256) k_bwd_parts_lds(DevSym Y, DevParts P, const T *blk, double *x, Ctl *ctl,
                                                       int gate) {
    if (gate && !ctl->step) return;
    extern __shared__ __align__(16) char smem[];
    const int T_ = Y.part_levels, g = (int)blockIdx.x;
    const int u0 = Y.part_ptr[(size_t)g * T_], u1 = Y.part_ptr[(size_t)(g + 1) * T_], nn = u1 - u0;
    const int e0 = P.bi_ptr[u0], ne = P.bi_ptr[u1] - e0;
    const int tid = (int)threadIdx.x, bd = (int)blockDim.x, grp = tid / 16, lane = tid & 15, ng = bd / 16;
    const int U0 = Y.n + Y.nnzL;
    double2 *xl = reinterpret_cast<double2 *>(smem);
    T *dg = reinterpret_cast<T *>(xl + P.max_nodes);
    T *eb = dg + 4 * (size_t)P.max_nodes;
    int *ek = reinterpret_cast<int *>(eb + 4 * (size_t)P.max_bi);
    // the entries above the group first (their x are final; a node has few):
    // a thread per node, with its inverse pivot
    for (int a = tid; a < nn; a += bd) {
        const int j = Y.part_nodes[u0 + a];
        double s0 = 0, s1 = 0;
        for (int e = P.bx_ptr[u0 + a]; e < P.bx_ptr[u0 + a + 1]; e++) {
            const B4<T> U = load_b4(blk, U0 + P.bx_p[e]);
            const int i = P.bx_i[e];
            const double x0 = x[2 * i], x1 = x[2 * i + 1];
            s0 += (double)U.a * x0 + (double)U.b * x1;
            s1 += (double)U.c * x0 + (double)U.d * x1;
        }
        xl[a] = make_double2(x[2 * j] - s0, x[2 * j + 1] - s1);
        store_b4(dg, a, load_b4(blk, Y.di0 + j));
    }
    for (int e = tid; e < ne; e += bd) {
        ek[e] = P.bi_i[e0 + e];
        store_b4(eb, e, load_b4(blk, U0 + P.bi_p[e0 + e]));
    }
    __syncthreads();
    for (int l = T_ - 1; l >= 0; l--) {
        const int a0 = Y.part_ptr[(size_t)g * T_ + l] - u0, a1 = Y.part_ptr[(size_t)g * T_ + l + 1] - u0;
        for (int a = a0 + grp; a < a1; a += ng) {
            double s0 = 0, s1 = 0;
            for (int e = P.bi_ptr[u0 + a] - e0 + lane; e < P.bi_ptr[u0 + a + 1] - e0; e += 16) {
                const B4<T> U = load_b4(eb, e);
                const double2 xi = xl[ek[e]];
                s0 += (double)U.a * xi.x + (double)U.b * xi.y;
                s1 += (double)U.c * xi.x + (double)U.d * xi.y;
            }
            sum16(s0, s1);
            if (lane == 0) {
                const double y0 = xl[a].x - s0, y1 = xl[a].y - s1;
                const B4<T> D = load_b4(dg, a);
                xl[a] = make_double2((double)D.a * y0 + (double)D.b * y1, (double)D.c * y0 + (double)D.d * y1);
            }
        }
        __syncthreads();
    }
    for (int a = tid; a < nn; a += bd) {
        const int j = Y.part_nodes[u0 + a];
        x[2 * j] = xl[a].x;
        x[2 * j + 1] = xl[a].y;
    }
}

// ------------------------------------------------------- branch flows
// Flows as the reference reports them (SpCalcFlux,
// CPU/esparso/potencia/spCalcFluxPQ.cpp:3-36): for the direction k -> m,
// a = phif(k,m) + th_k - th_m and G + jB = Ybus(k,m),
//   P = V_k^2 G - V_k V_m G cos a - V_k V_m B sin a
//   Q = -V_k^2 (B + bshf(k,m)) + V_k V_m B cos a - V_k V_m G sin a
// evaluated left to right without contraction.  One thread per branch;
// out = [Pdp | Qdp | Ppd | Qpd], nL each.
__device__ __forceinline__ void flow_dir(const DevNet &N, const DevState &S, int k, int m, int slot,
                                         double phi, double bsh, double &P, double &Q) {
#pragma clang fp contract(off)
    const double2 y = N.y[slot];
    const double G = y.x, B = y.y;
    const double a = phi + S.th[k] - S.th[m];
    double sn, cs;
    sincos(a, &sn, &cs);
    const double vkk = S.V[k] * S.V[k], vkm = S.V[k] * S.V[m];
    P = vkk * G - vkm * G * cs - vkm * B * sn;
    Q = -vkk * (B + bsh) + vkm * B * cs - vkm * G * sn;
}

__global__ void k_flows(DevNet N, DevState S, const FlowBranch *br, int nL, double *out) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= nL) return;
    const FlowBranch b = br[l];
    double P, Q;
    flow_dir(N, S, b.from, b.to, b.slot_ft, b.phi_ft, b.bsh_ft, P, Q);
    out[l] = P;
    out[nL + l] = Q;
    flow_dir(N, S, b.to, b.from, b.slot_tf, b.phi_tf, b.bsh_tf, P, Q);
    out[2 * nL + l] = P;
    out[3 * nL + l] = Q;
}

// --------------------------------------------------------------- vectors
__global__ void k_copy(double *dst, const double *src, int n, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

__global__ void k_axpy(double *x, const double *d, int n, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

__global__ void k_fill(double *x, size_t n, double v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

// FP64 refinement residual r = g - J x, with J x formed from the complex
// voltages (dE_m = E_m (dV_m / V_m + j dtheta_m), dS_k = dE_k conj(I_k) +
// E_k conj(sum_m Y_km dE_m)); a bus without a Q equation contributes dV.
template <int LPR>
__global__ void __launch_bounds__(256) k_residual(DevNet N, DevState S, int gate) {
    if (gate && !S.ctl->step) return;
    constexpr int ROWS = 256 / LPR;
    const int row = blockIdx.x * ROWS + (int)threadIdx.x / LPR;
    const int lane = (int)threadIdx.x % LPR;
    double ar = 0.0, ai = 0.0;
    const bool active = row < N.nB && row != N.slack;
    if (active) {
        for (int e = N.rowptr[row] + lane; e < N.rowptr[row + 1]; e += LPR) {
            const int m = N.colind[e];
            if (m == N.slack) continue;
            const int jm = N.node_of_bus[m];
            const double dth = S.x[2 * jm];
            const double dv = S.isq[m] ? S.x[2 * jm + 1] : 0.0;
            const double2 Em = S.E[m];
            const double f = dv / S.V[m];
            const double dr = Em.x * f - Em.y * dth, di = Em.y * f + Em.x * dth;
            const double2 y = N.y[e];
            ar += y.x * dr - y.y * di;
            ai += y.x * di + y.y * dr;
        }
    }
    ar = xor_sum<LPR>(ar);
    ai = xor_sum<LPR>(ai);
    if (active && lane == 0) {
        const int j = N.node_of_bus[row];
        const double dth = S.x[2 * j];
        const bool q = S.isq[row];
        const double dv = q ? S.x[2 * j + 1] : 0.0;
        const double2 Ek = S.E[row], Ik = S.I[row];
        const double f = dv / S.V[row];
        const double dr = Ek.x * f - Ek.y * dth, di = Ek.y * f + Ek.x * dth;   // dE_k
        // dS = dE conj(I) + E conj(Y dE)
        const double sr = dr * Ik.x + di * Ik.y + Ek.x * ar + Ek.y * ai;
        const double si = di * Ik.x - dr * Ik.y + Ek.y * ar - Ek.x * ai;
        S.r[2 * j] = S.g[2 * j] - sr;
        S.r[2 * j + 1] = q ? S.g[2 * j + 1] - si : -S.x[2 * j + 1];
    }
}

// ----------------------------------------------------------------- update
__global__ void __launch_bounds__(256) k_update(DevNet N, DevState S, int gate) {
    if (gate && !S.ctl->step) return;
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (k < N.nB && k != N.slack) {
        const int j = N.node_of_bus[k];
        const double dth = S.x[2 * j];
        double v = S.V[k], th = S.th[k] + dth;
        m = fabs(dth);
        if (S.isq[k]) {
            const double dv = S.x[2 * j + 1];
            v += dv;
            m = dmax_nan(m, fabs(dv));
        }
        S.V[k] = v;
        S.th[k] = th;
        double s, c;
        sincos(th, &s, &c);
        S.E[k] = make_double2(v * c, v * s);
    }
    m = wave_max(m);
    if ((threadIdx.x & 63) == 0) atomicMax(&S.ctl->maxDx_bits, double_as_u64(m));
}

// ------------------------------------------------------------ dense path
// metodo denso: the block Jacobian scattered into a dense column-major
// matrix (node order, 2n x 2n) and factored by one workgroup with partial
// pivoting; the right-hand side is eliminated alongside.
__global__ void k_dense_scatter(DevSym Y, const double *blk, double *A, int ld, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= Y.nslot) return;
    int r, c;
    if (s < Y.n) {
        r = c = s;
    } else if (s < Y.n + Y.nnzL) {
        int p = s - Y.n;
        r = Y.rowind[p];
        c = Y.pcol[p];
    } else {
        int p = s - Y.n - Y.nnzL;
        r = Y.pcol[p];
        c = Y.rowind[p];
    }
    B4<double> v = load_b4(blk, s);
    A[(size_t)(2 * c) * ld + 2 * r] = v.a;
    A[(size_t)(2 * c + 1) * ld + 2 * r] = v.b;
    A[(size_t)(2 * c) * ld + 2 * r + 1] = v.c;
    A[(size_t)(2 * c + 1) * ld + 2 * r + 1] = v.d;
}

__global__ void __launch_bounds__(1024) k_dense(double *A, double *b, int n, Ctl *ctl, int gate) {
    if (gate && !ctl->step) return;
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ int s_piv;
    const int tid = threadIdx.x, w = tid / 64, lane = tid & 63;
    for (int k = 0; k < n; k++) {
        double best = -1.0;
        int bi = k;
        for (int i = k + tid; i < n; i += 1024) {
            double a = fabs(A[(size_t)k * n + i]);
            if (a > best) { best = a; bi = i; }
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            double ob = __shfl_xor(best, off);
            int oi = __shfl_xor(bi, off);
            if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (lane == 0) { sv[w] = best; si[w] = bi; }
        __syncthreads();
        if (tid == 0) {
            double bb = sv[0];
            int ii = si[0];
            for (int q = 1; q < 16; q++)
                if (sv[q] > bb || (sv[q] == bb && si[q] < ii)) { bb = sv[q]; ii = si[q]; }
            s_piv = (bb > 0.0 && isfinite(bb)) ? ii : -1;
        }
        __syncthreads();
        const int p = s_piv;
        if (p < 0) {
            if (tid == 0) atomicOr(&ctl->singular, 1);
            return;   // uniform: every thread read the same s_piv
        }
        if (p != k) {
            for (int c = tid; c < n; c += 1024) {
                double t = A[(size_t)c * n + k];
                A[(size_t)c * n + k] = A[(size_t)c * n + p];
                A[(size_t)c * n + p] = t;
            }
            if (tid == 0) { double t = b[k]; b[k] = b[p]; b[p] = t; }
        }
        __syncthreads();
        const double piv = A[(size_t)k * n + k];
        for (int i = k + 1 + tid; i < n; i += 1024) A[(size_t)k * n + i] /= piv;
        __syncthreads();
        const double bk = b[k];
        for (int i = k + 1 + tid; i < n; i += 1024) {
            const double li = A[(size_t)k * n + i];
            for (int c = k + 1; c < n; c++) A[(size_t)c * n + i] -= li * A[(size_t)c * n + k];
            b[i] -= li * bk;
        }
        __syncthreads();
    }
    for (int k = n - 1; k >= 0; k--) {
        if (tid == 0) b[k] /= A[(size_t)k * n + k];
        __syncthreads();
        const double bk = b[k];
        for (int i = tid; i < k; i += 1024) b[i] -= A[(size_t)k * n + i] * bk;
        __syncthreads();
    }
}

// explicit instantiations
template __global__ void k_pq<4>(DevNet, DevState, int);
template __global__ void k_residual<4>(DevNet, DevState, int);
template __global__ void k_jfill<double>(DevNet, DevState, DevSym, double *, int);
template __global__ void k_jfill<float>(DevNet, DevState, DevSym, float *, int);
template __global__ void k_factor_level<double>(DevSym, double *, Ctl *, int, int, int);
template __global__ void k_factor_level<float>(DevSym, float *, Ctl *, int, int, int);
template __global__ void k_factor_parts<double>(DevSym, double *, Ctl *, int);
template __global__ void k_factor_parts<float>(DevSym, float *, Ctl *, int);
template __global__ void k_fwd_parts<double>(DevSym, const double *, double *, Ctl *, int);
template __global__ void k_fwd_parts<float>(DevSym, const float *, double *, Ctl *, int);
template __global__ void k_bwd_parts<double>(DevSym, const double *, double *, Ctl *, int);
template __global__ void k_bwd_parts<float>(DevSym, const float *, double *, Ctl *, int);
template __global__ void k_chain_schur<double>(DevSym, double *, Ctl *, int);
template __global__ void k_chain_schur<float>(DevSym, float *, Ctl *, int);
template __global__ void k_chain_invert<double>(DevSym, double *, Ctl *, int);
template __global__ void k_chain_invert<float>(DevSym, float *, Ctl *, int);
template __global__ void k_factor_tail<double>(DevSym, double *, Ctl *, int, int, int);
template __global__ void k_factor_tail<float>(DevSym, float *, Ctl *, int, int, int);
template __global__ void k_tail_factor_ext<double>(DevSym, DevFlat, double *, Ctl *, int, int);
template __global__ void k_tail_factor_ext<float>(DevSym, DevFlat, float *, Ctl *, int, int);
template __global__ void k_factor_tail_flat<double>(DevSym, DevFlat, double *, Ctl *, int, int, int);
template __global__ void k_factor_tail_flat<float>(DevSym, DevFlat, float *, Ctl *, int, int, int);
template __global__ void k_factor_parts_flat<double>(DevSym, DevFlat, double *, Ctl *, int);
template __global__ void k_factor_parts_flat<float>(DevSym, DevFlat, float *, Ctl *, int);
template __global__ void k_fwd_parts_lds<double>(DevSym, DevParts, const double *, double *, Ctl *, int);
template __global__ void k_fwd_parts_lds<float>(DevSym, DevParts, const float *, double *, Ctl *, int);
template __global__ void k_bwd_parts_lds<double>(DevSym, DevParts, const double *, double *, Ctl *, int);
template __global__ void k_bwd_parts_lds<float>(DevSym, DevParts, const float *, double *, Ctl *, int);
template __global__ void k_fwd_level<double>(DevSym, const double *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_level<float>(DevSym, const float *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_level<double>(DevSym, const double *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_level<float>(DevSym, const float *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_tail<double>(DevSym, const double *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_tail<float>(DevSym, const float *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail<double>(DevSym, const double *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail<float>(DevSym, const float *, double *, Ctl *, int, int, int);
template __global__ void k_tail_fwd_ext<double>(DevSym, DevTail, const double *, double *, Ctl *, int);
template __global__ void k_tail_fwd_ext<float>(DevSym, DevTail, const float *, double *, Ctl *, int);
template __global__ void k_tail_bwd_ext<double>(DevSym, DevTail, const double *, double *, Ctl *, int);
template __global__ void k_tail_bwd_ext<float>(DevSym, DevTail, const float *, double *, Ctl *, int);
template __global__ void k_fwd_tail_lds<double, true>(DevSym, DevTail, const double *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_tail_lds<double, false>(DevSym, DevTail, const double *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_tail_lds<float, true>(DevSym, DevTail, const float *, double *, Ctl *, int, int, int);
template __global__ void k_fwd_tail_lds<float, false>(DevSym, DevTail, const float *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail_lds<double, true>(DevSym, DevTail, const double *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail_lds<double, false>(DevSym, DevTail, const double *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail_lds<float, true>(DevSym, DevTail, const float *, double *, Ctl *, int, int, int);
template __global__ void k_bwd_tail_lds<float, false>(DevSym, DevTail, const float *, double *, Ctl *, int, int, int);

}  // namespace pf
