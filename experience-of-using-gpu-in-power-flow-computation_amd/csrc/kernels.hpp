// Device kernels of libpf (gfx950, wave64).  Shared by the single-network
// solver (solver.hip) and the stage entry points.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "pf_internal.hpp"

namespace pf {

constexpr double kFolga = 1e-4;   // tol_limInjReat (CPU/global.cpp:11)

// NR control block, device resident.  Accumulators are reset by k_control
// after being read, so one captured iteration can be replayed many times.
struct Ctl {
    unsigned long long maxG_bits;    // max |g| of this iteration (bit pattern of a double >= 0)
    unsigned long long maxDx_bits;   // max |dx| of the last step
    unsigned int simple_bad;         // some PV bus outside [Qmin-folga, Qmax+folga]
    int it;                          // passes completed (steps + zero steps)
    int done;                        // converged
    int step;                        // this pass takes a Newton step
    int post;                        // this pass is a switch taken at convergence (zero step)
    int flips;                       // PV<->PQ flips so far
    int singular;                    // a zero/non-finite pivot was met
    int max_iter;
    double tol;
    double maxG;                     // last max |g|
};

// Topology and per-bus data of one network on the device.
struct DevNet {
    int nB, nL, slack, nPV, n;       // n = nB - 1 block unknowns
    const int *rowptr, *colind, *coorow, *ydiag;     // Ybus CSR
    const int *yc_ptr, *yc;                          // Ybus assembly contributions
    const double *br_r, *br_x, *br_b, *br_tr, *br_ti;
    const double *gsh, *bsh;
    double2 *y;                                      // Ybus values
    const int *node_of_bus;                          // elimination position, -1 = slack
    const int *bus_of_node;
    const int *pv_bus, *pv_of_bus;                   // PV list / bus -> PV index (-1)
    const double *qmin, *qmax, *vset;
};

struct DevState {
    double *V, *th;
    double2 *E, *I;                  // E = V e^{j theta}, I = Y E
    double *P, *Q;
    const double *Pliq, *Qliq, *Qload;
    double *QliqLim;
    unsigned char *isq;              // bus carries a Q equation (PQ, or PV switched to PQ)
    unsigned char *limQ;             // per PV
    double *g;                       // mismatch, 2 per node (node order)
    double *x;                       // solution / step, 2 per node
    double *r;                       // refinement residual / correction, 2 per node
    Ctl *ctl;
};

// trailing dense chain of the single-network solver: at most kChainMax
// nodes (pf_internal.hpp), S held in TR x TC register tiles of blocks by one
// workgroup of 1024 threads
template <typename T> struct ChainTiles;
template <> struct ChainTiles<float> { static constexpr int threads = 1024, tr = 4, tc = 4, cap = kChainMax; };
template <> struct ChainTiles<double> { static constexpr int threads = 1024, tr = 4, tc = 3, cap = kChainMaxF64; };
template <typename T>
constexpr bool chain_tiles_fit() {
    constexpr int m = ChainTiles<T>::cap;
    return m <= kChainMax && ((m + ChainTiles<T>::tr - 1) / ChainTiles<T>::tr) *
                                     ((m + ChainTiles<T>::tc - 1) / ChainTiles<T>::tc) <= ChainTiles<T>::threads;
}
static_assert(chain_tiles_fit<float>() && chain_tiles_fit<double>(), "dense chain does not fit the registers");
// lanes per destination block in the cooperative narrow-level factorisation
constexpr int kCoopLanes = 4;
// most chunks of update pairs on one tail level of the single-network solver
constexpr int kMaxChunks = 1024;

// Per-branch constants of the reported flows: the Ybus slots of (from, to)
// and (to, from), and the reference's phif / bshf lookups in both directions.
struct FlowBranch {
    int from, to, slot_ft, slot_tf;
    double phi_ft, phi_tf, bsh_ft, bsh_tf;
};

// Symbolic factor program on the device.
struct DevSym {
    int n, nnzL, nslot, nlev;
    const int *colptr, *rowind, *pcol;      // pcol[p] = column of L position p
    const int *lvl_ptr, *lvl_nodes;
    const int *fd_ptr, *fc_ptr, *fc_a, *fc_b;
    const int *rs_ptr, *rs_k, *rs_p;
    const int *lr_pos;                      // L position -> row-ordered copy (block n + 2 nnzL + q)
    // trailing dense chain [ch_c0, n) on levels [nlev_s, nlev); ch_m = 0: none
    int ch_m, ch_c0, nlev_s;
    const int *ch_split;
    int di0;                                // block offset of the inverse pivots
    // groups of whole subtrees below the tail, one workgroup each
    int n_parts, part_levels;
    const int *part_ptr, *part_nodes;
    // tail levels as chunks of update pairs (null: not chunked); te0 =
    // ld_ptr[tail_start], the first flattened destination of the tail
    const int *ck_lvl, *ck_q0, *ck_q1, *te_node, *te_p, *te_ck;
    int te0, tail_start;
    const int *slot_e;
    const int *ld_ptr, *ld_node, *ld_d;     // flattened destinations per level
};

// The tail's triangular sweeps with its unknowns in LDS (TailSweeps on the
// host): nt = t1 - t0 tail nodes, nfi forward and nbi backward entries on
// tail nodes.
struct DevTail {
    int t0, t1, nfi, nbi;
    const int *fx_ptr, *fx_k, *fx_q, *fi_ptr, *fi_k, *fi_q;
    const int *bx_ptr, *bx_i, *bx_p, *bi_ptr, *bi_i, *bi_p;
};

// A flat pair program on the device (FlatTail on the host): level steps,
// each its destinations [de_ptr[t], de_ptr[t+1]) and pairs
// [fl_ptr[t], fl_ptr[t+1]) (fl_s: a pair's staged slots in the previous
// step, or -1); lv0: the first level of a level-by-level program.
struct DevFlat {
    int lv0, max_pairs, max_col, de_sentinel;
    const int *fl_ptr, *fl_split, *fl_a, *fl_b, *fl_s, *de_ptr;
    const int *de_node, *de_p, *de_q0, *de_ext, *de_lo, *de_hi, *de_dlo, *de_dhi, *de_lr, *st_g;
};

// The subtree groups' sweeps with their unknowns in LDS (PartSweeps on the
// host); the group ranges are the DevSym's part_ptr / part_nodes.
struct DevParts {
    int max_nodes, max_fi, max_bi;
    const int *fi_ptr, *fi_k, *fi_q, *bi_ptr, *bi_i, *bi_p, *bx_ptr, *bx_i, *bx_p;
};

// bytes of dynamic LDS of a subtree group's sweep: its unknowns, its
// entries' local indices and blocks, and (backward) its inverse pivots
template <typename T>
inline long long part_sweep_lds(int max_nodes, int max_entries, bool backward) {
    long long b = 16LL * max_nodes + (4LL + 4LL * (long long)sizeof(T)) * max_entries;
    if (backward) b += 4LL * (long long)sizeof(T) * max_nodes;
    return b;
}

// bytes of LDS of the flat factorisation: the pair products of one level
// (the second product only for column destinations)
template <typename T>
__host__ __device__ inline long long flat_tail_lds(int max_pairs, int max_col) {
    return (long long)sizeof(T) * 4 * ((long long)max_pairs + max_col);
}

// bytes of LDS that stage the L and U blocks of nd destinations of a flat
// step for the next one
template <typename T>
__host__ __device__ inline long long flat_stage_lds(int nd) {
    return (long long)sizeof(T) * 8 * nd;
}

// bytes of dynamic LDS of the tail sweeps: x of the tail nodes (2 doubles
// each), the level bounds, the entry pointers and local indices, and the
// entries' factor blocks (staged) or their slots (read from memory); the
// backward sweep also stages the inverse pivots
template <typename T>
inline long long tail_sweep_lds(int nt, int nlev_tail, int nent, bool staged, bool backward) {
    long long b = 16LL * nt;
    if (staged) b += (long long)sizeof(T) * 4 * nent;
    if (backward) b += (long long)sizeof(T) * 4 * nt;
    b += 4LL * (nlev_tail + 1) + 4LL * (nt + 1) + 4LL * nent;
    if (!staged) b += 4LL * nent;
    return b;
}

}  // namespace pf
