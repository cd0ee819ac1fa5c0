// Internal host-side data model of libpf (not part of the C ABI).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/pf.h"

// Network as read from a CDF/CDFX file: 0-based structure of arrays, the
// layout of the reference's sistemaType/barraType/ramoType
// (CPU/estrutura/sistema.h:23-59, barra.h:5-22, ramo.h:13-28).
struct pf_network {
    int nB = 0, nL = 0, nPV = 0, nPQ = 0, slack = -1, n_duplicates = 0;
    double baseMVA = 0;
    std::vector<int> id;
    std::vector<double> V, theta, Pliq, Qliq, Pload, Qload, Pg, Qg, Vbase, gsh, bsh;
    std::vector<int> pv, pq;                  // ascending bus indices
    std::vector<double> Qmin, Qmax, Vset;     // per PV
    std::vector<int> from, to;
    std::vector<double> r, x, bch, tap_re, tap_im, phi;
};

namespace pf {

// most nodes of the dense top block the single-network solver factors as
// one block: FP32 factor (kChainMax) and FP64 factor (kChainMaxF64), as many
// as its register tiles hold
constexpr int kChainMax = 128;
// most destinations of a flat step whose blocks the next step may read from
// LDS staging (slots are 15-bit in FlatTail::fl_s)
constexpr int kFlatStageMax = 32767;
constexpr int kChainMaxF64 = 101;
// widest elimination-tree level the dense top block may take
constexpr int kTopWidth = 1;
// elimination-tree levels narrower than this run inside one workgroup (the
// single-network solver's tail); the levels below it form subtree groups
constexpr int kTailWidth = 64;
// the same threshold for the straightforward level schedule (lu_blocked = 0)
constexpr int kPlainTailWidth = 16;

// Ybus pattern (host, once per network): diagonal + both directions of every
// branch, columns sorted; plus, per branch, the four CSR slots its pi-model
// terms land in.  Values are assembled on the GPU.
struct YPattern {
    int nB = 0, nnz = 0;
    std::vector<int> rowptr, colind, coorow, diag;
    std::vector<int> br_ff, br_ft, br_tf, br_tt;   // per branch
    // deterministic segmented sum: for each CSR slot, the list of
    // contributions (branch term codes) in a fixed order
    std::vector<int> contrib_ptr, contrib;        // code = 4*branch + {0 ff,1 ft,2 tf,3 tt}
};

YPattern build_ypattern(const pf_network &net);

// Symbolic analysis of the 2x2-block Jacobian on the bus graph minus the
// slack.  Node j (0..n-1) is a block unknown in elimination order.
struct Symbolic {
    int n = 0;                  // nB - 1
    int nnzL = 0;               // strict-lower 2x2 blocks (= strict-upper blocks)
    std::vector<int> perm;      // perm[j] = block index (bus order minus slack) of node j
    std::vector<int> iperm;     // inverse
    std::vector<int> colptr, rowind;   // struct(L(:,j)), ascending rows > j
    std::vector<int> parent, level;
    int nlev = 0;
    std::vector<int> lvl_ptr, lvl_nodes;   // nodes grouped by level (the identity after analyze)
    int tail_start = 0;         // first level processed by the single-workgroup tail
    // factor program: per node, destinations (diag first, then column entries)
    std::vector<int> fd_ptr;    // [n+1] into destination list
    std::vector<int> fc_ptr;    // [ndest+1] into contribution pairs
    std::vector<int> fc_a, fc_b;    // pairs (p_ik, p_jk): positions in L/U arrays
    // destinations of every level, flattened (cooperative factorisation):
    // entry e of level l (ld_ptr[l] <= e < ld_ptr[l+1]) is destination
    // ld_d[e] (an index into fc_ptr) of node ld_node[e]
    std::vector<int> ld_ptr, ld_node, ld_d;
    // forward-solve rows: per node j, k in rowstruct(j) and position p_jk
    std::vector<int> rs_ptr, rs_k, rs_p;
    // the factor also stores L row by row for the forward solve: L position
    // p is copied to block n + 2 nnzL + lr_pos[p] (lr_pos inverts rs_p)
    std::vector<int> lr_pos;
    // dense top ("chain"): the last ch_m nodes [ch_c0, n), the levels
    // [nlev_s, nlev) of the elimination tree.  Their Schur complement S
    // (ch_m x ch_m blocks, zero outside the pattern of L) is formed with the
    // updates of the nodes before ch_c0 and inverted densely; S^-1 is stored
    // row-major at block n + 3 nnzL.  ch_split[e]: for chain destination
    // fd_ptr[ch_c0] + e, the end of its update pairs from nodes < ch_c0.
    int ch_m = 0, ch_c0 = 0, nlev_s = 0;
    std::vector<int> ch_split;
    // the levels below tail_start split into groups of whole subtrees, one
    // workgroup per group (no dependence between groups): group g owns the
    // nodes part_nodes[part_ptr[g * tail_start + l] .. part_ptr[g * tail_start + l + 1])
    // of level l
    int n_parts = 0;
    std::vector<int> part_ptr, part_nodes;
    // the tail levels [tail_start, nlev_s) as chunks of at most kChunkPairs
    // update pairs, each inside one destination: level tail_start + t has the
    // chunks ck_lvl[t] .. ck_lvl[t+1]); chunk c covers pairs ck_q0[c] ..
    // (ck_q1[c] & kChunkMask) of a diagonal destination when bit 30 is set.
    // Destination E (a flattened ld entry, counted from the tail's first)
    // belongs to node te_node[E], column slot te_p[E] (-1: the pivot) and
    // owns the chunks te_ck[E] .. te_ck[E+1]).  ck_max: most chunks on a level.
    std::vector<int> ck_lvl, ck_q0, ck_q1, te_node, te_p, te_ck;
    int ck_max = 0;
    // Jacobian slot map: block slot -> Y entry (-1: fill-in, zero)
    //  slots [0, n): diagonal, [n, n+nnzL): L, [n+nnzL, n+2nnzL): U
    std::vector<int> slot_e;
    int nslot() const { return n + 2 * nnzL; }
    // device factor storage, in blocks: Jacobian slots [0, n + 2 nnzL) (the
    // diagonal slots keep the updated pivots), the row-ordered copy of L,
    // the dense chain inverse, then the inverse pivots at di0()
    int di0() const { return n + 3 * nnzL + ch_m * ch_m; }
    int nblk() const { return di0() + n; }
    // nodes numbered level by level: level l is the node range
    // [lvl_ptr[l], lvl_ptr[l+1]); the GPU kernels rely on it
    bool level_ordered() const {
        for (int i = 0; i < n; i++)
            if (lvl_nodes[i] != i) return false;
        return true;
    }
};

// chain_cap: most nodes of the dense top factored as one block (0: none);
// top_width: widest elimination-tree level it may take
Symbolic analyze(const pf_network &net, const YPattern &yp, int tail_width, int chain_cap, int top_width);

constexpr int kChunkPairs = 16;
constexpr int kChunkDiag = 1 << 30, kChunkMask = kChunkDiag - 1;
// Builds the chunk lists of the tail levels [tail_start, nlev_s), chunks of
// at most chunk_pairs update pairs.
void chunk_tail(Symbolic &s, int chunk_pairs);

// The triangular sweeps of the tail nodes [t0, t1) (t0 = lvl_ptr[tail_start],
// t1 = lvl_ptr[nlev_s]) with the tail's own unknowns held by one workgroup:
// every forward row and backward column of a tail node is split into the
// entries on nodes outside the tail, whose values are final before the sweep
// starts (forward: nodes < t0; backward: the dense top, nodes >= t1), and the
// entries on tail nodes, walked level by level.  Per tail node a = j - t0:
//   forward  fx[fx_ptr[a], fx_ptr[a+1]): (node k < t0, row-ordered L slot q)
//            fi[fi_ptr[a], fi_ptr[a+1]): (tail-local k - t0, q)
//   backward bx[bx_ptr[a], bx_ptr[a+1]): (node i >= t1, L position p)
//            bi[bi_ptr[a], bi_ptr[a+1]): (tail-local i - t0, p)
// in the order of the row (ascending k) and of the column (ascending i).
struct TailSweeps {
    int t0 = 0, t1 = 0;
    std::vector<int> fx_ptr, fx_k, fx_q, fi_ptr, fi_k, fi_q;
    std::vector<int> bx_ptr, bx_i, bx_p, bi_ptr, bi_i, bi_p;
};
TailSweeps tail_sweeps(const Symbolic &s);

// Levels factored as flat pair lists, one pair per lane.  A program is a
// sequence of level steps t: step t owns the destinations
// [de_ptr[t], de_ptr[t+1]) and the update pairs fl_a/fl_b[fl_ptr[t],
// fl_ptr[t+1]) (pairs (p_ik, p_jk)), the pivots' pairs first, the column
// destinations' from fl_ptr[t] + fl_split[t] on.  Destination E: node
// de_node[E], L position de_p[E] (-1: the pivot), its pairs at
// [de_lo[E], de_hi[E]) of its step's list and those of its node's pivot at
// [de_dlo[E], de_dhi[E]); its pairs from sources below the program, applied
// before it, are fc_a/fc_b[de_q0[E], de_ext[E]); de_lr[E] = lr_pos of its
// L position (0 for a pivot).  A destination's products are summed by
// st_g[t] lanes (1, 4 or 16, a function of the step alone: see
// flat_group_width): lane i adds products lo + i, lo + i + G, ... in order
// and the lanes' sums are added as a balanced tree in lane order
// (flat_group_sum restates it).  A pair whose source node belongs to the
// previous step (its blocks were formed there) has fl_s = sa | sb << 16,
// the indices of its two positions' destinations in that step's list, so
// a kernel can read them from where that step staged them; every other
// pair has fl_s = -1 (also a fresh pair after a step of more than
// kFlatStageMax destinations).  max_pairs / max_col bound one step.  The
// pair and destination lists end with one sentinel entry past the program
// (pair 0; a pivot of node 0 without pairs), which lanes without a pair or
// a destination read when they prefetch; de_ptr.back() is its index.
//   flat_tail: levels [lv0, nlev_s), step t = level lv0 + t, destinations
//   in ld order; with split_below (single-network solver, lv0 = tail_start)
//   the pairs from sources below lv0 (k < lvl_ptr[lv0]) are left to a
//   grid-wide pass, without it (one workgroup per scenario) the lists hold
//   every pair.
//   flat_parts: the subtree groups below tail_start, step g * tail_start + l
//   = the nodes of group g on level l (every pair is inside the group).
struct FlatTail {
    int lv0 = 0;
    std::vector<int> fl_ptr, fl_split, fl_a, fl_b, fl_s, de_ptr;
    std::vector<int> de_node, de_p, de_q0, de_ext, de_lo, de_hi, de_dlo, de_dhi, de_lr, st_g;
    int max_pairs = 0, max_col = 0;
    bool stage_ok = true;   // fl_s is usable (a pair's two positions always share a step)
};
FlatTail flat_tail(const Symbolic &s, int lv0, bool split_below);
FlatTail flat_parts(const Symbolic &s);
// Starting state of every scenario of a batch from rows of an array on GPU
// `device` (row s: scenario s; |V| = row[v_col + i] * sd[v_col + i] +
// mu[v_col + i], theta likewise from th_col), written on `stream` (a
// hipStream_t).  Rows have ld floats and must hold v_col + nB and
// th_col + nB of each scenario's network (checked against ld).
pf_status batch_start_from_rows(pf_batch *b, int device, int rows, const float *Y, long long ld, int v_col,
                                int th_col, const float *mu, const float *sd, void *stream);

// lanes per destination of a step with nd destinations of at most maxpd
// products each
int flat_group_width(int nd, int maxpd);
// the sum of v[lo, hi) (blocks of 4 doubles) in the order G lanes take it
void flat_group_sum(const double *v, int lo, int hi, int G, double *out);

// The subtree groups' sweeps with each group's unknowns held by its
// workgroup.  Per node u of part_nodes (group g owns u in [part_ptr[g T],
// part_ptr[(g+1) T])), in that group's local numbering (u - part_ptr[g T]):
//   forward  fi[fi_ptr[u], fi_ptr[u+1]): (local k, row-ordered L slot q),
//            every entry of the row (its sources are the group's own);
//   backward bi[bi_ptr[u], bi_ptr[u+1]): (local i, L position p) for the
//            column's entries inside the group, bx[bx_ptr[u], ...): (node i,
//            p) for those above it (final before the group's sweep).
// max_nodes / max_fi / max_bi bound one group.
struct PartSweeps {
    std::vector<int> fi_ptr, fi_k, fi_q, bi_ptr, bi_i, bi_p, bx_ptr, bx_i, bx_p;
    int max_nodes = 0, max_fi = 0, max_bi = 0;
};
PartSweeps part_sweeps(const Symbolic &s);

// Splits the levels below s.tail_start into at most max_parts groups of
// whole subtrees, balanced by factorisation work (longest processing time
// first).
void partition_subtrees(Symbolic &s, int max_parts);

// Minimum-degree ordering of a symmetric pattern (CSR, diagonal ignored).
std::vector<int> min_degree(int n, const std::vector<int> &rowptr, const std::vector<int> &colind);

// Host numeric factor/solve over the same symbolic program (metodo hibridoC).
// blocks: nslot() 2x2 blocks (row-major 4 doubles each), overwritten by
// the factors (diagonal slots hold the inverse pivots).  Returns false on a
// zero/non-finite pivot.
bool host_block_factor(const Symbolic &s, double *blocks);
void host_block_solve(const Symbolic &s, const double *blocks, double *x /* 2n, node order */);

// The same factor and solve in the order of the device's blocked program:
// the subtree groups' flat steps (parts: flat_parts, or empty), the narrow
// top's pairs from below it and its flat steps (top: flat_tail with
// split_below), the dense chain; the solve with the subtree groups' and the
// narrow top's sweep splits (part_sweeps, tail_sweeps).  Same layout and
// results as host_block_factor / host_block_solve up to the order of the
// sums.
bool host_flat_factor(const Symbolic &s, const FlatTail &parts, const FlatTail &top, double *blocks);
void host_split_solve(const Symbolic &s, const TailSweeps &ts, const PartSweeps &ps, const double *blocks,
                      double *x);

// text of the last failure on this thread, read back by pf_last_error()
void set_last_error(const std::string &what);

// data.txt parsing
void options_default(pf_options &o);
pf_status options_parse(const std::string &text, int cpu_layout, pf_options &o);

// CDF / CDFX reader, dispatching .m (MATPOWER) and .pwf (ANAREDE) files by
// extension (CPU/leitura/arquivo.cpp:3-51)
pf_status read_case(const std::string &path, double v0, double theta0, pf_network &net);
pf_status read_matpower(const std::string &path, double v0, double theta0, pf_network &net);
pf_status read_pwf(const std::string &path, double v0, double theta0, pf_network &net);

}  // namespace pf
