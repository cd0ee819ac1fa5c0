// Single-network NR driver of libpf: device buffers, the per-pass launch
// sequence (eager or one captured hipGraph), the three metodo variants and
// the stage entry points.
//
// One NR pass (CPU/newtonRaphson.cpp:180-305, GPU/newtonRaphson.cuh:283-560):
//   k_pq -> [k_limits: trigger] -> k_mismatch -> k_control -> [k_limits: switch
//   at convergence] -> step: k_jfill -> factor -> solves (-> refinement) -> k_update
// Every kernel after k_control is gated on the device flag ctl->step, and
// every kernel before it on ctl->done, so a captured pass can be replayed
// until the device says it converged; the host polls one small struct.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "kernels_decl.hpp"
#include "pf_internal.hpp"

using namespace pf;

namespace {

thread_local std::string g_last_error;

#define HIPCHK(expr)                                                               \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            g_last_error = std::string(#expr) + ": " + hipGetErrorString(e_);      \
            return PF_DEVICE_ERROR;                                                \
        }                                                                          \
    } while (0)

#define PFCHK(expr)                         \
    do {                                    \
        pf_status s_ = (expr);              \
        if (s_ != PF_OK) return s_;         \
    } while (0)


}  // namespace

void pf::set_last_error(const std::string &what) { g_last_error = what; }

struct pf_solver {
    pf_network net;
    pf_options opt;
    int device = 0;
    YPattern yp;
    Symbolic sym;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    DevNet dn{};
    DevState ds{};
    DevSym dy{};
    // tail sweeps with x in LDS (blocked program): per precision [double,
    // float] 0 = not used (the level walk through memory), 1 = entry slots
    // in LDS, 2 = entry blocks staged in LDS; bytes of dynamic LDS per sweep
    DevTail dtl{};
    // flat tail factorisation (blocked program): per precision [double,
    // float] its dynamic LDS bytes, 0 = not used (the chunked walk)
    DevFlat dfl{}, dfp{};
    // subtree groups' sweeps with x in LDS: per precision, forward and backward LDS bytes (0: not used)
    DevParts dps{};
    int psw_lds[2][2] = {{0, 0}, {0, 0}};
    int flat_ne = 0;
    int flat_lds[2] = {0, 0}, parts_lds[2] = {0, 0};
    int flat_cap = 0;   // FP32 tail: most destinations of a step whose blocks the next step reads from LDS
    int tail_mode[2] = {0, 0};
    int tail_lds_f[2] = {0, 0}, tail_lds_b[2] = {0, 0};
    double *blk64 = nullptr;
    float *blk32 = nullptr;
    double *dense = nullptr;
    Ctl *ctl = nullptr;
    Ctl *ctl_host = nullptr;      // pinned, 2 slots
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    std::vector<int> e2slot;
    std::vector<double> Pliq, Qliq, Qload;   // current injections (host copy)
    std::vector<double> hblk, hx;
    FlowBranch *flow_br = nullptr;   // branch-flow constants (built on first use)
    double *flow_out = nullptr;
    // stage timing (the reference's timers, CPU/benchmarks.h:41-52): HIP
    // events around the timed phases of every pass, eager solves only
    bool timing = false;
    hipEvent_t tev[6] = {};
    pf_timing times{};
    double hybrid_linear_ms = 0;   // host factor + solve of the last hibridoC pass
    // per-pass trace for the caller (max |g|, sum g, PV<->PQ flips)
    pf_pass_callback trace_cb = nullptr;
    void *trace_user = nullptr;
    std::vector<unsigned char> trace_lq;
    std::vector<double> trace_ql, trace_g;
    void mark(int i) {
        if (timing) (void)hipEventRecord(tev[i], stream);
    }

    template <typename T>
    pf_status alloc(T **p, size_t n) {
        void *q = nullptr;
        HIPCHK(hipMalloc(&q, sizeof(T) * std::max<size_t>(n, 1)));
        allocs.push_back(q);
        *p = (T *)q;
        return PF_OK;
    }
    template <typename T>
    pf_status upload(T **p, const std::vector<T> &v) {
        PFCHK(alloc(p, v.size()));
        if (!v.empty()) HIPCHK(hipMemcpy(*p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
        return PF_OK;
    }
    template <typename T>
    pf_status put(T *d, const T *h, size_t n) {
        if (n) HIPCHK(hipMemcpyAsync(d, h, sizeof(T) * n, hipMemcpyHostToDevice, stream));
        return PF_OK;
    }
    template <typename T>
    pf_status get(T *h, const T *d, size_t n) {
        if (n) HIPCHK(hipMemcpyAsync(h, d, sizeof(T) * n, hipMemcpyDeviceToHost, stream));
        HIPCHK(hipStreamSynchronize(stream));
        return PF_OK;
    }

    ~pf_solver() {
        if (gexec) (void)hipGraphExecDestroy(gexec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        for (auto &e : tev)
            if (e) (void)hipEventDestroy(e);
        for (void *p : allocs) (void)hipFree(p);
        if (ctl_host) (void)hipHostFree(ctl_host);
        if (stream) (void)hipStreamDestroy(stream);
    }

    int lim() const { return opt.lim_inj_reat ? 1 : 0; }
    // the solve path of the metodo: hibridoA takes the sparse GPU path,
    // hibridoB the GPU Jacobian + dense GPU LU path (GPU/newtonRaphson.cuh:480-530)
    int path() const {
        return opt.metodo == PF_HIBRIDO_A ? PF_ESPARSO : opt.metodo == PF_HIBRIDO_B ? PF_DENSO : opt.metodo;
    }
    bool mixed() const { return opt.precision == PF_MIXED && path() == PF_ESPARSO; }

    pf_status build();
    pf_status reset_state();
    pf_status flows(double *Pdp, double *Qdp, double *Ppd, double *Qpd);
    pf_status set_isq_from_limq(const unsigned char *limQ);
    pf_status reset_ctl();
    pf_status enqueue_head(int gate);
    pf_status enqueue_factor_solve(int gate);
    pf_status setup_tail_sweeps();
    pf_status setup_flat_tail();
    pf_status upload_flat(const FlatTail &ft, DevFlat &F);
    template <typename T> pf_status enqueue_lu(T *blk, int gate);
    template <typename T> pf_status enqueue_solve(const T *blk, double *v, int gate);
    pf_status enqueue_step(int gate);
    pf_status solve(pf_result *r);
    pf_status run_hybrid_step();
    pf_status ensure_dense();
    pf_status account_pass(bool step, double *pass_ms);
    pf_status trace_pass(int pass, const Ctl &c, const double *pass_ms);
};

// ------------------------------------------------------------------ build

pf_status pf_solver::build() {
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void **)&ctl_host, 2 * sizeof(Ctl), hipHostMallocDefault));
    yp = build_ypattern(net);
    // lu_blocked = 0, the straightforward level schedule: one launch per
    // elimination-tree level, levels narrower than kPlainTailWidth nodes in
    // one workgroup, no dense chain and no subtree groups
    const int chain_cap = opt.precision == PF_MIXED ? ChainTiles<float>::cap : ChainTiles<double>::cap;
    sym = analyze(net, yp, opt.lu_blocked ? kTailWidth : kPlainTailWidth, opt.lu_blocked ? chain_cap : 0,
                  kTopWidth);
    if (!sym.level_ordered() || (sym.ch_m > 0 && sym.tail_start > sym.nlev_s)) {
        g_last_error = "symbolic analysis: nodes not numbered level by level, or chain outside the tail";
        return PF_INVALID_ARG;
    }
    const int nB = net.nB, n = sym.n, sl = net.slack;
    // slot of every Y entry (for the reduced-layout Jacobian export)
    e2slot.assign(yp.nnz, -1);
    for (int s = 0; s < sym.nslot(); s++)
        if (sym.slot_e[s] >= 0) e2slot[sym.slot_e[s]] = s;
    std::vector<int> node_of_bus(nB, -1), bus_of_node(n), pv_of_bus(nB, -1);
    for (int k = 0; k < nB; k++)
        if (k != sl) node_of_bus[k] = sym.iperm[k - (k > sl)];
    for (int j = 0; j < n; j++) {
        int b = sym.perm[j];
        bus_of_node[j] = b + (b >= sl ? 1 : 0);
    }
    for (int i = 0; i < net.nPV; i++) pv_of_bus[net.pv[i]] = i;
    std::vector<int> pcol(sym.nnzL);
    for (int j = 0; j < n; j++)
        for (int p = sym.colptr[j]; p < sym.colptr[j + 1]; p++) pcol[p] = j;

    dn.nB = nB;
    dn.nL = net.nL;
    dn.slack = sl;
    dn.nPV = net.nPV;
    dn.n = n;
    PFCHK(upload((int **)&dn.rowptr, yp.rowptr));
    PFCHK(upload((int **)&dn.colind, yp.colind));
    PFCHK(upload((int **)&dn.coorow, yp.coorow));
    PFCHK(upload((int **)&dn.ydiag, yp.diag));
    PFCHK(upload((int **)&dn.yc_ptr, yp.contrib_ptr));
    PFCHK(upload((int **)&dn.yc, yp.contrib));
    PFCHK(upload((double **)&dn.br_r, net.r));
    PFCHK(upload((double **)&dn.br_x, net.x));
    PFCHK(upload((double **)&dn.br_b, net.bch));
    PFCHK(upload((double **)&dn.br_tr, net.tap_re));
    PFCHK(upload((double **)&dn.br_ti, net.tap_im));
    PFCHK(upload((double **)&dn.gsh, net.gsh));
    PFCHK(upload((double **)&dn.bsh, net.bsh));
    PFCHK(alloc(&dn.y, yp.nnz));
    PFCHK(upload((int **)&dn.node_of_bus, node_of_bus));
    PFCHK(upload((int **)&dn.bus_of_node, bus_of_node));
    PFCHK(upload((int **)&dn.pv_bus, net.pv));
    PFCHK(upload((int **)&dn.pv_of_bus, pv_of_bus));
    PFCHK(upload((double **)&dn.qmin, net.Qmin));
    PFCHK(upload((double **)&dn.qmax, net.Qmax));
    PFCHK(upload((double **)&dn.vset, net.Vset));

    PFCHK(alloc(&ds.V, nB));
    PFCHK(alloc(&ds.th, nB));
    PFCHK(alloc(&ds.E, nB));
    PFCHK(alloc(&ds.I, nB));
    PFCHK(alloc(&ds.P, nB));
    PFCHK(alloc(&ds.Q, nB));
    PFCHK(upload((double **)&ds.Pliq, Pliq));
    PFCHK(upload((double **)&ds.Qliq, Qliq));
    PFCHK(upload((double **)&ds.Qload, Qload));
    PFCHK(alloc(&ds.QliqLim, nB));
    PFCHK(alloc(&ds.isq, nB));
    PFCHK(alloc(&ds.limQ, net.nPV));
    PFCHK(alloc(&ds.g, 2 * (size_t)n));
    PFCHK(alloc(&ds.x, 2 * (size_t)n));
    PFCHK(alloc(&ds.r, 2 * (size_t)n));
    PFCHK(alloc(&ctl, 1));
    ds.ctl = ctl;

    dy.n = n;
    dy.nnzL = sym.nnzL;
    dy.nslot = sym.nslot();
    dy.nlev = sym.nlev;
    dy.di0 = sym.di0();
    dy.ch_m = sym.ch_m;
    dy.ch_c0 = sym.ch_c0;
    dy.nlev_s = sym.nlev_s;
    PFCHK(upload((int **)&dy.ch_split, sym.ch_split));
    if (opt.lu_blocked) {   // one workgroup per group of subtrees, one group per CU
        int cus = 0;
        HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
        partition_subtrees(sym, std::max(1, cus));
    }
    chunk_tail(sym, kChunkPairs);
    const bool chunked = sym.ck_max <= kMaxChunks && !sym.ck_q0.empty();
    dy.tail_start = sym.tail_start;
    dy.te0 = sym.ld_ptr[sym.tail_start];
    dy.ck_lvl = dy.ck_q0 = dy.ck_q1 = dy.te_node = dy.te_p = dy.te_ck = nullptr;
    if (chunked) {
        PFCHK(upload((int **)&dy.ck_lvl, sym.ck_lvl));
        PFCHK(upload((int **)&dy.ck_q0, sym.ck_q0));
        PFCHK(upload((int **)&dy.ck_q1, sym.ck_q1));
        PFCHK(upload((int **)&dy.te_node, sym.te_node));
        PFCHK(upload((int **)&dy.te_p, sym.te_p));
        PFCHK(upload((int **)&dy.te_ck, sym.te_ck));
    }
    if (opt.lu_blocked && sym.tail_start < sym.nlev_s) {
        PFCHK(setup_tail_sweeps());
        PFCHK(setup_flat_tail());
    }
    dy.n_parts = sym.n_parts;
    dy.part_levels = sym.tail_start;
    PFCHK(upload((int **)&dy.part_ptr, sym.part_ptr));
    PFCHK(upload((int **)&dy.part_nodes, sym.part_nodes));
    PFCHK(upload((int **)&dy.colptr, sym.colptr));
    PFCHK(upload((int **)&dy.rowind, sym.rowind));
    PFCHK(upload((int **)&dy.pcol, pcol));
    PFCHK(upload((int **)&dy.lvl_ptr, sym.lvl_ptr));
    PFCHK(upload((int **)&dy.lvl_nodes, sym.lvl_nodes));
    PFCHK(upload((int **)&dy.fd_ptr, sym.fd_ptr));
    PFCHK(upload((int **)&dy.fc_ptr, sym.fc_ptr));
    PFCHK(upload((int **)&dy.fc_a, sym.fc_a));
    PFCHK(upload((int **)&dy.fc_b, sym.fc_b));
    PFCHK(upload((int **)&dy.rs_ptr, sym.rs_ptr));
    PFCHK(upload((int **)&dy.rs_k, sym.rs_k));
    PFCHK(upload((int **)&dy.rs_p, sym.rs_p));
    PFCHK(upload((int **)&dy.lr_pos, sym.lr_pos));
    PFCHK(upload((int **)&dy.slot_e, sym.slot_e));
    PFCHK(upload((int **)&dy.ld_ptr, sym.ld_ptr));
    PFCHK(upload((int **)&dy.ld_node, sym.ld_node));
    PFCHK(upload((int **)&dy.ld_d, sym.ld_d));
    PFCHK(alloc(&blk64, 4 * (size_t)sym.nblk()));
    if (opt.precision == PF_MIXED) PFCHK(alloc(&blk32, 4 * (size_t)sym.nblk()));

    // Ybus on the device (once per network), timed like the reference's
    // admitancia phase (CPU/main.cpp:47-56)
    for (auto &e : tev) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(tev[0], stream));
    k_ybus<<<(yp.nnz + 255) / 256, 256, 0, stream>>>(dn, yp.nnz);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(tev[1], stream));
    HIPCHK(hipEventSynchronize(tev[1]));
    {
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, tev[0], tev[1]));
        times.ybus_ms = ms;
    }
    PFCHK(reset_state());
    HIPCHK(hipStreamSynchronize(stream));
    return PF_OK;
}

pf_status pf_solver::ensure_dense() {
    if (dense) return PF_OK;
    size_t m = 2 * (size_t)sym.n;
    if (m > 4096) return PF_INVALID_ARG;   // dense path is for systems up to 4096 unknowns
    return alloc(&dense, m * m);
}

pf_status pf_solver::set_isq_from_limq(const unsigned char *limQ) {
    std::vector<unsigned char> isq(net.nB, 0);
    for (int k : net.pq) isq[k] = 1;
    for (int i = 0; i < net.nPV; i++) isq[net.pv[i]] = limQ ? limQ[i] : 0;
    std::vector<unsigned char> lq(net.nPV, 0);
    if (limQ) std::copy(limQ, limQ + net.nPV, lq.begin());
    PFCHK(put(ds.isq, isq.data(), isq.size()));
    PFCHK(put(ds.limQ, lq.data(), lq.size()));
    HIPCHK(hipStreamSynchronize(stream));
    return PF_OK;
}

// flat start (CPU/estrutura/barra.cpp:20-25 + PV/slack fields) and initLim
// (CPU/limInjReat.cpp:177-207)
pf_status pf_solver::reset_state() {
    PFCHK(put(ds.V, net.V.data(), net.nB));
    PFCHK(put(ds.th, net.theta.data(), net.nB));
    PFCHK(put(ds.QliqLim, Qliq.data(), net.nB));
    PFCHK(set_isq_from_limq(nullptr));
    return PF_OK;
}

pf_status pf_solver::reset_ctl() {
    Ctl c{};
    c.tol = opt.tolerancia;
    c.max_iter = opt.no_max_iter;
    ctl_host[0] = c;
    PFCHK(put(ctl, ctl_host, 1));
    HIPCHK(hipStreamSynchronize(stream));
    return PF_OK;
}

// ------------------------------------------------------- launch sequences

pf_status pf_solver::enqueue_head(int gate) {
    const int nB = net.nB;
    mark(0);
    k_pq<4><<<(nB + 63) / 64, 256, 0, stream>>>(dn, ds, gate);
    mark(1);
    if (lim()) k_limits<<<1, 1024, 0, stream>>>(dn, ds, 0, 1);
    k_mismatch<<<(nB + 255) / 256, 256, 0, stream>>>(dn, ds, gate);
    k_control<<<1, 64, 0, stream>>>(ds, lim());
    if (lim()) k_limits<<<1, 1024, 0, stream>>>(dn, ds, 1, 1);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

// The tail's sweeps with its unknowns in LDS, per precision as the LDS
// holds them: the entries' factor blocks staged, else only their slots.
pf_status pf_solver::setup_tail_sweeps() {
    const TailSweeps ts = tail_sweeps(sym);
    const int nt = ts.t1 - ts.t0, nlt = sym.nlev_s - sym.tail_start;
    dtl.t0 = ts.t0;
    dtl.t1 = ts.t1;
    dtl.nfi = (int)ts.fi_k.size();
    dtl.nbi = (int)ts.bi_i.size();
    PFCHK(upload((int **)&dtl.fx_ptr, ts.fx_ptr));
    PFCHK(upload((int **)&dtl.fx_k, ts.fx_k));
    PFCHK(upload((int **)&dtl.fx_q, ts.fx_q));
    PFCHK(upload((int **)&dtl.fi_ptr, ts.fi_ptr));
    PFCHK(upload((int **)&dtl.fi_k, ts.fi_k));
    PFCHK(upload((int **)&dtl.fi_q, ts.fi_q));
    PFCHK(upload((int **)&dtl.bx_ptr, ts.bx_ptr));
    PFCHK(upload((int **)&dtl.bx_i, ts.bx_i));
    PFCHK(upload((int **)&dtl.bx_p, ts.bx_p));
    PFCHK(upload((int **)&dtl.bi_ptr, ts.bi_ptr));
    PFCHK(upload((int **)&dtl.bi_i, ts.bi_i));
    PFCHK(upload((int **)&dtl.bi_p, ts.bi_p));
    int max_lds = 0;
    HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const long long cap = max_lds - 8192;   // room for the kernels' static LDS
    auto pick = [&](int pi, long long f2, long long b2, long long f1, long long b1) {
        if (f2 <= cap && b2 <= cap) {
            tail_mode[pi] = 2, tail_lds_f[pi] = (int)f2, tail_lds_b[pi] = (int)b2;
        } else if (f1 <= cap && b1 <= cap) {
            tail_mode[pi] = 1, tail_lds_f[pi] = (int)f1, tail_lds_b[pi] = (int)b1;
        }
    };
    pick(0, tail_sweep_lds<double>(nt, nlt, dtl.nfi, true, false), tail_sweep_lds<double>(nt, nlt, dtl.nbi, true, true),
         tail_sweep_lds<double>(nt, nlt, dtl.nfi, false, false), tail_sweep_lds<double>(nt, nlt, dtl.nbi, false, true));
    pick(1, tail_sweep_lds<float>(nt, nlt, dtl.nfi, true, false), tail_sweep_lds<float>(nt, nlt, dtl.nbi, true, true),
         tail_sweep_lds<float>(nt, nlt, dtl.nfi, false, false), tail_sweep_lds<float>(nt, nlt, dtl.nbi, false, true));
    auto allow = [&](const void *k, int bytes) -> pf_status {
        HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, bytes));
        return PF_OK;
    };
    if (tail_mode[0] == 2) {
        PFCHK(allow((const void *)k_fwd_tail_lds<double, true>, tail_lds_f[0]));
        PFCHK(allow((const void *)k_bwd_tail_lds<double, true>, tail_lds_b[0]));
    } else if (tail_mode[0] == 1) {
        PFCHK(allow((const void *)k_fwd_tail_lds<double, false>, tail_lds_f[0]));
        PFCHK(allow((const void *)k_bwd_tail_lds<double, false>, tail_lds_b[0]));
    }
    if (tail_mode[1] == 2) {
        PFCHK(allow((const void *)k_fwd_tail_lds<float, true>, tail_lds_f[1]));
        PFCHK(allow((const void *)k_bwd_tail_lds<float, true>, tail_lds_b[1]));
    } else if (tail_mode[1] == 1) {
        PFCHK(allow((const void *)k_fwd_tail_lds<float, false>, tail_lds_f[1]));
        PFCHK(allow((const void *)k_bwd_tail_lds<float, false>, tail_lds_b[1]));
    }
    return PF_OK;
}

// The tail levels' factorisation with flat pair lists, per precision as
// the LDS holds one level's products.
pf_status pf_solver::upload_flat(const FlatTail &ft, DevFlat &F) {
    F.lv0 = ft.lv0;
    F.max_pairs = ft.max_pairs;
    F.max_col = ft.max_col;
    F.de_sentinel = ft.de_ptr.back();
    PFCHK(upload((int **)&F.fl_ptr, ft.fl_ptr));
    PFCHK(upload((int **)&F.fl_split, ft.fl_split));
    PFCHK(upload((int **)&F.fl_a, ft.fl_a));
    PFCHK(upload((int **)&F.fl_b, ft.fl_b));
    PFCHK(upload((int **)&F.fl_s, ft.fl_s));
    PFCHK(upload((int **)&F.de_ptr, ft.de_ptr));
    PFCHK(upload((int **)&F.de_node, ft.de_node));
    PFCHK(upload((int **)&F.de_p, ft.de_p));
    PFCHK(upload((int **)&F.de_q0, ft.de_q0));
    PFCHK(upload((int **)&F.de_ext, ft.de_ext));
    PFCHK(upload((int **)&F.de_lo, ft.de_lo));
    PFCHK(upload((int **)&F.de_hi, ft.de_hi));
    PFCHK(upload((int **)&F.de_dlo, ft.de_dlo));
    PFCHK(upload((int **)&F.de_dhi, ft.de_dhi));
    PFCHK(upload((int **)&F.de_lr, ft.de_lr));
    PFCHK(upload((int **)&F.st_g, ft.st_g));
    return PF_OK;
}

// The tail levels and the subtree groups as flat pair programs, per
// precision as the LDS holds one step's products.
pf_status pf_solver::setup_flat_tail() {
    int max_lds = 0;
    HIPCHK(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    const long long cap = max_lds - 8192;
    auto allow = [&](const void *k, long long bytes, int &out) -> pf_status {
        if (bytes > cap) return PF_OK;
        HIPCHK(hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        out = (int)bytes;
        return PF_OK;
    };
    const FlatTail ft = flat_tail(sym, sym.tail_start, true);
    flat_ne = ft.de_ptr.back();
    PFCHK(upload_flat(ft, dfl));
    PFCHK(allow((const void *)k_factor_tail_flat<double>,
                std::max<long long>(flat_tail_lds<double>(ft.max_pairs, ft.max_col), 16), flat_lds[0]));
    {   // FP32: the products, then the staged blocks of as many destinations as fit
        const long long pf = std::max<long long>(flat_tail_lds<float>(ft.max_pairs, ft.max_col), 16);
        int max_ne = 0;
        for (size_t t = 0; t + 1 < ft.de_ptr.size(); t++) max_ne = std::max(max_ne, ft.de_ptr[t + 1] - ft.de_ptr[t]);
        flat_cap = !ft.stage_ok ? 0 : (int)std::max<long long>(
            0, std::min<long long>({(long long)max_ne, (long long)kFlatStageMax, (cap - pf) / flat_stage_lds<float>(1)}));
        PFCHK(allow((const void *)k_factor_tail_flat<float>, pf + flat_stage_lds<float>(flat_cap), flat_lds[1]));
    }
    if (sym.n_parts > 0) {
        const PartSweeps ps = part_sweeps(sym);
        dps.max_nodes = ps.max_nodes;
        dps.max_fi = ps.max_fi;
        dps.max_bi = ps.max_bi;
        PFCHK(upload((int **)&dps.fi_ptr, ps.fi_ptr));
        PFCHK(upload((int **)&dps.fi_k, ps.fi_k));
        PFCHK(upload((int **)&dps.fi_q, ps.fi_q));
        PFCHK(upload((int **)&dps.bi_ptr, ps.bi_ptr));
        PFCHK(upload((int **)&dps.bi_i, ps.bi_i));
        PFCHK(upload((int **)&dps.bi_p, ps.bi_p));
        PFCHK(upload((int **)&dps.bx_ptr, ps.bx_ptr));
        PFCHK(upload((int **)&dps.bx_i, ps.bx_i));
        PFCHK(upload((int **)&dps.bx_p, ps.bx_p));
        PFCHK(allow((const void *)k_fwd_parts_lds<double>,
                    std::max<long long>(part_sweep_lds<double>(ps.max_nodes, ps.max_fi, false), 16), psw_lds[0][0]));
        PFCHK(allow((const void *)k_bwd_parts_lds<double>,
                    std::max<long long>(part_sweep_lds<double>(ps.max_nodes, ps.max_bi, true), 16), psw_lds[0][1]));
        PFCHK(allow((const void *)k_fwd_parts_lds<float>,
                    std::max<long long>(part_sweep_lds<float>(ps.max_nodes, ps.max_fi, false), 16), psw_lds[1][0]));
        PFCHK(allow((const void *)k_bwd_parts_lds<float>,
                    std::max<long long>(part_sweep_lds<float>(ps.max_nodes, ps.max_bi, true), 16), psw_lds[1][1]));
        const FlatTail fp = flat_parts(sym);
        PFCHK(upload_flat(fp, dfp));
        PFCHK(allow((const void *)k_factor_parts_flat<double>,
                    std::max<long long>(flat_tail_lds<double>(fp.max_pairs, fp.max_col), 16), parts_lds[0]));
        PFCHK(allow((const void *)k_factor_parts_flat<float>,
                    std::max<long long>(flat_tail_lds<float>(fp.max_pairs, fp.max_col), 16), parts_lds[1]));
    }
    return PF_OK;
}

template <typename T>
pf_status pf_solver::enqueue_lu(T *blk, int gate) {
    // the threads past the slots clear the dense top block
    const int nfill = sym.nslot() + sym.ch_m * sym.ch_m;
    mark(2);
    k_jfill<T><<<(nfill + 255) / 256, 256, 0, stream>>>(dn, ds, dy, blk, gate);
    mark(3);
    const int pl = parts_lds[sizeof(T) == 4 ? 1 : 0];
    if (sym.n_parts > 0 && pl) {
        k_factor_parts_flat<T><<<sym.n_parts, 256, pl, stream>>>(dy, dfp, blk, ctl, gate);
    } else if (sym.n_parts > 0) {
        k_factor_parts<T><<<sym.n_parts, 1024, 0, stream>>>(dy, blk, ctl, gate);
    } else {
        for (int l = 0; l < sym.tail_start; l++) {
            int c = sym.lvl_ptr[l + 1] - sym.lvl_ptr[l];
            k_factor_level<T><<<(c + 3) / 4, 256, 0, stream>>>(dy, blk, ctl, sym.lvl_ptr[l], c, gate);
        }
    }
    const int fl = sym.tail_start < sym.nlev_s ? flat_lds[sizeof(T) == 4 ? 1 : 0] : 0;
    if (fl) {   // narrow levels below the chain: pairs from below the tail across the GPU first
        k_tail_factor_ext<T><<<(flat_ne + 31) / 32, 256, 0, stream>>>(dy, dfl, blk, ctl, flat_ne, gate);
        k_factor_tail_flat<T><<<1, 1024, fl, stream>>>(dy, dfl, blk, ctl, sym.nlev_s - sym.tail_start, gate,
                                                       sizeof(T) == 4 ? flat_cap : 0);
    } else if (sym.tail_start < sym.nlev_s) {
        k_factor_tail<T><<<1, 1024, 0, stream>>>(dy, blk, ctl, sym.tail_start, sym.nlev_s, gate);
    }
    if (sym.ch_m > 0) {   // the dense chain: Schur complement across the GPU, then its inverse
        const int ne = sym.fd_ptr[sym.n] - sym.fd_ptr[sym.ch_c0];
        k_chain_schur<T><<<(ne + 15) / 16, 256, 0, stream>>>(dy, blk, ctl, gate);
        k_chain_invert<T><<<1, ChainTiles<T>::threads, 0, stream>>>(dy, blk, ctl, gate);
    }
    HIPCHK(hipGetLastError());
    return PF_OK;
}

template <typename T>
pf_status pf_solver::enqueue_solve(const T *blk, double *v, int gate) {
    const int *pl = psw_lds[sizeof(T) == 4 ? 1 : 0];
    if (sym.n_parts > 0 && pl[0]) {
        k_fwd_parts_lds<T><<<sym.n_parts, 256, pl[0], stream>>>(dy, dps, blk, v, ctl, gate);
    } else if (sym.n_parts > 0) {
        k_fwd_parts<T><<<sym.n_parts, 1024, 0, stream>>>(dy, blk, v, ctl, gate);
    } else {
        for (int l = 0; l < sym.tail_start; l++) {
            int c = sym.lvl_ptr[l + 1] - sym.lvl_ptr[l];
            k_fwd_level<T><<<(c + 15) / 16, 256, 0, stream>>>(dy, blk, v, ctl, sym.lvl_ptr[l], c, gate);
        }
    }
    const int pi = sizeof(T) == 4 ? 1 : 0, tm = sym.tail_start < sym.nlev_s ? tail_mode[pi] : 0;
    const int lv0 = sym.tail_start, lv1 = sym.nlev_s, text = (dtl.t1 - dtl.t0 + 15) / 16;
    if (tm) {
        k_tail_fwd_ext<T><<<text, 256, 0, stream>>>(dy, dtl, blk, v, ctl, gate);
        if (tm == 2) k_fwd_tail_lds<T, true><<<1, 1024, tail_lds_f[pi], stream>>>(dy, dtl, blk, v, ctl, lv0, lv1, gate);
        else k_fwd_tail_lds<T, false><<<1, 1024, tail_lds_f[pi], stream>>>(dy, dtl, blk, v, ctl, lv0, lv1, gate);
        if (sym.ch_m > 0) k_tail_bwd_ext<T><<<text, 256, 0, stream>>>(dy, dtl, blk, v, ctl, gate);
        if (tm == 2) k_bwd_tail_lds<T, true><<<1, 1024, tail_lds_b[pi], stream>>>(dy, dtl, blk, v, ctl, lv0, lv1, gate);
        else k_bwd_tail_lds<T, false><<<1, 1024, tail_lds_b[pi], stream>>>(dy, dtl, blk, v, ctl, lv0, lv1, gate);
    } else {
        if (sym.tail_start < sym.nlev)
            k_fwd_tail<T><<<1, 1024, 0, stream>>>(dy, blk, v, ctl, sym.tail_start, sym.nlev_s, gate);
        if (sym.tail_start < sym.nlev_s)
            k_bwd_tail<T><<<1, 1024, 0, stream>>>(dy, blk, v, ctl, sym.tail_start, sym.nlev_s, gate);
    }
    if (sym.n_parts > 0 && pl[1]) {
        k_bwd_parts_lds<T><<<sym.n_parts, 256, pl[1], stream>>>(dy, dps, blk, v, ctl, gate);
    } else if (sym.n_parts > 0) {
        k_bwd_parts<T><<<sym.n_parts, 1024, 0, stream>>>(dy, blk, v, ctl, gate);
    } else {
        for (int l = sym.tail_start - 1; l >= 0; l--) {
            int c = sym.lvl_ptr[l + 1] - sym.lvl_ptr[l];
            k_bwd_level<T><<<(c + 15) / 16, 256, 0, stream>>>(dy, blk, v, ctl, sym.lvl_ptr[l], c, gate);
        }
    }
    HIPCHK(hipGetLastError());
    return PF_OK;
}

// factor at the current state and solve J x = g (x, in node order)
pf_status pf_solver::enqueue_factor_solve(int gate) {
    const int m = 2 * sym.n;
    if (path() == PF_DENSO) {
        PFCHK(ensure_dense());
        mark(2);
        k_jfill<double><<<(sym.nslot() + 255) / 256, 256, 0, stream>>>(dn, ds, dy, blk64, gate);
        mark(3);
        size_t mm = (size_t)m * m;
        k_fill<<<(unsigned)((mm + 255) / 256), 256, 0, stream>>>(dense, mm, 0.0);
        k_dense_scatter<<<(sym.nslot() + 255) / 256, 256, 0, stream>>>(dy, blk64, dense, m, ctl, gate);
        k_copy<<<(m + 255) / 256, 256, 0, stream>>>(ds.x, ds.g, m, ctl, gate);
        k_dense<<<1, 1024, 0, stream>>>(dense, ds.x, m, ctl, gate);
    } else if (mixed()) {
        PFCHK(enqueue_lu(blk32, gate));
        k_copy<<<(m + 255) / 256, 256, 0, stream>>>(ds.x, ds.g, m, ctl, gate);
        PFCHK(enqueue_solve((const float *)blk32, ds.x, gate));
        for (int r = 0; r < opt.refine_steps; r++) {
            k_residual<4><<<(net.nB + 63) / 64, 256, 0, stream>>>(dn, ds, gate);
            PFCHK(enqueue_solve((const float *)blk32, ds.r, gate));
            k_axpy<<<(m + 255) / 256, 256, 0, stream>>>(ds.x, ds.r, m, ctl, gate);
        }
    } else {
        PFCHK(enqueue_lu(blk64, gate));
        k_copy<<<(m + 255) / 256, 256, 0, stream>>>(ds.x, ds.g, m, ctl, gate);
        PFCHK(enqueue_solve((const double *)blk64, ds.x, gate));
    }
    HIPCHK(hipGetLastError());
    return PF_OK;
}

pf_status pf_solver::enqueue_step(int gate) {
    PFCHK(enqueue_factor_solve(gate));
    mark(4);
    k_update<<<(net.nB + 255) / 256, 256, 0, stream>>>(dn, ds, gate);
    mark(5);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

// metodo hibridoC: Jacobian on the GPU, factor and solve on the host over the
// same symbolic program (the reference: PARDISO, GPU/solverCPU.cpp:218-330).
pf_status pf_solver::run_hybrid_step() {
    const size_t ns = sym.nslot();
    mark(2);
    k_jfill<double><<<(ns + 255) / 256, 256, 0, stream>>>(dn, ds, dy, blk64, 1);
    HIPCHK(hipGetLastError());
    mark(3);
    auto h0 = std::chrono::steady_clock::now();
    hblk.assign(4 * (size_t)sym.nblk(), 0.0);   // slots, then room for the dense chain
    hx.resize(2 * (size_t)sym.n);
    PFCHK(get(hblk.data(), blk64, 4 * ns));
    PFCHK(get(hx.data(), ds.g, hx.size()));
    if (!host_block_factor(sym, hblk.data())) {
        ctl_host[0].singular = 1;
        int one = 1;
        PFCHK(put(&ctl->singular, &one, 1));
    } else {
        host_block_solve(sym, hblk.data(), hx.data());
    }
    PFCHK(put(ds.x, hx.data(), hx.size()));
    HIPCHK(hipStreamSynchronize(stream));
    hybrid_linear_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - h0).count();
    mark(4);
    k_update<<<(net.nB + 255) / 256, 256, 0, stream>>>(dn, ds, 1);
    HIPCHK(hipGetLastError());
    mark(5);
    return PF_OK;
}

// stage times of one eager pass whose events have completed (P/Q,
// Jacobian, linear system, update), added to the totals
pf_status pf_solver::account_pass(bool step, double *pass_ms) {
    for (int i = 0; i < 4; i++) pass_ms[i] = 0;
    if (!timing) return PF_OK;
    auto el = [&](int a, int b) {
        float ms = 0;
        return hipEventElapsedTime(&ms, tev[a], tev[b]) == hipSuccess ? (double)ms : 0.0;
    };
    pass_ms[0] = el(0, 1);
    if (step) {
        pass_ms[1] = el(2, 3);
        pass_ms[2] = path() == PF_HIBRIDO_C ? hybrid_linear_ms : el(3, 4);
        pass_ms[3] = el(4, 5);
    }
    times.pq_ms += pass_ms[0];
    times.jacobian_ms += pass_ms[1];
    times.linear_ms += pass_ms[2];
    times.update_ms += pass_ms[3];
    times.passes++;
    return PF_OK;
}

// hands one pass to the trace callback: max |g|, the sum of g and the
// PV<->PQ flips of the pass (compared with the state before it), with the
// limit each flip concerns (0 PV->PQ at Qmin, 1 PV->PQ at Qmax, 2 back to
// PV from Qmax, 3 back to PV from Qmin: CPU/limInjReat.cpp:38-91)
pf_status pf_solver::trace_pass(int pass, const Ctl &c, const double *pass_ms) {
    if (!trace_cb) return PF_OK;
    const int nPV = net.nPV;
    std::vector<unsigned char> lq(nPV);
    std::vector<double> ql(net.nB), g(2 * (size_t)sym.n);
    if (nPV) PFCHK(get(lq.data(), ds.limQ, nPV));
    PFCHK(get(ql.data(), ds.QliqLim, ql.size()));
    PFCHK(get(g.data(), ds.g, g.size()));
    double sum = 0;
    for (double v : g) sum += v;
    std::vector<int> idx, kind;
    for (int i = 0; i < nPV; i++) {
        if (lq[i] == trace_lq[i]) continue;
        const int b = net.pv[i];
        const bool at_max = (lq[i] ? ql[b] : trace_ql[b]) == net.Qmax[i] - Qload[b];
        idx.push_back(i);
        kind.push_back(lq[i] ? (at_max ? 1 : 0) : (at_max ? 2 : 3));
    }
    trace_cb(trace_user, pass, c.maxG, sum, c.post, (int)idx.size(), idx.data(), kind.data(), pass_ms);
    trace_lq = lq;
    trace_ql = ql;
    return PF_OK;
}

pf_status pf_solver::solve(pf_result *res) {
    HIPCHK(hipSetDevice(device));
    PFCHK(reset_ctl());
    if (timing) {
        const double yb = times.ybus_ms, fl = times.flows_ms;
        times = pf_timing{};
        times.ybus_ms = yb;
        times.flows_ms = fl;
    }
    auto t0 = std::chrono::steady_clock::now();
    k_init_E<<<(net.nB + 255) / 256, 256, 0, stream>>>(dn, ds);
    HIPCHK(hipGetLastError());
    const int maxit = opt.no_max_iter;
    Ctl c{};
    if (opt.use_graph && path() == PF_ESPARSO && !timing && !trace_cb) {
        if (!gexec) {
            HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            pf_status st = enqueue_head(1);
            if (st == PF_OK) st = enqueue_step(1);
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamEndCapture(stream, &g);
            if (st != PF_OK) return st;
            HIPCHK(e);
            graph = g;
            HIPCHK(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        }
        // pipelined polling: replay i+1 is queued before the host waits on the
        // copy of the control block that follows replay i
        int launched = 0, checked = 0;
        while (checked < maxit) {
            while (launched < maxit && launched < checked + 2) {
                HIPCHK(hipGraphLaunch(gexec, stream));
                HIPCHK(hipMemcpyAsync(&ctl_host[launched & 1], ctl, sizeof(Ctl), hipMemcpyDeviceToHost, stream));
                HIPCHK(hipEventRecord(ev[launched & 1], stream));
                launched++;
            }
            HIPCHK(hipEventSynchronize(ev[checked & 1]));
            c = ctl_host[checked & 1];
            checked++;
            if (c.done || c.singular) break;
        }
        HIPCHK(hipStreamSynchronize(stream));
        c = ctl_host[(checked - 1) & 1];
        PFCHK(get(&c, ctl, 1));
    } else {
        if (trace_cb) {
            trace_lq.assign(net.nPV, 0);
            trace_ql = Qliq;
            if (net.nPV) PFCHK(get(trace_lq.data(), ds.limQ, net.nPV));
            PFCHK(get(trace_ql.data(), ds.QliqLim, trace_ql.size()));
        }
        double pass_ms[4];
        for (int it = 0; it < maxit; it++) {
            PFCHK(enqueue_head(1));
            if (path() == PF_HIBRIDO_C) {
                PFCHK(get(&c, ctl, 1));
                const Ctl head = c;
                if (c.step) PFCHK(run_hybrid_step());
                PFCHK(get(&c, ctl, 1));
                PFCHK(account_pass(head.step != 0, pass_ms));
                PFCHK(trace_pass(it, head, pass_ms));
                if (c.done || c.singular) break;
                continue;
            }
            PFCHK(enqueue_step(1));
            PFCHK(get(&c, ctl, 1));
            PFCHK(account_pass(c.step != 0, pass_ms));
            PFCHK(trace_pass(it, c, pass_ms));
            if (c.done || c.singular) break;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (timing) times.iterative_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    res->iterations = c.it;
    res->converged = c.done && !c.singular;
    res->singular = c.singular;
    res->max_mismatch = c.maxG;
    res->n_switch_events = c.flips;
    res->solve_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
    if (c.singular) return PF_SINGULAR;
    return res->converged ? PF_OK : PF_NOT_CONVERGED;
}

// ===================================================================== ABI

extern "C" {

pf_status pf_solver_create(const pf_network *net, const pf_options *opt, int device, pf_solver **out) {
    if (!net || !opt || !out) return PF_INVALID_ARG;
    *out = nullptr;
    if (net->nB < 2 || net->slack < 0) return PF_INVALID_ARG;
    if (opt->metodo < PF_ESPARSO || opt->metodo > PF_HIBRIDO_B) return PF_INVALID_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        g_last_error = "no HIP device";
        return PF_DEVICE_ERROR;
    }
    auto s = std::make_unique<pf_solver>();
    s->net = *net;
    s->opt = *opt;
    s->device = device;
    s->Pliq = net->Pliq;
    s->Qliq = net->Qliq;
    s->Qload = net->Qload;
    pf_status st = s->build();
    if (st != PF_OK) return st;
    *out = s.release();
    return PF_OK;
}

void pf_solver_destroy(pf_solver *s) {
    if (!s) return;
    (void)hipSetDevice(s->device);
    delete s;
}

pf_status pf_solve(pf_solver *s, pf_result *res) {
    if (!s || !res) return PF_INVALID_ARG;
    std::memset(res, 0, sizeof(*res));
    return s->solve(res);
}

pf_status pf_solver_reset(pf_solver *s) {
    if (!s) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    return s->reset_state();
}

pf_status pf_solver_set_state(pf_solver *s, const double *V, const double *theta) {
    if (!s || !V || !theta) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    PFCHK(s->put(s->ds.V, V, s->net.nB));
    PFCHK(s->put(s->ds.th, theta, s->net.nB));
    HIPCHK(hipStreamSynchronize(s->stream));
    return PF_OK;
}

pf_status pf_solver_set_injections(pf_solver *s, const double *Pliq, const double *Qliq,
                                   const double *Qload) {
    if (!s || !Pliq || !Qliq || !Qload) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    const int nB = s->net.nB;
    s->Pliq.assign(Pliq, Pliq + nB);
    s->Qliq.assign(Qliq, Qliq + nB);
    s->Qload.assign(Qload, Qload + nB);
    PFCHK(s->put((double *)s->ds.Pliq, Pliq, nB));
    PFCHK(s->put((double *)s->ds.Qliq, Qliq, nB));
    PFCHK(s->put((double *)s->ds.Qload, Qload, nB));
    HIPCHK(hipStreamSynchronize(s->stream));
    return PF_OK;
}

pf_status pf_solver_get_state(pf_solver *s, double *V, double *theta, unsigned char *limQ) {
    if (!s) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    if (V) PFCHK(s->get(V, s->ds.V, s->net.nB));
    if (theta) PFCHK(s->get(theta, s->ds.th, s->net.nB));
    if (limQ) PFCHK(s->get(limQ, s->ds.limQ, s->net.nPV));
    return PF_OK;
}

pf_status pf_solver_lu_info(pf_solver *s, int *n, int *nnzL, int *levels, int *tail_levels) {
    if (!s) return PF_INVALID_ARG;
    if (n) *n = s->sym.n;
    if (nnzL) *nnzL = s->sym.nnzL;
    if (levels) *levels = s->sym.nlev;
    if (tail_levels) *tail_levels = s->sym.nlev - s->sym.tail_start;
    return PF_OK;
}

pf_status pf_solver_set_timing(pf_solver *s, int on) {
    if (!s) return PF_INVALID_ARG;
    s->timing = on != 0;
    return PF_OK;
}

pf_status pf_solver_get_timing(pf_solver *s, pf_timing *t) {
    if (!s || !t) return PF_INVALID_ARG;
    *t = s->times;
    return PF_OK;
}

pf_status pf_solver_set_trace(pf_solver *s, pf_pass_callback cb, void *user) {
    if (!s) return PF_INVALID_ARG;
    s->trace_cb = cb;
    s->trace_user = user;
    return PF_OK;
}

const char *pf_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------ stages

pf_status pf_stage_ybus(pf_solver *s, int *rowptr, int *colind, double *y, int *nnz) {
    if (!s) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    if (nnz) *nnz = s->yp.nnz;
    if (rowptr) std::copy(s->yp.rowptr.begin(), s->yp.rowptr.end(), rowptr);
    if (colind) std::copy(s->yp.colind.begin(), s->yp.colind.end(), colind);
    if (y) PFCHK(s->get(y, (const double *)s->dn.y, 2 * (size_t)s->yp.nnz));
    return PF_OK;
}

static pf_status stage_state(pf_solver *s, const double *V, const double *theta,
                             const unsigned char *limQ, const double *QliqLim) {
    HIPCHK(hipSetDevice(s->device));
    PFCHK(s->reset_ctl());
    PFCHK(s->put(s->ds.V, V, s->net.nB));
    PFCHK(s->put(s->ds.th, theta, s->net.nB));
    PFCHK(s->set_isq_from_limq(limQ));
    if (QliqLim) PFCHK(s->put(s->ds.QliqLim, QliqLim, s->net.nB));
    k_init_E<<<(s->net.nB + 255) / 256, 256, 0, s->stream>>>(s->dn, s->ds);
    k_pq<4><<<(s->net.nB + 63) / 64, 256, 0, s->stream>>>(s->dn, s->ds, 0);
    HIPCHK(hipGetLastError());
    return PF_OK;
}

// Branch flows at the device state (V, theta, Ybus).  The per-branch lookups
// follow the reference: phif(k,m) (CPU/phifBshf.cpp:31-46 over the triplet
// matrix of CPU/estrutura/ramo.cpp:4-13, duplicates summed) is the phase of
// the branches stored as (k,m), else minus that of (m,k), else 0; bshf(k,m)
// (CPU/phifBshf.cpp:48-63) is the charging of the first branch joining k, m.
pf_status pf_solver::flows(double *Pdp, double *Qdp, double *Ppd, double *Qpd) {
    const int nL = net.nL;
    if (nL == 0) return PF_OK;
    if (!flow_br) {
        std::map<std::pair<int, int>, double> phisum;
        std::map<std::pair<int, int>, int> first;
        for (int l = 0; l < nL; l++) {
            phisum[{net.from[l], net.to[l]}] += net.phi[l];
            auto key = std::minmax(net.from[l], net.to[l]);
            first.emplace(std::make_pair(key.first, key.second), l);
        }
        auto phif = [&](int k, int m) {
            auto it = phisum.find({k, m});
            if (it != phisum.end()) return it->second;
            it = phisum.find({m, k});
            return it != phisum.end() ? -it->second : 0.0;
        };
        auto bshf = [&](int k, int m) {
            auto key = std::minmax(k, m);
            auto it = first.find(std::make_pair(key.first, key.second));
            return it != first.end() ? net.bch[it->second] : 0.0;
        };
        std::vector<FlowBranch> h(nL);
        for (int l = 0; l < nL; l++) {
            const int f = net.from[l], t = net.to[l];
            h[l] = FlowBranch{f, t, yp.br_ft[l], yp.br_tf[l], phif(f, t), phif(t, f), bshf(f, t), bshf(t, f)};
        }
        PFCHK(upload(&flow_br, h));
        PFCHK(alloc(&flow_out, 4 * (size_t)nL));
    }
    k_flows<<<(nL + 255) / 256, 256, 0, stream>>>(dn, ds, flow_br, nL, flow_out);
    HIPCHK(hipGetLastError());
    if (Pdp) PFCHK(get(Pdp, flow_out, nL));
    if (Qdp) PFCHK(get(Qdp, flow_out + nL, nL));
    if (Ppd) PFCHK(get(Ppd, flow_out + 2 * (size_t)nL, nL));
    if (Qpd) PFCHK(get(Qpd, flow_out + 3 * (size_t)nL, nL));
    HIPCHK(hipStreamSynchronize(stream));
    return PF_OK;
}

pf_status pf_solver_flows(pf_solver *s, double *Pdp, double *Qdp, double *Ppd, double *Qpd) {
    if (!s) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    auto t0 = std::chrono::steady_clock::now();   // the reference's fluxo phase (CPU/main.cpp:103-118)
    pf_status st = s->flows(Pdp, Qdp, Ppd, Qpd);
    if (s->timing) s->times.flows_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return st;
}

pf_status pf_stage_flows(pf_solver *s, const double *V, const double *theta, double *Pdp, double *Qdp,
                         double *Ppd, double *Qpd) {
    if (!s || !V || !theta) return PF_INVALID_ARG;
    PFCHK(stage_state(s, V, theta, nullptr, nullptr));
    return s->flows(Pdp, Qdp, Ppd, Qpd);
}

pf_status pf_stage_pq(pf_solver *s, const double *V, const double *theta, double *P, double *Q) {
    if (!s || !V || !theta) return PF_INVALID_ARG;
    PFCHK(stage_state(s, V, theta, nullptr, nullptr));
    if (P) PFCHK(s->get(P, s->ds.P, s->net.nB));
    if (Q) PFCHK(s->get(Q, s->ds.Q, s->net.nB));
    return PF_OK;
}

// reference layout of the reduced system: rows/cols [non-slack buses][PQlim buses]
static void reduced_lists(const pf_solver *s, const unsigned char *limQ, std::vector<int> &pqlim) {
    std::vector<unsigned char> isq(s->net.nB, 0);
    for (int k : s->net.pq) isq[k] = 1;
    if (limQ)
        for (int i = 0; i < s->net.nPV; i++) isq[s->net.pv[i]] = limQ[i];
    pqlim.clear();
    for (int k = 0; k < s->net.nB; k++)
        if (isq[k]) pqlim.push_back(k);
}

pf_status pf_stage_mismatch(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, const double *QliqLim, double *g,
                            double *max_abs) {
    if (!s || !V || !theta || !QliqLim) return PF_INVALID_ARG;
    PFCHK(stage_state(s, V, theta, limQ, QliqLim));
    k_mismatch<<<(s->net.nB + 255) / 256, 256, 0, s->stream>>>(s->dn, s->ds, 0);
    HIPCHK(hipGetLastError());
    std::vector<double> gn(2 * (size_t)s->sym.n);
    PFCHK(s->get(gn.data(), s->ds.g, gn.size()));
    Ctl c;
    PFCHK(s->get(&c, s->ctl, 1));
    double m;
    std::memcpy(&m, &c.maxG_bits, sizeof(m));
    if (max_abs) *max_abs = m;
    if (g) {
        std::vector<int> pqlim;
        reduced_lists(s, limQ, pqlim);
        const int nB = s->net.nB, sl = s->net.slack;
        std::vector<int> node(nB, -1);
        for (int k = 0; k < nB; k++)
            if (k != sl) node[k] = s->sym.iperm[k - (k > sl)];
        for (int k = 0; k < nB; k++)
            if (k != sl) g[k - (k > sl)] = gn[2 * node[k]];
        for (size_t i = 0; i < pqlim.size(); i++) g[nB - 1 + i] = gn[2 * node[pqlim[i]] + 1];
    }
    return PF_OK;
}

pf_status pf_stage_jacobian(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, int *nJ, int *rowptr, int *col,
                            double *val, int cap, int *nnz) {
    if (!s || !V || !theta) return PF_INVALID_ARG;
    PFCHK(stage_state(s, V, theta, limQ, nullptr));
    const size_t ns = s->sym.nslot();
    k_jfill<double><<<(ns + 255) / 256, 256, 0, s->stream>>>(s->dn, s->ds, s->dy, s->blk64, 0);
    HIPCHK(hipGetLastError());
    std::vector<double> b(4 * ns);
    PFCHK(s->get(b.data(), s->blk64, b.size()));
    std::vector<int> pqlim;
    reduced_lists(s, limQ, pqlim);
    const int nB = s->net.nB, sl = s->net.slack;
    std::vector<int> pos(nB, -1);
    for (size_t i = 0; i < pqlim.size(); i++) pos[pqlim[i]] = (int)i;
    const auto &yp = s->yp;
    int count = 0, row = 0;
    if (rowptr) rowptr[0] = 0;
    auto push = [&](int c, double v) {
        if (count < cap && col && val) { col[count] = c; val[count] = v; }
        count++;
    };
    auto block = [&](int e) { return &b[4 * (size_t)s->e2slot[e]]; };
    for (int k = 0; k < nB; k++) {          // rows dP_k: H then N (CPU/esparso/spCalcJac.cpp:409-504)
        if (k == sl) continue;
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (m == sl) continue;
            push(m - (m > sl), block(e)[0]);
        }
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (pos[m] < 0) continue;
            push(nB - 1 + pos[m], block(e)[1]);
        }
        if (rowptr) rowptr[++row] = count;
    }
    for (int k : pqlim) {                    // rows dQ_k: M then L (:506-553)
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (m == sl) continue;
            push(m - (m > sl), block(e)[2]);
        }
        for (int e = yp.rowptr[k]; e < yp.rowptr[k + 1]; e++) {
            int m = yp.colind[e];
            if (pos[m] < 0) continue;
            push(nB - 1 + pos[m], block(e)[3]);
        }
        if (rowptr) rowptr[++row] = count;
    }
    if (nJ) *nJ = nB - 1 + (int)pqlim.size();
    if (nnz) *nnz = count;
    return count <= cap ? PF_OK : PF_INVALID_ARG;
}

pf_status pf_stage_lu_solve(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, const double *rhs, double *x) {
    if (!s || !V || !theta || !rhs || !x) return PF_INVALID_ARG;
    PFCHK(stage_state(s, V, theta, limQ, nullptr));
    std::vector<int> pqlim;
    reduced_lists(s, limQ, pqlim);
    const int nB = s->net.nB, sl = s->net.slack, n = s->sym.n;
    std::vector<int> node(nB, -1);
    for (int k = 0; k < nB; k++)
        if (k != sl) node[k] = s->sym.iperm[k - (k > sl)];
    std::vector<double> gn(2 * (size_t)n, 0.0);
    for (int k = 0; k < nB; k++)
        if (k != sl) gn[2 * node[k]] = rhs[k - (k > sl)];
    for (size_t i = 0; i < pqlim.size(); i++) gn[2 * node[pqlim[i]] + 1] = rhs[nB - 1 + i];
    PFCHK(s->put(s->ds.g, gn.data(), gn.size()));
    // force the gated kernels of one step to run
    Ctl c{};
    c.step = 1;
    c.tol = s->opt.tolerancia;
    s->ctl_host[0] = c;
    PFCHK(s->put(s->ctl, s->ctl_host, 1));
    std::vector<double> xn(2 * (size_t)n);
    if (s->path() == PF_HIBRIDO_C) {
        const size_t ns = s->sym.nslot();
        k_jfill<double><<<(ns + 255) / 256, 256, 0, s->stream>>>(s->dn, s->ds, s->dy, s->blk64, 0);
        HIPCHK(hipGetLastError());
        std::vector<double> b(4 * (size_t)s->sym.nblk());
        PFCHK(s->get(b.data(), s->blk64, 4 * ns));
        if (!host_block_factor(s->sym, b.data())) return PF_SINGULAR;
        xn = gn;
        host_block_solve(s->sym, b.data(), xn.data());
    } else {
        PFCHK(s->enqueue_factor_solve(1));
        PFCHK(s->get(xn.data(), s->ds.x, xn.size()));
        PFCHK(s->get(&c, s->ctl, 1));
        if (c.singular) return PF_SINGULAR;
    }
    for (int k = 0; k < nB; k++)
        if (k != sl) x[k - (k > sl)] = xn[2 * node[k]];
    for (size_t i = 0; i < pqlim.size(); i++) x[nB - 1 + i] = xn[2 * node[pqlim[i]] + 1];
    return PF_OK;
}

pf_status pf_stage_limits(pf_solver *s, int mode, const double *V, const double *Q,
                          unsigned char *limQ, double *QliqLim, int *flips) {
    if (!s || !V || !Q || !limQ || !QliqLim || (mode != 0 && mode != 1)) return PF_INVALID_ARG;
    HIPCHK(hipSetDevice(s->device));
    PFCHK(s->reset_ctl());
    PFCHK(s->put(s->ds.V, V, s->net.nB));
    PFCHK(s->put(s->ds.Q, Q, s->net.nB));
    PFCHK(s->set_isq_from_limq(limQ));
    PFCHK(s->put(s->ds.QliqLim, QliqLim, s->net.nB));
    k_limits<<<1, 1024, 0, s->stream>>>(s->dn, s->ds, mode == 0 ? 2 : 3, 1);
    HIPCHK(hipGetLastError());
    PFCHK(s->get(limQ, s->ds.limQ, s->net.nPV));
    PFCHK(s->get(QliqLim, s->ds.QliqLim, s->net.nB));
    Ctl c;
    PFCHK(s->get(&c, s->ctl, 1));
    if (flips) *flips = c.flips;
    return PF_OK;
}

}  // extern "C"
