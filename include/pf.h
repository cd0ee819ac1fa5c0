/*
 * pf.h -- C ABI of the MI355X Newton-Raphson power-flow library (libpf.so).
 *
 * Plain C types only: ints, doubles, pointers and sizes.  Every entry point
 * returns a pf_status; no entry point exits, asserts or throws.
 *
 * Reference interfaces each entry point replaces (joaovdd/Experience-of-Using-
 * GPU-in-Power-Flow-Computation, paths relative to the repository root):
 *
 *   pf_options_parse / pf_options_load   loadFile()          CPU/leitura/config.cpp:5-166,
 *                                                            GPU/sistInfo.h:62-242
 *   pf_read_case                         lerArquivoEAlocarMemoria() CPU/leitura/arquivo.cpp:3-51
 *                                        + readCDF/readCDFX  CPU/leitura/cdf.cpp:163-493
 *   pf_solver_create                     calcYbusSp_Matpower() CPU/esparso/spCalcYbus.cpp:4-90,
 *                                        initGPU/d_init*      GPU/sistema.cuh:12-469,
 *                                        initLim             CPU/limInjReat.cpp:177-207
 *   pf_solve                             nR()                CPU/newtonRaphson.cpp:173-314,
 *                                        mNR()               GPU/newtonRaphson.cuh:171-623
 *   pf_stage_ybus                        calcYbusSp_Matpower CPU/esparso/spCalcYbus.cpp:4-90
 *   pf_stage_pq                          spCalcP/spCalcQ     CPU/esparso/potencia/spCalcPQliq.cpp:3-45,
 *                                        calcP/Qeficiente_0based GPU/PQcalc.cuh:177-219
 *   pf_stage_mismatch                    calcResLim          CPU/limInjReat.cpp:158-175,
 *                                        calcResLimf         GPU/newtonRaphson.cuh:35-74
 *   pf_stage_jacobian                    spCalcJ_eficiente   CPU/esparso/spCalcJac.cpp:402-568,
 *                                        SpCalcJac           GPU/Jacobiano.cuh:382-415
 *   pf_stage_lu_solve                    solver()/spSolveMKL CPU/solver.cpp:3-24,
 *                                        CPU/esparso/spSolver.cpp:84-189; cusolverSpDcsrlsvqr
 *                                        GPU/newtonRaphson.cuh:466-468
 *   pf_stage_limits                      chkLimQ/undo_chkLimQ/geraVetPV/geraVetPQ
 *                                        CPU/limInjReat.cpp:3-131, CPU/newtonRaphson.cpp:64-154
 *   pf_solver_flows / pf_stage_flows     SpCalcFlux          CPU/esparso/potencia/spCalcFluxPQ.cpp:3-36,
 *                                        phif/bshf           CPU/phifBshf.cpp:31-63
 *   pf_batch_*                           (no counterpart: the reference solves one case per run)
 *   pf_warmstart_*                       (no counterpart: learned initial state)
 */
#ifndef PF_H
#define PF_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum pf_status {
    PF_OK = 0,
    PF_NOT_CONVERGED = 1,   /* no_max_iter reached (the reference only printed this) */
    PF_SINGULAR = 2,        /* zero or non-finite pivot in the Jacobian factorisation */
    PF_PARSE_ERROR = 3,     /* malformed CDF/CDFX or data.txt                         */
    PF_DEVICE_ERROR = 4,    /* HIP runtime failure, or no GPU                          */
    PF_INVALID_ARG = 5,
    PF_IO_ERROR = 6         /* file cannot be opened                                  */
} pf_status;

/* metodo (CPU/estrutura/metodo.h:2-10, GPU/sistInfo.h:94-118) */
typedef enum pf_metodo {
    PF_ESPARSO = 0,         /* GPU Jacobian + GPU sparse LU                           */
    PF_DENSO = 1,           /* GPU Jacobian + GPU dense LU with partial pivoting      */
    PF_HIBRIDO_C = 2,       /* GPU Jacobian + host sparse LU (the reference: PARDISO) */
    PF_HIBRIDO_A = 3,       /* the reference's Ybus-to-CSR-on-GPU variant: here the sparse
                               GPU path (the reference takes no linear solve for it,
                               GPU/newtonRaphson.cuh:452-535; this library solves J dx = g) */
    PF_HIBRIDO_B = 4        /* GPU sparse Jacobian, made dense, GPU dense LU
                               (GPU/newtonRaphson.cuh:480-530): the denso path here     */
} pf_metodo;

typedef enum pf_precision {
    PF_FP64 = 0,            /* FP64 factor                                            */
    PF_MIXED = 1            /* FP32 factor + FP64 iterative refinement                */
} pf_precision;

/* data.txt keys; the last three are this library's own additions. */
typedef struct pf_options {
    int metodo;             /* pf_metodo                                              */
    int lim_inj_reat;       /* PV->PQ switching (default OFF, CPU/global.cpp:14)      */
    int openmp;
    double v_inicial, theta_inicial, tolerancia;
    int no_max_iter;
    char arq_entrada[1024];
    int verbose_mode, laconic_mode;
    int output_bench;       /* 0 ALL, 1 SCREEN, 2 FILE                                */
    int output_ans, output_iter, streams;
    int precision;          /* pf_precision                                           */
    int refine_steps;       /* FP64 refinement steps after an FP32 solve (default 2)  */
    int use_graph;          /* 1: one captured hipGraph per NR iteration, no per-iteration host sync */
    int lu_blocked;         /* sparse LU schedule.  1 (default): the subtrees below the narrow
                               top of the elimination tree in one launch (a workgroup per group
                               of subtrees), the narrow top in one workgroup, both as flat lists
                               of update pairs level by level, and the tree's trailing dense
                               chain as one dense block (Gauss-Jordan inverse of its Schur
                               complement); 0: one launch per elimination-tree level, node by
                               node, and the narrow top in one workgroup                       */
} pf_options;

const char *pf_status_str(pf_status s);

void pf_options_default(pf_options *o);
/* cpu_layout = 1 prefixes "sistemas/" to arq_entrada, as the CPU build does
 * (CPU/leitura/config.cpp:108); 0 keeps it as written (GPU/sistInfo.h:153). */
pf_status pf_options_parse(const char *text, int cpu_layout, pf_options *out);
pf_status pf_options_load(const char *path, int cpu_layout, pf_options *out);

/* ---------------------------------------------------------------- network */
typedef struct pf_network pf_network;

typedef struct pf_network_info {
    int nB, nL, nPV, nPQ;
    int slack;              /* 0-based index of the Vtheta bus                        */
    int n_duplicates;       /* parallel branches merged by the reader                 */
    double baseMVA;
} pf_network_info;

enum pf_bus_field {
    PF_BUS_V = 0, PF_BUS_THETA, PF_BUS_PLIQ, PF_BUS_QLIQ, PF_BUS_PLOAD, PF_BUS_QLOAD,
    PF_BUS_PG, PF_BUS_QG, PF_BUS_VBASE, PF_BUS_GSH, PF_BUS_BSH, PF_BUS_NFIELDS
};

/* By extension (CPU/leitura/arquivo.cpp:3-51): CDF (.cdf), CDFX (.cdfx),
 * MATPOWER (.m) or ANAREDE (.pwf); v0/theta0 = flat-start values.           */
pf_status pf_read_case(const char *path, double v0, double theta0, pf_network **out);
void pf_network_free(pf_network *net);
pf_status pf_network_get_info(const pf_network *net, pf_network_info *info);
pf_status pf_network_get_bus(const pf_network *net, int field, double *out /* nB */);
/* bus numbers as written in the file [nB] */
pf_status pf_network_get_ids(const pf_network *net, int *id);
/* any pointer may be NULL */
pf_status pf_network_get_lists(const pf_network *net, int *pv, int *pq, double *qmin,
                               double *qmax, double *vset);
pf_status pf_network_get_branches(const pf_network *net, int *from, int *to, double *r,
                                  double *x, double *bch, double *tap_re, double *tap_im);

/* ----------------------------------------------------------------- solver */
typedef struct pf_solver pf_solver;

typedef struct pf_result {
    int iterations;         /* updates applied (the reference's iterativo->iteracao)  */
    int converged;
    int singular;
    double max_mismatch;    /* max |g| at the last mismatch evaluation (pu)           */
    int n_switch_events;    /* PV<->PQ flips over the run                             */
    double solve_ms;        /* wall time of the NR loop (host clock)                   */
} pf_result;

/* Builds Ybus on the GPU, the ordering and symbolic factorisation (once),
 * uploads everything to `device` and sets the flat-start state. */
pf_status pf_solver_create(const pf_network *net, const pf_options *opt, int device,
                           pf_solver **out);
void pf_solver_destroy(pf_solver *s);
/* NR from the current state (flat start after create / reset). */
pf_status pf_solve(pf_solver *s, pf_result *res);
pf_status pf_solver_reset(pf_solver *s);                 /* back to flat start */
pf_status pf_solver_set_state(pf_solver *s, const double *V, const double *theta);
pf_status pf_solver_set_injections(pf_solver *s, const double *Pliq, const double *Qliq,
                                   const double *Qload);
pf_status pf_solver_get_state(pf_solver *s, double *V, double *theta,
                              unsigned char *limQ /* nPV, may be NULL */);
/* sizes of the factorisation: n = nB-1 block unknowns, nnz of strict L in
 * 2x2 blocks, elimination-tree levels, levels run by one workgroup */
pf_status pf_solver_lu_info(pf_solver *s, int *n, int *nnzL, int *levels, int *tail_levels);

/* ------------------------------------------------ timing and iteration trace */
/* The reference's phase timers (timerType, CPU/benchmarks.h:41-89; phases
 * timed in CPU/main.cpp:41-118 and CPU/newtonRaphson.cpp:179-305), in ms:
 * Ybus assembly at pf_solver_create, then per pf_solve summed over the
 * passes (P/Q, Jacobian, linear system with its refinement, state update),
 * the whole NR loop, and the last pf_solver_flows.  GPU phases are timed
 * with HIP events, so timing runs the NR loop with one host sync per pass
 * (use_graph is ignored while it is on).                                    */
typedef struct pf_timing {
    double ybus_ms, pq_ms, jacobian_ms, linear_ms, update_ms, iterative_ms, flows_ms;
    int passes;             /* NR passes timed (updates applied + the last check)  */
} pf_timing;
pf_status pf_solver_set_timing(pf_solver *s, int on);
pf_status pf_solver_get_timing(pf_solver *s, pf_timing *t);
/* Called after every NR pass of pf_solve (one host sync per pass while set):
 * the pass, max|g| and the sum of g at its mismatch evaluation
 * (CPU/newtonRaphson.cpp:248-251), whether it is a switch taken at
 * convergence (CPU/newtonRaphson.cpp:253-257), and its PV<->PQ flips: PV
 * index and kind (0 PV->PQ at Qmin, 1 PV->PQ at Qmax, 2 PQ at Qmax -> PV,
 * 3 PQ at Qmin -> PV; the messages of CPU/limInjReat.cpp:46-80), and the
 * pass's stage times in ms (P/Q, Jacobian, linear system, update; zeros
 * when timing is off).  NULL: off.                                          */
typedef void (*pf_pass_callback)(void *user, int pass, double max_g, double sum_g, int at_convergence,
                                 int n_flips, const int *pv_index, const int *kind, const double *stage_ms);
pf_status pf_solver_set_trace(pf_solver *s, pf_pass_callback cb, void *user);

/* ---------------------------------------------------- stage entry points */
/* Host buffers in and out; each stage runs its own kernels on the GPU.      */
/* Ybus CSR (0-based, sorted columns), y interleaved re/im: 2*nnz doubles.   */
pf_status pf_stage_ybus(pf_solver *s, int *rowptr, int *colind, double *y, int *nnz);
pf_status pf_stage_pq(pf_solver *s, const double *V, const double *theta, double *P,
                      double *Q);
/* Branch flows as the reference reports them, per branch [nL]: Pdp/Qdp from
 * -> to, Ppd/Qpd to -> from (any output may be NULL).  The formula and its
 * lookups are the reference's, including the sign of the flows and the full
 * line charging in Q.  pf_solver_flows: at the solver's current state.      */
pf_status pf_stage_flows(pf_solver *s, const double *V, const double *theta, double *Pdp,
                         double *Qdp, double *Ppd, double *Qpd);
pf_status pf_solver_flows(pf_solver *s, double *Pdp, double *Qdp, double *Ppd, double *Qpd);
/* limQ [nPV] and QliqLim [nB] give the PV->PQ state; g gets nB-1+nPQlim
 * entries in the reference's order; returns max|g|.                         */
pf_status pf_stage_mismatch(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, const double *QliqLim, double *g,
                            double *max_abs);
/* Jacobian in the reference's reduced CSR layout (rows: P of the non-slack
 * buses, Q of barrasPQlim; columns: theta of the non-slack buses, |V| of
 * barrasPQlim).  rowptr has nJ+1 entries, col/val hold nnz (cap entries).   */
pf_status pf_stage_jacobian(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, int *nJ, int *rowptr, int *col,
                            double *val, int cap, int *nnz);
/* factor the Jacobian at (V, theta, limQ) and solve J x = rhs (reduced layout) */
pf_status pf_stage_lu_solve(pf_solver *s, const double *V, const double *theta,
                            const unsigned char *limQ, const double *rhs, double *x);
/* one limInjReat decision pass (mode 0: chkLimQ, 1: undo_chkLimQ) at state
 * (V, Q); limQ/QliqLim in/out; returns the number of flips.                 */
pf_status pf_stage_limits(pf_solver *s, int mode, const double *V, const double *Q,
                          unsigned char *limQ, double *QliqLim, int *flips);

/* text of the last HIP runtime failure seen by this thread ("" if none)      */
const char *pf_last_error(void);

/* ------------------------------------------------------ host-only probes */
/* The ordering / symbolic factor program without a GPU (CPU tests, and the
 * host factorisation behind metodo hibridoC).  Slots: [0,n) diagonal blocks,
 * [n, n+nnzL) strict-lower blocks, [n+nnzL, n+2nnzL) strict-upper blocks.   */
typedef struct pf_symbolic_probe pf_symbolic_probe;
pf_status pf_probe_create(const pf_network *net, pf_symbolic_probe **out, int *n, int *nnzL,
                          int *nslot, int *levels);
void pf_probe_free(pf_symbolic_probe *p);
/* perm[j] = block (bus order without the slack) eliminated j-th; slot_e[s] =
 * Ybus entry whose Jacobian block lands in slot s (-1: fill-in)             */
pf_status pf_probe_get(const pf_symbolic_probe *p, int *perm, int *slot_e, int *rowptr,
                       int *colind, int *ynnz);
/* elimination-tree levels: lvl_ptr [levels+1] (nodes of level l are the
 * positions lvl_ptr[l] .. lvl_ptr[l+1]-1), column pointers of L [n+1]      */
pf_status pf_probe_levels(const pf_symbolic_probe *p, int *lvl_ptr, int *colptr);
/* trailing dense chain of the elimination tree (nodes c0 .. n-1, one per
 * level, full columns) that the solver factors as one dense block; m = 0:
 * none (shorter than 4 nodes)                                               */
pf_status pf_probe_chain(const pf_symbolic_probe *p, int *m, int *c0);
/* groups of whole subtrees below the narrow top of the tree, one workgroup
 * each (as the solver builds them for 256 CUs): group g owns the nodes
 * part_nodes[part_ptr[g*part_levels + l] .. part_ptr[g*part_levels + l + 1])
 * of level l; part_ptr has n_parts*part_levels + 1 entries, part_nodes one
 * per node below the top; parent: elimination tree (n, -1 at the root).
 * Any pointer may be NULL.                                                  */
pf_status pf_probe_parts(const pf_symbolic_probe *p, int *n_parts, int *part_levels, int *part_ptr,
                         int *part_nodes, int *parent);
/* blocks: nslot 2x2 blocks (row-major), x: 2n right-hand side in node order */
pf_status pf_probe_factor_solve(const pf_symbolic_probe *p, double *blocks, double *x);
/* The same factor and solve in the order of the GPU's blocked program (the
 * subtree groups' and the narrow top's flat pair programs, the narrow top's
 * split sweeps), restated on the host.                                      */
pf_status pf_probe_flat_factor_solve(const pf_symbolic_probe *p, double *blocks, double *x);

/* ---------------------------------------------------------------- batches */
/* Many scenarios of one or more topologies solved in one persistent launch:
 * one workgroup owns one scenario for its whole NR solve.                   */
typedef struct pf_batch pf_batch;

pf_status pf_batch_create(const pf_network *const *nets, int n_nets, const pf_options *opt,
                          int device, pf_batch **out);
void pf_batch_destroy(pf_batch *b);
/* scenario s uses network net_of[s]; Pliq/Qliq/Qload rows of that network's
 * nB, packed back to back in scenario order.  V0/theta0 (may be NULL = flat
 * start) are packed the same way. */
pf_status pf_batch_set_scenarios(pf_batch *b, int n_scen, const int *net_of,
                                 const double *Pliq, const double *Qliq, const double *Qload,
                                 const double *V0, const double *theta0);
pf_status pf_batch_solve(pf_batch *b, double *elapsed_ms);
pf_status pf_batch_get(pf_batch *b, double *V, double *theta, int *iterations,
                       int *converged, double *max_mismatch);
/* The PV->PQ decision state at the end of each scenario's solve: limQ packed
 * per scenario (nPV of its network, scenarios in order; the reference's
 * iterativo->limQ, CPU/limInjReat.cpp:38-91) and the PV<->PQ flips of that
 * solve (flat-start fallback: of the fallback solve).  Either may be NULL.  */
pf_status pf_batch_get_limq(pf_batch *b, unsigned char *limQ, int *flips);
/* Allocation-free batched solve.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  Solves n_scen scenarios of
 * the batch's networks with every input and output in caller device memory,
 * packed as in pf_batch_set_scenarios / pf_batch_get / pf_batch_get_limq:
 * net_of [n_scen], Pliq / Qliq / Qload (and optionally V0 / theta0) [sum nB],
 * V / theta [sum nB], iterations / converged / flips [n_scen],
 * max_mismatch [n_scen], limQ [sum nPV].  Per-call tables, work counters and
 * the per-workgroup working-set slices live in a fixed device array of
 * 64 MiB per GPU (held until the process ends); a call whose tables and
 * slices do not fit is refused with PF_INVALID_ARG.  The call allocates
 * nothing.  stream NULL: the batch's own stream, and the call waits for the
 * results (an index outside the networks then gives PF_INVALID_ARG);
 * otherwise a hipStream_t: the work is queued there and not waited for.
 * Calls that overlap on one GPU race on the fixed array.                    */
pf_status pf_batch_solve_noalloc(pf_batch *b, int n_scen, const int *net_of, const double *Pliq,
                                 const double *Qliq, const double *Qload, const double *V0,
                                 const double *theta0, double *V, double *theta, int *iterations,
                                 int *converged, double *max_mismatch, unsigned char *limQ, int *flips,
                                 void *stream);
/* Autotuner (no counterpart in the reference): for every topology of the
 * current scenario set, times one launch of its persistent kernel for each
 * candidate launch shape (working set fully in LDS / step vectors and level
 * metadata in LDS / all global; workgroup size; resident workgroups per CU),
 * keeps the fastest (minimum over `reps` timed launches after one warm-up)
 * and appends it to the text cache at cache_path (may be NULL: no cache).
 * A shape found in the cache for the same device, topology sizes, options and
 * scenario count is applied without measuring.  Outputs of the batch are
 * overwritten; call pf_batch_solve again afterwards.                        */
pf_status pf_batch_autotune(pf_batch *b, const char *cache_path, int reps, int *n_measured);
/* fix the launch shape of network `net` (after pf_batch_set_scenarios):
 * mode 2 needs the working set to fit in LDS, 1 the step vectors and level
 * metadata, 0 always fits; block a multiple of 64 up to 1024 (modes 1 and
 * 0: at least 256); per_cu <= 0:
 * as many workgroups per CU as fit.  PF_INVALID_ARG when it cannot run.     */
pf_status pf_batch_set_shape(pf_batch *b, int net, int mode, int block, int per_cu);
/* launch shape in use for network `net`: kernel mode (2 LDS, 1 partial LDS,
 * 0 global), workgroup size, workgroups, dynamic LDS bytes                   */
pf_status pf_batch_get_shape(pf_batch *b, int net, int *mode, int *block, int *grid,
                             long long *lds_bytes);

/* ------------------------------------------------------------ warm start */
/* MLP: injections + setpoints -> |V|, theta.  BF16 weights/activations,
 * FP32 accumulation on the matrix cores, FP32 master weights and AdamW.
 * Parameter layout: per layer W [K x N] row-major then bias [N].
 * n_in, n_out and hidden must be multiples of 8 (16-byte rows of bf16);
 * pad features and targets with zero columns otherwise.                    */
typedef struct pf_warmstart pf_warmstart;
pf_status pf_warmstart_create(int n_in, int n_out, int hidden, int n_hidden_layers,
                              unsigned long long seed, int device, pf_warmstart **out);
void pf_warmstart_destroy(pf_warmstart *w);
pf_status pf_warmstart_info(pf_warmstart *w, size_t *n_params, int *n_layers);
/* input/target normalisation (x' = (x - mean) / std; std <= 0 means 1)     */
pf_status pf_warmstart_set_norm(pf_warmstart *w, const float *x_mean, const float *x_std,
                                const float *y_mean, const float *y_std);
pf_status pf_warmstart_get_params(pf_warmstart *w, float *out);
pf_status pf_warmstart_set_params(pf_warmstart *w, const float *in);
/* forward + backward of the mean squared error on normalised targets       */
pf_status pf_warmstart_grad(pf_warmstart *w, const float *X, const float *Y, int B,
                            float *grads, double *loss);
/* AdamW over a resident data set (n rows), minibatches of B (B % 8 == 0)   */
pf_status pf_warmstart_fit(pf_warmstart *w, const float *X, const float *Y, int n, int epochs,
                           int B, float lr, float weight_decay, unsigned long long seed,
                           double *epoch_loss);
pf_status pf_warmstart_predict(pf_warmstart *w, const float *X, int B, float *Y);
/* device-resident inference into a batch: pf_warmstart_stage copies B rows
 * of features to the GPU once (like pf_batch_set_scenarios, outside a timed
 * solve); pf_warmstart_predict_into runs the network on them and writes, for
 * batch scenario s (B must equal the batch's scenario count) and bus i of
 * its network, |V| = Y[s][v_col + i] and theta = Y[s][th_col + i] as the
 * scenario's starting state, with no host round trip.  The batch keeps that
 * start until pf_batch_set_scenarios is called again.  Returns after the
 * start is written.  A later predict, grad or fit call replaces the staged
 * rows (predict_into then fails until the next stage).                     */
pf_status pf_warmstart_stage(pf_warmstart *w, const float *X, int B);
pf_status pf_warmstart_predict_into(pf_warmstart *w, pf_batch *b, int v_col, int th_col);
/* one product through the training GEMM kernel: C = op(A) op(B)            */
pf_status pf_warmstart_gemm_probe(pf_warmstart *w, int ta, int tb, int M, int N, int K,
                                  const float *A, const float *B, float *C);

#ifdef __cplusplus
}
#endif
#endif
