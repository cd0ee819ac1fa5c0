/*
 * pf_oracle.h -- FP64 CPU restatement of the reference's CPU sparse
 * Newton-Raphson power-flow path (metodo = esparso).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Deviations from
 * the reference are listed at the top of pf_oracle.c.
 */
#ifndef PF_ORACLE_H
#define PF_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Network as the reference's reader leaves it (CPU/estrutura/sistema.h, barra.h, ramo.h), 0-based. */
typedef struct {
    int nB, nL, nPV, nPQ;
    int slack;            /* 0-based index of the Vtheta bus (barraVO - 1)    */
    int n_duplicates;     /* parallel branches merged by the reader           */
    double baseMVA;
    int *id;              /* [nB] bus numbers from the file                    */
    double *V, *theta;    /* [nB] initial state (flat start + PV/slack fields) */
    double *Pliq, *Qliq, *Pload, *Qload, *Pg, *Qg, *Vbase, *gsh, *bsh; /* [nB] */
    int *pv, *pq;         /* [nPV], [nPQ] 0-based bus indices, file order      */
    double *Qmin, *Qmax, *Vset; /* [nPV]                                       */
    int *from, *to;       /* [nL] 0-based                                      */
    double *r, *x, *bch, *tap_re, *tap_im, *phi; /* [nL]                       */
} orc_case;

typedef struct {
    double tol;           /* tolerancia                                        */
    int max_iter;         /* no_max_iter                                       */
    int lim_inj_reat;     /* PV->PQ switching on/off                           */
    int openmp;           /* OpenMP loops on/off (CPU/leitura/config.cpp:66)   */
    int reuse_symbolic;   /* 0 = ordering+symbolic every iteration (reference) */
} orc_opts;

typedef struct {
    int iterations;       /* updates applied (reference iterativo->iteracao)   */
    int converged;        /* 1 if max|mismatch| <= tol with a stable limQ set  */
    int singular;         /* 1 if a zero pivot stopped the solve               */
    double max_mismatch;  /* max|g| at the last mismatch evaluation            */
    int n_switch_events;  /* number of PV<->PQ flag flips over the run         */
} orc_result;

/* --- reader (CPU/leitura/cdf.cpp, CPU/leitura/id2i.cpp) --------------------- */
/* wide = 0: CDF columns; 1: CDFX columns; -1: choose by extension.           */
int  orc_case_read(const char *path, int wide, double v0, double theta0, orc_case **out);
void orc_case_free(orc_case *c);
/* copies the case's arrays into caller buffers (any pointer may be NULL)     */
void orc_case_get(const orc_case *c, int *ints6, double *bus11, int *pv, int *pq,
                  double *pvlim3, int *br_ft, double *br6);

/* --- Ybus, MATPOWER form (CPU/esparso/spCalcYbus.cpp:4-90) ------------------- */
int  orc_ybus_nnz(const orc_case *c);
void orc_ybus(const orc_case *c, int *rowptr, int *colind, double *y /* 2*nnz */);

/* --- injections (CPU/esparso/potencia/spCalcPQliq.cpp:3-45) ------------------ */
void orc_pq(int nB, const int *rowptr, const int *colind, const double *y,
            const double *V, const double *theta, double *P, double *Q, int openmp);

/* --- mismatch (CPU/limInjReat.cpp:158-175) ----------------------------------- */
void orc_mismatch(int nB, int slack, const double *Pliq, const double *QliqLim,
                  const double *P, const double *Q, int nPQlim, const int *pqlim,
                  double *g);

/* --- Jacobian (CPU/esparso/spCalcJac.cpp:402-568, isPQ fixed) ---------------- */
/* returns nnzJ; rowptrJ has nB-1+nPQlim+1 entries; colJ/valJ capacity cap.    */
int  orc_jacobian(int nB, int slack, const int *rowptr, const int *colind,
                  const double *y, const double *V, const double *theta,
                  const double *P, const double *Q, int nPQlim, const int *pqlim,
                  int *rowptrJ, int *colJ, double *valJ, int cap);

/* --- sparse direct solve A x = b (replaces PARDISO, CPU/esparso/spSolver.cpp:84-189) */
/* minimum-degree ordering of A+A^T, left-looking LU with diagonal-preference  */
/* threshold pivoting.  Returns 0 on success, 1 on a zero pivot.              */
int  orc_sparse_solve(int n, const int *rowptr, const int *colind, const double *val,
                      const double *b, double *x);
/* the ordering alone (for tests): perm[k] = original index eliminated k-th   */
void orc_min_degree(int n, const int *rowptr, const int *colind, int *perm);

/* --- one limInjReat decision pass (CPU/limInjReat.cpp:38-131): mode 0 chkLimQ,
 *     1 undo_chkLimQ; limQ [nPV] and QliqLim [nB] in/out; returns flips     */
int  orc_limits(const orc_case *c, int mode, const double *Qload, const double *Qliq,
                const double *Q, const double *V, unsigned char *limQ, double *QliqLim);

/* --- Newton-Raphson driver (CPU/newtonRaphson.cpp:173-314) ------------------- */
/* V/theta in: initial state (NULL = the case's), out: final state.           */
/* limQ_out [nPV] (may be NULL).  Pliq/Qliq/Qload override the case (NULL=case) */
int  orc_nr(const orc_case *c, const orc_opts *o, const double *Pliq,
            const double *Qliq, const double *Qload, double *V, double *theta,
            unsigned char *limQ_out, orc_result *res);

/* scenario batch on the CPU: nsc scenarios of the same network, each with its */
/* own Pliq/Qliq/Qload rows; scenario-parallel over nthreads OpenMP threads.  */
/* --- branch flows as reported (CPU/esparso/potencia/spCalcFluxPQ.cpp:3-36) -- */
/* Pdp/Qdp: from -> to, Ppd/Qpd: to -> from, per branch [nL]                  */
void orc_flows(const orc_case *c, const double *V, const double *theta, double *Pdp, double *Qdp,
               double *Ppd, double *Qpd);

/* limQ [nsc x nPV] may be NULL.  inner = 0: scenario-parallel over nthreads  */
/* threads; inner = 1: scenarios one after another, OpenMP over nthreads     */
/* threads inside each solve (the reference's shape).                        */
int  orc_nr_batch(const orc_case *c, const orc_opts *o, int nsc, const double *Pliq,
                  const double *Qliq, const double *Qload, double *V, double *theta,
                  int *iters, int *converged, unsigned char *limQ, int nthreads, int inner);

#ifdef __cplusplus
}
#endif
#endif
