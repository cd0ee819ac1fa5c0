/*
 * pf_oracle.c -- FP64 CPU restatement of the reference's CPU sparse NR path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path (see
 * pf_oracle.h).  Nothing in the product links or calls this file.
 *
 * Follows, function by function:
 *   reader      CPU/leitura/cdf.cpp:3-493, CPU/leitura/id2i.cpp:3-10,
 *               CPU/estrutura/barra.cpp:20-25 (flat start)
 *   Ybus        CPU/esparso/spCalcYbus.cpp:4-90 (MATPOWER branch model)
 *   P, Q        CPU/esparso/potencia/spCalcPQliq.cpp:3-45
 *   mismatch    CPU/limInjReat.cpp:158-175
 *   Jacobian    CPU/esparso/spCalcJac.cpp:402-568
 *   solve       CPU/esparso/spSolver.cpp:84-189 (PARDISO; semantics only)
 *   limits      CPU/limInjReat.cpp:3-131, CPU/newtonRaphson.cpp:64-171
 *   NR loop     CPU/newtonRaphson.cpp:173-314
 *
 * Deviations from the reference (each documented where it applies):
 *   D1  isPQ (CPU/esparso/spCalcJac.cpp:391-400) stops scanning at nPQ rather
 *       than nPQlim, so after a PV->PQ switch the last switched buses lose
 *       their N/L columns and J turns structurally singular.  The oracle tests
 *       membership in barrasPQlim properly (as the reference's GPU rebuild,
 *       GPU/esparso.h:79-146, does).
 *   D2  PARDISO (reorder + factor + solve, pivot perturbation 1e-13) is
 *       replaced by the oracle's own sparse LU: minimum-degree ordering of
 *       A+A^T, left-looking LU with diagonal-preference threshold pivoting.
 *       The step solves the same J dx = g; only last-bit rounding differs.
 *   D3  After a switch taken at convergence (CPU/newtonRaphson.cpp:251-256)
 *       the reference solves the new J against the mismatch vector computed
 *       before the switch, whose entries are all <= tol; the state moves by
 *       O(tol).  The oracle takes a zero step there (the iteration is
 *       counted), keeping the reference's iteration count and decision path.
 *   D4  Complex division uses Smith's algorithm (CACM Algorithm 116, 1962)
 *       instead of cuComplex's cuCdiv; last-bit differences in Ybus only.
 *   D5  id2i and the duplicate-branch scan keep the reference's "first match"
 *       semantics with hashed/sorted lookups instead of linear scans.
 */
#define _GNU_SOURCE
#include "pf_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define PI_REF 3.14159265358979323846  /* CPU/leitura/cdf.cpp:215 */
#define FOLGA 1e-4                     /* tol_limInjReat, CPU/global.cpp:11 */

/* ------------------------------------------------------------------------- */
/* reader                                                                     */
/* ------------------------------------------------------------------------- */

/* std::string::substr(pos, len) followed by atoi/atof: pos past the end is an
 * error (std::out_of_range in the reference), a short tail is taken as is.  */
static int sub(const char *line, size_t len, size_t pos, size_t n, char *out) {
    if (pos > len) return -1;
    size_t m = len - pos < n ? len - pos : n;
    memcpy(out, line + pos, m);
    out[m] = 0;
    return 0;
}

static int sub_i(const char *line, size_t len, size_t pos, size_t n, int *v) {
    char b[64];
    if (sub(line, len, pos, n, b)) return -1;
    *v = atoi(b);
    return 0;
}

static int sub_d(const char *line, size_t len, size_t pos, size_t n, double *v) {
    char b[64];
    if (sub(line, len, pos, n, b)) return -1;
    *v = atof(b);
    return 0;
}

typedef struct {
    int id, idx;
} idpair;

static int cmp_idpair(const void *a, const void *b) {
    const idpair *x = a, *y = b;
    if (x->id != y->id) return x->id < y->id ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* first bus with this id (CPU/leitura/id2i.cpp:3-10); -1 if none */
static int id2i(const idpair *s, int n, int id) {
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (s[mid].id < id) lo = mid + 1;
        else hi = mid;
    }
    return (lo < n && s[lo].id == id) ? s[lo].idx : -1;
}

/* open-addressing map (from, to) -> first branch index */
typedef struct {
    long long *key;
    int *val;
    int cap;
} pairmap;

static void pm_init(pairmap *m, int n) {
    m->cap = 16;
    while (m->cap < 2 * n + 16) m->cap <<= 1;
    m->key = malloc(sizeof(long long) * m->cap);
    m->val = malloc(sizeof(int) * m->cap);
    for (int i = 0; i < m->cap; i++) m->key[i] = -1;
}

static int pm_find_or_add(pairmap *m, int a, int b, int v) {
    long long k = ((long long)a << 32) | (unsigned)b;
    unsigned long long h = (unsigned long long)k * 0x9E3779B97F4A7C15ull;
    int i = (int)(h >> 40) & (m->cap - 1);
    while (m->key[i] != -1) {
        if (m->key[i] == k) return m->val[i];
        i = (i + 1) & (m->cap - 1);
    }
    m->key[i] = k;
    m->val[i] = v;
    return -1;
}

static void pm_free(pairmap *m) {
    free(m->key);
    free(m->val);
}

static void cdiv(double ar, double ai, double br, double bi, double *cr, double *ci) {
    /* Smith's algorithm (CACM Algorithm 116) */
    if (fabs(br) >= fabs(bi)) {
        double r = bi / br, d = br + bi * r;
        *cr = (ar + ai * r) / d;
        *ci = (ai - ar * r) / d;
    } else {
        double r = br / bi, d = bi + br * r;
        *cr = (ar * r + ai) / d;
        *ci = (ai * r - ar) / d;
    }
}

static int is_end(const char *line) { return strstr(line, "-999") != NULL; }

static void chomp(char *line, ssize_t *len) {
    if (*len > 0 && line[*len - 1] == '\n') line[--(*len)] = 0;
}

int orc_case_read(const char *path, int wide, double v0, double theta0, orc_case **out) {
    *out = NULL;
    if (wide < 0) {
        const char *dot = strrchr(path, '.');
        wide = (dot && (strcmp(dot, ".cdfx") == 0 || strcmp(dot, ".CDFX") == 0));
    }
    FILE *f = fopen(path, "r");
    if (!f) return 6;
    /* column map: CPU/leitura/cdf.cpp:185-233 (CDF), :351-399 (CDFX) */
    const size_t o = wide ? 1 : 0;   /* every bus field after the id moves by one */
    const size_t ob = wide ? 2 : 0;  /* branch fields after both ids move by two  */
    char *line = NULL;
    size_t cap = 0;
    ssize_t len;
    int rc = 0;
    /* pass 1: sizes (CPU/leitura/cdf.cpp:3-81) */
    int nB = 0, nPQ = 0, nPV = 0, nSl = 0, nLraw = 0;
    if (getline(&line, &cap, f) < 0 || getline(&line, &cap, f) < 0) { rc = 3; goto done; }
    while ((len = getline(&line, &cap, f)) >= 0) {
        chomp(line, &len);
        if (is_end(line)) break;
        int t;
        if (sub_i(line, len, 24 + o, 2, &t)) { rc = 3; goto done; }
        nB++;
        if (t == 0 || t == 1) nPQ++;
        else if (t == 2) nPV++;
        else if (t == 3) nSl++;
        else { rc = 3; goto done; }   /* "valor de tipo de barra invalido" */
    }
    if (nSl != 1) { rc = 3; goto done; } /* no or several swing buses */
    if (getline(&line, &cap, f) < 0) { rc = 3; goto done; }
    while ((len = getline(&line, &cap, f)) >= 0) {
        chomp(line, &len);
        if (is_end(line)) break;
        nLraw++;
    }
    rewind(f);

    orc_case *c = calloc(1, sizeof(orc_case));
    c->nB = nB;
    c->nPV = nPV;
    c->nPQ = nPQ;
    double **bus[] = {&c->V, &c->theta, &c->Pliq, &c->Qliq, &c->Pload, &c->Qload,
                      &c->Pg, &c->Qg, &c->Vbase, &c->gsh, &c->bsh};
    for (int k = 0; k < 11; k++) *bus[k] = calloc(nB > 0 ? nB : 1, sizeof(double));
    c->id = calloc(nB > 0 ? nB : 1, sizeof(int));
    c->pv = calloc(nPV > 0 ? nPV : 1, sizeof(int));
    c->pq = calloc(nPQ > 0 ? nPQ : 1, sizeof(int));
    c->Qmin = calloc(nPV > 0 ? nPV : 1, sizeof(double));
    c->Qmax = calloc(nPV > 0 ? nPV : 1, sizeof(double));
    c->Vset = calloc(nPV > 0 ? nPV : 1, sizeof(double));
    int nLcap = nLraw > 0 ? nLraw : 1;
    c->from = calloc(nLcap, sizeof(int));
    c->to = calloc(nLcap, sizeof(int));
    double **br[] = {&c->r, &c->x, &c->bch, &c->tap_re, &c->tap_im, &c->phi};
    for (int k = 0; k < 6; k++) *br[k] = calloc(nLcap, sizeof(double));
    for (int i = 0; i < nB; i++) { c->V[i] = v0; c->theta[i] = theta0; }  /* barra.cpp:20-25 */

    /* pass 2: values (CPU/leitura/cdf.cpp:163-327) */
    if ((len = getline(&line, &cap, f)) < 0) { rc = 3; goto fail; }
    chomp(line, &len);
    { int mva; if (sub_i(line, len, 31, 6, &mva)) { rc = 3; goto fail; } c->baseMVA = mva; }
    if (c->baseMVA == 0) { rc = 3; goto fail; }
    getline(&line, &cap, f);
    int i = 0, ipv = 0, ipq = 0;
    while ((len = getline(&line, &cap, f)) >= 0) {
        chomp(line, &len);
        if (is_end(line)) break;
        int t = 0, e = 0;
        e |= sub_i(line, len, 0, 4 + o, &c->id[i]);
        e |= sub_i(line, len, 24 + o, 2, &t);
        if (t == 0 || t == 1) {
            c->pq[ipq++] = i;
        } else if (t == 2) {
            c->pv[ipv] = i;
            e |= sub_d(line, len, 27 + o, 6, &c->V[i]);
            c->Vset[ipv] = c->V[i];
            e |= sub_d(line, len, 90 + o, 7, &c->Qmax[ipv]);
            e |= sub_d(line, len, 98 + o, 7, &c->Qmin[ipv]);
            c->Qmax[ipv] /= c->baseMVA;
            c->Qmin[ipv] /= c->baseMVA;
            ipv++;
        } else {
            c->slack = i;
            e |= sub_d(line, len, 27 + o, 6, &c->V[i]);
            e |= sub_d(line, len, 33 + o, 7, &c->theta[i]);
            c->theta[i] = c->theta[i] * PI_REF / 180.;
        }
        e |= sub_d(line, len, 40 + o, 9, &c->Pload[i]);
        e |= sub_d(line, len, 49 + o, 10, &c->Qload[i]);
        e |= sub_d(line, len, 59 + o, 8, &c->Pg[i]);
        e |= sub_d(line, len, 67 + o, 8, &c->Qg[i]);
        e |= sub_d(line, len, 76 + o, 7, &c->Vbase[i]);
        e |= sub_d(line, len, 114 + o, 8, &c->bsh[i]);
        e |= sub_d(line, len, 106 + o, 8, &c->gsh[i]);
        if (e) { rc = 3; goto fail; }
        c->Pload[i] /= c->baseMVA;
        c->Qload[i] /= c->baseMVA;
        c->Pg[i] /= c->baseMVA;
        c->Qg[i] /= c->baseMVA;
        c->Pliq[i] = c->Pg[i] - c->Pload[i];
        c->Qliq[i] = c->Qg[i] - c->Qload[i];
        i++;
    }
    getline(&line, &cap, f);
    {
        idpair *ids = malloc(sizeof(idpair) * (nB > 0 ? nB : 1));
        for (int k = 0; k < nB; k++) { ids[k].id = c->id[k]; ids[k].idx = k; }
        qsort(ids, nB, sizeof(idpair), cmp_idpair);
        pairmap pm;
        pm_init(&pm, nLraw);
        int nL = 0;
        while ((len = getline(&line, &cap, f)) >= 0) {
            chomp(line, &len);
            if (is_end(line)) break;
            int a, b, e = 0;
            double r, x, bc, tapA, ph;
            e |= sub_i(line, len, 0, 4 + o, &a);
            e |= sub_i(line, len, 4 + ob, 5, &b);
            e |= sub_d(line, len, 19 + ob, 10, &r);
            e |= sub_d(line, len, 29 + ob, 11, &x);
            e |= sub_d(line, len, 40 + ob, 10, &bc);
            e |= sub_d(line, len, 76 + ob, 6, &tapA);
            e |= sub_d(line, len, 83 + ob, 7, &ph);
            if (e) { rc = 3; break; }
            int de = id2i(ids, nB, a), pa = id2i(ids, nB, b);
            if (de < 0 || pa < 0) { rc = 3; break; }  /* "Barra inexistente!" */
            int prev = pm_find_or_add(&pm, de, pa, nL);
            if (prev < 0) {
                c->from[nL] = de;
                c->to[nL] = pa;
                c->r[nL] = r;
                c->x[nL] = x;
                c->phi[nL] = ph * PI_REF / 180.;
                if (tapA != 0.) {
                    c->tap_re[nL] = tapA * cos(c->phi[nL]);
                    c->tap_im[nL] = tapA * sin(c->phi[nL]);
                } else {
                    c->tap_re[nL] = cos(c->phi[nL]);
                    c->tap_im[nL] = sin(c->phi[nL]);
                }
                c->bch[nL] = bc;
                nL++;
            } else {
                /* parallel equivalent: z = z1 z2 / (z1 + z2); bsh = 1/(b1+b2)
                 * (CPU/leitura/cdf.cpp:302-313, reciprocal kept as in the reference) */
                double nr = r * c->r[prev] - x * c->x[prev];
                double ni = r * c->x[prev] + x * c->r[prev];
                cdiv(nr, ni, c->r[prev] + r, c->x[prev] + x, &c->r[prev], &c->x[prev]);
                if (bc != 0 || c->bch[prev] != 0) c->bch[prev] = 1 / (bc + c->bch[prev]);
                else c->bch[prev] = 0;
                c->n_duplicates++;
            }
        }
        c->nL = nL;
        pm_free(&pm);
        free(ids);
    }
    if (rc) goto fail;
    *out = c;
    goto done;
fail:
    orc_case_free(c);
done:
    free(line);
    fclose(f);
    return rc;
}

void orc_case_free(orc_case *c) {
    if (!c) return;
    double *d[] = {c->V, c->theta, c->Pliq, c->Qliq, c->Pload, c->Qload, c->Pg, c->Qg,
                   c->Vbase, c->gsh, c->bsh, c->Qmin, c->Qmax, c->Vset,
                   c->r, c->x, c->bch, c->tap_re, c->tap_im, c->phi};
    for (size_t k = 0; k < sizeof(d) / sizeof(d[0]); k++) free(d[k]);
    free(c->id);
    free(c->pv);
    free(c->pq);
    free(c->from);
    free(c->to);
    free(c);
}

void orc_case_get(const orc_case *c, int *ints6, double *bus11, int *pv, int *pq,
                  double *pvlim3, int *br_ft, double *br6) {
    if (ints6) {
        ints6[0] = c->nB; ints6[1] = c->nL; ints6[2] = c->nPV; ints6[3] = c->nPQ;
        ints6[4] = c->slack; ints6[5] = c->n_duplicates;
    }
    if (bus11) {
        const double *src[] = {c->V, c->theta, c->Pliq, c->Qliq, c->Pload, c->Qload,
                               c->Pg, c->Qg, c->Vbase, c->gsh, c->bsh};
        for (int k = 0; k < 11; k++) memcpy(bus11 + (size_t)k * c->nB, src[k], sizeof(double) * c->nB);
    }
    if (pv) memcpy(pv, c->pv, sizeof(int) * c->nPV);
    if (pq) memcpy(pq, c->pq, sizeof(int) * c->nPQ);
    if (pvlim3) {
        memcpy(pvlim3, c->Qmin, sizeof(double) * c->nPV);
        memcpy(pvlim3 + c->nPV, c->Qmax, sizeof(double) * c->nPV);
        memcpy(pvlim3 + 2 * c->nPV, c->Vset, sizeof(double) * c->nPV);
    }
    if (br_ft) {
        memcpy(br_ft, c->from, sizeof(int) * c->nL);
        memcpy(br_ft + c->nL, c->to, sizeof(int) * c->nL);
    }
    if (br6) {
        const double *src[] = {c->r, c->x, c->bch, c->tap_re, c->tap_im, c->phi};
        for (int k = 0; k < 6; k++) memcpy(br6 + (size_t)k * c->nL, src[k], sizeof(double) * c->nL);
    }
}

/* ------------------------------------------------------------------------- */
/* Ybus (CPU/esparso/spCalcYbus.cpp:4-90)                                     */
/* ------------------------------------------------------------------------- */

static int cmp_int(const void *a, const void *b) {
    int x = *(const int *)a, y = *(const int *)b;
    return (x > y) - (x < y);
}

/* structural pattern: diagonal + both directions of every branch, sorted */
static int ybus_pattern(const orc_case *c, int *rowptr, int *colind) {
    int nB = c->nB;
    int *cnt = calloc(nB + 1, sizeof(int));
    for (int i = 0; i < nB; i++) cnt[i] = 1;
    for (int l = 0; l < c->nL; l++) { cnt[c->from[l]]++; cnt[c->to[l]]++; }
    int *ptr = malloc(sizeof(int) * (nB + 1));
    ptr[0] = 0;
    for (int i = 0; i < nB; i++) ptr[i + 1] = ptr[i] + cnt[i];
    int *tmp = malloc(sizeof(int) * (ptr[nB] > 0 ? ptr[nB] : 1));
    int *fill = malloc(sizeof(int) * (nB > 0 ? nB : 1));
    for (int i = 0; i < nB; i++) { fill[i] = ptr[i]; tmp[fill[i]++] = i; }
    for (int l = 0; l < c->nL; l++) {
        tmp[fill[c->from[l]]++] = c->to[l];
        tmp[fill[c->to[l]]++] = c->from[l];
    }
    int nnz = 0;
    if (rowptr) rowptr[0] = 0;
    for (int i = 0; i < nB; i++) {
        qsort(tmp + ptr[i], ptr[i + 1] - ptr[i], sizeof(int), cmp_int);
        int last = -1;
        for (int p = ptr[i]; p < ptr[i + 1]; p++) {
            if (tmp[p] == last) continue;
            last = tmp[p];
            if (colind) colind[nnz] = last;
            nnz++;
        }
        if (rowptr) rowptr[i + 1] = nnz;
    }
    free(cnt); free(ptr); free(tmp); free(fill);
    return nnz;
}

int orc_ybus_nnz(const orc_case *c) { return ybus_pattern(c, NULL, NULL); }

static int find_col(const int *rowptr, const int *colind, int i, int j) {
    int lo = rowptr[i], hi = rowptr[i + 1];
    while (lo < hi) {
        int mid = (lo + hi) / 2;
        if (colind[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

void orc_ybus(const orc_case *c, int *rowptr, int *colind, double *y) {
    int nnz = ybus_pattern(c, rowptr, colind);
    memset(y, 0, sizeof(double) * 2 * nnz);
    for (int l = 0; l < c->nL; l++) {
        int f = c->from[l], t = c->to[l];
        double yr, yi;
        cdiv(1.0, 0.0, c->r[l], c->x[l], &yr, &yi);                  /* 1/z */
        /* 1/|t|^2 as 1/(re^2 + im^2): the reference squares cuCabs(t); the
         * product form has no library call, so it rounds the same everywhere */
        double s = 1.0 / (c->tap_re[l] * c->tap_re[l] + c->tap_im[l] * c->tap_im[l]);
        double ffr = yr * s, ffi = (yi + c->bch[l] / 2.) * s;        /* (y + jb/2)/|t|^2 */
        double a_r, a_i, ftr, fti, tfr, tfi;
        cdiv(-1.0, 0.0, c->tap_re[l], -c->tap_im[l], &a_r, &a_i);   /* -1/conj(t) */
        cdiv(a_r, a_i, c->r[l], c->x[l], &ftr, &fti);               /* ... / z    */
        cdiv(-1.0, 0.0, c->tap_re[l], c->tap_im[l], &a_r, &a_i);    /* -1/t       */
        cdiv(a_r, a_i, c->r[l], c->x[l], &tfr, &tfi);
        double ttr = yr, tti = yi + c->bch[l] / 2.;                  /* y + jb/2   */
        int p;
        p = find_col(rowptr, colind, f, f); y[2 * p] += ffr; y[2 * p + 1] += ffi;
        p = find_col(rowptr, colind, f, t); y[2 * p] += ftr; y[2 * p + 1] += fti;
        p = find_col(rowptr, colind, t, f); y[2 * p] += tfr; y[2 * p + 1] += tfi;
        p = find_col(rowptr, colind, t, t); y[2 * p] += ttr; y[2 * p + 1] += tti;
    }
    for (int i = 0; i < c->nB; i++) {
        int p = find_col(rowptr, colind, i, i);
        y[2 * p] += c->gsh[i];
        y[2 * p + 1] += c->bsh[i];
    }
}

/* ------------------------------------------------------------------------- */
/* P, Q (CPU/esparso/potencia/spCalcPQliq.cpp:3-45)                           */
/* ------------------------------------------------------------------------- */

void orc_pq(int nB, const int *rowptr, const int *colind, const double *y,
            const double *V, const double *theta, double *P, double *Q, int openmp) {
    (void)openmp;
#pragma omp parallel for if (openmp)
    for (int k = 0; k < nB; k++) {
        double accp = 0, accq = 0;
        for (int e = rowptr[k]; e < rowptr[k + 1]; e++) {
            int m = colind[e];
            double a = theta[k] - theta[m];
            double G = y[2 * e], B = y[2 * e + 1];
            accp += (G * cos(a) + B * sin(a)) * V[m];
            accq += (G * sin(a) - B * cos(a)) * V[m];
        }
        P[k] = accp * V[k];
        Q[k] = accq * V[k];
    }
}

/* ------------------------------------------------------------------------- */
/* mismatch (CPU/limInjReat.cpp:158-175)                                      */
/* ------------------------------------------------------------------------- */

void orc_mismatch(int nB, int slack, const double *Pliq, const double *QliqLim,
                  const double *P, const double *Q, int nPQlim, const int *pqlim,
                  double *g) {
    for (int k = 0; k < nB; k++) {
        if (k == slack) continue;
        g[k - (k > slack)] = Pliq[k] - P[k];
    }
    for (int i = 0; i < nPQlim; i++) {
        int b = pqlim[i];
        g[nB - 1 + i] = QliqLim[b] - Q[b];
    }
}

/* ------------------------------------------------------------------------- */
/* Jacobian (CPU/esparso/spCalcJac.cpp:402-568)                               */
/* rows:  [P rows of the non-slack buses, ascending] [Q rows of barrasPQlim]   */
/* cols:  [theta of the non-slack buses]             [|V| of barrasPQlim]      */
/* ------------------------------------------------------------------------- */

int orc_jacobian(int nB, int slack, const int *rowptr, const int *colind,
                 const double *y, const double *V, const double *theta,
                 const double *P, const double *Q, int nPQlim, const int *pqlim,
                 int *rowptrJ, int *colJ, double *valJ, int cap) {
    int *pos = malloc(sizeof(int) * (nB > 0 ? nB : 1));
    for (int k = 0; k < nB; k++) pos[k] = -1;
    for (int i = 0; i < nPQlim; i++) pos[pqlim[i]] = i;   /* D1: proper membership */
    int nnz = 0, row = 0;
    rowptrJ[0] = 0;
#define PUSH(cc, vv) do { if (nnz < cap) { colJ[nnz] = (cc); valJ[nnz] = (vv); } nnz++; } while (0)
    for (int k = 0; k < nB; k++) {
        if (k == slack) continue;
        for (int e = rowptr[k]; e < rowptr[k + 1]; e++) {          /* H */
            int m = colind[e];
            if (m == slack) continue;
            double G = y[2 * e], B = y[2 * e + 1];
            double v;
            if (m != k) {
                double a = theta[k] - theta[m];
                v = V[k] * V[m] * (G * sin(a) - B * cos(a));
            } else {
                v = -Q[m] - V[m] * V[m] * B;
            }
            PUSH(m - (m > slack), v);
        }
        for (int e = rowptr[k]; e < rowptr[k + 1]; e++) {          /* N */
            int m = colind[e];
            if (pos[m] < 0) continue;
            double G = y[2 * e], B = y[2 * e + 1];
            double v;
            if (m != k) {
                double a = theta[k] - theta[m];
                v = V[k] * (G * cos(a) + B * sin(a));
            } else {
                v = (P[m] + V[m] * V[m] * G) / V[m];
            }
            PUSH(nB - 1 + pos[m], v);
        }
        rowptrJ[++row] = nnz;
    }
    for (int i = 0; i < nPQlim; i++) {
        int k = pqlim[i];
        for (int e = rowptr[k]; e < rowptr[k + 1]; e++) {          /* M */
            int m = colind[e];
            if (m == slack) continue;
            double G = y[2 * e], B = y[2 * e + 1];
            double v;
            if (m != k) {
                double a = theta[k] - theta[m];
                v = -V[k] * V[m] * (G * cos(a) + B * sin(a));
            } else {
                v = P[m] - V[m] * V[m] * G;
            }
            PUSH(m - (m > slack), v);
        }
        for (int e = rowptr[k]; e < rowptr[k + 1]; e++) {          /* L */
            int m = colind[e];
            if (pos[m] < 0) continue;
            double G = y[2 * e], B = y[2 * e + 1];
            double v;
            if (m != k) {
                double a = theta[k] - theta[m];
                v = V[k] * (G * sin(a) - B * cos(a));
            } else {
                v = (Q[m] - V[m] * V[m] * B) / V[m];
            }
            PUSH(nB - 1 + pos[m], v);
        }
        rowptrJ[++row] = nnz;
    }
#undef PUSH
    free(pos);
    return nnz;
}

/* ------------------------------------------------------------------------- */
/* sparse LU (D2)                                                             */
/* ------------------------------------------------------------------------- */

/* binary min-heap of (degree, node) with lazy deletion */
typedef struct { int *d, *v, n, cap; } heap;

static void hpush(heap *h, int d, int v) {
    if (h->n == h->cap) {
        h->cap = h->cap ? 2 * h->cap : 64;
        h->d = realloc(h->d, sizeof(int) * h->cap);
        h->v = realloc(h->v, sizeof(int) * h->cap);
    }
    int i = h->n++;
    while (i > 0) {
        int p = (i - 1) / 2;
        if (h->d[p] < d || (h->d[p] == d && h->v[p] < v)) break;
        h->d[i] = h->d[p]; h->v[i] = h->v[p]; i = p;
    }
    h->d[i] = d; h->v[i] = v;
}

static void hpop(heap *h, int *d, int *v) {
    *d = h->d[0]; *v = h->v[0];
    int ld = h->d[--h->n], lv = h->v[h->n], i = 0;
    for (;;) {
        int c = 2 * i + 1;
        if (c >= h->n) break;
        if (c + 1 < h->n && (h->d[c + 1] < h->d[c] || (h->d[c + 1] == h->d[c] && h->v[c + 1] < h->v[c]))) c++;
        if (ld < h->d[c] || (ld == h->d[c] && lv < h->v[c])) break;
        h->d[i] = h->d[c]; h->v[i] = h->v[c]; i = c;
    }
    h->d[i] = ld; h->v[i] = lv;
}

typedef struct { int *a, n, cap; } ivec;

static void iv_push(ivec *x, int v) {
    if (x->n == x->cap) { x->cap = x->cap ? 2 * x->cap : 8; x->a = realloc(x->a, sizeof(int) * x->cap); }
    x->a[x->n++] = v;
}

/* Minimum degree on the explicit elimination graph of A+A^T; ties -> lowest
 * index.  perm[k] = vertex eliminated k-th. */
void orc_min_degree(int n, const int *rowptr, const int *colind, int *perm) {
    ivec *adj = calloc(n > 0 ? n : 1, sizeof(ivec));
    for (int i = 0; i < n; i++)
        for (int p = rowptr[i]; p < rowptr[i + 1]; p++) {
            int j = colind[p];
            if (j == i) continue;
            iv_push(&adj[i], j);
            iv_push(&adj[j], i);
        }
    for (int i = 0; i < n; i++) {       /* sort + unique */
        qsort(adj[i].a, adj[i].n, sizeof(int), cmp_int);
        int m = 0;
        for (int p = 0; p < adj[i].n; p++)
            if (m == 0 || adj[i].a[m - 1] != adj[i].a[p]) adj[i].a[m++] = adj[i].a[p];
        adj[i].n = m;
    }
    char *dead = calloc(n > 0 ? n : 1, 1);
    heap h = {0};
    for (int i = 0; i < n; i++) hpush(&h, adj[i].n, i);
    int *tmp = NULL, tcap = 0;
    for (int k = 0; k < n; k++) {
        int d, v;
        do { hpop(&h, &d, &v); } while (dead[v] || d != adj[v].n);
        perm[k] = v;
        dead[v] = 1;
        ivec nv = adj[v];
        for (int q = 0; q < nv.n; q++) {
            int u = nv.a[q];
            /* adj[u] <- (adj[u] U N(v)) \ {u, v} */
            int need = adj[u].n + nv.n;
            if (need > tcap) { tcap = 2 * need; tmp = realloc(tmp, sizeof(int) * tcap); }
            int i1 = 0, i2 = 0, m = 0;
            while (i1 < adj[u].n || i2 < nv.n) {
                int x;
                if (i2 >= nv.n || (i1 < adj[u].n && adj[u].a[i1] < nv.a[i2])) x = adj[u].a[i1++];
                else if (i1 >= adj[u].n || nv.a[i2] < adj[u].a[i1]) x = nv.a[i2++];
                else { x = adj[u].a[i1]; i1++; i2++; }
                if (x == u || x == v) continue;
                tmp[m++] = x;
            }
            if (m > adj[u].cap) { adj[u].cap = m; adj[u].a = realloc(adj[u].a, sizeof(int) * m); }
            memcpy(adj[u].a, tmp, sizeof(int) * m);
            adj[u].n = m;
            hpush(&h, m, u);
        }
        free(adj[v].a);
        adj[v].a = NULL; adj[v].n = adj[v].cap = 0;
    }
    for (int i = 0; i < n; i++) free(adj[i].a);
    free(adj); free(dead); free(tmp); free(h.d); free(h.v);
}

typedef struct { int *p, *i; double *x; int nz, cap; } ccol;

static void cc_push(ccol *m, int i, double x) {
    if (m->nz == m->cap) {
        m->cap = m->cap ? 2 * m->cap : 256;
        m->i = realloc(m->i, sizeof(int) * m->cap);
        m->x = realloc(m->x, sizeof(double) * m->cap);
    }
    m->i[m->nz] = i; m->x[m->nz++] = x;
}

/* Left-looking LU (Gilbert-Peierls column algorithm) of A(:, q) with
 * diagonal-preference threshold partial pivoting.  A in CSC. */
static int lu_factor(int n, const int *Ap, const int *Ai, const double *Ax, const int *q,
                     ccol *L, ccol *U, int *pinv) {
    const double tol = 1e-3;
    double *x = calloc(n > 0 ? n : 1, sizeof(double));
    int *xi = malloc(sizeof(int) * (2 * n + 1));
    int *mark = malloc(sizeof(int) * (n > 0 ? n : 1));
    int *pstack = malloc(sizeof(int) * (n > 0 ? n : 1));
    L->p = malloc(sizeof(int) * (n + 1));
    U->p = malloc(sizeof(int) * (n + 1));
    for (int i = 0; i < n; i++) { pinv[i] = -1; mark[i] = -1; }
    int rc = 0;
    for (int k = 0; k < n; k++) {
        L->p[k] = L->nz;
        U->p[k] = U->nz;
        int col = q[k];
        /* reach of A(:,col) in the graph of L (columns of pivoted rows), topological */
        int top = n;
        for (int p = Ap[col]; p < Ap[col + 1]; p++) {
            int s = Ai[p];
            if (mark[s] == k) continue;
            int head = 0;
            xi[0] = s;
            while (head >= 0) {
                int j = xi[head];
                int J = pinv[j];
                if (mark[j] != k) {
                    mark[j] = k;
                    pstack[head] = (J < 0) ? 0 : L->p[J] + 1;  /* skip the pivot entry */
                }
                int done = 1;
                int pend = (J < 0) ? 0 : L->p[J + 1];   /* pivoted => column J < k is complete */
                for (int pp = pstack[head]; pp < pend; pp++) {
                    int ii = L->i[pp];
                    if (mark[ii] == k) continue;
                    pstack[head] = pp + 1;
                    xi[++head] = ii;
                    done = 0;
                    break;
                }
                if (done) { head--; xi[--top] = j; }
            }
        }
        for (int p = Ap[col]; p < Ap[col + 1]; p++) x[Ai[p]] = Ax[p];
        for (int px = top; px < n; px++) {
            int j = xi[px];
            int J = pinv[j];
            if (J < 0) continue;
            double xj = x[j];
            for (int p = L->p[J] + 1; p < L->p[J + 1]; p++) x[L->i[p]] -= L->x[p] * xj;
        }
        int ipiv = -1;
        double amax = -1;
        for (int px = top; px < n; px++) {
            int i = xi[px];
            if (pinv[i] < 0) {
                double a = fabs(x[i]);
                if (a > amax) { amax = a; ipiv = i; }
            } else {
                cc_push(U, pinv[i], x[i]);
            }
        }
        if (ipiv < 0 || amax <= 0) { rc = 1; break; }
        if (pinv[col] < 0 && mark[col] == k && fabs(x[col]) >= tol * amax) ipiv = col;
        double piv = x[ipiv];
        cc_push(U, k, piv);
        pinv[ipiv] = k;
        cc_push(L, ipiv, 1.0);
        for (int px = top; px < n; px++) {
            int i = xi[px];
            if (pinv[i] < 0) cc_push(L, i, x[i] / piv);
            x[i] = 0;
        }
    }
    if (!rc) {
        L->p[n] = L->nz;
        U->p[n] = U->nz;
        for (int p = 0; p < L->nz; p++) L->i[p] = pinv[L->i[p]];
    }
    free(x); free(xi); free(mark); free(pstack);
    return rc;
}

static void csr_to_csc(int n, const int *rp, const int *ci, const double *v,
                       int *cp, int *ri, double *cv) {
    int *w = calloc(n + 1, sizeof(int));
    for (int p = 0; p < rp[n]; p++) w[ci[p]]++;
    cp[0] = 0;
    for (int j = 0; j < n; j++) cp[j + 1] = cp[j] + w[j];
    for (int j = 0; j < n; j++) w[j] = cp[j];
    for (int i = 0; i < n; i++)
        for (int p = rp[i]; p < rp[i + 1]; p++) {
            int d = w[ci[p]]++;
            ri[d] = i;
            cv[d] = v[p];
        }
    free(w);
}

static int sparse_solve_q(int n, const int *rowptr, const int *colind, const double *val,
                          const double *b, double *xout, const int *q) {
    int nnz = rowptr[n];
    int *cp = malloc(sizeof(int) * (n + 1));
    int *ri = malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double *cv = malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    csr_to_csc(n, rowptr, colind, val, cp, ri, cv);
    ccol L = {0}, U = {0};
    int *pinv = malloc(sizeof(int) * (n > 0 ? n : 1));
    int rc = lu_factor(n, cp, ri, cv, q, &L, &U, pinv);
    if (!rc) {
        double *y = malloc(sizeof(double) * (n > 0 ? n : 1));
        for (int i = 0; i < n; i++) y[pinv[i]] = b[i];
        for (int j = 0; j < n; j++)
            for (int p = L.p[j] + 1; p < L.p[j + 1]; p++) y[L.i[p]] -= L.x[p] * y[j];
        for (int j = n - 1; j >= 0; j--) {
            y[j] /= U.x[U.p[j + 1] - 1];
            for (int p = U.p[j]; p < U.p[j + 1] - 1; p++) y[U.i[p]] -= U.x[p] * y[j];
        }
        for (int k = 0; k < n; k++) xout[q[k]] = y[k];
        free(y);
    }
    free(cp); free(ri); free(cv); free(pinv);
    free(L.p); free(L.i); free(L.x); free(U.p); free(U.i); free(U.x);
    return rc;
}

int orc_sparse_solve(int n, const int *rowptr, const int *colind, const double *val,
                     const double *b, double *x) {
    int *q = malloc(sizeof(int) * (n > 0 ? n : 1));
    orc_min_degree(n, rowptr, colind, q);
    int rc = sparse_solve_q(n, rowptr, colind, val, b, x, q);
    free(q);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* reactive-limit logic (CPU/limInjReat.cpp, CPU/newtonRaphson.cpp:64-171)    */
/* ------------------------------------------------------------------------- */

typedef struct {
    int nPQlim, nPVlim;
    int *pqlim, *pvlim;
    unsigned char *limQ;
    double *QliqLim;
} limstate;

/* geraVetPV / geraVetPQ (CPU/limInjReat.cpp:3-36): sorted merge */
static void gera_vets(const orc_case *c, limstate *s) {
    s->nPVlim = 0;
    for (int i = 0; i < c->nPV; i++)
        if (!s->limQ[i]) s->pvlim[s->nPVlim++] = c->pv[i];
    int a = 0, m = 0;
    for (int i = 0; i < c->nPV; i++) {
        if (!s->limQ[i]) continue;
        while (a < c->nPQ && c->pq[a] < c->pv[i]) s->pqlim[m++] = c->pq[a++];
        s->pqlim[m++] = c->pv[i];
    }
    while (a < c->nPQ) s->pqlim[m++] = c->pq[a++];
    s->nPQlim = m;
}

/* chkLimQ (CPU/limInjReat.cpp:38-91) */
static int chk_lim_q(const orc_case *c, const double *Qload, const double *Qliq,
                     const double *Q, const double *V, limstate *s) {
    int flips = 0;
    for (int i = 0; i < c->nPV; i++) {
        int b = c->pv[i];
        double qg = Q[b] + Qload[b];
        if (!s->limQ[i]) {
            if (qg < c->Qmin[i] - FOLGA) {
                s->QliqLim[b] = c->Qmin[i] - Qload[b];
                s->limQ[i] = 1; flips++;
            } else if (qg > c->Qmax[i] + FOLGA) {
                s->limQ[i] = 1; flips++;
                s->QliqLim[b] = c->Qmax[i] - Qload[b];
            }
        } else {
            if (s->QliqLim[b] == c->Qmax[i] - Qload[b]) {
                if (V[b] > c->Vset[i] + FOLGA) {
                    s->limQ[i] = 0; flips++;
                    s->QliqLim[b] = Qliq[b];
                }
            } else {
                if (V[b] < c->Vset[i] - FOLGA) {
                    s->limQ[i] = 0; flips++;
                    s->QliqLim[b] = Qliq[b];
                }
            }
        }
    }
    return flips;
}

/* undo_chkLimQ (CPU/limInjReat.cpp:93-131): back-switch tests only */
static int undo_chk_lim_q(const orc_case *c, const double *Qload, const double *Qliq,
                          const double *V, limstate *s) {
    int flips = 0;
    for (int i = 0; i < c->nPV; i++) {
        if (!s->limQ[i]) continue;
        int b = c->pv[i];
        if (s->QliqLim[b] == c->Qmax[i] - Qload[b]) {
            if (V[b] > c->Vset[i] + FOLGA) { s->limQ[i] = 0; s->QliqLim[b] = Qliq[b]; flips++; }
        } else {
            if (V[b] < c->Vset[i] - FOLGA) { s->limQ[i] = 0; s->QliqLim[b] = Qliq[b]; flips++; }
        }
    }
    return flips;
}

/* simpleChkLimQ (CPU/newtonRaphson.cpp:140-154) */
static int simple_chk_lim_q(const orc_case *c, const double *Qload, const double *Q) {
    for (int i = 0; i < c->nPV; i++) {
        int b = c->pv[i];
        double qg = Q[b] + Qload[b];
        if (qg < c->Qmin[i] - FOLGA) return 0;
        if (qg > c->Qmax[i] + FOLGA) return 0;
    }
    return 1;
}

static double maxabs(const double *v, int n) {
    double m = 0;
    for (int i = 0; i < n; i++) if (fabs(v[i]) > m) m = fabs(v[i]);
    return m;
}

/* exported single decision passes (stage tests): mode 0 chkLimQ, 1 undo_chkLimQ */
int orc_limits(const orc_case *c, int mode, const double *Qload, const double *Qliq,
               const double *Q, const double *V, unsigned char *limQ, double *QliqLim) {
    limstate s;
    s.limQ = limQ;
    s.QliqLim = QliqLim;
    if (mode == 0) return chk_lim_q(c, Qload, Qliq, Q, V, &s);
    return undo_chk_lim_q(c, Qload, Qliq, V, &s);
}

/* ------------------------------------------------------------------------- */
/* NR driver (CPU/newtonRaphson.cpp:173-314)                                  */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n;
    int *q;          /* cached ordering (reuse_symbolic) */
    int nPQlim;
} symcache;

int orc_nr(const orc_case *c, const orc_opts *o, const double *Pliq_in,
           const double *Qliq_in, const double *Qload_in, double *V, double *theta,
           unsigned char *limQ_out, orc_result *res) {
    const int nB = c->nB, nPV = c->nPV, nPQ = c->nPQ, sl = c->slack;
    const double *Pliq = Pliq_in ? Pliq_in : c->Pliq;
    const double *Qliq = Qliq_in ? Qliq_in : c->Qliq;
    const double *Qload = Qload_in ? Qload_in : c->Qload;
    memset(res, 0, sizeof(*res));
    int nnzY = orc_ybus_nnz(c);
    int *rowptr = malloc(sizeof(int) * (nB + 1));
    int *colind = malloc(sizeof(int) * (nnzY > 0 ? nnzY : 1));
    double *y = malloc(sizeof(double) * 2 * (nnzY > 0 ? nnzY : 1));
    orc_ybus(c, rowptr, colind, y);
    double *P = calloc(nB, sizeof(double)), *Q = calloc(nB, sizeof(double));
    int nmax = nB - 1 + nPQ + nPV;
    double *g = calloc(nmax > 0 ? nmax : 1, sizeof(double));
    double *dx = calloc(nmax > 0 ? nmax : 1, sizeof(double));
    int capJ = 4 * nnzY + 16;
    int *rpJ = malloc(sizeof(int) * (nmax + 1));
    int *ciJ = malloc(sizeof(int) * capJ);
    double *vJ = malloc(sizeof(double) * capJ);
    limstate s;
    s.pqlim = malloc(sizeof(int) * (nPQ + nPV + 1));
    s.pvlim = malloc(sizeof(int) * (nPV + 1));
    s.limQ = calloc(nPV + 1, 1);
    s.QliqLim = malloc(sizeof(double) * nB);
    double *Vown = NULL, *Town = NULL;
    if (!V) { V = Vown = malloc(sizeof(double) * nB); memcpy(V, c->V, sizeof(double) * nB); }
    if (!theta) { theta = Town = malloc(sizeof(double) * nB); memcpy(theta, c->theta, sizeof(double) * nB); }
    /* initLim (CPU/limInjReat.cpp:177-207) */
    for (int i = 0; i < nPQ; i++) s.pqlim[i] = c->pq[i];
    s.nPQlim = nPQ;
    for (int i = 0; i < nPV; i++) s.pvlim[i] = c->pv[i];
    s.nPVlim = nPV;
    for (int i = 0; i < nB; i++) s.QliqLim[i] = Qliq[i];
    int ngLim = nB - 1 + s.nPQlim;
    symcache sc = {0, NULL, -1};
    int it;
    for (it = 0; it < o->max_iter; it++) {
        orc_pq(nB, rowptr, colind, y, V, theta, P, Q, o->openmp);
        if (it >= 1 && o->lim_inj_reat) {
            int flips = 0;
            if (maxabs(dx, ngLim) < 0.1) {                 /* :202-206 evalLimInjReat */
                flips = chk_lim_q(c, Qload, Qliq, Q, V, &s);
                gera_vets(c, &s);
            } else {
                int any = 0;
                for (int i = 0; i < nPV; i++) any |= s.limQ[i];
                if (any) {                                  /* :207-211 undo_LimInjReat */
                    flips = undo_chk_lim_q(c, Qload, Qliq, V, &s);
                    if (flips) gera_vets(c, &s);
                }
            }
            res->n_switch_events += flips;
        }
        ngLim = nB - 1 + s.nPQlim;
        orc_mismatch(nB, sl, Pliq, s.QliqLim, P, Q, s.nPQlim, s.pqlim, g);
        double maxG = maxabs(g, ngLim);
        res->max_mismatch = maxG;
        if (maxG <= o->tol) {
            if (o->lim_inj_reat && !simple_chk_lim_q(c, Qload, Q)) {   /* :252-256 */
                res->n_switch_events += chk_lim_q(c, Qload, Qliq, Q, V, &s);
                gera_vets(c, &s);
                ngLim = nB - 1 + s.nPQlim;
                memset(dx, 0, sizeof(double) * ngLim);                  /* D3: zero step */
                continue;
            }
            res->converged = 1;
            break;
        }
        int nnzJ = orc_jacobian(nB, sl, rowptr, colind, y, V, theta, P, Q, s.nPQlim, s.pqlim,
                                rpJ, ciJ, vJ, capJ);
        if (nnzJ > capJ) {
            capJ = nnzJ;
            ciJ = realloc(ciJ, sizeof(int) * capJ);
            vJ = realloc(vJ, sizeof(double) * capJ);
            orc_jacobian(nB, sl, rowptr, colind, y, V, theta, P, Q, s.nPQlim, s.pqlim,
                         rpJ, ciJ, vJ, capJ);
        }
        int rc;
        if (o->reuse_symbolic) {
            if (sc.nPQlim != s.nPQlim || !sc.q) {
                free(sc.q);
                sc.q = malloc(sizeof(int) * (ngLim > 0 ? ngLim : 1));
                orc_min_degree(ngLim, rpJ, ciJ, sc.q);
                sc.nPQlim = s.nPQlim;
            }
            rc = sparse_solve_q(ngLim, rpJ, ciJ, vJ, g, dx, sc.q);
        } else {
            rc = orc_sparse_solve(ngLim, rpJ, ciJ, vJ, g, dx);
        }
        if (rc) { res->singular = 1; it = o->max_iter; break; }
        /* attVOlim (CPU/limInjReat.cpp:209-232) */
        for (int k = 0; k < nB; k++) {
            if (k == sl) continue;
            theta[k] += dx[k - (k > sl)];
        }
        for (int i = 0; i < s.nPQlim; i++) V[s.pqlim[i]] += dx[nB - 1 + i];
    }
    res->iterations = it < o->max_iter ? it : o->max_iter;
    if (limQ_out) memcpy(limQ_out, s.limQ, nPV);
    free(Vown); free(Town);
    free(rowptr); free(colind); free(y); free(P); free(Q); free(g); free(dx);
    free(rpJ); free(ciJ); free(vJ);
    free(s.pqlim); free(s.pvlim); free(s.limQ); free(s.QliqLim); free(sc.q);
    return res->converged ? 0 : (res->singular ? 2 : 1);
}

int orc_nr_batch(const orc_case *c, const orc_opts *o, int nsc, const double *Pliq,
                 const double *Qliq, const double *Qload, double *V, double *theta,
                 int *iters, int *converged, unsigned char *limQ, int nthreads, int inner) {
    int nB = c->nB, nPV = c->nPV;
    int nt = nthreads > 0 ? nthreads : 1;
    if (inner) {
        /* the reference's own shape: one case at a time, OpenMP inside the
           solve on the loops it parallelises (CPU/esparso/potencia/spCalcPQliq.cpp) */
        orc_opts oi = *o;
        oi.openmp = nt > 1;
#ifdef _OPENMP
        omp_set_num_threads(nt);
#endif
        for (int s = 0; s < nsc; s++) {
            orc_result r;
            double *Vs = V + (size_t)s * nB, *Ts = theta + (size_t)s * nB;
            memcpy(Vs, c->V, sizeof(double) * nB);
            memcpy(Ts, c->theta, sizeof(double) * nB);
            orc_nr(c, &oi, Pliq + (size_t)s * nB, Qliq + (size_t)s * nB, Qload + (size_t)s * nB, Vs, Ts,
                   limQ ? limQ + (size_t)s * nPV : NULL, &r);
            if (iters) iters[s] = r.iterations;
            if (converged) converged[s] = r.converged;
        }
        return 0;
    }
    orc_opts oi = *o;
    oi.openmp = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(nt)
    for (int s = 0; s < nsc; s++) {
        orc_result r;
        double *Vs = V + (size_t)s * nB, *Ts = theta + (size_t)s * nB;
        memcpy(Vs, c->V, sizeof(double) * nB);
        memcpy(Ts, c->theta, sizeof(double) * nB);
        orc_nr(c, &oi, Pliq + (size_t)s * nB, Qliq + (size_t)s * nB, Qload + (size_t)s * nB,
               Vs, Ts, limQ ? limQ + (size_t)s * nPV : NULL, &r);
        if (iters) iters[s] = r.iterations;
        if (converged) converged[s] = r.converged;
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Branch flows as the reference reports them                                 */
/* SpCalcFlux CPU/esparso/potencia/spCalcFluxPQ.cpp:17-36, SpP/SpQ :3-15:     */
/*   a = phif(k,m) + th_k - th_m, G + jB = Ybus(k,m)                          */
/*   P_km = V_k^2 G - V_k V_m G cos a - V_k V_m B sin a                       */
/*   Q_km = -V_k^2 (B + bshf(k,m)) + V_k V_m B cos a - V_k V_m G sin a        */
/* phif CPU/phifBshf.cpp:31-46 over eigen_phi, built from the (de, para, phi) */
/* triplets with duplicates summed (CPU/estrutura/ramo.cpp:4-13): the stored  */
/* (k,m) entry, else minus the stored (m,k) entry, else 0.  bshf              */
/* CPU/phifBshf.cpp:48-63: bsh of the first branch joining k and m.           */
/* Ybus(k,m) is 0 where the pattern has no entry (Eigen coeff()).             */
/* ------------------------------------------------------------------------- */

static double flow_phif(const orc_case *c, int k, int m) {
    double s = 0;
    int found = 0;
    for (int l = 0; l < c->nL; l++)
        if (c->from[l] == k && c->to[l] == m) { s += c->phi[l]; found = 1; }
    if (found) return s;
    for (int l = 0; l < c->nL; l++)
        if (c->from[l] == m && c->to[l] == k) { s += c->phi[l]; found = 1; }
    return found ? -s : 0.0;
}

static double flow_bshf(const orc_case *c, int a, int b) {
    for (int l = 0; l < c->nL; l++) {
        if (c->from[l] == a) {
            if (c->to[l] == b) return c->bch[l];
        } else if (c->to[l] == a && c->from[l] == b) {
            return c->bch[l];
        }
    }
    return 0.0;
}

static void flow_km(const orc_case *c, const int *rowptr, const int *colind, const double *y,
                    const double *V, const double *th, int k, int m, double *P, double *Q) {
    double G = 0, B = 0;
    for (int e = rowptr[k]; e < rowptr[k + 1]; e++)
        if (colind[e] == m) { G = y[2 * e]; B = y[2 * e + 1]; break; }
    const double a = flow_phif(c, k, m) + th[k] - th[m];
    const double vkk = V[k] * V[k], vkm = V[k] * V[m];
    *P = vkk * G - vkm * G * cos(a) - vkm * B * sin(a);
    *Q = -vkk * (B + flow_bshf(c, k, m)) + vkm * B * cos(a) - vkm * G * sin(a);
}

void orc_flows(const orc_case *c, const double *V, const double *theta, double *Pdp, double *Qdp,
               double *Ppd, double *Qpd) {
    int nnz = orc_ybus_nnz(c);
    int *rowptr = malloc(sizeof(int) * (c->nB + 1)), *colind = malloc(sizeof(int) * (nnz > 0 ? nnz : 1));
    double *y = malloc(sizeof(double) * 2 * (nnz > 0 ? nnz : 1));
    orc_ybus(c, rowptr, colind, y);
    for (int l = 0; l < c->nL; l++) {
        flow_km(c, rowptr, colind, y, V, theta, c->from[l], c->to[l], &Pdp[l], &Qdp[l]);
        flow_km(c, rowptr, colind, y, V, theta, c->to[l], c->from[l], &Ppd[l], &Qpd[l]);
    }
    free(rowptr); free(colind); free(y);
}
