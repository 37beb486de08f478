/*
 * fsx_score.c -- host side of the scoring path: argument checks, the tables
 * the GPU pipeline consumes (constant cross terms, Wigner-d per beta grid
 * point, d-symbols, twiddles) and the drop-in sxs_compute_saxs_scores.
 *
 * Reference: src/fftsaxs.c:608-986.  The GPU evaluates the same cross-term
 * profiles per requested pose with exact twiddles (the reference's dense-slab
 * FFTW path, fftsaxs.c:899-900) instead of first summing whole (z, beta1,
 * beta2) slabs; where the reference uses its truncated-pi direct DFT
 * (slabs with < 30 poses, fftsaxs.c:529-606) the two differ by ~1e-12
 * relative in I(q).
 */
#include "../fsx_internal.h"
#include "fftsaxs.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

static double now_ms(void)
{
    struct timeval tv;
    gettimeofday(&tv, NULL);
    return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

const char *fsx_error_string(int code)
{
    switch (code) {
    case FSX_OK: return "ok";
    case FSX_ERR_ARG: return "invalid argument";
    case FSX_ERR_HIP: return "HIP runtime error";
    case FSX_ERR_NOMEM: return "out of memory";
    case FSX_ERR_NODEV: return "no HIP device";
    case FSX_ERR_SCRATCH: return "problem too large for the fixed working memory";
    default: return "unknown error";
    }
}

int fsx_device_count(void) { return fsx_hip_device_count(); }

/* fftsaxs.c:27-50: sum_lm Re(A_X conj A_Y) + Re(B_X conj B_Y) per q */
static void const_terms(const fsx_problem *pb, double *cint)
{
    static const int px[6] = {0, 0, 0, 1, 1, 2}, py[6] = {0, 1, 2, 1, 2, 2};
    static const double fac[6] = {1.0, 2.0, 2.0, 1.0, 2.0, 1.0};   /* fftsaxs.c:76-81 */
    int K2 = (pb->L + 1) * (pb->L + 1), Q = pb->qnum;
    size_t blk = (size_t)Q * K2;
    for (int k = 0; k < 6; k++)
        for (int q = 0; q < Q; q++) {
            const double *a1r = pb->A_re + px[k] * blk + (size_t)q * K2;
            const double *a1i = pb->A_im + px[k] * blk + (size_t)q * K2;
            const double *a2r = pb->A_re + py[k] * blk + (size_t)q * K2;
            const double *a2i = pb->A_im + py[k] * blk + (size_t)q * K2;
            const double *b1r = pb->B_re + px[k] * blk + (size_t)q * K2;
            const double *b1i = pb->B_im + px[k] * blk + (size_t)q * K2;
            const double *b2r = pb->B_re + py[k] * blk + (size_t)q * K2;
            const double *b2i = pb->B_im + py[k] * blk + (size_t)q * K2;
            double v = 0.0;
            for (int i = 0; i < K2; i++) {
                v += a1r[i] * a2r[i] + a1i[i] * a2i[i];
                v += b1r[i] * b2r[i] + b1i[i] * b2i[i];
            }
            cint[k * Q + q] = v * fac[k];
        }
}

/* Wigner-d tables of recent beta grids, kept across calls (they depend only
 * on L, nbeta and pi); fsx_score calls are serialised on this lock. */
#define FSX_DW_CACHE 8
static pthread_mutex_t g_dw_lock = PTHREAD_MUTEX_INITIALIZER;
static struct {
    int L, nbeta;
    double pi;
    double *dW;
    unsigned long long key, used;
} g_dw[FSX_DW_CACHE];
static unsigned long long g_dw_serial, g_dw_clock;

/* Device buffers and the host tables (Wigner-d, d-symbols); the next call
 * builds and uploads them again.  Waits for calls in flight. */
void fsx_release(void)
{
    pthread_mutex_lock(&g_dw_lock);
    for (int i = 0; i < FSX_DW_CACHE; i++) {
        free(g_dw[i].dW);
        g_dw[i].dW = NULL;
        g_dw[i].key = 0;
    }
    fsx_d_symb_drop();
    fsx_hip_release();
    pthread_mutex_unlock(&g_dw_lock);
}

/* the Wigner-d table of (L, nbeta, pi), built on a miss in place of the
 * least recently used entry; *key identifies its content */
static const double *dw_table(int L, int nbeta, double pi, unsigned long long *key)
{
    int hit = -1, lru = 0;
    for (int i = 0; i < FSX_DW_CACHE; i++) {
        if (g_dw[i].dW && g_dw[i].L == L && g_dw[i].nbeta == nbeta &&
            memcmp(&g_dw[i].pi, &pi, sizeof(double)) == 0)
            hit = i;
        if (g_dw[i].used < g_dw[lru].used) lru = i;
    }
    if (hit < 0) {
        const size_t dsz = (size_t)(L + 1) * (2 * L + 1) * (2 * L + 1);
        double *dW = (double *)malloc(sizeof(double) * dsz * nbeta);
        if (!dW) return NULL;
        /* fftsaxs.c:632-636, 706-708: beta_i = 0 + i * (pi - 0)/(nbeta - 1) */
        const double bstep = (pi - 0.0) / (nbeta - 1);
        for (int b = 0; b < nbeta; b++) fsx_wigner_d(L, 0.0 + b * bstep, dW + b * dsz);
        hit = lru;
        free(g_dw[hit].dW);
        g_dw[hit].dW = dW;
        g_dw[hit].L = L;
        g_dw[hit].nbeta = nbeta;
        g_dw[hit].pi = pi;
        g_dw[hit].key = ++g_dw_serial;
    }
    g_dw[hit].used = ++g_dw_clock;
    *key = g_dw[hit].key;
    return g_dw[hit].dW;
}

/* per-call host tables of one complex */
typedef struct {
    double *cint, *twa, *twg;
} call_tabs;

static void tabs_free(call_tabs *h)
{
    free(h->cint);
    free(h->twa);
    free(h->twg);
    h->cint = h->twa = h->twg = NULL;
}

/* Checks one complex's arguments and fills the GPU pipeline's arguments
 * (tables in h, owned by the caller of this function).  Returns FSX_OK,
 * an error code, or 1 when there is nothing to score. */
static int prepare(const fsx_problem *pb, const int64_t *poses, int n, double *score, double *c1,
                   double *c2, double *cross, unsigned flags, fsx_hip_args *a, call_tabs *h)
{
    memset(h, 0, sizeof(*h));
    if (!pb || n < 0 || (n > 0 && (!poses || !score || !c1 || !c2))) return FSX_ERR_ARG;
    if (pb->L < 1 || pb->L > 40 || pb->nbeta < 2 || pb->nalpha < 1 || pb->ngamma < 1 ||
        pb->qnum < 2 || pb->znum < 1 || !pb->qvals || !pb->zvals || !pb->A_re ||
        !pb->A_im || !pb->B_re || !pb->B_im || !pb->a)
        return FSX_ERR_ARG;
    if (n == 0) return 1;
    const int L = pb->L;
    const double *ds = fsx_d_symb_cached(L);
    if (!ds) return FSX_ERR_NOMEM;
    unsigned long long dw_key = 0;
    const double *dW = dw_table(L, pb->nbeta, pb->pi, &dw_key);
    h->cint = (double *)malloc(sizeof(double) * 6 * pb->qnum);
    h->twa = (double *)malloc(sizeof(double) * 2 * pb->nalpha);
    h->twg = (double *)malloc(sizeof(double) * 2 * pb->ngamma);
    if (!dW || !h->cint || !h->twa || !h->twg) {
        tabs_free(h);
        return FSX_ERR_NOMEM;
    }
    const_terms(pb, h->cint);
    /* exact twiddles e^{-2 pi i k / N} (the FFTW transform of fftsaxs.c:680) */
    const double two_pi = 6.283185307179586476925286766559;
    for (int k = 0; k < pb->nalpha; k++) {
        h->twa[2 * k] = cos(two_pi * k / pb->nalpha);
        h->twa[2 * k + 1] = -sin(two_pi * k / pb->nalpha);
    }
    for (int k = 0; k < pb->ngamma; k++) {
        h->twg[2 * k] = cos(two_pi * k / pb->ngamma);
        h->twg[2 * k + 1] = -sin(two_pi * k / pb->ngamma);
    }
    memset(a, 0, sizeof(*a));
    a->L = L; a->nbeta = pb->nbeta; a->nalpha = pb->nalpha; a->ngamma = pb->ngamma;
    a->qnum = pb->qnum; a->znum = pb->znum; a->qvals = pb->qvals; a->zvals = pb->zvals;
    a->A_re = pb->A_re; a->A_im = pb->A_im; a->B_re = pb->B_re; a->B_im = pb->B_im;
    a->a = pb->a; a->cint = h->cint; a->dW = dW; a->dsymb = ds; a->tw_a = h->twa; a->tw_g = h->twg;
    a->dw_key = dw_key;
    a->ds_key = (unsigned long long)L + 1;   /* the d-symbol table of L never changes */
    a->mult = pb->mult; a->peak = pb->peak;
    a->c1_lo = pb->c1_lo; a->c1_hi = pb->c1_hi; a->c2_lo = pb->c2_lo; a->c2_hi = pb->c2_hi;
    a->poses = poses; a->n = n; a->score = score; a->c1 = c1; a->c2 = c2; a->cross = cross;
    a->flags = flags;
    return FSX_OK;
}

int fsx_score(const fsx_problem *pb, const int64_t *poses, int n, double *score,
              double *c1, double *c2, double *cross, unsigned flags, fsx_stats *st)
{
    const double t0 = now_ms();
    fsx_stats local;
    if (!st) st = &local;
    memset(st, 0, sizeof(*st));
    pthread_mutex_lock(&g_dw_lock);
    fsx_hip_args a;
    call_tabs h;
    int rc = prepare(pb, poses, n, score, c1, c2, cross, flags, &a, &h);
    if (rc == FSX_OK) {
        st->ms_tables = now_ms() - t0;
        rc = fsx_hip_run(&a, st);
    } else if (rc == 1) {
        rc = FSX_OK;
    }
    tabs_free(&h);
    pthread_mutex_unlock(&g_dw_lock);
    st->ms_total = now_ms() - t0;
    return rc;
}

/* The complexes are prepared on the host, then scored by one batched GPU
 * call (fsx_hip_run_batch): per-complex stage times are not measured
 * there, and every record's ms_total is the whole call's. */
int fsx_score_batch(int nprob, const fsx_problem *pbs, const int64_t *const *poses,
                    const int *n, double *const *score, double *const *c1, double *const *c2,
                    double *const *cross, unsigned flags, fsx_stats *stats)
{
    if (nprob < 0 || (nprob > 0 && (!pbs || !poses || !n || !score || !c1 || !c2)))
        return FSX_ERR_ARG;
    const double t0 = now_ms();
    if (stats) memset(stats, 0, sizeof(fsx_stats) * (size_t)nprob);
    if (nprob == 0) return FSX_OK;
    fsx_hip_args *a = (fsx_hip_args *)calloc(nprob, sizeof(fsx_hip_args));
    call_tabs *h = (call_tabs *)calloc(nprob, sizeof(call_tabs));
    int *idx = (int *)calloc(nprob, sizeof(int));   /* complex of each batch entry */
    fsx_stats *st = (fsx_stats *)calloc(nprob, sizeof(fsx_stats));
    if (!a || !h || !idx || !st) {
        free(a); free(h); free(idx); free(st);
        return FSX_ERR_NOMEM;
    }
    pthread_mutex_lock(&g_dw_lock);
    int rc = FSX_OK, nb = 0;
    for (int i = 0; i < nprob && rc == FSX_OK; i++) {
        const int r = prepare(&pbs[i], poses[i], n[i], score[i], c1[i], c2[i],
                              cross ? cross[i] : NULL, flags, &a[nb], &h[nb]);
        if (r == FSX_OK) idx[nb++] = i;
        else if (r != 1) rc = r;
    }
    const double t1 = now_ms();
    if (rc == FSX_OK && nb > 0) rc = fsx_hip_run_batch(nb, a, st);
    for (int k = 0; k < nprob; k++) tabs_free(&h[k]);
    pthread_mutex_unlock(&g_dw_lock);
    const double t2 = now_ms();
    if (stats && rc == FSX_OK)
        for (int k = 0; k < nb; k++) {
            stats[idx[k]] = st[k];
            stats[idx[k]].ms_tables = t1 - t0;
        }
    if (stats)
        for (int i = 0; i < nprob; i++) stats[i].ms_total = t2 - t0;
    free(a); free(h); free(idx); free(st);
    return rc;
}

/* Flatten the reference structs into [3][qnum][(L+1)^2] arrays. */
static void flatten_spf(const struct sxs_spf_full *s, int L, int qnum, double *re, double *im)
{
    int K2 = (L + 1) * (L + 1);
    size_t blk = (size_t)qnum * K2;
    int Ls = s->L < L ? s->L : L;
    memset(re, 0, sizeof(double) * 3 * blk);
    memset(im, 0, sizeof(double) * 3 * blk);
    for (int q = 0; q < qnum; q++) {
        struct sxs_spf_sing *x[3] = {s->V[q], s->D[q], s->W[q]};
        for (int X = 0; X < 3; X++)
            for (int i = 0; i < (Ls + 1) * (Ls + 1); i++) {
                re[X * blk + (size_t)q * K2 + i] = x[X]->re[i];
                im[X * blk + (size_t)q * K2 + i] = x[X]->im[i];
            }
    }
}

void sxs_compute_saxs_scores(double *scores_list, double *c1_list, double *c2_list,
                             int *index_list, int nout, struct sxs_spf_full *A,
                             struct sxs_spf_full *B, struct sxs_opt_params *params,
                             double *qvals, int qnum, double *zvals, int znum, int L,
                             int skip)
{
    (void)skip;  /* skip = 0 only changes the reference's transform choice (exact twiddles) */
    if (!A || !B || !params || !qvals || !zvals || !index_list || !scores_list || !c1_list ||
        !c2_list) {
        fprintf(stderr, "[Error] sxs_compute_saxs_scores: Null pointer\n");
        exit(EXIT_FAILURE);
    }
    int K2 = (L + 1) * (L + 1);
    size_t blk = (size_t)qnum * K2;
    double *buf = (double *)malloc(sizeof(double) * 12 * blk);
    int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nout > 0 ? nout : 1));
    if (!buf || !idx) {
        fprintf(stderr, "[Error] sxs_compute_saxs_scores: out of memory\n");
        exit(EXIT_FAILURE);
    }
    flatten_spf(A, L, qnum, buf, buf + 3 * blk);
    flatten_spf(B, L, qnum, buf + 6 * blk, buf + 9 * blk);
    for (int i = 0; i < nout; i++) idx[i] = index_list[i];
    fsx_problem pb;
    memset(&pb, 0, sizeof(pb));
    pb.L = L; pb.nbeta = L + 1; pb.nalpha = 2 * L + 1; pb.ngamma = 2 * L + 1;
    pb.qnum = qnum; pb.qvals = qvals; pb.znum = znum; pb.zvals = zvals;
    pb.A_re = buf; pb.A_im = buf + 3 * blk; pb.B_re = buf + 6 * blk; pb.B_im = buf + 9 * blk;
    pb.a = params->a; pb.mult = params->mult; pb.peak = params->peak;
    pb.c1_lo = 0.96; pb.c1_hi = 1.04; pb.c2_lo = -2.00; pb.c2_hi = 4.00;
    pb.pi = FSX_PI_REFERENCE;
    int rc = fsx_score(&pb, idx, nout, scores_list, c1_list, c2_list, NULL, 0, NULL);
    free(buf);
    free(idx);
    if (rc != FSX_OK) {
        fprintf(stderr, "[Error] sxs_compute_saxs_scores: %s\n", fsx_error_string(rc));
        exit(EXIT_FAILURE);
    }
}

int fsx_fit_batch(const double *ct, int n, const double *qvals, int qnum, const double *a,
                  double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                  double *out)
{
    if (n < 0 || (n > 0 && (!ct || !qvals || !a || !out)) || qnum < 2) return FSX_ERR_ARG;
    if (n == 0) return FSX_OK;
    return fsx_hip_fit_batch(ct, n, qvals, qnum, a, mult, c1_lo, c1_hi, c2_lo, c2_hi, out);
}

int fsx_sbessel_batch(const double *x, int n, int pmax, double *out)
{
    if (n < 0 || pmax < 0 || (n > 0 && (!x || !out))) return FSX_ERR_ARG;
    if (n == 0) return FSX_OK;
    return fsx_hip_sbessel_batch(x, n, pmax, out);
}
