/*
 * fsx_api.c -- the reference's data-model API on the host: SPF and profile
 * lifecycles and text formats, fit parameters, the single-profile fitting
 * entry points and the grid index.  Same names, argument meaning, text
 * formats and ownership as src/pdb2spf.c, src/profile.c, src/min_saxs.c,
 * src/index.c; fatal errors print and exit(EXIT_FAILURE) like common.h:26-36.
 */
#include "../fsx_internal.h"
#include "../fsx_lbfgsb.h"
#include "../fsx_lb2.h"
#include "fftsaxs.h"
#include "min_saxs.h"
#include "pdb2spf.h"
#include "profile.h"
#include "index.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FSX_REL_ERR 0.05   /* define.h:23 */

static void fsx_fatal(const char *func, const char *msg)
{
    fprintf(stderr, "[Error] %s: %s\n", func, msg);
    exit(EXIT_FAILURE);
}

static void *xcalloc(size_t n, size_t s)
{
    void *p = calloc(n ? n : 1, s);
    if (!p) fsx_fatal("calloc", "out of memory");
    return p;
}

/* ---- pdb2spf.c:178-350 --------------------------------------------------- */
struct sxs_spf_sing *sxs_spf_sing_create(int L)
{
    struct sxs_spf_sing *s = (struct sxs_spf_sing *)xcalloc(1, sizeof(*s));
    sxs_spf_sing_init(s, L);
    return s;
}

void sxs_spf_sing_init(struct sxs_spf_sing *s, int L)
{
    if (!s) return;
    s->L = L;
    s->re = (double *)xcalloc((size_t)(L + 1) * (L + 1), sizeof(double));
    s->im = (double *)xcalloc((size_t)(L + 1) * (L + 1), sizeof(double));
}

void sxs_spf_sing_destroy(struct sxs_spf_sing *s)
{
    if (!s) return;
    free(s->re);
    free(s->im);
}

void sxs_spf_sing_free(struct sxs_spf_sing *s)
{
    sxs_spf_sing_destroy(s);
    free(s);
}

struct sxs_spf_full *sxs_spf_full_create(int L, int qnum)
{
    struct sxs_spf_full *s = (struct sxs_spf_full *)xcalloc(1, sizeof(*s));
    sxs_spf_full_init(s, L, qnum);
    return s;
}

void sxs_spf_full_init(struct sxs_spf_full *s, int L, int qnum)
{
    if (!s) return;
    s->L = L;
    s->qnum = qnum;
    s->V = (struct sxs_spf_sing **)xcalloc((size_t)qnum, sizeof(void *));
    s->D = (struct sxs_spf_sing **)xcalloc((size_t)qnum, sizeof(void *));
    s->W = (struct sxs_spf_sing **)xcalloc((size_t)qnum, sizeof(void *));
    for (int i = 0; i < qnum; i++) {
        s->V[i] = sxs_spf_sing_create(L);
        s->D[i] = sxs_spf_sing_create(L);
        s->W[i] = sxs_spf_sing_create(L);
    }
}

void sxs_spf_full_destroy(struct sxs_spf_full *s)
{
    if (!s) return;
    for (int i = 0; i < s->qnum; i++) {
        sxs_spf_sing_free(s->V[i]);
        sxs_spf_sing_free(s->D[i]);
        sxs_spf_sing_free(s->W[i]);
    }
    free(s->V);
    free(s->D);
    free(s->W);
    s->V = s->D = s->W = NULL;
}

void sxs_spf_full_free(struct sxs_spf_full *s)
{
    sxs_spf_full_destroy(s);
    free(s);
}

void sxs_spf_full_reset(struct sxs_spf_full *s)
{
    size_t sz = sizeof(double) * (size_t)(s->L + 1) * (s->L + 1);
    for (int i = 0; i < s->qnum; i++) {
        memset(s->V[i]->re, 0, sz); memset(s->V[i]->im, 0, sz);
        memset(s->D[i]->re, 0, sz); memset(s->D[i]->im, 0, sz);
        memset(s->W[i]->re, 0, sz); memset(s->W[i]->im, 0, sz);
    }
}

void sxs_spf_full_write(char *path, struct sxs_spf_full *s)
{
    FILE *f = fopen(path, "w");
    if (!f) fsx_fatal("sxs_spf_full_write", "unable to open file");
    fprintf(f, "% i % i % .4f\n", s->L, s->qnum, s->rm);
    for (int q = 0; q < s->qnum; q++)
        for (int l = 0; l <= s->L; l++)
            for (int m = -l; m <= l; m++) {
                int i = l * (l + 1) + m;
                fprintf(f, "% .4e % .4e % .4e % .4e % .4e % .4e\n", s->V[q]->re[i],
                        s->V[q]->im[i], s->D[q]->re[i], s->D[q]->im[i], s->W[q]->re[i],
                        s->W[q]->im[i]);
            }
    fclose(f);
}

struct sxs_spf_full *sxs_spf_full_fread(FILE *f)
{
    if (!f) return NULL;
    int L, qnum;
    double rm;
    if (fscanf(f, "%i %i %lf\n", &L, &qnum, &rm) != 3) fsx_fatal("sxs_spf_full_fread", "Wrong input file format.");
    struct sxs_spf_full *s = sxs_spf_full_create(L, qnum);
    s->rm = rm;
    for (int q = 0; q < qnum; q++)
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++) {
                int i = l * (l + 1) + m;
                if (fscanf(f, "%lf %lf %lf %lf %lf %lf\n", &s->V[q]->re[i], &s->V[q]->im[i],
                           &s->D[q]->re[i], &s->D[q]->im[i], &s->W[q]->re[i],
                           &s->W[q]->im[i]) != 6)
                    fsx_fatal("sxs_spf_full_fread", "Wrong input file format.");
            }
    return s;
}

struct sxs_spf_full *sxs_spf_full_read(char *path)
{
    FILE *f = fopen(path, "r");
    struct sxs_spf_full *s = sxs_spf_full_fread(f);
    if (f) fclose(f);
    return s;
}

/* ---- profile.c ------------------------------------------------------------ */
struct sxs_profile *sxs_profile_create(double *qvals, int qnum, int cross_terms_flag)
{
    struct sxs_profile *p = (struct sxs_profile *)xcalloc(1, sizeof(*p));
    sxs_profile_init(p, qvals, qnum, cross_terms_flag);
    return p;
}

void sxs_profile_init(struct sxs_profile *p, double *qvals, int qnum, int cross_terms_flag)
{
    if (!p) return;
    p->qnum = qnum;
    p->rerr = FSX_REL_ERR;
    p->qvals = qvals;
    p->in = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->err = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->score = INFINITY;
    p->c1 = 1.0;
    p->c2 = 0.0;
    p->scale = 1.0;
    if (cross_terms_flag == 1) {
        sxs_profile_alloc_cross_terms(p, qnum);
    } else {
        p->VV = p->VD = p->VW = p->DD = p->DW = p->WW = NULL;
    }
}

void sxs_profile_alloc_cross_terms(struct sxs_profile *p, int qnum)
{
    p->VV = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->VD = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->VW = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->DD = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->DW = (double *)xcalloc((size_t)qnum, sizeof(double));
    p->WW = (double *)xcalloc((size_t)qnum, sizeof(double));
}

void sxs_profile_destroy(struct sxs_profile *p)
{
    p->qvals = NULL;
    free(p->in); free(p->err);
    free(p->VV); free(p->VD); free(p->VW); free(p->DD); free(p->DW); free(p->WW);
    p->in = p->err = p->VV = p->VD = p->VW = p->DD = p->DW = p->WW = NULL;
}

void sxs_profile_free(struct sxs_profile *p)
{
    sxs_profile_destroy(p);
    free(p);
}

void sxs_profile_write(char *path, struct sxs_profile *p)
{
    FILE *f = fopen(path, "w");
    if (!f) {
        fprintf(stderr, "Unable to open %s\n", path);
        exit(EXIT_FAILURE);
    }
    for (int i = 0; i < p->qnum; i++) {
        double err = p->err[i] > 0.0 ? p->err[i] : p->in[i] * p->rerr;
        fprintf(f, "%.4f %.4f %.4f\n", p->qvals[i], p->in[i], err);
    }
    fclose(f);
}

struct sxs_profile *sxs_profile_fread(FILE *f)
{
    if (!f) return NULL;
    int c, qnum = 0;
    double b1, b2, b3;
    while ((c = fscanf(f, "%lf %lf %lf", &b1, &b2, &b3)) != EOF) {
        if (c != 3) fsx_fatal("sxs_profile_fread", "Wrong input file format.");
        qnum++;
    }
    rewind(f);
    double *qvals = (double *)xcalloc((size_t)qnum, sizeof(double));
    struct sxs_profile *p = sxs_profile_create(qvals, qnum, 0);
    for (int q = 0; q < qnum; q++)
        if (fscanf(f, "%lf %lf %lf\n", &qvals[q], &p->in[q], &p->err[q]) != 3)
            fsx_fatal("sxs_profile_fread", "Wrong input file format.");
    return p;
}

struct sxs_profile *sxs_profile_read(char *path)
{
    FILE *f = fopen(path, "r");
    struct sxs_profile *p = sxs_profile_fread(f);
    if (f) fclose(f);
    return p;
}

static double mult_of(double rm)
{
    return pow(4.0 * FSX_PI_REFERENCE / 3.0, 3.0 / 2.0) * rm * rm / (16.0 * FSX_PI_REFERENCE);
}

void sxs_profile_from_spf(struct sxs_profile *p, struct sxs_spf_full *s, double c1, double c2)
{
    if (!p || !s) fsx_fatal("sxs_profile_from_spf", "Null pointer");
    double corr = -mult_of(s->rm) * (c1 * c1 - 1.0);
    for (int q = 0; q < p->qnum; q++) {
        const double *vr = s->V[q]->re, *vi = s->V[q]->im, *dr = s->D[q]->re;
        const double *di = s->D[q]->im, *wr = s->W[q]->re, *wi = s->W[q]->im;
        /* G at qvals[0] for every q, as in profile.c:214 */
        double G = c1 * c1 * c1 * exp(corr * p->qvals[0] * p->qvals[0]);
        double in = 0.0;
        for (int l = 0; l <= s->L; l++)
            for (int m = -l; m <= l; m++) {
                int i = l * (l + 1) + m;
                double ar = vr[i] - G * dr[i] + c2 * wr[i];
                double ai = vi[i] - G * di[i] + c2 * wi[i];
                in += ar * ar + ai * ai;
            }
        p->in[q] = in;
        p->err[q] = in * p->rerr;
    }
    p->c1 = c1;
    p->c2 = c2;
}

/* ---- min_saxs.c ----------------------------------------------------------- */
double *scoring_helper(struct sxs_profile *exp, int qnum, double *qvals)
{
    double *a = (double *)xcalloc((size_t)6 * qnum, sizeof(double));
    int j = 0, nexp = exp->qnum;
    for (int i = 0; i < qnum; i++) {
        double lo = i > 0 ? qvals[i - 1] : qvals[0], hi = qvals[i];
        /* the reference scans without a bound (min_saxs.c:369); stop at the end */
        while (j < nexp && exp->qvals[j] >= lo && exp->qvals[j] <= hi) {
            double e2 = exp->err[j] * exp->err[j];
            a[i * 6 + 0] += exp->in[j] * exp->in[j] / e2;
            a[i * 6 + 1] += exp->in[j] / e2;
            a[i * 6 + 2] += exp->qvals[j] * exp->in[j] / e2;
            a[i * 6 + 3] += 1.0 / e2;
            a[i * 6 + 4] += exp->qvals[j] / e2;
            a[i * 6 + 5] += exp->qvals[j] * exp->qvals[j] / e2;
            j++;
        }
    }
    double norm = (double)j;
    for (int i = 0; i < 6 * qnum; i++) a[i] /= norm;
    return a;
}

struct sxs_opt_params *sxs_opt_params_create(struct sxs_profile *exp, double *qvals, int qnum, double rm)
{
    struct sxs_opt_params *p = (struct sxs_opt_params *)xcalloc(1, sizeof(*p));
    sxs_opt_params_init(p, exp, qvals, qnum, rm);
    return p;
}

void sxs_opt_params_init(struct sxs_opt_params *p, struct sxs_profile *exp, double *qvals, int qnum, double rm)
{
    if (!p) return;
    p->a = scoring_helper(exp, qnum, qvals);
    p->rm = rm;
    p->mult = pow(4.0 * FSX_PI_REFERENCE / 3.0, 1.5) * rm * rm / (16.0 * FSX_PI_REFERENCE);
    p->peak = exp->in[0];
}

void sxs_opt_params_destroy(struct sxs_opt_params *p)
{
    if (!p) return;
    free(p->a);
    p->a = NULL;
}

void sxs_opt_params_free(struct sxs_opt_params *p)
{
    sxs_opt_params_destroy(p);
    free(p);
}

static fsx_fitpar fitpar_of(const struct sxs_profile *pr, const struct sxs_opt_params *p)
{
    fsx_fitpar fp;
    fp.a = p->a;
    fp.q = pr->qvals;
    fp.rdq = NULL;
    fp.qnum = pr->qnum;
    fp.mult = p->mult;
    fp.lo0 = 0.96; fp.hi0 = 1.04; fp.lo1 = -2.00; fp.hi1 = 4.00;   /* define.h:28-31 */
    fp.factr = 1e+7;
    fp.pgtol = 1e-5;
    fp.onepass = 0;
    fp.a0sum = 0.0;
    return fp;
}

static fsx_ct ct_packed(const double *c, int qnum)
{
    fsx_ct t;
    t.c = c; t.v = NULL; t.cks = qnum; t.cqs = 1; t.vks = 0; t.vqs = 0; t.scale = 1.0;
    return t;
}

static double *pack_ct(const struct sxs_profile *pr)
{
    int Q = pr->qnum;
    double *c = (double *)xcalloc((size_t)6 * Q, sizeof(double));
    memcpy(c + 0 * Q, pr->VV, sizeof(double) * Q);
    memcpy(c + 1 * Q, pr->VD, sizeof(double) * Q);
    memcpy(c + 2 * Q, pr->VW, sizeof(double) * Q);
    memcpy(c + 3 * Q, pr->DD, sizeof(double) * Q);
    memcpy(c + 4 * Q, pr->DW, sizeof(double) * Q);
    memcpy(c + 5 * Q, pr->WW, sizeof(double) * Q);
    return c;
}

double sxs_best_scale(struct sxs_profile *pr, struct sxs_opt_params *p, double c1, double c2)
{
    double *c = pack_ct(pr);
    fsx_ct t = ct_packed(c, pr->qnum);
    fsx_fitpar fp = fitpar_of(pr, p);
    double k = fsx_best_scale(&t, &fp, c1, c2);
    free(c);
    return k;
}

double sxs_point_score(struct sxs_profile *pr, struct sxs_opt_params *p, double c1, double c2)
{
    double *c = pack_ct(pr);
    fsx_ct t = ct_packed(c, pr->qnum);
    fsx_fitpar fp = fitpar_of(pr, p);
    double k = fsx_best_scale(&t, &fp, c1, c2);
    double x[2] = {c1, c2}, g[2], f;
    fsx_objective(&t, &fp, k, x, g, &f);
    free(c);
    return f;
}

void sxs_compile_intensity(struct sxs_profile *pr, double rm, double c1, double c2)
{
    double corr = -mult_of(rm) * (c1 * c1 - 1.0);
    for (int i = 0; i < pr->qnum; i++) {
        double G = c1 * c1 * c1 * exp(corr * pr->qvals[i] * pr->qvals[i]);
        pr->in[i] = pr->VV[i] - G * pr->VD[i] + c2 * pr->VW[i] + G * G * pr->DD[i]
                  - G * c2 * pr->DW[i] + c2 * c2 * pr->WW[i];
        pr->err[i] = pr->rerr * pr->in[i];
    }
}

void sxs_lbfgs_fitting(struct sxs_profile *pr, struct sxs_opt_params *p)
{
    double *c = pack_ct(pr);
    fsx_ct t = ct_packed(c, pr->qnum);
    fsx_fitpar fp = fitpar_of(pr, p);
    double x[2] = {1.0, 0.0}, f;
    int nev;
    fsx_lbfgsb_fit(&t, &fp, x, &f, &nev);
    free(c);
    pr->score = sqrt(f);
    pr->scale = sxs_best_scale(pr, p, x[0], x[1]);
    pr->c1 = x[0];
    pr->c2 = x[1];
    sxs_compile_intensity(pr, p->rm, x[0], x[1]);
}

void sxs_fit_params(struct sxs_profile **profiles, struct sxs_opt_params *p, int *mask, int n)
{
    for (int i = 0; i < n; i++) {
        if (mask[i] != 1) continue;
        struct sxs_profile *pr = profiles[i];
        double s = pr->VV[0] + pr->DD[0] + pr->WW[0] + pr->VW[0] - pr->VD[0] - pr->DW[0];
        s = p->peak / s;
        for (int q = 0; q < pr->qnum; q++) {
            pr->VV[q] *= s; pr->DD[q] *= s; pr->WW[q] *= s;
            pr->VW[q] *= s; pr->VD[q] *= s; pr->DW[q] *= s;
        }
        sxs_lbfgs_fitting(pr, p);
    }
}

void sxs_spf2cross_terms(struct sxs_profile *pr, struct sxs_spf_full *s)
{
    for (int q = 0; q < pr->qnum; q++) {
        const double *vr = s->V[q]->re, *vi = s->V[q]->im, *dr = s->D[q]->re;
        const double *di = s->D[q]->im, *wr = s->W[q]->re, *wi = s->W[q]->im;
        double vv = 0, vd = 0, vw = 0, dd = 0, dw = 0, ww = 0;
        int K = (s->L + 1) * (s->L + 1);
        for (int i = 0; i < K; i++) {
            vv += vr[i] * vr[i] + vi[i] * vi[i];
            dd += dr[i] * dr[i] + di[i] * di[i];
            ww += wr[i] * wr[i] + wi[i] * wi[i];
            vd += vr[i] * dr[i] + vi[i] * di[i];
            vw += vr[i] * wr[i] + vi[i] * wi[i];
            dw += dr[i] * wr[i] + di[i] * wi[i];
        }
        pr->VV[q] = vv; pr->DD[q] = dd; pr->WW[q] = ww;
        pr->VD[q] = vd * 2.0; pr->VW[q] = vw * 2.0; pr->DW[q] = dw * 2.0;
    }
}

void sxs_spf2fitted_profile(struct sxs_profile *pr, struct sxs_spf_full *s, struct sxs_opt_params *p)
{
    sxs_spf2cross_terms(pr, s);
    sxs_lbfgs_fitting(pr, p);
}

int fsx_fit_host(const double *ct, const double *qvals, int qnum, const double *a,
                 double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                 double *out4)
{
    return fsx_fit_host_form(ct, qvals, qnum, a, mult, c1_lo, c1_hi, c2_lo, c2_hi, 0u, out4);
}

int fsx_fit_host_form(const double *ct, const double *qvals, int qnum, const double *a,
                      double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                      unsigned flags, double *out4)
{
    if (!ct || !qvals || !a || !out4 || qnum < 2) return FSX_ERR_ARG;
    fsx_ct t = ct_packed(ct, qnum);
    fsx_fitpar fp;
    fp.a = a; fp.q = qvals; fp.rdq = NULL; fp.qnum = qnum; fp.mult = mult;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    fp.onepass = (flags & FSX_MIXED_FP32) != 0;
    fp.a0sum = fsx_a0sum(a, qnum);
    double x[2] = {1.0, 0.0}, f;
    int nev;
    if ((flags & FSX_MIXED_FP32) && !(flags & FSX_FIT_COMPACT))
        fsx_lb2_fit(&t, &fp, x, &f, &nev);
    else
        fsx_lbfgsb_fit(&t, &fp, x, &f, &nev);
    out4[0] = sqrt(f);
    out4[1] = x[0];
    out4[2] = x[1];
    out4[3] = nev;
    return FSX_OK;
}

/* ---- index.c:3-29 ---------------------------------------------------------- */
size_t sxs_assemble_index(struct sxs_index *id, int nbeta, int L)
{
    int n = 2 * L + 1;
    return ((((id->z * nbeta + id->b1) * nbeta + id->b2) * n + id->a2) * n + id->g1) * n + id->g2;
}

void sxs_disassemble_index(struct sxs_index *id, size_t index, int nbeta, int L)
{
    int n = 2 * L + 1;
    int t;
    id->g2 = (int)(index % n);
    t = (int)(index / n);
    id->g1 = t % n;
    t = t / n;
    id->a2 = t % n;
    t = t / n;
    id->b2 = t % nbeta;
    t = t / nbeta;
    id->b1 = t % nbeta;
    id->z = t / nbeta;
}
