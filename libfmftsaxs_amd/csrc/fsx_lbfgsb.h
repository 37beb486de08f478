/*
 * fsx_lbfgsb.h -- the per-pose c1/c2 chi^2 fit: bounded L-BFGS-B (n = 2,
 * m = 3) on the interpolated chi^2 objective with the closed-form scale.
 *
 * One implementation compiled twice: as plain C11 for the host entry points
 * (sxs_lbfgs_fitting, sxs_fit_params) and as HIP device code for the fused
 * per-pose fit kernel.  State lives in a fixed-size struct (no allocation),
 * so a GPU thread owns one fit.
 *
 * Semantics follow the reference:
 *   objective    min_saxs.c:3-105 (gradient), :261-319 (sxs_best_scale)
 *   wrapper      min_saxs.c:196-259 (x0 = (1, 0), m = 3, factr 1e7, pgtol 1e-5,
 *                scale refreshed before every f/g evaluation)
 *   optimiser    lbfgsb/src (L-BFGS-B 3.0): setulb/mainlb lbfgsb.c:52-515,
 *                cauchy/freev/formk/formt/cmprlb/subsm/matupd/projgr/hpsolb/bmv
 *                in subalgorithms.c, lnsrlb/dcsrch/dcstep in linesearch.c,
 *                dpofa/dtrsl in linpack.c.
 * The reverse-communication state machine of setulb is run as one direct
 * loop (the caller in min_saxs.c resumes every NEW_X immediately, so the
 * visited points are the same).  Every arithmetic expression keeps the
 * reference's evaluation order; compile with FP contraction off so that the
 * iterates are bit-identical to the reference's for identical inputs.
 * The reference loop has no iteration cap (min_saxs.c:234-248);
 * FSX_LB_MAX_EVALS bounds it here.
 */
#ifndef FSX_LBFGSB_H
#define FSX_LBFGSB_H

#include <math.h>
#include <float.h>

#ifdef __HIPCC__
#define FSX_HD __host__ __device__ static inline
#else
#define FSX_HD static inline
#endif

#define FSX_LB_N 2
#define FSX_LB_M 3
#define FSX_LB_MAX_EVALS 2000

/* task codes (lbfgsb.h:80-130 numbering) */
#define FSX_T_START 1
#define FSX_T_NEW_X 2
#define FSX_T_ABNORMAL 3
#define FSX_T_FG 10
#define FSX_T_FG_LN 11
#define FSX_T_CONVERGENCE 20
#define FSX_T_CONV_GRAD 21
#define FSX_T_CONV_F 22
#define FSX_T_STOP_ITER 32
#define FSX_T_W_ROUND 101
#define FSX_T_W_XTOL 102
#define FSX_T_W_STPMAX 103
#define FSX_T_W_STPMIN 104
#define FSX_T_E_SMALLSTP 201
#define FSX_T_E_LARGESTP 202
#define FSX_T_E_INITIAL 203
#define FSX_T_E_STP0 207
#define FSX_T_E_STP1 208

/* Cross-term profile of one pose: value of term k (0..5 = VV,VD,VW,DD,DW,WW)
 * at q index i is scale * (c[k*cks + i*cqs] + v[k*vks + i*vqs]); v may be
 * NULL.  This lets the GPU read (constant part + pose part) in place and
 * apply the peak rescale on the fly without a private copy. */
typedef struct fsx_ct {
    const double *c;
    const double *v;
    long cks, cqs, vks, vqs;
    double scale;
} fsx_ct;

FSX_HD double fsx_ct_get(const fsx_ct *t, int k, int i)
{
    double x = t->c[k * t->cks + i * t->cqs];
    if (t->v) x = x + t->v[k * t->vks + i * t->vqs];
    return x * t->scale;
}

typedef struct fsx_fitpar {
    const double *a;      /* 6*qnum moments (scoring_helper) */
    const double *q;      /* model q grid */
    int qnum;
    double mult;          /* (4pi/3)^1.5 rm^2 / (16 pi) */
    double lo0, hi0, lo1, hi1;
    double factr, pgtol;
} fsx_fitpar;

/* The six terms at q index i, each (c + v) * scale as fsx_ct_get.  Fetched
 * together (and one q ahead in the loops below) so that a GPU thread has
 * all its loads of a step in flight at once. */
FSX_HD void fsx_ct_get6(const fsx_ct *t, int i, double *o)
{
    const double *c = t->c + i * t->cqs;
    if (t->v) {
        const double *v = t->v + i * t->vqs;
        for (int k = 0; k < 6; k++) {
            double x = c[k * t->cks];
            x = x + v[k * t->vks];
            o[k] = x * t->scale;
        }
    } else {
        for (int k = 0; k < 6; k++) o[k] = c[k * t->cks] * t->scale;
    }
}

/* Terms at q = i0 .. i0+nb-1 (nb <= FSX_QB) into blk: all raw loads first,
 * then (c + v) * scale per value exactly as fsx_ct_get. */
#define FSX_QB 4
FSX_HD void fsx_ct_block(const fsx_ct *t, int i0, int nb, double blk[FSX_QB][6])
{
    double cr[FSX_QB][6];
    const double *c = t->c + i0 * t->cqs;
    const long cks = t->cks, cqs = t->cqs;
#ifdef __HIPCC__
#pragma unroll
#endif
    for (int u = 0; u < FSX_QB; u++)
        for (int k = 0; k < 6; k++) cr[u][k] = u < nb ? c[u * cqs + k * cks] : 0.0;
    if (t->v) {
        double vr[FSX_QB][6];
        const double *v = t->v + i0 * t->vqs;
        const long vks = t->vks, vqs = t->vqs;
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) vr[u][k] = u < nb ? v[u * vqs + k * vks] : 0.0;
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) {
                double x = cr[u][k];
                x = x + vr[u][k];
                blk[u][k] = x * t->scale;
            }
    } else {
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) blk[u][k] = cr[u][k] * t->scale;
    }
}

/* I(q) at G = c1^3 exp(-mult (c1^2 - 1) q^2) and c2 (min_saxs.c:285-287) */
#define FSX_IN(v, G, c2)                                                              \
    ((v)[0] - (G) * (v)[1] + (c2) * (v)[2] + (G) * (G) * (v)[3] - (G) * (c2) * (v)[4] \
     + (c2) * (c2) * (v)[5])

FSX_HD double fsx_in(const fsx_ct *t, int i, double G, double c2)
{
    double v[6];
    fsx_ct_get6(t, i, v);
    return FSX_IN(v, G, c2);
}

/* The q loops below fetch the profile FSX_QB q values at a time (all loads
 * of a block in flight together), then run the steps of the block in order;
 * the arithmetic and its order are those of the reference loops. */

/* min_saxs.c:261-319 */
FSX_HD double fsx_best_scale(const fsx_ct *t, const fsx_fitpar *p, double c1, double c2)
{
    const double *a = p->a, *q = p->q;
    const int qnum = p->qnum;
    double v0[6];
    fsx_ct_get6(t, 0, v0);
    double corr = -p->mult * (c1 * c1 - 1.0);
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double in_prev = FSX_IN(v0, G, c2);
    double c1c = c1 * c1 * c1;
    double q_prev = -1.0, up = 0.0, down = 0.0;
    for (int i0 = 0; i0 < qnum; i0 += FSX_QB) {
        double blk[FSX_QB][6];
        fsx_ct_block(t, i0, qnum - i0 < FSX_QB ? qnum - i0 : FSX_QB, blk);
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++) {
            const int i = i0 + u;
            if (i < qnum) {
                const double *cur = blk[u];
                double qc = q[i];
                G = c1c * exp(corr * qc * qc);
                double in = FSX_IN(cur, G, c2);
                double tan_ = (in - in_prev) / (qc - q_prev);
                double buf = in - tan_ * qc;
                up += buf * a[i * 6 + 1] + tan_ * a[i * 6 + 2];
                down += buf * buf * a[i * 6 + 3] + 2.0 * tan_ * buf * a[i * 6 + 4]
                      + tan_ * tan_ * a[i * 6 + 5];
                in_prev = in;
                q_prev = qc;
            }
        }
    }
    return up / down;
}

/* min_saxs.c:3-105 */
FSX_HD void fsx_objective(const fsx_ct *t, const fsx_fitpar *p, double k,
                          const double *x, double *g, double *fval)
{
    const double *a = p->a, *q = p->q;
    const int qnum = p->qnum;
    double v0[6];
    fsx_ct_get6(t, 0, v0);
    double c1 = x[0], c2 = x[1];
    double g0 = 0.0, g1 = 0.0, mult = p->mult;
    double corr = -mult * (c1 * c1 - 1.0);
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double Gd = G * (3.0 / c1 - 2.0 * c1 * mult * q[0] * q[0]);
    double in_prev = FSX_IN(v0, G, c2);
    double d1p = -Gd * v0[1] + 2.0 * G * Gd * v0[3] - Gd * c2 * v0[4];
    double d2p = v0[2] - G * v0[4] + 2.0 * c2 * v0[5];
    double q_prev = -1.0, score = 0.0, c1c = c1 * c1 * c1;
    for (int i0 = 0; i0 < qnum; i0 += FSX_QB) {
        double blk[FSX_QB][6];
        fsx_ct_block(t, i0, qnum - i0 < FSX_QB ? qnum - i0 : FSX_QB, blk);
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++) {
            const int i = i0 + u;
            if (i < qnum) {
                const double *cur = blk[u];
                double qc = q[i];
                G = c1c * exp(corr * qc * qc);
                Gd = G * (3.0 / c1 - 2.0 * c1 * mult * qc * qc);
                double vd = cur[1], vw = cur[2], dd = cur[3];
                double dw = cur[4], ww = cur[5];
                double in = cur[0] - G * vd + c2 * vw + G * G * dd - G * c2 * dw + c2 * c2 * ww;
                double d1 = Gd * (-vd + 2.0 * G * dd - c2 * dw);
                double d2 = vw - G * dw + 2.0 * c2 * ww;
                double buf = 1.0 / (qc - q_prev);
                double tan_ = (in - in_prev) * buf;
                double t1 = (d1 - d1p) * buf;
                double t2 = (d2 - d2p) * buf;
                const double *ai = a + i * 6;
                g0 += 2.0 * k * (-(d1 - t1 * qc) * ai[1] - t1 * ai[2]
                                 + k * ((in - tan_ * qc) * (d1 - t1 * qc) * ai[3]
                                        + (in * t1 + d1 * tan_ - 2.0 * tan_ * t1 * qc) * ai[4]
                                        + tan_ * t1 * ai[5]));
                g1 += 2.0 * k * (-(d2 - t2 * qc) * ai[1] - t2 * ai[2]
                                 + k * ((in - tan_ * qc) * (d2 - t2 * qc) * ai[3]
                                        + (in * t2 + d2 * tan_ - 2.0 * tan_ * t2 * qc) * ai[4]
                                        + tan_ * t2 * ai[5]));
                buf = in - tan_ * qc;
                score += ai[0] + k * (-2.0 * buf * ai[1] - 2.0 * tan_ * ai[2]
                                      + k * (buf * buf * ai[3] + 2.0 * buf * tan_ * ai[4]
                                             + tan_ * tan_ * ai[5]));
                in_prev = in;
                d1p = d1;
                d2p = d2;
                q_prev = qc;
            }
        }
    }
    g[0] = g0;
    g[1] = g1;
    *fval = score;
}

/* ---------------------------------------------------------------------- */
/* L-BFGS-B state for n = 2, m = 3; matrices column-major, 1-based macros. */
typedef struct fsx_lb {
    double ws[FSX_LB_N * FSX_LB_M], wy[FSX_LB_N * FSX_LB_M];
    double sy[FSX_LB_M * FSX_LB_M], ss[FSX_LB_M * FSX_LB_M], wt[FSX_LB_M * FSX_LB_M];
    double wn[4 * FSX_LB_M * FSX_LB_M], snd[4 * FSX_LB_M * FSX_LB_M];
    double z[FSX_LB_N], r[FSX_LB_N], d[FSX_LB_N], t[FSX_LB_N], xp[FSX_LB_N];
    double wa[8 * FSX_LB_M];
    int index[FSX_LB_N], iwhere[FSX_LB_N], indx2[FSX_LB_N];
} fsx_lb;

#define LB_WS(i, j) w->ws[((i) - 1) + ((j) - 1) * FSX_LB_N]
#define LB_WY(i, j) w->wy[((i) - 1) + ((j) - 1) * FSX_LB_N]
#define LB_SY(i, j) w->sy[((i) - 1) + ((j) - 1) * FSX_LB_M]
#define LB_SS(i, j) w->ss[((i) - 1) + ((j) - 1) * FSX_LB_M]
#define LB_WT(i, j) w->wt[((i) - 1) + ((j) - 1) * FSX_LB_M]
#define LB_WN(i, j) w->wn[((i) - 1) + ((j) - 1) * 2 * FSX_LB_M]
#define LB_WN1(i, j) w->snd[((i) - 1) + ((j) - 1) * 2 * FSX_LB_M]

FSX_HD double lb_max(double a, double b) { return a >= b ? a : b; }
FSX_HD double lb_min(double a, double b) { return a <= b ? a : b; }
FSX_HD double lb_abs(double a) { return a >= 0 ? a : -a; }

/* miniCBLAS.c: ddot is a left-to-right sum from 0 */
FSX_HD double lb_dot(int n, const double *x, const double *y)
{
    double s = 0.;
    for (int i = 0; i < n; i++) s = s + x[i] * y[i];
    return s;
}
FSX_HD void lb_axpy(int n, double a, const double *x, double *y)
{
    if (n <= 0 || a == 0.) return;
    for (int i = 0; i < n; i++) y[i] += a * x[i];
}

/* linpack.c dpofa on a(lda, n) */
FSX_HD int lb_chol(double *a, int lda, int n)
{
    for (int j = 1; j <= n; ++j) {
        double s = 0.;
        for (int k = 1; k <= j - 1; ++k) {
            double t = a[(k - 1) + (j - 1) * lda]
                     - lb_dot(k - 1, &a[(k - 1) * lda], &a[(j - 1) * lda]);
            t /= a[(k - 1) + (k - 1) * lda];
            a[(k - 1) + (j - 1) * lda] = t;
            s += t * t;
        }
        s = a[(j - 1) + (j - 1) * lda] - s;
        if (s <= 0.) return j;
        a[(j - 1) + (j - 1) * lda] = sqrt(s);
    }
    return 0;
}

/* linpack.c dtrsl, job 01 (T x = b) and job 11 (T' x = b), T upper */
FSX_HD int lb_trsl(const double *t, int ldt, int n, double *b, int job)
{
    for (int i = 1; i <= n; ++i)
        if (t[(i - 1) + (i - 1) * ldt] == 0.) return i;
    if (job == 1) {
        b[n - 1] /= t[(n - 1) + (n - 1) * ldt];
        for (int jj = 2; jj <= n; ++jj) {
            int j = n - jj + 1;
            double tmp = -b[j];
            lb_axpy(j, tmp, &t[j * ldt], b);
            b[j - 1] /= t[(j - 1) + (j - 1) * ldt];
        }
    } else {
        b[0] /= t[0];
        for (int j = 2; j <= n; ++j) {
            b[j - 1] -= lb_dot(j - 1, &t[(j - 1) * ldt], b);
            b[j - 1] /= t[(j - 1) + (j - 1) * ldt];
        }
    }
    return 0;
}

/* subalgorithms.c bmv */
FSX_HD int lb_bmv(fsx_lb *w, int col, const double *v, double *p)
{
    if (col == 0) return 0;
    p[col] = v[col];
    for (int i = 2; i <= col; ++i) {
        int i2 = col + i;
        double sum = 0.;
        for (int k = 1; k <= i - 1; ++k) sum += LB_SY(i, k) * v[k - 1] / LB_SY(k, k);
        p[i2 - 1] = v[i2 - 1] + sum;
    }
    int info = lb_trsl(w->wt, FSX_LB_M, col, &p[col], 11);
    if (info) return info;
    for (int i = 1; i <= col; ++i) p[i - 1] = v[i - 1] / sqrt(LB_SY(i, i));
    info = lb_trsl(w->wt, FSX_LB_M, col, &p[col], 1);
    if (info) return info;
    for (int i = 1; i <= col; ++i) p[i - 1] = -p[i - 1] / sqrt(LB_SY(i, i));
    for (int i = 1; i <= col; ++i) {
        double sum = 0.;
        for (int k = i + 1; k <= col; ++k) sum += LB_SY(k, i) * p[col + k - 1] / LB_SY(i, i);
        p[i - 1] += sum;
    }
    return 0;
}

/* subalgorithms.c hpsolb on 1-based views */
FSX_HD void lb_hpsolb(int n, double *t, int *iorder, int iheap)
{
    if (iheap == 0) {
        for (int k = 2; k <= n; ++k) {
            double dd = t[(k) - 1];
            int in = iorder[(k) - 1];
            int i = k;
            while (i > 1) {
                int j = i / 2;
                if (dd < t[(j) - 1]) { t[(i) - 1] = t[(j) - 1]; iorder[(i) - 1] = iorder[(j) - 1]; i = j; }
                else break;
            }
            t[(i) - 1] = dd;
            iorder[(i) - 1] = in;
        }
    }
    if (n > 1) {
        int i = 1;
        double out = t[(1) - 1];
        int iou = iorder[(1) - 1];
        double dd = t[(n) - 1];
        int in = iorder[(n) - 1];
        for (;;) {
            int j = i + i;
            if (j <= n - 1) {
                if (t[(j + 1) - 1] < t[(j) - 1]) ++j;
                if (t[(j) - 1] < dd) { t[(i) - 1] = t[(j) - 1]; iorder[(i) - 1] = iorder[(j) - 1]; i = j; continue; }
            }
            break;
        }
        t[(i) - 1] = dd;
        iorder[(i) - 1] = in;
        t[(n) - 1] = out;
        iorder[(n) - 1] = iou;
    }
}

/* subalgorithms.c cauchy (all bounds of type 2 here: nbd = 2) */
FSX_HD int lb_cauchy(fsx_lb *w, const double *x0, const double *l0, const double *u0,
                     const double *g0, double theta, int col, int head, double sbgnrm,
                     int *nseg)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    const double *x = x0, *l = l0, *u = u0, *g = g0;
    int *iorder = w->indx2, *iwhere = w->iwhere;
    double *t = w->t, *d = w->d, *xcp = w->z;
    double *p = w->wa, *c = w->wa + 2 * m, *wbp = w->wa + 4 * m, *v = w->wa + 6 * m;
    const double epsmch = DBL_EPSILON;
    if (sbgnrm <= 0.) {
        for (int i = 0; i < n; i++) w->z[i] = x0[i];
        return 0;
    }
    int bnded = 1, nfree = n + 1, nbreak = 0, ibkmin = 0;
    double bkmin = 0., f1 = 0., tl = 0., tu = 0.;
    int col2 = col << 1;
    for (int i = 0; i < col2; i++) p[i] = 0.;
    for (int i = 1; i <= n; ++i) {
        double neggi = -g[(i) - 1];
        if (iwhere[(i) - 1] != 3 && iwhere[(i) - 1] != -1) {
            tl = x[(i) - 1] - l[(i) - 1];
            tu = u[(i) - 1] - x[(i) - 1];
            int xlower = tl <= 0.;
            int xupper = tu <= 0.;
            iwhere[(i) - 1] = 0;
            if (xlower) {
                if (neggi <= 0.) iwhere[(i) - 1] = 1;
            } else if (xupper) {
                if (neggi >= 0.) iwhere[(i) - 1] = 2;
            } else {
                if (lb_abs(neggi) <= 0.) iwhere[(i) - 1] = -3;
            }
        }
        int pointr = head;
        if (iwhere[(i) - 1] != 0 && iwhere[(i) - 1] != -1) {
            d[(i) - 1] = 0.;
        } else {
            d[(i) - 1] = neggi;
            f1 -= neggi * neggi;
            for (int j = 1; j <= col; ++j) {
                p[j - 1] += LB_WY(i, pointr) * neggi;
                p[col + j - 1] += LB_WS(i, pointr) * neggi;
                pointr = pointr % m + 1;
            }
            if (neggi < 0.) {
                ++nbreak;
                iorder[(nbreak) - 1] = i;
                t[(nbreak) - 1] = tl / (-neggi);
                if (nbreak == 1 || t[(nbreak) - 1] < bkmin) { bkmin = t[(nbreak) - 1]; ibkmin = nbreak; }
            } else if (neggi > 0.) {
                ++nbreak;
                iorder[(nbreak) - 1] = i;
                t[(nbreak) - 1] = tu / neggi;
                if (nbreak == 1 || t[(nbreak) - 1] < bkmin) { bkmin = t[(nbreak) - 1]; ibkmin = nbreak; }
            } else {
                --nfree;
                iorder[(nfree) - 1] = i;
                if (lb_abs(neggi) > 0.) bnded = 0;
            }
        }
    }
    if (theta != 1.)
        for (int i = 0; i < col; i++) p[col + i] = theta * p[col + i];
    for (int i = 0; i < n; i++) w->z[i] = x0[i];
    if (nbreak == 0 && nfree == n + 1) return 0;
    for (int j = 0; j < col2; j++) c[j] = 0.;
    double f2 = -theta * f1;
    double f2_org = f2;
    if (col > 0) {
        int info = lb_bmv(w, col, p, v);
        if (info) return info;
        f2 -= lb_dot(col2, v, p);
    }
    double dtm = -f1 / f2;
    double tsum = 0.;
    *nseg = 1;
    int reached_end = 0;
    if (nbreak != 0) {
        int nleft = nbreak, iter = 1, ibp;
        double tj = 0.;
        for (;;) {
            double tj0 = tj;
            if (iter == 1) {
                tj = bkmin;
                ibp = iorder[(ibkmin) - 1];
            } else {
                if (iter == 2 && ibkmin != nbreak) {
                    t[(ibkmin) - 1] = t[(nbreak) - 1];
                    iorder[(ibkmin) - 1] = iorder[(nbreak) - 1];
                }
                lb_hpsolb(nleft, t, iorder, iter - 2);
                tj = t[(nleft) - 1];
                ibp = iorder[(nleft) - 1];
            }
            double dt = tj - tj0;
            if (dtm < dt) break;
            tsum += dt;
            --nleft;
            ++iter;
            double dibp = d[(ibp) - 1];
            d[(ibp) - 1] = 0.;
            double zibp;
            if (dibp > 0.) {
                zibp = u[(ibp) - 1] - x[(ibp) - 1];
                xcp[(ibp) - 1] = u[(ibp) - 1];
                iwhere[(ibp) - 1] = 2;
            } else {
                zibp = l[(ibp) - 1] - x[(ibp) - 1];
                xcp[(ibp) - 1] = l[(ibp) - 1];
                iwhere[(ibp) - 1] = 1;
            }
            if (nleft == 0 && nbreak == n) {
                dtm = dt;
                reached_end = 1;
                break;
            }
            ++(*nseg);
            double dibp2 = dibp * dibp;
            f1 = f1 + dt * f2 + dibp2 - theta * dibp * zibp;
            f2 -= theta * dibp2;
            if (col > 0) {
                lb_axpy(col2, dt, p, c);
                int pointr = head;
                for (int j = 1; j <= col; ++j) {
                    wbp[j - 1] = LB_WY(ibp, pointr);
                    wbp[col + j - 1] = theta * LB_WS(ibp, pointr);
                    pointr = pointr % m + 1;
                }
                int info = lb_bmv(w, col, wbp, v);
                if (info) return info;
                double wmc = lb_dot(col2, c, v);
                double wmp = lb_dot(col2, p, v);
                double wmw = lb_dot(col2, wbp, v);
                lb_axpy(col2, -dibp, wbp, p);
                f1 += dibp * wmc;
                f2 = f2 + dibp * 2. * wmp - dibp2 * wmw;
            }
            f2 = lb_max(epsmch * f2_org, f2);
            if (nleft > 0) {
                dtm = -f1 / f2;
                continue;
            } else if (bnded) {
                f1 = 0.; f2 = 0.; dtm = 0.;
            } else {
                dtm = -f1 / f2;
            }
            break;
        }
    }
    if (!reached_end) {
        if (dtm <= 0.) dtm = 0.;
        tsum += dtm;
        lb_axpy(n, tsum, w->d, w->z);
    }
    if (col > 0) lb_axpy(col2, dtm, p, c);
    return 0;
}

/* subalgorithms.c freev */
FSX_HD void lb_freev(fsx_lb *w, int *nfree, int *nenter, int *ileave, int *wrk,
                     int updatd, int iter)
{
    const int n = FSX_LB_N;
    int *index = w->index, *indx2 = w->indx2, *iwhere = w->iwhere;
    *nenter = 0;
    *ileave = n + 1;
    if (iter > 0) {
        for (int i = 1; i <= *nfree; ++i) {
            int k = index[(i) - 1];
            if (iwhere[(k) - 1] > 0) { --(*ileave); indx2[(*ileave) - 1] = k; }
        }
        for (int i = *nfree + 1; i <= n; ++i) {
            int k = index[(i) - 1];
            if (iwhere[(k) - 1] <= 0) { ++(*nenter); indx2[(*nenter) - 1] = k; }
        }
    }
    *wrk = *ileave < n + 1 || *nenter > 0 || updatd;
    *nfree = 0;
    int iact = n + 1;
    for (int i = 1; i <= n; ++i) {
        if (iwhere[(i) - 1] <= 0) { ++(*nfree); index[(*nfree) - 1] = i; }
        else { --iact; index[(iact) - 1] = i; }
    }
}

/* subalgorithms.c formk */
FSX_HD int lb_formk(fsx_lb *w, int nsub, int nenter, int ileave, int iupdat,
                    int updatd, double theta, int col, int head)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    const int *ind = w->index, *indx2 = w->indx2;
    int upcl;
    if (updatd) {
        if (iupdat > m) {
            for (int jy = 1; jy <= m - 1; ++jy) {
                int js = m + jy;
                for (int k = 0; k < m - jy; k++) LB_WN1(jy + k, jy) = LB_WN1(jy + 1 + k, jy + 1);
                for (int k = 0; k < m - jy; k++) LB_WN1(js + k, js) = LB_WN1(js + 1 + k, js + 1);
                for (int k = 0; k < m - 1; k++) LB_WN1(m + 1 + k, jy) = LB_WN1(m + 2 + k, jy + 1);
            }
        }
        int iy = col, is = m + col;
        int ipntr = head + col - 1;
        if (ipntr > m) ipntr -= m;
        int jpntr = head;
        for (int jy = 1; jy <= col; ++jy) {
            int js = m + jy;
            double t1 = 0., t2 = 0., t3 = 0.;
            for (int k = 1; k <= nsub; ++k) {
                int k1 = ind[(k) - 1];
                t1 += LB_WY(k1, ipntr) * LB_WY(k1, jpntr);
            }
            for (int k = nsub + 1; k <= n; ++k) {
                int k1 = ind[(k) - 1];
                t2 += LB_WS(k1, ipntr) * LB_WS(k1, jpntr);
                t3 += LB_WS(k1, ipntr) * LB_WY(k1, jpntr);
            }
            LB_WN1(iy, jy) = t1;
            LB_WN1(is, js) = t2;
            LB_WN1(is, jy) = t3;
            jpntr = jpntr % m + 1;
        }
        int jy = col;
        jpntr = head + col - 1;
        if (jpntr > m) jpntr -= m;
        ipntr = head;
        for (int i = 1; i <= col; ++i) {
            int is2 = m + i;
            double t3 = 0.;
            for (int k = 1; k <= nsub; ++k) {
                int k1 = ind[(k) - 1];
                t3 += LB_WS(k1, ipntr) * LB_WY(k1, jpntr);
            }
            ipntr = ipntr % m + 1;
            LB_WN1(is2, jy) = t3;
        }
        upcl = col - 1;
    } else {
        upcl = col;
    }
    int ipntr = head;
    for (int iy = 1; iy <= upcl; ++iy) {
        int is = m + iy;
        int jpntr = head;
        for (int jy = 1; jy <= iy; ++jy) {
            int js = m + jy;
            double t1 = 0., t2 = 0., t3 = 0., t4 = 0.;
            for (int k = 1; k <= nenter; ++k) {
                int k1 = indx2[(k) - 1];
                t1 += LB_WY(k1, ipntr) * LB_WY(k1, jpntr);
                t2 += LB_WS(k1, ipntr) * LB_WS(k1, jpntr);
            }
            for (int k = ileave; k <= n; ++k) {
                int k1 = indx2[(k) - 1];
                t3 += LB_WY(k1, ipntr) * LB_WY(k1, jpntr);
                t4 += LB_WS(k1, ipntr) * LB_WS(k1, jpntr);
            }
            LB_WN1(iy, jy) = LB_WN1(iy, jy) + t1 - t3;
            LB_WN1(is, js) = LB_WN1(is, js) - t2 + t4;
            jpntr = jpntr % m + 1;
        }
        ipntr = ipntr % m + 1;
    }
    ipntr = head;
    for (int is = m + 1; is <= m + upcl; ++is) {
        int jpntr = head;
        for (int jy = 1; jy <= upcl; ++jy) {
            double t1 = 0., t3 = 0.;
            for (int k = 1; k <= nenter; ++k) {
                int k1 = indx2[(k) - 1];
                t1 += LB_WS(k1, ipntr) * LB_WY(k1, jpntr);
            }
            for (int k = ileave; k <= n; ++k) {
                int k1 = indx2[(k) - 1];
                t3 += LB_WS(k1, ipntr) * LB_WY(k1, jpntr);
            }
            if (is <= jy + m) LB_WN1(is, jy) = LB_WN1(is, jy) + t1 - t3;
            else LB_WN1(is, jy) = LB_WN1(is, jy) - t1 + t3;
            jpntr = jpntr % m + 1;
        }
        ipntr = ipntr % m + 1;
    }
    const int m2 = m << 1;
    for (int iy = 1; iy <= col; ++iy) {
        int is = col + iy, is1 = m + iy;
        for (int jy = 1; jy <= iy; ++jy) {
            int js = col + jy, js1 = m + jy;
            LB_WN(jy, iy) = LB_WN1(iy, jy) / theta;
            LB_WN(js, is) = LB_WN1(is1, js1) * theta;
        }
        for (int jy = 1; jy <= iy - 1; ++jy) LB_WN(jy, is) = -LB_WN1(is1, jy);
        for (int jy = iy; jy <= col; ++jy) LB_WN(jy, is) = LB_WN1(is1, jy);
        LB_WN(iy, iy) += LB_SY(iy, iy);
    }
    if (lb_chol(w->wn, m2, col)) return -1;
    int col2 = col << 1;
    for (int js = col + 1; js <= col2; ++js) lb_trsl(w->wn, m2, col, &LB_WN(1, js), 11);
    for (int is = col + 1; is <= col2; ++is)
        for (int js = is; js <= col2; ++js)
            LB_WN(is, js) += lb_dot(col, &LB_WN(1, is), &LB_WN(1, js));
    if (lb_chol(&LB_WN(col + 1, col + 1), m2, col)) return -2;
    return 0;
}

/* subalgorithms.c formt */
FSX_HD int lb_formt(fsx_lb *w, int col, double theta)
{
    for (int j = 1; j <= col; ++j) LB_WT(1, j) = theta * LB_SS(1, j);
    for (int i = 2; i <= col; ++i)
        for (int j = i; j <= col; ++j) {
            int k1 = (i <= j ? i : j) - 1;
            double dd = 0.;
            for (int k = 1; k <= k1; ++k) dd += LB_SY(i, k) * LB_SY(j, k) / LB_SY(k, k);
            LB_WT(i, j) = dd + theta * LB_SS(i, j);
        }
    return lb_chol(w->wt, FSX_LB_M, col) ? -3 : 0;
}

/* subalgorithms.c cmprlb (constrained problem) */
FSX_HD int lb_cmprlb(fsx_lb *w, const double *x0, const double *g0, double theta,
                     int col, int head, int nfree)
{
    const int m = FSX_LB_M;
    const double *x = x0, *g = g0, *z = w->z;
    double *r = w->r;
    const int *index = w->index;
    for (int i = 1; i <= nfree; ++i) {
        int k = index[(i) - 1];
        r[(i) - 1] = -theta * (z[(k) - 1] - x[(k) - 1]) - g[(k) - 1];
    }
    if (lb_bmv(w, col, w->wa + 2 * m, w->wa)) return -8;
    int pointr = head;
    for (int j = 1; j <= col; ++j) {
        double a1 = w->wa[j - 1], a2 = theta * w->wa[col + j - 1];
        for (int i = 1; i <= nfree; ++i) {
            int k = index[(i) - 1];
            r[(i) - 1] = r[(i) - 1] + LB_WY(k, pointr) * a1 + LB_WS(k, pointr) * a2;
        }
        pointr = pointr % m + 1;
    }
    return 0;
}

/* subalgorithms.c subsm (3.0, with the projected backtracking safeguard) */
FSX_HD int lb_subsm(fsx_lb *w, int nsub, const double *l0, const double *u0,
                    const double *xx0, const double *gg0, double theta, int col,
                    int head)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    const double *l = l0, *u = u0, *xx = xx0, *gg = gg0;
    const int *ind = w->index;
    double *x = w->z, *d = w->r, *wv = w->wa;
    if (nsub <= 0) return 0;
    int pointr = head;
    for (int i = 1; i <= col; ++i) {
        double t1 = 0., t2 = 0.;
        for (int j = 1; j <= nsub; ++j) {
            int k = ind[(j) - 1];
            t1 += LB_WY(k, pointr) * d[(j) - 1];
            t2 += LB_WS(k, pointr) * d[(j) - 1];
        }
        wv[i - 1] = t1;
        wv[col + i - 1] = theta * t2;
        pointr = pointr % m + 1;
    }
    int info = lb_trsl(w->wn, 2 * m, 2 * col, wv, 11);
    if (info) return info;
    for (int i = 0; i < col; ++i) wv[i] = -wv[i];
    info = lb_trsl(w->wn, 2 * m, 2 * col, wv, 1);
    if (info) return info;
    pointr = head;
    for (int jy = 1; jy <= col; ++jy) {
        int js = col + jy;
        for (int i = 1; i <= nsub; ++i) {
            int k = ind[(i) - 1];
            d[(i) - 1] = d[(i) - 1] + LB_WY(k, pointr) * wv[jy - 1] / theta + LB_WS(k, pointr) * wv[js - 1];
        }
        pointr = pointr % m + 1;
    }
    double rt = 1. / theta;
    for (int i = 0; i < nsub; i++) w->r[i] = rt * w->r[i];
    int iword = 0;
    for (int i = 0; i < n; i++) w->xp[i] = w->z[i];
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[(i) - 1];
        double dk = d[(i) - 1], xk = x[(k) - 1];
        xk = lb_max(l[(k) - 1], xk + dk);
        x[(k) - 1] = lb_min(u[(k) - 1], xk);
        if (x[(k) - 1] == l[(k) - 1] || x[(k) - 1] == u[(k) - 1]) iword = 1;
    }
    if (iword == 0) return 0;
    double dd_p = 0.;
    for (int i = 1; i <= n; ++i) dd_p += (x[(i) - 1] - xx[(i) - 1]) * gg[(i) - 1];
    if (dd_p > 0.) {
        for (int i = 0; i < n; i++) w->z[i] = w->xp[i];
    } else {
        return 0;
    }
    double alpha = 1., t1 = alpha;
    int ibd = 0;
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[(i) - 1];
        double dk = d[(i) - 1];
        if (dk < 0.) {
            double t2 = l[(k) - 1] - x[(k) - 1];
            if (t2 >= 0.) t1 = 0.;
            else if (dk * alpha < t2) t1 = t2 / dk;
        } else if (dk > 0.) {
            double t2 = u[(k) - 1] - x[(k) - 1];
            if (t2 <= 0.) t1 = 0.;
            else if (dk * alpha > t2) t1 = t2 / dk;
        }
        if (t1 < alpha) { alpha = t1; ibd = i; }
    }
    if (alpha < 1.) {
        double dk = d[(ibd) - 1];
        int k = ind[(ibd) - 1];
        if (dk > 0.) { x[(k) - 1] = u[(k) - 1]; d[(ibd) - 1] = 0.; }
        else if (dk < 0.) { x[(k) - 1] = l[(k) - 1]; d[(ibd) - 1] = 0.; }
    }
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[(i) - 1];
        x[(k) - 1] += alpha * d[(i) - 1];
    }
    return 0;
}

/* subalgorithms.c matupd */
FSX_HD void lb_matupd(fsx_lb *w, int *itail, int iupdat, int *col, int *head,
                      double *theta, double rr, double dr, double stp, double dtd)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    if (iupdat <= m) {
        *col = iupdat;
        *itail = (*head + iupdat - 2) % m + 1;
    } else {
        *itail = *itail % m + 1;
        *head = *head % m + 1;
    }
    for (int i = 0; i < n; i++) {
        LB_WS(i + 1, *itail) = w->d[i];
        LB_WY(i + 1, *itail) = w->r[i];
    }
    *theta = rr / dr;
    if (iupdat > m) {
        for (int j = 1; j <= *col - 1; ++j) {
            for (int k = 0; k < j; k++) LB_SS(1 + k, j) = LB_SS(2 + k, j + 1);
            for (int k = 0; k < *col - j; k++) LB_SY(j + k, j) = LB_SY(j + 1 + k, j + 1);
        }
    }
    int pointr = *head;
    for (int j = 1; j <= *col - 1; ++j) {
        LB_SY(*col, j) = lb_dot(n, w->d, &LB_WY(1, pointr));
        LB_SS(j, *col) = lb_dot(n, &LB_WS(1, pointr), w->d);
        pointr = pointr % m + 1;
    }
    if (stp == 1.) LB_SS(*col, *col) = dtd;
    else LB_SS(*col, *col) = stp * stp * dtd;
    LB_SY(*col, *col) = dr;
}

/* linesearch.c dcstep */
FSX_HD void lb_dcstep(double *stx, double *fx, double *dx, double *sty, double *fy,
                      double *dy, double *stp, double fp, double dp, int *brackt,
                      double stpmin, double stpmax)
{
    double sgnd = dp * (*dx / lb_abs(*dx));
    double stpf, stpc, stpq, theta, s, gamma, p, q, r, tt;
    if (fp > *fx) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(*dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(tt * tt - *dx / s * (dp / s));
        if (*stp < *stx) gamma = -gamma;
        p = gamma - *dx + theta;
        q = gamma - *dx + gamma + dp;
        r = p / q;
        stpc = *stx + r * (*stp - *stx);
        stpq = *stx + *dx / ((*fx - fp) / (*stp - *stx) + *dx) / 2. * (*stp - *stx);
        if (lb_abs(stpc - *stx) < lb_abs(stpq - *stx)) stpf = stpc;
        else stpf = stpc + (stpq - stpc) / 2.;
        *brackt = 1;
    } else if (sgnd < 0.) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(*dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(tt * tt - *dx / s * (dp / s));
        if (*stp > *stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma - dp + gamma + *dx;
        r = p / q;
        stpc = *stp + r * (*stx - *stp);
        stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
        if (lb_abs(stpc - *stp) > lb_abs(stpq - *stp)) stpf = stpc;
        else stpf = stpq;
        *brackt = 1;
    } else if (lb_abs(dp) < lb_abs(*dx)) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(*dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(lb_max(0., tt * tt - *dx / s * (dp / s)));
        if (*stp > *stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma + (*dx - dp) + gamma;
        r = p / q;
        if (r < 0. && gamma != 0.) stpc = *stp + r * (*stx - *stp);
        else if (*stp > *stx) stpc = stpmax;
        else stpc = stpmin;
        stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
        if (*brackt) {
            if (lb_abs(stpc - *stp) < lb_abs(stpq - *stp)) stpf = stpc;
            else stpf = stpq;
            if (*stp > *stx) stpf = lb_min(*stp + (*sty - *stp) * .66, stpf);
            else stpf = lb_max(*stp + (*sty - *stp) * .66, stpf);
        } else {
            if (lb_abs(stpc - *stp) > lb_abs(stpq - *stp)) stpf = stpc;
            else stpf = stpq;
            stpf = lb_min(stpmax, stpf);
            stpf = lb_max(stpmin, stpf);
        }
    } else {
        if (*brackt) {
            theta = (fp - *fy) * 3. / (*sty - *stp) + *dy + dp;
            s = lb_max(lb_max(lb_abs(theta), lb_abs(*dy)), lb_abs(dp));
            tt = theta / s;
            gamma = s * sqrt(tt * tt - *dy / s * (dp / s));
            if (*stp > *sty) gamma = -gamma;
            p = gamma - dp + theta;
            q = gamma - dp + gamma + *dy;
            r = p / q;
            stpc = *stp + r * (*sty - *stp);
            stpf = stpc;
        } else if (*stp > *stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > *fx) {
        *sty = *stp; *fy = fp; *dy = dp;
    } else {
        if (sgnd < 0.) { *sty = *stx; *fy = *fx; *dy = *dx; }
        *stx = *stp; *fx = fp; *dx = dp;
    }
    *stp = stpf;
}

typedef struct fsx_ls {
    int brackt, stage;
    double ginit, gtest, gx, gy, finit, fx, fy, stx, sty, stmin, stmax, width, width1;
} fsx_ls;

/* linesearch.c dcsrch with ftol 1e-3, gtol 0.9, xtol 0.1, stpmin 0 */
FSX_HD void lb_dcsrch(double f, double g, double *stp, double stpmin, double stpmax,
                      int *task, fsx_ls *s)
{
    const double ftol = .001, gtol = .9, xtol = .1;
    if (*task == FSX_T_START) {
        if (*stp < stpmin) *task = FSX_T_E_SMALLSTP;
        if (*stp > stpmax) *task = FSX_T_E_LARGESTP;
        if (g >= 0.) *task = FSX_T_E_INITIAL;
        if (stpmin < 0.) *task = FSX_T_E_STP0;
        if (stpmax < stpmin) *task = FSX_T_E_STP1;
        if (*task >= 200 && *task <= 240) return;
        s->brackt = 0;
        s->stage = 1;
        s->finit = f;
        s->ginit = g;
        s->gtest = ftol * s->ginit;
        s->width = stpmax - stpmin;
        s->width1 = s->width / .5;
        s->stx = 0.; s->fx = s->finit; s->gx = s->ginit;
        s->sty = 0.; s->fy = s->finit; s->gy = s->ginit;
        s->stmin = 0.;
        s->stmax = *stp + *stp * 4.;
        *task = FSX_T_FG;
        return;
    }
    double ftest = s->finit + *stp * s->gtest;
    if (s->stage == 1 && f <= ftest && g >= 0.) s->stage = 2;
    if (s->brackt && (*stp <= s->stmin || *stp >= s->stmax)) *task = FSX_T_W_ROUND;
    if (s->brackt && s->stmax - s->stmin <= xtol * s->stmax) *task = FSX_T_W_XTOL;
    if (*stp == stpmax && f <= ftest && g <= s->gtest) *task = FSX_T_W_STPMAX;
    if (*stp == stpmin && (f > ftest || g >= s->gtest)) *task = FSX_T_W_STPMIN;
    if (f <= ftest && lb_abs(g) <= gtol * (-s->ginit)) *task = FSX_T_CONVERGENCE;
    if ((*task >= 100 && *task <= 110) || (*task >= 20 && *task <= 25)) return;
    if (s->stage == 1 && f <= s->fx && f > ftest) {
        double fm = f - *stp * s->gtest;
        double fxm = s->fx - s->stx * s->gtest;
        double fym = s->fy - s->sty * s->gtest;
        double gm = g - s->gtest;
        double gxm = s->gx - s->gtest;
        double gym = s->gy - s->gtest;
        lb_dcstep(&s->stx, &fxm, &gxm, &s->sty, &fym, &gym, stp, fm, gm, &s->brackt,
                  s->stmin, s->stmax);
        s->fx = fxm + s->stx * s->gtest;
        s->fy = fym + s->sty * s->gtest;
        s->gx = gxm + s->gtest;
        s->gy = gym + s->gtest;
    } else {
        lb_dcstep(&s->stx, &s->fx, &s->gx, &s->sty, &s->fy, &s->gy, stp, f, g, &s->brackt,
                  s->stmin, s->stmax);
    }
    if (s->brackt) {
        if (lb_abs(s->sty - s->stx) >= s->width1 * .66) *stp = s->stx + (s->sty - s->stx) * .5;
        s->width1 = s->width;
        s->width = lb_abs(s->sty - s->stx);
    }
    if (s->brackt) {
        s->stmin = lb_min(s->stx, s->sty);
        s->stmax = lb_max(s->stx, s->sty);
    } else {
        s->stmin = *stp + (*stp - s->stx) * 1.1;
        s->stmax = *stp + (*stp - s->stx) * 4.;
    }
    *stp = lb_max(*stp, stpmin);
    *stp = lb_min(*stp, stpmax);
    if ((s->brackt && (*stp <= s->stmin || *stp >= s->stmax))
        || (s->brackt && s->stmax - s->stmin <= xtol * s->stmax))
        *stp = s->stx;
    *task = FSX_T_FG;
}

FSX_HD void lb_zero(fsx_lb *w)
{
    for (int i = 0; i < FSX_LB_N * FSX_LB_M; i++) { w->ws[i] = 0.0; w->wy[i] = 0.0; }
    for (int i = 0; i < FSX_LB_M * FSX_LB_M; i++) { w->sy[i] = 0.0; w->ss[i] = 0.0; w->wt[i] = 0.0; }
    for (int i = 0; i < 4 * FSX_LB_M * FSX_LB_M; i++) { w->wn[i] = 0.0; w->snd[i] = 0.0; }
    for (int i = 0; i < FSX_LB_N; i++) {
        w->z[i] = w->r[i] = w->d[i] = w->t[i] = w->xp[i] = 0.0;
        w->index[i] = w->iwhere[i] = w->indx2[i] = 0;
    }
    for (int i = 0; i < 8 * FSX_LB_M; i++) w->wa[i] = 0.0;
}

/* The objective evaluation requested by the wrapper on every FG task
 * (min_saxs.c:238-241): refresh the scale, then chi^2 and gradient. */
FSX_HD void lb_eval(const fsx_ct *t, const fsx_fitpar *p, const double *x, double *f,
                    double *g, double *kscale)
{
    double k = fsx_best_scale(t, p, x[0], x[1]);
    *kscale = k;
    fsx_objective(t, p, k, x, g, f);
}

/* ---------------------------------------------------------------------- */
/* Resumable driver.  setulb's reverse communication (lbfgsb.c:52-515, driven
 * by min_saxs.c:234-248) is kept as an explicit state machine:
 * fsx_lb_start() sets the first trial point s->x; the caller evaluates f
 * and g there into s->f / s->g and calls fsx_lb_resume(), which runs the
 * optimiser up to the next point it needs evaluated (returns 1, point in
 * s->x) or to the end of the fit (returns 0, reason in s->task).  Since
 * every NEW_X is resumed at once, the visited points are those of the
 * reference.  Many fits can share one evaluation loop this way, and the
 * slot of a finished fit can start the next one. */
#define FSX_PH_INIT 0
#define FSX_PH_LS 1
#define FSX_PH_DONE 2

typedef struct fsx_lbrc {
    fsx_lb w;
    fsx_ls ls;
    double x[FSX_LB_N], g[FSX_LB_N], f;
    double l[FSX_LB_N], u[FSX_LB_N];
    double theta, fold, gd, stpmx, sbgnrm, stp, gdold, dtd, tol, pgtol;
    int col, head, iupdat, updatd, iback, itail, iter, nseg, nfree, ifun, nenter, ileave;
    int wrk, csave, nev, task, phase;
} fsx_lbrc;

/* x0 = (c1, c2) projected into the box [lo, hi] (active, lbfgsb.c:540-600,
 * every variable bounded on both sides); the first evaluation is due at
 * s->x. */
FSX_HD void fsx_lb_start(fsx_lbrc *s, const fsx_fitpar *p, double x0, double x1)
{
    lb_zero(&s->w);
    s->ls.brackt = 0; s->ls.stage = 0;
    s->ls.ginit = s->ls.gtest = s->ls.gx = s->ls.gy = s->ls.finit = s->ls.fx = s->ls.fy = 0.0;
    s->ls.stx = s->ls.sty = s->ls.stmin = s->ls.stmax = s->ls.width = s->ls.width1 = 0.0;
    s->l[0] = p->lo0; s->l[1] = p->lo1;
    s->u[0] = p->hi0; s->u[1] = p->hi1;
    s->x[0] = x0; s->x[1] = x1;
    s->g[0] = 0.0; s->g[1] = 0.0; s->f = 0.0;
    s->col = 0; s->head = 1; s->iupdat = 0; s->updatd = 0; s->iback = 0; s->itail = 0;
    s->iter = 0; s->nseg = 0; s->nfree = FSX_LB_N; s->ifun = 0; s->nenter = 0; s->ileave = 0;
    s->theta = 1.; s->fold = 0.; s->gd = 0.; s->stpmx = 0.; s->sbgnrm = 0.; s->stp = 0.;
    s->gdold = 0.; s->dtd = 0.;
    s->tol = p->factr * DBL_EPSILON;
    s->pgtol = p->pgtol;
    s->wrk = 0; s->csave = 0; s->nev = 0; s->task = FSX_T_FG; s->phase = FSX_PH_INIT;
    for (int i = 0; i < FSX_LB_N; i++) {
        if (s->x[i] <= s->l[i]) {
            if (s->x[i] < s->l[i]) s->x[i] = s->l[i];
        } else if (s->x[i] >= s->u[i]) {
            if (s->x[i] > s->u[i]) s->x[i] = s->u[i];
        }
        s->w.iwhere[i] = (s->u[i] - s->l[i] <= 0.) ? 3 : 0;
    }
}

/* mainlb from the point after an evaluation to the next one (or the end). */
FSX_HD int fsx_lb_resume(fsx_lbrc *s)
{
    fsx_lb *w = &s->w;
    const int n = FSX_LB_N;
    double *x = s->x, *g = s->g;
    const double *l = s->l, *u = s->u;
    const double epsmch = DBL_EPSILON;
    int info = 0, restart, skip_search;
    double rr, dr, ddum;
    if (s->phase == FSX_PH_DONE) return 0;
    s->nev++;
    if (s->phase == FSX_PH_LS) goto ls_resume;
#define LB_PROJGR(out)                                                          \
    do {                                                                        \
        double s_ = 0.;                                                         \
        for (int i_ = 0; i_ < n; i_++) {                                        \
            double gi_ = g[i_];                                                 \
            if (gi_ < 0.) gi_ = lb_max(x[i_] - u[i_], gi_);                     \
            else gi_ = lb_min(x[i_] - l[i_], gi_);                              \
            s_ = lb_max(s_, lb_abs(gi_));                                       \
        }                                                                       \
        (out) = s_;                                                             \
    } while (0)
#define LB_RESET()                                                              \
    do {                                                                        \
        info = 0; s->col = 0; s->head = 1; s->theta = 1.;                       \
        s->iupdat = 0; s->updatd = 0;                                           \
    } while (0)
    /* the first evaluation (at the projected x0) is done */
    LB_PROJGR(s->sbgnrm);
    if (s->sbgnrm <= s->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
    for (;;) {
        info = lb_cauchy(w, x, l, u, g, s->theta, s->col, s->head, s->sbgnrm, &s->nseg);
        if (info) {
            LB_RESET();
            continue;
        }
        lb_freev(w, &s->nfree, &s->nenter, &s->ileave, &s->wrk, s->updatd, s->iter);
        if (!(s->nfree == 0 || s->col == 0)) {
            if (s->wrk)
                info = lb_formk(w, s->nfree, s->nenter, s->ileave, s->iupdat, s->updatd,
                                s->theta, s->col, s->head);
            if (info) {
                LB_RESET();
                continue;
            }
            info = lb_cmprlb(w, x, g, s->theta, s->col, s->head, s->nfree);
            if (!info) info = lb_subsm(w, s->nfree, l, u, x, g, s->theta, s->col, s->head);
            if (info) {
                LB_RESET();
                continue;
            }
        }
        for (int i = 0; i < n; i++) w->d[i] = w->z[i] - x[i];
        /* lnsrlb entry (every variable boxed) */
        s->dtd = lb_dot(n, w->d, w->d);
        if (s->iter == 0) {
            s->stpmx = 1.;
        } else {
            s->stpmx = 1e10;
            for (int i = 0; i < n; i++) {
                double a1 = w->d[i];
                if (a1 < 0.) {
                    double a2 = l[i] - x[i];
                    if (a2 >= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx < a2) s->stpmx = a2 / a1;
                } else if (a1 > 0.) {
                    double a2 = u[i] - x[i];
                    if (a2 <= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx > a2) s->stpmx = a2 / a1;
                }
            }
        }
        s->stp = 1.;
        for (int i = 0; i < n; i++) { w->t[i] = x[i]; w->r[i] = g[i]; }
        s->fold = s->f;
        s->ifun = 0;
        s->iback = 0;
        s->csave = FSX_T_START;
        restart = 0;
        for (;;) {
            s->gd = lb_dot(n, g, w->d);
            skip_search = 0;
            if (s->ifun == 0) {
                s->gdold = s->gd;
                if (s->gd >= 0.) { info = -4; skip_search = 1; }
            }
            if (!skip_search) {
                lb_dcsrch(s->f, s->gd, &s->stp, 0., s->stpmx, &s->csave, &s->ls);
                if (!(s->csave >= 100 && s->csave <= 110) && !(s->csave >= 20 && s->csave <= 25)) {
                    s->task = FSX_T_FG_LN;
                    ++s->ifun;
                    s->iback = s->ifun - 1;
                    if (s->stp == 1.) {
                        for (int i = 0; i < n; i++) x[i] = w->z[i];
                    } else {
                        for (int i = 0; i < n; i++) x[i] = s->stp * w->d[i] + w->t[i];
                    }
                } else {
                    s->task = FSX_T_NEW_X;
                }
            }
            if (info != 0 || s->iback >= 20) {
                for (int i = 0; i < n; i++) { x[i] = w->t[i]; g[i] = w->r[i]; }
                s->f = s->fold;
                if (s->col == 0) {
                    s->task = FSX_T_ABNORMAL;
                    ++s->iter;
                    goto done;
                }
                LB_RESET();
                restart = 1;
                break;
            } else if (s->task == FSX_T_FG_LN) {
                if (s->nev >= FSX_LB_MAX_EVALS) { s->task = FSX_T_STOP_ITER; goto done; }
                s->phase = FSX_PH_LS;
                return 1;
            ls_resume:
                continue;
            } else {
                ++s->iter;
                LB_PROJGR(s->sbgnrm);
                break;
            }
        }
        if (restart) continue;
        if (s->sbgnrm <= s->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
        ddum = lb_max(lb_max(lb_abs(s->fold), lb_abs(s->f)), 1.);
        if (s->fold - s->f <= s->tol * ddum) { s->task = FSX_T_CONV_F; goto done; }
        for (int i = 0; i < n; i++) w->r[i] = g[i] - w->r[i];
        rr = lb_dot(n, w->r, w->r);
        if (s->stp == 1.) {
            dr = s->gd - s->gdold;
            ddum = -s->gdold;
        } else {
            dr = (s->gd - s->gdold) * s->stp;
            for (int i = 0; i < n; i++) w->d[i] = s->stp * w->d[i];
            ddum = -s->gdold * s->stp;
        }
        if (dr <= epsmch * ddum) {
            s->updatd = 0;
            continue;
        }
        s->updatd = 1;
        ++s->iupdat;
        lb_matupd(w, &s->itail, s->iupdat, &s->col, &s->head, &s->theta, rr, dr, s->stp,
                  s->dtd);
        if (lb_formt(w, s->col, s->theta)) LB_RESET();
    }
#undef LB_PROJGR
#undef LB_RESET
done:
    s->phase = FSX_PH_DONE;
    return 0;
}

/* One complete fit from x = (c1, c2).  Returns the final task; x is the
 * result, *fout = chi^2, *nevals = objective evaluations. */
FSX_HD int fsx_lbfgsb_fit(const fsx_ct *ct, const fsx_fitpar *p, double *x, double *fout,
                          int *nevals)
{
    fsx_lbrc S;
    double k;
    fsx_lb_start(&S, p, x[0], x[1]);
    do {
        lb_eval(ct, p, S.x, &S.f, S.g, &k);
    } while (fsx_lb_resume(&S));
    x[0] = S.x[0];
    x[1] = S.x[1];
    *fout = S.f;
    *nevals = S.nev;
    return S.task;
}

#endif
