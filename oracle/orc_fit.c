/*
 * orc_fit.c -- oracle restatement of the reference's c1/c2 chi^2 fit
 * (src/min_saxs.c) and of the profile helpers that feed it
 * (src/profile.c, src/pdb2spf.c).
 *
 * TEST INFRASTRUCTURE (checker only; never linked into the product).
 * Expressions keep the reference's evaluation order (build with
 * -ffp-contract=off) so the L-BFGS-B iterates match bit for bit.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* cross-term slots of a per-pose profile block [6][qnum] */
enum { VV = 0, VD = 1, VW = 2, DD = 3, DW = 4, WW = 5 };

/* min_saxs.c:353-389.  Moments of the experimental curve per model q bin
 * [q_{i-1}, q_i] (inclusive), normalised by the number of consumed points.
 * The scan stops at n_exp (the reference's while loop has no bound and reads
 * past the experimental arrays when they end inside the last bin). */
void orc_scoring_helper(const double *eq, const double *ein, const double *eerr,
                        int n_exp, const double *qvals, int qnum, double *a)
{
    memset(a, 0, sizeof(double) * 6 * (size_t)qnum);
    int j = 0;
    for (int i = 0; i < qnum; i++) {
        double lo = qvals[0], hi = qvals[i];
        if (i > 0) lo = qvals[i - 1];
        while (j < n_exp && eq[j] >= lo && eq[j] <= hi) {
            double e2 = eerr[j] * eerr[j];
            a[i * 6 + 0] += ein[j] * ein[j] / e2;
            a[i * 6 + 1] += ein[j] / e2;
            a[i * 6 + 2] += eq[j] * ein[j] / e2;
            a[i * 6 + 3] += 1.0 / e2;
            a[i * 6 + 4] += eq[j] / e2;
            a[i * 6 + 5] += eq[j] * eq[j] / e2;
            j++;
        }
    }
    double norm = (double)j;
    for (int i = 0; i < 6 * qnum; i++) a[i] /= norm;
}

/* min_saxs.c:116-124 plus the wrapper constants of min_saxs.c:202-229 and
 * define.h:28-34. */
void orc_params_init(orc_fit_params *p, double *a, double rm, double peak)
{
    double pi = orc_get_pi();
    p->a = a;
    p->rm = rm;
    p->mult = pow(4.0 * pi / 3.0, 1.5) * rm * rm / (16.0 * pi);
    p->peak = peak;
    p->c1_lo = 0.96; p->c1_hi = 1.04;
    p->c2_lo = -2.00; p->c2_hi = 4.00;
    p->factr = 1e+7;
    p->pgtol = 1e-5;
    p->m = 3;
}

static inline double model_in(const double *ct, int qnum, int i, double G, double c2)
{
    const double *vv = ct + VV * qnum, *vd = ct + VD * qnum, *vw = ct + VW * qnum;
    const double *dd = ct + DD * qnum, *dw = ct + DW * qnum, *ww = ct + WW * qnum;
    return vv[i] - G * vd[i] + c2 * vw[i] + G * G * dd[i] - G * c2 * dw[i] + c2 * c2 * ww[i];
}

/* min_saxs.c:261-319: closed-form scale k of the model that minimises
 * chi^2 for fixed (c1, c2), using the piecewise-linear interpolant. */
double orc_best_scale(const double *ct, const double *q, int qnum,
                      const orc_fit_params *p, double c1, double c2)
{
    const double *a = p->a;
    double corr = -p->mult * (c1 * c1 - 1.0);
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double in_prev = model_in(ct, qnum, 0, G, c2);
    double c1c = c1 * c1 * c1;
    double q_prev = -1.0, up = 0.0, down = 0.0;
    for (int i = 0; i < qnum; i++) {
        double qc = q[i];
        G = c1c * exp(corr * qc * qc);
        double in = model_in(ct, qnum, i, G, c2);
        double tan = (in - in_prev) / (qc - q_prev);
        double buf = in - tan * qc;
        up += buf * a[i * 6 + 1] + tan * a[i * 6 + 2];
        down += buf * buf * a[i * 6 + 3] + 2.0 * tan * buf * a[i * 6 + 4] + tan * tan * a[i * 6 + 5];
        in_prev = in;
        q_prev = qc;
    }
    return up / down;
}

/* min_saxs.c:3-105: chi^2 and its (c1, c2) gradient at scale k. */
void orc_gradient(const double *ct, const double *q, int qnum,
                  const orc_fit_params *p, double k, const double *x,
                  double *grad, double *fval)
{
    const double *a = p->a;
    const double *VDp = ct + VD * qnum, *VWp = ct + VW * qnum;
    const double *DDp = ct + DD * qnum, *DWp = ct + DW * qnum, *WWp = ct + WW * qnum;
    double c1 = x[0], c2 = x[1];
    double g0 = 0.0, g1 = 0.0;
    double mult = p->mult;
    double corr = -mult * (c1 * c1 - 1.0);
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double G_der = G * (3.0 / c1 - 2.0 * c1 * mult * q[0] * q[0]);
    double in_prev = model_in(ct, qnum, 0, G, c2);
    double dc1_prev = -G_der * VDp[0] + 2.0 * G * G_der * DDp[0] - G_der * c2 * DWp[0];
    double dc2_prev = VWp[0] - G * DWp[0] + 2.0 * c2 * WWp[0];
    double q_prev = -1.0, score = 0.0, c1c = c1 * c1 * c1;
    for (int i = 0; i < qnum; i++) {
        double qc = q[i];
        G = c1c * exp(corr * qc * qc);
        G_der = G * (3.0 / c1 - 2.0 * c1 * mult * qc * qc);
        double in = model_in(ct, qnum, i, G, c2);
        double dc1 = G_der * (-VDp[i] + 2.0 * G * DDp[i] - c2 * DWp[i]);
        double dc2 = VWp[i] - G * DWp[i] + 2.0 * c2 * WWp[i];
        double buf = 1.0 / (qc - q_prev);
        double tan = (in - in_prev) * buf;
        double t1 = (dc1 - dc1_prev) * buf;
        double t2 = (dc2 - dc2_prev) * buf;
        g0 += 2.0 * k * (-(dc1 - t1 * qc) * a[i * 6 + 1] - t1 * a[i * 6 + 2]
                         + k * ((in - tan * qc) * (dc1 - t1 * qc) * a[i * 6 + 3]
                                + (in * t1 + dc1 * tan - 2.0 * tan * t1 * qc) * a[i * 6 + 4]
                                + tan * t1 * a[i * 6 + 5]));
        g1 += 2.0 * k * (-(dc2 - t2 * qc) * a[i * 6 + 1] - t2 * a[i * 6 + 2]
                         + k * ((in - tan * qc) * (dc2 - t2 * qc) * a[i * 6 + 3]
                                + (in * t2 + dc2 * tan - 2.0 * tan * t2 * qc) * a[i * 6 + 4]
                                + tan * t2 * a[i * 6 + 5]));
        buf = in - tan * qc;
        score += a[i * 6] + k * (-2.0 * buf * a[i * 6 + 1] - 2.0 * tan * a[i * 6 + 2]
                                 + k * (buf * buf * a[i * 6 + 3] + 2.0 * buf * tan * a[i * 6 + 4]
                                        + tan * tan * a[i * 6 + 5]));
        in_prev = in;
        dc1_prev = dc1;
        dc2_prev = dc2;
        q_prev = qc;
    }
    grad[0] = g0;
    grad[1] = g1;
    *fval = score;
}

/* min_saxs.c:321-350 (intensity part). */
void orc_compile_intensity(const double *ct, const double *q, int qnum,
                           double rm, double c1, double c2, double *in)
{
    double pi = orc_get_pi();
    double mult = pow(4.0 * pi / 3.0, 3.0 / 2.0) * rm * rm / (16.0 * pi);
    double corr = -mult * (c1 * c1 - 1.0);
    for (int i = 0; i < qnum; i++) {
        double G = c1 * c1 * c1 * exp(corr * q[i] * q[i]);
        in[i] = model_in(ct, qnum, i, G, c2);
    }
}

typedef struct {
    const double *ct, *q;
    int qnum;
    const orc_fit_params *p;
} fit_ctx;

/* min_saxs.c:238-241: on every f/g request the scale is refreshed, then
 * chi^2 and its gradient are evaluated at that scale. */
static void fit_fg(void *vctx, const double *x, double *f, double *g)
{
    fit_ctx *c = (fit_ctx *)vctx;
    double k = orc_best_scale(c->ct, c->q, c->qnum, c->p, x[0], x[1]);
    orc_gradient(c->ct, c->q, c->qnum, c->p, k, x, g, f);
}

int orc_fit_one(double *ct, const double *q, int qnum, const orc_fit_params *p,
                int rescale, double *out3)
{
    if (rescale) {
        /* min_saxs.c:170-188 */
        double s = ct[VV * qnum] + ct[DD * qnum] + ct[WW * qnum] + ct[VW * qnum]
                 - ct[VD * qnum] - ct[DW * qnum];
        s = p->peak / s;
        for (int i = 0; i < qnum; i++) {
            ct[VV * qnum + i] *= s;
            ct[DD * qnum + i] *= s;
            ct[WW * qnum + i] *= s;
            ct[VW * qnum + i] *= s;
            ct[VD * qnum + i] *= s;
            ct[DW * qnum + i] *= s;
        }
    }
    double x[2] = {1.0, 0.0};
    double l[2] = {p->c1_lo, p->c2_lo};
    double u[2] = {p->c1_hi, p->c2_hi};
    int nbd[2] = {2, 2};
    fit_ctx c = {ct, q, qnum, p};
    orc_lbfgsb_result r;
    orc_lbfgsb_minimize(2, p->m, x, l, u, nbd, fit_fg, &c, p->factr, p->pgtol,
                        100000, &r);
    out3[0] = sqrt(r.f);
    out3[1] = x[0];
    out3[2] = x[1];
    return r.nevals;
}

/* ---- pdb2spf.c / profile.c / min_saxs.c profile helpers ---------------- */

static inline int lm_tab(int N, int l, int m) { return N * (m + N - 1) + l; }

/* pdb2spf.c:9-22 */
static void cart2sph(double x, double y, double z, double *r, double *th, double *fi)
{
    double pi = orc_get_pi();
    *r = sqrt(x * x + y * y + z * z);
    *th = acos(z / *r);
    if (y > 0.0) *fi = acos(x / sqrt(x * x + y * y));
    else *fi = -acos(x / sqrt(x * x + y * y)) + 2.0 * pi;
}

double orc_atoms2spf(int natoms, const double *xyz, const double *vac,
                     const double *dum, const double *sa, double h2o_ff,
                     const double *vdw, const double *qvals, int qnum, int L,
                     double *re, double *im)
{
    double pi = orc_get_pi();
    int K = (L + 1) * (L + 1);
    size_t blk = (size_t)qnum * K;
    memset(re, 0, sizeof(double) * 3 * blk);
    memset(im, 0, sizeof(double) * 3 * blk);
    /* saxs_utils.c:19-28 */
    double rm = 0.0;
    for (int i = 0; i < natoms; i++) rm += vdw[i];
    rm = rm / (double)(size_t)natoms;

    int N = L + 1;
    double *yn = (double *)calloc((size_t)N * (2 * N - 1), sizeof(double));
    double *leg = (double *)calloc((size_t)N * (2 * N - 1), sizeof(double));
    double *hre = (double *)calloc((size_t)K, sizeof(double));
    double *him = (double *)calloc((size_t)K, sizeof(double));
    orc_sph_norm(N, yn);
    double fp = 4.0 * pi;
    const double c_re[4] = {fp, 0.0, -fp, 0.0}, c_im[4] = {0.0, fp, 0.0, -fp};

    for (int a = 0; a < natoms; a++) {
        double r, th, fi;
        cart2sph(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2], &r, &th, &fi);
        memset(leg, 0, sizeof(double) * (size_t)N * (2 * N - 1));
        orc_legendre(N, cos(th), leg);
        double wf = 0.0;
        if (sa) wf = h2o_ff * sa[a];
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++) {
                int idx = l * (l + 1) + m;
                double h = yn[lm_tab(N, l, m)] * leg[lm_tab(N, l, abs(m))];
                hre[idx] = h * cos(fi * m);
                him[idx] = h * sin(-fi * m);
            }
        for (int q = 0; q < qnum; q++) {
            double *vr = re + (size_t)q * K, *vi = im + (size_t)q * K;
            double *dr = re + blk + (size_t)q * K, *di = im + blk + (size_t)q * K;
            double *wr = re + 2 * blk + (size_t)q * K, *wi = im + 2 * blk + (size_t)q * K;
            for (int l = 0; l <= L; l++) {
                double b = orc_sbessel(l, qvals[q] * r);
                for (int m = -l; m <= l; m++) {
                    int idx = l * (l + 1) + m;
                    double v_re = b * hre[idx], v_im = b * him[idx];
                    vr[idx] += vac[a] * v_re;
                    vi[idx] += vac[a] * v_im;
                    dr[idx] += dum[a] * v_re;
                    di[idx] += dum[a] * v_im;
                    wr[idx] += wf * v_re;
                    wi[idx] += wf * v_im;
                }
            }
        }
    }
    for (int X = 0; X < 3; X++)
        for (int q = 0; q < qnum; q++) {
            double *xr = re + X * blk + (size_t)q * K, *xi = im + X * blk + (size_t)q * K;
            for (int l = 0; l <= L; l++) {
                double cr = c_re[l % 4], ci = c_im[l % 4];
                for (int m = -l; m <= l; m++) {
                    int idx = l * (l + 1) + m;
                    double vr0 = xr[idx], vi0 = xi[idx];
                    xr[idx] = cr * vr0 - ci * vi0;
                    xi[idx] = cr * vi0 + ci * vr0;
                }
            }
        }
    free(yn); free(leg); free(hre); free(him);
    return rm;
}

/* profile.c:186-233; note G uses qvals[0] for every q (profile.c:214). */
void orc_profile_from_spf(const double *re, const double *im, int L,
                          const double *qvals, int qnum, double rm,
                          double c1, double c2, double rerr,
                          double *in, double *err)
{
    double pi = orc_get_pi();
    int K = (L + 1) * (L + 1);
    size_t blk = (size_t)qnum * K;
    double mult = pow(4.0 * pi / 3.0, 3.0 / 2.0) * rm * rm / (16.0 * pi);
    double corr = -mult * (c1 * c1 - 1.0);
    for (int q = 0; q < qnum; q++) {
        const double *vr = re + (size_t)q * K, *vi = im + (size_t)q * K;
        const double *dr = re + blk + (size_t)q * K, *di = im + blk + (size_t)q * K;
        const double *wr = re + 2 * blk + (size_t)q * K, *wi = im + 2 * blk + (size_t)q * K;
        double G = c1 * c1 * c1 * exp(corr * qvals[0] * qvals[0]);
        double s = 0.0;
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++) {
                int idx = l * (l + 1) + m;
                double ar = vr[idx] - G * dr[idx] + c2 * wr[idx];
                double ai = vi[idx] - G * di[idx] + c2 * wi[idx];
                s += ar * ar + ai * ai;
            }
        in[q] = s;
        err[q] = s * rerr;
    }
}

/* min_saxs.c:450-505 */
void orc_spf2cross_terms(const double *re, const double *im, int L, int qnum,
                         double *out)
{
    int K = (L + 1) * (L + 1);
    size_t blk = (size_t)qnum * K;
    for (int q = 0; q < qnum; q++) {
        const double *vr = re + (size_t)q * K, *vi = im + (size_t)q * K;
        const double *dr = re + blk + (size_t)q * K, *di = im + blk + (size_t)q * K;
        const double *wr = re + 2 * blk + (size_t)q * K, *wi = im + 2 * blk + (size_t)q * K;
        double s[6] = {0, 0, 0, 0, 0, 0};
        for (int id = 0; id < K; id++) {
            s[VV] += vr[id] * vr[id] + vi[id] * vi[id];
            s[DD] += dr[id] * dr[id] + di[id] * di[id];
            s[WW] += wr[id] * wr[id] + wi[id] * wi[id];
            s[VD] += vr[id] * dr[id] + vi[id] * di[id];
            s[VW] += vr[id] * wr[id] + vi[id] * wi[id];
            s[DW] += dr[id] * wr[id] + di[id] * wi[id];
        }
        s[VD] *= 2.0;
        s[VW] *= 2.0;
        s[DW] *= 2.0;
        for (int k = 0; k < 6; k++) out[k * qnum + q] = s[k];
    }
}
