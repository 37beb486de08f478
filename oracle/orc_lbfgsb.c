/*
 * orc_lbfgsb.c -- oracle restatement of L-BFGS-B 3.0 (Zhu, Byrd, Lu,
 * Nocedal; Morales-Nocedal 2011 revision) as vendored by the reference in
 * lbfgsb/src (C translation by S. Becker).
 *
 * TEST INFRASTRUCTURE (checker only; never linked into the product).
 *
 * The reference drives the optimiser by reverse communication (setulb
 * returns with task = FG / NEW_X, lbfgsb.c:52-316, mainlb 327-...).  Here
 * the same state machine runs as one direct loop that calls an objective
 * callback, which is equivalent because every NEW_X return is immediately
 * resumed by the caller (min_saxs.c:234-248).  The arithmetic of every
 * sub-step follows the reference operation for operation so that the
 * iterates are bit-identical (pinned against oracle/_ref, which is built
 * from lbfgsb/src).  Matrices are column-major with 1-based accessors, as
 * in the reference, to keep the index algebra checkable line by line.
 */
/*
 * This file restates L-BFGS-B 3.0 (Ciyou Zhu, Richard Byrd, Peihuang Lu,
 * Jorge Nocedal; revision by Jose Luis Morales and Jorge Nocedal, 2011),
 * following the C translation vendored by libfmftsaxs in lbfgsb/src, which
 * carries this licence:
 *
 *   BSD 3-clause license
 *   Copyright (c) 2015, Stephen Becker
 *   All rights reserved.
 *
 *   Redistribution and use in source and binary forms, with or without
 *   modification, are permitted provided that the following conditions are
 *   met:
 *
 *   * Redistributions of source code must retain the above copyright notice,
 *     this list of conditions and the following disclaimer.
 *
 *   * Redistributions in binary form must reproduce the above copyright
 *     notice, this list of conditions and the following disclaimer in the
 *     documentation and/or other materials provided with the distribution.
 *
 *   * Neither the name of the copyright holder nor the names of its
 *     contributors may be used to endorse or promote products derived from
 *     this software without specific prior written permission.
 *
 *   THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS
 *   IS" AND ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED
 *   TO, THE IMPLIED WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A
 *   PARTICULAR PURPOSE ARE DISCLAIMED. IN NO EVENT SHALL THE COPYRIGHT
 *   HOLDER OR CONTRIBUTORS BE LIABLE FOR ANY DIRECT, INDIRECT, INCIDENTAL,
 *   SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES (INCLUDING, BUT NOT LIMITED
 *   TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES; LOSS OF USE, DATA, OR
 *   PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND ON ANY THEORY OF
 *   LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT (INCLUDING
 *   NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF THIS
 *   SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* task codes, lbfgsb.h:80-130 */
enum { T_START = 1, T_NEW_X = 2, T_ABNORMAL = 3, T_RESTART = 4,
       T_FG = 10, T_FG_LN = 11, T_FG_ST = 12,
       T_CONVERGENCE = 20, T_CONV_GRAD = 21, T_CONV_F = 22,
       T_STOP_ITER = 32,
       T_WARNING_ROUND = 101, T_WARNING_XTOL = 102, T_WARNING_STPMAX = 103,
       T_WARNING_STPMIN = 104,
       T_ERROR_SMALLSTP = 201, T_ERROR_LARGESTP = 202, T_ERROR_INITIAL = 203,
       T_ERROR_FTOL = 204, T_ERROR_GTOL = 205, T_ERROR_XTOL = 206,
       T_ERROR_STP0 = 207, T_ERROR_STP1 = 208 };

#define IS_WARN(x) ((x) >= 100 && (x) <= 110)
#define IS_CONV(x) ((x) >= 20 && (x) <= 25)
#define IS_ERR(x)  ((x) >= 200 && (x) <= 240)

/* the reference's comparison macros (lbfgsb.h:60-68): ties pick the first */
static inline double mx(double a, double b) { return a >= b ? a : b; }
static inline double mn(double a, double b) { return a <= b ? a : b; }
static inline double ab(double a) { return a >= 0 ? a : -a; }

/* miniCBLAS.c restated: ddot's 5-way unrolled loop is a left-to-right
 * sum from 0.0, which is what this loop computes. */
static double v_dot(int n, const double *x, const double *y)
{
    double s = 0.;
    for (int i = 0; i < n; i++) s = s + x[i] * y[i];
    return s;
}
static void v_axpy(int n, double a, const double *x, double *y)
{
    if (n <= 0 || a == 0.) return;   /* miniCBLAS.c daxpy early exit */
    for (int i = 0; i < n; i++) y[i] += a * x[i];
}
static void v_scal(int n, double a, double *x)
{
    for (int i = 0; i < n; i++) x[i] = a * x[i];
}
static void v_copy(int n, const double *x, double *y)
{
    for (int i = 0; i < n; i++) y[i] = x[i];
}

/* linpack.c dpofa: upper Cholesky of a(lda, n), column-major, 1-based. */
static int chol_upper(double *a, int lda, int n)
{
#define A(i, j) a[((i) - 1) + ((j) - 1) * lda]
    for (int j = 1; j <= n; ++j) {
        double s = 0.;
        for (int k = 1; k <= j - 1; ++k) {
            double t = A(k, j) - v_dot(k - 1, &A(1, k), &A(1, j));
            t /= A(k, k);
            A(k, j) = t;
            s += t * t;
        }
        s = A(j, j) - s;
        if (s <= 0.) return j;
        A(j, j) = sqrt(s);
    }
    return 0;
#undef A
}

/* linpack.c dtrsl, the two cases L-BFGS-B uses on an upper-triangular T:
 * job 01 solves T x = b, job 11 solves T' x = b. */
static int tri_solve(const double *t, int ldt, int n, double *b, int job)
{
#define T_(i, j) t[((i) - 1) + ((j) - 1) * ldt]
    for (int i = 1; i <= n; ++i)
        if (T_(i, i) == 0.) return i;
    if (job == 1) {
        b[n - 1] /= T_(n, n);
        for (int jj = 2; jj <= n; ++jj) {
            int j = n - jj + 1;
            double tmp = -b[j];             /* b(j+1) */
            v_axpy(j, tmp, &T_(1, j + 1), b);
            b[j - 1] /= T_(j, j);
        }
    } else { /* 11 */
        b[0] /= T_(1, 1);
        for (int j = 2; j <= n; ++j) {
            b[j - 1] -= v_dot(j - 1, &T_(1, j), b);
            b[j - 1] /= T_(j, j);
        }
    }
    return 0;
#undef T_
}

typedef struct {
    int n, m;
    double *ws, *wy, *sy, *ss, *wt, *wn, *snd;
    double *z, *r, *d, *t, *xp, *wa;
    int *index, *iwhere, *indx2;
} lbws;

#define WS(i, j) w->ws[((i) - 1) + ((j) - 1) * w->n]
#define WY(i, j) w->wy[((i) - 1) + ((j) - 1) * w->n]
#define SY(i, j) w->sy[((i) - 1) + ((j) - 1) * w->m]
#define SS(i, j) w->ss[((i) - 1) + ((j) - 1) * w->m]
#define WT(i, j) w->wt[((i) - 1) + ((j) - 1) * w->m]
#define WN(i, j) w->wn[((i) - 1) + ((j) - 1) * 2 * w->m]
#define WN1(i, j) w->snd[((i) - 1) + ((j) - 1) * 2 * w->m]

/* subalgorithms.c bmv: product of the 2m x 2m middle matrix of the compact
 * L-BFGS formula with v (1-based v, p of length 2*col). */
static int mid_mat_vec(lbws *w, int col, const double *v, double *p)
{
    if (col == 0) return 0;
    p[col + 1 - 1] = v[col + 1 - 1];
    for (int i = 2; i <= col; ++i) {
        int i2 = col + i;
        double sum = 0.;
        for (int k = 1; k <= i - 1; ++k)
            sum += SY(i, k) * v[k - 1] / SY(k, k);
        p[i2 - 1] = v[i2 - 1] + sum;
    }
    int info = tri_solve(w->wt, w->m, col, &p[col], 11);
    if (info) return info;
    for (int i = 1; i <= col; ++i)
        p[i - 1] = v[i - 1] / sqrt(SY(i, i));
    info = tri_solve(w->wt, w->m, col, &p[col], 1);
    if (info) return info;
    for (int i = 1; i <= col; ++i)
        p[i - 1] = -p[i - 1] / sqrt(SY(i, i));
    for (int i = 1; i <= col; ++i) {
        double sum = 0.;
        for (int k = i + 1; k <= col; ++k)
            sum += SY(k, i) * p[col + k - 1] / SY(i, i);
        p[i - 1] += sum;
    }
    return 0;
}

/* subalgorithms.c active: project x0 into the box, classify variables. */
static void init_active(lbws *w, const double *l, const double *u,
                        const int *nbd, double *x, int *prjctd, int *cnstnd,
                        int *boxed)
{
    int n = w->n;
    *prjctd = 0; *cnstnd = 0; *boxed = 1;
    for (int i = 0; i < n; i++) {
        if (nbd[i] > 0) {
            if (nbd[i] <= 2 && x[i] <= l[i]) {
                if (x[i] < l[i]) { *prjctd = 1; x[i] = l[i]; }
            } else if (nbd[i] >= 2 && x[i] >= u[i]) {
                if (x[i] > u[i]) { *prjctd = 1; x[i] = u[i]; }
            }
        }
    }
    for (int i = 0; i < n; i++) {
        if (nbd[i] != 2) *boxed = 0;
        if (nbd[i] == 0) {
            w->iwhere[i] = -1;
        } else {
            *cnstnd = 1;
            w->iwhere[i] = (nbd[i] == 2 && u[i] - l[i] <= 0.) ? 3 : 0;
        }
    }
}

/* subalgorithms.c projgr: infinity norm of the projected gradient. */
static double proj_grad_norm(int n, const double *l, const double *u,
                             const int *nbd, const double *x, const double *g)
{
    double s = 0.;
    for (int i = 0; i < n; i++) {
        double gi = g[i];
        if (nbd[i] != 0) {
            if (gi < 0.) {
                if (nbd[i] >= 2) gi = mx(x[i] - u[i], gi);
            } else {
                if (nbd[i] <= 2) gi = mn(x[i] - l[i], gi);
            }
        }
        s = mx(s, ab(gi));
    }
    return s;
}

/* subalgorithms.c hpsolb: heap sort step on the breakpoints (1-based). */
static void heap_pop(int n, double *t, int *iorder, int iheap)
{
    /* t, iorder are 1-based views */
    if (iheap == 0) {
        for (int k = 2; k <= n; ++k) {
            double dd = t[k];
            int in = iorder[k];
            int i = k;
            while (i > 1) {
                int j = i / 2;
                if (dd < t[j]) { t[i] = t[j]; iorder[i] = iorder[j]; i = j; }
                else break;
            }
            t[i] = dd;
            iorder[i] = in;
        }
    }
    if (n > 1) {
        int i = 1;
        double out = t[1];
        int iou = iorder[1];
        double dd = t[n];
        int in = iorder[n];
        for (;;) {
            int j = i + i;
            if (j <= n - 1) {
                if (t[j + 1] < t[j]) ++j;
                if (t[j] < dd) { t[i] = t[j]; iorder[i] = iorder[j]; i = j; continue; }
            }
            break;
        }
        t[i] = dd;
        iorder[i] = in;
        t[n] = out;
        iorder[n] = iou;
    }
}

/* subalgorithms.c cauchy: generalised Cauchy point along the projected
 * steepest descent path.  Arrays x,l,u,g,d,xcp,t,iorder,iwhere are used
 * through 1-based views; p, c, wbp, v are the four 2m blocks of wa. */
static int cauchy_point(lbws *w, const double *x0, const double *l0,
                        const double *u0, const int *nbd0, const double *g0,
                        double theta, int col, int head, double sbgnrm,
                        int *nseg, double epsmch)
{
    int n = w->n, m = w->m;
    const double *x = x0 - 1, *l = l0 - 1, *u = u0 - 1, *g = g0 - 1;
    const int *nbd = nbd0 - 1;
    int *iorder = w->indx2 - 1, *iwhere = w->iwhere - 1;
    double *t = w->t - 1, *d = w->d - 1, *xcp = w->z - 1;
    double *p = w->wa, *c = w->wa + 2 * m, *wbp = w->wa + 4 * m, *v = w->wa + 6 * m;
    int info = 0;

    if (sbgnrm <= 0.) {
        v_copy(n, x0, w->z);
        return 0;
    }
    int bnded = 1, nfree = n + 1, nbreak = 0, ibkmin = 0;
    double bkmin = 0., f1 = 0., tl = 0., tu = 0.;
    int col2 = col << 1;
    for (int i = 0; i < col2; i++) p[i] = 0.;
    for (int i = 1; i <= n; ++i) {
        double neggi = -g[i];
        if (iwhere[i] != 3 && iwhere[i] != -1) {
            if (nbd[i] <= 2) tl = x[i] - l[i];
            if (nbd[i] >= 2) tu = u[i] - x[i];
            int xlower = nbd[i] <= 2 && tl <= 0.;
            int xupper = nbd[i] >= 2 && tu <= 0.;
            iwhere[i] = 0;
            if (xlower) {
                if (neggi <= 0.) iwhere[i] = 1;
            } else if (xupper) {
                if (neggi >= 0.) iwhere[i] = 2;
            } else {
                if (ab(neggi) <= 0.) iwhere[i] = -3;
            }
        }
        int pointr = head;
        if (iwhere[i] != 0 && iwhere[i] != -1) {
            d[i] = 0.;
        } else {
            d[i] = neggi;
            f1 -= neggi * neggi;
            for (int j = 1; j <= col; ++j) {
                p[j - 1] += WY(i, pointr) * neggi;
                p[col + j - 1] += WS(i, pointr) * neggi;
                pointr = pointr % m + 1;
            }
            if (nbd[i] <= 2 && nbd[i] != 0 && neggi < 0.) {
                ++nbreak;
                iorder[nbreak] = i;
                t[nbreak] = tl / (-neggi);
                if (nbreak == 1 || t[nbreak] < bkmin) { bkmin = t[nbreak]; ibkmin = nbreak; }
            } else if (nbd[i] >= 2 && neggi > 0.) {
                ++nbreak;
                iorder[nbreak] = i;
                t[nbreak] = tu / neggi;
                if (nbreak == 1 || t[nbreak] < bkmin) { bkmin = t[nbreak]; ibkmin = nbreak; }
            } else {
                --nfree;
                iorder[nfree] = i;
                if (ab(neggi) > 0.) bnded = 0;
            }
        }
    }
    if (theta != 1.) v_scal(col, theta, &p[col]);
    v_copy(n, x0, w->z);
    if (nbreak == 0 && nfree == n + 1) return 0;
    for (int j = 0; j < col2; j++) c[j] = 0.;
    double f2 = -theta * f1;
    double f2_org = f2;
    if (col > 0) {
        info = mid_mat_vec(w, col, p, v);
        if (info) return info;
        f2 -= v_dot(col2, v, p);
    }
    double dtm = -f1 / f2;
    double tsum = 0.;
    *nseg = 1;
    if (nbreak == 0) goto gcp_found;
    {
        int nleft = nbreak, iter = 1, ibp;
        double tj = 0.;
        for (;;) {
            double tj0 = tj;
            if (iter == 1) {
                tj = bkmin;
                ibp = iorder[ibkmin];
            } else {
                if (iter == 2 && ibkmin != nbreak) {
                    t[ibkmin] = t[nbreak];
                    iorder[ibkmin] = iorder[nbreak];
                }
                heap_pop(nleft, t, iorder, iter - 2);
                tj = t[nleft];
                ibp = iorder[nleft];
            }
            double dt = tj - tj0;
            if (dtm < dt) goto gcp_found;
            tsum += dt;
            --nleft;
            ++iter;
            double dibp = d[ibp];
            d[ibp] = 0.;
            double zibp;
            if (dibp > 0.) {
                zibp = u[ibp] - x[ibp];
                xcp[ibp] = u[ibp];
                iwhere[ibp] = 2;
            } else {
                zibp = l[ibp] - x[ibp];
                xcp[ibp] = l[ibp];
                iwhere[ibp] = 1;
            }
            if (nleft == 0 && nbreak == n) {
                dtm = dt;
                goto done;
            }
            ++(*nseg);
            double dibp2 = dibp * dibp;
            f1 = f1 + dt * f2 + dibp2 - theta * dibp * zibp;
            f2 -= theta * dibp2;
            if (col > 0) {
                v_axpy(col2, dt, p, c);
                int pointr = head;
                for (int j = 1; j <= col; ++j) {
                    wbp[j - 1] = WY(ibp, pointr);
                    wbp[col + j - 1] = theta * WS(ibp, pointr);
                    pointr = pointr % m + 1;
                }
                info = mid_mat_vec(w, col, wbp, v);
                if (info) return info;
                double wmc = v_dot(col2, c, v);
                double wmp = v_dot(col2, p, v);
                double wmw = v_dot(col2, wbp, v);
                v_axpy(col2, -dibp, wbp, p);
                f1 += dibp * wmc;
                f2 = f2 + dibp * 2. * wmp - dibp2 * wmw;
            }
            f2 = mx(epsmch * f2_org, f2);
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
gcp_found:
    if (dtm <= 0.) dtm = 0.;
    tsum += dtm;
    v_axpy(n, tsum, w->d, w->z);
done:
    if (col > 0) v_axpy(col2, dtm, p, c);
    return 0;
}

/* subalgorithms.c freev: free/active sets at the Cauchy point. */
static void free_vars(lbws *w, int *nfree, int *nenter, int *ileave, int *wrk,
                      int updatd, int cnstnd, int iter)
{
    int n = w->n;
    int *index = w->index - 1, *indx2 = w->indx2 - 1, *iwhere = w->iwhere - 1;
    *nenter = 0;
    *ileave = n + 1;
    if (iter > 0 && cnstnd) {
        for (int i = 1; i <= *nfree; ++i) {
            int k = index[i];
            if (iwhere[k] > 0) { --(*ileave); indx2[*ileave] = k; }
        }
        for (int i = *nfree + 1; i <= n; ++i) {
            int k = index[i];
            if (iwhere[k] <= 0) { ++(*nenter); indx2[*nenter] = k; }
        }
    }
    *wrk = *ileave < n + 1 || *nenter > 0 || updatd;
    *nfree = 0;
    int iact = n + 1;
    for (int i = 1; i <= n; ++i) {
        if (iwhere[i] <= 0) { ++(*nfree); index[*nfree] = i; }
        else { --iact; index[iact] = i; }
    }
}

/* subalgorithms.c formk: LEL' factorisation of the 2col x 2col middle
 * matrix restricted to the free variables (wn), maintaining N in snd. */
static int form_k(lbws *w, int nsub, int nenter, int ileave, int iupdat,
                  int updatd, double theta, int col, int head)
{
    int n = w->n, m = w->m;
    const int *ind = w->index - 1, *indx2 = w->indx2 - 1;
    int upcl;
    if (updatd) {
        if (iupdat > m) {
            for (int jy = 1; jy <= m - 1; ++jy) {
                int js = m + jy;
                for (int k = 0; k < m - jy; k++) WN1(jy + k, jy) = WN1(jy + 1 + k, jy + 1);
                for (int k = 0; k < m - jy; k++) WN1(js + k, js) = WN1(js + 1 + k, js + 1);
                for (int k = 0; k < m - 1; k++) WN1(m + 1 + k, jy) = WN1(m + 2 + k, jy + 1);
            }
        }
        int pbegin = 1, pend = nsub, dbegin = nsub + 1, dend = n;
        int iy = col, is = m + col;
        int ipntr = head + col - 1;
        if (ipntr > m) ipntr -= m;
        int jpntr = head;
        for (int jy = 1; jy <= col; ++jy) {
            int js = m + jy;
            double t1 = 0., t2 = 0., t3 = 0.;
            for (int k = pbegin; k <= pend; ++k) {
                int k1 = ind[k];
                t1 += WY(k1, ipntr) * WY(k1, jpntr);
            }
            for (int k = dbegin; k <= dend; ++k) {
                int k1 = ind[k];
                t2 += WS(k1, ipntr) * WS(k1, jpntr);
                t3 += WS(k1, ipntr) * WY(k1, jpntr);
            }
            WN1(iy, jy) = t1;
            WN1(is, js) = t2;
            WN1(is, jy) = t3;
            jpntr = jpntr % m + 1;
        }
        int jy = col;
        jpntr = head + col - 1;
        if (jpntr > m) jpntr -= m;
        ipntr = head;
        for (int i = 1; i <= col; ++i) {
            int is2 = m + i;
            double t3 = 0.;
            for (int k = pbegin; k <= pend; ++k) {
                int k1 = ind[k];
                t3 += WS(k1, ipntr) * WY(k1, jpntr);
            }
            ipntr = ipntr % m + 1;
            WN1(is2, jy) = t3;
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
                int k1 = indx2[k];
                t1 += WY(k1, ipntr) * WY(k1, jpntr);
                t2 += WS(k1, ipntr) * WS(k1, jpntr);
            }
            for (int k = ileave; k <= n; ++k) {
                int k1 = indx2[k];
                t3 += WY(k1, ipntr) * WY(k1, jpntr);
                t4 += WS(k1, ipntr) * WS(k1, jpntr);
            }
            WN1(iy, jy) = WN1(iy, jy) + t1 - t3;
            WN1(is, js) = WN1(is, js) - t2 + t4;
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
                int k1 = indx2[k];
                t1 += WS(k1, ipntr) * WY(k1, jpntr);
            }
            for (int k = ileave; k <= n; ++k) {
                int k1 = indx2[k];
                t3 += WS(k1, ipntr) * WY(k1, jpntr);
            }
            if (is <= jy + m) WN1(is, jy) = WN1(is, jy) + t1 - t3;
            else WN1(is, jy) = WN1(is, jy) - t1 + t3;
            jpntr = jpntr % m + 1;
        }
        ipntr = ipntr % m + 1;
    }
    int m2 = m << 1;
    for (int iy = 1; iy <= col; ++iy) {
        int is = col + iy, is1 = m + iy;
        for (int jy = 1; jy <= iy; ++jy) {
            int js = col + jy, js1 = m + jy;
            WN(jy, iy) = WN1(iy, jy) / theta;
            WN(js, is) = WN1(is1, js1) * theta;
        }
        for (int jy = 1; jy <= iy - 1; ++jy) WN(jy, is) = -WN1(is1, jy);
        for (int jy = iy; jy <= col; ++jy) WN(jy, is) = WN1(is1, jy);
        WN(iy, iy) += SY(iy, iy);
    }
    if (chol_upper(w->wn, m2, col)) return -1;
    int col2 = col << 1;
    for (int js = col + 1; js <= col2; ++js)
        tri_solve(w->wn, m2, col, &WN(1, js), 11);
    for (int is = col + 1; is <= col2; ++is)
        for (int js = is; js <= col2; ++js)
            WN(is, js) += v_dot(col, &WN(1, is), &WN(1, js));
    if (chol_upper(&WN(col + 1, col + 1), m2, col)) return -2;
    return 0;
}

/* subalgorithms.c formt: Cholesky of theta*S'S + L D^-1 L'. */
static int form_t(lbws *w, int col, double theta)
{
    for (int j = 1; j <= col; ++j) WT(1, j) = theta * SS(1, j);
    for (int i = 2; i <= col; ++i) {
        for (int j = i; j <= col; ++j) {
            int k1 = (i <= j ? i : j) - 1;
            double dd = 0.;
            for (int k = 1; k <= k1; ++k)
                dd += SY(i, k) * SY(j, k) / SY(k, k);
            WT(i, j) = dd + theta * SS(i, j);
        }
    }
    return chol_upper(w->wt, w->m, col) ? -3 : 0;
}

/* subalgorithms.c cmprlb: reduced gradient at the Cauchy point. */
static int reduced_grad(lbws *w, const double *x0, const double *g0,
                        double theta, int col, int head, int nfree, int cnstnd)
{
    int n = w->n, m = w->m;
    const double *x = x0 - 1, *g = g0 - 1, *z = w->z - 1;
    double *r = w->r - 1;
    const int *index = w->index - 1;
    if (!cnstnd && col > 0) {
        for (int i = 1; i <= n; ++i) r[i] = -g[i];
        return 0;
    }
    for (int i = 1; i <= nfree; ++i) {
        int k = index[i];
        r[i] = -theta * (z[k] - x[k]) - g[k];
    }
    if (mid_mat_vec(w, col, w->wa + 2 * m, w->wa)) return -8;
    int pointr = head;
    for (int j = 1; j <= col; ++j) {
        double a1 = w->wa[j - 1], a2 = theta * w->wa[col + j - 1];
        for (int i = 1; i <= nfree; ++i) {
            int k = index[i];
            r[i] = r[i] + WY(k, pointr) * a1 + WS(k, pointr) * a2;
        }
        pointr = pointr % m + 1;
    }
    return 0;
}

/* subalgorithms.c subsm (3.0): subspace minimisation, then projection of
 * the Newton point into the box with the backtracking safeguard. */
static int subspace_min(lbws *w, int nsub, const double *l0, const double *u0,
                        const int *nbd0, const double *xx0, const double *gg0,
                        double theta, int col, int head, int *iword)
{
    int n = w->n, m = w->m;
    const double *l = l0 - 1, *u = u0 - 1, *xx = xx0 - 1, *gg = gg0 - 1;
    const int *nbd = nbd0 - 1, *ind = w->index - 1;
    double *x = w->z - 1, *d = w->r - 1, *wv = w->wa;
    if (nsub <= 0) return 0;
    int pointr = head;
    for (int i = 1; i <= col; ++i) {
        double t1 = 0., t2 = 0.;
        for (int j = 1; j <= nsub; ++j) {
            int k = ind[j];
            t1 += WY(k, pointr) * d[j];
            t2 += WS(k, pointr) * d[j];
        }
        wv[i - 1] = t1;
        wv[col + i - 1] = theta * t2;
        pointr = pointr % m + 1;
    }
    int m2 = m << 1, col2 = col << 1;
    int info = tri_solve(w->wn, m2, col2, wv, 11);
    if (info) return info;
    for (int i = 0; i < col; ++i) wv[i] = -wv[i];
    info = tri_solve(w->wn, m2, col2, wv, 1);
    if (info) return info;
    pointr = head;
    for (int jy = 1; jy <= col; ++jy) {
        int js = col + jy;
        for (int i = 1; i <= nsub; ++i) {
            int k = ind[i];
            d[i] = d[i] + WY(k, pointr) * wv[jy - 1] / theta + WS(k, pointr) * wv[js - 1];
        }
        pointr = pointr % m + 1;
    }
    v_scal(nsub, 1. / theta, w->r);
    *iword = 0;
    v_copy(n, w->z, w->xp);
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[i];
        double dk = d[i], xk = x[k];
        if (nbd[k] != 0) {
            if (nbd[k] == 1) {
                x[k] = mx(l[k], xk + dk);
                if (x[k] == l[k]) *iword = 1;
            } else if (nbd[k] == 2) {
                xk = mx(l[k], xk + dk);
                x[k] = mn(u[k], xk);
                if (x[k] == l[k] || x[k] == u[k]) *iword = 1;
            } else if (nbd[k] == 3) {
                x[k] = mn(u[k], xk + dk);
                if (x[k] == u[k]) *iword = 1;
            }
        } else {
            x[k] = xk + dk;
        }
    }
    if (*iword == 0) return 0;
    double dd_p = 0.;
    for (int i = 1; i <= n; ++i) dd_p += (x[i] - xx[i]) * gg[i];
    if (dd_p > 0.) {
        v_copy(n, w->xp, w->z);
    } else {
        return 0;
    }
    double alpha = 1., t1 = alpha;
    int ibd = 0;
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[i];
        double dk = d[i];
        if (nbd[k] != 0) {
            if (dk < 0. && nbd[k] <= 2) {
                double t2 = l[k] - x[k];
                if (t2 >= 0.) t1 = 0.;
                else if (dk * alpha < t2) t1 = t2 / dk;
            } else if (dk > 0. && nbd[k] >= 2) {
                double t2 = u[k] - x[k];
                if (t2 <= 0.) t1 = 0.;
                else if (dk * alpha > t2) t1 = t2 / dk;
            }
            if (t1 < alpha) { alpha = t1; ibd = i; }
        }
    }
    if (alpha < 1.) {
        double dk = d[ibd];
        int k = ind[ibd];
        if (dk > 0.) { x[k] = u[k]; d[ibd] = 0.; }
        else if (dk < 0.) { x[k] = l[k]; d[ibd] = 0.; }
    }
    for (int i = 1; i <= nsub; ++i) {
        int k = ind[i];
        x[k] += alpha * d[i];
    }
    return 0;
}

/* subalgorithms.c matupd: append the newest s, y pair and update S'Y, S'S. */
static void update_pairs(lbws *w, int *itail, int iupdat, int *col, int *head,
                         double *theta, double rr, double dr, double stp,
                         double dtd)
{
    int n = w->n, m = w->m;
    if (iupdat <= m) {
        *col = iupdat;
        *itail = (*head + iupdat - 2) % m + 1;
    } else {
        *itail = *itail % m + 1;
        *head = *head % m + 1;
    }
    v_copy(n, w->d, &WS(1, *itail));
    v_copy(n, w->r, &WY(1, *itail));
    *theta = rr / dr;
    if (iupdat > m) {
        for (int j = 1; j <= *col - 1; ++j) {
            for (int k = 0; k < j; k++) SS(1 + k, j) = SS(2 + k, j + 1);
            for (int k = 0; k < *col - j; k++) SY(j + k, j) = SY(j + 1 + k, j + 1);
        }
    }
    int pointr = *head;
    for (int j = 1; j <= *col - 1; ++j) {
        SY(*col, j) = v_dot(n, w->d, &WY(1, pointr));
        SS(j, *col) = v_dot(n, &WS(1, pointr), w->d);
        pointr = pointr % m + 1;
    }
    if (stp == 1.) SS(*col, *col) = dtd;
    else SS(*col, *col) = stp * stp * dtd;
    SY(*col, *col) = dr;
}

/* linesearch.c dcstep: safeguarded cubic/quadratic step. */
static void ls_step(double *stx, double *fx, double *dx, double *sty,
                    double *fy, double *dy, double *stp, double fp, double dp,
                    int *brackt, double stpmin, double stpmax)
{
    double sgnd = dp * (*dx / ab(*dx));
    double stpf, stpc, stpq, theta, s, gamma, p, q, r;
    if (fp > *fx) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = mx(mx(ab(theta), ab(*dx)), ab(dp));
        double t = theta / s;
        gamma = s * sqrt(t * t - *dx / s * (dp / s));
        if (*stp < *stx) gamma = -gamma;
        p = gamma - *dx + theta;
        q = gamma - *dx + gamma + dp;
        r = p / q;
        stpc = *stx + r * (*stp - *stx);
        stpq = *stx + *dx / ((*fx - fp) / (*stp - *stx) + *dx) / 2. * (*stp - *stx);
        if (ab(stpc - *stx) < ab(stpq - *stx)) stpf = stpc;
        else stpf = stpc + (stpq - stpc) / 2.;
        *brackt = 1;
    } else if (sgnd < 0.) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = mx(mx(ab(theta), ab(*dx)), ab(dp));
        double t = theta / s;
        gamma = s * sqrt(t * t - *dx / s * (dp / s));
        if (*stp > *stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma - dp + gamma + *dx;
        r = p / q;
        stpc = *stp + r * (*stx - *stp);
        stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
        if (ab(stpc - *stp) > ab(stpq - *stp)) stpf = stpc;
        else stpf = stpq;
        *brackt = 1;
    } else if (ab(dp) < ab(*dx)) {
        theta = (*fx - fp) * 3. / (*stp - *stx) + *dx + dp;
        s = mx(mx(ab(theta), ab(*dx)), ab(dp));
        double t = theta / s;
        gamma = s * sqrt(mx(0., t * t - *dx / s * (dp / s)));
        if (*stp > *stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma + (*dx - dp) + gamma;
        r = p / q;
        if (r < 0. && gamma != 0.) stpc = *stp + r * (*stx - *stp);
        else if (*stp > *stx) stpc = stpmax;
        else stpc = stpmin;
        stpq = *stp + dp / (dp - *dx) * (*stx - *stp);
        if (*brackt) {
            if (ab(stpc - *stp) < ab(stpq - *stp)) stpf = stpc;
            else stpf = stpq;
            if (*stp > *stx) stpf = mn(*stp + (*sty - *stp) * .66, stpf);
            else stpf = mx(*stp + (*sty - *stp) * .66, stpf);
        } else {
            if (ab(stpc - *stp) > ab(stpq - *stp)) stpf = stpc;
            else stpf = stpq;
            stpf = mn(stpmax, stpf);
            stpf = mx(stpmin, stpf);
        }
    } else {
        if (*brackt) {
            theta = (fp - *fy) * 3. / (*sty - *stp) + *dy + dp;
            s = mx(mx(ab(theta), ab(*dy)), ab(dp));
            double t = theta / s;
            gamma = s * sqrt(t * t - *dy / s * (dp / s));
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

/* dcsrch state, persisted across calls like isave/dsave (linesearch.c). */
typedef struct {
    int brackt, stage;
    double ginit, gtest, gx, gy, finit, fx, fy, stx, sty, stmin, stmax,
           width, width1;
} lsstate;

/* linesearch.c dcsrch (More'-Thuente), ftol=1e-3, gtol=0.9, xtol=0.1. */
static void ls_search(double f, double g, double *stp, double stpmin,
                      double stpmax, int *task, lsstate *s)
{
    const double ftol = .001, gtol = .9, xtol = .1;
    if (*task == T_START) {
        if (*stp < stpmin) *task = T_ERROR_SMALLSTP;
        if (*stp > stpmax) *task = T_ERROR_LARGESTP;
        if (g >= 0.) *task = T_ERROR_INITIAL;
        if (stpmin < 0.) *task = T_ERROR_STP0;
        if (stpmax < stpmin) *task = T_ERROR_STP1;
        if (IS_ERR(*task)) return;
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
        *task = T_FG;
        return;
    }
    double ftest = s->finit + *stp * s->gtest;
    if (s->stage == 1 && f <= ftest && g >= 0.) s->stage = 2;
    if (s->brackt && (*stp <= s->stmin || *stp >= s->stmax)) *task = T_WARNING_ROUND;
    if (s->brackt && s->stmax - s->stmin <= xtol * s->stmax) *task = T_WARNING_XTOL;
    if (*stp == stpmax && f <= ftest && g <= s->gtest) *task = T_WARNING_STPMAX;
    if (*stp == stpmin && (f > ftest || g >= s->gtest)) *task = T_WARNING_STPMIN;
    if (f <= ftest && ab(g) <= gtol * (-s->ginit)) *task = T_CONVERGENCE;
    if (IS_WARN(*task) || IS_CONV(*task)) return;
    if (s->stage == 1 && f <= s->fx && f > ftest) {
        double fm = f - *stp * s->gtest;
        double fxm = s->fx - s->stx * s->gtest;
        double fym = s->fy - s->sty * s->gtest;
        double gm = g - s->gtest;
        double gxm = s->gx - s->gtest;
        double gym = s->gy - s->gtest;
        ls_step(&s->stx, &fxm, &gxm, &s->sty, &fym, &gym, stp, fm, gm,
                &s->brackt, s->stmin, s->stmax);
        s->fx = fxm + s->stx * s->gtest;
        s->fy = fym + s->sty * s->gtest;
        s->gx = gxm + s->gtest;
        s->gy = gym + s->gtest;
    } else {
        ls_step(&s->stx, &s->fx, &s->gx, &s->sty, &s->fy, &s->gy, stp, f, g,
                &s->brackt, s->stmin, s->stmax);
    }
    if (s->brackt) {
        if (ab(s->sty - s->stx) >= s->width1 * .66)
            *stp = s->stx + (s->sty - s->stx) * .5;
        s->width1 = s->width;
        s->width = ab(s->sty - s->stx);
    }
    if (s->brackt) {
        s->stmin = mn(s->stx, s->sty);
        s->stmax = mx(s->stx, s->sty);
    } else {
        s->stmin = *stp + (*stp - s->stx) * 1.1;
        s->stmax = *stp + (*stp - s->stx) * 4.;
    }
    *stp = mx(*stp, stpmin);
    *stp = mn(*stp, stpmax);
    if ((s->brackt && (*stp <= s->stmin || *stp >= s->stmax))
        || (s->brackt && s->stmax - s->stmin <= xtol * s->stmax))
        *stp = s->stx;
    *task = T_FG;
}

int orc_lbfgsb_minimize(int n, int m, double *x, const double *l,
                        const double *u, const int *nbd, orc_fg_fn fg,
                        void *ctx, double factr, double pgtol, int max_evals,
                        orc_lbfgsb_result *res)
{
    lbws W, *w = &W;
    w->n = n; w->m = m;
    size_t nd = (size_t)2 * m * n + (size_t)3 * m * m + (size_t)8 * m * m
              + (size_t)5 * n + (size_t)8 * m;
    double *buf = (double *)calloc(nd, sizeof(double));
    int *ibuf = (int *)calloc((size_t)3 * n, sizeof(int));
    double *g = (double *)calloc((size_t)n, sizeof(double));
    double *q = buf;
    w->ws = q; q += (size_t)m * n;
    w->wy = q; q += (size_t)m * n;
    w->sy = q; q += (size_t)m * m;
    w->ss = q; q += (size_t)m * m;
    w->wt = q; q += (size_t)m * m;
    w->wn = q; q += (size_t)4 * m * m;
    w->snd = q; q += (size_t)4 * m * m;
    w->z = q; q += n;
    w->r = q; q += n;
    w->d = q; q += n;
    w->t = q; q += n;
    w->xp = q; q += n;
    w->wa = q;
    w->index = ibuf; w->iwhere = ibuf + n; w->indx2 = ibuf + 2 * n;

    const double epsmch = DBL_EPSILON;
    int col = 0, head = 1, iupdat = 0, updatd = 0, iback = 0, itail = 0;
    int iword = 0, nact = 0, ileave = 0, nenter = 0, iter = 0, nfgv = 0;
    int nseg = 0, nintol = 0, nskip = 0, nfree = n, ifun = 0, info = 0;
    double theta = 1., fold = 0., dnorm = 0., gd = 0., stpmx = 0.;
    double sbgnrm = 0., stp = 0., gdold = 0., dtd = 0., f = 0.;
    double tol = factr * epsmch;
    int prjctd, cnstnd, boxed, wrk = 0, task, csave = 0, nevals = 0;
    lsstate ls;
    memset(&ls, 0, sizeof(ls));
    (void)nact; (void)nintol; (void)nskip; (void)iword; (void)prjctd;

    init_active(w, l, u, nbd, x, &prjctd, &cnstnd, &boxed);
    fg(ctx, x, &f, g);
    nevals++;
    nfgv = 1;
    sbgnrm = proj_grad_norm(n, l, u, nbd, x, g);
    if (sbgnrm <= pgtol) { task = T_CONV_GRAD; goto finish; }

    for (;;) {
        /* L222: one iteration */
        iword = -1;
        if (!cnstnd && col > 0) {
            v_copy(n, x, w->z);
            wrk = updatd;
            nseg = 0;
        } else {
            info = cauchy_point(w, x, l, u, nbd, g, theta, col, head, sbgnrm,
                                &nseg, epsmch);
            if (info) {
                info = 0; col = 0; head = 1; theta = 1.; iupdat = 0; updatd = 0;
                continue;
            }
            nintol += nseg;
            free_vars(w, &nfree, &nenter, &ileave, &wrk, updatd, cnstnd, iter);
            nact = n - nfree;
        }
        /* L333 */
        if (!(nfree == 0 || col == 0)) {
            if (wrk) info = form_k(w, nfree, nenter, ileave, iupdat, updatd,
                                   theta, col, head);
            if (info) {
                info = 0; col = 0; head = 1; theta = 1.; iupdat = 0; updatd = 0;
                continue;
            }
            info = reduced_grad(w, x, g, theta, col, head, nfree, cnstnd);
            if (!info)
                info = subspace_min(w, nfree, l, u, nbd, x, g, theta, col,
                                    head, &iword);
            if (info) {
                info = 0; col = 0; head = 1; theta = 1.; iupdat = 0; updatd = 0;
                continue;
            }
        }
        /* L555: search direction */
        for (int i = 0; i < n; i++) w->d[i] = w->z[i] - x[i];

        /* lnsrlb, first entry (linesearch.c:3-105) */
        dtd = v_dot(n, w->d, w->d);
        dnorm = sqrt(dtd);
        stpmx = 1e10;
        if (cnstnd) {
            if (iter == 0) {
                stpmx = 1.;
            } else {
                for (int i = 0; i < n; i++) {
                    double a1 = w->d[i];
                    if (nbd[i] != 0) {
                        if (a1 < 0. && nbd[i] <= 2) {
                            double a2 = l[i] - x[i];
                            if (a2 >= 0.) stpmx = 0.;
                            else if (a1 * stpmx < a2) stpmx = a2 / a1;
                        } else if (a1 > 0. && nbd[i] >= 2) {
                            double a2 = u[i] - x[i];
                            if (a2 <= 0.) stpmx = 0.;
                            else if (a1 * stpmx > a2) stpmx = a2 / a1;
                        }
                    }
                }
            }
        }
        if (iter == 0 && !boxed) stp = mn(1. / dnorm, stpmx);
        else stp = 1.;
        v_copy(n, x, w->t);
        v_copy(n, g, w->r);
        fold = f;
        ifun = 0;
        iback = 0;
        csave = T_START;
        int restart = 0;
        for (;;) {
            /* L556 */
            gd = v_dot(n, g, w->d);
            int ls_done = 0;
            if (ifun == 0) {
                gdold = gd;
                if (gd >= 0.) { info = -4; ls_done = 1; }
            }
            if (!ls_done) {
                ls_search(f, gd, &stp, 0., stpmx, &csave, &ls);
                if (!IS_WARN(csave) && !IS_CONV(csave)) {
                    task = T_FG_LN;
                    ++ifun;
                    ++nfgv;
                    iback = ifun - 1;
                    if (stp == 1.) v_copy(n, w->z, x);
                    else for (int i = 0; i < n; i++) x[i] = stp * w->d[i] + w->t[i];
                } else {
                    task = T_NEW_X;
                }
            }
            if (info != 0 || iback >= 20) {
                v_copy(n, w->t, x);
                v_copy(n, w->r, g);
                f = fold;
                if (col == 0) {
                    if (info == 0) { info = -9; --nfgv; --ifun; --iback; }
                    task = T_ABNORMAL;
                    ++iter;
                    goto finish;
                }
                if (info == 0) --nfgv;
                info = 0; col = 0; head = 1; theta = 1.; iupdat = 0; updatd = 0;
                task = T_RESTART;
                restart = 1;
                break;
            } else if (task == T_FG_LN) {
                if (max_evals > 0 && nevals >= max_evals) { task = T_STOP_ITER; goto finish; }
                fg(ctx, x, &f, g);
                nevals++;
                continue;
            } else {
                ++iter;
                sbgnrm = proj_grad_norm(n, l, u, nbd, x, g);
                break;
            }
        }
        if (restart) continue;
        /* L777 */
        if (sbgnrm <= pgtol) { task = T_CONV_GRAD; goto finish; }
        double ddum = mx(mx(ab(fold), ab(f)), 1.);
        if (fold - f <= tol * ddum) {
            task = T_CONV_F;
            if (iback >= 10) info = -5;
            goto finish;
        }
        for (int i = 0; i < n; i++) w->r[i] = g[i] - w->r[i];
        double rr = v_dot(n, w->r, w->r), dr;
        if (stp == 1.) {
            dr = gd - gdold;
            ddum = -gdold;
        } else {
            dr = (gd - gdold) * stp;
            v_scal(n, stp, w->d);
            ddum = -gdold * stp;
        }
        if (dr <= epsmch * ddum) {
            ++nskip;
            updatd = 0;
            continue;
        }
        updatd = 1;
        ++iupdat;
        update_pairs(w, &itail, iupdat, &col, &head, &theta, rr, dr, stp, dtd);
        if (form_t(w, col, theta)) {
            info = 0; col = 0; head = 1; theta = 1.; iupdat = 0; updatd = 0;
        }
    }
finish:
    if (res) {
        res->task = task;
        res->iter = iter;
        res->nfgv = nfgv;
        res->nevals = nevals;
        res->f = f;
    }
    free(buf);
    free(ibuf);
    free(g);
    return task;
}
