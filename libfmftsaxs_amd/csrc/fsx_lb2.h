/*
 * fsx_lb2.h -- L-BFGS-B for the two fit variables (c1, c2) with the
 * limited-memory matrix held explicitly: the steps of lbfgsb.c's mainlb
 * (Cauchy point, subspace minimisation with the 3.0 projection, the
 * dcsrch line search, the correction-pair update, the stopping tests;
 * lbfgsb.c:327-515, subalgorithms.c, linesearch.c) with every product by
 * the L-BFGS matrix B taken with B itself.  For n = 2, B is a 2 x 2
 * matrix: theta I updated by BFGS with the held pairs, oldest first, which
 * is the matrix the reference's compact form theta I - W M W' represents
 * (Byrd, Nocedal and Schnabel 1994, Theorem 2.3).  Its state is a few
 * dozen values, against the hundreds of the compact form's factors
 * (fsx_lbfgsb.h), so that a GPU lane group can run it beside the
 * evaluation.
 *
 * The values equal the reference's up to rounding, with one difference:
 * the reference restarts its memory where a Cholesky factorisation of the
 * compact form's middle matrices fails (formk, formt, bmv: with three
 * pairs in two variables those matrices are nearly singular, and whether a
 * pivot comes out positive is decided by rounding); here B is always
 * formed, and the memory restarts only where a pair no longer gives a
 * positive definite update.  The fits end at the reference's end point
 * wherever the two paths meet the same minimum.
 */
#ifndef FSX_LB2_H
#define FSX_LB2_H

#include "fsx_lbfgsb.h"

typedef struct fsx_lb2 {
    double x[2], g[2], f;          /* iterate (trial point during a line search) */
    double S[FSX_LB_M][2], Y[FSX_LB_M][2];   /* correction pairs, oldest first */
    double theta;
    double z[2], d[2], t[2], r[2]; /* target point, direction, x and g at the search start */
    double fold, gd, gdold, stp, stpmx, dtd, sbgnrm;
    fsx_ls ls;
    int col, iter, ifun, iback, csave, nev, task, phase, iw[2];
} fsx_lb2;

/* B = theta I updated by BFGS with the pairs 0 .. col-1; b = (b00, b01,
 * b11).  Returns nonzero where a pair does not give a positive definite
 * update (s'y <= 0 or s'Bs <= 0). */
FSX_HD int lb2_bmat(const fsx_lb2 *s, double *b)
{
    double b00 = s->theta, b01 = 0.0, b11 = s->theta;
    for (int j = 0; j < FSX_LB_M; j++) {
        if (j >= s->col) break;
        const double s0 = s->S[j][0], s1 = s->S[j][1], y0 = s->Y[j][0], y1 = s->Y[j][1];
        const double bs0 = b00 * s0 + b01 * s1, bs1 = b01 * s0 + b11 * s1;
        const double sbs = s0 * bs0 + s1 * bs1, sy = s0 * y0 + s1 * y1;
        if (!(sbs > 0.0) || !(sy > 0.0)) return 1;
        b00 = b00 - bs0 * bs0 / sbs + y0 * y0 / sy;
        b01 = b01 - bs0 * bs1 / sbs + y0 * y1 / sy;
        b11 = b11 - bs1 * bs1 / sbs + y1 * y1 / sy;
    }
    b[0] = b00;
    b[1] = b01;
    b[2] = b11;
    return 0;
}

/* Element i (0 or 1) of a 2-vector, and a store to it, as selects: a GPU
 * thread then keeps the whole state in registers (no indexed addressing). */
#define LB2_AT(a, i) ((i) ? (a)[1] : (a)[0])
#define LB2_PUT(a, i, v)                                                      \
    do {                                                                      \
        const double v_ = (v);                                                \
        if (i) (a)[1] = v_;                                                   \
        else (a)[0] = v_;                                                     \
    } while (0)

/* subalgorithms.c cauchy: the generalized Cauchy point z along the
 * projected gradient path, the model's derivatives from B; iw as the
 * reference's iwhere (0 free, 1 at lower, 2 at upper, -3 free with a zero
 * gradient, 3 fixed).  With two variables there are at most two
 * breakpoints, taken smallest first (ties: the lower variable). */
FSX_HD void lb2_cauchy(fsx_lb2 *s, const double *l, const double *u, const double *b)
{
    const double *x = s->x, *g = s->g;
    double d[2];
    double tb0 = 0.0, tb1 = 0.0;
    int ord0 = 0, ord1 = 0, nbreak = 0, ibk = 0;
    double f1 = 0.0, bkmin = 0.0;
    s->z[0] = x[0];
    s->z[1] = x[1];
    if (s->sbgnrm <= 0.0) return;
    for (int i = 0; i < 2; i++) {
        const double neggi = -g[i];
        double tl = x[i] - l[i], tu = u[i] - x[i];
        if (s->iw[i] != 3) {
            const int xlower = tl <= 0.0, xupper = tu <= 0.0;
            s->iw[i] = 0;
            if (xlower) {
                if (neggi <= 0.0) s->iw[i] = 1;
            } else if (xupper) {
                if (neggi >= 0.0) s->iw[i] = 2;
            } else if (lb_abs(neggi) <= 0.0) {
                s->iw[i] = -3;
            }
        }
        if (s->iw[i] != 0) {
            d[i] = 0.0;
        } else {
            d[i] = neggi;
            f1 -= neggi * neggi;
            if (neggi != 0.0) {
                const double tt = neggi < 0.0 ? tl / (-neggi) : tu / neggi;
                if (nbreak == 0) {
                    tb0 = tt;
                    ord0 = i;
                    bkmin = tt;
                    ibk = 0;
                } else {
                    tb1 = tt;
                    ord1 = i;
                    if (tt < bkmin) {
                        bkmin = tt;
                        ibk = 1;
                    }
                }
                nbreak++;
            }
        }
    }
    if (nbreak == 0) return;   /* d = 0: the GCP is x */
    double f2 = d[0] * (b[0] * d[0] + b[1] * d[1]) + d[1] * (b[1] * d[0] + b[2] * d[1]);
    const double f2_org = f2;
    double dtm = -f1 / f2, tsum = 0.0, tj = 0.0;
    int nleft = nbreak;
    for (int it = 0; it < 2; it++) {
        if (it >= nbreak) break;
        const double tj0 = tj;
        const int kb = it == 0 ? ibk : 1 - ibk;
        tj = kb ? tb1 : tb0;
        const int ibp = kb ? ord1 : ord0;
        const double dt = tj - tj0;
        if (dtm < dt) goto gcp;
        tsum += dt;
        nleft--;
        const double dibp = LB2_AT(d, ibp);
        LB2_PUT(d, ibp, 0.0);
        if (dibp > 0.0) {
            LB2_PUT(s->z, ibp, LB2_AT(u, ibp));
            if (ibp) s->iw[1] = 2;
            else s->iw[0] = 2;
        } else {
            LB2_PUT(s->z, ibp, LB2_AT(l, ibp));
            if (ibp) s->iw[1] = 1;
            else s->iw[0] = 1;
        }
        if (nleft == 0 && nbreak == 2) return;   /* every variable at a bound */
        /* the model's slope and curvature along the remaining direction at
         * the breakpoint (displacement z - x there: the bound of ibp, tsum d
         * of the other variable) */
        const double zd0 = ibp ? tsum * d[0] : s->z[0] - x[0];
        const double zd1 = ibp ? s->z[1] - x[1] : tsum * d[1];
        const double bz0 = b[0] * zd0 + b[1] * zd1, bz1 = b[1] * zd0 + b[2] * zd1;
        f1 = g[0] * d[0] + g[1] * d[1] + d[0] * bz0 + d[1] * bz1;
        f2 = d[0] * (b[0] * d[0] + b[1] * d[1]) + d[1] * (b[1] * d[0] + b[2] * d[1]);
        f2 = lb_max(DBL_EPSILON * f2_org, f2);
        if (nleft > 0) {
            dtm = -f1 / f2;
        } else {
            /* every variable bounded: the path ends at the last breakpoint */
            dtm = 0.0;
        }
    }
gcp:
    if (dtm <= 0.0) dtm = 0.0;
    tsum += dtm;
    s->z[0] += tsum * d[0];
    s->z[1] += tsum * d[1];
}

/* cmprlb + subsm (3.0): the Newton step of the model over the free
 * variables (iw <= 0, in index order as freev lists them) from the Cauchy
 * point, projected on the box, with the reference's backtracking where the
 * projection is not a descent direction.  Returns nonzero where the reduced
 * matrix is not positive definite. */
FSX_HD int lb2_subsm(fsx_lb2 *s, const double *l, const double *u, const double *b)
{
    const double *x = s->x, *g = s->g;
    const int fr0 = s->iw[0] <= 0, fr1 = s->iw[1] <= 0;
    if (!fr0 && !fr1) return 0;
    /* r = -Z'(g + B (z - x)); dd per variable (0 where fixed) */
    const double zx0 = s->z[0] - x[0], zx1 = s->z[1] - x[1];
    const double r0 = -g[0] - (b[0] * zx0 + b[1] * zx1);
    const double r1 = -g[1] - (b[1] * zx0 + b[2] * zx1);
    double dd[2] = {0.0, 0.0};
    if (fr0 && fr1) {
        const double det = b[0] * b[2] - b[1] * b[1];
        if (!(b[0] > 0.0) || !(det > 0.0)) return 1;
        dd[0] = (b[2] * r0 - b[1] * r1) / det;
        dd[1] = (b[0] * r1 - b[1] * r0) / det;
    } else if (fr0) {
        if (!(b[0] > 0.0)) return 1;
        dd[0] = r0 / b[0];
    } else {
        if (!(b[2] > 0.0)) return 1;
        dd[1] = r1 / b[2];
    }
    const int fr[2] = {fr0, fr1};
    const double xp0 = s->z[0], xp1 = s->z[1];
    int iword = 0;
    for (int k = 0; k < 2; k++) {
        if (!fr[k]) continue;
        const double xk = lb_max(l[k], s->z[k] + dd[k]);
        s->z[k] = lb_min(u[k], xk);
        if (s->z[k] == l[k] || s->z[k] == u[k]) iword = 1;
    }
    if (iword == 0) return 0;
    const double dd_p = (s->z[0] - x[0]) * g[0] + (s->z[1] - x[1]) * g[1];
    if (!(dd_p > 0.0)) return 0;
    s->z[0] = xp0;
    s->z[1] = xp1;
    double alpha = 1.0, t1 = alpha;
    int ibd = -1;
    for (int k = 0; k < 2; k++) {
        if (!fr[k]) continue;
        const double dk = dd[k];
        if (dk < 0.0) {
            const double t2 = l[k] - s->z[k];
            if (t2 >= 0.0) t1 = 0.0;
            else if (dk * alpha < t2) t1 = t2 / dk;
        } else if (dk > 0.0) {
            const double t2 = u[k] - s->z[k];
            if (t2 <= 0.0) t1 = 0.0;
            else if (dk * alpha > t2) t1 = t2 / dk;
        }
        if (t1 < alpha) {
            alpha = t1;
            ibd = k;
        }
    }
    if (alpha < 1.0 && ibd >= 0) {
        const double dk = LB2_AT(dd, ibd);
        if (dk > 0.0) {
            LB2_PUT(s->z, ibd, LB2_AT(u, ibd));
            LB2_PUT(dd, ibd, 0.0);
        } else if (dk < 0.0) {
            LB2_PUT(s->z, ibd, LB2_AT(l, ibd));
            LB2_PUT(dd, ibd, 0.0);
        }
    }
    for (int k = 0; k < 2; k++)
        if (fr[k]) s->z[k] += alpha * dd[k];
    return 0;
}

/* x0 projected into the box (active, lbfgsb.c:540-600); the first
 * evaluation is due at s->x */
FSX_HD void fsx_lb2_start(fsx_lb2 *s, const fsx_fitpar *p, double x0, double x1)
{
    const double l[2] = {p->lo0, p->lo1}, u[2] = {p->hi0, p->hi1};
    s->x[0] = x0;
    s->x[1] = x1;
    for (int i = 0; i < 2; i++) {
        if (s->x[i] < l[i]) s->x[i] = l[i];
        else if (s->x[i] > u[i]) s->x[i] = u[i];
        s->iw[i] = (u[i] - l[i] <= 0.0) ? 3 : 0;
    }
    s->g[0] = s->g[1] = s->f = 0.0;
    s->col = 0;
    s->theta = 1.0;
    s->iter = 0;
    s->nev = 0;
    s->ifun = 0;
    s->iback = 0;
    s->csave = 0;
    s->task = FSX_T_FG;
    s->phase = FSX_PH_INIT;
    s->fold = s->gd = s->gdold = s->stp = s->stpmx = s->dtd = s->sbgnrm = 0.0;
}

/* mainlb from an evaluation (s->f, s->g at s->x) to the next point it
 * needs (returns 1, point in s->x) or to the end (returns 0, s->task). */
FSX_HD int fsx_lb2_resume(fsx_lb2 *s, const fsx_fitpar *p)
{
    const double l[2] = {p->lo0, p->lo1}, u[2] = {p->hi0, p->hi1};
    const double tol = p->factr * DBL_EPSILON;
    double *x = s->x, *g = s->g;
    double b[3];
    int nfree, restart = 0, info = 0, skip = 0;
    if (s->phase == FSX_PH_DONE) return 0;
    s->nev++;
    if (s->phase == FSX_PH_LS) goto ls_resume;
#define LB2_PROJGR(out)                                                         \
    do {                                                                        \
        double s_ = 0.;                                                         \
        for (int i_ = 0; i_ < 2; i_++) {                                        \
            double gi_ = g[i_];                                                 \
            if (gi_ < 0.) gi_ = lb_max(x[i_] - u[i_], gi_);                     \
            else gi_ = lb_min(x[i_] - l[i_], gi_);                              \
            s_ = lb_max(s_, lb_abs(gi_));                                       \
        }                                                                       \
        (out) = s_;                                                             \
    } while (0)
#define LB2_RESET()                                                             \
    do {                                                                        \
        s->col = 0;                                                             \
        s->theta = 1.;                                                          \
    } while (0)
    LB2_PROJGR(s->sbgnrm);
    if (s->sbgnrm <= p->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
    for (;;) {
        if (lb2_bmat(s, b)) {
            LB2_RESET();
            lb2_bmat(s, b);
        }
        lb2_cauchy(s, l, u, b);
        nfree = (s->iw[0] <= 0) + (s->iw[1] <= 0);
        if (!(nfree == 0 || s->col == 0)) {
            if (lb2_subsm(s, l, u, b)) {
                LB2_RESET();
                continue;
            }
        }
        s->d[0] = s->z[0] - x[0];
        s->d[1] = s->z[1] - x[1];
        /* lnsrlb entry (every variable boxed) */
        s->dtd = s->d[0] * s->d[0] + s->d[1] * s->d[1];
        if (s->iter == 0) {
            s->stpmx = 1.;
        } else {
            s->stpmx = 1e10;
            for (int i = 0; i < 2; i++) {
                const double a1 = s->d[i];
                if (a1 < 0.) {
                    const double a2 = l[i] - x[i];
                    if (a2 >= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx < a2) s->stpmx = a2 / a1;
                } else if (a1 > 0.) {
                    const double a2 = u[i] - x[i];
                    if (a2 <= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx > a2) s->stpmx = a2 / a1;
                }
            }
        }
        s->stp = 1.;
        s->t[0] = x[0]; s->t[1] = x[1];
        s->r[0] = g[0]; s->r[1] = g[1];
        s->fold = s->f;
        s->ifun = 0;
        s->iback = 0;
        s->csave = FSX_T_START;
        restart = 0;
        info = 0;
        for (;;) {
            s->gd = g[0] * s->d[0] + g[1] * s->d[1];
            skip = 0;
            if (s->ifun == 0) {
                s->gdold = s->gd;
                if (s->gd >= 0.) { info = -4; skip = 1; }
            }
            if (!skip) {
                lb_dcsrch(s->f, s->gd, &s->stp, 0., s->stpmx, &s->csave, &s->ls);
                if (!(s->csave >= 100 && s->csave <= 110) && !(s->csave >= 20 && s->csave <= 25)) {
                    s->task = FSX_T_FG_LN;
                    ++s->ifun;
                    s->iback = s->ifun - 1;
                    if (s->stp == 1.) {
                        x[0] = s->z[0]; x[1] = s->z[1];
                    } else {
                        x[0] = s->stp * s->d[0] + s->t[0];
                        x[1] = s->stp * s->d[1] + s->t[1];
                    }
                } else {
                    s->task = FSX_T_NEW_X;
                }
            }
            if (info != 0 || s->iback >= 20) {
                x[0] = s->t[0]; x[1] = s->t[1];
                g[0] = s->r[0]; g[1] = s->r[1];
                s->f = s->fold;
                if (s->col == 0) {
                    s->task = FSX_T_ABNORMAL;
                    ++s->iter;
                    goto done;
                }
                LB2_RESET();
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
                LB2_PROJGR(s->sbgnrm);
                break;
            }
        }
        if (restart) continue;
        if (s->sbgnrm <= p->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
        double ddum = lb_max(lb_max(lb_abs(s->fold), lb_abs(s->f)), 1.);
        if (s->fold - s->f <= tol * ddum) { s->task = FSX_T_CONV_F; goto done; }
        /* the pair: y = g - g_old, s = stp d */
        const double y0 = g[0] - s->r[0], y1 = g[1] - s->r[1];
        const double rr = y0 * y0 + y1 * y1;
        double dr;
        double s0 = s->d[0], s1 = s->d[1];
        if (s->stp == 1.) {
            dr = s->gd - s->gdold;
            ddum = -s->gdold;
        } else {
            dr = (s->gd - s->gdold) * s->stp;
            s0 = s->stp * s0;
            s1 = s->stp * s1;
            ddum = -s->gdold * s->stp;
        }
        if (dr <= DBL_EPSILON * ddum) continue;   /* update skipped */
        if (s->col < FSX_LB_M) {
            s->col++;
        } else {
            for (int j = 0; j < FSX_LB_M - 1; j++) {
                s->S[j][0] = s->S[j + 1][0]; s->S[j][1] = s->S[j + 1][1];
                s->Y[j][0] = s->Y[j + 1][0]; s->Y[j][1] = s->Y[j + 1][1];
            }
        }
        for (int j = 0; j < FSX_LB_M; j++) {   /* column col - 1, by selects */
            if (j == s->col - 1) {
                s->S[j][0] = s0; s->S[j][1] = s1;
                s->Y[j][0] = y0; s->Y[j][1] = y1;
            }
        }
        s->theta = rr / dr;
    }
#undef LB2_PROJGR
#undef LB2_RESET
done:
    s->phase = FSX_PH_DONE;
    return 0;
}

/* one complete fit from x = (c1, c2) with the evaluation ev(x, &f, g) */
FSX_HD int fsx_lb2_fit(const fsx_ct *ct, const fsx_fitpar *p, double *x, double *fout,
                       int *nevals)
{
    fsx_lb2 S;
    fsx_lb2_start(&S, p, x[0], x[1]);
    do {
        lb_eval_1p(ct, p, S.x, &S.f, S.g);
    } while (fsx_lb2_resume(&S, p));
    x[0] = S.x[0];
    x[1] = S.x[1];
    *fout = S.f;
    *nevals = S.nev;
    return S.task;
}

#endif
