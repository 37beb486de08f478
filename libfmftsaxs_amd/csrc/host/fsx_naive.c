/*
 * fsx_naive.c -- the per-pose brute-force scorer of tools/score_ft_naive.c
 * (:156-212) as a library call, with the atom expansion on the GPU: for each
 * pose the two molecules are placed in real space, joined, centred, expanded
 * to L (atom_grp2spf_inplace, pdb2spf.c:24-152, through fsx_atoms2spf) and the
 * complex's cross-term profile fitted as sxs_spf2fitted_profile
 * (min_saxs.c:507-511).  It shares no code path with the FFT correlation, so
 * it measures that path's truncation at L independently.
 */
#include "../fsx_internal.h"
#include "../fsx_lbfgsb.h"
#include "saxs_utils.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

int fsx_score_naive(int nrec, int nlig, const double *xyz, const double *vac,
                    const double *dum, const double *sa, double h2o_ff, const double *vdw,
                    const double *qvals, int qnum, int L, const double *a, double mult,
                    double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                    const double *poses, int n, double *out, double *cross)
{
    if (nrec < 1 || nlig < 1 || !xyz || !vac || !dum || !vdw || !qvals || qnum < 2 || L < 0 ||
        L > 40 || !a || n < 0 || (n > 0 && (!poses || !out)))
        return FSX_ERR_ARG;
    const int N = nrec + nlig, K2 = (L + 1) * (L + 1);
    const size_t blk = (size_t)qnum * K2;
    double *x = (double *)malloc(sizeof(double) * 3 * (size_t)N);
    double *re = (double *)malloc(sizeof(double) * 3 * blk);
    double *im = (double *)malloc(sizeof(double) * 3 * blk);
    double *ct = (double *)malloc(sizeof(double) * 6 * (size_t)qnum);
    if (!x || !re || !im || !ct) {
        free(x); free(re); free(im); free(ct);
        return FSX_ERR_NOMEM;
    }
    int rc = FSX_OK;
    for (int i = 0; i < n && rc == FSX_OK; i++) {
        const double *p = poses + 6 * (size_t)i;   /* z, b1, g1, a2, b2, g2 */
        struct mol_matrix3 r1, r2;
        sxs_fill_active_rotation_matrix(&r1, 0.0, p[1], p[2]);
        sxs_fill_active_rotation_matrix(&r2, p[3], p[4], p[5]);
        /* receptor R(0, b1, g1) x; ligand R(a2, b2, g2) x + (0, 0, z)
         * (score_ft_naive.c:160-164) */
        for (int k = 0; k < N; k++) {
            const struct mol_matrix3 *m = k < nrec ? &r1 : &r2;
            const double *s = xyz + 3 * (size_t)k;
            double *d = x + 3 * (size_t)k;
            d[0] = m->m11 * s[0] + m->m12 * s[1] + m->m13 * s[2];
            d[1] = m->m21 * s[0] + m->m22 * s[1] + m->m23 * s[2];
            d[2] = m->m31 * s[0] + m->m32 * s[1] + m->m33 * s[2];
            if (k >= nrec) d[2] += p[0];
        }
        /* the complex moved to its centroid (score_ft_naive.c:175-178) */
        double cx = 0.0, cy = 0.0, cz = 0.0;
        for (int k = 0; k < N; k++) {
            cx += x[3 * k];
            cy += x[3 * k + 1];
            cz += x[3 * k + 2];
        }
        cx /= N;
        cy /= N;
        cz /= N;
        for (int k = 0; k < N; k++) {
            x[3 * k] -= cx;
            x[3 * k + 1] -= cy;
            x[3 * k + 2] -= cz;
        }
        rc = fsx_atoms2spf(N, x, vac, dum, sa, h2o_ff, vdw, qvals, qnum, L, re, im, NULL);
        if (rc != FSX_OK) break;
        /* cross terms of the complex (sxs_spf2cross_terms, min_saxs.c:450-505) */
        for (int q = 0; q < qnum; q++) {
            const double *vr = re + (size_t)q * K2, *vi = im + (size_t)q * K2;
            const double *dr = re + blk + (size_t)q * K2, *di = im + blk + (size_t)q * K2;
            const double *wr = re + 2 * blk + (size_t)q * K2, *wi = im + 2 * blk + (size_t)q * K2;
            double vv = 0, vd = 0, vw = 0, dd = 0, dw = 0, ww = 0;
            for (int e = 0; e < K2; e++) {
                vv += vr[e] * vr[e] + vi[e] * vi[e];
                dd += dr[e] * dr[e] + di[e] * di[e];
                ww += wr[e] * wr[e] + wi[e] * wi[e];
                vd += vr[e] * dr[e] + vi[e] * di[e];
                vw += vr[e] * wr[e] + vi[e] * wi[e];
                dw += dr[e] * wr[e] + di[e] * wi[e];
            }
            ct[0 * qnum + q] = vv;
            ct[1 * qnum + q] = vd * 2.0;
            ct[2 * qnum + q] = vw * 2.0;
            ct[3 * qnum + q] = dd;
            ct[4 * qnum + q] = dw * 2.0;
            ct[5 * qnum + q] = ww;
        }
        if (cross) memcpy(cross + 6 * (size_t)qnum * i, ct, sizeof(double) * 6 * (size_t)qnum);
        /* the fit of sxs_lbfgs_fitting (min_saxs.c:196-259), no peak rescale */
        fsx_ct t;
        t.c = ct; t.v = NULL; t.cks = qnum; t.cqs = 1; t.vks = 0; t.vqs = 0; t.scale = 1.0;
        fsx_fitpar fp;
        fp.a = a; fp.q = qvals; fp.rdq = NULL; fp.qnum = qnum; fp.mult = mult;
        fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
        fp.factr = 1e+7; fp.pgtol = 1e-5;
        fp.onepass = 0; fp.a0sum = 0.0;
        double xs[2] = {1.0, 0.0}, f;
        int nev;
        fsx_lbfgsb_fit(&t, &fp, xs, &f, &nev);
        out[3 * (size_t)i] = sqrt(f);
        out[3 * (size_t)i + 1] = xs[0];
        out[3 * (size_t)i + 2] = xs[1];
    }
    free(x); free(re); free(im); free(ct);
    return rc;
}
