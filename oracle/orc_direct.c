/*
 * orc_direct.c -- per-pose direct evaluation of the six cross-term
 * profiles of one rigid dimer pose, used to pin the slab/DFT restatement
 * in orc_fftsaxs.c against the underlying identity
 *
 *   F^{XY}(a2,g1,g2) = sum_m e^{-i m alpha2} sum_l  R^X_{ml}(beta1,gamma1)
 *                                               * U^Y_{ml}(z,beta2,gamma2)
 *   R^X_{ml} = sum_{m1} e^{-i m1 gamma1} d^l_{m m1}(beta1) A^X_{l m1}
 *   U^Y_{ml} = sum_{l1} conj(T^{|m|}_{l l1}(qz))
 *                       sum_{m2} e^{-i m2 gamma2} d^{l1}_{m m2}(beta2) conj(B^Y_{l1 m2})
 *
 * i.e. the receptor rotated by (0, beta1, gamma1) paired, order by order,
 * with the ligand rotated by (., beta2, gamma2) and translated by z, then
 * rotated about the axis by alpha2 (fftsaxs.c:251-333, 416-524, 529-606
 * summed in a different order).  Angles are continuous; twiddles exact.
 *
 * TEST INFRASTRUCTURE (checker only; never linked into the product).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int dix(int L, int l, int m, int m1)
{
    return (2 * L + 1) * ((2 * L + 1) * l + (L + m)) + (L + m1);
}

void orc_direct_pose_cross_terms(const double *A_re, const double *A_im,
                                 const double *B_re, const double *B_im,
                                 int L, const double *qvals, int qnum,
                                 double z, double beta1, double gamma1,
                                 double alpha2, double beta2, double gamma2,
                                 double *out)
{
    int N = 2 * L + 1, K = L + 1, KK = K * K;
    size_t blk = (size_t)qnum * KK;
    double *d1 = (double *)malloc(sizeof(double) * (size_t)K * N * N);
    double *d2 = (double *)malloc(sizeof(double) * (size_t)K * N * N);
    double *ds = (double *)malloc(sizeof(double) * (size_t)K * K * K * N);
    orc_wigner_d(L, beta1, d1);
    orc_wigner_d(L, beta2, d2);
    orc_d_symb(L, ds);
    double *Rr = (double *)calloc((size_t)3 * N * K, sizeof(double));
    double *Ri = (double *)calloc((size_t)3 * N * K, sizeof(double));
    double *Ur = (double *)calloc((size_t)3 * N * K, sizeof(double));
    double *Ui = (double *)calloc((size_t)3 * N * K, sizeof(double));
    double *Br_ = (double *)calloc((size_t)N * K, sizeof(double));
    double *Bi_ = (double *)calloc((size_t)N * K, sizeof(double));
    double *Tr = (double *)calloc((size_t)K * K * K, sizeof(double));
    double *Ti = (double *)calloc((size_t)K * K * K, sizeof(double));
    static const int px[6] = {0, 0, 0, 1, 1, 2}, py[6] = {0, 1, 2, 1, 2, 2};

    for (int q = 0; q < qnum; q++) {
        /* const part: sum_lm Re(A^X conj A^Y) + Re(B^X conj B^Y), doubled for X != Y */
        for (int k = 0; k < 6; k++) {
            double s = 0.0;
            for (int id = 0; id < KK; id++) {
                size_t ox = px[k] * blk + (size_t)q * KK + id, oy = py[k] * blk + (size_t)q * KK + id;
                s += A_re[ox] * A_re[oy] + A_im[ox] * A_im[oy];
                s += B_re[ox] * B_re[oy] + B_im[ox] * B_im[oy];
            }
            out[k * qnum + q] = (px[k] == py[k]) ? s : 2.0 * s;
        }
        /* translation operator */
        double x = z * qvals[q];
        for (int m = 0; m <= L; m++)
            for (int l = m; l <= L; l++)
                for (int l1 = m; l1 <= L; l1++) {
                    double re = 0.0, im = 0.0;
                    for (int p = abs(l - l1); p <= l + l1; p++) {
                        double v = ((m % 2) ? -1.0 : 1.0)
                                 * ds[((size_t)(l * (l + 1) + m) * K + l1) * N + p]
                                 * orc_sbessel(p, x);
                        if (p % 4 == 0) re += v;
                        else if (p % 4 == 1) im += v;
                        else if (p % 4 == 2) re -= v;
                        else im -= v;
                    }
                    Tr[(m * K + l) * K + l1] = re;
                    Ti[(m * K + l) * K + l1] = im;
                }
        for (int X = 0; X < 3; X++) {
            const double *ar = A_re + X * blk + (size_t)q * KK, *ai = A_im + X * blk + (size_t)q * KK;
            const double *br = B_re + X * blk + (size_t)q * KK, *bi = B_im + X * blk + (size_t)q * KK;
            for (int m = -L; m <= L; m++)
                for (int l = abs(m); l <= L; l++) {
                    double rr = 0, ri = 0, sr = 0, si = 0;
                    for (int m1 = -l; m1 <= l; m1++) {
                        double c = cos(m1 * gamma1), s = -sin(m1 * gamma1);
                        double d = d1[dix(L, l, m, m1)];
                        double vr = d * ar[l * (l + 1) + m1], vi = d * ai[l * (l + 1) + m1];
                        rr += c * vr - s * vi;
                        ri += c * vi + s * vr;
                        double c2 = cos(m1 * gamma2), s2 = -sin(m1 * gamma2);
                        double e = d2[dix(L, l, m, m1)];
                        double wr = e * br[l * (l + 1) + m1], wi = -e * bi[l * (l + 1) + m1];
                        sr += c2 * wr - s2 * wi;
                        si += c2 * wi + s2 * wr;
                    }
                    Rr[(X * N + m + L) * K + l] = rr;
                    Ri[(X * N + m + L) * K + l] = ri;
                    Br_[(m + L) * K + l] = sr;
                    Bi_[(m + L) * K + l] = si;
                }
            for (int m = -L; m <= L; m++) {
                int am = abs(m);
                for (int l = am; l <= L; l++) {
                    double ur = 0, ui = 0;
                    for (int l1 = am; l1 <= L; l1++) {
                        double tr = Tr[(am * K + l) * K + l1], ti = -Ti[(am * K + l) * K + l1];
                        double br2 = Br_[(m + L) * K + l1], bi2 = Bi_[(m + L) * K + l1];
                        ur += tr * br2 - ti * bi2;
                        ui += tr * bi2 + ti * br2;
                    }
                    Ur[(X * N + m + L) * K + l] = ur;
                    Ui[(X * N + m + L) * K + l] = ui;
                }
            }
        }
        for (int k = 0; k < 6; k++) {
            int X = px[k], Y = py[k];
            double Fr = 0.0;
            for (int m = -L; m <= L; m++) {
                double gr = 0, gi = 0;
                for (int l = abs(m); l <= L; l++) {
                    size_t ix = (X * N + m + L) * K + l, iy = (Y * N + m + L) * K + l;
                    gr += Rr[ix] * Ur[iy] - Ri[ix] * Ui[iy];
                    gi += Rr[ix] * Ui[iy] + Ri[ix] * Ur[iy];
                    if (X != Y) {
                        gr += Rr[iy] * Ur[ix] - Ri[iy] * Ui[ix];
                        gi += Rr[iy] * Ui[ix] + Ri[iy] * Ur[ix];
                    }
                }
                double c = cos(m * alpha2), s = -sin(m * alpha2);
                Fr += c * gr - s * gi;
            }
            out[k * qnum + q] += 2.0 * Fr;
        }
    }
    free(d1); free(d2); free(ds); free(Rr); free(Ri); free(Ur); free(Ui);
    free(Br_); free(Bi_); free(Tr); free(Ti);
}
