/*
 * fsx_spf.c -- host side of the spherical-harmonic amplitude expansion of
 * an atom group (src/pdb2spf.c:24-152):
 *   A^X_lm(q) = 4 pi i^l sum_atoms f^X_a j_l(q r_a) conj(Y_lm(theta_a, phi_a))
 * with f^V, f^D the scalar vacuum / excluded-volume form factors and
 * f^W = zero_ff(OH2) * SASA fraction.  Per atom, this side evaluates the
 * spherical angles (cart2sph, pdb2spf.c:9-22, phi wrapped with the reference
 * pi), the normalised associated Legendre values (borrowed.c:19-65) and
 * conj(Y_lm); the Bessel table and the sums over atoms run on the GPU.
 */
#include "../fsx_internal.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

static inline int lm_tab(int N, int l, int m) { return N * (m + N - 1) + l; }

/* borrowed.c:19-38 */
static void sph_norm(int N, double *y)
{
    for (int l = 0; l < N; ++l) {
        double f = 1.;
        double base = (2 * l + 1.) / (4. * FSX_PI_REFERENCE);
        for (int m = 0; m <= l; ++m) {
            if (m > 0) f *= (l - m + 1) * (l + m);
            double v = sqrt(base / f);
            y[lm_tab(N, l, m)] = v;
            y[lm_tab(N, l, -m)] = ((m % 2 == 0) ? 1 : -1) * v;
        }
    }
}

/* borrowed.c:40-65 */
static void legendre(int N, double x, double *P)
{
    double y = sqrt(1. - x * x);
    P[lm_tab(N, 0, 0)] = 1;
    if (N == 1) return;
    P[lm_tab(N, 1, 0)] = x;
    for (int l = 2; l < N; ++l)
        P[lm_tab(N, l, 0)] = (2. * l - 1.) / l * x * P[lm_tab(N, l - 1, 0)]
                           - (l - 1.) / l * P[lm_tab(N, l - 2, 0)];
    for (int m = 1; m < N - 1; ++m) {
        P[lm_tab(N, m, m)] = -1. * (2. * m - 1.) * y * P[lm_tab(N, m - 1, m - 1)];
        P[lm_tab(N, m + 1, m)] = (2. * m + 1.) * x * P[lm_tab(N, m, m)];
        for (int l = m + 2; l < N; ++l)
            P[lm_tab(N, l, m)] = (2. * l - 1.) / (l - m) * x * P[lm_tab(N, l - 1, m)]
                               - (l + m - 1.) / (l - m) * P[lm_tab(N, l - 2, m)];
    }
    P[lm_tab(N, N - 1, N - 1)] = -1. * (2. * (N - 1.) - 1.) * y * P[lm_tab(N, N - 2, N - 2)];
}

int fsx_atoms2spf(int natoms, const double *xyz, const double *vac, const double *dum,
                  const double *sa, double h2o_ff, const double *vdw, const double *qvals,
                  int qnum, int L, double *re, double *im, double *rm_out)
{
    if (natoms < 1 || !xyz || !vac || !dum || !vdw || !qvals || qnum < 1 || L < 0 || L > 40 ||
        !re || !im)
        return FSX_ERR_ARG;
    int N = L + 1, K2 = N * N;
    double *r = (double *)malloc(sizeof(double) * natoms);
    double *f = (double *)malloc(sizeof(double) * 3 * (size_t)natoms);
    double *Y = (double *)malloc(sizeof(double) * 2 * (size_t)K2 * natoms);
    double *yn = (double *)calloc((size_t)N * (2 * N - 1), sizeof(double));
    double *P = (double *)calloc((size_t)N * (2 * N - 1), sizeof(double));
    if (!r || !f || !Y || !yn || !P) {
        free(r); free(f); free(Y); free(yn); free(P);
        return FSX_ERR_NOMEM;
    }
    /* saxs_utils.c:19-28 */
    double av = 0.0;
    for (int a = 0; a < natoms; a++) av += vdw[a];
    if (rm_out) *rm_out = av / (double)(size_t)natoms;
    sph_norm(N, yn);
    for (int a = 0; a < natoms; a++) {
        double x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
        double rr = sqrt(x * x + y * y + z * z);
        double th = acos(z / rr), fi;
        if (y > 0.0) fi = acos(x / sqrt(x * x + y * y));
        else fi = -acos(x / sqrt(x * x + y * y)) + 2.0 * FSX_PI_REFERENCE;
        r[a] = rr;
        memset(P, 0, sizeof(double) * (size_t)N * (2 * N - 1));
        legendre(N, cos(th), P);
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++) {
                int i = l * (l + 1) + m;
                double h = yn[lm_tab(N, l, m)] * P[lm_tab(N, l, abs(m))];
                Y[(size_t)i * natoms + a] = h * cos(fi * m);
                Y[(size_t)(K2 + i) * natoms + a] = h * sin(-fi * m);
            }
        f[a] = vac[a];
        f[natoms + a] = dum[a];
        f[2 * natoms + a] = sa ? h2o_ff * sa[a] : 0.0;
    }
    int rc = fsx_hip_atoms2spf(natoms, r, f, Y, qvals, qnum, L, re, im);
    free(r); free(f); free(Y); free(yn); free(P);
    return rc;
}
