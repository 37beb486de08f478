/*
 * fsx_math.c -- host-side special functions and the tables the GPU path
 * uploads: the reference's spherical Bessel series (sfbessel.c:40-60), the
 * Wigner small-d table (borrowed.c:243-313) and the d-symbol table of the
 * translation operator (fftsaxs.c:138-174).
 *
 * The Bessel series (a truncated series whose large-argument values depend on
 * its exact rounding) and the Wigner recurrence feed a 1e-10 parity bar, so
 * their evaluation order is kept alike the reference's; this file is built with
 * FP contraction off.  The 3j symbols are computed here with binary128
 * arithmetic (Racah's sum) and truncated toward zero like the reference's
 * mpf_get_d; the table is built once per L per process.
 */
#include "../fsx_internal.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static double odd_fact(int x)
{
    double r = 1.0;
    for (int i = 1; i <= x; i += 2) r *= (double)i;
    return r;
}

double fsx_sbessel_host(int l, double x)
{
    if (x > 0.0) {
        double tl1 = 2.0 * (double)l + 1.0;
        double term = 1.0 / odd_fact((int)tl1);
        double sum = term;
        for (int k = 1; fabs(term / sum) > 0.00001; k++) {
            double ratio = (-1.0) * x * x / (2.0 * (double)k * (2.0 * (double)k + tl1));
            term *= ratio;
            sum += term;
        }
        return sum * pow(x, l);
    }
    return l == 0 ? 1.0 : 0.0;
}

double sxs_sbessel(int order, double x) { return fsx_sbessel_host(order, x); }

static inline size_t wd_ix(int L, int l, int m, int m1)
{
    return (size_t)(2 * L + 1) * ((size_t)(2 * L + 1) * l + (L + m)) + (L + m1);
}

void fsx_wigner_d(int L, double beta, double *d)
{
    memset(d, 0, sizeof(double) * (size_t)(L + 1) * (2 * L + 1) * (2 * L + 1));
    double cb = cos(beta), sb = sin(beta), s2 = sqrt(2.0);
    d[wd_ix(L, 0, 0, 0)] = 1.0;
    if (L < 1) return;
    d[wd_ix(L, 1, -1, -1)] = 1.0 * (1.0 + cb) / 2.0;
    d[wd_ix(L, 1, -1, 0)] = 1.0 * sb / s2;
    d[wd_ix(L, 1, -1, 1)] = 1.0 * (1.0 - cb) / 2.0;
    d[wd_ix(L, 1, 0, -1)] = -1.0 * sb / s2;
    d[wd_ix(L, 1, 0, 0)] = 1.0 * cb;
    d[wd_ix(L, 1, 0, 1)] = 1.0 * sb / s2;
    d[wd_ix(L, 1, 1, -1)] = 1.0 * (1.0 - cb) / 2.0;
    d[wd_ix(L, 1, 1, 0)] = -1.0 * sb / s2;
    d[wd_ix(L, 1, 1, 1)] = 1.0 * (1.0 + cb) / 2.0;
    double ch = cos(beta / 2.0), sh = sin(beta / 2.0);
    for (int l = 2; l <= L; ++l) {
        for (int m = -l; m <= l; ++m) {
            /* edges m or m1 = +-l: product formula (Kostelec-Rockmore eq. 26) */
            double e[4] = {1, 1, 1, 1};
            for (int k = 1; k <= 2 * l; ++k) {
                int a = (k <= l + m), b = (k <= l - m);
                double f = sqrt(k / ((a ? k : 1.0) * (b ? k : 1.0)));
                e[0] *= f * (a ? ch : 1.0) * (b ? -sh : 1.0);
                e[1] *= f * (b ? ch : 1.0) * (a ? sh : 1.0);
                e[2] *= f * (a ? ch : 1.0) * (b ? sh : 1.0);
                e[3] *= f * (b ? ch : 1.0) * (a ? -sh : 1.0);
            }
            d[wd_ix(L, l, l, m)] = e[0];
            d[wd_ix(L, l, -l, m)] = e[1];
            d[wd_ix(L, l, m, l)] = e[2];
            d[wd_ix(L, l, m, -l)] = e[3];
            if (m == -l || m == l) continue;
            /* interior: three-term recurrence in l (eq. 28) */
            int j = l - 1;
            for (int m1 = 1 - l; m1 <= l - 1; ++m1) {
                double v = 1.0;
                v *= (j + 1) * (2.0 * j + 1)
                     / sqrt(((j + 1) * (j + 1) - m * m) * ((j + 1) * (j + 1) - m1 * m1));
                v *= (cb - (double)(m * m1) / (double)(j * (j + 1))) * d[wd_ix(L, j, m, m1)];
                double u = -1.0;
                u *= sqrt((j * j - m * m) * (j * j - m1 * m1));
                u *= (double)(j + 1.0) * d[wd_ix(L, j - 1, m, m1)] / (double)j;
                u /= sqrt(((j + 1.0) * (j + 1.0) - (double)(m * m))
                          * ((j + 1.0) * (j + 1.0) - (double)(m1 * m1)));
                d[wd_ix(L, l, m, m1)] = v + u;
            }
        }
    }
}

/* ---- Wigner 3j, binary128 Racah sum ------------------------------------ */
typedef __float128 q128;

static q128 *g_fact;      /* n! for n <= g_fact_n */
static int g_fact_n = -1;
static pthread_mutex_t g_lock = PTHREAD_MUTEX_INITIALIZER;

static void fact_table(int n)
{
    if (n <= g_fact_n) return;
    q128 *f = (q128 *)malloc(sizeof(q128) * (size_t)(n + 1));
    f[0] = 1;
    for (int i = 1; i <= n; i++) f[i] = f[i - 1] * (q128)i;
    free(g_fact);
    g_fact = f;
    g_fact_n = n;
}

static q128 q_sqrt(q128 v)
{
    if (v <= 0) return 0;
    q128 s = (q128)sqrtl((long double)v);
    s = (s + v / s) * (q128)0.5;
    s = (s + v / s) * (q128)0.5;
    return s;
}

/* caller holds g_lock and has sized the factorial table */
static double w3j_locked(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0 || j3 > j1 + j2 || j3 < abs(j1 - j2)) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    int t1 = j2 - m1 - j3, t2 = j1 + m2 - j3, t3 = j1 + j2 - j3, t4 = j1 - m1, t5 = j2 + m2;
    int lo = 0, hi = t3;
    if (t1 > lo) lo = t1;
    if (t2 > lo) lo = t2;
    if (t4 < hi) hi = t4;
    if (t5 < hi) hi = t5;
    const q128 *F = g_fact;
    q128 pos = 0, neg = 0;
    for (int t = lo; t <= hi; t++) {
        q128 den = F[t] * F[t - t1] * F[t - t2] * F[t3 - t] * F[t4 - t] * F[t5 - t];
        if (t % 2 == 0) pos += 1 / den;
        else neg += 1 / den;
    }
    q128 s = pos - neg;
    /* odd j1+j2+j3 with all m = 0 vanishes exactly */
    if (m1 == 0 && m2 == 0 && ((j1 + j2 + j3) & 1)) return 0.0;
    q128 tri = F[j1 + j2 - j3] * F[j1 - j2 + j3] * F[-j1 + j2 + j3] / F[j1 + j2 + j3 + 1];
    q128 v = tri * F[j1 + m1] * F[j1 - m1] * F[j2 + m2] * F[j2 - m2] * F[j3 + m3] * F[j3 - m3];
    q128 r = q_sqrt(v) * s;
    if (((j1 - j2 - m3) % 2) != 0) r = -r;
    q128 a = r < 0 ? -r : r;
    double d = (double)a;
    if ((q128)d > a) d = nextafter(d, 0.0);   /* truncate toward zero (mpf_get_d) */
    return r < 0 ? -d : d;
}

double fsx_w3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    pthread_mutex_lock(&g_lock);
    fact_table(j1 + j2 + j3 + 1);
    double r = w3j_locked(j1, j2, j3, m1, m2, m3);
    pthread_mutex_unlock(&g_lock);
    return r;
}

#define FSX_MAX_L 64
static double *g_dsymb[FSX_MAX_L + 1];

const double *fsx_d_symb_cached(int L)
{
    if (L < 0 || L > FSX_MAX_L) return NULL;
    pthread_mutex_lock(&g_lock);
    if (!g_dsymb[L]) {
        fact_table(4 * L + 1);
        int K = L + 1, N = 2 * L + 1;
        double *ds = (double *)calloc((size_t)K * K * K * N, sizeof(double));
        for (int l = 0; l <= L; l++) {
            double k1 = 2 * l + 1;
            for (int l1 = 0; l1 <= L; l1++) {
                double k2 = sqrt((2 * l1 + 1) * k1);
                for (int p = abs(l - l1); p <= l + l1; p++) {
                    double k3 = (2 * p + 1) * k2 * w3j_locked(l, p, l1, 0, 0, 0);
                    int mm = l < l1 ? l : l1;
                    for (int m = -mm; m <= mm; m++) {
                        double k4 = w3j_locked(l, p, l1, -m, 0, m);
                        ds[((size_t)(l * (l + 1) + m) * K + l1) * N + p] = k3 * k4;
                    }
                }
            }
        }
        g_dsymb[L] = ds;
    }
    const double *r = g_dsymb[L];
    pthread_mutex_unlock(&g_lock);
    return r;
}

double *sxs_mkarray(double begin, double end, int qnum)
{
    if (begin > end || begin < 0 || end < 0) {
        fprintf(stderr, "WRONG Q VALUES WHEN CREATING ARRAY");
        exit(EXIT_FAILURE);
    }
    double step = (end - begin) / (qnum - 1);
    double *q = (double *)calloc((size_t)qnum, sizeof(double));
    q[0] = begin;
    for (int i = 1; i < qnum; i++) q[i] = q[i - 1] + step;
    return q;
}

void fsx_d_symb_drop(void)
{
    pthread_mutex_lock(&g_lock);
    for (int L = 0; L <= FSX_MAX_L; L++) {
        free(g_dsymb[L]);
        g_dsymb[L] = NULL;
    }
    pthread_mutex_unlock(&g_lock);
}
