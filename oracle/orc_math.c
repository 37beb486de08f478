/*
 * orc_math.c -- oracle restatement of the reference's L1 math:
 * spherical Bessel series, Y_lm normalisation, associated Legendre
 * recurrence, Wigner small-d recurrence, Wigner 3j symbols and the
 * d-symbol table of the translation operator.
 *
 * TEST INFRASTRUCTURE (checker only; never linked into the product).
 *
 * Arithmetic is written operation-for-operation in the order the reference
 * evaluates it (build with -ffp-contract=off), because these values feed a
 * parity bar of 1e-10 relative on I(q) and bit-exact pins against the
 * reference-built oracle/_ref objects.
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

static double g_pi = ORC_PI_TRUNC;

void orc_set_pi(double pi) { g_pi = pi; }
double orc_get_pi(void) { return g_pi; }

/* (2l+1)!! as the reference builds it: product of odd integers in
 * increasing order (sfbessel.c:21-37). */
static double orc_odd_factorial(int x)
{
    double res = 1.0;
    if (x <= 0)
        return 1.0;
    for (int i = 1; i <= x; i += 2)
        res *= (double)i;
    return res;
}

/* sfbessel.c:40-60.  Truncated Taylor series of j_l(x), stopped when the
 * last term is <= 1e-5 of the running sum, then multiplied by x^l.  For
 * x <= 0 the reference returns j_0 = 1 and j_l = 0 (l > 0). */
double orc_sbessel(int l, double x)
{
    if (x > 0.0) {
        double tl1 = 2.0 * (double)l + 1.0;
        double term = 1.0 / orc_odd_factorial((int)tl1);
        double sum = term;
        for (int k = 1; fabs(term / sum) > 0.00001; k++) {
            double num = (-1.0) * x * x;
            double den = 2.0 * (double)k * (2.0 * (double)k + tl1);
            term *= num / den;
            sum += term;
        }
        return sum * pow(x, l);
    }
    return (l == 0) ? 1.0 : 0.0;
}

/* index_LM, borrowed.h:22-25 */
static inline int lm_tab(int N, int l, int m) { return N * (m + N - 1) + l; }

/* borrowed.c:19-38: sqrt((2l+1)/(4 pi) / ((l-m+1)...(l+m))) with the
 * Condon-Shortley sign on negative m. */
void orc_sph_norm(int N, double *y)
{
    memset(y, 0, sizeof(double) * (size_t)N * (size_t)(2 * N - 1));
    for (int l = 0; l < N; ++l) {
        double f = 1.;
        double base = (2 * l + 1.) / (4. * g_pi);
        for (int m = 0; m <= l; ++m) {
            if (m > 0)
                f *= (l - m + 1) * (l + m);
            double v = sqrt(base / f);
            double s = (m % 2 == 0) ? (1) : (-1);
            y[lm_tab(N, l, m)] = v;
            y[lm_tab(N, l, -m)] = s * v;
        }
    }
}

/* borrowed.c:40-65: P_l^m(x) by the standard three-term recurrences
 * (m = 0 column, diagonal P_m^m with the -(2m-1) y factor, first
 * off-diagonal, then upward in l).  Only m >= 0 entries are written. */
void orc_legendre(int N, double x, double *P)
{
    double y = sqrt(1. - x * x);
    P[lm_tab(N, 0, 0)] = 1;
    if (N == 1)
        return;
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

/* index_d_array, borrowed.h:78-81 */
static inline int dix(int L, int l, int m, int m1)
{
    return (2 * L + 1) * ((2 * L + 1) * l + (L + m)) + (L + m1);
}

/* borrowed.c:243-313.  Wigner small-d d^l_{m m1}(beta): l = 0, 1 written
 * explicitly; for l >= 2 the four edges (m or m1 = +-l) come from the
 * closed product of Kostelec-Rockmore eq. 26, the interior from the
 * two-term recurrence of their eq. 28.  Entries with |m| or |m1| > l stay 0.
 * Requires L >= 1 (the reference writes l = 1 entries unconditionally). */
void orc_wigner_d(int L, double beta, double *d)
{
    memset(d, 0, sizeof(double) * (size_t)(L + 1) * (2 * L + 1) * (2 * L + 1));
    double cb = cos(beta), sb = sin(beta), r2 = sqrt(2.0);
    d[dix(L, 0, 0, 0)] = 1.0;
    d[dix(L, 1, -1, -1)] = 1.0 * (1.0 + cb) / 2.0;
    d[dix(L, 1, -1, 0)] = 1.0 * sb / r2;
    d[dix(L, 1, -1, 1)] = 1.0 * (1.0 - cb) / 2.0;
    d[dix(L, 1, 0, -1)] = -1.0 * sb / r2;
    d[dix(L, 1, 0, 0)] = 1.0 * cb;
    d[dix(L, 1, 0, 1)] = 1.0 * sb / r2;
    d[dix(L, 1, 1, -1)] = 1.0 * (1.0 - cb) / 2.0;
    d[dix(L, 1, 1, 0)] = -1.0 * sb / r2;
    d[dix(L, 1, 1, 1)] = 1.0 * (1.0 + cb) / 2.0;

    double ch = cos(beta / 2.0), sh = sin(beta / 2.0);
    for (int l = 2; l <= L; ++l) {
        for (int m = -l; m <= l; ++m) {
            double e1 = 1, e2 = 1, e3 = 1, e4 = 1;
            for (int k = 1; k <= 2 * l; ++k) {
                int up = (k <= l + m), dn = (k <= l - m);
                double f = sqrt(k / ((up ? k : 1.0) * (dn ? k : 1.0)));
                e1 *= f * (up ? ch : 1.0) * (dn ? -sh : 1.0);
                e2 *= f * (dn ? ch : 1.0) * (up ? sh : 1.0);
                e3 *= f * (up ? ch : 1.0) * (dn ? sh : 1.0);
                e4 *= f * (dn ? ch : 1.0) * (up ? -sh : 1.0);
            }
            d[dix(L, l, l, m)] = e1;
            d[dix(L, l, -l, m)] = e2;
            d[dix(L, l, m, l)] = e3;
            d[dix(L, l, m, -l)] = e4;

            if (m <= -l || m >= l)
                continue;
            int j = l - 1;
            for (int m1 = -l + 1; m1 < l; ++m1) {
                double v = 1.0;
                v *= (j + 1) * (2.0 * j + 1)
                     / sqrt(((j + 1) * (j + 1) - m * m) * ((j + 1) * (j + 1) - m1 * m1));
                v *= (cb - (double)(m * m1) / (double)(j * (j + 1))) * d[dix(L, j, m, m1)];
                double w = -1.0;
                w *= sqrt((j * j - m * m) * (j * j - m1 * m1));
                w *= (double)(j + 1.0) * d[dix(L, j - 1, m, m1)] / (double)j;
                w /= sqrt(((j + 1.0) * (j + 1.0) - (double)(m * m))
                          * ((j + 1.0) * (j + 1.0) - (double)(m1 * m1)));
                v += w;
                d[dix(L, l, m, m1)] = v;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* Wigner 3j by the Racah formula (borrowed.c:86-222), evaluated exactly:
 * the alternating sum is accumulated as an exact rational with small
 * unsigned big integers (Horner form over consecutive-term ratios), the
 * square-root prefactor and the final quotient in binary128.  The result is
 * truncated toward zero to double, as mpf_get_d does in the reference.
 * Exact zeros stay exact (the reference's GMP sum leaves residues of
 * ~1e-39 where the parity selection rule makes the value zero). */

#define BN_LIMBS 96
typedef struct { uint32_t w[BN_LIMBS]; int n; } bn_t; /* magnitude, little endian */

static void bn_set(bn_t *a, uint32_t v) { a->w[0] = v; a->n = v ? 1 : 0; }
static void bn_mul_small(bn_t *a, uint32_t v)
{
    uint64_t c = 0;
    for (int i = 0; i < a->n; i++) {
        uint64_t t = (uint64_t)a->w[i] * v + c;
        a->w[i] = (uint32_t)t;
        c = t >> 32;
    }
    if (c) {
        if (a->n >= BN_LIMBS) abort();
        a->w[a->n++] = (uint32_t)c;
    }
    if (v == 0) a->n = 0;
}
static int bn_cmp(const bn_t *a, const bn_t *b)
{
    if (a->n != b->n) return a->n < b->n ? -1 : 1;
    for (int i = a->n - 1; i >= 0; i--)
        if (a->w[i] != b->w[i]) return a->w[i] < b->w[i] ? -1 : 1;
    return 0;
}
static void bn_add(bn_t *r, const bn_t *a, const bn_t *b) /* r may alias */
{
    int n = a->n > b->n ? a->n : b->n;
    uint64_t c = 0;
    for (int i = 0; i < n; i++) {
        uint64_t t = (uint64_t)(i < a->n ? a->w[i] : 0) + (i < b->n ? b->w[i] : 0) + c;
        r->w[i] = (uint32_t)t;
        c = t >> 32;
    }
    r->n = n;
    if (c) { if (n >= BN_LIMBS) abort(); r->w[r->n++] = (uint32_t)c; }
}
static void bn_sub(bn_t *r, const bn_t *a, const bn_t *b) /* a >= b */
{
    int64_t br = 0;
    for (int i = 0; i < a->n; i++) {
        int64_t t = (int64_t)a->w[i] - (i < b->n ? b->w[i] : 0) - br;
        br = t < 0;
        r->w[i] = (uint32_t)(t + (br ? ((int64_t)1 << 32) : 0));
    }
    r->n = a->n;
    while (r->n > 0 && r->w[r->n - 1] == 0) r->n--;
}
static __float128 bn_to_f128(const bn_t *a)
{
    __float128 r = 0;
    for (int i = a->n - 1; i >= 0; i--)
        r = r * (__float128)4294967296.0 + (__float128)a->w[i];
    return r;
}
/* signed value = sign * mag */
typedef struct { bn_t mag; int sign; } sbn_t;
static void sbn_addmul(sbn_t *r, const sbn_t *a, int sa, const sbn_t *b, int sb_)
{   /* r = sa*a + sb*b where a, b already multiplied as needed */
    int s1 = a->sign * sa, s2 = b->sign * sb_;
    if (a->mag.n == 0) { r->mag = b->mag; r->sign = s2; return; }
    if (b->mag.n == 0) { r->mag = a->mag; r->sign = s1; return; }
    if (s1 == s2) { bn_add(&r->mag, &a->mag, &b->mag); r->sign = s1; return; }
    int c = bn_cmp(&a->mag, &b->mag);
    if (c == 0) { r->mag.n = 0; r->sign = 1; return; }
    if (c > 0) { bn_sub(&r->mag, &a->mag, &b->mag); r->sign = s1; }
    else { bn_sub(&r->mag, &b->mag, &a->mag); r->sign = s2; }
}

static __float128 f128_fact(int n)
{
    __float128 r = 1;
    for (int i = 2; i <= n; i++) r *= i;
    return r;
}

static __float128 f128_sqrt(__float128 v)
{
    if (v <= 0) return 0;
    __float128 s = (__float128)sqrtl((long double)v);
    s = (s + v / s) * (__float128)0.5;
    s = (s + v / s) * (__float128)0.5;
    return s;
}

/* truncate a positive binary128 to the largest double <= it */
static double f128_trunc_double(__float128 v)
{
    double d = (double)v;
    if ((__float128)d > v) d = nextafter(d, 0.0);
    return d;
}

double orc_w3j(int j1, int j2, int j3, int m1, int m2, int m3)
{
    if (m1 + m2 + m3 != 0) return 0.0;
    if (j3 > j1 + j2 || j3 < abs(j1 - j2)) return 0.0;
    if (abs(m1) > j1 || abs(m2) > j2 || abs(m3) > j3) return 0.0;
    int t1 = j2 - m1 - j3, t2 = j1 + m2 - j3, t3 = j1 + j2 - j3;
    int t4 = j1 - m1, t5 = j2 + m2;
    int tmin = 0;
    if (t1 > tmin) tmin = t1;
    if (t2 > tmin) tmin = t2;
    int tmax = t3;
    if (t4 < tmax) tmax = t4;
    if (t5 < tmax) tmax = t5;
    if (tmin > tmax) return 0.0;

    /* sum_{t} (-1)^t / P_t = (-1)^tmin / P_tmin * X_0 where
     * X_k = 1 - (a_k/b_k) X_{k+1}, a_k/b_k = |term_{t+1}/term_t| at t = tmin+k,
     * a_k = (t3-t)(t4-t)(t5-t), b_k = (t+1)(t+1-t1)(t+1-t2).  X = p/q exactly. */
    sbn_t p; bn_t q;
    bn_set(&p.mag, 1); p.sign = 1; bn_set(&q, 1);
    for (int t = tmax - 1; t >= tmin; t--) {
        uint64_t a = (uint64_t)(t3 - t) * (uint64_t)(t4 - t) * (uint64_t)(t5 - t);
        uint64_t b = (uint64_t)(t + 1) * (uint64_t)(t + 1 - t1) * (uint64_t)(t + 1 - t2);
        if (a > 0xffffffffu || b > 0xffffffffu) abort();
        /* p' = q*b - a*p ; q' = q*b */
        sbn_t qb, ap;
        qb.mag = q; qb.sign = 1; bn_mul_small(&qb.mag, (uint32_t)b);
        ap = p; bn_mul_small(&ap.mag, (uint32_t)a);
        sbn_addmul(&p, &qb, 1, &ap, -1);
        bn_mul_small(&q, (uint32_t)b);
    }
    if (p.mag.n == 0) return 0.0;

    __float128 Pt = f128_fact(tmin) * f128_fact(tmin - t1) * f128_fact(tmin - t2)
                  * f128_fact(t3 - tmin) * f128_fact(t4 - tmin) * f128_fact(t5 - tmin);
    __float128 delta = f128_fact(j1 + j2 - j3) * f128_fact(j1 - j2 + j3)
                     * f128_fact(-j1 + j2 + j3) / f128_fact(j1 + j2 + j3 + 1)
                     * f128_fact(j1 + m1) * f128_fact(j1 - m1) * f128_fact(j2 + m2)
                     * f128_fact(j2 - m2) * f128_fact(j3 + m3) * f128_fact(j3 - m3);
    __float128 mag = f128_sqrt(delta) * bn_to_f128(&p.mag) / (bn_to_f128(&q) * Pt);
    int sign = p.sign * ((tmin % 2 == 0) ? 1 : -1) * (((j1 - j2 - m3) % 2 == 0) ? 1 : -1);
    double d = f128_trunc_double(mag);
    return sign < 0 ? -d : d;
}

/* fftsaxs.c:138-174: d_lm(l1, p) = (2p+1) sqrt((2l+1)(2l1+1))
 *   * (l p l1; 0 0 0) * (l p l1; -m 0 m), stored at
 *   [(l(l+1)+m)(L+1) + l1](2L+1) + p.  Entries outside the loops stay 0. */
void orc_d_symb(int L, double *ds)
{
    size_t n = (size_t)(L + 1) * (L + 1) * (L + 1) * (2 * L + 1);
    memset(ds, 0, n * sizeof(double));
    for (int l = 0; l < L + 1; l++) {
        double k1 = 2 * l + 1;
        for (int l1 = 0; l1 < L + 1; l1++) {
            double k2 = sqrt((2 * l1 + 1) * k1);
            for (int p = abs(l - l1); p <= l + l1; p++) {
                double k3 = (2 * p + 1) * k2 * orc_w3j(l, p, l1, 0, 0, 0);
                int mm = l < l1 ? l : l1;
                for (int m = -mm; m <= mm; m++) {
                    double k4 = orc_w3j(l, p, l1, -m, 0, m);
                    size_t idx = ((size_t)(l * (l + 1) + m) * (L + 1) + l1) * (2 * L + 1) + p;
                    ds[idx] = k3 * k4;
                }
            }
        }
    }
}
