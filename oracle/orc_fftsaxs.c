/*
 * orc_fftsaxs.c -- oracle restatement of sxs_compute_saxs_scores
 * (src/fftsaxs.c:608-986) and of its static kernels, in the reference's
 * own loop structure: per z the Bessel table and translation matrix, per
 * (z, beta2) the first sums, per (z, beta1, beta2) slab and q the second
 * sums and the transform to the Euler-angle grid, then one fit per
 * requested grid point.
 *
 * TEST INFRASTRUCTURE (checker only; never linked into the product).
 *
 * Differences from the reference that do not change any computed value:
 * poses are bucketed by slab once instead of rescanned per slab,
 * only requested grid points are materialised, slabs of one z run in
 * parallel (OpenMP), and the dense-slab transform (FFTW in the reference,
 * fftsaxs.c:899-900) is an exact-twiddle direct DFT at the requested points
 * (same twiddles, different rounding order than FFTW's plan).
 * Generalised grids (nbeta != L+1 or alpha/gamma grids != 2L+1 points) are
 * evaluated with exact twiddles on signed orders m in [-L, L].
 */
#include "oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { VV = 0, VD = 1, VW = 2, DD = 3, DW = 4, WW = 5 };

static inline int iabs_(int v) { return v < 0 ? -v : v; }
static inline int imax_(int a, int b) { return a > b ? a : b; }

/* fftsaxs.c:27-50 */
static void const_int(const double *r1, const double *i1, const double *r2,
                      const double *i2, const double *s1, const double *t1,
                      const double *s2, const double *t2, int qnum, int L,
                      double *out)
{
    int K = (L + 1) * (L + 1);
    for (int q = 0; q < qnum; q++) {
        double v = 0.0;
        for (int l = 0; l <= L; l++)
            for (int m = -l; m <= l; m++) {
                int id = q * K + l * (l + 1) + m;
                v += r1[id] * r2[id] + i1[id] * i2[id];
                v += s1[id] * s2[id] + t1[id] * t2[id];
            }
        out[q] = v;
    }
}

/* fftsaxs.c:182-243: T^{|m|}_{l l1}(qz) = sum_p (-1)^m i^p d_lm(l1,p) j_p(qz),
 * symmetric in (l, l1); stored [q][m][l][l1] for m >= 0. */
static void t_matrix(double *tre, double *tim, const double *ds,
                     const double *bes, int qnum, int L)
{
    static const double ipr[4] = {1.0, 0.0, -1.0, 0.0};
    static const double ipi[4] = {0.0, 1.0, 0.0, -1.0};
    static const double bin[2] = {1.0, -1.0};
    int K = L + 1, N = 2 * L + 1;
    for (int q = 0; q < qnum; q++) {
        int bs = q * N;
        for (int m = 0; m <= L; m++) {
            double sg = bin[m % 2];
            int t1 = (q * K + m) * K;
            for (int l = m; l <= L; l++)
                for (int l1 = l; l1 <= L; l1++) {
                    double re = 0.0, im = 0.0;
                    size_t di = ((size_t)(l * (l + 1) + m) * K + l1) * N;
                    for (int p = iabs_(l - l1); p <= l + l1; p++) {
                        double v = sg * ds[di + p] * bes[bs + p];
                        re += ipr[p % 4] * v;
                        im += ipi[p % 4] * v;
                    }
                    tre[(t1 + l) * K + l1] = re;
                    tim[(t1 + l) * K + l1] = im;
                    tre[(t1 + l1) * K + l] = re;
                    tim[(t1 + l1) * K + l] = im;
                }
        }
    }
}

static inline int dix(int L, int l, int m, int m1)
{
    return (2 * L + 1) * ((2 * L + 1) * l + (L + m)) + (L + m1);
}

/* fftsaxs.c:251-333: U[q][m][m2][l] = sum_{l1} d^{l1}_{m m2} conj(B_{l1 m2} T_{q|m| l l1}) */
static void first_sum(double *sre, double *sim, const double *tre,
                      const double *tim, const double *dw, const double *Bre,
                      const double *Bim, int qnum, int L)
{
    int K = L + 1, N = 2 * L + 1, KK = K * K, cube = KK * K;
    for (int q = 0; q < qnum; q++) {
        const double *br = Bre + (size_t)q * KK, *bi = Bim + (size_t)q * KK;
        const double *tr = tre + (size_t)q * cube, *ti = tim + (size_t)q * cube;
        for (int m = -L; m <= L; m++) {
            int tx1 = iabs_(m) * K;
            size_t s2 = ((size_t)q * N + m + L) * N;
            for (int m2 = -L; m2 <= L; m2++) {
                size_t s3 = (s2 + m2 + L) * K;
                for (int l = iabs_(m); l <= L; l++) {
                    int tx2 = (tx1 + l) * K;
                    double vr = 0.0, vi = 0.0;
                    for (int l1 = imax_(iabs_(m2), iabs_(m)); l1 <= L; l1++) {
                        double d = dw[dix(L, l1, m, m2)];
                        int b = l1 * (l1 + 1) + m2;
                        double bre = br[b], bim = bi[b];
                        int t3 = tx2 + l1;
                        vr += d * (bre * tr[t3] - bim * ti[t3]);
                        vi -= d * (bre * ti[t3] + bim * tr[t3]);
                    }
                    sre[s3 + l] = vr;
                    sim[s3 + l] = vi;
                }
            }
        }
    }
}

/* fftsaxs.c:416-524: the six receptor-ligand cross sums for one q,
 * written in FFT order (negative orders wrapped to N + m) for the native
 * grid; out: [6][N][N][N] complex interleaved. */
static void second_sum(double *out, const double *const *U, const double *dw,
                       const double *const *Ar, const double *const *Ai,
                       int q, int L)
{
    int K = L + 1, N = 2 * L + 1, KK = K * K;
    size_t fsz = (size_t)N * N * N;
    const double *Vr = Ar[0] + (size_t)q * KK, *Vi = Ai[0] + (size_t)q * KK;
    const double *Dr = Ar[1] + (size_t)q * KK, *Di = Ai[1] + (size_t)q * KK;
    const double *Wr = Ar[2] + (size_t)q * KK, *Wi = Ai[2] + (size_t)q * KK;
    const double *svr = U[0], *svi = U[1], *sdr = U[2], *sdi = U[3], *swr = U[4], *swi = U[5];
    size_t s1 = (size_t)q * N;
    for (int m = -L; m <= L; m++) {
        size_t s2 = (s1 + m + L) * N;
        int o1 = (m < 0 ? N + m : m) * N;
        for (int m1 = -L; m1 <= L; m1++) {
            int o2 = ((m1 < 0 ? N + m1 : m1) + o1) * N;
            for (int m2 = -L; m2 <= L; m2++) {
                size_t s3 = (s2 + m2 + L) * K;
                size_t o3 = (size_t)((m2 < 0 ? N + m2 : m2) + o2);
                int didx = (L + m) * N + (L + m1);
                double r[6] = {0, 0, 0, 0, 0, 0}, im[6] = {0, 0, 0, 0, 0, 0};
                for (int l = imax_(iabs_(m1), iabs_(m)); l < L + 1; l++) {
                    double d = dw[didx + l * N * N];
                    int a = l * (l + 1) + m1;
                    double vr = Vr[a], vi = Vi[a], dr = Dr[a], di = Di[a], wr = Wr[a], wi = Wi[a];
                    size_t s4 = s3 + l;
                    r[0] += d * (vr * svr[s4] - vi * svi[s4]);
                    im[0] += d * (vr * svi[s4] + vi * svr[s4]);
                    r[1] += d * (vr * sdr[s4] + dr * svr[s4] - vi * sdi[s4] - di * svi[s4]);
                    im[1] += d * (vr * sdi[s4] + dr * svi[s4] + vi * sdr[s4] + di * svr[s4]);
                    r[2] += d * (vr * swr[s4] + wr * svr[s4] - vi * swi[s4] - wi * svi[s4]);
                    im[2] += d * (vr * swi[s4] + wr * svi[s4] + vi * swr[s4] + wi * svr[s4]);
                    r[3] += d * (dr * sdr[s4] - di * sdi[s4]);
                    im[3] += d * (dr * sdi[s4] + di * sdr[s4]);
                    r[4] += d * (dr * swr[s4] + wr * sdr[s4] - di * swi[s4] - wi * sdi[s4]);
                    im[4] += d * (dr * swi[s4] + wr * sdi[s4] + di * swr[s4] + wi * sdr[s4]);
                    r[5] += d * (wr * swr[s4] - wi * swi[s4]);
                    im[5] += d * (wr * swi[s4] + wi * swr[s4]);
                }
                for (int k = 0; k < 6; k++) {
                    out[2 * (k * fsz + o3)] = r[k];
                    out[2 * (k * fsz + o3) + 1] = im[k];
                }
            }
        }
    }
}

/* fftsaxs.c:529-606 (simple_ft) with a caller-chosen twiddle table: real
 * part of the forward 3-D DFT at grid point (a2, g1, g2), accumulated in
 * the reference's m, m1, m2 order. */
static void dft_at_point(const double *S, const double *ere, const double *eim,
                         int N, int a2, int g1, int g2, double *val)
{
    size_t fsz = (size_t)N * N * N;
    for (int k = 0; k < 6; k++) val[k] = 0.0;
    int eid = 0;
    size_t mm = 0;
    for (int m = 0; m < N; m++) {
        for (int m1 = 0; m1 < N; m1++) {
            for (int m2 = 0; m2 < N; m2++) {
                double er = ere[eid % N], ei = eim[eid % N];
                for (int k = 0; k < 6; k++)
                    val[k] += er * S[2 * (k * fsz + mm)] - ei * S[2 * (k * fsz + mm) + 1];
                eid += g2;
                mm++;
            }
            eid += g1;
        }
        eid += a2;
    }
}

/* generalised grid: S stored in FFT order of N = 2L+1 (negative m wrapped),
 * transform at angles 2 pi a2/na, 2 pi g1/ng, 2 pi g2/ng on signed m. */
static void dft_general(const double *S, int L, int na, int ng, int a2, int g1,
                        int g2, double *val)
{
    int N = 2 * L + 1;
    size_t fsz = (size_t)N * N * N;
    const double tp = 2.0 * ORC_PI_EXACT;
    for (int k = 0; k < 6; k++) val[k] = 0.0;
    for (int m = -L; m <= L; m++) {
        double pa = tp * (double)(((long)m * a2 % na + na) % na) / na;
        for (int m1 = -L; m1 <= L; m1++) {
            double p1 = tp * (double)(((long)m1 * g1 % ng + ng) % ng) / ng;
            for (int m2 = -L; m2 <= L; m2++) {
                double p2 = tp * (double)(((long)m2 * g2 % ng + ng) % ng) / ng;
                double ph = pa + p1 + p2;
                double er = cos(ph), ei = -sin(ph);
                size_t mm = ((size_t)(m < 0 ? N + m : m) * N + (m1 < 0 ? N + m1 : m1)) * N
                          + (m2 < 0 ? N + m2 : m2);
                for (int k = 0; k < 6; k++)
                    val[k] += er * S[2 * (k * fsz + mm)] - ei * S[2 * (k * fsz + mm) + 1];
            }
        }
    }
}

typedef struct { int64_t key; int pose; } keyed;
static int cmp_keyed(const void *a, const void *b)
{
    const keyed *x = (const keyed *)a, *y = (const keyed *)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->pose < y->pose ? -1 : (x->pose > y->pose);
}

void orc_compute_saxs_scores(double *scores, double *c1s, double *c2s,
                             const int64_t *index_list, int nout,
                             const double *A_re, const double *A_im,
                             const double *B_re, const double *B_im,
                             const orc_fit_params *p,
                             const double *qvals, int qnum,
                             const double *zvals, int znum,
                             const orc_grid *grid, int skip,
                             double *cross_out, int *twiddle_out)
{
    int L = grid->L, N = 2 * L + 1, K = L + 1, KK = K * K;
    int nb = grid->nbeta, na = grid->na, ng = grid->ng;
    int native = (na == N && ng == N);
    int thr = grid->dense_threshold > 0 ? grid->dense_threshold : 30;
    size_t blk = (size_t)qnum * KK;
    double pi = orc_get_pi();

    const double *Ar[3] = {A_re, A_re + blk, A_re + 2 * blk};
    const double *Ai[3] = {A_im, A_im + blk, A_im + 2 * blk};
    const double *Br[3] = {B_re, B_re + blk, B_re + 2 * blk};
    const double *Bi[3] = {B_im, B_im + blk, B_im + 2 * blk};

    /* fftsaxs.c:638-643 */
    static const int pr[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};
    double *cint = (double *)calloc((size_t)6 * qnum, sizeof(double));
    for (int k = 0; k < 6; k++)
        const_int(Ar[pr[k][0]], Ai[pr[k][0]], Ar[pr[k][1]], Ai[pr[k][1]],
                  Br[pr[k][0]], Bi[pr[k][0]], Br[pr[k][1]], Bi[pr[k][1]],
                  qnum, L, cint + (size_t)k * qnum);

    /* tables, fftsaxs.c:632-708 */
    double beta_step = (pi - 0.0) / (nb - 1);
    size_t dsz = (size_t)K * N * N;
    double *dw = (double *)malloc(sizeof(double) * dsz * nb);
    for (int b = 0; b < nb; b++) orc_wigner_d(L, 0.0 + b * beta_step, dw + b * dsz);
    double *ds = (double *)malloc(sizeof(double) * (size_t)K * K * K * N);
    orc_d_symb(L, ds);

    double step = 2 * pi / N;
    double *etr = (double *)malloc(sizeof(double) * N), *eti = (double *)malloc(sizeof(double) * N);
    double *eer = (double *)malloc(sizeof(double) * N), *eei = (double *)malloc(sizeof(double) * N);
    for (int i = 0; i < N; i++) {
        etr[i] = cos(i * step);
        eti[i] = -sin(i * step);
        eer[i] = cos(2.0 * ORC_PI_EXACT * i / N);
        eei[i] = -sin(2.0 * ORC_PI_EXACT * i / N);
    }

    /* bucket poses by slab (z, b1, b2) once */
    int64_t pts = (int64_t)na * ng * ng;
    int64_t size5d = pts * nb * nb;
    keyed *kp = (keyed *)malloc(sizeof(keyed) * (size_t)(nout > 0 ? nout : 1));
    int nk = 0;
    for (int i = 0; i < nout; i++) {
        int64_t id = index_list[i];
        if (id < 0) continue;
        if (id / size5d >= znum) continue;
        kp[nk].key = id;
        kp[nk].pose = i;
        nk++;
    }
    qsort(kp, (size_t)nk, sizeof(keyed), cmp_keyed);

    size_t s1sz = (size_t)qnum * N * N * K;
    double *bes = (double *)malloc(sizeof(double) * (size_t)qnum * N);
    double *tre = (double *)calloc((size_t)qnum * K * K * K, sizeof(double));
    double *tim = (double *)calloc((size_t)qnum * K * K * K, sizeof(double));
    double **U = (double **)calloc((size_t)nb, sizeof(double *));

    int i0 = 0;
    while (i0 < nk) {
        int z = (int)(kp[i0].key / size5d);
        int i1 = i0;
        while (i1 < nk && kp[i1].key / size5d == z) i1++;
        double zv = zvals[z];
        for (int q = 0; q < qnum; q++)
            for (int pp = 0; pp < N; pp++) bes[q * N + pp] = orc_sbessel(pp, zv * qvals[q]);
        t_matrix(tre, tim, ds, bes, qnum, L);

        /* active beta2 for this z */
        char *act2 = (char *)calloc((size_t)nb, 1);
        for (int i = i0; i < i1; i++) act2[(kp[i].key % size5d) / pts % nb] = 1;
        for (int b2 = 0; b2 < nb; b2++) {
            if (!act2[b2]) continue;
            if (!U[b2]) U[b2] = (double *)calloc(6 * s1sz, sizeof(double));
            for (int X = 0; X < 3; X++)
                first_sum(U[b2] + (2 * X) * s1sz, U[b2] + (2 * X + 1) * s1sz, tre, tim,
                          dw + b2 * dsz, Br[X], Bi[X], qnum, L);
        }
        free(act2);

        /* slab boundaries within [i0, i1) */
        int nslab = 0;
        int *sb = (int *)malloc(sizeof(int) * (size_t)(i1 - i0 + 1));
        for (int i = i0; i < i1; i++)
            if (i == i0 || kp[i].key / pts != kp[i - 1].key / pts) sb[nslab++] = i;
        sb[nslab] = i1;

#ifdef _OPENMP
        int nth = grid->nthreads > 0 ? grid->nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nth)
#endif
        for (int s = 0; s < nslab; s++) {
            int ja = sb[s], jb = sb[s + 1];
            int64_t slab = (kp[ja].key % size5d) / pts;
            int b1 = (int)(slab / nb), b2 = (int)(slab % nb);
            int nprof = jb - ja;
            int exact = (skip == 0) || nprof >= thr;
            /* distinct points in ascending order */
            int npt = 0;
            int64_t *ptid = (int64_t *)malloc(sizeof(int64_t) * (size_t)nprof);
            int *pmap = (int *)malloc(sizeof(int) * (size_t)nprof);
            for (int j = ja; j < jb; j++) {
                int64_t pt = kp[j].key % pts;
                if (npt == 0 || ptid[npt - 1] != pt) ptid[npt++] = pt;
                pmap[j - ja] = npt - 1;
            }
            double *ct = (double *)malloc(sizeof(double) * 6 * (size_t)qnum * npt);
            for (int t = 0; t < npt; t++)
                for (int q = 0; q < qnum; q++) {
                    double *c = ct + (size_t)t * 6 * qnum;
                    c[VV * qnum + q] = cint[VV * qnum + q];
                    c[VD * qnum + q] = cint[VD * qnum + q] * 2.0;
                    c[VW * qnum + q] = cint[VW * qnum + q] * 2.0;
                    c[DD * qnum + q] = cint[DD * qnum + q];
                    c[DW * qnum + q] = cint[DW * qnum + q] * 2.0;
                    c[WW * qnum + q] = cint[WW * qnum + q];
                }
            double *S = (double *)malloc(sizeof(double) * 2 * 6 * (size_t)N * N * N);
            const double *Uq[6];
            for (int q = 0; q < qnum; q++) {
                for (int k = 0; k < 6; k++) Uq[k] = U[b2] + k * s1sz;
                second_sum(S, Uq, dw + b1 * dsz, Ar, Ai, q, L);
                for (int t = 0; t < npt; t++) {
                    int g2 = (int)(ptid[t] % ng);
                    int g1 = (int)(ptid[t] / ng % ng);
                    int a2 = (int)(ptid[t] / ((int64_t)ng * ng));
                    double val[6];
                    if (native)
                        dft_at_point(S, exact ? eer : etr, exact ? eei : eti, N, a2, g1, g2, val);
                    else
                        dft_general(S, L, na, ng, a2, g1, g2, val);
                    double *c = ct + (size_t)t * 6 * qnum;
                    for (int k = 0; k < 6; k++) c[k * qnum + q] += 2.0 * val[k];
                }
            }
            free(S);
            double *res = (double *)malloc(sizeof(double) * 3 * (size_t)npt);
            double *tmp = (double *)malloc(sizeof(double) * 6 * (size_t)qnum);
            for (int t = 0; t < npt; t++) {
                double *c = ct + (size_t)t * 6 * qnum;
                for (int j = ja; j < jb; j++)
                    if (pmap[j - ja] == t && cross_out)
                        memcpy(cross_out + (size_t)kp[j].pose * 6 * qnum, c,
                               sizeof(double) * 6 * qnum);
                memcpy(tmp, c, sizeof(double) * 6 * qnum);
                orc_fit_one(tmp, qvals, qnum, p, 1, res + 3 * t);
            }
            for (int j = ja; j < jb; j++) {
                int t = pmap[j - ja], ps = kp[j].pose;
                scores[ps] = res[3 * t];
                c1s[ps] = res[3 * t + 1];
                c2s[ps] = res[3 * t + 2];
                if (twiddle_out) twiddle_out[ps] = exact;
            }
            free(tmp); free(res); free(ct); free(ptid); free(pmap);
        }
        free(sb);
        i0 = i1;
    }

    for (int b = 0; b < nb; b++) free(U[b]);
    free(U); free(bes); free(tre); free(tim); free(kp);
    free(etr); free(eti); free(eer); free(eei);
    free(dw); free(ds); free(cint);
}

/* ---------------------------------------------------------------------- */
/* Index helpers, index.c:3-29 and correlate.c:215-251. */
int64_t orc_assemble_index(int z, int b1, int b2, int a2, int g1, int g2,
                           int nbeta, int na, int ng)
{
    return (((((int64_t)z * nbeta + b1) * nbeta + b2) * na + a2) * ng + g1) * ng + g2;
}

void orc_disassemble_index(int64_t index, int nbeta, int na, int ng, int *o)
{
    o[5] = (int)(index % ng); index /= ng;
    o[4] = (int)(index % ng); index /= ng;
    o[3] = (int)(index % na); index /= na;
    o[2] = (int)(index % nbeta); index /= nbeta;
    o[1] = (int)(index % nbeta);
    o[0] = (int)(index / nbeta);
}

int orc_euler2index(double z, double b1, double g1, double a2, double b2,
                    double g2, const double *zvals, int znum, int L,
                    int *out, int max_out)
{
    double pi = orc_get_pi();
    int nbeta = L + 1, N = 2 * L + 1, cnt = 0;
    double b_step = pi / L;
    double a_step = 2.0 * pi / (2 * L + 1);
    for (int j = 0; j < znum; j++) {
        if (zvals[j] > z - 0.001 && zvals[j] < z + 0.001) {
            a2 = 2 * pi - a2;
            g2 = 2 * pi - g2;
            int id = j * nbeta;
            id = (id + (int)(round(b1 / b_step))) * nbeta;
            id = (id + (int)(round(b2 / b_step))) * N;
            id = (id + (int)(round(a2 / a_step))) * N;
            id = (id + (int)(round(g1 / a_step))) * N;
            id = id + (int)(round(g2 / a_step));
            if (cnt < max_out) out[cnt] = id;
            cnt++;
        }
    }
    return cnt;
}
