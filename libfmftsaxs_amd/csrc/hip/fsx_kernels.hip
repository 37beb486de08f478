/*
 * fsx_kernels.hip -- gfx950 kernels of the FFT dimer SAXS scoring path and
 * the C-ABI launchers the host side calls (fsx_hip_*).
 *
 * Pipeline for one call (reference: src/fftsaxs.c:608-986):
 *   k_bessel      j_p(z q) by the reference's truncated series        (fftsaxs.c:110-121, sfbessel.c:40-60)
 *   k_tmat        translation operator T^{|m|}_{l l1}(q z)            (fftsaxs.c:182-243)
 *   k_rotate      receptor  R^X_lm(b1,g1,q) = sum_m1 e^{-i m1 g1} d^l_{m m1}(b1) A^X_{l m1}
 *                 ligand    B^Y_lm(b2,g2,q) = sum_m2 e^{-i m2 g2} d^l_{m m2}(b2) conj(B^Y_{l m2})
 *                 (the beta/gamma factors of compute_sum1/compute_sum2 and of the
 *                 3-D DFT, fftsaxs.c:251-333, 416-524, 680, folded per pose axis)
 *   k_count/k_scan/k_scatter   bucket poses by (z, beta2, gamma2)
 *   k_cross       per bucket and q: U^Y_lm = sum_l1 conj(T_{|m| l l1}) B^Y_{l1 m} in LDS,
 *                 then per pose 2 Re sum_lm e^{-i m a2} (R^X U^Y [+ R^Y U^X]) for the six
 *                 cross terms VV,VD,VW,DD,DW,WW (fill_var, fftsaxs.c:87-108)
 *   k_fit         constant terms + peak rescale + bounded L-BFGS-B chi^2 fit per pose
 *                 (fill_const fftsaxs.c:52-85, sxs_fit_params min_saxs.c:153-259)
 *
 * The Bessel series, the translation operator and the fit run with FP
 * contraction off (bit-level agreement with the reference's arithmetic);
 * the contraction kernels use FMA.
 */
#include <hip/hip_runtime.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../fsx_internal.h"
#include "../fsx_lbfgsb.h"

#define FSX_CROSS_THREADS 256
#define FSX_WAVE 64
#define FSX_NXCD 8

struct cplx { double re, im; };

/* ---------------------------------------------------------------------- */
/* device helpers */

__device__ __forceinline__ int lm_l(int lm)
{
    int l = (int)sqrt((double)lm);
    while (l * l > lm) l--;
    while ((l + 1) * (l + 1) <= lm) l++;
    return l;
}

__device__ __forceinline__ int pos_mod(long v, int n)
{
    long r = v % n;
    return (int)(r < 0 ? r + n : r);
}

/* x^n for x > 0, integer n >= 0, correctly rounded: binary exponentiation
 * in double-double (each product exact to ~2^-104 through fma), rounded
 * once at the end.  glibc's pow, which the reference calls, is correctly
 * rounded for these arguments in practice; the device pow is not. */
__device__ __forceinline__ void dd_mul(double ah, double al, double bh, double bl, double *rh,
                                       double *rl)
{
    double p = ah * bh;
    double e = fma(ah, bh, -p);
    e = e + (ah * bl + al * bh);
    double h = p + e;
    *rl = e - (h - p);
    *rh = h;
}

__device__ double pow_int_cr(double x, int n)
{
    double rh = 1.0, rl = 0.0, bh = x, bl = 0.0;
    while (n) {
        if (n & 1) dd_mul(rh, rl, bh, bl, &rh, &rl);
        n >>= 1;
        if (n) dd_mul(bh, bl, bh, bl, &bh, &bl);
    }
    return rh + rl;
}

/* j_l(x) exactly as sfbessel.c:40-60 (contraction off) */
__device__ double dev_sbessel(int l, double x)
{
#pragma clang fp contract(off)
    if (x > 0.0) {
        double tl1 = 2.0 * (double)l + 1.0;
        int n = (int)tl1;
        double f = 1.0;
        for (int i = 1; i <= n; i += 2) f *= (double)i;
        double term = 1.0 / f;
        double sum = term;
        for (int k = 1; fabs(term / sum) > 0.00001; k++) {
            double ratio = (-1.0) * x * x / (2.0 * (double)k * (2.0 * (double)k + tl1));
            term *= ratio;
            sum += term;
        }
        return sum * pow_int_cr(x, l);
    }
    return l == 0 ? 1.0 : 0.0;
}

__global__ void k_bessel(double *bes, const double *zvals, const double *qvals, int znum,
                         int Q, int P)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot = (long)znum * Q * P;
    if (i >= tot) return;
    int p = (int)(i % P);
    int q = (int)((i / P) % Q);
    int z = (int)(i / ((long)P * Q));
    bes[i] = dev_sbessel(p, zvals[z] * qvals[q]);
}

__global__ void k_sbessel_batch(const double *x, int n, int P, double *out)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)n * P) return;
    out[i] = dev_sbessel((int)(i % P), x[i / P]);
}

/* T[z][q][m][l][l1], m >= 0, computed from (min(l,l1), max(l,l1)) so the
 * mirrored entries are bitwise equal, as in fftsaxs.c:213-238. */
__global__ void k_tmat(cplx *T, const double *bes, const double *ds, int znum, int Q, int L)
{
#pragma clang fp contract(off)
    const int K = L + 1, N = 2 * L + 1;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long per = (long)K * K * K;
    if (i >= (long)znum * Q * per) return;
    int r = (int)(i % per);
    long zq = i / per;
    int l1 = r % K, l = (r / K) % K, m = r / (K * K);
    cplx out = {0.0, 0.0};
    if (l >= m && l1 >= m) {
        int lo = l < l1 ? l : l1, hi = l < l1 ? l1 : l;
        const double *b = bes + zq * N;
        const double *d = ds + ((long)(lo * (lo + 1) + m) * K + hi) * N;
        double sg = (m % 2) ? -1.0 : 1.0;
        const double ipr[4] = {1.0, 0.0, -1.0, 0.0}, ipi[4] = {0.0, 1.0, 0.0, -1.0};
        double re = 0.0, im = 0.0;
        for (int p = hi - lo; p <= lo + hi; p++) {
            double v = sg * d[p] * b[p];
            re += ipr[p % 4] * v;
            im += ipi[p % 4] * v;
        }
        out.re = re;
        out.im = im;
    }
    T[i] = out;
}

/* rotated coefficients: out[((b*ng + g)*Q + q)*3 + X][lm] */
__global__ void k_rotate(cplx *out, const double *src_re, const double *src_im,
                         const double *dW, const double *twg, int nbeta, int ng, int Q,
                         int L, int conj_src)
{
    const int K2 = (L + 1) * (L + 1), N = 2 * L + 1;
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long tot = (long)nbeta * ng * Q * 3 * K2;
    if (i >= tot) return;
    int lm = (int)(i % K2);
    long r = i / K2;
    int X = (int)(r % 3); r /= 3;
    int q = (int)(r % Q); r /= Q;
    int g = (int)(r % ng);
    int b = (int)(r / ng);
    int l = lm_l(lm), m = lm - l * (l + 1);
    const double *d = dW + (long)b * (L + 1) * N * N + ((long)l * N + (L + m)) * N + L;
    const double *sr = src_re + ((long)X * Q + q) * K2 + l * (l + 1);
    const double *si = src_im + ((long)X * Q + q) * K2 + l * (l + 1);
    double sgn = conj_src ? -1.0 : 1.0;
    double re = 0.0, im = 0.0;
    for (int m1 = -l; m1 <= l; m1++) {
        int k = pos_mod((long)m1 * g, ng);
        double wr = twg[2 * k], wi = twg[2 * k + 1];
        double dv = d[m1];
        double ar = dv * sr[m1], ai = sgn * dv * si[m1];
        re += wr * ar - wi * ai;
        im += wr * ai + wi * ar;
    }
    out[i].re = re;
    out[i].im = im;
}

/* ---- pose bucketing --------------------------------------------------- */
__device__ __forceinline__ void decode_pose(long long idx, int nbeta, int na, int ng, int *o)
{
    o[5] = (int)(idx % ng); idx /= ng;
    o[4] = (int)(idx % ng); idx /= ng;
    o[3] = (int)(idx % na); idx /= na;
    o[2] = (int)(idx % nbeta); idx /= nbeta;
    o[1] = (int)(idx % nbeta);
    o[0] = (int)(idx / nbeta);
}

__global__ void k_fill_int(int *p, long n, int v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_count(const long long *poses, int n, int znum, int nbeta, int na, int ng,
                        int *keys, int *counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    long long idx = poses[i];
    int key = -1;
    if (idx >= 0) {
        int o[6];
        decode_pose(idx, nbeta, na, ng, o);
        if (o[0] < znum) {
            key = (o[0] * nbeta + o[2]) * ng + o[5];
            atomicAdd(&counts[key], 1);
        }
    }
    keys[i] = key;
}

/* exclusive scan of counts[0..nk) into offs[0..nk]; one block of 1024 */
__global__ void k_scan(const int *counts, int nk, int *offs)
{
    __shared__ int part[1024];
    int t = threadIdx.x;
    int chunk = (nk + 1023) / 1024;
    int b = t * chunk, e = b + chunk < nk ? b + chunk : nk;
    int s = 0;
    for (int i = b; i < e; i++) s += counts[i];
    part[t] = s;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += v;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; i++) {
        offs[i] = run;
        run += counts[i];
    }
    if (t == 1023) offs[nk] = part[1023];
}

__global__ void k_scatter(const int *keys, int n, const int *offs, int *fill, int *sorted)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = keys[i];
    if (k < 0) return;
    int p = offs[k] + atomicAdd(&fill[k], 1);
    sorted[p] = i;
}

/* ---- per-pose cross terms ------------------------------------------------ */
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, FSX_WAVE);
    return v;
}

/* cross2[(q*6 + k)*n + s] = 2 Re F_k(q) of the pose at sorted position s */
__global__ void __launch_bounds__(FSX_CROSS_THREADS)
k_cross(double *cross2, const cplx *Rt, const cplx *Bt, const cplx *T, const double *twa,
        const int *counts, const int *offs, const int *sorted, const long long *poses, int n,
        int nkeys, int Q, int L, int nbeta, int na, int ng, long total)
{
    extern __shared__ cplx smem[];
    const int K = L + 1, K2 = K * K;
    /* XCD-aware remap: the dispatcher deals consecutive blocks round-robin to
     * the 8 XCDs; give each XCD a contiguous range of (q, bucket) items so the
     * T(z, q) and B(beta2, gamma2, q) blocks they share stay in its L2. */
    long nb = gridDim.x;
    long per = nb / FSX_NXCD;
    long w = (long)(blockIdx.x % FSX_NXCD) * per + blockIdx.x / FSX_NXCD;
    if (w >= total) return;
    int key = (int)(w % nkeys);
    int q = (int)(w / nkeys);
    int cnt = counts[key];
    if (cnt == 0) return;
    int start = offs[key];
    int g2 = key % ng;
    int b2 = (key / ng) % nbeta;
    int z = key / (ng * nbeta);
    cplx *sB = smem;            /* [3][K2] */
    cplx *sU = smem + 3 * K2;   /* [3][K2] */
    const cplx *bsrc = Bt + (((long)b2 * ng + g2) * Q + q) * 3 * K2;
    for (int i = threadIdx.x; i < 3 * K2; i += blockDim.x) sB[i] = bsrc[i];
    __syncthreads();
    const cplx *Tq = T + ((long)z * Q + q) * K * K2;
    for (int lm = threadIdx.x; lm < K2; lm += blockDim.x) {
        int l = lm_l(lm), m = lm - l * (l + 1), am = m < 0 ? -m : m;
        const cplx *Tr = Tq + ((long)am * K + l) * K;
        double ur[3] = {0, 0, 0}, ui[3] = {0, 0, 0};
        for (int l1 = am; l1 <= L; l1++) {
            cplx t = Tr[l1];
            int j = l1 * (l1 + 1) + m;
#pragma unroll
            for (int Y = 0; Y < 3; Y++) {
                cplx b = sB[Y * K2 + j];
                /* conj(t) * b */
                ur[Y] += t.re * b.re + t.im * b.im;
                ui[Y] += t.re * b.im - t.im * b.re;
            }
        }
#pragma unroll
        for (int Y = 0; Y < 3; Y++) {
            sU[Y * K2 + lm].re = ur[Y];
            sU[Y * K2 + lm].im = ui[Y];
        }
    }
    __syncthreads();
    const int wave = threadIdx.x / FSX_WAVE, lane = threadIdx.x % FSX_WAVE;
    const int nwaves = blockDim.x / FSX_WAVE;
    for (int pi = wave; pi < cnt; pi += nwaves) {
        int s = start + pi;
        int pose = sorted[s];
        int o[6];
        decode_pose(poses[pose], nbeta, na, ng, o);
        int b1 = o[1], a2 = o[3], g1 = o[4];
        const cplx *R = Rt + (((long)b1 * ng + g1) * Q + q) * 3 * K2;
        double acc[6] = {0, 0, 0, 0, 0, 0};
        for (int lm = lane; lm < K2; lm += FSX_WAVE) {
            int l = lm_l(lm), m = lm - l * (l + 1);
            int kk = pos_mod((long)m * a2, na);
            double wr = twa[2 * kk], wi = twa[2 * kk + 1];
            cplx rv = R[lm], rd = R[K2 + lm], rw = R[2 * K2 + lm];
            cplx uv = sU[lm], ud = sU[K2 + lm], uw = sU[2 * K2 + lm];
            /* twiddled receptor factor: w * R^X */
            double vr = wr * rv.re - wi * rv.im, vi = wr * rv.im + wi * rv.re;
            double dr = wr * rd.re - wi * rd.im, di = wr * rd.im + wi * rd.re;
            double xr = wr * rw.re - wi * rw.im, xi = wr * rw.im + wi * rw.re;
            /* Re of products */
            acc[0] += vr * uv.re - vi * uv.im;
            acc[1] += (vr * ud.re - vi * ud.im) + (dr * uv.re - di * uv.im);
            acc[2] += (vr * uw.re - vi * uw.im) + (xr * uv.re - xi * uv.im);
            acc[3] += dr * ud.re - di * ud.im;
            acc[4] += (dr * uw.re - di * uw.im) + (xr * ud.re - xi * ud.im);
            acc[5] += xr * uw.re - xi * uw.im;
        }
#pragma unroll
        for (int k = 0; k < 6; k++) acc[k] = wave_sum(acc[k]);
        if (lane < 6) {
            double v = acc[0];
#pragma unroll
            for (int k = 1; k < 6; k++) v = (lane == k) ? acc[k] : v;
            cross2[((long)q * 6 + lane) * n + s] = 2.0 * v;
        }
    }
}

/* ---- per-pose fit -------------------------------------------------------- */
__global__ void __launch_bounds__(256)
k_fit(const double *cross2, const double *cint, const int *sorted, const int *offs, int nkeys,
      int n, const double *qvals, int Q, const double *a, double mult, double peak,
      double c1_lo, double c1_hi, double c2_lo, double c2_hi, double *score, double *c1o,
      double *c2o, double *cross_out, int *nev_out)
{
#pragma clang fp contract(off)
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    int nvalid = offs[nkeys];
    if (s >= nvalid) return;
    int pose = sorted[s];
    fsx_ct t;
    t.c = cint; t.cks = Q; t.cqs = 1;
    t.v = cross2 + s; t.vks = n; t.vqs = 6L * n;
    t.scale = 1.0;
    /* min_saxs.c:170-179 */
    double sc = fsx_ct_get(&t, 0, 0) + fsx_ct_get(&t, 3, 0) + fsx_ct_get(&t, 5, 0)
              + fsx_ct_get(&t, 2, 0) - fsx_ct_get(&t, 1, 0) - fsx_ct_get(&t, 4, 0);
    if (cross_out) {
        for (int k = 0; k < 6; k++)
            for (int q = 0; q < Q; q++)
                cross_out[((long)pose * 6 + k) * Q + q] = fsx_ct_get(&t, k, q);
    }
    t.scale = peak / sc;
    fsx_fitpar fp;
    fp.a = a; fp.q = qvals; fp.qnum = Q; fp.mult = mult;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    double x[2] = {1.0, 0.0}, f;
    int nev;
    fsx_lbfgsb_fit(&t, &fp, x, &f, &nev);
    score[pose] = sqrt(f);
    c1o[pose] = x[0];
    c2o[pose] = x[1];
    if (nev_out) nev_out[s] = nev;
}

/* fits of caller-supplied (rescaled) profiles ct[n][6][Q] */
__global__ void __launch_bounds__(256)
k_fit_batch(const double *ct, int n, const double *qvals, int Q, const double *a, double mult,
            double c1_lo, double c1_hi, double c2_lo, double c2_hi, double *out)
{
#pragma clang fp contract(off)
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    fsx_ct t;
    t.c = ct + (long)s * 6 * Q; t.cks = Q; t.cqs = 1;
    t.v = NULL; t.vks = 0; t.vqs = 0; t.scale = 1.0;
    fsx_fitpar fp;
    fp.a = a; fp.q = qvals; fp.qnum = Q; fp.mult = mult;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    double x[2] = {1.0, 0.0}, f;
    int nev;
    fsx_lbfgsb_fit(&t, &fp, x, &f, &nev);
    out[4L * s] = sqrt(f);
    out[4L * s + 1] = x[0];
    out[4L * s + 2] = x[1];
    out[4L * s + 3] = nev;
}

/* ---- atom expansion (pdb2spf.c:24-152) ------------------------------------ */
__global__ void k_atom_bessel(double *J, const double *r, const double *qvals, int natoms, int Q,
                              int K)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)natoms * Q * K) return;
    int a = (int)(i % natoms);
    int l = (int)((i / natoms) % K);
    int q = (int)(i / ((long)natoms * K));
    J[i] = dev_sbessel(l, qvals[q] * r[a]);
}

/* one block per (q, lm): sum over atoms of f^X_a j_l(q r_a) conj(Y_lm), then
 * the 4 pi i^l factor (pdb2spf.c:118-147) */
__global__ void __launch_bounds__(256)
k_atom_sum(double *re, double *im, const double *J, const double *Y, const double *f,
           int natoms, int Q, int L, double four_pi)
{
    __shared__ double red[6][4];
    const int K = L + 1, K2 = K * K;
    int lm = blockIdx.x % K2, q = blockIdx.x / K2;
    int l = lm_l(lm);
    const double *Jq = J + ((long)q * K + l) * natoms;
    const double *yr = Y + (long)lm * natoms, *yi = Y + (long)(K2 + lm) * natoms;
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (int a = threadIdx.x; a < natoms; a += blockDim.x) {
        double j = Jq[a];
        double vr = j * yr[a], vi = j * yi[a];
#pragma unroll
        for (int X = 0; X < 3; X++) {
            double fx = f[(long)X * natoms + a];
            acc[2 * X] += fx * vr;
            acc[2 * X + 1] += fx * vi;
        }
    }
    int wave = threadIdx.x / FSX_WAVE, lane = threadIdx.x % FSX_WAVE;
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double v = wave_sum(acc[k]);
        if (lane == 0) red[k][wave] = v;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        int X = threadIdx.x;
        double vr = 0.0, vi = 0.0;
        for (int w = 0; w < (int)(blockDim.x / FSX_WAVE); w++) {
            vr += red[2 * X][w];
            vi += red[2 * X + 1][w];
        }
        const double cr[4] = {four_pi, 0.0, -four_pi, 0.0}, ci[4] = {0.0, four_pi, 0.0, -four_pi};
        double c_r = cr[l % 4], c_i = ci[l % 4];
        long o = ((long)X * Q + q) * K2 + lm;
        re[o] = c_r * vr - c_i * vi;
        im[o] = c_r * vi + c_i * vr;
    }
}

/* ---------------------------------------------------------------------- */
/* host side: device buffers, stream, launches */

static pthread_mutex_t g_mtx = PTHREAD_MUTEX_INITIALIZER;
static int g_init = 0;
static hipStream_t g_stream;
enum {
    B_BES, B_T, B_RT, B_BT, B_A, B_B, B_DW, B_DS, B_TWA, B_TWG, B_Q, B_Z, B_AM, B_CINT,
    B_POSES, B_KEYS, B_COUNTS, B_OFFS, B_FILL, B_SORTED, B_CROSS2, B_SCORE, B_C1, B_C2,
    B_CROSSOUT, B_NEV, B_FITCT, B_FITOUT, B_BX, B_BOUT, B_AR, B_AF, B_AY, B_AJ, B_AQ,
    B_ARE, B_AIM, B_NBUF
};
static void *g_buf[B_NBUF];
static size_t g_cap[B_NBUF];

#define HIPCHK(x)                                                                     \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "[fsx] %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                              \
            return FSX_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

static int ensure(int slot, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (g_cap[slot] >= bytes) return FSX_OK;
    if (g_buf[slot]) HIPCHK(hipFree(g_buf[slot]));
    g_buf[slot] = NULL;
    g_cap[slot] = 0;
    size_t cap = bytes + bytes / 8;
    if (hipMalloc(&g_buf[slot], cap) != hipSuccess) {
        g_buf[slot] = NULL;
        return FSX_ERR_NOMEM;
    }
    g_cap[slot] = cap;
    return FSX_OK;
}

static int init_once(void)
{
    if (g_init) return FSX_OK;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return FSX_ERR_NODEV;
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    g_init = 1;
    return FSX_OK;
}

extern "C" int fsx_hip_device_count(void)
{
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess) return 0;
    return nd;
}

extern "C" void fsx_hip_release(void)
{
    pthread_mutex_lock(&g_mtx);
    for (int i = 0; i < B_NBUF; i++) {
        if (g_buf[i]) (void)hipFree(g_buf[i]);
        g_buf[i] = NULL;
        g_cap[i] = 0;
    }
    pthread_mutex_unlock(&g_mtx);
}

static inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

#define RC(x)                          \
    do {                               \
        int r_ = (x);                  \
        if (r_ != FSX_OK) return r_;   \
    } while (0)

static int run_locked(const fsx_hip_args *A, fsx_stats *st)
{
    RC(init_once());
    const int L = A->L, K = L + 1, K2 = K * K, N = 2 * L + 1, Q = A->qnum, n = A->n;
    const int nb = A->nbeta, na = A->nalpha, ng = A->ngamma, zn = A->znum;
    const long nkeys = (long)zn * nb * ng;
    if (nkeys > 0x7fffffffL) return FSX_ERR_ARG;
    const size_t dsz = (size_t)K * N * N;
    const size_t spf = (size_t)3 * Q * K2;
    size_t lds = sizeof(cplx) * 6 * K2;
    if (lds > 160 * 1024) return FSX_ERR_ARG;
    hipStream_t s = g_stream;
    RC(ensure(B_BES, sizeof(double) * zn * Q * N));
    RC(ensure(B_T, sizeof(cplx) * zn * Q * (size_t)K * K2));
    RC(ensure(B_RT, sizeof(cplx) * (size_t)nb * ng * spf));
    RC(ensure(B_BT, sizeof(cplx) * (size_t)nb * ng * spf));
    RC(ensure(B_A, sizeof(double) * 2 * spf));
    RC(ensure(B_B, sizeof(double) * 2 * spf));
    RC(ensure(B_DW, sizeof(double) * dsz * nb));
    RC(ensure(B_DS, sizeof(double) * (size_t)K * K2 * N));
    RC(ensure(B_TWA, sizeof(double) * 2 * na));
    RC(ensure(B_TWG, sizeof(double) * 2 * ng));
    RC(ensure(B_Q, sizeof(double) * Q));
    RC(ensure(B_Z, sizeof(double) * zn));
    RC(ensure(B_AM, sizeof(double) * 6 * Q));
    RC(ensure(B_CINT, sizeof(double) * 6 * Q));
    RC(ensure(B_KEYS, sizeof(int) * n));
    RC(ensure(B_COUNTS, sizeof(int) * nkeys));
    RC(ensure(B_OFFS, sizeof(int) * (nkeys + 1)));
    RC(ensure(B_FILL, sizeof(int) * nkeys));
    RC(ensure(B_SORTED, sizeof(int) * n));
    RC(ensure(B_CROSS2, sizeof(double) * 6 * Q * (size_t)n));
    RC(ensure(B_NEV, sizeof(int) * n));
    double *dA = (double *)g_buf[B_A], *dB = (double *)g_buf[B_B];
    HIPCHK(hipMemcpyAsync(dA, A->A_re, sizeof(double) * spf, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dA + spf, A->A_im, sizeof(double) * spf, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dB, A->B_re, sizeof(double) * spf, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(dB + spf, A->B_im, sizeof(double) * spf, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_DW], A->dW, sizeof(double) * dsz * nb, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_DS], A->dsymb, sizeof(double) * K * K2 * N, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_TWA], A->tw_a, sizeof(double) * 2 * na, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_TWG], A->tw_g, sizeof(double) * 2 * ng, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_Q], A->qvals, sizeof(double) * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_Z], A->zvals, sizeof(double) * zn, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_AM], A->a, sizeof(double) * 6 * Q, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(g_buf[B_CINT], A->cint, sizeof(double) * 6 * Q, hipMemcpyHostToDevice, s));

    const long long *dposes;
    if (A->flags & FSX_DEVICE_POSES) {
        dposes = (const long long *)A->poses;
    } else {
        RC(ensure(B_POSES, sizeof(long long) * n));
        HIPCHK(hipMemcpyAsync(g_buf[B_POSES], A->poses, sizeof(long long) * n, hipMemcpyHostToDevice, s));
        dposes = (const long long *)g_buf[B_POSES];
    }
    double *dscore, *dc1, *dc2, *dcross = NULL;
    int dev_out = (A->flags & FSX_DEVICE_OUTPUT) != 0;
    if (dev_out) {
        dscore = A->score; dc1 = A->c1; dc2 = A->c2; dcross = A->cross;
    } else {
        /* device mirrors start from the caller's values so that poses the
         * reference leaves untouched stay untouched */
        RC(ensure(B_SCORE, sizeof(double) * n));
        RC(ensure(B_C1, sizeof(double) * n));
        RC(ensure(B_C2, sizeof(double) * n));
        dscore = (double *)g_buf[B_SCORE]; dc1 = (double *)g_buf[B_C1]; dc2 = (double *)g_buf[B_C2];
        HIPCHK(hipMemcpyAsync(dscore, A->score, sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dc1, A->c1, sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dc2, A->c2, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (A->cross) {
            RC(ensure(B_CROSSOUT, sizeof(double) * 6 * Q * (size_t)n));
            dcross = (double *)g_buf[B_CROSSOUT];
            HIPCHK(hipMemcpyAsync(dcross, A->cross, sizeof(double) * 6 * Q * (size_t)n, hipMemcpyHostToDevice, s));
        }
    }

    hipEvent_t ev[7];
    for (int i = 0; i < 7; i++) HIPCHK(hipEventCreate(&ev[i]));
    HIPCHK(hipEventRecord(ev[0], s));
    /* translation tables */
    long nbes = (long)zn * Q * N;
    hipLaunchKernelGGL(k_bessel, dim3(nblk(nbes, 256)), dim3(256), 0, s, (double *)g_buf[B_BES],
                       (const double *)g_buf[B_Z], (const double *)g_buf[B_Q], zn, Q, N);
    long nt = (long)zn * Q * K * K2;
    hipLaunchKernelGGL(k_tmat, dim3(nblk(nt, 256)), dim3(256), 0, s, (cplx *)g_buf[B_T],
                       (const double *)g_buf[B_BES], (const double *)g_buf[B_DS], zn, Q, L);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[1], s));
    /* rotated receptor and ligand coefficients */
    long nr = (long)nb * ng * (long)spf;
    hipLaunchKernelGGL(k_rotate, dim3(nblk(nr, 256)), dim3(256), 0, s, (cplx *)g_buf[B_RT],
                       (const double *)dA, (const double *)(dA + spf), (const double *)g_buf[B_DW],
                       (const double *)g_buf[B_TWG], nb, ng, Q, L, 0);
    hipLaunchKernelGGL(k_rotate, dim3(nblk(nr, 256)), dim3(256), 0, s, (cplx *)g_buf[B_BT],
                       (const double *)dB, (const double *)(dB + spf), (const double *)g_buf[B_DW],
                       (const double *)g_buf[B_TWG], nb, ng, Q, L, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[2], s));
    /* bucket poses by (z, beta2, gamma2) */
    int *counts = (int *)g_buf[B_COUNTS], *offs = (int *)g_buf[B_OFFS], *fill = (int *)g_buf[B_FILL];
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nkeys, 256)), dim3(256), 0, s, counts, nkeys, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nkeys, 256)), dim3(256), 0, s, fill, nkeys, 0);
    hipLaunchKernelGGL(k_count, dim3(nblk(n, 256)), dim3(256), 0, s, dposes, n, zn, nb, na, ng,
                       (int *)g_buf[B_KEYS], counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const int *)counts, (int)nkeys, offs);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, (const int *)g_buf[B_KEYS],
                       n, (const int *)offs, fill, (int *)g_buf[B_SORTED]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[3], s));
    /* per-pose cross terms */
    long total = (long)Q * nkeys;
    long grid = (total + FSX_NXCD - 1) / FSX_NXCD * FSX_NXCD;
    if (grid > 0x7fffffffL) return FSX_ERR_ARG;
    hipLaunchKernelGGL(k_cross, dim3((unsigned)grid), dim3(FSX_CROSS_THREADS), lds, s,
                       (double *)g_buf[B_CROSS2], (const cplx *)g_buf[B_RT],
                       (const cplx *)g_buf[B_BT], (const cplx *)g_buf[B_T],
                       (const double *)g_buf[B_TWA], (const int *)counts, (const int *)offs,
                       (const int *)g_buf[B_SORTED], dposes, n, (int)nkeys, Q, L, nb, na, ng, total);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[4], s));
    /* per-pose fits */
    hipLaunchKernelGGL(k_fit, dim3(nblk(n, 256)), dim3(256), 0, s, (const double *)g_buf[B_CROSS2],
                       (const double *)g_buf[B_CINT], (const int *)g_buf[B_SORTED],
                       (const int *)offs, (int)nkeys, n, (const double *)g_buf[B_Q], Q,
                       (const double *)g_buf[B_AM], A->mult, A->peak, A->c1_lo, A->c1_hi,
                       A->c2_lo, A->c2_hi, dscore, dc1, dc2, dcross, (int *)g_buf[B_NEV]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(ev[5], s));
    if (!dev_out) {
        HIPCHK(hipMemcpyAsync(A->score, dscore, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(A->c1, dc1, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(A->c2, dc2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (A->cross)
            HIPCHK(hipMemcpyAsync(A->cross, dcross, sizeof(double) * 6 * Q * (size_t)n, hipMemcpyDeviceToHost, s));
    }
    int nvalid = 0;
    HIPCHK(hipMemcpyAsync(&nvalid, offs + nkeys, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPCHK(hipEventRecord(ev[6], s));
    HIPCHK(hipStreamSynchronize(s));
    float ms[6];
    for (int i = 0; i < 6; i++) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
    for (int i = 0; i < 7; i++) HIPCHK(hipEventDestroy(ev[i]));
    st->ms_bessel_t = ms[0];
    st->ms_rotate = ms[1];
    st->ms_group = ms[2];
    st->ms_cross = ms[3];
    st->ms_fit = ms[4];
    st->n_scored = nvalid;
    /* algorithmic traffic of k_cross: per (bucket, q) the ligand block and the
     * translation rows, per pose and q the receptor block and 6 outputs */
    long nonempty_q = 0;
    {
        /* count non-empty buckets from the sorted keys: nvalid poses, and the
         * buckets themselves are bounded by min(nkeys, nvalid) */
        nonempty_q = (long)Q * (nkeys < nvalid ? nkeys : nvalid);
    }
    st->cross_bytes = (double)nonempty_q * (3.0 * K2 * 16.0 + (double)K * K2 * 16.0 / 2.0)
                    + (double)nvalid * Q * (3.0 * K2 * 16.0 + 6.0 * 8.0);
    st->cross_flops = (double)nonempty_q * 3.0 * 8.0 * (K2 * (K + 1) / 3.0)
                    + (double)nvalid * Q * K2 * (12.0 + 6.0 * 4.0 + 9.0 * 2.0);
    return FSX_OK;
}

extern "C" int fsx_hip_run(const fsx_hip_args *A, fsx_stats *st)
{
    pthread_mutex_lock(&g_mtx);
    int rc = run_locked(A, st);
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

extern "C" int fsx_hip_fit_batch(const double *ct, int n, const double *qvals, int Q,
                                 const double *a, double mult, double c1_lo, double c1_hi,
                                 double c2_lo, double c2_hi, double *out)
{
    pthread_mutex_lock(&g_mtx);
    int rc = init_once();
    if (rc == FSX_OK) rc = ensure(B_FITCT, sizeof(double) * 6 * Q * (size_t)n);
    if (rc == FSX_OK) rc = ensure(B_FITOUT, sizeof(double) * 4 * (size_t)n);
    if (rc == FSX_OK) rc = ensure(B_Q, sizeof(double) * Q);
    if (rc == FSX_OK) rc = ensure(B_AM, sizeof(double) * 6 * Q);
    if (rc == FSX_OK) {
        hipStream_t s = g_stream;
        hipError_t e = hipMemcpyAsync(g_buf[B_FITCT], ct, sizeof(double) * 6 * Q * (size_t)n,
                                      hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_Q], qvals, sizeof(double) * Q, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_AM], a, sizeof(double) * 6 * Q, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_fit_batch, dim3(nblk(n, 256)), dim3(256), 0, s,
                               (const double *)g_buf[B_FITCT], n, (const double *)g_buf[B_Q], Q,
                               (const double *)g_buf[B_AM], mult, c1_lo, c1_hi, c2_lo, c2_hi,
                               (double *)g_buf[B_FITOUT]);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out, g_buf[B_FITOUT], sizeof(double) * 4 * (size_t)n, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            fprintf(stderr, "[fsx] fit batch: %s\n", hipGetErrorString(e));
            rc = FSX_ERR_HIP;
        }
    }
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

extern "C" int fsx_hip_sbessel_batch(const double *x, int n, int pmax, double *out)
{
    pthread_mutex_lock(&g_mtx);
    int P = pmax + 1;
    int rc = init_once();
    if (rc == FSX_OK) rc = ensure(B_BX, sizeof(double) * n);
    if (rc == FSX_OK) rc = ensure(B_BOUT, sizeof(double) * (size_t)n * P);
    if (rc == FSX_OK) {
        hipStream_t s = g_stream;
        hipError_t e = hipMemcpyAsync(g_buf[B_BX], x, sizeof(double) * n, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            hipLaunchKernelGGL(k_sbessel_batch, dim3(nblk((long)n * P, 256)), dim3(256), 0, s,
                               (const double *)g_buf[B_BX], n, P, (double *)g_buf[B_BOUT]);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(out, g_buf[B_BOUT], sizeof(double) * (size_t)n * P, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            fprintf(stderr, "[fsx] sbessel batch: %s\n", hipGetErrorString(e));
            rc = FSX_ERR_HIP;
        }
    }
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

extern "C" int fsx_hip_atoms2spf(int natoms, const double *r, const double *f, const double *Y,
                                 const double *qvals, int Q, int L, double *re, double *im)
{
    pthread_mutex_lock(&g_mtx);
    const int K = L + 1, K2 = K * K;
    int rc = init_once();
    if (rc == FSX_OK) rc = ensure(B_AR, sizeof(double) * natoms);
    if (rc == FSX_OK) rc = ensure(B_AF, sizeof(double) * 3 * (size_t)natoms);
    if (rc == FSX_OK) rc = ensure(B_AY, sizeof(double) * 2 * (size_t)K2 * natoms);
    if (rc == FSX_OK) rc = ensure(B_AJ, sizeof(double) * (size_t)Q * K * natoms);
    if (rc == FSX_OK) rc = ensure(B_AQ, sizeof(double) * Q);
    if (rc == FSX_OK) rc = ensure(B_ARE, sizeof(double) * 3 * (size_t)Q * K2);
    if (rc == FSX_OK) rc = ensure(B_AIM, sizeof(double) * 3 * (size_t)Q * K2);
    if (rc == FSX_OK) {
        hipStream_t s = g_stream;
        hipError_t e = hipMemcpyAsync(g_buf[B_AR], r, sizeof(double) * natoms, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_AF], f, sizeof(double) * 3 * (size_t)natoms, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_AY], Y, sizeof(double) * 2 * (size_t)K2 * natoms, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_AQ], qvals, sizeof(double) * Q, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            long nj = (long)Q * K * natoms;
            hipLaunchKernelGGL(k_atom_bessel, dim3(nblk(nj, 256)), dim3(256), 0, s,
                               (double *)g_buf[B_AJ], (const double *)g_buf[B_AR],
                               (const double *)g_buf[B_AQ], natoms, Q, K);
            hipLaunchKernelGGL(k_atom_sum, dim3((unsigned)(Q * K2)), dim3(256), 0, s,
                               (double *)g_buf[B_ARE], (double *)g_buf[B_AIM],
                               (const double *)g_buf[B_AJ], (const double *)g_buf[B_AY],
                               (const double *)g_buf[B_AF], natoms, Q, L,
                               4.0 * FSX_PI_REFERENCE);
            e = hipGetLastError();
        }
        if (e == hipSuccess) e = hipMemcpyAsync(re, g_buf[B_ARE], sizeof(double) * 3 * (size_t)Q * K2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipMemcpyAsync(im, g_buf[B_AIM], sizeof(double) * 3 * (size_t)Q * K2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            fprintf(stderr, "[fsx] atoms2spf: %s\n", hipGetErrorString(e));
            rc = FSX_ERR_HIP;
        }
    }
    pthread_mutex_unlock(&g_mtx);
    return rc;
}
