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
#include "../fsx_pose.h"
/* the fit kernels keep the L-BFGS-B blocks K11, K12, K22 and WN1 in LDS, one
 * column of 256 lanes (a fit block) per value */
#define FSX_LB_KSHARED
#define FSX_LB_KSTRIDE 256
#include "../fsx_lbfgsb.h"
#include "../fsx_lb2.h"

#define FSX_CROSS_THREADS 256
#define FSX_WAVE 64
#define FSX_NXCD 8
/* counters of the fit kernel: evaluations, capped fits, loop iterations summed
 * over wavefronts, of those wavefront-wide ones, the longest wavefront's */
#define FSX_TALLY_N 5

struct cplx { double re, im; };
struct cplx32 { float re, im; };   /* element of the mixed-precision (FP32) tables */
typedef double fsx_v4d __attribute__((ext_vector_type(4)));   /* FP64 MFMA 16x16x4 result */

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

/* Packed translation blocks: per (z, q) the rows (|m|, l) for l >= |m| in
 * order of |m| then l, each holding l1 = |m|..L.  Row (am, l) starts at
 * S(K) - S(K - am) + (l - am)(K - am), S(n) = n(n+1)(2n+1)/6; the block has
 * S(K) entries. */
__host__ __device__ __forceinline__ int tpk_sq(int n) { return n * (n + 1) * (2 * n + 1) / 6; }
/* Entry stride (in elements) of the cross-term kernel's staged ligand rows.
 * FP32 tables keep the 3G values of an entry together; when 3G * 8 bytes is
 * a multiple of 32 (G = 4: 96 bytes) two elements of padding make it an odd
 * multiple of 16, so that the entries that one wavefront reads at a U step
 * fall on twice as many LDS bank offsets.  FP64 tables are [G][3][nc]. */
template <typename E, int G>
__host__ __device__ constexpr int cross_sb()
{
    return sizeof(E) == 8 ? ((3 * G) % 4 == 0 ? 3 * G + 2 : 3 * G) : 3 * G;
}
/* LDS bytes of the cross-term kernel's translation block (Tn values of type F),
 * rounded up to 8 so that the pose records after it stay aligned */
template <typename F>
__host__ __device__ __forceinline__ size_t cross_tst_bytes(int Tn)
{
    return ((size_t)Tn * sizeof(F) + 7) & ~(size_t)7;
}
__host__ __device__ __forceinline__ int tpk_row(int am, int l, int K)
{
    return tpk_sq(K) - tpk_sq(K - am) + (l - am) * (K - am);
}

/* T^{|m|}_{l l1}(q z) (fftsaxs.c:182-243), computed from (min(l,l1), max(l,l1))
 * so the mirrored entries are bitwise equal, as in fftsaxs.c:213-238.  One
 * thread per (z, q, packed entry) in the packed layout, per |m| block
 * [l1][l]: element (am, l, l1) at tpk_row(am, am, K) + (l1 - am) * (K - am) +
 * (l - am), the order in which the cross-term kernel's U step reads them
 * (consecutive l, consecutive threads). */
template <typename E>
__global__ void k_tmat(decltype(E::re) *T, const double *bes, const double *ds, int znum, int Q, int L)
{
#pragma clang fp contract(off)
    const int K = L + 1, N = 2 * L + 1, Tn = tpk_sq(K);
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)znum * Q * Tn) return;
    const int r = (int)(i % Tn);
    const long zq = i / Tn;
    /* |m| block of packed entry r: block am holds (K - am)^2 entries from
     * tpk_row(am, am, K) */
    int m = 0;
    while (m < L && r >= tpk_row(m + 1, m + 1, K)) m++;
    const int off = r - tpk_row(m, m, K), w = K - m;
    const int l1 = m + off / w, l = m + off % w;
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
    /* only p of the parity of l + l1 contribute (the 3j symbol (l p l1; 0 0 0)
     * vanishes exactly otherwise), so T = i^(l1 - l) tau with tau real, and
     * the other part is an exact zero: stored is tau = i^(l - l1) T */
    const int ph = ((l - l1) % 4 + 4) % 4;
    const double tau = ph == 0 ? re : ph == 1 ? -im : ph == 2 ? -re : im;
    T[zq * Tn + r] = (decltype(E::re))tau;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, FSX_WAVE);
    return v;
}

/* Cross-lane steps of the 64-lane transpose-reduce, on the two 32-bit halves
 * of a double.  v_permlane32_swap exchanges lanes 32..63 of its first operand
 * with lanes 0..31 of its second, v_permlane16_swap the odd 16-lane rows of
 * the first with the even rows of the second, so that the sum of the two
 * results is, in every lane, its own slot's value plus the partner lane's
 * value of the same slot (lanes below the split keep slot a, lanes above it
 * slot b). */
__device__ __forceinline__ double tr_swap32(double a, double b)
{
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
    const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
    const auto lo = __builtin_amdgcn_permlane32_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32),
                                                     false, false);
    const double r0 = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
    const double r1 = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
    return r0 + r1;
}
__device__ __forceinline__ double tr_swap16(double a, double b)
{
    const unsigned long long ua = __builtin_bit_cast(unsigned long long, a);
    const unsigned long long ub = __builtin_bit_cast(unsigned long long, b);
    const auto lo = __builtin_amdgcn_permlane16_swap((unsigned)ua, (unsigned)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((unsigned)(ua >> 32), (unsigned)(ub >> 32),
                                                     false, false);
    const double r0 = __builtin_bit_cast(double, ((unsigned long long)hi[0] << 32) | lo[0]);
    const double r1 = __builtin_bit_cast(double, ((unsigned long long)hi[1] << 32) | lo[1]);
    return r0 + r1;
}
__device__ __forceinline__ float tr_swap32(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane32_swap(__builtin_bit_cast(unsigned, a),
                                                    __builtin_bit_cast(unsigned, b), false, false);
    return __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
}
__device__ __forceinline__ float tr_swap16(float a, float b)
{
    const auto r = __builtin_amdgcn_permlane16_swap(__builtin_bit_cast(unsigned, a),
                                                    __builtin_bit_cast(unsigned, b), false, false);
    return __builtin_bit_cast(float, (unsigned)r[0]) + __builtin_bit_cast(float, (unsigned)r[1]);
}
/* a double moved between lanes by a DPP pattern (row_ror:8 = lane xor 8,
 * row_half_mirror = lane 7 - i in each group of 8, quad_perm) */
template <int CTRL>
__device__ __forceinline__ double dpp_mov(double v)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)u, CTRL, 0xf, 0xf, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_update_dpp(0, (int)(unsigned)(u >> 32), CTRL, 0xf, 0xf, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_mov(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL,
                                                                 0xf, 0xf, false));
}
#define FSX_DPP_ROR8 0x128
#define FSX_DPP_HMIRROR 0x141
#define FSX_DPP_XOR1 0xB1
#define FSX_DPP_XOR2 0x4E

/* x mod n in [0, n) for |x| < 2^24, n >= 1: a float quotient estimate
 * (off by at most one) and one correction. */
__device__ __forceinline__ int mod_small(int x, int n, float inv_n)
{
    int k = x - n * (int)floorf((float)x * inv_n);
    if (k < 0) k += n;
    else if (k >= n) k -= n;
    return k;
}

/* ---- rotated receptor / ligand coefficients -------------------------------
 * k_twiddle: src~[q][X][lm][g] = e^{-i m1 g} src^X_{l m1}(q) (conjugated first
 * for the ligand), gamma innermost.
 * k_rotate: out^X_{lm}(beta, gamma, q) = sum_{m1} d^l_{m m1}(beta) src~[q][X][l m1][gamma]
 * with one wavefront per (beta, q, X, entry) and the lanes over gamma, so the
 * Wigner-d value of a term is the same for the whole wavefront and the
 * source loads are contiguous.  Output layout per m-chunk c (entries e in
 * m-major order): out[coff[c] + ((q*nrows + beta*ng + gamma)*3 + X)*csize[c] + e - cstart[c]]. */
__global__ void k_twiddle(cplx *dst, const double *src_re, const double *src_im,
                          const double *twg, int ng, int Q, int L, int conj_src)
{
    const int K2 = (L + 1) * (L + 1);
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)Q * 3 * K2 * ng) return;
    int g = (int)(i % ng);
    long r = i / ng;
    int lm = (int)(r % K2);
    long xq = r / K2;   /* q*3 + X */
    int X = (int)(xq % 3), q = (int)(xq / 3);
    int l = lm_l(lm), m1 = lm - l * (l + 1);
    int k = mod_small(m1 * g, ng, 1.0f / (float)ng);
    double wr = twg[2 * k], wi = twg[2 * k + 1];
    long so = ((long)X * Q + q) * K2 + lm;
    double sr = src_re[so], si = (conj_src ? -1.0 : 1.0) * src_im[so];
    cplx v;
    v.re = wr * sr - wi * si;
    v.im = wr * si + wi * sr;
    dst[i] = v;
}

/* One block per (q, X, l, chunk of FSX_ROT_BC betas, gamma chunk of gc): the
 * 2l+1 source rows of l, gamma in the chunk, are staged in LDS
 * ([m1][gamma]) together with the blocks d^l_{m m1}(beta) of the chunk's
 * betas; the wavefronts then take (beta, pair of m) items and form
 * out^X_{lm}(beta, gamma) = sum_m1 d^l_{m m1}(beta) src_{l m1}(gamma), lanes
 * over gamma, m1 ascending.  The output goes to a staging table with rows
 * innermost, per chunk c [q][X][entry][row] at coff[c] (lanes store
 * consecutive rows); k_tab_transpose then writes the row-major table the
 * cross-term kernel reads. */
#define FSX_ROT_BC 2
template <typename E>
__global__ void __launch_bounds__(256)
k_rotate(E *out, const cplx *srct, const double *dW, const int *ent, const long *coff,
         int nbeta, int ng, int Q, int L, int gc, int mlo)
{
#pragma clang fp contract(fast)
    extern __shared__ cplx sR[];   /* [2l+1][gc], then FSX_ROT_BC blocks [2l+1][2l+1] of d */
    const int K = L + 1, K2 = K * K, N = 2 * L + 1;
    const int *ent_c = ent + 2 * K2, *ent_ec = ent + 3 * K2, *csize = ent + 4 * K2;
    const int *mstart = ent + 5 * K2;
    const int nrows = nbeta * ng;
    const int lane = threadIdx.x % FSX_WAVE, wave = threadIdx.x / FSX_WAVE;
    const int nwave = blockDim.x / FSX_WAVE;
    const int nch = (ng + gc - 1) / gc, nbc = (nbeta + FSX_ROT_BC - 1) / FSX_ROT_BC;
    int r = blockIdx.x;
    const int ch = r % nch; r /= nch;
    const int bc = r % nbc; r /= nbc;
    const int l = r % K; r /= K;
    const int X = r % 3;
    const int q = r / 3;
    const int g0 = ch * gc, gn = (ng - g0) < gc ? (ng - g0) : gc;
    const int b0 = bc * FSX_ROT_BC, nbh = nbeta - b0 < FSX_ROT_BC ? nbeta - b0 : FSX_ROT_BC;
    /* output rows m = mb..l (all m, or m >= 0 for the half entry set) */
    const int w = 2 * l + 1, mb = mlo > -l ? mlo : -l, nm = l - mb + 1, ww = nm * w;
    double *sD = reinterpret_cast<double *>(sR + w * gc);
    const cplx *src = srct + (((long)q * 3 + X) * K2 + l * (l + 1) - l) * ng + g0;
    for (int i = threadIdx.x; i < w * gc; i += blockDim.x) {
        int m1 = i / gc, g = i - m1 * gc;
        if (g < gn) sR[i] = src[(long)m1 * ng + g];
    }
    for (int i = threadIdx.x; i < nbh * ww; i += blockDim.x) {
        const int bb = i / ww, rem = i - bb * ww, mr = rem / w, k = rem - mr * w;
        sD[i] = dW[(long)(b0 + bb) * K * N * N + ((long)l * N + (L + mb + mr)) * N + (L - l + k)];
    }
    __syncthreads();
    const int np = (nm + 1) / 2;   /* pairs of m */
    for (int it = wave; it < nbh * np; it += nwave) {
        const int bb = it / np, mi = 2 * (it - bb * np), b = b0 + bb;
        const bool two = mi + 1 < nm;
        const double *d0 = sD + bb * ww + mi * w, *d1 = two ? d0 + w : d0;
        E *o[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
            /* entry (l, m) in m-major order */
            const int m = mb + (two ? mi + u : mi);
            const int e = mstart[m + L] + (l - (m < 0 ? -m : m));
            const int c = ent_c[e];
            const long cs = csize[c];
            o[u] = out + coff[c] + ((long)(q * 3 + X) * cs + ent_ec[e]) * nrows + (long)b * ng + g0;
        }
        for (int gw = 0; gw < gn; gw += FSX_WAVE) {
            const int g = gw + lane;
            if (g < gn) {
                double r0 = 0.0, i0 = 0.0, r1 = 0.0, i1 = 0.0;
#pragma unroll 4
                for (int k = 0; k < w; k++) {
                    const cplx v = sR[k * gc + g];
                    const double a0 = d0[k], a1 = d1[k];
                    r0 += a0 * v.re;
                    i0 += a0 * v.im;
                    r1 += a1 * v.re;
                    i1 += a1 * v.im;
                }
                E v0, v1;
                v0.re = r0;
                v0.im = i0;
                v1.re = r1;
                v1.im = i1;
                o[0][g] = v0;
                if (two) o[1][g] = v1;
            }
        }
    }
}

/* One chunk of a rotated table from the staging layout [q][X][entry][row]
 * to [q][row][X][entry] (cs entries), through a 64-entry x TR-row LDS tile
 * per block: loads run along rows, stores along entries. */
template <typename E, int TR>
__global__ void __launch_bounds__(256)
k_tab_transpose(E *out, const E *tmp, int cs, int nrows, int nte, int ntr, int npk,
                const int *lrot, int lsgn)
{
    __shared__ E tile[64][TR + 1];
    int bid = blockIdx.x;
    const int tr = bid % ntr;
    bid /= ntr;
    const int te = bid % nte;
    const int qx = bid / nte;   /* q * 3 + X */
    const int X = qx % 3, q = qx / 3;
    const int e0 = te * 64, r0 = tr * TR;
    const E *src = tmp + (long)qx * cs * nrows;
    for (int k = threadIdx.x; k < 64 * TR; k += 256) {
        const int i = k / TR, j = k - (k / TR) * TR;
        const int e = e0 + i, r = r0 + j;
        if (e < cs && r < nrows) tile[i][j] = src[(long)e * nrows + r];
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 64 * TR; k += 256) {
        const int j = k >> 6, i = k & 63;
        const int e = e0 + i, r = r0 + j;
        if (e < cs && r < nrows) {
            E *d = out + (((long)q * nrows + r) * 3 + X) * cs;
            E v = tile[i][j];
            if (lsgn != 0) {   /* times i^(lsgn l), l of the entry */
                const int ph = (lsgn > 0 ? lrot[e] : -lrot[e]) & 3;
                const decltype(v.re) a = v.re, b = v.im;
                v.re = ph == 0 ? a : ph == 1 ? -b : ph == 2 ? -a : b;
                v.im = ph == 0 ? b : ph == 1 ? a : ph == 2 ? -b : -a;
            }
            /* packed m = 0 pairs: i^l R_l0 is real for every l */
            if (e >= 2 * npk) d[e - npk] = v;
            else if (e & 1) d[e >> 1].im = v.re;
            else d[e >> 1].re = v.re;
        }
    }
}

/* Rotated table on FP64 MFMA, written in the cross-term kernel's layout in
 * one pass (no staging table, no transpose).  One block of four waves per
 * (q, X, beta, tile of GT gamma): for every l, the (m x m1) block of
 * d^l(beta) times the (m1 x gamma, re/im) block of twiddled source rows
 * src[q][X][l m1][gamma] (16x16x4 FP64 MFMA: lane i feeds A[i % 16][i / 16],
 * B[i / 16][i % 16]; result register j of lane i is D[i / 16 + 4j][i % 16]),
 * rows m = max(mlo, -l)..l, each block's m rows and 16 columns (8 gamma, real
 * and imaginary parts) a job of one wave.  The results go to an LDS tile of
 * the block's GT rows of the table, all entries; the tile is written out row
 * by row, turned by i^(lsgn l) and with the npk packed m = 0 slot pairs of
 * chunk 0, exactly as k_tab_transpose writes them. */
template <typename E, int GT>
__global__ void __launch_bounds__(256)
k_rotate_mfma(E *out, const cplx *srct, const double *dW, const int *ent, const long *coff,
              int nbeta, int ng, int Q, int L, int mlo, int ne, int npk, int lsgn)
{
    extern __shared__ double smem_raw[];
    E *tile = reinterpret_cast<E *>(smem_raw);   /* [GT][ne] */
    const int K = L + 1, K2 = K * K, N = 2 * L + 1;
    const int *ent_l = ent, *ent_c = ent + 2 * K2, *ent_ec = ent + 3 * K2, *csize = ent + 4 * K2;
    const int *mstart = ent + 5 * K2;
    const int nrows = nbeta * ng, ngt = (ng + GT - 1) / GT;
    int r = blockIdx.x;
    const int gt = r % ngt; r /= ngt;
    const int b = r % nbeta; r /= nbeta;
    const int X = r % 3;
    const int q = r / 3;
    const int g0 = gt * GT;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    constexpr int NT = GT / 8;   /* 16-column tiles: 8 gamma x (re, im) each */
    const double *dWb = dW + (size_t)b * K * N * N;
    const cplx *srcq = srct + (size_t)(q * 3 + X) * K2 * ng;
    /* jobs (l, m-tile, column tile), dealt to the waves in turn */
    int j = 0;
    for (int l = 0; l <= L; l++) {
        const int mb = mlo > -l ? mlo : -l, nm = l - mb + 1, nmt = (nm + 15) / 16;
        const int njob = nmt * NT;
        for (int jj = 0; jj < njob; jj++, j++) {
            if ((j & 3) != w) continue;
            const int mt = jj / NT, nt = jj - mt * NT;
            const int i = lane & 15, k = lane >> 4;
            const int m = mb + mt * 16 + i;
            const int gl = nt * 8 + (i >> 1), part = i & 1, g = g0 + gl;
            const bool arow = m <= l, bcol = gl < GT && g < ng;
            const double *drow = dWb + ((size_t)l * N + (L + (arow ? m : l))) * N + (L - l);
            const double *scol = reinterpret_cast<const double *>(srcq + (size_t)l * l * ng +
                                                                  (bcol ? g : g0)) + part;
            fsx_v4d acc = {0.0, 0.0, 0.0, 0.0};
            /* the job's operands of up to KS steps loaded together (one
             * round trip), then its MFMAs; the rest step by step.  KS = 8
             * (2l + 1 <= 32) for FP64 tables, whose LDS tile bounds the
             * blocks per CU anyway; 4 for FP32 tables, which keep more
             * blocks per CU with fewer registers (measured: the FP32 rotate
             * stage 0.52 against 0.60 ms at L = 15) */
            constexpr int KS = sizeof(E) == 8 ? 4 : 8;
            const int nks = (2 * l + 1 + 3) / 4;
            double av[KS], bv[KS];
#pragma unroll
            for (int st = 0; st < KS; st++) {
                const int m1i = 4 * st + k;
                const bool kin = st < nks && m1i <= 2 * l;
                av[st] = arow && kin ? drow[m1i] : 0.0;
                bv[st] = bcol && kin ? scol[(size_t)m1i * ng * 2] : 0.0;
            }
#pragma unroll
            for (int st = 0; st < KS; st++)
                if (st < nks) acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av[st], bv[st], acc, 0, 0, 0);
            for (int kk = 4 * KS; kk < 2 * l + 1; kk += 4) {
                const int m1i = kk + k;
                const bool kin = m1i <= 2 * l;
                const double a1 = arow && kin ? drow[m1i] : 0.0;
                const double b1 = bcol && kin ? scol[(size_t)m1i * ng * 2] : 0.0;
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc, 0, 0, 0);
            }
            const int c = lane & 15;
            const int gl2 = nt * 8 + (c >> 1), part2 = c & 1;
#pragma unroll
            for (int rr = 0; rr < 4; rr++) {
                const int m2 = mb + mt * 16 + (lane >> 4) + 4 * rr;
                if (m2 <= l && gl2 < GT) {
                    const int e = mstart[m2 + L] + (l - (m2 < 0 ? -m2 : m2));
                    decltype(E::re) *t = reinterpret_cast<decltype(E::re) *>(&tile[gl2 * ne + e]);
                    t[part2] = (decltype(E::re))acc[rr];
                }
            }
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < GT * ne; idx += blockDim.x) {
        const int gl = idx / ne, e = idx - gl * ne, g = g0 + gl;
        if (g >= ng) continue;
        E v = tile[idx];
        if (lsgn != 0) {   /* times i^(lsgn l), l of the entry */
            const int ph = (lsgn > 0 ? ent_l[e] : -ent_l[e]) & 3;
            const decltype(v.re) x = v.re, y = v.im;
            v.re = ph == 0 ? x : ph == 1 ? -y : ph == 2 ? -x : y;
            v.im = ph == 0 ? y : ph == 1 ? x : ph == 2 ? -y : -x;
        }
        const int c = ent_c[e], ec = ent_ec[e], cs = csize[c];
        E *d = out + coff[c] + ((size_t)((size_t)q * nrows + b * ng + g) * 3 + X) * cs;
        const int pk = c == 0 ? npk : 0;
        /* packed m = 0 pairs: i^l R_l0 is real for every l */
        if (ec >= 2 * pk) d[ec - pk] = v;
        else if (ec & 1) d[ec >> 1].im = v.re;
        else d[ec >> 1].re = v.re;
    }
}

/* ---- pose bucketing --------------------------------------------------- */
__global__ void k_fill_int(int *p, long n, int v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

/* bucket key (z, beta2, gamma2) and the pose's receptor row / alpha index */
__global__ void k_count(const long long *poses, int n, int znum, int nbeta, int na, int ng,
                        int *keys, int2 *info, int *counts)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int row, alpha;
    const int key = fsx_pose_key(poses[i], znum, nbeta, na, ng, &row, &alpha);
    if (key >= 0) atomicAdd(&counts[key], 1);
    keys[i] = key;
    info[i] = make_int2(row, alpha);
}

/* exclusive scan of counts[0..nk) into offs[0..nk]; offs[nk+1] = number of
 * non-empty buckets.  One block of 1024. */
__global__ void k_scan(const int *counts, int nk, int *offs)
{
    __shared__ int part[1024];
    __shared__ int nz[1024];
    int t = threadIdx.x;
    int chunk = (nk + 1023) / 1024;
    int b = t * chunk, e = b + chunk < nk ? b + chunk : nk;
    int s = 0, z = 0;
    for (int i = b; i < e; i++) {
        s += counts[i];
        z += counts[i] != 0;
    }
    part[t] = s;
    nz[t] = z;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        int v = t >= off ? part[t - off] : 0;
        int w = t >= off ? nz[t - off] : 0;
        __syncthreads();
        part[t] += v;
        nz[t] += w;
        __syncthreads();
    }
    int run = part[t] - s;
    for (int i = b; i < e; i++) {
        offs[i] = run;
        run += counts[i];
    }
    if (t == 1023) {
        offs[nk] = part[1023];
        offs[nk + 1] = nz[1023];
    }
}

/* sorted pose records: x = input position, y = receptor row, z = alpha index */
__global__ void k_scatter(const int *keys, const int2 *info, int n, int ng, const int *offs,
                          int *fill, int4 *recs)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = keys[i];
    if (k < 0) return;
    int p = offs[k] + atomicAdd(&fill[k], 1);
    int2 inf = info[i];
    /* input position, receptor row, alpha2, gamma2 (key % ng) */
    recs[p] = make_int4(i, inf.x, inf.y, k % ng);
}

/* ---- per-pose cross terms, one m-chunk per launch --------------------------
 * Work tile = (q, beta2, G consecutive gamma2) with all z inside.  The G
 * ligand row chunks (3 * nc entries each) are staged in LDS once per tile;
 * the packed translation block of each (z, q) is staged once per z (TST;
 * read from global memory when it does not fit) and serves the G rows; each
 * wavefront prefetches the receptor row of its first pose before U is
 * formed.  Per z with requested poses in the tile's G buckets
 * (z, beta2, gamma2), which are adjacent in the bucket order:
 *   U^Y_e = sum_{l1 >= |m|} conj(T^{|m|}_{l l1}(q z)) B^Y_{(l1, m)}
 * (one entry e per thread, for each of the G rows), then per pose (one
 * wavefront, J entries per lane)
 *   2 Re sum_e e^{-i m a2} (R^X_e U^Y_e [+ R^Y_e U^X_e])
 * over the chunk's entries e = (l, m), stored (first chunk) or added into
 * acc[(s*Q + q)*6 + k]. */
/* Workgroup barrier that orders LDS accesses only: unlike __syncthreads it
 * does not wait for the block's outstanding global stores (the cross-term
 * accumulator writes, which no other thread of the launch reads), so they
 * drain while the block goes on. */
__device__ __forceinline__ void lds_barrier()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

#define FSX_CROSS_BLOCK 256
#define FSX_CROSS_PB 64   /* poses per batch of the cross-term pose loop */
/* s_waitcnt vmcnt(0) (expcnt and lgkmcnt left at their maxima) */
#define FSX_WAITCNT_VM0 0x0F70
/* FP32 tables: at most 128 VGPRs, four blocks of four wavefronts per CU
 * (five blocks of 96 VGPRs were slower: 8.1 against 5.9 ms at L = 15);
 * FP64 tables: the LDS of three blocks fills a CU, no register bound */
template <int J, int TST, int G, typename E, int PW>
__device__ __forceinline__ void
cross_body(double *acc, const E *Rc, const E *Bc, const decltype(E::re) *T, const E *tw2,
           const int *ent, int cstart, int nc, const int *counts, const int *offs,
           const int4 *recs, long ntiles, int nrows, int Q, int L, int nbeta, int nx, int ng,
           int zn, int first, int npk)
{
#pragma clang fp contract(fast)
    typedef decltype(E::re) F;   /* arithmetic type of the contraction */
    static_assert(G <= 8, "the tile row of a pose is kept in three bits of its record");
    extern __shared__ double smem_raw[];
    E *smem = reinterpret_cast<E *>(smem_raw);
    /* ligand rows: FP32 tables [nc][SB] (the 3G values of one entry together,
     * read as 16-byte pairs, then cross_sb's padding), FP64 tables [G][3][nc] */
    constexpr bool BIL = sizeof(E) == 8;
    E *sB = smem;
    constexpr int SB = cross_sb<E, G>();   /* sB: [nc][SB] (BIL) or [G][3][nc] */
    E *sU = smem + SB * nc;   /* [G][3][nc] */
    F *sT = reinterpret_cast<F *>(sU + 3 * G * nc);   /* T block, [l1][l] per |m| as in T (TST) */
    const int K = L + 1, K2 = K * K, N = 2 * L + 1, Tn = tpk_sq(K);
    /* the current batch of up to FSX_CROSS_PB of the tile's poses: their
     * (receptor row, 8 alpha2 + gamma2 row within the tile), read from LDS
     * in the pose loop so that a pose
     * waits for one global round trip (its receptor row), not two, and their
     * six lane sums (type F), written to acc by the whole block after the
     * batch, so that no store is in flight in the pose loop */
    int2 *sP = reinterpret_cast<int2 *>(reinterpret_cast<char *>(sU + 3 * G * nc) +
                                        (TST ? cross_tst_bytes<F>(Tn) : 0));
    F *sO = reinterpret_cast<F *>(sP + FSX_CROSS_PB);
    const int *ent_l = ent, *ent_m = ent + K2, *mstart = ent + 5 * K2;
    const int tid = threadIdx.x, lane = tid & (FSX_WAVE - 1), wave = tid / FSX_WAVE;
    const int nwave = FSX_CROSS_BLOCK / FSX_WAVE;
    /* U items (entries ordered by decreasing row length K - |m|, so that
     * the lanes of a wavefront run loops of nearly equal length).  The
     * nsplit longest entries are split into the even and the odd steps of
     * their l1 range (adjacent entries, whose LDS reads fall on different
     * banks), run by lanes 2i and 2i + 1 and summed across the pair; every
     * other entry takes one thread.  nsplit fills the block: with nc + nsplit
     * items the longest loop is about half the longest row. */
    const int *uperm = ent + 7 * K2 + 2;
    int nlen1 = 0;   /* trailing entries of row length 1 (|m| = L): not split */
    for (int i = nc - 1; i >= 0 && nlen1 < 2; i--) {
        const int mm = ent_m[cstart + uperm[cstart + i]];
        if ((mm < 0 ? -mm : mm) != L) break;
        nlen1++;
    }
    const int nsplit = nc < FSX_CROSS_BLOCK ? min(FSX_CROSS_BLOCK - nc, nc - nlen1) : 0;
    const bool uact = tid < nc + nsplit;
    const bool usplit = tid < 2 * nsplit;
    int ue = 0, ul = 0, um = 0, uam = 0, uk0 = 0, uk1 = 0, ukst = 1;
    if (uact) {
        ue = uperm[cstart + (usplit ? tid >> 1 : tid - nsplit)];
        ul = ent_l[cstart + ue];
        um = ent_m[cstart + ue];
        uam = um < 0 ? -um : um;
        uk0 = usplit && (tid & 1) ? 1 : 0;
        uk1 = L - uam + 1;
        ukst = usplit ? 2 : 1;
    }
    const bool ustore = uact && !(usplit && (tid & 1));
    const int ub0 = uact ? mstart[um + L] - cstart : 0;
    /* T (in global memory, and its LDS copy when TST) is stored per |m|
     * block as [l1][l] (consecutive l, i.e. consecutive threads, at
     * consecutive addresses): element (am, l, l1) at tpk_row(am, am, K) +
     * (l1 - am) * (K - am) + (l - am) */
    const int urow = uact ? tpk_row(uam, uam, K) + (ul - uam) : 0;
    const int ustep = K - uam;
    /* pairing slots: with npk > 0 (conjugate-symmetric coefficients, the
     * chunk of m = 0) the m = 0 entries l = 2i and 2i + 1 share slot i, the
     * real part of the first (real for even l) and the imaginary part of the
     * second (imaginary for odd l); their twiddle is 1, so Re(w R U) of the
     * slot is the sum of the two entries' terms.  Entry e >= 2 npk has slot
     * e - npk. */
    const int ns = nc - npk;
    /* PW poses per wavefront, LP lanes each: lane sl of a pose takes slots
     * sl, sl + LP, ... (J per lane) */
    constexpr int LP = FSX_WAVE / PW;
    const int sl = lane % LP, half = lane / LP;
    /* bytes per receptor row and per row of U in LDS (three parts of nc
     * entries each), the lane's slot within a row, and the step to its last
     * slot (back to slot 0 when that one is past the last slot) */
    const unsigned rstride = 3u * (unsigned)nc * (unsigned)sizeof(E);
    const unsigned soff = (unsigned)(sl * sizeof(E));
    const unsigned dlast = sl + LP * (J - 1) < ns ? (unsigned)(LP * (J - 1) * sizeof(E))
                                                  : (unsigned)0 - soff;
    /* byte offset of each slot's twiddle e^(-i m alpha2) within a row of tw2
     * (rows of 2L + 2 entries, m = -L..L and a zero at 2L + 1, which a slot
     * past the last one reads) */
    unsigned emo[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        int ec = sl + LP * j;
        const int m = ec < ns ? ent_m[cstart + (ec < npk ? 2 * ec : ec + npk)] : L + 1;
        emo[j] = (unsigned)(m + L) * (unsigned)sizeof(E);
    }
    /* after the lane reduction: one pose: lanes 0, 8, .., 40 hold the six
     * sums; two poses: lanes 0, 8, 4, 16, 24, 20 of each half (bits 4, 3, 2
     * of the lane select sum 3 b4 + (b2 ? 2 : b3)) */
    const int rb4 = (lane >> 4) & 1, rb3 = (lane >> 3) & 1, rb2 = (lane >> 2) & 1;
    const int slot = PW == 1 ? lane >> 3 : 3 * rb4 + (rb2 ? 2 : rb3);
    const bool writer = PW == 1 ? (lane & 7) == 0 && slot < 6 : (lane & 3) == 0 && !(rb2 && rb3);
    const int ngg = (ng + G - 1) / G;   /* gamma2 groups per beta2 */
    const int nbg = nbeta * ngg;
    /* Tiles in nx contiguous ranges of the (q, beta2, gamma2 group) order,
     * one per XCD: workgroups are dispatched to the XCDs round-robin, so
     * block b runs on XCD b % nx, and the blocks of one XCD walk one range
     * together.  At any time they then read the receptor rows of about one
     * q, which that XCD's L2 can hold.  nx = 1: one range, plain grid
     * stride.  The grid is a multiple of nx. */
    const long per = (ntiles + nx - 1) / nx;
    const int xb = blockIdx.x % nx, gb = gridDim.x / nx;
    for (long j = blockIdx.x / nx; j < per; j += gb) {
        const long tile = xb * per + j;
        if (tile >= ntiles) break;
        const int q = (int)(tile / nbg), bgg = (int)(tile - (long)q * nbg);
        const int b2 = bgg / ngg, g20 = (bgg - b2 * ngg) * G;
        const int nr = ng - g20 < G ? ng - g20 : G;   /* rows of this tile */
        const E *Bq = Bc + (((long)q * nrows + b2 * ng + g20) * 3) * nc;
        const E *Rq = Rc + (long)q * nrows * 3 * nc;   /* receptor rows of this q */
        const char *RqB = reinterpret_cast<const char *>(Rq);
        bool staged = false;
        /* Each z's bucket header, translation block (TST) and first pose
         * records are read one z ahead: the next z's header is requested
         * before the U step, its translation block and records after it,
         * so that they arrive while this z's U
         * step and pose loop run instead of stalling the block between z
         * values.  Header: the tile's buckets are adjacent keys, their poses
         * adjacent in recs; lane r < nr of every wavefront reads the count
         * of row r and lane G the first record's offset (vector loads, so
         * that waiting for them does not wait for the LDS traffic of the U
         * step), read back by lane broadcast. */
        auto hload = [&](int zz) {
            const int key = (zz * nbeta + b2) * ng + g20;
            int v = 0;
            if (lane < nr) v = counts[key + lane];
            else if (lane == G) v = offs[key];
            return v;
        };
        /* first pose of each row of the tile (rows past nr: none) */
        auto hread = [&](int v, int (&cm)[G + 1], int &st) {
            cm[0] = 0;
#pragma unroll
            for (int r = 0; r < G; r++)
                cm[r + 1] = cm[r] + (r < nr ? __builtin_amdgcn_readlane(v, r) : 0);
            st = __builtin_amdgcn_readlane(v, G);
        };
        /* a translation block straight into sT (TST): each wavefront copies
         * 64 consecutive 4-byte words per instruction with global_load_lds,
         * so the copy runs in the background without registers; it is
         * waited for (vmcnt) before the barrier that precedes the U step */
        auto tdma = [&](int zz) {
            const unsigned *Tz = reinterpret_cast<const unsigned *>(T + ((long)zz * Q + q) * Tn);
            const int nw = Tn * (int)(sizeof(F) / 4);
            unsigned *dst = reinterpret_cast<unsigned *>(sT);
            for (int b = wave * FSX_WAVE; b < nw; b += nwave * FSX_WAVE)
                if (b + lane < nw)
                    __builtin_amdgcn_global_load_lds(
                        Tz + b + lane, (__attribute__((address_space(3))) void *)(dst + b), 4, 0, 0);
        };
        int2 rpf = make_int2(0, 0);   /* a first record (tid < FSX_CROSS_PB) */
        auto rload = [&](int st, int n) {
            if (n > 0 && tid < FSX_CROSS_PB) {
                const int4 r4 = recs[st + min(tid, n - 1)];
                rpf = make_int2(r4.y, (r4.z << 3) | (r4.w - g20));
            }
        };
        int cum[G + 1], start = 0;
        hread(hload(0), cum, start);
        int cnt = cum[G];
        if (TST && cnt > 0) tdma(0);
        rload(start, cnt);
        for (int z = 0; z < zn; z++) {
            const bool zmore = z + 1 < zn;
            int cum2[G + 1], start2 = 0;
#pragma unroll
            for (int r = 0; r <= G; r++) cum2[r] = 0;
            if (cnt == 0) {
                /* an empty z: straight on to the next one */
                if (zmore) {
                    hread(hload(z + 1), cum2, start2);
                    const int cnt2 = cum2[G];
                    if (TST && cnt2 > 0) tdma(z + 1);
                    rload(start2, cnt2);
#pragma unroll
                    for (int r = 0; r <= G; r++) cum[r] = cum2[r];
                    start = start2;
                    cnt = cnt2;
                }
                continue;
            }
            const F *Tq = T + ((long)z * Q + q) * Tn;
            /* pose loader: receptor row entries, twiddles and row of pose p,
             * whose record is sP[p - pb].  One register set (a second one for
             * the next pose's loads, in flight during this pose's pairing,
             * measured slower: the kernel then holds three blocks per CU
             * instead of four); the first pose is loaded after the U step
             * (loading it before kept its registers live through U and cost
             * FP64 tables a wavefront per SIMD) */
            E rrc[3][J];
            F wrc[J], wic[J];
            int prowc = 0;
            auto load_pose = [&](int p, int pb, E (&rr)[3][J], F (&wr)[J], F (&wi)[J],
                                 int &prow) {
                const int2 rec = sP[p - pb];
                prow = rec.y & 7;
                /* no branches, so that the J slots' loads are all in flight
                 * together: a lane's slot past the last one (only the last
                 * of the J can be) reads slot 0, and the pairing gives it a
                 * zero twiddle.  Receptor entries: the part's base (uniform,
                 * in scalar registers) plus one 32-bit byte offset of the
                 * row and the lane's slot (the host keeps a q's rows below
                 * 4 GiB and their count below 2^24), the slot step LP j as
                 * the load's immediate offset */
                const unsigned roff = __umul24((unsigned)rec.x, rstride) + soff;
                const unsigned twr = __umul24((unsigned)rec.y >> 3, (unsigned)((N + 1) * sizeof(E)));
#pragma unroll
                for (int j = 0; j < J; j++) {
                    const E w = *reinterpret_cast<const E *>(
                        reinterpret_cast<const char *>(tw2) + (twr + emo[j]));
                    wr[j] = w.re;
                    wi[j] = w.im;
#pragma unroll
                    for (int X = 0; X < 3; X++) {
                        const char *base = RqB + (size_t)X * nc * sizeof(E);
                        rr[X][j] = j < J - 1
                                       ? *reinterpret_cast<const E *>(base + (size_t)roff +
                                                                      (size_t)(LP * j * sizeof(E)))
                                       : *reinterpret_cast<const E *>(base + (size_t)(roff + dlast));
                    }
                }
            };
            /* the translation block copied in the background (TST) has
             * landed in sT (the barrier before the U step makes it visible
             * to the other wavefronts) */
            if (TST) __builtin_amdgcn_s_waitcnt(FSX_WAITCNT_VM0);
            /* the next z's header, requested after that wait (which would
             * otherwise wait for it too) and read after the U step */
            const int hv = zmore ? hload(z + 1) : 0;
            __builtin_amdgcn_sched_barrier(0);

            /* all FSX_CROSS_PB records of a batch are valid: past the
             * z's last pose they repeat it, so the pose loop reads them
             * without a clamp */
            if (tid < FSX_CROSS_PB) sP[tid] = rpf;
            /* staging: four loads per thread in flight (an index past the end
             * reloads element b0, whose store is then skipped), then the stores */
            if (!staged) {
                /* global (row r, part Y, entry e) -> LDS e * 3G + 3r + Y (BIL) */
                const int nb3 = 3 * nc * nr;
                auto bdst = [&](int b) {
                    if (!BIL) return b;
                    const int rY = b / nc;
                    return (b - rY * nc) * SB + rY;
                };
                for (int b0 = tid; b0 < nb3; b0 += 4 * FSX_CROSS_BLOCK) {
                    const int i1 = b0 + FSX_CROSS_BLOCK, i2 = b0 + 2 * FSX_CROSS_BLOCK;
                    const int i3 = b0 + 3 * FSX_CROSS_BLOCK;
                    E t0 = Bq[b0];
                    E t1 = Bq[i1 < nb3 ? i1 : b0];
                    E t2 = Bq[i2 < nb3 ? i2 : b0];
                    E t3 = Bq[i3 < nb3 ? i3 : b0];
                    sB[bdst(b0)] = t0;
                    if (i1 < nb3) sB[bdst(i1)] = t1;
                    if (i2 < nb3) sB[bdst(i2)] = t2;
                    if (i3 < nb3) sB[bdst(i3)] = t3;
                }
                staged = true;
            }
            lds_barrier();
            {
                /* T = i^(l1 - l) tau with tau real and the ligand rows stored
                 * as i^(-l1) B, so that i^(-l) conj(T) B = tau (i^(-l1) B): a
                 * real-by-complex sum */
                const F *Tr = (TST ? sT : Tq) + urow;
                const E *bp = sB + ub0 * (BIL ? SB : 1);
                F ur[G][3], ui[G][3];
#pragma unroll
                for (int r = 0; r < G; r++)
#pragma unroll
                    for (int Y = 0; Y < 3; Y++) ur[r][Y] = ui[r][Y] = 0;
                for (int k = uk0; k < uk1; k += ukst) {
                    /* the step's LDS reads first, then its FMAs */
                    const F tv = Tr[k * ustep];
                    E bv[3 * G];
#pragma unroll
                    for (int i = 0; i < 3 * G; i++) bv[i] = BIL ? bp[k * SB + i] : bp[i * nc + k];
#pragma unroll
                    for (int r = 0; r < G; r++)
#pragma unroll
                        for (int Y = 0; Y < 3; Y++) {
                            ur[r][Y] += tv * bv[r * 3 + Y].re;
                            ui[r][Y] += tv * bv[r * 3 + Y].im;
                        }
                }
                /* split entries: the even steps (even lane) plus the odd
                 * steps (its odd neighbour), in that order */
                if (nsplit > 0) {
#pragma unroll
                    for (int r = 0; r < G; r++)
#pragma unroll
                        for (int Y = 0; Y < 3; Y++) {
                            const F orr = dpp_mov<FSX_DPP_XOR1>(ur[r][Y]);
                            const F oi = dpp_mov<FSX_DPP_XOR1>(ui[r][Y]);
                            if (usplit) {
                                ur[r][Y] += orr;
                                ui[r][Y] += oi;
                            }
                        }
                }
                /* stored is i^(-l) U (the receptor rows hold i^l R, so the
                 * pairing's products are unchanged); in a packed m = 0 slot
                 * both values are real, the second one negated so that
                 * Re(R U) of the slot adds the two entries' products */
                if (ustore) {
                    const int us = ue < 2 * npk ? ue >> 1 : ue - npk;   /* pairing slot */
#pragma unroll
                    for (int r = 0; r < G; r++)
#pragma unroll
                        for (int Y = 0; Y < 3; Y++) {
                            E *d = sU + (r * 3 + Y) * nc + us;
                            if (ue >= 2 * npk) {
                                d->re = ur[r][Y];
                                d->im = ui[r][Y];
                            } else if (ue & 1) {
                                d->im = -ur[r][Y];
                            } else {
                                d->re = ur[r][Y];
                            }
                        }
                }
            }
            lds_barrier();
            /* the next z's header, and its first records and translation
             * block in flight during this z's pose loop (U has read sT) */
            int cnt2 = 0;
            if (zmore) {
                hread(hv, cum2, start2);
                cnt2 = cum2[G];
                rload(start2, cnt2);
                if (TST && cnt2 > 0) tdma(z + 1);
            }
            /* batches of FSX_CROSS_PB poses; the records of the first are
             * staged with the ligand rows */
            for (int pb = 0; pb < cnt; pb += FSX_CROSS_PB) {
            if (pb > 0) {
                if (tid < FSX_CROSS_PB) {
                    const int4 r4 = recs[start + min(pb + tid, cnt - 1)];
                    sP[tid] = make_int2(r4.y, (r4.z << 3) | (r4.w - g20));
                }
                lds_barrier();
            }
            const int pe = min(cnt, pb + FSX_CROSS_PB);
            /* pairing of one pose (this lane's half of the wavefront) with
             * the pose record's receptor row and twiddles in rr, wr, wi */
            auto pair_pose = [&](const E (&rr)[3][J], const F (&wr)[J], const F (&wi)[J],
                                 int prow, int p) {
                const bool act = p < pe;
                const E *sUp = reinterpret_cast<const E *>(reinterpret_cast<const char *>(sU) +
                                                           __umul24((unsigned)prow, rstride));
                F va[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
                for (int j = 0; j < J; j++) {
                    const bool ok = j < J - 1 || sl + LP * j < ns;
                    const int ec = ok ? sl + LP * j : 0;
                    {
                        E rv = rr[0][j], rd = rr[1][j], rw = rr[2][j];
                        E uv = sUp[ec], ud = sUp[nc + ec], uw = sUp[2 * nc + ec];
                        const F w_r = wr[j], w_i = wi[j];   /* zero past the last slot */
                        /* twiddled receptor factor: w * R^X */
                        F vr = w_r * rv.re - w_i * rv.im, vi = w_r * rv.im + w_i * rv.re;
                        F dr = w_r * rd.re - w_i * rd.im, di = w_r * rd.im + w_i * rd.re;
                        F xr = w_r * rw.re - w_i * rw.im, xi = w_r * rw.im + w_i * rw.re;
                        /* Re of each product, accumulated term by term (one
                         * fused multiply-add per term) */
                        va[0] += vr * uv.re; va[0] -= vi * uv.im;
                        va[1] += vr * ud.re; va[1] -= vi * ud.im;
                        va[1] += dr * uv.re; va[1] -= di * uv.im;
                        va[2] += vr * uw.re; va[2] -= vi * uw.im;
                        va[2] += xr * uv.re; va[2] -= xi * uv.im;
                        va[3] += dr * ud.re; va[3] -= di * ud.im;
                        va[4] += dr * uw.re; va[4] -= di * uw.im;
                        va[4] += xr * ud.re; va[4] -= xi * ud.im;
                        va[5] += xr * uw.re; va[5] -= xi * uw.im;
                    }
                }
                /* the cross-lane sums run in the contraction's type */
                F v[8] = {va[0], va[1], va[2], va[3], va[4], va[5], 0, 0};
                if (PW == 1) {
                    /* transpose-reduce over the 64 lanes: after the three
                     * halving steps lane bits 5,4,3 select the slot, then a
                     * sum over the 8 lanes of bits 2,1,0; lanes 0, 8, .., 40
                     * hold slots 0..5 */
#pragma unroll
                    for (int i = 0; i < 4; i++) v[i] = tr_swap32(v[i], v[i + 4]);
#pragma unroll
                    for (int i = 0; i < 2; i++) v[i] = tr_swap16(v[i], v[i + 2]);
                    const bool hi = (lane & 8) != 0;
                    const F send = hi ? v[0] : v[1];
                    const F keep = hi ? v[1] : v[0];
                    v[0] = keep + dpp_mov<FSX_DPP_ROR8>(send);
                    v[0] += dpp_mov<FSX_DPP_HMIRROR>(v[0]);
                    v[0] += dpp_mov<FSX_DPP_XOR1>(v[0]);
                    v[0] += dpp_mov<FSX_DPP_XOR2>(v[0]);
                } else {
                    /* the same within each 32-lane half (one pose each), on
                     * the six sums: lane bit 4 (16-row swap) pairs sums i and
                     * i + 3; bit 3 (row_ror:8) pairs the first two of those
                     * and adds the third across itself; bit 2
                     * (row_half_mirror: lane j <-> 7 - j) pairs the two
                     * results; then a sum over the quad of bits 1,0 */
#pragma unroll
                    for (int i = 0; i < 3; i++) v[i] = tr_swap16(v[i], v[i + 3]);
                    const bool b3 = rb3 != 0, b2 = rb2 != 0;
                    {
                        const F send = b3 ? v[0] : v[1];
                        const F keep = b3 ? v[1] : v[0];
                        v[0] = keep + dpp_mov<FSX_DPP_ROR8>(send);
                        v[1] = v[2] + dpp_mov<FSX_DPP_ROR8>(v[2]);
                    }
                    const F send = b2 ? v[0] : v[1];
                    const F keep = b2 ? v[1] : v[0];
                    v[0] = keep + dpp_mov<FSX_DPP_HMIRROR>(send);
                    v[0] += dpp_mov<FSX_DPP_XOR1>(v[0]);
                    v[0] += dpp_mov<FSX_DPP_XOR2>(v[0]);
                }
                if (writer && act) sO[(p - pb) * 6 + slot] = v[0];
            };
            /* the wavefront's poses of this batch, pp = wave, wave + nwave,
             * ...: PW consecutive poses per step (a half past the batch's
             * last pose repeats it and does not write) */
            for (int pp = wave; pb + PW * pp < pe; pp += nwave) {
                const int p = pb + PW * pp + half;
                load_pose(p, pb, rrc, wrc, wic, prowc);
                /* every load of the pose is issued before the first use */
                __builtin_amdgcn_sched_barrier(0);
                pair_pose(rrc, wrc, wic, prowc, p);
            }
            /* the batch's sums to acc (pose-major [pose][q][6]), added to
             * those of the earlier m-chunks */
            lds_barrier();
            for (int i = tid; i < (pe - pb) * 6; i += FSX_CROSS_BLOCK) {
                const int ip = i / 6;
                double *d = acc + ((long)(start + pb + ip) * Q + q) * 6 + (i - ip * 6);
                const double x = 2.0 * (double)sO[i];
                *d = first ? x : *d + x;
            }
            }
            /* no barrier here: until the next z's first barrier a wavefront
             * only writes the next records (sP), the staged ligand rows (sB,
             * at a new tile) and the translation block (sT, read by the U
             * step before its closing barrier), none of which the batch's
             * writing of sO reads; the next pose loop writes sO only after
             * that barrier */
#pragma unroll
            for (int r = 0; r <= G; r++) cum[r] = cum2[r];
            start = start2;
            cnt = cnt2;
        }
    }
}

template <int J, int TST, int G, typename E, int PW>
__global__ void __launch_bounds__(FSX_CROSS_BLOCK, sizeof(E) == 8 ? 4 : 1)
k_cross(double *acc, const E *Rc, const E *Bc, const decltype(E::re) *T, const E *tw2,
        const int *ent, int cstart, int nc, const int *counts, const int *offs,
        const int4 *recs, long ntiles, int nrows, int Q, int L, int nbeta, int nx, int ng,
        int zn, int first, int npk)
{
    cross_body<J, TST, G, E, PW>(acc, Rc, Bc, T, tw2, ent, cstart, nc, counts, offs, recs, ntiles,
                                 nrows, Q, L, nbeta, nx, ng, zn, first, npk);
}

/* One m-chunk of several complexes in one persistent launch (the batched
 * form): every block takes its share of the tiles of job 0, then of job 1,
 * and so on, so that a block that runs out of one complex's tiles goes on
 * to the next complex instead of idling until the launch drains.  The jobs
 * of one launch share the template parameters; each complex appears at most
 * once (its m-chunks are launched in order, one launch per chunk index). */
struct fsx_cross_job {
    double *acc;
    const void *Rc, *Bc, *T, *tw2;
    const int *ent, *counts, *offs;
    const int4 *recs;
    long ntiles;
    int cstart, nc, nrows, Q, L, nbeta, nx, ng, zn, first, npk;
};

template <int J, int TST, int G, typename E, int PW>
__global__ void __launch_bounds__(FSX_CROSS_BLOCK, sizeof(E) == 8 ? 4 : 1)
k_cross_jobs(const fsx_cross_job *jobs, int njobs)
{
    typedef decltype(E::re) F;
    for (int k = 0; k < njobs; k++) {
        const fsx_cross_job &jb = jobs[k];
        cross_body<J, TST, G, E, PW>(jb.acc, (const E *)jb.Rc, (const E *)jb.Bc, (const F *)jb.T,
                                     (const E *)jb.tw2, jb.ent, jb.cstart, jb.nc, jb.counts,
                                     jb.offs, jb.recs, jb.ntiles, jb.nrows, jb.Q, jb.L, jb.nbeta,
                                     jb.nx, jb.ng, jb.zn, jb.first, jb.npk);
        /* the next job lays out the LDS for its own chunk size */
        __syncthreads();
    }
}

/* ---- per-pose fit, lockstep over a pose queue ------------------------------
 * Each lane owns one fit at a time.  All busy lanes of a wavefront evaluate
 * chi^2 and its gradient together (the same instruction stream over q), then
 * each advances its own L-BFGS-B state to the next point it needs; a lane
 * whose fit ended stores the result and takes the next pose from the queue.
 * Profiles: pose-major acc[(s*Q + q)*6 + k] plus the constant parts cint[k*Q
 * + q], rescaled by peak / I(q0) (min_saxs.c:170-179). */
/* Scale of each pose's profile, peak / I(q0) with I(q0) the six terms at
 * q index 0 summed as min_saxs.c:170-179 (one thread per pose slot). */
__global__ void k_fit_scale(const double *acc, const double *cint, const int *nvalid_p, int Q,
                            double peak, double *scale)
{
#pragma clang fp contract(off)
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= *nvalid_p) return;
    const double *v = acc + (long)s * Q * 6;
    double g[6];
    for (int k = 0; k < 6; k++) {
        double x = cint[k * Q];
        x = x + v[k];
        g[k] = x * 1.0;
    }
    const double sc = g[0] + g[3] + g[5] + g[2] - g[1] - g[4];
    scale[s] = peak / sc;
}

/* The fit's profiles, (constant part + pose part) * scale, in place of the
 * pose part, and the unscaled profiles to cross_out when asked; one thread
 * per (pose slot, q): its six terms, contiguous and 16-byte aligned, are read
 * and written as three pairs. */
/* accf (the mixed form's lane-group fit): the rescaled profiles in FP32 to
 * accf [pose][q][6] instead of in place */
__global__ void k_fit_prep(double *acc, const double *cint, const int4 *recs,
                           const int *nvalid_p, int Q, const double *scale, double *cross_out,
                           float *accf)
{
#pragma clang fp contract(off)
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)*nvalid_p * Q) return;
    const int s = (int)(i / Q), q = (int)(i - (long)s * Q);
    double2 *a2 = reinterpret_cast<double2 *>(acc + i * 6);
    double x[6];
#pragma unroll
    for (int h = 0; h < 3; h++) {
        const double2 p = a2[h];
        x[2 * h] = p.x;
        x[2 * h + 1] = p.y;
    }
#pragma unroll
    for (int k = 0; k < 6; k++) {
        double y = cint[k * Q + q];
        y = y + x[k];
        x[k] = y;
    }
    if (cross_out) {
        const long o = (long)recs[s].x * 6 * Q + q;
#pragma unroll
        for (int k = 0; k < 6; k++) cross_out[o + (long)k * Q] = x[k] * 1.0;
    }
    const double sc = scale[s];
    if (accf) {
        float2 *f2 = reinterpret_cast<float2 *>(accf + i * 6);
#pragma unroll
        for (int h = 0; h < 3; h++) f2[h] = make_float2((float)(x[2 * h] * sc), (float)(x[2 * h + 1] * sc));
        return;
    }
#pragma unroll
    for (int h = 0; h < 3; h++) a2[h] = make_double2(x[2 * h] * sc, x[2 * h + 1] * sc);
}

__device__ __forceinline__ void fit_begin(fsx_ct *t, fsx_lbrc *S, const fsx_fitpar *fp,
                                          const double *acc, int s, int Q)
{
    t->v = acc + (long)s * Q * 6;
    fsx_lb_start(S, fp, 1.0, 0.0);
}

/* LDS of one fit block (256 lanes): the L-BFGS-B K blocks, the sums of the
 * wavefront-wide evaluation (FSX_WIDE_SUMS x 64 per wavefront), then (when
 * they fit, aq_lds) the moments, the q grid and the reciprocals
 * 1/(q_i - q_{i-1}) */
#define FSX_FIT_LDS (160 * 1024)
#define FSX_WIDE_SUMS 3
static inline size_t fit_lds_k(void)
{
    return (size_t)(FSX_LB_KN * FSX_LB_KSTRIDE + 4 * FSX_WIDE_SUMS * 64) * sizeof(double);
}
static inline int fit_aq_lds(int Q) { return fit_lds_k() + 8 * (size_t)Q * sizeof(double) <= FSX_FIT_LDS; }
static inline size_t fit_lds(int Q) { return fit_lds_k() + (fit_aq_lds(Q) ? 8 * (size_t)Q * sizeof(double) : 0); }

/* lane l's value of v, l the same in every lane */
__device__ __forceinline__ double lane_d(double v, int l)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, l);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

/* LDS written by some lanes of this wavefront, read by others */
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* sums[c] += row c of sW (nq values, in order) in lane c < nc: the q-ordered
 * sums of the wide evaluation */
__device__ __forceinline__ void wide_sum(const double *sW, int nq, int nc, double *sum)
{
    const int lane = threadIdx.x & 63;
    if (lane < nc) {
        const double *row = sW + lane * 64;
        double c = *sum;
        int i = 0;
        for (; i + 8 <= nq; i += 8) {
            double t[8];
#pragma unroll
            for (int u = 0; u < 8; u++) t[u] = row[i + u];
#pragma unroll
            for (int u = 0; u < 8; u++) c = c + t[u];
        }
        for (; i < nq; i++) c = c + row[i];
        *sum = c;
    }
}

/* The six one-pass sums of the wide evaluation, two rows of three at a
 * time: terms tt[0..5] of this lane's q (rows 0..nq-1 used), added in q
 * order into sa (sums 0..2, lanes 0..2) and sb (sums 3..5, lanes 0..2). */
__device__ __forceinline__ void wide_sum6(double *sW, const double *tt, int nq, double *sa,
                                          double *sb)
{
    const int lane = threadIdx.x & 63;
    sW[lane] = tt[0];
    sW[64 + lane] = tt[1];
    sW[128 + lane] = tt[2];
    wave_lds_sync();
    wide_sum(sW, nq, 3, sa);
    wave_lds_sync();
    sW[lane] = tt[3];
    sW[64 + lane] = tt[4];
    sW[128 + lane] = tt[5];
    wave_lds_sync();
    wide_sum(sW, nq, 3, sb);
    wave_lds_sync();
}

/* The wide evaluation of the mixed-precision form's one-pass objective
 * (FSX_1P_TERMS): the profile of a fit is read once per evaluation and each
 * q's G is formed once.  Lane i holds q = 64 r + i in round r; the sums add
 * the terms in q order, as fsx_eval_1p_blk does, so the results are
 * bit-identical to the lane-per-fit evaluation's. */
__device__ __forceinline__ void fit_eval_wide_1p(unsigned long long amask, const double *acc,
                                                 int slot, const fsx_fitpar *fp, const double *x,
                                                 double *f, double *g, double *sW)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, Q = fp->qnum;
    const double *qv = fp->q, *av = fp->a, *rdq = fp->rdq;
    const double mult = fp->mult;
    if (Q <= 64) {
        const int qs = lane < Q ? lane : 0;
        const double qc = qv[qs], qp = qs > 0 ? qv[qs - 1] : -1.0;
        const double rq = rdq ? rdq[qs] : 1.0 / (qc - qp);
        const double *ai = av + qs * 6;
        double vn[6];
        auto load = [&](int jj) {
            const double2 *v2 = reinterpret_cast<const double2 *>(
                acc + ((long)__builtin_amdgcn_readlane(slot, jj) * Q + qs) * 6);
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const double2 p = v2[h];
                vn[2 * h] = p.x;
                vn[2 * h + 1] = p.y;
            }
        };
        load(__builtin_ctzll(amask));
        while (amask) {
            const int j = __builtin_ctzll(amask);
            amask &= amask - 1;
            double v[6];
#pragma unroll
            for (int h = 0; h < 6; h++) v[h] = vn[h];
            if (amask) load(__builtin_ctzll(amask));
            const double c1 = lane_d(x[0], j), c2 = lane_d(x[1], j);
            const double corr = -mult * (c1 * c1 - 1.0), c1c = c1 * c1 * c1;
            const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
            const double G = c1c * exp(corr * qc * qc);
            const double Gd = G * (r3c1 - c1m2 * qc * qc);
            double in, d1, d2;
            FSX_OBJ_VALS(v, G, Gd, c2, in, d1, d2);
            double inp = __shfl_up(in, 1), d1p = __shfl_up(d1, 1), d2p = __shfl_up(d2, 1);
            if (lane == 0) {
                inp = in;
                d1p = d1;
                d2p = d2;
            }
            double tt[6];
            FSX_1P_TERMS(in, d1, d2, inp, d1p, d2p, rq, qc, ai, tt);
            double sa = 0.0, sb = 0.0;
            wide_sum6(sW, tt, Q, &sa, &sb);
            const double sm[6] = {lane_d(sa, 0), lane_d(sa, 1), lane_d(sa, 2),
                                  lane_d(sb, 0), lane_d(sb, 1), lane_d(sb, 2)};
            if (lane == j) fsx_1p_finish(sm, fp->a0sum, f, g);
        }
        return;
    }
    while (amask) {
        const int j = __builtin_ctzll(amask);
        amask &= amask - 1;
        const double c1 = lane_d(x[0], j), c2 = lane_d(x[1], j);
        const double *vp = acc + (long)__builtin_amdgcn_readlane(slot, j) * Q * 6;
        const double corr = -mult * (c1 * c1 - 1.0), c1c = c1 * c1 * c1;
        const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
        double in_l = 0.0, d1_l = 0.0, d2_l = 0.0, sa = 0.0, sb = 0.0;
        for (int r0 = 0; r0 < Q; r0 += 64) {
            const int nq = Q - r0 < 64 ? Q - r0 : 64, q = r0 + lane, ok = lane < nq;
            const int qs = ok ? q : r0;    /* a valid q for the unused lanes */
            const double2 *v2 = reinterpret_cast<const double2 *>(vp + (long)qs * 6);
            double v[6];
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const double2 p = v2[h];
                v[2 * h] = p.x;
                v[2 * h + 1] = p.y;
            }
            const double qc = qv[qs], qp = qs > 0 ? qv[qs - 1] : -1.0;
            const double G = c1c * exp(corr * qc * qc);
            const double Gd = G * (r3c1 - c1m2 * qc * qc);
            double in, d1, d2;
            FSX_OBJ_VALS(v, G, Gd, c2, in, d1, d2);
            double inp = __shfl_up(in, 1), d1p = __shfl_up(d1, 1), d2p = __shfl_up(d2, 1);
            if (lane == 0) {
                inp = r0 == 0 ? in : in_l;
                d1p = r0 == 0 ? d1 : d1_l;
                d2p = r0 == 0 ? d2 : d2_l;
            }
            const double rq = rdq ? rdq[qs] : 1.0 / (qc - qp);
            double tt[6];
            FSX_1P_TERMS(in, d1, d2, inp, d1p, d2p, rq, qc, av + qs * 6, tt);
            wide_sum6(sW, tt, nq, &sa, &sb);
            in_l = lane_d(in, 63);
            d1_l = lane_d(d1, 63);
            d2_l = lane_d(d2, 63);
        }
        const double sm[6] = {lane_d(sa, 0), lane_d(sa, 1), lane_d(sa, 2),
                              lane_d(sb, 0), lane_d(sb, 1), lane_d(sb, 2)};
        if (lane == j) fsx_1p_finish(sm, fp->a0sum, f, g);
    }
}

/* Evaluation of the fits of the lanes in amask (all of this wavefront) by
 * the whole wavefront, one fit at a time: lane i holds q = 64 r + i in
 * round r, forms that q's terms (FSX_BS_TERMS, FSX_OBJ_VALS, FSX_OBJ_TERMS,
 * with the previous q's values from its neighbour lane) and writes them to
 * sW; lane c then adds row c to its running sum in q order.  Every value is
 * therefore formed by the operations of the sequential loop and every sum
 * adds the same terms in the same order: the results are bit-identical to
 * lb_eval's.  x, slot: each lane's point and profile slot; f, g: the results
 * of lane j land in lane j. */
template <int ONEP>
__device__ __forceinline__ void fit_eval_wide(unsigned long long amask, const double *acc, int slot,
                              const fsx_fitpar *fp, const double *x, double *f, double *g,
                              double *sW)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63, Q = fp->qnum;
    const double *qv = fp->q, *av = fp->a, *rdq = fp->rdq;
    const double mult = fp->mult;
    if (ONEP) {
        fit_eval_wide_1p(amask, acc, slot, fp, x, f, g, sW);
        return;
    }
    if (Q <= 64) {
        /* one round: each lane's q, its neighbour and 1/(q - q_prev) once;
         * the profile of a fit read once for both passes (the next fit's
         * read issued before this one's sums); G and I of the scale pass
         * reused by the objective */
        const int qs = lane < Q ? lane : 0;
        const double qc = qv[qs], qp = qs > 0 ? qv[qs - 1] : -1.0;
        const double rq = rdq ? rdq[qs] : 1.0 / (qc - qp);
        const double *ai = av + qs * 6;
        double vn[6];
        auto load = [&](int jj) {
            const double2 *v2 = reinterpret_cast<const double2 *>(
                acc + ((long)__builtin_amdgcn_readlane(slot, jj) * Q + qs) * 6);
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const double2 p = v2[h];
                vn[2 * h] = p.x;
                vn[2 * h + 1] = p.y;
            }
        };
        load(__builtin_ctzll(amask));
        while (amask) {
            const int j = __builtin_ctzll(amask);
            amask &= amask - 1;
            double v[6];
#pragma unroll
            for (int h = 0; h < 6; h++) v[h] = vn[h];
            if (amask) load(__builtin_ctzll(amask));
            const double c1 = lane_d(x[0], j), c2 = lane_d(x[1], j);
            const double corr = -mult * (c1 * c1 - 1.0), c1c = c1 * c1 * c1;
            const double G = c1c * exp(corr * qc * qc);
            const double in = FSX_IN(v, G, c2);
            double inp = __shfl_up(in, 1);
            if (lane == 0) inp = in;
            double ut, dt;
            FSX_BS_TERMS(in, inp, qc, qp, ai, ut, dt);
            sW[lane] = ut;
            sW[64 + lane] = dt;
            wave_lds_sync();
            double sum = 0.0;
            wide_sum(sW, Q, 2, &sum);
            wave_lds_sync();
            const double k = lane_d(sum, 0) / lane_d(sum, 1);
            const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
            const double Gd = G * (r3c1 - c1m2 * qc * qc);
            double in2, d1, d2;
            FSX_OBJ_VALS(v, G, Gd, c2, in2, d1, d2);
            double d1p = __shfl_up(d1, 1), d2p = __shfl_up(d2, 1);
            inp = __shfl_up(in2, 1);
            if (lane == 0) {
                /* the values before the loop (min_saxs.c:28-41) */
                inp = FSX_IN(v, G, c2);
                d1p = -Gd * v[1] + 2.0 * G * Gd * v[3] - Gd * c2 * v[4];
                d2p = v[2] - G * v[4] + 2.0 * c2 * v[5];
            }
            double tg0, tg1, tsc;
            FSX_OBJ_TERMS(in2, d1, d2, inp, d1p, d2p, rq, qc, ai, k, tg0, tg1, tsc);
            sW[lane] = tg0;
            sW[64 + lane] = tg1;
            sW[128 + lane] = tsc;
            wave_lds_sync();
            sum = 0.0;
            wide_sum(sW, Q, 3, &sum);
            wave_lds_sync();
            const double g0 = lane_d(sum, 0), g1 = lane_d(sum, 1), sc = lane_d(sum, 2);
            if (lane == j) {
                g[0] = g0;
                g[1] = g1;
                *f = sc;
            }
        }
        return;
    }
    while (amask) {
        const int j = __builtin_ctzll(amask);
        amask &= amask - 1;
        const double c1 = lane_d(x[0], j), c2 = lane_d(x[1], j);
        const double *vp = acc + (long)__builtin_amdgcn_readlane(slot, j) * Q * 6;
        const double corr = -mult * (c1 * c1 - 1.0), c1c = c1 * c1 * c1;
        /* scale pass (min_saxs.c:261-319): up, down in lanes 0, 1 */
        double sum = 0.0, in_last = 0.0;
        for (int r0 = 0; r0 < Q; r0 += 64) {
            const int nq = Q - r0 < 64 ? Q - r0 : 64, q = r0 + lane, ok = lane < nq;
            const int qs = ok ? q : r0;    /* a valid q for the unused lanes */
            const double2 *v2 = reinterpret_cast<const double2 *>(vp + (long)qs * 6);
            double v[6];
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const double2 p = v2[h];
                v[2 * h] = p.x;
                v[2 * h + 1] = p.y;
            }
            const double qc = qv[qs], qp = qs > 0 ? qv[qs - 1] : -1.0;
            const double G = c1c * exp(corr * qc * qc);
            const double in = FSX_IN(v, G, c2);
            double inp = __shfl_up(in, 1);
            if (lane == 0) inp = r0 == 0 ? in : in_last;
            double ut, dt;
            FSX_BS_TERMS(in, inp, qc, qp, av + qs * 6, ut, dt);
            sW[lane] = ut;
            sW[64 + lane] = dt;
            wave_lds_sync();
            wide_sum(sW, nq, 2, &sum);
            wave_lds_sync();
            in_last = lane_d(in, 63);
        }
        const double up = lane_d(sum, 0), down = lane_d(sum, 1);
        const double k = up / down;
        /* objective (min_saxs.c:3-105): g0, g1, score in lanes 0, 1, 2 */
        const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
        double in_l = 0.0, d1_l = 0.0, d2_l = 0.0;
        sum = 0.0;
        for (int r0 = 0; r0 < Q; r0 += 64) {
            const int nq = Q - r0 < 64 ? Q - r0 : 64, q = r0 + lane, ok = lane < nq;
            const int qs = ok ? q : r0;
            const double2 *v2 = reinterpret_cast<const double2 *>(vp + (long)qs * 6);
            double v[6];
#pragma unroll
            for (int h = 0; h < 3; h++) {
                const double2 p = v2[h];
                v[2 * h] = p.x;
                v[2 * h + 1] = p.y;
            }
            const double qc = qv[qs], qp = qs > 0 ? qv[qs - 1] : -1.0;
            const double G = c1c * exp(corr * qc * qc);
            const double Gd = G * (r3c1 - c1m2 * qc * qc);
            double in, d1, d2;
            FSX_OBJ_VALS(v, G, Gd, c2, in, d1, d2);
            double inp = __shfl_up(in, 1), d1p = __shfl_up(d1, 1), d2p = __shfl_up(d2, 1);
            if (lane == 0) {
                if (r0 == 0) {
                    /* the values before the loop (min_saxs.c:28-41) */
                    inp = FSX_IN(v, G, c2);
                    d1p = -Gd * v[1] + 2.0 * G * Gd * v[3] - Gd * c2 * v[4];
                    d2p = v[2] - G * v[4] + 2.0 * c2 * v[5];
                } else {
                    inp = in_l;
                    d1p = d1_l;
                    d2p = d2_l;
                }
            }
            const double rq = rdq ? rdq[qs] : 1.0 / (qc - qp);
            double tg0, tg1, tsc;
            FSX_OBJ_TERMS(in, d1, d2, inp, d1p, d2p, rq, qc, av + qs * 6, k, tg0, tg1, tsc);
            sW[lane] = tg0;
            sW[64 + lane] = tg1;
            sW[128 + lane] = tsc;
            wave_lds_sync();
            wide_sum(sW, nq, 3, &sum);
            wave_lds_sync();
            in_l = lane_d(in, 63);
            d1_l = lane_d(d1, 63);
            d2_l = lane_d(d2, 63);
        }
        const double g0 = lane_d(sum, 0), g1 = lane_d(sum, 1), sc = lane_d(sum, 2);
        if (lane == j) {
            g[0] = g0;
            g[1] = g1;
            *f = sc;
        }
    }
}

template <int AQ_LDS, int ONEP>
__global__ void __launch_bounds__(256)
k_fit(const double *acc, const int4 *recs, const int *nvalid_p, int Q, const double *qvals,
      const double *a, double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
      double *score, double *c1o, double *c2o, int *next, unsigned long long *tally,
      int wide_max)
{
#pragma clang fp contract(off)
    /* LDS: the K blocks [FSX_LB_KN][256], the wide evaluation's sums
     * [4 wavefronts][FSX_WIDE_SUMS][64], then a [6][Q], q [Q] and
     * 1/(q_i - q_{i-1}) [Q] when aq_lds (read by every lane at the same
     * address); otherwise a and q are read from global memory and the
     * reciprocals formed per step */
    extern __shared__ double sKf[];
    double *sW = sKf + FSX_LB_KN * FSX_LB_KSTRIDE + (threadIdx.x >> 6) * FSX_WIDE_SUMS * 64;
    const int nvalid = *nvalid_p;
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    fsx_fitpar fp;
    if (AQ_LDS) {
        double *sA = sKf + FSX_LB_KN * FSX_LB_KSTRIDE + 4 * FSX_WIDE_SUMS * 64;
        double *sQ = sA + 6 * Q, *sR = sQ + Q;
        for (int i = threadIdx.x; i < 6 * Q; i += blockDim.x) sA[i] = a[i];
        for (int i = threadIdx.x; i < Q; i += blockDim.x) {
            const double qc = qvals[i];
            sQ[i] = qc;
            sR[i] = 1.0 / (qc - (i > 0 ? qvals[i - 1] : -1.0));
        }
        __syncthreads();
        fp.a = sA; fp.q = sQ; fp.rdq = sR;
    } else {
        fp.a = a; fp.q = qvals; fp.rdq = NULL;
    }
    fp.qnum = Q; fp.mult = mult;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    fp.onepass = ONEP;
    fp.a0sum = ONEP ? fsx_a0sum(fp.a, Q) : 0.0;
    /* profiles already rescaled in place (k_fit_prep) */
    fsx_ct t;
    t.c = NULL; t.cks = 0; t.cqs = 0;
    t.v = acc; t.vks = 1; t.vqs = 6;
    t.scale = 1.0;
    fsx_lbrc S;
    S.w.kb = sKf + threadIdx.x;
    unsigned long long nev = 0, ncap = 0, steps = 0, wsteps = 0;
    int act = s < nvalid;
    int pose = act ? recs[s].x : 0;
    if (act) fit_begin(&t, &S, &fp, acc, s, Q);
    while (__any(act)) {
        /* once the queue is empty and at most wide_max lanes of the
         * wavefront still fit, the wavefront evaluates their fits together
         * (q across the lanes) instead of each lane running its whole q loop */
        const unsigned long long amask = __ballot(act);
        steps++;
        if (__popcll(amask) <= wide_max) {
            wsteps++;
            fit_eval_wide<ONEP>(amask, acc, s, &fp, S.x, &S.f, S.g, sW);
        } else if (act) {
            if (ONEP) {
                lb_eval_1p(&t, &fp, S.x, &S.f, S.g);
            } else {
                double k;
                lb_eval(&t, &fp, S.x, &S.f, S.g, &k);
            }
        }
        if (act && !fsx_lb_resume(&S)) {
            score[pose] = sqrt(S.f);
            c1o[pose] = S.x[0];
            c2o[pose] = S.x[1];
            nev += S.nev;
            ncap += S.task == FSX_T_STOP_ITER;
            s = atomicAdd(next, 1);
            act = s < nvalid;
            if (act) {
                pose = recs[s].x;
                fit_begin(&t, &S, &fp, acc, s, Q);
            }
        }
    }
    if (nev) atomicAdd(&tally[0], nev);
    if (ncap) atomicAdd(&tally[1], ncap);
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(&tally[2], steps);
        atomicAdd(&tally[3], wsteps);
        atomicMax(&tally[4], steps);
    }
}

/* ---------------------------------------------------------------------- */
/* The mixed-precision form's fit: the two-variable L-BFGS-B of fsx_lb2.h
 * (the reference's steps with the 2 x 2 L-BFGS matrix held explicitly),
 * one pose per group of fpl lanes (a power of two).  Lane j of a group
 * evaluates the q range [j qpl, (j + 1) qpl), qpl = ceil(Q / fpl), reading
 * the pose's profile at each evaluation (one-pass form, fsx_lbfgsb.h
 * FSX_1P_TERMS): its six sums run in q order from its second q on, its
 * first q pairs with the previous lane's last (a lane shift), and the sums
 * meet by an xor butterfly, after which every lane of the group holds the
 * same value and gradient and runs the same solver step.  Between steps the
 * solver state lives in LDS, one copy per group (every lane of the group
 * writes the same values), so that the kernel fits in 128 registers with
 * 144 bytes of spills (reloaded once per evaluation or per fit, none in the
 * loop over q): four wavefronts per SIMD, against one for the
 * compact form's state (on the 210,000-pose L = 15 workload 1.50 ms,
 * against 1.60 ms at three wavefronts in 168 registers without spills).
 * With four lanes per fit (Q up to 64) a wavefront runs 16 fits, so the
 * solver's divergent code is shared by many fits while the evaluation of
 * each is split four ways (at three wavefronts per SIMD: 1.64 ms, against
 * 1.83 ms with two lanes and 1.87 ms with eight).  The moments, q and the
 * step reciprocals sit in LDS, slot
 * ii * fpl + j for q = j qpl + ii, so the lanes of a group read consecutive
 * doubles.  Groups take poses from a queue (next) until none is left. */
#define FSX_FG_ARR 8     /* LDS arrays: a0..a5, q, 1 / (q_i - q_{i-1}) */
#define FSX_FG_QLANE 16  /* q per lane the default lanes per fit aim at */
#ifndef FSX_FG_WAVES
#define FSX_FG_WAVES 4   /* wavefronts per SIMD (the register budget of the launch bounds) */
#endif

__device__ __forceinline__ void fitg_eval(const float *pr, const double *sNt, int ns, int qpl,
                                          int fpl, int j, int Q, double mult, double a0sum,
                                          const double *x, double *fv, double *g)
{
#pragma clang fp contract(fast)
    const double c1 = x[0], c2 = x[1];
    const double corr = -mult * (c1 * c1 - 1.0), c1c = c1 * c1 * c1;
    const int q0 = j * qpl, nq = Q - q0 < qpl ? Q - q0 : qpl;
    double s[6];
#pragma unroll
    for (int k = 0; k < 6; k++) s[k] = 0.0;
    double vf[3] = {0.0, 0.0, 0.0}, vp[3] = {0.0, 0.0, 0.0};
    /* the sums' terms of step ii from I and its derivatives there (vv) and
     * at the previous q (pp) */
    auto terms_at = [&](const double *vv, const double *pp, int ii) {
        const int slot = ii * fpl + j;
        double ai[6], tt[6];
#pragma unroll
        for (int k = 0; k < 6; k++) ai[k] = sNt[k * ns + slot];
        FSX_1P_TERMS(vv[0], vv[1], vv[2], pp[0], pp[1], pp[2], sNt[7 * ns + slot],
                     sNt[6 * ns + slot], ai, tt);
#pragma unroll
        for (int k = 0; k < 6; k++) s[k] += tt[k];
    };
    for (int ii = 0; ii < nq; ii++) {
        const float2 *p2 = reinterpret_cast<const float2 *>(pr + 6 * (q0 + ii));
        const float2 a = p2[0], b = p2[1], c = p2[2];
        const double cur[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
        /* I and its derivatives in c1, c2 (min_saxs.c:52-68) */
        const double qc = sNt[6 * ns + ii * fpl + j];
        const double G = c1c * exp(corr * qc * qc);
        const double Gd = G * (3.0 / c1 - 2.0 * c1 * mult * qc * qc);
        double vv[3];
        FSX_OBJ_VALS(cur, G, Gd, c2, vv[0], vv[1], vv[2]);
        if (ii == 0) {
#pragma unroll
            for (int k = 0; k < 3; k++) vf[k] = vv[k];
        } else {
            terms_at(vv, vp, ii);
        }
#pragma unroll
        for (int k = 0; k < 3; k++) vp[k] = vv[k];
    }
    /* the first q of a lane pairs with the previous lane's last (lane 0:
     * q = 0 with itself, slope 0, as min_saxs.c:30-35 starts) */
    double pl[3];
#pragma unroll
    for (int k = 0; k < 3; k++) pl[k] = __shfl_up(vp[k], 1, fpl);
    if (nq > 0) terms_at(vf, j == 0 ? vf : pl, 0);
    for (int o = 1; o < fpl; o <<= 1) {
#pragma unroll
        for (int k = 0; k < 6; k++) s[k] += __shfl_xor(s[k], o, fpl);
    }
    fsx_1p_finish(s, a0sum, fv, g);
}

__device__ __forceinline__ void
fitg_body(const float *acc, const int4 *recs, const int *nvalid_p, int Q, const double *qvals,
          const double *a, double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
          double *score, double *c1o, double *c2o, int *next, unsigned long long *tally, int fpl)
{
    extern __shared__ double sNt[];
    const int qpl = (Q + fpl - 1) / fpl;
    const int ns = qpl * fpl;
    fsx_lb2 *sSt = reinterpret_cast<fsx_lb2 *>(sNt + FSX_FG_ARR * ns);
    for (int i = threadIdx.x; i < ns; i += blockDim.x) {
        const int q = (i % fpl) * qpl + i / fpl;
        const bool ok = q < Q;
#pragma unroll
        for (int k = 0; k < 6; k++) sNt[k * ns + i] = ok ? a[6 * q + k] : 0.0;
        sNt[6 * ns + i] = ok ? qvals[q] : 0.0;
        /* min_saxs.c:60: 1 / (q_i - q_{i-1}) with q_{-1} = -1 */
        sNt[7 * ns + i] = ok ? 1.0 / (qvals[q] - (q > 0 ? qvals[q - 1] : -1.0)) : 0.0;
    }
    __syncthreads();
    fsx_fitpar fp;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    const double a0sum = fsx_a0sum(a, Q);
    const int nvalid = *nvalid_p;
    const int lane = threadIdx.x & (FSX_WAVE - 1), j = lane & (fpl - 1);
    const int leader = lane & ~(fpl - 1);
    int s = (int)(((long)blockIdx.x * blockDim.x + threadIdx.x) / fpl);
    unsigned long long nev = 0, ncap = 0, nmax = 0;
    fsx_lb2 *S = sSt + threadIdx.x / fpl;
    while (s < nvalid) {
        const float *pr = acc + (long)s * 6 * Q;
        fsx_lb2_start(S, &fp, 1.0, 0.0);   /* define.h:33-34 */
        int more;
        do {
            double xt[2] = {S->x[0], S->x[1]}, fv, gv[2];
            /* the state is read back from LDS after the evaluation rather
             * than kept in registers across it */
            asm volatile("" ::: "memory");
            fitg_eval(pr, sNt, ns, qpl, fpl, j, Q, mult, a0sum, xt, &fv, gv);
            asm volatile("" ::: "memory");
            S->f = fv;
            S->g[0] = gv[0];
            S->g[1] = gv[1];
            more = fsx_lb2_resume(S, &fp);
        } while (more);
        if (j == 0) {
            const int o = recs ? recs[s].x : s;
            score[o] = sqrt(S->f);
            c1o[o] = S->x[0];
            c2o[o] = S->x[1];
        }
        nev += S->nev;
        ncap += S->task == FSX_T_STOP_ITER;
        nmax = nmax > (unsigned long long)S->nev ? nmax : (unsigned long long)S->nev;
        int sn = 0;
        if (j == 0) sn = atomicAdd(next, 1);
        s = __shfl(sn, leader, FSX_WAVE);
    }
    if (j == 0) {
        /* evaluations; fits stopped at the evaluation cap; evaluations again
         * as the loop count of the lane groups; none wavefront-wide; the
         * longest fit's evaluations */
        if (nev) {
            atomicAdd(&tally[0], nev);
            atomicAdd(&tally[2], nev);
        }
        if (ncap) atomicAdd(&tally[1], ncap);
        if (nmax) atomicMax(&tally[4], nmax);
    }
}

__global__ void __launch_bounds__(256, FSX_FG_WAVES)
k_fitg(const float *acc, const int4 *recs, const int *nvalid_p, int Q, const double *qvals,
       const double *a, double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
       double *score, double *c1o, double *c2o, int *next, unsigned long long *tally, int fpl)
{
    fitg_body(acc, recs, nvalid_p, Q, qvals, a, mult, c1_lo, c1_hi, c2_lo, c2_hi, score, c1o,
              c2o, next, tally, fpl);
}

/* The fits of several complexes in one persistent launch (the batched
 * form): every block runs the lane groups of job 0 until its pose queue is
 * empty, then those of job 1, and so on (each job with its own queue, q
 * grid, moments, bounds and lanes per fit). */
struct fsx_fit_job {
    const float *acc;
    const double *qvals, *a;
    const int4 *recs;
    const int *nvalid_p;
    double mult, c1_lo, c1_hi, c2_lo, c2_hi;
    double *score, *c1o, *c2o;
    int *next;
    unsigned long long *tally;
    int Q, fpl;
};

__global__ void __launch_bounds__(256, FSX_FG_WAVES)
k_fitg_jobs(const fsx_fit_job *jobs, int njobs)
{
    for (int k = 0; k < njobs; k++) {
        const fsx_fit_job &jb = jobs[k];
        fitg_body(jb.acc, jb.recs, jb.nvalid_p, jb.Q, jb.qvals, jb.a, jb.mult, jb.c1_lo,
                  jb.c1_hi, jb.c2_lo, jb.c2_hi, jb.score, jb.c1o, jb.c2o, jb.next, jb.tally,
                  jb.fpl);
        /* the next job refills the LDS arrays */
        __syncthreads();
    }
}

/* LDS of a k_fitg block: the moment arrays and one solver state per group */
static inline size_t fitg_lds(int Q, int fpl)
{
    const size_t qpl = ((size_t)Q + fpl - 1) / fpl;
    return (size_t)FSX_FG_ARR * qpl * fpl * sizeof(double) + (size_t)(256 / fpl) * sizeof(fsx_lb2);
}

/* fits of caller-supplied (rescaled) profiles ct[n][6][Q] */
__global__ void __launch_bounds__(256)
k_fit_batch(const double *ct, int n, const double *qvals, int Q, const double *a, double mult,
            double c1_lo, double c1_hi, double c2_lo, double c2_hi, double *out)
{
#pragma clang fp contract(off)
    extern __shared__ double sKb[];   /* K blocks [FSX_LB_KN][256] */
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    fsx_ct t;
    t.c = ct + (long)s * 6 * Q; t.cks = Q; t.cqs = 1;
    t.v = NULL; t.vks = 0; t.vqs = 0; t.scale = 1.0;
    fsx_fitpar fp;
    fp.a = a; fp.q = qvals; fp.rdq = NULL; fp.qnum = Q; fp.mult = mult;
    fp.lo0 = c1_lo; fp.hi0 = c1_hi; fp.lo1 = c2_lo; fp.hi1 = c2_hi;
    fp.factr = 1e+7; fp.pgtol = 1e-5;
    fp.onepass = 0; fp.a0sum = 0.0;
    double x[2] = {1.0, 0.0}, f;
    int nev;
    fsx_lbfgsb_fit(&t, &fp, x, &f, &nev, sKb + threadIdx.x);
    out[4L * s] = sqrt(f);
    out[4L * s + 1] = x[0];
    out[4L * s + 2] = x[1];
    out[4L * s + 3] = nev;
}

/* ---- atom expansion (pdb2spf.c:24-152) ------------------------------------
 * A^X_lm(q) = 4 pi i^l sum_a f^X_a j_l(q r_a) conj(Y_lm(theta_a, phi_a)) for
 * X = V, D, W (f^V, f^D the scalar form factors, f^W = zero_ff(OH2) * SASA).
 * With real f, conj(Y_{l,-m}) = (-1)^m Y_lm, so the sums are formed for
 * m >= 0 only and A_{l,-m} = (-1)^m 4 pi i^l conj(S_lm), S_lm the m >= 0 sum.
 * Stages, per pass over a range of atoms:
 *   k_atom_geom   r, f^X and conj(Y_lm) (m >= 0) per atom, the reference's
 *                 angles (cart2sph), norms and Legendre recurrence
 *                 (borrowed.c:19-65), contraction off
 *   k_atom_bes    J[l][atom][q] = j_l(q r) by the reference's series, q padded
 *                 to a multiple of 16 with zeros
 *   k_atom_mfma   per (chunk of 1024 atoms, l, 16 q): the sums over the chunk's
 *                 atoms as FP64 MFMA products of the (q x atoms) Bessel tile and
 *                 the (atoms x (X, m, re/im)) tile f^X conj(Y_lm); four waves
 *                 take 256 atoms each and are summed in LDS in a fixed order;
 *                 one partial per chunk
 * then k_atom_finish sums the chunks in order and applies 4 pi i^l. */
#define FSX_AT_CHUNK 1024   /* atoms per k_atom_mfma block (4 waves x 256) */
#define FSX_AT_MAXT 16      /* column tiles: 3 * 2 (L + 1) <= 16 * 16 for L <= 41 */

__global__ void k_atom_geom(const double *xyz, const double *vac, const double *dum,
                            const double *sa, double h2o_ff, int na, int L, double pi_ref,
                            double *r_out, double *f_out, double *Y)
{
#pragma clang fp contract(off)
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= na) return;
    const int N = L + 1, Kh = N * (N + 1) / 2;
    /* cart2sph (pdb2spf.c:9-22) */
    const double x = xyz[3 * a], y = xyz[3 * a + 1], z = xyz[3 * a + 2];
    const double rr = sqrt(x * x + y * y + z * z);
    const double th = acos(z / rr);
    double fi;
    if (y > 0.0) fi = acos(x / sqrt(x * x + y * y));
    else fi = -acos(x / sqrt(x * x + y * y)) + 2.0 * pi_ref;
    r_out[a] = rr;
    f_out[a] = vac[a];
    f_out[na + a] = dum[a];
    f_out[2 * na + a] = sa ? h2o_ff * sa[a] : 0.0;
    /* associated Legendre values (borrowed.c:40-65), written as conj(Y_lm)
     * = yn(l, m) P_l^m (cos m phi, sin(-m phi)) as soon as each is formed;
     * yn(l, m) = sqrt((2l + 1) / (4 pi) / prod_{k=1..m} (l - k + 1)(l + k))
     * (borrowed.c:19-38) */
    double *Ya = Y + (size_t)a * Kh * 2;
    const double xc = cos(th), yc = sqrt(1. - xc * xc);
    auto put = [&](int l, int m, double P) {
        double f = 1.;
        const double base = (2 * l + 1.) / (4. * pi_ref);
        for (int k = 1; k <= m; k++) f *= (l - k + 1) * (l + k);
        const double h = sqrt(base / f) * P;
        const int e = l * (l + 1) / 2 + m;
        Ya[2 * e] = h * cos(fi * m);
        Ya[2 * e + 1] = h * sin(-fi * m);
    };
    /* m = 0 column */
    double p2 = 1, p1 = xc;
    put(0, 0, 1.0);
    if (N > 1) put(1, 0, xc);
    for (int l = 2; l < N; ++l) {
        const double p = (2. * l - 1.) / l * xc * p1 - (l - 1.) / l * p2;
        put(l, 0, p);
        p2 = p1;
        p1 = p;
    }
    /* m >= 1: the diagonal from the previous one, then upwards in l */
    double pmm = 1.0;
    for (int m = 1; m < N - 1; ++m) {
        pmm = -1. * (2. * m - 1.) * yc * pmm;
        put(m, m, pmm);
        double q2 = pmm, q1 = (2. * m + 1.) * xc * pmm;
        put(m + 1, m, q1);
        for (int l = m + 2; l < N; ++l) {
            const double p = (2. * l - 1.) / (l - m) * xc * q1 - (l + m - 1.) / (l - m) * q2;
            put(l, m, p);
            q2 = q1;
            q1 = p;
        }
    }
    if (N > 1) put(N - 1, N - 1, -1. * (2. * (N - 1.) - 1.) * yc * pmm);
}

/* J[l][a][q] = j_l(q r_a) for q < Q, 0 for Q <= q < Qp */
__global__ void k_atom_bes(double *J, const double *r, const double *qvals, int na, int Q, int Qp,
                           int K)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (long)K * na * Qp) return;
    const int q = (int)(i % Qp);
    const long la = i / Qp;
    const int a = (int)(la % na), l = (int)(la / na);
    J[i] = q < Q ? dev_sbessel(l, qvals[q] * r[a]) : 0.0;
}

/* One block of four waves per (chunk, l, 16-q tile).  Rows are q, columns
 * c = X * 2(l+1) + 2m + part (part 0: real, 1: imaginary) of f^X conj(Y_lm).
 * FP64 MFMA 16x16x4 operand layout on gfx950: lane i feeds A[i % 16][i / 16]
 * and B[i / 16][i % 16]; result register j of lane i is D[i / 16 + 4j][i % 16].
 * Writes the chunk's partial sums P[chunk][X][Qp][Kh][2]. */
__global__ void __launch_bounds__(256)
k_atom_mfma(const double *J, const double *f, const double *Y, double *P, int na, int Qp, int L,
            int chunk0)
{
    __shared__ fsx_v4d red[3][64];
    const int K = L + 1, Kh = K * (K + 1) / 2, nqt = Qp / 16;
    int bid = blockIdx.x;
    const int qt = bid % nqt;
    bid /= nqt;
    const int l = bid % K;
    const int ch = bid / K;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ncol = 3 * 2 * (l + 1), nct = (ncol + 15) / 16;
    /* this lane's column of each tile: X, entry offset and part */
    int cx[FSX_AT_MAXT], coff[FSX_AT_MAXT];
#pragma unroll
    for (int t = 0; t < FSX_AT_MAXT; t++) {
        const int c = t * 16 + (lane & 15);
        const int X = c / (2 * (l + 1)), rem = c - X * 2 * (l + 1);
        cx[t] = c < ncol ? X : -1;
        coff[t] = c < ncol ? l * (l + 1) / 2 * 2 + rem : 0;   /* (entry, part) as a double index */
    }
    fsx_v4d acc[FSX_AT_MAXT];
#pragma unroll
    for (int t = 0; t < FSX_AT_MAXT; t++) acc[t] = (fsx_v4d){0.0, 0.0, 0.0, 0.0};
    const int abeg = ch * FSX_AT_CHUNK + w * (FSX_AT_CHUNK / 4);
    const int aend = min(abeg + FSX_AT_CHUNK / 4, na);
    const double *Jl = J + (size_t)l * na * Qp + qt * 16 + (lane & 15);
    for (int a0 = abeg; a0 < aend; a0 += 4) {
        const int a = a0 + (lane >> 4);
        const bool ok = a < aend;
        const double av = ok ? Jl[(size_t)a * Qp] : 0.0;
        const double *Ya = Y + (size_t)(ok ? a : 0) * Kh * 2;
#pragma unroll
        for (int t = 0; t < FSX_AT_MAXT; t++) {
            if (t < nct) {
                const double bv = ok && cx[t] >= 0 ? f[(size_t)cx[t] * na + a] * Ya[coff[t]] : 0.0;
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
            }
        }
    }
    /* wave 0 + wave 1 + wave 2 + wave 3, in that order, per tile */
    const int q0 = qt * 16;
    double *Pc = P + (size_t)(chunk0 + ch) * 3 * Qp * Kh * 2;
#pragma unroll
    for (int t = 0; t < FSX_AT_MAXT; t++) {
        if (t < nct) {
            if (w > 0) red[w - 1][lane] = acc[t];
            __syncthreads();
            if (w == 0) {
                fsx_v4d v = acc[t];
                v += red[0][lane];
                v += red[1][lane];
                v += red[2][lane];
                if (cx[t] >= 0) {
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int q = q0 + (lane >> 4) + 4 * j;
                        Pc[((size_t)cx[t] * Qp + q) * Kh * 2 + coff[t]] = v[j];
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* ---- the atom expansion in split BF16 on MFMA (the mixed form) ----------
 * The same sums over atoms as k_atom_mfma, each FP64 operand x written as
 * three BF16 parts x = h + m + l + O(2^-24 x) (h = bf16(x), m = bf16(x - h),
 * l = bf16(x - h - m)), and every product as the six part products of at
 * least that order (hh, hm, mh, hl, lh, mm) on v_mfma_f32_16x16x32_bf16 with
 * FP32 accumulation; the chunk partials go to FP64 and k_atom_finish sums
 * them in order.  The parts are formed once per value by k_atom_bes_split
 * (j_l(q r), [part][l][q][atom]) and k_atom_bsplit (f^X conj(Y_lm),
 * [part][column][atom]), atoms innermost so that a lane's 8 atoms of a step
 * are one 16-byte load; atoms are padded to whole chunks with zeros. */
typedef __bf16 fsx_bf16x8 __attribute__((ext_vector_type(8)));
typedef float fsx_v4f __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void split3(double x, __bf16 &h, __bf16 &m, __bf16 &l)
{
    h = (__bf16)(float)x;
    const double r1 = x - (double)(float)h;
    m = (__bf16)(float)r1;
    const double r2 = r1 - (double)(float)m;
    l = (__bf16)(float)r2;
}

/* Js[part][l][q][a] = parts of j_l(q r_a), 0 for q >= Q or a >= na */
__global__ void k_atom_bes_split(__bf16 *Js, const double *r, const double *qvals, int na, int nap,
                                 int Q, int Qp, int K)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)K * Qp * nap;
    if (i >= tot) return;
    const int a = (int)(i % nap);
    const long lq = i / nap;
    const int q = (int)(lq % Qp), l = (int)(lq / Qp);
    const double v = q < Q && a < na ? dev_sbessel(l, qvals[q] * r[a]) : 0.0;
    __bf16 h, m, lo;
    split3(v, h, m, lo);
    Js[i] = h;
    Js[tot + i] = m;
    Js[2 * tot + i] = lo;
}

/* Bs[part][3 l (l+1) + c][a] = parts of f^X_a conj(Y_lm)(a) for column c =
 * X * 2 (l+1) + 2 m + (0 real, 1 imaginary) of order l */
__global__ void k_atom_bsplit(__bf16 *Bs, const double *f, const double *Y, int na, int nap, int L)
{
    const int K = L + 1, Kh = K * (K + 1) / 2, ncolt = 6 * Kh;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)ncolt * nap;
    if (i >= tot) return;
    const int a = (int)(i % nap), cg = (int)(i / nap);
    int l = 0;
    while (3 * (l + 1) * (l + 2) <= cg) l++;
    const int c = cg - 3 * l * (l + 1);
    const int X = c / (2 * (l + 1)), rem = c - X * 2 * (l + 1);
    const double v = a < na ? f[(size_t)X * na + a] * Y[(size_t)a * Kh * 2 + l * (l + 1) / 2 * 2 + rem]
                            : 0.0;
    __bf16 h, m, lo;
    split3(v, h, m, lo);
    Bs[i] = h;
    Bs[tot + i] = m;
    Bs[2 * tot + i] = lo;
}

/* One block of four waves per (chunk, l, 16-q tile), as k_atom_mfma.  BF16
 * MFMA 16x16x32 operand layout on gfx950: lane i feeds A[i % 16][8 (i / 16)
 * + e] and B[8 (i / 16) + e][i % 16], e = 0..7; result register j of lane i
 * is D[4 (i / 16) + j][i % 16].  Writes the chunk's partial sums
 * P[chunk][X][Qp][Kh][2] in FP64. */
__global__ void __launch_bounds__(256)
k_atom_bf16(const __bf16 *Js, const __bf16 *Bs, double *P, int nap, int Qp, int L, int chunk0)
{
    __shared__ fsx_v4f red[3][64];
    const int K = L + 1, Kh = K * (K + 1) / 2, nqt = Qp / 16;
    int bid = blockIdx.x;
    const int qt = bid % nqt;
    bid /= nqt;
    const int l = bid % K;
    const int ch = bid / K;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int ncol = 3 * 2 * (l + 1), nct = (ncol + 15) / 16;
    const long jtot = (long)K * Qp * nap, btot = (long)6 * Kh * nap;
    fsx_v4f acc[FSX_AT_MAXT];
#pragma unroll
    for (int t = 0; t < FSX_AT_MAXT; t++) acc[t] = (fsx_v4f){0.f, 0.f, 0.f, 0.f};
    const int abeg = ch * FSX_AT_CHUNK + w * (FSX_AT_CHUNK / 4);
    const __bf16 *Ja = Js + ((long)l * Qp + qt * 16 + (lane & 15)) * nap + 8 * (lane >> 4);
    const __bf16 *Bl = Bs + (long)3 * l * (l + 1) * nap + 8 * (lane >> 4);
    for (int a0 = abeg; a0 < abeg + FSX_AT_CHUNK / 4; a0 += 32) {
        const fsx_bf16x8 ah = *reinterpret_cast<const fsx_bf16x8 *>(Ja + a0);
        const fsx_bf16x8 am = *reinterpret_cast<const fsx_bf16x8 *>(Ja + jtot + a0);
        const fsx_bf16x8 al = *reinterpret_cast<const fsx_bf16x8 *>(Ja + 2 * jtot + a0);
#pragma unroll
        for (int t = 0; t < FSX_AT_MAXT; t++) {
            if (t < nct) {
                const int c = t * 16 + (lane & 15);
                fsx_bf16x8 bh = {}, bm = {}, bl = {};
                if (c < ncol) {
                    const __bf16 *Bc = Bl + (long)c * nap + a0;
                    bh = *reinterpret_cast<const fsx_bf16x8 *>(Bc);
                    bm = *reinterpret_cast<const fsx_bf16x8 *>(Bc + btot);
                    bl = *reinterpret_cast<const fsx_bf16x8 *>(Bc + 2 * btot);
                }
                fsx_v4f v = acc[t];
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, v, 0, 0, 0);
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, v, 0, 0, 0);
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, v, 0, 0, 0);
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, v, 0, 0, 0);
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, v, 0, 0, 0);
                v = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, v, 0, 0, 0);
                acc[t] = v;
            }
        }
    }
    /* wave 0 + wave 1 + wave 2 + wave 3, in that order, per tile */
    const int q0 = qt * 16;
    double *Pc = P + (size_t)(chunk0 + ch) * 3 * Qp * Kh * 2;
#pragma unroll
    for (int t = 0; t < FSX_AT_MAXT; t++) {
        if (t < nct) {
            if (w > 0) red[w - 1][lane] = acc[t];
            __syncthreads();
            if (w == 0) {
                const int c = t * 16 + (lane & 15);
                const int X = c / (2 * (l + 1)), rem = c - X * 2 * (l + 1);
                double v[4];
#pragma unroll
                for (int j = 0; j < 4; j++)
                    v[j] = (double)acc[t][j] + (double)red[0][lane][j] + (double)red[1][lane][j] +
                           (double)red[2][lane][j];
                if (c < ncol) {
                    const int coff = l * (l + 1) / 2 * 2 + rem;
#pragma unroll
                    for (int j = 0; j < 4; j++) {
                        const int q = q0 + 4 * (lane >> 4) + j;
                        Pc[((size_t)X * Qp + q) * Kh * 2 + coff] = v[j];
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* A^X_lm = 4 pi i^l S_lm and A^X_{l,-m} = (-1)^m 4 pi i^l conj(S_lm), S the
 * sum of the nch chunk partials in chunk order; one thread per (X, q, l, m >= 0) */
__global__ void k_atom_finish(const double *P, int nch, double *re, double *im, int Q, int Qp,
                              int L, double four_pi)
{
#pragma clang fp contract(off)
    const int K = L + 1, Kh = K * (K + 1) / 2, K2 = K * K;
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 3L * Q * Kh) return;
    const int e = (int)(i % Kh);
    const int xq = (int)(i / Kh), q = xq % Q, X = xq / Q;
    int l = 0;
    while ((l + 1) * (l + 2) / 2 <= e) l++;
    const int m = e - l * (l + 1) / 2;
    double sr = 0.0, si = 0.0;
    const size_t off = (((size_t)X * Qp + q) * Kh + e) * 2, stride = (size_t)3 * Qp * Kh * 2;
    for (int c = 0; c < nch; c++) {
        sr += P[c * stride + off];
        si += P[c * stride + off + 1];
    }
    const double cr[4] = {four_pi, 0.0, -four_pi, 0.0}, ci[4] = {0.0, four_pi, 0.0, -four_pi};
    const double c_r = cr[l % 4], c_i = ci[l % 4];
    const size_t o = ((size_t)X * Q + q) * K2;
    re[o + l * (l + 1) + m] = c_r * sr - c_i * si;
    im[o + l * (l + 1) + m] = c_r * si + c_i * sr;
    if (m > 0) {
        const double sg = (m & 1) ? -1.0 : 1.0, mr = sg * sr, mi = -sg * si;
        re[o + l * (l + 1) - m] = c_r * mr - c_i * mi;
        im[o + l * (l + 1) - m] = c_r * mi + c_i * mr;
    }
}

/* ---------------------------------------------------------------------- */
/* host side: device buffers, stream, launches */

static pthread_mutex_t g_mtx = PTHREAD_MUTEX_INITIALIZER;
static int g_init = 0;
static hipStream_t g_stream;

/* Working memory of the allocation-free atom expansion
 * (fsx_atoms2spf_noalloc): a fixed device array, one copy per GPU, held until
 * the process ends.  In host code the name scratch is only a host-side
 * stand-in for the symbol: kernels and copies get g_scratch, its device
 * address on the library's device, taken once by hipGetSymbolAddress in
 * init_once. */
#define FSX_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
static __device__ __align__(256) unsigned char scratch[FSX_NOALLOC_SCRATCH_BYTES];
static void *g_scratch = NULL;
enum {
    B_BES, B_T, B_RT, B_BT, B_A, B_B, B_DW, B_DS, B_TWA, B_TWG, B_Q, B_Z, B_AM, B_CINT,
    B_POSES, B_KEYS, B_COUNTS, B_OFFS, B_FILL, B_RECS, B_INFO, B_ACC, B_SCORE, B_C1, B_C2,
    B_CROSSOUT, B_ENT, B_COFF, B_NEXT, B_TALLY, B_SRCT, B_TW2, B_FITCT, B_FITOUT, B_BX, B_BOUT, B_AXYZ,
    B_AFF, B_AWS, B_AQ, B_ARE, B_AIM, B_RTMP, B_FITSC, B_ACCF, B_IN, B_NBUF
};
static void *g_buf[B_NBUF];
static size_t g_cap[B_NBUF];
/* content key of what a buffer holds (fsx_hip_args dw_key / ds_key); 0 =
 * unknown.  Reset whenever the buffer is (re)allocated or released. */
static unsigned long long g_key[B_NBUF];

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
    g_key[slot] = 0;
    size_t cap = bytes + bytes / 8;
    if (hipMalloc(&g_buf[slot], cap) != hipSuccess) {
        g_buf[slot] = NULL;
        return FSX_ERR_NOMEM;
    }
    g_cap[slot] = cap;
    return FSX_OK;
}

static int g_ncu = 256;
static int g_nxcd = 1;   /* XCDs (accelerator complex dies): workgroups go round-robin */

/* launch-shape knobs: environment override for tuning, defaults otherwise */
static int env_int(const char *name, int dflt, int lo, int hi)
{
    const char *v = getenv(name);
    if (!v || !*v) return dflt;
    int x = atoi(v);
    return x < lo ? lo : (x > hi ? hi : x);
}

static int init_once(void)
{
    if (g_init) return FSX_OK;
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess || nd < 1) return FSX_ERR_NODEV;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&g_ncu, hipDeviceAttributeMultiprocessorCount, dev));
    if (g_ncu < 1) g_ncu = 1;
    if (hipDeviceGetAttribute(&g_nxcd, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess ||
        g_nxcd < 1)
        g_nxcd = 1;
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
    HIPCHK(hipGetSymbolAddress(&g_scratch, HIP_SYMBOL(scratch)));
    g_init = 1;
    return FSX_OK;
}

extern "C" int fsx_hip_device_count(void)
{
    int nd = 0;
    if (hipGetDeviceCount(&nd) != hipSuccess) return 0;
    return nd;
}

static void batch_release(void);
static void stage_release(void);
extern "C" void fsx_hip_release(void)
{
    pthread_mutex_lock(&g_mtx);
    batch_release();
    stage_release();
    for (int i = 0; i < B_NBUF; i++) {
        if (g_buf[i]) (void)hipFree(g_buf[i]);
        g_buf[i] = NULL;
        g_cap[i] = 0;
        g_key[i] = 0;
    }
    pthread_mutex_unlock(&g_mtx);
}

static inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

#define RC(x)                          \
    do {                               \
        int r_ = (x);                  \
        if (r_ != FSX_OK) return r_;   \
    } while (0)

/* launches of the element-type dependent stages (FP64 or FP32 tables);
 * g: the buffers of the plan being launched (indices P_*) */
enum { G_DW = 7, G_ENT = 23, G_COFF = 25 };
template <typename E>
static void launch_rotate(hipStream_t s, void *out, void *tmp, const cplx *srct, long nrot,
                          size_t lds, int nb, int ng, int Q, int L, int gc, int C,
                          const int *csz, const int *cst, const long *coff, int mlo, int npk0,
                          int lrot, void *const *g)
{
    hipLaunchKernelGGL(k_rotate<E>, dim3((unsigned)nrot), dim3(256), lds, s, (E *)tmp, srct,
                       (const double *)g[G_DW], (const int *)g[G_ENT],
                       (const long *)g[G_COFF], nb, ng, Q, L, gc, mlo);
    /* 64-row tiles of FP32 tables, 32-row tiles of FP64 tables (33 KB of LDS) */
    const int TR = sizeof(E) == 8 ? 64 : 32;
    const int nrows = nb * ng;
    for (int c = 0; c < C; c++) {
        if (csz[c] == 0) continue;
        const int nte = (csz[c] + 63) / 64, ntr = (nrows + TR - 1) / TR;
        const long nblkT = (long)Q * 3 * nte * ntr;
        if (sizeof(E) == 8)
            hipLaunchKernelGGL((k_tab_transpose<E, 64>), dim3((unsigned)nblkT), dim3(256), 0, s,
                               (E *)out + coff[c], (const E *)tmp + coff[c], csz[c], nrows, nte, ntr,
                               c == 0 ? npk0 : 0, (const int *)g[G_ENT] + cst[c], lrot);
        else
            hipLaunchKernelGGL((k_tab_transpose<E, 32>), dim3((unsigned)nblkT), dim3(256), 0, s,
                               (E *)out + coff[c], (const E *)tmp + coff[c], csz[c], nrows, nte, ntr,
                               c == 0 ? npk0 : 0, (const int *)g[G_ENT] + cst[c], lrot);
    }
}

/* plan buffer indices the cross-term launch reads (kept equal to the P_*
 * enumeration below) */
enum { G_T = 1, G_RT = 2, G_BT = 3, G_COUNTS = 17, G_OFFS = 18, G_RECS = 20, G_ACC = 21,
       G_TW2 = 24 };
template <typename E, int G>
static void launch_cross_g(hipStream_t s, int C, const int *csz, const int *cst,
                           const long *coff, int Tn, long grid, int nrows, int Q, int L,
                           int nb, int npk0, int ng, int zn, void *const *g)
{
    const int tst = (long)Tn * (long)(sizeof(E) / 2) <= 40 * 1024;
    const long ntiles = (long)Q * nb * ((ng + G - 1) / G);
    for (int c = 0; c < C; c++) {
        if (csz[c] == 0) continue;
        const int npk = c == 0 ? npk0 : 0;
        const int ns = csz[c] - npk;   /* pairing slots */
        /* FP32 tables: two poses per wavefront (32 lanes each) when the
         * slots fit four per lane (FSX_CROSS_PW=1: one pose per wavefront) */
        const int pw = sizeof(E) == 8 && ns <= 4 * (FSX_WAVE / 2) &&
                       env_int("FSX_CROSS_PW", 2, 1, 2) == 2 ? 2 : 1;
        int jn = (ns + FSX_WAVE / pw - 1) / (FSX_WAVE / pw);
        void (*kern)(double *, const E *, const E *, const decltype(E::re) *, const E *, const int *,
                     int, int, const int *, const int *, const int4 *, long, int, int, int,
                     int, int, int, int, int, int);
        if (pw == 2)
            kern = tst ? (jn <= 1 ? k_cross<1, 1, G, E, 2> : jn <= 2 ? k_cross<2, 1, G, E, 2>
                          : jn <= 3 ? k_cross<3, 1, G, E, 2> : k_cross<4, 1, G, E, 2>)
                       : (jn <= 1 ? k_cross<1, 0, G, E, 2> : jn <= 2 ? k_cross<2, 0, G, E, 2>
                          : jn <= 3 ? k_cross<3, 0, G, E, 2> : k_cross<4, 0, G, E, 2>);
        else if (tst)
            kern = jn <= 1 ? k_cross<1, 1, G, E, 1> : jn <= 2 ? k_cross<2, 1, G, E, 1>
                 : jn <= 3 ? k_cross<3, 1, G, E, 1> : k_cross<4, 1, G, E, 1>;
        else
            kern = jn <= 1 ? k_cross<1, 0, G, E, 1> : jn <= 2 ? k_cross<2, 0, G, E, 1>
                 : jn <= 3 ? k_cross<3, 0, G, E, 1> : k_cross<4, 0, G, E, 1>;
        size_t lds = sizeof(E) * (size_t)(cross_sb<E, G>() + 3 * G) * csz[c] +
                     (tst ? cross_tst_bytes<decltype(E::re)>(Tn) : 0) +
                     (sizeof(decltype(E::re)) * 6 + sizeof(int2)) * FSX_CROSS_PB;
        /* persistent-style grid: the blocks that fit the chip at once
         * (FSX_CROSS_BLOCKS_PER_CU overrides the per-CU count) */
        long gsz = grid;
        if (gsz <= 0) {
            int per = 0;
            if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kern, FSX_CROSS_BLOCK, lds)
                    != hipSuccess || per < 1)
                per = 2;
            gsz = (long)g_ncu * per;
        }
        if (gsz > ntiles) gsz = ntiles;
        /* XCD-contiguous tile ranges when the grid divides evenly
         * (FSX_CROSS_XCD=0 turns them off) */
        int nx = env_int("FSX_CROSS_XCD", 1, 0, 1) ? g_nxcd : 1;
        if (nx > 1 && (gsz % nx != 0 || ntiles < 4L * gsz)) nx = 1;
        hipLaunchKernelGGL(kern, dim3((unsigned)gsz), dim3(FSX_CROSS_BLOCK), lds, s,
                           (double *)g[G_ACC], (const E *)g[G_RT] + coff[c],
                           (const E *)g[G_BT] + coff[c], (const decltype(E::re) *)g[G_T],
                           (const E *)g[G_TW2], (const int *)g[G_ENT], cst[c], csz[c],
                           (const int *)g[G_COUNTS], (const int *)g[G_OFFS],
                           (const int4 *)g[G_RECS], ntiles, nrows, Q, L, nb, nx, ng, zn,
                           c == 0, npk);
    }
}

/* The persistent form of the cross-term stage (fsx_score_batch): each
 * complex's m-chunks become jobs, grouped into launches of k_cross_jobs by
 * chunk index and kernel shape, with the shape choices of launch_cross_g. */
struct fsx_cross_group {
    int mixed, G, tst, jn, pw, chunk;   /* the kernel shape and the m-chunk index */
    size_t lds;                         /* the largest job's LDS */
    int njobs, cap;
    long grid;                          /* blocks of the launch */
    fsx_cross_job *jobs;
};

template <typename E, int G>
static const void *cross_jobs_kernel(int tst, int jn, int pw)
{
    if (pw == 2)
        return tst ? (jn <= 1 ? (const void *)k_cross_jobs<1, 1, G, E, 2>
                      : jn <= 2 ? (const void *)k_cross_jobs<2, 1, G, E, 2>
                      : jn <= 3 ? (const void *)k_cross_jobs<3, 1, G, E, 2>
                                : (const void *)k_cross_jobs<4, 1, G, E, 2>)
                   : (jn <= 1 ? (const void *)k_cross_jobs<1, 0, G, E, 2>
                      : jn <= 2 ? (const void *)k_cross_jobs<2, 0, G, E, 2>
                      : jn <= 3 ? (const void *)k_cross_jobs<3, 0, G, E, 2>
                                : (const void *)k_cross_jobs<4, 0, G, E, 2>);
    return tst ? (jn <= 1 ? (const void *)k_cross_jobs<1, 1, G, E, 1>
                  : jn <= 2 ? (const void *)k_cross_jobs<2, 1, G, E, 1>
                  : jn <= 3 ? (const void *)k_cross_jobs<3, 1, G, E, 1>
                            : (const void *)k_cross_jobs<4, 1, G, E, 1>)
               : (jn <= 1 ? (const void *)k_cross_jobs<1, 0, G, E, 1>
                  : jn <= 2 ? (const void *)k_cross_jobs<2, 0, G, E, 1>
                  : jn <= 3 ? (const void *)k_cross_jobs<3, 0, G, E, 1>
                            : (const void *)k_cross_jobs<4, 0, G, E, 1>);
}

static const void *cross_group_kernel(const fsx_cross_group *g)
{
    if (g->mixed) {
        if (g->G >= 4) return cross_jobs_kernel<cplx32, 4>(g->tst, g->jn, g->pw);
        if (g->G >= 2) return cross_jobs_kernel<cplx32, 2>(g->tst, g->jn, g->pw);
        return cross_jobs_kernel<cplx32, 1>(g->tst, g->jn, g->pw);
    }
    if (g->G >= 4) return cross_jobs_kernel<cplx, 4>(g->tst, g->jn, g->pw);
    if (g->G >= 2) return cross_jobs_kernel<cplx, 2>(g->tst, g->jn, g->pw);
    return cross_jobs_kernel<cplx, 1>(g->tst, g->jn, g->pw);
}

/* the jobs of one complex's chunks appended to their groups (groups grown
 * as needed); nx is set at launch */
template <typename E, int G>
static int cross_collect_g(fsx_cross_group **gr, int *ngr, int C, const int *csz, const int *cst,
                           const long *coff, int Tn, int nrows, int Q, int L, int nb, int npk0,
                           int ng, int zn, void *const *g)
{
    typedef decltype(E::re) F;
    const int tst = (long)Tn * (long)(sizeof(E) / 2) <= 40 * 1024;
    const long ntiles = (long)Q * nb * ((ng + G - 1) / G);
    for (int c = 0; c < C; c++) {
        if (csz[c] == 0) continue;
        const int npk = c == 0 ? npk0 : 0;
        const int ns = csz[c] - npk;
        const int pw = sizeof(E) == 8 && ns <= 4 * (FSX_WAVE / 2) &&
                       env_int("FSX_CROSS_PW", 2, 1, 2) == 2 ? 2 : 1;
        const int jn = (ns + FSX_WAVE / pw - 1) / (FSX_WAVE / pw);
        const size_t lds = sizeof(E) * (size_t)(cross_sb<E, G>() + 3 * G) * csz[c] +
                           (tst ? cross_tst_bytes<F>(Tn) : 0) +
                           (sizeof(F) * 6 + sizeof(int2)) * FSX_CROSS_PB;
        const int mixed = sizeof(E) == 8, jq = jn <= 1 ? 1 : jn <= 2 ? 2 : jn <= 3 ? 3 : 4;
        fsx_cross_group *t = NULL;
        for (int i = 0; i < *ngr; i++) {
            fsx_cross_group *u = &(*gr)[i];
            if (u->mixed == mixed && u->G == G && u->tst == tst && u->jn == jq && u->pw == pw &&
                u->chunk == c)
                t = u;
        }
        if (!t) {
            fsx_cross_group *ng2 = (fsx_cross_group *)realloc(*gr, sizeof(fsx_cross_group) * (*ngr + 1));
            if (!ng2) return FSX_ERR_NOMEM;
            *gr = ng2;
            t = &(*gr)[(*ngr)++];
            memset(t, 0, sizeof(*t));
            t->mixed = mixed; t->G = G; t->tst = tst; t->jn = jq; t->pw = pw; t->chunk = c;
        }
        if (t->njobs == t->cap) {
            const int cap = t->cap ? 2 * t->cap : 8;
            fsx_cross_job *nj = (fsx_cross_job *)realloc(t->jobs, sizeof(fsx_cross_job) * cap);
            if (!nj) return FSX_ERR_NOMEM;
            t->jobs = nj;
            t->cap = cap;
        }
        if (lds > t->lds) t->lds = lds;
        fsx_cross_job *jb = &t->jobs[t->njobs++];
        jb->acc = (double *)g[G_ACC];
        jb->Rc = (const E *)g[G_RT] + coff[c];
        jb->Bc = (const E *)g[G_BT] + coff[c];
        jb->T = g[G_T];
        jb->tw2 = g[G_TW2];
        jb->ent = (const int *)g[G_ENT];
        jb->counts = (const int *)g[G_COUNTS];
        jb->offs = (const int *)g[G_OFFS];
        jb->recs = (const int4 *)g[G_RECS];
        jb->ntiles = ntiles;
        jb->cstart = cst[c]; jb->nc = csz[c]; jb->nrows = nrows; jb->Q = Q; jb->L = L;
        jb->nbeta = nb; jb->nx = 1; jb->ng = ng; jb->zn = zn; jb->first = c == 0; jb->npk = npk;
    }
    return FSX_OK;
}

/* G = gamma2 rows per tile: tile_rows when set (the batch tuner's choice),
 * else FSX_CROSS_ROWS, else 4 for FP32 tables (the translation block and the
 * pose loop shared by more rows) and 2 for FP64 tables (LDS-bound) */
template <typename E>
static void launch_cross(hipStream_t s, int C, const int *csz, const int *cst, const long *coff,
                         int Tn, long grid, int nrows, int Q, int L, int nb, int npk0, int ng,
                         int zn, void *const *g, int tile_rows)
{
    const int G = tile_rows > 0 ? tile_rows : env_int("FSX_CROSS_ROWS", sizeof(E) == 8 ? 4 : 2, 1, 4);
    if (G >= 4) launch_cross_g<E, 4>(s, C, csz, cst, coff, Tn, grid, nrows, Q, L, nb, npk0, ng, zn, g);
    else if (G >= 2) launch_cross_g<E, 2>(s, C, csz, cst, coff, Tn, grid, nrows, Q, L, nb, npk0, ng, zn, g);
    else launch_cross_g<E, 1>(s, C, csz, cst, coff, Tn, grid, nrows, Q, L, nb, npk0, ng, zn, g);
}

/* Entry tables of the m-major coefficient order used by the rotated tables
 * and the cross-term kernel: e runs over m = mlo..L, then l = |m|..L (mlo =
 * -L for the full set, 0 for the half set of conjugate-symmetric inputs).
 * The entries are cut into C chunks of whole m values with about ne/C
 * entries each.  ent layout (ints): [0,K2) l, [K2,2K2) m, [2K2,3K2) chunk,
 * [3K2,4K2) index within the chunk, [4K2, 4K2+C) chunk sizes,
 * [5K2, 5K2+2L+1) first entry of each m, [6K2, 6K2+C) chunk starts. */
static int entry_tables(int L, int C, int mlo, int *ent)
{
    const int K = L + 1, K2 = K * K;
    const int ne = mlo == 0 ? K * (K + 1) / 2 : K2;
    int *el = ent, *em = ent + K2, *ec = ent + 2 * K2, *ei = ent + 3 * K2;
    int *csz = ent + 4 * K2, *mst = ent + 5 * K2, *cst = ent + 6 * K2;
    int e = 0, c = 0, cfirst = 0;
    for (int m = -L; m <= L; m++) mst[m + L] = 0;
    for (int m = mlo; m <= L; m++) {
        int am = m < 0 ? -m : m;
        /* close the chunk once it holds its share of the entries */
        if (e > cfirst && c < C - 1 && (long)e * C >= (long)(c + 1) * ne) {
            csz[c] = e - cfirst;
            cst[c] = cfirst;
            c++;
            cfirst = e;
        }
        mst[m + L] = e;
        for (int l = am; l <= L; l++, e++) {
            el[e] = l;
            em[e] = m;
            ec[e] = c;
            ei[e] = e - cfirst;
        }
    }
    csz[c] = e - cfirst;
    cst[c] = cfirst;
    for (int k = c + 1; k < C; k++) { csz[k] = 0; cst[k] = e; }
    return ne;
}

/* A^X_{l,-m}(q) = (-1)^(l+m) conj(A^X_{lm}(q)) for m >= 0, to 1e-13 of the
 * largest coefficient of each (X, q): the amplitude of a real density.
 * Then the terms of m and -m of every cross-term sum are complex
 * conjugates of each other, and the sums run over m >= 0 with the m > 0
 * terms counted twice. */
static int spf_conj_symmetric(const double *re, const double *im, int Q, int L)
{
    const int K2 = (L + 1) * (L + 1);
    for (long xq = 0; xq < 3L * Q; xq++) {
        const double *r = re + xq * K2, *i = im + xq * K2;
        double mx = 0.0;
        for (int e = 0; e < K2; e++) {
            const double a = fabs(r[e]), b = fabs(i[e]);
            mx = a > mx ? a : mx;
            mx = b > mx ? b : mx;
        }
        const double tol = 1e-13 * mx;
        /* m = 0: A_l0 = (-1)^l conj(A_l0), i.e. real for even l and imaginary
         * for odd l (the packed m = 0 slots keep only that part of i^l A_l0) */
        for (int l = 0; l <= L; l++) {
            const int e0 = l * l + l;
            if (!(fabs((l & 1) ? r[e0] : i[e0]) <= tol)) return 0;
        }
        for (int l = 1; l <= L; l++)
            for (int m = 1; m <= l; m++) {
                const int ep = l * l + l + m, en = l * l + l - m;
                const double sg = ((l + m) & 1) ? -1.0 : 1.0;
                if (!(fabs(r[en] - sg * r[ep]) <= tol) || !(fabs(i[en] + sg * i[ep]) <= tol)) return 0;
            }
    }
    return 1;
}

/* ---- one complex's pipeline: plan, buffers, uploads, launches ---------------
 * plan_init derives sizes and the host tables of a call; plan_bind gets its
 * device buffers (the process-wide slots for single calls, per-complex sets
 * for batches); plan_upload copies the inputs and tables; plan_enqueue
 * launches every stage on one stream without waiting; plan_finish fills the
 * statistics once the stream has drained. */
enum {
    P_BES, P_T, P_RT, P_BT, P_RTMP, P_A, P_B, P_DW, P_DS, P_TWA, P_TWG, P_Q, P_Z, P_AM, P_CINT,
    P_KEYS, P_INFO, P_COUNTS, P_OFFS, P_FILL, P_RECS, P_ACC, P_FITSC, P_ENT, P_TW2, P_COFF,
    P_NEXT, P_TALLY, P_SRCT, P_POSES, P_SCORE, P_C1, P_C2, P_CROSSOUT, P_ACCF, P_IN, P_NBUF
};
/* the process-wide slot of each plan buffer (single calls) */
static const int k_plan_slot[P_NBUF] = {
    B_BES, B_T, B_RT, B_BT, B_RTMP, B_A, B_B, B_DW, B_DS, B_TWA, B_TWG, B_Q, B_Z, B_AM, B_CINT,
    B_KEYS, B_INFO, B_COUNTS, B_OFFS, B_FILL, B_RECS, B_ACC, B_FITSC, B_ENT, B_TW2, B_COFF,
    B_NEXT, B_TALLY, B_SRCT, B_POSES, B_SCORE, B_C1, B_C2, B_CROSSOUT, B_ACCF, B_IN};
/* The inputs and host tables of a call, packed into one device buffer
 * (P_IN, each at a 256-byte aligned offset) and uploaded by one copy from
 * pinned memory: separate copies from pageable memory cost 10 to 40 us of
 * host time each. */
static const int k_packed[] = {P_A, P_B, P_TWA, P_TWG, P_Q, P_Z, P_AM, P_CINT, P_TW2, P_ENT, P_COFF};
static inline bool plan_packed(int i)
{
    for (int k = 0; k < (int)(sizeof(k_packed) / sizeof(k_packed[0])); k++)
        if (k_packed[k] == i) return true;
    return false;
}

typedef void (*fit_kernel_t)(const double *, const int4 *, const int *, int, const double *,
                             const double *, double, double, double, double, double, double *,
                             double *, double *, int *, unsigned long long *, int);

struct fsx_plan {
    fsx_hip_args A;
    int L, K, K2, N, Q, n, nb, na, ng, zn, nrows, mixed, half, mlo, npk, C, Tn, ne, dev_out;
    long nkeys;
    size_t dsz, spf, es, nent;
    int *ent;     /* host tables (malloc) */
    long *coff;
    double *tw2;
    size_t sz[P_NBUF];   /* bytes of each device buffer (0: not used) */
    size_t inoff[P_NBUF];   /* offset of a packed input within P_IN */
    void *d[P_NBUF];     /* device buffers */
    const long long *dposes;
    double *dscore, *dc1, *dc2, *dcross;
    int aq_lds;
    size_t lds_fit;
    long lanes;
    fit_kernel_t kfit;   /* the compact-form fit (grp = 0) */
    int grp;             /* the lane-group fit k_fitg (mixed form) */
    int fit_wide;        /* lanes with a fit at or below which a fit wavefront
                            evaluates them together (L-BFGS-B fits); lanes per
                            fit (Newton fits) */
    long fit_next0;      /* first queue entry not taken at launch */
    int tile_rows;       /* gamma2 rows per cross-term tile (0: the form's default) */
    int rot_gt;          /* gamma per block of the MFMA rotation (16 or 8), 0: the
                            rotation-then-transpose kernels */
};

static void plan_free_host(fsx_plan *p)
{
    free(p->ent);
    free(p->coff);
    free(p->tw2);
    p->ent = NULL;
    p->coff = NULL;
    p->tw2 = NULL;
}

/* tile_rows, slot_cap: the cross-term tile shape (gamma2 rows per tile,
 * pairing slots per m-chunk); 0 takes the defaults (environment knobs, then
 * the form's defaults) */
static int plan_init(fsx_plan *p, const fsx_hip_args *A, int tile_rows, int slot_cap)
{
    memset(p, 0, sizeof(*p));
    p->A = *A;
    p->tile_rows = tile_rows;
    const int L = A->L, K = L + 1, K2 = K * K, N = 2 * L + 1, Q = A->qnum, n = A->n;
    const int nb = A->nbeta, na = A->nalpha, ng = A->ngamma, zn = A->znum;
    p->L = L; p->K = K; p->K2 = K2; p->N = N; p->Q = Q; p->n = n;
    p->nb = nb; p->na = na; p->ng = ng; p->zn = zn;
    p->nrows = nb * ng;
    p->nkeys = (long)zn * nb * ng;
    if (p->nkeys > 0x3fffffffL || (long)p->nrows * Q > 0x7fffffffL) return FSX_ERR_ARG;
    if ((long)L * na >= (1L << 24) || (long)L * ng >= (1L << 24)) return FSX_ERR_ARG;
    p->dsz = (size_t)K * N * N;
    p->spf = (size_t)3 * Q * K2;
    p->mixed = (A->flags & FSX_MIXED_FP32) != 0;
    p->dev_out = (A->flags & FSX_DEVICE_OUTPUT) != 0;
    /* half entry set (m >= 0) when both molecules' coefficients are
     * conjugate symmetric; FSX_CROSS_HALF=0 forces the full set */
    p->half = env_int("FSX_CROSS_HALF", 1, 0, 1) && spf_conj_symmetric(A->A_re, A->A_im, Q, L) &&
              spf_conj_symmetric(A->B_re, A->B_im, Q, L);
    p->mlo = p->half ? 0 : -L;
    /* half set: the m = 0 entries of the receptor table and of U share
     * pairing slots two by two (FSX_CROSS_PACK=0 keeps one slot each) */
    p->npk = p->half && env_int("FSX_CROSS_PACK", 1, 0, 1) ? K / 2 : 0;
    p->es = p->mixed ? sizeof(cplx32) : sizeof(cplx);   /* table element size */
    /* m-chunks: one launch per chunk of entries, starting from one chunk
     * (FSX_CROSS_CHUNKS: the smallest count to try); more below while a chunk
     * is too large */
    int C = env_int("FSX_CROSS_CHUNKS", 1, 1, N);
    p->Tn = tpk_sq(K);
    p->nent = 8 * (size_t)K2 + 2;
    p->ent = (int *)malloc(sizeof(int) * p->nent);
    p->coff = (long *)malloc(sizeof(long) * (N + 1));
    p->tw2 = (double *)calloc(2 * (size_t)na * (N + 1), sizeof(double));
    if (!p->ent || !p->coff || !p->tw2) {
        plan_free_host(p);
        return FSX_ERR_NOMEM;
    }
    int *ent = p->ent;
    /* more chunks until the largest fits: at most 4 entries per lane, and
     * for FP32 tables at most 128 pairing slots (the two-poses-per-wavefront
     * pairing); FSX_CROSS_CHUNK overrides the slot limit */
    const int cmax = slot_cap > 0 ? slot_cap
                                  : env_int("FSX_CROSS_CHUNK", p->mixed ? 2 * FSX_WAVE : 4 * FSX_WAVE,
                                            64, 4 * FSX_WAVE);
    int ncmax, nsmax, ne;
    for (;;) {
        ne = entry_tables(L, C, p->mlo, ent);
        ncmax = nsmax = 0;
        for (int c = 0; c < C; c++) {
            const int nc = ent[4 * K2 + c], ns = nc - (c == 0 ? p->npk : 0);
            if (nc > ncmax) ncmax = nc;
            if (ns > nsmax) nsmax = ns;
        }
        if ((ncmax <= 4 * FSX_WAVE && nsmax <= cmax) || C == N) break;
        C++;
    }
    if (ncmax > 4 * FSX_WAVE) {
        plan_free_host(p);
        return FSX_ERR_ARG;
    }
    /* the cross-term kernel addresses a q's receptor rows and a row of the
     * twiddle table by 32-bit byte offsets from 24-bit products */
    if ((size_t)p->nrows * 3 * ncmax * p->es >= ((size_t)1 << 32) || p->nrows >= (1 << 24) ||
        na >= (1 << 24) || (size_t)na * (2 * L + 2) * p->es >= ((size_t)1 << 32)) {
        plan_free_host(p);
        return FSX_ERR_ARG;
    }
    p->C = C;
    p->ne = ne;
    /* U-phase thread -> entry of each chunk: entries by increasing |m|
     * (decreasing row length K - |m|), in entry order within one |m| */
    {
        int *uperm = ent + 7 * K2 + 2;
        const int *em = ent + K2, *csz = ent + 4 * K2, *cst = ent + 6 * K2;
        for (int c = 0; c < C; c++) {
            int t = 0;
            for (int am = 0; am <= L; am++)
                for (int i = 0; i < csz[c]; i++) {
                    const int m = em[cst[c] + i];
                    if ((m < 0 ? -m : m) == am) uperm[cst[c] + t++] = i;
                }
        }
    }
    /* twiddle e^{-i m a2 2pi/na} per (a2, m): the same table entries the
     * reference's exact transform uses, gathered by (m a2) mod na */
    for (int a2 = 0; a2 < na; a2++)
        for (int m = -L; m <= L; m++) {
            long k = ((long)m * a2) % na;
            if (k < 0) k += na;
            /* half set: the m > 0 terms stand for m and -m (an exact
             * factor 2) */
            const double f = p->half && m > 0 ? 2.0 : 1.0;
            p->tw2[2 * ((long)a2 * (N + 1) + m + L)] = f * A->tw_a[2 * k];
            p->tw2[2 * ((long)a2 * (N + 1) + m + L) + 1] = f * A->tw_a[2 * k + 1];
        }
    /* the mixed form reads them as FP32 pairs, in place */
    if (p->mixed) {
        float *f32 = (float *)p->tw2;
        for (long i = 0; i < 2L * na * (N + 1); i++) f32[i] = (float)p->tw2[i];
    }
    const int *csz = ent + 4 * K2;
    p->coff[0] = 0;
    for (int c = 0; c < C; c++) p->coff[c + 1] = p->coff[c] + (long)Q * p->nrows * 3 * csz[c];
    /* buffer sizes */
    const size_t es = p->es, spfe = (size_t)3 * Q * ne;   /* rotated-table entries per row */
    size_t *z = p->sz;
    z[P_BES] = sizeof(double) * zn * Q * N;
    z[P_T] = es * zn * Q * (size_t)p->Tn;
    z[P_RT] = z[P_BT] = z[P_RTMP] = es * (size_t)p->nrows * spfe;
    /* the MFMA rotation when a tile of 16 (else 8) rows of the table fits
     * 64 KB of LDS (FSX_ROT_MFMA=0: the rotation-then-transpose kernels) */
    p->rot_gt = 0;
    if (env_int("FSX_ROT_MFMA", 1, 0, 1)) {
        if ((size_t)16 * ne * es <= 64 * 1024) p->rot_gt = 16;
        else if ((size_t)8 * ne * es <= 64 * 1024) p->rot_gt = 8;
    }
    if (p->rot_gt) z[P_RTMP] = 0;
    z[P_A] = z[P_B] = sizeof(double) * 2 * p->spf;
    z[P_DW] = sizeof(double) * p->dsz * nb;
    z[P_DS] = sizeof(double) * (size_t)K * K2 * N;
    z[P_TWA] = sizeof(double) * 2 * na;
    z[P_TWG] = sizeof(double) * 2 * ng;
    z[P_Q] = sizeof(double) * Q;
    z[P_Z] = sizeof(double) * zn;
    z[P_AM] = z[P_CINT] = sizeof(double) * 6 * Q;
    z[P_KEYS] = sizeof(int) * n;
    z[P_INFO] = sizeof(int2) * n;
    z[P_COUNTS] = z[P_FILL] = sizeof(int) * p->nkeys;
    z[P_OFFS] = sizeof(int) * (p->nkeys + 2);
    z[P_RECS] = sizeof(int4) * n;
    z[P_ACC] = sizeof(double) * 6 * Q * (size_t)n;
    z[P_FITSC] = sizeof(double) * (size_t)(n > 0 ? n : 1);
    z[P_ENT] = sizeof(int) * p->nent;
    z[P_TW2] = sizeof(double) * 2 * (size_t)na * (N + 1);   /* rows of 2L + 2 entries */
    z[P_COFF] = sizeof(long) * (C + 1);
    z[P_NEXT] = sizeof(int);
    z[P_TALLY] = FSX_TALLY_N * sizeof(unsigned long long);
    z[P_SRCT] = sizeof(cplx) * (size_t)Q * 3 * K2 * ng;
    z[P_POSES] = (A->flags & FSX_DEVICE_POSES) ? 0 : sizeof(long long) * n;
    z[P_SCORE] = z[P_C1] = z[P_C2] = p->dev_out ? 0 : sizeof(double) * n;
    z[P_CROSSOUT] = (!p->dev_out && A->cross) ? sizeof(double) * 6 * Q * (size_t)n : 0;
    size_t inb = 0;
    for (int k = 0; k < (int)(sizeof(k_packed) / sizeof(k_packed[0])); k++) {
        p->inoff[k_packed[k]] = inb;
        inb += (z[k_packed[k]] + 255) & ~(size_t)255;
    }
    z[P_IN] = inb;
    /* the fit launch: as many lanes as are resident at once (the fit
     * state lives in registers, one 256-thread block per CU) */
    p->aq_lds = fit_aq_lds(Q);
    p->lds_fit = fit_lds(Q);
    /* the mixed-precision form's fit: the two-variable L-BFGS-B on lane
     * groups (k_fitg), or with FSX_FIT_COMPACT (or a q grid whose arrays do
     * not fit the block's LDS) the compact-form L-BFGS-B lane per fit, both
     * evaluated in one pass over q; the FP64 form: the reference's
     * L-BFGS-B, lane per fit.  Lanes per fit of the group fit: about
     * FSX_FG_QLANE q per lane, at least two (FSX_FIT_GROUP_LANES
     * overrides), doubled until the block's LDS fits */
    int fpl = 2;
    while (fpl < FSX_WAVE && fpl * FSX_FG_QLANE < Q) fpl *= 2;
    fpl = env_int("FSX_FIT_GROUP_LANES", fpl, 1, FSX_WAVE);
    while (fpl & (fpl - 1)) fpl &= fpl - 1;   /* a power of two */
    while (fpl < FSX_WAVE && fitg_lds(Q, fpl) > FSX_FIT_LDS) fpl *= 2;
    const int grp = p->mixed && !(A->flags & FSX_FIT_COMPACT) && fitg_lds(Q, fpl) <= FSX_FIT_LDS;
    p->kfit = p->aq_lds ? (p->mixed ? k_fit<1, 1> : k_fit<1, 0>)
                        : (p->mixed ? k_fit<0, 1> : k_fit<0, 0>);
    p->fit_wide = env_int("FSX_FIT_WIDE", 24, 0, 64);
    p->grp = grp;
    if (grp) {
        p->fit_wide = fpl;
        p->lds_fit = fitg_lds(Q, fpl);
        p->sz[P_ACCF] = sizeof(float) * 6 * Q * (size_t)n;
        int bpc = env_int("FSX_FIT_BLOCKS_PER_CU", 0, 0, 16);
        if (bpc == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fitg, 256,
                                                                      p->lds_fit) != hipSuccess ||
                         bpc < 1))
            bpc = 1;
        p->lanes = (long)g_ncu * 256 * bpc;
        if (p->lanes > (long)n * fpl) p->lanes = ((long)n * fpl + 255) / 256 * 256;
        p->fit_next0 = p->lanes / fpl;
        return FSX_OK;
    }
    int fit_bpc = env_int("FSX_FIT_BLOCKS_PER_CU", 0, 0, 16);
    if (fit_bpc == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit_bpc, p->kfit, 256,
                                                                      p->lds_fit) != hipSuccess ||
                         fit_bpc < 1))
        fit_bpc = 1;
    p->lanes = (long)g_ncu * 256 * fit_bpc;
    const int fit_blocks = env_int("FSX_FIT_BLOCKS", 0, 0, 1 << 20);
    if (fit_blocks > 0) p->lanes = 256L * fit_blocks;
    if (p->lanes > n) p->lanes = ((long)n + 255) / 256 * 256;
    p->fit_next0 = p->lanes;
    return FSX_OK;
}

/* Device pointers of the plan from get(buffer index, bytes, ctx); the
 * device-resident inputs and outputs of the call are used in place. */
static int plan_bind(fsx_plan *p, int (*get)(int, size_t, void *, void **), void *ctx)
{
    for (int i = 0; i < P_NBUF; i++) {
        p->d[i] = NULL;
        if (p->sz[i] == 0 || plan_packed(i)) continue;
        RC(get(i, p->sz[i], ctx, &p->d[i]));
    }
    for (int i = 0; i < P_NBUF; i++)
        if (plan_packed(i)) p->d[i] = (char *)p->d[P_IN] + p->inoff[i];
    const fsx_hip_args *A = &p->A;
    p->dposes = (A->flags & FSX_DEVICE_POSES) ? (const long long *)A->poses
                                              : (const long long *)p->d[P_POSES];
    if (p->dev_out) {
        p->dscore = A->score; p->dc1 = A->c1; p->dc2 = A->c2; p->dcross = A->cross;
    } else {
        p->dscore = (double *)p->d[P_SCORE];
        p->dc1 = (double *)p->d[P_C1];
        p->dc2 = (double *)p->d[P_C2];
        p->dcross = (double *)p->d[P_CROSSOUT];
    }
    return FSX_OK;
}

#define HIPR(x)                                                                       \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "[fsx] %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                    __FILE__, __LINE__);                                              \
            return FSX_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

/* Pinned staging memory of the packed inputs: stage_begin(bytes) before a
 * call's uploads waits until the previous call's copies out of it are done
 * (an event recorded by stage_end) and grows it to bytes; stage_take hands
 * out 256-byte aligned pieces in order. */
static void *g_stage = NULL;
static size_t g_stage_cap = 0, g_stage_used = 0;
static hipEvent_t g_stage_ev = NULL;
static int g_stage_pending = 0;

static int stage_begin(size_t bytes)
{
    if (g_stage_pending) {
        HIPCHK(hipEventSynchronize(g_stage_ev));
        g_stage_pending = 0;
    }
    if (!g_stage_ev) HIPCHK(hipEventCreateWithFlags(&g_stage_ev, hipEventDisableTiming));
    g_stage_used = 0;
    if (g_stage_cap >= bytes) return FSX_OK;
    if (g_stage) HIPCHK(hipHostFree(g_stage));
    g_stage = NULL;
    g_stage_cap = 0;
    const size_t cap = bytes + bytes / 8 + 256;
    if (hipHostMalloc(&g_stage, cap, hipHostMallocDefault) != hipSuccess) {
        g_stage = NULL;
        return FSX_ERR_NOMEM;
    }
    g_stage_cap = cap;
    return FSX_OK;
}

static char *stage_take(size_t bytes)
{
    bytes = (bytes + 255) & ~(size_t)255;
    if (!g_stage || g_stage_used + bytes > g_stage_cap) return NULL;
    char *h = (char *)g_stage + g_stage_used;
    g_stage_used += bytes;
    return h;
}

static int stage_end(hipStream_t s)
{
    HIPCHK(hipEventRecord(g_stage_ev, s));
    g_stage_pending = 1;
    return FSX_OK;
}

static void stage_release(void)
{
    if (g_stage_pending) (void)hipEventSynchronize(g_stage_ev);
    g_stage_pending = 0;
    if (g_stage) (void)hipHostFree(g_stage);
    g_stage = NULL;
    g_stage_cap = g_stage_used = 0;
}

/* Inputs and tables to the plan's buffers: the packed ones through the
 * pinned staging memory (stage_begin must have made room for sz[P_IN]) in
 * one copy.  up_dw / up_ds: whether the Wigner-d and d-symbol tables must
 * be copied (the buffers may already hold them). */
static int plan_upload(fsx_plan *p, hipStream_t s, int up_dw, int up_ds)
{
    const fsx_hip_args *A = &p->A;
    const size_t spf = p->spf;
    const int Q = p->Q, n = p->n, K = p->K, K2 = p->K2, N = p->N;
    char *h = stage_take(p->sz[P_IN]);
    if (!h) return FSX_ERR_NOMEM;
    auto put = [&](int i, size_t at, const void *src, size_t bytes) {
        memcpy(h + p->inoff[i] + at, src, bytes);
    };
    put(P_A, 0, A->A_re, sizeof(double) * spf);
    put(P_A, sizeof(double) * spf, A->A_im, sizeof(double) * spf);
    put(P_B, 0, A->B_re, sizeof(double) * spf);
    put(P_B, sizeof(double) * spf, A->B_im, sizeof(double) * spf);
    put(P_TWA, 0, A->tw_a, p->sz[P_TWA]);
    put(P_TWG, 0, A->tw_g, p->sz[P_TWG]);
    put(P_Q, 0, A->qvals, sizeof(double) * Q);
    put(P_Z, 0, A->zvals, sizeof(double) * p->zn);
    put(P_AM, 0, A->a, sizeof(double) * 6 * Q);
    put(P_CINT, 0, A->cint, sizeof(double) * 6 * Q);
    put(P_ENT, 0, p->ent, sizeof(int) * p->nent);
    put(P_TW2, 0, p->tw2, p->sz[P_TW2]);
    put(P_COFF, 0, p->coff, sizeof(long) * (p->C + 1));
    HIPR(hipMemcpyAsync(p->d[P_IN], h, p->sz[P_IN], hipMemcpyHostToDevice, s));
    if (up_dw) HIPR(hipMemcpyAsync(p->d[P_DW], A->dW, p->sz[P_DW], hipMemcpyHostToDevice, s));
    if (up_ds)
        HIPR(hipMemcpyAsync(p->d[P_DS], A->dsymb, sizeof(double) * K * K2 * N, hipMemcpyHostToDevice, s));
    if (!(A->flags & FSX_DEVICE_POSES))
        HIPR(hipMemcpyAsync(p->d[P_POSES], A->poses, sizeof(long long) * n, hipMemcpyHostToDevice, s));
    if (!p->dev_out) {
        /* device mirrors start from the caller's values so that poses the
         * reference leaves untouched stay untouched */
        HIPR(hipMemcpyAsync(p->dscore, A->score, sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPR(hipMemcpyAsync(p->dc1, A->c1, sizeof(double) * n, hipMemcpyHostToDevice, s));
        HIPR(hipMemcpyAsync(p->dc2, A->c2, sizeof(double) * n, hipMemcpyHostToDevice, s));
        if (A->cross)
            HIPR(hipMemcpyAsync(p->dcross, A->cross, sizeof(double) * 6 * Q * (size_t)n,
                                hipMemcpyHostToDevice, s));
    }
    return FSX_OK;
}

/* Every stage of the plan on stream s, no waiting; ev (may be NULL): seven
 * events recorded at the stage boundaries (after the tables, rotations,
 * bucketing, cross terms, fit preparation and fit). */
/* stages of plan_enqueue */
#define FSX_ST_PREP 1      /* translation tables, rotated tables, pose buckets */
#define FSX_ST_CROSS 2     /* cross terms */
#define FSX_ST_FITPREP 4   /* fit queue and counters, profile rescale */
#define FSX_ST_FIT 8       /* the fit */
#define FSX_ST_ALL 15

static int plan_enqueue(fsx_plan *p, hipStream_t s, hipEvent_t *ev, int stages = FSX_ST_ALL)
{
    const fsx_hip_args *A = &p->A;
    const int L = p->L, K = p->K, K2 = p->K2, N = p->N, Q = p->Q, n = p->n;
    const int nb = p->nb, na = p->na, ng = p->ng, zn = p->zn, nrows = p->nrows, C = p->C;
    const long nkeys = p->nkeys;
    const int *csz = p->ent + 4 * K2, *cst = p->ent + 6 * K2;
    const long *coff = p->coff;
    void **g = p->d;
    int *counts = (int *)g[P_COUNTS], *offs = (int *)g[P_OFFS], *fill = (int *)g[P_FILL];
    (void)na;
    if (stages & FSX_ST_PREP) {
    /* translation tables */
    const long nbes = (long)zn * Q * N;
    hipLaunchKernelGGL(k_bessel, dim3(nblk(nbes, 256)), dim3(256), 0, s, (double *)g[P_BES],
                       (const double *)g[P_Z], (const double *)g[P_Q], zn, Q, N);
    const long nt = (long)zn * Q * tpk_sq(K);
    if (p->mixed)
        hipLaunchKernelGGL(k_tmat<cplx32>, dim3(nblk(nt, 256)), dim3(256), 0, s, (float *)g[P_T],
                           (const double *)g[P_BES], (const double *)g[P_DS], zn, Q, L);
    else
        hipLaunchKernelGGL(k_tmat<cplx>, dim3(nblk(nt, 256)), dim3(256), 0, s, (double *)g[P_T],
                           (const double *)g[P_BES], (const double *)g[P_DS], zn, Q, L);
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[1], s));
    /* rotated receptor and ligand coefficients, m-chunked */
    {
        const long ntw = (long)Q * 3 * K2 * ng;
        cplx *srct = (cplx *)g[P_SRCT];
        const double *dA = (const double *)g[P_A], *dB = (const double *)g[P_B];
        const size_t spf = p->spf;
        /* gamma chunk: the staged rows (2L+1) x gc complex stay within 48 KB */
        int gc = ng < FSX_WAVE ? ng : FSX_WAVE;
        while (gc > 1 && (size_t)N * gc * sizeof(cplx) > 48 * 1024) gc /= 2;
        const long nrot = (long)Q * 3 * K * ((nb + FSX_ROT_BC - 1) / FSX_ROT_BC) * ((ng + gc - 1) / gc);
        const size_t lds_rot = (size_t)N * gc * sizeof(cplx) + (size_t)FSX_ROT_BC * N * N * sizeof(double);
        for (int side = 0; side < 2; side++) {
            const double *src = side == 0 ? dA : dB;
            void *out = side == 0 ? g[P_RT] : g[P_BT];
            hipLaunchKernelGGL(k_twiddle, dim3(nblk(ntw, 256)), dim3(256), 0, s, srct, src,
                               src + spf, (const double *)g[P_TWG], ng, Q, L, side);
            const int npk = side == 0 ? p->npk : 0, lrot = side == 0 ? 1 : -1;
            if (p->rot_gt) {
                const int gt = p->rot_gt;
                const long nbr = (long)Q * 3 * nb * ((ng + gt - 1) / gt);
                const size_t lds = (size_t)gt * p->ne * p->es;
                const double *dw = (const double *)g[P_DW];
                const int *en = (const int *)g[P_ENT];
                const long *co = (const long *)g[P_COFF];
                if (p->mixed && gt == 16)
                    hipLaunchKernelGGL((k_rotate_mfma<cplx32, 16>), dim3((unsigned)nbr), dim3(256), lds, s,
                                       (cplx32 *)out, (const cplx *)srct, dw, en, co, nb, ng, Q, L,
                                       p->mlo, p->ne, npk, lrot);
                else if (p->mixed)
                    hipLaunchKernelGGL((k_rotate_mfma<cplx32, 8>), dim3((unsigned)nbr), dim3(256), lds, s,
                                       (cplx32 *)out, (const cplx *)srct, dw, en, co, nb, ng, Q, L,
                                       p->mlo, p->ne, npk, lrot);
                else if (gt == 16)
                    hipLaunchKernelGGL((k_rotate_mfma<cplx, 16>), dim3((unsigned)nbr), dim3(256), lds, s,
                                       (cplx *)out, (const cplx *)srct, dw, en, co, nb, ng, Q, L,
                                       p->mlo, p->ne, npk, lrot);
                else
                    hipLaunchKernelGGL((k_rotate_mfma<cplx, 8>), dim3((unsigned)nbr), dim3(256), lds, s,
                                       (cplx *)out, (const cplx *)srct, dw, en, co, nb, ng, Q, L,
                                       p->mlo, p->ne, npk, lrot);
                continue;
            }
            if (p->mixed)
                launch_rotate<cplx32>(s, out, g[P_RTMP], srct, nrot, lds_rot, nb, ng, Q, L, gc, C,
                                      csz, cst, coff, p->mlo, npk, lrot, g);
            else
                launch_rotate<cplx>(s, out, g[P_RTMP], srct, nrot, lds_rot, nb, ng, Q, L, gc, C,
                                    csz, cst, coff, p->mlo, npk, lrot, g);
        }
    }
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[2], s));
    /* bucket poses by (z, beta2, gamma2) */
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nkeys, 256)), dim3(256), 0, s, counts, nkeys, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(nblk(nkeys, 256)), dim3(256), 0, s, fill, nkeys, 0);
    hipLaunchKernelGGL(k_count, dim3(nblk(n, 256)), dim3(256), 0, s, p->dposes, n, zn, nb, p->na,
                       ng, (int *)g[P_KEYS], (int2 *)g[P_INFO], counts);
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, s, (const int *)counts, (int)nkeys, offs);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n, 256)), dim3(256), 0, s, (const int *)g[P_KEYS],
                       (const int2 *)g[P_INFO], n, ng, (const int *)offs, fill, (int4 *)g[P_RECS]);
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[3], s));
    }
    /* per-pose cross terms, one launch per m-chunk, accumulated in order */
    if (stages & FSX_ST_CROSS) {
    const long grid = (long)g_ncu * env_int("FSX_CROSS_BLOCKS_PER_CU", 0, 0, 16);
    if (p->mixed)
        launch_cross<cplx32>(s, C, csz, cst, coff, p->Tn, grid, nrows, Q, L, nb, p->npk, ng, zn, g,
                             p->tile_rows);
    else
        launch_cross<cplx>(s, C, csz, cst, coff, p->Tn, grid, nrows, Q, L, nb, p->npk, ng, zn, g,
                           p->tile_rows);
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[4], s));
    }
    if (stages & FSX_ST_FITPREP) {
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, (int *)g[P_NEXT], 1L, (int)p->fit_next0);
    hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, (int *)g[P_TALLY], 2L * FSX_TALLY_N, 0);
    /* rescaled profiles in place of the pose parts (bounded by n slots;
     * the kernels read the number of valid slots on the device) */
    hipLaunchKernelGGL(k_fit_scale, dim3(nblk(n, 256)), dim3(256), 0, s, (const double *)g[P_ACC],
                       (const double *)g[P_CINT], (const int *)(offs + nkeys), Q, A->peak,
                       (double *)g[P_FITSC]);
    hipLaunchKernelGGL(k_fit_prep, dim3(nblk((long)n * Q, 256)), dim3(256), 0, s,
                       (double *)g[P_ACC], (const double *)g[P_CINT], (const int4 *)g[P_RECS],
                       (const int *)(offs + nkeys), Q, (const double *)g[P_FITSC], p->dcross,
                       p->grp ? (float *)g[P_ACCF] : (float *)NULL);
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[5], s));   /* the fit kernel alone from here */
    }
    if (stages & FSX_ST_FIT) {
    if (p->grp)
        hipLaunchKernelGGL(k_fitg, dim3((unsigned)(p->lanes / 256)), dim3(256), p->lds_fit, s,
                           (const float *)g[P_ACCF], (const int4 *)g[P_RECS],
                           (const int *)(offs + nkeys), Q, (const double *)g[P_Q],
                           (const double *)g[P_AM], A->mult, A->c1_lo, A->c1_hi, A->c2_lo,
                           A->c2_hi, p->dscore, p->dc1, p->dc2, (int *)g[P_NEXT],
                           (unsigned long long *)g[P_TALLY], p->fit_wide);
    else
        hipLaunchKernelGGL(p->kfit, dim3((unsigned)(p->lanes / 256)), dim3(256), p->lds_fit, s,
                           (const double *)g[P_ACC], (const int4 *)g[P_RECS],
                           (const int *)(offs + nkeys), Q, (const double *)g[P_Q],
                           (const double *)g[P_AM], A->mult, A->c1_lo, A->c1_hi, A->c2_lo,
                           A->c2_hi, p->dscore, p->dc1, p->dc2, (int *)g[P_NEXT],
                           (unsigned long long *)g[P_TALLY], p->fit_wide);
    HIPR(hipGetLastError());
    if (ev) HIPR(hipEventRecord(ev[6], s));
    }
    return FSX_OK;
}

/* Results to the host (host outputs) and the call's counters to tail /
 * tally, on stream s. */
static int plan_download(fsx_plan *p, hipStream_t s, int *tail, unsigned long long *tally)
{
    const fsx_hip_args *A = &p->A;
    const int n = p->n, Q = p->Q;
    if (!p->dev_out) {
        HIPR(hipMemcpyAsync(A->score, p->dscore, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPR(hipMemcpyAsync(A->c1, p->dc1, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        HIPR(hipMemcpyAsync(A->c2, p->dc2, sizeof(double) * n, hipMemcpyDeviceToHost, s));
        if (A->cross)
            HIPR(hipMemcpyAsync(A->cross, p->dcross, sizeof(double) * 6 * Q * (size_t)n,
                                hipMemcpyDeviceToHost, s));
    }
    HIPR(hipMemcpyAsync(tail, (int *)p->d[P_OFFS] + p->nkeys, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    HIPR(hipMemcpyAsync(tally, p->d[P_TALLY], FSX_TALLY_N * sizeof(unsigned long long),
                        hipMemcpyDeviceToHost, s));
    return FSX_OK;
}

/* The statistics of a finished plan: counters, algorithmic work and (ms not
 * NULL) the six stage times. */
static void plan_stats(const fsx_plan *p, fsx_stats *st, const float *ms, const int *tail,
                       const unsigned long long *tally)
{
    const int K = p->K, K2 = p->K2, Q = p->Q, ne = p->ne, npk = p->npk, C = p->C;
    const int nvalid = tail[0];
    const long items = (long)Q * tail[1];
    if (ms) {
        st->ms_bessel_t = ms[0];
        st->ms_rotate = ms[1];
        st->ms_group = ms[2];
        st->ms_cross = ms[3];
        st->ms_fit_prep = ms[4];
        st->ms_fit = ms[5];
    }
    st->n_scored = nvalid;
    st->fit_evals = (long)tally[0];
    st->n_capped = (int)tally[1];
    st->fit_steps = (long)tally[2];
    st->fit_wide_steps = (long)tally[3];
    st->fit_max_steps = (long)tally[4];
    st->cross_items = items;
    st->cross_chunks = C;
    /* useful FP64 flops of the cross-term launches: 8 per complex
     * multiply-add of the translated ligand (3 parts, sum_e (L+1-|m_e|)
     * terms per item) and 54 per (pose, q, pairing slot); the npk packed
     * slots carry two m = 0 entries whose other halves are zero */
    double tpair = 0.0;
    for (int e = 0; e < ne; e++) tpair += K - (p->ent[K2 + e] < 0 ? -p->ent[K2 + e] : p->ent[K2 + e]);
    const double nslot = (double)(ne - npk);
    st->cross_flops = (double)items * 3.0 * 8.0 * tpair + (double)nvalid * Q * nslot * 54.0;
    /* algorithmic bytes: per pose and q one receptor row (3 parts, one
     * value per slot) and the six-term accumulator (read and written by
     * every chunk after the first); per item one ligand row */
    st->cross_bytes = (double)nvalid * Q * (3.0 * nslot * (double)p->es + (2.0 * C - 1.0) * 48.0)
                    + (double)items * 3.0 * ne * (double)p->es;
    st->cross_entries = ne;
}

static_assert(G_T == P_T && G_RT == P_RT && G_BT == P_BT && G_DW == P_DW && G_COUNTS == P_COUNTS &&
                  G_OFFS == P_OFFS && G_RECS == P_RECS && G_ACC == P_ACC && G_ENT == P_ENT &&
                  G_TW2 == P_TW2 && G_COFF == P_COFF,
              "plan buffer indices of the launch helpers");

/* single calls: the process-wide buffer slots */
static int get_slot(int i, size_t bytes, void *ctx, void **out)
{
    (void)ctx;
    RC(ensure(k_plan_slot[i], bytes));
    *out = g_buf[k_plan_slot[i]];
    return FSX_OK;
}

static int run_locked(const fsx_hip_args *A, fsx_stats *st)
{
    RC(init_once());
    fsx_plan p;
    RC(plan_init(&p, A, 0, 0));
    hipStream_t s = g_stream;
    hipEvent_t ev[7];
    int nev_made = 0, rc = FSX_OK;
    int tail[2] = {0, 0};
    unsigned long long tally[FSX_TALLY_N] = {0};
    float ms[6];
    rc = plan_bind(&p, get_slot, NULL);
    if (rc == FSX_OK) {
        /* the Wigner-d and d-symbol tables are not copied again when the
         * buffer already holds the same table (content keys) */
        const int up_dw = !(A->dw_key && g_key[B_DW] == A->dw_key);
        const int up_ds = !(A->ds_key && g_key[B_DS] == A->ds_key);
        /* the staging event is recorded whatever the upload returns, so
         * that the next call waits for any copy it did enqueue */
        rc = stage_begin(p.sz[P_IN]);
        if (rc == FSX_OK) {
            rc = plan_upload(&p, s, up_dw, up_ds);
            const int r2 = stage_end(s);
            if (rc == FSX_OK) rc = r2;
        }
        if (rc == FSX_OK) {
            g_key[B_DW] = A->dw_key;
            g_key[B_DS] = A->ds_key;
        }
    }
    for (; rc == FSX_OK && nev_made < 7; nev_made++)
        if (hipEventCreate(&ev[nev_made]) != hipSuccess) rc = FSX_ERR_HIP;
    if (rc == FSX_OK && hipEventRecord(ev[0], s) != hipSuccess) rc = FSX_ERR_HIP;
    if (rc == FSX_OK) rc = plan_enqueue(&p, s, ev);
    if (rc == FSX_OK) rc = plan_download(&p, s, tail, tally);
    if (rc == FSX_OK && hipStreamSynchronize(s) != hipSuccess) rc = FSX_ERR_HIP;
    for (int i = 0; rc == FSX_OK && i < 6; i++)
        if (hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]) != hipSuccess) rc = FSX_ERR_HIP;
    if (rc == FSX_OK) plan_stats(&p, st, ms, tail, tally);
    for (int i = 0; i < nev_made; i++) (void)hipEventDestroy(ev[i]);
    plan_free_host(&p);
    return rc;
}

extern "C" int fsx_hip_run(const fsx_hip_args *A, fsx_stats *st)
{
    pthread_mutex_lock(&g_mtx);
    int rc = run_locked(A, st);
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

/* ---- batches: several complexes in one call --------------------------------
 * Each complex gets its own plan and buffer set.  The inputs are uploaded on
 * the library's stream; then every complex's stages are launched on one of
 * FSX_BSTREAMS streams (complex i on stream i mod FSX_BSTREAMS, forked from
 * and joined back into the library's stream), so that the stages of small
 * complexes run side by side.  That launch sequence is captured once into a
 * HIP graph and replayed by one graph launch on later calls whose launch
 * parameters (sizes, tile shapes, scalars, buffer addresses) are all the same.
 * The cross-term tile shape per expansion-order bucket and form comes from a
 * tuner that times the candidates once on the first complex of the bucket. */
#define FSX_BSTREAMS 4
struct fsx_bset {
    void *d[P_NBUF];
    size_t cap[P_NBUF];
    unsigned long long dw_key, ds_key;   /* tables held by d[P_DW], d[P_DS] */
};
static fsx_bset *g_bat = NULL;
static int g_nbat = 0;
/* intermediates (written and read by the stages of one complex only) are
 * shared by the complexes of one stream, which run one after another;
 * inputs uploaded before the stages run, outputs and the counters read
 * back afterwards are per complex */
static fsx_bset g_bstr[FSX_BSTREAMS];
/* persist: the persistent launches read the translation and rotated tables,
 * the buckets, the accumulators and the fit queue of every complex after all
 * complexes' table stages, so those are per complex */
static inline int per_stream_buffer(int i, int persist)
{
    switch (i) {
    case P_T: case P_RT: case P_BT: case P_COUNTS: case P_RECS: case P_ACC: case P_NEXT:
    case P_ACCF:
        return !persist;
    case P_BES: case P_RTMP: case P_KEYS: case P_INFO: case P_FILL: case P_FITSC: case P_SRCT:
        return 1;
    default:
        return 0;
    }
}
struct fsx_bctx {
    fsx_bset *cplx, *strm;
    int persist;
};
static hipStream_t g_bs[FSX_BSTREAMS];
static hipEvent_t g_bfork, g_bjoin[FSX_BSTREAMS];
static int g_binit = 0;
static hipGraph_t g_graph = NULL;
static hipGraphExec_t g_gexec = NULL;
static void *g_gkey = NULL;
static size_t g_gkey_len = 0;

/* the launch parameters a captured graph depends on, per complex */
struct fsx_gkey {
    int L, nb, na, ng, Q, zn, n, flags, half, npk, C, tile_rows, aq_lds;
    long lanes;
    double mult, peak, c1_lo, c1_hi, c2_lo, c2_hi;
    void *d[P_NBUF];
    const void *dposes;
    void *dscore, *dc1, *dc2, *dcross;
};

static int batch_init(void)
{
    if (g_binit) return FSX_OK;
    for (int i = 0; i < FSX_BSTREAMS; i++) {
        HIPCHK(hipStreamCreateWithFlags(&g_bs[i], hipStreamNonBlocking));
        HIPCHK(hipEventCreateWithFlags(&g_bjoin[i], hipEventDisableTiming));
    }
    HIPCHK(hipEventCreateWithFlags(&g_bfork, hipEventDisableTiming));
    g_binit = 1;
    return FSX_OK;
}

static void graph_drop(void)
{
    if (g_gexec) (void)hipGraphExecDestroy(g_gexec);
    if (g_graph) (void)hipGraphDestroy(g_graph);
    g_gexec = NULL;
    g_graph = NULL;
    free(g_gkey);
    g_gkey = NULL;
    g_gkey_len = 0;
}

static int get_bat(int i, size_t bytes, void *ctx, void **out)
{
    const fsx_bctx *c = (const fsx_bctx *)ctx;
    fsx_bset *b = per_stream_buffer(i, c->persist) ? c->strm : c->cplx;
    if (b->cap[i] < bytes) {
        if (b->d[i]) HIPCHK(hipFree(b->d[i]));
        b->d[i] = NULL;
        b->cap[i] = 0;
        if (i == P_DW) b->dw_key = 0;
        if (i == P_DS) b->ds_key = 0;
        const size_t cap = bytes + bytes / 8;
        if (hipMalloc(&b->d[i], cap) != hipSuccess) {
            b->d[i] = NULL;
            return FSX_ERR_NOMEM;
        }
        b->cap[i] = cap;
    }
    *out = b->d[i];
    return FSX_OK;
}


/* ---- tile-shape tuner ----
 * Per expansion-order bucket (L <= 10, 15, 20, 25, 30, 40) and form, the
 * gamma2 rows per cross-term tile with the shortest cross-term stage on the
 * first complex of that bucket; FSX_TUNE_FILE names a text file the table is
 * read from and appended to ("bucket mixed rows slots ms" lines, slots 0:
 * the form's default pairing-slot cap). */
#define FSX_TUNE_BUCKETS 6
static int g_tune[FSX_TUNE_BUCKETS][2][2];   /* rows, slots; 0: not tuned */
static int g_tune_loaded = 0;
static fsx_cross_job *g_cjob_dev;
static size_t g_cjob_cap;
static fsx_fit_job *g_fjob_dev;
static size_t g_fjob_cap;
static void batch_release(void)
{
    graph_drop();
    if (g_cjob_dev) (void)hipFree(g_cjob_dev);
    if (g_fjob_dev) (void)hipFree(g_fjob_dev);
    g_cjob_dev = NULL;
    g_fjob_dev = NULL;
    g_cjob_cap = g_fjob_cap = 0;
    memset(g_tune, 0, sizeof(g_tune));
    g_tune_loaded = 0;
    for (int k = 0; k < g_nbat; k++)
        for (int i = 0; i < P_NBUF; i++)
            if (g_bat[k].d[i]) (void)hipFree(g_bat[k].d[i]);
    for (int k = 0; k < FSX_BSTREAMS; k++)
        for (int i = 0; i < P_NBUF; i++)
            if (g_bstr[k].d[i]) (void)hipFree(g_bstr[k].d[i]);
    memset(g_bstr, 0, sizeof(g_bstr));
    free(g_bat);
    g_bat = NULL;
    g_nbat = 0;
}

static inline int tune_bucket(int L)
{
    return L <= 10 ? 0 : L <= 15 ? 1 : L <= 20 ? 2 : L <= 25 ? 3 : L <= 30 ? 4 : 5;
}

static void tune_load(void)
{
    if (g_tune_loaded) return;
    g_tune_loaded = 1;
    const char *path = getenv("FSX_TUNE_FILE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "r");
    if (!f) return;
    int b, m, r, c;
    double ms;
    while (fscanf(f, "%d %d %d %d %lf", &b, &m, &r, &c, &ms) == 5)
        if (b >= 0 && b < FSX_TUNE_BUCKETS && m >= 0 && m <= 1 && r >= 1 && r <= 4 &&
            (c == 0 || (c >= 64 && c <= 4 * FSX_WAVE))) {
            g_tune[b][m][0] = r;
            g_tune[b][m][1] = c;
        }
    fclose(f);
}

static void tune_save(int b, int m, int r, int c, double ms)
{
    const char *path = getenv("FSX_TUNE_FILE");
    if (!path || !*path) return;
    FILE *f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%d %d %d %d %.4f\n", b, m, r, c, ms);
    fclose(f);
}

/* Time the cross-term stage of complex A for each candidate shape in the
 * process-wide buffers and record the fastest for its bucket. */
static int tune_one(const fsx_hip_args *A)
{
    const int mixed = (A->flags & FSX_MIXED_FP32) != 0, b = tune_bucket(A->L);
    /* gamma2 rows per tile; the pairing-slot cap stays the form's default,
     * since it sets the m-chunks and with them the order of the FP32 sums
     * (rows change no arithmetic: a tuned batch scores every pose exactly as
     * a single call does) */
    static const int cand32[][2] = {{4, 0}, {2, 0}, {1, 0}};
    static const int cand64[][2] = {{2, 0}, {1, 0}, {4, 0}};
    const int nc = 3;
    hipEvent_t ev[7];
    int nev = 0, rc = FSX_OK, best = -1;
    float best_ms = 0.f;
    for (; nev < 7; nev++)
        if (hipEventCreate(&ev[nev]) != hipSuccess) {
            rc = FSX_ERR_HIP;
            break;
        }
    for (int k = 0; rc == FSX_OK && k < nc; k++) {
        const int rows = mixed ? cand32[k][0] : cand64[k][0];
        const int slots = mixed ? cand32[k][1] : cand64[k][1];
        fsx_plan p;
        rc = plan_init(&p, A, rows, slots);
        if (rc != FSX_OK) break;
        rc = plan_bind(&p, get_slot, NULL);
        if (rc == FSX_OK) {
            rc = stage_begin(p.sz[P_IN]);
            if (rc == FSX_OK) {
                rc = plan_upload(&p, g_stream, 1, 1);
                const int r2 = stage_end(g_stream);
                if (rc == FSX_OK) rc = r2;
            }
        }
        if (rc == FSX_OK) {
            g_key[B_DW] = 0;   /* the buffers now hold this complex's tables */
            g_key[B_DS] = 0;
        }
        /* twice: the first run warms the caches and code */
        float ms = 0.f;
        for (int rep = 0; rc == FSX_OK && rep < 2; rep++) {
            if (hipEventRecord(ev[0], g_stream) != hipSuccess) rc = FSX_ERR_HIP;
            if (rc == FSX_OK) rc = plan_enqueue(&p, g_stream, ev, FSX_ST_PREP | FSX_ST_CROSS);
            if (rc == FSX_OK && hipStreamSynchronize(g_stream) != hipSuccess) rc = FSX_ERR_HIP;
            if (rc == FSX_OK && hipEventElapsedTime(&ms, ev[3], ev[4]) != hipSuccess) rc = FSX_ERR_HIP;
        }
        plan_free_host(&p);
        if (rc == FSX_OK && (best < 0 || ms < best_ms)) {
            best = k;
            best_ms = ms;
        }
    }
    for (int i = 0; i < nev; i++) (void)hipEventDestroy(ev[i]);
    if (rc == FSX_OK && best >= 0) {
        g_tune[b][mixed][0] = mixed ? cand32[best][0] : cand64[best][0];
        g_tune[b][mixed][1] = mixed ? cand32[best][1] : cand64[best][1];
        tune_save(b, mixed, g_tune[b][mixed][0], g_tune[b][mixed][1], best_ms);
    }
    return rc;
}


/* the cross-term jobs of plan p appended to the groups (the tile rows of
 * launch_cross) */
static int plan_cross_collect(const fsx_plan *p, fsx_cross_group **gr, int *ngr)
{
    const int G = p->tile_rows > 0 ? p->tile_rows
                                   : env_int("FSX_CROSS_ROWS", p->mixed ? 4 : 2, 1, 4);
    const int *csz = p->ent + 4 * p->K2, *cst = p->ent + 6 * p->K2;
    void *const *g = p->d;
#define FSX_COLLECT(E, GG)                                                                   \
    cross_collect_g<E, GG>(gr, ngr, p->C, csz, cst, p->coff, p->Tn, p->nrows, p->Q, p->L,    \
                           p->nb, p->npk, p->ng, p->zn, g)
    if (p->mixed) return G >= 4 ? FSX_COLLECT(cplx32, 4) : G >= 2 ? FSX_COLLECT(cplx32, 2)
                                                                : FSX_COLLECT(cplx32, 1);
    return G >= 4 ? FSX_COLLECT(cplx, 4) : G >= 2 ? FSX_COLLECT(cplx, 2) : FSX_COLLECT(cplx, 1);
#undef FSX_COLLECT
}

static int grow_dev(void **buf, size_t *cap, size_t bytes)
{
    if (*cap >= bytes) return FSX_OK;
    if (*buf) HIPCHK(hipFree(*buf));
    *buf = NULL;
    *cap = 0;
    if (hipMalloc(buf, bytes) != hipSuccess) {
        *buf = NULL;
        return FSX_ERR_NOMEM;
    }
    *cap = bytes;
    return FSX_OK;
}

static int run_batch_locked(int nprob, const fsx_hip_args *args, fsx_stats *st)
{
    RC(init_once());
    RC(batch_init());
    tune_load();
    const int use_graph = env_int("FSX_BATCH_GRAPH", 1, 0, 1);
    const int tune = env_int("FSX_TUNE", 1, 0, 1);
    /* FSX_BATCH_PERSIST=1: the persistent form, one cross-term launch per
     * (m-chunk index, kernel shape) and one fit launch over all complexes,
     * after every complex's table stages.  Default: each complex's whole
     * pipeline on one of four streams, which on the 64-complex bench batch
     * is faster (674 against 873 ms per mixed step): the streams overlap one
     * complex's tables, another's cross terms and a third's fit, where the
     * persistent form runs those stages one after another (96 ms of tables,
     * 601 ms of cross terms, 189 ms of fits, the last with a block-level
     * tail per complex) */
    const int persist = env_int("FSX_BATCH_PERSIST", 0, 0, 1);
    /* the tuner first (it runs in the process-wide buffers) */
    for (int i = 0; tune && i < nprob; i++) {
        const int m = (args[i].flags & FSX_MIXED_FP32) != 0, b = tune_bucket(args[i].L);
        if (!g_tune[b][m][0] && args[i].n > 0) RC(tune_one(&args[i]));
    }
    if (g_nbat < nprob) {
        fsx_bset *nb = (fsx_bset *)realloc(g_bat, sizeof(fsx_bset) * nprob);
        if (!nb) return FSX_ERR_NOMEM;
        memset(nb + g_nbat, 0, sizeof(fsx_bset) * (nprob - g_nbat));
        g_bat = nb;
        g_nbat = nprob;
    }
    fsx_plan *pl = (fsx_plan *)calloc(nprob, sizeof(fsx_plan));
    int *tail = (int *)calloc(2 * (size_t)nprob, sizeof(int));
    unsigned long long *tally =
        (unsigned long long *)calloc(FSX_TALLY_N * (size_t)nprob, sizeof(unsigned long long));
    /* one key per complex and one for the batch's own launches */
    fsx_gkey *key = (fsx_gkey *)calloc((size_t)nprob + 1, sizeof(fsx_gkey));
    fsx_cross_group *cg = NULL;
    int ncg = 0, nfj = 0;
    fsx_fit_job *fj = NULL;
    size_t fj_lds = 0;
    long fj_lanes = 0;
    int rc = (pl && tail && tally && key) ? FSX_OK : FSX_ERR_NOMEM;
    int ninit = 0;
    for (; rc == FSX_OK && ninit < nprob; ninit++) {
        const fsx_hip_args *A = &args[ninit];
        const int m = (A->flags & FSX_MIXED_FP32) != 0, b = tune_bucket(A->L);
        rc = plan_init(&pl[ninit], A, g_tune[b][m][0], g_tune[b][m][1]);
        if (rc != FSX_OK) break;
    }
    /* the shared buffers of each stream at the size its largest complex
     * needs, before any plan takes their addresses */
    for (int j = 0; rc == FSX_OK && j < FSX_BSTREAMS; j++)
        for (int i = 0; rc == FSX_OK && i < P_NBUF; i++) {
            if (!per_stream_buffer(i, persist)) continue;
            size_t need = 0;
            for (int k = j; k < ninit; k += FSX_BSTREAMS)
                if (pl[k].sz[i] > need) need = pl[k].sz[i];
            if (need == 0) continue;
            fsx_bctx bc = {NULL, &g_bstr[j], persist};
            void *dummy;
            rc = get_bat(i, need, &bc, &dummy);
        }
    size_t stage_bytes = 0;
    for (int k = 0; k < ninit; k++) stage_bytes += (pl[k].sz[P_IN] + 255) & ~(size_t)255;
    const int staged = rc == FSX_OK;
    if (rc == FSX_OK) rc = stage_begin(stage_bytes);
    for (int k = 0; rc == FSX_OK && k < ninit; k++) {
        const fsx_hip_args *A = &args[k];
        fsx_bctx bc = {&g_bat[k], &g_bstr[k % FSX_BSTREAMS], persist};
        rc = plan_bind(&pl[k], get_bat, &bc);
        if (rc != FSX_OK) break;
        fsx_bset *bs = &g_bat[k];
        const int up_dw = !(A->dw_key && bs->dw_key == A->dw_key);
        const int up_ds = !(A->ds_key && bs->ds_key == A->ds_key);
        rc = plan_upload(&pl[k], g_stream, up_dw, up_ds);
        if (rc == FSX_OK) {
            bs->dw_key = A->dw_key;
            bs->ds_key = A->ds_key;
        }
    }
    if (staged && g_stage_ev) {
        const int r2 = stage_end(g_stream);
        if (rc == FSX_OK) rc = r2;
    }
    /* the persistent launches' jobs */
    int fitjobs = 0;
    if (rc == FSX_OK && persist && nprob > 0) {
        for (int k = 0; rc == FSX_OK && k < ninit; k++)
            if (pl[k].n > 0) rc = plan_cross_collect(&pl[k], &cg, &ncg);
        /* launches in m-chunk order, so that each complex's chunks add up in
         * order (one chunk of a complex per launch) */
        for (int i = 1; i < ncg; i++)
            for (int j = i; j > 0 && cg[j].chunk < cg[j - 1].chunk; j--) {
                const fsx_cross_group t = cg[j];
                cg[j] = cg[j - 1];
                cg[j - 1] = t;
            }
        size_t ntot = 0;
        for (int i = 0; i < ncg; i++) ntot += cg[i].njobs;
        const long bpc_env = env_int("FSX_CROSS_BLOCKS_PER_CU", 0, 0, 16);
        const int xcd = env_int("FSX_CROSS_XCD", 1, 0, 1);
        for (int i = 0; rc == FSX_OK && i < ncg; i++) {
            int per = (int)bpc_env;
            if (per == 0 && (hipOccupancyMaxActiveBlocksPerMultiprocessor(
                                 &per, cross_group_kernel(&cg[i]), FSX_CROSS_BLOCK, cg[i].lds) !=
                                 hipSuccess ||
                             per < 1))
                per = 2;
            const long gsz = (long)g_ncu * per;
            for (int j = 0; j < cg[i].njobs; j++) {
                fsx_cross_job *jb = &cg[i].jobs[j];
                jb->nx = xcd && g_nxcd > 1 && gsz % g_nxcd == 0 && jb->ntiles >= 4L * gsz ? g_nxcd : 1;
            }
            cg[i].grid = gsz;
        }
        if (rc == FSX_OK) rc = grow_dev((void **)&g_cjob_dev, &g_cjob_cap,
                                        sizeof(fsx_cross_job) * (ntot > 0 ? ntot : 1));
        for (int i = 0, o = 0; rc == FSX_OK && i < ncg; o += cg[i].njobs, i++)
            if (hipMemcpy(g_cjob_dev + o, cg[i].jobs, sizeof(fsx_cross_job) * cg[i].njobs,
                          hipMemcpyHostToDevice) != hipSuccess)
                rc = FSX_ERR_HIP;
        /* one fit launch when every complex with poses takes the lane-group
         * fit */
        fitjobs = 1;
        for (int k = 0; k < ninit; k++)
            if (pl[k].n > 0 && !pl[k].grp) fitjobs = 0;
        if (fitjobs && rc == FSX_OK) {
            fj = (fsx_fit_job *)calloc((size_t)ninit, sizeof(fsx_fit_job));
            if (!fj) rc = FSX_ERR_NOMEM;
            for (int k = 0; rc == FSX_OK && k < ninit; k++) {
                if (pl[k].n == 0) continue;
                if (pl[k].lds_fit > fj_lds) fj_lds = pl[k].lds_fit;
            }
            int bpc = env_int("FSX_FIT_BLOCKS_PER_CU", 0, 0, 16);
            if (rc == FSX_OK && bpc == 0 &&
                (hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpc, k_fitg_jobs, 256, fj_lds) !=
                     hipSuccess ||
                 bpc < 1))
                bpc = 1;
            fj_lanes = (long)g_ncu * 256 * bpc;
            for (int k = 0; rc == FSX_OK && k < ninit; k++) {
                fsx_plan *p = &pl[k];
                if (p->n == 0) continue;
                const fsx_hip_args *A = &p->A;
                void **g = p->d;
                fsx_fit_job *f = &fj[nfj++];
                f->acc = (const float *)g[P_ACCF];
                f->qvals = (const double *)g[P_Q];
                f->a = (const double *)g[P_AM];
                f->recs = (const int4 *)g[P_RECS];
                f->nvalid_p = (const int *)g[P_OFFS] + p->nkeys;
                f->mult = A->mult;
                f->c1_lo = A->c1_lo; f->c1_hi = A->c1_hi; f->c2_lo = A->c2_lo; f->c2_hi = A->c2_hi;
                f->score = p->dscore; f->c1o = p->dc1; f->c2o = p->dc2;
                f->next = (int *)g[P_NEXT];
                f->tally = (unsigned long long *)g[P_TALLY];
                f->Q = p->Q;
                f->fpl = p->fit_wide;
                /* each job's queue starts past the groups the launch holds */
                p->fit_next0 = fj_lanes / p->fit_wide;
            }
            if (rc == FSX_OK) rc = grow_dev((void **)&g_fjob_dev, &g_fjob_cap,
                                            sizeof(fsx_fit_job) * (nfj > 0 ? nfj : 1));
            if (rc == FSX_OK && nfj > 0 &&
                hipMemcpy(g_fjob_dev, fj, sizeof(fsx_fit_job) * nfj, hipMemcpyHostToDevice) !=
                    hipSuccess)
                rc = FSX_ERR_HIP;
        }
    }
    for (int k = 0; rc == FSX_OK && k < ninit; k++) {
        const fsx_hip_args *A = &args[k];
        fsx_gkey *ky = &key[k];
        const fsx_plan *p = &pl[k];
        ky->L = p->L; ky->nb = p->nb; ky->na = p->na; ky->ng = p->ng; ky->Q = p->Q; ky->zn = p->zn;
        ky->n = p->n; ky->flags = (int)A->flags; ky->half = p->half; ky->npk = p->npk; ky->C = p->C;
        ky->tile_rows = p->tile_rows; ky->aq_lds = p->aq_lds; ky->lanes = p->lanes;
        ky->mult = A->mult; ky->peak = A->peak; ky->c1_lo = A->c1_lo; ky->c1_hi = A->c1_hi;
        ky->c2_lo = A->c2_lo; ky->c2_hi = A->c2_hi;
        memcpy(ky->d, p->d, sizeof(ky->d));
        ky->dposes = p->dposes; ky->dscore = p->dscore; ky->dc1 = p->dc1; ky->dc2 = p->dc2;
        ky->dcross = p->dcross;
    }
    if (rc == FSX_OK) {
        /* the batch's own launches: the form, the job tables and their grids */
        fsx_gkey *ky = &key[nprob];
        ky->n = persist;
        ky->flags = fitjobs;
        ky->lanes = fj_lanes;
        ky->d[0] = g_cjob_dev;
        ky->d[1] = g_fjob_dev;
        ky->C = ncg;
        ky->Q = nfj;
    }
    /* the stages of every complex: one graph launch, or the launches */
    if (rc == FSX_OK && nprob > 0) {
        const size_t klen = sizeof(fsx_gkey) * ((size_t)nprob + 1);
        const int replay = use_graph && g_gexec && g_gkey_len == klen && memcmp(g_gkey, key, klen) == 0;
        if (!replay) {
            graph_drop();
            hipError_t e = use_graph ? hipStreamBeginCapture(g_stream, hipStreamCaptureModeRelaxed)
                                     : hipSuccess;
            /* fork to the streams, stages, join back on g_stream */
            auto fork = [&]() {
                hipError_t r = hipEventRecord(g_bfork, g_stream);
                for (int j = 0; r == hipSuccess && j < FSX_BSTREAMS; j++)
                    r = hipStreamWaitEvent(g_bs[j], g_bfork, 0);
                return r;
            };
            auto join = [&]() {
                hipError_t r = hipSuccess;
                for (int j = 0; r == hipSuccess && j < FSX_BSTREAMS; j++) {
                    r = hipEventRecord(g_bjoin[j], g_bs[j]);
                    if (r == hipSuccess) r = hipStreamWaitEvent(g_stream, g_bjoin[j], 0);
                }
                return r;
            };
            if (e == hipSuccess) e = fork();
            if (e != hipSuccess) rc = FSX_ERR_HIP;
            for (int i = 0; rc == FSX_OK && i < nprob; i++)
                rc = plan_enqueue(&pl[i], g_bs[i % FSX_BSTREAMS], NULL,
                                  persist ? FSX_ST_PREP : FSX_ST_ALL);
            if (rc == FSX_OK && join() != hipSuccess) rc = FSX_ERR_HIP;
            if (persist) {
                for (int i = 0, o = 0; rc == FSX_OK && i < ncg; o += cg[i].njobs, i++) {
                    const fsx_cross_job *jobs = g_cjob_dev + o;
                    int nj = cg[i].njobs;
                    void *kargs[2] = {(void *)&jobs, (void *)&nj};
                    if (hipLaunchKernel(cross_group_kernel(&cg[i]), dim3((unsigned)cg[i].grid),
                                        dim3(FSX_CROSS_BLOCK), kargs, cg[i].lds, g_stream) !=
                        hipSuccess)
                        rc = FSX_ERR_HIP;
                }
                if (rc == FSX_OK && fork() != hipSuccess) rc = FSX_ERR_HIP;
                for (int i = 0; rc == FSX_OK && i < nprob; i++)
                    rc = plan_enqueue(&pl[i], g_bs[i % FSX_BSTREAMS], NULL,
                                      FSX_ST_FITPREP | (fitjobs ? 0 : FSX_ST_FIT));
                if (rc == FSX_OK && join() != hipSuccess) rc = FSX_ERR_HIP;
                if (rc == FSX_OK && fitjobs && nfj > 0) {
                    hipLaunchKernelGGL(k_fitg_jobs, dim3((unsigned)(fj_lanes / 256)), dim3(256),
                                       fj_lds, g_stream, (const fsx_fit_job *)g_fjob_dev, nfj);
                    if (hipGetLastError() != hipSuccess) rc = FSX_ERR_HIP;
                }
            }
            if (use_graph) {
                hipGraph_t gr = NULL;
                const hipError_t ee = hipStreamEndCapture(g_stream, &gr);
                if (rc == FSX_OK && ee != hipSuccess) rc = FSX_ERR_HIP;
                if (rc == FSX_OK && hipGraphInstantiate(&g_gexec, gr, NULL, NULL, 0) != hipSuccess) {
                    g_gexec = NULL;
                    rc = FSX_ERR_HIP;
                }
                if (rc == FSX_OK) {
                    g_graph = gr;
                    g_gkey = malloc(klen);
                    if (g_gkey) {
                        memcpy(g_gkey, key, klen);
                        g_gkey_len = klen;
                    }
                } else if (gr) {
                    (void)hipGraphDestroy(gr);
                }
            }
        }
        if (rc == FSX_OK && use_graph && hipGraphLaunch(g_gexec, g_stream) != hipSuccess)
            rc = FSX_ERR_HIP;
    }
    for (int i = 0; rc == FSX_OK && i < nprob; i++)
        rc = plan_download(&pl[i], g_stream, tail + 2 * i, tally + FSX_TALLY_N * i);
    if (hipStreamSynchronize(g_stream) != hipSuccess && rc == FSX_OK) rc = FSX_ERR_HIP;
    for (int i = 0; rc == FSX_OK && i < nprob; i++)
        plan_stats(&pl[i], &st[i], NULL, tail + 2 * i, tally + FSX_TALLY_N * i);
    if (rc != FSX_OK) graph_drop();
    for (int i = 0; i < ninit; i++) plan_free_host(&pl[i]);
    for (int i = 0; i < ncg; i++) free(cg[i].jobs);
    free(cg);
    free(fj);
    free(pl);
    free(tail);
    free(tally);
    free(key);
    return rc;
}

extern "C" int fsx_hip_run_batch(int nprob, const fsx_hip_args *args, fsx_stats *st)
{
    pthread_mutex_lock(&g_mtx);
    int rc = run_batch_locked(nprob, args, st);
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
            hipLaunchKernelGGL(k_fit_batch, dim3(nblk(n, 256)), dim3(256),
                               (size_t)FSX_LB_KN * FSX_LB_KSTRIDE * sizeof(double), s,
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

/* Atom-expansion workspace: the per-chunk partial sums of all natoms atoms,
 * then the tables of one pass of atoms (Bessel table, f, r, conj(Y)). */
static inline int at_qp(int Q) { return (Q + 15) / 16 * 16; }
static inline size_t at_part_bytes(int natoms, int Q, int L)
{
    const size_t Kh = (size_t)(L + 1) * (L + 2) / 2, nch = ((size_t)natoms + FSX_AT_CHUNK - 1) / FSX_AT_CHUNK;
    return (nch * 3 * at_qp(Q) * Kh * 2 * sizeof(double) + 255) & ~(size_t)255;
}
static inline size_t at_atom_bytes(int Q, int L)
{
    const size_t K = L + 1, Kh = K * (K + 1) / 2;
    return sizeof(double) * (K * at_qp(Q) + 4 + 2 * Kh);
}

/* The expansion of natoms atoms (device inputs) into re/im [3][Q][(L+1)^2]
 * (device), on stream s, in ws (ws_bytes): passes of as many whole chunks of
 * atoms as the workspace holds. */
static int atoms_run(hipStream_t s, char *ws, size_t ws_bytes, int natoms, const double *xyz,
                     const double *vac, const double *dum, const double *sa, double h2o_ff,
                     const double *qv, int Q, int L, double *re, double *im)
{
    const int K = L + 1, Kh = K * (K + 1) / 2, Qp = at_qp(Q);
    const int nch = (natoms + FSX_AT_CHUNK - 1) / FSX_AT_CHUNK;
    const size_t pb = at_part_bytes(natoms, Q, L), per = at_atom_bytes(Q, L);
    if (ws_bytes < pb + per * FSX_AT_CHUNK && ws_bytes < pb + per * (size_t)natoms) return FSX_ERR_ARG;
    long npass = (long)((ws_bytes - pb) / per) / FSX_AT_CHUNK * FSX_AT_CHUNK;
    if (npass < 1) npass = natoms;   /* a single pass of fewer than one chunk of atoms */
    if (npass > natoms) npass = natoms;
    double *P = reinterpret_cast<double *>(ws);
    double *J = reinterpret_cast<double *>(ws + pb);
    double *f = J + (size_t)K * npass * Qp, *r = f + 3 * (size_t)npass, *Y = r + npass;
    for (long a0 = 0; a0 < natoms; a0 += npass) {
        const int na = (int)(natoms - a0 < npass ? natoms - a0 : npass);
        hipLaunchKernelGGL(k_atom_geom, dim3(nblk(na, 256)), dim3(256), 0, s, xyz + 3 * a0,
                           vac + a0, dum + a0, sa ? sa + a0 : (const double *)NULL, h2o_ff, na, L,
                           FSX_PI_REFERENCE, r, f, Y);
        hipLaunchKernelGGL(k_atom_bes, dim3(nblk((long)K * na * Qp, 256)), dim3(256), 0, s, J,
                           (const double *)r, qv, na, Q, Qp, K);
        const int nchp = (na + FSX_AT_CHUNK - 1) / FSX_AT_CHUNK;
        hipLaunchKernelGGL(k_atom_mfma, dim3((unsigned)(nchp * K * (Qp / 16))), dim3(256), 0, s,
                           (const double *)J, (const double *)f, (const double *)Y, P, na, Qp, L,
                           (int)(a0 / FSX_AT_CHUNK));
    }
    hipLaunchKernelGGL(k_atom_finish, dim3(nblk(3L * Q * Kh, 256)), dim3(256), 0, s,
                       (const double *)P, nch, re, im, Q, Qp, L, 4.0 * FSX_PI_REFERENCE);
    return hipGetLastError() == hipSuccess ? FSX_OK : FSX_ERR_HIP;
}

/* The mixed form of atoms_run in one pass over all atoms: the BF16 parts of
 * the Bessel values and of f conj(Y_lm), then the split products on BF16
 * MFMA.  Workspace: the chunk partials, then per atom (padded to whole
 * chunks) f, r, Y in FP64 and the parts. */
static inline size_t at_bf16_atom_bytes(int Q, int L)
{
    const size_t K = L + 1, Kh = K * (K + 1) / 2;
    return sizeof(double) * (4 + 2 * Kh) + 3 * sizeof(__bf16) * (K * at_qp(Q) + 6 * Kh);
}
static int atoms_run_bf16(hipStream_t s, char *ws, size_t ws_bytes, int natoms,
                          const double *xyz, const double *vac, const double *dum,
                          const double *sa, double h2o_ff, const double *qv, int Q, int L,
                          double *re, double *im)
{
    const int K = L + 1, Kh = K * (K + 1) / 2, Qp = at_qp(Q);
    const int nch = (natoms + FSX_AT_CHUNK - 1) / FSX_AT_CHUNK, nap = nch * FSX_AT_CHUNK;
    const size_t pb = at_part_bytes(natoms, Q, L);
    if (ws_bytes < pb + at_bf16_atom_bytes(Q, L) * (size_t)nap) return FSX_ERR_ARG;
    double *P = reinterpret_cast<double *>(ws);
    double *f = reinterpret_cast<double *>(ws + pb), *r = f + 3 * (size_t)natoms;
    double *Y = r + natoms;
    __bf16 *Js = reinterpret_cast<__bf16 *>(Y + (size_t)natoms * Kh * 2);
    __bf16 *Bs = Js + 3 * (size_t)K * Qp * nap;
    hipLaunchKernelGGL(k_atom_geom, dim3(nblk(natoms, 256)), dim3(256), 0, s, xyz, vac, dum, sa,
                       h2o_ff, natoms, L, FSX_PI_REFERENCE, r, f, Y);
    hipLaunchKernelGGL(k_atom_bes_split, dim3(nblk((long)K * Qp * nap, 256)), dim3(256), 0, s, Js,
                       (const double *)r, qv, natoms, nap, Q, Qp, K);
    hipLaunchKernelGGL(k_atom_bsplit, dim3(nblk(6L * Kh * nap, 256)), dim3(256), 0, s, Bs,
                       (const double *)f, (const double *)Y, natoms, nap, L);
    hipLaunchKernelGGL(k_atom_bf16, dim3((unsigned)(nch * K * (Qp / 16))), dim3(256), 0, s,
                       (const __bf16 *)Js, (const __bf16 *)Bs, P, nap, Qp, L, 0);
    hipLaunchKernelGGL(k_atom_finish, dim3(nblk(3L * Q * Kh, 256)), dim3(256), 0, s,
                       (const double *)P, nch, re, im, Q, Qp, L, 4.0 * FSX_PI_REFERENCE);
    return hipGetLastError() == hipSuccess ? FSX_OK : FSX_ERR_HIP;
}

extern "C" int fsx_hip_atoms2spf(int natoms, const double *xyz, const double *vac,
                                 const double *dum, const double *sa, double h2o_ff,
                                 const double *qvals, int Q, int L, double *re, double *im,
                                 int bf16)
{
    pthread_mutex_lock(&g_mtx);
    const int K = L + 1, K2 = K * K;
    const size_t nap = ((size_t)natoms + FSX_AT_CHUNK - 1) / FSX_AT_CHUNK * FSX_AT_CHUNK;
    const size_t wsb = at_part_bytes(natoms, Q, L) +
                       (bf16 ? at_bf16_atom_bytes(Q, L) * nap : at_atom_bytes(Q, L) * (size_t)natoms);
    int rc = init_once();
    if (rc == FSX_OK) rc = ensure(B_AXYZ, sizeof(double) * 3 * (size_t)natoms);
    if (rc == FSX_OK) rc = ensure(B_AFF, sizeof(double) * 3 * (size_t)natoms);
    if (rc == FSX_OK) rc = ensure(B_AWS, wsb);
    if (rc == FSX_OK) rc = ensure(B_AQ, sizeof(double) * Q);
    if (rc == FSX_OK) rc = ensure(B_ARE, sizeof(double) * 3 * (size_t)Q * K2);
    if (rc == FSX_OK) rc = ensure(B_AIM, sizeof(double) * 3 * (size_t)Q * K2);
    if (rc == FSX_OK) {
        hipStream_t s = g_stream;
        double *dff = (double *)g_buf[B_AFF];
        const size_t nb = sizeof(double) * (size_t)natoms;
        hipError_t e = hipMemcpyAsync(g_buf[B_AXYZ], xyz, 3 * nb, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dff, vac, nb, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(dff + natoms, dum, nb, hipMemcpyHostToDevice, s);
        if (e == hipSuccess && sa) e = hipMemcpyAsync(dff + 2 * (size_t)natoms, sa, nb, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) e = hipMemcpyAsync(g_buf[B_AQ], qvals, sizeof(double) * Q, hipMemcpyHostToDevice, s);
        if (e == hipSuccess) {
            rc = (bf16 ? atoms_run_bf16 : atoms_run)(
                s, (char *)g_buf[B_AWS], g_cap[B_AWS], natoms, (const double *)g_buf[B_AXYZ], dff,
                dff + natoms, sa ? dff + 2 * (size_t)natoms : NULL, h2o_ff,
                (const double *)g_buf[B_AQ], Q, L, (double *)g_buf[B_ARE], (double *)g_buf[B_AIM]);
            if (rc == FSX_ERR_HIP) e = hipErrorUnknown;
        }
        if (e == hipSuccess && rc == FSX_OK)
            e = hipMemcpyAsync(re, g_buf[B_ARE], sizeof(double) * 3 * (size_t)Q * K2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess && rc == FSX_OK)
            e = hipMemcpyAsync(im, g_buf[B_AIM], sizeof(double) * 3 * (size_t)Q * K2, hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            fprintf(stderr, "[fsx] atoms2spf: %s\n", hipGetErrorString(e));
            rc = FSX_ERR_HIP;
        }
    }
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

/* The atom expansion on caller-owned device arrays with no allocation: its
 * tables and chunk partials live in the fixed scratch array.  Nothing there
 * needs clearing first: every pass writes its tables whole, and every chunk
 * partial is written (all of (X, q, l, m, part)) before the final sum reads
 * it.  The library's lock orders calls from several threads of one process. */
extern "C" int fsx_atoms2spf_noalloc(int natoms, const double *coords, const double *vac,
                                     const double *dum, const double *sa, double h2o_ff,
                                     const double *qvals, int qnum, int L, double *re, double *im)
{
    if (natoms < 1 || !coords || !vac || !dum || !qvals || qnum < 1 || L < 0 || L > 40 || !re ||
        !im)
        return FSX_ERR_ARG;
    const size_t need = at_part_bytes(natoms, qnum, L) +
                        at_atom_bytes(qnum, L) * (size_t)(natoms < FSX_AT_CHUNK ? natoms : FSX_AT_CHUNK);
    if (need > FSX_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "[fsx] atoms2spf_noalloc: %d atoms at L = %d with %d q need %zu bytes of "
                "working memory, more than the %zu-byte scratch array\n", natoms, L, qnum, need,
                FSX_NOALLOC_SCRATCH_BYTES);
        return FSX_ERR_SCRATCH;
    }
    pthread_mutex_lock(&g_mtx);
    int rc = init_once();
    if (rc == FSX_OK) {
        rc = atoms_run(g_stream, (char *)g_scratch, FSX_NOALLOC_SCRATCH_BYTES, natoms, coords, vac,
                       dum, sa, sa ? h2o_ff : 0.0, qvals, qnum, L, re, im);
        const hipError_t e = hipStreamSynchronize(g_stream);
        if (rc == FSX_OK && e != hipSuccess) {
            fprintf(stderr, "[fsx] atoms2spf_noalloc: %s\n", hipGetErrorString(e));
            rc = FSX_ERR_HIP;
        }
    }
    pthread_mutex_unlock(&g_mtx);
    return rc;
}

