/*
 * fsx_lbfgsb.h -- the per-pose c1/c2 chi^2 fit: bounded L-BFGS-B (n = 2,
 * m = 3) on the interpolated chi^2 objective with the closed-form scale.
 *
 * One implementation compiled twice: as plain C11 for the host entry points
 * (sxs_lbfgs_fitting, sxs_fit_params) and as HIP device code for the fused
 * per-pose fit kernel.  State lives in a fixed-size struct (no allocation)
 * whose every subscript is a compile-time constant after unrolling, so a
 * GPU thread owns one fit and keeps its state in registers.
 *
 * Semantics follow the reference:
 *   objective    min_saxs.c:3-105 (gradient), :261-319 (sxs_best_scale)
 *   wrapper      min_saxs.c:196-259 (x0 = (1, 0), m = 3, factr 1e7, pgtol 1e-5,
 *                scale refreshed before every f/g evaluation)
 *   optimiser    lbfgsb/src (L-BFGS-B 3.0): setulb/mainlb lbfgsb.c:52-515,
 *                cauchy/freev/formk/formt/cmprlb/subsm/matupd/projgr/hpsolb/bmv
 *                in subalgorithms.c, lnsrlb/dcsrch/dcstep in linesearch.c,
 *                dpofa/dtrsl in linpack.c.
 * The reverse-communication state machine of setulb is run as one direct
 * loop (the caller in min_saxs.c resumes every NEW_X immediately, so the
 * visited points are the same).  Every arithmetic expression keeps the
 * reference's evaluation order; compile with FP contraction off so that the
 * iterates are bit-identical to the reference's for identical inputs.
 * The reference loop has no iteration cap (min_saxs.c:234-248);
 * FSX_LB_MAX_EVALS bounds it here.
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
#ifndef FSX_LBFGSB_H
#define FSX_LBFGSB_H

#include <math.h>
#include <float.h>

#ifdef __HIPCC__
/* inlined before any of these bodies is optimised on its own, so that the
 * optimiser sees the fit state as one local object of the kernel */
#define FSX_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define FSX_HD static inline
#endif

#define FSX_LB_N 2
#define FSX_LB_M 3
#define FSX_LB_MAX_EVALS 2000

/* task codes (lbfgsb.h:80-130 numbering) */
#define FSX_T_START 1
#define FSX_T_NEW_X 2
#define FSX_T_ABNORMAL 3
#define FSX_T_FG 10
#define FSX_T_FG_LN 11
#define FSX_T_CONVERGENCE 20
#define FSX_T_CONV_GRAD 21
#define FSX_T_CONV_F 22
#define FSX_T_STOP_ITER 32
#define FSX_T_W_ROUND 101
#define FSX_T_W_XTOL 102
#define FSX_T_W_STPMAX 103
#define FSX_T_W_STPMIN 104
#define FSX_T_E_SMALLSTP 201
#define FSX_T_E_LARGESTP 202
#define FSX_T_E_INITIAL 203
#define FSX_T_E_STP0 207
#define FSX_T_E_STP1 208

/* Cross-term profile of one pose: value of term k (0..5 = VV,VD,VW,DD,DW,WW)
 * at q index i is scale * (c[k*cks + i*cqs] + v[k*vks + i*vqs]); v may be
 * NULL.  This lets the GPU read (constant part + pose part) in place and
 * apply the peak rescale on the fly without a private copy.  With c NULL,
 * v holds the final values (already summed and rescaled). */
typedef struct fsx_ct {
    const double *c;
    const double *v;
    long cks, cqs, vks, vqs;
    double scale;
} fsx_ct;

FSX_HD double fsx_ct_get(const fsx_ct *t, int k, int i)
{
    if (!t->c) return t->v[k * t->vks + i * t->vqs];
    double x = t->c[k * t->cks + i * t->cqs];
    if (t->v) x = x + t->v[k * t->vks + i * t->vqs];
    return x * t->scale;
}

typedef struct fsx_fitpar {
    const double *a;      /* 6*qnum moments (scoring_helper) */
    const double *q;      /* model q grid */
    /* optional (may be NULL): rdq[i] = 1.0 / (q[i] - q[i-1]) with q[-1] =
     * -1.0, the reciprocal the objective forms per q (min_saxs.c:60);
     * precomputed by the same division, so the values are identical */
    const double *rdq;
    int qnum;
    double mult;          /* (4pi/3)^1.5 rm^2 / (16 pi) */
    double lo0, hi0, lo1, hi1;
    double factr, pgtol;
    /* onepass (the mixed-precision form): the evaluation by fsx_eval_1p, in
     * one pass over q, with a0sum = the sum of a[6 i] over q in q order */
    int onepass;
    double a0sum;
} fsx_fitpar;

/* The six terms at q index i, each (c + v) * scale as fsx_ct_get. */
FSX_HD void fsx_ct_get6(const fsx_ct *t, int i, double *o)
{
    if (!t->c) {
        const double *v = t->v + i * t->vqs;
        for (int k = 0; k < 6; k++) o[k] = v[k * t->vks];
        return;
    }
    const double *c = t->c + i * t->cqs;
    if (t->v) {
        const double *v = t->v + i * t->vqs;
        for (int k = 0; k < 6; k++) {
            double x = c[k * t->cks];
            x = x + v[k * t->vks];
            o[k] = x * t->scale;
        }
    } else {
        for (int k = 0; k < 6; k++) o[k] = c[k * t->cks] * t->scale;
    }
}

/* Terms at q = i0 .. i0+nb-1 (nb <= FSX_QB) into blk: all raw loads first,
 * then (c + v) * scale per value exactly as fsx_ct_get.  One q per block:
 * the GPU fit kernel runs the evaluation and the whole L-BFGS-B update in
 * one loop, and every copy of a q step the unrolled loops make adds to code
 * that the instruction cache (64 KB per pair of CUs) must hold; with four q
 * per block the kernel's loop no longer fitted and the fit ran 5.5 ms
 * instead of 5.2 ms on the 210,000-pose workload. */
#ifndef FSX_QB
#define FSX_QB 1
#endif
FSX_HD void fsx_ct_block(const fsx_ct *t, int i0, int nb, double blk[FSX_QB][6])
{
    if (!t->c) {
        const double *v = t->v + i0 * t->vqs;
        const long vks = t->vks, vqs = t->vqs;
#if defined(__HIP_DEVICE_COMPILE__)
        if (vks == 1 && vqs == 6) {
            /* pose-major profile (the fit kernel's accumulator): the six
             * terms of one q are contiguous and 16-byte aligned, so they are
             * read as three pairs */
            const double2 *v2 = reinterpret_cast<const double2 *>(v);
#pragma unroll
            for (int u = 0; u < FSX_QB; u++)
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double2 p = u < nb ? v2[u * 3 + k] : make_double2(0.0, 0.0);
                    blk[u][2 * k] = p.x;
                    blk[u][2 * k + 1] = p.y;
                }
            return;
        }
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) blk[u][k] = u < nb ? v[u * vqs + k * vks] : 0.0;
        return;
    }
    double cr[FSX_QB][6];
    const double *c = t->c + i0 * t->cqs;
    const long cks = t->cks, cqs = t->cqs;
#ifdef __HIPCC__
#pragma unroll
#endif
    for (int u = 0; u < FSX_QB; u++)
        for (int k = 0; k < 6; k++) cr[u][k] = u < nb ? c[u * cqs + k * cks] : 0.0;
    if (t->v) {
        double vr[FSX_QB][6];
        const double *v = t->v + i0 * t->vqs;
        const long vks = t->vks, vqs = t->vqs;
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) vr[u][k] = u < nb ? v[u * vqs + k * vks] : 0.0;
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) {
                double x = cr[u][k];
                x = x + vr[u][k];
                blk[u][k] = x * t->scale;
            }
    } else {
#ifdef __HIPCC__
#pragma unroll
#endif
        for (int u = 0; u < FSX_QB; u++)
            for (int k = 0; k < 6; k++) blk[u][k] = cr[u][k] * t->scale;
    }
}

/* I(q) at G = c1^3 exp(-mult (c1^2 - 1) q^2) and c2 (min_saxs.c:285-287) */
#define FSX_IN(v, G, c2)                                                              \
    ((v)[0] - (G) * (v)[1] + (c2) * (v)[2] + (G) * (G) * (v)[3] - (G) * (c2) * (v)[4] \
     + (c2) * (c2) * (v)[5])

FSX_HD double fsx_in(const fsx_ct *t, int i, double G, double c2)
{
    double v[6];
    fsx_ct_get6(t, i, v);
    return FSX_IN(v, G, c2);
}

/* The q loops below walk the profile in blocks of FSX_QB q values, two
 * blocks at a time: the loads of the next block are issued before the steps
 * of the current one run, so that a GPU thread's profile reads overlap its
 * arithmetic.  Each q step's arithmetic and the order of the sums are those
 * of the reference loops.  BODY(cur, i) runs the step of q index i on the six
 * terms cur[0..5]. */
#define FSX_QB_MIN(a, b) ((a) < (b) ? (a) : (b))
#define FSX_QLOOP(t, qnum, bA, BODY)                                                  \
    do {                                                                              \
        double bB_[FSX_QB][6];                                                        \
        for (int i0_ = 0; i0_ < (qnum); i0_ += 2 * FSX_QB) {                          \
            if (i0_ + FSX_QB < (qnum))                                                \
                fsx_ct_block((t), i0_ + FSX_QB, FSX_QB_MIN((qnum) - i0_ - FSX_QB, FSX_QB), bB_); \
            FSX_QLOOP_UNROLL                                                          \
            for (int u_ = 0; u_ < FSX_QB; u_++)                                       \
                if (i0_ + u_ < (qnum)) { BODY(bA[u_], i0_ + u_); }                    \
            if (i0_ + 2 * FSX_QB < (qnum))                                            \
                fsx_ct_block((t), i0_ + 2 * FSX_QB,                                   \
                             FSX_QB_MIN((qnum) - i0_ - 2 * FSX_QB, FSX_QB), bA);      \
            FSX_QLOOP_UNROLL                                                          \
            for (int u_ = 0; u_ < FSX_QB; u_++)                                       \
                if (i0_ + FSX_QB + u_ < (qnum)) { BODY(bB_[u_], i0_ + FSX_QB + u_); } \
        }                                                                             \
    } while (0)
#ifdef __HIPCC__
#define FSX_QLOOP_UNROLL _Pragma("unroll")
#else
#define FSX_QLOOP_UNROLL
#endif

/* The per-q terms of the two passes, written once so that every evaluation
 * (the sequential q loops below and the GPU's wavefront-wide evaluation,
 * which forms the terms of many q at once and sums them in q order) forms
 * them by the same operations in the same order.
 *
 * FSX_BS_TERMS: what one q step of the scale pass adds to up and down
 * (min_saxs.c:302-310), from I at this q and the previous one. */
#define FSX_BS_TERMS(in, in_prev, qc, q_prev, ai, ut, dt)                             \
    do {                                                                              \
        const double tan_ = ((in) - (in_prev)) / ((qc) - (q_prev));                   \
        const double buf_ = (in) - tan_ * (qc);                                       \
        (ut) = buf_ * (ai)[1] + tan_ * (ai)[2];                                       \
        (dt) = buf_ * buf_ * (ai)[3] + 2.0 * tan_ * buf_ * (ai)[4]                    \
             + tan_ * tan_ * (ai)[5];                                                 \
    } while (0)

/* FSX_OBJ_VALS: I and its derivatives in c1 and c2 at one q (min_saxs.c:
 * 55-68), from the six terms cur, G, Gd = dG/dc1 and c2. */
#define FSX_OBJ_VALS(cur, G, Gd, c2, in, d1, d2)                                      \
    do {                                                                              \
        const double vd_ = (cur)[1], vw_ = (cur)[2], dd_ = (cur)[3];                  \
        const double dw_ = (cur)[4], ww_ = (cur)[5];                                  \
        (in) = (cur)[0] - (G) * vd_ + (c2) * vw_ + (G) * (G) * dd_ - (G) * (c2) * dw_ \
             + (c2) * (c2) * ww_;                                                     \
        (d1) = (Gd) * (-vd_ + 2.0 * (G) * dd_ - (c2) * dw_);                          \
        (d2) = vw_ - (G) * dw_ + 2.0 * (c2) * ww_;                                    \
    } while (0)

/* FSX_OBJ_TERMS: what one q step of the objective adds to g0, g1 and the
 * score (min_saxs.c:70-93), from the values at this q and the previous
 * one, rq = 1 / (q_i - q_{i-1}) and the scale k. */
#define FSX_OBJ_TERMS(in, d1, d2, in_prev, d1p, d2p, rq, qc, ai, k, tg0, tg1, tsc)     \
    do {                                                                              \
        const double tan_ = ((in) - (in_prev)) * (rq);                                \
        const double t1_ = ((d1) - (d1p)) * (rq);                                     \
        const double t2_ = ((d2) - (d2p)) * (rq);                                     \
        (tg0) = 2.0 * (k) * (-((d1) - t1_ * (qc)) * (ai)[1] - t1_ * (ai)[2]           \
                             + (k) * (((in) - tan_ * (qc)) * ((d1) - t1_ * (qc)) * (ai)[3] \
                                      + ((in) * t1_ + (d1) * tan_ - 2.0 * tan_ * t1_ * (qc)) \
                                            * (ai)[4]                                 \
                                      + tan_ * t1_ * (ai)[5]));                       \
        (tg1) = 2.0 * (k) * (-((d2) - t2_ * (qc)) * (ai)[1] - t2_ * (ai)[2]           \
                             + (k) * (((in) - tan_ * (qc)) * ((d2) - t2_ * (qc)) * (ai)[3] \
                                      + ((in) * t2_ + (d2) * tan_ - 2.0 * tan_ * t2_ * (qc)) \
                                            * (ai)[4]                                 \
                                      + tan_ * t2_ * (ai)[5]));                       \
        const double b_ = (in) - tan_ * (qc);                                         \
        (tsc) = (ai)[0] + (k) * (-2.0 * b_ * (ai)[1] - 2.0 * tan_ * (ai)[2]           \
                                 + (k) * (b_ * b_ * (ai)[3] + 2.0 * b_ * tan_ * (ai)[4] \
                                          + tan_ * tan_ * (ai)[5]));                  \
    } while (0)

/* FSX_1P_TERMS: the six terms one q step adds to the sums of the one-pass
 * evaluation (fsx_eval_1p), from the values at this q and the previous one
 * (as FSX_OBJ_TERMS takes them).  With tan, t_j the slopes of I and of its
 * derivative d_j in c_j over the step (the same slope serves the scale pass,
 * min_saxs.c:302-303), b = I - tan q and e_j = d_j - t_j q, the reference's
 * per-q sums (min_saxs.c:70-93, 302-310) are
 *   up    = sum (b a1 + tan a2)
 *   down  = sum (b^2 a3 + 2 tan b a4 + tan^2 a5)
 *   score = sum a0 - 2 k up + k^2 down
 *   g_j   = 2 k sum (-e_j a1 - t_j a2)
 *         + 2 k^2 sum (b e_j a3 + (I t_j + d_j tan - 2 tan t_j q) a4 + tan t_j a5)
 * with k = up / down.  None of the six sums depends on k, so one pass over
 * q forms all of them. */
#define FSX_1P_TERMS(in, d1, d2, in_prev, d1p, d2p, rq, qc, ai, tt)                   \
    do {                                                                              \
        const double tan_ = ((in) - (in_prev)) * (rq);                                \
        const double t1_ = ((d1) - (d1p)) * (rq);                                     \
        const double t2_ = ((d2) - (d2p)) * (rq);                                     \
        const double b_ = (in) - tan_ * (qc);                                         \
        const double e1_ = (d1) - t1_ * (qc), e2_ = (d2) - t2_ * (qc);                \
        (tt)[0] = b_ * (ai)[1] + tan_ * (ai)[2];                                      \
        (tt)[1] = b_ * b_ * (ai)[3] + 2.0 * tan_ * b_ * (ai)[4] + tan_ * tan_ * (ai)[5]; \
        (tt)[2] = -e1_ * (ai)[1] - t1_ * (ai)[2];                                     \
        (tt)[3] = b_ * e1_ * (ai)[3]                                                  \
                + ((in) * t1_ + (d1) * tan_ - 2.0 * tan_ * t1_ * (qc)) * (ai)[4]      \
                + tan_ * t1_ * (ai)[5];                                               \
        (tt)[4] = -e2_ * (ai)[1] - t2_ * (ai)[2];                                     \
        (tt)[5] = b_ * e2_ * (ai)[3]                                                  \
                + ((in) * t2_ + (d2) * tan_ - 2.0 * tan_ * t2_ * (qc)) * (ai)[4]      \
                + tan_ * t2_ * (ai)[5];                                               \
    } while (0)

/* the value and gradient from the six one-pass sums s and sum a0 */
FSX_HD void fsx_1p_finish(const double *s, double a0sum, double *fval, double *g)
{
    const double k = s[0] / s[1];
    *fval = a0sum + k * (-2.0 * s[0] + k * s[1]);
    g[0] = 2.0 * k * (s[2] + k * s[3]);
    g[1] = 2.0 * k * (s[4] + k * s[5]);
}

/* min_saxs.c:261-319.  bA holds the first block of the profile (q = 0 ..
 * FSX_QB-1) on entry; it is overwritten. */
FSX_HD double fsx_best_scale_blk(const fsx_ct *t, const fsx_fitpar *p, double c1, double c2,
                                 double bA[FSX_QB][6])
{
    const double *a = p->a, *q = p->q;
    const int qnum = p->qnum;
    double corr = -p->mult * (c1 * c1 - 1.0);
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double in_prev = FSX_IN(bA[0], G, c2);
    double c1c = c1 * c1 * c1;
    double q_prev = -1.0, up = 0.0, down = 0.0;
#define FSX_BS_STEP(cur, i)                                                           \
    do {                                                                              \
        double qc = q[i];                                                             \
        G = c1c * exp(corr * qc * qc);                                                \
        double in = FSX_IN(cur, G, c2);                                               \
        double ut, dt;                                                                \
        FSX_BS_TERMS(in, in_prev, qc, q_prev, a + (i) * 6, ut, dt);                   \
        up += ut;                                                                     \
        down += dt;                                                                   \
        in_prev = in;                                                                 \
        q_prev = qc;                                                                  \
    } while (0)
    FSX_QLOOP(t, qnum, bA, FSX_BS_STEP);
#undef FSX_BS_STEP
    return up / down;
}

FSX_HD double fsx_best_scale(const fsx_ct *t, const fsx_fitpar *p, double c1, double c2)
{
    double bA[FSX_QB][6];
    fsx_ct_block(t, 0, FSX_QB_MIN(p->qnum, FSX_QB), bA);
    return fsx_best_scale_blk(t, p, c1, c2, bA);
}

/* min_saxs.c:3-105.  bA holds the first block of the profile on entry; it is
 * overwritten. */
FSX_HD void fsx_objective_blk(const fsx_ct *t, const fsx_fitpar *p, double k, const double *x,
                              double *g, double *fval, double bA[FSX_QB][6])
{
    const double *a = p->a, *q = p->q, *rdq = p->rdq;
    const int qnum = p->qnum;
    double v0[6];
    for (int j = 0; j < 6; j++) v0[j] = bA[0][j];
    double c1 = x[0], c2 = x[1];
    double g0 = 0.0, g1 = 0.0, mult = p->mult;
    double corr = -mult * (c1 * c1 - 1.0);
    /* 3 / c1 and 2 c1 mult of dG/dc1, the same at every q */
    const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double Gd = G * (r3c1 - c1m2 * q[0] * q[0]);
    double in_prev = FSX_IN(v0, G, c2);
    double d1p = -Gd * v0[1] + 2.0 * G * Gd * v0[3] - Gd * c2 * v0[4];
    double d2p = v0[2] - G * v0[4] + 2.0 * c2 * v0[5];
    double q_prev = -1.0, score = 0.0, c1c = c1 * c1 * c1;
#define FSX_OBJ_STEP(cur, i)                                                          \
    do {                                                                              \
        double qc = q[i];                                                             \
        G = c1c * exp(corr * qc * qc);                                                \
        Gd = G * (r3c1 - c1m2 * qc * qc);                                             \
        double in, d1, d2, tg0, tg1, tsc;                                             \
        FSX_OBJ_VALS(cur, G, Gd, c2, in, d1, d2);                                     \
        const double rq = rdq ? rdq[i] : 1.0 / (qc - q_prev);                         \
        FSX_OBJ_TERMS(in, d1, d2, in_prev, d1p, d2p, rq, qc, a + (i) * 6, k, tg0, tg1, tsc); \
        g0 += tg0;                                                                    \
        g1 += tg1;                                                                    \
        score += tsc;                                                                 \
        in_prev = in;                                                                 \
        d1p = d1;                                                                     \
        d2p = d2;                                                                     \
        q_prev = qc;                                                                  \
    } while (0)
    FSX_QLOOP(t, qnum, bA, FSX_OBJ_STEP);
#undef FSX_OBJ_STEP
    g[0] = g0;
    g[1] = g1;
    *fval = score;
}

FSX_HD void fsx_objective(const fsx_ct *t, const fsx_fitpar *p, double k,
                          const double *x, double *g, double *fval)
{
    double bA[FSX_QB][6];
    fsx_ct_block(t, 0, FSX_QB_MIN(p->qnum, FSX_QB), bA);
    fsx_objective_blk(t, p, k, x, g, fval, bA);
}

/* The evaluation of the mixed-precision form: the scale, the objective and
 * its gradient (min_saxs.c:3-105, 261-319) from one pass over q, by the sums
 * of FSX_1P_TERMS: one read of the profile and one exp per q instead of two.
 * The values equal the two-pass ones up to rounding, so the fits end at the
 * reference's optimum within its stopping tolerance, not on its exact
 * iterates.  bA holds the first block of the profile on entry. */
FSX_HD void fsx_eval_1p_blk(const fsx_ct *t, const fsx_fitpar *p, const double *x, double *g,
                            double *fval, double bA[FSX_QB][6])
{
    const double *a = p->a, *q = p->q, *rdq = p->rdq;
    const int qnum = p->qnum;
    double v0[6];
    for (int j = 0; j < 6; j++) v0[j] = bA[0][j];
    const double c1 = x[0], c2 = x[1], mult = p->mult;
    const double corr = -mult * (c1 * c1 - 1.0);
    const double r3c1 = 3.0 / c1, c1m2 = 2.0 * c1 * mult;
    double G = c1 * c1 * c1 * exp(corr * q[0] * q[0]);
    double Gd = G * (r3c1 - c1m2 * q[0] * q[0]);
    double in_prev, d1p, d2p;
    FSX_OBJ_VALS(v0, G, Gd, c2, in_prev, d1p, d2p);
    double q_prev = -1.0, c1c = c1 * c1 * c1;
    double s[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
#define FSX_1P_STEP(cur, i)                                                           \
    do {                                                                              \
        double qc = q[i];                                                             \
        G = c1c * exp(corr * qc * qc);                                                \
        Gd = G * (r3c1 - c1m2 * qc * qc);                                             \
        double in, d1, d2, tt[6];                                                     \
        FSX_OBJ_VALS(cur, G, Gd, c2, in, d1, d2);                                     \
        const double rq = rdq ? rdq[i] : 1.0 / (qc - q_prev);                         \
        FSX_1P_TERMS(in, d1, d2, in_prev, d1p, d2p, rq, qc, a + (i) * 6, tt);         \
        for (int j_ = 0; j_ < 6; j_++) s[j_] += tt[j_];                               \
        in_prev = in;                                                                 \
        d1p = d1;                                                                     \
        d2p = d2;                                                                     \
        q_prev = qc;                                                                  \
    } while (0)
    FSX_QLOOP(t, qnum, bA, FSX_1P_STEP);
#undef FSX_1P_STEP
    fsx_1p_finish(s, p->a0sum, fval, g);
}

/* sum of a[6 i] over q, in q order (the constant part of the objective) */
FSX_HD double fsx_a0sum(const double *a, int qnum)
{
    double s = 0.0;
    for (int i = 0; i < qnum; i++) s += a[6 * i];
    return s;
}

/* ---------------------------------------------------------------------- */
/* L-BFGS-B for n = 2 variables and m = 3 correction pairs, the sizes the
 * wrapper always uses (min_saxs.c:196-259).  Every subscript into the state
 * is a compile-time constant once the fixed-bound loops below are unrolled,
 * so a GPU thread keeps the whole state in registers:
 *  - the correction pairs S (ws) and Y (wy) are stored oldest first: column
 *    j is the reference's ring-buffer column ((head + j - 2) mod m) + 1, so
 *    its pointer walks (pointr, ipntr, jpntr) become plain column loops;
 *  - the 2col-vectors of bmv, cauchy, cmprlb and subsm are held as their two
 *    col-blocks (elements 1..col and col+1..2col), except c, which cauchy
 *    leaves for cmprlb and which keeps the reference's flat layout;
 *  - the 2col x 2col middle matrix K (wn) is held as its blocks K11, K12,
 *    K22 (wn11, wn12, wn22);
 *  - loops bounded by col, nsub, nenter, ileave or nbreak run to their
 *    fixed maximum under a guard;
 *  - data-dependent subscripts into 2-vectors (ind[k], ibp, ...) are
 *    two-way selects.
 * Values, operations and their order are those of the reference routines,
 * so the iterates are bit-identical to the reference's. */
#ifdef __HIPCC__
#define LB_UNROLL _Pragma("unroll")
#else
#define LB_UNROLL
#endif

/* Element i (1-based, i in {1, 2}) of a 2-vector.  Both elements are read
 * (written) every time and the value is selected, so that no address
 * depends on data: the optimiser then keeps such arrays in registers. */
#define LB_AT(a, i)                                                   \
    ({                                                                \
        __typeof__((a)[0]) a0_ = (a)[0], a1_ = (a)[1];                \
        (i) == 1 ? a0_ : a1_;                                         \
    })
#define LB_PUT(a, i, v)                                               \
    do {                                                              \
        __typeof__((a)[0]) v_ = (v), a0_ = (a)[0], a1_ = (a)[1];      \
        const int one_ = (i) == 1;                                    \
        (a)[0] = one_ ? v_ : a0_;                                     \
        (a)[1] = one_ ? a1_ : v_;                                     \
    } while (0)

typedef struct fsx_lb {
    double ws[FSX_LB_N][FSX_LB_M], wy[FSX_LB_N][FSX_LB_M];   /* [variable][pair] */
    double sy[FSX_LB_M][FSX_LB_M], ss[FSX_LB_M][FSX_LB_M];
    /* the blocks K11, K12, K22 of the middle matrix K (wn), WN1 and the
     * factor of T (wt), in that order (FSX_LB_KN values): in the thread's own
     * array, or, where
     * FSX_LB_KSHARED is defined (the GPU fit kernels), at kb in LDS with a
     * stride of FSX_LB_KSTRIDE lanes, which keeps them out of the registers
     * that the rest of the update needs */
#ifdef FSX_LB_KSHARED
    double *kb;
#else
    double kb[4 * FSX_LB_M * FSX_LB_M + 4 * FSX_LB_M * FSX_LB_M];
#endif
    double wc[2 * FSX_LB_M];     /* c of cauchy, flat: c(1..col), c(col+1..2col) */
    double z[FSX_LB_N], r[FSX_LB_N], d[FSX_LB_N], t[FSX_LB_N];
    int index[FSX_LB_N], iwhere[FSX_LB_N], indx2[FSX_LB_N];
} fsx_lb;

/* 1-based accessors; LB_WSK / LB_WYK take a variable number k in {1, 2}
 * that may be data dependent */
#define LB_WS(i, j) w->ws[(i) - 1][(j) - 1]
#define LB_WY(i, j) w->wy[(i) - 1][(j) - 1]
#define LB_WSK(k, j)                                                  \
    ({                                                                \
        double s0_ = w->ws[0][(j) - 1], s1_ = w->ws[1][(j) - 1];      \
        (k) == 1 ? s0_ : s1_;                                         \
    })
#define LB_WYK(k, j)                                                  \
    ({                                                                \
        double y0_ = w->wy[0][(j) - 1], y1_ = w->wy[1][(j) - 1];      \
        (k) == 1 ? y0_ : y1_;                                         \
    })
#define LB_SY(i, j) w->sy[(i) - 1][(j) - 1]
#define LB_SS(i, j) w->ss[(i) - 1][(j) - 1]
#define LB_WT(i, j) LB_WT0((i) - 1, (j) - 1)
#define FSX_LB_KN (4 * FSX_LB_M * FSX_LB_M + 4 * FSX_LB_M * FSX_LB_M)
#ifdef FSX_LB_KSHARED
#define LB_KV(w, o) ((w)->kb[(long)(o) * FSX_LB_KSTRIDE])
#define FSX_LB_KARG , double *kbuf
#define FSX_LB_KSET(s) ((s).w.kb = kbuf)
#else
#define LB_KV(w, o) ((w)->kb[o])
#define FSX_LB_KARG
#define FSX_LB_KSET(s) ((void)0)
#endif
/* 0-based K blocks, 1-based WN1 */
#define LB_K11(a, b) LB_KV(w, (a) * FSX_LB_M + (b))
#define LB_K12(a, b) LB_KV(w, FSX_LB_M * FSX_LB_M + (a) * FSX_LB_M + (b))
#define LB_K22(a, b) LB_KV(w, 2 * FSX_LB_M * FSX_LB_M + (a) * FSX_LB_M + (b))
#define LB_WN1(i, j) LB_KV(w, 3 * FSX_LB_M * FSX_LB_M + ((i) - 1) * 2 * FSX_LB_M + ((j) - 1))
#define LB_WT0(a, b) LB_KV(w, 7 * FSX_LB_M * FSX_LB_M + (a) * FSX_LB_M + (b))

FSX_HD double lb_max(double a, double b) { return a >= b ? a : b; }
FSX_HD double lb_min(double a, double b) { return a <= b ? a : b; }
FSX_HD double lb_abs(double a) { return a >= 0 ? a : -a; }

/* ddot (miniCBLAS.c) is a left-to-right sum from 0.0; daxpy returns at once
 * when the factor is 0.  Both are written out in place below, apart from
 * the n-vector dot product. */
FSX_HD double lb_dot2(const double x[FSX_LB_N], const double y[FSX_LB_N])
{
    double s = 0.;
    s = s + x[0] * y[0];
    s = s + x[1] * y[1];
    return s;
}

/* linpack.c dpofa: upper Cholesky factor of the leading n x n block of a
 * (a[i-1][j-1] = A(i, j)), n <= m.  0, or the column that failed. */
FSX_HD int lb_chol(double a[FSX_LB_M][FSX_LB_M], int n)
{
    LB_UNROLL
    for (int j = 1; j <= FSX_LB_M; ++j) {
        if (j <= n) {
            double s = 0.;
            LB_UNROLL
            for (int k = 1; k <= j - 1; ++k) {
                double dt = 0.;
                LB_UNROLL
                for (int i = 1; i <= k - 1; ++i) dt = dt + a[i - 1][k - 1] * a[i - 1][j - 1];
                double t = a[k - 1][j - 1] - dt;
                t /= a[k - 1][k - 1];
                a[k - 1][j - 1] = t;
                s += t * t;
            }
            s = a[j - 1][j - 1] - s;
            if (s <= 0.) return j;
            a[j - 1][j - 1] = sqrt(s);
        }
    }
    return 0;
}

/* linpack.c dtrsl with T the leading n x n upper triangle of t: job 11
 * solves T' x = b, job 01 solves T x = b, in place.  0, or the first zero
 * diagonal. */
FSX_HD int lb_trsl(double t[FSX_LB_M][FSX_LB_M], int n, double b[FSX_LB_M], int job)
{
    LB_UNROLL
    for (int i = 1; i <= FSX_LB_M; ++i)
        if (i <= n && t[i - 1][i - 1] == 0.) return i;
    if (job == 1) {
        LB_UNROLL
        for (int c = 1; c <= FSX_LB_M; ++c)
            if (c == n) b[c - 1] /= t[c - 1][c - 1];
        /* columns n .. 2: b(1..j) += -b(j+1) T(1..j, j+1), b(j) /= T(j, j) */
        LB_UNROLL
        for (int j = FSX_LB_M - 1; j >= 1; --j) {
            if (j <= n - 1) {
                double tmp = -b[j];
                if (tmp != 0.) {
                    LB_UNROLL
                    for (int i = 0; i < j; i++) b[i] += tmp * t[i][j];
                }
                b[j - 1] /= t[j - 1][j - 1];
            }
        }
    } else {
        b[0] /= t[0][0];
        LB_UNROLL
        for (int j = 2; j <= FSX_LB_M; ++j) {
            if (j <= n) {
                double dt = 0.;
                LB_UNROLL
                for (int i = 0; i < j - 1; i++) dt = dt + t[i][j - 1] * b[i];
                b[j - 1] -= dt;
                b[j - 1] /= t[j - 1][j - 1];
            }
        }
    }
    return 0;
}

/* dtrsl on the 2col x 2col upper triangle K = [K11 K12; 0 K22] held as
 * its blocks, b = (b1, b2): job 11 solves K' x = b, job 01 solves K x = b.
 * The element sequence of every sum and update is that of dtrsl on the
 * assembled matrix. */
FSX_HD int lb_trsl_k(const fsx_lb *w, int col, double b1[FSX_LB_M], double b2[FSX_LB_M],
                     int job)
{
    const int m = FSX_LB_M;
    LB_UNROLL
    for (int i = 1; i <= m; ++i)
        if (i <= col && LB_K11(i - 1, i - 1) == 0.) return i;
    LB_UNROLL
    for (int i = 1; i <= m; ++i)
        if (i <= col && LB_K22(i - 1, i - 1) == 0.) return col + i;
    if (job == 1) {
        LB_UNROLL
        for (int c = 1; c <= m; ++c)
            if (c == col) b2[c - 1] /= LB_K22(c - 1, c - 1);
        /* columns 2col .. col+1 */
        LB_UNROLL
        for (int cp = m; cp >= 1; --cp) {
            if (cp <= col) {
                double tmp = -b2[cp - 1];
                if (tmp != 0.) {
                    LB_UNROLL
                    for (int i = 0; i < m; i++)
                        if (i < col) b1[i] += tmp * LB_K12(i, cp - 1);
                    LB_UNROLL
                    for (int i = 0; i < cp - 1; i++) b2[i] += tmp * LB_K22(i, cp - 1);
                }
                if (cp >= 2) {
                    b2[cp - 2] /= LB_K22(cp - 2, cp - 2);
                } else {
                    LB_UNROLL
                    for (int c = 1; c <= m; ++c)
                        if (c == col) b1[c - 1] /= LB_K11(c - 1, c - 1);
                }
            }
        }
        /* columns col .. 2 */
        LB_UNROLL
        for (int c = m; c >= 2; --c) {
            if (c <= col) {
                double tmp = -b1[c - 1];
                if (tmp != 0.) {
                    LB_UNROLL
                    for (int i = 0; i < c - 1; i++) b1[i] += tmp * LB_K11(i, c - 1);
                }
                b1[c - 2] /= LB_K11(c - 2, c - 2);
            }
        }
    } else {
        b1[0] /= LB_K11(0, 0);
        LB_UNROLL
        for (int j = 2; j <= m; ++j) {
            if (j <= col) {
                double dt = 0.;
                LB_UNROLL
                for (int i = 0; i < j - 1; i++) dt = dt + LB_K11(i, j - 1) * b1[i];
                b1[j - 1] -= dt;
                b1[j - 1] /= LB_K11(j - 1, j - 1);
            }
        }
        LB_UNROLL
        for (int jp = 1; jp <= m; ++jp) {
            if (jp <= col) {
                double dt = 0.;
                LB_UNROLL
                for (int i = 0; i < m; i++)
                    if (i < col) dt = dt + LB_K12(i, jp - 1) * b1[i];
                LB_UNROLL
                for (int i = 0; i < jp - 1; i++) dt = dt + LB_K22(i, jp - 1) * b2[i];
                b2[jp - 1] -= dt;
                b2[jp - 1] /= LB_K22(jp - 1, jp - 1);
            }
        }
    }
    return 0;
}

/* subalgorithms.c bmv: (p1, p2) = M (v1, v2), M the 2col x 2col middle
 * matrix of the compact L-BFGS representation */
FSX_HD int lb_bmv(fsx_lb *w, int col, const double v1[FSX_LB_M], const double v2[FSX_LB_M],
                  double p1[FSX_LB_M], double p2[FSX_LB_M])
{
    if (col == 0) return 0;
    p2[0] = v2[0];
    LB_UNROLL
    for (int i = 2; i <= FSX_LB_M; ++i) {
        if (i <= col) {
            double sum = 0.;
            LB_UNROLL
            for (int k = 1; k <= i - 1; ++k) sum += LB_SY(i, k) * v1[k - 1] / LB_SY(k, k);
            p2[i - 1] = v2[i - 1] + sum;
        }
    }
    double wt[FSX_LB_M][FSX_LB_M];   /* the factor of T, read once */
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) wt[i][j] = LB_WT0(i, j);
    int info = lb_trsl(wt, col, p2, 11);
    if (info) return info;
    LB_UNROLL
    for (int i = 1; i <= FSX_LB_M; ++i)
        if (i <= col) p1[i - 1] = v1[i - 1] / sqrt(LB_SY(i, i));
    info = lb_trsl(wt, col, p2, 1);
    if (info) return info;
    LB_UNROLL
    for (int i = 1; i <= FSX_LB_M; ++i)
        if (i <= col) p1[i - 1] = -p1[i - 1] / sqrt(LB_SY(i, i));
    LB_UNROLL
    for (int i = 1; i <= FSX_LB_M; ++i) {
        if (i <= col) {
            double sum = 0.;
            LB_UNROLL
            for (int k = i + 1; k <= FSX_LB_M; ++k)
                if (k <= col) sum += LB_SY(k, i) * p2[k - 1] / LB_SY(i, i);
            p1[i - 1] += sum;
        }
    }
    return 0;
}

/* element j (0-based, < col) of the second block of the flat c */
FSX_HD double lb_c2(const fsx_lb *w, int col, int j)
{
    double e1 = w->wc[1 + j], e2 = w->wc[2 + j], e3 = w->wc[3 + j];
    return col == 1 ? e1 : col == 2 ? e2 : e3;
}
/* c2[j] += a * x */
FSX_HD void lb_c2_axpy(fsx_lb *w, int col, int j, double a, double x)
{
    double e1 = w->wc[1 + j], e2 = w->wc[2 + j], e3 = w->wc[3 + j];
    w->wc[1 + j] = col == 1 ? e1 + a * x : e1;
    w->wc[2 + j] = col == 2 ? e2 + a * x : e2;
    w->wc[3 + j] = col == 3 ? e3 + a * x : e3;
}

/* subalgorithms.c cauchy, every variable bounded on both sides (nbd = 2):
 * the generalized Cauchy point into w->z and c = W'(xcp - x) into w->wc.
 * With n = 2 there are at most two breakpoints; the heap of the second one
 * (hpsolb on one element) leaves it in place. */
FSX_HD int lb_cauchy(fsx_lb *w, const double x[FSX_LB_N], const double l[FSX_LB_N],
                     const double u[FSX_LB_N], const double g[FSX_LB_N], double theta,
                     int col, double sbgnrm, int *nseg)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    int *iorder = w->indx2, *iwhere = w->iwhere;
    double *t = w->t, *d = w->d, *xcp = w->z;
    double p1[FSX_LB_M], p2[FSX_LB_M], v1[FSX_LB_M], v2[FSX_LB_M];
    double wbp1[FSX_LB_M], wbp2[FSX_LB_M];
    const double epsmch = DBL_EPSILON;
    if (sbgnrm <= 0.) {
        xcp[0] = x[0];
        xcp[1] = x[1];
        return 0;
    }
    int bnded = 1, nfree = n + 1, nbreak = 0, ibkmin = 0;
    double bkmin = 0., f1 = 0., tl = 0., tu = 0.;
    LB_UNROLL
    for (int j = 0; j < m; j++) {
        p1[j] = 0.; p2[j] = 0.; v1[j] = 0.; v2[j] = 0.; wbp1[j] = 0.; wbp2[j] = 0.;
    }
    LB_UNROLL
    for (int i = 1; i <= n; ++i) {
        double neggi = -g[i - 1];
        if (iwhere[i - 1] != 3 && iwhere[i - 1] != -1) {
            tl = x[i - 1] - l[i - 1];
            tu = u[i - 1] - x[i - 1];
            int xlower = tl <= 0.;
            int xupper = tu <= 0.;
            iwhere[i - 1] = 0;
            if (xlower) {
                if (neggi <= 0.) iwhere[i - 1] = 1;
            } else if (xupper) {
                if (neggi >= 0.) iwhere[i - 1] = 2;
            } else {
                if (lb_abs(neggi) <= 0.) iwhere[i - 1] = -3;
            }
        }
        if (iwhere[i - 1] != 0 && iwhere[i - 1] != -1) {
            d[i - 1] = 0.;
        } else {
            d[i - 1] = neggi;
            f1 -= neggi * neggi;
            LB_UNROLL
            for (int j = 1; j <= m; ++j) {
                if (j <= col) {
                    p1[j - 1] += LB_WY(i, j) * neggi;
                    p2[j - 1] += LB_WS(i, j) * neggi;
                }
            }
            if (neggi < 0.) {
                ++nbreak;
                LB_PUT(iorder, nbreak, i);
                LB_PUT(t, nbreak, tl / (-neggi));
                if (nbreak == 1 || LB_AT(t, nbreak) < bkmin) {
                    bkmin = LB_AT(t, nbreak);
                    ibkmin = nbreak;
                }
            } else if (neggi > 0.) {
                ++nbreak;
                LB_PUT(iorder, nbreak, i);
                LB_PUT(t, nbreak, tu / neggi);
                if (nbreak == 1 || LB_AT(t, nbreak) < bkmin) {
                    bkmin = LB_AT(t, nbreak);
                    ibkmin = nbreak;
                }
            } else {
                --nfree;
                LB_PUT(iorder, nfree, i);
                if (lb_abs(neggi) > 0.) bnded = 0;
            }
        }
    }
    if (theta != 1.) {
        LB_UNROLL
        for (int j = 0; j < m; j++)
            if (j < col) p2[j] = theta * p2[j];
    }
    xcp[0] = x[0];
    xcp[1] = x[1];
    if (nbreak == 0 && nfree == n + 1) return 0;
    LB_UNROLL
    for (int j = 0; j < 2 * m; j++)
        if (j < 2 * col) w->wc[j] = 0.;
    double f2 = -theta * f1;
    double f2_org = f2;
    if (col > 0) {
        int info = lb_bmv(w, col, p1, p2, v1, v2);
        if (info) return info;
        double dt = 0.;
        LB_UNROLL
        for (int j = 0; j < m; j++)
            if (j < col) dt = dt + v1[j] * p1[j];
        LB_UNROLL
        for (int j = 0; j < m; j++)
            if (j < col) dt = dt + v2[j] * p2[j];
        f2 -= dt;
    }
    double dtm = -f1 / f2;
    double tsum = 0.;
    *nseg = 1;
    int reached_end = 0;
    if (nbreak != 0) {
        int nleft = nbreak, ibp = 0;
        double tj = 0.;
        LB_UNROLL
        for (int iter = 1; iter <= n; ++iter) {
            double tj0 = tj;
            if (iter == 1) {
                tj = bkmin;
                ibp = LB_AT(iorder, ibkmin);
            } else {
                if (ibkmin != nbreak) {
                    LB_PUT(t, ibkmin, LB_AT(t, nbreak));
                    LB_PUT(iorder, ibkmin, LB_AT(iorder, nbreak));
                }
                tj = LB_AT(t, nleft);
                ibp = LB_AT(iorder, nleft);
            }
            double dt = tj - tj0;
            if (dtm < dt) break;
            tsum += dt;
            --nleft;
            double dibp = LB_AT(d, ibp);
            LB_PUT(d, ibp, 0.);
            double zibp;
            if (dibp > 0.) {
                zibp = LB_AT(u, ibp) - LB_AT(x, ibp);
                LB_PUT(xcp, ibp, LB_AT(u, ibp));
                LB_PUT(iwhere, ibp, 2);
            } else {
                zibp = LB_AT(l, ibp) - LB_AT(x, ibp);
                LB_PUT(xcp, ibp, LB_AT(l, ibp));
                LB_PUT(iwhere, ibp, 1);
            }
            if (nleft == 0 && nbreak == n) {
                dtm = dt;
                reached_end = 1;
                break;
            }
            ++(*nseg);
            double dibp2 = dibp * dibp;
            f1 = f1 + dt * f2 + dibp2 - theta * dibp * zibp;
            f2 -= theta * dibp2;
            if (col > 0) {
                if (dt != 0.) {
                    LB_UNROLL
                    for (int j = 0; j < m; j++)
                        if (j < col) w->wc[j] += dt * p1[j];
                    LB_UNROLL
                    for (int j = 0; j < m; j++)
                        if (j < col) lb_c2_axpy(w, col, j, dt, p2[j]);
                }
                LB_UNROLL
                for (int j = 1; j <= m; ++j) {
                    if (j <= col) {
                        wbp1[j - 1] = LB_WYK(ibp, j);
                        wbp2[j - 1] = theta * LB_WSK(ibp, j);
                    }
                }
                int info = lb_bmv(w, col, wbp1, wbp2, v1, v2);
                if (info) return info;
                double wmc = 0., wmp = 0., wmw = 0.;
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmc = wmc + w->wc[j] * v1[j];
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmc = wmc + lb_c2(w, col, j) * v2[j];
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmp = wmp + p1[j] * v1[j];
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmp = wmp + p2[j] * v2[j];
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmw = wmw + wbp1[j] * v1[j];
                LB_UNROLL
                for (int j = 0; j < m; j++)
                    if (j < col) wmw = wmw + wbp2[j] * v2[j];
                const double a = -dibp;
                if (a != 0.) {
                    LB_UNROLL
                    for (int j = 0; j < m; j++)
                        if (j < col) { p1[j] += a * wbp1[j]; p2[j] += a * wbp2[j]; }
                }
                f1 += dibp * wmc;
                f2 = f2 + dibp * 2. * wmp - dibp2 * wmw;
            }
            f2 = lb_max(epsmch * f2_org, f2);
            if (nleft > 0) {
                dtm = -f1 / f2;
                continue;
            } else if (bnded) {
                f1 = 0.;
                f2 = 0.;
                dtm = 0.;
            } else {
                dtm = -f1 / f2;
            }
            break;
        }
    }
    if (!reached_end) {
        if (dtm <= 0.) dtm = 0.;
        tsum += dtm;
        if (tsum != 0.) {
            xcp[0] += tsum * d[0];
            xcp[1] += tsum * d[1];
        }
    }
    if (col > 0 && dtm != 0.) {
        LB_UNROLL
        for (int j = 0; j < m; j++)
            if (j < col) w->wc[j] += dtm * p1[j];
        LB_UNROLL
        for (int j = 0; j < m; j++)
            if (j < col) lb_c2_axpy(w, col, j, dtm, p2[j]);
    }
    return 0;
}

/* subalgorithms.c freev */
FSX_HD void lb_freev(fsx_lb *w, int *nfree, int *nenter, int *ileave, int *wrk,
                     int updatd, int iter)
{
    const int n = FSX_LB_N;
    int *index = w->index, *indx2 = w->indx2, *iwhere = w->iwhere;
    *nenter = 0;
    *ileave = n + 1;
    if (iter > 0) {
        LB_UNROLL
        for (int i = 1; i <= n; ++i) {
            if (i <= *nfree) {
                int k = index[i - 1];
                if (LB_AT(iwhere, k) > 0) {
                    --(*ileave);
                    LB_PUT(indx2, *ileave, k);
                }
            }
        }
        LB_UNROLL
        for (int i = 1; i <= n; ++i) {
            if (i >= *nfree + 1) {
                int k = index[i - 1];
                if (LB_AT(iwhere, k) <= 0) {
                    ++(*nenter);
                    LB_PUT(indx2, *nenter, k);
                }
            }
        }
    }
    *wrk = *ileave < n + 1 || *nenter > 0 || updatd;
    *nfree = 0;
    int iact = n + 1;
    LB_UNROLL
    for (int i = 1; i <= n; ++i) {
        if (iwhere[i - 1] <= 0) {
            ++(*nfree);
            LB_PUT(index, *nfree, i);
        } else {
            --iact;
            LB_PUT(index, iact, i);
        }
    }
}

/* subalgorithms.c formk: update WN1 (the inner products of the pairs over
 * the free and the active variables) and factor K = [K11 K12; 0 K22]. */
FSX_HD int lb_formk(fsx_lb *w, int nsub, int nenter, int ileave, int iupdat, int updatd,
                    double theta, int col)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    const int *ind = w->index, *indx2 = w->indx2;
    int upcl;
    if (updatd) {
        if (iupdat > m) {
            LB_UNROLL
            for (int jy = 1; jy <= m - 1; ++jy) {
                const int js = m + jy;
                LB_UNROLL
                for (int k = 0; k < m - jy; k++) LB_WN1(jy + k, jy) = LB_WN1(jy + 1 + k, jy + 1);
                LB_UNROLL
                for (int k = 0; k < m - jy; k++) LB_WN1(js + k, js) = LB_WN1(js + 1 + k, js + 1);
                LB_UNROLL
                for (int k = 0; k < m - 1; k++) LB_WN1(m + 1 + k, jy) = LB_WN1(m + 2 + k, jy + 1);
            }
        }
        /* the row and column of the newest pair (column col) */
        LB_UNROLL
        for (int cc = 1; cc <= m; ++cc) {
            if (cc == col) {
                const int iy = cc, is = m + cc;
                LB_UNROLL
                for (int jy = 1; jy <= cc; ++jy) {
                    const int js = m + jy;
                    double t1 = 0., t2 = 0., t3 = 0.;
                    LB_UNROLL
                    for (int k = 1; k <= n; ++k) {
                        if (k <= nsub) {
                            int k1 = ind[k - 1];
                            t1 += LB_WYK(k1, cc) * LB_WYK(k1, jy);
                        }
                    }
                    LB_UNROLL
                    for (int k = 1; k <= n; ++k) {
                        if (k >= nsub + 1) {
                            int k1 = ind[k - 1];
                            t2 += LB_WSK(k1, cc) * LB_WSK(k1, jy);
                            t3 += LB_WSK(k1, cc) * LB_WYK(k1, jy);
                        }
                    }
                    LB_WN1(iy, jy) = t1;
                    LB_WN1(is, js) = t2;
                    LB_WN1(is, jy) = t3;
                }
                LB_UNROLL
                for (int i = 1; i <= cc; ++i) {
                    const int is2 = m + i;
                    double t3 = 0.;
                    LB_UNROLL
                    for (int k = 1; k <= n; ++k) {
                        if (k <= nsub) {
                            int k1 = ind[k - 1];
                            t3 += LB_WSK(k1, i) * LB_WYK(k1, cc);
                        }
                    }
                    LB_WN1(is2, cc) = t3;
                }
            }
        }
        upcl = col - 1;
    } else {
        upcl = col;
    }
    /* the older pairs: variables that entered or left the free set */
    LB_UNROLL
    for (int iy = 1; iy <= m; ++iy) {
        if (iy <= upcl) {
            const int is = m + iy;
            LB_UNROLL
            for (int jy = 1; jy <= iy; ++jy) {
                const int js = m + jy;
                double t1 = 0., t2 = 0., t3 = 0., t4 = 0.;
                LB_UNROLL
                for (int k = 1; k <= n; ++k) {
                    if (k <= nenter) {
                        int k1 = indx2[k - 1];
                        t1 += LB_WYK(k1, iy) * LB_WYK(k1, jy);
                        t2 += LB_WSK(k1, iy) * LB_WSK(k1, jy);
                    }
                }
                LB_UNROLL
                for (int k = 1; k <= n; ++k) {
                    if (k >= ileave) {
                        int k1 = indx2[k - 1];
                        t3 += LB_WYK(k1, iy) * LB_WYK(k1, jy);
                        t4 += LB_WSK(k1, iy) * LB_WSK(k1, jy);
                    }
                }
                LB_WN1(iy, jy) = LB_WN1(iy, jy) + t1 - t3;
                LB_WN1(is, js) = LB_WN1(is, js) - t2 + t4;
            }
        }
    }
    LB_UNROLL
    for (int is = m + 1; is <= 2 * m; ++is) {
        if (is <= m + upcl) {
            LB_UNROLL
            for (int jy = 1; jy <= m; ++jy) {
                if (jy <= upcl) {
                    double t1 = 0., t3 = 0.;
                    LB_UNROLL
                    for (int k = 1; k <= n; ++k) {
                        if (k <= nenter) {
                            int k1 = indx2[k - 1];
                            t1 += LB_WSK(k1, is - m) * LB_WYK(k1, jy);
                        }
                    }
                    LB_UNROLL
                    for (int k = 1; k <= n; ++k) {
                        if (k >= ileave) {
                            int k1 = indx2[k - 1];
                            t3 += LB_WSK(k1, is - m) * LB_WYK(k1, jy);
                        }
                    }
                    if (is <= jy + m) LB_WN1(is, jy) = LB_WN1(is, jy) + t1 - t3;
                    else LB_WN1(is, jy) = LB_WN1(is, jy) - t1 + t3;
                }
            }
        }
    }
    /* K11 = Y'Y / theta + D (upper), K22 = theta S'S (upper), K12 = -L_a' + R_z */
    LB_UNROLL
    for (int iy = 1; iy <= m; ++iy) {
        if (iy <= col) {
            const int is1 = m + iy;
            LB_UNROLL
            for (int jy = 1; jy <= iy; ++jy) {
                const int js1 = m + jy;
                LB_K11(jy - 1, iy - 1) = LB_WN1(iy, jy) / theta;
                LB_K22(jy - 1, iy - 1) = LB_WN1(is1, js1) * theta;
            }
            LB_UNROLL
            for (int jy = 1; jy <= iy - 1; ++jy) LB_K12(jy - 1, iy - 1) = -LB_WN1(is1, jy);
            LB_UNROLL
            for (int jy = iy; jy <= m; ++jy)
                if (jy <= col) LB_K12(jy - 1, iy - 1) = LB_WN1(is1, jy);
            LB_K11(iy - 1, iy - 1) += LB_SY(iy, iy);
        }
    }
    /* the Cholesky factor of K11 on a local copy, written back */
    double k11[FSX_LB_M][FSX_LB_M];
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) k11[i][j] = LB_K11(i, j);
    const int c11 = lb_chol(k11, col);
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) LB_K11(i, j) = k11[i][j];
    if (c11) return -1;
    /* K12 <- L11^{-1} K12, column by column (dtrsl job 11) */
    LB_UNROLL
    for (int js = 1; js <= m; ++js) {
        if (js <= col) {
            double b[FSX_LB_M];
            LB_UNROLL
            for (int i = 0; i < m; i++) b[i] = LB_K12(i, js - 1);
            lb_trsl(k11, col, b, 11);
            LB_UNROLL
            for (int i = 0; i < m; i++) LB_K12(i, js - 1) = b[i];
        }
    }
    LB_UNROLL
    for (int is = 1; is <= m; ++is) {
        if (is <= col) {
            LB_UNROLL
            for (int js = is; js <= m; ++js) {
                if (js <= col) {
                    double dt = 0.;
                    LB_UNROLL
                    for (int i = 0; i < m; i++)
                        if (i < col) dt = dt + LB_K12(i, is - 1) * LB_K12(i, js - 1);
                    LB_K22(is - 1, js - 1) += dt;
                }
            }
        }
    }
    double k22[FSX_LB_M][FSX_LB_M];
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) k22[i][j] = LB_K22(i, j);
    const int c22 = lb_chol(k22, col);
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) LB_K22(i, j) = k22[i][j];
    if (c22) return -2;
    return 0;
}

/* subalgorithms.c formt: T = theta S'S + L D^{-1} L', factored (upper) */
FSX_HD int lb_formt(fsx_lb *w, int col, double theta)
{
    LB_UNROLL
    for (int j = 1; j <= FSX_LB_M; ++j)
        if (j <= col) LB_WT(1, j) = theta * LB_SS(1, j);
    LB_UNROLL
    for (int i = 2; i <= FSX_LB_M; ++i) {
        if (i <= col) {
            LB_UNROLL
            for (int j = i; j <= FSX_LB_M; ++j) {
                if (j <= col) {
                    const int k1 = i - 1;   /* min(i, j) - 1 */
                    double dd = 0.;
                    LB_UNROLL
                    for (int k = 1; k <= k1; ++k) dd += LB_SY(i, k) * LB_SY(j, k) / LB_SY(k, k);
                    LB_WT(i, j) = dd + theta * LB_SS(i, j);
                }
            }
        }
    }
    double wt[FSX_LB_M][FSX_LB_M];   /* factored on a local copy, written back */
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) wt[i][j] = LB_WT0(i, j);
    const int ct = lb_chol(wt, col);
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++)
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) LB_WT0(i, j) = wt[i][j];
    return ct ? -3 : 0;
}

/* subalgorithms.c cmprlb (constrained problem): reduced gradient of the
 * quadratic model at the Cauchy point, over the free variables */
FSX_HD int lb_cmprlb(fsx_lb *w, const double x[FSX_LB_N], const double g[FSX_LB_N],
                     double theta, int col, int nfree)
{
    const int m = FSX_LB_M;
    const double *z = w->z;
    double *r = w->r;
    const int *index = w->index;
    LB_UNROLL
    for (int i = 1; i <= FSX_LB_N; ++i) {
        if (i <= nfree) {
            int k = index[i - 1];
            r[i - 1] = -theta * (LB_AT(z, k) - LB_AT(x, k)) - LB_AT(g, k);
        }
    }
    double c1[FSX_LB_M], c2[FSX_LB_M], o1[FSX_LB_M], o2[FSX_LB_M];
    LB_UNROLL
    for (int j = 0; j < m; j++) {
        c1[j] = w->wc[j];
        c2[j] = lb_c2(w, col, j);
        o1[j] = 0.;
        o2[j] = 0.;
    }
    if (lb_bmv(w, col, c1, c2, o1, o2)) return -8;
    LB_UNROLL
    for (int j = 1; j <= m; ++j) {
        if (j <= col) {
            double a1 = o1[j - 1], a2 = theta * o2[j - 1];
            LB_UNROLL
            for (int i = 1; i <= FSX_LB_N; ++i) {
                if (i <= nfree) {
                    int k = index[i - 1];
                    r[i - 1] = r[i - 1] + LB_WYK(k, j) * a1 + LB_WSK(k, j) * a2;
                }
            }
        }
    }
    return 0;
}

/* subalgorithms.c subsm (3.0, with the projected backtracking safeguard):
 * subspace minimisation over the free variables, result in w->z */
FSX_HD int lb_subsm(fsx_lb *w, int nsub, const double l[FSX_LB_N], const double u[FSX_LB_N],
                    const double xx[FSX_LB_N], const double gg[FSX_LB_N], double theta,
                    int col)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    const int *ind = w->index;
    double *x = w->z, *d = w->r;
    double wv1[FSX_LB_M], wv2[FSX_LB_M], xp[FSX_LB_N];
    if (nsub <= 0) return 0;
    LB_UNROLL
    for (int i = 1; i <= m; ++i) {
        wv1[i - 1] = 0.;
        wv2[i - 1] = 0.;
        if (i <= col) {
            double t1 = 0., t2 = 0.;
            LB_UNROLL
            for (int j = 1; j <= n; ++j) {
                if (j <= nsub) {
                    int k = ind[j - 1];
                    t1 += LB_WYK(k, i) * d[j - 1];
                    t2 += LB_WSK(k, i) * d[j - 1];
                }
            }
            wv1[i - 1] = t1;
            wv2[i - 1] = theta * t2;
        }
    }
    int info = lb_trsl_k(w, col, wv1, wv2, 11);
    if (info) return info;
    LB_UNROLL
    for (int i = 0; i < m; ++i)
        if (i < col) wv1[i] = -wv1[i];
    info = lb_trsl_k(w, col, wv1, wv2, 1);
    if (info) return info;
    LB_UNROLL
    for (int jy = 1; jy <= m; ++jy) {
        if (jy <= col) {
            LB_UNROLL
            for (int i = 1; i <= n; ++i) {
                if (i <= nsub) {
                    int k = ind[i - 1];
                    d[i - 1] = d[i - 1] + LB_WYK(k, jy) * wv1[jy - 1] / theta
                             + LB_WSK(k, jy) * wv2[jy - 1];
                }
            }
        }
    }
    double rt = 1. / theta;
    LB_UNROLL
    for (int i = 0; i < n; i++)
        if (i < nsub) w->r[i] = rt * w->r[i];
    int iword = 0;
    xp[0] = w->z[0];
    xp[1] = w->z[1];
    LB_UNROLL
    for (int i = 1; i <= n; ++i) {
        if (i <= nsub) {
            int k = ind[i - 1];
            double dk = d[i - 1], xk = LB_AT(x, k);
            xk = lb_max(LB_AT(l, k), xk + dk);
            LB_PUT(x, k, lb_min(LB_AT(u, k), xk));
            if (LB_AT(x, k) == LB_AT(l, k) || LB_AT(x, k) == LB_AT(u, k)) iword = 1;
        }
    }
    if (iword == 0) return 0;
    double dd_p = 0.;
    LB_UNROLL
    for (int i = 1; i <= n; ++i) dd_p += (x[i - 1] - xx[i - 1]) * gg[i - 1];
    if (dd_p > 0.) {
        w->z[0] = xp[0];
        w->z[1] = xp[1];
    } else {
        return 0;
    }
    double alpha = 1., t1 = alpha;
    int ibd = 0;
    LB_UNROLL
    for (int i = 1; i <= n; ++i) {
        if (i <= nsub) {
            int k = ind[i - 1];
            double dk = d[i - 1];
            if (dk < 0.) {
                double t2 = LB_AT(l, k) - LB_AT(x, k);
                if (t2 >= 0.) t1 = 0.;
                else if (dk * alpha < t2) t1 = t2 / dk;
            } else if (dk > 0.) {
                double t2 = LB_AT(u, k) - LB_AT(x, k);
                if (t2 <= 0.) t1 = 0.;
                else if (dk * alpha > t2) t1 = t2 / dk;
            }
            if (t1 < alpha) {
                alpha = t1;
                ibd = i;
            }
        }
    }
    if (alpha < 1.) {
        double dk = LB_AT(d, ibd);
        int k = LB_AT(ind, ibd);
        if (dk > 0.) {
            LB_PUT(x, k, LB_AT(u, k));
            LB_PUT(d, ibd, 0.);
        } else if (dk < 0.) {
            LB_PUT(x, k, LB_AT(l, k));
            LB_PUT(d, ibd, 0.);
        }
    }
    LB_UNROLL
    for (int i = 1; i <= n; ++i) {
        if (i <= nsub) {
            int k = ind[i - 1];
            LB_PUT(x, k, LB_AT(x, k) + alpha * d[i - 1]);
        }
    }
    return 0;
}

/* subalgorithms.c matupd: append the pair (s, y) = (d, r), dropping the
 * oldest when m are held (the reference advances its ring-buffer head and
 * tail; here the pairs move one column down), and update S'Y, S'S. */
FSX_HD void lb_matupd(fsx_lb *w, int iupdat, int *col, double *theta, double rr, double dr,
                      double stp, double dtd)
{
    const int n = FSX_LB_N, m = FSX_LB_M;
    if (iupdat <= m) {
        *col = iupdat;
    } else {
        LB_UNROLL
        for (int j = 1; j <= m - 1; ++j) {
            LB_UNROLL
            for (int i = 1; i <= n; ++i) {
                LB_WS(i, j) = LB_WS(i, j + 1);
                LB_WY(i, j) = LB_WY(i, j + 1);
            }
        }
    }
    const int c = *col;
    *theta = rr / dr;
    /* when iupdat > m, col = m: move the old parts of S'S and S'Y */
    if (iupdat > m) {
        LB_UNROLL
        for (int j = 1; j <= m - 1; ++j) {
            LB_UNROLL
            for (int k = 0; k < j; k++) LB_SS(1 + k, j) = LB_SS(2 + k, j + 1);
            LB_UNROLL
            for (int k = 0; k < m - j; k++) LB_SY(j + k, j) = LB_SY(j + 1 + k, j + 1);
        }
    }
    LB_UNROLL
    for (int cc = 1; cc <= m; ++cc) {
        if (cc == c) {
            LB_UNROLL
            for (int i = 1; i <= n; ++i) {
                LB_WS(i, cc) = w->d[i - 1];
                LB_WY(i, cc) = w->r[i - 1];
            }
            LB_UNROLL
            for (int j = 1; j <= cc - 1; ++j) {
                double sy = 0., ss = 0.;
                LB_UNROLL
                for (int i = 1; i <= n; ++i) sy = sy + w->d[i - 1] * LB_WY(i, j);
                LB_UNROLL
                for (int i = 1; i <= n; ++i) ss = ss + LB_WS(i, j) * w->d[i - 1];
                LB_SY(cc, j) = sy;
                LB_SS(j, cc) = ss;
            }
            if (stp == 1.) LB_SS(cc, cc) = dtd;
            else LB_SS(cc, cc) = stp * stp * dtd;
            LB_SY(cc, cc) = dr;
        }
    }
}

/* linesearch.c dcstep.  The in/out arguments are copied into locals and
 * back, so that merged branches never select between their addresses. */
FSX_HD void lb_dcstep(double *stx_, double *fx_, double *dx_, double *sty_, double *fy_,
                      double *dy_, double *stp_, double fp, double dp, int *brackt_,
                      double stpmin, double stpmax)
{
    double stx = *stx_, fx = *fx_, dx = *dx_, sty = *sty_, fy = *fy_, dy = *dy_;
    double stp = *stp_;
    int brackt = *brackt_;
    double sgnd = dp * (dx / lb_abs(dx));
    double stpf, stpc, stpq, theta, s, gamma, p, q, r, tt;
    if (fp > fx) {
        theta = (fx - fp) * 3. / (stp - stx) + dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(tt * tt - dx / s * (dp / s));
        if (stp < stx) gamma = -gamma;
        p = gamma - dx + theta;
        q = gamma - dx + gamma + dp;
        r = p / q;
        stpc = stx + r * (stp - stx);
        stpq = stx + dx / ((fx - fp) / (stp - stx) + dx) / 2. * (stp - stx);
        if (lb_abs(stpc - stx) < lb_abs(stpq - stx)) stpf = stpc;
        else stpf = stpc + (stpq - stpc) / 2.;
        brackt = 1;
    } else if (sgnd < 0.) {
        theta = (fx - fp) * 3. / (stp - stx) + dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(tt * tt - dx / s * (dp / s));
        if (stp > stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma - dp + gamma + dx;
        r = p / q;
        stpc = stp + r * (stx - stp);
        stpq = stp + dp / (dp - dx) * (stx - stp);
        if (lb_abs(stpc - stp) > lb_abs(stpq - stp)) stpf = stpc;
        else stpf = stpq;
        brackt = 1;
    } else if (lb_abs(dp) < lb_abs(dx)) {
        theta = (fx - fp) * 3. / (stp - stx) + dx + dp;
        s = lb_max(lb_max(lb_abs(theta), lb_abs(dx)), lb_abs(dp));
        tt = theta / s;
        gamma = s * sqrt(lb_max(0., tt * tt - dx / s * (dp / s)));
        if (stp > stx) gamma = -gamma;
        p = gamma - dp + theta;
        q = gamma + (dx - dp) + gamma;
        r = p / q;
        if (r < 0. && gamma != 0.) stpc = stp + r * (stx - stp);
        else if (stp > stx) stpc = stpmax;
        else stpc = stpmin;
        stpq = stp + dp / (dp - dx) * (stx - stp);
        if (brackt) {
            if (lb_abs(stpc - stp) < lb_abs(stpq - stp)) stpf = stpc;
            else stpf = stpq;
            if (stp > stx) stpf = lb_min(stp + (sty - stp) * .66, stpf);
            else stpf = lb_max(stp + (sty - stp) * .66, stpf);
        } else {
            if (lb_abs(stpc - stp) > lb_abs(stpq - stp)) stpf = stpc;
            else stpf = stpq;
            stpf = lb_min(stpmax, stpf);
            stpf = lb_max(stpmin, stpf);
        }
    } else {
        if (brackt) {
            theta = (fp - fy) * 3. / (sty - stp) + dy + dp;
            s = lb_max(lb_max(lb_abs(theta), lb_abs(dy)), lb_abs(dp));
            tt = theta / s;
            gamma = s * sqrt(tt * tt - dy / s * (dp / s));
            if (stp > sty) gamma = -gamma;
            p = gamma - dp + theta;
            q = gamma - dp + gamma + dy;
            r = p / q;
            stpc = stp + r * (sty - stp);
            stpf = stpc;
        } else if (stp > stx) {
            stpf = stpmax;
        } else {
            stpf = stpmin;
        }
    }
    if (fp > fx) {
        sty = stp; fy = fp; dy = dp;
    } else {
        if (sgnd < 0.) { sty = stx; fy = fx; dy = dx; }
        stx = stp; fx = fp; dx = dp;
    }
    stp = stpf;
    *stx_ = stx; *fx_ = fx; *dx_ = dx;
    *sty_ = sty; *fy_ = fy; *dy_ = dy;
    *stp_ = stp;
    *brackt_ = brackt;
}

typedef struct fsx_ls {
    int brackt, stage;
    double ginit, gtest, gx, gy, finit, fx, fy, stx, sty, stmin, stmax, width, width1;
} fsx_ls;

/* linesearch.c dcsrch with ftol 1e-3, gtol 0.9, xtol 0.1, stpmin 0 */
FSX_HD void lb_dcsrch(double f, double g, double *stp, double stpmin, double stpmax,
                      int *task, fsx_ls *s)
{
    const double ftol = .001, gtol = .9, xtol = .1;
    if (*task == FSX_T_START) {
        if (*stp < stpmin) *task = FSX_T_E_SMALLSTP;
        if (*stp > stpmax) *task = FSX_T_E_LARGESTP;
        if (g >= 0.) *task = FSX_T_E_INITIAL;
        if (stpmin < 0.) *task = FSX_T_E_STP0;
        if (stpmax < stpmin) *task = FSX_T_E_STP1;
        if (*task >= 200 && *task <= 240) return;
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
        *task = FSX_T_FG;
        return;
    }
    double ftest = s->finit + *stp * s->gtest;
    if (s->stage == 1 && f <= ftest && g >= 0.) s->stage = 2;
    if (s->brackt && (*stp <= s->stmin || *stp >= s->stmax)) *task = FSX_T_W_ROUND;
    if (s->brackt && s->stmax - s->stmin <= xtol * s->stmax) *task = FSX_T_W_XTOL;
    if (*stp == stpmax && f <= ftest && g <= s->gtest) *task = FSX_T_W_STPMAX;
    if (*stp == stpmin && (f > ftest || g >= s->gtest)) *task = FSX_T_W_STPMIN;
    if (f <= ftest && lb_abs(g) <= gtol * (-s->ginit)) *task = FSX_T_CONVERGENCE;
    if ((*task >= 100 && *task <= 110) || (*task >= 20 && *task <= 25)) return;
    /* one dcstep, on the modified function (stage 1, lower function value
     * with sufficient decrease not yet met) or on f itself */
    const int modf = s->stage == 1 && f <= s->fx && f > ftest;
    double fv, gv, fxv, fyv, gxv, gyv;
    if (modf) {
        fv = f - *stp * s->gtest;
        fxv = s->fx - s->stx * s->gtest;
        fyv = s->fy - s->sty * s->gtest;
        gv = g - s->gtest;
        gxv = s->gx - s->gtest;
        gyv = s->gy - s->gtest;
    } else {
        fv = f; fxv = s->fx; fyv = s->fy;
        gv = g; gxv = s->gx; gyv = s->gy;
    }
    lb_dcstep(&s->stx, &fxv, &gxv, &s->sty, &fyv, &gyv, stp, fv, gv, &s->brackt, s->stmin,
              s->stmax);
    if (modf) {
        s->fx = fxv + s->stx * s->gtest;
        s->fy = fyv + s->sty * s->gtest;
        s->gx = gxv + s->gtest;
        s->gy = gyv + s->gtest;
    } else {
        s->fx = fxv; s->fy = fyv;
        s->gx = gxv; s->gy = gyv;
    }
    if (s->brackt) {
        if (lb_abs(s->sty - s->stx) >= s->width1 * .66) *stp = s->stx + (s->sty - s->stx) * .5;
        s->width1 = s->width;
        s->width = lb_abs(s->sty - s->stx);
    }
    if (s->brackt) {
        s->stmin = lb_min(s->stx, s->sty);
        s->stmax = lb_max(s->stx, s->sty);
    } else {
        s->stmin = *stp + (*stp - s->stx) * 1.1;
        s->stmax = *stp + (*stp - s->stx) * 4.;
    }
    *stp = lb_max(*stp, stpmin);
    *stp = lb_min(*stp, stpmax);
    if ((s->brackt && (*stp <= s->stmin || *stp >= s->stmax))
        || (s->brackt && s->stmax - s->stmin <= xtol * s->stmax))
        *stp = s->stx;
    *task = FSX_T_FG;
}

FSX_HD void lb_zero(fsx_lb *w)
{
    LB_UNROLL
    for (int i = 0; i < FSX_LB_N; i++) {
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) { w->ws[i][j] = 0.0; w->wy[i][j] = 0.0; }
    }
    LB_UNROLL
    for (int i = 0; i < FSX_LB_M; i++) {
        LB_UNROLL
        for (int j = 0; j < FSX_LB_M; j++) {
            w->sy[i][j] = 0.0; w->ss[i][j] = 0.0; LB_WT0(i, j) = 0.0;
            LB_K11(i, j) = 0.0; LB_K12(i, j) = 0.0; LB_K22(i, j) = 0.0;
        }
    }
    LB_UNROLL
    for (int i = 0; i < 2 * FSX_LB_M; i++) {
        w->wc[i] = 0.0;
        LB_UNROLL
        for (int j = 0; j < 2 * FSX_LB_M; j++) LB_WN1(i + 1, j + 1) = 0.0;
    }
    LB_UNROLL
    for (int i = 0; i < FSX_LB_N; i++) {
        w->z[i] = w->r[i] = w->d[i] = w->t[i] = 0.0;
        w->index[i] = w->iwhere[i] = w->indx2[i] = 0;
    }
}

/* The objective evaluation requested by the wrapper on every FG task
 * (min_saxs.c:238-241): refresh the scale, then chi^2 and gradient. */
FSX_HD void lb_eval(const fsx_ct *t, const fsx_fitpar *p, const double *x, double *f,
                    double *g, double *kscale)
{
    double bA[FSX_QB][6];
    fsx_ct_block(t, 0, FSX_QB_MIN(p->qnum, FSX_QB), bA);
    double k = fsx_best_scale_blk(t, p, x[0], x[1], bA);
    *kscale = k;
    fsx_ct_block(t, 0, FSX_QB_MIN(p->qnum, FSX_QB), bA);
    fsx_objective_blk(t, p, k, x, g, f, bA);
}

FSX_HD void lb_eval_1p(const fsx_ct *t, const fsx_fitpar *p, const double *x, double *f,
                       double *g)
{
    double bA[FSX_QB][6];
    fsx_ct_block(t, 0, FSX_QB_MIN(p->qnum, FSX_QB), bA);
    fsx_eval_1p_blk(t, p, x, g, f, bA);
}

/* ---------------------------------------------------------------------- */
/* Resumable driver.  setulb's reverse communication (lbfgsb.c:52-515, driven
 * by min_saxs.c:234-248) is kept as an explicit state machine:
 * fsx_lb_start() sets the first trial point s->x; the caller evaluates f
 * and g there into s->f / s->g and calls fsx_lb_resume(), which runs the
 * optimiser up to the next point it needs evaluated (returns 1, point in
 * s->x) or to the end of the fit (returns 0, reason in s->task).  Since
 * every NEW_X is resumed at once, the visited points are those of the
 * reference.  Many fits can share one evaluation loop this way, and the
 * slot of a finished fit can start the next one. */
#define FSX_PH_INIT 0
#define FSX_PH_LS 1
#define FSX_PH_DONE 2

typedef struct fsx_lbrc {
    fsx_lb w;
    fsx_ls ls;
    double x[FSX_LB_N], g[FSX_LB_N], f;
    double l[FSX_LB_N], u[FSX_LB_N];
    double theta, fold, gd, stpmx, sbgnrm, stp, gdold, dtd, tol, pgtol;
    int col, iupdat, updatd, iback, iter, nseg, nfree, ifun, nenter, ileave;
    int wrk, csave, nev, task, phase;
} fsx_lbrc;

/* x0 = (c1, c2) projected into the box [lo, hi] (active, lbfgsb.c:540-600,
 * every variable bounded on both sides); the first evaluation is due at
 * s->x. */
FSX_HD void fsx_lb_start(fsx_lbrc *s, const fsx_fitpar *p, double x0, double x1)
{
    lb_zero(&s->w);
    s->ls.brackt = 0; s->ls.stage = 0;
    s->ls.ginit = s->ls.gtest = s->ls.gx = s->ls.gy = s->ls.finit = s->ls.fx = s->ls.fy = 0.0;
    s->ls.stx = s->ls.sty = s->ls.stmin = s->ls.stmax = s->ls.width = s->ls.width1 = 0.0;
    s->l[0] = p->lo0; s->l[1] = p->lo1;
    s->u[0] = p->hi0; s->u[1] = p->hi1;
    s->x[0] = x0; s->x[1] = x1;
    s->g[0] = 0.0; s->g[1] = 0.0; s->f = 0.0;
    s->col = 0; s->iupdat = 0; s->updatd = 0; s->iback = 0;
    s->iter = 0; s->nseg = 0; s->nfree = FSX_LB_N; s->ifun = 0; s->nenter = 0; s->ileave = 0;
    s->theta = 1.; s->fold = 0.; s->gd = 0.; s->stpmx = 0.; s->sbgnrm = 0.; s->stp = 0.;
    s->gdold = 0.; s->dtd = 0.;
    s->tol = p->factr * DBL_EPSILON;
    s->pgtol = p->pgtol;
    s->wrk = 0; s->csave = 0; s->nev = 0; s->task = FSX_T_FG; s->phase = FSX_PH_INIT;
    for (int i = 0; i < FSX_LB_N; i++) {
        if (s->x[i] <= s->l[i]) {
            if (s->x[i] < s->l[i]) s->x[i] = s->l[i];
        } else if (s->x[i] >= s->u[i]) {
            if (s->x[i] > s->u[i]) s->x[i] = s->u[i];
        }
        s->w.iwhere[i] = (s->u[i] - s->l[i] <= 0.) ? 3 : 0;
    }
}

/* mainlb from the point after an evaluation to the next one (or the end). */
FSX_HD int fsx_lb_resume(fsx_lbrc *s)
{
    fsx_lb *w = &s->w;
    const int n = FSX_LB_N;
    double *x = s->x, *g = s->g;
    const double *l = s->l, *u = s->u;
    const double epsmch = DBL_EPSILON;
    int info = 0, restart, skip_search;
    double rr, dr, ddum;
    if (s->phase == FSX_PH_DONE) return 0;
    s->nev++;
    if (s->phase == FSX_PH_LS) goto ls_resume;
#define LB_PROJGR(out)                                                          \
    do {                                                                        \
        double s_ = 0.;                                                         \
        for (int i_ = 0; i_ < n; i_++) {                                        \
            double gi_ = g[i_];                                                 \
            if (gi_ < 0.) gi_ = lb_max(x[i_] - u[i_], gi_);                     \
            else gi_ = lb_min(x[i_] - l[i_], gi_);                              \
            s_ = lb_max(s_, lb_abs(gi_));                                       \
        }                                                                       \
        (out) = s_;                                                             \
    } while (0)
#define LB_RESET()                                                              \
    do {                                                                        \
        info = 0; s->col = 0; s->theta = 1.;                                    \
        s->iupdat = 0; s->updatd = 0;                                           \
    } while (0)
    /* the first evaluation (at the projected x0) is done */
    LB_PROJGR(s->sbgnrm);
    if (s->sbgnrm <= s->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
    for (;;) {
        info = lb_cauchy(w, x, l, u, g, s->theta, s->col, s->sbgnrm, &s->nseg);
        if (info) {
            LB_RESET();
            continue;
        }
        lb_freev(w, &s->nfree, &s->nenter, &s->ileave, &s->wrk, s->updatd, s->iter);
        if (!(s->nfree == 0 || s->col == 0)) {
            if (s->wrk)
                info = lb_formk(w, s->nfree, s->nenter, s->ileave, s->iupdat, s->updatd,
                                s->theta, s->col);
            if (info) {
                LB_RESET();
                continue;
            }
            info = lb_cmprlb(w, x, g, s->theta, s->col, s->nfree);
            if (!info) info = lb_subsm(w, s->nfree, l, u, x, g, s->theta, s->col);
            if (info) {
                LB_RESET();
                continue;
            }
        }
        for (int i = 0; i < n; i++) w->d[i] = w->z[i] - x[i];
        /* lnsrlb entry (every variable boxed) */
        s->dtd = lb_dot2(w->d, w->d);
        if (s->iter == 0) {
            s->stpmx = 1.;
        } else {
            s->stpmx = 1e10;
            for (int i = 0; i < n; i++) {
                double a1 = w->d[i];
                if (a1 < 0.) {
                    double a2 = l[i] - x[i];
                    if (a2 >= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx < a2) s->stpmx = a2 / a1;
                } else if (a1 > 0.) {
                    double a2 = u[i] - x[i];
                    if (a2 <= 0.) s->stpmx = 0.;
                    else if (a1 * s->stpmx > a2) s->stpmx = a2 / a1;
                }
            }
        }
        s->stp = 1.;
        for (int i = 0; i < n; i++) { w->t[i] = x[i]; w->r[i] = g[i]; }
        s->fold = s->f;
        s->ifun = 0;
        s->iback = 0;
        s->csave = FSX_T_START;
        restart = 0;
        for (;;) {
            s->gd = lb_dot2(g, w->d);
            skip_search = 0;
            if (s->ifun == 0) {
                s->gdold = s->gd;
                if (s->gd >= 0.) { info = -4; skip_search = 1; }
            }
            if (!skip_search) {
                lb_dcsrch(s->f, s->gd, &s->stp, 0., s->stpmx, &s->csave, &s->ls);
                if (!(s->csave >= 100 && s->csave <= 110) && !(s->csave >= 20 && s->csave <= 25)) {
                    s->task = FSX_T_FG_LN;
                    ++s->ifun;
                    s->iback = s->ifun - 1;
                    if (s->stp == 1.) {
                        for (int i = 0; i < n; i++) x[i] = w->z[i];
                    } else {
                        for (int i = 0; i < n; i++) x[i] = s->stp * w->d[i] + w->t[i];
                    }
                } else {
                    s->task = FSX_T_NEW_X;
                }
            }
            if (info != 0 || s->iback >= 20) {
                for (int i = 0; i < n; i++) { x[i] = w->t[i]; g[i] = w->r[i]; }
                s->f = s->fold;
                if (s->col == 0) {
                    s->task = FSX_T_ABNORMAL;
                    ++s->iter;
                    goto done;
                }
                LB_RESET();
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
                LB_PROJGR(s->sbgnrm);
                break;
            }
        }
        if (restart) continue;
        if (s->sbgnrm <= s->pgtol) { s->task = FSX_T_CONV_GRAD; goto done; }
        ddum = lb_max(lb_max(lb_abs(s->fold), lb_abs(s->f)), 1.);
        if (s->fold - s->f <= s->tol * ddum) { s->task = FSX_T_CONV_F; goto done; }
        for (int i = 0; i < n; i++) w->r[i] = g[i] - w->r[i];
        rr = lb_dot2(w->r, w->r);
        if (s->stp == 1.) {
            dr = s->gd - s->gdold;
            ddum = -s->gdold;
        } else {
            dr = (s->gd - s->gdold) * s->stp;
            for (int i = 0; i < n; i++) w->d[i] = s->stp * w->d[i];
            ddum = -s->gdold * s->stp;
        }
        if (dr <= epsmch * ddum) {
            s->updatd = 0;
            continue;
        }
        s->updatd = 1;
        ++s->iupdat;
        lb_matupd(w, s->iupdat, &s->col, &s->theta, rr, dr, s->stp, s->dtd);
        if (lb_formt(w, s->col, s->theta)) LB_RESET();
    }
#undef LB_PROJGR
#undef LB_RESET
done:
    s->phase = FSX_PH_DONE;
    return 0;
}

/* One complete fit from x = (c1, c2).  Returns the final task; x is the
 * result, *fout = chi^2, *nevals = objective evaluations. */
FSX_HD int fsx_lbfgsb_fit(const fsx_ct *ct, const fsx_fitpar *p, double *x, double *fout,
                          int *nevals FSX_LB_KARG)
{
    fsx_lbrc S;
    double k;
    FSX_LB_KSET(S);
    fsx_lb_start(&S, p, x[0], x[1]);
    do {
        if (p->onepass)
            lb_eval_1p(ct, p, S.x, &S.f, S.g);
        else
            lb_eval(ct, p, S.x, &S.f, S.g, &k);
    } while (fsx_lb_resume(&S));
    x[0] = S.x[0];
    x[1] = S.x[1];
    *fout = S.f;
    *nevals = S.nev;
    return S.task;
}

#endif
