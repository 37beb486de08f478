/*
 * fsx_internal.h -- declarations shared by the host C side and the HIP side
 * of libfmftsaxs_amd (not installed).
 */
#ifndef FSX_INTERNAL_H
#define FSX_INTERNAL_H

#include <stdint.h>
#include "fsx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* host tables (fsx_math.c) */
double fsx_sbessel_host(int l, double x);
void fsx_wigner_d(int L, double beta, double *d);   /* (L+1)(2L+1)^2, reference layout */
double fsx_w3j(int j1, int j2, int j3, int m1, int m2, int m3);
const double *fsx_d_symb_cached(int L);             /* (L+1)^3 (2L+1), process-wide cache */
void fsx_d_symb_drop(void);                         /* empty that cache */

/* everything the GPU pipeline needs, prepared on the host */
typedef struct fsx_hip_args {
    int L, nbeta, nalpha, ngamma, qnum, znum;
    const double *qvals, *zvals;
    const double *A_re, *A_im, *B_re, *B_im;   /* [3][qnum][(L+1)^2] */
    const double *a;                           /* [6*qnum] */
    const double *cint;                        /* [6][qnum]: constant parts, x2 for mixed pairs */
    const double *dW;                          /* [nbeta][(L+1)(2L+1)^2] */
    const double *dsymb;                       /* (L+1)^3 (2L+1) */
    const double *tw_a, *tw_g;                 /* [nalpha][2], [ngamma][2]: cos, -sin of 2 pi k/N */
    double mult, peak, c1_lo, c1_hi, c2_lo, c2_hi;
    const int64_t *poses;
    int n;
    double *score, *c1, *c2, *cross;
    unsigned flags;
    /* nonzero keys identify the content of dW / dsymb: a table whose key
     * matches the one last uploaded into the same device buffer is not
     * uploaded again */
    unsigned long long dw_key, ds_key;
} fsx_hip_args;

int fsx_hip_run(const fsx_hip_args *args, fsx_stats *st);
/* nprob complexes in one call (stats: nprob records, stage times not set) */
int fsx_hip_run_batch(int nprob, const fsx_hip_args *args, fsx_stats *st);
int fsx_hip_fit_batch(const double *ct, int n, const double *qvals, int qnum, const double *a,
                      double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                      double *out);
int fsx_hip_sbessel_batch(const double *x, int n, int pmax, double *out);
/* bf16: the sums over atoms in split BF16 on MFMA (the mixed form) instead
 * of FP64 MFMA */
int fsx_hip_atoms2spf(int natoms, const double *xyz, const double *vac, const double *dum,
                      const double *sa, double h2o_ff, const double *qvals, int qnum, int L,
                      double *re, double *im, int bf16);
int fsx_hip_device_count(void);
void fsx_hip_release(void);

#ifdef __cplusplus
}
#endif
#endif
