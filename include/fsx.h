/*
 * fsx.h -- extended, re-entrant C ABI of the MI355X scoring path.
 *
 * No counterpart in the reference: it is the flat-array form of
 * sxs_compute_saxs_scores (src/fftsaxs.h:95-105) that the drop-in wrapper
 * calls, generalised to
 *   - any beta grid size (beta_i = i*pi/(nbeta-1); the reference fixes L+1),
 *   - any alpha_B and gamma_A/gamma_B grid sizes (reference: 2L+1),
 *   - 64-bit pose indices (the reference's int index overflows for L >= 20
 *     with 80 z values),
 *   - caller-chosen c1/c2 bounds (reference: define.h:28-31),
 *   - inputs and outputs in host or device (HBM) memory,
 * and it returns error codes instead of exiting.
 *
 * Pose index: mixed radix [z][b1][b2][a2][g1][g2] with radices
 * (znum, nbeta, nbeta, nalpha, ngamma, ngamma), exactly the reference
 * encoding when nbeta = L+1 and nalpha = ngamma = 2L+1 (index.c:3-29).
 *
 * SPF arrays: [3][qnum][(L+1)^2] doubles, block 0 = V (vacuum), 1 = D
 * (excluded volume), 2 = W (water shell); element l(l+1)+m holds A_lm
 * (pdb2spf.h:41-53 flattened).
 *
 * Threading: calls may come from several host threads; calls on one device
 * are serialised internally.
 */
#ifndef FSX_H
#define FSX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FSX_OK 0
#define FSX_ERR_ARG (-1)     /* invalid argument (sizes, NULL pointers) */
#define FSX_ERR_HIP (-2)     /* a HIP runtime call failed */
#define FSX_ERR_NOMEM (-3)   /* host or device allocation failed */
#define FSX_ERR_NODEV (-4)   /* no HIP device visible */
#define FSX_ERR_SCRATCH (-5) /* problem too large for the fixed working memory of
                                fsx_atoms2spf_noalloc */

/* flags for fsx_score */
#define FSX_DEVICE_POSES 1u    /* poses points to device memory */
#define FSX_DEVICE_OUTPUT 2u   /* score/c1/c2/cross point to device memory */
#define FSX_ALL_POSES_VALID 4u /* caller guarantees every index is in range */
#define FSX_MIXED_FP32 8u      /* mixed-precision form: rotated coefficients, translation
                                  blocks and the cross-term contraction in FP32 (Bessel,
                                  Wigner-d, the lane reductions, the per-pose accumulators
                                  and the chi^2 fit stay FP64); the fit is L-BFGS-B with
                                  the 2 x 2 L-BFGS matrix held explicitly, run by lane
                                  groups that hold the profile in registers */
#define FSX_FIT_COMPACT 16u    /* with FSX_MIXED_FP32: fit with the compact-form L-BFGS-B
                                  (the reference's factorisations, lane per fit) instead */

/* pi of the reference CMake build (-std=c11 leaves M_PI undefined, so
 * define.h:13-15 supplies this truncated value) */
#define FSX_PI_REFERENCE 3.14159265358

typedef struct fsx_problem {
    int L;               /* expansion order, 1..40 */
    int nbeta;           /* beta grid points on [0, pi] (reference: L+1) */
    int nalpha;          /* alpha_B grid points on [0, 2pi) (reference: 2L+1) */
    int ngamma;          /* gamma_A / gamma_B grid points (reference: 2L+1) */
    int qnum;
    const double *qvals;         /* host, qnum */
    int znum;
    const double *zvals;         /* host, znum (Angstrom) */
    const double *A_re, *A_im;   /* receptor SPF, host [3][qnum][(L+1)^2] */
    const double *B_re, *B_im;   /* ligand SPF, host */
    const double *a;             /* host [6*qnum], scoring_helper moments */
    double mult;                 /* (4pi/3)^1.5 rm^2 / (16pi), min_saxs.c:121 */
    double peak;                 /* experimental I(q0), min_saxs.c:122 */
    double c1_lo, c1_hi, c2_lo, c2_hi;   /* reference: 0.96, 1.04, -2, 4 */
    double pi;                   /* pi of the beta grid (FSX_PI_REFERENCE) */
} fsx_problem;

typedef struct fsx_stats {
    int n_scored;        /* poses with an in-range z slot */
    int n_capped;        /* fits that hit the evaluation cap (never in parity runs) */
    long fit_evals;      /* total objective evaluations */
    double ms_tables;    /* host tables + uploads */
    double ms_bessel_t;  /* Bessel series + translation matrices (GPU) */
    double ms_rotate;    /* rotated receptor / ligand coefficients (GPU) */
    double ms_group;     /* pose bucketing by (z, beta2, gamma2) (GPU) */
    double ms_cross;     /* per-pose cross-term profiles (GPU) */
    double ms_fit;       /* per-pose chi^2 fits: the fit kernel alone (GPU) */
    double ms_total;     /* whole call, wall clock */
    double cross_bytes;  /* algorithmic bytes of the cross-term launches (receptor rows,
                            ligand rows, per-pose accumulators) */
    double cross_flops;  /* useful FP64 flops of the cross-term launches */
    long cross_items;    /* non-empty (z, beta2, gamma2) buckets x q */
    int cross_chunks;    /* m-chunks (cross-term launches) used */
    int cross_entries;   /* coefficient entries per table row: (L+1)^2, or
                            (L+1)(L+2)/2 when both molecules' coefficients are
                            conjugate symmetric and only m >= 0 is formed */
    double ms_fit_prep;  /* profile rescale (peak / I(q0)) before the fits (GPU) */
    long fit_steps;      /* fit loop iterations (one evaluation and one optimiser advance
                            each) summed over the fit kernel's wavefronts */
    long fit_wide_steps; /* of those, iterations in which the wavefront evaluated its last
                            few fits together (q across the lanes) */
    long fit_max_steps;  /* iterations of the longest-running fit wavefront */
} fsx_stats;

/* Score n poses.  score/c1/c2 receive one value per pose whose z slot is
 * in range; other entries are not written.  cross (optional, may be NULL)
 * receives the un-rescaled cross-term profiles [n][6][qnum] (VV, VD, VW,
 * DD, DW, WW; the values the reference fits, fftsaxs.c:877-905).  stats
 * may be NULL.  Returns FSX_OK or a negative error code. */
int fsx_score(const fsx_problem *pb, const int64_t *poses, int n, double *score,
              double *c1, double *c2, double *cross, unsigned flags, fsx_stats *stats);

/* Batched form of fsx_score: nprob complexes, each with its own problem
 * (L, grids, q, coefficients, fit parameters) and pose set, scored in one
 * call on one device, in order, with the device tables of each complex built
 * in turn (no counterpart in the reference, which scores one complex per
 * call).  poses[i], n[i], score[i], c1[i], c2[i] and, if cross is not NULL,
 * cross[i] (may be NULL) are those of fsx_score for problem i; stats (may be
 * NULL) receives nprob records.  Stops at the first failing problem and
 * returns its error code. */
int fsx_score_batch(int nprob, const fsx_problem *pbs, const int64_t *const *poses,
                    const int *n, double *const *score, double *const *c1, double *const *c2,
                    double *const *cross, unsigned flags, fsx_stats *stats);

/* Cross-term profile of SPF coefficients and fit of one profile on the GPU
 * path's fit code compiled for the host (min_saxs.c:196-259 semantics).
 * ct: [6][qnum] (already rescaled); out: score, c1, c2, evaluations. */
int fsx_fit_host(const double *ct, const double *qvals, int qnum, const double *a,
                 double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                 double *out4);

/* fsx_fit_host in either form: flags 0 (the reference's evaluation order,
 * iterates bit-identical to sxs_lbfgs_fitting) or FSX_MIXED_FP32 (the
 * mixed-precision form's fit: value and gradient from one pass over q, equal
 * to the two-pass ones up to rounding, and L-BFGS-B with the 2 x 2 L-BFGS
 * matrix held explicitly, whose steps are the reference's up to rounding;
 * with FSX_FIT_COMPACT as well, the reference's compact-form L-BFGS-B on the
 * one-pass values).  out4[3] is the number of evaluations.  No counterpart in
 * the reference; it lets a caller reproduce either form's fit on the host. */
int fsx_fit_host_form(const double *ct, const double *qvals, int qnum, const double *a,
                      double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                      unsigned flags, double *out4);

/* Fit many profiles on the GPU: ct [n][6][qnum] (host), already rescaled.
 * out [n][4]: score, c1, c2, evaluations. */
int fsx_fit_batch(const double *ct, int n, const double *qvals, int qnum, const double *a,
                  double mult, double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                  double *out);

/* Spherical-harmonic amplitude expansion of an atom group, the array form of
 * atom_grp2spf_inplace (src/pdb2spf.c:24-152): coords [natoms][3] already
 * centred; vac/dum per-atom scalar form factors; sa per-atom SASA fraction
 * (or NULL: no water shell) weighted by h2o_ff (zero_ff of OH2, 3.50968);
 * vdw per-atom radii (their mean is returned in *rm).  Output re/im
 * [3][qnum][(L+1)^2] (V, D, W) on the host.  Bessel values and atom sums run
 * on the GPU. */
int fsx_atoms2spf(int natoms, const double *coords, const double *vac, const double *dum,
                  const double *sa, double h2o_ff, const double *vdw, const double *qvals,
                  int qnum, int L, double *re, double *im, double *rm);

/* fsx_atoms2spf in either form: flags 0 (the sums over atoms as FP64 MFMA
 * products, within 1e-12 of the reference's FP64 loop) or FSX_MIXED_FP32
 * (the mixed-precision form: each j_l(q r) and f conj(Y_lm) split into three
 * BF16 parts, the six leading part products on BF16 MFMA with FP32
 * accumulation over each 1024-atom chunk, the chunks summed in FP64; about
 * 1e-6 of each (X, q, l) block's largest amplitude).  No counterpart in the
 * reference, which has one (FP64) form. */
int fsx_atoms2spf_form(int natoms, const double *coords, const double *vac, const double *dum,
                       const double *sa, double h2o_ff, const double *vdw, const double *qvals,
                       int qnum, int L, unsigned flags, double *re, double *im, double *rm);

/* Allocation-free form of fsx_atoms2spf.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid allocation).  Every pointer
 * is device memory: coords [natoms][3] (already centred), vac, dum, sa
 * ([natoms]; sa may be NULL: no water shell), qvals [qnum]; re/im
 * [3][qnum][(L+1)^2] receive the amplitudes (the mean radius is not
 * formed).  Working memory is a fixed device array of 64 MiB per GPU, held
 * until the process ends; a problem that does not fit it (it needs
 * 24 qnum16 (L+1)(L+2) bytes per started 1024 atoms for the partial sums,
 * plus 8 (L+1) qnum16 + 8 (L+1)(L+2) + 32 bytes per atom for a pass of
 * min(natoms, 1024) atoms; qnum16 is qnum rounded up to a multiple of 16)
 * returns FSX_ERR_SCRATCH.
 * The call runs on the library's stream and returns when the results are
 * written; device inputs must be ready.  Calls share the one array per GPU:
 * the library's lock orders calls made from several threads of one process,
 * and any other overlap of calls on one GPU would race.  The library's
 * first call (of any entry point) creates its stream. */
int fsx_atoms2spf_noalloc(int natoms, const double *coords, const double *vac,
                          const double *dum, const double *sa, double h2o_ff,
                          const double *qvals, int qnum, int L, double *re, double *im);

/* Brute-force per-pose scoring, the loop of tools/score_ft_naive.c:156-212
 * as a library call.  xyz [nrec + nlig][3]: receptor atoms first, then ligand
 * atoms, each molecule in its starting frame (the tool centres both).  Pose
 * i = poses[6i .. 6i+5] = (z, b1, g1, a2, b2, g2): the receptor is rotated by
 * R(0, b1, g1), the ligand by R(a2, b2, g2) and moved by (0, 0, z)
 * (sxs_fill_active_rotation_matrix), the complex is centred at its centroid,
 * expanded to L (fsx_atoms2spf, sa of the starting complex) and its profile
 * fitted as sxs_spf2fitted_profile (cross terms, then L-BFGS-B from
 * (c1, c2) = (1, 0) in the given box, no peak rescale).  a [6][qnum] and mult
 * are those of sxs_opt_params for the experimental profile and the joint mean
 * radius.  out [n][3]: score, c1, c2; cross (optional, may be NULL)
 * [n][6][qnum]: VV, VD, VW, DD, DW, WW of the complex.  Independent of the
 * FFT correlation: grid pose (b1, g1, a2, b2, g2) of fsx_score is the real-space
 * pose (b1, g1, 2pi - a2, b2, 2pi - g2) here, the mirroring of the Euler-list
 * reader, and the two agree to the truncation error at L. */
int fsx_score_naive(int nrec, int nlig, const double *xyz, const double *vac,
                    const double *dum, const double *sa, double h2o_ff, const double *vdw,
                    const double *qvals, int qnum, int L, const double *a, double mult,
                    double c1_lo, double c1_hi, double c2_lo, double c2_hi,
                    const double *poses, int n, double *out, double *cross);

/* Bessel table j_p(x_i) for p = 0..pmax on the GPU, x host [n]; out [n][pmax+1] */
int fsx_sbessel_batch(const double *x, int n, int pmax, double *out);

/* Correlate front end (tools/correlate.c:202-251): read an Euler list
 * ("ft_id z b1 g1 a2 b2 g2" per line) and snap every line to the grid; a
 * line yields one index per zvals entry within 0.001 A of its z, with the
 * digits rounded separately and not reduced (a full turn carries into the
 * next digit, as in the reference).  Grid: beta step pi/(nbeta-1), alpha
 * and gamma steps 2pi/nalpha and 2pi/ngamma, alpha2 and gamma2 mirrored to
 * 2pi - angle; pi = FSX_PI_REFERENCE reproduces the reference.  *index,
 * *ft_id and *order (input line number) are allocated here; release them
 * with fsx_free. */
int fsx_euler_list_read(const char *path, const double *zvals, int znum, int L, int nbeta,
                        int nalpha, int ngamma, double pi, int64_t **index, int **ft_id,
                        int **order, int *n);
/* Score file of tools/correlate.c:371-381: "order ft_id score c1 c2" lines */
int fsx_scores_write(const char *path, const int *order, const int *ft_id, const double *score,
                     const double *c1, const double *c2, int n);
void fsx_free(void *p);

int fsx_device_count(void);
const char *fsx_error_string(int code);
/* Release the cached device buffers and host tables (Wigner-d, d-symbols)
 * of the calling process; the next call builds and uploads them again (a
 * cold call).  Waits for scoring calls in flight. */
void fsx_release(void);

#ifdef __cplusplus
}
#endif
#endif
