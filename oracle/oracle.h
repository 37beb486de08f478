/*
 * oracle.h -- FP64 CPU restatement of libfmftsaxs's FFT dimer SAXS scoring
 * path, used ONLY as a test/benchmark checker.
 *
 * TEST INFRASTRUCTURE: nothing in the shipped library (libfmftsaxs_amd/)
 * links, imports or calls this code.  Only tests/, the smoke entry point
 * and bench.py's cpu_baseline leg use it, as the checker or as the timed CPU
 * baseline.  Every function cites the reference file:line it restates.
 *
 * Parity pins (see the tests named beside each group):
 *   orc_sbessel            bit-exact vs oracle/_ref (sfbessel.c) ........ tests/test_oracle_ref.py
 *   orc_wigner_d           bit-exact vs oracle/_ref (borrowed.c) ........ tests/test_oracle_ref.py
 *   orc_sph_norm/legendre  bit-exact vs oracle/_ref (borrowed.c) ........ tests/test_oracle_ref.py
 *   orc_w3j/orc_d_symb     vs oracle/_ref (borrowed.c, GMP) and exact sympy 3j
 *   orc_lbfgsb_*           bit-exact iterates vs oracle/_ref (lbfgsb/src)
 *   fit objective          chi^2 expansion identity vs brute-force interpolated chi^2
 *   correlation driver     per-pose direct rotation/translation evaluation on tiny L,
 *                          closed forms (T(z=0)=I, rotation invariance), and the
 *                          reference's golden score table tests/data/ref_chi
 *   atom expansion         Debye-sum limit, tests/data/ref_spf / ref_profile goldens
 *
 * Conventions: all arrays are plain C arrays (no structs), 0-based.
 * SPF coefficients are stored as [X][q][lm] with X in {V,D,W}, lm = l(l+1)+m,
 * separately for real and imaginary parts (pdb2spf.h:30-53 layout, flattened).
 */
#ifndef FSX_ORACLE_H
#define FSX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pi used by the reference CMake build (-std=c11 => define.h:13-15). */
#define ORC_PI_TRUNC 3.14159265358
#define ORC_PI_EXACT 3.14159265358979323846

/* Set/get the pi constant used by every oracle routine that the reference
 * computes with M_PI (the -std=c11 build takes define.h's truncated value).
 * Default: ORC_PI_TRUNC. */
void orc_set_pi(double pi);
double orc_get_pi(void);

/* ---- L1 math -------------------------------------------------------- */
double orc_sbessel(int l, double x);                               /* sfbessel.c:40-60 */
void   orc_sph_norm(int N, double *y_norm);                        /* borrowed.c:19-38, N*(2N-1) */
void   orc_legendre(int N, double x, double *P);                   /* borrowed.c:40-65, N*(2N-1) */
void   orc_wigner_d(int L, double beta, double *d);                /* borrowed.c:243-313, (L+1)(2L+1)^2 */
double orc_w3j(int j1, int j2, int j3, int m1, int m2, int m3);   /* borrowed.c:86-222 (exact, truncated) */
void   orc_d_symb(int L, double *d_symb);                          /* fftsaxs.c:138-174, (L+1)^3 (2L+1) */

/* ---- atom expansion (pdb2spf.c:24-152) ------------------------------ */
/* coords: natoms*3; vac/dum: per-atom scalar form factors; sa: per-atom
 * fractional SASA or NULL; vdw: per-atom vdW radius (rm).  Output re/im:
 * [3][qnum][(L+1)^2].  Returns rm. */
double orc_atoms2spf(int natoms, const double *coords, const double *vac,
                     const double *dum, const double *sa, double h2o_ff,
                     const double *vdw, const double *qvals, int qnum, int L,
                     double *re, double *im);

/* I(q) of one SPF at (c1, c2) (profile.c:186-233, G uses qvals[0]: profile.c:214). */
void orc_profile_from_spf(const double *re, const double *im, int L,
                          const double *qvals, int qnum, double rm,
                          double c1, double c2, double rerr,
                          double *in, double *err);
/* Cross terms of one SPF (min_saxs.c:450-505). out: [6][qnum] VV,VD,VW,DD,DW,WW */
void orc_spf2cross_terms(const double *re, const double *im, int L, int qnum,
                         double *out);

/* ---- fit (min_saxs.c) ------------------------------------------------ */
typedef struct orc_fit_params {
    double *a;      /* 6*qnum moments (scoring_helper) */
    double rm;
    double mult;
    double peak;
    double c1_lo, c1_hi, c2_lo, c2_hi;
    double factr, pgtol;
    int m;          /* L-BFGS-B memory (3) */
} orc_fit_params;

/* scoring_helper, min_saxs.c:353-389.  n_exp bounds the experimental
 * arrays (the reference's loop has no bound and reads past the end; here the
 * scan stops). */
void orc_scoring_helper(const double *exp_q, const double *exp_in,
                        const double *exp_err, int n_exp,
                        const double *qvals, int qnum, double *a);
/* sxs_opt_params_init, min_saxs.c:116-124 (mult, peak). */
void orc_params_init(orc_fit_params *p, double *a, double rm, double peak);
double orc_best_scale(const double *ct, const double *q, int qnum,
                      const orc_fit_params *p, double c1, double c2); /* :261-319 */
void orc_gradient(const double *ct, const double *q, int qnum,
                  const orc_fit_params *p, double k, const double *x,
                  double *g, double *f);                               /* :3-105 */
void orc_compile_intensity(const double *ct, const double *q, int qnum,
                           double rm, double c1, double c2, double *in); /* :321-350 */

/* Per-pose fit: peak rescale (min_saxs.c:170-188) + L-BFGS-B
 * (min_saxs.c:196-259).  ct is modified in place (rescaled).  Returns the
 * number of f/g evaluations; out3 = {score, c1, c2}. */
int orc_fit_one(double *ct, const double *q, int qnum, const orc_fit_params *p,
                int rescale, double *out3);

/* ---- L-BFGS-B 3.0 restatement (lbfgsb/src) ---------------------------- */
typedef void (*orc_fg_fn)(void *ctx, const double *x, double *f, double *g);
typedef struct orc_lbfgsb_result {
    int task;       /* final task code, same numbering as lbfgsb.h:80-130 */
    int iter;       /* iterations */
    int nfgv;       /* f/g evaluations as counted by mainlb */
    int nevals;     /* actual callback calls */
    double f;
} orc_lbfgsb_result;
int orc_lbfgsb_minimize(int n, int m, double *x, const double *l,
                        const double *u, const int *nbd, orc_fg_fn fg,
                        void *ctx, double factr, double pgtol, int max_evals,
                        orc_lbfgsb_result *res);

/* ---- grid index (index.c:3-29) ------------------------------------- */
int64_t orc_assemble_index(int z, int b1, int b2, int a2, int g1, int g2,
                           int nbeta, int na, int ng);
void orc_disassemble_index(int64_t index, int nbeta, int na, int ng, int *out6);
/* correlate.c:215-251: snap one Euler line to the native grid index.
 * Returns 1 and writes *index when the pose's z matches zvals[j]; the
 * first matching j is used (the reference emits one entry per match). */
int orc_euler2index(double z, double b1, double g1, double a2, double b2,
                    double g2, const double *zvals, int znum, int L,
                    int *index_out, int max_out);

/* ---- correlation driver (fftsaxs.c:608-986) ------------------------- */
typedef struct orc_grid {
    int L;
    int nbeta;      /* reference: L+1 */
    int na;         /* alpha_B points; reference: 2L+1 */
    int ng;         /* gamma_A, gamma_B points; reference: 2L+1 */
    int dense_threshold; /* reference: 30 (fftsaxs.c:899) */
    int nthreads;   /* OpenMP threads over slabs (0 = default) */
} orc_grid;

/* A_re/A_im/B_re/B_im: [3][qnum][(L+1)^2].  index_list uses the mixed radix
 * [z][b1][b2][a2][g1][g2] with radices (znum, nbeta, nbeta, na, ng, ng).
 * Outputs for poses whose z slot is out of range are left untouched
 * (fftsaxs.c:922-936).  cross_out (optional, [nout][6][qnum]) receives the
 * un-rescaled cross-term profiles of each pose. twiddle_out (optional,
 * [nout]) receives 1 if the pose's slab used the exact-twiddle transform.
 * skip: 1 = only slabs holding requested poses (reference skip=1);
 *       0 = every slab is treated as dense (reference skip=0). */
void orc_compute_saxs_scores(double *scores, double *c1s, double *c2s,
                             const int64_t *index_list, int nout,
                             const double *A_re, const double *A_im,
                             const double *B_re, const double *B_im,
                             const orc_fit_params *p,
                             const double *qvals, int qnum,
                             const double *zvals, int znum,
                             const orc_grid *grid, int skip,
                             double *cross_out, int *twiddle_out);

/* Cross terms of one pose by direct rotation+translation (independent
 * check, O(L^4) per q): rotate A by D(0,b1,g1), B by D(a2,b2,g2) after
 * translating B by z along the axis, then pair sum_lm.  Angles in radians;
 * exact (non-truncated) twiddles are implied. out: [6][qnum]. */
void orc_direct_pose_cross_terms(const double *A_re, const double *A_im,
                                 const double *B_re, const double *B_im,
                                 int L, const double *qvals, int qnum,
                                 double z, double beta1, double gamma1,
                                 double alpha2, double beta2, double gamma2,
                                 double *out);

#ifdef __cplusplus
}
#endif
#endif
