/*
 * fftsaxs.h -- FFT-accelerated dimer SAXS scoring.  Drop-in for the
 * reference's src/fftsaxs.h:95-105: same name, signature, index encoding
 * (index.c:3-29, radices znum, L+1, L+1, 2L+1, 2L+1, 2L+1), output
 * semantics (entries whose z slot is outside [0, znum) are left untouched,
 * fftsaxs.c:922-936) and exit-on-error behaviour.  The work runs on the GPU
 * through the fsx_score entry point of fsx.h.
 */
#ifndef FSX_FFTSAXS_H
#define FSX_FFTSAXS_H

#include "sfbessel.h"
#include "pdb2spf.h"
#include "min_saxs.h"
#include "profile.h"
#include "index.h"

#ifdef __cplusplus
extern "C" {
#endif

void sxs_compute_saxs_scores(double *scores_list, double *c1_list, double *c2_list,
                             int *index_list, int nout,
                             struct sxs_spf_full *A, struct sxs_spf_full *B,
                             struct sxs_opt_params *params,
                             double *qvals, int qnum, double *zvals, int znum,
                             int L, int skip);

#ifdef __cplusplus
}
#endif
#endif
