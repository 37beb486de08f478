/*
 * min_saxs.h -- chi^2 fit of (c1, c2) against an experimental curve.
 * Drop-in for the reference's src/min_saxs.h:32-179.  sxs_opt_params keeps
 * the reference layout, including the L-BFGS-B workspaces it embeds
 * (min_saxs.h:41-45; integer/logical are long as in lbfgsb.h:43-45); this
 * implementation keeps its optimiser state on the stack instead, so the
 * fitting entry points are re-entrant.
 */
#ifndef FSX_MIN_SAXS_H
#define FSX_MIN_SAXS_H

#include "profile.h"
#include "pdb2spf.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifndef FSX_LBFGSB_INTEGER_DEFINED
#define FSX_LBFGSB_INTEGER_DEFINED
typedef long int integer;
typedef long int logical;
#endif

/* min_saxs.h:32-46 */
struct sxs_opt_params {
    double rm;
    double mult;
    double peak;
    double *a;
    double wa[43251];
    integer iwa[3072];
    double dsave[29];
    integer isave[44];
    logical lsave[4];
};

struct sxs_opt_params *sxs_opt_params_create(struct sxs_profile *exp, double *qvals, int qnum, double rm); /* min_saxs.c:108-114 */
void sxs_opt_params_init(struct sxs_opt_params *params, struct sxs_profile *exp, double *qvals, int qnum, double rm); /* :116-124 */
void sxs_opt_params_destroy(struct sxs_opt_params *params);  /* min_saxs.c:126-144 */
void sxs_opt_params_free(struct sxs_opt_params *params);     /* min_saxs.c:146-150 */
/* min_saxs.c:153-194: peak-rescale and fit every profile with mask[i] == 1 */
void sxs_fit_params(struct sxs_profile **profiles, struct sxs_opt_params *params, int *mask, int n);
void sxs_lbfgs_fitting(struct sxs_profile *profile, struct sxs_opt_params *params); /* :196-259 */
double *scoring_helper(struct sxs_profile *exp, int qnum, double *qvals);         /* :353-389 */
void sxs_compile_intensity(struct sxs_profile *profile, double rm, double c1, double c2); /* :321-350 */
double sxs_best_scale(struct sxs_profile *profile, struct sxs_opt_params *params, double c1, double c2); /* :261-319 */
/* declared at min_saxs.h:161; defined in the reference as point_score (min_saxs.c:391-448) */
double sxs_point_score(struct sxs_profile *profile, struct sxs_opt_params *params, double c1, double c2);
void sxs_spf2cross_terms(struct sxs_profile *profile, struct sxs_spf_full *s);    /* :450-505 */
void sxs_spf2fitted_profile(struct sxs_profile *profile, struct sxs_spf_full *s, struct sxs_opt_params *params); /* :507-511 */

#ifdef __cplusplus
}
#endif
#endif
