/*
 * profile.h -- scattering profile with its six cross-term curves.
 * Drop-in for the reference's src/profile.h:23-120 (same layout, names,
 * ownership: a profile borrows qvals; one read from a file owns its own).
 */
#ifndef FSX_PROFILE_H
#define FSX_PROFILE_H

#include <stdio.h>
#include "pdb2spf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* profile.h:23-49 */
struct sxs_profile {
    int qnum;
    double rerr;
    double *in;
    double *err;
    double *qvals;
    double *VV;
    double *VD;
    double *VW;
    double *DD;
    double *DW;
    double *WW;
    double score;
    double scale;
    double c1;
    double c2;
    double b1;
    double g1;
    double a2;
    double b2;
    double g2;
};

struct sxs_profile *sxs_profile_create(double *qvals, int qnum, int cross_terms_flag); /* profile.c:3-9 */
void sxs_profile_init(struct sxs_profile *p, double *qvals, int qnum, int cross_terms_flag); /* :11-37 */
void sxs_profile_alloc_cross_terms(struct sxs_profile *p, int qnum);  /* profile.c:39-47 */
void sxs_profile_destroy(struct sxs_profile *p);                      /* profile.c:49-61 */
void sxs_profile_free(struct sxs_profile *p);                         /* profile.c:63-67 */
void sxs_profile_write(char *path, struct sxs_profile *p);            /* profile.c:69-85 */
struct sxs_profile *sxs_profile_read(char *path);                     /* profile.c:87-96 */
struct sxs_profile *sxs_profile_fread(FILE *f);                       /* profile.c:98-124 */
/* profile.c:186-233: I(q) = sum_lm |V - G D + c2 W|^2 with G taken at qvals[0] */
void sxs_profile_from_spf(struct sxs_profile *p, struct sxs_spf_full *spf, double c1, double c2);

#ifdef __cplusplus
}
#endif
#endif
