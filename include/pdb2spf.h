/*
 * pdb2spf.h -- spherical-harmonic amplitude expansion types and their
 * lifecycle/text I/O.  Drop-in for the reference's src/pdb2spf.h:30-172
 * (same struct layouts, names, argument meaning and ownership; fatal I/O
 * errors print and exit(EXIT_FAILURE) as the reference does, common.h:26-36).
 */
#ifndef FSX_PDB2SPF_H
#define FSX_PDB2SPF_H

#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* pdb2spf.h:30-35: A_lm for one q; re/im hold (L+1)^2 values at l(l+1)+m */
struct sxs_spf_sing {
    int L;
    double *re;
    double *im;
};

/* pdb2spf.h:41-53: vacuum (V), excluded-volume dummy (D) and water-shell (W)
 * amplitudes for every q, plus the molecule's mean atomic radius rm. */
struct sxs_spf_full {
    int L;
    int qnum;
    double rm;
    struct sxs_spf_sing **V;
    struct sxs_spf_sing **D;
    struct sxs_spf_sing **W;
};

struct sxs_spf_sing *sxs_spf_sing_create(int L);                 /* pdb2spf.c:178-184 */
void sxs_spf_sing_init(struct sxs_spf_sing *s, int L);           /* pdb2spf.c:186-193 */
void sxs_spf_sing_free(struct sxs_spf_sing *s);                  /* pdb2spf.c:195-199 */
void sxs_spf_sing_destroy(struct sxs_spf_sing *s);               /* pdb2spf.c:201-207 */
struct sxs_spf_full *sxs_spf_full_create(int L, int qnum);       /* pdb2spf.c:209-215 */
void sxs_spf_full_init(struct sxs_spf_full *s, int L, int qnum); /* pdb2spf.c:217-233 */
void sxs_spf_full_free(struct sxs_spf_full *s);                  /* pdb2spf.c:235-239 */
void sxs_spf_full_destroy(struct sxs_spf_full *s);               /* pdb2spf.c:241-250 */
void sxs_spf_full_reset(struct sxs_spf_full *s);                 /* pdb2spf.c:252-265 */
void sxs_spf_full_write(char *path, struct sxs_spf_full *s);     /* pdb2spf.c:267-295 */
struct sxs_spf_full *sxs_spf_full_fread(FILE *f);                /* pdb2spf.c:297-341 */
struct sxs_spf_full *sxs_spf_full_read(char *path);              /* pdb2spf.c:343-350 */

#ifdef __cplusplus
}
#endif
#endif
