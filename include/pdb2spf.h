/*
 * pdb2spf.h -- spherical-harmonic amplitude expansion types and their
 * lifecycle/text I/O.  Drop-in for the reference's src/pdb2spf.h:30-172
 * (same struct layouts, names, argument meaning and ownership; fatal I/O
 * errors print and exit(EXIT_FAILURE) as the reference does, common.h:26-36).
 */
#ifndef FSX_PDB2SPF_H
#define FSX_PDB2SPF_H

#include <stdio.h>

#include "form_factor_table.h"
#include "fsx_mol.h"

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

/* pdb2spf.c:24-152: SPF of an atom group (coordinates already centred) into
 * spf_coefs (allocated for at least L and qnum): rm = mean vdW radius, per
 * atom the vacuum / dummy form factors from ff_table, the water shell
 * weighted by zero_ff(OH2) * saxs_sa[i] (no water shell if saxs_sa is
 * NULL).  The atom sums run on the GPU.  An atom without a form factor is
 * fatal. */
void atom_grp2spf_inplace(struct sxs_spf_full *spf_coefs, struct mol_atom_group *ag,
                          struct saxs_form_factor_table *ff_table, double *qvals, int qnum,
                          int L, double *saxs_sa);
/* pdb2spf.c:154-175: as above into a new object; water == 1 computes the
 * fractional SASA (probe 1.4 A, sxs_faccs) for the water shell */
struct sxs_spf_full *atom_grp2spf(struct mol_atom_group *ag,
                                  struct saxs_form_factor_table *ff_table, double *qvals,
                                  int qnum, int L, int water);

#ifdef __cplusplus
}
#endif
#endif
