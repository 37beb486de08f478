/*
 * saxs_utils.h -- helpers of the reference's src/saxs_utils.h used on the
 * scoring path: mean atomic radius, fractional solvent-accessible surface,
 * the q grid, and the rotation-matrix helpers of the Euler conversion.
 */
#ifndef FSX_SAXS_UTILS_H
#define FSX_SAXS_UTILS_H

#include "fsx_mol.h"

#ifdef __cplusplus
extern "C" {
#endif

/* saxs_utils.c:19-28: mean van der Waals radius */
double mol_atom_group_average_radius(const struct mol_atom_group *ag);

/* saxs_utils.c:30-48: fraction of each atom's surface accessible to a
 * solvent sphere of radius r_solv, as contact area over 4 pi r_vdw^2 (0 for
 * radius-0 atoms).  The reference gets the area from libmol2's accs; here it
 * is Shrake-Rupley with FSX_SASA_POINTS Fibonacci points per atom. */
#define FSX_SASA_POINTS 1500
void sxs_faccs(double *fractional_sa, const struct mol_atom_group *ag, double r_solv);

/* saxs_utils.c:50-63: qnum values from begin to end by cumulative addition */
double *sxs_mkarray(double begin, double end, int qnum);

/* saxs_utils.c:65-79: active rotation R(alpha, beta, gamma) (zyz) */
void sxs_fill_active_rotation_matrix(struct mol_matrix3 *rm, double alpha, double beta,
                                     double gamma);
/* saxs_utils.c:81-94: c = a b */
void sxs_mult_rot_mats(struct mol_matrix3 *c, struct mol_matrix3 *a, struct mol_matrix3 *b);

#ifdef __cplusplus
}
#endif
#endif
