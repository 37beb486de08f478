/*
 * index.h -- the mixed-radix grid index of a dimer pose
 * [z][beta1][beta2][alpha2][gamma1][gamma2].  Drop-in for the reference's
 * src/index.h:25-68.
 */
#ifndef FSX_INDEX_H
#define FSX_INDEX_H

#include <stddef.h>

#include "fsx_mol.h"

#ifdef __cplusplus
extern "C" {
#endif

struct sxs_euler { double z, b1, g1, a2, b2, g2; };  /* index.h:25-33 */
struct sxs_index { int z, b1, g1, a2, b2, g2; };     /* index.h:35-43 */

size_t sxs_assemble_index(struct sxs_index *id, int nbeta, int L);               /* index.c:3-7 */
void sxs_disassemble_index(struct sxs_index *id, size_t index, int nbeta, int L); /* index.c:9-29 */

/* index.c:38-79: Euler coordinates of a dimer pose given the ligand
 * translation tv, its rotation rm about its centre and the reference ligand
 * position ref_lig relative to the receptor centre (z rounded to whole A) */
void sxs_ft2euler(struct sxs_euler *euler, struct mol_vector3 *tv, struct mol_matrix3 *rm,
                  struct mol_vector3 *ref_lig);
/* index.c:81-121: ft-file ("id tx ty tz" + 6 more numbers per line) and
 * rotation list -> Euler file ("%d\t% .3f" x 6 per line).  The rotation
 * list holds one matrix per line, "[index] m11 m12 m13 m21 m22 m23 m31 m32
 * m33"; a malformed line or unreadable file is fatal. */
void sxs_ft_file2euler_file(const char *eu_path, const char *ft_path, const char *rm_path,
                            struct mol_vector3 *ref_lig);

#ifdef __cplusplus
}
#endif
#endif
