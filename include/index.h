/*
 * index.h -- the mixed-radix grid index of a dimer pose
 * [z][beta1][beta2][alpha2][gamma1][gamma2].  Drop-in for the reference's
 * src/index.h:25-52 (the ft-file -> Euler conversion of index.h:54-68 needs
 * the rotation-matrix reader and is provided by the correlate front end).
 */
#ifndef FSX_INDEX_H
#define FSX_INDEX_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct sxs_euler { double z, b1, g1, a2, b2, g2; };  /* index.h:25-33 */
struct sxs_index { int z, b1, g1, a2, b2, g2; };     /* index.h:35-43 */

size_t sxs_assemble_index(struct sxs_index *id, int nbeta, int L);               /* index.c:3-7 */
void sxs_disassemble_index(struct sxs_index *id, size_t index, int nbeta, int L); /* index.c:9-29 */

#ifdef __cplusplus
}
#endif
#endif
