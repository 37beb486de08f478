/*
 * fsx_mol.h -- the few libmol2 types the scoring path's entry points take
 * (mol2/atom_group.h, mol2/vector.h, mol2/matrix.h as used by
 * src/pdb2spf.c:24-175, src/saxs_utils.c:19-48, src/index.c:38-121 and
 * src/form_factor_table.c:294-339).
 *
 * libmol2 is not part of this library.  struct mol_atom_group here declares
 * only the members the path reads (atom count, coordinates, van der Waals
 * radii, residue and atom names), so code that fills those members compiles
 * unchanged; it is not binary compatible with libmol2's own struct, whose
 * objects must not be passed to this library.
 */
#ifndef FSX_MOL_H
#define FSX_MOL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

struct mol_vector3 {
    double X, Y, Z;
};

struct mol_matrix3 {
    double m11, m12, m13;
    double m21, m22, m23;
    double m31, m32, m33;
};

struct mol_atom_group {
    size_t natoms;
    struct mol_vector3 *coords;   /* [natoms] */
    double *vdw_radius;           /* [natoms] */
    char **residue_name;          /* [natoms], e.g. "ALA" (blank padding allowed) */
    char **atom_name;             /* [natoms], e.g. " CA " */
};

#ifdef __cplusplus
}
#endif
#endif
