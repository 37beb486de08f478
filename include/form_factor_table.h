/*
 * form_factor_table.h -- per-atom-type zero-angle form factors and the
 * (residue, atom) -> type mapping that atom_grp2spf uses.  Drop-in for the
 * parts of the reference's src/form_factor_table.h:11-99 on the scoring path
 * (types, default_ff_table, get_ff, vacuum_ff/dummy_ff, get_vacuum_ff /
 * get_dummy_ff, ff_type_to_string, ff_table_destroy).  The q-dependent
 * coefficient tables (read_ff_table, compute_form_factor,
 * build_form_factor_tables) are not on the path and are not provided.
 */
#ifndef FSX_FORM_FACTOR_TABLE_H
#define FSX_FORM_FACTOR_TABLE_H

#include <stddef.h>
#include "fsx_mol.h"

#ifdef __cplusplus
extern "C" {
#endif

/* form_factor_table.h:26-37 (same enumerators and values) */
enum saxs_ff_type {
    H, He,
    Li, Be, B, C, N, O, F, Ne,
    Na, Mg, Al, Si, P, S, Cl, Ar,
    K, Ca, Cr, Mn, Fe, Co,
    Ni, Cu, Zn, Se, Br,
    Io, Ir, Pt, Au, Hg, SINGLE_ATOM_SIZE = 34,
    CH = 34, CH2 = 35, CH3 = 36, NH = 37, NH2 = 38, NH3 = 39, OH = 40, s_OH2 = 41, SH = 42,
    PO4 = 43,
    HEAVY_ATOM_SIZE = 43, s_UNK = 99
};

struct saxs_form_factor {
    double zero_ff;
    double vacuum_ff;
    double dummy_ff;
    double *values;   /* unused on this path, NULL */
};

/* one (residue, atom) -> type entry of the mapping file */
struct saxs_resi_atom_type {
    char residue_name[8];
    char atom_name[8];
    enum saxs_ff_type type;
};

struct saxs_form_factor_table {
    struct saxs_form_factor factors[HEAVY_ATOM_SIZE];
    struct saxs_resi_atom_type *map;   /* sorted by (residue, atom) */
    size_t nmap;
    struct saxs_form_factor *water_ff;
};

/* The default zero-angle table plus the mapping read from type_mapping_file
 * (lines "RES ATOM TYPE", '#' comments; form_factor_table.c:138-188).  Like
 * the reference, the first table built is cached and returned by every
 * later call whatever its argument; an unreadable file is fatal. */
struct saxs_form_factor_table *default_ff_table(const char *type_mapping_file);
void ff_table_destroy(struct saxs_form_factor_table *table);
const char *ff_type_to_string(enum saxs_ff_type type);

/* form_factor_table.c:294-339: the factors of atom atom_index, looked up by
 * its residue and atom names with surrounding blanks removed; NULL (with a
 * message on stderr) when unmapped or mapped to UNK */
const struct saxs_form_factor *get_ff(const struct saxs_form_factor_table *table,
                                      const struct mol_atom_group *ag, size_t atom_index);
double dummy_ff(const struct saxs_form_factor_table *table, const struct mol_atom_group *ag,
                size_t atom_index);
double vacuum_ff(const struct saxs_form_factor_table *table, const struct mol_atom_group *ag,
                 size_t atom_index);
void get_dummy_ff(double *dummy_ffs, const struct mol_atom_group *ag,
                  const struct saxs_form_factor_table *table);
void get_vacuum_ff(double *vacuum_ffs, const struct mol_atom_group *ag,
                   const struct saxs_form_factor_table *table);

#ifdef __cplusplus
}
#endif
#endif
