/*
 * fsx_front.c -- the atom-group side of the scoring path and the correlate
 * front end, on the host:
 *   form-factor table and (residue, atom) mapping   form_factor_table.c:8-188, 294-339, 509-545
 *   mean radius, fractional SASA, rotations          saxs_utils.c:19-48, 65-94
 *   atom_grp2spf_inplace / atom_grp2spf              pdb2spf.c:24-175 (the sums run on the GPU
 *                                                    through fsx_atoms2spf)
 *   ft-file -> Euler coordinates                     index.c:31-121
 *   Euler list -> grid indices, score file           tools/correlate.c:202-251, 371-381
 */
/*
 * The zero-angle form-factor values of k_zero_ff and the residue/atom type
 * names below reproduce data of libfmftsaxs (src/form_factor_table.c),
 * which is available under this licence:
 *
 *   Copyright (c) 2017 Mikhail Ignatov, Andrey Kazennov and Dima Kozakov
 *
 *   Permission is hereby granted, free of charge, to any person obtaining a
 *   copy of this software and associated documentation files (the
 *   "Software"), to deal in the Software without restriction, including
 *   without limitation the rights to use, copy, modify, merge, publish,
 *   distribute, sublicense, and/or sell copies of the Software, and to
 *   permit persons to whom the Software is furnished to do so, subject to
 *   the following conditions:
 *
 *   The above copyright notice and this permission notice shall be included
 *   in all copies or substantial portions of the Software.
 *
 *   THE SOFTWARE IS PROVIDED "AS IS", WITHOUT WARRANTY OF ANY KIND, EXPRESS
 *   OR IMPLIED, INCLUDING BUT NOT LIMITED TO THE WARRANTIES OF
 *   MERCHANTABILITY, FITNESS FOR A PARTICULAR PURPOSE AND NONINFRINGEMENT.
 *   IN NO EVENT SHALL THE AUTHORS OR COPYRIGHT HOLDERS BE LIABLE FOR ANY
 *   CLAIM, DAMAGES OR OTHER LIABILITY, WHETHER IN AN ACTION OF CONTRACT,
 *   TORT OR OTHERWISE, ARISING FROM, OUT OF OR IN CONNECTION WITH THE
 *   SOFTWARE OR THE USE OR OTHER DEALINGS IN THE SOFTWARE.
 */
#include "../fsx_internal.h"
#include "form_factor_table.h"
#include "index.h"
#include "pdb2spf.h"
#include "saxs_utils.h"

#include <ctype.h>
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* ---- zero-angle form factors per type ----------------------------------
 * The default table of form_factor_table.c:8-49 (values per enum
 * saxs_ff_type, H .. SH); the data must be identical to reproduce the
 * reference's amplitudes. */
static const double k_zero_ff[HEAVY_ATOM_SIZE] = {
    -0.720147, -0.720228,                                                      /* H He */
    1.591, 2.591, 3.591, 0.50824, 6.16294, 4.94998, 7.591, 6.993,              /* Li..Ne */
    7.9864, 8.9805, 9.984, 10.984, 13.0855, 9.36656, 13.984, 16.591,           /* Na..Ar */
    15.984, 14.9965, 20.984, 21.984, 20.9946, 23.984,                          /* K..Co */
    24.984, 25.984, 24.9936, 30.9825, 31.984, 49.16,                           /* Ni..I */
    70.35676, 71.35676, 72.324, 73.35676,                                      /* Ir..Hg */
    -0.211907, -0.932054, -1.6522, 5.44279, 4.72265, 4.0025, 4.22983, 3.50968, /* CH..OH2 */
    8.64641                                                                    /* SH */
};
static const double k_vacuum_ff[HEAVY_ATOM_SIZE] = {
    0.999953, 0.999872, 2.99, 3.99, 4.99, 5.9992, 6.9946, 7.9994, 8.99, 9.999,
    10.9924, 11.9865, 12.99, 13.99, 14.9993, 15.9998, 16.99, 17.99,
    18.99, 18.0025, 23.99, 24.99, 24.0006, 26.99,
    27.99, 28.99, 27.9996, 33.99, 34.99, 52.99, 76.99, 77.99, 78.9572, 79.99,
    6.99915, 7.99911, 8.99906, 7.99455, 8.99451, 9.99446, 8.99935, 9.9993, 16.9998
};
static const double k_dummy_ff[HEAVY_ATOM_SIZE] = {
    1.7201, 1.7201, 1.399, 1.399, 1.399, 5.49096, 0.83166, 3.04942, 1.399, 3.006,
    3.006, 3.006, 3.006, 3.006, 1.91382, 6.63324, 3.006, 1.399,
    3.006, 3.006, 3.006, 3.006, 3.006, 3.006,
    3.006, 3.006, 3.006, 3.006, 3.006, 3.83, 6.63324, 6.63324, 6.63324, 6.63324,
    7.21106, 8.93116, 10.6513, 2.55176, 4.27186, 5.99196, 4.76952, 6.48962, 8.35334
};

/* type names of the mapping file (form_factor_table.c:76-107; the
 * comparison uses the name as written, upper case) */
static const struct {
    const char *name;
    enum saxs_ff_type type;
} k_type_names[] = {
    {"H", H}, {"HE", He}, {"C", C}, {"N", N}, {"O", O}, {"NE", Ne}, {"SOD+", Na},
    {"MG2+", Mg}, {"P", P}, {"S", S}, {"K", K}, {"CAL2+", Ca}, {"FE2+", Fe}, {"ZN2+", Zn},
    {"SE", Se}, {"AU", Au}, {"CH", CH}, {"CH2", CH2}, {"CH3", CH3}, {"NH", NH},
    {"NH2", NH2}, {"NH3", NH3}, {"OH", OH}, {"SH", SH},
};
#define N_TYPE_NAMES (sizeof(k_type_names) / sizeof(k_type_names[0]))

static enum saxs_ff_type type_of_name(const char *name)
{
    for (size_t i = 0; i < N_TYPE_NAMES; i++)
        if (!strcmp(name, k_type_names[i].name)) return k_type_names[i].type;
    return s_UNK;
}

const char *ff_type_to_string(enum saxs_ff_type type)
{
    for (size_t i = 0; i < N_TYPE_NAMES; i++)
        if (k_type_names[i].type == type) return k_type_names[i].name;
    return NULL;
}

static int cmp_key(const char *r1, const char *a1, const char *r2, const char *a2)
{
    int c = strncmp(r1, r2, 7);
    return c ? c : strncmp(a1, a2, 7);
}

static int cmp_entry(const void *x, const void *y)
{
    const struct saxs_resi_atom_type *a = x, *b = y;
    return cmp_key(a->residue_name, a->atom_name, b->residue_name, b->atom_name);
}

static int blank_line(const char *s)
{
    for (; *s; s++)
        if (!isspace((unsigned char)*s)) return 0;
    return 1;
}

static pthread_mutex_t g_ff_mtx = PTHREAD_MUTEX_INITIALIZER;
static struct saxs_form_factor_table *g_ff_default = NULL;

struct saxs_form_factor_table *default_ff_table(const char *type_mapping_file)
{
    pthread_mutex_lock(&g_ff_mtx);
    if (g_ff_default) {
        pthread_mutex_unlock(&g_ff_mtx);
        return g_ff_default;
    }
    FILE *fp = fopen(type_mapping_file, "r");
    if (!fp) {
        fprintf(stderr, "Could not open file %s\n", type_mapping_file);
        exit(1);
    }
    struct saxs_form_factor_table *t = calloc(1, sizeof(*t));
    if (!t) {
        fprintf(stderr, "[Error] default_ff_table: out of memory\n");
        exit(EXIT_FAILURE);
    }
    for (int i = 0; i < HEAVY_ATOM_SIZE; i++) {
        t->factors[i].zero_ff = k_zero_ff[i];
        t->factors[i].vacuum_ff = k_vacuum_ff[i];
        t->factors[i].dummy_ff = k_dummy_ff[i];
        t->factors[i].values = NULL;
    }
    size_t cap = 256;
    t->map = malloc(cap * sizeof(*t->map));
    char *line = NULL;
    size_t len = 0;
    while (getline(&line, &len, fp) != -1) {
        if (line[0] == '#' || blank_line(line)) continue;
        char res[8] = {0}, atom[8] = {0}, tname[16] = {0};
        if (sscanf(line, "%7s %7s %15s", res, atom, tname) < 3) continue;
        enum saxs_ff_type type = type_of_name(tname);
        /* a repeated key keeps the last type, as a hash put would */
        size_t k;
        for (k = 0; k < t->nmap; k++)
            if (!cmp_key(t->map[k].residue_name, t->map[k].atom_name, res, atom)) break;
        if (k == t->nmap) {
            if (t->nmap == cap) {
                cap *= 2;
                t->map = realloc(t->map, cap * sizeof(*t->map));
            }
            if (!t->map) {
                fprintf(stderr, "[Error] default_ff_table: out of memory\n");
                exit(EXIT_FAILURE);
            }
            memcpy(t->map[k].residue_name, res, 8);
            memcpy(t->map[k].atom_name, atom, 8);
            t->nmap++;
        }
        t->map[k].type = type;
    }
    free(line);
    fclose(fp);
    qsort(t->map, t->nmap, sizeof(*t->map), cmp_entry);
    t->water_ff = NULL;
    g_ff_default = t;
    pthread_mutex_unlock(&g_ff_mtx);
    return t;
}

void ff_table_destroy(struct saxs_form_factor_table *table)
{
    if (!table) return;
    pthread_mutex_lock(&g_ff_mtx);
    if (table == g_ff_default) g_ff_default = NULL;
    pthread_mutex_unlock(&g_ff_mtx);
    free(table->map);
    free(table);
}

/* the name without leading blanks, up to the next blank (at most 7 chars) */
static void trim_name(const char *src, char *dst)
{
    memset(dst, 0, 8);
    if (!src) return;
    while (*src == ' ') src++;
    for (int i = 0; i < 7 && *src && *src != ' '; i++) dst[i] = *src++;
}

const struct saxs_form_factor *get_ff(const struct saxs_form_factor_table *table,
                                      const struct mol_atom_group *ag, size_t atom_index)
{
    struct saxs_resi_atom_type key;
    trim_name(ag->residue_name[atom_index], key.residue_name);
    trim_name(ag->atom_name[atom_index], key.atom_name);
    const struct saxs_resi_atom_type *hit =
        bsearch(&key, table->map, table->nmap, sizeof(key), cmp_entry);
    if (!hit) {
        fprintf(stderr, "mapping for (%s, %s) not found\n", key.residue_name, key.atom_name);
        return NULL;
    }
    if (hit->type == s_UNK) {
        printf("(%s, %s) maps to UNK\n", key.residue_name, key.atom_name);
        return NULL;
    }
    return &table->factors[hit->type];
}

double dummy_ff(const struct saxs_form_factor_table *table, const struct mol_atom_group *ag,
                size_t atom_index)
{
    const struct saxs_form_factor *f = get_ff(table, ag, atom_index);
    return f ? f->dummy_ff : 0.0;
}

double vacuum_ff(const struct saxs_form_factor_table *table, const struct mol_atom_group *ag,
                 size_t atom_index)
{
    const struct saxs_form_factor *f = get_ff(table, ag, atom_index);
    return f ? f->vacuum_ff : 0.0;
}

void get_dummy_ff(double *out, const struct mol_atom_group *ag,
                  const struct saxs_form_factor_table *table)
{
    for (size_t i = 0; i < ag->natoms; i++) out[i] = dummy_ff(table, ag, i);
}

void get_vacuum_ff(double *out, const struct mol_atom_group *ag,
                   const struct saxs_form_factor_table *table)
{
    for (size_t i = 0; i < ag->natoms; i++) out[i] = vacuum_ff(table, ag, i);
}

/* ---- saxs_utils -------------------------------------------------------- */
double mol_atom_group_average_radius(const struct mol_atom_group *ag)
{
    double r = 0.0;
    for (size_t i = 0; i < ag->natoms; i++) r += ag->vdw_radius[i];
    return r / ag->natoms;
}

/* Shrake-Rupley: points of a Fibonacci sphere on radius r_vdw + r_solv,
 * a point is buried when it lies strictly inside another atom's expanded
 * sphere; neighbours come from a cell list of edge 2 max(r_vdw + r_solv). */
void sxs_faccs(double *frac, const struct mol_atom_group *ag, double r_solv)
{
    const size_t n = ag->natoms;
    const int np = FSX_SASA_POINTS;
    if (n == 0) return;
    double *sph = malloc(sizeof(double) * 3 * np);
    double *R = malloc(sizeof(double) * n);
    long *key = malloc(sizeof(long) * 3 * n);
    size_t *nb = malloc(sizeof(size_t) * n);
    if (!sph || !R || !key || !nb) {
        fprintf(stderr, "[Error] sxs_faccs: out of memory\n");
        exit(EXIT_FAILURE);
    }
    const double golden = 3.14159265358979323846 * (1.0 + sqrt(5.0));
    for (int k = 0; k < np; k++) {
        double kk = k + 0.5;
        double phi = acos(1.0 - 2.0 * kk / np);
        double th = golden * kk;
        sph[3 * k] = cos(th) * sin(phi);
        sph[3 * k + 1] = sin(th) * sin(phi);
        sph[3 * k + 2] = cos(phi);
    }
    double rmax = 0.0;
    for (size_t i = 0; i < n; i++) {
        R[i] = ag->vdw_radius[i] + r_solv;
        if (R[i] > rmax) rmax = R[i];
    }
    const double cell = 2.0 * rmax;
    for (size_t i = 0; i < n; i++) {
        key[3 * i] = (long)floor(ag->coords[i].X / cell);
        key[3 * i + 1] = (long)floor(ag->coords[i].Y / cell);
        key[3 * i + 2] = (long)floor(ag->coords[i].Z / cell);
    }
    for (size_t i = 0; i < n; i++) {
        size_t nn = 0;
        for (size_t j = 0; j < n; j++) {
            if (j == i) continue;
            if (labs(key[3 * j] - key[3 * i]) <= 1 && labs(key[3 * j + 1] - key[3 * i + 1]) <= 1
                && labs(key[3 * j + 2] - key[3 * i + 2]) <= 1)
                nb[nn++] = j;
        }
        int free_pts = 0;
        for (int k = 0; k < np; k++) {
            double px = ag->coords[i].X + R[i] * sph[3 * k];
            double py = ag->coords[i].Y + R[i] * sph[3 * k + 1];
            double pz = ag->coords[i].Z + R[i] * sph[3 * k + 2];
            int buried = 0;
            for (size_t t = 0; t < nn && !buried; t++) {
                size_t j = nb[t];
                double dx = px - ag->coords[j].X, dy = py - ag->coords[j].Y;
                double dz = pz - ag->coords[j].Z;
                buried = dx * dx + dy * dy + dz * dz < R[j] * R[j];
            }
            free_pts += !buried;
        }
        frac[i] = ag->vdw_radius[i] == 0.0 ? 0.0 : (double)free_pts / np;
    }
    free(nb);
    free(key);
    free(R);
    free(sph);
}

void sxs_fill_active_rotation_matrix(struct mol_matrix3 *rm, double alpha, double beta,
                                     double gamma)
{
    double ca = cos(alpha), sa = sin(alpha), cb = cos(beta), sb = sin(beta);
    double cg = cos(gamma), sg = sin(gamma);
    rm->m11 = cg * cb * ca - sg * sa;
    rm->m21 = cg * cb * sa + sg * ca;
    rm->m31 = -cg * sb;
    rm->m12 = -sg * cb * ca - cg * sa;
    rm->m22 = -sg * cb * sa + cg * ca;
    rm->m32 = sg * sb;
    rm->m13 = sb * ca;
    rm->m23 = sb * sa;
    rm->m33 = cb;
}

void sxs_mult_rot_mats(struct mol_matrix3 *c, struct mol_matrix3 *a, struct mol_matrix3 *b)
{
    struct mol_matrix3 r;
    r.m11 = a->m11 * b->m11 + a->m12 * b->m21 + a->m13 * b->m31;
    r.m12 = a->m11 * b->m12 + a->m12 * b->m22 + a->m13 * b->m32;
    r.m13 = a->m11 * b->m13 + a->m12 * b->m23 + a->m13 * b->m33;
    r.m21 = a->m21 * b->m11 + a->m22 * b->m21 + a->m23 * b->m31;
    r.m22 = a->m21 * b->m12 + a->m22 * b->m22 + a->m23 * b->m32;
    r.m23 = a->m21 * b->m13 + a->m22 * b->m23 + a->m23 * b->m33;
    r.m31 = a->m31 * b->m11 + a->m32 * b->m21 + a->m33 * b->m31;
    r.m32 = a->m31 * b->m12 + a->m32 * b->m22 + a->m33 * b->m32;
    r.m33 = a->m31 * b->m13 + a->m32 * b->m23 + a->m33 * b->m33;
    *c = r;
}

/* ---- atom group -> SPF (pdb2spf.c:24-175) -------------------------------- */
void atom_grp2spf_inplace(struct sxs_spf_full *spf, struct mol_atom_group *ag,
                          struct saxs_form_factor_table *ff_table, double *qvals, int qnum,
                          int L, double *saxs_sa)
{
    if (!spf || !ag || !ff_table || !qvals || qnum < 1 || L < 0 || spf->L < L
        || spf->qnum < qnum) {
        fprintf(stderr, "[Error] atom_grp2spf_inplace: invalid arguments\n");
        exit(EXIT_FAILURE);
    }
    sxs_spf_full_reset(spf);
    spf->rm = mol_atom_group_average_radius(ag);
    const int n = (int)ag->natoms, K2 = (L + 1) * (L + 1);
    double h2o_ff = saxs_sa ? ff_table->factors[s_OH2].zero_ff : 0.0;
    double *xyz = malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    double *vac = malloc(sizeof(double) * (n > 0 ? n : 1));
    double *dum = malloc(sizeof(double) * (n > 0 ? n : 1));
    double *re = malloc(sizeof(double) * 3 * (size_t)qnum * K2);
    double *im = malloc(sizeof(double) * 3 * (size_t)qnum * K2);
    if (!xyz || !vac || !dum || !re || !im) {
        fprintf(stderr, "[Error] atom_grp2spf_inplace: out of memory\n");
        exit(EXIT_FAILURE);
    }
    for (int i = 0; i < n; i++) {
        const struct saxs_form_factor *ff = get_ff(ff_table, ag, i);
        if (!ff) {
            fprintf(stderr, "[Error] atom_grp2spf_inplace: no form factor for atom %d\n", i);
            exit(EXIT_FAILURE);
        }
        xyz[3 * i] = ag->coords[i].X;
        xyz[3 * i + 1] = ag->coords[i].Y;
        xyz[3 * i + 2] = ag->coords[i].Z;
        vac[i] = ff->vacuum_ff;
        dum[i] = ff->dummy_ff;
    }
    double rm = 0.0;
    int rc = fsx_atoms2spf(n, xyz, vac, dum, saxs_sa, h2o_ff, ag->vdw_radius, qvals, qnum, L,
                           re, im, &rm);
    if (rc != FSX_OK) {
        fprintf(stderr, "[Error] atom_grp2spf_inplace: %s\n", fsx_error_string(rc));
        exit(EXIT_FAILURE);
    }
    size_t blk = (size_t)qnum * K2;
    for (int q = 0; q < qnum; q++) {
        struct sxs_spf_sing *x[3] = {spf->V[q], spf->D[q], spf->W[q]};
        for (int X = 0; X < 3; X++)
            for (int lm = 0; lm < K2; lm++) {
                x[X]->re[lm] = re[X * blk + (size_t)q * K2 + lm];
                x[X]->im[lm] = im[X * blk + (size_t)q * K2 + lm];
            }
    }
    free(xyz);
    free(vac);
    free(dum);
    free(re);
    free(im);
}

struct sxs_spf_full *atom_grp2spf(struct mol_atom_group *ag,
                                  struct saxs_form_factor_table *ff_table, double *qvals,
                                  int qnum, int L, int water)
{
    struct sxs_spf_full *spf = sxs_spf_full_create(L, qnum);
    double *sa = NULL;
    if (water == 1) {
        sa = calloc(ag->natoms ? ag->natoms : 1, sizeof(double));
        if (!sa) {
            fprintf(stderr, "[Error] atom_grp2spf: out of memory\n");
            exit(EXIT_FAILURE);
        }
        sxs_faccs(sa, ag, 1.4);
    }
    atom_grp2spf_inplace(spf, ag, ff_table, qvals, qnum, L, sa);
    free(sa);
    return spf;
}

/* ---- ft-file -> Euler coordinates (index.c:31-121) ------------------------ */
static double clamp1(double a)
{
    if (a < -1.0) return -1.0;
    if (a > 1.0) return 1.0;
    return a;
}

void sxs_ft2euler(struct sxs_euler *euler, struct mol_vector3 *tv, struct mol_matrix3 *rm,
                  struct mol_vector3 *ref_lig)
{
    const double pi = FSX_PI_REFERENCE;
    struct mol_vector3 v;
    v.X = tv->X + ref_lig->X;
    v.Y = tv->Y + ref_lig->Y;
    v.Z = tv->Z + ref_lig->Z;
    double z = round(sqrt(v.X * v.X + v.Y * v.Y + v.Z * v.Z));
    double b1 = acos(clamp1(v.Z / z));
    double g1 = acos(clamp1(-v.X / (z * sin(b1))));
    if (v.Y / (z * sin(b1)) < 0.0) g1 = 2 * pi - g1;
    struct mol_matrix3 rec, lig;
    sxs_fill_active_rotation_matrix(&rec, 0.0, b1, g1);
    sxs_mult_rot_mats(&lig, &rec, rm);
    double b2 = acos(clamp1(lig.m33));
    double a2 = acos(clamp1(lig.m13 / sin(b2)));
    if (lig.m23 / sin(b2) < 0.0) a2 = 2 * pi - a2;
    double g2 = acos(clamp1(-lig.m31 / sin(b2)));
    if (lig.m32 / sin(b2) < 0.0) g2 = 2 * pi - g2;
    euler->z = z;
    euler->b1 = b1;
    euler->g1 = g1;
    euler->a2 = a2;
    euler->b2 = b2;
    euler->g2 = g2;
}

/* Rotation list: one matrix per non-comment line, "[index] m11 m12 m13 m21
 * m22 m23 m31 m32 m33" (a leading integer index is skipped); the i-th
 * matrix is rotation i of the ft-file. */
static struct mol_matrix3 *read_rotations(const char *path, int *count)
{
    FILE *fp = fopen(path, "r");
    if (!fp) return NULL;
    int cap = 1024, n = 0;
    struct mol_matrix3 *m = malloc(sizeof(*m) * cap);
    char *line = NULL;
    size_t len = 0;
    while (m && getline(&line, &len, fp) != -1) {
        if (line[0] == '#' || blank_line(line)) continue;
        double v[10];
        int k = sscanf(line, "%lf %lf %lf %lf %lf %lf %lf %lf %lf %lf", &v[0], &v[1], &v[2],
                       &v[3], &v[4], &v[5], &v[6], &v[7], &v[8], &v[9]);
        const double *a;
        if (k == 10) a = v + 1;
        else if (k == 9) a = v;
        else continue;
        if (n == cap) {
            cap *= 2;
            m = realloc(m, sizeof(*m) * cap);
            if (!m) break;
        }
        m[n].m11 = a[0]; m[n].m12 = a[1]; m[n].m13 = a[2];
        m[n].m21 = a[3]; m[n].m22 = a[4]; m[n].m23 = a[5];
        m[n].m31 = a[6]; m[n].m32 = a[7]; m[n].m33 = a[8];
        n++;
    }
    free(line);
    fclose(fp);
    *count = n;
    return m;
}

void sxs_ft_file2euler_file(const char *eu_path, const char *ft_path, const char *rm_path,
                            struct mol_vector3 *ref_lig)
{
    FILE *ft = fopen(ft_path, "r");
    FILE *eu = ft ? fopen(eu_path, "w") : NULL;
    int nrot = 0;
    struct mol_matrix3 *rot = eu ? read_rotations(rm_path, &nrot) : NULL;
    if (!ft || !eu || !rot) {
        fprintf(stderr, "[Error] sxs_ft_file2euler_file: cannot open %s\n",
                !ft ? ft_path : (!eu ? eu_path : rm_path));
        exit(EXIT_FAILURE);
    }
    int id;
    double f[9];
    int nread;
    while ((nread = fscanf(ft, "%d %lf %lf %lf %lf %lf %lf %lf %lf %lf\n", &id, &f[0], &f[1],
                           &f[2], &f[3], &f[4], &f[5], &f[6], &f[7], &f[8])) != EOF) {
        if (nread != 10 || id < 0 || id >= nrot) {
            fprintf(stderr, "[Error] Wrong input file format.\n");
            exit(EXIT_FAILURE);
        }
        struct mol_vector3 tv = {f[0], f[1], f[2]};
        struct sxs_euler e;
        sxs_ft2euler(&e, &tv, &rot[id], ref_lig);
        fprintf(eu, "%d\t% .3f\t% .3f\t% .3f\t% .3f\t% .3f\t% .3f\n", id, e.z, e.b1, e.g1, e.a2,
                e.b2, e.g2);
    }
    free(rot);
    fclose(eu);
    fclose(ft);
}

/* ---- Euler list -> grid indices (tools/correlate.c:202-251) -------------- */
int fsx_euler_list_read(const char *path, const double *zvals, int znum, int L, int nbeta,
                        int nalpha, int ngamma, double pi, int64_t **index, int **ft_id,
                        int **order, int *n)
{
    if (!path || !zvals || znum < 1 || L < 1 || nbeta < 2 || nalpha < 1 || ngamma < 1 || !index
        || !ft_id || !order || !n)
        return FSX_ERR_ARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return FSX_ERR_ARG;
    const double b_step = pi / (nbeta - 1);
    const double a_step = 2.0 * pi / nalpha, g_step = 2.0 * pi / ngamma;
    int cap = 1024, cnt = 0, line = 0;
    int64_t *ix = malloc(sizeof(int64_t) * cap);
    int *id = malloc(sizeof(int) * cap), *ord = malloc(sizeof(int) * cap);
    int fid;
    struct sxs_euler e;
    int rc = FSX_OK;
    while (ix && id && ord
           && fscanf(fp, "%d %lf %lf %lf %lf %lf %lf", &fid, &e.z, &e.b1, &e.g1, &e.a2, &e.b2,
                     &e.g2) == 7) {
        for (int j = 0; j < znum; j++) {
            if (!(zvals[j] > e.z - 0.001 && zvals[j] < e.z + 0.001)) continue;
            double a2 = 2 * pi - e.a2, g2 = 2 * pi - e.g2;
            /* digits are not reduced: a rounded angle of a full turn carries
             * into the next digit, as in the reference */
            int64_t t = j;
            t = t * nbeta + (int64_t)round(e.b1 / b_step);
            t = t * nbeta + (int64_t)round(e.b2 / b_step);
            t = t * nalpha + (int64_t)round(a2 / a_step);
            t = t * ngamma + (int64_t)round(e.g1 / g_step);
            t = t * ngamma + (int64_t)round(g2 / g_step);
            if (cnt == cap) {
                cap *= 2;
                ix = realloc(ix, sizeof(int64_t) * cap);
                id = realloc(id, sizeof(int) * cap);
                ord = realloc(ord, sizeof(int) * cap);
                if (!ix || !id || !ord) break;
            }
            ix[cnt] = t;
            id[cnt] = fid;
            ord[cnt] = line;
            cnt++;
        }
        line++;
    }
    fclose(fp);
    if (!ix || !id || !ord) {
        free(ix);
        free(id);
        free(ord);
        return FSX_ERR_NOMEM;
    }
    *index = ix;
    *ft_id = id;
    *order = ord;
    *n = cnt;
    return rc;
}

/* tools/correlate.c:371-381 */
int fsx_scores_write(const char *path, const int *order, const int *ft_id, const double *score,
                     const double *c1, const double *c2, int n)
{
    FILE *fp = fopen(path, "w");
    if (!fp) return FSX_ERR_ARG;
    for (int i = 0; i < n; i++)
        fprintf(fp, "%-6d\t%d\t%.3lf\t%.3lf\t%.3lf\n", order[i], ft_id[i], score[i], c1[i], c2[i]);
    fclose(fp);
    return FSX_OK;
}

void fsx_free(void *p) { free(p); }
