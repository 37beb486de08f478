/*
 * fsx_spf.c -- host entry of the spherical-harmonic amplitude expansion of
 * an atom group (src/pdb2spf.c:24-152):
 *   A^X_lm(q) = 4 pi i^l sum_atoms f^X_a j_l(q r_a) conj(Y_lm(theta_a, phi_a))
 * with f^V, f^D the scalar vacuum / excluded-volume form factors and
 * f^W = zero_ff(OH2) * SASA fraction.  The angles, the normalised Legendre
 * values, the Bessel table and the sums over atoms all run on the GPU
 * (fsx_kernels.hip: k_atom_geom, k_atom_bes, k_atom_mfma, k_atom_finish;
 * the mixed form k_atom_bes_split, k_atom_bsplit, k_atom_bf16); this side
 * checks the arguments and forms the mean radius (saxs_utils.c:19-28).
 */
#include "../fsx_internal.h"

#include <stddef.h>

int fsx_atoms2spf(int natoms, const double *xyz, const double *vac, const double *dum,
                  const double *sa, double h2o_ff, const double *vdw, const double *qvals,
                  int qnum, int L, double *re, double *im, double *rm_out)
{
    return fsx_atoms2spf_form(natoms, xyz, vac, dum, sa, h2o_ff, vdw, qvals, qnum, L, 0u, re, im,
                              rm_out);
}

int fsx_atoms2spf_form(int natoms, const double *xyz, const double *vac, const double *dum,
                       const double *sa, double h2o_ff, const double *vdw, const double *qvals,
                       int qnum, int L, unsigned flags, double *re, double *im, double *rm_out)
{
    if (natoms < 1 || !xyz || !vac || !dum || !vdw || !qvals || qnum < 1 || L < 0 || L > 40 ||
        !re || !im)
        return FSX_ERR_ARG;
    double av = 0.0;
    for (int a = 0; a < natoms; a++) av += vdw[a];
    if (rm_out) *rm_out = av / (double)(size_t)natoms;
    return fsx_hip_atoms2spf(natoms, xyz, vac, dum, sa, sa ? h2o_ff : 0.0, qvals, qnum, L, re,
                             im, (flags & FSX_MIXED_FP32) != 0);
}
