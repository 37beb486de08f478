/*
 * fsx_pose.h -- the pose-index decode of the bucketing kernel, shared by the
 * HIP side and host C (so that it can be exercised on the CPU).
 *
 * Pose index: mixed radix [z][b1][b2][a2][g1][g2] with radices (znum,
 * nbeta, nbeta, nalpha, ngamma, ngamma): the encoding of
 * sxs_assemble_index / sxs_disassemble_index (src/index.c:3-29) with the
 * generalised grid sizes of fsx_score.
 */
#ifndef FSX_POSE_H
#define FSX_POSE_H

#ifdef __HIPCC__
#define FSX_POSE_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define FSX_POSE_HD static inline
#endif

/* Digits of a non-negative pose index into o[0..5] = (z, b1, b2, a2, g1, g2).
 * Returns the z quotient as a 64-bit value: an index far past the grid has a
 * z quotient beyond int range, and only a quotient below znum is a scored
 * pose.  o[0] is that quotient when it is below znum and -1 otherwise, so
 * every digit written is within its radix. */
FSX_POSE_HD long long fsx_decode_pose(long long idx, int znum, int nbeta, int na, int ng, int *o)
{
    o[5] = (int)(idx % ng); idx /= ng;
    o[4] = (int)(idx % ng); idx /= ng;
    o[3] = (int)(idx % na); idx /= na;
    o[2] = (int)(idx % nbeta); idx /= nbeta;
    o[1] = (int)(idx % nbeta);
    const long long zz = idx / nbeta;
    o[0] = zz < znum ? (int)zz : -1;
    return zz;
}

/* Bucket key (z, b2, g2) of a pose and its receptor row b1 * ng + g1 and
 * alpha digit, or -1 for a pose outside the z range (negative indices
 * included): the pose is then not scored and its outputs are not
 * written. */
FSX_POSE_HD int fsx_pose_key(long long idx, int znum, int nbeta, int na, int ng, int *row,
                             int *alpha)
{
    *row = 0;
    *alpha = 0;
    if (idx < 0) return -1;
    int o[6];
    if (fsx_decode_pose(idx, znum, nbeta, na, ng, o) >= znum) return -1;
    *row = o[1] * ng + o[4];
    *alpha = o[3];
    return (o[0] * nbeta + o[2]) * ng + o[5];
}

#endif
