/*
 * sfbessel.h -- spherical Bessel j_l(x) exactly as the reference defines it:
 * truncated Taylor series stopped at a 1e-5 relative term, times x^l
 * (src/sfbessel.h:19, sfbessel.c:40-60).
 */
#ifndef FSX_SFBESSEL_H
#define FSX_SFBESSEL_H
#ifdef __cplusplus
extern "C" {
#endif
double sxs_sbessel(int order, double x);
/* saxs_utils.c:50-63: qnum values from begin to end by cumulative addition */
double *sxs_mkarray(double begin, double end, int qnum);
#ifdef __cplusplus
}
#endif
#endif
