/* tpr_rocsparse.h -- the vendor comparator slot: rocSPARSE's batched
 * tridiagonal solver (no pivoting) behind the same argument convention as
 * tprh_solve_batched.  Built as its own library (libtpr_rocsparse.so) so
 * the solver library does not depend on rocSPARSE.
 *
 * Counterpart of the reference's cuSPARSE comparator
 * (src/reference_cusparse.cu:113-145, gtsv2_nopivot on one system, and its
 * nstab slot src/numerical_stability.cpp:191-193); rocSPARSE's strided
 * batch form solves a whole batch per call. */
#ifndef TPR_ROCSPARSE_H
#define TPR_ROCSPARSE_H

#ifdef __cplusplus
extern "C" {
#endif

/* nsys systems of n equations back to back in device arrays: sub-diagonal a
 * (a[0] = 0), diagonal b, super-diagonal c (c[n-1] = 0), right-hand side d;
 * the answers go to x (d is not modified).  Asynchronous on `stream`
 * (a hipStream_t, NULL = default stream).  Returns 0 on success, else a
 * rocSPARSE status (> 0) or -1 for a HIP error (see tprh_vendor_last_error). */
int tprh_vendor_gtsv_batch(const float *a, const float *b, const float *c, const float *d, float *x, int n, int nsys,
                           void *stream);
const char *tprh_vendor_last_error(void);
/* rocSPARSE version as major * 100000 + minor * 100 + patch. */
int tprh_vendor_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TPR_ROCSPARSE_H */
