/* rbh_testing.h - entry points of librbh.so for tests only.
 *
 * No counterpart in the reference.  They run one internal stage of the step
 * on caller data, so that a test can check the stage's edge cases directly
 * (sizes the step rarely produces).  They allocate device memory per call and
 * are not meant for production paths.
 */
#ifndef RBH_TESTING_H
#define RBH_TESTING_H

#include "rbh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The step's device-wide exclusive prefix sum, on a host array: out[i] =
 * in[0] + ... + in[i-1] for i = 0..n (out holds n + 1 ints) and *total =
 * out[n].  in holds n elements of elem_bytes = 1 (unsigned char) or 4 (int).
 * cap (n <= cap <= 2^28) is the capacity the launch is sized for, as in the
 * step: the grid covers cap, and the live length n is read on the device.
 * Below 32768 the scan is one workgroup; from there on it is the single-pass
 * look-back scan (a wait for a predecessor that times out returns RBH_E_HIP). */
int rbh_prefix_sum(rbh_ctx* ctx, const void* in, int elem_bytes, int n, int cap, int* out, int* total);

#ifdef __cplusplus
}
#endif
#endif
