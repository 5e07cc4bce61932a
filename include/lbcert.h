/* lbcert.h -- C ABI of the host part of the sparse FP64 certificate (liblbcert.so).
 *
 * Replaces the factorization step of the reference's lifted certificate (factor_price with the cholcert / CHOLMOD
 * backend, certify_lifted_vc.py:391-449, cholcert.c:24-76) and of its verifier's own LDL^T (psd_ldl,
 * verify_cert.py:292-386), in the conventions of cholcert.c: plain C, 64-bit sizes, caller-owned host arrays, int
 * status (0 ok, < 0 usage error, > 0 numerical failure). The dense core of the factorization runs on the device
 * (lb_cert_core in lbhip.h).
 */
#ifndef LBCERT_H
#define LBCERT_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct lbc_fact lbc_fact;

/* Order and factor the sparse part of a symmetric matrix given in full (both triangles and the diagonal) as CSR
 * (ptr N+1, idx, val). order NULL: minimum-degree elimination on the quotient graph (lazy exact degrees), rows of
 * degree above dense_deg never pivoted, elimination stopped when the smallest exact degree exceeds core_deg. order
 * given (a permutation of 0..N-1): elimination in that order, each row whose exact degree exceeds core_deg at its turn
 * deferred. The rows not eliminated are the dense core. Right-looking LDL^T of the sparse pivots. *out receives the
 * factorization object (also on a numerical failure, for lbc_info); free it with lbc_free. Returns 0, -1 (bad
 * arguments) or k > 0 (sparse pivot at position k - 1 not positive). */
int lbc_factor(int64_t N, const int64_t *ptr, const int32_t *idx, const double *val, int64_t dense_deg,
               int64_t core_deg, const int64_t *order, lbc_fact **out);
void lbc_free(lbc_fact *f);
/* The elimination order alone (same rules as lbc_factor) for a symmetric pattern: order (N original ids), *nsparse the
 * number of pivots before the core. */
int lbc_md_order(int64_t N, const int64_t *ptr, const int32_t *idx, int64_t dense_deg, int64_t core_deg,
                 int64_t *order, int64_t *nsparse);
/* iout[9] = {N, sparse pivots, core rows m, nnz of the sparse columns of L, h_max (most products in one entry + 1),
 * largest column count, core entries of A (lower), core terms of the sparse columns, status};
 * dout[3] = {sum |A| over all stored entries, sum_k D_k (sum_i |L_ik|)^2 over the sparse pivots (L_kk = 1 included),
 * smallest sparse pivot}. */
int lbc_info(const lbc_fact *f, int64_t *iout, double *dout);
/* The elimination order (N original row ids: sparse pivots, then the core rows). */
int lbc_order(const lbc_fact *f, int64_t *order);
/* The core's inputs (core row numbers 0..m-1 in the order's tail): the entries of A among core rows, lower triangle
 * (ai >= aj), and per sparse pivot k the core part of its column, CSR (cptr nsp+1, cidx, cval = L_ik) with the
 * pivots cd (nsp). Any pointer may be NULL. */
int lbc_core(const lbc_fact *f, int32_t *ai, int32_t *aj, double *av, int64_t *cptr, int32_t *cidx, double *cval,
             double *cd);

#ifdef __cplusplus
}
#endif
#endif
