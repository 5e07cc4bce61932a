// liblbverify: the independent verifier's PSD proof (host, C ABI). No counterpart exists as a C interface in the
// reference: its verifier's LDL is a numba function (verify_cert.py:292-386, `_etree`, `_colcounts`, `_ldl`,
// `psd_ldl`), which lagbound_amd.verify calls this library in place of.
#ifndef LBVERIFY_H
#define LBVERIFY_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Up-looking LDL^T of P A P^T (A: N x N symmetric, full CSR ptr/idx/val with both triangles; perm[k] = the row of A
// that becomes row k) with a per-entry running residual bound (verify_cert.py:315-346 states the same bound).
// Outputs: out[0] = sum over the full matrix of the residual bounds (the PSD price before the allowance for its own
// evaluation), out[2] = sum_j colcount_j^2 (the flop measure), out[3] = sum |A|; iout[0] = first failed pivot (1-based,
// 0 if none), iout[1] = nnz(L) (strictly lower), iout[2] = the largest h_k, iout[3] = rows factored. Dout (N values,
// may be null) receives D; Lp_out (N + 1) / Li_out / Lx_out (nnz(L), may be null) receive L by columns (call once with
// nulls to learn nnz(L)). Returns 0 (D > 0), 1 (a pivot not positive or not finite), -1 (bad arguments) or -2
// (nnz(L) above max_nnz; 0 = no limit).
int lbv_ldl(int64_t N, const int64_t *ptr, const int32_t *idx, const double *val, const int64_t *perm,
            int64_t max_nnz, double *Dout, double *out, int64_t *iout, int64_t *Lp_out, int32_t *Li_out,
            double *Lx_out);

#ifdef __cplusplus
}
#endif
#endif
