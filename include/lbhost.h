// liblbhost: host bookkeeping of the search between device calls (C ABI). The reference does these steps with numpy
// inside its main loop (alm_vcs.py:1735-1758 for the triangle pool, vcsops.py:439-464 `update_ws` / `_merge` for the
// VCS working sets); lagbound_amd.alm calls this library for them. Plain C, 64-bit sizes, caller-owned arrays.
#ifndef LBHOST_H
#define LBHOST_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

// Working-set merge (update_ws, vcsops.py:439-464): the kept entries of a CSR working set (n rows, entries ptr/col,
// keep[e] != 0) and m new entries (ni[k], nR[k]) as one CSR sorted by the key i P + R, in the order of a stable sort
// of [kept entries in storage order, new entries in the given order]. Outputs ptr_out (n + 1), col_out and src_out
// (tot = kept + m entries each): src_out[e] = the old entry index of a kept entry, -1 for a new one. Returns tot, or
// -1 on bad arguments (a row or cluster out of range).
int64_t lbh_ws_merge(int64_t n, int64_t P, const int64_t *ptr, const int32_t *col, const uint8_t *keep, int64_t m,
                     const int64_t *ni, const int64_t *nR, int64_t *ptr_out, int32_t *col_out, int64_t *src_out);

// out[e] = a[src[e]] for src[e] >= 0, 0 for src[e] < 0 (FP64 and byte values). Returns 0, or -1 if src reaches past
// na.
int lbh_take_f64(int64_t cnt, const int64_t *src, const double *a, int64_t na, double *out);
int lbh_take_u8(int64_t cnt, const int64_t *src, const uint8_t *a, int64_t na, uint8_t *out);

// Stable argsorts (the permutation numpy's argsort(kind="stable") returns): of keys in [0, kmax) by counting (-1 if
// a key is out of range), and of any int64 keys by a radix sort. The search's set-up uses them for the degree order,
// the CSR transpose of the relabelled graph and the CSR of the edge keys (instances.py, vertex_order.py).
int lbh_argsort_bounded(int64_t m, const int64_t *keys, int64_t kmax, int64_t *perm_out);
int lbh_argsort_i64(int64_t m, const int64_t *keys, int64_t *perm_out);

// Rows of a CSR (n rows, ptr) taken in `order` (a permutation of 0..n-1), each row's entries kept in their order:
// the new row pointers (nptr_out, n + 1) and, for each entry of the new layout, its entry index in the old one
// (idx_out, ptr[n] - ptr[0] entries). Returns the entry count, or -1 on bad arguments (an order entry out of range,
// or a total that differs from ptr[n] - ptr[0], as a repeated row makes it; idx_out is then not written).
int64_t lbh_row_gather(int64_t n, const int64_t *ptr, const int64_t *order, int64_t *nptr_out, int64_t *idx_out);

// COO triplets (r, c, v; m of them, indices in [0, N)) as a CSR with rows and, inside a row, columns ascending and
// duplicates combined in input order: indptr_out (N + 1), indices_out and, per stored entry, the sum of its values,
// the sum of their absolute values and their count (each output may be null; v null: every value 1). Outputs hold
// up to m entries. Returns the number of stored entries, or -1 on bad arguments. Used for the lifted certificate's
// assembly (with the counts and absolute sums of its rounding price), its symmetric copy and the vertex graph.
int64_t lbh_coo_sum(int64_t N, int64_t m, const int64_t *r, const int64_t *c, const double *v, int64_t *indptr_out,
                    int64_t *indices_out, double *sum_out, double *abs_out, double *cnt_out);

// The full symmetric CSR of a lower-triangular CSR (N rows; columns of a row strictly ascending and at most the row
// index), with no arithmetic: rows and columns ascending, every off-diagonal value stored twice. Outputs hold
// 2 ptr[N] entries at most. Returns the entry count, or -1 if the input is not such a lower triangle.
int64_t lbh_sym_from_lower(int64_t N, const int64_t *ptr, const int64_t *idx, const double *val, int64_t *ptr_out,
                           int64_t *idx_out, double *val_out);

// Order of VCS separation candidates given by descending score sc (m entries): perm_out = the permutation that sorts
// every run of equal scores by (vi, vR, position) ascending and keeps the runs in place -- np.lexsort((vR, vi, -sc)) for
// descending sc. Returns 0, 1 if sc is not descending (nothing written: sort in full), -1 on bad arguments.
int lbh_tie_order(int64_t m, const double *sc, const int64_t *vi, const int64_t *vR, int64_t *perm_out);

// Triangle candidates not in the pool (alm_vcs.py:1751-1758): the packed key (a n + b) n + c of every candidate
// (rows of cand, m x 3) and of every pool triangle (rows of pool, T x 3), in int64 arithmetic with its wrap-around;
// writes to idx_out the index of the first occurrence of every distinct candidate key absent from the pool, in
// ascending key order. Returns the number written.
int64_t lbh_new_triangles(int64_t n, int64_t m, const int64_t *cand, int64_t T, const int64_t *pool, int64_t *idx_out);

#ifdef __cplusplus
}
#endif
#endif
