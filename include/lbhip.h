/* lbhip.h -- C ABI of the MI355X (gfx950) ALM search library (liblbhip.so).
 *
 * Replaces the numba kernels that lagbound's solver stage runs on the CPU (reference: src/lagbound/_core/alm_vcs.py,
 * aggops.py, vcsops.py, est_dual_vcs.py). The reference has no C ABI for this path: its only FFI precedent is
 * cholcert.c (src/lagbound/_core/cholcert.c:24-76, bound with ctypes at certify_lifted_vc.py:395-408), whose
 * conventions this header keeps: plain C, 64-bit sizes, caller-owned host arrays borrowed for the call only, an int
 * status (0 ok, < 0 usage / allocation / HIP error, > 0 numerical), no exceptions across the ABI.
 *
 * A context owns two device streams (the second runs the evaluation's hub-segment kernel beside the cluster-sum and
 * VCS kernels, joined by events) and every device buffer (graph, levels, working sets, triangle pool, V, G,
 * multipliers, scratch). V and the multipliers stay resident on the device between calls. Calls on one context are
 * serialised by the caller (one context per host thread); several contexts may share a device.
 *
 * Storage types of V (`vdtype`): LB_F32 (FP32 V, FP32 arithmetic, FP64 accumulation of every global sum),
 * LB_BF16 (BF16-stored V, FP32 arithmetic, FP64 accumulation), LB_F64 (FP64 throughout).
 */
#ifndef LBHIP_H
#define LBHIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define LB_F32 0
#define LB_BF16 1
#define LB_F64 2

#define LB_NPARTS 12
/* parts[] written by lb_eval_fgrad (same layout as the FP64 oracle):
 * [0] f  [1] quadA (rho > 0) or sum_i G_i.V_i (rho = 0)  [2] |d^T V|^2 / L  [3] sum of pair-level PHR terms
 * [4] triangle PHR term  [5] VCS PHR term  [6] primal <B, V V^T> + C0  [7] C0  [8] #(h_QR < -1e-9)
 * [9] min triangle slack 1 - tau  [10] #(g_iR < -1e-9)  [11] sum(lam) */

typedef struct lb_ctx lb_ctx;

/* Create a context on HIP device `device` for rank `r` (8, 16 or 32) embeddings stored as `vdtype`.
 * (No reference counterpart: the reference keeps everything in one numba process, alm_vcs.py:1103-1117.) */
int lb_create(lb_ctx **out, int device, int vdtype, int r);
void lb_destroy(lb_ctx *ctx);
/* Last error message of the context (static text owned by the context). */
const char *lb_last_error(const lb_ctx *ctx);

/* Graph: symmetric CSR with sorted rows (rowptr int64 n+1, col int32), optional FP64 edge weights (NULL =
 * unweighted), vertex volumes d (FP64), L = original 2m, C0 = 2 * inside edges.
 * Replaces the instance load of alm_vcs.py:1103-1117 (A, d, L, C0). */
int lb_set_graph(lb_ctx *ctx, int64_t n, const int64_t *rowptr, const int32_t *col, const double *weights,
                 const double *d, double L, double C0);

/* Pair levels (aggops.PairLevel, aggops.py:51-68; rq, r: alm_vcs.py:1254-1259): labels (nlev x n, int32, values in
 * [0, P_l)), P (nlev), sig (concatenated P_l x P_l FP64, symmetric), rq and dsq (concatenated P_l FP64),
 * pairlvl_weight w (r_QR = rq_Q rq_R w). */
int lb_set_levels(lb_ctx *ctx, int nlev, const int32_t *labs, const int32_t *P, const double *sig,
                  const double *rq, const double *dsq, double pairlvl_weight);
/* Replace the pair-level multipliers (concatenated P_l^2). */
int lb_set_level_sig(lb_ctx *ctx, const double *sig);

/* Sparse vertex-cluster levels (vcsops.VCSLevel, vcsops.py:31-70): labels (nvcs x n), P (nvcs), rq (concatenated),
 * working sets as row CSR: ptr (nvcs x (n+1), each starting at 0), col (concatenated, sorted per row), sig
 * (concatenated FP64); wv = vc_weight * pairlvl_weight (r_iR = wv rq_R / d_i^2). */
int lb_set_vcs(lb_ctx *ctx, int nvcs, const int32_t *labs, const int32_t *P, const double *rq, const int64_t *ptr,
               const int32_t *col, const double *sig, double wv);
int lb_set_vcs_sig(lb_ctx *ctx, const double *sig);
/* VCS entry ages (consecutive outers with multiplier 0 and slack, uint8 per entry, vcsops.py:300-305): set / read
 * together with the multipliers. lb_phr_update advances the ages from the last evaluation's slack flags. */
int lb_set_vcs_age(lb_ctx *ctx, const uint8_t *age);
int lb_get_vcs_state(lb_ctx *ctx, double *sig, uint8_t *age);
/* Change the pair weight (r_QR = rq_Q rq_R w) and the VCS weight wv = vc_weight * w everywhere they enter
 * (set_pair_weight, alm_vcs.py:1270-1275). */
int lb_set_weights(lb_ctx *ctx, double pairlvl_weight, double wv);

/* Triangle pool: TI (T x 3 int32, (a, centre, c)), multipliers lam (FP64). Builds the vertex -> (t, role)
 * incidence in t order (alm_vcs.py:94-104) and the incidence-ordered records (inc_abc / inc_lam,
 * alm_vcs.py:353-360) on the device side. */
int lb_set_triangles(lb_ctx *ctx, int64_t T, const int32_t *TI, const double *lam);
int lb_set_lam(lb_ctx *ctx, const double *lam);

/* V (n x r row-major) from / to host FP64 or FP32 (`host_dtype` LB_F64 / LB_F32). With LB_BF16 storage lb_set_V
 * rounds to nearest even and lb_get_V returns the stored (dequantised) values. */
int lb_set_V(lb_ctx *ctx, const void *V, int host_dtype);
int lb_get_V(lb_ctx *ctx, void *V, int host_dtype);
/* Row order of the V / G transfers: device row k is the caller's row order[k] (a permutation of 0..n-1), applied by
 * lb_set_V, lb_get_V and lb_get_G in their host copy loops; NULL restores the identity, and lb_set_graph clears
 * it. For a graph uploaded with its vertices renumbered (new id k = old vertex order[k]), the caller then keeps V
 * and G in the old numbering. Refused on batch contexts. No reference counterpart (the reference keeps the given
 * numbering). */
int lb_set_row_order(lb_ctx *ctx, const int64_t *order);

/* One Lagrangian f+grad evaluation at the resident V and multipliers (the metric's unit of work):
 *   rho > 0: f = augmented(V, rho) (alm_vcs.py:1496-1533) with rt = rho * tri_weight;
 *   rho = 0: f = C0 + sum(lam) + sum_i (C V)_i.V_i (est_dual_vcs.value, est_dual_vcs.py:245-246);
 *   G = C(y) V (CV, alm_vcs.py:1299-1312) with y the PHR effective multipliers (pp, pt, pv), kept on the device.
 * parts (LB_NPARTS doubles, host) may be NULL to leave the scalars on the device (asynchronous call). */
int lb_eval_fgrad(lb_ctx *ctx, double rho, double rt, double *parts);
/* Copy G (n x r) to host as FP64 or FP32. */
int lb_get_G(lb_ctx *ctx, void *G, int host_dtype);
/* Copy the effective multipliers of the last evaluation: pp (concatenated P_l^2), pt (T), pv (concatenated nnz);
 * any pointer may be NULL. */
int lb_get_multipliers(lb_ctx *ctx, double *pp, double *pt, double *pv);

/* PHR multiplier update on the device: sig <- pp, lam <- pt, VCS sig <- pv of the last rho > 0 evaluation, and the
 * VCS ages from its slack flags (alm_vcs.py:1704-1711,1723; vcsops.age_update). */
int lb_phr_update(lb_ctx *ctx);

/* `nsw` sweeps of batched Gauss-Seidel row steps on the resident V (sweeps + batch_step6, alm_vcs.py:1360-1436,
 * 255-351; pair_all alm_vcs.py:378-415; kq_level alm_vcs.py:417-428; VCS pre/post vcsops.py:127-213) at penalty
 * rho, triangle penalty rt = rho * tri_weight, damping damp and VCS curvature factor kap2 (vc_kap2). orders holds
 * nsw permutations of 0..n-1 (int64, e.g. numpy default_rng(seed).permutation(n) per sweep, alm_vcs.py:1387);
 * the batches are np.array_split(order, nbatch) (alm_vcs.py:1408). Every row of a batch steps from the V and
 * aggregates of the batch start (the reference's rows read each other's fresh values within a batch). */
int lb_sweeps(lb_ctx *ctx, int nsw, double rho, double rt, double damp, double kap2, const int64_t *orders,
              int nbatch);

/* Triangle separation on the resident V (separate_ref fam 0, wedges (a, centre, c); separate1_ref fam 1, one-edge
 * arm (h, centre, k); alm_vcs.py:965-984, kernels 554-666): arms = neighbours with X_ij > thresh (top maxh by value),
 * violation vm = X_aj + X_cj - X_ac - 1 in FP64 summed in the reference's order, a per-chunk (nchunk chunks of
 * centres) nb-bin histogram of vm > vtol, threshold vmin of the first bin whose tail holds at most cap candidates,
 * and the candidates vm >= vmin in the reference's generation order. TI_out: cap x 3 int64 (min, centre, max),
 * val_out: cap; *count = number written; *vmin = threshold (-1 if none). Bit-exact with the reference on the same V;
 * tie rule at the top-maxh / top-maxk cuts: descending value, ties in prior order. fam 1 needs hub2 * cap2 <= 2048. */
int lb_separate_triangles(lb_ctx *ctx, int fam, double thresh, int64_t maxh, int64_t maxk, int64_t hub2,
                          int64_t cap2, double vtol, int64_t cap, int nchunk, int nb, int64_t *TI_out,
                          double *val_out, int64_t *count, double *vmin);

/* VCS separation of level m on the resident V (vcsops.separate, vcsops.py:395-436): the nadd best violated
 * vertex-cluster entries outside the working set (mean-X violation s / vol < -tol; score mode 0 pen = wv rq_R s^2,
 * 1 mean, 2 mass), score descending with ties in (i, R) order; out_* hold nadd entries, *count the number written;
 * stats[4] = (#violated, violation mass, #violated in the working set, mass in the working set). */
int lb_separate_vcs(lb_ctx *ctx, int m, double tol, int mode, int64_t nadd, int64_t *out_i, int32_t *out_R,
                    double *out_score, int64_t *count, double *stats);

/* Wait for the context's stream. */
int lb_sync(lb_ctx *ctx);

/* Timing of the evaluation: runs `iters` evaluations back to back on the context's stream and returns the mean
 * milliseconds per evaluation (whole evaluation, HIP events) and per kernel (`kernel_ms`, LB_NKERNELS entries,
 * events around each launch in a separate pass; NULL to skip). No reference counterpart (the reference logs
 * wall-clock accumulators, alm_vcs.py:1350-1351,1779). */
#define LB_NKERNELS 8
int lb_time_eval(lb_ctx *ctx, double rho, double rt, int iters, double *ms_per_eval, double *kernel_ms);
/* Name of kernel slot k of lb_time_eval (static string). */
const char *lb_kernel_name(int k);

/* Tuning knobs of the row kernels: rows with more than `heavy_edges` neighbours or `heavy_tris` triangle records
 * are split into segments of `seg_len` items processed by whole wavefronts (takes effect at the next lb_set_graph /
 * lb_set_triangles); `row_grid` workgroups run the grid-stride row kernel (0 keeps the current value). Defaults
 * 128, 128, 1024, 8192: within 2% of the best measured on degree-numbered power-law graphs of 1M rows (mean degree
 * 13) and 2.4M rows (mean degree 50). No reference
 * counterpart (the reference's parallel shapes are numba prange and chunk sizes, alm_vcs.py:1420-1423). */
int lb_set_tuning(lb_ctx *ctx, int heavy_edges, int heavy_tris, int seg_len, int row_grid);
/* Named switches of the kernels (no reference counterpart): "row_order" (default 0) 1 processes the rows of the f+grad
 * row kernel in descending order of their work instead of index order (same results up to summation order; on a
 * graph numbered by degree index order is faster);
 * "vcs_kernel" (default -1: by storage type, 1 for FP32 / FP64 V and 0 for BF16 V) 1 runs the evaluation's VCS pass
 * as one sweep with the Z sums on the FP64 matrix cores, 0 as two passes (multipliers, then Z by LDS adds), same
 * multipliers and G up to summation order; "vcs_grid" (default 0: 512 for the one-pass kernel, 1536 otherwise, at
 * most 1536) its workgroups; "vcs_tile" (default 32, else 16) the rows per wavefront tile of the one-pass kernel (32
 * only where its LDS fits); "sweep_graph" (default 0) 1 captures a sweep's per-batch launches as one HIP graph and
 * replays it while the launch arguments are unchanged (same kernels and results as the eager launches);
 * "sweep_fused" (default 0; with "step_wave" 1) 1 runs a sweep batch as two launches -- the previous batch's VCS post
 * step with this batch's pair rows, then the row steps with the hub segments, each row adding its increment to the
 * cluster sums itself -- in place of five (same arithmetic; the sums' FP64 adds in another order); off by default
 * since one serving run with it ended in a GPU memory fault of unestablished cause; "sweep_pair_segs" (default 1)
 * runs a batch's pair-level rows and hub segments in one launch (0: two launches; same kernels' arithmetic);
 * "sweep_post_merge" (default 1) runs the aggregate increments and the VCS post step of batches of at most 256 rows
 * in one launch, the post step forming the VCS cluster rows after the batch from a snapshot and the batch's
 * increments (0: two launches); "sweep_two" (default 1, eager sweeps with the three switches above on) runs a batch
 * as two launches: its row steps (and a snapshot of Y), then its aggregate increments and VCS post step with the
 * next batch's pair rows and hub segments, which read the snapshot plus the batch's increments and, for a row just
 * stepped, its staged new value (0: the forms above); "sweep_delta" (default 1; batches above 128 rows in that form)
 * has the rows add their increments to a per-batch delta buffer with FP64 atomics, read by the next launch, in place
 * of each of its workgroups summing the batch in LDS (0: always the LDS sums);
 * "sweep_seg_grid"
 * (default 0: 64) caps the workgroups that compute the sweeps' hub segments;
 * "overlap" (default 1) runs the evaluation's hub-segment kernel on a second stream of the context, concurrently
 * with the cluster-sum and VCS kernels, joined before the row kernel (identical results; lb_time_eval's per-kernel
 * figures always come from the one-stream order); "seg_grid" (default 0) caps the workgroups of that hub-segment
 * kernel (0: one per four segments); "sweep_heavy" / "sweep_seg" (defaults 32 / 128) are the hub threshold and
 * segment length of the sweeps' own segment tables; "step_wave" (default 1) runs the sweeps' row step and VCS post
 * step with one wavefront per batch row (0: one slot of r/4 lanes per row); batch contexts: "batch_light" (default
 * 128) is the most edges and triangle records a row may have to run as one slot of the light-row kernel (longer rows
 * take one wavefront each), "batch_light_grid" / "batch_heavy_grid" (defaults 8192) cap the workgroups of those two
 * kernels (lagbound_amd.tune.autotune_batch picks all three). */
int lb_set_option(lb_ctx *ctx, const char *name, int64_t value);

/* Device row work of the candidate bound's Riemannian gradient polish (polish_rgd, est_dual_vcs.py:276-308), with
 * the step rule kept by the caller. The context holds the instance and its multipliers; the iterate W (n x r, unit
 * rows, caller row order) lives on the device in FP64, and each evaluation reads its storage-type copy at rho = 0.
 *   lb_polish_begin(W): out[0] = sum_i g_i (g_i = (C W)_i . W_i), out[1] = sum Rr . R with Rr = C W - g W and
 *     R = Rr / max(|(C W)_i|, 1e-12);
 *   lb_polish_trial(alpha): candidate Wn = normalise(W + alpha R), evaluated; *gsum = sum_i g_i of Wn;
 *   lb_polish_accept: W <- Wn (the last candidate), new R; out4 = (sum Rr . R, sy, ss, yy), the Barzilai-Borwein
 *     sums sum S.(Y nr), S.(S nr), Y.(Y nr) with S = Wn - W, Y = Rn - R, nr the new row norms of C Wn;
 *   lb_polish_get: W and g of the current iterate (either pointer may be NULL).
 * A trial overwrites the context's V. lb_set_graph discards the polish state (the polish calls then fail until the
 * next lb_polish_begin). No reference counterpart beyond the polish itself (its loop is numpy). */
int lb_polish_begin(lb_ctx *ctx, const double *W, double *out);
int lb_polish_trial(lb_ctx *ctx, double alpha, double *gsum);
int lb_polish_accept(lb_ctx *ctx, double *out4);
int lb_polish_get(lb_ctx *ctx, double *W, double *g);

/* Library version string. */
const char *lb_version(void);

/* Exact optimum-preserving reduction (reduce_exact.py:12-62): merge leaves into their neighbour and adjacent true
 * twins (N[u] = N[v], d^2 < 2m, d > 1). keys: m sorted unique u*n + v (u < v < n). rhash: n hash weights of the
 * candidate filter (the reference's are numpy default_rng(11).integers(1, 2**62, n)) or NULL for the library's own;
 * the result does not depend on them. Outputs (caller-allocated): group_of (n), vol (n; first N used), keys_out and
 * w_out (m; first nkeys used: sorted unique group-pair keys a*N + b, a < b, and their edge counts), stats =
 * {N, nkeys, leaf_unions, twin_unions, inside_edges}. Bit-exact with the reference. Uses the context's device and
 * stream only (not its graph). */
int lb_reduce_exact(lb_ctx *ctx, int64_t n, int64_t m, const int64_t *keys, const uint64_t *rhash, int64_t *group_of,
                    int64_t *vol, int64_t *keys_out, double *w_out, int64_t *stats);

/* Dense FP64 certificate step on the device (the dense analogue of the reference's LDL certificate,
 * certify_lifted_vc.py / verify_cert.py psd_ldl, verify_cert.py:92-104): assembles M = Diag(dg) - C(y) from the
 * context's graph, levels, working sets and triangle pool (FP64 context only, so the multipliers are the exact
 * values), factors it with rocSOLVER dpotrf and returns out[0] = info (0: factored; k > 0: leading minor k not
 * positive), out[1] = sum of R_ij^2 of the factor, out[2] = sum over the off-diagonal entries of (sum of the
 * magnitudes of the terms of C_ij)^2, from which the caller prices the rounding: lambda_min(M) >= -(gamma(n+1)
 * out[1] + ||F||_F). Needs 16 n^2 bytes of free device memory. */
int lb_certify_dense(lb_ctx *ctx, const double *dg, double *out);

/* Dense core of the lifted sparse certificate on device `device` (the factorization's last, dense block; the
 * reference factors the whole lifted matrix with CHOLMOD on the host, certify_lifted_vc.py:391-449, cholcert.c:24-76):
 * forms the Schur complement S = A22 - sum_k cd_k c_k c_k^T (m x m FP64) from the entries of A among the m core rows
 * (na, lower: ai >= aj) and, per sparse pivot k, the core part c_k of its column of L (CSR cptr nsp+1 / cidx
 * ascending / cval) with its pivot cd_k, factors S = R R^T and returns out[0] = info (0: factored; k > 0: leading
 * minor k not positive) and out[1] = sum_j (sum_i |R_ij|)^2, from which the caller prices the factor's rounding
 * (Higham, Thm 10.3). method 1: rocSOLVER dpotrf; 2: the library's own blocked right-looking Cholesky (128-column
 * panels, vector FMA trailing updates, 64-bit offsets); 3: the same with the trailing updates on the FP64 matrix
 * cores; 0: rocSOLVER while m^2 < 2^31, method 2 above (rocSOLVER's dpotrf_64 reported a non-positive leading minor at
 * m = 131,287 on a matrix whose leading blocks factor, so it is not used there). Host
 * arrays in; allocates 8 m^2 bytes of device memory on a stream of its own. Returns 0, -1 (bad arguments), -2 (not
 * enough free device memory) or -3 (HIP / rocSOLVER error). */
int lb_cert_core(int device, int64_t m, int64_t na, const int32_t *ai, const int32_t *aj, const double *av,
                 int64_t nsp, const int64_t *cptr, const int32_t *cidx, const double *cval, const double *cd,
                 int method, double *out);

/* Batch mode (no reference counterpart: the reference solves one graph per process): the graph set by lb_set_graph
 * is the block-diagonal union of G graphs with contiguous row ranges row_off (G+1, from 0 to n), each with its own
 * L_g and C0_g. Edges and triangles must stay inside one graph (checked). lb_batch_set_levels gives each graph one
 * pair level: lab (n, cluster ids local to the row's graph), P (G), sig (concatenated P_g x P_g), rq and dsq
 * (concatenated P_g). lb_batch_eval evaluates every graph in one launch of each kernel and writes parts (G x
 * LB_NPARTS, the layout of lb_eval_fgrad per graph) and G = C(y)V for all rows (lb_get_G). Triangle separation of
 * the wedge family (fam 0) on a batch context selects over the union (a wedge stays inside its graph); VCS levels,
 * sweeps and the one-edge family act on single-graph contexts only. */
int lb_batch_set(lb_ctx *ctx, int G, const int64_t *row_off, const double *L, const double *C0);
int lb_batch_set_levels(lb_ctx *ctx, const int32_t *lab, const int32_t *P, const double *sig, const double *rq,
                        const double *dsq, double pairlvl_weight);
int lb_batch_eval(lb_ctx *ctx, double rho, double rt, double *parts);
/* Batched search (no reference counterpart; the reference's inner solver is its sweeps, alm_vcs.py:1360-1436, one
 * graph per process): `iters` iterations of [lb_batch_eval at (rho, rt); one Riemannian ascent step of every graph
 * on the unit rows with its own Barzilai-Borwein step, clamped to [amin, amax] / rmax_g, a0 / rmax_g on its first
 * step, rmax_g the graph's largest row gradient norm (so amax bounds the turn of any row)], recorded once
 * into a HIP graph on the context's stream and replayed (re-recorded when rho, rt, a0, amin or amax change, or after
 * any lb_set_* call). reset = 1 forgets every graph's previous iterate first (after a multiplier update). state
 * (4 per graph, may be NULL): step, have-previous flag, <s,s>, <s,y>. FP32 or FP64 storage. The last evaluation
 * precedes the last step: evaluate again before lb_phr_update. */
int lb_batch_rgd(lb_ctx *ctx, int iters, double rho, double rt, double a0, double amin, double amax, int reset,
                 double *state);
/* The batch pair-level multipliers (concatenated P_g x P_g, the layout of lb_batch_set_levels); lb_phr_update
 * updates them from the last rho > 0 batch evaluation. */
int lb_batch_get_sig(lb_ctx *ctx, double *sig);

/* Captured evaluation (no reference counterpart): record the evaluation chain (lb_eval_fgrad, or lb_batch_eval on
 * a batch context) at fixed rho, rt into a HIP graph on the context's stream, then replay it `count` times with one
 * launch each (parts as in the eager call, may be NULL). Any lb_set_*, lb_sweeps, lb_phr_update or separation call
 * ends the capture; lb_set_V does not (replays read the current V). */
int lb_capture_eval(lb_ctx *ctx, double rho, double rt);
int lb_replay_eval(lb_ctx *ctx, int count, double *parts);

/* Device eigensolver blocks (the pencil LOBPCG of the candidate bound, est_dual_vcs.py:325-364, with its n-sized
 * blocks kept on the device; no reference counterpart at this granularity: the reference runs scipy's LOBPCG on
 * the host). FP64 contexts only. A block is n x r FP64, passed in the caller's numbering (lb_set_row_order applies).
 * lb_blk_alloc reserves nblk blocks and ndiag n-vectors (diagonals). lb_blk_op: dst = Diag(w) src - C(y) src with
 * the instance's multipliers (the evaluation at rho = 0; it overwrites the context's V; diag < 0 drops the Diag
 * term). lb_blk_elem kind 1: dst = Diag(w) src; kind 2: dst = other - Diag(w) src Diag(theta) (theta: r values).
 * lb_blk_gram: out = [A_a0 .. A_a(na-1)]^T Diag(w) [B_b0 ..] as (na r) x (nb r) row-major (diag < 0: no weights),
 * summed in a fixed order. lb_blk_comb: dst (+)= sum_j src_j coef_j, coef_j r x r row-major; dst must not be a
 * source. At most 6 blocks per side. lb_set_graph releases the reservation: the block calls then fail (-1) until the
 * next lb_blk_alloc. */
int lb_blk_alloc(lb_ctx *ctx, int nblk, int ndiag);
int lb_blk_set(lb_ctx *ctx, int k, const double *block);
int lb_blk_get(lb_ctx *ctx, int k, double *block);
int lb_blk_set_diag(lb_ctx *ctx, int k, const double *diag);
int lb_blk_op(lb_ctx *ctx, int dst, int src, int diag);
int lb_blk_elem(lb_ctx *ctx, int kind, int dst, int src, int other, int diag, const double *theta);
int lb_blk_gram(lb_ctx *ctx, int na, const int *a, int nb, const int *b, int diag, double *out);
int lb_blk_comb(lb_ctx *ctx, int dst, int ns, const int *src, const double *coef, int accumulate);

/* Gram matrix without any allocation. No counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops that forbid hipMalloc). The no-allocation variant of
 * lb_blk_gram: out = [A_0 .. A_(na-1)]^T Diag(w) [B_0 .. B_(nb-1)] for caller-owned device blocks A_j, B_j (n x r
 * FP64 each, row-major, r 8 / 16 / 32), w a device n-vector or NULL (no weights), out a caller-owned device array of
 * (na r) x (nb r) doubles (row-major), summed in the same fixed order as lb_blk_gram. Launched on `stream` (a
 * hipStream_t of device `device`; NULL: its null stream) without waiting; the calling thread's current device is
 * left unchanged. The per-row-block partial sums live in a library-owned static device array of 128 x 96 x 96
 * doubles (9.4 MB, one per device until the process ends), so na r and nb r must not exceed 96. Calls that overlap
 * on one device race on that array. Returns 0, -1 (bad arguments), -2 (na r or nb r above 96) or -3 (HIP error). */
int lb_gram_noalloc(int device, void *stream, int64_t n, int r, int na, const double *const *A, int nb,
                    const double *const *B, const double *w, double *out);

#ifdef __cplusplus
}
#endif
#endif
