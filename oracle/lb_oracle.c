/* FP64 CPU restatement of lagbound's ALM search kernels -- TEST INFRASTRUCTURE ONLY.
 *
 * This file is the oracle: the parity checker for the HIP path and the timed CPU baseline. It is never linked into
 * or called by the product library. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Every function restates the loop structure of the reference's numba kernel it cites (prange -> omp parallel for,
 * same per-row / per-entry arithmetic order, FP64 throughout). Citations are into the reference repository
 * (src/lagbound/_core/...).
 *
 * Pinning (see tests/test_oracle.py): the evaluation (or_eval) is checked against a dense brute-force evaluation of
 * the mathematical definitions on small graphs (C(y) assembled entry by entry), against central finite differences
 * of f (gradient = 2 C(y) V in the tangent space), and against a vectorised numpy/scipy restatement at medium size.
 * The row step (or_batch_step) is checked against its closed form (unit rows, minorant step direction) and against
 * the numpy restatement.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct {
    int64_t n;
    int32_t r;
    const int64_t *rowptr;  /* n+1 */
    const int32_t *col;     /* nnz */
    const double *w;        /* nnz or NULL (unweighted) */
    const double *d;        /* n */
    double L, C0;
    /* pair levels */
    int32_t nlev;
    const int32_t *labs;    /* nlev x n */
    const int32_t *P;       /* nlev */
    const double *sig;      /* concat P_l^2 */
    const double *rq;       /* concat P_l */
    const double *dsq;      /* concat P_l */
    double pw;              /* pairlvl_weight */
    /* sparse VC levels */
    int32_t nvcs;
    const int32_t *vlabs;   /* nvcs x n */
    const int32_t *vP;      /* nvcs */
    const double *vrq;      /* concat vP */
    const int64_t *vptr;    /* nvcs x (n+1), each relative to its own level's entries */
    const int32_t *vcol;    /* concat nnz_m */
    const double *vsig;     /* concat nnz_m */
    double wv;              /* vc_weight * pairlvl_weight */
    /* triangles */
    int64_t T;
    const int32_t *TI;      /* T x 3 */
    const double *lam;      /* T */
} or_inst;

static int64_t sumP(const int32_t *P, int32_t k) { int64_t s = 0; for (int32_t i = 0; i < k; i++) s += P[i]; return s; }
static int64_t sumP2(const int32_t *P, int32_t k) { int64_t s = 0; for (int32_t i = 0; i < k; i++) s += (int64_t)P[i] * P[i]; return s; }

/* vertex -> (t, role) incidence CSR, codes 3t+role in t order (alm_vcs.py:94-104) */
void or_incidence(int64_t n, int64_t T, const int32_t *TI, int64_t *cnt, int64_t *codes) {
    memset(cnt, 0, sizeof(int64_t) * (n + 1));
    for (int64_t t = 0; t < T; t++)
        for (int r = 0; r < 3; r++) cnt[TI[3 * t + r] + 1] += 1;
    for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    for (int64_t i = 0; i < n; i++) pos[i] = cnt[i];
    for (int64_t t = 0; t < T; t++)
        for (int r = 0; r < 3; r++) { int64_t i = TI[3 * t + r]; codes[pos[i]++] = 3 * t + r; }
    free(pos);
}

/* out = A X (alm_vcs.py:528-535) */
void or_spmm(const or_inst *I, const double *X, double *out) {
    const int k = I->r;
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < I->n; i++) {
        double *o = out + i * k;
        for (int c = 0; c < k; c++) o[c] = 0.0;
        for (int64_t e = I->rowptr[i]; e < I->rowptr[i + 1]; e++) {
            int64_t j = I->col[e]; double w = I->w ? I->w[e] : 1.0;
            for (int c = 0; c < k; c++) o[c] += w * X[j * k + c];
        }
    }
}

/* sum_ij A_ij V_i.V_j (alm_vcs.py:463-474) */
double or_quadA(const or_inst *I, const double *V) {
    const int k = I->r; double tot = 0.0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot)
    for (int64_t i = 0; i < I->n; i++) {
        double s = 0.0;
        for (int64_t e = I->rowptr[i]; e < I->rowptr[i + 1]; e++) {
            int64_t j = I->col[e]; double dt = 0.0;
            for (int c = 0; c < k; c++) dt += V[i * k + c] * V[j * k + c];
            s += (I->w ? I->w[e] : 1.0) * dt;
        }
        tot += s;
    }
    return tot;
}

/* out[lab(i)] += d_i V_i (cluster_sums, alm_vcs.py:511-526): per-thread partial sums over contiguous row ranges,
 * then a fixed-order reduction over the partials. */
void or_cluster_sums(int64_t n, int k, const int32_t *lab, int32_t P, const double *d, const double *V, double *out) {
    int nthr = 1;
#ifdef _OPENMP
    nthr = omp_get_max_threads();
    if (nthr > 16) nthr = 16;
#endif
    double *part = (double *)calloc((size_t)nthr * P * k, sizeof(double));
    int64_t cs = (n + nthr - 1) / nthr;
#pragma omp parallel for num_threads(nthr)
    for (int t = 0; t < nthr; t++) {
        double *pt = part + (size_t)t * P * k;
        int64_t i1 = (t + 1) * cs < n ? (t + 1) * cs : n;
        for (int64_t i = t * cs; i < i1; i++) {
            double di = d[i]; int64_t Q = lab[i];
            for (int c = 0; c < k; c++) pt[Q * k + c] += di * V[i * k + c];
        }
    }
    for (int64_t q = 0; q < (int64_t)P * k; q++) {
        double s = 0.0;
        for (int t = 0; t < nthr; t++) s += part[(size_t)t * P * k + q];
        out[q] = s;
    }
    free(part);
}

/* level term of augmented() (level_aug, alm_vcs.py:476-496): h = sym(H), rr = rho (rq_Q rq_R) w,
 * pp = max(0, sig - rr h), returns sum (pp^2 - sig^2)/(2 rr); *nviol = #(h < -1e-9). */
double or_level_aug(int32_t P, const double *H, const double *sig, const double *rq, double w, double rho,
                    double *pp, int64_t *nviol) {
    double tot = 0.0; int64_t cnt = 0;
#pragma omp parallel for reduction(+ : tot, cnt)
    for (int32_t Q = 0; Q < P; Q++) {
        double s = 0.0; int64_t c = 0;
        for (int32_t R = 0; R < P; R++) {
            double h = 0.5 * (H[(int64_t)Q * P + R] + H[(int64_t)R * P + Q]);
            double rr = rho * ((rq[Q] * rq[R]) * w), sg = sig[(int64_t)Q * P + R];
            double p = sg - rr * h;
            if (p < 0.0) p = 0.0;
            pp[(int64_t)Q * P + R] = p; s += (p * p - sg * sg) / (2 * rr);
            if (h < -1e-9) c++;
        }
        tot += s; cnt += c;
    }
    *nviol = cnt;
    return tot;
}

/* tri_aug (alm_vcs.py:498-509): gt = 1 - tau, pt = max(0, lam - rt gt); returns sum (pt^2 - lam^2)/(2 rt). */
double or_tri_aug(int64_t T, const int32_t *TI, int k, const double *V, const double *lam, double rt, double *gt,
                  double *pt, double *gmin_out) {
    double tot = 0.0, gmin = INFINITY;
#pragma omp parallel for reduction(+ : tot) reduction(min : gmin)
    for (int64_t t = 0; t < T; t++) {
        int64_t a = TI[3 * t], b = TI[3 * t + 1], c = TI[3 * t + 2]; double s = 0.0;
        for (int q = 0; q < k; q++) s += V[a * k + q] * V[b * k + q] + V[b * k + q] * V[c * k + q] - V[a * k + q] * V[c * k + q];
        double g = 1.0 - s; if (gt) gt[t] = g;
        double p = lam[t] - rt * g; if (p < 0.0) p = 0.0;
        pt[t] = p; tot += (p * p - lam[t] * lam[t]) / (2 * rt);
        if (g < gmin) gmin = g;
    }
    *gmin_out = gmin;
    return tot;
}

/* out += T(w) V via the incidence (tri_apply, alm_vcs.py:106-120) */
void or_tri_apply(int64_t n, int k, const int32_t *TI, const double *w, const int64_t *ptr, const int64_t *codes,
                  const double *V, double *out) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            int64_t code = codes[e], t = code / 3; int r = (int)(code % 3); double h = 0.5 * w[t];
            if (h == 0.0) continue;
            int64_t a = TI[3 * t], b = TI[3 * t + 1], c = TI[3 * t + 2];
            if (r == 0) for (int q = 0; q < k; q++) out[i * k + q] += h * (V[c * k + q] - V[b * k + q]);
            else if (r == 1) for (int q = 0; q < k; q++) out[i * k + q] -= h * (V[a * k + q] + V[c * k + q]);
            else for (int q = 0; q < k; q++) out[i * k + q] += h * (V[a * k + q] - V[b * k + q]);
        }
    }
}

/* VCS augmented term over the working set (_vcs_aug_rows, vcsops.py:266-288), Y = W V of the level.
 * pp = max(0, sig - rho r g), r = wv rq_R / d_i^2, g = d_i (V_i.Y_R - [R = R(i)] d_i). */
double or_vcs_aug(int64_t n, int k, const double *V, const double *d, const int32_t *lab, const double *Y,
                  const int64_t *ptr, const int32_t *col, const double *sig, const double *rq, double wv, double rho,
                  double *pp, int64_t *nviol) {
    double tot = 0.0; int64_t cnt = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : tot, cnt)
    for (int64_t i = 0; i < n; i++) {
        double di = d[i], acc = 0.0; int32_t Ri = lab[i]; int64_t cc = 0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            int32_t R = col[e]; double s = 0.0;
            for (int c = 0; c < k; c++) s += V[i * k + c] * Y[(int64_t)R * k + c];
            double g = di * s;
            if (R == Ri) g = di * (s - di);
            double sg = sig[e], p = sg;
            if (rho > 0.0 && di > 0.0) {
                double rr = rho * (wv * rq[R] / (di * di)); p = sg - rr * g; if (p < 0.0) p = 0.0;
                acc += (p * p - sg * sg) / (2 * rr);
            }
            pp[e] = p;
            if (g < -1e-9) cc++;
        }
        tot += acc; cnt += cc;
    }
    *nviol = cnt;
    return tot;
}

/* out += S_vc X = 1/2 d (Sig (W X) + (Sig^T (d X))[lab]) - own d^2 X (vcs_apply, vcsops.py:326-330) */
void or_vcs_apply(int64_t n, int k, int32_t P, const int32_t *lab, const double *d, const int64_t *ptr,
                  const int32_t *col, const double *sig, const double *X, double *out) {
    double *YX = (double *)calloc((size_t)P * k, sizeof(double));
    double *ZX = (double *)calloc((size_t)P * k, sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        for (int c = 0; c < k; c++) YX[(int64_t)lab[i] * k + c] += d[i] * X[i * k + c];
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++)
            for (int c = 0; c < k; c++) ZX[(int64_t)col[e] * k + c] += sig[e] * d[i] * X[i * k + c];
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < n; i++) {
        double di = d[i], own = 0.0; int32_t Ri = lab[i];
        double acc[64];
        for (int c = 0; c < k; c++) acc[c] = 0.0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            int32_t R = col[e];
            if (R == Ri) own = sig[e];
            for (int c = 0; c < k; c++) acc[c] += sig[e] * YX[(int64_t)R * k + c];
        }
        for (int c = 0; c < k; c++)
            out[i * k + c] += 0.5 * di * (acc[c] + ZX[(int64_t)Ri * k + c]) - own * di * di * X[i * k + c];
    }
    free(YX); free(ZX);
}

/* One f+grad evaluation.
 *   rho > 0: f = augmented(V, rho) (alm_vcs.py:1496-1533): quadA - |d^T V|^2/L - sum_l level_aug - tri_aug - vcs_aug
 *            + C0, effective multipliers pp (levels), pt (triangles, rt = rho * tri_weight), pv (VCS);
 *   rho = 0: f = C0 + sum(lam) + sum_i (C V)_i . V_i  (est_dual_vcs.value, est_dual_vcs.py:245-246), multipliers as
 *            stored.
 *   G = C(y) V with y the effective multipliers (CV, alm_vcs.py:1299-1312; est_dual_vcs.py:227-243).
 * parts: [0] f, [1] quadA (rho > 0) or sum_i G_i.V_i (rho = 0), [2] |z|^2/L, [3] sum of level terms, [4] triangle
 * term, [5] VCS term, [6] primal (quadA - |z|^2/L + C0), [7] C0, [8] level #(h<-1e-9), [9] min triangle slack,
 * [10] VCS #(g<-1e-9), [11] sum(lam).
 * Returns 0 on success. */
int or_eval(const or_inst *I, const double *V, double rho, double rt, double *parts, double *G, double *pp,
            double *pt, double *pv, const int64_t *inc_ptr, const int64_t *inc_codes) {
    const int64_t n = I->n; const int k = I->r;
    for (int q = 0; q < 12; q++) parts[q] = 0.0;
    double *z = (double *)calloc(k, sizeof(double));
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < k; c++) z[c] += I->d[i] * V[i * k + c];
    double zz = 0.0;
    for (int c = 0; c < k; c++) zz += z[c] * z[c];
    double lamsum = 0.0;
    for (int64_t t = 0; t < I->T; t++) lamsum += I->lam[t];
    parts[11] = lamsum; parts[7] = I->C0; parts[2] = zz / I->L;

    /* effective multipliers */
    int64_t so = 0, po = 0;
    double levtot = 0.0; int64_t levviol = 0;
    double **Ys = (double **)calloc(I->nlev > 0 ? I->nlev : 1, sizeof(double *));
    for (int32_t l = 0; l < I->nlev; l++) {
        int32_t P = I->P[l];
        double *Y = (double *)calloc((size_t)P * k, sizeof(double));
        or_cluster_sums(n, k, I->labs + (int64_t)l * n, P, I->d, V, Y);
        Ys[l] = Y;
        if (rho > 0.0) {
            double *H = (double *)malloc(sizeof(double) * (size_t)P * P);
            for (int32_t Q = 0; Q < P; Q++)
                for (int32_t R = 0; R < P; R++) {
                    double s = 0.0;
                    for (int c = 0; c < k; c++) s += Y[(int64_t)Q * k + c] * Y[(int64_t)R * k + c];
                    H[(int64_t)Q * P + R] = s;
                }
            for (int32_t Q = 0; Q < P; Q++) H[(int64_t)Q * P + Q] -= I->dsq[po + Q];
            int64_t nv = 0;
            levtot += or_level_aug(P, H, I->sig + so, I->rq + po, I->pw, rho, pp + so, &nv);
            levviol += nv;
            free(H);
        } else {
            memcpy(pp + so, I->sig + so, sizeof(double) * (size_t)P * P);
        }
        so += (int64_t)P * P; po += P;
    }
    parts[3] = levtot; parts[8] = (double)levviol;

    double tritot = 0.0, gmin = INFINITY;
    if (I->T) {
        if (rho > 0.0) tritot = or_tri_aug(I->T, I->TI, k, V, I->lam, rt, NULL, pt, &gmin);
        else memcpy(pt, I->lam, sizeof(double) * I->T);
    }
    parts[4] = tritot; parts[9] = gmin;

    double vcstot = 0.0; int64_t vcsviol = 0; int64_t eo = 0, vpo = 0;
    for (int32_t m = 0; m < I->nvcs; m++) {
        int32_t P = I->vP[m]; const int32_t *lab = I->vlabs + (int64_t)m * n; const int64_t *ptr = I->vptr + (int64_t)m * (n + 1);
        int64_t nnz = ptr[n];
        double *Y = (double *)calloc((size_t)P * k, sizeof(double));
        or_cluster_sums(n, k, lab, P, I->d, V, Y);
        int64_t nv = 0;
        if (rho > 0.0) vcstot += or_vcs_aug(n, k, V, I->d, lab, Y, ptr, I->vcol + eo, I->vsig + eo, I->vrq + vpo, I->wv, rho, pv + eo, &nv);
        else {
            or_vcs_aug(n, k, V, I->d, lab, Y, ptr, I->vcol + eo, I->vsig + eo, I->vrq + vpo, I->wv, 0.0, pv + eo, &nv);
        }
        vcsviol += nv;
        free(Y);
        eo += nnz; vpo += P;
    }
    parts[5] = vcstot; parts[10] = (double)vcsviol;

    /* G = C(y) V */
    or_spmm(I, V, G);
#pragma omp parallel for
    for (int64_t i = 0; i < n; i++)
        for (int c = 0; c < k; c++) G[i * k + c] -= I->d[i] * z[c] / I->L;
    so = 0; po = 0;
    for (int32_t l = 0; l < I->nlev; l++) {
        int32_t P = I->P[l]; const int32_t *lab = I->labs + (int64_t)l * n; const double *s = pp + so; double *Y = Ys[l];
        double *SY = (double *)calloc((size_t)P * k, sizeof(double));
        for (int32_t Q = 0; Q < P; Q++)
            for (int32_t R = 0; R < P; R++)
                for (int c = 0; c < k; c++) SY[(int64_t)Q * k + c] += s[(int64_t)Q * P + R] * Y[(int64_t)R * k + c];
#pragma omp parallel for
        for (int64_t i = 0; i < n; i++) {
            int32_t Q = lab[i]; double di = I->d[i];
            for (int c = 0; c < k; c++)
                G[i * k + c] += di * SY[(int64_t)Q * k + c] - (s[(int64_t)Q * P + Q] * di * di) * V[i * k + c];
        }
        free(SY);
        so += (int64_t)P * P; po += P;
    }
    eo = 0;
    for (int32_t m = 0; m < I->nvcs; m++) {
        const int64_t *ptr = I->vptr + (int64_t)m * (n + 1);
        or_vcs_apply(n, k, I->vP[m], I->vlabs + (int64_t)m * n, I->d, ptr, I->vcol + eo, pv + eo, V, G);
        eo += ptr[n];
    }
    if (I->T) {
        /* the incidence is rebuilt only when the pool changes (alm_vcs.py:1228,1762); callers pass it in */
        if (inc_ptr && inc_codes) {
            or_tri_apply(n, k, I->TI, pt, inc_ptr, inc_codes, V, G);
        } else {
            int64_t *cnt = (int64_t *)malloc(sizeof(int64_t) * (n + 1));
            int64_t *codes = (int64_t *)malloc(sizeof(int64_t) * 3 * I->T);
            or_incidence(n, I->T, I->TI, cnt, codes);
            or_tri_apply(n, k, I->TI, pt, cnt, codes, V, G);
            free(cnt); free(codes);
        }
    }

    if (rho > 0.0) {
        double qa = or_quadA(I, V);
        parts[1] = qa;
        parts[6] = qa - zz / I->L + I->C0;
        parts[0] = qa - zz / I->L - levtot - tritot - vcstot + I->C0;
    } else {
        double gv = 0.0;
#pragma omp parallel for reduction(+ : gv)
        for (int64_t i = 0; i < n; i++) {
            double s = 0.0;
            for (int c = 0; c < k; c++) s += G[i * k + c] * V[i * k + c];
            gv += s;
        }
        parts[1] = gv;
        parts[6] = or_quadA(I, V) - zz / I->L + I->C0;
        parts[0] = I->C0 + lamsum + gv;
    }
    for (int32_t l = 0; l < I->nlev; l++) free(Ys[l]);
    free(Ys); free(z);
    return 0;
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void or_set_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* ------------------------------------------------------------------------------------------------ sweeps
 * Batched Gauss-Seidel row steps (sweeps, alm_vcs.py:1360-1436, fast bookkeeping with pair_refresh = 1) with the
 * per-batch snapshot semantics of the device path: every row of a batch reads V, z, Y, M, pdiag, Kq and the VCS
 * state as they were when the batch started (the reference lets rows of one batch read each other's new values,
 * Hogwild, alm_vcs.py:13-14); the batch's new rows, z += sum DD, Y[Q] += DD (_acc_rows, alm_vcs.py:430-437) and the
 * VCS Y / Z / cnt / stZ updates (vcs_yupd, vcs_post, vcsops.py:159-213) are applied after the batch.
 * orders: nsw x n row permutations; batches are np.array_split(order, nbatch) (the first n % nbatch batches hold
 * one extra row). kap2 = vc_kap2 (default 1). */

/* kq_level (alm_vcs.py:417-428): Kq_Q = 2 rho (sum_R [pp_QR > 0] r_QR y2_R + [pp_QQ > 0] r_QQ y2_Q) */
static void or_kq_level(int32_t P, int k, const double *Y, const double *sig, const double *rq, const double *dsq,
                        double w, double rho, double *out) {
    double *y2 = (double *)malloc(sizeof(double) * (P > 0 ? P : 1));
    for (int32_t R = 0; R < P; R++) { double s = 0.0; for (int c = 0; c < k; c++) s += Y[(int64_t)R * k + c] * Y[(int64_t)R * k + c]; y2[R] = s; }
    for (int32_t Q = 0; Q < P; Q++) {
        double s = 0.0, pq = 0.0;
        for (int32_t R = 0; R < P; R++) {
            double H = 0.0;
            for (int c = 0; c < k; c++) H += Y[(int64_t)Q * k + c] * Y[(int64_t)R * k + c];
            if (Q == R) H -= dsq[Q];
            double r = (rq[Q] * rq[R]) * w;
            double p = sig[(int64_t)Q * P + R] - rho * r * H;
            if (p > 0.0) s += r * y2[R];
            if (Q == R) pq = p;
        }
        double dg = pq > 0.0 ? (rq[Q] * rq[Q]) * w * y2[Q] : 0.0;
        out[Q] = 2 * rho * (s + dg);
    }
    free(y2);
}

/* pair_all (alm_vcs.py:378-415): fresh pp row of cluster Q from the current Y, M_Q = sum_R pp_QR Y_R, pdiag_Q */
static void or_pair_row(int32_t P, int k, int32_t Q, const double *Y, const double *sig, const double *rq,
                        const double *dsq, double w, double rho, double *M, double *pdiag) {
    double pq = 0.0;
    for (int c = 0; c < k; c++) M[(int64_t)Q * k + c] = 0.0;
    for (int32_t R = 0; R < P; R++) {
        double p = sig[(int64_t)Q * P + R];
        if (rho > 0.0) {
            double h = 0.0;
            for (int c = 0; c < k; c++) h += Y[(int64_t)Q * k + c] * Y[(int64_t)R * k + c];
            if (R == Q) h -= dsq[Q];
            p = p - (rho * ((rq[Q] * rq[R]) * w)) * h;
            if (p < 0.0) p = 0.0;
        }
        if (R == Q) pq = p;
        if (p != 0.0)
            for (int c = 0; c < k; c++) M[(int64_t)Q * k + c] += p * Y[(int64_t)R * k + c];
    }
    pdiag[Q] = pq;
}

/* vcs_sigt (vcsops.py:94-108) for one row: effective multipliers at (V, Y) */
static double or_vcs_p(double sg, double s, double di, int own, double rho, double wv, double rqR) {
    if (rho <= 0.0 || di <= 0.0) return sg;
    double g = own ? di * (s - di) : di * s;
    double p = sg - rho * (wv * rqR / (di * di)) * g;
    return p > 0.0 ? p : 0.0;
}

int or_sweeps(const or_inst *I, double *V, int nsw, double rho, double rt, double damp, const int64_t *orders,
              int nbatch, double kap2, const int64_t *inc_ptr, const int64_t *inc_codes) {
    const int64_t n = I->n; const int k = I->r;
    if (nbatch < 1 || k > 64) return -1;
    int64_t sP = 0, sPv = 0, nnzv = 0;
    for (int32_t l = 0; l < I->nlev; l++) sP += I->P[l];
    for (int32_t m = 0; m < I->nvcs; m++) { sPv += I->vP[m]; nnzv += I->vptr[(int64_t)m * (n + 1) + n]; }
    double *z = (double *)calloc(k, sizeof(double));
    double *Y = (double *)calloc((size_t)(sP ? sP : 1) * k, sizeof(double));
    double *Yv = (double *)calloc((size_t)(sPv ? sPv : 1) * k, sizeof(double));
    double *Kq = (double *)calloc(sP ? sP : 1, sizeof(double));
    double *M = (double *)calloc((size_t)(sP ? sP : 1) * k, sizeof(double));
    double *pdiag = (double *)calloc(sP ? sP : 1, sizeof(double));
    double *Z = (double *)calloc((size_t)(sPv ? sPv : 1) * k, sizeof(double));
    double *cnt = (double *)calloc(sPv ? sPv : 1, sizeof(double));
    double *stZ = (double *)calloc(nnzv ? nnzv : 1, sizeof(double));
    int64_t maxb = n / nbatch + 1;
    double *Vn = (double *)malloc(sizeof(double) * (size_t)maxb * k);
    double *Vo = (double *)malloc(sizeof(double) * (size_t)maxb * k);
    double *DD = (double *)malloc(sizeof(double) * (size_t)maxb * k);
    int64_t *cinc = NULL, *ccodes = NULL;
    if (I->T && !(inc_ptr && inc_codes)) {
        cinc = (int64_t *)malloc(sizeof(int64_t) * (n + 1)); ccodes = (int64_t *)malloc(sizeof(int64_t) * 3 * I->T);
        or_incidence(n, I->T, I->TI, cinc, ccodes);
        inc_ptr = cinc; inc_codes = ccodes;
    }
    for (int sw = 0; sw < nsw; sw++) {
        const int64_t *order = orders + (int64_t)sw * n;
        /* sweep start: z, Y (all pair levels), Kq, VCS state */
        for (int c = 0; c < k; c++) z[c] = 0.0;
        for (int64_t i = 0; i < n; i++) for (int c = 0; c < k; c++) z[c] += I->d[i] * V[i * k + c];
        int64_t po = 0, so = 0;
        for (int32_t l = 0; l < I->nlev; l++) {
            or_cluster_sums(n, k, I->labs + (int64_t)l * n, I->P[l], I->d, V, Y + po * k);
            if (rho > 0.0) or_kq_level(I->P[l], k, Y + po * k, I->sig + so, I->rq + po, I->dsq + po, I->pw, rho, Kq + po);
            else for (int32_t Q = 0; Q < I->P[l]; Q++) Kq[po + Q] = 0.0;
            po += I->P[l]; so += (int64_t)I->P[l] * I->P[l];
        }
        int64_t vpo = 0, eo = 0;
        for (int32_t m = 0; m < I->nvcs; m++) {
            const int32_t *lab = I->vlabs + (int64_t)m * n; const int64_t *ptr = I->vptr + (int64_t)m * (n + 1);
            double *Ym = Yv + vpo * k;
            or_cluster_sums(n, k, lab, I->vP[m], I->d, V, Ym);
            for (int32_t R = 0; R < I->vP[m]; R++) { cnt[vpo + R] = 0.0; for (int c = 0; c < k; c++) Z[(vpo + R) * k + c] = 0.0; }
            for (int64_t i = 0; i < n; i++) {
                double di = I->d[i];
                for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
                    int32_t R = I->vcol[eo + e]; double s = 0.0;
                    for (int c = 0; c < k; c++) s += V[i * k + c] * Ym[(int64_t)R * k + c];
                    double p = or_vcs_p(I->vsig[eo + e], s, di, R == lab[i], rho, I->wv, I->vrq[vpo + R]);
                    stZ[eo + e] = p;
                    if (p != 0.0) for (int c = 0; c < k; c++) Z[(vpo + R) * k + c] += p * (di * V[i * k + c]);
                    if (p > 0.0) cnt[vpo + R] += 1.0;
                }
            }
            vpo += I->vP[m]; eo += ptr[n];
        }
        int64_t start = 0;
        for (int b = 0; b < nbatch; b++) {
            int64_t bs = n / nbatch + (b < n % nbatch ? 1 : 0);
            const int64_t *rows = order + start;
            /* pair_all for every cluster (equal to the touched-only refresh for the rows that read it) */
            po = 0; so = 0;
            for (int32_t l = 0; l < I->nlev; l++) {
                for (int32_t Q = 0; Q < I->P[l]; Q++)
                    or_pair_row(I->P[l], k, Q, Y + po * k, I->sig + so, I->rq + po, I->dsq + po, I->pw, rho, M + po * k, pdiag + po);
                po += I->P[l]; so += (int64_t)I->P[l] * I->P[l];
            }
            /* row steps from the batch snapshot (batch_step6, alm_vcs.py:255-351; vcs_pre, vcsops.py:127-156) */
#pragma omp parallel for schedule(dynamic, 16)
            for (int64_t idx = 0; idx < bs; idx++) {
                int64_t i = rows[idx]; double di = I->d[i];
                double g[64], q[64];
                const double *vi = V + i * k;
                for (int c = 0; c < k; c++) g[c] = -(di / I->L) * (z[c] - di * vi[c]);
                for (int64_t e = I->rowptr[i]; e < I->rowptr[i + 1]; e++) {
                    int64_t j = I->col[e]; double w = I->w ? I->w[e] : 1.0;
                    for (int c = 0; c < k; c++) g[c] += w * V[j * k + c];
                }
                double kap = 0.0;
                int64_t pl = 0;
                for (int32_t l = 0; l < I->nlev; l++) {
                    int64_t Q = pl + I->labs[(int64_t)l * n + i]; double pd = pdiag[Q] * di;
                    for (int c = 0; c < k; c++) g[c] += di * (M[Q * k + c] - pd * vi[c]);
                    kap += di * di * Kq[Q];
                    pl += I->P[l];
                }
                int64_t vp = 0, eo2 = 0;
                for (int32_t m = 0; m < I->nvcs; m++) {
                    const int32_t *lab = I->vlabs + (int64_t)m * n; const int64_t *ptr = I->vptr + (int64_t)m * (n + 1);
                    const double *Ym = Yv + vp * k; int32_t Ri = lab[i];
                    double acc[64]; for (int c = 0; c < k; c++) acc[c] = 0.0;
                    double kp = 0.0, pown = 0.0, sz = 0.0;
                    for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
                        int32_t R = I->vcol[eo2 + e]; double s = 0.0;
                        for (int c = 0; c < k; c++) s += vi[c] * Ym[(int64_t)R * k + c];
                        double p;
                        if (rho > 0.0 && di > 0.0) {
                            double g2 = (R == Ri) ? di * (s - di) : di * s;
                            double rr = I->wv * I->vrq[vp + R] / (di * di);
                            p = I->vsig[eo2 + e] - rho * rr * g2;
                            if (p > 0.0) {
                                double y2 = 0.0;
                                for (int c = 0; c < k; c++) y2 += Ym[(int64_t)R * k + c] * Ym[(int64_t)R * k + c];
                                if (R == Ri) y2 = y2 - 2 * di * s + di * di;
                                kp += rho * rr * di * di * y2;
                            } else p = 0.0;
                        } else p = I->vsig[eo2 + e];
                        if (R == Ri) { pown = p; sz = stZ[eo2 + e]; }
                        if (p != 0.0) for (int c = 0; c < k; c++) acc[c] += p * Ym[(int64_t)R * k + c];
                    }
                    for (int c = 0; c < k; c++)
                        g[c] += 0.5 * di * (acc[c] - pown * di * vi[c]) + 0.5 * di * (Z[(vp + Ri) * k + c] - sz * di * vi[c]);
                    if (rho > 0.0) kp += kap2 * rho * di * di * I->wv * I->vrq[vp + Ri] * (cnt[vp + Ri] - (sz > 0.0 ? 1.0 : 0.0));
                    kap += kp;
                    vp += I->vP[m]; eo2 += ptr[n];
                }
                if (I->T) {
                    for (int64_t e = inc_ptr[i]; e < inc_ptr[i + 1]; e++) {
                        int64_t code = inc_codes[e], t = code / 3; int r = (int)(code % 3);
                        const double *va = V + (int64_t)I->TI[3 * t] * k, *vb = V + (int64_t)I->TI[3 * t + 1] * k, *vc = V + (int64_t)I->TI[3 * t + 2] * k;
                        double tau = 0.0, q2 = 0.0;
                        for (int c = 0; c < k; c++) {
                            tau += va[c] * vb[c] + vb[c] * vc[c] - va[c] * vc[c];
                            q[c] = r == 0 ? vb[c] - vc[c] : (r == 1 ? va[c] + vc[c] : vb[c] - va[c]);
                            q2 += q[c] * q[c];
                        }
                        double pt = I->lam[t] - rt * (1.0 - tau);
                        if (pt > 0.0) { for (int c = 0; c < k; c++) g[c] -= 0.5 * pt * q[c]; kap += rt * q2; }
                    }
                }
                double gn = 0.0;
                for (int c = 0; c < k; c++) gn += g[c] * g[c];
                double w = 0.5 * kap + damp * sqrt(gn), nv = 0.0;
                for (int c = 0; c < k; c++) { q[c] = g[c] + w * vi[c]; nv += q[c] * q[c]; }
                for (int c = 0; c < k; c++) Vo[idx * k + c] = vi[c];
                if (nv <= 0.0) {
                    for (int c = 0; c < k; c++) { Vn[idx * k + c] = vi[c]; DD[idx * k + c] = 0.0; }
                    continue;
                }
                nv = 1.0 / sqrt(nv);
                for (int c = 0; c < k; c++) { double vn = q[c] * nv; Vn[idx * k + c] = vn; DD[idx * k + c] = di * (vn - vi[c]); }
            }
            /* apply the batch, then the aggregate updates in batch-row order */
            for (int64_t idx = 0; idx < bs; idx++) {
                int64_t i = rows[idx];
                for (int c = 0; c < k; c++) V[i * k + c] = Vn[idx * k + c];
            }
            for (int64_t idx = 0; idx < bs; idx++) {
                int64_t i = rows[idx];
                for (int c = 0; c < k; c++) {
                    double x = DD[idx * k + c]; z[c] += x;
                    int64_t pl = 0;
                    for (int32_t l = 0; l < I->nlev; l++) { Y[(pl + I->labs[(int64_t)l * n + i]) * k + c] += x; pl += I->P[l]; }
                    int64_t vp = 0;
                    for (int32_t m = 0; m < I->nvcs; m++) { Yv[(vp + I->vlabs[(int64_t)m * n + i]) * k + c] += x; vp += I->vP[m]; }
                }
            }
            int64_t vp = 0, eo3 = 0;
            for (int32_t m = 0; m < I->nvcs; m++) {   /* vcs_post: new effective multipliers of the moved rows */
                const int32_t *lab = I->vlabs + (int64_t)m * n; const int64_t *ptr = I->vptr + (int64_t)m * (n + 1);
                const double *Ym = Yv + vp * k;
                for (int64_t idx = 0; idx < bs; idx++) {
                    int64_t i = rows[idx]; double di = I->d[i];
                    for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
                        int32_t R = I->vcol[eo3 + e]; double s = 0.0;
                        for (int c = 0; c < k; c++) s += V[i * k + c] * Ym[(int64_t)R * k + c];
                        double sn = or_vcs_p(I->vsig[eo3 + e], s, di, R == lab[i], rho, I->wv, I->vrq[vp + R]);
                        double so2 = stZ[eo3 + e];
                        if (sn != 0.0 || so2 != 0.0)
                            for (int c = 0; c < k; c++) Z[(vp + R) * k + c] += di * (sn * V[i * k + c] - so2 * Vo[idx * k + c]);
                        if (sn > 0.0) { if (!(so2 > 0.0)) cnt[vp + R] += 1.0; }
                        else if (so2 > 0.0) cnt[vp + R] -= 1.0;
                        stZ[eo3 + e] = sn;
                    }
                }
                vp += I->vP[m]; eo3 += ptr[n];
            }
            start += bs;
        }
    }
    free(z); free(Y); free(Yv); free(Kq); free(M); free(pdiag); free(Z); free(cnt); free(stZ); free(Vn); free(Vo); free(DD);
    free(cinc); free(ccodes);
    return 0;
}

/* ------------------------------------------------------------------------------------------------ separation
 * Triangle separation with the reference's semantics (separate_ref / separate1_ref, alm_vcs.py:965-984): FP64 dot
 * products summed over q in order without contraction, 512 chunks of consecutive centres, a 2000-bin histogram of
 * the violations vm > vtol, the first bin b0 whose tail count is <= cap, vmin = vtol + (2 - vtol) b0 / nb, and per
 * chunk the candidates with vm >= vmin in generation order (centre, then arm pairs), at most the chunk's histogram
 * tail count. Tie rule for the top-maxh / top-maxk cuts: descending value, ties by position (stable); the
 * reference's unstable argsort agrees whenever the values at the cut are distinct. */

typedef struct { double x; int64_t id; int64_t pos; } or_arm;

static int or_cmp_desc(const void *a, const void *b) {
    const or_arm *p = (const or_arm *)a, *q = (const or_arm *)b;
    if (p->x > q->x) return -1;
    if (p->x < q->x) return 1;
    return p->pos < q->pos ? -1 : (p->pos > q->pos ? 1 : 0);
}
static int or_cmp_id(const void *a, const void *b) {
    const or_arm *p = (const or_arm *)a, *q = (const or_arm *)b;
    if (p->id < q->id) return -1;
    if (p->id > q->id) return 1;
    return p->pos < q->pos ? -1 : (p->pos > q->pos ? 1 : 0);
}

static double or_dot(const double *a, const double *b, int k) {
    double s = 0.0;
    for (int q = 0; q < k; q++) s += a[q] * b[q];
    return s;
}

/* _centre (alm_vcs.py:554-565): arms of centre j; returns the count, arms in H (CSR order, or value-descending
 * after the top-maxh cut) */
static int64_t or_centre(int64_t j, const int64_t *rowptr, const int32_t *col, int k, const double *V, double thresh,
                         int64_t maxh, or_arm *H) {
    int64_t m = 0;
    for (int64_t e = rowptr[j]; e < rowptr[j + 1]; e++) {
        int64_t i = col[e];
        double s = or_dot(V + i * k, V + j * k, k);
        if (s > thresh) { H[m].x = s; H[m].id = i; H[m].pos = m; m++; }
    }
    if (m > maxh) {
        qsort(H, (size_t)m, sizeof(or_arm), or_cmp_desc);
        m = maxh;
    }
    return m;
}

static int or_isadj(const int64_t *rowptr, const int32_t *col, int64_t j, int64_t x) {
    int64_t lo = rowptr[j], hi = rowptr[j + 1];
    while (lo < hi) { int64_t mid = (lo + hi) / 2; if (col[mid] < x) lo = mid + 1; else hi = mid; }
    return lo < rowptr[j + 1] && col[lo] == x;
}

/* _centre1 (alm_vcs.py:608-632): non-adjacent 2-hop arms K of centre j through the first hub2 arms of H (at most
 * cap2 neighbours each), value > thresh, sorted by id and deduplicated, then the top-maxk by value */
static int64_t or_centre1(int64_t j, const int64_t *rowptr, const int32_t *col, int k, const double *V, double thresh,
                          const or_arm *H, int64_t mh, int64_t maxk, int64_t hub2, int64_t cap2, or_arm *K) {
    int64_t nh = mh < hub2 ? mh : hub2, m = 0;
    for (int64_t a = 0; a < nh; a++) {
        int64_t i2 = H[a].id, lo = rowptr[i2], hi = rowptr[i2 + 1];
        if (hi > lo + cap2) hi = lo + cap2;
        for (int64_t e = lo; e < hi; e++) {
            int64_t c = col[e];
            if (c == j || or_isadj(rowptr, col, j, c)) continue;
            double s = or_dot(V + c * k, V + j * k, k);
            if (s > thresh) { K[m].x = s; K[m].id = c; K[m].pos = m; m++; }
        }
    }
    if (m > 1) {
        qsort(K, (size_t)m, sizeof(or_arm), or_cmp_id);
        int64_t u = 1;
        for (int64_t t = 1; t < m; t++) if (K[t].id != K[u - 1].id) K[u++] = K[t];
        m = u;
    }
    for (int64_t t = 0; t < m; t++) K[t].pos = t;
    if (m > maxk) {
        qsort(K, (size_t)m, sizeof(or_arm), or_cmp_desc);
        m = maxk;
    }
    return m;
}

static int or_bin(double vm, double vtol, int nb) {
    int b = (int)((vm - vtol) / (2.0 - vtol) * nb);
    return b < nb - 1 ? b : nb - 1;
}

/* one family over every centre of chunk ch: fam 0 = wedges (a, j, c), a, c in H; fam 1 = one-edge arm (h, j, k),
 * h in H, k in K. mode 0: histogram (hist[nb]); mode 1: collect candidates with vm >= vmin into out/val from *m up
 * to end. */
static void or_sep_chunk(int fam, int mode, int64_t ch, int64_t cs, int64_t n, const int64_t *rowptr,
                         const int32_t *col, int k, const double *V, double thresh, int64_t maxh, int64_t maxk,
                         int64_t hub2, int64_t cap2, double vtol, int nb, int64_t *hist, double vmin, int64_t *m,
                         int64_t end, int64_t *out, double *val, or_arm *H, or_arm *K) {
    int64_t j1 = (ch + 1) * cs < n ? (ch + 1) * cs : n;
    for (int64_t j = ch * cs; j < j1; j++) {
        int64_t dj = rowptr[j + 1] - rowptr[j];
        if (dj < (fam == 0 ? 2 : 1)) continue;
        int64_t mh = or_centre(j, rowptr, col, k, V, thresh, maxh, H);
        if (fam == 0) {
            for (int64_t x = 0; x < mh; x++)
                for (int64_t y = x + 1; y < mh; y++) {
                    double s = or_dot(V + H[x].id * k, V + H[y].id * k, k);
                    double vm = H[x].x + H[y].x - s - 1.0;
                    if (mode == 0) { if (vm > vtol) hist[or_bin(vm, vtol, nb)]++; }
                    else if (vm >= vmin && *m < end) {
                        int64_t a = H[x].id < H[y].id ? H[x].id : H[y].id, c = H[x].id < H[y].id ? H[y].id : H[x].id;
                        out[3 * *m] = a; out[3 * *m + 1] = j; out[3 * *m + 2] = c; val[*m] = vm; (*m)++;
                    }
                }
        } else {
            int64_t mk = or_centre1(j, rowptr, col, k, V, thresh, H, mh, maxk, hub2, cap2, K);
            for (int64_t x = 0; x < mh; x++)
                for (int64_t y = 0; y < mk; y++) {
                    if (H[x].id == K[y].id) continue;
                    double s = or_dot(V + H[x].id * k, V + K[y].id * k, k);
                    double vm = H[x].x + K[y].x - s - 1.0;
                    if (mode == 0) { if (vm > vtol) hist[or_bin(vm, vtol, nb)]++; }
                    else if (vm >= vmin && *m < end) {
                        int64_t a = H[x].id < K[y].id ? H[x].id : K[y].id, c = H[x].id < K[y].id ? K[y].id : H[x].id;
                        out[3 * *m] = a; out[3 * *m + 1] = j; out[3 * *m + 2] = c; val[*m] = vm; (*m)++;
                    }
                }
        }
    }
}

/* separate_ref (fam 0) / separate1_ref (fam 1). out: cap x 3 int64, val: cap. Returns the number of candidates
 * kept (val > 0), in the reference's output order; *vmin_out = vmin (or -1 if nothing is selected). */
int64_t or_separate(int fam, int64_t n, const int64_t *rowptr, const int32_t *col, int k, const double *V,
                    double thresh, int64_t maxh, int64_t maxk, int64_t hub2, int64_t cap2, double vtol, int64_t cap,
                    int nchunk, int nb, int64_t *out, double *val, double *vmin_out, int64_t *hist_out) {
    int64_t cs = (n + nchunk - 1) / nchunk;
    int64_t maxdeg = 1;
    for (int64_t j = 0; j < n; j++) if (rowptr[j + 1] - rowptr[j] > maxdeg) maxdeg = rowptr[j + 1] - rowptr[j];
    /* 2-hop candidates of one centre: at most min(hub2, maxdeg) arms x min(cap2, maxdeg) neighbours */
    int64_t kcap = (hub2 < maxdeg ? hub2 : maxdeg) * (cap2 < maxdeg ? cap2 : maxdeg);
    if (kcap < 1) kcap = 1;
    int64_t *hist = (int64_t *)calloc((size_t)nchunk * nb, sizeof(int64_t));
#pragma omp parallel
    {
        or_arm *H = (or_arm *)malloc(sizeof(or_arm) * (size_t)maxdeg);
        or_arm *K = (or_arm *)malloc(sizeof(or_arm) * (size_t)kcap);
#pragma omp for schedule(dynamic, 1)
        for (int64_t ch = 0; ch < nchunk; ch++)
            or_sep_chunk(fam, 0, ch, cs, n, rowptr, col, k, V, thresh, maxh, maxk, hub2, cap2, vtol, nb,
                         hist + ch * nb, 0.0, NULL, 0, NULL, NULL, H, K);
        free(H); free(K);
    }
    if (hist_out) memcpy(hist_out, hist, sizeof(int64_t) * (size_t)nchunk * nb);
    int64_t *cum = (int64_t *)calloc((size_t)nb + 1, sizeof(int64_t));
    for (int b = nb - 1; b >= 0; b--) {
        int64_t t = 0;
        for (int64_t ch = 0; ch < nchunk; ch++) t += hist[ch * nb + b];
        cum[b] = cum[b + 1] + t;
    }
    int b0 = nb;
    for (int b = 0; b < nb; b++) if (cum[b] <= cap) { b0 = b; break; }
    free(cum);
    if (b0 >= nb) { free(hist); *vmin_out = -1.0; return 0; }
    double vmin = vtol + (2.0 - vtol) * b0 / nb;
    *vmin_out = vmin;
    int64_t *counts = (int64_t *)calloc((size_t)nchunk, sizeof(int64_t));
    int64_t *offs = (int64_t *)calloc((size_t)nchunk + 1, sizeof(int64_t));
    for (int64_t ch = 0; ch < nchunk; ch++) {
        for (int b = b0; b < nb; b++) counts[ch] += hist[ch * nb + b];
        offs[ch + 1] = offs[ch] + counts[ch];
    }
    int64_t tot = offs[nchunk];
    int64_t *o2 = (int64_t *)calloc((size_t)(tot > 0 ? tot : 1) * 3, sizeof(int64_t));
    double *v2 = (double *)calloc((size_t)(tot > 0 ? tot : 1), sizeof(double));
#pragma omp parallel
    {
        or_arm *H = (or_arm *)malloc(sizeof(or_arm) * (size_t)maxdeg);
        or_arm *K = (or_arm *)malloc(sizeof(or_arm) * (size_t)kcap);
#pragma omp for schedule(dynamic, 1)
        for (int64_t ch = 0; ch < nchunk; ch++) {
            int64_t m = offs[ch];
            or_sep_chunk(fam, 1, ch, cs, n, rowptr, col, k, V, thresh, maxh, maxk, hub2, cap2, vtol, nb, NULL, vmin,
                         &m, offs[ch] + counts[ch], o2, v2, H, K);
        }
        free(H); free(K);
    }
    int64_t kept = 0;
    for (int64_t t = 0; t < tot; t++)
        if (v2[t] > 0.0) {
            out[3 * kept] = o2[3 * t]; out[3 * kept + 1] = o2[3 * t + 1]; out[3 * kept + 2] = o2[3 * t + 2];
            val[kept] = v2[t]; kept++;
        }
    free(hist); free(counts); free(offs); free(o2); free(v2);
    return kept;
}

/* ------------------------------------------------------------------------------------------------ exact reduction
 * reduce_exact.py:12-62. keys: m sorted keys u*n + v (u < v); rhash: the n hash weights of the twin candidate filter
 * (reduce_exact.py:31, numpy default_rng(11).integers(1, 2**62, n)). Union-find with par[max] = min (27-29), leaf
 * rule over the edges in key order (31-33), candidate twins (h[u] == h[v], d[u] == d[v], d^2 < 2m, d > 1) checked
 * exactly on the closed neighbourhoods in key order (34-37), groups = rank of the root among the sorted roots (38),
 * volumes, inside edges and the sorted unique group-pair keys with their counts (39-42).
 * Outputs: group_of (n), vol (n, first N used), kk / w (m, first nkeys used), stats = {N, nkeys, leaf_unions,
 * twin_unions, inside_edges}. Returns 0. */
static int64_t or_find(int64_t *par, int64_t x) {
    int64_t r = x;
    while (par[r] != r) r = par[r];
    while (par[x] != r) { int64_t nx = par[x]; par[x] = r; x = nx; }
    return r;
}

static void or_union(int64_t *par, int64_t a, int64_t b) {
    int64_t ra = or_find(par, a), rb = or_find(par, b);
    if (ra != rb) par[ra > rb ? ra : rb] = ra < rb ? ra : rb;
}

static int or_cmp_i64(const void *a, const void *b) {
    int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
    return x < y ? -1 : x > y;
}

int or_reduce(int64_t n, int64_t m, const int64_t *keys, const uint64_t *rhash, int64_t *group_of, int64_t *vol,
              int64_t *kk, double *w, int64_t *stats) {
    int64_t *d = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t *ptr = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t *adj = (int64_t *)malloc(sizeof(int64_t) * (size_t)(2 * m + 1));
    int64_t *fill = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    int64_t *par = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n + 1));
    uint64_t *h = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n + 1));
    for (int64_t e = 0; e < m; e++) { d[keys[e] / n]++; d[keys[e] % n]++; }
    for (int64_t i = 0; i < n; i++) ptr[i + 1] = ptr[i] + d[i];
    for (int64_t e = 0; e < m; e++) {
        int64_t u = keys[e] / n, v = keys[e] % n;
        adj[ptr[u] + fill[u]++] = v;
        adj[ptr[v] + fill[v]++] = u;
    }
    for (int64_t i = 0; i < n; i++) qsort(adj + ptr[i], (size_t)d[i], sizeof(int64_t), or_cmp_i64);
    for (int64_t i = 0; i < n; i++) par[i] = i;
    int64_t nL = 0, nT = 0;
    for (int64_t e = 0; e < m; e++) {                       /* rule L */
        int64_t u = keys[e] / n, v = keys[e] % n;
        if (d[u] == 1 || d[v] == 1) { or_union(par, u, v); nL++; }
    }
    for (int64_t i = 0; i < n; i++) h[i] = rhash[i];
    for (int64_t e = 0; e < m; e++) {
        int64_t u = keys[e] / n, v = keys[e] % n;
        h[u] += rhash[v]; h[v] += rhash[u];
    }
    for (int64_t e = 0; e < m; e++) {                       /* rule T */
        int64_t u = keys[e] / n, v = keys[e] % n;
        if (!(h[u] == h[v] && d[u] == d[v] && (double)d[u] * (double)d[u] < 2.0 * (double)m && d[u] > 1)) continue;
        /* N[u] == N[v] for adjacent u, v: the sorted rows agree once v is dropped from row u and u from row v */
        int64_t a = ptr[u], ae = ptr[u + 1], b = ptr[v], be = ptr[v + 1], same = 1;
        while (1) {
            if (a < ae && adj[a] == v) a++;
            if (b < be && adj[b] == u) b++;
            if (a == ae || b == be) { same = (a == ae && b == be); break; }
            if (adj[a] != adj[b]) { same = 0; break; }
            a++; b++;
        }
        if (same) { or_union(par, u, v); nT++; }
    }
    int64_t *root = fill;                                   /* reuse */
    for (int64_t x = 0; x < n; x++) root[x] = or_find(par, x);
    int64_t *rank = ptr;                                    /* reuse: rank of each root id */
    int64_t N = 0;
    for (int64_t x = 0; x < n; x++) rank[x] = (root[x] == x) ? N++ : -1;
    for (int64_t x = 0; x < n; x++) group_of[x] = rank[root[x]];
    for (int64_t g = 0; g < N; g++) vol[g] = 0;
    for (int64_t x = 0; x < n; x++) vol[group_of[x]] += d[x];
    int64_t C0e = 0, nc = 0;
    for (int64_t e = 0; e < m; e++) {
        int64_t gu = group_of[keys[e] / n], gv = group_of[keys[e] % n];
        if (gu == gv) { C0e++; continue; }
        adj[nc++] = (gu < gv ? gu : gv) * N + (gu < gv ? gv : gu);
    }
    qsort(adj, (size_t)nc, sizeof(int64_t), or_cmp_i64);
    int64_t nk = 0;
    for (int64_t t = 0; t < nc; t++) {
        if (nk > 0 && kk[nk - 1] == adj[t]) { w[nk - 1] += 1.0; continue; }
        kk[nk] = adj[t]; w[nk] = 1.0; nk++;
    }
    stats[0] = N; stats[1] = nk; stats[2] = nL; stats[3] = nT; stats[4] = C0e;
    free(d); free(ptr); free(adj); free(fill); free(par); free(h);
    return 0;
}
