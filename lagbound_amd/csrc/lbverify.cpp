// liblbverify: the verifier's own PSD proof of a lifted certificate matrix (the role of verify_cert.py:292-386,
// `_etree` / `_colcounts` / `_ldl` / `psd_ldl`): a simplicial up-looking LDL^T of P A P^T that carries, for every entry
// of the factor, a running bound on its residual, so that
//
//     |A - L D L^T|_kj <= gamma(h_k + 3) (|A_kj| + sum_i |L_ji| |y_i|) + u |y_j|        (j < k)
//     |A - L D L^T|_kk <= gamma(h_k + 3) (|A_kk| + sum_j |L_kj| |y_j|)
//
// where y_j is the value of row k's entry j before the division by D_j (y_j = L_kj D_j up to one rounding) and h_k is
// the number of nonzeros of row k of L plus one: every such entry is a sum of at most h_k rounded products. D > 0 then
// makes L D L^T positive semidefinite exactly, and the sum of the bounds over the full matrix is the PSD price.
//
// The algorithm is the up-looking factorization of T. A. Davis (Direct Methods for Sparse Linear Systems, SIAM 2006,
// ch. 4; the LDL package, ACM TOMS 31(4), 2005): the elimination tree, the row patterns by reaching up the tree from
// the nonzeros of each column of the upper triangle, and each row of L by a sparse triangular solve against the rows
// before it. It shares no code with the certifier's right-looking factorization (lbcert.cpp).
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/lbverify.h"

namespace {

const double U = 1.0 / 9007199254740992.0;   // 2^-53

inline double gam(double k) { return k * U / (1.0 - k * U); }

}  // namespace

extern "C" {

int lbv_ldl(int64_t N, const int64_t *ptr, const int32_t *idx, const double *val, const int64_t *perm,
            int64_t max_nnz, double *Dout, double *out, int64_t *iout, int64_t *Lp_out, int32_t *Li_out,
            double *Lx_out) {
    if (N < 0 || !ptr || (N > 0 && (!idx || !val || !perm || !out || !iout))) return -1;
    std::memset(out, 0, 4 * sizeof(double));
    std::memset(iout, 0, 4 * sizeof(int64_t));
    if (N == 0) return 0;
    std::vector<int64_t> pinv((size_t)N, -1);
    for (int64_t k = 0; k < N; k++) {
        const int64_t o = perm[k];
        if (o < 0 || o >= N || pinv[(size_t)o] >= 0) return -1;   // not a permutation
        pinv[(size_t)o] = k;
    }
    // upper triangle of P A P^T by columns: column k holds (i, a_ik) for i <= k
    std::vector<int64_t> cp((size_t)N + 1, 0);
    for (int64_t k = 0; k < N; k++) {
        const int64_t o = perm[k];
        for (int64_t p = ptr[o]; p < ptr[o + 1]; p++) {
            const int64_t j = idx[p];
            if (j < 0 || j >= N) return -1;
            if (pinv[(size_t)j] <= k) cp[(size_t)k + 1]++;
        }
    }
    for (int64_t k = 0; k < N; k++) cp[(size_t)k + 1] += cp[(size_t)k];
    std::vector<int64_t> ci((size_t)cp[(size_t)N]);
    std::vector<double> cx((size_t)cp[(size_t)N]);
    {
        std::vector<int64_t> fill(cp.begin(), cp.end() - 1);
        for (int64_t k = 0; k < N; k++) {
            const int64_t o = perm[k];
            for (int64_t p = ptr[o]; p < ptr[o + 1]; p++) {
                const int64_t i = pinv[(size_t)idx[p]];
                if (i <= k) { ci[(size_t)fill[(size_t)k]] = i; cx[(size_t)fill[(size_t)k]++] = val[p]; }
            }
        }
    }
    // elimination tree (ancestor path compression) and the column counts of L by reaching up from each column
    std::vector<int64_t> parent((size_t)N, -1), anc((size_t)N, -1), flag((size_t)N, -1), cnt((size_t)N, 0);
    for (int64_t k = 0; k < N; k++) {
        for (int64_t p = cp[(size_t)k]; p < cp[(size_t)k + 1]; p++) {
            int64_t i = ci[(size_t)p];
            while (i != -1 && i < k) {
                const int64_t nx = anc[(size_t)i];
                anc[(size_t)i] = k;
                if (nx == -1) { parent[(size_t)i] = k; break; }
                i = nx;
            }
        }
    }
    int64_t nnzL = 0, hmax = 0;
    std::vector<int64_t> rowlen((size_t)N, 0);
    for (int64_t k = 0; k < N; k++) {
        flag[(size_t)k] = k;
        int64_t len = 0;
        for (int64_t p = cp[(size_t)k]; p < cp[(size_t)k + 1]; p++) {
            for (int64_t i = ci[(size_t)p]; i < k && flag[(size_t)i] != k; i = parent[(size_t)i]) {
                flag[(size_t)i] = k;
                cnt[(size_t)i]++;
                len++;
            }
        }
        rowlen[(size_t)k] = len;
        nnzL += len;
        if (len + 1 > hmax) hmax = len + 1;
    }
    iout[1] = nnzL;
    iout[2] = hmax;
    double flops = 0.0;
    for (int64_t j = 0; j < N; j++) flops += (double)cnt[(size_t)j] * (double)cnt[(size_t)j];
    out[2] = flops;
    if (max_nnz > 0 && nnzL > max_nnz) return -2;
    std::vector<int64_t> Lp((size_t)N + 1, 0), Lnz((size_t)N, 0);
    for (int64_t j = 0; j < N; j++) Lp[(size_t)j + 1] = Lp[(size_t)j] + cnt[(size_t)j];
    std::vector<int32_t> Li((size_t)nnzL);
    std::vector<double> Lx((size_t)nnzL), D((size_t)N), y((size_t)N, 0.0), ya((size_t)N, 0.0);
    std::vector<int64_t> stack((size_t)N);
    std::fill(flag.begin(), flag.end(), -1);
    double price_off = 0.0, price_diag = 0.0, sum_abs_a = 0.0;
    for (int64_t k = 0; k < N; k++) {
        // row pattern of L (topological order: each reach is pushed onto the end of the stack in reverse)
        int64_t top = N;
        flag[(size_t)k] = k;
        for (int64_t p = cp[(size_t)k]; p < cp[(size_t)k + 1]; p++) {
            const int64_t i0 = ci[(size_t)p];
            y[(size_t)i0] += cx[(size_t)p];
            ya[(size_t)i0] += std::fabs(cx[(size_t)p]);
            sum_abs_a += (i0 == k ? 1.0 : 2.0) * std::fabs(cx[(size_t)p]);
            int64_t len = 0;
            for (int64_t i = i0; i < k && flag[(size_t)i] != k; i = parent[(size_t)i]) {
                stack[(size_t)len++] = i;
                flag[(size_t)i] = k;
            }
            while (len > 0) stack[(size_t)--top] = stack[(size_t)--len];
        }
        const double g = gam((double)(rowlen[(size_t)k] + 4));   // h_k + 3, h_k = row length + 1
        double dk = y[(size_t)k], dka = ya[(size_t)k];
        y[(size_t)k] = 0.0;
        ya[(size_t)k] = 0.0;
        for (int64_t t = top; t < N; t++) {
            const int64_t j = stack[(size_t)t];
            const double yj = y[(size_t)j], yaj = ya[(size_t)j];
            y[(size_t)j] = 0.0;
            ya[(size_t)j] = 0.0;
            for (int64_t p = Lp[(size_t)j]; p < Lp[(size_t)j] + Lnz[(size_t)j]; p++) {
                const double pr = Lx[(size_t)p] * yj;
                y[(size_t)Li[(size_t)p]] -= pr;
                ya[(size_t)Li[(size_t)p]] += std::fabs(pr);
            }
            const double lkj = yj / D[(size_t)j];
            const double pr = lkj * yj;
            dk -= pr;
            dka += std::fabs(pr);
            price_off += 2.0 * (g * yaj + U * std::fabs(yj));
            const int64_t q = Lp[(size_t)j] + Lnz[(size_t)j]++;
            Li[(size_t)q] = (int32_t)k;
            Lx[(size_t)q] = lkj;
        }
        price_diag += g * dka;
        if (!(dk > 0.0) || !std::isfinite(dk) || !std::isfinite(dka)) {
            out[0] = price_off + price_diag;
            out[3] = sum_abs_a;
            iout[0] = k + 1;
            return (int)1;
        }
        D[(size_t)k] = dk;
    }
    out[0] = price_off + price_diag;
    out[1] = 0.0;
    out[3] = sum_abs_a;
    iout[3] = N;
    if (Dout) std::memcpy(Dout, D.data(), sizeof(double) * (size_t)N);
    if (Lp_out && Li_out && Lx_out) {
        std::memcpy(Lp_out, Lp.data(), sizeof(int64_t) * ((size_t)N + 1));
        std::memcpy(Li_out, Li.data(), sizeof(int32_t) * (size_t)nnzL);
        std::memcpy(Lx_out, Lx.data(), sizeof(double) * (size_t)nnzL);
    }
    return 0;
}

}  // extern "C"
