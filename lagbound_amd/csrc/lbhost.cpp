// liblbhost: host bookkeeping of the search (include/lbhost.h). Each function has one pass over its inputs plus
// sorts of the new items only, in place of numpy's chains of temporaries.
#include "../../include/lbhost.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// Stable LSD radix sort of 64-bit keys (11-bit digits; a pass is skipped when every key has the same digit there),
// carrying an int64 payload when one is given. Signed order: the keys are taken with their sign bit flipped.
void radix_sort(std::vector<uint64_t> &k, std::vector<int64_t> *v) {
    const size_t n = k.size();
    if (n < 2) return;
    constexpr int B = 11, NB = 1 << B, PASSES = (64 + B - 1) / B;
    std::vector<uint64_t> hist((size_t)PASSES * NB, 0);
    for (size_t t = 0; t < n; t++) {
        const uint64_t x = k[t];
        for (int p = 0; p < PASSES; p++) hist[(size_t)p * NB + ((x >> (p * B)) & (NB - 1))]++;
    }
    std::vector<uint64_t> k2(n);
    std::vector<int64_t> v2(v ? n : 0);
    for (int p = 0; p < PASSES; p++) {
        uint64_t *h = hist.data() + (size_t)p * NB;
        bool trivial = false;
        for (int b = 0; b < NB; b++)
            if (h[b] == n) { trivial = true; break; }
        if (trivial) continue;
        uint64_t run = 0;
        for (int b = 0; b < NB; b++) { const uint64_t c = h[b]; h[b] = run; run += c; }
        const int sh = p * B;
        if (v) {
            for (size_t t = 0; t < n; t++) {
                const uint64_t o = h[(k[t] >> sh) & (NB - 1)]++;
                k2[o] = k[t];
                v2[o] = (*v)[t];
            }
            v->swap(v2);
        } else {
            for (size_t t = 0; t < n; t++) k2[h[(k[t] >> sh) & (NB - 1)]++] = k[t];
        }
        k.swap(k2);
    }
}

}  // namespace

extern "C" {

int64_t lbh_ws_merge(int64_t n, int64_t P, const int64_t *ptr, const int32_t *col, const uint8_t *keep, int64_t m,
                     const int64_t *ni, const int64_t *nR, int64_t *ptr_out, int32_t *col_out, int64_t *src_out) {
    if (n < 0 || P < 1 || m < 0 || !ptr || !ptr_out || (m && (!ni || !nR))) return -1;
    for (int64_t k = 0; k < m; k++)
        if (ni[k] < 0 || ni[k] >= n || nR[k] < 0 || nR[k] >= P) return -1;
    // per row: kept count, then the new entries bucketed by row in their given order (a counting sort: stable)
    std::vector<int64_t> cnt_new((size_t)n + 1, 0);
    for (int64_t k = 0; k < m; k++) cnt_new[(size_t)ni[k] + 1]++;
    for (int64_t i = 0; i < n; i++) cnt_new[(size_t)i + 1] += cnt_new[(size_t)i];
    std::vector<int64_t> newby((size_t)m);
    {
        std::vector<int64_t> at(cnt_new.begin(), cnt_new.end() - 1);
        for (int64_t k = 0; k < m; k++) newby[(size_t)at[(size_t)ni[k]]++] = k;
    }
    ptr_out[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        int64_t kept = 0;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) kept += keep[e] ? 1 : 0;
        ptr_out[i + 1] = ptr_out[i] + kept + (cnt_new[(size_t)i + 1] - cnt_new[(size_t)i]);
    }
    // per row: [kept in storage order, new in given order], stably sorted by R, in place in the output (an insertion
    // sort: rows hold a few entries each, and the kept ones are already in order)
    for (int64_t i = 0; i < n; i++) {
        int64_t o = ptr_out[i];
        const int64_t o0 = o;
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++)
            if (keep[e]) {
                if (col[e] < 0 || col[e] >= P) return -1;
                col_out[o] = col[e];
                src_out[o++] = e;
            }
        for (int64_t q = cnt_new[(size_t)i]; q < cnt_new[(size_t)i + 1]; q++) {
            const int64_t k = newby[(size_t)q];
            col_out[o] = (int32_t)nR[k];
            src_out[o++] = -1;
        }
        for (int64_t a = o0 + 1; a < o; a++) {
            const int32_t c = col_out[a];
            const int64_t sr = src_out[a];
            int64_t b = a;
            while (b > o0 && col_out[b - 1] > c) {      // strictly greater: equal keys keep their order (stable)
                col_out[b] = col_out[b - 1];
                src_out[b] = src_out[b - 1];
                b--;
            }
            col_out[b] = c;
            src_out[b] = sr;
        }
    }
    return ptr_out[n];
}

int lbh_take_f64(int64_t cnt, const int64_t *src, const double *a, int64_t na, double *out) {
    for (int64_t e = 0; e < cnt; e++) {
        if (src[e] >= na) return -1;
        out[e] = src[e] >= 0 ? a[src[e]] : 0.0;
    }
    return 0;
}

int lbh_take_u8(int64_t cnt, const int64_t *src, const uint8_t *a, int64_t na, uint8_t *out) {
    for (int64_t e = 0; e < cnt; e++) {
        if (src[e] >= na) return -1;
        out[e] = src[e] >= 0 ? a[src[e]] : 0;
    }
    return 0;
}

int lbh_argsort_bounded(int64_t m, const int64_t *keys, int64_t kmax, int64_t *perm_out) {
    if (m < 0 || kmax < 1 || (m && (!keys || !perm_out))) return -1;
    std::vector<int64_t> at((size_t)kmax + 1, 0);
    for (int64_t t = 0; t < m; t++) {
        if (keys[t] < 0 || keys[t] >= kmax) return -1;
        at[(size_t)keys[t] + 1]++;
    }
    for (int64_t k = 0; k < kmax; k++) at[(size_t)k + 1] += at[(size_t)k];
    for (int64_t t = 0; t < m; t++) perm_out[at[(size_t)keys[t]]++] = t;
    return 0;
}

int lbh_argsort_i64(int64_t m, const int64_t *keys, int64_t *perm_out) {
    if (m < 0 || (m && (!keys || !perm_out))) return -1;
    const uint64_t flip = (uint64_t)1 << 63;
    std::vector<uint64_t> k((size_t)m);
    std::vector<int64_t> v((size_t)m);
    for (int64_t t = 0; t < m; t++) { k[(size_t)t] = (uint64_t)keys[t] ^ flip; v[(size_t)t] = t; }
    radix_sort(k, &v);
    std::memcpy(perm_out, v.data(), sizeof(int64_t) * (size_t)m);
    return 0;
}

int64_t lbh_row_gather(int64_t n, const int64_t *ptr, const int64_t *order, int64_t *nptr_out, int64_t *idx_out) {
    if (n < 0 || !ptr || (n && (!order || !nptr_out))) return -1;
    nptr_out[0] = 0;
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = order[k];
        if (i < 0 || i >= n || ptr[i + 1] < ptr[i]) return -1;
        nptr_out[k + 1] = nptr_out[k] + (ptr[i + 1] - ptr[i]);
    }
    // a repeated row (order not a permutation) would write past the caller's ptr[n] - ptr[0] entries
    if (nptr_out[n] != ptr[n] - ptr[0] || (nptr_out[n] && !idx_out)) return -1;
    int64_t o = 0;
    for (int64_t k = 0; k < n; k++) {
        const int64_t i = order[k];
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) idx_out[o++] = e;
    }
    return o;
}

int64_t lbh_coo_sum(int64_t N, int64_t m, const int64_t *r, const int64_t *c, const double *v, int64_t *indptr_out,
                    int64_t *indices_out, double *sum_out, double *abs_out, double *cnt_out) {
    if (N < 0 || m < 0 || !indptr_out || (m && (!r || !c || !indices_out))) return -1;
    std::vector<int64_t> at((size_t)N + 1, 0);
    for (int64_t t = 0; t < m; t++) {
        if (r[t] < 0 || r[t] >= N || c[t] < 0 || c[t] >= N) return -1;
        at[(size_t)r[t] + 1]++;
    }
    for (int64_t i = 0; i < N; i++) at[(size_t)i + 1] += at[(size_t)i];
    // (column, input position) of every entry, bucketed by row in input order
    std::vector<std::pair<int64_t, int64_t>> cp((size_t)m);
    {
        std::vector<int64_t> fill(at.begin(), at.end() - 1);
        for (int64_t t = 0; t < m; t++) cp[(size_t)fill[(size_t)r[t]]++] = {c[t], t};
    }
    int64_t w = 0;
    indptr_out[0] = 0;
    for (int64_t i = 0; i < N; i++) {
        std::pair<int64_t, int64_t> *b = cp.data() + at[(size_t)i], *e = cp.data() + at[(size_t)i + 1];
        if (e - b <= 32) {                                // short rows: insertion sort, equal columns keep their order
            for (std::pair<int64_t, int64_t> *q = b + 1; q < e; q++) {
                const std::pair<int64_t, int64_t> x = *q;
                std::pair<int64_t, int64_t> *y = q;
                while (y > b && (y - 1)->first > x.first) { *y = *(y - 1); y--; }
                *y = x;
            }
        } else {
            std::sort(b, e);                              // by column, then input position
        }
        for (std::pair<int64_t, int64_t> *q = b; q < e;) {
            const int64_t col = q->first;
            double sv = 0.0, sa = 0.0, sn = 0.0;
            for (; q < e && q->first == col; q++) {       // duplicates summed in input order
                const double x = v ? v[q->second] : 1.0;
                sv += x;
                sa += std::fabs(x);
                sn += 1.0;
            }
            indices_out[w] = col;
            if (sum_out) sum_out[w] = sv;
            if (abs_out) abs_out[w] = sa;
            if (cnt_out) cnt_out[w] = sn;
            w++;
        }
        indptr_out[i + 1] = w;
    }
    return w;
}

int64_t lbh_sym_from_lower(int64_t N, const int64_t *ptr, const int64_t *idx, const double *val, int64_t *ptr_out,
                           int64_t *idx_out, double *val_out) {
    if (N < 0 || !ptr || !ptr_out) return -1;
    // row k of the result: row k of the lower triangle (columns <= k, ascending), then the entries (i, k) with i > k
    // in ascending i, which a pass over the rows in order drops into bucket k already sorted
    std::vector<int64_t> up((size_t)N + 1, 0);
    for (int64_t i = 0; i < N; i++)
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            const int64_t j = idx[e];
            if (j < 0 || j > i) return -1;
            if (e > ptr[i] && idx[e - 1] >= j) return -1;  // columns strictly ascending
            if (j < i) up[(size_t)j + 1]++;
        }
    ptr_out[0] = 0;
    for (int64_t k = 0; k < N; k++) ptr_out[k + 1] = ptr_out[k] + (ptr[k + 1] - ptr[k]) + up[(size_t)k + 1];
    std::vector<int64_t> fill((size_t)N);
    for (int64_t k = 0; k < N; k++) {
        int64_t o = ptr_out[k];
        for (int64_t e = ptr[k]; e < ptr[k + 1]; e++) { idx_out[o] = idx[e]; val_out[o] = val[e]; o++; }
        fill[(size_t)k] = o;
    }
    for (int64_t i = 0; i < N; i++)
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            const int64_t j = idx[e];
            if (j < i) {
                const int64_t o = fill[(size_t)j]++;
                idx_out[o] = i;
                val_out[o] = val[e];
            }
        }
    return ptr_out[N];
}

int lbh_tie_order(int64_t m, const double *sc, const int64_t *vi, const int64_t *vR, int64_t *perm_out) {
    if (m < 0 || (m && (!sc || !vi || !vR || !perm_out))) return -1;
    for (int64_t t = 0; t + 1 < m; t++)
        if (!(sc[t] >= sc[t + 1])) return 1;         // not descending (or a NaN): the caller sorts in full
    for (int64_t t = 0; t < m; t++) perm_out[t] = t;
    for (int64_t a = 0; a < m;) {
        int64_t b = a + 1;
        while (b < m && sc[b] == sc[a]) b++;
        if (b - a > 1)
            std::sort(perm_out + a, perm_out + b, [vi, vR](int64_t x, int64_t y) {
                return vi[x] != vi[y] ? vi[x] < vi[y] : (vR[x] != vR[y] ? vR[x] < vR[y] : x < y);
            });
        a = b;
    }
    return 0;
}

int64_t lbh_new_triangles(int64_t n, int64_t m, const int64_t *cand, int64_t T, const int64_t *pool, int64_t *idx_out) {
    if (n < 0 || m < 0 || T < 0) return -1;
    // the packed key in int64 arithmetic with its wrap-around (computed in uint64, read back as int64)
    // (sign bit flipped, so that unsigned order is the int64 order)
    const uint64_t un = (uint64_t)n, flip = (uint64_t)1 << 63;
    auto key = [un, flip](const int64_t *t) -> uint64_t {
        return (((uint64_t)t[0] * un + (uint64_t)t[1]) * un + (uint64_t)t[2]) ^ flip;
    };
    std::vector<uint64_t> kc((size_t)m);
    std::vector<int64_t> ic((size_t)m);
    for (int64_t k = 0; k < m; k++) { kc[(size_t)k] = key(cand + 3 * k); ic[(size_t)k] = k; }
    radix_sort(kc, &ic);                             // stable: the first occurrence leads each run of equal keys
    std::vector<uint64_t> kp((size_t)T);
    for (int64_t t = 0; t < T; t++) kp[(size_t)t] = key(pool + 3 * t);
    radix_sort(kp, nullptr);
    int64_t w = 0;
    size_t p = 0;
    for (size_t a = 0; a < kc.size();) {
        size_t b = a + 1;
        while (b < kc.size() && kc[b] == kc[a]) b++;
        while (p < kp.size() && kp[p] < kc[a]) p++;   // both ascending: one merge pass
        if (!(p < kp.size() && kp[p] == kc[a])) idx_out[w++] = ic[a];
        a = b;
    }
    return w;
}

}  // extern "C"
