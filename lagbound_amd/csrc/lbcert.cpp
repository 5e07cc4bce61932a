// lbcert.cpp -- host part of the sparse FP64 certificate (liblbcert.so, plain C ABI, see include/lbcert.h).
//
// The certificate proves that the lifted matrix Mt (symmetric, N x N, assembled by lagbound_amd/sparse_cert.py) is
// positive semidefinite up to a priced residual, as the reference's certifier and verifier do with CHOLMOD or their
// own simplicial LDL^T (certify_lifted_vc.py:391-449; verify_cert.py:292-386). Here the factorization is split:
//   1. a fill-reducing elimination order (minimum degree on a quotient graph with an upper-bound degree, in the
//      spirit of Amestoy, Davis & Duff, SIAM J. Matrix Anal. Appl. 17(4), 1996): rows of very high degree are never
//      pivoted, and the elimination stops when every remaining row's degree bound exceeds a cap -- those rows are the
//      dense core;
//   2. a right-looking LDL^T of the sparse pivots on the host, whose column patterns are the element member sets of
//      the ordering (exact: a pivot's element holds exactly its neighbours in the elimination graph);
//   3. the dense core's Schur complement S = A22 - L21 D1 L21^T, formed and Cholesky-factored on the device
//      (lb_cert_core in liblbhip).
// Rounding: every entry of L D L^T - A in the sparse part, and of S_computed - S_exact, is a sum of at most h products
// of at most three factors, so |error| <= gamma(h + 4) (|A_ij| + (|L| |D| |L^T|)_ij) for any summation order; summed
// over all entries this is gamma(h_max + 4) (sum |A| + sum_k |D_k| (sum_i |L_ik|)^2) (L_kk = 1 included). This file
// returns the two sums and h_max; the caller rounds them up and adds the dense core's own bound.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "../../include/lbcert.h"

namespace {

struct Fact {
    int64_t N = 0;
    std::vector<int64_t> order;          // elimination order: sparse pivots, then the core (original ids)
    int64_t nsp = 0;                     // number of sparse pivots
    std::vector<int64_t> pos;            // original id -> position in `order`
    // column patterns of the sparse pivots (positions > the pivot's, ascending) and their values (L, after the solve)
    std::vector<int64_t> lptr;
    std::vector<int32_t> lidx;
    std::vector<double> lval;
    std::vector<double> D;               // pivots (sparse part)
    // core: entries of A among core rows (lower, positions relative to the core start) and per pivot the core part of
    // its column
    std::vector<int32_t> ai, aj;
    std::vector<double> av;
    std::vector<int64_t> cptr;
    std::vector<int32_t> cidx;
    std::vector<double> cval;
    int status = 0;                      // 0 ok, > 0: sparse pivot (1-based position) not positive
    double sumA = 0.0, sumLDL = 0.0;     // sum |A| over all entries; sum_k |D_k| (sum_i |L_ik|)^2 (sparse pivots)
    int64_t hmax = 0, nnzL = 0, maxcol = 0;
    double minD = INFINITY;
};

// Minimum-degree order on the quotient graph (variables, elements). A pivot's element is the set of variables adjacent
// to it at its elimination; elements adjacent to the pivot are absorbed into it. Degree of a variable: |A_i| +
// sum over its elements e of (|L_e| - 1), an upper bound on its external degree. Inert rows (initial degree above
// dense_deg) are kept in the adjacency of the others (so the patterns are complete) but never pivoted or updated.
void md_order(int64_t N, const int64_t *ptr, const int32_t *idx, int64_t dense_deg, int64_t core_deg, Fact &F,
              const int64_t *given = nullptr) {
    std::vector<std::vector<int32_t>> A(N), E(N), Le(N);
    std::vector<uint8_t> st(N, 0);                      // 0 variable, 1 element (live), 2 absorbed, 3 inert
    std::vector<int64_t> deg(N, 0), mark(N, -1);
    std::vector<int64_t> lcnt(N, 0);                     // an element's members that are not inert (fixed once made)
    std::vector<uint8_t> exact(N, 0);                    // deg[i] is i's exact external degree (since its last update)
    for (int64_t i = 0; i < N; i++) {
        const int64_t d = ptr[i + 1] - ptr[i];
        if (d > dense_deg) st[i] = 3;
    }
    for (int64_t i = 0; i < N; i++) {
        if (st[i] == 3) continue;
        A[i].reserve(ptr[i + 1] - ptr[i]);
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++)
            if (idx[e] != i) A[i].push_back(idx[e]);
        int64_t dg = 0;
        for (int32_t j : A[i]) dg += st[j] != 3;
        deg[i] = dg;
    }
    // degree buckets (doubly linked lists); degrees are capped at N
    std::vector<int64_t> head(N + 1, -1), nxt(N, -1), prv(N, -1);
    auto insert = [&](int64_t i) {
        const int64_t d = std::min<int64_t>(deg[i], N);
        nxt[i] = head[d]; prv[i] = -1;
        if (head[d] >= 0) prv[head[d]] = i;
        head[d] = i;
    };
    auto remove = [&](int64_t i) {
        const int64_t d = std::min<int64_t>(deg[i], N);
        if (prv[i] >= 0) nxt[prv[i]] = nxt[i]; else head[d] = nxt[i];
        if (nxt[i] >= 0) prv[nxt[i]] = prv[i];
        nxt[i] = prv[i] = -1;
    };
    for (int64_t i = 0; i < N; i++) if (st[i] == 0) { exact[i] = 1; insert(i); }
    F.order.clear();
    F.order.reserve(N);
    F.lptr.assign(1, 0);
    F.lidx.clear();
    std::vector<int32_t> Lp;
    int64_t mindeg = 0, tag = 0, gi = 0;
    for (;;) {
        int64_t p;
        if (given) {
            // the caller's order: the next variable; one whose exact degree is above the cap goes to the core
            while (gi < N && st[given[gi]] != 0) gi++;
            if (gi >= N) break;
            p = given[gi++];
            remove(p);
            tag++;
            mark[p] = tag;
            int64_t dx = 0;
            for (int32_t v : A[p]) if (st[v] == 0 && mark[v] != tag) { mark[v] = tag; dx++; }
            for (int32_t e : E[p]) for (int32_t v : Le[e]) if (st[v] == 0 && mark[v] != tag) { mark[v] = tag; dx++; }
            if (dx > core_deg) { st[p] = 3; continue; }
        } else {
        while (mindeg <= N && head[mindeg] < 0) mindeg++;
        if (mindeg > N) break;
        p = head[mindeg];
        remove(p);
        if (!exact[p]) {
            // the key is an upper bound: replace it by the exact external degree (non-inert variables adjacent to p
            // directly or through its elements) and take p only if it is still the smallest key
            tag++;
            mark[p] = tag;
            int64_t dx = 0;
            for (int32_t v : A[p]) if (st[v] == 0 && mark[v] != tag) { mark[v] = tag; dx++; }
            for (int32_t e : E[p]) for (int32_t v : Le[e]) if (st[v] == 0 && mark[v] != tag) { mark[v] = tag; dx++; }
            exact[p] = 1;
            if (dx < deg[p]) {
                deg[p] = dx;
                insert(p);
                if (dx < mindeg) mindeg = dx;
                continue;
            }
        }
        if (deg[p] > core_deg) { insert(p); break; }
        }
        // element of p: variables adjacent to p directly or through its elements
        tag++;
        mark[p] = tag;
        Lp.clear();
        for (int32_t v : A[p]) if (st[v] == 0 || st[v] == 3) { if (mark[v] != tag) { mark[v] = tag; Lp.push_back(v); } }
        for (int32_t e : E[p]) {
            for (int32_t v : Le[e]) if ((st[v] == 0 || st[v] == 3) && mark[v] != tag) { mark[v] = tag; Lp.push_back(v); }
            st[e] = 2;                                  // absorbed into p
            std::vector<int32_t>().swap(Le[e]);
        }
        st[p] = 1;
        std::vector<int32_t>().swap(A[p]);
        std::vector<int32_t>().swap(E[p]);
        F.order.push_back(p);
        F.lidx.insert(F.lidx.end(), Lp.begin(), Lp.end());   // the pivot's column pattern (original ids for now)
        F.lptr.push_back((int64_t)F.lidx.size());
        Le[p] = Lp;
        int64_t lc = 0;
        for (int32_t v : Lp) lc += st[v] != 3;
        lcnt[p] = lc;
        // neighbours: drop absorbed elements and the adjacency p's element now covers, add p, new degree bound
        for (int32_t v : Lp) {
            if (st[v] == 3) continue;
            remove(v);
            auto &Ev = E[v];
            size_t w = 0;
            for (int32_t e : Ev) if (st[e] == 1) Ev[w++] = e;
            Ev.resize(w);
            Ev.push_back((int32_t)p);
            auto &Av = A[v];
            w = 0;
            int64_t dA = 0;
            for (int32_t u : Av) {
                if (st[u] == 1 || st[u] == 2 || mark[u] == tag) continue;   // eliminated, or covered by element p
                Av[w++] = u;
                dA += st[u] != 3;
            }
            Av.resize(w);
            int64_t dg = dA;
            for (int32_t e : Ev) dg += lcnt[e] - 1;                 // inert rows do not count
            deg[v] = std::min<int64_t>(dg, N);
            exact[v] = 0;
            insert(v);
            if (deg[v] < mindeg) mindeg = deg[v];
        }
    }
    F.nsp = (int64_t)F.order.size();
    for (int64_t i = 0; i < N; i++) if (st[i] == 0) F.order.push_back(i);
    for (int64_t i = 0; i < N; i++) if (st[i] == 3) F.order.push_back(i);
    F.pos.assign(N, 0);
    for (int64_t k = 0; k < N; k++) F.pos[F.order[k]] = k;
    // column patterns in positions, ascending
    for (int64_t k = 0; k < F.nsp; k++) {
        int32_t *b = F.lidx.data() + F.lptr[k], *e = F.lidx.data() + F.lptr[k + 1];
        for (int32_t *q = b; q < e; q++) *q = (int32_t)F.pos[*q];
        std::sort(b, e);
    }
}

// Right-looking LDL^T of the sparse pivots; updates between two core rows are left to the device (the core part of
// each column is exported with its pivot).
void numeric(int64_t N, const int64_t *ptr, const int32_t *idx, const double *val, Fact &F) {
    const int64_t nsp = F.nsp;
    F.lval.assign(F.lidx.size(), 0.0);
    F.D.assign(nsp, 0.0);
    std::vector<int64_t> rowcnt(N, 0);
    // scatter A: diagonal to D, below-diagonal entries (in positions) to their column's slot (binary search), core
    // entries to the core list; sum |A| over all stored entries (the full symmetric matrix)
    F.sumA = 0.0;
    for (int64_t i = 0; i < N; i++) {
        const int64_t pi = F.pos[i];
        for (int64_t e = ptr[i]; e < ptr[i + 1]; e++) {
            const int64_t j = idx[e], pj = F.pos[j];
            const double a = val[e];
            F.sumA += std::fabs(a);
            if (pj > pi) continue;                         // each unordered pair once: from its later row
            if (pj == pi) {
                if (pi < nsp) F.D[pi] += a;
                else { F.ai.push_back((int32_t)(pi - nsp)); F.aj.push_back((int32_t)(pi - nsp)); F.av.push_back(a); }
                continue;
            }
            if (pj < nsp) {                                // column pj holds row pi
                const int32_t *b = F.lidx.data() + F.lptr[pj], *en = F.lidx.data() + F.lptr[pj + 1];
                const int32_t *hit = std::lower_bound(b, en, (int32_t)pi);
                F.lval[hit - F.lidx.data()] += a;
            } else {
                F.ai.push_back((int32_t)(pi - nsp)); F.aj.push_back((int32_t)(pj - nsp)); F.av.push_back(a);
            }
        }
    }
    F.cptr.assign(nsp + 1, 0);
    F.sumLDL = 0.0;
    F.status = 0;
    for (int64_t k = 0; k < nsp; k++) {
        const double dk = F.D[k];
        if (!(dk > 0.0) || !std::isfinite(dk)) { F.status = (int)std::min<int64_t>(k + 1, INT32_MAX); return; }
        F.minD = std::min(F.minD, dk);
        const int64_t b = F.lptr[k], e = F.lptr[k + 1];
        double colsum = 1.0;                               // L_kk
        for (int64_t t = b; t < e; t++) {
            F.lval[t] /= dk;
            colsum += std::fabs(F.lval[t]);
            rowcnt[F.lidx[t]]++;
        }
        F.sumLDL += dk * colsum * colsum;
        // updates: row i, column j (positions, j <= i), both in the pattern of k
        for (int64_t t = b; t < e; t++) {
            const int32_t j = F.lidx[t];
            if (j >= nsp) break;                           // the rest of the pattern is core (sorted positions)
            const double lj = F.lval[t] * dk;
            F.D[j] -= lj * F.lval[t];
            const int32_t *cb = F.lidx.data() + F.lptr[j], *ce = F.lidx.data() + F.lptr[j + 1];
            const int32_t *cur = cb;
            for (int64_t s = t + 1; s < e; s++) {
                const int32_t i = F.lidx[s];
                cur = std::lower_bound(cur, ce, i);        // i is in column j's pattern (elimination-graph closure)
                F.lval[cur - F.lidx.data()] -= F.lval[s] * lj;
            }
        }
        // the core part of column k, for the device
        int64_t c0 = b;
        while (c0 < e && F.lidx[c0] < nsp) c0++;
        F.cptr[k + 1] = F.cptr[k] + (e - c0);
    }
    F.cidx.resize(F.cptr[nsp]);
    F.cval.resize(F.cptr[nsp]);
    for (int64_t k = 0; k < nsp; k++) {
        const int64_t b = F.lptr[k], e = F.lptr[k + 1];
        int64_t c0 = b;
        while (c0 < e && F.lidx[c0] < nsp) c0++;
        for (int64_t t = c0, o = F.cptr[k]; t < e; t++, o++) { F.cidx[o] = F.lidx[t] - (int32_t)nsp; F.cval[o] = F.lval[t]; }
    }
    int64_t hm = 0, mc = 0;
    for (int64_t i = 0; i < N; i++) hm = std::max(hm, rowcnt[i]);
    for (int64_t k = 0; k < nsp; k++) mc = std::max(mc, F.lptr[k + 1] - F.lptr[k]);
    F.hmax = hm + 1;                                       // products per entry, plus the entry of A
    F.maxcol = mc;
    F.nnzL = (int64_t)F.lidx.size();
}

}  // namespace

extern "C" {

int lbc_factor(int64_t N, const int64_t *ptr, const int32_t *idx, const double *val, int64_t dense_deg,
               int64_t core_deg, const int64_t *order, lbc_fact **out) {
    if (!out) return -1;
    *out = nullptr;
    if (N < 1 || !ptr || !idx || !val || ptr[0] != 0) return -1;
    for (int64_t i = 0; i < N; i++) if (ptr[i + 1] < ptr[i]) return -1;
    for (int64_t e = 0; e < ptr[N]; e++) if (idx[e] < 0 || idx[e] >= N) return -1;
    if (order) {
        std::vector<uint8_t> seen(N, 0);
        for (int64_t k = 0; k < N; k++) {
            if (order[k] < 0 || order[k] >= N || seen[order[k]]) return -1;
            seen[order[k]] = 1;
        }
    }
    Fact *F = new Fact();
    F->N = N;
    md_order(N, ptr, idx, dense_deg, core_deg, *F, order);
    numeric(N, ptr, idx, val, *F);
    *out = reinterpret_cast<lbc_fact *>(F);
    return F->status;
}

int lbc_md_order(int64_t N, const int64_t *ptr, const int32_t *idx, int64_t dense_deg, int64_t core_deg,
                 int64_t *order, int64_t *nsparse) {
    if (N < 1 || !ptr || !idx || !order || ptr[0] != 0) return -1;
    for (int64_t i = 0; i < N; i++) if (ptr[i + 1] < ptr[i]) return -1;
    for (int64_t e = 0; e < ptr[N]; e++) if (idx[e] < 0 || idx[e] >= N) return -1;
    Fact F;
    F.N = N;
    md_order(N, ptr, idx, dense_deg, core_deg, F);
    std::copy(F.order.begin(), F.order.end(), order);
    if (nsparse) *nsparse = F.nsp;
    return 0;
}

void lbc_free(lbc_fact *f) { delete reinterpret_cast<Fact *>(f); }

int lbc_info(const lbc_fact *f, int64_t *iout, double *dout) {
    if (!f) return -1;
    const Fact *F = reinterpret_cast<const Fact *>(f);
    if (iout) {
        iout[0] = F->N; iout[1] = F->nsp; iout[2] = F->N - F->nsp; iout[3] = F->nnzL; iout[4] = F->hmax;
        iout[5] = F->maxcol; iout[6] = (int64_t)F->av.size(); iout[7] = F->cptr.empty() ? 0 : F->cptr.back();
        iout[8] = F->status;
    }
    if (dout) { dout[0] = F->sumA; dout[1] = F->sumLDL; dout[2] = F->minD; }
    return 0;
}

int lbc_order(const lbc_fact *f, int64_t *order) {
    if (!f || !order) return -1;
    const Fact *F = reinterpret_cast<const Fact *>(f);
    std::copy(F->order.begin(), F->order.end(), order);
    return 0;
}

int lbc_core(const lbc_fact *f, int32_t *ai, int32_t *aj, double *av, int64_t *cptr, int32_t *cidx, double *cval,
             double *cd) {
    if (!f) return -1;
    const Fact *F = reinterpret_cast<const Fact *>(f);
    if (F->status) return -1;
    if (ai) std::copy(F->ai.begin(), F->ai.end(), ai);
    if (aj) std::copy(F->aj.begin(), F->aj.end(), aj);
    if (av) std::copy(F->av.begin(), F->av.end(), av);
    if (cptr) std::copy(F->cptr.begin(), F->cptr.end(), cptr);
    if (cidx) std::copy(F->cidx.begin(), F->cidx.end(), cidx);
    if (cval) std::copy(F->cval.begin(), F->cval.end(), cval);
    if (cd) std::copy(F->D.begin(), F->D.end(), cd);
    return 0;
}

}  // extern "C"
