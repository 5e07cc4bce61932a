// liblbhip: MI355X (gfx950) kernels of lagbound's augmented-Lagrangian low-rank SDP search, behind the C ABI of
// include/lbhip.h.
//
// One f+grad evaluation (alm_vcs.py:1496-1533 `augmented` + alm_vcs.py:1299-1312 `CV`) is a chain of kernels on the
// context's stream:
//   lb_k_colsum   z = d^T V and the cluster sums Y = W V of every label set (pair + VCS levels), FP64, one pass
//                 over V (cluster_sums, alm_vcs.py:511-526); per-block partials in LDS, fixed-order reduction.
//   lb_k_levels   per pair-level cluster row Q: h_QR, PHR multipliers pp, PHR value, M_Q = sum_R pp_QR Y_R
//                 (level_aug alm_vcs.py:476-496; pair_all's M alm_vcs.py:378-415), FP64.
//   lb_k_vcs_z    per vertex: VCS effective multipliers pv over the working set, PHR value, and the transposed
//                 sums Z_R = sum_j pv_jR d_j V_j (_vcs_aug_rows vcsops.py:266-288; vcs_apply vcsops.py:326-330).
//   lb_k_segs     wavefront-per-segment pass over power-law hubs: rows whose neighbour list or triangle
//                 incidence is long are cut into segments; each wavefront sums one segment.
//   lb_k_rows     the hot kernel: per vertex, one slot of S lanes (16 bytes of the V row per lane) computes
//                 G_i = (A V)_i - d_i z / L + sum_l d_i (M_Q - pp_QQ d_i V_i) + VCS row terms + T(pt) V rows,
//                 quadA_i = V_i.(A V)_i, and (for role-0 triangle records) pt_t and the triangle PHR value.
//   lb_k_finalize fixed-order FP64 reduction of every per-block partial into the parts[] scalars.
// Global sums are FP64; per-row arithmetic is FP32 (LB_F32, LB_BF16) or FP64 (LB_F64).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <climits>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>
#include <array>
#include <mutex>

#include "../../include/lbhip.h"

// gathers kept in flight per lane: row kernel edges / triangle records, batch row kernel edges / records
// LDS row stride padding (doubles) of per-cluster rows accumulated or read by many slots at once (colsum, the VCS
// passes): two doubles shift consecutive cluster rows by 16 B, so slots on different clusters spread over the banks
#define LB_VPAD 2
#define LB_BLK_MAXS 6      // blocks per Gram side / per combination of the device eigensolver
#ifndef LB_SEB
#define LB_SEB 8           // sweep hub segments: edge gathers per slot in flight
#endif
#ifndef LB_SRB
#define LB_SRB 4           // sweep hub segments: triangle records per slot in flight
#endif
#ifndef LB_POST_ROWS
#define LB_POST_ROWS 32    // sweep aggregate increments (lb_k_post_acc): batch rows per workgroup
#endif
#ifndef LB_EB
#define LB_EB 16
#endif
#ifndef LB_RB
#define LB_RB 4
#endif
#ifndef LB_BEB
#define LB_BEB 8
#endif
// batch mode: rows with at most this many edges and triangle records run one slot per row (default of the
// "batch_light" option)
#ifndef LB_BLIGHT
#define LB_BLIGHT 128
#endif
#ifndef LB_BRB
#define LB_BRB 4
#endif

#define LB_MAXLEV 8
#define LB_MAXVCS 4
#ifndef LB_RED_BLOCKS
#define LB_RED_BLOCKS 1024   // cluster-sum workgroups: 1024 measured 0.5% faster than 512 and 256 on the 1M-row graph
#endif
#ifndef LB_VCS_BLOCKS
#define LB_VCS_BLOCKS 1536   // VCS pass workgroups: 1536 measured best of 512 / 1024 / 1536 / 2048 on the 1M-row graph
#endif
#define LB_BLOCK 256
#define LB_ZPART 32        // fused sweeps: copies of z the rows of a batch add into (a power of two)
// merged post step (lb_k_post_both) only for batches of at most this many rows: each of its workgroups sums the whole
// batch's increments per VCS cluster in LDS, one batch row per thread, so larger batches take a second pass
// (measured on a 40k-vertex request, 76 rows: 0.304 s of sweeps against 0.321 s in two post launches; on a
// 150k-vertex request, 282 rows: 0.400 s against 0.389 s)
#define LB_POST_MERGE_ROWS LB_BLOCK

typedef uint16_t bf16raw;

// ------------------------------------------------------------------------------------------------ storage traits
template <typename TV> struct VT;
template <> struct VT<float> {
    typedef float TC;
    static constexpr int CPL = 4;  // columns per lane: 16 bytes of the row
    __device__ static inline void unpack(const uint4 &u, float *x) {
        x[0] = __uint_as_float(u.x); x[1] = __uint_as_float(u.y); x[2] = __uint_as_float(u.z); x[3] = __uint_as_float(u.w);
    }
};
template <> struct VT<bf16raw> {
    typedef float TC;
    static constexpr int CPL = 8;
    __device__ static inline void unpack(const uint4 &u, float *x) {
        x[0] = __uint_as_float(u.x << 16); x[1] = __uint_as_float(u.x & 0xffff0000u);
        x[2] = __uint_as_float(u.y << 16); x[3] = __uint_as_float(u.y & 0xffff0000u);
        x[4] = __uint_as_float(u.z << 16); x[5] = __uint_as_float(u.z & 0xffff0000u);
        x[6] = __uint_as_float(u.w << 16); x[7] = __uint_as_float(u.w & 0xffff0000u);
    }
};
template <> struct VT<double> {
    typedef double TC;
    static constexpr int CPL = 2;
    __device__ static inline void unpack(const uint4 &u, double *x) {
        x[0] = __hiloint2double((int)u.y, (int)u.x); x[1] = __hiloint2double((int)u.w, (int)u.z);
    }
};

template <typename TC> struct alignas(16) TriRec { int32_t u, w, code; TC lam; };
template <> struct alignas(8) TriRec<double> { int32_t u, w, code, pad; double lam; };

struct Seg { int32_t row, kind; int64_t b, e; };

// cross-lane exchange inside a quad (lanes 4k..4k+3) by DPP quad_perm: xor 1 = [1,0,3,2], xor 2 = [2,3,0,1]
template <int M> __device__ inline float dpp_xor(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), M == 1 ? 0xB1 : 0x4E, 0xF, 0xF, false));
}
template <int M> __device__ inline double dpp_xor(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(u & 0xffffffffu), M == 1 ? 0xB1 : 0x4E, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(u >> 32), M == 1 ? 0xB1 : 0x4E, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// sum over the S consecutive lanes of a slot (S a power of two <= 64): DPP inside quads, shuffles beyond
template <int S, typename T> __device__ inline T slot_sum(T v) {
    if (S >= 2) v += dpp_xor<1>(v);
    if (S >= 4) v += dpp_xor<2>(v);
#pragma unroll
    for (int m = 4; m < S; m <<= 1) v += __shfl_xor(v, m);
    return v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m);
    return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmin(v, __shfl_xor(v, m));
    return v;
}
// Block sum in a fixed order (wave trees, then waves in order); valid in thread 0.
__device__ inline double block_sum(double v, double *sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;
}
__device__ inline double block_min(double v, double *sh) {
    v = wave_min(v);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = INFINITY;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s = fmin(s, sh[k]);
    return s;
}

// ------------------------------------------------------------------------------------------------ kernel args
struct LabelSets {             // every label set whose cluster sums are needed (pair levels, then VCS levels)
    int nsets;
    int off[LB_MAXLEV + LB_MAXVCS + 1];   // cluster offsets into Y (rows of R doubles)
    const int32_t *lab[LB_MAXLEV + LB_MAXVCS];
};

struct ColsumArgs {
    const uint4 *V; int64_t n; const double *d;
    LabelSets ls; int sumP; int use_lds;
    double *part;              // gridDim x (R + sumP R) when use_lds, else gridDim x R
    double *Yglob;             // sumP x R (atomic path)
};

struct LevelArgs {
    int nlev; int R;
    int Poff[LB_MAXLEV + 1];   // cluster offsets (global ids, rows of Y)
    int64_t soff[LB_MAXLEV + 1];
    const double *Y; const double *sig; double *pp; const double *rq; const double *dsq;
    double w, rho;
    double *M; double *pdiag; double *tot; double *viol;   // tot/viol per cluster row
};

struct VcsArgs {
    const uint4 *V; int64_t n; const double *d;
    int nvcs; int Yoff[LB_MAXVCS]; int Zoff[LB_MAXVCS + 1];
    const int32_t *lab[LB_MAXVCS]; const double *rq[LB_MAXVCS];
    const int64_t *ptr[LB_MAXVCS];   // absolute entry offsets
    const int32_t *col; const void *sig; void *pv; uint8_t *slack;
    const double *Y; double wv, rho; int sumPv; int use_lds;
    double *part;              // gridDim x (2 + sumPv R) when use_lds, else gridDim x 2
    double *Zglob;
    int zacc;                  // 1: lb_k_vcs_z also accumulates Z; 0: lb_k_vcs_acc does it in a second pass
    const int32_t *vrow;       // entry -> row (entry-parallel passes)
    int64_t eoff[LB_MAXVCS + 1];   // entry range of each level: [eoff[m], eoff[m + 1])
};

struct SegArgs {
    const uint4 *V; const int64_t *rowptr; const int32_t *col; const void *w;
    const void *rec; const Seg *segs; int64_t nseg;
    void *segout; void *pt; double rho, rt;
    double *tpart;             // per block: tri tot, tri gmin
};

struct RowArgs {
    const uint4 *V; int64_t n; const int64_t *rowptr; const int32_t *col; const void *w; const double *d; double L;
    const double *z;
    int nlev; int Poff[LB_MAXLEV]; const int32_t *lab[LB_MAXLEV]; const double *M; const double *pdiag;
    int nvcs; int Yoff[LB_MAXVCS]; int Zoff[LB_MAXVCS]; const int32_t *vlab[LB_MAXVCS]; const int64_t *vptr[LB_MAXVCS];
    const int32_t *vcol; const void *pv; const double *Y; const double *Z;
    const int64_t *iptr; const void *rec; void *pt; double rho, rt;
    int64_t heavyE, heavyT;
    const int32_t *hE_rows; const int32_t *hE_seg0; int nhE;
    const int32_t *hT_rows; const int32_t *hT_seg0; int nhT;
    const void *segout;
    void *G;
    double *part;              // per block: quad, gv, tri tot, tri gmin
    int stage, sumPpair, sumPv;  // stage [z | M | pdiag | Y_vcs | Z] in LDS (as the compute type)
    const int32_t *order;        // processing order of the rows (similar work next to each other), or null
};

struct FinArgs {
    int R; double L, C0; const double *z; const double *lamsum;
    const double *lev_tot; const double *lev_viol; int sumPpair;
    const double *vcs_part; int nvcs_blk; int vcs_width;
    const double *row_part; int nrow_blk;
    const double *seg_part; int nseg_blk;
    double rho; double *parts;
};

// ------------------------------------------------------------------------------------------------ lb_k_colsum
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_colsum(ColsumArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shY[];
    __shared__ double zred[LB_BLOCK / 64][R];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int width = a.use_lds ? a.sumP * R : 0;
    constexpr int RS = R + LB_VPAD;                  // padded LDS row stride
    for (int e = threadIdx.x; e < a.sumP * RS && a.use_lds; e += LB_BLOCK) shY[e] = 0.0;
    __syncthreads();
    double zacc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) zacc[c] = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        TC x[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
        const double di = a.d[i];
        double v[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) { v[c] = di * (double)x[c]; zacc[c] += v[c]; }
        for (int s = 0; s < a.ls.nsets; s++) {
            const int Q = a.ls.off[s] + a.ls.lab[s][i];
            double *dst = a.use_lds ? (shY + (int64_t)Q * RS + q * CPL) : (a.Yglob + (int64_t)Q * R + q * CPL);
#pragma unroll
            for (int c = 0; c < CPL; c++) atomicAdd(dst + c, v[c]);
        }
    }
    // z: reduce over the slots of each wave (lanes with equal q), then over waves
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        double t = zacc[c];
        for (int m = S; m < 64; m <<= 1) t += __shfl_xor(t, m);
        zacc[c] = t;
    }
    const int wv = threadIdx.x >> 6, ln = threadIdx.x & 63;
    if (ln < S) {
#pragma unroll
        for (int c = 0; c < CPL; c++) zred[wv][ln * CPL + c] = zacc[c];
    }
    __syncthreads();
    const int pw = R + width;
    double *out = a.part + (int64_t)blockIdx.x * pw;
    if (threadIdx.x < R) {
        double s = 0.0;
        for (int k = 0; k < LB_BLOCK / 64; k++) s += zred[k][threadIdx.x];
        out[threadIdx.x] = s;
    }
    for (int e = threadIdx.x; e < width; e += LB_BLOCK) out[R + e] = shY[(e / R) * RS + e % R];
}

// fixed-order sum of per-block partials: out[e] = sum_b part[b * stride + e]. 64 columns per workgroup (one per
// lane, coalesced), the 16 wavefronts split the blocks, then a fixed-order combine in LDS.
__global__ __launch_bounds__(1024) void lb_k_reduce(const double *part, int nblk, int width, int stride, double *out) {
    __shared__ double sh[16][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int e = blockIdx.x * 64 + lane;
    double s0 = 0.0, s1 = 0.0;
    if (e < width) {
        int b = w;
        for (; b + 16 < nblk; b += 32) { s0 += part[(int64_t)b * stride + e]; s1 += part[(int64_t)(b + 16) * stride + e]; }
        for (; b < nblk; b += 16) s0 += part[(int64_t)b * stride + e];
    }
    sh[w][lane] = s0 + s1;
    __syncthreads();
    if (w == 0 && e < width) {
        double t = 0.0;
        for (int k = 0; k < 16; k++) t += sh[k][lane];
        out[e] = t;
    }
}

__global__ void lb_k_fill(double *p, int64_t cnt, double v) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) p[e] = v;
}

// ------------------------------------------------------------------------------------------------ lb_k_levels
// One wavefront per cluster row Q of a pair level (global row id = blockIdx.x); the rank is a compile-time constant
// so the per-lane M accumulator stays in registers.
template <int R>
__global__ __launch_bounds__(64) void lb_k_levels(LevelArgs a) {
    const int Qg = blockIdx.x;
    int l = 0;
    while (l + 1 < a.nlev && Qg >= a.Poff[l + 1]) l++;
    const int P = a.Poff[l + 1] - a.Poff[l], Q = Qg - a.Poff[l];
    const double *YQ = a.Y + (int64_t)Qg * R;
    const double *sig = a.sig + a.soff[l] + (int64_t)Q * P;
    double *pp = a.pp + a.soff[l] + (int64_t)Q * P;
    const double rqQ = a.rq[Qg];
    double Macc[R];
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = 0.0;
    double tot = 0.0, viol = 0.0, pq = 0.0;
    for (int Rr = threadIdx.x; Rr < P; Rr += 64) {
        const double *YR = a.Y + (int64_t)(a.Poff[l] + Rr) * R;
        double p = sig[Rr];
        if (a.rho > 0.0) {
            double h = 0.0;
#pragma unroll
            for (int c = 0; c < R; c++) h += YQ[c] * YR[c];
            if (Rr == Q) h -= a.dsq[Qg];
            const double rr = a.rho * ((rqQ * a.rq[a.Poff[l] + Rr]) * a.w);
            const double sg = p;
            p = sg - rr * h;
            if (p < 0.0) p = 0.0;
            tot += (p * p - sg * sg) / (2 * rr);
            if (h < -1e-9) viol += 1.0;
        }
        pp[Rr] = p;
        if (Rr == Q) pq = p;
        if (p != 0.0) {
#pragma unroll
            for (int c = 0; c < R; c++) Macc[c] += p * YR[c];
        }
    }
    tot = wave_sum(tot); viol = wave_sum(viol); pq = wave_sum(pq);
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = wave_sum(Macc[c]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < R; c++) a.M[(int64_t)Qg * R + c] = Macc[c];
        a.pdiag[Qg] = pq; a.tot[Qg] = tot; a.viol[Qg] = viol;
    }
}

// ------------------------------------------------------------------------------------------------ lb_k_vcs_z
// LDS path (use_lds): the VCS cluster sums Y are staged in LDS (FP64) and Z is accumulated in one LDS copy per
// wavefront in the row arithmetic type (FP32, FP64 for LB_F64), combined in FP64 at the end of the workgroup.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_vcs_z(VcsArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S, NW = LB_BLOCK / 64;
    extern __shared__ double shY[];
    __shared__ double red[LB_BLOCK / 64];
    const int slot = threadIdx.x / S, q = threadIdx.x % S, wv = threadIdx.x >> 6;
    const int width = a.use_lds ? a.sumPv * R : 0;
    TC *shZ = (TC *)(shY + width);
    const double *Yb = a.Y + (int64_t)a.Yoff[0] * R;   // the VCS cluster sums (all VCS levels, consecutive)
    if (a.use_lds) {
        for (int e = threadIdx.x; e < width; e += LB_BLOCK) shY[e] = Yb[e];
        if (a.zacc)
            for (int e = threadIdx.x; e < NW * width; e += LB_BLOCK) shZ[e] = (TC)0;
        __syncthreads();
        Yb = shY;
    }
    TC *myZ = shZ + wv * width;
    const TC *sig = (const TC *)a.sig;
    TC *pv = (TC *)a.pv;
    double tot = 0.0, viol = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        TC x[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
        const double di = a.d[i];
        for (int m = 0; m < a.nvcs; m++) {
            const int Ri = a.lab[m][i];
            const int64_t e0 = a.ptr[m][i], e1 = a.ptr[m][i + 1];
            for (int64_t eb = e0; eb < e1; eb += 4) {
              // the entries' columns and multipliers are loaded 4 at a time before any is used
              int Rq[4];
              TC sq[4];
#pragma unroll
              for (int t = 0; t < 4; t++) {
                  const int64_t ec = eb + t < e1 ? eb + t : e1 - 1;
                  Rq[t] = a.col[ec];
                  sq[t] = sig[ec];
              }
              const int nb = (int)(e1 - eb < 4 ? e1 - eb : 4);
              for (int t = 0; t < nb; t++) {
                const int64_t e = eb + t;
                const int Rc = Rq[t];
                const double *YR = Yb + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < CPL; c++) s += (double)x[c] * YR[c];
                s = slot_sum<S>(s);
                const double g = (Rc == Ri) ? di * (s - di) : di * s;
                const double sg = (double)sq[t];
                double p = sg;
                if (a.rho > 0.0 && di > 0.0) {
                    const double rr = a.rho * (a.wv * a.rq[m][Rc] / (di * di));
                    p = sg - rr * g;
                    if (p < 0.0) p = 0.0;
                    if (q == 0) tot += (p * p - sg * sg) / (2 * rr);
                }
                if (q == 0) {
                    pv[e] = (TC)p;
                    if (g < -1e-9) viol += 1.0;
                    a.slack[e] = (p == 0.0 && g > 0.0) ? 1 : 0;   // inactive with slack (vcsops.py:285)
                }
                const TC pc = (TC)p;  // the stored multiplier is what G uses
                if (a.zacc && pc != (TC)0) {
                    if (a.use_lds) {
                        TC *dst = myZ + (a.Zoff[m] + Rc) * R + q * CPL;
#pragma unroll
                        for (int c = 0; c < CPL; c++) atomicAdd(dst + c, (TC)((double)pc * (di * (double)x[c])));
                    } else {
                        double *dst = a.Zglob + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
#pragma unroll
                        for (int c = 0; c < CPL; c++) atomicAdd(dst + c, (double)pc * (di * (double)x[c]));
                    }
                }
              }
            }
        }
    }
    tot = block_sum(tot, red);
    viol = block_sum(viol, red);
    __syncthreads();
    const int pw = 2 + width;
    double *out = a.part + (int64_t)blockIdx.x * pw;
    if (threadIdx.x == 0) { out[0] = tot; out[1] = viol; }
    if (a.zacc)
        for (int e = threadIdx.x; e < width; e += LB_BLOCK) {
            double z = 0.0;
            for (int k = 0; k < NW; k++) z += (double)shZ[k * width + e];
            out[2 + e] = z;
        }
}

// Entry-parallel form of the multipliers pass (used when the VCS cluster sums fit in LDS): one slot per working-set
// entry (i, R) -- its row comes from vrow -- and LB_VU entries per slot in flight: every entry's loads are issued
// before any is used (the pass is bound by dependent load latency, not by bandwidth or ALU). Same arithmetic as
// lb_k_vcs_z; Z is left to lb_k_vcs_acc.
#define LB_VU 4
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_vcs_pv(VcsArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shY[];
    __shared__ double red[LB_BLOCK / 64];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int width = a.sumPv * R;
    const double *Yg = a.Y + (int64_t)a.Yoff[0] * R;   // the VCS cluster sums (all VCS levels, consecutive)
    constexpr int RS = R + LB_VPAD;
    for (int e = threadIdx.x; e < width; e += LB_BLOCK) shY[(e / R) * RS + e % R] = Yg[e];
    __syncthreads();
    const TC *sig = (const TC *)a.sig;
    TC *pv = (TC *)a.pv;
    const int64_t ntot = a.eoff[a.nvcs], stride = (int64_t)gridDim.x * NSL;
    double tot = 0.0, viol = 0.0;
    for (int64_t e0 = (int64_t)blockIdx.x * NSL + slot; e0 < ntot; e0 += LB_VU * stride) {
        int64_t ee[LB_VU], ii[LB_VU];
        int Rq[LB_VU];
        TC sq[LB_VU];
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            const int64_t e = e0 + t * stride;
            ee[t] = e < ntot ? e : ntot - 1;           // past the end: a valid entry, loaded and not used
            ii[t] = a.vrow[ee[t]];
            Rq[t] = a.col[ee[t]];
            sq[t] = sig[ee[t]];
        }
        TC xq[LB_VU][CPL];
        double dq[LB_VU];
        int mq[LB_VU], Rl[LB_VU];
        double rqq[LB_VU];
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            int m = 0;
            while (m + 1 < a.nvcs && ee[t] >= a.eoff[m + 1]) m++;
            mq[t] = m;
            VT<TV>::unpack(a.V[ii[t] * S + q], xq[t]);
            dq[t] = a.d[ii[t]];
            Rl[t] = a.lab[m][ii[t]];
            rqq[t] = a.rq[m][Rq[t]];
        }
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            if (e0 + t * stride >= ntot) break;          // slot-uniform
            const int Rc = Rq[t];
            const double di = dq[t];
            const double *YR = shY + (int64_t)(a.Zoff[mq[t]] + Rc) * RS + q * CPL;
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < CPL; c++) s += (double)xq[t][c] * YR[c];
            s = slot_sum<S>(s);
            const double g = (Rc == Rl[t]) ? di * (s - di) : di * s;
            const double sg = (double)sq[t];
            double p = sg;
            if (a.rho > 0.0 && di > 0.0) {
                const double rr = a.rho * (a.wv * rqq[t] / (di * di));
                p = sg - rr * g;
                if (p < 0.0) p = 0.0;
                if (q == 0) tot += (p * p - sg * sg) / (2 * rr);
            }
            if (q == 0) {
                pv[ee[t]] = (TC)p;
                if (g < -1e-9) viol += 1.0;
                a.slack[ee[t]] = (p == 0.0 && g > 0.0) ? 1 : 0;   // inactive with slack (vcsops.py:285)
            }
        }
    }
    tot = block_sum(tot, red);
    viol = block_sum(viol, red);
    if (threadIdx.x == 0) {
        double *out = a.part + (int64_t)blockIdx.x * (2 + width);
        out[0] = tot; out[1] = viol;
    }
}

// Z_R = sum over the working-set entries (i, R) of pv_iR d_i V_i, after lb_k_vcs_pv wrote pv: one slot per entry,
// LB_VU entries in flight per slot, one FP64 copy of Z per workgroup in LDS, written to the same per-workgroup
// partial rows as the multipliers pass.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_vcs_acc(VcsArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shZ[];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int width = a.sumPv * R;
    constexpr int RS = R + LB_VPAD;
    for (int e = threadIdx.x; e < a.sumPv * RS; e += LB_BLOCK) shZ[e] = 0.0;
    __syncthreads();
    const TC *pv = (const TC *)a.pv;
    const int64_t ntot = a.eoff[a.nvcs], stride = (int64_t)gridDim.x * NSL;
    for (int64_t e0 = (int64_t)blockIdx.x * NSL + slot; e0 < ntot; e0 += LB_VU * stride) {
        int64_t ee[LB_VU], ii[LB_VU];
        int Rq[LB_VU];
        double pq[LB_VU];
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            const int64_t e = e0 + t * stride;
            ee[t] = e < ntot ? e : ntot - 1;
            ii[t] = a.vrow[ee[t]];
            Rq[t] = a.col[ee[t]];
            pq[t] = e < ntot ? (double)pv[ee[t]] : 0.0;
        }
        TC xq[LB_VU][CPL];
        double dq[LB_VU];
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            VT<TV>::unpack(a.V[ii[t] * S + q], xq[t]);
            dq[t] = a.d[ii[t]];
        }
#pragma unroll
        for (int t = 0; t < LB_VU; t++) {
            if (pq[t] == 0.0) continue;
            int m = 0;
            while (m + 1 < a.nvcs && ee[t] >= a.eoff[m + 1]) m++;
            double *dst = shZ + (int64_t)(a.Zoff[m] + Rq[t]) * RS + q * CPL;
#pragma unroll
            for (int c = 0; c < CPL; c++) atomicAdd(dst + c, pq[t] * (dq[t] * (double)xq[t][c]));
        }
    }
    __syncthreads();
    double *out = a.part + (int64_t)blockIdx.x * (2 + width);
    for (int e = threadIdx.x; e < width; e += LB_BLOCK) out[2 + e] = shZ[(e / R) * RS + e % R];
}

// ------------------------------------------------------------------------------------------------ lb_k_vcs_mf
// The VCS pass in one sweep over the rows, with the Z sums on the FP64 matrix cores. A wavefront takes tiles of 16
// consecutive rows:
//   1. the tile's V rows and d (as FP64) and the tile's working-set bounds go to the wavefront's part of LDS;
//   2. one lane per working-set entry (i, R) of the tile: g = d_i (V_i . Y_R - [R = R_i] d_i) against LDS-staged Y,
//      the PHR multiplier pv, its PHR value, the violation count and the slack flag (vcsops.py:266-288), and pv
//      written into a dense 16 x SP block of the tile (row of the tile x cluster; zero where a row has no entry;
//      every (i, R) is one entry, so the writes never collide);
//   3. Z^T (R x SP) += (diag(d) V)_tile^T (R x 16) P_tile (16 x SP) with v_mfma_f64_16x16x4f64 over the tile's four
//      4-row steps (vcs_apply's Z = sum over entries of pv_iR d_i V_i, vcsops.py:326-330).
// V, the entries and their outputs stream through once; there are no atomics. The wavefronts' Z^T tiles are summed
// in a fixed order into the workgroup's partial row (2 + sumPv R doubles, as lb_k_vcs_z writes), which lb_k_reduce
// sums over the workgroups in a fixed order. Used when sumPv <= 128 and ceil(R / 16) ceil(sumPv / 16) <= 8 (the
// accumulators stay in registers). MFMA lane layout as in lb_k_chol_update_mfma: lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16]; register v of the result is D[4 v + l / 16][l % 16].
typedef double lb_vd4 __attribute__((ext_vector_type(4)));
#define LB_VMF_WAVES 4
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}
// LDS bytes of lb_k_vcs_mf: Y and rq of the VCS clusters, then per wavefront the multiplier tile (compute type),
// the V tile and d (FP64), the bounds and the labels; at the end the FP64 Z^T tiles of all wavefronts reuse the
// space from the multiplier tiles on, which must therefore hold NW x 16 MB x SP doubles.
template <typename TC> inline size_t vcs_mf_lds(int sumPv, int R, int TR) {
    const size_t NB = (sumPv + 15) / 16, SP = NB * 16, MB = (R + 15) / 16, NW = LB_VMF_WAVES;
    const size_t head = sizeof(double) * ((size_t)sumPv * (R + LB_VPAD) + sumPv);
    const size_t tiles = sizeof(TC) * (NW * TR * SP + NW * TR * R) + sizeof(double) * NW * TR +
                         sizeof(int64_t) * NW * LB_MAXVCS * (TR + 1) + sizeof(int32_t) * NW * LB_MAXVCS * TR;
    const size_t zt = sizeof(double) * NW * 16 * MB * SP;
    return head + std::max(tiles, zt);
}
template <typename TV, int R, int NACC, int TR>
__device__ __forceinline__ void vcs_mf_body(const VcsArgs &a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, MB = (R + 15) / 16, RS = R + LB_VPAD, NW = LB_VMF_WAVES;
    constexpr int RPL = 64 / S;                                 // tile rows per wavefront-wide load of V
    constexpr int NLD = (TR + RPL - 1) / RPL;                   // loads per lane for the TR rows of a tile
    constexpr int TB = TR + 1;                                  // working-set bounds per level and tile
    constexpr int KB = 64 / TB, KL = 64 / TR;                   // levels whose bounds / labels one load covers
    extern __shared__ double smem[];
    __shared__ double red[NW];
    const int NB = (a.sumPv + 15) / 16, SP = NB * 16;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    double *shY = smem;                                         // sumPv x RS: the VCS cluster sums
    double *shq = shY + (int64_t)a.sumPv * RS;                  // sumPv: rq of every VCS cluster
    TC *dense = reinterpret_cast<TC *>(shq + a.sumPv);          // NW x TR x SP: the tiles' multipliers
    TC *tV = dense + NW * TR * SP;                              // NW x TR x R: the tile's V rows
    double *td = reinterpret_cast<double *>(tV + NW * TR * R);  // NW x TR: the tile's d (8-byte aligned: R even)
    int64_t *tp = reinterpret_cast<int64_t *>(td + NW * TR);    // NW x LB_MAXVCS x TB: working-set bounds
    int32_t *tl = reinterpret_cast<int32_t *>(tp + NW * LB_MAXVCS * TB);   // NW x LB_MAXVCS x TR: labels
    const double *Yg = a.Y + (int64_t)a.Yoff[0] * R;
    for (int e = threadIdx.x; e < a.sumPv * R; e += 64 * NW) shY[(e / R) * RS + e % R] = Yg[e];
    for (int m = 0; m < a.nvcs; m++)
        for (int e = threadIdx.x; e < a.Zoff[m + 1] - a.Zoff[m]; e += 64 * NW) shq[a.Zoff[m] + e] = a.rq[m][e];
    TC *myd = dense + w * TR * SP;
    TC *mV = tV + w * TR * R;
    double *mdd = td + w * TR;
    int64_t *mp = tp + w * LB_MAXVCS * TB;
    int32_t *ml = tl + w * LB_MAXVCS * TR;
    for (int e = lane; e < TR * SP; e += 64) myd[e] = (TC)0;
    __syncthreads();
    const TC *sig = (const TC *)a.sig;
    TC *pv = (TC *)a.pv;
    double tot = 0.0, viol = 0.0;
    lb_vd4 acc[NACC];
#pragma unroll
    for (int k = 0; k < NACC; k++) acc[k] = (lb_vd4){0.0, 0.0, 0.0, 0.0};
    const int64_t ntile = (a.n + TR - 1) / TR;
    // the tile data (V rows, d, working-set bounds, labels) is loaded one tile ahead into registers
    uint4 pV[NLD];
    double pd = 0.0;
    int64_t pp0 = 0, pp1 = 0;
    int32_t pl = 0;
    auto fetch = [&](int64_t t) {
        const int64_t r0 = t * TR;
#pragma unroll
        for (int j = 0; j < NLD; j++) {
            const int k = j * RPL + lane / S;
            const int64_t row = r0 + k < a.n ? r0 + k : a.n - 1;   // rows past n: a valid row, zeroed when stored
            pV[j] = a.V[row * S + lane % S];
        }
        pd = lane < TR && r0 + lane < a.n ? a.d[r0 + lane] : 0.0;
        // bounds: TB lanes per level for the first KB levels; labels: TR lanes per level for the first KL levels;
        // further levels are loaded when the tile is stored
        const int mb_ = lane / TB, kb_ = lane % TB;
        pp0 = mb_ < a.nvcs && mb_ < KB ? a.ptr[mb_][min(r0 + kb_, a.n)] : 0;
        pp1 = 0;
        const int ml_ = lane / TR, kl_ = lane % TR;
        pl = ml_ < a.nvcs && ml_ < KL && r0 + kl_ < a.n ? a.lab[ml_][r0 + kl_] : 0;
    };
    auto store = [&](int64_t t) {
        const int64_t r0 = t * TR;
#pragma unroll
        for (int j = 0; j < NLD; j++) {
            const int k = j * RPL + lane / S, q = lane % S;
            if (k < TR) {
                TC x[CPL];
                VT<TV>::unpack(pV[j], x);
#pragma unroll
                for (int c = 0; c < CPL; c++) mV[k * R + q * CPL + c] = r0 + k < a.n ? x[c] : (TC)0;
            }
        }
        if (lane < TR) mdd[lane] = pd;
        const int mb_ = lane / TB, kb_ = lane % TB;
        if (mb_ < a.nvcs && mb_ < KB) mp[mb_ * TB + kb_] = pp0;
        for (int e = KB * TB + lane; e < a.nvcs * TB; e += 64) mp[e] = a.ptr[e / TB][min(r0 + e % TB, a.n)];
        const int ml_ = lane / TR, kl_ = lane % TR;
        if (ml_ < a.nvcs && ml_ < KL) ml[ml_ * TR + kl_] = pl;
        for (int e = KL * TR + lane; e < a.nvcs * TR; e += 64)
            ml[e] = r0 + e % TR < a.n ? a.lab[e / TR][r0 + e % TR] : 0;
    };
    int64_t t = (int64_t)blockIdx.x * NW + w;
    const int64_t tstride = (int64_t)gridDim.x * NW;
    if (t < ntile) fetch(t);
    for (; t < ntile; t += tstride) {
        const int64_t r0 = t * TR;
        // 1. this tile's data to LDS; the next tile's loads go out now and land while this one is computed
        store(t);
        if (t + tstride < ntile) fetch(t + tstride);
        wave_lds_sync();
        // 2. the entries' multipliers, one lane each, into the dense tile
        for (int m = 0; m < a.nvcs; m++) {
            const int64_t E0 = mp[m * TB], E1 = mp[m * TB + TR];
            for (int64_t e = E0 + lane; e < E1; e += 64) {
                const int Rc = a.col[e];
                const double sg = (double)sig[e];
                int k = 0;                                  // the entry's row: mp[k] <= e < mp[k + 1]
#pragma unroll
                for (int h = TR / 2; h >= 1; h >>= 1) k += mp[m * TB + k + h] <= e ? h : 0;
                const int Ri = ml[m * TR + k];
                const double di = mdd[k];
                const double *YR = shY + (int64_t)(a.Zoff[m] + Rc) * RS;
                double s = 0.0;
#pragma unroll 4
                for (int c = 0; c < R; c++) s += (double)mV[k * R + c] * YR[c];
                const double g = (Rc == Ri) ? di * (s - di) : di * s;
                double p = sg;
                if (a.rho > 0.0 && di > 0.0) {
                    const double rr = a.rho * (a.wv * shq[a.Zoff[m] + Rc] / (di * di));
                    p = sg - rr * g;
                    if (p < 0.0) p = 0.0;
                    tot += (p * p - sg * sg) / (2 * rr);
                }
                const TC pc = (TC)p;  // the stored multiplier is what Z and G use
                pv[e] = pc;
                if (g < -1e-9) viol += 1.0;
                a.slack[e] = (p == 0.0 && g > 0.0) ? 1 : 0;   // inactive with slack (vcsops.py:285)
                myd[k * SP + a.Zoff[m] + Rc] = pc;
            }
        }
        wave_lds_sync();
        // 3. Z^T += (D V)_tile^T P_tile, TR / 4 steps of 4 rows
        const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
        for (int st = 0; st < TR / 4; st++) {
            const int k = 4 * st + lk;
            double av[MB];
#pragma unroll
            for (int mb = 0; mb < MB; mb++) {
                const int c = 16 * mb + li;
                av[mb] = c < R ? mdd[k] * (double)mV[k * R + c] : 0.0;
            }
            for (int b = 0; b < NB; b++) {
                const double bv = (double)myd[k * SP + 16 * b + li];
#pragma unroll
                for (int mb = 0; mb < MB; mb++) {
                    const int slot = b * MB + mb;
                    // slot < NB MB <= NACC: the accumulators are indexed with compile-time constants
#pragma unroll
                    for (int u = 0; u < NACC; u++)
                        if (u == slot) acc[u] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[mb], bv, acc[u], 0, 0, 0);
                }
            }
        }
        wave_lds_sync();
        for (int e = lane; e < TR * SP; e += 64) myd[e] = (TC)0;
        (void)r0;
    }
    // the wavefronts' Z^T tiles go to LDS (over the multiplier tiles and the V tiles, which are no longer needed)
    // and are summed over the waves in a fixed order
    __syncthreads();
    double *zt = reinterpret_cast<double *>(dense);             // NW x (16 MB) x SP doubles fit in that space
    double *myz = zt + (int64_t)w * 16 * MB * SP;
#pragma unroll
    for (int u = 0; u < NACC; u++) {
        if (u >= NB * MB) break;
        const int b = u / MB, mb = u % MB;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int c = 16 * mb + 4 * v + (lane >> 4);
            myz[c * SP + 16 * b + (lane & 15)] = acc[u][v];
        }
    }
    tot = block_sum(tot, red);
    __syncthreads();
    viol = block_sum(viol, red);
    __syncthreads();
    const int width = a.sumPv * R;
    double *out = a.part + (int64_t)blockIdx.x * (2 + width);
    if (threadIdx.x == 0) { out[0] = tot; out[1] = viol; }
    for (int e = threadIdx.x; e < width; e += 64 * NW) {
        const int cl = e / R, c = e % R;
        double z = 0.0;
        for (int k = 0; k < NW; k++) z += zt[(int64_t)k * 16 * MB * SP + c * SP + cl];
        out[2 + e] = z;
    }
}

// four accumulator tiles (R <= 16 and sumPv <= 64): 128 registers, four wavefronts per SIMD without spills
template <typename TV, int R, int TR>
__global__ __launch_bounds__(64 * LB_VMF_WAVES) __attribute__((amdgpu_waves_per_eu(4))) void lb_k_vcs_mf4(VcsArgs a) {
    vcs_mf_body<TV, R, 4, TR>(a);
}
template <typename TV, int R, int TR>
__global__ __launch_bounds__(64 * LB_VMF_WAVES) void lb_k_vcs_mf8(VcsArgs a) {
    vcs_mf_body<TV, R, 8, TR>(a);
}

// ------------------------------------------------------------------------------------------------ triangle record
// One incidence record of vertex i (x = its V chunk, xu / xw = the chunks of the two other vertices): tau of the
// triangle in the canonical (a, b, c) order (the same arithmetic for all three owners, as batch_step6 does,
// alm_vcs.py:297-312), pt, and the row contribution of T(pt) V (tri_apply, alm_vcs.py:106-120).
template <typename TV, int R>
__device__ inline void tri_record_x(int code, typename VT<TV>::TC lam, const typename VT<TV>::TC *x,
                                    const typename VT<TV>::TC *xu, const typename VT<TV>::TC *xw, double rho, double rt,
                                    typename VT<TV>::TC *g, typename VT<TV>::TC &pt_out, typename VT<TV>::TC &tau_out) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int role = code % 3;
    const TC *pa = role == 0 ? x : xu;
    const TC *pb = role == 0 ? xu : (role == 1 ? x : xw);
    const TC *pc = role == 2 ? x : xw;
    TC tl = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) tl += pa[c] * pb[c] + pb[c] * pc[c] - pa[c] * pc[c];
    const TC tau = slot_sum<S>(tl);
    TC p = lam;
    if (rho > 0.0) {
        p = lam - (TC)rt * ((TC)1 - tau);
        if (p < (TC)0) p = (TC)0;
    }
    if (p != (TC)0) {
        const TC h = (TC)0.5 * p;
        if (role == 0) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += h * (xw[c] - xu[c]);
        } else if (role == 1) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] -= h * (xu[c] + xw[c]);
        } else {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += h * (xu[c] - xw[c]);
        }
    }
    pt_out = p;
    tau_out = tau;
}

template <typename TV, int R>
__device__ inline void tri_record(const uint4 *V, const TriRec<typename VT<TV>::TC> &rc, const typename VT<TV>::TC *x,
                                  int q, double rho, double rt, typename VT<TV>::TC *g, typename VT<TV>::TC &pt_out,
                                  typename VT<TV>::TC &tau_out) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    TC xu[CPL], xw[CPL];
    VT<TV>::unpack(V[(int64_t)rc.u * S + q], xu);
    VT<TV>::unpack(V[(int64_t)rc.w * S + q], xw);
    tri_record_x<TV, R>(rc.code, rc.lam, x, xu, xw, rho, rt, g, pt_out, tau_out);
}

// ------------------------------------------------------------------------------------------------ lb_k_segs
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_segs(SegArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSW = 64 / S;
    __shared__ double red[LB_BLOCK / 64];
    const int lane = threadIdx.x & 63, ws = lane / S, q = lane % S;
    double tot = 0.0, gmin = INFINITY;
    // one wavefront per segment; a grid smaller than nseg / 4 blocks strides over the segments (a fixed assignment,
    // so the per-block partials do not depend on timing)
    for (int64_t sg = (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6); sg < a.nseg;
         sg += (int64_t)gridDim.x * (LB_BLOCK / 64)) {
        const Seg s = a.segs[sg];
        TC acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) acc[c] = 0;
        if (s.kind == 0) {
            // a slot's items are e = b + ws + k NSW; LB_EB of them gathered together, accumulated in item order
            // (items past the segment read the hub row with weight 0: an exact zero)
            const TC *wt = (const TC *)a.w;
            for (int64_t e0 = s.b + ws; e0 < s.e; e0 += (int64_t)LB_EB * NSW) {
                int64_t jj[LB_EB];
                TC ww[LB_EB];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) {
                    const int64_t e = e0 + (int64_t)t * NSW;
                    const bool in = e < s.e;
                    jj[t] = in ? (int64_t)a.col[e] : (int64_t)s.row;
                    ww[t] = in ? (wt ? wt[e] : (TC)1) : (TC)0;
                }
                uint4 u[LB_EB];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) u[t] = a.V[jj[t] * S + q];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) {
                    TC xj[CPL];
                    VT<TV>::unpack(u[t], xj);
#pragma unroll
                    for (int c = 0; c < CPL; c++) acc[c] += ww[t] * xj[c];
                }
            }
        } else {
            TC x[CPL];
            VT<TV>::unpack(a.V[(int64_t)s.row * S + q], x);
            const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
            TC *pt = (TC *)a.pt;
            // LB_RB records of the slot (strided by NSW) with their two rows gathered together, in record order
            for (int64_t k0 = s.b + ws; k0 < s.e; k0 += (int64_t)LB_RB * NSW) {
                TriRec<TC> rq[LB_RB];
#pragma unroll
                for (int t = 0; t < LB_RB; t++) {
                    const int64_t k = k0 + (int64_t)t * NSW;
                    rq[t] = rec[k < s.e ? k : k0];
                }
                uint4 uu[LB_RB], uw[LB_RB];
#pragma unroll
                for (int t = 0; t < LB_RB; t++) { uu[t] = a.V[(int64_t)rq[t].u * S + q]; uw[t] = a.V[(int64_t)rq[t].w * S + q]; }
#pragma unroll
                for (int t = 0; t < LB_RB; t++) {
                    if (k0 + (int64_t)t * NSW >= s.e) break;     // slot-uniform
                    const TriRec<TC> &rc = rq[t];
                    TC xu[CPL], xw[CPL], p, tau;
                    VT<TV>::unpack(uu[t], xu);
                    VT<TV>::unpack(uw[t], xw);
                    tri_record_x<TV, R>(rc.code, rc.lam, x, xu, xw, a.rho, a.rt, acc, p, tau);
                    if (rc.code % 3 == 0 && q == 0) {
                        pt[rc.code / 3] = p;
                        if (a.rho > 0.0) {
                            tot += ((double)p * p - (double)rc.lam * rc.lam) / (2 * a.rt);
                            gmin = fmin(gmin, (double)((TC)1 - tau));
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            TC t = acc[c];
            for (int m = S; m < 64; m <<= 1) t += __shfl_xor(t, m);
            acc[c] = t;
        }
        if (ws == 0) {
            TC *out = (TC *)a.segout + sg * R + q * CPL;
#pragma unroll
            for (int c = 0; c < CPL; c++) out[c] = acc[c];
        }
    }
    tot = block_sum(tot, red);
    __syncthreads();
    gmin = block_min(gmin, red);
    if (threadIdx.x == 0) { a.tpart[2 * blockIdx.x] = tot; a.tpart[2 * blockIdx.x + 1] = gmin; }
}

__device__ inline int find_row(const int32_t *rows, int cnt, int64_t i) {
    int lo = 0, hi = cnt;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (rows[mid] < i) lo = mid + 1; else hi = mid;
    }
    return lo;
}

// ------------------------------------------------------------------------------------------------ lb_k_rows
// Grid-stride over vertices, one slot of S lanes per vertex (16 bytes of the V row per lane). The small per-cluster
// arrays are staged in LDS. Neighbour gathers are issued 8 at a time and triangle records 4 at a time (all row loads
// of a group in flight before any is used), so each slot keeps several independent gathers outstanding.
template <typename TV, int R>
__device__ __forceinline__ void rows_body(const RowArgs &a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double sm[];
    __shared__ double red[LB_BLOCK / 64];
    const double *Yg = a.Y + (int64_t)a.sumPpair * R;
    // staged copies in the compute type (FP32 for FP32 / BF16 V): half the LDS of FP64 staging, so more resident
    // workgroups per CU; the global arrays stay FP64 and are read directly when the staging does not fit
    TC *s0 = reinterpret_cast<TC *>(sm), *s1 = s0 + R, *s2 = s1 + a.sumPpair * R, *s3 = s2 + a.sumPpair,
       *s4 = s3 + a.sumPv * R;
    if (a.stage) {
        for (int e = threadIdx.x; e < R; e += LB_BLOCK) s0[e] = (TC)a.z[e];
        for (int e = threadIdx.x; e < a.sumPpair * R; e += LB_BLOCK) s1[e] = (TC)a.M[e];
        for (int e = threadIdx.x; e < a.sumPpair; e += LB_BLOCK) s2[e] = (TC)a.pdiag[e];
        for (int e = threadIdx.x; e < a.sumPv * R; e += LB_BLOCK) { s3[e] = (TC)Yg[e]; s4[e] = (TC)a.Z[e]; }
        __syncthreads();
    }
    const bool stg = a.stage != 0;
    auto ldz = [&](int64_t k) -> double { return stg ? (double)s0[k] : a.z[k]; };
    auto ldM = [&](int64_t k) -> double { return stg ? (double)s1[k] : a.M[k]; };
    auto ldpd = [&](int64_t k) -> double { return stg ? (double)s2[k] : a.pdiag[k]; };
    auto ldY = [&](int64_t k) -> double { return stg ? (double)s3[k] : Yg[k]; };
    auto ldZ = [&](int64_t k) -> double { return stg ? (double)s4[k] : a.Z[k]; };
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const TC *wt = (const TC *)a.w;
    const TC *pv = (const TC *)a.pv;
    const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
    TC *pt = (TC *)a.pt;
    double quad = 0.0, gv = 0.0, tot = 0.0, gmin = INFINITY;
    for (int64_t ii = (int64_t)blockIdx.x * NSL + slot; ii < a.n; ii += (int64_t)gridDim.x * NSL) {
        const int64_t i = a.order ? (int64_t)a.order[ii] : ii;
        TC x[CPL], g[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
#pragma unroll
        for (int c = 0; c < CPL; c++) g[c] = 0;
        // (A V)_i
        const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
        if (e1 - e0 > a.heavyE) {
            const int k = find_row(a.hE_rows, a.nhE, i);
            const TC *so = (const TC *)a.segout;
            for (int s = a.hE_seg0[k]; s < a.hE_seg0[k + 1]; s++) {
#pragma unroll
                for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
            }
        } else {
            for (int64_t e = e0; e < e1; e += LB_EB) {
                const int64_t nE = e1 - e;
                int32_t jj[LB_EB];
                TC ww[LB_EB];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) {   // clamped index: always in bounds; lanes past the row reuse row i
                    const int64_t ec = t < nE ? e + t : e1 - 1;
                    const int32_t cj = a.col[ec];
                    jj[t] = t < nE ? cj : (int32_t)i;
                    ww[t] = wt ? wt[ec] : (TC)1;
                    if (t >= nE) ww[t] = (TC)0;
                }
                uint4 u[LB_EB];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) u[t] = a.V[(int64_t)jj[t] * S + q];
#pragma unroll
                for (int t = 0; t < LB_EB; t++) {
                    TC xj[CPL];
                    VT<TV>::unpack(u[t], xj);
#pragma unroll
                    for (int c = 0; c < CPL; c++) g[c] += ww[t] * xj[c];
                }
            }
        }
        TC qd = 0;
#pragma unroll
        for (int c = 0; c < CPL; c++) qd += x[c] * g[c];
        qd = slot_sum<S>(qd);
        if (q == 0) quad += (double)qd;
        const double di = a.d[i];
        // rank-1 correction -d_i z / L
#pragma unroll
        for (int c = 0; c < CPL; c++) g[c] -= (TC)(di * ldz(q * CPL + c) / a.L);
        // pair levels: d_i (M_Q - pp_QQ d_i V_i)
        for (int l = 0; l < a.nlev; l++) {
            const int Q = a.Poff[l] + a.lab[l][i];
            const int64_t mo = (int64_t)Q * R + q * CPL;
            const double pd = ldpd(Q) * di;
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += (TC)(di * (ldM(mo + c) - pd * (double)x[c]));
        }
        // VCS: 1/2 d_i (sum_R pv_iR Y_R + Z_R(i)) - pv_own d_i^2 V_i
        for (int m = 0; m < a.nvcs; m++) {
            const int Ri = a.vlab[m][i];
            double acc[CPL];
#pragma unroll
            for (int c = 0; c < CPL; c++) acc[c] = 0.0;
            double pown = 0.0;
            const int64_t v0 = a.vptr[m][i], v1 = a.vptr[m][i + 1];
            for (int64_t eb = v0; eb < v1; eb += 4) {
                int Rq[4];
                double pq[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {           // 4 entries in flight; entries past the row add 0
                    const int64_t ec = eb + t < v1 ? eb + t : v1 - 1;
                    Rq[t] = a.vcol[ec];
                    pq[t] = eb + t < v1 ? (double)pv[ec] : 0.0;
                }
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    if (Rq[t] == Ri && eb + t < v1) pown = pq[t];
                    const int64_t yo = (int64_t)(a.Zoff[m] + Rq[t]) * R + q * CPL;
#pragma unroll
                    for (int c = 0; c < CPL; c++) acc[c] += pq[t] * ldY(yo + c);
                }
            }
            const int64_t zo = (int64_t)(a.Zoff[m] + Ri) * R + q * CPL;
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += (TC)(0.5 * di * (acc[c] + ldZ(zo + c)) - pown * di * di * (double)x[c]);
        }
        // triangles: T(pt) V rows through the incidence records
        const int64_t r0 = a.iptr[i], r1 = a.iptr[i + 1];
        if (r1 - r0 > a.heavyT) {
            const int k = find_row(a.hT_rows, a.nhT, i);
            const TC *so = (const TC *)a.segout;
            for (int s = a.hT_seg0[k]; s < a.hT_seg0[k + 1]; s++) {
#pragma unroll
                for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
            }
        } else {
            for (int64_t k = r0; k < r1; k += LB_RB) {
                const int64_t nK = r1 - k;
                TriRec<TC> rc[LB_RB];
#pragma unroll
                for (int t = 0; t < LB_RB; t++) rc[t] = rec[t < nK ? k + t : r1 - 1];
                uint4 uu[LB_RB], uw[LB_RB];
#pragma unroll
                for (int t = 0; t < LB_RB; t++) { uu[t] = a.V[(int64_t)rc[t].u * S + q]; uw[t] = a.V[(int64_t)rc[t].w * S + q]; }
#pragma unroll
                for (int t = 0; t < LB_RB; t++) {
                    // records past the row's end are processed with lam = 0 and rho = 0: they add nothing
                    const bool valid = t < nK;
                    TC xu[CPL], xw[CPL], p, tau;
                    VT<TV>::unpack(uu[t], xu);
                    VT<TV>::unpack(uw[t], xw);
                    tri_record_x<TV, R>(rc[t].code, valid ? rc[t].lam : (TC)0, x, xu, xw, valid ? a.rho : 0.0, a.rt, g,
                                        p, tau);
                    if (valid && rc[t].code % 3 == 0 && q == 0) {
                        pt[rc[t].code / 3] = p;
                        if (a.rho > 0.0) {
                            tot += ((double)p * p - (double)rc[t].lam * rc[t].lam) / (2 * a.rt);
                            gmin = fmin(gmin, (double)((TC)1 - tau));
                        }
                    }
                }
            }
        }
        // G row and <G_i, V_i>
        TC *G = (TC *)a.G + i * R + q * CPL;
        TC gd = 0;
#pragma unroll
        for (int c = 0; c < CPL; c++) { G[c] = g[c]; gd += g[c] * x[c]; }
        gd = slot_sum<S>(gd);
        if (q == 0) gv += (double)gd;
    }
    quad = block_sum(quad, red);
    __syncthreads();
    gv = block_sum(gv, red);
    __syncthreads();
    tot = block_sum(tot, red);
    __syncthreads();
    gmin = block_min(gmin, red);
    if (threadIdx.x == 0) {
        double *o = a.part + 4 * (int64_t)blockIdx.x;
        o[0] = quad; o[1] = gv; o[2] = tot; o[3] = gmin;
    }
}

template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_rows(RowArgs a) {
    rows_body<TV, R>(a);
}
// BF16 V: eight columns per lane take 169 registers, one step over three wavefronts per SIMD; capped at 168 (a
// 12-byte spill) the row kernel runs 271 against 331 us on the 1M-row graph. FP64 V loses from the same cap.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) __attribute__((amdgpu_waves_per_eu(3))) void lb_k_rows_w3(RowArgs a) {
    rows_body<TV, R>(a);
}

// ------------------------------------------------------------------------------------------------ lb_k_finalize
__global__ __launch_bounds__(1024) void lb_k_finalize(FinArgs a) {
    double quad = 0, gv = 0, ttot = 0, gmin = INFINITY, vtot = 0, vviol = 0, ltot = 0, lviol = 0;
    // per-thread strided partial sums, combined in a fixed order
    for (int b = threadIdx.x; b < a.nrow_blk; b += 1024) {
        const double *o = a.row_part + 4 * (int64_t)b;
        quad += o[0]; gv += o[1]; ttot += o[2]; gmin = fmin(gmin, o[3]);
    }
    for (int b = threadIdx.x; b < a.nseg_blk; b += 1024) {
        ttot += a.seg_part[2 * b]; gmin = fmin(gmin, a.seg_part[2 * b + 1]);
    }
    for (int b = threadIdx.x; b < a.nvcs_blk; b += 1024) {
        vtot += a.vcs_part[(int64_t)b * a.vcs_width]; vviol += a.vcs_part[(int64_t)b * a.vcs_width + 1];
    }
    for (int b = threadIdx.x; b < a.sumPpair; b += 1024) { ltot += a.lev_tot[b]; lviol += a.lev_viol[b]; }
    // the eight block reductions in one pass: each value keeps block_sum's (block_min's) wave butterfly and its
    // fixed order over the waves
    __shared__ double red8[16][8];
    {
        double v8[8] = {quad, gv, ttot, gmin, vtot, vviol, ltot, lviol};
#pragma unroll
        for (int j = 0; j < 8; j++) v8[j] = j == 3 ? wave_min(v8[j]) : wave_sum(v8[j]);
        const int w = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
#pragma unroll
            for (int j = 0; j < 8; j++) red8[w][j] = v8[j];
        __syncthreads();
        if (threadIdx.x == 0) {
            double t8[8] = {0, 0, 0, INFINITY, 0, 0, 0, 0};
            for (int k = 0; k < (int)(blockDim.x >> 6); k++)
#pragma unroll
                for (int j = 0; j < 8; j++) t8[j] = j == 3 ? fmin(t8[j], red8[k][j]) : t8[j] + red8[k][j];
            quad = t8[0]; gv = t8[1]; ttot = t8[2]; gmin = t8[3]; vtot = t8[4]; vviol = t8[5]; ltot = t8[6]; lviol = t8[7];
        }
    }
    if (threadIdx.x == 0) {
        double zz = 0.0;
        for (int c = 0; c < a.R; c++) zz += a.z[c] * a.z[c];
        double *p = a.parts;
        const double lamsum = *a.lamsum;
        p[2] = zz / a.L; p[7] = a.C0; p[11] = lamsum;
        p[6] = quad - zz / a.L + a.C0;
        if (a.rho > 0.0) {
            p[1] = quad; p[3] = ltot; p[4] = ttot; p[5] = vtot; p[9] = gmin;
            p[0] = quad - zz / a.L - ltot - ttot - vtot + a.C0;
        } else {
            p[1] = gv; p[3] = 0.0; p[4] = 0.0; p[5] = 0.0; p[9] = INFINITY;
            p[0] = a.C0 + lamsum + gv;
        }
        p[8] = lviol; p[10] = vviol;
    }
}

// ------------------------------------------------------------------------------------------------ small kernels
// per-block partial sums (fixed order) of x, reduced by lb_k_reduce
template <typename TC>
__global__ void lb_k_sum(const TC *x, int64_t cnt, double *part) {
    __shared__ double red[LB_BLOCK / 64];
    double s = 0.0;
    for (int64_t e = (int64_t)blockIdx.x * LB_BLOCK + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * LB_BLOCK) s += (double)x[e];
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// age_update (vcsops.py:300-305): consecutive outers an entry stayed inactive with slack
__global__ void lb_k_age(uint8_t *age, const uint8_t *slack, int64_t cnt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
        if (slack[e]) { if (age[e] < 255) age[e] = age[e] + 1; }
        else age[e] = 0;
    }
}

template <typename TC>
__global__ void lb_k_inc_lam(TriRec<TC> *rec, int64_t nrec, const TC *lam) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nrec; k += (int64_t)gridDim.x * blockDim.x)
        rec[k].lam = lam[rec[k].code / 3];
}

template <typename TC>
__global__ void lb_k_copy(TC *dst, const TC *src, int64_t cnt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) dst[e] = src[e];
}

// ------------------------------------------------------------------------------------------------ sweeps (device)
// Batched Gauss-Seidel row steps (sweeps, alm_vcs.py:1360-1436 with pair_refresh = 1). Per sweep: z, Y (pair and
// VCS levels), the curvature Kq (kq_level, alm_vcs.py:417-428), the VCS state stZ / Z / cnt (vcs_state,
// vcsops.py:308-323) and the batch -> touched-cluster mask. Per batch: fresh pair rows of the touched clusters
// (pair_all, alm_vcs.py:378-415), the row steps of the batch from a snapshot of V and the aggregates
// (batch_step6, alm_vcs.py:255-351, with vcs_pre, vcsops.py:127-156), then the aggregate updates z, Y
// (_acc_rows, alm_vcs.py:430-437) and the VCS post-step (vcs_yupd / vcs_post, vcsops.py:159-213) with the new rows
// written back.

__device__ inline uint16_t lb_f2bf(float f) {   // round to nearest even
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

template <typename TV> struct Store;
template <> struct Store<float> {
    __device__ static inline uint4 pack(const float *x) { return make_uint4(__float_as_uint(x[0]), __float_as_uint(x[1]), __float_as_uint(x[2]), __float_as_uint(x[3])); }
};
template <> struct Store<bf16raw> {
    __device__ static inline uint4 pack(const float *x) {
        uint4 u;
        u.x = (uint32_t)lb_f2bf(x[0]) | ((uint32_t)lb_f2bf(x[1]) << 16);
        u.y = (uint32_t)lb_f2bf(x[2]) | ((uint32_t)lb_f2bf(x[3]) << 16);
        u.z = (uint32_t)lb_f2bf(x[4]) | ((uint32_t)lb_f2bf(x[5]) << 16);
        u.w = (uint32_t)lb_f2bf(x[6]) | ((uint32_t)lb_f2bf(x[7]) << 16);
        return u;
    }
};
template <> struct Store<double> {
    __device__ static inline uint4 pack(const double *x) {
        const unsigned long long a = __double_as_longlong(x[0]), b = __double_as_longlong(x[1]);
        return make_uint4((uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32));
    }
};

struct KqArgs {
    int nlev, R; int Poff[LB_MAXLEV + 1]; int64_t soff[LB_MAXLEV + 1];
    const double *Y; const double *sig; const double *rq; const double *dsq; double w, rho; double *Kq;
};

// Kq_Q = 2 rho (sum_R [pp_QR > 0] r_QR |Y_R|^2 + [pp_QQ > 0] r_QQ |Y_Q|^2), pp = sig - rho r H (one wave per Q)
__global__ __launch_bounds__(64) void lb_k_kq(KqArgs a) {
    const int Qg = blockIdx.x;
    int l = 0;
    while (l + 1 < a.nlev && Qg >= a.Poff[l + 1]) l++;
    const int P = a.Poff[l + 1] - a.Poff[l], Q = Qg - a.Poff[l], R = a.R;
    const double *YQ = a.Y + (int64_t)Qg * R;
    const double *sig = a.sig + a.soff[l] + (int64_t)Q * P;
    double s = 0.0, dg = 0.0;
    for (int Rr = threadIdx.x; Rr < P; Rr += 64) {
        const double *YR = a.Y + (int64_t)(a.Poff[l] + Rr) * R;
        double H = 0.0, y2 = 0.0;
        for (int c = 0; c < R; c++) { H += YQ[c] * YR[c]; y2 += YR[c] * YR[c]; }
        if (Rr == Q) H -= a.dsq[Qg];
        const double r = (a.rq[Qg] * a.rq[a.Poff[l] + Rr]) * a.w;
        const double p = sig[Rr] - a.rho * r * H;
        if (p > 0.0) s += r * y2;
        if (Rr == Q && p > 0.0) dg = r * y2;
    }
    s = wave_sum(s);
    dg = wave_sum(dg);
    if (threadIdx.x == 0) a.Kq[Qg] = 2 * a.rho * (s + dg);
}

struct PairRowArgs {
    int nlev, R; int Poff[LB_MAXLEV + 1]; int64_t soff[LB_MAXLEV + 1];
    const double *Y; const double *sig; const double *rq; const double *dsq; double w, rho;
    const uint8_t *tmask;    // touched clusters of this batch (sumP bytes)
    double *M; double *pdiag;
    // dd set (lb_k_post_next): Y holds the pair cluster rows before the previous batch, whose increments dd (dnb
    // rows drows, R each, labels lab) each workgroup adds per cluster in LDS (sumP R doubles) first
    const double *dd; const int32_t *drows; int64_t dnb; const int32_t *lab[LB_MAXLEV];
    const double *dPg;          // set (sweep_delta): the previous batch's pair cluster-row increments, read with Y
};

// pair_all for the clusters the batch touches: pp row from the current Y, M_Q = sum_R pp_QR Y_R, pdiag_Q = pp_QQ
template <int R>
__device__ inline void pair_row_one(const PairRowArgs &a, int Qg, int lane, const double *dP = nullptr);
template <int R>
__global__ __launch_bounds__(64) void lb_k_pair_rows(PairRowArgs a) {
    pair_row_one<R>(a, (int)blockIdx.x, (int)threadIdx.x);
}
// one wavefront per cluster (lane = the wavefront's lane); dP (optional): increments added to every Y row read
template <int R>
__device__ inline void pair_row_one(const PairRowArgs &a, int Qg, int lane, const double *dP) {
    if (!a.tmask[Qg]) return;
    int l = 0;
    while (l + 1 < a.nlev && Qg >= a.Poff[l + 1]) l++;
    const int P = a.Poff[l + 1] - a.Poff[l], Q = Qg - a.Poff[l];
    const double *YQ = a.Y + (int64_t)Qg * R;
    double yq[R];
#pragma unroll
    for (int c = 0; c < R; c++) yq[c] = dP ? YQ[c] + dP[(int64_t)Qg * R + c] : YQ[c];
    const double *sig = a.sig + a.soff[l] + (int64_t)Q * P;
    double Macc[R];
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = 0.0;
    double pq = 0.0;
    for (int Rr = lane; Rr < P; Rr += 64) {
        const int64_t Rg = a.Poff[l] + Rr;
        const double *YR = a.Y + Rg * R;
        double yr[R];
#pragma unroll
        for (int c = 0; c < R; c++) yr[c] = dP ? YR[c] + dP[Rg * R + c] : YR[c];
        double p = sig[Rr];
        if (a.rho > 0.0) {
            double h = 0.0;
#pragma unroll
            for (int c = 0; c < R; c++) h += yq[c] * yr[c];
            if (Rr == Q) h -= a.dsq[Qg];
            p = p - (a.rho * ((a.rq[Qg] * a.rq[a.Poff[l] + Rr]) * a.w)) * h;
            if (p < 0.0) p = 0.0;
        }
        if (Rr == Q) pq = p;
        if (p != 0.0) {
#pragma unroll
            for (int c = 0; c < R; c++) Macc[c] += p * yr[c];
        }
    }
    pq = wave_sum(pq);
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = wave_sum(Macc[c]);
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < R; c++) a.M[(int64_t)Qg * R + c] = Macc[c];
        a.pdiag[Qg] = pq;
    }
}

// Workgroup blk of the pair rows (four clusters, one wavefront each); with a.dd the previous batch's increments are
// first summed per pair cluster in LDS (one batch row per thread) and added to every Y row read. Every thread of the
// workgroup calls it (it synchronises the workgroup).
template <int R>
__device__ inline void pair_rows_blk(const PairRowArgs &a, int blk, int nclu) {
    extern __shared__ double shP[];
    const double *dP = nullptr;
    if (a.dd) {
        const int wp = a.Poff[a.nlev] * R;
        for (int e = threadIdx.x; e < wp; e += LB_BLOCK) shP[e] = 0.0;
        __syncthreads();
        for (int64_t p = threadIdx.x; p < a.dnb; p += LB_BLOCK) {
            const int32_t i = a.drows[p];
            double xv[R];
#pragma unroll
            for (int c = 0; c < R; c++) xv[c] = a.dd[p * R + c];
            for (int l = 0; l < a.nlev; l++) {
                double *dst = shP + (int64_t)(a.Poff[l] + a.lab[l][i]) * R;
#pragma unroll
                for (int c = 0; c < R; c++) atomicAdd(dst + c, xv[c]);
            }
        }
        __syncthreads();
        dP = shP;
    }
    if (a.dPg) dP = a.dPg;
    const int Qg = blk * (LB_BLOCK / 64) + (int)(threadIdx.x >> 6);
    if (Qg < nclu) pair_row_one<R>(a, Qg, (int)(threadIdx.x & 63), dP);
}

// batch -> touched clusters: tmask[b * sumP + Q] = 1 (b = batch of position p of the sweep's order)
struct TmaskArgs {
    const int32_t *order; int64_t n; int nbatch; int nlev; int Poff[LB_MAXLEV]; const int32_t *lab[LB_MAXLEV];
    int sumP; uint8_t *tmask;
};
__global__ void lb_k_tmask(TmaskArgs a) {
    const int64_t q = a.n / a.nbatch, rem = a.n % a.nbatch;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < a.n; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t b = p < (q + 1) * rem ? p / (q + 1) : rem + (p - (q + 1) * rem) / q;
        const int32_t i = a.order[p];
        for (int l = 0; l < a.nlev; l++) a.tmask[b * a.sumP + a.Poff[l] + a.lab[l][i]] = 1;
    }
}

struct VcsStateArgs {
    const uint4 *V; int64_t n; const double *d;
    int nvcs; int Yoff[LB_MAXVCS]; int Zoff[LB_MAXVCS];
    const int32_t *lab[LB_MAXVCS]; const double *rq[LB_MAXVCS]; const int64_t *ptr[LB_MAXVCS];
    const int32_t *col; const void *sig; const double *Y; double wv, rho;
    double *stZ; double *Z; double *cnt;     // without LDS: Z, cnt zeroed before the launch; FP64 atomics
    int sumPv; int use_lds;     // with LDS: per-workgroup sums of [Z | cnt] written to part (reduced afterwards)
    double *part;
};

// vcs_state: stZ = effective multipliers at (V, Y); Z_R = sum stZ d_j v_j; cnt_R = #(stZ > 0). With LDS, each
// workgroup sums its rows' contributions in LDS and writes one partial row [Z | cnt] (reduced in a fixed order
// afterwards); the entries of a row are loaded four at a time before any is used.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_vcs_state(VcsStateArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shS[];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const TC *sig = (const TC *)a.sig;
    const int wz = a.sumPv * R, wt = wz + a.sumPv;
    double *Zacc = a.use_lds ? shS : a.Z;
    double *Cacc = a.use_lds ? shS + wz : a.cnt;
    if (a.use_lds) {
        for (int e = threadIdx.x; e < wt; e += LB_BLOCK) shS[e] = 0.0;
        __syncthreads();
    }
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        TC x[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
        const double di = a.d[i];
        for (int m = 0; m < a.nvcs; m++) {
            const int Ri = a.lab[m][i];
            const int64_t ea = a.ptr[m][i], eb = a.ptr[m][i + 1];
            for (int64_t e4 = ea; e4 < eb; e4 += 4) {
                int Rq[4];
                double sq[4];
#pragma unroll
                for (int t = 0; t < 4; t++) {
                    const int64_t ec = e4 + t < eb ? e4 + t : eb - 1;
                    Rq[t] = a.col[ec];
                    sq[t] = (double)sig[ec];
                }
                const int nq = (int)(eb - e4 < 4 ? eb - e4 : 4);
                for (int t = 0; t < nq; t++) {
                    const int64_t e = e4 + t;
                    const int Rc = Rq[t];
                    const double *YR = a.Y + (int64_t)(a.Yoff[m] + Rc) * R + q * CPL;
                    double s = 0.0;
#pragma unroll
                    for (int c = 0; c < CPL; c++) s += (double)x[c] * YR[c];
                    s = slot_sum<S>(s);
                    double p = sq[t];
                    if (a.rho > 0.0 && di > 0.0) {
                        const double g = (Rc == Ri) ? di * (s - di) : di * s;
                        p = p - a.rho * (a.wv * a.rq[m][Rc] / (di * di)) * g;
                        if (p < 0.0) p = 0.0;
                    }
                    if (q == 0) {
                        a.stZ[e] = p;
                        if (p > 0.0) atomicAdd(Cacc + a.Zoff[m] + Rc, 1.0);
                    }
                    if (p != 0.0) {
                        double *dst = Zacc + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
#pragma unroll
                        for (int c = 0; c < CPL; c++) atomicAdd(dst + c, p * (di * (double)x[c]));
                    }
                }
            }
        }
    }
    if (a.use_lds) {
        __syncthreads();
        double *out = a.part + (int64_t)blockIdx.x * wt;
        for (int e = threadIdx.x; e < wt; e += LB_BLOCK) out[e] = shS[e];
    }
}

struct StepArgs {
    const uint4 *V; const int64_t *rowptr; const int32_t *col; const void *w; const double *d; double L;
    const double *z;
    int nlev; int Poff[LB_MAXLEV]; const int32_t *lab[LB_MAXLEV]; const double *M; const double *pdiag; const double *Kq;
    int nvcs; int Yoff[LB_MAXVCS]; int Zoff[LB_MAXVCS]; const int32_t *vlab[LB_MAXVCS]; const int64_t *vptr[LB_MAXVCS];
    const double *vrq[LB_MAXVCS];
    const int32_t *vcol; const void *vsig; const double *Y; const double *Z; const double *cnt; const double *stZ;
    double wv, kap2;
    const int64_t *iptr; const void *rec; double rho, rt, damp;
    int64_t heavyE, heavyT;
    const int32_t *hE_rows; const int32_t *hE_seg0; int nhE;
    const int32_t *hT_rows; const int32_t *hT_seg0; int nhT;
    const int32_t *hE_idx; const int32_t *hT_idx;   // row -> hub index (or -1)
    const void *segout; const double *segkap;
    const int32_t *rows; int64_t nb;            // the rows of this batch
    uint4 *Vstage; double *DD;                  // batch-position indexed
    // fused sweeps (lb_k_step_fused): zYn = [z | Y] after this batch, a copy of [z | Y] before it into which every
    // row adds its increment (FP64 atomics; DD is then not written); per hub row h (edge hubs by their hE index, the
    // other triangle hubs at nhE + their hT index): arrival count, batch position, own terms g[R] and kap
    double *zYn; int32_t *hubcnt; int64_t *hubpos; double *hubown;
    double *zpart;     // z increments spread over LB_ZPART copies of R doubles (by batch position), folded into z later
    // posb set (lb_k_step_copy): each row stepped records posb[i] = tag, posp[i] = its batch position
    int32_t *posb; int32_t *posp; int tag;
    // dnext set: the rows add their increments to [LB_ZPART copies of z | Y] deltas of this batch (FP64 atomics) in
    // place of writing DD
    double *dnext;
};

// one triangle record of the row step: tau (canonical order), q, pt = lam - rt (1 - tau); if pt > 0 the record adds
// -pt/2 q to g and rt |q|^2 to the curvature (batch_step6, alm_vcs.py:297-316)
template <typename TV, int R>
__device__ inline void step_record_x(const TriRec<typename VT<TV>::TC> &rc, const typename VT<TV>::TC *x,
                                     const typename VT<TV>::TC *xu, const typename VT<TV>::TC *xw, double rt,
                                     typename VT<TV>::TC *g, double &kap) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int role = rc.code % 3;
    const TC *pa = role == 0 ? x : xu;
    const TC *pb = role == 0 ? xu : (role == 1 ? x : xw);
    const TC *pc = role == 2 ? x : xw;
    TC tl = 0, ql = 0;
    TC qv[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        tl += pa[c] * pb[c] + pb[c] * pc[c] - pa[c] * pc[c];
        qv[c] = role == 0 ? xu[c] - xw[c] : (role == 1 ? xu[c] + xw[c] : xw[c] - xu[c]);
        ql += qv[c] * qv[c];
    }
    const TC tau = slot_sum<S>(tl);
    const TC q2 = slot_sum<S>(ql);
    const TC pt = rc.lam - (TC)rt * ((TC)1 - tau);
    if (pt > (TC)0) {
#pragma unroll
        for (int c = 0; c < CPL; c++) g[c] -= (TC)0.5 * pt * qv[c];
        kap += rt * (double)q2;
    }
}
template <typename TV, int R>
__device__ inline void step_record(const uint4 *V, const TriRec<typename VT<TV>::TC> &rc,
                                   const typename VT<TV>::TC *x, int q, double rt, typename VT<TV>::TC *g, double &kap) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    TC xu[CPL], xw[CPL];
    VT<TV>::unpack(V[(int64_t)rc.u * S + q], xu);
    VT<TV>::unpack(V[(int64_t)rc.w * S + q], xw);
    step_record_x<TV, R>(rc, x, xu, xw, rt, g, kap);
}

struct StepSegArgs {
    const uint4 *V; const int32_t *col; const void *w; const void *rec; const Seg *segs;
    const int32_t *list; const int64_t *lpos; int b;   // segment ids of batch b's hub rows: list[lpos[b] .. lpos[b + 1])
    double rt; void *segout; double *segkap;
    // posb set (lb_k_post_next): a row j with posb[j] == tag was stepped in the previous batch and its new value is
    // read from Vstage[posp[j]] (V of such rows is being rewritten by the same launch)
    const int32_t *posb; const int32_t *posp; const uint4 *Vstage; int tag;
};

// wavefront per hub segment: partial sum of A V (edge segments) or of the triangle records' step terms
// The batch's segment range is read from device memory (the lists change every sweep, the launch does not), and a
// fixed grid strides over it, so a sweep's launches can be captured once and replayed.
template <typename TV, int R>
__device__ inline void step_seg_one(const StepSegArgs &a, int32_t sid);
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_step_segs(StepSegArgs a) {
    const int64_t l0 = a.lpos[a.b], l1 = a.lpos[a.b + 1];
    for (int64_t li = l0 + (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6); li < l1;
         li += (int64_t)gridDim.x * (LB_BLOCK / 64))
        step_seg_one<TV, R>(a, a.list[li]);
}
// A batch's pair-level rows and its hub segments in one launch: the two read disjoint inputs and write disjoint
// outputs (M, pdiag; segout, segkap). Workgroups [0, npair) take four clusters each (one wavefront per cluster, as
// lb_k_pair_rows), the next nsegb stride over the batch's hub segments (as lb_k_step_segs), any others copy cw
// doubles from csrc to cdst (a snapshot no other part touches).
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_pair_segs(PairRowArgs pa, int npair, int nclu, StepSegArgs a, int nsegb,
                                                           const double *csrc, double *cdst, int64_t cw) {
    const int blk = (int)blockIdx.x;
    if (blk < npair) {
        const int Qg = blk * (LB_BLOCK / 64) + (int)(threadIdx.x >> 6);
        if (Qg < nclu) pair_row_one<R>(pa, Qg, (int)(threadIdx.x & 63));
        return;
    }
    if (blk < npair + nsegb) {
        const int64_t l0 = a.lpos[a.b], l1 = a.lpos[a.b + 1];
        const int64_t nsw = (int64_t)nsegb * (LB_BLOCK / 64);
        for (int64_t li = l0 + (int64_t)(blk - npair) * (LB_BLOCK / 64) + (threadIdx.x >> 6); li < l1; li += nsw)
            step_seg_one<TV, R>(a, a.list[li]);
        return;
    }
    // the rest: copy cw doubles (the VCS cluster rows at the batch start, for lb_k_post_both)
    const int64_t nthr = (int64_t)(gridDim.x - npair - nsegb) * LB_BLOCK;
    for (int64_t e = (int64_t)(blk - npair - nsegb) * LB_BLOCK + threadIdx.x; e < cw; e += nthr) cdst[e] = csrc[e];
}
template <typename TV, int R>
__device__ inline void step_seg_one(const StepSegArgs &a, int32_t sid) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSW = 64 / S;
    const int lane = threadIdx.x & 63, ws = lane / S, q = lane % S;
    const Seg s = a.segs[sid];
    TC acc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) acc[c] = 0;
    double kap = 0.0;
    if (s.kind == 0) {
        // LB_SEB edges per slot in flight (items past the segment read the hub row with weight 0: an exact zero)
        const TC *wt = (const TC *)a.w;
        for (int64_t e0 = s.b + ws; e0 < s.e; e0 += (int64_t)LB_SEB * NSW) {
            int64_t jj[LB_SEB];
            TC ww[LB_SEB];
#pragma unroll
            for (int t = 0; t < LB_SEB; t++) {
                const int64_t e = e0 + (int64_t)t * NSW;
                const bool in = e < s.e;
                jj[t] = in ? (int64_t)a.col[e] : (int64_t)s.row;
                ww[t] = in ? (wt ? wt[e] : (TC)1) : (TC)0;
            }
            uint4 u[LB_SEB];
#pragma unroll
            for (int t = 0; t < LB_SEB; t++) u[t] = a.V[jj[t] * S + q];
            if (a.posb) {
                int32_t tb[LB_SEB];
#pragma unroll
                for (int t = 0; t < LB_SEB; t++) tb[t] = a.posb[jj[t]];
#pragma unroll
                for (int t = 0; t < LB_SEB; t++)
                    if (tb[t] == a.tag) u[t] = a.Vstage[(int64_t)a.posp[jj[t]] * S + q];
            }
#pragma unroll
            for (int t = 0; t < LB_SEB; t++) {
                TC xj[CPL];
                VT<TV>::unpack(u[t], xj);
#pragma unroll
                for (int c = 0; c < CPL; c++) acc[c] += ww[t] * xj[c];
            }
        }
    } else {
        TC x[CPL];
        VT<TV>::unpack(a.V[(int64_t)s.row * S + q], x);
        const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
        // LB_SRB records per slot in flight (their two rows gathered together), applied in record order
        for (int64_t k0 = s.b + ws; k0 < s.e; k0 += (int64_t)LB_SRB * NSW) {
            TriRec<TC> rq[LB_SRB];
#pragma unroll
            for (int t = 0; t < LB_SRB; t++) {
                const int64_t k = k0 + (int64_t)t * NSW;
                rq[t] = rec[k < s.e ? k : k0];
            }
            uint4 uu[LB_SRB], uw[LB_SRB];
#pragma unroll
            for (int t = 0; t < LB_SRB; t++) { uu[t] = a.V[(int64_t)rq[t].u * S + q]; uw[t] = a.V[(int64_t)rq[t].w * S + q]; }
            if (a.posb) {
                int32_t tu[LB_SRB], tw[LB_SRB];
#pragma unroll
                for (int t = 0; t < LB_SRB; t++) { tu[t] = a.posb[rq[t].u]; tw[t] = a.posb[rq[t].w]; }
#pragma unroll
                for (int t = 0; t < LB_SRB; t++) {
                    if (tu[t] == a.tag) uu[t] = a.Vstage[(int64_t)a.posp[rq[t].u] * S + q];
                    if (tw[t] == a.tag) uw[t] = a.Vstage[(int64_t)a.posp[rq[t].w] * S + q];
                }
            }
#pragma unroll
            for (int t = 0; t < LB_SRB; t++) {
                if (k0 + (int64_t)t * NSW >= s.e) break;      // slot-uniform
                TC xu[CPL], xw[CPL];
                VT<TV>::unpack(uu[t], xu);
                VT<TV>::unpack(uw[t], xw);
                step_record_x<TV, R>(rq[t], x, xu, xw, a.rt, acc, kap);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        TC t = acc[c];
        for (int m = S; m < 64; m <<= 1) t += __shfl_xor(t, m);
        acc[c] = t;
    }
    for (int m = S; m < 64; m <<= 1) kap += __shfl_xor(kap, m);
    if (ws == 0) {
        TC *out = (TC *)a.segout + (int64_t)sid * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++) out[c] = acc[c];
        if (q == 0) a.segkap[sid] = kap;
    }
}

// rank-one, pair levels, VCS pre-step terms, then the normalised minorant step; writes Vstage[pos], DD[pos]
template <typename TV, int R>
__device__ inline void step_finish(const StepArgs &a, const typename VT<TV>::TC *x, int64_t i, int64_t pos, int q,
                                   typename VT<TV>::TC *g, double kap, bool writer = true) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const double di = a.d[i];
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] += (TC)(-(di / a.L) * (a.z[q * CPL + c] - di * (double)x[c]));
    for (int l = 0; l < a.nlev; l++) {
        const int Q = a.Poff[l] + a.lab[l][i];
        const double *M = a.M + (int64_t)Q * R + q * CPL;
        const double pd = a.pdiag[Q] * di;
#pragma unroll
        for (int c = 0; c < CPL; c++) g[c] += (TC)(di * (M[c] - pd * (double)x[c]));
        kap += di * di * a.Kq[Q];
    }
    const TC *vsig = (const TC *)a.vsig;
    for (int m = 0; m < a.nvcs; m++) {
        const int Ri = a.vlab[m][i];
        double acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) acc[c] = 0.0;
        double kp = 0.0, pown = 0.0, sz = 0.0;
        const int64_t ea = a.vptr[m][i], eb = a.vptr[m][i + 1];
        for (int64_t e4 = ea; e4 < eb; e4 += 4) {
          // four entries at a time: their columns and multipliers, then their cluster rows, before any is used
          int Rq[4];
          double sq[4], zq[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
              const int64_t ec = e4 + t < eb ? e4 + t : eb - 1;
              Rq[t] = a.vcol[ec];
              sq[t] = (double)vsig[ec];
              zq[t] = a.stZ[ec];
          }
          double yq[4][CPL];
#pragma unroll
          for (int t = 0; t < 4; t++) {
              const double *YR = a.Y + (int64_t)(a.Yoff[m] + Rq[t]) * R + q * CPL;
#pragma unroll
              for (int c = 0; c < CPL; c++) yq[t][c] = YR[c];
          }
          const int nq = (int)(eb - e4 < 4 ? eb - e4 : 4);
#pragma unroll
          for (int t = 0; t < 4; t++) {
            if (t >= nq) break;                              // slot-uniform
            const int Rc = Rq[t];
            const double *YR = yq[t];
            double s = 0.0, y2 = 0.0;
#pragma unroll
            for (int c = 0; c < CPL; c++) { s += (double)x[c] * YR[c]; y2 += YR[c] * YR[c]; }
            s = slot_sum<S>(s);
            double p;
            if (a.rho > 0.0 && di > 0.0) {
                const double g2 = (Rc == Ri) ? di * (s - di) : di * s;
                const double rr = a.wv * a.vrq[m][Rc] / (di * di);
                p = sq[t] - a.rho * rr * g2;
                if (p > 0.0) {
                    y2 = slot_sum<S>(y2);
                    if (Rc == Ri) y2 = y2 - 2 * di * s + di * di;
                    kp += a.rho * rr * di * di * y2;
                } else p = 0.0;
            } else p = sq[t];
            if (Rc == Ri) { pown = p; sz = zq[t]; }
            if (p != 0.0) {
#pragma unroll
                for (int c = 0; c < CPL; c++) acc[c] += p * YR[c];
            }
          }
        }
        const double *ZR = a.Z + (int64_t)(a.Zoff[m] + Ri) * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++)
            g[c] += (TC)(0.5 * di * (acc[c] - pown * di * (double)x[c]) + 0.5 * di * (ZR[c] - sz * di * (double)x[c]));
        if (a.rho > 0.0) kp += a.kap2 * a.rho * di * di * a.wv * a.vrq[m][Ri] * (a.cnt[a.Zoff[m] + Ri] - (sz > 0.0 ? 1.0 : 0.0));
        kap += kp;
    }
    TC gl = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) gl += g[c] * g[c];
    const TC gn = slot_sum<S>(gl);
    const TC w = (TC)(0.5 * kap) + (TC)a.damp * sqrt(gn);
    TC qv[CPL], nl = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) { qv[c] = g[c] + w * x[c]; nl += qv[c] * qv[c]; }
    const TC nv = slot_sum<S>(nl);
    TC xn[CPL];
    if (nv > (TC)0) {
        const TC inv = (TC)1 / sqrt(nv);
#pragma unroll
        for (int c = 0; c < CPL; c++) xn[c] = qv[c] * inv;
    } else {
#pragma unroll
        for (int c = 0; c < CPL; c++) xn[c] = x[c];
    }
    // round to the storage type first: DD tracks the stored rows exactly
    const uint4 u = Store<TV>::pack(xn);
    if (!writer) return;
    a.Vstage[pos * S + q] = u;
    TC xs[CPL];
    VT<TV>::unpack(u, xs);
    double *dd = a.DD + pos * R + q * CPL;
#pragma unroll
    for (int c = 0; c < CPL; c++) dd[c] = di * ((double)xs[c] - (double)x[c]);
}

template <typename TV, int R>
__device__ inline void step_finish_wave(const StepArgs &a, const typename VT<TV>::TC *x, int64_t i, int64_t pos, int q,
                                        int ws, typename VT<TV>::TC *g, double kap) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const double di = a.d[i];
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] += (TC)(-(di / a.L) * (a.z[q * CPL + c] - di * (double)x[c]));
    for (int l = 0; l < a.nlev; l++) {
        const int Q = a.Poff[l] + a.lab[l][i];
        const double *M = a.M + (int64_t)Q * R + q * CPL;
        const double pd = a.pdiag[Q] * di;
#pragma unroll
        for (int c = 0; c < CPL; c++) g[c] += (TC)(di * (M[c] - pd * (double)x[c]));
        kap += di * di * a.Kq[Q];
    }
    const TC *vsig = (const TC *)a.vsig;
    for (int m = 0; m < a.nvcs; m++) {
        const int Ri = a.vlab[m][i];
        double acc[CPL];
#pragma unroll
        for (int c = 0; c < CPL; c++) acc[c] = 0.0;
        double kp = 0.0, pown = 0.0, sz = 0.0;
        const int64_t ea = a.vptr[m][i], eb = a.vptr[m][i + 1];
        for (int64_t e = ea + ws; e < eb; e += 64 / S) {       // the row's entries dealt to the slots
            const int Rc = a.vcol[e];
            const double sqe = (double)vsig[e], zqe = a.stZ[e];
            const double *YR = a.Y + (int64_t)(a.Yoff[m] + Rc) * R + q * CPL;
            double yv[CPL];
#pragma unroll
            for (int c = 0; c < CPL; c++) yv[c] = YR[c];
            double s = 0.0, y2 = 0.0;
#pragma unroll
            for (int c = 0; c < CPL; c++) { s += (double)x[c] * yv[c]; y2 += yv[c] * yv[c]; }
            s = slot_sum<S>(s);
            double p;
            if (a.rho > 0.0 && di > 0.0) {
                const double g2 = (Rc == Ri) ? di * (s - di) : di * s;
                const double rr = a.wv * a.vrq[m][Rc] / (di * di);
                p = sqe - a.rho * rr * g2;
                if (p > 0.0) {
                    y2 = slot_sum<S>(y2);
                    if (Rc == Ri) y2 = y2 - 2 * di * s + di * di;
                    kp += a.rho * rr * di * di * y2;
                } else p = 0.0;
            } else p = sqe;
            if (Rc == Ri) { pown = p; sz = zqe; }
            if (p != 0.0) {
#pragma unroll
                for (int c = 0; c < CPL; c++) acc[c] += p * yv[c];
            }
        }
        // sum over the slots (the own-cluster entry sits in one slot; the others hold zeros there)
#pragma unroll
        for (int c = 0; c < CPL; c++)
            for (int mm = S; mm < 64; mm <<= 1) acc[c] += __shfl_xor(acc[c], mm);
        for (int mm = S; mm < 64; mm <<= 1) {
            kp += __shfl_xor(kp, mm);
            pown += __shfl_xor(pown, mm);
            sz += __shfl_xor(sz, mm);
        }
        const double *ZR = a.Z + (int64_t)(a.Zoff[m] + Ri) * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++)
            g[c] += (TC)(0.5 * di * (acc[c] - pown * di * (double)x[c]) + 0.5 * di * (ZR[c] - sz * di * (double)x[c]));
        if (a.rho > 0.0) kp += a.kap2 * a.rho * di * di * a.wv * a.vrq[m][Ri] * (a.cnt[a.Zoff[m] + Ri] - (sz > 0.0 ? 1.0 : 0.0));
        kap += kp;
    }
    TC gl = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) gl += g[c] * g[c];
    const TC gn = slot_sum<S>(gl);
    const TC w = (TC)(0.5 * kap) + (TC)a.damp * sqrt(gn);
    TC qv[CPL], nl = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) { qv[c] = g[c] + w * x[c]; nl += qv[c] * qv[c]; }
    const TC nv = slot_sum<S>(nl);
    TC xn[CPL];
    if (nv > (TC)0) {
        const TC inv = (TC)1 / sqrt(nv);
#pragma unroll
        for (int c = 0; c < CPL; c++) xn[c] = qv[c] * inv;
    } else {
#pragma unroll
        for (int c = 0; c < CPL; c++) xn[c] = x[c];
    }
    // round to the storage type first: DD tracks the stored rows exactly
    const uint4 u = Store<TV>::pack(xn);
    if (ws != 0) return;
    a.Vstage[pos * S + q] = u;
    TC xs[CPL];
    VT<TV>::unpack(u, xs);
    if (a.zYn) {
        // the increment straight into the next [z | Y]: the row's cluster of every pair and VCS level, and z through
        // one of LB_ZPART copies (every row of a batch adds to z: one address would serialise them)
        double *zy = a.zYn;
        double *zp = a.zpart + (int64_t)(pos & (LB_ZPART - 1)) * R + q * CPL;
        double *yl[LB_MAXLEV + LB_MAXVCS];
        int ns = 0;
        for (int l = 0; l < a.nlev; l++) yl[ns++] = zy + R + (int64_t)(a.Poff[l] + a.lab[l][i]) * R + q * CPL;
        for (int m = 0; m < a.nvcs; m++) yl[ns++] = zy + R + (int64_t)(a.Yoff[m] + a.vlab[m][i]) * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            const double dv = di * ((double)xs[c] - (double)x[c]);
            atomicAdd(zp + c, dv);
            for (int t = 0; t < ns; t++) atomicAdd(yl[t] + c, dv);
        }
        return;
    }
    if (a.posb && q == 0) { a.posb[i] = a.tag; a.posp[i] = (int32_t)pos; }
    if (a.dnext) {
        double *zp = a.dnext + (int64_t)(pos & (LB_ZPART - 1)) * R + q * CPL;
        double *yb = a.dnext + LB_ZPART * R;
        double *yl[LB_MAXLEV + LB_MAXVCS];
        int ns = 0;
        for (int l = 0; l < a.nlev; l++) yl[ns++] = yb + (int64_t)(a.Poff[l] + a.lab[l][i]) * R + q * CPL;
        for (int m = 0; m < a.nvcs; m++) yl[ns++] = yb + (int64_t)(a.Yoff[m] + a.vlab[m][i]) * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++) {
            const double dv = di * ((double)xs[c] - (double)x[c]);
            atomicAdd(zp + c, dv);
            for (int t = 0; t < ns; t++) atomicAdd(yl[t] + c, dv);
        }
        return;
    }
    double *dd = a.DD + pos * R + q * CPL;
#pragma unroll
    for (int c = 0; c < CPL; c++) dd[c] = di * ((double)xs[c] - (double)x[c]);
}

// one slot per batch row: SpMM row (or its hub segments), triangle records (or their segments), rank-one term,
// pair levels, VCS pre-step terms, then the normalised minorant step; writes Vstage[pos] and DD[pos]
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_step(StepArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int64_t pos = (int64_t)blockIdx.x * NSL + slot;
    if (pos >= a.nb) return;
    const int64_t i = a.rows[pos];
    TC x[CPL], g[CPL];
    VT<TV>::unpack(a.V[i * S + q], x);
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] = 0;
    double kap = 0.0;
    const TC *so = (const TC *)a.segout;
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    if (e1 - e0 > a.heavyE) {
        const int k = find_row(a.hE_rows, a.nhE, i);
        for (int s = a.hE_seg0[k]; s < a.hE_seg0[k + 1]; s++) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
        }
    } else {
        // LB_EB gathers in flight; lanes past the row read row i with weight 0 (adds an exact zero, same order)
        const TC *wt = (const TC *)a.w;
        for (int64_t e = e0; e < e1; e += LB_EB) {
            const int64_t nE = e1 - e;
            int32_t jj[LB_EB];
            TC ww[LB_EB];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) {
                const int64_t ec = t < nE ? e + t : e1 - 1;
                const int32_t cj = a.col[ec];
                jj[t] = t < nE ? cj : (int32_t)i;
                ww[t] = t < nE ? (wt ? wt[ec] : (TC)1) : (TC)0;
            }
            uint4 u[LB_EB];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) u[t] = a.V[(int64_t)jj[t] * S + q];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) {
                TC xj[CPL];
                VT<TV>::unpack(u[t], xj);
#pragma unroll
                for (int c = 0; c < CPL; c++) g[c] += ww[t] * xj[c];
            }
        }
    }
    const int64_t r0 = a.iptr[i], r1 = a.iptr[i + 1];
    if (r1 - r0 > a.heavyT) {
        const int k = find_row(a.hT_rows, a.nhT, i);
        for (int s = a.hT_seg0[k]; s < a.hT_seg0[k + 1]; s++) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
            kap += a.segkap[s];
        }
    } else {
        // LB_RB records in flight (their two rows gathered together), applied in record order
        const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
        for (int64_t k = r0; k < r1; k += LB_RB) {
            const int64_t nK = r1 - k;
            TriRec<TC> rc[LB_RB];
#pragma unroll
            for (int t = 0; t < LB_RB; t++) rc[t] = rec[t < nK ? k + t : r1 - 1];
            uint4 uu[LB_RB], uw[LB_RB];
#pragma unroll
            for (int t = 0; t < LB_RB; t++) { uu[t] = a.V[(int64_t)rc[t].u * S + q]; uw[t] = a.V[(int64_t)rc[t].w * S + q]; }
#pragma unroll
            for (int t = 0; t < LB_RB; t++) {
                if (t >= nK) break;                          // slot-uniform
                TC xu[CPL], xw[CPL];
                VT<TV>::unpack(uu[t], xu);
                VT<TV>::unpack(uw[t], xw);
                step_record_x<TV, R>(rc[t], x, xu, xw, a.rt, g, kap);
            }
        }
    }
    step_finish<TV, R>(a, x, i, pos, q, g, kap);
}

// One wavefront per batch row (a batch holds a few hundred rows, far fewer than the GPU's wavefront slots): the
// row's edges and triangle records are dealt to its 64 / S slots, so all their gathers are in flight at once, and the
// slot partials are summed by a butterfly over the slots. Every slot then holds the same g and curvature and runs
// the rest of the step (the slot-0 lanes write). Hub rows take their segment sums as in lb_k_step.
// The row's own terms of the wavefront-per-row step: its edges unless it is an edge hub, its triangle records unless
// it is a triangle hub, summed over the slots (every slot ends with the totals in g and kap).
template <typename TV, int R>
__device__ inline void step_row_own(const StepArgs &a, int64_t i, int q, int ws, const typename VT<TV>::TC *x,
                                    typename VT<TV>::TC *g, double &kap, bool &hubE, bool &hubT) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSW = 64 / S;
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] = 0;
    kap = 0.0;
    const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
    hubE = e1 - e0 > a.heavyE;
    if (!hubE) {
        // two edges per slot in flight; items past the row read row i with weight 0 (an exact zero)
        const TC *wt = (const TC *)a.w;
        for (int64_t eb = e0 + ws; eb < e1; eb += 2 * NSW) {
            int32_t jj[2];
            TC ww[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int64_t e = eb + (int64_t)t * NSW;
                const bool in = e < e1;
                jj[t] = in ? a.col[e] : (int32_t)i;
                ww[t] = in ? (wt ? wt[e] : (TC)1) : (TC)0;
            }
            uint4 u[2];
#pragma unroll
            for (int t = 0; t < 2; t++) u[t] = a.V[(int64_t)jj[t] * S + q];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                TC xj[CPL];
                VT<TV>::unpack(u[t], xj);
#pragma unroll
                for (int c = 0; c < CPL; c++) g[c] += ww[t] * xj[c];
            }
        }
    }
    const int64_t r0 = a.iptr[i], r1 = a.iptr[i + 1];
    hubT = r1 - r0 > a.heavyT;
    if (!hubT) {
        // two records per slot in flight (their four rows gathered together)
        const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
        for (int64_t kb = r0 + ws; kb < r1; kb += 2 * NSW) {
            TriRec<TC> rc[2];
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const int64_t k = kb + (int64_t)t * NSW;
                rc[t] = rec[k < r1 ? k : kb];
            }
            uint4 uu[2], uw[2];
#pragma unroll
            for (int t = 0; t < 2; t++) { uu[t] = a.V[(int64_t)rc[t].u * S + q]; uw[t] = a.V[(int64_t)rc[t].w * S + q]; }
#pragma unroll
            for (int t = 0; t < 2; t++) {
                if (kb + (int64_t)t * NSW >= r1) break;          // slot-uniform
                TC xu[CPL], xw[CPL];
                VT<TV>::unpack(uu[t], xu);
                VT<TV>::unpack(uw[t], xw);
                step_record_x<TV, R>(rc[t], x, xu, xw, a.rt, g, kap);
            }
        }
    }
    // sum over the slots (lanes with equal q): every slot ends with the row's totals
#pragma unroll
    for (int c = 0; c < CPL; c++) {
        TC t = g[c];
        for (int m = S; m < 64; m <<= 1) t += __shfl_xor(t, m);
        g[c] = t;
    }
    for (int m = S; m < 64; m <<= 1) kap += __shfl_xor(kap, m);
}

// a hub row's segment sums, added in segment order: edge segments, then triangle segments
template <typename TV, int R>
__device__ inline void step_add_segs(const StepArgs &a, int64_t i, int q, bool hubE, bool hubT, typename VT<TV>::TC *g,
                                     double &kap) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL;
    const TC *so = (const TC *)a.segout;
    if (hubE) {
        const int k = a.hE_idx[i];
        for (int s = a.hE_seg0[k]; s < a.hE_seg0[k + 1]; s++) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
        }
    }
    if (hubT) {
        const int k = a.hT_idx[i];
        for (int s = a.hT_seg0[k]; s < a.hT_seg0[k + 1]; s++) {
#pragma unroll
            for (int c = 0; c < CPL; c++) g[c] += so[(int64_t)s * R + q * CPL + c];
            kap += a.segkap[s];
        }
    }
}

template <typename TV, int R>
__device__ inline void step_wave_blk(const StepArgs &a, int blk);
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_step_wave(StepArgs a) {
    step_wave_blk<TV, R>(a, (int)blockIdx.x);
}
// workgroup blk of lb_k_step_wave (four batch rows, one wavefront each; no workgroup synchronisation)
template <typename TV, int R>
__device__ inline void step_wave_blk(const StepArgs &a, int blk) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int lane = threadIdx.x & 63, ws = lane / S, q = lane % S;
    const int64_t pos = (int64_t)blk * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    if (pos >= a.nb) return;                           // wavefront-uniform
    const int64_t i = a.rows[pos];
    TC x[CPL], g[CPL];
    VT<TV>::unpack(a.V[i * S + q], x);
    double kap;
    bool hubE, hubT;
    step_row_own<TV, R>(a, i, q, ws, x, g, kap, hubE, hubT);
    step_add_segs<TV, R>(a, i, q, hubE, hubT, g, kap);
    step_finish_wave<TV, R>(a, x, i, pos, q, ws, g, kap);
}

// Fused sweeps, hub rows: the row's own wavefront and the wavefronts that computed its segments arrive with the
// number of parts they hold (a release fence -- an L2 write-back, no invalidation -- then one atomic add); the last
// to arrive takes an acquire fence, adds the parts in the fixed order of lb_k_step_wave -- own terms, edge segments,
// triangle segments -- finishes the row, and clears the count for the next batch. No wavefront waits for another, so
// the kernel needs no co-residency.
template <typename TV, int R>
__device__ inline void hub_arrive(const StepArgs &a, int64_t i, int parts, int lane, int q, int ws) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int ke = a.hE_idx[i], kt = a.hT_idx[i];
    const int h = ke >= 0 ? ke : a.nhE + kt;
    const int total = 1 + (ke >= 0 ? a.hE_seg0[ke + 1] - a.hE_seg0[ke] : 0) +
                      (kt >= 0 ? a.hT_seg0[kt + 1] - a.hT_seg0[kt] : 0);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    int last = 0;
    if (lane == 0) last = atomicAdd(a.hubcnt + h, parts) + parts == total ? 1 : 0;
    last = __shfl(last, 0);
    if (!last) return;                                   // wavefront-uniform
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int64_t pos = a.hubpos[h];
    const double *ow = a.hubown + (int64_t)h * (R + 1);
    TC x[CPL], g[CPL];
    VT<TV>::unpack(a.V[i * S + q], x);
#pragma unroll
    for (int c = 0; c < CPL; c++) g[c] = (TC)ow[q * CPL + c];
    double kap = ow[R];
    step_add_segs<TV, R>(a, i, q, ke >= 0, kt >= 0, g, kap);
    if (lane == 0) a.hubcnt[h] = 0;
    step_finish_wave<TV, R>(a, x, i, pos, q, ws, g, kap);
}

// One launch per batch for the row steps and the hub segments: workgroups [0, nrow_blk) take four batch rows each
// (one wavefront per row, as lb_k_step_wave), the rest stride over the batch's hub segments (as lb_k_step_segs).
// Light rows finish at once; a hub row's own wavefront stores its own terms and position, and every part arrives
// through hub_arrive. Row increments go into a.zYn (no DD, no separate aggregate pass).
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_step_fused(StepArgs a, StepSegArgs ga, int nrow_blk) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int lane = threadIdx.x & 63, ws = lane / S, q = lane % S;
    if ((int)blockIdx.x < nrow_blk) {
        const int64_t pos = (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6);
        if (pos >= a.nb) return;                       // wavefront-uniform
        const int64_t i = a.rows[pos];
        TC x[CPL], g[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
        double kap;
        bool hubE, hubT;
        step_row_own<TV, R>(a, i, q, ws, x, g, kap, hubE, hubT);
        if (!hubE && !hubT) {
            step_finish_wave<TV, R>(a, x, i, pos, q, ws, g, kap);
            return;
        }
        const int ke = a.hE_idx[i];
        const int h = ke >= 0 ? ke : a.nhE + a.hT_idx[i];
        double *ow = a.hubown + (int64_t)h * (R + 1);
        if (ws == 0) {
#pragma unroll
            for (int c = 0; c < CPL; c++) ow[q * CPL + c] = (double)g[c];
        }
        if (lane == 0) { ow[R] = kap; a.hubpos[h] = pos; }
        hub_arrive<TV, R>(a, i, 1, lane, q, ws);
        return;
    }
    // each segment wavefront takes a contiguous share of the batch's list (a hub's segments are adjacent there) and
    // arrives once per hub row it worked for
    const int64_t l0 = ga.lpos[ga.b], l1 = ga.lpos[ga.b + 1];
    const int64_t nsw = (int64_t)(gridDim.x - nrow_blk) * (LB_BLOCK / 64);
    const int64_t per = (l1 - l0 + nsw - 1) / nsw;
    const int64_t wv = (int64_t)(blockIdx.x - nrow_blk) * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    const int64_t lb = l0 + wv * per, le = lb + per < l1 ? lb + per : l1;
    int64_t row = -1;
    int parts = 0;
    for (int64_t li = lb; li < le; li++) {
        const int32_t sid = ga.list[li];
        const int64_t rs = ga.segs[sid].row;
        if (rs != row && parts) {
            hub_arrive<TV, R>(a, row, parts, lane, q, ws);
            parts = 0;
        }
        row = rs;
        step_seg_one<TV, R>(ga, sid);
        parts++;
    }
    if (parts) hub_arrive<TV, R>(a, row, parts, lane, q, ws);
}

struct PostArgs {
    const int32_t *rows; int64_t nb; const double *DD; int R;
    LabelSets ls; int sumP; int use_lds;
    double *zY;                 // [z | Y(all label sets)] updated in place (FP64 atomics)
};

// z += sum DD, Y[Q] += DD for every label set (per-block LDS partials, then one atomic per block and entry)
__device__ inline void post_acc_blk(const PostArgs &a, int blk, int nblk);
__global__ __launch_bounds__(LB_BLOCK) void lb_k_post_acc(PostArgs a) {
    post_acc_blk(a, (int)blockIdx.x, (int)gridDim.x);
}
// workgroup blk of nblk of lb_k_post_acc (every thread of the workgroup calls it: it synchronises the workgroup)
__device__ inline void post_acc_blk(const PostArgs &a, int blk, int nblk) {
    extern __shared__ double sh[];
    const int width = a.R + a.sumP * a.R;
    if (a.use_lds) {
        for (int e = threadIdx.x; e < width; e += LB_BLOCK) sh[e] = 0.0;
        __syncthreads();
    }
    const int c = threadIdx.x % a.R, rs = threadIdx.x / a.R, nrs = LB_BLOCK / a.R;
    const int64_t per = (a.nb + nblk - 1) / nblk;
    const int64_t b0 = (int64_t)blk * per, b1 = b0 + per < a.nb ? b0 + per : a.nb;
    double zc = 0.0;
    for (int64_t p = b0 + rs; p < b1; p += nrs) {
        const double x = a.DD[p * a.R + c];
        zc += x;
        const int32_t i = a.rows[p];
        for (int s = 0; s < a.ls.nsets; s++) {
            const int Q = a.ls.off[s] + a.ls.lab[s][i];
            if (a.use_lds) atomicAdd(sh + a.R + (int64_t)Q * a.R + c, x);
            else atomicAdd(a.zY + a.R + (int64_t)Q * a.R + c, x);
        }
    }
    if (a.use_lds) atomicAdd(sh + c, zc); else atomicAdd(a.zY + c, zc);
    if (a.use_lds) {
        __syncthreads();
        for (int e = threadIdx.x; e < width; e += LB_BLOCK)
            if (sh[e] != 0.0) atomicAdd(a.zY + e, sh[e]);
    }
}

struct Post2Args {
    uint4 *V; const uint4 *Vstage; const int32_t *rows; int64_t nb; const double *d;
    int nvcs; int Yoff[LB_MAXVCS]; int Zoff[LB_MAXVCS]; const int32_t *lab[LB_MAXVCS]; const double *rq[LB_MAXVCS];
    const int64_t *ptr[LB_MAXVCS]; const int32_t *col; const void *sig; const double *Y; double wv, rho;
    double *stZ; double *Z; double *cnt;
    int sumPv; int use_lds;     // Z / cnt increments gathered per workgroup in LDS first (sumPv (R + 1) doubles)
    // dd set (lb_k_post_both): Y holds the VCS cluster rows at the batch start (Yoff relative to it) and every
    // workgroup adds the batch's increments dd (batch-position indexed, R each) per VCS cluster in LDS (sumPv R more
    // doubles), so the step reads the cluster rows after the batch without waiting for the aggregate pass
    const double *dd;
    const double *dYg;          // set (sweep_delta): the batch's VCS cluster-row increments, read with Y (Zoff-indexed)
};

// vcs_post: new effective multipliers SN of the moved rows (new v_i, updated Y), Z_R += d_i (SN v_new - stZ v_old),
// cnt_R += [SN > 0] - [stZ > 0], stZ = SN; then the new rows are written to V.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_post_vcs(Post2Args a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shZc[];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const int64_t pos = (int64_t)blockIdx.x * NSL + slot;
    const int wz = a.sumPv * R;
    double *Zacc = a.use_lds ? shZc : a.Z;
    double *Cacc = a.use_lds ? shZc + wz : a.cnt;
    if (a.use_lds) {
        for (int e = threadIdx.x; e < wz + a.sumPv; e += LB_BLOCK) shZc[e] = 0.0;
        __syncthreads();
    }
    if (pos < a.nb) {
    const int64_t i = a.rows[pos];
    const uint4 un = a.Vstage[pos * S + q];
    TC xn[CPL], xo[CPL];
    VT<TV>::unpack(un, xn);
    VT<TV>::unpack(a.V[i * S + q], xo);
    const double di = a.d[i];
    const TC *sig = (const TC *)a.sig;
    for (int m = 0; m < a.nvcs; m++) {
        const int Ri = a.lab[m][i];
        const int64_t ea = a.ptr[m][i], eb = a.ptr[m][i + 1];
        for (int64_t e4 = ea; e4 < eb; e4 += 4) {
          // the entries' columns, multipliers and snapshot values are loaded 4 at a time before any is used
          int Rq[4];
          double sq[4], oq[4];
#pragma unroll
          for (int t = 0; t < 4; t++) {
              const int64_t ec = e4 + t < eb ? e4 + t : eb - 1;
              Rq[t] = a.col[ec];
              sq[t] = (double)sig[ec];
              oq[t] = a.stZ[ec];
          }
          double yq[4][CPL];
#pragma unroll
          for (int t = 0; t < 4; t++) {
              const double *YR = a.Y + (int64_t)(a.Yoff[m] + Rq[t]) * R + q * CPL;
#pragma unroll
              for (int c = 0; c < CPL; c++) yq[t][c] = YR[c];
          }
          const int nq = (int)(eb - e4 < 4 ? eb - e4 : 4);
#pragma unroll
          for (int t = 0; t < 4; t++) {
            if (t >= nq) break;                              // slot-uniform
            const int64_t e = e4 + t;
            const int Rc = Rq[t];
            const double *YR = yq[t];
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < CPL; c++) s += (double)xn[c] * YR[c];
            s = slot_sum<S>(s);
            double sn = sq[t];
            if (a.rho > 0.0 && di > 0.0) {
                const double g = (Rc == Ri) ? di * (s - di) : di * s;
                sn = sn - a.rho * (a.wv * a.rq[m][Rc] / (di * di)) * g;
                if (sn < 0.0) sn = 0.0;
            }
            const double so = oq[t];
            if (sn != 0.0 || so != 0.0) {
                double *dst = Zacc + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
#pragma unroll
                for (int c = 0; c < CPL; c++) atomicAdd(dst + c, di * (sn * (double)xn[c] - so * (double)xo[c]));
            }
            if (q == 0) {
                const double dc = (sn > 0.0 ? 1.0 : 0.0) - (so > 0.0 ? 1.0 : 0.0);
                if (dc != 0.0) atomicAdd(Cacc + a.Zoff[m] + Rc, dc);
            }
            if (q == 0) a.stZ[e] = sn;
          }
        }
    }
    a.V[i * S + q] = un;
    }
    if (a.use_lds) {                                     // one global add per workgroup and nonzero entry
        __syncthreads();
        for (int e = threadIdx.x; e < wz; e += LB_BLOCK)
            if (shZc[e] != 0.0) atomicAdd(a.Z + e, shZc[e]);
        for (int e = threadIdx.x; e < a.sumPv; e += LB_BLOCK)
            if (shZc[wz + e] != 0.0) atomicAdd(a.cnt + e, shZc[wz + e]);
    }
}

// One wavefront per batch row (four rows per workgroup): the row's working-set entries are dealt to its 64 / S slots,
// one entry per slot at a time, with the same arithmetic and the same LDS staging of the Z / count increments as
// lb_k_post_vcs. The slot-0 lanes write the new V row after every slot has read the old one.
template <typename TV, int R>
__device__ inline void post_vcs_wave_blk(const Post2Args &a, int blk);
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_post_vcs_wave(Post2Args a) {
    post_vcs_wave_blk<TV, R>(a, (int)blockIdx.x);
}
// workgroup blk of lb_k_post_vcs_wave (every thread of the workgroup calls it: it synchronises the workgroup)
template <typename TV, int R>
__device__ inline void post_vcs_wave_blk(const Post2Args &a, int blk) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSW = 64 / S;
    extern __shared__ double shZc[];
    const int lane = threadIdx.x & 63, ws = lane / S, q = lane % S;
    const int64_t pos = (int64_t)blk * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    const int wz = a.sumPv * R;
    double *Zacc = a.use_lds ? shZc : a.Z;
    double *Cacc = a.use_lds ? shZc + wz : a.cnt;
    double *dY = shZc + wz + a.sumPv;                   // with dd: the batch's increments of the VCS cluster rows
    if (a.use_lds) {
        for (int e = threadIdx.x; e < wz + a.sumPv + (a.dd ? wz : 0); e += LB_BLOCK) shZc[e] = 0.0;
        __syncthreads();
    }
    if (a.dd) {
        // one batch row per thread: its cluster per VCS level, then its R increments added there
        for (int64_t p = threadIdx.x; p < a.nb; p += LB_BLOCK) {
            const int32_t i = a.rows[p];
            double xv[R];
#pragma unroll
            for (int c = 0; c < R; c++) xv[c] = a.dd[p * R + c];
            for (int m = 0; m < a.nvcs; m++) {
                double *dst = dY + (int64_t)(a.Zoff[m] + a.lab[m][i]) * R;
#pragma unroll
                for (int c = 0; c < R; c++) atomicAdd(dst + c, xv[c]);
            }
        }
        __syncthreads();
    }
    if (pos < a.nb) {                                  // wavefront-uniform
        const int64_t i = a.rows[pos];
        const uint4 un = a.Vstage[pos * S + q];
        TC xn[CPL], xo[CPL];
        VT<TV>::unpack(un, xn);
        VT<TV>::unpack(a.V[i * S + q], xo);
        const double di = a.d[i];
        const TC *sig = (const TC *)a.sig;
        for (int m = 0; m < a.nvcs; m++) {
            const int Ri = a.lab[m][i];
            const int64_t ea = a.ptr[m][i], eb = a.ptr[m][i + 1];
            for (int64_t e = ea + ws; e < eb; e += NSW) {
                const int Rc = a.col[e];
                const double sq = (double)sig[e], so = a.stZ[e];
                const double *YR = a.Y + (int64_t)(a.Yoff[m] + Rc) * R + q * CPL;
                const double *dR = (a.dYg ? a.dYg : dY) + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
                const bool add = a.dd || a.dYg;
                double s = 0.0;
#pragma unroll
                for (int c = 0; c < CPL; c++) s += (double)xn[c] * (add ? YR[c] + dR[c] : YR[c]);
                s = slot_sum<S>(s);
                double sn = sq;
                if (a.rho > 0.0 && di > 0.0) {
                    const double g = (Rc == Ri) ? di * (s - di) : di * s;
                    sn = sn - a.rho * (a.wv * a.rq[m][Rc] / (di * di)) * g;
                    if (sn < 0.0) sn = 0.0;
                }
                if (sn != 0.0 || so != 0.0) {
                    double *dst = Zacc + (int64_t)(a.Zoff[m] + Rc) * R + q * CPL;
#pragma unroll
                    for (int c = 0; c < CPL; c++) atomicAdd(dst + c, di * (sn * (double)xn[c] - so * (double)xo[c]));
                }
                if (q == 0) {
                    const double dc = (sn > 0.0 ? 1.0 : 0.0) - (so > 0.0 ? 1.0 : 0.0);
                    if (dc != 0.0) atomicAdd(Cacc + a.Zoff[m] + Rc, dc);
                    a.stZ[e] = sn;
                }
            }
        }
        if (ws == 0) a.V[i * S + q] = un;
    }
    if (a.use_lds) {                                     // one global add per workgroup and nonzero entry
        __syncthreads();
        for (int e = threadIdx.x; e < wz; e += LB_BLOCK)
            if (shZc[e] != 0.0) atomicAdd(a.Z + e, shZc[e]);
        for (int e = threadIdx.x; e < a.sumPv; e += LB_BLOCK)
            if (shZc[wz + e] != 0.0) atomicAdd(a.cnt + e, shZc[wz + e]);
    }
}

// Fused sweeps, the launch before a batch's step: workgroups [0, npost) run lb_k_post_vcs_wave for the previous batch
// (its rows' new values into V, their VCS multipliers, Z and counts at the Y after that batch), the next npair take
// the pair-level rows of the clusters this batch touches (lb_k_pair_rows, one wavefront per cluster, from the same
// Y), the rest complete z of that [z | Y] (zsrc) with the previous batch's LB_ZPART partial copies, clear those, and
// copy [z | Y] (cw doubles) to zdst, the buffer this batch's steps add into (none: zdst null). The parts touch
// disjoint memory: the first two read Y only.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_sweep_pre(Post2Args pb, int npost, PairRowArgs pa, int npair, int nclu,
                                                           double *zsrc, double *zdst, int64_t cw, double *zpart) {
    const int blk = (int)blockIdx.x;
    if (blk < npost) {
        post_vcs_wave_blk<TV, R>(pb, blk);
        return;
    }
    if (blk < npost + npair) {
        const int Qg = (blk - npost) * (LB_BLOCK / 64) + (int)(threadIdx.x >> 6);
        if (Qg < nclu) pair_row_one<R>(pa, Qg, (int)(threadIdx.x & 63));
        return;
    }
    const int64_t nthr = (int64_t)(gridDim.x - npost - npair) * LB_BLOCK;
    const int64_t ce = zdst ? cw : R;
    for (int64_t e = (int64_t)(blk - npost - npair) * LB_BLOCK + threadIdx.x; e < ce; e += nthr) {
        double v = zsrc[e];
        if (e < R) {
            for (int k = 0; k < LB_ZPART; k++) {
                v += zpart[k * R + e];
                zpart[k * R + e] = 0.0;
            }
            zsrc[e] = v;
        }
        if (zdst) zdst[e] = v;
    }
}

// A batch's aggregate increments and its VCS post step in one launch: workgroups [0, npa) run lb_k_post_acc (z and Y
// += the batch's increments, FP64 atomics), the rest lb_k_post_vcs_wave with pb.dd set (the VCS cluster rows after the
// batch formed in LDS from the batch-start snapshot and the same increments). The two write disjoint memory and the
// second reads none of what the first writes.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_post_both(PostArgs qa, int npa, Post2Args pb) {
    if ((int)blockIdx.x < npa) post_acc_blk(qa, (int)blockIdx.x, npa);
    else post_vcs_wave_blk<TV, R>(pb, (int)blockIdx.x - npa);
}

// Two launches per sweep batch (option sweep_two). lb_k_step_copy: workgroups [0, nrb) step the batch's rows
// (lb_k_step_wave, each row also recording its batch position under the batch's tag), the rest copy [Y] (cw doubles)
// to a snapshot. lb_k_post_next: the batch's aggregate increments (lb_k_post_acc), its VCS post step and the next
// batch's pair rows, both reading the snapshot plus the batch's increments summed in LDS, and the next batch's hub
// segments, which read a row stepped in this batch from Vstage (its V row is being rewritten by the same launch).
// The parts of each launch write disjoint memory, and none reads what another part of the same launch writes.
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_step_copy(StepArgs sa, int nrb, const double *src, double *dst,
                                                           int64_t cw) {
    const int blk = (int)blockIdx.x;
    if (blk < nrb) {
        step_wave_blk<TV, R>(sa, blk);
        return;
    }
    const int64_t nthr = (int64_t)(gridDim.x - nrb) * LB_BLOCK;
    for (int64_t e = (int64_t)(blk - nrb) * LB_BLOCK + threadIdx.x; e < cw; e += nthr) dst[e] = src[e];
}

// sweep_delta: [z | Y] += the batch's deltas (z from its LB_ZPART copies), and the other delta buffer cleared for the
// next batch
struct ApplyArgs { double *zY; const double *D; double *Dz; int64_t wY; };
__device__ inline void apply_blk(const ApplyArgs &a, int R, int blk, int nblk) {
    const int64_t wd = (int64_t)LB_ZPART * R + a.wY;
    for (int64_t e = (int64_t)blk * LB_BLOCK + threadIdx.x; e < wd; e += (int64_t)nblk * LB_BLOCK) {
        if (e < R) {
            double v = 0.0;
            for (int k = 0; k < LB_ZPART; k++) v += a.D[(int64_t)k * R + e];
            a.zY[e] += v;
        } else if (e >= (int64_t)LB_ZPART * R) {
            const int64_t y = e - (int64_t)LB_ZPART * R;
            a.zY[R + y] += a.D[e];
        }
        a.Dz[e] = 0.0;
    }
}

template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_post_next(PostArgs qa, int npa, Post2Args pb, int npv, PairRowArgs pa,
                                                           int npair, int nclu, StepSegArgs ga, int nsegb,
                                                           ApplyArgs ap, int nap) {
    int blk = (int)blockIdx.x;
    if (blk < npa) { post_acc_blk(qa, blk, npa); return; }
    blk -= npa;
    if (blk < nap) { apply_blk(ap, R, blk, nap); return; }
    blk -= nap;
    if (blk < npv) { post_vcs_wave_blk<TV, R>(pb, blk); return; }
    blk -= npv;
    if (blk < npair) { pair_rows_blk<R>(pa, blk, nclu); return; }
    blk -= npair;
    const int64_t l0 = ga.lpos[ga.b], l1 = ga.lpos[ga.b + 1];
    const int64_t nsw = (int64_t)nsegb * (LB_BLOCK / 64);
    for (int64_t li = l0 + (int64_t)blk * (LB_BLOCK / 64) + (threadIdx.x >> 6); li < l1; li += nsw)
        step_seg_one<TV, R>(ga, ga.list[li]);
}

// ------------------------------------------------------------------------------------------------ triangle separation
// Device form of separate_ref / separate1_ref (alm_vcs.py:965-984; _centre 554-565, _centre1 608-632, sep_hist
// 567-580, sep_collect 582-597, sep1_hist/sep1_collect 634-666). Bit-exact with the reference on the same V: every dot
// product is the FP64 sum over q in order (no contraction), the histogram bins, b0, vmin and the per-chunk caps use the
// reference's FP64 expressions, and candidates are written in the reference's generation order (chunk, centre, arm
// pair). Tie rule at the top-maxh / top-maxk cuts: descending value, ties in the prior order (stable).

#pragma clang fp contract(off)
template <typename TV, int R>
__device__ inline double dot_seq(const uint4 *V, int64_t a, int64_t b) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    double s = 0.0;
#pragma unroll
    for (int q = 0; q < S; q++) {
        TC xa[CPL], xb[CPL];
        VT<TV>::unpack(V[a * S + q], xa);
        VT<TV>::unpack(V[b * S + q], xb);
#pragma unroll
        for (int c = 0; c < CPL; c++) s = s + (double)xa[c] * (double)xb[c];
    }
    return s;
}
#pragma clang fp contract(on)

struct SepParams {
    const uint4 *V; int64_t n; const int64_t *rowptr; const int32_t *col;
    double thresh, vtol; int64_t maxh, maxk, hub2, cap2; int nchunk, nb; int64_t cs;
    // arms (edge space compacted): arm range of centre j = [abeg[j], abeg[j] + amh[j]), amh = min(m_j, maxh)
    const int64_t *apos;       // exclusive scan of arm flags over edges (n_edges + 1)
    const int32_t *arm_id; const double *arm_x;
    // one-edge arms: K lists (maxk per centre)
    const int32_t *kid; const double *kx; const int32_t *kcnt;
    // outputs / state
    unsigned long long *hist;  // nchunk x nb
    const double *vminp;       // device scalar
    int32_t *ccount;           // per centre candidate count (>= vmin)
    const int64_t *cpos;       // exclusive scan of ccount over centres
    const int64_t *chunk_cnt;  // histogram tail count per chunk (cap of the chunk)
    const int64_t *chunk_off;  // output offset per chunk
    int64_t *out; double *val; // slot arrays (sum of chunk_cnt)
};

__device__ inline int sep_bin(double vm, double vtol, int nb) {
    const int b = (int)((vm - vtol) / (2.0 - vtol) * nb);
    return b < nb - 1 ? b : nb - 1;
}

// arm values per edge: x_e = V_col[e] . V_j (reference order: s += V[i, q] * V[j, q]), flag = x_e > thresh
template <typename TV, int R>
__global__ void lb_k_sep_edges(SepParams p, double *xe, int32_t *flag) {
    const int64_t nnz = p.rowptr[p.n];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = p.n;            // row of edge e: last j with rowptr[j] <= e
        while (hi - lo > 1) {
            const int64_t mid = (lo + hi) >> 1;
            if (p.rowptr[mid] <= e) lo = mid; else hi = mid;
        }
        const double x = dot_seq<TV, R>(p.V, (int64_t)p.col[e], lo);
        xe[e] = x;
        flag[e] = x > p.thresh ? 1 : 0;
    }
}

__global__ void lb_k_sep_compact(const int64_t *rowptr, int64_t n, const int32_t *col, const double *xe,
                                 const int64_t *flag, const int64_t *apos, int32_t *arm_id, double *arm_x) {
    const int64_t nnz = rowptr[n];
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += (int64_t)gridDim.x * blockDim.x)
        if (flag[e]) { arm_id[apos[e]] = col[e]; arm_x[apos[e]] = xe[e]; }
}

// after the segmented sort of hub centres (value-descending, stable): copy their first maxh arms back in place
__global__ void lb_k_sep_cutcopy(const int64_t *seg_beg, int nseg, const int32_t *sid, const double *sx,
                                 int64_t maxh, int32_t *arm_id, double *arm_x) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < (int64_t)nseg * maxh; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t s = t / maxh, o = t % maxh, b = seg_beg[s];
        arm_id[b + o] = sid[b + o];
        arm_x[b + o] = sx[b + o];
    }
}

__device__ inline int64_t sep_mh(const SepParams &p, int64_t j, int64_t &abeg) {
    abeg = p.apos[p.rowptr[j]];
    const int64_t m = p.apos[p.rowptr[j + 1]] - abeg;
    return m < p.maxh ? m : p.maxh;
}

// one wavefront per centre, arm pairs in generation order; mode 0: histogram of vm > vtol; 1: count vm >= vmin;
// 2: write the candidates (slot = chunk offset + rank inside the chunk, if below the chunk's cap)
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_sep_wedges(SepParams p, int mode) {
    const int64_t j = (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    if (j >= p.n) return;
    const int lane = threadIdx.x & 63;
    if (p.rowptr[j + 1] - p.rowptr[j] < 2) { if (mode == 1 && lane == 0) p.ccount[j] = 0; return; }
    int64_t ab;
    const int64_t mh = sep_mh(p, j, ab);
    const int64_t ch = j / p.cs;
    const double vmin = mode ? *p.vminp : 0.0;
    int64_t cnt = 0;
    int64_t rank0 = 0, cap_end = 0;
    if (mode == 2) { rank0 = p.cpos[j] - p.cpos[ch * p.cs]; cap_end = p.chunk_cnt[ch]; }
    for (int64_t x = 0; x + 1 < mh; x++) {
        const int64_t hx = p.arm_id[ab + x];
        const double xx = p.arm_x[ab + x];
        for (int64_t y0 = x + 1; y0 < mh; y0 += 64) {
            const int64_t y = y0 + lane;
            double vm = -1.0;
            int64_t hy = 0;
            if (y < mh) {
                hy = p.arm_id[ab + y];
                const double s = dot_seq<TV, R>(p.V, hx, hy);
                vm = xx + p.arm_x[ab + y];
                vm = vm - s;
                vm = vm - 1.0;
            }
            if (mode == 0) {
                if (y < mh && vm > p.vtol) atomicAdd(p.hist + ch * p.nb + sep_bin(vm, p.vtol, p.nb), 1ull);
            } else {
                const bool ok = y < mh && vm >= vmin;
                const unsigned long long bal = __ballot(ok);
                if (mode == 2 && ok) {
                    const int64_t r = rank0 + cnt + __popcll(bal & ((1ull << lane) - 1ull));
                    if (r < cap_end) {
                        const int64_t o = p.chunk_off[ch] + r;
                        p.out[3 * o] = hx < hy ? hx : hy; p.out[3 * o + 1] = j; p.out[3 * o + 2] = hx < hy ? hy : hx;
                        p.val[o] = vm;
                    }
                }
                cnt += __popcll(bal);
            }
        }
    }
    if (mode == 1 && lane == 0) p.ccount[j] = (int32_t)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
}

// bitonic sort of nk (<= cap, power-of-two padded) (key, id) pairs in LDS by one wavefront; order: key
// ascending then id ascending (desc = false), or key descending then position ascending (desc = true)
__device__ inline void wave_bitonic(double *key, int32_t *id, int32_t *pos, int np2, bool desc) {
    const int lane = threadIdx.x & 63;
    for (int k = 2; k <= np2; k <<= 1) {
        for (int s = k >> 1; s > 0; s >>= 1) {
            for (int t = lane; t < np2; t += 64) {
                const int u = t ^ s;
                if (u > t) {
                    const bool up = (t & k) == 0;
                    bool gt;   // element t ordered after element u
                    if (desc) gt = key[t] < key[u] || (key[t] == key[u] && pos[t] > pos[u]);
                    else gt = id[t] > id[u];
                    if (gt == up) {
                        double tk = key[t]; key[t] = key[u]; key[u] = tk;
                        int32_t ti = id[t]; id[t] = id[u]; id[u] = ti;
                        int32_t tp = pos[t]; pos[t] = pos[u]; pos[u] = tp;
                    }
                }
            }
            __syncthreads();
        }
    }
}

#define LB_SEP1_CAP 2048   // 2-hop candidates per centre held in LDS (hub2 * cap2 must not exceed it)

// one-edge-arm lists K of every centre (_centre1): one workgroup of one wavefront per centre
template <typename TV, int R>
__global__ __launch_bounds__(64) void lb_k_sep_k(SepParams p, int32_t *kid, double *kx, int32_t *kcnt) {
    __shared__ double key[LB_SEP1_CAP];
    __shared__ int32_t id[LB_SEP1_CAP];
    __shared__ int32_t pos[LB_SEP1_CAP];
    __shared__ int32_t nk_s;
    const int64_t j = blockIdx.x;
    const int lane = threadIdx.x;
    if (p.rowptr[j + 1] - p.rowptr[j] < 1) { if (lane == 0) kcnt[j] = 0; return; }
    int64_t ab;
    const int64_t mh = sep_mh(p, j, ab);
    const int64_t nh = mh < p.hub2 ? mh : p.hub2;
    if (lane == 0) nk_s = 0;
    __syncthreads();
    const int64_t jlo = p.rowptr[j], jhi = p.rowptr[j + 1];
    for (int64_t a = 0; a < nh; a++) {
        const int64_t i2 = p.arm_id[ab + a];
        const int64_t lo = p.rowptr[i2];
        int64_t hi = p.rowptr[i2 + 1];
        if (hi > lo + p.cap2) hi = lo + p.cap2;
        for (int64_t e = lo + lane; e < hi; e += 64) {
            const int32_t c = p.col[e];
            bool keep = c != j;
            if (keep) {       // adjacency of c to j: binary search in j's sorted row
                int64_t l2 = jlo, h2 = jhi;
                while (l2 < h2) { const int64_t mid = (l2 + h2) >> 1; if (p.col[mid] < c) l2 = mid + 1; else h2 = mid; }
                keep = !(l2 < jhi && p.col[l2] == c);
            }
            if (keep) {
                const int t = atomicAdd(&nk_s, 1);
                if (t < LB_SEP1_CAP) { id[t] = c; key[t] = 0.0; pos[t] = t; }
            }
        }
    }
    __syncthreads();
    int nk = nk_s < LB_SEP1_CAP ? nk_s : LB_SEP1_CAP;
    int np2 = 1;
    while (np2 < nk) np2 <<= 1;
    for (int t = nk + lane; t < np2; t += 64) { id[t] = 0x7fffffff; key[t] = 0.0; pos[t] = t; }
    __syncthreads();
    wave_bitonic(key, id, pos, np2, false);           // by id (duplicates carry identical values)
    __syncthreads();
    // unique ids with value > thresh, in id order
    if (lane == 0) {
        int m = 0;
        for (int t = 0; t < nk; t++) {
            if (t > 0 && id[t] == id[t - 1]) continue;
            id[m] = id[t]; m++;
        }
        nk_s = m;
    }
    __syncthreads();
    nk = nk_s;
    for (int t = lane; t < nk; t += 64) key[t] = dot_seq<TV, R>(p.V, (int64_t)id[t], j);
    __syncthreads();
    if (lane == 0) {
        int m = 0;
        for (int t = 0; t < nk; t++)
            if (key[t] > p.thresh) { key[m] = key[t]; id[m] = id[t]; pos[m] = m; m++; }
        nk_s = m;
    }
    __syncthreads();
    nk = nk_s;
    if (nk > p.maxk) {
        np2 = 1;
        while (np2 < nk) np2 <<= 1;
        for (int t = nk + lane; t < np2; t += 64) { key[t] = -INFINITY; id[t] = 0x7fffffff; pos[t] = t; }
        __syncthreads();
        wave_bitonic(key, id, pos, np2, true);        // value descending, ties by id order (stable)
        __syncthreads();
        nk = (int)p.maxk;
    }
    for (int t = lane; t < nk; t += 64) { kid[j * p.maxk + t] = id[t]; kx[j * p.maxk + t] = key[t]; }
    if (lane == 0) kcnt[j] = nk;
}

// one wavefront per centre: pairs (arm h, 2-hop arm k) in generation order, modes as lb_k_sep_wedges
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_sep_one_edge(SepParams p, int mode) {
    const int64_t j = (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    if (j >= p.n) return;
    const int lane = threadIdx.x & 63;
    const int64_t mk = p.kcnt[j];
    if (p.rowptr[j + 1] - p.rowptr[j] < 1 || mk == 0) { if (mode == 1 && lane == 0) p.ccount[j] = 0; return; }
    int64_t ab;
    const int64_t mh = sep_mh(p, j, ab);
    const int64_t ch = j / p.cs;
    const double vmin = mode ? *p.vminp : 0.0;
    int64_t cnt = 0, rank0 = 0, cap_end = 0;
    if (mode == 2) { rank0 = p.cpos[j] - p.cpos[ch * p.cs]; cap_end = p.chunk_cnt[ch]; }
    for (int64_t x = 0; x < mh; x++) {
        const int64_t hx = p.arm_id[ab + x];
        const double xx = p.arm_x[ab + x];
        for (int64_t y0 = 0; y0 < mk; y0 += 64) {
            const int64_t y = y0 + lane;
            double vm = -1.0;
            int64_t ky = 0;
            bool valid = false;
            if (y < mk) {
                ky = p.kid[j * p.maxk + y];
                if (ky != hx) {
                    valid = true;
                    const double s = dot_seq<TV, R>(p.V, hx, ky);
                    vm = xx + p.kx[j * p.maxk + y];
                    vm = vm - s;
                    vm = vm - 1.0;
                }
            }
            if (mode == 0) {
                if (valid && vm > p.vtol) atomicAdd(p.hist + ch * p.nb + sep_bin(vm, p.vtol, p.nb), 1ull);
            } else {
                const bool ok = valid && vm >= vmin;
                const unsigned long long bal = __ballot(ok);
                if (mode == 2 && ok) {
                    const int64_t r = rank0 + cnt + __popcll(bal & ((1ull << lane) - 1ull));
                    if (r < cap_end) {
                        const int64_t o = p.chunk_off[ch] + r;
                        p.out[3 * o] = hx < ky ? hx : ky; p.out[3 * o + 1] = j; p.out[3 * o + 2] = hx < ky ? ky : hx;
                        p.val[o] = vm;
                    }
                }
                cnt += __popcll(bal);
            }
        }
    }
    if (mode == 1 && lane == 0) p.ccount[j] = (int32_t)(cnt < 0x7fffffff ? cnt : 0x7fffffff);
}

#define LB_THRESH_MAXNB 4096
// threshold from the histogram (single workgroup): b0 = first bin whose tail count <= cap, vmin, per-chunk caps and
// offsets; *total = sum of the chunk caps; *vminp = -1 if no bin qualifies
__global__ __launch_bounds__(LB_BLOCK) void lb_k_sep_thresh(const unsigned long long *hist, int nchunk, int nb,
                                                          double vtol, int64_t cap, double *vminp, int64_t *chunk_cnt,
                                                          int64_t *chunk_off, int64_t *total, int *b0p) {
    __shared__ int b0s;
    extern __shared__ unsigned long long btot[];    // per-bin totals over the chunks (nb entries) when they fit
    const bool par = nb <= LB_THRESH_MAXNB;
    if (par) {
        for (int b = threadIdx.x; b < nb; b += LB_BLOCK) {
            unsigned long long t = 0;
            for (int c = 0; c < nchunk; c++) t += hist[(int64_t)c * nb + b];
            btot[b] = t;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        unsigned long long cum = 0;
        int b0 = nb;
        // cum[b] is non-increasing in b: scan from the top bin down, remember the lowest bin with cum <= cap
        for (int b = nb - 1; b >= 0; b--) {
            unsigned long long t = 0;
            if (par) t = btot[b];
            else for (int c = 0; c < nchunk; c++) t += hist[(int64_t)c * nb + b];
            cum += t;
            if (cum <= (unsigned long long)cap) b0 = b; else break;
        }
        b0s = b0;
        *b0p = b0;
        *vminp = b0 >= nb ? -1.0 : vtol + (2.0 - vtol) * b0 / nb;
    }
    __syncthreads();
    const int b0 = b0s;
    for (int c = threadIdx.x; c < nchunk; c += LB_BLOCK) {
        unsigned long long t = 0;
        if (b0 < nb)
            for (int b = b0; b < nb; b++) t += hist[(int64_t)c * nb + b];
        chunk_cnt[c] = (int64_t)t;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t o = 0;
        for (int c = 0; c < nchunk; c++) { chunk_off[c] = o; o += chunk_cnt[c]; }
        *total = o;
    }
}


__global__ void lb_k_widen(const int32_t *in, int64_t cnt, int64_t *out) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) out[e] = in[e];
}

// arm range starts of every centre: out[j] = apos[rowptr[j]] (j = 0..n)
__global__ void lb_k_gather_pos(const int64_t *rowptr, int64_t n, const int64_t *apos, int64_t *out) {
    for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j <= n; j += (int64_t)gridDim.x * blockDim.x) out[j] = apos[rowptr[j]];
}


// ------------------------------------------------------------------------------------------------ VCS separation
// vcsops.separate (vcsops.py:395-436; _sep_rows 334-366): per vertex i and cluster R, s = V_i . Y_R (FP64), own
// cluster s -= d_i and vol -= d_i, violated if s / vol < -tol; entries outside the working set are candidates with
// score pen = wv rq_R s^2 | mean = -s / vol | mass = -d_i s (> 0). One thread per vertex; Y staged in LDS.
struct VcsSepArgs {
    const uint4 *V; int64_t n; const double *d; const int32_t *lab; const int64_t *ptr; const int32_t *col;
    const double *Y; const double *vol; const double *rq; int P; double wv, tol; int mode, use_lds;
    const int64_t *base;      // write pass: output offset per vertex
    int64_t *cnt;             // count pass: candidates per vertex
    int32_t *oi; int32_t *oR; double *os;
    double *stats;            // per block: #viol, mass, #viol in ws, mass in ws
};

template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_vcs_sep(VcsSepArgs a, int write) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    extern __shared__ double shY[];
    __shared__ double red[LB_BLOCK / 64];
    const double *Y = a.Y;
    if (a.use_lds) {
        for (int e = threadIdx.x; e < a.P * R; e += LB_BLOCK) shY[e] = a.Y[e];
        __syncthreads();
        Y = shY;
    }
    double nv = 0.0, mv = 0.0, nw = 0.0, mw = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * LB_BLOCK + threadIdx.x; i < a.n; i += (int64_t)gridDim.x * LB_BLOCK) {
        double x[R];
#pragma unroll
        for (int qq = 0; qq < S; qq++) {
            TC t[CPL];
            VT<TV>::unpack(a.V[i * S + qq], t);
#pragma unroll
            for (int c = 0; c < CPL; c++) x[qq * CPL + c] = (double)t[c];
        }
        const double di = a.d[i];
        const int Ri = a.lab[i];
        int64_t q = a.ptr[i];
        const int64_t qe = a.ptr[i + 1];
        int64_t cnt = 0;
        int64_t o = write ? a.base[i] : 0;
        for (int Rc = 0; Rc < a.P; Rc++) {
            const double *YR = Y + (int64_t)Rc * R;
            double s = 0.0;
#pragma unroll
            for (int c = 0; c < R; c++) s += x[c] * YR[c];
            double vR = a.vol[Rc];
            if (Rc == Ri) { s = s - di; vR = vR - di; }
            if (vR <= 0.0) continue;
            const double xb = s / vR;
            if (xb >= -a.tol) continue;
            while (q < qe && a.col[q] < Rc) q++;
            const bool inws = q < qe && a.col[q] == Rc;
            nv += 1.0; mv += -di * s;
            if (inws) { nw += 1.0; mw += -di * s; continue; }
            const double sc = a.mode == 0 ? a.wv * a.rq[Rc] * s * s : (a.mode == 1 ? -xb : -di * s);
            if (sc <= 0.0) continue;
            if (write) { a.oi[o] = (int32_t)i; a.oR[o] = Rc; a.os[o] = sc; o++; }
            cnt++;
        }
        if (!write) a.cnt[i] = cnt;
    }
    if (!write) {
        nv = block_sum(nv, red); __syncthreads();
        mv = block_sum(mv, red); __syncthreads();
        nw = block_sum(nw, red); __syncthreads();
        mw = block_sum(mw, red);
        if (threadIdx.x == 0) {
            double *st = a.stats + 4 * (int64_t)blockIdx.x;
            st[0] = nv; st[1] = mv; st[2] = nw; st[3] = mw;
        }
    }
}

__global__ void lb_k_iota64(int64_t *x, int64_t cnt) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) x[e] = e;
}

__global__ void lb_k_take(const int64_t *idx, int64_t cnt, const int32_t *oi, const int32_t *oR, const double *os,
                          int32_t *ti, int32_t *tR, double *ts) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t k = idx[e];
        ti[e] = oi[k]; tR[e] = oR[k]; ts[e] = os[k];
    }
}

// ================================================================================================== host side
namespace {

const char *kKernelNames[LB_NKERNELS] = {"lb_k_colsum", "lb_k_reduce(z,Y)", "lb_k_levels", "lb_k_vcs",
                                         "lb_k_reduce(Z)", "lb_k_segs", "lb_k_rows", "lb_k_finalize"};

struct DBuf {
    void *p = nullptr; size_t bytes = 0;
};

}  // namespace

// ------------------------------------------------------------------------------------------------ device RGD polish
// Row work of the estimate's Riemannian gradient polish (polish_rgd, est_dual_vcs.py:276-308) on an FP64 master copy
// of the iterate: the step-and-normalise, the row values g_i = (C W)_i . W_i with their sum, and the new direction
// R = (C W - g W) / max(|C W|, 1e-12) with the Barzilai-Borwein sums. Two FP64 columns per lane (S = R / 2 lanes per
// row); products and sums written without contraction so each rounds as the host form does.
struct PolArgs {
    int64_t n; double alpha; int init;
    const double *V;           // current iterate (FP64 master, n x r)
    double *Rd;                // current direction (n x r), replaced by the new one in lb_k_pol_post
    double *Vn;                // candidate iterate (FP64)
    void *Vst;                 // the context's V storage (the evaluation reads it)
    const void *G;             // C(y) Vn from the evaluation (TC)
    double *g;                 // row values of the candidate
    double *part;              // per-block partial sums
};

template <typename TV> __device__ inline void pol_store(void *Vst, int64_t e, double x);
template <> __device__ inline void pol_store<float>(void *Vst, int64_t e, double x) { ((float *)Vst)[e] = (float)x; }
template <> __device__ inline void pol_store<bf16raw>(void *Vst, int64_t e, double x) { ((uint16_t *)Vst)[e] = lb_f2bf((float)x); }
template <> __device__ inline void pol_store<double>(void *Vst, int64_t e, double x) { ((double *)Vst)[e] = x; }

// Vn = (V + alpha R) / |V + alpha R| row by row, and its storage copy
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_pol_step(PolArgs a) {
    constexpr int S = R / 2, NSL = LB_BLOCK / S;
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        const int64_t e = i * R + 2 * q;
        const double x0 = __dadd_rn(a.V[e], __dmul_rn(a.alpha, a.Rd[e]));
        const double x1 = __dadd_rn(a.V[e + 1], __dmul_rn(a.alpha, a.Rd[e + 1]));
        const double nrm = sqrt(slot_sum<S>(__dadd_rn(__dmul_rn(x0, x0), __dmul_rn(x1, x1))));
        const double y0 = x0 / nrm, y1 = x1 / nrm;
        a.Vn[e] = y0; a.Vn[e + 1] = y1;
        pol_store<TV>(a.Vst, e, y0);
        pol_store<TV>(a.Vst, e + 1, y1);
    }
}

// g_i = G_i . Vn_i (FP64) and the per-block sum of the g_i
template <typename TC, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_pol_gsum(PolArgs a) {
    constexpr int S = R / 2, NSL = LB_BLOCK / S;
    __shared__ double red[LB_BLOCK / 64];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const TC *G = (const TC *)a.G;
    double acc = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        const int64_t e = i * R + 2 * q;
        const double gi = slot_sum<S>(__dadd_rn(__dmul_rn((double)G[e], a.Vn[e]), __dmul_rn((double)G[e + 1], a.Vn[e + 1])));
        if (q == 0) { a.g[i] = gi; acc += gi; }
    }
    acc = block_sum(acc, red);
    if (threadIdx.x == 0) a.part[blockIdx.x] = acc;
}

// New direction from G at Vn: nr = max(|G_i|, 1e-12), Rr = G - g Vn, Rn = Rr / nr; per-block sums of
// Rr . Rn and (unless init) of S.(Y nr), S.(S nr), Y.(Y nr) with S = Vn - V, Y = Rn - R; Rd <- Rn.
template <typename TC, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_pol_post(PolArgs a) {
    constexpr int S = R / 2, NSL = LB_BLOCK / S;
    __shared__ double red[LB_BLOCK / 64];
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const TC *G = (const TC *)a.G;
    double gr = 0.0, sy = 0.0, ss = 0.0, yy = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NSL + slot; i < a.n; i += (int64_t)gridDim.x * NSL) {
        const int64_t e = i * R + 2 * q;
        const double c0 = (double)G[e], c1 = (double)G[e + 1];
        const double nr = fmax(sqrt(slot_sum<S>(__dadd_rn(__dmul_rn(c0, c0), __dmul_rn(c1, c1)))), 1e-12);
        const double gi = a.g[i];
        const double v0 = a.Vn[e], v1 = a.Vn[e + 1];
        const double r0 = __dsub_rn(c0, __dmul_rn(gi, v0)), r1 = __dsub_rn(c1, __dmul_rn(gi, v1));
        const double n0 = r0 / nr, n1 = r1 / nr;
        gr += __dadd_rn(__dmul_rn(r0, n0), __dmul_rn(r1, n1));
        if (!a.init) {
            const double s0 = __dsub_rn(v0, a.V[e]), s1 = __dsub_rn(v1, a.V[e + 1]);
            const double y0 = __dsub_rn(n0, a.Rd[e]), y1 = __dsub_rn(n1, a.Rd[e + 1]);
            sy += __dadd_rn(__dmul_rn(s0, __dmul_rn(y0, nr)), __dmul_rn(s1, __dmul_rn(y1, nr)));
            ss += __dadd_rn(__dmul_rn(s0, __dmul_rn(s0, nr)), __dmul_rn(s1, __dmul_rn(s1, nr)));
            yy += __dadd_rn(__dmul_rn(y0, __dmul_rn(y0, nr)), __dmul_rn(y1, __dmul_rn(y1, nr)));
        }
        a.Rd[e] = n0; a.Rd[e + 1] = n1;
    }
    gr = block_sum(gr, red);
    sy = block_sum(sy, red);
    ss = block_sum(ss, red);
    yy = block_sum(yy, red);
    if (threadIdx.x == 0) {
        double *o = a.part + (int64_t)blockIdx.x * 4;
        o[0] = gr; o[1] = sy; o[2] = ss; o[3] = yy;
    }
}

// Hub-segment tables of the sweeps. A sweep batch holds a few hundred rows and runs as long as its longest row, so
// the sweeps split rows at a lower threshold into shorter segments than the evaluation (which is tuned for
// throughput over all rows).
struct SweepSegs {
    int64_t heavyE = 32, heavyT = 32; int seg_len = 128;
    std::vector<int32_t> hE_rows, hE_seg0, hT_rows, hT_seg0, hE_idx, hT_idx;
    std::vector<Seg> segE, segT;
    DBuf hErows, hEseg0, hTrows, hTseg0, segs, segout, segkap;
    DBuf hEidx, hTidx;                  // row -> its index in hE_rows / hT_rows, or -1 (one load, no search)
};

struct lb_ctx {
    int device = 0, vdtype = LB_F32, r = 16;
    hipStream_t stream = nullptr;
    // the evaluation's hub-segment kernel runs on `side` while the cluster-sum / VCS kernels run on `stream`
    // (fork / join through the two events); option "overlap" 0 keeps the whole chain on `stream`
    hipStream_t side = nullptr;
    hipEvent_t fork_ev = nullptr, join_ev = nullptr;
    bool overlap = true;
    int64_t seg_grid = 0;              // workgroups of the evaluation's hub-segment kernel (0: one per 4 segments)
    bool step_wave = true;             // sweeps: one wavefront per batch row (lb_k_step_wave, lb_k_post_vcs_wave) or one
                                       // slot per row (lb_k_step, lb_k_post_vcs)
    std::string err;
    size_t vbytes = 4, cbytes = 4;   // storage bytes of V elements, of TC values
    int heavy_edges = 128, heavy_tris = 128, seg_len = 1024;  // rows above these counts are split into segments
    int64_t row_grid = 8192;                         // workgroups of the grid-stride row kernel
    int64_t built_heavyE = 64, built_heavyT = 64;   // thresholds the current segment tables were built with
    // graph
    int64_t n = 0, nnz = 0; double L = 1.0, C0 = 0.0; bool weighted = false;
    std::vector<int64_t> row_order;   // empty, or: device row k holds the caller's row row_order[k] (V, G transfers)
    void *hstage = nullptr; size_t hstage_bytes = 0;   // pinned host staging of the V / G transfers
    std::vector<int64_t> h_rowptr;
    // levels
    int nlev = 0; std::vector<int> P; std::vector<int64_t> soff; double pw = 1.0; int sumPpair = 0;
    // host copies of the pair-level and VCS-level labels; vdup[m] = a pair level with the same labels and cluster
    // count as VCS level m (its cluster sums are then copied, not accumulated twice), or -1
    std::vector<int32_t> h_labs, h_vlabs; std::vector<int> vdup;
    // vcs
    int nvcs = 0; std::vector<int> vP; std::vector<int64_t> vnnz, voff; double wv = 1.0; int sumPv = 0;
    // triangles
    int64_t T = 0;
    std::vector<int32_t> hE_rows, hE_seg0, hT_rows, hT_seg0;
    std::vector<Seg> segE, segT;
    bool have_eval = false; double last_rho = 0.0;
    std::vector<int32_t> hE_idx, hT_idx;   // row -> index into hE_rows / hT_rows, or -1
    SweepSegs sw;                          // the sweeps' own hub-segment tables
    std::vector<int64_t> h_iptr;            // triangle incidence pointer (host copy)
    // device buffers
    DBuf vrow;                        // VCS working-set entry -> its row (entry-parallel VCS passes)
    DBuf vst_part;                    // per-workgroup [Z | cnt] partials of the sweeps' VCS state pass
    DBuf b_light, b_heavy;            // batch rows by work: one slot per row / one wavefront per row
    int64_t b_nlight = 0, b_nheavy = 0; bool b_lists_ok = false;
    DBuf pol_V, pol_Vn, pol_R, pol_g, pol_part, pol_sum;   // device RGD polish state (n x r FP64, n, partials)
    // block workspace of the device eigensolver (lb_blk_*): nblk blocks of n x r FP64 and ndiag n-vectors, device
    // row order; gram partials and combination coefficients
    DBuf blk, blkdiag, blkpart, blkcoef, blkout; int nblk = 0, ndiag = 0;
    bool pol_ok = false;
    DBuf rowptr, col, w, d, labs, sig, pp, rq, dsq, vlabs, vrq, vptr, vcol, vsig, pv, TI, lam, pt, iptr, rec,
        V, G, Y, z, M, pdiag, levtot, levviol, Z, colsum_part, vcs_part, row_part, seg_part, segs, segout,
        hErows, hEseg0, hTrows, hTseg0, parts, lamsum,
        order, tmask, Kq, stZ, Zs, cnts, Vstage, DDb, segkap, seglist, vslack, vage, vvol;
    std::vector<double> h_d;                // vertex volumes (host copy)
    DBuf rorder;                            // row processing order of lb_k_rows (descending work)
    // option row_order 1: lb_k_rows takes the rows by descending work (rorder); 0 (default): in index order. With
    // the vertices numbered by degree (the Engine's default) index order is also by work, and a wavefront's rows,
    // their CSR and record ranges share cache lines: on the 1M-row graph 0.598 against 0.618 ms per evaluation,
    // lb_k_rows hitting L2 on 51% of its requests against 40%
    bool use_order = false;
    // VCS pass with LDS-resident sums: 1 lb_k_vcs_mf (one pass, Z on the matrix cores), 0 lb_k_vcs_pv + lb_k_vcs_acc;
    // -1 picks by storage type (measured on the 1M-row graph: one pass for FP32 and FP64 V, two passes for BF16 V)
    int vcs_kernel = -1;
    // workgroups of the VCS pass, at most LB_VCS_BLOCKS (the partials' allocation); 0 picks by kernel (measured:
    // 512 for the one-pass kernel, LB_VCS_BLOCKS for the others)
    int64_t vcs_grid = 0;
    // rows per wavefront tile of the one-pass VCS kernel: 32 when its LDS fits 64 KB (FP32 / BF16 V at P <= 64), else
    // 16; measured on the 1M-row graph: 140 against 150 us
    int vcs_tile = 32;
    // batch mode: the graph is the block-diagonal union of bG graphs with contiguous row ranges
    int bG = 0; std::vector<int64_t> hboff; std::vector<int> bP; int bsumP = 0; bool blev = false;
    DBuf b_off, b_gid, b_L, b_C0, b_Poff, b_soff, b_cg, b_lab, b_sig, b_pp, b_rq, b_dsq, b_zY, b_M, b_pdiag, b_ltot,
        b_lviol, b_rowv, b_parts, b_lamsum, b_tptr, b_tidx;
    double bw = 1.0;
    void *gexec = nullptr; double g_rho = 0.0;   // captured evaluation (hipGraphExec_t)
    bool b_lamsum_ok = false;                    // batch: per-graph lam sums up to date
    int64_t b_nsig = 0;                          // batch: pair-level multipliers (sum_g P_g^2)
    int64_t b_light_max = LB_BLIGHT, b_lgrid = 8192, b_hgrid = 8192;   // batch launch shapes (lb_set_option)
    DBuf b_R, b_Vp, b_Rp, b_rgst;                // batched search: gradient, previous iterate / gradient, per-graph step
    void *rgd_exec = nullptr;                    // captured search iteration (hipGraphExec_t) and its parameters
    // option "sweep_graph" 1: a sweep's per-batch launches captured as one graph (hipGraphExec_t), replayed while the
    // launch arguments (the key: every argument block, the batch bounds and grids) are unchanged. Off by default: on
    // a 150k-vertex request (512 batches per sweep, 24 sweeps) the replayed sweeps took 0.49 s against 0.44 s eager
    // (measured with five launches per batch), since a batch is bound by its kernels' own durations (about 33 of its
    // 36 us), not by the launches.
    void *sw_exec = nullptr; std::vector<unsigned char> sw_key; bool sweep_graph = false;
    DBuf lposd;                                  // device copy of the sweep's per-batch segment-list bounds
    // option "sweep_fused" 1 (with step_wave): two launches per batch (lb_k_sweep_pre, lb_k_step_fused) in place of
    // five; the batches' [z | Y] alternate between Y and Yalt; hub rows meet through hubcnt / hubpos / hubown. Off by
    // default: one run of a 16-request serving stream ended in a GPU memory fault inside a sweep of this form (after
    // two clean runs of the same stream) and its cause is not established
    bool sweep_fused = false;
    DBuf Yalt, hubcnt, hubpos, hubown, zpart;
    // workgroups of the sweeps' hub segments (0: 64); on a 150k-vertex request's search 0.326 s of sweeps at 64
    // against 0.335 s at 16
    int64_t sweep_seg_grid = 0;
    // option "sweep_pair_segs" 1 (default): a batch's pair-level rows and hub segments share one launch
    // (lb_k_pair_segs) in place of lb_k_pair_rows + lb_k_step_segs
    bool sweep_pair_segs = true;
    // option "sweep_post_merge" 1 (default, with sweep_pair_segs and step_wave): a batch's aggregate increments and
    // VCS post step in one launch (lb_k_post_both), the VCS cluster rows at the batch start copied to Ysnap by the
    // batch's first launch; used for batches of at most LB_POST_MERGE_ROWS (256) rows where the post step's LDS
    // (sumPv (2R + 1) doubles) fits 48 KB
    bool sweep_post_merge = true;
    DBuf Ysnap;
    // option "sweep_two" 1 (default, with step_wave, sweep_pair_segs and sweep_post_merge, eager sweeps): two launches
    // per batch (lb_k_step_copy, lb_k_post_next); posb / posp: each row's last batch tag and position (tags unique
    // within one lb_sweeps call, posb reset to -1 at its start)
    bool sweep_two = true;
    DBuf posb, posp;
    int sw_tag = 0;
    // option "sweep_delta" 1 (default, with sweep_two, for batches above 128 rows): the rows add their increments to a
    // per-batch delta buffer (two, alternating by tag) with FP64 atomics; the next launch reads snapshot + delta and
    // applies the delta to [z | Y]
    bool sweep_delta = true;
    DBuf dbuf;
    double rgd_key[5] = {0, 0, 0, 0, 0};
};

namespace {

int fail(lb_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    return code;
}

// any change of the context's buffers or of the scalars passed to the kernels ends a captured evaluation
void drop_capture(lb_ctx *c) {
    if (c && c->gexec) {
        hipGraphExecDestroy((hipGraphExec_t)c->gexec);
        c->gexec = nullptr;
    }
    if (c && c->rgd_exec) {
        hipGraphExecDestroy((hipGraphExec_t)c->rgd_exec);
        c->rgd_exec = nullptr;
    }
}

#define HIPCHK(ctx, call)                                                                                   \
    do {                                                                                                    \
        hipError_t e_ = (call);                                                                             \
        if (e_ != hipSuccess) return fail(ctx, -3, std::string(#call) + ": " + hipGetErrorString(e_));      \
    } while (0)

int dalloc(lb_ctx *c, DBuf &b, size_t bytes) {
    if (bytes == 0) bytes = 16;
    if (b.p && b.bytes >= bytes) return 0;
    if (b.p) { hipFree(b.p); b.p = nullptr; b.bytes = 0; }
    hipError_t e = hipMalloc(&b.p, bytes);
    if (e != hipSuccess) return fail(c, -2, std::string("hipMalloc of ") + std::to_string(bytes) + " bytes: " + hipGetErrorString(e));
    b.bytes = bytes;
    return 0;
}

void dfree(DBuf &b) {
    if (b.p) hipFree(b.p);
    b.p = nullptr; b.bytes = 0;
}

int upload(lb_ctx *c, DBuf &b, const void *src, size_t bytes) {
    int rc = dalloc(c, b, bytes);
    if (rc) return rc;
    if (bytes && src) HIPCHK(c, hipMemcpyAsync(b.p, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// host FP64 values -> device TC values (float or double)
int upload_tc(lb_ctx *c, DBuf &b, const double *src, int64_t cnt) {
    if (c->cbytes == 8) return upload(c, b, src, sizeof(double) * (size_t)cnt);
    std::vector<float> tmp((size_t)std::max<int64_t>(cnt, 1));
    for (int64_t e = 0; e < cnt; e++) tmp[e] = (float)src[e];
    return upload(c, b, tmp.data(), sizeof(float) * (size_t)cnt);
}

int download_tc(lb_ctx *c, const DBuf &b, double *dst, int64_t cnt) {
    if (!cnt) return 0;
    if (c->cbytes == 8) {
        HIPCHK(c, hipMemcpyAsync(dst, b.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return 0;
    }
    std::vector<float> tmp(cnt);
    HIPCHK(c, hipMemcpyAsync(tmp.data(), b.p, sizeof(float) * cnt, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int64_t e = 0; e < cnt; e++) dst[e] = tmp[e];
    return 0;
}

uint16_t f2bf(float f) {  // round to nearest even
    uint32_t u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf2f(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// Pinned host staging buffer of at least `bytes`, kept by the context (the V / G copies go through it at DMA speed).
int host_stage(lb_ctx *c, size_t bytes, void **out) {
    if (c->hstage_bytes < bytes) {
        if (c->hstage) hipHostFree(c->hstage);
        c->hstage = nullptr; c->hstage_bytes = 0;
        HIPCHK(c, hipHostMalloc(&c->hstage, bytes, hipHostMallocDefault));
        c->hstage_bytes = bytes;
    }
    *out = c->hstage;
    return 0;
}

// f(a, b) over row ranges [a, b) of [0, n), on up to 8 host threads for large n.
template <class F>
void par_rows(int64_t n, F f) {
    const int nt = (int)std::min<int64_t>(8, std::max<int64_t>(1, n / 32768));
    if (nt <= 1) { f((int64_t)0, n); return; }
    const int64_t cs = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++) {
        const int64_t a = t * cs, b = std::min<int64_t>(n, a + cs);
        if (a < b) th.emplace_back(f, a, b);
    }
    for (auto &x : th) x.join();
}

// Host rows -> device storage rows: device row k <- host row ord[k] (or k), converted to `dt`.
template <typename HT>
void rows_to_storage(int64_t a, int64_t b, int r, const HT *src, int dt, void *dst, const int64_t *ord) {
    for (int64_t k = a; k < b; k++) {
        const HT *s = src + (ord ? ord[k] : k) * r;
        if (dt == LB_F64) { double *d = (double *)dst + k * r; for (int j = 0; j < r; j++) d[j] = (double)s[j]; }
        else if (dt == LB_F32) { float *d = (float *)dst + k * r; for (int j = 0; j < r; j++) d[j] = (float)s[j]; }
        else { uint16_t *d = (uint16_t *)dst + k * r; for (int j = 0; j < r; j++) d[j] = f2bf((float)s[j]); }
    }
}

// Device storage rows (type `st`: LB_F64, LB_F32 or LB_BF16) -> host rows: host row ord[k] (or k) <- device row k.
template <typename HT>
void storage_to_rows(int64_t a, int64_t b, int r, const void *src, int st, HT *dst, const int64_t *ord) {
    for (int64_t k = a; k < b; k++) {
        HT *d = dst + (ord ? ord[k] : k) * r;
        if (st == LB_F64) { const double *s = (const double *)src + k * r; for (int j = 0; j < r; j++) d[j] = (HT)s[j]; }
        else if (st == LB_F32) { const float *s = (const float *)src + k * r; for (int j = 0; j < r; j++) d[j] = (HT)s[j]; }
        else { const uint16_t *s = (const uint16_t *)src + k * r; for (int j = 0; j < r; j++) d[j] = (HT)bf2f(s[j]); }
    }
}

void build_segments(int64_t n, const int64_t *ptr, int64_t heavy, int seg_len, int kind, std::vector<int32_t> &rows,
                    std::vector<int32_t> &seg0, std::vector<Seg> &segs, int32_t base) {
    rows.clear(); seg0.clear(); segs.clear();
    for (int64_t i = 0; i < n; i++) {
        const int64_t b = ptr[i], e = ptr[i + 1];
        if (e - b <= heavy) continue;
        rows.push_back((int32_t)i);
        seg0.push_back(base + (int32_t)segs.size());
        for (int64_t s = b; s < e; s += seg_len) segs.push_back(Seg{(int32_t)i, kind, s, std::min<int64_t>(e, s + seg_len)});
    }
    seg0.push_back(base + (int32_t)segs.size());
}

// Row processing order of lb_k_rows: descending per-row work (light edges + 2 x light triangle records, or the
// number of hub segments summed), stable in the row index (counting sort). Rows of similar work then share a
// wavefront, so its slots finish together.
int build_row_order(lb_ctx *c) {
    const int64_t n = c->n;
    if (!n) return 0;
    const int64_t W = 2048;
    std::vector<int32_t> w((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const int64_t de = c->h_rowptr[i + 1] - c->h_rowptr[i];
        const int64_t dt = c->h_iptr.size() == (size_t)n + 1 ? c->h_iptr[i + 1] - c->h_iptr[i] : 0;
        const int64_t we = de > c->built_heavyE ? (de + c->seg_len - 1) / c->seg_len : de;
        const int64_t wt = dt > c->built_heavyT ? (dt + c->seg_len - 1) / c->seg_len : dt;
        w[i] = (int32_t)std::min<int64_t>(W - 1, we + 2 * wt);
    }
    std::vector<int64_t> cnt(W + 1, 0);
    for (int64_t i = 0; i < n; i++) cnt[W - 1 - w[i] + 1]++;
    for (int64_t b = 0; b < W; b++) cnt[b + 1] += cnt[b];
    std::vector<int32_t> ord((size_t)n);
    for (int64_t i = 0; i < n; i++) ord[cnt[W - 1 - w[i]]++] = (int32_t)i;
    return upload(c, c->rorder, ord.data(), sizeof(int32_t) * (size_t)n);
}

// the sweeps' tables from the host row pointers and the triangle incidence counts (h_rowptr, h_iptr)
int build_sweep_segs(lb_ctx *c) {
    SweepSegs &w = c->sw;
    const int64_t n = c->n;
    int rc;
    build_segments(n, c->h_rowptr.data(), w.heavyE, w.seg_len, 0, w.hE_rows, w.hE_seg0, w.segE, 0);
    if (c->h_iptr.size() == (size_t)n + 1)
        build_segments(n, c->h_iptr.data(), w.heavyT, w.seg_len, 1, w.hT_rows, w.hT_seg0, w.segT, (int32_t)w.segE.size());
    else { w.hT_rows.clear(); w.hT_seg0.assign(1, (int32_t)w.segE.size()); w.segT.clear(); }
    w.hE_idx.assign(n, -1);
    for (size_t k = 0; k < w.hE_rows.size(); k++) w.hE_idx[w.hE_rows[k]] = (int32_t)k;
    w.hT_idx.assign(n, -1);
    for (size_t k = 0; k < w.hT_rows.size(); k++) w.hT_idx[w.hT_rows[k]] = (int32_t)k;
    std::vector<Seg> all(w.segE);
    all.insert(all.end(), w.segT.begin(), w.segT.end());
    const size_t ns = std::max<size_t>(all.size(), 1);
    if ((rc = dalloc(c, w.hErows, sizeof(int32_t) * std::max<size_t>(w.hE_rows.size(), 1))) ||
        (rc = dalloc(c, w.hTrows, sizeof(int32_t) * std::max<size_t>(w.hT_rows.size(), 1))))
        return rc;
    if (!w.hE_rows.empty() && (rc = upload(c, w.hErows, w.hE_rows.data(), sizeof(int32_t) * w.hE_rows.size()))) return rc;
    if (!w.hT_rows.empty() && (rc = upload(c, w.hTrows, w.hT_rows.data(), sizeof(int32_t) * w.hT_rows.size()))) return rc;
    if ((rc = upload(c, w.hEseg0, w.hE_seg0.data(), sizeof(int32_t) * w.hE_seg0.size()))) return rc;
    if ((rc = upload(c, w.hTseg0, w.hT_seg0.data(), sizeof(int32_t) * w.hT_seg0.size()))) return rc;
    if ((rc = upload(c, w.hEidx, w.hE_idx.data(), sizeof(int32_t) * w.hE_idx.size()))) return rc;
    if ((rc = upload(c, w.hTidx, w.hT_idx.data(), sizeof(int32_t) * w.hT_idx.size()))) return rc;
    if ((rc = dalloc(c, w.segs, sizeof(Seg) * ns))) return rc;
    if (!all.empty() && (rc = upload(c, w.segs, all.data(), sizeof(Seg) * all.size()))) return rc;
    if ((rc = dalloc(c, w.segout, ns * c->r * c->cbytes))) return rc;
    if ((rc = dalloc(c, w.segkap, sizeof(double) * ns))) return rc;
    return 0;
}

int rebuild_segs(lb_ctx *c) {
    std::vector<Seg> all(c->segE);
    all.insert(all.end(), c->segT.begin(), c->segT.end());
    int rc;
    if ((rc = build_row_order(c))) return rc;
    if ((rc = build_sweep_segs(c))) return rc;
    if ((rc = upload(c, c->segs, all.empty() ? nullptr : all.data(), sizeof(Seg) * all.size()))) return rc;
    if ((rc = dalloc(c, c->segout, (size_t)std::max<size_t>(all.size(), 1) * c->r * c->cbytes))) return rc;
    const int nblk = (int)((all.size() + 3) / 4);
    if ((rc = dalloc(c, c->seg_part, sizeof(double) * 2 * std::max(nblk, 1)))) return rc;
    return 0;
}

template <typename TV, int R>
struct Launch {
    typedef typename VT<TV>::TC TC;
    static constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;

    // z = d^T V and the cluster sums of every label set into c->Y = [z | Y]
    static void zY(lb_ctx *c, hipEvent_t *ev = nullptr, int *ek = nullptr) {
        const int64_t n = c->n;
        const int sumPall = c->sumPpair + c->sumPv;
        ColsumArgs ca{};
        ca.V = (const uint4 *)c->V.p; ca.n = n; ca.d = (const double *)c->d.p;
        // a VCS level with the labels of a pair level is not accumulated again: its rows of Y stay zero in the
        // partials and are copied from the pair level's after the reduction
        int off = 0, ns = 0;
        std::vector<int> poff(c->nlev);
        for (int l = 0; l < c->nlev; l++) {
            poff[l] = off;
            ca.ls.off[ns] = off; ca.ls.lab[ns++] = (const int32_t *)c->labs.p + (int64_t)l * n; off += c->P[l];
        }
        std::vector<std::array<int, 3>> dups;
        for (int m = 0; m < c->nvcs; m++) {
            const int dl = m < (int)c->vdup.size() ? c->vdup[m] : -1;
            if (dl >= 0) dups.push_back({poff[dl], off, c->vP[m]});
            else { ca.ls.off[ns] = off; ca.ls.lab[ns++] = (const int32_t *)c->vlabs.p + (int64_t)m * n; }
            off += c->vP[m];
        }
        ca.ls.nsets = ns;
        ca.ls.off[ns] = off;
        ca.sumP = sumPall;
        const size_t ldsY = sizeof(double) * (size_t)sumPall * (R + LB_VPAD);
        ca.use_lds = ldsY <= 48 * 1024;
        ca.part = (double *)c->colsum_part.p;
        ca.Yglob = (double *)c->Y.p + R;
        if (!ca.use_lds && sumPall) hipLaunchKernelGGL(lb_k_fill, dim3(256), dim3(256), 0, c->stream, (double *)c->Y.p + R, (int64_t)sumPall * R, 0.0);
        hipLaunchKernelGGL((lb_k_colsum<TV, R>), dim3(LB_RED_BLOCKS), dim3(LB_BLOCK), ca.use_lds ? ldsY : 0, c->stream, ca);
        if (ev) hipEventRecord(ev[(*ek)++], c->stream);
        const int width = R + (ca.use_lds ? sumPall * R : 0);
        hipLaunchKernelGGL(lb_k_reduce, dim3((width + 63) / 64), dim3(1024), 0, c->stream, (const double *)c->colsum_part.p,
                           LB_RED_BLOCKS, width, width, (double *)c->Y.p);
        double *Yb = (double *)c->Y.p + R;
        for (const auto &dp : dups) {                       // (pair-level offset, VCS-level offset, clusters)
            const int64_t cnt = (int64_t)dp[2] * R;
            hipLaunchKernelGGL((lb_k_copy<double>), dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream,
                               Yb + (int64_t)dp[1] * R, (const double *)(Yb + (int64_t)dp[0] * R), cnt);
        }
    }

    static int eval(lb_ctx *c, double rho, double rt, hipEvent_t *ev) {
        const int64_t n = c->n;
        const int sumPall = c->sumPpair + c->sumPv;
        int ek = 0;
        auto mark = [&]() { if (ev) hipEventRecord(ev[ek++], c->stream); };
        const int64_t nseg = (int64_t)(c->segE.size() + c->segT.size());
        const int nseg_blk = (int)(c->seg_grid > 0 ? std::min<int64_t>((nseg + 3) / 4, c->seg_grid) : (nseg + 3) / 4);
        auto launch_segs = [&](hipStream_t st) {
            SegArgs sa{};
            sa.V = (const uint4 *)c->V.p; sa.rowptr = (const int64_t *)c->rowptr.p; sa.col = (const int32_t *)c->col.p;
            sa.w = c->weighted ? c->w.p : nullptr; sa.rec = c->rec.p; sa.segs = (const Seg *)c->segs.p; sa.nseg = nseg;
            sa.segout = c->segout.p; sa.pt = c->pt.p; sa.rho = rho; sa.rt = rt; sa.tpart = (double *)c->seg_part.p;
            hipLaunchKernelGGL((lb_k_segs<TV, R>), dim3(nseg_blk), dim3(LB_BLOCK), 0, st, sa);
        };
        // the hub segments read only V, the CSR and the triangle records, none of which the kernels before lb_k_rows
        // write: they run on the side stream from the start (per-kernel timing keeps one stream)
        const bool fork = nseg && c->overlap && !ev;
        if (fork) {
            HIPCHK(c, hipEventRecord(c->fork_ev, c->stream));
            HIPCHK(c, hipStreamWaitEvent(c->side, c->fork_ev, 0));
            launch_segs(c->side);
            HIPCHK(c, hipEventRecord(c->join_ev, c->side));
        }
        mark();
        zY(c, ev, &ek);
        mark();
        const double *zp = (const double *)c->Y.p;
        const double *Yp = zp + R;
        // --- pair levels
        if (c->nlev) {
            LevelArgs la{};
            la.nlev = c->nlev; la.R = R;
            int o = 0;
            for (int l = 0; l <= c->nlev; l++) { la.Poff[l] = o; la.soff[l] = c->soff[l]; if (l < c->nlev) o += c->P[l]; }
            la.Y = Yp; la.sig = (const double *)c->sig.p; la.pp = (double *)c->pp.p; la.rq = (const double *)c->rq.p;
            la.dsq = (const double *)c->dsq.p; la.w = c->pw; la.rho = rho;
            la.M = (double *)c->M.p; la.pdiag = (double *)c->pdiag.p; la.tot = (double *)c->levtot.p; la.viol = (double *)c->levviol.p;
            hipLaunchKernelGGL(lb_k_levels<R>, dim3(c->sumPpair), dim3(64), 0, c->stream, la);
        }
        mark();
        // --- VCS multipliers and Z
        const size_t ldsZ = (sizeof(double) + (LB_BLOCK / 64) * sizeof(TC)) * (size_t)c->sumPv * R;
        const int vcs_lds = ldsZ <= 48 * 1024;
        const int vcs_width = 2 + (vcs_lds ? c->sumPv * R : 0);
        int64_t vgrid = LB_VCS_BLOCKS;
        if (c->nvcs) {
            VcsArgs va{};
            va.V = (const uint4 *)c->V.p; va.n = n; va.d = (const double *)c->d.p; va.nvcs = c->nvcs;
            int zo = 0;
            for (int m = 0; m < c->nvcs; m++) {
                va.Yoff[m] = c->sumPpair + zo; va.Zoff[m] = zo; zo += c->vP[m];
                va.lab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
                va.rq[m] = (const double *)c->vrq.p + (va.Zoff[m]);
                va.ptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
            }
            va.Zoff[c->nvcs] = zo;
            va.col = (const int32_t *)c->vcol.p; va.sig = c->vsig.p; va.pv = c->pv.p; va.slack = (uint8_t *)c->vslack.p;
            va.Y = Yp; va.wv = c->wv; va.rho = rho;
            va.vrow = (const int32_t *)c->vrow.p;
            va.eoff[0] = 0;
            for (int m = 0; m < c->nvcs; m++) va.eoff[m + 1] = va.eoff[m] + c->vnnz[m];
            va.sumPv = c->sumPv; va.use_lds = vcs_lds; va.part = (double *)c->vcs_part.p; va.Zglob = (double *)c->Z.p;
            // with the LDS path, the multipliers and the Z sums are two passes (the second has only the adds)
            va.zacc = vcs_lds ? 0 : 1;
            const size_t ldsY = sizeof(double) * (size_t)c->sumPv * R;
            if (!vcs_lds) hipLaunchKernelGGL(lb_k_fill, dim3(256), dim3(256), 0, c->stream, (double *)c->Z.p, (int64_t)c->sumPv * R, 0.0);
            const int NB = (c->sumPv + 15) / 16, MB = (R + 15) / 16;
            // tiles of 32 rows when they fit the LDS budget (twice the work per wavefront between waits), else 16
            const int TRw = (c->vcs_tile == 32 && vcs_mf_lds<TC>(c->sumPv, R, 32) <= 64 * 1024) ? 32 : 16;
            const size_t ldsMf = vcs_mf_lds<TC>(c->sumPv, R, TRw);
            const bool mf = vcs_lds && (c->vcs_kernel == 1 || (c->vcs_kernel < 0 && c->vdtype != LB_BF16)) &&
                            c->sumPv <= 128 && NB * MB <= 8 && ldsMf <= 64 * 1024;
            vgrid = c->vcs_grid > 0 ? c->vcs_grid : (mf ? 512 : LB_VCS_BLOCKS);
            if (mf) {
                const dim3 gv((unsigned)vgrid), bv(64 * LB_VMF_WAVES);
                if (NB * MB <= 4) {
                    if (TRw == 32) hipLaunchKernelGGL((lb_k_vcs_mf4<TV, R, 32>), gv, bv, ldsMf, c->stream, va);
                    else hipLaunchKernelGGL((lb_k_vcs_mf4<TV, R, 16>), gv, bv, ldsMf, c->stream, va);
                } else {
                    if (TRw == 32) hipLaunchKernelGGL((lb_k_vcs_mf8<TV, R, 32>), gv, bv, ldsMf, c->stream, va);
                    else hipLaunchKernelGGL((lb_k_vcs_mf8<TV, R, 16>), gv, bv, ldsMf, c->stream, va);
                }
            } else if (vcs_lds) {
                const size_t ldsYp = sizeof(double) * (size_t)c->sumPv * (R + LB_VPAD);
                hipLaunchKernelGGL((lb_k_vcs_pv<TV, R>), dim3(vgrid), dim3(LB_BLOCK), ldsYp, c->stream, va);
                hipLaunchKernelGGL((lb_k_vcs_acc<TV, R>), dim3(vgrid), dim3(LB_BLOCK), ldsYp, c->stream, va);
            } else {
                hipLaunchKernelGGL((lb_k_vcs_z<TV, R>), dim3(vgrid), dim3(LB_BLOCK), 0, c->stream, va);
            }
        }
        mark();
        if (c->nvcs && vcs_lds) {
            const int width = c->sumPv * R;
            // Z entries sit after the two scalars of each block's partial row
            hipLaunchKernelGGL(lb_k_reduce, dim3((width + 63) / 64), dim3(1024), 0, c->stream,
                               (const double *)c->vcs_part.p + 2, (int)vgrid, width, vcs_width, (double *)c->Z.p);
        }
        mark();
        // --- hub segments (joined here when they ran on the side stream)
        if (fork) HIPCHK(c, hipStreamWaitEvent(c->stream, c->join_ev, 0));
        else if (nseg) launch_segs(c->stream);
        mark();
        // --- rows
        const int64_t nrow_blk = std::min<int64_t>((n + NSL - 1) / NSL, c->row_grid);
        {
            RowArgs ra{};
            ra.V = (const uint4 *)c->V.p; ra.n = n; ra.rowptr = (const int64_t *)c->rowptr.p; ra.col = (const int32_t *)c->col.p;
            ra.w = c->weighted ? c->w.p : nullptr; ra.d = (const double *)c->d.p; ra.L = c->L; ra.z = zp;
            ra.nlev = c->nlev;
            int o = 0;
            for (int l = 0; l < c->nlev; l++) { ra.Poff[l] = o; o += c->P[l]; ra.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n; }
            ra.M = (const double *)c->M.p; ra.pdiag = (const double *)c->pdiag.p;
            ra.nvcs = c->nvcs;
            int zo = 0;
            for (int m = 0; m < c->nvcs; m++) {
                ra.Yoff[m] = c->sumPpair + zo; ra.Zoff[m] = zo; zo += c->vP[m];
                ra.vlab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
                ra.vptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
            }
            ra.vcol = (const int32_t *)c->vcol.p; ra.pv = c->pv.p; ra.Y = Yp; ra.Z = (const double *)c->Z.p;
            ra.iptr = (const int64_t *)c->iptr.p; ra.rec = c->rec.p; ra.pt = c->pt.p; ra.rho = rho; ra.rt = rt;
            ra.heavyE = c->built_heavyE; ra.heavyT = c->built_heavyT;
            ra.hE_rows = (const int32_t *)c->hErows.p; ra.hE_seg0 = (const int32_t *)c->hEseg0.p; ra.nhE = (int)c->hE_rows.size();
            ra.hT_rows = (const int32_t *)c->hTrows.p; ra.hT_seg0 = (const int32_t *)c->hTseg0.p; ra.nhT = (int)c->hT_rows.size();
            ra.segout = c->segout.p; ra.G = c->G.p; ra.part = (double *)c->row_part.p;
            ra.sumPpair = c->sumPpair; ra.sumPv = c->sumPv;
            const size_t stage_bytes = sizeof(TC) * ((size_t)R + (size_t)c->sumPpair * R + c->sumPpair + 2 * (size_t)c->sumPv * R);
            ra.stage = stage_bytes <= 48 * 1024;
            ra.order = c->use_order ? (const int32_t *)c->rorder.p : nullptr;
            if (c->vdtype == LB_BF16)
                hipLaunchKernelGGL((lb_k_rows_w3<TV, R>), dim3((unsigned)nrow_blk), dim3(LB_BLOCK), ra.stage ? stage_bytes : 0, c->stream, ra);
            else
                hipLaunchKernelGGL((lb_k_rows<TV, R>), dim3((unsigned)nrow_blk), dim3(LB_BLOCK), ra.stage ? stage_bytes : 0, c->stream, ra);
        }
        mark();
        // --- scalars
        FinArgs fa{};
        fa.R = R; fa.L = c->L; fa.C0 = c->C0; fa.z = zp; fa.lamsum = (const double *)c->lamsum.p;
        fa.lev_tot = (const double *)c->levtot.p; fa.lev_viol = (const double *)c->levviol.p; fa.sumPpair = c->sumPpair;
        fa.vcs_part = (const double *)c->vcs_part.p; fa.nvcs_blk = c->nvcs ? (int)vgrid : 0; fa.vcs_width = vcs_width;
        fa.row_part = (const double *)c->row_part.p; fa.nrow_blk = (int)nrow_blk;
        fa.seg_part = (const double *)c->seg_part.p; fa.nseg_blk = nseg ? nseg_blk : 0;
        fa.rho = rho; fa.parts = (double *)c->parts.p;
        hipLaunchKernelGGL(lb_k_finalize, dim3(1), dim3(1024), 0, c->stream, fa);
        mark();
        HIPCHK(c, hipGetLastError());
        return 0;
    }

    // one sweep of batched row steps over the uploaded order (c->order) with per-batch segment lists
    static int sweep(lb_ctx *c, double rho, double rt, double damp, double kap2, int nbatch,
                     const std::vector<int64_t> &bpos) {
        const int64_t n = c->n;
        const int sumPall = c->sumPpair + c->sumPv;
        zY(c);
        const double *zp = (const double *)c->Y.p;
        const double *Yp = zp + R;
        int Poff[LB_MAXLEV + 1];
        {
            int o = 0;
            for (int l = 0; l <= c->nlev; l++) { Poff[l] = o; if (l < c->nlev) o += c->P[l]; }
        }
        if (c->nlev) {
            if (rho > 0.0) {
                KqArgs ka{};
                ka.nlev = c->nlev; ka.R = R;
                for (int l = 0; l <= c->nlev; l++) { ka.Poff[l] = Poff[l]; ka.soff[l] = c->soff[l]; }
                ka.Y = Yp; ka.sig = (const double *)c->sig.p; ka.rq = (const double *)c->rq.p; ka.dsq = (const double *)c->dsq.p;
                ka.w = c->pw; ka.rho = rho; ka.Kq = (double *)c->Kq.p;
                hipLaunchKernelGGL(lb_k_kq, dim3(c->sumPpair), dim3(64), 0, c->stream, ka);
            } else {
                hipLaunchKernelGGL(lb_k_fill, dim3(64), dim3(256), 0, c->stream, (double *)c->Kq.p, (int64_t)c->sumPpair, 0.0);
            }
            TmaskArgs ta{};
            ta.order = (const int32_t *)c->order.p; ta.n = n; ta.nbatch = nbatch; ta.nlev = c->nlev;
            for (int l = 0; l < c->nlev; l++) { ta.Poff[l] = Poff[l]; ta.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n; }
            ta.sumP = c->sumPpair; ta.tmask = (uint8_t *)c->tmask.p;
            HIPCHK(c, hipMemsetAsync(c->tmask.p, 0, (size_t)nbatch * c->sumPpair, c->stream));
            hipLaunchKernelGGL(lb_k_tmask, dim3(1024), dim3(256), 0, c->stream, ta);
        }
        VcsStateArgs vs{};
        int Yoff[LB_MAXVCS], Zoff[LB_MAXVCS];
        {
            int zo = 0;
            for (int m = 0; m < c->nvcs; m++) { Yoff[m] = c->sumPpair + zo; Zoff[m] = zo; zo += c->vP[m]; }
        }
        if (c->nvcs) {
            vs.V = (const uint4 *)c->V.p; vs.n = n; vs.d = (const double *)c->d.p; vs.nvcs = c->nvcs;
            for (int m = 0; m < c->nvcs; m++) {
                vs.Yoff[m] = Yoff[m]; vs.Zoff[m] = Zoff[m];
                vs.lab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
                vs.rq[m] = (const double *)c->vrq.p + Zoff[m];
                vs.ptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
            }
            vs.col = (const int32_t *)c->vcol.p; vs.sig = c->vsig.p; vs.Y = Yp; vs.wv = c->wv; vs.rho = rho;
            vs.stZ = (double *)c->stZ.p; vs.Z = (double *)c->Zs.p; vs.cnt = (double *)c->cnts.p;
            const int wz = c->sumPv * R, wt = wz + c->sumPv;
            const size_t ldsS = sizeof(double) * (size_t)wt;
            vs.sumPv = c->sumPv; vs.use_lds = ldsS <= 48 * 1024; vs.part = (double *)c->vst_part.p;
            if (vs.use_lds) {
                hipLaunchKernelGGL((lb_k_vcs_state<TV, R>), dim3(LB_RED_BLOCKS), dim3(LB_BLOCK), ldsS, c->stream, vs);
                hipLaunchKernelGGL(lb_k_reduce, dim3((wz + 63) / 64), dim3(1024), 0, c->stream, (const double *)c->vst_part.p,
                                   LB_RED_BLOCKS, wz, wt, (double *)c->Zs.p);
                hipLaunchKernelGGL(lb_k_reduce, dim3((c->sumPv + 63) / 64), dim3(1024), 0, c->stream,
                                   (const double *)c->vst_part.p + wz, LB_RED_BLOCKS, c->sumPv, wt, (double *)c->cnts.p);
            } else {
                hipLaunchKernelGGL(lb_k_fill, dim3(64), dim3(256), 0, c->stream, (double *)c->Zs.p, (int64_t)c->sumPv * R, 0.0);
                hipLaunchKernelGGL(lb_k_fill, dim3(64), dim3(256), 0, c->stream, (double *)c->cnts.p, (int64_t)c->sumPv, 0.0);
                hipLaunchKernelGGL((lb_k_vcs_state<TV, R>), dim3(LB_RED_BLOCKS), dim3(LB_BLOCK), 0, c->stream, vs);
            }
        }
        StepArgs sa{};
        sa.V = (const uint4 *)c->V.p; sa.rowptr = (const int64_t *)c->rowptr.p; sa.col = (const int32_t *)c->col.p;
        sa.w = c->weighted ? c->w.p : nullptr; sa.d = (const double *)c->d.p; sa.L = c->L; sa.z = zp;
        sa.nlev = c->nlev;
        for (int l = 0; l < c->nlev; l++) { sa.Poff[l] = Poff[l]; sa.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n; }
        sa.M = (const double *)c->M.p; sa.pdiag = (const double *)c->pdiag.p; sa.Kq = (const double *)c->Kq.p;
        sa.nvcs = c->nvcs;
        for (int m = 0; m < c->nvcs; m++) {
            sa.Yoff[m] = Yoff[m]; sa.Zoff[m] = Zoff[m];
            sa.vlab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
            sa.vptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
            sa.vrq[m] = (const double *)c->vrq.p + Zoff[m];
        }
        sa.vcol = (const int32_t *)c->vcol.p; sa.vsig = c->vsig.p; sa.Y = Yp; sa.Z = (const double *)c->Zs.p;
        sa.cnt = (const double *)c->cnts.p; sa.stZ = (const double *)c->stZ.p; sa.wv = c->wv; sa.kap2 = kap2;
        sa.iptr = (const int64_t *)c->iptr.p; sa.rec = c->rec.p; sa.rho = rho; sa.rt = rt; sa.damp = damp;
        const SweepSegs &w = c->sw;
        sa.heavyE = w.heavyE; sa.heavyT = w.heavyT;
        sa.hE_rows = (const int32_t *)w.hErows.p; sa.hE_seg0 = (const int32_t *)w.hEseg0.p; sa.nhE = (int)w.hE_rows.size();
        sa.hT_rows = (const int32_t *)w.hTrows.p; sa.hT_seg0 = (const int32_t *)w.hTseg0.p; sa.nhT = (int)w.hT_rows.size();
        sa.hE_idx = (const int32_t *)w.hEidx.p; sa.hT_idx = (const int32_t *)w.hTidx.p;
        sa.segout = w.segout.p; sa.segkap = (const double *)w.segkap.p;
        sa.Vstage = (uint4 *)c->Vstage.p; sa.DD = (double *)c->DDb.p;
        StepSegArgs ga{};
        ga.V = (const uint4 *)c->V.p; ga.col = (const int32_t *)c->col.p; ga.w = c->weighted ? c->w.p : nullptr;
        ga.rec = c->rec.p; ga.segs = (const Seg *)w.segs.p; ga.rt = rt; ga.segout = w.segout.p; ga.segkap = (double *)w.segkap.p;
        ga.list = (const int32_t *)c->seglist.p; ga.lpos = (const int64_t *)c->lposd.p;
        const int64_t nsegs = (int64_t)(w.segE.size() + w.segT.size());
        // hub segments per batch vary from sweep to sweep: a fixed grid strides over the batch's list
        const unsigned gsegs = (unsigned)std::max<int64_t>(1, std::min<int64_t>(c->sweep_seg_grid > 0 ? c->sweep_seg_grid : 64,
                                                                                (nsegs + 3) / 4));
        PairRowArgs pa{};
        pa.nlev = c->nlev; pa.R = R;
        for (int l = 0; l <= c->nlev; l++) { pa.Poff[l] = Poff[l]; pa.soff[l] = c->soff[l]; }
        for (int l = 0; l < c->nlev; l++) pa.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n;
        pa.Y = Yp; pa.sig = (const double *)c->sig.p; pa.rq = (const double *)c->rq.p; pa.dsq = (const double *)c->dsq.p;
        pa.w = c->pw; pa.rho = rho; pa.M = (double *)c->M.p; pa.pdiag = (double *)c->pdiag.p;
        PostArgs qa{};
        qa.R = R; qa.ls.nsets = c->nlev + c->nvcs;
        {
            int off = 0;
            for (int l = 0; l < c->nlev; l++) { qa.ls.off[l] = off; qa.ls.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n; off += c->P[l]; }
            for (int m = 0; m < c->nvcs; m++) { qa.ls.off[c->nlev + m] = off; qa.ls.lab[c->nlev + m] = (const int32_t *)c->vlabs.p + (int64_t)m * n; off += c->vP[m]; }
            qa.ls.off[qa.ls.nsets] = off;
        }
        qa.sumP = sumPall; qa.DD = (const double *)c->DDb.p; qa.zY = (double *)c->Y.p;
        const size_t ldsP = sizeof(double) * (size_t)(R + sumPall * R);
        qa.use_lds = ldsP <= 48 * 1024;
        Post2Args pb{};
        pb.V = (uint4 *)c->V.p; pb.Vstage = (const uint4 *)c->Vstage.p; pb.d = (const double *)c->d.p; pb.nvcs = c->nvcs;
        for (int m = 0; m < c->nvcs; m++) {
            pb.Yoff[m] = Yoff[m]; pb.Zoff[m] = Zoff[m];
            pb.lab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
            pb.rq[m] = (const double *)c->vrq.p + Zoff[m];
            pb.ptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
        }
        pb.col = (const int32_t *)c->vcol.p; pb.sig = c->vsig.p; pb.Y = Yp; pb.wv = c->wv; pb.rho = rho;
        pb.stZ = (double *)c->stZ.p; pb.Z = (double *)c->Zs.p; pb.cnt = (double *)c->cnts.p;
        pb.sumPv = c->sumPv;
        const size_t ldsV = sizeof(double) * (size_t)c->sumPv * (R + 1);
        pb.use_lds = c->sumPv > 0 && ldsV <= 48 * 1024;
        const int32_t *order = (const int32_t *)c->order.p;
        const bool fused = c->sweep_fused && c->step_wave;
        if (fused) {
            sa.hubcnt = (int32_t *)c->hubcnt.p; sa.hubpos = (int64_t *)c->hubpos.p; sa.hubown = (double *)c->hubown.p;
            sa.zpart = (double *)c->zpart.p;
        }
        // the key of a captured sweep: every argument block before the per-batch fields are set, the batch bounds,
        // the kernel forms and the hub-segment grid
        std::vector<unsigned char> key;
        auto put = [&](const void *p_, size_t nb_) {
            const unsigned char *b_ = (const unsigned char *)p_;
            key.insert(key.end(), b_, b_ + nb_);
        };
        put(&pa, sizeof(pa)); put(&ga, sizeof(ga)); put(&sa, sizeof(sa)); put(&qa, sizeof(qa)); put(&pb, sizeof(pb));
        put(bpos.data(), sizeof(int64_t) * bpos.size());
        const int flags[7] = {c->step_wave ? 1 : 0, (int)gsegs, (int)c->sumPpair, c->nlev, fused ? 1 : 0,
                              c->sweep_pair_segs ? 1 : 0, c->sweep_post_merge ? 1 : 0};
        put(flags, sizeof(flags));
        put(&c->stream, sizeof(c->stream)); put(&c->tmask.p, sizeof(void *)); put(&order, sizeof(order));
        put(&c->Y.p, sizeof(void *)); put(&c->Yalt.p, sizeof(void *)); put(&c->Ysnap.p, sizeof(void *));
        put(&c->DDb.p, sizeof(void *));
        // merged post step: the VCS cluster rows at each batch start copied to Ysnap by the batch's first launch
        const size_t ldsV2 = ldsV + sizeof(double) * (size_t)c->sumPv * R;
        const bool pmerge = c->sweep_post_merge && c->sweep_pair_segs && c->step_wave && c->sumPv > 0 && pb.use_lds &&
                            ldsV2 <= 48 * 1024;
        const int64_t cwv = (int64_t)c->sumPv * R;
        const unsigned ncopy = pmerge ? (unsigned)std::max<int64_t>(1, std::min<int64_t>(8, (cwv + 1023) / 1024)) : 0u;
        Post2Args pbm = pb;
        if (pmerge) {
            pbm.Y = (const double *)c->Ysnap.p;
            for (int m = 0; m < c->nvcs; m++) pbm.Yoff[m] = Zoff[m];
            pbm.dd = (const double *)c->DDb.p;
        }
        const size_t ldsBoth = std::max(qa.use_lds ? ldsP : (size_t)0, ldsV2);
        auto batches = [&](PairRowArgs pa, StepSegArgs ga, StepArgs sa, PostArgs qa, Post2Args pb) {
        for (int b = 0; b < nbatch; b++) {
            const int64_t p0 = bpos[b], nb = bpos[b + 1] - bpos[b];
            if (nb == 0) continue;
            if (c->nlev) pa.tmask = (const uint8_t *)c->tmask.p + (int64_t)b * c->sumPpair;
            ga.b = b;
            const bool pm = pmerge && nb <= LB_POST_MERGE_ROWS;
            if (c->sweep_pair_segs && ((c->nlev && nsegs) || pm)) {
                const int npair = c->nlev ? (c->sumPpair + LB_BLOCK / 64 - 1) / (LB_BLOCK / 64) : 0;
                const int nsegb = nsegs ? (int)gsegs : 0;
                hipLaunchKernelGGL((lb_k_pair_segs<TV, R>), dim3((unsigned)(npair + nsegb) + (pm ? ncopy : 0u)), dim3(LB_BLOCK),
                                   0, c->stream, pa, npair, c->sumPpair, ga, nsegb,
                                   (const double *)(Yp + (int64_t)c->sumPpair * R), pm ? (double *)c->Ysnap.p : nullptr,
                                   pm ? cwv : (int64_t)0);
            } else {
                if (c->nlev) hipLaunchKernelGGL(lb_k_pair_rows<R>, dim3(c->sumPpair), dim3(64), 0, c->stream, pa);
                if (nsegs) hipLaunchKernelGGL((lb_k_step_segs<TV, R>), dim3(gsegs), dim3(LB_BLOCK), 0, c->stream, ga);
            }
            sa.rows = order + p0; sa.nb = nb;
            if (c->step_wave)
                hipLaunchKernelGGL((lb_k_step_wave<TV, R>), dim3((unsigned)((nb + 3) / 4)), dim3(LB_BLOCK), 0, c->stream, sa);
            else
                hipLaunchKernelGGL((lb_k_step<TV, R>), dim3((unsigned)((nb + NSL - 1) / NSL)), dim3(LB_BLOCK), 0, c->stream, sa);
            qa.rows = order + p0; qa.nb = nb;
            const unsigned pg = (unsigned)std::min<int64_t>(64, (nb + LB_POST_ROWS - 1) / LB_POST_ROWS);
            if (pm) {
                Post2Args pbb = pbm;
                pbb.rows = order + p0; pbb.nb = nb;
                hipLaunchKernelGGL((lb_k_post_both<TV, R>), dim3(pg + (unsigned)((nb + 3) / 4)), dim3(LB_BLOCK), ldsBoth,
                                   c->stream, qa, (int)pg, pbb);
                continue;
            }
            hipLaunchKernelGGL(lb_k_post_acc, dim3(pg), dim3(LB_BLOCK), qa.use_lds ? ldsP : 0, c->stream, qa);
            pb.rows = order + p0; pb.nb = nb;
            if (c->step_wave)
                hipLaunchKernelGGL((lb_k_post_vcs_wave<TV, R>), dim3((unsigned)((nb + 3) / 4)), dim3(LB_BLOCK),
                                   pb.use_lds ? ldsV : 0, c->stream, pb);
            else
                hipLaunchKernelGGL((lb_k_post_vcs<TV, R>), dim3((unsigned)((nb + NSL - 1) / NSL)), dim3(LB_BLOCK),
                                   pb.use_lds ? ldsV : 0, c->stream, pb);
        }
        };
        // fused: per batch b, lb_k_sweep_pre (the previous batch's post-step, this batch's pair rows, [z | Y] copied
        // to the buffer this batch adds into), then lb_k_step_fused; the batches' [z | Y] alternate between Y and Yalt
        auto batches_fused = [&](PairRowArgs pa, StepSegArgs ga, StepArgs sa, Post2Args pb) {
            const int64_t cw = R + (int64_t)sumPall * R;
            double *zb[2] = {(double *)c->Y.p, (double *)c->Yalt.p};
            int cur = 0;
            int64_t pprev = 0, nbprev = 0;
            const unsigned ncopy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(16, (cw + 4 * LB_BLOCK - 1) / (4 * LB_BLOCK)));
            const int npair = c->nlev ? (c->sumPpair + LB_BLOCK / 64 - 1) / (LB_BLOCK / 64) : 0;
            const size_t ldsPre = pb.use_lds ? ldsV : 0;
            auto pre = [&](int npost, int np, const uint8_t *tm, double *dst) {
                pb.rows = order + pprev; pb.nb = nbprev; pb.Y = zb[cur] + R;
                pa.Y = zb[cur] + R; pa.tmask = tm;
                hipLaunchKernelGGL((lb_k_sweep_pre<TV, R>), dim3((unsigned)(npost + np) + ncopy), dim3(LB_BLOCK), ldsPre,
                                   c->stream, pb, npost, pa, np, c->sumPpair, zb[cur], dst, cw, (double *)c->zpart.p);
            };
            for (int b = 0; b < nbatch; b++) {
                const int64_t p0 = bpos[b], nb = bpos[b + 1] - bpos[b];
                if (nb == 0) continue;
                pre(nbprev ? (int)((nbprev + 3) / 4) : 0, npair,
                    c->nlev ? (const uint8_t *)c->tmask.p + (int64_t)b * c->sumPpair : nullptr, zb[cur ^ 1]);
                sa.z = zb[cur]; sa.Y = zb[cur] + R; sa.zYn = zb[cur ^ 1]; sa.rows = order + p0; sa.nb = nb;
                ga.b = b;
                const int nrb = (int)((nb + 3) / 4);
                hipLaunchKernelGGL((lb_k_step_fused<TV, R>), dim3((unsigned)nrb + (nsegs ? gsegs : 0u)), dim3(LB_BLOCK), 0,
                                   c->stream, sa, ga, nrb);
                pprev = p0; nbprev = nb; cur ^= 1;
            }
            // the last batch's post-step, its z completed, and its [z | Y] back into Y
            if (nbprev) pre((int)((nbprev + 3) / 4), 0, nullptr, cur ? zb[0] : nullptr);
        };
        // two launches per batch (lb_k_step_copy, lb_k_post_next): eager sweeps only (the batch tags are host values)
        const size_t ldsPair = sizeof(double) * (size_t)c->sumPpair * R;
        const bool two = !fused && c->sweep_two && c->step_wave && c->sweep_pair_segs && c->sweep_post_merge &&
                         !c->sweep_graph && (c->sumPv == 0 || (pb.use_lds && ldsV2 <= 48 * 1024)) &&
                         ldsPair <= 48 * 1024;
        auto batches_two = [&](PairRowArgs pa, StepSegArgs ga, StepArgs sa, PostArgs qa, Post2Args pb) {
            std::vector<int> nz;
            for (int b = 0; b < nbatch; b++) if (bpos[b + 1] > bpos[b]) nz.push_back(b);
            if (nz.empty()) return;
            const int npair = c->nlev ? (c->sumPpair + LB_BLOCK / 64 - 1) / (LB_BLOCK / 64) : 0;
            const int nsegb = nsegs ? (int)gsegs : 0;
            const int64_t cw = (int64_t)(c->sumPpair + c->sumPv) * R;
            const unsigned ncp = (unsigned)std::max<int64_t>(1, std::min<int64_t>(8, (cw + 1023) / 1024));
            const size_t ldsNext = std::max(std::max(qa.use_lds ? ldsP : (size_t)0, c->sumPv ? ldsV2 : (size_t)0),
                                            c->nlev ? ldsPair : (size_t)0);
            double *ysnap = (double *)c->Ysnap.p;
            // the first batch's pair rows and hub segments, from the aggregates and V as they are
            if (npair + nsegb > 0) {
                if (c->nlev) pa.tmask = (const uint8_t *)c->tmask.p + (int64_t)nz[0] * c->sumPpair;
                ga.b = nz[0];
                hipLaunchKernelGGL((lb_k_pair_segs<TV, R>), dim3((unsigned)(npair + nsegb)), dim3(LB_BLOCK), 0, c->stream,
                                   pa, npair, c->sumPpair, ga, nsegb, (const double *)nullptr, (double *)nullptr, (int64_t)0);
            }
            // per-batch delta buffers for batches above 128 rows (every workgroup of the next launch would otherwise
            // sum the whole batch in LDS: 150k-vertex request, 282 rows, 0.330 against 0.352 s of sweeps; 40k, 76
            // rows, 0.257 against 0.250 s); one choice per call, as the buffers are cleared batch to batch
            const bool dl = c->sweep_delta && n / nbatch > 128;
            const int64_t wY = (int64_t)(c->sumPpair + c->sumPv) * R, wd = (int64_t)LB_ZPART * R + wY;
            const unsigned nap = dl ? (unsigned)std::max<int64_t>(1, std::min<int64_t>(16, (wd + 1023) / 1024)) : 0u;
            for (size_t k = 0; k < nz.size(); k++) {
                const int b = nz[k];
                const int64_t p0 = bpos[b], nb = bpos[b + 1] - bpos[b];
                const int tag = c->sw_tag++;
                double *Dcur = (double *)c->dbuf.p + (int64_t)(tag & 1) * wd;
                double *Doth = (double *)c->dbuf.p + (int64_t)((tag + 1) & 1) * wd;
                StepArgs st = sa;
                st.rows = order + p0; st.nb = nb;
                st.posb = (int32_t *)c->posb.p; st.posp = (int32_t *)c->posp.p; st.tag = tag;
                if (dl) st.dnext = Dcur;
                const int nrb = (int)((nb + 3) / 4);
                hipLaunchKernelGGL((lb_k_step_copy<TV, R>), dim3((unsigned)nrb + ncp), dim3(LB_BLOCK), 0, c->stream, st, nrb,
                                   Yp, ysnap, cw);
                PostArgs qb = qa;
                qb.rows = order + p0; qb.nb = nb;
                const unsigned pg = (unsigned)std::min<int64_t>(64, (nb + LB_POST_ROWS - 1) / LB_POST_ROWS);
                Post2Args pv = pb;
                pv.rows = order + p0; pv.nb = nb;
                if (c->sumPv) {
                    pv.Y = ysnap + (int64_t)c->sumPpair * R;
                    for (int m = 0; m < c->nvcs; m++) pv.Yoff[m] = Zoff[m];
                    if (dl) pv.dYg = Dcur + (int64_t)LB_ZPART * R + (int64_t)c->sumPpair * R;
                    else pv.dd = (const double *)c->DDb.p;
                }
                const bool next = k + 1 < nz.size();
                PairRowArgs pn = pa;
                StepSegArgs gn = ga;
                if (next) {
                    if (c->nlev) pn.tmask = (const uint8_t *)c->tmask.p + (int64_t)nz[k + 1] * c->sumPpair;
                    pn.Y = ysnap;
                    if (dl) pn.dPg = Dcur + (int64_t)LB_ZPART * R;
                    else { pn.dd = (const double *)c->DDb.p; pn.drows = order + p0; pn.dnb = nb; }
                    gn.b = nz[k + 1];
                    gn.posb = (const int32_t *)c->posb.p; gn.posp = (const int32_t *)c->posp.p;
                    gn.Vstage = (const uint4 *)c->Vstage.p; gn.tag = tag;
                }
                const int np2 = next ? npair : 0, ns2 = next ? nsegb : 0;
                ApplyArgs ap{};
                ap.zY = (double *)c->Y.p; ap.D = Dcur; ap.Dz = Doth; ap.wY = wY;
                const unsigned npa = dl ? 0u : pg;
                hipLaunchKernelGGL((lb_k_post_next<TV, R>), dim3(npa + nap + (unsigned)nrb + (unsigned)(np2 + ns2)),
                                   dim3(LB_BLOCK), dl ? ((c->sumPv && pb.use_lds) ? ldsV : (size_t)0) : ldsNext, c->stream, qb, (int)npa,
                                   pv, nrb, pn, np2, c->sumPpair, gn, ns2, ap, (int)nap);
            }
        };
        auto run = [&]() {
            if (fused) batches_fused(pa, ga, sa, pb);
            else if (two) batches_two(pa, ga, sa, qa, pb);
            else batches(pa, ga, sa, qa, pb);
        };
        if (!c->sweep_graph) {
            run();
            HIPCHK(c, hipGetLastError());
            return 0;
        }
        if (!c->sw_exec || key != c->sw_key) {
            if (c->sw_exec) { hipGraphExecDestroy((hipGraphExec_t)c->sw_exec); c->sw_exec = nullptr; }
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            run();
            hipGraph_t g = nullptr;
            const hipError_t le = hipGetLastError();
            hipError_t e = hipStreamEndCapture(c->stream, &g);
            if (le != hipSuccess) { if (g) hipGraphDestroy(g); return fail(c, -3, std::string("sweep capture: ") + hipGetErrorString(le)); }
            if (e != hipSuccess) return fail(c, -3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
            hipGraphExec_t ge = nullptr;
            e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            hipGraphDestroy(g);
            if (e != hipSuccess) return fail(c, -3, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
            c->sw_exec = ge;
            c->sw_key.swap(key);
        }
        HIPCHK(c, hipGraphLaunch((hipGraphExec_t)c->sw_exec, c->stream));
        return 0;
    }

    // triangle separation of one family (fam 0 wedges, fam 1 one-edge arms) on the resident V; fills out (m x 3)
    // and val (m), returns the count through *m_out (m <= cap)
    static int separate(lb_ctx *c, int fam, double thresh, int64_t maxh, int64_t maxk, int64_t hub2, int64_t cap2,
                        double vtol, int64_t cap, int nchunk, int nb, std::vector<int64_t> &out, std::vector<double> &val,
                        double *vmin_out) {
        const int64_t n = c->n, nnz = c->nnz;
        int rc;
        DBuf xe, flag, apos, arm_id, arm_x, mcnt, sx, sid, segb, sege, tmp, kid, kx, kcnt, hist, ccount, cpos,
            chunk_cnt, chunk_off, dout, dval, scal;
        auto cleanup = [&]() {
            DBuf *bs[] = {&xe, &flag, &apos, &arm_id, &arm_x, &mcnt, &sx, &sid, &segb, &sege, &tmp, &kid, &kx, &kcnt,
                          &hist, &ccount, &cpos, &chunk_cnt, &chunk_off, &dout, &dval, &scal};
            for (DBuf *b : bs) dfree(*b);
        };
#define SEPCHK(x) do { if ((rc = (x))) { cleanup(); return rc; } } while (0)
#define SEPHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return fail(c, -3, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
        SEPCHK(dalloc(c, xe, sizeof(double) * nnz));
        SEPCHK(dalloc(c, flag, sizeof(int64_t) * (nnz + 1)));
        SEPCHK(dalloc(c, apos, sizeof(int64_t) * (nnz + 1)));
        SEPCHK(dalloc(c, scal, 64));
        SepParams p{};
        p.V = (const uint4 *)c->V.p; p.n = n; p.rowptr = (const int64_t *)c->rowptr.p; p.col = (const int32_t *)c->col.p;
        p.thresh = thresh; p.vtol = vtol; p.maxh = maxh; p.maxk = maxk; p.hub2 = hub2; p.cap2 = cap2;
        p.nchunk = nchunk; p.nb = nb; p.cs = (n + nchunk - 1) / nchunk;
        // per-edge arm values (int64 flags: the scan output is the arm position)
        SEPHIP(hipMemsetAsync(flag.p, 0, sizeof(int64_t) * (nnz + 1), c->stream));
        {
            // the edge kernel writes int32 flags into the first half; widen them below
            hipLaunchKernelGGL((lb_k_sep_edges<TV, R>), dim3(2048), dim3(256), 0, c->stream, p, (double *)xe.p, (int32_t *)apos.p);
            hipLaunchKernelGGL(lb_k_widen, dim3(2048), dim3(256), 0, c->stream, (const int32_t *)apos.p, nnz, (int64_t *)flag.p);
        }
        size_t tb = 0;
        SEPHIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (const int64_t *)flag.p, (int64_t *)apos.p, (int)(nnz + 1), c->stream));
        SEPCHK(dalloc(c, tmp, tb));
        SEPHIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, (const int64_t *)flag.p, (int64_t *)apos.p, (int)(nnz + 1), c->stream));
        int64_t narms = 0;
        SEPHIP(hipMemcpyAsync(&narms, (const int64_t *)apos.p + nnz, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipStreamSynchronize(c->stream));
        SEPCHK(dalloc(c, arm_id, sizeof(int32_t) * std::max<int64_t>(narms, 1)));
        SEPCHK(dalloc(c, arm_x, sizeof(double) * std::max<int64_t>(narms, 1)));
        hipLaunchKernelGGL(lb_k_sep_compact, dim3(2048), dim3(256), 0, c->stream, (const int64_t *)c->rowptr.p, n,
                           (const int32_t *)c->col.p, (const double *)xe.p, (const int64_t *)flag.p, (const int64_t *)apos.p,
                           (int32_t *)arm_id.p, (double *)arm_x.p);
        // hub centres with more than maxh arms: stable value-descending sort of their arm segments, keep maxh
        SEPCHK(dalloc(c, mcnt, sizeof(int64_t) * (n + 1)));
        hipLaunchKernelGGL(lb_k_gather_pos, dim3(1024), dim3(256), 0, c->stream, (const int64_t *)c->rowptr.p, n,
                           (const int64_t *)apos.p, (int64_t *)mcnt.p);
        std::vector<int64_t> ab(n + 1);
        SEPHIP(hipMemcpyAsync(ab.data(), mcnt.p, sizeof(int64_t) * (n + 1), hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipStreamSynchronize(c->stream));
        std::vector<int64_t> hb, he;
        for (int64_t j = 0; j < n; j++)
            if (ab[j + 1] - ab[j] > maxh) { hb.push_back(ab[j]); he.push_back(ab[j + 1]); }
        if (!hb.empty()) {
            SEPCHK(upload(c, segb, hb.data(), sizeof(int64_t) * hb.size()));
            SEPCHK(upload(c, sege, he.data(), sizeof(int64_t) * he.size()));
            SEPCHK(dalloc(c, sx, sizeof(double) * narms));
            SEPCHK(dalloc(c, sid, sizeof(int32_t) * narms));
            size_t tb2 = 0;
            SEPHIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(nullptr, tb2, (const double *)arm_x.p, (double *)sx.p,
                   (const int32_t *)arm_id.p, (int32_t *)sid.p, (int)narms, (int)hb.size(), (const int64_t *)segb.p,
                   (const int64_t *)sege.p, 0, 64, c->stream));
            SEPCHK(dalloc(c, tmp, tb2));
            SEPHIP(hipcub::DeviceSegmentedRadixSort::SortPairsDescending(tmp.p, tb2, (const double *)arm_x.p, (double *)sx.p,
                   (const int32_t *)arm_id.p, (int32_t *)sid.p, (int)narms, (int)hb.size(), (const int64_t *)segb.p,
                   (const int64_t *)sege.p, 0, 64, c->stream));
            hipLaunchKernelGGL(lb_k_sep_cutcopy, dim3(256), dim3(256), 0, c->stream, (const int64_t *)segb.p, (int)hb.size(),
                               (const int32_t *)sid.p, (const double *)sx.p, maxh, (int32_t *)arm_id.p, (double *)arm_x.p);
        }
        p.apos = (const int64_t *)apos.p; p.arm_id = (const int32_t *)arm_id.p; p.arm_x = (const double *)arm_x.p;
        if (fam == 1) {
            SEPCHK(dalloc(c, kid, sizeof(int32_t) * (size_t)n * maxk));
            SEPCHK(dalloc(c, kx, sizeof(double) * (size_t)n * maxk));
            SEPCHK(dalloc(c, kcnt, sizeof(int32_t) * n));
            hipLaunchKernelGGL((lb_k_sep_k<TV, R>), dim3((unsigned)n), dim3(64), 0, c->stream, p, (int32_t *)kid.p,
                               (double *)kx.p, (int32_t *)kcnt.p);
            p.kid = (const int32_t *)kid.p; p.kx = (const double *)kx.p; p.kcnt = (const int32_t *)kcnt.p;
        }
        SEPCHK(dalloc(c, hist, sizeof(unsigned long long) * (size_t)nchunk * nb));
        SEPHIP(hipMemsetAsync(hist.p, 0, sizeof(unsigned long long) * (size_t)nchunk * nb, c->stream));
        p.hist = (unsigned long long *)hist.p;
        const unsigned g = (unsigned)((n + 3) / 4);
        if (fam == 0) hipLaunchKernelGGL((lb_k_sep_wedges<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 0);
        else hipLaunchKernelGGL((lb_k_sep_one_edge<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 0);
        SEPCHK(dalloc(c, chunk_cnt, sizeof(int64_t) * nchunk));
        SEPCHK(dalloc(c, chunk_off, sizeof(int64_t) * nchunk));
        double *dvmin = (double *)scal.p;
        int64_t *dtotal = (int64_t *)((char *)scal.p + 8);
        int *db0 = (int *)((char *)scal.p + 16);
        const size_t lds_t = nb <= LB_THRESH_MAXNB ? sizeof(unsigned long long) * (size_t)nb : 0;
        hipLaunchKernelGGL(lb_k_sep_thresh, dim3(1), dim3(LB_BLOCK), lds_t, c->stream, (const unsigned long long *)hist.p,
                           nchunk, nb, vtol, cap, dvmin, (int64_t *)chunk_cnt.p, (int64_t *)chunk_off.p, dtotal, db0);
        double vmin = -1.0;
        int64_t total = 0;
        SEPHIP(hipMemcpyAsync(&vmin, dvmin, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipMemcpyAsync(&total, dtotal, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipStreamSynchronize(c->stream));
        *vmin_out = vmin;
        out.clear(); val.clear();
        if (vmin < 0.0 || total == 0) { cleanup(); return 0; }
        p.vminp = dvmin; p.chunk_cnt = (const int64_t *)chunk_cnt.p; p.chunk_off = (const int64_t *)chunk_off.p;
        SEPCHK(dalloc(c, ccount, sizeof(int64_t) * (n + 1)));
        SEPCHK(dalloc(c, cpos, sizeof(int64_t) * (n + 1)));
        SEPHIP(hipMemsetAsync(ccount.p, 0, sizeof(int64_t) * (n + 1), c->stream));
        SEPCHK(dalloc(c, mcnt, sizeof(int64_t) * (n + 1)));
        p.ccount = (int32_t *)mcnt.p;
        SEPHIP(hipMemsetAsync(mcnt.p, 0, sizeof(int64_t) * (n + 1), c->stream));
        if (fam == 0) hipLaunchKernelGGL((lb_k_sep_wedges<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 1);
        else hipLaunchKernelGGL((lb_k_sep_one_edge<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 1);
        hipLaunchKernelGGL(lb_k_widen, dim3(1024), dim3(256), 0, c->stream, (const int32_t *)mcnt.p, n, (int64_t *)ccount.p);
        size_t tb3 = 0;
        SEPHIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, (const int64_t *)ccount.p, (int64_t *)cpos.p, (int)(n + 1), c->stream));
        SEPCHK(dalloc(c, tmp, tb3));
        SEPHIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb3, (const int64_t *)ccount.p, (int64_t *)cpos.p, (int)(n + 1), c->stream));
        p.cpos = (const int64_t *)cpos.p;
        SEPCHK(dalloc(c, dout, sizeof(int64_t) * 3 * total));
        SEPCHK(dalloc(c, dval, sizeof(double) * total));
        SEPHIP(hipMemsetAsync(dout.p, 0, sizeof(int64_t) * 3 * total, c->stream));
        SEPHIP(hipMemsetAsync(dval.p, 0, sizeof(double) * total, c->stream));
        p.out = (int64_t *)dout.p; p.val = (double *)dval.p;
        if (fam == 0) hipLaunchKernelGGL((lb_k_sep_wedges<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 2);
        else hipLaunchKernelGGL((lb_k_sep_one_edge<TV, R>), dim3(g), dim3(LB_BLOCK), 0, c->stream, p, 2);
        SEPHIP(hipGetLastError());
        std::vector<int64_t> o3(3 * total);
        std::vector<double> v1(total);
        SEPHIP(hipMemcpyAsync(o3.data(), dout.p, sizeof(int64_t) * 3 * total, hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipMemcpyAsync(v1.data(), dval.p, sizeof(double) * total, hipMemcpyDeviceToHost, c->stream));
        SEPHIP(hipStreamSynchronize(c->stream));
        for (int64_t t = 0; t < total; t++)
            if (v1[t] > 0.0) { out.push_back(o3[3 * t]); out.push_back(o3[3 * t + 1]); out.push_back(o3[3 * t + 2]); val.push_back(v1[t]); }
        cleanup();
#undef SEPCHK
#undef SEPHIP
        return 0;
    }

    // VCS separation of level m: the nadd best candidates (score descending, ties in (i, R) order) and statistics
    static int vcs_separate(lb_ctx *c, int m, double tol, int mode, int64_t nadd, std::vector<int64_t> &oi,
                            std::vector<int32_t> &oR, std::vector<double> &os, double *stats) {
        const int64_t n = c->n;
        int rc;
        zY(c);
        DBuf cnt, base, ci, cR, cs, st, tmp, idx, idx2, sk, ti, tR, ts;
        auto cleanup = [&]() { DBuf *bs[] = {&cnt, &base, &ci, &cR, &cs, &st, &tmp, &idx, &idx2, &sk, &ti, &tR, &ts}; for (DBuf *b : bs) dfree(*b); };
#define VCHK(x) do { if ((rc = (x))) { cleanup(); return rc; } } while (0)
#define VHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return fail(c, -3, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
        int zo = 0;
        for (int k = 0; k < m; k++) zo += c->vP[k];
        const int P = c->vP[m];
        VcsSepArgs a{};
        a.V = (const uint4 *)c->V.p; a.n = n; a.d = (const double *)c->d.p; a.lab = (const int32_t *)c->vlabs.p + (int64_t)m * n;
        a.ptr = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1); a.col = (const int32_t *)c->vcol.p;
        a.Y = (const double *)c->Y.p + R + (int64_t)(c->sumPpair + zo) * R; a.vol = (const double *)c->vvol.p + zo;
        a.rq = (const double *)c->vrq.p + zo; a.P = P; a.wv = c->wv; a.tol = tol; a.mode = mode;
        const size_t lds = sizeof(double) * (size_t)P * R;
        a.use_lds = lds <= 48 * 1024;
        const unsigned grid = (unsigned)std::min<int64_t>((n + LB_BLOCK - 1) / LB_BLOCK, 4096);
        VCHK(dalloc(c, cnt, sizeof(int64_t) * (n + 1)));
        VCHK(dalloc(c, base, sizeof(int64_t) * (n + 1)));
        VCHK(dalloc(c, st, sizeof(double) * 4 * grid));
        VHIP(hipMemsetAsync(cnt.p, 0, sizeof(int64_t) * (n + 1), c->stream));
        a.cnt = (int64_t *)cnt.p; a.stats = (double *)st.p;
        hipLaunchKernelGGL((lb_k_vcs_sep<TV, R>), dim3(grid), dim3(LB_BLOCK), a.use_lds ? lds : 0, c->stream, a, 0);
        size_t tb = 0;
        VHIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (const int64_t *)cnt.p, (int64_t *)base.p, (int)(n + 1), c->stream));
        VCHK(dalloc(c, tmp, tb));
        VHIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, (const int64_t *)cnt.p, (int64_t *)base.p, (int)(n + 1), c->stream));
        int64_t total = 0;
        std::vector<double> sth(4 * grid);
        VHIP(hipMemcpyAsync(&total, (const int64_t *)base.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
        VHIP(hipMemcpyAsync(sth.data(), st.p, sizeof(double) * 4 * grid, hipMemcpyDeviceToHost, c->stream));
        VHIP(hipStreamSynchronize(c->stream));
        for (int k = 0; k < 4; k++) { double t = 0.0; for (unsigned b = 0; b < grid; b++) t += sth[4 * b + k]; stats[k] = t; }
        oi.clear(); oR.clear(); os.clear();
        if (total == 0 || nadd <= 0) { cleanup(); return 0; }
        VCHK(dalloc(c, ci, sizeof(int32_t) * total));
        VCHK(dalloc(c, cR, sizeof(int32_t) * total));
        VCHK(dalloc(c, cs, sizeof(double) * total));
        a.base = (const int64_t *)base.p; a.oi = (int32_t *)ci.p; a.oR = (int32_t *)cR.p; a.os = (double *)cs.p;
        hipLaunchKernelGGL((lb_k_vcs_sep<TV, R>), dim3(grid), dim3(LB_BLOCK), a.use_lds ? lds : 0, c->stream, a, 1);
        const int64_t keep = std::min<int64_t>(nadd, total);
        const int32_t *pi = (const int32_t *)ci.p, *pR = (const int32_t *)cR.p;
        const double *ps = (const double *)cs.p;
        if (total > nadd) {   // stable descending sort of the scores: ties stay in (i, R) order
            VCHK(dalloc(c, idx, sizeof(int64_t) * total));
            VCHK(dalloc(c, idx2, sizeof(int64_t) * total));
            VCHK(dalloc(c, sk, sizeof(double) * total));
            hipLaunchKernelGGL(lb_k_iota64, dim3(1024), dim3(256), 0, c->stream, (int64_t *)idx.p, total);
            size_t tb2 = 0;
            VHIP(hipcub::DeviceRadixSort::SortPairsDescending(nullptr, tb2, ps, (double *)sk.p, (const int64_t *)idx.p,
                                                              (int64_t *)idx2.p, (int)total, 0, 64, c->stream));
            VCHK(dalloc(c, tmp, tb2));
            VHIP(hipcub::DeviceRadixSort::SortPairsDescending(tmp.p, tb2, ps, (double *)sk.p, (const int64_t *)idx.p,
                                                              (int64_t *)idx2.p, (int)total, 0, 64, c->stream));
            VCHK(dalloc(c, ti, sizeof(int32_t) * keep));
            VCHK(dalloc(c, tR, sizeof(int32_t) * keep));
            VCHK(dalloc(c, ts, sizeof(double) * keep));
            hipLaunchKernelGGL(lb_k_take, dim3(1024), dim3(256), 0, c->stream, (const int64_t *)idx2.p, keep, pi, pR, ps,
                               (int32_t *)ti.p, (int32_t *)tR.p, (double *)ts.p);
            pi = (const int32_t *)ti.p; pR = (const int32_t *)tR.p; ps = (const double *)ts.p;
        }
        std::vector<int32_t> hi(keep);
        oR.resize(keep); os.resize(keep);
        VHIP(hipMemcpyAsync(hi.data(), pi, sizeof(int32_t) * keep, hipMemcpyDeviceToHost, c->stream));
        VHIP(hipMemcpyAsync(oR.data(), pR, sizeof(int32_t) * keep, hipMemcpyDeviceToHost, c->stream));
        VHIP(hipMemcpyAsync(os.data(), ps, sizeof(double) * keep, hipMemcpyDeviceToHost, c->stream));
        VHIP(hipGetLastError());
        VHIP(hipStreamSynchronize(c->stream));
        oi.assign(hi.begin(), hi.end());
        cleanup();
#undef VCHK
#undef VHIP
        return 0;
    }
};

typedef int (*eval_fn)(lb_ctx *, double, double, hipEvent_t *);
typedef int (*sep_fn)(lb_ctx *, int, double, int64_t, int64_t, int64_t, int64_t, double, int64_t, int, int,
                      std::vector<int64_t> &, std::vector<double> &, double *);
typedef int (*vsep_fn)(lb_ctx *, int, double, int, int64_t, std::vector<int64_t> &, std::vector<int32_t> &,
                       std::vector<double> &, double *);
typedef int (*sweep_fn)(lb_ctx *, double, double, double, double, int, const std::vector<int64_t> &);

eval_fn pick_eval(int vdtype, int r) {
#define LB_PICK(TV)                                     \
    if (r == 8) return &Launch<TV, 8>::eval;            \
    if (r == 16) return &Launch<TV, 16>::eval;          \
    if (r == 32) return &Launch<TV, 32>::eval;
    if (vdtype == LB_F32) { LB_PICK(float) }
    if (vdtype == LB_BF16) { LB_PICK(bf16raw) }
    if (vdtype == LB_F64) { LB_PICK(double) }
#undef LB_PICK
    return nullptr;
}

sweep_fn pick_sweep(int vdtype, int r) {
#define LB_PICK(TV)                                     \
    if (r == 8) return &Launch<TV, 8>::sweep;           \
    if (r == 16) return &Launch<TV, 16>::sweep;         \
    if (r == 32) return &Launch<TV, 32>::sweep;
    if (vdtype == LB_F32) { LB_PICK(float) }
    if (vdtype == LB_BF16) { LB_PICK(bf16raw) }
    if (vdtype == LB_F64) { LB_PICK(double) }
#undef LB_PICK
    return nullptr;
}

sep_fn pick_sep(int vdtype, int r) {
#define LB_PICK(TV)                                     \
    if (r == 8) return &Launch<TV, 8>::separate;        \
    if (r == 16) return &Launch<TV, 16>::separate;      \
    if (r == 32) return &Launch<TV, 32>::separate;
    if (vdtype == LB_F32) { LB_PICK(float) }
    if (vdtype == LB_BF16) { LB_PICK(bf16raw) }
    if (vdtype == LB_F64) { LB_PICK(double) }
#undef LB_PICK
    return nullptr;
}

vsep_fn pick_vsep(int vdtype, int r) {
#define LB_PICK(TV)                                     \
    if (r == 8) return &Launch<TV, 8>::vcs_separate;    \
    if (r == 16) return &Launch<TV, 16>::vcs_separate;  \
    if (r == 32) return &Launch<TV, 32>::vcs_separate;
    if (vdtype == LB_F32) { LB_PICK(float) }
    if (vdtype == LB_BF16) { LB_PICK(bf16raw) }
    if (vdtype == LB_F64) { LB_PICK(double) }
#undef LB_PICK
    return nullptr;
}

int ensure_eval_buffers(lb_ctx *c) {
    const int R = c->r;
    const int sumPall = c->sumPpair + c->sumPv;
    int rc;
    if ((rc = dalloc(c, c->Y, sizeof(double) * (size_t)(R + sumPall * R)))) return rc;
    const bool ylds = sizeof(double) * (size_t)sumPall * (R + LB_VPAD) <= 48 * 1024;   // as lb_k_colsum's launch
    if ((rc = dalloc(c, c->colsum_part, sizeof(double) * (size_t)LB_RED_BLOCKS * (R + (ylds ? sumPall * R : 0))))) return rc;
    if ((rc = dalloc(c, c->M, sizeof(double) * (size_t)std::max(c->sumPpair, 1) * R))) return rc;
    if ((rc = dalloc(c, c->pdiag, sizeof(double) * std::max(c->sumPpair, 1)))) return rc;
    if ((rc = dalloc(c, c->levtot, sizeof(double) * std::max(c->sumPpair, 1)))) return rc;
    if ((rc = dalloc(c, c->levviol, sizeof(double) * std::max(c->sumPpair, 1)))) return rc;
    if ((rc = dalloc(c, c->Z, sizeof(double) * (size_t)std::max(c->sumPv, 1) * R))) return rc;
    const bool zlds = (sizeof(double) + (LB_BLOCK / 64) * c->cbytes) * (size_t)c->sumPv * R <= 48 * 1024;
    if ((rc = dalloc(c, c->vcs_part, sizeof(double) * (size_t)LB_VCS_BLOCKS * (2 + (zlds ? c->sumPv * R : 0))))) return rc;
    const int S = (int)(R * c->vbytes / 16), NSL = LB_BLOCK / S;
    const int64_t nrow_blk = (c->n + NSL - 1) / NSL;
    if ((rc = dalloc(c, c->row_part, sizeof(double) * 4 * (size_t)std::max<int64_t>(nrow_blk, 1)))) return rc;
    if ((rc = dalloc(c, c->parts, sizeof(double) * LB_NPARTS))) return rc;
    if (!c->lamsum.p) {
        if ((rc = dalloc(c, c->lamsum, sizeof(double)))) return rc;
        HIPCHK(c, hipMemsetAsync(c->lamsum.p, 0, sizeof(double), c->stream));
    }
    return 0;
}

int refresh_lamsum(lb_ctx *c) {
    int rc;
    if ((rc = dalloc(c, c->lamsum, sizeof(double)))) return rc;
    if (c->T == 0) {
        HIPCHK(c, hipMemsetAsync(c->lamsum.p, 0, sizeof(double), c->stream));
        return 0;
    }
    if ((rc = dalloc(c, c->colsum_part, sizeof(double) * LB_RED_BLOCKS))) return rc;
    if (c->cbytes == 8) hipLaunchKernelGGL(lb_k_sum<double>, dim3(LB_RED_BLOCKS), dim3(LB_BLOCK), 0, c->stream, (const double *)c->lam.p, c->T, (double *)c->colsum_part.p);
    else hipLaunchKernelGGL(lb_k_sum<float>, dim3(LB_RED_BLOCKS), dim3(LB_BLOCK), 0, c->stream, (const float *)c->lam.p, c->T, (double *)c->colsum_part.p);
    hipLaunchKernelGGL(lb_k_reduce, dim3(1), dim3(1024), 0, c->stream, (const double *)c->colsum_part.p, LB_RED_BLOCKS, 1, 1, (double *)c->lamsum.p);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int empty_triangles(lb_ctx *c) {
    // zero-triangle pool: incidence pointer of zeros so the row kernels see no records
    std::vector<int64_t> z(c->n + 1, 0);
    int rc;
    if ((rc = upload(c, c->iptr, z.data(), sizeof(int64_t) * (c->n + 1)))) return rc;
    if ((rc = dalloc(c, c->rec, 32))) return rc;
    if ((rc = dalloc(c, c->pt, 16))) return rc;
    if ((rc = dalloc(c, c->lam, 16))) return rc;
    c->hT_rows.clear(); c->hT_seg0.assign(1, (int32_t)c->segE.size()); c->segT.clear();
    if ((rc = upload(c, c->hTrows, c->hT_rows.data(), 0))) return rc;
    if ((rc = upload(c, c->hTseg0, c->hT_seg0.data(), sizeof(int32_t)))) return rc;
    return 0;
}

void find_label_dups(lb_ctx *c) {
    c->vdup.assign(c->nvcs, -1);
    const size_t n = (size_t)c->n;
    if (c->h_labs.size() != (size_t)c->nlev * n || c->h_vlabs.size() != (size_t)c->nvcs * n) return;
    for (int m = 0; m < c->nvcs; m++)
        for (int l = 0; l < c->nlev; l++)
            if (c->P[l] == c->vP[m] &&
                std::memcmp(c->h_labs.data() + l * n, c->h_vlabs.data() + m * n, n * sizeof(int32_t)) == 0) {
                c->vdup[m] = l;
                break;
            }
}

int empty_vcs(lb_ctx *c) {
    c->nvcs = 0; c->sumPv = 0; c->vP.clear(); c->vnnz.clear(); c->voff.clear();
    c->h_vlabs.clear(); c->vdup.clear();
    int rc;
    if ((rc = dalloc(c, c->vlabs, 16))) return rc;
    if ((rc = dalloc(c, c->vptr, 16))) return rc;
    if ((rc = dalloc(c, c->vcol, 16))) return rc;
    if ((rc = dalloc(c, c->vrow, 16))) return rc;
    if ((rc = dalloc(c, c->vsig, 16))) return rc;
    if ((rc = dalloc(c, c->pv, 16))) return rc;
    if ((rc = dalloc(c, c->vrq, 16))) return rc;
    return 0;
}

int empty_levels(lb_ctx *c) {
    c->nlev = 0; c->sumPpair = 0; c->P.clear(); c->soff.assign(1, 0);
    c->h_labs.clear();
    find_label_dups(c);
    int rc;
    if ((rc = dalloc(c, c->labs, 16))) return rc;
    if ((rc = dalloc(c, c->sig, 16))) return rc;
    if ((rc = dalloc(c, c->pp, 16))) return rc;
    if ((rc = dalloc(c, c->rq, 16))) return rc;
    if ((rc = dalloc(c, c->dsq, 16))) return rc;
    return 0;
}

}  // namespace

// ================================================================================================== C ABI
extern "C" {

const char *lb_version(void) { return "lbhip 0.1 (gfx950)"; }

const char *lb_kernel_name(int k) { return (k >= 0 && k < LB_NKERNELS) ? kKernelNames[k] : ""; }

int lb_create(lb_ctx **out, int device, int vdtype, int r) {
    if (!out) return -1;
    *out = nullptr;
    if (r != 8 && r != 16 && r != 32) return -1;
    if (vdtype != LB_F32 && vdtype != LB_BF16 && vdtype != LB_F64) return -1;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return -3;
    if (hipSetDevice(device) != hipSuccess) return -3;
    lb_ctx *c = new lb_ctx();
    c->device = device; c->vdtype = vdtype; c->r = r;
    c->vbytes = vdtype == LB_F32 ? 4 : (vdtype == LB_BF16 ? 2 : 8);
    c->cbytes = vdtype == LB_F64 ? 8 : 4;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return -3; }
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->fork_ev, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->join_ev, hipEventDisableTiming) != hipSuccess) {
        lb_destroy(c);
        return -3;
    }
    *out = c;
    return 0;
}

void lb_destroy(lb_ctx *c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    if (c->side) hipStreamSynchronize(c->side);
    DBuf *bufs[] = {&c->rowptr, &c->col, &c->w, &c->d, &c->labs, &c->sig, &c->pp, &c->rq, &c->dsq, &c->vlabs, &c->vrq,
                    &c->vptr, &c->vcol, &c->vsig, &c->pv, &c->vrow, &c->pol_V, &c->pol_Vn, &c->pol_R, &c->pol_g,
                    &c->pol_part, &c->pol_sum, &c->blk, &c->blkdiag, &c->blkpart, &c->blkcoef, &c->blkout, &c->vst_part, &c->b_light, &c->b_heavy, &c->sw.hErows, &c->sw.hEseg0, &c->sw.hTrows, &c->sw.hTseg0, &c->sw.hEidx, &c->sw.hTidx,
                    &c->sw.segs, &c->sw.segout, &c->sw.segkap, &c->TI, &c->lam, &c->pt, &c->iptr, &c->rec, &c->V, &c->G,
                    &c->Y, &c->z, &c->M, &c->pdiag, &c->levtot, &c->levviol, &c->Z, &c->colsum_part, &c->vcs_part,
                    &c->row_part, &c->seg_part, &c->segs, &c->segout, &c->hErows, &c->hEseg0, &c->hTrows,
                    &c->hTseg0, &c->parts, &c->lamsum, &c->order, &c->tmask, &c->Kq, &c->stZ, &c->Zs, &c->cnts,
                    &c->Vstage, &c->DDb, &c->segkap, &c->seglist, &c->vslack, &c->vage, &c->vvol, &c->rorder,
                    &c->Yalt, &c->hubcnt, &c->hubpos, &c->hubown, &c->zpart, &c->Ysnap, &c->posb, &c->posp, &c->dbuf,
                    &c->b_off, &c->b_gid, &c->b_L, &c->b_C0, &c->b_Poff, &c->b_soff, &c->b_cg, &c->b_lab, &c->b_sig,
                    &c->b_pp, &c->b_rq, &c->b_dsq, &c->b_zY, &c->b_M, &c->b_pdiag, &c->b_ltot, &c->b_lviol, &c->b_rowv,
                    &c->b_parts, &c->b_lamsum, &c->b_tptr, &c->b_tidx, &c->b_R, &c->b_Vp, &c->b_Rp, &c->b_rgst};
    for (DBuf *b : bufs) dfree(*b);
    if (c->hstage) hipHostFree(c->hstage);
    c->hstage = nullptr; c->hstage_bytes = 0;
    drop_capture(c);
    if (c->sw_exec) hipGraphExecDestroy((hipGraphExec_t)c->sw_exec);
    c->sw_exec = nullptr;
    dfree(c->lposd);
    if (c->fork_ev) hipEventDestroy(c->fork_ev);
    if (c->join_ev) hipEventDestroy(c->join_ev);
    if (c->side) hipStreamDestroy(c->side);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

const char *lb_last_error(const lb_ctx *c) { return c ? c->err.c_str() : "no context"; }

int lb_set_option(lb_ctx *c, const char *name, int64_t value) {
    drop_capture(c);
    if (!c || !name) return fail(c, -1, "lb_set_option: bad arguments");
    if (std::strcmp(name, "row_order") == 0) { c->use_order = value != 0; return 0; }
    if (std::strcmp(name, "overlap") == 0) { c->overlap = value != 0; return 0; }
    if (std::strcmp(name, "sweep_graph") == 0) { c->sweep_graph = value != 0; return 0; }
    if (std::strcmp(name, "sweep_fused") == 0) { c->sweep_fused = value != 0; return 0; }
    if (std::strcmp(name, "sweep_pair_segs") == 0) { c->sweep_pair_segs = value != 0; return 0; }
    if (std::strcmp(name, "sweep_post_merge") == 0) { c->sweep_post_merge = value != 0; return 0; }
    if (std::strcmp(name, "sweep_two") == 0) { c->sweep_two = value != 0; return 0; }
    if (std::strcmp(name, "sweep_delta") == 0) { c->sweep_delta = value != 0; return 0; }
    if (std::strcmp(name, "sweep_seg_grid") == 0) {
        if (value < 0 || value > 4096) return fail(c, -1, "lb_set_option: sweep_seg_grid is 0..4096");
        c->sweep_seg_grid = value;
        return 0;
    }
    if (std::strcmp(name, "vcs_tile") == 0) {
        if (value != 16 && value != 32) return fail(c, -1, "lb_set_option: vcs_tile is 16 or 32");
        c->vcs_tile = (int)value;
        return 0;
    }
    if (std::strcmp(name, "vcs_grid") == 0) {
        if (value < 0 || value > LB_VCS_BLOCKS) return fail(c, -1, "lb_set_option: vcs_grid must be in [0, 1536]");
        c->vcs_grid = value;
        return 0;
    }
    if (std::strcmp(name, "vcs_kernel") == 0) {
        if (value < -1 || value > 1) return fail(c, -1, "lb_set_option: vcs_kernel is -1 (by storage type), 0 or 1");
        c->vcs_kernel = (int)value;
        return 0;
    }
    if (std::strcmp(name, "step_wave") == 0) { c->step_wave = value != 0; return 0; }
    if (std::strcmp(name, "seg_grid") == 0) {
        if (value < 0 || value > ((int64_t)1 << 30)) return fail(c, -1, "lb_set_option: seg_grid must be >= 0");
        c->seg_grid = value;
        return 0;
    }
    if (std::strcmp(name, "batch_light") == 0 || std::strcmp(name, "batch_light_grid") == 0 ||
        std::strcmp(name, "batch_heavy_grid") == 0) {
        if (value < 1 || value > ((int64_t)1 << 30)) return fail(c, -1, "lb_set_option: batch shapes must be positive");
        if (name[6] == 'l' && name[11] == '\0') { c->b_light_max = value; c->b_lists_ok = false; }
        else if (name[6] == 'l') c->b_lgrid = value;
        else c->b_hgrid = value;
        return 0;
    }
    if (std::strcmp(name, "sweep_heavy") == 0 || std::strcmp(name, "sweep_seg") == 0) {
        if (value < 1 || value > ((int64_t)1 << 30)) return fail(c, -1, "lb_set_option: sweep segment values must be positive");
        if (name[6] == 'h') { c->sw.heavyE = value; c->sw.heavyT = value; } else c->sw.seg_len = (int)value;
        return c->n ? build_sweep_segs(c) : 0;
    }
    return fail(c, -1, std::string("lb_set_option: unknown option ") + name);
}

int lb_set_tuning(lb_ctx *c, int heavy_edges, int heavy_tris, int seg_len, int row_grid) {
    drop_capture(c);
    if (!c) return -1;
    if (heavy_edges < 1 || heavy_tris < 1 || seg_len < 1 || row_grid < 0) return fail(c, -1, "tuning values must be positive");
    c->heavy_edges = heavy_edges; c->heavy_tris = heavy_tris; c->seg_len = seg_len;
    if (row_grid > 0) c->row_grid = row_grid;
    return 0;
}

int lb_set_graph(lb_ctx *c, int64_t n, const int64_t *rowptr, const int32_t *col, const double *weights,
                 const double *d, double L, double C0) {
    drop_capture(c);
    if (!c) return -1;
    if (n <= 0 || n >= ((int64_t)1 << 31) || !rowptr || !col || !d || !(L > 0)) return fail(c, -1, "lb_set_graph: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    if (rowptr[0] != 0) return fail(c, -1, "lb_set_graph: rowptr[0] != 0");
    for (int64_t i = 0; i < n; i++)
        if (rowptr[i + 1] < rowptr[i]) return fail(c, -1, "lb_set_graph: rowptr not monotone");
    const int64_t nnz = rowptr[n];
    for (int64_t e = 0; e < nnz; e++)
        if (col[e] < 0 || col[e] >= n) return fail(c, -1, "lb_set_graph: column index out of range");
    c->n = n; c->nnz = nnz; c->L = L; c->C0 = C0; c->weighted = weights != nullptr;
    c->bG = 0; c->blev = false; c->bsumP = 0;
    // the eigensolver blocks and the polish state are sized by n: a new graph invalidates both
    c->nblk = 0; c->ndiag = 0; c->pol_ok = false;
    c->row_order.clear();
    c->b_lists_ok = false;
    c->h_rowptr.assign(rowptr, rowptr + n + 1);
    c->h_d.assign(d, d + n);
    int rc;
    if ((rc = upload(c, c->rowptr, rowptr, sizeof(int64_t) * (n + 1)))) return rc;
    if ((rc = upload(c, c->col, col, sizeof(int32_t) * nnz))) return rc;
    if (weights) { if ((rc = upload_tc(c, c->w, weights, nnz))) return rc; }
    if ((rc = upload(c, c->d, d, sizeof(double) * n))) return rc;
    build_segments(n, rowptr, c->heavy_edges, c->seg_len, 0, c->hE_rows, c->hE_seg0, c->segE, 0);
    c->built_heavyE = c->heavy_edges;
    c->hE_idx.assign(n, -1);
    for (size_t k = 0; k < c->hE_rows.size(); k++) c->hE_idx[c->hE_rows[k]] = (int32_t)k;
    c->hT_idx.assign(n, -1);
    c->h_iptr.assign(n + 1, 0);
    if ((rc = upload(c, c->hErows, c->hE_rows.data(), sizeof(int32_t) * c->hE_rows.size()))) return rc;
    if ((rc = upload(c, c->hEseg0, c->hE_seg0.data(), sizeof(int32_t) * c->hE_seg0.size()))) return rc;
    if ((rc = dalloc(c, c->V, (size_t)n * c->r * c->vbytes))) return rc;
    if ((rc = dalloc(c, c->G, (size_t)n * c->r * c->cbytes))) return rc;
    // the graph defines n: reset levels, working sets and the triangle pool
    c->T = 0; c->segT.clear();
    if ((rc = empty_levels(c)) || (rc = empty_vcs(c)) || (rc = empty_triangles(c))) return rc;
    if ((rc = rebuild_segs(c))) return rc;
    if ((rc = refresh_lamsum(c))) return rc;
    if ((rc = ensure_eval_buffers(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_eval = false;
    return 0;
}

int lb_set_levels(lb_ctx *c, int nlev, const int32_t *labs, const int32_t *P, const double *sig, const double *rq,
                  const double *dsq, double pairlvl_weight) {
    drop_capture(c);
    if (!c || !c->n) return fail(c, -1, "lb_set_levels: set the graph first");
    if (nlev < 0 || nlev > LB_MAXLEV) return fail(c, -1, "lb_set_levels: at most 8 pair levels");
    HIPCHK(c, hipSetDevice(c->device));
    if (nlev == 0) { int rc = empty_levels(c); return rc ? rc : ensure_eval_buffers(c); }
    int64_t sP = 0, sP2 = 0;
    std::vector<int64_t> soff(1, 0);
    for (int l = 0; l < nlev; l++) {
        if (P[l] <= 0) return fail(c, -1, "lb_set_levels: P must be positive");
        sP += P[l]; sP2 += (int64_t)P[l] * P[l]; soff.push_back(sP2);
    }
    if (sP > (1 << 20)) return fail(c, -1, "lb_set_levels: too many clusters");
    for (int l = 0; l < nlev; l++)
        for (int64_t i = 0; i < c->n; i++)
            if (labs[(int64_t)l * c->n + i] < 0 || labs[(int64_t)l * c->n + i] >= P[l]) return fail(c, -1, "lb_set_levels: label out of range");
    c->nlev = nlev; c->P.assign(P, P + nlev); c->soff = soff; c->sumPpair = (int)sP; c->pw = pairlvl_weight;
    c->h_labs.assign(labs, labs + (size_t)nlev * c->n);
    find_label_dups(c);
    int rc;
    if ((rc = upload(c, c->labs, labs, sizeof(int32_t) * (size_t)nlev * c->n))) return rc;
    if ((rc = upload(c, c->sig, sig, sizeof(double) * sP2))) return rc;
    if ((rc = dalloc(c, c->pp, sizeof(double) * sP2))) return rc;
    if ((rc = upload(c, c->rq, rq, sizeof(double) * sP))) return rc;
    if ((rc = upload(c, c->dsq, dsq, sizeof(double) * sP))) return rc;
    c->have_eval = false;
    return ensure_eval_buffers(c);
}

int lb_set_level_sig(lb_ctx *c, const double *sig) {
    drop_capture(c);
    if (!c || !c->nlev) return fail(c, -1, "lb_set_level_sig: no pair levels");
    HIPCHK(c, hipSetDevice(c->device));
    return upload(c, c->sig, sig, sizeof(double) * c->soff.back());
}

int lb_set_vcs(lb_ctx *c, int nvcs, const int32_t *labs, const int32_t *P, const double *rq, const int64_t *ptr,
               const int32_t *col, const double *sig, double wv) {
    drop_capture(c);
    if (!c || !c->n) return fail(c, -1, "lb_set_vcs: set the graph first");
    if (nvcs < 0 || nvcs > LB_MAXVCS) return fail(c, -1, "lb_set_vcs: at most 4 VCS levels");
    HIPCHK(c, hipSetDevice(c->device));
    if (nvcs == 0) { int rc = empty_vcs(c); return rc ? rc : ensure_eval_buffers(c); }
    const int64_t n = c->n;
    std::vector<int64_t> absptr((size_t)nvcs * (n + 1));
    std::vector<int64_t> vnnz, voff;
    int64_t tot = 0, sP = 0;
    for (int m = 0; m < nvcs; m++) {
        const int64_t *pm = ptr + (int64_t)m * (n + 1);
        if (P[m] <= 0) return fail(c, -1, "lb_set_vcs: P must be positive");
        if (pm[0] != 0) return fail(c, -1, "lb_set_vcs: ptr must start at 0");
        for (int64_t i = 0; i < n; i++) {
            if (pm[i + 1] < pm[i]) return fail(c, -1, "lb_set_vcs: ptr not monotone");
            const int32_t lb = labs[(int64_t)m * n + i];
            if (lb < 0 || lb >= P[m]) return fail(c, -1, "lb_set_vcs: label out of range");
        }
        for (int64_t e = 0; e < pm[n]; e++)
            if (col[tot + e] < 0 || col[tot + e] >= P[m]) return fail(c, -1, "lb_set_vcs: working-set cluster out of range");
        for (int64_t i = 0; i <= n; i++) absptr[(size_t)m * (n + 1) + i] = pm[i] + tot;
        voff.push_back(tot); vnnz.push_back(pm[n]); tot += pm[n]; sP += P[m];
    }
    c->nvcs = nvcs; c->vP.assign(P, P + nvcs); c->vnnz = vnnz; c->voff = voff; c->sumPv = (int)sP; c->wv = wv;
    c->h_vlabs.assign(labs, labs + (size_t)nvcs * n);
    find_label_dups(c);
    int rc;
    if ((rc = upload(c, c->vlabs, labs, sizeof(int32_t) * (size_t)nvcs * n))) return rc;
    if ((rc = upload(c, c->vptr, absptr.data(), sizeof(int64_t) * absptr.size()))) return rc;
    if ((rc = upload(c, c->vcol, col, sizeof(int32_t) * std::max<int64_t>(tot, 1)))) return rc;
    {
        std::vector<int32_t> vrow((size_t)std::max<int64_t>(tot, 1), 0);
        for (int m = 0; m < nvcs; m++) {
            const int64_t *pm = absptr.data() + (size_t)m * (n + 1);
            for (int64_t i = 0; i < n; i++)
                for (int64_t e = pm[i]; e < pm[i + 1]; e++) vrow[(size_t)e] = (int32_t)i;
        }
        if ((rc = upload(c, c->vrow, vrow.data(), sizeof(int32_t) * vrow.size()))) return rc;
    }
    if ((rc = upload_tc(c, c->vsig, sig, tot))) return rc;
    if ((rc = dalloc(c, c->pv, c->cbytes * std::max<int64_t>(tot, 1)))) return rc;
    if ((rc = upload(c, c->vrq, rq, sizeof(double) * sP))) return rc;
    if ((rc = dalloc(c, c->vslack, std::max<int64_t>(tot, 1)))) return rc;
    if ((rc = dalloc(c, c->vage, std::max<int64_t>(tot, 1)))) return rc;
    HIPCHK(c, hipMemsetAsync(c->vage.p, 0, std::max<int64_t>(tot, 1), c->stream));
    HIPCHK(c, hipMemsetAsync(c->vslack.p, 0, std::max<int64_t>(tot, 1), c->stream));
    // cluster volumes max(sum_{i in R} d_i, 1) (vcsops.py:37)
    std::vector<double> vol(sP, 0.0);
    {
        int64_t o = 0;
        for (int m = 0; m < nvcs; m++) {
            for (int64_t i = 0; i < n; i++) vol[o + labs[(int64_t)m * n + i]] += c->h_d[i];
            o += P[m];
        }
        for (auto &x : vol) x = std::max(x, 1.0);
    }
    if ((rc = upload(c, c->vvol, vol.data(), sizeof(double) * sP))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_eval = false;
    return ensure_eval_buffers(c);
}

int lb_set_vcs_age(lb_ctx *c, const uint8_t *age) {
    drop_capture(c);
    if (!c || !c->nvcs || !age) return fail(c, -1, "lb_set_vcs_age: no VCS levels");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t tot = 0;
    for (int64_t x : c->vnnz) tot += x;
    return upload(c, c->vage, age, tot);
}

int lb_get_vcs_state(lb_ctx *c, double *sig, uint8_t *age) {
    if (!c || !c->nvcs) return fail(c, -1, "lb_get_vcs_state: no VCS levels");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t tot = 0;
    for (int64_t x : c->vnnz) tot += x;
    int rc;
    if (sig && (rc = download_tc(c, c->vsig, sig, tot))) return rc;
    if (age && tot) {
        HIPCHK(c, hipMemcpyAsync(age, c->vage.p, tot, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int lb_set_weights(lb_ctx *c, double pairlvl_weight, double wv) {
    drop_capture(c);
    if (!c) return -1;
    if (!(pairlvl_weight > 0.0) || !(wv > 0.0)) return fail(c, -1, "lb_set_weights: weights must be positive");
    c->pw = pairlvl_weight; c->wv = wv;
    return 0;
}

int lb_set_vcs_sig(lb_ctx *c, const double *sig) {
    drop_capture(c);
    if (!c || !c->nvcs) return fail(c, -1, "lb_set_vcs_sig: no VCS levels");
    HIPCHK(c, hipSetDevice(c->device));
    int64_t tot = 0;
    for (int64_t x : c->vnnz) tot += x;
    return upload_tc(c, c->vsig, sig, tot);
}

int lb_set_triangles(lb_ctx *c, int64_t T, const int32_t *TI, const double *lam) {
    drop_capture(c);
    if (c) { c->b_lamsum_ok = false; c->b_lists_ok = false; }
    if (!c || !c->n) return fail(c, -1, "lb_set_triangles: set the graph first");
    if (T < 0 || T > ((int64_t)1 << 29)) return fail(c, -1, "lb_set_triangles: bad T");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    int rc;
    if (T == 0) {
        c->T = 0;
        c->hT_idx.assign(n, -1);
        c->h_iptr.assign(n + 1, 0);
        if ((rc = empty_triangles(c)) || (rc = rebuild_segs(c)) || (rc = refresh_lamsum(c))) return rc;
        c->have_eval = false;
        return 0;
    }
    for (int64_t e = 0; e < 3 * T; e++)
        if (TI[e] < 0 || TI[e] >= n) return fail(c, -1, "lb_set_triangles: vertex out of range");
    if (c->bG) {                                      // batch mode: a triangle lies inside one graph
        auto graph_of = [&](int64_t v) { return std::upper_bound(c->hboff.begin(), c->hboff.end(), v) - c->hboff.begin() - 1; };
        std::vector<int64_t> tptr(c->bG + 1, 0);
        std::vector<int32_t> tg((size_t)T);
        for (int64_t t = 0; t < T; t++) {
            const auto g0 = graph_of(TI[3 * t]);
            if (graph_of(TI[3 * t + 1]) != g0 || graph_of(TI[3 * t + 2]) != g0)
                return fail(c, -1, "lb_set_triangles: a triangle joins two graphs of the batch");
            tg[t] = (int32_t)g0;
            tptr[g0 + 1]++;
        }
        for (int g = 0; g < c->bG; g++) tptr[g + 1] += tptr[g];
        std::vector<int64_t> fillp(tptr.begin(), tptr.end() - 1);
        std::vector<int32_t> tidx((size_t)T);
        for (int64_t t = 0; t < T; t++) tidx[fillp[tg[t]]++] = (int32_t)t;
        int rc2;
        if ((rc2 = upload(c, c->b_tptr, tptr.data(), sizeof(int64_t) * tptr.size()))) return rc2;
        if ((rc2 = upload(c, c->b_tidx, tidx.data(), sizeof(int32_t) * tidx.size()))) return rc2;
    }
    // vertex -> (t, role) incidence in t order (alm_vcs.py:94-104)
    std::vector<int64_t> cnt(n + 1, 0);
    for (int64_t e = 0; e < 3 * T; e++) cnt[TI[e] + 1]++;
    for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    const bool dbl = c->cbytes == 8;
    const size_t rsz = dbl ? sizeof(TriRec<double>) : sizeof(TriRec<float>);
    std::vector<unsigned char> recs((size_t)3 * T * rsz);
    for (int64_t t = 0; t < T; t++) {
        const int32_t a = TI[3 * t], b = TI[3 * t + 1], cc = TI[3 * t + 2];
        for (int r = 0; r < 3; r++) {
            const int32_t i = TI[3 * t + r];
            const int64_t k = pos[i]++;
            int32_t u, w;
            if (r == 0) { u = b; w = cc; } else if (r == 1) { u = a; w = cc; } else { u = a; w = b; }
            if (dbl) {
                TriRec<double> x{u, w, (int32_t)(3 * t + r), 0, lam[t]};
                std::memcpy(&recs[k * rsz], &x, rsz);
            } else {
                TriRec<float> x{u, w, (int32_t)(3 * t + r), (float)lam[t]};
                std::memcpy(&recs[k * rsz], &x, rsz);
            }
        }
    }
    c->T = T;
    if ((rc = upload(c, c->iptr, cnt.data(), sizeof(int64_t) * (n + 1)))) return rc;
    if ((rc = upload(c, c->rec, recs.data(), recs.size()))) return rc;
    if ((rc = upload(c, c->TI, TI, sizeof(int32_t) * 3 * T))) return rc;
    if ((rc = upload_tc(c, c->lam, lam, T))) return rc;
    if ((rc = dalloc(c, c->pt, c->cbytes * T))) return rc;
    build_segments(n, cnt.data(), c->heavy_tris, c->seg_len, 1, c->hT_rows, c->hT_seg0, c->segT, (int32_t)c->segE.size());
    c->built_heavyT = c->heavy_tris;
    c->hT_idx.assign(n, -1);
    for (size_t k = 0; k < c->hT_rows.size(); k++) c->hT_idx[c->hT_rows[k]] = (int32_t)k;
    c->h_iptr = cnt;
    if ((rc = upload(c, c->hTrows, c->hT_rows.data(), sizeof(int32_t) * c->hT_rows.size()))) return rc;
    if ((rc = upload(c, c->hTseg0, c->hT_seg0.data(), sizeof(int32_t) * c->hT_seg0.size()))) return rc;
    if ((rc = rebuild_segs(c))) return rc;
    if ((rc = refresh_lamsum(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_eval = false;
    return 0;
}

int lb_set_lam(lb_ctx *c, const double *lam) {
    drop_capture(c);
    if (c) c->b_lamsum_ok = false;
    if (!c || !c->n) return fail(c, -1, "lb_set_lam: no graph");
    if (!c->T) return 0;
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = upload_tc(c, c->lam, lam, c->T))) return rc;
    const int64_t nrec = 3 * c->T;
    if (c->cbytes == 8) hipLaunchKernelGGL(lb_k_inc_lam<double>, dim3(1024), dim3(256), 0, c->stream, (TriRec<double> *)c->rec.p, nrec, (const double *)c->lam.p);
    else hipLaunchKernelGGL(lb_k_inc_lam<float>, dim3(1024), dim3(256), 0, c->stream, (TriRec<float> *)c->rec.p, nrec, (const float *)c->lam.p);
    HIPCHK(c, hipGetLastError());
    if ((rc = refresh_lamsum(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_set_V(lb_ctx *c, const void *V, int host_dtype) {
    if (!c || !c->n || !V) return fail(c, -1, "lb_set_V: no graph or V");
    if (host_dtype != LB_F64 && host_dtype != LB_F32) return fail(c, -1, "lb_set_V: host dtype must be LB_F64 or LB_F32");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r, dt = c->vdtype;
    const size_t bytes = (size_t)n * r * c->vbytes;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    int rc;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    par_rows(n, [&](int64_t a, int64_t b) {
        if (host_dtype == LB_F64) rows_to_storage(a, b, r, (const double *)V, dt, st, ord);
        else rows_to_storage(a, b, r, (const float *)V, dt, st, ord);
    });
    return upload(c, c->V, st, bytes);
}

int lb_get_V(lb_ctx *c, void *V, int host_dtype) {
    if (!c || !c->n || !V) return fail(c, -1, "lb_get_V: no graph or V");
    if (host_dtype != LB_F64 && host_dtype != LB_F32) return fail(c, -1, "lb_get_V: host dtype must be LB_F64 or LB_F32");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r, dt = c->vdtype;
    const size_t bytes = (size_t)n * r * c->vbytes;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    int rc;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    HIPCHK(c, hipMemcpyAsync(st, c->V.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    par_rows(n, [&](int64_t a, int64_t b) {
        if (host_dtype == LB_F64) storage_to_rows(a, b, r, st, dt, (double *)V, ord);
        else storage_to_rows(a, b, r, st, dt, (float *)V, ord);
    });
    return 0;
}

int lb_eval_fgrad(lb_ctx *c, double rho, double rt, double *parts) {
    if (!c || !c->n) return fail(c, -1, "lb_eval_fgrad: no graph");
    if (c->bG) return fail(c, -1, "lb_eval_fgrad: the context holds a batch (use lb_batch_eval)");
    if (!(rho >= 0.0) || (rho > 0.0 && c->T && !(rt > 0.0))) return fail(c, -1, "lb_eval_fgrad: need rho >= 0 and rt > 0 when rho > 0");
    HIPCHK(c, hipSetDevice(c->device));
    eval_fn f = pick_eval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_eval_fgrad: unsupported rank / dtype");
    int rc = f(c, rho, rt, nullptr);
    if (rc) return rc;
    c->have_eval = true; c->last_rho = rho;
    if (parts) {
        HIPCHK(c, hipMemcpyAsync(parts, c->parts.p, sizeof(double) * LB_NPARTS, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    return 0;
}

int lb_get_G(lb_ctx *c, void *G, int host_dtype) {
    if (!c || !c->n || !G) return fail(c, -1, "lb_get_G: no graph or G");
    if (!c->have_eval) return fail(c, -1, "lb_get_G: no evaluation yet");
    if (host_dtype != LB_F64 && host_dtype != LB_F32) return fail(c, -1, "lb_get_G: host dtype must be LB_F64 or LB_F32");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r, gt = c->cbytes == 8 ? LB_F64 : LB_F32;
    const size_t bytes = (size_t)n * r * c->cbytes;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    int rc;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    HIPCHK(c, hipMemcpyAsync(st, c->G.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    par_rows(n, [&](int64_t a, int64_t b) {
        if (host_dtype == LB_F64) storage_to_rows(a, b, r, st, gt, (double *)G, ord);
        else storage_to_rows(a, b, r, st, gt, (float *)G, ord);
    });
    return 0;
}

int lb_set_row_order(lb_ctx *c, const int64_t *order) {
    if (!c || !c->n) return fail(c, -1, "lb_set_row_order: set the graph first");
    if (c->bG) return fail(c, -1, "lb_set_row_order: not available on a batch context");
    if (!order) { c->row_order.clear(); return 0; }
    std::vector<unsigned char> seen((size_t)c->n, 0);
    for (int64_t k = 0; k < c->n; k++) {
        if (order[k] < 0 || order[k] >= c->n || seen[(size_t)order[k]]) return fail(c, -1, "lb_set_row_order: not a permutation of the rows");
        seen[(size_t)order[k]] = 1;
    }
    c->row_order.assign(order, order + c->n);
    return 0;
}

int lb_get_multipliers(lb_ctx *c, double *pp, double *pt, double *pv) {
    if (!c || !c->have_eval) return fail(c, -1, "lb_get_multipliers: no evaluation yet");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if (pp && c->nlev) {
        HIPCHK(c, hipMemcpyAsync(pp, c->pp.p, sizeof(double) * c->soff.back(), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    if (pt && c->T) { if ((rc = download_tc(c, c->pt, pt, c->T))) return rc; }
    if (pv && c->nvcs) {
        int64_t tot = 0;
        for (int64_t x : c->vnnz) tot += x;
        if ((rc = download_tc(c, c->pv, pv, tot))) return rc;
    }
    return 0;
}

int lb_phr_update(lb_ctx *c) {
    drop_capture(c);
    if (c) c->b_lamsum_ok = false;
    if (!c || !c->have_eval) return fail(c, -1, "lb_phr_update: no evaluation yet");
    if (!(c->last_rho > 0.0)) return fail(c, -1, "lb_phr_update: the last evaluation had rho = 0");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->nlev) HIPCHK(c, hipMemcpyAsync(c->sig.p, c->pp.p, sizeof(double) * c->soff.back(), hipMemcpyDeviceToDevice, c->stream));
    if (c->bG && c->blev && c->b_nsig)           // batch mode: every graph's pair level
        HIPCHK(c, hipMemcpyAsync(c->b_sig.p, c->b_pp.p, sizeof(double) * c->b_nsig, hipMemcpyDeviceToDevice, c->stream));
    if (c->nvcs) {
        int64_t tot = 0;
        for (int64_t x : c->vnnz) tot += x;
        HIPCHK(c, hipMemcpyAsync(c->vsig.p, c->pv.p, c->cbytes * tot, hipMemcpyDeviceToDevice, c->stream));
        if (tot) hipLaunchKernelGGL(lb_k_age, dim3(512), dim3(256), 0, c->stream, (uint8_t *)c->vage.p, (const uint8_t *)c->vslack.p, tot);
    }
    if (c->T) {
        HIPCHK(c, hipMemcpyAsync(c->lam.p, c->pt.p, c->cbytes * c->T, hipMemcpyDeviceToDevice, c->stream));
        const int64_t nrec = 3 * c->T;
        if (c->cbytes == 8) hipLaunchKernelGGL(lb_k_inc_lam<double>, dim3(1024), dim3(256), 0, c->stream, (TriRec<double> *)c->rec.p, nrec, (const double *)c->lam.p);
        else hipLaunchKernelGGL(lb_k_inc_lam<float>, dim3(1024), dim3(256), 0, c->stream, (TriRec<float> *)c->rec.p, nrec, (const float *)c->lam.p);
        HIPCHK(c, hipGetLastError());
        int rc = refresh_lamsum(c);
        if (rc) return rc;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_eval = false;
    return 0;
}

int lb_sweeps(lb_ctx *c, int nsw, double rho, double rt, double damp, double kap2, const int64_t *orders, int nbatch) {
    drop_capture(c);
    if (c && c->bG) return fail(c, -1, "lb_sweeps: not available on a batch context");
    if (!c || !c->n) return fail(c, -1, "lb_sweeps: no graph");
    if (nsw < 0 || nbatch < 1 || nbatch > c->n || !orders) return fail(c, -1, "lb_sweeps: bad arguments");
    if (!(rho >= 0.0) || !(rt >= 0.0) || !(damp >= 0.0)) return fail(c, -1, "lb_sweeps: rho, rt, damp must be >= 0");
    HIPCHK(c, hipSetDevice(c->device));
    sweep_fn f = pick_sweep(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_sweeps: unsupported rank / dtype");
    const int64_t n = c->n;
    const int R = c->r;
    const int S = (int)(R * c->vbytes / 16);
    const int64_t maxb = n / nbatch + 1;
    int64_t nnzv = 0;
    for (int64_t x : c->vnnz) nnzv += x;
    int rc;
    if ((rc = dalloc(c, c->order, sizeof(int32_t) * n))) return rc;
    if ((rc = dalloc(c, c->tmask, (size_t)nbatch * std::max(c->sumPpair, 1)))) return rc;
    if ((rc = dalloc(c, c->Kq, sizeof(double) * std::max(c->sumPpair, 1)))) return rc;
    if ((rc = dalloc(c, c->stZ, sizeof(double) * std::max<int64_t>(nnzv, 1)))) return rc;
    if ((rc = dalloc(c, c->Zs, sizeof(double) * (size_t)std::max(c->sumPv, 1) * R))) return rc;
    if ((rc = dalloc(c, c->cnts, sizeof(double) * std::max(c->sumPv, 1)))) return rc;
    if ((rc = dalloc(c, c->vst_part, sizeof(double) * (size_t)LB_RED_BLOCKS * std::max(c->sumPv * (R + 1), 1)))) return rc;
    if ((rc = dalloc(c, c->Vstage, (size_t)maxb * S * 16))) return rc;
    if ((rc = dalloc(c, c->DDb, sizeof(double) * (size_t)maxb * R))) return rc;
    {
        // fused sweeps: the alternate [z | Y] and the hub-row meeting state (counts start at zero; each batch's last
        // arrival clears its row's count)
        const size_t nhub = std::max<size_t>(1, c->sw.hE_rows.size() + c->sw.hT_rows.size());
        if ((rc = dalloc(c, c->Yalt, sizeof(double) * (size_t)(R + (c->sumPpair + c->sumPv) * R)))) return rc;
        if ((rc = dalloc(c, c->hubcnt, sizeof(int32_t) * nhub))) return rc;
        if ((rc = dalloc(c, c->hubpos, sizeof(int64_t) * nhub))) return rc;
        if ((rc = dalloc(c, c->hubown, sizeof(double) * nhub * (size_t)(R + 1)))) return rc;
        HIPCHK(c, hipMemsetAsync(c->hubcnt.p, 0, sizeof(int32_t) * nhub, c->stream));
        if ((rc = dalloc(c, c->zpart, sizeof(double) * LB_ZPART * (size_t)R))) return rc;
        if ((rc = dalloc(c, c->Ysnap, sizeof(double) * (size_t)std::max(c->sumPpair + c->sumPv, 1) * R))) return rc;
        if ((rc = dalloc(c, c->posb, sizeof(int32_t) * (size_t)n))) return rc;
        if ((rc = dalloc(c, c->posp, sizeof(int32_t) * (size_t)n))) return rc;
        HIPCHK(c, hipMemsetAsync(c->posb.p, 0xFF, sizeof(int32_t) * (size_t)n, c->stream));   // -1: no batch yet
        c->sw_tag = 0;
        const size_t wd = (size_t)LB_ZPART * R + (size_t)(c->sumPpair + c->sumPv) * R;
        if ((rc = dalloc(c, c->dbuf, sizeof(double) * 2 * wd))) return rc;
        HIPCHK(c, hipMemsetAsync(c->dbuf.p, 0, sizeof(double) * 2 * wd, c->stream));
        HIPCHK(c, hipMemsetAsync(c->zpart.p, 0, sizeof(double) * LB_ZPART * (size_t)R, c->stream));
    }
    std::vector<int32_t> o32(n);
    std::vector<uint8_t> seen(n);
    std::vector<int64_t> bpos(nbatch + 1), lpos(nbatch + 1);
    std::vector<int32_t> list;
    // batches = np.array_split(order, nbatch): the first n % nbatch batches hold one extra row
    const int64_t q = n / nbatch, rem = n % nbatch;
    for (int b = 0; b <= nbatch; b++) bpos[b] = (int64_t)b * q + std::min<int64_t>(b, rem);
    for (int sw = 0; sw < nsw; sw++) {
        const int64_t *ord = orders + (int64_t)sw * n;
        if (c->sw_tag > INT32_MAX - nbatch - 1) {        // batch tags about to wrap: clear the row map and the delta
            HIPCHK(c, hipMemsetAsync(c->posb.p, 0xFF, sizeof(int32_t) * (size_t)n, c->stream));   // buffers (whose
            HIPCHK(c, hipMemsetAsync(c->dbuf.p, 0, c->dbuf.bytes, c->stream));                     // parity restarts)
            c->sw_tag = 0;
        }
        std::fill(seen.begin(), seen.end(), 0);
        for (int64_t p = 0; p < n; p++) {
            const int64_t i = ord[p];
            if (i < 0 || i >= n || seen[i]) return fail(c, -1, "lb_sweeps: order is not a permutation of 0..n-1");
            seen[i] = 1;
            o32[p] = (int32_t)i;
        }
        list.clear();
        for (int b = 0; b < nbatch; b++) {
            lpos[b] = (int64_t)list.size();
            for (int64_t p = bpos[b]; p < bpos[b + 1]; p++) {
                const int32_t i = o32[p];
                const SweepSegs &w = c->sw;
                const int32_t ke = w.hE_idx[i], kt = w.hT_idx[i];
                if (ke >= 0) for (int32_t s = w.hE_seg0[ke]; s < w.hE_seg0[ke + 1]; s++) list.push_back(s);
                if (kt >= 0) for (int32_t s = w.hT_seg0[kt]; s < w.hT_seg0[kt + 1]; s++) list.push_back(s);
            }
        }
        lpos[nbatch] = (int64_t)list.size();
        HIPCHK(c, hipMemcpyAsync(c->order.p, o32.data(), sizeof(int32_t) * n, hipMemcpyHostToDevice, c->stream));
        // every hub segment is listed once per sweep, so the list keeps its size (and its buffer) from sweep to sweep
        if ((rc = dalloc(c, c->seglist, sizeof(int32_t) * std::max<size_t>(list.size(), 1)))) return rc;
        if ((rc = dalloc(c, c->lposd, sizeof(int64_t) * (size_t)(nbatch + 1)))) return rc;
        if (!list.empty()) HIPCHK(c, hipMemcpyAsync(c->seglist.p, list.data(), sizeof(int32_t) * list.size(), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->lposd.p, lpos.data(), sizeof(int64_t) * (size_t)(nbatch + 1), hipMemcpyHostToDevice, c->stream));
        if ((rc = f(c, rho, rt, damp, kap2, nbatch, bpos))) return rc;
        // host vectors are reused by the next sweep: wait for the uploads (and the sweep) to finish
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->have_eval = false;
    return 0;
}

int lb_separate_triangles(lb_ctx *c, int fam, double thresh, int64_t maxh, int64_t maxk, int64_t hub2, int64_t cap2,
                          double vtol, int64_t cap, int nchunk, int nb, int64_t *TI_out, double *val_out,
                          int64_t *count, double *vmin) {
    drop_capture(c);
    // on a batch context only the wedge family: its candidates (two edges at a centre) stay inside one graph
    if (c && c->bG && fam != 0) return fail(c, -1, "lb_separate_triangles: only fam 0 on a batch context");
    if (!c || !c->n) return fail(c, -1, "lb_separate_triangles: no graph");
    if ((fam != 0 && fam != 1) || maxh < 1 || maxk < 1 || hub2 < 0 || cap2 < 0 || cap < 0 || nchunk < 1 || nb < 1 ||
        !(vtol > 0.0) || !(vtol < 2.0) || !TI_out || !val_out || !count || !vmin)
        return fail(c, -1, "lb_separate_triangles: bad arguments");
    if (fam == 1 && hub2 * cap2 > LB_SEP1_CAP) return fail(c, -1, "lb_separate_triangles: hub2 * cap2 exceeds 2048");
    if (fam == 1 && maxk > LB_SEP1_CAP) return fail(c, -1, "lb_separate_triangles: maxk exceeds 2048");
    HIPCHK(c, hipSetDevice(c->device));
    sep_fn f = pick_sep(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_separate_triangles: unsupported rank / dtype");
    std::vector<int64_t> out;
    std::vector<double> val;
    int rc = f(c, fam, thresh, maxh, maxk, hub2, cap2, vtol, cap, nchunk, nb, out, val, vmin);
    if (rc) return rc;
    const int64_t m = (int64_t)val.size();
    if (m > cap) return fail(c, 1, "lb_separate_triangles: more candidates than cap");
    std::memcpy(TI_out, out.data(), sizeof(int64_t) * 3 * m);
    std::memcpy(val_out, val.data(), sizeof(double) * m);
    *count = m;
    return 0;
}

int lb_separate_vcs(lb_ctx *c, int m, double tol, int mode, int64_t nadd, int64_t *out_i, int32_t *out_R,
                    double *out_score, int64_t *count, double *stats) {
    drop_capture(c);
    if (c && c->bG) return fail(c, -1, "lb_separate_vcs: not available on a batch context");
    if (!c || !c->n || m < 0 || m >= c->nvcs) return fail(c, -1, "lb_separate_vcs: no such VCS level");
    if (mode < 0 || mode > 2 || nadd < 0 || !out_i || !out_R || !out_score || !count || !stats)
        return fail(c, -1, "lb_separate_vcs: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    vsep_fn f = pick_vsep(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_separate_vcs: unsupported rank / dtype");
    std::vector<int64_t> oi;
    std::vector<int32_t> oR;
    std::vector<double> os;
    int rc = f(c, m, tol, mode, nadd, oi, oR, os, stats);
    if (rc) return rc;
    const int64_t k = (int64_t)oi.size();
    std::memcpy(out_i, oi.data(), sizeof(int64_t) * k);
    std::memcpy(out_R, oR.data(), sizeof(int32_t) * k);
    std::memcpy(out_score, os.data(), sizeof(double) * k);
    *count = k;
    return 0;
}

int lb_sync(lb_ctx *c) {
    if (!c) return -1;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_time_eval(lb_ctx *c, double rho, double rt, int iters, double *ms_per_eval, double *kernel_ms) {
    if (c && c->bG) return fail(c, -1, "lb_time_eval: not available on a batch context");
    if (!c || !c->n || iters <= 0) return fail(c, -1, "lb_time_eval: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    eval_fn f = pick_eval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_time_eval: unsupported rank / dtype");
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    HIPCHK(c, hipEventRecord(e0, c->stream));
    for (int it = 0; it < iters; it++) {
        int rc = f(c, rho, rt, nullptr);
        if (rc) return rc;
    }
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(c, hipEventElapsedTime(&ms, e0, e1));
    if (ms_per_eval) *ms_per_eval = ms / iters;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if (kernel_ms) {
        hipEvent_t ev[LB_NKERNELS + 1];
        for (int k = 0; k <= LB_NKERNELS; k++) HIPCHK(c, hipEventCreate(&ev[k]));
        for (int k = 0; k < LB_NKERNELS; k++) kernel_ms[k] = 0.0;
        for (int it = 0; it < iters; it++) {
            int rc = f(c, rho, rt, ev);
            if (rc) return rc;
            HIPCHK(c, hipEventSynchronize(ev[LB_NKERNELS]));
            for (int k = 0; k < LB_NKERNELS; k++) {
                float t = 0.f;
                HIPCHK(c, hipEventElapsedTime(&t, ev[k], ev[k + 1]));
                kernel_ms[k] += t / iters;
            }
        }
        for (int k = 0; k <= LB_NKERNELS; k++) hipEventDestroy(ev[k]);
    }
    c->have_eval = true; c->last_rho = rho;
    return 0;
}

}  // extern "C"


// ================================================================================================ device eigensolver blocks
// Building blocks of a LOBPCG whose n-sized data stay on the device (the candidate bound's pencil eigenproblem,
// est_dual_vcs.py:325-364): blocks of n x r FP64 in the context's row order, the operator M X = Diag(g) X - C(y) X
// through the evaluation at rho = 0 (FP64 contexts), Gram matrices A^T Diag(w) B, block combinations, residuals and
// diagonal scalings. The small Rayleigh-Ritz problems are solved by the caller.
namespace {

#define LB_BLK_ROWBLK 128      // row blocks of the Gram kernel
#define LB_BLK_CH 64           // rows staged per step

struct GramArgs {
    const double *A[LB_BLK_MAXS]; const double *B[LB_BLK_MAXS]; int na, nb, r; const double *w; int64_t n;
    double *part;              // per row block: (na r) x (nb r), row-major
};

// one 16 x 16 tile of the Gram matrix per workgroup (blockIdx.y), over a range of rows (blockIdx.x); the tile's A and
// B columns of LB_BLK_CH rows are staged in LDS, each thread owns one entry
__global__ __launch_bounds__(256) void lb_k_blk_gram(GramArgs a) {
    __shared__ double sa[LB_BLK_CH][16], sb[LB_BLK_CH][16];
    const int wa = a.na * a.r, wb = a.nb * a.r;
    const int ntj = (wb + 15) / 16;
    const int ti = blockIdx.y / ntj, tj = blockIdx.y % ntj;
    const int ii = threadIdx.x / 16, jj = threadIdx.x % 16;
    const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per, r1 = r0 + per < a.n ? r0 + per : a.n;
    double acc = 0.0;
    for (int64_t c0 = r0; c0 < r1; c0 += LB_BLK_CH) {
        for (int e = threadIdx.x; e < LB_BLK_CH * 16; e += 256) {
            const int rr = e / 16, cc = e % 16;
            const int64_t row = c0 + rr;
            const int ca = ti * 16 + cc, cb = tj * 16 + cc;
            double va = 0.0, vb = 0.0;
            if (row < r1 && ca < wa) va = a.A[ca / a.r][row * a.r + ca % a.r];
            if (row < r1 && cb < wb) {
                vb = a.B[cb / a.r][row * a.r + cb % a.r];
                if (a.w) vb *= a.w[row];
            }
            sa[rr][cc] = va;
            sb[rr][cc] = vb;
        }
        __syncthreads();
#pragma unroll 8
        for (int rr = 0; rr < LB_BLK_CH; rr++) acc += sa[rr][ii] * sb[rr][jj];
        __syncthreads();
    }
    const int gi = ti * 16 + ii, gj = tj * 16 + jj;
    if (gi < wa && gj < wb) a.part[(int64_t)blockIdx.x * wa * wb + (int64_t)gi * wb + gj] = acc;
}

struct CombArgs {
    const double *S[LB_BLK_MAXS]; int ns, r; int64_t n; const double *coef;   // coef: ns blocks of r x r (row-major)
    double *dst; int accumulate;
};

// dst[row, col] (+)= sum_j sum_k S_j[row, k] coef_j[k, col]; the coefficients are staged in LDS
__global__ __launch_bounds__(256) void lb_k_blk_comb(CombArgs a) {
    extern __shared__ double sc[];
    const int nc = a.ns * a.r * a.r;
    for (int e = threadIdx.x; e < nc; e += 256) sc[e] = a.coef[e];
    __syncthreads();
    const int64_t cnt = a.n * a.r;
    for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * 256) {
        const int64_t row = e / a.r;
        const int col = (int)(e % a.r);
        double v = a.accumulate ? a.dst[e] : 0.0;
        for (int j = 0; j < a.ns; j++) {
            const double *src = a.S[j] + row * a.r;
            const double *cf = sc + j * a.r * a.r + col;
            for (int k = 0; k < a.r; k++) v += src[k] * cf[k * a.r];
        }
        a.dst[e] = v;
    }
}

struct ThetaArg { double t[32]; };

// kind 0: dst = w o src - G (operator finish); 1: dst = w o src; 2: dst = ax - w o x o theta (columns)
__global__ void lb_k_blk_elem(int kind, const double *src, const double *G, const double *w, int64_t n, int r,
                              ThetaArg th, double *dst) {
    const int64_t cnt = n * r;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t row = e / r;
        const double wi = w ? w[row] : 0.0;
        if (kind == 0) dst[e] = wi * src[e] - G[e];
        else if (kind == 1) dst[e] = wi * src[e];
        else dst[e] = G[e] - wi * src[e] * th.t[e % r];
    }
}

double *blk_ptr(lb_ctx *c, int k) { return (double *)c->blk.p + (size_t)k * c->n * c->r; }
double *diag_ptr(lb_ctx *c, int k) { return k < 0 ? nullptr : (double *)c->blkdiag.p + (size_t)k * c->n; }

int blk_check(lb_ctx *c, const char *what) {
    if (!c || !c->n) return fail(c, -1, std::string(what) + ": no graph");
    if (c->bG) return fail(c, -1, std::string(what) + ": not available on a batch context");
    if (!c->nblk) return fail(c, -1, std::string(what) + ": call lb_blk_alloc first");
    return 0;
}

}  // namespace

extern "C" {

int lb_blk_alloc(lb_ctx *c, int nblk, int ndiag) {
    drop_capture(c);
    if (!c || !c->n) return fail(c, -1, "lb_blk_alloc: no graph");
    if (c->vdtype != LB_F64) return fail(c, -1, "lb_blk_alloc: the block operator needs an FP64 context");
    if (nblk < 1 || nblk > 64 || ndiag < 0 || ndiag > 16) return fail(c, -1, "lb_blk_alloc: 1..64 blocks, 0..16 diagonals");
    HIPCHK(c, hipSetDevice(c->device));
    const size_t nr = (size_t)c->n * c->r;
    int rc;
    if ((rc = dalloc(c, c->blk, sizeof(double) * nr * nblk)) ||
        (rc = dalloc(c, c->blkdiag, sizeof(double) * (size_t)c->n * std::max(ndiag, 1))))
        return rc;
    c->nblk = nblk; c->ndiag = ndiag;
    return 0;
}

int lb_blk_set(lb_ctx *c, int k, const double *host) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_set"))) return rc;
    if (k < 0 || k >= c->nblk || !host) return fail(c, -1, "lb_blk_set: bad block or data");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r;
    const size_t bytes = sizeof(double) * (size_t)n * r;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    par_rows(n, [&](int64_t a, int64_t b) { rows_to_storage(a, b, r, host, LB_F64, st, ord); });
    HIPCHK(c, hipMemcpyAsync(blk_ptr(c, k), st, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_blk_get(lb_ctx *c, int k, double *host) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_get"))) return rc;
    if (k < 0 || k >= c->nblk || !host) return fail(c, -1, "lb_blk_get: bad block or output");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r;
    const size_t bytes = sizeof(double) * (size_t)n * r;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    HIPCHK(c, hipMemcpyAsync(st, blk_ptr(c, k), bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    par_rows(n, [&](int64_t a, int64_t b) { storage_to_rows(a, b, r, st, LB_F64, host, ord); });
    return 0;
}

int lb_blk_set_diag(lb_ctx *c, int k, const double *host) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_set_diag"))) return rc;
    if (k < 0 || k >= c->ndiag || !host) return fail(c, -1, "lb_blk_set_diag: bad diagonal or data");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    std::vector<double> h((size_t)n);
    if (c->row_order.empty()) std::copy(host, host + n, h.begin());
    else for (int64_t i = 0; i < n; i++) h[(size_t)i] = host[c->row_order[(size_t)i]];
    // on the context's stream, after any block work still queued there, and complete before h goes away
    HIPCHK(c, hipMemcpyAsync(diag_ptr(c, k), h.data(), sizeof(double) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_blk_op(lb_ctx *c, int dst, int src, int diag) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_op"))) return rc;
    if (dst < 0 || dst >= c->nblk || src < 0 || src >= c->nblk || dst == src || diag >= c->ndiag)
        return fail(c, -1, "lb_blk_op: bad block or diagonal index");
    HIPCHK(c, hipSetDevice(c->device));
    eval_fn f = pick_eval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_blk_op: unsupported rank / dtype");
    const int64_t n = c->n;
    const int r = c->r;
    HIPCHK(c, hipMemcpyAsync(c->V.p, blk_ptr(c, src), sizeof(double) * (size_t)n * r, hipMemcpyDeviceToDevice, c->stream));
    if ((rc = f(c, 0.0, 0.0, nullptr))) return rc;
    c->have_eval = true; c->last_rho = 0.0;
    ThetaArg th{};
    hipLaunchKernelGGL(lb_k_blk_elem, dim3(2048), dim3(256), 0, c->stream, 0, (const double *)blk_ptr(c, src),
                       (const double *)c->G.p, (const double *)diag_ptr(c, diag), n, r, th, blk_ptr(c, dst));
    HIPCHK(c, hipGetLastError());
    return 0;
}

int lb_blk_elem(lb_ctx *c, int kind, int dst, int src, int other, int diag, const double *theta) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_elem"))) return rc;
    if (kind < 1 || kind > 2 || dst < 0 || dst >= c->nblk || src < 0 || src >= c->nblk || diag < 0 || diag >= c->ndiag ||
        (kind == 2 && (other < 0 || other >= c->nblk || !theta)))
        return fail(c, -1, "lb_blk_elem: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    ThetaArg th{};
    if (kind == 2) for (int j = 0; j < c->r; j++) th.t[j] = theta[j];
    hipLaunchKernelGGL(lb_k_blk_elem, dim3(2048), dim3(256), 0, c->stream, kind, (const double *)blk_ptr(c, src),
                       kind == 2 ? (const double *)blk_ptr(c, other) : nullptr, (const double *)diag_ptr(c, diag), c->n,
                       c->r, th, blk_ptr(c, dst));
    HIPCHK(c, hipGetLastError());
    return 0;
}

int lb_blk_gram(lb_ctx *c, int na, const int *a, int nb, const int *b, int diag, double *out) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_gram"))) return rc;
    if (na < 1 || na > LB_BLK_MAXS || nb < 1 || nb > LB_BLK_MAXS || !a || !b || !out || diag >= c->ndiag)
        return fail(c, -1, "lb_blk_gram: bad arguments");
    for (int j = 0; j < na; j++) if (a[j] < 0 || a[j] >= c->nblk) return fail(c, -1, "lb_blk_gram: bad block index");
    for (int j = 0; j < nb; j++) if (b[j] < 0 || b[j] >= c->nblk) return fail(c, -1, "lb_blk_gram: bad block index");
    HIPCHK(c, hipSetDevice(c->device));
    GramArgs g{};
    for (int j = 0; j < na; j++) g.A[j] = blk_ptr(c, a[j]);
    for (int j = 0; j < nb; j++) g.B[j] = blk_ptr(c, b[j]);
    g.na = na; g.nb = nb; g.r = c->r; g.w = diag_ptr(c, diag); g.n = c->n;
    const int wa = na * c->r, wb = nb * c->r, m = wa * wb;
    if ((rc = dalloc(c, c->blkpart, sizeof(double) * (size_t)LB_BLK_ROWBLK * m)) ||
        (rc = dalloc(c, c->blkout, sizeof(double) * (size_t)m)))
        return rc;
    g.part = (double *)c->blkpart.p;
    const int tiles = ((wa + 15) / 16) * ((wb + 15) / 16);
    hipLaunchKernelGGL(lb_k_blk_gram, dim3(LB_BLK_ROWBLK, tiles), dim3(256), 0, c->stream, g);
    hipLaunchKernelGGL(lb_k_reduce, dim3((m + 63) / 64), dim3(1024), 0, c->stream, (const double *)c->blkpart.p,
                       LB_BLK_ROWBLK, m, m, (double *)c->blkout.p);
    HIPCHK(c, hipMemcpyAsync(out, c->blkout.p, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_blk_comb(lb_ctx *c, int dst, int ns, const int *src, const double *coef, int accumulate) {
    int rc;
    if ((rc = blk_check(c, "lb_blk_comb"))) return rc;
    if (dst < 0 || dst >= c->nblk || ns < 1 || ns > LB_BLK_MAXS || !src || !coef) return fail(c, -1, "lb_blk_comb: bad arguments");
    for (int j = 0; j < ns; j++)
        if (src[j] < 0 || src[j] >= c->nblk || src[j] == dst) return fail(c, -1, "lb_blk_comb: bad source (or the destination)");
    HIPCHK(c, hipSetDevice(c->device));
    const int r = c->r;
    const size_t nc = (size_t)ns * r * r;
    if ((rc = dalloc(c, c->blkcoef, sizeof(double) * nc))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->blkcoef.p, coef, sizeof(double) * nc, hipMemcpyHostToDevice, c->stream));
    CombArgs ca{};
    for (int j = 0; j < ns; j++) ca.S[j] = blk_ptr(c, src[j]);
    ca.ns = ns; ca.r = r; ca.n = c->n; ca.coef = (const double *)c->blkcoef.p; ca.dst = blk_ptr(c, dst);
    ca.accumulate = accumulate ? 1 : 0;
    hipLaunchKernelGGL(lb_k_blk_comb, dim3(2048), dim3(256), sizeof(double) * nc, c->stream, ca);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));     // the coefficients are the caller's memory
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------------------------- Gram without allocation
// lb_gram_noalloc keeps the Gram kernel's per-row-block partial sums in this file-scope device array instead of a
// context buffer: each device holds its own copy, alive until the process ends. In host code the name is a host-side
// stand-in, never device memory: the device address comes from hipGetSymbolAddress, looked up once per device with
// that device current, and only that address is passed to the kernels.
#define LB_GRAM_MAXW 96                                      // na r and nb r at most
#define LB_GRAM_SCRATCH_DOUBLES ((size_t)LB_BLK_ROWBLK * LB_GRAM_MAXW * LB_GRAM_MAXW)   // 9.4 MB
static __device__ __align__(256) double lb_gram_scratch[LB_GRAM_SCRATCH_DOUBLES];

namespace {

#define LB_MAX_DEVICES 64
std::mutex g_gram_mu;
double *g_gram_ptr[LB_MAX_DEVICES] = {};

// device address of lb_gram_scratch on the current device `device` (the first call per device looks it up)
hipError_t gram_scratch(int device, double **out) {
    std::lock_guard<std::mutex> lk(g_gram_mu);
    if (!g_gram_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(lb_gram_scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_gram_ptr[device] = (double *)p;
    }
    *out = g_gram_ptr[device];
    return hipSuccess;
}

}  // namespace

extern "C" {

int lb_gram_noalloc(int device, void *stream, int64_t n, int r, int na, const double *const *A, int nb,
                    const double *const *B, const double *w, double *out) {
    if (device < 0 || device >= LB_MAX_DEVICES || n < 1 || (r != 8 && r != 16 && r != 32) || na < 1 ||
        na > LB_BLK_MAXS || nb < 1 || nb > LB_BLK_MAXS || !A || !B || !out)
        return -1;
    for (int j = 0; j < na; j++) if (!A[j]) return -1;
    for (int j = 0; j < nb; j++) if (!B[j]) return -1;
    if (na * r > LB_GRAM_MAXW || nb * r > LB_GRAM_MAXW) return -2;        // the partials would not fit the array
    int ndev = 0, prev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) return -3;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return -3;
    int rc = 0;
    double *part = nullptr;
    if (gram_scratch(device, &part) != hipSuccess) {
        rc = -3;
    } else {
        GramArgs g{};
        for (int j = 0; j < na; j++) g.A[j] = A[j];
        for (int j = 0; j < nb; j++) g.B[j] = B[j];
        g.na = na; g.nb = nb; g.r = r; g.w = w; g.n = n; g.part = part;
        const int wa = na * r, wb = nb * r, m = wa * wb;
        const int tiles = ((wa + 15) / 16) * ((wb + 15) / 16);
        const hipStream_t st = (hipStream_t)stream;
        // every row block writes every entry of its partial tile (zero for an empty row range), so the array needs
        // no clearing between calls
        hipLaunchKernelGGL(lb_k_blk_gram, dim3(LB_BLK_ROWBLK, tiles), dim3(256), 0, st, g);
        hipLaunchKernelGGL(lb_k_reduce, dim3((m + 63) / 64), dim3(1024), 0, st, (const double *)part, LB_BLK_ROWBLK,
                           m, m, out);
        if (hipGetLastError() != hipSuccess) rc = -3;
    }
    if (hipSetDevice(prev) != hipSuccess && rc == 0) rc = -3;
    return rc;
}

}  // extern "C"

// ================================================================================================ exact reduction
// reduce_exact.py:12-62 on the device. The groups are the connected components of the eligible edges (leaf edges,
// and adjacent true twins with d^2 < 2m, d > 1), each numbered by the rank of its smallest vertex: the reference's
// union-find links par[max root] = min root, so its roots are the component minima, and np.unique(root) numbers them
// in increasing order. The same components come out of a lock-free union-find that only ever hooks a larger root
// under a smaller one (atomic compare-and-swap), so the result does not depend on the order of the unions. The twin
// test compares the sorted adjacency rows exactly; the hash only filters candidates (a true twin pair always passes
// it), so the hash weights do not change the result.
namespace {

__global__ void lb_k_red_deg(const int64_t *keys, int64_t m, int64_t n, unsigned long long *deg, int64_t *adjkey) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = keys[e] / n, v = keys[e] % n;
        atomicAdd(&deg[u], 1ull);
        atomicAdd(&deg[v], 1ull);
        adjkey[2 * e] = u * n + v;                  // both directions; a radix sort orders them by (row, column)
        adjkey[2 * e + 1] = v * n + u;
    }
}

__global__ void lb_k_red_hash(const int64_t *ptr, const int64_t *adjkey, int64_t n, const uint64_t *rh, uint64_t *h) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        uint64_t s = rh[x];
        for (int64_t k = ptr[x]; k < ptr[x + 1]; k++) s += rh[adjkey[k] % n];
        h[x] = s;
    }
}

__global__ void lb_k_red_splitmix(uint64_t *rh, int64_t n) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = (uint64_t)x + 0x9E3779B97F4A7C15ull;   // splitmix64 (Steele, Lea, Flood 2014)
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        rh[x] = (z ^ (z >> 31)) | 1ull;
    }
}

// flag[e]: 1 = leaf edge, 2 = verified twin pair, 0 = neither; cnt[0] += leaves, cnt[1] += twins
__global__ void lb_k_red_edges(const int64_t *keys, int64_t m, int64_t n, const unsigned long long *deg,
                               const int64_t *ptr, const int64_t *adjkey, const uint64_t *h, int8_t *flag,
                               unsigned long long *cnt) {
    unsigned long long nl = 0, nt = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t u = keys[e] / n, v = keys[e] % n;
        const int64_t du = (int64_t)deg[u], dv = (int64_t)deg[v];
        int8_t f = 0;
        if (du == 1 || dv == 1) {
            f = 1;
            nl++;
        } else if (h[u] == h[v] && du == dv && (double)du * (double)du < 2.0 * (double)m && du > 1) {
            // N[u] == N[v]: the sorted rows agree once v is dropped from row u and u from row v
            int64_t a = ptr[u], ae = ptr[u + 1], b = ptr[v], be = ptr[v + 1];
            bool same = true;
            while (true) {
                if (a < ae && adjkey[a] % n == v) a++;
                if (b < be && adjkey[b] % n == u) b++;
                if (a == ae || b == be) { same = (a == ae && b == be); break; }
                if (adjkey[a] % n != adjkey[b] % n) { same = false; break; }
                a++;
                b++;
            }
            if (same) { f = 2; nt++; }
        }
        flag[e] = f;
    }
    if (nl) atomicAdd(&cnt[0], nl);
    if (nt) atomicAdd(&cnt[1], nt);
}

__device__ __forceinline__ int64_t red_load(const int64_t *p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// root of x: parents strictly decrease along a path, so the walk ends within `limit` steps unless the forest is
// broken, which is reported through bad[0]
__device__ __forceinline__ int64_t red_find(const int64_t *par, int64_t x, int64_t limit, unsigned int *bad) {
    for (int64_t s = 0; s <= limit; s++) {
        const int64_t p = red_load(par + x);
        if (p == x) return x;
        x = p;
    }
    atomicOr(bad, 1u);
    return x;
}

__global__ void lb_k_red_union(const int64_t *keys, int64_t m, int64_t n, const int8_t *flag, int64_t *par,
                               unsigned int *bad) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        if (!flag[e]) continue;
        const int64_t a = keys[e] / n, b = keys[e] % n;
        // each failed compare-and-swap means another thread hooked that root: the loop is lock-free; the bound
        // only guards against a broken forest
        for (int64_t tries = 0; tries <= n; tries++) {
            const int64_t ra = red_find(par, a, n, bad), rb = red_find(par, b, n, bad);
            if (ra == rb || *bad) break;
            const int64_t lo = ra < rb ? ra : rb, hi = ra < rb ? rb : ra;
            unsigned long long *slot = (unsigned long long *)(par + hi);
            if (atomicCAS(slot, (unsigned long long)hi, (unsigned long long)lo) == (unsigned long long)hi) break;
            if (tries == n) atomicOr(bad, 1u);
        }
    }
}

__global__ void lb_k_red_roots(const int64_t *par, int64_t n, int64_t *root, int64_t *isroot, unsigned int *bad) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t r = red_find(par, x, n, bad);
        root[x] = r;
        isroot[x] = (r == x) ? 1 : 0;
    }
}

__global__ void lb_k_red_groups(const int64_t *root, const int64_t *rank, const unsigned long long *deg, int64_t n,
                                int64_t *grp, unsigned long long *vol) {
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (int64_t)gridDim.x * blockDim.x) {
        const int64_t g = rank[root[x]];
        grp[x] = g;
        if (deg[x]) atomicAdd(&vol[g], deg[x]);
    }
}

// group-pair key of each cross edge; inside edges get INT64_MAX (sorted to the end) and are counted
__global__ void lb_k_red_pairs(const int64_t *keys, int64_t m, int64_t n, const int64_t *grp, int64_t N,
                               int64_t *pk, unsigned long long *cnt) {
    unsigned long long ins = 0;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < m; e += (int64_t)gridDim.x * blockDim.x) {
        const int64_t gu = grp[keys[e] / n], gv = grp[keys[e] % n];
        if (gu == gv) { pk[e] = INT64_MAX; ins++; }
        else pk[e] = (gu < gv ? gu : gv) * N + (gu < gv ? gv : gu);
    }
    if (ins) atomicAdd(&cnt[2], ins);
}

__global__ void lb_k_red_counts(const int *rl, int64_t k, double *w) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < k; t += (int64_t)gridDim.x * blockDim.x)
        w[t] = (double)rl[t];
}

int bits_for(uint64_t x) {
    int b = 1;
    while (b < 64 && (x >> b)) b++;
    return b;
}

}  // namespace

extern "C" {

int lb_reduce_exact(lb_ctx *c, int64_t n, int64_t m, const int64_t *keys, const uint64_t *rhash, int64_t *group_of,
                    int64_t *vol, int64_t *keys_out, double *w_out, int64_t *stats) {
    if (!c) return -1;
    if (n < 1 || m < 0 || (m && !keys) || !group_of || !vol || !stats || (m && (!keys_out || !w_out)))
        return fail(c, -1, "lb_reduce_exact: bad arguments");
    if ((double)n * (double)n >= 9.2e18 || 2 * m > (int64_t)INT32_MAX)
        return fail(c, -1, "lb_reduce_exact: graph too large for 64-bit pair keys / 32-bit sort counts");
    for (int64_t e = 0; e < m; e++) {                    // keys must be sorted, unique, u < v, in range
        const int64_t u = keys[e] / n, v = keys[e] % n;
        if (keys[e] < 0 || u >= n || u >= v || (e && keys[e] <= keys[e - 1]))
            return fail(c, -1, "lb_reduce_exact: keys must be sorted unique u*n+v with u < v < n");
    }
    HIPCHK(c, hipSetDevice(c->device));
    DBuf dkeys, deg, adjkey, adjkey2, ptr, rh, h, flag, par, root, isroot, rank, grp, dvol, pk, pk2, uk, rl, nrun, dw,
        cnt, tmp;
    auto cleanup = [&]() {
        DBuf *bs[] = {&dkeys, &deg, &adjkey, &adjkey2, &ptr, &rh, &h, &flag, &par, &root, &isroot, &rank, &grp, &dvol,
                      &pk, &pk2, &uk, &rl, &nrun, &dw, &cnt, &tmp};
        for (DBuf *b : bs) dfree(*b);
    };
    int rc;
#define RCHK(x) do { if ((rc = (x))) { cleanup(); return rc; } } while (0)
#define RHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return fail(c, -3, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
    hipStream_t s = c->stream;
    const int G = 2048, B = 256;
    const size_t M2 = (size_t)std::max<int64_t>(2 * m, 1);
    RCHK(upload(c, dkeys, keys, sizeof(int64_t) * (size_t)m));
    RCHK(dalloc(c, deg, sizeof(uint64_t) * (size_t)(n + 1)));
    RCHK(dalloc(c, adjkey, sizeof(int64_t) * M2));
    RCHK(dalloc(c, adjkey2, sizeof(int64_t) * M2));
    RCHK(dalloc(c, ptr, sizeof(int64_t) * (size_t)(n + 1)));
    RCHK(dalloc(c, cnt, 64));
    RHIP(hipMemsetAsync(deg.p, 0, sizeof(uint64_t) * (size_t)(n + 1), s));
    RHIP(hipMemsetAsync(cnt.p, 0, 64, s));
    if (m) hipLaunchKernelGGL(lb_k_red_deg, dim3(G), dim3(B), 0, s, (const int64_t *)dkeys.p, m, n, (unsigned long long *)deg.p, (int64_t *)adjkey.p);
    RHIP(hipGetLastError());
    size_t tb = 0, tb2 = 0;
    const int endbit = bits_for((uint64_t)n * (uint64_t)n);
    // adjacency sorted by (row, column): one radix sort of the 2m directed keys
    RHIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, (const int64_t *)adjkey.p, (int64_t *)adjkey2.p, (int)(2 * m), 0, endbit, s));
    RHIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, (const unsigned long long *)deg.p, (unsigned long long *)ptr.p, (int)(n + 1), s));
    RCHK(dalloc(c, tmp, std::max(tb, tb2)));
    if (m) RHIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, (const int64_t *)adjkey.p, (int64_t *)adjkey2.p, (int)(2 * m), 0, endbit, s));
    RHIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb2, (const unsigned long long *)deg.p, (unsigned long long *)ptr.p, (int)(n + 1), s));
    // hash filter and the per-edge rules
    RCHK(dalloc(c, rh, sizeof(uint64_t) * (size_t)n));
    RCHK(dalloc(c, h, sizeof(uint64_t) * (size_t)n));
    if (rhash) RHIP(hipMemcpyAsync(rh.p, rhash, sizeof(uint64_t) * (size_t)n, hipMemcpyHostToDevice, s));
    else hipLaunchKernelGGL(lb_k_red_splitmix, dim3(G), dim3(B), 0, s, (uint64_t *)rh.p, n);
    hipLaunchKernelGGL(lb_k_red_hash, dim3(G), dim3(B), 0, s, (const int64_t *)ptr.p, (const int64_t *)adjkey2.p, n, (const uint64_t *)rh.p, (uint64_t *)h.p);
    RCHK(dalloc(c, flag, (size_t)std::max<int64_t>(m, 1)));
    if (m) hipLaunchKernelGGL(lb_k_red_edges, dim3(G), dim3(B), 0, s, (const int64_t *)dkeys.p, m, n, (const unsigned long long *)deg.p,
                              (const int64_t *)ptr.p, (const int64_t *)adjkey2.p, (const uint64_t *)h.p, (int8_t *)flag.p, (unsigned long long *)cnt.p);
    RHIP(hipGetLastError());
    // components: parents start as the identity
    RCHK(dalloc(c, par, sizeof(int64_t) * (size_t)n));
    RCHK(dalloc(c, root, sizeof(int64_t) * (size_t)n));
    RCHK(dalloc(c, isroot, sizeof(int64_t) * (size_t)(n + 1)));
    RCHK(dalloc(c, rank, sizeof(int64_t) * (size_t)(n + 1)));
    hipLaunchKernelGGL(lb_k_iota64, dim3(G), dim3(B), 0, s, (int64_t *)par.p, n);
    unsigned int *bad = (unsigned int *)((char *)cnt.p + 32);
    if (m) hipLaunchKernelGGL(lb_k_red_union, dim3(G), dim3(B), 0, s, (const int64_t *)dkeys.p, m, n, (const int8_t *)flag.p, (int64_t *)par.p, bad);
    RHIP(hipMemsetAsync(isroot.p, 0, sizeof(int64_t) * (size_t)(n + 1), s));
    hipLaunchKernelGGL(lb_k_red_roots, dim3(G), dim3(B), 0, s, (const int64_t *)par.p, n, (int64_t *)root.p, (int64_t *)isroot.p, bad);
    RHIP(hipGetLastError());
    size_t tb3 = 0;
    RHIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tb3, (const int64_t *)isroot.p, (int64_t *)rank.p, (int)(n + 1), s));
    RCHK(dalloc(c, tmp, tb3));
    RHIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb3, (const int64_t *)isroot.p, (int64_t *)rank.p, (int)(n + 1), s));
    int64_t N = 0;
    unsigned int hbad = 0;
    RHIP(hipMemcpyAsync(&N, (const int64_t *)rank.p + n, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    RHIP(hipMemcpyAsync(&hbad, bad, sizeof(unsigned int), hipMemcpyDeviceToHost, s));
    RHIP(hipStreamSynchronize(s));
    if (hbad || N < 1 || N > n) { cleanup(); return fail(c, 1, "lb_reduce_exact: union-find did not converge"); }
    RCHK(dalloc(c, grp, sizeof(int64_t) * (size_t)n));
    RCHK(dalloc(c, dvol, sizeof(uint64_t) * (size_t)N));
    RHIP(hipMemsetAsync(dvol.p, 0, sizeof(uint64_t) * (size_t)N, s));
    hipLaunchKernelGGL(lb_k_red_groups, dim3(G), dim3(B), 0, s, (const int64_t *)root.p, (const int64_t *)rank.p, (const unsigned long long *)deg.p, n,
                       (int64_t *)grp.p, (unsigned long long *)dvol.p);
    RHIP(hipGetLastError());
    // cross-group edges -> sorted unique group-pair keys with counts
    int64_t nk = 0;
    unsigned long long hc[3] = {0, 0, 0};
    if (m) {
        RCHK(dalloc(c, pk, sizeof(int64_t) * (size_t)m));
        RCHK(dalloc(c, pk2, sizeof(int64_t) * (size_t)m));
        RCHK(dalloc(c, uk, sizeof(int64_t) * (size_t)m));
        RCHK(dalloc(c, rl, sizeof(int) * (size_t)m));
        RCHK(dalloc(c, nrun, 16));
        RCHK(dalloc(c, dw, sizeof(double) * (size_t)m));
        hipLaunchKernelGGL(lb_k_red_pairs, dim3(G), dim3(B), 0, s, (const int64_t *)dkeys.p, m, n, (const int64_t *)grp.p, N, (int64_t *)pk.p,
                           (unsigned long long *)cnt.p);
        RHIP(hipGetLastError());
        size_t t4 = 0, t5 = 0;
        RHIP(hipcub::DeviceRadixSort::SortKeys(nullptr, t4, (const int64_t *)pk.p, (int64_t *)pk2.p, (int)m, 0, 64, s));
        RHIP(hipcub::DeviceRunLengthEncode::Encode(nullptr, t5, (const int64_t *)pk2.p, (int64_t *)uk.p, (int *)rl.p, (int *)nrun.p, (int)m, s));
        RCHK(dalloc(c, tmp, std::max(t4, t5)));
        RHIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, t4, (const int64_t *)pk.p, (int64_t *)pk2.p, (int)m, 0, 64, s));
        RHIP(hipMemcpyAsync(hc, cnt.p, sizeof(hc), hipMemcpyDeviceToHost, s));
        RHIP(hipStreamSynchronize(s));
        const int64_t ncross = m - (int64_t)hc[2];
        if (ncross > 0) {
            RHIP(hipcub::DeviceRunLengthEncode::Encode(tmp.p, t5, (const int64_t *)pk2.p, (int64_t *)uk.p, (int *)rl.p, (int *)nrun.p, (int)ncross, s));
            int hn = 0;
            RHIP(hipMemcpyAsync(&hn, nrun.p, sizeof(int), hipMemcpyDeviceToHost, s));
            RHIP(hipStreamSynchronize(s));
            nk = hn;
            hipLaunchKernelGGL(lb_k_red_counts, dim3(G), dim3(B), 0, s, (const int *)rl.p, nk, (double *)dw.p);
            RHIP(hipGetLastError());
            RHIP(hipMemcpyAsync(keys_out, uk.p, sizeof(int64_t) * (size_t)nk, hipMemcpyDeviceToHost, s));
            RHIP(hipMemcpyAsync(w_out, dw.p, sizeof(double) * (size_t)nk, hipMemcpyDeviceToHost, s));
        }
    }
    RHIP(hipMemcpyAsync(group_of, grp.p, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, s));
    RHIP(hipMemcpyAsync(vol, dvol.p, sizeof(int64_t) * (size_t)N, hipMemcpyDeviceToHost, s));
    RHIP(hipStreamSynchronize(s));
    stats[0] = N; stats[1] = nk; stats[2] = (int64_t)hc[0]; stats[3] = (int64_t)hc[1]; stats[4] = (int64_t)hc[2];
    cleanup();
#undef RCHK
#undef RHIP
    return 0;
}

}  // extern "C"

// ================================================================================================ dense certificate
// M = Diag(dg) - C(y) assembled in device memory (column-major, lower triangle), with the entrywise magnitudes
// |terms| of C summed into a second matrix for the rounding bound, then factored by rocSOLVER (dpotrf). The caller
// turns (info, sum R_ij^2, sum of the magnitudes squared) into a proved shift: R^T R = M + E with
// |E| <= gamma(n+1) |R^T||R| for any conventional Cholesky (Higham, Accuracy and Stability of Numerical Algorithms,
// Thm 10.3), so lambda_min(M) >= -(gamma(n+1) ||R||_F^2 + ||F||_F).
#include <dlfcn.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>
namespace {

// rocBLAS / rocSOLVER are opened on first use only (dlopen), so processes that never certify do not load them
struct SolverApi {
    bool ok = false; std::string err;
    rocblas_status (*create)(rocblas_handle *) = nullptr;
    rocblas_status (*destroy)(rocblas_handle) = nullptr;
    rocblas_status (*set_stream)(rocblas_handle, hipStream_t) = nullptr;
    rocblas_status (*potrf)(rocblas_handle, rocblas_fill, rocblas_int, double *, rocblas_int, rocblas_int *) = nullptr;
    // 64-bit sizes (dense cores of more than 2^31 entries); optional
    rocblas_status (*potrf64)(rocblas_handle, rocblas_fill, int64_t, double *, int64_t, int64_t *) = nullptr;
};

SolverApi load_solver_api() {
    SolverApi a;
    void *hb = dlopen("librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
    if (!hb) hb = dlopen("/opt/rocm/lib/librocblas.so.5", RTLD_NOW | RTLD_LOCAL);
    void *hs = dlopen("librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!hs) hs = dlopen("/opt/rocm/lib/librocsolver.so.0", RTLD_NOW | RTLD_LOCAL);
    if (!hb || !hs) { a.err = std::string("cannot open rocBLAS / rocSOLVER: ") + dlerror(); return a; }
    a.create = (rocblas_status(*)(rocblas_handle *))dlsym(hb, "rocblas_create_handle");
    a.destroy = (rocblas_status(*)(rocblas_handle))dlsym(hb, "rocblas_destroy_handle");
    a.set_stream = (rocblas_status(*)(rocblas_handle, hipStream_t))dlsym(hb, "rocblas_set_stream");
    a.potrf = (rocblas_status(*)(rocblas_handle, rocblas_fill, rocblas_int, double *, rocblas_int, rocblas_int *))dlsym(hs, "rocsolver_dpotrf");
    a.potrf64 = (rocblas_status(*)(rocblas_handle, rocblas_fill, int64_t, double *, int64_t, int64_t *))dlsym(hs, "rocsolver_dpotrf_64");
    a.ok = a.create && a.destroy && a.set_stream && a.potrf;
    if (!a.ok) a.err = "rocBLAS / rocSOLVER entry points not found";
    return a;
}

const SolverApi &solver_api() {
    static SolverApi a = load_solver_api();
    return a;
}

struct CertArgs {
    int64_t n; const double *d; double L; const double *dg;
    int nlev; int Poff[LB_MAXLEV]; int64_t soff[LB_MAXLEV]; const int32_t *lab[LB_MAXLEV]; const double *sig;
    int nvcs; const int32_t *vlab[LB_MAXVCS]; const int64_t *vptr[LB_MAXVCS]; const int32_t *vcol; const double *vsig;
    double *M; double *Ab;
};

__device__ __forceinline__ double vcs_lookup(const int64_t *ptr, const int32_t *col, const double *sig, int64_t row,
                                             int32_t R) {
    int64_t lo = ptr[row], hi = ptr[row + 1];
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if (col[mid] < R) lo = mid + 1; else hi = mid;
    }
    return (lo < ptr[row + 1] && col[lo] == R) ? sig[lo] : 0.0;
}

// lower triangle (i >= j) of M = Diag(dg) - (-d d^T / L + levels + VCS); the A and triangle terms are scattered next
__global__ void lb_k_cert_fill(CertArgs a) {
    const int64_t n = a.n, tot = n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % n, j = idx / n;
        if (i < j) continue;
        const double di = a.d[i], dj = a.d[j];
        if (i == j) {
            a.M[idx] = a.dg[i] + di * di / a.L;
            a.Ab[idx] = 0.0;
            continue;
        }
        const double dd = di * dj;
        double c = -dd / a.L, ab = dd / a.L;
        for (int l = 0; l < a.nlev; l++) {
            const double s = dd * a.sig[a.soff[l] + (int64_t)a.lab[l][i] * (a.Poff[l + 1] - a.Poff[l]) + a.lab[l][j]];
            c += s;
            ab += fabs(s);
        }
        for (int m = 0; m < a.nvcs; m++) {
            const double s = 0.5 * dd * (vcs_lookup(a.vptr[m], a.vcol, a.vsig, i, a.vlab[m][j]) +
                                         vcs_lookup(a.vptr[m], a.vcol, a.vsig, j, a.vlab[m][i]));
            c += s;
            ab += fabs(s);
        }
        a.M[idx] = -c;
        a.Ab[idx] = ab;
    }
}

__global__ void lb_k_cert_edges(const int64_t *rowptr, const int32_t *col, const double *w, int64_t n, double *M,
                                double *Ab) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        for (int64_t e = rowptr[i]; e < rowptr[i + 1]; e++) {
            const int64_t j = col[e];
            if (j >= i) continue;                     // each lower entry (i > j) has one writer: row i
            const double x = w ? w[e] : 1.0;
            M[j * n + i] -= x;
            Ab[j * n + i] += fabs(x);
        }
    }
}

__global__ void lb_k_cert_tris(const int32_t *TI, const double *lam, int64_t T, int64_t n, double *M, double *Ab) {
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < T; t += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = TI[3 * t], b = TI[3 * t + 1], c = TI[3 * t + 2];
        const double h = 0.5 * lam[t];
        // C_ab -= h, C_bc -= h, C_ac += h  ->  M = -C
        const int64_t p[3][2] = {{a, b}, {b, c}, {a, c}};
        const double sgn[3] = {1.0, 1.0, -1.0};
        for (int q = 0; q < 3; q++) {
            const int64_t hi = p[q][0] > p[q][1] ? p[q][0] : p[q][1], lo = p[q][0] > p[q][1] ? p[q][1] : p[q][0];
            atomicAdd(M + lo * n + hi, sgn[q] * h);
            atomicAdd(Ab + lo * n + hi, h);
        }
    }
}

// per-block partial sums of squares over the strictly lower (diag = 0) or lower-with-diagonal (diag = 1) triangle
__global__ __launch_bounds__(256) void lb_k_cert_sumsq(const double *X, int64_t n, int diag, double *part) {
    __shared__ double red[4];
    double s = 0.0;
    const int64_t tot = n * n;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < tot; idx += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = idx % n, j = idx / n;
        if (i > j || (diag && i == j)) s += X[idx] * X[idx];
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

}  // namespace

extern "C" {

int lb_certify_dense(lb_ctx *c, const double *dg, double *out) {
    if (!c) return -1;
    if (!c->n || !dg || !out) return fail(c, -1, "lb_certify_dense: bad arguments");
    if (c->vdtype != LB_F64) return fail(c, -1, "lb_certify_dense: needs an FP64 context (multipliers stored exactly)");
    const int64_t n = c->n;
    HIPCHK(c, hipSetDevice(c->device));
    size_t fr = 0, tt = 0;
    HIPCHK(c, hipMemGetInfo(&fr, &tt));
    const size_t need = 2 * sizeof(double) * (size_t)n * (size_t)n + (64u << 20);
    if (need > fr) return fail(c, -2, "lb_certify_dense: " + std::to_string(need >> 20) + " MiB needed, " +
                                      std::to_string(fr >> 20) + " MiB free");
    DBuf M, Ab, ddg, part, dinfo;
    auto cleanup = [&]() { dfree(M); dfree(Ab); dfree(ddg); dfree(part); dfree(dinfo); };
    int rc;
#define CCHK(x) do { if ((rc = (x))) { cleanup(); return rc; } } while (0)
#define CHIP(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { cleanup(); return fail(c, -3, std::string(#x) + ": " + hipGetErrorString(e_)); } } while (0)
    hipStream_t s = c->stream;
    const int G = 4096, B = 256;
    CCHK(dalloc(c, M, sizeof(double) * (size_t)n * n));
    CCHK(dalloc(c, Ab, sizeof(double) * (size_t)n * n));
    CCHK(upload(c, ddg, dg, sizeof(double) * (size_t)n));
    CCHK(dalloc(c, part, sizeof(double) * G));
    CCHK(dalloc(c, dinfo, sizeof(int)));
    CertArgs ca{};
    ca.n = n; ca.d = (const double *)c->d.p; ca.L = c->L; ca.dg = (const double *)ddg.p;
    ca.nlev = c->nlev;
    int o = 0;
    for (int l = 0; l < c->nlev; l++) { ca.Poff[l] = o; o += c->P[l]; ca.soff[l] = c->soff[l]; ca.lab[l] = (const int32_t *)c->labs.p + (int64_t)l * n; }
    ca.Poff[c->nlev] = o;
    ca.sig = (const double *)c->sig.p;
    ca.nvcs = c->nvcs;
    for (int m = 0; m < c->nvcs; m++) {
        ca.vlab[m] = (const int32_t *)c->vlabs.p + (int64_t)m * n;
        ca.vptr[m] = (const int64_t *)c->vptr.p + (int64_t)m * (n + 1);
    }
    ca.vcol = (const int32_t *)c->vcol.p; ca.vsig = (const double *)c->vsig.p;
    ca.M = (double *)M.p; ca.Ab = (double *)Ab.p;
    hipLaunchKernelGGL(lb_k_cert_fill, dim3(G), dim3(B), 0, s, ca);
    hipLaunchKernelGGL(lb_k_cert_edges, dim3(G), dim3(B), 0, s, (const int64_t *)c->rowptr.p, (const int32_t *)c->col.p,
                       c->weighted ? (const double *)c->w.p : nullptr, n, (double *)M.p, (double *)Ab.p);
    if (c->T) hipLaunchKernelGGL(lb_k_cert_tris, dim3(G), dim3(B), 0, s, (const int32_t *)c->TI.p, (const double *)c->lam.p, c->T, n,
                                 (double *)M.p, (double *)Ab.p);
    CHIP(hipGetLastError());
    std::vector<double> hp(G);
    hipLaunchKernelGGL(lb_k_cert_sumsq, dim3(G), dim3(B), 0, s, (const double *)Ab.p, n, 0, (double *)part.p);
    CHIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * G, hipMemcpyDeviceToHost, s));
    CHIP(hipStreamSynchronize(s));
    double sab = 0.0;
    for (double x : hp) sab += x;
    // factor
    const SolverApi &api = solver_api();
    if (!api.ok) { cleanup(); return fail(c, -3, "lb_certify_dense: " + api.err); }
    rocblas_handle hb = nullptr;
    if (api.create(&hb) != rocblas_status_success) { cleanup(); return fail(c, -3, "rocblas_create_handle failed"); }
    api.set_stream(hb, s);
    const rocblas_status st = api.potrf(hb, rocblas_fill_lower, (rocblas_int)n, (double *)M.p, (rocblas_int)n, (rocblas_int *)dinfo.p);
    int info = -1;
    hipError_t e1 = hipMemcpyAsync(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, s);
    hipError_t e2 = hipStreamSynchronize(s);
    api.destroy(hb);
    if (st != rocblas_status_success) { cleanup(); return fail(c, -3, "rocsolver_dpotrf failed with status " + std::to_string((int)st)); }
    CHIP(e1);
    CHIP(e2);
    double sr = 0.0;
    if (info == 0) {
        hipLaunchKernelGGL(lb_k_cert_sumsq, dim3(G), dim3(B), 0, s, (const double *)M.p, n, 1, (double *)part.p);
        CHIP(hipMemcpyAsync(hp.data(), part.p, sizeof(double) * G, hipMemcpyDeviceToHost, s));
        CHIP(hipStreamSynchronize(s));
        for (double x : hp) sr += x;
    }
    out[0] = (double)info; out[1] = sr; out[2] = 2.0 * sab;
    cleanup();
#undef CCHK
#undef CHIP
    return 0;
}

}  // extern "C"

// Dense core of the sparse certificate: S = A22 - sum_k d_k c_k c_k^T (lower triangle, column-major m x m), where c_k
// is the core part of sparse pivot k's column of L (core rows ascending), then its Cholesky factor in place.
namespace {

__global__ void lb_k_core_a(const int32_t *ai, const int32_t *aj, const double *av, int64_t na, int64_t m, double *S) {
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < na; e += (int64_t)gridDim.x * blockDim.x)
        atomicAdd(S + (int64_t)aj[e] * m + ai[e], av[e]);
}

// one workgroup per pivot column (grid-stride): every pair x >= y of its core entries, t = x (x + 1) / 2 + y
__global__ __launch_bounds__(256) void lb_k_core_upd(const int64_t *cptr, const int32_t *cidx, const double *cval,
                                                     const double *cd, int64_t nsp, int64_t m, double *S) {
    for (int64_t k = blockIdx.x; k < nsp; k += gridDim.x) {
        const int64_t b = cptr[k], c = cptr[k + 1] - b;
        if (c == 0) continue;
        const double d = cd[k];
        const int64_t npair = c * (c + 1) / 2;
        for (int64_t t = threadIdx.x; t < npair; t += blockDim.x) {
            int64_t x = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
            while (x * (x + 1) / 2 > t) x--;
            while ((x + 1) * (x + 2) / 2 <= t) x++;
            const int64_t y = t - x * (x + 1) / 2;
            const int64_t rx = cidx[b + x], ry = cidx[b + y];          // rx >= ry: the lower triangle
            atomicAdd(S + ry * m + rx, -(d * cval[b + x]) * cval[b + y]);
        }
    }
}

// colsum[j] = sum_{i >= j} |R_ij| of the lower-triangular factor (one workgroup per column, grid-stride)
__global__ __launch_bounds__(256) void lb_k_core_colsum(const double *S, int64_t m, double *colsum) {
    __shared__ double red[4];
    for (int64_t j = blockIdx.x; j < m; j += gridDim.x) {
        double a = 0.0;
        for (int64_t i = j + threadIdx.x; i < m; i += 256) a += fabs(S[j * m + i]);
        for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o);
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
        __syncthreads();
        if (threadIdx.x == 0) colsum[j] = red[0] + red[1] + red[2] + red[3];
        __syncthreads();
    }
}

// Blocked right-looking Cholesky S = R R^T in place (lower, column-major, lda m, 64-bit offsets), for cores whose
// m^2 entries exceed what the library factorization handles. Panels of 2 LB_CH columns, each factored as two halves of
// LB_CH: the diagonal block in LDS, the rows below solved against it (one thread per row, the row in registers); the
// second half is first updated by the first; then the trailing lower triangle takes one rank-2 LB_CH update on
// 128 x 128 tiles, so that it is read and written once per 128 columns. Each entry of R is a - sum of products / pivot, accumulated panel by panel, so the componentwise
// bound |S - R R^T| <= gamma(m + 1) |R| |R^T| holds as for any Cholesky. info (device, 0 on entry): 1-based index of
// the first pivot that is not positive (or not finite); every kernel returns at once once it is set.
#define LB_CH 64

__global__ __launch_bounds__(256) void lb_k_chol_diag(double *S, int64_t m, int64_t k, int64_t *info) {
    __shared__ double A[LB_CH][LB_CH + 1];
    __shared__ int bad;
    if (*info) return;
    const int b = (int)min<int64_t>(LB_CH, m - k);
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int i = e % b, j = e / b;
        A[i][j] = i >= j ? S[(k + j) * m + k + i] : 0.0;
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int j = 0; j < b; j++) {
        if (threadIdx.x == 0) {
            const double dj = A[j][j];
            if (!(dj > 0.0) || !isfinite(dj)) { bad = j + 1; A[j][j] = 1.0; }
            else A[j][j] = sqrt(dj);
        }
        __syncthreads();
        if (bad) break;
        const double piv = A[j][j];
        for (int i = j + 1 + threadIdx.x; i < b; i += 256) A[i][j] /= piv;
        __syncthreads();
        const int t = b - 1 - j;                          // trailing size
        for (int e = threadIdx.x; e < t * t; e += 256) {
            const int i = j + 1 + e % t, l = j + 1 + e / t;
            if (i >= l) A[i][l] -= A[i][j] * A[l][j];
        }
        __syncthreads();
    }
    if (bad) { if (threadIdx.x == 0) *info = k + bad; return; }
    for (int e = threadIdx.x; e < b * b; e += 256) {
        const int i = e % b, j = e / b;
        if (i >= j) S[(k + j) * m + k + i] = A[i][j];
    }
}

// rows r >= k + LB_CH: x L11^T = a (forward substitution over the panel's LB_CH columns), L11 staged in LDS
__global__ __launch_bounds__(256) void lb_k_chol_trsm(double *S, int64_t m, int64_t k, const int64_t *info) {
    __shared__ double Lt[LB_CH][LB_CH + 1];
    if (*info) return;
    for (int e = threadIdx.x; e < LB_CH * LB_CH; e += 256) {
        const int i = e % LB_CH, j = e / LB_CH;
        Lt[i][j] = i >= j ? S[(k + j) * m + k + i] : 0.0;
    }
    __syncthreads();
    const int64_t r = k + LB_CH + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    double x[LB_CH];
#pragma unroll
    for (int j = 0; j < LB_CH; j++) x[j] = S[(k + j) * m + r];
#pragma unroll
    for (int j = 0; j < LB_CH; j++) {
        double a = x[j];
#pragma unroll
        for (int l = 0; l < j; l++) a -= x[l] * Lt[j][l];
        x[j] = a / Lt[j][j];
    }
#pragma unroll
    for (int j = 0; j < LB_CH; j++) S[(k + j) * m + r] = x[j];
}

// The rank-K update C -= P_r P_c^T on 128 x 128 tiles, P = columns [kb, kb + K) (the solved panel): rows from rb,
// columns from cb up to ce (exclusive), lower triangle only (r >= c). tri = 1: tiles (I, J <= I) of the square
// trailing block, blockIdx.x = I (I + 1) / 2 + J; tri = 0: one column of tiles (J = 0), blockIdx.x = I. 256 threads,
// each owning an 8 x 8 block (rows ti + 16 u, columns tj + 16 v, so that a wavefront's LDS reads are broadcasts or
// consecutive); the panel columns are staged in LDS in chunks of LB_CK.
#define LB_CK 16
__global__ __launch_bounds__(256) void lb_k_chol_update(double *S, int64_t m, int64_t rb, int64_t cb, int64_t ce,
                                                        int64_t kb, int K, int tri, const int64_t *info) {
    __shared__ double PA[LB_CK][128 + 2], PB[LB_CK][128 + 2];
    if (*info) return;
    const int64_t t = blockIdx.x;
    int64_t I = t, J = 0;
    if (tri) {
        I = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (I * (I + 1) / 2 > t) I--;
        while ((I + 1) * (I + 2) / 2 <= t) I++;
        J = t - I * (I + 1) / 2;
    }
    const int64_t r0 = rb + I * 128, c0 = cb + J * 128;
    const int ti = threadIdx.x % 16, tj = threadIdx.x / 16;
    double acc[8][8];
#pragma unroll
    for (int u = 0; u < 8; u++)
#pragma unroll
        for (int v = 0; v < 8; v++) acc[u][v] = 0.0;
    // staging: thread x loads row x % 128 of columns x / 128 + 2 q of both panels from clamped (in-bounds) addresses,
    // so the loads carry no branches and are in flight together; rows past the matrix are zeroed after the load
    const int si = threadIdx.x & 127, sh = threadIdx.x >> 7;
    const bool ina = r0 + si < m, inb = c0 + si < ce;
    const int64_t ia = ina ? r0 + si : m - 1, ib = inb ? c0 + si : ce - 1;
    for (int kc = 0; kc < K; kc += LB_CK) {
        double xa[LB_CK / 2], xb[LB_CK / 2];
#pragma unroll
        for (int q = 0; q < LB_CK / 2; q++) {
            const double *col = S + (kb + kc + sh + 2 * q) * m;
            xa[q] = col[ia];
            xb[q] = col[ib];
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LB_CK / 2; q++) {
            PA[sh + 2 * q][si] = ina ? xa[q] : 0.0;
            PB[sh + 2 * q][si] = inb ? xb[q] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int l = 0; l < LB_CK; l++) {
            double a[8], b[8];
#pragma unroll
            for (int u = 0; u < 8; u++) { a[u] = PA[l][ti + 16 * u]; b[u] = PB[l][tj + 16 * u]; }
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int v = 0; v < 8; v++) acc[u][v] = fma(a[u], b[v], acc[u][v]);
        }
    }
    // write-back, two columns (16 entries) at a time: the loads from clamped addresses first, then the masked stores
#pragma unroll
    for (int v = 0; v < 8; v += 2) {
        double old[2][8];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int64_t c = min(c0 + tj + 16 * (v + h), ce - 1);
#pragma unroll
            for (int u = 0; u < 8; u++) old[h][u] = S[c * m + min(r0 + ti + 16 * u, m - 1)];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int64_t c = c0 + tj + 16 * (v + h);
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int64_t r = r0 + ti + 16 * u;
                if (c < ce && r < m && r >= c) S[c * m + r] = old[h][u] - acc[u][v + h];
            }
        }
    }
}

// The same rank-K update on the FP64 matrix cores (v_mfma_f64_16x16x4f64; lane l holds A[l % 16][l / 16] and
// B[l / 16][l % 16], and register v of the result is D[4 v + l / 16][l % 16], as measured on gfx950). The tile is
// computed transposed, D = P_c P_r^T (D rows = columns of C), so that the 16 lanes of a result row hold 16
// consecutive rows of one column of S and every store is a contiguous 128-byte run. Wave w of the workgroup owns the
// 64 x 64 quarter (rows 64 (w % 2), columns 64 (w / 2)) as 4 x 4 blocks of 16 x 16; per step of 4 panel columns each
// lane reads 8 operands from LDS for 16 matrix instructions. Same tiles, masks and arguments as lb_k_chol_update.
typedef double lb_d4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void lb_k_chol_update_mfma(double *S, int64_t m, int64_t rb, int64_t cb,
                                                             int64_t ce, int64_t kb, int K, int tri,
                                                             const int64_t *info) {
    __shared__ double PA[LB_CK][128 + 2], PB[LB_CK][128 + 2];
    if (*info) return;
    const int64_t t = blockIdx.x;
    int64_t I = t, J = 0;
    if (tri) {
        I = (int64_t)((sqrt(8.0 * (double)t + 1.0) - 1.0) * 0.5);
        while (I * (I + 1) / 2 > t) I--;
        while ((I + 1) * (I + 2) / 2 <= t) I++;
        J = t - I * (I + 1) / 2;
    }
    const int64_t r0 = rb + I * 128, c0 = cb + J * 128;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int wr = (w & 1) * 64, wc = (w >> 1) * 64;       // the wave's quarter: rows wr.., columns wc..
    const int li = lane & 15, lk = lane >> 4;
    lb_d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = (lb_d4){0.0, 0.0, 0.0, 0.0};
    // the next chunk of panel columns is loaded into registers while the current one is multiplied: thread x stages
    // row x % 128 of columns x / 128 + 2 q (q < LB_CK / 2) of both panels
    // (rows past the matrix read a clamped, in-bounds address and are zeroed after the load, so the loads carry no
    // branches and are all in flight together)
    const int si = threadIdx.x & 127, sh = threadIdx.x >> 7;
    const bool ina = r0 + si < m, inb = c0 + si < ce;
    const int64_t ia = ina ? r0 + si : m - 1, ib = inb ? c0 + si : ce - 1;
    double xa[LB_CK / 2], xb[LB_CK / 2];
    auto fetch = [&](int kc) {
#pragma unroll
        for (int q = 0; q < LB_CK / 2; q++) {
            const double *col = S + (kb + kc + sh + 2 * q) * m;
            xa[q] = col[ia];
            xb[q] = col[ib];
        }
#pragma unroll
        for (int q = 0; q < LB_CK / 2; q++) {
            xa[q] = ina ? xa[q] : 0.0;
            xb[q] = inb ? xb[q] : 0.0;
        }
    };
    fetch(0);
    for (int kc = 0; kc < K; kc += LB_CK) {
        __syncthreads();
#pragma unroll
        for (int q = 0; q < LB_CK / 2; q++) { PA[sh + 2 * q][si] = xa[q]; PB[sh + 2 * q][si] = xb[q]; }
        __syncthreads();
        if (kc + LB_CK < K) fetch(kc + LB_CK);
#pragma unroll
        for (int ks = 0; ks < LB_CK; ks += 4) {
            double opc[4], opr[4];
#pragma unroll
            for (int a = 0; a < 4; a++) {
                opc[a] = PB[ks + lk][wc + 16 * a + li];     // A operand: D row (a column of C) x panel column
                opr[a] = PA[ks + lk][wr + 16 * a + li];     // B operand: panel column x D column (a row of C)
            }
#pragma unroll
            for (int a = 0; a < 4; a++)
#pragma unroll
                for (int b = 0; b < 4; b++)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(opc[a], opr[b], acc[a][b], 0, 0, 0);
        }
    }
    // write-back, 16 entries at a time: all 16 loads (clamped addresses) in flight, then the masked stores
#pragma unroll
    for (int a = 0; a < 4; a++) {
        double old[4][4];
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t c = min(c0 + wc + 16 * a + 4 * v + lk, ce - 1);
#pragma unroll
            for (int b = 0; b < 4; b++) old[v][b] = S[c * m + min(r0 + wr + 16 * b + li, m - 1)];
        }
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t c = c0 + wc + 16 * a + 4 * v + lk;
#pragma unroll
            for (int b = 0; b < 4; b++) {
                const int64_t r = r0 + wr + 16 * b + li;
                if (c < ce && r < m && r >= c) S[c * m + r] = old[v][b] - acc[a][b][v];
            }
        }
    }
}

}  // namespace

extern "C" {

int lb_cert_core(int device, int64_t m, int64_t na, const int32_t *ai, const int32_t *aj, const double *av,
                 int64_t nsp, const int64_t *cptr, const int32_t *cidx, const double *cval, const double *cd,
                 int method, double *out) {
    if (m < 1 || na < 0 || nsp < 0 || !out || (na && (!ai || !aj || !av)) || (nsp && (!cptr || !cidx || !cval || !cd)) ||
        method < 0 || method > 3)
        return -1;
    for (int64_t e = 0; e < na; e++)
        if (ai[e] < 0 || ai[e] >= m || aj[e] < 0 || aj[e] > ai[e]) return -1;
    if (nsp) {
        if (cptr[0] != 0) return -1;
        for (int64_t k = 0; k < nsp; k++) {
            if (cptr[k + 1] < cptr[k]) return -1;
            for (int64_t t = cptr[k]; t < cptr[k + 1]; t++)
                if (cidx[t] < 0 || cidx[t] >= m || (t > cptr[k] && cidx[t] <= cidx[t - 1])) return -1;
        }
    }
    const int64_t nc = nsp ? cptr[nsp] : 0;
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(device) != hipSuccess) return -3;
    // method 0: the library factorization while m^2 < 2^31, the blocked kernels above; 1: library; 2: blocked kernels
    // (vector FMA trailing updates); 3: blocked kernels with the trailing updates on the matrix cores (one wave per
    // SIMD at 128 accumulator registers). At m = 90,000: library 6.5 s, method 2 7.2 s, method 3 9.1 s, the same
    // factor sum; the library's 64-bit dpotrf is still not used by default above 2^31 entries (see lbhip.h). The
    // blocked kernels' panels with rocBLAS dsyrk_64 as the trailing update ran 22-25% faster at m = 90,000-120,000,
    // but at m = 131,287 reported a non-positive leading minor (850) on a core that method 2 factors -- the size at
    // which the library's dpotrf_64 fails too -- so the trailing updates stay in this library's own kernels.
    const bool own = method >= 2 || (method == 0 && m * m >= ((int64_t)1 << 31));
    const SolverApi &api = solver_api();
    if (!own && (!api.ok || (!api.potrf64 && m * m >= ((int64_t)1 << 31)))) { hipSetDevice(prev); return -3; }
    size_t fr = 0, tt = 0;
    if (hipMemGetInfo(&fr, &tt) != hipSuccess) { hipSetDevice(prev); return -3; }
    const size_t need = sizeof(double) * ((size_t)m * m + m) + 16 * (size_t)(na + nc + nsp) + (256u << 20);
    if (need > fr) { hipSetDevice(prev); return -2; }
    hipStream_t s = nullptr;
    void *S = nullptr, *dai = nullptr, *daj = nullptr, *dav = nullptr, *dcp = nullptr, *dci = nullptr, *dcv = nullptr,
         *dcd = nullptr, *dcs = nullptr, *dinfo = nullptr;
    auto cleanup = [&]() {
        for (void *q : {S, dai, daj, dav, dcp, dci, dcv, dcd, dcs, dinfo}) if (q) hipFree(q);
        if (s) hipStreamDestroy(s);
        hipSetDevice(prev);
    };
    int rc = 0;
#define KHIP(x) do { if ((x) != hipSuccess) { cleanup(); return -3; } } while (0)
    KHIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    KHIP(hipMalloc(&S, sizeof(double) * (size_t)m * m));
    KHIP(hipMalloc(&dcs, sizeof(double) * (size_t)m));
    KHIP(hipMalloc(&dinfo, sizeof(int64_t)));
    KHIP(hipMemsetAsync(S, 0, sizeof(double) * (size_t)m * m, s));
    if (na) {
        KHIP(hipMalloc(&dai, sizeof(int32_t) * na)); KHIP(hipMalloc(&daj, sizeof(int32_t) * na));
        KHIP(hipMalloc(&dav, sizeof(double) * na));
        KHIP(hipMemcpyAsync(dai, ai, sizeof(int32_t) * na, hipMemcpyHostToDevice, s));
        KHIP(hipMemcpyAsync(daj, aj, sizeof(int32_t) * na, hipMemcpyHostToDevice, s));
        KHIP(hipMemcpyAsync(dav, av, sizeof(double) * na, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(lb_k_core_a, dim3(4096), dim3(256), 0, s, (const int32_t *)dai, (const int32_t *)daj,
                           (const double *)dav, na, m, (double *)S);
    }
    if (nsp && nc) {
        KHIP(hipMalloc(&dcp, sizeof(int64_t) * (nsp + 1))); KHIP(hipMalloc(&dci, sizeof(int32_t) * nc));
        KHIP(hipMalloc(&dcv, sizeof(double) * nc)); KHIP(hipMalloc(&dcd, sizeof(double) * nsp));
        KHIP(hipMemcpyAsync(dcp, cptr, sizeof(int64_t) * (nsp + 1), hipMemcpyHostToDevice, s));
        KHIP(hipMemcpyAsync(dci, cidx, sizeof(int32_t) * nc, hipMemcpyHostToDevice, s));
        KHIP(hipMemcpyAsync(dcv, cval, sizeof(double) * nc, hipMemcpyHostToDevice, s));
        KHIP(hipMemcpyAsync(dcd, cd, sizeof(double) * nsp, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(lb_k_core_upd, dim3(8192), dim3(256), 0, s, (const int64_t *)dcp, (const int32_t *)dci,
                           (const double *)dcv, (const double *)dcd, nsp, m, (double *)S);
    }
    KHIP(hipGetLastError());
    int64_t info = -1;
    if (own) {
        KHIP(hipMemsetAsync(dinfo, 0, sizeof(int64_t), s));
        // panels of 2 LB_CH columns: the first half (diagonal block, rows below), its rank-LB_CH update of the
        // second half's columns, the second half, then one rank-2 LB_CH update of the trailing lower triangle
        double *Sd = (double *)S;
        int64_t *inf = (int64_t *)dinfo;
        auto upd = method == 3 ? lb_k_chol_update_mfma : lb_k_chol_update;
        auto half = [&](int64_t k) {
            hipLaunchKernelGGL(lb_k_chol_diag, dim3(1), dim3(256), 0, s, Sd, m, k, inf);
            const int64_t rest = m - k - LB_CH;
            if (rest > 0)
                hipLaunchKernelGGL(lb_k_chol_trsm, dim3((unsigned)((rest + 255) / 256)), dim3(256), 0, s, Sd, m, k,
                                   (const int64_t *)inf);
        };
        for (int64_t k = 0; k < m; k += 2 * LB_CH) {
            half(k);
            const int64_t k2 = k + LB_CH;
            if (k2 < m) {
                const int64_t ce = std::min<int64_t>(m, k2 + LB_CH);
                hipLaunchKernelGGL(upd, dim3((unsigned)((m - k2 + 127) / 128)), dim3(256), 0, s, Sd, m, k2, k2, ce,
                                   k, LB_CH, 0, (const int64_t *)inf);
                half(k2);
            }
            const int64_t base = k + 2 * LB_CH, rest = m - base;
            if (rest > 0) {
                const int64_t T = (rest + 127) / 128;
                hipLaunchKernelGGL(upd, dim3((unsigned)(T * (T + 1) / 2)), dim3(256), 0, s, Sd, m, base, base, m, k,
                                   2 * LB_CH, 1, (const int64_t *)inf);
            }
            if ((k / (2 * LB_CH)) % 32 == 31) KHIP(hipStreamSynchronize(s));   // bounded queue depth
        }
        KHIP(hipGetLastError());
        KHIP(hipMemcpyAsync(&info, dinfo, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        KHIP(hipStreamSynchronize(s));
    } else {
    rocblas_handle hb = nullptr;
    if (api.create(&hb) != rocblas_status_success) { cleanup(); return -3; }
    api.set_stream(hb, s);
    rocblas_status st;
    if (api.potrf64) {
        st = api.potrf64(hb, rocblas_fill_lower, m, (double *)S, m, (int64_t *)dinfo);
        if (hipMemcpyAsync(&info, dinfo, sizeof(int64_t), hipMemcpyDeviceToHost, s) != hipSuccess) rc = -3;
    } else {
        st = api.potrf(hb, rocblas_fill_lower, (rocblas_int)m, (double *)S, (rocblas_int)m, (rocblas_int *)dinfo);
        int i32 = -1;
        if (hipMemcpyAsync(&i32, dinfo, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess) rc = -3;
        if (hipStreamSynchronize(s) != hipSuccess) rc = -3;
        info = i32;
    }
    if (hipStreamSynchronize(s) != hipSuccess) rc = -3;
    api.destroy(hb);
    if (st != rocblas_status_success || rc) { cleanup(); return -3; }
    }
    double sr = 0.0;
    if (info == 0) {
        std::vector<double> cs((size_t)m);
        hipLaunchKernelGGL(lb_k_core_colsum, dim3((unsigned)std::min<int64_t>(m, 8192)), dim3(256), 0, s,
                           (const double *)S, m, (double *)dcs);
        KHIP(hipGetLastError());
        KHIP(hipMemcpyAsync(cs.data(), dcs, sizeof(double) * (size_t)m, hipMemcpyDeviceToHost, s));
        KHIP(hipStreamSynchronize(s));
        for (double x : cs) sr += x * x;
    }
    out[0] = (double)info;
    out[1] = sr;
#undef KHIP
    cleanup();
    return 0;
}

}  // extern "C"

// ================================================================================================ batch mode
// Many independent graphs in one context: the graph uploaded by lb_set_graph is the block-diagonal union of bG
// graphs with contiguous row ranges; each graph g has its own L_g, C0_g, rank-1 term z_g = d_g^T V_g and one pair
// level of P_g clusters (the triangle pool holds triangles inside the graphs). One launch of each kernel below covers
// every graph: per-graph column sums (one workgroup per graph), per-graph cluster rows, one wavefront per row over
// all rows of all graphs, and per-graph fixed-order finals. The arithmetic per row and per cluster pair is that of
// the single-graph evaluation (augmented / CV, alm_vcs.py:1496-1533, 1299-1312).
namespace {

struct BArgs {
    const uint4 *V; int64_t n; const int64_t *rowptr; const int32_t *col; const void *w; const double *d;
    const int32_t *gid; const int64_t *boff; int G; const double *Lg; const double *C0g;
    int lev, ylds; const int32_t *lab; const int32_t *Poff; const int64_t *soff; const int32_t *cg; int sumP;
    const double *sig; double *pp; const double *rq; const double *dsq; double wpair;
    double *zY; double *M; double *pdiag; double *ltot; double *lviol;
    const int64_t *iptr; const void *rec; void *pt; double rho, rt; int64_t T; const int32_t *TI; const void *lam;
    void *Gout; double *rowv; double *parts; double *lamsum;
    const int64_t *tptr; const int32_t *tidx;   // triangles of each graph (CSR over graphs)
    const int32_t *light; int64_t nlight;       // rows with few edges and records: one slot per row
    const int32_t *heavy; int64_t nheavy;       // the others: one wavefront per row
};


// z_g and the cluster sums of graph g = blockIdx.x (Y in LDS when it fits, else global atomics)
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_bcolsum(BArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    extern __shared__ double shY[];
    __shared__ double zpart[LB_BLOCK * CPL];
    const int g = blockIdx.x;
    const int64_t r0 = a.boff[g], r1 = a.boff[g + 1];
    const int P = a.lev ? a.Poff[g + 1] - a.Poff[g] : 0;
    const bool lds = a.lev && a.ylds;              // the launch sized the LDS for the largest P of the batch
    double *Yg = a.zY + (int64_t)a.G * R + (a.lev ? (int64_t)a.Poff[g] * R : 0);
    if (lds) {
        for (int e = threadIdx.x; e < P * R; e += LB_BLOCK) shY[e] = 0.0;
        __syncthreads();
    }
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    double zc[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) zc[c] = 0.0;
    for (int64_t i = r0 + slot; i < r1; i += NSL) {
        TC x[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
        const double di = a.d[i];
#pragma unroll
        for (int c = 0; c < CPL; c++) zc[c] += di * (double)x[c];
        if (a.lev) {
            const int Q = a.lab[i];
            double *dst = lds ? shY + Q * R + q * CPL : Yg + (int64_t)Q * R + q * CPL;
#pragma unroll
            for (int c = 0; c < CPL; c++) atomicAdd(dst + c, di * (double)x[c]);
        }
    }
#pragma unroll
    for (int c = 0; c < CPL; c++) zpart[threadIdx.x * CPL + c] = zc[c];
    __syncthreads();
    // z: column k = q*CPL + c is held by threads with threadIdx.x % S == q; summed over the slots in a fixed order
    for (int k = threadIdx.x; k < R; k += LB_BLOCK) {
        const int qq = k / CPL, cc = k % CPL;
        double z = 0.0;
        for (int sl = 0; sl < NSL; sl++) z += zpart[(sl * S + qq) * CPL + cc];
        a.zY[(int64_t)g * R + k] = z;
    }
    if (lds)
        for (int e = threadIdx.x; e < P * R; e += LB_BLOCK) Yg[e] = shY[e];
}

// one wavefront per cluster row Q (global id blockIdx.x) of its graph's pair level (lb_k_levels per graph)
template <int R>
__global__ __launch_bounds__(64) void lb_k_blevels(BArgs a) {
    const int Qg = blockIdx.x, g = a.cg[Qg];
    const int P = a.Poff[g + 1] - a.Poff[g], Q = Qg - a.Poff[g];
    const double *Yb = a.zY + (int64_t)a.G * R + (int64_t)a.Poff[g] * R;
    const double *YQ = Yb + (int64_t)Q * R;
    const double *sig = a.sig + a.soff[g] + (int64_t)Q * P;
    double *pp = a.pp + a.soff[g] + (int64_t)Q * P;
    const double rqQ = a.rq[Qg];
    double Macc[R];
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = 0.0;
    double tot = 0.0, viol = 0.0, pq = 0.0;
    for (int Rr = threadIdx.x; Rr < P; Rr += 64) {
        const double *YR = Yb + (int64_t)Rr * R;
        double p = sig[Rr];
        if (a.rho > 0.0) {
            double h = 0.0;
#pragma unroll
            for (int c = 0; c < R; c++) h += YQ[c] * YR[c];
            if (Rr == Q) h -= a.dsq[Qg];
            const double rr = a.rho * ((rqQ * a.rq[a.Poff[g] + Rr]) * a.wpair);
            const double sg = p;
            p = sg - rr * h;
            if (p < 0.0) p = 0.0;
            tot += (p * p - sg * sg) / (2 * rr);
            if (h < -1e-9) viol += 1.0;
        }
        pp[Rr] = p;
        if (Rr == Q) pq = p;
        if (p != 0.0) {
#pragma unroll
            for (int c = 0; c < R; c++) Macc[c] += p * YR[c];
        }
    }
    tot = wave_sum(tot); viol = wave_sum(viol); pq = wave_sum(pq);
#pragma unroll
    for (int c = 0; c < R; c++) Macc[c] = wave_sum(Macc[c]);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int c = 0; c < R; c++) a.M[(int64_t)Qg * R + c] = Macc[c];
        a.pdiag[Qg] = pq; a.ltot[Qg] = tot; a.lviol[Qg] = viol;
    }
}

// the end of a batch row: G_i = (A V)_i + triangle terms - d_i z_g / L_g + level terms, rowv[4 i ..] =
// (V_i.(AV)_i, G_i.V_i, triangle PHR term, min 1 - tau); written by the lanes with `writer` set
template <typename TV, int R>
__device__ inline void brow_finish(const BArgs &a, int64_t i, int q, bool writer, const typename VT<TV>::TC *x,
                                   const typename VT<TV>::TC *ga, const typename VT<TV>::TC *gt, double ttot,
                                   double tmin) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL;
    const int g = a.gid[i];
    const double di = a.d[i], Lg = a.Lg[g];
    TC qd = 0, g2[CPL];
#pragma unroll
    for (int c = 0; c < CPL; c++) qd += x[c] * ga[c];
    qd = slot_sum<S>(qd);
#pragma unroll
    for (int c = 0; c < CPL; c++) g2[c] = ga[c] + gt[c] - (TC)(di * a.zY[(int64_t)g * R + q * CPL + c] / Lg);
    if (a.lev) {
        const int Qg = a.Poff[g] + a.lab[i];
        const double pd = a.pdiag[Qg] * di;
        const double *M = a.M + (int64_t)Qg * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++) g2[c] += (TC)(di * (M[c] - pd * (double)x[c]));
    }
    TC gd = 0;
#pragma unroll
    for (int c = 0; c < CPL; c++) gd += g2[c] * x[c];
    gd = slot_sum<S>(gd);
    if (writer) {
        TC *Go = (TC *)a.Gout + i * R + q * CPL;
#pragma unroll
        for (int c = 0; c < CPL; c++) Go[c] = g2[c];
        if (q == 0) {
            double *o = a.rowv + 4 * i;
            o[0] = (double)qd; o[1] = (double)gd; o[2] = ttot; o[3] = tmin;
        }
    }
}

// one wavefront per row: NG = 64 / S groups of S lanes take the row's edges and triangle records NG at a time,
// then an xor-shuffle sums the groups; rowv[4 i ..] = (V_i.(AV)_i, G_i.V_i, triangle PHR term, min 1 - tau)
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_brows(BArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NG = 64 / S;
    const int lane = threadIdx.x & 63, grp = lane / S, q = lane % S;
    const int64_t w0 = (int64_t)blockIdx.x * (LB_BLOCK / 64) + (threadIdx.x >> 6);
    const int64_t nw = (int64_t)gridDim.x * (LB_BLOCK / 64);
    const TC *wt = (const TC *)a.w;
    const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
    TC *pt = (TC *)a.pt;
    for (int64_t w = w0; w < a.nheavy; w += nw) {
        const int64_t i = a.heavy[w];
        TC x[CPL], ga[CPL], gt[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
#pragma unroll
        for (int c = 0; c < CPL; c++) { ga[c] = 0; gt[c] = 0; }
        const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
        // edges: each group takes 8 of its edges at a time (8 gathers in flight per lane); indices past the row are
        // clamped to the row's last edge and weighted 0
        for (int64_t eb = e0 + grp; eb < e1; eb += LB_BEB * NG) {
            int64_t jj[LB_BEB];
            TC ww[LB_BEB];
#pragma unroll
            for (int t = 0; t < LB_BEB; t++) {
                const int64_t e = eb + (int64_t)t * NG;
                const int64_t ec = e < e1 ? e : e1 - 1;
                jj[t] = a.col[ec];
                ww[t] = e < e1 ? (wt ? wt[ec] : (TC)1) : (TC)0;
            }
            uint4 u[LB_BEB];
#pragma unroll
            for (int t = 0; t < LB_BEB; t++) u[t] = a.V[jj[t] * S + q];
#pragma unroll
            for (int t = 0; t < LB_BEB; t++) {
                TC xj[CPL];
                VT<TV>::unpack(u[t], xj);
#pragma unroll
                for (int c = 0; c < CPL; c++) ga[c] += ww[t] * xj[c];
            }
        }
        double ttot = 0.0, tmin = INFINITY;
        const int64_t k0 = a.iptr[i], k1 = a.iptr[i + 1];
        // triangle records: 4 per group at a time; records past the row are processed with lam = 0 and rho = 0
        for (int64_t kb = k0 + grp; kb < k1; kb += LB_BRB * NG) {
            TriRec<TC> rc[LB_BRB];
#pragma unroll
            for (int t = 0; t < LB_BRB; t++) {
                const int64_t k = kb + (int64_t)t * NG;
                rc[t] = rec[k < k1 ? k : k1 - 1];
            }
            uint4 uu[LB_BRB], uw[LB_BRB];
#pragma unroll
            for (int t = 0; t < LB_BRB; t++) { uu[t] = a.V[(int64_t)rc[t].u * S + q]; uw[t] = a.V[(int64_t)rc[t].w * S + q]; }
#pragma unroll
            for (int t = 0; t < LB_BRB; t++) {
                const bool valid = kb + (int64_t)t * NG < k1;
                TC xu[CPL], xw[CPL], p, tau;
                VT<TV>::unpack(uu[t], xu);
                VT<TV>::unpack(uw[t], xw);
                tri_record_x<TV, R>(rc[t].code, valid ? rc[t].lam : (TC)0, x, xu, xw, valid ? a.rho : 0.0, a.rt, gt,
                                    p, tau);
                if (valid && rc[t].code % 3 == 0 && q == 0) {
                    pt[rc[t].code / 3] = p;
                    if (a.rho > 0.0) {
                        ttot += ((double)p * p - (double)rc[t].lam * rc[t].lam) / (2 * a.rt);
                        tmin = fmin(tmin, (double)((TC)1 - tau));
                    }
                }
            }
        }
        for (int m = S; m < 64; m <<= 1) {
#pragma unroll
            for (int c = 0; c < CPL; c++) { ga[c] += __shfl_xor(ga[c], m); gt[c] += __shfl_xor(gt[c], m); }
            ttot += __shfl_xor(ttot, m);
            tmin = fmin(tmin, __shfl_xor(tmin, m));
        }
        // every group now holds the sums; group 0 writes the row
        brow_finish<TV, R>(a, i, q, grp == 0, x, ga, gt, ttot, tmin);
    }
}

// one slot of S lanes per light row (LB_BLIGHT): its edges LB_EB and its triangle records LB_RB at a time, in order
template <typename TV, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_brows_slot(BArgs a) {
    typedef typename VT<TV>::TC TC;
    constexpr int CPL = VT<TV>::CPL, S = R / CPL, NSL = LB_BLOCK / S;
    const int slot = threadIdx.x / S, q = threadIdx.x % S;
    const TC *wt = (const TC *)a.w;
    const TriRec<TC> *rec = (const TriRec<TC> *)a.rec;
    TC *pt = (TC *)a.pt;
    for (int64_t p = (int64_t)blockIdx.x * NSL + slot; p < a.nlight; p += (int64_t)gridDim.x * NSL) {
        const int64_t i = a.light[p];
        TC x[CPL], ga[CPL], gt[CPL];
        VT<TV>::unpack(a.V[i * S + q], x);
#pragma unroll
        for (int c = 0; c < CPL; c++) { ga[c] = 0; gt[c] = 0; }
        const int64_t e0 = a.rowptr[i], e1 = a.rowptr[i + 1];
        for (int64_t e = e0; e < e1; e += LB_EB) {
            const int64_t nE = e1 - e;
            int64_t jj[LB_EB];
            TC ww[LB_EB];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) {
                const int64_t ec = t < nE ? e + t : e1 - 1;
                jj[t] = t < nE ? (int64_t)a.col[ec] : i;
                ww[t] = t < nE ? (wt ? wt[ec] : (TC)1) : (TC)0;
            }
            uint4 u[LB_EB];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) u[t] = a.V[jj[t] * S + q];
#pragma unroll
            for (int t = 0; t < LB_EB; t++) {
                TC xj[CPL];
                VT<TV>::unpack(u[t], xj);
#pragma unroll
                for (int c = 0; c < CPL; c++) ga[c] += ww[t] * xj[c];
            }
        }
        double ttot = 0.0, tmin = INFINITY;
        const int64_t k0 = a.iptr[i], k1 = a.iptr[i + 1];
        for (int64_t k = k0; k < k1; k += LB_RB) {
            const int64_t nK = k1 - k;
            TriRec<TC> rc[LB_RB];
#pragma unroll
            for (int t = 0; t < LB_RB; t++) rc[t] = rec[t < nK ? k + t : k1 - 1];
            uint4 uu[LB_RB], uw[LB_RB];
#pragma unroll
            for (int t = 0; t < LB_RB; t++) { uu[t] = a.V[(int64_t)rc[t].u * S + q]; uw[t] = a.V[(int64_t)rc[t].w * S + q]; }
#pragma unroll
            for (int t = 0; t < LB_RB; t++) {
                if (t >= nK) break;                            // slot-uniform
                TC xu[CPL], xw[CPL], p, tau;
                VT<TV>::unpack(uu[t], xu);
                VT<TV>::unpack(uw[t], xw);
                tri_record_x<TV, R>(rc[t].code, rc[t].lam, x, xu, xw, a.rho, a.rt, gt, p, tau);
                if (rc[t].code % 3 == 0 && q == 0) {
                    pt[rc[t].code / 3] = p;
                    if (a.rho > 0.0) {
                        ttot += ((double)p * p - (double)rc[t].lam * rc[t].lam) / (2 * a.rt);
                        tmin = fmin(tmin, (double)((TC)1 - tau));
                    }
                }
            }
        }
        brow_finish<TV, R>(a, i, q, true, x, ga, gt, ttot, tmin);
    }
}

// sum of lam over each graph's triangles (graph = blockIdx.x; its triangle ids in tptr/tidx), in a fixed order
template <typename TC>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_blamsum(BArgs a) {
    __shared__ double red[LB_BLOCK / 64];
    const TC *lam = (const TC *)a.lam;
    const int g = blockIdx.x;
    double s = 0.0;
    for (int64_t k = a.tptr[g] + threadIdx.x; k < a.tptr[g + 1]; k += LB_BLOCK) s += (double)lam[a.tidx[k]];
    s = block_sum(s, red);
    if (threadIdx.x == 0) a.lamsum[g] = s;
}

// per-graph scalars (graph = blockIdx.x), fixed-order sums over its rows and clusters
__global__ __launch_bounds__(LB_BLOCK) void lb_k_bfinal(BArgs a, int R) {
    __shared__ double red[LB_BLOCK / 64];
    const int g = blockIdx.x;
    double quad = 0, gv = 0, tt = 0, tmin = INFINITY, lt = 0, lv = 0;
    for (int64_t i = a.boff[g] + threadIdx.x; i < a.boff[g + 1]; i += LB_BLOCK) {
        const double *o = a.rowv + 4 * i;
        quad += o[0]; gv += o[1]; tt += o[2]; tmin = fmin(tmin, o[3]);
    }
    if (a.lev)
        for (int Q = a.Poff[g] + threadIdx.x; Q < a.Poff[g + 1]; Q += LB_BLOCK) { lt += a.ltot[Q]; lv += a.lviol[Q]; }
    quad = block_sum(quad, red); __syncthreads();
    gv = block_sum(gv, red); __syncthreads();
    tt = block_sum(tt, red); __syncthreads();
    tmin = block_min(tmin, red); __syncthreads();
    lt = block_sum(lt, red); __syncthreads();
    lv = block_sum(lv, red);
    if (threadIdx.x == 0) {
        double zz = 0.0;
        for (int c = 0; c < R; c++) zz += a.zY[(int64_t)g * R + c] * a.zY[(int64_t)g * R + c];
        const double L = a.Lg[g], C0 = a.C0g[g], ls = a.lamsum[g];
        double *p = a.parts + (int64_t)g * LB_NPARTS;
        p[2] = zz / L; p[7] = C0; p[11] = ls;
        p[6] = quad - zz / L + C0;
        if (a.rho > 0.0) {
            p[1] = quad; p[3] = lt; p[4] = tt; p[5] = 0.0; p[9] = tmin;
            p[0] = quad - zz / L - lt - tt + C0;
        } else {
            p[1] = gv; p[3] = 0.0; p[4] = 0.0; p[5] = 0.0; p[9] = INFINITY;
            p[0] = C0 + ls + gv;
        }
        p[8] = lv; p[10] = 0.0;
    }
}

template <typename TV, int R>
struct BLaunch {
    typedef typename VT<TV>::TC TC;
    static int eval(lb_ctx *c, double rho, double rt, const BArgs &a0) {
        BArgs a = a0;
        a.rho = rho; a.rt = rt;
        hipStream_t s = c->stream;
        const int64_t ylen = (int64_t)c->bG * R + (int64_t)c->bsumP * R;
        hipLaunchKernelGGL(lb_k_fill, dim3(256), dim3(256), 0, s, a.zY, ylen, 0.0);
        // per-graph sums of lam: recomputed only after lam changed (lb_set_triangles / lb_set_lam / lb_phr_update)
        if (!c->b_lamsum_ok) {
            if (c->T) {
                if (c->cbytes == 8) hipLaunchKernelGGL(lb_k_blamsum<double>, dim3(c->bG), dim3(LB_BLOCK), 0, s, a);
                else hipLaunchKernelGGL(lb_k_blamsum<float>, dim3(c->bG), dim3(LB_BLOCK), 0, s, a);
            } else {
                hipLaunchKernelGGL(lb_k_fill, dim3(64), dim3(256), 0, s, a.lamsum, (int64_t)c->bG, 0.0);
            }
            c->b_lamsum_ok = true;
        }
        int maxP = 0;
        for (int P : c->bP) maxP = std::max(maxP, P);
        const size_t lds = (c->blev && (size_t)maxP * R * 8 <= 32 * 1024) ? (size_t)maxP * R * 8 : 0;
        a.ylds = lds > 0 ? 1 : 0;
        hipLaunchKernelGGL((lb_k_bcolsum<TV, R>), dim3(c->bG), dim3(LB_BLOCK), lds, s, a);
        if (c->blev && c->bsumP) hipLaunchKernelGGL(lb_k_blevels<R>, dim3(c->bsumP), dim3(64), 0, s, a);
        constexpr int NSLb = LB_BLOCK / (R / VT<TV>::CPL);   // slots per workgroup of the light-row kernel
        if (a.nlight) {
            const int64_t lblk = std::min<int64_t>((a.nlight + NSLb - 1) / NSLb, c->b_lgrid);
            hipLaunchKernelGGL((lb_k_brows_slot<TV, R>), dim3((unsigned)lblk), dim3(LB_BLOCK), 0, s, a);
        }
        if (a.nheavy) {
            const int64_t nblk = std::min<int64_t>((a.nheavy + 3) / 4, c->b_hgrid);
            hipLaunchKernelGGL((lb_k_brows<TV, R>), dim3((unsigned)nblk), dim3(LB_BLOCK), 0, s, a);
        }
        hipLaunchKernelGGL(lb_k_bfinal, dim3(c->bG), dim3(LB_BLOCK), 0, s, a, R);
        HIPCHK(c, hipGetLastError());
        return 0;
    }
};

typedef int (*beval_fn)(lb_ctx *, double, double, const BArgs &);

// light / heavy row lists of a batch context (rebuilt after lb_batch_set or lb_set_triangles)
int ensure_batch_lists(lb_ctx *c) {
    if (c->b_lists_ok) return 0;
    const int64_t n = c->n;
    const bool has_t = c->h_iptr.size() == (size_t)n + 1;
    std::vector<int32_t> L, H;
    for (int64_t i = 0; i < n; i++) {
        const int64_t de = c->h_rowptr[i + 1] - c->h_rowptr[i];
        const int64_t dt = has_t ? c->h_iptr[i + 1] - c->h_iptr[i] : 0;
        (de <= c->b_light_max && dt <= c->b_light_max ? L : H).push_back((int32_t)i);
    }
    int rc;
    if ((rc = dalloc(c, c->b_light, sizeof(int32_t) * std::max<size_t>(L.size(), 1))) ||
        (rc = dalloc(c, c->b_heavy, sizeof(int32_t) * std::max<size_t>(H.size(), 1))))
        return rc;
    if (!L.empty() && (rc = upload(c, c->b_light, L.data(), sizeof(int32_t) * L.size()))) return rc;
    if (!H.empty() && (rc = upload(c, c->b_heavy, H.data(), sizeof(int32_t) * H.size()))) return rc;
    c->b_nlight = (int64_t)L.size(); c->b_nheavy = (int64_t)H.size();
    c->b_lists_ok = true;
    return 0;
}
beval_fn pick_beval(int vdtype, int r) {
#define LB_PICK(TV)                                     \
    if (r == 8) return &BLaunch<TV, 8>::eval;           \
    if (r == 16) return &BLaunch<TV, 16>::eval;         \
    if (r == 32) return &BLaunch<TV, 32>::eval;
    if (vdtype == LB_F32) { LB_PICK(float) }
    if (vdtype == LB_BF16) { LB_PICK(bf16raw) }
    if (vdtype == LB_F64) { LB_PICK(double) }
#undef LB_PICK
    return nullptr;
}

}  // namespace

extern "C" {

int lb_batch_set(lb_ctx *c, int G, const int64_t *row_off, const double *L, const double *C0) {
    drop_capture(c);
    if (c) { c->b_lamsum_ok = false; c->b_lists_ok = false; }
    if (!c || !c->n) return fail(c, -1, "lb_batch_set: set the graph (the union of the batch) first");
    if (G < 1 || !row_off || !L || !C0) return fail(c, -1, "lb_batch_set: bad arguments");
    if (row_off[0] != 0 || row_off[G] != c->n) return fail(c, -1, "lb_batch_set: row offsets must run from 0 to n");
    for (int g = 0; g < G; g++) {
        if (row_off[g + 1] <= row_off[g]) return fail(c, -1, "lb_batch_set: every graph needs at least one row");
        if (!(L[g] > 0)) return fail(c, -1, "lb_batch_set: L must be positive");
    }
    if (c->nvcs) return fail(c, -1, "lb_batch_set: VCS levels are not supported in batch mode");
    if (c->T) return fail(c, -1, "lb_batch_set: set the triangles after lb_batch_set (they are indexed per graph)");
    std::vector<int32_t> gid((size_t)c->n);
    for (int g = 0; g < G; g++)
        for (int64_t i = row_off[g]; i < row_off[g + 1]; i++) gid[i] = g;
    // edges must stay inside their graph (the union is block-diagonal)
    std::vector<int32_t> col((size_t)c->nnz);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpy(col.data(), c->col.p, sizeof(int32_t) * (size_t)c->nnz, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < c->n; i++)
        for (int64_t e = c->h_rowptr[i]; e < c->h_rowptr[i + 1]; e++)
            if (gid[col[e]] != gid[i]) return fail(c, -1, "lb_batch_set: an edge joins two graphs");
    int rc;
    c->bG = G;
    c->hboff.assign(row_off, row_off + G + 1);
    c->blev = false; c->bP.assign(G, 0); c->bsumP = 0;
    if ((rc = upload(c, c->b_off, row_off, sizeof(int64_t) * (G + 1)))) return rc;
    if ((rc = upload(c, c->b_gid, gid.data(), sizeof(int32_t) * gid.size()))) return rc;
    if ((rc = upload(c, c->b_L, L, sizeof(double) * G))) return rc;
    if ((rc = upload(c, c->b_C0, C0, sizeof(double) * G))) return rc;
    if ((rc = dalloc(c, c->b_rowv, sizeof(double) * 4 * (size_t)c->n))) return rc;
    if ((rc = dalloc(c, c->b_parts, sizeof(double) * LB_NPARTS * (size_t)G))) return rc;
    if ((rc = dalloc(c, c->b_lamsum, sizeof(double) * (size_t)G))) return rc;
    if ((rc = dalloc(c, c->b_zY, sizeof(double) * (size_t)G * c->r))) return rc;
    c->have_eval = false;
    return 0;
}

int lb_batch_set_levels(lb_ctx *c, const int32_t *lab, const int32_t *P, const double *sig, const double *rq,
                        const double *dsq, double pairlvl_weight) {
    drop_capture(c);
    if (!c || !c->bG) return fail(c, -1, "lb_batch_set_levels: call lb_batch_set first");
    if (!lab || !P || !sig || !rq || !dsq) return fail(c, -1, "lb_batch_set_levels: bad arguments");
    const int G = c->bG;
    std::vector<int32_t> Poff(G + 1, 0), cg;
    std::vector<int64_t> soff(G + 1, 0);
    for (int g = 0; g < G; g++) {
        if (P[g] < 1) return fail(c, -1, "lb_batch_set_levels: every graph needs at least one cluster");
        Poff[g + 1] = Poff[g] + P[g];
        soff[g + 1] = soff[g] + (int64_t)P[g] * P[g];
        for (int q = 0; q < P[g]; q++) cg.push_back(g);
        for (int64_t i = c->hboff[g]; i < c->hboff[g + 1]; i++)
            if (lab[i] < 0 || lab[i] >= P[g]) return fail(c, -1, "lb_batch_set_levels: label out of range");
    }
    int rc;
    HIPCHK(c, hipSetDevice(c->device));
    const int sumP = Poff[G];
    c->bP.assign(P, P + G); c->bsumP = sumP; c->blev = true; c->bw = pairlvl_weight;
    if ((rc = upload(c, c->b_lab, lab, sizeof(int32_t) * (size_t)c->n))) return rc;
    if ((rc = upload(c, c->b_Poff, Poff.data(), sizeof(int32_t) * (G + 1)))) return rc;
    if ((rc = upload(c, c->b_soff, soff.data(), sizeof(int64_t) * (G + 1)))) return rc;
    if ((rc = upload(c, c->b_cg, cg.data(), sizeof(int32_t) * cg.size()))) return rc;
    if ((rc = upload(c, c->b_sig, sig, sizeof(double) * soff[G]))) return rc;
    if ((rc = dalloc(c, c->b_pp, sizeof(double) * soff[G]))) return rc;
    c->b_nsig = soff[G];
    if ((rc = upload(c, c->b_rq, rq, sizeof(double) * sumP))) return rc;
    if ((rc = upload(c, c->b_dsq, dsq, sizeof(double) * sumP))) return rc;
    if ((rc = dalloc(c, c->b_zY, sizeof(double) * ((size_t)G + sumP) * c->r))) return rc;
    if ((rc = dalloc(c, c->b_M, sizeof(double) * (size_t)sumP * c->r))) return rc;
    if ((rc = dalloc(c, c->b_pdiag, sizeof(double) * sumP))) return rc;
    if ((rc = dalloc(c, c->b_ltot, sizeof(double) * sumP))) return rc;
    if ((rc = dalloc(c, c->b_lviol, sizeof(double) * sumP))) return rc;
    c->have_eval = false;
    return 0;
}

}  // extern "C"

// the batch kernels' arguments for the context's current buffers
static BArgs batch_args(lb_ctx *c) {
    BArgs a{};
    a.V = (const uint4 *)c->V.p; a.n = c->n; a.rowptr = (const int64_t *)c->rowptr.p; a.col = (const int32_t *)c->col.p;
    a.w = c->weighted ? c->w.p : nullptr; a.d = (const double *)c->d.p;
    a.gid = (const int32_t *)c->b_gid.p; a.boff = (const int64_t *)c->b_off.p; a.G = c->bG;
    a.Lg = (const double *)c->b_L.p; a.C0g = (const double *)c->b_C0.p;
    a.lev = c->blev ? 1 : 0; a.lab = (const int32_t *)c->b_lab.p; a.Poff = (const int32_t *)c->b_Poff.p;
    a.soff = (const int64_t *)c->b_soff.p; a.cg = (const int32_t *)c->b_cg.p; a.sumP = c->bsumP;
    a.sig = (const double *)c->b_sig.p; a.pp = (double *)c->b_pp.p; a.rq = (const double *)c->b_rq.p;
    a.dsq = (const double *)c->b_dsq.p; a.wpair = c->bw;
    a.zY = (double *)c->b_zY.p; a.M = (double *)c->b_M.p; a.pdiag = (double *)c->b_pdiag.p;
    a.ltot = (double *)c->b_ltot.p; a.lviol = (double *)c->b_lviol.p;
    a.iptr = (const int64_t *)c->iptr.p; a.rec = c->rec.p; a.pt = c->pt.p; a.T = c->T; a.TI = (const int32_t *)c->TI.p;
    a.lam = c->lam.p;
    a.Gout = c->G.p; a.rowv = (double *)c->b_rowv.p; a.parts = (double *)c->b_parts.p; a.lamsum = (double *)c->b_lamsum.p;
    a.tptr = (const int64_t *)c->b_tptr.p; a.tidx = (const int32_t *)c->b_tidx.p;
    a.light = (const int32_t *)c->b_light.p; a.nlight = c->b_nlight;
    a.heavy = (const int32_t *)c->b_heavy.p; a.nheavy = c->b_nheavy;
    return a;
}

// ------------------------------------------------------------------------------------------------ batched search
// One Riemannian ascent step of every graph of a batch on the oblique manifold (unit rows), after an evaluation
// wrote G = C(y)V: the tangent gradient R_i = G_i - (G_i . V_i) V_i, a per-graph Barzilai-Borwein step
// alpha_g = <s, s> / |<s, y>| (s = V - V_prev, y = R - R_prev over the graph's rows; Barzilai & Borwein, IMA J.
// Numer. Anal. 8, 1988), clamped to [amin, amax] (a0 on a graph's first step), and the retraction
// V_i <- (V_i + alpha_g R_i) / |V_i + alpha_g R_i|. The step is scaled by the graph's largest row gradient
// rmax_g = max_i |R_i|: a graph's first step is a0 / rmax_g, and every step is clamped to [amin, amax] / rmax_g, so
// that no row turns by much more than amax radians whatever the graph's degrees. One workgroup per graph: its sums
// are fixed-order block reductions, so a graph's step does not depend on the other graphs. st (4 per graph): alpha,
// have-previous flag, <s,s>, <s,y>.
struct BRgdArgs {
    void *V; const void *Gm; void *Rb; void *Vp; void *Rp;
    const int64_t *boff; double *st; double a0, amin, amax;
};

template <typename T, int R>
__global__ __launch_bounds__(LB_BLOCK) void lb_k_brgd(BRgdArgs a) {
    __shared__ double red[LB_BLOCK / 64];
    __shared__ double alpha_s;
    const int g = blockIdx.x;
    const int64_t r0 = a.boff[g], r1 = a.boff[g + 1];
    double *st = a.st + 4 * (int64_t)g;
    const bool prev = st[1] > 0.5;
    T *V = (T *)a.V, *Rb = (T *)a.Rb, *Vp = (T *)a.Vp, *Rp = (T *)a.Rp;
    const T *Gm = (const T *)a.Gm;
    double ss = 0.0, sy = 0.0, rmax = 0.0;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += LB_BLOCK) {
        T v[R], gr[R];
#pragma unroll
        for (int c = 0; c < R; c++) { v[c] = V[i * R + c]; gr[c] = Gm[i * R + c]; }
        double gv = 0.0, rn = 0.0;
#pragma unroll
        for (int c = 0; c < R; c++) gv += (double)gr[c] * (double)v[c];
#pragma unroll
        for (int c = 0; c < R; c++) {
            const T rr = (T)((double)gr[c] - gv * (double)v[c]);
            rn += (double)rr * (double)rr;
            if (prev) {
                const double sd = (double)v[c] - (double)Vp[i * R + c], yd = (double)rr - (double)Rp[i * R + c];
                ss += sd * sd;
                sy += sd * yd;
            }
            Rb[i * R + c] = rr;
        }
        rmax = fmax(rmax, sqrt(rn));
    }
    ss = block_sum(ss, red);
    __syncthreads();
    sy = block_sum(sy, red);
    __syncthreads();
    rmax = -block_min(-rmax, red);
    if (threadIdx.x == 0) {
        const double sc = rmax > 0.0 ? 1.0 / rmax : 0.0;
        double al = prev ? st[0] : a.a0 * sc;
        if (prev && fabs(sy) > 0.0) al = ss / fabs(sy);
        al = fmin(fmax(al, a.amin * sc), a.amax * sc);
        st[0] = al; st[1] = 1.0; st[2] = ss; st[3] = sy;
        alpha_s = al;
    }
    __syncthreads();
    const double al = alpha_s;
    for (int64_t i = r0 + threadIdx.x; i < r1; i += LB_BLOCK) {   // the rows this thread wrote R for above
        double w[R], nrm = 0.0;
#pragma unroll
        for (int c = 0; c < R; c++) {
            const T v = V[i * R + c], rr = Rb[i * R + c];
            Vp[i * R + c] = v;
            Rp[i * R + c] = rr;
            w[c] = (double)v + al * (double)rr;
            nrm += w[c] * w[c];
        }
        const double inv = nrm > 0.0 ? 1.0 / sqrt(nrm) : 0.0;
#pragma unroll
        for (int c = 0; c < R; c++) V[i * R + c] = (T)(w[c] * inv);
    }
}

template <typename T, int R> void brgd_launch(lb_ctx *c, const BRgdArgs &a) {
    hipLaunchKernelGGL((lb_k_brgd<T, R>), dim3(c->bG), dim3(LB_BLOCK), 0, c->stream, a);
}

// one search iteration on the context's stream: the batch evaluation at (rho, rt), then the step
static int brgd_iteration(lb_ctx *c, beval_fn f, double rho, double rt, const BRgdArgs &ra) {
    int rc = f(c, rho, rt, batch_args(c));
    if (rc) return rc;
    const bool dbl = c->vdtype == LB_F64;
    switch (c->r) {
        case 8: dbl ? brgd_launch<double, 8>(c, ra) : brgd_launch<float, 8>(c, ra); break;
        case 16: dbl ? brgd_launch<double, 16>(c, ra) : brgd_launch<float, 16>(c, ra); break;
        default: dbl ? brgd_launch<double, 32>(c, ra) : brgd_launch<float, 32>(c, ra); break;
    }
    HIPCHK(c, hipGetLastError());
    return 0;
}

extern "C" {

int lb_batch_rgd(lb_ctx *c, int iters, double rho, double rt, double a0, double amin, double amax, int reset,
                 double *state) {
    if (!c || !c->bG) return fail(c, -1, "lb_batch_rgd: not in batch mode (lb_batch_set)");
    if (c->vdtype != LB_F32 && c->vdtype != LB_F64) return fail(c, -1, "lb_batch_rgd: FP32 or FP64 storage only");
    if (iters < 0 || !(a0 > 0.0) || !(amin > 0.0) || !(amax >= amin) || !(rho >= 0.0) ||
        (rho > 0.0 && c->T && !(rt > 0.0)))
        return fail(c, -1, "lb_batch_rgd: bad arguments");
    if (c->T && !c->b_tptr.p) return fail(c, -1, "lb_batch_rgd: triangles were set before lb_batch_set");
    beval_fn f = pick_beval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_batch_rgd: unsupported rank");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    const size_t vb = (size_t)c->n * c->r * c->cbytes;
    if ((rc = dalloc(c, c->b_R, vb)) || (rc = dalloc(c, c->b_Vp, vb)) || (rc = dalloc(c, c->b_Rp, vb)) ||
        (rc = dalloc(c, c->b_rgst, sizeof(double) * 4 * (size_t)c->bG)))
        return rc;
    if ((rc = ensure_batch_lists(c))) return rc;
    if (reset) hipLaunchKernelGGL(lb_k_fill, dim3(64), dim3(256), 0, c->stream, (double *)c->b_rgst.p,
                                  (int64_t)4 * c->bG, 0.0);
    BRgdArgs ra{};
    ra.V = c->V.p; ra.Gm = c->G.p; ra.Rb = c->b_R.p; ra.Vp = c->b_Vp.p; ra.Rp = c->b_Rp.p;
    ra.boff = (const int64_t *)c->b_off.p; ra.st = (double *)c->b_rgst.p; ra.a0 = a0; ra.amin = amin; ra.amax = amax;
    const double key[5] = {rho, rt, a0, amin, amax};
    if (c->rgd_exec && std::memcmp(key, c->rgd_key, sizeof(key)) != 0) {
        hipGraphExecDestroy((hipGraphExec_t)c->rgd_exec);
        c->rgd_exec = nullptr;
    }
    if (!c->rgd_exec && iters > 0) {             // record one iteration; replays launch it as one graph
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = brgd_iteration(c, f, rho, rt, ra);
        hipGraph_t gr = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &gr);
        if (rc) { if (gr) hipGraphDestroy(gr); return rc; }
        if (e != hipSuccess) return fail(c, -3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        hipGraphExec_t ge = nullptr;
        e = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
        hipGraphDestroy(gr);
        if (e != hipSuccess) return fail(c, -3, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->rgd_exec = ge;
        std::memcpy(c->rgd_key, key, sizeof(key));
    }
    for (int k = 0; k < iters; k++) HIPCHK(c, hipGraphLaunch((hipGraphExec_t)c->rgd_exec, c->stream));
    if (state) HIPCHK(c, hipMemcpyAsync(state, c->b_rgst.p, sizeof(double) * 4 * c->bG, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->have_eval = false;                        // V moved after the last evaluation
    return 0;
}

int lb_batch_get_sig(lb_ctx *c, double *sig) {
    if (!c || !c->bG || !c->blev) return fail(c, -1, "lb_batch_get_sig: no batch pair levels");
    if (!sig) return fail(c, -1, "lb_batch_get_sig: bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(sig, c->b_sig.p, sizeof(double) * c->b_nsig, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

int lb_batch_eval(lb_ctx *c, double rho, double rt, double *parts) {
    if (!c || !c->bG) return fail(c, -1, "lb_batch_eval: not in batch mode (lb_batch_set)");
    if (!(rho >= 0.0) || (rho > 0.0 && c->T && !(rt > 0.0))) return fail(c, -1, "lb_batch_eval: need rho >= 0 and rt > 0 when rho > 0");
    HIPCHK(c, hipSetDevice(c->device));
    beval_fn f = pick_beval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "lb_batch_eval: unsupported rank / dtype");
    if (c->T && !c->b_tptr.p) return fail(c, -1, "lb_batch_eval: triangles were set before lb_batch_set");
    int rcl;
    if ((rcl = ensure_batch_lists(c))) return rcl;
    int rc = f(c, rho, rt, batch_args(c));
    if (rc) return rc;
    if (parts) {
        HIPCHK(c, hipMemcpyAsync(parts, c->b_parts.p, sizeof(double) * LB_NPARTS * (size_t)c->bG, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->have_eval = true; c->last_rho = rho;
    return 0;
}

}  // extern "C"

// ================================================================================================ captured evaluation
// The evaluation chain (single graph or batch) recorded once into a HIP graph on the context's stream and replayed:
// no host work and one launch per replay. The chain clears its accumulators with fill kernels (no memset nodes).
extern "C" {

int lb_capture_eval(lb_ctx *c, double rho, double rt) {
    if (!c || !c->n) return fail(c, -1, "lb_capture_eval: no graph");
    if (!(rho >= 0.0) || (rho > 0.0 && c->T && !(rt > 0.0))) return fail(c, -1, "lb_capture_eval: need rho >= 0 and rt > 0 when rho > 0");
    HIPCHK(c, hipSetDevice(c->device));
    if (c->gexec) { hipGraphExecDestroy((hipGraphExec_t)c->gexec); c->gexec = nullptr; }
    if (c->bG) {                                       // the batch row lists are uploaded before the capture
        int rcl;
        if ((rcl = ensure_batch_lists(c))) return rcl;
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = 0;
    if (c->bG) {
        beval_fn f = pick_beval(c->vdtype, c->r);
        if (!f) rc = fail(c, -1, "lb_capture_eval: unsupported rank / dtype");
        else if (c->T && !c->b_tptr.p) rc = fail(c, -1, "lb_capture_eval: triangles were set before lb_batch_set");
        else rc = f(c, rho, rt, batch_args(c));
    } else {
        eval_fn f = pick_eval(c->vdtype, c->r);
        if (!f) rc = fail(c, -1, "lb_capture_eval: unsupported rank / dtype");
        else rc = f(c, rho, rt, nullptr);
    }
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(c, -3, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    hipGraphExec_t ge = nullptr;
    e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(c, -3, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->gexec = ge; c->g_rho = rho;
    return 0;
}

int lb_replay_eval(lb_ctx *c, int count, double *parts) {
    if (!c || !c->gexec) return fail(c, -1, "lb_replay_eval: nothing captured (lb_capture_eval)");
    if (count < 1) return fail(c, -1, "lb_replay_eval: count must be positive");
    HIPCHK(c, hipSetDevice(c->device));
    for (int k = 0; k < count; k++) HIPCHK(c, hipGraphLaunch((hipGraphExec_t)c->gexec, c->stream));
    if (parts) {
        const DBuf &src = c->bG ? c->b_parts : c->parts;
        const size_t cnt = (size_t)LB_NPARTS * (c->bG ? c->bG : 1);
        HIPCHK(c, hipMemcpyAsync(parts, src.p, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
    }
    c->have_eval = true; c->last_rho = c->g_rho;
    return 0;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------------ device RGD polish
#define LB_POL_BLOCKS 1024

template <typename TV, int R>
int pol_launch(lb_ctx *c, int which, double alpha) {
    typedef typename VT<TV>::TC TC;
    PolArgs a{};
    a.n = c->n; a.alpha = alpha; a.init = which == 3;
    a.V = (const double *)c->pol_V.p; a.Rd = (double *)c->pol_R.p;
    a.Vn = which == 3 ? (double *)c->pol_V.p : (double *)c->pol_Vn.p;
    a.Vst = c->V.p; a.G = c->G.p; a.g = (double *)c->pol_g.p; a.part = (double *)c->pol_part.p;
    if (which == 0) hipLaunchKernelGGL((lb_k_pol_step<TV, R>), dim3(LB_POL_BLOCKS), dim3(LB_BLOCK), 0, c->stream, a);
    else if (which == 1) hipLaunchKernelGGL((lb_k_pol_gsum<TC, R>), dim3(LB_POL_BLOCKS), dim3(LB_BLOCK), 0, c->stream, a);
    else hipLaunchKernelGGL((lb_k_pol_post<TC, R>), dim3(LB_POL_BLOCKS), dim3(LB_BLOCK), 0, c->stream, a);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// which: 0 step, 1 g and its sum, 2 post, 3 post at the current iterate (init)
int pol_kernels(lb_ctx *c, int which, double alpha) {
    const int v = c->vdtype, r = c->r;
    if (v == LB_F32) return r == 8 ? pol_launch<float, 8>(c, which, alpha) : r == 16 ? pol_launch<float, 16>(c, which, alpha) : pol_launch<float, 32>(c, which, alpha);
    if (v == LB_BF16) return r == 8 ? pol_launch<bf16raw, 8>(c, which, alpha) : r == 16 ? pol_launch<bf16raw, 16>(c, which, alpha) : pol_launch<bf16raw, 32>(c, which, alpha);
    return r == 8 ? pol_launch<double, 8>(c, which, alpha) : r == 16 ? pol_launch<double, 16>(c, which, alpha) : pol_launch<double, 32>(c, which, alpha);
}

// evaluation at rho = 0 of the storage copy, then g and sum(g) -> *gsum
int pol_value(lb_ctx *c, double *gsum) {
    eval_fn f = pick_eval(c->vdtype, c->r);
    if (!f) return fail(c, -1, "polish: unsupported rank / dtype");
    int rc;
    if ((rc = f(c, 0.0, 0.0, nullptr))) return rc;
    c->have_eval = true; c->last_rho = 0.0;
    if ((rc = pol_kernels(c, 1, 0.0))) return rc;
    hipLaunchKernelGGL(lb_k_reduce, dim3(1), dim3(1024), 0, c->stream, (const double *)c->pol_part.p, LB_POL_BLOCKS, 1, 1,
                       (double *)c->pol_sum.p);
    HIPCHK(c, hipMemcpyAsync(gsum, c->pol_sum.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// post pass (init or accept): direction update and its four sums -> out4
int pol_post(lb_ctx *c, int init, double *out4) {
    int rc;
    if ((rc = pol_kernels(c, init ? 3 : 2, 0.0))) return rc;
    hipLaunchKernelGGL(lb_k_reduce, dim3(1), dim3(1024), 0, c->stream, (const double *)c->pol_part.p, LB_POL_BLOCKS, 4, 4,
                       (double *)c->pol_sum.p);
    HIPCHK(c, hipMemcpyAsync(out4, c->pol_sum.p, 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

extern "C" {

int lb_polish_begin(lb_ctx *c, const double *V, double *out) {
    drop_capture(c);
    if (!c || !c->n || !V || !out) return fail(c, -1, "lb_polish_begin: no graph, V or output");
    if (c->bG) return fail(c, -1, "lb_polish_begin: not available on a batch context");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r;
    const size_t bytes = sizeof(double) * (size_t)n * r;
    int rc;
    if ((rc = dalloc(c, c->pol_V, bytes)) || (rc = dalloc(c, c->pol_Vn, bytes)) || (rc = dalloc(c, c->pol_R, bytes)) ||
        (rc = dalloc(c, c->pol_g, sizeof(double) * (size_t)n)) ||
        (rc = dalloc(c, c->pol_part, sizeof(double) * 4 * LB_POL_BLOCKS)) || (rc = dalloc(c, c->pol_sum, sizeof(double) * 4)))
        return rc;
    // the iterate in device row order (FP64), then its storage copy through lb_set_V's path
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    void *st = nullptr;
    if ((rc = host_stage(c, bytes, &st))) return rc;
    par_rows(n, [&](int64_t a, int64_t b) { rows_to_storage(a, b, r, V, LB_F64, st, ord); });
    HIPCHK(c, hipMemcpyAsync(c->pol_V.p, st, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if ((rc = lb_set_V(c, V, LB_F64))) return rc;
    HIPCHK(c, hipMemcpyAsync(c->pol_Vn.p, c->pol_V.p, bytes, hipMemcpyDeviceToDevice, c->stream));
    double gs = 0.0, s4[4];
    if ((rc = pol_value(c, &gs))) return rc;
    if ((rc = pol_post(c, 1, s4))) return rc;
    out[0] = gs; out[1] = s4[0];
    c->pol_ok = true;
    return 0;
}

int lb_polish_trial(lb_ctx *c, double alpha, double *gsum) {
    if (!c || !c->pol_ok || !gsum) return fail(c, -1, "lb_polish_trial: no polish state (lb_polish_begin)");
    drop_capture(c);
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = pol_kernels(c, 0, alpha))) return rc;
    return pol_value(c, gsum);
}

int lb_polish_accept(lb_ctx *c, double *out4) {
    if (!c || !c->pol_ok || !out4) return fail(c, -1, "lb_polish_accept: no polish state (lb_polish_begin)");
    HIPCHK(c, hipSetDevice(c->device));
    int rc;
    if ((rc = pol_post(c, 0, out4))) return rc;
    std::swap(c->pol_V, c->pol_Vn);                 // the candidate becomes the iterate
    return 0;
}

int lb_polish_get(lb_ctx *c, double *V, double *g) {
    if (!c || !c->pol_ok) return fail(c, -1, "lb_polish_get: no polish state (lb_polish_begin)");
    HIPCHK(c, hipSetDevice(c->device));
    const int64_t n = c->n;
    const int r = c->r;
    const int64_t *ord = c->row_order.empty() ? nullptr : c->row_order.data();
    int rc;
    void *st = nullptr;
    if ((rc = host_stage(c, sizeof(double) * (size_t)n * r, &st))) return rc;
    if (V) {
        HIPCHK(c, hipMemcpyAsync(st, c->pol_V.p, sizeof(double) * (size_t)n * r, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        par_rows(n, [&](int64_t a, int64_t b) { storage_to_rows(a, b, r, st, LB_F64, V, ord); });
    }
    if (g) {
        HIPCHK(c, hipMemcpyAsync(st, c->pol_g.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const double *gs = (const double *)st;
        for (int64_t k = 0; k < n; k++) g[ord ? ord[k] : k] = gs[k];
    }
    return 0;
}

}  // extern "C"
