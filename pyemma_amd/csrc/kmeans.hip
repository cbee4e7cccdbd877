// kmeans++ seeding and Lloyd iterations.
//
// The reference hands both to deeptime (coordinates/clustering/kmeans.py:211-212,
// 254-261), whose source is outside the reference tree.  This file implements
// the documented rules shared with the CPU oracle:
//
// kmeans++ (Arthur & Vassilvitskii 2007, D^2 sampling):
//   round 0 picks row floor(u_0 n); round r >= 1 picks a row with probability
//   proportional to D^2 = squared FP32 distance (sequential loop, no FMA) to the
//   nearest chosen center.  u_r is draw r of a SplitMix64 stream of the seed.
//   Sums are FP64 in a fixed order: blocks of 4096 frames are summed lane-
//   strided by a 64-lane wavefront and combined with a butterfly; superblocks
//   of 64 blocks and segments of 64 frames are summed sequentially.  The pick
//   walks superblock -> block -> segment -> frame with running FP64 prefixes.
//
// Lloyd: labels = assign(X, C); cost = blocked FP64 sum of each frame's FP32
//   distance to its center; new C_j = FP64 mean of its members taken in frame
//   order in chunks of 256 (empty centers unchanged); stop when
//   |cost - prev| / cost <= tol.  Every iteration's kernels read a device flag
//   and return immediately once converged, so the host enqueues max_iter
//   iterations without synchronizing.
#include <math.h>

#include "emma_common.hpp"

namespace emma {

size_t assign_workspace_bytes(int64_t n, int d, int k, int mode);
int assign_run(const float *X, int64_t n, int d, const float *C, int k, int32_t *labels, int mode,
               int64_t *n_rechecked, void *workspace, const int32_t *done, hipStream_t s);

namespace km {

constexpr int BLOCK = 4096;
static_assert(BLOCK % 64 == 0, "a wave's frames lie in one block");
constexpr int SUPER = 64;
constexpr int SEG = 64;
constexpr int CHUNK = 256;

__device__ __forceinline__ float dist_seq(const float *__restrict__ x, const float *__restrict__ c, int d) {
    float acc = 0.0f;
    for (int i = 0; i < d; ++i) {
        const float t = __fsub_rn(x[i], c[i]);
        acc = __fadd_rn(acc, __fmul_rn(t, t));
    }
    return acc;
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__device__ __forceinline__ double uniform(uint64_t seed, uint64_t counter) {
    const uint64_t z = mix64(seed + (counter + 1ULL) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

// ------------------------------------------------------------------
// kmeans++
// ------------------------------------------------------------------
// round 0: first pick
__global__ void k_kmpp_first(const float *__restrict__ X, int64_t n, int d, uint64_t seed, float *__restrict__ C,
                             int64_t *__restrict__ idx) {
    int64_t pick = (int64_t)(uniform(seed, 0) * (double)n);
    if (pick >= n) pick = n - 1;
    if (threadIdx.x == 0) idx[0] = pick;
    for (int i = threadIdx.x; i < d; i += blockDim.x) C[i] = X[pick * d + i];
}

// D^2 update against center c (row of C): one thread per frame, the
// sequential FP32 loop over the dims in order.  VEC: rows 16-byte aligned
// (d % 4 == 0 and X aligned), read as float4 (the adds stay sequential).
template <bool VEC>
__global__ __launch_bounds__(256) void k_kmpp_d2(const float *__restrict__ X, int64_t n, int d,
                                                 const float *__restrict__ c, int first, float *__restrict__ d2) {
    extern __shared__ float s_c[];
    for (int i = threadIdx.x; i < d; i += blockDim.x) s_c[i] = c[i];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *x = X + i * d;
    float v = first ? 0.0f : d2[i];
    float di;
    if constexpr (VEC) {
        float acc = 0.0f;
        auto step4 = [&](const float4 xv, const float4 cv) __attribute__((always_inline)) {
            float t = __fsub_rn(xv.x, cv.x);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.y, cv.y);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.z, cv.z);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.w, cv.w);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
        };
        // dims in order; the frame loads of 16 dims are issued before their terms
        int q = 0;
        for (; q + 16 <= d; q += 16) {
            float4 xv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) xv[u] = *reinterpret_cast<const float4 *>(x + q + 4 * u);
#pragma unroll
            for (int u = 0; u < 4; ++u) step4(xv[u], *reinterpret_cast<const float4 *>(s_c + q + 4 * u));
        }
        for (; q < d; q += 4)
            step4(*reinterpret_cast<const float4 *>(x + q), *reinterpret_cast<const float4 *>(s_c + q));
        di = acc;
    } else {
        di = dist_seq(x, s_c, d);
    }
    if (first || di < v) v = di;
    d2[i] = v;
}

// Block sums of per-frame values (D^2 here, Lloyd's distances below): one
// wavefront per block of BLOCK frames; lane l sums frames l, l + 64, ... of
// the block in FP64 in that order, then a butterfly.  The loads of a lane are
// issued before its adds.  done (may be null): skip once Lloyd has converged.
__global__ __launch_bounds__(256) void k_block_sums(int64_t n, const float *__restrict__ d2,
                                                    double *__restrict__ bsum, const int32_t *__restrict__ done) {
    if (done && *done) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t b = (int64_t)blockIdx.x * 4 + wave;
    const int64_t nb = (n + BLOCK - 1) / BLOCK;
    if (b >= nb) return;
    const int64_t lo = b * BLOCK;
    double s = 0.0;
    constexpr int U = 16;
    for (int m0 = 0; m0 < BLOCK / 64; m0 += U) {
        float v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = lo + (m0 + u) * 64 + lane;
            v[u] = i < n ? d2[i] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int64_t i = lo + (m0 + u) * 64 + lane;
            if (i < n) s = s + (double)v[u];
        }
    }
    s = wave_butterfly_sum(s);
    if (lane == 0) bsum[b] = s;
}

// Pick of round r (one wavefront), exactly the oracle's oracle_kmpp_pick.
__global__ __launch_bounds__(64) void k_kmpp_pick(const float *__restrict__ X, int64_t n, int d, uint64_t seed,
                                                  int r, const float *__restrict__ d2,
                                                  const double *__restrict__ bsum, double *__restrict__ sbsum,
                                                  float *__restrict__ C, int64_t *__restrict__ idx) {
    __shared__ double s_seg[SEG];
    __shared__ int64_t s_pick;
    const int lane = threadIdx.x;
    const int64_t nb = (n + BLOCK - 1) / BLOCK;
    const int64_t nsb = (nb + SUPER - 1) / SUPER;
    for (int64_t sb = lane; sb < nsb; sb += 64) {
        double a = 0.0;
        for (int64_t b = sb * SUPER; b < nb && b < (sb + 1) * SUPER; ++b) a = a + bsum[b];
        sbsum[sb] = a;
    }
    __syncthreads();
    const double u = uniform(seed, (uint64_t)r);
    // lane 0 walks superblocks and blocks
    int64_t blk = -1;
    double before = 0.0, total = 0.0;
    if (lane == 0) {
        for (int64_t sb = 0; sb < nsb; ++sb) total = total + sbsum[sb];
        if (!(total > 0.0)) {
            int64_t i = (int64_t)(u * (double)n);
            s_pick = i < n ? i : n - 1;
        } else {
            s_pick = -1;
            const double T = u * total;
            double run = 0.0, last_before = 0.0;
            int64_t sel = -1, last_pos = -1;
            for (int64_t sb = 0; sb < nsb; ++sb) {
                const double a = sbsum[sb];
                const double prev = run;
                run = run + a;
                if (a > 0.0) { last_pos = sb; last_before = prev; }
                if (run > T) { sel = sb; before = prev; break; }
            }
            if (sel < 0) { sel = last_pos; before = last_before; }
            const int64_t sb = sel;
            run = before; sel = -1; last_pos = -1;
            for (int64_t b = sb * SUPER; b < nb && b < (sb + 1) * SUPER; ++b) {
                const double prev = run;
                run = run + bsum[b];
                if (bsum[b] > 0.0) { last_pos = b; last_before = prev; }
                if (run > T) { sel = b; before = prev; break; }
            }
            if (sel < 0) { sel = last_pos; before = last_before; }
            blk = sel;
        }
        s_seg[0] = before;  // pass through shared memory
        s_seg[1] = total > 0.0 ? u * total : 0.0;
        s_seg[2] = (double)blk;
    }
    __syncthreads();
    const double before_blk = s_seg[0], T = s_seg[1];
    blk = (int64_t)s_seg[2];
    __syncthreads();
    if (blk >= 0) {
        // segment sums of the chosen block in parallel (sequential inside each segment)
        const int64_t blo = blk * BLOCK, bhi = blo + BLOCK < n ? blo + BLOCK : n;
        const int64_t g = blo + (int64_t)lane * SEG;
        double a = 0.0;
        if (g < bhi)
            for (int64_t i = g; i < bhi && i < g + SEG; ++i) a = a + (double)d2[i];
        s_seg[lane] = a;
        __syncthreads();
        if (lane == 0) {
            double run = before_blk, last_before = 0.0, before_seg = 0.0;
            int64_t sel = -1, last_pos = -1;
            for (int q = 0; q < SEG; ++q) {
                const int64_t gq = blo + (int64_t)q * SEG;
                if (gq >= bhi) break;
                const double prev = run;
                run = run + s_seg[q];
                if (s_seg[q] > 0.0) { last_pos = gq; last_before = prev; }
                if (run > T) { sel = gq; before_seg = prev; break; }
            }
            if (sel < 0) { sel = last_pos; before_seg = last_before; }
            const int64_t gs = sel, ghi = gs + SEG < bhi ? gs + SEG : bhi;
            run = before_seg; sel = -1; last_pos = -1;
            for (int64_t i = gs; i < ghi; ++i) {
                run = run + (double)d2[i];
                if (d2[i] > 0.0f) last_pos = i;
                if (run > T) { sel = i; break; }
            }
            if (sel < 0) sel = last_pos;
            s_pick = sel;
        }
    }
    __syncthreads();
    const int64_t pick = s_pick;
    if (lane == 0) idx[r] = pick;
    for (int i = lane; i < d; i += 64) C[(int64_t)r * d + i] = X[pick * d + i];
}

struct PpWs {
    float *d2;
    double *bsum, *sbsum;
    size_t bytes;
};

static PpWs carve_pp(void *ws, int64_t n) {
    PpWs w{};
    Carver cv(ws);
    const int64_t nb = cdiv(n, BLOCK);
    w.d2 = cv.take<float>(n);
    w.bsum = cv.take<double>(nb);
    w.sbsum = cv.take<double>(cdiv(nb, SUPER));
    w.bytes = cv.bytes();
    return w;
}

// ------------------------------------------------------------------
// Lloyd
// ------------------------------------------------------------------
struct LloydWs {
    int32_t *done;        // [1]
    double *prev_cost;    // [1]
    double *bcost;        // [nb]
    int64_t *cnt;         // [k]
    int64_t *off;         // [k + 1] member offsets
    int64_t *coff;        // [k + 1] chunk offsets
    int32_t *blk_cnt;     // [nb * k] per-block label counts -> block prefixes
    int64_t *order;       // [n] members sorted by (label, frame)
    float *dist;          // [n] distance of each frame to its center
    double *chunk_sums;   // [n / CHUNK + k][d]
    void *assign_ws;
    size_t bytes;
};

static LloydWs carve_lloyd(void *ws, int64_t n, int d, int k, int mode) {
    LloydWs w{};
    Carver cv(ws);
    const int64_t nb = cdiv(n, BLOCK);
    w.done = cv.take<int32_t>(1);
    w.prev_cost = cv.take<double>(1);
    w.bcost = cv.take<double>(nb);
    w.cnt = cv.take<int64_t>(k);
    w.off = cv.take<int64_t>((size_t)k + 1);
    w.coff = cv.take<int64_t>((size_t)k + 1);
    w.blk_cnt = cv.take<int32_t>((size_t)nb * k);
    w.order = cv.take<int64_t>(n);
    w.dist = cv.take<float>(n);
    w.chunk_sums = cv.take<double>((size_t)(n / CHUNK + k + 1) * d);
    w.assign_ws = cv.take<char>(assign_workspace_bytes(n, d, k, mode));
    w.bytes = cv.bytes();
    return w;
}

__global__ void k_lloyd_init(int32_t *done, double *prev_cost, int32_t *n_iter, int32_t *converged) {
    *done = 0;
    *prev_cost = 0.0;
    *n_iter = 0;
    *converged = 0;
}

// One thread per frame.  HIST: the frame's count in its block's per-label
// counts (before the center update).  Otherwise: the exact distance to the
// center of its label (the sequential FP32 loop; VEC as in k_kmpp_d2, with C
// 16-byte aligned too), run after the update so that the cost is that of the
// new clusters against their means.  The block sums of the distances (the
// oracle's cost order) follow in k_block_sums.
template <bool VEC, bool HIST>
__global__ __launch_bounds__(256) void k_lloyd_dist(const float *__restrict__ X, int64_t n, int d,
                                                    const float *__restrict__ C, int k,
                                                    const int32_t *__restrict__ labels, float *__restrict__ dist,
                                                    int32_t *__restrict__ blk_cnt, const int32_t *__restrict__ done) {
    if (*done) return;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int lab = labels[i];
    if constexpr (HIST) {
        // histogram of labels per block (zeroed by k_fill_i32 beforehand), one
        // atomic per run of equal labels in the wave: consecutive frames of a
        // trajectory mostly share their label, and 64 lanes adding 1 to one
        // counter serialize.  A wave's frames lie in one block (BLOCK % 64 == 0);
        // the active lanes are 0 .. m - 1 (only a final partial wave has m < 64).
        const int lane = threadIdx.x & 63;
        const unsigned long long active = __ballot(1);
        const int prev = __shfl_up(lab, 1, 64);
        const bool start = lane == 0 || prev != lab;
        const unsigned long long starts = __ballot(start);
        if (start) {
            const unsigned long long later = starts & ~((2ull << lane) - 1ull);  // starts after this lane
            const int end = later ? __builtin_ctzll(later) : __popcll(active);
            atomicAdd(&blk_cnt[(i / BLOCK) * (int64_t)k + lab], end - lane);
        }
        return;
    }
    const float *x = X + i * d;
    const float *c = C + (int64_t)lab * d;
    float di;
    if constexpr (VEC) {
        float acc = 0.0f;
        auto step4 = [&](const float4 xv, const float4 cv) __attribute__((always_inline)) {
            float t = __fsub_rn(xv.x, cv.x);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.y, cv.y);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.z, cv.z);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
            t = __fsub_rn(xv.w, cv.w);
            acc = __fadd_rn(acc, __fmul_rn(t, t));
        };
        // dims in order; the loads of 16 dims are issued before their terms
        int q = 0;
        for (; q + 16 <= d; q += 16) {
            float4 xv[4], cv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                xv[u] = *reinterpret_cast<const float4 *>(x + q + 4 * u);
                cv[u] = *reinterpret_cast<const float4 *>(c + q + 4 * u);
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) step4(xv[u], cv[u]);
        }
        for (; q < d; q += 4)
            step4(*reinterpret_cast<const float4 *>(x + q), *reinterpret_cast<const float4 *>(c + q));
        di = acc;
    } else {
        di = dist_seq(x, c, d);
    }
    dist[i] = di;
}

// Per label: turn the blk_cnt column into exclusive prefixes over blocks and
// total it (cnt); then member offsets off[j] and chunk offsets coff[j]
// (exclusive scans over k).  One workgroup.
__global__ __launch_bounds__(1024) void k_lloyd_scan(int k, int64_t nb, int64_t *__restrict__ cnt,
                                                     int64_t *__restrict__ off, int64_t *__restrict__ coff,
                                                     int32_t *__restrict__ blk_cnt, const int32_t *__restrict__ done) {
    if (*done) return;
    for (int j = threadIdx.x; j < k; j += blockDim.x) {
        int32_t run = 0;
#pragma unroll 8
        for (int64_t b = 0; b < nb; ++b) {
            const int32_t v = blk_cnt[b * k + j];
            blk_cnt[b * k + j] = run;
            run += v;
        }
        cnt[j] = run;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t a = 0, c = 0;
        for (int j = 0; j < k; ++j) {
            off[j] = a;
            coff[j] = c;
            a += cnt[j];
            c += (cnt[j] + CHUNK - 1) / CHUNK;
        }
        off[k] = a;
        coff[k] = c;
    }
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *p, int64_t n, const int32_t *done) {
    if (*done) return;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// Stable scatter: position of frame i = off[label] + (members of label in
// earlier blocks) + (members in this block before i).  One wavefront per
// block walks its frames in groups of 64 in order.
__global__ __launch_bounds__(64) void k_lloyd_scatter(int64_t n, int k, const int32_t *__restrict__ labels,
                                                      const int64_t *__restrict__ off,
                                                      const int32_t *__restrict__ blk_cnt,
                                                      int64_t *__restrict__ order, const int32_t *__restrict__ done) {
    if (*done) return;
    extern __shared__ int32_t s_run[];  // k running counters of this block
    const int lane = threadIdx.x;
    const int64_t b = blockIdx.x;
    for (int j = lane; j < k; j += 64) s_run[j] = 0;
    __syncthreads();
    const int64_t lo = b * BLOCK;
    const int32_t *bp = blk_cnt + b * (int64_t)k;
    for (int m = 0; m < BLOCK / 64; ++m) {
        const int64_t i = lo + m * 64 + lane;
        const int lab = i < n ? labels[i] : -1;
        int rank = 0;
        for (int src = 0; src < 64; ++src) {
            const int ol = __shfl(lab, src, 64);
            rank += (src < lane && ol == lab) ? 1 : 0;
        }
        if (lab >= 0) order[off[lab] + bp[lab] + s_run[lab] + rank] = i;
        __syncthreads();
        if (lab >= 0) atomicAdd(&s_run[lab], 1);
        __syncthreads();
    }
}

// Chunk sums: one wavefront per chunk, lanes over dimensions, members in order.
__global__ __launch_bounds__(256) void k_lloyd_chunks(const float *__restrict__ X, int d, int k,
                                                      const int64_t *__restrict__ off,
                                                      const int64_t *__restrict__ coff,
                                                      const int64_t *__restrict__ order,
                                                      double *__restrict__ chunk_sums,
                                                      const int32_t *__restrict__ done) {
    if (*done) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t ch = (int64_t)blockIdx.x * 4 + wave;
    if (ch >= coff[k]) return;
    // center owning this chunk: largest j with coff[j] <= ch
    int lo = 0, hi = k - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (coff[mid] <= ch) lo = mid; else hi = mid - 1;
    }
    const int j = lo;
    const int64_t m0 = off[j] + (ch - coff[j]) * CHUNK;
    const int64_t m1 = m0 + CHUNK < off[j + 1] ? m0 + CHUNK : off[j + 1];
    for (int i0 = 0; i0 < d; i0 += 64) {
        const int i = i0 + lane;
        double a = 0.0;
        if (i < d) {
            // members in order; the loads of 8 members are issued before
            // their adds (same sum, one memory round trip per 8)
            int64_t m = m0;
            for (; m + 8 <= m1; m += 8) {
                float v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u) v[u] = X[order[m + u] * d + i];
#pragma unroll
                for (int u = 0; u < 8; ++u) a = a + (double)v[u];
            }
            for (; m < m1; ++m) a = a + (double)X[order[m] * d + i];
        }
        if (i < d) chunk_sums[ch * d + i] = a;
    }
}

// New centers (one wavefront per center) and the convergence bookkeeping.
__global__ __launch_bounds__(256) void k_lloyd_update(int d, int k, const int64_t *__restrict__ cnt,
                                                      const int64_t *__restrict__ coff,
                                                      const double *__restrict__ chunk_sums, float *__restrict__ C,
                                                      const int32_t *__restrict__ done) {
    if (*done) return;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + wave;
    if (j >= k || cnt[j] == 0) return;
    const double cntd = (double)cnt[j];
    for (int i0 = 0; i0 < d; i0 += 64) {
        const int i = i0 + lane;
        if (i >= d) continue;
        double a = 0.0;
        for (int64_t ch = coff[j]; ch < coff[j + 1]; ++ch) a = a + chunk_sums[ch * d + i];
        C[(int64_t)j * d + i] = (float)(a / cntd);
    }
}

__global__ void k_lloyd_cost_final(int64_t nb, const double *__restrict__ bcost, double tol, int it,
                                   double *__restrict__ prev_cost, double *__restrict__ inertias,
                                   int32_t *__restrict__ n_iter, int32_t *__restrict__ converged,
                                   int32_t *__restrict__ done) {
    if (*done) return;
    double cost = 0.0;
    for (int64_t b = 0; b < nb; ++b) cost = cost + bcost[b];
    const double prev = *prev_cost;
    const double rel = cost != 0.0 ? fabs(cost - prev) / cost : 0.0;
    *prev_cost = cost;
    inertias[it] = cost;
    *n_iter = it + 1;
    if (rel <= tol) {
        *converged = 1;
        *done = 1;
    }
}

}  // namespace km
}  // namespace emma

using namespace emma;

extern "C" size_t emma_kmeanspp_workspace_size(int64_t n, int32_t d, int32_t k) {
    (void)d;
    (void)k;
    if (n <= 0) return 0;
    return km::carve_pp(nullptr, n).bytes;
}

extern "C" int emma_kmeanspp(const float *X, int64_t n, int32_t d, int32_t k, uint64_t seed, float *C_out,
                             int64_t *idx_out, void *workspace, size_t workspace_bytes, void *stream) {
    EMMA_REQUIRE(n > 0 && d > 0 && k > 0, "emma_kmeanspp: bad sizes n=%lld d=%d k=%d", (long long)n, d, k);
    EMMA_REQUIRE(X && C_out && idx_out, "emma_kmeanspp: NULL pointer");
    EMMA_REQUIRE(d <= 16384, "emma_kmeanspp: d=%d too large", d);
    const size_t need = km::carve_pp(nullptr, n).bytes;
    EMMA_REQUIRE(workspace && workspace_bytes >= need, "emma_kmeanspp: workspace too small (%zu < %zu)",
                 workspace_bytes, need);
    km::PpWs w = km::carve_pp(workspace, n);
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = cdiv(n, km::BLOCK);
    km::k_kmpp_first<<<1, 256, 0, s>>>(X, n, d, seed, C_out, idx_out);
    EMMA_TRY(check_launch("kmeans++ first"));
    const bool vec = (d & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0;
    const size_t shm = sizeof(float) * (size_t)d;
    for (int r = 1; r < k; ++r) {
        const float *c = C_out + (int64_t)(r - 1) * d;
        if (vec)
            km::k_kmpp_d2<true><<<(unsigned)cdiv(n, 256), 256, shm, s>>>(X, n, d, c, r == 1 ? 1 : 0, w.d2);
        else
            km::k_kmpp_d2<false><<<(unsigned)cdiv(n, 256), 256, shm, s>>>(X, n, d, c, r == 1 ? 1 : 0, w.d2);
        km::k_block_sums<<<(unsigned)cdiv(nb, 4), 256, 0, s>>>(n, w.d2, w.bsum, nullptr);
        km::k_kmpp_pick<<<1, 64, 0, s>>>(X, n, d, seed, r, w.d2, w.bsum, w.sbsum, C_out, idx_out);
    }
    return check_launch("kmeans++ rounds");
}

extern "C" size_t emma_lloyd_workspace_size(int64_t n, int32_t d, int32_t k, int32_t assign_mode) {
    if (n <= 0 || d <= 0 || k <= 0) return 0;
    return km::carve_lloyd(nullptr, n, d, k, assign_mode).bytes;
}

extern "C" int emma_lloyd(const float *X, int64_t n, int32_t d, float *C, int32_t k, int32_t max_iter, double tol,
                          int32_t assign_mode, int32_t *labels, double *inertias, int32_t *n_iter,
                          int32_t *converged, void *workspace, size_t workspace_bytes, void *stream) {
    EMMA_REQUIRE(n > 0 && d > 0 && k > 0 && max_iter >= 0, "emma_lloyd: bad sizes n=%lld d=%d k=%d max_iter=%d",
                 (long long)n, d, k, max_iter);
    EMMA_REQUIRE(assign_mode == EMMA_ASSIGN_FP32_PLAIN || assign_mode == EMMA_ASSIGN_BF16X3_RECHECK,
                 "emma_lloyd: unknown assign mode %d", assign_mode);
    EMMA_REQUIRE(X && C && labels && n_iter && converged && (inertias || max_iter == 0), "emma_lloyd: NULL pointer");
    EMMA_REQUIRE(k <= 12288, "emma_lloyd: k=%d too large for the scatter kernel", k);
    const size_t need = km::carve_lloyd(nullptr, n, d, k, assign_mode).bytes;
    EMMA_REQUIRE(workspace && workspace_bytes >= need, "emma_lloyd: workspace too small (%zu < %zu)",
                 workspace_bytes, need);
    km::LloydWs w = km::carve_lloyd(workspace, n, d, k, assign_mode);
    hipStream_t s = (hipStream_t)stream;
    const int64_t nb = cdiv(n, km::BLOCK);
    km::k_lloyd_init<<<1, 1, 0, s>>>(w.done, w.prev_cost, n_iter, converged);
    const bool vec = (d & 3) == 0 && (reinterpret_cast<uintptr_t>(X) & 15) == 0 &&
                     (reinterpret_cast<uintptr_t>(C) & 15) == 0;
    for (int it = 0; it < max_iter; ++it) {
        EMMA_TRY(assign_run(X, n, d, C, k, labels, assign_mode, nullptr, w.assign_ws, w.done, s));
        km::k_fill_i32<<<1024, 256, 0, s>>>(w.blk_cnt, nb * (int64_t)k, w.done);
        km::k_lloyd_dist<false, true><<<(unsigned)cdiv(n, 256), 256, 0, s>>>(X, n, d, C, k, labels, w.dist,
                                                                            w.blk_cnt, w.done);
        km::k_lloyd_scan<<<1, 1024, 0, s>>>(k, nb, w.cnt, w.off, w.coff, w.blk_cnt, w.done);
        km::k_lloyd_scatter<<<(unsigned)nb, 64, sizeof(int32_t) * k, s>>>(n, k, labels, w.off, w.blk_cnt, w.order,
                                                                          w.done);
        const int64_t max_chunks = n / km::CHUNK + k + 1;
        km::k_lloyd_chunks<<<(unsigned)cdiv(max_chunks, 4), 256, 0, s>>>(X, d, k, w.off, w.coff, w.order,
                                                                         w.chunk_sums, w.done);
        // move the centers to the member means, then take the cost of the new
        // clusters against them and record it (which may raise the convergence flag)
        km::k_lloyd_update<<<(unsigned)cdiv(k, 4), 256, 0, s>>>(d, k, w.cnt, w.coff, w.chunk_sums, C, w.done);
        if (vec)
            km::k_lloyd_dist<true, false><<<(unsigned)cdiv(n, 256), 256, 0, s>>>(X, n, d, C, k, labels, w.dist,
                                                                                w.blk_cnt, w.done);
        else
            km::k_lloyd_dist<false, false><<<(unsigned)cdiv(n, 256), 256, 0, s>>>(X, n, d, C, k, labels, w.dist,
                                                                                 w.blk_cnt, w.done);
        km::k_block_sums<<<(unsigned)cdiv(nb, 4), 256, 0, s>>>(n, w.dist, w.bcost, w.done);
        km::k_lloyd_cost_final<<<1, 1, 0, s>>>(nb, w.bcost, tol, it, w.prev_cost, inertias, n_iter, converged,
                                               w.done);
        EMMA_TRY(check_launch("lloyd iteration"));
    }
    return EMMA_OK;
}
