// Lagged (co)moments of many trajectories on MFMA.
//
// Algorithm (restating moments_XXXY / moments_block / moments_XX and the Chan
// merge of the reference, moments.py:601-973, running_moments.py:35-61, in a
// form that needs only additive partials):
//
//   shift K (a frame average), x' = x - K, y' = y - K for every lagged pair
//   reversible:  U = x' + y', V = x - y;  UU = sum U U^T, VV = sum V V^T
//                raw_xx = (UU + VV)/2 = sum x'x'^T + y'y'^T
//                raw_xy = (UU - VV)/2 = sum x'y'^T + y'x'^T
//   otherwise:   raw_xx = sum x'x'^T, raw_xy = sum x'y'^T, raw_yy = sum y'y'^T
//   S = column sums of x' (and y'); the mean-free matrices are
//   raw - S S^T / w (Chan's formula collapsed to one global shift), the raw
//   sums are S + w K.
//
// Parallel layout: the d x d outputs are cut into TB x TB tiles (tasks); the
// pair index space is cut into slices.  One workgroup = (task, slice).
//
// Precisions and kernels:
//   BF16X4  k_rev_bf16 (every kind of moments): operands split into bf16 hi +
//           lo, products hh, hl, lh, ll on v_mfma_f32_32x32x16_bf16, FP32 tile
//           accumulation (two-level for BF16X4_ACC), TB = 128, 8 waves, two
//           LDS buffers, frame loads four stages ahead
//   FP32    k_main: v_mfma_f32_32x32x2_f32, FP32 tile accumulation, TB = 128
//   FP64    k_main: v_mfma_f64_16x16x4_f64 on FP64 operands (x - K exact in
//           FP64), FP64 accumulation, TB = 64
// k_main stages the operand columns of its two column blocks in LDS per stage
// of BT = 16 pairs, then 4 waves (2 x 2) run MFMA tiles; the frame loads of
// stage s+1 are in flight while stage s runs on the matrix cores.
// FP32 accumulators are flushed every FLUSH pairs into the workgroup's private
// FP64 slot in the workspace; a final kernel sums the slots in slice order
// (deterministic) and applies the FP64 corrections.
#include <type_traits>

#include "emma_common.hpp"

namespace emma {
namespace mom {

constexpr int NTHR = 256;    // threads per workgroup
constexpr int FLUSH = 1024;  // pairs per FP32 accumulation span
constexpr int BT = 16;       // pairs per stage

enum Kind { REV = 0, NR_SYM = 1, NR_XY = 2, LAG0 = 3 };

typedef double doublex4 __attribute__((ext_vector_type(4)));

struct Params {
    const float *X;
    int64_t n_rows;
    int d;
    const int64_t *traj_off;
    int ntraj;
    int lag, stride, skip;
    int tb;               // tile edge
    int acc2;             // bf16 kernel: two-level FP32 accumulation
    int frag;             // slot layout: 1 = MFMA fragment order (TB 128), 0 = row-major
    int64_t slot_doubles; // doubles per slot: 2 matrices + 2 column-sum vectors
    int nb;               // column blocks = ceil(d / tb)
    int n_sym;            // number of I <= J tasks
    int n_tasks;
    int n_slices;
    int rev;              // reversible
    int ctt;              // compute Y'Y in NR_SYM tasks
    int has_xy;           // non-reversible XY tasks present
    // workspace
    int64_t *P_total;     // [1]
    int64_t *pair_off;    // [ntraj + 1]
    float *K;             // [d]
    const double *wts;    // per-row frame weights (nullptr: unweighted)
    double *wpart;        // [ntraj] weight sum of each trajectory's pairs
    float *wpair;         // [P] weight of each pair as FP32, in pair order (weighted)
    double *W;            // [1] total pair weight
    double *slots;        // n_tasks * n_slices * slot_doubles
};

__device__ __forceinline__ void task_of(const Params &p, int t, int &I, int &J) {
    if (t < p.n_sym) {
        int i = 0, rem = t;
        while (rem >= p.nb - i) {
            rem -= p.nb - i;
            ++i;
        }
        I = i;
        J = i + rem;
    } else {
        const int u = t - p.n_sym;
        I = u / p.nb;
        J = u - I * p.nb;
    }
}

// ------------------------------------------------------------------
// prep: pair offsets (exclusive scan of per-trajectory pair counts), shift K
// ------------------------------------------------------------------
__device__ __forceinline__ int64_t traj_pairs(const Params &p, int t) {
    const int64_t T = p.traj_off[t + 1] - p.traj_off[t];
    const int64_t start = (int64_t)p.skip + p.lag;
    if (T <= start) return 0;
    return (T - start + p.stride - 1) / p.stride;
}

__global__ __launch_bounds__(1024) void k_prep(Params p) {
    __shared__ int64_t s_part[1024];
    __shared__ int64_t s_carry;
    const int tid = threadIdx.x;
    if (tid == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < p.ntraj; base += 1024) {
        const int t = base + tid;
        s_part[tid] = t < p.ntraj ? traj_pairs(p, t) : 0;
        __syncthreads();
        for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
            const int64_t a = tid >= off ? s_part[tid - off] : 0;
            __syncthreads();
            s_part[tid] += a;
            __syncthreads();
        }
        if (t < p.ntraj) p.pair_off[t + 1] = s_carry + s_part[tid];
        __syncthreads();
        if (tid == 1023) s_carry += s_part[1023];
        __syncthreads();
    }
    if (tid == 0) {
        p.pair_off[0] = 0;
        *p.P_total = s_carry;
    }
    __syncthreads();
    // shift: average of up to 64 x-frames spread over the pair index space
    const int64_t P = s_carry;
    const int ns = P < 64 ? (int)P : 64;
    for (int c = tid; c < p.d; c += 1024) {
        double acc = 0.0;
        int t = 0;
        for (int s = 0; s < ns; ++s) {
            const int64_t q = (int64_t)s * P / ns;
            while (p.pair_off[t + 1] <= q) ++t;
            const int64_t row = p.traj_off[t] + p.skip + (q - p.pair_off[t]) * p.stride;
            acc += (double)p.X[row * p.d + c];
        }
        p.K[c] = ns > 0 ? (float)(acc / ns) : 0.0f;
    }
}

// ------------------------------------------------------------------
// main kernel
// ------------------------------------------------------------------
template <int PREC>
struct Cfg;
template <>
struct Cfg<EMMA_PREC_FP32> {
    static constexpr int TB = 128;
    static constexpr int PITCH = BT + 1;  // odd: conflict-free column reads
    typedef float elem;
    static constexpr int PARTS = 1;
};
template <>
struct Cfg<EMMA_PREC_FP64> {
    static constexpr int TB = 64;
    static constexpr int PITCH = BT + 1;
    typedef double elem;
    static constexpr int PARTS = 1;
};

template <int PREC, int KIND, int NMAT, bool WGT>
__global__ __launch_bounds__(NTHR) void k_main(Params p, int task_base, int n_launch_tasks) {
    using C = Cfg<PREC>;
    constexpr int TB = C::TB;
    constexpr int PITCH = C::PITCH;
    constexpr int TPC = NTHR / (2 * TB);  // loader threads per column
    constexpr int PPT = BT / TPC;         // pairs per loader thread
    constexpr bool SYM = KIND != NR_XY;
    constexpr bool F64 = PREC == EMMA_PREC_FP64;
    typedef typename C::elem elem;

    __shared__ __align__(16) elem s_op[4][C::PARTS][TB * PITCH];  // slots A1, B1, A2, B2
    __shared__ int64_t s_xo[2][BT];                               // element offset of the x row (-1: none)
    __shared__ int s_tr[2][BT];                                   // trajectory of each pair
    __shared__ double s_w[2][BT];                                 // weight of each pair (WGT)
    __shared__ double s_sum[2][TPC][TB];                          // column-sum partials (diagonal tasks)

    // XCD-aware numbering: workgroups are dealt round-robin to the 8 XCDs, so
    // give each XCD a contiguous range of (slice, task) pairs; the tasks of one
    // slice then share frame rows through that XCD's L2.
    const int per_xcd = gridDim.x >> 3;
    const int wg = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (wg >= n_launch_tasks * p.n_slices) return;
    const int task = task_base + wg % n_launch_tasks;
    const int slice = wg / n_launch_tasks;
    int I, J;
    task_of(p, task, I, J);
    const bool diag = SYM && (I == J);

    const int64_t P = *p.P_total;
    const int64_t per = (P + p.n_slices - 1) / p.n_slices;
    const int64_t per_r = (per + BT - 1) / BT * BT;
    const int64_t p_lo = (int64_t)slice * per_r;
    const int64_t p_hi = p_lo + per_r < P ? p_lo + per_r : P;
    double *slot = p.slots + ((int64_t)slice * p.n_tasks + task) * p.slot_doubles;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int wr = wave >> 1, wc = wave & 1;

    // loader role: column lcol of block I (half 0) or J (half 1), pairs pq*PPT ..
    const int half = tid / (TB * TPC);
    const int lcol = (tid / TPC) % TB;
    const int pq = tid % TPC;
    const int gcol_raw = (half == 0 ? I : J) * TB + lcol;
    const bool col_ok = gcol_raw < p.d;
    const int gcol = col_ok ? gcol_raw : p.d - 1;  // always a valid address
    const float Kc = p.K[gcol];
    const int64_t lagd = (int64_t)p.lag * p.d;
    double sum_x = 0.0, sum_y = 0.0;

    // A wave computes all subtiles of its (TB/2 x TB/2) block or none: its block
    // must intersect the valid d x d region and, on a diagonal tile, lie on or
    // above the diagonal.  (Per-subtile predicates inside the stage loop make
    // the compiler shuttle the accumulators between register files.)
    constexpr int WTE = TB / 2;  // wave tile edge
    const bool wave_act = (I * TB + wr * WTE < p.d) && (J * TB + wc * WTE < p.d) && !(diag && wr > wc);

    // ---- pair table: offsets of the x rows of one stage ----
    auto build_table = [&](int b, int64_t s0, int t_start) __attribute__((always_inline)) {
        if (tid < BT) {
            const int64_t q = s0 + tid;
            int t = t_start;
            int64_t xo = -1;
            double wq = 0.0;
            if (q < p_hi) {
                while (p.pair_off[t + 1] <= q) ++t;
                const int64_t row = p.traj_off[t] + p.skip + (q - p.pair_off[t]) * p.stride;
                xo = row * p.d;
                if constexpr (WGT) wq = p.wts[row];
            }
            s_xo[b][tid] = xo;
            s_tr[b][tid] = t;
            if constexpr (WGT) s_w[b][tid] = wq;
        }
    };
    if (tid == 0) {
        int lo = 0, hi = p.ntraj - 1;  // trajectory of the slice's first pair
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.pair_off[mid] <= p_lo) lo = mid; else hi = mid - 1;
        }
        s_tr[1][BT - 1] = lo;
    }
    __syncthreads();
    build_table(0, p_lo, s_tr[1][BT - 1]);
    __syncthreads();

    // ---- prefetch registers ----
    float xr[PPT], yr[PPT];
    double wgt[WGT ? PPT : 1];  // pair weights of the prefetched stage
    uint32_t vmask = 0;
    auto issue_loads = [&](int b) __attribute__((always_inline)) {
        vmask = 0;
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            if constexpr (WGT) wgt[q] = s_w[b][pq * PPT + q];
            const int64_t xo = s_xo[b][pq * PPT + q];
            const bool v = xo >= 0;
            vmask |= (v ? 1u : 0u) << q;
            xr[q] = p.X[(v ? xo : 0) + gcol];
            if constexpr (KIND != LAG0) yr[q] = p.X[(v ? xo + lagd : 0) + gcol];
            else yr[q] = 0.0f;
        }
    };
    issue_loads(0);
    build_table(1, p_lo + BT, s_tr[0][BT - 1]);

    const int sl1 = half == 0 ? 0 : 1;  // slots: 0 = A1 (I), 1 = B1 (J), 2 = A2 (I), 3 = B2 (J)
    const int sl2 = half == 0 ? 2 : 3;
    int stage = 0;

    // ---- operands of one stage from the prefetched values -> LDS ----
    auto store_operands = [&]() __attribute__((always_inline)) {
        float o1[PPT], o2[PPT];
        double d1[PPT], d2[PPT];
        float bx = 0.0f, by = 0.0f;
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            const bool ok = ((vmask >> q) & 1u) && col_ok;
            const float x = ok ? xr[q] : 0.0f;
            const float y = ok ? yr[q] : 0.0f;
            // frame weights multiply the A operand (block I loaders) and the sums
            if constexpr (F64) {
                const double xs = ok ? (double)x - (double)Kc : 0.0;
                const double ys = ok ? (double)y - (double)Kc : 0.0;
                const double wa = (WGT && half == 0) ? wgt[WGT ? q : 0] : 1.0;
                sum_x += WGT ? wgt[WGT ? q : 0] * xs : xs;
                sum_y += WGT ? wgt[WGT ? q : 0] * ys : ys;
                if constexpr (KIND == REV) {
                    d1[q] = wa * (xs + ys);
                    d2[q] = wa * ((double)x - (double)y);
                } else if constexpr (KIND == NR_XY) {
                    d1[q] = half == 0 ? wa * xs : ys;
                    d2[q] = 0.0;
                } else {
                    d1[q] = wa * xs;
                    d2[q] = wa * ys;
                }
            } else {
                const float xs = ok ? x - Kc : 0.0f;
                const float ys = ok ? y - Kc : 0.0f;
                const float wq = WGT ? (float)wgt[WGT ? q : 0] : 1.0f;
                const float wa = (WGT && half == 0) ? wq : 1.0f;
                bx += WGT ? wq * xs : xs;
                by += WGT ? wq * ys : ys;
                if constexpr (KIND == REV) {
                    o1[q] = wa * (xs + ys);
                    o2[q] = wa * (x - y);
                } else if constexpr (KIND == NR_XY) {
                    o1[q] = half == 0 ? wa * xs : ys;
                    o2[q] = 0.0f;
                } else {
                    o1[q] = wa * xs;
                    o2[q] = wa * ys;
                }
            }
        }
        if constexpr (!F64) {
            sum_x += (double)bx;
            sum_y += (double)by;
        }
        elem *dst1 = &s_op[sl1][0][lcol * PITCH + pq * PPT];
        elem *dst2 = &s_op[sl2][0][lcol * PITCH + pq * PPT];
        if constexpr (PREC == EMMA_PREC_FP32) {
#pragma unroll
            for (int q = 0; q < PPT; ++q) dst1[q] = o1[q];
            if constexpr (NMAT == 2) {
#pragma unroll
                for (int q = 0; q < PPT; ++q) dst2[q] = o2[q];
            }
        } else {
#pragma unroll
            for (int q = 0; q < PPT; ++q) dst1[q] = d1[q];
            if constexpr (NMAT == 2) {
#pragma unroll
                for (int q = 0; q < PPT; ++q) dst2[q] = d2[q];
            }
        }
    };

    // ---- MFMA of one stage; FIRST starts the accumulation from a zero C operand ----
    floatx16 acc[NMAT][2][2];
    doublex4 acc64[NMAT][2][2];
    auto mfma_stage = [&](auto first_tag) __attribute__((always_inline)) {
        constexpr bool FIRST = decltype(first_tag)::value;
#pragma unroll
        for (int m = 0; m < NMAT; ++m) {
            const int sa = m == 0 ? 0 : 2, sb = m == 0 ? 1 : 3;
            if constexpr (PREC == EMMA_PREC_FP32) {
#pragma unroll
                for (int kk = 0; kk < BT; kk += 2) {
                    float av[2], bv[2];
#pragma unroll
                    for (int r = 0; r < 2; ++r)
                        av[r] = s_op[sa][0][(wr * 64 + r * 32 + (lane & 31)) * PITCH + kk + (lane >> 5)];
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        bv[c] = s_op[sb][0][(wc * 64 + c * 32 + (lane & 31)) * PITCH + kk + (lane >> 5)];
#pragma unroll
                    for (int r = 0; r < 2; ++r)
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            if (FIRST && kk == 0) {
                                const floatx16 z = {};
                                acc[m][r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[r], bv[c], z, 0, 0, 0);
                            } else {
                                acc[m][r][c] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[r], bv[c], acc[m][r][c], 0, 0, 0);
                            }
                        }
                }
            } else {
#pragma unroll
                for (int kk = 0; kk < BT; kk += 4) {
                    double av[2], bv[2];
#pragma unroll
                    for (int r = 0; r < 2; ++r)
                        av[r] = s_op[sa][0][(wr * 32 + r * 16 + (lane & 15)) * PITCH + kk + (lane >> 4)];
#pragma unroll
                    for (int c = 0; c < 2; ++c)
                        bv[c] = s_op[sb][0][(wc * 32 + c * 16 + (lane & 15)) * PITCH + kk + (lane >> 4)];
#pragma unroll
                    for (int r = 0; r < 2; ++r)
#pragma unroll
                        for (int c = 0; c < 2; ++c) {
                            if (FIRST && kk == 0) {
                                const doublex4 z = {};
                                acc64[m][r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], z, 0, 0, 0);
                            } else {
                                acc64[m][r][c] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[r], bv[c], acc64[m][r][c], 0, 0, 0);
                            }
                        }
                }
            }
        }
    };

    // ---- one stage: operands -> LDS, loads of the next stage, MFMA, table ----
    auto run_stage = [&](int64_t s0, auto mf_tag, auto first_tag) __attribute__((always_inline)) {
        const int b = stage & 1;
        store_operands();
        __syncthreads();     // operands of stage s and the table of stage s+1 are visible
        issue_loads(b ^ 1);  // stage s+1 in flight during the MFMAs below
        if constexpr (decltype(mf_tag)::value) mfma_stage(first_tag);
        build_table(b, s0 + 2 * BT, s_tr[b ^ 1][BT - 1]);  // stage s+2 into the buffer of stage s
        __syncthreads();     // MFMA reads of the operands done before they are overwritten
        ++stage;
    };

    // ---- flush into the workgroup's FP64 slot ----
    auto flush = [&](bool first_flush) __attribute__((always_inline)) {
        if constexpr (F64) {
            // row-major slot, written once at the end.  16x16x4 f64 output layout:
            // lane l, element e holds row (l >> 4) + 4 e, column l & 15.
#pragma unroll
            for (int m = 0; m < NMAT; ++m)
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const int j = wc * 32 + c * 16 + (lane & 15);
                        double *dst = slot + (int64_t)m * TB * TB + (int64_t)(wr * 32 + r * 16 + (lane >> 4)) * TB + j;
#pragma unroll
                        for (int e = 0; e < 4; ++e) dst[(int64_t)(4 * e) * TB] = acc64[m][r][c][e];
                    }
        } else {
            // fragment-order slot: element e of the 32x32 tile's lanes is 64
            // consecutive doubles, so one instruction covers whole cache lines
#pragma unroll
            for (int m = 0; m < NMAT; ++m)
#pragma unroll
                for (int r = 0; r < 2; ++r)
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        double *dst = slot + (int64_t)m * TB * TB + (int64_t)(wave * 4 + r * 2 + c) * 1024 + lane;
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
                            const double v = (double)acc[m][r][c][e];
                            // the slot is private to this workgroup and every element has one
                            // writer lane: the add needs no atomicity; the L2-side FP64 add
                            // keeps the old value out of registers
                            if (first_flush)
                                dst[e * 64] = v;
                            else
                                unsafeAtomicAdd(dst + e * 64, v);
                        }
                    }
        }
    };

    // ---- slice loop: spans of FLUSH pairs (FP64: one span) ----
    typedef std::integral_constant<bool, true> T_;
    typedef std::integral_constant<bool, false> F_;
    if (wave_act) {
        bool first_flush = true;
        int64_t f0 = p_lo;
        do {
            const int64_t f1 = F64 ? p_hi : (f0 + FLUSH < p_hi ? f0 + FLUSH : p_hi);
            run_stage(f0, T_(), T_());
            for (int64_t s0 = f0 + BT; s0 < f1; s0 += BT) run_stage(s0, T_(), F_());
            flush(first_flush);
            first_flush = false;
            f0 = f1;
        } while (f0 < p_hi);
    } else {
        // same loads and barriers, no matrix work
        int64_t f0 = p_lo;
        do {
            const int64_t f1 = F64 ? p_hi : (f0 + FLUSH < p_hi ? f0 + FLUSH : p_hi);
            run_stage(f0, F_(), T_());
            for (int64_t s0 = f0 + BT; s0 < f1; s0 += BT) run_stage(s0, F_(), F_());
            f0 = f1;
        } while (f0 < p_hi);
    }

    // column sums (diagonal tasks own them), combined over the loader threads of a column
    if (diag) {
        if (half == 0) {
            s_sum[0][pq][lcol] = sum_x;
            s_sum[1][pq][lcol] = sum_y;
        }
        __syncthreads();
        if (tid < TB) {
            double ax = 0.0, ay = 0.0;
#pragma unroll
            for (int u = 0; u < TPC; ++u) {
                ax += s_sum[0][u][tid];
                ay += s_sum[1][u][tid];
            }
            double *sums = slot + 2LL * TB * TB;
            sums[tid] = ax;
            sums[TB + tid] = ay;
        }
    }
}

// ------------------------------------------------------------------
// bf16 hi/lo split moments: the pipelined kernel (TICA, VAMP, PCA, weighted)
// ------------------------------------------------------------------
// One workgroup = (task, slice) as above, 8 waves.  Per stage of RBT = 16
// pairs the loader threads read x and y of their columns, form U = x' + y',
// V = x - y, split both into bf16 hi + lo and write them to one of two LDS
// buffers while the waves run the MFMAs of the previous stage from the other
// buffer; one barrier per stage.  The raw frame values of a stage are loaded
// RSETS = 4 stages ahead into one of four register sets (two with frame
// weights, which need more registers).  Inside a stage the
// 16 MFMAs of a wave are issued between the pieces of the split arithmetic
// (scheduling barriers fix the order), so that the vector ALU works while the
// matrix core runs.  Wave roles:
//   off-diagonal task (I < J): each wave computes one matrix (UU | VV) on one
//     64 x 64 block of the tile (the blocks with matrix work on waves 0, 1, 2,
//     ... first); loader thread t owns column t & 127 of block I (group
//     (t >> 7) even) or J (odd), pairs 8 ((t >> 7) >> 1) ..
//   diagonal task (I == J): only block I is staged and serves as both MFMA
//     operands; the blocks below the diagonal are skipped; loader thread t
//     owns column t & 127, pairs 4 (t >> 7) ..
// The rows of a pair come from a per-wave cursor over the trajectory table
// (the pairs of a wave are the same for all its lanes and increase
// monotonically), so no pair table is staged.
constexpr int RNT = 512;             // threads per workgroup
constexpr int RTB = 128;             // tile edge
constexpr int RBT = 16;              // pairs per stage
constexpr int RPITCH = RBT + 8;      // uint16 per column row (48 B: conflict-free 16-byte reads)
constexpr int RSUB = 4;              // stages per loop iteration = per FP32 sub-span (two-level)
// pairs per FP32 accumulation span: the two-level mode adds one sub-span sum
// per RSUB stages to its span sum, the single-level mode every product
constexpr int RFLUSH_TL = 4096;
constexpr int RFLUSH_SL = 1024;

struct RevLds {
    uint16_t op[2][4][2][RTB * RPITCH];  // [buffer][U_I, U_J, V_I, V_J][hi, lo][column row]
};

// scalar (wave-uniform) cursor: pair index -> x row
struct PairCursor {
    int t;
    int64_t p0, p1, base;
};

// MODE 0 (REV): UU and VV as above.  MODE 1 (NRS): the non-reversible
// symmetric matrices instead (XX from x' = x - K in place of U, YY from
// y' = y - K in place of V; column sums of x' and y'), the C00 / Ctt part of
// the non-reversible and VAMP moments; without the Ctt flag the YY waves only
// load and split.  MODE 2 (XY): the lagged products x'y'^T of the tasks
// (I, J) and (J, I) from the same loads: waves 0-3 compute x'_I^T y'_J, waves
// 4-7 x'_J^T y'_I (on a diagonal tile only waves 0-3, the whole tile), each
// into the slot of its own non-symmetric task; no column sums.
// WT: frame weights (the weight of a pair's x frame) multiply the A operand
// and the column sums: off-diagonal tiles weight the block-I operands; a
// diagonal tile writes weighted copies of U and V (x' and y') into the LDS
// slots it does not otherwise use and takes A from them; mode XY weights x'
// (the A operand of both of its tasks) and never y'.
template <bool DIAG, bool TL, int MODE, bool WT>
__device__ __forceinline__ void rev_body(const Params &p, RevLds &lds, double (&s_sum)[2][4][RTB], int I,
                                         int J, int task, int slice) {
    constexpr bool NRS = MODE != 0;  // operands x', y' (not U, V)
    constexpr bool XY = MODE == 2;
    constexpr int PPT = DIAG ? 4 : 8;  // pairs per loader thread per stage
    constexpr int TB = RTB;

    const int64_t P = *p.P_total;
    const int64_t per = (P + p.n_slices - 1) / p.n_slices;
    const int64_t per_r = (per + RBT - 1) / RBT * RBT;
    const int64_t p_lo = (int64_t)slice * per_r < P ? (int64_t)slice * per_r : P;
    const int64_t p_hi = p_lo + per_r < P ? p_lo + per_r : P;
    const int n_stages = (int)((p_hi - p_lo + RBT - 1) / RBT);
    double *slot = p.slots + ((int64_t)slice * p.n_tasks + task) * p.slot_doubles;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    // ---- loader role ----
    const int grp = tid >> 7;  // wave-uniform
    const int lcol = tid & (TB - 1);
    const int blk = DIAG ? I : ((grp & 1) ? J : I);
    // first pair of this thread in a stage (wave-uniform: the cursor lives in scalar registers)
    const int pfirst = __builtin_amdgcn_readfirstlane(DIAG ? grp * PPT : (grp >> 1) * PPT);
    const int su = DIAG ? 0 : (grp & 1);                     // LDS slot of U (V at su + 2)
    const int gcol_raw = blk * TB + lcol;
    // columns past d read column d - 1: the rows and columns of the tile
    // beyond d are never read back, so their values do not matter
    const int gcol = gcol_raw < p.d ? gcol_raw : p.d - 1;
    const float Kc = p.K[gcol];
    const int64_t lagd = (int64_t)p.lag * p.d;

    // ---- MFMA role ----
    // The 8 blocks (matrix m, row half wr, column half wc) of the tile, in the
    // order b -> (b >> 2, (b >> 1) & 1, b & 1) (diagonal tiles: (b & 1,
    // b >> 2, (b >> 1) & 1)); a block has matrix work when it lies inside d
    // (and on or above the diagonal, and is wanted).  The blocks with matrix
    // work go to waves 0, 1, 2, ... and the others after them, so that a tile
    // with fewer than 8 such blocks (the last column block less than half
    // full, or one matrix not wanted) spreads them over the four SIMDs (wave w
    // of the workgroup runs on SIMD w % 4) instead of doubling up on two.
    auto block_of = [&](int b, int &bm, int &bwr, int &bwc) __attribute__((always_inline)) {
        if (DIAG && !XY) {
            bm = b & 1;
            bwr = b >> 2;
            bwc = (b >> 1) & 1;
        } else {
            bm = b >> 2;
            bwr = (b >> 1) & 1;
            bwc = b & 1;
        }
    };
    auto block_active = [&](int bm, int bwr, int bwc) __attribute__((always_inline)) -> bool {
        if constexpr (XY) {
            const int ri = bm == 0 ? I : J, ci = bm == 0 ? J : I;
            return (ri * TB + bwr * 64 < p.d) && (ci * TB + bwc * 64 < p.d) && !(DIAG && bm == 1);
        } else {
            return (I * TB + bwr * 64 < p.d) && (J * TB + bwc * 64 < p.d) && !(DIAG && bwr > bwc) &&
                   !(NRS && bm == 1 && !p.ctt);
        }
    };
    int m, wr, wc;
    bool wave_act;
    {
        int n_act = 0;
        for (int b = 0; b < 8; ++b) {
            int bm, bwr, bwc;
            block_of(b, bm, bwr, bwc);
            n_act += block_active(bm, bwr, bwc) ? 1 : 0;
        }
        int ia = 0, ii = n_act, pick = 0;
        for (int b = 0; b < 8; ++b) {
            int bm, bwr, bwc;
            block_of(b, bm, bwr, bwc);
            const int pos = block_active(bm, bwr, bwc) ? ia++ : ii++;
            if (pos == wave) pick = b;
        }
        block_of(pick, m, wr, wc);
        wave_act = wave < n_act;
    }
    int sa, sb;
    double *mslot = slot + (int64_t)m * TB * TB;  // output matrix of this wave
    if constexpr (XY) {
        const int ri = m == 0 ? I : J, ci = m == 0 ? J : I;  // row (x') and column (y') blocks
        sa = DIAG ? 0 : (m == 0 ? 0 : 1);  // x' of the row block
        sb = DIAG ? 2 : (m == 0 ? 3 : 2);  // y' of the column block
        const int txy = p.n_sym + ri * p.nb + ci;
        mslot = p.slots + ((int64_t)slice * p.n_tasks + txy) * p.slot_doubles;
    } else {
        sa = m == 0 ? 0 : 2;
        sb = DIAG ? sa : sa + 1;
        if (WT && DIAG) sa += 1;  // the weighted copy
    }

    // ---- cursor: trajectory of the slice's first pair ----
    // (the tables are read-only here: uniform reads go through the scalar cache)
    typedef const __attribute__((address_space(4))) int64_t *cptr;
    const cptr poff_c = (cptr)p.pair_off;
    const cptr toff_c = (cptr)p.traj_off;
    PairCursor cur;
    {
        int lo = 0, hi = p.ntraj - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (p.pair_off[mid] <= p_lo) lo = mid; else hi = mid - 1;
        }
        cur.t = lo;
        cur.p0 = p.pair_off[lo];
        cur.p1 = p.pair_off[lo + 1];
        cur.base = p.traj_off[lo] + p.skip;
    }
    // x row (frame index) of pair q (wave-uniform q; -1 if q is past the slice)
    auto row_of = [&](int64_t q) __attribute__((always_inline)) -> int64_t {
        if (q >= p_hi) return -1;
        while (q >= cur.p1) {
            ++cur.t;
            cur.p0 = cur.p1;
            cur.p1 = poff_c[cur.t + 1];
            cur.base = toff_c[cur.t] + p.skip;
        }
        return cur.base + (q - cur.p0) * p.stride;
    };

    // Frame loads of one stage.  The x and y rows of the stage's pairs are
    // wave-uniform pointers (scalar registers), found before the stage's
    // matrix block, and the lane offset is a 32-bit VGPR, so the loads inside
    // the block are one form for every stage.  Common case: the stage's pairs
    // lie in one trajectory and the rows are base + q * stride * d.
    // Otherwise every pair gets its own row; a pair past the slice reads the
    // shift vector K for both x and y, which makes its U and V exactly zero.
    const int64_t step_b = (int64_t)p.stride * p.d * 4;
    const uint32_t coff = (uint32_t)gcol * 4u;
    const char *Xb = reinterpret_cast<const char *>(p.X);
    const char *Kb = reinterpret_cast<const char *>(p.K);
    // first pair of stage t, and whether all its pairs lie in the cursor's trajectory
    auto stage_q0 = [&](int t) __attribute__((always_inline)) -> int64_t { return p_lo + (int64_t)t * RBT + pfirst; };
    auto stage_fast = [&](int64_t q0) __attribute__((always_inline)) -> bool {
        return q0 >= cur.p0 && q0 + PPT <= cur.p1 && q0 + PPT <= p_hi;
    };
    // register sets of raw frame values: stage t is loaded into set t % RSETS
    // (loads RSETS stages ahead; at most 63 vector loads may be in flight)
    constexpr int RSETS = WT ? 2 : 4;
    float xs[RSETS][PPT], ys[RSETS][PPT];
    // WT: lane l of a register set's weight holds the weight of the set's
    // pair l (l < PPT; pairs are consecutive in the pair-ordered weights), so
    // one vector load per stage brings them with the frames; the split reads
    // them back with readlane
    float ws[RSETS][1];
    const char *rx[PPT], *ry[PPT];  // rows of the stage being issued (wave-uniform)
    int64_t q0n = 0;                // first pair of the stage being issued
    // A run of common-case stages: when a stage's pairs lie in the cursor's
    // trajectory, so do those of the next fast_left stages (whose rows follow
    // at RBT * stride rows per stage), so their first rows come from one add
    // instead of the pair -> row multiplications and the range checks.
    int fast_left = 0;
    const char *bx_run = Xb;  // x row of the next stage of the run
    const int64_t stage_b = (int64_t)RBT * step_b;
    auto stage_rows = [&](int t) __attribute__((always_inline)) {
        const int64_t q0 = stage_q0(t);
        q0n = q0;
        bool fast = fast_left > 0;
        const char *bx = bx_run;
        if (fast) {
            --fast_left;
        } else if (stage_fast(q0)) {
            fast = true;
            bx = Xb + (cur.base + (q0 - cur.p0) * p.stride) * p.d * 4;
            const int64_t lim = cur.p1 < p_hi ? cur.p1 : p_hi;
            fast_left = (int)((lim - q0 - PPT) / RBT);
        }
        if (fast) {
            bx_run = bx + stage_b;
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                rx[q] = bx + q * step_b;
                ry[q] = rx[q] + lagd * 4;
            }
        } else {
#pragma unroll
            for (int q = 0; q < PPT; ++q) {
                const int64_t r = row_of(q0 + q);
                rx[q] = r >= 0 ? Xb + r * p.d * 4 : Kb;
                ry[q] = r >= 0 ? rx[q] + lagd * 4 : Kb;
            }
        }
    };
    auto issue_rows = [&](float *xr, float *yr, float *wr) __attribute__((always_inline)) {
        // the 32-bit column offset is made opaque here, so that each load is
        // (scalar row pointer + 32-bit lane offset) and needs no 64-bit vector add
        uint32_t co = coff;
        asm volatile("" : "+v"(co));
#pragma unroll
        for (int q = 0; q < PPT; ++q) {
            xr[q] = *reinterpret_cast<const float *>(rx[q] + co);
            yr[q] = *reinterpret_cast<const float *>(ry[q] + co);
        }
        if constexpr (WT) {
            // pairs past the slice (U = V = 0) read any valid weight
            const int64_t qw = q0n + (lane & (PPT - 1));
            wr[0] = P > 0 ? p.wpair[qw < P ? qw : P - 1] : 0.0f;
        }
    };
    double sum_u = 0.0, sum_v = 0.0;
    // weighted A operand of this loader thread (WT, off-diagonal modes REV / NRS:
    // block I; mode XY: every x')
    const bool wscale = WT && !(DIAG && !XY) && (XY || su == 0);
    // split of the stage being prepared: bf16 hi / lo pairs of U and V (x', y')
    uint32_t uh[PPT / 2], ul[PPT / 2], vh[PPT / 2], vl[PPT / 2];
    uint32_t uwh[PPT / 2], uwl[PPT / 2], vwh[PPT / 2], vwl[PPT / 2];  // weighted copies (WT, diagonal)
    float bu = 0.0f, bv = 0.0f;  // the stage's column sums (diagonal tasks)
    // pairs q, q + 1 of a register set (q even)
    auto prep_pair = [&](const float *xr, const float *yr, const float *wr, int q) __attribute__((always_inline)) {
        float u0, u1, v0, v1;
        if constexpr (NRS) {
            u0 = __fsub_rn(xr[q], Kc);
            u1 = __fsub_rn(xr[q + 1], Kc);
            v0 = __fsub_rn(yr[q], Kc);
            v1 = __fsub_rn(yr[q + 1], Kc);
        } else {
            u0 = __fadd_rn(__fsub_rn(xr[q], Kc), __fsub_rn(yr[q], Kc));
            u1 = __fadd_rn(__fsub_rn(xr[q + 1], Kc), __fsub_rn(yr[q + 1], Kc));
            v0 = __fsub_rn(xr[q], yr[q]);
            v1 = __fsub_rn(xr[q + 1], yr[q + 1]);
        }
        if constexpr (WT) {
            const float w0 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wr[0]), q));
            const float w1 = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(wr[0]), q + 1));
            if constexpr (DIAG && !XY) {
                const float a0 = __fmul_rn(w0, u0), a1 = __fmul_rn(w1, u1);
                const float b0 = __fmul_rn(w0, v0), b1 = __fmul_rn(w1, v1);
                bu = __fadd_rn(bu, __fadd_rn(a0, a1));
                if constexpr (NRS) bv = __fadd_rn(bv, __fadd_rn(b0, b1));
                split_bf16x2(a0, a1, uwh[q / 2], uwl[q / 2]);
                split_bf16x2(b0, b1, vwh[q / 2], vwl[q / 2]);
            } else if (wscale) {
                u0 = __fmul_rn(w0, u0);
                u1 = __fmul_rn(w1, u1);
                if constexpr (!XY) {
                    v0 = __fmul_rn(w0, v0);
                    v1 = __fmul_rn(w1, v1);
                }
            }
        } else if constexpr (DIAG) {
            bu = __fadd_rn(bu, __fadd_rn(u0, u1));
            if constexpr (NRS) bv = __fadd_rn(bv, __fadd_rn(v0, v1));
        }
        split_bf16x2(u0, u1, uh[q / 2], ul[q / 2]);
        split_bf16x2(v0, v1, vh[q / 2], vl[q / 2]);
    };
    auto prep_store = [&](int buf) __attribute__((always_inline)) {
        if constexpr (DIAG) {
            sum_u += (double)bu;
            if constexpr (NRS) sum_v += (double)bv;
            bu = 0.0f;
            bv = 0.0f;
        }
        const int off = lcol * RPITCH + pfirst;
        if constexpr (WT && DIAG && !XY) {  // PPT == 4: the weighted copies into slots 1 and 3
            *reinterpret_cast<uint2 *>(&lds.op[buf][1][0][off]) = make_uint2(uwh[0], uwh[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][1][1][off]) = make_uint2(uwl[0], uwl[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][3][0][off]) = make_uint2(vwh[0], vwh[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][3][1][off]) = make_uint2(vwl[0], vwl[1]);
        }
        if constexpr (PPT == 8) {
            *reinterpret_cast<uint4 *>(&lds.op[buf][su][0][off]) = make_uint4(uh[0], uh[1], uh[2], uh[3]);
            *reinterpret_cast<uint4 *>(&lds.op[buf][su][1][off]) = make_uint4(ul[0], ul[1], ul[2], ul[3]);
            *reinterpret_cast<uint4 *>(&lds.op[buf][su + 2][0][off]) = make_uint4(vh[0], vh[1], vh[2], vh[3]);
            *reinterpret_cast<uint4 *>(&lds.op[buf][su + 2][1][off]) = make_uint4(vl[0], vl[1], vl[2], vl[3]);
        } else {
            *reinterpret_cast<uint2 *>(&lds.op[buf][su][0][off]) = make_uint2(uh[0], uh[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][su][1][off]) = make_uint2(ul[0], ul[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][su + 2][0][off]) = make_uint2(vh[0], vh[1]);
            *reinterpret_cast<uint2 *>(&lds.op[buf][su + 2][1][off]) = make_uint2(vl[0], vl[1]);
        }
    };

    // TL (two-level FP32 accumulation): the MFMAs accumulate a sub-span of
    // RSUB stages from zero (acc: the first product of the sub-span takes a
    // zero accumulator operand), which is then added to the span sum (sacc);
    // the span sum goes to the FP64 slot every RFLUSH_TL pairs.  An element of
    // a span then carries about RSUB * 4 + RFLUSH_TL / (RBT * RSUB) sequential
    // FP32 roundings instead of RFLUSH_TL / RBT * 4, for 64 more registers per
    // lane.  Without TL, acc is the span sum itself (RFLUSH_SL pairs).
    floatx16 acc[2][2], sacc[2][2];
    auto zero_acc = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                acc[r][c] = floatx16{};
                if constexpr (TL) sacc[r][c] = floatx16{};
            }
    };
    auto fold = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c)
#pragma unroll
                for (int e = 0; e < 16; ++e) sacc[r][c][e] = __fadd_rn(sacc[r][c][e], acc[r][c][e]);
    };
    bf16x8 ah[2], al[2], bh[2], bl[2];
    auto load_ops = [&](int buf) __attribute__((always_inline)) {
        // operand reads in the order the MFMAs consume them (lo parts first),
        // so that the first products wait for two reads, not eight
        const int offr0 = (wr * 64 + (lane & 31)) * RPITCH + (lane >> 5) * 8;
        const int offc0 = (wc * 64 + (lane & 31)) * RPITCH + (lane >> 5) * 8;
#pragma unroll
        for (int r = 0; r < 2; ++r) {
            al[r] = *reinterpret_cast<const bf16x8 *>(&lds.op[buf][sa][1][offr0 + r * 32 * RPITCH]);
            bl[r] = *reinterpret_cast<const bf16x8 *>(&lds.op[buf][sb][1][offc0 + r * 32 * RPITCH]);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r) ah[r] = *reinterpret_cast<const bf16x8 *>(&lds.op[buf][sa][0][offr0 + r * 32 * RPITCH]);
#pragma unroll
        for (int c = 0; c < 2; ++c) bh[c] = *reinterpret_cast<const bf16x8 *>(&lds.op[buf][sb][0][offc0 + c * 32 * RPITCH]);
    };
    // MFMA i of a stage (i = 0..15): product i >> 2 (lo lo, hi lo, lo hi, hi hi:
    // the small products first) on the 32 x 32 block ((i >> 1) & 1, i & 1); ZC:
    // the stage starts a sub-span, its first product starts from zero
    auto mfma_i = [&](auto zc_tag, int i) __attribute__((always_inline)) {
        constexpr bool ZC = decltype(zc_tag)::value;
        const int pr = i >> 2, r = (i >> 1) & 1, c = i & 1;
        const bf16x8 a = (pr & 1) ? ah[r] : al[r];
        const bf16x8 b = (pr & 2) ? bh[c] : bl[c];
        if (ZC && pr == 0)
            acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, floatx16{}, 0, 0, 0);
        else
            acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[r][c], 0, 0, 0);
    };
    auto flush_as = [&](auto add_tag) __attribute__((always_inline)) {
        constexpr bool ADD = decltype(add_tag)::value;
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                // element-major fragment order: one instruction covers whole cache
                // lines; wave-uniform base + 32-bit lane offset (no 64-bit vector
                // addresses to keep live across the loop)
                char *base = reinterpret_cast<char *>(mslot +
                                                      (int64_t)((wr * 2 + wc) * 4 + r * 2 + c) * 1024);
                uint32_t loff = (uint32_t)lane * 8u;
                asm volatile("" : "+v"(loff));  // recomputed per flush: keeps the addresses out of the loop
#pragma unroll
                for (int e = 0; e < 16; ++e) {
                    const double v = TL ? (double)sacc[r][c][e] : (double)acc[r][c][e];
                    double *dst = reinterpret_cast<double *>(base + e * 512 + loff);
                    // private slot, one writer lane per element: no atomicity needed;
                    // the L2-side add keeps the old value out of registers
                    if constexpr (ADD)
                        unsafeAtomicAdd(dst, v);
                    else
                        *dst = v;
                }
            }
        // drain the slot writes here: the loop's frame loads then are the only
        // vector memory operations in flight, and each stage waits only for
        // the loads of its own register set
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    };
    auto flush = [&](bool first_flush) __attribute__((always_inline)) {
        if (first_flush)
            flush_as(std::integral_constant<bool, false>());
        else
            flush_as(std::integral_constant<bool, true>());
    };

    // ---- pipeline: stage t runs its MFMAs from LDS buffer t & 1 while stage
    // t + 1 is split from register set (t + 1) % RSETS into the other buffer;
    // that set is then refilled with stage t + 1 + RSETS ----
    // The loop runs whole iterations of RU stages (a multiple of RSETS and of
    // RSUB, so that every set and buffer index is a constant) and at least one
    // (zero) span, so that every slot is written; stages past the slice load
    // the shift vector for x and y (U = V = 0).
    constexpr int RU = RSETS * RSUB / (RSETS % RSUB == 0 ? RSUB : (RSETS % 2 == 0 ? 2 : 1));
    static_assert(RU % RSETS == 0 && RU % RSUB == 0, "loop unroll");
    const int rspan = (TL ? RFLUSH_TL : RFLUSH_SL) / RBT;  // stages per span, a multiple of RSUB
    const int n_run = n_stages > 0 ? (n_stages + RU - 1) / RU * RU : RU;
    auto issue = [&](float *xr, float *yr, float *wr, int t) __attribute__((always_inline)) {
        stage_rows(t);
        issue_rows(xr, yr, wr);
    };
    // prologue (inside each of the two loop versions below, so that no load is
    // moved across the branch between them): the loads in stage order
    // (scheduling barriers), stage 0 split into buffer 0, then every load
    // drained once, so that the loop's first stage waits only for its own
    // register set (a register the prologue loads into may serve as an
    // accumulator inside the loop)
    auto prologue = [&]() __attribute__((always_inline)) {
#pragma unroll
        for (int k = 0; k < RSETS; ++k) {
            issue(xs[k], ys[k], ws[k], k);
            __builtin_amdgcn_sched_barrier(0);
        }
#pragma unroll
        for (int q = 0; q < PPT; q += 2) prep_pair(xs[0], ys[0], ws[0], q);
        prep_store(0);
        __builtin_amdgcn_sched_barrier(0);
        issue(xs[0], ys[0], ws[0], RSETS);
        __builtin_amdgcn_sched_barrier(0);
        zero_acc();
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
        __syncthreads();
    };
    bool first_flush = true;
    // one stage: the MFMAs of the stage in buffer mb (ACT), the split of the
    // register set (xr, yr, wr) into buffer pb in PPT / 2 pieces, each issued
    // after its share of the MFMAs, then the set's refill with stage t.  The
    // rows of stage t are found first (scalar code with branches), so that
    // the matrix block is one basic block, the same for every stage.
    auto stage = [&](auto act_tag, auto zc_tag, int mb, float *xr, float *yr, float *wr, int pb, int t)
                     __attribute__((always_inline)) {
        constexpr bool ACT = decltype(act_tag)::value;
        constexpr int NP = PPT / 2;        // split pieces
        constexpr int MPP = 16 / NP;       // MFMAs per piece
        constexpr int VPM = DIAG ? 3 : 5;  // vector instructions after each MFMA
        stage_rows(t);
        // raised wave priority while the stage's MFMAs are issued (the two
        // waves of a SIMD then tend to alternate: measured 1 % faster)
        if constexpr (ACT) __builtin_amdgcn_s_setprio(1);
        if constexpr (ACT) load_ops(mb);
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            if constexpr (ACT) {
#pragma unroll
                for (int i = j * MPP; i < (j + 1) * MPP; ++i) mfma_i(zc_tag, i);
            }
            prep_pair(xr, yr, wr, 2 * j);
            if (j == NP - 1) {
                prep_store(pb);
                issue_rows(xr, yr, wr);
            }
            if constexpr (ACT) {
#pragma unroll
                for (int i = 0; i < MPP; ++i) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);    // 1 MFMA
                    __builtin_amdgcn_sched_group_barrier(0x002, VPM, 0);  // vector ALU
                }
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if constexpr (ACT) __builtin_amdgcn_s_setprio(0);
        __syncthreads();
    };
    // waves without matrix work run the same loads, splits and barriers
    auto run = [&](auto act_tag) __attribute__((always_inline)) {
        constexpr bool ACT = decltype(act_tag)::value;
        prologue();
        for (int s = 0; s < n_run; s += RU) {
            static_for<0, RU>([&](auto kc) __attribute__((always_inline)) {
                constexpr int k = decltype(kc)::value;
                constexpr int ns = (k + 1) % RSETS;  // set of stage s + k + 1
                stage(act_tag, std::integral_constant<bool, TL && k % RSUB == 0>(), k & 1, xs[ns], ys[ns], ws[ns],
                      (k + 1) & 1, s + k + 1 + RSETS);
                if constexpr (k % RSUB == RSUB - 1) {
                    if constexpr (ACT && TL) fold();
                    const int done = s + k + 1;  // stages accumulated
                    if (done % rspan == 0 || done >= n_run) {
                        if constexpr (ACT) {
                            flush(first_flush);
                            zero_acc();
                        }
                        first_flush = false;
                    }
                }
            });
        }
    };
    if (wave_act)
        run(std::integral_constant<bool, true>());
    else
        run(std::integral_constant<bool, false>());

    // ---- column sums, diagonal tasks only: of U (= shifted x + shifted y) and
    // zero, or (NRS) of x' and y' ----
    if (DIAG && !XY) {
        s_sum[0][grp][lcol] = sum_u;
        s_sum[1][grp][lcol] = sum_v;
        __syncthreads();
        if (tid < TB) {
            const double a = s_sum[0][0][tid] + s_sum[0][1][tid] + s_sum[0][2][tid] + s_sum[0][3][tid];
            const double b = s_sum[1][0][tid] + s_sum[1][1][tid] + s_sum[1][2][tid] + s_sum[1][3][tid];
            double *sums = slot + 2LL * TB * TB;
            sums[tid] = a;
            sums[TB + tid] = NRS ? b : 0.0;
        }
    }
}

// All I <= J tile tasks of one kind of bf16 moments (MODE: reversible,
// non-reversible symmetric, lagged products) in one launch.  The workgroups of one pair slice are
// consecutive in one XCD's dispatch order (off-diagonal tasks first, then the
// diagonal ones), so the tasks that read the same frame rows run side by side
// and share the XCD's L2: each frame comes from HBM about once per call
// instead of once per kind of task.
template <bool TL, int MODE, bool WT>
__global__ __launch_bounds__(RNT) void k_rev_bf16(Params p) {
    __shared__ __align__(16) RevLds lds;
    __shared__ double s_sum[2][4][RTB];
    const int per_xcd = gridDim.x >> 3;
    const int wg = (blockIdx.x & 7) * per_xcd + (blockIdx.x >> 3);
    if (wg >= p.n_sym * p.n_slices) return;
    const int slice = wg / p.n_sym;
    const int lt = wg - slice * p.n_sym;
    const int n_off = p.n_sym - p.nb;
    int I, J;
    if (lt >= n_off) {
        I = J = lt - n_off;
    } else {
        int i = 0, rem = lt;  // lt-th pair I < J in row-major order
        while (rem >= p.nb - 1 - i) {
            rem -= p.nb - 1 - i;
            ++i;
        }
        I = i;
        J = i + 1 + rem;
    }
    int task = 0;
    for (int i = 0; i < I; ++i) task += p.nb - i;
    task += J - I;
    if (I == J)
        rev_body<true, TL, MODE, WT>(p, lds, s_sum, I, J, task, slice);
    else
        rev_body<false, TL, MODE, WT>(p, lds, s_sum, I, J, task, slice);
}

// ------------------------------------------------------------------
// frame weights: sum of the weights of every trajectory's pairs, then the
// total (both in a fixed order)
// ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_wsum(Params p) {
    __shared__ double s_part[256];
    const int t = blockIdx.x, tid = threadIdx.x;
    const int64_t np = p.pair_off[t + 1] - p.pair_off[t];
    const int64_t base = p.traj_off[t] + p.skip;
    double a = 0.0;
    const int64_t q0 = p.pair_off[t];
    for (int64_t i = tid; i < np; i += 256) {
        const double wv = p.wts[base + i * p.stride];
        a += wv;
        p.wpair[q0 + i] = (float)wv;
    }
    s_part[tid] = a;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (tid < off) s_part[tid] += s_part[tid + off];
        __syncthreads();
    }
    if (tid == 0) p.wpart[t] = s_part[0];
}

__global__ void k_wtotal(Params p) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double a = 0.0;
        for (int t = 0; t < p.ntraj; ++t) a += p.wpart[t];
        *p.W = a;
    }
}

// ------------------------------------------------------------------
// reduction + FP64 corrections
// ------------------------------------------------------------------
__device__ __forceinline__ int sym_task(const Params &p, int I, int J) {
    int t = 0;
    for (int i = 0; i < I; ++i) t += p.nb - i;
    return t + (J - I);
}

// Position of tile element (li, lj) inside a slot matrix.  Fragment layout:
// wave (li/64, lj/64) of the 2x2 wave grid, 32x32 subtile (r, c), and inside
// it the 32x32 MFMA output layout lane = jj + 32*((ii>>2)&1), e = 4*(ii>>3) + (ii&3),
// stored element-major (64 lanes of one element are consecutive).
__device__ __forceinline__ int64_t slot_index(const Params &p, int li, int lj) {
    if (!p.frag) return (int64_t)li * p.tb + lj;
    const int wave = (li >> 6) * 2 + (lj >> 6);
    const int r = (li >> 5) & 1, c = (lj >> 5) & 1;
    const int ii = li & 31, jj = lj & 31;
    const int lane = jj + 32 * ((ii >> 2) & 1);
    const int e = 4 * (ii >> 3) + (ii & 3);
    return (int64_t)(wave * 4 + r * 2 + c) * 1024 + e * 64 + lane;
}

__device__ __forceinline__ double slot_sum(const Params &p, int task, int m, int li, int lj) {
    const int64_t off = (int64_t)m * p.tb * p.tb + slot_index(p, li, lj);
    double a = 0.0;
    for (int s = 0; s < p.n_slices; ++s) a += p.slots[((int64_t)s * p.n_tasks + task) * p.slot_doubles + off];
    return a;
}

// sums: S_x[c], S_y[c] (shifted) -> out sx/sy and workspace copies
__global__ __launch_bounds__(256) void k_reduce_sums(Params p, double *Sx, double *Sy, double *w_out,
                                                     double *sx_out, double *sy_out) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t P = *p.P_total;
    const double Wp = p.wts ? *p.W : (double)P;  // total pair weight
    const double w = p.rev ? 2.0 * Wp : Wp;
    if (c == 0) *w_out = w;
    if (c >= p.d) return;
    const int I = c / p.tb, lc = c - I * p.tb;
    const int t = sym_task(p, I, I);
    double ax = 0.0, ay = 0.0;
    for (int s = 0; s < p.n_slices; ++s) {
        const double *sums = p.slots + ((int64_t)s * p.n_tasks + t) * p.slot_doubles + 2LL * p.tb * p.tb;
        ax += sums[lc];
        ay += sums[p.tb + lc];
    }
    const double K = (double)p.K[c];
    Sx[c] = ax;
    Sy[c] = ay;
    if (p.lag == 0) {
        sx_out[c] = ax + w * K;
        if (sy_out) sy_out[c] = ax + w * K;
    } else if (p.rev) {
        const double s = ax + ay + w * K;
        sx_out[c] = s;
        if (sy_out) sy_out[c] = s;
    } else {
        sx_out[c] = ax + Wp * K;
        if (sy_out) sy_out[c] = ay + Wp * K;
    }
}

// one thread per output element (i, j)
__global__ __launch_bounds__(256) void k_reduce_mats(Params p, const double *Sx, const double *Sy,
                                                     int remove_mean, double *Mxx, double *Mxy,
                                                     double *Myy) {
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int d = p.d, tb = p.tb;
    if (e >= (int64_t)d * d) return;
    const int i = (int)(e / d), j = (int)(e - (int64_t)i * d);
    const int64_t P = *p.P_total;
    const int I = i / tb, J = j / tb, li = i - I * tb, lj = j - J * tb;
    // symmetric matrices are taken from the upper tile triangle (mirrored)
    const int is = i <= j ? i : j, js = i <= j ? j : i;
    const int Is = is / tb, Js = js / tb;
    const int ts = sym_task(p, Is, Js);
    const int lis = is - Is * tb, ljs = js - Js * tb;
    const double K_i = (double)p.K[i], K_j = (double)p.K[j];
    const double Wp = p.wts ? *p.W : (double)P;  // total pair weight
    if (P == 0) {
        if (Mxx) Mxx[e] = 0.0;
        if (Mxy) Mxy[e] = 0.0;
        if (Myy) Myy[e] = 0.0;
        return;
    }
    // the correction terms are grouped so that (i, j) and (j, i) round alike:
    // the symmetric matrices come out exactly symmetric
    if (p.lag == 0) {
        const double raw = slot_sum(p, ts, 0, lis, ljs);
        const double w = Wp;
        Mxx[e] = remove_mean ? raw - Sx[i] * Sx[j] / w : raw + ((Sx[i] * K_j + K_i * Sx[j]) + w * (K_i * K_j));
        return;
    }
    if (p.rev) {
        const double uu = slot_sum(p, ts, 0, lis, ljs);
        const double vv = slot_sum(p, ts, 1, lis, ljs);
        const double w = 2.0 * Wp;
        const double Si = Sx[i] + Sy[i], Sj = Sx[j] + Sy[j];
        const double rxx = 0.5 * (uu + vv), rxy = 0.5 * (uu - vv);
        const double corr = remove_mean ? -(Si * Sj) / w : (Si * K_j + K_i * Sj) + w * (K_i * K_j);
        Mxx[e] = rxx + corr;
        if (Mxy) Mxy[e] = rxy + corr;
        return;
    }
    const double w = Wp;
    const double rxx = slot_sum(p, ts, 0, lis, ljs);
    Mxx[e] = remove_mean ? rxx - Sx[i] * Sx[j] / w : rxx + ((Sx[i] * K_j + K_i * Sx[j]) + w * (K_i * K_j));
    if (Myy) {
        const double ryy = slot_sum(p, ts, 1, lis, ljs);
        Myy[e] = remove_mean ? ryy - Sy[i] * Sy[j] / w : ryy + ((Sy[i] * K_j + K_i * Sy[j]) + w * (K_i * K_j));
    }
    if (Mxy) {
        const int tx = p.n_sym + I * p.nb + J;
        const double rxy = slot_sum(p, tx, 0, li, lj);
        Mxy[e] = remove_mean ? rxy - Sx[i] * Sy[j] / w : rxy + ((Sx[i] * K_j + K_i * Sy[j]) + w * (K_i * K_j));
    }
}

template <int PREC, bool WT>
static void launch_kind_w(const Params &p, hipStream_t s) {
    const int n_sym = p.n_sym;
    const unsigned g_sym = (unsigned)(cdiv((int64_t)n_sym * p.n_slices, 8) * 8);
    if constexpr (PREC == EMMA_PREC_BF16X4) {
        // the pipelined kernel: MODE 0 reversible, MODE 1 the symmetric
        // non-reversible (or lag-0) matrices, MODE 2 the lagged products
        if (p.rev) {
            if (p.acc2)
                k_rev_bf16<true, 0, WT><<<g_sym, RNT, 0, s>>>(p);
            else
                k_rev_bf16<false, 0, WT><<<g_sym, RNT, 0, s>>>(p);
            return;
        }
        if (p.acc2)
            k_rev_bf16<true, 1, WT><<<g_sym, RNT, 0, s>>>(p);
        else
            k_rev_bf16<false, 1, WT><<<g_sym, RNT, 0, s>>>(p);
        if (p.has_xy) {
            if (p.acc2)
                k_rev_bf16<true, 2, WT><<<g_sym, RNT, 0, s>>>(p);
            else
                k_rev_bf16<false, 2, WT><<<g_sym, RNT, 0, s>>>(p);
        }
    } else {
        // FP32 and FP64 tiles: the general kernel
        if (p.lag == 0)
            k_main<PREC, LAG0, 1, WT><<<g_sym, NTHR, 0, s>>>(p, 0, n_sym);
        else if (p.rev)
            k_main<PREC, REV, 2, WT><<<g_sym, NTHR, 0, s>>>(p, 0, n_sym);
        else if (p.ctt)
            k_main<PREC, NR_SYM, 2, WT><<<g_sym, NTHR, 0, s>>>(p, 0, n_sym);
        else
            k_main<PREC, NR_SYM, 1, WT><<<g_sym, NTHR, 0, s>>>(p, 0, n_sym);
        if (p.has_xy) {
            const int n_xy = p.nb * p.nb;
            const unsigned g_xy = (unsigned)(cdiv((int64_t)n_xy * p.n_slices, 8) * 8);
            k_main<PREC, NR_XY, 1, WT><<<g_xy, NTHR, 0, s>>>(p, n_sym, n_xy);
        }
    }
}

template <int PREC>
static void launch_kind(const Params &p, hipStream_t s) {
    if (p.wts)
        launch_kind_w<PREC, true>(p, s);
    else
        launch_kind_w<PREC, false>(p, s);
}

static int launch_main(const Params &p, int precision, hipStream_t s) {
    if (precision == EMMA_PREC_BF16X4)
        launch_kind<EMMA_PREC_BF16X4>(p, s);
    else if (precision == EMMA_PREC_FP32)
        launch_kind<EMMA_PREC_FP32>(p, s);
    else
        launch_kind<EMMA_PREC_FP64>(p, s);
    return check_launch("moments main");
}

struct Plan {
    Params p;
    size_t ws_bytes;
    double *Sx, *Sy;
};

static void make_plan(Plan &pl, const float *X, int64_t n_rows, int d, const int64_t *traj_off, int ntraj,
                      int lag, int stride, int skip, uint32_t flags, int precision, void *ws) {
    Params &p = pl.p;
    p.X = X;
    p.n_rows = n_rows;
    p.d = d;
    p.traj_off = traj_off;
    p.ntraj = ntraj;
    p.lag = lag;
    p.stride = stride;
    p.skip = skip;
    p.acc2 = 0;
    p.tb = precision == EMMA_PREC_FP64 ? 64 : 128;
    p.frag = precision == EMMA_PREC_FP64 ? 0 : 1;
    p.slot_doubles = 2LL * p.tb * p.tb + 2LL * p.tb;
    p.nb = (int)cdiv(d, p.tb);
    p.n_sym = p.nb * (p.nb + 1) / 2;
    p.rev = (flags & EMMA_MOM_SYMMETRIZE) && lag > 0 ? 1 : 0;
    p.ctt = (flags & EMMA_MOM_CTT) && lag > 0 && !p.rev ? 1 : 0;
    p.has_xy = lag > 0 && !p.rev && (flags & EMMA_MOM_C0T) ? 1 : 0;
    p.n_tasks = p.n_sym + (p.has_xy ? p.nb * p.nb : 0);
    const bool weighted = (flags & EMMA_MOM_WEIGHTED) != 0;
    if (precision == EMMA_PREC_BF16X4) {
        // one 8-wave workgroup per CU: 256 slices make the launch of
        // k_rev_bf16 a whole number of rounds (one per task) over the 256 CUs
        const int64_t cap = cdiv(n_rows > 0 ? n_rows : 1, 8192);
        p.n_slices = (int)(cap < 256 ? cap : 256);
    } else {
        // about 3 workgroups per CU on 256 CUs, at least ~4k pairs per slice
        const int64_t want = cdiv(768, p.n_tasks);
        const int64_t cap = cdiv(n_rows > 0 ? n_rows : 1, 4096);
        p.n_slices = (int)(want < cap ? want : cap);
    }
    if (p.n_slices < 1) p.n_slices = 1;
    Carver cv(ws);
    p.P_total = cv.take<int64_t>(1);
    p.pair_off = cv.take<int64_t>((size_t)ntraj + 1);
    p.K = cv.take<float>(d);
    pl.Sx = cv.take<double>(d);
    pl.Sy = cv.take<double>(d);
    p.slots = cv.take<double>((size_t)p.n_tasks * p.n_slices * p.slot_doubles);
    p.wts = nullptr;
    p.wpart = weighted ? cv.take<double>((size_t)(ntraj > 0 ? ntraj : 1)) : nullptr;
    p.W = weighted ? cv.take<double>(1) : nullptr;
    p.wpair = weighted ? cv.take<float>((size_t)(n_rows > 0 ? n_rows : 1)) : nullptr;  // P <= n_rows
    pl.ws_bytes = cv.bytes();
}

}  // namespace mom
}  // namespace emma

using namespace emma;

namespace emma {
namespace mom {
// EMMA_PREC_BF16X4_ACC: the bf16 kernel with two-level accumulation (every
// kind of moments runs the pipelined kernel in bf16).
static int effective_precision(int precision, int *acc2) {
    *acc2 = precision == EMMA_PREC_BF16X4_ACC ? 1 : 0;
    return *acc2 ? EMMA_PREC_BF16X4 : precision;
}
}  // namespace mom
}  // namespace emma

extern "C" size_t emma_moments_workspace_size(int64_t n_rows, int32_t d, int32_t ntraj, uint32_t flags,
                                              int32_t precision) {
    if (d <= 0 || ntraj < 0) return 0;
    if (precision == EMMA_PREC_BF16X4_ACC) precision = EMMA_PREC_BF16X4;  // the same plan
    mom::Plan pl, pl0;
    // lag only matters through lag == 0; size for the larger task set
    mom::make_plan(pl, nullptr, n_rows, d, nullptr, ntraj, 1, 1, 0, flags, precision, nullptr);
    mom::make_plan(pl0, nullptr, n_rows, d, nullptr, ntraj, 0, 1, 0, flags, precision, nullptr);
    return pl.ws_bytes > pl0.ws_bytes ? pl.ws_bytes : pl0.ws_bytes;
}

extern "C" int emma_moments_lagged_weighted(const float *X, int64_t n_rows, int32_t d, const int64_t *traj_offsets,
                                            int32_t ntraj, int32_t lag, int32_t stride, int32_t skip, uint32_t flags,
                                            int32_t precision, const double *row_weights, double *w, double *sx,
                                            double *sy, double *Mxx, double *Mxy, double *Myy, void *workspace,
                                            size_t workspace_bytes, void *stream) {
    EMMA_REQUIRE(d > 0 && ntraj > 0 && n_rows >= 0, "emma_moments_lagged: bad sizes d=%d ntraj=%d", d, ntraj);
    EMMA_REQUIRE(lag >= 0 && stride >= 1 && skip >= 0, "emma_moments_lagged: bad lag/stride/skip");
    EMMA_REQUIRE(precision == EMMA_PREC_FP32 || precision == EMMA_PREC_BF16X4 || precision == EMMA_PREC_FP64 ||
                     precision == EMMA_PREC_BF16X4_ACC,
                 "emma_moments_lagged: unknown precision %d", precision);
    EMMA_REQUIRE(X && traj_offsets && w && sx && Mxx, "emma_moments_lagged: NULL pointer");
    EMMA_REQUIRE(((flags & EMMA_MOM_WEIGHTED) != 0) == (row_weights != nullptr),
                 "emma_moments_lagged: EMMA_MOM_WEIGHTED and row_weights must come together");
    const bool rev = (flags & EMMA_MOM_SYMMETRIZE) && lag > 0;
    EMMA_REQUIRE(!((flags & EMMA_MOM_CTT) && rev), "emma_moments_lagged: CTT with SYMMETRIZE is meaningless");
    EMMA_REQUIRE(lag > 0 || !(flags & (EMMA_MOM_C0T | EMMA_MOM_CTT)),
                 "emma_moments_lagged: lag must be positive if c0t or ctt are requested");
    EMMA_REQUIRE(!(flags & EMMA_MOM_C0T) || Mxy, "emma_moments_lagged: Mxy is NULL");
    EMMA_REQUIRE(!(flags & EMMA_MOM_CTT) || Myy, "emma_moments_lagged: Myy is NULL");
    int acc2 = 0;
    precision = mom::effective_precision(precision, &acc2);
    mom::Plan pl;
    mom::make_plan(pl, X, n_rows, d, traj_offsets, ntraj, lag, stride, skip, flags, precision, nullptr);
    EMMA_REQUIRE(workspace && workspace_bytes >= pl.ws_bytes,
                 "emma_moments_lagged: workspace too small (%zu < %zu)", workspace_bytes, pl.ws_bytes);
    mom::make_plan(pl, X, n_rows, d, traj_offsets, ntraj, lag, stride, skip, flags, precision, workspace);
    pl.p.wts = row_weights;
    pl.p.acc2 = acc2;
    hipStream_t s = (hipStream_t)stream;
    const mom::Params &p = pl.p;
    mom::k_prep<<<1, 1024, 0, s>>>(p);
    EMMA_TRY(check_launch("moments prep"));
    if (p.wts) {
        mom::k_wsum<<<(unsigned)ntraj, 256, 0, s>>>(p);
        mom::k_wtotal<<<1, 64, 0, s>>>(p);
        EMMA_TRY(check_launch("moments weight sums"));
    }
    EMMA_TRY(mom::launch_main(p, precision, s));
    mom::k_reduce_sums<<<(unsigned)cdiv(d, 256), 256, 0, s>>>(p, pl.Sx, pl.Sy, w, sx, sy);
    EMMA_TRY(check_launch("moments reduce sums"));
    const int rm = (flags & EMMA_MOM_REMOVE_MEAN) ? 1 : 0;
    mom::k_reduce_mats<<<(unsigned)cdiv((int64_t)d * d, 256), 256, 0, s>>>(
        p, pl.Sx, pl.Sy, rm, Mxx, p.lag > 0 && (flags & EMMA_MOM_C0T) ? Mxy : nullptr, p.ctt ? Myy : nullptr);
    return check_launch("moments reduce");
}

extern "C" int emma_moments_lagged(const float *X, int64_t n_rows, int32_t d, const int64_t *traj_offsets,
                                   int32_t ntraj, int32_t lag, int32_t stride, int32_t skip, uint32_t flags,
                                   int32_t precision, double *w, double *sx, double *sy, double *Mxx,
                                   double *Mxy, double *Myy, void *workspace, size_t workspace_bytes,
                                   void *stream) {
    EMMA_REQUIRE(!(flags & EMMA_MOM_WEIGHTED), "emma_moments_lagged: use emma_moments_lagged_weighted for weights");
    return emma_moments_lagged_weighted(X, n_rows, d, traj_offsets, ntraj, lag, stride, skip, flags, precision,
                                        nullptr, w, sx, sy, Mxx, Mxy, Myy, workspace, workspace_bytes, stream);
}
