// Assignment of frames to their nearest cluster center.
//
// Reference semantics: label(x) = argmin_j sum_i (x_i - c_ji)^2 with the sum
// evaluated in FP32, in index order, without fused multiply-add, strict '<'
// (the first minimum wins) -- ClusterModel.transform as called from
// coordinates/clustering/interface.py:157-167.
//
// Two device paths produce exactly these labels:
//
//  * k_assign_plain: the sequential FP32 loop itself on the vector ALUs
//    (4 frames x 4 centers per thread, LDS-staged tiles).
//  * k_assign_mfma, then k_assign_plain on a list: distances from bf16 hi/lo MFMA products on
//    shifted data, d~_j = |a|^2 + |b_j|^2 - 2 a.b_j with a = fl(x - mu),
//    b_j = fl(c_j - mu).  With E_j = CE (|a| + |b_j|)^2 a rigorous bound on
//    |d~_j - d_j| (CE from u = 2^-24, the bf16 split residual 2^-16 and the
//    FP32 accumulation of 3 d products, see assign_error_coeff) and
//    g = gamma_{d+2} the relative rounding of the reference loop, a frame is
//    decided when
//        (1 - g) min_{j != best} (d~_j - E_j)  >  (1 + g)(d~_best + E_best);
//    then no other center can reach the reference loop's minimum.  All other
//    frames are appended to a list and re-evaluated by k_assign_plain with the
//    sequential FP32 loop over every center.
#include <math.h>

#include <mutex>

#include "emma_common.hpp"

namespace emma {
namespace asg {

// ------------------------------------------------------------------
// exact sequential FP32 distance pieces (library built with -ffp-contract=off)
// ------------------------------------------------------------------
__device__ __forceinline__ float sq_step(float acc, float x, float c) {
    const float t = __fsub_rn(x, c);
    return __fadd_rn(acc, __fmul_rn(t, t));
}

// ------------------------------------------------------------------
// plain kernel: 64 frames x all centers per workgroup
// ------------------------------------------------------------------
constexpr int PF = 64;      // frames per workgroup
constexpr int PC = 64;      // centers per tile
constexpr int PD = 32;      // dims per staged chunk
constexpr int PPITCH = 68;  // floats per staged row (16-byte aligned)

__global__ __launch_bounds__(256) void k_assign_plain(const float *__restrict__ X, int64_t n, int d,
                                                      const float *__restrict__ C, int k,
                                                      int32_t *__restrict__ labels,
                                                      const int64_t *__restrict__ list,
                                                      const int64_t *__restrict__ list_count,
                                                      const int32_t *__restrict__ done) {
    if (done && *done) return;
    __shared__ __align__(16) float xs[PD][PPITCH];
    __shared__ __align__(16) float cs[PD][PPITCH];
    __shared__ float s_best[16][PF];
    __shared__ int s_bidx[16][PF];
    __shared__ int64_t s_row[PF];
    __shared__ int s_nan0[PF];  // distance to center 0 is NaN

    const int tid = threadIdx.x;
    const int fg = tid >> 4;  // frames fg*4 .. +3
    const int cg = tid & 15;  // centers cg*4 .. +3 within the tile
    const int64_t nrows = list ? *list_count : n;
    // grid-stride over blocks of PF frames (a listed run sizes its grid for the
    // whole device and reads the list length on the device)
    for (int64_t f0 = (int64_t)blockIdx.x * PF; f0 < nrows; f0 += (int64_t)gridDim.x * PF) {
    __syncthreads();  // s_row / s_best of the previous block are consumed
    if (tid < PF) {
        const int64_t f = f0 + tid;
        s_row[tid] = f < nrows ? (list ? list[f] : f) : -1;
        s_nan0[tid] = 0;
    }
    __syncthreads();

    float best[4];
    int bidx[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        best[a] = INFINITY;
        bidx[a] = 0x7fffffff;
    }
    // loader mapping: row r = tid / 4 (0..63), dims (tid % 4) * 8 .. +7
    const int lr = tid >> 2, ld0 = (tid & 3) * 8;
    const int64_t xrow = s_row[lr];

    for (int c0 = 0; c0 < k; c0 += PC) {
        float acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = 0.0f;
        const int crow = c0 + lr;
        for (int i0 = 0; i0 < d; i0 += PD) {
            __syncthreads();
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = i0 + ld0 + q;
                xs[ld0 + q][lr] = (xrow >= 0 && i < d) ? X[xrow * d + i] : 0.0f;
                cs[ld0 + q][lr] = (crow < k && i < d) ? C[(int64_t)crow * d + i] : 0.0f;
            }
            __syncthreads();
            const int ni = d - i0 < PD ? d - i0 : PD;
            for (int ii = 0; ii < ni; ++ii) {
                const float4 xv = *reinterpret_cast<const float4 *>(&xs[ii][fg * 4]);
                const float4 cv = *reinterpret_cast<const float4 *>(&cs[ii][cg * 4]);
                const float xa[4] = {xv.x, xv.y, xv.z, xv.w};
                const float ca[4] = {cv.x, cv.y, cv.z, cv.w};
#pragma unroll
                for (int a = 0; a < 4; ++a)
#pragma unroll
                    for (int b = 0; b < 4; ++b) acc[a][b] = sq_step(acc[a][b], xa[a], ca[b]);
            }
        }
        // the reference loop starts from best = d(x, c_0): a NaN there is
        // never replaced, so the label stays 0
        if (c0 == 0 && cg == 0) {
#pragma unroll
            for (int a = 0; a < 4; ++a)
                if (acc[a][0] != acc[a][0]) s_nan0[fg * 4 + a] = 1;
        }
        // centers of this thread in increasing index order: strict '<'
#pragma unroll
        for (int b = 0; b < 4; ++b) {
            const int j = c0 + cg * 4 + b;
            if (j < k) {
#pragma unroll
                for (int a = 0; a < 4; ++a)
                    if (acc[a][b] < best[a]) {
                        best[a] = acc[a][b];
                        bidx[a] = j;
                    }
            }
        }
    }
    // reduce over the 16 center groups: smallest distance, then smallest index
#pragma unroll
    for (int a = 0; a < 4; ++a) {
        s_best[cg][fg * 4 + a] = best[a];
        s_bidx[cg][fg * 4 + a] = bidx[a];
    }
    __syncthreads();
    if (tid < PF) {
        float bv = s_best[0][tid];
        int bi = s_bidx[0][tid];
        for (int g = 1; g < 16; ++g) {
            const float v = s_best[g][tid];
            const int vi = s_bidx[g][tid];
            if (v < bv || (v == bv && vi < bi)) {
                bv = v;
                bi = vi;
            }
        }
        // every distance +inf, or the distance to center 0 NaN: the
        // reference loop keeps center 0
        if (s_row[tid] >= 0) labels[s_row[tid]] = (bi == 0x7fffffff || s_nan0[tid]) ? 0 : bi;
    }
    }
}

__global__ void k_copy_i64(int64_t *dst, const int64_t *src) { *dst = *src; }
__global__ void k_add_i64(int64_t *dst, const int64_t *src) { *dst += *src; }

// ------------------------------------------------------------------
// center preparation for the MFMA path
// ------------------------------------------------------------------
// Centers are packed once per call into "tile images": for every tile of 32
// centers one contiguous block that is byte-for-byte the LDS layout the MFMA
// kernel reads,
//   uint16 hi[32][pitch], lo[32][pitch]   bf16 bits of b = fl(c - mu), rows
//                                          padded from d_pad to pitch = d_pad + 8
//   float  nb[32], q[32], r[32]           |b|^2 (FP32 sequential), nb - CE bn^2, 2 CE bn
// so that a tile moves from L2 to LDS with asynchronous 16-byte copies.
struct CenterPack {
    uint8_t *img;    // k_pad / 32 tile images of img_bytes each
    uint8_t *img16;  // k_pad / 32 fp16 tile images of img16_bytes each (the coarse pass)
    float *bn;       // [k_pad] |b| rounded up
    float *mu;       // [d_pad] shift
    float *bmax;     // max_j |b_j| (zeroed before the packing, then an atomic max)
    float *sc16;     // [4] coarse pass: 2^t, -2^(1 - 2t), the coefficient c3 of its absolute error term
};

__host__ __device__ constexpr int tile_pitch(int d_pad) { return d_pad + 8; }
__host__ __device__ constexpr int tile_img_bytes(int d_pad) { return 128 * tile_pitch(d_pad) + 3 * 32 * 4; }
// fp16 image of the coarse pass: uint16 h[32][pitch] (fp16 bits of 2^t b),
// then the same constants nb, q, r
__host__ __device__ constexpr int tile_img16_bytes(int d_pad) { return 64 * tile_pitch(d_pad) + 3 * 32 * 4; }

// mu = mean of the centers (FP64 in a fixed order, rounded to FP32):
// 64 dims per workgroup, wave w sums centers w, w + 4, ...
__global__ __launch_bounds__(256) void k_center_mean(const float *__restrict__ C, int k, int d, int d_pad,
                                                     float *__restrict__ mu) {
    __shared__ double s_part[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int i = blockIdx.x * 64 + lane;
    double acc = 0.0;
    if (i < d)
        for (int j = wave; j < k; j += 4) acc += (double)C[(int64_t)j * d + i];
    s_part[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && i < d_pad)
        mu[i] = i < d ? (float)(((s_part[0][lane] + s_part[1][lane]) + (s_part[2][lane] + s_part[3][lane])) / k) : 0.0f;
}

// one wavefront per (padded) center
__global__ __launch_bounds__(256) void k_center_pack(const float *__restrict__ C, int k, int d, int d_pad,
                                                     int k_pad, float2 CE, CenterPack cp) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + wave;
    if (j >= k_pad) return;
    const int pitch = tile_pitch(d_pad);
    uint8_t *img = cp.img + (int64_t)(j >> 5) * tile_img_bytes(d_pad);
    uint16_t *hi = reinterpret_cast<uint16_t *>(img) + (j & 31) * pitch;
    uint16_t *lo = hi + 32 * pitch;
    float *cst = reinterpret_cast<float *>(img + 128 * pitch);
    double n2 = 0.0;
    for (int i = lane; i < pitch; i += 64) {
        float b = 0.0f;
        if (j < k && i < d) b = __fsub_rn(C[(int64_t)j * d + i], cp.mu[i]);
        uint32_t h, l;
        split_bf16(b, h, l);
        hi[i] = (uint16_t)h;  // the pad columns d_pad .. pitch - 1 hold zeros
        lo[i] = (uint16_t)l;
        n2 += (double)b * (double)b;
    }
    // FP32 |b|^2 in index order (one lane walks the row; the error bound only
    // needs gamma_d, any order would do, this keeps it deterministic)
    double n2s = wave_butterfly_sum(n2);
    if (lane == 0) {
        float nb, q, r, bn;
        if (j < k) {
            float acc = 0.0f;
            for (int i = 0; i < d; ++i) {
                const float b = __fsub_rn(C[(int64_t)j * d + i], cp.mu[i]);
                acc = __fadd_rn(acc, __fmul_rn(b, b));
            }
            bn = (float)(sqrt(n2s) * (1.0 + 1e-6));  // rounded up
            nb = acc;
            q = acc - CE.x * bn * bn;
            r = 2.0f * CE.y * bn;
        } else {
            // padding: a finite value beyond every real d~ (the epilogue's
            // index bits must not turn it into a NaN); never a candidate
            nb = 3.0e38f;
            bn = 0.0f;
            q = 3.0e38f;
            r = 0.0f;
        }
        cp.bn[j] = bn;
        if (j < k) atomicMax(reinterpret_cast<unsigned int *>(cp.bmax), __float_as_uint(bn));  // bn >= 0
        cst[j & 31] = nb;
        cst[32 + (j & 31)] = q;
        cst[64 + (j & 31)] = r;
    }
}

// Scale of the coarse pass: 2^t with t = 9 - e, bmax = m 2^e (m in [1/2, 1)),
// clamped to |t| <= 60, so that every |2^t b_ji| < 2^10 (fp16 holds up to
// 65504: frames with |a_i| < 64 max|b| stay finite; larger ones overflow to
// inf and go to the fine pass).  c3 bounds the effect of fp16 underflow
// (absolute error <= 2^-25 of a scaled component) on d~:
// 2 * 1.01 * 2^-25 2^-t sqrt(d) (|a| + |b_j|), with a 25 % margin.
__global__ void k_center_scale16(int d, CenterPack cp) {
    const float bm = *cp.bmax;
    int e = 0;
    frexpf(bm, &e);
    int t = bm > 0.0f && bm == bm && bm < INFINITY ? 9 - e : 0;
    t = t > 60 ? 60 : (t < -60 ? -60 : t);
    cp.sc16[0] = ldexpf(1.0f, t);
    cp.sc16[1] = -ldexpf(1.0f, 1 - 2 * t);
    cp.sc16[2] = (float)(1.25 * 2.0 * 1.01 * ldexp(1.0, -25 - t) * sqrt((double)d));
    cp.sc16[3] = 0.0f;
}

// fp16 tile images of 2^t b (the same b = fl(c - mu) as the bf16 images, and
// a copy of their constants); one wavefront per (padded) center
__global__ __launch_bounds__(256) void k_center_pack16(const float *__restrict__ C, int k, int d, int d_pad,
                                                       int k_pad, CenterPack cp) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int j = blockIdx.x * 4 + wave;
    if (j >= k_pad) return;
    const int pitch = tile_pitch(d_pad);
    const float sc = cp.sc16[0];
    uint8_t *img = cp.img16 + (int64_t)(j >> 5) * tile_img16_bytes(d_pad);
    _Float16 *h = reinterpret_cast<_Float16 *>(img) + (j & 31) * pitch;
    for (int i = lane; i < pitch; i += 64) {
        float b = 0.0f;
        if (j < k && i < d) b = __fsub_rn(C[(int64_t)j * d + i], cp.mu[i]);
        h[i] = (_Float16)__fmul_rn(b, sc);  // the pad columns hold zeros
    }
    if (lane < 3) {
        const float *src = reinterpret_cast<const float *>(cp.img + (int64_t)(j >> 5) * tile_img_bytes(d_pad) +
                                                           128 * pitch);
        float *dst = reinterpret_cast<float *>(img + 64 * pitch);
        dst[32 * lane + (j & 31)] = src[32 * lane + (j & 31)];
    }
}

__global__ void k_zero_i64n(int64_t *p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

__global__ void k_zero_i32n(int32_t *p, int64_t n) {
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// ------------------------------------------------------------------
// the recheck of undecided frames
// ------------------------------------------------------------------
// Pass 1 (the MFMA kernel) appends every frame it cannot decide to `list`
// together with `upper`, a rigorous upper bound (rounded up to FP32) of the
// reference loop's distance to the frame's best candidate; NaN marks a
// frame whose distances are not finite.  Pass 2 (the MFMA kernel in
// candidate mode) recomputes the listed frames' distance bounds and records,
// per list position p < slots, every center j whose rigorous lower bound
// does not exceed upper[p] (at most CAND_MAX of them in cand[p][.], their
// number in ccnt[p]); those centers contain the reference loop's choice.
// Pass 3 runs the reference loop over the recorded candidates only; frames
// beyond the slots, with more than CAND_MAX candidates or without a finite
// bound go to `list2` and the full loop over every center.
constexpr int CAND_MAX = 8;

struct RecheckLists {
    int64_t *listA;  // [n] frames the coarse pass leaves undecided
    int64_t *list;   // [n] frames the fine pass leaves undecided
    int64_t *count;  // [8]: list length, bits of max |b_j|, list2 length, listA length, unused
    float *upper;    // [n] upper bound per list position
    int32_t *ccnt;   // [slots] candidates found per list position
    int32_t *cand;   // [slots][CAND_MAX]
    int64_t *list2;  // [n] frames for the full loop
    int64_t slots;
};

// float >= v (v a finite double or NaN)
__device__ __forceinline__ float f32_round_up(double v) {
    float f = (float)v;
    if ((double)f < v) f = __uint_as_float(f >= 0.0f ? __float_as_uint(f) + 1u
                                                   : (__float_as_uint(f) == 0x80000000u ? 1u : __float_as_uint(f) - 1u));
    return f;
}

__device__ __forceinline__ void list_frame_coarse(const RecheckLists &rl, int64_t fr) {
    const unsigned long long pos = atomicAdd((unsigned long long *)(rl.count + 3), 1ull);
    rl.listA[pos] = fr;
}

__device__ __forceinline__ void list_frame(const RecheckLists &rl, int64_t fr, float upper) {
    const unsigned long long pos = atomicAdd((unsigned long long *)rl.count, 1ull);
    rl.list[pos] = fr;
    rl.upper[pos] = upper;
}

// ------------------------------------------------------------------
// MFMA kernel
// ------------------------------------------------------------------
// Wave w of a workgroup owns 32 frames; their bf16 hi/lo fragments stay in
// registers for all centers.  Center tile images stream through two LDS
// buffers: while the waves compute tile t from one buffer, the copy engine
// path (global_load_lds, 16 bytes per lane) fills the other with tile t + 1.
// v_mfma_f32_32x32x16_bf16: A = 32 centers x 16 dims, B = 16 dims x 32
// frames; lane l holds frame l % 32 and centers 8(e>>2) + 4(l>>5) + (e&3).
typedef __attribute__((address_space(3))) void *lds_ptr_t;

// vmcnt(n): wait until at most n vector-memory operations of this wave are
// outstanding (n a small run-time value; the instruction needs a constant)
__device__ __forceinline__ void wait_vmcnt(int n) {
    switch (n) {
        case 0: __builtin_amdgcn_s_waitcnt(0x0F70); break;
        case 1: __builtin_amdgcn_s_waitcnt(0x0F71); break;
        case 2: __builtin_amdgcn_s_waitcnt(0x0F72); break;
        case 3: __builtin_amdgcn_s_waitcnt(0x0F73); break;
        case 4: __builtin_amdgcn_s_waitcnt(0x0F74); break;
        case 5: __builtin_amdgcn_s_waitcnt(0x0F75); break;
        case 6: __builtin_amdgcn_s_waitcnt(0x0F76); break;
        case 7: __builtin_amdgcn_s_waitcnt(0x0F77); break;
        case 8: __builtin_amdgcn_s_waitcnt(0x0F78); break;
        case 9: __builtin_amdgcn_s_waitcnt(0x0F79); break;
        case 10: __builtin_amdgcn_s_waitcnt(0x0F7A); break;
        case 11: __builtin_amdgcn_s_waitcnt(0x0F7B); break;
        case 12: __builtin_amdgcn_s_waitcnt(0x0F7C); break;
        case 13: __builtin_amdgcn_s_waitcnt(0x0F7D); break;
        case 14: __builtin_amdgcn_s_waitcnt(0x0F7E); break;
        default: __builtin_amdgcn_s_waitcnt(0x0F70); break;  // larger counts: wait for everything
    }
}

// Running state of the epilogue per frame block and lane: the two smallest
// values v_j = d~_j - |a|^2 = |b_j|^2 - 2 a.b_j seen so far (m1 <= m2 as a
// multiset), each with the element e (0..15) of its 32 x 32 tile in the 4 low
// mantissa bits, and the tile of m1.  Per element: one fma for v, one
// and-or for the index, med3 and min for the pair (m2 = med3(m1, m2, v) is
// the new second smallest when m1 <= m2).
__device__ __forceinline__ float pack_elem(float v, int e) {
    return __uint_as_float((__float_as_uint(v) & 0xFFFFFFF0u) | (uint32_t)e);
}
// min(m1, vp) as the median of (m1, vp, -inf): v_min_f32 would first
// canonicalize the bit-built vp (one more instruction per element), and the
// compiler folds a median with a literal -inf back into that min, so the -inf
// comes from opaque_neg_inf() (an SGPR set once per kernel).
__device__ __forceinline__ float opaque_neg_inf() {
    float v = -INFINITY;
    asm volatile("" : "+s"(v));
    return v;
}
__device__ __forceinline__ void track2(float &m1, float &m2, float vp, float ninf) {
    m2 = __builtin_amdgcn_fmed3f(m1, m2, vp);
    m1 = __builtin_amdgcn_fmed3f(m1, vp, ninf);
}

// Label of one frame from the running state of its two half-waves (lane and
// lane ^ 32 hold the same frame, different centers), and whether it needs
// the recheck (lanes < 32 write).  With E_max = c1 (|a|^2 + max|b|^2) +
// 2 c2 |a| max|b| >= E_j for every center: for j != best,
// d_j >= d~_j - E_j >= |a|^2 + m2 - E_max, and d_best <= |a|^2 + m1 + E_best
// (both through the reference loop's relative rounding g).  A listed frame
// carries the upper bound of its best candidate to the candidate pass.
template <bool COARSE>
__device__ __forceinline__ void decide_frame(int lane, int64_t fr, bool fvalid, float m1, float m2, int mt, int k,
                                             float na, float sa, float2 CE, float c3, float gref, float bmax,
                                             const float *__restrict__ bn, int32_t *__restrict__ labels,
                                             const RecheckLists &rl) {
    int half = lane >> 5;
    {
        const float o1 = __shfl_xor(m1, 32, 64);
        const float o2 = __shfl_xor(m2, 32, 64);
        const int ot = __shfl_xor(mt, 32, 64);
        m2 = fminf(fminf(m2, o2), fmaxf(m1, o1));
        if (o1 < m1 || (o1 == m1 && (ot < mt || (ot == mt && half == 1)))) {
            m1 = o1;
            mt = ot;
            half ^= 1;
        }
    }
    const int e = (int)(__float_as_uint(m1) & 15u);
    const int imin = (mt >= 0 && m1 == m1 && m1 < 1e38f) ? mt * 32 + 8 * (e >> 2) + 4 * half + (e & 3) : 0x7fffffff;
    if (lane < 32 && fvalid && imin >= k) {
        labels[fr] = 0;  // non-finite distances (or, in the coarse pass, fp16 overflow): the next pass
        if (COARSE)
            list_frame_coarse(rl, fr);
        else
            list_frame(rl, fr, __uint_as_float(0x7fc00000u));
    } else if (lane < 32 && fvalid) {
        const double a2 = (double)sa * sa;
        const double bnb = (double)bn[imin];
        const double E1 = (double)CE.x * (a2 + bnb * bnb) + 2.0 * (double)CE.y * (double)sa * bnb +
                          (double)c3 * ((double)sa + bnb);
        const double Em = (double)CE.x * (a2 + (double)bmax * bmax) + 2.0 * (double)CE.y * (double)sa * bmax +
                          (double)c3 * ((double)sa + (double)bmax);
        const double upper = ((double)na + (double)m1 + E1) * (1.0 + (double)gref);
        const double lower = ((double)na + (double)m2 - Em) * (1.0 - (double)gref);
        labels[fr] = imin;
        if (!(lower > upper)) {
            if (COARSE)
                list_frame_coarse(rl, fr);
            else
                list_frame(rl, fr, f32_round_up(upper));
        }
    }
}

// The MFMA distance kernel in its three passes:
//   PASS_COARSE: every frame; one fp16 product per K-step of 2^t a and 2^t b
//     (fp16 images); decides with the coarse bound, lists the rest in listA.
//   PASS_FINE: the frames of listA; bf16 hi/lo of a and b, three products
//     per K-step; decides with the fine bound, lists the rest with their
//     upper bound.
//   PASS_CAND: the frames listed by PASS_FINE (positions below rl.slots);
//     the epilogue records the centers whose lower bound reaches the frame's
//     upper bound instead of tracking the minimum.
//   PASS_FULL: PASS_FINE over every frame (pass 1 without a coarse pass).
constexpr int PASS_COARSE = 0, PASS_FINE = 1, PASS_CAND = 2, PASS_FULL = 3;
template <int KSTEPS, int NW, int MINB, int FB, int PASS>
__global__ __launch_bounds__(NW * 64, MINB) void k_assign_mfma(const float *__restrict__ X, int64_t n, int d,
                                                            CenterPack cp, int k, int k_pad, float2 CE,
                                                            float gref, int32_t *__restrict__ labels,
                                                            RecheckLists rl, const int32_t *__restrict__ done) {
    constexpr bool COARSE = PASS == PASS_COARSE, CAND = PASS == PASS_CAND;
    if (done && *done) return;
    int64_t nf = n;  // frames: all (coarse), listA (fine), the candidate slots (cand)
    if constexpr (PASS == PASS_FINE) {
        nf = rl.count[3];
        if ((int64_t)blockIdx.x * NW * FB * 32 >= nf) return;
    }
    if constexpr (CAND) {
        const int64_t c = *rl.count;
        nf = c < rl.slots ? c : rl.slots;
        if ((int64_t)blockIdx.x * NW * FB * 32 >= nf) return;
    }
    constexpr int DP = KSTEPS * 16;
    constexpr int PITCH = tile_pitch(DP);
    constexpr int IMG = COARSE ? tile_img16_bytes(DP) : tile_img_bytes(DP);
    constexpr int MAT = (COARSE ? 64 : 128) * PITCH;  // bytes of the center fragments in an image
    constexpr int MCH = MAT / 16;           // 16-byte chunks of hi + lo
    constexpr int CCH = 3 * 32 * 4 / 16;    // 16-byte chunks of nb, q, r
    constexpr int NT = NW * 64;
    constexpr int ROUNDS = (MCH + NT - 1) / NT;
    // LDS buffer b holds hi + lo of a tile t with t & 1 == b and the constants
    // (nb, q, r) of tile t - 1: the MFMAs of tile t and the epilogue of tile
    // t - 1 run together and read the same buffer, while the copies fill the
    // other one.  Two separate objects, so that the compiler can tell the
    // buffer being read from the one being filled.
    // three buffers (tiles prefetched two ahead) where they fit next to the
    // occupancy of the variant, otherwise two
    constexpr int NB = (3 * IMG * MINB <= 160 * 1024) ? 3 : 2;
    __shared__ __align__(16) uint8_t s_img0[IMG];
    __shared__ __align__(16) uint8_t s_img1[IMG];
    __shared__ __align__(16) uint8_t s_img2[NB == 3 ? IMG : 16];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // FB blocks of 32 frames per wave: every center fragment read from LDS
    // feeds FB MFMAs, and a tile's copies are hidden behind FB times the work
    int64_t frame[FB], frow[FB], fpos[FB];
    bool fvalid[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b) {
        fpos[b] = (((int64_t)blockIdx.x * NW + wave) * FB + b) * 32 + (lane & 31);
        fvalid[b] = fpos[b] < nf;
        if constexpr (CAND)
            frame[b] = fvalid[b] ? rl.list[fpos[b]] : 0;
        else if constexpr (PASS == PASS_FINE)
            frame[b] = fvalid[b] ? rl.listA[fpos[b]] : 0;
        else
            frame[b] = fpos[b];
        frow[b] = (fvalid[b] ? frame[b] : 0) * (int64_t)d;  // always a valid row
    }
    const int kh = (lane >> 5) * 8;  // dim offset of this half-wave inside a K-step
    const int ntiles = k_pad / 32;

    // asynchronous copies into an LDS buffer: hi + lo of tile tm (tm < 0: none)
    // and the constants of tile tc (tc < 0: none)
    auto fetch = [&](int tm, int tc, uint8_t *dst) __attribute__((always_inline)) {
        const uint8_t *base = COARSE ? cp.img16 : cp.img;
        if (tm >= 0) {
            const uint8_t *src = base + (int64_t)tm * IMG;
#pragma unroll
            for (int i = 0; i < ROUNDS; ++i) {
                const int c0 = i * NT + wave * 64;  // first chunk of this wave's 1 KB block
                if (c0 + lane < MCH)
                    __builtin_amdgcn_global_load_lds(src + (int64_t)(c0 + lane) * 16, (lds_ptr_t)(dst + c0 * 16), 16, 0, 0);
            }
        }
        if (tc >= 0 && wave == 0 && lane < CCH)
            __builtin_amdgcn_global_load_lds(base + (int64_t)tc * IMG + MAT + lane * 16, (lds_ptr_t)(dst + MAT), 16, 0, 0);
    };
    fetch(0, -1, s_img0);

    // ---- frame fragments: a = fl(x - mu) split into bf16 hi/lo, or (coarse)
    // 2^t a rounded to fp16 (branch-free) ----
    bf16x8 bhi[FB][KSTEPS], blo[FB][KSTEPS];
    half8 fh[FB][KSTEPS];
    float na[FB], sa[FB];
    const float sc16 = COARSE ? cp.sc16[0] : 1.0f;
    // a lane's 8 dims of a K-step are consecutive in the frame row: with d a
    // multiple of 4 they are two 16-byte loads, each wholly inside the row or
    // wholly in the padding (loaded from the row's last 4 dims and masked).
    // Only for small K-step counts: with many steps the compiler keeps every
    // loaded vector live at once and runs out of registers.
    const bool vec = KSTEPS <= 8 && (d & 3) == 0 && ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
    // More K-steps (16-byte aligned rows; the bf16 passes only in tile
    // shapes of one wave per SIMD, whose registers hold the load groups next
    // to the hi / lo fragments): 16-byte loads of the frames in groups of GK K-steps, the next group's loads issued
    // before the current group is converted, so that the frame loads of a
    // wave are not one dword at a time (about 10 loads in flight otherwise:
    // the frame phase of a workgroup is then latency-bound, and up to
    // k = 512 it takes longer than the workgroup's tiles).
    // mu of a group is loaded before the next group's frames (loads complete
    // in order), and the values and the order of the |a|^2 terms are those of
    // the general loop below.
    constexpr int GK = 4;
    constexpr int NG = (KSTEPS + GK - 1) / GK;
    const bool gvec = (COARSE || NW * MINB <= 4) && KSTEPS > 8 && (d & 3) == 0 &&
                      ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
    if (gvec) {
#pragma unroll
        for (int b = 0; b < FB; ++b) {
            float na_part = 0.0f;
            float4 xg[2][GK][2];
            auto load_group = [&](int g, int slot) __attribute__((always_inline)) {
#pragma unroll
                for (int j = 0; j < GK; ++j) {
                    const int s = g * GK + j;
                    if (s < KSTEPS) {
#pragma unroll
                        for (int hv = 0; hv < 2; ++hv) {
                            const int i0 = s * 16 + kh + 4 * hv;
                            xg[slot][j][hv] = *reinterpret_cast<const float4 *>(X + frow[b] + (i0 < d ? i0 : d - 4));
                        }
                    }
                }
            };
            load_group(0, 0);
            float4 mg[GK][2];
            static_for<0, NG>([&](auto gc) __attribute__((always_inline)) {
                constexpr int g = decltype(gc)::value;
#pragma unroll
                for (int j = 0; j < GK; ++j) {
                    const int s = g * GK + j;
                    if (s < KSTEPS) {
                        mg[j][0] = *reinterpret_cast<const float4 *>(cp.mu + s * 16 + kh);
                        mg[j][1] = *reinterpret_cast<const float4 *>(cp.mu + s * 16 + kh + 4);
                    }
                }
                if constexpr (g + 1 < NG) load_group(g + 1, (g + 1) & 1);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int j = 0; j < GK; ++j) {
                    const int s = g * GK + j;
                    if (s < KSTEPS) {
                        const float4 *xv = xg[g & 1][j];
                        const float xs[8] = {xv[0].x, xv[0].y, xv[0].z, xv[0].w, xv[1].x, xv[1].y, xv[1].z, xv[1].w};
                        const float ms[8] = {mg[j][0].x, mg[j][0].y, mg[j][0].z, mg[j][0].w,
                                             mg[j][1].x, mg[j][1].y, mg[j][1].z, mg[j][1].w};
                        _Float16 f[8];
                        uint32_t h[8], l[8];
#pragma unroll
                        for (int q = 0; q < 8; ++q) {
                            const int i = s * 16 + kh + q;
                            const bool ok = fvalid[b] && i < d;
                            const uint32_t keep = ok ? 0xffffffffu : 0u;
                            const float a = __uint_as_float(__float_as_uint(__fsub_rn(xs[q], ms[q])) & keep);
                            na_part = __fadd_rn(na_part, __fmul_rn(a, a));
                            if constexpr (COARSE)
                                f[q] = (_Float16)__fmul_rn(a, sc16);
                            else
                                split_bf16(a, h[q], l[q]);
                        }
                        if constexpr (COARSE) {
                            fh[b][s] = half8{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]};
                        } else {
                            uint4 hv = make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7]));
                            uint4 lv = make_uint4(pack2(l[0], l[1]), pack2(l[2], l[3]), pack2(l[4], l[5]), pack2(l[6], l[7]));
                            bhi[b][s] = __builtin_bit_cast(bf16x8, hv);
                            blo[b][s] = __builtin_bit_cast(bf16x8, lv);
                        }
                    }
                }
                __builtin_amdgcn_sched_barrier(0);
            });
            na[b] = __fadd_rn(na_part, __shfl_xor(na_part, 32, 64));
            sa[b] = sqrtf(na[b]) * (1.0f + 1e-6f);  // |a| rounded up
        }
    }
#pragma unroll
    for (int b = 0; b < FB; ++b) {
        if (gvec) break;
        float na_part = 0.0f;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            uint32_t h[8], l[8];
            _Float16 f[8];
            float xs[8], ms[8];
            if (vec) {
#pragma unroll
                for (int hv = 0; hv < 2; ++hv) {
                    const int i0 = s * 16 + kh + 4 * hv;
                    const float4 x4 = *reinterpret_cast<const float4 *>(X + frow[b] + (i0 < d ? i0 : d - 4));
                    const float4 m4 = *reinterpret_cast<const float4 *>(cp.mu + i0);
                    xs[4 * hv] = x4.x;
                    xs[4 * hv + 1] = x4.y;
                    xs[4 * hv + 2] = x4.z;
                    xs[4 * hv + 3] = x4.w;
                    ms[4 * hv] = m4.x;
                    ms[4 * hv + 1] = m4.y;
                    ms[4 * hv + 2] = m4.z;
                    ms[4 * hv + 3] = m4.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int i = s * 16 + kh + q;
                    xs[q] = X[frow[b] + (i < d ? i : d - 1)];
                    ms[q] = cp.mu[i];
                }
            }
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = s * 16 + kh + q;
                const bool ok = fvalid[b] && i < d;
                const uint32_t keep = ok ? 0xffffffffu : 0u;
                const float a = __uint_as_float(__float_as_uint(__fsub_rn(xs[q], ms[q])) & keep);
                na_part = __fadd_rn(na_part, __fmul_rn(a, a));
                if constexpr (COARSE)
                    f[q] = (_Float16)__fmul_rn(a, sc16);
                else
                    split_bf16(a, h[q], l[q]);
            }
            if constexpr (COARSE) {
                fh[b][s] = half8{f[0], f[1], f[2], f[3], f[4], f[5], f[6], f[7]};
            } else {
                uint4 hv = make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7]));
                uint4 lv = make_uint4(pack2(l[0], l[1]), pack2(l[2], l[3]), pack2(l[4], l[5]), pack2(l[6], l[7]));
                bhi[b][s] = __builtin_bit_cast(bf16x8, hv);
                blo[b][s] = __builtin_bit_cast(bf16x8, lv);
            }
        }
        na[b] = __fadd_rn(na_part, __shfl_xor(na_part, 32, 64));
        sa[b] = sqrtf(na[b]) * (1.0f + 1e-6f);  // |a| rounded up
    }
    // candidate mode: center j is recorded when its lower bound reaches the
    // frame's upper bound, (na - CE sa^2 + lb_j)(1 - g) <= upper, as the
    // threshold lb_j <= T on the epilogue's FP32 lb_j = q_j - 2 p - sa r_j,
    // raised by 8 u (sa + max|b|)^2 so that rounding only ever adds
    // candidates; no threshold (NaN upper): no candidate, the frame goes to
    // the full loop
    float thr[FB];
#pragma unroll
    for (int b = 0; b < FB; ++b) {
        thr[b] = -INFINITY;
        if constexpr (CAND) {
            const float up = fvalid[b] ? rl.upper[fpos[b]] : __uint_as_float(0x7fc00000u);
            if (up == up) {
                const double sb = (double)sa[b] + (double)*cp.bmax;
                const double T = (double)up / (1.0 - (double)gref) - ((double)na[b] - (double)CE.x * (double)sa[b] * (double)sa[b]) +
                                 8.0 * 5.9604644775390625e-8 * sb * sb;
                thr[b] = f32_round_up(T);
            }
        }
    }

    // per frame block: the two smallest d~ - |a|^2 with the element index in
    // the low bits, and the tile of the smallest (see track2)
    float m1[FB], m2[FB];
    int mt[FB];
    const float ninf = opaque_neg_inf();
    // d~ - |a|^2 = |b|^2 + pscale * (MFMA sum): -2, or -2^(1 - 2t) for the
    // coarse pass's scaled operands
    const float pscale = COARSE ? cp.sc16[1] : -2.0f;
#pragma unroll
    for (int b = 0; b < FB; ++b) {
        m1[b] = m2[b] = INFINITY;
        mt[b] = -1;
    }

    struct AccFB {
        floatx16 v[FB];
    };
    auto chain = [&](const uint8_t *img) __attribute__((always_inline)) -> AccFB {
        const uint16_t *ah_base = reinterpret_cast<const uint16_t *>(img) + (lane & 31) * PITCH + kh;
        const uint16_t *al_base = ah_base + 32 * PITCH;
        AccFB acc;
#pragma unroll
        for (int b = 0; b < FB; ++b) acc.v[b] = floatx16{};
        if constexpr (COARSE) {
            half8 ch = *reinterpret_cast<const half8 *>(ah_base);
#pragma unroll
            for (int s = 0; s < KSTEPS; ++s) {
                half8 chn = ch;
                if (s + 1 < KSTEPS) chn = *reinterpret_cast<const half8 *>(ah_base + (s + 1) * 16);
#pragma unroll
                for (int b = 0; b < FB; ++b) acc.v[b] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ch, fh[b][s], acc.v[b], 0, 0, 0);
                ch = chn;
            }
            return acc;
        }
        // center operands read one K-step ahead of their MFMAs
        bf16x8 ah = *reinterpret_cast<const bf16x8 *>(ah_base);
        bf16x8 al = *reinterpret_cast<const bf16x8 *>(al_base);
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            bf16x8 ahn = ah, aln = al;
            if (s + 1 < KSTEPS) {
                ahn = *reinterpret_cast<const bf16x8 *>(ah_base + (s + 1) * 16);
                aln = *reinterpret_cast<const bf16x8 *>(al_base + (s + 1) * 16);
            }
#pragma unroll
            for (int b = 0; b < FB; ++b) acc.v[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, blo[b][s], acc.v[b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < FB; ++b) acc.v[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bhi[b][s], acc.v[b], 0, 0, 0);
#pragma unroll
            for (int b = 0; b < FB; ++b) acc.v[b] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bhi[b][s], acc.v[b], 0, 0, 0);
            ah = ahn;
            al = aln;
        }
        return acc;
    };
    auto epilogue = [&](const AccFB &acc, const uint8_t *img, int t) __attribute__((always_inline)) {
        const float *cst = reinterpret_cast<const float *>(img + MAT);
        if constexpr (CAND) {
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int c0 = 8 * g + 4 * (lane >> 5);
                const float4 qv = *reinterpret_cast<const float4 *>(&cst[32 + c0]);
                const float4 rv = *reinterpret_cast<const float4 *>(&cst[64 + c0]);
                const float qa[4] = {qv.x, qv.y, qv.z, qv.w};
                const float ra[4] = {rv.x, rv.y, rv.z, rv.w};
#pragma unroll
                for (int b = 0; b < FB; ++b)
#pragma unroll
                    for (int h = 0; h < 4; ++h) {
                        const float lb = fmaf(-sa[b], ra[h], fmaf(-2.0f, acc.v[b][4 * g + h], qa[h]));
                        if (lb <= thr[b]) {
                            const int j = t * 32 + c0 + h;
                            if (j < k) {
                                const int c = atomicAdd(&rl.ccnt[fpos[b]], 1);
                                if (c < CAND_MAX) rl.cand[fpos[b] * CAND_MAX + c] = j;
                            }
                        }
                    }
            }
            return;
        }
        float prev[FB];
#pragma unroll
        for (int b = 0; b < FB; ++b) prev[b] = m1[b];
#pragma unroll
        for (int g = 0; g < 4; ++g) {
            const int c0 = 8 * g + 4 * (lane >> 5);
            const float4 nbv = *reinterpret_cast<const float4 *>(&cst[c0]);
            const float nba[4] = {nbv.x, nbv.y, nbv.z, nbv.w};
#pragma unroll
            for (int b = 0; b < FB; ++b)
#pragma unroll
                for (int h = 0; h < 4; ++h)
                    track2(m1[b], m2[b], pack_elem(fmaf(pscale, acc.v[b][4 * g + h], nba[h]), 4 * g + h), ninf);
        }
#pragma unroll
        for (int b = 0; b < FB; ++b) mt[b] = m1[b] != prev[b] ? t : mt[b];
    };
    // MFMAs of one tile interleaved with the epilogue of the previous one
    // (about 4 VALU per element, 16 elements per frame block, spread over the
    // K-steps)
    constexpr int EV = (68 * FB + KSTEPS - 1) / KSTEPS;
    auto interleave = [&]() __attribute__((always_inline)) {
        __builtin_amdgcn_sched_group_barrier(0x100, COARSE ? 1 : 2, 0);  // operand reads of K-step 0
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            if (s + 1 < KSTEPS) __builtin_amdgcn_sched_group_barrier(0x100, COARSE ? 1 : 2, 0);  // operands of K-step s + 1
            __builtin_amdgcn_sched_group_barrier(0x008, (COARSE ? 1 : 3) * FB, 0);   // the MFMAs of K-step s
            __builtin_amdgcn_sched_group_barrier(0x002, EV, 0);                     // epilogue VALU
        }
    };

    if constexpr (NB == 3) {
        // Tiles prefetched two ahead: buffer u % 3 holds hi + lo of tile u and
        // the constants of tile u - 1.  Iteration t waits until the copies of
        // tile t are complete (those of tile t + 1 may still be in flight: the
        // count of this wave's copies per fetch is known), synchronizes, starts
        // the copies of tile t + 2 into the buffer read in iteration t - 1, and
        // runs the MFMAs of tile t with the epilogue of tile t - 1.
        int rw = 0;  // hi + lo copy instructions this wave issues per tile
#pragma unroll
        for (int i = 0; i < ROUNDS; ++i) rw += (i * NT + wave * 64 < MCH) ? 1 : 0;
        auto copies = [&](int u) -> int {  // copy instructions of this wave in fetch(u)
            return (u < ntiles ? rw : 0) + ((wave == 0 && u >= 1 && u - 1 < ntiles) ? 1 : 0);
        };
        fetch(ntiles > 1 ? 1 : -1, 0, s_img1);
        // tile 0: no epilogue yet
        wait_vmcnt(copies(1));
        __syncthreads();
        fetch(2 < ntiles ? 2 : -1, 1 < ntiles ? 1 : -1, s_img2);
        AccFB acc = chain(s_img0);
        auto iter = [&](int t, uint8_t *cur, uint8_t *tgt) __attribute__((always_inline)) {
            wait_vmcnt(copies(t + 1));
            __syncthreads();
            fetch(t + 2 < ntiles ? t + 2 : -1, t + 1 < ntiles ? t + 1 : -1, tgt);
            const AccFB nacc = chain(cur);
            epilogue(acc, cur, t - 1);
            interleave();
            acc = nacc;
        };
        for (int t = 1; t < ntiles; t += 3) {
            iter(t, s_img1, s_img0);
            if (t + 1 >= ntiles) break;
            iter(t + 1, s_img2, s_img1);
            if (t + 2 >= ntiles) break;
            iter(t + 2, s_img0, s_img2);
        }
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        const int rb = ntiles % 3;
        epilogue(acc, rb == 0 ? s_img0 : (rb == 1 ? s_img1 : s_img2), ntiles - 1);
    } else {
    // tile 0, then tiles 1 .. ntiles - 1 with the epilogue one tile behind.
    // Each barrier follows a wait for this wave's copies, so the buffer about
    // to be read is complete for every wave and every wave has finished
    // reading the buffer that the next copies overwrite.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    fetch(ntiles > 1 ? 1 : -1, 0, s_img1);
    AccFB acc = chain(s_img0);
    int t = 1;
    while (t < ntiles) {
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        fetch(t + 1 < ntiles ? t + 1 : -1, t, s_img0);
        {
            const AccFB nacc = chain(s_img1);
            epilogue(acc, s_img1, t - 1);
            interleave();
            acc = nacc;
        }
        if (++t >= ntiles) break;
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        fetch(t + 1 < ntiles ? t + 1 : -1, t, s_img1);
        {
            const AccFB nacc = chain(s_img0);
            epilogue(acc, s_img0, t - 1);
            interleave();
            acc = nacc;
        }
        ++t;
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);
    __syncthreads();
    epilogue(acc, (ntiles & 1) ? s_img1 : s_img0, ntiles - 1);
    }

    // per frame block: merge the two half-waves that hold the same frame and decide
    if constexpr (CAND) return;
    const float bmax = *cp.bmax;
    const float c3 = COARSE ? cp.sc16[2] : 0.0f;
#pragma unroll
    for (int b = 0; b < FB; ++b)
        decide_frame<COARSE>(lane, frame[b], fvalid[b], m1[b], m2[b], mt[b], k, na[b], sa[b], CE, c3, gref, bmax,
                             cp.bn, labels, rl);
}

// Resident-center form for small k d (every center tile image fits in LDS):
// one workgroup per CU copies all tile images into LDS once, then each wave
// walks 32-frame blocks with a grid stride, running every tile from LDS with
// no further copies or barriers.  Where the streaming kernel waits for a
// tile's copies every few MFMAs (few K-steps per tile), this one only streams
// the frames (the next block's raw values are loaded while the current block
// runs its tiles), so its time is set by the frames' HBM traffic, the MFMAs
// and the epilogue.
constexpr int RES_LDS = 144 * 1024;  // bytes of resident tile images

__host__ __device__ constexpr int res_max_tiles(int d_pad) { return RES_LDS / tile_img_bytes(d_pad); }

template <int KSTEPS, int NW>
__global__ __launch_bounds__(NW * 64, 1) void k_assign_res(const float *__restrict__ X, int64_t n, int d,
                                                           CenterPack cp, int k, int k_pad, float2 CE, float gref,
                                                           int32_t *__restrict__ labels, RecheckLists rl,
                                                           const int32_t *__restrict__ done) {
    if (done && *done) return;
    constexpr int DP = KSTEPS * 16;
    constexpr int PITCH = tile_pitch(DP);
    constexpr int IMG = tile_img_bytes(DP);
    constexpr int MAT = 128 * PITCH;
    constexpr int NT = NW * 64;
    __shared__ __align__(16) uint8_t s_res[RES_LDS];

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int kh = (lane >> 5) * 8;
    const int ntiles = k_pad / 32;  // ntiles * IMG <= RES_LDS (checked by the host)

    // every tile image (hi, lo and the constants nb, q, r) in one asynchronous copy
    const int nch = ntiles * (IMG / 16);
    for (int c0 = wave * 64; c0 < nch; c0 += NT)
        if (c0 + lane < nch)
            __builtin_amdgcn_global_load_lds(cp.img + (int64_t)(c0 + lane) * 16, (lds_ptr_t)(s_res + c0 * 16), 16, 0, 0);
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    __syncthreads();
    const float bmax = *cp.bmax;

    auto chain = [&](const uint8_t *img, const bf16x8 *bhi, const bf16x8 *blo) __attribute__((always_inline)) -> floatx16 {
        const uint16_t *ah_base = reinterpret_cast<const uint16_t *>(img) + (lane & 31) * PITCH + kh;
        const uint16_t *al_base = ah_base + 32 * PITCH;
        floatx16 acc = {};
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            const bf16x8 ah = *reinterpret_cast<const bf16x8 *>(ah_base + s * 16);
            const bf16x8 al = *reinterpret_cast<const bf16x8 *>(al_base + s * 16);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, blo[s], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, bhi[s], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah, bhi[s], acc, 0, 0, 0);
        }
        return acc;
    };

    const int64_t nblk = (n + 31) / 32;
    const int64_t stride = (int64_t)gridDim.x * NW;
    // raw frame values of the wave's next block are loaded while the current
    // block runs its tiles (lanes: frame lane & 31, dims s * 16 + kh .. + 7)
    const bool vec = KSTEPS <= 8 && (d & 3) == 0 && ((reinterpret_cast<uintptr_t>(X) & 15) == 0);
    float xr[KSTEPS][8];
    auto load_raw = [&](int64_t b) __attribute__((always_inline)) {
        const int64_t fr = b * 32 + (lane & 31);
        const int64_t row = (fr < n ? fr : 0) * (int64_t)d;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            if (vec) {
#pragma unroll
                for (int hv = 0; hv < 2; ++hv) {
                    const int i0 = s * 16 + kh + 4 * hv;
                    const float4 x4 = *reinterpret_cast<const float4 *>(X + row + (i0 < d ? i0 : d - 4));
                    xr[s][4 * hv] = x4.x;
                    xr[s][4 * hv + 1] = x4.y;
                    xr[s][4 * hv + 2] = x4.z;
                    xr[s][4 * hv + 3] = x4.w;
                }
            } else {
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    const int i = s * 16 + kh + q;
                    xr[s][q] = X[row + (i < d ? i : d - 1)];
                }
            }
        }
    };
    int64_t blk = (int64_t)blockIdx.x * NW + wave;
    if (blk < nblk) load_raw(blk);
    for (; blk < nblk; blk += stride) {
        const int64_t frame = blk * 32 + (lane & 31);
        const bool fvalid = frame < n;
        bf16x8 bhi[KSTEPS], blo[KSTEPS];
        float na_part = 0.0f;
#pragma unroll
        for (int s = 0; s < KSTEPS; ++s) {
            uint32_t h[8], l[8];
#pragma unroll
            for (int q = 0; q < 8; ++q) {
                const int i = s * 16 + kh + q;
                const uint32_t keep = (fvalid && i < d) ? 0xffffffffu : 0u;
                const float a = __uint_as_float(__float_as_uint(__fsub_rn(xr[s][q], cp.mu[i])) & keep);
                na_part = __fadd_rn(na_part, __fmul_rn(a, a));
                split_bf16(a, h[q], l[q]);
            }
            bhi[s] = __builtin_bit_cast(bf16x8, make_uint4(pack2(h[0], h[1]), pack2(h[2], h[3]), pack2(h[4], h[5]), pack2(h[6], h[7])));
            blo[s] = __builtin_bit_cast(bf16x8, make_uint4(pack2(l[0], l[1]), pack2(l[2], l[3]), pack2(l[4], l[5]), pack2(l[6], l[7])));
        }
        const float na = __fadd_rn(na_part, __shfl_xor(na_part, 32, 64));
        const float sa = sqrtf(na) * (1.0f + 1e-6f);  // |a| rounded up
        if (blk + stride < nblk) load_raw(blk + stride);

        // the two smallest d~ - |a|^2 and the tile of the smallest (track2)
        const float ninf = opaque_neg_inf();
        float m1 = INFINITY, m2 = INFINITY;
        int mt = -1;
        for (int t = 0; t < ntiles; ++t) {
            const uint8_t *img = s_res + t * IMG;
            const floatx16 acc = chain(img, bhi, blo);
            const float *cst = reinterpret_cast<const float *>(img + MAT);
            const float prev = m1;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                const int c0 = 8 * g + 4 * (lane >> 5);
                const float4 nbv = *reinterpret_cast<const float4 *>(&cst[c0]);
                const float nba[4] = {nbv.x, nbv.y, nbv.z, nbv.w};
#pragma unroll
                for (int h = 0; h < 4; ++h)
                    track2(m1, m2, pack_elem(fmaf(-2.0f, acc[4 * g + h], nba[h]), 4 * g + h), ninf);
            }
            mt = m1 != prev ? t : mt;
        }
        decide_frame<false>(lane, frame, fvalid, m1, m2, mt, k, na, sa, CE, 0.0f, gref, bmax, cp.bn, labels, rl);
    }
}

// Coefficients (c1, c2) = CE of the per-candidate bound
//   E_j = c1 (|a|^2 + |b_j|^2) + 2 c2 |a| |b_j|   >=   |d~_j - d_j|,
// with u = 2^-24, gamma_n = n u / (1 - n u):
//   shift rounding of a and b:            2.1 u (|a|+|b|)^2
//   final combination and fmaf rounding:  2.1 u (|a|+|b|)^2
//   element index in the low mantissa bits: 32 u (|a|+|b|)^2
//   FP32 norms |a|^2, |b|^2:              gamma_d (|a|^2 + |b|^2)
//   omitted bf16 split terms:             3.1 * 2^-16 * 2 sum|a_i b_i|
//   FP32 MFMA accumulation of 3 d_pad terms (each addition rounded to
//   nearest, as the device probe of the matrix core shows):
//                                         1.02 gamma_{3 d_pad} * 2 sum|a_i b_i|
// using (|a|+|b|)^2 = |a|^2 + |b|^2 + 2|a||b| and sum|a_i b_i| <= |a||b|, so
// the product terms are charged on 2|a||b| only.  A 25 % margin covers the
// rounding of the FP32 bound arithmetic itself.
static float2 assign_error_coeffs(int d, int d_pad) {
    const double u = ldexp(1.0, -24);
    auto gam = [&](double m) { return m * u / (1.0 - m * u); };
    // + 32 u: the epilogue keeps an element index in the 4 low mantissa bits
    // of d~ - |a|^2 (|d~ - |a|^2| <= (|a|+|b|)^2), a relative change < 2^-19
    const double c0 = 4.2 * u + 32.0 * u;
    float2 ce;
    ce.x = (float)(1.25 * (c0 + gam(d)));
    ce.y = (float)(1.25 * (c0 + 3.1 * ldexp(1.0, -16) + 1.02 * gam(3.0 * d_pad)));
    return ce;
}

// The coarse pass: one fp16 product of ah = fp16(2^t a) and bh = fp16(2^t b)
// per dimension.  Normal fp16 values carry a relative rounding <= 2^-11
// (|ah bh - 2^2t a_i b_i| <= (2^-10 + 2^-22) 2^2t |a_i b_i|), the FP32 MFMA
// accumulation of d_pad exact products adds 1.01 gamma_{d_pad}; the norms and
// the other terms are those of the fine bound (c1 unchanged).  Underflow to
// fp16 subnormals is the absolute term c3 (|a| + |b_j|) of k_center_scale16.
static float2 assign_error_coeffs_coarse(int d, int d_pad) {
    const double u = ldexp(1.0, -24);
    auto gam = [&](double m) { return m * u / (1.0 - m * u); };
    const double c0 = 4.2 * u + 32.0 * u;
    float2 ce;
    ce.x = (float)(1.25 * (c0 + gam(d)));
    ce.y = (float)(1.25 * (c0 + ldexp(1.0, -10) + ldexp(1.0, -22) + 1.01 * gam(d_pad)));
    return ce;
}

static double gamma_ref(int d) {
    const double u = ldexp(1.0, -24);
    const double m = d + 2;
    return 1.25 * m * u / (1.0 - m * u);
}

static int pick_ksteps(int d) {
    const int ks = (int)cdiv(d, 16);
    static const int sizes[] = {2, 4, 8, 12, 16, 19, 24, 32};
    for (int s : sizes)
        if (ks <= s) return s;
    return -1;
}

struct AssignWs {
    CenterPack cp;
    RecheckLists rl;
    size_t bytes;
};

// candidate slots: one per frame that can be listed (near-tie-heavy center
// sets, such as the cluster means after a Lloyd update with many centers per
// mode of the data, list about half of the frames)
static int64_t cand_slots(int64_t n) { return n > 0 ? n : 1; }

static AssignWs carve_assign(void *ws, int64_t n, int d, int k, int mode) {
    AssignWs w{};
    Carver cv(ws);
    const size_t nl = n > 0 ? (size_t)n : 1;
    // list length, bits of max |b_j|, list2 length, listA length (zeroed per call)
    w.rl.count = cv.take<int64_t>(8);
    w.rl.list = cv.take<int64_t>(nl);
    w.cp.bmax = w.rl.count ? reinterpret_cast<float *>(w.rl.count + 1) : nullptr;
    if (mode == EMMA_ASSIGN_BF16X3_RECHECK) {
        w.rl.slots = cand_slots(n);
        w.rl.upper = cv.take<float>(nl);
        w.rl.ccnt = cv.take<int32_t>((size_t)w.rl.slots);
        w.rl.cand = cv.take<int32_t>((size_t)w.rl.slots * CAND_MAX);
        // listA is consumed by the fine pass before the candidate recheck
        // writes list2: one buffer serves both
        w.rl.listA = cv.take<int64_t>(nl);
        w.rl.list2 = w.rl.listA;
        const int ks = pick_ksteps(d);
        const int d_pad = (ks > 0 ? ks : 1) * 16;
        const int k_pad = (int)cdiv(k, 32) * 32;
        w.cp.img = cv.take<uint8_t>((size_t)(k_pad / 32) * tile_img_bytes(d_pad));
        w.cp.img16 = cv.take<uint8_t>((size_t)(k_pad / 32) * tile_img16_bytes(d_pad));
        w.cp.bn = cv.take<float>(k_pad);
        w.cp.mu = cv.take<float>(d_pad);
        w.cp.sc16 = cv.take<float>(4);
    }
    w.bytes = cv.bytes();
    return w;
}

// Pass 3 of the recheck: the reference loop over the recorded candidates.
// CAND_MAX consecutive threads serve one list position (thread c: candidate
// c); the smallest distance wins, the smallest center index among equal
// ones (the reference's first minimum: the candidates contain every center
// the loop could choose, see RecheckLists).  Positions without usable
// candidates go to list2.  VEC: rows read as 16-byte vectors (d % 4 == 0,
// aligned X and C).
template <bool VEC>
__global__ __launch_bounds__(256) void k_assign_cand_exact(const float *__restrict__ X, int d,
                                                           const float *__restrict__ C, int32_t *__restrict__ labels,
                                                           RecheckLists rl, const int32_t *__restrict__ done) {
    if (done && *done) return;
    const int64_t total = *rl.count * CAND_MAX;  // whole groups enter each iteration (the stride is a multiple)
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const int64_t pos = idx / CAND_MAX;
        const int c = (int)(idx % CAND_MAX);
        const int64_t fr = rl.list[pos];
        int nc = -1;  // -1: the full loop
        if (pos < rl.slots) {
            const int cc = rl.ccnt[pos];
            if (cc >= 1 && cc <= CAND_MAX) nc = cc;
        }
        float dist = INFINITY;
        int j = 0x7fffffff;
        if (c < nc) {
            j = rl.cand[pos * CAND_MAX + c];
            const float *x = X + fr * (int64_t)d;
            const float *cr = C + (int64_t)j * d;
            float acc = 0.0f;
            if (VEC) {
                for (int i = 0; i < d; i += 4) {
                    const float4 xv = *reinterpret_cast<const float4 *>(x + i);
                    const float4 cv = *reinterpret_cast<const float4 *>(cr + i);
                    acc = sq_step(acc, xv.x, cv.x);
                    acc = sq_step(acc, xv.y, cv.y);
                    acc = sq_step(acc, xv.z, cv.z);
                    acc = sq_step(acc, xv.w, cv.w);
                }
            } else {
                for (int i = 0; i < d; ++i) acc = sq_step(acc, x[i], cr[i]);
            }
            dist = acc;
        }
#pragma unroll
        for (int off = 1; off < CAND_MAX; off <<= 1) {
            const float od = __shfl_xor(dist, off, 64);
            const int oj = __shfl_xor(j, off, 64);
            if (od < dist || (od == dist && oj < j)) {
                dist = od;
                j = oj;
            }
        }
        if (c == 0) {
            if (nc < 0) {
                const unsigned long long p2 = atomicAdd((unsigned long long *)(rl.count + 2), 1ull);
                rl.list2[p2] = fr;
            } else {
                labels[fr] = j;
            }
        }
    }
}

// Tile variants of the MFMA kernel for each K-step count: every variant fits
// the registers and the LDS tile buffers of its occupancy.  Several 32-frame
// blocks per wave (small d, where the fragments of a block take few
// registers) reuse each center fragment for more MFMAs and hide a tile's
// copies behind more work.  Variant 0 is the default; for K steps <= 8 and
// k small enough that every tile image fits in LDS it runs the
// resident-center kernel instead.  The autotuner (emma_assign_select) may
// pick another per (K steps, k) bucket.
// (waves per workgroup, workgroups per CU, 32-frame blocks per wave)
struct Variant {
    int nw, minb, fb;
};
static const Variant kVarSmall[] = {{4, 2, 1}, {8, 1, 1}, {16, 1, 1}, {4, 2, 2}, {8, 1, 2}, {4, 1, 4}};  // KSTEPS <= 4
static const Variant kVarK8[] = {{4, 2, 1}, {8, 1, 1}, {4, 1, 2}};                                      // KSTEPS = 8
static const Variant kVarMid[] = {{4, 2, 1}, {8, 1, 1}};                                              // 12 .. 19
static const Variant kVarLarge[] = {{4, 1, 1}};                                                       // 24, 32

static int n_shapes(int ks) {
    if (ks <= 4) return (int)(sizeof(kVarSmall) / sizeof(Variant));
    if (ks <= 8) return (int)(sizeof(kVarK8) / sizeof(Variant));
    if (ks <= 19) return (int)(sizeof(kVarMid) / sizeof(Variant));
    return (int)(sizeof(kVarLarge) / sizeof(Variant));
}
// variants 0 .. n_shapes - 1: the coarse pass in tile shape v (variant 0 of a
// small K-step count: the resident fine kernel where its images fit);
// n_shapes .. 2 n_shapes - 1: the fine pass over every frame in tile shape
// v - n_shapes.  Which pipeline is faster depends on d, k and the data (how
// many frames the coarse bound leaves to the fine pass), so the autotuner
// times both.
static int n_variants(int ks) { return 2 * n_shapes(ks); }

// process-wide selection table: [K-step class][ceil(log2(k_pad / 32))]
static std::mutex g_sel_mu;
static int g_sel[8][16];
static bool g_sel_init = false;

static int ks_index(int ks) {
    static const int sizes[] = {2, 4, 8, 12, 16, 19, 24, 32};
    for (int i = 0; i < 8; ++i)
        if (sizes[i] == ks) return i;
    return -1;
}
static int k_bucket(int k_pad) {
    int b = 0;
    while ((32 << b) < k_pad && b < 15) ++b;
    return b;
}
static int selected_variant(int ks, int k_pad) {
    std::lock_guard<std::mutex> lk(g_sel_mu);
    if (!g_sel_init) return 0;
    const int v = g_sel[ks_index(ks)][k_bucket(k_pad)];
    return v < n_variants(ks) ? v : 0;
}

// pass 1 over every frame (PASS_COARSE or PASS_FULL), tile shape (NW, MINB, FB)
template <int KS, int PASS, int NW, int MINB, int FB = 1>
static void launch_v(const float *X, int64_t n, int d, const AssignWs &w, int k, int k_pad, float2 CE, float gref,
                     int32_t *labels, const int32_t *done, hipStream_t s) {
    const int64_t blocks = cdiv(n, 32 * NW * FB);
    k_assign_mfma<KS, NW, MINB, FB, PASS><<<(unsigned)blocks, NW * 64, 0, s>>>(X, n, d, w.cp, k, k_pad, CE, gref,
                                                                             labels, w.rl, done);
}

template <int KS, int PASS>
static void launch_shape(int v, const float *X, int64_t n, int d, const AssignWs &w, int k, int k_pad, float2 CE,
                         float gref, int32_t *labels, const int32_t *done, hipStream_t s) {
    if constexpr (KS <= 4) {
        switch (v) {
            case 1: launch_v<KS, PASS, 8, 1>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            case 2: launch_v<KS, PASS, 16, 1>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            case 3: launch_v<KS, PASS, 4, 2, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            case 4: launch_v<KS, PASS, 8, 1, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            case 5: launch_v<KS, PASS, 4, 1, 4>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            default: launch_v<KS, PASS, 4, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
        }
    } else if constexpr (KS <= 8) {
        switch (v) {
            case 1: launch_v<KS, PASS, 8, 1>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            case 2: launch_v<KS, PASS, 4, 1, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
            default: launch_v<KS, PASS, 4, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s); return;
        }
    } else if constexpr (KS <= 19) {
        if (v == 1)
            launch_v<KS, PASS, 8, 1>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
        else
            launch_v<KS, PASS, 4, 2>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
    } else {
        launch_v<KS, PASS, 4, 1>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
    }
}

static int device_cus() {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        cus = 256;
    return cus;
}

// pass 1: variant 0 of a small K-step count with every bf16 tile image
// resident in LDS is the resident kernel with the fine bound (it lists to the
// fine pass's output directly, listA stays empty); every other variant is the
// coarse streaming pass
template <int KS>
static void launch_mfma_ks(const float *X, int64_t n, int d, const AssignWs &w, int k, int k_pad, float2 CE,
                           float2 CEc, float gref, int32_t *labels, const int32_t *done, hipStream_t s) {
    const int v = selected_variant(KS, k_pad);
    if constexpr (KS <= 8) {
        if (v == 0 && k_pad / 32 <= res_max_tiles(KS * 16)) {
            constexpr int NW = KS <= 4 ? 8 : 4;  // 8 K-steps: 4 waves, the registers of one wave per SIMD
            int64_t blocks = cdiv(n, 32 * NW);
            const int64_t cap = device_cus();
            if (blocks > cap) blocks = cap;
            k_assign_res<KS, NW><<<(unsigned)blocks, NW * 64, 0, s>>>(X, n, d, w.cp, k, k_pad, CE, gref, labels,
                                                                    w.rl, done);
            return;
        }
    }
    const int ns = n_shapes(KS);
    if (v < ns)
        launch_shape<KS, PASS_COARSE>(v, X, n, d, w, k, k_pad, CEc, gref, labels, done, s);
    else
        launch_shape<KS, PASS_FULL>(v - ns, X, n, d, w, k, k_pad, CE, gref, labels, done, s);
}

// the fine pass over listA and the candidate pass over the fine pass's list:
// grids for the most frames the lists can hold, blocks past a list's end exit
template <int KS, int PASS>
static void launch_list_ks(const float *X, int64_t n, int d, const AssignWs &w, int k, int k_pad, float2 CE,
                           float gref, int32_t *labels, const int32_t *done, hipStream_t s) {
    constexpr int NW = 4, MINB = KS <= 19 ? 2 : 1;
    const int64_t blocks = cdiv(PASS == PASS_CAND ? w.rl.slots : n, 32 * NW);
    k_assign_mfma<KS, NW, MINB, 1, PASS><<<(unsigned)blocks, NW * 64, 0, s>>>(X, n, d, w.cp, k, k_pad, CE, gref,
                                                                            labels, w.rl, done);
}

}  // namespace asg

// shared with kmeans.hip (Lloyd)
size_t assign_workspace_bytes(int64_t n, int d, int k, int mode) { return asg::carve_assign(nullptr, n, d, k, mode).bytes; }

int assign_run(const float *X, int64_t n, int d, const float *C, int k, int32_t *labels, int mode,
               int64_t *n_rechecked, void *workspace, const int32_t *done, hipStream_t s) {
    using namespace asg;
    AssignWs w = carve_assign(workspace, n, d, k, mode);
    if (mode == EMMA_ASSIGN_FP32_PLAIN) {
        k_assign_plain<<<(unsigned)cdiv(n, PF), 256, 0, s>>>(X, n, d, C, k, labels, nullptr, nullptr, done);
        EMMA_TRY(check_launch("assign plain"));
        if (n_rechecked) k_zero_i64n<<<1, 64, 0, s>>>(n_rechecked, 1);
        return check_launch("assign plain");
    }
    const int ks = pick_ksteps(d);
    EMMA_REQUIRE(ks > 0, "emma_assign: d=%d exceeds the MFMA path limit (512)", d);
    const int d_pad = ks * 16;
    const int k_pad = (int)cdiv(k, 32) * 32;
    const float2 CE = assign_error_coeffs(d, d_pad);
    const float2 CEc = assign_error_coeffs_coarse(d, d_pad);
    const float gref = (float)gamma_ref(d);
    k_zero_i64n<<<1, 64, 0, s>>>(w.rl.count, 4);
    k_zero_i32n<<<(unsigned)(cdiv(w.rl.slots, 256) < 1024 ? cdiv(w.rl.slots, 256) : 1024), 256, 0, s>>>(w.rl.ccnt,
                                                                                                   w.rl.slots);
    k_center_mean<<<(unsigned)cdiv(d_pad, 64), 256, 0, s>>>(C, k, d, d_pad, w.cp.mu);
    k_center_pack<<<(unsigned)cdiv(k_pad, 4), 256, 0, s>>>(C, k, d, d_pad, k_pad, CE, w.cp);
    k_center_scale16<<<1, 1, 0, s>>>(d, w.cp);
    k_center_pack16<<<(unsigned)cdiv(k_pad, 4), 256, 0, s>>>(C, k, d, d_pad, k_pad, w.cp);
    EMMA_TRY(check_launch("assign center pack"));
    switch (ks) {
        case 2: launch_mfma_ks<2>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 4: launch_mfma_ks<4>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 8: launch_mfma_ks<8>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 12: launch_mfma_ks<12>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 16: launch_mfma_ks<16>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 19: launch_mfma_ks<19>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 24: launch_mfma_ks<24>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
        case 32: launch_mfma_ks<32>(X, n, d, w, k, k_pad, CE, CEc, gref, labels, done, s); break;
    }
    EMMA_TRY(check_launch("assign mfma"));
    // the frames the coarse pass left: the fine pass (bf16 x 3), then the
    // recheck of those it leaves: candidates by the MFMA kernel, the
    // reference loop over them, the full loop over every center for the rest;
    // the grids cover the device, the kernels read the list lengths themselves
    switch (ks) {
        case 2:
            launch_list_ks<2, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<2, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 4:
            launch_list_ks<4, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<4, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 8:
            launch_list_ks<8, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<8, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 12:
            launch_list_ks<12, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<12, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 16:
            launch_list_ks<16, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<16, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 19:
            launch_list_ks<19, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<19, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 24:
            launch_list_ks<24, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<24, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
        case 32:
            launch_list_ks<32, PASS_FINE>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            launch_list_ks<32, PASS_CAND>(X, n, d, w, k, k_pad, CE, gref, labels, done, s);
            break;
    }
    EMMA_TRY(check_launch("assign candidates"));
    const int64_t eb = cdiv(n * CAND_MAX, 256) < 4096 ? cdiv(n * CAND_MAX, 256) : 4096;
    const bool vec = d % 4 == 0 && (((uintptr_t)X | (uintptr_t)C) & 15) == 0;
    if (vec)
        k_assign_cand_exact<true><<<(unsigned)eb, 256, 0, s>>>(X, d, C, labels, w.rl, done);
    else
        k_assign_cand_exact<false><<<(unsigned)eb, 256, 0, s>>>(X, d, C, labels, w.rl, done);
    EMMA_TRY(check_launch("assign candidate recheck"));
    const int64_t rb = cdiv(n, PF) < 2048 ? cdiv(n, PF) : 2048;
    k_assign_plain<<<(unsigned)rb, 256, 0, s>>>(X, n, d, C, k, labels, w.rl.list2, w.rl.count + 2, done);
    EMMA_TRY(check_launch("assign recheck"));
    if (n_rechecked) k_copy_i64<<<1, 1, 0, s>>>(n_rechecked, w.rl.count);
    return check_launch("assign copy count");
}

}  // namespace emma

using namespace emma;

extern "C" size_t emma_assign_workspace_size(int64_t n, int32_t d, int32_t k, int32_t mode) {
    if (n < 0 || d <= 0 || k <= 0) return 0;
    return assign_workspace_bytes(n, d, k, mode);
}

extern "C" int emma_assign(const float *X, int64_t n, int32_t d, const float *C, int32_t k, int32_t *labels,
                           int32_t mode, int64_t *n_rechecked, void *workspace, size_t workspace_bytes,
                           void *stream) {
    EMMA_REQUIRE(n >= 0 && d > 0 && k > 0, "emma_assign: bad sizes n=%lld d=%d k=%d", (long long)n, d, k);
    EMMA_REQUIRE(mode == EMMA_ASSIGN_FP32_PLAIN || mode == EMMA_ASSIGN_BF16X3_RECHECK,
                 "emma_assign: unknown mode %d", mode);
    if (n == 0) return EMMA_OK;
    EMMA_REQUIRE(X && C && labels, "emma_assign: NULL pointer");
    const size_t need = assign_workspace_bytes(n, d, k, mode);
    EMMA_REQUIRE(workspace && workspace_bytes >= need, "emma_assign: workspace too small (%zu < %zu)",
                 workspace_bytes, need);
    return assign_run(X, n, d, C, k, labels, mode, n_rechecked, workspace, nullptr, (hipStream_t)stream);
}

extern "C" int32_t emma_assign_variants(int32_t d) {
    const int ks = asg::pick_ksteps(d);
    return ks > 0 ? asg::n_variants(ks) : 0;
}

extern "C" int emma_assign_select(int32_t d, int32_t k, int32_t variant) {
    using namespace asg;
    EMMA_REQUIRE(d > 0 && k > 0, "emma_assign_select: bad sizes d=%d k=%d", d, k);
    const int ks = pick_ksteps(d);
    EMMA_REQUIRE(ks > 0, "emma_assign_select: d=%d exceeds the MFMA path limit (512)", d);
    EMMA_REQUIRE(variant < n_variants(ks), "emma_assign_select: variant %d of %d", variant, n_variants(ks));
    std::lock_guard<std::mutex> lk(g_sel_mu);
    if (!g_sel_init) {
        for (auto &row : g_sel)
            for (int &v : row) v = 0;
        g_sel_init = true;
    }
    g_sel[ks_index(ks)][k_bucket((int)cdiv(k, 32) * 32)] = variant < 0 ? 0 : variant;
    return EMMA_OK;
}

// ------------------------------------------------------------------
// allocation-free assignment: the workspace of emma_assign carved from a
// file-scope device array instead of a caller buffer
// ------------------------------------------------------------------
namespace emma {
namespace asg {

// Every GPU the library runs on holds its own copy of this array until the
// process ends; host code reaches it only through the address that
// hipGetSymbolAddress returns on that device (the array's name on the host
// is not a device address).
static __device__ __align__(256) uint8_t scratch[EMMA_ASSIGN_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices];

// Device address of `scratch` on the current device, resolved once per
// device (with that device current) and kept.
static int scratch_on_current_device(void **out) {
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        set_error("emma_assign_noalloc: hipGetDevice: %s", hipGetErrorString(e));
        return EMMA_ERR_HIP;
    }
    EMMA_REQUIRE(dev >= 0 && dev < kMaxDevices, "emma_assign_noalloc: device %d beyond %d", dev, kMaxDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("emma_assign_noalloc: hipGetSymbolAddress(scratch) on device %d: %s", dev,
                      hipGetErrorString(e));
            return EMMA_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return EMMA_OK;
}

}  // namespace asg
}  // namespace emma

extern "C" int emma_assign_noalloc(const float *X, int64_t n, int32_t d, const float *C, int32_t k,
                                   int32_t *labels, int32_t mode, int64_t *n_rechecked, void *stream) {
    using namespace emma::asg;
    EMMA_REQUIRE(n >= 0 && d > 0 && k > 0, "emma_assign_noalloc: bad sizes n=%lld d=%d k=%d", (long long)n, d, k);
    EMMA_REQUIRE(mode == EMMA_ASSIGN_FP32_PLAIN || mode == EMMA_ASSIGN_BF16X3_RECHECK,
                 "emma_assign_noalloc: unknown mode %d", mode);
    if (n == 0) return EMMA_OK;
    EMMA_REQUIRE(X && C && labels, "emma_assign_noalloc: NULL pointer");
    hipStream_t s = (hipStream_t)stream;
    if (mode == EMMA_ASSIGN_FP32_PLAIN)  // the plain loop needs no workspace
        return assign_run(X, n, d, C, k, labels, mode, n_rechecked, nullptr, nullptr, s);
    // the center images and constants take a fixed part of the scratch; the
    // recheck arrays (56 bytes per frame) the rest, so the frames run
    // in chunks
    const size_t fixed = assign_workspace_bytes(0, d, k, mode);
    constexpr size_t cap = EMMA_ASSIGN_NOALLOC_SCRATCH_BYTES;
    // per frame: list (8 bytes), its upper bound (4), list2 (8), a candidate
    // slot (36), with room for the 256-byte alignments
    const int64_t chunk_max = fixed + 8192 < cap ? (int64_t)((cap - fixed - 8192) / 64) : 0;
    EMMA_REQUIRE(chunk_max >= EMMA_ASSIGN_NOALLOC_MIN_CHUNK,
                 "emma_assign_noalloc: k=%d centers of d=%d take %zu of the %zu scratch bytes; at most %lld bytes "
                 "leave room for %d frames per chunk (use emma_assign with a workspace)",
                 k, d, fixed, cap, (long long)(cap - 8192 - 64 * (size_t)EMMA_ASSIGN_NOALLOC_MIN_CHUNK),
                 EMMA_ASSIGN_NOALLOC_MIN_CHUNK);
    const int64_t chunk = chunk_max < ((int64_t)1 << 22) ? chunk_max : ((int64_t)1 << 22);
    EMMA_REQUIRE(assign_workspace_bytes(chunk, d, k, mode) <= cap, "emma_assign_noalloc: chunk of %lld frames "
                 "exceeds the scratch", (long long)chunk);
    void *ws = nullptr;
    EMMA_TRY(scratch_on_current_device(&ws));
    if (n_rechecked) k_zero_i64n<<<1, 64, 0, s>>>(n_rechecked, 1);
    // the list length of each chunk is the first int64 of the carved workspace
    const AssignWs w = carve_assign(ws, chunk, d, k, mode);
    for (int64_t f0 = 0; f0 < n; f0 += chunk) {
        const int64_t m = n - f0 < chunk ? n - f0 : chunk;
        EMMA_TRY(assign_run(X + f0 * (int64_t)d, m, d, C, k, labels + f0, mode, nullptr, ws, nullptr, s));
        if (n_rechecked) k_add_i64<<<1, 1, 0, s>>>(n_rechecked, w.rl.count);
    }
    return check_launch("emma_assign_noalloc");
}
