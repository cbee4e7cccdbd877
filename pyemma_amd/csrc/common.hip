// Library-wide entry points: version, error state, input checks, TICA projection.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include "emma_common.hpp"

namespace emma {

static thread_local char g_err[512] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("%s: %s", what, hipGetErrorString(e));
        return EMMA_ERR_HIP;
    }
    return EMMA_OK;
}

// ------------------------------------------------------------------
// non-finite count: grid-stride, one int64 atomic per workgroup
// ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_count_nonfinite(const float *__restrict__ X, int64_t n,
                                                         unsigned long long *__restrict__ count) {
    __shared__ unsigned int s_cnt;
    if (threadIdx.x == 0) s_cnt = 0;
    __syncthreads();
    unsigned int local = 0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint32_t u = __float_as_uint(X[i]);
        local += ((u & 0x7F800000u) == 0x7F800000u) ? 1u : 0u;
    }
    if (local) atomicAdd(&s_cnt, local);
    __syncthreads();
    if (threadIdx.x == 0 && s_cnt) atomicAdd(count, (unsigned long long)s_cnt);
}

__global__ void k_zero_i64(int64_t *p) { *p = 0; }

// ------------------------------------------------------------------
// TICA projection: Y[i, :m] = (X[i] - mean) V on FP64 MFMA
// (v_mfma_f64_16x16x4f64: D[16 frames][16 outputs] += A[16][4 dims] B[4][16]).
// A workgroup holds 256 frames (4 waves x 4 row tiles of 16) and, per pass,
// NT column tiles of 16 outputs.  The frames' dims are staged in chunks of
// PROJ_DC: coalesced loads, written transposed into LDS, together with the
// chunk's rows of V (zero past d and m) and of the mean; the next chunk's
// loads are issued before this chunk's MFMAs.  Lane layout (l = lane):
// A[i][k] = x[l % 16][l / 16] - mean, B[k][j] = V[l / 16][l % 16], D holds
// rows 4 r + l / 16, r = 0..3, of column l % 16.  Every product and sum is
// FP64; the outputs are rounded to float32 once.
// ------------------------------------------------------------------
typedef double v4d __attribute__((ext_vector_type(4)));
constexpr int PROJ_ROWS = 256;                // frames per workgroup
constexpr int PROJ_DC = 32;                   // dims per staged chunk
constexpr int PROJ_XS = PROJ_ROWS + 17;       // LDS row stride of the staged frames (floats)
// LDS row stride of the staged V rows (doubles): NT tiles of 16 plus 2
// (rows of one read start 2 banks apart)
template <int NT>
constexpr int proj_vs() { return 16 * NT + 2; }

// VEC: frames loaded as 16-byte vectors (d % 4 == 0, X 16-byte aligned):
// thread t loads dims 4 (t % 8) .. + 3 of the chunk for rows t / 8 + 32 it
template <int NT, bool VEC>
// one output tile: at least 3 workgroups per CU (the register budget of 3
// waves per SIMD holds without spills); two tiles: the compiler's choice
__global__ __launch_bounds__(PROJ_ROWS, NT == 1 ? 3 : 1) void k_project(const float *__restrict__ X, int64_t n,
                                                                       int d, const double *__restrict__ mean,
                                                                       const double *__restrict__ V, int ldv, int m,
                                                                       float *__restrict__ Y) {
    __shared__ float s_x[PROJ_DC][PROJ_XS];
    __shared__ double s_v[PROJ_DC][proj_vs<NT>()];
    __shared__ double s_mean[PROJ_DC];
    constexpr int PER = PROJ_DC;                            // frame values per thread per chunk
    constexpr int NJ = 16 * NT;                             // outputs per pass
    constexpr int VPER = PROJ_DC * NJ / PROJ_ROWS;          // V loads per thread per chunk
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int64_t r0 = (int64_t)blockIdx.x * PROJ_ROWS;
    const int nr = (int)((n - r0) < PROJ_ROWS ? (n - r0) : PROJ_ROWS);
    const int n_chunks = (d + PROJ_DC - 1) / PROJ_DC;
    // staging map: thread t loads dim t % 32 of rows t / 32 + 8 it
    const int lcol = tid & (PROJ_DC - 1);
    const int lrow0 = tid / PROJ_DC;
    const float *xb = X + r0 * (int64_t)d;
    const uint32_t rstep = (uint32_t)(PROJ_ROWS / PROJ_DC) * (uint32_t)d;  // 8 rows
    // last element of the workgroup's nr frames: a load of a row >= nr is
    // clamped to it (every load stays inside the block; those rows belong to
    // outputs that are not stored); dims >= d are staged as zero
    const uint32_t off_max = (uint32_t)(nr * d - 1);
    float pre[PER];
    double prev[VPER], prem = 0.0;
    // vector staging map: float4 q = t % 8 of rows t / 8 + 32 it
    const int vq = tid & 7, vrow0 = tid >> 3;
    const uint32_t vstep = 32u * (uint32_t)d;
    auto load_chunk = [&](int ch, int j0) __attribute__((always_inline)) {
        const int c0 = ch * PROJ_DC;
        if constexpr (VEC) {
            const bool cin = c0 + 4 * vq < d;  // d % 4 == 0: a vector is all in or all out
            const uint32_t vmax = off_max - 3;  // last vector of the block
            uint32_t off = (uint32_t)(vrow0 * d + c0 + 4 * vq);
#pragma unroll
            for (int it = 0; it < PER / 4; ++it) {
                const float4 v = *reinterpret_cast<const float4 *>(xb + (off < vmax ? off : vmax));
                pre[4 * it + 0] = cin ? v.x : 0.0f;
                pre[4 * it + 1] = cin ? v.y : 0.0f;
                pre[4 * it + 2] = cin ? v.z : 0.0f;
                pre[4 * it + 3] = cin ? v.w : 0.0f;
                off += vstep;
                asm volatile("" : "+v"(off));
            }
        } else {
            const bool cin = c0 + lcol < d;
            uint32_t off = (uint32_t)(lrow0 * d + c0 + lcol);  // < 256 d elements from xb
#pragma unroll
            for (int it = 0; it < PER; ++it) {
                const float v = xb[off < off_max ? off : off_max];
                pre[it] = cin ? v : 0.0f;
                off += rstep;
                asm volatile("" : "+v"(off));
            }
        }
#pragma unroll
        for (int u = 0; u < VPER; ++u) {
            const int e = u * PROJ_ROWS + tid, cc = e / NJ, jj = e - cc * NJ;
            prev[u] = (c0 + cc < d && j0 + jj < m) ? V[(int64_t)(c0 + cc) * ldv + j0 + jj] : 0.0;
        }
        if (tid < PROJ_DC) prem = c0 + tid < d ? mean[c0 + tid] : 0.0;
    };
    const int li = lane & 15, lk = lane >> 4;
    for (int j0 = 0; j0 < m; j0 += NJ) {
        v4d acc[4][NT];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) acc[rt][ct] = v4d{0.0, 0.0, 0.0, 0.0};
        load_chunk(0, j0);
        for (int ch = 0; ch < n_chunks; ++ch) {
            __syncthreads();  // the previous chunk's LDS reads are done
            if constexpr (VEC) {
#pragma unroll
                for (int it = 0; it < PER / 4; ++it)
#pragma unroll
                    for (int e = 0; e < 4; ++e) s_x[4 * vq + e][vrow0 + 32 * it] = pre[4 * it + e];
            } else {
#pragma unroll
                for (int it = 0; it < PER; ++it) s_x[lcol][lrow0 + it * (PROJ_ROWS / PROJ_DC)] = pre[it];
            }
#pragma unroll
            for (int u = 0; u < VPER; ++u) {
                const int e = u * PROJ_ROWS + tid, cc = e / NJ, jj = e - cc * NJ;
                s_v[cc][jj] = prev[u];
            }
            if (tid < PROJ_DC) s_mean[tid] = prem;
            __syncthreads();
            if (ch + 1 < n_chunks) load_chunk(ch + 1, j0);
#pragma unroll 2
            for (int ks = 0; ks < PROJ_DC / 4; ++ks) {
                const int c = 4 * ks + lk;
                const double mu = s_mean[c];
                double a[4], b[NT];
#pragma unroll
                for (int rt = 0; rt < 4; ++rt) a[rt] = (double)s_x[c][wave * 64 + rt * 16 + li] - mu;
#pragma unroll
                for (int ct = 0; ct < NT; ++ct) b[ct] = s_v[c][ct * 16 + li];
#pragma unroll
                for (int rt = 0; rt < 4; ++rt)
#pragma unroll
                    for (int ct = 0; ct < NT; ++ct)
                        acc[rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt], b[ct], acc[rt][ct], 0, 0, 0);
            }
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
            for (int ct = 0; ct < NT; ++ct) {
                const int j = j0 + ct * 16 + li;
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int f = wave * 64 + rt * 16 + 4 * r + lk;
                    if (f < nr && j < m) Y[(r0 + f) * (int64_t)m + j] = (float)acc[rt][ct][r];
                }
            }
    }
}

}  // namespace emma

using namespace emma;

extern "C" const char *emma_version(void) { return "emma_amd 0.1.0 (gfx950)"; }

extern "C" const char *emma_last_error(void) { return g_err; }

extern "C" int emma_count_nonfinite(const float *X, int64_t n, int64_t *count, void *stream) {
    EMMA_REQUIRE(n >= 0, "emma_count_nonfinite: n < 0");
    EMMA_REQUIRE(count != nullptr, "emma_count_nonfinite: count is NULL");
    hipStream_t s = (hipStream_t)stream;
    k_zero_i64<<<1, 1, 0, s>>>(count);
    if (n > 0) {
        EMMA_REQUIRE(X != nullptr, "emma_count_nonfinite: X is NULL");
        int64_t blocks = cdiv(n, 256 * 8);
        if (blocks > 4096) blocks = 4096;
        k_count_nonfinite<<<(int)blocks, 256, 0, s>>>(X, n, (unsigned long long *)count);
    }
    return check_launch("emma_count_nonfinite");
}

extern "C" int emma_project(const float *X, int64_t n, int32_t d, const double *mean,
                            const double *V, int32_t ldv, int32_t m, float *Y, void *stream) {
    EMMA_REQUIRE(n >= 0 && d > 0 && m >= 0 && ldv >= m, "emma_project: bad sizes n=%lld d=%d m=%d ldv=%d",
                 (long long)n, d, m, ldv);
    if (n == 0 || m == 0) return EMMA_OK;
    EMMA_REQUIRE(X && mean && V && Y, "emma_project: NULL pointer");
    // 32-bit element offsets inside a workgroup's block of 256 frames
    EMMA_REQUIRE(d <= (1 << 22), "emma_project: d=%d too large (max %d)", d, 1 << 22);
    hipStream_t s = (hipStream_t)stream;
    const unsigned grid = (unsigned)cdiv(n, PROJ_ROWS);
    const bool vec = d % 4 == 0 && ((uintptr_t)X & 15) == 0;
    if (m <= 16) {
        if (vec)
            k_project<1, true><<<grid, PROJ_ROWS, 0, s>>>(X, n, d, mean, V, ldv, m, Y);
        else
            k_project<1, false><<<grid, PROJ_ROWS, 0, s>>>(X, n, d, mean, V, ldv, m, Y);
    } else {
        if (vec)
            k_project<2, true><<<grid, PROJ_ROWS, 0, s>>>(X, n, d, mean, V, ldv, m, Y);
        else
            k_project<2, false><<<grid, PROJ_ROWS, 0, s>>>(X, n, d, mean, V, ldv, m, Y);
    }
    return check_launch("emma_project");
}

// ------------------------------------------------------------------
// row affine map: out[i] = X[i,:] . u + u0 in FP64, one wavefront per row
// ------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_row_affine(const float *__restrict__ X, int64_t n, int d,
                                                    const double *__restrict__ u, double u0,
                                                    double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= n) return;
    const float *x = X + row * d;
    double a = 0.0;
    for (int c = lane; c < d; c += 64) a = fma((double)x[c], u[c], a);
    a = emma::wave_butterfly_sum(a);
    if (lane == 0) out[row] = a + u0;
}

extern "C" int emma_row_affine(const float *X, int64_t n, int32_t d, const double *u, double u0, double *out,
                               void *stream) {
    EMMA_REQUIRE(n >= 0 && d > 0, "emma_row_affine: bad sizes n=%lld d=%d", (long long)n, d);
    if (n == 0) return EMMA_OK;
    EMMA_REQUIRE(X && u && out, "emma_row_affine: NULL pointer");
    k_row_affine<<<(unsigned)cdiv(n, 4), 256, 0, (hipStream_t)stream>>>(X, n, d, u, u0, out);
    return check_launch("row affine");
}

// ------------------------------------------------------------------
// matrix-core rounding probe: one v_mfma_f32_32x32x16_bf16, one wave.
// Operand layout of the instruction (lane l): A row l % 32, dims
// 8 (l / 32) .. + 7; B column l % 32, the same dims; D column l % 32, rows
// 8 (e / 4) + 4 (l / 32) + e % 4 for e = 0..15.
// ------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_probe_mfma_bf16(const uint16_t *__restrict__ A,
                                                        const uint16_t *__restrict__ B,
                                                        const float *__restrict__ C, float *__restrict__ D) {
    const int lane = threadIdx.x;
    const int r = lane & 31, kh = (lane >> 5) * 8;
    emma::bf16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        a[i] = __builtin_bit_cast(__bf16, A[r * 16 + kh + i]);
        b[i] = __builtin_bit_cast(__bf16, B[(kh + i) * 32 + r]);
    }
    emma::floatx16 acc;
#pragma unroll
    for (int e = 0; e < 16; ++e) acc[e] = C[(8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)) * 32 + r];
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int e = 0; e < 16; ++e) D[(8 * (e >> 2) + 4 * (lane >> 5) + (e & 3)) * 32 + r] = acc[e];
}

extern "C" int emma_probe_mfma_bf16(const uint16_t *A, const uint16_t *B, const float *C, float *D, void *stream) {
    EMMA_REQUIRE(A && B && C && D, "emma_probe_mfma_bf16: NULL pointer");
    k_probe_mfma_bf16<<<1, 64, 0, (hipStream_t)stream>>>(A, B, C, D);
    return check_launch("mfma probe");
}
