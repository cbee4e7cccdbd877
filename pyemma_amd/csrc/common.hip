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
// TICA projection: Y[i, :m] = (X[i] - mean) V, FP64 accumulation.
// One thread per frame (PROJ_ROWS = 256 frames per workgroup) holds MB
// output accumulators; each output is a sequential FP64 fma chain over the
// dims in order, x_c = (double)X[i, c] - mean[c].  The frames' dims are
// staged in chunks of PROJ_DC: coalesced loads, written transposed into LDS
// so that a thread reads its own frame's dims conflict-free; the chunk's V
// rows and mean sit in LDS and are read as broadcasts.  The next chunk's
// frames, V rows and mean are loaded while this chunk computes.
// ------------------------------------------------------------------
constexpr int PROJ_ROWS = 256;  // frames per workgroup (one per thread)
constexpr int PROJ_DC = 32;     // dims per staged chunk
constexpr int PROJ_MB = 8;      // outputs per pass over the dims

// MB outputs per pass over the dims; the chunk's V rows are loaded with its
// frames, one chunk ahead.  With MB = 8 a pass reads the frames at about
// 3.6 TB/s (10^7 x 300 frames: 3.4 ms); wider blocks make the broadcast V
// reads from LDS the limit (16 outputs: 7 ms per pass), so m > 8 takes
// ceil(m / 8) passes.
template <int MB>
__global__ __launch_bounds__(PROJ_ROWS) void k_project(const float *__restrict__ X, int64_t n, int d,
                                                       const double *__restrict__ mean,
                                                       const double *__restrict__ V, int ldv, int m,
                                                       float *__restrict__ Y) {
    __shared__ float s_x[PROJ_DC][PROJ_ROWS + 1];
    __shared__ double s_v[PROJ_DC][MB];
    __shared__ double s_mean[PROJ_DC];
    constexpr int PER = PROJ_DC;                           // frame loads per thread per chunk
    constexpr int VPER = (PROJ_DC * MB + PROJ_ROWS - 1) / PROJ_ROWS;  // V loads per thread per chunk
    const int tid = threadIdx.x;
    const int64_t r0 = (int64_t)blockIdx.x * PROJ_ROWS;
    const int nr = (int)((n - r0) < PROJ_ROWS ? (n - r0) : PROJ_ROWS);
    const int n_chunks = (d + PROJ_DC - 1) / PROJ_DC;
    // staging map: thread t loads dim t % 32 of rows t / 32 + 8 it
    const int lcol = tid & (PROJ_DC - 1);
    const int lrow0 = tid / PROJ_DC;
    const float *xb = X + r0 * (int64_t)d;
    float pre[PER];
    double prev[VPER], prem = 0.0;
    auto load_chunk = [&](int ch, int j0, int mb) __attribute__((always_inline)) {
        const int c0 = ch * PROJ_DC;
        const int c = c0 + lcol;
#pragma unroll
        for (int it = 0; it < PER; ++it) {
            const int r = lrow0 + it * (PROJ_ROWS / PROJ_DC);
            pre[it] = (r < nr && c < d) ? xb[(uint32_t)(r * d + c)] : 0.0f;  // < 256 d elements
        }
#pragma unroll
        for (int u = 0; u < VPER; ++u) {
            const int e = u * PROJ_ROWS + tid, cc = e / MB, jj = e - cc * MB;
            prev[u] = (e < PROJ_DC * MB && c0 + cc < d && jj < mb) ? V[(int64_t)(c0 + cc) * ldv + j0 + jj] : 0.0;
        }
        if (tid < PROJ_DC) prem = c0 + tid < d ? mean[c0 + tid] : 0.0;
    };
    for (int j0 = 0; j0 < m; j0 += MB) {
        const int mb = (m - j0) < MB ? (m - j0) : MB;
        double acc[MB];
#pragma unroll
        for (int j = 0; j < MB; ++j) acc[j] = 0.0;
        load_chunk(0, j0, mb);
        for (int ch = 0; ch < n_chunks; ++ch) {
            const int nc = (d - ch * PROJ_DC) < PROJ_DC ? (d - ch * PROJ_DC) : PROJ_DC;
            __syncthreads();  // the previous chunk's LDS reads are done
#pragma unroll
            for (int it = 0; it < PER; ++it) s_x[lcol][lrow0 + it * (PROJ_ROWS / PROJ_DC)] = pre[it];
#pragma unroll
            for (int u = 0; u < VPER; ++u) {
                const int e = u * PROJ_ROWS + tid;
                if (e < PROJ_DC * MB) s_v[e / MB][e % MB] = prev[u];
            }
            if (tid < PROJ_DC) s_mean[tid] = prem;
            __syncthreads();
            if (ch + 1 < n_chunks) load_chunk(ch + 1, j0, mb);
#pragma unroll 2
            for (int c = 0; c < nc; ++c) {
                const double x = (double)s_x[c][tid] - s_mean[c];
#pragma unroll
                for (int j = 0; j < MB; ++j) acc[j] = fma(x, s_v[c][j], acc[j]);
            }
        }
        if (tid < nr) {
            float *y = Y + (r0 + tid) * (int64_t)m + j0;
            for (int j = 0; j < mb; ++j) y[j] = (float)acc[j];
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
    hipStream_t s = (hipStream_t)stream;
    k_project<PROJ_MB><<<(unsigned)cdiv(n, PROJ_ROWS), PROJ_ROWS, 0, s>>>(X, n, d, mean, V, ldv, m, Y);
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
