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
// One wavefront per frame row block; lanes own output columns.
// ------------------------------------------------------------------
constexpr int PROJ_ROWS = 16;  // frames per workgroup

__global__ __launch_bounds__(256) void k_project(const float *__restrict__ X, int64_t n, int d,
                                                 const double *__restrict__ mean,
                                                 const double *__restrict__ V, int ldv, int m,
                                                 float *__restrict__ Y) {
    extern __shared__ double s_x[];  // PROJ_ROWS x d centered frames
    const int64_t r0 = (int64_t)blockIdx.x * PROJ_ROWS;
    const int nr = (int)((n - r0) < PROJ_ROWS ? (n - r0) : PROJ_ROWS);
    for (int e = threadIdx.x; e < nr * d; e += blockDim.x) {
        int r = e / d, c = e - r * d;
        s_x[e] = (double)X[(r0 + r) * (int64_t)d + c] - mean[c];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nr * m; e += blockDim.x) {
        int r = e / m, j = e - r * m;
        const double *x = s_x + (int64_t)r * d;
        double acc = 0.0;
        for (int c = 0; c < d; ++c) acc = fma(x[c], V[(int64_t)c * ldv + j], acc);
        Y[(r0 + r) * (int64_t)m + j] = (float)acc;
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
    size_t lds = sizeof(double) * PROJ_ROWS * (size_t)d;
    EMMA_REQUIRE(lds <= 64 * 1024, "emma_project: d=%d too large (max %d)", d, 64 * 1024 / (8 * PROJ_ROWS));
    hipStream_t s = (hipStream_t)stream;
    k_project<<<(unsigned)cdiv(n, PROJ_ROWS), 256, lds, s>>>(X, n, d, mean, V, ldv, m, Y);
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
