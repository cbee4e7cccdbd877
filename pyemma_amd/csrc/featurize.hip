// Pairwise-distance features of particle coordinates (the producer of the
// 300-d distance features the moments and assignment kernels consume).
//
// out[f, p] = sqrt((dx * dx + dy * dy) + dz * dz) with (dx, dy, dz) = x[f, j] - x[f, i]
// for pair p = (i, j), every operation a correctly rounded FP32 operation in
// that order (no fused multiply-add): the non-periodic distance of
// mdtraj.compute_distances as DistanceFeature.transform calls it; with the
// INVERSE flag 1 / distance (InverseDistanceFeature).  A pair with an index
// outside [0, n_atoms) gives NaN instead of reading out of bounds.
//
// HBM-bound: one workgroup stages a block of frames (coordinates, coalesced)
// in LDS and writes the block's n_frames x n_pairs features with consecutive
// threads on consecutive output elements.
#include "emma_common.hpp"

namespace emma {
namespace feat {

constexpr int NT = 256;
constexpr int LDS_FLOATS = 12288;  // 48 KiB of staged coordinates per workgroup

__global__ __launch_bounds__(NT) void k_pair_distances(const float *__restrict__ xyz, int64_t n_frames,
                                                       int n_atoms, const int32_t *__restrict__ pairs,
                                                       int n_pairs, int frames_per_block, int inverse,
                                                       float *__restrict__ out) {
    __shared__ float s_xyz[LDS_FLOATS];
    const int64_t f0 = (int64_t)blockIdx.x * frames_per_block;
    const int nf = (int)((n_frames - f0) < frames_per_block ? (n_frames - f0) : frames_per_block);
    const int per_frame = 3 * n_atoms;
    const int nval = nf * per_frame;  // <= LDS_FLOATS by the host's choice of frames_per_block
    const float *src = xyz + f0 * per_frame;
    for (int e = threadIdx.x; e < nval; e += NT) s_xyz[e] = src[e];
    __syncthreads();
    const int nout = nf * n_pairs;  // < 2^31 by the host's choice of frames_per_block
    float *dst = out + f0 * n_pairs;
    for (int e = threadIdx.x; e < nout; e += NT) {
        const int f = e / n_pairs;
        const int p = e - f * n_pairs;
        const int i = pairs[2 * p], j = pairs[2 * p + 1];
        float r;
        if (i < 0 || i >= n_atoms || j < 0 || j >= n_atoms) {
            r = __int_as_float(0x7fc00000);
        } else {
            const float *a = s_xyz + f * per_frame + 3 * i;
            const float *b = s_xyz + f * per_frame + 3 * j;
            const float dx = __fsub_rn(b[0], a[0]);
            const float dy = __fsub_rn(b[1], a[1]);
            const float dz = __fsub_rn(b[2], a[2]);
            const float s = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
            r = sqrtf(s);  // correctly rounded expansion (approximate root + one-ulp fma correction)
            if (inverse) r = __fdiv_rn(1.0f, r);
        }
        dst[e] = r;
    }
}

}  // namespace feat
}  // namespace emma

using namespace emma;
using namespace emma::feat;

extern "C" int emma_pair_distances(const float *xyz, int64_t n_frames, int32_t n_atoms, const int32_t *pairs,
                                   int32_t n_pairs, uint32_t flags, float *out, void *stream) {
    EMMA_REQUIRE(n_frames >= 0 && n_atoms >= 1 && n_pairs >= 1 && n_pairs <= (1 << 30),
                 "emma_pair_distances: bad sizes n_frames=%lld n_atoms=%d n_pairs=%d", (long long)n_frames,
                 n_atoms, n_pairs);
    EMMA_REQUIRE(3 * n_atoms <= LDS_FLOATS, "emma_pair_distances: n_atoms=%d too large (max %d)", n_atoms,
                 LDS_FLOATS / 3);
    EMMA_REQUIRE((flags & ~EMMA_DIST_INVERSE) == 0, "emma_pair_distances: unknown flags 0x%x", flags);
    if (n_frames == 0) return EMMA_OK;
    EMMA_REQUIRE(xyz && pairs && out, "emma_pair_distances: NULL pointer");
    // frames per workgroup: fill the staging buffer, but keep >= 1024 workgroups when there is work for them
    int fpb = LDS_FLOATS / (3 * n_atoms);
    const int64_t want = cdiv(n_frames, 1024);
    if (want < fpb) fpb = (int)(want < 1 ? 1 : want);
    if ((int64_t)fpb * n_pairs > INT32_MAX - NT) fpb = (int)((INT32_MAX - NT) / n_pairs);
    if (fpb < 1) fpb = 1;
    hipStream_t s = (hipStream_t)stream;
    k_pair_distances<<<(unsigned)cdiv(n_frames, fpb), NT, 0, s>>>(xyz, n_frames, n_atoms, pairs, n_pairs, fpb,
                                                                  (flags & EMMA_DIST_INVERSE) ? 1 : 0, out);
    return check_launch("emma_pair_distances");
}
