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
constexpr int LDS_FLOATS = 8192;   // 32 KiB of staged coordinates per workgroup
constexpr int PAIR_CHUNK = 4096;   // pairs staged per pass (16 KiB)
constexpr uint32_t BAD_PAIR = 0xFFFFFFFFu;

// One workgroup: a block of frames staged in LDS, then the pairs in chunks
// of PAIR_CHUNK, each pair packed as the LDS offsets (3 i) | (3 j) << 16.  The
// element loop walks (frame, pair) with incremental counters, so it has no
// division and no global loads: LDS reads, VALU and one coalesced store.
__global__ __launch_bounds__(NT) void k_pair_distances(const float *__restrict__ xyz, int64_t n_frames,
                                                       int n_atoms, const int32_t *__restrict__ pairs,
                                                       int n_pairs, int frames_per_block, int inverse,
                                                       float *__restrict__ out) {
    __shared__ float s_xyz[LDS_FLOATS];
    __shared__ uint32_t s_pair[PAIR_CHUNK];
    const int tid = threadIdx.x;
    const int64_t f0 = (int64_t)blockIdx.x * frames_per_block;
    const int nf = (int)((n_frames - f0) < frames_per_block ? (n_frames - f0) : frames_per_block);
    const int per_frame = 3 * n_atoms;
    const int nval = nf * per_frame;  // <= LDS_FLOATS by the host's choice of frames_per_block
    const float *src = xyz + f0 * per_frame;
    for (int e = tid; e < nval; e += NT) s_xyz[e] = src[e];
    float *dst = out + f0 * n_pairs;
    for (int pc = 0; pc < n_pairs; pc += PAIR_CHUNK) {
        const int np = (n_pairs - pc) < PAIR_CHUNK ? (n_pairs - pc) : PAIR_CHUNK;
        __syncthreads();  // coordinates staged; the previous chunk's pairs no longer read
        for (int q = tid; q < np; q += NT) {
            const int i = pairs[2 * (pc + q)], j = pairs[2 * (pc + q) + 1];
            const bool ok = i >= 0 && i < n_atoms && j >= 0 && j < n_atoms;
            s_pair[q] = ok ? (uint32_t)(3 * i) | ((uint32_t)(3 * j) << 16) : BAD_PAIR;
        }
        __syncthreads();
        const int total = nf * np;  // < 2^31 by the host's choice of frames_per_block
        const int df = NT / np, dp = NT - df * np;
        int f = tid / np, p = tid - f * np;
        for (int e = tid; e < total; e += NT) {
            const uint32_t pk = s_pair[p];
            float r;
            if (pk == BAD_PAIR) {
                r = __int_as_float(0x7fc00000);
            } else {
                const float *fr = s_xyz + f * per_frame;
                const float *a = fr + (pk & 0xFFFFu);
                const float *b = fr + (pk >> 16);
                const float dx = __fsub_rn(b[0], a[0]);
                const float dy = __fsub_rn(b[1], a[1]);
                const float dz = __fsub_rn(b[2], a[2]);
                const float s2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
                r = sqrtf(s2);  // correctly rounded expansion (approximate root + one-ulp fma correction)
                if (inverse) r = __fdiv_rn(1.0f, r);
            }
            dst[(int64_t)f * n_pairs + pc + p] = r;
            p += dp;
            f += df;
            if (p >= np) {
                p -= np;
                ++f;
            }
        }
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
    if ((int64_t)fpb * PAIR_CHUNK > INT32_MAX - NT) fpb = (int)((INT32_MAX - NT) / PAIR_CHUNK);
    if (fpb < 1) fpb = 1;
    hipStream_t s = (hipStream_t)stream;
    k_pair_distances<<<(unsigned)cdiv(n_frames, fpb), NT, 0, s>>>(xyz, n_frames, n_atoms, pairs, n_pairs, fpb,
                                                                  (flags & EMMA_DIST_INVERSE) ? 1 : 0, out);
    return check_launch("emma_pair_distances");
}
