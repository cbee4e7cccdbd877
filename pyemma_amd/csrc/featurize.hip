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
// threads on consecutive output elements.  LDS is sized per launch (dynamic):
// the block's coordinates plus at most PAIR_CHUNK packed pairs.
#include "emma_common.hpp"

namespace emma {
namespace feat {

constexpr int NT = 256;
constexpr int LDS_FLOATS = 8192;   // at most 32 KiB of staged coordinates per workgroup
constexpr int PAIR_CHUNK = 4096;   // at most this many pairs staged per pass (16 KiB)
// coordinates staged per workgroup when frames are small (16 KiB): several
// workgroups per CU hide one another's staging round trip (measured on 10^7
// frames x 25 atoms: 4.0 ms at 16 KiB, 4.4 / 4.1 / 4.2 ms at 4 / 8 / 32 KiB)
constexpr int TARGET_FLOATS = 4096;

// One workgroup: a block of frames staged in LDS, then the pairs in chunks
// of PAIR_CHUNK, each pair packed as the LDS offsets (3 i) | (3 j) << 16 (a
// pair with an atom out of range as offsets 0 | 0 plus a NaN flag in bit 31).
// The staging loads are issued in batches of SB per thread before any is
// written to LDS, so a batch costs one memory round trip.  The element loop
// walks (frame, pair) with incremental counters, so it has no division and
// no global loads: LDS reads, VALU and one coalesced store at a 32-bit
// offset from the workgroup's output block.
constexpr int SB = 16;                 // staging loads in flight per thread
constexpr uint32_t NAN_FLAG = 0x80000000u;

template <bool INV>
__global__ __launch_bounds__(NT) void k_pair_distances(const float *__restrict__ xyz, int64_t n_frames,
                                                       int n_atoms, const int32_t *__restrict__ pairs,
                                                       int n_pairs, int frames_per_block, int pair_slots,
                                                       float *__restrict__ out) {
    // dynamic LDS: frames_per_block * 3 n_atoms coordinates, then pair_slots packed pairs
    extern __shared__ float s_dyn[];
    float *s_xyz = s_dyn;
    uint32_t *s_pair = reinterpret_cast<uint32_t *>(s_dyn + frames_per_block * 3 * n_atoms);
    const int tid = threadIdx.x;
    const int64_t f0 = (int64_t)blockIdx.x * frames_per_block;
    const int nf = (int)((n_frames - f0) < frames_per_block ? (n_frames - f0) : frames_per_block);
    const int per_frame = 3 * n_atoms;
    const int nval = nf * per_frame;  // <= LDS_FLOATS by the host's choice of frames_per_block
    const float *src = xyz + f0 * per_frame;
    for (int e0 = 0; e0 < nval; e0 += SB * NT) {
        float v[SB];
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int e = e0 + u * NT + tid;
            v[u] = e < nval ? src[e] : 0.0f;
        }
#pragma unroll
        for (int u = 0; u < SB; ++u) {
            const int e = e0 + u * NT + tid;
            if (e < nval) s_xyz[e] = v[u];
        }
    }
    float *dst = out + f0 * n_pairs;  // nf * n_pairs <= 2^30 elements by the host's choice
    for (int pc = 0; pc < n_pairs; pc += pair_slots) {
        const int np = (n_pairs - pc) < pair_slots ? (n_pairs - pc) : pair_slots;
        __syncthreads();  // coordinates staged; the previous chunk's pairs no longer read
        for (int q0 = 0; q0 < np; q0 += SB * NT) {
            int pi[SB], pj[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int q = q0 + u * NT + tid;
                pi[u] = q < np ? pairs[2 * (pc + q)] : 0;
                pj[u] = q < np ? pairs[2 * (pc + q) + 1] : 0;
            }
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int q = q0 + u * NT + tid;
                const int i = pi[u], j = pj[u];
                const bool ok = i >= 0 && i < n_atoms && j >= 0 && j < n_atoms;
                if (q < np) s_pair[q] = ok ? (uint32_t)(3 * i) | ((uint32_t)(3 * j) << 16) : NAN_FLAG;
            }
        }
        __syncthreads();
        const int total = nf * np;
        const int df = NT / np, dp = NT - df * np;
        int f = tid / np, p = tid - f * np;
        uint32_t o = (uint32_t)(f * n_pairs + pc + p);  // output offset of (f, p)
        const uint32_t do_ = (uint32_t)(df * n_pairs + dp), wrap = (uint32_t)(n_pairs - np);
        for (int e = tid; e < total; e += NT) {
            const uint32_t pk = s_pair[p];
            const float *fr = s_xyz + f * per_frame;
            const float *a = fr + (pk & 0xFFFFu);
            const float *b = fr + ((pk >> 16) & 0x7FFFu);
            const float dx = __fsub_rn(b[0], a[0]);
            const float dy = __fsub_rn(b[1], a[1]);
            const float dz = __fsub_rn(b[2], a[2]);
            const float s2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
            float r = sqrtf(s2);  // correctly rounded expansion (approximate root + one-ulp fma correction)
            if constexpr (INV) r = __fdiv_rn(1.0f, r);
            r = (pk & NAN_FLAG) ? __int_as_float(0x7fc00000) : r;
            dst[o] = r;
            p += dp;
            f += df;
            o += do_;
            if (p >= np) {
                p -= np;
                ++f;
                o += wrap;
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
    // frames per workgroup: about TARGET_FLOATS staged coordinates (at least one
    // frame), but keep >= 1024 workgroups when there is work for them
    int fpb = TARGET_FLOATS / (3 * n_atoms);
    if (fpb < 1) fpb = 1;
    const int64_t want = cdiv(n_frames, 1024);
    if (want < fpb) fpb = (int)(want < 1 ? 1 : want);
    if ((int64_t)fpb * PAIR_CHUNK > INT32_MAX - NT) fpb = (int)((INT32_MAX - NT) / PAIR_CHUNK);
    // 32-bit output byte offsets inside a workgroup's block of fpb x n_pairs elements
    if ((int64_t)fpb * n_pairs > (1 << 30)) fpb = (int)((1 << 30) / n_pairs);
    if (fpb < 1) fpb = 1;
    hipStream_t s = (hipStream_t)stream;
    const int pair_slots = n_pairs < PAIR_CHUNK ? n_pairs : PAIR_CHUNK;
    const size_t shm = ((size_t)fpb * 3 * n_atoms + pair_slots) * 4;
    const unsigned grid = (unsigned)cdiv(n_frames, fpb);
    if (flags & EMMA_DIST_INVERSE)
        k_pair_distances<true><<<grid, NT, shm, s>>>(xyz, n_frames, n_atoms, pairs, n_pairs, fpb, pair_slots, out);
    else
        k_pair_distances<false><<<grid, NT, shm, s>>>(xyz, n_frames, n_atoms, pairs, n_pairs, fpb, pair_slots, out);
    return check_launch("emma_pair_distances");
}
