// Transition counts of discrete trajectories (the consumer of assign's labels).
//
// C[i, j] = number of t with s[t] = i and s[t + lag] = j inside one trajectory,
// for every t (sliding window) or for t = 0, lag, 2 lag, ... (sample mode);
// hist[i] = number of frames in state i.  Frames or pairs with a negative
// state (-1: unassigned) are not counted; states below -1 or at or above
// nstates are counted in n_bad and not written.
//
// One thread per frame.  Metastable trajectories repeat the same pair for
// long stretches, so each wave merges runs of equal keys over consecutive
// lanes (a ballot of run heads) and issues one 64-bit atomic per run.
#include "emma_common.hpp"

namespace emma {
namespace cnt {

constexpr int NT = 256;

__global__ __launch_bounds__(NT) void k_fill_u64(unsigned long long *__restrict__ p, int64_t n,
                                                 unsigned long long v) {
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Adds the run length of each run of equal keys (consecutive lanes) at base[key];
// negative keys are skipped.  Every lane of the wave takes part.
__device__ __forceinline__ void add_runs(int64_t key, unsigned long long *__restrict__ base, int lane) {
    const int64_t prev = __shfl_up(key, 1);
    const bool head = (lane == 0) || (key != prev);
    const uint64_t heads = __ballot(head);
    const uint64_t above = (lane == 63) ? 0ull : (heads >> (lane + 1));
    const int len = above ? (__builtin_ctzll(above) + 1) : (64 - lane);
    if (head && key >= 0) atomicAdd(base + key, (unsigned long long)len);
}

__global__ __launch_bounds__(NT) void k_counts(const int32_t *__restrict__ S, int64_t n,
                                               const int64_t *__restrict__ off, int ntraj, int lag,
                                               int sliding, int nstates, unsigned long long *__restrict__ C,
                                               unsigned long long *__restrict__ hist,
                                               unsigned long long *__restrict__ n_bad) {
    const int lane = threadIdx.x & 63;
    const int64_t wave0 = (int64_t)blockIdx.x * NT + (threadIdx.x & ~63);
    const int64_t g = wave0 + lane;
    // trajectory of the wave's first frame (binary search, same for all lanes),
    // then each lane walks forward over the few boundaries inside the wave
    int lo = 0, hi = ntraj;  // off[lo] <= wave0 < off[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= wave0) lo = mid;
        else hi = mid;
    }
    int64_t key = -1, hkey = -1;
    unsigned int bad = 0;
    if (g < n) {
        int tr = lo;
        while (tr + 1 < ntraj && off[tr + 1] <= g) ++tr;
        const int64_t t = g - off[tr];
        const int64_t T = off[tr + 1] - off[tr];
        const int s0 = S[g];
        bad += (s0 < -1 || s0 >= nstates) ? 1u : 0u;
        if (s0 >= 0 && s0 < nstates) hkey = s0;
        if (t + lag < T && g + lag < n && (sliding || (t % lag) == 0)) {
            const int s1 = S[g + lag];
            // the pair's second frame is checked where it is the first frame
            if (s0 >= 0 && s0 < nstates && s1 >= 0 && s1 < nstates) key = (int64_t)s0 * nstates + s1;
        }
    }
    add_runs(key, C, lane);
    if (hist) add_runs(hkey, hist, lane);
    const uint64_t badm = __ballot(bad != 0);
    if (lane == 0 && badm) atomicAdd(n_bad, (unsigned long long)__builtin_popcountll(badm));
}

__global__ __launch_bounds__(NT) void k_max_state(const int32_t *__restrict__ S, int64_t n, int *__restrict__ out) {
    int m = -1;
    const int64_t stride = (int64_t)gridDim.x * NT;
    for (int64_t i = (int64_t)blockIdx.x * NT + threadIdx.x; i < n; i += stride) m = max(m, S[i]);
    for (int o = 32; o > 0; o >>= 1) m = max(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ void k_set_i32(int *p, int v) { *p = v; }

}  // namespace cnt
}  // namespace emma

using namespace emma;
using namespace emma::cnt;

extern "C" int emma_count_transitions(const int32_t *dtraj, int64_t n, const int64_t *traj_offsets,
                                      int32_t ntraj, int32_t lag, int32_t sliding, int32_t nstates, int64_t *C,
                                      int64_t *hist, int64_t *n_bad, void *stream) {
    EMMA_REQUIRE(n >= 0 && ntraj >= 1 && lag >= 1 && nstates >= 1,
                 "emma_count_transitions: bad sizes n=%lld ntraj=%d lag=%d nstates=%d", (long long)n, ntraj, lag,
                 nstates);
    EMMA_REQUIRE((int64_t)nstates * nstates <= ((int64_t)1 << 31),
                 "emma_count_transitions: nstates=%d too large for a dense count matrix", nstates);
    EMMA_REQUIRE(traj_offsets && C && n_bad, "emma_count_transitions: NULL pointer");
    EMMA_REQUIRE(n == 0 || dtraj, "emma_count_transitions: dtraj is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int64_t nc = (int64_t)nstates * nstates;
    k_fill_u64<<<(unsigned)std::min<int64_t>(cdiv(nc, NT * 4), 8192), NT, 0, s>>>((unsigned long long *)C, nc, 0ull);
    if (hist) k_fill_u64<<<(unsigned)std::min<int64_t>(cdiv(nstates, NT), 1024), NT, 0, s>>>(
        (unsigned long long *)hist, nstates, 0ull);
    k_fill_u64<<<1, NT, 0, s>>>((unsigned long long *)n_bad, 1, 0ull);
    if (n > 0)
        k_counts<<<(unsigned)cdiv(n, NT), NT, 0, s>>>(dtraj, n, traj_offsets, ntraj, lag, sliding ? 1 : 0, nstates,
                                                      (unsigned long long *)C, (unsigned long long *)hist,
                                                      (unsigned long long *)n_bad);
    return check_launch("emma_count_transitions");
}

extern "C" int emma_max_state(const int32_t *dtraj, int64_t n, int32_t *out, void *stream) {
    EMMA_REQUIRE(n >= 0 && out, "emma_max_state: bad arguments");
    EMMA_REQUIRE(n == 0 || dtraj, "emma_max_state: dtraj is NULL");
    hipStream_t s = (hipStream_t)stream;
    k_set_i32<<<1, 1, 0, s>>>(out, -1);
    if (n > 0) k_max_state<<<(unsigned)std::min<int64_t>(cdiv(n, NT * 8), 2048), NT, 0, s>>>(dtraj, n, out);
    return check_launch("emma_max_state");
}
