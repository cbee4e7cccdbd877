/*
 * CPU oracle for the k-means half of the discretization path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker that the HIP kernels
 * are compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load the library built from it; the
 * product path (pyemma_amd) never calls it.
 *
 * The reference delegates the k-means arithmetic to deeptime
 * (pyemma/coordinates/clustering/kmeans.py:39, interface.py:164-165), whose
 * source is not part of the reference tree.  The rules restated here are the
 * ones fixed here and documented in the functions below:
 *
 *  - assign: label(x) = argmin_j sum_i (x_i - c_ji)^2, the sum evaluated in
 *    FP32 sequentially over i with no fused multiply-add, strict '<' so the
 *    first minimum wins (call site interface.py:157-167, api.py:2060-2156).
 *  - kmeans++: D^2 sampling (Arthur & Vassilvitskii 2007) with a documented
 *    counter-based RNG and a documented FP64 summation order, so that a
 *    parallel implementation can reproduce every draw exactly
 *    (call site kmeans.py:211-212, 254-258; observed entry
 *    clustering/tests/test_kmeans.py:304).
 *  - Lloyd: assign, member means in a documented FP64 order, cost and the
 *    relative-change stopping rule of kmeans.py:67-75 / api.py:1757-1762.
 *
 * Build: oracle/Makefile (gcc -O3 -ffp-contract=off -fopenmp).  The flag
 * -ffp-contract=off is essential: it keeps x*x+acc from being fused.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define EMMA_BLOCK 4096      /* frames per block in every blocked FP64 sum */
#define EMMA_LANES 64        /* lanes per block in the lane-strided partial sums */
#define EMMA_SUPER 64        /* blocks per superblock in the kmeans++ search */
#define EMMA_SEG 64          /* frames per segment in the kmeans++ search */
#define EMMA_CHUNK 256       /* members per chunk in the Lloyd center sums */

/* ------------------------------------------------------------------ */
/* exact reference distance                                            */
/* ------------------------------------------------------------------ */

/* Sequential FP32 squared Euclidean distance, index order, no FMA. */
static inline float emma_dist_seq(const float *x, const float *c, int d) {
    float acc = 0.0f;
    for (int i = 0; i < d; ++i) {
        float t = x[i] - c[i];
        float s = t * t;
        acc = acc + s;
    }
    return acc;
}

float oracle_dist(const float *x, const float *c, int d) { return emma_dist_seq(x, c, d); }

/* Assign every frame to its nearest center (first minimum wins).
 * Follows the semantics of AbstractClustering._transform_array
 * (clustering/interface.py:157-167) with the sequential FP32 loop above. */
void oracle_assign(const float *X, int64_t n, int d, const float *C, int k,
                   int32_t *labels, float *dists, int nthreads) {
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int64_t i = 0; i < n; ++i) {
        const float *x = X + i * (int64_t)d;
        float best = emma_dist_seq(x, C, d);
        int bi = 0;
        for (int j = 1; j < k; ++j) {
            float dj = emma_dist_seq(x, C + (int64_t)j * d, d);
            if (dj < best) {
                best = dj;
                bi = j;
            }
        }
        labels[i] = bi;
        if (dists) dists[i] = best;
    }
}

/* ------------------------------------------------------------------ */
/* documented RNG                                                      */
/* ------------------------------------------------------------------ */

/* SplitMix64 finalizer (Steele, Lea, Flood 2014). */
static inline uint64_t emma_mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

/* Uniform double in [0,1) for draw number `counter` of stream `seed`. */
double oracle_uniform(uint64_t seed, uint64_t counter) {
    uint64_t z = emma_mix64(seed + (counter + 1ULL) * 0x9E3779B97F4A7C15ULL);
    return (double)(z >> 11) * 0x1.0p-53;
}

/* ------------------------------------------------------------------ */
/* documented FP64 block sum                                           */
/* ------------------------------------------------------------------ */

/* Sum of v[lo..hi) (hi-lo <= EMMA_BLOCK) in FP64: lane l (0..63) adds the
 * elements lo+m*64+l, m = 0,1,..., in order starting from 0.0; the 64 lane
 * sums are then combined by a butterfly: for off = 32,16,8,4,2,1 every lane
 * replaces its value s_l by s_l + s_(l xor off).  The result is lane 0's
 * value.  A 64-wide wavefront computes exactly this with __shfl_xor. */
double oracle_block_sum(const float *v, int64_t lo, int64_t hi) {
    double s[EMMA_LANES], t[EMMA_LANES];
    for (int l = 0; l < EMMA_LANES; ++l) {
        double a = 0.0;
        for (int64_t i = lo + l; i < hi; i += EMMA_LANES) a = a + (double)v[i];
        s[l] = a;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        for (int l = 0; l < EMMA_LANES; ++l) t[l] = s[l] + s[l ^ off];
        memcpy(s, t, sizeof(s));
    }
    return s[0];
}

/* Sum of a per-frame FP32 array: block sums as above, then a sequential
 * FP64 sum over the blocks in order. */
double oracle_blocked_sum(const float *v, int64_t n) {
    double tot = 0.0;
    for (int64_t lo = 0; lo < n; lo += EMMA_BLOCK) {
        int64_t hi = lo + EMMA_BLOCK < n ? lo + EMMA_BLOCK : n;
        tot = tot + oracle_block_sum(v, lo, hi);
    }
    return tot;
}

/* ------------------------------------------------------------------ */
/* kmeans++                                                            */
/* ------------------------------------------------------------------ */

/* Pick the next center index from the D^2 weights.
 *
 * Hierarchy: frames -> segments of 64 frames -> blocks of 4096 frames ->
 * superblocks of 64 blocks.  Block sums use oracle_block_sum; superblock
 * sums and segment sums are sequential FP64 sums in natural order.  With
 * total = sequential sum of superblock sums and T = u * total, each level
 * scans its children in order with a running FP64 prefix (starting from the
 * prefix before the parent) and takes the first child whose running prefix
 * exceeds T; if rounding leaves none, it takes the last child with a
 * positive weight.  When total == 0 the pick is floor(u * n). */
int64_t oracle_kmpp_pick(const float *d2, int64_t n, const double *bsum, double u) {
    int64_t nb = (n + EMMA_BLOCK - 1) / EMMA_BLOCK;
    int64_t nsb = (nb + EMMA_SUPER - 1) / EMMA_SUPER;
    double total = 0.0;
    for (int64_t s = 0; s < nsb; ++s) {
        double a = 0.0;
        for (int64_t b = s * EMMA_SUPER; b < nb && b < (s + 1) * EMMA_SUPER; ++b) a = a + bsum[b];
        total = total + a;
    }
    if (!(total > 0.0)) {
        int64_t i = (int64_t)(u * (double)n);
        return i < n ? i : n - 1;
    }
    double T = u * total;
    /* superblock level */
    double run = 0.0, before = 0.0;
    int64_t sel = -1, last_pos = -1;
    double last_before = 0.0;
    for (int64_t s = 0; s < nsb; ++s) {
        double a = 0.0;
        for (int64_t b = s * EMMA_SUPER; b < nb && b < (s + 1) * EMMA_SUPER; ++b) a = a + bsum[b];
        double prev = run;
        run = run + a;
        if (a > 0.0) { last_pos = s; last_before = prev; }
        if (run > T) { sel = s; before = prev; break; }
    }
    if (sel < 0) { sel = last_pos; before = last_before; }
    /* block level */
    int64_t s = sel;
    run = before; sel = -1; last_pos = -1;
    for (int64_t b = s * EMMA_SUPER; b < nb && b < (s + 1) * EMMA_SUPER; ++b) {
        double prev = run;
        run = run + bsum[b];
        if (bsum[b] > 0.0) { last_pos = b; last_before = prev; }
        if (run > T) { sel = b; before = prev; break; }
    }
    if (sel < 0) { sel = last_pos; before = last_before; }
    /* segment level */
    int64_t b = sel;
    int64_t blo = b * EMMA_BLOCK, bhi = blo + EMMA_BLOCK < n ? blo + EMMA_BLOCK : n;
    run = before; sel = -1; last_pos = -1;
    for (int64_t g = blo; g < bhi; g += EMMA_SEG) {
        int64_t ghi = g + EMMA_SEG < bhi ? g + EMMA_SEG : bhi;
        double a = 0.0;
        for (int64_t i = g; i < ghi; ++i) a = a + (double)d2[i];
        double prev = run;
        run = run + a;
        if (a > 0.0) { last_pos = g; last_before = prev; }
        if (run > T) { sel = g; before = prev; break; }
    }
    if (sel < 0) { sel = last_pos; before = last_before; }
    /* frame level */
    int64_t g = sel;
    int64_t ghi = g + EMMA_SEG < bhi ? g + EMMA_SEG : bhi;
    run = before; sel = -1; last_pos = -1;
    for (int64_t i = g; i < ghi; ++i) {
        run = run + (double)d2[i];
        if (d2[i] > 0.0f) last_pos = i;
        if (run > T) { sel = i; break; }
    }
    if (sel < 0) sel = last_pos;
    return sel;
}

/* kmeans++ seeding.  Round 0 picks floor(u_0 * n); round r >= 1 picks by
 * D^2 weights with u_r, where u_r = oracle_uniform(seed, r) and D^2(x) is
 * the minimum over the chosen centers of emma_dist_seq (FP32).
 * Returns 0, or -1 on bad arguments. */
int oracle_kmeanspp(const float *X, int64_t n, int d, int k, uint64_t seed,
                    float *C, int64_t *idx, int nthreads) {
    if (n <= 0 || d <= 0 || k <= 0) return -1;
    if (nthreads < 1) nthreads = 1;
    int64_t nb = (n + EMMA_BLOCK - 1) / EMMA_BLOCK;
    float *d2 = (float *)malloc(sizeof(float) * n);
    double *bsum = (double *)malloc(sizeof(double) * nb);
    if (!d2 || !bsum) { free(d2); free(bsum); return -1; }
    int64_t pick = (int64_t)(oracle_uniform(seed, 0) * (double)n);
    if (pick >= n) pick = n - 1;
    for (int r = 0; r < k; ++r) {
        if (r > 0) pick = oracle_kmpp_pick(d2, n, bsum, oracle_uniform(seed, (uint64_t)r));
        idx[r] = pick;
        const float *c = X + pick * (int64_t)d;
        memcpy(C + (int64_t)r * d, c, sizeof(float) * d);
        if (r == k - 1) break;
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < n; ++i) {
            float di = emma_dist_seq(X + i * (int64_t)d, c, d);
            if (r == 0 || di < d2[i]) d2[i] = di;
        }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t b = 0; b < nb; ++b) {
            int64_t lo = b * EMMA_BLOCK, hi = lo + EMMA_BLOCK < n ? lo + EMMA_BLOCK : n;
            bsum[b] = oracle_block_sum(d2, lo, hi);
        }
    }
    free(d2);
    free(bsum);
    return 0;
}

/* ------------------------------------------------------------------ */
/* Lloyd                                                               */
/* ------------------------------------------------------------------ */

/* New centers from labels.  Members of center j are taken in increasing
 * frame order and grouped into consecutive chunks of 256; each chunk is
 * summed sequentially in FP64, the chunk sums are summed sequentially, and
 * the center becomes (float)(sum / count).  Empty centers keep their old
 * coordinates.  Returns 0 or -1 on allocation failure. */
int oracle_update_centers(const float *X, int64_t n, int d, const int32_t *labels,
                          float *C, int k, int64_t *counts_out) {
    int64_t *cnt = (int64_t *)calloc(k + 1, sizeof(int64_t));
    int64_t *order = (int64_t *)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    double *acc = (double *)malloc(sizeof(double) * d);
    double *chunk = (double *)malloc(sizeof(double) * d);
    if (!cnt || !order || !acc || !chunk) { free(cnt); free(order); free(acc); free(chunk); return -1; }
    for (int64_t i = 0; i < n; ++i) cnt[labels[i] + 1]++;
    for (int j = 0; j < k; ++j) cnt[j + 1] += cnt[j];
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (k > 0 ? k : 1));
    for (int j = 0; j < k; ++j) pos[j] = cnt[j];
    for (int64_t i = 0; i < n; ++i) order[pos[labels[i]]++] = i;
    for (int j = 0; j < k; ++j) {
        int64_t lo = cnt[j], hi = cnt[j + 1];
        if (counts_out) counts_out[j] = hi - lo;
        if (hi == lo) continue;
        for (int c = 0; c < d; ++c) acc[c] = 0.0;
        for (int64_t q = lo; q < hi; q += EMMA_CHUNK) {
            int64_t qh = q + EMMA_CHUNK < hi ? q + EMMA_CHUNK : hi;
            for (int c = 0; c < d; ++c) chunk[c] = 0.0;
            for (int64_t m = q; m < qh; ++m) {
                const float *x = X + order[m] * (int64_t)d;
                for (int c = 0; c < d; ++c) chunk[c] = chunk[c] + (double)x[c];
            }
            for (int c = 0; c < d; ++c) acc[c] = acc[c] + chunk[c];
        }
        double cntd = (double)(hi - lo);
        for (int c = 0; c < d; ++c) C[(int64_t)j * d + c] = (float)(acc[c] / cntd);
    }
    free(pos); free(cnt); free(order); free(acc); free(chunk);
    return 0;
}

/* Lloyd iterations.  Iteration it: labels = assign(X, C); C = member means
 * (oracle_update_centers); cost = blocked FP64 sum (oracle_blocked_sum) of
 * each frame's FP32 distance to its center after the update, i.e. the
 * k-means cost C(S) = sum_i sum_{x in S_i} ||x - mu_i||^2 of the clusters
 * S_i just formed with mu_i their means (kmeans.py:67-75);
 * rel = |cost - prev| / cost (0 when cost == 0), prev starting at 0;
 * stop with converged = 1 when rel <= tol.  inertias[it] = cost.
 * Returns the number of iterations run, or -1 on failure. */
int oracle_lloyd(const float *X, int64_t n, int d, float *C, int k, int max_iter, double tol,
                 int32_t *labels, double *inertias, int *converged, int nthreads) {
    float *dist = (float *)malloc(sizeof(float) * (n > 0 ? n : 1));
    if (!dist) return -1;
    if (nthreads < 1) nthreads = 1;
    double prev = 0.0;
    int it = 0;
    *converged = 0;
    for (; it < max_iter; ++it) {
        oracle_assign(X, n, d, C, k, labels, NULL, nthreads);
        if (oracle_update_centers(X, n, d, labels, C, k, NULL) != 0) { free(dist); return -1; }
#pragma omp parallel for num_threads(nthreads) schedule(static)
        for (int64_t i = 0; i < n; ++i)
            dist[i] = emma_dist_seq(X + i * (int64_t)d, C + (int64_t)labels[i] * d, d);
        double cost = oracle_blocked_sum(dist, n);
        inertias[it] = cost;
        double rel = cost != 0.0 ? fabs(cost - prev) / cost : 0.0;
        prev = cost;
        if (rel <= tol) { *converged = 1; ++it; break; }
    }
    free(dist);
    return it;
}
