/*
 * emma_amd.h — C ABI of the MI355X discretization library (libemma_amd.so).
 *
 * The library computes, on one AMD Instinct MI355X (gfx950):
 *   - TICA's lagged mean-free moments (mean, C00, C0t, optionally Ctt),
 *   - assignment of frames to cluster centers (argmin of squared distance),
 *   - kmeans++ seeding and Lloyd iterations,
 *   - the TICA projection (X - mean) V.
 *
 * Conventions
 *   - Every pointer argument is a DEVICE pointer owned by the caller unless
 *     the comment says otherwise.  Frame matrices are row-major, C-contiguous
 *     float32 (the layout of np.require(X, float32, 'C'),
 *     clustering/interface.py:159).  Labels are int32, moments FP64.
 *   - `stream` is a hipStream_t (NULL = default stream).  Every call only
 *     enqueues work on that stream; nothing synchronizes the host.
 *   - Scratch memory is passed in as `workspace`; its size comes from the
 *     matching *_workspace_size() call.  The library itself allocates no
 *     device memory.
 *   - Return value: EMMA_OK (0) or a negative EMMA_ERR_* code; the message of
 *     the last failure on the calling thread is emma_last_error().
 *   - Results are deterministic: the same inputs give bit-identical outputs
 *     on every run.
 *
 * Reference interfaces each entry point replaces are cited per function
 * (paths relative to the PyEMMA source tree).
 */
#ifndef EMMA_AMD_H
#define EMMA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define EMMA_OK 0
#define EMMA_ERR_INVALID_ARG (-1)
#define EMMA_ERR_HIP (-2)
#define EMMA_ERR_WORKSPACE (-3)
#define EMMA_ERR_UNSUPPORTED (-4)

/* moment flags */
#define EMMA_MOM_SYMMETRIZE 1u   /* reversible: sx=sy=sum X+sum Y, Cxx=X'X+Y'Y, Cxy=X'Y+Y'X */
#define EMMA_MOM_REMOVE_MEAN 2u  /* mean-free (co)moments */
#define EMMA_MOM_C0T 4u          /* compute Mxy */
#define EMMA_MOM_CTT 8u          /* compute Myy = Y'Y (not with SYMMETRIZE) */
#define EMMA_MOM_WEIGHTED 16u    /* frame weights given (emma_moments_lagged_weighted) */

/* arithmetic of the covariance tiles */
#define EMMA_PREC_FP32 0    /* FP32 MFMA (v_mfma_f32_32x32x2_f32), FP32 tile accumulation */
#define EMMA_PREC_BF16X4 1  /* operands split into bf16 hi+lo, 4 bf16 MFMA products, FP32 tile accumulation */
#define EMMA_PREC_FP64 2    /* FP64 operands on v_mfma_f64_16x16x4_f64, FP64 accumulation */
#define EMMA_PREC_BF16X4_ACC 3 /* bf16 hi+lo, 4 products, FP32 accumulation restarted every 64 pairs and summed in FP32
                                  over 4096 pairs (every kind of moments) */

/* assignment modes */
#define EMMA_ASSIGN_FP32_PLAIN 0    /* sequential FP32 loop per (frame, center) on the vector ALUs */
#define EMMA_ASSIGN_BF16X3_RECHECK 1 /* bf16x3 MFMA distances + near-tie recheck with the FP32 loop */

/* Library version string (host pointer, static storage). */
const char *emma_version(void);

/* Message of the last failed call on this thread (host pointer). */
const char *emma_last_error(void);

/* ------------------------------------------------------------------ */
/* Lagged moments                                                      */
/* ------------------------------------------------------------------ */

/* Workspace bytes needed by emma_moments_lagged for these arguments. */
size_t emma_moments_workspace_size(int64_t n_rows, int32_t d, int32_t ntraj, uint32_t flags,
                                   int32_t precision);

/*
 * Unnormalized lagged moments of many trajectories in one call.
 *
 * Replaces the chunk loop of LaggedCovariance._estimate
 * (coordinates/estimation/covariance.py:219-244) over RunningCovar.add
 * (_ext/variational/estimators/running_moments.py:233-308) with
 * moments_XXXY / moments_block / moments_XX (moments.py:601-973) and the
 * Chan merge Moments.combine (running_moments.py:35-61).
 *
 * X            frames of all trajectories, rows concatenated: (n_rows, d) float32
 * traj_offsets device int64[ntraj+1], trajectory t is rows [off[t], off[t+1])
 * lag, stride, skip  pairing of iterable.py:95-160: pair i of a trajectory is
 *              (row skip+i*stride, row skip+lag+i*stride); lag == 0 means
 *              instantaneous moments of rows skip, skip+stride, ... (C00 only)
 * flags        EMMA_MOM_*
 * precision    EMMA_PREC_*
 * outputs (FP64, device): w[1]; sx[d]; sy[d] (may be NULL); Mxx[d*d];
 *              Mxy[d*d] (needs EMMA_MOM_C0T); Myy[d*d] (needs EMMA_MOM_CTT).
 *              The values follow the reference: with SYMMETRIZE w = 2P, sx = sy
 *              = sum of X and Y rows; otherwise w = P, sx = sum X, sy = sum Y.
 *              With REMOVE_MEAN the matrices are centered at sx/w (and sy/w).
 * P == 0 (no trajectory longer than skip+lag) writes w = 0 and zeros.
 */
int emma_moments_lagged(const float *X, int64_t n_rows, int32_t d, const int64_t *traj_offsets,
                        int32_t ntraj, int32_t lag, int32_t stride, int32_t skip, uint32_t flags,
                        int32_t precision, double *w, double *sx, double *sy, double *Mxx,
                        double *Mxy, double *Myy, void *workspace, size_t workspace_bytes,
                        void *stream);

/*
 * The same moments with frame weights: row_weights (device FP64, one per row
 * of X) weigh every term of a pair by the weight of its x frame, and w becomes
 * the weight sum (times 2 with SYMMETRIZE), as the reference's weighted
 * estimator does (moments.py:267-303, 540-553; weights iterated with the data,
 * covariance.py:198-240).  flags must include EMMA_MOM_WEIGHTED; the workspace
 * size is queried with the same flags.
 */
int emma_moments_lagged_weighted(const float *X, int64_t n_rows, int32_t d, const int64_t *traj_offsets,
                                 int32_t ntraj, int32_t lag, int32_t stride, int32_t skip, uint32_t flags,
                                 int32_t precision, const double *row_weights, double *w, double *sx,
                                 double *sy, double *Mxx, double *Mxy, double *Myy, void *workspace,
                                 size_t workspace_bytes, void *stream);

/* ------------------------------------------------------------------ */
/* Assignment                                                          */
/* ------------------------------------------------------------------ */

/* Workspace bytes needed by emma_assign for these arguments. */
size_t emma_assign_workspace_size(int64_t n, int32_t d, int32_t k, int32_t mode);

/*
 * labels[i] = argmin_j sum_c (X[i,c] - C[j,c])^2, the sum taken in FP32 in
 * index order without fused multiply-add, first minimum wins.  As in the
 * reference loop (best = distance to center 0, replaced only by a smaller
 * one): a NaN distance never wins, a NaN distance to center 0 keeps label 0,
 * and a row whose distances are all +inf gets label 0.
 *
 * Replaces ClusterModel(cluster_centers, metric='euclidean').transform(X)
 * as called from AbstractClustering._transform_array
 * (coordinates/clustering/interface.py:157-167; API assign_to_centers,
 * coordinates/api.py:2060-2156).
 *
 * mode EMMA_ASSIGN_BF16X3_RECHECK computes candidate distances with bf16 hi/lo
 * MFMA products and re-evaluates, with the FP32 loop, every frame whose two
 * best candidates lie within a rigorous error bound; labels are identical to
 * EMMA_ASSIGN_FP32_PLAIN.  n_rechecked (device int64, may be NULL) receives the
 * number of re-evaluated frames.
 */
int emma_assign(const float *X, int64_t n, int32_t d, const float *C, int32_t k, int32_t *labels,
                int32_t mode, int64_t *n_rechecked, void *workspace, size_t workspace_bytes,
                void *stream);

/*
 * emma_assign without a workspace argument and without any allocation: no
 * counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same labels and n_rechecked as emma_assign.  Its
 * scratch (center images, recheck list) is a fixed device array of
 * EMMA_ASSIGN_NOALLOC_SCRATCH_BYTES that every GPU the library runs on holds
 * until the process ends; frames are processed in chunks that fit it.
 * Limit: the k centers' tile images, ceil(k/32) * (192 * (d_pad + 8) + 768)
 * bytes (bf16 and fp16 images) with d_pad = d rounded up to the kernel's
 * K-step width (32, 64, 128, 192, 256, 304, 384 or 512), must leave room for
 * EMMA_ASSIGN_NOALLOC_MIN_CHUNK frames (64 bytes each): k <= 9,300 at
 * d = 512, k <= 15,400 at d = 300; larger sets return EMMA_ERR_INVALID_ARG
 * with a message.  Every
 * call on one GPU shares that GPU's array: calls that overlap on one GPU
 * (different streams, threads) race and must be serialized by the caller.
 */
#define EMMA_ASSIGN_NOALLOC_SCRATCH_BYTES (32u << 20)
#define EMMA_ASSIGN_NOALLOC_MIN_CHUNK 65536
int emma_assign_noalloc(const float *X, int64_t n, int32_t d, const float *C, int32_t k, int32_t *labels,
                        int32_t mode, int64_t *n_rechecked, void *stream);

/*
 * Variants of the MFMA assignment, for the autotuner: the pipeline of its
 * first pass (a coarse fp16 pass over every frame whose undecided frames go
 * to the fine bf16 x 3 pass, or the fine pass over every frame) and the tile
 * shape (waves per workgroup, workgroups per CU, frame blocks per wave).
 * emma_assign_variants(d) is the
 * number of variants at dimension d (0 if d is beyond the MFMA path);
 * emma_assign_select(d, k, v) makes variant v the one used for every later
 * call whose dimension and center count fall in the same bucket as (d, k)
 * (process-wide, thread-safe; v = -1 restores the default).  Variants only
 * change speed: labels are identical for every variant.
 * No reference counterpart: the reference has no tunable kernels.
 */
int32_t emma_assign_variants(int32_t d);
int emma_assign_select(int32_t d, int32_t k, int32_t variant);

/* ------------------------------------------------------------------ */
/* k-means                                                             */
/* ------------------------------------------------------------------ */

size_t emma_kmeanspp_workspace_size(int64_t n, int32_t d, int32_t k);

/*
 * kmeans++ seeding: C_out (k,d) float32 rows of X, idx_out (k,) int64 row
 * indices.  Round 0 takes row floor(u_0 n); round r takes a row with
 * probability proportional to its squared FP32 distance to the nearest
 * chosen center, u_r = SplitMix64 draw r of `seed`, with the FP64 summation
 * order documented in oracle/emma_oracle.c (oracle_kmpp_pick).
 *
 * Replaces deeptime's kmeans++ initialization invoked by
 * KmeansClustering._estimate (coordinates/clustering/kmeans.py:211-212,
 * 254-258; entry observed as init_centers_kmpp, clustering/tests/
 * test_kmeans.py:304).
 */
int emma_kmeanspp(const float *X, int64_t n, int32_t d, int32_t k, uint64_t seed, float *C_out,
                  int64_t *idx_out, void *workspace, size_t workspace_bytes, void *stream);

size_t emma_lloyd_workspace_size(int64_t n, int32_t d, int32_t k, int32_t assign_mode);

/*
 * Lloyd iterations on centers C (k,d) float32, updated in place.  Each
 * iteration: labels = emma_assign(X, C); C_j = mean of its members (FP64,
 * documented order; empty centers unchanged); cost = blocked FP64 sum of
 * every frame's FP32 distance to the updated center of its label (the
 * k-means cost of the new clusters against their means, kmeans.py:67-75);
 * stop when |cost - prev| / cost <= tol (prev starts at 0).
 * Device outputs: labels (n,) of the last assignment, inertias[max_iter]
 * (cost per iteration run), n_iter[1] (int32), converged[1] (int32).
 *
 * Replaces deeptime KMeans.fit's Lloyd loop as called at
 * coordinates/clustering/kmeans.py:254-261 (tolerance rule kmeans.py:67-75).
 */
int emma_lloyd(const float *X, int64_t n, int32_t d, float *C, int32_t k, int32_t max_iter,
               double tol, int32_t assign_mode, int32_t *labels, double *inertias,
               int32_t *n_iter, int32_t *converged, void *workspace, size_t workspace_bytes,
               void *stream);

/* ------------------------------------------------------------------ */
/* TICA projection and input checks                                    */
/* ------------------------------------------------------------------ */

/*
 * Y = (X - mean) V[:, :m], computed in FP64 and rounded to float32.
 * mean (d,) FP64, V (d, ldv) FP64 row-major (columns 0..m-1 used), Y (n, m);
 * d <= 2^22.
 *
 * Replaces TICABase._transform_array
 * (coordinates/transform/_tica_base.py:117-133).
 */
int emma_project(const float *X, int64_t n, int32_t d, const double *mean, const double *V,
                 int32_t ldv, int32_t m, float *Y, void *stream);

/*
 * out[i] = sum_c X[i,c] u[c] + u0 in FP64 (out device FP64, n entries): the
 * frame weights of a linear weight model, _KoopmanWeights.weights
 * (coordinates/estimation/koopman.py:54-60).
 */
int emma_row_affine(const float *X, int64_t n, int32_t d, const double *u, double u0, double *out,
                    void *stream);

/*
 * count[0] = number of non-finite entries of X (n elements, float32).
 * Replaces the per-chunk np.isfinite check of DataSourceIterator.__next__
 * (coordinates/data/_base/datasource.py:1067-1075).  count is device int64.
 */
int emma_count_nonfinite(const float *X, int64_t n, int64_t *count, void *stream);

/*
 * Transition counts of discrete trajectories: C[i * nstates + j] (device
 * int64, nstates^2 entries, overwritten) = number of t with dtraj[t] = i and
 * dtraj[t + lag] = j within one trajectory, for every t (sliding != 0) or for
 * t = 0, lag, 2 lag, ... (sliding == 0); hist (device int64, nstates entries,
 * may be NULL) = frames per state.  Replaces deeptime's
 * count_matrix(dtrajs, lag, sliding, nstates) called at
 * msm/estimators/_dtraj_stats.py:138-145 (and count_states, :60) on the
 * int32 labels that assignment produces (coordinates/clustering/interface.py:
 * 157-231).  Frames with state -1 (unassigned) are not counted; the number of
 * entries below -1 or at/above nstates goes to n_bad (device int64) and those
 * entries are not counted.  traj_offsets: ntraj + 1 device int64 CSR offsets
 * into dtraj (traj_offsets[ntraj] == n).
 */
int emma_count_transitions(const int32_t *dtraj, int64_t n, const int64_t *traj_offsets, int32_t ntraj,
                           int32_t lag, int32_t sliding, int32_t nstates, int64_t *C, int64_t *hist,
                           int64_t *n_bad, void *stream);

/*
 * out[0] (device int32) = largest entry of dtraj, -1 if n == 0: the state
 * count of deeptime's number_of_states (msm/estimators/_dtraj_stats.py:64).
 */
int emma_max_state(const int32_t *dtraj, int64_t n, int32_t *out, void *stream);

/*
 * D = A B + C by exactly one v_mfma_f32_32x32x16_bf16 instruction (one
 * wave): A (32 x 16) and B (16 x 32) row-major bf16 bit patterns, C and D
 * (32 x 32) row-major float32, all device memory.  No counterpart in the
 * reference: it exposes the rounding of the matrix core's FP32
 * accumulation, the property the error bound of EMMA_ASSIGN_BF16X3_RECHECK
 * rests on (round to nearest, unit 2^-24), so that a test can check it on the
 * device the library runs on.
 */
int emma_probe_mfma_bf16(const uint16_t *A, const uint16_t *B, const float *C, float *D, void *stream);

#define EMMA_DIST_INVERSE 1u /* 1 / distance (InverseDistanceFeature) */

/*
 * out[f * n_pairs + p] (device float32) = distance between atoms pairs[2p]
 * and pairs[2p + 1] of frame f of xyz (device float32, n_frames x n_atoms x
 * 3, row-major): sqrt((dx*dx + dy*dy) + dz*dz) with d = x_j - x_i, each step
 * a correctly rounded FP32 operation, or its reciprocal with
 * EMMA_DIST_INVERSE.  Replaces DistanceFeature.transform /
 * InverseDistanceFeature.transform (coordinates/data/featurization/
 * distances.py:53-54, 73-74: mdtraj.compute_distances without a unit cell)
 * behind Featurizer.add_distances / add_inverse_distances
 * (featurizer.py:305-342).  Pairs (device int32, n_pairs x 2) with an index
 * outside [0, n_atoms) give NaN.  n_atoms <= 2730.
 */
int emma_pair_distances(const float *xyz, int64_t n_frames, int32_t n_atoms, const int32_t *pairs, int32_t n_pairs,
                        uint32_t flags, float *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* EMMA_AMD_H */
