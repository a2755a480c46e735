/*
 * kmc.h -- C ABI of the MI355X (gfx950) matrix-centric k-means engine.
 *
 * Plain C: opaque handle, plain pointers and sizes, status codes, no HIP or
 * torch types. Host buffers are owned by the caller; device buffers are owned
 * by the handle and freed by kmc_destroy. The library never calls exit().
 *
 * Each entry point names the reference interface it replaces. Paths are
 * relative to the reference repository (HicrestLaboratory/Matrix-Centric-K-Means).
 */
#ifndef KMC_H
#define KMC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. 1..4 keep the meaning of the reference CLI exit codes
 * (src/include/utils.hpp:34-37: EXIT_ARGS, EXIT_CUDA_DEV, EXIT_POINT_IOB,
 * EXIT_INVALID_INFILE); the reference exits on every runtime error
 * (src/cuda_utils.cu:10-41), the library returns the codes from 5 on. */
typedef enum {
  KMC_OK = 0,
  KMC_ERR_ARG = 1,
  KMC_ERR_NO_DEVICE = 2,
  KMC_ERR_OUT_OF_BOUNDS = 3,
  KMC_ERR_INFILE = 4,
  KMC_ERR_HIP = 5,
  KMC_ERR_OOM = 6,
  KMC_ERR_STATE = 7,
  KMC_ERR_UNSUPPORTED = 8
} kmc_status;

/* Kernel functions: Kmeans::Kernel of src/kmeans.cuh:56-61 (linear,
 * polynomial (x.y+1)^2, sigmoid tanh(x.y+1); src/kernels/kernels.cuh:43-118).
 * KMC_KERNEL_RBF (exp(-gamma ||x-y||^2)) has no reference counterpart. */
typedef enum {
  KMC_KERNEL_LINEAR = 0,
  KMC_KERNEL_POLYNOMIAL = 1,
  KMC_KERNEL_SIGMOID = 2,
  KMC_KERNEL_RBF = 3
} kmc_kernel;

/* Arithmetic of the distance GEMM X C^T. Every form finishes with the FP32
 * margin-refinement pass, so labels equal those of the FP32 direct form
 * sum_j (x_j - c_j)^2 (comparisons/CPU_Kmeans/src/kmeans.cpp:121-131). */
typedef enum {
  KMC_PREC_FP32 = 0,   /* v_mfma_f32_32x32x2_f32                          */
  KMC_PREC_BF16X3 = 1, /* hi+lo split bf16, 3 x v_mfma_f32_32x32x16_bf16  */
  KMC_PREC_BF16 = 2,   /* single bf16 product, wide refinement band       */
  KMC_PREC_FP8 = 3     /* e4m3 operands with one power-of-two scale per
                          point and per centroid, applied by the matrix core
                          (v_mfma_scale_f32_32x32x64_f8f6f4); widest band.
                          Rows are padded to at least 64 dimensions. Not
                          available for batches. */
} kmc_precision;

/* Centroid update summation order. */
typedef enum {
  /* per cluster, float sums in point-index order: bit-identical to
   * comparisons/CPU_Kmeans/src/kmeans.cpp:166-180 */
  KMC_UPDATE_ORDERED = 0,
  /* per cluster, FP64 partial sums over fixed chunks of the sorted member
   * list, combined in chunk order and rounded once: deterministic, at least
   * as accurate, every CU busy; not bit-identical to the float sums */
  KMC_UPDATE_TREE = 1,
  /* one streaming pass over the points in index order (no sort): exact 64-bit
   * fixed-point sums (x scaled by a power of two chosen from max|x| and n;
   * exact for |x| >= 2^-14 relative to that scale), means rounded once;
   * deterministic. Falls back to KMC_UPDATE_TREE when the per-workgroup
   * accumulators (K x 16 x 8 bytes at least) exceed the LDS */
  KMC_UPDATE_FIXED = 2
} kmc_update_mode;

/* kmc_run flags */
#define KMC_RUN_CHECK_CONVERGED 0x1u /* stop at the first iteration >= 2 whose
                                        centroid shift test passes            */
#define KMC_RUN_TIME_KERNELS 0x2u    /* record HIP events around the distance
                                        kernel of every iteration             */
#define KMC_RUN_CONVERGE_SCORE 0x4u  /* with KMC_RUN_CHECK_CONVERGED: stop when
                                        |score_t - score_{t-1}| < tol instead
                                        (the reference GPU engine's test,
                                        src/kmeans.cu:1064-1073)              */
#define KMC_RUN_GRAPH 0x20u          /* capture the iterations between host
                                        convergence checks into a HIP graph
                                        (kept on the handle and replayed by
                                        later runs of the same shape)        */
#define KMC_RUN_EXCLUDE_EMPTY 0x8u   /* a cluster left empty by an assignment
                                        is never chosen again (the reference
                                        GPU engine gives it distance +inf,
                                        src/kernels/distances.cu:588-598); by
                                        default it becomes the zero vector and
                                        stays selectable, as in the reference
                                        CPU k-means                           */
/* The explicit-Lloyd rules of the reference's RAFT benchmark
 * (experiments/raft-bench/src/kmeans.cu:28-265):                            */
#define KMC_RUN_CONVERGE_SHIFT_SUM 0x40u /* with KMC_RUN_CHECK_CONVERGED: stop
                                        when sum_c ||c_new - c_old||^2 < tol,
                                        tested every iteration (kmeans.cu:
                                        168-212)                              */
#define KMC_RUN_COST_RATIO 0x80u     /* stop when cost_t / cost_{t-1} > 1 - tol
                                        from iteration 2, cost = FP64 inertia of
                                        the iteration's assignment (its
                                        inertia_check, kmeans.cu:186-204)    */
#define KMC_RUN_KEEP_EMPTY 0x100u    /* a cluster left empty keeps its
                                        centroid (RAFT's update_centroids)    */

typedef struct {
  size_t n;          /* points                                   */
  uint32_t d;        /* dimensions                               */
  uint32_t k;        /* clusters                                 */
  float tol;         /* centroid-shift tolerance (L2, per centroid) */
  int32_t precision; /* kmc_precision                            */
  int32_t update;    /* kmc_update_mode                          */
  int32_t device;    /* HIP device ordinal                       */
  int32_t reserved;
} kmc_config;

typedef struct {
  uint64_t iterations;   /* iterations run                                   */
  uint64_t converged_at; /* first iteration passing the shift test, or maxiter
                            (the return value of Kmeans::run,
                            src/kmeans.cuh:128-134)                          */
  double inertia;        /* FP64 sum_i ||x_i - c_{l_i}||^2 of the last assign */
  double score;          /* reference "Score" = inertia - sum_i ||x_i||^2
                            (src/kmeans.cu:1053-1061 omits ||x||^2)          */
  double max_shift;      /* FP64 max_c ||c_new - c_old|| of the last update  */
  uint64_t refined_two;  /* points refined on two candidates (last assign)   */
  uint64_t refined_full; /* points refined on all k centroids (last assign)  */
  double ms_total;       /* wall time of the iteration loop, ms              */
  double ms_assign;      /* sum of distance-kernel times (KMC_RUN_TIME_KERNELS) */
  uint64_t assign_launches; /* distance-kernel launches timed                */
  double ms_update;      /* sum of centroid-update kernel times (same flag)   */
  uint64_t update_launches; /* update-kernel launches timed                   */
  /* KMC_UPDATE_FIXED keeps exact cluster totals across the iterations of a
     run: the first iteration streams every point, the others update the
     totals from the moved points only (the batch engine's exact tree form:
     once at most an eighth of the points moved in the previous iteration): */
  uint64_t full_passes;  /* iterations whose update streamed every point      */
  uint64_t moved_points; /* points moved, summed over the other iterations    */
} kmc_result;

typedef struct kmc_handle kmc_handle;

/* Replaces the Kmeans constructor's set-up (src/kmeans.cu:71-305): device
 * selection, handle and stream creation. */
kmc_status kmc_create(const kmc_config* cfg, kmc_handle** out);

/* Replaces the point packing + H2D copy of src/kmeans.cu:94-99,125-126.
 * X is host row-major n x d with leading dimension ld >= d (floats). Also
 * builds the per-point norms and, for the bf16 forms, the hi/lo operands. */
kmc_status kmc_set_points(kmc_handle* h, const float* X, size_t ld);

/* Same, from a device pointer (no PCIe copy). */
kmc_status kmc_set_points_device(kmc_handle* h, const float* dX, size_t ld);

/* Explicit initial centroids, host k x d row-major (the reference CPU
 * k-means seeds with k distinct input rows, comparisons/CPU_Kmeans/src/
 * kmeans.cpp:84-119). */
kmc_status kmc_set_centroids(kmc_handle* h, const float* C0);

/* k-means++ seeding (D^2 sampling) of the K initial centroids from the
 * loaded points. Replaces init_centroids_plus_plus (src/kmeans.cu:447-769),
 * which the reference compiles out and which reads freed buffers when
 * enabled; this one is new: round 0 takes point floor(u0 * n), round r the
 * first point (in index order) whose running sum of FP64 squared distances
 * to the chosen set exceeds u_r times their total, with u_r the r-th
 * splitmix64(seed) uniform. Deterministic for a seed. chosen: host K point
 * indices (may be NULL). */
kmc_status kmc_init_plusplus(kmc_handle* h, uint64_t seed, int64_t* chosen);

/* Greedy k-means++: each round draws `trials` candidates by D^2 sampling
 * (round r >= 1, trial t uses uniform 1 + (r-1)*trials + t of the splitmix64
 * sequence) and keeps the one whose potential sum_i min(D2_i, ||x_i - c||^2)
 * (FP64, blocked fixed order) is smallest, the first on ties. This is the
 * scheme of the reference's init_centroids_plus_plus (src/kmeans.cu:463-769,
 * s = 2 + ceil(ln k) trials, compiled out there). trials = 0 selects that
 * s; trials = 1 is kmc_init_plusplus; at most 16. */
kmc_status kmc_init_plusplus_trials(kmc_handle* h, uint64_t seed, uint32_t trials,
                                    int64_t* chosen);

/* Initial labels (host, n values in [0,k)); the initial centroids become the
 * label means. Replaces init_centroids_rand (src/kmeans.cu:365-444), which
 * uses labels i % k. */
kmc_status kmc_set_labels(kmc_handle* h, const int32_t* labels0);

/* Replaces Kmeans::run (src/kmeans.cu:772-1122): up to maxiter Lloyd
 * iterations (assign -> update -> shift test). flags: KMC_RUN_*. */
kmc_status kmc_run(kmc_handle* h, uint64_t maxiter, uint32_t flags,
                   kmc_result* res);

/* Replaces the label write-back (src/kmeans.cu:1084-1103). */
kmc_status kmc_get_labels(kmc_handle* h, int32_t* labels);

/* Current centroids, host k x d. */
kmc_status kmc_get_centroids(kmc_handle* h, float* C);

/* Stage hook: one assignment of every point against host centroids C (k x d):
 * the fused distance GEMM + top-3 + margin band, then the refinement pass.
 * Replaces compute_distances + argmin (src/kmeans.cu:846-966;
 * comparisons/CPU_Kmeans/src/kmeans.cpp:121-154). labels: host n. */
kmc_status kmc_assign(kmc_handle* h, const float* C, int32_t* labels,
                      kmc_result* res);

/* Stage hook: centroid update from host labels and the centroids those labels
 * were assigned against (C_old, for the inertia). Writes C_new (host k x d)
 * and counts (host k). Replaces compute_centroids_spmm
 * (src/kernels/centroids.cu:177-223) and the CPU compute_centroids
 * (comparisons/CPU_Kmeans/src/kmeans.cpp:166-180). */
kmc_status kmc_update(kmc_handle* h, const int32_t* labels, const float* C_old,
                      float* C_new, uint32_t* counts, kmc_result* res);

/* Device times of the handle's set-up steps, ms (HIP events): the H2D copy of
 * the points (0 for kmc_set_points_device), the derived point operands
 * (norms, bf16/fp8 planes; the reference's B = -2 X X^T build, "b-mult"),
 * and the last initial-centroid step (kmc_set_centroids, kmc_set_labels or
 * k-means++). Replace the memcpy, b-mult and init_centroids timers of
 * src/kmeans.cu:117-136, 148-205, 267-297. */
typedef struct {
  double memcpy_ms;
  double operands_ms;
  double init_ms;
} kmc_setup_times;

kmc_status kmc_get_setup_times(const kmc_handle* h, kmc_setup_times* out);

/* Per-iteration stage times, ms, of the last kmc_run with
 * KMC_RUN_TIME_KERNELS (its first 4096 iterations): the distance stage (fused
 * distance GEMM + top-4), the argmin stage (exact refinement of the margin
 * band) and the centroid stage (sort/update/finish). Replace the
 * compute_distances, clusters_argmin_shfl and compute_centroids timers of
 * src/kmeans.cu:837-904, 938-958, 968-1044. Writes min(cap, count) entries;
 * *count (may be NULL) receives the number of timed iterations. */
typedef struct {
  float distances_ms;
  float argmin_ms;
  float centroids_ms;
} kmc_stage_times;

kmc_status kmc_get_stage_times(kmc_handle* h, kmc_stage_times* out, uint64_t cap,
                               uint64_t* count);

/* Allocation-free assignment. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc). The same stage as
 * kmc_assign in the fp32 form (fused distance GEMM + top-4 + margin band,
 * then the exact refinement), with the same labels, without a handle and
 * without any allocation: everything besides the inputs and the output
 * lives in a fixed device array of the library (under 20 MiB, one per GPU,
 * held until the process ends).
 *   X       device, n rows of ld floats; C device, k rows of ld floats;
 *           ld = d rounded up to 16, 32, 64, 128 or 256 (KMC_ERR_ARG
 *           otherwise) and the columns d..ld-1 of both must be zero
 *   labels  device, n int32 (output)
 *   stream  a hipStream_t, or NULL for the null stream; the call only
 *           enqueues work (synchronise the stream before reading labels)
 * Limits: n <= 524288 and k <= 1024 (KMC_ERR_UNSUPPORTED beyond). Uses the
 * current HIP device. Calls that overlap on one GPU share the array and race;
 * calls on one stream are ordered and safe. */
kmc_status kmc_assign_static(const float* X, size_t n, uint32_t d, size_t ld, const float* C,
                             uint32_t k, int32_t* labels, void* stream);

/* Diagnostic, no counterpart in the reference: one 32 x 32 block of the
 * distance GEMM's matrix-core products, out[i*32 + j] = A_i . B_j, for host
 * fp32 operands A, B (32 x d each) stored as the form stores them (fp32 as
 * is; bf16 round to nearest even; fp8 e4m3 round to nearest even after one
 * power-of-two scale per row) and multiplied by the distance kernel's MFMA
 * chain (same instructions, k-step order and lane layout) from a zero
 * accumulator. Aq, Bq (32 x d) receive the stored operand values. It lets a
 * caller check the error model behind the margin band (operand rounding and
 * accumulation bounds per form). precision: KMC_PREC_FP32, _BF16 or _FP8;
 * d in {16, 32, 64, 128, 256} (fp8: d >= 64). */
kmc_status kmc_probe_mfma(int32_t precision, uint32_t d, const float* A, const float* B,
                          float* out, float* Aq, float* Bq);

/* Launch-variant autotuner for the fused distance kernel (no counterpart in
 * the reference, whose stages are fixed library calls): times every variant
 * (waves per workgroup, point sets per wave, register prefetch, ranking
 * lists per lane and values kept per list) on the handle's points and
 * current centroids -- the distance
 * kernel plus the refinement of the points its band leaves -- and keeps the
 * fastest for later kmc_run/kmc_assign. Variants whose registers spill, or
 * that need every centroid in LDS where they do not fit, are not offered
 * (ms = -1). */
#define KMC_MAX_VARIANTS 16
typedef struct {
  int32_t nvariants;
  int32_t chosen;
  float ms[KMC_MAX_VARIANTS];     /* assignment time per variant (best of two rounds' means
                                     of three), -1 if not offered */
  int32_t threads[KMC_MAX_VARIANTS];
  int32_t tile[KMC_MAX_VARIANTS]; /* points per workgroup tile */
} kmc_tuning;

kmc_status kmc_autotune(kmc_handle* h, kmc_tuning* out);

/* Select a launch variant of the fused distance kernel directly. */
kmc_status kmc_set_assign_variant(kmc_handle* h, int variant);

/* Points whose nearest MFMA candidates all fall inside the margin band are
 * refined against every centroid in the reference's float arithmetic. Where
 * kmc_set_short_prune's single-sweep form applies (its shapes, on by
 * default) it takes every such queue; otherwise, from
 * min_points such points per assignment on (default 32768), that pass first
 * narrows each point's centroids with a bf16x3 MFMA sweep and evaluates only
 * the survivors; below it every centroid is evaluated. The pruned pass is
 * launched in a run only when the handle's previous run ended with an all-k
 * queue of at least min_points / 4 points (or none ran yet). 1 selects the
 * pruned pass always, 0 never. Labels are the same either way. A tuning
 * knob of this library's refinement, with no counterpart in the reference. */
kmc_status kmc_set_prune_min(kmc_handle* h, int32_t min_points);

/* With K <= 512 at D <= 64, or K <= 256 at 64 < D <= 128, the all-k queue
 * takes a single-sweep form of the same MFMA pruning, run in one launch beside
 * the two/three-candidate refinement on a share of the grid sized by the two
 * queues (default on); 0 returns these shapes to kmc_set_prune_min's choice
 * between the two-sweep pruned pass and the exact pass over every centroid.
 * Labels are the same either way. A tuning knob with no counterpart in the
 * reference. */
kmc_status kmc_set_short_prune(kmc_handle* h, int32_t enable);

/* Release every device buffer (src/kmeans.cu:307-359). */
kmc_status kmc_destroy(kmc_handle* h);

/* Message of the last error on this handle (never NULL). */
const char* kmc_last_error(const kmc_handle* h);

/* Static description of a status code. */
const char* kmc_status_string(kmc_status s);

/* Number of visible HIP devices (0 when none); replaces the device query of
 * src/main.cu:40-52. */
int kmc_device_count(void);

/* Device memory in use (total - free) on a device; replaces the
 * MEMORY FOOTPRINT report of src/kmeans.cu:1094-1099. */
kmc_status kmc_memory_footprint(int device, size_t* used_bytes);

/* ------------------------------------------------------------------------
 * Kernel k-means (the reference's Popcorn path, src/kmeans.cu:71-305,772-1122)
 * ------------------------------------------------------------------------ */

/* Storage type of the n x n kernel matrix. */
typedef enum { KMC_STORE_FP32 = 0, KMC_STORE_BF16 = 1 } kmc_kstore;

/* Arithmetic of the Gram product X X^T that the kernel matrix is built from
 * (the reference's cublasSgemm / cublasSsyrk, src/kernels/kernels.cuh:
 * 567-649); every form computes only the tiles on and below the diagonal and
 * mirrors them, as the reference's SYRK path does for n/d <= 100. */
typedef enum {
  KMC_KBUILD_FP32 = 0,   /* v_mfma_f32_32x32x2_f32                         */
  KMC_KBUILD_BF16X3 = 1, /* hi+lo split bf16, 3 x v_mfma_f32_32x32x16_bf16 */
  KMC_KBUILD_BF16 = 2    /* one bf16 product                               */
} kmc_kbuild;

#define KMC_RUN_STOP_STABLE 0x10u /* with KMC_RUN_CHECK_CONVERGED (kernel
                                     k-means): stop when no label changes,
                                     instead of the score test              */

/* How S = K V^T (the kernel-space inner products) is formed each iteration. */
typedef enum {
  KMC_KK_AUTO = 0,    /* cluster-sorted (measured faster at every k)      */
  KMC_KK_DENSE = 1,   /* one-hot MFMA product, k <= 64                    */
  KMC_KK_SORTED = 2   /* cluster-sorted row sums (the reference's REORDER
                         level, src/kmeans.cu:1125-1261), k <= 1024      */
} kmc_kk_product;

typedef struct {
  size_t n;
  uint32_t d;
  uint32_t k;        /* clusters: at most 64 dense, 1024 sorted       */
  float tol;         /* score tolerance (src/kmeans.cu:1068-1073)    */
  int32_t kernel;    /* kmc_kernel                                   */
  int32_t store;     /* kmc_kstore                                   */
  float gamma;       /* RBF width; <= 0 selects 1/d                  */
  int32_t device;
  int32_t product;   /* kmc_kk_product                               */
  int32_t build;     /* kmc_kbuild (0: fp32)                          */
} kmc_kk_config;

typedef struct {
  uint64_t iterations;
  uint64_t converged_at;
  double objective;  /* FP64 sum_i ||phi(x_i) - mu_{l_i}||^2 of the last assign */
  double score;      /* reference "Score": objective - sum_i K_ii       */
  uint64_t changed;  /* labels changed by the last assignment          */
  double ms_total;   /* iteration loop, ms                             */
  double ms_build;   /* kernel-matrix build at kmc_kk_set_points, ms   */
  double ms_spmm;    /* sum of K-streaming kernel times (KMC_RUN_TIME_KERNELS) */
  uint64_t spmm_launches;
} kmc_kk_result;

typedef struct kmc_kk_handle kmc_kk_handle;

/* Replaces the Kmeans constructor's kernel-matrix path (src/kmeans.cu:71-305). */
kmc_status kmc_kk_create(const kmc_kk_config* cfg, kmc_kk_handle** out);

/* Copies X (host n x d, leading dimension ld) and builds K = f(X X^T):
 * replaces init_kernel_mtx / _gemm / _syrk + Kernel::function
 * (src/kernels/kernels.cuh:43-118,567-718; src/kmeans.cu:158-171). */
kmc_status kmc_kk_set_points(kmc_kk_handle* h, const float* X, size_t ld);

/* Initial labels (host n); replaces init_centroids_rand (src/kmeans.cu:365-444). */
kmc_status kmc_kk_set_labels(kmc_kk_handle* h, const int32_t* labels0);

/* Replaces Kmeans::run on the kernel path (src/kmeans.cu:772-1122).
 * flags: KMC_RUN_CHECK_CONVERGED (score test), KMC_RUN_STOP_STABLE,
 * KMC_RUN_TIME_KERNELS. */
kmc_status kmc_kk_run(kmc_kk_handle* h, uint64_t maxiter, uint32_t flags, kmc_kk_result* res);

kmc_status kmc_kk_get_labels(kmc_kk_handle* h, int32_t* labels);

/* Set-up times of a kernel k-means handle (memcpy: the H2D copy of X;
 * operands: the kernel-matrix build, the reference's b-mult; init: the last
 * kmc_kk_set_labels), and per-iteration stage times of the last kmc_kk_run
 * with KMC_RUN_TIME_KERNELS (distances: the K product; argmin: assignment and
 * statistics; centroids: the rebuild of V, i.e. the sort by label). Same
 * reference timers as kmc_get_setup_times / kmc_get_stage_times. */
kmc_status kmc_kk_get_setup_times(const kmc_kk_handle* h, kmc_setup_times* out);
kmc_status kmc_kk_get_stage_times(kmc_kk_handle* h, kmc_stage_times* out, uint64_t cap,
                                  uint64_t* count);

/* Stage hook: kernel-space distances for host labels, host n x k, as the
 * reference materialises them (no K_ii term): replaces
 * compute_distances_popcorn_spmv (src/kernels/distances.cu:895-1002). */
kmc_status kmc_kk_distances(kmc_kk_handle* h, const int32_t* labels, float* D);

/* Rows [row0, row0+nrows) of K as stored (bf16 widened to float), host
 * nrows x n; replaces the LOG_KERNEL dump of src/kmeans.cu:173-194. */
kmc_status kmc_kk_get_kernel_rows(kmc_kk_handle* h, size_t row0, size_t nrows, float* out);

const char* kmc_kk_last_error(const kmc_kk_handle* h);

kmc_status kmc_kk_destroy(kmc_kk_handle* h);

/* ---- Batches of independent Lloyd problems -------------------------------
 * Replaces the reference's `-r` multi-run loop (src/main.cu:89-105: one
 * Kmeans object and one run per repetition) for many independent
 * clusterings of uneven n, d and k: every stage of an iteration is one launch
 * for the whole batch, the distance stage a persistent kernel over a packed
 * (problem, tile) queue. Each problem follows the reference CPU k-means
 * (comparisons/CPU_Kmeans/src/kmeans.cpp:86-200): exact first-min labels,
 * its own shift test from iteration 2 (a converged problem is frozen while
 * the others go on); centroids are exact fixed-point sums rounded once, as
 * KMC_UPDATE_FIXED with the problem's own scale (a full pass over the sorted
 * members, then the moved points only while at most an eighth move). Limits: d <= 256, k <= 256, and the centroid operands of
 * the largest problem must fit the LDS (k x d <= 256 x 128 in fp32 or
 * bf16x3). */
typedef struct kmc_problem {
  uint64_t n;        /* points */
  uint32_t d, k;     /* dimensions, clusters */
  const float* X;    /* host, n x d row-major, leading dimension ld */
  size_t ld;
  const float* C0;   /* host, k x d initial centroids */
} kmc_problem;

typedef struct kmc_batch_config {
  float tol;          /* shift tolerance (KMC_RUN_CHECK_CONVERGED) */
  int32_t precision;  /* kmc_precision, the same for every problem */
  int32_t device;
  int32_t reserved;
} kmc_batch_config;

typedef struct kmc_batch_result {
  uint64_t iterations;   /* iterations run by this problem */
  uint64_t converged_at; /* iteration whose shift test passed, or maxiter (the
                            value Kmeans::run returns, src/kmeans.cu:1122) */
  double inertia;        /* FP64 sum of ||x - c_label||^2, final state */
} kmc_batch_result;

typedef struct kmc_batch kmc_batch;

/* Packs the problems into one device layout and uploads them. */
kmc_status kmc_batch_create(const kmc_problem* probs, uint32_t count,
                            const kmc_batch_config* cfg, kmc_batch** out);
/* Runs every problem from its initial centroids for at most maxiter
 * iterations. flags: KMC_RUN_CHECK_CONVERGED. res: count entries (may be
 * NULL). ms_total: device time of the run (may be NULL). */
kmc_status kmc_batch_run(kmc_batch* b, uint64_t maxiter, uint32_t flags,
                         kmc_batch_result* res, double* ms_total);
kmc_status kmc_batch_get_labels(kmc_batch* b, uint32_t problem, int32_t* labels);
kmc_status kmc_batch_get_centroids(kmc_batch* b, uint32_t problem, float* C);
const char* kmc_batch_last_error(const kmc_batch* b);
kmc_status kmc_batch_destroy(kmc_batch* b);

#ifdef __cplusplus
}
#endif

#endif /* KMC_H */
