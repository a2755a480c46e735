// kmc_kernels.h -- parameter blocks and launchers of the Lloyd kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

namespace kmc {

enum { PREC_FP32 = 0, PREC_BF16X3 = 1, PREC_BF16 = 2, PREC_FP8 = 3 };
// fp8 centroid operand row: DP e4m3 bytes, then the row's E8M0 scale as an
// int at byte DP, padded to DP + kFp8RowPad bytes
constexpr int kFp8RowPad = 16;
enum { UPDATE_ORDERED = 0, UPDATE_TREE = 1, UPDATE_FIXED = 2 };
constexpr int kTreeRows = 256;  // sorted rows per tree-update chunk
constexpr int kMaxRegions = 4096;  // assign workgroups (refinement queue regions)
constexpr int kRefineBlocks = 1024;  // refinement grid
constexpr int kMaxDP = 256;     // largest padded dimension of the Lloyd path
// ||c||^2 of padding rows and excluded clusters: finite (so that index bits
// can be packed into it) and above any real distance
constexpr float kFar = 3.0e38f;
constexpr float kFarTest = 1.0e38f;

// Device-resident loop state (one per handle).
struct DevState {
  unsigned long long iter;          // iterations completed
  unsigned long long maxiter;       // stop after this many
  unsigned long long converged_at;  // first iteration passing the shift test
  int done;                         // set: every kernel returns at once
  int pad;
  double inertia;                   // FP64 inertia of the last assignment
  double max_shift2;                // FP64 max_c ||c_new - c_old||^2
  unsigned long long refined_two;
  unsigned long long refined_all;
  double sum_x2;                    // FP64 sum_i ||x_i||^2 (set-up)
  double last_score;                // score of the previous iteration
  double sum_shift2;                // FP64 sum_c ||c_new - c_old||^2
  double prior_cost;                // FP64 inertia of the previous assignment
  // fixed-point update, incremental form: points whose label changed in the
  // last iteration, the count being formed by this one, and whether the
  // cluster totals match the labels of the last iteration (0 at run start)
  unsigned long long moved;
  unsigned long long moved_acc;
  int sums_valid;
  int pad2;
  unsigned long long full_passes;  // iterations of this run with a full pass
  unsigned long long moved_total;  // points moved in the other iterations
};

// Exponent of the fixed-point scale 2^e of x -> round(x * 2^e): the largest
// that keeps n * max|x| * 2^e below 2^61, so no cluster sum can overflow.
inline int fx_scale_exp(size_t n, float amax) {
  const double nn = (double)(n > 0 ? n : 1);
  const int e = (int)std::floor(61.0 - std::log2(nn * std::max<double>(amax, 1e-30)));
  return std::max(-120, std::min(120, e));
}

// The batch engine's exact tree update of an iteration is a full pass over
// the sorted members (and it rebuilds the cluster totals) unless the totals
// are valid and at most an eighth of the points changed cluster in the last
// iteration; then only the moved points' rows are subtracted from and added
// to the totals.
__host__ __device__ inline bool fx_full_pass(const DevState* s, long long n) {
  return !s->sums_valid || s->moved > (unsigned long long)(n / 8);
}

// The single-problem fixed-point update (k_update_fx) chooses among three
// forms: a full streaming pass while the cluster totals are not valid (the
// start of a run); otherwise only the points whose label changed are taken
// out of their old cluster's totals and added to the new one's -- through
// the workgroups' LDS accumulators and partials when more than
// n / kFxMovedDiv points moved in the last iteration (a pass that reads
// every label but only the moved points' rows), else by global atomics on
// the totals' delta copies.
enum { FX_FULL = 0, FX_MOVED = 1, FX_ATOMIC = 2 };
constexpr long long kFxMovedDiv = 64;
__host__ __device__ inline int fx_mode(const DevState* s, long long n) {
  if (!s->sums_valid) return FX_FULL;
  return s->moved > (unsigned long long)(n / kFxMovedDiv) ? FX_MOVED : FX_ATOMIC;
}

// convergence tests of the k_stats kernel
enum { CONV_NONE = 0, CONV_SHIFT = 1, CONV_SCORE = 2, CONV_STABLE = 3, CONV_SHIFT_SUM = 4 };
// per-run rules of the update and stop test (bit mask)
enum { RULE_EXCLUDE_EMPTY = 1, RULE_KEEP_EMPTY = 2, RULE_COST_RATIO = 4 };

struct PrepParams {
  const float* C;  // K x DP master centroids
  float* Cw;       // Kp x DP  (-2c, fp32 form)
  uint16_t* Chi;   // Kp x DP  (bf16 forms)
  uint16_t* Clo;
  uint8_t* Cq;     // Kp x (DP + kFp8RowPad): e4m3 of -2c * 2^-e, scale e + 127 (fp8 form)
  float* cdn;      // Kp: ||stored operand - (-2c)||, rounded up (single-product forms)
  float* cn;       // Kp, ||c||^2 (kFar for padding rows and excluded clusters)
  float* CT;       // DP x Kp transposed centroids (all-k refinement)
  // Kp x DP bf16 hi and lo of -2c for the pruned all-k passes (null: not
  // written; the bf16x3 form's Chi / Clo hold the same values)
  uint16_t* Phi;
  uint16_t* Plo;
  const int* alive;  // K: 0 = cluster excluded from assignment (+inf)
  const int* done;
  int K, Kp, DP, prec;
};

// One problem of a batch (kmc_batch_*): points [poff, poff + n) of the packed
// point space (poff a multiple of kBatchAlign; the rows up to the next
// problem are zero and belong to the problem's extra "dead" cluster k),
// master clusters [coff, coff + k] (k real + the dead one), distance-operand
// rows [cpoff, cpoff + kpad) (kpad = k rounded up to 64, the padding rows
// carry kFar), distance-kernel tiles [tile0, tile0 + ceil(n / tile)).
struct BatchProblem {
  long long poff;
  int n, k, kpad, coff, cpoff, tile0, nblk;  // nblk: kBatchAlign blocks of the problem
  long long bcoff;  // offset of the problem's (k + 1) x nblk block histogram
};
constexpr int kBatchAlign = 1024;  // point-space alignment of a batch problem
constexpr int kBatchAssignThreads = 512;  // batch distance kernel: 8 waves,
constexpr int kBatchTile = 256;           // 256 points per tile

struct AssignParams {
  const float* X;       // n x DP
  const uint16_t* Xhi;  // n x DP
  const uint16_t* Xlo;
  const uint8_t* Xq;    // n x DP e4m3 of x * 2^-e (fp8 form)
  const int* xsc;       // n, e + 127 (the E8M0 scale of the row)
  const float* xn2;     // n, ||x||^2
  long long n;
  const float* Cw;
  const uint16_t* Chi;
  const uint16_t* Clo;
  const uint8_t* Cq;
  const float* cn;
  // measured operand rounding (single-product forms): per-point ||stored x -
  // x||, per-centroid ||stored -2c - (-2c)||, and the factor applied to the
  // operand term of the band (0: the worst-case constant in alpha is used)
  const float* xdn;
  const float* cdn;
  float hop;
  int K, Kp, kchunk;
  uint32_t idx_mask;        // low mantissa bits that carry the centroid index
  float alpha, beta, gref;  // margin-band constants
  float xpk;  // packing bound relative to ||x||^2 (forms that accumulate without it)
  int* labels;
  int* q2;          // (point, cand1, cand2, cand3 or -1), one region per workgroup
  int* qa;          // points needing all-k refinement, one region per workgroup
  unsigned* q2cnt;  // entries per region
  unsigned* qacnt;
  long long qstride;  // entries per region
  const int* done;
  // batch mode (k_assign<..., BATCH=true>): the packed (problem, tile) queue
  const BatchProblem* bp;
  const int* item_prob;  // problem of each tile
  const int* pdone;      // per-problem stop flags (converged problems are frozen)
  int nitems;
};

struct RefineParams {
  const float* X;
  const float* C;
  const float* CT;
  const int* q2;
  const int* qa;
  const unsigned* q2cnt;
  const unsigned* qacnt;
  long long qstride;
  int nregions;
  const int* alive;
  int* labels;
  const int* done;
  int K, Kp, D, DP;
  // batch mode: the all-k pass is restricted to the point's own problem
  const BatchProblem* bp;  // null for a single problem
  const int* blk_prob;     // problem of each kBatchAlign-point block
  // MFMA-pruned all-k pass (single problem; null cn: every centroid is
  // evaluated exactly): ||c||^2 as prepared (kFar for padding rows and
  // excluded clusters), ||x||^2, and the bf16x3 form's margin constants
  const float* cn;
  const float* xn2;
  const uint16_t* Phi;  // Kp x DP bf16 hi and lo of -2c
  const uint16_t* Plo;
  float alpha3, beta3, gref3;
  int prune_min;  // all-k queues this long or longer take the pruned pass (0: never)
  int pq;         // shorter all-k queues take the single-sweep pruned pass (pq_shape)
};
// shapes of the single-sweep pruned all-k pass: K <= 512 up to padded D = 64,
// K <= 256 at padded D = 128
__host__ __device__ constexpr bool pq_shape(int K, int DP) {
  return DP <= 64 ? K <= 512 : (DP == 128 && K <= 256);
}
// all-k queue length from which the pruned pass runs: one 128-point group
// per CU of the chip
constexpr int kPruneMinPoints = 256 * 128;

struct SortParams {
  const int* labels;
  int* bcount;   // K x nblk (cluster-major) -> exclusive offsets
  int* count;    // K cluster sizes
  int* start;    // K exclusive scan of count
  int* members;  // n point ids sorted by (label, id)
  const int* done;
  long long n;
  int K, CH, nblk, lbits;
};

struct UpdateParams {
  const float* X;
  const float* C;   // centroids the labels were assigned against
  float* Cnext;
  const int* labels;
  const int* members;
  const int* start;
  const int* count;
  double* ipart;    // partial inertias (ordered: K*S, tree: nchunks*S)
  double* tpart;    // tree partials, (nchunks + K) x DP
  const int* done;
  long long n;
  int K, D, DP;
  // fixed-point streaming update (UPDATE_FIXED): workgroup w of dimension
  // slice s (ds dims) writes exact integer sums fpart[s][w][Kp][ds], and
  // (slice 0) the fixed-point sum of ||x||^2 and the member count per cluster
  const double* xn64;  // FP64 ||x||^2 per point
  long long* fpart;
  long long* qpart;    // [w][Kp]
  int* cpart;          // [w][Kp]
  float fscale;        // x -> round(x * fscale), a power of two
  int magic;           // max |x| * fscale < 2^50: the conversion by the double-add form
  double qscale;       // ||x||^2 -> round(||x||^2 * qscale), a power of two
  int Kp, nslice, wgps, ds;
  // incremental form (null Stot: every iteration is a full pass): labels of
  // the last update, exact cluster totals (Kp x DP, Kp, Kp) and the state
  // that chooses between a full pass and the moved points only
  int* prev;
  long long* Stot;
  long long* Qtot;
  int* Ntot;
  DevState* state;
  // exact tree form (batches): per master cluster the fixed-point scale of
  // its problem. The tree update then sums the converted values as 64-bit
  // integers per segment (tpart holds them), keeps the labels in prev and
  // counts the moved points; iterations that are not a full pass take the
  // moved points only (Qtot may be null: no fixed-point ||x||^2).
  const float* cl_scale;
  // incremental fixed-point update: kDeltaCopies copies of the totals'
  // changes (Kp x DP, Kp, Kp each; null: the moved rows go to the totals
  // directly); a wave adds to copy (its index) mod kDeltaCopies, so the
  // atomics on a much-moved cluster's row spread over the copies, and
  // k_finish folds them into the totals and clears them
  long long* Sdel;
  long long* Qdel;
  int* Ndel;
};
constexpr int kDeltaCopies = 8;


struct StatsParams {
  const double* ipart;
  int nparts;
  const double* shift2;
  const int* conv;
  const unsigned* q2cnt;
  const unsigned* qacnt;
  int nregions;
  DevState* state;
  const int* done;
  float tol;
  int K, mode;     // mode: CONV_*
  int cost_ratio;  // also stop when cost_t / cost_{t-1} > 1 - tol (t >= 2)
  int incremental;  // fixed-point update keeps cluster totals across iterations
  long long n;      // points (the full-pass test)
  int fx_modes;     // the update is k_update_fx: fx_mode decides what a full pass is
};

struct FinishParams {
  float* C;
  float* Cnext;
  const double* tpart;  // tree partials (tree update only)
  const int* start;
  int tree;
  int* conv;
  double* shift2;
  const int* count;  // cluster sizes of the assignment just used
  int* alive;        // updated when exclude_empty is set
  const int* done;
  float tol;
  int K, D, DP, exclude_empty;
  int keep_empty;     // an empty cluster keeps its centroid
  int do_prep;        // write the next iteration's distance operands (prep)
  // fixed-point update: the partials to add; count and inertia are written
  int fixed;
  const long long* fpart;
  const long long* qpart;
  const int* cpart;
  int* count_out;      // cluster sizes
  double* ipart_out;   // FP64 inertia of the assignment, per cluster
  int Kp, nslice, wgps, ds;
  double inv_fscale, inv_qscale;
  // incremental fixed-point update: the cluster totals (written after a full
  // pass, read otherwise) and the number of points
  long long* Stot;
  long long* Qtot;
  int* Ntot;
  long long n;
  // exact tree form (tree == 2): integer tree partials of a full pass or the
  // cluster totals, each cluster's inverse fixed-point scale, and the state
  // that chooses between them
  const double* cl_inv_scale;
  const DevState* fx_state;
  PrepParams prep;
  StatsParams stats;  // k_stats, launched after k_finish
  // the incremental update's delta copies (null: none), folded into the
  // totals and cleared here
  long long* Sdel;
  long long* Qdel;
  int* Ndel;
  // labels of the iteration and the labels the totals stand for: after an
  // update of the FX_MOVED form k_finish copies the former into the latter
  const int* labels;
  int* prev;
};

struct LoadParams {
  const float* src;
  size_t ld;
  long long n;
  int D, DP;
  float* X;
  uint16_t* Xhi;
  uint16_t* Xlo;
  uint8_t* Xq;  // fp8 form: e4m3 plane and per-row scales
  int* xsc;
  float* xdn;   // ||stored single-product operand - x|| per row, rounded up
  float* xn2;
  double* xn2_64;
};

// margin-band constants of a precision form (host; kmc_api.hip); with
// measured = true the operand rounding is left out of alpha (the distance
// kernel adds the measured operand term instead) and *hop is its factor
void band_constants(int prec, int D, int idx_bits, float* alpha, float* beta, float* gref,
                    float* xpk, bool measured = false, float* hop = nullptr);

hipError_t launch_prep(const PrepParams& P, hipStream_t s);
const void* assign_batch_kernel_ptr(int DP, int prec);
hipError_t launch_assign_batch(int DP, int prec, const AssignParams& P, int grid, size_t lds,
                               hipStream_t s);
hipError_t launch_assign(int DP, int prec, int variant, const AssignParams& P, int grid,
                         size_t lds, hipStream_t s);
size_t assign_lds(int DP, int prec, int kchunk);
int assign_variants();
// the variant stages every centroid in LDS (no chunked form)
bool assign_variant_resident_only(int variant);
int assign_threads(int variant);
int assign_tile(int variant);
const void* assign_kernel_ptr(int DP, int prec, int variant);
// raise (never lower) a kernel's dynamic-LDS limit to bytes
hipError_t raise_lds_limit(const void* fn, size_t bytes);
hipError_t launch_refine(const RefineParams& P, hipStream_t s);
hipError_t launch_sort(const SortParams& P, hipStream_t s);
hipError_t launch_update(const UpdateParams& P, int mode, hipStream_t s);
hipError_t launch_finish(const FinishParams& F, hipStream_t s);
// fixed-point streaming update: LDS bytes of one workgroup, 0 when it does
// not fit (then the tree update is used)
size_t update_fx_lds(int Kp, int ds);
hipError_t launch_absmax(const float* X, long long count, const double* xn64, long long n,
                         unsigned* amax, unsigned long long* qmax, hipStream_t s);
hipError_t launch_load_points(const LoadParams& P, hipStream_t s);
// k-means++ seeding of K centroid rows of C from the padded points X
// (uni: 1 + (K-1) * max(trials, 1) uniforms in [0,1); D2: n doubles; bsum:
// ceil(n/1024) doubles; chosen: K point indices). trials > 1: greedy
// k-means++ with that many D^2-sampled trials per round (at most 16; cand:
// trials indices, tsum: trials * ceil(n/1024) doubles).
hipError_t launch_plusplus(const float* X, long long n, int DP, float* C, int K,
                           const double* uni, double* D2, double* bsum, long long* chosen,
                           hipStream_t s, int trials = 1, long long* cand = nullptr,
                           double* tsum = nullptr);
hipError_t launch_sum_f64(const double* v, long long n, double* out, double* part256,
                          hipStream_t s);
hipError_t launch_fill_i32(int* p, long long n, int v, hipStream_t s);

}  // namespace kmc
