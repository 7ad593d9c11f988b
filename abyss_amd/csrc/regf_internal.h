// Internal interface between the runtime (regf_runtime.hip) and the kernels
// (regf_kernels.hip).  Not part of the public ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>

namespace abyss {

// Device-resident j-set, SoA with stride `stride` (>= n).  The reference
// packs the same seven arrays [x|y|z|vx|vy|vz|m] with stride N into one
// buffer (src/cuda/cuda_my_acceleration.cu:399-405) plus r_nb^2 in a second.
struct JState {
    const double *x, *y, *z, *vx, *vy, *vz;
    const double *m;      // FP64 mass (FP64 kernel)
    const float *m32;     // FP32 mass (split-precision kernel)
    const double *rnb2;   // r_nb^2 per particle
    int n;
};

// Output of one query launch.  With S j-splits the arrays are indexed
// [s][k] (k = target position); with S == 1 they are the final results.
struct QueryOut {
    double *acc;    // [S][nt][3]
    double *adot;   // [S][nt][3]
    int *cnt;       // [S][nt]  true neighbour count of the split's j-range
    int *list;      // [S][nt][cap]
    // [S][nt][2] sums over the FP32 runs of |a run partial|^2 and |jdot run
    // partial|^2 (checked precision only; null otherwise)
    float *spread = nullptr;
};

// Optional second copy of the final results, written by the finalize into
// mapped host memory (device addresses; all null when unused).
struct FinalMirror {
    double *acc = nullptr;   // [nt][3]
    double *adot = nullptr;  // [nt][3]
    int *nnb = nullptr;      // [nt]
    int *list = nullptr;     // [nt][cap]
};

// Checked precision's flagging done by the finalize of a split launch
// (counter null: none): targets that need recomputation are appended to fq
// (position) / fi (j-index, from tgt) and counted in *counter, which must
// start at 0.  k2 = (ABYSS_CHECK_ERR_SCALE / tol)^2.
struct CheckFlag {
    double k2 = 0.0;
    int *counter = nullptr;
    int *fq = nullptr, *fi = nullptr;
    const int *tgt = nullptr;
};

struct Plan {
    int ipl = 1;     // targets per lane
    int splits = 1;  // j-splits (grid.y)
    int jspan = 0;   // j per split (a multiple of kTile whatever the tile)
    int grid_x = 0;  // target blocks
    int tile = 256;  // j per LDS tile
    int unroll = 2;  // j per unrolled inner-loop step
    int jl = 1;      // lanes per target (1: one target per lane; 4/16/64)
    int wpb = 4;     // wave64s per workgroup (lanes-per-target kernels)
    // checked precision of a small block: every target in FP64 terms in one
    // pass (kPrecRecheck arithmetic, with lists), no flag or recomputation
    bool fp64_terms = false;
};
// Checked-precision blocks of at most this many targets (outside batches)
// run in FP64 terms in one pass: below ~1000 targets that is faster than
// the FP32 pass, its flag and its recomputation together.
constexpr int kCheckedDirectMax = 256;
// Checked-precision blocks above that and up to this many targets (outside
// batches) read the flagged count on the device: the recomputation is a
// grid-stride loop over the flagged targets' work items, queued right behind
// the flag, instead of a host wait for the count and a launch planned for it.
constexpr int kRecheckOnDeviceMax = 4096;

// Plan choice measured by the autotuner for one target-count bucket.
struct TuneEntry {
    int jl;         // lanes per target
    int unroll;     // j-unroll of the one-target-per-lane kernel
    int wg_per_cu;  // workgroups per CU the j-split count aims for
    int tile = 256; // j per LDS tile          } lanes-per-target kernels,
    int wpb = 4;    // wave64s per workgroup   } split and checked precision
};
// The tuned entry for nt targets on `device`, if a table is installed.
bool tuned_plan(int device, int nt, TuneEntry *out);

// Whether a (precision, targets-per-lane, unroll, lanes-per-target) kernel
// variant exists.
bool variant_supported(int prec, int ipl, int unroll, int jl);
// Targets handled by one workgroup under plan p.
int targets_per_block(const Plan &p);
// Whether the lanes-per-target kernels of term mode 0 (split and checked
// precision) exist with j-tile `tile` and `wpb` wave64s per workgroup: tile
// 64, 128 or 256, wpb 1, 2 or 4, and tile <= 64 wpb (every thread loads at
// most one j record of a tile); or the one-shot tile (kTileOneShot, 4
// waves), which holds a whole split of at most that many j and exists for
// 16 and 64 lanes per target.  The other modes run (256, 4) only.
bool shape_supported(int tile, int wpb);
constexpr int kTileOneShot = 1024;

constexpr int kBlock = 256;   // threads per workgroup (4 wave64)
// j per LDS tile of the default shape, the largest tile, and the unit of
// every j-split span: with splits a multiple of every tile, a lane's FP32
// runs fold at the same j whatever the tile, so the shape changes no result.
constexpr int kTile = 256;
constexpr int kFold = 64;     // j per FP32 partial before the FP64 fold
constexpr int kMaxSplits = 2048;  // bound of the finalize kernel's LDS prefix

// ---- batched ensembles ---------------------------------------------------
// One cluster of a batched launch.  `splits` contiguous j-ranges of
// `jspan`; with splits == 1 the pairwise bodies write the final outputs
// (acc/adot/nnb/list) directly, otherwise `part` is this slice's region of
// the batch scratch, indexed [s][q - q_base] with stride nt_stride, and the
// batched finalize merges it into the final outputs.
struct ClusterDesc {
    JState js;
    const int *tgt;
    int nt, jl;
    int tile;        // j per LDS tile (64, 128 or 256)
    int splits, jspan;
    int q_base, nt_stride;
    QueryOut part;
    double *acc, *adot;
    int *nnb, *list;
    // checked precision (null when off): final spread [nt][2] (part.spread
    // holds a split slice's partials); flagged positions / j-indices and
    // their counter; the slice of targets this batch evaluates
    float *spread;
    int *fq, *fi, *count;
    int chk_q0, chk_nq;
    int chk_reset;   // zero *count before this batch's flags (a new pass)
};

// (cluster, target block, j-split): one workgroup's worth of work.
struct WorkItem {
    int cluster, tblock, split, pad;
};

// Per-launch sizes, read by the persistent kernels from device memory so a
// captured graph can replay with a different mix of clusters.
struct BatchHeader {
    int nitems;    // work items in `items`
    int nfin;      // targets of split cluster slices (finalize work)
    int nfc;       // number of split cluster slices
    int pad;
};

constexpr int kMaxSplitsBatch = 512;   // j-splits of one cluster in a batch

// counters[0..1] := 0 (a kernel, so that graph replays need no memset node)
hipError_t launch_batch_reset(int *counters, hipStream_t st);
// Persistent pairwise kernel over the work queue; exits when the queue index
// passes hdr->nitems.  batch_blocks_per_cu(): its resident workgroups per CU
// (the grid is that many per CU).
int batch_blocks_per_cu();
hipError_t launch_batch(const ClusterDesc *cd, const WorkItem *items, const BatchHeader *hdr,
                        int *counter, int cap, int grid, bool est, hipStream_t st);
// Checked precision of a batch: zeroes *cd[a].count of every descriptor
// with chk_reset set (the first batch of a pass over the cluster's target
// blocks); the counts of later batches of the pass keep accumulating.
hipError_t launch_batch_count_reset(const ClusterDesc *cd, int ncd, hipStream_t st);
// Checked precision of a batch: over the chk_nq targets of each of the ncd
// descriptors (ck_off[k]: prefix of their counts, ck_off[ncd] = total),
// appends the flagged ones to the descriptor's fq / fi and counts them in
// *count (which must start at 0); the rule of launch_check_flag.
hipError_t launch_batch_check_flag(const ClusterDesc *cd, const int *ck_off, int ncd, int total,
                                   double k2, hipStream_t st);
// Checked precision of a batch, the recomputation: counter reset, the
// persistent recompute over the items of the ncd descriptors (tgt = flagged
// j-indices, nt = their count, part = split partials with stride nt), then
// the merge of every flagged target (off[k]: prefix of the descriptors'
// counts, off[ncd] = total) over the cluster's results at fq.
hipError_t launch_batch_recheck(const ClusterDesc *cd, const WorkItem *items,
                                const BatchHeader *hdr, int *counter, const int *off, int ncd,
                                int total, int grid, int fin_grid, hipStream_t st);
// p[0..n) := 0
hipError_t launch_zero_ints(int *p, int n, hipStream_t st);
// Persistent finalize over the targets of split slices: fin_off[k] (k <=
// nfc) is the prefix of slice target counts, fin_cl[k] the descriptor index.
hipError_t launch_batch_finalize(const ClusterDesc *cd, const int *fin_cl, const int *fin_off,
                                 const BatchHeader *hdr, int cap, int grid, hipStream_t st);

// Launch the pairwise kernel for nt targets against js (all j), split into
// plan.splits contiguous j-ranges.  prec: an ABYSS_PREC_* mode, or
// kPrecRecheck (the recomputation of checked precision).
// reset (checked precision, lanes-per-target plans): a counter the kernel
// zeroes before any later launch on the stream reads it.
hipError_t launch_regf(const JState &js, const int *tgt, int nt, const Plan &plan,
                       const QueryOut &out, int cap, int prec, hipStream_t st,
                       int *reset = nullptr);

// Internal mode of launch_regf: the neighbour test of split precision (FP32
// r^2 of the rounded differences, so the regular set is the same) with FP64
// terms and a Newton-refined rsqrt, accumulated in FP64.  Used with cap = 0
// (no list is written).
constexpr int kPrecRecheck = 100;

// Sum the S partial results in split order and concatenate the neighbour
// sub-lists in split (= ascending j) order into the final outputs; with
// `spread` non-null also the S spread partials.
hipError_t launch_finalize(const QueryOut &part, int S, int nt, int cap,
                           double *acc, double *adot, int *nnb, int *list,
                           const FinalMirror &mirror, float *spread, hipStream_t st,
                           const CheckFlag &flag = CheckFlag{});

// Checked precision.  counters[0] must be 0.  Appends to (fq, fi) the
// position and j-index of every target q whose a or jdot may be off by
// more than the tolerance: spread[q][0] * k2 > |acc[q]|^2 or spread[q][1] *
// k2 > |adot[q]|^2.  counters[0] ends as their number (order of the
// entries unspecified).
hipError_t launch_check_flag(const double *acc, const double *adot, const float *spread,
                             const int *tgt, int nt, double k2, int *counters, int *fq, int *fi,
                             hipStream_t st);
// Sums the recomputation's S partials of each of the m flagged targets (in
// split order) and writes them over acc/adot at fq[k].
hipError_t launch_recheck_finalize(const QueryOut &part, int S, int m, const int *fq,
                                   double *acc, double *adot, hipStream_t st);
// The same two steps with the flagged count read on the device (*count, as
// left by launch_check_flag) and no host wait: plan p is for all nt targets
// (partials [S][nt]); workgroups and waves past the count exit at once.
// Writes the count to *m_report (host memory) and the results also to the
// mirror when it is set.
hipError_t launch_recheck_dev(const JState &js, const int *fi, const int *count, int nt,
                              const Plan &p, const QueryOut &out, const int *fq, double *acc,
                              double *adot, const FinalMirror &mirror, int *m_report,
                              hipStream_t st);

// The recomputation of the first *count targets of fi (count on the device,
// at most nt): `grid` workgroups loop over the (target block, split) items of
// plan p (lanes per target, splits and span; 256-record tiles, 4 waves), then
// the merge.  Partials: rows s * nt + k of out.  Writes the count to
// *m_report (mapped host memory) when it is set.
hipError_t launch_recheck_loop(const JState &js, const int *fi, const int *count, int nt,
                               const Plan &p, int grid, const QueryOut &out, const int *fq,
                               double *acc, double *adot, int *m_report, hipStream_t st);

}  // namespace abyss
