/*
 * abyss_regf.h -- C ABI of the MI355X (gfx950) regular-force library.
 *
 * Part 1 keeps the six entry points ABYSS's host code calls, with their
 * signatures unchanged and CUDA_REAL = double (the serial build's choice,
 * src/defs.h:97-102; src_mpi's callers have matching double branches,
 * src_mpi/cuda/CalculateRegularAcceleration.cpp:152-153,380-429).
 * Part 2 adds explicit contexts (one per cluster / per thread), a
 * device-resident launch path, tuning controls and error reporting.
 * Part 3 is the batched-ensemble and served-ensemble interface.
 * Part 4 is one query that allocates nothing, for callers that cannot.
 *
 * Semantics shared by every query (reference: src_mpi/cuda/cuda_kernels.cu
 * :269-450 and the callers cited below):
 *   - acc/adot are the REGULAR part only: the sum over j with
 *     r_ij^2 > r_nb,i^2 of m_j r_ij/|r_ij|^3 and
 *     m_j [v_ij/|r_ij|^3 - 3 (r_ij . v_ij) r_ij/|r_ij|^5]  (G = 1, no softening),
 *     with r_ij = x_j - x_i, v_ij = v_j - v_i.
 *   - j with r_ij^2 <= r_nb,i^2 and j != i (by index) are neighbours; they are
 *     listed in ascending j order at row k of NeighborList, row stride
 *     NumNeighborMax; NumNeighbor[k] is their count.  A coincident j != i
 *     (r = 0) is a neighbour and adds no force.
 *   - r_nb,i^2 is the TARGET's own value from the last upload (r2[i]).
 *   - neighbour indices and targets index the j-set of the last upload.
 *   - If a list would exceed NumNeighborMax, the first NumNeighborMax
 *     neighbours are written, NumNeighbor[k] = -(true count), and the call's
 *     status is ABYSS_ERR_NEIGHBOR_OVERFLOW (the reference asserts instead,
 *     src_mpi/cuda/cuda_kernels.cu:363,432).
 *
 * Precision (default ABYSS_PREC_CHECKED, every target's a and adot within
 * 1e-6 relative of FP64).  ABYSS_PREC_SPLIT: positions and velocities are
 * FP64 on the device and the i-j differences are taken in FP64; r^2, rsqrt
 * and the force/jerk terms are FP32, summed in FP32 over short runs of j and
 * folded into FP64 per-particle sums (a target whose jerk sum cancels can
 * miss 1e-6).  ABYSS_PREC_MIXED keeps r^2, the neighbour
 * test and the terms in FP64 as well (sums in FP64); only rsqrt is the FP32
 * hardware estimate (~1e-7 relative).  ABYSS_PREC_REFINED adds one FP64
 * Newton step to that estimate (~1e-14).  ABYSS_PREC_FP64 evaluates
 * everything in FP64 with a correctly rounded rsqrt.  ABYSS_PREC_CHECKED is
 * ABYSS_PREC_SPLIT plus a per-target check: targets whose jerk sum cancels
 * so far that its FP32 rounding may exceed the check tolerance (default
 * 1e-6 relative) are recomputed with FP64 terms (as ABYSS_PREC_REFINED) over
 * the same regular set; every other result is bitwise that of
 * ABYSS_PREC_SPLIT, and the lists always are.
 *
 * Threading: the legacy entry points share one process-wide default context
 * and are not re-entrant (as in the reference, whose state is file-static,
 * src_mpi/cuda/cuda_my_acceleration.cu:23-88).  A context of part 2 may be
 * used by one host thread at a time; distinct contexts are independent.
 * All calls are synchronous unless named *_launch / *_async.
 */
#ifndef ABYSS_REGF_H
#define ABYSS_REGF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Neighbour-list row stride of the legacy entry points: a compile-time
 * contract shared with the caller, which allocates NumTarget*NumNeighborMax
 * ints (src_mpi/cuda/CalculateRegularAcceleration.cpp:91).  1000 matches
 * src_mpi/def.h:20; override at build time with -DABYSS_NUM_NEIGHBOR_MAX=...
 * or at run time (before OpenDevice) with abyss_set_num_neighbor_max(). */
#ifndef ABYSS_NUM_NEIGHBOR_MAX
#define ABYSS_NUM_NEIGHBOR_MAX 1000
#endif

/* ---- status codes --------------------------------------------------- */
#define ABYSS_OK                    0
#define ABYSS_ERR_NOT_OPEN          1  /* device not initialised / opened */
#define ABYSS_ERR_NO_DEVICE         2  /* no usable HIP device */
#define ABYSS_ERR_INVALID_ARG       3  /* null pointer, bad size or count */
#define ABYSS_ERR_TARGET_RANGE      4  /* target index outside [0, N) */
#define ABYSS_ERR_INVALID_INPUT     5  /* non-finite x/v/m, m < 0, r2 < 0 or NaN */
#define ABYSS_ERR_NO_JSET           6  /* query before any successful upload */
#define ABYSS_ERR_HIP               7  /* a HIP runtime call failed */
#define ABYSS_ERR_OUT_OF_MEMORY     8
#define ABYSS_ERR_NEIGHBOR_OVERFLOW 9  /* results valid; some NumNeighbor < 0 */
#define ABYSS_ERR_ADMISSION        10  /* served ensemble: HBM budget exhausted */
#define ABYSS_ERR_BUSY             11  /* served ensemble: no free slot, or request turned away */
#define ABYSS_ERR_SCRATCH_LIMIT    12  /* no-allocation query: sizes exceed its static scratch */

#define ABYSS_PREC_SPLIT   0
#define ABYSS_PREC_FP64    1
#define ABYSS_PREC_MIXED   2
#define ABYSS_PREC_REFINED 3
#define ABYSS_PREC_CHECKED 4

/* ===================================================================== */
/* Part 1: legacy entry points (src_mpi/cuda/cuda_functions.h:3-10,      */
/* defined at src_mpi/cuda/cuda_my_acceleration.cu:1087-1106;            */
/* identical list in src/cuda/cuda_functions.h:3-10).                    */
/* ===================================================================== */

/* Scalar type of the legacy entry points, the reference's CUDA_REAL.
 * libabyss_regf.so exports them with double (src/defs.h:101; src_mpi with
 * `#define CUDA_FLOAT` removed, src_mpi/def.h:69).  libabyss_regf_f32.so
 * exports the same six symbols with float for an src_mpi build that keeps
 * CUDA_FLOAT (src_mpi/def.h:69-74); a caller of that library compiles with
 * -DABYSS_LEGACY_FLOAT.  The float library widens the caller's inputs to
 * FP64 and rounds the FP64 results to float; everything in between is the
 * FP64-input path. */
#ifdef ABYSS_LEGACY_FLOAT
typedef float abyss_legacy_real;
#else
typedef double abyss_legacy_real;
#endif

/* Selects device (*irank mod device count), creates the stream and marks
 * the library open.  Replaces _InitializeDevice
 * (src/cuda/cuda_my_acceleration.cu:415-455, which also sets is_open). */
void InitializeDevice(int *irank);

/* Initialises (if needed) and opens; a second open only reports it.
 * Replaces _OpenDevice (src_mpi/cuda/cuda_my_acceleration.cu:987-1008). */
void OpenDevice(const int *irank);

/* Closes the library and frees its device and pinned buffers (the reference
 * only flips is_open, src_mpi/cuda/cuda_my_acceleration.cu:1012-1047). */
void CloseDevice(void);

/* Prints and resets the call/interaction/time counters, with the 60-flop
 * interaction convention (src_mpi/cuda/cuda_my_acceleration.cu:1051-1062).
 * grav(s) is the wall-clock time spent in CalculateAccelerationOnDevice, as
 * the reference's host timers measured it. */
void ProfileDevice(int *irank);

/* Replaces the whole j-set: N = *NNB particles, masses m[N], positions
 * x[N][3], velocities v[N][3], r2[N] = r_nb^2 per particle; mdot is ignored.
 * Inputs are copied before return.  A rejected upload (for example a
 * non-finite position) leaves no j-set: later queries fail with
 * ABYSS_ERR_NO_JSET until an upload succeeds.  Replaces _ReceiveFromHost
 * (src_mpi/cuda/cuda_my_acceleration.cu:603-778; src/...:313-411). */
void SendToDevice(int *NNB, abyss_legacy_real m[], abyss_legacy_real x[][3],
                  abyss_legacy_real v[][3], abyss_legacy_real r2[], abyss_legacy_real mdot[]);

/* Regular acc/adot and neighbour lists of *NumTarget targets
 * h_target_list[k] in [0, N), 0 < *NumTarget <= N.  Outputs for k <
 * *NumTarget: acc[k][3], adot[k][3], NumNeighbor[k],
 * NeighborList[k*NumNeighborMax + 0..count).  Replaces GetAcceleration
 * (src/cuda/cuda_my_acceleration.cu:66-302; src_mpi/...:95-344). */
void CalculateAccelerationOnDevice(int *NumTarget, int *h_target_list,
                                   abyss_legacy_real acc[][3], abyss_legacy_real adot[][3],
                                   int NumNeighbor[], int *NeighborList);

/* ===================================================================== */
/* Part 2: status, contexts, device-resident path, tuning.              */
/* No counterpart in the reference (which has no return codes).         */
/* ===================================================================== */

/* Status of the last legacy call on this process (ABYSS_OK or an error). */
int abyss_last_error(void);
const char *abyss_error_string(int code);
/* Human-readable detail of the last error of the legacy API. */
const char *abyss_last_error_detail(void);
/* Legacy list stride; setter only allowed while the library is closed. */
int abyss_num_neighbor_max(void);
int abyss_set_num_neighbor_max(int cap);
/* Legacy API precision: any ABYSS_PREC_* mode (default ABYSS_PREC_CHECKED,
 * which keeps every target's a and adot within 1e-6 relative of FP64). */
int abyss_set_precision(int mode);
/* What a failing legacy call does.  ABYSS_LEGACY_EXIT (the default) prints
 * the error to stderr and ends the process with exit(1), as the reference
 * does on neighbour overflow (src/cuda/cuda_my_acceleration.cu:221-226,
 * src_mpi/cuda/cuda_my_acceleration.cu:328) and through its asserts
 * (:103-104); its callers never check a status.  ABYSS_LEGACY_RETURN
 * prints, records the status for abyss_last_error() and returns; an
 * overflowed list then comes back truncated with NumNeighbor[k] =
 * -(true count). */
#define ABYSS_LEGACY_EXIT   0
#define ABYSS_LEGACY_RETURN 1
int abyss_set_legacy_errors(int mode);
/* Number of HIP devices visible (0 without a GPU; never fails). */
int abyss_device_count(void);

typedef struct abyss_ctx_s *abyss_ctx_t;

/* Creates a context on `device` with neighbour-list stride `cap` (>0),
 * precision ABYSS_PREC_CHECKED. */
int abyss_ctx_create(int device, int cap, abyss_ctx_t *out);
int abyss_ctx_destroy(abyss_ctx_t ctx);
int abyss_ctx_set_precision(abyss_ctx_t ctx, int mode);
/* Forces targets per lane (1, 2 or 4), the j-split count (1..2048), the
 * inner-loop j unroll (2 or 4) and the lanes sharing one target (1, 4, 16
 * or 64; > 1 implies one target per lane group); 0 = automatic for each. */
int abyss_ctx_set_tuning(abyss_ctx_t ctx, int i_per_lane, int j_splits, int j_unroll,
                         int lanes_per_target);
/* Forces the workgroup shape of the lanes-per-target kernels under split and
 * checked precision: j per LDS tile (64, 128 or 256) and wave64s per
 * workgroup (1, 2 or 4; a workgroup serves waves * 64 / lanes_per_target
 * targets), with tile <= 64 * waves; 0 = automatic for each.  Tile 1024 is
 * the one-shot tile of small blocks (4 waves; 16 or 64 lanes per target,
 * j-splits of at most 1024 j; other plans fall back to 256), also for the
 * one FP64-term pass of a checked block of at most 256 targets, which
 * otherwise ignores a forced shape.  The
 * reference's compile-time BatchSize and NNB_per_block (src_mpi/def.h:77,79).
 * No shape changes a result bit. */
int abyss_ctx_set_shape(abyss_ctx_t ctx, int tile, int waves_per_block);
/* Upper bound on the split-j scratch, S * m * (52 + 4 * cap) bytes for S
 * j-splits of m targets (bytes; default 2 GiB).  A smaller bound caps S. */
int abyss_ctx_set_scratch_limit(abyss_ctx_t ctx, unsigned long long bytes);

/* Upload a j-set: x, v as [n][3], m and r2 as [n] (FP64, host memory). */
int abyss_ctx_upload(abyss_ctx_t ctx, int n, const double *m, const double *x,
                     const double *v, const double *r2);
/* Synchronous query with host buffers (list row stride = ctx cap). */
int abyss_ctx_query(abyss_ctx_t ctx, int nt, const int *targets, double *acc,
                    double *adot, int *nnb, int *list);

/* Device-resident path: upload targets once, then enqueue any number of
 * launches on the context stream (no host synchronisation, results stay in
 * HBM), then fetch the last launch's results. */
int abyss_ctx_set_targets(abyss_ctx_t ctx, int nt, const int *targets);
int abyss_ctx_launch(abyss_ctx_t ctx);
int abyss_ctx_sync(abyss_ctx_t ctx);
int abyss_ctx_fetch(abyss_ctx_t ctx, double *acc, double *adot, int *nnb,
                    int *list);
/* hipStream_t of the context (for event timing by the caller). */
void *abyss_ctx_stream(abyss_ctx_t ctx);
/* Plan chosen for the current targets: targets per lane, j splits, target
 * blocks, j per LDS tile, j unroll, lanes per target. */
int abyss_ctx_plan(abyss_ctx_t ctx, int *i_per_lane, int *j_splits,
                   int *grid_x, int *tile, int *j_unroll, int *lanes_per_target);
/* Workgroup of the plan: wave64s per workgroup and targets per workgroup. */
int abyss_ctx_plan_block(abyss_ctx_t ctx, int *waves_per_block, int *targets_per_block);
/* Detail string of the context's last error. */
const char *abyss_ctx_error_detail(abyss_ctx_t ctx);
/* Force-kernel timing: with it on, every launch records a HIP event pair
 * on the context stream around its pairwise force kernel (not the merge of
 * the j-splits, nor checked precision's flagging and recomputation);
 * kernel_timing returns the summed milliseconds and the number of launches
 * since it was last switched (it waits for the last recorded launch).
 * The GPU-time, per-kernel form of the timing ProfileDevice reports
 * (src_mpi/cuda/cuda_my_acceleration.cu:1051-1062). */
int abyss_ctx_set_kernel_timing(abyss_ctx_t ctx, int enable);
int abyss_ctx_kernel_timing(abyss_ctx_t ctx, double *force_ms, long long *launches);

/* Checked precision (ABYSS_PREC_CHECKED).  A launch sums, for each target,
 * the squared norms of its FP32 run partials of a and of jdot (Sa, Sj; the
 * root of each is the scale of that sum's rounding error, since the largest
 * terms dominate both) and recomputes with FP64 terms every target with
 *   ABYSS_CHECK_ERR_SCALE * sqrt(Sa) > tol * |a|  or
 *   ABYSS_CHECK_ERR_SCALE * sqrt(Sj) > tol * |adot|,
 * before the results are fetched.  A block of at most 256 targets (outside
 * batched ensembles) instead evaluates every target with FP64 terms in one
 * pass, which is faster there and waits for nothing; its lists are those of
 * the FP32 pass as well.  A block of 257 to 4096 targets reads the flagged
 * count on the device (the recomputation loops over it), so its launch does
 * not wait either.  A larger launch waits once on the host (for the flagged
 * count, to plan the recomputation).  No counterpart in the reference (whose
 * FP32 build has no accuracy control).
 *   set_check_tolerance: tol (> 0; default 1e-6);
 *   last_recheck: targets recomputed by the last launch (all of a block of
 *     at most 256); waits for the launch when its count was read on the
 *     device;
 *   fetch_spread: [nt][2] = sqrt(Sa), sqrt(Sj) per target of the last
 *     launch (a launch of more than 256 targets). */
#define ABYSS_CHECK_ERR_SCALE 1.0e-6
int abyss_ctx_set_check_tolerance(abyss_ctx_t ctx, double tol);
int abyss_ctx_last_recheck(abyss_ctx_t ctx, int *count);
int abyss_ctx_fetch_spread(abyss_ctx_t ctx, float *spread);

/* Device-resident particle state: j prediction on the device instead of
 * the caller's O(N) predict + pack + copy before every regular step
 * (src_mpi/cuda/CalculateRegularAcceleration.cpp:330-449).
 *   upload_state: x, v at each particle's own time t, its total
 *     acceleration a and jerk adot there ([n][3] each), t, m, r2 ([n]);
 *   update_state: replaces `count` particles idx[k] (e.g. after their
 *     corrector), arrays indexed by k;
 *   predict: every particle to t_new with the second-order predictor
 *     x + h(v + h/2(a + h/3 adot)), v + h(a + h/2 adot), h = t_new - t
 *     (src_mpi/Particle/UpdateParticle.cpp:37-59), written to the j-set
 *     the next query uses.  The arithmetic is plain IEEE double in the
 *     reference's order, so it equals a host prediction bit for bit.
 * A plain abyss_ctx_upload discards the resident state. */
int abyss_ctx_upload_state(abyss_ctx_t ctx, int n, const double *m, const double *x,
                           const double *v, const double *a, const double *adot,
                           const double *t, const double *r2);
int abyss_ctx_update_state(abyss_ctx_t ctx, int count, const int *idx, const double *m,
                           const double *x, const double *v, const double *a,
                           const double *adot, const double *t, const double *r2);
int abyss_ctx_predict(abyss_ctx_t ctx, double t_new);

/* Irregular force on the device.  The reference sums the irregular force
 * and jerk over each particle's neighbour list on the host, in FP64, with
 * m_r3 = m_j/(r2*sqrt(r2)), a += m_r3 r, adot += m_r3 (v - 3 r (r.v)/r2)
 * (Particle::computeAccelerationIrr,
 * src_mpi/Particle/ComputeAcceleration.cpp:14-190).
 *   keep_lists(ctx, 1): from then on every abyss_ctx_launch / _query also
 *     stores its neighbour lists in HBM by particle index (n x cap ints);
 *     a particle's list is the one from its last regular step, and a
 *     particle never targeted has an empty list.  A j-set of another size
 *     empties all lists.  keep_lists(ctx, 0) frees them.
 *   irregular(ctx, nt, targets, acc, adot): for each target i, the FP64
 *     sum over its resident list evaluated on the current j-set (e.g. after
 *     abyss_ctx_predict), acc/adot [nt][3].  A neighbour at r = 0 adds
 *     nothing.  Lists truncated by overflow sum their first cap entries.
 *     Synchronous. */
int abyss_ctx_keep_lists(abyss_ctx_t ctx, int enable);
int abyss_ctx_irregular(abyss_ctx_t ctx, int nt, const int *targets, double *acc,
                        double *adot);

/* One regular step with the neighbour-set correction and the fourth-order
 * regular corrector fused on the device, in place of the caller's per-
 * particle loop after CalculateAccelerationOnDevice
 * (Particle::updateRegularParticleCuda,
 * src_mpi/Particle/ComputeAcceleration.cpp:400-661).  Needs keep_lists on
 * and the j-set predicted to the step's end time t_new (e.g. by
 * abyss_ctx_predict).  For each target k (particle i):
 *   - acc/adot/nnb/list: the new regular force and list, as abyss_ctx_query;
 *   - with L_old = i's resident list (its last regular step; empty if none)
 *     and L_new = the list just found, terms f_j on the current j-set:
 *       a_irr = sum_{L_new} f_j,
 *       a_tmp = sum_{L_new \ L_old} f_j - sum_{L_old \ L_new} f_j;
 *   - with dt = dt[k] (> 0), a0 = a_old[k][0..2], j0 = a_old[k][3..5] (the
 *     previous regular force and jerk), a1 = acc[k], j1 = adot[k]:
 *       da_dt2 = (a0 - a1 - a_tmp)/dt^2, adot_dt = (j0 + j1 + adot_tmp)/dt,
 *       a2 = -6 da_dt2 - 2 adot_dt - 2 j0/dt, a3 = (12 da_dt2 + 6 adot_dt)/dt,
 *       x_new = x + a2 dt^4/24 + a3 dt^5/120, v_new = v + a2 dt^3/6 + a3 dt^4/24
 *     with x, v the target's own j-set entry (at t_new);
 *   - out[k][18] = x_new[3], v_new[3], a2[3], a3[3], a_irr[3], adot_irr[3];
 *   - i's resident list becomes L_new.
 * A neighbour at r = 0 adds nothing (the reference would divide by zero). */
int abyss_ctx_regular_step(abyss_ctx_t ctx, int nt, const int *targets, const double *dt,
                           const double *a_old, double *acc, double *adot, int *nnb,
                           int *list, double *out);

/* Start-up second and third time derivatives of the acceleration, split
 * into regular (r^2 > r_nb,i^2) and irregular parts, summed over all j != i
 * in FP64 on the resident state of abyss_ctx_upload_state (x, v, a = total
 * acceleration, adot = total jerk, m, r2), in place of the caller's O(N^2)
 * start-up loop (CalculateAcceleration23,
 * src_mpi/Particle/Initialize.cpp:102-185).
 *   out[k][12] = a_reg2[3], a_reg3[3], a_irr2[3], a_irr3[3] of targets[k]
 *   (the reference adds them to zero-initialised a_reg/a_irr[.][2..3]).
 * reference_signs = 1 reproduces the reference's expressions, whose
 * projections x.(a_i - a_j)/r^2, v.(a_i - a_j)/r^2 and x.(adot_i - adot_j)/r^2
 * enter B and C with the opposite sign to the relative acceleration
 * a_j - a_i used by its leading terms; reference_signs = 0 uses
 * x.(a_j - a_i) etc., which gives the exact derivatives of the pair forces.
 * A j at r = 0 adds nothing. */
int abyss_ctx_derivs23(abyss_ctx_t ctx, int nt, const int *targets, int reference_signs,
                       double *out);

/* Neighbour-radius adaptation after a regular step (Particle::updateRadius,
 * src_mpi/Particle/UpdateParticle.cpp:108-139) for the targets of the last
 * launch, from their neighbour counts (|count| when overflowed): r2 (the
 * squared radius, the upload's r2) is scaled by (1 - a), x = count - fix,
 * c = nmax - fix, a = 0.9 |x| x / c^2 above fix, a = 0.5 x^3 / c^3 below
 * (the radius grows), unchanged at fix.  The new r2 replaces the j-set's
 * (the next launch uses it) and the resident state's, and is written to
 * r2_out[nt].  The reference uses fix = FixNumNeighbor (100) and
 * nmax = NumNeighborMax (1000), src_mpi/def.h:18,20. */
int abyss_ctx_update_radius(abyss_ctx_t ctx, int fix, int nmax, double *r2_out);

/* Plan autotuner.  Times candidate plans of the FP32 pass (lanes per target
 * 4, 16 or 64, j per LDS tile, wave64s per workgroup, workgroups per CU ->
 * j-split count) for target counts 1, 4, 16, ... n on a synthetic
 * n-particle cluster (uniform cube, n >= 256) and installs the fastest per
 * power-of-two target-count bucket for every later automatic plan of that
 * pass on `device` (split precision, and checked precision above 256
 * targets; forced tunings still win).  Writes the table to
 * `cache_path` when not NULL (one line per bucket: bucket, lanes per
 * target, j-unroll, workgroups per CU, tile, waves).  Takes seconds to a
 * minute; results of a query depend on its plan only through the FP64
 * summation order of its j-splits.  Replaces the reference's compile-time
 * BatchSize / GridDimY / NNB_per_block (src_mpi/def.h:77-79). */
int abyss_autotune(int device, int n, const char *cache_path);
/* The same on the caller's own n-particle j-set (the abyss_ctx_upload
 * arrays), so the table is measured on the workload it will serve. */
int abyss_autotune_jset(int device, int n, const double *m, const double *x,
                        const double *v, const double *r2, const char *cache_path);
/* Installs a table written by abyss_autotune. */
int abyss_autotune_load(int device, const char *cache_path);
/* Removes the table: plans use the built-in rules again. */
int abyss_autotune_clear(int device);

/* ===================================================================== */
/* Part 3: batched ensembles and the served ensemble.                   */
/* No counterpart in the reference (one cluster per process there).     */
/* ===================================================================== */

/* An ensemble holds up to max_clusters independent clusters, each with its
 * own device-resident FP64 j-set.  abyss_ens_launch evaluates every cluster
 * that has targets set, in ONE persistent kernel launch replayed from a
 * captured HIP graph (work items of all clusters share a device queue).  A
 * cluster's decomposition depends only on its own N and target count, so
 * its results are bitwise identical alone and in any mix. */
typedef struct abyss_ens_s *abyss_ens_t;

int abyss_ens_create(int device, int cap, int max_clusters, abyss_ens_t *out);
int abyss_ens_destroy(abyss_ens_t ens);
/* (Re)load cluster slot `cluster` (0 <= cluster < max_clusters). */
int abyss_ens_upload(abyss_ens_t ens, int cluster, int n, const double *m, const double *x,
                     const double *v, const double *r2);
/* Frees a cluster slot and its device memory. */
int abyss_ens_release(abyss_ens_t ens, int cluster);
/* Targets of a cluster for the next launch; nt = 0 leaves it out. */
int abyss_ens_set_targets(abyss_ens_t ens, int cluster, int nt, const int *targets);
/* One batched step over all clusters with targets (asynchronous). */
int abyss_ens_launch(abyss_ens_t ens);
int abyss_ens_sync(abyss_ens_t ens);
/* Target blocks (work items) of a cluster's current targets and the
 * targets per block; the launch covers blocks [first, first + count) when a
 * range is set (count < 0: all; set_targets resets to all).  Splitting a
 * cluster's blocks over several launches gives the same bits. */
int abyss_ens_blocks(abyss_ens_t ens, int cluster, int *nblocks, int *targets_per_block);
int abyss_ens_set_block_range(abyss_ens_t ens, int cluster, int first, int count);
/* Lanes per target and j per LDS tile of a cluster's work items (from the
 * tuned table when one is installed; workgroups of the batch are 4 waves). */
int abyss_ens_plan(abyss_ens_t ens, int cluster, int *lanes_per_target, int *tile);
/* Results of a cluster from the last launch (list row stride = cap). */
int abyss_ens_fetch(abyss_ens_t ens, int cluster, double *acc, double *adot, int *nnb,
                    int *list);
void *abyss_ens_stream(abyss_ens_t ens);
/* Precision of every cluster: ABYSS_PREC_CHECKED (default) or
 * ABYSS_PREC_SPLIT (ABYSS_PREC_CHECKED: the flag rule of abyss_ctx_set_check_tolerance, over
 * every target a launch evaluates).  Under ABYSS_PREC_CHECKED a launch
 * waits once on the host after the batch (for the flagged counts), then
 * queues each cluster's recomputation; last_recheck reports a cluster's
 * recomputed targets of the last launch. */
int abyss_ens_set_precision(abyss_ens_t ens, int mode);
int abyss_ens_set_check_tolerance(abyss_ens_t ens, double tol);
int abyss_ens_last_recheck(abyss_ens_t ens, int cluster, int *count);
/* Device bytes held by the ensemble (all clusters + descriptors). */
unsigned long long abyss_ens_device_bytes(abyss_ens_t ens);
const char *abyss_ens_error_detail(abyss_ens_t ens);

/* Served ensemble: clusters stay resident; requests (one regular step of
 * one cluster) are submitted from any host thread.  A worker thread runs
 * time-sliced continuous batching: each batch carries at least one target
 * block of every running request, smallest remaining work first, and more
 * blocks until a pair budget is used; requests that arrive meanwhile join
 * the next batch.  A completion thread copies finished results to the host
 * on its own stream while the next batch runs.  At most one request per
 * cluster runs at a time (later ones queue).  New clusters are admitted
 * while their estimated HBM footprint fits the budget. */
typedef struct abyss_srv_s *abyss_srv_t;

int abyss_srv_create(int device, int cap, int max_clusters, unsigned long long hbm_budget,
                     abyss_srv_t *out);
int abyss_srv_destroy(abyss_srv_t srv);
/* Work budget of one batch in pair-interactions (default 3e10, ~25 ms). */
int abyss_srv_set_batch_budget(abyss_srv_t srv, double pairs);
/* Precision of the served clusters (abyss_ens_set_precision; takes effect
 * from the next batch).  Under ABYSS_PREC_CHECKED a request whose blocks
 * span several batches gets each batch's flagged targets recomputed with a
 * plan for that batch's count: the same values alone and in the mix to
 * ~1e-13 relative (unflagged targets bit for bit). */
int abyss_srv_set_precision(abyss_srv_t srv, int mode);
int abyss_srv_set_check_tolerance(abyss_srv_t srv, double tol);
/* Admits a cluster (ABYSS_ERR_ADMISSION if it does not fit the budget,
 * ABYSS_ERR_BUSY if no slot is free); returns its id in *cluster. */
int abyss_srv_add_cluster(abyss_srv_t srv, int n, const double *m, const double *x,
                          const double *v, const double *r2, int *cluster);
int abyss_srv_remove_cluster(abyss_srv_t srv, int cluster);
/* Queues one regular step of `cluster` for `nt` targets (copied). */
int abyss_srv_submit(abyss_srv_t srv, int cluster, int nt, const int *targets,
                     long long *ticket);
/* Blocks until the request is done; copies its results (list row stride =
 * cap) and its latency (submit -> results on the host) in milliseconds. */
int abyss_srv_wait(abyss_srv_t srv, long long ticket, double *acc, double *adot, int *nnb,
                   int *list, double *latency_ms);
/* Completed requests, batches, summed interactions and latency percentiles
 * (p50, p99; ms) since creation or the last reset. */
int abyss_srv_stats(abyss_srv_t srv, long long *completed, long long *batches,
                    double *interactions, double *p50_ms, double *p99_ms);
int abyss_srv_reset_stats(abyss_srv_t srv);
/* Request admission: limit on the outstanding work (pair-interactions,
 * nt * (N - 1) summed over submitted requests not yet completed); 0 (the
 * default) means unlimited.  With a limit, abyss_srv_submit returns
 * ABYSS_ERR_BUSY (no ticket) for a request that would take the outstanding
 * work past it, unless nothing is outstanding. */
int abyss_srv_set_admission(abyss_srv_t srv, double max_outstanding_pairs);
/* Requests turned away by admission, the most requests waiting to start at
 * once, and the peak outstanding work (pairs) since creation or the last
 * reset. */
int abyss_srv_queue_stats(abyss_srv_t srv, long long *rejected, long long *max_pending,
                          double *peak_outstanding_pairs);
/* GPU time of the batches run since creation or the last reset (HIP events
 * around each batch on the ensemble stream; a checked batch's host wait for
 * its flagged count falls inside) and the wall-clock time over the same
 * span, in ms: their ratio is the fraction of time the GPU was busy. */
int abyss_srv_gpu_busy(abyss_srv_t srv, double *busy_ms, double *wall_ms);
/* Estimated device bytes a cluster of n particles needs (admission rule). */
unsigned long long abyss_srv_cluster_footprint(abyss_srv_t srv, int n);

/* ===================================================================== */
/* Part 4: no-allocation query.  No counterpart in the reference: it is  */
/* for callers that cannot allocate device memory (code under HIP graph  */
/* capture, tight serving loops, harnesses that forbid hipMalloc).       */
/* ===================================================================== */

/* Bytes of the library's static device scratch (one array per GPU the
 * library runs on, held until the process ends), and the part of it the
 * plan may give to j-split partials (a context given the same limit with
 * abyss_ctx_set_scratch_limit plans the same launch). */
#define ABYSS_NOALLOC_SCRATCH_BYTES (128ull << 20)
#define ABYSS_NOALLOC_SPLIT_BYTES (64ull << 20)

/* One regular-force + neighbour query with every buffer in device memory
 * owned by the caller, enqueued on `stream` (a hipStream_t; null = the null
 * stream) and not waited for; it allocates nothing.  Same results as
 * abyss_ctx_set_targets + abyss_ctx_launch on a context holding the same
 * j-set, with the same precision and scratch limit ABYSS_NOALLOC_SPLIT_BYTES
 * (checked precision: targets it recomputes agree to FP64 rounding, since
 * the recomputation here is planned for every target and reads the flagged
 * count on the device).
 *   jset: device SoA of n particles, 8 arrays of n doubles
 *         [x | y | z | vx | vy | vz | m | r_nb^2];
 *   targets: device int[nt], indices into the j-set (0 < nt <= n);
 *   acc, adot: device double[nt][3]; nnb: device int[nt];
 *   list: device int[nt][cap] (rows of cap entries, ascending j);
 *   precision: ABYSS_PREC_SPLIT or ABYSS_PREC_CHECKED (tolerance 1e-6).
 * A target outside [0, n) gets NaN forces and nnb = INT_MIN (the device
 * cannot report it without a host wait).  Returns ABYSS_ERR_SCRATCH_LIMIT
 * when n and nt need more than ABYSS_NOALLOC_SCRATCH_BYTES of scratch
 * (about 4 n + 20 nt bytes plus the recomputation's partials), before
 * anything is enqueued.  Every call on a device uses that device's one
 * scratch array: calls that overlap on one device (two streams, two host
 * threads) race and must be ordered by the caller. */
int abyss_query_noalloc(int device, int n, const double *jset, int nt, const int *targets,
                        int cap, int precision, double *acc, double *adot, int *nnb, int *list,
                        void *stream);

#ifdef __cplusplus
}
#endif

#endif /* ABYSS_REGF_H */
