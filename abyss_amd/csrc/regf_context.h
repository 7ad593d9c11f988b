// Internal: the per-cluster context behind abyss_ctx_t, shared by the
// single-cluster runtime (regf_runtime.hip) and the ensemble / serving
// runtime (regf_ensemble.hip).  Not part of the public ABI.
#pragma once
#include "../../include/abyss_regf.h"
#include "regf_internal.h"

struct abyss_ctx_s {
    int device = -1;
    hipStream_t stream = nullptr;
    bool own_stream = true;   // false: the stream belongs to an ensemble
    int cu_count = 0;
    int cap = ABYSS_NUM_NEIGHBOR_MAX;
    int precision = ABYSS_PREC_CHECKED;
    int force_ipl = 0, force_splits = 0, force_unroll = 0, force_jl = 0;
    int force_tile = 0, force_wpb = 0;   // lanes-per-target shape (0: automatic)
    bool batch_member = false;   // an ensemble's cluster: lane-group plans only
    unsigned long long scratch_limit = 2ull << 30;   // split-j scratch bound

    // j-set: device SoA [x|y|z|vx|vy|vz|m|rnb2] (stride n) + FP32 mass
    int n = 0;
    size_t j_bytes = 0;
    char *d_j = nullptr;
    char *h_j = nullptr;  // pinned staging, same layout
    abyss::JState js{};

    // targets and outputs
    int nt = 0;
    bool counts_valid = false;   // d_nnb holds the counts of the current targets
    size_t t_capacity = 0;  // targets the output buffers can hold
    int *d_tgt = nullptr, *h_tgt = nullptr;
    // h_tgt is mapped, coherent pinned memory; h_tgt_dev is its device
    // address.  `tgt` is the target array the kernels read: h_tgt_dev for
    // small blocks (no copy), d_tgt otherwise.
    const int *h_tgt_dev = nullptr, *tgt = nullptr;
    // per-target results (device block and pinned host mirror, same
    // layout): acc [nt][3] | adot [nt][3] | nnb [nt] | pad to 256 B | list
    // rows [nt][cap]; the pointers below are views of d_out / h_out for the
    // current nt (the host mirror holds rows only up to the one-copy size)
    char *d_out = nullptr, *h_out = nullptr;
    double *d_acc = nullptr, *d_adot = nullptr;
    int *d_nnb = nullptr, *d_list = nullptr;
    double *h_res = nullptr;  // acc then adot
    int *h_nnb = nullptr, *h_rows = nullptr;
    // h_out is mapped, coherent pinned memory; for small split launches the
    // finalize also writes the results there (device view `mirror`), and
    // mirror_valid tells fetch that no copy is needed
    char *h_out_dev = nullptr;
    abyss::FinalMirror mirror{};
    bool mirror_valid = false;
    int *h_list = nullptr;    // pinned, populated columns of larger results
    size_t h_list_bytes = 0;

    // device-resident particle state (regf_state.hip): SoA x,v,a,adot (3n
    // each), t, m, r2 (n each); staging for incremental updates
    double *d_state = nullptr, *h_state = nullptr;
    size_t state_bytes = 0;
    int state_n = 0;
    char *d_upd = nullptr, *h_upd = nullptr;
    size_t upd_bytes = 0;

    // resident neighbour lists (regf_irregular.hip): [nnb: n ints, padded
    // to 64][list: n rows of cap ints], row = particle index; filled after
    // each launch while keep_lists is on
    bool keep_lists = false;
    char *d_rlists = nullptr;
    size_t rlist_bytes = 0;
    int rlist_n = 0;
    char *d_irr = nullptr, *h_irr = nullptr;  // targets + [nt][6] results
    size_t irr_bytes = 0;
    // 64-byte AoS copy of the j-set {x,y,z,vx,vy,vz,m,0} for the list
    // gathers (one cache line per neighbour instead of seven); rebuilt when
    // jset_version (bumped by every upload / predict) has moved on
    char *d_aos = nullptr;
    size_t aos_bytes = 0;
    long long jset_version = 0, aos_version = -1;
    // start-up derivatives (regf_derivs.hip): records, split partials,
    // targets, results; pinned targets + results; radius results
    char *d_drv = nullptr, *h_drv = nullptr;
    size_t drv_bytes = 0, h_drv_bytes = 0;
    char *d_rad = nullptr;
    size_t rad_bytes = 0;

    // split-j scratch
    size_t scratch_bytes = 0;
    char *d_scratch = nullptr;
    abyss::QueryOut part{};

    // checked precision: device block of spread [t][2] | spread partials
    // [S][t][2] | counters [2] | flagged positions [t] | their j-indices [t];
    // the recomputation's split partials; pinned readback of the count
    char *d_chk = nullptr;
    size_t chk_bytes = 0;
    float *chk_spread = nullptr, *chk_pspread = nullptr;
    int *chk_counters = nullptr, *chk_fq = nullptr, *chk_fi = nullptr;
    char *d_rchk = nullptr;
    size_t rchk_bytes = 0;
    int *h_chk = nullptr;       // pinned, mapped: [0] flagged-count readback,
                                // [1] the count a device-read recomputation reports
    int *h_chk_dev = nullptr;   // device address of h_chk
    double check_tol = 1.0e-6;
    int n_rechecked = 0;        // targets recomputed by the last launch
    bool recheck_pending = false;   // n_rechecked is h_chk[1] once the stream drains
    bool spread_valid = false;  // chk_spread holds the last launch's spread

    abyss::Plan plan{};
    bool plan_valid = false;

    // counters (ProfileDevice); t_grav_ms is the legacy entry points'
    // wall-clock time in CalculateAccelerationOnDevice
    long long n_send = 0, n_grav = 0, sum_ni = 0;
    double interactions = 0.0, t_grav_ms = 0.0;

    // force-kernel timing (abyss_ctx_set_kernel_timing): an event pair
    // around the force kernel of every launch, in a ring that is read when
    // it fills and at sync
    static constexpr int kTimingRing = 64;
    bool ktiming = false;
    hipEvent_t kev[2 * kTimingRing] = {};
    int kev_used = 0;
    double kforce_ms = 0.0;
    long long klaunches = 0;

    char detail[512] = {0};
};


namespace abyss {
// Creates a context that issues all its work on `stream` (owned by the
// caller) instead of a stream of its own.
int ctx_create_on_stream(int device, int cap, hipStream_t stream, abyss_ctx_t *out);
// abyss_ctx_fetch with the device->host work issued on `st` (which must not
// race with a launch writing this context's outputs).
int ctx_fetch_on_stream(abyss_ctx_t c, hipStream_t st, double *acc, double *adot, int *nnb,
                        int *list);
// Device bytes a context holds (j-set, targets, outputs, scratch).
unsigned long long ctx_device_bytes(abyss_ctx_t c);
// abyss_ctx_launch without keeping the resident lists (force kernels only).
int ctx_launch_force(abyss_ctx_t c);
// After a launch: copies its neighbour rows into the resident per-particle
// lists when they are enabled (no-op otherwise).
int keep_launch_lists(abyss_ctx_t c);
// Checked precision: the spread / flag buffers for nt targets and S splits.
int ctx_ensure_check(abyss_ctx_t c, int nt, int S);
// abyss_ctx_set_targets with (sync) or without the wait for the context's
// stream; _idle: the caller guarantees that no launch using this context
// is in flight (an ensemble member whose last batch has finished).
int ctx_set_targets_sync(abyss_ctx_t c, int nt, const int *targets, bool sync);
int ctx_set_targets_idle(abyss_ctx_t c, int nt, const int *targets);
}  // namespace abyss
