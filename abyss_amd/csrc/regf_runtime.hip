// Runtime behind the C ABI (include/abyss_regf.h): device contexts, the
// device-resident j-set, pinned staging, launch planning, result transfer,
// input validation and the six legacy entry points.
//
// Reference counterpart: src_mpi/cuda/cuda_my_acceleration.cu (file-static
// state :23-88, _ReceiveFromHost :603-778, GetAcceleration :95-344, open /
// close / profile :987-1062, extern "C" wrappers :1087-1106).  Differences by
// design: validation happens before any HIP call and is reported through
// status codes; device buffers are sized from the actual N and targets
// (grow-only) instead of next-pow2 staging of GridDimY*NNB_per_block*N ints;
// small results (all list rows <= 1 MB) come back in one copy, larger ones
// move only the populated columns of the neighbour lists; small target lists
// are read by the kernels from mapped host memory.
#include "regf_context.h"
#include "regf_pack.h"

#include <limits.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>

using abyss::JState;
using abyss::Plan;
using abyss::QueryOut;

namespace {

// Below this many bytes of full-width list rows, fetch copies them with the
// other results in one round trip instead of first reading the widest row.
constexpr size_t kOneTripListBytes = 1u << 20;
// Up to this many targets the kernels read the target list straight from
// mapped host memory instead of a device copy (saves one copy per query).
constexpr int kMappedTargets = 64;
// a checked block with a mapped mirror runs in FP64 terms in one pass, so no
// recomputation ever has to update the mirror
static_assert(abyss::kCheckedDirectMax >= kMappedTargets, "mirrored blocks are direct");
// Most j-splits of the recomputation that reads its count on the device
// (scratch: this many rows per target of the block, 52 bytes each).
constexpr int kRecheckLoopSplits = 256;

int set_detail(abyss_ctx_t c, int code, const char *fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->detail, sizeof(c->detail), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(ctx, call)                                                            \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            return set_detail(ctx,                                                   \
                              e_ == hipErrorOutOfMemory ? ABYSS_ERR_OUT_OF_MEMORY    \
                                                        : ABYSS_ERR_HIP,             \
                              "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                              __FILE__, __LINE__);                                   \
        }                                                                            \
    } while (0)

template <class T>
void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}
template <class T>
void hfree(T *&p)
{
    if (p) (void)hipHostFree(p);
    p = nullptr;
}

size_t grow(size_t need) { return need + need / 2 + 4096; }

// Bytes of the j-set layout for n particles.
size_t jset_bytes(int n) { return (size_t)n * (8 * sizeof(double) + sizeof(float)); }

void bind_jstate(abyss_ctx_t c, char *base, int n)
{
    double *d = reinterpret_cast<double *>(base);
    c->js.x = d + 0 * (size_t)n;
    c->js.y = d + 1 * (size_t)n;
    c->js.z = d + 2 * (size_t)n;
    c->js.vx = d + 3 * (size_t)n;
    c->js.vy = d + 4 * (size_t)n;
    c->js.vz = d + 5 * (size_t)n;
    c->js.m = d + 6 * (size_t)n;
    c->js.rnb2 = d + 7 * (size_t)n;
    c->js.m32 = reinterpret_cast<const float *>(d + 8 * (size_t)n);
    c->js.n = n;
}

// Bytes of the packed per-target results of nt targets:
// acc [nt][3] | adot [nt][3] (FP64) | nnb [nt] (int32), padded to 256 B.
// The list rows [nt][cap] follow at that offset in the same device block.
size_t packed_bytes(size_t nt)
{
    const size_t b = nt * (6 * sizeof(double) + sizeof(int));
    return (b + 255) & ~(size_t)255;
}

// Views of the output block for the current target count: acc, adot, nnb,
// then the list rows, so that one copy can move every result of a small
// query.  The host block mirrors the layout up to kOneTripListBytes of rows.
void set_output_views(abyss_ctx_t c, int nt)
{
    const size_t n = (size_t)nt;
    c->d_acc = reinterpret_cast<double *>(c->d_out);
    c->d_adot = c->d_acc + 3 * n;
    c->d_nnb = reinterpret_cast<int *>(c->d_adot + 3 * n);
    c->d_list = reinterpret_cast<int *>(c->d_out + packed_bytes(n));
    c->h_res = reinterpret_cast<double *>(c->h_out);
    c->h_nnb = reinterpret_cast<int *>(c->h_res + 6 * n);
    c->h_rows = reinterpret_cast<int *>(c->h_out + packed_bytes(n));
    double *m = reinterpret_cast<double *>(c->h_out_dev);
    c->mirror.acc = m;
    c->mirror.adot = m + 3 * n;
    c->mirror.nnb = reinterpret_cast<int *>(m + 6 * n);
    c->mirror.list = reinterpret_cast<int *>(c->h_out_dev + packed_bytes(n));
}

int ensure_outputs(abyss_ctx_t c, int nt)
{
    if ((size_t)nt <= c->t_capacity) {
        set_output_views(c, nt);
        return ABYSS_OK;
    }
    size_t cap_t = grow((size_t)nt);
    dfree(c->d_tgt); dfree(c->d_out);
    hfree(c->h_tgt); hfree(c->h_out);
    c->t_capacity = 0;
    c->d_list = nullptr;
    c->h_tgt_dev = c->tgt = nullptr;
    c->h_out_dev = nullptr;
    c->mirror = abyss::FinalMirror{};
    c->mirror_valid = false;
    const size_t rows = cap_t * (size_t)c->cap * sizeof(int);
    HIPCHK(c, hipMalloc(&c->d_tgt, cap_t * sizeof(int)));
    HIPCHK(c, hipMalloc(&c->d_out, packed_bytes(cap_t) + rows));
    HIPCHK(c, hipHostMalloc(&c->h_tgt, cap_t * sizeof(int),
                            hipHostMallocMapped | hipHostMallocCoherent));
    void *dev = nullptr;
    HIPCHK(c, hipHostGetDevicePointer(&dev, c->h_tgt, 0));
    c->h_tgt_dev = static_cast<const int *>(dev);
    HIPCHK(c, hipHostMalloc(&c->h_out, packed_bytes(cap_t) + std::min(rows, kOneTripListBytes),
                            hipHostMallocMapped | hipHostMallocCoherent));
    void *out_dev = nullptr;
    HIPCHK(c, hipHostGetDevicePointer(&out_dev, c->h_out, 0));
    c->h_out_dev = static_cast<char *>(out_dev);
    c->t_capacity = cap_t;
    set_output_views(c, nt);
    return ABYSS_OK;
}

// Plan for nt targets against n j under precision `prec`; list rows of
// `cap` ints bound the split count through the scratch limit.  `forced`:
// honour the context's forced tuning and the tuned table (the context's own
// queries), or not (the recomputation of checked precision).
Plan plan_for(abyss_ctx_t c, int nt, int n, int prec, int cap, bool forced)
{
    Plan p;
    p.tile = abyss::kTile;
    const int cu = c->cu_count > 0 ? c->cu_count : 256;
    // Lanes per target: 4 lanes share a target from 256 targets up (its
    // inner loop is one VALU op per pair shorter than one target per lane
    // and it holds 5 waves per SIMD instead of 4: 47.8% vs 45.4% of FP32
    // peak at N = m = 262144); smaller blocks spread a target over 16 lanes
    // (16..255 targets) or 64 (below 16), measured best at N = 262144.
    abyss::TuneEntry te{0, 0, 16};
    // a small checked block runs in FP64 terms in one pass (term mode 3,
    // the recomputation's arithmetic, with lists): Plummer + Kroupa
    // N = 65536, m = 1 / 64 / 256: 15.7 -> 12.7, 25.2 -> 21.1, 54.4 -> 35.3
    // us; King N = 262144, m = 16 / 256: 42.2 -> 23.9, 111.6 -> 102.6 us;
    // m = 1024 is a tie or slower
    const bool direct = prec == ABYSS_PREC_CHECKED && nt <= abyss::kCheckedDirectMax &&
                        !c->batch_member;
    p.fp64_terms = direct;
    const bool mode0 = prec == ABYSS_PREC_SPLIT || (prec == ABYSS_PREC_CHECKED && !direct);
    const bool tuned = forced && mode0 && abyss::tuned_plan(c->device, nt, &te);
    const int force_jl = forced ? c->force_jl : 0;
    const int force_splits = forced ? c->force_splits : 0;
    // Workgroup shape of the lanes-per-target kernels (term mode 0 only; a
    // batch member runs in the 4-wave workgroups of the persistent batch
    // kernel): forced, else tuned, else 256 j per tile and 4 waves (the
    // one-shot tile for 64 lanes per target, below).
    int tile = abyss::kTile, wpb = 4;
    if (mode0) {
        if (forced && c->force_tile) tile = c->force_tile;
        else if (tuned) tile = te.tile;
        if (forced && c->force_wpb) wpb = c->force_wpb;
        else if (tuned) wpb = te.wpb;
        if (c->batch_member) wpb = 4;
        if (tile == abyss::kTileOneShot) {
            // the one-shot tile's only workgroup is 4 waves; a forced other
            // block wins over a tuned one-shot tile
            if (wpb != 4 && forced && c->force_wpb && !c->force_tile) tile = 64 * wpb;
            else wpb = 4;
        } else if (tile > 64 * wpb) {
            // a forced tile wins over a tuned or default block, and v.v.
            if (forced && c->force_wpb && !c->force_tile) tile = 64 * wpb;
            else wpb = tile / 64;
        }
    }
    int jl = force_jl;
    if (jl == 0) {
        if (tuned) jl = te.jl;
        else if (prec == ABYSS_PREC_FP64) jl = 1;
        else if (nt >= 256) jl = 4;
        else if (nt >= 16) jl = 16;
        else jl = 64;
    }
    // A lane group wide enough to fill the chip: the j-split count is capped
    // (>= 4 tiles per split, kMaxSplits, scratch), so on a small j-set the
    // 4-lane kernel may get too few workgroups.  Move to 16 lanes below 4
    // workgroups per CU, and from 16 to 64 below one per CU (Plummer +
    // Kroupa N = 65536: m = 256 53.3 -> 35.7 us, m = 16 26.0 -> 19.9 us;
    // King N = 262144 keeps its choices).
    if (jl > 1 && force_jl == 0 && !tuned && force_splits == 0 && !c->batch_member &&
        prec != ABYSS_PREC_FP64) {
        int s_max = std::min(std::max(1, n / (4 * abyss::kTile)), abyss::kMaxSplits);
        const double per = (double)nt * ((double)cap * 4.0 + 6 * 8.0 + 4.0);
        s_max = std::min(s_max, (int)std::max(1.0, std::floor((double)c->scratch_limit / per)));
        auto groups = [&](int l) {
            const int tpb = wpb * 64 / l;
            return (double)((nt + tpb - 1) / tpb) * s_max;
        };
        if (jl == 4 && groups(4) < 4.0 * cu) jl = 16;
        if (jl == 16 && groups(16) < 1.0 * cu) jl = 64;
    }
    if (prec == ABYSS_PREC_FP64) jl = 1;
    // the FP64-term modes exist only in the lanes-per-target kernel
    if ((prec == ABYSS_PREC_MIXED || prec == ABYSS_PREC_REFINED ||
         prec == ABYSS_PREC_CHECKED || c->batch_member) && jl == 1)
        jl = 4;
    // One target per lane (ipl = 1, unroll 2) measured fastest at every
    // block size on MI355X; ipl 2/4 keep more targets in registers but cut
    // occupancy to 3 and 1 waves per SIMD.
    int ipl = forced && c->force_ipl ? c->force_ipl : 1;
    int unroll = forced && c->force_unroll ? c->force_unroll : (tuned ? te.unroll : 2);
    if (prec == ABYSS_PREC_FP64) {
        ipl = std::min(ipl, 2);
        unroll = 2;
    }
    if (jl > 1) ipl = 1;
    if (jl == 1 || !mode0) {
        tile = abyss::kTile;
        wpb = 4;
    }
    // A checked block in FP64 terms in one pass (term mode 3) takes the
    // one-shot tile too: when forced, or built in for 64 lanes per target
    // (below 16 targets), as split precision does.
    if (direct && !c->batch_member) {
        const bool forced_shape = forced && (c->force_tile || c->force_wpb);
        if (forced_shape ? c->force_tile == abyss::kTileOneShot &&
                               (c->force_wpb == 0 || c->force_wpb == 4)
                         : jl == 64) {
            tile = abyss::kTileOneShot;
            wpb = 4;
        }
    }
    // Built-in shape of blocks below 16 targets (64 lanes per target): the
    // one-shot tile, whose splits load their j at once (Plummer + Kroupa
    // N = 65536, m = 1..16: 12.8 -> 12.1 ... 14.3 -> 13.3 us per split
    // launch; 16 lanes, m = 64..256, gain nothing).  Kept only when the
    // splits fit it (below).
    if (mode0 && jl == 64 && !tuned && !c->batch_member &&
        !(forced && (c->force_tile || c->force_wpb))) {
        tile = abyss::kTileOneShot;
        wpb = 4;
    }
    p.tile = tile;
    p.wpb = wpb;
    p.jl = jl;
    p.ipl = ipl;
    p.unroll = unroll;
    const int tpb = abyss::targets_per_block(p);
    p.grid_x = (nt + tpb - 1) / tpb;
    int S = force_splits;
    if (S == 0) {
        // Workgroups per CU over the whole launch (or the tuned count), which
        // sets the tail of the last round of workgroups: for the 1- and
        // 4-lane kernels 96 from 65536 targets up and 48 below (at N = 256K:
        // m = 65536..262144 run 1.0-1.3 % faster with 96 than with 48,
        // m = 4096 and 16384 best at 48), 16 for the 16/64-lane ones.
        const int wg_rule = jl <= 4 ? (nt >= 65536 ? 96 : 48) : 16;
        const int want_wg = cu * (tuned ? te.wg_per_cu : wg_rule);
        S = (want_wg + p.grid_x - 1) / p.grid_x;
        // every split keeps at least 4 tiles (1024 j): finer splits cost
        // more in finalize than they gain (m = 1 at N = 262144: S = 256
        // takes 21.7 us, S = 1024 31.1 us)
        const int max_by_n = std::max(1, n / (4 * abyss::kTile));
        S = std::min(S, std::min(max_by_n, abyss::kMaxSplits));
        // large blocks keep longer spans: where splits of 16 tiles (4096 j)
        // still give 8 workgroups per CU, no more splits than that
        // (Plummer + Kroupa N = 65536, m = 16384 / 65536: S = 43 -> 16 and
        // 24 -> 16, 936 -> 910 and 3640 -> 3519 us per split launch; at
        // N = 262144 the rule above already gives fewer)
        const int s_long = std::max(1, n / (16 * abyss::kTile));
        if (!tuned && !c->batch_member && (long long)p.grid_x * s_long >= 8LL * cu)
            S = std::min(S, s_long);
    }
    if (S > 1) {
        const double per = (double)nt * ((double)cap * 4.0 + 6 * 8.0 + 4.0);
        const int max_by_mem = (int)std::max(1.0, std::floor((double)c->scratch_limit / per));
        S = std::min(S, max_by_mem);
    }
    S = std::max(1, std::min(S, n));
    // spans are whole kTile blocks whatever the tile (every tile divides
    // kTile), so the tile changes no fold point of a lane's FP32 runs
    int jspan = (n + S - 1) / S;
    jspan = ((jspan + abyss::kTile - 1) / abyss::kTile) * abyss::kTile;
    p.jspan = jspan;
    p.splits = (n + jspan - 1) / jspan;
    // the one-shot tile holds a whole split, with 16 or 64 lanes per target
    // (4 lanes would fold their FP32 runs inside it), outside batches
    if (p.tile == abyss::kTileOneShot &&
        (p.jl < 16 || p.jspan > abyss::kTileOneShot || c->batch_member))
        p.tile = abyss::kTile;
    return p;
}

Plan make_plan(abyss_ctx_t c, int nt, int n)
{
    return plan_for(c, nt, n, c->precision, c->cap, true);
}

int ensure_scratch(abyss_ctx_t c, const Plan &p, int nt)
{
    if (p.splits <= 1) return ABYSS_OK;
    const size_t S = (size_t)p.splits;
    const size_t b_acc = S * nt * 3 * sizeof(double);
    const size_t b_cnt = S * nt * sizeof(int);
    const size_t b_list = S * nt * (size_t)c->cap * sizeof(int);
    const size_t need = 2 * b_acc + b_cnt + b_list + 4 * 256;
    if (need > c->scratch_bytes) {
        // a previous launch may still read the old scratch
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c->d_scratch);
        c->scratch_bytes = 0;
        HIPCHK(c, hipMalloc(&c->d_scratch, need));
        c->scratch_bytes = need;
    }
    char *b = c->d_scratch;
    auto take = [&](size_t bytes) {
        char *r = b;
        b += (bytes + 255) & ~size_t(255);
        return r;
    };
    c->part.acc = reinterpret_cast<double *>(take(b_acc));
    c->part.adot = reinterpret_cast<double *>(take(b_acc));
    c->part.cnt = reinterpret_cast<int *>(take(b_cnt));
    c->part.list = reinterpret_cast<int *>(take(b_list));
    return ABYSS_OK;
}

// Checked precision: the spread arrays, counters and flagged-target arrays
// for nt targets and S splits.
int ensure_check(abyss_ctx_t c, int nt, int S)
{
    auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
    const size_t t = (size_t)nt;
    const size_t b_sp = al(2 * t * sizeof(float));
    const size_t b_ps = al(2 * (size_t)S * t * sizeof(float));
    const size_t b_ix = al(t * sizeof(int));
    const size_t need = b_sp + b_ps + 256 + 2 * b_ix;
    if (need > c->chk_bytes) {
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c->d_chk);
        c->chk_bytes = 0;
        HIPCHK(c, hipMalloc(&c->d_chk, grow(need)));
        c->chk_bytes = grow(need);
    }
    if (!c->h_chk) {
        HIPCHK(c, hipHostMalloc(&c->h_chk, 256, hipHostMallocMapped | hipHostMallocCoherent));
        void *dev = nullptr;
        HIPCHK(c, hipHostGetDevicePointer(&dev, c->h_chk, 0));
        c->h_chk_dev = static_cast<int *>(dev);
    }
    char *b = c->d_chk;
    c->chk_spread = reinterpret_cast<float *>(b);
    b += b_sp;
    c->chk_pspread = reinterpret_cast<float *>(b);
    b += b_ps;
    c->chk_counters = reinterpret_cast<int *>(b);
    b += 256;
    c->chk_fq = reinterpret_cast<int *>(b);
    b += b_ix;
    c->chk_fi = reinterpret_cast<int *>(b);
    return ABYSS_OK;
}

// Device buffer of the recomputation's split partials for `rows` (split,
// target) rows.
int ensure_recheck_scratch(abyss_ctx_t c, size_t rows, QueryOut *rpart)
{
    const size_t b_acc = (rows * 3 * sizeof(double) + 255) & ~size_t(255);
    const size_t b_cnt = (rows * sizeof(int) + 255) & ~size_t(255);
    const size_t need = 2 * b_acc + b_cnt;
    if (need > c->rchk_bytes) {
        // an earlier launch's recomputation may still read it
        HIPCHK(c, hipStreamSynchronize(c->stream));
        dfree(c->d_rchk);
        c->rchk_bytes = 0;
        HIPCHK(c, hipMalloc(&c->d_rchk, grow(need)));
        c->rchk_bytes = grow(need);
    }
    *rpart = QueryOut{};
    rpart->acc = reinterpret_cast<double *>(c->d_rchk);
    rpart->adot = reinterpret_cast<double *>(c->d_rchk + b_acc);
    rpart->cnt = reinterpret_cast<int *>(c->d_rchk + 2 * b_acc);
    rpart->list = nullptr;   // cap 0: nothing is listed
    return ABYSS_OK;
}

// The recomputation of m flagged targets (c->chk_fq / chk_fi hold them),
// planned for m; the stream need not be idle.
int recompute_flagged(abyss_ctx_t c, int m)
{
    c->n_rechecked = m;
    if (m <= 0) return ABYSS_OK;
    if (m > c->nt) return set_detail(c, ABYSS_ERR_HIP, "check: %d flagged of %d", m, c->nt);
    const Plan rp = plan_for(c, m, c->n, ABYSS_PREC_REFINED, 0, false);
    QueryOut rpart;
    int rc = ensure_recheck_scratch(c, (size_t)rp.splits * m, &rpart);
    if (rc != ABYSS_OK) return rc;
    HIPCHK(c, abyss::launch_regf(c->js, c->chk_fi, m, rp, rpart, 0, abyss::kPrecRecheck,
                                 c->stream));
    HIPCHK(c, abyss::launch_recheck_finalize(rpart, rp.splits, m, c->chk_fq, c->d_acc, c->d_adot,
                                             c->stream));
    return ABYSS_OK;
}

// The checked precision's k2 = (ABYSS_CHECK_ERR_SCALE / tol)^2.
double check_k2(abyss_ctx_t c)
{
    const double ratio = ABYSS_CHECK_ERR_SCALE / c->check_tol;
    return ratio * ratio;
}

// After a checked launch's force and finalize: flags the targets whose a or
// jdot may miss the tolerance (unless the finalize of a split launch has
// flagged them: `flagged`) and recomputes them with FP64 terms over the
// same regular set: the flagged count is read back (a host wait) and the
// recomputation planned for it.  (Blocks small enough to skip that wait run
// in FP64 terms in one pass instead, Plan::fp64_terms; a finalize mirror is
// written only for them, so the recomputed targets never have to update it.)
int recheck(abyss_ctx_t c, bool flagged)
{
    if (!flagged) {
        // one-split launches: a flag pass over the final results (the force
        // kernel zeroed the counter)
        HIPCHK(c, abyss::launch_check_flag(c->d_acc, c->d_adot, c->chk_spread, c->tgt, c->nt,
                                           check_k2(c), c->chk_counters, c->chk_fq, c->chk_fi,
                                           c->stream));
    }
    if (c->nt <= abyss::kRecheckOnDeviceMax && !c->batch_member) {
        // The count stays on the device: lanes per target planned for a
        // typical flagged count (1/16 of the block), a grid-stride loop over
        // the actual one.  Its items are short (j-ranges of 1 tile where N
        // allows, up to 256 splits), since a few flagged targets give few
        // items and each item's tiles run one after another.
        Plan rp = plan_for(c, std::max(1, c->nt / 16), c->n, ABYSS_PREC_REFINED, 0, false);
        const int tiles = (c->n + abyss::kTile - 1) / abyss::kTile;
        rp.splits = std::min(tiles, kRecheckLoopSplits);
        rp.jspan = (tiles + rp.splits - 1) / rp.splits * abyss::kTile;
        rp.splits = (c->n + rp.jspan - 1) / rp.jspan;
        const int tpb = 4 * (64 / rp.jl);
        const int cu = c->cu_count > 0 ? c->cu_count : 256;
        const long long items = (long long)((c->nt + tpb - 1) / tpb) * rp.splits;
        const int grid = (int)std::min<long long>(items, 5LL * cu);
        QueryOut rpart;
        int rc = ensure_recheck_scratch(c, (size_t)rp.splits * c->nt, &rpart);
        if (rc != ABYSS_OK) return rc;
        HIPCHK(c, abyss::launch_recheck_loop(c->js, c->chk_fi, c->chk_counters, c->nt, rp, grid,
                                             rpart, c->chk_fq, c->d_acc, c->d_adot,
                                             c->h_chk_dev + 1, c->stream));
        c->recheck_pending = true;
        return ABYSS_OK;
    }
    HIPCHK(c, hipMemcpyAsync(c->h_chk, c->chk_counters, sizeof(int), hipMemcpyDeviceToHost,
                             c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return recompute_flagged(c, c->h_chk[0]);
}

// Force-kernel timing: adds up the event pairs recorded so far (waits for
// the last of them) and empties the ring.
int collect_kernel_timing(abyss_ctx_t c)
{
    for (int k = 0; k < c->kev_used; ++k) {
        float ms = 0.0f;
        HIPCHK(c, hipEventSynchronize(c->kev[2 * k + 1]));
        HIPCHK(c, hipEventElapsedTime(&ms, c->kev[2 * k], c->kev[2 * k + 1]));
        c->kforce_ms += ms;
        c->klaunches += 1;
    }
    c->kev_used = 0;
    return ABYSS_OK;
}

void release_all(abyss_ctx_t c)
{
    if (c->device >= 0) (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->d_j); hfree(c->h_j);
    dfree(c->d_tgt); dfree(c->d_out);
    c->d_list = nullptr;
    c->h_tgt_dev = c->tgt = nullptr;
    dfree(c->d_scratch); dfree(c->d_state); dfree(c->d_upd);
    dfree(c->d_chk); dfree(c->d_rchk); hfree(c->h_chk);
    dfree(c->d_rlists); dfree(c->d_irr); dfree(c->d_aos); dfree(c->d_drv); dfree(c->d_rad);
    hfree(c->h_drv);
    hfree(c->h_state); hfree(c->h_upd); hfree(c->h_irr);
    hfree(c->h_tgt); hfree(c->h_out); hfree(c->h_list);
    for (hipEvent_t &e : c->kev)
        if (e) {
            (void)hipEventDestroy(e);
            e = nullptr;
        }
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// Context API
// ---------------------------------------------------------------------------
extern "C" {

int abyss_device_count(void)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

const char *abyss_error_string(int code)
{
    switch (code) {
        case ABYSS_OK: return "ok";
        case ABYSS_ERR_NOT_OPEN: return "device not open";
        case ABYSS_ERR_NO_DEVICE: return "no HIP device";
        case ABYSS_ERR_INVALID_ARG: return "invalid argument";
        case ABYSS_ERR_TARGET_RANGE: return "target index out of range";
        case ABYSS_ERR_INVALID_INPUT: return "invalid particle data";
        case ABYSS_ERR_NO_JSET: return "no j-particles uploaded";
        case ABYSS_ERR_HIP: return "HIP runtime error";
        case ABYSS_ERR_OUT_OF_MEMORY: return "out of device memory";
        case ABYSS_ERR_NEIGHBOR_OVERFLOW: return "neighbour list overflow";
        case ABYSS_ERR_ADMISSION: return "admission refused (memory budget)";
        case ABYSS_ERR_BUSY: return "busy: no free slot, or request turned away by admission";
        case ABYSS_ERR_SCRATCH_LIMIT:
            return "no-allocation query: sizes exceed the static scratch";
        default: return "unknown error";
    }
}

}  // extern "C"

namespace abyss {

int ctx_create_on_stream(int device, int cap, hipStream_t stream, abyss_ctx_t *out)
{
    if (!out || cap <= 0) return ABYSS_ERR_INVALID_ARG;
    *out = nullptr;
    const int ndev = abyss_device_count();
    if (ndev <= 0) return ABYSS_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = new (std::nothrow) abyss_ctx_s();
    if (!c) return ABYSS_ERR_OUT_OF_MEMORY;
    c->device = device;
    c->cap = cap;
    int rc = ABYSS_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = ABYSS_ERR_NO_DEVICE; break; }
        if (stream) {
            c->stream = stream;
            c->own_stream = false;
        } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = ABYSS_ERR_HIP; break;
        }
        if (hipDeviceGetAttribute(&c->cu_count, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess) { rc = ABYSS_ERR_HIP; break; }
    } while (0);
    if (rc != ABYSS_OK) {
        release_all(c);
        delete c;
        return rc;
    }
    *out = c;
    return ABYSS_OK;
}

unsigned long long ctx_device_bytes(abyss_ctx_t c)
{
    if (!c) return 0;
    return (unsigned long long)c->j_bytes + c->scratch_bytes + c->state_bytes + c->upd_bytes +
           c->rlist_bytes + c->irr_bytes + c->aos_bytes + c->drv_bytes + c->rad_bytes +
           (unsigned long long)c->t_capacity * (sizeof(int) * 2 + 6 * sizeof(double) +
                                                (size_t)c->cap * sizeof(int)) + 256;
}

}  // namespace abyss

extern "C" {

int abyss_ctx_create(int device, int cap, abyss_ctx_t *out)
{
    return abyss::ctx_create_on_stream(device, cap, nullptr, out);
}

int abyss_ctx_destroy(abyss_ctx_t c)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    release_all(c);
    delete c;
    return ABYSS_OK;
}

int abyss_ctx_set_precision(abyss_ctx_t c, int mode)
{
    if (!c || mode < ABYSS_PREC_SPLIT || mode > ABYSS_PREC_CHECKED) return ABYSS_ERR_INVALID_ARG;
    c->precision = mode;
    c->plan_valid = false;
    return ABYSS_OK;
}

int abyss_ctx_set_tuning(abyss_ctx_t c, int ipl, int splits, int unroll, int jl)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (ipl != 0 && ipl != 1 && ipl != 2 && ipl != 4)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "i_per_lane must be 0, 1, 2 or 4");
    if (unroll != 0 && unroll != 2 && unroll != 4)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "j_unroll must be 0, 2 or 4");
    if (jl != 0 && jl != 1 && jl != 4 && jl != 16 && jl != 64)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "lanes_per_target must be 0, 1, 4, 16 or 64");
    if (splits < 0 || splits > abyss::kMaxSplits)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "j_splits must be in [0, %d]",
                          abyss::kMaxSplits);
    c->force_ipl = ipl;
    c->force_splits = splits;
    c->force_unroll = unroll;
    c->force_jl = jl;
    c->plan_valid = false;
    return ABYSS_OK;
}

int abyss_ctx_set_shape(abyss_ctx_t c, int tile, int waves_per_block)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (tile != 0 && tile != 64 && tile != 128 && tile != 256 && tile != abyss::kTileOneShot)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "tile must be 0, 64, 128, 256 or 1024");
    if (waves_per_block != 0 && waves_per_block != 1 && waves_per_block != 2 &&
        waves_per_block != 4)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "waves_per_block must be 0, 1, 2 or 4");
    if (tile != 0 && waves_per_block != 0 && !abyss::shape_supported(tile, waves_per_block))
        return set_detail(c, ABYSS_ERR_INVALID_ARG,
                          "a tile of %d j needs at least %d waves per block", tile, tile / 64);
    c->force_tile = tile;
    c->force_wpb = waves_per_block;
    c->plan_valid = false;
    return ABYSS_OK;
}

int abyss_ctx_set_scratch_limit(abyss_ctx_t c, unsigned long long bytes)
{
    if (!c || bytes == 0) return ABYSS_ERR_INVALID_ARG;
    c->scratch_limit = bytes;
    c->plan_valid = false;
    return ABYSS_OK;
}

const char *abyss_ctx_error_detail(abyss_ctx_t c) { return c ? c->detail : ""; }

int abyss_ctx_set_check_tolerance(abyss_ctx_t c, double tol)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!(tol > 0.0) || !isfinite(tol))
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "check tolerance must be finite and > 0");
    c->check_tol = tol;
    return ABYSS_OK;
}

int abyss_ctx_last_recheck(abyss_ctx_t c, int *count)
{
    if (!c || !count) return ABYSS_ERR_INVALID_ARG;
    if (c->recheck_pending) {
        // the recomputation read its count on the device and reported it
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->n_rechecked = c->h_chk[1];
        c->recheck_pending = false;
    }
    *count = c->n_rechecked;
    return ABYSS_OK;
}

int abyss_ctx_fetch_spread(abyss_ctx_t c, float *spread)
{
    if (!c || !spread) return ABYSS_ERR_INVALID_ARG;
    if (!c->spread_valid)
        return set_detail(c, ABYSS_ERR_INVALID_ARG,
                          "no spread to read: the last launch was not a checked FP32 pass "
                          "(checked blocks of at most %d targets run in FP64 terms)",
                          abyss::kCheckedDirectMax);
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(spread, c->chk_spread, 2 * (size_t)c->nt * sizeof(float),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < 2 * c->nt; ++k) spread[k] = sqrtf(spread[k]);
    return ABYSS_OK;
}

void *abyss_ctx_stream(abyss_ctx_t c) { return c ? (void *)c->stream : nullptr; }

int abyss_ctx_upload(abyss_ctx_t c, int n, const double *m, const double *x, const double *v,
                     const double *r2)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    // A rejected upload leaves no j-set: queries fail with ABYSS_ERR_NO_JSET
    // until an upload succeeds, instead of silently using the previous one.
    const int old_n = c->n, old_nt = c->nt;
    const bool old_plan = c->plan_valid;
    c->n = 0;
    c->nt = 0;
    c->plan_valid = false;
    c->state_n = 0;
    c->jset_version += 1;
    if (n <= 0 || !m || !x || !v || !r2)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "upload: n=%d or null array", n);
    HIPCHK(c, hipSetDevice(c->device));
    // The previous upload / query may still read the staging buffer.
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const size_t need = jset_bytes(n);
    if (need > c->j_bytes) {
        dfree(c->d_j);
        hfree(c->h_j);
        c->j_bytes = 0;
        const size_t cap_b = grow(need);
        HIPCHK(c, hipMalloc(&c->d_j, cap_b));
        HIPCHK(c, hipHostMalloc(&c->h_j, cap_b, hipHostMallocDefault));
        c->j_bytes = cap_b;
    }
    // Validation and AoS -> SoA pack of the caller's arrays, on a few host
    // threads for large N, before anything of them is used
    // (src/cuda/cuda_my_acceleration.cu:399-405 packs the same arrays with
    // stride N).  The positions are packed first; their three arrays are
    // copied while the other threads pack the rest, then the rest is copied
    // (both copies contiguous).  Packing and copying in particle chunks was
    // slower: the strided chunk copies run as blit kernels (192 vs 168 us
    // for a 65536-particle upload + query).  After a failure an enqueued
    // position copy is harmless: the j-set is dropped, and the next upload
    // waits for the stream before it writes the staging buffer.
    double *h = reinterpret_cast<double *>(c->h_j);
    float *h32 = reinterpret_cast<float *>(h + 8 * (size_t)n);
    const size_t pos_bytes = 3 * (size_t)n * sizeof(double);
    hipError_t pos_copy = hipSuccess;
    const abyss::PackStatus st = abyss::pack_jset_staged(n, m, x, v, r2, h, h32, [&] {
        pos_copy = hipMemcpyAsync(c->d_j, c->h_j, pos_bytes, hipMemcpyHostToDevice, c->stream);
    });
    if (st.index >= 0) {
        const int j = st.index;
        if (st.kind == abyss::kPackBadState)
            return set_detail(c, ABYSS_ERR_INVALID_INPUT,
                              "upload: non-finite position/velocity at j=%d", j);
        if (st.kind == abyss::kPackBadMass)
            return set_detail(c, ABYSS_ERR_INVALID_INPUT, "upload: bad mass %g at j=%d", m[j],
                              j);
        return set_detail(c, ABYSS_ERR_INVALID_INPUT, "upload: bad r_nb^2 %g at j=%d", r2[j], j);
    }
    HIPCHK(c, pos_copy);
    HIPCHK(c, hipMemcpyAsync(c->d_j + pos_bytes, c->h_j + pos_bytes, need - pos_bytes,
                             hipMemcpyHostToDevice, c->stream));
    bind_jstate(c, c->d_j, n);
    c->jset_version += 1;
    c->n = n;
    if (old_n == n) {
        // the same size right after a good j-set: its targets stay set
        c->nt = old_nt;
        c->plan_valid = old_plan;
    }
    c->n_send += 1;
    return ABYSS_OK;
}

int abyss_ctx_set_targets(abyss_ctx_t c, int nt, const int *targets)
{
    return abyss::ctx_set_targets_sync(c, nt, targets, true);
}

}  // extern "C"

namespace abyss {

int ctx_set_targets_idle(abyss_ctx_t c, int nt, const int *targets)
{
    return ctx_set_targets_sync(c, nt, targets, false);
}

// `sync`: wait for the context's stream first (the previous launch may
// still read the target list and outputs); without it the caller
// guarantees that no launch using this context is in flight.
int ctx_set_targets_sync(abyss_ctx_t c, int nt, const int *targets, bool sync)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->n <= 0) return set_detail(c, ABYSS_ERR_NO_JSET, "no j-set uploaded");
    if (!targets || nt <= 0 || nt > c->n)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "targets: need 0 < NumTarget=%d <= N=%d", nt,
                          c->n);
    for (int k = 0; k < nt; ++k)
        if (targets[k] < 0 || targets[k] >= c->n)
            return set_detail(c, ABYSS_ERR_TARGET_RANGE, "target[%d]=%d outside [0,%d)", k,
                              targets[k], c->n);
    HIPCHK(c, hipSetDevice(c->device));
    if (sync) HIPCHK(c, hipStreamSynchronize(c->stream));
    int rc = ensure_outputs(c, nt);
    if (rc != ABYSS_OK) return rc;
    memcpy(c->h_tgt, targets, (size_t)nt * sizeof(int));
    if (nt <= kMappedTargets && !c->batch_member) {
        // small block: kernels read the (coherent) pinned copy directly; no
        // launch still reads the previous list
        c->tgt = c->h_tgt_dev;
    } else {
        HIPCHK(c, hipMemcpyAsync(c->d_tgt, c->h_tgt, (size_t)nt * sizeof(int),
                                 hipMemcpyHostToDevice, c->stream));
        c->tgt = c->d_tgt;
    }
    c->nt = nt;
    c->counts_valid = false;   // counts belong to the next launch
    c->mirror_valid = false;
    c->spread_valid = false;
    c->plan = make_plan(c, nt, c->n);
    rc = ensure_scratch(c, c->plan, nt);
    if (rc != ABYSS_OK) return rc;
    c->plan_valid = true;
    return ABYSS_OK;
}

}  // namespace abyss

extern "C" {

int abyss_ctx_plan(abyss_ctx_t c, int *ipl, int *splits, int *grid_x, int *tile, int *unroll,
                   int *jl)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!c->plan_valid) return set_detail(c, ABYSS_ERR_INVALID_ARG, "no targets set");
    if (ipl) *ipl = c->plan.ipl;
    if (splits) *splits = c->plan.splits;
    if (grid_x) *grid_x = c->plan.grid_x;
    if (tile) *tile = c->plan.tile;
    if (unroll) *unroll = c->plan.unroll;
    if (jl) *jl = c->plan.jl;
    return ABYSS_OK;
}

int abyss_ctx_plan_block(abyss_ctx_t c, int *waves_per_block, int *targets_per_block)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!c->plan_valid) return set_detail(c, ABYSS_ERR_INVALID_ARG, "no targets set");
    if (waves_per_block) *waves_per_block = c->plan.jl > 1 ? c->plan.wpb : abyss::kBlock / 64;
    if (targets_per_block) *targets_per_block = abyss::targets_per_block(c->plan);
    return ABYSS_OK;
}

}  // extern "C"

namespace abyss {

int ctx_ensure_check(abyss_ctx_t c, int nt, int S) { return ensure_check(c, nt, S); }

int ctx_launch_force(abyss_ctx_t c)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->n <= 0) return set_detail(c, ABYSS_ERR_NO_JSET, "no j-set uploaded");
    if (c->nt <= 0) return set_detail(c, ABYSS_ERR_INVALID_ARG, "no targets set");
    HIPCHK(c, hipSetDevice(c->device));
    if (!c->plan_valid) {
        c->plan = make_plan(c, c->nt, c->n);
        int rc = ensure_scratch(c, c->plan, c->nt);
        if (rc != ABYSS_OK) return rc;
        c->plan_valid = true;
    }
    const Plan &p = c->plan;
    // checked precision: the FP32 pass with its flag and recomputation, or
    // for a small block every target in FP64 terms (p.fp64_terms)
    const bool direct = c->precision == ABYSS_PREC_CHECKED && p.fp64_terms;
    const bool checked = c->precision == ABYSS_PREC_CHECKED && !direct;
    c->spread_valid = false;
    c->n_rechecked = direct ? c->nt : 0;
    c->recheck_pending = false;
    if (checked) {
        int rc = ensure_check(c, c->nt, p.splits);
        if (rc != ABYSS_OK) return rc;
    }
    QueryOut fin{c->d_acc, c->d_adot, c->d_nnb, c->d_list};
    QueryOut part = c->part;
    if (checked) {
        fin.spread = c->chk_spread;
        part.spread = c->chk_pspread;
    }
    c->mirror_valid = false;
    if (c->ktiming) {
        if (c->kev_used == abyss_ctx_s::kTimingRing) {
            int rc = collect_kernel_timing(c);
            if (rc != ABYSS_OK) return rc;
        }
        HIPCHK(c, hipEventRecord(c->kev[2 * c->kev_used], c->stream));
    }
    // checked precision: the force kernel zeroes the flagged-target counter
    // and a split launch's finalize flags the targets (no launches of their
    // own)
    HIPCHK(c, abyss::launch_regf(c->js, c->tgt, c->nt, p, p.splits == 1 ? fin : part, c->cap,
                                 direct ? abyss::kPrecRecheck : c->precision, c->stream,
                                 checked ? c->chk_counters : nullptr));
    if (c->ktiming) {
        HIPCHK(c, hipEventRecord(c->kev[2 * c->kev_used + 1], c->stream));
        c->kev_used += 1;
    }
    if (p.splits > 1) {
        // small results also go straight into the mapped host block, so a
        // following fetch needs no copy (a checked recomputation updates it
        // too)
        const bool mirror = c->nt <= kMappedTargets && !c->batch_member &&
                            (size_t)c->nt * c->cap * sizeof(int) <= kOneTripListBytes;
        abyss::CheckFlag flag;
        if (checked) {
            flag.k2 = check_k2(c);
            flag.counter = c->chk_counters;
            flag.fq = c->chk_fq;
            flag.fi = c->chk_fi;
            flag.tgt = c->tgt;
        }
        HIPCHK(c, abyss::launch_finalize(part, p.splits, c->nt, c->cap, c->d_acc, c->d_adot,
                                         c->d_nnb, c->d_list,
                                         mirror ? c->mirror : abyss::FinalMirror{},
                                         fin.spread, c->stream, flag));
        c->mirror_valid = mirror;
    }
    if (checked) {
        int rc = recheck(c, p.splits > 1);
        if (rc != ABYSS_OK) return rc;
        c->spread_valid = true;
    }
    c->n_grav += 1;
    c->counts_valid = true;
    c->sum_ni += c->nt;
    c->interactions += (double)c->nt * (double)(c->n - 1);
    return ABYSS_OK;
}

}  // namespace abyss

extern "C" {

int abyss_ctx_launch(abyss_ctx_t c)
{
    int rc = abyss::ctx_launch_force(c);
    if (rc != ABYSS_OK) return rc;
    rc = abyss::keep_launch_lists(c);
    if (rc != ABYSS_OK) return set_detail(c, rc, "keeping resident neighbour lists failed");
    return ABYSS_OK;
}

int abyss_ctx_sync(abyss_ctx_t c)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return collect_kernel_timing(c);
}

int abyss_ctx_set_kernel_timing(abyss_ctx_t c, int enable)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    if (enable && !c->kev[0])
        for (hipEvent_t &e : c->kev) HIPCHK(c, hipEventCreate(&e));
    int rc = collect_kernel_timing(c);   // pairs of earlier launches
    if (rc != ABYSS_OK) return rc;
    c->ktiming = enable != 0;
    c->kforce_ms = 0.0;
    c->klaunches = 0;
    return ABYSS_OK;
}

int abyss_ctx_kernel_timing(abyss_ctx_t c, double *force_ms, long long *launches)
{
    if (!c || !force_ms || !launches) return ABYSS_ERR_INVALID_ARG;
    HIPCHK(c, hipSetDevice(c->device));
    int rc = collect_kernel_timing(c);
    if (rc != ABYSS_OK) return rc;
    *force_ms = c->kforce_ms;
    *launches = c->klaunches;
    return ABYSS_OK;
}

}  // extern "C"

namespace abyss {

int ctx_fetch_on_stream(abyss_ctx_t c, hipStream_t st, double *acc, double *adot, int *nnb,
                        int *list)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->nt <= 0) return set_detail(c, ABYSS_ERR_INVALID_ARG, "no targets set");
    if (!acc || !adot || !nnb || !list)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "fetch: null output array");
    HIPCHK(c, hipSetDevice(c->device));
    const int nt = c->nt;
    const int cap = c->cap;
    // Small results: acc, adot, counts and the full list rows are one
    // contiguous block, moved by one copy.  Larger ones: the counts first,
    // then only the populated list columns.
    const size_t full_bytes = (size_t)nt * cap * sizeof(int);
    const bool one_trip = full_bytes <= kOneTripListBytes;
    if (c->mirror_valid && st == c->stream) {
        // the launch's finalize already wrote every result into h_out
        HIPCHK(c, hipStreamSynchronize(st));
    } else {
        HIPCHK(c, hipMemcpyAsync(c->h_out, c->d_out,
                                 one_trip ? packed_bytes((size_t)nt) + full_bytes
                                          : (size_t)nt * (6 * sizeof(double) + sizeof(int)),
                                 hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
    }
    // populated width and overflow count, from the counts on the host
    int width = 0, overflowed = 0;
    for (int k = 0; k < nt; ++k) {
        const int cnt = c->h_nnb[k];
        overflowed += cnt < 0;
        width = std::max(width, std::min(cnt < 0 ? -cnt : cnt, cap));
    }
    const int *rows = c->h_rows;
    int stride = cap;   // row stride of `rows`
    if (!one_trip && width > 0) {
        const size_t need = (size_t)nt * width * sizeof(int);
        if (need > c->h_list_bytes) {
            hfree(c->h_list);
            c->h_list_bytes = 0;
            HIPCHK(c, hipHostMalloc(&c->h_list, grow(need), hipHostMallocDefault));
            c->h_list_bytes = grow(need);
        }
        HIPCHK(c, hipMemcpy2DAsync(c->h_list, (size_t)width * sizeof(int), c->d_list,
                                   (size_t)cap * sizeof(int), (size_t)width * sizeof(int), nt,
                                   hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        rows = c->h_list;
        stride = width;
    }
    memcpy(acc, c->h_res, (size_t)nt * 3 * sizeof(double));
    memcpy(adot, c->h_res + (size_t)nt * 3, (size_t)nt * 3 * sizeof(double));
    memcpy(nnb, c->h_nnb, (size_t)nt * sizeof(int));
    for (int k = 0; k < nt; ++k) {
        const int cnt = c->h_nnb[k];
        const int ncopy = std::min(cnt < 0 ? -cnt : cnt, cap);
        if (ncopy > 0)
            memcpy(list + (size_t)k * cap, rows + (size_t)k * stride,
                   (size_t)ncopy * sizeof(int));
    }
    if (overflowed > 0)
        return set_detail(c, ABYSS_ERR_NEIGHBOR_OVERFLOW,
                          "%d target(s) exceeded the neighbour cap %d", overflowed, cap);
    return ABYSS_OK;
}

}  // namespace abyss

extern "C" {

int abyss_ctx_fetch(abyss_ctx_t c, double *acc, double *adot, int *nnb, int *list)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    return abyss::ctx_fetch_on_stream(c, c->stream, acc, adot, nnb, list);
}

int abyss_ctx_query(abyss_ctx_t c, int nt, const int *targets, double *acc, double *adot,
                    int *nnb, int *list)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!acc || !adot || !nnb || !list)
        return set_detail(c, ABYSS_ERR_INVALID_ARG, "query: null output array");
    int rc = abyss_ctx_set_targets(c, nt, targets);
    if (rc != ABYSS_OK) return rc;
    rc = abyss_ctx_launch(c);
    if (rc != ABYSS_OK) return rc;
    return abyss_ctx_fetch(c, acc, adot, nnb, list);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// No-allocation query.  Every device buffer the query needs beyond the
// caller's inputs and outputs (FP32 masses, a checked copy of the target
// list, split partials, the checked-precision spread, flags and
// recomputation partials) lives in `scratch`, one static device array per
// GPU.  Its device address is looked up once per device with
// hipGetSymbolAddress; in host code the array's name is only a host-side
// stand-in, so kernels are always given the looked-up address.  Concurrent
// calls on one device share the array and race.
// ---------------------------------------------------------------------------
namespace {

__device__ __align__(256) char scratch[ABYSS_NOALLOC_SCRATCH_BYTES];

constexpr int kNoAllocDevices = 64;
std::mutex g_na_mutex;
char *g_na_ptr[kNoAllocDevices] = {};   // device address of `scratch` per device
int g_na_cu[kNoAllocDevices] = {};

// With `device` current: the per-device address of `scratch` and the CU count.
int noalloc_init(int device, char **ptr, int *cu)
{
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return ABYSS_ERR_HIP;
        int n_cu = 0;
        if (hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess)
            return ABYSS_ERR_HIP;
        g_na_cu[device] = n_cu;
        g_na_ptr[device] = static_cast<char *>(p);
    }
    *ptr = g_na_ptr[device];
    *cu = g_na_cu[device];
    return ABYSS_OK;
}

// FP32 masses (the rounding of the context's upload) and the checked target
// copy: an index outside [0, n) becomes 0 in the copy and is counted.
__global__ void noalloc_prep_kernel(const double *__restrict__ m, float *__restrict__ m32, int n,
                                    const int *__restrict__ tgt, int *__restrict__ tgt_s, int nt,
                                    int *__restrict__ bad)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) m32[i] = (float)m[i];
    if (i < nt) {
        const int t = tgt[i];
        const bool ok = t >= 0 && t < n;
        tgt_s[i] = ok ? t : 0;
        if (!ok) atomicAdd(bad, 1);
    }
}

// Results of targets outside [0, n): NaN forces and INT_MIN as the count.
__global__ void noalloc_mark_kernel(const int *__restrict__ tgt, int nt, int n,
                                    const int *__restrict__ bad, double *__restrict__ acc,
                                    double *__restrict__ adot, int *__restrict__ nnb)
{
    if (*bad == 0) return;
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nt) return;
    const int t = tgt[q];
    if (t >= 0 && t < n) return;
    const double nan = __builtin_nan("");
    for (int c = 0; c < 3; ++c) {
        acc[3 * (size_t)q + c] = nan;
        adot[3 * (size_t)q + c] = nan;
    }
    nnb[q] = INT_MIN;
}

// Bump allocator over the static array (256-byte aligned pieces).
struct Carve {
    char *base;
    size_t used = 0;
    char *take(size_t bytes)
    {
        char *r = base + used;
        used += (bytes + 255) & ~size_t(255);
        return r;
    }
};

}  // namespace

extern "C" {

int abyss_query_noalloc(int device, int n, const double *jset, int nt, const int *targets,
                        int cap, int precision, double *acc, double *adot, int *nnb, int *list,
                        void *stream)
{
    if (n <= 0 || nt <= 0 || nt > n || cap <= 0 || !jset || !targets || !acc || !adot || !nnb ||
        !list)
        return ABYSS_ERR_INVALID_ARG;
    if (precision != ABYSS_PREC_SPLIT && precision != ABYSS_PREC_CHECKED)
        return ABYSS_ERR_INVALID_ARG;
    if (device < 0 || device >= kNoAllocDevices || device >= abyss_device_count())
        return ABYSS_ERR_INVALID_ARG;
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess) return ABYSS_ERR_HIP;
    if (prev != device && hipSetDevice(device) != hipSuccess) return ABYSS_ERR_HIP;
    auto done = [&](int rc) {
        if (prev != device) (void)hipSetDevice(prev);
        return rc;
    };
    char *base = nullptr;
    int cu = 0;
    int rc = noalloc_init(device, &base, &cu);
    if (rc != ABYSS_OK) return done(rc);
    hipStream_t st = static_cast<hipStream_t>(stream);

    // the context's plan for these sizes under a split-scratch limit of
    // ABYSS_NOALLOC_SPLIT_BYTES (a context given the same limit plans alike)
    abyss_ctx_s tmp;
    tmp.device = device;
    tmp.cu_count = cu;
    tmp.cap = cap;
    tmp.precision = precision;
    tmp.scratch_limit = ABYSS_NOALLOC_SPLIT_BYTES;
    const Plan p = plan_for(&tmp, nt, n, precision, cap, true);
    // a small checked block: every target in FP64 terms, as a context does
    const bool direct = precision == ABYSS_PREC_CHECKED && p.fp64_terms;
    const bool checked = precision == ABYSS_PREC_CHECKED && !direct;
    const int kprec = direct ? abyss::kPrecRecheck : precision;
    const size_t S = (size_t)p.splits, t = (size_t)nt;

    Carve cv{base};
    int *counters = reinterpret_cast<int *>(cv.take(256));   // [0] flagged, [1] bad targets
    int *tgt_s = reinterpret_cast<int *>(cv.take(t * sizeof(int)));
    float *m32 = reinterpret_cast<float *>(cv.take((size_t)n * sizeof(float)));
    QueryOut part{};
    if (S > 1) {
        part.acc = reinterpret_cast<double *>(cv.take(S * t * 3 * sizeof(double)));
        part.adot = reinterpret_cast<double *>(cv.take(S * t * 3 * sizeof(double)));
        part.cnt = reinterpret_cast<int *>(cv.take(S * t * sizeof(int)));
        part.list = reinterpret_cast<int *>(cv.take(S * t * (size_t)cap * sizeof(int)));
    }
    float *spread = nullptr;
    int *fq = nullptr, *fi = nullptr;
    if (checked) {
        spread = reinterpret_cast<float *>(cv.take(2 * t * sizeof(float)));
        if (S > 1) part.spread = reinterpret_cast<float *>(cv.take(2 * S * t * sizeof(float)));
        fq = reinterpret_cast<int *>(cv.take(t * sizeof(int)));
        fi = reinterpret_cast<int *>(cv.take(t * sizeof(int)));
    }
    if (cv.used > ABYSS_NOALLOC_SCRATCH_BYTES) return done(ABYSS_ERR_SCRATCH_LIMIT);
    // the recomputation of checked precision: planned for every target (the
    // flagged count stays on the device), its splits bounded by what is left
    Plan rp{};
    QueryOut rpart{};
    if (checked) {
        const size_t row = 2 * 3 * sizeof(double) + sizeof(int);
        const size_t left = ABYSS_NOALLOC_SCRATCH_BYTES - cv.used;
        if (left < t * row + 3 * 256) return done(ABYSS_ERR_SCRATCH_LIMIT);
        tmp.scratch_limit = left - 3 * 256;
        rp = plan_for(&tmp, nt, n, ABYSS_PREC_REFINED, 0, false);
        const size_t rows = (size_t)rp.splits * t;
        rpart.acc = reinterpret_cast<double *>(cv.take(rows * 3 * sizeof(double)));
        rpart.adot = reinterpret_cast<double *>(cv.take(rows * 3 * sizeof(double)));
        rpart.cnt = reinterpret_cast<int *>(cv.take(rows * sizeof(int)));
        if (cv.used > ABYSS_NOALLOC_SCRATCH_BYTES) return done(ABYSS_ERR_SCRATCH_LIMIT);
    }

    JState js{};
    js.x = jset + 0 * (size_t)n;
    js.y = jset + 1 * (size_t)n;
    js.z = jset + 2 * (size_t)n;
    js.vx = jset + 3 * (size_t)n;
    js.vy = jset + 4 * (size_t)n;
    js.vz = jset + 5 * (size_t)n;
    js.m = jset + 6 * (size_t)n;
    js.rnb2 = jset + 7 * (size_t)n;
    js.m32 = m32;
    js.n = n;

#define NACHK(call)                                  \
    do {                                             \
        if ((call) != hipSuccess) return done(ABYSS_ERR_HIP); \
    } while (0)
    NACHK(abyss::launch_batch_reset(counters, st));   // both counters := 0
    const int w = std::max(n, nt);
    hipLaunchKernelGGL(noalloc_prep_kernel, dim3((w + 255) / 256), dim3(256), 0, st, js.m, m32, n,
                       targets, tgt_s, nt, counters + 1);
    NACHK(hipGetLastError());
    QueryOut fin{acc, adot, nnb, list};
    fin.spread = spread;
    if (S == 1) {
        NACHK(abyss::launch_regf(js, tgt_s, nt, p, fin, cap, kprec, st));
    } else {
        NACHK(abyss::launch_regf(js, tgt_s, nt, p, part, cap, kprec, st));
        NACHK(abyss::launch_finalize(part, p.splits, nt, cap, acc, adot, nnb, list,
                                     abyss::FinalMirror{}, spread, st));
    }
    if (checked) {
        const double ratio = ABYSS_CHECK_ERR_SCALE / 1.0e-6;
        NACHK(abyss::launch_check_flag(acc, adot, spread, tgt_s, nt, ratio * ratio, counters, fq,
                                       fi, st));
        NACHK(abyss::launch_recheck_dev(js, fi, counters, nt, rp, rpart, fq, acc, adot,
                                        abyss::FinalMirror{}, nullptr, st));
    }
    hipLaunchKernelGGL(noalloc_mark_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, targets,
                       nt, n, counters + 1, acc, adot, nnb);
    NACHK(hipGetLastError());
#undef NACHK
    return done(ABYSS_OK);
}

}  // extern "C"
