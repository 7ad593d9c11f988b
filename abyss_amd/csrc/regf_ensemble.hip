// Batched ensembles and the served ensemble (include/abyss_regf.h, Part 3).
//
// Ensemble: every cluster is a context (regf_context.h) whose work is issued
// on the ensemble's stream.  A launch builds a work list of (cluster, target
// block) items, copies it with the cluster descriptors to the device, and
// replays a captured HIP graph of two kernels: a counter reset and the
// persistent batch_kernel, which drains the queue, and a persistent finalize
// that merges the j-split partials.  Each cluster is split in j so that one
// item is ~4M pairs whatever N is (large clusters would otherwise give items
// of 1e8+ pairs that no scheduler can balance); split partials live in a
// scratch pool sized by the batch, not by the clusters.  Items are queued
// longest-first so the long ones start early and short ones fill in.
//
// Served ensemble: time-sliced continuous batching.  A worker thread starts
// pending requests (one per cluster at a time), and every batch carries at
// least one target block of each running request -- smallest remaining work
// first -- plus more blocks until a pair budget (~25 ms of work) is used.  So
// a small request that arrives while a huge one runs waits at most about one
// slice instead of the whole huge step.  The worker enqueues the next batch
// while the current one runs (two alternating pinned staging sets), so the
// GPU does not wait for the host between batches.  Finished requests go to
// a completion thread that copies results to the host on its own stream
// while later batches run.  Splitting a request's target blocks over batches does not
// change its bits.  Admission compares a cluster's estimated device
// footprint with the remaining HBM budget.
#include "regf_context.h"

#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

using abyss::BatchHeader;
using abyss::ClusterDesc;
using abyss::WorkItem;

struct abyss_ens_s {
    int device = -1;
    hipStream_t stream = nullptr;
    int cap = 0;
    int cu_count = 0;
    int max_clusters = 0;
    std::vector<abyss_ctx_t> cl;     // nullptr: free slot
    std::vector<char> active;        // has targets for the next launch
    std::vector<int> blk_first, blk_count;   // target-block range (count < 0: all)
    // checked precision: target blocks of the current targets evaluated
    // since the cluster's flagged count was last reset; a pass ends (and
    // its flagged targets are recomputed) once every block is done
    std::vector<int> chk_done;
    // pairs per work item of the main pass (the server uses shorter items,
    // so that its short batches drain sooner)
    double item_pairs = 4194304.0;

    // Pinned host staging of a launch's descriptors, work items, header and
    // prefixes.  Two sets alternate: a launch fills one while the copies of
    // the launch before it may still read the other, so launches need no
    // host wait between them.  `done` is recorded after the launch that
    // used the set (its number `seq`, 0: none).  h_cd ... h_ck_off below
    // point into the set of the launch being formed.
    struct Staging {
        ClusterDesc *h_cd = nullptr;
        WorkItem *h_items = nullptr;
        size_t items_cap = 0;
        BatchHeader *h_hdr = nullptr;
        int *h_fin_cl = nullptr, *h_fin_off = nullptr, *h_ck_off = nullptr;
        hipEvent_t done = nullptr;
        long long seq = 0;
    };
    Staging stg[2];
    int cur = 0;
    long long seq = 0;               // launches enqueued
    std::vector<long long> last_seq; // per cluster: the last launch it was in

    ClusterDesc *d_cd = nullptr, *h_cd = nullptr;
    WorkItem *d_items = nullptr, *h_items = nullptr;
    size_t items_cap = 0;            // device work items
    BatchHeader *d_hdr = nullptr, *h_hdr = nullptr;
    int *d_counter = nullptr;
    int *d_fin_cl = nullptr, *h_fin_cl = nullptr;     // [max_clusters]
    int *d_fin_off = nullptr, *h_fin_off = nullptr;   // [max_clusters + 1]
    char *d_pool = nullptr;                           // split scratch of one batch
    size_t pool_bytes = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int grid = 0, fin_grid = 0;

    // checked precision: per-descriptor flagged counts and the prefix of the
    // targets each descriptor checks ([max_clusters (+1)])
    bool checked = true;   // ABYSS_PREC_CHECKED unless set otherwise
    double check_tol = 1.0e-6;
    int *d_ckcount = nullptr, *h_ckcount = nullptr;
    int *d_ck_off = nullptr, *h_ck_off = nullptr;

    char detail[512] = {0};
};

namespace {

int ens_detail(abyss_ens_t e, int code, const char *fmt, ...)
{
    if (e) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(e->detail, sizeof(e->detail), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define ENSCHK(e, call)                                                                \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess)                                                          \
            return ens_detail(e,                                                       \
                              e_ == hipErrorOutOfMemory ? ABYSS_ERR_OUT_OF_MEMORY      \
                                                        : ABYSS_ERR_HIP,               \
                              "%s failed: %s", #call, hipGetErrorString(e_));          \
    } while (0)

// Copies a context's error into the ensemble's detail string.
int from_ctx(abyss_ens_t e, abyss_ctx_t c, int rc)
{
    if (rc != ABYSS_OK && rc != ABYSS_ERR_NEIGHBOR_OVERFLOW && c)
        ens_detail(e, rc, "%s", abyss_ctx_error_detail(c));
    return rc;
}

void drop_graph(abyss_ens_t e)
{
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// Work-item room for n items: the current staging set's pinned buffer (its
// previous launch has finished, see begin_stage) and the device buffer.
int ensure_items(abyss_ens_t e, size_t n)
{
    abyss_ens_s::Staging &st = e->stg[e->cur];
    if (n > st.items_cap) {
        const size_t cap = std::max<size_t>(n + n / 2, 4096);
        if (st.h_items) (void)hipHostFree(st.h_items);
        st.h_items = nullptr;
        st.items_cap = 0;
        ENSCHK(e, hipHostMalloc(&st.h_items, cap * sizeof(WorkItem), hipHostMallocDefault));
        st.items_cap = cap;
        e->h_items = st.h_items;
    }
    if (n <= e->items_cap) return ABYSS_OK;
    const size_t cap = std::max<size_t>(n + n / 2, 4096);
    ENSCHK(e, hipStreamSynchronize(e->stream));
    if (e->d_items) (void)hipFree(e->d_items);
    e->d_items = nullptr;
    e->items_cap = 0;
    drop_graph(e);   // the graph's kernel arguments point at the old buffer
    ENSCHK(e, hipMalloc(&e->d_items, cap * sizeof(WorkItem)));
    e->items_cap = cap;
    return ABYSS_OK;
}

// Start forming a launch: switch to the other staging set, wait until the
// launch that last used it has finished (two launches back), and point the
// h_* buffers into it.
int begin_stage(abyss_ens_t e)
{
    e->cur ^= 1;
    abyss_ens_s::Staging &st = e->stg[e->cur];
    if (st.seq > 0) ENSCHK(e, hipEventSynchronize(st.done));
    e->h_cd = st.h_cd;
    e->h_items = st.h_items;
    e->h_hdr = st.h_hdr;
    e->h_fin_cl = st.h_fin_cl;
    e->h_fin_off = st.h_fin_off;
    e->h_ck_off = st.h_ck_off;
    return ABYSS_OK;
}

// A launch is enqueued: number it, mark the current set with its completion
// and remember it as the last launch of each of its clusters.
int end_stage(abyss_ens_t e, const std::vector<int> &clusters)
{
    abyss_ens_s::Staging &st = e->stg[e->cur];
    e->seq += 1;
    ENSCHK(e, hipEventRecord(st.done, e->stream));
    st.seq = e->seq;
    for (int k : clusters) e->last_seq[k] = e->seq;
    return ABYSS_OK;
}

// Waits until every launch that included `cluster` has finished.  Launches
// up to seq - 2 have (a launch waits for the one two back before it forms).
int wait_cluster(abyss_ens_t e, int cluster)
{
    const long long ls = e->last_seq[cluster];
    if (ls <= 0 || ls < e->seq - 1) return ABYSS_OK;
    for (auto &st : e->stg)
        if (st.seq == ls) ENSCHK(e, hipEventSynchronize(st.done));
    return ABYSS_OK;
}

int ensure_pool(abyss_ens_t e, size_t bytes)
{
    if (bytes <= e->pool_bytes) return ABYSS_OK;
    const size_t cap = bytes + bytes / 2;
    ENSCHK(e, hipStreamSynchronize(e->stream));
    if (e->d_pool) (void)hipFree(e->d_pool);
    e->d_pool = nullptr;
    e->pool_bytes = 0;
    ENSCHK(e, hipMalloc(&e->d_pool, cap));
    e->pool_bytes = cap;
    return ABYSS_OK;
}

int capture_graph(abyss_ens_t e)
{
    ENSCHK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    hipError_t r1 = abyss::launch_batch_reset(e->d_counter, e->stream);
    hipError_t r2 = abyss::launch_batch(e->d_cd, e->d_items, e->d_hdr, e->d_counter, e->cap,
                                        e->grid, e->checked, e->stream);
    hipError_t r4 = abyss::launch_batch_finalize(e->d_cd, e->d_fin_cl, e->d_fin_off, e->d_hdr,
                                                 e->cap, e->fin_grid, e->stream);
    hipGraph_t g = nullptr;
    hipError_t r3 = hipStreamEndCapture(e->stream, &g);
    ENSCHK(e, r1);
    ENSCHK(e, r2);
    ENSCHK(e, r4);
    ENSCHK(e, r3);
    e->graph = g;
    ENSCHK(e, hipGraphInstantiate(&e->exec, e->graph, nullptr, nullptr, 0));
    return ABYSS_OK;
}

}  // namespace

extern "C" {

int abyss_ens_create(int device, int cap, int max_clusters, abyss_ens_t *out)
{
    if (!out || cap <= 0 || max_clusters <= 0) return ABYSS_ERR_INVALID_ARG;
    *out = nullptr;
    const int ndev = abyss_device_count();
    if (ndev <= 0) return ABYSS_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return ABYSS_ERR_INVALID_ARG;
    abyss_ens_t e = new (std::nothrow) abyss_ens_s();
    if (!e) return ABYSS_ERR_OUT_OF_MEMORY;
    e->device = device;
    e->cap = cap;
    e->max_clusters = max_clusters;
    e->cl.assign(max_clusters, nullptr);
    e->active.assign(max_clusters, 0);
    e->blk_first.assign(max_clusters, 0);
    e->blk_count.assign(max_clusters, -1);
    e->chk_done.assign(max_clusters, 0);
    e->last_seq.assign(max_clusters, 0);
    int rc = ABYSS_OK;
    do {
        if (hipSetDevice(device) != hipSuccess) { rc = ABYSS_ERR_NO_DEVICE; break; }
        if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
            rc = ABYSS_ERR_HIP; break;
        }
        if (hipDeviceGetAttribute(&e->cu_count, hipDeviceAttributeMultiprocessorCount, device) !=
            hipSuccess) { rc = ABYSS_ERR_HIP; break; }
        // as many 256-thread workgroups per CU as the batch kernel keeps
        // resident (5 at its 96 VGPRs): one persistent wave of workgroups
        e->grid = std::max(1, e->cu_count) * abyss::batch_blocks_per_cu();
        e->fin_grid = std::max(1, e->cu_count) * 8;
        if (hipMalloc(&e->d_cd, (size_t)max_clusters * sizeof(ClusterDesc)) != hipSuccess ||
            hipMalloc(&e->d_hdr, sizeof(BatchHeader)) != hipSuccess ||
            hipMalloc(&e->d_counter, 4 * sizeof(int)) != hipSuccess ||
            hipMalloc(&e->d_fin_cl, (size_t)max_clusters * sizeof(int)) != hipSuccess ||
            hipMalloc(&e->d_fin_off, (size_t)(max_clusters + 1) * sizeof(int)) != hipSuccess ||
            hipMalloc(&e->d_ckcount, (size_t)max_clusters * sizeof(int)) != hipSuccess ||
            hipHostMalloc(&e->h_ckcount, (size_t)max_clusters * sizeof(int),
                          hipHostMallocDefault) != hipSuccess ||
            hipMalloc(&e->d_ck_off, (size_t)(max_clusters + 1) * sizeof(int)) != hipSuccess) {
            rc = ABYSS_ERR_OUT_OF_MEMORY; break;
        }
        for (auto &st : e->stg) {
            if (hipHostMalloc(&st.h_cd, (size_t)max_clusters * sizeof(ClusterDesc),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&st.h_hdr, sizeof(BatchHeader), hipHostMallocDefault) !=
                    hipSuccess ||
                hipHostMalloc(&st.h_fin_cl, (size_t)max_clusters * sizeof(int),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&st.h_fin_off, (size_t)(max_clusters + 1) * sizeof(int),
                              hipHostMallocDefault) != hipSuccess ||
                hipHostMalloc(&st.h_ck_off, (size_t)(max_clusters + 1) * sizeof(int),
                              hipHostMallocDefault) != hipSuccess) {
                rc = ABYSS_ERR_OUT_OF_MEMORY; break;
            }
            if (hipEventCreateWithFlags(&st.done, hipEventDisableTiming) != hipSuccess) {
                rc = ABYSS_ERR_HIP; break;
            }
            memset(st.h_hdr, 0, sizeof(BatchHeader));
        }
        if (rc != ABYSS_OK) break;
    } while (0);
    if (rc != ABYSS_OK) {
        abyss_ens_destroy(e);
        return rc;
    }
    *out = e;
    return ABYSS_OK;
}

int abyss_ens_destroy(abyss_ens_t e)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    if (e->device >= 0) (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    for (auto &c : e->cl)
        if (c) abyss_ctx_destroy(c);
    drop_graph(e);
    if (e->d_cd) (void)hipFree(e->d_cd);
    if (e->d_items) (void)hipFree(e->d_items);
    if (e->d_hdr) (void)hipFree(e->d_hdr);
    if (e->d_counter) (void)hipFree(e->d_counter);
    if (e->d_fin_cl) (void)hipFree(e->d_fin_cl);
    if (e->d_fin_off) (void)hipFree(e->d_fin_off);
    if (e->d_pool) (void)hipFree(e->d_pool);
    if (e->d_ckcount) (void)hipFree(e->d_ckcount);
    if (e->h_ckcount) (void)hipHostFree(e->h_ckcount);
    if (e->d_ck_off) (void)hipFree(e->d_ck_off);
    for (auto &st : e->stg) {
        if (st.h_cd) (void)hipHostFree(st.h_cd);
        if (st.h_items) (void)hipHostFree(st.h_items);
        if (st.h_hdr) (void)hipHostFree(st.h_hdr);
        if (st.h_fin_cl) (void)hipHostFree(st.h_fin_cl);
        if (st.h_fin_off) (void)hipHostFree(st.h_fin_off);
        if (st.h_ck_off) (void)hipHostFree(st.h_ck_off);
        if (st.done) (void)hipEventDestroy(st.done);
    }
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return ABYSS_OK;
}

int abyss_ens_set_precision(abyss_ens_t e, int mode)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    if (mode != ABYSS_PREC_SPLIT && mode != ABYSS_PREC_CHECKED)
        return ens_detail(e, ABYSS_ERR_INVALID_ARG,
                          "ensembles run split or checked precision (mode %d)", mode);
    const bool checked = mode == ABYSS_PREC_CHECKED;
    if (checked != e->checked) {
        ENSCHK(e, hipSetDevice(e->device));
        ENSCHK(e, hipStreamSynchronize(e->stream));
        drop_graph(e);   // the captured batch kernel depends on the mode
        e->checked = checked;
        e->chk_done.assign(e->max_clusters, 0);   // passes start afresh
    }
    return ABYSS_OK;
}

int abyss_ens_set_check_tolerance(abyss_ens_t e, double tol)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    if (!(tol > 0.0) || !std::isfinite(tol))
        return ens_detail(e, ABYSS_ERR_INVALID_ARG, "check tolerance must be finite and > 0");
    e->check_tol = tol;
    return ABYSS_OK;
}

int abyss_ens_last_recheck(abyss_ens_t e, int cluster, int *count)
{
    if (!e || !count || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = e->cl[cluster];
    if (!c) return ens_detail(e, ABYSS_ERR_NO_JSET, "cluster %d is empty", cluster);
    *count = c->n_rechecked;
    return ABYSS_OK;
}

const char *abyss_ens_error_detail(abyss_ens_t e) { return e ? e->detail : ""; }
void *abyss_ens_stream(abyss_ens_t e) { return e ? (void *)e->stream : nullptr; }

int abyss_ens_upload(abyss_ens_t e, int cluster, int n, const double *m, const double *x,
                     const double *v, const double *r2)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    if (cluster < 0 || cluster >= e->max_clusters)
        return ens_detail(e, ABYSS_ERR_INVALID_ARG, "cluster %d outside [0,%d)", cluster,
                          e->max_clusters);
    abyss_ctx_t &c = e->cl[cluster];
    if (!c) {
        int rc = abyss::ctx_create_on_stream(e->device, e->cap, e->stream, &c);
        if (rc != ABYSS_OK) {
            c = nullptr;
            return ens_detail(e, rc, "context for cluster %d: %s", cluster,
                              abyss_error_string(rc));
        }
        // batched clusters run unsplit with a lane group per target (the
        // batch kernel's bodies: 4, 16 or 64 lanes)
        abyss_ctx_set_tuning(c, 1, 1, 2, 0);
        c->batch_member = true;
    }
    e->active[cluster] = 0;
    e->chk_done[cluster] = 0;
    return from_ctx(e, c, abyss_ctx_upload(c, n, m, x, v, r2));
}

int abyss_ens_release(abyss_ens_t e, int cluster)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    if (e->cl[cluster]) {
        ENSCHK(e, hipSetDevice(e->device));
        ENSCHK(e, hipStreamSynchronize(e->stream));
        abyss_ctx_destroy(e->cl[cluster]);
    }
    e->cl[cluster] = nullptr;
    e->active[cluster] = 0;
    return ABYSS_OK;
}

int abyss_ens_set_targets(abyss_ens_t e, int cluster, int nt, const int *targets)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = e->cl[cluster];
    if (!c) return ens_detail(e, ABYSS_ERR_NO_JSET, "cluster %d is empty", cluster);
    if (nt == 0) {
        e->active[cluster] = 0;
        return ABYSS_OK;
    }
    // the cluster's target and output buffers are free once its own last
    // launch has finished; other clusters' launches may still run
    int rc = wait_cluster(e, cluster);
    if (rc != ABYSS_OK) return rc;
    rc = abyss::ctx_set_targets_idle(c, nt, targets);
    e->active[cluster] = rc == ABYSS_OK ? 1 : 0;
    e->chk_done[cluster] = 0;
    e->blk_first[cluster] = 0;
    e->blk_count[cluster] = -1;
    return from_ctx(e, c, rc);
}

int abyss_ens_blocks(abyss_ens_t e, int cluster, int *nblocks, int *targets_per_block)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = e->cl[cluster];
    if (!c || !e->active[cluster] || !c->plan_valid)
        return ens_detail(e, ABYSS_ERR_INVALID_ARG, "cluster %d has no targets", cluster);
    if (nblocks) *nblocks = c->plan.grid_x;
    if (targets_per_block) *targets_per_block = abyss::targets_per_block(c->plan);
    return ABYSS_OK;
}

int abyss_ens_plan(abyss_ens_t e, int cluster, int *lanes_per_target, int *tile)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = e->cl[cluster];
    if (!c || !e->active[cluster] || !c->plan_valid)
        return ens_detail(e, ABYSS_ERR_INVALID_ARG, "cluster %d has no targets", cluster);
    if (lanes_per_target) *lanes_per_target = c->plan.jl;
    if (tile) *tile = c->plan.tile;
    return ABYSS_OK;
}

int abyss_ens_set_block_range(abyss_ens_t e, int cluster, int first, int count)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters || first < 0)
        return ABYSS_ERR_INVALID_ARG;
    e->blk_first[cluster] = first;
    e->blk_count[cluster] = count;
    return ABYSS_OK;
}

int abyss_ens_launch(abyss_ens_t e)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    ENSCHK(e, hipSetDevice(e->device));
    // Per active cluster: its target-block range and a j-split so that one
    // work item is about kItemPairs pairs whatever N is.  The split depends
    // only on the cluster's own N and targets per block.
    constexpr double kItemPairs = 4194304.0;
    constexpr int kItemTilesSmall = 16;
    struct Entry {
        int cluster, first, nblocks, splits, jspan, nt_slice;
        double item_cost;
    };
    std::vector<Entry> act;
    size_t nitems = 0, pool = 0;
    const size_t per_target = (size_t)e->cap * sizeof(int) + 6 * sizeof(double) + sizeof(int) +
                              (e->checked ? 2 * sizeof(float) : 0);
    for (int k = 0; k < e->max_clusters; ++k) {
        abyss_ctx_t c = e->cl[k];
        if (!c || !e->active[k] || c->nt <= 0 || !c->plan_valid) continue;
        const int tpb = abyss::targets_per_block(c->plan);
        const int first = std::min(e->blk_first[k], c->plan.grid_x);
        const int avail = c->plan.grid_x - first;
        const int nb = e->blk_count[k] < 0 ? avail : std::min(e->blk_count[k], avail);
        if (nb <= 0) continue;
        const int eff = std::min(tpb, c->nt);
        const int tile = abyss::kTile;
        const int nround = ((c->n + tile - 1) / tile) * tile;
        int jspan = (int)std::ceil(e->item_pairs / eff / tile) * tile;
        // 16/64-lane groups do only 16 or 4 steps per tile, too little to
        // hide the next tile's load: an item's time is ~1-2 us per tile
        // whatever its pairs.  Cap their items at kItemTilesSmall tiles so a
        // small block of a large cluster spreads over many short items
        // (8 targets of a 500K-particle cluster: one 2 ms item before).
        if (c->plan.jl >= 16) jspan = std::min(jspan, kItemTilesSmall * tile);
        jspan = std::max(tile, std::min(jspan, nround));
        int S = (c->n + jspan - 1) / jspan;
        if (S > abyss::kMaxSplitsBatch) {
            S = abyss::kMaxSplitsBatch;
            jspan = (((c->n + S - 1) / S + tile - 1) / tile) * tile;
            S = (c->n + jspan - 1) / jspan;
        }
        const int nt_slice = std::min(c->nt, (first + nb) * tpb) - first * tpb;
        act.push_back({k, first, nb, S, jspan, nt_slice, (double)eff * jspan});
        nitems += (size_t)nb * S;
        if (S > 1) pool += (size_t)S * nt_slice * per_target + 4 * 256;
    }
    if (act.empty()) return ABYSS_OK;
    std::sort(act.begin(), act.end(), [](const Entry &a, const Entry &b) {
        return a.item_cost != b.item_cost ? a.item_cost > b.item_cost : a.cluster < b.cluster;
    });
    // this launch's staging set: the launch two back (which used it) has
    // finished; whatever is enqueued below is marked with this launch's
    // completion on every return path
    int rc = begin_stage(e);
    if (rc != ABYSS_OK) return rc;
    std::vector<int> members;
    for (const Entry &en : act) members.push_back(en.cluster);
    struct StageEnd {
        abyss_ens_t e;
        const std::vector<int> &members;
        ~StageEnd() { (void)end_stage(e, members); }
    } stage_end{e, members};
    if (e->checked) {
        for (const Entry &en : act) {
            abyss_ctx_t c = e->cl[en.cluster];
            int rc = abyss::ctx_ensure_check(c, c->nt, 1);
            if (rc != ABYSS_OK) return from_ctx(e, c, rc);
            c->spread_valid = false;
            c->n_rechecked = 0;
        }
    }
    rc = ensure_items(e, nitems);
    if (rc != ABYSS_OK) return rc;
    rc = ensure_pool(e, pool);
    if (rc != ABYSS_OK) return rc;
    size_t it = 0;
    char *bump = e->d_pool;
    auto take = [&](size_t bytes) {
        char *r = bump;
        bump += (bytes + 255) & ~size_t(255);
        return r;
    };
    int nfc = 0, nfin = 0, nck = 0;
    for (size_t a = 0; a < act.size(); ++a) {
        const Entry &en = act[a];
        abyss_ctx_t c = e->cl[en.cluster];
        const int tpb = abyss::targets_per_block(c->plan);
        ClusterDesc &d = e->h_cd[a];
        d.spread = e->checked ? c->chk_spread : nullptr;
        d.fq = c->chk_fq;
        d.fi = c->chk_fi;
        // the cluster's own flagged count: a request whose target blocks
        // span several batches accumulates its flagged targets over them
        d.count = e->d_ckcount + en.cluster;
        d.chk_reset = e->chk_done[en.cluster] == 0 ? 1 : 0;
        d.chk_q0 = en.first * tpb;
        d.chk_nq = en.nt_slice;
        e->h_ck_off[a] = nck;
        nck += en.nt_slice;
        d.js = c->js;
        d.tgt = c->tgt;
        d.nt = c->nt;
        d.jl = c->plan.jl;
        d.tile = c->plan.tile;
        d.splits = en.splits;
        d.jspan = en.jspan;
        d.acc = c->d_acc;
        d.adot = c->d_adot;
        d.nnb = c->d_nnb;
        d.list = c->d_list;
        if (en.splits > 1) {
            const size_t m = (size_t)en.splits * en.nt_slice;
            d.q_base = en.first * tpb;
            d.nt_stride = en.nt_slice;
            d.part.acc = reinterpret_cast<double *>(take(m * 3 * sizeof(double)));
            d.part.adot = reinterpret_cast<double *>(take(m * 3 * sizeof(double)));
            d.part.cnt = reinterpret_cast<int *>(take(m * sizeof(int)));
            d.part.list = reinterpret_cast<int *>(take(m * (size_t)e->cap * sizeof(int)));
            d.part.spread =
                e->checked ? reinterpret_cast<float *>(take(m * 2 * sizeof(float))) : nullptr;
            e->h_fin_cl[nfc] = (int)a;
            e->h_fin_off[nfc] = nfin;
            ++nfc;
            nfin += en.nt_slice;
        } else {
            d.q_base = 0;
            d.nt_stride = c->nt;
            d.part = abyss::QueryOut{};
        }
        for (int b = 0; b < en.nblocks; ++b)
            for (int sp = 0; sp < en.splits; ++sp)
                e->h_items[it++] = WorkItem{(int)a, en.first + b, sp, 0};
        c->n_grav += 1;
    }
    e->h_fin_off[nfc] = nfin;
    e->h_ck_off[act.size()] = nck;
    e->h_hdr->nitems = (int)nitems;
    e->h_hdr->nfin = nfin;
    e->h_hdr->nfc = nfc;
    ENSCHK(e, hipMemcpyAsync(e->d_cd, e->h_cd, act.size() * sizeof(ClusterDesc),
                             hipMemcpyHostToDevice, e->stream));
    ENSCHK(e, hipMemcpyAsync(e->d_items, e->h_items, nitems * sizeof(WorkItem),
                             hipMemcpyHostToDevice, e->stream));
    ENSCHK(e, hipMemcpyAsync(e->d_fin_cl, e->h_fin_cl, std::max(nfc, 1) * sizeof(int),
                             hipMemcpyHostToDevice, e->stream));
    ENSCHK(e, hipMemcpyAsync(e->d_fin_off, e->h_fin_off, (nfc + 1) * sizeof(int),
                             hipMemcpyHostToDevice, e->stream));
    ENSCHK(e, hipMemcpyAsync(e->d_hdr, e->h_hdr, sizeof(BatchHeader), hipMemcpyHostToDevice,
                             e->stream));
    if (e->checked) {
        ENSCHK(e, hipMemcpyAsync(e->d_ck_off, e->h_ck_off, (act.size() + 1) * sizeof(int),
                                 hipMemcpyHostToDevice, e->stream));
        ENSCHK(e, abyss::launch_batch_count_reset(e->d_cd, (int)act.size(), e->stream));
    }
    if (!e->exec) {
        rc = capture_graph(e);
        if (rc != ABYSS_OK) return rc;
    }
    ENSCHK(e, hipGraphLaunch(e->exec, e->stream));
    if (e->checked) {
        // flag every checked target of the batch, read the per-cluster
        // counts back (one host wait), then recompute each cluster's
        // flagged targets with its own plan
        const double ratio = ABYSS_CHECK_ERR_SCALE / e->check_tol;
        ENSCHK(e, abyss::launch_batch_check_flag(e->d_cd, e->d_ck_off, (int)act.size(), nck,
                                                 ratio * ratio, e->stream));
        ENSCHK(e, hipMemcpyAsync(e->h_ckcount, e->d_ckcount, e->max_clusters * sizeof(int),
                                 hipMemcpyDeviceToHost, e->stream));
        ENSCHK(e, hipStreamSynchronize(e->stream));
        // The recomputation of the flagged targets of every cluster whose
        // target blocks this batch completes, in one persistent launch.  A
        // cluster's items follow the rule above for its total flagged count
        // (over all the batches of its blocks) and N, so its results do not
        // depend on the rest of the batch nor on how its blocks were split
        // over batches.  The batch has finished: its descriptor, item,
        // header, prefix and pool buffers can be refilled.
        struct Rec {
            int a, m, jl, nb, splits, jspan;
        };
        std::vector<Rec> rs;
        size_t nitems2 = 0, pool2 = 0;
        for (size_t a = 0; a < act.size(); ++a) {
            abyss_ctx_t c = e->cl[act[a].cluster];
            c->spread_valid = true;
            int &done = e->chk_done[act[a].cluster];
            done += act[a].nblocks;
            if (done < c->plan.grid_x) {
                c->n_rechecked = 0;   // more blocks of this pass to come
                continue;
            }
            done = 0;                 // the pass is complete
            const int m = e->h_ckcount[act[a].cluster];
            c->n_rechecked = m;
            if (m <= 0) continue;
            if (m > c->nt)
                return ens_detail(e, ABYSS_ERR_HIP, "check: %d flagged of %d", m, c->nt);
            const int jl = m >= 256 ? 4 : (m >= 16 ? 16 : 64);
            const int tpb = 4 * (64 / jl);
            const int tile = abyss::kTile;
            const int nround = ((c->n + tile - 1) / tile) * tile;
            int jspan = (int)std::ceil(kItemPairs / std::min(tpb, m) / tile) * tile;
            // every lane group: the recomputation is a few percent of the
            // batch's pairs, so its items must be short to spread over the
            // chip (a 4-lane item of kItemPairs FP64-term pairs took ~5 ms,
            // the whole recomputation's length)
            jspan = std::min(jspan, kItemTilesSmall * tile);
            jspan = std::max(tile, std::min(jspan, nround));
            int S = (c->n + jspan - 1) / jspan;
            if (S > abyss::kMaxSplitsBatch) {
                S = abyss::kMaxSplitsBatch;
                jspan = (((c->n + S - 1) / S + tile - 1) / tile) * tile;
                S = (c->n + jspan - 1) / jspan;
            }
            const int nb = (m + tpb - 1) / tpb;
            rs.push_back({(int)a, m, jl, nb, S, jspan});
            nitems2 += (size_t)nb * S;
            pool2 += (size_t)S * m * (6 * sizeof(double) + sizeof(int)) + 3 * 256;
        }
        if (!rs.empty()) {
            rc = ensure_items(e, nitems2);
            if (rc != ABYSS_OK) return rc;
            rc = ensure_pool(e, pool2);
            if (rc != ABYSS_OK) return rc;
            bump = e->d_pool;
            size_t it2 = 0;
            int total = 0;
            // items longest first (as the main pass), so short ones fill the
            // tail; an item's slot of the partials does not depend on when
            // it runs
            struct Item2 {
                double cost;
                int r, b, sp;
            };
            std::vector<Item2> order;
            order.reserve(nitems2);
            for (size_t r = 0; r < rs.size(); ++r) {
                const Rec &x = rs[r];
                const int tpb = 4 * (64 / x.jl);
                abyss_ctx_t c = e->cl[act[x.a].cluster];
                for (int b = 0; b < x.nb; ++b)
                    for (int sp = 0; sp < x.splits; ++sp) {
                        const int tb = std::min(tpb, x.m - b * tpb);
                        const int js = std::min(x.jspan, c->n - sp * x.jspan);
                        order.push_back({(double)tb * js, (int)r, b, sp});
                    }
            }
            std::sort(order.begin(), order.end(), [](const Item2 &p, const Item2 &q) {
                if (p.cost != q.cost) return p.cost > q.cost;
                if (p.r != q.r) return p.r < q.r;
                return p.b != q.b ? p.b < q.b : p.sp < q.sp;
            });
            for (const Item2 &o : order) e->h_items[it2++] = WorkItem{o.r, o.b, o.sp, 0};
            for (size_t r = 0; r < rs.size(); ++r) {
                const Rec &x = rs[r];
                abyss_ctx_t c = e->cl[act[x.a].cluster];
                ClusterDesc &d = e->h_cd[r];
                d = ClusterDesc{};
                d.js = c->js;
                d.tgt = c->chk_fi;
                d.nt = x.m;
                d.jl = x.jl;
                d.tile = abyss::kTile;
                d.splits = x.splits;
                d.jspan = x.jspan;
                d.nt_stride = x.m;
                const size_t mm = (size_t)x.splits * x.m;
                d.part.acc = reinterpret_cast<double *>(take(mm * 3 * sizeof(double)));
                d.part.adot = reinterpret_cast<double *>(take(mm * 3 * sizeof(double)));
                d.part.cnt = reinterpret_cast<int *>(take(mm * sizeof(int)));
                d.acc = c->d_acc;
                d.adot = c->d_adot;
                d.fq = c->chk_fq;
                e->h_ck_off[r] = total;
                total += x.m;
            }
            e->h_ck_off[rs.size()] = total;
            e->h_hdr->nitems = (int)nitems2;
            ENSCHK(e, hipMemcpyAsync(e->d_cd, e->h_cd, rs.size() * sizeof(ClusterDesc),
                                     hipMemcpyHostToDevice, e->stream));
            ENSCHK(e, hipMemcpyAsync(e->d_items, e->h_items, nitems2 * sizeof(WorkItem),
                                     hipMemcpyHostToDevice, e->stream));
            ENSCHK(e, hipMemcpyAsync(e->d_ck_off, e->h_ck_off, (rs.size() + 1) * sizeof(int),
                                     hipMemcpyHostToDevice, e->stream));
            ENSCHK(e, hipMemcpyAsync(e->d_hdr, e->h_hdr, sizeof(BatchHeader),
                                     hipMemcpyHostToDevice, e->stream));
            ENSCHK(e, abyss::launch_batch_recheck(e->d_cd, e->d_items, e->d_hdr, e->d_counter,
                                                  e->d_ck_off, (int)rs.size(), total, e->grid,
                                                  e->fin_grid, e->stream));
        }
    }
    return ABYSS_OK;
}

int abyss_ens_sync(abyss_ens_t e)
{
    if (!e) return ABYSS_ERR_INVALID_ARG;
    ENSCHK(e, hipSetDevice(e->device));
    ENSCHK(e, hipStreamSynchronize(e->stream));
    return ABYSS_OK;
}

int abyss_ens_fetch(abyss_ens_t e, int cluster, double *acc, double *adot, int *nnb, int *list)
{
    if (!e || cluster < 0 || cluster >= e->max_clusters) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = e->cl[cluster];
    if (!c) return ens_detail(e, ABYSS_ERR_NO_JSET, "cluster %d is empty", cluster);
    return from_ctx(e, c, abyss_ctx_fetch(c, acc, adot, nnb, list));
}

unsigned long long abyss_ens_device_bytes(abyss_ens_t e)
{
    if (!e) return 0;
    unsigned long long b = (unsigned long long)e->max_clusters * sizeof(ClusterDesc) +
                           e->items_cap * sizeof(WorkItem) + e->pool_bytes;
    for (auto c : e->cl) b += abyss::ctx_device_bytes(c);
    return b;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Served ensemble: time-sliced continuous batching
// ---------------------------------------------------------------------------
namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id = 0;
    int cluster = 0;
    std::vector<int> targets;
    Clock::time_point t_submit;
    double work = 0.0;   // pair-interactions: nt * (N - 1)
    // progress through the cluster's target blocks (worker thread only)
    int nblocks = 0, next_block = 0;
    double block_cost = 0.0;   // pairs per block
    bool done = false;
    int status = ABYSS_OK;
    std::vector<double> acc, adot;
    std::vector<int> nnb, list;
    double latency_ms = 0.0;
};

using ReqPtr = std::shared_ptr<Request>;

}  // namespace

struct abyss_srv_s {
    abyss_ens_t ens = nullptr;
    int device = 0;
    int cap = 0;
    int max_clusters = 0;
    unsigned long long budget = 0, used = 0;
    double batch_pairs = 3.0e10;   // work budget of one batch (pairs)
    std::vector<unsigned long long> footprint;   // 0: free slot
    std::vector<int> nparticles;
    std::vector<char> busy;        // a request of the cluster is running / being fetched

    std::mutex mu;        // queues, request map, statistics, cluster table
    std::mutex ens_mu;    // every call into the ensemble
    std::condition_variable cv_work, cv_done, cv_fetch;
    std::deque<ReqPtr> pending;    // submitted, not started
    std::deque<ReqPtr> to_fetch;   // all blocks done (or failed), results not on the host
    std::unordered_map<long long, ReqPtr> reqs;
    std::thread worker, completer;
    bool stop = false;
    bool worker_done = false;   // set by the worker when it has drained everything
    long long next_id = 1;

    std::vector<double> latencies;
    double interactions = 0.0;
    long long batches = 0, completed = 0;
    // GPU time of the batches (events around each on the ensemble stream)
    // and the wall-clock start of the statistics
    hipEvent_t ev_b0[2] = {}, ev_b1[2] = {};   // two batches in flight
    double busy_ms = 0.0;
    Clock::time_point t_stats = Clock::now();

    // request admission: submitted, not yet completed work (pairs) and its
    // limit (0: unlimited); rejections and queue high-water marks
    double max_outstanding = 0.0;
    double outstanding = 0.0;
    double peak_outstanding = 0.0;
    long long rejected = 0, max_pending = 0;
};

namespace {

// Pairs per work item of the server's batches (the batched ensemble uses
// 4M): its time slices are ~25 ms, so shorter items drain them sooner.
constexpr double kServeItemPairs = 1048576.0;

// Worker: forms batches.  Every running request gets at least one target
// block per batch; requests with the least remaining work go first, then
// blocks are handed out until the batch's pair budget is used.  Batches are
// pipelined: batch k + 1 is formed and enqueued while batch k runs (the
// ensemble's staging sets alternate, and a cluster's new targets wait only
// for the cluster's own last batch), and the requests batch k completes go
// to the completion thread once it has finished.  At most two batches are
// in flight.
void serve_loop(abyss_srv_t s)
{
    (void)hipSetDevice(s->device);
    std::vector<ReqPtr> running;
    struct Flight {
        int slot = 0;                // event pair of the batch
        bool recorded = false;       // its end event is on the stream
        bool timed = false;          // both events are
        std::vector<ReqPtr> done;    // requests complete after this batch
    };
    std::deque<Flight> flights;
    int slot = 0;
    // Waits for the oldest batch in flight, adds its GPU time and hands the
    // requests it completed to the completion thread.
    auto retire = [&]() {
        Flight f = std::move(flights.front());
        flights.pop_front();
        bool ok = f.recorded ? hipEventSynchronize(s->ev_b1[f.slot]) == hipSuccess
                             : hipStreamSynchronize(s->ens->stream) == hipSuccess;
        float ms = 0.0f;
        if (!ok || !f.timed ||
            hipEventElapsedTime(&ms, s->ev_b0[f.slot], s->ev_b1[f.slot]) != hipSuccess)
            ms = 0.0f;
        std::lock_guard<std::mutex> lk(s->mu);
        s->busy_ms += ms;
        for (auto &r : f.done) {
            if (!ok && r->status == ABYSS_OK) r->status = ABYSS_ERR_HIP;
            s->to_fetch.push_back(r);
        }
        if (!f.done.empty()) s->cv_fetch.notify_one();
    };
    for (;;) {
        std::vector<ReqPtr> starting;
        bool idle = false;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            auto startable = [&] {
                for (auto &r : s->pending)
                    if (!s->busy[r->cluster]) return true;
                return false;
            };
            if (flights.empty())
                s->cv_work.wait(lk, [&] {
                    return !running.empty() || startable() || (s->stop && s->pending.empty());
                });
            if (s->stop && s->pending.empty() && running.empty() && flights.empty()) {
                s->worker_done = true;
                s->cv_fetch.notify_all();
                return;
            }
            for (auto it = s->pending.begin(); it != s->pending.end();) {
                const int c = (*it)->cluster;
                if (!s->busy[c]) {
                    s->busy[c] = 1;
                    starting.push_back(*it);
                    it = s->pending.erase(it);
                } else {
                    ++it;
                }
            }
            idle = running.empty() && starting.empty();
        }
        if (idle) {
            retire();   // nothing to add: wait for the batch in flight
            continue;
        }
        std::vector<ReqPtr> failed;
        {
            std::lock_guard<std::mutex> elk(s->ens_mu);
            for (auto &r : starting) {
                r->status = abyss_ens_set_targets(s->ens, r->cluster, (int)r->targets.size(),
                                                  r->targets.data());
                int tpb = 0;
                if (r->status == ABYSS_OK)
                    r->status = abyss_ens_blocks(s->ens, r->cluster, &r->nblocks, &tpb);
                if (r->status != ABYSS_OK) {
                    failed.push_back(r);
                    continue;
                }
                r->block_cost = (double)std::min<size_t>(tpb, r->targets.size()) *
                                (double)s->nparticles[r->cluster];
                running.push_back(r);
            }
            if (!running.empty()) {
                std::sort(running.begin(), running.end(), [](const ReqPtr &a, const ReqPtr &b) {
                    const double ra = (a->nblocks - a->next_block) * a->block_cost;
                    const double rb = (b->nblocks - b->next_block) * b->block_cost;
                    return ra != rb ? ra < rb : a->id < b->id;
                });
                double left = s->batch_pairs;
                std::vector<int> take(running.size());
                for (size_t k = 0; k < running.size(); ++k) {
                    const ReqPtr &r = running[k];
                    const int remain = r->nblocks - r->next_block;
                    int nb = 1;
                    if (left > r->block_cost)
                        nb = (int)std::min<double>(remain, std::floor(left / r->block_cost));
                    nb = std::max(1, std::min(nb, remain));
                    left -= nb * r->block_cost;
                    take[k] = nb;
                    abyss_ens_set_block_range(s->ens, r->cluster, r->next_block, nb);
                }
                Flight f;
                f.slot = slot;
                slot ^= 1;
                const bool t0 = hipEventRecord(s->ev_b0[f.slot], s->ens->stream) == hipSuccess;
                const int rc = abyss_ens_launch(s->ens);
                f.recorded = hipEventRecord(s->ev_b1[f.slot], s->ens->stream) == hipSuccess;
                f.timed = t0 && f.recorded;
                std::vector<ReqPtr> still;
                for (size_t k = 0; k < running.size(); ++k) {
                    const ReqPtr &r = running[k];
                    if (rc != ABYSS_OK) r->status = rc;
                    r->next_block += take[k];
                    if (r->status != ABYSS_OK || r->next_block >= r->nblocks) {
                        abyss_ens_set_targets(s->ens, r->cluster, 0, nullptr);
                        f.done.push_back(r);   // to the completion thread after the batch
                    } else {
                        still.push_back(r);
                    }
                }
                running.swap(still);
                flights.push_back(std::move(f));
                std::lock_guard<std::mutex> lk(s->mu);
                s->batches += 1;
            }
        }
        if (!failed.empty()) {
            std::lock_guard<std::mutex> lk(s->mu);
            for (auto &r : failed) s->to_fetch.push_back(r);
            s->cv_fetch.notify_one();
        }
        while (flights.size() > 1) retire();   // one batch queued behind the running one
    }
}

// Completion thread: copies finished requests' results to the host on its
// own stream while the worker's next batch runs, then frees the cluster.
void complete_loop(abyss_srv_t s)
{
    (void)hipSetDevice(s->device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    for (;;) {
        ReqPtr r;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_fetch.wait(lk, [&] { return !s->to_fetch.empty() || s->worker_done; });
            if (s->to_fetch.empty()) break;   // the worker is done and nothing is left
            r = s->to_fetch.front();
            s->to_fetch.pop_front();
        }
        double inter = 0.0;
        if (r->status == ABYSS_OK) {
            const size_t nt = r->targets.size();
            r->acc.resize(3 * nt);
            r->adot.resize(3 * nt);
            r->nnb.resize(nt);
            r->list.resize(nt * (size_t)s->cap);
            abyss_ctx_t c = s->ens->cl[r->cluster];
            r->status = abyss::ctx_fetch_on_stream(c, st ? st : c->stream, r->acc.data(),
                                                   r->adot.data(), r->nnb.data(), r->list.data());
            inter = (double)nt * (double)(s->nparticles[r->cluster] - 1);
        }
        const auto now = Clock::now();
        {
            std::lock_guard<std::mutex> lk(s->mu);
            r->latency_ms = std::chrono::duration<double, std::milli>(now - r->t_submit).count();
            r->done = true;
            s->busy[r->cluster] = 0;
            s->outstanding -= r->work;
            if (s->outstanding < 0.5) s->outstanding = 0.0;   // sums of integers
            s->latencies.push_back(r->latency_ms);
            s->completed += 1;
            s->interactions += inter;
        }
        s->cv_done.notify_all();
        s->cv_work.notify_one();
    }
    if (st) (void)hipStreamDestroy(st);
}

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = p * (double)(v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - (double)i;
    return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int abyss_srv_create(int device, int cap, int max_clusters, unsigned long long hbm_budget,
                     abyss_srv_t *out)
{
    if (!out || hbm_budget == 0) return ABYSS_ERR_INVALID_ARG;
    *out = nullptr;
    abyss_ens_t e = nullptr;
    int rc = abyss_ens_create(device, cap, max_clusters, &e);
    if (rc != ABYSS_OK) return rc;
    abyss_srv_t s = new (std::nothrow) abyss_srv_s();
    if (!s) {
        abyss_ens_destroy(e);
        return ABYSS_ERR_OUT_OF_MEMORY;
    }
    s->ens = e;
    e->item_pairs = kServeItemPairs;
    s->device = device;
    s->cap = cap;
    s->max_clusters = max_clusters;
    s->budget = hbm_budget;
    s->footprint.assign(max_clusters, 0);
    s->nparticles.assign(max_clusters, 0);
    s->busy.assign(max_clusters, 0);
    bool events_ok = hipSetDevice(device) == hipSuccess;
    for (int k = 0; k < 2 && events_ok; ++k)
        events_ok = hipEventCreate(&s->ev_b0[k]) == hipSuccess &&
                    hipEventCreate(&s->ev_b1[k]) == hipSuccess;
    if (!events_ok) {
        for (int k = 0; k < 2; ++k) {
            if (s->ev_b0[k]) (void)hipEventDestroy(s->ev_b0[k]);
            if (s->ev_b1[k]) (void)hipEventDestroy(s->ev_b1[k]);
        }
        abyss_ens_destroy(e);
        delete s;
        return ABYSS_ERR_HIP;
    }
    s->worker = std::thread(serve_loop, s);
    s->completer = std::thread(complete_loop, s);
    *out = s;
    return ABYSS_OK;
}

int abyss_srv_destroy(abyss_srv_t s)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();   // drains every submitted request
    if (s->completer.joinable()) s->completer.join();
    abyss_ens_destroy(s->ens);
    for (int k = 0; k < 2; ++k) {
        if (s->ev_b0[k]) (void)hipEventDestroy(s->ev_b0[k]);
        if (s->ev_b1[k]) (void)hipEventDestroy(s->ev_b1[k]);
    }
    delete s;
    return ABYSS_OK;
}

int abyss_srv_set_batch_budget(abyss_srv_t s, double pairs)
{
    if (!s || !(pairs > 0.0)) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> elk(s->ens_mu);
    s->batch_pairs = pairs;
    return ABYSS_OK;
}

int abyss_srv_set_precision(abyss_srv_t s, int mode)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> elk(s->ens_mu);   // between two batches
    return abyss_ens_set_precision(s->ens, mode);
}

int abyss_srv_set_check_tolerance(abyss_srv_t s, double tol)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> elk(s->ens_mu);
    return abyss_ens_set_check_tolerance(s->ens, tol);
}

unsigned long long abyss_srv_cluster_footprint(abyss_srv_t s, int n)
{
    if (!s || n <= 0) return 0;
    // j-set (68 B/particle) + targets and results for up to n targets, with
    // the 1.5x growth headroom of the context's buffers
    const double per = 68.0 + 1.5 * (4.0 + 48.0 + 4.0 + 4.0 * s->cap);
    return (unsigned long long)(per * n) + (1ull << 20);
}

int abyss_srv_add_cluster(abyss_srv_t s, int n, const double *m, const double *x,
                          const double *v, const double *r2, int *cluster)
{
    if (!s || !cluster || n <= 0) return ABYSS_ERR_INVALID_ARG;
    const unsigned long long need = abyss_srv_cluster_footprint(s, n);
    int slot = -1;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (int k = 0; k < s->max_clusters; ++k)
            if (s->footprint[k] == 0) {
                slot = k;
                break;
            }
        if (slot < 0) return ABYSS_ERR_BUSY;
        if (s->used + need > s->budget) return ABYSS_ERR_ADMISSION;
        s->footprint[slot] = need;   // reserve before the upload
        s->used += need;
        s->nparticles[slot] = n;
    }
    int rc;
    {
        std::lock_guard<std::mutex> elk(s->ens_mu);
        rc = abyss_ens_upload(s->ens, slot, n, m, x, v, r2);
        if (rc != ABYSS_OK) abyss_ens_release(s->ens, slot);
    }
    if (rc != ABYSS_OK) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->used -= s->footprint[slot];
        s->footprint[slot] = 0;
        return rc;
    }
    *cluster = slot;
    return ABYSS_OK;
}

int abyss_srv_remove_cluster(abyss_srv_t s, int cluster)
{
    if (!s || cluster < 0 || cluster >= s->max_clusters) return ABYSS_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->footprint[cluster] == 0) return ABYSS_ERR_INVALID_ARG;
        if (s->busy[cluster]) return ABYSS_ERR_BUSY;
        for (auto &r : s->pending)
            if (r->cluster == cluster) return ABYSS_ERR_BUSY;
        s->footprint[cluster] = ~0ull;   // keep the slot while it is released
    }
    int rc;
    {
        std::lock_guard<std::mutex> elk(s->ens_mu);
        rc = abyss_ens_release(s->ens, cluster);
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->used -= abyss_srv_cluster_footprint(s, s->nparticles[cluster]);
    s->footprint[cluster] = 0;
    s->nparticles[cluster] = 0;
    return rc;
}

int abyss_srv_submit(abyss_srv_t s, int cluster, int nt, const int *targets, long long *ticket)
{
    if (!s || !ticket || !targets || nt <= 0 || cluster < 0 || cluster >= s->max_clusters)
        return ABYSS_ERR_INVALID_ARG;
    auto r = std::make_shared<Request>();
    r->cluster = cluster;
    r->targets.assign(targets, targets + nt);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->footprint[cluster] == 0 || s->footprint[cluster] == ~0ull)
            return ABYSS_ERR_NO_JSET;
        const int n = s->nparticles[cluster];
        if (nt > n) return ABYSS_ERR_INVALID_ARG;
        for (int k = 0; k < nt; ++k)
            if (targets[k] < 0 || targets[k] >= n) return ABYSS_ERR_TARGET_RANGE;
        r->work = (double)nt * (double)(n - 1);
        // Admission: with a limit set, a request that would take the
        // outstanding work past it is turned away, unless nothing is
        // outstanding (so a request larger than the limit still runs alone).
        if (s->max_outstanding > 0.0 && s->outstanding > 0.0 &&
            s->outstanding + r->work > s->max_outstanding) {
            s->rejected += 1;
            return ABYSS_ERR_BUSY;
        }
        s->outstanding += r->work;
        s->peak_outstanding = std::max(s->peak_outstanding, s->outstanding);
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        s->reqs[r->id] = r;
        s->pending.push_back(r);
        s->max_pending = std::max<long long>(s->max_pending, (long long)s->pending.size());
        *ticket = r->id;
    }
    s->cv_work.notify_one();
    return ABYSS_OK;
}

int abyss_srv_wait(abyss_srv_t s, long long ticket, double *acc, double *adot, int *nnb,
                   int *list, double *latency_ms)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    ReqPtr r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) return ABYSS_ERR_INVALID_ARG;
        r = it->second;
        s->cv_done.wait(lk, [&] { return r->done; });
        s->reqs.erase(ticket);
    }
    if (latency_ms) *latency_ms = r->latency_ms;
    if (r->status != ABYSS_OK && r->status != ABYSS_ERR_NEIGHBOR_OVERFLOW) return r->status;
    if (acc) memcpy(acc, r->acc.data(), r->acc.size() * sizeof(double));
    if (adot) memcpy(adot, r->adot.data(), r->adot.size() * sizeof(double));
    if (nnb) memcpy(nnb, r->nnb.data(), r->nnb.size() * sizeof(int));
    if (list) memcpy(list, r->list.data(), r->list.size() * sizeof(int));
    return r->status;
}

int abyss_srv_stats(abyss_srv_t s, long long *completed, long long *batches,
                    double *interactions, double *p50_ms, double *p99_ms)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (completed) *completed = s->completed;
    if (batches) *batches = s->batches;
    if (interactions) *interactions = s->interactions;
    if (p50_ms) *p50_ms = percentile(s->latencies, 0.50);
    if (p99_ms) *p99_ms = percentile(s->latencies, 0.99);
    return ABYSS_OK;
}

int abyss_srv_reset_stats(abyss_srv_t s)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->latencies.clear();
    s->interactions = 0.0;
    s->batches = s->completed = 0;
    s->rejected = s->max_pending = 0;
    s->peak_outstanding = s->outstanding;
    s->busy_ms = 0.0;
    s->t_stats = Clock::now();
    return ABYSS_OK;
}

int abyss_srv_gpu_busy(abyss_srv_t s, double *busy_ms, double *wall_ms)
{
    if (!s || !busy_ms || !wall_ms) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    *busy_ms = s->busy_ms;
    *wall_ms = std::chrono::duration<double, std::milli>(Clock::now() - s->t_stats).count();
    return ABYSS_OK;
}

int abyss_srv_set_admission(abyss_srv_t s, double max_outstanding_pairs)
{
    if (!s || !(max_outstanding_pairs >= 0.0)) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->max_outstanding = max_outstanding_pairs;
    return ABYSS_OK;
}

int abyss_srv_queue_stats(abyss_srv_t s, long long *rejected, long long *max_pending,
                          double *peak_outstanding_pairs)
{
    if (!s) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    if (rejected) *rejected = s->rejected;
    if (max_pending) *max_pending = s->max_pending;
    if (peak_outstanding_pairs) *peak_outstanding_pairs = s->peak_outstanding;
    return ABYSS_OK;
}

}  // extern "C"
