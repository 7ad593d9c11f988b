// Plan autotuner (include/abyss_regf.h, abyss_autotune*).
//
// For target counts m = 1, 4, 16, ... up to n on an n-particle cluster (a
// synthetic uniform cube with r_nb for ~48 neighbours, or the caller's own
// j-set), every candidate plan of the FP32 pass is timed under split
// precision with HIP events (median of 5 launches, 3 above 1 ms):
//   lanes per target  JL in {4, 16, 64} (the lane-group kernels; the
//                     one-target-per-lane kernel runs only when forced)
//   shape             j per LDS tile and wave64s per workgroup (targets per
//                     workgroup = waves * 64 / JL): (256, 4) (128, 4)
//                     (64, 4) (128, 2) (64, 2) (64, 1), and for JL >= 16
//                     the one-shot (1024, 4) of splits up to 1024 j
//   workgroups per CU W in {1, 2, 4, 8, 16, 32, 48, 64, 96}, which sets the j-split
//                     count S = ceil(W * CUs / target blocks)
// These are the compile-time BatchSize (j per shared-memory load),
// NNB_per_block (targets per block) and GridDimY (j-splits) of the
// reference, src_mpi/def.h:77-79.  The fastest (JL, unroll, W, tile, waves)
// per power-of-two target-count bucket is kept in a per-device table that
// plan selection consults for every later query on that device (buckets
// between measured points take the nearest one).  The table can be written
// to and read from a small text file.  Batched ensembles take the lanes per
// target and the tile of each cluster from the same table (their persistent
// workgroups are 4 waves).  The table governs the FP32 pass: split
// precision, and checked precision above 256 targets (smaller checked
// blocks run in FP64 terms with the built-in rules).  No choice changes a
// result bit: the plan's j-split count is the only one that changes the
// order of the FP64 sums.
#include "regf_context.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

namespace abyss {

namespace {

constexpr int kBuckets = 32;        // floor(log2(nt)), nt < 2^31
constexpr int kMaxDevices = 64;

struct Table {
    bool valid = false;
    TuneEntry e[kBuckets];
};

std::mutex g_mu;
Table g_tab[kMaxDevices];

int bucket_of(int nt)
{
    int b = 0;
    while ((1 << (b + 1)) <= nt && b + 1 < kBuckets) ++b;
    return b;
}

// xorshift64* stream for the synthetic cluster
struct Rng {
    unsigned long long s;
    double next()
    {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return (double)((s * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
    }
};

}  // namespace

bool tuned_plan(int device, int nt, TuneEntry *out)
{
    if (device < 0 || device >= kMaxDevices || nt <= 0) return false;
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_tab[device].valid) return false;
    *out = g_tab[device].e[bucket_of(nt)];
    return true;
}

}  // namespace abyss

extern "C" {

int abyss_autotune_clear(int device)
{
    if (device < 0 || device >= abyss::kMaxDevices) return ABYSS_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(abyss::g_mu);
    abyss::g_tab[device].valid = false;
    return ABYSS_OK;
}

int abyss_autotune_load(int device, const char *path)
{
    if (device < 0 || device >= abyss::kMaxDevices || !path) return ABYSS_ERR_INVALID_ARG;
    FILE *f = fopen(path, "r");
    if (!f) return ABYSS_ERR_INVALID_ARG;
    abyss::Table t;
    int got = 0;
    char line[256];
    while (fgets(line, sizeof(line), f)) {
        int b, jl, un, w, tile = abyss::kTile, wpb = 4;
        if (line[0] == '#') continue;
        // six columns; a four-column line (an older table) has the default shape
        const int k = sscanf(line, "%d %d %d %d %d %d", &b, &jl, &un, &w, &tile, &wpb);
        if ((k == 4 || k == 6) && b >= 0 && b < abyss::kBuckets &&
            (jl == 1 || jl == 4 || jl == 16 || jl == 64) && (un == 2 || un == 4) && w > 0 &&
            abyss::shape_supported(tile, wpb)) {
            t.e[b] = abyss::TuneEntry{jl, un, w, tile, wpb};
            ++got;
        }
    }
    fclose(f);
    if (got != abyss::kBuckets) return ABYSS_ERR_INVALID_INPUT;
    t.valid = true;
    std::lock_guard<std::mutex> lk(abyss::g_mu);
    abyss::g_tab[device] = t;
    return ABYSS_OK;
}

int abyss_autotune_jset(int device, int n, const double *mass, const double *pos,
                        const double *vel, const double *r2, const char *path)
{
    if (device < 0 || device >= abyss::kMaxDevices || n < 256 || !mass || !pos || !vel || !r2)
        return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = nullptr;
    int rc = abyss_ctx_create(device, 256, &c);
    if (rc != ABYSS_OK) return rc;
    abyss::Rng rng{0x9E3779B97F4A7C15ull};
    rc = abyss_ctx_set_precision(c, ABYSS_PREC_SPLIT);
    if (rc == ABYSS_OK) rc = abyss_ctx_upload(c, n, mass, pos, vel, r2);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (rc == ABYSS_OK && (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = ABYSS_ERR_HIP;
    std::vector<int> idx(n);
    for (int k = 0; k < n; ++k) idx[k] = k;
    for (int k = n - 1; k > 0; --k) std::swap(idx[k], idx[(int)(rng.next() * (k + 1))]);

    struct Meas {
        int bucket;
        abyss::TuneEntry best;
    };
    std::vector<Meas> meas;
    const int cu = c ? (c->cu_count > 0 ? c->cu_count : 256) : 256;
    // candidates: every (tile, waves per workgroup) shape of each lane group
    struct Cand {
        int jl, un, tile, wpb;
    };
    std::vector<Cand> cands;
    constexpr int kShapes[][2] = {{256, 4}, {128, 4}, {64, 4}, {128, 2}, {64, 2}, {64, 1}};
    for (int jl : {4, 16, 64}) {
        for (const auto &sh : kShapes) cands.push_back({jl, 2, sh[0], sh[1]});
        // the one-shot tile (plans whose splits exceed it fall back to 256)
        if (jl >= 16) cands.push_back({jl, 2, abyss::kTileOneShot, 4});
    }
    std::vector<int> counts;
    for (int mt = 1; mt < n; mt *= 4) counts.push_back(mt);
    counts.push_back(n);
    for (int mt : counts) {
        if (rc != ABYSS_OK) break;
        rc = abyss_ctx_set_targets(c, mt, idx.data());
        if (rc != ABYSS_OK) break;
        double best_ms = 1e30;
        abyss::TuneEntry best{4, 2, 16};
        for (const Cand &cd : cands) {
            int prev_S = -1;
            for (int w : {1, 2, 4, 8, 16, 32, 48, 64, 96}) {
                const int tpb = cd.jl > 1 ? cd.wpb * (64 / cd.jl) : abyss::kBlock;
                const int gx = (mt + tpb - 1) / tpb;
                const int S = std::max(1, std::min((w * cu + gx - 1) / gx,
                                                   std::min(n / abyss::kTile, abyss::kMaxSplits)));
                if (S == prev_S) continue;   // same plan as the previous w
                prev_S = S;
                rc = abyss_ctx_set_tuning(c, 1, S, cd.un, cd.jl);
                if (rc == ABYSS_OK)
                    rc = abyss_ctx_set_shape(c, cd.jl > 1 ? cd.tile : 0, cd.jl > 1 ? cd.wpb : 0);
                if (rc == ABYSS_OK) rc = abyss_ctx_launch(c);   // warm-up (sizes the scratch)
                float t[5];
                int reps = 5;
                for (int r = 0; r < reps && rc == ABYSS_OK; ++r) {
                    if (hipEventRecord(e0, c->stream) != hipSuccess) rc = ABYSS_ERR_HIP;
                    if (rc == ABYSS_OK) rc = abyss_ctx_launch(c);
                    if (rc == ABYSS_OK && hipEventRecord(e1, c->stream) != hipSuccess)
                        rc = ABYSS_ERR_HIP;
                    if (rc == ABYSS_OK && hipEventSynchronize(e1) != hipSuccess)
                        rc = ABYSS_ERR_HIP;
                    if (rc == ABYSS_OK && hipEventElapsedTime(&t[r], e0, e1) != hipSuccess)
                        rc = ABYSS_ERR_HIP;
                    if (r == 0 && t[0] > 1.0f) reps = 3;   // long launches: 3 suffice
                }
                if (rc != ABYSS_OK) break;
                std::sort(t, t + reps);
                if (t[reps / 2] < best_ms) {
                    best_ms = t[reps / 2];
                    best = abyss::TuneEntry{cd.jl, cd.un, w, cd.tile, cd.wpb};
                }
            }
            if (rc != ABYSS_OK) break;
        }
        if (rc == ABYSS_OK) meas.push_back({abyss::bucket_of(mt), best});
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    abyss_ctx_destroy(c);
    if (rc != ABYSS_OK) return rc;
    if (meas.empty()) return ABYSS_ERR_INVALID_ARG;
    abyss::Table t;
    for (int b = 0; b < abyss::kBuckets; ++b) {
        const Meas *nearest = &meas[0];
        for (auto &mm : meas)
            if (abs(mm.bucket - b) < abs(nearest->bucket - b)) nearest = &mm;
        t.e[b] = nearest->best;
    }
    t.valid = true;
    {
        std::lock_guard<std::mutex> lk(abyss::g_mu);
        abyss::g_tab[device] = t;
    }
    if (path) {
        FILE *f = fopen(path, "w");
        if (!f) return ABYSS_ERR_INVALID_ARG;
        fprintf(f, "# abyss_regf plan table: bucket(floor log2 targets) lanes_per_target "
                   "j_unroll workgroups_per_cu j_tile waves_per_workgroup  (tuned at n=%d)\n", n);
        for (int b = 0; b < abyss::kBuckets; ++b)
            fprintf(f, "%d %d %d %d %d %d\n", b, t.e[b].jl, t.e[b].unroll, t.e[b].wg_per_cu,
                    t.e[b].tile, t.e[b].wpb);
        fclose(f);
    }
    return ABYSS_OK;
}

}  // extern "C"

extern "C" int abyss_autotune(int device, int n, const char *path)
{
    if (device < 0 || device >= abyss::kMaxDevices || n < 256) return ABYSS_ERR_INVALID_ARG;
    // synthetic cluster: uniform unit cube, r_nb for ~48 neighbours
    std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n), r2(n);
    abyss::Rng rng{0x9E3779B97F4A7C15ull};
    for (size_t k = 0; k < 3 * (size_t)n; ++k) {
        x[k] = rng.next();
        v[k] = rng.next() - 0.5;
    }
    const double rnb = cbrt(48.0 / (n * 4.18879020478639));
    std::fill(r2.begin(), r2.end(), rnb * rnb);
    return abyss_autotune_jset(device, n, m.data(), x.data(), v.data(), r2.data(), path);
}
