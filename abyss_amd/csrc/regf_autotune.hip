// Plan autotuner (include/abyss_regf.h, abyss_autotune*).
//
// For target counts m = 1, 4, 16, ... up to n on a synthetic n-particle
// cluster (uniform cube, r_nb for ~48 neighbours), every candidate plan is
// timed with HIP events (median of 3 launches):
//   lanes per target  JL in {1, 4, 16, 64} (JL = 1 also with j-unroll 2, 4)
//   workgroups per CU W in {1, 2, 4, 8, 16, 32, 48, 64, 96}, which sets the j-split
//                     count S = ceil(W * CUs / target blocks)
// The fastest (JL, unroll, W) per power-of-two target-count bucket is kept in
// a per-device table that plan selection consults for every later query on
// that device (buckets between measured points take the nearest one).  The
// table can be written to and read from a small text file.  Batched
// ensembles take the lanes-per-target choice from the same table.
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
        int b, jl, un, w;
        if (line[0] == '#') continue;
        if (sscanf(line, "%d %d %d %d", &b, &jl, &un, &w) == 4 && b >= 0 && b < abyss::kBuckets &&
            (jl == 1 || jl == 4 || jl == 16 || jl == 64) && (un == 2 || un == 4) && w > 0) {
            t.e[b] = abyss::TuneEntry{jl, un, w};
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

int abyss_autotune(int device, int n, const char *path)
{
    if (device < 0 || device >= abyss::kMaxDevices || n < 256) return ABYSS_ERR_INVALID_ARG;
    abyss_ctx_t c = nullptr;
    int rc = abyss_ctx_create(device, 256, &c);
    if (rc != ABYSS_OK) return rc;
    // synthetic cluster: uniform unit cube, r_nb for ~48 neighbours
    std::vector<double> m(n, 1.0 / n), x(3 * (size_t)n), v(3 * (size_t)n), r2(n);
    abyss::Rng rng{0x9E3779B97F4A7C15ull};
    for (size_t k = 0; k < 3 * (size_t)n; ++k) {
        x[k] = rng.next();
        v[k] = rng.next() - 0.5;
    }
    const double rnb = cbrt(48.0 / (n * 4.18879020478639));
    std::fill(r2.begin(), r2.end(), rnb * rnb);
    rc = abyss_ctx_upload(c, n, m.data(), x.data(), v.data(), r2.data());
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
    const int cu = c->cu_count > 0 ? c->cu_count : 256;
    std::vector<int> counts;
    for (int mt = 1; mt < n; mt *= 4) counts.push_back(mt);
    counts.push_back(n);
    for (int mt : counts) {
        if (rc != ABYSS_OK) break;
        rc = abyss_ctx_set_targets(c, mt, idx.data());
        if (rc != ABYSS_OK) break;
        double best_ms = 1e30;
        abyss::TuneEntry best{1, 2, 16};
        const int jls[] = {1, 4, 16, 64};
        for (int jl : jls) {
            for (int un : {2, 4}) {
                if (jl > 1 && un == 4) continue;
                int prev_S = -1;
                for (int w : {1, 2, 4, 8, 16, 32, 48, 64, 96}) {
                    const int tpb = jl > 1 ? 4 * (64 / jl) : abyss::kBlock;
                    const int gx = (mt + tpb - 1) / tpb;
                    const int S = std::max(1, std::min((w * cu + gx - 1) / gx,
                                                       std::min(n / abyss::kTile,
                                                                abyss::kMaxSplits)));
                    if (S == prev_S) continue;   // same plan as the previous w
                    prev_S = S;
                    rc = abyss_ctx_set_tuning(c, 1, S, un, jl);
                    if (rc != ABYSS_OK) break;
                    rc = abyss_ctx_launch(c);   // warm-up (also sizes the scratch)
                    if (rc != ABYSS_OK) break;
                    float t[3];
                    for (int r = 0; r < 3 && rc == ABYSS_OK; ++r) {
                        if (hipEventRecord(e0, c->stream) != hipSuccess) rc = ABYSS_ERR_HIP;
                        if (rc == ABYSS_OK) rc = abyss_ctx_launch(c);
                        if (rc == ABYSS_OK && hipEventRecord(e1, c->stream) != hipSuccess)
                            rc = ABYSS_ERR_HIP;
                        if (rc == ABYSS_OK && hipEventSynchronize(e1) != hipSuccess)
                            rc = ABYSS_ERR_HIP;
                        if (rc == ABYSS_OK && hipEventElapsedTime(&t[r], e0, e1) != hipSuccess)
                            rc = ABYSS_ERR_HIP;
                    }
                    if (rc != ABYSS_OK) break;
                    std::sort(t, t + 3);
                    if (t[1] < best_ms) {
                        best_ms = t[1];
                        best = abyss::TuneEntry{jl, un, w};
                    }
                }
                if (rc != ABYSS_OK) break;
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
                   "j_unroll workgroups_per_cu  (tuned at n=%d)\n", n);
        for (int b = 0; b < abyss::kBuckets; ++b)
            fprintf(f, "%d %d %d %d\n", b, t.e[b].jl, t.e[b].unroll, t.e[b].wg_per_cu);
        fclose(f);
    }
    return ABYSS_OK;
}

}  // extern "C"
