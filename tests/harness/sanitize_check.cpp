// Host code under the sanitizers (built by tests/test_sanitize_cpu.py with
// AddressSanitizer + UndefinedBehaviorSanitizer, and ThreadSanitizer for the
// pack): the j-set validation and pack of an upload (regf_pack.cpp) against
// a plain serial restatement, its error reporting, concurrent callers; and
// the oracle's C restatement on a small random j-set.
//   sanitize_check pack     -- pack checks only
//   sanitize_check all      -- pack + oracle
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <random>
#include <thread>
#include <vector>

#include "../../abyss_amd/csrc/regf_pack.h"

extern "C" {
int oracle_regf(int n, const double *x, const double *v, const double *m, const double *rnb2,
                int nt, const int *targets, double *acc, double *adot, int *nnb, int *list,
                int cap, double *scale_a, double *scale_j, int nthreads);
void oracle_band_pairs(int n, const double *x, const double *rnb2, int nt, const int *targets,
                       double rel, int *bcnt, int *blist, int bcap, int nthreads);
void oracle_predict2(int n, const double *x, const double *v, const double *a,
                     const double *adot, const double *dt, double *xp, double *vp);
void oracle_irregular(int n, const double *x, const double *v, const double *m, int nt,
                      const int *targets, const int *nnb, const int *list, int cap, double *acc,
                      double *adot);
void oracle_update_radius(int count, const int *nnb, int fix, int nmax, double *r2);
}

static int fails = 0;
#define CHECK(c)                                                       \
    do {                                                               \
        if (!(c)) {                                                    \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            ++fails;                                                   \
        }                                                              \
    } while (0)

struct JSet {
    int n;
    std::vector<double> m, x, v, r2;
};

static JSet make(int n, unsigned seed)
{
    std::mt19937_64 g(seed);
    std::normal_distribution<double> nd;
    JSet s{n, std::vector<double>(n), std::vector<double>(3 * n), std::vector<double>(3 * n),
           std::vector<double>(n)};
    for (int j = 0; j < n; ++j) {
        s.m[j] = 1.0 / n * (0.5 + std::abs(nd(g)));
        s.r2[j] = 0.01 * (0.5 + std::abs(nd(g)));
        for (int d = 0; d < 3; ++d) {
            s.x[3 * j + d] = nd(g);
            s.v[3 * j + d] = 0.5 * nd(g);
        }
    }
    return s;
}

static void check_pack(int n, unsigned seed)
{
    JSet s = make(n, seed);
    std::vector<double> soa(8 * (size_t)n, -1.0);
    std::vector<float> m32(n, -1.0f);
    abyss::PackStatus st = abyss::pack_jset(n, 0, n, s.m.data(), s.x.data(), s.v.data(),
                                            s.r2.data(), soa.data(), m32.data());
    CHECK(st.index == -1);
    for (int j = 0; j < n; ++j) {
        for (int d = 0; d < 3; ++d) {
            CHECK(soa[(size_t)d * n + j] == s.x[3 * j + d]);
            CHECK(soa[(size_t)(3 + d) * n + j] == s.v[3 * j + d]);
        }
        CHECK(soa[6 * (size_t)n + j] == s.m[j]);
        CHECK(soa[7 * (size_t)n + j] == s.r2[j]);
        CHECK(m32[j] == (float)s.m[j]);
    }
    // one bad value at a time, then all three: the lowest index is reported
    const int bad[3] = {n - 1, n / 2, n / 3 + 1};
    const int kinds[3] = {abyss::kPackBadRadius, abyss::kPackBadState, abyss::kPackBadMass};
    auto spoil = [&](int k, bool on) {
        if (k == 0) s.r2[bad[0]] = on ? -1.0 : 0.01;
        if (k == 1) s.v[3 * bad[1] + 2] = on ? NAN : 0.0;
        if (k == 2) s.m[bad[2]] = on ? -2.0 : 1.0;
    };
    auto pack = [&] {
        return abyss::pack_jset(n, 0, n, s.m.data(), s.x.data(), s.v.data(), s.r2.data(),
                                soa.data(), m32.data());
    };
    // the staged form: the callback sees every position packed (it reads
    // them, as the upload's copy does, while the rest is being packed)
    int calls = 0;
    auto staged = [&] {
        calls = 0;
        return abyss::pack_jset_staged(n, s.m.data(), s.x.data(), s.v.data(), s.r2.data(),
                                       soa.data(), m32.data(), [&] {
                                           ++calls;
                                           for (int j = 0; j < n; ++j)
                                               for (int d = 0; d < 3; ++d)
                                                   CHECK(soa[(size_t)d * n + j] ==
                                                         s.x[3 * j + d]);
                                       });
    };
    for (int k = 0; k < 3; ++k) {
        spoil(k, true);
        st = pack();
        CHECK(st.index == bad[k] && st.kind == kinds[k]);
        const abyss::PackStatus st2 = staged();
        CHECK(st2.index == st.index && st2.kind == st.kind && calls == 1);
        spoil(k, false);
    }
    // a bad position: no callback, and the lowest index over both groups
    s.x[3 * (n / 4) + 1] = INFINITY;
    s.m[n / 2] = -1.0;
    st = staged();
    CHECK(st.index == n / 4 && st.kind == abyss::kPackBadState && calls == 0);
    s.m[n / 4] = -1.0;   // the same index: the position is reported
    st = staged();
    CHECK(st.index == n / 4 && st.kind == abyss::kPackBadState && calls == 0);
    s.x[3 * (n / 4) + 1] = 0.0;
    st = staged();
    CHECK(st.index == n / 4 && st.kind == abyss::kPackBadMass && calls == 1);
    s.m[n / 4] = 1.0;
    s.m[n / 2] = 1.0;
    {
        CHECK(pack().index == -1);
        const std::vector<double> one = soa;
        const std::vector<float> one32 = m32;
        std::fill(soa.begin(), soa.end(), -1.0);
        std::fill(m32.begin(), m32.end(), -1.0f);
        CHECK(staged().index == -1 && calls == 1);
        CHECK(soa == one && m32 == one32);   // the contents of one pass
    }
    for (int k = 0; k < 3; ++k) spoil(k, true);
    st = pack();
    CHECK(st.index == bad[2] && st.kind == kinds[2]);   // n/3 + 1 is the lowest
    for (int k = 0; k < 3; ++k) spoil(k, false);
    CHECK(pack().index == -1);
    // in consecutive ranges (as an upload pipelines them): the same contents
    std::vector<double> soa2(8 * (size_t)n, -1.0);
    std::vector<float> m322(n, -1.0f);
    for (int j0 = 0; j0 < n; j0 += 20000) {
        const int j1 = std::min(n, j0 + 20000);
        CHECK(abyss::pack_jset(n, j0, j1, s.m.data(), s.x.data(), s.v.data(), s.r2.data(),
                               soa2.data(), m322.data()).index == -1);
    }
    CHECK(soa2 == soa && m322 == m32);
}

static void check_pack_concurrent()
{
    // four callers at once (the pool serves one; the others pack alone)
    std::vector<std::thread> th;
    for (int t = 0; t < 4; ++t) th.emplace_back([t] { check_pack(50000 + 1000 * t, 100 + t); });
    for (auto &x : th) x.join();
}

static void check_oracle()
{
    const int n = 300, cap = 64;
    JSet s = make(n, 7);
    std::vector<int> tgt;
    for (int k = 0; k < n; k += 3) tgt.push_back(k);
    const int nt = (int)tgt.size();
    std::vector<double> acc(3 * nt), adot(3 * nt), sa(nt), sj(nt);
    std::vector<int> nnb(nt), list((size_t)nt * cap);
    oracle_regf(n, s.x.data(), s.v.data(), s.m.data(), s.r2.data(), nt, tgt.data(), acc.data(),
                adot.data(), nnb.data(), list.data(), cap, sa.data(), sj.data(), 2);
    for (int k = 0; k < nt; ++k) CHECK(isfinite(acc[3 * k]) && nnb[k] >= -n && nnb[k] < n);
    std::vector<int> bcnt(nt), blist((size_t)nt * 8);
    oracle_band_pairs(n, s.x.data(), s.r2.data(), nt, tgt.data(), 1e-6, bcnt.data(),
                      blist.data(), 8, 2);
    std::vector<double> a(3 * n, 0.1), ad(3 * n, 0.2), dt(n, 1.0 / 64), xp(3 * n), vp(3 * n);
    oracle_predict2(n, s.x.data(), s.v.data(), a.data(), ad.data(), dt.data(), xp.data(),
                    vp.data());
    std::vector<double> iacc(3 * nt), iadot(3 * nt);
    oracle_irregular(n, s.x.data(), s.v.data(), s.m.data(), nt, tgt.data(), nnb.data(),
                     list.data(), cap, iacc.data(), iadot.data());
    std::vector<double> r2 = s.r2;
    std::vector<int> cnt(n, 50);
    oracle_update_radius(n, cnt.data(), 100, 1000, r2.data());
    CHECK(isfinite(xp[0]) && isfinite(iacc[0]) && isfinite(r2[0]));
}

int main(int argc, char **argv)
{
    const bool all = argc < 2 || strcmp(argv[1], "all") == 0;
    check_pack(1000, 1);             // serial
    check_pack(70001, 2);            // split over the pool
    check_pack_concurrent();
    if (all) check_oracle();
    printf("%s: %d failure(s)\n", fails ? "FAIL" : "OK", fails);
    return fails ? 1 : 0;
}
