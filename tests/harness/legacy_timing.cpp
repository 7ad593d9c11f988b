// Wall-clock time of the reference's calling pattern through the six legacy
// entry points, from C as ABYSS calls them: SendToDevice of the whole j-set,
// then CalculateAccelerationOnDevice of one active block (the reference's
// src_mpi/cuda/CalculateRegularAcceleration.cpp:151,439).  Used by
// bench.py's active-block sweep; reads the workload from a file it writes:
//   int32 n; double m[n], x[n][3], v[n][3], r2[n];
//   int32 nsizes; per size: int32 m; int32 targets[m].
// Prints one JSON line per block size with the medians (us) of the upload
// call, of the query call and of both, over `reps` timed pairs, at the
// given precision mode (ABYSS_PREC_*).
#include "abyss_regf.h"

#include <stdio.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <vector>

namespace {

template <typename T>
bool read_n(FILE *f, T *p, size_t n)
{
    return fread(p, sizeof(T), n, f) == n;
}

double median(std::vector<double> v)
{
    std::sort(v.begin(), v.end());
    return v[v.size() / 2];
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 4) {
        fprintf(stderr, "usage: legacy_timing <workload file> <cap> <precision> [reps]\n");
        return 2;
    }
    const int cap = atoi(argv[2]);
    const int prec = atoi(argv[3]);
    const int reps = argc > 4 ? atoi(argv[4]) : 200;
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int n = 0;
    if (!read_n(f, &n, 1) || n <= 0) return 2;
    std::vector<double> m(n), x(3 * (size_t)n), v(3 * (size_t)n), r2(n), mdot(n, 0.0);
    if (!read_n(f, m.data(), m.size()) || !read_n(f, x.data(), x.size()) ||
        !read_n(f, v.data(), v.size()) || !read_n(f, r2.data(), r2.size()))
        return 2;
    int nsizes = 0;
    if (!read_n(f, &nsizes, 1) || nsizes <= 0) return 2;
    std::vector<std::vector<int>> blocks(nsizes);
    for (auto &b : blocks) {
        int mt = 0;
        if (!read_n(f, &mt, 1) || mt <= 0 || mt > n) return 2;
        b.resize(mt);
        if (!read_n(f, b.data(), b.size())) return 2;
    }
    fclose(f);

    if (abyss_set_num_neighbor_max(cap) != ABYSS_OK || abyss_set_precision(prec) != ABYSS_OK)
        return 2;
    int rank = 0;
    OpenDevice(&rank);
    for (const auto &b : blocks) {
        const int mt = (int)b.size();
        std::vector<double> acc(3 * (size_t)mt), adot(3 * (size_t)mt);
        std::vector<int> nnb(mt), list((size_t)mt * cap);
        std::vector<double> ts, tc, tt;
        for (int r = 0; r < reps + 10; ++r) {
            int nn = n, nt = mt;
            const auto t0 = std::chrono::steady_clock::now();
            SendToDevice(&nn, m.data(), reinterpret_cast<double(*)[3]>(x.data()),
                         reinterpret_cast<double(*)[3]>(v.data()), r2.data(), mdot.data());
            const auto t1 = std::chrono::steady_clock::now();
            CalculateAccelerationOnDevice(&nt, const_cast<int *>(b.data()),
                                          reinterpret_cast<double(*)[3]>(acc.data()),
                                          reinterpret_cast<double(*)[3]>(adot.data()),
                                          nnb.data(), list.data());
            const auto t2 = std::chrono::steady_clock::now();
            if (abyss_last_error() != ABYSS_OK && abyss_last_error() != ABYSS_ERR_NEIGHBOR_OVERFLOW)
                return 1;
            if (r < 10) continue;   // warm-up
            ts.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
            tc.push_back(std::chrono::duration<double, std::micro>(t2 - t1).count());
            tt.push_back(std::chrono::duration<double, std::micro>(t2 - t0).count());
        }
        printf("{\"m\": %d, \"send_us\": %.2f, \"call_us\": %.2f, \"total_us\": %.2f}\n", mt,
               median(ts), median(tc), median(tt));
        fflush(stdout);
    }
    CloseDevice();
    return 0;
}
