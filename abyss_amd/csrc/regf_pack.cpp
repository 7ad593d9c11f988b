// Host side of a j-set upload: validation of the caller's arrays and the
// AoS -> SoA pack into the device layout, spread over a small pool of host
// threads for large N.
//
// Reference counterpart: the pack loop of _ReceiveFromHost
// (src/cuda/cuda_my_acceleration.cu:399-405, src_mpi/cuda/
// cuda_my_acceleration.cu:603-778), which copies m, x, v with stride N into
// pinned staging on one thread and validates nothing.
#include "regf_pack.h"

#include <math.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

namespace abyss {
namespace {

// Parts [1, parts) of a job run on parked worker threads, part 0 on the
// caller; run() returns when every part is done.  One job at a time: a
// caller that finds the pool busy does the whole job itself.
class Pool {
public:
    ~Pool()
    {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }

    // Workers available to a job (0: the pool could not start).
    int workers()
    {
        std::call_once(once_, [this] {
            const unsigned hw = std::thread::hardware_concurrency();
            const int want = (int)std::min<unsigned>(kMaxWorkers, hw > 1 ? hw - 1 : 0);
            for (int k = 0; k < want; ++k) th_.emplace_back([this, k] { loop(k + 1); });
        });
        return (int)th_.size();
    }

    bool run(int parts, const std::function<void(int)> &f)
    {
        std::unique_lock<std::mutex> busy(job_mu_, std::try_to_lock);
        if (!busy.owns_lock()) return false;
        parts = std::min(parts, workers() + 1);
        {
            std::lock_guard<std::mutex> lk(mu_);
            job_ = &f;
            parts_ = parts;
            left_ = parts - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        std::unique_lock<std::mutex> lk(mu_);
        done_cv_.wait(lk, [this] { return left_ == 0; });
        job_ = nullptr;
        return true;
    }

private:
    static constexpr unsigned kMaxWorkers = 7;

    void loop(int part)
    {
        long long seen = 0;
        for (;;) {
            const std::function<void(int)> *f = nullptr;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_.wait(lk, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                if (part >= parts_) continue;   // not needed for this job
                f = job_;
            }
            (*f)(part);
            std::lock_guard<std::mutex> lk(mu_);
            if (--left_ == 0) done_cv_.notify_one();
        }
    }

    std::once_flag once_;
    std::vector<std::thread> th_;
    std::mutex job_mu_;   // one job at a time
    std::mutex mu_;
    std::condition_variable cv_, done_cv_;
    const std::function<void(int)> *job_ = nullptr;
    int parts_ = 0, left_ = 0;
    long long gen_ = 0;
    bool stop_ = false;
};

Pool &pool()
{
    static Pool p;
    return p;
}

// Particles [j0, j1): the first invalid one, or index -1.
PackStatus pack_range(int n, int j0, int j1, const double *m, const double *x, const double *v,
                      const double *r2, double *soa, float *m32)
{
    const size_t N = (size_t)n;
    for (int j = j0; j < j1; ++j) {
        const double *xj = x + 3 * (size_t)j, *vj = v + 3 * (size_t)j;
        if (!(isfinite(xj[0]) && isfinite(xj[1]) && isfinite(xj[2]) && isfinite(vj[0]) &&
              isfinite(vj[1]) && isfinite(vj[2])))
            return {j, kPackBadState};
        if (!isfinite(m[j]) || m[j] < 0.0) return {j, kPackBadMass};
        if (isnan(r2[j]) || r2[j] < 0.0) return {j, kPackBadRadius};
        soa[0 * N + j] = xj[0];
        soa[1 * N + j] = xj[1];
        soa[2 * N + j] = xj[2];
        soa[3 * N + j] = vj[0];
        soa[4 * N + j] = vj[1];
        soa[5 * N + j] = vj[2];
        soa[6 * N + j] = m[j];
        soa[7 * N + j] = r2[j];
        m32[j] = (float)m[j];
    }
    return {};
}

// Particles [j0, j1): positions only (POS) or the rest only (!POS).
template <bool POS>
PackStatus pack_part(int n, int j0, int j1, const double *m, const double *x, const double *v,
                     const double *r2, double *soa, float *m32)
{
    const size_t N = (size_t)n;
    for (int j = j0; j < j1; ++j) {
        if constexpr (POS) {
            const double *xj = x + 3 * (size_t)j;
            if (!(isfinite(xj[0]) && isfinite(xj[1]) && isfinite(xj[2])))
                return {j, kPackBadState};
            soa[0 * N + j] = xj[0];
            soa[1 * N + j] = xj[1];
            soa[2 * N + j] = xj[2];
        } else {
            const double *vj = v + 3 * (size_t)j;
            if (!(isfinite(vj[0]) && isfinite(vj[1]) && isfinite(vj[2])))
                return {j, kPackBadState};
            if (!isfinite(m[j]) || m[j] < 0.0) return {j, kPackBadMass};
            if (isnan(r2[j]) || r2[j] < 0.0) return {j, kPackBadRadius};
            soa[3 * N + j] = vj[0];
            soa[4 * N + j] = vj[1];
            soa[5 * N + j] = vj[2];
            soa[6 * N + j] = m[j];
            soa[7 * N + j] = r2[j];
            m32[j] = (float)m[j];
        }
    }
    return {};
}

// The first failure of the two field groups of the same range, as one pass
// over all fields would report it: the lower index, and at an equal index
// the position (the first field checked).
PackStatus first_of(const PackStatus &pos, const PackStatus &rest)
{
    if (pos.index < 0) return rest;
    if (rest.index < 0 || pos.index <= rest.index) return pos;
    return rest;
}

// The first failure over parts in index order.
PackStatus first_part(const std::vector<PackStatus> &st)
{
    for (const PackStatus &s : st)
        if (s.index >= 0) return s;
    return {};
}

}  // namespace

PackStatus pack_jset_staged(int n, const double *m, const double *x, const double *v,
                            const double *r2, double *soa, float *m32,
                            const std::function<void()> &positions_done)
{
    const auto serial = [&] {
        const PackStatus a = pack_part<true>(n, 0, n, m, x, v, r2, soa, m32);
        if (a.index < 0) positions_done();
        return first_of(a, pack_part<false>(n, 0, n, m, x, v, r2, soa, m32));
    };
    if (n < kPackParallelMin) return serial();
    const int w = pool().workers();
    const int parts = std::min(w + 1, (n + kPackParallelMin / 2 - 1) / (kPackParallelMin / 2));
    if (parts <= 1) return serial();
    std::vector<PackStatus> pos(parts), rest(parts);
    std::atomic<int> packed{0};
    const auto job = [&](int p) {
        const int a = (int)((long long)n * p / parts);
        const int b = (int)((long long)n * (p + 1) / parts);
        pos[p] = pack_part<true>(n, a, b, m, x, v, r2, soa, m32);
        packed.fetch_add(1, std::memory_order_release);
        if (p == 0) {
            // only part 0 waits: the others go on to the rest, so every
            // part it waits for is running or done
            while (packed.load(std::memory_order_acquire) < parts) std::this_thread::yield();
            if (first_part(pos).index < 0) positions_done();
        }
        rest[p] = pack_part<false>(n, a, b, m, x, v, r2, soa, m32);
    };
    if (!pool().run(parts, job)) return serial();
    return first_of(first_part(pos), first_part(rest));
}

PackStatus pack_jset(int n, int j0, int j1, const double *m, const double *x, const double *v,
                     const double *r2, double *soa, float *m32)
{
    const int len = j1 - j0;
    if (len < kPackParallelMin) return pack_range(n, j0, j1, m, x, v, r2, soa, m32);
    const int w = pool().workers();
    const int parts = std::min(w + 1, (len + kPackParallelMin / 2 - 1) / (kPackParallelMin / 2));
    std::vector<PackStatus> st(parts);
    const auto job = [&](int p) {
        const int a = j0 + (int)((long long)len * p / parts);
        const int b = j0 + (int)((long long)len * (p + 1) / parts);
        st[p] = pack_range(n, a, b, m, x, v, r2, soa, m32);
    };
    if (parts <= 1 || !pool().run(parts, job)) return pack_range(n, j0, j1, m, x, v, r2, soa, m32);
    for (const PackStatus &s : st)   // parts in index order: the first is the lowest
        if (s.index >= 0) return s;
    return {};
}

}  // namespace abyss
