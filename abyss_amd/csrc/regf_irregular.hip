// Irregular (neighbour) force on the device over resident neighbour lists
// (include/abyss_regf.h: abyss_ctx_keep_lists, abyss_ctx_irregular).
//
// ABYSS computes the irregular force on the host at every irregular step:
// for particle i, the FP64 sum over its neighbour list of
// m_j r/|r|^3 and m_j (v - 3 r (r.v)/r^2)/|r|^3 with m_r3 = m/(r2*sqrt(r2))
// (Particle::computeAccelerationIrr, src_mpi/Particle/ComputeAcceleration.cpp
// :14-190; the list comes from the last regular step).  Here the lists of
// every launch are kept resident per particle (row = particle index), and
// any block of particles gets its irregular force and jerk from one kernel:
// one wave per particle, lanes over its list entries (lane l takes entries
// l, l+64, ...), partial sums combined by a fixed xor-butterfly, so results
// are bitwise reproducible.  A coincident neighbour (r = 0) adds nothing,
// as in the regular kernels.
#include "regf_context.h"

#include <stdio.h>
#include <string.h>

#include <cmath>

#include <algorithm>

namespace {

__global__ __launch_bounds__(256) void fill_int_kernel(int *__restrict__ p, size_t count, int v)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = v;
}

// Fourth-order regular corrector of one dimension, in the reference's
// expression order, contraction off (out: x_new, v_new, a2, a3).
#pragma clang fp contract(off)
__device__ __forceinline__ void corrector(double dt, double a0, double j0, double a1, double j1,
                                       double a_tmp, double adot_tmp, double x, double v,
                                       double *o)
{
    const double dt2 = dt * dt, dt3 = dt2 * dt, dt4 = dt3 * dt, dt5 = dt4 * dt;
    const double da_dt2 = (a0 - a1 - a_tmp) / dt2;
    const double adot_dt = (j0 + j1 + adot_tmp) / dt;
    const double a2 = -6 * da_dt2 - 2 * adot_dt - 2 * j0 / dt;
    const double a3 = (12 * da_dt2 + 6 * adot_dt) / dt;
    o[0] = x + a2 * dt4 / 24 + a3 * dt5 / 120;
    o[1] = v + a2 * dt3 / 6 + a3 * dt4 / 24;
    o[2] = a2;
    o[3] = a3;
}
#pragma clang fp contract(on)

// Copies the lists of one launch (rows by target position) into the
// per-particle table (rows by particle index).  One wave per target.
__global__ __launch_bounds__(256) void keep_lists_kernel(const int *__restrict__ tgt, int nt,
                                                         const int *__restrict__ nnb,
                                                         const int *__restrict__ list, int cap,
                                                         int *__restrict__ rnnb,
                                                         int *__restrict__ rlist)
{
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nt) return;
    const int i = tgt[q];
    const int c = nnb[q];
    const int n = min(c < 0 ? -c : c, cap);
    for (int e = lane; e < n; e += 64) rlist[(size_t)i * cap + e] = list[(size_t)q * cap + e];
    if (lane == 0) rnnb[i] = c;
}

struct __align__(64) JRec {
    double x, y, z, vx, vy, vz, m, pad;
};

__global__ __launch_bounds__(256) void pack_aos_kernel(abyss::JState js, int n,
                                                       JRec *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = JRec{js.x[i], js.y[i], js.z[i], js.vx[i], js.vy[i], js.vz[i], js.m[i], 0.0};
}

// Irregular-force term of j on a target at (xi, vi), FP64: t = (acc, jerk);
// a coincident j (r2 = 0) gives zero.  Every kernel adds whole terms with
// contraction off, so the same list in the same lane order gives the same
// bits in every kernel, and each term is the plain IEEE evaluation.
#pragma clang fp contract(off)
__device__ __forceinline__ void term(const JRec *__restrict__ jr, int j, const double *xi,
                                     const double *vi, double *t)
{
    const JRec r = jr[j];
    const double x = r.x - xi[0], y = r.y - xi[1], z = r.z - xi[2];
    const double vx = r.vx - vi[0], vy = r.vy - vi[1], vz = r.vz - vi[2];
    const double r2 = x * x + y * y + z * z;
    if (r2 == 0.0) {
#pragma unroll
        for (int k = 0; k < 6; ++k) t[k] = 0.0;
        return;
    }
    const double rv = vx * x + vy * y + vz * z;
    const double m_r3 = r.m / (r2 * sqrt(r2));
    t[0] = m_r3 * x;
    t[1] = m_r3 * y;
    t[2] = m_r3 * z;
    t[3] = m_r3 * (vx - 3 * x * rv / r2);
    t[4] = m_r3 * (vy - 3 * y * rv / r2);
    t[5] = m_r3 * (vz - 3 * z * rv / r2);
}

template <int K>
__device__ __forceinline__ void wave_sum(double *a)
{
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int k = 0; k < K; ++k) a[k] += __shfl_xor(a[k], d, 64);
}

// Ascending list membership (lists come from the force kernels, which
// emit neighbours in ascending j).
__device__ __forceinline__ bool in_list(const int *__restrict__ l, int n, int j)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (l[mid] < j) lo = mid + 1;
        else hi = mid;
    }
    return lo < n && l[lo] == j;
}

__device__ __forceinline__ double pick(double v0, double v1, double v2, int d)
{
    return d == 0 ? v0 : (d == 1 ? v1 : v2);
}

__device__ __forceinline__ int list_len(int c, int cap)
{
    return min(c < 0 ? -c : c, cap);
}

__global__ __launch_bounds__(256) void irregular_kernel(const JRec *__restrict__ jr,
                                                        const int *__restrict__ tgt, int nt,
                                                        const int *__restrict__ rnnb,
                                                        const int *__restrict__ rlist, int cap,
                                                        double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nt) return;   // wave-uniform; no barriers below
    const int i = tgt[q];
    const int n = list_len(rnnb[i], cap);
    const JRec ri = jr[i];
    const double xi[3] = {ri.x, ri.y, ri.z};
    const double vi[3] = {ri.vx, ri.vy, ri.vz};
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int e = lane; e < n; e += 64) {
        double t[6];
        term(jr, rlist[(size_t)i * cap + e], xi, vi, t);
#pragma unroll
        for (int k = 0; k < 6; ++k) a[k] += t[k];
    }
    wave_sum<6>(a);
    if (lane < 6) out[(size_t)q * 6 + lane] = a[lane];
}

// Neighbour-set correction and fourth-order regular corrector of one
// regular step (Particle::updateRegularParticleCuda,
// src_mpi/Particle/ComputeAcceleration.cpp:400-661), one wave per target.
// L_new = this launch's list (row q), L_old = the resident list of i.
//   a_irr = sum over L_new;  a_tmp = sum over L_new \ L_old - sum over L_old \ L_new
// then, per dimension, with a0/j0 the previous regular force/jerk and
// a1/j1 this launch's:
//   da_dt2 = (a0 - a1 - a_tmp)/dt^2,  adot_dt = (j0 + j1 + adot_tmp)/dt
//   a2 = -6 da_dt2 - 2 adot_dt - 2 j0/dt,  a3 = (12 da_dt2 + 6 adot_dt)/dt
//   x_new = x + a2 dt^4/24 + a3 dt^5/120,  v_new = v + a2 dt^3/6 + a3 dt^4/24
// out row (18): x_new, v_new, a2, a3, a_irr, adot_irr.
// STAGED: both lists of the wave's target are first copied into LDS
// (dynamic, 2 * cap ints per wave), so the membership binary searches read
// LDS instead of paying global-memory latency at every step.
template <bool STAGED>
__global__ __launch_bounds__(256) void correct_kernel(
    const JRec *__restrict__ jr, const int *__restrict__ tgt, int nt, const int *__restrict__ new_nnb,
    const int *__restrict__ new_list, const double *__restrict__ new_acc,
    const double *__restrict__ new_adot, int cap, const int *__restrict__ rnnb,
    const int *__restrict__ rlist, const double *__restrict__ dts,
    const double *__restrict__ a_old, double *__restrict__ out)
{
    extern __shared__ int s_lists[];   // STAGED: [4 waves][new cap | old cap]
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    const bool active = q < nt;
    const int i = active ? tgt[q] : 0;
    const int cn = active ? list_len(new_nnb[q], cap) : 0;
    const int co = active ? list_len(rnnb[i], cap) : 0;
    const int *ln = new_list + (size_t)q * cap;
    const int *lo = rlist + (size_t)i * cap;
    if constexpr (STAGED) {
        int *sn = s_lists + (threadIdx.x >> 6) * 2 * cap;
        int *so = sn + cap;
        for (int e = lane; e < cn; e += 64) sn[e] = ln[e];
        for (int e = lane; e < co; e += 64) so[e] = lo[e];
        __syncthreads();   // every wave of the block reaches it
        ln = sn;
        lo = so;
    }
    if (!active) return;   // wave-uniform; no barriers below
    const JRec ri = jr[i];
    const double xi[3] = {ri.x, ri.y, ri.z};
    const double vi[3] = {ri.vx, ri.vy, ri.vz};
    double a[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // a_irr[6], a_tmp[6]
    for (int e = lane; e < cn; e += 64) {
        const int j = ln[e];
        double t[6];
        term(jr, j, xi, vi, t);
#pragma unroll
        for (int k = 0; k < 6; ++k) a[k] += t[k];
        if (!in_list(lo, co, j)) {
#pragma unroll
            for (int k = 0; k < 6; ++k) a[6 + k] += t[k];
        }
    }
    for (int e = lane; e < co; e += 64) {
        const int j = lo[e];
        if (!in_list(ln, cn, j)) {
            double t[6];
            term(jr, j, xi, vi, t);
#pragma unroll
            for (int k = 0; k < 6; ++k) a[6 + k] -= t[k];
        }
    }
    wave_sum<12>(a);
    if (lane < 3) {
        // lane d takes dimension d (selects, not a lane-indexed array)
        const int d = lane;
        const double dt = dts[q];
        const double a0 = a_old[(size_t)q * 6 + d], j0 = a_old[(size_t)q * 6 + 3 + d];
        const double a1 = new_acc[(size_t)q * 3 + d], j1 = new_adot[(size_t)q * 3 + d];
        const double irr = pick(a[0], a[1], a[2], d), irr_dot = pick(a[3], a[4], a[5], d);
        const double tmp = pick(a[6], a[7], a[8], d), tmp_dot = pick(a[9], a[10], a[11], d);
        double corr[4];
        corrector(dt, a0, j0, a1, j1, tmp, tmp_dot, pick(xi[0], xi[1], xi[2], d),
                  pick(vi[0], vi[1], vi[2], d), corr);
        double *o = out + (size_t)q * 18;
        o[d] = corr[0];
        o[3 + d] = corr[1];
        o[6 + d] = corr[2];
        o[9 + d] = corr[3];
        o[12 + d] = irr;
        o[15 + d] = irr_dot;
    }
}

#pragma clang fp contract(on)

int ensure_resident_lists(abyss_ctx_t c)
{
    const size_t n = (size_t)c->n;
    const size_t need = n * (size_t)c->cap * sizeof(int) + n * sizeof(int) + 256;
    if (c->rlist_n == c->n && c->rlist_bytes >= need) return ABYSS_OK;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    if (need > c->rlist_bytes) {
        if (c->d_rlists) (void)hipFree(c->d_rlists);
        c->d_rlists = nullptr;
        c->rlist_bytes = 0;
        if (hipMalloc(&c->d_rlists, need) != hipSuccess) return ABYSS_ERR_OUT_OF_MEMORY;
        c->rlist_bytes = need;
    }
    // new j-set size: every particle starts with an empty list
    int *rnnb = reinterpret_cast<int *>(c->d_rlists);
    hipLaunchKernelGGL(fill_int_kernel, dim3(std::min<size_t>((n + 255) / 256, 4096)), dim3(256),
                       0, c->stream, rnnb, n, 0);
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    c->rlist_n = c->n;
    return ABYSS_OK;
}

// The AoS gather copy of the current j-set (rebuilt after an upload or a
// prediction).
int ensure_aos(abyss_ctx_t c)
{
    const size_t need = (size_t)c->n * sizeof(JRec);
    if (c->aos_version == c->jset_version && c->aos_bytes >= need) return ABYSS_OK;
    if (need > c->aos_bytes) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
        if (c->d_aos) (void)hipFree(c->d_aos);
        c->d_aos = nullptr;
        c->aos_bytes = 0;
        if (hipMalloc(&c->d_aos, need) != hipSuccess) return ABYSS_ERR_OUT_OF_MEMORY;
        c->aos_bytes = need;
    }
    hipLaunchKernelGGL(pack_aos_kernel, dim3((c->n + 255) / 256), dim3(256), 0, c->stream,
                       c->js, c->n, reinterpret_cast<JRec *>(c->d_aos));
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    c->aos_version = c->jset_version;
    return ABYSS_OK;
}

// Device + pinned staging of the irregular / corrector entry points (the
// stream must be idle: the previous contents may still be in flight).
int ensure_staging(abyss_ctx_t c, size_t need)
{
    if (need <= c->irr_bytes) return ABYSS_OK;
    if (c->d_irr) (void)hipFree(c->d_irr);
    if (c->h_irr) (void)hipHostFree(c->h_irr);
    c->d_irr = nullptr;
    c->h_irr = nullptr;
    c->irr_bytes = 0;
    const size_t cap = need + need / 2 + 4096;
    if (hipMalloc(&c->d_irr, cap) != hipSuccess ||
        hipHostMalloc(&c->h_irr, cap, hipHostMallocDefault) != hipSuccess)
        return ABYSS_ERR_OUT_OF_MEMORY;
    c->irr_bytes = cap;
    return ABYSS_OK;
}

}  // namespace

namespace abyss {

// Called after every single-cluster launch when resident lists are on.
int keep_launch_lists(abyss_ctx_t c)
{
    if (!c->keep_lists) return ABYSS_OK;
    int rc = ensure_resident_lists(c);
    if (rc != ABYSS_OK) return rc;
    int *rnnb = reinterpret_cast<int *>(c->d_rlists);
    int *rlist = rnnb + (((size_t)c->n + 63) & ~size_t(63));
    hipLaunchKernelGGL(keep_lists_kernel, dim3((c->nt + 3) / 4), dim3(256), 0, c->stream,
                       c->tgt, c->nt, c->d_nnb, c->d_list, c->cap, rnnb, rlist);
    return hipGetLastError() == hipSuccess ? ABYSS_OK : ABYSS_ERR_HIP;
}

}  // namespace abyss

extern "C" {

int abyss_ctx_keep_lists(abyss_ctx_t c, int enable)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    if (!enable) {
        if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
        if (c->d_rlists) (void)hipFree(c->d_rlists);
        c->d_rlists = nullptr;
        c->rlist_bytes = 0;
        c->rlist_n = 0;
        c->keep_lists = false;
        return ABYSS_OK;
    }
    c->keep_lists = true;
    if (c->n > 0) return ensure_resident_lists(c);
    return ABYSS_OK;
}

int abyss_ctx_irregular(abyss_ctx_t c, int nt, const int *targets, double *acc, double *adot)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!c->keep_lists || c->n <= 0) {
        snprintf(c->detail, sizeof(c->detail), "irregular: resident lists are off or no j-set");
        return ABYSS_ERR_NO_JSET;
    }
    if (nt <= 0 || !targets || !acc || !adot) {
        snprintf(c->detail, sizeof(c->detail), "irregular: nt=%d or null array", nt);
        return ABYSS_ERR_INVALID_ARG;
    }
    for (int k = 0; k < nt; ++k)
        if (targets[k] < 0 || targets[k] >= c->n) {
            snprintf(c->detail, sizeof(c->detail), "irregular: target %d outside [0,%d)",
                     targets[k], c->n);
            return ABYSS_ERR_TARGET_RANGE;
        }
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    int rc = ensure_resident_lists(c);   // a j-set of another size: empty lists
    if (rc != ABYSS_OK) return rc;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    const size_t tb = (((size_t)nt * sizeof(int)) + 255) & ~size_t(255);
    rc = ensure_staging(c, tb + (size_t)nt * 6 * sizeof(double));
    if (rc != ABYSS_OK) return rc;
    memcpy(c->h_irr, targets, (size_t)nt * sizeof(int));
    if (hipMemcpyAsync(c->d_irr, c->h_irr, (size_t)nt * sizeof(int), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    const int *rnnb = reinterpret_cast<const int *>(c->d_rlists);
    const int *rlist = rnnb + (((size_t)c->n + 63) & ~size_t(63));
    double *d_out = reinterpret_cast<double *>(c->d_irr + tb);
    rc = ensure_aos(c);
    if (rc != ABYSS_OK) return rc;
    hipLaunchKernelGGL(irregular_kernel, dim3((nt + 3) / 4), dim3(256), 0, c->stream,
                       reinterpret_cast<const JRec *>(c->d_aos),
                       reinterpret_cast<const int *>(c->d_irr), nt, rnnb, rlist, c->cap, d_out);
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    double *h_out = reinterpret_cast<double *>(c->h_irr + tb);
    if (hipMemcpyAsync(h_out, d_out, (size_t)nt * 6 * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    for (int k = 0; k < nt; ++k)
        for (int d = 0; d < 3; ++d) {
            acc[3 * k + d] = h_out[6 * (size_t)k + d];
            adot[3 * k + d] = h_out[6 * (size_t)k + 3 + d];
        }
    return ABYSS_OK;
}

int abyss_ctx_regular_step(abyss_ctx_t c, int nt, const int *targets, const double *dt,
                           const double *a_old, double *acc, double *adot, int *nnb, int *list,
                           double *out)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (!c->keep_lists || c->n <= 0) {
        snprintf(c->detail, sizeof(c->detail), "regular_step: resident lists are off or no j-set");
        return ABYSS_ERR_NO_JSET;
    }
    if (nt <= 0 || !targets || !dt || !a_old || !acc || !adot || !nnb || !list || !out) {
        snprintf(c->detail, sizeof(c->detail), "regular_step: nt=%d or null array", nt);
        return ABYSS_ERR_INVALID_ARG;
    }
    for (int k = 0; k < nt; ++k) {
        if (!(dt[k] > 0.0) || !std::isfinite(dt[k])) {
            snprintf(c->detail, sizeof(c->detail), "regular_step: dt[%d] = %g", k, dt[k]);
            return ABYSS_ERR_INVALID_INPUT;
        }
        for (int e = 0; e < 6; ++e)
            if (!std::isfinite(a_old[6 * (size_t)k + e])) {
                snprintf(c->detail, sizeof(c->detail), "regular_step: non-finite a_old, k=%d", k);
                return ABYSS_ERR_INVALID_INPUT;
            }
    }
    int rc = abyss_ctx_set_targets(c, nt, targets);   // validates the targets
    if (rc != ABYSS_OK) return rc;
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    rc = ensure_resident_lists(c);
    if (rc != ABYSS_OK) return rc;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    // staging: dt [nt] | a_old [nt][6] | out [nt][18]
    const size_t in_bytes = (size_t)nt * 7 * sizeof(double);
    rc = ensure_staging(c, in_bytes + (size_t)nt * 18 * sizeof(double));
    if (rc != ABYSS_OK) return rc;
    double *h = reinterpret_cast<double *>(c->h_irr);
    memcpy(h, dt, (size_t)nt * sizeof(double));
    memcpy(h + nt, a_old, (size_t)nt * 6 * sizeof(double));
    if (hipMemcpyAsync(c->d_irr, c->h_irr, in_bytes, hipMemcpyHostToDevice, c->stream) !=
        hipSuccess)
        return ABYSS_ERR_HIP;
    rc = abyss::ctx_launch_force(c);
    if (rc != ABYSS_OK) return rc;
    const double *d_in = reinterpret_cast<const double *>(c->d_irr);
    double *d_out = reinterpret_cast<double *>(c->d_irr) + (size_t)nt * 7;
    const int *rnnb = reinterpret_cast<const int *>(c->d_rlists);
    const int *rlist = rnnb + (((size_t)c->n + 63) & ~size_t(63));
    rc = ensure_aos(c);
    if (rc != ABYSS_OK) return rc;
    // lists staged in LDS when both fit (2 * cap ints per wave, 64 KB per
    // workgroup at most)
    const size_t lds = 4 * 2 * (size_t)c->cap * sizeof(int);
    if (lds <= 65536)
        hipLaunchKernelGGL(correct_kernel<true>, dim3((nt + 3) / 4), dim3(256), lds, c->stream,
                           reinterpret_cast<const JRec *>(c->d_aos), c->tgt, nt, c->d_nnb,
                           c->d_list, c->d_acc, c->d_adot, c->cap, rnnb, rlist, d_in, d_in + nt,
                           d_out);
    else
        hipLaunchKernelGGL(correct_kernel<false>, dim3((nt + 3) / 4), dim3(256), 0, c->stream,
                           reinterpret_cast<const JRec *>(c->d_aos), c->tgt, nt, c->d_nnb,
                           c->d_list, c->d_acc, c->d_adot, c->cap, rnnb, rlist, d_in, d_in + nt,
                           d_out);
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    // the resident lists become this step's (after the correction read them)
    rc = abyss::keep_launch_lists(c);
    if (rc != ABYSS_OK) return rc;
    double *h_out = h + (size_t)nt * 7;
    if (hipMemcpyAsync(h_out, d_out, (size_t)nt * 18 * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    rc = abyss_ctx_fetch(c, acc, adot, nnb, list);   // synchronises the stream
    if (rc != ABYSS_OK && rc != ABYSS_ERR_NEIGHBOR_OVERFLOW) return rc;
    memcpy(out, h_out, (size_t)nt * 18 * sizeof(double));
    return rc;
}

}  // extern "C"
