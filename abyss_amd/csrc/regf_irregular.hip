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

#include <algorithm>

namespace {

__global__ __launch_bounds__(256) void fill_int_kernel(int *__restrict__ p, size_t count, int v)
{
    for (size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x; k < count;
         k += (size_t)gridDim.x * blockDim.x)
        p[k] = v;
}

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

__global__ __launch_bounds__(256) void irregular_kernel(abyss::JState js,
                                                        const int *__restrict__ tgt, int nt,
                                                        const int *__restrict__ rnnb,
                                                        const int *__restrict__ rlist, int cap,
                                                        double *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= nt) return;   // wave-uniform; no barriers below
    const int i = tgt[q];
    const int c = rnnb[i];
    const int n = min(c < 0 ? -c : c, cap);
    const double xi = js.x[i], yi = js.y[i], zi = js.z[i];
    const double vxi = js.vx[i], vyi = js.vy[i], vzi = js.vz[i];
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int e = lane; e < n; e += 64) {
        const int j = rlist[(size_t)i * cap + e];
        const double x = js.x[j] - xi, y = js.y[j] - yi, z = js.z[j] - zi;
        const double vx = js.vx[j] - vxi, vy = js.vy[j] - vyi, vz = js.vz[j] - vzi;
        const double r2 = x * x + y * y + z * z;
        if (r2 == 0.0) continue;
        const double rv = vx * x + vy * y + vz * z;
        const double m_r3 = js.m[j] / (r2 * sqrt(r2));
        a[0] += m_r3 * x;
        a[1] += m_r3 * y;
        a[2] += m_r3 * z;
        a[3] += m_r3 * (vx - 3 * x * rv / r2);
        a[4] += m_r3 * (vy - 3 * y * rv / r2);
        a[5] += m_r3 * (vz - 3 * z * rv / r2);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int k = 0; k < 6; ++k) a[k] += __shfl_xor(a[k], d, 64);
    if (lane < 6) out[(size_t)q * 6 + lane] = a[lane];
}

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
                       c->d_tgt, c->nt, c->d_nnb, c->d_list, c->cap, rnnb, rlist);
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
    const size_t need = tb + (size_t)nt * 6 * sizeof(double);
    if (need > c->irr_bytes) {
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
    }
    memcpy(c->h_irr, targets, (size_t)nt * sizeof(int));
    if (hipMemcpyAsync(c->d_irr, c->h_irr, (size_t)nt * sizeof(int), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    const int *rnnb = reinterpret_cast<const int *>(c->d_rlists);
    const int *rlist = rnnb + (((size_t)c->n + 63) & ~size_t(63));
    double *d_out = reinterpret_cast<double *>(c->d_irr + tb);
    hipLaunchKernelGGL(irregular_kernel, dim3((nt + 3) / 4), dim3(256), 0, c->stream, c->js,
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

}  // extern "C"
