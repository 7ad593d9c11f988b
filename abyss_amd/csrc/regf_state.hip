// Device-resident particle state and j prediction (include/abyss_regf.h,
// abyss_ctx_upload_state / abyss_ctx_update_state / abyss_ctx_predict).
//
// The reference caller predicts every particle on the host, packs the
// arrays and copies all N of them before each regular step
// (src_mpi/cuda/CalculateRegularAcceleration.cpp:330-449, predictor
// src_mpi/Particle/UpdateParticle.cpp:37-59).  Here the state lives in HBM:
// the full state is uploaded once, after each corrector only the changed
// particles are scattered in, and a kernel predicts all j to the new time
// straight into the j-set the force kernels read.  The predictor evaluates
// the reference's expressions in the same order with FMA contraction off,
// so the predicted j-set is bitwise identical to a host prediction with
// plain IEEE double arithmetic.
#include "regf_context.h"

#include <stdio.h>
#include <string.h>

#include <cmath>

#include <algorithm>

namespace {

// State SoA (stride n): x[3][n], v[3][n], a[3][n], ad[3][n], t[n], m[n],
// r2[n] (all FP64).  Component-major so each kernel access is coalesced.
constexpr int kStateDoubles = 15;

struct StateView {
    const double *x, *v, *a, *ad, *t, *m, *r2;
};

StateView state_view(const double *base, int n)
{
    const size_t N = (size_t)n;
    return StateView{base, base + 3 * N, base + 6 * N, base + 9 * N, base + 12 * N, base + 13 * N,
                     base + 14 * N};
}

#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void predict_kernel(StateView st, int n, double t_new,
                                                      double *__restrict__ jx, float *__restrict__ m32)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t N = (size_t)n;
    const double h = t_new - st.t[i];
    for (int d = 0; d < 3; ++d) {
        const double x = st.x[d * N + i], v = st.v[d * N + i];
        double xp, vp;
        if (h == 0) {
            xp = x;
            vp = v;
        } else {
            const double a0 = st.a[d * N + i], a1 = st.ad[d * N + i];
            xp = ((a1 * h / 3 + a0) * h / 2 + v) * h + x;
            vp = (a1 * h / 2 + a0) * h + v;
        }
        jx[d * N + i] = xp;
        jx[(3 + d) * N + i] = vp;
    }
    jx[6 * N + i] = st.m[i];
    jx[7 * N + i] = st.r2[i];
    m32[i] = (float)st.m[i];
}
#pragma clang fp contract(on)

// Scatter `count` updated particles (AoS staging, 17 doubles each: x3 v3 a3
// ad3 t m r2 + 2 pad) into the SoA state.
__global__ __launch_bounds__(256) void scatter_kernel(double *__restrict__ state, int n,
                                                      const int *__restrict__ idx,
                                                      const double *__restrict__ rec, int count)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const int i = idx[k];
    const size_t N = (size_t)n;
    const double *r = rec + (size_t)k * 17;
    for (int c = 0; c < 12; ++c) state[(size_t)(c / 3) * 3 * N + (size_t)(c % 3) * N + i] = r[c];
    state[12 * N + i] = r[12];
    state[13 * N + i] = r[13];
    state[14 * N + i] = r[14];
}

}  // namespace

extern "C" {

int abyss_ctx_upload_state(abyss_ctx_t c, int n, const double *m, const double *x,
                           const double *v, const double *a, const double *adot, const double *t,
                           const double *r2)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (n <= 0 || !m || !x || !v || !a || !adot || !t || !r2) {
        snprintf(c->detail, sizeof(c->detail), "upload_state: n=%d or null array", n);
        return ABYSS_ERR_INVALID_ARG;
    }
    for (int j = 0; j < n; ++j) {
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(x[3 * j + d]) || !std::isfinite(v[3 * j + d]) || !std::isfinite(a[3 * j + d]) ||
                !std::isfinite(adot[3 * j + d])) {
                snprintf(c->detail, sizeof(c->detail), "upload_state: non-finite data at j=%d", j);
                return ABYSS_ERR_INVALID_INPUT;
            }
        if (!std::isfinite(m[j]) || m[j] < 0.0 || !std::isfinite(t[j]) || std::isnan(r2[j]) || r2[j] < 0.0) {
            snprintf(c->detail, sizeof(c->detail), "upload_state: bad m/t/r2 at j=%d", j);
            return ABYSS_ERR_INVALID_INPUT;
        }
    }
    // an ordinary upload of the current positions sizes and fills the j-set
    int rc = abyss_ctx_upload(c, n, m, x, v, r2);
    if (rc != ABYSS_OK) return rc;
    const size_t bytes = (size_t)n * kStateDoubles * sizeof(double);
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    if (bytes > c->state_bytes) {
        if (c->d_state) (void)hipFree(c->d_state);
        if (c->h_state) (void)hipHostFree(c->h_state);
        c->d_state = nullptr;
        c->h_state = nullptr;
        c->state_bytes = 0;
        if (hipMalloc(&c->d_state, bytes) != hipSuccess ||
            hipHostMalloc(&c->h_state, bytes, hipHostMallocDefault) != hipSuccess)
            return ABYSS_ERR_OUT_OF_MEMORY;
        c->state_bytes = bytes;
    }
    double *h = c->h_state;
    const size_t N = (size_t)n;
    for (int j = 0; j < n; ++j) {
        for (int d = 0; d < 3; ++d) {
            h[d * N + j] = x[3 * j + d];
            h[(3 + d) * N + j] = v[3 * j + d];
            h[(6 + d) * N + j] = a[3 * j + d];
            h[(9 + d) * N + j] = adot[3 * j + d];
        }
        h[12 * N + j] = t[j];
        h[13 * N + j] = m[j];
        h[14 * N + j] = r2[j];
    }
    if (hipMemcpyAsync(c->d_state, h, bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    c->state_n = n;
    return ABYSS_OK;
}

int abyss_ctx_update_state(abyss_ctx_t c, int count, const int *idx, const double *m,
                           const double *x, const double *v, const double *a,
                           const double *adot, const double *t, const double *r2)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->state_n <= 0) {
        snprintf(c->detail, sizeof(c->detail), "update_state before upload_state");
        return ABYSS_ERR_NO_JSET;
    }
    if (count <= 0 || !idx || !m || !x || !v || !a || !adot || !t || !r2) {
        snprintf(c->detail, sizeof(c->detail), "update_state: count=%d or null array", count);
        return ABYSS_ERR_INVALID_ARG;
    }
    for (int k = 0; k < count; ++k) {
        if (idx[k] < 0 || idx[k] >= c->state_n) {
            snprintf(c->detail, sizeof(c->detail), "update_state: index %d outside [0,%d)",
                     idx[k], c->state_n);
            return ABYSS_ERR_TARGET_RANGE;
        }
        for (int d = 0; d < 3; ++d)
            if (!std::isfinite(x[3 * k + d]) || !std::isfinite(v[3 * k + d]) || !std::isfinite(a[3 * k + d]) ||
                !std::isfinite(adot[3 * k + d])) {
                snprintf(c->detail, sizeof(c->detail), "update_state: non-finite data, k=%d", k);
                return ABYSS_ERR_INVALID_INPUT;
            }
        if (!std::isfinite(m[k]) || m[k] < 0.0 || !std::isfinite(t[k]) || std::isnan(r2[k]) || r2[k] < 0.0) {
            snprintf(c->detail, sizeof(c->detail), "update_state: bad m/t/r2, k=%d", k);
            return ABYSS_ERR_INVALID_INPUT;
        }
    }
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    // staging from a previous update may still be in flight
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    const size_t rec_bytes = (size_t)count * 17 * sizeof(double);
    const size_t idx_bytes = ((size_t)count * sizeof(int) + 255) & ~size_t(255);
    const size_t need = rec_bytes + idx_bytes;
    if (need > c->upd_bytes) {
        if (c->d_upd) (void)hipFree(c->d_upd);
        if (c->h_upd) (void)hipHostFree(c->h_upd);
        c->d_upd = nullptr;
        c->h_upd = nullptr;
        c->upd_bytes = 0;
        const size_t cap = need + need / 2 + 4096;
        if (hipMalloc(&c->d_upd, cap) != hipSuccess ||
            hipHostMalloc(&c->h_upd, cap, hipHostMallocDefault) != hipSuccess)
            return ABYSS_ERR_OUT_OF_MEMORY;
        c->upd_bytes = cap;
    }
    int *hidx = reinterpret_cast<int *>(c->h_upd);
    double *hrec = reinterpret_cast<double *>(c->h_upd + idx_bytes);
    memcpy(hidx, idx, (size_t)count * sizeof(int));
    for (int k = 0; k < count; ++k) {
        double *r = hrec + (size_t)k * 17;
        for (int d = 0; d < 3; ++d) {
            r[d] = x[3 * k + d];
            r[3 + d] = v[3 * k + d];
            r[6 + d] = a[3 * k + d];
            r[9 + d] = adot[3 * k + d];
        }
        r[12] = t[k];
        r[13] = m[k];
        r[14] = r2[k];
        r[15] = r[16] = 0.0;
    }
    if (hipMemcpyAsync(c->d_upd, c->h_upd, need, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    hipLaunchKernelGGL(scatter_kernel, dim3((count + 255) / 256), dim3(256), 0, c->stream,
                       c->d_state, c->state_n, reinterpret_cast<const int *>(c->d_upd),
                       reinterpret_cast<const double *>(c->d_upd + idx_bytes), count);
    return hipGetLastError() == hipSuccess ? ABYSS_OK : ABYSS_ERR_HIP;
}

int abyss_ctx_predict(abyss_ctx_t c, double t_new)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->state_n <= 0 || c->n != c->state_n) {
        snprintf(c->detail, sizeof(c->detail), "predict before upload_state");
        return ABYSS_ERR_NO_JSET;
    }
    if (!std::isfinite(t_new)) return ABYSS_ERR_INVALID_ARG;
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    const int n = c->state_n;
    double *jx = reinterpret_cast<double *>(c->d_j);
    float *m32 = reinterpret_cast<float *>(jx + 8 * (size_t)n);
    hipLaunchKernelGGL(predict_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       state_view(c->d_state, n), n, t_new, jx, m32);
    c->jset_version += 1;
    return hipGetLastError() == hipSuccess ? ABYSS_OK : ABYSS_ERR_HIP;
}

}  // extern "C"
