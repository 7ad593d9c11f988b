// Start-up derivatives and neighbour-radius adaptation on the device
// (include/abyss_regf.h: abyss_ctx_derivs23, abyss_ctx_update_radius).
//
// ABYSS computes the second and third time derivatives of every particle's
// acceleration once at start-up, by an O(N^2) host loop per particle split
// into regular and irregular parts (CalculateAcceleration23,
// src_mpi/Particle/Initialize.cpp:102-185), and rescales each particle's
// neighbour radius after every regular step from its neighbour count
// (Particle::updateRadius, src_mpi/Particle/UpdateParticle.cpp:108-139).
// Both run here on the resident particle state (abyss_ctx_upload_state).
//
// derivs23: one target per lane; every lane of a wave walks the same j, so
// the j record (x, v, a, adot, m; 128-byte AoS copy of the state) is read at
// a wave-uniform address and reaches the lanes through the scalar cache.
// The j range is split over grid.y so that small target blocks still fill
// the chip; partial sums are added in split order by a second kernel
// (bitwise reproducible).
#include "regf_context.h"

#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cmath>

namespace {

struct __align__(128) DRec {
    double x, y, z, vx, vy, vz, ax, ay, az, jx, jy, jz, m, pad0, pad1, pad2;
};

// State SoA (stride n): x[3][n], v[3][n], a[3][n], ad[3][n], t[n], m[n], r2[n].
__global__ __launch_bounds__(256) void pack_drec_kernel(const double *__restrict__ st, int n,
                                                        DRec *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const size_t N = (size_t)n;
    out[i] = DRec{st[i],          st[N + i],      st[2 * N + i],  st[3 * N + i],
                  st[4 * N + i],  st[5 * N + i],  st[6 * N + i],  st[7 * N + i],
                  st[8 * N + i],  st[9 * N + i],  st[10 * N + i], st[11 * N + i],
                  st[13 * N + i], 0.0,            0.0,            0.0};
}

// One split of the j range for 256 targets.  REF: the reference's signs of
// the projections x.(a_i - a_j)/r2, v.(a_i - a_j)/r2, x.(adot_i - adot_j)/r2
// in B and C; !REF: x.(a_j - a_i) etc., the exact derivatives of
// m x/|x|^3 along the relative orbit (x'' = a_j - a_i).
template <bool REF>
__global__ __launch_bounds__(256) void derivs23_kernel(const DRec *__restrict__ jr, int n,
                                                       const int *__restrict__ tgt, int nt,
                                                       const double *__restrict__ rnb2,
                                                       int jspan, double *__restrict__ part)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int s = blockIdx.y;
    const int j0 = s * jspan;
    const int j1 = min(n, j0 + jspan);
    const bool active = q < nt;
    const int i = active ? tgt[q] : 0;
    const DRec ri = jr[i];
    const double rn = rnb2[i];
    const double ps = REF ? 1.0 : -1.0;
    double reg2[3] = {0, 0, 0}, reg3[3] = {0, 0, 0}, irr2[3] = {0, 0, 0}, irr3[3] = {0, 0, 0};
    for (int j = j0; j < j1; ++j) {
        const DRec rj = jr[j];   // wave-uniform address
        const double x[3] = {rj.x - ri.x, rj.y - ri.y, rj.z - ri.z};
        const double v[3] = {rj.vx - ri.vx, rj.vy - ri.vy, rj.vz - ri.vz};
        const double da[3] = {ri.ax - rj.ax, ri.ay - rj.ay, ri.az - rj.az};      // a1 - a2
        const double dj[3] = {ri.jx - rj.jx, ri.jy - rj.jy, ri.jz - rj.jz};      // a1dot - a2dot
        const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
        if (r2 == 0.0) continue;   // self (and coincident j)
        const double vr = v[0] * x[0] + v[1] * x[1] + v[2] * x[2];
        const double v2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        const double r3 = r2 * sqrt(r2);
        const double m_r3 = rj.m / r3;
        const double inv_r2 = 1.0 / r2;
        double a21[3], a21dot[3];
        double rdf = 0, vdf = 0, rdfdot = 0;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            a21[d] = m_r3 * x[d];
            a21dot[d] = m_r3 * (v[d] - 3 * x[d] * vr * inv_r2);
            rdf += x[d] * da[d];
            vdf += v[d] * da[d];
            rdfdot += x[d] * dj[d];
        }
        rdf *= ps * inv_r2;
        vdf *= ps * inv_r2;
        rdfdot *= ps * inv_r2;
        const double A = vr * inv_r2;
        const double B = v2 * inv_r2 + rdf + A * A;
        const double C = 3 * vdf + rdfdot + A * (3 * B - 4 * A * A);
        const bool reg = r2 > rn;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const double adot2 = -rj.m * da[d] / r3 - 6 * A * a21dot[d] - 3 * B * a21[d];
            const double adot3 =
                -rj.m * dj[d] / r3 - 9 * A * adot2 - 9 * B * a21dot[d] - 3 * C * a21[d];
            if (reg) {
                reg2[d] += adot2;
                reg3[d] += adot3;
            } else {
                irr2[d] += adot2;
                irr3[d] += adot3;
            }
        }
    }
    if (!active) return;
    double *o = part + ((size_t)s * nt + q) * 12;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        o[d] = reg2[d];
        o[3 + d] = reg3[d];
        o[6 + d] = irr2[d];
        o[9 + d] = irr3[d];
    }
}

// out[q][c] = sum over splits s (in order) of part[s][q][c].
__global__ __launch_bounds__(256) void derivs_sum_kernel(const double *__restrict__ part, int S,
                                                         int nt, double *__restrict__ out)
{
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t total = (size_t)nt * 12;
    if (e >= total) return;
    double acc = 0.0;
    for (int s = 0; s < S; ++s) acc += part[(size_t)s * total + e];
    out[e] = acc;
}

// Radius rule of the reference on r2 (the squared radius) for the targets
// of the last launch, from their neighbour counts (|count| if overflowed);
// contraction off, so it equals the host rule bit for bit.
#pragma clang fp contract(off)
__global__ __launch_bounds__(256) void update_radius_kernel(const int *__restrict__ tgt, int nt,
                                                            const int *__restrict__ nnb, int fix,
                                                            int nmax, double *__restrict__ st_r2,
                                                            double *__restrict__ j_r2,
                                                            double *__restrict__ r2_out)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nt) return;
    const int i = tgt[q];
    const int c = nnb[q] < 0 ? -nnb[q] : nnb[q];
    double r2 = j_r2[i];
    const double cc = (double)(nmax - fix);
    const double x = (double)(c - fix);
    if (c > fix) r2 *= (1. - 0.9 / (cc * cc) * fabs(x) * x);
    else if (c < fix) r2 *= (1. - 0.5 / (cc * cc * cc) * (x * x * x));
    j_r2[i] = r2;
    if (st_r2) st_r2[i] = r2;
    r2_out[q] = r2;
}
#pragma clang fp contract(on)

int ensure_buf(char **d, size_t *bytes, size_t need, hipStream_t st)
{
    if (need <= *bytes) return ABYSS_OK;
    if (hipStreamSynchronize(st) != hipSuccess) return ABYSS_ERR_HIP;
    if (*d) (void)hipFree(*d);
    *d = nullptr;
    *bytes = 0;
    if (hipMalloc(d, need) != hipSuccess) return ABYSS_ERR_OUT_OF_MEMORY;
    *bytes = need;
    return ABYSS_OK;
}

}  // namespace

extern "C" {

int abyss_ctx_derivs23(abyss_ctx_t c, int nt, const int *targets, int reference_signs,
                       double *out)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->state_n <= 0 || c->n != c->state_n) {
        snprintf(c->detail, sizeof(c->detail), "derivs23 needs abyss_ctx_upload_state first");
        return ABYSS_ERR_NO_JSET;
    }
    if (nt <= 0 || !targets || !out) {
        snprintf(c->detail, sizeof(c->detail), "derivs23: nt=%d or null array", nt);
        return ABYSS_ERR_INVALID_ARG;
    }
    const int n = c->state_n;
    for (int k = 0; k < nt; ++k)
        if (targets[k] < 0 || targets[k] >= n) {
            snprintf(c->detail, sizeof(c->detail), "derivs23: target %d outside [0,%d)",
                     targets[k], n);
            return ABYSS_ERR_TARGET_RANGE;
        }
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    // j splits: ~8 workgroups per CU, each split at least 1024 j
    const int gx = (nt + 255) / 256;
    const int cu = c->cu_count > 0 ? c->cu_count : 256;
    int S = std::max(1, (cu * 8 + gx - 1) / gx);
    S = std::min(S, std::max(1, n / 1024));
    S = std::min(S, 1024);
    const int jspan = (n + S - 1) / S;
    S = (n + jspan - 1) / jspan;
    const size_t rec_bytes = (size_t)n * sizeof(DRec);
    const size_t part_bytes = (size_t)S * nt * 12 * sizeof(double);
    const size_t tgt_bytes = (((size_t)nt * sizeof(int)) + 255) & ~size_t(255);
    const size_t out_bytes = (size_t)nt * 12 * sizeof(double);
    int rc = ensure_buf(&c->d_drv, &c->drv_bytes, rec_bytes + part_bytes + tgt_bytes + out_bytes,
                        c->stream);
    if (rc != ABYSS_OK) return rc;
    if (hipStreamSynchronize(c->stream) != hipSuccess) return ABYSS_ERR_HIP;
    if (out_bytes + tgt_bytes > c->h_drv_bytes) {
        if (c->h_drv) (void)hipHostFree(c->h_drv);
        c->h_drv = nullptr;
        c->h_drv_bytes = 0;
        if (hipHostMalloc(&c->h_drv, out_bytes + tgt_bytes, hipHostMallocDefault) != hipSuccess)
            return ABYSS_ERR_OUT_OF_MEMORY;
        c->h_drv_bytes = out_bytes + tgt_bytes;
    }
    DRec *d_rec = reinterpret_cast<DRec *>(c->d_drv);
    double *d_part = reinterpret_cast<double *>(c->d_drv + rec_bytes);
    int *d_tgt = reinterpret_cast<int *>(c->d_drv + rec_bytes + part_bytes);
    double *d_out = reinterpret_cast<double *>(c->d_drv + rec_bytes + part_bytes + tgt_bytes);
    memcpy(c->h_drv, targets, (size_t)nt * sizeof(int));
    if (hipMemcpyAsync(d_tgt, c->h_drv, (size_t)nt * sizeof(int), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    hipLaunchKernelGGL(pack_drec_kernel, dim3((n + 255) / 256), dim3(256), 0, c->stream,
                       c->d_state, n, d_rec);
    const double *st_r2 = c->d_state + 14 * (size_t)n;
    if (reference_signs)
        hipLaunchKernelGGL(derivs23_kernel<true>, dim3(gx, S), dim3(256), 0, c->stream, d_rec, n,
                           d_tgt, nt, st_r2, jspan, d_part);
    else
        hipLaunchKernelGGL(derivs23_kernel<false>, dim3(gx, S), dim3(256), 0, c->stream, d_rec,
                           n, d_tgt, nt, st_r2, jspan, d_part);
    const size_t total = (size_t)nt * 12;
    hipLaunchKernelGGL(derivs_sum_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       c->stream, d_part, S, nt, d_out);
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    double *h_out = reinterpret_cast<double *>(c->h_drv + tgt_bytes);
    if (hipMemcpyAsync(h_out, d_out, out_bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    memcpy(out, h_out, out_bytes);
    return ABYSS_OK;
}

int abyss_ctx_update_radius(abyss_ctx_t c, int fix, int nmax, double *r2_out)
{
    if (!c) return ABYSS_ERR_INVALID_ARG;
    if (c->n <= 0 || c->nt <= 0 || !c->counts_valid) {
        snprintf(c->detail, sizeof(c->detail), "update_radius: no launch to take counts from");
        return ABYSS_ERR_NO_JSET;
    }
    if (fix < 0 || nmax <= fix || !r2_out) {
        snprintf(c->detail, sizeof(c->detail), "update_radius: fix=%d nmax=%d or null output",
                 fix, nmax);
        return ABYSS_ERR_INVALID_ARG;
    }
    if (hipSetDevice(c->device) != hipSuccess) return ABYSS_ERR_HIP;
    const int nt = c->nt;
    int rc = ensure_buf(&c->d_rad, &c->rad_bytes, (size_t)nt * sizeof(double), c->stream);
    if (rc != ABYSS_OK) return rc;
    double *j_r2 = reinterpret_cast<double *>(c->d_j) + 7 * (size_t)c->n;   // j-set r_nb^2
    double *st_r2 = (c->state_n == c->n && c->d_state) ? c->d_state + 14 * (size_t)c->n : nullptr;
    double *d_r2 = reinterpret_cast<double *>(c->d_rad);
    hipLaunchKernelGGL(update_radius_kernel, dim3((nt + 255) / 256), dim3(256), 0, c->stream,
                       c->tgt, nt, c->d_nnb, fix, nmax, st_r2, j_r2, d_r2);
    if (hipGetLastError() != hipSuccess) return ABYSS_ERR_HIP;
    if (hipMemcpyAsync(r2_out, d_r2, (size_t)nt * sizeof(double), hipMemcpyDeviceToHost,
                       c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        return ABYSS_ERR_HIP;
    return ABYSS_OK;
}

}  // extern "C"
