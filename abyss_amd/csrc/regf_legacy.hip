// The six legacy entry points (src_mpi/cuda/cuda_functions.h:3-10, defined
// at src_mpi/cuda/cuda_my_acceleration.cu:1087-1106) over one process-wide
// default context, plus the legacy status / precision / error-mode controls.
//
// This file is compiled twice: as is for libabyss_regf.so (CUDA_REAL =
// double, src/defs.h:101) and with -DABYSS_LEGACY_FLOAT for
// libabyss_regf_f32.so (CUDA_REAL = float, src_mpi/def.h:69-74), whose
// SendToDevice widens the caller's arrays to FP64 and whose query rounds the
// FP64 results to float.  Everything else is the same code path.
//
// Errors: the reference stops the process on neighbour overflow
// (fprintf + exit(1), src/cuda/cuda_my_acceleration.cu:221-226,
// src_mpi/cuda/cuda_my_acceleration.cu:328) and on a closed device or a bad
// target count (asserts, :103-104); its callers never look at a status.  By
// default every failing call here prints its error and calls exit(1) too;
// abyss_set_legacy_errors(ABYSS_LEGACY_RETURN) records the status for
// abyss_last_error() and returns instead.
#include "regf_context.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

namespace {
abyss_ctx_t g_ctx = nullptr;
bool g_open = false;
int g_cap = ABYSS_NUM_NEIGHBOR_MAX;
int g_prec = ABYSS_PREC_CHECKED;
int g_errors = ABYSS_LEGACY_EXIT;
int g_last = ABYSS_OK;
char g_detail[512] = {0};
char g_init_detail[256] = {0};

#ifdef ABYSS_LEGACY_FLOAT
// FP64 staging of the float ABI: j-set inputs (m, x, v, r2) and results
std::vector<double> g_in, g_res;
#endif

void legacy_status(int rc)
{
    g_last = rc;
    if (rc == ABYSS_OK) {
        g_detail[0] = 0;
        return;
    }
    const char *d = g_ctx ? abyss_ctx_error_detail(g_ctx) : g_init_detail;
    snprintf(g_detail, sizeof(g_detail), "%s%s%s", abyss_error_string(rc), d[0] ? ": " : "", d);
    fprintf(stderr, "abyss_regf: %s\n", g_detail);
    if (g_errors == ABYSS_LEGACY_EXIT) {
        fflush(stderr);
        exit(1);
    }
}

int legacy_init(int irank)
{
    if (g_ctx) return ABYSS_OK;
    const int ndev = abyss_device_count();
    if (ndev <= 0) {
        snprintf(g_init_detail, sizeof(g_init_detail), "no HIP device visible");
        return ABYSS_ERR_NO_DEVICE;
    }
    const int dev = ((irank % ndev) + ndev) % ndev;
    int rc = abyss_ctx_create(dev, g_cap, &g_ctx);
    if (rc != ABYSS_OK) {
        snprintf(g_init_detail, sizeof(g_init_detail), "context creation on device %d failed",
                 dev);
        g_ctx = nullptr;
        return rc;
    }
    g_init_detail[0] = 0;
    abyss_ctx_set_precision(g_ctx, g_prec);
    return ABYSS_OK;
}

int set_ctx_detail(const char *msg)
{
    if (g_ctx) snprintf(g_ctx->detail, sizeof(g_ctx->detail), "%s", msg);
    return ABYSS_ERR_INVALID_ARG;
}
}  // namespace

extern "C" {

int abyss_last_error(void) { return g_last; }
const char *abyss_last_error_detail(void) { return g_detail; }
int abyss_num_neighbor_max(void) { return g_cap; }

int abyss_set_num_neighbor_max(int cap)
{
    if (cap <= 0) return ABYSS_ERR_INVALID_ARG;
    if (g_ctx) return ABYSS_ERR_INVALID_ARG;  // only while closed
    g_cap = cap;
    return ABYSS_OK;
}

int abyss_set_precision(int mode)
{
    if (mode < ABYSS_PREC_SPLIT || mode > ABYSS_PREC_CHECKED) return ABYSS_ERR_INVALID_ARG;
    g_prec = mode;
    if (g_ctx) abyss_ctx_set_precision(g_ctx, mode);
    return ABYSS_OK;
}

int abyss_set_legacy_errors(int mode)
{
    if (mode != ABYSS_LEGACY_EXIT && mode != ABYSS_LEGACY_RETURN) return ABYSS_ERR_INVALID_ARG;
    g_errors = mode;
    return ABYSS_OK;
}

void InitializeDevice(int *irank)
{
    const int r = irank ? *irank : 0;
    int rc = legacy_init(r);
    if (rc == ABYSS_OK) g_open = true;  // src/cuda/cuda_my_acceleration.cu:454
    legacy_status(rc);
}

void OpenDevice(const int *irank)
{
    if (g_open) {
        fprintf(stderr, "abyss_regf: it is already open\n");
        legacy_status(ABYSS_OK);
        return;
    }
    int rc = legacy_init(irank ? *irank : 0);
    if (rc == ABYSS_OK) g_open = true;
    legacy_status(rc);
}

void CloseDevice(void)
{
    if (!g_open) {
        fprintf(stderr, "abyss_regf: it is already closed\n");
        legacy_status(ABYSS_OK);
        return;
    }
    if (g_ctx) abyss_ctx_destroy(g_ctx);
    g_ctx = nullptr;
    g_open = false;
#ifdef ABYSS_LEGACY_FLOAT
    std::vector<double>().swap(g_in);
    std::vector<double>().swap(g_res);
#endif
    legacy_status(ABYSS_OK);
}

void ProfileDevice(int *irank)
{
    if (!g_ctx) {
        legacy_status(ABYSS_ERR_NOT_OPEN);
        return;
    }
    abyss_ctx_t c = g_ctx;
    if (c->n_grav) {
        const double gflops = c->t_grav_ms > 0 ? 60e-9 * c->interactions / (c->t_grav_ms * 1e-3)
                                               : 0.0;
        fprintf(stderr,
                "[R.%d-D.%d GPU Reg.F ] Nsend %lld  Ngrav %lld  <Ni> %lld  grav(s) %f  "
                "interactions %.6e  Perf.(Gflops) %f\n",
                irank ? *irank : 0, c->device, c->n_send, c->n_grav,
                c->n_send ? c->sum_ni / std::max(1LL, c->n_send) : 0LL, c->t_grav_ms * 1e-3,
                c->interactions, gflops);
    }
    c->n_send = c->n_grav = c->sum_ni = 0;
    c->interactions = 0.0;
    c->t_grav_ms = 0.0;
    legacy_status(ABYSS_OK);
}

void SendToDevice(int *NNB, abyss_legacy_real m[], abyss_legacy_real x[][3],
                  abyss_legacy_real v[][3], abyss_legacy_real r2[], abyss_legacy_real mdot[])
{
    (void)mdot;  // ignored, as in the reference
    if (!g_open || !g_ctx) {
        legacy_status(ABYSS_ERR_NOT_OPEN);
        return;
    }
    if (!NNB) {
        legacy_status(set_ctx_detail("NNB is null"));
        return;
    }
#ifdef ABYSS_LEGACY_FLOAT
    const int n = *NNB;
    if (n <= 0 || !m || !x || !v || !r2) {
        // the context reports the argument error (and drops its j-set)
        legacy_status(abyss_ctx_upload(g_ctx, n, nullptr, nullptr, nullptr, nullptr));
        return;
    }
    // [m | x | v | r2] widened to FP64 (exact)
    g_in.resize((size_t)n * 8);
    double *dm = g_in.data(), *dx = dm + n, *dv = dx + 3 * (size_t)n, *dr = dv + 3 * (size_t)n;
    for (int j = 0; j < n; ++j) {
        dm[j] = m[j];
        dr[j] = r2[j];
        for (int d = 0; d < 3; ++d) {
            dx[3 * (size_t)j + d] = x[j][d];
            dv[3 * (size_t)j + d] = v[j][d];
        }
    }
    legacy_status(abyss_ctx_upload(g_ctx, n, dm, dx, dv, dr));
#else
    legacy_status(abyss_ctx_upload(g_ctx, *NNB, m, x ? &x[0][0] : nullptr,
                                   v ? &v[0][0] : nullptr, r2));
#endif
}

void CalculateAccelerationOnDevice(int *NumTarget, int *h_target_list,
                                   abyss_legacy_real acc[][3], abyss_legacy_real adot[][3],
                                   int NumNeighbor[], int *NeighborList)
{
    if (!g_open || !g_ctx) {
        legacy_status(ABYSS_ERR_NOT_OPEN);
        return;
    }
    if (!NumTarget) {
        legacy_status(set_ctx_detail("NumTarget is null"));
        return;
    }
    // grav(s) of ProfileDevice: the wall-clock time of the call, as the
    // reference's host timers measured it (src/cuda/old/
    // cuda_acceleration.cu_nbody:756)
    const auto t0 = std::chrono::steady_clock::now();
#ifdef ABYSS_LEGACY_FLOAT
    const int nt = *NumTarget;
    if (nt <= 0 || !acc || !adot) {
        legacy_status(abyss_ctx_query(g_ctx, nt, h_target_list, nullptr, nullptr, NumNeighbor,
                                      NeighborList));
        return;
    }
    g_res.resize((size_t)nt * 6);
    double *ra = g_res.data(), *rj = ra + 3 * (size_t)nt;
    const int rc = abyss_ctx_query(g_ctx, nt, h_target_list, ra, rj, NumNeighbor, NeighborList);
    if (rc == ABYSS_OK || rc == ABYSS_ERR_NEIGHBOR_OVERFLOW) {
        for (int k = 0; k < nt; ++k)
            for (int d = 0; d < 3; ++d) {
                acc[k][d] = (float)ra[3 * (size_t)k + d];
                adot[k][d] = (float)rj[3 * (size_t)k + d];
            }
    }
#else
    const int rc = abyss_ctx_query(g_ctx, *NumTarget, h_target_list,
                                   acc ? &acc[0][0] : nullptr, adot ? &adot[0][0] : nullptr,
                                   NumNeighbor, NeighborList);
#endif
    g_ctx->t_grav_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    legacy_status(rc);
}

}  // extern "C"
