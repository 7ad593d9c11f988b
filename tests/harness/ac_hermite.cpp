// Test harness (not product): a block-timestep 4th-order Hermite integrator
// whose force is split the Ahmad-Cohen way -- regular part from the regular
// force provider, irregular part summed on the host over the provider's
// neighbour list -- driven through the legacy C ABI exactly as ABYSS's
// caller does:
//   predict every particle to the block time (second-order predictor,
//   src_mpi/Particle/UpdateParticle.cpp:37-59), SendToDevice(all j),
//   CalculateAccelerationOnDevice(active block)
//   (src/cuda/CalculateRegularAcceleration.cpp:130-149), then the Hermite
//   corrector (src_mpi/Particle/ComputeAcceleration.cpp:310-346) and an
//   Aarseth time step quantised to powers of two
//   (src_mpi/TimeStepRoutines.cpp:10-25, :100-115).
// Regular and irregular parts share each particle's step (no separate
// regular time scale), so the run checks the split force and jerk and the
// neighbour lists in a dynamical setting through energy conservation.
//
// Providers: "gpu" = libabyss_regf.so through the six legacy entry points
// (host prediction, full upload every block); "resident" = the same library
// with the particle state resident on the device (abyss_ctx_upload_state,
// abyss_ctx_predict on the device, abyss_ctx_update_state of the active
// block after its corrector); "oracle" = the FP64 oracle
// (oracle/liboracle.so).  Output: one JSON line.
//
// usage: ac_hermite <gpu|resident|oracle> <ic.bin> <t_end> <eta>
//   ic.bin: int32 n, then x[n][3], v[n][3], m[n], r2[n] as float64
// Built with -DAC_NO_ORACLE (tests/harness/ac_hermite_gpu, which bench.py
// runs) it has no oracle provider and does not link the oracle library.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "abyss_regf.h"

#ifndef AC_NO_ORACLE
extern "C" int oracle_regf(int n, const double *x, const double *v, const double *m,
                           const double *rnb2, int nt, const int *targets, double *acc,
                           double *adot, int *nnb, int *list, int cap, double *scale_a,
                           double *scale_j, int nthreads);
#endif

namespace {

struct System {
    int n = 0;
    std::vector<double> x, v, m, r2;     // [n][3], [n][3], [n], [n]
    std::vector<double> a, ad;           // total force and jerk at t_i
    std::vector<double> a2, a3;          // 2nd, 3rd derivative estimates at t_i
    std::vector<double> t, dt;           // last update time, step
};

enum Provider { GPU, RESIDENT, ORACLE };
Provider prov = GPU;
abyss_ctx_t rctx = nullptr;   // RESIDENT: the device-resident context
int cap = ABYSS_NUM_NEIGHBOR_MAX;
long long reg_calls = 0, reg_targets = 0;

// Regular force + jerk and neighbour lists of targets[k] from the provider,
// with j predicted to time tnow; then irregular part over the neighbours.
void total_force(const System &s, double tn, const std::vector<double> &xp,
                 const std::vector<double> &vp, const std::vector<int> &tg,
                 std::vector<double> &acc, std::vector<double> &jerk)
{
    const int nt = (int)tg.size();
    std::vector<double> ar(3 * nt), jr(3 * nt);
    std::vector<int> nnb(nt), list((size_t)nt * cap);
    if (prov == RESIDENT) {
        int rc = abyss_ctx_predict(rctx, tn);
        if (rc == ABYSS_OK)
            rc = abyss_ctx_query(rctx, nt, tg.data(), ar.data(), jr.data(), nnb.data(),
                                 list.data());
        if (rc != ABYSS_OK) {
            fprintf(stderr, "resident regular force failed: %s\n",
                    abyss_ctx_error_detail(rctx));
            exit(2);
        }
    } else if (prov == GPU) {
        int n = s.n, ntt = nt;
        SendToDevice(&n, const_cast<double *>(s.m.data()),
                     reinterpret_cast<double(*)[3]>(const_cast<double *>(xp.data())),
                     reinterpret_cast<double(*)[3]>(const_cast<double *>(vp.data())),
                     const_cast<double *>(s.r2.data()), nullptr);
        CalculateAccelerationOnDevice(&ntt, const_cast<int *>(tg.data()),
                                      reinterpret_cast<double(*)[3]>(ar.data()),
                                      reinterpret_cast<double(*)[3]>(jr.data()), nnb.data(),
                                      list.data());
        if (abyss_last_error() != ABYSS_OK) {
            fprintf(stderr, "regular force failed: %s\n", abyss_last_error_detail());
            exit(2);
        }
    } else {
#ifdef AC_NO_ORACLE
        fprintf(stderr, "built without the oracle provider\n");
        exit(2);
#else
        oracle_regf(s.n, xp.data(), vp.data(), s.m.data(), s.r2.data(), nt, tg.data(), ar.data(),
                    jr.data(), nnb.data(), list.data(), cap, nullptr, nullptr, 0);
#endif
    }
    reg_calls += 1;
    reg_targets += nt;
    acc.assign(3 * nt, 0.0);
    jerk.assign(3 * nt, 0.0);
    for (int k = 0; k < nt; ++k) {
        const int i = tg[k];
        if (nnb[k] < 0) {
            fprintf(stderr, "neighbour overflow at target %d\n", i);
            exit(3);
        }
        double ax = ar[3 * k], ay = ar[3 * k + 1], az = ar[3 * k + 2];
        double jx = jr[3 * k], jy = jr[3 * k + 1], jz = jr[3 * k + 2];
        for (int e = 0; e < nnb[k]; ++e) {   // irregular part, FP64
            const int j = list[(size_t)k * cap + e];
            const double dx = xp[3 * j] - xp[3 * i], dy = xp[3 * j + 1] - xp[3 * i + 1],
                         dz = xp[3 * j + 2] - xp[3 * i + 2];
            const double dvx = vp[3 * j] - vp[3 * i], dvy = vp[3 * j + 1] - vp[3 * i + 1],
                         dvz = vp[3 * j + 2] - vp[3 * i + 2];
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (r2 == 0.0) continue;
            const double rv = dx * dvx + dy * dvy + dz * dvz;
            const double mr3 = s.m[j] / r2 / sqrt(r2);
            ax += mr3 * dx;
            ay += mr3 * dy;
            az += mr3 * dz;
            jx += mr3 * (dvx - 3 * dx * rv / r2);
            jy += mr3 * (dvy - 3 * dy * rv / r2);
            jz += mr3 * (dvz - 3 * dz * rv / r2);
        }
        acc[3 * k] = ax; acc[3 * k + 1] = ay; acc[3 * k + 2] = az;
        jerk[3 * k] = jx; jerk[3 * k + 1] = jy; jerk[3 * k + 2] = jz;
    }
}

// Second-order predictor of every particle to time tn.
void predict_all(const System &s, double tn, std::vector<double> &xp, std::vector<double> &vp)
{
    xp.resize(3 * (size_t)s.n);
    vp.resize(3 * (size_t)s.n);
    for (int i = 0; i < s.n; ++i) {
        const double h = tn - s.t[i];
        for (int d = 0; d < 3; ++d) {
            const int q = 3 * i + d;
            if (h == 0) {
                xp[q] = s.x[q];
                vp[q] = s.v[q];
            } else {
                xp[q] = ((s.ad[q] * h / 3 + s.a[q]) * h / 2 + s.v[q]) * h + s.x[q];
                vp[q] = (s.ad[q] * h / 2 + s.a[q]) * h + s.v[q];
            }
        }
    }
}

double norm3(const double *p) { return sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// Aarseth criterion, quantised to a power of two that divides the time.
double block_step(double a, double j, double s2, double s3, double eta, double tnow, double dtold,
                  double dtmax)
{
    double dt = sqrt(eta * (a * s2 + j * j) / (j * s3 + s2 * s2));
    if (!(dt > 0) || !isfinite(dt)) dt = dtmax;
    double q = pow(2.0, floor(log2(dt)));
    q = std::min(q, std::min(2.0 * dtold, dtmax));
    while (fmod(tnow, q) != 0.0) q *= 0.5;   // keep the block grid
    return q;
}

double energy(const System &s)
{
    double K = 0, U = 0;
    for (int i = 0; i < s.n; ++i) {
        K += 0.5 * s.m[i] * (s.v[3 * i] * s.v[3 * i] + s.v[3 * i + 1] * s.v[3 * i + 1] +
                             s.v[3 * i + 2] * s.v[3 * i + 2]);
        for (int j = i + 1; j < s.n; ++j) {
            const double dx = s.x[3 * j] - s.x[3 * i], dy = s.x[3 * j + 1] - s.x[3 * i + 1],
                         dz = s.x[3 * j + 2] - s.x[3 * i + 2];
            U -= s.m[i] * s.m[j] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    return K + U;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 5) {
        fprintf(stderr, "usage: %s <gpu|oracle> <ic.bin> <t_end> <eta>\n", argv[0]);
        return 1;
    }
    prov = strcmp(argv[1], "gpu") == 0 ? GPU : strcmp(argv[1], "resident") == 0 ? RESIDENT : ORACLE;
    const double t_end = atof(argv[3]);
    const double eta = atof(argv[4]);
    System s;
    FILE *f = fopen(argv[2], "rb");
    if (!f || fread(&s.n, sizeof(int), 1, f) != 1 || s.n <= 1) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 1;
    }
    const int n = s.n;
    s.x.resize(3 * n); s.v.resize(3 * n); s.m.resize(n); s.r2.resize(n);
    if (fread(s.x.data(), 8, 3 * n, f) != (size_t)3 * n ||
        fread(s.v.data(), 8, 3 * n, f) != (size_t)3 * n || fread(s.m.data(), 8, n, f) != (size_t)n ||
        fread(s.r2.data(), 8, n, f) != (size_t)n) {
        fprintf(stderr, "short read\n");
        return 1;
    }
    fclose(f);
    s.a.assign(3 * n, 0); s.ad.assign(3 * n, 0); s.a2.assign(3 * n, 0); s.a3.assign(3 * n, 0);
    s.t.assign(n, 0.0); s.dt.assign(n, 0.0);
    if (prov == GPU) {
        int rank = 0;
        OpenDevice(&rank);
        if (abyss_last_error() != ABYSS_OK) {
            fprintf(stderr, "OpenDevice: %s\n", abyss_last_error_detail());
            return 2;
        }
    } else if (prov == RESIDENT) {
        std::vector<double> zero(3 * n, 0.0);
        if (abyss_ctx_create(0, cap, &rctx) != ABYSS_OK ||
            abyss_ctx_upload_state(rctx, n, s.m.data(), s.x.data(), s.v.data(), zero.data(),
                                   zero.data(), s.t.data(), s.r2.data()) != ABYSS_OK) {
            fprintf(stderr, "resident context setup failed\n");
            return 2;
        }
    }
    const auto wall0 = std::chrono::steady_clock::now();
    const double E0 = energy(s);
    // start-up: forces of everyone, initial steps from |a|/|jerk|
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;
    std::vector<double> xp, vp, acc, jerk;
    predict_all(s, 0.0, xp, vp);
    total_force(s, 0.0, xp, vp, all, acc, jerk);
    const double dtmax = 1.0 / 16.0;
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            s.a[3 * i + d] = acc[3 * i + d];
            s.ad[3 * i + d] = jerk[3 * i + d];
        }
        const double dt0 = 0.01 * norm3(&s.a[3 * i]) / std::max(norm3(&s.ad[3 * i]), 1e-300);
        double q = pow(2.0, floor(log2(std::max(dt0, 1e-12))));
        s.dt[i] = std::min(q, dtmax);
    }
    auto push_state = [&](const std::vector<int> &idx) {
        if (prov != RESIDENT) return;
        const int k = (int)idx.size();
        std::vector<double> m(k), x(3 * k), v(3 * k), a(3 * k), ad(3 * k), t(k), r2(k);
        for (int q = 0; q < k; ++q) {
            const int i = idx[q];
            m[q] = s.m[i];
            t[q] = s.t[i];
            r2[q] = s.r2[i];
            for (int d = 0; d < 3; ++d) {
                x[3 * q + d] = s.x[3 * i + d];
                v[3 * q + d] = s.v[3 * i + d];
                a[3 * q + d] = s.a[3 * i + d];
                ad[3 * q + d] = s.ad[3 * i + d];
            }
        }
        if (abyss_ctx_update_state(rctx, k, idx.data(), m.data(), x.data(), v.data(), a.data(),
                                   ad.data(), t.data(), r2.data()) != ABYSS_OK) {
            fprintf(stderr, "update_state failed: %s\n", abyss_ctx_error_detail(rctx));
            exit(2);
        }
    };
    push_state(all);
    long long steps = 0, blocks = 0;
    for (;;) {
        // next block time among particles that have not reached t_end
        double tn = 1e300;
        for (int i = 0; i < n; ++i)
            if (s.t[i] < t_end) tn = std::min(tn, s.t[i] + s.dt[i]);
        if (tn > t_end) break;   // everyone is at t_end
        std::vector<int> act;
        for (int i = 0; i < n; ++i)
            if (s.t[i] < t_end && s.t[i] + s.dt[i] == tn) act.push_back(i);
        predict_all(s, tn, xp, vp);
        total_force(s, tn, xp, vp, act, acc, jerk);
        for (size_t k = 0; k < act.size(); ++k) {
            const int i = act[k];
            const double h = tn - s.t[i];
            double s2v[3], s3v[3];
            for (int d = 0; d < 3; ++d) {
                const int q = 3 * i + d;
                const double a0 = s.a[q], j0 = s.ad[q], a1 = acc[3 * k + d], j1 = jerk[3 * k + d];
                const double A2 = (-6 * (a0 - a1) - h * (4 * j0 + 2 * j1)) / (h * h);
                const double A3 = (12 * (a0 - a1) + 6 * h * (j0 + j1)) / (h * h * h);
                s.x[q] = xp[q] + A2 * h * h * h * h / 24 + A3 * h * h * h * h * h / 120;
                s.v[q] = vp[q] + A2 * h * h * h / 6 + A3 * h * h * h * h / 24;
                s.a[q] = a1;
                s.ad[q] = j1;
                s.a2[q] = A2 + A3 * h;   // 2nd derivative at the new time
                s.a3[q] = A3;
                s2v[d] = s.a2[q];
                s3v[d] = A3;
            }
            s.t[i] = tn;
            // power-of-two steps tile [0, t_end] exactly when t_end is a
            // multiple of dtmax, so t + dt never passes t_end
            s.dt[i] = block_step(norm3(&s.a[3 * i]), norm3(&s.ad[3 * i]), norm3(s2v), norm3(s3v),
                                 eta, tn, s.dt[i], dtmax);
        }
        push_state(act);
        steps += (long long)act.size();
        blocks += 1;
    }
    const double E1 = energy(s);
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    if (prov == GPU) {
        int rank = 0;
        ProfileDevice(&rank);
        CloseDevice();
    } else if (prov == RESIDENT) {
        abyss_ctx_destroy(rctx);
    }
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; ++i) {
        cx += s.m[i] * s.x[3 * i];
        cy += s.m[i] * s.x[3 * i + 1];
        cz += s.m[i] * s.x[3 * i + 2];
    }
    printf("{\"provider\": \"%s\", \"n\": %d, \"t_end\": %.6f, \"eta\": %g, \"E0\": %.15e, "
           "\"E1\": %.15e, \"dE_over_E\": %.6e, \"blocks\": %lld, \"particle_steps\": %lld, "
           "\"regular_calls\": %lld, \"mean_block\": %.2f, \"com\": [%.6e, %.6e, %.6e], "
           "\"x0\": [%.15e, %.15e, %.15e], \"wall_s\": %.3f}\n",
           argv[1], n, t_end, eta, E0, E1, (E1 - E0) / fabs(E0), blocks, steps,
           reg_calls, (double)reg_targets / std::max(1LL, reg_calls), cx, cy, cz, s.x[0], s.x[1],
           s.x[2], wall);
    return 0;
}
