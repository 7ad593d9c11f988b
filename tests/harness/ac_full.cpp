// Test harness (not product): an Ahmad-Cohen integrator with separate
// irregular and regular time steps, the scheme ABYSS runs
// (src_mpi/Particle/ComputeAcceleration.cpp, UpdateParticle.cpp,
// TimeStepRoutines.cpp), driven through the device-resident entry points of
// libabyss_regf.so or through the FP64 oracle:
//   start-up   query (regular force + lists), irregular force over the
//              lists, start-up derivatives a2/a3 (abyss_ctx_derivs23);
//   irregular  every block: predict all j on the device (abyss_ctx_predict),
//              irregular force at the block time over each particle's
//              resident list (abyss_ctx_irregular), Hermite corrector of the
//              irregular part on the host, the regular force extrapolated
//              linearly from the last regular step;
//   regular    when a particle's regular step is due: regular force, new
//              list, neighbour-set correction and regular corrector in one
//              call (abyss_ctx_regular_step), which also gives the irregular
//              force over the new list.
// The corrected particles go back with abyss_ctx_update_state.  Time steps
// are Aarseth's criterion on the irregular / regular derivatives, quantised
// to powers of two, the irregular step dividing the regular one.
// Providers: "gpu" (the library) and "oracle" (oracle/liboracle.so: the same
// steps on the host, in FP64).  Output: one JSON line.
//
// After every regular step the neighbour radii are adapted by the
// reference's rule (abyss_ctx_update_radius) towards <fix> neighbours, with
// nmax = the list capacity; overflowed lists are counted and reported.
//
// usage: ac_full <gpu|oracle> <ic.bin> <t_end> <eta_irr> <eta_reg> <fix>
//   ic.bin: int32 n, then x[n][3], v[n][3], m[n], r2[n] as float64
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "abyss_regf.h"

extern "C" {
int oracle_regf(int n, const double *x, const double *v, const double *m, const double *rnb2,
                int nt, const int *targets, double *acc, double *adot, int *nnb, int *list,
                int cap, double *scale_a, double *scale_j, int nthreads);
void oracle_irregular(int n, const double *x, const double *v, const double *m, int nt,
                      const int *targets, const int *nnb, const int *list, int cap, double *acc,
                      double *adot);
void oracle_regular_correct(int n, const double *x, const double *v, const double *m, int nt,
                            const int *targets, const double *dts, const double *a_old,
                            const double *new_acc, const double *new_adot, const int *old_nnb,
                            const int *old_list, const int *new_nnb, const int *new_list,
                            int cap, double *out);
void oracle_derivs23(int n, const double *X, const double *V, const double *Acc,
                     const double *Adot, const double *m, const double *rnb2, int nt,
                     const int *targets, int ref_signs, int nthreads, double *out);
void oracle_update_radius(int count, const int *nnb, int fix, int nmax, double *r2);
}

namespace {

constexpr int kCap = 256;

void die(const char *what, const char *detail)
{
    fprintf(stderr, "%s failed: %s\n", what, detail ? detail : "");
    exit(2);
}

// The steps the integrator needs from a provider.  State arrays: [n][3]
// x, v, a (total force), ad (total jerk) at each particle's time t.
struct Provider {
    virtual ~Provider() {}
    virtual void upload_state(int n, const double *m, const double *x, const double *v,
                              const double *a, const double *ad, const double *t,
                              const double *r2) = 0;
    virtual void update_state(int k, const int *idx, const double *m, const double *x,
                              const double *v, const double *a, const double *ad, const double *t,
                              const double *r2) = 0;
    virtual void predict(double tn) = 0;
    virtual void query(int nt, const int *tg, double *acc, double *adot, int *nnb) = 0;
    virtual void irregular(int nt, const int *tg, double *acc, double *adot) = 0;
    virtual void regular_step(int nt, const int *tg, const double *dt, const double *a_old,
                              double *acc, double *adot, int *nnb, double *out) = 0;
    virtual void derivs23(int nt, const int *tg, double *out) = 0;
    // radius rule for the targets of the last regular step (counts nnb)
    virtual void update_radius(int nt, const int *tg, const int *nnb, int fix, int nmax,
                               double *r2_out) = 0;
};

struct GpuProvider : Provider {
    abyss_ctx_t c = nullptr;
    std::vector<int> list;
    GpuProvider()
    {
        if (abyss_ctx_create(0, kCap, &c) != ABYSS_OK) die("abyss_ctx_create", "");
        if (abyss_ctx_keep_lists(c, 1) != ABYSS_OK) die("keep_lists", abyss_ctx_error_detail(c));
    }
    ~GpuProvider() { abyss_ctx_destroy(c); }
    void chk(int rc, const char *what)
    {
        if (rc != ABYSS_OK) die(what, abyss_ctx_error_detail(c));
    }
    void upload_state(int n, const double *m, const double *x, const double *v, const double *a,
                      const double *ad, const double *t, const double *r2) override
    {
        chk(abyss_ctx_upload_state(c, n, m, x, v, a, ad, t, r2), "upload_state");
    }
    void update_state(int k, const int *idx, const double *m, const double *x, const double *v,
                      const double *a, const double *ad, const double *t,
                      const double *r2) override
    {
        chk(abyss_ctx_update_state(c, k, idx, m, x, v, a, ad, t, r2), "update_state");
    }
    void predict(double tn) override { chk(abyss_ctx_predict(c, tn), "predict"); }
    void query(int nt, const int *tg, double *acc, double *adot, int *nnb) override
    {
        list.resize((size_t)nt * kCap);
        chk(abyss_ctx_query(c, nt, tg, acc, adot, nnb, list.data()), "query");
    }
    void irregular(int nt, const int *tg, double *acc, double *adot) override
    {
        chk(abyss_ctx_irregular(c, nt, tg, acc, adot), "irregular");
    }
    void regular_step(int nt, const int *tg, const double *dt, const double *a_old, double *acc,
                      double *adot, int *nnb, double *out) override
    {
        list.resize((size_t)nt * kCap);
        chk(abyss_ctx_regular_step(c, nt, tg, dt, a_old, acc, adot, nnb, list.data(), out),
            "regular_step");
    }
    void derivs23(int nt, const int *tg, double *out) override
    {
        chk(abyss_ctx_derivs23(c, nt, tg, 0, out), "derivs23");
    }
    void update_radius(int, const int *, const int *, int fix, int nmax, double *r2_out) override
    {
        chk(abyss_ctx_update_radius(c, fix, nmax, r2_out), "update_radius");
    }
};

// The same steps on the host with the FP64 oracle; lists kept per particle.
struct OracleProvider : Provider {
    int n = 0;
    std::vector<double> m, x, v, a, ad, t, r2;   // state
    std::vector<double> xp, vp;                  // predicted j-set
    std::vector<int> rnnb, rlist;                // resident lists [n][kCap]
    void upload_state(int n_, const double *m_, const double *x_, const double *v_,
                      const double *a_, const double *ad_, const double *t_,
                      const double *r2_) override
    {
        n = n_;
        m.assign(m_, m_ + n); x.assign(x_, x_ + 3 * n); v.assign(v_, v_ + 3 * n);
        a.assign(a_, a_ + 3 * n); ad.assign(ad_, ad_ + 3 * n); t.assign(t_, t_ + n);
        r2.assign(r2_, r2_ + n);
        rnnb.assign(n, 0);
        rlist.assign((size_t)n * kCap, 0);
        xp = x;
        vp = v;
    }
    void update_state(int k, const int *idx, const double *m_, const double *x_,
                      const double *v_, const double *a_, const double *ad_, const double *t_,
                      const double *r2_) override
    {
        for (int q = 0; q < k; ++q) {
            const int i = idx[q];
            m[i] = m_[q]; t[i] = t_[q]; r2[i] = r2_[q];
            for (int d = 0; d < 3; ++d) {
                x[3 * i + d] = x_[3 * q + d]; v[3 * i + d] = v_[3 * q + d];
                a[3 * i + d] = a_[3 * q + d]; ad[3 * i + d] = ad_[3 * q + d];
            }
        }
    }
    void predict(double tn) override
    {
        for (int i = 0; i < n; ++i) {
            const double h = tn - t[i];
            for (int d = 0; d < 3; ++d) {
                const int q = 3 * i + d;
                if (h == 0) {
                    xp[q] = x[q];
                    vp[q] = v[q];
                } else {
                    xp[q] = ((ad[q] * h / 3 + a[q]) * h / 2 + v[q]) * h + x[q];
                    vp[q] = (ad[q] * h / 2 + a[q]) * h + v[q];
                }
            }
        }
    }
    void keep(int nt, const int *tg, const int *nnb, const int *list)
    {
        for (int k = 0; k < nt; ++k) {
            rnnb[tg[k]] = nnb[k];
            memcpy(&rlist[(size_t)tg[k] * kCap], &list[(size_t)k * kCap], kCap * sizeof(int));
        }
    }
    void query(int nt, const int *tg, double *acc, double *adot, int *nnb) override
    {
        std::vector<int> list((size_t)nt * kCap);
        oracle_regf(n, xp.data(), vp.data(), m.data(), r2.data(), nt, tg, acc, adot, nnb,
                    list.data(), kCap, nullptr, nullptr, 0);
        keep(nt, tg, nnb, list.data());
    }
    void gather(int nt, const int *tg, std::vector<int> &nnb, std::vector<int> &list)
    {
        nnb.resize(nt);
        list.resize((size_t)nt * kCap);
        for (int k = 0; k < nt; ++k) {
            nnb[k] = rnnb[tg[k]];
            memcpy(&list[(size_t)k * kCap], &rlist[(size_t)tg[k] * kCap], kCap * sizeof(int));
        }
    }
    void irregular(int nt, const int *tg, double *acc, double *adot) override
    {
        std::vector<int> nnb, list;
        gather(nt, tg, nnb, list);
        oracle_irregular(n, xp.data(), vp.data(), m.data(), nt, tg, nnb.data(), list.data(), kCap,
                         acc, adot);
    }
    void regular_step(int nt, const int *tg, const double *dt, const double *a_old, double *acc,
                      double *adot, int *nnb, double *out) override
    {
        std::vector<int> onnb, olist, nlist((size_t)nt * kCap);
        gather(nt, tg, onnb, olist);
        oracle_regf(n, xp.data(), vp.data(), m.data(), r2.data(), nt, tg, acc, adot, nnb,
                    nlist.data(), kCap, nullptr, nullptr, 0);
        oracle_regular_correct(n, xp.data(), vp.data(), m.data(), nt, tg, dt, a_old, acc, adot,
                               onnb.data(), olist.data(), nnb, nlist.data(), kCap, out);
        keep(nt, tg, nnb, nlist.data());
    }
    void derivs23(int nt, const int *tg, double *out) override
    {
        oracle_derivs23(n, x.data(), v.data(), a.data(), ad.data(), m.data(), r2.data(), nt, tg,
                        0, 0, out);
    }
    void update_radius(int nt, const int *tg, const int *nnb, int fix, int nmax,
                       double *r2_out) override
    {
        std::vector<int> c(nt);
        for (int k = 0; k < nt; ++k) {
            c[k] = nnb[k] < 0 ? -nnb[k] : nnb[k];
            r2_out[k] = r2[tg[k]];
        }
        oracle_update_radius(nt, c.data(), fix, nmax, r2_out);
        for (int k = 0; k < nt; ++k) r2[tg[k]] = r2_out[k];
    }
};

double norm3(const double *p) { return sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]); }

// Aarseth's step from force F and its derivatives F1, F2, F3 (norms).
double aarseth(double f, double f1, double f2, double f3, double eta)
{
    const double dt = sqrt(eta * (f * f2 + f1 * f1) / (f1 * f3 + f2 * f2));
    return (dt > 0 && isfinite(dt)) ? dt : 1e300;
}

// Largest power of two <= dt, <= cap, that divides t (t on the block grid).
double quantise(double dt, double cap, double t)
{
    double q = pow(2.0, floor(log2(std::max(std::min(dt, cap), 1e-30))));
    while (q > 1e-30 && fmod(t, q) != 0.0) q *= 0.5;
    return q;
}

struct Particle {
    double x[3], v[3];
    double FI[4][3], FR[4][3];   // irregular at t, regular at t_reg (force .. 3rd derivative)
    double t = 0, dt = 0, t_reg = 0, dt_reg = 0;
    int nnb = 0;
};

// Regular force at time tn, extrapolated linearly from the last regular step.
void reg_at(const Particle &p, double tn, double *f, double *f1)
{
    const double d = tn - p.t_reg;
    for (int k = 0; k < 3; ++k) {
        f[k] = p.FR[0][k] + p.FR[1][k] * d;
        f1[k] = p.FR[1][k];
    }
}

double energy(int n, const std::vector<Particle> &P, const std::vector<double> &m)
{
    double K = 0, U = 0;
    for (int i = 0; i < n; ++i) {
        K += 0.5 * m[i] * (P[i].v[0] * P[i].v[0] + P[i].v[1] * P[i].v[1] + P[i].v[2] * P[i].v[2]);
        for (int j = i + 1; j < n; ++j) {
            const double dx = P[j].x[0] - P[i].x[0], dy = P[j].x[1] - P[i].x[1],
                         dz = P[j].x[2] - P[i].x[2];
            U -= m[i] * m[j] / sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    return K + U;
}

}  // namespace

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s <gpu|oracle> <ic.bin> <t_end> <eta_irr> <eta_reg> <fix>\n",
                argv[0]);
        return 1;
    }
    const bool gpu = strcmp(argv[1], "gpu") == 0;
    const double t_end = atof(argv[3]), eta_i = atof(argv[4]), eta_r = atof(argv[5]);
    const int fix = atoi(argv[6]);
    int n = 0;
    FILE *f = fopen(argv[2], "rb");
    if (!f || fread(&n, sizeof(int), 1, f) != 1 || n <= 1) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 1;
    }
    std::vector<double> x(3 * n), v(3 * n), m(n), r2(n);
    if (fread(x.data(), 8, 3 * n, f) != (size_t)3 * n ||
        fread(v.data(), 8, 3 * n, f) != (size_t)3 * n || fread(m.data(), 8, n, f) != (size_t)n ||
        fread(r2.data(), 8, n, f) != (size_t)n) {
        fprintf(stderr, "short read\n");
        return 1;
    }
    fclose(f);
    Provider *pv = gpu ? static_cast<Provider *>(new GpuProvider())
                       : static_cast<Provider *>(new OracleProvider());
    const auto wall0 = std::chrono::steady_clock::now();
    std::vector<Particle> P(n);
    for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) {
            P[i].x[d] = x[3 * i + d];
            P[i].v[d] = v[3 * i + d];
        }
    const double E0 = energy(n, P, m);
    std::vector<int> all(n);
    for (int i = 0; i < n; ++i) all[i] = i;

    // state push of a set of particles (x, v, total force and jerk at t)
    auto push = [&](const std::vector<int> &idx, bool full) {
        const int k = (int)idx.size();
        std::vector<double> mm(k), xx(3 * k), vv(3 * k), aa(3 * k), jj(3 * k), tt(k), rr(k);
        for (int q = 0; q < k; ++q) {
            const Particle &p = P[idx[q]];
            double fr[3], fr1[3];
            reg_at(p, p.t, fr, fr1);
            mm[q] = m[idx[q]];
            tt[q] = p.t;
            rr[q] = r2[idx[q]];
            for (int d = 0; d < 3; ++d) {
                xx[3 * q + d] = p.x[d];
                vv[3 * q + d] = p.v[d];
                aa[3 * q + d] = p.FI[0][d] + fr[d];
                jj[3 * q + d] = p.FI[1][d] + fr1[d];
            }
        }
        if (full)
            pv->upload_state(n, mm.data(), xx.data(), vv.data(), aa.data(), jj.data(), tt.data(),
                             rr.data());
        else
            pv->update_state(k, idx.data(), mm.data(), xx.data(), vv.data(), aa.data(),
                             jj.data(), tt.data(), rr.data());
    };

    // ---- start-up
    for (auto &p : P) memset(p.FI, 0, sizeof(p.FI)), memset(p.FR, 0, sizeof(p.FR));
    push(all, true);
    pv->predict(0.0);
    std::vector<double> acc(3 * n), jerk(3 * n), ai(3 * n), ji(3 * n), d23(12 * (size_t)n);
    std::vector<int> nnb(n);
    pv->query(n, all.data(), acc.data(), jerk.data(), nnb.data());
    pv->irregular(n, all.data(), ai.data(), ji.data());
    for (int i = 0; i < n; ++i) {
        P[i].nnb = nnb[i];
        for (int d = 0; d < 3; ++d) {
            P[i].FR[0][d] = acc[3 * i + d];
            P[i].FR[1][d] = jerk[3 * i + d];
            P[i].FI[0][d] = ai[3 * i + d];
            P[i].FI[1][d] = ji[3 * i + d];
        }
    }
    push(all, false);   // total force and jerk for the derivatives
    pv->derivs23(n, all.data(), d23.data());
    const double dtmax = 1.0 / 32.0;
    for (int i = 0; i < n; ++i) {
        Particle &p = P[i];
        for (int d = 0; d < 3; ++d) {
            p.FR[2][d] = d23[12 * (size_t)i + d];
            p.FR[3][d] = d23[12 * (size_t)i + 3 + d];
            p.FI[2][d] = d23[12 * (size_t)i + 6 + d];
            p.FI[3][d] = d23[12 * (size_t)i + 9 + d];
        }
        p.dt_reg = quantise(aarseth(norm3(p.FR[0]), norm3(p.FR[1]), norm3(p.FR[2]),
                                    norm3(p.FR[3]), eta_r), dtmax, 0.0);
        p.dt = p.nnb > 0 ? quantise(aarseth(norm3(p.FI[0]), norm3(p.FI[1]), norm3(p.FI[2]),
                                            norm3(p.FI[3]), eta_i), p.dt_reg, 0.0)
                         : p.dt_reg;
    }

    long long blocks = 0, irr_steps = 0, reg_steps = 0, nnb_sum = 0, overflows = 0;
    std::vector<int> B, R;
    for (;;) {
        double tn = 1e300;
        for (int i = 0; i < n; ++i)
            if (P[i].t < t_end) tn = std::min(tn, P[i].t + P[i].dt);
        if (tn > t_end) break;
        B.clear();
        R.clear();
        for (int i = 0; i < n; ++i)
            if (P[i].t < t_end && P[i].t + P[i].dt == tn) {
                B.push_back(i);
                if (P[i].t_reg + P[i].dt_reg == tn) R.push_back(i);
            }
        const int nb = (int)B.size();
        // irregular step of the block over the old lists
        pv->predict(tn);
        pv->irregular(nb, B.data(), ai.data(), ji.data());
        for (int k = 0; k < nb; ++k) {
            Particle &p = P[B[k]];
            const double h = tn - p.t;
            double fr[3], fr1[3];
            reg_at(p, p.t, fr, fr1);
            for (int d = 0; d < 3; ++d) {
                const double A0 = p.FI[0][d] + fr[d], J0 = p.FI[1][d] + fr1[d];
                const double xpd = ((J0 * h / 3 + A0) * h / 2 + p.v[d]) * h + p.x[d];
                const double vpd = (J0 * h / 2 + A0) * h + p.v[d];
                const double f0 = p.FI[0][d], f1 = p.FI[1][d], g0 = ai[3 * k + d],
                             g1 = ji[3 * k + d];
                const double a2 = (-6 * (f0 - g0) - h * (4 * f1 + 2 * g1)) / (h * h);
                const double a3 = (12 * (f0 - g0) + 6 * h * (f1 + g1)) / (h * h * h);
                p.x[d] = xpd + a2 * h * h * h * h / 24 + a3 * h * h * h * h * h / 120;
                p.v[d] = vpd + a2 * h * h * h / 6 + a3 * h * h * h * h / 24;
                p.FI[0][d] = g0;
                p.FI[1][d] = g1;
                p.FI[2][d] = a2 + a3 * h;
                p.FI[3][d] = a3;
            }
            p.t = tn;
        }
        push(B, false);
        irr_steps += nb;
        // regular steps due now
        const int nr = (int)R.size();
        if (nr > 0) {
            pv->predict(tn);   // corrected block positions into the j-set
            std::vector<double> dtr(nr), a_old(6 * (size_t)nr), out(18 * (size_t)nr);
            for (int k = 0; k < nr; ++k) {
                const Particle &p = P[R[k]];
                dtr[k] = p.dt_reg;
                for (int d = 0; d < 3; ++d) {
                    a_old[6 * (size_t)k + d] = p.FR[0][d];
                    a_old[6 * (size_t)k + 3 + d] = p.FR[1][d];
                }
            }
            pv->regular_step(nr, R.data(), dtr.data(), a_old.data(), acc.data(), jerk.data(),
                             nnb.data(), out.data());
            std::vector<double> r2new(nr);
            pv->update_radius(nr, R.data(), nnb.data(), fix, kCap, r2new.data());
            for (int k = 0; k < nr; ++k) {
                r2[R[k]] = r2new[k];
                overflows += nnb[k] < 0;
            }
            for (int k = 0; k < nr; ++k) {
                Particle &p = P[R[k]];
                const double *o = &out[18 * (size_t)k];
                const double hr = p.dt_reg;
                for (int d = 0; d < 3; ++d) {
                    p.x[d] = o[d];
                    p.v[d] = o[3 + d];
                    p.FR[0][d] = acc[3 * k + d];
                    p.FR[1][d] = jerk[3 * k + d];
                    p.FR[2][d] = o[6 + d] + o[9 + d] * hr;
                    p.FR[3][d] = o[9 + d];
                    p.FI[0][d] = o[12 + d];
                    p.FI[1][d] = o[15 + d];
                }
                p.nnb = nnb[k] < 0 ? -nnb[k] : nnb[k];
                nnb_sum += p.nnb;
                p.t_reg = tn;
                p.dt_reg = quantise(aarseth(norm3(p.FR[0]), norm3(p.FR[1]), norm3(p.FR[2]),
                                            norm3(p.FR[3]), eta_r),
                                    std::min(dtmax, 2 * p.dt_reg), tn);
            }
            push(R, false);
            reg_steps += nr;
        }
        // next irregular steps (dividing what is left of the regular step)
        for (int k = 0; k < nb; ++k) {
            Particle &p = P[B[k]];
            const double left = p.t_reg + p.dt_reg - tn;
            p.dt = p.nnb > 0 ? quantise(aarseth(norm3(p.FI[0]), norm3(p.FI[1]), norm3(p.FI[2]),
                                                norm3(p.FI[3]), eta_i),
                                        std::min(left, 2 * p.dt), tn)
                             : left;
        }
        blocks += 1;
    }
    const double E1 = energy(n, P, m);
    const double wall =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
    delete pv;
    double cx = 0, cy = 0, cz = 0;
    for (int i = 0; i < n; ++i) {
        cx += m[i] * P[i].x[0];
        cy += m[i] * P[i].x[1];
        cz += m[i] * P[i].x[2];
    }
    printf("{\"provider\": \"%s\", \"n\": %d, \"t_end\": %.6f, \"eta_irr\": %g, \"eta_reg\": %g, "
           "\"E0\": %.15e, \"E1\": %.15e, \"dE_over_E\": %.6e, \"blocks\": %lld, "
           "\"irregular_steps\": %lld, \"regular_steps\": %lld, \"mean_nnb\": %.2f, "
           "\"overflows\": %lld, "
           "\"com\": [%.6e, %.6e, %.6e], \"x0\": [%.15e, %.15e, %.15e], \"wall_s\": %.3f}\n",
           argv[1], n, t_end, eta_i, eta_r, E0, E1, (E1 - E0) / fabs(E0), blocks, irr_steps,
           reg_steps, (double)nnb_sum / std::max(1LL, reg_steps), overflows, cx, cy, cz, P[0].x[0],
           P[0].x[1], P[0].x[2], wall);
    return 0;
}
