/*
 * ORACLE -- test infrastructure only.
 *
 * Plain-C FP64 restatement of ABYSS's regular-force + neighbour-list step.
 * It is the checker for the HIP library and the CPU baseline that bench.py
 * times; nothing in the product path links, loads or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may use it.
 *
 * What it restates (reference paths relative to the reference checkout):
 *   - pair formula and split rule:
 *       src/cuda/cuda_kernels.cu:283-317 (FP64 compute_forces, serial build)
 *       src_mpi/cuda/cuda_kernels.cu:331-364 (same rule, float build)
 *       src_mpi/Particle/ComputeAcceleration.cpp:253-299 (CPU FP64 loop:
 *         m_r3 = m/r2/sqrt(r2); a += m_r3*x; adot += m_r3*(v - 3*x*vx/r2))
 *     j is REGULAR when r2 > r_nb,i^2; otherwise, if j != i by index, j is a
 *     NEIGHBOUR (src_mpi/Particle/Initialize.cpp:67 uses the same "> is
 *     regular, <= is neighbour" split).  Self-exclusion is by index only, so a
 *     coincident j != i (r2 == 0) is a neighbour and adds no force.
 *   - output meaning (src_mpi/cuda/cuda_functions.h:3-10, reduction and gather
 *     at src_mpi/cuda/cuda_kernels.cu:384-450): acc/adot hold the regular part
 *     only; the neighbour list of target k is written at row k with stride
 *     `cap`, in ascending j order (gather concatenates the j-blocks in order).
 *   - the target's own r_nb^2 is used (r2[i_ptcl], src_mpi/cuda/cuda_kernels.cu:290).
 *
 * Overflow (count > cap) is an error in the reference (device assert,
 * src_mpi/cuda/cuda_kernels.cu:363,432).  Here the first `cap` entries are
 * written and nnb[k] = -(true count), the convention the HIP library uses.
 *
 * Parity pinning (tests/test_oracle.py): two-body closed forms,
 * regular + irregular = the unsplit direct sum (oracle_irregular), the
 * corrector's cubic closed form and split invariance (oracle_regular_correct),
 * finite-difference derivatives of the pair force (oracle_derivs23), the
 * radius rule at its end points (oracle_update_radius),
 * a = -grad(phi)
 * and jdot = d a/dt by finite differences, momentum conservation, homogeneity
 * scalings, r_nb -> 0 / inf limits, the exact "<=" boundary case, coincident
 * particles, and a vectorised numpy brute force on small N.
 *
 * Build: oracle/build_oracle.py (gcc -O3 -fopenmp -ffp-contract=off).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* One target against all j.  Returns the true neighbour count. */
static int oracle_one(int n, const double *x, const double *v, const double *m,
                      int i, double rnb2, double *a_out, double *j_out,
                      int *list, int cap, double *scale_a, double *scale_j)
{
    const double xi = x[3 * i + 0], yi = x[3 * i + 1], zi = x[3 * i + 2];
    const double vxi = v[3 * i + 0], vyi = v[3 * i + 1], vzi = v[3 * i + 2];
    double ax = 0, ay = 0, az = 0, jx = 0, jy = 0, jz = 0;
    double sa = 0, sj = 0;
    int cnt = 0;
    for (int j = 0; j < n; ++j) {
        const double dx = x[3 * j + 0] - xi;
        const double dy = x[3 * j + 1] - yi;
        const double dz = x[3 * j + 2] - zi;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 > rnb2) {
            const double dvx = v[3 * j + 0] - vxi;
            const double dvy = v[3 * j + 1] - vyi;
            const double dvz = v[3 * j + 2] - vzi;
            const double vx = dvx * dx + dvy * dy + dvz * dz;
            const double m_r3 = m[j] / r2 / sqrt(r2);
            ax += m_r3 * dx;
            ay += m_r3 * dy;
            az += m_r3 * dz;
            jx += m_r3 * (dvx - 3 * dx * vx / r2);
            jy += m_r3 * (dvy - 3 * dy * vx / r2);
            jz += m_r3 * (dvz - 3 * dz * vx / r2);
            /* magnitude scales used to normalise errors: sum_j |term_j| */
            const double r = sqrt(r2);
            sa += fabs(m[j]) / r2;
            sj += fabs(m_r3) * (sqrt(dvx * dvx + dvy * dvy + dvz * dvz) + 3.0 * fabs(vx) / r);
        } else if (j != i) {
            if (cnt < cap) list[cnt] = j;
            ++cnt;
        }
    }
    a_out[0] = ax; a_out[1] = ay; a_out[2] = az;
    j_out[0] = jx; j_out[1] = jy; j_out[2] = jz;
    if (scale_a) *scale_a = sa;
    if (scale_j) *scale_j = sj;
    return cnt;
}

/*
 * Regular force + neighbour lists for `nt` targets.
 *   x, v: [n][3] FP64, m: [n], rnb2: [n] (r_nb^2 per particle, as uploaded by
 *   SendToDevice), targets: [nt] indices into the j-set.
 *   acc, adot: [nt][3]; nnb: [nt]; list: [nt][cap]
 *   scale_a, scale_j: optional [nt] sums of |term| (may be NULL)
 * Returns the number of targets whose list overflowed `cap`.
 */
int oracle_regf(int n, const double *x, const double *v, const double *m,
                const double *rnb2, int nt, const int *targets,
                double *acc, double *adot, int *nnb, int *list, int cap,
                double *scale_a, double *scale_j, int nthreads)
{
    int overflow = 0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : overflow)
#endif
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        int c = oracle_one(n, x, v, m, i, rnb2[i], acc + 3 * k, adot + 3 * k,
                           list + (int64_t)k * cap, cap,
                           scale_a ? scale_a + k : NULL,
                           scale_j ? scale_j + k : NULL);
        if (c > cap) {
            nnb[k] = -c;
            ++overflow;
        } else {
            nnb[k] = c;
        }
    }
    (void)nthreads;
    return overflow;
}

/*
 * Pairs whose r^2 lies within `rel` of r_nb,i^2 (|r2 - rnb2| <= rel*rnb2),
 * j != i.  For each target, up to `bcap` such j are written (ascending) and
 * the count is returned in bcnt[k].  These are the only pairs whose
 * neighbour/regular classification may legitimately differ between a
 * reduced-precision evaluation and this FP64 oracle.
 */
void oracle_band_pairs(int n, const double *x, const double *rnb2, int nt,
                       const int *targets, double rel, int *bcnt, int *blist,
                       int bcap, int nthreads)
{
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 16)
#endif
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        const double xi = x[3 * i], yi = x[3 * i + 1], zi = x[3 * i + 2];
        const double r2nb = rnb2[i];
        int c = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double dx = x[3 * j] - xi, dy = x[3 * j + 1] - yi,
                         dz = x[3 * j + 2] - zi;
            const double r2 = dx * dx + dy * dy + dz * dz;
            if (fabs(r2 - r2nb) <= rel * r2nb) {
                if (c < bcap) blist[(int64_t)k * bcap + c] = j;
                ++c;
            }
        }
        bcnt[k] = c;
    }
    (void)nthreads;
}

/*
 * Second-order j prediction used by the reference caller before upload:
 *   x_p = ((adot*dt/3 + a)*dt/2 + v)*dt + x,  v_p = (adot*dt/2 + a)*dt + v
 * (src_mpi/Particle/UpdateParticle.cpp:37-59; dt == 0 returns x, v).
 */
void oracle_predict2(int n, const double *x, const double *v, const double *a,
                     const double *adot, const double *dt, double *xp,
                     double *vp)
{
    for (int i = 0; i < n; ++i) {
        const double h = dt[i];
        for (int d = 0; d < 3; ++d) {
            const int q = 3 * i + d;
            if (h == 0) {
                xp[q] = x[q];
                vp[q] = v[q];
            } else {
                xp[q] = ((adot[q] * h / 3 + a[q]) * h / 2 + v[q]) * h + x[q];
                vp[q] = (adot[q] * h / 2 + a[q]) * h + v[q];
            }
        }
    }
}

/*
 * Irregular force and jerk of each target over its neighbour list, FP64:
 *   m_r3 = m_j/(r2*sqrt(r2)); a += m_r3*x; adot += m_r3*(v - 3*x*(x.v)/r2)
 * with x = x_j - x_i, v = v_j - v_i, summed in list order
 * (Particle::computeAccelerationIrr, src_mpi/Particle/ComputeAcceleration.cpp:14-190).
 * Row k of `list` (stride cap) holds min(|nnb[k]|, cap) indices.  An entry at
 * r2 == 0 adds nothing, as in the regular split.
 */
void oracle_irregular(int n, const double *x, const double *v, const double *m, int nt,
                      const int *targets, const int *nnb, const int *list, int cap,
                      double *acc, double *adot)
{
    (void)n;
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        int c = nnb[k] < 0 ? -nnb[k] : nnb[k];
        if (c > cap) c = cap;
        double a[3] = {0, 0, 0}, ad[3] = {0, 0, 0};
        for (int e = 0; e < c; ++e) {
            const int j = list[(size_t)k * cap + e];
            double dx[3], dv[3];
            for (int d = 0; d < 3; ++d) {
                dx[d] = x[3 * j + d] - x[3 * i + d];
                dv[d] = v[3 * j + d] - v[3 * i + d];
            }
            const double r2 = dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2];
            if (r2 == 0.0) continue;
            const double rv = dx[0] * dv[0] + dx[1] * dv[1] + dx[2] * dv[2];
            const double m_r3 = m[j] / (r2 * sqrt(r2));
            for (int d = 0; d < 3; ++d) {
                a[d] += m_r3 * dx[d];
                ad[d] += m_r3 * (dv[d] - 3 * dx[d] * rv / r2);
            }
        }
        for (int d = 0; d < 3; ++d) {
            acc[3 * k + d] = a[d];
            adot[3 * k + d] = ad[d];
        }
    }
}

/*
 * Neighbour-set correction + fourth-order regular corrector of one regular
 * step, per target, in the reference's loop and expression order
 * (Particle::updateRegularParticleCuda,
 * src_mpi/Particle/ComputeAcceleration.cpp:400-661):
 *   for e < max(|L_old|, |L_new|):
 *     e < |L_old| and L_old[e] not in L_new: a_tmp -= f(L_old[e])
 *     e < |L_new|: a_irr += f(L_new[e]); if L_new[e] not in L_old: a_tmp += f
 *   f: m_r3 = m/dr2/sqrt(dr2); a += m_r3 dx; adot += m_r3 (dv - 3 dx dxdv/dr2)
 *   (dr2 == 0 adds nothing here; the reference would divide by zero)
 *   da_dt2 = (a0 - a1 - a_tmp)/dt^2; adot_dt = (j0 + j1 + adot_tmp)/dt
 *   a2 = -6 da_dt2 - 2 adot_dt - 2 j0/dt; a3 = (12 da_dt2 + 6 adot_dt)/dt
 *   x_new = x + a2 dt^4/24 + a3 dt^5/120; v_new = v + a2 dt^3/6 + a3 dt^4/24
 * x, v: the current (predicted) j-set; a_old [nt][6] = a0, j0; new_acc /
 * new_adot [nt][3] = a1, j1; lists as rows of stride cap with counts
 * (negative = overflowed, first cap entries used).
 * out [nt][18] = x_new, v_new, a2, a3, a_irr, adot_irr.
 */
static int in_row(const int *row, int n, int j)
{
    for (int e = 0; e < n; ++e)
        if (row[e] == j) return 1;
    return 0;
}

static void corr_term(const double *x, const double *v, const double *m, int i, int j,
                      double sgn, double *a, double *ad)
{
    double dx[3], dv[3], dr2 = 0.0, dxdv = 0.0;
    for (int d = 0; d < 3; ++d) {
        dx[d] = x[3 * j + d] - x[3 * i + d];
        dv[d] = v[3 * j + d] - v[3 * i + d];
        dr2 += dx[d] * dx[d];
        dxdv += dx[d] * dv[d];
    }
    if (dr2 == 0.0) return;
    const double m_r3 = m[j] / dr2 / sqrt(dr2);
    for (int d = 0; d < 3; ++d) {
        a[d] += sgn * (m_r3 * dx[d]);
        ad[d] += sgn * (m_r3 * (dv[d] - 3 * dx[d] * dxdv / dr2));
    }
}

void oracle_regular_correct(int n, const double *x, const double *v, const double *m, int nt,
                            const int *targets, const double *dts, const double *a_old,
                            const double *new_acc, const double *new_adot, const int *old_nnb,
                            const int *old_list, const int *new_nnb, const int *new_list,
                            int cap, double *out)
{
    (void)n;
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        int co = old_nnb[k] < 0 ? -old_nnb[k] : old_nnb[k];
        int cn = new_nnb[k] < 0 ? -new_nnb[k] : new_nnb[k];
        if (co > cap) co = cap;
        if (cn > cap) cn = cap;
        const int *lo = old_list + (size_t)k * cap;
        const int *ln = new_list + (size_t)k * cap;
        double a_tmp[3] = {0, 0, 0}, adot_tmp[3] = {0, 0, 0};
        double a_irr[3] = {0, 0, 0}, adot_irr[3] = {0, 0, 0};
        const int size = co > cn ? co : cn;
        for (int e = 0; e < size; ++e) {
            if (e < co && !in_row(ln, cn, lo[e]))
                corr_term(x, v, m, i, lo[e], -1.0, a_tmp, adot_tmp);
            if (e < cn) {
                corr_term(x, v, m, i, ln[e], 1.0, a_irr, adot_irr);
                if (!in_row(lo, co, ln[e])) corr_term(x, v, m, i, ln[e], 1.0, a_tmp, adot_tmp);
            }
        }
        const double dt = dts[k];
        const double dt2 = dt * dt, dt3 = dt2 * dt, dt4 = dt3 * dt, dt5 = dt4 * dt;
        double *o = out + (size_t)k * 18;
        for (int d = 0; d < 3; ++d) {
            const double a0 = a_old[6 * (size_t)k + d], j0 = a_old[6 * (size_t)k + 3 + d];
            const double da_dt2 = (a0 - new_acc[3 * (size_t)k + d] - a_tmp[d]) / dt2;
            const double adot_dt = (j0 + new_adot[3 * (size_t)k + d] + adot_tmp[d]) / dt;
            const double a2 = -6 * da_dt2 - 2 * adot_dt - 2 * j0 / dt;
            const double a3 = (12 * da_dt2 + 6 * adot_dt) / dt;
            o[d] = x[3 * i + d] + a2 * dt4 / 24 + a3 * dt5 / 120;
            o[3 + d] = v[3 * i + d] + a2 * dt3 / 6 + a3 * dt4 / 24;
            o[6 + d] = a2;
            o[9 + d] = a3;
            o[12 + d] = a_irr[d];
            o[15 + d] = adot_irr[d];
        }
    }
}

/*
 * Start-up second and third derivatives of the acceleration, split into
 * regular (r2 > r_nb,i^2) and irregular parts, FP64, in the reference's
 * expression order (CalculateAcceleration23,
 * src_mpi/Particle/Initialize.cpp:102-185).  x, v, a, adot: [n][3] at the
 * common start time (a, adot: total acceleration and jerk); every j != i is
 * summed; r2 == 0 adds nothing (the reference would divide by zero).
 * out [nt][12] = a_reg2[3], a_reg3[3], a_irr2[3], a_irr3[3].
 *
 * ref_signs = 1 keeps the reference's sign in the three projections
 * rdf_r2 = x.(a1 - a2)/r2, vdf_r2 = v.(a1 - a2)/r2, rdfdot_r2 =
 * x.(a1dot - a2dot)/r2 (a1: the target's, a2: j's, x = x_j - x_i), which
 * enter B and C with the opposite sign to the relative acceleration
 * x'' = a2 - a1 that the leading terms -m (a1 - a2)/r^3 use; ref_signs = 0
 * uses x.(a2 - a1) etc., which makes adot2 and adot3 the exact second and
 * third time derivatives of m x/|x|^3 along the relative orbit.
 */
void oracle_derivs23(int n, const double *X, const double *V, const double *Acc,
                     const double *Adot, const double *m, const double *rnb2, int nt,
                     const int *targets, int ref_signs, int nthreads, double *out)
{
    const double ps = ref_signs ? 1.0 : -1.0;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (int k = 0; k < nt; ++k) {
        const int i = targets[k];
        double reg2[3] = {0, 0, 0}, reg3[3] = {0, 0, 0}, irr2[3] = {0, 0, 0}, irr3[3] = {0, 0, 0};
        const double *a1 = Acc + 3 * (size_t)i, *a1dot = Adot + 3 * (size_t)i;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const double *a2 = Acc + 3 * (size_t)j, *a2dot = Adot + 3 * (size_t)j;
            double x[3], v[3], r2 = 0, vr = 0, v2 = 0;
            for (int d = 0; d < 3; ++d) {
                x[d] = X[3 * (size_t)j + d] - X[3 * (size_t)i + d];
                v[d] = V[3 * (size_t)j + d] - V[3 * (size_t)i + d];
                r2 += x[d] * x[d];
                vr += v[d] * x[d];
                v2 += v[d] * v[d];
            }
            if (r2 == 0.0) continue;
            const double r3 = r2 * sqrt(r2);
            const double m_r3 = m[j] / r3;
            double a21[3], a21dot[3], rdf_r2 = 0, vdf_r2 = 0, rdfdot_r2 = 0;
            for (int d = 0; d < 3; ++d) {
                a21[d] = m_r3 * x[d];
                a21dot[d] = m_r3 * (v[d] - 3 * x[d] * vr / r2);
                rdf_r2 += ps * (x[d] * (a1[d] - a2[d]) / r2);
                vdf_r2 += ps * (v[d] * (a1[d] - a2[d]) / r2);
                rdfdot_r2 += ps * (x[d] * (a1dot[d] - a2dot[d]) / r2);
            }
            const double A = vr / r2;
            const double B = v2 / r2 + rdf_r2 + A * A;
            const double C = 3 * vdf_r2 + rdfdot_r2 + A * (3 * B - 4 * A * A);
            const int reg = r2 > rnb2[i];
            for (int d = 0; d < 3; ++d) {
                const double adot2 = -m[j] * (a1[d] - a2[d]) / r3 - 6 * A * a21dot[d] - 3 * B * a21[d];
                const double adot3 = -m[j] * (a1dot[d] - a2dot[d]) / r3 - 9 * A * adot2 -
                                     9 * B * a21dot[d] - 3 * C * a21[d];
                if (reg) {
                    reg2[d] += adot2;
                    reg3[d] += adot3;
                } else {
                    irr2[d] += adot2;
                    irr3[d] += adot3;
                }
            }
        }
        double *o = out + (size_t)k * 12;
        for (int d = 0; d < 3; ++d) {
            o[d] = reg2[d];
            o[3 + d] = reg3[d];
            o[6 + d] = irr2[d];
            o[9 + d] = irr3[d];
        }
    }
}

/*
 * Neighbour-radius adaptation (Particle::updateRadius,
 * src_mpi/Particle/UpdateParticle.cpp:108-139).  r2 (the squared radius,
 * RadiusOfNeighbor) is scaled by (1 - a) with x = nnb - fix, c = nmax - fix:
 *   nnb > fix: a = 0.9/c^2 * |x| * x      (n = 2)
 *   nnb < fix: a = 0.5/c^3 * x^3          (n = 3; a < 0, the radius grows)
 *   nnb = fix: unchanged.
 */
void oracle_update_radius(int count, const int *nnb, int fix, int nmax, double *r2)
{
    for (int k = 0; k < count; ++k) {
        const double c = (double)(nmax - fix);
        const double x = (double)(nnb[k] - fix);
        if (nnb[k] > fix) {
            const double b = 0.9 / pow(c, 2);
            r2[k] *= (1. - b * fabs(x) * pow(x, 1));
        } else if (nnb[k] < fix) {
            const double b = 0.5 / pow(c, 3);
            r2[k] *= (1. - b * pow(x, 3));
        }
    }
}

int oracle_max_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
