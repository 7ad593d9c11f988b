// Regular-force + neighbour-list kernels for gfx950 (CDNA4, wave64).
//
// Work decomposition (one launch = one query):
//   grid.x : blocks of kBlock*IPL targets; each lane owns IPL targets and
//            keeps their FP64 position/velocity, r_nb^2, FP64 sums and
//            neighbour cursor in registers.
//   grid.y : S contiguous j-ranges ("splits", a multiple of the tile each).
//            S > 1 exposes parallelism when few targets are active; the
//            partial sums and per-split sub-lists are merged in split order
//            by finalize_kernel, so results do not depend on scheduling.
// Each workgroup streams its j-range through LDS in tiles of kTile j
// (double-buffered: the next tile's coalesced SoA loads are in flight while
// the current tile is evaluated; one barrier per tile).  Every lane of a
// wave reads the same j record from LDS (broadcast), so j costs 3 b128 + 1
// b32 LDS reads per wave and is reused by the wave's 64*IPL targets.
//
// Pair evaluation, split precision (the reference formula is
// src_mpi/cuda/cuda_kernels.cu:331-354, FP64 form src/cuda/cuda_kernels.cu
// :283-311):
//   dx = (float)(x_j - x_i)  [FP64 subtract, one rounding]   (same for v)
//   r2 = dx.dx (FP32), rinv = v_rsq_f32(r2)
//   regular  <=> r2 > r_nb,i^2 ; otherwise neighbour if j != i (index)
//   a    += m_j rinv^3 dx
//   jdot += m_j rinv^3 (dv - 3 (dx.dv) rinv^2 dx)
// A neighbour (or the target itself, or a coincident j) gets rinv = 0, so it
// adds exactly zero and never produces inf/NaN.  The FP32 terms are summed
// over kFold consecutive j and each such partial is added to an FP64 sum.
// Neighbours are appended to the lane's own list in ascending j order; the
// append branch is taken only for neighbours, so the self test (j != i)
// costs nothing on the common path.
#include "../../include/abyss_regf.h"
#include "regf_internal.h"

namespace abyss {
namespace {

template <bool F64>
struct MassT { using type = float; };
template <>
struct MassT<true> { using type = double; };

// Loads one j record per thread from the SoA j-set (coalesced) and stores
// it into an LDS tile as 3 x double2 {x,y} {z,vx} {vy,vz} plus the mass.
// Records past the end of the j-range are zero-filled so that kernels which
// read whole tiles never see stale LDS contents.
template <typename MT>
struct TileLoader {
    double rx = 0, ry = 0, rz = 0, rvx = 0, rvy = 0, rvz = 0;
    MT rm = 0;
    __device__ __forceinline__ void load(const JState &js, int j, int j1)
    {
        if (j < j1) {
            rx = js.x[j]; ry = js.y[j]; rz = js.z[j];
            rvx = js.vx[j]; rvy = js.vy[j]; rvz = js.vz[j];
            if constexpr (sizeof(MT) == 8) rm = js.m[j]; else rm = js.m32[j];
        } else {
            rx = ry = rz = rvx = rvy = rvz = 0.0;
            rm = 0;
        }
    }
    __device__ __forceinline__ void store(double2 *pos, MT *mass, int r) const
    {
        pos[3 * r + 0] = make_double2(rx, ry);
        pos[3 * r + 1] = make_double2(rz, rvx);
        pos[3 * r + 2] = make_double2(rvy, rvz);
        mass[r] = rm;
    }
};

template <int IPL, bool F64>
struct Targets {
    double x[IPL], y[IPL], z[IPL], vx[IPL], vy[IPL], vz[IPL];
    double acc[IPL][6];       // FP64 per-particle sums (a, jdot)
    float rnbf[IPL];          // r_nb^2 (FP32 compare, split kernel)
    double rnbd[IPL];         // r_nb^2 (FP64 compare, FP64 kernel)
    int idx[IPL];             // j-index of the target (self exclusion)
    int cnt[IPL];             // neighbours found so far in this j-range
    int *lst[IPL];            // this (split, target)'s list row
};

// Append j to target k's list.  SELF: the tile may contain the target
// itself, which must be skipped (self-exclusion is by index).
template <int IPL, bool F64, bool SELF>
__device__ __forceinline__ void append(Targets<IPL, F64> &T, int k, int jg, int cap)
{
    if (!SELF || jg != T.idx[k]) {
        if (T.cnt[k] < cap) T.lst[k][T.cnt[k]] = jg;
        T.cnt[k] += 1;
    }
}

// Evaluate j records [jb, je) of LDS buffer (pos, mass) against all targets.
// jg0 is the global index of record 0 of the buffer.
template <int IPL, int U, bool F64, bool FULL, bool SELF>
__device__ __forceinline__ void run_tile(Targets<IPL, F64> &T, const double2 *__restrict__ pos,
                                         const typename MassT<F64>::type *__restrict__ mass,
                                         int count, int jg0, int cap)
{
    if constexpr (!F64) {
        for (int f = 0; f < (FULL ? kTile : count); f += kFold) {
            float pa[IPL][6];
#pragma unroll
            for (int k = 0; k < IPL; ++k)
#pragma unroll
                for (int c = 0; c < 6; ++c) pa[k][c] = 0.0f;
            const int fe = FULL ? kFold : min(kFold, count - f);
            // partial tiles: a run-time trip count, which the neighbour
            // branch keeps from being unrolled with a remainder loop
            constexpr int kUnroll = FULL ? U : 1;
#pragma unroll kUnroll
            for (int jj = 0; jj < fe; ++jj) {
                const int r = f + jj;
                const double2 p0 = pos[3 * r + 0];   // x, y
                const double2 p1 = pos[3 * r + 1];   // z, vx
                const double2 p2 = pos[3 * r + 2];   // vy, vz
                const float mj = mass[r];
#pragma unroll
                for (int k = 0; k < IPL; ++k) {
                    const float dx = (float)(p0.x - T.x[k]);
                    const float dy = (float)(p0.y - T.y[k]);
                    const float dz = (float)(p1.x - T.z[k]);
                    const float dvx = (float)(p1.y - T.vx[k]);
                    const float dvy = (float)(p2.x - T.vy[k]);
                    const float dvz = (float)(p2.y - T.vz[k]);
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    float rinv = __builtin_amdgcn_rsqf(r2);
                    // Neighbour (or self, or coincident j): rare branch that
                    // records j and zeroes rinv, so the pair adds exactly 0.
                    if (!(r2 > T.rnbf[k])) {
                        // volatile keeps the zeroing inside the rare branch
                        // (a select would cost every pair an instruction)
                        asm volatile("v_mov_b32 %0, 0" : "=v"(rinv));
                        append<IPL, F64, SELF>(T, k, jg0 + r, cap);
                    }
                    const float rinv2 = rinv * rinv;
                    const float mr3 = mj * rinv * rinv2;
                    const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
                    const float al = -3.0f * rv * rinv2;
                    pa[k][0] = fmaf(mr3, dx, pa[k][0]);
                    pa[k][1] = fmaf(mr3, dy, pa[k][1]);
                    pa[k][2] = fmaf(mr3, dz, pa[k][2]);
                    pa[k][3] = fmaf(mr3, fmaf(al, dx, dvx), pa[k][3]);
                    pa[k][4] = fmaf(mr3, fmaf(al, dy, dvy), pa[k][4]);
                    pa[k][5] = fmaf(mr3, fmaf(al, dz, dvz), pa[k][5]);
                }
            }
#pragma unroll
            for (int k = 0; k < IPL; ++k)
#pragma unroll
                for (int c = 0; c < 6; ++c) T.acc[k][c] += (double)pa[k][c];
        }
    } else {
        const int c_end = FULL ? kTile : count;
#pragma unroll U
        for (int r = 0; r < c_end; ++r) {
            const double2 p0 = pos[3 * r + 0];
            const double2 p1 = pos[3 * r + 1];
            const double2 p2 = pos[3 * r + 2];
            const double mj = mass[r];
#pragma unroll
            for (int k = 0; k < IPL; ++k) {
                const double dx = p0.x - T.x[k];
                const double dy = p0.y - T.y[k];
                const double dz = p1.x - T.z[k];
                const double dvx = p1.y - T.vx[k];
                const double dvy = p2.x - T.vy[k];
                const double dvz = p2.y - T.vz[k];
                const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
                double rinv = rsqrt(r2);
                if (!(r2 > T.rnbd[k])) {
                    rinv = 0.0;
                    append<IPL, F64, SELF>(T, k, jg0 + r, cap);
                }
                const double rinv2 = rinv * rinv;
                const double mr3 = mj * rinv * rinv2;
                const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
                const double al = -3.0 * rv * rinv2;
                T.acc[k][0] = fma(mr3, dx, T.acc[k][0]);
                T.acc[k][1] = fma(mr3, dy, T.acc[k][1]);
                T.acc[k][2] = fma(mr3, dz, T.acc[k][2]);
                T.acc[k][3] = fma(mr3, fma(al, dx, dvx), T.acc[k][3]);
                T.acc[k][4] = fma(mr3, fma(al, dy, dvy), T.acc[k][4]);
                T.acc[k][5] = fma(mr3, fma(al, dz, dvz), T.acc[k][5]);
            }
        }
    }
}

// One workgroup's share of a query: target block bx against j-split s.
// Shared by regf_kernel (one launch per query) and batch_kernel (work items
// of many clusters in one persistent launch), so both evaluate every pair
// with the same instructions in the same order.
template <int IPL, int U, bool F64>
// Partial outputs of target q in split s live at index s*nt_stride + q -
// q_base (q_base = 0, nt_stride = nt for a whole query; a batch slice of a
// split cluster uses its first target and its target count).
__device__ __forceinline__ void regf_body(const JState &js, const int *__restrict__ tgt, int nt,
                                          int jspan, const QueryOut &out, int cap,
                                          int final_counts, int bx, int s, int q_base,
                                          int nt_stride,
                                          double2 (*s_pos)[kTile * 3],
                                          typename MassT<F64>::type (*s_m)[kTile])
{
    using MT = typename MassT<F64>::type;
    const int tid = threadIdx.x;
    const int j0 = s * jspan;
    const int j1 = min(js.n, j0 + jspan);

    Targets<IPL, F64> T;
#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int q = bx * (kBlock * IPL) + k * kBlock + tid;
        if (q < nt) {
            const int i = tgt[q];
            T.idx[k] = i;
            T.x[k] = js.x[i]; T.y[k] = js.y[i]; T.z[k] = js.z[i];
            T.vx[k] = js.vx[i]; T.vy[k] = js.vy[i]; T.vz[k] = js.vz[i];
            T.rnbd[k] = js.rnb2[i];
            T.lst[k] = out.list + ((size_t)s * nt_stride + (q - q_base)) * (size_t)cap;
        } else {
            // Inactive lane: r_nb^2 = -1 makes every j regular, so it never
            // appends; its sums are discarded.
            T.idx[k] = -1;
            T.x[k] = T.y[k] = T.z[k] = 0.0;
            T.vx[k] = T.vy[k] = T.vz[k] = 0.0;
            T.rnbd[k] = -1.0;
            T.lst[k] = nullptr;
        }
        T.rnbf[k] = (float)T.rnbd[k];
        T.cnt[k] = 0;
#pragma unroll
        for (int c = 0; c < 6; ++c) T.acc[k][c] = 0.0;
    }

    const int ntile = (j1 - j0 + kTile - 1) / kTile;
    TileLoader<MT> ld;
    static_assert(kTile == kBlock, "one j per thread per tile load");
    // A wave none of whose lanes holds a target (the tail of a small block)
    // still loads its share of every tile and meets every barrier, but skips
    // the pair loop.
    bool any_target = false;
#pragma unroll
    for (int k = 0; k < IPL; ++k) any_target |= T.idx[k] >= 0;
    const bool wave_busy = __any(any_target);

    if (ntile > 0) {
        ld.load(js, j0 + tid, j1);
        ld.store(s_pos[0], s_m[0], tid);
    }
    __syncthreads();
    for (int t = 0; t < ntile; ++t) {
        const int b = t & 1;
        if (t + 1 < ntile) ld.load(js, j0 + (t + 1) * kTile + tid, j1);
        const int jt = j0 + t * kTile;
        const int count = min(kTile, j1 - jt);
        // Wave vote: does this tile hold the target of any lane?  Only then
        // does the tile need the per-neighbour self test.
        bool mine = false;
#pragma unroll
        for (int k = 0; k < IPL; ++k) mine |= (unsigned)(T.idx[k] - jt) < (unsigned)count;
        if (!wave_busy) {
            // no target in this wave: nothing to evaluate
        } else if (__any(mine)) {
            run_tile<IPL, U, F64, false, true>(T, s_pos[b], s_m[b], count, jt, cap);
        } else if (count == kTile) {
            run_tile<IPL, U, F64, true, false>(T, s_pos[b], s_m[b], count, jt, cap);
        } else {
            run_tile<IPL, U, F64, false, false>(T, s_pos[b], s_m[b], count, jt, cap);
        }
        if (t + 1 < ntile) ld.store(s_pos[b ^ 1], s_m[b ^ 1], tid);
        __syncthreads();
    }

#pragma unroll
    for (int k = 0; k < IPL; ++k) {
        const int q = bx * (kBlock * IPL) + k * kBlock + tid;
        if (q < nt) {
            const size_t o = (size_t)s * nt_stride + (q - q_base);
            out.acc[3 * o + 0] = T.acc[k][0];
            out.acc[3 * o + 1] = T.acc[k][1];
            out.acc[3 * o + 2] = T.acc[k][2];
            out.adot[3 * o + 0] = T.acc[k][3];
            out.adot[3 * o + 1] = T.acc[k][4];
            out.adot[3 * o + 2] = T.acc[k][5];
            const int c = T.cnt[k];
            out.cnt[o] = (final_counts && c > cap) ? -c : c;
        }
    }
}

template <int IPL, int U, bool F64>
__global__ __launch_bounds__(kBlock) void regf_kernel(JState js, const int *__restrict__ tgt,
                                                      int nt, int jspan, QueryOut out, int cap,
                                                      int final_counts)
{
    __shared__ double2 s_pos[2][kTile * 3];
    __shared__ typename MassT<F64>::type s_m[2][kTile];
    regf_body<IPL, U, F64>(js, tgt, nt, jspan, out, cap, final_counts, blockIdx.x, blockIdx.y,
                           0, nt, s_pos, s_m);
}

// Lanes-per-target variant: JL lanes share one target (JL = 4, 16 or 64),
// so a wave serves 64/JL targets and a few active particles still fill the
// chip.  In a tile, step t hands the group's lanes the JL consecutive
// records t*JL + sub (sub = lane mod JL).  Neighbours of a step are
// compacted with a wave ballot: a lane's slot is the group's running count
// plus the number of neighbour lanes below it in the group, which keeps
// every list in ascending j.  At the end the JL FP64 sums of a group are
// combined by a fixed xor-butterfly (bitwise reproducible).
//
// Term modes TM:
//   0  FP64 differences rounded to FP32; r^2, rsqrt and terms in FP32; FP32
//      partials folded into FP64 sums at the end of the tile in which a lane
//      reaches kFold pairs
//   1  FP64 differences, r^2 and terms, accumulated in FP64; rsqrt is the
//      FP32 v_rsq_f32 of the rounded r^2 (about 1e-7 relative); the
//      neighbour test compares FP64 r^2 with FP64 r_nb^2
//   2  as 1, with one Newton step y += y (1 - r^2 y^2) / 2 in FP64 on the
//      rsqrt (relative error ~1e-14)
//   3  terms as 2, but the neighbour test is that of mode 0 (FP32 r^2 of
//      the differences rounded to FP32 against FP32 r_nb^2), so a pair is
//      regular exactly when mode 0 finds it so; the recomputation of
//      checked precision
// FP64 vector instructions issue at the FP32 rate on MI355X, so modes 1 to
// 3 trade the six FP64->FP32 conversions for FP64 arithmetic.
template <int TM>
struct TermMass { using type = double; };
template <>
struct TermMass<0> { using type = float; };

// Per-lane state of the lanes-per-target kernel.
struct JLLane {
    double xi, yi, zi, vxi, vyi, vzi;
    double rnbd;     // r_nb^2 (modes 1, 2); r_nb^2 (1 + 1e-6) (mode 3)
    float rnbf;      // r_nb^2 (mode 0)
    int idx, sub, cnt, cap;
    int *lst;
    unsigned long long gmask, below;
    float pa[6];     // mode 0: FP32 partials
    double acc[6];   // FP64 sums
};

// Rare branch of a step: zero the near lanes' rsqrt here rather than with a
// select on every pair (volatile keeps it inside the branch), and append the
// listed neighbours in lane order.
template <bool FULL, bool SELF>
__device__ __forceinline__ void jl_near(JLLane &L, bool near, unsigned long long nb, int r,
                                        int count, int jt, float &rinv)
{
    if (near) asm volatile("v_mov_b32 %0, 0" : "=v"(rinv));
    bool listed = near;
    if (!FULL) listed = listed && r < count;          // zero-filled records
    if (SELF) listed = listed && jt + r != L.idx;     // self-exclusion by index
    const unsigned long long bal = (FULL && !SELF) ? nb : __ballot(listed);
    const unsigned long long gb = bal & L.gmask;
    if (listed) {
        const int p = L.cnt + __popcll(gb & L.below);
        if (p < L.cap) L.lst[p] = jt + r;
    }
    L.cnt += __popcll(gb);
}

// Steps of one tile of TILE records.  FULL: all TILE records are valid;
// SELF: the tile may hold a lane's own target.  The common case (FULL,
// !SELF) lists exactly the near pairs, so its only extra instruction over
// the force is the compare.
template <int JL, int TM, bool FULL, bool SELF, int TILE>
__device__ __forceinline__ void jl_tile(JLLane &L, const double2 *__restrict__ pos,
                                        const typename TermMass<TM>::type *__restrict__ mass,
                                        int count, int jt)
{
    const int steps = FULL ? TILE / JL : (count + JL - 1) / JL;
    constexpr int kUnroll = FULL ? 2 : 1;   // partial tiles: run-time trip count
#pragma unroll kUnroll
    for (int st = 0; st < steps; ++st) {
        const int r = st * JL + L.sub;
        const double2 p0 = pos[3 * r + 0];
        const double2 p1 = pos[3 * r + 1];
        const double2 p2 = pos[3 * r + 2];
        if constexpr (TM == 0) {
            const float mj = mass[r];
            const float dx = (float)(p0.x - L.xi);
            const float dy = (float)(p0.y - L.yi);
            const float dz = (float)(p1.x - L.zi);
            const float dvx = (float)(p1.y - L.vxi);
            const float dvy = (float)(p2.x - L.vyi);
            const float dvz = (float)(p2.y - L.vzi);
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            float rinv = __builtin_amdgcn_rsqf(r2);
            const bool near = !(r2 > L.rnbf);
            const unsigned long long nb = __ballot(near);
            if (nb) jl_near<FULL, SELF>(L, near, nb, r, count, jt, rinv);
            const float rinv2 = rinv * rinv;
            const float mr3 = mj * rinv * rinv2;
            const float rv = fmaf(dz, dvz, fmaf(dy, dvy, dx * dvx));
            const float al = -3.0f * rv * rinv2;
            L.pa[0] = fmaf(mr3, dx, L.pa[0]);
            L.pa[1] = fmaf(mr3, dy, L.pa[1]);
            L.pa[2] = fmaf(mr3, dz, L.pa[2]);
            L.pa[3] = fmaf(mr3, fmaf(al, dx, dvx), L.pa[3]);
            L.pa[4] = fmaf(mr3, fmaf(al, dy, dvy), L.pa[4]);
            L.pa[5] = fmaf(mr3, fmaf(al, dz, dvz), L.pa[5]);
        } else {
            const double mj = mass[r];
            const double dx = p0.x - L.xi;
            const double dy = p0.y - L.yi;
            const double dz = p1.x - L.zi;
            const double dvx = p1.y - L.vxi;
            const double dvy = p2.x - L.vyi;
            const double dvz = p2.y - L.vzi;
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            // rsqrt seed of the FP64 r^2 (modes 2 and 3 refine it below)
            float rf = __builtin_amdgcn_rsqf((float)r2);
            if constexpr (TM == 3) {
                // The neighbour test must be mode 0's: the FP32 r^2 of the
                // rounded differences against the FP32 r_nb^2.  That r^2 is
                // within 5 u (u = 2^-24) of the FP64 one and the FP32 r_nb^2
                // within u of r_nb^2, so an FP64 r^2 above r_nb^2 (1 + 1e-6)
                // (L.rnbd in this mode) is regular under mode 0 as well; only
                // steps with a pair below that margin evaluate mode 0's
                // expression, in this rare branch.
                if (__any(!(r2 > L.rnbd))) {
                    float fx, fy, fz;
                    // volatile: the converts (and the test after them) stay
                    // in the branch instead of being computed for every pair
                    asm volatile("v_cvt_f32_f64 %0, %1" : "=v"(fx) : "v"(dx));
                    asm volatile("v_cvt_f32_f64 %0, %1" : "=v"(fy) : "v"(dy));
                    asm volatile("v_cvt_f32_f64 %0, %1" : "=v"(fz) : "v"(dz));
                    const float r2f = fmaf(fz, fz, fmaf(fy, fy, fx * fx));
                    const bool near = !(r2f > L.rnbf);
                    const unsigned long long nb = __ballot(near);
                    if (nb) jl_near<FULL, SELF>(L, near, nb, r, count, jt, rf);
                }
            } else {
                const bool near = !(r2 > L.rnbd);
                const unsigned long long nb = __ballot(near);
                if (nb) jl_near<FULL, SELF>(L, near, nb, r, count, jt, rf);
            }
            double rinv = (double)rf;
            if constexpr (TM >= 2) {
                const double h = fma(-(r2 * rinv), rinv, 1.0);
                rinv = fma(0.5 * rinv, h, rinv);
            }
            const double rinv2 = rinv * rinv;
            const double mr3 = mj * rinv * rinv2;
            const double rv = fma(dz, dvz, fma(dy, dvy, dx * dvx));
            const double al = -3.0 * rv * rinv2;
            L.acc[0] = fma(mr3, dx, L.acc[0]);
            L.acc[1] = fma(mr3, dy, L.acc[1]);
            L.acc[2] = fma(mr3, dz, L.acc[2]);
            L.acc[3] = fma(mr3, fma(al, dx, dvx), L.acc[3]);
            L.acc[4] = fma(mr3, fma(al, dy, dvy), L.acc[4]);
            L.acc[5] = fma(mr3, fma(al, dz, dvz), L.acc[5]);
        }
    }
}

// The common tile of term mode 0 (all TILE records valid, no lane's own
// target in the tile).  G consecutive steps share one neighbour test: the
// minimum of their r^2 (v_min3 + v_min) against r_nb^2, one compare and one
// branch instead of G of each.  r^2 is a sum of squares of finite
// differences, so the minimum is below r_nb^2 exactly when one of the G
// tests is.  Inside the rare branch each step is tested and listed in step
// order, so lists stay ascending.  Every pair is evaluated with the same
// instructions and summed in the same order as in jl_tile, so the results
// are bitwise those of jl_tile.  Between the test and the terms a step keeps
// only its FP32 differences (x velocity's taken with the positions) and
// rsqrt; the rare branch recomputes r^2 from the same differences with the
// same expression, hence the same bits.  That keeps the group's state at 5
// registers per step.
template <int JL, int G, int TILE>
__device__ __forceinline__ void jl_tile_common(JLLane &L, const double2 *__restrict__ pos,
                                               const float *__restrict__ mass, int jt)
{
    constexpr int steps = TILE / JL;
    static_assert(steps % G == 0, "the group divides the tile");
#pragma unroll 1
    for (int st = 0; st < steps; st += G) {
        float dx[G], dy[G], dz[G], rinv[G], dvxg[G];
        float mn = 0.0f;
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int r = (st + g) * JL + L.sub;
            const double2 p0 = pos[3 * r + 0];
            const double2 p1 = pos[3 * r + 1];
            dx[g] = (float)(p0.x - L.xi);
            dy[g] = (float)(p0.y - L.yi);
            dz[g] = (float)(p1.x - L.zi);
            dvxg[g] = (float)(p1.y - L.vxi);
            const float r2 = fmaf(dz[g], dz[g], fmaf(dy[g], dy[g], dx[g] * dx[g]));
            rinv[g] = __builtin_amdgcn_rsqf(r2);
            mn = g == 0 ? r2 : fminf(mn, r2);
        }
        if (__any(!(mn > L.rnbf))) {
#pragma unroll
            for (int g = 0; g < G; ++g) {
                const float r2g = fmaf(dz[g], dz[g], fmaf(dy[g], dy[g], dx[g] * dx[g]));
                const bool near = !(r2g > L.rnbf);
                const unsigned long long nb = __ballot(near);
                if (nb) {
                    // The lane offset passes through an empty asm so that the
                    // list index is formed here, in the rare branch, rather
                    // than kept in registers for every step of the tile.
                    int sub = L.sub;
                    asm volatile("" : "+v"(sub));
                    if (near) asm volatile("v_mov_b32 %0, 0" : "=v"(rinv[g]));
                    const unsigned long long gb = nb & L.gmask;
                    if (near) {
                        const int p = L.cnt + __popcll(gb & L.below);
                        if (p < L.cap) L.lst[p] = jt + (st + g) * JL + sub;
                    }
                    L.cnt += __popcll(gb);
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) {
            const int r = (st + g) * JL + L.sub;
            const double2 p2 = pos[3 * r + 2];
            const float mj = mass[r];
            const float dvx = dvxg[g];
            const float dvy = (float)(p2.x - L.vyi);
            const float dvz = (float)(p2.y - L.vzi);
            const float rinv2 = rinv[g] * rinv[g];
            const float mr3 = mj * rinv[g] * rinv2;
            const float rv = fmaf(dz[g], dvz, fmaf(dy[g], dvy, dx[g] * dvx));
            const float al = -3.0f * rv * rinv2;
            L.pa[0] = fmaf(mr3, dx[g], L.pa[0]);
            L.pa[1] = fmaf(mr3, dy[g], L.pa[1]);
            L.pa[2] = fmaf(mr3, dz[g], L.pa[2]);
            L.pa[3] = fmaf(mr3, fmaf(al, dx[g], dvx), L.pa[3]);
            L.pa[4] = fmaf(mr3, fmaf(al, dy[g], dvy), L.pa[4]);
            L.pa[5] = fmaf(mr3, fmaf(al, dz[g], dvz), L.pa[5]);
        }
    }
}

// EST (mode 0 only): also sum, over the lane's FP32 runs, the squared norms
// of each run's acceleration and jdot partials, and write the target's two
// totals to out.spread.  The rounding error of each sum scales with the
// root of its total (the largest terms dominate both), which the checked
// precision compares with |a| and |jdot|.  The pair arithmetic and the sums
// are those of EST = false.
// Shape: WPB wave64s per workgroup (WPB * 64 / JL targets) stream the
// j-range through double-buffered LDS tiles of TILE records; the first TILE
// threads load one record each.  A full tile is TILE / JL steps, a divisor
// of kFold, so the FP32 runs fold at the same j for every TILE (spans are
// multiples of kTile): the shape changes neither a sum nor a list.
// One-shot tiles (TILE > 64 * WPB, small blocks): the whole j-range of the
// split is one tile (the host guarantees jspan <= TILE), loaded at once
// with TILE / (64 WPB) records per thread into a single LDS buffer, so a
// split waits for one load instead of one per 256 j.
template <int JL, int TM, bool EST = false, int TILE = kTile, int WPB = 4>
__device__ __forceinline__ void regf_body_jl(const JState &js, const int *__restrict__ tgt, int nt,
                                             int jspan, const QueryOut &out, int cap,
                                             int final_counts, int bx, int s, int q_base,
                                             int nt_stride, double2 (*s_pos)[TILE * 3],
                                             typename TermMass<TM>::type (*s_m)[TILE])
{
    constexpr bool ONE_SHOT = TILE > 64 * WPB;
    constexpr int R = ONE_SHOT ? TILE / (64 * WPB) : 1;   // records per thread
    static_assert(!ONE_SHOT || TILE % (64 * WPB) == 0, "whole records per thread");
    static_assert(TILE % JL == 0 && kFold % (TILE / JL) == 0, "runs fold at tile ends");
    static_assert(ONE_SHOT ? TILE % kTile == 0 : kTile % TILE == 0,
                  "split spans are whole tiles, or one one-shot tile");
    using MT = typename TermMass<TM>::type;
    constexpr int TW = 64 / JL;            // targets per wave
    constexpr int G = TILE / JL < 4 ? TILE / JL : 4;   // steps per neighbour test
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int g = lane / JL;
    const int j0 = s * jspan;
    const int j1 = min(js.n, j0 + jspan);
    const int q = bx * (WPB * TW) + (tid >> 6) * TW + g;
    const bool loader = TILE == 64 * WPB || tid < TILE;

    JLLane L;
    L.sub = lane % JL;
    L.gmask = (JL == 64) ? ~0ull : (((1ull << JL) - 1ull) << (g * JL));
    L.below = (1ull << lane) - 1ull;   // lanes < this lane
    L.cap = cap;
    L.cnt = 0;
    L.xi = L.yi = L.zi = L.vxi = L.vyi = L.vzi = 0.0;
    L.rnbf = -1.0f;   // inactive target: every j regular, nothing listed
    L.rnbd = -1.0;
    L.idx = -1;
    L.lst = nullptr;
    if (q < nt) {
        L.idx = tgt[q];
        L.xi = js.x[L.idx]; L.yi = js.y[L.idx]; L.zi = js.z[L.idx];
        L.vxi = js.vx[L.idx]; L.vyi = js.vy[L.idx]; L.vzi = js.vz[L.idx];
        L.rnbd = js.rnb2[L.idx];
        L.rnbf = (float)L.rnbd;
        // mode 3 compares the FP64 r^2 with this bound first (see jl_tile)
        if constexpr (TM == 3) L.rnbd *= 1.0 + 1.0e-6;
        L.lst = out.list + ((size_t)s * nt_stride + (q - q_base)) * (size_t)cap;
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
        L.acc[c] = 0.0;
        L.pa[c] = 0.0f;
    }
    int since_fold = 0;
    float spa = 0.0f, spj = 0.0f;   // EST: sums of squared a / jdot run partials
    auto run_spread = [&]() {
        if constexpr (EST) {
            spa = fmaf(L.pa[2], L.pa[2], fmaf(L.pa[1], L.pa[1], fmaf(L.pa[0], L.pa[0], spa)));
            spj = fmaf(L.pa[5], L.pa[5], fmaf(L.pa[4], L.pa[4], fmaf(L.pa[3], L.pa[3], spj)));
        }
    };

    const int ntile = (j1 - j0 + TILE - 1) / TILE;
    // a wave without targets loads its share of every tile and meets every
    // barrier, but evaluates nothing
    const bool wave_busy = __any(q < nt);
    TileLoader<MT> ld;
    if constexpr (ONE_SHOT) {
        // every load in flight before the first store (ntile <= 1)
        TileLoader<MT> lr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) lr[r].load(js, j0 + r * 64 * WPB + tid, j1);
#pragma unroll
        for (int r = 0; r < R; ++r) lr[r].store(s_pos[0], s_m[0], r * 64 * WPB + tid);
    } else if (ntile > 0 && loader) {
        ld.load(js, j0 + tid, j1);
        ld.store(s_pos[0], s_m[0], tid);
    }
    __syncthreads();
    for (int t = 0; t < ntile; ++t) {
        const int b = ONE_SHOT ? 0 : (t & 1);
        if (!ONE_SHOT && t + 1 < ntile && loader) ld.load(js, j0 + (t + 1) * TILE + tid, j1);
        const int jt = j0 + t * TILE;
        const int count = min(TILE, j1 - jt);
        if (!wave_busy) {
            // no target in this wave (the tail of a small block)
        } else if (__any((unsigned)(L.idx - jt) < (unsigned)count))
            jl_tile<JL, TM, false, true, TILE>(L, s_pos[b], s_m[b], count, jt);
        else if (count == TILE) {
            if constexpr (TM == 0)
                jl_tile_common<JL, G, TILE>(L, s_pos[b], s_m[b], jt);
            else
                jl_tile<JL, TM, true, false, TILE>(L, s_pos[b], s_m[b], count, jt);
        }
        else
            jl_tile<JL, TM, false, false, TILE>(L, s_pos[b], s_m[b], count, jt);
        if constexpr (TM == 0) {
            // fold the FP32 partials once a lane has summed >= kFold pairs
            // (every tile for JL = 4, every 4 / 16 tiles for JL = 16 / 64)
            since_fold += (count + JL - 1) / JL;
            if (since_fold >= kFold) {
                run_spread();
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    L.acc[c] += (double)L.pa[c];
                    L.pa[c] = 0.0f;
                }
                since_fold = 0;
            }
        }
        if (!ONE_SHOT && t + 1 < ntile && loader) ld.store(s_pos[b ^ 1], s_m[b ^ 1], tid);
        __syncthreads();
    }
    double acc[6];
#pragma unroll
    for (int c = 0; c < 6; ++c)
        acc[c] = TM == 0 ? L.acc[c] + (double)L.pa[c] : L.acc[c];
#pragma unroll
    for (int d = JL / 2; d >= 1; d >>= 1)
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[c] += __shfl_xor(acc[c], d, 64);
    if constexpr (EST) {
        run_spread();
#pragma unroll
        for (int d = JL / 2; d >= 1; d >>= 1) {
            spa += __shfl_xor(spa, d, 64);
            spj += __shfl_xor(spj, d, 64);
        }
        if (q < nt && L.sub == 0) {
            const size_t o = (size_t)s * nt_stride + (q - q_base);
            out.spread[2 * o + 0] = spa;
            out.spread[2 * o + 1] = spj;
        }
    }
    if (q < nt && L.sub == 0) {
        const size_t o = (size_t)s * nt_stride + (q - q_base);
        out.acc[3 * o + 0] = acc[0];
        out.acc[3 * o + 1] = acc[1];
        out.acc[3 * o + 2] = acc[2];
        out.adot[3 * o + 0] = acc[3];
        out.adot[3 * o + 1] = acc[4];
        out.adot[3 * o + 2] = acc[5];
        out.cnt[o] = (final_counts && L.cnt > cap) ? -L.cnt : L.cnt;
    }
}

// XCD-aware placement of a launch's gx * S (target block, split) pairs on
// its 1-D grid.  The hardware deals workgroup ids round-robin to the 8 XCDs,
// each with its own L2: id runs on XCD id % 8 as that XCD's (id / 8)-th
// workgroup.  XCD x takes the contiguous run [start(x), start(x) + n(x)) of
// the split-major order L = s * gx + b (n(x) = total / 8, plus one for the
// first total % 8 XCDs), so the workgroups one XCD runs at a time share one
// or two splits' j-ranges, which its L2 keeps, instead of every XCD
// streaming every split.  A bijection of [0, total): every pair runs once,
// with the same arithmetic as on a 2-D grid.
__device__ __forceinline__ void xcd_place(int id, int total, int gx, int &bx, int &s)
{
    const int x = id & 7, k = id >> 3;
    const int q = total >> 3, r = total & 7;
    const int L = x * q + min(x, r) + k;
    s = L / gx;
    bx = L - s * gx;
}

// 5 waves per SIMD (<= 96 VGPRs); the FP64-term modes would otherwise take
// 98 registers and drop to 4.
// reset (checked precision, may be null): the flagged-target counter, zeroed
// here so that no separate launch is needed before the flagging after it.
// A one-shot tile's 52 KB of LDS (56 KB with FP64 masses) holds 3 (2)
// workgroups per CU, so it asks for 3 (2).
// Grid: gx * splits workgroups on x (xcd_place).
template <int JL, int TM, bool EST, int TILE = kTile, int WPB = 4>
__global__ __launch_bounds__(64 * WPB)
__attribute__((amdgpu_waves_per_eu(TILE > 64 * WPB ? (TM == 0 ? 3 : 2) : 5, 8)))
void regf_kernel_jl(JState js, const int *__restrict__ tgt, int nt, int jspan, QueryOut out,
                    int cap, int final_counts, int *__restrict__ reset, int gx, int splits)
{
    if (reset && blockIdx.x == 0 && threadIdx.x == 0) *reset = 0;
    // a one-shot tile (TILE > 64 WPB) has a single buffer
    constexpr int NB = TILE > 64 * WPB ? 1 : 2;
    __shared__ double2 s_pos[NB][TILE * 3];
    __shared__ typename TermMass<TM>::type s_m[NB][TILE];
    int bx, s;
    xcd_place(blockIdx.x, gx * splits, gx, bx, s);
    regf_body_jl<JL, TM, EST, TILE, WPB>(js, tgt, nt, jspan, out, cap, final_counts, bx, s, 0,
                                         nt, s_pos, s_m);
}

// One wave per target.  Partial sums: lane l adds splits l, l+64, ... in
// order, then a fixed xor-butterfly combines the 64 lane sums, so the total
// is bitwise reproducible for a given S.  Neighbour sub-lists: the split
// counts are loaded lane-parallel and prefix-summed across the wave; every
// output entry e then finds its split by binary search over the prefix and
// is copied independently.  Splits are contiguous ascending j-ranges, so the
// concatenation in split order is ascending in j.
// Called by every thread of the workgroup (it contains a barrier); waves with
// active == false only take part in the barrier.  `off` is this wave's LDS
// prefix array of S + 1 ints.
// Partials of this target are read at s*nt_stride + q_rel; the final
// outputs are written at q (absolute target position).
// The S split partials of one target summed by one wave: lane l adds splits
// l, l+64, ... in order, then a fixed xor-butterfly (every lane ends with
// the total).
__device__ __forceinline__ void sum_splits(const QueryOut &part, int S, int nt_stride, int q_rel,
                                           bool active, double (&a)[6])
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int c = 0; c < 6; ++c) a[c] = 0.0;
    if (active) {
        for (int s = lane; s < S; s += 64) {
            const size_t o = (size_t)s * nt_stride + q_rel;
            a[0] += part.acc[3 * o + 0];
            a[1] += part.acc[3 * o + 1];
            a[2] += part.acc[3 * o + 2];
            a[3] += part.adot[3 * o + 0];
            a[4] += part.adot[3 * o + 1];
            a[5] += part.adot[3 * o + 2];
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
#pragma unroll
        for (int c = 0; c < 6; ++c) a[c] += __shfl_xor(a[c], d, 64);
}

// The checked precision's rule: a target is recomputed when either spread,
// scaled by k2 = (ABYSS_CHECK_ERR_SCALE / tol)^2, exceeds the squared norm
// of its sum.  One expression (explicit FMAs) for every kernel that applies
// it, so they flag the same targets.
__device__ __forceinline__ bool check_rule(float sa, float sj, double ax, double ay, double az,
                                           double jx, double jy, double jz, double k2)
{
    return (double)sa * k2 > fma(az, az, fma(ay, ay, ax * ax)) ||
           (double)sj * k2 > fma(jz, jz, fma(jy, jy, jx * jx));
}

__device__ __forceinline__ void finalize_body(const QueryOut &part, int S, int nt_stride, int cap,
                                              int q_rel, int q, bool active, int *off,
                                              double *__restrict__ acc,
                                              double *__restrict__ adot,
                                              int *__restrict__ nnb, int *__restrict__ list,
                                              const FinalMirror &mir = FinalMirror{},
                                              float *__restrict__ spread = nullptr,
                                              const CheckFlag &flag = CheckFlag{})
{
    const int lane = threadIdx.x & 63;

    double a[6];
    sum_splits(part, S, nt_stride, q_rel, active, a);
    if (spread) {
        float ea = 0.0f, ej = 0.0f;
        if (active)
            for (int s = lane; s < S; s += 64) {
                const size_t o = (size_t)s * nt_stride + q_rel;
                ea += part.spread[2 * o + 0];
                ej += part.spread[2 * o + 1];
            }
#pragma unroll
        for (int d = 32; d >= 1; d >>= 1) {
            ea += __shfl_xor(ea, d, 64);
            ej += __shfl_xor(ej, d, 64);
        }
        if (active && lane == 0) {
            spread[2 * (size_t)q + 0] = ea;
            spread[2 * (size_t)q + 1] = ej;
            // checked precision: flag the target here instead of in a
            // separate pass (the rule of check_flag_kernel)
            if (flag.counter && check_rule(ea, ej, a[0], a[1], a[2], a[3], a[4], a[5], flag.k2)) {
                const int k = atomicAdd(flag.counter, 1);
                flag.fq[k] = q;
                flag.fi[k] = flag.tgt[q];
            }
        }
    }

    int running = 0;
    for (int base = 0; base < S; base += 64) {
        const int s = base + lane;
        const int c = (active && s < S) ? part.cnt[(size_t)s * nt_stride + q_rel] : 0;
        int incl = c;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) incl += t;
        }
        if (s < S) off[s] = running + incl - c;
        running += __shfl(incl, 63, 64);
    }
    if (lane == 0) off[S] = running;
    __syncthreads();

    if (!active) return;
    const bool mirror = mir.acc != nullptr;
    if (lane < 3) {
        acc[(size_t)q * 3 + lane] = a[lane];
        if (mirror) mir.acc[(size_t)q * 3 + lane] = a[lane];
    } else if (lane < 6) {
        adot[(size_t)q * 3 + lane - 3] = a[lane];
        if (mirror) mir.adot[(size_t)q * 3 + lane - 3] = a[lane];
    }
    const int total = running;
    const int ncopy = min(total, cap);
    int *dst = list + (size_t)q * cap;
    for (int e = lane; e < ncopy; e += 64) {
        // largest s with off[s] <= e (off is non-decreasing, off[0] = 0)
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const int j = part.list[((size_t)lo * nt_stride + q_rel) * (size_t)cap + (e - off[lo])];
        dst[e] = j;
        if (mirror) mir.list[(size_t)q * cap + e] = j;
    }
    if (lane == 0) {
        nnb[q] = total <= cap ? total : -total;
        if (mirror) mir.nnb[q] = total <= cap ? total : -total;
    }
}

__global__ __launch_bounds__(256) void finalize_kernel(QueryOut part, int S, int nt, int cap,
                                                       double *__restrict__ acc,
                                                       double *__restrict__ adot,
                                                       int *__restrict__ nnb,
                                                       int *__restrict__ list, FinalMirror mir,
                                                       float *__restrict__ spread, CheckFlag flag)
{
    extern __shared__ int s_off_all[];   // [4 waves][S + 1]
    const int w = threadIdx.x >> 6;
    const int q = blockIdx.x * 4 + w;
    finalize_body(part, S, nt, cap, q, q, q < nt, s_off_all + w * (S + 1), acc, adot, nnb, list,
                  mir, spread, flag);
}

// finalize_kernel for S <= G j-splits: G lanes per target, 64 / G targets
// per wave.  Lane l of a group adds split l to zero and the butterfly runs
// over the group's G lanes; in finalize_kernel's 64-lane butterfly the other
// lanes hold zeros, which add exactly, so the sums (and spreads) have the
// same bits.  Lists are merged the same way, G entries at a time.
template <int G>
__global__ __launch_bounds__(256) void finalize_group_kernel(QueryOut part, int S, int nt, int cap,
                                                             double *__restrict__ acc,
                                                             double *__restrict__ adot,
                                                             int *__restrict__ nnb,
                                                             int *__restrict__ list,
                                                             FinalMirror mir,
                                                             float *__restrict__ spread,
                                                             CheckFlag flag)
{
    static_assert(G >= 8 && G <= 64 && (G & (G - 1)) == 0, "a group holds the 6 sums");
    __shared__ int s_off[256 / G][G + 1];
    const int sub = threadIdx.x & (G - 1);
    const int slot = threadIdx.x / G;
    const int q = blockIdx.x * (256 / G) + slot;
    const bool active = q < nt;
    const bool mine = active && sub < S;
    const size_t o = (size_t)sub * nt + q;
    double a[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) a[c] = 0.0;
    if (mine) {
        a[0] += part.acc[3 * o + 0];
        a[1] += part.acc[3 * o + 1];
        a[2] += part.acc[3 * o + 2];
        a[3] += part.adot[3 * o + 0];
        a[4] += part.adot[3 * o + 1];
        a[5] += part.adot[3 * o + 2];
    }
#pragma unroll
    for (int d = G / 2; d >= 1; d >>= 1)
#pragma unroll
        for (int c = 0; c < 6; ++c) a[c] += __shfl_xor(a[c], d, 64);
    if (spread) {
        float ea = 0.0f, ej = 0.0f;
        if (mine) {
            ea += part.spread[2 * o + 0];
            ej += part.spread[2 * o + 1];
        }
#pragma unroll
        for (int d = G / 2; d >= 1; d >>= 1) {
            ea += __shfl_xor(ea, d, 64);
            ej += __shfl_xor(ej, d, 64);
        }
        if (active && sub == 0) {
            spread[2 * (size_t)q + 0] = ea;
            spread[2 * (size_t)q + 1] = ej;
            if (flag.counter && check_rule(ea, ej, a[0], a[1], a[2], a[3], a[4], a[5], flag.k2)) {
                const int k = atomicAdd(flag.counter, 1);
                flag.fq[k] = q;
                flag.fi[k] = flag.tgt[q];
            }
        }
    }

    // split offsets of this target's list entries: a prefix over its group
    const int c0 = mine ? part.cnt[o] : 0;
    int incl = c0;
#pragma unroll
    for (int d = 1; d < G; d <<= 1) {
        const int t = __shfl_up(incl, d, G);
        if (sub >= d) incl += t;
    }
    const int total = __shfl(incl, G - 1, G);
    if (sub < S) s_off[slot][sub] = incl - c0;
    if (sub == 0) s_off[slot][S] = total;
    __syncthreads();

    if (!active) return;
    const bool mirror = mir.acc != nullptr;
    if (sub < 3) {
        acc[(size_t)q * 3 + sub] = a[sub];
        if (mirror) mir.acc[(size_t)q * 3 + sub] = a[sub];
    } else if (sub < 6) {
        adot[(size_t)q * 3 + sub - 3] = a[sub];
        if (mirror) mir.adot[(size_t)q * 3 + sub - 3] = a[sub];
    }
    const int *off = s_off[slot];
    const int ncopy = min(total, cap);
    int *dst = list + (size_t)q * cap;
    for (int e = sub; e < ncopy; e += G) {
        // largest s with off[s] <= e (off is non-decreasing, off[0] = 0)
        int lo = 0, hi = S - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= e) lo = mid; else hi = mid - 1;
        }
        const int j = part.list[((size_t)lo * nt + q) * (size_t)cap + (e - off[lo])];
        dst[e] = j;
        if (mirror) mir.list[(size_t)q * cap + e] = j;
    }
    if (sub == 0) {
        nnb[q] = total <= cap ? total : -total;
        if (mirror) mir.nnb[q] = total <= cap ? total : -total;
    }
}

// ---- checked precision ---------------------------------------------------

// One thread per target.  A wave appends its flagged targets with one
// atomic add (of its flag count) and their ranks within the wave.
__global__ __launch_bounds__(256) void check_flag_kernel(const double *__restrict__ acc,
                                                         const double *__restrict__ adot,
                                                         const float *__restrict__ spread,
                                                         const int *__restrict__ tgt, int nt,
                                                         double k2, int *__restrict__ counter,
                                                         int *__restrict__ fq,
                                                         int *__restrict__ fi)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool flag = false;
    if (q < nt) {
        const size_t o = 3 * (size_t)q;
        flag = check_rule(spread[2 * (size_t)q + 0], spread[2 * (size_t)q + 1], acc[o],
                          acc[o + 1], acc[o + 2], adot[o], adot[o + 1], adot[o + 2], k2);
    }
    const unsigned long long b = __ballot(flag);
    if (b == 0) return;   // the whole wave
    const int first = __ffsll((long long)b) - 1;
    int base = 0;
    if (lane == first) base = atomicAdd(counter, __popcll(b));
    base = __shfl(base, first, 64);
    if (flag) {
        const int k = base + __popcll(b & ((1ull << lane) - 1ull));
        fq[k] = q;
        fi[k] = tgt[q];
    }
}

// The recomputation with the flagged count read on the device (the
// no-allocation query: no host wait).  The grid covers every target; a
// workgroup whose
// target block starts at or past the count exits before its first barrier.
// Partials are indexed s * nt_stride + k.
template <int JL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void recheck_kernel_dev(JState js, const int *__restrict__ tgt, const int *__restrict__ count,
                        int nt_stride, int jspan, QueryOut out, int gx, int splits)
{
    __shared__ double2 s_pos[2][kTile * 3];
    __shared__ double s_m[2][kTile];
    const int m = *count;
    int bx, s;
    xcd_place(blockIdx.x, gx * splits, gx, bx, s);
    if (bx * (4 * (64 / JL)) >= m) return;   // the whole workgroup
    regf_body_jl<JL, 3>(js, tgt, m, jspan, out, 0, 0, bx, s, 0, nt_stride, s_pos, s_m);
}

// The recomputation with the flagged count read on the device, as a
// grid-stride loop: workgroup b takes items b, b + gridDim.x, ... of the
// (target block, split) items of the first m = *count targets, split-minor.
// The trip count is the same for every thread of a workgroup, so the
// barriers inside each item are met by all of them; the last barrier of an
// item precedes the next item's first LDS store.
template <int JL>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void recheck_kernel_loop(JState js, const int *__restrict__ tgt, const int *__restrict__ count,
                         int nt_stride, int jspan, QueryOut out, int splits)
{
    __shared__ double2 s_pos[2][kTile * 3];
    __shared__ double s_m[2][kTile];
    const int m = min(*count, nt_stride);
    constexpr int TPB = 4 * (64 / JL);   // targets per workgroup
    const int items = (m + TPB - 1) / TPB * splits;
    for (int it = blockIdx.x; it < items; it += gridDim.x)
        regf_body_jl<JL, 3>(js, tgt, m, jspan, out, 0, 0, it / splits, it % splits, 0, nt_stride,
                            s_pos, s_m);
}

// One wave per flagged target k < m (m = *m_dev when m_dev is set; then
// thread 0 also writes m to *m_report, host memory the caller reads after
// the stream has drained).  The mirror, when set, receives the same values.
__global__ __launch_bounds__(256) void recheck_finalize_kernel(QueryOut part, int S, int nt_stride,
                                                               int m_host,
                                                               const int *__restrict__ m_dev,
                                                               int *__restrict__ m_report,
                                                               const int *__restrict__ fq,
                                                               double *__restrict__ acc,
                                                               double *__restrict__ adot,
                                                               FinalMirror mir)
{
    const int lane = threadIdx.x & 63;
    const int m = min(m_dev ? *m_dev : m_host, nt_stride);
    if (m_report && blockIdx.x == 0 && threadIdx.x == 0) *m_report = m;
    const int k = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (k >= m) return;   // the whole wave
    double a[6];
    sum_splits(part, S, nt_stride, k, true, a);
    const size_t q = (size_t)fq[k];
    if (lane < 3) {
        acc[q * 3 + lane] = a[lane];
        if (mir.acc) mir.acc[q * 3 + lane] = a[lane];
    } else if (lane < 6) {
        adot[q * 3 + lane - 3] = a[lane];
        if (mir.adot) mir.adot[q * 3 + lane - 3] = a[lane];
    }
}

// ---- batched ensembles ---------------------------------------------------

__global__ void batch_reset_kernel(int *counters)
{
    if (threadIdx.x < 2) counters[threadIdx.x] = 0;
}

// One work item of a batch with the cluster's lanes per target and j-tile
// in 4-wave workgroups: smaller tiles use the start of the LDS buffers.
template <int JL, bool EST, int TILE>
__device__ __forceinline__ void batch_body(const ClusterDesc &d, const WorkItem &w,
                                           const QueryOut &out, int cap, int fc,
                                           double2 (*s_pos)[kTile * 3], float (*s_m)[kTile])
{
    regf_body_jl<JL, 0, EST, TILE, 4>(d.js, d.tgt, d.nt, d.jspan, out, cap, fc, w.tblock,
                                      w.split, d.q_base, d.nt_stride,
                                      reinterpret_cast<double2 (*)[TILE * 3]>(s_pos),
                                      reinterpret_cast<float (*)[TILE]>(s_m));
}

// Persistent: each workgroup takes work items from a device queue until it
// is empty, so small and large clusters share one launch and the chip stays
// busy.  A cluster's items use the same plan and body as a single-cluster
// query, so its results do not depend on what else is in the batch.
// Every workgroup leaves the loop once the queue index passes nitems.
// Lanes-per-target bodies only (ensemble members never plan one target per
// lane), so the kernel fits 5 waves per SIMD like regf_kernel_jl.
// EST: checked precision (the spread estimate of regf_body_jl).
template <bool EST>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void batch_kernel(const ClusterDesc *__restrict__ cd, const WorkItem *__restrict__ items,
                  const BatchHeader *__restrict__ hdr, int *__restrict__ counter, int cap)
{
    __shared__ double2 s_pos[2][kTile * 3];
    __shared__ float s_m[2][kTile];
    __shared__ int s_item;
    const int nitems = hdr->nitems;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const ClusterDesc &d = cd[w.cluster];
        const int fc = d.splits == 1 ? 1 : 0;
        const QueryOut out =
            fc ? QueryOut{d.acc, d.adot, d.nnb, d.list, d.spread} : d.part;
        switch (d.jl * 1024 + d.tile) {
            case 4 * 1024 + 64: batch_body<4, EST, 64>(d, w, out, cap, fc, s_pos, s_m); break;
            case 4 * 1024 + 128: batch_body<4, EST, 128>(d, w, out, cap, fc, s_pos, s_m); break;
            case 4 * 1024 + 256: batch_body<4, EST, 256>(d, w, out, cap, fc, s_pos, s_m); break;
            case 16 * 1024 + 64: batch_body<16, EST, 64>(d, w, out, cap, fc, s_pos, s_m); break;
            case 16 * 1024 + 128: batch_body<16, EST, 128>(d, w, out, cap, fc, s_pos, s_m); break;
            case 16 * 1024 + 256: batch_body<16, EST, 256>(d, w, out, cap, fc, s_pos, s_m); break;
            case 64 * 1024 + 64: batch_body<64, EST, 64>(d, w, out, cap, fc, s_pos, s_m); break;
            case 64 * 1024 + 128: batch_body<64, EST, 128>(d, w, out, cap, fc, s_pos, s_m); break;
            default: batch_body<64, EST, 256>(d, w, out, cap, fc, s_pos, s_m); break;
        }
    }
}

// Checked precision of a batch: the recomputation of every cluster's
// flagged targets in one persistent launch.  A descriptor's tgt holds the
// flagged j-indices and nt their count; results always go to the split
// partials d.part (stride d.nt), merged by batch_recheck_finalize_kernel.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5, 8)))
void batch_recheck_kernel(const ClusterDesc *__restrict__ cd, const WorkItem *__restrict__ items,
                          const BatchHeader *__restrict__ hdr, int *__restrict__ counter)
{
    __shared__ double2 s_pos[2][kTile * 3];
    __shared__ double s_m[2][kTile];
    __shared__ int s_item;
    const int nitems = hdr->nitems;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
        __syncthreads();
        const int it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const WorkItem w = items[it];
        const ClusterDesc &d = cd[w.cluster];
        switch (d.jl) {
            case 4:
                regf_body_jl<4, 3>(d.js, d.tgt, d.nt, d.jspan, d.part, 0, 0, w.tblock, w.split, 0,
                                   d.nt, s_pos, s_m);
                break;
            case 16:
                regf_body_jl<16, 3>(d.js, d.tgt, d.nt, d.jspan, d.part, 0, 0, w.tblock, w.split,
                                    0, d.nt, s_pos, s_m);
                break;
            default:
                regf_body_jl<64, 3>(d.js, d.tgt, d.nt, d.jspan, d.part, 0, 0, w.tblock, w.split,
                                    0, d.nt, s_pos, s_m);
                break;
        }
    }
}

// One wave per flagged target, persistent over their flat index (descriptor
// k by binary search over off, off[ncd] = total): the S partials summed in
// split order and written over the cluster's results at fq.
__global__ __launch_bounds__(256) void batch_recheck_finalize_kernel(
    const ClusterDesc *__restrict__ cd, const int *__restrict__ off, int ncd, int total)
{
    const int lane = threadIdx.x & 63;
    for (int base = blockIdx.x * 4; base < total; base += gridDim.x * 4) {
        const int g = base + (threadIdx.x >> 6);
        if (g >= total) continue;   // the whole wave; no barrier below
        int lo = 0, hi = ncd - 1;   // largest k with off[k] <= g
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (off[mid] <= g) lo = mid; else hi = mid - 1;
        }
        const ClusterDesc &d = cd[lo];
        const int k = g - off[lo];
        double a[6];
        sum_splits(d.part, d.splits, d.nt, k, true, a);
        const size_t q = (size_t)d.fq[k];
        if (lane < 3)
            d.acc[q * 3 + lane] = a[lane];
        else if (lane < 6)
            d.adot[q * 3 + lane - 3] = a[lane];
    }
}

// Persistent over the flat index of (split slice, target); the loop bound is
// the same for every wave of a workgroup, so the body's barrier is uniform.
__global__ __launch_bounds__(256) void batch_finalize_kernel(const ClusterDesc *__restrict__ cd,
                                                             const int *__restrict__ fin_cl,
                                                             const int *__restrict__ fin_off,
                                                             const BatchHeader *__restrict__ hdr,
                                                             int cap)
{
    __shared__ int s_off_all[4 * (kMaxSplitsBatch + 1)];
    const int w = threadIdx.x >> 6;
    const int total = hdr->nfin;
    const int nfc = hdr->nfc;
    int *off = s_off_all + w * (kMaxSplitsBatch + 1);
    for (int base = blockIdx.x * 4; base < total; base += gridDim.x * 4) {
        const int g = base + w;
        const bool active = g < total;
        int k = 0;
        if (active) {
            int lo = 0, hi = nfc - 1;   // largest k with fin_off[k] <= g
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (fin_off[mid] <= g) lo = mid; else hi = mid - 1;
            }
            k = lo;
        }
        const ClusterDesc &d = cd[fin_cl[k]];
        const int q_rel = active ? g - fin_off[k] : 0;
        finalize_body(d.part, active ? d.splits : 1, d.nt_stride, cap, q_rel, d.q_base + q_rel,
                      active, off, d.acc, d.adot, d.nnb, d.list, FinalMirror{},
                      d.part.spread ? d.spread : nullptr);
    }
}

// Flat over the checked targets of every descriptor (descriptor k by binary
// search over ck_off); one thread per target.  Lanes of a wave may belong
// to different clusters, so each flagged lane appends with its own atomic.
__global__ __launch_bounds__(256) void batch_check_flag_kernel(const ClusterDesc *__restrict__ cd,
                                                               const int *__restrict__ ck_off,
                                                               int ncd, int total, double k2)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total) return;
    int lo = 0, hi = ncd - 1;   // largest k with ck_off[k] <= g
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ck_off[mid] <= g) lo = mid; else hi = mid - 1;
    }
    const ClusterDesc &d = cd[lo];
    const size_t q = (size_t)d.chk_q0 + (g - ck_off[lo]);
    if (check_rule(d.spread[2 * q + 0], d.spread[2 * q + 1], d.acc[3 * q], d.acc[3 * q + 1],
                   d.acc[3 * q + 2], d.adot[3 * q], d.adot[3 * q + 1], d.adot[3 * q + 2], k2)) {
        const int k = atomicAdd(d.count, 1);
        d.fq[k] = (int)q;
        d.fi[k] = d.tgt[q];
    }
}

// Zeroes the flagged count of every descriptor that starts a pass over its
// cluster's target blocks; later batches of the pass keep appending to it.
__global__ void batch_count_reset_kernel(const ClusterDesc *__restrict__ cd, int ncd)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a < ncd && cd[a].chk_reset) *cd[a].count = 0;
}

__global__ void zero_ints_kernel(int *__restrict__ p, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

template <int IPL, int U, bool F64>
hipError_t launch_t(const JState &js, const int *tgt, int nt, const Plan &p, const QueryOut &out,
                    int cap, hipStream_t st)
{
    dim3 grid(p.grid_x, p.splits);
    hipLaunchKernelGGL((regf_kernel<IPL, U, F64>), grid, dim3(kBlock), 0, st, js, tgt, nt,
                       p.jspan, out, cap, p.splits == 1 ? 1 : 0);
    return hipGetLastError();
}

// Term mode 0 with the plan's shape (j-tile, wave64s per workgroup).
template <int JL, bool EST, int TILE, int WPB>
void launch_shape(const JState &js, const int *tgt, int nt, const Plan &p, const QueryOut &out,
                  int cap, int fc, int *reset, hipStream_t st)
{
    hipLaunchKernelGGL((regf_kernel_jl<JL, 0, EST, TILE, WPB>), dim3(p.grid_x * p.splits),
                       dim3(64 * WPB), 0, st, js, tgt, nt, p.jspan, out, cap, fc, reset, p.grid_x,
                       p.splits);
}

template <int JL, bool EST>
hipError_t launch_mode0(const JState &js, const int *tgt, int nt, const Plan &p,
                        const QueryOut &out, int cap, int fc, int *reset, hipStream_t st)
{
    switch (p.tile * 8 + p.wpb) {
        case 256 * 8 + 4: launch_shape<JL, EST, 256, 4>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 128 * 8 + 4: launch_shape<JL, EST, 128, 4>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 64 * 8 + 4: launch_shape<JL, EST, 64, 4>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 128 * 8 + 2: launch_shape<JL, EST, 128, 2>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 64 * 8 + 2: launch_shape<JL, EST, 64, 2>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 64 * 8 + 1: launch_shape<JL, EST, 64, 1>(js, tgt, nt, p, out, cap, fc, reset, st); break;
        case 1024 * 8 + 4:
            // one-shot tiles: every split within one tile (fold at 16 or 64 steps)
            if constexpr (JL >= 16) {
                if (p.jspan > 1024) return hipErrorInvalidValue;
                launch_shape<JL, EST, 1024, 4>(js, tgt, nt, p, out, cap, fc, reset, st);
                break;
            }
            return hipErrorInvalidValue;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// tm: term mode; est: the spread estimate of checked precision (mode 0)
template <int JL>
hipError_t launch_jl(const JState &js, const int *tgt, int nt, const Plan &p, const QueryOut &out,
                     int cap, int tm, bool est, int *reset, hipStream_t st)
{
    // one workgroup per (target block, split) on a 1-D grid (xcd_place)
    if ((long long)p.grid_x * p.splits > 0x7fffffffLL) return hipErrorInvalidValue;
    dim3 grid(p.grid_x * p.splits);
    const int fc = p.splits == 1 ? 1 : 0;
    if (tm == 3 && p.tile == kTileOneShot && p.wpb == 4) {
        // a checked block in FP64 terms in one pass, every split one tile
        if constexpr (JL >= 16) {
            if (p.jspan > kTileOneShot) return hipErrorInvalidValue;
            hipLaunchKernelGGL((regf_kernel_jl<JL, 3, false, kTileOneShot, 4>), grid,
                               dim3(kBlock), 0, st, js, tgt, nt, p.jspan, out, cap, fc, reset,
                               p.grid_x, p.splits);
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    if (tm != 0 && (p.tile != kTile || p.wpb != 4)) return hipErrorInvalidValue;
    if (tm == 0 && est)
        return launch_mode0<JL, true>(js, tgt, nt, p, out, cap, fc, reset, st);
    else if (tm == 0)
        return launch_mode0<JL, false>(js, tgt, nt, p, out, cap, fc, reset, st);
    else if (tm == 1)
        hipLaunchKernelGGL((regf_kernel_jl<JL, 1, false>), grid, dim3(kBlock), 0, st, js, tgt,
                           nt, p.jspan, out, cap, fc, reset, p.grid_x, p.splits);
    else if (tm == 2)
        hipLaunchKernelGGL((regf_kernel_jl<JL, 2, false>), grid, dim3(kBlock), 0, st, js, tgt,
                           nt, p.jspan, out, cap, fc, reset, p.grid_x, p.splits);
    else
        hipLaunchKernelGGL((regf_kernel_jl<JL, 3, false>), grid, dim3(kBlock), 0, st, js, tgt,
                           nt, p.jspan, out, cap, fc, reset, p.grid_x, p.splits);
    return hipGetLastError();
}

// Term mode of the lanes-per-target kernel for a precision mode.
int term_mode(int prec)
{
    switch (prec) {
        case ABYSS_PREC_MIXED: return 1;
        case ABYSS_PREC_REFINED: return 2;
        case kPrecRecheck: return 3;
        default: return 0;
    }
}

}  // namespace

bool variant_supported(int prec, int ipl, int unroll, int jl)
{
    if (prec == ABYSS_PREC_MIXED || prec == ABYSS_PREC_REFINED || prec == ABYSS_PREC_CHECKED ||
        prec == kPrecRecheck)
        return jl == 4 || jl == 16 || jl == 64;
    if (jl != 1) return prec == ABYSS_PREC_SPLIT && (jl == 4 || jl == 16 || jl == 64);
    if (prec == ABYSS_PREC_FP64) return (ipl == 1 || ipl == 2) && unroll == 2;
    return (ipl == 1 || ipl == 2 || ipl == 4) && (unroll == 2 || unroll == 4);
}

int targets_per_block(const Plan &p)
{
    return p.jl > 1 ? p.wpb * (64 / p.jl) : kBlock * p.ipl;
}

bool shape_supported(int tile, int wpb)
{
    if (tile == kTileOneShot) return wpb == 4;
    return (tile == 64 || tile == 128 || tile == 256) && (wpb == 1 || wpb == 2 || wpb == 4) &&
           tile <= 64 * wpb;
}

hipError_t launch_regf(const JState &js, const int *tgt, int nt, const Plan &p,
                       const QueryOut &out, int cap, int prec, hipStream_t st, int *reset)
{
    if (!variant_supported(prec, p.ipl, p.unroll, p.jl)) return hipErrorInvalidValue;
    const int tm = term_mode(prec);
    const bool est = prec == ABYSS_PREC_CHECKED;
    if (est && !out.spread) return hipErrorInvalidValue;
    if (p.jl > 1 && !shape_supported(p.tile, p.wpb)) return hipErrorInvalidValue;
    if (p.jl == 4) return launch_jl<4>(js, tgt, nt, p, out, cap, tm, est, reset, st);
    if (p.jl == 16) return launch_jl<16>(js, tgt, nt, p, out, cap, tm, est, reset, st);
    if (p.jl == 64) return launch_jl<64>(js, tgt, nt, p, out, cap, tm, est, reset, st);
    if (reset || p.tile != kTile || p.wpb != 4)
        return hipErrorInvalidValue;   // the lanes-per-target kernels only
    if (prec == ABYSS_PREC_FP64) {
        if (p.ipl == 1) return launch_t<1, 2, true>(js, tgt, nt, p, out, cap, st);
        return launch_t<2, 2, true>(js, tgt, nt, p, out, cap, st);
    }
    switch (p.ipl * 10 + p.unroll) {
        case 12: return launch_t<1, 2, false>(js, tgt, nt, p, out, cap, st);
        case 14: return launch_t<1, 4, false>(js, tgt, nt, p, out, cap, st);
        case 22: return launch_t<2, 2, false>(js, tgt, nt, p, out, cap, st);
        case 24: return launch_t<2, 4, false>(js, tgt, nt, p, out, cap, st);
        case 42: return launch_t<4, 2, false>(js, tgt, nt, p, out, cap, st);
        case 44: return launch_t<4, 4, false>(js, tgt, nt, p, out, cap, st);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_finalize(const QueryOut &part, int S, int nt, int cap, double *acc,
                           double *adot, int *nnb, int *list, const FinalMirror &mirror,
                           float *spread, hipStream_t st, const CheckFlag &flag)
{
    if (spread && !part.spread) return hipErrorInvalidValue;
    if (flag.counter && !spread) return hipErrorInvalidValue;
    if (S <= 8) {
        hipLaunchKernelGGL(finalize_group_kernel<8>, dim3((nt + 31) / 32), dim3(256), 0, st, part,
                           S, nt, cap, acc, adot, nnb, list, mirror, spread, flag);
        return hipGetLastError();
    }
    if (S <= 16) {
        hipLaunchKernelGGL(finalize_group_kernel<16>, dim3((nt + 15) / 16), dim3(256), 0, st, part,
                           S, nt, cap, acc, adot, nnb, list, mirror, spread, flag);
        return hipGetLastError();
    }
    dim3 grid((nt + 3) / 4);
    const size_t lds = 4 * (size_t)(S + 1) * sizeof(int);
    hipLaunchKernelGGL(finalize_kernel, grid, dim3(256), lds, st, part, S, nt, cap, acc, adot, nnb,
                       list, mirror, spread, flag);
    return hipGetLastError();
}

hipError_t launch_check_flag(const double *acc, const double *adot, const float *spread,
                             const int *tgt, int nt, double k2, int *counters, int *fq, int *fi,
                             hipStream_t st)
{
    hipLaunchKernelGGL(check_flag_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, acc, adot,
                       spread, tgt, nt, k2, counters, fq, fi);
    return hipGetLastError();
}

hipError_t launch_recheck_finalize(const QueryOut &part, int S, int m, const int *fq,
                                   double *acc, double *adot, hipStream_t st)
{
    hipLaunchKernelGGL(recheck_finalize_kernel, dim3((m + 3) / 4), dim3(256), 0, st, part, S, m,
                       m, (const int *)nullptr, (int *)nullptr, fq, acc, adot, FinalMirror{});
    return hipGetLastError();
}

hipError_t launch_recheck_dev(const JState &js, const int *fi, const int *count, int nt,
                              const Plan &p, const QueryOut &out, const int *fq, double *acc,
                              double *adot, const FinalMirror &mirror, int *m_report,
                              hipStream_t st)
{
    if ((p.jl != 4 && p.jl != 16 && p.jl != 64) || p.tile != kTile || p.wpb != 4)
        return hipErrorInvalidValue;
    if ((long long)p.grid_x * p.splits > 0x7fffffffLL) return hipErrorInvalidValue;
    dim3 grid(p.grid_x * p.splits);
    if (p.jl == 4)
        hipLaunchKernelGGL(recheck_kernel_dev<4>, grid, dim3(kBlock), 0, st, js, fi, count, nt,
                           p.jspan, out, p.grid_x, p.splits);
    else if (p.jl == 16)
        hipLaunchKernelGGL(recheck_kernel_dev<16>, grid, dim3(kBlock), 0, st, js, fi, count, nt,
                           p.jspan, out, p.grid_x, p.splits);
    else
        hipLaunchKernelGGL(recheck_kernel_dev<64>, grid, dim3(kBlock), 0, st, js, fi, count, nt,
                           p.jspan, out, p.grid_x, p.splits);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(recheck_finalize_kernel, dim3((nt + 3) / 4), dim3(256), 0, st, out,
                       p.splits, nt, 0, count, m_report, fq, acc, adot, mirror);
    return hipGetLastError();
}

hipError_t launch_recheck_loop(const JState &js, const int *fi, const int *count, int nt,
                               const Plan &p, int grid, const QueryOut &out, const int *fq,
                               double *acc, double *adot, int *m_report, hipStream_t st)
{
    if ((p.jl != 4 && p.jl != 16 && p.jl != 64) || p.tile != kTile || p.wpb != 4 || grid <= 0 ||
        nt <= 0 || p.splits <= 0 || (long long)nt * p.splits > 0x7fffffffLL)
        return hipErrorInvalidValue;
    if (p.jl == 4)
        hipLaunchKernelGGL(recheck_kernel_loop<4>, dim3(grid), dim3(kBlock), 0, st, js, fi, count,
                           nt, p.jspan, out, p.splits);
    else if (p.jl == 16)
        hipLaunchKernelGGL(recheck_kernel_loop<16>, dim3(grid), dim3(kBlock), 0, st, js, fi,
                           count, nt, p.jspan, out, p.splits);
    else
        hipLaunchKernelGGL(recheck_kernel_loop<64>, dim3(grid), dim3(kBlock), 0, st, js, fi,
                           count, nt, p.jspan, out, p.splits);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(recheck_finalize_kernel, dim3((nt + 3) / 4), dim3(256), 0, st, out,
                       p.splits, nt, 0, count, m_report, fq, acc, adot, FinalMirror{});
    return hipGetLastError();
}

hipError_t launch_batch_reset(int *counters, hipStream_t st)
{
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, st, counters);
    return hipGetLastError();
}

int batch_blocks_per_cu()
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, batch_kernel<false>, kBlock, 0) !=
            hipSuccess ||
        nb < 1)
        nb = 4;
    return nb;
}

hipError_t launch_batch(const ClusterDesc *cd, const WorkItem *items, const BatchHeader *hdr,
                        int *counter, int cap, int grid, bool est, hipStream_t st)
{
    if (est)
        hipLaunchKernelGGL(batch_kernel<true>, dim3(grid), dim3(kBlock), 0, st, cd, items, hdr,
                           counter, cap);
    else
        hipLaunchKernelGGL(batch_kernel<false>, dim3(grid), dim3(kBlock), 0, st, cd, items, hdr,
                           counter, cap);
    return hipGetLastError();
}

hipError_t launch_batch_count_reset(const ClusterDesc *cd, int ncd, hipStream_t st)
{
    if (ncd <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_count_reset_kernel, dim3((ncd + 255) / 256), dim3(256), 0, st, cd,
                       ncd);
    return hipGetLastError();
}

hipError_t launch_batch_check_flag(const ClusterDesc *cd, const int *ck_off, int ncd, int total,
                                   double k2, hipStream_t st)
{
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_check_flag_kernel, dim3((total + 255) / 256), dim3(256), 0, st, cd,
                       ck_off, ncd, total, k2);
    return hipGetLastError();
}

hipError_t launch_batch_recheck(const ClusterDesc *cd, const WorkItem *items,
                                const BatchHeader *hdr, int *counter, const int *off, int ncd,
                                int total, int grid, int fin_grid, hipStream_t st)
{
    hipError_t e = launch_batch_reset(counter, st);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_recheck_kernel, dim3(grid), dim3(kBlock), 0, st, cd, items, hdr,
                       counter);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(batch_recheck_finalize_kernel, dim3(fin_grid), dim3(256), 0, st, cd, off,
                       ncd, total);
    return hipGetLastError();
}

hipError_t launch_zero_ints(int *p, int n, hipStream_t st)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(zero_ints_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_batch_finalize(const ClusterDesc *cd, const int *fin_cl, const int *fin_off,
                                 const BatchHeader *hdr, int cap, int grid, hipStream_t st)
{
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(grid), dim3(256), 0, st, cd, fin_cl, fin_off,
                       hdr, cap);
    return hipGetLastError();
}

}  // namespace abyss
