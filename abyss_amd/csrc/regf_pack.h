// Internal: host-side validation and pack of a j-set upload (regf_pack.cpp).
// Not part of the public ABI.
#pragma once

#include <functional>

namespace abyss {

constexpr int kPackBadState = 1;    // non-finite position or velocity
constexpr int kPackBadMass = 2;     // non-finite or negative mass
constexpr int kPackBadRadius = 3;   // NaN or negative r_nb^2
// Below this many particles the pack runs on the calling thread alone.
constexpr int kPackParallelMin = 16384;

// The lowest-indexed invalid particle (index -1: every particle is valid)
// and what is wrong with it.
struct PackStatus {
    int index = -1;
    int kind = 0;
};

// Validates particles [j0, j1) of n (x, v: [n][3]; m, r2: [n]) and packs
// them into the device layout: soa = 8 arrays of n doubles
// [x|y|z|vx|vy|vz|m|r2], m32 = the masses rounded to FP32.  After a failure
// soa / m32 hold partial contents.
PackStatus pack_jset(int n, int j0, int j1, const double *m, const double *x, const double *v,
                     const double *r2, double *soa, float *m32);

// The same over all n particles, positions first: once every position is
// packed and valid, `positions_done` runs (once, on the calling thread)
// while the other threads pack the rest (velocities, masses, radii), so a
// caller can copy the three position arrays meanwhile.  It does not run
// when a position is invalid.  Returns the first invalid particle as
// pack_jset does.
PackStatus pack_jset_staged(int n, const double *m, const double *x, const double *v,
                            const double *r2, double *soa, float *m32,
                            const std::function<void()> &positions_done);

}  // namespace abyss
