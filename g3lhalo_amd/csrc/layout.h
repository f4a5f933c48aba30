// layout.h -- data layout shared by the host engine and the gfx950 kernels.
//
// Region pool (HBM, structure of arrays): field f of slot s lives at
// geo[f * stride + s], with
//   f = 0..6    centre of the region in each integration dimension
//   f = 7..13   half-width in each dimension
//   f = 14      volume (kept per region: the lower half of a split halves its
//               parent's volume, the upper half recomputes the product, as in
//               src/hcubature.cpp:206-218)
//   f = 15, 16  rule value and error estimate
// plus split[s] (uint8), the dimension chosen for the next cut.
// Each integral owns a contiguous range [base, base + cap) of slots.
#pragma once
#include <stdint.h>

namespace g3l {

constexpr int kMaxDim = 7;
constexpr int kFieldC = 0;
constexpr int kFieldH = 7;
constexpr int kFieldVol = 14;
constexpr int kFieldVal = 15;
constexpr int kFieldErr = 16;
constexpr int kNumFields = 17;

// Integrand kinds: the n_bar integrand and the six statistics terms.
enum Kind : int {
  kNbar = 0,   // 1-D, Gauss-Kronrod 15
  kNN1 = 1,    // NNMap 1-halo, 5-D
  kNN2 = 2,    // NNMap 2-halo, 6-D
  kNN3 = 3,    // NNMap 3-halo, 7-D
  kNM1 = 4,    // NMapMap 1-halo, 5-D
  kNM2 = 5,    // NMapMap 2-halo, 6-D
  kNM3 = 6,    // NMapMap 3-halo, 7-D
  kNumKinds = 7
};

__host__ __device__ constexpr int kind_dim(int k) {
  return k == kNbar ? 1 : (k == kNN1 || k == kNM1) ? 5 : (k == kNN2 || k == kNM2) ? 6 : 7;
}
// Genz-Malik 7/5: 1 + 4d + 2d(d-1) + 2^d points; GK15: 15 (src/hcubature.cpp:416-419).
__host__ __device__ constexpr int rule_points(int d) {
  return d == 1 ? 15 : 1 + 4 * d + 2 * d * (d - 1) + (1 << d);
}

// Job flags.
enum : uint32_t {
  kJobDone = 1u,        // finished (converged, capped or failed)
  kJobConverged = 2u,   // error bound met
  kJobCapped = 4u,      // evaluation budget exhausted
  kJobNonfinite = 8u,   // a NaN/Inf reached the totals
  kJobPoolFull = 16u,   // region capacity exhausted
};

// Static description of one adaptive integral.
struct JobDesc {
  int32_t kind;
  int32_t dim;
  int32_t pts;
  int32_t args;       // index into the TermArgs table
  uint32_t base;      // first pool slot
  uint32_t cap;       // slot capacity
  double reltol;
  uint64_t maxeval;   // 0 = unlimited
};

// Mutable state of one adaptive integral.
struct JobState {
  uint64_t neval;
  uint32_t n;         // live regions
  uint32_t flags;
  uint32_t rounds;
  uint32_t pad;
  double val, err;
};

// Entry of an evaluation list: pool slot and job index.
struct EvalItem {
  uint32_t slot;
  uint32_t job;
};

}  // namespace g3l
