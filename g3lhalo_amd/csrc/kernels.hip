// kernels.hip -- gfx950 kernels of the adaptive cubature engine.
//
//   k_eval<KIND>   fused: generate the rule points of a group of regions,
//                  evaluate the integrand, reduce each region to (value,
//                  error, split dimension) from LDS.  The hot kernel.
//   k_roots        one root region per integral (src/hcubature.cpp:160-172).
//   k_select       per integral and refinement round: totals, convergence,
//                  Gladwell batch selection and region cuts
//                  (src/hcubature.cpp:938-989), appending the children to
//                  the evaluation lists.
//   k_zero         clears the round counters (a kernel rather than a memset
//                  so the round can be captured in a graph).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"
#include "physics.h"
#include "rules.h"
#include "select_math.h"

#pragma clang fp contract(off)

namespace g3l {

// ------------------------------------------------------------ evaluation
template <int KIND, class P>
__device__ __forceinline__ double integrand(const DevTables &T, const TermArgs &a,
                                            const double *x) {
  using Q = typename EvalPolicy<KIND, P>::type;
  if constexpr (KIND == kNbar) {
    return nbar_kernel<Q>(T, a, x);
  } else {
    DirectSrc<KIND, Q> s(T, a, x);
    return integrand_src<KIND, Q>(T, a, s);
  }
}

// Launch variants: a kernel template parameter KIND carries the integrand
// kind in its low four bits and a block-shape variant above them; kind_of_kv
// and variant_of_kv take it apart.  Every variant computes the same values.
constexpr int kVariantShift = 4;
constexpr int kNumVariants = 4;
__host__ __device__ constexpr int kind_of_kv(int kv) { return kv & ((1 << kVariantShift) - 1); }
__host__ __device__ constexpr int variant_of_kv(int kv) { return kv >> kVariantShift; }
__host__ __device__ constexpr int make_kv(int kind, int v) { return kind | (v << kVariantShift); }

// G regions per block and NT threads: every pass gives each thread one
// item per round, in straight-line code -- no loop for the compiler to hoist
// the FP64 constants of the special functions out of, which otherwise pins
// them in registers for the whole kernel and drops occupancy to one wave per
// SIMD.  Four shapes per dimension, picked per kind at run time
// (g3l_autotune): variant 0 fills one block of up to 1024 threads, variants
// 1 and 2 take fewer regions so that more blocks share a compute unit and
// one block's barriers overlap another's work, and variant 3 is a group
// size whose points fill their waves closely:
//   n_bar  (1-D): 68 regions x 15 points = 1020 threads (all variants);
//   1-halo (5-D): 11 x 93 = 1023 points, 5 x 93 = 465 of 512, 2 x 93 = 186 of 192,
//                 4 x 93 = 372 of 384;
//   2-halo (6-D):  6 x 149 = 894 points of 896 (u_NFW pass 6 x 318 = 1908 in
//                  three rounds), 2 x 149 = 298 of 320, 1 x 149 of 192,
//                  3 x 149 = 447 of 512;
//   3-halo (7-D):  4 x 241 = 964 points, 2 x 241 = 482 of 512, 1 x 241 of 256,
//                  3 x 241 = 723 of 768.
template <int KV>
struct KGroup {
  static constexpr int D = kind_dim(kind_of_kv(KV));
  static constexpr int V = variant_of_kv(KV);
  static constexpr int P = rule_points(D);
  static constexpr int G = D == 1   ? 68
                           : D == 5 ? (V == 3 ? 4 : V == 2 ? 2 : V ? 5 : 11)
                           : D == 6 ? (V == 3 ? 3 : V == 2 ? 1 : V ? 2 : 6)
                                    : (V == 3 ? 3 : V == 2 ? 1 : V ? 2 : 4);
  // the 2-halo kinds at three regions take 512 threads rather than 448, so
  // that their 954 u_NFW items fit two per thread
  static constexpr int NT = D == 6 && V == 3 ? 512 : (G * P + 63) / 64 * 64;
  static_assert(NT <= 1024, "one block holds every point of its regions");
};

// Region staging shared by both evaluation kernels: centre, half-width,
// volume, slot and integral parameters of the block's regions.
template <int D, int G>
struct Stage {
  double c[G][D];
  double h[G][D];
  double vol[G];
  uint32_t slot[G];
  TermArgs args[G];
};

// The group of G list entries (regions) this block evaluates.  A grid's
// blocks are dealt to the eight XCDs in turn (block b to XCD b mod 8), each
// XCD with its own L2, and neighbouring list entries mostly hold neighbouring
// slots of every geometry field; so each XCD takes runs of C consecutive
// groups (one run per XCD in turn) rather than every eighth group, and the
// runs stay short enough that every XCD still sees a mix of the list's
// integrals.  A bijection on [0, gridDim.x) whatever the dispatch order (the
// tail beyond the last full round of runs keeps its order): the results do
// not depend on it.
__device__ __forceinline__ uint32_t block_group() {
  constexpr uint32_t X = 8, C = 16;
  const uint32_t b = blockIdx.x, full = gridDim.x / (X * C) * (X * C);
  if (b >= full) return b;
  const uint32_t x = b % X, i = b / X;
  return ((i / C) * X + x) * C + i % C;
}

template <int D, int G, int NT>
__device__ __forceinline__ int stage_regions(const EvalLaunch &L, Stage<D, G> &S) {
  const int tid = threadIdx.x;
  const uint32_t first = block_group() * G;
  const int ng = (L.count - first) < (uint32_t)G ? (int)(L.count - first) : G;
  const uint64_t St = L.stride;
  if (tid < ng) {
    EvalItem it = L.list[first + tid];
    bool ok = it.slot < St && it.job < L.njobs;
    S.slot[tid] = ok ? it.slot : 0xffffffffu;
    if (ok) {
      S.args[tid] = L.args[L.jobs[it.job].args];
      S.vol[tid] = L.geo[kFieldVol * St + it.slot];
    }
  }
  __syncthreads();
  for (int i = tid; i < ng * D; i += NT) {
    int g = i / D, d = i - g * D;
    uint32_t sl = S.slot[g];
    if (sl == 0xffffffffu) continue;
    S.c[g][d] = L.geo[(kFieldC + d) * St + sl];
    S.h[g][d] = L.geo[(kFieldH + d) * St + sl];
  }
  __syncthreads();
  return ng;
}

// Rule reduction of the block's regions from the point values in f.  For the
// Genz-Malik rule lanes 2g and 2g + 1 of the first wave sum region g's 4th
// and 5th point groups (gm_group_sum), and lane 2g finishes the region.
template <int D, int G, int NP>
__device__ __forceinline__ void reduce_regions(const EvalLaunch &L, const Stage<D, G> &S,
                                               const double *f, int ng) {
  const int tid = threadIdx.x;
  const uint64_t St = L.stride;
  if constexpr (D == 1) {
    if (tid < ng && S.slot[tid] != 0xffffffffu) {
      double val, err;
      gk_reduce(f + tid * NP, S.h[tid][0], val, err);
      const uint32_t sl = S.slot[tid];
      L.geo[kFieldVal * St + sl] = val;
      L.geo[kFieldErr * St + sl] = err;
      L.split[sl] = 0;
    }
  } else {
    static_assert(2 * G <= 64, "a region's two lanes are in the first wave");
    if (tid < 2 * ng) {
      const int g = tid >> 1;
      const bool fifth = tid & 1;
      const double *v = f + g * NP;
      const double s = gm_group_sum<D>(v, fifth);
      const double s5 = __shfl_xor(s, 1);   // every lane below 2 ng is here
      if (!fifth && S.slot[g] != 0xffffffffu) {
        double val, err;
        int split = 0;
        gm_finish<D>(v, S.vol[g], s, s5, val, err, split);
        const uint32_t sl = S.slot[g];
        L.geo[kFieldVal * St + sl] = val;
        L.geo[kFieldErr * St + sl] = err;
        L.split[sl] = (uint8_t)split;
      }
    }
  }
}

// Direct evaluation (n_bar): every point evaluates its integrand in full.
// Held to 64 VGPRs (8 waves per SIMD, two 1024-thread blocks per CU) as
// k_eval is: the 256-HOD n_bar sweep runs 148 ms instead of 158 ms at 88.
template <int KIND, class P>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_eval_direct(EvalLaunch L) {
  constexpr int D = kind_dim(kind_of_kv(KIND));
  constexpr int NP = KGroup<KIND>::P;
  constexpr int G = KGroup<KIND>::G;
  constexpr int NT = KGroup<KIND>::NT;
  __shared__ Stage<D, G> S;
  __shared__ double s_f[G * NP];
  if (block_group() * G >= L.count) return;
  const int ng = stage_regions<D, G, NT>(L, S);
  const int tid = threadIdx.x;
  const DevTables T = L.T;
  if (tid < ng * NP) {
    const int g = tid / NP, q = tid - g * NP;
    if (S.slot[g] != 0xffffffffu) {
      double x[D];
      rule_point<D>(q, S.c[g], S.h[g], x);
      s_f[tid] = integrand<kind_of_kv(KIND), P>(T, S.args[g], x);
    }
  }
  __syncthreads();
  reduce_regions<D, G, NP>(L, S, s_f, ng);
}

// ---------------------------------------------------- factored evaluation
// The NNMap/NMapMap integrands evaluated in three passes per block:
//   1. per region, each distinct value of the one- and two-coordinate
//      quantities: l1, l2 and their filters, cos phi, the redshift look-ups,
//      and per mass coordinate the mass, its indices, r_200 and the HOD
//      occupations (a function of (m, z));
//   2. per region, each distinct value of the u_NFW calls that depend on
//      part of the coordinates (33 of 93 points for u(k1, m, z) in 5-D) and
//      of the tree-level bispectrum (57 of 241 points in 7-D);
//   3. per point, the rest (u_NFW of k3 where it depends on every
//      coordinate, the products) from the cached values.
// Each cached value is computed by the same expressions from the same
// coordinates as in DirectSrc, so the point values are bit-identical.

// 72 bytes: it and zt (table indices below 2^16) share a word, so that a
// record's stride across the LDS banks stays 18 dwords.
struct MassC {
  double m, L, r200;
  int ir;
  uint16_t it, zt;
  double nc1, ns1, nc2, ns2, nss;
};
static_assert(sizeof(MassC) == 72, "MassC stride");

// u_NFW normalisation (c, ln(1+c), F) per f selector (1, f1, f2).
struct NfwC {
  double c[3], lc[3], F[3];
};

// Whether pass 1 caches the u_NFW normalisations per (mass, z) class (5-D and
// 6-D kinds).  The 7-D kinds compute them per u_NFW item: their classes are
// few, their pass 1 is the longest, and the larger cache costs them more than
// the logarithms it saves.
template <int KIND>
constexpr bool cache_nfw() { return Spec<kind_of_kv(KIND)>::D <= 6; }

// Whether pass 1 runs in two stages (the kinds that cache the NFW
// normalisations): the masses first, then the occupations and the
// normalisations as items of their own.  A (mass, z) item that does all of it
// is the pass's longest by far (pow, log, cbrt, two erf/exp occupations,
// three normalisations), and the block's other waves wait for it at the
// barrier; staged, the longest item is about a third as long, with one more
// barrier and no work repeated (NNMap 1h launches 3 %, 2h 1 % faster).
template <int KIND>
constexpr bool staged_p1() { return Spec<kind_of_kv(KIND)>::D <= 6; }

template <int KIND, int I>
struct UF {
  static constexpr int D = Spec<kind_of_kv(KIND)>::D;
  static constexpr USpec s = Spec<kind_of_kv(KIND)>::u[I];
  static constexpr int n = ProjTab<D, s.mask>::n;
};
// Offset of u factor I in the flat cache, in units of regions x classes.
template <int KIND, int I>
struct UOff {
  static constexpr int v = UOff<KIND, I - 1>::v + KGroup<KIND>::G * UF<KIND, I - 1>::n;
};
template <int KIND>
struct UOff<KIND, 0> {
  static constexpr int v = 0;
};

template <int KIND>
struct Cache {
  using S = Spec<kind_of_kv(KIND)>;
  static constexpr int D = S::D, G = KGroup<KIND>::G;
  static constexpr int NU_TOTAL = UOff<KIND, S::NU>::v;
  static constexpr int NB = S::has_bilin ? ProjTab<D, bilin_mask<D>()>::n : 1;
  static_assert(ProjTab<D, kL1>::n == 7 && ProjTab<D, kL2>::n == 7 &&
                    ProjTab<D, bit(2)>::n == 7 && ProjTab<D, bit(D - 1)>::n == 7,
                "seven coordinates per dimension");
  static_assert(ProjTab<D, kL1 | bit(D - 1)>::n == 17 && ProjTab<D, kL2 | bit(D - 1)>::n == 17,
                "seventeen (l, z) classes");
  static_assert(ProjTab<D, bit(S::mdim[0]) | bit(D - 1)>::n == 17 &&
                    ProjTab<D, bit(S::mdim[1]) | bit(D - 1)>::n == 17 &&
                    ProjTab<D, bit(S::mdim[2]) | bit(D - 1)>::n == 17,
                "seventeen (mass, z) classes");
  double l1[G][7], ap1[G][7], l2[G][7], ap2[G][7], cphi[G][7], z[G][7], w[G][7];
  double zpre[G][7], zfac[G][7];
  double rw[G][7];   // 1/w, correctly rounded: k = l / w by div_rn
  // NNMap 3-halo: dwdz, 1/dwdz, g, p1, p2, 1/w per redshift (proj_3h)
  double ztab[kind_of_kv(KIND) == kNN3 ? G : 1][7][6];
  int zr[G][7], zt[G][7];
  MassC m[S::NM][G][17];
  NfwC nf[cache_nfw<KIND>() ? S::NM : 1][cache_nfw<KIND>() ? G : 1][17];
  double pk[S::has_pk ? 2 : 1][G][17];  // P(k1), P(k2) per (l, z) class
  static constexpr int N3 = ProjTab<D, kL3>::n;             // (l1, l2, phi) classes
  static constexpr int N3Z = ProjTab<D, kL3 | bit(D - 1)>::n;  // (l1, l2, phi, z) classes
  double l3[G][N3], ap3[G][N3];
  double pk3[S::has_pk ? G : 1][N3Z];                      // P(k3), 2-halo terms
  double u[NU_TOTAL];
  double bl[G * NB];
};

// Per-point view of the cache (the value source of the integrands).
template <int KIND, class P>
struct CachedSrc {
  using S = Spec<kind_of_kv(KIND)>;
  static constexpr int D = S::D, G = KGroup<KIND>::G;
  const Cache<KIND> &C;
  const TermArgs &a;
  int g, q;
  Geo G_;
  __device__ __forceinline__ CachedSrc(const Cache<KIND> &C_, const TermArgs &a_, int g_, int q_)
      : C(C_), a(a_), g(g_), q(q_) {
    const int c1 = kProj<D, kL1>.cls[q], c2 = kProj<D, kL2>.cls[q];
    const int cp = kProj<D, bit(2)>.cls[q], cz = kProj<D, bit(D - 1)>.cls[q];
    G_.l1 = C.l1[g][c1];
    G_.l2 = C.l2[g][c2];
    G_.cphi = C.cphi[g][cp];
    G_.z = C.z[g][cz];
    G_.zr = C.zr[g][cz];
    G_.zt = C.zt[g][cz];
    G_.w = C.w[g][cz];
    G_.l3 = C.l3[g][kProj<D, kL3>.cls[q]];
    G_.k1 = G_.l1 / G_.w;
    G_.k2 = G_.l2 / G_.w;
    G_.k3 = G_.l3 / G_.w;
  }
  __device__ __forceinline__ const MassC &mc(int i) const {
    if (i == 0) return C.m[0][g][kProj<D, bit(S::mdim[0]) | bit(D - 1)>.cls[q]];
    if (S::NM > 1 && i == 1) return C.m[S::NM > 1 ? 1 : 0][g][kProj<D, bit(S::mdim[1]) | bit(D - 1)>.cls[q]];
    return C.m[S::NM > 2 ? 2 : 0][g][kProj<D, bit(S::mdim[2]) | bit(D - 1)>.cls[q]];
  }
  __device__ __forceinline__ const Geo &geo() const { return G_; }
  __device__ __forceinline__ Mass mass(int i) const {
    const MassC &e = mc(i);
    Mass M;
    M.m = e.m;
    M.L = e.L;
    M.ir = e.ir;
    M.it = e.it;
    M.r200 = e.r200;
    return M;
  }
  __device__ __forceinline__ Occ occ1(int i) const {
    const MassC &e = mc(i);
    return Occ{e.nc1, e.ns1};
  }
  __device__ __forceinline__ Occ occ2(int i) const {
    const MassC &e = mc(i);
    return Occ{e.nc2, e.ns2};
  }
  __device__ __forceinline__ double nss(int i) const { return mc(i).nss; }
  __device__ __forceinline__ Nfw nfw(int i, int fsel) const {
    static_assert(cache_nfw<KIND>(), "per-point u_NFW normalisations come from the cache");
    const int cls = i == 0   ? kProj<D, bit(S::mdim[0]) | bit(D - 1)>.cls[q]
                    : i == 1 ? kProj<D, bit(S::mdim[S::NM > 1 ? 1 : 0]) | bit(D - 1)>.cls[q]
                             : kProj<D, bit(S::mdim[S::NM > 2 ? 2 : 0]) | bit(D - 1)>.cls[q];
    const NfwC &e = C.nf[i][g][cls];
    return Nfw{mc(i).r200, e.c[fsel], e.lc[fsel], e.F[fsel]};
  }
  template <int I>
  __device__ __forceinline__ double u() const {
    using F = UF<KIND, I>;
    return C.u[UOff<KIND, I>::v + g * F::n + kProj<D, F::s.mask>.cls[q]];
  }
  __device__ __forceinline__ double ap1() const { return C.ap1[g][kProj<D, kL1>.cls[q]]; }
  __device__ __forceinline__ double ap2() const { return C.ap2[g][kProj<D, kL2>.cls[q]]; }
  __device__ __forceinline__ double bl() const {
    return C.bl[g * Cache<KIND>::NB + kProj<D, bilin_mask<D>()>.cls[q]];
  }
  __device__ __forceinline__ double ap3() const { return C.ap3[g][kProj<D, kL3>.cls[q]]; }
  __device__ __forceinline__ double pk3() const {
    return C.pk3[S::has_pk ? g : 0][kProj<D, kL3 | bit(D - 1)>.cls[q]];
  }
  __device__ __forceinline__ double pk1() const {
    return C.pk[0][g][kProj<D, kL1 | bit(D - 1)>.cls[q]];
  }
  __device__ __forceinline__ double pk2() const {
    return C.pk[S::has_pk ? 1 : 0][g][kProj<D, kL2 | bit(D - 1)>.cls[q]];
  }
  __device__ __forceinline__ double zpre() const {
    return C.zpre[g][kProj<D, bit(D - 1)>.cls[q]];
  }
  __device__ __forceinline__ double proj(double B) const {
    const int cz = kProj<D, bit(D - 1)>.cls[q];
    if constexpr (kind_of_kv(KIND) == kNN3) {
      const double *t = C.ztab[g][cz];
      return proj_3h(B, t[0], t[1], t[2], t[3], t[4], C.w[g][cz], t[5], C.z[g][cz]);
    } else {
      return C.zfac[g][cz] * B;
    }
  }
};

// Pass 1 item layout: wave-aligned segments of items that run the same code:
// l1 and l2 with their filters (2 x G x 7), cos phi (G x 7), the redshift
// look-ups and z-only factors (G x 7), every mass coordinate's (m, z) class
// (NM x G x 17; the coordinates of one kind need the same u_NFW
// normalisations), and for the 2-halo terms P(k1), P(k2) per (l, z) class
// (2 x G x 17).
template <int KIND>
struct P1Layout {
  using S = Spec<kind_of_kv(KIND)>;
  static constexpr int G = KGroup<KIND>::G;
  static constexpr int pad(int n) { return (n + 63) / 64 * 64; }
  static constexpr int L = 0, PHI = pad(2 * G * 7), Z = PHI + pad(G * 7), MZ = Z + pad(G * 7);
  static constexpr int PK = MZ + pad(S::NM * G * 17);
  static constexpr int END = PK + (S::has_pk ? pad(2 * G * 17) : 0);
};

template <int KIND, class P>
__device__ __forceinline__ void pass1(const DevTables &T, const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                      Cache<KIND> &C, int i, int ng) {
  using S = Spec<kind_of_kv(KIND)>;
  using Lay = P1Layout<KIND>;
  constexpr int D = S::D, G = KGroup<KIND>::G;
  double x[D];
  if (i < Lay::PHI) {
    const int r = i - Lay::L;
    if (r >= 2 * G * 7) return;
    const int which = r / (G * 7), rr = r - which * G * 7, g = rr / 7, c = rr - g * 7;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    const TermArgs &a = St.args[g];
    // l1 and l2 items share one wave: one code path with selected operands
    rule_point<D>(which == 0 ? kProj<D, kL1>.rep[c] : kProj<D, kL2>.rep[c], St.c[g], St.h[g], x);
    const double l = P::l(which == 0 ? x[0] : x[1]);
    const double ap = P::ap((which == 0 ? a.t1 : a.t2) * l);
    if (which == 0) {
      C.l1[g][c] = l;
      C.ap1[g][c] = ap;
    } else {
      C.l2[g][c] = l;
      C.ap2[g][c] = ap;
    }
  } else if (i < Lay::Z) {
    const int r = i - Lay::PHI;
    if (r >= G * 7) return;
    const int g = r / 7, c = r - g * 7;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    rule_point<D>(kProj<D, bit(2)>.rep[c], St.c[g], St.h[g], x);
    C.cphi[g][c] = cos(x[2]);
  } else if (i < Lay::MZ) {
    const int r = i - Lay::Z;
    if (r >= G * 7) return;
    const int g = r / 7, c = r - g * 7;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    rule_point<D>(kProj<D, bit(D - 1)>.rep[c], St.c[g], St.h[g], x);
    const double z = x[D - 1];
    const int zr = zix_round(T, z);
    const double w = T.w[zr];
    C.z[g][c] = z;
    C.zr[g][c] = zr;
    C.zt[g][c] = zix_trunc(T, z);
    C.w[g][c] = w;
    C.rw[g][c] = 1. / w;
    zfactors<kind_of_kv(KIND)>(T, St.args[g], zr, z, w, C.zpre[g][c], C.zfac[g][c]);
    if constexpr (kind_of_kv(KIND) == kNN3) {
      double *t = C.ztab[g][c];
      t[0] = T.dwdz[zr];
      t[1] = 1. / t[0];
      t[2] = T.g[zr];
      t[3] = T.p1[zr];
      t[4] = T.p2[zr];
      t[5] = 1. / w;
    }
  } else if (i < Lay::PK) {
    const int r = i - Lay::MZ;
    if (r >= S::NM * G * 17) return;
    const int slot = r / (G * 17), rr = r - slot * G * 17, g = rr / 17, c = rr - g * 17;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    const TermArgs &a = St.args[g];
    const int md = slot == 0 ? S::mdim[0] : slot == 1 ? S::mdim[1] : S::mdim[2];
    int rep;
    if (slot == 0) rep = kProj<D, bit(S::mdim[0]) | bit(D - 1)>.rep[c];
    else if (slot == 1) rep = kProj<D, bit(S::mdim[S::NM > 1 ? 1 : 0]) | bit(D - 1)>.rep[c];
    else rep = kProj<D, bit(S::mdim[S::NM > 2 ? 2 : 0]) | bit(D - 1)>.rep[c];
    rule_point<D>(rep, St.c[g], St.h[g], x);
    const int zt = zix_trunc(T, x[D - 1]);
    const Mass M = make_mass<P>(T, x[md], zt);
    MassC &e = C.m[slot][g][c];
    e.m = M.m;
    e.L = M.L;
    e.r200 = M.r200;
    e.ir = M.ir;
    e.it = (uint16_t)M.it;
    e.zt = (uint16_t)zt;
    if constexpr (staged_p1<KIND>()) return;   // the rest in pass1b
    if constexpr (kind_of_kv(KIND) == kNN3) {
      // the NNMap 3-halo term needs the first HOD at m1 and the second at m2
      // only: one occupation per item, the HOD selected by the mass slot
      const bool second = slot == 1;
      const Occ o = P::occ(second ? a.h2 : a.h1, second ? a.lth2 : a.lth1,
                           second ? a.lmp2 : a.lmp1, M);
      e.nc1 = e.nc2 = o.nc;
      e.ns1 = e.ns2 = o.ns;
    } else {
      const Occ o1 = occ1<P>(a, M);
      e.nc1 = o1.nc;
      e.ns1 = o1.ns;
      if constexpr (S::two_hods) {
        const Occ o2 = occ2<P>(a, M);
        e.nc2 = o2.nc;
        e.ns2 = o2.ns;
        if (slot == S::nss_slot) e.nss = nsatnsat<P>(a, o1.ns, o2.ns, M);
      }
    }
    if constexpr (cache_nfw<KIND>()) {
      NfwC &en = C.nf[slot][g][c];
#pragma unroll
      for (int fs = 0; fs < 3; ++fs) {
        const bool need = slot == 0 ? uses_nfw<kind_of_kv(KIND)>(0, fs)
                                    : slot == 1 ? uses_nfw<kind_of_kv(KIND)>(1, fs) : uses_nfw<kind_of_kv(KIND)>(2, fs);
        if (need) {
          const Nfw nf = make_nfw(T, M, zt, f_of(a, fs));
          en.c[fs] = nf.c;
          en.lc[fs] = nf.lc;
          en.F[fs] = nf.F;
        }
      }
    }
  } else if constexpr (S::has_pk) {
    const int r = i - Lay::PK;
    if (r >= 2 * G * 17) return;
    const int which = r / (G * 17), rr = r - which * G * 17, g = rr / 17, c = rr - g * 17;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    // P(k) at k = l / w, the same expressions as the per-point Geo
    rule_point<D>(which == 0 ? kProj<D, kL1 | bit(D - 1)>.rep[c] : kProj<D, kL2 | bit(D - 1)>.rep[c],
                  St.c[g], St.h[g], x);
    const double l = P::l(x[which]);
    const int zr = zix_round(T, x[D - 1]);
    C.pk[which][g][c] = pk(T, zr, l / T.w[zr]);
  }
}

// Second stage of pass 1 (staged_p1 kinds): item i is one of a (mass, z)
// class's NFW normalisations (3 x NM x G x 17 items, one per f selector) or
// HOD occupations (the first HOD's, and the second's for the two-HOD kinds:
// NM x G x 17 items each), from the masses of the first stage, or l3 with its
// aperture filter per (l1, l2, phi) class (G x N3 items), from the first
// stage's l1, l2 and cos phi.  NsatNsat, which needs both occupations, is a
// pass-2 item.  The normalisations come first: where a thread takes a second
// item, its first is one of the shortest.
template <int KIND>
struct P1bLayout {
  using S = Spec<kind_of_kv(KIND)>;
  static constexpr int G = KGroup<KIND>::G;
  static constexpr int pad(int n) { return (n + 63) / 64 * 64; }
  static constexpr int NMZ = S::NM * G * 17;
  static constexpr int OCC1 = pad(3 * NMZ), OCC2 = OCC1 + pad(NMZ);
  static constexpr int L3 = OCC2 + (S::two_hods ? pad(NMZ) : 0);
  static constexpr int END = L3 + pad(G * ProjTab<S::D, kL3>::n);
};

__device__ __forceinline__ Mass mass_of(const MassC &e) {
  Mass M;
  M.m = e.m;
  M.L = e.L;
  M.ir = e.ir;
  M.it = e.it;
  M.r200 = e.r200;
  return M;
}

template <int KIND, class P>
__device__ __forceinline__ void pass1b(const DevTables &T, const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                       Cache<KIND> &C, int i, int ng) {
  using S = Spec<kind_of_kv(KIND)>;
  using Lay = P1bLayout<KIND>;
  constexpr int D = S::D, G = KGroup<KIND>::G;
  if (i < Lay::OCC1) {
    if (i >= 3 * Lay::NMZ) return;
    const int fs = i / Lay::NMZ, rm = i - fs * Lay::NMZ;
    const int slot = rm / (G * 17), rr = rm - slot * G * 17, g = rr / 17, c = rr - g * 17;
    const bool need = slot == 0 ? uses_nfw<kind_of_kv(KIND)>(0, fs)
                                : slot == 1 ? uses_nfw<kind_of_kv(KIND)>(1, fs) : uses_nfw<kind_of_kv(KIND)>(2, fs);
    if (!need || g >= ng || St.slot[g] == 0xffffffffu) return;
    const MassC &e = C.m[slot][g][c];
    const Nfw nf = make_nfw(T, mass_of(e), e.zt, f_of(St.args[g], fs));
    NfwC &en = C.nf[slot][g][c];
    en.c[fs] = nf.c;
    en.lc[fs] = nf.lc;
    en.F[fs] = nf.F;
    return;
  }
  if (i < Lay::L3) {
    const bool second = i >= Lay::OCC2;
    const int r = i - (second ? Lay::OCC2 : Lay::OCC1);
    if (r >= Lay::NMZ) return;
    const int slot = r / (G * 17), rr = r - slot * G * 17, g = rr / 17, c = rr - g * 17;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    MassC &e = C.m[slot][g][c];
    const Mass M = mass_of(e);
    const TermArgs &a = St.args[g];
    if (!second) {
      const Occ o = occ1<P>(a, M);
      e.nc1 = o.nc;
      e.ns1 = o.ns;
    } else if constexpr (S::two_hods) {
      const Occ o = occ2<P>(a, M);
      e.nc2 = o.nc;
      e.ns2 = o.ns;
    }
    return;
  }
  // l3 and its filter per (l1, l2, phi) class, the expressions of the Geo
  constexpr int N3 = Cache<KIND>::N3;
  const int r = i - Lay::L3;
  if (r >= G * N3) return;
  const int g = r / N3, c = r - g * N3;
  if (g >= ng || St.slot[g] == 0xffffffffu) return;
  const int q = kProj<D, kL3>.rep[c];
  const double l1 = C.l1[g][kProj<D, kL1>.cls[q]], l2 = C.l2[g][kProj<D, kL2>.cls[q]];
  const double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * C.cphi[g][kProj<D, bit(2)>.cls[q]]);
  C.l3[g][c] = l3;
  C.ap3[g][c] = P::ap(St.args[g].t3 * l3);
}

// Pass 2: the inputs of u-factor item i (factor I or later).
template <int KIND, class P, int I>
__device__ __forceinline__ bool u_inputs(const DevTables &T,
                                         const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                         const Cache<KIND> &C, int i, int ng, int &dst, Nfw &n,
                                         double &k, uint32_t &nfe) {
  using S = Spec<kind_of_kv(KIND)>;
  if constexpr (I >= S::NU) {
    return false;
  } else {
    constexpr int D = S::D;
    using F = UF<KIND, I>;
    constexpr int lo = UOff<KIND, I>::v, hi = UOff<KIND, I + 1>::v;
    if (i >= hi) return u_inputs<KIND, P, I + 1>(T, St, C, i, ng, dst, n, k, nfe);
    const int r = i - lo, g = r / F::n, c = r - g * F::n;
    if (g >= ng || St.slot[g] == 0xffffffffu) return false;
    const int q = kProj<D, F::s.mask>.rep[c];
    const int cz = kProj<D, bit(D - 1)>.cls[q];
    const double w = C.w[g][cz];
    double l;
    if constexpr (F::s.kidx == 1) {
      l = C.l1[g][kProj<D, kL1>.cls[q]];
    } else if constexpr (F::s.kidx == 2) {
      l = C.l2[g][kProj<D, kL2>.cls[q]];
    } else if constexpr (staged_p1<KIND>()) {
      l = C.l3[g][kProj<D, kL3>.cls[q]];   // from pass 1, the same expression
    } else {
      const double l1 = C.l1[g][kProj<D, kL1>.cls[q]], l2 = C.l2[g][kProj<D, kL2>.cls[q]];
      const double cp = C.cphi[g][kProj<D, bit(2)>.cls[q]];
      l = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * cp);
    }
    k = div_rn(l, w, C.rw[g][cz]);   // = l / w (IEEE), in three instructions
    const int cm = kProj<D, bit(S::mdim[F::s.slot]) | bit(D - 1)>.cls[q];
    const MassC &e = C.m[F::s.slot][g][cm];
    if constexpr (cache_nfw<KIND>()) {
      const NfwC &en = C.nf[F::s.slot][g][cm];
      n = Nfw{e.r200, en.c[F::s.fsel], en.lc[F::s.fsel], en.F[F::s.fsel]};
      // the normalisations' entry in C.nf (flat), times 4, plus the f selector
      nfe = (uint32_t)(((F::s.slot * KGroup<KIND>::G + g) * 17 + cm) * 4 + F::s.fsel);
    } else {
      Mass M;
      M.m = e.m;
      M.L = e.L;
      M.ir = e.ir;
      M.it = e.it;
      M.r200 = e.r200;
      n = make_nfw(T, M, C.zt[g][cz], f_of(St.args[g], F::s.fsel));
    }
    dst = i;
    return true;
  }
}

// Pass 2 item layout: the u_NFW items, then (64-aligned) the bispectrum
// items, l3 with its filter per (l1, l2, phi) class, and for the 2-halo terms
// P(k3) per (l1, l2, phi, z) class.
template <int KIND>
struct P2Layout {
  using S = Spec<kind_of_kv(KIND)>;
  static constexpr int G = KGroup<KIND>::G;
  static constexpr int pad(int n) { return (n + 63) / 64 * 64; }
  static constexpr int UPAD = pad(Cache<KIND>::NU_TOTAL);
  static constexpr int L3 = UPAD + (S::has_bilin ? pad(G * Cache<KIND>::NB) : 0);
  static constexpr int PK3 = L3 + (staged_p1<KIND>() ? 0 : pad(G * Cache<KIND>::N3));
  // NsatNsat per (mass, z) class of its mass coordinate, where pass 1 runs
  // in two stages (the occupations are then separate items)
  static constexpr int NSS = PK3 + (S::has_pk ? pad(G * Cache<KIND>::N3Z) : 0);
  static constexpr bool kNss = staged_p1<KIND>() && S::two_hods && S::nss_slot >= 0;
  static constexpr int END = NSS + (kNss ? pad(G * 17) : 0);
};

// Pass 2 items after the u_NFW ones (i >= the u items' padded count).
template <int KIND, class P>
__device__ __forceinline__ void pass2_rest(const DevTables &T, const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                           Cache<KIND> &C, int i, int ng) {
  using S = Spec<kind_of_kv(KIND)>;
  constexpr int D = S::D, G = KGroup<KIND>::G;
  constexpr int NU_TOTAL = Cache<KIND>::NU_TOTAL;
  constexpr int UPAD = (NU_TOTAL + 63) / 64 * 64;
  using Lay = P2Layout<KIND>;
  if constexpr (Lay::kNss) {
    if (i >= Lay::NSS) {
      const int r = i - Lay::NSS;
      if (r >= G * 17) return;
      const int g = r / 17, c = r - g * 17;
      if (g >= ng || St.slot[g] == 0xffffffffu) return;
      MassC &e = C.m[S::nss_slot > 0 ? S::nss_slot : 0][g][c];
      e.nss = nsatnsat<P>(St.args[g], e.ns1, e.ns2, mass_of(e));
      return;
    }
  }
  if (i < Lay::L3) {
    if constexpr (S::has_bilin) {
      constexpr int NB = Cache<KIND>::NB;
      const int r = i - UPAD;
      if (r < 0 || r >= G * NB) return;
      const int g = r / NB, c = r - g * NB;
      if (g >= ng || St.slot[g] == 0xffffffffu) return;
      const int q = kProj<D, bilin_mask<D>()>.rep[c];
      const int cz = kProj<D, bit(D - 1)>.cls[q];
      const double w = C.w[g][cz];
      const double k1 = C.l1[g][kProj<D, kL1>.cls[q]] / w, k2 = C.l2[g][kProj<D, kL2>.cls[q]] / w;
      C.bl[r] = bilin(T, k1, k2, C.cphi[g][kProj<D, bit(2)>.cls[q]], C.z[g][cz]);
    }
    return;
  }
  if (i < Lay::PK3) {
    // l3 and its filter per (l1, l2, phi) class, the expressions of the Geo
    constexpr int N3 = Cache<KIND>::N3;
    const int r = i - Lay::L3;
    if (r >= G * N3) return;
    const int g = r / N3, c = r - g * N3;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    const int q = kProj<D, kL3>.rep[c];
    const double l1 = C.l1[g][kProj<D, kL1>.cls[q]], l2 = C.l2[g][kProj<D, kL2>.cls[q]];
    const double l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * C.cphi[g][kProj<D, bit(2)>.cls[q]]);
    C.l3[g][c] = l3;
    C.ap3[g][c] = P::ap(St.args[g].t3 * l3);
    return;
  }
  if constexpr (S::has_pk) {
    // P(k3) per (l1, l2, phi, z) class
    constexpr int N3Z = Cache<KIND>::N3Z;
    const int r = i - Lay::PK3;
    if (r >= G * N3Z) return;
    const int g = r / N3Z, c = r - g * N3Z;
    if (g >= ng || St.slot[g] == 0xffffffffu) return;
    const int q = kProj<D, kL3 | bit(D - 1)>.rep[c];
    const int cz = kProj<D, bit(D - 1)>.cls[q];
    double l3;
    if constexpr (staged_p1<KIND>()) {
      l3 = C.l3[g][kProj<D, kL3>.cls[q]];
    } else {
      const double l1 = C.l1[g][kProj<D, kL1>.cls[q]], l2 = C.l2[g][kProj<D, kL2>.cls[q]];
      l3 = sqrt(l1 * l1 + l2 * l2 + 2 * l1 * l2 * C.cphi[g][kProj<D, bit(2)>.cls[q]]);
    }
    C.pk3[g][c] = pk(T, C.zr[g][cz], l3 / C.w[g][cz]);
  }
}

template <int KIND, class P>
__device__ __forceinline__ void pass2(const DevTables &T, const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                      Cache<KIND> &C, int i, int ng) {
  if (i < Cache<KIND>::NU_TOTAL) {
    int dst;
    Nfw n;
    double k;
    uint32_t nfe;
    if (!u_inputs<KIND, P, 0>(T, St, C, i, ng, dst, n, k, nfe)) return;
    C.u[dst] = P::u(n, k);
    return;
  }
  pass2_rest<KIND, P>(T, St, C, i, ng);
}

// u_NFW items in branch order.  u_NFW has three branches (a1 >= 4,
// a1 < 4 <= a2, a2 < 4) and a wave runs every branch one of its lanes takes:
// on the 550-triplet workload about half of the waves of u items mix two or
// three branches and run about a third more instructions than their lanes
// need.  For the kinds that cache the NFW normalisations, pass 2 therefore (1) computes each u item's a1, keeps it
// in the item's own output slot and appends the item to its branch's list
// (one LDS atomic per wave), and (2) evaluates the items from the lists in
// branch order, so that a wave's lanes mostly share one branch.  Every item runs the same expressions on
// the same inputs either way: the values are bit-identical.
template <int KIND>
constexpr bool sort_u() { return cache_nfw<KIND>(); }

// Step 1 for u item i: a1 into C.u[i], its normalisations' entry into
// unf[i], and i into its branch's list: a1 >= 4 from the front of `ends`,
// a2 < 4 from the back of `ends`, a1 < 4 <= a2 from the front of `mid`.  The
// three branch counts of a wave are added to the block's counters (21-bit
// fields of one 64-bit word) by a single LDS atomic, whose old value gives
// each branch's base.
template <int KIND, class P>
__device__ __forceinline__ void u_classify(const DevTables &T,
                                           const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                           Cache<KIND> &C, int i, int ng, unsigned long long *cnt,
                                           uint16_t *unf, uint16_t *ends, uint16_t *mid) {
  constexpr int NU = Cache<KIND>::NU_TOTAL;
  int dst = 0;
  Nfw n;
  double k = 0;
  uint32_t nfe = 0;
  const bool ok = u_inputs<KIND, P, 0>(T, St, C, i, ng, dst, n, k, nfe);
  int b = 3;
  if (ok) {
    const double a1 = nfw_a1(n, k);
    C.u[dst] = a1;
    unf[dst] = (uint16_t)nfe;
    b = a1 >= 4 ? 0 : (a1 * (1 + n.c) >= 4 ? 1 : 2);
  }
  const uint64_t m0 = __ballot(b == 0), m1 = __ballot(b == 1), m2 = __ballot(b == 2);
  const uint64_t add = (uint64_t)__popcll(m0) | ((uint64_t)__popcll(m1) << 21) |
                       ((uint64_t)__popcll(m2) << 42);
  const int first = __ffsll((unsigned long long)__ballot(1)) - 1;
  uint64_t old = 0;
  if ((int)__lane_id() == first) old = atomicAdd(cnt, (unsigned long long)add);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)old, first);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(old >> 32), first);
  if (b == 3) return;
  const uint64_t o = ((uint64_t)hi << 32) | lo;
  const uint32_t base = (uint32_t)(o >> (21 * b)) & 0x1fffffu;
  const uint64_t m = b == 0 ? m0 : b == 1 ? m1 : m2;
  const uint32_t pos = base + __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
  if (b == 1) mid[pos] = (uint16_t)dst;
  else ends[b == 0 ? pos : NU - 1 - pos] = (uint16_t)dst;
}

// Step 1, round r of R over the pass-2 items tid + r NT: each round holds
// only the code its items can reach (u items, the rest, or both).
template <int KIND, class P, int r, int R>
__device__ __forceinline__ void sort_round(const DevTables &T,
                                           const Stage<Spec<kind_of_kv(KIND)>::D, KGroup<KIND>::G> &St,
                                           Cache<KIND> &C, int tid, int ng, unsigned long long *cnt,
                                           uint16_t *unf, uint16_t *ends, uint16_t *mid) {
  if constexpr (r < R) {
    constexpr int NT = KGroup<KIND>::NT, NU = Cache<KIND>::NU_TOTAL, P2 = P2Layout<KIND>::END;
    const int i = tid + r * NT;
    if constexpr ((r + 1) * NT <= NU) {
      u_classify<KIND, P>(T, St, C, i, ng, cnt, unf, ends, mid);
    } else if constexpr (r * NT >= NU) {
      if (i < P2) pass2_rest<KIND, P>(T, St, C, i, ng);
    } else {
      if (i < NU) u_classify<KIND, P>(T, St, C, i, ng, cnt, unf, ends, mid);
      else if (i < P2) pass2_rest<KIND, P>(T, St, C, i, ng);
    }
    sort_round<KIND, P, r + 1, R>(T, St, C, tid, ng, cnt, unf, ends, mid);
  }
}

// Step 3 for u item i, whose a1 is in C.u[i] and normalisations' entry in
// unf[i].
template <int KIND, class P>
__device__ __forceinline__ void u_eval_sorted(Cache<KIND> &C, int i, const uint16_t *unf) {
  const uint32_t e = unf[i];
  const NfwC &en = (&C.nf[0][0][0])[e >> 2];
  const int fs = e & 3;
  C.u[i] = P::u_a1(C.u[i], Nfw{0.0, en.c[fs], en.lc[fs], en.F[fs]});
}

// At most 64 VGPRs (8 waves per SIMD): the FP64 5-D and 7-D kinds otherwise
// compile to 70-80 and run 6-7 waves per SIMD.  The few registers this
// spills cost less than the lost occupancy: NNMap 1h 4.12 -> 3.69 ms and
// 3h 4.26 -> 3.93 ms per launch on the 550-triplet workload, 2h unchanged.
template <int KIND, class P>
__global__ __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(8))) void k_eval(EvalLaunch L) {
  using S = Spec<kind_of_kv(KIND)>;
  constexpr int D = S::D;
  constexpr int NP = KGroup<KIND>::P;
  constexpr int G = KGroup<KIND>::G;
  constexpr int NT = KGroup<KIND>::NT;
  static_assert(P1Layout<KIND>::END <= 2 * NT, "pass 1 fits two items per thread");
  constexpr int P2 = P2Layout<KIND>::END;
  static_assert(P2 <= 3 * NT, "pass 2 fits three items per thread");
  static_assert(G * NP <= NT, "pass 3 fits one point per thread");
  using Q = typename EvalPolicy<kind_of_kv(KIND), P>::type;
  constexpr bool kSort = sort_u<KIND>();
  constexpr int NU = Cache<KIND>::NU_TOTAL;
  // s_f holds the point values of pass 3, and before that the branch lists
  // of the u items (2 NU entries of 16 bits)
  constexpr int NF = kSort && (4 * NU + 7) / 8 > G * NP ? (4 * NU + 7) / 8 : G * NP;
  __shared__ Stage<D, G> St;
  __shared__ Cache<KIND> C;
  __shared__ double s_f[NF];
  __shared__ unsigned long long s_ucnt;   // u items per branch, 21-bit fields
  __shared__ uint16_t s_unf[kSort ? NU : 1];
  if (block_group() * G >= L.count) return;
  const int ng = stage_regions<D, G, NT>(L, St);
  const int tid = threadIdx.x;
  const DevTables T = L.T;
  if (kSort && tid == 0) s_ucnt = 0;
  pass1<KIND, Q>(T, St, C, tid, ng);
  if constexpr (P1Layout<KIND>::END > NT) pass1<KIND, Q>(T, St, C, tid + NT, ng);
  __syncthreads();
  if constexpr (staged_p1<KIND>()) {
    static_assert(P1bLayout<KIND>::END <= 2 * NT, "pass 1b fits two items per thread");
    pass1b<KIND, Q>(T, St, C, tid, ng);
    if constexpr (P1bLayout<KIND>::END > NT) pass1b<KIND, Q>(T, St, C, tid + NT, ng);
    __syncthreads();
  }
  if constexpr (kSort) {
    constexpr int R = (P2 + NT - 1) / NT;
    uint16_t *ends = reinterpret_cast<uint16_t *>(s_f), *mid = ends + NU;
    sort_round<KIND, Q, 0, R>(T, St, C, tid, ng, &s_ucnt, s_unf, ends, mid);
    __syncthreads();
    const unsigned long long cw = s_ucnt;
    const uint32_t nb = (uint32_t)cw & 0x1fffffu, nbm = nb + ((uint32_t)(cw >> 21) & 0x1fffffu);
    const uint32_t nu = nbm + (uint32_t)(cw >> 42);
    // fixed rounds rather than a loop: in a loop the compiler keeps the
    // polynomial coefficients in registers across iterations and spills them
    constexpr int RU = (NU + NT - 1) / NT;
#pragma unroll
    for (int r = 0; r < RU; ++r) {
      const uint32_t j = tid + r * NT;
      if (j < nu) {
        const uint16_t item = j < nb ? ends[j] : j < nbm ? mid[j - nb] : ends[NU - 1 - (j - nbm)];
        u_eval_sorted<KIND, Q>(C, item, s_unf);
      }
    }
  } else {
    pass2<KIND, Q>(T, St, C, tid, ng);
    if constexpr (P2 > NT) pass2<KIND, Q>(T, St, C, tid + NT, ng);
    if constexpr (P2 > 2 * NT) pass2<KIND, Q>(T, St, C, tid + 2 * NT, ng);
  }
  __syncthreads();
  if (tid < ng * NP) {
    const int g = tid / NP, q = tid - g * NP;
    if (St.slot[g] != 0xffffffffu) {
      CachedSrc<KIND, Q> src(C, St.args[g], g, q);
      s_f[tid] = integrand_src<kind_of_kv(KIND), Q>(T, St.args[g], src);
    }
  }
  __syncthreads();
  reduce_regions<D, G, NP>(L, St, s_f, ng);
}

// ----------------------------------------------------------------- roots
__device__ __forceinline__ double box_volume(int D, const double *h) {
  double v = 1;
  for (int i = 0; i < D; ++i) v *= 2 * h[i];
  return v;
}

__global__ void k_roots(RootsLaunch L) {
  uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= L.njobs) return;
  JobDesc d = L.jobs[j];
  const double *lo = L.boxes + (size_t)j * 2 * kMaxDim;
  const double *hi = lo + kMaxDim;
  const uint64_t S = L.stride;
  double h[kMaxDim];
  for (int i = 0; i < d.dim; ++i) {
    double c = 0.5 * (lo[i] + hi[i]);
    h[i] = 0.5 * (hi[i] - lo[i]);
    L.geo[(kFieldC + i) * S + d.base] = c;
    L.geo[(kFieldH + i) * S + d.base] = h[i];
  }
  L.geo[kFieldVol * S + d.base] = box_volume(d.dim, h);
  L.split[d.base] = 0;
  JobState st;
  st.neval = (uint64_t)d.pts;
  st.n = 1;
  st.flags = 0;
  st.rounds = 0;
  st.pad = 0;
  st.val = 0;
  st.err = 0;
  L.state[j] = st;
  uint32_t at = atomicAdd(&L.counters[d.kind], 1u);
  L.lists[d.kind][at] = EvalItem{d.base, j};
}

__global__ void k_admit(AdmitLaunch L) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= L.n) return;
  const uint32_t j = L.ids[t];
  const JobDesc d = L.desc[t];
  L.jobs[j] = d;
  L.args_table[j] = L.args[t];
  const double *lo = L.boxes + (size_t)t * 2 * kMaxDim;
  const double *hi = lo + kMaxDim;
  const uint64_t S = L.stride;
  double h[kMaxDim];
  for (int i = 0; i < d.dim; ++i) {
    double c = 0.5 * (lo[i] + hi[i]);
    h[i] = 0.5 * (hi[i] - lo[i]);
    L.geo[(kFieldC + i) * S + d.base] = c;
    L.geo[(kFieldH + i) * S + d.base] = h[i];
  }
  L.geo[kFieldVol * S + d.base] = box_volume(d.dim, h);
  L.split[d.base] = 0;
  JobState st;
  st.neval = (uint64_t)d.pts;
  st.n = 1;
  st.flags = 0;
  st.rounds = 0;
  st.pad = 0;
  st.val = 0;
  st.err = 0;
  L.state[j] = st;
  uint32_t at = atomicAdd(&L.counters[d.kind], 1u);
  L.lists[d.kind][at] = EvalItem{d.base, j};
}

__global__ void k_gather(GatherLaunch L) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < L.n) L.out[t] = L.state[L.ids[t]];
}

__global__ void k_zero(uint32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- select
// One 256-thread block (4 waves of 64 lanes) per integral.  The passes over
// the integral's regions are latency-bound (one block per integral, a few
// hundred iterations per thread for the largest), so each iteration issues
// four independent loads, and the scans use wave shuffles rather than
// thread-serial loops.
constexpr int kSelThreads = 256;

__device__ __forceinline__ uint64_t shfl_up_u64(uint64_t v, unsigned d) {
  const uint32_t lo = __shfl_up((uint32_t)v, d, 64), hi = __shfl_up((uint32_t)(v >> 32), d, 64);
  return ((uint64_t)hi << 32) | lo;
}

// Inclusive prefix sums over the block in thread order of (a, b); s_wa, s_wb
// hold the wave totals (4 each).  Ends with a barrier, so the buffers can be
// reused at once; ta, tb receive the block totals.
__device__ __forceinline__ void block_scan2(uint64_t &a, uint64_t &b, uint64_t *s_wa,
                                            uint64_t *s_wb, uint64_t &ta, uint64_t &tb) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (unsigned d = 1; d < 64; d <<= 1) {
    const uint64_t xa = shfl_up_u64(a, d), xb = shfl_up_u64(b, d);
    if (lane >= (int)d) {
      a += xa;
      b += xb;
    }
  }
  if (lane == 63) {
    s_wa[wave] = a;
    s_wb[wave] = b;
  }
  __syncthreads();
  uint64_t oa = 0, ob = 0;
  for (int w = 0; w < wave; ++w) {
    oa += s_wa[w];
    ob += s_wb[w];
  }
  ta = s_wa[0] + s_wa[1] + s_wa[2] + s_wa[3];
  tb = s_wb[0] + s_wb[1] + s_wb[2] + s_wb[3];
  a += oa;
  b += ob;
  __syncthreads();
}

// Block-wide reductions over an integral's error keys for the exact stopping
// point: the FP64 sum of the errors whose key is below T (Neumaier, each
// thread over its strided entries, then a fixed tree), and the largest key
// below T or smallest key above T with its multiplicity.  Every thread of the
// block calls them; results are valid in every thread.
__device__ double block_sum_below(const double *err, uint32_t n, uint64_t T, double *s_a,
                                  double *s_b) {
  const int tid = threadIdx.x;
  double s = 0, c = 0;
  for (uint32_t i = tid; i < n; i += kSelThreads) {
    const uint64_t k = err_key(err[i]);
    if (k >= T) continue;
    const double e = __builtin_bit_cast(double, k);
    const double t = s + e;
    c += fabs(s) >= fabs(e) ? (s - t) + e : (e - t) + s;
    s = t;
  }
  s_a[tid] = s;
  s_b[tid] = c;
  __syncthreads();
  for (int w = kSelThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      const double a = s_a[tid], b = s_a[tid + w];
      const double t = a + b;
      s_a[tid] = t;
      s_b[tid] += s_b[tid + w] + (fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a);
    }
    __syncthreads();
  }
  const double r = s_a[0] + s_b[0];
  __syncthreads();
  return r;
}

// below: the largest key < T; otherwise the smallest key > T.  Returns the
// key (0 / ~0 when none) and sets cnt to its multiplicity.
__device__ uint64_t block_next_key(const double *err, uint32_t n, uint64_t T, bool below,
                                   uint64_t *s_k, uint32_t *s_c, uint32_t &cnt) {
  const int tid = threadIdx.x;
  uint64_t best = below ? 0ull : ~0ull;
  bool any = false;
  for (uint32_t i = tid; i < n; i += kSelThreads) {
    const uint64_t k = err_key(err[i]);
    if (below ? (k < T && (!any || k > best)) : (k > T && (!any || k < best))) {
      best = k;
      any = true;
    }
  }
  s_k[tid] = best;
  s_c[tid] = any ? 1u : 0u;
  __syncthreads();
  for (int w = kSelThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      const uint64_t a = s_k[tid], b = s_k[tid + w];
      const uint32_t ha = s_c[tid], hb = s_c[tid + w];
      if (hb && (!ha || (below ? b > a : b < a))) s_k[tid] = b;
      s_c[tid] = ha | hb;
    }
    __syncthreads();
  }
  const uint64_t key = s_k[0];
  const bool found = s_c[0] != 0;
  __syncthreads();
  uint32_t m = 0;
  if (found)
    for (uint32_t i = tid; i < n; i += kSelThreads) m += err_key(err[i]) == key ? 1u : 0u;
  s_c[tid] = m;
  __syncthreads();
  for (int w = kSelThreads / 2; w > 0; w >>= 1) {
    if (tid < w) s_c[tid] += s_c[tid + w];
    __syncthreads();
  }
  cnt = s_c[0];
  __syncthreads();
  return found ? key : (below ? 0ull : ~0ull);
}

// Moves the batch boundary (key T with multiplicity cntT, `above` keys
// larger than T, r of the T-ties taken) until the count K = above + r is the
// first at which the remaining error meets thr (or K reaches kcap or n),
// computing the remaining error in FP64.  Every thread calls it.
__device__ void select_exact_count(const double *err, uint32_t n, uint64_t kcap, double thr,
                                   uint64_t &T, uint32_t &cntT, uint32_t &above, uint32_t &r,
                                   double *s_a, double *s_b, uint64_t *s_k, uint32_t *s_c) {
  const int tid = threadIdx.x;
  __shared__ double s_tail;
  __shared__ int s_act;
  const double below = block_sum_below(err, n, T, s_a, s_b);
  if (tid == 0) s_tail = below + (double)(cntT - r) * __builtin_bit_cast(double, T);
  // the loop ends: every step moves K by one towards the unique stopping count
  for (int guard = 0; guard < 64; ++guard) {
    if (tid == 0) {
      double tail = s_tail;
      int act = 0;
      for (;;) {
        const double eT = __builtin_bit_cast(double, T);
        const uint64_t K = (uint64_t)above + r;
        // not yet converged after K pops: pop one more (budget and regions left)
        if (tail > 0 && tail > thr && K < kcap && K < n) {
          if (r < cntT) {
            ++r;
            tail -= eT;
            continue;
          }
          act = 1;  // the next key below T
          break;
        }
        // converged already after K - 1 pops: one pop fewer
        const double prev = tail + eT;
        if (K > 1 && !(prev > 0 && prev > thr)) {
          tail = prev;
          if (r > 1) {
            --r;
            continue;
          }
          r = 0;
          act = 2;  // the next key above T, all of its ties taken
          break;
        }
        break;
      }
      s_tail = tail;
      s_act = act;
    }
    __syncthreads();
    const int act = s_act;
    if (act == 0) break;
    uint32_t c = 0;
    const uint64_t nk = block_next_key(err, n, T, act == 1, s_k, s_c, c);
    if (tid == 0) {
      if (act == 1) {
        above += cntT;
        T = nk;
        cntT = c;
        r = 1;
        s_tail -= __builtin_bit_cast(double, nk);
      } else {
        T = nk;
        cntT = c;
        above -= c;
        r = c;
      }
    }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSelThreads) void k_select(SelectLaunch L) {
  const uint32_t J = L.active ? L.active[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ JobState s_st;
  __shared__ double s_red[kSelThreads];
  __shared__ double s_red2[kSelThreads];
  __shared__ uint64_t s_redk[kSelThreads];
  __shared__ uint32_t s_cnt[kSelThreads];
  __shared__ unsigned long long s_sum[kSelThreads];
  __shared__ uint64_t s_wa[4], s_wb[4];
  __shared__ int s_flag;
  __shared__ uint64_t s_prefix, s_mask, s_cumsum, s_need;
  __shared__ uint32_t s_cumcnt, s_lastcnt, s_r, s_K, s_listbase;

  if (tid == 0) s_st = L.state[J];
  __syncthreads();
  if (s_st.flags & kJobDone) return;
  const JobDesc d = L.jobs[J];
  const uint64_t S = L.stride;
  const uint32_t n = s_st.n;
  const double *val = L.geo + kFieldVal * S + d.base;
  const double *err = L.geo + kFieldErr * S + d.base;

  // 1. totals: value with Neumaier compensation, error plain, each thread
  //    over i = tid, tid + 256, ... in that order, then a fixed tree: the
  //    result does not depend on scheduling.  Loads four at a time; absent
  //    entries enter as zeros, which leave every sum unchanged.
  double vs = 0, vc = 0, es = 0;
  uint64_t kmax = 0;
  bool bad = false;
  for (uint32_t i0 = tid; i0 < n; i0 += 4 * kSelThreads) {
    double v[4], e[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const uint32_t i = i0 + u * kSelThreads;
      v[u] = i < n ? val[i] : 0.0;
      e[u] = i < n ? err[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      bad |= !isfinite(v[u]) || !isfinite(e[u]);
      const double t = vs + v[u];
      vc += fabs(vs) >= fabs(v[u]) ? (vs - t) + v[u] : (v[u] - t) + vs;
      vs = t;
      es += e[u];
      const uint64_t k = err_key(e[u]);
      kmax = k > kmax ? k : kmax;
    }
  }
  s_red[tid] = vs;
  s_red2[tid] = vc;
  s_redk[tid] = kmax;
  s_cnt[tid] = bad ? 1u : 0u;
  __syncthreads();
  for (int w = kSelThreads / 2; w > 0; w >>= 1) {
    if (tid < w) {
      double a = s_red[tid], b = s_red[tid + w];
      double t = a + b;
      double c = fabs(a) >= fabs(b) ? (a - t) + b : (b - t) + a;
      s_red[tid] = t;
      s_red2[tid] += s_red2[tid + w] + c;
      s_redk[tid] = s_redk[tid] > s_redk[tid + w] ? s_redk[tid] : s_redk[tid + w];
      s_cnt[tid] |= s_cnt[tid + w];
    }
    __syncthreads();
  }
  const double V = s_red[0] + s_red2[0];
  const uint64_t maxkey = s_redk[0];
  const bool nonfinite = s_cnt[0] != 0;
  __syncthreads();
  s_red[tid] = es;
  __syncthreads();
  for (int w = kSelThreads / 2; w > 0; w >>= 1) {
    if (tid < w) s_red[tid] += s_red[tid + w];
    __syncthreads();
  }
  const double E = s_red[0];
  __syncthreads();

  // 2. stop tests in the order of src/hcubature.cpp:938-939.
  if (tid == 0) {
    JobState st = s_st;
    st.val = V;
    st.err = E;
    if (nonfinite) st.flags |= kJobNonfinite;
    int flag = 0;
    if (d.maxeval != 0 && st.neval >= d.maxeval) {
      st.flags |= kJobDone | kJobCapped;
      flag = 1;
    } else if (!(E > 0 && E > fabs(V) * d.reltol)) {
      st.flags |= kJobDone | kJobConverged;
      flag = 1;
    }
    s_st = st;
    s_flag = flag;
  }
  __syncthreads();
  if (s_flag) {
    if (tid == 0) {
      L.state[J] = s_st;
      if (L.done) L.done[atomicAdd(&L.counters[kNumKinds + 1], 1u)] = J;
    }
    return;
  }

  // 3. batch size: pops continue until the remaining error meets the
  //    tolerance or the evaluation budget is spent (src/hcubature.cpp:971-985).
  //    The pops take the regions in decreasing error; the batch is found by
  //    a radix select over the 64-bit error keys, 8 bits per level from the
  //    top, in fixed-point error units (sel_shift maps maxkey low enough that
  //    a sum over all n keys stays below 2^63; integer sums are exact and
  //    order-free).
  const uint64_t twoP = 2ull * (uint64_t)d.pts;
  uint64_t kcap = n;
  if (d.maxeval != 0) {
    uint64_t room = d.maxeval - s_st.neval;
    uint64_t kc = (room + twoP - 1) / twoP;
    kcap = kc < kcap ? kc : kcap;
  }
  const int shift = sel_shift(maxkey, n);
  const double thr = ldexp(fabs(V) * d.reltol, shift);
  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_cumsum = 0;
    s_cumcnt = 0;
  }
  for (int level = 0; level < 8; ++level) {
    const int sh = 56 - 8 * level;
    s_cnt[tid] = 0;
    s_sum[tid] = 0;
    __syncthreads();
    const uint64_t prefix = s_prefix, mask = s_mask;
    for (uint32_t i0 = tid; i0 < n; i0 += 4 * kSelThreads) {
      uint64_t k[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const uint32_t i = i0 + u * kSelThreads;
        k[u] = i < n ? err_key(err[i]) : ~0ull;  // ~0 matches no prefix
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        if ((k[u] & mask) == prefix && k[u] != ~0ull) {
          const int dgt = (int)((k[u] >> sh) & 0xff);
          atomicAdd(&s_cnt[dgt], 1u);
          atomicAdd(&s_sum[dgt], fixed_err(k[u], shift));
        }
      }
    }
    __syncthreads();
    // bins in decreasing digit order, thread t <-> digit 255 - t: running
    // sums from the top digit down
    const int bin = kSelThreads - 1 - tid;
    uint64_t cs = s_sum[bin], cc = s_cnt[bin], tot_s, tot_c;
    block_scan2(cs, cc, s_wa, s_wb, tot_s, tot_c);
    if (level == 0) {
      // the level-0 histogram holds every key: its sum is the total error
      if (tid == 0) s_need = batch_need(tot_s, thr);
      __syncthreads();
    }
    const uint64_t need = s_need, cs0 = s_cumsum, cc0 = s_cumcnt;
    // the pops end in the first non-empty bin (from the top) at which the
    // error sum reaches need or the count reaches kcap; the running sums only
    // grow in thread order, so the threads that pass the test are a suffix
    // and the found bin is its first
    const bool hit = cs0 + cs >= need || cc0 + cc >= kcap;
    const uint64_t hb = __ballot(hit);
    if ((tid & 63) == 0) s_wa[tid >> 6] = (uint64_t)__popcll(hb);
    __syncthreads();
    const int nhit = (int)(s_wa[0] + s_wa[1] + s_wa[2] + s_wa[3]);
    // thread of the found bin: the first one that hits (bin 0 if none does,
    // which the total meeting need rules out)
    const int tf = nhit > 0 ? kSelThreads - nhit : kSelThreads - 1;
    if (tid == tf) {
      const int found = kSelThreads - 1 - tf;
      s_prefix = prefix | ((uint64_t)found << sh);
      s_mask = mask | (0xffull << sh);
      s_cumsum = cs0 + cs - s_sum[found];
      s_cumcnt = (uint32_t)(cc0 + cc - s_cnt[found]);
      s_lastcnt = s_cnt[found];
    }
    __syncthreads();
  }
  // s_prefix is now the exact key of the last popped region; take r of the
  // regions carrying that key, lowest slot first.
  if (tid == 0) {
    uint64_t T = s_prefix;
    unsigned long long fxT = fixed_err(T, shift);
    const uint64_t need = s_need;
    uint64_t rest = need > s_cumsum ? need - s_cumsum : 0;
    uint64_t r = fxT > 0 ? (rest + fxT - 1) / fxT : (uint64_t)s_lastcnt;
    if (r < 1) r = 1;
    if (r > s_lastcnt) r = s_lastcnt;
    if ((uint64_t)s_cumcnt + r > kcap) r = kcap - s_cumcnt;
    s_r = (uint32_t)r;
  }
  __syncthreads();

  // 3b. The exact stopping point.  The reference pops in decreasing error,
  //     subtracting each popped error from the running total in FP64, and
  //     stops at the first count whose remaining error meets the tolerance
  //     (src/hcubature.cpp:971-985, src/converged.h:10-14).  The fixed-point
  //     keys above are truncated, by up to one unit (2^-46 of the largest
  //     error) per region, so on integrals whose errors span many decades
  //     their sum over thousands of small-error regions can misplace that
  //     count by a region or two.  The remaining error is recomputed here in
  //     FP64 (compensated) and the count stepped, one tie at a time and one
  //     key at a time, until it is the reference's.
  select_exact_count(err, n, kcap, fabs(V) * d.reltol, s_prefix, s_lastcnt, s_cumcnt, s_r,
                     s_red, s_red2, s_redk, s_cnt);

  if (tid == 0) {
    const uint32_t K = s_cumcnt + s_r;
    s_K = K;
    JobState st = s_st;
    if (st.n + K > d.cap) {
      st.flags |= kJobDone | kJobPoolFull;
      s_flag = 1;
    } else {
      s_listbase = atomicAdd(&L.counters[d.kind], 2u * K);
      if (s_listbase + 2u * K > L.list_cap[d.kind]) {
        st.flags |= kJobDone | kJobPoolFull;
        s_flag = 1;
      } else {
        atomicAdd(&L.counters[kNumKinds], 1u);
      }
    }
    s_st = st;
  }
  __syncthreads();
  if (s_flag) {
    if (tid == 0) {
      L.state[J] = s_st;
      if (L.done) L.done[atomicAdd(&L.counters[kNumKinds + 1], 1u)] = J;
    }
    return;
  }

  // 4. mark and cut in slot order; children go to the kind's list.  Each
  //    thread marks four consecutive slots per chunk of 1024; the ranks among
  //    the tied keys and among the selected regions come from block scans of
  //    the per-thread counts (carried across chunks in registers), and the
  //    chunk's selected slots are listed in LDS in rank order, so that
  //    consecutive lanes cut consecutive regions (contiguous child slots)
  //    and each cut reads its region's geometry before writing any.  The
  //    list holds the K halves that keep their parents' slots first and the
  //    K new slots after them, both in slot order, so that the regions an
  //    evaluation block stages sit in neighbouring slots of each field.
  __shared__ uint32_t s_selp[4 * kSelThreads];
  const uint64_t T = s_prefix;
  const uint32_t r = s_r;
  const uint32_t K = s_K;
  const int D = d.dim;
  EvalItem *list = L.lists[d.kind] + s_listbase;
  uint64_t tie_carry = 0, sel_carry = 0;
  for (uint32_t chunk = 0; chunk < n; chunk += 4 * kSelThreads) {
    const uint32_t i0 = chunk + 4 * tid;
    uint64_t k[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) k[u] = i0 + u < n ? err_key(err[i0 + u]) : 0;
    uint64_t nt = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) nt += (i0 + u < n && k[u] == T) ? 1 : 0;
    // ties before this thread's, then the selections it makes
    uint64_t tie_incl = nt, dummy = 0, tot_t, tot_d;
    block_scan2(tie_incl, dummy, s_wa, s_wb, tot_t, tot_d);
    uint64_t trank = tie_carry + tie_incl - nt;
    uint32_t selm = 0;
    uint64_t ns = 0;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const bool in = i0 + u < n;
      const bool tie = in && k[u] == T;
      const bool sel = in && (k[u] > T || (tie && trank < r));
      if (tie) ++trank;
      selm |= sel ? 1u << u : 0u;
      ns += sel ? 1 : 0;
    }
    uint64_t sel_incl = ns, dummy2 = 0, tot_s2, tot_d2;
    block_scan2(sel_incl, dummy2, s_wa, s_wb, tot_s2, tot_d2);
    uint32_t at = (uint32_t)(sel_incl - ns);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (selm >> u & 1u) s_selp[at++] = i0 + u;
    __syncthreads();
    for (uint32_t j = tid; j < (uint32_t)tot_s2; j += kSelThreads) {
      const uint32_t p = d.base + s_selp[j];
      const uint64_t a = sel_carry + j;
      const uint32_t q = d.base + n + (uint32_t)a;
      const int sd = L.split[p];
      double c[kMaxDim], h[kMaxDim], hd = 0;
#pragma unroll
      for (int e = 0; e < kMaxDim; ++e) {
        if (e < D) {
          c[e] = L.geo[(kFieldC + e) * S + p];
          h[e] = L.geo[(kFieldH + e) * S + p];
        }
      }
#pragma unroll
      for (int e = 0; e < kMaxDim; ++e)
        if (e < D && e == sd) hd = h[e] * 0.5;
      const double vol = L.geo[kFieldVol * S + p];
      double vq = 1;  // box_volume of the child, the same products in order
#pragma unroll
      for (int e = 0; e < kMaxDim; ++e) {
        if (e >= D) continue;
        const double he = (e == sd) ? hd : h[e];
        vq *= 2 * he;
        L.geo[(kFieldC + e) * S + q] = (e == sd) ? c[e] + hd : c[e];
        L.geo[(kFieldH + e) * S + q] = he;
        if (e == sd) {
          L.geo[(kFieldC + e) * S + p] = c[e] - hd;
          L.geo[(kFieldH + e) * S + p] = hd;
        }
      }
      L.geo[kFieldVol * S + p] = vol * 0.5;
      L.geo[kFieldVol * S + q] = vq;
      list[a] = EvalItem{p, (uint32_t)J};
      list[K + a] = EvalItem{q, (uint32_t)J};
    }
    tie_carry += tot_t;
    sel_carry += tot_s2;
    __syncthreads();
  }
  if (tid == 0) {
    JobState st = s_st;
    if (L.trace && st.rounds < L.trace_cap) L.trace[(size_t)J * L.trace_cap + st.rounds] = 2u * s_K;
    st.n = n + s_K;
    st.neval += twoP * s_K;
    st.rounds += 1;
    L.state[J] = st;
  }
}


// ---------------------------------------------------------- point values
// The integrand at caller-given points x[npts][dim] (one thread per point),
// the device counterpart of the reference's integrand callbacks.
template <int KIND, class P>
__global__ void k_points(DevTables T, const TermArgs *args, const double *x, uint32_t npts,
                         double *f) {
  constexpr int D = kind_dim(kind_of_kv(KIND));
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  double p[D];
#pragma unroll
  for (int d = 0; d < D; ++d) p[d] = x[(size_t)i * D + d];
  f[i] = integrand<KIND, P>(T, *args, p);
}

template <class P>
static void points_kind(int kind, dim3 g, dim3 b, hipStream_t s, const DevTables &T,
                        const TermArgs *args, const double *x, uint32_t npts, double *f) {
  switch (kind) {
    case kNbar: hipLaunchKernelGGL((k_points<kNbar, P>), g, b, 0, s, T, args, x, npts, f); break;
    case kNN1: hipLaunchKernelGGL((k_points<kNN1, P>), g, b, 0, s, T, args, x, npts, f); break;
    case kNN2: hipLaunchKernelGGL((k_points<kNN2, P>), g, b, 0, s, T, args, x, npts, f); break;
    case kNN3: hipLaunchKernelGGL((k_points<kNN3, P>), g, b, 0, s, T, args, x, npts, f); break;
    case kNM1: hipLaunchKernelGGL((k_points<kNM1, P>), g, b, 0, s, T, args, x, npts, f); break;
    case kNM2: hipLaunchKernelGGL((k_points<kNM2, P>), g, b, 0, s, T, args, x, npts, f); break;
    default: hipLaunchKernelGGL((k_points<kNM3, P>), g, b, 0, s, T, args, x, npts, f); break;
  }
}

hipError_t launch_points(int kind, int prec, const DevTables &T, const TermArgs *args,
                         const double *x, uint32_t npts, double *f, hipStream_t s) {
  if (npts == 0) return hipSuccess;
  if (kind < 0 || kind >= kNumKinds) return hipErrorInvalidValue;
  dim3 g((npts + 255) / 256), b(256);
  if (prec == 0 || kind == kNbar) points_kind<P64>(kind, g, b, s, T, args, x, npts, f);
  else points_kind<PMix>(kind, g, b, s, T, args, x, npts, f);
  return hipGetLastError();
}

// ------------------------------------------------- precomputation tables
// The closed-form 2-D tables of the precomputation from its integrals, one
// thread per (z, m or k) entry: linear P(k, z) (src/Powerspectrum.cpp:200-207),
// c(m, z) = 9/(1+z) (m/m*)^-0.13 (:128-135), the Sheth-Tormen mass function
// and bias (src/HMF.cpp:7-49, with pi = 3.14159 of src/constants.h:15).
__global__ void k_precomp_tables(PrecompTablesLaunch L) {
  const int n = L.nb * L.nb;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int iz = i / L.nb, j = i - iz * L.nb;
  const PrecompParams &p = *reinterpret_cast<const PrecompParams *>(L.aux);
  const double z = L.zs[iz], m = L.ms[j];
  L.plin[i] = eh_spectrum(p, L.ks[j], L.growth[iz]);
  const double c0 = 9, alpha = 0.13;
  L.conc[i] = c0 / (1. + z) * pow(m / L.mstar[iz], -alpha);
  const double A = 0.322, q = 0.707, pp = 0.3, pi = 3.14159;
  const double sigma2 = L.sigma2[i], dsigma2 = L.dsigma2[i], delta_c = L.delta_c[iz];
  const double rho_mean = L.rho_mean[iz];
  const double nuSq = delta_c * delta_c / sigma2;
  L.hmf[i] = -rho_mean / m / sigma2 * dsigma2 * A * (1 + pow(q * nuSq, -pp)) *
             sqrt(q * nuSq / 2 / pi) * exp(-0.5 * q * nuSq);
  L.bh[i] = 1 + 1. / delta_c *
                    (q * delta_c * delta_c / sigma2 - 1 +
                     2 * pp / (1 + pow(q * delta_c * delta_c / sigma2, pp)));
}

hipError_t launch_precomp_tables(const PrecompTablesLaunch &L, hipStream_t s) {
  const int n = L.nb * L.nb;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_precomp_tables, dim3((n + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}

// ----------------------------------------------------------------- launch
template <int KV, class P>
static hipError_t launch_eval_kind(const EvalLaunch &L, hipStream_t s) {
  using Gr = KGroup<KV>;
  uint32_t blocks = (L.count + Gr::G - 1) / Gr::G;
  if (blocks == 0) return hipSuccess;
  if constexpr (kind_of_kv(KV) == kNbar) {
    hipLaunchKernelGGL((k_eval_direct<KV, P>), dim3(blocks), dim3(Gr::NT), 0, s, L);
  } else {
    hipLaunchKernelGGL((k_eval<KV, P>), dim3(blocks), dim3(Gr::NT), 0, s, L);
  }
  return hipGetLastError();
}

template <int V, class P>
static hipError_t launch_eval_variant(int kind, const EvalLaunch &L, hipStream_t s) {
  switch (kind) {
    case kNN1: return launch_eval_kind<make_kv(kNN1, V), P>(L, s);
    case kNN2: return launch_eval_kind<make_kv(kNN2, V), P>(L, s);
    case kNN3: return launch_eval_kind<make_kv(kNN3, V), P>(L, s);
    case kNM1: return launch_eval_kind<make_kv(kNM1, V), P>(L, s);
    case kNM2: return launch_eval_kind<make_kv(kNM2, V), P>(L, s);
    case kNM3: return launch_eval_kind<make_kv(kNM3, V), P>(L, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_eval(int kind, int prec, int variant, const EvalLaunch &L, hipStream_t s) {
  if (kind == kNbar) return launch_eval_kind<kNbar, P64>(L, s);
  if (variant < 0 || variant >= kNumVariants) return hipErrorInvalidValue;
  if (prec == 0) {
    if (variant == 3) return launch_eval_variant<3, P64>(kind, L, s);
    if (variant == 2) return launch_eval_variant<2, P64>(kind, L, s);
    return variant ? launch_eval_variant<1, P64>(kind, L, s) : launch_eval_variant<0, P64>(kind, L, s);
  }
  if (variant == 3) return launch_eval_variant<3, PMix>(kind, L, s);
  if (variant == 2) return launch_eval_variant<2, PMix>(kind, L, s);
  return variant ? launch_eval_variant<1, PMix>(kind, L, s) : launch_eval_variant<0, PMix>(kind, L, s);
}

int eval_num_variants() { return kNumVariants; }

// Kernel symbols and block shapes, for launches recorded in a graph.
template <int KV, class P>
static KernelRef eval_ref_kind() {
  using Gr = KGroup<KV>;
  KernelRef r;
  if constexpr (kind_of_kv(KV) == kNbar)
    r.fn = reinterpret_cast<const void *>(&k_eval_direct<KV, P>);
  else
    r.fn = reinterpret_cast<const void *>(&k_eval<KV, P>);
  r.threads = (unsigned)Gr::NT;
  r.group = (unsigned)Gr::G;
  return r;
}

template <int V, class P>
static int eval_ref_variant(int kind, KernelRef *out) {
  switch (kind) {
    case kNN1: *out = eval_ref_kind<make_kv(kNN1, V), P>(); return 0;
    case kNN2: *out = eval_ref_kind<make_kv(kNN2, V), P>(); return 0;
    case kNN3: *out = eval_ref_kind<make_kv(kNN3, V), P>(); return 0;
    case kNM1: *out = eval_ref_kind<make_kv(kNM1, V), P>(); return 0;
    case kNM2: *out = eval_ref_kind<make_kv(kNM2, V), P>(); return 0;
    case kNM3: *out = eval_ref_kind<make_kv(kNM3, V), P>(); return 0;
    default: return -1;
  }
}

int eval_kernel_ref(int kind, int prec, int variant, KernelRef *out) {
  if (kind == kNbar) {
    *out = eval_ref_kind<kNbar, P64>();
    return 0;
  }
  if (variant < 0 || variant >= kNumVariants) return -1;
  if (prec == 0) {
    if (variant == 3) return eval_ref_variant<3, P64>(kind, out);
    if (variant == 2) return eval_ref_variant<2, P64>(kind, out);
    return variant ? eval_ref_variant<1, P64>(kind, out) : eval_ref_variant<0, P64>(kind, out);
  }
  if (variant == 3) return eval_ref_variant<3, PMix>(kind, out);
  if (variant == 2) return eval_ref_variant<2, PMix>(kind, out);
  return variant ? eval_ref_variant<1, PMix>(kind, out) : eval_ref_variant<0, PMix>(kind, out);
}

KernelRef select_kernel_ref() {
  return KernelRef{reinterpret_cast<const void *>(&k_select), (unsigned)kSelThreads, 1u};
}
KernelRef zero_kernel_ref() { return KernelRef{reinterpret_cast<const void *>(&k_zero), 256u, 1u}; }

int eval_group_size(int kind, int variant) {
  const int d = kind_dim(kind);
  if (d == 1) return KGroup<kNbar>::G;
  const int k = d == 5 ? kNN1 : d == 6 ? kNN2 : kNN3;
  if (variant == 3) {
    return k == kNN1 ? KGroup<make_kv(kNN1, 3)>::G
                     : k == kNN2 ? KGroup<make_kv(kNN2, 3)>::G : KGroup<make_kv(kNN3, 3)>::G;
  }
  if (variant == 2) {
    return k == kNN1 ? KGroup<make_kv(kNN1, 2)>::G
                     : k == kNN2 ? KGroup<make_kv(kNN2, 2)>::G : KGroup<make_kv(kNN3, 2)>::G;
  }
  if (variant == 1) {
    return k == kNN1 ? KGroup<make_kv(kNN1, 1)>::G
                     : k == kNN2 ? KGroup<make_kv(kNN2, 1)>::G : KGroup<make_kv(kNN3, 1)>::G;
  }
  return k == kNN1 ? KGroup<kNN1>::G : k == kNN2 ? KGroup<kNN2>::G : KGroup<kNN3>::G;
}

hipError_t launch_roots(const RootsLaunch &L, hipStream_t s) {
  if (L.njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_roots, dim3((L.njobs + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_select(const SelectLaunch &L, uint32_t njobs, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_select, dim3(njobs), dim3(kSelThreads), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_admit(const AdmitLaunch &L, hipStream_t s) {
  if (L.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_admit, dim3((L.n + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_gather(const GatherLaunch &L, hipStream_t s) {
  if (L.n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_gather, dim3((L.n + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace g3l
