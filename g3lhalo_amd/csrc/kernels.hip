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

#pragma clang fp contract(off)

namespace g3l {

// ------------------------------------------------------------ evaluation
template <int KIND, class P>
__device__ __forceinline__ double integrand(const DevTables &T, const TermArgs &a,
                                            const double *x) {
  if constexpr (KIND == kNbar) {
    return nbar_kernel<P>(T, a, x);
  } else if constexpr (KIND == kNN1) {
    return nnmap_1h<P>(T, a, x);
  } else if constexpr (KIND == kNN2) {
    return nnmap_2h<P>(T, a, x);
  } else if constexpr (KIND == kNN3) {
    return nnmap_3h<P>(T, a, x);
  } else if constexpr (KIND == kNM1) {
    return nmapmap_1h<P>(T, a, x);
  } else if constexpr (KIND == kNM2) {
    return nmapmap_2h<P>(T, a, x);
  } else {
    return nmapmap_3h<P>(T, a, x);
  }
}

// G regions per block, one rule point per thread: G * P <= 1024, as close
// to a multiple of 64 as the point counts allow (15 * 68 = 1020,
// 93 * 11 = 1023, 149 * 6 = 894, 241 * 4 = 964).  With one point per thread
// the integrand is straight-line code: no loop for the compiler to hoist the
// FP64 constants of the special functions out of, which otherwise pins them
// in registers for the whole kernel and drops occupancy to one wave per SIMD.
template <int D>
struct Group {
  static constexpr int P = rule_points(D);
  static constexpr int G = D == 1 ? 68 : D == 5 ? 11 : D == 6 ? 6 : 4;
  static constexpr int kThreads = (G * P + 63) / 64 * 64;
  static_assert(kThreads <= 1024, "one block holds every point of its regions");
};

template <int KIND, class P>
__global__ __launch_bounds__(1024) void k_eval(EvalLaunch L) {
  constexpr int D = kind_dim(KIND);
  constexpr int NP = Group<D>::P;
  constexpr int G = Group<D>::G;
  constexpr int NT = Group<D>::kThreads;
  __shared__ double s_c[G][D];
  __shared__ double s_h[G][D];
  __shared__ double s_vol[G];
  __shared__ uint32_t s_slot[G];
  __shared__ TermArgs s_args[G];
  __shared__ double s_f[G * NP];

  const int tid = threadIdx.x;
  const uint32_t first = blockIdx.x * G;
  if (first >= L.count) return;
  const int ng = (L.count - first) < (uint32_t)G ? (int)(L.count - first) : G;
  const uint64_t S = L.stride;

  if (tid < ng) {
    EvalItem it = L.list[first + tid];
    bool ok = it.slot < S && it.job < L.njobs;
    s_slot[tid] = ok ? it.slot : 0xffffffffu;
    if (ok) {
      s_args[tid] = L.args[L.jobs[it.job].args];
      s_vol[tid] = L.geo[kFieldVol * S + it.slot];
    }
  }
  __syncthreads();
  for (int i = tid; i < ng * D; i += NT) {
    int g = i / D, d = i - g * D;
    uint32_t s = s_slot[g];
    if (s == 0xffffffffu) continue;
    s_c[g][d] = L.geo[(kFieldC + d) * S + s];
    s_h[g][d] = L.geo[(kFieldH + d) * S + s];
  }
  __syncthreads();

  const DevTables T = L.T;
  if (tid < ng * NP) {
    const int g = tid / NP, q = tid - g * NP;
    if (s_slot[g] != 0xffffffffu) {
      double x[D];
      rule_point<D>(q, s_c[g], s_h[g], x);
      s_f[tid] = integrand<KIND, P>(T, s_args[g], x);
    }
  }
  __syncthreads();

  if (tid < ng && s_slot[tid] != 0xffffffffu) {
    double val, err;
    int split = 0;
    const double *v = s_f + tid * NP;
    if constexpr (D == 1) {
      gk_reduce(v, s_h[tid][0], val, err);
    } else {
      gm_reduce<D>(v, s_vol[tid], val, err, split);
    }
    uint32_t s = s_slot[tid];
    L.geo[kFieldVal * S + s] = val;
    L.geo[kFieldErr * S + s] = err;
    L.split[s] = (uint8_t)split;
  }
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

__global__ void k_zero(uint32_t *p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0;
}

// ---------------------------------------------------------------- select
// Block-wide exclusive scan of a 0/1 predicate over 256 threads (4 waves of
// 64 lanes); returns the prefix and writes the block total.
__device__ __forceinline__ uint32_t block_scan_bit(bool pred, uint32_t *s_wave, uint32_t &total) {
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  uint64_t mask = __ballot(pred);
  uint32_t before = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
  if (lane == 0) s_wave[wave] = (uint32_t)__popcll(mask);
  __syncthreads();
  uint32_t off = 0;
  for (int w = 0; w < wave; ++w) off += s_wave[w];
  total = s_wave[0] + s_wave[1] + s_wave[2] + s_wave[3];
  __syncthreads();
  return off + before;
}

__device__ __forceinline__ uint64_t err_key(double e) {
  // errmax of src/hcubature.cpp:119-126: max(0, err), NaN -> 0
  double k = e > 0 ? e : 0.0;
  return (uint64_t)__double_as_longlong(k);
}

__global__ __launch_bounds__(256) void k_select(SelectLaunch L) {
  const int J = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ JobState s_st;
  __shared__ double s_red[256];
  __shared__ double s_red2[256];
  __shared__ uint64_t s_redk[256];
  __shared__ uint32_t s_cnt[256];
  __shared__ unsigned long long s_sum[256];
  __shared__ uint32_t s_wave[4];
  __shared__ int s_flag;
  __shared__ uint64_t s_prefix, s_mask, s_cumsum;
  __shared__ uint32_t s_cumcnt, s_lastcnt, s_r, s_K, s_listbase, s_tiecarry, s_selcarry;

  if (tid == 0) s_st = L.state[J];
  __syncthreads();
  if (s_st.flags & kJobDone) return;
  const JobDesc d = L.jobs[J];
  const uint64_t S = L.stride;
  const uint32_t n = s_st.n;
  const double *val = L.geo + kFieldVal * S + d.base;
  const double *err = L.geo + kFieldErr * S + d.base;

  // 1. totals: value with Neumaier compensation, error plain, both in a
  //    fixed order so the result does not depend on scheduling.
  double vs = 0, vc = 0, es = 0;
  uint64_t kmax = 0;
  bool bad = false;
  for (uint32_t i = tid; i < n; i += 256) {
    double v = val[i], e = err[i];
    bad |= !isfinite(v) || !isfinite(e);
    double t = vs + v;
    vc += fabs(vs) >= fabs(v) ? (vs - t) + v : (v - t) + vs;
    vs = t;
    es += e;
    uint64_t k = err_key(e);
    kmax = k > kmax ? k : kmax;
  }
  s_red[tid] = vs;
  s_red2[tid] = vc;
  s_redk[tid] = kmax;
  s_cnt[tid] = bad ? 1u : 0u;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
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
  for (int w = 128; w > 0; w >>= 1) {
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
    if (tid == 0) L.state[J] = s_st;
    return;
  }

  // 3. batch size: pops continue until the remaining error meets the
  //    tolerance or the evaluation budget is spent (src/hcubature.cpp:971-985).
  const uint64_t twoP = 2ull * (uint64_t)d.pts;
  uint64_t kcap = n;
  if (d.maxeval != 0) {
    uint64_t room = d.maxeval - s_st.neval;
    uint64_t kc = (room + twoP - 1) / twoP;
    kcap = kc < kcap ? kc : kcap;
  }
  // fixed-point error units: maxkey maps below 2^47, so a sum over 2^16
  // regions cannot overflow; integer sums are exact and order-free.
  int ex = (int)((maxkey >> 52) & 0x7ff) - 1023;
  int shift = 46 - ex;
  uint64_t Efx = 0;
  {
    unsigned long long part = 0;
    for (uint32_t i = tid; i < n; i += 256) {
      double k = __longlong_as_double((long long)err_key(err[i]));
      part += (unsigned long long)ldexp(k, shift);
    }
    s_sum[tid] = part;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) s_sum[tid] += s_sum[tid + w];
      __syncthreads();
    }
    Efx = s_sum[0];
    __syncthreads();
  }
  double thr = ldexp(fabs(V) * d.reltol, shift);
  uint64_t need;
  if (thr >= (double)Efx) need = 1;  // at least one region is popped
  else need = Efx - (uint64_t)floor(thr);
  if (need == 0) need = 1;

  // radix select over the 64-bit keys, 8 bits per level, from the top
  if (tid == 0) {
    s_prefix = 0;
    s_mask = 0;
    s_cumsum = 0;
    s_cumcnt = 0;
  }
  __syncthreads();
  for (int level = 0; level < 8; ++level) {
    const int sh = 56 - 8 * level;
    s_cnt[tid] = 0;
    s_sum[tid] = 0;
    __syncthreads();
    const uint64_t prefix = s_prefix, mask = s_mask;
    for (uint32_t i = tid; i < n; i += 256) {
      uint64_t k = err_key(err[i]);
      if ((k & mask) == prefix) {
        int dgt = (int)((k >> sh) & 0xff);
        atomicAdd(&s_cnt[dgt], 1u);
        atomicAdd(&s_sum[dgt],
                  (unsigned long long)ldexp(__longlong_as_double((long long)k), shift));
      }
    }
    __syncthreads();
    if (tid == 0) {
      uint64_t cs = s_cumsum;
      uint32_t cc = s_cumcnt;
      int found = -1;
      for (int b = 255; b >= 0; --b) {
        if (s_cnt[b] == 0) continue;
        if (cs + s_sum[b] >= need || (uint64_t)cc + s_cnt[b] >= kcap) {
          found = b;
          break;
        }
        cs += s_sum[b];
        cc += s_cnt[b];
      }
      if (found < 0) found = 0;  // not reachable: the total meets need
      s_prefix = prefix | ((uint64_t)found << sh);
      s_mask = mask | (0xffull << sh);
      s_cumsum = cs;
      s_cumcnt = cc;
      s_lastcnt = s_cnt[found];
    }
    __syncthreads();
  }
  // s_prefix is now the exact key of the last popped region; take r of the
  // regions carrying that key, lowest slot first.
  if (tid == 0) {
    uint64_t T = s_prefix;
    unsigned long long fxT = (unsigned long long)ldexp(__longlong_as_double((long long)T), shift);
    uint64_t rest = need > s_cumsum ? need - s_cumsum : 0;
    uint64_t r = fxT > 0 ? (rest + fxT - 1) / fxT : (uint64_t)s_lastcnt;
    if (r < 1) r = 1;
    if (r > s_lastcnt) r = s_lastcnt;
    if ((uint64_t)s_cumcnt + r > kcap) r = kcap - s_cumcnt;
    s_r = (uint32_t)r;
    uint32_t K = s_cumcnt + (uint32_t)r;
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
    s_tiecarry = 0;
    s_selcarry = 0;
  }
  __syncthreads();
  if (s_flag) {
    if (tid == 0) L.state[J] = s_st;
    return;
  }

  // 4. mark and cut in slot order; children go to the kind's list.
  const uint64_t T = s_prefix;
  const uint32_t r = s_r;
  const int D = d.dim;
  EvalItem *list = L.lists[d.kind] + s_listbase;
  for (uint32_t chunk = 0; chunk < n; chunk += 256) {
    uint32_t i = chunk + tid;
    uint64_t k = i < n ? err_key(err[i]) : 0;
    bool tie = i < n && k == T;
    uint32_t tot;
    uint32_t trank = s_tiecarry + block_scan_bit(tie, s_wave, tot);
    bool sel = i < n && (k > T || (tie && trank < r));
    uint32_t tot2;
    uint32_t at = s_selcarry + block_scan_bit(sel, s_wave, tot2);
    if (sel) {
      uint32_t p = d.base + i;
      uint32_t q = d.base + n + at;
      int sd = L.split[p];
      double hd = L.geo[(kFieldH + sd) * S + p] * 0.5;
      double vol = L.geo[kFieldVol * S + p];
      double hq[kMaxDim];
      for (int e = 0; e < D; ++e) {
        double c = L.geo[(kFieldC + e) * S + p];
        double h = (e == sd) ? hd : L.geo[(kFieldH + e) * S + p];
        hq[e] = h;
        L.geo[(kFieldC + e) * S + q] = (e == sd) ? c + hd : c;
        L.geo[(kFieldH + e) * S + q] = h;
        if (e == sd) {
          L.geo[(kFieldC + e) * S + p] = c - hd;
          L.geo[(kFieldH + e) * S + p] = hd;
        }
      }
      L.geo[kFieldVol * S + p] = vol * 0.5;
      L.geo[kFieldVol * S + q] = box_volume(D, hq);
      list[2 * at] = EvalItem{p, (uint32_t)J};
      list[2 * at + 1] = EvalItem{q, (uint32_t)J};
    }
    __syncthreads();
    if (tid == 0) {
      s_tiecarry += tot;
      s_selcarry += tot2;
    }
    __syncthreads();
  }
  if (tid == 0) {
    JobState st = s_st;
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
  constexpr int D = kind_dim(KIND);
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

// ----------------------------------------------------------------- launch
template <int KIND, class P>
static hipError_t launch_eval_kind(const EvalLaunch &L, hipStream_t s) {
  using Gr = Group<kind_dim(KIND)>;
  uint32_t blocks = (L.count + Gr::G - 1) / Gr::G;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL((k_eval<KIND, P>), dim3(blocks), dim3(Gr::kThreads), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_eval(int kind, int prec, const EvalLaunch &L, hipStream_t s) {
  if (prec == 0 || kind == kNbar) {
    switch (kind) {
      case kNbar: return launch_eval_kind<kNbar, P64>(L, s);
      case kNN1: return launch_eval_kind<kNN1, P64>(L, s);
      case kNN2: return launch_eval_kind<kNN2, P64>(L, s);
      case kNN3: return launch_eval_kind<kNN3, P64>(L, s);
      case kNM1: return launch_eval_kind<kNM1, P64>(L, s);
      case kNM2: return launch_eval_kind<kNM2, P64>(L, s);
      case kNM3: return launch_eval_kind<kNM3, P64>(L, s);
      default: return hipErrorInvalidValue;
    }
  }
  switch (kind) {
    case kNN1: return launch_eval_kind<kNN1, PMix>(L, s);
    case kNN2: return launch_eval_kind<kNN2, PMix>(L, s);
    case kNN3: return launch_eval_kind<kNN3, PMix>(L, s);
    case kNM1: return launch_eval_kind<kNM1, PMix>(L, s);
    case kNM2: return launch_eval_kind<kNM2, PMix>(L, s);
    case kNM3: return launch_eval_kind<kNM3, PMix>(L, s);
    default: return hipErrorInvalidValue;
  }
}

int eval_group_size(int kind) {
  switch (kind_dim(kind)) {
    case 1: return Group<1>::G;
    case 5: return Group<5>::G;
    case 6: return Group<6>::G;
    default: return Group<7>::G;
  }
}

hipError_t launch_roots(const RootsLaunch &L, hipStream_t s) {
  if (L.njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_roots, dim3((L.njobs + 255) / 256), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_select(const SelectLaunch &L, uint32_t njobs, hipStream_t s) {
  if (njobs == 0) return hipSuccess;
  hipLaunchKernelGGL(k_select, dim3(njobs), dim3(256), 0, s, L);
  return hipGetLastError();
}

hipError_t launch_zero(uint32_t *p, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_zero, dim3((n + 255) / 256), dim3(256), 0, s, p, n);
  return hipGetLastError();
}

}  // namespace g3l
