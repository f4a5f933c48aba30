// engine.cpp -- host side of the gfx950 NNMap / NMapMap engine: context and
// tables in HBM, HOD sets with device n_bar(z), the plain host-driven
// cubature and the device-resident batched cubature, behind the C ABI of
// include/g3l.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "engine.h"

using namespace g3l;

namespace {

// Constants of the reference (src/constants.h:15-39, src/Cosmology.h:27-28).
const double kPi = 3.14159;
const double kLn10 = std::log(10.0);
const double kC = 2.99792e5;
const double kGMsun = 1.327e11;
const double kMpcKm = 3.086e19;

DevHod to_dev(const g3l_hod &h) { return DevHod{h.f, h.alpha, h.mth, h.sigma, h.mprime, h.beta}; }

}  // namespace

namespace g3l {

int kind_of(int stat, int term) { return stat == G3L_NNMAP ? term : 3 + term; }

KindLaunch kind_launch(const g3l_ctx *c, int prec) {
  KindLaunch L;
  for (int k = 0; k < kNumKinds; ++k) {
    const bool mixed = prec == G3L_MIXED && (c->mixed_fp32_kinds >> k & 1u);
    L.prec[k] = mixed ? G3L_MIXED : G3L_FP64;
    L.variant[k] = c->variant[mixed][k];
  }
  return L;
}

const double *nbar_ptr(const g3l_ctx *c, int set, int which) {
  return c->nbar_chunks[set / kSetsPerChunk].as<double>() +
         ((size_t)(set % kSetsPerChunk) * 2 + which) * c->nb;
}

// Log-space HOD constants: log(mmin/M_th), log(mmin/M'), log(mmin).
void fill_logs(const g3l_ctx *c, TermArgs &a) {
  a.lmmin = std::log(c->mmin);
  a.lth1 = std::log(c->mmin / a.h1.mth);
  a.lmp1 = std::log(c->mmin / a.h1.mprime);
  a.lth2 = std::log(c->mmin / a.h2.mth);
  a.lmp2 = std::log(c->mmin / a.h2.mprime);
}

// HOD and n_bar selection by type (src/NNMap_Model.cu:437-475); NMapMap uses
// its one HOD (slot 1).
TermArgs make_args(const g3l_ctx *c, const g3l_request &r) {
  const HodSet &s = c->sets[r.hodset];
  TermArgs a{};
  a.t1 = r.theta[0];
  a.t2 = r.theta[1];
  a.t3 = r.theta[2];
  a.A = s.A;
  a.eps = s.eps;
  if (r.stat == G3L_NMAPMAP) {
    a.h1 = a.h2 = to_dev(s.h1);
    a.nb1 = a.nb2 = nbar_ptr(c, r.hodset, 0);
    a.same = 1;
  } else {
    bool a1 = r.type1 == 1, a2 = r.type2 == 1;
    a.h1 = to_dev(a1 ? s.h1 : s.h2);
    a.nb1 = nbar_ptr(c, r.hodset, a1 ? 0 : 1);
    a.h2 = to_dev(a2 ? s.h1 : s.h2);
    a.nb2 = nbar_ptr(c, r.hodset, a2 ? 0 : 1);
    a.same = (a1 == a2) || s.identical;
  }
  fill_logs(c, a);
  return a;
}

// Integration box of a term (src/NNMap_Model.cu:313-314, 338-339, 363-364).
int term_box(const g3l_ctx *c, int term, double *lo, double *hi) {
  int d = 0;
  lo[d] = std::log10(0.1);
  hi[d++] = std::log10(1e6);
  lo[d] = std::log10(0.1);
  hi[d++] = std::log10(1e6);
  lo[d] = 0;
  hi[d++] = 2 * kPi;
  for (int i = 0; i < term; ++i) {
    lo[d] = std::log10(c->mmin_i);
    hi[d++] = std::log10(c->mmax_i);
  }
  lo[d] = c->zmin_i;
  hi[d++] = c->zmax_i;
  return d;
}

// Prefactors (src/NNMap_Model.cu:321-322, 346-347, 371-372;
// src/NMapMap_Model.cu:188-190, 212-214, 235-237).
double prefactor(const g3l_ctx *c, int stat, int term, double r) {
  const double tp = 2 * kPi;
  double v;
  if (stat == G3L_NNMAP) v = r * 3 * c->H0 * c->H0 * c->Om / 2 / kC / kC;
  else v = r * 9 * std::pow(c->H0, 4) * std::pow(c->Om, 2) / 4 / std::pow(kC, 4);
  for (int i = 0; i < 2 + term; ++i) v = v * kLn10;
  return v / tp / tp / tp;
}

int check_request(const g3l_ctx *c, const g3l_request &r) {
  if (r.stat != G3L_NNMAP && r.stat != G3L_NMAPMAP) return G3L_EINVAL;
  if (r.term < 1 || r.term > 3) return G3L_EINVAL;
  if (r.hodset < 0 || r.hodset >= (int)c->sets.size()) return G3L_EINVAL;
  if (!(r.reltol >= 0) || !std::isfinite(r.theta[0]) || !std::isfinite(r.theta[1]) ||
      !std::isfinite(r.theta[2]))
    return G3L_EINVAL;
  return G3L_OK;
}

int status_of(uint32_t flags) {
  if (flags & kJobPoolFull) return G3L_ENOMEM_POOL;
  if (flags & kJobNonfinite) return G3L_NONFINITE;
  if (flags & kJobCapped) return G3L_MAXEVAL_HIT;
  return G3L_OK;
}

// Slot capacity of one integral: with a budget, the last pop happens while
// numEval < maxEval, so n <= 1 + (maxEval + pts) / (2 pts).
uint32_t slot_cap(uint64_t maxeval, int pts, uint32_t unlimited_cap) {
  if (maxeval == 0) return unlimited_cap;
  uint64_t c = 2 + (maxeval + (uint64_t)pts) / (2ull * (uint64_t)pts);
  return (uint32_t)std::min<uint64_t>(c, 1ull << 26);
}

}  // namespace g3l

namespace g3l {

int ctx_create_bare(g3l_ctx **out) {
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return G3L_ENODEV;
  g3l_ctx *c = new (std::nothrow) g3l_ctx();
  if (!c) return G3L_ENOMEM;
  if (hipGetDevice(&c->device) != hipSuccess ||
      hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return G3L_EHIP;
  }
  bool ok = true;
  for (auto &e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
  for (auto &e : c->kev) ok = ok && hipEventCreate(&e) == hipSuccess;
  if (!ok) {
    g3l_ctx_destroy(c);
    return G3L_EHIP;
  }
  // the allocation-free entry point's array on this device (a failure is
  // reported here and again by that entry point, which retries)
  (void)scratch_resolve(c->device);
  *out = c;
  return G3L_OK;
}

// ---------------------------------------------------------------------------
// Batched engine: all jobs refine together; per round one selection launch
// and one evaluation launch per integrand kind.

int refine_jobs(g3l_ctx *c, const EngineBufs &B, uint32_t nj, bool timed, int prec,
                bool use_graph) {
  hipStream_t s = c->stream;
  uint32_t *dcnt = B.dcnt;
  uint32_t *hcnt = B.hcnt;
  EvalItem *lists[kNumKinds];
  for (int k = 0; k < kNumKinds; ++k) lists[k] = B.lists[k];
  const uint64_t *per_kind = B.per_kind;
  const uint64_t slots = B.slots;
  int rc;

  RootsLaunch R{};
  R.jobs = B.jobs;
  R.boxes = B.boxes;
  R.state = B.state;
  R.geo = B.geo;
  R.split = B.split;
  R.stride = slots;
  for (int k = 0; k < kNumKinds; ++k) R.lists[k] = lists[k];
  R.counters = dcnt;
  R.njobs = nj;
  G3L_HIP(launch_zero(dcnt, kNumKinds + 1, s));
  G3L_HIP(launch_roots(R, s));

  EvalLaunch EL{};
  EL.T = c->T;
  EL.args = B.args;
  EL.jobs = B.jobs;
  EL.geo = B.geo;
  EL.split = B.split;
  EL.stride = slots;
  EL.njobs = nj;

  SelectLaunch SL{};
  SL.jobs = B.jobs;
  SL.state = B.state;
  SL.geo = B.geo;
  SL.split = B.split;
  SL.stride = slots;
  for (int k = 0; k < kNumKinds; ++k) {
    SL.lists[k] = lists[k];
    SL.list_cap[k] = (uint32_t)std::min<uint64_t>(per_kind[k], 0xffffffffull);
  }
  SL.counters = dcnt;
  SL.trace = B.trace;
  SL.trace_cap = B.trace ? B.trace_cap : 0;

  double eval_ms = 0, select_ms = 0;
  double kind_ms[kNumKinds] = {0};
  uint64_t kind_launches[kNumKinds] = {0};
  bool kind_used[kNumKinds];
  uint64_t launches = 0;
  uint32_t rounds = 0;
  G3L_HIP(hipMemcpyAsync(hcnt, dcnt, (kNumKinds + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
  G3L_HIP(hipStreamSynchronize(s));
  // graph mode: the round's launches and the counter copy are one graph
  // (every node records its events; they are read only when timed)
  RoundGraph &graph = c->graph;
  const KindLaunch KL = kind_launch(c, prec);
  if (use_graph) {
    // the executable graph is kept while the kinds, variants, precision,
    // job count and buffers stay the same (repeated calls of one workload)
    bool in_run[kNumKinds];
    std::vector<uint64_t> key = {nj, slots, (uint64_t)(uintptr_t)B.geo,
                                 (uint64_t)(uintptr_t)lists[0], (uint64_t)(uintptr_t)B.jobs,
                                 (uint64_t)(uintptr_t)B.state, (uint64_t)(uintptr_t)B.args,
                                 (uint64_t)(uintptr_t)B.split, (uint64_t)(uintptr_t)dcnt,
                                 (uint64_t)(uintptr_t)hcnt, (uint64_t)(uintptr_t)SL.trace,
                                 SL.trace_cap};
    for (int k = 0; k < kNumKinds; ++k) {
      in_run[k] = per_kind[k] != 0;
      key.push_back(per_kind[k]);
      key.push_back((uint64_t)KL.prec[k]);
      key.push_back((uint64_t)KL.variant[k]);
    }
    if (!graph.exec || key != c->graph_key) {
      c->graph_key.clear();
      if ((rc = graph.build(in_run, KL.prec, KL.variant, EL, SL, dcnt, kNumKinds + 1, hcnt, c->ev,
                            c->kev)))
        return rc;
      c->graph_key = key;
    }
  }
  for (;;) {
    uint32_t cnt[kNumKinds];
    for (int k = 0; k < kNumKinds; ++k) {
      cnt[k] = (uint32_t)std::min<uint64_t>(hcnt[k], per_kind[k]);
      kind_used[k] = cnt[k] != 0;
      if (!kind_used[k]) continue;
      ++launches;
      ++kind_launches[k];
    }
    if (use_graph) {
      if ((rc = graph.launch(cnt, lists, nj, s))) return rc;
    } else {
      if (timed) G3L_HIP(hipEventRecord(c->ev[0], s));
      for (int k = 0; k < kNumKinds; ++k) {
        if (!kind_used[k]) continue;
        EL.list = lists[k];
        EL.count = cnt[k];
        if (timed) G3L_HIP(hipEventRecord(c->kev[2 * k], s));
        G3L_HIP(launch_eval(k, KL.prec[k], KL.variant[k], EL, s));
        if (timed) G3L_HIP(hipEventRecord(c->kev[2 * k + 1], s));
      }
      if (timed) G3L_HIP(hipEventRecord(c->ev[1], s));
      G3L_HIP(launch_zero(dcnt, kNumKinds + 1, s));
      G3L_HIP(launch_select(SL, nj, s));
      if (timed) G3L_HIP(hipEventRecord(c->ev[2], s));
      G3L_HIP(hipMemcpyAsync(hcnt, dcnt, (kNumKinds + 1) * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             s));
    }
    G3L_HIP(hipStreamSynchronize(s));
    if (timed) {
      float a = 0, b = 0;
      G3L_HIP(hipEventElapsedTime(&a, c->ev[0], c->ev[1]));
      G3L_HIP(hipEventElapsedTime(&b, c->ev[1], c->ev[2]));
      eval_ms += a;
      select_ms += b;
      for (int k = 0; k < kNumKinds; ++k) {
        if (!kind_used[k]) continue;
        float m = 0;
        G3L_HIP(hipEventElapsedTime(&m, c->kev[2 * k], c->kev[2 * k + 1]));
        kind_ms[k] += m;
      }
    }
    ++rounds;
    if (hcnt[kNumKinds] == 0) break;
  }
  c->last.eval_ms = eval_ms;
  c->last.select_ms = select_ms;
  c->last.eval_launches = launches;
  c->last.rounds = rounds;
  for (int k = 0; k < 8; ++k) {
    c->last.kind_ms[k] = k < kNumKinds ? kind_ms[k] : 0;
    c->last.kind_launches[k] = k < kNumKinds ? kind_launches[k] : 0;
  }
  return G3L_OK;
}

void finish_stats(g3l_ctx *c, const JobDesc *desc, const JobState *st, uint32_t nj,
                  double wall_ms) {
  uint64_t evals = 0;
  uint64_t kind_evals[kNumKinds] = {0};
  for (uint32_t j = 0; j < nj; ++j) {
    evals += st[j].neval;
    kind_evals[desc[j].kind] += st[j].neval;
  }
  c->last.wall_ms = wall_ms;
  c->last.evals = evals;
  for (int k = 0; k < 8; ++k) c->last.kind_evals[k] = k < kNumKinds ? kind_evals[k] : 0;
}

int run_engine(g3l_ctx *c, EngineJobs &E, std::vector<JobState> &out, bool timed, int prec,
               bool trace) {
  const uint32_t nj = (uint32_t)E.desc.size();
  out.assign(nj, JobState{});
  if (nj == 0) return G3L_OK;
  auto t0 = std::chrono::steady_clock::now();
  EngineBufs B{};
  for (auto &d : E.desc) {
    d.base = (uint32_t)B.slots;
    B.slots += d.cap;
    B.per_kind[d.kind] += 2ull * d.cap;
  }
  if (B.slots >= (1ull << 32) - 1) return G3L_ENOMEM;
  uint64_t list_total = 0;
  for (int k = 0; k < kNumKinds; ++k) list_total += B.per_kind[k];
  int rc;
  if ((rc = c->pool.ensure(B.slots * kNumFields * sizeof(double)))) return rc;
  if ((rc = c->split.ensure(B.slots))) return rc;
  if ((rc = c->lists.ensure(list_total * sizeof(EvalItem)))) return rc;
  if ((rc = c->jobs.ensure(nj * sizeof(JobDesc)))) return rc;
  if ((rc = c->state.ensure(nj * sizeof(JobState)))) return rc;
  if ((rc = c->args.ensure(E.args.size() * sizeof(TermArgs)))) return rc;
  if ((rc = c->boxes.ensure(E.boxes.size() * sizeof(double)))) return rc;
  if ((rc = c->counters.ensure(64))) return rc;
  if ((rc = c->hcounters.ensure(64))) return rc;
  hipStream_t s = c->stream;
  if (trace && c->trace_cap > 0) {
    const size_t entries = (size_t)nj * c->trace_cap;
    if (entries > 0x7fffffffull) return G3L_EINVAL;
    if ((rc = c->trace.ensure(entries * sizeof(uint32_t)))) return rc;
    G3L_HIP(launch_zero(c->trace.as<uint32_t>(), (int)entries, s));
    B.trace = c->trace.as<uint32_t>();
    B.trace_cap = c->trace_cap;
  }
  G3L_HIP(hipMemcpyAsync(c->jobs.p, E.desc.data(), nj * sizeof(JobDesc), hipMemcpyHostToDevice, s));
  G3L_HIP(hipMemcpyAsync(c->args.p, E.args.data(), E.args.size() * sizeof(TermArgs),
                         hipMemcpyHostToDevice, s));
  G3L_HIP(hipMemcpyAsync(c->boxes.p, E.boxes.data(), E.boxes.size() * sizeof(double),
                         hipMemcpyHostToDevice, s));
  {
    uint64_t off = 0;
    for (int k = 0; k < kNumKinds; ++k) {
      B.lists[k] = c->lists.as<EvalItem>() + off;
      off += B.per_kind[k];
    }
  }
  B.geo = c->pool.as<double>();
  B.split = c->split.as<uint8_t>();
  B.jobs = c->jobs.as<JobDesc>();
  B.state = c->state.as<JobState>();
  B.args = c->args.as<TermArgs>();
  B.boxes = c->boxes.as<double>();
  B.dcnt = c->counters.as<uint32_t>();
  B.hcnt = c->hcounters.as<uint32_t>();
  if ((rc = refine_jobs(c, B, nj, timed, prec, c->launch_mode == G3L_LAUNCH_GRAPH))) return rc;
  G3L_HIP(hipMemcpyAsync(out.data(), B.state, nj * sizeof(JobState), hipMemcpyDeviceToHost, s));
  G3L_HIP(hipStreamSynchronize(s));
  if (B.trace) {
    c->trace_jobs = nj;
    c->trace_rounds.resize(nj);
    for (uint32_t j = 0; j < nj; ++j) c->trace_rounds[j] = out[j].rounds;
  }
  auto t1 = std::chrono::steady_clock::now();
  finish_stats(c, E.desc.data(), out.data(), nj,
               std::chrono::duration<double, std::milli>(t1 - t0).count());
  return G3L_OK;
}

}  // namespace g3l

namespace {

// ---------------------------------------------------------------------------
// Plain engine: the reference's driver (src/hcubature.cpp:904-1021) on the
// host, each batch of regions evaluated by one k_eval launch.
struct HRegion {
  double c[kMaxDim], h[kMaxDim];
  double vol, val, err, key;
  int split;
};

double box_volume(int D, const double *h) {
  double v = 1;
  for (int i = 0; i < D; ++i) v *= 2 * h[i];
  return v;
}

void cut(int D, HRegion &R, HRegion &R2) {
  int d = R.split;
  R2 = R;
  R.h[d] *= 0.5;
  R.vol *= 0.5;
  for (int i = 0; i < D; ++i) {
    R2.c[i] = R.c[i];
    R2.h[i] = R.h[i];
  }
  R2.vol = box_volume(D, R2.h);
  R.c[d] -= R.h[d];
  R2.c[d] += R.h[d];
}

struct HHeap {
  std::vector<HRegion> a;
  double tval = 0, terr = 0;
  void push(const HRegion &r) {
    tval += r.val;
    terr += r.err;
    size_t at = a.size();
    a.push_back(r);
    while (at) {
      size_t parent = (at - 1) / 2;
      if (r.key <= a[parent].key) break;
      a[at] = a[parent];
      at = parent;
    }
    a[at] = r;
  }
  HRegion pop() {
    HRegion top = a[0];
    size_t n = a.size() - 1;
    a[0] = a[n];
    a.pop_back();
    size_t i = 0;
    for (;;) {
      size_t child = 2 * i + 1;
      if (child >= n) break;
      size_t largest = (a[child].key <= a[i].key) ? i : child;
      ++child;
      if (child < n && a[largest].key < a[child].key) largest = child;
      if (largest == i) break;
      std::swap(a[i], a[largest]);
      i = largest;
    }
    tval -= top.val;
    terr -= top.err;
    return top;
  }
};

int plain_ensure(g3l_ctx *c, uint32_t nR) {
  if (nR <= c->pcap) return G3L_OK;
  uint32_t cap = std::max<uint32_t>(4096, c->pcap);
  while (cap < nR) cap *= 2;
  int rc;
  if ((rc = c->pgeo.ensure((size_t)cap * kNumFields * sizeof(double)))) return rc;
  if ((rc = c->psplit.ensure(cap))) return rc;
  if ((rc = c->plist.ensure((size_t)cap * sizeof(EvalItem)))) return rc;
  if ((rc = c->hgeo.ensure((size_t)cap * kNumFields * sizeof(double)))) return rc;
  if ((rc = c->hsplit.ensure(cap))) return rc;
  std::vector<EvalItem> ident(cap);
  for (uint32_t i = 0; i < cap; ++i) ident[i] = EvalItem{i, 0};
  G3L_HIP(hipMemcpy(c->plist.p, ident.data(), cap * sizeof(EvalItem), hipMemcpyHostToDevice));
  c->pcap = cap;
  return G3L_OK;
}

int plain_eval_batch(g3l_ctx *c, int kind, int D, std::vector<HRegion> &R, size_t nR) {
  int rc = plain_ensure(c, (uint32_t)nR);
  if (rc) return rc;
  const size_t cap = c->pcap;
  double *hg = c->hgeo.as<double>();
  for (size_t i = 0; i < nR; ++i) {
    for (int d = 0; d < D; ++d) {
      hg[(kFieldC + d) * cap + i] = R[i].c[d];
      hg[(kFieldH + d) * cap + i] = R[i].h[d];
    }
    hg[kFieldVol * cap + i] = R[i].vol;
  }
  hipStream_t s = c->stream;
  G3L_HIP(hipMemcpy2DAsync(c->pgeo.p, cap * sizeof(double), hg, cap * sizeof(double),
                           nR * sizeof(double), kFieldVol + 1, hipMemcpyHostToDevice, s));
  EvalLaunch EL{};
  EL.T = c->T;
  EL.args = c->parg.as<TermArgs>();
  EL.jobs = c->pjob.as<JobDesc>();
  EL.list = c->plist.as<EvalItem>();
  EL.count = (uint32_t)nR;
  EL.geo = c->pgeo.as<double>();
  EL.split = c->psplit.as<uint8_t>();
  EL.stride = cap;
  EL.njobs = 1;
  G3L_HIP(launch_eval(kind, G3L_FP64, c->variant[0][kind], EL, s));
  G3L_HIP(hipMemcpy2DAsync(hg + kFieldVal * cap, cap * sizeof(double),
                           c->pgeo.as<double>() + kFieldVal * cap, cap * sizeof(double),
                           nR * sizeof(double), 2, hipMemcpyDeviceToHost, s));
  G3L_HIP(hipMemcpyAsync(c->hsplit.p, c->psplit.p, nR, hipMemcpyDeviceToHost, s));
  G3L_HIP(hipStreamSynchronize(s));
  const uint8_t *hs = c->hsplit.as<uint8_t>();
  for (size_t i = 0; i < nR; ++i) {
    R[i].val = hg[kFieldVal * cap + i];
    R[i].err = hg[kFieldErr * cap + i];
    R[i].split = hs[i];
    R[i].key = R[i].err > 0 ? R[i].err : 0;
  }
  return G3L_OK;
}

bool converged(double val, double err, double reltol) {
  return !(err > 0 && err > std::fabs(val) * reltol);
}

int plain_cubature(g3l_ctx *c, int kind, const double *lo, const double *hi, uint64_t maxEval,
                   double reltol, double *val, double *err, uint64_t *neval, uint32_t *rounds,
                   bool *capped) {
  const int D = kind_dim(kind);
  const int np = rule_points(D);
  std::vector<HRegion> R(2);
  HRegion root{};
  for (int i = 0; i < D; ++i) {
    root.c[i] = 0.5 * (lo[i] + hi[i]);
    root.h[i] = 0.5 * (hi[i] - lo[i]);
  }
  root.vol = box_volume(D, root.h);
  R[0] = root;
  uint64_t numEval = 0;
  uint32_t nb = 0;
  HHeap heap;
  int rc = plain_eval_batch(c, kind, D, R, 1);
  if (rc) return rc;
  ++nb;
  heap.push(R[0]);
  numEval += np;
  *capped = false;
  while (numEval < maxEval || !maxEval) {
    if (converged(heap.tval, heap.terr, reltol)) break;
    size_t nR = 0;
    double eval_ = heap.tval, eerr = heap.terr;
    do {
      if (nR + 2 > R.size()) R.resize((nR + 2) * 2);
      R[nR] = heap.pop();
      eerr -= R[nR].err;
      cut(D, R[nR], R[nR + 1]);
      numEval += (uint64_t)np * 2;
      nR += 2;
      if (converged(eval_, eerr, reltol)) break;
    } while (!heap.a.empty() && (numEval < maxEval || !maxEval));
    if ((rc = plain_eval_batch(c, kind, D, R, nR))) return rc;
    ++nb;
    for (size_t i = 0; i < nR; ++i) heap.push(R[i]);
  }
  if (maxEval && numEval >= maxEval) *capped = true;
  double v = 0, e = 0;
  for (const HRegion &r : heap.a) {
    v += r.val;
    e += r.err;
  }
  *val = v;
  *err = e;
  *neval = numEval;
  *rounds = nb;
  return G3L_OK;
}

// n_bar(z_i) = ln10 * int m n(m,z) (Ncen + Nsat) dlog10 m for every grid z
// and both HODs of each new set (src/NNMap_Model.cu:244-282), all on the
// device in one batched run.
int compute_nbar(g3l_ctx *c, int first, int count) {
  const int nb = c->nb;
  const size_t need_chunks = ((size_t)first + count + kSetsPerChunk - 1) / kSetsPerChunk;
  while (c->nbar_chunks.size() < need_chunks) {
    DevBuf b;
    int rc = b.ensure((size_t)kSetsPerChunk * 2 * nb * sizeof(double));
    if (rc) return rc;
    c->nbar_chunks.push_back(b);
  }
  // run in chunks of sets to bound the pool
  const int chunk = 64;
  for (int s0 = first; s0 < first + count; s0 += chunk) {
    int s1 = std::min(first + count, s0 + chunk);
    EngineJobs E;
    for (int s = s0; s < s1; ++s) {
      for (int which = 0; which < 2; ++which) {
        for (int i = 0; i < nb; ++i) {
          JobDesc d{};
          d.kind = kNbar;
          d.dim = 1;
          d.pts = 15;
          d.args = (int)E.args.size();
          d.cap = 2048;
          d.reltol = 1e-4;
          d.maxeval = 0;
          TermArgs a{};
          a.t1 = c->zmin + i * (c->zmax - c->zmin) / nb;
          a.h1 = a.h2 = to_dev(which == 0 ? c->sets[s].h1 : c->sets[s].h2);
          fill_logs(c, a);
          E.args.push_back(a);
          E.desc.push_back(d);
          double box[2 * kMaxDim] = {0};
          box[0] = std::log10(c->mmin_i);
          box[kMaxDim] = std::log10(c->mmax_i);
          E.boxes.insert(E.boxes.end(), box, box + 2 * kMaxDim);
        }
      }
    }
    std::vector<JobState> st;
    g3l_batch_stats keep = c->last;
    int rc = run_engine(c, E, st, false, G3L_FP64);
    c->last = keep;
    if (rc) return rc;
    std::vector<double> nbv(st.size());
    for (size_t j = 0; j < st.size(); ++j) {
      if (st[j].flags & (kJobPoolFull | kJobNonfinite)) return status_of(st[j].flags);
      nbv[j] = kLn10 * st[j].val;
    }
    // copy run by run: sets s0..s1 may straddle two storage chunks
    for (int s = s0; s < s1;) {
      int e = std::min(s1, (s / kSetsPerChunk + 1) * kSetsPerChunk);
      G3L_HIP(hipMemcpy(const_cast<double *>(nbar_ptr(c, s, 0)), nbv.data() + (size_t)(s - s0) * 2 * nb,
                        (size_t)(e - s) * 2 * nb * sizeof(double), hipMemcpyHostToDevice));
      s = e;
    }
  }
  return G3L_OK;
}

}  // namespace

// ===========================================================================
// C ABI
extern "C" {

const char *g3l_version(void) { return "g3lhalo_amd 0.1 (gfx950)"; }

int g3l_ctx_create(const g3l_tables *t, g3l_ctx **out) {
  if (!t || !out || t->nbins < 2 || t->nbins > 46340 || !t->g || !t->p_lens1 || !t->w ||
      !t->dwdz || !t->hmf ||
      !t->P_lin || !t->b_h || !t->concentration)
    return G3L_EINVAL;
  if (!(t->zmax > t->zmin) || !(t->kmax > t->kmin) || !(t->mmax > t->mmin) || t->mmin <= 0 ||
      t->kmin <= 0)
    return G3L_EINVAL;
  g3l_ctx *c = nullptr;
  if (int rc = ctx_create_bare(&c)) return rc;
  const int nb = t->nbins;
  c->nb = nb;
  c->zmin = t->zmin; c->zmax = t->zmax; c->kmin = t->kmin; c->kmax = t->kmax;
  c->mmin = t->mmin; c->mmax = t->mmax;
  c->H0 = t->H0; c->Om = t->Omega_m; c->OL = t->Omega_Lambda;
  c->rho_crit = 3 * c->H0 * c->H0 * kMpcKm / 8 / kPi / kGMsun;
  // integral borders (src/NNMap_Model.cu:46-49)
  c->zmin_i = c->zmin;
  c->zmax_i = c->zmax - (c->zmax - c->zmin) / nb;
  c->mmin_i = c->mmin;
  c->mmax_i = std::exp(std::log(c->mmin) + std::log(c->mmax / c->mmin) * (nb - 1) / nb);
  // rho_bar(z) = Omega_m(z) rho_crit (src/NNMap_Model.cu:256, src/Cosmology.cpp:79-87)
  c->rho.resize(nb);
  for (int i = 0; i < nb; ++i) {
    double z = c->zmin + i * (c->zmax - c->zmin) / nb;
    double E = c->Om * (1 + z) * (1 + z) * (1 + z) + c->OL;
    c->rho[i] = c->Om * (1 + z) * (1 + z) * (1 + z) / E * c->rho_crit;
  }
  const size_t n1 = nb, n2 = (size_t)nb * nb;
  const size_t total = 6 * n1 + 4 * n2 + n1;
  if (c->tables.ensure(total * sizeof(double))) {
    g3l_ctx_destroy(c);
    return G3L_ENOMEM;
  }
  std::vector<double> h(total, 0.0);
  double *p = h.data();
  auto put = [&](const double *src, size_t n) {
    double *at = p;
    if (src) std::memcpy(p, src, n * sizeof(double));
    p += n;
    return (size_t)(at - h.data());
  };
  size_t og = put(t->g, n1), op1 = put(t->p_lens1, n1),
         op2 = put(t->p_lens2 ? t->p_lens2 : t->p_lens1, n1), ow = put(t->w, n1),
         odw = put(t->dwdz, n1), orho = put(c->rho.data(), n1), ohmf = put(t->hmf, n2),
         opl = put(t->P_lin, n2);
  // P_lin keeps nb trailing zeros (h is zero-filled; the caller's array has
  // n2 entries): the look-up has no upper guard (src/NNMap_Model.cu:826-839),
  // so the last row's overrun reads zeros.
  put(nullptr, n1);
  size_t obh = put(t->b_h, n2), oc = put(t->concentration, n2);
  if (hipMemcpy(c->tables.p, h.data(), total * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess) {
    g3l_ctx_destroy(c);
    return G3L_EHIP;
  }
  double *d = c->tables.as<double>();
  DevTables &T = c->T;
  T.nb = nb;
  T.zmin = c->zmin; T.zmax = c->zmax; T.kmin = c->kmin; T.kmax = c->kmax;
  T.mmin = c->mmin; T.mmax = c->mmax;
  T.lmr = std::log(c->mmax / c->mmin);
  T.lkr = std::log(c->kmax / c->kmin);
  T.g = d + og; T.p1 = d + op1; T.p2 = d + op2; T.w = d + ow; T.dwdz = d + odw;
  T.rho = d + orho; T.hmf = d + ohmf; T.plin = d + opl; T.bh = d + obh; T.conc = d + oc;
  *out = c;
  return G3L_OK;
}

int g3l_ctx_destroy(g3l_ctx *c) {
  if (!c) return G3L_EINVAL;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  c->tables.release();
  for (auto &b : c->nbar_chunks) b.release();
  c->nbar_chunks.clear();
  c->pool.release(); c->split.release(); c->lists.release(); c->jobs.release();
  c->state.release(); c->args.release(); c->boxes.release(); c->counters.release();
  c->hcounters.release();
  c->trace.release();
  c->pgeo.release(); c->psplit.release(); c->plist.release(); c->pjob.release();
  c->parg.release(); c->hgeo.release(); c->hsplit.release();
  for (auto &e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &e : c->kev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return G3L_OK;
}

int g3l_hodset_create_many(g3l_ctx *c, int32_t n, const g3l_hod *hod1, const g3l_hod *hod2,
                           const double *A, const double *eps, const int32_t *identical,
                           int32_t *first_id) {
  if (!c || n <= 0 || !hod1 || !first_id) return G3L_EINVAL;
  int first = (int)c->sets.size();
  for (int i = 0; i < n; ++i) {
    HodSet s;
    s.h1 = hod1[i];
    s.h2 = hod2 ? hod2[i] : hod1[i];
    s.A = A ? A[i] : 0.0;
    s.eps = eps ? eps[i] : 0.0;
    s.identical = identical ? identical[i] : 0;
    for (const g3l_hod *h : {&s.h1, &s.h2})
      if (!(h->mth > 0) || !(h->sigma > 0) || !(h->mprime > 0)) {
        c->sets.resize(first);
        return G3L_EINVAL;
      }
    c->sets.push_back(s);
  }
  int rc = compute_nbar(c, first, n);
  if (rc) {
    c->sets.resize(first);
    return rc;
  }
  *first_id = first;
  return G3L_OK;
}

int g3l_hodset_create(g3l_ctx *c, const g3l_hod *hod1, const g3l_hod *hod2, double A,
                      double epsilon, int32_t identical, int32_t *set_id) {
  return g3l_hodset_create_many(c, 1, hod1, hod2, &A, &epsilon, &identical, set_id);
}

int g3l_hodset_densities(g3l_ctx *c, int32_t set, double *rho_bar, double *n_bar1,
                         double *n_bar2) {
  if (!c || set < 0 || set >= (int)c->sets.size()) return G3L_EINVAL;
  if (rho_bar) std::memcpy(rho_bar, c->rho.data(), c->nb * sizeof(double));
  if (n_bar1) G3L_HIP(hipMemcpy(n_bar1, nbar_ptr(c, set, 0), c->nb * sizeof(double),
                                hipMemcpyDeviceToHost));
  if (n_bar2) G3L_HIP(hipMemcpy(n_bar2, nbar_ptr(c, set, 1), c->nb * sizeof(double),
                                hipMemcpyDeviceToHost));
  return G3L_OK;
}

int g3l_hodset_clear(g3l_ctx *c) {
  if (!c) return G3L_EINVAL;
  c->sets.clear();
  return G3L_OK;
}

int g3l_eval_plain(g3l_ctx *c, const g3l_request *r, g3l_result *out) {
  if (!c || !r || !out) return G3L_EINVAL;
  int rc = check_request(c, *r);
  if (rc) return rc;
  const int kind = kind_of(r->stat, r->term);
  TermArgs a = make_args(c, *r);
  JobDesc d{};
  d.kind = kind;
  d.dim = kind_dim(kind);
  d.pts = rule_points(d.dim);
  d.args = 0;
  if ((rc = c->parg.ensure(sizeof(TermArgs)))) return rc;
  if ((rc = c->pjob.ensure(sizeof(JobDesc)))) return rc;
  G3L_HIP(hipMemcpy(c->parg.p, &a, sizeof(a), hipMemcpyHostToDevice));
  G3L_HIP(hipMemcpy(c->pjob.p, &d, sizeof(d), hipMemcpyHostToDevice));
  double lo[kMaxDim], hi[kMaxDim];
  term_box(c, r->term, lo, hi);
  double val = 0, err = 0;
  uint64_t ne = 0;
  uint32_t nbt = 0;
  bool capped = false;
  rc = plain_cubature(c, kind, lo, hi, r->maxeval, r->reltol, &val, &err, &ne, &nbt, &capped);
  if (rc) return rc;
  out->raw = val;
  out->err = err;
  out->value = prefactor(c, r->stat, r->term, val);
  out->neval = ne;
  out->rounds = nbt;
  out->status = !std::isfinite(val) || !std::isfinite(err) ? G3L_NONFINITE
                : capped                                     ? G3L_MAXEVAL_HIT
                                                             : G3L_OK;
  return G3L_OK;
}

int g3l_eval_batch(g3l_ctx *c, const g3l_request *reqs, size_t n, int32_t precision,
                   g3l_result *out) {
  if (!c || (!reqs && n) || (!out && n)) return G3L_EINVAL;
  if (precision != G3L_FP64 && precision != G3L_MIXED) return G3L_EINVAL;
  EngineJobs E;
  E.desc.reserve(n);
  E.args.reserve(n);
  E.boxes.reserve(n * 2 * kMaxDim);
  for (size_t i = 0; i < n; ++i) {
    int rc = check_request(c, reqs[i]);
    if (rc) return rc;
    const int kind = kind_of(reqs[i].stat, reqs[i].term);
    JobDesc d{};
    d.kind = kind;
    d.dim = kind_dim(kind);
    d.pts = rule_points(d.dim);
    d.args = (int)i;
    d.reltol = reqs[i].reltol;
    d.maxeval = reqs[i].maxeval;
    d.cap = slot_cap(d.maxeval, d.pts, 1u << 16);
    E.desc.push_back(d);
    E.args.push_back(make_args(c, reqs[i]));
    double box[2 * kMaxDim] = {0};
    term_box(c, reqs[i].term, box, box + kMaxDim);
    E.boxes.insert(E.boxes.end(), box, box + 2 * kMaxDim);
  }
  std::vector<JobState> st;
  int rc = run_engine(c, E, st, true, precision, true);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) {
    out[i].raw = st[i].val;
    out[i].err = st[i].err;
    out[i].value = prefactor(c, reqs[i].stat, reqs[i].term, st[i].val);
    out[i].neval = st[i].neval;
    out[i].rounds = st[i].rounds + 1;
    out[i].status = status_of(st[i].flags);
  }
  return G3L_OK;
}


int g3l_integrand(g3l_ctx *c, const g3l_request *r, int32_t precision, size_t npts,
                  const double *x, double *f) {
  if (!c || !r || (npts && (!x || !f)) || npts > 0xffffffffull) return G3L_EINVAL;
  if (precision != G3L_FP64 && precision != G3L_MIXED) return G3L_EINVAL;
  int rc;
  int kind;
  TermArgs a{};
  if (r->term == 0) {
    if (r->hodset < 0 || r->hodset >= (int)c->sets.size()) return G3L_EINVAL;
    kind = kNbar;
    a.t1 = r->theta[0];
    const HodSet &s = c->sets[r->hodset];
    a.h1 = a.h2 = to_dev(r->type1 == 2 ? s.h2 : s.h1);
    fill_logs(c, a);
  } else {
    if ((rc = check_request(c, *r))) return rc;
    kind = kind_of(r->stat, r->term);
    a = make_args(c, *r);
  }
  if (npts == 0) return G3L_OK;
  const int D = kind_dim(kind);
  DevBuf dx, df, da;
  if (dx.ensure(npts * D * sizeof(double)) || df.ensure(npts * sizeof(double)) ||
      da.ensure(sizeof(TermArgs))) {
    dx.release(); df.release(); da.release();
    return G3L_ENOMEM;
  }
  rc = G3L_OK;
  if (hipMemcpy(dx.p, x, npts * D * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(da.p, &a, sizeof(a), hipMemcpyHostToDevice) != hipSuccess ||
      launch_points(kind, precision, c->T, da.as<TermArgs>(), dx.as<double>(), (uint32_t)npts,
                    df.as<double>(), c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess ||
      hipMemcpy(f, df.p, npts * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    rc = G3L_EHIP;
  dx.release(); df.release(); da.release();
  return rc;
}

int g3l_set_trace(g3l_ctx *c, uint32_t cap) {
  if (!c || cap > 4096) return G3L_EINVAL;
  c->trace_cap = cap;
  c->trace_jobs = 0;
  c->trace_rounds.clear();
  return G3L_OK;
}

int g3l_last_trace(g3l_ctx *c, size_t i, uint32_t *regions, uint32_t cap, uint32_t *len) {
  if (!c || !len || (cap && !regions)) return G3L_EINVAL;
  if (c->trace_cap == 0 || i >= c->trace_jobs) return G3L_EINVAL;
  const uint32_t rounds = c->trace_rounds[i] + 1;  // the root's evaluation included
  uint32_t n = std::min<uint32_t>(std::min<uint32_t>(rounds, c->trace_cap + 1), cap);
  if (n > 0) {
    regions[0] = 1;
    if (n > 1)
      G3L_HIP(hipMemcpy(regions + 1, c->trace.as<uint32_t>() + i * c->trace_cap,
                        (n - 1) * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  *len = n;
  return G3L_OK;
}

int g3l_last_batch_stats(g3l_ctx *c, g3l_batch_stats *out) {
  if (!c || !out) return G3L_EINVAL;
  *out = c->last;
  return G3L_OK;
}

int g3l_autotune(g3l_ctx *c, int32_t hodset, int32_t precision, g3l_tune_report *out) {
  if (!c || hodset < 0 || hodset >= (int)c->sets.size()) return G3L_EINVAL;
  if (precision != G3L_FP64 && precision != G3L_MIXED) return G3L_EINVAL;
  // calibration: 55 aperture triplets x 1h/2h/3h for each statistic, at the
  // production tolerance and budget (launch sizes like the real workloads)
  const double arcmin = 2.909e-4;
  std::vector<g3l_request> reqs[2];
  for (int st = 0; st < 2; ++st)
    for (int i = 0; i < 55; ++i) {
      const int a = i % 10, b = (i * 3) % 10, d = (i * 7) % 10;
      for (int term = 1; term <= 3; ++term) {
        g3l_request r{};
        r.stat = st;
        r.term = term;
        r.hodset = hodset;
        r.type1 = 1;
        r.type2 = 2;
        r.theta[0] = 0.5 * std::pow(64.0, a / 9.0) * arcmin;
        r.theta[1] = 0.5 * std::pow(64.0, b / 9.0) * arcmin;
        r.theta[2] = 0.5 * std::pow(64.0, d / 9.0) * arcmin;
        r.reltol = 1e-4;
        r.maxeval = 5000000;
        reqs[st].push_back(r);
      }
    }
  const int pi = precision == G3L_MIXED;
  const int nv = eval_num_variants();
  double ms[kNumKinds][8] = {};
  const g3l_batch_stats keep = c->last;
  for (int v = 0; v < nv && v < 8; ++v) {
    for (int k = 1; k < kNumKinds; ++k) c->variant[pi][k] = v;
    for (int st = 0; st < 2; ++st) {
      EngineJobs E;
      for (const g3l_request &r : reqs[st]) {
        const int kind = kind_of(r.stat, r.term);
        JobDesc d{};
        d.kind = kind;
        d.dim = kind_dim(kind);
        d.pts = rule_points(d.dim);
        d.args = (int)E.args.size();
        d.reltol = r.reltol;
        d.maxeval = r.maxeval;
        d.cap = slot_cap(d.maxeval, d.pts, 1u << 16);
        E.desc.push_back(d);
        E.args.push_back(make_args(c, r));
        double box[2 * kMaxDim] = {0};
        term_box(c, r.term, box, box + kMaxDim);
        E.boxes.insert(E.boxes.end(), box, box + 2 * kMaxDim);
      }
      std::vector<JobState> jst;
      for (int rep = 0; rep < 2; ++rep) {  // the first run warms up
        int rc = run_engine(c, E, jst, true, precision);
        if (rc) {
          c->last = keep;
          return rc;
        }
      }
      for (int k = 1; k < kNumKinds; ++k) ms[k][v] += c->last.kind_ms[k];
    }
  }
  c->last = keep;
  for (int k = 1; k < kNumKinds; ++k) {
    int best = 0;
    for (int v = 1; v < nv; ++v)
      if (ms[k][v] < ms[k][best]) best = v;
    c->variant[pi][k] = best;
  }
  if (out) {
    *out = g3l_tune_report{};
    for (int k = 0; k < kNumKinds && k < 8; ++k) {
      out->variant[k] = c->variant[pi][k];
      for (int v = 0; v < nv && v < 4; ++v) {
        out->ms[k][v] = ms[k][v];
        out->group[k][v] = eval_group_size(k, v);
      }
    }
  }
  return G3L_OK;
}

int g3l_set_launch_mode(g3l_ctx *c, int32_t mode) {
  if (!c || (mode != G3L_LAUNCH_STREAM && mode != G3L_LAUNCH_GRAPH)) return G3L_EINVAL;
  c->launch_mode = mode;
  return G3L_OK;
}

int g3l_set_mixed_policy(g3l_ctx *c, uint32_t fp32_kinds) {
  if (!c || (fp32_kinds & ~0x7eu)) return G3L_EINVAL;  // kinds 1-6 only
  c->mixed_fp32_kinds = fp32_kinds;
  return G3L_OK;
}

int g3l_get_mixed_policy(g3l_ctx *c, uint32_t *fp32_kinds) {
  if (!c || !fp32_kinds) return G3L_EINVAL;
  *fp32_kinds = c->mixed_fp32_kinds;
  return G3L_OK;
}

int g3l_set_variant(g3l_ctx *c, int32_t kind, int32_t precision, int32_t variant) {
  if (!c || kind < 1 || kind >= kNumKinds || variant < 0 || variant >= eval_num_variants())
    return G3L_EINVAL;
  if (precision != G3L_FP64 && precision != G3L_MIXED) return G3L_EINVAL;
  c->variant[precision == G3L_MIXED][kind] = variant;
  return G3L_OK;
}

}  // extern "C"
