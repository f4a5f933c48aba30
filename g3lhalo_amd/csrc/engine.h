// engine.h -- internal declarations shared by the engine and the serving
// loop: device/host buffers, the context, and the request helpers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdio>
#include <vector>

#include "../../include/g3l.h"
#include "kernels.h"
#include "layout.h"
#include "physics.h"

#define G3L_HIP(expr)                                           \
  do {                                                          \
    hipError_t e_ = (expr);                                     \
    if (e_ != hipSuccess) {                                     \
      std::fprintf(stderr, "g3l: %s failed: %s (%s:%d)\n", #expr, \
                   hipGetErrorString(e_), __FILE__, __LINE__);  \
      return G3L_EHIP;                                          \
    }                                                           \
  } while (0)

namespace g3l {

// Device buffer that only grows.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return G3L_OK;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    if (hipMalloc(&p, want) != hipSuccess) {
      p = nullptr;
      return G3L_ENOMEM;
    }
    bytes = want;
    return G3L_OK;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

// Pinned host buffer that only grows.
struct HostBuf {
  void *p = nullptr;
  size_t bytes = 0;
  int ensure(size_t want) {
    if (want <= bytes) return G3L_OK;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      return G3L_ENOMEM;
    }
    bytes = want;
    return G3L_OK;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T *as() const {
    return reinterpret_cast<T *>(p);
  }
};

struct HodSet {
  g3l_hod h1, h2;
  double A, eps;
  int identical;
};

// One refinement round recorded as an instantiated HIP graph (round_graph.cpp):
// an evaluation node per kind in use, each between two event records (kev),
// the counter clear, the selection over the integrals in flight, an event
// record, and the copy of ncounters round counters to pinned host memory.
// launch() rewrites the evaluation and selection grids and arguments in the
// executable graph, then launches it on s.
struct RoundGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool in_use[kNumKinds] = {};
  KernelRef eval_ref[kNumKinds] = {};
  hipGraphNode_t eval_node[kNumKinds] = {};
  EvalLaunch eval_args[kNumKinds] = {};
  void *eval_params[kNumKinds][1] = {};
  hipGraphNode_t select_node = nullptr;
  SelectLaunch sel_args{};
  void *sel_params[1] = {};
  uint32_t *zero_ptr = nullptr;
  int zero_n = 0;
  void *zero_params[2] = {};
  // ev[0..2]: round start, evaluations done, selection done; kev[2k], kev[2k+1]
  // bracket kind k's evaluation
  int build(const bool used[kNumKinds], const int prec[kNumKinds], const int variant[kNumKinds],
            const EvalLaunch &eval_base, const SelectLaunch &sel_base, uint32_t *dcounters,
            int ncounters, uint32_t *hcounters, const hipEvent_t *ev, const hipEvent_t *kev);
  int launch(const uint32_t count[kNumKinds], EvalItem *const lists[kNumKinds], uint32_t nselect,
             hipStream_t s);
  void destroy();
  ~RoundGraph() { destroy(); }
};

}  // namespace g3l

struct g3l_ctx {
  using DevBuf = g3l::DevBuf;
  using HostBuf = g3l::HostBuf;
  using HodSet = g3l::HodSet;
  static constexpr int kNumKinds = g3l::kNumKinds;
  int device = 0;
  hipStream_t stream = nullptr;
  int nb = 0;
  double zmin, zmax, kmin, kmax, mmin, mmax;
  double H0, Om, OL, rho_crit;
  double zmin_i, zmax_i, mmin_i, mmax_i;
  std::vector<double> rho;
  DevBuf tables;     // g p1 p2 w dwdz rho | hmf plin(+nb) bh conc
  g3l::DevTables T{};
  std::vector<HodSet> sets;
  // n_bar per HOD set, [set][2][nb], in chunks of kSetsPerChunk sets that
  // never move: integrals in flight hold pointers into them.
  std::vector<DevBuf> nbar_chunks;
  // batched engine
  DevBuf pool, split, lists, jobs, state, args, boxes, counters;
  HostBuf hcounters;
  // plain engine
  DevBuf pgeo, psplit, plist, pjob, parg;
  HostBuf hgeo, hsplit;
  uint32_t pcap = 0;
  // launch variant (block shape) per precision and integrand kind (n_bar,
  // NNMap 1h/2h/3h, NMapMap 1h/2h/3h): g3l_autotune's choices on MI355X for
  // the 550-triplet workloads, 11, 3 and 1 regions per block
  int variant[2][kNumKinds] = {{0, 0, 3, 2, 0, 3, 2}, {0, 0, 3, 2, 0, 3, 2}};
  // kinds evaluated with the FP32-bulk policy under G3L_MIXED (bit k = kind
  // k); every other kind runs FP64.  The default keeps FP32 only for the
  // NNMap 2-halo term, the one kind whose mixed values stayed within 1e-4 of
  // the FP64 oracle on every integral of the 550-triplet workloads
  // (g3l_set_mixed_policy).
  uint32_t mixed_fp32_kinds = 1u << g3l::kNN2;
  // refinement rounds of the batched engine and the server: G3L_LAUNCH_GRAPH
  // (one graph launch per round) or G3L_LAUNCH_STREAM (launches on the stream)
  int launch_mode = G3L_LAUNCH_GRAPH;
  g3l::RoundGraph graph;               // the batched engine's round graph
  std::vector<uint64_t> graph_key;     // what it was built for
  hipEvent_t ev[4]{};
  hipEvent_t kev[2 * kNumKinds]{};
  g3l_batch_stats last{};
  // batch-size trace of the last g3l_eval_batch (g3l_set_trace): per job,
  // trace_cap entries of regions per selection round, and its round count
  DevBuf trace;
  uint32_t trace_cap = 0;
  uint32_t trace_jobs = 0;
  std::vector<uint32_t> trace_rounds;
};

namespace g3l {

constexpr int kSetsPerChunk = 64;

// Jobs of one batched-engine run: descriptors, integrand parameters and
// per-job boxes lo[kMaxDim], hi[kMaxDim].
struct EngineJobs {
  std::vector<JobDesc> desc;
  std::vector<TermArgs> args;
  std::vector<double> boxes;
};
// Refines every job to convergence or its budget on c's device buffers and
// stream; out receives each job's final state (engine.cpp).  trace records
// the batch-size trace when g3l_set_trace enabled it.
int run_engine(g3l_ctx *c, EngineJobs &E, std::vector<JobState> &out, bool timed, int prec,
               bool trace = false);

// Device buffers of one batched run: region pool (slots slots, field-major),
// split bytes, one evaluation list per kind (per_kind entries each), job
// descriptors, states, integrand parameters and boxes, the round counters
// (dcnt on the device, hcnt readable by the host after a stream
// synchronisation) and an optional trace (SelectLaunch::trace).
struct EngineBufs {
  double *geo = nullptr;
  uint8_t *split = nullptr;
  EvalItem *lists[kNumKinds] = {};
  uint64_t per_kind[kNumKinds] = {};
  uint64_t slots = 0;
  JobDesc *jobs = nullptr;
  JobState *state = nullptr;
  TermArgs *args = nullptr;
  double *boxes = nullptr;
  uint32_t *dcnt = nullptr;
  uint32_t *hcnt = nullptr;
  uint32_t *trace = nullptr;
  uint32_t trace_cap = 0;
};
// Roots and refinement rounds of nj jobs whose descriptors (with their slot
// bases), parameters and boxes are on the device in B; rounds as one graph
// launch each (use_graph) or as stream launches.  Fills the timing fields of
// c->last; the states stay on the device.
int refine_jobs(g3l_ctx *c, const EngineBufs &B, uint32_t nj, bool timed, int prec,
                bool use_graph);
// Evaluation counts and wall time of c->last from the final states.
void finish_stats(g3l_ctx *c, const JobDesc *desc, const JobState *st, uint32_t nj,
                  double wall_ms);
// A context with a stream and timing events but no tables, for engine runs
// whose integrands need none (the table precomputation).
int ctx_create_bare(g3l_ctx **out);

// Resolves the device address of the allocation-free entry point's scratch
// array on `device` (the current device); once per device (noalloc.cpp).
int scratch_resolve(int device);

// Per integrand kind, the precision policy and launch variant of a run at
// precision prec (G3L_MIXED: PMix for the kinds of mixed_fp32_kinds, P64
// for the rest).
struct KindLaunch {
  int prec[kNumKinds];
  int variant[kNumKinds];
};
KindLaunch kind_launch(const g3l_ctx *c, int prec);

int kind_of(int stat, int term);
const double *nbar_ptr(const g3l_ctx *c, int set, int which);
TermArgs make_args(const g3l_ctx *c, const g3l_request &r);
int term_box(const g3l_ctx *c, int term, double *lo, double *hi);
double prefactor(const g3l_ctx *c, int stat, int term, double r);
int check_request(const g3l_ctx *c, const g3l_request &r);
int status_of(uint32_t flags);
uint32_t slot_cap(uint64_t maxeval, int pts, uint32_t unlimited_cap);

}  // namespace g3l
