// kernels.h -- launch descriptors of the gfx950 cubature kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "layout.h"
#include "physics.h"

namespace g3l {

// Fused point generation + integrand + rule reduction over an evaluation list.
struct EvalLaunch {
  DevTables T;
  const TermArgs *args;
  const JobDesc *jobs;
  const EvalItem *list;
  uint32_t count;
  double *geo;        // region pool, field f of slot s at geo[f * stride + s]
  uint8_t *split;
  uint64_t stride;
  uint32_t njobs;     // entries naming a slot >= stride or a job >= njobs are skipped
};

struct RootsLaunch {
  const JobDesc *jobs;
  const double *boxes;  // per job: lo[kMaxDim], hi[kMaxDim]
  JobState *state;
  double *geo;
  uint8_t *split;
  uint64_t stride;
  EvalItem *lists[kNumKinds];
  uint32_t *counters;   // [kNumKinds + 1]
  uint32_t njobs;
};

struct SelectLaunch {
  const JobDesc *jobs;
  JobState *state;
  double *geo;
  uint8_t *split;
  uint64_t stride;
  EvalItem *lists[kNumKinds];
  uint32_t *counters;   // per-kind list fill, [kNumKinds] = jobs still refining,
                        // [kNumKinds + 1] = entries written to done
  uint32_t list_cap[kNumKinds];
  const uint32_t *active;  // block b selects for job active[b]; null: job b
  uint32_t *done;          // receives each job that finishes this round; may be null
  // optional batch-size trace: regions cut into by job J's r-th selection
  // (2 per popped region) at trace[J * trace_cap + r], r < trace_cap; may be null
  uint32_t *trace;
  uint32_t trace_cap;
};

// New integrals joining a running engine: their descriptors and parameters
// go to table rows ids[t], their root regions to the evaluation lists.
struct AdmitLaunch {
  const uint32_t *ids;
  const JobDesc *desc;
  const TermArgs *args;
  const double *boxes;  // per new job: lo[kMaxDim], hi[kMaxDim]
  JobDesc *jobs;
  TermArgs *args_table;
  JobState *state;
  double *geo;
  uint8_t *split;
  uint64_t stride;
  EvalItem *lists[kNumKinds];
  uint32_t *counters;
  uint32_t n;
};

// Copies the states of the jobs named in ids into out (n entries).
struct GatherLaunch {
  const uint32_t *ids;
  const JobState *state;
  JobState *out;
  uint32_t n;
};

// Closed-form precomputation tables from the integrals (nb x nb, z-major).
struct PrecompTablesLaunch {
  const double *aux;                                // PrecompParams
  const double *zs, *ms, *ks;                       // grids [nb]
  const double *growth, *delta_c, *rho_mean, *mstar;  // per redshift [nb]
  const double *sigma2, *dsigma2;                   // [nb * nb]
  double *plin, *conc, *hmf, *bh;                   // [nb * nb]
  int nb;
};
hipError_t launch_precomp_tables(const PrecompTablesLaunch &L, hipStream_t s);

// A kernel symbol with its block size and regions per block, for launches
// recorded as graph nodes.
struct KernelRef {
  const void *fn;
  unsigned threads;
  unsigned group;
};
int eval_kernel_ref(int kind, int prec, int variant, KernelRef *out);
KernelRef select_kernel_ref();   // k_select(SelectLaunch), one block per job
KernelRef zero_kernel_ref();     // k_zero(uint32_t *p, int n)

// variant: block shape, 0 .. eval_num_variants() - 1 (the same values for all)
hipError_t launch_eval(int kind, int prec, int variant, const EvalLaunch &L, hipStream_t s);
int eval_num_variants();
hipError_t launch_roots(const RootsLaunch &L, hipStream_t s);
hipError_t launch_select(const SelectLaunch &L, uint32_t njobs, hipStream_t s);
hipError_t launch_zero(uint32_t *p, int n, hipStream_t s);
hipError_t launch_admit(const AdmitLaunch &L, hipStream_t s);
hipError_t launch_gather(const GatherLaunch &L, hipStream_t s);
int eval_group_size(int kind, int variant);
hipError_t launch_points(int kind, int prec, const DevTables &T, const TermArgs *args,
                         const double *x, uint32_t npts, double *f, hipStream_t s);

}  // namespace g3l
