/*
 * g3l.h -- C ABI of the MI355X (gfx950) NNMap / NMapMap integration engine.
 *
 * Plain pointers and sizes only.  Every function returns a g3l_status; none
 * calls exit().  Angles are radians, masses M_sun, distances Mpc, as in the
 * reference.
 *
 * The reference has no C ABI: its C++ classes call the vendored integrator
 * hcubature_v (src/cubature.h:90-94) with integrand callbacks
 * (src/cubature.h:55-57).  Each entry point below names the reference
 * interface it replaces.  The C++ classes of include/g3lhalo/ keep the
 * reference's class API on top of these functions.
 */
#ifndef G3L_H
#define G3L_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum g3l_status {
  G3L_OK = 0,
  G3L_EINVAL = 1,        /* bad argument */
  G3L_ENOMEM = 2,        /* device or host allocation failed */
  G3L_EHIP = 3,          /* HIP runtime error */
  G3L_ENOMEM_POOL = 4,   /* an integral ran out of region slots */
  G3L_MAXEVAL_HIT = 5,   /* warning: the evaluation budget ended the integral
                            (the reference returns silently in that case) */
  G3L_NONFINITE = 6,     /* a NaN/Inf reached an integral's totals */
  G3L_ENODEV = 7,        /* no HIP device */
  G3L_ETOOBIG = 8        /* the workload exceeds the fixed scratch array of
                            g3l_eval_batch_noalloc */
} g3l_status;

typedef enum g3l_stat { G3L_NNMAP = 0, G3L_NMAPMAP = 1 } g3l_stat;

/* Precision of the integrand evaluation in g3l_eval_batch.  G3L_FP64: every
 * integrand in FP64.  G3L_MIXED: the kinds selected by g3l_set_mixed_policy
 * with the FP32-bulk integrand (FP32 u_NFW where k r_s is not small and FP32
 * aperture exponentials; FP64 small-k r_s u_NFW, HOD occupations, indices,
 * table values, products and all region sums), the others in FP64. */
typedef enum g3l_precision { G3L_FP64 = 0, G3L_MIXED = 1 } g3l_precision;

typedef struct g3l_ctx g3l_ctx;

/* Lookup tables and binning of NNMap_Model's constructor
 * (src/NNMap_Model.h:124-128, src/NMapMap_Model.h:102-106): 1-D arrays of
 * nbins values on z_i = zmin + i (zmax-zmin)/nbins, 2-D arrays z-major,
 * nbins*nbins values (2 <= nbins <= 46340, so that nbins^2 indexes an
 * int).  For NMapMap pass the one lens distribution as p_lens1.  The
 * context copies all tables into HBM. */
typedef struct g3l_tables {
  int32_t nbins;
  double zmin, zmax, kmin, kmax, mmin, mmax;
  const double *g, *p_lens1, *p_lens2, *w, *dwdz;
  const double *hmf, *P_lin, *b_h, *concentration;
  double H0, Omega_m, Omega_Lambda; /* Cosmology fields (src/Cosmology.h) */
} g3l_tables;

/* One HOD: {f, alpha, M_th, sigma, M', beta} -- the order of
 * HOD::param_arr (src/HOD.cu:15-20). */
typedef struct g3l_hod {
  double f, alpha, mth, sigma, mprime, beta;
} g3l_hod;

/* One adaptive integral of one term (the arguments of
 * NNMap_Model::NNMap_1h/2h/3h, src/NNMap_Model.h:197-219, and of
 * NMapMap_Model::NMapMap_1h/2h/3h, src/NMapMap_Model.h:127-150). */
typedef struct g3l_request {
  int32_t stat;      /* g3l_stat */
  int32_t term;      /* 1, 2, 3 */
  int32_t hodset;    /* from g3l_hodset_create */
  int32_t type1;     /* NNMap: 1 selects hod1/n_bar1, otherwise hod2/n_bar2 */
  int32_t type2;
  int32_t pad;
  double theta[3];   /* apertures [rad] */
  double reltol;     /* reference: 1e-3 (NNMap 1h, 2h; NMapMap), 1e-4 (NNMap 3h) */
  uint64_t maxeval;  /* reference: 5000000; 0 = unlimited */
} g3l_request;

typedef struct g3l_result {
  double value;      /* prefactored term, what NNMap_1h etc. return */
  double raw;        /* the integral itself */
  double err;        /* its error estimate */
  uint64_t neval;    /* integrand evaluations */
  uint32_t rounds;   /* refinement rounds (Gladwell batches) */
  int32_t status;    /* g3l_status of this integral */
} g3l_result;

/* Replaces the table/device set-up of NNMap_Model::NNMap_Model and
 * NMapMap_Model::NMapMap_Model (src/NNMap_Model.cu:15-101,
 * src/NMapMap_Model.cu:14-79). */
int g3l_ctx_create(const g3l_tables *tables, g3l_ctx **out);
int g3l_ctx_destroy(g3l_ctx *ctx);

/* Replaces NNMap_Model::updateHODs/updateDensities
 * (src/NNMap_Model.cu:237-297) and NMapMap_Model::updateHOD
 * (src/NMapMap_Model.cu:116-167): stores the HOD pair and computes rho_bar(z),
 * n_bar1(z), n_bar2(z) on the device.  identical != 0 marks hod1 and hod2 as
 * one object, which selects the same-type NsatNsat branch
 * (src/HOD.cu:81-83) for every type pair. */
int g3l_hodset_create(g3l_ctx *ctx, const g3l_hod *hod1, const g3l_hod *hod2, double A,
                      double epsilon, int32_t identical, int32_t *set_id);
/* n HOD pairs at once; all n_bar integrals refine together on the device. */
int g3l_hodset_create_many(g3l_ctx *ctx, int32_t n, const g3l_hod *hod1, const g3l_hod *hod2,
                           const double *A, const double *epsilon, const int32_t *identical,
                           int32_t *first_id);
/* Reads rho_bar, n_bar1, n_bar2 (nbins each; any pointer may be NULL). */
int g3l_hodset_densities(g3l_ctx *ctx, int32_t set_id, double *rho_bar, double *n_bar1,
                         double *n_bar2);
/* Drops all HOD sets of the context. */
int g3l_hodset_clear(g3l_ctx *ctx);

/* Plain form: one adaptive integral, host-driven exactly as hcubature_v
 * (src/hcubature.cpp:904-1021) -- the region heap on the host, every
 * Gladwell batch evaluated by one gfx950 kernel launch.  Replaces one call
 * of NNMap_Model::NNMap_1h/2h/3h or NMapMap_Model::NMapMap_1h/2h/3h. */
int g3l_eval_plain(g3l_ctx *ctx, const g3l_request *req, g3l_result *out);

/* Batched form: n independent integrals refine together from one HBM region
 * pool, selection and cuts on the device.  Replaces a loop of the calls
 * above, e.g. NNMap_Model::calculateAll (src/NNMap_Model.cu:382-415). */
int g3l_eval_batch(g3l_ctx *ctx, const g3l_request *reqs, size_t n, int32_t precision,
                   g3l_result *out);

/* Allocation-free form of g3l_eval_batch.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory during the call --
 * code under HIP graph capture, serving loops and harnesses that forbid
 * hipMalloc.  Same arguments, kernels and results as g3l_eval_batch, with
 * every device buffer of the batched engine (region pool, evaluation lists,
 * job descriptors and states, integrand parameters, round counters) in one
 * fixed array of G3L_NOALLOC_SCRATCH_BYTES per device, held until the process
 * ends.  The call makes no hipMalloc, hipFree or other allocation (the
 * context and the HOD sets are made beforehand); its refinement rounds are
 * launched on the context's stream.  A workload whose buffers exceed the
 * array (about 64 bytes per integral plus 153 bytes per region slot; an
 * integral at maxeval m holds 2 + (m + pts)/(2 pts) slots, pts = 93, 149,
 * 241 for terms 1, 2, 3) returns G3L_ETOOBIG.  Every call on one device
 * shares that device's array: calls that overlap in time on one device race
 * and must be serialised by the caller. */
#define G3L_NOALLOC_SCRATCH_BYTES ((size_t)64 << 20)
int g3l_eval_batch_noalloc(g3l_ctx *ctx, const g3l_request *reqs, size_t n, int32_t precision,
                           g3l_result *out);

/* The integrand of a request at npts points x[npts][dim] (dim 5/6/7 for
 * terms 1/2/3, coordinates log10 l1, log10 l2, phi, log10 m..., z), written
 * to f[npts].  term = 0 selects the n_bar integrand (dim 1, x = log10 m,
 * theta[0] = z, type1 = HOD slot 1 or 2 of the set), in the given
 * precision (g3l_precision; n_bar is always FP64).  Replaces the
 * reference's integrand callbacks NNM::integrand_1halo/2halo/3halo,
 * NMM::integrand_1halo/2halo/3halo and integrand_nbar
 * (src/NNMap_Model.h:336,347,446; src/NMapMap_Model.h:247,327,411;
 * src/generalFunctions.h:127). */
int g3l_integrand(g3l_ctx *ctx, const g3l_request *req, int32_t precision, size_t npts,
                  const double *x, double *f);

/* Optional timing of the last g3l_eval_batch: device time of the dominant
 * (evaluation) kernels, total kernel count and algorithmic evaluations. */
typedef struct g3l_batch_stats {
  double eval_ms;        /* summed duration of the evaluation kernels */
  double select_ms;      /* summed duration of the selection kernels */
  double wall_ms;        /* host wall time of the call */
  uint64_t eval_launches;
  uint64_t evals;        /* integrand evaluations */
  uint32_t rounds;
  uint32_t pad;
  /* per integrand kind (0 n_bar, 1-3 NNMap 1h-3h, 4-6 NMapMap 1h-3h) */
  double kind_ms[8];
  uint64_t kind_launches[8];
  uint64_t kind_evals[8];
} g3l_batch_stats;
int g3l_last_batch_stats(g3l_ctx *ctx, g3l_batch_stats *out);

/* Batch-size trace of g3l_eval_batch (a test and diagnosis hook).  The
 * reference's integrator evaluates one Gladwell batch of nR regions per
 * integrand call (src/hcubature.cpp:971-988); its call count and the nR
 * sequence record which regions were refined.  cap > 0 makes the following
 * g3l_eval_batch calls record, per integral, the regions of each of its
 * first cap selection rounds; 0 (the default) stops recording. */
int g3l_set_trace(g3l_ctx *ctx, uint32_t cap);
/* Regions evaluated per integrand call by integral i of the last traced
 * g3l_eval_batch: regions[0] = 1 (the root region), then twice the regions
 * popped by each selection; *len = min(rounds, trace cap + 1, cap) entries,
 * where rounds is that integral's g3l_result.rounds. */
int g3l_last_trace(g3l_ctx *ctx, size_t i, uint32_t *regions, uint32_t cap, uint32_t *len);

/* Block-shape tuning.  No counterpart in the reference (its CUDA build fixes
 * <<<184,256>>>, src/NNMap_Model.cu:596-631).  The evaluation kernels come
 * in launch variants (regions per block) that give identical results;
 * g3l_autotune times each on a calibration batch (55 aperture triplets x
 * 1h/2h/3h for both statistics, tol 1e-4, maxEval 5e6, HOD set `hodset`)
 * and keeps the fastest per integrand kind for `precision`.  ms[k][v] is
 * kind k's kernel time under variant v, group[k][v] its regions per block. */
typedef struct g3l_tune_report {
  int32_t variant[8];    /* chosen variant per kind (1-3 NNMap 1h-3h, 4-6 NMapMap) */
  double ms[8][4];
  int32_t group[8][4];
} g3l_tune_report;
int g3l_autotune(g3l_ctx *ctx, int32_t hodset, int32_t precision, g3l_tune_report *out);
/* Which integrand kinds G3L_MIXED evaluates with the FP32-bulk integrand:
 * bit k for kind k (1-3 NNMap 1h-3h, 4-6 NMapMap 1h-3h); the others run FP64.
 * No counterpart in the reference (FP64 throughout).  The default is the
 * NNMap 2-halo term only (bit 2): integrals that stop at the evaluation cap
 * unconverged change with any FP32 rounding in their point values (near-tied
 * regions are split in another order), and on the 550-triplet workloads only
 * that kind's FP32 values stayed within 1e-4 of the FP64 oracle for every
 * integral. */
int g3l_set_mixed_policy(g3l_ctx *ctx, uint32_t fp32_kinds);
int g3l_get_mixed_policy(g3l_ctx *ctx, uint32_t *fp32_kinds);
/* Sets one kind's variant directly (kind 1-6, variant 0-3). */
int g3l_set_variant(g3l_ctx *ctx, int32_t kind, int32_t precision, int32_t variant);

/* How the refinement rounds of g3l_eval_batch and of the servers made on
 * this context reach the device.  No counterpart in the reference (its host
 * loop issues one synchronous launch per batch, src/NNMap_Model.cu:596-631).
 * G3L_LAUNCH_GRAPH (the default): each round's evaluation, counter-clear,
 * selection and counter-copy operations are one instantiated HIP graph whose
 * kernel nodes are re-sized per round; G3L_LAUNCH_STREAM: the same kernels
 * launched one by one on the context's stream.  Results are identical. */
typedef enum g3l_launch_mode { G3L_LAUNCH_STREAM = 0, G3L_LAUNCH_GRAPH = 1 } g3l_launch_mode;
int g3l_set_launch_mode(g3l_ctx *ctx, int32_t mode);

/* ---------------------------------------------------------------------
 * Table precomputation.  Replaces doPrecomputations.x
 * (src/doPrecomputations.cpp:7-135 with src/Cosmology.cpp,
 * src/Powerspectrum.cpp, src/HMF.cpp): the nine input tables of the model
 * from a cosmology and a source redshift distribution, on the grids
 * z_i = zmin + i (zmax - zmin)/N, m_j = exp(log mmin + j log(mmax/mmin)/N),
 * k_j = exp(log kmin + j log(kmax/kmin)/N).  Every adaptive integral (the
 * comoving distances, the sigma8 normalisation, sigma^2(m, z) and
 * d sigma^2/dm on the model grid and on the 256-mass m* grid) refines on the
 * device's batched Gauss-Kronrod engine at the reference's tolerance (1e-4
 * relative); the closed-form tables are computed on the device from them.
 * Outputs are caller-allocated host arrays; returns a g3l_status.  Needs no
 * context. */
typedef struct g3l_cosmology {
  double Omega_m, Omega_Lambda, Omega_b, sigma8, H0, ns;  /* H0 in km/s/Mpc */
} g3l_cosmology;
typedef struct g3l_precompute_in {
  g3l_cosmology cosmo;
  double zmin, zmax, mmin, mmax, kmin, kmax;  /* [Msun], [1/Mpc] */
  int32_t nbins;
  const double *nz_z, *nz_n;  /* source n(z): nz_count samples (src/Function.cpp:6-37) */
  int32_t nz_count;
} g3l_precompute_in;
typedef struct g3l_precompute_out {
  double *z, *m, *k;                            /* grids [nbins] */
  double *g, *w, *dwdz;                         /* [nbins] */
  double *hmf, *P_lin, *b_h, *concentration;    /* [nbins * nbins], z-major */
} g3l_precompute_out;
int g3l_precompute(const g3l_precompute_in *in, g3l_precompute_out *out);

/* Version string of this build. */
const char *g3l_version(void);

/* ---------------------------------------------------------------------
 * Serving: continuous batching.  No counterpart in the reference, whose
 * callers (the drivers, and the fit loop of
 * src/Fitstuff - not fixed/fitHelpers.cu:194-271) issue one
 * NNMap_Model::calculateAll after another.  A server keeps a region pool
 * and a job table in HBM; requests (groups of integrals, each a g3l_request
 * as for g3l_eval_batch) may be submitted at any time, and each
 * g3l_server_step admits waiting integrals in arrival order as pool memory
 * and job rows free up, runs one refinement round over every integral in
 * flight, and retires the finished ones.  An integral's result does not
 * depend on what else is in flight: it equals g3l_eval_batch's for it.
 * One server per context at a time; calls on one server are not
 * thread-safe; HOD sets referenced by waiting integrals must not be
 * cleared. */
typedef struct g3l_server g3l_server;

typedef struct g3l_server_config {
  uint64_t pool_slots;     /* region slots shared by the integrals in flight,
                              249 bytes of HBM each; 0 = 2^24 */
  uint32_t max_integrals;  /* integrals in flight at most; 0 = 65536 */
  int32_t precision;       /* g3l_precision */
} g3l_server_config;

typedef struct g3l_server_status {
  uint32_t in_flight;   /* integrals refining */
  uint32_t pending;     /* integrals waiting for admission */
  uint32_t completed;   /* finished requests not yet polled */
  uint32_t admitted;    /* integrals admitted by this step */
  uint64_t rounds;      /* refinement rounds run so far */
} g3l_server_status;

int g3l_server_create(g3l_ctx *ctx, const g3l_server_config *cfg, g3l_server **out);
/* Queues n integrals as one request; out[0..n) receives their results, in
 * place, by the step that completes the request (out must stay valid until
 * then).  EINVAL for an integral whose slot need exceeds the pool. */
int g3l_server_submit(g3l_server *srv, const g3l_request *reqs, size_t n, g3l_result *out,
                      uint64_t tag);
/* One admission + refinement round; status may be NULL. */
int g3l_server_step(g3l_server *srv, g3l_server_status *status);
/* Tags of completed requests, in completion order (at most cap). */
int g3l_server_poll(g3l_server *srv, uint64_t *tags, size_t cap, size_t *n);
/* The same, with each request's round stamps: submit_round = rounds run
 * before it was submitted, admit_round = rounds run before the round that
 * admitted its last integral, done_round = rounds run when it completed
 * (g3l_server_status.rounds counts).  A client that notes the time of each
 * step splits the request's latency into admission wait and refinement. */
typedef struct g3l_server_request_info {
  uint64_t tag;
  uint64_t submit_round;
  uint64_t admit_round;
  uint64_t done_round;
} g3l_server_request_info;
int g3l_server_poll_info(g3l_server *srv, g3l_server_request_info *info, size_t cap, size_t *n);
int g3l_server_destroy(g3l_server *srv);

#ifdef __cplusplus
}
#endif
#endif /* G3L_H */
