// noalloc.cpp -- g3l_eval_batch_noalloc: the batched engine over one fixed,
// file-scope device array instead of allocated buffers, for callers that
// cannot allocate (code under HIP graph capture, serving loops that forbid
// hipMalloc).  Same kernels and the same refinement as g3l_eval_batch, with
// the rounds launched on the context's stream.
//
// `scratch` is a per-device array: each GPU the library runs on holds its own
// copy, at its own address, until the process ends.  Host code never passes
// the array's name to a kernel or a copy (on the host it is only a stand-in
// for the symbol); every use goes through the device address that
// hipGetSymbolAddress returned for that device (scratch_base), resolved once
// per device with that device current, at context creation.
//
// Every call on one device shares that device's array and leaves the last
// call's contents in it.  The engine writes every region, list entry, job
// state and counter before reading it (the counters are cleared by a kernel
// at the start of every round), so nothing carries over from one call to the
// next; calls that overlap in time on one device (two threads, or two
// contexts on one device) race on the array and must be serialised by the
// caller.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <mutex>

#include "engine.h"

using namespace g3l;

namespace {

constexpr size_t kScratchBytes = G3L_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char *g_scratch_ptr[kMaxDevices];  // device address per device, null until resolved

// Jobs staged per host-to-device copy (stack buffers, no allocation).
constexpr size_t kChunk = 32;

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

namespace g3l {

// Resolves the device address of `scratch` on `device`, which must be the
// current device.  Called at context creation; idempotent.
int scratch_resolve(int device) {
  if (device < 0 || device >= kMaxDevices) return G3L_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return G3L_OK;
  void *p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    std::fprintf(stderr, "g3l: hipGetSymbolAddress(scratch) failed on device %d: %s\n", device,
                 hipGetErrorString(e));
    return G3L_EHIP;
  }
  g_scratch_ptr[device] = static_cast<unsigned char *>(p);
  return G3L_OK;
}

}  // namespace g3l

namespace {

unsigned char *scratch_base(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return g_scratch_ptr[device];
}

// The descriptor of request r with its slot base.
JobDesc desc_of(const g3l_request &r, int args_index, uint32_t base) {
  const int kind = kind_of(r.stat, r.term);
  JobDesc d{};
  d.kind = kind;
  d.dim = kind_dim(kind);
  d.pts = rule_points(d.dim);
  d.args = args_index;
  d.reltol = r.reltol;
  d.maxeval = r.maxeval;
  d.cap = slot_cap(d.maxeval, d.pts, 1u << 16);
  d.base = base;
  return d;
}

}  // namespace

extern "C" int g3l_eval_batch_noalloc(g3l_ctx *c, const g3l_request *reqs, size_t n,
                                      int32_t precision, g3l_result *out) {
  if (!c || (!reqs && n) || (!out && n)) return G3L_EINVAL;
  if (precision != G3L_FP64 && precision != G3L_MIXED) return G3L_EINVAL;
  if (n == 0) return G3L_OK;
  if (n > 0xffffffffull) return G3L_EINVAL;
  const auto t0 = std::chrono::steady_clock::now();
  unsigned char *base = scratch_base(c->device);
  if (!base) {
    // not resolved at context creation (another device was current): do it now
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != c->device) return G3L_EINVAL;
    if (int rc = scratch_resolve(c->device)) return rc;
    base = scratch_base(c->device);
  }
  // pass 1: validate, and size the slots and the per-kind lists
  EngineBufs B{};
  for (size_t i = 0; i < n; ++i) {
    if (int rc = check_request(c, reqs[i])) return rc;
    const JobDesc d = desc_of(reqs[i], 0, 0);
    B.slots += d.cap;
    B.per_kind[d.kind] += 2ull * d.cap;
  }
  uint64_t list_total = 0;
  for (int k = 0; k < kNumKinds; ++k) list_total += B.per_kind[k];
  // layout of the array: counters, descriptors, states, parameters, boxes,
  // split bytes, evaluation lists, region pool (each 256-byte aligned)
  size_t off = 0;
  const size_t o_cnt = off;
  off = align256(off + 64);
  const size_t o_jobs = off;
  off = align256(off + n * sizeof(JobDesc));
  const size_t o_state = off;
  off = align256(off + n * sizeof(JobState));
  const size_t o_args = off;
  off = align256(off + n * sizeof(TermArgs));
  const size_t o_boxes = off;
  off = align256(off + n * 2 * kMaxDim * sizeof(double));
  const size_t o_split = off;
  off = align256(off + B.slots);
  const size_t o_lists = off;
  off = align256(off + list_total * sizeof(EvalItem));
  const size_t o_pool = off;
  off = off + B.slots * kNumFields * sizeof(double);
  if (off > kScratchBytes || B.slots >= (1ull << 32) - 1) {
    std::fprintf(stderr,
                 "g3l_eval_batch_noalloc: %zu integrals need %zu bytes of device scratch "
                 "(%llu region slots); the fixed array holds %zu bytes\n",
                 n, off, (unsigned long long)B.slots, kScratchBytes);
    return G3L_ETOOBIG;
  }
  B.dcnt = reinterpret_cast<uint32_t *>(base + o_cnt);
  B.jobs = reinterpret_cast<JobDesc *>(base + o_jobs);
  B.state = reinterpret_cast<JobState *>(base + o_state);
  B.args = reinterpret_cast<TermArgs *>(base + o_args);
  B.boxes = reinterpret_cast<double *>(base + o_boxes);
  B.split = base + o_split;
  B.geo = reinterpret_cast<double *>(base + o_pool);
  {
    uint64_t lo = 0;
    for (int k = 0; k < kNumKinds; ++k) {
      B.lists[k] = reinterpret_cast<EvalItem *>(base + o_lists) + lo;
      lo += B.per_kind[k];
    }
  }
  uint32_t hcnt[16] = {0};  // round counters read back into host stack memory
  B.hcnt = hcnt;
  hipStream_t s = c->stream;
  // pass 2: descriptors (with slot bases), parameters and boxes, staged in
  // chunks through stack buffers
  {
    JobDesc dj[kChunk];
    TermArgs aj[kChunk];
    double bj[kChunk * 2 * kMaxDim];
    uint32_t slot = 0;
    for (size_t i0 = 0; i0 < n; i0 += kChunk) {
      const size_t m = std::min(kChunk, n - i0);
      for (size_t t = 0; t < m; ++t) {
        const g3l_request &r = reqs[i0 + t];
        dj[t] = desc_of(r, (int)(i0 + t), slot);
        slot += dj[t].cap;
        aj[t] = make_args(c, r);
        double *box = bj + t * 2 * kMaxDim;
        std::fill(box, box + 2 * kMaxDim, 0.0);
        term_box(c, r.term, box, box + kMaxDim);
      }
      G3L_HIP(hipMemcpyAsync(B.jobs + i0, dj, m * sizeof(JobDesc), hipMemcpyHostToDevice, s));
      G3L_HIP(hipMemcpyAsync(B.args + i0, aj, m * sizeof(TermArgs), hipMemcpyHostToDevice, s));
      G3L_HIP(hipMemcpyAsync(B.boxes + i0 * 2 * kMaxDim, bj, m * 2 * kMaxDim * sizeof(double),
                             hipMemcpyHostToDevice, s));
      G3L_HIP(hipStreamSynchronize(s));  // the stack buffers are reused
    }
  }
  if (int rc = refine_jobs(c, B, (uint32_t)n, true, precision, false)) return rc;
  // results, read back in chunks
  uint64_t evals = 0;
  uint64_t kind_evals[kNumKinds] = {0};
  {
    JobState sj[kChunk];
    for (size_t i0 = 0; i0 < n; i0 += kChunk) {
      const size_t m = std::min(kChunk, n - i0);
      G3L_HIP(hipMemcpyAsync(sj, B.state + i0, m * sizeof(JobState), hipMemcpyDeviceToHost, s));
      G3L_HIP(hipStreamSynchronize(s));
      for (size_t t = 0; t < m; ++t) {
        const g3l_request &r = reqs[i0 + t];
        const JobState &st = sj[t];
        g3l_result &o = out[i0 + t];
        o.raw = st.val;
        o.err = st.err;
        o.value = prefactor(c, r.stat, r.term, st.val);
        o.neval = st.neval;
        o.rounds = st.rounds + 1;
        o.status = status_of(st.flags);
        evals += st.neval;
        kind_evals[kind_of(r.stat, r.term)] += st.neval;
      }
    }
  }
  c->last.wall_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  c->last.evals = evals;
  for (int k = 0; k < 8; ++k) c->last.kind_evals[k] = k < kNumKinds ? kind_evals[k] : 0;
  return G3L_OK;
}
