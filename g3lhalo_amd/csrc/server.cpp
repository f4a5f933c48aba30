// server.cpp -- continuous batching over the device-resident cubature
// engine.  Requests (groups of integrals) are submitted at any time; each
// step admits waiting integrals as region-pool memory and job rows free up,
// runs one refinement round (evaluation launches per integrand kind, then
// one selection launch over the integrals in flight) and retires the
// integrals that finished.  Integrals of different requests refine side by
// side in the same launches; each one's result is the one g3l_eval_batch
// gives it alone, because its refinement depends only on its own regions.
#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <vector>

#include "engine.h"

using namespace g3l;

struct g3l_server {
  g3l_ctx *c = nullptr;
  int precision = G3L_FP64;
  uint64_t slots = 0;     // region-pool capacity
  uint32_t max_jobs = 0;  // job-table rows
  // device state
  DevBuf pool, split, lists, jobs, state, args, counters, active, done, gstate;
  DevBuf st_ids, st_desc, st_args, st_boxes;
  // pinned host mirrors and staging
  HostBuf hcounters, hdone, hgstate, hactive, hs_ids, hs_desc, hs_args, hs_boxes;
  EvalItem *lists_k[kNumKinds] = {};
  // host bookkeeping
  std::map<uint64_t, uint64_t> free_seg;  // start -> length
  std::vector<uint32_t> free_ids;
  struct JobInfo {
    uint32_t req, pos;
    uint64_t base, cap;
    int32_t stat, term;
  };
  std::vector<JobInfo> info;
  std::vector<uint32_t> active_ids;
  std::vector<char> gone;
  bool active_dirty = false;
  struct Req {
    uint64_t tag;
    g3l_result *out;
    uint32_t remaining;
    uint64_t submit_round, admit_round;
  };
  std::vector<Req> reqs;
  std::vector<uint32_t> free_reqs;
  struct Pending {
    uint32_t req, pos;
    g3l_request r;
  };
  std::deque<Pending> pending;
  std::deque<g3l_server_request_info> completed;
  uint32_t fill[kNumKinds] = {};  // evaluation-list entries for the next round
  uint64_t rounds = 0;
  // the round as a graph (launch mode G3L_LAUNCH_GRAPH), built at the first
  // step and again whenever the context's launch variants change
  RoundGraph graph;
  int graph_variant[kNumKinds] = {};
  int graph_prec[kNumKinds] = {};
};

namespace {

bool seg_alloc(g3l_server *s, uint64_t len, uint64_t &base) {
  for (auto it = s->free_seg.begin(); it != s->free_seg.end(); ++it) {
    if (it->second < len) continue;
    base = it->first;
    uint64_t rest = it->second - len;
    s->free_seg.erase(it);
    if (rest) s->free_seg[base + len] = rest;
    return true;
  }
  return false;
}

void seg_free(g3l_server *s, uint64_t base, uint64_t len) {
  auto next = s->free_seg.lower_bound(base);
  if (next != s->free_seg.end() && base + len == next->first) {
    len += next->second;
    next = s->free_seg.erase(next);
  }
  if (next != s->free_seg.begin()) {
    auto prev = std::prev(next);
    if (prev->first + prev->second == base) {
      prev->second += len;
      return;
    }
  }
  s->free_seg[base] = len;
}

// Stores one integral's result; the request completes with its last one.
void finish(g3l_server *s, uint32_t req, uint32_t pos, const g3l_result &r) {
  g3l_server::Req &rq = s->reqs[req];
  rq.out[pos] = r;
  if (--rq.remaining == 0) {
    // rounds already counts the round that finished it
    s->completed.push_back(
        g3l_server_request_info{rq.tag, rq.submit_round, rq.admit_round, s->rounds});
    s->free_reqs.push_back(req);
  }
}

// Admits waiting integrals in arrival order while a job row and a pool
// segment are free; launches their root regions.
int admit(g3l_server *s, uint32_t &admitted) {
  g3l_ctx *c = s->c;
  admitted = 0;
  std::vector<uint32_t> ids;
  std::vector<JobDesc> desc;
  std::vector<TermArgs> args;
  std::vector<double> boxes;
  while (!s->pending.empty() && !s->free_ids.empty()) {
    const g3l_server::Pending &p = s->pending.front();
    if (check_request(c, p.r) != G3L_OK) {  // its HOD set was cleared meanwhile
      finish(s, p.req, p.pos, g3l_result{0, 0, 0, 0, 0, G3L_EINVAL});
      s->pending.pop_front();
      continue;
    }
    const int kind = kind_of(p.r.stat, p.r.term);
    JobDesc d{};
    d.kind = kind;
    d.dim = kind_dim(kind);
    d.pts = rule_points(d.dim);
    d.reltol = p.r.reltol;
    d.maxeval = p.r.maxeval;
    d.cap = slot_cap(d.maxeval, d.pts, 1u << 16);
    uint64_t base;
    if (!seg_alloc(s, d.cap, base)) break;  // wait for memory, keep the order
    const uint32_t id = s->free_ids.back();
    s->free_ids.pop_back();
    d.base = (uint32_t)base;
    d.args = (int)id;
    s->info[id] = g3l_server::JobInfo{p.req, p.pos, base, d.cap, p.r.stat, p.r.term};
    s->reqs[p.req].admit_round = s->rounds;
    ids.push_back(id);
    desc.push_back(d);
    args.push_back(make_args(c, p.r));
    double box[2 * kMaxDim] = {0};
    term_box(c, p.r.term, box, box + kMaxDim);
    boxes.insert(boxes.end(), box, box + 2 * kMaxDim);
    s->active_ids.push_back(id);
    s->fill[kind] += 1;
    s->pending.pop_front();
  }
  const uint32_t n = (uint32_t)ids.size();
  if (n == 0) return G3L_OK;
  s->active_dirty = true;
  admitted = n;
  int rc;
  if ((rc = s->hs_ids.ensure(n * sizeof(uint32_t))) || (rc = s->st_ids.ensure(n * sizeof(uint32_t))) ||
      (rc = s->hs_desc.ensure(n * sizeof(JobDesc))) || (rc = s->st_desc.ensure(n * sizeof(JobDesc))) ||
      (rc = s->hs_args.ensure(n * sizeof(TermArgs))) ||
      (rc = s->st_args.ensure(n * sizeof(TermArgs))) ||
      (rc = s->hs_boxes.ensure(boxes.size() * sizeof(double))) ||
      (rc = s->st_boxes.ensure(boxes.size() * sizeof(double))))
    return rc;
  hipStream_t st = c->stream;
  // the stream is idle here (every step ends with a synchronisation), so the
  // pinned staging buffers are free to overwrite
  std::memcpy(s->hs_ids.p, ids.data(), n * sizeof(uint32_t));
  std::memcpy(s->hs_desc.p, desc.data(), n * sizeof(JobDesc));
  std::memcpy(s->hs_args.p, args.data(), n * sizeof(TermArgs));
  std::memcpy(s->hs_boxes.p, boxes.data(), boxes.size() * sizeof(double));
  G3L_HIP(hipMemcpyAsync(s->st_ids.p, s->hs_ids.p, n * sizeof(uint32_t), hipMemcpyHostToDevice, st));
  G3L_HIP(hipMemcpyAsync(s->st_desc.p, s->hs_desc.p, n * sizeof(JobDesc), hipMemcpyHostToDevice, st));
  G3L_HIP(hipMemcpyAsync(s->st_args.p, s->hs_args.p, n * sizeof(TermArgs), hipMemcpyHostToDevice, st));
  G3L_HIP(hipMemcpyAsync(s->st_boxes.p, s->hs_boxes.p, boxes.size() * sizeof(double),
                         hipMemcpyHostToDevice, st));
  AdmitLaunch A{};
  A.ids = s->st_ids.as<uint32_t>();
  A.desc = s->st_desc.as<JobDesc>();
  A.args = s->st_args.as<TermArgs>();
  A.boxes = s->st_boxes.as<double>();
  A.jobs = s->jobs.as<JobDesc>();
  A.args_table = s->args.as<TermArgs>();
  A.state = s->state.as<JobState>();
  A.geo = s->pool.as<double>();
  A.split = s->split.as<uint8_t>();
  A.stride = s->slots;
  for (int k = 0; k < kNumKinds; ++k) A.lists[k] = s->lists_k[k];
  A.counters = s->counters.as<uint32_t>();
  A.n = n;
  G3L_HIP(launch_admit(A, st));
  return G3L_OK;
}

void fill_status(const g3l_server *s, uint32_t admitted, g3l_server_status *out) {
  if (!out) return;
  out->in_flight = (uint32_t)s->active_ids.size();
  out->pending = (uint32_t)s->pending.size();
  out->completed = (uint32_t)s->completed.size();
  out->admitted = admitted;
  out->rounds = s->rounds;
}

}  // namespace

extern "C" {

int g3l_server_create(g3l_ctx *c, const g3l_server_config *cfg, g3l_server **out) {
  if (!c || !out) return G3L_EINVAL;
  *out = nullptr;
  g3l_server_config d{};
  if (cfg) d = *cfg;
  if (d.precision != G3L_FP64 && d.precision != G3L_MIXED) return G3L_EINVAL;
  const uint64_t slots = d.pool_slots ? d.pool_slots : (1ull << 24);
  const uint32_t max_jobs = d.max_integrals ? d.max_integrals : 65536u;
  if (slots >= 0xffffffffull || max_jobs > (1u << 24)) return G3L_EINVAL;
  g3l_server *s = new (std::nothrow) g3l_server;
  if (!s) return G3L_ENOMEM;
  s->c = c;
  s->precision = d.precision;
  s->slots = slots;
  s->max_jobs = max_jobs;
  int rc = G3L_OK;
  if ((rc = s->pool.ensure(slots * kNumFields * sizeof(double))) ||
      (rc = s->split.ensure(slots)) ||
      (rc = s->lists.ensure(2 * slots * kNumKinds * sizeof(EvalItem))) ||
      (rc = s->jobs.ensure((size_t)max_jobs * sizeof(JobDesc))) ||
      (rc = s->state.ensure((size_t)max_jobs * sizeof(JobState))) ||
      (rc = s->args.ensure((size_t)max_jobs * sizeof(TermArgs))) ||
      (rc = s->counters.ensure(64)) || (rc = s->hcounters.ensure(64)) ||
      (rc = s->active.ensure((size_t)max_jobs * sizeof(uint32_t))) ||
      (rc = s->hactive.ensure((size_t)max_jobs * sizeof(uint32_t))) ||
      (rc = s->done.ensure((size_t)max_jobs * sizeof(uint32_t))) ||
      (rc = s->hdone.ensure((size_t)max_jobs * sizeof(uint32_t))) ||
      (rc = s->gstate.ensure((size_t)max_jobs * sizeof(JobState))) ||
      (rc = s->hgstate.ensure((size_t)max_jobs * sizeof(JobState)))) {
    g3l_server_destroy(s);
    return rc;
  }
  // a round appends at most two entries per occupied slot of the kind
  for (int k = 0; k < kNumKinds; ++k) s->lists_k[k] = s->lists.as<EvalItem>() + (size_t)2 * k * slots;
  s->free_seg[0] = slots;
  s->free_ids.resize(max_jobs);
  for (uint32_t i = 0; i < max_jobs; ++i) s->free_ids[i] = max_jobs - 1 - i;
  s->info.resize(max_jobs);
  if (launch_zero(s->counters.as<uint32_t>(), kNumKinds + 2, c->stream) != hipSuccess) {
    g3l_server_destroy(s);
    return G3L_EHIP;
  }
  *out = s;
  return G3L_OK;
}

int g3l_server_submit(g3l_server *s, const g3l_request *r, size_t n, g3l_result *out,
                      uint64_t tag) {
  if (!s || (n && (!r || !out)) || n > 0xffffffffull) return G3L_EINVAL;
  for (size_t i = 0; i < n; ++i) {
    int rc = check_request(s->c, r[i]);
    if (rc) return rc;
    const int kind = kind_of(r[i].stat, r[i].term);
    if (slot_cap(r[i].maxeval, rule_points(kind_dim(kind)), 1u << 16) > s->slots)
      return G3L_EINVAL;  // could never be admitted
  }
  if (n == 0) {
    s->completed.push_back(g3l_server_request_info{tag, s->rounds, s->rounds, s->rounds});
    return G3L_OK;
  }
  uint32_t q;
  if (!s->free_reqs.empty()) {
    q = s->free_reqs.back();
    s->free_reqs.pop_back();
  } else {
    q = (uint32_t)s->reqs.size();
    s->reqs.push_back({});
  }
  s->reqs[q] = g3l_server::Req{tag, out, (uint32_t)n, s->rounds, s->rounds};
  for (size_t i = 0; i < n; ++i) s->pending.push_back(g3l_server::Pending{q, (uint32_t)i, r[i]});
  return G3L_OK;
}

int g3l_server_step(g3l_server *s, g3l_server_status *status) {
  if (!s) return G3L_EINVAL;
  g3l_ctx *c = s->c;
  hipStream_t st = c->stream;
  uint32_t admitted = 0;
  int rc = admit(s, admitted);
  if (rc) return rc;
  if (s->active_ids.empty()) {
    fill_status(s, admitted, status);
    return G3L_OK;
  }
  uint32_t *dcnt = s->counters.as<uint32_t>();
  const uint32_t na = (uint32_t)s->active_ids.size();
  if (s->active_dirty) {
    // the stream is idle here: every step ends with a synchronisation
    std::memcpy(s->hactive.p, s->active_ids.data(), na * sizeof(uint32_t));
    G3L_HIP(hipMemcpyAsync(s->active.p, s->hactive.p, na * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st));
    s->active_dirty = false;
  }
  EvalLaunch EL{};
  EL.T = c->T;
  EL.args = s->args.as<TermArgs>();
  EL.jobs = s->jobs.as<JobDesc>();
  EL.geo = s->pool.as<double>();
  EL.split = s->split.as<uint8_t>();
  EL.stride = s->slots;
  EL.njobs = s->max_jobs;
  SelectLaunch SL{};
  SL.jobs = s->jobs.as<JobDesc>();
  SL.state = s->state.as<JobState>();
  SL.geo = s->pool.as<double>();
  SL.split = s->split.as<uint8_t>();
  SL.stride = s->slots;
  for (int k = 0; k < kNumKinds; ++k) {
    SL.lists[k] = s->lists_k[k];
    SL.list_cap[k] = (uint32_t)std::min<uint64_t>(2 * s->slots, 0xffffffffull);
  }
  SL.counters = dcnt;
  SL.active = s->active.as<uint32_t>();
  SL.done = s->done.as<uint32_t>();
  uint32_t *hc = s->hcounters.as<uint32_t>();
  const KindLaunch KL = kind_launch(c, s->precision);
  const int *var = KL.variant;
  if (c->launch_mode == G3L_LAUNCH_GRAPH) {
    if (!s->graph.exec || !std::equal(var, var + kNumKinds, s->graph_variant) ||
        !std::equal(KL.prec, KL.prec + kNumKinds, s->graph_prec)) {
      bool in_use[kNumKinds];
      for (int k = 0; k < kNumKinds; ++k) in_use[k] = k != kNbar;  // the six terms
      if ((rc = s->graph.build(in_use, KL.prec, var, EL, SL, dcnt, kNumKinds + 2, hc, c->ev,
                               c->kev)))
        return rc;
      std::copy(var, var + kNumKinds, s->graph_variant);
      std::copy(KL.prec, KL.prec + kNumKinds, s->graph_prec);
    }
    if ((rc = s->graph.launch(s->fill, s->lists_k, na, st))) return rc;
  } else {
    for (int k = 0; k < kNumKinds; ++k) {
      if (!s->fill[k]) continue;
      EL.list = s->lists_k[k];
      EL.count = s->fill[k];
      G3L_HIP(launch_eval(k, KL.prec[k], var[k], EL, st));
    }
    G3L_HIP(launch_zero(dcnt, kNumKinds + 2, st));
    G3L_HIP(launch_select(SL, na, st));
    G3L_HIP(hipMemcpyAsync(hc, dcnt, (kNumKinds + 2) * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
  }
  G3L_HIP(hipStreamSynchronize(st));
  ++s->rounds;
  for (int k = 0; k < kNumKinds; ++k) s->fill[k] = hc[k];
  const uint32_t nd = hc[kNumKinds + 1];
  if (nd) {
    GatherLaunch GL{s->done.as<uint32_t>(), s->state.as<JobState>(), s->gstate.as<JobState>(), nd};
    G3L_HIP(launch_gather(GL, st));
    G3L_HIP(hipMemcpyAsync(s->hdone.p, s->done.p, nd * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    G3L_HIP(hipMemcpyAsync(s->hgstate.p, s->gstate.p, nd * sizeof(JobState),
                           hipMemcpyDeviceToHost, st));
    G3L_HIP(hipStreamSynchronize(st));
    const uint32_t *ids = s->hdone.as<uint32_t>();
    const JobState *js = s->hgstate.as<JobState>();
    std::vector<char> &gone = s->gone;
    gone.assign(s->max_jobs, 0);
    for (uint32_t t = 0; t < nd; ++t) {
      const uint32_t id = ids[t];
      const g3l_server::JobInfo &ji = s->info[id];
      g3l_result o;
      o.raw = js[t].val;
      o.err = js[t].err;
      o.value = prefactor(c, ji.stat, ji.term, js[t].val);
      o.neval = js[t].neval;
      o.rounds = js[t].rounds + 1;
      o.status = status_of(js[t].flags);
      finish(s, ji.req, ji.pos, o);
      seg_free(s, ji.base, ji.cap);
      s->free_ids.push_back(id);
      gone[id] = 1;
    }
    s->active_ids.erase(std::remove_if(s->active_ids.begin(), s->active_ids.end(),
                                       [&](uint32_t id) { return gone[id] != 0; }),
                        s->active_ids.end());
    s->active_dirty = true;
  }
  fill_status(s, admitted, status);
  return G3L_OK;
}

int g3l_server_poll(g3l_server *s, uint64_t *tags, size_t cap, size_t *n) {
  if (!s || !n || (cap && !tags)) return G3L_EINVAL;
  size_t k = 0;
  while (k < cap && !s->completed.empty()) {
    tags[k++] = s->completed.front().tag;
    s->completed.pop_front();
  }
  *n = k;
  return G3L_OK;
}

int g3l_server_poll_info(g3l_server *s, g3l_server_request_info *info, size_t cap, size_t *n) {
  if (!s || !n || (cap && !info)) return G3L_EINVAL;
  size_t k = 0;
  while (k < cap && !s->completed.empty()) {
    info[k++] = s->completed.front();
    s->completed.pop_front();
  }
  *n = k;
  return G3L_OK;
}

int g3l_server_destroy(g3l_server *s) {
  if (!s) return G3L_EINVAL;
  if (s->c && s->c->stream) (void)hipStreamSynchronize(s->c->stream);
  for (DevBuf *b : {&s->pool, &s->split, &s->lists, &s->jobs, &s->state, &s->args, &s->counters,
                    &s->active, &s->done, &s->gstate, &s->st_ids, &s->st_desc, &s->st_args,
                    &s->st_boxes})
    b->release();
  for (HostBuf *b : {&s->hcounters, &s->hdone, &s->hgstate, &s->hactive, &s->hs_ids, &s->hs_desc,
                     &s->hs_args, &s->hs_boxes})
    b->release();
  delete s;
  return G3L_OK;
}

}  // extern "C"
