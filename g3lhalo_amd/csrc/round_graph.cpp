// round_graph.cpp -- one refinement round of the batched engine and of the
// serving loop as an instantiated HIP graph.
//
// A round is always the same chain of device work: one evaluation launch per
// integrand kind that has regions to evaluate, the clear of the round
// counters, one selection launch over the integrals in flight, and the copy
// of the counters to pinned host memory (the host needs them to size the next
// round and to stop).  The chain is built once per engine run or server, with
// an evaluation node for every kind the run can hold, and each round only
// rewrites the grid sizes and launch arguments of the kernel nodes in the
// executable graph before launching it: one graph launch per round instead
// of up to nine stream operations.  A kind without regions in a round keeps
// its node with a one-block grid over an empty list, which exits at once.
// Event-record nodes around each evaluation node time the kinds separately.
// No memset nodes: the counters are cleared by a kernel.
#include <algorithm>

#include "engine.h"

namespace g3l {

#define G3L_GR(expr)                                                        \
  do {                                                                      \
    hipError_t e_ = (expr);                                                 \
    if (e_ != hipSuccess) {                                                 \
      std::fprintf(stderr, "g3l: %s failed: %s (%s:%d)\n", #expr,           \
                   hipGetErrorString(e_), __FILE__, __LINE__);              \
      return G3L_EHIP;                                                      \
    }                                                                       \
  } while (0)

namespace {

hipKernelNodeParams kparams(const KernelRef &k, unsigned blocks, void **args) {
  hipKernelNodeParams p{};
  p.func = const_cast<void *>(k.fn);
  p.gridDim = dim3(blocks);
  p.blockDim = dim3(k.threads);
  p.sharedMemBytes = 0;
  p.kernelParams = args;
  p.extra = nullptr;
  return p;
}

}  // namespace

int RoundGraph::build(const bool used[kNumKinds], const int prec[kNumKinds],
                      const int variant[kNumKinds],
                      const EvalLaunch &eval_base, const SelectLaunch &sel_base, uint32_t *dcounters,
                      int ncounters, uint32_t *hcounters, const hipEvent_t *ev,
                      const hipEvent_t *kev) {
  destroy();
  G3L_GR(hipGraphCreate(&graph, 0));
  hipGraphNode_t last = nullptr;
  auto dep = [&](void) -> size_t { return last ? 1u : 0u; };
  G3L_GR(hipGraphAddEventRecordNode(&last, graph, nullptr, 0, ev[0]));
  for (int k = 0; k < kNumKinds; ++k) {
    in_use[k] = used[k];
    if (!used[k]) continue;
    if (eval_kernel_ref(k, prec[k], variant[k], &eval_ref[k]) != 0) return G3L_EINVAL;
    eval_args[k] = eval_base;
    eval_args[k].count = 0;
    eval_params[k][0] = &eval_args[k];
    hipGraphNode_t n;
    G3L_GR(hipGraphAddEventRecordNode(&n, graph, &last, dep(), kev[2 * k]));
    last = n;
    hipKernelNodeParams p = kparams(eval_ref[k], 1, eval_params[k]);
    G3L_GR(hipGraphAddKernelNode(&eval_node[k], graph, &last, dep(), &p));
    last = eval_node[k];
    G3L_GR(hipGraphAddEventRecordNode(&n, graph, &last, dep(), kev[2 * k + 1]));
    last = n;
  }
  {
    hipGraphNode_t n;
    G3L_GR(hipGraphAddEventRecordNode(&n, graph, &last, dep(), ev[1]));
    last = n;
  }
  zero_ptr = dcounters;
  zero_n = ncounters;
  zero_params[0] = &zero_ptr;
  zero_params[1] = &zero_n;
  {
    hipKernelNodeParams p = kparams(zero_kernel_ref(), 1, zero_params);
    hipGraphNode_t n;
    G3L_GR(hipGraphAddKernelNode(&n, graph, &last, dep(), &p));
    last = n;
  }
  sel_args = sel_base;
  sel_params[0] = &sel_args;
  {
    hipKernelNodeParams p = kparams(select_kernel_ref(), 1, sel_params);
    G3L_GR(hipGraphAddKernelNode(&select_node, graph, &last, dep(), &p));
    last = select_node;
  }
  {
    hipGraphNode_t n;
    G3L_GR(hipGraphAddEventRecordNode(&n, graph, &last, dep(), ev[2]));
    last = n;
    G3L_GR(hipGraphAddMemcpyNode1D(&n, graph, &last, dep(), hcounters, dcounters,
                                   (size_t)ncounters * sizeof(uint32_t), hipMemcpyDeviceToHost));
  }
  G3L_GR(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  return G3L_OK;
}

int RoundGraph::launch(const uint32_t count[kNumKinds], EvalItem *const lists[kNumKinds],
                       uint32_t nselect, hipStream_t s) {
  if (!exec) return G3L_EINVAL;
  for (int k = 0; k < kNumKinds; ++k) {
    if (!in_use[k]) {
      if (count[k]) return G3L_EINVAL;  // a kind the graph was not built for
      continue;
    }
    const uint64_t blocks = std::max<uint64_t>(1, (count[k] + eval_ref[k].group - 1) / eval_ref[k].group);
    if (blocks > 0x7fffffffull) return G3L_EINVAL;
    eval_args[k].count = count[k];
    eval_args[k].list = lists[k];
    hipKernelNodeParams p = kparams(eval_ref[k], (unsigned)blocks, eval_params[k]);
    G3L_GR(hipGraphExecKernelNodeSetParams(exec, eval_node[k], &p));
  }
  if (nselect == 0 || nselect > 0x7fffffffu) return G3L_EINVAL;
  hipKernelNodeParams p = kparams(select_kernel_ref(), nselect, sel_params);
  G3L_GR(hipGraphExecKernelNodeSetParams(exec, select_node, &p));
  G3L_GR(hipGraphLaunch(exec, s));
  return G3L_OK;
}

void RoundGraph::destroy() {
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  exec = nullptr;
  graph = nullptr;
  for (int k = 0; k < kNumKinds; ++k) in_use[k] = false;
}

}  // namespace g3l
