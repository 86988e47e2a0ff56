// graph.hip — tiled runs recorded once as HIP graphs and replayed.
//
// A run of a circuit is many kernel launches (one per tile pass).  For a
// large state each pass streams gigabytes and launch cost does not matter;
// for states of up to ~22 qubits a pass takes microseconds and the host's
// launch and planning work is a large part of a run.  A run graph records the
// launches once (hipStreamBeginCapture on the state's stream) and replays
// them with one hipGraphLaunch.  Its kernels read their register programs and
// operand blobs from a device buffer the graph owns, so a circuit of the same
// structure with other angles only changes that buffer: nqpu_graph_update
// plans it on the host, checks that it launches the same kernels with the
// same arguments, and copies the new image in stream order (after every replay
// queued before it).  Another structure records the graph again.
// Nothing in a graph writes memory with hipMemsetAsync: a graph that starts
// from |0...0> has its first register pass generate that state (or, for LDS
// tile passes, a fill kernel before them).  A replay leaves the state in the
// qubit layout of the run's last pass (State::perm); a graph that does not
// start from |0...0> restores the canonical layout before its replay.
#include "runtime.h"

#include <chrono>
#include <cstring>
#include <memory>

namespace nqpu {

namespace {

struct RunGraph {
  State* s = nullptr;
  NqpuRunOptions o{};
  uint32_t gflags = 0;
  PreparedRun run;  // the plan the graph's kernels and image follow
  void* image = nullptr;  // device copy of run.image
  size_t image_bytes = 0;
  void* stage = nullptr;  // pinned staging of image uploads
  size_t stage_bytes = 0;
  hipEvent_t stage_ev = nullptr;
  bool stage_pending = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  NqpuGraphInfo info{};

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  ~RunGraph() {
    if (s) (void)hipStreamSynchronize(s->stream);
    drop_graph();
    if (image) (void)hipFree(image);
    if (stage) (void)hipHostFree(stage);
    if (stage_ev) (void)hipEventDestroy(stage_ev);
  }
};

// Same kernels with the same arguments: the recorded graph stays valid for
// the new image.  Compiled kernels are cached by their source text, so equal
// function handles mean equal program structure; the interpreting kernels
// read every op from the image.
bool same_launches(const PreparedRun& a, const PreparedRun& b) {
  if (a.prec != b.prec || a.n != b.n || a.image.size() != b.image.size() || a.launches.size() != b.launches.size())
    return false;
  for (size_t i = 0; i < a.launches.size(); ++i) {
    const RunLaunch &x = a.launches[i], &y = b.launches[i];
    if (x.kind != y.kind || x.eb != y.eb || x.blob_at != y.blob_at || x.fn != y.fn ||
        std::memcmp(&x.pa, &y.pa, sizeof(PassArgs)) != 0)
      return false;
  }
  return true;
}

NqpuError plan(RunGraph* g, const Circuit& c, PreparedRun* out) {
  const auto t0 = std::chrono::steady_clock::now();
  PlanOptions po = plan_options(g->o, g->s->prec, g->s->n);
  po.keep_perm = true;
  NqpuError e = prepare_tiled(c, g->o, g->s->prec, g->s->n, po, out);
  if (e == NQPU_OK && (g->gflags & NQPU_GRAPH_FROM_ZERO)) (void)start_from_zero(out);
  g->info.plan_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return e;
}

// Copy an image into the graph's device buffer, in stream order.
NqpuError upload(RunGraph* g, const std::vector<char>& img) {
  if (g->stage_pending) {
    HIPTRY(hipEventSynchronize(g->stage_ev), "wait for the previous graph image upload");
    g->stage_pending = false;
  }
  if (img.size() > g->stage_bytes) {
    if (g->stage) (void)hipHostFree(g->stage);
    g->stage = nullptr;
    g->stage_bytes = 0;
    HIPTRY(hipHostMalloc(&g->stage, img.size()), "pinned graph image stage");
    g->stage_bytes = img.size();
  }
  if (!g->stage_ev) HIPTRY(hipEventCreateWithFlags(&g->stage_ev, hipEventDisableTiming), "graph stage event");
  std::memcpy(g->stage, img.data(), img.size());
  HIPTRY(hipMemcpyAsync(g->image, g->stage, img.size(), hipMemcpyHostToDevice, g->s->stream), "graph image upload");
  HIPTRY(hipEventRecord(g->stage_ev, g->s->stream), "graph stage event record");
  g->stage_pending = true;
  return NQPU_OK;
}

// Record g->run's launches into a new graph (the device image is in place).
NqpuError capture(RunGraph* g) {
  State* s = g->s;
  HIPTRY(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal), "begin graph capture");
  NqpuError e = NQPU_OK;
  // from |0...0>: a fill kernel unless the first pass generates the state
  const bool fill = (g->gflags & NQPU_GRAPH_FROM_ZERO) && (g->run.launches.empty() || !g->run.launches[0].pa.from_zero);
  if (fill) {
    if (s->prec == NQPU_PRECISION_C64) Launch<float>::init_zero(s->amps, s->n, s->stream);
    else Launch<double>::init_zero(s->amps, s->n, s->stream);
    e = hip_check(hipGetLastError(), "|0...0> fill launch");
  }
  if (e == NQPU_OK) e = launch_prepared(g->run, s->amps, g->image, s->stream, nullptr);
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(s->stream, &graph);
  if (e != NQPU_OK) {
    if (graph) (void)hipGraphDestroy(graph);
    return e;
  }
  HIPTRY(end, "end graph capture");
  g->graph = graph;
  HIPTRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0), "instantiate graph");
  g->info.launches = g->run.launches.size() + (fill ? 1 : 0);
  g->info.passes = g->run.launches.size();
  return NQPU_OK;
}

// Make the graph follow a newly planned run: new device image, and the
// launches recorded again.  Waits for the replays in flight (they use the old
// buffer and modules).
NqpuError rebuild(RunGraph* g, PreparedRun&& r) {
  HIPTRY(hipStreamSynchronize(g->s->stream), "wait for graph replays");
  g->drop_graph();
  const size_t need = std::max<size_t>(r.image.size(), 256);
  if (need > g->image_bytes) {
    if (g->image) (void)hipFree(g->image);
    g->image = nullptr;
    g->image_bytes = 0;
    if (hipMalloc(&g->image, need) != hipSuccess) {
      (void)hipGetLastError();
      set_error("graph image: out of device memory");
      return NQPU_OUT_OF_DEVICE_MEMORY;
    }
    g->image_bytes = need;
  }
  g->run = std::move(r);
  NqpuError e = upload(g, g->run.image);
  if (e != NQPU_OK) return e;
  g->info.image_bytes = g->run.image.size();
  return capture(g);
}

}  // namespace

}  // namespace nqpu

using namespace nqpu;

namespace {
RunGraph* G(NqpuRunGraph* p) { return reinterpret_cast<RunGraph*>(p); }
const RunGraph* G(const NqpuRunGraph* p) { return reinterpret_cast<const RunGraph*>(p); }
const Circuit* Cg(const NqpuCircuit* p) { return reinterpret_cast<const Circuit*>(p); }
}  // namespace

extern "C" {

NqpuError nqpu_graph_create(const NqpuCircuit* circuit, NqpuState* state, const NqpuRunOptions* opts,
                            uint32_t graph_flags, NqpuRunGraph** out) {
  if (!circuit || !state || !out) return NQPU_NULL_POINTER;
  *out = nullptr;
  State* s = reinterpret_cast<State*>(state);
  const Circuit& c = *Cg(circuit);
  if (c.n > s->n || (graph_flags & ~NQPU_GRAPH_FROM_ZERO)) return NQPU_INVALID_PARAMETER;
  NqpuRunOptions o{NQPU_FUSION_TILED, NQPU_SEMANTICS_STANDARD, 0, 0, 0};
  if (opts) o = *opts;
  if (o.fusion != NQPU_FUSION_TILED || o.semantics > NQPU_SEMANTICS_REFERENCE ||
      (o.flags & (NQPU_RUN_ASYNC | NQPU_RUN_TIME_PASSES))) {
    set_error("run graphs record tiled runs (NQPU_FUSION_TILED) without NQPU_RUN_ASYNC or NQPU_RUN_TIME_PASSES");
    return NQPU_INVALID_PARAMETER;
  }
  // work queued on the state before the graph exists is waited for, so the
  // capture starts on an idle stream
  NqpuError e = synchronize_state(s);
  if (e != NQPU_OK) return e;
  std::unique_ptr<RunGraph> g(new RunGraph);
  g->s = s;
  g->o = o;
  g->gflags = graph_flags;
  PreparedRun r;
  if ((e = plan(g.get(), c, &r)) != NQPU_OK) return e;
  if ((e = rebuild(g.get(), std::move(r))) != NQPU_OK) return e;
  *out = reinterpret_cast<NqpuRunGraph*>(g.release());
  return NQPU_OK;
}

NqpuError nqpu_graph_update(NqpuRunGraph* graph, const NqpuCircuit* circuit) {
  if (!graph || !circuit) return NQPU_NULL_POINTER;
  RunGraph* g = G(graph);
  const Circuit& c = *Cg(circuit);
  if (c.n > g->s->n) return NQPU_INVALID_PARAMETER;
  PreparedRun r;
  NqpuError e = plan(g, c, &r);
  if (e != NQPU_OK) return e;
  g->info.updates++;
  if (same_launches(g->run, r)) {
    // the recorded kernels read the new numbers: one copy, ordered after the
    // replays already queued
    if ((e = upload(g, r.image)) != NQPU_OK) return e;
    g->run = std::move(r);
    return NQPU_OK;
  }
  g->info.recaptures++;
  return rebuild(g, std::move(r));
}

NqpuError nqpu_graph_launch(NqpuRunGraph* graph, uint32_t flags) {
  if (!graph) return NQPU_NULL_POINTER;
  RunGraph* g = G(graph);
  if (flags & ~NQPU_RUN_ASYNC) return NQPU_INVALID_PARAMETER;
  if (!g->exec) {
    set_error("run graph has no recorded launches (a failed update)");
    return NQPU_INVALID_PARAMETER;
  }
  // the recorded passes start from the canonical layout (or from |0...0>,
  // which is the same in every layout)
  NqpuError e = NQPU_OK;
  if (!(g->gflags & NQPU_GRAPH_FROM_ZERO) && (e = settle_state(g->s)) != NQPU_OK) return e;
  HIPTRY(hipGraphLaunch(g->exec, g->s->stream), "graph launch");
  g->s->perm = g->run.perm;
  g->s->perm_k = g->run.perm_k;
  g->info.replays++;
  g->s->stats = g->run.stats;
  if (flags & NQPU_RUN_ASYNC) return NQPU_OK;
  HIPTRY(hipStreamSynchronize(g->s->stream), "graph replay");
  return NQPU_OK;
}

NqpuError nqpu_graph_info(const NqpuRunGraph* graph, NqpuGraphInfo* out) {
  if (!graph || !out) return NQPU_NULL_POINTER;
  *out = G(graph)->info;
  return NQPU_OK;
}

void nqpu_graph_destroy(NqpuRunGraph* graph) { delete G(graph); }

}  // extern "C"
