// runtime.h — host-side runtime of the engine: device state handles,
// circuits, gate lowering, the fusion passes and the pass planner.
#pragma once

#include "../../include/nqpu_hip.h"
#include "engine.h"

#include <functional>
#include <memory>
#include <string>
#include <vector>

namespace nqpu {

// A gate in the circuit IR (gates::Gate, sdk/rust/src/core/gates/gates.rs:632).
struct Gate {
  int32_t kind = 0;
  int32_t nt = 0, nc = 0;
  int32_t t[3] = {0, 0, 0};
  int32_t c[3] = {0, 0, 0};
  double p[4] = {0, 0, 0, 0};
  cd m[16];
  uint64_t qubit_mask() const;
};

struct Circuit {
  int n = 0;
  std::vector<Gate> gates;
};

// Item of the reference fusion pass (FusedOrOriginal, gate_fusion.rs:86-93).
struct FusedItem {
  enum Kind { FUSED, FUSED_TWO, DIAGONAL, ORIGINAL } kind = ORIGINAL;
  int q = 0, qlo = 0, qhi = 0;
  cd m[16];
  std::vector<std::pair<int, std::pair<cd, cd>>> single;  // (q, (p0, p1))
  std::vector<std::pair<std::pair<int, int>, cd>> pairs;  // ((lo, hi), p11)
  Gate gate;
  int fused_count = 1;
};

struct FusionResult {
  std::vector<FusedItem> items;
  size_t original_count = 0, fused_count = 0, gates_eliminated = 0;
};

// One planned tile pass.
struct Pass {
  std::vector<int> qubits;  // ascending physical positions, size K
  std::vector<HostOp> ops;  // physical numbering
  // the store position of the amplitude bit loaded from qubits[j] (a
  // permutation of qubits); empty = the same positions
  std::vector<int> out;
};

struct PlanOptions {
  int tile_qubits = 12;
  int low_qubits = 6;
  bool registers = true;  // register-resident tile kernel where it applies
  // Controls and diagonal factors on qubits outside a tile are resolved per
  // tile, so only the targets of non-diagonal ops need to be tile bits.
  bool global_ops = true;
  // fuse runs of gates on four tile bits into dense blocks applied on the
  // matrix cores (complex64 register tiles)
  bool mfma = true;
  // run complex64 register tile passes as straight-line kernels compiled at
  // run time (jit.hip) instead of the interpreting kernel
  bool compile = false;
  // merge runs of diagonal gates in a register tile pass into one quadratic
  // phase op (R_QPHASE)
  bool qphase = true;
  // complex64 register passes: the physical low positions may hold any
  // qubits of the previous tile (plan_tiles)
  bool move_low = false;
  // register tile passes: tiles per workgroup (the launch grid is the tile
  // count divided by this)
  int tiles_per_group = 1;
  // with move_low: end the run with the qubits where its last pass left
  // them (no closing permutation passes); plan_tiles reports that map
  bool keep_perm = false;
};

// A measured plan choice for states of n qubits at one precision (tune.hip).
struct TunedPlan {
  int tile_qubits = 0, low_qubits = 0;
  uint32_t flags = 0;  // NQPU_RUN_COMPILE / NQPU_RUN_FIXED_LOW as measured
  int tiles_per_group = 1;
  double ms = 0, default_ms = 0;
};
bool tuned_lookup(int n, int prec, TunedPlan* out);
// Measure the candidate plans for an n-qubit state (probe: the circuit to
// time, or null for a random U3 + CNOT brickwork of depth 8) and make the
// fastest the default for (n, prec).
NqpuError autotune(int n, int prec, const Circuit* probe, TunedPlan* out);
// Forget every tuned plan of this process (the cache file is not read again).
void tune_forget();
// Read tuned plans from a file in the NQPU_TUNE_CACHE format.
NqpuError tune_load(const char* path);

// lowering and planning (pure host code; unit-testable without a device)
std::vector<HostOp> lower_gate(const Gate& g, bool reference_semantics);
std::vector<HostOp> lower_circuit(const Circuit& c, bool reference_semantics);
std::vector<HostOp> merge_single_qubit(const std::vector<HostOp>& ops);
// final_pi (optional): where circuit qubit q is at the end of the plan
// (empty: at q; only with po.keep_perm can it differ)
std::vector<Pass> plan_tiles(const std::vector<HostOp>& ops, int n, const PlanOptions& po,
                             std::vector<int>* final_pi = nullptr);
// Permutation-only passes of K-qubit tiles (L = 5 low positions) that move
// circuit qubit q from position pi[q] back to position q (appended to out);
// false if they do not get there.
bool closing_passes(std::vector<int> pi, int n, int K, int L, std::vector<Pass>* out);
FusionResult fuse_gates_reference(const std::vector<Gate>& gates);
PlanOptions plan_options(const NqpuRunOptions& o, int prec, int n = 0);  // n: state qubits (0: unknown)
// What run_circuit would launch, computed on the host.
NqpuRunStats plan_stats(const Circuit& c, const NqpuRunOptions& o, int prec);
// Host emulation of the tiled execution (plan, per-pass programs, kernel
// semantics) on a complex128 vector.
NqpuError emulate_tiled(const Circuit& c, const NqpuRunOptions& o, int prec, std::vector<cd>& a);

// Kernels of the given register tile passes (EB = eb, complex64) as
// straight-line code, compiled with hipRTC for the current device on first use
// and cached for the life of the process by program structure (jit.hip).
// The modules behind *fns are appended to *hold; they stay loaded while held.
template <typename R>
NqpuError jit_pass_functions(const std::vector<RegOp<R>>& prog, const std::vector<PassArgs>& args, int eb,
                             std::vector<hipFunction_t>* fns, std::vector<std::shared_ptr<void>>* hold);
// Register tile passes compiled so far in this process.
uint64_t jit_compiled_passes();
// Generate and compile (no device needed, nothing loaded or cached) the
// straight-line kernels nqpu_run with NQPU_RUN_COMPILE would use for c.
NqpuError jit_check_circuit(const Circuit& c, const NqpuRunOptions& o, int prec = NQPU_PRECISION_C64);
template <typename R>
NqpuError jit_check_compile(const std::vector<RegOp<R>>& prog, const std::vector<PassArgs>& args, int eb);

struct DeviceScratch {
  void* p = nullptr;
  size_t bytes = 0;
  ~DeviceScratch();
  bool ensure(size_t b);
};

// HIP events bracketing each launch of a run (when the run asks for timing).
struct PassTimer {
  std::vector<hipEvent_t> pool;
  size_t used = 0;
  bool on = false;
  void begin(bool enable) { on = enable; used = 0; }
  void mark(hipStream_t s);
  double total_ms();  // sum over (start, stop) pairs; call after the stream is synchronised
  ~PassTimer();
};

struct State {
  int n = 0;
  int prec = NQPU_PRECISION_C128;
  void* amps = nullptr;
  bool owns_amps = true;  // false: amps lives in memory the owner of the State manages (a server's arena)
  hipStream_t stream = nullptr;
  int device = 0;
  NqpuRunStats stats{};
  DeviceScratch scratch, scratch2, opbuf;
  // pinned staging of register tile programs: their upload is an
  // asynchronous copy in stream order, so the host plans the next run while
  // the device works; stage_ev marks the copy out of it done
  void* stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_ev = nullptr;
  bool stage_pending = false;
  PassTimer timer;
  // run-time compiled modules the launches in flight on `stream` use; released
  // after the stream is synchronised
  std::vector<std::shared_ptr<void>> jit_hold;
  // runs were queued with NQPU_RUN_ASYNC and not yet waited for
  bool async_open = false;
  // Qubit layout: amplitude bit q of the state's logical index is stored at
  // bit perm[q] of its position in `amps` (empty: at bit q).  A tiled run
  // leaves the qubits where its last pass put them; perm_k is that run's
  // tile width, which the passes that restore the layout (settle_state) use.
  std::vector<int> perm;
  int perm_k = 0;
  ~State();
  size_t amp_size() const { return prec == NQPU_PRECISION_C64 ? 8 : 16; }
  uint64_t dim() const { return 1ull << n; }
};

// One kernel launch of a tiled run.
struct RunLaunch {
  enum Kind { LDS = 0, REG = 1, COMPILED = 2 };
  int kind = LDS;        // k_tile, k_tile_reg, or a run-time compiled pass kernel
  PassArgs pa{};
  int eb = 0;            // register passes: register bits per thread
  size_t blob_at = 0;    // register passes: byte offset of the pass's operand blobs in the image
  hipFunction_t fn = nullptr;
};
// A tiled run planned on the host: the image its kernels read from device
// memory (register programs and operand blobs, or the LDS kernel's ops) and
// its launches, which read the image at whatever device address it is
// uploaded to.
struct PreparedRun {
  int prec = NQPU_PRECISION_C64, n = 0;
  std::vector<char> image;
  size_t prog_bytes = 0;  // programs; operand blobs follow
  std::vector<RunLaunch> launches;
  std::vector<std::shared_ptr<void>> hold;  // compiled modules the launches use
  NqpuRunStats stats{};
  // the state's qubit layout after the run (State::perm) and its tile width
  std::vector<int> perm;
  int perm_k = 0;
};
// Plan circuit c (NQPU_FUSION_TILED; c.n <= n) for an n-qubit state of precision prec.
NqpuError prepare_tiled(const Circuit& c, const NqpuRunOptions& o, int prec, int n, const PlanOptions& po,
                        PreparedRun* out);
// Make the run start from |0...0>: its first register pass generates that
// state instead of loading it.  False (run unchanged) when its first launch
// is an LDS tile pass or it has none.
bool start_from_zero(PreparedRun* r);
// Enqueue the run's launches on st over the state at amps, the image at
// `image` in device memory (timer: optional HIP events around each launch).
NqpuError launch_prepared(const PreparedRun& r, void* amps, const void* image, hipStream_t st, PassTimer* timer);

// Error plumbing: the last error detail is thread-local.
void set_error(const std::string& s);
const char* last_error();
NqpuError hip_check(hipError_t e, const char* what);
#define HIPTRY(expr, what)                              \
  do {                                                  \
    NqpuError e_ = ::nqpu::hip_check((expr), what);     \
    if (e_ != NQPU_OK) return e_;                       \
  } while (0)

int device_count();
NqpuError create_state(int n, int prec, State** out);
NqpuError clone_state(const State* s, State** out);
NqpuError reset_state(State* s);
// Restore the state's canonical layout (State::perm) with permutation-only
// passes queued on its stream; a no-op when it has it.  Every operation that
// reads or writes amplitudes by index calls this first; the norm and scaling
// do not depend on the layout.
NqpuError settle_state(const State* s);
// The positions in the state's current layout of the qubits in mask m.
inline uint64_t layout_mask(const State* s, uint64_t m) {
  if (s->perm.empty()) return m;
  uint64_t r = 0;
  for (; m; m &= m - 1) r |= 1ull << s->perm[__builtin_ctzll(m)];
  return r;
}
NqpuError run_circuit(State* s, const Circuit& c, const NqpuRunOptions& o);
// Wait for the runs queued on the state; completes their statistics.
NqpuError synchronize_state(State* s);
// run_circuit with explicit plan options for the tiled path
NqpuError run_circuit_with(State* s, const Circuit& c, const NqpuRunOptions& o, const PlanOptions& po);
NqpuError apply_gate_now(State* s, const Gate& g);
NqpuError get_amplitudes(const State* s, uint64_t off, uint64_t count, NqpuComplex* out);
NqpuError set_amplitudes(State* s, uint64_t off, uint64_t count, const NqpuComplex* in);
NqpuError probabilities(const State* s, double* out);
NqpuError norm(const State* s, double* out);
NqpuError zsum_all(const State* s, double* tot, std::vector<double>* s1);
NqpuError pauli(const State* s, uint64_t x, uint64_t z, double* out);
NqpuError scale(State* s, double f);
NqpuError overlap(const State* a, const State* b, NqpuComplex* out);
NqpuError sample(const State* s, const std::vector<double>& r, uint64_t* out);
// Reductions that share passes (reduce.hip).
struct PauliTerm;
// <P_t> of every term (no coefficients), terms with equal X masks in one pass.
// phase[t] (optional): the term is i^phase[t] * (X on x & ~z, Y on x & z as
// X Z products, Z on z & ~x) evaluated as Re(i^phase[t] * sum_i (-1)^popc(i & z)
// conj(a_{i^x}) a_i); without it, phase = the number of Y factors.
NqpuError pauli_values(const State* s, const std::vector<PauliTerm>& terms, std::vector<double>* vals,
                       const std::vector<int>* phase = nullptr);
// sum_t coeffs[t] <P_t>
NqpuError pauli_sum(const State* s, const uint64_t* xm, const uint64_t* zm, const double* coeffs, uint64_t count,
                    double* out);
// The same sum with every device buffer in a fixed device array of
// reduce.hip (nothing is allocated); at most kPauliSumNoAllocMaxTerms terms.
constexpr uint64_t kPauliSumNoAllocMaxTerms = 1024;
NqpuError pauli_sum_noalloc(const State* s, const uint64_t* xm, const uint64_t* zm, const double* coeffs,
                            uint64_t count, double* out);
// Look up the fixed scratch array's address on device `dev` (once per device).
NqpuError scratch_init(int dev);
NqpuError measure_qubit(State* s, int q, double r, uint32_t* outcome, double* prob);
// Keep the amplitudes whose bit q equals `outcome`, multiplied by f; zero the rest.
NqpuError collapse(State* s, int q, uint32_t outcome, double f);
// lib.rs:1705-1730 and lib.rs:1918-1935 (reduce.hip)
NqpuError reset_qubit(State* s, int q);
NqpuError initialize_amplitudes(State* s, const NqpuComplex* in, uint64_t len);

// Batched expectation values over many complex64 circuits (k_batch).
struct PauliTerm {
  uint64_t x, z;
  double coeff;
};
// One batch in flight: its tables on the host (pinned image) and device,
// its results, and the events that mark its upload and completion.
struct BatchSlot {
  DeviceScratch tables, out;
  char* pinned = nullptr;  // pinned host image of the tables, reused across batches
  size_t pinned_bytes = 0;
  double* results = nullptr;  // pinned host copy of the results
  size_t results_cap = 0;
  double* dest = nullptr;     // the caller's buffer, written when the batch is retired
  size_t count = 0;
  hipEvent_t uploaded = nullptr, k0 = nullptr, k1 = nullptr, done = nullptr;
  bool busy = false;
  NqpuBatchStats st{};
};
// Batched execution context: one stream, one state region, and two slots so
// that the host plans batch k+1 while the device runs batch k.
struct Batch {
  hipStream_t stream = nullptr;
  hipStream_t copy_stream = nullptr;  // table uploads: batch k+1's overlaps batch k's launch
  DeviceScratch states;
  BatchSlot slot[2];
  int next = 0;  // slot of the next batch
  NqpuBatchStats stats{};
  ~Batch();
};
// Host half of a batch: the tables k_batch reads, in one upload image.
// Host image: [circuits][passes][programs][terms][counter] (head bytes), the
// phase-table units, then one complex64 16x16 matrix per matrix-core block.
// Device tables: the head, the blocks' operand blobs (built on the device
// from the matrices), the phase-table units, the matrices.  Programs address
// both kinds of operand unit from the start of the blob area.
struct BatchPlan {
  char* host = nullptr;  // from the buffer callback of plan_batch (not owned)
  size_t oc = 0, op = 0, og = 0, ot = 0, ocnt = 0, total = 0;  // total: host image bytes
  size_t head = 0;                            // host = device offset of the blob area
  size_t nblk = 0, qunits = 0;                // matrix-core blocks, phase-table units
  // device offsets: phase tables, the blocks' word index and row-operation
  // words, matrices (composed there on the device); size
  size_t dev_q = 0, dev_bi = 0, dev_bw = 0, dev_m = 0, dev_total = 0;
  uint64_t slot_amps = 0;  // largest queued HBM-resident circuit: each workgroup's state slot
  size_t ncirc = 0, program_ops = 0;
  // split circuits (kBatchSplitQubits and more) follow the nqueue queued ones
  // in the circuit table.  They run in waves whose states fit
  // kBatchSplitWaveAmps amplitudes (split_amps: the largest wave) after the
  // workgroup slots; each wave's launches run its gate passes (one launch
  // per pass index) and then its observable passes, which write one partial
  // sum per item.  Work items at oit; per split circuit its first observable
  // item (counted over all waves) and count at ofirst / ocount.
  struct SplitLaunch {
    size_t first, count;  // work items
    bool obs;             // observable items (else gate items)
    size_t part0;         // observable items: index of the first one's partial sum
  };
  size_t nqueue = 0;
  uint64_t split_amps = 0;
  size_t oit = 0, ofirst = 0, ocount = 0;
  std::vector<SplitLaunch> split_launches;
  size_t split_parts = 0;  // observable items over all waves
  NqpuBatchStats st{};
};
// buffer(bytes) returns at least `bytes` of host memory for the table image
NqpuError plan_batch(const std::vector<const Circuit*>& circuits, const std::vector<std::vector<PauliTerm>>& terms,
                     const NqpuRunOptions& o, BatchPlan& plan, const std::function<char*(size_t)>& buffer);
NqpuError create_batch(Batch** out);
// With NQPU_RUN_ASYNC in o.flags the call returns once the batch is queued;
// `out` is written when batch_synchronize returns (or a later batch needs
// the slot), and must stay valid until then.
NqpuError batch_expectation(Batch* b, const std::vector<const Circuit*>& circuits,
                            const std::vector<std::vector<PauliTerm>>& terms, const NqpuRunOptions& o,
                            double* out);
// Wait for every queued batch and write its results.
NqpuError batch_synchronize(Batch* b);

// The C ABI's thread-local xorshift64* stream (c_ffi.rs:48-67).
// The C ABI's xorshift64* generator (sdk/rust/src/infra/bindings/c_ffi.rs:48-67):
// a thread-local stream for the nqpu_* calls and explicit streams for callers
// that own one (a server request).
void rng_seed(uint64_t seed);
double rng_next();
uint64_t rng_init(uint64_t seed);
double rng_step(uint64_t& state);

// A stream of default priority that does not synchronise with the null stream.
hipError_t new_stream(hipStream_t* st);

// Continuous-batching request server (server.hip).
struct Server;
NqpuError create_server(const NqpuServerConfig& cfg, Server** out);
void destroy_server(Server* s);
NqpuError server_submit(Server* s, const Circuit& c, const std::vector<PauliTerm>& terms, uint32_t shots,
                        uint64_t seed, uint64_t* id);
NqpuError server_poll(Server* s, uint64_t id, NqpuRequestStatus* out);
NqpuError server_wait(Server* s, uint64_t id, uint32_t timeout_ms, NqpuRequestStatus* out);
NqpuError server_take(Server* s, uint64_t id, double* expectations, uint32_t num_terms, uint64_t* samples,
                      uint32_t shots);
NqpuError server_stats(const Server* s, NqpuServerStats* out);
// qubits, term count and shots of a known request
NqpuError server_request_shape(Server* s, uint64_t id, int* n, uint32_t* num_terms, uint32_t* shots);

}  // namespace nqpu
