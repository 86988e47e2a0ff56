// engine.h — internal interface between the host runtime (engine.hip,
// capi.hip) and the gfx950 kernels (kernels.hip).  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include "devtypes.h"

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace nqpu {

using cd = std::complex<double>;

// Lowered operation kinds.  A gate of any kind is lowered to one or more of
// these (engine.hip: lower_gate); all carry a control mask whose bits must be
// 1 for the operation to act.
enum OpKind : int32_t {
  OP_U1 = 0,    // dense 2x2 on qubit a
  OP_D1 = 1,    // diag(m[0], m[1]) on qubit a
  OP_U2 = 2,    // dense 4x4 on (a, b); sub-index (bit_b << 1) | bit_a
  OP_SWAP = 3,  // exchange |..a=1..b=0..> with |..a=0..b=1..>
  OP_D2 = 4,    // diag(m[0..3]) on (a, b); index (bit_b << 1) | bit_a
  OP_FLIP = 5,  // X on qubit a (a permutation; CNOT/Toffoli with controls)
  OP_DIAGK = 6  // phase LUT over up to 16 qubits (fused diagonal item)
};

// Operation in global qubit numbering, FP64 coefficients.
struct HostOp {
  int32_t kind = OP_U1;
  int32_t a = 0, b = 0;
  uint64_t cmask = 0;
  cd m[16];
  std::vector<int32_t> dq;   // OP_DIAGK qubits
  std::vector<cd> lut;       // OP_DIAGK phases, index = bits of dq
  uint64_t qubit_mask() const;
  // Qubits that must be bits of a tile for the op to run inside a tile pass:
  // the targets of non-diagonal ops.  Controls and diagonal factors on other
  // qubits are fixed per tile and resolved from the tile's base index.
  uint64_t tile_mask() const;
};

// Operation as the tile kernel reads it: qubit positions are tile-local bits.
template <typename R>
struct DevOp {
  int32_t kind, a, b, pad;  // OP_D1/OP_D2: a, b >= kGlobalBit name qubit a-kGlobalBit outside the tile
  uint64_t cmask;
  uint64_t gc;              // controls outside the tile (global qubit mask)
  R m[32];  // interleaved re/im, up to 4x4
};

// ------------------------------------------------------ batched execution
// k_batch: one persistent launch over many complex64 circuits.  Each
// workgroup takes whole circuits from an atomic queue; a circuit of at most
// 8 + kBatchEB qubits lives in registers/LDS for its whole life, a larger
// one streams register-tile passes over the HBM slot of the workgroup that
// runs it (slot_amps amplitudes per workgroup, reused circuit after circuit).  Observables
// (Pauli sums) are evaluated on LDS tiles in FP64.
constexpr int kBatchEB = 4;

struct BatchCircuit {
  int32_t n;
  int32_t resident;      // 1: the whole state is one tile
  int32_t pass0, npass;  // gate passes (BatchPass table)
  int32_t red0, nred;    // observable passes (BatchPass table; resident: the gate pass)
  int32_t out;           // output slot
  int32_t pad;
  uint64_t amps;         // amplitudes of the HBM state (not resident)
  uint64_t off;          // split circuits: offset of its own state in the split-state region
};

// Circuits of at least kBatchSplitQubits qubits are "split": their passes
// run as ordinary grids, one launch per pass index over every split
// circuit, each workgroup taking kBatchSplitTiles consecutive tiles of one
// circuit's pass (instead of one workgroup streaming a whole circuit, which
// made the largest circuits the tail of the persistent launch).
constexpr int kBatchSplitQubits = 18;
constexpr int kBatchSplitTiles = 8;
// split circuits run in waves whose states fit this many complex64
// amplitudes (8 GiB), so the HBM a batch needs does not grow with its count
constexpr uint64_t kBatchSplitWaveAmps = 1ull << 30;

// A unit of a split circuit's work: `ntile` consecutive tiles of gate pass
// or observable pass `pass` of circuit `circ` (table index).
struct BatchItem {
  int32_t circ, pass, tile0, ntile;
};

struct BatchPass {
  int32_t K, tb;         // tile qubits, thread bits (K = tb + kBatchEB for gate passes)
  int32_t op0, nops;     // register program
  int32_t term0, nterm;  // observable terms evaluated on this tile set
  int32_t q[16];         // tile qubits, ascending
  // gate passes over HBM: first operand unit (kMfmaBlobBytes) of the
  // program's matrix-core blocks and phase tables in the batch's operand
  // buffer, and nonzero when the amplitudes go through the blocks scaled
  // by 2^15 (as in the standalone passes)
  int32_t blob_unit, scaled;
};

struct BatchTerm {
  uint32_t xl, zl;       // X/Z bits inside the tile (tile-local numbering)
  uint64_t zg;           // Z bits outside the tile (global numbering)
  double coeff;
  int32_t ny;            // number of Y factors mod 4
  int32_t pad;
};

void launch_batch(void* states, uint64_t slot_amps, const BatchCircuit* circs, int ncirc, const BatchPass* passes,
                  const RegOp<float>* prog, const BatchTerm* terms, const unsigned char* blob, unsigned* counter,
                  double* out, int grid, hipStream_t s);
// Split circuits: one gate pass index over `nitems` items (states: the
// split-state region); all observable passes into part[item]; out[C.out] =
// the sum of each split circuit's parts, in item order (first[k], count[k]
// for split circuit k of the table's split circuits circ0, circ0 + 1, ...).
void launch_batch_split_gate(void* states, const BatchCircuit* circs, const BatchPass* passes,
                             const RegOp<float>* prog, const unsigned char* blob, const BatchItem* items,
                             int nitems, hipStream_t s);
void launch_batch_split_obs(const void* states, const BatchCircuit* circs, const BatchPass* passes,
                            const BatchTerm* terms, const BatchItem* items, int nitems, double* part, hipStream_t s);
void launch_batch_split_out(const BatchCircuit* circs, int circ0, int nsplit, const int32_t* first,
                            const int32_t* count, const double* part, double* out, hipStream_t s);


// Operand blobs of nblk matrix-core blocks from their complex64 matrices
// (kMfmaMatrixBytes each, row-major in block-index order) into consecutive
// kMfmaBlobBytes units: the layout and fp16 high/low split reg_mfma reads.
void launch_blob_split(const float2* mats, unsigned char* blobs, size_t nblk, hipStream_t s);
// Matrices of nblk blocks from their row-operation words (BR_* records).
void launch_block_compose(const uint32_t* index, const uint64_t* words, float2* mats, size_t nblk, hipStream_t s);

// ----------------------------------------------------------- kernel launches
// All launches are asynchronous on `s`; R = float (c64) or double (c128).
template <typename R>
struct Launch {
  static void init_zero(void* st, int n, hipStream_t s);
  static void fill_basis(void* st, int n, uint64_t idx, hipStream_t s);
  static void u1(void* st, int n, int a, uint64_t cmask, const cd* m, hipStream_t s);
  static void d1(void* st, int n, int a, uint64_t cmask, cd p0, cd p1, hipStream_t s);
  static void flip(void* st, int n, int a, uint64_t cmask, hipStream_t s);
  static void u2(void* st, int n, int a, int b, uint64_t cmask, const cd* m, hipStream_t s);
  static void d2(void* st, int n, int a, int b, uint64_t cmask, const cd* p, hipStream_t s);
  static void swap(void* st, int n, int a, int b, uint64_t cmask, hipStream_t s);
  static void diagk(void* st, int n, int k, const int32_t* q, const void* dlut, hipStream_t s);
  static void tile_pass(void* st, const PassArgs& pa, const void* dops, hipStream_t s);
  // register-resident tile pass; pa.K = 8 + eb, program = RegOp<R>[pa.op0 ..]
  static void tile_pass_reg(void* st, const PassArgs& pa, int eb, const void* prog, const void* blob,
                            hipStream_t s);
  static void scale(void* st, int n, double f, hipStream_t s);
  // FP64 reductions: per-chunk partials, then deterministic final sums.
  static void sumsq_chunks(const void* st, int n, int chunk_bits, double* out, hipStream_t s);
  static void zsum_chunks(const void* st, int n, int chunk_bits, double* out, hipStream_t s);
  static void to_c128(const void* st, uint64_t count, uint64_t off, void* out, hipStream_t s);
  static void from_c128(void* st, uint64_t count, uint64_t off, const void* in, hipStream_t s);
  static void collapse(void* st, int n, int q, uint32_t outcome, double renorm, hipStream_t s);
};

// Reduce per-chunk double partials [nchunks][width] -> [width] in a fixed
// order (deterministic for a given n).
void launch_reduce_rows(const double* in, uint64_t rows, int width, double* out, hipStream_t s);
// dst[i] = src[i] over 2^n amplitudes (a plain streaming copy).
void launch_copy(void* dst, const void* src, uint64_t dim, bool c64, hipStream_t s);
// Per-chunk [Re, Im] of sum conj(a_i) b_i; a64/b64 select complex64 storage.
void launch_overlap(const void* a, bool a64, const void* b, bool b64, int n, int cb, double* out,
                    hipStream_t s);

}  // namespace nqpu
