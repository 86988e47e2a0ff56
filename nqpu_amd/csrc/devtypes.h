// devtypes.h — register tile programs as the gfx950 kernels read them.
// Shared by the host runtime, the precompiled kernels and the per-pass
// kernels generated at run time (whose source embeds this file), so it uses
// fixed-width integer types only and includes nothing.
#pragma once

namespace nqpu {

// Register-resident tile program (k_tile_reg).  Each of the 256 threads of a
// tile workgroup holds 2^EB amplitudes; tile bits are "thread slots" (0..7,
// the thread index bits) or "register slots" (0..EB-1, the register index
// bits).  map[0..7] / map[8..8+EB) give the tile-local bit held by each slot.
enum RegOpKind : int32_t {
  R_REMAP = 0,  // exchange through LDS: slots change from map_old to map_new
  R_U1 = 1,     // 2x2 on register slot s0
  R_FLIP = 2,   // X on register slot s0
  R_D1 = 3,     // diag(m[0], m[1]) on position s0 (>= kGlobalBit: qubit s0-kGlobalBit outside the
                //   tile; >= 16: register slot s0-16; else thread slot)
  R_D2 = 4,     // diag(m[0..3]) on positions s0 (index bit 0), s1 (index bit 1)
  R_U2 = 5,     // 4x4 on register slots s0 (index bit 0), s1 (index bit 1)
  R_SWAP = 6,   // exchange register slots s0 and s1 (controlled SWAP)
  // Dense 16x16 block on four tile bits, on the matrix cores (complex64, EB = 4
  // only): block index bits a0, a1, a3 are register slots 0, 1, 2 and a2 is
  // thread slot 5 (lane bit 5); register slot 3 is not in the block.  The
  // result leaves a1 on thread slot 5 and a2 on register slot 1.  The block
  // depends on pad (<= 2) qubits outside the tile, at positions s1 & 255 and
  // s1 >> 8; their bits in the tile's base index select the variant, blob
  // s0 + variant.  s0 indexes the op's operand blob (kMfmaBlobBytes each): per lane l and K-step s, the
  // fp16 high and low parts of A[l % 32][16 s + 8 (l / 32) + i], i = 0..7,
  // where A is the 32x32 real form of the block matrix (row 2a'+c', column
  // 2a+c; c = 0 real, 1 imaginary part).
  R_MFMA = 7,
  // Exchange register slot s0 with thread slot s1 < 6 (a lane bit: DPP or
  // lane-half swaps inside the wavefront, no LDS) or with register slot
  // s1 - 16 (register moves).  Afterwards each slot holds the other's bit.
  R_XCHG = 8,
  // Phase exp(2 pi i Phi(x)) with Phi a quadratic form in the index bits, in
  // revolutions (a run of one- and two-qubit diagonal gates).  Per thread
  // (tid) and register index r:
  //   Phi = C[tid] + sum_c K0_c + sum_t tid_t Kt_c[t] + sum_j r_j (L_j[tid] + sum_c Kr_c[j])
  //         + sum_{j<k} r_j r_k m[pair(j, k)]
  // Tables start at operand unit s0 (kMfmaBlobBytes each) of the pass: C[256]
  // then L_j[256] for j < EB (floats), then one table per chunk c < pad of
  // qubits outside the tile, 256 x 16 floats indexed by the chunk's bits of
  // the tile's base index (lowest qubit = bit 0): [K0, Kt[0..7], Kr[0..EB-1]].
  // Chunk c's qubits are the set bits of the 64-bit mask qphase_chunk(op, c)
  // (bytes 8c.. of map_old ++ map_new).  m[pair(j, k)]: register-slot pair
  // coefficients, pair(j, k) = k (k - 1) / 2 + j.
  R_QPHASE = 9
};

constexpr int kMfmaBlobBytes = 64 * 64;  // 64 lanes x (2 K-steps x (hi, lo) x 8 fp16)
constexpr int kMfmaMatrixBytes = 256 * 8;  // a block's 16x16 complex64 matrix (the blob's source)

// A matrix-core block recorded as its row operations (batched circuits): a
// stream of 64-bit words per block, composed into the 16x16 matrix on the
// device in FP64.  Each op is a header word, then `nvals` FP64 values
// (re, im interleaved).  Header: kind | pa << 8 | pb << 12 | cmk << 16 |
// nvals << 32.  pa, pb: block index bits (0..3); the op acts on the rows i
// with (i & cmk) == cmk.
enum BlockRowKind : uint32_t {
  BR_U1 = 0,     // 2x2 on bit pa (4 values)
  BR_U2 = 1,     // 4x4 on bits (pa, pb), sub-index bit_pa | bit_pb << 1 (16 values)
  BR_FLIP = 2,   // X on bit pa
  BR_SWAP = 3,   // exchange bits pa and pb
  BR_SCALE0 = 4, // every row times one factor (1 value)
  BR_SCALE1 = 5, // row i times f[bit_pa(i)] (2 values)
  BR_SCALE2 = 6  // row i times f[bit_pa(i) | bit_pb(i) << 1] (4 values)
};
constexpr uint64_t block_row_header(uint32_t kind, int pa, int pb, uint32_t cmk, int nvals) {
  return (uint64_t)kind | (uint64_t)pa << 8 | (uint64_t)pb << 12 | (uint64_t)cmk << 16 | (uint64_t)nvals << 32;
}
constexpr int kQPhaseChunkFloats = 256 * 16;  // one chunk table

template <typename R>
struct RegOp {
  int32_t kind, s0, s1, pad;
  uint32_t ct;  // controls on thread slots (bits of the thread index)
  uint32_t cr;  // controls on register slots (bits of the register index)
  uint8_t map_old[16];
  uint8_t map_new[16];
  uint64_t gc;  // controls outside the tile (global qubit mask): the op runs on tiles that satisfy them
  R m[32];
};

// R_QPHASE: mask of the qubits outside the tile in chunk c < 4 (read as two
// 32-bit words: the kernels load it with scalar loads)
template <typename R>
__host__ __device__ inline uint64_t qphase_chunk(const RegOp<R>& op, int c) {
  const uint32_t* w = reinterpret_cast<const uint32_t*>(c < 2 ? op.map_old : op.map_new) + 2 * (c & 1);
  return (uint64_t)w[0] | ((uint64_t)w[1] << 32);
}
template <typename R>
inline void set_qphase_chunk(RegOp<R>& op, int c, uint64_t m) {
  uint8_t* b = (c < 2 ? op.map_old : op.map_new) + 8 * (c & 1);
  for (int i = 0; i < 8; ++i) b[i] = (uint8_t)(m >> (8 * i));
}

// Diagonal-op positions at or above kGlobalBit name qubit (pos - kGlobalBit)
// outside the tile; its value is the same for the whole tile.
constexpr int kGlobalBit = 64;

constexpr int kMaxTile = 14;  // maximum tile qubits (LDS: 2^13 * 16 B = 128 KiB)


// One tile pass: tile qubits (global positions, ascending), the ops it
// applies, and where they live in the uploaded op buffer.
struct PassArgs {
  int32_t K;
  int32_t n;
  int32_t nops;
  int32_t op0;
  int32_t q[kMaxTile];  // tile qubits, ascending (the tile's base index skips them)
  // register tile pass: qubit held by each slot at load (ql) and at store
  // (qf); slots 0..7 are thread-index bits, 8.. register-index bits
  int32_t ql[16];
  int32_t qf[16];
  // nonzero: the pass holds R_MFMA ops, so amplitudes are kept multiplied by
  // 2^15 in registers (loaded x 2^15, stored x 2^-15) and their fp16 parts
  // stay clear of the subnormal range
  int32_t scaled;
  // register tile pass: bytes of LDS for R_REMAP (0 when the program has none)
  int32_t lds_remap;
  // register tile pass: nonzero when a wavefront stores amplitudes that
  // another wavefront of the workgroup loaded (the pass permutes the
  // positions held by the wave slots) and no R_REMAP barrier separates the
  // loads from the stores: all loads of the workgroup complete first
  int32_t sync_store;
  // tiles each workgroup processes (grid = tiles / tiles_per_group); 0 or 1: one
  int32_t tiles_per_group;
  // register tile pass: nonzero: the pass starts from |0...0> (amplitude 1 at
  // position 0) instead of loading the state; the first pass of a run from
  // the start state reads nothing
  int32_t from_zero;
};
constexpr float kMfmaScale = 32768.0f;

}  // namespace nqpu
