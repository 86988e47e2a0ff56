// reduce.hip — FP64 reductions that read the state once for many results.
//
//  * Grouped Pauli sums: terms that share their X/Y mask share one pass over
//    the state (all diagonal Z strings of a Hamiltonian in one pass; every
//    term with the same flip pattern in one more), with one FP64 accumulator
//    per distinct pattern of Z bits above the grid stride; strings with
//    different X masks that fit one 12-qubit tile share an LDS tile pass.
//    The reference evaluates each string with its own serial loop
//    over all 2^n indices (expectation_pauli_string, sdk/rust/src/lib.rs:465-514;
//    expectation_hamiltonian :518).
//  * Sampling: sub-block masses in one coalesced pass (one wavefront per
//    sub-block), chunk prefix sums on the device, then one wavefront per shot
//    that walks chunk -> sub-block -> amplitude with wavefront prefix scans.
//    It returns the C ABI's outcome, the first index i with r < cumsum(|a|^2)
//    (c_ffi.rs:506-525), without a host round trip between the passes.
//  * Probabilities of any size computed on the device (|a|^2 in FP64).
//  * A Pauli-sum entry point that allocates nothing: its partial sums live
//    in a fixed device array of this file.
//
// Every sum runs in a fixed order for a given qubit count (fixed grids,
// fixed block reductions, fixed final sums), so results are bitwise
// reproducible and do not depend on what else runs on the device.
#include "runtime.h"
#include "regtile.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace nqpu {

namespace {

template <typename R> struct V2r;
template <> struct V2r<float> { using T = float2; };
template <> struct V2r<double> { using T = double2; };

template <typename T>
__device__ __forceinline__ double nsq2(T a) {
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Inclusive prefix sum over the 64 lanes of a wavefront (Hillis-Steele).
__device__ __forceinline__ double wscan(double v, int lane) {
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  return v;
}

// ------------------------------------------------------------ Pauli groups

constexpr int kPauliGroup = 32;  // terms evaluated in one pass
constexpr int kPauliKeys = 16;   // distinct accumulators of one pass
constexpr int kPauliThreads = 256;

struct PauliGroupArgs {
  uint64_t x;     // X/Y mask shared by the group's terms (0: diagonal strings)
  int32_t h;      // highest bit of x
  int32_t nt;     // terms in this launch
  int32_t nk;     // keys in this launch; key 0 is (no iteration bits, real part)
  int32_t width;  // doubles per partial row (>= nt)
  uint32_t kodd;  // bit j: key j accumulates Im(conj(a_j) a_i) (an odd number of Y factors)
  uint32_t pad;
  uint64_t u[kPauliKeys];      // iteration bits of each key's Z mask
  uint64_t zt[kPauliGroup];    // thread bits of each term's Z mask
  uint32_t key[kPauliGroup];   // key of each term
};

// Per workgroup and term t: A_t = sum over the workgroup's indices of
//   x == 0: (-1)^popc(k & z_t) |a_k|^2
//   x != 0: (-1)^popc(i & z_t) w_t(i)  over pairs (i, j = i ^ x), bit h of i clear,
//           w_t = Re(conj(a_j) a_i) (even Y count) or Im(conj(a_j) a_i) (odd)
// <P_t> follows on the host: x == 0: A; else 2A with the i^#Y phase.
// With i = ins0(k, h) and zc_t = z_t with bit h removed, popc(i & z_t) =
// popc(k & zc_t).  The grid is a power of two, so k = r * stride + base with
// base = block * 256 + thread < stride: the sign splits into a per-thread
// part popc(base & zt_t), applied once at the end, and a part popc(r & u_t) of
// the iteration r, uniform over the grid (scalar).  Terms with the same
// iteration bits u_t and the same real/imaginary choice share one
// accumulator (a key): the loop costs one FP64 FMA and one scalar parity per
// key, not per term, and key 0 (no iteration bits: every string on the
// qubits below the iteration bits) needs no sign.  The accumulator of a term
// equals that of its key bit for bit, so the sums do not depend on how terms
// are keyed.  Four loads are kept in flight.
template <typename T, int NK>
__global__ __launch_bounds__(kPauliThreads) void k_pauli_group(const T* __restrict__ st, uint64_t work,
                                                              PauliGroupArgs g, double* __restrict__ part) {
  double acc[NK];
#pragma unroll
  for (int j = 0; j < NK; ++j) acc[j] = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * kPauliThreads;
  const uint64_t base = (uint64_t)blockIdx.x * kPauliThreads + threadIdx.x;
  const uint64_t hmask = (1ull << g.h) - 1;
  auto one = [&](uint64_t r0) __attribute__((always_inline)) {
    const uint64_t k = r0 + base;
    if (k >= work) return;
    double w0, w1;
    if (g.x == 0) {
      w0 = w1 = nsq2(st[k]);
    } else {
      const uint64_t i = ((k & ~hmask) << 1) | (k & hmask), j = i ^ g.x;
      const T ai = st[i], aj = st[j];
      w0 = (double)aj.x * ai.x + (double)aj.y * ai.y;  // Re(conj(a_j) a_i)
      w1 = (double)aj.x * ai.y - (double)aj.y * ai.x;  // Im(conj(a_j) a_i)
    }
    acc[0] += w0;
#pragma unroll
    for (int j = 1; j < NK; ++j) {
      const double sg = (__popcll(r0 & g.u[j]) & 1) ? -1.0 : 1.0;  // uniform over the grid
      acc[j] = fma(((g.kodd >> j) & 1) ? w1 : w0, sg, acc[j]);
    }
  };
  uint64_t r = 0;
  for (; r + 3 * stride < work; r += 4 * stride) {
    one(r);
    one(r + stride);
    one(r + 2 * stride);
    one(r + 3 * stride);
  }
  for (; r < work; r += stride) one(r);
  // Each thread stages its keys in its own LDS column and reads them back
  // by term (a uniform index), so no barrier is needed before the reads.
  __shared__ double stage[NK][kPauliThreads];
  __shared__ double red[kPauliGroup][kPauliThreads / 64];
#pragma unroll
  for (int j = 0; j < NK; ++j) stage[j][threadIdx.x] = acc[j];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int t = 0; t < g.nt; ++t) {
    const double a = stage[g.key[t]][threadIdx.x];
    const double v = (__popcll(base & g.zt[t]) & 1) ? -a : a;  // the per-thread part of the sign
    const double w = wsum(v);
    if (lane == 0) red[t][wave] = w;
  }
  __syncthreads();
  if ((int)threadIdx.x < g.nt) {
    double s = 0;
#pragma unroll
    for (int w = 0; w < kPauliThreads / 64; ++w) s += red[threadIdx.x][w];
    part[(uint64_t)blockIdx.x * g.width + threadIdx.x] = s;
  }
}

// ------------------------------------------------ Pauli strings on tiles
//
// Strings with different X masks need a pass each in k_pauli_group.  When
// their X bits all fit one tile of K <= 12 qubits (the transverse-field
// terms X_i of a chain, for one), one pass over the state evaluates them all:
// a workgroup loads a tile of 2^K amplitudes into LDS (the state's lowest
// qubits are tile bits, so the loads are contiguous runs), and every pair
// (i, i ^ x) of a string lies inside one tile.  The raw sum A_t is the same
// as k_pauli_group's: the sign (-1)^popc(i & z) splits into the tile-local Z
// bits and the Z bits outside the tile (one parity per tile).

constexpr int kPauliTileMaxBits = 12;
constexpr int kPauliTileGroup = 16;  // strings of one tile pass

struct PauliTileArgs {
  int32_t K, nt;
  int32_t q[kPauliTileMaxBits];  // tile qubits, ascending
  uint32_t xl[kPauliGroup];      // X/Y bits of each string, tile-local
  uint32_t zl[kPauliGroup];      // Z/Y bits inside the tile, tile-local
  uint64_t zg[kPauliGroup];      // Z/Y bits outside the tile
  uint32_t odd;                  // bit t: string t has an odd count of bits in x & z
  uint32_t pad;
};

// A tile's amplitudes between HBM, registers and LDS: element tid + 256 r
// of a tile of E amplitudes (E <= 256 R) is (vx[r], vy[r]) in registers.
// Full tiles take a straight-line path with no per-element guard, so all
// of a tile's loads are issued back to back.
template <int R, typename T, typename C>
__device__ __forceinline__ void tile_regs_load(const T* __restrict__ st, uint64_t b, const uint64_t (&dhi)[R],
                                               C (&vx)[R], C (&vy)[R], int tid, int E) {
  if (E == 256 * R) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const T a = st[b | dhi[r]];
      vx[r] = a.x;
      vy[r] = a.y;
    }
  } else {
#pragma unroll
    for (int r = 0; r < R; ++r)
      if (r * 256 + tid < E) {
        const T a = st[b | dhi[r]];
        vx[r] = a.x;
        vy[r] = a.y;
      }
  }
}

template <int R, typename T, typename C>
__device__ __forceinline__ void tile_regs_store(T* sm, const C (&vx)[R], const C (&vy)[R], int tid, int E) {
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (E == 256 * R || r * 256 + tid < E) {
      T a;
      a.x = vx[r];
      a.y = vy[r];
      sm[r * 256 + tid] = a;
    }
}

// Sum over thread tid's pairs p = tid + 256 i of a tile of E amplitudes in
// LDS of (-1)^popc(off_i & z) w(e0 | off_i, j0 ^ off_i), off_i = 256 i with a
// zero inserted at bit h (hm = 2^h - 1), w = Re(conj(a_j) a_e), or Im for IM.
// Straight-line over the pairs of a full tile (no index guard), so the LDS
// reads of all pairs are in flight together.
template <bool IM, int R, typename T>
__device__ __forceinline__ double tile_pair_sum(const T* sm, uint32_t e0, uint32_t j0, uint32_t hm, uint32_t z,
                                                int tid, int E) {
  double s = 0.0;
  auto pair = [&](int i) __attribute__((always_inline)) {
    const uint32_t pi = 256u * i, off = ((pi & ~hm) << 1) | (pi & hm);
    // +-1.0 built from its bits: a select of the two constants is turned
    // into a branch between an add and a subtract
    const double sg = __longlong_as_double(0x3FF0000000000000ll | ((long long)(__popc(off & z) & 1) << 63));
    const T ai = sm[e0 | off], aj = sm[j0 ^ off];
    const double w = IM ? (double)aj.x * ai.y - (double)aj.y * ai.x   // Im(conj(a_j) a_i)
                        : (double)aj.x * ai.x + (double)aj.y * ai.y;  // Re(conj(a_j) a_i)
    s = fma(w, sg, s);
  };
  if (E == 256 * R) {
#pragma unroll
    for (int i = 0; i < R / 2; ++i) pair(i);
  } else {  // tiles of states below the largest tile: a short loop
#pragma unroll 1
    for (int i = 0; 256 * i < E / 2; ++i)
      if (256 * i + tid < E / 2) pair(i);
  }
  return s;
}

// Each thread holds up to R amplitudes of a tile in registers (R * 256 =
// the largest tile); all of a tile's loads are issued together, and the next
// tile's loads are in flight while the current tile's pair sums run.
template <typename T, int NT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_pauli_tile(
    const T* __restrict__ st, uint64_t ntiles, PauliTileArgs g, double* __restrict__ part) {
  constexpr int R = (sizeof(T) == 8 ? (1 << kPauliTileMaxBits) : (1 << (kPauliTileMaxBits - 1))) / 256;
  extern __shared__ unsigned char smraw[];
  T* sm = reinterpret_cast<T*>(smraw);
  const int tid = threadIdx.x, E = 1 << g.K;
  // tile element e = tid + 256 r sits at global index base | dlo | dhi[r]
  uint64_t dlo = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j)
    if (j < g.K) dlo |= (uint64_t)((tid >> j) & 1) << g.q[j];
  uint64_t dhi[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    uint64_t d = 0;
    for (int j = 8; j < g.K; ++j) d |= (uint64_t)((r >> (j - 8)) & 1) << g.q[j];
    dhi[r] = d;
  }
  auto tile_base = [&](uint64_t tile) {
    uint64_t b = tile;
    for (int j = 0; j < g.K; ++j) b = ins0(b, g.q[j]);
    return b;
  };
  // the amplitudes in flight as separate components (kept in registers)
  using C = decltype(T{}.x);
  C vx[R], vy[R];
  double acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = 0.0;
  uint64_t tile = blockIdx.x;
  if (tile < ntiles) tile_regs_load<R>(st, tile_base(tile) | dlo, dhi, vx, vy, tid, E);
  for (; tile < ntiles; tile += gridDim.x) {
    tile_regs_store<R>(sm, vx, vy, tid, E);
    __syncthreads();
    const uint64_t base = tile_base(tile);
    if (tile + gridDim.x < ntiles) tile_regs_load<R>(st, tile_base(tile + gridDim.x) | dlo, dhi, vx, vy, tid, E);
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      if (t < g.nt) {
        // Pair i of thread tid is p = tid + 256 i, its lower element e =
        // p with a zero inserted at bit h (the highest X bit).  Inserting a
        // bit maps bit positions one to one, so e = e0 | off_i with e0 =
        // ins(tid) per thread and off_i = ins(256 i) uniform over the
        // workgroup, on disjoint bits: the partner is (e0 ^ x) ^ off_i and
        // the sign parity popc(e & z) = popc(e0 & z) + popc(off_i & z).
        // The uniform part of the sign multiplies each pair's product; the
        // per-thread part is applied once.
        int h = 31 - __clz(g.xl[t]);
        asm volatile("" : "+s"(h));  // recomputed per tile, not held for every string
        const uint32_t x = g.xl[t], z = g.zl[t], hm = (1u << h) - 1;
        const uint32_t e0 = (((uint32_t)tid & ~hm) << 1) | ((uint32_t)tid & hm), j0 = e0 ^ x;
        const double s = ((g.odd >> t) & 1) ? tile_pair_sum<true, R>(sm, e0, j0, hm, z, tid, E)
                                            : tile_pair_sum<false, R>(sm, e0, j0, hm, z, tid, E);
        acc[t] += ((__popc(e0 & z) + __popcll(base & g.zg[t])) & 1) ? -s : s;
      }
    }
    __syncthreads();
  }
  __shared__ double red[NT][4];
  const int lane = tid & 63, wave = tid >> 6;
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const double w = wsum(acc[t]);
    if (lane == 0) red[t][wave] = w;
  }
  __syncthreads();
  if (tid < NT) part[(uint64_t)blockIdx.x * NT + tid] = (red[tid][0] + red[tid][1]) + (red[tid][2] + red[tid][3]);
}

template <typename T>
void launch_pauli_tile(const T* st, uint64_t ntiles, const PauliTileArgs& g, int grid, int width, double* part,
                       hipStream_t s) {
  const dim3 gr(grid), bl(256);
  const size_t lds = sizeof(T) << g.K;
  switch (width) {
    case 2: hipLaunchKernelGGL((k_pauli_tile<T, 2>), gr, bl, lds, s, st, ntiles, g, part); break;
    case 4: hipLaunchKernelGGL((k_pauli_tile<T, 4>), gr, bl, lds, s, st, ntiles, g, part); break;
    case 8: hipLaunchKernelGGL((k_pauli_tile<T, 8>), gr, bl, lds, s, st, ntiles, g, part); break;
    default: hipLaunchKernelGGL((k_pauli_tile<T, 16>), gr, bl, lds, s, st, ntiles, g, part); break;
  }
}

// out[c] = sum over rows r of in[r * width + c] for c < cols: one workgroup
// per column; thread t adds rows t, t + 256, ... in order, then a fixed tree.
__global__ __launch_bounds__(256) void k_sum_columns(const double* __restrict__ in, int rows, int width, int cols,
                                                     double* __restrict__ out) {
  const int c = blockIdx.x;
  if (c >= cols) return;
  double s = 0;
  for (int r = threadIdx.x; r < rows; r += 256) s += in[(size_t)r * width + c];
  __shared__ double red[4];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) out[c] = (red[0] + red[1]) + (red[2] + red[3]);
}

constexpr int kPauliMaxGrid = 4096;  // partial rows of one group pass: kPauliMaxGrid * kPauliGroup doubles

// A power of two (work is one), so the grid stride splits the index bits.
int pauli_grid(uint64_t work) {
  const uint64_t g = work / (kPauliThreads * 16ull);
  return (int)std::max<uint64_t>(1, std::min<uint64_t>(g, kPauliMaxGrid));
}

template <typename T>
void launch_pauli_group(const T* st, uint64_t work, const PauliGroupArgs& g, int grid, double* part,
                        hipStream_t s) {
  const int nk = g.nk;
  const dim3 gr(grid), bl(kPauliThreads);
  if (nk <= 1) hipLaunchKernelGGL((k_pauli_group<T, 1>), gr, bl, 0, s, st, work, g, part);
  else if (nk <= 2) hipLaunchKernelGGL((k_pauli_group<T, 2>), gr, bl, 0, s, st, work, g, part);
  else if (nk <= 4) hipLaunchKernelGGL((k_pauli_group<T, 4>), gr, bl, 0, s, st, work, g, part);
  else if (nk <= 6) hipLaunchKernelGGL((k_pauli_group<T, 6>), gr, bl, 0, s, st, work, g, part);
  else if (nk <= 8) hipLaunchKernelGGL((k_pauli_group<T, 8>), gr, bl, 0, s, st, work, g, part);
  else if (nk <= 12) hipLaunchKernelGGL((k_pauli_group<T, 12>), gr, bl, 0, s, st, work, g, part);
  else hipLaunchKernelGGL((k_pauli_group<T, 16>), gr, bl, 0, s, st, work, g, part);
}

int pow2_at_least(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// Launch the passes for terms (xm[t], zm[t]), sorted so that equal X masks
// are adjacent, and leave the raw sums A_t in res[t].  part holds
// kPauliMaxGrid * kPauliGroup doubles.
NqpuError pauli_launch(const State* s, const uint64_t* xm, const uint64_t* zm, uint64_t count, double* part,
                       double* res) {
  // tiles of 32 KiB of LDS: 2^12 complex64 or 2^11 complex128 amplitudes
  const int K = std::min(s->n, s->prec == NQPU_PRECISION_C64 ? kPauliTileMaxBits : kPauliTileMaxBits - 1);
  const uint64_t low = (1ull << std::min(s->n, 5)) - 1;  // the state's lowest qubits: contiguous tile loads
  for (uint64_t a = 0; a < count;) {
    // strings with at least two different X masks whose X bits fit one tile
    // (with the lowest qubits): one tile pass for all of them
    if (xm[a] != 0) {
      uint64_t qm = low, b = a, last = 0;
      int distinct = 0;
      while (b < count && b - a < (uint64_t)kPauliTileGroup) {
        if (xm[b] != last) {
          if (__builtin_popcountll(qm | xm[b]) > K) break;
          qm |= xm[b];
          last = xm[b];
          ++distinct;
        }
        ++b;
      }
      if (distinct >= 2) {
        for (int q = 0; __builtin_popcountll(qm) < K; ++q) qm |= 1ull << q;
        PauliTileArgs g{};
        g.K = K;
        g.nt = (int32_t)(b - a);
        int loc[64];
        for (int q = 0, j = 0; q < 64; ++q) {
          loc[q] = -1;
          if ((qm >> q) & 1) {
            g.q[j] = q;
            loc[q] = j++;
          }
        }
        auto local = [&](uint64_t m) {
          uint32_t r = 0;
          for (int q = 0; q < 64; ++q)
            if (((m >> q) & 1) && loc[q] >= 0) r |= 1u << loc[q];
          return r;
        };
        for (uint64_t t = a; t < b; ++t) {
          g.xl[t - a] = local(xm[t]);
          g.zl[t - a] = local(zm[t]);
          g.zg[t - a] = zm[t] & ~qm;
          if (__builtin_popcountll(xm[t] & zm[t]) & 1) g.odd |= 1u << (t - a);
        }
        const uint64_t ntiles = s->dim() >> K;
        const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>(ntiles, kPauliMaxGrid));
        const int width = std::max(2, pow2_at_least(g.nt));
        if (s->prec == NQPU_PRECISION_C64)
          launch_pauli_tile((const float2*)s->amps, ntiles, g, grid, width, part, s->stream);
        else
          launch_pauli_tile((const double2*)s->amps, ntiles, g, grid, width, part, s->stream);
        HIPTRY(hipGetLastError(), "pauli tile launch");
        hipLaunchKernelGGL(k_sum_columns, dim3(g.nt), dim3(256), 0, s->stream, (const double*)part, grid, width,
                           g.nt, res + a);
        HIPTRY(hipGetLastError(), "pauli column sums");
        a = b;
        continue;
      }
    }
    const uint64_t x = xm[a];
    PauliGroupArgs g{};
    g.x = x;
    g.h = x ? 63 - __builtin_clzll(x) : 0;
    const uint64_t work = x ? (s->dim() >> 1) : s->dim();
    const int grid = pauli_grid(work);
    // bits of the kernel's index k below the grid stride belong to the
    // thread, the bits above it to the iteration
    const uint64_t tmask = (uint64_t)grid * kPauliThreads - 1;
    g.nk = 1;
    uint64_t b = a;
    while (b < count && xm[b] == x && b - a < (uint64_t)kPauliGroup) {
      // the kernel indexes pairs by k with i = k with a zero inserted at bit
      // h, so it takes z with bit h removed
      const uint64_t z = zm[b];
      const uint64_t lo = g.h ? (z & ((1ull << g.h) - 1)) : 0;
      const uint64_t zc = x ? (lo | ((z >> (g.h + 1)) << g.h)) : z;
      const uint32_t odd = __builtin_popcountll(x & z) & 1;
      const uint64_t u = zc & ~tmask;
      int j = 0;
      if (u != 0 || odd) {
        for (j = 1; j < g.nk && !(g.u[j] == u && ((g.kodd >> j) & 1) == odd); ++j) {
        }
        if (j == g.nk) {
          if (g.nk == kPauliKeys) break;  // the next pass takes this term
          g.u[j] = u;
          g.kodd |= odd << j;
          ++g.nk;
        }
      }
      g.zt[b - a] = zc & tmask;
      g.key[b - a] = (uint32_t)j;
      ++b;
    }
    g.nt = (int32_t)(b - a);
    const int width = pow2_at_least(g.nt);
    g.width = width;
    if (s->prec == NQPU_PRECISION_C64) launch_pauli_group((const float2*)s->amps, work, g, grid, part, s->stream);
    else launch_pauli_group((const double2*)s->amps, work, g, grid, part, s->stream);
    HIPTRY(hipGetLastError(), "pauli group launch");
    // raw sums of this group, columns in term order of the group
    hipLaunchKernelGGL(k_sum_columns, dim3(g.nt), dim3(256), 0, s->stream, (const double*)part, grid, width,
                       g.nt, res + a);
    HIPTRY(hipGetLastError(), "pauli column sums");
    a = b;
  }
  return NQPU_OK;
}

// Re(i^m S) for a term whose raw sum is A (see k_pauli_group): S = 2A for
// an even count of bits in x & z (S real), S = 2iA for an odd count (S
// imaginary); for x == 0, S = A.  m is the power of i the term carries: the
// number of Y factors for a plain Pauli string.
double pauli_value(uint64_t x, uint64_t z, int m, double a) {
  const bool imag = x != 0 && (__builtin_popcountll(x & z) & 1);
  if (x != 0) a *= 2.0;
  const int e = ((imag ? m + 1 : m) % 4 + 4) % 4;  // Re(i^e * |S|-part)
  switch (e) {
    case 0: return a;
    case 2: return -a;
    default: return 0.0;
  }
}

// Term order sorted by X mask (stable), so equal masks are adjacent.
// in_place: an insertion sort that needs no temporary buffer.
void group_order(const uint64_t* xm, uint64_t count, uint32_t* idx, bool in_place = false) {
  for (uint64_t t = 0; t < count; ++t) idx[t] = (uint32_t)t;
  if (!in_place) {
    std::stable_sort(idx, idx + count, [&](uint32_t p, uint32_t q) { return xm[p] < xm[q]; });
    return;
  }
  for (uint64_t t = 1; t < count; ++t) {
    const uint32_t v = idx[t];
    uint64_t u = t;
    while (u > 0 && xm[idx[u - 1]] > xm[v]) {
      idx[u] = idx[u - 1];
      --u;
    }
    idx[u] = v;
  }
}

// ---------------------------------------------------------------- sampling

constexpr int kSubBits = 10;    // amplitudes per sub-block: 2^10
constexpr int kChunkBits = 16;  // amplitudes per chunk: 2^16

// out[b] = sum |a|^2 over sub-block b (2^sb amplitudes), one wavefront per
// sub-block: lane l adds amplitudes l, l + 64, ... (coalesced).
template <typename T>
__global__ __launch_bounds__(256) void k_subsum_wave(const T* __restrict__ st, uint64_t nsub, int sb,
                                                     double* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const int per = sb >= 6 ? 1 << (sb - 6) : 1;
  for (uint64_t b = w0; b < nsub; b += nw) {
    const T* p = st + (b << sb);
    double s = 0;
    if (sb >= 6) {
      for (int r = 0; r < per; ++r) s += nsq2(p[lane + 64 * r]);
    } else if (lane < (1 << sb)) {
      s = nsq2(p[lane]);
    }
    s = wsum(s);
    if (lane == 0) out[b] = s;
  }
}

// pre[c + 1] = chunk c's sum of its per_chunk sub-block sums: one
// wavefront per chunk, its lanes reading the sub-block sums coalesced.
__global__ __launch_bounds__(256) void k_chunk_sums(const double* __restrict__ subs, uint64_t nch, int per_chunk,
                                                    double* __restrict__ pre) {
  const uint64_t c = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);  // the same for a whole wavefront
  const int lane = threadIdx.x & 63;
  if (c >= nch) return;
  double s = 0;
  for (int b = lane; b < per_chunk; b += 64) s += subs[c * per_chunk + b];
  s = wsum(s);
  if (lane == 0) pre[c + 1] = s;
}

// pre[c + 1] = sum of chunk sums 0..c (in place over k_chunk_sums' output),
// pre[0] = 0: one workgroup of 1024 threads, each owning a contiguous run of
// chunks; the runs' offsets come from wavefront scans.
__global__ __launch_bounds__(1024) void k_chunk_prefix(uint64_t nch, double* __restrict__ pre) {
  __shared__ double wtot[16];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const uint64_t m = (nch + 1023) / 1024;
  const uint64_t c0 = std::min<uint64_t>(nch, (uint64_t)t * m), c1 = std::min<uint64_t>(nch, c0 + m);
  double s = 0;
  for (uint64_t c = c0; c < c1; ++c) s += pre[c + 1];
  const double inc = wscan(s, lane);
  if (lane == 63) wtot[wave] = inc;
  __syncthreads();
  double run = inc - s;  // exclusive within the wavefront
  for (int w = 0; w < wave; ++w) run += wtot[w];
  if (t == 0) pre[0] = 0;
  for (uint64_t c = c0; c < c1; ++c) {
    run += pre[c + 1];
    pre[c + 1] = run;
  }
}

// One wavefront per shot: first chunk c with r < pre[c + 1] (binary search),
// then the first sub-block and amplitude where the running sum exceeds r.
// r at or beyond the total mass returns dim - 1, as the C ABI's scan does.
template <typename T>
__global__ __launch_bounds__(256) void k_sample_wave(const T* __restrict__ st, int n, int cb, int sb,
                                                     const double* __restrict__ pre, const double* __restrict__ subs,
                                                     const double* __restrict__ rr, uint64_t shots,
                                                     uint64_t* __restrict__ out) {
  const int lane = threadIdx.x & 63;
  const uint64_t w0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t dim = 1ull << n, nch = dim >> cb;
  const int nsc = 1 << (cb - sb);                    // sub-blocks per chunk (<= 64)
  const int per = sb >= 6 ? 1 << (sb - 6) : 1;       // amplitudes per lane in a sub-block
  const int used = sb >= 6 ? 64 : 1 << sb;           // lanes holding amplitudes
  for (uint64_t s = w0; s < shots; s += nw) {
    const double r = rr[s];
    uint64_t lo = 0, hi = nch;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (r < pre[mid + 1]) hi = mid;
      else lo = mid + 1;
    }
    uint64_t res;
    if (lo >= nch) {
      res = dim - 1;
    } else {
      const double c0 = pre[lo];
      const uint64_t sub0 = (lo << cb) >> sb;
      const double ss = lane < nsc ? subs[sub0 + lane] : 0.0;
      const double inc = c0 + wscan(ss, lane);
      const uint64_t hit = __ballot(lane < nsc && r < inc);
      if (!hit) {
        res = ((lo + 1) << cb) - 1;  // rounding between the chunk prefix and its sub-block sums
      } else {
        const int kb = __builtin_ctzll(hit);
        const double cb0 = __shfl(inc - ss, kb);  // mass before sub-block kb
        const uint64_t b = (sub0 + kb) << sb;
        double ps[16];
        double ls = 0;
        if (lane < used) {
          for (int i = 0; i < per && i < 16; ++i) {
            ps[i] = nsq2(st[b + (uint64_t)lane * per + i]);
            ls += ps[i];
          }
        }
        const double inc2 = cb0 + wscan(ls, lane);
        const uint64_t hit2 = __ballot(lane < used && r < inc2);
        if (!hit2) {
          res = b + (1ull << sb) - 1;  // rounding between the sub-block sum and its amplitudes
        } else {
          const int kl = __builtin_ctzll(hit2);
          uint64_t found = b + (uint64_t)kl * per + per - 1;
          if (lane == kl) {
            double cum = inc2 - ls;
            for (int i = 0; i < per && i < 16; ++i) {
              cum += ps[i];
              if (r < cum) {
                found = b + (uint64_t)lane * per + i;
                break;
              }
            }
          }
          res = __shfl(found, kl);
        }
      }
    }
    if (lane == 0) out[s] = res;
  }
}

template <typename T>
__global__ void k_probs_slice(const T* __restrict__ st, uint64_t off, uint64_t count, double* __restrict__ out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x)
    out[i] = nsq2(st[off + i]);
}

// Per workgroup: sum |in_i|^2 over a grid-stride share of count amplitudes.
template <typename T>
__global__ __launch_bounds__(256) void k_sumsq_part(const T* __restrict__ in, uint64_t count, double* __restrict__ part) {
  double s = 0;
  for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < count; i += (uint64_t)gridDim.x * 256)
    s += nsq2(in[i]);
  __shared__ double red[4];
  s = wsum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

// st[off + i] = in[i] * f, converted to the state's precision.
template <typename T>
__global__ void k_import_scaled(T* __restrict__ st, uint64_t off, uint64_t count, const double2* __restrict__ in,
                                double f) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    const double2 v = in[i];
    T w;
    w.x = (decltype(w.x))(v.x * f);
    w.y = (decltype(w.y))(v.y * f);
    st[off + i] = w;
  }
}

unsigned grid_of(uint64_t threads, unsigned cap) {
  const uint64_t g = (threads + 255) / 256;
  return (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(g, cap));
}

}  // namespace

// ------------------------------------------------------------------ hosts

NqpuError pauli_values(const State* s0, const std::vector<PauliTerm>& terms, std::vector<double>* vals,
                       const std::vector<int>* phase) {
  State* s = const_cast<State*>(s0);
  const uint64_t count = terms.size();
  vals->assign(count, 0.0);
  if (!count) return NQPU_OK;
  // a Pauli string's value does not depend on the layout: its masks are
  // taken to the positions that hold its qubits
  std::vector<uint64_t> xm(count), zm(count);
  for (uint64_t t = 0; t < count; ++t) {
    xm[t] = layout_mask(s, terms[t].x);
    zm[t] = layout_mask(s, terms[t].z);
  }
  std::vector<uint32_t> idx(count);
  group_order(xm.data(), count, idx.data());
  const size_t part = (size_t)kPauliMaxGrid * kPauliGroup;
  if (!s->scratch.ensure((part + count) * sizeof(double))) return NQPU_OUT_OF_DEVICE_MEMORY;
  double* dpart = (double*)s->scratch.p;
  double* dres = dpart + part;
  std::vector<uint64_t> xs(count), zs(count);
  for (uint64_t t = 0; t < count; ++t) {
    xs[t] = xm[idx[t]];
    zs[t] = zm[idx[t]];
  }
  NqpuError e = pauli_launch(s, xs.data(), zs.data(), count, dpart, dres);
  if (e != NQPU_OK) return e;
  std::vector<double> raw(count);
  HIPTRY(hipMemcpyAsync(raw.data(), dres, count * sizeof(double), hipMemcpyDeviceToHost, s->stream),
         "pauli results copy");
  HIPTRY(hipStreamSynchronize(s->stream), "pauli sums");
  for (uint64_t t = 0; t < count; ++t) {
    const int m = phase ? (*phase)[idx[t]] : __builtin_popcountll(xs[t] & zs[t]);
    (*vals)[idx[t]] = pauli_value(xs[t], zs[t], m, raw[t]);
  }
  return NQPU_OK;
}

NqpuError pauli(const State* s, uint64_t x, uint64_t z, double* out) {
  std::vector<double> v;
  NqpuError e = pauli_values(s, {PauliTerm{x, z, 1.0}}, &v);
  if (e != NQPU_OK) return e;
  *out = v[0];
  return NQPU_OK;
}

NqpuError pauli_sum(const State* s, const uint64_t* xm, const uint64_t* zm, const double* coeffs, uint64_t count,
                    double* out) {
  std::vector<PauliTerm> terms(count);
  for (uint64_t t = 0; t < count; ++t) terms[t] = PauliTerm{xm[t], zm[t], coeffs[t]};
  std::vector<double> v;
  NqpuError e = pauli_values(s, terms, &v);
  if (e != NQPU_OK) return e;
  double acc = 0;
  for (uint64_t t = 0; t < count; ++t) acc += coeffs[t] * v[t];
  *out = acc;
  return NQPU_OK;
}

// ------------------------------------------- the no-allocation Pauli sum
//
// Partial sums and raw results live in `scratch`, one fixed array per device
// (the HIP runtime gives each device its own copy, held until the process
// ends).  Its address is looked up once per device with hipGetSymbolAddress;
// host code only ever passes that address, never the array's name, to
// kernels and copies.  Every word a call reads was written earlier in the same
// call (each group pass writes all partial rows its column sums read, and the
// column sums write every result read back), so nothing needs clearing
// between calls.  Calls that overlap on one device share the array and race.
// Size: kPauliMaxGrid x kPauliGroup partial sums + kPauliSumNoAllocMaxTerms
// results = 132096 doubles (1.0 MiB).

constexpr size_t kStaticPart = (size_t)kPauliMaxGrid * kPauliGroup;  // partial rows of one group pass
constexpr size_t kStaticScratchDoubles = kStaticPart + kPauliSumNoAllocMaxTerms;
static __device__ __align__(256) double scratch[kStaticScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

NqpuError scratch_ptr(int dev, double** out) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index out of range for the fixed scratch array");
    return NQPU_DEVICE_ERROR;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = 0;
    HIPTRY(hipGetDevice(&cur), "hipGetDevice");
    if (cur != dev) HIPTRY(hipSetDevice(dev), "hipSetDevice");
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) (void)hipSetDevice(cur);
    if (rc != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
      return NQPU_DEVICE_ERROR;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return NQPU_OK;
}
}  // namespace

NqpuError scratch_init(int dev) {
  double* p = nullptr;
  return scratch_ptr(dev, &p);
}

NqpuError pauli_sum_noalloc(const State* s, const uint64_t* xm, const uint64_t* zm, const double* coeffs,
                            uint64_t count, double* out) {
  // the caller (nqpu_expectation_pauli_sum_noalloc) has checked count <= kPauliSumNoAllocMaxTerms
  if (count == 0) {
    *out = 0.0;
    return NQPU_OK;
  }
  double* sp = nullptr;
  NqpuError e = scratch_ptr(s->device, &sp);
  if (e != NQPU_OK) return e;
  // host arrays on the stack: nothing here allocates.  The masks are taken
  // to the positions of the state's layout, as in pauli_values.
  uint32_t idx[kPauliSumNoAllocMaxTerms];
  uint64_t xl[kPauliSumNoAllocMaxTerms], zl[kPauliSumNoAllocMaxTerms];
  uint64_t xs[kPauliSumNoAllocMaxTerms], zs[kPauliSumNoAllocMaxTerms];
  double raw[kPauliSumNoAllocMaxTerms];
  for (uint64_t t = 0; t < count; ++t) {
    xl[t] = layout_mask(s, xm[t]);
    zl[t] = layout_mask(s, zm[t]);
  }
  group_order(xl, count, idx, true);
  for (uint64_t t = 0; t < count; ++t) {
    xs[t] = xl[idx[t]];
    zs[t] = zl[idx[t]];
  }
  double* dpart = sp;
  double* dres = sp + kStaticPart;
  e = pauli_launch(s, xs, zs, count, dpart, dres);
  if (e != NQPU_OK) return e;
  HIPTRY(hipMemcpyAsync(raw, dres, count * sizeof(double), hipMemcpyDeviceToHost, s->stream), "pauli results copy");
  HIPTRY(hipStreamSynchronize(s->stream), "pauli sums");
  double v[kPauliSumNoAllocMaxTerms];
  for (uint64_t t = 0; t < count; ++t) v[idx[t]] = pauli_value(xs[t], zs[t], __builtin_popcountll(xs[t] & zs[t]), raw[t]);
  double acc = 0;
  for (uint64_t t = 0; t < count; ++t) acc += coeffs[t] * v[t];
  *out = acc;
  return NQPU_OK;
}

// ------------------------------------------------------------- sampling

NqpuError sample(const State* s0, const std::vector<double>& r, uint64_t* out) {
  State* s = const_cast<State*>(s0);
  if (r.empty()) return NQPU_OK;
  // shots follow the cumulative distribution in index order
  NqpuError e = settle_state(s);
  if (e != NQPU_OK) return e;
  const int n = s->n;
  const int cb = std::min(n, kChunkBits), sb = std::min(cb, kSubBits);
  const uint64_t nch = s->dim() >> cb, nsub = s->dim() >> sb;
  const uint64_t shots = r.size();
  // scratch: [pre (nch + 1)] [subs (nsub)] [r (shots)] [out (shots)]
  const size_t bytes = (nch + 1 + nsub + shots) * 8 + shots * 8;
  if (!s->scratch.ensure(bytes)) return NQPU_OUT_OF_DEVICE_MEMORY;
  double* dpre = (double*)s->scratch.p;
  double* dsub = dpre + (nch + 1);
  double* dr = dsub + nsub;
  uint64_t* dout = (uint64_t*)(dr + shots);
  HIPTRY(hipMemcpyAsync(dr, r.data(), shots * 8, hipMemcpyHostToDevice, s->stream), "uniforms upload");
  const unsigned gsub = grid_of(nsub * 64, 8192);
  const unsigned gshot = grid_of(shots * 64, 4096);
  if (s->prec == NQPU_PRECISION_C64) {
    hipLaunchKernelGGL(k_subsum_wave<float2>, dim3(gsub), dim3(256), 0, s->stream, (const float2*)s->amps, nsub, sb,
                       dsub);
  } else {
    hipLaunchKernelGGL(k_subsum_wave<double2>, dim3(gsub), dim3(256), 0, s->stream, (const double2*)s->amps, nsub,
                       sb, dsub);
  }
  HIPTRY(hipGetLastError(), "sub-block sums");
  hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)((nch + 3) / 4)), dim3(256), 0, s->stream, (const double*)dsub, nch,
                     1 << (cb - sb), dpre);
  HIPTRY(hipGetLastError(), "chunk sums");
  hipLaunchKernelGGL(k_chunk_prefix, dim3(1), dim3(1024), 0, s->stream, nch, dpre);
  HIPTRY(hipGetLastError(), "chunk prefix");
  if (s->prec == NQPU_PRECISION_C64) {
    hipLaunchKernelGGL(k_sample_wave<float2>, dim3(gshot), dim3(256), 0, s->stream, (const float2*)s->amps, n, cb,
                       sb, (const double*)dpre, (const double*)dsub, (const double*)dr, shots, dout);
  } else {
    hipLaunchKernelGGL(k_sample_wave<double2>, dim3(gshot), dim3(256), 0, s->stream, (const double2*)s->amps, n,
                       cb, sb, (const double*)dpre, (const double*)dsub, (const double*)dr, shots, dout);
  }
  HIPTRY(hipGetLastError(), "sample search");
  HIPTRY(hipMemcpyAsync(out, dout, shots * 8, hipMemcpyDeviceToHost, s->stream), "sample copy");
  HIPTRY(hipStreamSynchronize(s->stream), "sample");
  return NQPU_OK;
}

// ------------------------------------------------------- state mutations

// initialize_from_amplitudes (sdk/rust/src/lib.rs:1918-1935): the length
// must be 2^n and the norm at least 1e-10, else the state is left as it was
// (NQPU_INVALID_PARAMETER); then the state becomes in / ||in||.  The norm is
// summed in FP64 on the device from the uploaded complex128 values; a state
// of more than 2^24 amplitudes is uploaded twice (norm, then import), so the
// old state survives a rejected call.
NqpuError initialize_amplitudes(State* s, const NqpuComplex* in, uint64_t len) {
  const uint64_t dim = s->dim();
  if (len != dim) return NQPU_INVALID_PARAMETER;
  const uint64_t slice = std::min<uint64_t>(dim, 1ull << 24);
  constexpr int kGrid = 512;
  if (!s->scratch.ensure(slice * 16 + (kGrid + 8) * 8)) return NQPU_OUT_OF_DEVICE_MEMORY;
  double2* stage = (double2*)s->scratch.p;
  double* part = (double*)((char*)s->scratch.p + slice * 16);
  double* tot = part + kGrid;
  HIPTRY(hipStreamSynchronize(s->stream), "sync");
  double norm2 = 0;
  for (uint64_t off = 0; off < dim; off += slice) {
    const uint64_t c = std::min(slice, dim - off);
    HIPTRY(hipMemcpyAsync(stage, in + off, c * 16, hipMemcpyHostToDevice, s->stream), "amplitude upload");
    hipLaunchKernelGGL(k_sumsq_part<double2>, dim3(kGrid), dim3(256), 0, s->stream, (const double2*)stage, c, part);
    hipLaunchKernelGGL(k_sum_columns, dim3(1), dim3(256), 0, s->stream, (const double*)part, kGrid, 1, 1, tot);
    HIPTRY(hipGetLastError(), "norm of the new amplitudes");
    double v = 0;
    HIPTRY(hipMemcpyAsync(&v, tot, 8, hipMemcpyDeviceToHost, s->stream), "norm copy");
    HIPTRY(hipStreamSynchronize(s->stream), "norm");
    norm2 += v;
  }
  const double nrm = std::sqrt(norm2);
  if (!(nrm >= 1e-10)) {
    set_error("initialize_from_amplitudes: norm below 1e-10");
    return NQPU_INVALID_PARAMETER;
  }
  const double inv = 1.0 / nrm;
  s->perm.clear();  // every amplitude is written, in the canonical layout
  for (uint64_t off = 0; off < dim; off += slice) {
    const uint64_t c = std::min(slice, dim - off);
    if (slice < dim)  // the single-slice case still holds the upload
      HIPTRY(hipMemcpyAsync(stage, in + off, c * 16, hipMemcpyHostToDevice, s->stream), "amplitude upload");
    const unsigned g = grid_of(c, 65536);
    if (s->prec == NQPU_PRECISION_C64)
      hipLaunchKernelGGL(k_import_scaled<float2>, dim3(g), dim3(256), 0, s->stream, (float2*)s->amps, off, c,
                         (const double2*)stage, inv);
    else
      hipLaunchKernelGGL(k_import_scaled<double2>, dim3(g), dim3(256), 0, s->stream, (double2*)s->amps, off, c,
                         (const double2*)stage, inv);
    HIPTRY(hipGetLastError(), "import");
    HIPTRY(hipStreamSynchronize(s->stream), "import");
  }
  return NQPU_OK;
}

// reset_qubit (sdk/rust/src/lib.rs:1705-1730): zero the amplitudes with bit
// q set and divide the rest by the square root of their mass when it is
// positive (FP64 mass; one reduction pass and one collapse pass).
NqpuError reset_qubit(State* s, int q) {
  double tot;
  std::vector<double> s1;
  NqpuError e = zsum_all(s, &tot, &s1);
  if (e != NQPU_OK) return e;
  const double kept = tot - s1[q];
  const double f = kept > 0.0 ? 1.0 / std::sqrt(kept) : 1.0;
  return collapse(s, q, 0, f);
}

// ---------------------------------------------------------- probabilities

NqpuError probabilities(const State* s0, double* out) {
  State* s = const_cast<State*>(s0);
  NqpuError e = settle_state(s);
  if (e != NQPU_OK) return e;
  const uint64_t dim = s->dim();
  const uint64_t slice = std::min<uint64_t>(dim, 1ull << 24);
  if (!s->scratch.ensure(slice * 8)) return NQPU_OUT_OF_DEVICE_MEMORY;
  for (uint64_t off = 0; off < dim; off += slice) {
    const uint64_t c = std::min(slice, dim - off);
    const unsigned g = grid_of(c, 65536);
    if (s->prec == NQPU_PRECISION_C64)
      hipLaunchKernelGGL(k_probs_slice<float2>, dim3(g), dim3(256), 0, s->stream, (const float2*)s->amps, off, c,
                         (double*)s->scratch.p);
    else
      hipLaunchKernelGGL(k_probs_slice<double2>, dim3(g), dim3(256), 0, s->stream, (const double2*)s->amps, off, c,
                         (double*)s->scratch.p);
    HIPTRY(hipGetLastError(), "probs");
    HIPTRY(hipMemcpyAsync(out + off, s->scratch.p, c * 8, hipMemcpyDeviceToHost, s->stream), "probs copy");
    HIPTRY(hipStreamSynchronize(s->stream), "probs");
  }
  return NQPU_OK;
}

}  // namespace nqpu
