// kernels.hip — gfx950 kernels of the statevector engine.
//
// State layout in HBM: 2^n interleaved complex amplitudes (float2 for c64,
// double2 for c128), qubit q = bit q of the index (LSB order, as the
// reference's QuantumState, sdk/rust/src/lib.rs:326-346).
//
// Two kernel families apply gates:
//  * direct kernels: one full-state (or touched-subset) sweep per operation,
//    one thread per amplitude pair / group enumerated with the reference's
//    insert-zero-bit scheme (lib.rs:575-632).  These are the per-gate path.
//  * k_tile: the multi-gate pass.  A workgroup stages a tile of 2^K
//    amplitudes (K chosen qubits, the low ones contiguous so loads coalesce)
//    in LDS, applies every operation of the pass there, and writes the tile
//    back: one HBM read + write of the state for any number of gates.
// Reductions accumulate in FP64 whatever the amplitude precision.
#include "engine.h"
#include "regtile.h"

#include <algorithm>

namespace nqpu {

template <typename R> struct V2;
template <> struct V2<float> { using T = float2; };
template <> struct V2<double> { using T = double2; };


__device__ __forceinline__ uint32_t ins0u(uint32_t v, int q) {
  uint32_t lo = v & ((1u << q) - 1);
  return ((v >> q) << (q + 1)) | lo;
}

struct Fixed {
  int32_t nf;
  int32_t p[8];
};
__device__ __forceinline__ uint64_t deposit(uint64_t k, const Fixed& f) {
  for (int j = 0; j < f.nf; ++j) k = ins0(k, f.p[j]);
  return k;
}

struct M2 { double v[8]; };
struct M4 { double v[32]; };

constexpr int kBlock = 256;

static inline unsigned grid_for(uint64_t work) {
  uint64_t g = (work + kBlock - 1) / kBlock;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return (unsigned)g;
}

static Fixed make_fixed(std::initializer_list<int> qs, uint64_t cmask) {
  std::vector<int> v(qs);
  for (int q = 0; q < 64; ++q)
    if ((cmask >> q) & 1) v.push_back(q);
  std::sort(v.begin(), v.end());
  v.erase(std::unique(v.begin(), v.end()), v.end());
  Fixed f{};
  f.nf = (int)v.size();
  for (size_t i = 0; i < v.size() && i < 8; ++i) f.p[i] = v[i];
  return f;
}

// ------------------------------------------------------------------ direct

template <typename R, typename T>
__global__ void k_init_zero(T* st, uint64_t dim) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T z;
    z.x = (i == 0) ? R(1) : R(0);
    z.y = R(0);
    st[i] = z;
  }
}

template <typename R, typename T>
__global__ void k_fill_basis(T* st, uint64_t dim, uint64_t idx) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T z;
    z.x = (i == idx) ? R(1) : R(0);
    z.y = R(0);
    st[i] = z;
  }
}

// Dense 2x2 on qubit a over the amplitude pairs whose control bits are set.
template <typename R, typename T>
__global__ void k_u1(T* st, uint64_t work, Fixed f, uint64_t cmask, int a, M2 mm) {
  const R m0r = (R)mm.v[0], m0i = (R)mm.v[1], m1r = (R)mm.v[2], m1i = (R)mm.v[3];
  const R m2r = (R)mm.v[4], m2i = (R)mm.v[5], m3r = (R)mm.v[6], m3i = (R)mm.v[7];
  const uint64_t am = 1ull << a;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i0 = deposit(k, f) | cmask, i1 = i0 | am;
    T x = st[i0], y = st[i1];
    st[i0] = mac2<R>(m0r, m0i, m1r, m1i, x, y);
    st[i1] = mac2<R>(m2r, m2i, m3r, m3i, x, y);
  }
}

template <typename R, typename T>
__global__ void k_flip(T* st, uint64_t work, Fixed f, uint64_t cmask, int a) {
  const uint64_t am = 1ull << a;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i0 = deposit(k, f) | cmask, i1 = i0 | am;
    T x = st[i0], y = st[i1];
    st[i0] = y;
    st[i1] = x;
  }
}

// Diagonal on qubit a: `only1` touches only the |1> amplitudes (p0 == 1).
template <typename R, typename T>
__global__ void k_d1(T* st, uint64_t work, Fixed f, uint64_t cmask, int a, M2 mm, int only1) {
  T p0, p1;
  p0.x = (R)mm.v[0]; p0.y = (R)mm.v[1];
  p1.x = (R)mm.v[2]; p1.y = (R)mm.v[3];
  const uint64_t am = 1ull << a;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t i0 = deposit(k, f) | cmask, i1 = i0 | am;
    if (!only1) st[i0] = cmul_(p0, st[i0]);
    st[i1] = cmul_(p1, st[i1]);
  }
}

// Dense 4x4 on (a, b), sub-index (bit_b << 1) | bit_a.
template <typename R, typename T>
__global__ void k_u2(T* st, uint64_t work, Fixed f, uint64_t cmask, int a, int b, M4 mm) {
  R m[32];
#pragma unroll
  for (int i = 0; i < 32; ++i) m[i] = (R)mm.v[i];
  const uint64_t am = 1ull << a, bm = 1ull << b;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t base = deposit(k, f) | cmask;
    uint64_t idx[4] = {base, base | am, base | bm, base | am | bm};
    T v[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) v[s] = st[idx[s]];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      T acc;
      acc.x = 0;
      acc.y = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        R mr = m[2 * (4 * r + c)], mi = m[2 * (4 * r + c) + 1];
        acc.x += mr * v[c].x - mi * v[c].y;
        acc.y += mr * v[c].y + mi * v[c].x;
      }
      st[idx[r]] = acc;
    }
  }
}

template <typename R, typename T>
__global__ void k_d2(T* st, uint64_t work, Fixed f, uint64_t cmask, int a, int b, M4 mm) {
  T p[4];
#pragma unroll
  for (int s = 0; s < 4; ++s) {
    p[s].x = (R)mm.v[2 * s];
    p[s].y = (R)mm.v[2 * s + 1];
  }
  const uint64_t am = 1ull << a, bm = 1ull << b;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t base = deposit(k, f) | cmask;
    uint64_t idx[4] = {base, base | am, base | bm, base | am | bm};
#pragma unroll
    for (int s = 0; s < 4; ++s) st[idx[s]] = cmul_(p[s], st[idx[s]]);
  }
}

template <typename R, typename T>
__global__ void k_swap(T* st, uint64_t work, Fixed f, uint64_t cmask, int a, int b) {
  const uint64_t am = 1ull << a, bm = 1ull << b;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < work;
       k += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t base = deposit(k, f) | cmask;
    T x = st[base | am], y = st[base | bm];
    st[base | am] = y;
    st[base | bm] = x;
  }
}

struct QList {
  int32_t k;
  int32_t q[16];
};

// Fused diagonal: phase LUT indexed by the bits of up to 16 qubits
// (execute_diagonal, gate_fusion.rs:905-1033).
template <typename R, typename T>
__global__ void k_diagk(T* st, uint64_t dim, QList ql, const double2* lut) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t li = 0;
    for (int j = 0; j < ql.k; ++j) li |= (uint32_t)((i >> ql.q[j]) & 1) << j;
    double2 ph = lut[li];
    T p;
    p.x = (R)ph.x;
    p.y = (R)ph.y;
    st[i] = cmul_(p, st[i]);
  }
}

template <typename T>
__global__ void k_copy(T* __restrict__ dst, const T* __restrict__ src, uint64_t dim) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

void launch_copy(void* dst, const void* src, uint64_t dim, bool c64, hipStream_t s) {
  if (c64)
    hipLaunchKernelGGL(k_copy<float2>, dim3(grid_for(dim)), dim3(kBlock), 0, s, (float2*)dst, (const float2*)src, dim);
  else
    hipLaunchKernelGGL(k_copy<double2>, dim3(grid_for(dim)), dim3(kBlock), 0, s, (double2*)dst, (const double2*)src,
                       dim);
}

template <typename R, typename T>
__global__ void k_scale(T* st, uint64_t dim, double f) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T v = st[i];
    v.x = (R)(v.x * f);
    v.y = (R)(v.y * f);
    st[i] = v;
  }
}

// -------------------------------------------------------------- tile pass
//
// Tile element t (0 <= t < 2^K) lives at global index base | dep(t), where
// dep scatters bit j of t to qubit q[j] and base scatters the tile number to
// the remaining qubits.  Thread `tid` owns t = tid + r*NT for the load and
// the store; operations re-partition the tile among threads as pairs /
// groups, so every operation ends with a barrier.

// Bit x of tile element t, or of the tile's base index for x >= kGlobalBit.
__device__ __forceinline__ bool tile_bit(uint32_t t, int x, uint64_t base) {
  return x >= kGlobalBit ? ((base >> (x - kGlobalBit)) & 1) : ((t >> x) & 1);
}

template <typename R, typename T, int NT>
__device__ __forceinline__ void tile_apply_ops(T* sm, const int K, const DevOp<R>* ops, int nops, uint64_t base) {
  const uint32_t TT = 1u << K;
  const int tid = threadIdx.x;
  for (int o = 0; o < nops; ++o) {
    const DevOp<R>& op = ops[o];
    // a control outside the tile is 0 on this tile: the op is the identity
    // here (the test is uniform over the workgroup, so skipping the barrier
    // is too)
    if ((base & op.gc) != op.gc) continue;
    const int kind = op.kind, a = op.a, b = op.b;
    const uint32_t cm = (uint32_t)op.cmask;
    if (kind == OP_U1) {
      const R m0r = op.m[0], m0i = op.m[1], m1r = op.m[2], m1i = op.m[3];
      const R m2r = op.m[4], m2i = op.m[5], m3r = op.m[6], m3i = op.m[7];
      for (uint32_t p = tid; p < (TT >> 1); p += NT) {
        uint32_t i0 = ins0u(p, a), i1 = i0 | (1u << a);
        if ((i0 & cm) != cm) continue;
        T x = sm[i0], y = sm[i1];
        sm[i0] = mac2<R>(m0r, m0i, m1r, m1i, x, y);
        sm[i1] = mac2<R>(m2r, m2i, m3r, m3i, x, y);
      }
    } else if (kind == OP_FLIP) {
      for (uint32_t p = tid; p < (TT >> 1); p += NT) {
        uint32_t i0 = ins0u(p, a), i1 = i0 | (1u << a);
        if ((i0 & cm) != cm) continue;
        T x = sm[i0], y = sm[i1];
        sm[i0] = y;
        sm[i1] = x;
      }
    } else if (kind == OP_D1) {
      T p0, p1;
      p0.x = op.m[0]; p0.y = op.m[1];
      p1.x = op.m[2]; p1.y = op.m[3];
      for (uint32_t t = tid; t < TT; t += NT) {
        if ((t & cm) != cm) continue;
        sm[t] = cmul_(tile_bit(t, a, base) ? p1 : p0, sm[t]);
      }
    } else if (kind == OP_D2) {
      for (uint32_t t = tid; t < TT; t += NT) {
        if ((t & cm) != cm) continue;
        const bool ba = tile_bit(t, a, base), bb = tile_bit(t, b, base);
        const int s = (bb ? 4 : 0) + (ba ? 2 : 0);  // index of the phase's real part
        T ph;
        ph.x = op.m[s];
        ph.y = op.m[s + 1];
        sm[t] = cmul_(ph, sm[t]);
      }
    } else if (kind == OP_U2) {
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      for (uint32_t g = tid; g < (TT >> 2); g += NT) {
        uint32_t base = ins0u(ins0u(g, lo), hi);
        if ((base & cm) != cm) continue;
        uint32_t idx[4] = {base, base | (1u << a), base | (1u << b), base | (1u << a) | (1u << b)};
        T v[4];
#pragma unroll
        for (int s = 0; s < 4; ++s) v[s] = sm[idx[s]];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          T acc;
          acc.x = 0;
          acc.y = 0;
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            const R mr = op.m[2 * (4 * r + c)], mi = op.m[2 * (4 * r + c) + 1];
            acc.x += mr * v[c].x - mi * v[c].y;
            acc.y += mr * v[c].y + mi * v[c].x;
          }
          sm[idx[r]] = acc;
        }
      }
    } else if (kind == OP_SWAP) {
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      for (uint32_t g = tid; g < (TT >> 2); g += NT) {
        uint32_t base = ins0u(ins0u(g, lo), hi);
        if ((base & cm) != cm) continue;
        uint32_t ia = base | (1u << a), ib = base | (1u << b);
        T x = sm[ia], y = sm[ib];
        sm[ia] = y;
        sm[ib] = x;
      }
    }
    __syncthreads();
  }
}

template <typename R, typename T, int NT, int EPT>
__global__ __launch_bounds__(NT) void k_tile(T* __restrict__ st, PassArgs pa, uint64_t ntiles,
                                             const DevOp<R>* __restrict__ ops) {
  extern __shared__ unsigned char smraw[];
  T* sm = reinterpret_cast<T*>(smraw);
  const int K = pa.K;
  const int tid = threadIdx.x;
  // dep() of this thread's low index bits (t & (NT-1)); EPT = 2^K / NT.
  constexpr int LB = (NT == 256) ? 8 : (NT == 128 ? 7 : 6);
  uint64_t dlo = 0;
#pragma unroll
  for (int j = 0; j < LB; ++j) dlo |= (uint64_t)((tid >> j) & 1) << pa.q[j];
  uint64_t dhi[EPT];
#pragma unroll
  for (int r = 0; r < EPT; ++r) {
    uint64_t d = 0;
    for (int j = LB; j < K; ++j) d |= (uint64_t)((r >> (j - LB)) & 1) << pa.q[j];
    dhi[r] = d;
  }
  const DevOp<R>* myops = ops + pa.op0;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint64_t base = tile;
    for (int j = 0; j < K; ++j) base = ins0(base, pa.q[j]);
    T v[EPT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) v[r] = st[base | dlo | dhi[r]];
#pragma unroll
    for (int r = 0; r < EPT; ++r) sm[tid + r * NT] = v[r];
    __syncthreads();
    tile_apply_ops<R, T, NT>(sm, K, myops, pa.nops, base);
#pragma unroll
    for (int r = 0; r < EPT; ++r) v[r] = sm[tid + r * NT];
#pragma unroll
    for (int r = 0; r < EPT; ++r) st[base | dlo | dhi[r]] = v[r];
    __syncthreads();
  }
}

// Tiles smaller than the workgroup (2^K < NT): element-per-thread loops,
// any number of tiles.
template <typename R, typename T, int NT>
__global__ __launch_bounds__(NT) void k_tile_small(T* __restrict__ st, PassArgs pa, uint64_t ntiles,
                                                   const DevOp<R>* __restrict__ ops) {
  extern __shared__ unsigned char smraw[];
  T* sm = reinterpret_cast<T*>(smraw);
  const int K = pa.K;
  const uint32_t TT = 1u << K;
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint64_t base = tile;
    for (int j = 0; j < K; ++j) base = ins0(base, pa.q[j]);
    for (uint32_t t = threadIdx.x; t < TT; t += NT) {
      uint64_t d = base;
      for (int j = 0; j < K; ++j) d |= (uint64_t)((t >> j) & 1) << pa.q[j];
      sm[t] = st[d];
    }
    __syncthreads();
    tile_apply_ops<R, T, NT>(sm, K, ops + pa.op0, pa.nops, base);
    for (uint32_t t = threadIdx.x; t < TT; t += NT) {
      uint64_t d = base;
      for (int j = 0; j < K; ++j) d |= (uint64_t)((t >> j) & 1) << pa.q[j];
      st[d] = sm[t];
    }
    __syncthreads();
  }
}

// Register tile pass interpreting the pass's register program.
template <typename R, typename T, int EB>
__global__ __launch_bounds__(256, 1) void k_tile_reg(T* __restrict__ st, PassArgs pa, uint64_t ntiles,
                                                  const RegOp<R>* __restrict__ prog,
                                                  const unsigned char* __restrict__ blob) {
  const RegOp<R>* ops = prog + pa.op0;
  // forced inline: a called body would keep the amplitudes in scratch memory
  reg_tile_pass<R, T, EB>(st, pa, ntiles, [&](T (&v)[1 << EB], T* sm, uint64_t base, int tid)
                                              __attribute__((always_inline)) {
    reg_program<R, T, EB>(v, sm, ops, pa.nops, tid, true, base, blob);
  });
}

// ------------------------------------------------------------- reductions

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block sum (result valid on thread 0).  `red` holds NT/64 doubles.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  if (l == 0) red[w] = v;
  __syncthreads();
  double s = 0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NT / 64; ++i) s += red[i];
  __syncthreads();
  return s;
}

template <typename R, typename T>
__device__ __forceinline__ double nsq(T a) {
  return (double)a.x * (double)a.x + (double)a.y * (double)a.y;
}

// ------------------------------------------------------ batched circuits

// sum_k coeff_k <P_k> contributions of this thread over an LDS tile of TT
// amplitudes in plain (unswizzled) order whose global index is base | dep(t).
__device__ double tile_terms(const float2* sm, uint32_t TT, const BatchTerm* terms, int nterm, uint64_t base,
                             int tid) {
  // |a|^2 of this thread's amplitudes t = tid + 256 r, computed once for all
  // diagonal terms (a tile holds at most 2^12 amplitudes: r < 16)
  constexpr int kRows = 16;
  const uint32_t nr = TT >= 256 ? (TT >> 8) : ((uint32_t)tid < TT ? 1u : 0u);
  double p[kRows];
  bool have_p = false;
  double acc = 0;
  for (int k = 0; k < nterm; ++k) {
    const BatchTerm tm = terms[k];
    const double gs = (__popcll(base & tm.zg) & 1) ? -1.0 : 1.0;
    double sr = 0, si = 0;
    if (tm.xl == 0) {
      if (!have_p) {
#pragma unroll
        for (int r = 0; r < kRows; ++r) {
          p[r] = 0.0;
          if ((uint32_t)r < nr) {
            const float2 a = sm[tid + 256 * r];
            p[r] = (double)a.x * a.x + (double)a.y * a.y;
          }
        }
        have_p = true;
      }
#pragma unroll
      for (int r = 0; r < kRows; ++r)
        if ((uint32_t)r < nr) sr += (__popc((uint32_t)(tid + 256 * r) & tm.zl) & 1) ? -p[r] : p[r];
    } else if (tm.zl == 0) {
      // no Z factor inside the tile: s_i = s_j = 1, so sr = 2 Re sum and si = 0
      const int h = 31 - __clz(tm.xl);
      for (uint32_t k2 = tid; k2 < (TT >> 1); k2 += 256) {
        const uint32_t i = ins0u(k2, h), j = i ^ tm.xl;
        const float2 ai = sm[i], aj = sm[j];
        sr += (double)aj.x * ai.x + (double)aj.y * ai.y;
      }
      sr *= 2.0;
    } else {
      const int h = 31 - __clz(tm.xl);
      for (uint32_t k2 = tid; k2 < (TT >> 1); k2 += 256) {
        const uint32_t i = ins0u(k2, h), j = i ^ tm.xl;
        const float2 ai = sm[i], aj = sm[j];
        const double s_i = (__popc(i & tm.zl) & 1) ? -1.0 : 1.0;
        const double s_j = (__popc(j & tm.zl) & 1) ? -1.0 : 1.0;
        const double cr = (double)aj.x * ai.x + (double)aj.y * ai.y;
        const double ci = (double)aj.x * ai.y - (double)aj.y * ai.x;
        sr += (s_i + s_j) * cr;
        si += (s_i - s_j) * ci;
      }
    }
    const double val = tm.ny == 0 ? sr : (tm.ny == 1 ? -si : (tm.ny == 2 ? -sr : si));
    acc += tm.coeff * gs * val;
  }
  return acc;
}

// +-1.0 from a parity bit, built from its bits (a select of the two
// constants feeding an FMA is turned into a branch between add and subtract)
__device__ __forceinline__ double parity_sign(uint32_t odd) {
  return __longlong_as_double(0x3FF0000000000000ll | ((long long)(odd & 1) << 63));
}

// tile_terms for a full 12-qubit tile (TT = 4096: rows r = 0..15 of 256).
// A diagonal term's sign on element tid + 256 r splits into
// popc(tid & zl) (per thread) and popc(r & zh), zh = zl >> 8 (uniform): the
// row sum folds the 16 |a|^2 rows one bit of zh at a time, 15 FMAs with
// uniform signs.  An X/Y term without Z factors in the tile sums the pairs
// e = e0 | off_i, j = e ^ xl (e0 = ins(tid), off_i = ins(256 i), a zero
// inserted at the highest X bit; disjoint bits), straight-line.  Other
// terms take tile_terms' loop.
__device__ double tile_terms_full(const float2* sm, const BatchTerm* terms, int nterm, uint64_t base, int tid) {
  constexpr uint32_t TT = 4096;
  double p[16];
  bool have_p = false;
  double acc = 0;
  for (int k = 0; k < nterm; ++k) {
    const BatchTerm tm = terms[k];
    const double gs = (__popcll(base & tm.zg) & 1) ? -1.0 : 1.0;
    double sr = 0, si = 0;
    if (tm.xl == 0) {
      if (!have_p) {
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const float2 a = sm[tid + 256 * r];
          p[r] = (double)a.x * a.x + (double)a.y * a.y;
        }
        have_p = true;
      }
      const uint32_t zh = tm.zl >> 8;
      double q[8];
#pragma unroll
      for (int r = 0; r < 8; ++r) q[r] = fma(p[2 * r + 1], parity_sign(zh), p[2 * r]);
#pragma unroll
      for (int r = 0; r < 4; ++r) q[r] = fma(q[2 * r + 1], parity_sign(zh >> 1), q[2 * r]);
#pragma unroll
      for (int r = 0; r < 2; ++r) q[r] = fma(q[2 * r + 1], parity_sign(zh >> 2), q[2 * r]);
      q[0] = fma(q[1], parity_sign(zh >> 3), q[0]);
      sr = (__popc((uint32_t)tid & tm.zl & 0xFFu) & 1) ? -q[0] : q[0];
    } else if (tm.zl == 0) {
      const int h = 31 - __clz(tm.xl);
      const uint32_t hm = (1u << h) - 1, e0 = (((uint32_t)tid & ~hm) << 1) | ((uint32_t)tid & hm),
                     j0 = e0 ^ tm.xl;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const uint32_t pi = 256u * i, off = ((pi & ~hm) << 1) | (pi & hm);
        const float2 ai = sm[e0 | off], aj = sm[j0 ^ off];
        sr += (double)aj.x * ai.x + (double)aj.y * ai.y;
      }
      sr *= 2.0;
    } else {
      const int h = 31 - __clz(tm.xl);
      for (uint32_t k2 = tid; k2 < (TT >> 1); k2 += 256) {
        const uint32_t i = ins0u(k2, h), j = i ^ tm.xl;
        const float2 ai = sm[i], aj = sm[j];
        const double s_i = (__popc(i & tm.zl) & 1) ? -1.0 : 1.0;
        const double s_j = (__popc(j & tm.zl) & 1) ? -1.0 : 1.0;
        const double cr = (double)aj.x * ai.x + (double)aj.y * ai.y;
        const double ci = (double)aj.x * ai.y - (double)aj.y * ai.x;
        sr += (s_i + s_j) * cr;
        si += (s_i - s_j) * ci;
      }
    }
    const double val = tm.ny == 0 ? sr : (tm.ny == 1 ? -si : (tm.ny == 2 ? -sr : si));
    acc += tm.coeff * gs * val;
  }
  return acc;
}

// Observable passes over 12-qubit tiles: the positions, relative to a tile's
// base index, of this thread's 16 amplitudes t = tid + 256 r (tile bit j on
// qubit q[j]), computed once per pass; then each tile is 16 independent
// loads into LDS.
constexpr int kObsTileQubits = 12;
static_assert(kObsTileQubits == 8 + kBatchEB, "observable passes use the gate passes' tile width");
static_assert(kObsTileQubits == 12, "tile_terms_full takes tiles of 4096 amplitudes");

__device__ __forceinline__ void obs_offsets(const int32_t* q, int tid, uint64_t (&off)[16]) {
  uint64_t lo = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) lo |= (uint64_t)((tid >> j) & 1) << q[j];
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    uint64_t hi = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) hi |= (uint64_t)((r >> j) & 1) << q[8 + j];
    off[r] = lo | hi;
  }
}

__device__ __forceinline__ void obs_load_tile(const float2* __restrict__ st, uint64_t base, const uint64_t (&off)[16],
                                              float2* sm, int tid) {
  float2 v[16];
#pragma unroll
  for (int r = 0; r < 16; ++r) v[r] = st[base | off[r]];
#pragma unroll
  for (int r = 0; r < 16; ++r) sm[tid + 256 * r] = v[r];
}

template <int EB>
__global__ __launch_bounds__(256, 1) void k_batch(float2* __restrict__ states, uint64_t slot_amps,
                                                  const BatchCircuit* __restrict__ circs,
                                                  int ncirc, const BatchPass* __restrict__ passes,
                                                  const RegOp<float>* __restrict__ prog,
                                                  const BatchTerm* __restrict__ terms,
                                                  const unsigned char* __restrict__ blob, unsigned* counter,
                                                  double* __restrict__ out) {
  using T = float2;
  constexpr int E = 1 << EB;
  extern __shared__ unsigned char smraw[];
  T* sm = reinterpret_cast<T*>(smraw);
  __shared__ double red[4];
  __shared__ unsigned s_c;
  const int tid = threadIdx.x;
  // The queue index is made wave-uniform (readfirstlane): with a value the
  // compiler must treat as divergent, the loop exit around the barriers
  // below is restructured per lane and lanes other than the one doing the
  // atomic keep iterating on a stale index.
  if (tid == 0) s_c = atomicAdd(counter, 1u);
  __syncthreads();
  unsigned c = __builtin_amdgcn_readfirstlane(s_c);
  __syncthreads();
  while (c < (unsigned)ncirc) {
    const BatchCircuit C = circs[c];
    double acc = 0;
    if (C.resident) {
      const BatchPass P = passes[C.pass0];
      const bool active = tid < (1 << P.tb);
      T v[E];
#pragma unroll
      for (int r = 0; r < E; ++r) {
        v[r].x = (tid == 0 && r == 0) ? 1.f : 0.f;
        v[r].y = 0.f;
      }
      reg_program<float, T, EB>(v, sm, prog + P.op0, P.nops, tid, active, 0, nullptr);
      if (active) {
#pragma unroll
        for (int r = 0; r < E; ++r) sm[tid | (r << P.tb)] = v[r];
      }
      __syncthreads();
      acc = tile_terms(sm, 1u << P.K, terms + P.term0, P.nterm, 0, tid);
      __syncthreads();
    } else {
      T* st = states + (uint64_t)blockIdx.x * slot_amps;  // this workgroup's slot
      for (int p = 0; p < C.npass; ++p) {
        const BatchPass P = passes[C.pass0 + p];
        const uint64_t ntiles = 1ull << (C.n - P.K);
        uint64_t dlo = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) dlo |= (uint64_t)((tid >> j) & 1) << P.q[j];
        for (uint64_t tile = 0; tile < ntiles; ++tile) {
          uint64_t base = tile;
#pragma unroll
          for (int j = 0; j < 8 + EB; ++j) base = ins0(base, P.q[j]);
          T v[E];
          // register r's address: its parent's plus one stride (as in
          // batch_gate_tiles)
          T* ad[E];
          ad[0] = st + (base | dlo);
          opaque_addr(ad[0]);
#pragma unroll
          for (int r = 1; r < E; ++r) {
            ad[r] = ad[r & (r - 1)] + (1ull << P.q[8 + __builtin_ctz(r)]);
            opaque_addr(ad[r]);
          }
#pragma unroll
          for (int r = 0; r < E; ++r) {
            if (p == 0) {  // the circuit starts from |0...0>: amplitude 1 at index 0
              v[r].x = (ad[r] == st) ? 1.f : 0.f;
              v[r].y = 0.f;
            } else {
              v[r] = *ad[r];
            }
            if (P.scaled) {
              v[r].x *= kMfmaScale;
              v[r].y *= kMfmaScale;
            }
          }
          reg_program<float, T, EB>(v, sm, prog + P.op0, P.nops, tid, true, base,
                                    blob + (size_t)P.blob_unit * kMfmaBlobBytes);
          ad[0] = st + (base | dlo);
          opaque_addr(ad[0]);
#pragma unroll
          for (int r = 1; r < E; ++r) {
            ad[r] = ad[r & (r - 1)] + (1ull << P.q[8 + __builtin_ctz(r)]);
            opaque_addr(ad[r]);
          }
#pragma unroll
          for (int r = 0; r < E; ++r) {
            if (P.scaled) {
              v[r].x *= 1.0f / kMfmaScale;
              v[r].y *= 1.0f / kMfmaScale;
            }
            *ad[r] = v[r];
          }
        }
        __threadfence_block();
        __syncthreads();
      }
      for (int rp = 0; rp < C.nred; ++rp) {
        const BatchPass* P = passes + C.red0 + rp;  // P->K == kObsTileQubits
        const int32_t* q = P->q;
        const uint64_t ntiles = 1ull << (C.n - kObsTileQubits);
        uint64_t off[16];
        obs_offsets(q, tid, off);
        for (uint64_t tile = 0; tile < ntiles; ++tile) {
          uint64_t base = tile;
#pragma unroll
          for (int j = 0; j < kObsTileQubits; ++j) base = ins0(base, q[j]);
          obs_load_tile(st, base, off, sm, tid);
          __syncthreads();
          acc += tile_terms_full(sm, terms + P->term0, P->nterm, base, tid);
          __syncthreads();
        }
      }
    }
    const double s = block_sum<256>(acc, red);
    if (tid == 0) {
      out[C.out] = s;
      s_c = atomicAdd(counter, 1u);
    }
    __syncthreads();
    c = __builtin_amdgcn_readfirstlane(s_c);
    __syncthreads();
  }
}

void launch_batch(void* states, uint64_t slot_amps, const BatchCircuit* circs, int ncirc, const BatchPass* passes,
                  const RegOp<float>* prog, const BatchTerm* terms, const unsigned char* blob, unsigned* counter,
                  double* out, int grid, hipStream_t s) {
  const size_t lds = sizeof(float2) << (8 + kBatchEB);
  hipLaunchKernelGGL(k_batch<kBatchEB>, dim3(grid), dim3(256), lds, s, (float2*)states, slot_amps, circs, ncirc,
                     passes, prog, terms, blob, counter, out);
}

// Tiles [tile0, tile0 + ntile) of gate pass p of a batched circuit over its
// state at st: load (pass 0 starts from |0...0>), run the pass's register
// program, store.  The same steps as the persistent kernel's passes.
template <int EB>
__device__ __forceinline__ void batch_gate_tiles(float2* __restrict__ st, const BatchPass& P, bool first,
                                                 uint64_t tile0, uint64_t ntile, const RegOp<float>* prog,
                                                 const unsigned char* blob, float2* sm, int tid) {
  using T = float2;
  constexpr int E = 1 << EB;
  uint64_t dlo = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) dlo |= (uint64_t)((tid >> j) & 1) << P.q[j];
  for (uint64_t tile = tile0; tile < tile0 + ntile; ++tile) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < 8 + EB; ++j) base = ins0(base, P.q[j]);
    T v[E];
    // register r's address: its parent's (r without its lowest set bit) plus
    // one stride (per-lane values: no 2^EB scalar offsets)
    T* ad[E];
    ad[0] = st + (base | dlo);
    opaque_addr(ad[0]);
#pragma unroll
    for (int r = 1; r < E; ++r) {
      ad[r] = ad[r & (r - 1)] + (1ull << P.q[8 + __builtin_ctz(r)]);
      opaque_addr(ad[r]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if (first) {  // the circuit starts from |0...0>: amplitude 1 at index 0
        v[r].x = (ad[r] == st) ? 1.f : 0.f;
        v[r].y = 0.f;
      } else {
        v[r] = *ad[r];
      }
      if (P.scaled) {
        v[r].x *= kMfmaScale;
        v[r].y *= kMfmaScale;
      }
    }
    reg_program<float, T, EB>(v, sm, prog + P.op0, P.nops, tid, true, base,
                              blob + (size_t)P.blob_unit * kMfmaBlobBytes);
    // the same addresses again for the stores (not kept across the program)
    ad[0] = st + (base | dlo);
    opaque_addr(ad[0]);
#pragma unroll
    for (int r = 1; r < E; ++r) {
      ad[r] = ad[r & (r - 1)] + (1ull << P.q[8 + __builtin_ctz(r)]);
      opaque_addr(ad[r]);
    }
#pragma unroll
    for (int r = 0; r < E; ++r) {
      if (P.scaled) {
        v[r].x *= 1.0f / kMfmaScale;
        v[r].y *= 1.0f / kMfmaScale;
      }
      *ad[r] = v[r];
    }
  }
}

template <int EB>
__global__ __launch_bounds__(256, 1) void k_batch_split_gate(float2* __restrict__ states,
                                                             const BatchCircuit* __restrict__ circs,
                                                             const BatchPass* __restrict__ passes,
                                                             const RegOp<float>* __restrict__ prog,
                                                             const unsigned char* __restrict__ blob,
                                                             const BatchItem* __restrict__ items) {
  extern __shared__ unsigned char smraw[];
  const BatchItem it = items[blockIdx.x];
  const BatchCircuit C = circs[it.circ];
  const BatchPass P = passes[C.pass0 + it.pass];
  batch_gate_tiles<EB>(states + C.off, P, it.pass == 0, (uint64_t)it.tile0, (uint64_t)it.ntile, prog, blob,
                       reinterpret_cast<float2*>(smraw), threadIdx.x);
}

// Observable pass it.pass of a split circuit over tiles [tile0, tile0 + ntile):
// part[blockIdx.x] = the workgroup's sum of coeff * <term> contributions.
__global__ __launch_bounds__(256) void k_batch_split_obs(const float2* __restrict__ states,
                                                         const BatchCircuit* __restrict__ circs,
                                                         const BatchPass* __restrict__ passes,
                                                         const BatchTerm* __restrict__ terms,
                                                         const BatchItem* __restrict__ items,
                                                         double* __restrict__ part) {
  extern __shared__ unsigned char smraw[];
  float2* sm = reinterpret_cast<float2*>(smraw);
  __shared__ double red[4];
  const int tid = threadIdx.x;
  const BatchItem it = items[blockIdx.x];
  const BatchCircuit C = circs[it.circ];
  const BatchPass* P = passes + C.red0 + it.pass;
  const float2* st = states + C.off;
  const int32_t* q = P->q;  // P->K == kObsTileQubits
  uint64_t off[16];
  obs_offsets(q, tid, off);
  double acc = 0;
  for (uint64_t tile = (uint64_t)it.tile0; tile < (uint64_t)(it.tile0 + it.ntile); ++tile) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < kObsTileQubits; ++j) base = ins0(base, q[j]);
    obs_load_tile(st, base, off, sm, tid);
    __syncthreads();
    acc += tile_terms_full(sm, terms + P->term0, P->nterm, base, tid);
    __syncthreads();
  }
  const double s = block_sum<256>(acc, red);
  if (tid == 0) part[blockIdx.x] = s;
}

// out[C.out] of split circuit k = the sum of its observable items' parts, in order.
__global__ void k_batch_split_out(const BatchCircuit* __restrict__ circs, int circ0, int nsplit,
                                  const int32_t* __restrict__ first, const int32_t* __restrict__ count,
                                  const double* __restrict__ part, double* __restrict__ out) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nsplit) return;
  double s = 0;
  for (int i = 0; i < count[k]; ++i) s += part[first[k] + i];
  out[circs[circ0 + k].out] = s;
}

void launch_batch_split_gate(void* states, const BatchCircuit* circs, const BatchPass* passes,
                             const RegOp<float>* prog, const unsigned char* blob, const BatchItem* items,
                             int nitems, hipStream_t s) {
  const size_t lds = sizeof(float2) << (8 + kBatchEB);
  hipLaunchKernelGGL(k_batch_split_gate<kBatchEB>, dim3(nitems), dim3(256), lds, s, (float2*)states, circs, passes,
                     prog, blob, items);
}

void launch_batch_split_obs(const void* states, const BatchCircuit* circs, const BatchPass* passes,
                            const BatchTerm* terms, const BatchItem* items, int nitems, double* part, hipStream_t s) {
  const size_t lds = sizeof(float2) << (8 + kBatchEB);
  hipLaunchKernelGGL(k_batch_split_obs, dim3(nitems), dim3(256), lds, s, (const float2*)states, circs, passes, terms,
                     items, part);
}

void launch_batch_split_out(const BatchCircuit* circs, int circ0, int nsplit, const int32_t* first,
                            const int32_t* count, const double* part, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_batch_split_out, dim3((nsplit + 255) / 256), dim3(256), 0, s, circs, circ0, nsplit, first,
                     count, part, out);
}

// Matrix-core blocks of a batch composed from their row operations (BR_*
// word streams; block b's words are [index[b], index[b + 1])) in FP64, then
// stored as complex64 16x16 matrices (row-major, row = output index): 16
// threads per block, one column each, the column in LDS.  Row operations act
// on each column alone, so the threads never wait for each other.
constexpr int kComposeBlocks = 8;  // blocks per workgroup

__device__ __forceinline__ double2 cfma2(double2 a, double2 x, double2 b, double2 y) {  // a x + b y
  return make_double2(a.x * x.x - a.y * x.y + b.x * y.x - b.y * y.y, a.x * x.y + a.y * x.x + b.x * y.y + b.y * y.x);
}

__global__ __launch_bounds__(16 * kComposeBlocks) void k_block_compose(const uint32_t* __restrict__ index,
                                                                        const uint64_t* __restrict__ words,
                                                                        float2* __restrict__ mats, uint64_t nblk) {
  __shared__ double2 sm[kComposeBlocks * 256];
  const int lb = threadIdx.x >> 4, c = threadIdx.x & 15;
  const uint64_t blk = (uint64_t)blockIdx.x * kComposeBlocks + lb;
  if (blk >= nblk) return;  // (no barriers below)
  double2* M = sm + lb * 256;  // M[16 i + c]: row i of this thread's column c
  for (int i = 0; i < 16; ++i) M[16 * i + c] = make_double2(i == c ? 1.0 : 0.0, 0.0);
  const uint64_t* w = words + index[blk];
  const uint64_t* const end = words + index[blk + 1];
  while (w < end) {
    const uint64_t h = *w++;
    const uint32_t kind = (uint32_t)(h & 0xFF), cmk = (uint32_t)((h >> 16) & 0xFFFF);
    const int pa = (int)((h >> 8) & 3), pb = (int)((h >> 12) & 3), nv = (int)((h >> 32) & 0xFF);
    const double* v = reinterpret_cast<const double*>(w);
    w += nv;
    auto val = [&](int k) { return make_double2(v[2 * k], v[2 * k + 1]); };
    switch (kind) {
      case BR_U1: {
        const double2 m0 = val(0), m1 = val(1), m2 = val(2), m3 = val(3);
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk || ((i >> pa) & 1)) continue;
          const int j = i | (1 << pa);
          const double2 x = M[16 * i + c], y = M[16 * j + c];
          M[16 * i + c] = cfma2(m0, x, m1, y);
          M[16 * j + c] = cfma2(m2, x, m3, y);
        }
        break;
      }
      case BR_U2: {
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk || ((i >> pa) & 1) || ((i >> pb) & 1)) continue;
          const int idx[4] = {i, i | (1 << pa), i | (1 << pb), i | (1 << pa) | (1 << pb)};
          double2 x[4];
          for (int t = 0; t < 4; ++t) x[t] = M[16 * idx[t] + c];
          for (int r = 0; r < 4; ++r) {
            double sr = 0, si = 0;
            for (int t = 0; t < 4; ++t) {
              const double2 m = val(4 * r + t);
              sr += m.x * x[t].x - m.y * x[t].y;
              si += m.x * x[t].y + m.y * x[t].x;
            }
            M[16 * idx[r] + c] = make_double2(sr, si);
          }
        }
        break;
      }
      case BR_FLIP:
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk || ((i >> pa) & 1)) continue;
          const int j = i | (1 << pa);
          const double2 x = M[16 * i + c];
          M[16 * i + c] = M[16 * j + c];
          M[16 * j + c] = x;
        }
        break;
      case BR_SWAP:
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk || !((i >> pa) & 1) || ((i >> pb) & 1)) continue;
          const int j = (i ^ (1 << pa)) | (1 << pb);
          const double2 x = M[16 * i + c];
          M[16 * i + c] = M[16 * j + c];
          M[16 * j + c] = x;
        }
        break;
      case BR_SCALE0:
      case BR_SCALE1:
      case BR_SCALE2:
        for (int i = 0; i < 16; ++i) {
          if ((i & cmk) != cmk) continue;
          const int k = kind == BR_SCALE0 ? 0
                        : kind == BR_SCALE1 ? ((i >> pa) & 1)
                                            : (((i >> pa) & 1) | (((i >> pb) & 1) << 1));
          const double2 z = val(k), x = M[16 * i + c];
          M[16 * i + c] = make_double2(z.x * x.x - z.y * x.y, z.x * x.y + z.y * x.x);
        }
        break;
      default: break;
    }
  }
  float2* out = mats + blk * 256;
  for (int i = 0; i < 16; ++i) {
    const double2 z = M[16 * i + c];
    out[16 * i + c] = make_float2((float)z.x, (float)z.y);
  }
}

void launch_block_compose(const uint32_t* index, const uint64_t* words, float2* mats, size_t nblk, hipStream_t s) {
  const unsigned grid = (unsigned)((nblk + kComposeBlocks - 1) / kComposeBlocks);
  hipLaunchKernelGGL(k_block_compose, dim3(grid), dim3(16 * kComposeBlocks), 0, s, index, words, mats,
                     (uint64_t)nblk);
}

// Operand blob of block blockIdx.x from its 16x16 complex64 matrix M: lane l
// of the blob holds A(l % 32, 16 s + 8 (l / 32) + i) for K-steps s = 0, 1 and
// i = 0..7 of the 32x32 real form [[re, -im], [im, re]], as 8 fp16 high
// parts then 8 fp16 low parts (x - hi, rounded to nearest even again; the
// difference is exact in float).  One wavefront per block, one lane per blob
// lane, 64 bytes written per lane.
__global__ __launch_bounds__(64) void k_blob_split(const float2* __restrict__ mats, unsigned char* __restrict__ blobs) {
  const int l = threadIdx.x, m = l % 32;
  const float2* M = mats + (size_t)blockIdx.x * 256 + 16 * (m >> 1);
  uint32_t w[16];
#pragma unroll
  for (int s = 0; s < 2; ++s)
#pragma unroll
    for (int i = 0; i < 8; i += 2) {
      uint16_t hb[2], lb[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int k = 16 * s + 8 * (l / 32) + i + t;
        const float2 u = M[k >> 1];
        const float x = (m & 1) == 0 ? ((k & 1) == 0 ? u.x : -u.y) : ((k & 1) == 0 ? u.y : u.x);
        const _Float16 h = (_Float16)x;
        const _Float16 lo = (_Float16)(x - (float)h);
        hb[t] = __builtin_bit_cast(uint16_t, h);
        lb[t] = __builtin_bit_cast(uint16_t, lo);
      }
      w[8 * s + i / 2] = hb[0] | ((uint32_t)hb[1] << 16);
      w[8 * s + 4 + i / 2] = lb[0] | ((uint32_t)lb[1] << 16);
    }
  uint4* out = reinterpret_cast<uint4*>(blobs + (size_t)blockIdx.x * kMfmaBlobBytes + l * 64);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = make_uint4(w[4 * q], w[4 * q + 1], w[4 * q + 2], w[4 * q + 3]);
}

void launch_blob_split(const float2* mats, unsigned char* blobs, size_t nblk, hipStream_t s) {
  for (size_t b0 = 0; b0 < nblk; b0 += (1u << 30)) {  // grid.x stays far below its limit
    const size_t nb = std::min<size_t>(nblk - b0, 1u << 30);
    hipLaunchKernelGGL(k_blob_split, dim3((unsigned)nb), dim3(64), 0, s, mats + b0 * 256,
                       blobs + b0 * kMfmaBlobBytes);
  }
}

// out[c] = sum |a_i|^2 over chunk c of 2^cb amplitudes.
template <typename R, typename T>
__global__ __launch_bounds__(kBlock) void k_sumsq(const T* st, int cb, uint64_t dim, double* out) {
  __shared__ double red[kBlock / 64];
  const uint64_t ch = 1ull << cb, c = blockIdx.x, beg = c * ch, end = min(beg + ch, dim);
  double s = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += kBlock) s += nsq<R>(st[i]);
  s = block_sum<kBlock>(s, red);
  if (threadIdx.x == 0) out[c] = s;
}

// Row c: [total, S_0, ..., S_{cb-1}] with S_q = sum |a_i|^2 * bit_q(i) over
// chunk c (2^cb amplitudes; cb >= 8).
template <typename R, typename T>
__global__ __launch_bounds__(kBlock) void k_zsum(const T* st, int cb, double* out) {
  __shared__ double red[kBlock / 64];
  const uint64_t beg = (uint64_t)blockIdx.x << cb;
  const int nr = 1 << (cb - 8);
  double tot = 0, hi[16];  // hi[j]: rows of the chunk with bit j set
#pragma unroll
  for (int j = 0; j < 16; ++j) hi[j] = 0;
  if (nr >= 16) {
    // 16 rows at a time, their loads issued together: bits 0..3 of a row
    // are fixed by its place in the group, the higher bits are the same for
    // the whole group (one add of the group's sum)
    for (int r = 0; r < nr; r += 16) {
      double p[16];
#pragma unroll
      for (int k = 0; k < 16; ++k) p[k] = nsq<R>(st[beg + threadIdx.x + ((uint64_t)(r + k) << 8)]);
      double g = 0;
#pragma unroll
      for (int k = 0; k < 16; ++k) {
        g += p[k];
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if ((k >> j) & 1) hi[j] += p[k];
      }
      tot += g;
#pragma unroll
      for (int j = 4; j < 16; ++j)
        if (j < cb - 8 && ((r >> j) & 1)) hi[j] += g;
    }
  } else {
    for (int r = 0; r < nr; ++r) {
      const double p = nsq<R>(st[beg + threadIdx.x + ((uint64_t)r << 8)]);
      tot += p;
#pragma unroll
      for (int j = 0; j < 16; ++j)
        if (j < cb - 8 && ((r >> j) & 1)) hi[j] += p;
    }
  }
  double* row = out + (uint64_t)blockIdx.x * (cb + 1);
  double s = block_sum<kBlock>(tot, red);
  if (threadIdx.x == 0) row[0] = s;
  for (int q = 0; q < 8; ++q) {
    s = block_sum<kBlock>(((threadIdx.x >> q) & 1) ? tot : 0.0, red);
    if (threadIdx.x == 0) row[1 + q] = s;
  }
  for (int j = 0; j < cb - 8; ++j) {
    double h = 0;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj)
      if (jj == j) h = hi[jj];
    s = block_sum<kBlock>(h, red);
    if (threadIdx.x == 0) row[9 + j] = s;
  }
}

template <typename R, typename T>
__global__ void k_to_c128(const T* st, uint64_t count, uint64_t off, double2* out) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T v = st[off + i];
    out[i] = make_double2((double)v.x, (double)v.y);
  }
}

template <typename R, typename T>
__global__ void k_from_c128(T* st, uint64_t count, uint64_t off, const double2* in) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count;
       i += (uint64_t)gridDim.x * blockDim.x) {
    double2 v = in[i];
    T w;
    w.x = (R)v.x;
    w.y = (R)v.y;
    st[off + i] = w;
  }
}

template <typename R, typename T>
__global__ void k_collapse(T* st, uint64_t dim, int q, uint32_t outcome, double renorm) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < dim;
       i += (uint64_t)gridDim.x * blockDim.x) {
    T v = st[i];
    if ((uint32_t)((i >> q) & 1) == outcome) {
      v.x = (R)(v.x * renorm);
      v.y = (R)(v.y * renorm);
    } else {
      v.x = 0;
      v.y = 0;
    }
    st[i] = v;
  }
}

// Row c = [Re, Im] of sum conj(a_i) b_i over chunk c (FP64 accumulation).
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_overlap(const TA* a, const TB* b, int cb, uint64_t dim,
                                                     double* out) {
  __shared__ double red[kBlock / 64];
  const uint64_t ch = 1ull << cb, beg = (uint64_t)blockIdx.x * ch, end = min(beg + ch, dim);
  double sr = 0, si = 0;
  for (uint64_t i = beg + threadIdx.x; i < end; i += kBlock) {
    const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
    sr += ar * br + ai * bi;
    si += ar * bi - ai * br;
  }
  double x = block_sum<kBlock>(sr, red);
  double y = block_sum<kBlock>(si, red);
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = x;
    out[2 * blockIdx.x + 1] = y;
  }
}

void launch_overlap(const void* a, bool a64, const void* b, bool b64, int n, int cb, double* out,
                    hipStream_t s) {
  const uint64_t dim = 1ull << n, nch = (dim + (1ull << cb) - 1) >> cb;
  dim3 g((unsigned)nch), bl(kBlock);
  if (a64 && b64)
    hipLaunchKernelGGL((k_overlap<float2, float2>), g, bl, 0, s, (const float2*)a, (const float2*)b, cb, dim, out);
  else if (a64)
    hipLaunchKernelGGL((k_overlap<float2, double2>), g, bl, 0, s, (const float2*)a, (const double2*)b, cb, dim, out);
  else if (b64)
    hipLaunchKernelGGL((k_overlap<double2, float2>), g, bl, 0, s, (const double2*)a, (const float2*)b, cb, dim, out);
  else
    hipLaunchKernelGGL((k_overlap<double2, double2>), g, bl, 0, s, (const double2*)a, (const double2*)b, cb, dim,
                       out);
}

__global__ void k_reduce_rows(const double* in, uint64_t rows, int width, double* out) {
  for (int w = threadIdx.x; w < width; w += blockDim.x) {
    double s = 0;
    for (uint64_t r = 0; r < rows; ++r) s += in[r * width + w];
    out[w] = s;
  }
}

// ------------------------------------------------------------ launchers

#define T_ typename V2<R>::T

template <typename R>
void Launch<R>::init_zero(void* st, int n, hipStream_t s) {
  uint64_t dim = 1ull << n;
  hipLaunchKernelGGL((k_init_zero<R, T_>), dim3(grid_for(dim)), dim3(kBlock), 0, s, (T_*)st, dim);
}
template <typename R>
void Launch<R>::fill_basis(void* st, int n, uint64_t idx, hipStream_t s) {
  uint64_t dim = 1ull << n;
  hipLaunchKernelGGL((k_fill_basis<R, T_>), dim3(grid_for(dim)), dim3(kBlock), 0, s, (T_*)st, dim, idx);
}

static M2 to_m2(const cd* m) {
  M2 r;
  for (int i = 0; i < 4; ++i) {
    r.v[2 * i] = m[i].real();
    r.v[2 * i + 1] = m[i].imag();
  }
  return r;
}
static M4 to_m4(const cd* m) {
  M4 r;
  for (int i = 0; i < 16; ++i) {
    r.v[2 * i] = m[i].real();
    r.v[2 * i + 1] = m[i].imag();
  }
  return r;
}

template <typename R>
void Launch<R>::u1(void* st, int n, int a, uint64_t cmask, const cd* m, hipStream_t s) {
  Fixed f = make_fixed({a}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  hipLaunchKernelGGL((k_u1<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a,
                     to_m2(m));
}
template <typename R>
void Launch<R>::flip(void* st, int n, int a, uint64_t cmask, hipStream_t s) {
  Fixed f = make_fixed({a}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  hipLaunchKernelGGL((k_flip<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a);
}
template <typename R>
void Launch<R>::d1(void* st, int n, int a, uint64_t cmask, cd p0, cd p1, hipStream_t s) {
  Fixed f = make_fixed({a}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  cd m[4] = {p0, p1, 0, 0};
  int only1 = (p0 == cd(1, 0)) ? 1 : 0;
  hipLaunchKernelGGL((k_d1<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a,
                     to_m2(m), only1);
}
template <typename R>
void Launch<R>::u2(void* st, int n, int a, int b, uint64_t cmask, const cd* m, hipStream_t s) {
  Fixed f = make_fixed({a, b}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  hipLaunchKernelGGL((k_u2<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a, b,
                     to_m4(m));
}
template <typename R>
void Launch<R>::d2(void* st, int n, int a, int b, uint64_t cmask, const cd* p, hipStream_t s) {
  Fixed f = make_fixed({a, b}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  cd m[16];
  for (int i = 0; i < 4; ++i) m[i] = p[i];
  hipLaunchKernelGGL((k_d2<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a, b,
                     to_m4(m));
}
template <typename R>
void Launch<R>::swap(void* st, int n, int a, int b, uint64_t cmask, hipStream_t s) {
  Fixed f = make_fixed({a, b}, cmask);
  uint64_t work = (1ull << n) >> f.nf;
  hipLaunchKernelGGL((k_swap<R, T_>), dim3(grid_for(work)), dim3(kBlock), 0, s, (T_*)st, work, f, cmask, a, b);
}
template <typename R>
void Launch<R>::diagk(void* st, int n, int k, const int32_t* q, const void* dlut, hipStream_t s) {
  QList ql{};
  ql.k = k;
  for (int i = 0; i < k && i < 16; ++i) ql.q[i] = q[i];
  uint64_t dim = 1ull << n;
  hipLaunchKernelGGL((k_diagk<R, T_>), dim3(grid_for(dim)), dim3(kBlock), 0, s, (T_*)st, dim, ql,
                     (const double2*)dlut);
}
template <typename R>
void Launch<R>::scale(void* st, int n, double f, hipStream_t s) {
  uint64_t dim = 1ull << n;
  hipLaunchKernelGGL((k_scale<R, T_>), dim3(grid_for(dim)), dim3(kBlock), 0, s, (T_*)st, dim, f);
}

template <typename R, int EPT>
static void launch_tile(void* st, const PassArgs& pa, const void* dops, hipStream_t s) {
  const uint64_t ntiles = 1ull << (pa.n - pa.K);
  const size_t lds = sizeof(T_) << pa.K;
  uint64_t g = ntiles < (1ull << 20) ? ntiles : (1ull << 20);
  auto fn = k_tile<R, T_, kBlock, EPT>;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(fn, dim3((unsigned)g), dim3(kBlock), lds, s, (T_*)st, pa, ntiles,
                     (const DevOp<R>*)dops);
}

template <typename R>
void Launch<R>::tile_pass(void* st, const PassArgs& pa, const void* dops, hipStream_t s) {
  switch (pa.K) {
    case 8: launch_tile<R, 1>(st, pa, dops, s); return;
    case 9: launch_tile<R, 2>(st, pa, dops, s); return;
    case 10: launch_tile<R, 4>(st, pa, dops, s); return;
    case 11: launch_tile<R, 8>(st, pa, dops, s); return;
    case 12: launch_tile<R, 16>(st, pa, dops, s); return;
    case 13: launch_tile<R, 32>(st, pa, dops, s); return;
    default: break;
  }
  // K < 8: tiles smaller than a workgroup
  const uint64_t ntiles = 1ull << (pa.n - pa.K);
  const uint64_t g = ntiles < (1ull << 20) ? ntiles : (1ull << 20);
  const size_t lds = sizeof(T_) << pa.K;
  hipLaunchKernelGGL((k_tile_small<R, T_, kBlock>), dim3((unsigned)g), dim3(kBlock), lds, s, (T_*)st, pa, ntiles,
                     (const DevOp<R>*)dops);
}

template <typename R, int EB>
static void launch_tile_reg(void* st, const PassArgs& pa, const void* prog, const void* blob, hipStream_t s) {
  const uint64_t ntiles = 1ull << (pa.n - pa.K);
  const uint64_t tpg = pa.tiles_per_group > 1 ? (uint64_t)pa.tiles_per_group : 1;
  const uint64_t g = std::max<uint64_t>(1, std::min<uint64_t>(ntiles / tpg, 1ull << 20));
  const size_t lds = (size_t)pa.lds_remap;
  auto fn = k_tile_reg<R, T_, EB>;
  static bool attr_set = false;
  if (!attr_set) {
    (void)hipFuncSetAttribute((const void*)fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr_set = true;
  }
  hipLaunchKernelGGL(fn, dim3((unsigned)g), dim3(kBlock), lds, s, (T_*)st, pa, ntiles, (const RegOp<R>*)prog,
                     (const unsigned char*)blob);
}

template <typename R>
void Launch<R>::tile_pass_reg(void* st, const PassArgs& pa, int eb, const void* prog, const void* blob,
                              hipStream_t s) {
  switch (eb) {
    case 3: launch_tile_reg<R, 3>(st, pa, prog, blob, s); return;
    case 4: launch_tile_reg<R, 4>(st, pa, prog, blob, s); return;
    case 5: launch_tile_reg<R, 5>(st, pa, prog, blob, s); return;
    default: return;
  }
}

template <typename R>
void Launch<R>::sumsq_chunks(const void* st, int n, int cb, double* out, hipStream_t s) {
  uint64_t dim = 1ull << n;
  uint64_t nch = (dim + (1ull << cb) - 1) >> cb;
  hipLaunchKernelGGL((k_sumsq<R, T_>), dim3((unsigned)nch), dim3(kBlock), 0, s, (const T_*)st, cb, dim, out);
}
template <typename R>
void Launch<R>::zsum_chunks(const void* st, int n, int cb, double* out, hipStream_t s) {
  uint64_t nch = (1ull << n) >> cb;
  hipLaunchKernelGGL((k_zsum<R, T_>), dim3((unsigned)nch), dim3(kBlock), 0, s, (const T_*)st, cb, out);
}
template <typename R>
void Launch<R>::to_c128(const void* st, uint64_t count, uint64_t off, void* out, hipStream_t s) {
  hipLaunchKernelGGL((k_to_c128<R, T_>), dim3(grid_for(count)), dim3(kBlock), 0, s, (const T_*)st, count, off,
                     (double2*)out);
}
template <typename R>
void Launch<R>::from_c128(void* st, uint64_t count, uint64_t off, const void* in, hipStream_t s) {
  hipLaunchKernelGGL((k_from_c128<R, T_>), dim3(grid_for(count)), dim3(kBlock), 0, s, (T_*)st, count, off,
                     (const double2*)in);
}
template <typename R>
void Launch<R>::collapse(void* st, int n, int q, uint32_t outcome, double renorm, hipStream_t s) {
  uint64_t dim = 1ull << n;
  hipLaunchKernelGGL((k_collapse<R, T_>), dim3(grid_for(dim)), dim3(kBlock), 0, s, (T_*)st, dim, q, outcome,
                     renorm);
}
void launch_reduce_rows(const double* in, uint64_t rows, int width, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(64), 0, s, in, rows, width, out);
}

template struct Launch<float>;
template struct Launch<double>;

}  // namespace nqpu
