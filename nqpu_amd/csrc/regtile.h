// regtile.h — device code of register tile passes: each thread of a 256-thread
// workgroup holds 2^EB amplitudes of a tile, and register programs (RegOp
// sequences) act on them.  Used by the interpreting kernel k_tile_reg and by
// the straight-line per-pass kernels generated at run time, whose source
// embeds devtypes.h and this file; so it depends on the HIP device headers
// only.
#pragma once

namespace nqpu {

template <typename T>
struct nqpu_is_float2 {
  static constexpr bool value = false;
};
template <>
struct nqpu_is_float2<float2> {
  static constexpr bool value = true;
};

// on ? a : b, per component: a select of whole HIP double2 values makes the
// compiler keep the amplitude array in scratch memory
template <typename T> __device__ __forceinline__ T sel_(bool on, T a, T b) {
  T r;
  r.x = on ? a.x : b.x;
  r.y = on ? a.y : b.y;
  return r;
}
template <typename T> __device__ __forceinline__ T cmul_(T a, T b) {
  T r;
  r.x = a.x * b.x - a.y * b.y;
  r.y = a.x * b.y + a.y * b.x;
  return r;
}
__device__ __forceinline__ uint64_t ins0(uint64_t v, int q) {
  uint64_t lo = v & ((1ull << q) - 1);
  return ((v >> q) << (q + 1)) | lo;
}
// m0*x + m1*y for complex m0, m1 given as (re, im) pairs
template <typename R, typename T>
__device__ __forceinline__ T mac2(R m0r, R m0i, R m1r, R m1i, T x, T y) {
  T r;
  r.x = m0r * x.x - m0i * x.y + m1r * y.x - m1i * y.y;
  r.y = m0r * x.y + m0i * x.x + m1r * y.y + m1i * y.x;
  return r;
}

// Non-temporal state access for full-state passes: every amplitude is read
// once and written once per pass, so the lines are not kept in the caches.
template <typename T>
struct NatVec2;
template <>
struct NatVec2<float2> {
  typedef float type __attribute__((ext_vector_type(2)));
};
template <>
struct NatVec2<double2> {
  typedef double type __attribute__((ext_vector_type(2)));
};
template <typename T>
__device__ __forceinline__ T load_stream(const T* p) {
  typedef typename NatVec2<T>::type V;
  const V x = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
  T r;
  r.x = x[0];
  r.y = x[1];
  return r;
}
template <typename T>
__device__ __forceinline__ void store_stream(T* p, const T& v) {
  typedef typename NatVec2<T>::type V;
  const V x = {v.x, v.y};
  __builtin_nontemporal_store(x, reinterpret_cast<V*>(p));
}

// Sorted bit positions with fixed values (targets and controls of an op).
// ---------------------------------------------------- register tile pass
//
// Each thread keeps E = 2^EB amplitudes of the tile in registers: amplitude
// r of thread tid is tile element t = sum_i bit_i(tid) << map[i] +
// sum_j bit_j(r) << map[8 + j].  A gate whose targets are register slots is
// applied in registers; controls and diagonal factors may sit on any slot
// (they only select or scale elements); everything else waits for an
// R_REMAP, which moves the tile through LDS into a new slot assignment.
// LDS addresses are XOR-swizzled so that lanes spread over banks whichever
// tile bits they vary.

__device__ __forceinline__ uint32_t lds_swz(uint32_t t) { return t ^ (((t >> 5) ^ (t >> 10)) & 31u); }

// PLAIN: 1 = the op has no controls inside the tile, 0 = it has some, -1 =
// read from the op (the interpreting kernel); compiled passes know it.
template <typename R, typename T, int EB, int J, int PLAIN = -1>
__device__ __forceinline__ void reg_u1(T (&v)[1 << EB], const RegOp<R>& op, bool on) {
  const R m0r = op.m[0], m0i = op.m[1], m1r = op.m[2], m1i = op.m[3];
  const R m2r = op.m[4], m2i = op.m[5], m3r = op.m[6], m3i = op.m[7];
  const uint32_t cr = PLAIN == 1 ? 0u : op.cr;
  if (PLAIN < 0 ? (op.ct == 0 && cr == 0) : PLAIN == 1) {  // uncontrolled: straight-line math
#pragma unroll
    for (int r0 = 0; r0 < (1 << EB); ++r0) {
      if (r0 & (1 << J)) continue;
      const int r1 = r0 | (1 << J);
      T x = v[r0], y = v[r1];
      v[r0] = mac2<R>(m0r, m0i, m1r, m1i, x, y);
      v[r1] = mac2<R>(m2r, m2i, m3r, m3i, x, y);
    }
    return;
  }
#pragma unroll
  for (int r0 = 0; r0 < (1 << EB); ++r0) {
    if (r0 & (1 << J)) continue;
    if ((r0 & cr) != cr) continue;
    const int r1 = r0 | (1 << J);
    T x = v[r0], y = v[r1];
    T nx = mac2<R>(m0r, m0i, m1r, m1i, x, y);
    T ny = mac2<R>(m2r, m2i, m3r, m3i, x, y);
    v[r0] = sel_(on, nx, x);
    v[r1] = sel_(on, ny, y);
  }
}

// PLAIN: 1 = no controls on thread slots, 0 = some, -1 = read from the op.
template <typename R, typename T, int EB, int J, int PLAIN = -1>
__device__ __forceinline__ void reg_flip(T (&v)[1 << EB], const RegOp<R>& op, bool on) {
  const uint32_t cr = op.cr;
  if (PLAIN < 0 ? op.ct == 0 : PLAIN == 1) {  // controls only on register slots: a uniform permutation
#pragma unroll
    for (int r0 = 0; r0 < (1 << EB); ++r0) {
      if (r0 & (1 << J)) continue;
      if ((r0 & cr) != cr) continue;
      const int r1 = r0 | (1 << J);
      T x = v[r0];
      v[r0] = v[r1];
      v[r1] = x;
    }
    return;
  }
#pragma unroll
  for (int r0 = 0; r0 < (1 << EB); ++r0) {
    if (r0 & (1 << J)) continue;
    if ((r0 & cr) != cr) continue;
    const int r1 = r0 | (1 << J);
    T x = v[r0], y = v[r1];
    v[r0] = sel_(on, y, x);
    v[r1] = sel_(on, x, y);
  }
}

template <typename R, typename T, int EB, int J1, int J2>
__device__ __forceinline__ void reg_u2(T (&v)[1 << EB], const RegOp<R>& op, bool on) {
  const uint32_t cr = op.cr;
#pragma unroll
  for (int r = 0; r < (1 << EB); ++r) {
    if ((r & (1 << J1)) || (r & (1 << J2))) continue;
    if ((r & cr) != cr) continue;
    const int idx[4] = {r, r | (1 << J1), r | (1 << J2), r | (1 << J1) | (1 << J2)};
    T x[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) x[s] = v[idx[s]];
#pragma unroll
    for (int row = 0; row < 4; ++row) {
      T acc;
      acc.x = 0;
      acc.y = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const R mr = op.m[2 * (4 * row + c)], mi = op.m[2 * (4 * row + c) + 1];
        acc.x += mr * x[c].x - mi * x[c].y;
        acc.y += mr * x[c].y + mi * x[c].x;
      }
      v[idx[row]] = sel_(on, acc, x[row]);
    }
  }
}

template <typename R, typename T, int EB, int J1, int J2>
__device__ __forceinline__ void reg_swap(T (&v)[1 << EB], const RegOp<R>& op, bool on) {
  const uint32_t cr = op.cr;
#pragma unroll
  for (int r = 0; r < (1 << EB); ++r) {
    if (!(r & (1 << J1)) || (r & (1 << J2))) continue;
    if ((r & cr) != cr) continue;
    const int r2 = r ^ (1 << J1) ^ (1 << J2);
    T x = v[r], y = v[r2];
    v[r] = sel_(on, y, x);
    v[r2] = sel_(on, x, y);
  }
}

template <typename R, typename T, int EB>
__device__ __forceinline__ void reg_diag(T (&v)[1 << EB], const RegOp<R>& op, int tid, bool on, uint64_t base) {
  R p0r = op.m[0], p0i = op.m[1], p1r = op.m[2], p1i = op.m[3];
  R p2r = op.m[4], p2i = op.m[5], p3r = op.m[6], p3i = op.m[7];
  const bool two = op.kind == R_D2;
  int a = op.s0, b = two ? op.s1 : 16;
  if (!two) {  // a single factor: the table does not depend on position b
    p2r = p0r; p2i = p0i;
    p3r = p1r; p3i = p1i;
  }
  // a factor on a qubit outside the tile is a fixed bit here: fold it into
  // the phase table, leaving a diagonal on the other position (or a scalar,
  // applied through register slot 0 with equal phases)
  if (b >= kGlobalBit) {
    if ((base >> (b - kGlobalBit)) & 1) {
      p0r = p2r; p0i = p2i;
      p1r = p3r; p1i = p3i;
    }
    b = 16;
    p2r = p0r; p2i = p0i;
    p3r = p1r; p3i = p1i;
  }
  if (a >= kGlobalBit) {
    if ((base >> (a - kGlobalBit)) & 1) {
      p0r = p1r; p0i = p1i;
      p2r = p3r; p2i = p3i;
    }
    a = 16;
    p1r = p0r; p1i = p0i;
    p3r = p2r; p3i = p2i;
  }
  const bool areg = a >= 16, breg = b >= 16;
  const int as = a & 15, bs = b & 15;
  const uint32_t cr = op.cr;
  if (op.ct == 0 && areg && breg) {
    // every factor depends on register bits only: the phase of each element
    // is chosen per register index, uniformly across the wavefront
#pragma unroll
    for (int r = 0; r < (1 << EB); ++r) {
      if ((r & cr) != cr) continue;
      const bool ba = (r >> as) & 1;
      const bool bb = (r >> bs) & 1;
      T ph;
      ph.x = bb ? (ba ? p3r : p2r) : (ba ? p1r : p0r);
      ph.y = bb ? (ba ? p3i : p2i) : (ba ? p1i : p0i);
      v[r] = cmul_(ph, v[r]);
    }
    return;
  }
  const bool ta = (tid >> as) & 1, tb = (tid >> bs) & 1;
#pragma unroll
  for (int r = 0; r < (1 << EB); ++r) {
    if ((r & cr) != cr) continue;
    const bool ba = areg ? ((r >> as) & 1) : ta;
    const bool bb = breg ? ((r >> bs) & 1) : tb;
    T ph;
    ph.x = bb ? (ba ? p3r : p2r) : (ba ? p1r : p0r);
    ph.y = bb ? (ba ? p3i : p2i) : (ba ? p1i : p0i);
    T nv = cmul_(ph, v[r]);
    v[r] = sel_(on, nv, v[r]);
  }
}

// `active` is false for threads beyond the tile's thread slots (small tiles
// in batched execution); they only take part in the barriers.
template <typename R, typename T, int EB>
__device__ __forceinline__ void reg_remap(T (&v)[1 << EB], T* sm, const RegOp<R>& op, int tid, bool active) {
  // lds_swz is linear over GF(2), and the thread and register parts of a tile
  // index occupy different bits, so the swizzled byte offset of register r is
  // the thread's part XOR a combination of per-register-slot parts that is
  // the same for every lane: one vector XOR per register.
  char* lds = reinterpret_cast<char*>(sm);
  // the slot maps are read as words (scalar loads), never as bytes (vector loads)
  const uint32_t* mo = reinterpret_cast<const uint32_t*>(op.map_old);
  const uint32_t* mn = reinterpret_cast<const uint32_t*>(op.map_new);
  auto byte = [](const uint32_t* m, int i) { return (m[i >> 2] >> (8 * (i & 3))) & 0xffu; };
  uint32_t tp = 0, rs[EB];
#pragma unroll
  for (int i = 0; i < 8; ++i) tp |= (uint32_t)((tid >> i) & 1) << byte(mo, i);
#pragma unroll
  for (int j = 0; j < EB; ++j) rs[j] = lds_swz(1u << byte(mo, 8 + j)) * (uint32_t)sizeof(T);
  uint32_t ts = lds_swz(tp) * (uint32_t)sizeof(T);
  if (active) {
#pragma unroll
    for (int r = 0; r < (1 << EB); ++r) {
      uint32_t u = 0;
#pragma unroll
      for (int j = 0; j < EB; ++j)
        if ((r >> j) & 1) u ^= rs[j];
      *reinterpret_cast<T*>(lds + (ts ^ u)) = v[r];
    }
  }
  __syncthreads();
  tp = 0;
#pragma unroll
  for (int i = 0; i < 8; ++i) tp |= (uint32_t)((tid >> i) & 1) << byte(mn, i);
#pragma unroll
  for (int j = 0; j < EB; ++j) rs[j] = lds_swz(1u << byte(mn, 8 + j)) * (uint32_t)sizeof(T);
  ts = lds_swz(tp) * (uint32_t)sizeof(T);
  if (active) {
#pragma unroll
    for (int r = 0; r < (1 << EB); ++r) {
      uint32_t u = 0;
#pragma unroll
      for (int j = 0; j < EB; ++j)
        if ((r >> j) & 1) u ^= rs[j];
      v[r] = *reinterpret_cast<const T*>(lds + (ts ^ u));
    }
  }
  __syncthreads();
}

#define NQPU_SLOT_CASES(FN, ...)                                      \
  switch (op.s0) {                                                    \
    case 0: FN<R, T, EB, 0>(__VA_ARGS__); break;                      \
    case 1: if constexpr (EB > 1) FN<R, T, EB, 1>(__VA_ARGS__); break; \
    case 2: if constexpr (EB > 2) FN<R, T, EB, 2>(__VA_ARGS__); break; \
    case 3: if constexpr (EB > 3) FN<R, T, EB, 3>(__VA_ARGS__); break; \
    case 4: if constexpr (EB > 4) FN<R, T, EB, 4>(__VA_ARGS__); break; \
    default: break;                                                   \
  }

template <typename R, typename T, int EB, int J1>
__device__ __forceinline__ void reg_pair_op(T (&v)[1 << EB], const RegOp<R>& op, bool on, bool is_u2) {
#define NQPU_J2(J2)                                                           \
  case J2:                                                                    \
    if constexpr (J2 < EB && J2 != J1) {                                      \
      if (is_u2) reg_u2<R, T, EB, J1, J2>(v, op, on);                         \
      else reg_swap<R, T, EB, J1, J2>(v, op, on);                             \
    }                                                                         \
    break;
  switch (op.s1) {
    NQPU_J2(0) NQPU_J2(1) NQPU_J2(2) NQPU_J2(3) NQPU_J2(4)
    default: break;
  }
#undef NQPU_J2
}

typedef _Float16 nqpu_half8 __attribute__((ext_vector_type(8)));
typedef _Float16 nqpu_half2 __attribute__((ext_vector_type(2)));
typedef float nqpu_float2v __attribute__((ext_vector_type(2)));
typedef float nqpu_f32x16 __attribute__((ext_vector_type(16)));

// x - (float)h for the low (HI = false) or high half of the packed fp16 pair
// h, exact: one mixed-precision FMA, -h * 1 + x.  hipcc does not form
// v_fma_mix_f32 from the plain expression (it emits a convert and a subtract).
template <bool HI>
__device__ __forceinline__ float sub_half(unsigned h, float x) {
  float r;
  if constexpr (HI)
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel:[1,0,0] op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
  else
    asm("v_fma_mix_f32 %0, -%1, 1.0, %2 op_sel_hi:[1,0,0]" : "=v"(r) : "v"(h), "v"(x));
  return r;
}

// R_QPHASE: the phase of a run of diagonal gates as one quadratic form Phi of
// the index bits (layout in devtypes.h).  Table entries and coefficients are
// reduced to [-1/2, 1/2) revolutions on the host; the sums per thread are
// reduced again before they are combined, so Phi(r) stays within a few
// revolutions and keeps an absolute error of a few 1e-7.  v_sin_f32 and
// v_cos_f32 take their argument in revolutions.
// The summed chunk-table rows of one R_QPHASE op for the tile at `base`:
// all row indices first, then all row loads, so their latencies overlap.
// With masks known at compile time (generated pass kernels) the index
// gathers reduce to scalar bit extractions.
template <int NCH>
__device__ __forceinline__ void qphase_rows(const float* ch, const uint64_t (&mask)[4], uint64_t base,
                                            float (&ks)[16]) {
  uint32_t p[4] = {0, 0, 0, 0};
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    int i = 0;
#pragma unroll
    for (int b = 0; b < 64; ++b)
      if ((mask[k] >> b) & 1) p[k] |= (uint32_t)((base >> b) & 1) << i++;
  }
#pragma unroll
  for (int x = 0; x < 16; ++x) ks[x] = 0.0f;
#pragma unroll
  for (int k = 0; k < NCH; ++k) {
    const float* e = ch + (size_t)k * kQPhaseChunkFloats + 16 * p[k];
#pragma unroll
    for (int x = 0; x < 16; ++x) ks[x] += e[x];
  }
}

// R_QPHASE with NCH chunk tables whose qubit masks are `mask`.
template <typename R, typename T, int EB, int NCH>
__device__ __forceinline__ void reg_qphase_n(T (&v)[1 << EB], const RegOp<R>& op, const unsigned char* blob, int tid,
                                             uint64_t base, const uint64_t (&mask)[4]) {
  const float* tab = reinterpret_cast<const float*>(blob + (size_t)op.s0 * kMfmaBlobBytes);
  float c = tab[tid];
  float l[EB];
#pragma unroll
  for (int j = 0; j < EB; ++j) l[j] = tab[256 * (1 + j) + tid];
  // qubits outside the tile: one table row per chunk, chosen by the tile's
  // base index (uniform over the workgroup)
  if constexpr (NCH > 0) {
    float ks[16];
    qphase_rows<NCH>(tab + 256 * (1 + EB), mask, base, ks);
    float kt = ks[0];
#pragma unroll
    for (int t = 0; t < 8; ++t) kt += ((tid >> t) & 1) ? ks[1 + t] : 0.0f;
    c += kt;
#pragma unroll
    for (int j = 0; j < EB; ++j) l[j] += ks[9 + j];
  }
  c -= rintf(c);
#pragma unroll
  for (int j = 0; j < EB; ++j) l[j] -= rintf(l[j]);
  // Phi over the register index: r = rest + 2^j with j its lowest set bit
  float ph[1 << EB];
  ph[0] = c;
#pragma unroll
  for (int r = 1; r < (1 << EB); ++r) {
    const int j = __builtin_ctz(r), rest = r & (r - 1);
    float x = ph[rest] + l[j];
#pragma unroll
    for (int k = j + 1; k < EB; ++k)
      if ((rest >> k) & 1) x += (float)op.m[k * (k - 1) / 2 + j];
    ph[r] = x;
  }
#pragma unroll
  for (int r = 0; r < (1 << EB); ++r) {
    const float x = ph[r] - rintf(ph[r]);
    T z;
    z.x = __builtin_amdgcn_cosf(x);
    z.y = __builtin_amdgcn_sinf(x);
    v[r] = cmul_(v[r], z);
  }
}

// R_QPHASE with the chunk masks read from the op (the interpreting kernel).
template <typename R, typename T, int EB>
__device__ __forceinline__ void reg_qphase(T (&v)[1 << EB], const RegOp<R>& op, const unsigned char* blob, int tid,
                                           uint64_t base) {
  uint64_t mask[4] = {0, 0, 0, 0};
  for (int k = 0; k < op.pad && k < 4; ++k) mask[k] = qphase_chunk(op, k);
  switch (op.pad) {
    case 0: reg_qphase_n<R, T, EB, 0>(v, op, blob, tid, base, mask); break;
    case 1: reg_qphase_n<R, T, EB, 1>(v, op, blob, tid, base, mask); break;
    case 2: reg_qphase_n<R, T, EB, 2>(v, op, blob, tid, base, mask); break;
    case 3: reg_qphase_n<R, T, EB, 3>(v, op, blob, tid, base, mask); break;
    default: reg_qphase_n<R, T, EB, 4>(v, op, blob, tid, base, mask); break;
  }
}

// R_MFMA: a dense 16x16 complex block on the matrix cores.  Per group e of
// 8 registers (register bits 3.., one bit for 12-qubit tiles and two for
// 13-qubit tiles), the wavefront's 32 columns (lane bits 0..4) are 32
// vectors of 16 amplitudes: K-step s holds the 8 reals of a3 = s, a2 = lane
// bit 5, (a0, a1) = register bits 0, 1.  Amplitudes (already scaled by
// kMfmaScale) and the matrix are split into fp16 high and low parts (high =
// round to nearest, low = the exact fp32 remainder rounded to nearest); the
// products hi*hi + hi*lo + lo*hi accumulate in FP32, which leaves a relative
// error of about 2^-22 per element.  The split costs four vector instructions
// per amplitude: two packed converts and two mixed-precision FMAs.  The 16
// accumulator values of lane l are rows 8j + 4 (l / 32) + i: exactly the
// amplitudes a0 = i / 2, a1 = lane bit 5, a2 = j & 1, a3 = j >> 1 in register
// order (a0, a2, a3).
template <typename T, int EB>
__device__ __forceinline__ void reg_mfma(T (&v)[1 << EB], const unsigned char* blob, int s0, int lane) {
  if constexpr (nqpu_is_float2<T>::value && (EB == 4 || EB == 5)) {
    const nqpu_half8* ap = reinterpret_cast<const nqpu_half8*>(blob + (size_t)s0 * kMfmaBlobBytes + lane * 64);
    const nqpu_half8 ah0 = ap[0], al0 = ap[1], ah1 = ap[2], al1 = ap[3];
#pragma unroll
    for (int e = 0; e < (1 << (EB - 3)); ++e) {
      nqpu_half8 bh[2], bl[2];
#pragma unroll
      for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const float2 x = v[8 * e + 4 * s + q];
          const nqpu_half2 h = __builtin_convertvector((nqpu_float2v){x.x, x.y}, nqpu_half2);
          const unsigned hb = __builtin_bit_cast(unsigned, h);
          const nqpu_half2 l =
              __builtin_convertvector((nqpu_float2v){sub_half<false>(hb, x.x), sub_half<true>(hb, x.y)}, nqpu_half2);
          bh[s][2 * q] = h[0];
          bh[s][2 * q + 1] = h[1];
          bl[s][2 * q] = l[0];
          bl[s][2 * q + 1] = l[1];
        }
      nqpu_f32x16 d = {};
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(al0, bh[0], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bl[0], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(al1, bh[1], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bl[1], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah0, bh[0], d, 0, 0, 0);
      d = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah1, bh[1], d, 0, 0, 0);
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        v[8 * e + j].x = d[2 * j];
        v[8 * e + j].y = d[2 * j + 1];
      }
    }
  }
}

// One 32-bit component of a register pair (lo: register bit J = 0, hi: = 1)
// under R_XCHG with lane bit B: afterwards the value at (register bit y, lane
// bit z) is the one that was at (register bit z, lane bit y).  Lane bits 4 and
// 5 use the lane-half swaps (v_permlane16/32_swap); lower bits read the
// partner lane l ^ 2^B with DPP (quad_perm xor for B <= 1, row_shr/row_shl 4
// for B = 2, row_ror 8 for B = 3).
template <int B>
__device__ __forceinline__ void xchg_bits(unsigned& lo, unsigned& hi, bool y) {
  unsigned a = lo, b = hi;
  if constexpr (B == 5) {
    const auto p = __builtin_amdgcn_permlane32_swap(a, b, false, false);
    a = p[0];
    b = p[1];
  } else if constexpr (B == 4) {
    const auto p = __builtin_amdgcn_permlane16_swap(a, b, false, false);
    a = p[0];
    b = p[1];
  } else if constexpr (B >= 2) {
    // the lanes with bit B set are whole DPP banks (4 lanes): each new value is
    // one DPP move into a copy of the value it replaces, writing only the banks
    // that take the partner's value (row_ror:N reads lane i - N of the row)
    constexpr int kSetBanks = B == 2 ? 0xA : 0xC, kClearBanks = B == 2 ? 0x5 : 0x3;
    constexpr int kDown = B == 2 ? 0x124 : 0x128, kUp = B == 2 ? 0x12C : 0x128;  // row_ror:4 / :12 / :8
    (void)y;
    const unsigned na = (unsigned)__builtin_amdgcn_update_dpp((int)a, (int)b, kDown, 0xF, kSetBanks, false);
    const unsigned nb = (unsigned)__builtin_amdgcn_update_dpp((int)b, (int)a, kUp, 0xF, kClearBanks, false);
    a = na;
    b = nb;
  } else {
    // lanes with bit B set read lane - 2^B; lanes with it clear read lane + 2^B
    // (a lane whose row has no such partner keeps an undefined value, which
    // the select below never takes)
    constexpr int kDown = B == 0 ? 0xB1 : 0x4E;  // quad_perm xor 1 / xor 2
    constexpr int kUp = kDown;
    const unsigned from_hi = (unsigned)__builtin_amdgcn_mov_dpp((int)b, kDown, 0xF, 0xF, false);
    const unsigned from_lo = (unsigned)__builtin_amdgcn_mov_dpp((int)a, kUp, 0xF, 0xF, false);
    const unsigned na = y ? from_hi : a;
    const unsigned nb = y ? b : from_lo;
    a = na;
    b = nb;
  }
  lo = a;
  hi = b;
}

template <int B>
__device__ __forceinline__ void xchg_component(float& lo, float& hi, bool y) {
  unsigned a = __float_as_uint(lo), b = __float_as_uint(hi);
  xchg_bits<B>(a, b, y);
  lo = __uint_as_float(a);
  hi = __uint_as_float(b);
}

// A double component moves as its two 32-bit halves.
template <int B>
__device__ __forceinline__ void xchg_component(double& lo, double& hi, bool y) {
  const uint64_t a = (uint64_t)__double_as_longlong(lo), b = (uint64_t)__double_as_longlong(hi);
  unsigned a0 = (unsigned)a, a1 = (unsigned)(a >> 32), b0 = (unsigned)b, b1 = (unsigned)(b >> 32);
  xchg_bits<B>(a0, b0, y);
  xchg_bits<B>(a1, b1, y);
  lo = __longlong_as_double((long long)(((uint64_t)a1 << 32) | a0));
  hi = __longlong_as_double((long long)(((uint64_t)b1 << 32) | b0));
}

// complex64 and complex128 (compiled passes: the interpreting kernel keeps
// complex128 exchanges out, see reg_program)
template <typename T, int EB, int J, int B>
__device__ __forceinline__ void reg_xchg_lane(T (&v)[1 << EB], int lane) {
  if constexpr (J < EB && B < 6) {
    const bool y = (lane >> B) & 1;
#pragma unroll
    for (int r0 = 0; r0 < (1 << EB); ++r0) {
      if (r0 & (1 << J)) continue;
      const int r1 = r0 | (1 << J);
      xchg_component<B>(v[r0].x, v[r1].x, y);
      xchg_component<B>(v[r0].y, v[r1].y, y);
    }
  }
}

template <typename T, int EB, int J, int K>
__device__ __forceinline__ void reg_xchg_reg(T (&v)[1 << EB]) {
  if constexpr (J < EB && K < EB && J != K) {
#pragma unroll
    for (int r = 0; r < (1 << EB); ++r) {
      if (!((r >> J) & 1) || ((r >> K) & 1)) continue;
      const int r2 = r ^ (1 << J) ^ (1 << K);
      const T x = v[r];
      v[r] = v[r2];
      v[r2] = x;
    }
  }
}

template <typename R, typename T, int EB, int J>
__device__ __forceinline__ void reg_xchg(T (&v)[1 << EB], const RegOp<R>& op, int lane) {
  switch (op.s1) {
    case 0: reg_xchg_lane<T, EB, J, 0>(v, lane); break;
    case 1: reg_xchg_lane<T, EB, J, 1>(v, lane); break;
    case 2: reg_xchg_lane<T, EB, J, 2>(v, lane); break;
    case 3: reg_xchg_lane<T, EB, J, 3>(v, lane); break;
    case 4: reg_xchg_lane<T, EB, J, 4>(v, lane); break;
    case 5: reg_xchg_lane<T, EB, J, 5>(v, lane); break;
    case 16: reg_xchg_reg<T, EB, J, 0>(v); break;
    case 17: reg_xchg_reg<T, EB, J, 1>(v); break;
    case 18: reg_xchg_reg<T, EB, J, 2>(v); break;
    case 19: reg_xchg_reg<T, EB, J, 3>(v); break;
    case 20: reg_xchg_reg<T, EB, J, 4>(v); break;
    default: break;
  }
}

// Run a register-tile program on v (the standalone pass and the batched
// kernel share it).
template <typename R, typename T, int EB>
__device__ __forceinline__ void reg_program(T (&v)[1 << EB], T* sm, const RegOp<R>* ops, int nops, int tid,
                                            bool active, uint64_t base, const unsigned char* blob) {
  for (int o = 0; o < nops; ++o) {
    const RegOp<R>& op = ops[o];
    // controls outside the tile: uniform over the workgroup (remaps have none)
    if ((base & op.gc) != op.gc) continue;
    const bool on = (tid & op.ct) == op.ct;
    switch (op.kind) {
      case R_REMAP: reg_remap<R, T, EB>(v, sm, op, tid, active); break;
      case R_MFMA: {
        // the block's matrix variant: bits of the tile's base index at the
        // qubits outside the tile it depends on (uniform over the workgroup)
        int var = 0;
        if (op.pad > 0) var |= (int)((base >> (op.s1 & 255)) & 1);
        if (op.pad > 1) var |= (int)((base >> (op.s1 >> 8)) & 1) << 1;
        reg_mfma<T, EB>(v, blob, op.s0 + var, tid & 63);
        break;
      }
      case R_XCHG:  // complex64 programs only (the complex128 build keeps it out: its double2 swaps went to scratch)
        if constexpr (nqpu_is_float2<T>::value) {
          NQPU_SLOT_CASES(reg_xchg, v, op, tid & 63)
        }
        break;
      case R_QPHASE: reg_qphase<R, T, EB>(v, op, blob, tid, base); break;
      case R_U1: NQPU_SLOT_CASES(reg_u1, v, op, on) break;
      case R_FLIP: NQPU_SLOT_CASES(reg_flip, v, op, on) break;
      case R_D1:
      case R_D2: reg_diag<R, T, EB>(v, op, tid, on, base); break;
      case R_U2: NQPU_SLOT_CASES(reg_pair_op, v, op, on, true) break;
      case R_SWAP: NQPU_SLOT_CASES(reg_pair_op, v, op, on, false) break;
      default: break;
    }
  }
}

// An address the compiler must treat as an unknown per-lane value (an empty
// asm statement: no instruction), so that sums built from it are not
// reassociated into per-register scalar constants.
template <typename T>
__device__ __forceinline__ void opaque_addr(T*& p) {
  asm("" : "+v"(p));
}

// One register tile pass over the state: tile `tile` of 2^(8 + EB) amplitudes
// goes to workgroup tile % gridDim.x; it is loaded in the slot order pa.ql,
// scaled for matrix-core blocks if the pass has them, transformed by
// body(v, lds, base, tid), and stored in the slot order pa.qf.
template <typename R, typename T, int EB, typename F>
__device__ __forceinline__ void reg_tile_pass(T* __restrict__ st, const PassArgs& pa, uint64_t ntiles, F&& body) {
  constexpr int E = 1 << EB;
  extern __shared__ unsigned char smraw[];
  T* sm = reinterpret_cast<T*>(smraw);
  const int tid = threadIdx.x;
  uint64_t dlo = 0, dst = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    dlo |= (uint64_t)((tid >> j) & 1) << pa.ql[j];
    dst |= (uint64_t)((tid >> j) & 1) << pa.qf[j];
  }
  for (uint64_t tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    uint64_t base = tile;
#pragma unroll
    for (int j = 0; j < 8 + EB; ++j) base = ins0(base, pa.q[j]);
    T v[E];
    if (pa.from_zero) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        uint64_t d = base | dlo;
#pragma unroll
        for (int j = 0; j < EB; ++j)
          if ((r >> j) & 1) d |= 1ull << pa.ql[8 + j];
        v[r].x = d == 0 ? (R)1 : (R)0;
        v[r].y = (R)0;
      }
    } else {
      // register r's amplitude address: the address of r without its lowest
      // set bit j, plus the stride of the qubit on register slot j (one
      // 64-bit add each).  Left to itself the compiler keeps all 2^EB offsets
      // as 64-bit scalars, which overflows the scalar registers into
      // vector-lane spills; opaque() keeps each address a per-lane value.
      T* ad[E];
      ad[0] = st + (base | dlo);
      opaque_addr(ad[0]);
      v[0] = load_stream(ad[0]);
#pragma unroll
      for (int r = 1; r < E; ++r) {
        ad[r] = ad[r & (r - 1)] + (1ull << pa.ql[8 + __builtin_ctz(r)]);
        opaque_addr(ad[r]);
        v[r] = load_stream(ad[r]);
      }
    }
    if (pa.scaled) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        v[r].x *= (R)kMfmaScale;
        v[r].y *= (R)kMfmaScale;
      }
    }
    body(v, sm, base, tid);
    if (pa.sync_store) {
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    if (pa.scaled) {
#pragma unroll
      for (int r = 0; r < E; ++r) {
        v[r].x *= (R)(1.0f / kMfmaScale);
        v[r].y *= (R)(1.0f / kMfmaScale);
      }
    }
    {
      T* ad[E];  // as for the loads, with the store positions
      ad[0] = st + (base | dst);
      opaque_addr(ad[0]);
      store_stream(ad[0], v[0]);
#pragma unroll
      for (int r = 1; r < E; ++r) {
        ad[r] = ad[r & (r - 1)] + (1ull << pa.qf[8 + __builtin_ctz(r)]);
        opaque_addr(ad[r]);
        store_stream(ad[r], v[r]);
      }
    }
  }
}

}  // namespace nqpu
