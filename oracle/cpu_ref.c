/*
 * cpu_ref.c — FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; see cpu_ref.h).
 *
 * Each function names the reference routine it restates.  Loops over pair /
 * group indices use the reference's own enumeration (insert_zero_bit(s),
 * lib.rs:575-632) and are OpenMP-parallel where the reference is rayon-parallel;
 * reductions are serial, as they are in the reference (lib.rs:407-524).
 */
#include "cpu_ref.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef int64_t i64;

static inline orc_c cmul(orc_c a, orc_c b) {
  orc_c r = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
  return r;
}
static inline orc_c cadd(orc_c a, orc_c b) {
  orc_c r = {a.re + b.re, a.im + b.im};
  return r;
}
static inline orc_c C(double re, double im) {
  orc_c r = {re, im};
  return r;
}
static inline double nsq(orc_c a) { return a.re * a.re + a.im * a.im; }

/* insert_zero_bit (lib.rs:626) */
static inline i64 ins0(i64 v, int q) {
  i64 lo = v & ((((i64)1) << q) - 1);
  return ((v >> q) << (q + 1)) | lo;
}
/* insert_zero_bits (lib.rs:575), b0 < b1 */
static inline i64 ins00(i64 v, int b0, int b1) { return ins0(ins0(v, b0), b1); }
/* insert_zero_bits_3 (lib.rs:600), sorted */
static inline i64 ins000(i64 v, int b0, int b1, int b2) {
  return ins0(ins0(ins0(v, b0), b1), b2);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
void orc_set_num_threads(int t) {
#ifdef _OPENMP
  if (t > 0) omp_set_num_threads(t);
#else
  (void)t;
#endif
}

void orc_init_zero(orc_c* a, int n) {
  i64 dim = (i64)1 << n;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < dim; i++) a[i] = C(0, 0);
  a[0] = C(1, 0);
}

/* ---------------------------------------------------------------- matrices */

/* U{theta,phi,lambda} exactly as gates.rs:136-150 / ascii_viz.rs:135-149 */
static void u3_matrix(double th, double ph, double la, orc_c* m) {
  double c = cos(th / 2.0), s = sin(th / 2.0);
  m[0] = C(c, 0.0);
  m[1] = C(-s * cos(la), -s * sin(la));
  m[2] = C(s * cos(ph), s * sin(ph));
  m[3] = C(c * cos(ph + la), c * sin(ph + la));
}

static void zero_mat(orc_c* m, int d) {
  for (int i = 0; i < d * d; i++) m[i] = C(0, 0);
}

/* GateType::matrix() — gates.rs:70-598 */
void orc_gate_matrix(const orc_gate* g, orc_c* m, int* dim) {
  const double* p = g->p;
  const double r2 = 1.0 / sqrt(2.0);
  switch (g->kind) {
    case ORC_H: *dim = 2; m[0] = C(r2, 0); m[1] = C(r2, 0); m[2] = C(r2, 0); m[3] = C(-r2, 0); return;
    case ORC_X: *dim = 2; zero_mat(m, 2); m[1] = C(1, 0); m[2] = C(1, 0); return;
    case ORC_Y: *dim = 2; zero_mat(m, 2); m[1] = C(0, -1); m[2] = C(0, 1); return;
    case ORC_Z: *dim = 2; zero_mat(m, 2); m[0] = C(1, 0); m[3] = C(-1, 0); return;
    case ORC_S: *dim = 2; zero_mat(m, 2); m[0] = C(1, 0); m[3] = C(0, 1); return;
    case ORC_T: *dim = 2; zero_mat(m, 2); m[0] = C(1, 0); m[3] = C(cos(M_PI / 4), sin(M_PI / 4)); return;
    case ORC_RX: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      *dim = 2; m[0] = C(c, 0); m[1] = C(0, -s); m[2] = C(0, -s); m[3] = C(c, 0); return;
    }
    case ORC_RY: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      *dim = 2; m[0] = C(c, 0); m[1] = C(-s, 0); m[2] = C(s, 0); m[3] = C(c, 0); return;
    }
    case ORC_RZ:
      *dim = 2; zero_mat(m, 2);
      m[0] = C(cos(-p[0] / 2), sin(-p[0] / 2)); m[3] = C(cos(p[0] / 2), sin(p[0] / 2)); return;
    case ORC_U: *dim = 2; u3_matrix(p[0], p[1], p[2], m); return;
    case ORC_SX: *dim = 2; m[0] = C(.5, .5); m[1] = C(.5, -.5); m[2] = C(.5, -.5); m[3] = C(.5, .5); return;
    case ORC_PHASE: *dim = 2; zero_mat(m, 2); m[0] = C(1, 0); m[3] = C(cos(p[0]), sin(p[0])); return;
    case ORC_CUSTOM1: *dim = 2; for (int i = 0; i < 4; i++) m[i] = C(g->m[2 * i], g->m[2 * i + 1]); return;
    case ORC_CUSTOM2: *dim = 4; for (int i = 0; i < 16; i++) m[i] = C(g->m[2 * i], g->m[2 * i + 1]); return;
    default: break;
  }
  /* 4x4 and 8x8: start from identity */
  int d = (g->kind == ORC_TOFFOLI || g->kind == ORC_CCZ || g->kind == ORC_CSWAP) ? 8 : 4;
  *dim = d;
  zero_mat(m, d);
  for (int i = 0; i < d; i++) m[i * d + i] = C(1, 0);
#define M4(r, c) m[(r) * 4 + (c)]
  switch (g->kind) {
    case ORC_CNOT: M4(2, 2) = C(0, 0); M4(3, 3) = C(0, 0); M4(2, 3) = C(1, 0); M4(3, 2) = C(1, 0); break;
    case ORC_CZ: M4(3, 3) = C(-1, 0); break;
    case ORC_SWAP: M4(1, 1) = C(0, 0); M4(2, 2) = C(0, 0); M4(1, 2) = C(1, 0); M4(2, 1) = C(1, 0); break;
    case ORC_CRX: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      M4(2, 2) = C(c, 0); M4(2, 3) = C(0, -s); M4(3, 2) = C(0, -s); M4(3, 3) = C(c, 0); break;
    }
    case ORC_CRY: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      M4(2, 2) = C(c, 0); M4(2, 3) = C(-s, 0); M4(3, 2) = C(s, 0); M4(3, 3) = C(c, 0); break;
    }
    case ORC_CRZ:
      M4(2, 2) = C(cos(-p[0] / 2), sin(-p[0] / 2)); M4(3, 3) = C(cos(p[0] / 2), sin(p[0] / 2)); break;
    case ORC_CR: M4(3, 3) = C(cos(p[0]), sin(p[0])); break;
    case ORC_ISWAP: M4(1, 1) = C(0, 0); M4(2, 2) = C(0, 0); M4(1, 2) = C(0, 1); M4(2, 1) = C(0, 1); break;
    case ORC_RXX: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      for (int i = 0; i < 4; i++) M4(i, i) = C(c, 0);
      M4(0, 3) = C(0, -s); M4(1, 2) = C(0, -s); M4(2, 1) = C(0, -s); M4(3, 0) = C(0, -s); break;
    }
    case ORC_RYY: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      for (int i = 0; i < 4; i++) M4(i, i) = C(c, 0);
      M4(0, 3) = C(0, s); M4(1, 2) = C(0, -s); M4(2, 1) = C(0, -s); M4(3, 0) = C(0, s); break;
    }
    case ORC_RZZ: {
      double c = cos(p[0] / 2), s = sin(p[0] / 2);
      M4(0, 0) = C(c, -s); M4(1, 1) = C(c, s); M4(2, 2) = C(c, s); M4(3, 3) = C(c, -s); break;
    }
    case ORC_CU: {
      double th = p[0], ph = p[1], la = p[2], ga = p[3];
      double c = cos(th / 2), s = sin(th / 2);
      M4(2, 2) = C(cos(ga) * c, sin(ga) * c);
      M4(2, 3) = C(-cos(ga + la) * s, -sin(ga + la) * s);
      M4(3, 2) = C(cos(ga + ph) * s, sin(ga + ph) * s);
      M4(3, 3) = C(cos(ga + ph + la) * c, sin(ga + ph + la) * c);
      break;
    }
    case ORC_TOFFOLI:
      m[6 * 8 + 6] = C(0, 0); m[7 * 8 + 7] = C(0, 0); m[6 * 8 + 7] = C(1, 0); m[7 * 8 + 6] = C(1, 0); break;
    case ORC_CCZ: m[7 * 8 + 7] = C(-1, 0); break;
    case ORC_CSWAP:
      m[5 * 8 + 5] = C(0, 0); m[6 * 8 + 6] = C(0, 0); m[5 * 8 + 6] = C(1, 0); m[6 * 8 + 5] = C(1, 0); break;
    default: break;
  }
#undef M4
}

/* ------------------------------------------------------------ gate kernels */

/* GateOperations::u (lib.rs:1493) with apply_unitary_chunk_scalar arithmetic
 * (simd_ops.rs:312-330). */
void orc_u1(orc_c* a, int n, int q, const orc_c* m) {
  i64 pairs = ((i64)1 << n) >> 1, st = (i64)1 << q;
#pragma omp parallel for schedule(static)
  for (i64 k = 0; k < pairs; k++) {
    i64 i = ins0(k, q), j = i | st;
    orc_c x = a[i], y = a[j];
    a[i] = C(m[0].re * x.re - m[0].im * x.im + m[1].re * y.re - m[1].im * y.im,
             m[0].re * x.im + m[0].im * x.re + m[1].re * y.im + m[1].im * y.re);
    a[j] = C(m[2].re * x.re - m[2].im * x.im + m[3].re * y.re - m[3].im * y.im,
             m[2].re * x.im + m[2].im * x.re + m[3].re * y.im + m[3].im * y.re);
  }
}

/* GateOperations::u2 (lib.rs:1536-1622): sub-index = (bit_hi<<1)|bit_lo. */
void orc_u2(orc_c* a, int n, int qlo, int qhi, const orc_c* m) {
  i64 groups = ((i64)1 << n) >> 2;
  int b0 = qlo < qhi ? qlo : qhi, b1 = qlo < qhi ? qhi : qlo;
  i64 ml = (i64)1 << qlo, mh = (i64)1 << qhi;
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; g++) {
    i64 base = ins00(g, b0, b1);
    i64 idx[4] = {base, base | ml, base | mh, base | ml | mh};
    orc_c v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
    for (int s = 0; s < 4; s++) {
      orc_c acc = C(0, 0);
      for (int k = 0; k < 4; k++) acc = cadd(acc, cmul(m[s * 4 + k], v[k]));
      a[idx[s]] = acc;
    }
  }
}

/* Controlled 2x2: apply m to the target on the control=1 half (the rows 2-3
 * block of a controlled gate's gates.rs matrix).  Used for CRY (lib.rs:1361),
 * the standard CRX and CU. */
static void cu1(orc_c* a, int n, int c, int t, const orc_c* m) {
  i64 groups = ((i64)1 << n) >> 2;
  int b0 = c < t ? c : t, b1 = c < t ? t : c;
  i64 cm = (i64)1 << c, tm = (i64)1 << t;
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; g++) {
    i64 i = ins00(g, b0, b1) | cm, j = i | tm;
    orc_c x = a[i], y = a[j];
    a[i] = cadd(cmul(m[0], x), cmul(m[1], y));
    a[j] = cadd(cmul(m[2], x), cmul(m[3], y));
  }
}

/* Multiply every amplitude whose bits in `mask` are all set by ph. */
static void phase_on_mask(orc_c* a, int n, i64 mask, orc_c ph) {
  i64 dim = (i64)1 << n;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < dim; i++)
    if ((i & mask) == mask) a[i] = cmul(a[i], ph);
}

/* Swap amplitude pairs (i, i^ma^mb) where bit a=1, bit b=0 and all control
 * bits in cmask are set.  SWAP (lib.rs:1207), CSWAP, CNOT (lib.rs:835) and
 * Toffoli (lib.rs:1251) are all instances. */
static void swap_bits(orc_c* a, int n, int qa, int qb, i64 cmask) {
  i64 dim = (i64)1 << n, ma = (i64)1 << qa, mb = (i64)1 << qb;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < dim; i++) {
    if ((i & cmask) == cmask && (i & ma) && !(i & mb)) {
      i64 j = (i ^ ma) | mb;
      orc_c t = a[i];
      a[i] = a[j];
      a[j] = t;
    }
  }
}

static void flip_bit(orc_c* a, int n, int t, i64 cmask) {
  i64 dim = (i64)1 << n, tm = (i64)1 << t;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < dim; i++) {
    if ((i & cmask) == cmask && !(i & tm)) {
      orc_c x = a[i];
      a[i] = a[i | tm];
      a[i | tm] = x;
    }
  }
}

/* GateOperations::h (lib.rs:666), apply_hadamard_chunk_scalar arithmetic. */
static void gate_h(orc_c* a, int n, int q) {
  i64 pairs = ((i64)1 << n) >> 1, st = (i64)1 << q;
  const double r = 1.0 / sqrt(2.0);
#pragma omp parallel for schedule(static)
  for (i64 k = 0; k < pairs; k++) {
    i64 i = ins0(k, q), j = i | st;
    orc_c x = a[i], y = a[j];
    a[i] = C((x.re + y.re) * r, (x.im + y.im) * r);
    a[j] = C((x.re - y.re) * r, (x.im - y.im) * r);
  }
}

/* Diagonal on one qubit: |0> *= p0, |1> *= p1 (rz lib.rs:1146; z/s/t/phase). */
static void diag1(orc_c* a, int n, int q, orc_c p0, orc_c p1) {
  i64 pairs = ((i64)1 << n) >> 1, st = (i64)1 << q;
#pragma omp parallel for schedule(static)
  for (i64 k = 0; k < pairs; k++) {
    i64 i = ins0(k, q), j = i | st;
    a[i] = cmul(p0, a[i]);
    a[j] = cmul(p1, a[j]);
  }
}

/* CRX as executed by the reference (lib.rs:1292-1358): on the control=1 half,
 * |t=0> *= (cos - i sin), |t=1> *= (cos + i sin). */
static void crx_quirk(orc_c* a, int n, int c, int t, double th) {
  i64 groups = ((i64)1 << n) >> 2;
  int b0 = c < t ? c : t, b1 = c < t ? t : c;
  i64 cm = (i64)1 << c, tm = (i64)1 << t;
  double ch = cos(th / 2), sh = sin(th / 2);
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; g++) {
    i64 i = ins00(g, b0, b1) | cm, j = i ^ tm;
    orc_c x = a[i], y = a[j];
    a[i] = C(x.re * ch + x.im * sh, x.im * ch - x.re * sh);
    a[j] = C(y.re * ch - y.im * sh, y.im * ch + y.re * sh);
  }
}

/* CRZ parallel branch (lib.rs:1439-1473): only |c=1,t=0> *= (cos + i sin). */
static void crz_quirk(orc_c* a, int n, int c, int t, double th) {
  i64 groups = ((i64)1 << n) >> 2;
  int b0 = c < t ? c : t, b1 = c < t ? t : c;
  i64 cm = (i64)1 << c;
  double ch = cos(th / 2), sh = sin(th / 2);
#pragma omp parallel for schedule(static)
  for (i64 g = 0; g < groups; g++) {
    i64 i = ins00(g, b0, b1) | cm;
    orc_c v = a[i];
    a[i] = C(v.re * ch - v.im * sh, v.im * ch + v.re * sh);
  }
}

/* apply_gate_to_state — ascii_viz.rs:124-179. */
void orc_apply_gate(orc_c* a, int n, const orc_gate* g, int quirk) {
  const double* p = g->p;
  int t0 = g->t[0], c0 = g->c[0];
  orc_c m[64];
  int d;
  switch (g->kind) {
    case ORC_H: gate_h(a, n, t0); return;
    case ORC_X: flip_bit(a, n, t0, 0); return;
    case ORC_Y: case ORC_RX: case ORC_RY: case ORC_U: case ORC_SX: case ORC_CUSTOM1:
      if (g->kind == ORC_CUSTOM1 && quirk) return; /* Custom: no-op in the dispatcher */
      orc_gate_matrix(g, m, &d);
      orc_u1(a, n, t0, m);
      return;
    case ORC_Z: diag1(a, n, t0, C(1, 0), C(-1, 0)); return;
    case ORC_S: diag1(a, n, t0, C(1, 0), C(0, 1)); return;
    case ORC_T: diag1(a, n, t0, C(1, 0), C(cos(M_PI / 4), sin(M_PI / 4))); return;
    case ORC_RZ: {
      double ch = cos(p[0] / 2), sh = sin(p[0] / 2);
      diag1(a, n, t0, C(ch, -sh), C(ch, sh));
      return;
    }
    case ORC_PHASE: diag1(a, n, t0, C(1, 0), C(cos(p[0]), sin(p[0]))); return;
    case ORC_CNOT: flip_bit(a, n, t0, (i64)1 << c0); return;
    case ORC_CZ: phase_on_mask(a, n, ((i64)1 << c0) | ((i64)1 << t0), C(-1, 0)); return;
    case ORC_CR: phase_on_mask(a, n, ((i64)1 << c0) | ((i64)1 << t0), C(cos(p[0]), sin(p[0]))); return;
    case ORC_SWAP:
      if (g->t[0] == g->t[1]) return;
      swap_bits(a, n, g->t[0], g->t[1], 0);
      return;
    case ORC_TOFFOLI: flip_bit(a, n, t0, ((i64)1 << g->c[0]) | ((i64)1 << g->c[1])); return;
    case ORC_CRX:
      if (quirk) { crx_quirk(a, n, c0, t0, p[0]); return; }
      /* fallthrough */
    case ORC_CRY: {
      orc_gate_matrix(g, m, &d);
      orc_c u[4] = {m[10], m[11], m[14], m[15]};
      cu1(a, n, c0, t0, u);
      return;
    }
    case ORC_CRZ: {
      if (quirk) { crz_quirk(a, n, c0, t0, p[0]); return; }
      double ch = cos(p[0] / 2), sh = sin(p[0] / 2);
      orc_c u[4] = {C(ch, -sh), C(0, 0), C(0, 0), C(ch, sh)};
      cu1(a, n, c0, t0, u);
      return;
    }
    case ORC_ISWAP: {
      /* lib.rs:1652-1686: |01> -> i|10>, |10> -> i|01> */
      int lo = g->t[0] < g->t[1] ? g->t[0] : g->t[1], hi = g->t[0] < g->t[1] ? g->t[1] : g->t[0];
      i64 groups = ((i64)1 << n) >> 2, sl = (i64)1 << lo, sh = (i64)1 << hi;
#pragma omp parallel for schedule(static)
      for (i64 k = 0; k < groups; k++) {
        i64 b = ins00(k, lo, hi), i01 = b | sl, i10 = b | sh;
        orc_c x = a[i01], y = a[i10];
        a[i01] = C(-y.im, y.re);
        a[i10] = C(-x.im, x.re);
      }
      return;
    }
    case ORC_CCZ:
      phase_on_mask(a, n, ((i64)1 << g->c[0]) | ((i64)1 << g->c[1]) | ((i64)1 << t0), C(-1, 0));
      return;
    case ORC_RXX: case ORC_RYY: case ORC_RZZ:
      if (quirk) return; /* ascii_viz.rs:166-168 */
      orc_gate_matrix(g, m, &d);
      /* symmetric matrices: targets[0] -> sub-index bit 0 */
      orc_u2(a, n, g->t[0], g->t[1], m);
      return;
    case ORC_CSWAP:
      if (quirk) return; /* ascii_viz.rs:169-171 */
      swap_bits(a, n, g->t[0], g->t[1], (i64)1 << c0);
      return;
    case ORC_CU: {
      if (quirk) return; /* ascii_viz.rs:172-174 */
      orc_gate_matrix(g, m, &d);
      orc_c u[4] = {m[10], m[11], m[14], m[15]};
      cu1(a, n, c0, t0, u);
      return;
    }
    case ORC_CUSTOM2:
      if (quirk) return; /* ascii_viz.rs:175-177 */
      /* targets[0] is the high bit of the 4x4 index, as c_ffi.rs apply_2q(q1,q2) */
      orc_gate_matrix(g, m, &d);
      orc_u2(a, n, g->t[1], g->t[0], m);
      return;
    default: return;
  }
}

void orc_apply_circuit(orc_c* a, int n, const orc_gate* g, int count, int quirk) {
  for (int i = 0; i < count; i++) orc_apply_gate(a, n, &g[i], quirk);
}

/* execute_diagonal (gate_fusion.rs:905-1033), LUT form. */
void orc_execute_diagonal(orc_c* a, int n, const int32_t* sq, const orc_c* sp, int ns,
                          const int32_t* pq, const orc_c* pp, int np) {
  i64 dim = (i64)1 << n;
#pragma omp parallel for schedule(static)
  for (i64 i = 0; i < dim; i++) {
    orc_c ph = C(1, 0);
    for (int k = 0; k < ns; k++) ph = cmul(ph, ((i >> sq[k]) & 1) ? sp[2 * k + 1] : sp[2 * k]);
    for (int k = 0; k < np; k++)
      if (((i >> pq[2 * k]) & 1) && ((i >> pq[2 * k + 1]) & 1)) ph = cmul(ph, pp[k]);
    a[i] = cmul(a[i], ph);
  }
}

/* -------------------------------------------------------------- reductions */

/* c_ffi.rs:243-245 norm(); serial sum */
double orc_norm(const orc_c* a, int n) {
  i64 dim = (i64)1 << n;
  double s = 0;
  for (i64 i = 0; i < dim; i++) s += nsq(a[i]);
  return s;
}

void orc_probabilities(const orc_c* a, int n, double* out) {
  i64 dim = (i64)1 << n;
  for (i64 i = 0; i < dim; i++) out[i] = nsq(a[i]);
}

/* lib.rs:407-421 */
double orc_expectation_z(const orc_c* a, int n, int q) {
  i64 dim = (i64)1 << n, st = (i64)1 << q;
  double e = 0;
  for (i64 i = 0; i < dim; i++) e += (i & st) ? -nsq(a[i]) : nsq(a[i]);
  return e;
}

/* lib.rs:425-440 */
double orc_expectation_x(const orc_c* a, int n, int q) {
  i64 dim = (i64)1 << n, st = (i64)1 << q;
  double e = 0;
  for (i64 i = 0; i < dim; i++)
    if (!(i & st)) {
      orc_c x = a[i], y = a[i | st];
      e += 2.0 * (x.re * y.re + x.im * y.im);
    }
  return e;
}

/* lib.rs:444-460 computes 2*Im(conj(a1)*a0), which is -<Y>; the standard
 * value is 2*Im(conj(a0)*a1). */
double orc_expectation_y(const orc_c* a, int n, int q, int quirk) {
  i64 dim = (i64)1 << n, st = (i64)1 << q;
  double e = 0;
  for (i64 i = 0; i < dim; i++)
    if (!(i & st)) {
      orc_c x = a[i], y = a[i | st];
      e += 2.0 * (y.re * x.im - y.im * x.re);
    }
  return quirk ? e : -e;
}

/* expectation_pauli_string (lib.rs:465-514).  The reference accumulates
 * conj(a_i) * coeff(i) * a_{i^x}; the standard value (primitives.rs:683-728)
 * is conj(a_{i^x}) * coeff(i) * a_i.  They differ by (-1)^{#Y}. */
double orc_expectation_pauli(const orc_c* a, int n, uint64_t xmask, uint64_t zmask, int quirk) {
  i64 dim = (i64)1 << n;
  double r = 0;
  for (i64 i = 0; i < dim; i++) {
    i64 j = i ^ (i64)xmask;
    double cr = 1, ci = 0;
    for (int q = 0; q < n; q++) {
      int xb = (xmask >> q) & 1, zb = (zmask >> q) & 1, bit = (i >> q) & 1;
      if (xb && zb) { /* Y */
        double nr, ni;
        if (bit == 0) { nr = -ci; ni = cr; } else { nr = ci; ni = -cr; }
        cr = nr; ci = ni;
      } else if (zb) {
        if (bit) { cr = -cr; ci = -ci; }
      }
    }
    if (quirk) {
      orc_c ai = a[i], aj = a[j];
      double pr = cr * aj.re - ci * aj.im, pi = cr * aj.im + ci * aj.re;
      r += ai.re * pr + ai.im * pi;
    } else {
      orc_c ak = a[j], ai = a[i];
      double pr = cr * ai.re - ci * ai.im, pi = cr * ai.im + ci * ai.re;
      r += ak.re * pr + ak.im * pi; /* Re(conj(a_{i^x}) * coeff * a_i) */
    }
  }
  return r;
}

/* lib.rs:388-403 */
double orc_fidelity(const orc_c* a, const orc_c* b, int n) {
  i64 dim = (i64)1 << n;
  double re = 0, im = 0;
  for (i64 i = 0; i < dim; i++) {
    re += a[i].re * b[i].re + a[i].im * b[i].im;
    im += a[i].re * b[i].im - a[i].im * b[i].re;
  }
  return re * re + im * im;
}

/* ---------------------------------------------------------------- sampling */

/* Thread-local xorshift64* of the C ABI (c_ffi.rs:48-67, nqpu_set_seed 1444). */
static __thread uint64_t rng_state = 0x853c49e6748fea9bULL;

void orc_rng_seed(uint64_t seed) { rng_state = seed == 0 ? 0x853c49e6748fea9bULL : seed; }

/* The generator's raw 64-bit state (for tests of the state transition). */
uint64_t orc_rng_state(void) { return rng_state; }

double orc_rng_next(void) {
  uint64_t s = rng_state;
  s ^= s >> 12;
  s ^= s << 25;
  s ^= s >> 27;
  rng_state = s;
  uint64_t v = s * 0x2545F4914F6CDD1DULL;
  return (double)(v >> 11) / (double)(1ULL << 53);
}

/* c_ffi.rs:506-525: per shot, linear scan for the first i with r < cumsum. */
void orc_sample_cabi(const orc_c* a, int n, uint32_t shots, uint64_t* out) {
  i64 dim = (i64)1 << n;
  /* the cumulative sums are the same for every shot; a binary search over them
   * returns the same index as the reference's linear scan */
  double* cdf = (double*)malloc(sizeof(double) * dim);
  double s = 0;
  for (i64 i = 0; i < dim; i++) { s += nsq(a[i]); cdf[i] = s; }
  for (uint32_t k = 0; k < shots; k++) {
    double r = orc_rng_next();
    i64 lo = 0, hi = dim; /* first index with r < cdf[i] */
    while (lo < hi) {
      i64 mid = (lo + hi) >> 1;
      if (r < cdf[mid]) hi = mid; else lo = mid + 1;
    }
    out[k] = (uint64_t)(lo < dim ? lo : dim - 1);
  }
  free(cdf);
}

/* lib.rs:527-550: first index with cdf >= r, clamped to dim-1. */
void orc_sample_cdf(const orc_c* a, int n, const double* r, uint64_t shots, uint64_t* out) {
  i64 dim = (i64)1 << n;
  double* cdf = (double*)malloc(sizeof(double) * dim);
  double s = 0;
  for (i64 i = 0; i < dim; i++) { s += nsq(a[i]); cdf[i] = s; }
  for (uint64_t k = 0; k < shots; k++) {
    i64 lo = 0, hi = dim;
    while (lo < hi) {
      i64 mid = (lo + hi) >> 1;
      if (cdf[mid] < r[k]) lo = mid + 1; else hi = mid;
    }
    out[k] = (uint64_t)(lo < dim ? lo : dim - 1);
  }
  free(cdf);
}

/* c_ffi.rs:473-500 */
uint32_t orc_measure_qubit(orc_c* a, int n, int q, double* prob) {
  i64 dim = (i64)1 << n;
  double p0 = 0;
  for (i64 i = 0; i < dim; i++)
    if (!((i >> q) & 1)) p0 += nsq(a[i]);
  double p1 = 1.0 - p0;
  double r = orc_rng_next();
  uint32_t out = r < p0 ? 0 : 1;
  double po = out ? p1 : p0;
  double rn = 1.0 / sqrt(po);
  for (i64 i = 0; i < dim; i++) {
    if ((uint32_t)((i >> q) & 1) == out) a[i] = C(a[i].re * rn, a[i].im * rn);
    else a[i] = C(0, 0);
  }
  if (prob) *prob = po;
  return out;
}

/* ------------------------------------------------------- state mutations */

/* reset_qubit (lib.rs:1705-1730): zero every amplitude whose bit q is 1, then
 * divide by the square root of the remaining mass when it is positive. */
void orc_reset_qubit(orc_c* a, int n, int q) {
  i64 dim = (i64)1 << n, mask = (i64)1 << q;
  for (i64 i = 0; i < dim; i++)
    if (i & mask) a[i] = C(0, 0);
  double s = 0;
  for (i64 i = 0; i < dim; i++) s += nsq(a[i]);
  s = sqrt(s);
  if (s > 0.0) {
    double inv = 1.0 / s;
    for (i64 i = 0; i < dim; i++) a[i] = C(a[i].re * inv, a[i].im * inv);
  }
}

/* initialize_from_amplitudes (lib.rs:1918-1935): the length must be 2^n and
 * the norm at least 1e-10; then dst = src / ||src||.  Returns 0 (dst
 * untouched) when either check fails, 1 otherwise. */
int orc_initialize_from_amplitudes(orc_c* dst, int n, const orc_c* src, uint64_t len) {
  i64 dim = (i64)1 << n;
  if ((i64)len != dim) return 0;
  double s = 0;
  for (i64 i = 0; i < dim; i++) s += nsq(src[i]);
  s = sqrt(s);
  if (s < 1e-10) return 0;
  double inv = 1.0 / s;
  for (i64 i = 0; i < dim; i++) dst[i] = C(src[i].re * inv, src[i].im * inv);
  return 1;
}

/* primitives::pauli_expectation (primitives.rs:683-728): qubit 0 is the MOST
 * significant bit of the index (bit = n - 1 - qubit); for each basis index k
 * the string maps |k> to phase |k'> (X flips, Y flips with phase i on |0> and
 * -i on |1>, Z gives -1 on |1>), and the result is
 * coeff * Re(sum_k conj(a[k']) * phase * a[k]).  paulis[j] is one of
 * 'I', 'X', 'Y', 'Z' acting on qubits[j]. */
double orc_pauli_expectation_msb(const orc_c* a, int n, const int32_t* qubits, const char* paulis, int count,
                                 double coeff) {
  i64 dim = (i64)1 << n;
  double rr = 0, ri = 0;
  for (i64 k = 0; k < dim; k++) {
    orc_c ph = C(1, 0);
    i64 kp = k;
    for (int j = 0; j < count; j++) {
      int bit = n - 1 - qubits[j];
      int bv = (int)((k >> bit) & 1);
      switch (paulis[j]) {
        case 'X': kp ^= (i64)1 << bit; break;
        case 'Y':
          kp ^= (i64)1 << bit;
          ph = cmul(ph, bv == 0 ? C(0, 1) : C(0, -1));
          break;
        case 'Z':
          if (bv) ph = cmul(ph, C(-1, 0));
          break;
        default: break;
      }
    }
    orc_c ck = C(a[kp].re, -a[kp].im);
    orc_c t = cmul(cmul(ck, ph), a[k]);
    rr += t.re;
    ri += t.im;
  }
  (void)ri;
  return coeff * rr;
}
