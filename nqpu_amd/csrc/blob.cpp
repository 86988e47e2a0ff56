// blob.cpp — operand blobs of matrix-core blocks, built on the host CPU
// (host-only translation unit: it uses the x86 F16C conversions).
//
// A block's blob holds, for each of the 64 lanes, its part of the 32x32 real
// form of the 16x16 complex block matrix M in the layout
// v_mfma_f32_32x32x16_f16 reads as its A operand: two K-steps of 8 values,
// each as fp16 high part, then fp16 low part (x - hi, rounded again).  Values
// are rounded to float, then to binary16 with round-to-nearest-even; the low
// part is formed in double from the exact value of the high part.  F16C does
// both conversions for 8 values at once with exactly these roundings.
#include <immintrin.h>

#include <complex>
#include <cstdint>

namespace nqpu {

__attribute__((target("avx,f16c"))) static void blob_f16c(const std::complex<double>* M, uint16_t* out) {
  // the 32x32 real form of M, [[re, -im], [im, re]] per complex entry
  alignas(32) double A[32][32];
  for (int r = 0; r < 16; ++r)
    for (int c = 0; c < 16; ++c) {
      const std::complex<double> u = M[16 * r + c];
      A[2 * r][2 * c] = u.real();
      A[2 * r][2 * c + 1] = -u.imag();
      A[2 * r + 1][2 * c] = u.imag();
      A[2 * r + 1][2 * c + 1] = u.real();
    }
  for (int l = 0; l < 64; ++l)
    for (int s = 0; s < 2; ++s) {
      const double* x = &A[l % 32][16 * s + 8 * (l / 32)];
      alignas(32) float xf[8], hf[8], rf[8];
      for (int i = 0; i < 8; ++i) xf[i] = (float)x[i];
      const __m128i hi = _mm256_cvtps_ph(_mm256_load_ps(xf), _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC);
      _mm256_store_ps(hf, _mm256_cvtph_ps(hi));
      for (int i = 0; i < 8; ++i) rf[i] = (float)(x[i] - (double)hf[i]);
      const __m128i lo = _mm256_cvtps_ph(_mm256_load_ps(rf), _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC);
      _mm_storeu_si128(reinterpret_cast<__m128i*>(out + l * 32 + s * 16), hi);
      _mm_storeu_si128(reinterpret_cast<__m128i*>(out + l * 32 + s * 16 + 8), lo);
    }
}

// True when this CPU converts with F16C; the caller then uses mfma_blob_fast.
bool mfma_blob_fast_available() {
  static const bool ok = __builtin_cpu_supports("avx") && __builtin_cpu_supports("f16c");
  return ok;
}

void mfma_blob_fast(const std::complex<double>* M, uint16_t* out) { blob_f16c(M, out); }

}  // namespace nqpu
