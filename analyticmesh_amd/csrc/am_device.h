// Device-side building blocks shared by the marching kernels (gfx950).
//
// Every floating-point expression here is written out operation by operation
// (files are compiled with -ffp-contract=off; fused multiply-adds appear only
// where fma() is written).  The FP64 kernels therefore perform the same IEEE
// operation sequence for the plane chain, the distance keys, the 3x3 solve,
// the feasibility and turn tests as a sequential CPU implementation of the
// same formulas.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "am_limits.h"

namespace am {

// Tolerances of the reference engine (backend/inc/macro.h:16-24).
template <typename T> struct Tol;
template <> struct Tol<double> {
  static constexpr double EPS = 1e-20;
  static constexpr double SAME_EPS = 3e-14;
};
template <> struct Tol<float> {
  static constexpr float EPS = 1e-12f;
  static constexpr float SAME_EPS = 3e-7f;
};

__device__ __forceinline__ double fmaT(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmaT(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double absT(double a) { return __builtin_fabs(a); }
__device__ __forceinline__ float absT(float a) { return __builtin_fabsf(a); }

__device__ __forceinline__ bool mbit(const uint32_t* m, int g) { return (m[g >> 5] >> (g & 31)) & 1u; }

// bfloat16 bits of x (round to nearest even) and back.
__device__ __forceinline__ uint16_t bf16_bits(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_value(uint16_t h) { return __uint_as_float(uint32_t(h) << 16); }

// s = a.v + c, evaluated as fma(ax, vx, fma(ay, vy, fma(az, vz, c))).
template <typename T>
__device__ __forceinline__ T plane_eval(T ax, T ay, T az, T c, const T* v) {
  return fmaT(ax, v[0], fmaT(ay, v[1], fmaT(az, v[2], c)));
}
template <typename T>
__device__ __forceinline__ T plane_nn(T ax, T ay, T az) {
  return fmaT(ax, ax, fmaT(ay, ay, az * az));
}
// Geometric-distance test s/|a| > EPS written without the square root
// (backend/inc/kernel.h:1087-1096).
template <typename T>
__device__ __forceinline__ bool violates(T ax, T ay, T az, T c, const T* v) {
  const T s = plane_eval(ax, ay, az, c, v);
  return s > T(0) && s * s > (Tol<T>::EPS * Tol<T>::EPS) * plane_nn(ax, ay, az);
}

#define AM_DET3(a0, a1, a2, a3, a4, a5, a6, a7, a8) \
  ((a0) * ((a4) * (a8) - (a5) * (a7)) - (a1) * ((a3) * (a8) - (a5) * (a6)) + (a2) * ((a3) * (a7) - (a4) * (a6)))

// Cramer's rule for [p1; p2; p3] x = -[c1; c2; c3] (backend/inc/kernel.h:555-573).
template <typename T>
__device__ __forceinline__ void solve3(const T* p1, const T* p2, const T* p3, T* x) {
  const T A0 = p1[0], A1 = p1[1], A2 = p1[2], A3 = p2[0], A4 = p2[1], A5 = p2[2], A6 = p3[0], A7 = p3[1],
          A8 = p3[2];
  const T b0 = -p1[3], b1 = -p2[3], b2 = -p3[3];
  const T d0 = AM_DET3(A0, A1, A2, A3, A4, A5, A6, A7, A8);
  const T d1 = AM_DET3(b0, A1, A2, b1, A4, A5, b2, A7, A8);
  const T d2 = AM_DET3(A0, b0, A2, A3, b1, A5, A6, b2, A8);
  const T d3 = AM_DET3(A0, A1, b0, A3, A4, b1, A6, A7, b2);
  x[0] = d1 / d0;
  x[1] = d2 / d0;
  x[2] = d3 / d0;
}

#define AM_DOT3(a, b) ((a)[0] * (b)[0] + (a)[1] * (b)[1] + (a)[2] * (b)[2])
// Turn test of the projected normals (backend/inc/kernel.h:1103-1116).
template <typename T>
__device__ __forceinline__ bool pivot_valid(const T* neq, const T* n1, const T* n2, const T* n3) {
  const T A12 = AM_DOT3(n1, n2), A23 = AM_DOT3(n2, n3), A13 = AM_DOT3(n1, n3), A22 = AM_DOT3(n2, n2);
  const T C1 = AM_DOT3(n1, neq), C2 = AM_DOT3(n2, neq), C3 = AM_DOT3(n3, neq), B = AM_DOT3(neq, neq);
  return ((A12 * A23 - A22 * A13) - (A12 * C2 * C3 + A23 * C1 * C2 - A13 * C2 * C2 - A22 * C1 * C3) / B) > 0;
}

template <typename T>
__device__ __forceinline__ bool same_vertex(const T* a, const T* b) {
  return absT(a[0] - b[0]) + absT(a[1] - b[1]) + absT(a[2] - b[2]) < Tol<T>::SAME_EPS;
}

// x -> (i, j), i > j: smallest a with a(a + 3) >= 2k replaces the reference's
// float ceil(sqrt(2k + 2.25) - 1.5) (backend/inc/kernel.h:609-624).
__device__ __forceinline__ void get_i_j(int x, int ij_max, int& i, int& j) {
  const int k = x / 2;
  int a = 0;
  while (a * (a + 3) < 2 * k) ++a;
  const int r = x - a * (a + 1);
  const int A = 2 * a + r / (a + 1);
  const int B = r % (a + 1);
  i = (A + 1) / 2 + B + 1;
  j = A + 1 - i;
  i = i < ij_max ? i : ij_max - 1;
  j = j < ij_max ? j : ij_max - 1;
}

// 64-bit mix of a packed mask (visited-set hash).  Not the reference's FNV:
// the visited set here compares full keys, so any well-mixed hash will do.
__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  x ^= x >> 33;
  x *= 0xff51afd7ed558ccdULL;
  x ^= x >> 33;
  x *= 0xc4ceb9fe1a85ec53ULL;
  x ^= x >> 33;
  return x;
}
// Hash of word x at position w.  Keys hash to the sum (mod 2^64) of their
// words' hashes, so a key that differs from a known one in one word is hashed
// in O(1): the known sum with that word's term swapped.
__device__ __forceinline__ uint64_t word_hash(uint32_t x, int w) {
  return mix64((uint64_t(x) << 32) ^ (0x9e3779b97f4a7c15ULL * uint64_t(w + 1)));
}
__device__ __forceinline__ uint64_t key_hash_sum(const uint32_t* m, int nw) {
  uint64_t h = 0;
  for (int w = 0; w < nw; ++w) h += word_hash(m[w], w);
  return h;
}

__device__ __forceinline__ uint64_t hash_words(const uint32_t* m, int nw, uint64_t seed) {
  uint64_t h = seed;
  for (int w = 0; w < nw; ++w) h = mix64(h ^ (uint64_t(m[w]) + 0x9e3779b97f4a7c15ULL * uint64_t(w + 1)));
  return h;
}

}  // namespace am
