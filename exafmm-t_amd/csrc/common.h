// Shared device-side definitions of the gfx950 FMM engine.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace xfmm {

enum Kind { LAPLACE = 0, HELMHOLTZ = 1, YUKAWA = 2 };

constexpr double FOUR_PI_INV = 0.07957747154594766788;  // 1/(4 pi)
constexpr int NCHILD = 8;
constexpr int NM2L = 26;  // M2L relative positions of a target parent

template <typename R>
struct cplx {
  R x, y;
};

// Wave numbers as seen by the kernels.  Laplace ignores both, modified
// Helmholtz (Yukawa) uses re, Helmholtz uses re + i im.
struct Wavek {
  double re, im;
};

__device__ __forceinline__ double rsqrt_pos(double r2) { return r2 > 0.0 ? rsqrt(r2) : 0.0; }
__device__ __forceinline__ float rsqrt_pos(float r2) { return r2 > 0.0f ? rsqrtf(r2) : 0.0f; }

// 1/sqrt(r2) for r2 known to be positive and finite (well-separated surfaces):
// hardware estimate plus one second-order Newton correction,
// y (1 + e/2 + 3e^2/8) with e = 1 - r2 y^2.
__device__ __forceinline__ double rsqrt_sep(double r2) {
  const double y = __builtin_amdgcn_rsq(r2);
  const double e = fma(-r2 * y, y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}
__device__ __forceinline__ float rsqrt_sep(float r2) { return __builtin_amdgcn_rsqf(r2); }
__device__ __forceinline__ double dexp(double x) { return exp(x); }
__device__ __forceinline__ void dsincos(double x, double* s, double* c) { sincos(x, s, c); }
// FP32 (mixed / FLOAT near field and surface evaluations): exp through the
// hardware 2^x, sine and cosine on the hardware units, whose argument is in
// revolutions, after reducing x / (2 pi) to [-1/2, 1/2].  The reductions
// cost a relative error of |x| 2^-24 (phase error |x| 6e-8 rad), the FP32
// rounding of the arguments themselves.
__device__ __forceinline__ float dexp(float x) { return __builtin_amdgcn_exp2f(x * 1.44269504088896341f); }
__device__ __forceinline__ void dsincos(float x, float* s, float* c) {
  const float t = x * 0.159154943091895336f;
  const float f = t - __builtin_rintf(t);
  *s = __builtin_amdgcn_sinf(f);
  *c = __builtin_amdgcn_cosf(f);
}

// Compensated (Kahan) accumulation of a partial sum into (s, c).
template <typename R>
__device__ __forceinline__ void kahan_add(R& s, R& c, R v) {
  const R y = v - c;
  const R t = s + y;
  c = (t - s) - y;
  s = t;
}

}  // namespace xfmm
