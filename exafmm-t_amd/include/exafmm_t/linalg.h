// Host dense linear algebra and DFTs for the precomputation (FP64).
//
// Replaces the reference's BLAS/LAPACK/FFTW calls (include/math_wrapper.h:51-166,
// FFTW plans in fmm_scale_invariant.h:524-540 and fmm.h:527-606) with this
// project's own code, so the host library has no third-party math dependency:
//  * gemm: row-major C = A B (math_wrapper.h:76-95 semantics);
//  * svd: one-sided Jacobi (Hestenes) SVD, A = U diag(S) V^H with singular
//    values in descending order (math_wrapper.h:98-136 returns the same
//    factorisation from ?gesvd);
//  * dft3_r2c / dft3_c2c: forward 3-D DFTs with the FFTW conventions
//    (exp(-2 pi i jk/n), unnormalised, r2c output halved along the last,
//    fastest dimension), computed as separable dense DFTs.
#pragma once
#include "exafmm_t.h"

namespace exafmm_t {

template <typename T>
inline T conj_if(const T& v) { return v; }
inline complex_t conj_if(const complex_t& v) { return std::conj(v); }

// C (m x n) = A (m x k) * B (k x n), all row-major.
template <typename T>
void gemm(int m, int n, int k, const T* A, const T* B, T* C) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < m; i++) {
    T* c = C + size_t(i) * n;
    for (int j = 0; j < n; j++) c[j] = T(0);
    for (int l = 0; l < k; l++) {
      const T a = A[size_t(i) * k + l];
      const T* b = B + size_t(l) * n;
      for (int j = 0; j < n; j++) c[j] += a * b[j];
    }
  }
}

template <typename T>
std::vector<T> transpose(const std::vector<T>& v, int m, int n) {
  std::vector<T> t(v.size());
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) t[size_t(j) * m + i] = v[size_t(i) * n + j];
  return t;
}

template <typename T>
std::vector<T> conjugate_transpose(const std::vector<T>& v, int m, int n) {
  std::vector<T> t(v.size());
  for (int i = 0; i < m; i++)
    for (int j = 0; j < n; j++) t[size_t(j) * m + i] = conj_if(v[size_t(i) * n + j]);
  return t;
}

// One-sided Jacobi SVD of a square n x n matrix A (row-major, n even or odd).
// Outputs row-major U (n x n), S (n, descending), VH (n x n) with
// A = U diag(S) VH.  Column pairs are processed in round-robin (tournament)
// order so that each round's n/2 rotations are independent (OpenMP).
template <typename T>
void svd(int n, const std::vector<T>& A, std::vector<T>& U, std::vector<real_t>& S, std::vector<T>& VH) {
  const int np = n + (n & 1);  // padded to even for the tournament
  // column-major working copies: W = A (columns of A), V = I
  std::vector<T> W(size_t(np) * n, T(0)), V(size_t(np) * np, T(0));
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) W[size_t(j) * n + i] = A[size_t(i) * n + j];
  for (int j = 0; j < np; j++) V[size_t(j) * np + j] = T(1);
  std::vector<int> perm(np);
  for (int i = 0; i < np; i++) perm[i] = i;
  const real_t tol = real_t(n) * 2.220446049250313e-16;
  for (int sweep = 0; sweep < 60; sweep++) {
    int rotated = 0;
    for (int round = 0; round < np - 1; round++) {
#pragma omp parallel for schedule(static) reduction(+ : rotated)
      for (int pi = 0; pi < np / 2; pi++) {
        int a = perm[pi], b = perm[np - 1 - pi];
        if (a > b) std::swap(a, b);
        if (b >= n) continue;  // padding column
        T* wa = &W[size_t(a) * n];
        T* wb = &W[size_t(b) * n];
        real_t alpha = 0, beta = 0;
        T gamma = T(0);
        for (int i = 0; i < n; i++) {
          alpha += std::norm(wa[i]);
          beta += std::norm(wb[i]);
          gamma += conj_if(wa[i]) * wb[i];
        }
        const real_t g = std::abs(gamma);
        if (g == 0 || g <= tol * std::sqrt(alpha * beta)) continue;
        rotated++;
        const real_t zeta = (beta - alpha) / (2 * g);
        const real_t t = (zeta >= 0 ? 1.0 : -1.0) / (std::abs(zeta) + std::sqrt(1 + zeta * zeta));
        const real_t cs = 1 / std::sqrt(1 + t * t), sn = cs * t;
        const T ph = gamma / g;  // unit phase of gamma; rotate (w_a, w_b conj(ph))
        const T phc = conj_if(ph);
        for (int i = 0; i < n; i++) {
          const T x = wa[i], y = wb[i] * phc;
          wa[i] = cs * x - sn * y;
          wb[i] = sn * x + cs * y;
        }
        T* va = &V[size_t(a) * np];
        T* vb = &V[size_t(b) * np];
        for (int i = 0; i < n; i++) {
          const T x = va[i], y = vb[i] * phc;
          va[i] = cs * x - sn * y;
          vb[i] = sn * x + cs * y;
        }
      }
      // rotate the tournament: keep perm[0], shift the rest
      const int last = perm[np - 1];
      for (int i = np - 1; i > 1; i--) perm[i] = perm[i - 1];
      perm[1] = last;
    }
    if (rotated == 0) break;
  }
  std::vector<real_t> sv(n);
  for (int j = 0; j < n; j++) {
    real_t s = 0;
    for (int i = 0; i < n; i++) s += std::norm(W[size_t(j) * n + i]);
    sv[j] = std::sqrt(s);
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return sv[x] > sv[y]; });
  U.assign(size_t(n) * n, T(0));
  VH.assign(size_t(n) * n, T(0));
  S.assign(n, 0);
  for (int k = 0; k < n; k++) {
    const int j = order[k];
    S[k] = sv[j];
    for (int i = 0; i < n; i++) {
      U[size_t(i) * n + k] = sv[j] > 0 ? W[size_t(j) * n + i] / sv[j] : T(i == k ? 1 : 0);
      VH[size_t(k) * n + i] = conj_if(V[size_t(j) * np + i]);
    }
  }
}

// 1-D DFT along one axis of a row-major 3-D complex array (in place).
inline void dft_axis(std::vector<complex_t>& a, int n0, int n1, int n2, int axis, int sign) {
  const int n = axis == 0 ? n0 : (axis == 1 ? n1 : n2);
  std::vector<complex_t> w(n);
  for (int m = 0; m < n; m++) w[m] = std::polar(1.0, sign * 2.0 * PI * m / n);
  const size_t s2 = 1, s1 = size_t(n2), s0 = size_t(n1) * n2;
  const size_t stride = axis == 0 ? s0 : (axis == 1 ? s1 : s2);
  const int outer0 = axis == 0 ? 1 : n0, outer1 = axis == 1 ? 1 : n1, outer2 = axis == 2 ? 1 : n2;
#pragma omp parallel for collapse(2) schedule(static)
  for (int i0 = 0; i0 < outer0; i0++)
    for (int i1 = 0; i1 < outer1; i1++) {
      std::vector<complex_t> line(n), out(n);
      for (int i2 = 0; i2 < outer2; i2++) {
        const size_t base = size_t(i0) * s0 + size_t(i1) * s1 + size_t(i2) * s2;
        for (int j = 0; j < n; j++) line[j] = a[base + j * stride];
        for (int k = 0; k < n; k++) {
          complex_t acc = 0;
          for (int j = 0; j < n; j++) acc += line[j] * w[(size_t(k) * j) % n];
          out[k] = acc;
        }
        for (int k = 0; k < n; k++) a[base + k * stride] = out[k];
      }
    }
}

// Forward 3-D c2c DFT of an n x n x n row-major array.
inline std::vector<complex_t> dft3_c2c(const std::vector<complex_t>& in, int n) {
  std::vector<complex_t> a(in);
  for (int ax = 2; ax >= 0; ax--) dft_axis(a, n, n, n, ax, -1);
  return a;
}

// Forward 3-D r2c DFT of an n x n x n row-major real array; output n x n x (n/2+1).
inline std::vector<complex_t> dft3_r2c(const RealVec& in, int n) {
  std::vector<complex_t> a(in.begin(), in.end());
  for (int ax = 2; ax >= 0; ax--) dft_axis(a, n, n, n, ax, -1);
  const int nh = n / 2 + 1;
  std::vector<complex_t> out(size_t(n) * n * nh);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < nh; k++) out[(size_t(i) * n + j) * nh + k] = a[(size_t(i) * n + j) * n + k];
  return out;
}

}  // namespace exafmm_t
