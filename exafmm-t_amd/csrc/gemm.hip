// Row-gathered FP64 GEMM for the translation operators.
//
//   C[crow[i] + n] (=|+=) alpha[i] * sum_k A[arow[i] + k] * B[k*N + n]
//
// Used for the pseudo-inverse applications of P2M / L2P (reference
// fmm_scale_invariant.h:189-193, 215-219: two gemv per leaf, batched here over
// all leaves of a level), for M2M (fmm_scale_invariant.h:243-251: the 8 child
// vectors of a parent are contiguous, so a level is one GEMM with
// K = 8*nsurf against the stacked transposed M2M matrices) and for L2L
// (fmm_scale_invariant.h:260-268: N = 8*nsurf against the concatenated
// transposed L2L matrices, accumulating into the contiguous child vectors).
//
// Tiling: 256 threads, 128 x 64 (real) output tile, 8 x 4 outputs per thread,
// K staged through LDS in steps of 8.  Complex: 64 x 64 tile, 4 x 4 complex
// outputs per thread.
#include "kernels.h"

namespace xfmm {

template <int BM, int BN, int TM, int TN, int BK>
__global__ __launch_bounds__(256) void k_gemm_rows_real(int m, int n, int k, const double* __restrict__ A,
                                                       const long* __restrict__ arow, const double* __restrict__ B,
                                                       double* __restrict__ C, const long* __restrict__ crow,
                                                       const double* __restrict__ alpha, int beta) {
  static_assert((BM / TM) * (BN / TN) == 256, "thread layout");
  __shared__ double As[BK][BM + 1];
  __shared__ double Bs[BK][BN];
  __shared__ long arow_s[BM];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.y * BM;
  const int col0 = blockIdx.x * BN;
  for (int r = tid; r < BM; r += 256) arow_s[r] = (row0 + r < m) ? arow[row0 + r] : -1;
  const int tr = tid / (BN / TN);  // thread row group
  const int tc = tid % (BN / TN);  // thread col group
  double acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = 0;
  __syncthreads();
  for (int k0 = 0; k0 < k; k0 += BK) {
    // A tile: BM x BK, element e -> (row e / BK, kk e % BK)
    for (int e = tid; e < BM * BK; e += 256) {
      const int r = e / BK, kk = e % BK;
      const long ar = arow_s[r];
      As[kk][r] = (ar >= 0 && k0 + kk < k) ? A[ar + k0 + kk] : 0.0;
    }
    for (int e = tid; e < BK * BN; e += 256) {
      const int kk = e / BN, c = e % BN;
      Bs[kk][c] = (k0 + kk < k && col0 + c < n) ? B[size_t(k0 + kk) * n + col0 + c] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      double a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = As[kk][tr * TM + i];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bs[kk][tc * TN + j];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TM; i++) {
    const int r = row0 + tr * TM + i;
    if (r >= m) continue;
    const long cr = crow[r];
    const double al = alpha ? alpha[r] : 1.0;
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int c = col0 + tc * TN + j;
      if (c >= n) continue;
      double v = al * acc[i][j];
      if (beta) v += C[cr + c];
      C[cr + c] = v;
    }
  }
}

// complex: values are interleaved (re, im); offsets are in complex elements
template <int BM, int BN, int TM, int TN, int BK>
__global__ __launch_bounds__(256) void k_gemm_rows_cplx(int m, int n, int k, const double2* __restrict__ A,
                                                       const long* __restrict__ arow, const double2* __restrict__ B,
                                                       double2* __restrict__ C, const long* __restrict__ crow,
                                                       const double* __restrict__ alpha, int beta) {
  static_assert((BM / TM) * (BN / TN) == 256, "thread layout");
  __shared__ double2 As[BK][BM + 1];
  __shared__ double2 Bs[BK][BN];
  __shared__ long arow_s[BM];
  const int tid = threadIdx.x;
  const int row0 = blockIdx.y * BM;
  const int col0 = blockIdx.x * BN;
  for (int r = tid; r < BM; r += 256) arow_s[r] = (row0 + r < m) ? arow[row0 + r] : -1;
  const int tr = tid / (BN / TN);
  const int tc = tid % (BN / TN);
  double2 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; i++)
#pragma unroll
    for (int j = 0; j < TN; j++) acc[i][j] = make_double2(0.0, 0.0);
  __syncthreads();
  for (int k0 = 0; k0 < k; k0 += BK) {
    for (int e = tid; e < BM * BK; e += 256) {
      const int r = e / BK, kk = e % BK;
      const long ar = arow_s[r];
      As[kk][r] = (ar >= 0 && k0 + kk < k) ? A[ar + k0 + kk] : make_double2(0.0, 0.0);
    }
    for (int e = tid; e < BK * BN; e += 256) {
      const int kk = e / BN, c = e % BN;
      Bs[kk][c] = (k0 + kk < k && col0 + c < n) ? B[size_t(k0 + kk) * n + col0 + c] : make_double2(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk++) {
      double2 a[TM], b[TN];
#pragma unroll
      for (int i = 0; i < TM; i++) a[i] = As[kk][tr * TM + i];
#pragma unroll
      for (int j = 0; j < TN; j++) b[j] = Bs[kk][tc * TN + j];
#pragma unroll
      for (int i = 0; i < TM; i++)
#pragma unroll
        for (int j = 0; j < TN; j++) {
          acc[i][j].x = fma(a[i].x, b[j].x, acc[i][j].x);
          acc[i][j].x = fma(-a[i].y, b[j].y, acc[i][j].x);
          acc[i][j].y = fma(a[i].x, b[j].y, acc[i][j].y);
          acc[i][j].y = fma(a[i].y, b[j].x, acc[i][j].y);
        }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < TM; i++) {
    const int r = row0 + tr * TM + i;
    if (r >= m) continue;
    const long cr = crow[r];
    const double al = alpha ? alpha[r] : 1.0;
#pragma unroll
    for (int j = 0; j < TN; j++) {
      const int c = col0 + tc * TN + j;
      if (c >= n) continue;
      double2 v = make_double2(al * acc[i][j].x, al * acc[i][j].y);
      if (beta) { v.x += C[cr + c].x; v.y += C[cr + c].y; }
      C[cr + c] = v;
    }
  }
}

void launch_gemm_rows(bool cplx, int m, int n, int k, const double* A, const long* arow, const double* B,
                      double* C, const long* crow, const double* alpha, int beta, hipStream_t s) {
  if (m <= 0 || n <= 0) return;
  if (!cplx) {
    constexpr int BM = 128, BN = 64;
    dim3 grid((n + BN - 1) / BN, (m + BM - 1) / BM);
    hipLaunchKernelGGL((k_gemm_rows_real<BM, BN, 8, 4, 8>), grid, dim3(256), 0, s, m, n, k, A, arow, B, C, crow,
                       alpha, beta);
  } else {
    constexpr int BM = 64, BN = 64;
    dim3 grid((n + BN - 1) / BN, (m + BM - 1) / BM);
    hipLaunchKernelGGL((k_gemm_rows_cplx<BM, BN, 4, 4, 8>), grid, dim3(256), 0, s, m, n, k,
                       reinterpret_cast<const double2*>(A), arow, reinterpret_cast<const double2*>(B),
                       reinterpret_cast<double2*>(C), crow, alpha, beta);
  }
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
  size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) p[i] = 0.0;
}

void launch_zero(double* p, size_t n, hipStream_t s) {
  if (n == 0) return;
  size_t blocks = (n + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(k_zero, dim3(unsigned(blocks)), dim3(256), 0, s, p, n);
}

__global__ void k_scatter_values(const double* __restrict__ v, const int* __restrict__ perm, int n, int w,
                                 double* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(n) * w) return;
  const size_t t = i / w, e = i % w;
  out[size_t(perm[t]) * w + e] = v[i];
}

void launch_scatter_values(const double* sorted_val, const int* perm, int n, int vw, double* out, hipStream_t s) {
  const int w = 4 * vw;
  const size_t tot = size_t(n) * w;
  if (tot == 0) return;
  hipLaunchKernelGGL(k_scatter_values, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, s, sorted_val, perm, n, w, out);
}

}  // namespace xfmm
