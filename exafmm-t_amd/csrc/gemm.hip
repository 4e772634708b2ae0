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
// All of them run on the matrix cores (k_gemm_rows_mfma64, split-K when the
// output has too few tiles to fill the chip).  Complex operands (Helmholtz)
// come as real GEMMs: interleaved (re, im) rows times the operator expanded
// into real 2 x 2 blocks [[Re, Im], [-Im, Re]] (K and N doubled, built once
// by the engine when the operators are set).
#include <stdexcept>

#include "kernels.h"

namespace xfmm {

// MFMA path: BM x BN output tile per 256-thread workgroup, the 4 waves as
// 2 x 2 sub-tiles of (BM/2) x (BN/2) (a grid of 16 x 16 MFMA tiles each), K
// staged through LDS in steps of BK (16) with the next
// step's global loads held in registers while the current step computes.
// Tiles are ordered so that the column tiles of one BM-row band run on one
// XCD (workgroup w goes to XCD w % 8), which then reads that band of A once
// into its L2.  Optional split-K (grid.y): each split writes its partial
// product to `partial` and k_gemm_reduce combines them in a fixed order.
// Compute type of the MFMA path: FP64 (v_mfma_f64_16x16x4_f64, result
// register q of lane group g = row g + 4q) or FP32 (v_mfma_f32_16x16x4_f32,
// row 4g + q) for operators that tolerate it; A (the vectors) is FP64 in
// memory and rounded while staged, B (the operator) is read in the compute
// type, results are stored as FP64.
template <typename CT>
struct GemmMfma;
template <>
struct GemmMfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return g + 4 * q; }
};
template <>
struct GemmMfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int g, int q) { return 4 * g + q; }
};

// Without `partial` the kernel writes C itself: alpha scales the rows of A as
// they are staged (exact: alpha is a power of two) and BETA adds C in the
// epilogue.  BETA = false keeps the epilogue a plain store, so the
// accumulators never leave the AGPRs before it (BETA = true costs occupancy
// and is used only when no split-K scratch is available).
template <int BM, int BN, int BK, bool BETA, typename CT>
__global__ __launch_bounds__(256) void k_gemm_rows_mfma64(int m, int n, int k, int kper, int ncol,
                                                         const double* __restrict__ A, const long* __restrict__ arow,
                                                         const CT* __restrict__ B, double* __restrict__ C,
                                                         const long* __restrict__ crow, double alpha,
                                                         double* __restrict__ partial) {
  constexpr int WM = BM / 2, WN = BN / 2, TA = WM / 16, TB = WN / 16;
  constexpr int NQA = BM * BK / 256, NQB = BK * BN / 256;  // loader elements per thread
  constexpr int ARS = 256 / BK, BKS = 256 / BN;            // loader row / k strides
  static_assert(BM % 32 == 0 && BN % 32 == 0 && BN <= 256 && NQA >= 1 && NQB >= 1, "tile shape");
  typedef typename GemmMfma<CT>::v4 d4;
  __shared__ CT As[BK][BM + 2];
  __shared__ CT Bs[BK][BN + 2];
  __shared__ long arow_s[BM];
  // XCD ordering when there are at least 8 row bands (grid padded to a
  // multiple of 8 bands), plain row-major tile order otherwise
  const int w = blockIdx.x;
  const int nband = (m + BM - 1) / BM;
  int row0, col0;
  if (nband >= 8) {
    const int x = w & 7, i = w >> 3;
    row0 = ((i / ncol) * 8 + x) * BM;
    col0 = (i % ncol) * BN;
  } else {
    row0 = (w / ncol) * BM;
    col0 = (w % ncol) * BN;
  }
  if (row0 >= m) return;  // padding bands of the XCD ordering (uniform per workgroup)
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int wr = wv >> 1, wc = wv & 1;
  const int kbeg = blockIdx.y * kper, kend = min(k, kbeg + kper);
  for (int r = tid; r < BM; r += 256) arow_s[r] = (row0 + r < m) ? arow[row0 + r] : -1;
  const bool direct = partial == nullptr;
  const double scale = direct ? alpha : 1.0;
  d4 acc[TA][TB];
#pragma unroll
  for (int a = 0; a < TA; a++)
#pragma unroll
    for (int b = 0; b < TB; b++) acc[a][b] = d4{0, 0, 0, 0};
  __syncthreads();
  // loader mapping: A element e = tid + 256 q -> row e / BK, k e % BK;
  // B element e -> k e / BN, column e % BN
  const int akk = tid % BK, ar0 = tid / BK;
  const int bkk = tid / BN, bc = tid % BN;
  long arw[NQA];
#pragma unroll
  for (int q = 0; q < NQA; q++) arw[q] = arow_s[ar0 + ARS * q];
  const bool bcol_ok = col0 + bc < n;
  double ra[NQA];
  CT rb[NQB];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < NQA; q++) ra[q] = (arw[q] >= 0 && k0 + akk < kend) ? A[arw[q] + k0 + akk] : 0.0;
#pragma unroll
    for (int q = 0; q < NQB; q++) {
      const int kb = bkk + BKS * q;
      rb[q] = (bcol_ok && k0 + kb < kend) ? B[size_t(k0 + kb) * n + col0 + bc] : CT(0);
    }
  };
  if (kbeg < kend) fetch(kbeg);
  for (int k0 = kbeg; k0 < kend; k0 += BK) {
    // (the scale is applied here, when the loads have arrived anyway, so the
    // prefetch of the next stage is not waited for before the MFMAs)
#pragma unroll
    for (int q = 0; q < NQA; q++) As[akk][ar0 + ARS * q] = CT(scale * ra[q]);
#pragma unroll
    for (int q = 0; q < NQB; q++) Bs[bkk + BKS * q][bc] = rb[q];
    __syncthreads();
    if (k0 + BK < kend) fetch(k0 + BK);
#pragma unroll
    for (int s = 0; s < BK / 4; s++) {
      const int kk = 4 * s + g;
      CT av[TA], bv[TB];
#pragma unroll
      for (int t = 0; t < TA; t++) av[t] = As[kk][wr * WM + t * 16 + j];
#pragma unroll
      for (int t = 0; t < TB; t++) bv[t] = Bs[kk][wc * WN + t * 16 + j];
#pragma unroll
      for (int a = 0; a < TA; a++)
#pragma unroll
        for (int b = 0; b < TB; b++) acc[a][b] = GemmMfma<CT>::mma(av[a], bv[b], acc[a][b]);
    }
    __syncthreads();
  }
  // result register q of lane (g, j): row GemmMfma::row(g, q), column j of
  // the 16 x 16 tile
#pragma unroll
  for (int a = 0; a < TA; a++)
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int r = row0 + wr * WM + a * 16 + GemmMfma<CT>::row(g, q);
      if (r >= m) continue;
#pragma unroll
      for (int b = 0; b < TB; b++) {
        const int cc = col0 + wc * WN + b * 16 + j;
        if (cc >= n) continue;
        const double v = double(acc[a][b][q]);
        if (!direct) partial[(size_t(blockIdx.y) * m + r) * n + cc] = v;
        else if (BETA) C[crow[r] + cc] += v;
        else C[crow[r] + cc] = v;
      }
      // one accumulator tile at a time through the VGPRs in the BETA epilogue
      if (BETA) __builtin_amdgcn_sched_barrier(0);
    }
}

__global__ void k_to_float(const double* __restrict__ in, float* __restrict__ out, size_t n) {
  const size_t stride = size_t(gridDim.x) * blockDim.x;
  for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = float(in[i]);
}

void launch_to_float(const double* in, float* out, size_t n, hipStream_t s) {
  if (n == 0) return;
  const unsigned blocks = unsigned(std::min<size_t>((n + 255) / 256, 65536));
  hipLaunchKernelGGL(k_to_float, dim3(blocks), dim3(256), 0, s, in, out, n);
}

__global__ void k_gemm_reduce(int m, int n, int nsplit, const double* __restrict__ partial, double* __restrict__ C,
                              const long* __restrict__ crow, double alpha, int beta) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(m) * n) return;
  const int r = int(i / n), cc = int(i % n);
  double s = 0;
  for (int z = 0; z < nsplit; z++) s += partial[(size_t(z) * m + r) * n + cc];
  const long cr = crow[r];
  double v = alpha * s;
  if (beta) v += C[cr + cc];
  C[cr + cc] = v;
}

// Tile shape: 128 x 128 for large outputs, 64 x 64 for small levels (more
// workgroups, shorter per-workgroup K loop).
struct GemmTile {
  int bm, bn;
};
static GemmTile gemm_tile(int m) { return m >= 2048 ? GemmTile{128, 128} : GemmTile{64, 64}; }

// split-K factor: enough workgroups for every resident slot of the chip
// (256 CUs x 2 workgroups of 128 x 128, x 4 of 64 x 64), each split keeping
// at least 64 of K (4 LDS stages)
static int gemm_split(int m, int n, int k) {
  const GemmTile t = gemm_tile(m);
  const int tiles = ((n + t.bn - 1) / t.bn) * ((m + t.bm - 1) / t.bm);
  const int target = t.bm == 128 ? 512 : 1024;
  int split = 1;
  while (split < 64 && tiles * split < target && k / (split * 2) >= 64) split *= 2;
  return split;
}

size_t gemm_scratch_doubles(int m, int n, int k) {
  const int split = gemm_split(m, n, k);
  return split > 1 ? size_t(split) * m * n : 0;
}

// K step (measured on 32768x488x488 and 4096x488x3904: FP32 with 32 needs
// more registers and drops to 2 waves per SIMD, 5-7 % slower than 16)
template <typename CT>
constexpr int gemm_bk() { return 16; }

template <int BM, int BN, bool BETA, typename CT>
static void launch_mfma(int m, int n, int k, int split, const double* A, const long* arow, const CT* B, double* C,
                        const long* crow, double alpha, double* partial, hipStream_t s) {
  constexpr int BK = gemm_bk<CT>();
  int kper = (k + split - 1) / split;
  kper = (kper + BK - 1) / BK * BK;
  const int ncol = (n + BN - 1) / BN;
  const int nband = (m + BM - 1) / BM;
  const int nb = nband >= 8 ? (nband + 7) / 8 * 8 : nband;
  dim3 grid(unsigned(nb * ncol), split);
  hipLaunchKernelGGL((k_gemm_rows_mfma64<BM, BN, BK, BETA, CT>), grid, dim3(256), 0, s, m, n, k, kper, ncol, A,
                     arow, B, C, crow, alpha, partial);
}

template <int BM, int BN, typename CT>
static void launch_tile(int m, int n, int k, int split, bool part, const double* A, const long* arow, const CT* B,
                        double* C, const long* crow, double alpha, int beta, double* scratch, hipStream_t s) {
  // (the 128 x 128 += epilogue is never built: it would cost occupancy;
  // launch_real picks 128 x 128 only for split-K or plain stores)
  if (part) {
    launch_mfma<BM, BN, false, CT>(m, n, k, split, A, arow, B, C, crow, alpha, scratch, s);
  } else if constexpr (BM != 128) {
    if (beta) launch_mfma<BM, BN, true, CT>(m, n, k, 1, A, arow, B, C, crow, alpha, nullptr, s);
    else launch_mfma<BM, BN, false, CT>(m, n, k, 1, A, arow, B, C, crow, alpha, nullptr, s);
  } else {
    if (beta) throw std::logic_error("launch_tile: accumulating 128 x 128 tile");
    launch_mfma<BM, BN, false, CT>(m, n, k, 1, A, arow, B, C, crow, alpha, nullptr, s);
  }
}

template <typename CT>
static void launch_real(int m, int n, int k, const double* A, const long* arow, const CT* B, double* C,
                        const long* crow, double alpha, int beta, hipStream_t s, double* scratch,
                        size_t scratch_doubles) {
  // an accumulating product without split-K uses the 64 x 64 tile, whose
  // += epilogue keeps full occupancy
  int split = gemm_split(m, n, k);
  const bool part = scratch && split > 1 && size_t(split) * m * n <= scratch_doubles;
  if (!part) split = 1;
  if (gemm_tile(m).bm == 128 && (part || !beta))
    launch_tile<128, 128, CT>(m, n, k, split, part, A, arow, B, C, crow, alpha, beta, scratch, s);
  else
    launch_tile<64, 64, CT>(m, n, k, split, part, A, arow, B, C, crow, alpha, beta, scratch, s);
  if (part) {
    const size_t tot = size_t(m) * n;
    hipLaunchKernelGGL(k_gemm_reduce, dim3(unsigned((tot + 255) / 256)), dim3(256), 0, s, m, n, split, scratch, C,
                       crow, alpha, beta);
  }
}

void launch_gemm_rows(bool f32, int m, int n, int k, const double* A, const long* arow, const void* Bv, double* C,
                      const long* crow, double alpha, int beta, hipStream_t s, double* scratch, size_t scratch_doubles) {
  if (m <= 0 || n <= 0) return;
  if (f32) launch_real<float>(m, n, k, A, arow, static_cast<const float*>(Bv), C, crow, alpha, beta, s, scratch,
                              scratch_doubles);
  else launch_real<double>(m, n, k, A, arow, static_cast<const double*>(Bv), C, crow, alpha, beta, s, scratch,
                           scratch_doubles);
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

__global__ __launch_bounds__(256) void k_charge_sums(const double* __restrict__ q, int n, bool cplx,
                                                     double* __restrict__ out) {
  __shared__ double red[3][256];
  double a = 0, b = 0, c = 0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += CHARGE_SUM_BLOCKS * 256) {
    const double re = cplx ? q[2 * size_t(i)] : q[i], im = cplx ? q[2 * size_t(i) + 1] : 0.0;
    a += re;
    b += im;
    c += sqrt(re * re + im * im);
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = c;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w)
      for (int k = 0; k < 3; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 3) out[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_charge_sums_total(double* __restrict__ out) {
  if (threadIdx.x >= 3) return;
  double t = 0;
  for (int b = 0; b < CHARGE_SUM_BLOCKS; b++) t += out[3 * b + threadIdx.x];
  out[3 * CHARGE_SUM_BLOCKS + threadIdx.x] = t;
}

void launch_charge_sums(const double* q, int n, bool cplx, double* out, hipStream_t s) {
  hipLaunchKernelGGL(k_charge_sums, dim3(CHARGE_SUM_BLOCKS), dim3(256), 0, s, q, n, cplx, out);
  hipLaunchKernelGGL(k_charge_sums_total, dim3(1), dim3(64), 0, s, out);
}

__global__ void k_scale_values(double* __restrict__ val, const double* __restrict__ scale, int n, int vw) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const size_t w = size_t(4) * vw;
  if (i >= size_t(n) * w) return;
  const size_t t = i / w, a = (i % w) / vw;
  val[i] *= scale[2 * t + (a == 0 ? 0 : 1)];
}

void launch_scale_values(double* val, const double* scale, int n, int vw, hipStream_t s) {
  const size_t m = size_t(n) * 4 * vw;
  if (m == 0) return;
  hipLaunchKernelGGL(k_scale_values, dim3(unsigned((m + 255) / 256)), dim3(256), 0, s, val, scale, n, vw);
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
