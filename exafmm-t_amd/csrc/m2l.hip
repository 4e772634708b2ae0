// FFT-accelerated M2L for gfx950.
//
// Same algorithm as exafmm-t fmm_scale_invariant.h:407-440 / fmm.h:395-446:
//  fft_up_equiv  (fmm_scale_invariant.h:560-604): the upward equivalent
//     densities of the 8 children of every M2L source are placed on a (2p)^3
//     grid (surf2conv_up) and transformed, r2c for real kernels
//     (FFTW layout, x halved: f = kx + (p+1)*(ky + 2p*kz)), c2c for Helmholtz;
//  hadamard_product (fmm_scale_invariant.h:343-399): for every target parent,
//     frequency k and relative position r,
//     out[t][k][j1] += sum_j2 M[r][k][j2*8+j1] * in[s(t,r)][k][j2];
//  ifft_dn_check (fmm_scale_invariant.h:606-646): inverse transform
//     (unnormalised; 1/n^3 is folded into M) and gather through surf2conv_dn
//     into the children's downward check potentials, times ifft_scale.
//
// GPU structure.  Transforms: one workgroup per (node, child); the input is
// nonzero only on the [0,p)^3 corner and only the [p,2p)^3 corner of the
// output is needed, so each 1-D pass is a small dense DFT restricted to those
// index ranges, evaluated from LDS with a twiddle table.  Hadamard: a
// workgroup owns 64 target parents (consecutive in Hilbert order, so their
// sources overlap in L2) and 16 frequencies; each wave takes 4 frequencies,
// lanes are targets, the 8x8 matrix M[r][k] is wave-uniform (scalar loads)
// and the 26 relative positions are accumulated in registers.
#include <type_traits>

#include "kernels.h"

namespace xfmm {

template <typename R>
struct C2 { R x, y; };
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

template <typename R>
__device__ __forceinline__ C2<R> cmul(C2<R> a, C2<R> b) {
  return C2<R>{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R>
__device__ __forceinline__ void cfma(C2<R>& acc, C2<R> a, C2<R> b) {
  acc.x = fma(a.x, b.x, acc.x);
  acc.x = fma(-a.y, b.y, acc.x);
  acc.y = fma(a.x, b.y, acc.y);
  acc.y = fma(a.y, b.x, acc.y);
}

// twiddle table W[m] = exp(sign * 2 pi i m / n1) in LDS
template <typename R>
__device__ __forceinline__ void make_twiddles(C2<R>* w, int n1, int sign) {
  for (int m = threadIdx.x; m < n1; m += blockDim.x) {
    double s, c;
    sincospi(double(sign) * 2.0 * m / n1, &s, &c);
    w[m] = C2<R>{R(c), R(s)};
  }
}

// Work item -> (slot, child) with the 8 children of a node on the same XCD:
// consecutive workgroups are dispatched round-robin over the 8 XCDs, so the
// children of node s use workgroup ids 64*(s/8) + 8*c + s%8.
__device__ __forceinline__ void xcd_slot_child(int w, int nslots, int& slot, int& child) {
  const int grp = w / 64, rem = w % 64;
  child = rem / 8;
  slot = grp * 8 + rem % 8;
  (void)nslots;
}

// Largest |value| a thread stored, reduced over the wave and folded into the
// evaluation-wide maximum *amax (float bits; order-preserving for values
// >= 0).  Called by every lane of the wave.
__device__ __forceinline__ void wave_amax(float v, unsigned* amax) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(amax, __float_as_uint(v));
}

template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_fft_up(M2LDev m, DevTree t, const int* __restrict__ src_node, int nslots) {
  extern __shared__ char smem_raw[];
  const int p = m.p, n1 = m.n1;
  const int nh = CPLX ? n1 : n1 / 2 + 1;
  int slot, c;
  xcd_slot_child(blockIdx.x, nslots, slot, c);
  if (slot >= nslots) return;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);       // n1
  C2<R>* A = w + n1;                                   // p*p*nh
  C2<R>* B = A + p * p * nh;                           // p*n1*nh  (shares space with the input grid)
  R* g = reinterpret_cast<R*>(B);                      // p^3 (x2 for complex)
  make_twiddles(w, n1, -1);
  const int node = src_node[slot];
  const int child = t.child0[node] + c;
  const double* up = t.up + size_t(child) * m.nsurf * (CPLX ? 2 : 1);
  const int p3 = p * p * p;
  for (int e = threadIdx.x; e < p3 * (CPLX ? 2 : 1); e += blockDim.x) g[e] = R(0);
  __syncthreads();
  for (int i = threadIdx.x; i < m.nsurf; i += blockDim.x) {
    const int l = m.map_up[i];
    if (CPLX) { g[2 * l] = R(up[2 * i]); g[2 * l + 1] = R(up[2 * i + 1]); }
    else g[l] = R(up[i]);
  }
  __syncthreads();
  // x pass: A[z][y][kx] = sum_x g[z][y][x] W^(kx x)
  for (int e = threadIdx.x; e < p * p * nh; e += blockDim.x) {
    const int kx = e % nh, zy = e / nh;
    C2<R> acc{0, 0};
    for (int x = 0; x < p; x++) {
      const C2<R> tw = w[(kx * x) % n1];
      if (CPLX) cfma(acc, C2<R>{g[2 * (zy * p + x)], g[2 * (zy * p + x) + 1]}, tw);
      else { const R v = g[zy * p + x]; acc.x = fma(v, tw.x, acc.x); acc.y = fma(v, tw.y, acc.y); }
    }
    A[e] = acc;
  }
  __syncthreads();
  // y pass: B[z][ky][kx] = sum_y A[z][y][kx] W^(ky y)
  for (int e = threadIdx.x; e < p * n1 * nh; e += blockDim.x) {
    const int kx = e % nh, ky = (e / nh) % n1, z = e / (nh * n1);
    C2<R> acc{0, 0};
    for (int y = 0; y < p; y++) cfma(acc, A[(z * p + y) * nh + kx], w[(ky * y) % n1]);
    B[e] = acc;
  }
  __syncthreads();
  // z pass straight to global: f = kx + nh*(ky + n1*kz)
  C2<R>* out = reinterpret_cast<C2<R>*>(m.fft_in) + size_t(slot) * m.nfreq * NCHILD;
  float mx = 0.f;
  for (int f = threadIdx.x; f < m.nfreq; f += blockDim.x) {
    const int kx = f % nh, ky = (f / nh) % n1, kz = f / (nh * n1);
    C2<R> acc{0, 0};
    for (int z = 0; z < p; z++) cfma(acc, B[(z * n1 + ky) * nh + kx], w[(kz * z) % n1]);
    out[size_t(f) * NCHILD + c] = acc;
    mx = fmaxf(mx, fmaxf(fabsf(float(acc.x)), fabsf(float(acc.y))));
  }
  if (m.amax) wave_amax(mx, m.amax);
}

template <typename R, bool CPLX>
__global__ __launch_bounds__(256) void k_ifft_dn(M2LDev m, DevTree t, const int* __restrict__ trg_node,
                                                 const double* __restrict__ scale, int nslots) {
  extern __shared__ char smem_raw[];
  const int p = m.p, n1 = m.n1;
  const int nh = CPLX ? n1 : n1 / 2 + 1;
  int slot, c;
  xcd_slot_child(blockIdx.x, nslots, slot, c);
  if (slot >= nslots) return;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);  // n1, exp(+2 pi i m/n1)
  C2<R>* Bp = w + n1;                              // p*n1*nh : [zz][ky][kx]
  C2<R>* Ap = Bp + p * n1 * nh;                    // p*p*nh  : [zz][yy][kx]
  R* g = reinterpret_cast<R*>(Bp);                 // p^3 output cube (x2 complex), reuses Bp
  make_twiddles(w, n1, +1);
  __syncthreads();
  const C2<R>* in = reinterpret_cast<const C2<R>*>(m.fft_out) + size_t(slot) * m.nfreq * NCHILD;
  // inverse z pass for z' = p + zz
  for (int e = threadIdx.x; e < p * n1 * nh; e += blockDim.x) {
    const int kx = e % nh, ky = (e / nh) % n1, zz = e / (nh * n1);
    const int z = p + zz;
    C2<R> acc{0, 0};
    for (int kz = 0; kz < n1; kz++)
      cfma(acc, in[size_t(kx + nh * (ky + n1 * kz)) * NCHILD + c], w[(kz * z) % n1]);
    Bp[e] = acc;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < p * p * nh; e += blockDim.x) {
    const int kx = e % nh, yy = (e / nh) % p, zz = e / (nh * p);
    const int y = p + yy;
    C2<R> acc{0, 0};
    for (int ky = 0; ky < n1; ky++) cfma(acc, Bp[(zz * n1 + ky) * nh + kx], w[(ky * y) % n1]);
    Ap[e] = acc;
  }
  __syncthreads();
  const int p3 = p * p * p;
  for (int e = threadIdx.x; e < p3; e += blockDim.x) {
    const int xx = e % p, zy = e / p;
    const int x = p + xx;
    const C2<R>* a = Ap + zy * nh;
    if (CPLX) {
      C2<R> acc{0, 0};
      for (int kx = 0; kx < n1; kx++) cfma(acc, a[kx], w[(kx * x) % n1]);
      g[2 * e] = acc.x;
      g[2 * e + 1] = acc.y;
    } else {
      // c2r along x: half spectrum kx = 0..n1/2 of a Hermitian sequence
      R acc = a[0].x + ((x & 1) ? -a[nh - 1].x : a[nh - 1].x);
      R twice = 0;
      for (int kx = 1; kx < nh - 1; kx++) {
        const C2<R> tw = w[(kx * x) % n1];
        twice = fma(a[kx].x, tw.x, twice);
        twice = fma(-a[kx].y, tw.y, twice);
      }
      g[e] = acc + 2 * twice;
    }
  }
  __syncthreads();
  const int node = trg_node[slot];
  const int child = t.child0[node] + c;
  const double sc = scale[slot];
  double* dn = t.dn + size_t(child) * m.nsurf * (CPLX ? 2 : 1);
  for (int i = threadIdx.x; i < m.nsurf; i += blockDim.x) {
    const int l = m.map_dn[i];
    if (CPLX) {
      dn[2 * i] += sc * double(g[2 * l]);
      dn[2 * i + 1] += sc * double(g[2 * l + 1]);
    } else {
      dn[i] += sc * double(g[l]);
    }
  }
}

// ------------------------------------------- per-node transforms (real kernels)
//
// One workgroup per node handles all 8 children, slab by slab in kx (the
// halved r2c axis).  For a fixed kx the 8 children x (ky, kz) spectrum values
// are written as contiguous 8-child groups, i.e. full cache lines of the
// [node][freq][child] layout.  Each 1-D DFT is register blocked: a thread loads
// its p (or 2p) inputs once and produces all outputs of the line.

// LDS of the per-node MFMA transforms (rb = bytes of the compute type):
// forward: twiddles, the 8 children's densities on the surface of the
// [0,p)^3 grid (complex for Helmholtz), y-pass input [2p][8p] and z-pass
// input [2p][16p]; inverse: twiddles, y-pass input [4p][8p] and the y-pass
// output [p][p][8] complex.
static size_t fft_node_lds(const M2LDev& m, size_t rb) {
  const size_t p = m.p, n1 = m.n1, vw = m.cplx ? 2 : 1;
  const size_t fwd = n1 * 2 * rb + NCHILD * size_t(m.nsurf) * vw * rb + 2 * p * 8 * p * rb + 2 * p * 16 * p * rb;
  const size_t inv = n1 * 2 * rb + 4 * p * 8 * p * rb + p * p * NCHILD * 2 * rb;
  return fwd > inv ? fwd : inv;
}

// per-node transforms: p in [3, 12], LDS within the CU's 160 KiB with FP64
// arithmetic (so that the choice, and with it the frequency order of the
// spectra and expanded operators, does not depend on the arithmetic type,
// which the mixed path switches with the charges)
static bool fft_node_ok(const M2LDev& m) {
  return m.p >= 3 && m.p <= 12 && fft_node_lds(m, 8) <= 160 * 1024;
}

template <typename F>
static void node_dispatch(int p, F&& f) {
  switch (p) {
    case 3: f(std::integral_constant<int, 3>()); break;
    case 4: f(std::integral_constant<int, 4>()); break;
    case 5: f(std::integral_constant<int, 5>()); break;
    case 6: f(std::integral_constant<int, 6>()); break;
    case 7: f(std::integral_constant<int, 7>()); break;
    case 8: f(std::integral_constant<int, 8>()); break;
    case 9: f(std::integral_constant<int, 9>()); break;
    case 10: f(std::integral_constant<int, 10>()); break;
    case 11: f(std::integral_constant<int, 11>()); break;
    default: f(std::integral_constant<int, 12>()); break;
  }
}

// ---------------------------------------------------------------- Hadamard
//
// Per frequency k the product is a complex 8x8 times 8-vector for every
// (source, target) pair.  Written over the reals, [Re out; Im out] (16) =
// A_rk (16x16) [Re in; Im in] (16) with A = [[Re M^T, -Im M^T], [Im M^T, Re M^T]],
// so a tile of 16 targets is a 16 x 16 x (16 x 26) MFMA product (4 MFMA k-steps
// per relative position).  The A fragment of a lane needs M[j2][j1] for
// j2 in {g, g+4} (g = lane/16, j1 = lane%8); the B fragment needs the input
// children g and g+4 of the lane's target source.
//
// Workgroup = one block of <= 64 M2L targets (4 waves x 16-target tiles) and
// a chunk of kch frequencies (HAD_KCH unless tuned).  The union of the block's M2L sources (<= HAD_UCAP) is
// staged in LDS one frequency at a time, so each source vector is read from
// HBM once per block and reused by every target of the block that needs it.
// Grid order is XCD aware: the workgroups of one frequency chunk run on the
// same XCD, so that chunk's operator slice is fetched into one L2 once and
// reused by all target blocks.

template <typename R>
struct MfmaT;
template <>
struct MfmaT<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // row of result register r of lane group g (v_mfma_f32_16x16x4_f32)
  static __device__ __forceinline__ int row(int g, int r) { return 4 * g + r; }
};
template <>
struct MfmaT<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mma(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of result register r of lane group g (v_mfma_f64_16x16x4_f64)
  static __device__ __forceinline__ int row(int g, int r) { return g + 4 * r; }
};

template <typename R>
struct Quad;
template <>
struct Quad<float> {
  typedef float t __attribute__((ext_vector_type(4)));
};
template <>
struct Quad<double> {
  typedef double t __attribute__((ext_vector_type(4)));
};

// Operators in MFMA A-operand form: for operator level L, frequency k,
// relative position r and lane (g, j) of a wave, the four A values of the
// lane's four MFMAs (row j = (Re or Im, target child t) by HadRow, inner index
// g resp. g + 4 = source child):
//   Re row: { Re M[g][t],  Re M[g+4][t], -Im M[g][t], -Im M[g+4][t] }
//   Im row: { Im M[g][t],  Im M[g+4][t],  Re M[g][t],  Re M[g+4][t] },
// stored as X[((L nfreq + k) 26 + r) 64 + 16 g + j], so a wave reads one
// contiguous 1 KiB per (k, r).  Built once from the reference layout
// M[L][r][f][8 s + t] (complex, f = kx + (p+1) (ky + 2p kz)).  With
// slab = true, X (and the spectra of the per-node transforms) number the
// frequencies kx-slab first, k = kx (2p)^2 + ky + 2p kz, so that one kx slab
// of a node's spectrum is contiguous in memory.
// Row j of the 16-row Hadamard product: target child and Re (0) / Im (1),
// arranged so that the result registers of a lane hold whole (Re, Im) pairs:
// FP32 (register q of group g = row 4 g + q): children 2g, 2g+1, one 16-byte
// store; FP64 (row g + 4 q): children g and g + 4, two 16-byte stores.
template <typename R>
struct HadRow;
template <>
struct HadRow<float> {
  static __device__ __forceinline__ int child(int j) { return j >> 1; }
  static __device__ __forceinline__ int im(int j) { return j & 1; }
};
template <>
struct HadRow<double> {
  static __device__ __forceinline__ int child(int j) { return (j & 3) + 4 * (j >> 3); }
  static __device__ __forceinline__ int im(int j) { return (j >> 2) & 1; }
};

template <typename R>
__global__ void k_m2l_expand(int nlev, int nfreq, int n1, int nh, bool slab, const R* __restrict__ M,
                             R* __restrict__ X) {
  const size_t n = size_t(nlev) * nfreq * NM2L * 64;
  for (size_t e = size_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n; e += size_t(gridDim.x) * blockDim.x) {
    const int lane = int(e % 64), g = lane >> 4, j = lane & 15, t = HadRow<R>::child(j);
    const size_t q = e / 64;
    const int r = int(q % NM2L);
    const size_t lk = q / NM2L;  // L * nfreq + k
    const int k = int(lk % nfreq), L = int(lk / nfreq);
    int f = k;
    if (slab) {
      const int kx = k / (n1 * n1), rem = k % (n1 * n1);
      f = kx + nh * rem;  // rem = ky + n1 kz
    }
    const R* src = M + ((size_t(L) * NM2L + r) * nfreq + f) * 128;
    const R m0x = src[2 * (g * 8 + t)], m0y = src[2 * (g * 8 + t) + 1];
    const R m1x = src[2 * ((g + 4) * 8 + t)], m1y = src[2 * ((g + 4) * 8 + t) + 1];
    R* o = X + e * 4;
    if (!HadRow<R>::im(j)) { o[0] = m0x; o[1] = m1x; o[2] = -m0y; o[3] = -m1y; }
    else { o[0] = m0y; o[1] = m1y; o[2] = m0x; o[3] = m1x; }
  }
}

// Split-FP16 operator form for the mixed Hadamard (k_hadamard_h).  Per
// operator level L and frequency k the 26 relative positions are taken in
// pairs (2 pm, 2 pm + 1); for pair pm and lane (g, j) of a wave the 8 A
// values of a 16x16x32 MFMA are, for c' = 0..3 and source child
// s = 4 (g & 1) + c' of relative position r = 2 pm + (g >> 1), t = j >> 1:
//   j even: ( Re M[s][t], -Im M[s][t] )      (Re of target child t)
//   j odd:  ( Im M[s][t],  Re M[s][t] )      (Im of target child t)
// scaled by 2^e(L,k) so that the largest |entry| of (L, k) lies in
// [2^14, 2^15), and split a = hi + lo, hi = fp16(a), lo = fp16(a - hi).
// Stored as X[((L nfreq + k) 13 + pm) 2 + term][lane][8] (16 bytes per lane);
// hscale[L][k] = 2^-e(L,k).
__global__ __launch_bounds__(256) void k_m2l_expand_h(int nfreq, int n1, int nh, bool slab,
                                                      const float* __restrict__ M, h8* __restrict__ X,
                                                      float* __restrict__ hscale) {
  __shared__ float red[256];
  const int lk = blockIdx.x, k = lk % nfreq, L = lk / nfreq, tid = threadIdx.x;
  int f = k;
  if (slab) {
    const int kx = k / (n1 * n1), rem = k % (n1 * n1);
    f = kx + nh * rem;
  }
  auto src = [&](int r) { return M + ((size_t(L) * NM2L + r) * nfreq + f) * 128; };
  float mx = 0.f;
  for (int e = tid; e < NM2L * 128; e += 256) mx = fmaxf(mx, fabsf(src(e >> 7)[e & 127]));
  red[tid] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] = fmaxf(red[tid], red[tid + w]);
    __syncthreads();
  }
  int ex = 0;
  frexpf(red[0], &ex);
  if (!(red[0] <= 3.0e38f)) ex = 15;
  const float sc = ldexpf(1.f, 15 - ex);
  if (tid == 0) hscale[lk] = ldexpf(1.f, ex - 15);
  h8* o = X + size_t(lk) * (NM2L / 2) * 2 * 64;
  for (int e = tid; e < (NM2L / 2) * 64; e += 256) {
    const int pm = e >> 6, lane = e & 63, g = lane >> 4, j = lane & 15, t = j >> 1;
    const float* mr = src(2 * pm + (g >> 1));
    h8 hi, lo;
#pragma unroll
    for (int cp = 0; cp < 4; cp++) {
      const int sidx = 4 * (g & 1) + cp;
      const float re = mr[2 * (sidx * 8 + t)], im = mr[2 * (sidx * 8 + t) + 1];
      const bool re_row = (j & 1) == 0;
      const float a0 = sc * (re_row ? re : im), a1 = sc * (re_row ? -im : re);
      const _Float16 h0 = _Float16(a0), h1 = _Float16(a1);
      hi[2 * cp] = h0;
      hi[2 * cp + 1] = h1;
      lo[2 * cp] = _Float16(a0 - float(h0));
      lo[2 * cp + 1] = _Float16(a1 - float(h1));
    }
    o[(pm * 2) * 64 + lane] = hi;
    o[(pm * 2 + 1) * 64 + lane] = lo;
  }
}

void launch_m2l_expand(const M2LDev& m, int nlev, const void* M, void* X, float* hscale, hipStream_t s) {
  const size_t n = size_t(nlev) * m.nfreq * NM2L * 64;
  const unsigned blocks = unsigned(std::min<size_t>((n + 255) / 256, 65536));
  // the per-node transforms store frequencies kx-slab major: k = kx n1^2 +
  // ky + n1 kz, reference f = kx + nh (ky + n1 kz)
  const bool slab = fft_node_ok(m);
  const int nh = m.cplx ? m.n1 : m.n1 / 2 + 1;
  if (m.mixed && hscale)
    hipLaunchKernelGGL(k_m2l_expand_h, dim3(unsigned(nlev * m.nfreq)), dim3(256), 0, s, m.nfreq, m.n1, nh, slab,
                       static_cast<const float*>(M), static_cast<h8*>(X), hscale);
  else if (m.mixed)
    hipLaunchKernelGGL(k_m2l_expand<float>, dim3(blocks), dim3(256), 0, s, nlev, m.nfreq, m.n1, nh, slab,
                       static_cast<const float*>(M), static_cast<float*>(X));
  else
    hipLaunchKernelGGL(k_m2l_expand<double>, dim3(blocks), dim3(256), 0, s, nlev, m.nfreq, m.n1, nh, slab,
                       static_cast<const double*>(M), static_cast<double*>(X));
}

// Per block of <= 64 target slots and a chunk of kch frequencies.  For
// each frequency the block stages in LDS the spectra of its source union,
// arranged so that every lane fetches its B operands for one relative
// position with one 16-byte read: [g][union slot] = {Re c_g, Im c_g,
// Re c_(g+4), Im c_(g+4)}, and the 26 operators in the expanded form above
// (one 16-byte read per lane and relative position).  All staging moves
// 16-byte vectors (no LDS bank conflicts on the stores), and the next
// frequency's data is loaded into registers while the current one computes.
// Each wave owns 16 targets; per
// relative position the complex 8x8 product is one real 16x16x16 product
// (four 16x16x4 MFMAs) whose rows are (Re or Im, target child) in HadRow
// order.  A lane's union index for each relative position is fixed
// for the block and held in registers; an absent source points at a zero
// slot, so the 26 relative positions run without branches.
template <typename R>
__global__ __launch_bounds__(256) void k_hadamard(M2LDev m, const R* __restrict__ opsx, HadBlocks hb, int nchunk, int kch) {
  typedef typename MfmaT<R>::v4 v4;
  typedef typename Quad<R>::t q4;
  constexpr int US = HAD_UCAP + 1;
  // FP64 reads its operators straight from global memory (L1 / L2): staged,
  // they would not leave room for a second workgroup per CU
  constexpr bool STAGE_M = std::is_same<R, float>::value;
  __shared__ q4 Us[4 * US];                   // [g][union slot]; slot HAD_UCAP stays zero
  __shared__ q4 Ms[STAGE_M ? NM2L * 64 : 1];  // [r][lane]: A operands of the frequency
  // XCD-aware (chunk, block) mapping: XCD x = w % 8 owns chunks x, x+8, ...
  const int w = blockIdx.x;
  const int x = w & 7, i = w >> 3;
  const int chunk = (i / hb.nblk) * 8 + x;
  const int b = i % hb.nblk;
  if (chunk >= nchunk) return;
  const int kbeg = chunk * kch, kend = min(m.nfreq, kbeg + kch);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int nu = hb.nunion[b];
  if (tid < 4) Us[tid * US + HAD_UCAP] = q4{0, 0, 0, 0};
  int uoff[NM2L];
#pragma unroll
  for (int r = 0; r < NM2L; r++) {
    const int u = hb.loc[(size_t(b) * NM2L + r) * 64 + wave * 16 + j];
    uoff[r] = g * US + (u >= 0 ? u : HAD_UCAP);
  }
  const C2<R>* fin = reinterpret_cast<const C2<R>*>(m.fft_in);
  C2<R>* fout = reinterpret_cast<C2<R>*>(m.fft_out);
  const q4* X = reinterpret_cast<const q4*>(opsx);
  // staging: element e = tid + 256 q (q < 4) is slot e / 4, pair g' = e % 4
  // (children g' and g' + 4)
  constexpr int NU_Q = (HAD_UCAP * 4 + 255) / 256;
  const int gp = tid & 3;
  int usrc[NU_Q];
#pragma unroll
  for (int q = 0; q < NU_Q; q++) {
    const int e = tid + 256 * q;
    usrc[q] = (e < nu * 4) ? hb.uni[size_t(b) * HAD_UCAP + (e >> 2)] : -1;
  }
  constexpr int NM_Q = (NM2L * 64 + 255) / 256;
  C2<R> ra[NU_Q], rb[NU_Q];
  q4 rm[NM_Q];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) {
        const C2<R>* f = fin + (size_t(usrc[q]) * m.nfreq + k) * NCHILD;
        ra[q] = f[gp];
        rb[q] = f[gp + 4];
      }
    if constexpr (STAGE_M) {
      const q4* xk = X + size_t(k) * NM2L * 64;
#pragma unroll
      for (int q = 0; q < NM_Q; q++)
        if (tid + 256 * q < NM2L * 64) rm[q] = xk[tid + 256 * q];
    }
  };
  const bool im_out = j >= 8;
  (void)im_out;
  const int trg = hb.trg[size_t(b) * 64 + wave * 16 + j];
  fetch(kbeg);
  for (int k = kbeg; k < kend; k++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) Us[gp * US + ((tid + 256 * q) >> 2)] = q4{ra[q].x, ra[q].y, rb[q].x, rb[q].y};
    if constexpr (STAGE_M) {
#pragma unroll
      for (int q = 0; q < NM_Q; q++)
        if (tid + 256 * q < NM2L * 64) Ms[tid + 256 * q] = rm[q];
    }
    __syncthreads();
    if (k + 1 < kend) fetch(k + 1);
    // two accumulators (even / odd r); operands of r + 1 read while r computes
    v4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    const q4* Ma = STAGE_M ? Ms : X + size_t(k) * NM2L * 64;
    q4 mm = Ma[lane];
    q4 vv = Us[uoff[0]];
#pragma unroll
    for (int r = 0; r < NM2L; r++) {
      q4 mn = mm, vn = vv;
      if (r + 1 < NM2L) {
        mn = Ma[(r + 1) * 64 + lane];
        vn = Us[uoff[r + 1]];
      }
      v4& acc = (r & 1) ? acc1 : acc0;
      acc = MfmaT<R>::mma(mm.x, vv.x, acc);
      acc = MfmaT<R>::mma(mm.y, vv.z, acc);
      acc = MfmaT<R>::mma(mm.z, vv.y, acc);
      acc = MfmaT<R>::mma(mm.w, vv.w, acc);
      mm = mn;
      vv = vn;
    }
    const v4 acc = acc0 + acc1;
    // result: lane (g, j) register q holds row MfmaT::row(g, q) of target j,
    // registers (0, 1) and (2, 3) the (Re, Im) of one child each (HadRow)
    if (trg >= 0) {
      C2<R>* o = fout + (size_t(trg) * m.nfreq + k) * NCHILD;
      o[HadRow<R>::child(MfmaT<R>::row(g, 0))] = C2<R>{acc[0], acc[1]};
      o[HadRow<R>::child(MfmaT<R>::row(g, 2))] = C2<R>{acc[2], acc[3]};
    }
  }
}

// Mixed-precision Hadamard on the FP16 matrix cores.  Every FP32 operand a is
// carried as two FP16 numbers, a = hi + lo (22 significant bits), and a
// product as hi_a hi_b + hi_a lo_b + lo_a hi_b (three v_mfma_f32_16x16x32_f16
// with FP32 accumulation; the dropped lo_a lo_b is 2^-22 relative).  FP16
// range is kept by power-of-two scales: the operators per (level,
// frequency) at expansion time, the spectra by one scale per evaluation
// from the largest |spectrum value| that the forward transforms recorded in
// *m.amax; the result is multiplied back by both before it is stored.
// One 16x16x32 MFMA covers two relative positions: lane groups g = 0, 1
// take source children 0..3 / 4..7 of position 2 pm, groups 2, 3 those of
// 2 pm + 1 (A layout of k_m2l_expand_h); the B operand of a lane is the
// matching 4 children (re, im) of its target's source, staged in LDS as
// [term][h][union slot] 16-byte vectors.  Blocks, union staging, grid order
// and result layout are those of k_hadamard.
template <int GPB, int TPW>
__global__ __launch_bounds__(256 * GPB / TPW) void k_hadamard_h(M2LDev m, const h8* __restrict__ X,
                                                                const float* __restrict__ hscale, HadBlocks hb,
                                                                int nchunk, int kch) {
  typedef float v4 __attribute__((ext_vector_type(4)));
  constexpr int US = HAD_UCAP + 1;
  constexpr int NP = NM2L / 2;
  static_assert(TPW == 1 || TPW == 2, "tiles per wave");
  constexpr int TPB = 256 / TPW;    // threads per target block: 4 / TPW waves of TPW 16-target tiles
  constexpr int NT = TPB * GPB;
  __shared__ h8 Us[GPB][4 * US];    // per target block: [term][h][union slot]; slot HAD_UCAP stays zero
  __shared__ h8 Ms[NP * 2 * 64];    // [pair][term][lane], shared by the GPB target blocks
  const int nbg = (hb.nblk + GPB - 1) / GPB;
  const int w = blockIdx.x;
  const int x = w & 7, i = w >> 3;
  const int chunk = (i / nbg) * 8 + x;
  if (chunk >= nchunk) return;
  const int kbeg = chunk * kch, kend = min(m.nfreq, kbeg + kch);
  const int tid = threadIdx.x, sub = tid / TPB, st = tid % TPB;
  const int lane = tid & 63, wave = st >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int b = (i % nbg) * GPB + sub;
  const bool live = b < hb.nblk;  // the last workgroup may hold fewer than GPB target blocks
  const int nu = live ? hb.nunion[b] : 0;
  h8* U = Us[sub];
  if (st < 4) U[st * US + HAD_UCAP] = h8{0, 0, 0, 0, 0, 0, 0, 0};
  // tile t of this wave: targets (TPW wave + t) 16 + j of the block
  int uoff[TPW][NP];
  int trg[TPW];
#pragma unroll
  for (int t = 0; t < TPW; t++) {
    const int tt = (TPW * wave + t) * 16 + j;
#pragma unroll
    for (int pm = 0; pm < NP; pm++) {
      const int u = live ? hb.loc[(size_t(b) * NM2L + 2 * pm + (g >> 1)) * 64 + tt] : -1;
      uoff[t][pm] = (g & 1) * US + (u >= 0 ? u : HAD_UCAP);
    }
    trg[t] = live ? hb.trg[size_t(b) * 64 + tt] : -1;
  }
  // spectrum scale: largest |value| of this evaluation's spectra -> [2^14, 2^15)
  const float amax = __uint_as_float(*m.amax);
  int ex = 0;
  frexpf(amax, &ex);
  if (!(amax <= 3.0e38f)) ex = 15;
  const float sb = ldexpf(1.f, 15 - ex), isb = ldexpf(1.f, ex - 15);
  typedef float f4 __attribute__((ext_vector_type(4)));
  const f4* fin = reinterpret_cast<const f4*>(m.fft_in);  // [slot][k][2 halves of 4 children]
  C2<float>* fout = reinterpret_cast<C2<float>*>(m.fft_out);
  // staging of a target block's union by its TPB threads: element
  // e = st + TPB q is union slot e / 2, children 4 (e % 2) .. + 3
  constexpr int NU_Q = (HAD_UCAP * 2 + TPB - 1) / TPB;
  const int hh = st & 1;
  int usrc[NU_Q];
#pragma unroll
  for (int q = 0; q < NU_Q; q++) {
    const int e = st + TPB * q;
    usrc[q] = (e < nu * 2) ? hb.uni[size_t(b) * HAD_UCAP + (e >> 1)] : -1;
  }
  // the operators of a frequency, staged once for the GPB target blocks
  constexpr int NM_Q = (NP * 2 * 64 + NT - 1) / NT;
  f4 ra[NU_Q], rb[NU_Q];
  h8 rm[NM_Q];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) {
        const f4* f = fin + ((size_t(usrc[q]) * m.nfreq + k) * 2 + hh) * 2;
        ra[q] = f[0];
        rb[q] = f[1];
      }
    const h8* xk = X + size_t(k) * NP * 2 * 64;
#pragma unroll
    for (int q = 0; q < NM_Q; q++)
      if (tid + NT * q < NP * 2 * 64) rm[q] = xk[tid + NT * q];
  };
  fetch(kbeg);
  for (int k = kbeg; k < kend; k++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) {
        h8 hi, lo;
#pragma unroll
        for (int e = 0; e < 8; e++) {
          const float v = sb * (e < 4 ? ra[q][e] : rb[q][e - 4]);
          const _Float16 h = _Float16(v);
          hi[e] = h;
          lo[e] = _Float16(v - float(h));
        }
        const int slot = (st + TPB * q) >> 1;
        U[hh * US + slot] = hi;
        U[2 * US + hh * US + slot] = lo;
      }
#pragma unroll
    for (int q = 0; q < NM_Q; q++)
      if (tid + NT * q < NP * 2 * 64) Ms[tid + NT * q] = rm[q];
    __syncthreads();
    if (k + 1 < kend) fetch(k + 1);
    v4 ahh[TPW], ahl[TPW], alh[TPW];
#pragma unroll
    for (int t = 0; t < TPW; t++) ahh[t] = ahl[t] = alh[t] = v4{0, 0, 0, 0};
#pragma unroll
    for (int pm = 0; pm < NP; pm++) {
      // the operator pair is read once and applied to the wave's TPW tiles
      const h8 ah = Ms[(2 * pm) * 64 + lane], al = Ms[(2 * pm + 1) * 64 + lane];
#pragma unroll
      for (int t = 0; t < TPW; t++) {
        const h8 bh = U[uoff[t][pm]], bl = U[2 * US + uoff[t][pm]];
        ahh[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, ahh[t], 0, 0, 0);
        ahl[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, ahl[t], 0, 0, 0);
        alh[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, alh[t], 0, 0, 0);
      }
    }
    const float sc = hscale[k] * isb;
    // lane (g, j) register q holds row 4 g + q of target j: (Re, Im) of
    // children 2g and 2g + 1, one 16-byte store
#pragma unroll
    for (int t = 0; t < TPW; t++) {
      const v4 acc = (ahh[t] + (ahl[t] + alh[t])) * sc;
      if (trg[t] >= 0)
        *reinterpret_cast<v4*>(fout + (size_t(trg[t]) * m.nfreq + k) * NCHILD + 2 * g) = acc;
    }
  }
}

// ------------------------------------------------ per-node MFMA transforms
//
// One workgroup per M2L node (its 8 children), in kx slabs.  Each 1-D DFT
// restricted to the index ranges that matter (input [0,p), output [0,2p) for
// the forward transform; input [0,2p), output [p,2p) for the inverse) is a
// real matrix product: rows = (output frequency or point, re/im), inner =
// (input index, re/im), columns = (other index, child).  The DFT matrix is
// the MFMA A operand and stays in registers for the whole kernel; B comes
// from LDS (or, for the inverse z pass, straight from the spectra).  The x
// direction is done on the VALU, sparsely: forward, only the surface points
// of each x line contribute (line table); inverse, the c2r sum is evaluated
// only at the surface points.
//
// Row order inside a 16-row tile: chosen so that a lane's result registers
// hold the (re, im) pair of one output, RowMap::fa/ra for the A operand rows
// and RowMap::fd for the result registers (q & 1 = re/im).
template <typename R>
struct RowMap;
template <>
struct RowMap<float> {  // v_mfma_f32_16x16x4_f32: register q of lane group g is row 4g + q
  static __device__ __forceinline__ int fa(int j) { return j >> 1; }
  static __device__ __forceinline__ int ra(int j) { return j & 1; }
  static __device__ __forceinline__ int fd(int g, int q) { return 2 * g + (q >> 1); }
};
template <>
struct RowMap<double> {  // v_mfma_f64_16x16x4_f64: register q of lane group g is row g + 4q
  static __device__ __forceinline__ int fa(int j) { return (j & 3) + 4 * (j >> 3); }
  static __device__ __forceinline__ int ra(int j) { return (j >> 2) & 1; }
  static __device__ __forceinline__ int fd(int g, int q) { return g + 4 * (q >> 1); }
};

// A-operand entry of the real form of a complex DFT matrix: output (f, ro),
// input (v, ri), phase sign * 2 pi * ((f' v') mod n1) / n1.
__device__ __forceinline__ double dft_entry(int ph, int n1, int sign, int ro, int ri) {
  double sn, cs;
  sincospi(double(sign) * 2.0 * double(ph % n1) / double(n1), &sn, &cs);
  // (re, im) out = (c re - s im, s re + c im)
  return ro == 0 ? (ri == 0 ? cs : -sn) : (ri == 0 ? sn : cs);
}

// Compact LDS storage of a child's upward equivalent density on the
// [0,p)^3 grid: the density lives on the cube's surface points only, stored
// line by line along x, lines (z, y) in z-major order; a line on a face of
// the cube (y or z in {0, p-1}) holds all p points, an interior line only
// x = 0 and x = p - 1 -- 6 (p-1)^2 + 2 = nsurf values in all.
template <int P>
__device__ __forceinline__ int surf_line_off(int z, int y) {
  constexpr int SL = 4 * P - 4;  // values of an interior z slab
  if (z == 0) return y * P;
  if (z == P - 1) return P * P + (P - 2) * SL + y * P;
  const int b = P * P + (z - 1) * SL;
  return y == 0 ? b : (y == P - 1 ? b + P + 2 * (P - 2) : b + P + 2 * (y - 1));
}
// Face line f (0 <= f < 4p - 4) of the cube: lines (z, y) with z or y in
// {0, p - 1}; interior lines are (1 + s / (p - 2), 1 + s % (p - 2)).
template <int P>
__device__ __forceinline__ void face_line(int f, int& z, int& y) {
  if (f < P) {
    z = 0;
    y = f;
  } else if (f < 2 * P) {
    z = P - 1;
    y = f - P;
  } else {
    z = 1 + ((f - 2 * P) >> 1);
    y = ((f - 2 * P) & 1) ? P - 1 : 0;
  }
}
template <int P>
__device__ __forceinline__ int surf_compact(int v) {  // v = x + P (y + P z), a surface point
  const int x = v % P, y = (v / P) % P, z = v / (P * P);
  const bool full = z == 0 || z == P - 1 || y == 0 || y == P - 1;
  return surf_line_off<P>(z, y) + (full ? x : (x == 0 ? 0 : 1));
}

// fft_up_equiv for one node: out[(kx n1^2 + ky + n1 kz) * 8 + c] = sum over
// the child's surface points (x,y,z) of up[c][i] exp(-2 pi i (kx x + ky y +
// kz z) / n1), kx over the r2c half axis [0, p] for the real kernels
// (reference fmm_scale_invariant.h:560-604) and over [0, 2p) for Helmholtz,
// whose densities are complex (c2c, fmm.h:631-711).
template <typename RS, typename R, int P, bool CPLX>
__global__ __launch_bounds__(256) void k_fft_up_mfma(M2LDev m, DevTree t, const int* __restrict__ src_node,
                                                     int nslots) {
  typedef typename MfmaT<R>::v4 v4;
  constexpr int n1 = 2 * P, nh = P + 1;
  constexpr int NKX = CPLX ? n1 : nh;               // kx slabs
  constexpr int VW = CPLX ? 2 : 1;                  // reals per density value
  constexpr int KK = 2 * P, KS = (KK + 3) / 4;      // inner: (y or z, re/im)
  constexpr int MT = (2 * n1 + 15) / 16;             // rows: (ky or kz, re/im)
  constexpr int NCY = 8 * P, NTY = (NCY + 15) / 16;  // y pass columns (z, c)
  constexpr int NCZ = 8 * n1, NTZ = NCZ / 16;        // z pass columns (ky, c)
  constexpr int NS = 6 * (P - 1) * (P - 1) + 2;     // surface points per child
  constexpr int NFL = 4 * P - 4, NSL = (P - 2) * (P - 2);  // face lines, interior lines
  extern __shared__ char smem_raw[];
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);  // exp(-2 pi i m / n1)
  R* gx = reinterpret_cast<R*>(w + n1);           // [c][compact surface index] (x VW reals)
  R* Ay = gx + NCHILD * NS * VW;                  // [(y, re/im)][(z, c)]
  R* Bz = Ay + KK * NCY;                          // [(z, re/im)][(ky, c)]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  make_twiddles(w, n1, -1);
  const int node = src_node[slot];
  const double* up = t.up + size_t(t.child0[node]) * NS * VW;
  // every compact slot is one surface point (map_up is one-to-one onto the
  // cube's surface, checked when the context is created)
  for (int e = tid; e < NCHILD * NS; e += 256) {
    const int c = e / NS, i = e - c * NS;
    R* gp = gx + (c * NS + surf_compact<P>(m.map_up[i])) * VW;
    gp[0] = R(up[size_t(e) * VW]);
    if constexpr (CPLX) gp[1] = R(up[2 * size_t(e) + 1]);
  }
  R wf[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int f = mt * 8 + RowMap<R>::fa(j), k = ks * 4 + g, v = k >> 1;
      wf[mt][ks] = (f < n1 && k < KK) ? R(dft_entry(f * v, n1, -1, RowMap<R>::ra(j), k & 1)) : R(0);
    }
  __syncthreads();
  C2<RS>* out = reinterpret_cast<C2<RS>*>(m.fft_in) + size_t(slot) * m.nfreq * NCHILD;
  float mx = 0.f;
  for (int kx = 0; kx < NKX; kx++) {
    // x: Ay[(y, re/im)][(z, c)] = sum_x gx[c][z][y][x] exp(-2 pi i kx x / n1),
    // the face lines (all p points) first, then the interior lines (x = 0
    // and x = p - 1 only), so that no wave mixes the two
    C2<R> tw[P];
#pragma unroll
    for (int x = 0; x < P; x++) tw[x] = w[(kx * x) % n1];
    auto line_sum = [&](const R* gl, int x, C2<R>& acc, const C2<R>& tx) {
      if constexpr (CPLX) {
        cfma(acc, C2<R>{gl[2 * x], gl[2 * x + 1]}, tx);
      } else {
        acc.x = fma(gl[x], tx.x, acc.x);
        acc.y = fma(gl[x], tx.y, acc.y);
      }
    };
    for (int e = tid; e < NCHILD * NFL; e += 256) {
      const int c = e & 7;
      int z, y;
      face_line<P>(e >> 3, z, y);
      const R* gl = gx + (c * NS + surf_line_off<P>(z, y)) * VW;
      C2<R> acc{0, 0};
#pragma unroll
      for (int x = 0; x < P; x++) line_sum(gl, x, acc, tw[x]);
      Ay[(2 * y) * NCY + z * 8 + c] = acc.x;
      Ay[(2 * y + 1) * NCY + z * 8 + c] = acc.y;
    }
    for (int e = tid; e < NCHILD * NSL; e += 256) {
      const int c = e & 7, sl = e >> 3, z = 1 + sl / (P - 2), y = 1 + sl % (P - 2);
      const R* gl = gx + (c * NS + surf_line_off<P>(z, y)) * VW;
      C2<R> acc{0, 0};
      line_sum(gl, 0, acc, tw[0]);
      line_sum(gl, 1, acc, tw[P - 1]);
      Ay[(2 * y) * NCY + z * 8 + c] = acc.x;
      Ay[(2 * y + 1) * NCY + z * 8 + c] = acc.y;
    }
    __syncthreads();
    // y: Bz[(z, re/im)][(ky, c)] = sum_y W(ky, y) Ay[(y, .)][(z, c)]
    for (int nt = wave; nt < NTY; nt += 4) {
      const int col = nt * 16 + j;
      R b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int k = ks * 4 + g;
        b[ks] = (k < KK && col < NCY) ? Ay[k * NCY + col] : R(0);
      }
      const int z = col >> 3, c = col & 7;
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) acc = MfmaT<R>::mma(wf[mt][ks], b[ks], acc);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int ky = mt * 8 + RowMap<R>::fd(g, q);
          if (ky < n1 && col < NCY) Bz[(2 * z + (q & 1)) * NCZ + ky * 8 + c] = acc[q];
        }
      }
    }
    __syncthreads();
    // z: out(kx, ky, kz)[c] = sum_z W(kz, z) Bz[(z, .)][(ky, c)]
    for (int nt = wave; nt < NTZ; nt += 4) {
      const int col = nt * 16 + j, ky = col >> 3, c = col & 7;
      R b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int k = ks * 4 + g;
        b[ks] = (k < KK) ? Bz[k * NCZ + col] : R(0);
      }
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) acc = MfmaT<R>::mma(wf[mt][ks], b[ks], acc);
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int kz = mt * 8 + RowMap<R>::fd(g, 2 * h);
          if (kz < n1) {
            const C2<RS> v{RS(acc[2 * h]), RS(acc[2 * h + 1])};
            out[size_t(kx * n1 * n1 + ky + n1 * kz) * NCHILD + c] = v;
            mx = fmaxf(mx, fmaxf(fabsf(float(v.x)), fabsf(float(v.y))));
          }
        }
      }
    }
    // the next x pass writes Ay only: its last readers (y pass) are behind
    // the barrier above
  }
  if (m.amax) wave_amax(mx, m.amax);
}

// ifft_dn_check for one node: dn[c][i] += scale * sum over (kx, ky, kz) of
// w(kx) Re(in(kx,ky,kz)[c] exp(+2 pi i (kx x + ky y + kz z) / n1)) at the
// child's check-surface points (x,y,z) in [p,2p)^3, w = 1 for kx = 0 and the
// Nyquist plane, else 2 (c2r of the half spectrum; reference
// fmm_scale_invariant.h:606-646).  Helmholtz (CPLX): the complex sum over
// the full kx axis with w = 1 (c2c, fmm.h:713-788).
template <typename RS, typename R, int P, bool CPLX>
__global__ __launch_bounds__(256) void k_ifft_dn_mfma(M2LDev m, DevTree t, const int* __restrict__ trg_node,
                                                      const double* __restrict__ scale, int nslots) {
  typedef typename MfmaT<R>::v4 v4;
  constexpr int n1 = 2 * P, nh = P + 1;
  constexpr int NKX = CPLX ? n1 : nh;                // kx slabs
  constexpr int VW = CPLX ? 2 : 1;                   // reals per check potential
  constexpr int KK = 2 * n1;                         // inner: (kz or ky, re/im)
  constexpr int KS = 2 * ((n1 + 3) / 4);             // k-steps (pairs of steps share a frequency)
  constexpr int MT = (2 * P + 15) / 16;              // rows: (z or y in [p,2p), re/im)
  constexpr int NCZ = 8 * n1, NTZ = NCZ / 16;        // z pass columns (ky, c)
  constexpr int NTZW = (NTZ + 3) / 4;                // z tiles per wave (at most)
  constexpr int NCY = 8 * P, NTY = (NCY + 15) / 16;  // y pass columns (z, c)
  constexpr int NFL = 4 * P - 4, NSL = (P - 2) * (P - 2);  // face lines, interior lines of the cube
  constexpr int NLF = (NCHILD * NFL + 255) / 256;    // face lines per thread
  constexpr int NLS = (NCHILD * NSL + 255) / 256;    // interior lines per thread
  extern __shared__ char smem_raw[];
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  const int ns = m.nsurf;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);   // exp(+2 pi i m / n1)
  R* Bp = reinterpret_cast<R*>(w + n1);            // [(ky, re/im)][(z, c)]
  C2<R>* Ap = reinterpret_cast<C2<R>*>(Bp + KK * NCY);  // [z][y][c]
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  make_twiddles(w, n1, +1);
  // inner index of k-step ks, lane group g: frequency (ks / 2) 4 + g, re/im
  // ks % 2, so that one complex input feeds a lane's two consecutive k-steps
  R wf[MT][KS];
#pragma unroll
  for (int mt = 0; mt < MT; mt++)
#pragma unroll
    for (int ks = 0; ks < KS; ks++) {
      const int f = mt * 8 + RowMap<R>::fa(j), v = (ks >> 1) * 4 + g;
      wf[mt][ks] = (f < P && v < n1) ? R(dft_entry(v * (P + f), n1, +1, RowMap<R>::ra(j), ks & 1)) : R(0);
    }
  // x results of this thread, accumulated over the kx slabs in registers:
  // only the check surface is needed, i.e. the grid points x in [p, 2p) of
  // the face lines (z or y in {p, 2p - 1}) and the two end points of the
  // interior lines.  Face line items e = tid + 256 r (line e / 8, child e % 8)
  // keep p values, interior line items two.
  R gf[NLF][P][VW], gs[NLS][2][VW];
#pragma unroll
  for (int r = 0; r < NLF; r++)
#pragma unroll
    for (int x = 0; x < P; x++)
#pragma unroll
      for (int v = 0; v < VW; v++) gf[r][x][v] = R(0);
#pragma unroll
  for (int r = 0; r < NLS; r++)
#pragma unroll
    for (int x = 0; x < 2; x++)
#pragma unroll
      for (int v = 0; v < VW; v++) gs[r][x][v] = R(0);
  const C2<RS>* in = reinterpret_cast<const C2<RS>*>(m.fft_out) + size_t(slot) * m.nfreq * NCHILD;
  // z-pass input fragments of this wave's tiles, loaded one slab ahead
  RS bz[NTZW][KS];
  auto fetch = [&](int kx) {
#pragma unroll
    for (int s = 0; s < NTZW; s++) {
      const int nt = wave + 4 * s;
      const int col = nt * 16 + j, ky = col >> 3, c = col & 7;
#pragma unroll
      for (int kp = 0; kp < KS / 2; kp++) {
        C2<RS> v{RS(0), RS(0)};
        if (nt < NTZ && kp * 4 + g < n1) v = in[size_t(kx * n1 * n1 + ky + n1 * (kp * 4 + g)) * NCHILD + c];
        bz[s][2 * kp] = v.x;
        bz[s][2 * kp + 1] = v.y;
      }
    }
  };
  fetch(0);
  __syncthreads();
  for (int kx = 0; kx < NKX; kx++) {
    // z: Bp[(ky, re/im)][(z, c)] = sum_kz W(z, kz) in(kx, ky, kz)[c]
#pragma unroll
    for (int s = 0; s < NTZW; s++) {
      const int nt = wave + 4 * s;
      if (nt >= NTZ) continue;
      const int col = nt * 16 + j, ky = col >> 3, c = col & 7;
      R b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) b[ks] = R(bz[s][ks]);
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) acc = MfmaT<R>::mma(wf[mt][ks], b[ks], acc);
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int z = mt * 8 + RowMap<R>::fd(g, q);
          if (z < P) Bp[(2 * ky + (q & 1)) * NCY + z * 8 + c] = acc[q];
        }
      }
    }
    if (kx + 1 < NKX) fetch(kx + 1);
    __syncthreads();
    // y: Ap[z][y][c] = sum_ky W(y, ky) Bp[(ky, .)][(z, c)]
    for (int nt = wave; nt < NTY; nt += 4) {
      const int col = nt * 16 + j;
      R b[KS];
#pragma unroll
      for (int ks = 0; ks < KS; ks++) {
        const int v = (ks >> 1) * 4 + g;
        b[ks] = (col < NCY && v < n1) ? Bp[(2 * v + (ks & 1)) * NCY + col] : R(0);
      }
      const int z = col >> 3, c = col & 7;
#pragma unroll
      for (int mt = 0; mt < MT; mt++) {
        v4 acc = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) acc = MfmaT<R>::mma(wf[mt][ks], b[ks], acc);
#pragma unroll
        for (int h = 0; h < 2; h++) {
          const int y = mt * 8 + RowMap<R>::fd(g, 2 * h);
          if (y < P && col < NCY) Ap[(z * P + y) * NCHILD + c] = C2<R>{acc[2 * h], acc[2 * h + 1]};
        }
      }
    }
    __syncthreads();
    // x over each line: c2r with weight 1 for kx = 0 and the Nyquist plane,
    // else 2; c2c (Helmholtz) with weight 1 (the twiddles are wave-uniform)
    const R wt = (CPLX || kx == 0 || kx == nh - 1) ? R(1) : R(2);
    C2<R> tw[P];
#pragma unroll
    for (int x = 0; x < P; x++) {
      const C2<R> v = w[(kx * (P + x)) % n1];
      tw[x] = C2<R>{wt * v.x, wt * v.y};
    }
    auto acc_x = [&](R* gv, const C2<R>& a, const C2<R>& tx) {
      gv[0] += a.x * tx.x - a.y * tx.y;
      if constexpr (CPLX) gv[VW - 1] += a.x * tx.y + a.y * tx.x;
    };
#pragma unroll
    for (int r = 0; r < NLF; r++) {
      const int e = tid + 256 * r;
      if (e < NCHILD * NFL) {
        int z, y;
        face_line<P>(e >> 3, z, y);
        const C2<R> a = Ap[(z * P + y) * NCHILD + (e & 7)];
#pragma unroll
        for (int x = 0; x < P; x++) acc_x(gf[r][x], a, tw[x]);
      }
    }
#pragma unroll
    for (int r = 0; r < NLS; r++) {
      const int e = tid + 256 * r;
      if (e < NCHILD * NSL) {
        const int sl = e >> 3, z = 1 + sl / (P - 2), y = 1 + sl % (P - 2);
        const C2<R> a = Ap[(z * P + y) * NCHILD + (e & 7)];
        acc_x(gs[r][0], a, tw[0]);
        acc_x(gs[r][1], a, tw[P - 1]);
      }
    }
    // the next z pass writes Bp only (its readers are behind the barrier
    // above); Ap is next written after the following barrier
  }
  const int node = trg_node[slot];
  const double sc = scale[slot];
  double* dn = t.dn + size_t(t.child0[node]) * ns * VW;
  // scatter the surface points of the lines (each surface point is one
  // grid point of one line)
  auto put = [&](int c, int z, int y, int x, const R* gv) {
    const int i = m.inv_dn[(z * P + y) * P + x];
    if (i >= 0) {
#pragma unroll
      for (int v = 0; v < VW; v++) dn[(size_t(c) * ns + i) * VW + v] += sc * double(gv[v]);
    }
  };
#pragma unroll
  for (int r = 0; r < NLF; r++) {
    const int e = tid + 256 * r;
    if (e >= NCHILD * NFL) continue;
    int z, y;
    face_line<P>(e >> 3, z, y);
#pragma unroll
    for (int x = 0; x < P; x++) put(e & 7, z, y, x, gf[r][x]);
  }
#pragma unroll
  for (int r = 0; r < NLS; r++) {
    const int e = tid + 256 * r;
    if (e >= NCHILD * NSL) continue;
    const int sl = e >> 3, z = 1 + sl / (P - 2), y = 1 + sl % (P - 2);
    put(e & 7, z, y, 0, gs[r][0]);
    put(e & 7, z, y, P - 1, gs[r][1]);
  }
}

// --------------------------------------------------------------- launchers

static size_t fft_up_lds(const M2LDev& m, size_t rb) {
  const int p = m.p, n1 = m.n1, nh = m.cplx ? n1 : n1 / 2 + 1;
  const size_t cb = 2 * rb;
  const size_t gsz = size_t(p) * p * p * rb * (m.cplx ? 2 : 1);
  const size_t bsz = size_t(p) * n1 * nh * cb;
  return n1 * cb + size_t(p) * p * nh * cb + (bsz > gsz ? bsz : gsz);
}
static size_t ifft_dn_lds(const M2LDev& m, size_t rb) {
  const int p = m.p, n1 = m.n1, nh = m.cplx ? n1 : n1 / 2 + 1;
  const size_t cb = 2 * rb;
  const size_t gsz = size_t(p) * p * p * rb * (m.cplx ? 2 : 1);
  const size_t bsz = size_t(p) * n1 * nh * cb;
  return n1 * cb + (bsz > gsz ? bsz : gsz) + size_t(p) * p * nh * cb;
}

size_t m2l_lds_bytes(const M2LDev& m) {
  const size_t rb = m.mixed ? 4 : 8;
  if (fft_node_ok(m)) return fft_node_lds(m, (m.mixed && m.mixed_fft) ? 4 : 8);
  const size_t a = fft_up_lds(m, rb), b = ifft_dn_lds(m, rb);
  return a > b ? a : b;
}

static int xcd_grid(int nslots) { return ((nslots + 7) / 8) * 64; }

// Dynamic LDS above 64 KiB must be opted into per kernel.
template <typename K>
static void allow_lds(K kernel, size_t bytes) {
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              int(bytes));
}

void launch_fft_up(const M2LDev& m, const DevTree& t, const int* src_node, int nslots, hipStream_t s) {
  if (nslots <= 0) return;
  if (fft_node_ok(m)) {
    const size_t lds = fft_node_lds(m, (m.mixed && m.mixed_fft) ? 4 : 8);
    auto go = [&](auto kern) {
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3(nslots), dim3(256), lds, s, m, t, src_node, nslots);
    };
    node_dispatch(m.p, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      if (m.cplx) {
        if (m.mixed && m.mixed_fft) go(k_fft_up_mfma<float, float, PP, true>);
        else if (m.mixed) go(k_fft_up_mfma<float, double, PP, true>);
        else go(k_fft_up_mfma<double, double, PP, true>);
      } else {
        if (m.mixed && m.mixed_fft) go(k_fft_up_mfma<float, float, PP, false>);
        else if (m.mixed) go(k_fft_up_mfma<float, double, PP, false>);
        else go(k_fft_up_mfma<double, double, PP, false>);
      }
    });
    return;
  }
  const size_t lds = fft_up_lds(m, m.mixed ? 4 : 8);
  dim3 grid(xcd_grid(nslots));
  if (m.mixed) {
    if (m.cplx) allow_lds(k_fft_up<float, true>, lds); else allow_lds(k_fft_up<float, false>, lds);
    if (m.cplx) hipLaunchKernelGGL((k_fft_up<float, true>), grid, dim3(256), lds, s, m, t, src_node, nslots);
    else hipLaunchKernelGGL((k_fft_up<float, false>), grid, dim3(256), lds, s, m, t, src_node, nslots);
  } else {
    if (m.cplx) allow_lds(k_fft_up<double, true>, lds); else allow_lds(k_fft_up<double, false>, lds);
    if (m.cplx) hipLaunchKernelGGL((k_fft_up<double, true>), grid, dim3(256), lds, s, m, t, src_node, nslots);
    else hipLaunchKernelGGL((k_fft_up<double, false>), grid, dim3(256), lds, s, m, t, src_node, nslots);
  }
}

void launch_ifft_dn(const M2LDev& m, const DevTree& t, const int* trg_node, const double* scale, int nslots,
                    hipStream_t s) {
  if (nslots <= 0) return;
  if (fft_node_ok(m)) {
    const size_t lds = fft_node_lds(m, (m.mixed && m.mixed_fft) ? 4 : 8);
    auto go = [&](auto kern) {
      allow_lds(kern, lds);
      hipLaunchKernelGGL(kern, dim3(nslots), dim3(256), lds, s, m, t, trg_node, scale, nslots);
    };
    node_dispatch(m.p, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      if (m.cplx) {
        if (m.mixed && m.mixed_fft) go(k_ifft_dn_mfma<float, float, PP, true>);
        else if (m.mixed) go(k_ifft_dn_mfma<float, double, PP, true>);
        else go(k_ifft_dn_mfma<double, double, PP, true>);
      } else {
        if (m.mixed && m.mixed_fft) go(k_ifft_dn_mfma<float, float, PP, false>);
        else if (m.mixed) go(k_ifft_dn_mfma<float, double, PP, false>);
        else go(k_ifft_dn_mfma<double, double, PP, false>);
      }
    });
    return;
  }
  const size_t lds = ifft_dn_lds(m, m.mixed ? 4 : 8);
  dim3 grid(xcd_grid(nslots));
  if (m.mixed) {
    if (m.cplx) allow_lds(k_ifft_dn<float, true>, lds); else allow_lds(k_ifft_dn<float, false>, lds);
    if (m.cplx) hipLaunchKernelGGL((k_ifft_dn<float, true>), grid, dim3(256), lds, s, m, t, trg_node, scale, nslots);
    else hipLaunchKernelGGL((k_ifft_dn<float, false>), grid, dim3(256), lds, s, m, t, trg_node, scale, nslots);
  } else {
    if (m.cplx) allow_lds(k_ifft_dn<double, true>, lds); else allow_lds(k_ifft_dn<double, false>, lds);
    if (m.cplx) hipLaunchKernelGGL((k_ifft_dn<double, true>), grid, dim3(256), lds, s, m, t, trg_node, scale, nslots);
    else hipLaunchKernelGGL((k_ifft_dn<double, false>), grid, dim3(256), lds, s, m, t, trg_node, scale, nslots);
  }
}

void launch_hadamard(const M2LDev& m, const void* m2l_opsx, const float* hscale, const HadBlocks& hb, int kch,
                     int groups, int tiles, hipStream_t s) {
  if (hb.nblk <= 0) return;
  if (kch < 1) kch = HAD_KCH;
  const int nchunk = (m.nfreq + kch - 1) / kch;
  const int nchunk8 = (nchunk + 7) / 8 * 8;
  dim3 grid(unsigned(nchunk8) * unsigned(hb.nblk));
  if (m.mixed && hscale && m.amax) {
    auto go = [&](auto kern, int gp, int tpw) {
      dim3 gg(unsigned(nchunk8) * unsigned((hb.nblk + gp - 1) / gp));
      hipLaunchKernelGGL(kern, gg, dim3(256 * gp / tpw), 0, s, m, static_cast<const h8*>(m2l_opsx), hscale, hb,
                         nchunk, kch);
    };
    if (tiles >= 2) {
      if (groups >= 4) go(k_hadamard_h<4, 2>, 4, 2);
      else if (groups == 2) go(k_hadamard_h<2, 2>, 2, 2);
      else go(k_hadamard_h<1, 2>, 1, 2);
    } else {
      if (groups >= 4) go(k_hadamard_h<4, 1>, 4, 1);
      else if (groups == 2) go(k_hadamard_h<2, 1>, 2, 1);
      else go(k_hadamard_h<1, 1>, 1, 1);
    }
    return;
  }
  else if (m.mixed)
    hipLaunchKernelGGL((k_hadamard<float>), grid, dim3(256), 0, s, m, static_cast<const float*>(m2l_opsx), hb, nchunk,
                       kch);
  else
    hipLaunchKernelGGL((k_hadamard<double>), grid, dim3(256), 0, s, m, static_cast<const double*>(m2l_opsx), hb,
                       nchunk, kch);
}

}  // namespace xfmm
