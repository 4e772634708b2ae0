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
  for (int f = threadIdx.x; f < m.nfreq; f += blockDim.x) {
    const int kx = f % nh, ky = (f / nh) % n1, kz = f / (nh * n1);
    C2<R> acc{0, 0};
    for (int z = 0; z < p; z++) cfma(acc, B[(z * n1 + ky) * nh + kx], w[(kz * z) % n1]);
    out[size_t(f) * NCHILD + c] = acc;
  }
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

template <typename RS, typename R, int P>
__global__ __launch_bounds__(256) void k_fft_up_node(M2LDev m, DevTree t, const int* __restrict__ src_node,
                                                     int nslots) {
  extern __shared__ char smem_raw[];
  constexpr int p = P, n1 = 2 * P, nh = P + 1, p3 = P * P * P;
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);  // n1
  C2<R>* A = w + n1;                              // [c][z][y]   8 p^2
  C2<R>* B = A + NCHILD * p * p;                  // [c][z][ky]  8 p n1
  R* g = reinterpret_cast<R*>(B + NCHILD * p * n1);  // [c][z][y][x] 8 p^3
  make_twiddles(w, n1, -1);
  for (int e = threadIdx.x; e < NCHILD * p3; e += blockDim.x) g[e] = R(0);
  __syncthreads();
  const int node = src_node[slot];
  const double* up = t.up + size_t(t.child0[node]) * m.nsurf;
  for (int e = threadIdx.x; e < NCHILD * m.nsurf; e += blockDim.x) {
    const int c = e / m.nsurf, i = e % m.nsurf;
    g[c * p3 + m.map_up[i]] = R(up[e]);
  }
  __syncthreads();
  C2<RS>* out = reinterpret_cast<C2<RS>*>(m.fft_in) + size_t(slot) * m.nfreq * NCHILD;
  for (int kx = 0; kx < nh; kx++) {
    // x pass: A[c][z][y] = sum_x g[c][z][y][x] W^(kx x)
    for (int e = threadIdx.x; e < NCHILD * p * p; e += blockDim.x) {
      const R* gl = g + e * p;
      C2<R> acc{0, 0};
#pragma unroll
      for (int x = 0; x < p; x++) {
        const C2<R> tw = w[(kx * x) % n1];
        acc.x = fma(gl[x], tw.x, acc.x);
        acc.y = fma(gl[x], tw.y, acc.y);
      }
      A[e] = acc;
    }
    __syncthreads();
    // y pass: thread (c, z): B[c][z][ky] = sum_y A[c][z][y] W^(ky y)
    for (int e = threadIdx.x; e < NCHILD * p; e += blockDim.x) {
      C2<R> a[p];
#pragma unroll
      for (int y = 0; y < p; y++) a[y] = A[e * p + y];
      for (int ky = 0; ky < n1; ky++) {
        C2<R> acc{0, 0};
#pragma unroll
        for (int y = 0; y < p; y++) cfma(acc, a[y], w[(ky * y) % n1]);
        B[e * n1 + ky] = acc;
      }
    }
    __syncthreads();
    // z pass: thread (c, ky) with c fastest; out[kx + nh (ky + n1 kz)][c]
    for (int e = threadIdx.x; e < NCHILD * n1; e += blockDim.x) {
      const int c = e & 7, ky = e >> 3;
      C2<R> b[p];
#pragma unroll
      for (int z = 0; z < p; z++) b[z] = B[(c * p + z) * n1 + ky];
      for (int kz = 0; kz < n1; kz++) {
        C2<R> acc{0, 0};
#pragma unroll
        for (int z = 0; z < p; z++) cfma(acc, b[z], w[(kz * z) % n1]);
        out[size_t(kx + nh * (ky + n1 * kz)) * NCHILD + c] = C2<RS>{RS(acc.x), RS(acc.y)};
      }
    }
    __syncthreads();
  }
}

template <typename RS, typename R, int P>
__global__ __launch_bounds__(256) void k_ifft_dn_node(M2LDev m, DevTree t, const int* __restrict__ trg_node,
                                                      const double* __restrict__ scale, int nslots) {
  extern __shared__ char smem_raw[];
  constexpr int p = P, n1 = 2 * P, nh = P + 1, p3 = P * P * P;
  const int slot = blockIdx.x;
  if (slot >= nslots) return;
  C2<R>* w = reinterpret_cast<C2<R>*>(smem_raw);  // n1, exp(+2 pi i m / n1)
  C2<R>* Bp = w + n1;                             // [c][zz][ky]  8 p n1
  C2<R>* Ap = Bp + NCHILD * p * n1;               // [c][zz][yy]  8 p^2
  R* G = reinterpret_cast<R*>(Ap + NCHILD * p * p);  // [c][zz][yy][xx] 8 p^3
  make_twiddles(w, n1, +1);
  for (int e = threadIdx.x; e < NCHILD * p3; e += blockDim.x) G[e] = R(0);
  __syncthreads();
  const C2<RS>* in = reinterpret_cast<const C2<RS>*>(m.fft_out) + size_t(slot) * m.nfreq * NCHILD;
  for (int kx = 0; kx < nh; kx++) {
    // inverse z pass, thread (c, ky): z' = p + zz
    for (int e = threadIdx.x; e < NCHILD * n1; e += blockDim.x) {
      const int c = e & 7, ky = e >> 3;
      C2<R> v[n1];
#pragma unroll
      for (int kz = 0; kz < n1; kz++) {
        const C2<RS> u = in[size_t(kx + nh * (ky + n1 * kz)) * NCHILD + c];
        v[kz] = C2<R>{R(u.x), R(u.y)};
      }
      for (int zz = 0; zz < p; zz++) {
        C2<R> acc{0, 0};
#pragma unroll
        for (int kz = 0; kz < n1; kz++) cfma(acc, v[kz], w[(kz * (p + zz)) % n1]);
        Bp[(c * p + zz) * n1 + ky] = acc;
      }
    }
    __syncthreads();
    // inverse y pass, thread (c, zz): y' = p + yy
    for (int e = threadIdx.x; e < NCHILD * p; e += blockDim.x) {
      C2<R> v[n1];
#pragma unroll
      for (int ky = 0; ky < n1; ky++) v[ky] = Bp[e * n1 + ky];
      for (int yy = 0; yy < p; yy++) {
        C2<R> acc{0, 0};
#pragma unroll
        for (int ky = 0; ky < n1; ky++) cfma(acc, v[ky], w[(ky * (p + yy)) % n1]);
        Ap[e * p + yy] = acc;
      }
    }
    __syncthreads();
    // x (c2r) accumulation: weight 1 for kx = 0 and the Nyquist plane, else 2
    const R wt = (kx == 0 || kx == nh - 1) ? R(1) : R(2);
    for (int e = threadIdx.x; e < NCHILD * p * p; e += blockDim.x) {
      const C2<R> a = Ap[e];
      R* Gl = G + e * p;
#pragma unroll
      for (int xx = 0; xx < p; xx++) {
        const C2<R> tw = w[(kx * (p + xx)) % n1];
        Gl[xx] += wt * (a.x * tw.x - a.y * tw.y);
      }
    }
    __syncthreads();
  }
  const int node = trg_node[slot];
  const double sc = scale[slot];
  double* dn = t.dn + size_t(t.child0[node]) * m.nsurf;
  for (int e = threadIdx.x; e < NCHILD * m.nsurf; e += blockDim.x) {
    const int c = e / m.nsurf, i = e % m.nsurf;
    dn[e] += sc * double(G[c * p3 + m.map_dn[i]]);
  }
}

static size_t fft_node_lds(const M2LDev& m, size_t rb) {
  const size_t p = m.p, n1 = m.n1;
  return n1 * 2 * rb + NCHILD * p * p * 2 * rb + NCHILD * p * n1 * 2 * rb + NCHILD * p * p * p * rb;
}

// per-node transforms: real kernels, p in [3, 12], LDS within the CU's 160 KiB
static bool fft_node_ok(const M2LDev& m) {
  return !m.cplx && m.p >= 3 && m.p <= 12 && fft_node_lds(m, (m.mixed && m.mixed_fft) ? 4 : 8) <= 160 * 1024;
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
// HAD_KCH frequencies.  The union of the block's M2L sources (<= HAD_UCAP) is
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

// Per block of <= 64 target slots and a chunk of HAD_KCH frequencies.  For
// each frequency the block stages the spectra of its source union and the 26
// operator matrices in LDS, arranged so that every lane fetches its MFMA
// operands for one relative position with one vector read each: children g
// and g+4 sit side by side, {Re c_g, Im c_g, Re c_(g+4), Im c_(g+4)}.  The
// next frequency's spectra and operators are loaded into registers while the
// current one computes.  Each wave owns 16 targets; per relative position the
// complex 8x8 product is one real 16x16x16 product (four 16x16x4 MFMAs):
//   rows 0..7 of the result = Re of target children, rows 8..15 = Im,
//   A[row][s] = Re/Im of M[s][row & 7], B[s][t] = Re/Im of the source spectrum.
// A lane's union index for each relative position is fixed for the block and
// held in registers; an absent source points at a zero row of the staging
// array, so the 26 relative positions run without branches.
template <typename R>
__global__ __launch_bounds__(256) void k_hadamard(M2LDev m, const R* __restrict__ ops, HadBlocks hb, int nchunk) {
  typedef typename MfmaT<R>::v4 v4;
  typedef typename Quad<R>::t q4;
  __shared__ q4 Us[(HAD_UCAP + 1) * 4];  // [union slot][g]; slot HAD_UCAP stays zero
  __shared__ q4 Ms[NM2L * 8 * 4];        // [r][target child][g]
  // XCD-aware (chunk, block) mapping: XCD x = w % 8 owns chunks x, x+8, ...
  const int w = blockIdx.x;
  const int x = w & 7, i = w >> 3;
  const int chunk = (i / hb.nblk) * 8 + x;
  const int b = i % hb.nblk;
  if (chunk >= nchunk) return;
  const int kbeg = chunk * HAD_KCH, kend = min(m.nfreq, kbeg + HAD_KCH);
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = tid >> 6;
  const int j = lane & 15, g = lane >> 4;
  const int nu = hb.nunion[b];
  if (tid < 4) Us[HAD_UCAP * 4 + tid] = q4{0, 0, 0, 0};
  int uoff[NM2L];
#pragma unroll
  for (int r = 0; r < NM2L; r++) {
    const int u = hb.loc[(size_t(b) * NM2L + r) * 64 + wave * 16 + j];
    uoff[r] = (u >= 0 ? u : HAD_UCAP) * 4 + g;
  }
  const C2<R>* fin = reinterpret_cast<const C2<R>*>(m.fft_in);
  C2<R>* fout = reinterpret_cast<C2<R>*>(m.fft_out);
  const C2<R>* M = reinterpret_cast<const C2<R>*>(ops);
  // staging assignment: spectra element e = tid + 256 q (q < 8) is child
  // e % 8 of union slot e / 8; operator element e (q < 7) is entry e % 64 of
  // relative position e / 64
  constexpr int NU_Q = HAD_UCAP * NCHILD / 256, NM_Q = (NM2L * 64 + 255) / 256;
  const int c8 = tid & 7;
  int usrc[NU_Q];
#pragma unroll
  for (int q = 0; q < NU_Q; q++) {
    const int e = tid + 256 * q;
    usrc[q] = (e < nu * NCHILD) ? hb.uni[size_t(b) * HAD_UCAP + (e >> 3)] : -1;
  }
  C2<R> ru[NU_Q], rm[NM_Q];
  auto fetch = [&](int k) {
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) ru[q] = fin[(size_t(usrc[q]) * m.nfreq + k) * NCHILD + c8];
#pragma unroll
    for (int q = 0; q < NM_Q; q++) {
      const int e = tid + 256 * q;
      if (e < NM2L * 64) rm[q] = M[(size_t(e >> 6) * m.nfreq + k) * 64 + (e & 63)];
    }
  };
  R* Usr = reinterpret_cast<R*>(Us);
  R* Msr = reinterpret_cast<R*>(Ms);
  const int cpos = 4 * (c8 & 3) + 2 * (c8 >> 2);
  const int j1 = j & 7;
  const bool im_out = j >= 8;
  const int trg = hb.trg[size_t(b) * 64 + wave * 16 + j];
  fetch(kbeg);
  for (int k = kbeg; k < kend; k++) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NU_Q; q++)
      if (usrc[q] >= 0) {
        const int e = tid + 256 * q;
        Usr[(e >> 3) * 16 + cpos] = ru[q].x;
        Usr[(e >> 3) * 16 + cpos + 1] = ru[q].y;
      }
#pragma unroll
    for (int q = 0; q < NM_Q; q++) {
      const int e = tid + 256 * q;
      if (e < NM2L * 64) {
        // entry (s, t) = operator from source child s to target child t
        const int r = e >> 6, s = (e >> 3) & 7, t = e & 7;
        const int o = ((r * 8 + t) * 4 + (s & 3)) * 4 + 2 * (s >> 2);
        Msr[o] = rm[q].x;
        Msr[o + 1] = rm[q].y;
      }
    }
    __syncthreads();
    if (k + 1 < kend) fetch(k + 1);
    // two accumulators (even / odd r) and operands of r + 1 read while r
    // computes
    v4 acc0 = {0, 0, 0, 0}, acc1 = {0, 0, 0, 0};
    q4 mm = Ms[j1 * 4 + g];
    q4 vv = Us[uoff[0]];
#pragma unroll
    for (int r = 0; r < NM2L; r++) {
      q4 mn = mm, vn = vv;
      if (r + 1 < NM2L) {
        mn = Ms[((r + 1) * 8 + j1) * 4 + g];
        vn = Us[uoff[r + 1]];
      }
      v4& acc = (r & 1) ? acc1 : acc0;
      acc = MfmaT<R>::mma(im_out ? mm.y : mm.x, vv.x, acc);
      acc = MfmaT<R>::mma(im_out ? mm.w : mm.z, vv.z, acc);
      acc = MfmaT<R>::mma(im_out ? mm.x : -mm.y, vv.y, acc);
      acc = MfmaT<R>::mma(im_out ? mm.z : -mm.w, vv.w, acc);
      mm = mn;
      vv = vn;
    }
    const v4 acc = acc0 + acc1;
    // result: lane (g, j) register q holds row MfmaT::row(g, q) of target j;
    // rows 0..7 are Re of children 0..7, rows 8..15 Im.
    if (trg >= 0) {
      R* o = reinterpret_cast<R*>(fout + (size_t(trg) * m.nfreq + k) * NCHILD);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int rw = MfmaT<R>::row(g, q);
        o[2 * (rw & 7) + (rw >> 3)] = acc[q];
      }
    }
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
    node_dispatch(m.p, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      if (m.mixed && m.mixed_fft) {
        allow_lds(k_fft_up_node<float, float, PP>, lds);
        hipLaunchKernelGGL((k_fft_up_node<float, float, PP>), dim3(nslots), dim3(256), lds, s, m, t, src_node, nslots);
      } else if (m.mixed) {
        allow_lds(k_fft_up_node<float, double, PP>, lds);
        hipLaunchKernelGGL((k_fft_up_node<float, double, PP>), dim3(nslots), dim3(256), lds, s, m, t, src_node, nslots);
      } else {
        allow_lds(k_fft_up_node<double, double, PP>, lds);
        hipLaunchKernelGGL((k_fft_up_node<double, double, PP>), dim3(nslots), dim3(256), lds, s, m, t, src_node, nslots);
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
    node_dispatch(m.p, [&](auto P) {
      constexpr int PP = decltype(P)::value;
      if (m.mixed && m.mixed_fft) {
        allow_lds(k_ifft_dn_node<float, float, PP>, lds);
        hipLaunchKernelGGL((k_ifft_dn_node<float, float, PP>), dim3(nslots), dim3(256), lds, s, m, t, trg_node, scale,
                           nslots);
      } else if (m.mixed) {
        allow_lds(k_ifft_dn_node<float, double, PP>, lds);
        hipLaunchKernelGGL((k_ifft_dn_node<float, double, PP>), dim3(nslots), dim3(256), lds, s, m, t, trg_node, scale,
                           nslots);
      } else {
        allow_lds(k_ifft_dn_node<double, double, PP>, lds);
        hipLaunchKernelGGL((k_ifft_dn_node<double, double, PP>), dim3(nslots), dim3(256), lds, s, m, t, trg_node,
                           scale, nslots);
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

void launch_hadamard(const M2LDev& m, const void* m2l_ops, const HadBlocks& hb, hipStream_t s) {
  if (hb.nblk <= 0) return;
  const int nchunk = (m.nfreq + HAD_KCH - 1) / HAD_KCH;
  const int nchunk8 = (nchunk + 7) / 8 * 8;
  dim3 grid(unsigned(nchunk8) * unsigned(hb.nblk));
  if (m.mixed)
    hipLaunchKernelGGL((k_hadamard<float>), grid, dim3(256), 0, s, m, static_cast<const float*>(m2l_ops), hb, nchunk);
  else
    hipLaunchKernelGGL((k_hadamard<double>), grid, dim3(256), 0, s, m, static_cast<const double*>(m2l_ops), hb, nchunk);
}

}  // namespace xfmm
