// Particle/surface evaluation kernels for gfx950: P2P, M2P, L2P (particle
// targets), P2M check potentials and P2L (surface targets), and the direct
// summation used by verify().
//
// Kernel formulas follow exafmm-t's scalar loops (targets t, sources s,
// d = x_s - x_t, results scaled by 1/(4 pi) at the end):
//   Laplace   laplace.h:139-162:            phi += q/r,            grad += q d / r^3
//   Yukawa    modified_helmholtz.h:149-170: phi += q e^{-kr}/r,    grad += q e^{-kr}(kr+1) d / r^3
//   Helmholtz helmholtz.h:186-207:          phi += q e^{ikr}/r,    grad += (1/r^2 - ik/r) phi_ij d
// Pairs with r = 0 contribute nothing (laplace.h:57, `invr &= r2 > zero`).
//
// Layout: one wave (64 lanes) per target leaf / node; lane = one target (in
// chunks of 64); sources are staged through LDS in tiles of 64 and read as
// wave-uniform broadcasts.  Coordinates are made relative to the target leaf
// centre in FP64 before any FP32 conversion, so the mixed path keeps
// |d| accuracy at the scale of the leaf.  In the FP32 path each tile is
// summed in plain FP32 and tiles are combined with compensated (Kahan)
// summation.
#include "kernels.h"

namespace xfmm {

// Helmholtz with a real wave number (Im k = 0): the evaluation kernels are
// instantiated without the damping factor e^{-Im(k) r}, which is then exactly
// 1 (the values are bit-identical to the general instantiation).  Internal to
// the launchers of this file.
constexpr int HELMHOLTZ_REAL = 3;

template <int KIND>
constexpr bool kcplx() { return KIND == HELMHOLTZ || KIND == HELMHOLTZ_REAL; }

// 1/r of a pair.  RSQ_LIB: library rsqrt, r = 0 gives 0 (the direct sum used as
// accuracy check).  RSQ_MASK: hardware estimate refined to full precision
// (rsqrt_sep), r^2 <= rmin2 gives 0 (a leaf with itself, where a target may
// coincide with a source; the FP32 near field also leaves out the
// near-coincident pairs that k_close_fix adds in FP64).  RSQ_SEP: rsqrt_sep
// without the test (pairs that cannot coincide: distinct leaves own distinct
// points, and surfaces never touch particles).
enum RsqMode { RSQ_LIB = 0, RSQ_MASK = 1, RSQ_SEP = 2 };
template <int RM, typename R>
__device__ __forceinline__ R inv_r(R r2, R rmin2) {
  if constexpr (RM == RSQ_LIB) return rsqrt_pos(r2);
  else if constexpr (RM == RSQ_MASK) return r2 > rmin2 ? rsqrt_sep(r2) : R(0);
  else return rsqrt_sep(r2);
}

// potential + gradient of one pair, accumulated into a[] (4 or 8 values)
template <int KIND, typename R, int RM>
__device__ __forceinline__ void pair_grad(R dx, R dy, R dz, R qr, R qi, R kr, R ki, R* a, R rmin2 = R(0)) {
  const R r2 = dx * dx + dy * dy + dz * dz;
  const R invr = inv_r<RM>(r2, rmin2);
  if constexpr (KIND == LAPLACE) {
    const R qinv = qr * invr;
    a[0] += qinv;
    const R s = qinv * invr * invr;
    a[1] += s * dx;
    a[2] += s * dy;
    a[3] += s * dz;
  } else if constexpr (KIND == YUKAWA) {
    const R r = r2 * invr;
    const R qinv = qr * dexp(-kr * r) * invr;
    a[0] += qinv;
    const R s = qinv * (R(1) + kr * r) * invr * invr;
    a[1] += s * dx;
    a[2] += s * dy;
    a[3] += s * dz;
  } else {
    const R r = r2 * invr;
    const R e = KIND == HELMHOLTZ_REAL ? invr : dexp(-ki * r) * invr;
    R sn, cs;
    dsincos(kr * r, &sn, &cs);
    const R gr = e * cs, gi = e * sn;
    const R pr = qr * gr - qi * gi;
    const R pim = qr * gi + qi * gr;
    a[0] += pr;
    a[1] += pim;
    const R cgr = (KIND == HELMHOLTZ_REAL ? invr : (R(1) + ki * r) * invr) * invr;
    const R cgi = -kr * invr;
    const R fr = cgr * pr - cgi * pim;
    const R fi = cgr * pim + cgi * pr;
    a[2] += dx * fr;
    a[3] += dx * fi;
    a[4] += dy * fr;
    a[5] += dy * fi;
    a[6] += dz * fr;
    a[7] += dz * fi;
  }
}

// potential of one pair into a[0] (and a[1] for Helmholtz)
template <int KIND, typename R, int RM>
__device__ __forceinline__ void pair_pot(R dx, R dy, R dz, R qr, R qi, R kr, R ki, R* a) {
  const R r2 = dx * dx + dy * dy + dz * dz;
  const R invr = inv_r<RM>(r2, R(0));
  if constexpr (KIND == LAPLACE) {
    a[0] += qr * invr;
  } else if constexpr (KIND == YUKAWA) {
    a[0] += qr * dexp(-kr * (r2 * invr)) * invr;
  } else {
    const R r = r2 * invr;
    const R e = KIND == HELMHOLTZ_REAL ? invr : dexp(-ki * r) * invr;
    R sn, cs;
    dsincos(kr * r, &sn, &cs);
    const R gr = e * cs, gi = e * sn;
    a[0] += qr * gr - qi * gi;
    a[1] += qr * gi + qi * gr;
  }
}

template <typename R, int NA>
__device__ __forceinline__ void fold_tile(R* acc, R* comp, const R* part) {
  if constexpr (sizeof(R) == 4) {
#pragma unroll
    for (int a = 0; a < NA; a++) kahan_add(acc[a], comp[a], part[a]);
  } else {
#pragma unroll
    for (int a = 0; a < NA; a++) acc[a] += part[a];
  }
}

constexpr int TILE = 64;

// Sum the m (<= TILE) staged sources of a tile into part[], in source order,
// four pairs per loop trip so that the LDS reads of a trip are issued together.
template <int KIND, typename R, int RM, bool GRAD>
__device__ __forceinline__ void tile_sum(int m, const R* sx, const R* sy, const R* sz, const R* sq, const R* sqi, R tx,
                                         R ty, R tz, R kr, R ki, R* part, R rmin2 = R(0)) {
  constexpr bool C = kcplx<KIND>();
  int j = 0;
  for (; j + 4 <= m; j += 4) {
#pragma unroll
    for (int u = 0; u < 4; u++) {
      if constexpr (GRAD)
        pair_grad<KIND, R, RM>(sx[j + u] - tx, sy[j + u] - ty, sz[j + u] - tz, sq[j + u], C ? sqi[j + u] : R(0), kr, ki,
                               part, rmin2);
      else
        pair_pot<KIND, R, RM>(sx[j + u] - tx, sy[j + u] - ty, sz[j + u] - tz, sq[j + u], C ? sqi[j + u] : R(0), kr, ki,
                              part);
    }
  }
  for (; j < m; j++) {
    if constexpr (GRAD)
      pair_grad<KIND, R, RM>(sx[j] - tx, sy[j] - ty, sz[j] - tz, sq[j], C ? sqi[j] : R(0), kr, ki, part, rmin2);
    else
      pair_pot<KIND, R, RM>(sx[j] - tx, sy[j] - ty, sz[j] - tz, sq[j], C ? sqi[j] : R(0), kr, ki, part);
  }
}

// Squared distance below which the FP32 near field of a leaf with itself
// leaves a pair to k_close_fix: (side / 16)^2 of the target leaf when the
// near-coincident pairs are listed (t.close), else 0 (only r = 0 masked).
template <typename R>
__device__ __forceinline__ R close_rmin2(const DevTree& t, int leaf) {
  if constexpr (sizeof(R) == 4) {
    if (t.close) {
      const double side = 2.0 * t.r0 * ldexp(1.0, -t.level[leaf]);
      return R(side * side * (1.0 / 256.0));
    }
  }
  return R(0);
}

// ------------------------------------------------------------------ P2P
//
// Near field of the real kernels with the sources read as wave-uniform
// 16/32-byte vectors (scalar loads, no LDS and no barriers): each source is
// packed once per evaluation as (x, y, z, q) relative to its leaf centre; a
// target is moved into the frame of the source leaf with one subtraction per
// source leaf (the offset of the two centres is formed in FP64).  Pairs are
// summed per source leaf in chunks of 64 (plain sums, FP32 chunks combined
// with Kahan summation, as in the tiled kernel).
template <typename R>
struct P4;
template <>
struct P4<float> {
  typedef float4 t;
};
template <>
struct P4<double> {
  typedef double4 t;
};

template <typename R>
__global__ void k_pack_sources(int n, const double* __restrict__ rel, const double* __restrict__ q,
                               typename P4<R>::t* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  typename P4<R>::t v;
  v.x = R(rel[3 * size_t(i)]);
  v.y = R(rel[3 * size_t(i) + 1]);
  v.z = R(rel[3 * size_t(i) + 2]);
  v.w = R(q[i]);
  out[i] = v;
}

// Tail of a leaf with nt <= t.p2p_split (<= TILE / 2) targets: k = TILE / nt lanes per target,
// lane (target lane % nt, phase lane / nt) takes every k-th source from its
// phase on (vector loads); the k partial sums of a target are added in FP64
// through LDS by split_reduce.
// add (may be null): NA values per target added after the sum (the
// near field that an L2P adds to its result, see DevTree::near_add).
template <typename R, int NA>
__device__ __forceinline__ void split_reduce(const R* acc, const R* comp, int nt, int k, double* out,
                                             const double* add = nullptr) {
  __shared__ double red[TILE][NA];
  const int lane = threadIdx.x;
#pragma unroll
  for (int a = 0; a < NA; a++) red[lane][a] = double(acc[a]) - double(comp[a]);
  __syncthreads();
  if (lane < nt) {
#pragma unroll
    for (int a = 0; a < NA; a++) {
      double v = 0;
      for (int ph = 0; ph < k; ph++) v += red[lane + ph * nt][a];
      if (add) out[a] = (out[a] + v * FOUR_PI_INV) + add[a];
      else out[a] += v * FOUR_PI_INV;
    }
  }
}

// Near field of the real kernels (fmm_base.h:82-94): one wave per (leaf,
// tile of TILE targets), work[2 b] = leaf, work[2 b + 1] = tile.  A full tile
// has a target per lane and reads the packed sources (relative to their leaf
// centre) with scalar loads (constant address space); a tail tile of at most
// t.p2p_split (<= TILE / 2) targets splits the sources of each target over
// several lanes.
template <int KIND, typename R>
__global__ __launch_bounds__(64) void k_p2p_smem(DevTree t, Wavek wk, const int* __restrict__ work) {
  constexpr int NA = 4;
  // constant address space: the packed sources are read-only for the whole
  // launch, and wave-uniform loads from it become scalar loads
  typedef const R __attribute__((address_space(4))) * CR;
  typedef typename P4<R>::t V;
  const CR src = (CR)(t.src_pack);
  const V* srcv = static_cast<const V*>(t.src_pack);
  const int leaf = work[2 * blockIdx.x], t0 = TILE * work[2 * blockIdx.x + 1];
  const int lane = threadIdx.x;
  const double cx = t.center[3 * leaf], cy = t.center[3 * leaf + 1], cz = t.center[3 * leaf + 2];
  const int tb = t.trg_begin[leaf], nt = min(TILE, t.trg_count[leaf] - t0);
  const R kr = R(wk.re), ki = R(wk.im);
  const int g0 = t.p2p_off[leaf], g1 = t.p2p_off[leaf + 1];
  const bool split = nt <= t.p2p_split;
  const int k = split ? TILE / nt : 1;
  const int ti = split ? lane % nt : lane, ph = split ? lane / nt : 0;
  const bool act = split ? ph < k : ti < nt;
  const R rmin2 = close_rmin2<R>(t, leaf);
  R tx = 0, ty = 0, tz = 0;
  if (ti < nt) {
    const double* x = t.trg + 3 * size_t(tb + t0 + ti);
    tx = R(x[0] - cx);
    ty = R(x[1] - cy);
    tz = R(x[2] - cz);
  }
  R acc[NA], comp[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) { acc[a] = 0; comp[a] = 0; }
  for (int g = g0; g < g1; g++) {
    const int snode = t.p2p_src[g];
    const int sb = t.src_begin[snode], ns = t.src_count[snode];
    // target in the frame of the source leaf
    const R ux = tx - R(t.center[3 * snode] - cx);
    const R uy = ty - R(t.center[3 * snode + 1] - cy);
    const R uz = tz - R(t.center[3 * snode + 2] - cz);
    const bool self = snode == leaf;
    if (split) {
      R part[NA] = {0, 0, 0, 0};
      if (act) {
        if (self)
          for (int j = ph; j < ns; j += k) {
            const V v = srcv[sb + j];
            pair_grad<KIND, R, RSQ_MASK>(v.x - ux, v.y - uy, v.z - uz, v.w, R(0), kr, ki, part, rmin2);
          }
        else
          for (int j = ph; j < ns; j += k) {
            const V v = srcv[sb + j];
            pair_grad<KIND, R, RSQ_SEP>(v.x - ux, v.y - uy, v.z - uz, v.w, R(0), kr, ki, part);
          }
      }
      fold_tile<R, NA>(acc, comp, part);
      continue;
    }
    for (int s0 = 0; s0 < ns; s0 += TILE) {
      const int m = min(TILE, ns - s0);
      const CR S = src + 4 * size_t(sb + s0);
      R part[NA] = {0, 0, 0, 0};
      int j = 0;
      if (self) {
        for (; j + 4 <= m; j += 4) {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const CR v = S + 4 * (j + u);
            pair_grad<KIND, R, RSQ_MASK>(v[0] - ux, v[1] - uy, v[2] - uz, v[3], R(0), kr, ki, part, rmin2);
          }
        }
        for (; j < m; j++) {
          const CR v = S + 4 * j;
          pair_grad<KIND, R, RSQ_MASK>(v[0] - ux, v[1] - uy, v[2] - uz, v[3], R(0), kr, ki, part, rmin2);
        }
      } else {
        for (; j + 4 <= m; j += 4) {
#pragma unroll
          for (int u = 0; u < 4; u++) {
            const CR v = S + 4 * (j + u);
            pair_grad<KIND, R, RSQ_SEP>(v[0] - ux, v[1] - uy, v[2] - uz, v[3], R(0), kr, ki, part);
          }
        }
        for (; j < m; j++) {
          const CR v = S + 4 * j;
          pair_grad<KIND, R, RSQ_SEP>(v[0] - ux, v[1] - uy, v[2] - uz, v[3], R(0), kr, ki, part);
        }
      }
      fold_tile<R, NA>(acc, comp, part);
    }
  }
  double* o = t.trg_value + size_t(tb + t0) * 4;
  if (split) {
    split_reduce<R, NA>(acc, comp, nt, k, o + 4 * size_t(lane < nt ? lane : 0));
  } else if (act) {
#pragma unroll
    for (int a = 0; a < NA; a++) o[4 * lane + a] += (double(acc[a]) - double(comp[a])) * FOUR_PI_INV;
  }
}

// L2P of the real kernels (fmm_scale_invariant.h:198-230, fmm.h:188-218):
// one wave per (leaf, tile of TILE targets), work as for k_p2p_smem.  The
// sources, the leaf's equivalent surface points (per-level table, relative
// to the centre) and its local equivalent densities, are the same for every
// lane of a full tile and are read with scalar loads, four per unrolled
// step; a tail tile of at most TILE / 2 targets splits them over lanes.
template <int KIND, typename R>
__global__ __launch_bounds__(64) void k_l2p_smem(DevTree t, Wavek wk, const int* __restrict__ work) {
  constexpr int NA = 4;
  typedef const double __attribute__((address_space(4))) * CD;
  // surface points in the arithmetic type (no per-source conversions)
  typedef const R __attribute__((address_space(4))) * CRT;
  const int leaf = work[2 * blockIdx.x], t0 = TILE * work[2 * blockIdx.x + 1];
  const int lane = threadIdx.x;
  const double cx = t.center[3 * leaf], cy = t.center[3 * leaf + 1], cz = t.center[3 * leaf + 2];
  const int tb = t.trg_begin[leaf], nt = min(TILE, t.trg_count[leaf] - t0), ns = t.nsurf;
  const R kr = R(wk.re), ki = R(wk.im);
  const size_t tab0 = size_t(t.level[leaf]) * ns;
  const R* tabr;
  if constexpr (sizeof(R) == 4) tabr = reinterpret_cast<const R*>(t.surf_tab32);
  else tabr = reinterpret_cast<const R*>(t.surf_tab);
  const CRT sp = (CRT)(tabr) + tab0 * 4;
  const CD eq = (CD)(t.dn) + size_t(leaf) * ns;
  const bool split = nt <= t.p2p_split;
  const int k = split ? TILE / nt : 1;
  const int ti = split ? lane % nt : lane, ph = split ? lane / nt : 0;
  const bool act = split ? ph < k : ti < nt;
  R tx = 0, ty = 0, tz = 0;
  if (ti < nt) {
    const double* x = t.trg + 3 * size_t(tb + t0 + ti);
    tx = R(x[0] - cx);
    ty = R(x[1] - cy);
    tz = R(x[2] - cz);
  }
  R acc[NA], comp[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) { acc[a] = 0; comp[a] = 0; }
  double* o = t.trg_value + size_t(tb + t0) * 4;
  if (split) {
    const double4* tab = reinterpret_cast<const double4*>(t.surf_tab) + tab0;
    const double* eqv = t.dn + size_t(leaf) * ns;
    for (int s0 = 0; s0 < ns; s0 += TILE * k) {
      R part[NA] = {0, 0, 0, 0};
      const int e = min(ns, s0 + TILE * k);
      if (act)
        for (int j = s0 + ph; j < e; j += k) {
          const double4 v = tab[j];
          pair_grad<KIND, R, RSQ_SEP>(R(v.x) - tx, R(v.y) - ty, R(v.z) - tz, R(eqv[j]), R(0), kr, ki, part);
        }
      fold_tile<R, NA>(acc, comp, part);
    }
    const size_t ol = 4 * size_t(lane < nt ? lane : 0);
    split_reduce<R, NA>(acc, comp, nt, k, o + ol, t.near_add ? t.near_add + size_t(tb + t0) * 4 + ol : nullptr);
    return;
  }
  for (int s0 = 0; s0 < ns; s0 += TILE) {
    const int m = min(TILE, ns - s0);
    R part[NA] = {0, 0, 0, 0};
    int j = 0;
    for (; j + 4 <= m; j += 4) {
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const CRT v = sp + 4 * (s0 + j + u);
        pair_grad<KIND, R, RSQ_SEP>(v[0] - tx, v[1] - ty, v[2] - tz, R(eq[s0 + j + u]), R(0), kr, ki, part);
      }
    }
    for (; j < m; j++) {
      const CRT v = sp + 4 * (s0 + j);
      pair_grad<KIND, R, RSQ_SEP>(v[0] - tx, v[1] - ty, v[2] - tz, R(eq[s0 + j]), R(0), kr, ki, part);
    }
    fold_tile<R, NA>(acc, comp, part);
  }
  if (act) {
    const double* add = t.near_add ? t.near_add + size_t(tb + t0) * 4 + 4 * lane : nullptr;
#pragma unroll
    for (int a = 0; a < NA; a++) {
      const double v = (double(acc[a]) - double(comp[a])) * FOUR_PI_INV;
      if (add) o[4 * lane + a] = (o[4 * lane + a] + v) + add[a];
      else o[4 * lane + a] += v;
    }
  }
}

template <int KIND, typename R, int MODE>
__global__ __launch_bounds__(64) void k_eval_targets(DevTree t, Wavek wk, const int* __restrict__ work) {
  constexpr bool C = kcplx<KIND>();
  constexpr int NA = C ? 8 : 4;
  __shared__ R sx[TILE], sy[TILE], sz[TILE], sq[TILE], sqi[TILE];
  const int leaf = work[blockIdx.x];
  const int lane = threadIdx.x;
  const double cx = t.center[3 * leaf], cy = t.center[3 * leaf + 1], cz = t.center[3 * leaf + 2];
  const int tb = t.trg_begin[leaf], nt = t.trg_count[leaf];
  const R kr = R(wk.re), ki = R(wk.im);
  const int vw = t.vw;
  int g0 = 0, g1 = 1;
  if (MODE == MODE_P2P) { g0 = t.p2p_off[leaf]; g1 = t.p2p_off[leaf + 1]; }
  if (MODE == MODE_M2P) { g0 = t.m2p_off[leaf]; g1 = t.m2p_off[leaf + 1]; }
  for (int t0 = 0; t0 < nt; t0 += TILE) {
    const int ti = t0 + lane;
    const bool act = ti < nt;
    R tx = 0, ty = 0, tz = 0;
    if (act) {
      const double* x = t.trg + 3 * size_t(tb + ti);
      tx = R(x[0] - cx);
      ty = R(x[1] - cy);
      tz = R(x[2] - cz);
    }
    R acc[NA], comp[NA];
#pragma unroll
    for (int a = 0; a < NA; a++) { acc[a] = 0; comp[a] = 0; }
    for (int g = g0; g < g1; g++) {
      int snode, ns, sb = 0;
      double b = 0, ox = 0, oy = 0, oz = 0;
      const double* eq = nullptr;
      if (MODE == MODE_P2P) {
        snode = t.p2p_src[g];
        sb = t.src_begin[snode];
        ns = t.src_count[snode];
      } else {
        snode = (MODE == MODE_M2P) ? t.m2p_src[g] : leaf;
        ns = t.nsurf;
        b = ((MODE == MODE_M2P) ? 1.05 : 2.95) * t.r0 * ldexp(1.0, -t.level[snode]);
        ox = t.center[3 * snode] - cx;
        oy = t.center[3 * snode + 1] - cy;
        oz = t.center[3 * snode + 2] - cz;
        eq = ((MODE == MODE_M2P) ? t.up : t.dn) + size_t(snode) * t.nsurf * vw;
      }
      for (int s0 = 0; s0 < ns; s0 += TILE) {
        const int m = min(TILE, ns - s0);
        __syncthreads();
        if (lane < m) {
          const int si = s0 + lane;
          if (MODE == MODE_P2P) {
            const double* x = t.src + 3 * size_t(sb + si);
            sx[lane] = R(x[0] - cx);
            sy[lane] = R(x[1] - cy);
            sz[lane] = R(x[2] - cz);
            const double* qq = t.q + size_t(sb + si) * vw;
            sq[lane] = R(qq[0]);
            if (C) sqi[lane] = R(qq[1]);
          } else {
            sx[lane] = R(ox + t.unit_surf[3 * si] * b);
            sy[lane] = R(oy + t.unit_surf[3 * si + 1] * b);
            sz[lane] = R(oz + t.unit_surf[3 * si + 2] * b);
            sq[lane] = R(eq[size_t(si) * vw]);
            if (C) sqi[lane] = R(eq[size_t(si) * vw + 1]);
          }
        }
        __syncthreads();
        R part[NA];
#pragma unroll
        for (int a = 0; a < NA; a++) part[a] = 0;
        // only a leaf with itself can pair coinciding points; equivalent
        // surfaces never touch the targets
        if (MODE == MODE_P2P && snode == leaf)
          tile_sum<KIND, R, RSQ_MASK, true>(m, sx, sy, sz, sq, sqi, tx, ty, tz, kr, ki, part,
                                            close_rmin2<R>(t, leaf));
        else
          tile_sum<KIND, R, RSQ_SEP, true>(m, sx, sy, sz, sq, sqi, tx, ty, tz, kr, ki, part);
        fold_tile<R, NA>(acc, comp, part);
      }
    }
    if (act) {
      double* o = t.trg_value + size_t(tb + ti) * 4 * vw;
      const double* add = (MODE == MODE_L2P && t.near_add) ? t.near_add + size_t(tb + ti) * 4 * vw : nullptr;
#pragma unroll
      for (int a = 0; a < NA; a++) {
        const double v = (double(acc[a]) - double(comp[a])) * FOUR_PI_INV;
        if (add) o[a] = (o[a] + v) + add[a];
        else o[a] += v;
      }
    }
  }
}

template <int KIND, typename R, int MODE>
__global__ __launch_bounds__(64) void k_eval_surface(DevTree t, Wavek wk, const int* __restrict__ work,
                                                     double* __restrict__ check) {
  constexpr bool C = kcplx<KIND>();
  constexpr int NA = C ? 2 : 1;
  __shared__ R sx[TILE], sy[TILE], sz[TILE], sq[TILE], sqi[TILE];
  const int node = work[blockIdx.x];
  const int lane = threadIdx.x;
  const double cx = t.center[3 * node], cy = t.center[3 * node + 1], cz = t.center[3 * node + 2];
  const double b = ((MODE == MODE_P2M) ? 2.95 : 1.05) * t.r0 * ldexp(1.0, -t.level[node]);
  const R kr = R(wk.re), ki = R(wk.im);
  const int vw = t.vw;
  int g0 = 0, g1 = 1;
  if (MODE == MODE_P2L) { g0 = t.p2l_off[node]; g1 = t.p2l_off[node + 1]; }
  for (int i0 = 0; i0 < t.nsurf; i0 += TILE) {
    const int i = i0 + lane;
    const bool act = i < t.nsurf;
    R px = 0, py = 0, pz = 0;
    if (act) {
      px = R(t.unit_surf[3 * i] * b);
      py = R(t.unit_surf[3 * i + 1] * b);
      pz = R(t.unit_surf[3 * i + 2] * b);
    }
    R acc[NA], comp[NA];
#pragma unroll
    for (int a = 0; a < NA; a++) { acc[a] = 0; comp[a] = 0; }
    for (int g = g0; g < g1; g++) {
      const int snode = (MODE == MODE_P2L) ? t.p2l_src[g] : node;
      const int sb = t.src_begin[snode], ns = t.src_count[snode];
      for (int s0 = 0; s0 < ns; s0 += TILE) {
        const int m = min(TILE, ns - s0);
        __syncthreads();
        if (lane < m) {
          const double* x = t.src + 3 * size_t(sb + s0 + lane);
          sx[lane] = R(x[0] - cx);
          sy[lane] = R(x[1] - cy);
          sz[lane] = R(x[2] - cz);
          const double* qq = t.q + size_t(sb + s0 + lane) * vw;
          sq[lane] = R(qq[0]);
          if (C) sqi[lane] = R(qq[1]);
        }
        __syncthreads();
        R part[NA];
#pragma unroll
        for (int a = 0; a < NA; a++) part[a] = 0;
        // check surfaces are well separated from the source particles
        tile_sum<KIND, R, RSQ_SEP, false>(m, sx, sy, sz, sq, sqi, px, py, pz, kr, ki, part);
        fold_tile<R, NA>(acc, comp, part);
      }
    }
    if (act) {
      if (MODE == MODE_P2M) {
        double* o = check + (size_t(blockIdx.x) * t.nsurf + i) * vw;
#pragma unroll
        for (int a = 0; a < NA; a++) o[a] = (double(acc[a]) - double(comp[a])) * FOUR_PI_INV;
      } else {
        double* o = t.dn + (size_t(node) * t.nsurf + i) * vw;
#pragma unroll
        for (int a = 0; a < NA; a++) o[a] += (double(acc[a]) - double(comp[a])) * FOUR_PI_INV;
      }
    }
  }
}

// Direct sum: block (x = 64 targets, y = source split).  Differences are
// formed in FP64; tiles are combined with compensated summation in both
// precisions (this kernel is the repository's accuracy check).
template <int KIND, typename R, bool GRAD>
__global__ __launch_bounds__(64) void k_direct(const double* __restrict__ src, const double* __restrict__ q, int ns,
                                               const double* __restrict__ trg, int nt, double* __restrict__ partial,
                                               int chunk, Wavek wk) {
  constexpr bool C = kcplx<KIND>();
  constexpr int NA = GRAD ? (C ? 8 : 4) : (C ? 2 : 1);
  constexpr int vw = C ? 2 : 1;
  __shared__ double sx[TILE], sy[TILE], sz[TILE];
  __shared__ R sq[TILE], sqi[TILE];
  const int lane = threadIdx.x;
  const int ti = blockIdx.x * TILE + lane;
  const bool act = ti < nt;
  double tx = 0, ty = 0, tz = 0;
  if (act) { tx = trg[3 * size_t(ti)]; ty = trg[3 * size_t(ti) + 1]; tz = trg[3 * size_t(ti) + 2]; }
  const int sbeg = blockIdx.y * chunk;
  const int send = min(ns, sbeg + chunk);
  const R kr = R(wk.re), ki = R(wk.im);
  R acc[NA], comp[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) { acc[a] = 0; comp[a] = 0; }
  for (int s0 = sbeg; s0 < send; s0 += TILE) {
    const int m = min(TILE, send - s0);
    __syncthreads();
    if (lane < m) {
      const size_t si = size_t(s0 + lane);
      sx[lane] = src[3 * si];
      sy[lane] = src[3 * si + 1];
      sz[lane] = src[3 * si + 2];
      sq[lane] = R(q[si * vw]);
      if (C) sqi[lane] = R(q[si * vw + 1]);
    }
    __syncthreads();
    R part[NA];
#pragma unroll
    for (int a = 0; a < NA; a++) part[a] = 0;
    for (int j = 0; j < m; j++) {
      const R dx = R(sx[j] - tx), dy = R(sy[j] - ty), dz = R(sz[j] - tz);
      if (GRAD) pair_grad<KIND, R, RSQ_LIB>(dx, dy, dz, sq[j], C ? sqi[j] : R(0), kr, ki, part);
      else pair_pot<KIND, R, RSQ_LIB>(dx, dy, dz, sq[j], C ? sqi[j] : R(0), kr, ki, part);
    }
#pragma unroll
    for (int a = 0; a < NA; a++) kahan_add(acc[a], comp[a], part[a]);
  }
  if (act) {
    double* o = partial + (size_t(blockIdx.y) * nt + ti) * NA;
#pragma unroll
    for (int a = 0; a < NA; a++) o[a] = double(acc[a]) - double(comp[a]);
  }
}

__global__ void k_direct_reduce(const double* __restrict__ partial, int nsplit, int nt, int na,
                                double* __restrict__ out) {
  const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= size_t(nt) * na) return;
  double s = 0, c = 0;
  for (int k = 0; k < nsplit; k++) kahan_add(s, c, partial[size_t(k) * nt * na + i]);
  out[i] += (s - c) * FOUR_PI_INV;
}

// ------------------------------------------------- near-coincident pairs
//
// The FP32 near field forms a pair's difference vector from coordinates
// relative to leaf centres, each rounded to FP32: an absolute error of about
// 2^-24 of the leaf size per coordinate, so a relative error of about
// 2^-24 L / r in 1/r (and twice that in the gradient) for a pair at
// distance r.  Only pairs much closer than the leaf size are hurt, and they
// are rare: they are listed once per tree (k_close_find, FP64 distances) and
// after every FP32 near field k_close_fix replaces each one's FP32
// contribution, recomputed from the very operands the near-field kernel used,
// by its FP64 value.
__global__ __launch_bounds__(64) void k_close_find(DevTree t, const int* __restrict__ work, int* __restrict__ out,
                                                   int cap, int* __restrict__ count) {
  const int leaf = work[blockIdx.x];
  const int tb = t.trg_begin[leaf], nt = t.trg_count[leaf];
  const double side = 2.0 * t.r0 * ldexp(1.0, -t.level[leaf]);
  // (side / 16)^2 with a margin, so that every pair the FP32 near field
  // leaves out (FP32 r^2 <= (side / 16)^2, close_rmin2) is listed
  const double d2 = side * side * (1.002 / 256.0);
  const int g0 = t.p2p_off[leaf], g1 = t.p2p_off[leaf + 1];
  for (int i = threadIdx.x; i < nt; i += blockDim.x) {
    const double* x = t.trg + 3 * size_t(tb + i);
    const double x0 = x[0], x1 = x[1], x2 = x[2];
    for (int g = g0; g < g1; g++) {
      const int sn = t.p2p_src[g], sb = t.src_begin[sn], ns = t.src_count[sn];
      for (int j = 0; j < ns; j++) {
        const double* y = t.src + 3 * size_t(sb + j);
        const double dx = y[0] - x0, dy = y[1] - x1, dz = y[2] - x2;
        const double r2 = dx * dx + dy * dy + dz * dz;
        if (r2 < d2 && r2 > 0.0) {
          const int k = atomicAdd(count, 1);
          if (k < cap) {
            int* o = out + 4 * size_t(k);
            o[0] = tb + i;
            o[1] = sb + j;
            o[2] = leaf;
            o[3] = sn;
          }
        }
      }
    }
  }
}

void launch_close_find(const DevTree& t, const int* work, int nwork, int* out, int cap, int* count, hipStream_t s) {
  if (nwork <= 0) return;
  hipLaunchKernelGGL(k_close_find, dim3(nwork), dim3(64), 0, s, t, work, out, cap, count);
}

// The correction of each listed pair per unit charge, once per tree: one
// thread per pair, coef[k][a] = FP64 minus FP32 value of the pair with q = 1
// (real kernels: a = potential, gradient; Helmholtz: their (re, im) parts).
// The FP32 operands are formed as the near-field kernel forms them: PACKED
// (real kernels, k_p2p_smem) the source relative to its leaf centre as
// k_pack_sources rounds it and the target moved into the source leaf's frame,
// else (k_eval_targets) both relative to the target leaf centre.  A pair is
// linear in its charge, so per evaluation its correction is q coef (complex
// product for Helmholtz); the FP32 value of a charged pair differs from q
// times that of a unit charge by the rounding of the products, ~1e-7 of the
// pair's value.
template <int KIND, bool PACKED>
__global__ __launch_bounds__(256) void k_close_coef(DevTree t, Wavek wk, int npairs, double* __restrict__ coef) {
  constexpr bool C = kcplx<KIND>();
  constexpr int NA = C ? 8 : 4;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= npairs) return;
  const int* cp = t.close + 4 * size_t(k);
  const int it = cp[0], is = cp[1], tl = cp[2], sl = cp[3];
  const double cx = t.center[3 * tl], cy = t.center[3 * tl + 1], cz = t.center[3 * tl + 2];
  const double* x = t.trg + 3 * size_t(it);
  const double* y = t.src + 3 * size_t(is);
  const float tx = float(x[0] - cx), ty = float(x[1] - cy), tz = float(x[2] - cz);
  const float kr = float(wk.re), ki = float(wk.im);
  float a32[NA];
  double a64[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) {
    a32[a] = 0.f;
    a64[a] = 0.0;
  }
  float dx, dy, dz;
  if constexpr (PACKED) {
    const double* rel = t.src_rel + 3 * size_t(is);
    dx = float(rel[0]) - (tx - float(t.center[3 * sl] - cx));
    dy = float(rel[1]) - (ty - float(t.center[3 * sl + 1] - cy));
    dz = float(rel[2]) - (tz - float(t.center[3 * sl + 2] - cz));
  } else {
    dx = float(y[0] - cx) - tx;
    dy = float(y[1] - cy) - ty;
    dz = float(y[2] - cz) - tz;
  }
  // a leaf with itself: the near-field kernel left out the pairs with FP32
  // r^2 <= rmin2 (their FP32 value is 0 here too)
  if (tl == sl) pair_grad<KIND, float, RSQ_MASK>(dx, dy, dz, 1.f, 0.f, kr, ki, a32, close_rmin2<float>(t, tl));
  else pair_grad<KIND, float, RSQ_SEP>(dx, dy, dz, 1.f, 0.f, kr, ki, a32);
  pair_grad<KIND, double, RSQ_MASK>(y[0] - x[0], y[1] - x[1], y[2] - x[2], 1.0, 0.0, wk.re, wk.im, a64);
#pragma unroll
  for (int a = 0; a < NA; a++) coef[size_t(k) * NA + a] = a64[a] - double(a32[a]);
}

void launch_close_coef(int kind, const DevTree& t, Wavek k, int npairs, double* coef, hipStream_t s) {
  if (npairs <= 0) return;
  const dim3 g((npairs + 255) / 256), b(256);
  if (kind == LAPLACE) hipLaunchKernelGGL((k_close_coef<LAPLACE, true>), g, b, 0, s, t, k, npairs, coef);
  else if (kind == YUKAWA) hipLaunchKernelGGL((k_close_coef<YUKAWA, true>), g, b, 0, s, t, k, npairs, coef);
  else hipLaunchKernelGGL((k_close_coef<HELMHOLTZ, false>), g, b, 0, s, t, k, npairs, coef);
}

// One thread per target with listed pairs, its pairs in list order (so the
// result does not depend on scheduling): trg_value += sum of q_s coef / 4 pi.
template <int KIND>
__global__ __launch_bounds__(256) void k_close_fix(DevTree t) {
  constexpr bool C = kcplx<KIND>();
  constexpr int NA = C ? 8 : 4;
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= t.nclose) return;
  double fix[NA];
#pragma unroll
  for (int a = 0; a < NA; a++) fix[a] = 0.0;
  const int k0 = t.close_off[u], k1 = t.close_off[u + 1];
  for (int k = k0; k < k1; k++) {
    const int is = t.close[4 * size_t(k) + 1];
    const double* cf = t.close_coef + size_t(k) * NA;
    const double qr = t.q[size_t(is) * t.vw];
    if constexpr (C) {
      const double qi = t.q[size_t(is) * t.vw + 1];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        fix[2 * c] += qr * cf[2 * c] - qi * cf[2 * c + 1];
        fix[2 * c + 1] += qr * cf[2 * c + 1] + qi * cf[2 * c];
      }
    } else {
#pragma unroll
      for (int a = 0; a < NA; a++) fix[a] += qr * cf[a];
    }
  }
  double* o = t.trg_value + size_t(t.close[4 * size_t(k0)]) * 4 * t.vw;
#pragma unroll
  for (int a = 0; a < NA; a++) o[a] += fix[a] * FOUR_PI_INV;
}

// ---------------------------------------------------------------- launchers

template <int KIND, typename R>
static void eval_targets_t(int mode, const DevTree& t, Wavek k, const int* work, int nwork, const int* tiles,
                           int ntiles, hipStream_t s) {
  if (nwork <= 0) return;
  const bool real = !kcplx<KIND>() && tiles && ntiles > 0;
  if (mode == MODE_P2P && real && t.src_pack && t.src_rel) {
    if (t.nsrcs > 0)
      hipLaunchKernelGGL(k_pack_sources<R>, dim3((t.nsrcs + 255) / 256), dim3(256), 0, s, t.nsrcs, t.src_rel, t.q,
                         static_cast<typename P4<R>::t*>(t.src_pack));
    hipLaunchKernelGGL((k_p2p_smem<KIND, R>), dim3(ntiles), dim3(64), 0, s, t, k, tiles);
  } else if (mode == MODE_P2P) hipLaunchKernelGGL((k_eval_targets<KIND, R, MODE_P2P>), dim3(nwork), dim3(64), 0, s, t, k, work);
  else if (mode == MODE_M2P) hipLaunchKernelGGL((k_eval_targets<KIND, R, MODE_M2P>), dim3(nwork), dim3(64), 0, s, t, k, work);
  else if (real && t.surf_tab) hipLaunchKernelGGL((k_l2p_smem<KIND, R>), dim3(ntiles), dim3(64), 0, s, t, k, tiles);
  else hipLaunchKernelGGL((k_eval_targets<KIND, R, MODE_L2P>), dim3(nwork), dim3(64), 0, s, t, k, work);
  // FP32 near field: the near-coincident pairs in FP64 (same stream, after it)
  if constexpr (sizeof(R) == 4) {
    if (mode == MODE_P2P && t.nclose > 0 && t.close && t.close_off && t.close_coef)
      hipLaunchKernelGGL(k_close_fix<KIND == HELMHOLTZ_REAL ? int(HELMHOLTZ) : KIND>, dim3((t.nclose + 255) / 256),
                         dim3(256), 0, s, t);
  }
}

void launch_eval_targets(int kind, bool f32, int mode, const DevTree& t, Wavek k, const int* work, int nwork,
                         const int* tiles, int ntiles, hipStream_t s) {
  switch (kind) {
    case LAPLACE: f32 ? eval_targets_t<LAPLACE, float>(mode, t, k, work, nwork, tiles, ntiles, s)
                      : eval_targets_t<LAPLACE, double>(mode, t, k, work, nwork, tiles, ntiles, s); break;
    case YUKAWA: f32 ? eval_targets_t<YUKAWA, float>(mode, t, k, work, nwork, tiles, ntiles, s)
                     : eval_targets_t<YUKAWA, double>(mode, t, k, work, nwork, tiles, ntiles, s); break;
    default:
      if (k.im == 0.0) f32 ? eval_targets_t<HELMHOLTZ_REAL, float>(mode, t, k, work, nwork, tiles, ntiles, s)
                           : eval_targets_t<HELMHOLTZ_REAL, double>(mode, t, k, work, nwork, tiles, ntiles, s);
      else f32 ? eval_targets_t<HELMHOLTZ, float>(mode, t, k, work, nwork, tiles, ntiles, s)
               : eval_targets_t<HELMHOLTZ, double>(mode, t, k, work, nwork, tiles, ntiles, s);
      break;
  }
}

template <int KIND, typename R>
static void eval_surface_t(int mode, const DevTree& t, Wavek k, const int* work, int nwork, double* check,
                           hipStream_t s) {
  if (nwork <= 0) return;
  if (mode == MODE_P2M) hipLaunchKernelGGL((k_eval_surface<KIND, R, MODE_P2M>), dim3(nwork), dim3(64), 0, s, t, k, work, check);
  else hipLaunchKernelGGL((k_eval_surface<KIND, R, MODE_P2L>), dim3(nwork), dim3(64), 0, s, t, k, work, check);
}

void launch_eval_surface(int kind, bool f32, int mode, const DevTree& t, Wavek k, const int* work, int nwork,
                         double* check, hipStream_t s) {
  switch (kind) {
    case LAPLACE: f32 ? eval_surface_t<LAPLACE, float>(mode, t, k, work, nwork, check, s)
                      : eval_surface_t<LAPLACE, double>(mode, t, k, work, nwork, check, s); break;
    case YUKAWA: f32 ? eval_surface_t<YUKAWA, float>(mode, t, k, work, nwork, check, s)
                     : eval_surface_t<YUKAWA, double>(mode, t, k, work, nwork, check, s); break;
    default:
      if (k.im == 0.0) f32 ? eval_surface_t<HELMHOLTZ_REAL, float>(mode, t, k, work, nwork, check, s)
                           : eval_surface_t<HELMHOLTZ_REAL, double>(mode, t, k, work, nwork, check, s);
      else f32 ? eval_surface_t<HELMHOLTZ, float>(mode, t, k, work, nwork, check, s)
               : eval_surface_t<HELMHOLTZ, double>(mode, t, k, work, nwork, check, s);
      break;
  }
}

template <int KIND, typename R>
static void direct_t(Wavek k, const double* src, const double* q, int ns, const double* trg, int nt, double* out,
                     int gradient, double* scratch, int nsplit, hipStream_t s) {
  const int na = (gradient ? 4 : 1) * (KIND == HELMHOLTZ ? 2 : 1);
  int chunk = (ns + nsplit - 1) / nsplit;
  chunk = ((chunk + TILE - 1) / TILE) * TILE;
  if (chunk == 0) chunk = TILE;
  dim3 grid((nt + TILE - 1) / TILE, nsplit);
  if (gradient) hipLaunchKernelGGL((k_direct<KIND, R, true>), grid, dim3(64), 0, s, src, q, ns, trg, nt, scratch, chunk, k);
  else hipLaunchKernelGGL((k_direct<KIND, R, false>), grid, dim3(64), 0, s, src, q, ns, trg, nt, scratch, chunk, k);
  const size_t n = size_t(nt) * na;
  hipLaunchKernelGGL(k_direct_reduce, dim3((n + 255) / 256), dim3(256), 0, s, scratch, nsplit, nt, na, out);
}

void launch_direct(int kind, bool mixed, bool cplx, Wavek k, const double* src, const double* q, int ns,
                   const double* trg, int nt, double* out, int gradient, double* scratch, int nsplit,
                   hipStream_t s) {
  (void)cplx;
  if (nt <= 0) return;
  switch (kind) {
    case LAPLACE: mixed ? direct_t<LAPLACE, float>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s)
                        : direct_t<LAPLACE, double>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s); break;
    case YUKAWA: mixed ? direct_t<YUKAWA, float>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s)
                       : direct_t<YUKAWA, double>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s); break;
    default: mixed ? direct_t<HELMHOLTZ, float>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s)
                   : direct_t<HELMHOLTZ, double>(k, src, q, ns, trg, nt, out, gradient, scratch, nsplit, s); break;
  }
}

}  // namespace xfmm
