/*
 * oracle.c -- CPU restatement of exafmm-t's FMM evaluation phase.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the gfx950 engine:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * it (through ctypes).  The product never links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the reference algorithm of
 *   include/fmm_base.h            P2P (:82-94), M2P (:97-124), P2L (:127-154),
 *                                 upward_pass (:162-171), downward_pass (:179-200)
 *   include/fmm_scale_invariant.h P2M (:163-195), L2P (:198-230), M2M (:233-253),
 *                                 L2L (:256-275), hadamard_product (:343-399),
 *                                 M2L (:407-440), precompute_M2L (:517-558),
 *                                 fft_up_equiv (:560-604), ifft_dn_check (:606-646)
 *   include/fmm.h                 the same operators with per-level matrices
 *                                 (:156-263, :395-446, :524-788)
 *   include/laplace.h, helmholtz.h, modified_helmholtz.h   the scalar P2P loops
 *                                 (laplace.h:139-162, helmholtz.h:186-207,
 *                                 modified_helmholtz.h:149-170)
 *   include/geometry.h            surface (:21-63), convolution_grid (:75-92),
 *                                 surf2conv (:101-134)
 * over the flattened tree layout of include/exafmm_t_amd.h (same fields).
 * DFTs are plain separable DFTs with the FFTW conventions (forward e^{-i},
 * unnormalised backward, r2c halved along the fastest axis).
 *
 * Pinning (tests/test_oracle.py): P2P against closed forms, long-double brute
 * force and finite differences of the potential; DFTs against numpy.fft; the
 * FFT M2L against a dense kernel-matrix M2L on the actual surfaces; the whole
 * evaluation against an FP64 direct sum.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"


static const double INV4PI = 0.0795774715459476678844418816863;

/* ------------------------------------------------------------- kernels --- */

/* potential of one source at d = x_t - x_s (0 for r = 0) */
static double complex kern(int kind, double complex k, double dx, double dy, double dz) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 == 0) return 0;
  const double r = sqrt(r2);
  if (kind == LAPLACE) return 1.0 / r;
  if (kind == YUKAWA) return exp(-creal(k) * r) / r;
  return cexp(I * k * r) / r;
}

/* d(potential)/d(x_t) = g * (x_t - x_s): returns g */
static double complex kern_grad(int kind, double complex k, double dx, double dy, double dz) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  if (r2 == 0) return 0;
  const double r = sqrt(r2);
  if (kind == LAPLACE) return -1.0 / (r2 * r);                                   /* laplace.h:158-161 */
  if (kind == YUKAWA) return -exp(-creal(k) * r) * (creal(k) * r + 1) / (r2 * r); /* modified_helmholtz.h:159-164 */
  return (I * k / r - 1.0 / r2) * cexp(I * k * r) / r;                           /* helmholtz.h:199-203 */
}

/* values are T = double (vw=1) or double complex stored as (re,im) (vw=2) */
static double complex getv(const double* v, size_t i, int vw) {
  return vw == 2 ? v[2 * i] + I * v[2 * i + 1] : v[i];
}
static void addv(double* v, size_t i, int vw, double complex x) {
  if (vw == 2) { v[2 * i] += creal(x); v[2 * i + 1] += cimag(x); }
  else v[i] += creal(x);
}
static void setv(double* v, size_t i, int vw, double complex x) {
  if (vw == 2) { v[2 * i] = creal(x); v[2 * i + 1] = cimag(x); }
  else v[i] = creal(x);
}

/* trg_value += sum_s q_s K(x_t, x_s) [and gradient], textbook loop with
 * long-double accumulation.  Exported: the FP64 direct-sum checker. */
void orc_direct(int kind, double kre, double kim, const double* src, const double* q, int ns, const double* trg,
                int nt, double* out, int gradient) {
  const double complex k = kre + I * kim;
  const int vw = kind == HELMHOLTZ ? 2 : 1;
  const int w = gradient ? 4 : 1;
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; t++) {
    long double pr = 0, pi = 0, gr[3] = {0, 0, 0}, gi[3] = {0, 0, 0};
    for (int s = 0; s < ns; s++) {
      const double dx = trg[3 * t] - src[3 * s], dy = trg[3 * t + 1] - src[3 * s + 1], dz = trg[3 * t + 2] - src[3 * s + 2];
      const double complex qs = getv(q, s, vw);
      const double complex phi = kern(kind, k, dx, dy, dz) * qs;
      pr += creal(phi);
      pi += cimag(phi);
      if (gradient) {
        const double complex g = kern_grad(kind, k, dx, dy, dz) * qs;
        gr[0] += creal(g) * dx; gi[0] += cimag(g) * dx;
        gr[1] += creal(g) * dy; gi[1] += cimag(g) * dy;
        gr[2] += creal(g) * dz; gi[2] += cimag(g) * dz;
      }
    }
    addv(out, (size_t)t * w, vw, ((double)pr + I * (double)pi) * INV4PI);
    if (gradient)
      for (int d = 0; d < 3; d++) addv(out, (size_t)t * w + 1 + d, vw, ((double)gr[d] + I * (double)gi[d]) * INV4PI);
  }
}

/* ------------------------------------------------------------ geometry --- */

/* surface points (geometry.h:21-63), absolute coordinates */
void orc_surface(int p, double r0, int level, const double* c, double alpha, double* out) {
  const int n = 6 * (p - 1) * (p - 1) + 2;
  int cnt = 0;
  out[0] = -1.0; out[1] = -1.0; out[2] = -1.0;
  cnt = 1;
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++, cnt++) {
      out[3 * cnt] = -1.0;
      out[3 * cnt + 1] = (2.0 * (i + 1) - p + 1) / (p - 1);
      out[3 * cnt + 2] = (2.0 * j - p + 1) / (p - 1);
    }
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++, cnt++) {
      out[3 * cnt] = (2.0 * i - p + 1) / (p - 1);
      out[3 * cnt + 1] = -1.0;
      out[3 * cnt + 2] = (2.0 * (j + 1) - p + 1) / (p - 1);
    }
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++, cnt++) {
      out[3 * cnt] = (2.0 * (i + 1) - p + 1) / (p - 1);
      out[3 * cnt + 1] = (2.0 * j - p + 1) / (p - 1);
      out[3 * cnt + 2] = -1.0;
    }
  for (int i = 0; i < (n / 2) * 3; i++) out[3 * cnt + i] = -out[i];
  const double b = alpha * r0 * pow(0.5, level);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) out[3 * i + d] = out[3 * i + d] * b + c[d];
}

/* surf2conv_up (dn = 0) / surf2conv_dn (dn = 1) index maps (geometry.h:101-134) */
void orc_surf2conv(int p, int dn, int* map) {
  const int n = 6 * (p - 1) * (p - 1) + 2, n1 = 2 * p;
  double c[3] = {0.5 * (p - 1), 0.5 * (p - 1), 0.5 * (p - 1)};
  double* s = (double*)malloc(sizeof(double) * 3 * n);
  orc_surface(p, 0.5, 0, c, (double)(p - 1), s);
  const int off = dn ? 2 * p - 1 : p - 1;
  for (int i = 0; i < n; i++)
    map[i] = (int)(off - s[3 * i]) + (int)(off - s[3 * i + 1]) * n1 + (int)(off - s[3 * i + 2]) * n1 * n1;
  free(s);
}

/* ----------------------------------------------------------------- DFT --- */

/* in-place 1-D DFT along axis (0 = slowest) of an n^3 complex array */
static void dft_axis(double complex* a, int n, int axis, int sign) {
  double complex* w = (double complex*)malloc(sizeof(double complex) * n);
  for (int m = 0; m < n; m++) w[m] = cexp(sign * 2.0 * M_PI * I * m / n);
  const size_t stride = axis == 0 ? (size_t)n * n : (axis == 1 ? (size_t)n : 1);
  for (int u = 0; u < n; u++)
    for (int v = 0; v < n; v++) {
      size_t base;
      if (axis == 0) base = (size_t)u * n + v;
      else if (axis == 1) base = (size_t)u * n * n + v;
      else base = ((size_t)u * n + v) * n;
      double complex line[64], res[64];
      for (int j = 0; j < n; j++) line[j] = a[base + j * stride];
      for (int k = 0; k < n; k++) {
        double complex acc = 0;
        for (int j = 0; j < n; j++) acc += line[j] * w[((size_t)k * j) % n];
        res[k] = acc;
      }
      for (int k = 0; k < n; k++) a[base + k * stride] = res[k];
    }
  free(w);
}

/* forward r2c: in n^3 real (x fastest), out n*n*(n/2+1) complex (re,im) */
void orc_dft3_r2c(const double* in, int n, double* out) {
  const size_t n3 = (size_t)n * n * n;
  double complex* a = (double complex*)malloc(sizeof(double complex) * n3);
  for (size_t i = 0; i < n3; i++) a[i] = in[i];
  for (int ax = 0; ax < 3; ax++) dft_axis(a, n, ax, -1);
  const int nh = n / 2 + 1;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < nh; k++) {
        const double complex v = a[((size_t)i * n + j) * n + k];
        const size_t o = ((size_t)i * n + j) * nh + k;
        out[2 * o] = creal(v);
        out[2 * o + 1] = cimag(v);
      }
  free(a);
}

/* backward c2r (unnormalised), half spectrum in, n^3 real out: Hermitian
 * extension of the missing half, full backward DFT, real part */
void orc_dft3_c2r(const double* in, int n, double* out) {
  const size_t n3 = (size_t)n * n * n;
  const int nh = n / 2 + 1;
  double complex* a = (double complex*)malloc(sizeof(double complex) * n3);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++)
      for (int k = 0; k < n; k++) {
        double complex v;
        if (k < nh) {
          const size_t o = ((size_t)i * n + j) * nh + k;
          v = in[2 * o] + I * in[2 * o + 1];
        } else {
          const size_t o = ((size_t)((n - i) % n) * n + (n - j) % n) * nh + (n - k);
          v = in[2 * o] - I * in[2 * o + 1];
        }
        a[((size_t)i * n + j) * n + k] = v;
      }
  for (int ax = 0; ax < 3; ax++) dft_axis(a, n, ax, +1);
  for (size_t i = 0; i < n3; i++) out[i] = creal(a[i]);
  free(a);
}

void orc_dft3_c2c(const double* in, int n, int sign, double* out) {
  const size_t n3 = (size_t)n * n * n;
  double complex* a = (double complex*)malloc(sizeof(double complex) * n3);
  for (size_t i = 0; i < n3; i++) a[i] = in[2 * i] + I * in[2 * i + 1];
  for (int ax = 0; ax < 3; ax++) dft_axis(a, n, ax, sign);
  for (size_t i = 0; i < n3; i++) { out[2 * i] = creal(a[i]); out[2 * i + 1] = cimag(a[i]); }
  free(a);
}

/* ------------------------------------------------- M2L operator (pin) --- */

static int hash5(int x, int y, int z) { return ((z + 5) * 10 + (y + 5)) * 10 + (x + 5); }

/* M2L frequency matrices for one level (fmm_scale_invariant.h:517-558 with
 * unit = 2 r0 and child level 0 when scale_invariant, fmm.h:524-629 with
 * unit = r0 2^(1-l), child level l otherwise).  out: [26][nfreq][64][2]. */
void orc_m2l_operator(int kind, double kre, double kim, int p, double r0, int child_level, int scale_invariant,
                      double* out) {
  const int n1 = 2 * p, nconv = n1 * n1 * n1;
  const int nfreq = kind == HELMHOLTZ ? nconv : n1 * n1 * (n1 / 2 + 1);
  const double complex k = kre + I * kim;
  /* helper relative positions: [-3,3]^3 with max |c| >= 2, z outer, x inner */
  int hx[343], hy[343], hz[343], lut[2000];
  for (int i = 0; i < 2000; i++) lut[i] = -1;
  int nh = 0;
  for (int z = -3; z <= 3; z++)
    for (int y = -3; y <= 3; y++)
      for (int x = -3; x <= 3; x++)
        if (abs(x) >= 2 || abs(y) >= 2 || abs(z) >= 2) { hx[nh] = x; hy[nh] = y; hz[nh] = z; lut[hash5(x, y, z)] = nh++; }
  const double unit = scale_invariant ? r0 / 0.5 : r0 * pow(0.5, child_level - 1);
  const double a = 2 * r0 * pow(0.5, child_level) * 1.05;
  double* helper = (double*)malloc(sizeof(double) * 2 * (size_t)nfreq * nh);
#pragma omp parallel for schedule(dynamic)
  for (int h = 0; h < nh; h++) {
    double* val = (double*)malloc(sizeof(double) * 2 * nconv);
    double* re = (double*)malloc(sizeof(double) * nconv);
    const double c[3] = {hx[h] * unit, hy[h] * unit, hz[h] * unit};
    for (int kk = 0; kk < n1; kk++)
      for (int j = 0; j < n1; j++)
        for (int i = 0; i < n1; i++) {
          const size_t idx = (size_t)i + n1 * j + (size_t)n1 * n1 * kk;
          const double gx = (i - p) * a / (p - 1) + c[0], gy = (j - p) * a / (p - 1) + c[1], gz = (kk - p) * a / (p - 1) + c[2];
          const double complex v = kern(kind, k, -gx, -gy, -gz) * INV4PI;
          val[2 * idx] = creal(v);
          val[2 * idx + 1] = cimag(v);
          re[idx] = creal(v);
        }
    if (kind == HELMHOLTZ) orc_dft3_c2c(val, n1, -1, helper + 2 * (size_t)nfreq * h);
    else orc_dft3_r2c(re, n1, helper + 2 * (size_t)nfreq * h);
    free(val);
    free(re);
  }
  memset(out, 0, sizeof(double) * (size_t)NM2L * nfreq * 128);
  int r = 0;
  for (int z = -1; z <= 1; z++)
    for (int y = -1; y <= 1; y++)
      for (int x = -1; x <= 1; x++) {
        if (!x && !y && !z) continue;
        for (int j1 = 0; j1 < 8; j1++)
          for (int j2 = 0; j2 < 8; j2++) {
            const int cx = 2 * x - (j1 & 1) + ((j2 >> 0) & 1);
            const int cy = 2 * y - ((j1 >> 1) & 1) + ((j2 >> 1) & 1);
            const int cz = 2 * z - ((j1 >> 2) & 1) + ((j2 >> 2) & 1);
            const int h = lut[hash5(cx, cy, cz)];
            if (h < 0) continue;
            const int j = j2 * 8 + j1;
            for (int f = 0; f < nfreq; f++) {
              const size_t o = ((size_t)r * nfreq + f) * 128 + 2 * j;
              out[o] = helper[2 * ((size_t)nfreq * h + f)] / nconv;
              out[o + 1] = helper[2 * ((size_t)nfreq * h + f) + 1] / nconv;
            }
          }
        r++;
      }
  free(helper);
}

/* ------------------------------------------------------ evaluation ------ */


static int node_level(const orc_tree* t, int n) {
  int l = 0;
  while (t->level_offset[l + 1] <= n) l++;
  return l;
}

/* y = A x, A row-major ns x ns (math_wrapper.h:51-73) */
static void gemv(int ns, const double* A, int vw, const double complex* x, double complex* y) {
  for (int i = 0; i < ns; i++) {
    double complex acc = 0;
    for (int j = 0; j < ns; j++) acc += getv(A, (size_t)i * ns + j, vw) * x[j];
    y[i] = acc;
  }
}

/* potentials at pts from leaf particles of node s (potential_P2P) */
static void pot_from_leaf(const orc_tree* t, const orc_ops* o, const double* q, int vw, int s, const double* pts,
                          int npts, double complex* acc) {
  const double complex k = o->kre + I * o->kim;
  for (int i = 0; i < npts; i++) {
    double complex a = 0;
    for (int j = 0; j < t->src_count[s]; j++) {
      const int sj = t->src_begin[s] + j;
      a += kern(o->kind, k, pts[3 * i] - t->src_coord[3 * sj], pts[3 * i + 1] - t->src_coord[3 * sj + 1],
                pts[3 * i + 2] - t->src_coord[3 * sj + 2]) * getv(q, sj, vw);
    }
    acc[i] += a * INV4PI;
  }
}

/* gradient_P2P from point sources (coords, values) onto the targets of leaf l */
static void grad_to_leaf(const orc_tree* t, const orc_ops* o, int l, const double* sc, const double complex* sv,
                         int ns, double* trg_value, int vw) {
  const double complex k = o->kre + I * o->kim;
  for (int i = 0; i < t->trg_count[l]; i++) {
    const int ti = t->trg_begin[l] + i;
    double complex ph = 0, g[3] = {0, 0, 0};
    for (int j = 0; j < ns; j++) {
      const double dx = t->trg_coord[3 * ti] - sc[3 * j], dy = t->trg_coord[3 * ti + 1] - sc[3 * j + 1],
                   dz = t->trg_coord[3 * ti + 2] - sc[3 * j + 2];
      ph += kern(o->kind, k, dx, dy, dz) * sv[j];
      const double complex gg = kern_grad(o->kind, k, dx, dy, dz) * sv[j];
      g[0] += gg * dx; g[1] += gg * dy; g[2] += gg * dz;
    }
    addv(trg_value, (size_t)ti * 4, vw, ph * INV4PI);
    for (int d = 0; d < 3; d++) addv(trg_value, (size_t)ti * 4 + 1 + d, vw, g[d] * INV4PI);
  }
}

/* Runs the selected stages in the reference order, accumulating into up, dn
 * (T[nnodes][nsurf]) and trg_value (T[ntrgs][4]); q is T[nsrcs]. */
void orc_evaluate(const orc_tree* t, const orc_ops* o, const double* q, int stages, double* up, double* dn,
                  double* trg_value) {
  const int p = o->p, ns = 6 * (p - 1) * (p - 1) + 2, nn = t->nnodes;
  const int vw = o->kind == HELMHOLTZ ? 2 : 1;
  const size_t ns2 = (size_t)ns * ns;
  const int lap = o->kind == LAPLACE;
  int* lev = (int*)malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++) lev[i] = node_level(t, i);
#define OPL(l) (o->nlevels == 1 ? 0 : (l))

  if (stages & ST_P2M) { /* fmm_scale_invariant.h:163-195 */
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0 || (t->src_count[l] == 0 && t->trg_count[l] == 0)) continue;
      double* pts = (double*)malloc(sizeof(double) * 3 * ns);
      double complex *chk = (double complex*)calloc(ns, sizeof(double complex)), *tmp = (double complex*)malloc(sizeof(double complex) * ns),
                     *eq = (double complex*)malloc(sizeof(double complex) * ns);
      orc_surface(p, t->r0, lev[l], t->center + 3 * l, 2.95, pts);
      pot_from_leaf(t, o, q, vw, l, pts, ns, chk);
      const int L = OPL(lev[l]);
      gemv(ns, o->uc2e_u + L * ns2 * vw, vw, chk, tmp);
      gemv(ns, o->uc2e_v + L * ns2 * vw, vw, tmp, eq);
      const double sc = lap ? pow(0.5, lev[l]) : 1.0;
      for (int i = 0; i < ns; i++) setv(up, (size_t)l * ns + i, vw, sc * eq[i]);
      free(pts); free(chk); free(tmp); free(eq);
    }
  }
  if (stages & ST_M2M) { /* post-order; fmm_scale_invariant.h:233-253 */
    for (int L0 = t->depth - 1; L0 >= 0; L0--) {
#pragma omp parallel for schedule(dynamic)
      for (int n = t->level_offset[L0]; n < t->level_offset[L0 + 1]; n++) {
        if (t->child0[n] < 0) continue;
        double complex x[2048], y[2048];
        for (int c = 0; c < 8; c++) {
          const int ch = t->child0[n] + c;
          for (int i = 0; i < ns; i++) x[i] = getv(up, (size_t)ch * ns + i, vw);
          gemv(ns, o->m2m + ((size_t)OPL(L0) * 8 + c) * ns2 * vw, vw, x, y);
          for (int i = 0; i < ns; i++) addv(up, (size_t)n * ns + i, vw, y[i]);
        }
      }
    }
  }
  if (stages & ST_P2L) { /* fmm_base.h:127-154 */
#pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < nn; n++) {
      if (t->p2l_offset[n + 1] == t->p2l_offset[n]) continue;
      double* pts = (double*)malloc(sizeof(double) * 3 * ns);
      double complex* acc = (double complex*)calloc(ns, sizeof(double complex));
      orc_surface(p, t->r0, lev[n], t->center + 3 * n, 1.05, pts);
      for (int j = t->p2l_offset[n]; j < t->p2l_offset[n + 1]; j++) pot_from_leaf(t, o, q, vw, t->p2l_source[j], pts, ns, acc);
      for (int i = 0; i < ns; i++) addv(dn, (size_t)n * ns + i, vw, acc[i]);
      free(pts); free(acc);
    }
  }
  if (stages & ST_M2P) { /* fmm_base.h:97-124 */
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0) continue;
      double* pts = (double*)malloc(sizeof(double) * 3 * ns);
      double complex* sv = (double complex*)malloc(sizeof(double complex) * ns);
      for (int j = t->m2p_offset[l]; j < t->m2p_offset[l + 1]; j++) {
        const int s = t->m2p_source[j];
        orc_surface(p, t->r0, lev[s], t->center + 3 * s, 1.05, pts);
        for (int i = 0; i < ns; i++) sv[i] = getv(up, (size_t)s * ns + i, vw);
        grad_to_leaf(t, o, l, pts, sv, ns, trg_value, vw);
      }
      free(pts); free(sv);
    }
  }
  if (stages & ST_P2P) { /* fmm_base.h:82-94 */
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0) continue;
      for (int j = t->p2p_offset[l]; j < t->p2p_offset[l + 1]; j++) {
        const int s = t->p2p_source[j];
        const int m = t->src_count[s];
        double complex* sv = (double complex*)malloc(sizeof(double complex) * (m + 1));
        for (int i = 0; i < m; i++) sv[i] = getv(q, t->src_begin[s] + i, vw);
        grad_to_leaf(t, o, l, t->src_coord + 3 * (size_t)t->src_begin[s], sv, m, trg_value, vw);
        free(sv);
      }
    }
  }
  if (stages & ST_M2L) { /* fmm_scale_invariant.h:407-440, fmm.h:395-446 */
    const int n1 = 2 * p, nconv = n1 * n1 * n1;
    const int nfreq = vw == 2 ? nconv : n1 * n1 * (n1 / 2 + 1);
    int* mup = (int*)malloc(sizeof(int) * ns);
    int* mdn = (int*)malloc(sizeof(int) * ns);
    orc_surf2conv(p, 0, mup);
    orc_surf2conv(p, 1, mdn);
    /* fft_up_equiv for every M2L source */
    int* slot = (int*)malloc(sizeof(int) * nn);
    for (int i = 0; i < nn; i++) slot[i] = -1;
    int nsrc = 0;
    for (int n = 0; n < nn; n++)
      for (int r = 0; r < NM2L; r++) {
        const int s = t->m2l_source[(size_t)n * NM2L + r];
        if (s >= 0 && slot[s] < 0) slot[s] = 0;
      }
    for (int n = 0; n < nn; n++)
      if (slot[n] >= 0) slot[n] = nsrc++;
    double* fin = (double*)malloc(sizeof(double) * 2 * (size_t)nfreq * 8 * (nsrc + 1));
#pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < nn; n++) {
      if (slot[n] < 0) continue;
      double* grid = (double*)malloc(sizeof(double) * 2 * nconv);
      double* spec = (double*)malloc(sizeof(double) * 2 * nfreq);
      for (int c = 0; c < 8; c++) {
        const int ch = t->child0[n] + c;
        memset(grid, 0, sizeof(double) * 2 * nconv);
        for (int i = 0; i < ns; i++) {
          if (vw == 2) { grid[2 * mup[i]] = up[2 * ((size_t)ch * ns + i)]; grid[2 * mup[i] + 1] = up[2 * ((size_t)ch * ns + i) + 1]; }
          else grid[mup[i]] = up[(size_t)ch * ns + i];
        }
        if (vw == 2) orc_dft3_c2c(grid, n1, -1, spec);
        else orc_dft3_r2c(grid, n1, spec);
        for (int f = 0; f < nfreq; f++) {
          const size_t d = (((size_t)slot[n] * nfreq + f) * 8 + c) * 2;
          fin[d] = spec[2 * f];
          fin[d + 1] = spec[2 * f + 1];
        }
      }
      free(grid); free(spec);
    }
    /* hadamard_product + ifft_dn_check per target */
#pragma omp parallel for schedule(dynamic)
    for (int n = 0; n < nn; n++) {
      int any = 0;
      for (int r = 0; r < NM2L; r++) any |= t->m2l_source[(size_t)n * NM2L + r] >= 0;
      if (!any) continue;
      const double* M = o->m2l + (size_t)(o->nm2l == 1 ? 0 : lev[n]) * NM2L * nfreq * 128;
      double* fout = (double*)calloc(2 * (size_t)nfreq * 8, sizeof(double));
      for (int r = 0; r < NM2L; r++) {
        const int s = t->m2l_source[(size_t)n * NM2L + r];
        if (s < 0) continue;
        for (int f = 0; f < nfreq; f++) {
          const double* Mk = M + ((size_t)r * nfreq + f) * 128;
          const double* in = fin + ((size_t)slot[s] * nfreq + f) * 16;
          for (int j1 = 0; j1 < 8; j1++) {
            double complex acc = 0;
            for (int j2 = 0; j2 < 8; j2++)
              acc += (Mk[2 * (j2 * 8 + j1)] + I * Mk[2 * (j2 * 8 + j1) + 1]) * (in[2 * j2] + I * in[2 * j2 + 1]);
            fout[(f * 8 + j1) * 2] += creal(acc);
            fout[(f * 8 + j1) * 2 + 1] += cimag(acc);
          }
        }
      }
      double* spec = (double*)malloc(sizeof(double) * 2 * nfreq);
      double* grid = (double*)malloc(sizeof(double) * 2 * nconv);
      const double sc = lap ? pow(2.0, lev[n] + 1) : 1.0;
      for (int c = 0; c < 8; c++) {
        for (int f = 0; f < nfreq; f++) { spec[2 * f] = fout[(f * 8 + c) * 2]; spec[2 * f + 1] = fout[(f * 8 + c) * 2 + 1]; }
        const int ch = t->child0[n] + c;
        if (vw == 2) {
          orc_dft3_c2c(spec, n1, +1, grid);
          for (int i = 0; i < ns; i++) addv(dn, (size_t)ch * ns + i, vw, sc * (grid[2 * mdn[i]] + I * grid[2 * mdn[i] + 1]));
        } else {
          orc_dft3_c2r(spec, n1, grid);
          for (int i = 0; i < ns; i++) dn[(size_t)ch * ns + i] += sc * grid[mdn[i]];
        }
      }
      free(spec); free(grid); free(fout);
    }
    free(fin); free(slot); free(mup); free(mdn);
  }
  if (stages & ST_L2L) { /* pre-order; fmm_scale_invariant.h:256-275 */
    for (int L0 = 0; L0 < t->depth; L0++) {
#pragma omp parallel for schedule(dynamic)
      for (int n = t->level_offset[L0]; n < t->level_offset[L0 + 1]; n++) {
        if (t->child0[n] < 0) continue;
        double complex x[2048], y[2048];
        for (int i = 0; i < ns; i++) x[i] = getv(dn, (size_t)n * ns + i, vw);
        for (int c = 0; c < 8; c++) {
          const int ch = t->child0[n] + c;
          gemv(ns, o->l2l + ((size_t)OPL(L0) * 8 + c) * ns2 * vw, vw, x, y);
          for (int i = 0; i < ns; i++) addv(dn, (size_t)ch * ns + i, vw, y[i]);
        }
      }
    }
  }
  if (stages & ST_L2P) { /* fmm_scale_invariant.h:198-230 */
#pragma omp parallel for schedule(dynamic)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0 || (t->src_count[l] == 0 && t->trg_count[l] == 0)) continue;
      double* pts = (double*)malloc(sizeof(double) * 3 * ns);
      double complex *x = (double complex*)malloc(sizeof(double complex) * ns), *tmp = (double complex*)malloc(sizeof(double complex) * ns),
                     *eq = (double complex*)malloc(sizeof(double complex) * ns);
      for (int i = 0; i < ns; i++) x[i] = getv(dn, (size_t)l * ns + i, vw);
      const int L = OPL(lev[l]);
      gemv(ns, o->dc2e_u + L * ns2 * vw, vw, x, tmp);
      gemv(ns, o->dc2e_v + L * ns2 * vw, vw, tmp, eq);
      const double sc = lap ? pow(0.5, lev[l]) : 1.0;
      for (int i = 0; i < ns; i++) { eq[i] *= sc; setv(dn, (size_t)l * ns + i, vw, eq[i]); }
      orc_surface(p, t->r0, lev[l], t->center + 3 * l, 2.95, pts);
      grad_to_leaf(t, o, l, pts, eq, ns, trg_value, vw);
      free(pts); free(x); free(tmp); free(eq);
    }
  }
  free(lev);
#undef OPL
}
