/*
 * fast.c -- FFT-based CPU evaluation of the Laplace FMM (the CPU baseline).
 *
 * TEST INFRASTRUCTURE ONLY: bench.py's cpu_baseline leg times it and
 * tests/test_oracle.py pins it against orc_evaluate (oracle.c) to 1e-12; the
 * product never links it.
 *
 * Same algorithm and operators as orc_evaluate, which restates the reference
 * evaluation (include/fmm_base.h:162-200 and the operators it calls), but at
 * the cost of the reference's own CPU implementation rather than the plain
 * restatement's:
 *  - P2P and the particle/surface evaluations vectorised over targets
 *    (laplace.h:91-164 vectorises them over 8 targets with AVX-512);
 *  - P2M / L2P pseudo-inverse products, M2M and L2L as cache-blocked GEMMs
 *    over all leaves / non-leaves of a level (the reference runs one BLAS
 *    gemv per node, fmm_scale_invariant.h:163-275);
 *  - fft_up_equiv / ifft_dn_check with a mixed-radix FFT of length n1 = 2p
 *    written here (the reference calls FFTW r2c / c2r plans,
 *    fmm_scale_invariant.h:560-646), restricted to the lines that hold data
 *    (input on [0,p)^3, output needed on [p,2p)^3);
 *  - hadamard_product blocked over (64 targets x 32 frequencies) so the
 *    operator and source-spectrum chunks stay in cache
 *    (fmm_scale_invariant.h:343-399 blocks by 64 targets the same way).
 * Laplace only (the benchmark workload), FP64 throughout, OpenMP.
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <omp.h>
#include <string.h>

#include "oracle.h"

static const double INV4PI_F = 0.0795774715459476678844418816863;

typedef double v4d __attribute__((vector_size(32)));

/* ------------------------------------------------------------------ FFT --- */

/* Mixed-radix FFT of length n <= 64 (factors 4, 2, 3, 5), decimation in
 * time, applied to a batch of L independent lines at once: line element j of
 * batch member b sits at [j * ld + b] of separate real / imaginary planes, so
 * every butterfly is a loop over b that vectorises. */
typedef struct {
  int n, nf, f[8];
  double twr[64], twi[64]; /* exp(sign 2 pi i m / n) */
} fft_plan;

static void plan_init(fft_plan* pl, int n, int sign) {
  pl->n = n;
  pl->nf = 0;
  int m = n;
  while (m > 1) {
    int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : m;
    pl->f[pl->nf++] = r;
    m /= r;
  }
  for (int k = 0; k < n; k++) {
    pl->twr[k] = cos(2.0 * M_PI * k / n);
    pl->twi[k] = sign * sin(2.0 * M_PI * k / n);
  }
}

/* out rows [0, n) = DFT over the input rows in[0], in[s], ..., in[(n-1) s]
 * (sub-length n of the plan's length N, twiddle stride N / n); t: scratch of
 * 8 L doubles per plane */
static void fft_rec(const fft_plan* pl, const double* ir, const double* ii, int s, double* orr, double* oi, int n,
                    int fi, int tws, int L, int ld, double* tr, double* ti) {
  if (n == 1) {
    memcpy(orr, ir, sizeof(double) * L);
    memcpy(oi, ii, sizeof(double) * L);
    return;
  }
  const int r = pl->f[fi], m = n / r, N = pl->n;
  for (int q = 0; q < r; q++)
    fft_rec(pl, ir + (size_t)q * s * ld, ii + (size_t)q * s * ld, s * r, orr + (size_t)q * m * ld,
            oi + (size_t)q * m * ld, m, fi + 1, tws * r, L, ld, tr, ti);
  for (int k = 0; k < m; k++) {
    for (int q = 0; q < r; q++) {
      const int e = (q * k * tws) % N;
      const double wr = pl->twr[e], wi = pl->twi[e];
      const double* xr = orr + (size_t)(q * m + k) * ld;
      const double* xi = oi + (size_t)(q * m + k) * ld;
      double* yr = tr + (size_t)q * L;
      double* yi = ti + (size_t)q * L;
#pragma omp simd
      for (int b = 0; b < L; b++) {
        yr[b] = wr * xr[b] - wi * xi[b];
        yi[b] = wr * xi[b] + wi * xr[b];
      }
    }
    for (int j = 0; j < r; j++) {
      double* zr = orr + (size_t)(k + j * m) * ld;
      double* zi = oi + (size_t)(k + j * m) * ld;
#pragma omp simd
      for (int b = 0; b < L; b++) {
        zr[b] = tr[b];
        zi[b] = ti[b];
      }
      for (int q = 1; q < r; q++) {
        const int e = ((q * j) % r) * (N / r);
        const double wr = pl->twr[e], wi = pl->twi[e];
        const double* yr = tr + (size_t)q * L;
        const double* yi = ti + (size_t)q * L;
#pragma omp simd
        for (int b = 0; b < L; b++) {
          zr[b] += wr * yr[b] - wi * yi[b];
          zi[b] += wr * yi[b] + wi * yr[b];
        }
      }
    }
  }
}

/* batch transform: out row k = sum_j in row j * exp(sign 2 pi i j k / n) */
static void fft_rows(const fft_plan* pl, const double* ir, const double* ii, double* orr, double* oi, int L, int ld,
                     double* tr, double* ti) {
  fft_rec(pl, ir, ii, 1, orr, oi, pl->n, 0, 1, L, ld, tr, ti);
}

/* exported for the pin against numpy.fft: in/out (re, im) interleaved, n <= 64 */
void orc_fft(int n, int sign, const double* in, double* out) {
  fft_plan pl;
  plan_init(&pl, n, sign);
  double ir[64], ii[64], orr[64], oi[64], tr[8], ti[8];
  for (int j = 0; j < n; j++) {
    ir[j] = in[2 * j];
    ii[j] = in[2 * j + 1];
  }
  fft_rows(&pl, ir, ii, orr, oi, 1, 1, tr, ti);
  for (int k = 0; k < n; k++) {
    out[2 * k] = orr[k];
    out[2 * k + 1] = oi[k];
  }
}

/* ---------------------------------------------------------------- GEMM --- */

/* C[crow[i] + j] (beta ? += : =) alpha[i] * sum_k A[arow[i] + k] B[k n + j]
 * for i < m, j < n (B row-major k x n).  4 x 8 register tiles, rows in
 * blocks of 4 over the threads. */
static void gemm_rows(int m, int n, int k, const double* A, const long* arow, const double* B, double* C,
                      const long* crow, const double* alpha, int beta) {
#pragma omp parallel for schedule(dynamic, 4)
  for (int i0 = 0; i0 < m; i0 += 4) {
    const int mr = m - i0 < 4 ? m - i0 : 4;
    const double* a[4];
    double* c[4];
    double al[4];
    for (int r = 0; r < 4; r++) {
      const int i = i0 + (r < mr ? r : 0);
      a[r] = A + arow[i];
      c[r] = C + crow[i];
      al[r] = alpha ? alpha[i] : 1.0;
    }
    const int n8 = n & ~7;
    for (int j0 = 0; j0 < n8; j0 += 8) {
      v4d acc[4][2];
      for (int r = 0; r < 4; r++) acc[r][0] = acc[r][1] = (v4d){0, 0, 0, 0};
      const double* b = B + j0;
      for (int kk = 0; kk < k; kk++, b += n) {
        v4d b0, b1;
        memcpy(&b0, b, 32);
        memcpy(&b1, b + 4, 32);
        for (int r = 0; r < 4; r++) {
          const double av = a[r][kk];
          const v4d ar = {av, av, av, av};
          acc[r][0] += ar * b0;
          acc[r][1] += ar * b1;
        }
      }
      for (int r = 0; r < mr; r++)
        for (int j = 0; j < 8; j++) {
          const double v = al[r] * (j < 4 ? acc[r][0][j] : acc[r][1][j - 4]);
          if (beta) c[r][j0 + j] += v;
          else c[r][j0 + j] = v;
        }
    }
    /* remaining n % 8 columns */
    for (int r = 0; r < mr; r++)
      for (int j = n8; j < n; j++) {
        double v = 0;
        for (int kk = 0; kk < k; kk++) v += a[r][kk] * B[(size_t)kk * n + j];
        if (beta) c[r][j] += al[r] * v;
        else c[r][j] = al[r] * v;
      }
  }
}

/* ------------------------------------------------------- evaluations --- */

/* potentials at npts points (SoA px, py, pz) from ns point sources */
static void pot_points(int ns, const double* sx, const double* sy, const double* sz, const double* sq, int npts,
                       const double* px, const double* py, const double* pz, double* out) {
  for (int j = 0; j < ns; j++) {
    const double x = sx[j], y = sy[j], z = sz[j], q = sq[j];
#pragma omp simd
    for (int i = 0; i < npts; i++) {
      const double dx = px[i] - x, dy = py[i] - y, dz = pz[i] - z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double inv = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
      out[i] += q * inv;
    }
  }
}

/* potential and gradient (gradient_P2P, laplace.h:91-164) at nt targets */
static void grad_points(int ns, const double* sx, const double* sy, const double* sz, const double* sq, int nt,
                        const double* tx, const double* ty, const double* tz, double* ph, double* gx, double* gy,
                        double* gz) {
  for (int j = 0; j < ns; j++) {
    const double x = sx[j], y = sy[j], z = sz[j], q = sq[j];
#pragma omp simd
    for (int i = 0; i < nt; i++) {
      const double dx = tx[i] - x, dy = ty[i] - y, dz = tz[i] - z;
      const double r2 = dx * dx + dy * dy + dz * dz;
      const double inv = r2 > 0 ? 1.0 / sqrt(r2) : 0.0;
      const double qi = q * inv, qi3 = qi * inv * inv;
      ph[i] += qi;
      gx[i] -= qi3 * dx;
      gy[i] -= qi3 * dy;
      gz[i] -= qi3 * dz;
    }
  }
}

static int lvl_of(const orc_tree* t, int n) {
  int l = 0;
  while (t->level_offset[l + 1] <= n) l++;
  return l;
}

/* transposed copy: out[j][i] = in[i][j] of an n x n matrix */
static double* transpose(const double* in, int n) {
  double* o = (double*)malloc(sizeof(double) * n * n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) o[(size_t)j * n + i] = in[(size_t)i * n + j];
  return o;
}

/* Full Laplace evaluation (upward + downward pass) into up, dn (double
 * [nnodes][nsurf]) and trg_value (double [ntrgs][4]), all overwritten.
 * stage_s (may be NULL) receives the seconds of: setup, P2M, M2M, P2L+M2P,
 * P2P, M2L forward transforms, Hadamard, inverse transforms, L2L, L2P.
 * Returns 0, or -1 for an unsupported kernel. */
int orc_evaluate_fast(const orc_tree* t, const orc_ops* o, const double* q, double* up, double* dn,
                      double* trg_value, double* stage_s) {
  double tick = omp_get_wtime();
#define STAGE(i)                                     \
  do {                                               \
    const double now_ = omp_get_wtime();             \
    if (stage_s) stage_s[i] = now_ - tick;           \
    tick = now_;                                     \
  } while (0)
  if (o->kind != LAPLACE || o->nlevels != 1 || o->nm2l != 1) return -1;
  const int p = o->p, ns = 6 * (p - 1) * (p - 1) + 2, nn = t->nnodes, D = t->depth;
  const size_t ns2 = (size_t)ns * ns;
  memset(up, 0, sizeof(double) * (size_t)nn * ns);
  memset(dn, 0, sizeof(double) * (size_t)nn * ns);
  memset(trg_value, 0, sizeof(double) * (size_t)t->ntrgs * 4);
  int* lev = (int*)malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++) lev[i] = lvl_of(t, i);
  /* unit surfaces */
  double* usurf = (double*)malloc(sizeof(double) * 3 * ns);
  const double c0[3] = {0, 0, 0};
  orc_surface(p, 1.0, 0, c0, 1.0, usurf); /* the unit-cube surface points */
  /* operators as row-GEMM B matrices (out = in x B): B = Op^T */
  double* uuT = transpose(o->uc2e_u, ns);
  double* uvT = transpose(o->uc2e_v, ns);
  double* duT = transpose(o->dc2e_u, ns);
  double* dvT = transpose(o->dc2e_v, ns);
  double* m2mcat = (double*)malloc(sizeof(double) * 8 * ns2); /* [c ns + k][j] = M2M[c][j][k] */
  double* l2lcat = (double*)malloc(sizeof(double) * 8 * ns2); /* [k][c ns + j] = L2L[c][j][k] */
  for (int c = 0; c < 8; c++)
    for (int j = 0; j < ns; j++)
      for (int k = 0; k < ns; k++) {
        m2mcat[((size_t)c * ns + k) * ns + j] = o->m2m[(size_t)c * ns2 + (size_t)j * ns + k];
        l2lcat[(size_t)k * 8 * ns + (size_t)c * ns + j] = o->l2l[(size_t)c * ns2 + (size_t)j * ns + k];
      }
  /* leaves with bodies, by level */
  int nleaf = 0;
  int* leaf = (int*)malloc(sizeof(int) * nn);
  for (int i = 0; i < nn; i++)
    if (t->child0[i] < 0 && (t->src_count[i] > 0 || t->trg_count[i] > 0)) leaf[nleaf++] = i;
  long* lrow_chk = (long*)malloc(sizeof(long) * (nleaf + 1));
  long* lrow_node = (long*)malloc(sizeof(long) * (nleaf + 1));
  double* lscale = (double*)malloc(sizeof(double) * (nleaf + 1));
  for (int i = 0; i < nleaf; i++) {
    lrow_chk[i] = (long)i * ns;
    lrow_node[i] = (long)leaf[i] * ns;
    lscale[i] = ldexp(1.0, -lev[leaf[i]]);
  }
  double* chk = (double*)calloc((size_t)(nleaf + 1) * ns, sizeof(double));
  double* tmp = (double*)calloc((size_t)(nleaf + 1) * ns, sizeof(double));

  STAGE(0);
  /* ---- P2M (fmm_scale_invariant.h:163-195): check potentials, then two GEMMs */
#pragma omp parallel
  {
    double *px = (double*)malloc(sizeof(double) * ns), *py = (double*)malloc(sizeof(double) * ns),
           *pz = (double*)malloc(sizeof(double) * ns);
    int cap = 0;
    double *sx = NULL, *sy = NULL, *sz = NULL;
#pragma omp for schedule(dynamic, 16)
    for (int li = 0; li < nleaf; li++) {
      const int l = leaf[li], n = t->src_count[l], b = t->src_begin[l];
      const double rr = 2.95 * t->r0 * ldexp(1.0, -lev[l]);
      for (int i = 0; i < ns; i++) {
        px[i] = t->center[3 * l] + rr * usurf[3 * i];
        py[i] = t->center[3 * l + 1] + rr * usurf[3 * i + 1];
        pz[i] = t->center[3 * l + 2] + rr * usurf[3 * i + 2];
      }
      if (n > cap) {
        cap = n;
        sx = (double*)realloc(sx, sizeof(double) * cap);
        sy = (double*)realloc(sy, sizeof(double) * cap);
        sz = (double*)realloc(sz, sizeof(double) * cap);
      }
      for (int j = 0; j < n; j++) {
        sx[j] = t->src_coord[3 * (size_t)(b + j)];
        sy[j] = t->src_coord[3 * (size_t)(b + j) + 1];
        sz[j] = t->src_coord[3 * (size_t)(b + j) + 2];
      }
      double* c = chk + (size_t)li * ns;
      pot_points(n, sx, sy, sz, q + b, ns, px, py, pz, c);
      for (int i = 0; i < ns; i++) c[i] *= INV4PI_F;
    }
    free(px); free(py); free(pz); free(sx); free(sy); free(sz);
  }
  gemm_rows(nleaf, ns, ns, chk, lrow_chk, uuT, tmp, lrow_chk, NULL, 0);
  gemm_rows(nleaf, ns, ns, tmp, lrow_chk, uvT, up, lrow_node, lscale, 0);

  STAGE(1);
  /* ---- M2M (fmm_scale_invariant.h:233-253): per level, parent += [children] x M2Mcat */
  long* prow = (long*)malloc(sizeof(long) * (nn + 1));
  long* crow = (long*)malloc(sizeof(long) * (nn + 1));
  for (int L0 = D - 1; L0 >= 0; L0--) {
    int m = 0;
    for (int n = t->level_offset[L0]; n < t->level_offset[L0 + 1]; n++)
      if (t->child0[n] >= 0) {
        prow[m] = (long)n * ns;
        crow[m] = (long)t->child0[n] * ns;
        m++;
      }
    if (m) gemm_rows(m, ns, 8 * ns, up, crow, m2mcat, up, prow, NULL, 1);
  }

  STAGE(2);
  /* ---- P2L, M2P (fmm_base.h:97-154): direct loops (empty for full trees) */
#pragma omp parallel
  {
    double *px = (double*)malloc(sizeof(double) * ns), *py = (double*)malloc(sizeof(double) * ns),
           *pz = (double*)malloc(sizeof(double) * ns), *acc = (double*)malloc(sizeof(double) * ns);
    double *sx = (double*)malloc(sizeof(double) * ns), *sy = (double*)malloc(sizeof(double) * ns),
           *sz = (double*)malloc(sizeof(double) * ns);
#pragma omp for schedule(dynamic)
    for (int n = 0; n < nn; n++) {
      if (t->p2l_offset[n + 1] == t->p2l_offset[n]) continue;
      const double rr = 1.05 * t->r0 * ldexp(1.0, -lev[n]);
      for (int i = 0; i < ns; i++) {
        px[i] = t->center[3 * n] + rr * usurf[3 * i];
        py[i] = t->center[3 * n + 1] + rr * usurf[3 * i + 1];
        pz[i] = t->center[3 * n + 2] + rr * usurf[3 * i + 2];
        acc[i] = 0;
      }
      for (int e = t->p2l_offset[n]; e < t->p2l_offset[n + 1]; e++) {
        const int s = t->p2l_source[e], b = t->src_begin[s];
        for (int j = 0; j < t->src_count[s]; j++) {
          const double* x = t->src_coord + 3 * (size_t)(b + j);
          pot_points(1, x, x + 1, x + 2, q + b + j, ns, px, py, pz, acc);
        }
      }
      for (int i = 0; i < ns; i++) dn[(size_t)n * ns + i] += acc[i] * INV4PI_F;
    }
#pragma omp for schedule(dynamic)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0 || t->m2p_offset[l + 1] == t->m2p_offset[l]) continue;
      const int nt = t->trg_count[l], tb = t->trg_begin[l];
      double* w = (double*)calloc((size_t)7 * (nt + 1), sizeof(double));
      double *tx = w, *ty = w + nt, *tz = w + 2 * nt, *ph = w + 3 * nt, *gx = w + 4 * nt, *gy = w + 5 * nt,
             *gz = w + 6 * nt;
      for (int i = 0; i < nt; i++) {
        tx[i] = t->trg_coord[3 * (size_t)(tb + i)];
        ty[i] = t->trg_coord[3 * (size_t)(tb + i) + 1];
        tz[i] = t->trg_coord[3 * (size_t)(tb + i) + 2];
      }
      for (int e = t->m2p_offset[l]; e < t->m2p_offset[l + 1]; e++) {
        const int s = t->m2p_source[e];
        const double rr = 1.05 * t->r0 * ldexp(1.0, -lev[s]);
        for (int i = 0; i < ns; i++) {
          sx[i] = t->center[3 * s] + rr * usurf[3 * i];
          sy[i] = t->center[3 * s + 1] + rr * usurf[3 * i + 1];
          sz[i] = t->center[3 * s + 2] + rr * usurf[3 * i + 2];
        }
        grad_points(ns, sx, sy, sz, up + (size_t)s * ns, nt, tx, ty, tz, ph, gx, gy, gz);
      }
      for (int i = 0; i < nt; i++) {
        double* v = trg_value + 4 * (size_t)(tb + i);
        v[0] += ph[i] * INV4PI_F;
        v[1] += gx[i] * INV4PI_F;
        v[2] += gy[i] * INV4PI_F;
        v[3] += gz[i] * INV4PI_F;
      }
      free(w);
    }
    free(px); free(py); free(pz); free(acc); free(sx); free(sy); free(sz);
  }

  STAGE(3);
  /* ---- P2P (fmm_base.h:82-94) */
#pragma omp parallel
  {
    int cap = 0, scap = 0;
    double* w = NULL;
    double *sx = NULL, *sy = NULL, *sz = NULL;
#pragma omp for schedule(dynamic, 8)
    for (int l = 0; l < nn; l++) {
      if (t->child0[l] >= 0 || t->trg_count[l] == 0 || t->p2p_offset[l + 1] == t->p2p_offset[l]) continue;
      const int nt = t->trg_count[l], tb = t->trg_begin[l];
      if (nt > cap) {
        cap = nt;
        w = (double*)realloc(w, sizeof(double) * 7 * cap);
      }
      double *tx = w, *ty = w + cap, *tz = w + 2 * cap, *ph = w + 3 * cap, *gx = w + 4 * cap, *gy = w + 5 * cap,
             *gz = w + 6 * cap;
      for (int i = 0; i < nt; i++) {
        tx[i] = t->trg_coord[3 * (size_t)(tb + i)];
        ty[i] = t->trg_coord[3 * (size_t)(tb + i) + 1];
        tz[i] = t->trg_coord[3 * (size_t)(tb + i) + 2];
        ph[i] = gx[i] = gy[i] = gz[i] = 0;
      }
      for (int e = t->p2p_offset[l]; e < t->p2p_offset[l + 1]; e++) {
        const int s = t->p2p_source[e], n = t->src_count[s], b = t->src_begin[s];
        if (n > scap) {
          scap = n;
          sx = (double*)realloc(sx, sizeof(double) * scap);
          sy = (double*)realloc(sy, sizeof(double) * scap);
          sz = (double*)realloc(sz, sizeof(double) * scap);
        }
        for (int j = 0; j < n; j++) {
          sx[j] = t->src_coord[3 * (size_t)(b + j)];
          sy[j] = t->src_coord[3 * (size_t)(b + j) + 1];
          sz[j] = t->src_coord[3 * (size_t)(b + j) + 2];
        }
        grad_points(n, sx, sy, sz, q + b, nt, tx, ty, tz, ph, gx, gy, gz);
      }
      for (int i = 0; i < nt; i++) {
        double* v = trg_value + 4 * (size_t)(tb + i);
        v[0] += ph[i] * INV4PI_F;
        v[1] += gx[i] * INV4PI_F;
        v[2] += gy[i] * INV4PI_F;
        v[3] += gz[i] * INV4PI_F;
      }
    }
    free(w); free(sx); free(sy); free(sz);
  }

  STAGE(4);
  /* ---- M2L (fmm_scale_invariant.h:407-440) */
  {
    const int n1 = 2 * p, nh = p + 1, nfreq = n1 * n1 * nh;
    int *mup = (int*)malloc(sizeof(int) * ns), *mdn = (int*)malloc(sizeof(int) * ns);
    orc_surf2conv(p, 0, mup);
    orc_surf2conv(p, 1, mdn);
    int* slot = (int*)malloc(sizeof(int) * nn);
    for (int i = 0; i < nn; i++) slot[i] = -1;
    int nsrc = 0, ntrg = 0;
    int* trgs = (int*)malloc(sizeof(int) * nn);
    for (int n = 0; n < nn; n++) {
      int any = 0;
      for (int r = 0; r < NM2L; r++) {
        const int s = t->m2l_source[(size_t)n * NM2L + r];
        if (s >= 0) {
          any = 1;
          if (slot[s] < 0) slot[s] = 0;
        }
      }
      if (any) trgs[ntrg++] = n;
    }
    int* srcs = (int*)malloc(sizeof(int) * nn);
    for (int n = 0; n < nn; n++)
      if (slot[n] >= 0) {
        slot[n] = nsrc;
        srcs[nsrc++] = n;
      }
    /* spectra [slot][f][child] (re, im); operators split into re / im planes [r][f][j2][j1] */
    double complex* fin = (double complex*)malloc(sizeof(double complex) * (size_t)nfreq * 8 * (nsrc + 1));
    double complex* fout = (double complex*)malloc(sizeof(double complex) * (size_t)nfreq * 8 * (ntrg + 1));
    double* mre = (double*)malloc(sizeof(double) * (size_t)NM2L * nfreq * 64);
    double* mim = (double*)malloc(sizeof(double) * (size_t)NM2L * nfreq * 64);
    for (size_t e = 0; e < (size_t)NM2L * nfreq * 64; e++) {
      mre[e] = o->m2l[2 * e];
      mim[e] = o->m2l[2 * e + 1];
    }
    fft_plan pf, pb;
    plan_init(&pf, n1, -1);
    plan_init(&pb, n1, +1);
    /* fft_up_equiv: each child's grid (data on [0,p)^3, layout [x][y][z]),
     * transformed along x (kept kx < nh: the r2c half spectrum), y and z;
     * zero rows pad every input to n1 */
    const int P2 = p * p, NB = nh * n1; /* batch sizes of the x and z passes */
#pragma omp parallel
    {
      const size_t big = (size_t)n1 * (NB > P2 ? NB : P2);
      double *ar = (double*)calloc(big, sizeof(double)), *ai = (double*)calloc(big, sizeof(double));
      double *br = (double*)calloc(big, sizeof(double)), *bi = (double*)calloc(big, sizeof(double));
      double *tr = (double*)malloc(sizeof(double) * 8 * (NB > P2 ? NB : P2)), *ti = (double*)malloc(sizeof(double) * 8 * (NB > P2 ? NB : P2));
#pragma omp for schedule(dynamic)
      for (int si = 0; si < nsrc; si++) {
        const int node = srcs[si];
        for (int c = 0; c < 8; c++) {
          const double* u = up + (size_t)(t->child0[node] + c) * ns;
          /* x pass: rows x (n1, rows >= p zero), batch (y, z) of p^2 */
          memset(ar, 0, sizeof(double) * (size_t)n1 * P2);
          memset(ai, 0, sizeof(double) * (size_t)n1 * P2);
          for (int i = 0; i < ns; i++) {
            const int m = mup[i], x = m % n1, y = (m / n1) % n1, z = m / (n1 * n1);
            ar[(size_t)x * P2 + y * p + z] = u[i];
          }
          fft_rows(&pf, ar, ai, br, bi, P2, P2, tr, ti); /* b[kx][y][z] */
          /* y pass per kx < nh: rows y (n1, rows >= p zero), batch z of p -> a[kx][ky][z] */
          for (int kx = 0; kx < nh; kx++) {
            double *inr = ar + (size_t)kx * n1 * p, *ini = ai + (size_t)kx * n1 * p;
            /* gather rows y < p of kx (zero rows y >= p) into a scratch slab */
            double slr[64 * 64], sli[64 * 64];
            for (int y = 0; y < n1; y++)
              for (int z = 0; z < p; z++) {
                slr[y * p + z] = y < p ? br[(size_t)kx * P2 + y * p + z] : 0.0;
                sli[y * p + z] = y < p ? bi[(size_t)kx * P2 + y * p + z] : 0.0;
              }
            fft_rows(&pf, slr, sli, inr, ini, p, p, tr, ti);
          }
          /* transpose to b[z][kx n1 + ky] (rows z >= p zero), z pass, batch nh n1 */
          memset(br + (size_t)p * NB, 0, sizeof(double) * (size_t)(n1 - p) * NB);
          memset(bi + (size_t)p * NB, 0, sizeof(double) * (size_t)(n1 - p) * NB);
          for (int kx = 0; kx < nh; kx++)
            for (int ky = 0; ky < n1; ky++)
              for (int z = 0; z < p; z++) {
                br[(size_t)z * NB + kx * n1 + ky] = ar[((size_t)kx * n1 + ky) * p + z];
                bi[(size_t)z * NB + kx * n1 + ky] = ai[((size_t)kx * n1 + ky) * p + z];
              }
          fft_rows(&pf, br, bi, ar, ai, NB, NB, tr, ti); /* a[kz][kx n1 + ky] */
          for (int kz = 0; kz < n1; kz++)
            for (int kx = 0; kx < nh; kx++)
              for (int ky = 0; ky < n1; ky++) {
                const size_t e = (size_t)kz * NB + kx * n1 + ky;
                fin[((size_t)si * nfreq + kx + (size_t)nh * (ky + (size_t)n1 * kz)) * 8 + c] = ar[e] + I * ai[e];
              }
        }
      }
      free(ar); free(ai); free(br); free(bi); free(tr); free(ti);
    }
    STAGE(5);
    /* hadamard_product: tasks of 64 targets x 32 frequencies; per frequency
     * and relative position the 8 x 8 operator is applied to every target of
     * the task while it sits in L1 */
    const int TB = 64, FB = 16;
    const int ntb = (ntrg + TB - 1) / TB, nfb = (nfreq + FB - 1) / FB;
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int bt = 0; bt < ntb; bt++)
      for (int bf = 0; bf < nfb; bf++) {
        const int t0 = bt * TB, nb = (t0 + TB < ntrg ? t0 + TB : ntrg) - t0;
        const int f0 = bf * FB, f1 = f0 + FB < nfreq ? f0 + FB : nfreq;
        int sl[64][NM2L];
        for (int ti = 0; ti < nb; ti++)
          for (int r = 0; r < NM2L; r++) {
            const int sn = t->m2l_source[(size_t)trgs[t0 + ti] * NM2L + r];
            sl[ti][r] = sn >= 0 ? slot[sn] : -1;
          }
        double are[64][8], aim[64][8];
        for (int f = f0; f < f1; f++) {
          memset(are, 0, sizeof(are));
          memset(aim, 0, sizeof(aim));
          for (int r = 0; r < NM2L; r++) {
            const double* Mr = mre + ((size_t)r * nfreq + f) * 64;
            const double* Mi = mim + ((size_t)r * nfreq + f) * 64;
            for (int ti = 0; ti < nb; ti++) {
              if (sl[ti][r] < 0) continue;
              const double* x = (const double*)(fin + ((size_t)sl[ti][r] * nfreq + f) * 8);
              v4d r0, r1, i0, i1;
              memcpy(&r0, are[ti], 32);
              memcpy(&r1, are[ti] + 4, 32);
              memcpy(&i0, aim[ti], 32);
              memcpy(&i1, aim[ti] + 4, 32);
              for (int j2 = 0; j2 < 8; j2++) {
                const v4d xr = {x[2 * j2], x[2 * j2], x[2 * j2], x[2 * j2]};
                const v4d xi = {x[2 * j2 + 1], x[2 * j2 + 1], x[2 * j2 + 1], x[2 * j2 + 1]};
                v4d mr0, mr1, mi0, mi1;
                memcpy(&mr0, Mr + 8 * j2, 32);
                memcpy(&mr1, Mr + 8 * j2 + 4, 32);
                memcpy(&mi0, Mi + 8 * j2, 32);
                memcpy(&mi1, Mi + 8 * j2 + 4, 32);
                r0 += mr0 * xr - mi0 * xi;
                r1 += mr1 * xr - mi1 * xi;
                i0 += mr0 * xi + mi0 * xr;
                i1 += mr1 * xi + mi1 * xr;
              }
              memcpy(are[ti], &r0, 32);
              memcpy(are[ti] + 4, &r1, 32);
              memcpy(aim[ti], &i0, 32);
              memcpy(aim[ti] + 4, &i1, 32);
            }
          }
          for (int ti = 0; ti < nb; ti++) {
            double complex* y = fout + ((size_t)(t0 + ti) * nfreq + f) * 8;
            for (int j1 = 0; j1 < 8; j1++) y[j1] = are[ti][j1] + I * aim[ti][j1];
          }
        }
      }
    STAGE(6);
    /* ifft_dn_check: inverse along z (batch kx, ky), y (per kx, batch z) and
     * x (c2r: Hermitian extension of kx, batch (y, z)); values needed only on
     * [p,2p)^3 */
#pragma omp parallel
    {
      const size_t big = (size_t)n1 * (NB > P2 ? NB : P2);
      double *ar = (double*)malloc(sizeof(double) * big), *ai = (double*)malloc(sizeof(double) * big);
      double *br = (double*)malloc(sizeof(double) * big), *bi = (double*)malloc(sizeof(double) * big);
      double *tr = (double*)malloc(sizeof(double) * 8 * (NB > P2 ? NB : P2)), *ti = (double*)malloc(sizeof(double) * 8 * (NB > P2 ? NB : P2));
#pragma omp for schedule(dynamic)
      for (int tj = 0; tj < ntrg; tj++) {
        const int node = trgs[tj];
        const double sc = ldexp(1.0, lev[node] + 1);
        for (int c = 0; c < 8; c++) {
          /* a[kz][kx n1 + ky] */
          for (int kz = 0; kz < n1; kz++)
            for (int kx = 0; kx < nh; kx++)
              for (int ky = 0; ky < n1; ky++) {
                const double complex v = fout[((size_t)tj * nfreq + kx + (size_t)nh * (ky + (size_t)n1 * kz)) * 8 + c];
                ar[(size_t)kz * NB + kx * n1 + ky] = creal(v);
                ai[(size_t)kz * NB + kx * n1 + ky] = cimag(v);
              }
          fft_rows(&pb, ar, ai, br, bi, NB, NB, tr, ti); /* b[z][kx n1 + ky], rows z in [p, 2p) used */
          /* y pass per kx: rows ky (n1), batch z' < p -> a[kx][y][z'] */
          for (int kx = 0; kx < nh; kx++) {
            double slr[64 * 64], sli[64 * 64];
            for (int ky = 0; ky < n1; ky++)
              for (int z = 0; z < p; z++) {
                slr[ky * p + z] = br[(size_t)(z + p) * NB + kx * n1 + ky];
                sli[ky * p + z] = bi[(size_t)(z + p) * NB + kx * n1 + ky];
              }
            fft_rows(&pb, slr, sli, ar + (size_t)kx * n1 * p, ai + (size_t)kx * n1 * p, p, p, tr, ti);
          }
          /* x pass: rows kx (Hermitian extension), batch (y', z') of p^2 */
          for (int kx = 0; kx < n1; kx++) {
            const int src = kx < nh ? kx : n1 - kx;
            const double sg = kx < nh ? 1.0 : -1.0;
            for (int y = 0; y < p; y++)
              for (int z = 0; z < p; z++) {
                br[(size_t)kx * P2 + y * p + z] = ar[((size_t)src * n1 + y + p) * p + z];
                bi[(size_t)kx * P2 + y * p + z] = sg * ai[((size_t)src * n1 + y + p) * p + z];
              }
          }
          fft_rows(&pb, br, bi, ar, ai, P2, P2, tr, ti); /* a[x][y'][z'] real parts, rows x in [p, 2p) */
          double* dv = dn + (size_t)(t->child0[node] + c) * ns;
          for (int i = 0; i < ns; i++) {
            const int m = mdn[i], x = m % n1, y = (m / n1) % n1 - p, z = m / (n1 * n1) - p;
            dv[i] += sc * ar[(size_t)x * P2 + y * p + z];
          }
        }
      }
      free(ar); free(ai); free(br); free(bi); free(tr); free(ti);
    }
    free(fin); free(fout); free(mre); free(mim); free(slot); free(srcs); free(trgs); free(mup); free(mdn);
  }

  STAGE(7);
  /* ---- L2L (fmm_scale_invariant.h:256-275): per level, [children] += parent x L2Lcat */
  for (int L0 = 0; L0 < D; L0++) {
    int m = 0;
    for (int n = t->level_offset[L0]; n < t->level_offset[L0 + 1]; n++)
      if (t->child0[n] >= 0) {
        prow[m] = (long)n * ns;
        crow[m] = (long)t->child0[n] * ns;
        m++;
      }
    if (m) gemm_rows(m, 8 * ns, ns, dn, prow, l2lcat, dn, crow, NULL, 1);
  }

  STAGE(8);
  /* ---- L2P (fmm_scale_invariant.h:198-230): two GEMMs, then the surface -> target gradient */
  gemm_rows(nleaf, ns, ns, dn, lrow_node, duT, tmp, lrow_chk, NULL, 0);
  gemm_rows(nleaf, ns, ns, tmp, lrow_chk, dvT, dn, lrow_node, lscale, 0);
#pragma omp parallel
  {
    double *sx = (double*)malloc(sizeof(double) * ns), *sy = (double*)malloc(sizeof(double) * ns),
           *sz = (double*)malloc(sizeof(double) * ns);
    int cap = 0;
    double* w = NULL;
#pragma omp for schedule(dynamic, 8)
    for (int li = 0; li < nleaf; li++) {
      const int l = leaf[li], nt = t->trg_count[l], tb = t->trg_begin[l];
      if (nt == 0) continue;
      if (nt > cap) {
        cap = nt;
        w = (double*)realloc(w, sizeof(double) * 7 * cap);
      }
      double *tx = w, *ty = w + cap, *tz = w + 2 * cap, *ph = w + 3 * cap, *gx = w + 4 * cap, *gy = w + 5 * cap,
             *gz = w + 6 * cap;
      const double rr = 2.95 * t->r0 * ldexp(1.0, -lev[l]);
      for (int i = 0; i < ns; i++) {
        sx[i] = t->center[3 * l] + rr * usurf[3 * i];
        sy[i] = t->center[3 * l + 1] + rr * usurf[3 * i + 1];
        sz[i] = t->center[3 * l + 2] + rr * usurf[3 * i + 2];
      }
      for (int i = 0; i < nt; i++) {
        tx[i] = t->trg_coord[3 * (size_t)(tb + i)];
        ty[i] = t->trg_coord[3 * (size_t)(tb + i) + 1];
        tz[i] = t->trg_coord[3 * (size_t)(tb + i) + 2];
        ph[i] = gx[i] = gy[i] = gz[i] = 0;
      }
      grad_points(ns, sx, sy, sz, dn + (size_t)l * ns, nt, tx, ty, tz, ph, gx, gy, gz);
      for (int i = 0; i < nt; i++) {
        double* v = trg_value + 4 * (size_t)(tb + i);
        v[0] += ph[i] * INV4PI_F;
        v[1] += gx[i] * INV4PI_F;
        v[2] += gy[i] * INV4PI_F;
        v[3] += gz[i] * INV4PI_F;
      }
    }
    free(sx); free(sy); free(sz); free(w);
  }
  free(lev); free(usurf); free(uuT); free(uvT); free(duT); free(dvT); free(m2mcat); free(l2lcat); free(leaf);
  free(lrow_chk); free(lrow_node); free(lscale); free(chk); free(tmp); free(prow); free(crow);
  STAGE(9);
#undef STAGE
  return 0;
}
