/*
 * oracle.h -- shared declarations of the oracle (oracle.c) and its FFT-based
 * CPU evaluation (fast.c).  TEST INFRASTRUCTURE ONLY: loaded by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg; the product never
 * links it.
 */
#ifndef ORACLE_H
#define ORACLE_H

#define LAPLACE 0
#define HELMHOLTZ 1
#define YUKAWA 2
#define NM2L 26

/* flattened tree: the fields of xfmm_tree (include/exafmm_t_amd.h) */
typedef struct {
  int nnodes, depth;
  double r0;
  double x0[3];
  const int* level_offset;
  const double* center;
  const int* child0;
  const int* src_begin;
  const int* src_count;
  const int* trg_begin;
  const int* trg_count;
  int nsrcs, ntrgs;
  const double* src_coord;
  const double* trg_coord;
  const int* p2p_offset;
  const int* p2p_source;
  const int* m2p_offset;
  const int* m2p_source;
  const int* p2l_offset;
  const int* p2l_source;
  const int* m2l_source;
} orc_tree;

typedef struct {
  int kind, p;
  double kre, kim;
  int nlevels, nm2l;
  const double *uc2e_u, *uc2e_v, *dc2e_u, *dc2e_v, *m2m, *l2l, *m2l;
} orc_ops;

enum { ST_P2M = 1, ST_M2M = 2, ST_P2L = 4, ST_M2P = 8, ST_P2P = 16, ST_M2L = 32, ST_L2L = 64, ST_L2P = 128 };

void orc_surface(int p, double r0, int level, const double* c, double alpha, double* out);
void orc_surf2conv(int p, int dn, int* map);

#endif
