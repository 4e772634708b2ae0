// Surfaces, convolution grids and relative-position tables.
//
// Restates exafmm-t include/geometry.h:21-215.  The point order of `surface`,
// the grid index order of `convolution_grid` and the relative-position order of
// `init_rel_coord` are data formats shared with the reference's precomputed
// operators, so they are kept identical; the code is this project's own.
#pragma once
#include "exafmm_t.h"

namespace exafmm_t {

// Unit-cube surface points in the reference order (geometry.h:24-54): the
// corner (-1,-1,-1), the interior+edge points of the three faces x=-1, y=-1,
// z=-1, then the point reflection of that first half.
inline RealVec unit_surface(int p) {
  const int n = 6 * (p - 1) * (p - 1) + 2;
  RealVec u(3 * n);
  auto coord = [&](int m) { return (2.0 * m - p + 1) / (p - 1); };
  int c = 0;
  auto put = [&](real_t a, real_t b, real_t d) { u[3 * c] = a; u[3 * c + 1] = b; u[3 * c + 2] = d; c++; };
  put(-1.0, -1.0, -1.0);
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(-1.0, coord(i + 1), coord(j));
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(coord(i), -1.0, coord(j + 1));
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(coord(i + 1), coord(j), -1.0);
  const int half = c;  // == n/2
  for (int i = 0; i < 3 * half; i++) u[3 * half + i] = -u[i];
  return u;
}

// Surface points of a box: half-side alpha * r0 * 0.5^level around c
// (geometry.h:21-63; alpha 2.95 upward check / downward equivalent,
// 1.05 upward equivalent / downward check).
inline RealVec surface(int p, real_t r0, int level, const real_t* c, real_t alpha) {
  RealVec u = unit_surface(p);
  const real_t r = r0 * std::pow(0.5, level);
  const real_t b = alpha * r;
  const int n = int(u.size() / 3);
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) u[3 * i + d] = u[3 * i + d] * b + c[d];
  return u;
}

// (2p)^3 convolution grid of a box (geometry.h:75-92); index i + n1*j + n1^2*k
// with i along x.
inline RealVec convolution_grid(int p, real_t r0, int level, const real_t* c) {
  const real_t d = 2 * r0 * std::pow(0.5, level);
  const real_t a = d * 1.05;
  const int n1 = 2 * p;
  RealVec g(3 * n1 * n1 * n1);
  for (int k = 0; k < n1; k++)
    for (int j = 0; j < n1; j++)
      for (int i = 0; i < n1; i++) {
        const size_t idx = size_t(i + n1 * j + n1 * n1 * k);
        g[3 * idx + 0] = (i - p) * a / (p - 1) + c[0];
        g[3 * idx + 1] = (j - p) * a / (p - 1) + c[1];
        g[3 * idx + 2] = (k - p) * a / (p - 1) + c[2];
      }
  return g;
}

// Surface point -> convolution-grid index maps (geometry.h:101-134).  The
// surface is laid on the integer lattice [0,p-1]^3; upward equivalent points
// go to (p-1-x, p-1-y, p-1-z), downward check points to (2p-1-x, ...).
inline std::vector<int> surf2conv(int p, int offset) {
  const int n1 = 2 * p;
  real_t c[3];
  for (int d = 0; d < 3; d++) c[d] = 0.5 * (p - 1);
  RealVec s = surface(p, 0.5, 0, c, real_t(p - 1));
  std::vector<int> map(s.size() / 3);
  for (size_t i = 0; i < map.size(); i++)
    map[i] = int(offset - s[3 * i]) + int(offset - s[3 * i + 1]) * n1 + int(offset - s[3 * i + 2]) * n1 * n1;
  return map;
}
inline std::vector<int> generate_surf2conv_up(int p) { return surf2conv(p, p - 1); }
inline std::vector<int> generate_surf2conv_dn(int p) { return surf2conv(p, 2 * p - 1); }

// Relative position hash, x + 10y + 100z + 555 (geometry.h:143-146).
inline int hash(const ivec3& c) {
  const int n = 5;
  return ((c[2] + n) * (2 * n) + (c[1] + n)) * (2 * n) + (c[0] + n);
}

inline void init_rel_coord(int max_r, int min_r, int step, Mat_Type t) {
  HASH_LUT[t].assign(2000, -1);
  REL_COORD[t].clear();
  for (int k = -max_r; k <= max_r; k += step)
    for (int j = -max_r; j <= max_r; j += step)
      for (int i = -max_r; i <= max_r; i += step)
        if (std::abs(i) >= min_r || std::abs(j) >= min_r || std::abs(k) >= min_r) {
          ivec3 c;
          c[0] = i; c[1] = j; c[2] = k;
          REL_COORD[t].push_back(c);
          HASH_LUT[t][hash(c)] = int(REL_COORD[t].size()) - 1;
        }
}

// M2L_INDEX_MAP[parent relpos][j2*8 + j1] -> M2L_Helper index of the child
// pair (target child j1, source child j2), or -1 for adjacent children
// (geometry.h:176-195).
inline void generate_M2L_index_map() {
  const int npos = int(REL_COORD[M2L_Type].size());
  M2L_INDEX_MAP.assign(npos, std::vector<int>(NCHILD * NCHILD, -1));
  for (int i = 0; i < npos; i++)
    for (int j1 = 0; j1 < NCHILD; j1++)
      for (int j2 = 0; j2 < NCHILD; j2++) {
        const ivec3& P = REL_COORD[M2L_Type][i];
        ivec3 c;
        for (int d = 0; d < 3; d++) c[d] = 2 * P[d] - ((j1 >> d) & 1) + ((j2 >> d) & 1);
        M2L_INDEX_MAP[i][j2 * NCHILD + j1] = HASH_LUT[M2L_Helper_Type][hash(c)];
      }
}

// Same table set as geometry.h:198-215 (once-guarded).
inline void init_rel_coord() {
  static bool done = false;
  if (done) return;
  REL_COORD.assign(Type_Count, {});
  HASH_LUT.assign(Type_Count, {});
  init_rel_coord(1, 1, 2, M2M_Type);
  init_rel_coord(1, 1, 2, L2L_Type);
  init_rel_coord(3, 3, 2, P2P0_Type);
  init_rel_coord(1, 0, 1, P2P1_Type);
  init_rel_coord(3, 3, 2, P2P2_Type);
  init_rel_coord(3, 2, 1, M2L_Helper_Type);
  init_rel_coord(1, 1, 1, M2L_Type);
  init_rel_coord(5, 5, 2, M2P_Type);
  init_rel_coord(5, 5, 2, P2L_Type);
  generate_M2L_index_map();
  done = true;
}

}  // namespace exafmm_t
