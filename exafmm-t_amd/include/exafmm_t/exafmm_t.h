// Core types of the MI355X exafmm-t evaluator host layer.
//
// Mirrors the public types of exafmm-t's include/exafmm_t.h:16-148 (Body, Node,
// Nodes, NodePtrs, Mat_Type, REL_COORD/HASH_LUT/M2L_INDEX_MAP) so that code
// written against the reference compiles against this header set.  Unlike the
// reference, every global is `inline` (C++17), so the headers may be included
// from several translation units.
#pragma once
#include <algorithm>
#include <cassert>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <set>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

// Tree variant (reference: -DNON_ADAPTIVE selects build_non_adaptive_tree.h
// and the target-only child lists of build_list.h:131-136).  The non-adaptive
// builder is always available as build_uniform_tree; the macro only decides
// which builder the reference's name build_tree denotes.
#if defined(NON_ADAPTIVE) && NON_ADAPTIVE
#define EXAFMM_T_NON_ADAPTIVE 1
#else
#define EXAFMM_T_NON_ADAPTIVE 0
#endif

namespace exafmm_t {

// Host arithmetic is always FP64; the device precision mode (FP64 or mixed)
// is chosen per FMM object (reference: exafmm_t.h:21-51 selects float/double
// at compile time with -DFLOAT).
typedef double real_t;
typedef std::complex<real_t> complex_t;
const real_t EPS = 1e-16;               // exafmm_t.h:38
const real_t PI = M_PI;
const int NCHILD = 8;
const int CACHE_SIZE = 512;             // exafmm_t.h:18, used by M2L target blocking

// Small fixed-size vector with the operators the FMM host code needs
// (the reference's vec<N,T> in vec.h, scalar only).
template <int N, typename T>
struct vec {
  T data[N];
  vec() { for (int i = 0; i < N; i++) data[i] = T(0); }
  vec(T v) { for (int i = 0; i < N; i++) data[i] = v; }
  T& operator[](int i) { return data[i]; }
  const T& operator[](int i) const { return data[i]; }
  vec& operator+=(const vec& o) { for (int i = 0; i < N; i++) data[i] += o.data[i]; return *this; }
  vec& operator-=(const vec& o) { for (int i = 0; i < N; i++) data[i] -= o.data[i]; return *this; }
  vec& operator*=(T s) { for (int i = 0; i < N; i++) data[i] *= s; return *this; }
  vec& operator/=(T s) { for (int i = 0; i < N; i++) data[i] /= s; return *this; }
  vec operator+(const vec& o) const { vec r(*this); r += o; return r; }
  vec operator-(const vec& o) const { vec r(*this); r -= o; return r; }
  vec operator+(T s) const { vec r(*this); for (int i = 0; i < N; i++) r.data[i] += s; return r; }
  vec operator-(T s) const { vec r(*this); for (int i = 0; i < N; i++) r.data[i] -= s; return r; }
  vec operator*(T s) const { vec r(*this); r *= s; return r; }
  vec operator/(T s) const { vec r(*this); r /= s; return r; }
};
template <int N, typename T> inline T norm(const vec<N, T>& v) {
  T s = 0; for (int i = 0; i < N; i++) s += v[i] * v[i]; return s;
}
template <int N, typename T> inline vec<N, T> min(const vec<N, T>& a, const vec<N, T>& b) {
  vec<N, T> r; for (int i = 0; i < N; i++) r[i] = std::min(a[i], b[i]); return r;
}
template <int N, typename T> inline vec<N, T> max(const vec<N, T>& a, const vec<N, T>& b) {
  vec<N, T> r; for (int i = 0; i < N; i++) r[i] = std::max(a[i], b[i]); return r;
}
template <int N, typename T> inline T max(const vec<N, T>& a) {
  T m = a[0]; for (int i = 1; i < N; i++) m = std::max(m, a[i]); return m;
}

typedef vec<3, int> ivec3;
typedef vec<3, real_t> vec3;
typedef vec<3, complex_t> cvec3;
typedef std::vector<real_t> RealVec;
typedef std::vector<complex_t> ComplexVec;

// Interaction types; same numbering as exafmm_t.h:69-80.
typedef enum {
  M2M_Type = 0,
  L2L_Type = 1,
  M2L_Helper_Type = 2,
  M2L_Type = 3,
  P2P0_Type = 4,
  P2P1_Type = 5,
  P2P2_Type = 6,
  M2P_Type = 7,
  P2L_Type = 8,
  Type_Count = 9
} Mat_Type;

// Body (exafmm_t.h:87-94).
template <typename T>
struct Body {
  int ibody;
  vec3 X;
  T q;
  T p;
  vec<3, T> F;
};
template <typename T> using Bodies = std::vector<Body<T>>;

// Node (exafmm_t.h:102-129).  Same fields; `up_equiv`/`dn_equiv` are only
// filled on the host when the FMM object is asked to copy them back from the
// device (they live in HBM during evaluation).
template <typename T>
struct Node {
  size_t idx = 0;
  size_t idx_M2L = 0;
  bool is_leaf = false;
  int ntrgs = 0;
  int nsrcs = 0;
  vec3 x;
  real_t r = 0;
  uint64_t key = 0;
  int level = 0;
  int octant = 0;
  Node* parent = nullptr;
  std::vector<Node*> children;
  std::vector<Node*> colleagues;
  std::vector<Node*> P2L_list;
  std::vector<Node*> M2P_list;
  std::vector<Node*> P2P_list;
  std::vector<Node*> M2L_list;
  std::vector<int> isrcs;
  std::vector<int> itrgs;
  RealVec src_coord;
  RealVec trg_coord;
  std::vector<T> src_value;
  std::vector<T> trg_value;
  std::vector<T> up_equiv;
  std::vector<T> dn_equiv;
};
template <typename T> using Nodes = std::vector<Node<T>>;
template <typename T> using NodePtrs = std::vector<Node<T>*>;
using Keys = std::vector<std::set<uint64_t>>;

// Relative-position tables (exafmm_t.h:146-148), filled by init_rel_coord().
inline std::vector<std::vector<ivec3>> REL_COORD;
inline std::vector<std::vector<int>> HASH_LUT;
inline std::vector<std::vector<int>> M2L_INDEX_MAP;

}  // namespace exafmm_t
