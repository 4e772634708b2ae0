// Adaptive octree with 2:1 balance (host side).
//
// Same construction rules as exafmm-t include/build_tree.h:20-359 and the
// Morton key helpers of include/hilbert.h:15-214 (EXAFMM_HILBERT=0 there):
//  * bounding cube r0 *= 1.00001 (build_tree.h:34),
//  * octant bits from strict `>` against the node centre (build_tree.h:57),
//  * a node is a leaf iff nsrcs <= ncrit && ntrgs <= ncrit (and, during the
//    balanced rebuild, its key is a leaf key) (build_tree.h:102),
//  * bodies inside a node are stably counting-sorted by octant,
//  * nodes are stored depth-first with the 8 children of a node appended as
//    one block when the node is split (build_tree.h:147-165),
//  * 2:1 balance by refining the parents of the colleagues of every refined
//    node, level by level from the bottom (build_tree.h:247-299); the tree is
//    rebuilt with the balanced leaf keys only when that set differs from the
//    tree's own (the reference always rebuilds; the result is the same).
// The octant sorts and the node fill run as OpenMP tasks over subtrees; the
// node order is fixed by the subtree sizes, not by the schedule.
#pragma once
#include <array>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <queue>
#include <type_traits>
#include <unordered_map>
#include "fmm_base.h"

namespace exafmm_t {

inline uint64_t levelOffset(int level) { return ((uint64_t(1) << (3 * level)) - 1) / 7; }

inline int getLevel(uint64_t key) {
  int level = -1;
  uint64_t offset = 0;
  while (key >= offset) {
    level++;
    offset += uint64_t(1) << (3 * level);
  }
  return level;
}

inline uint64_t getParent(uint64_t key) {
  const int l = getLevel(key);
  return (key - levelOffset(l)) / 8 + levelOffset(l - 1);
}

inline uint64_t getChild(uint64_t key) {
  const int l = getLevel(key);
  return (key - levelOffset(l)) * 8 + levelOffset(l + 1);
}

// Morton key: per level bit b, z -> bit 3b, y -> bit 3b+1, x -> bit 3b+2.
inline uint64_t getKey(ivec3 iX, int level, bool offset = true) {
  uint64_t k = 0;
  for (int b = 0; b < level; b++) {
    k |= uint64_t((iX[2] >> b) & 1) << (3 * b);
    k |= uint64_t((iX[1] >> b) & 1) << (3 * b + 1);
    k |= uint64_t((iX[0] >> b) & 1) << (3 * b + 2);
  }
  return offset ? k + levelOffset(level) : k;
}

inline ivec3 get3DIndex(uint64_t key) {
  const int level = getLevel(key);
  key -= levelOffset(level);
  ivec3 iX(0);
  for (int b = 0; b < level; b++) {
    iX[2] |= int((key >> (3 * b)) & 1) << b;
    iX[1] |= int((key >> (3 * b + 1)) & 1) << b;
    iX[0] |= int((key >> (3 * b + 2)) & 1) << b;
  }
  return iX;
}

inline ivec3 get3DIndex(vec3 X, int level, vec3 x0, real_t r0) {
  const real_t dx = 2 * r0 / (1 << level);
  ivec3 iX;
  for (int d = 0; d < 3; d++) iX[d] = int(std::floor((X[d] - (x0[d] - r0)) / dx));
  return iX;
}

template <typename T>
void get_bounds(const Bodies<T>& sources, const Bodies<T>& targets, vec3& x0, real_t& r0) {
  if (sources.empty() && targets.empty()) throw std::runtime_error("get_bounds: no bodies");
  vec3 lo = sources.empty() ? targets[0].X : sources[0].X;
  vec3 hi = lo;
  const long ns = long(sources.size()), nt = long(targets.size());
#pragma omp parallel
  {
    vec3 tlo = lo, thi = hi;
#pragma omp for nowait schedule(static)
    for (long i = 0; i < ns + nt; i++) {
      const vec3& x = i < ns ? sources[i].X : targets[i - ns].X;
      tlo = min(tlo, x);
      thi = max(thi, x);
    }
#pragma omp critical
    {
      lo = min(lo, tlo);
      hi = max(hi, thi);
    }
  }
  x0 = (hi + lo) / 2;
  r0 = std::max(max(x0 - lo), max(hi - x0));
  r0 *= 1.00001;
}

namespace detail {

inline int octant_of(const vec3& x, const vec3& c) {
  return int(x[0] > c[0]) + (int(x[1] > c[1]) << 1) + (int(x[2] > c[2]) << 2);
}

// Stable counting sort of bodies [b,e) by octant of centre c; returns the 9
// octant boundaries.  Large ranges count and scatter in parallel (per-thread
// histograms over consecutive chunks keep the sort stable).
template <typename T>
std::vector<int> split_octants(Body<T>* bodies, Body<T>* tmp, int b, int e, const vec3& c) {
  static_assert(std::is_trivially_copyable<Body<T>>::value, "bodies are moved with memcpy");
  std::vector<int> off(9, b);
  const int n = e - b;
  if (n < 200000) {
    std::vector<int> cnt(8, 0);
    for (int i = b; i < e; i++) cnt[octant_of(bodies[i].X, c)]++;
    for (int o = 0; o < 8; o++) off[o + 1] = off[o] + cnt[o];
    std::vector<int> pos(off.begin(), off.begin() + 8);
    for (int i = b; i < e; i++) std::memcpy(tmp + pos[octant_of(bodies[i].X, c)]++, bodies + i, sizeof(Body<T>));
    std::memcpy(bodies + b, tmp + b, sizeof(Body<T>) * size_t(n));
    return off;
  }
  constexpr int kChunks = 64;
  std::vector<std::array<int, 8>> cnt(kChunks);
#define XFMM_CHUNK_LO(k) (b + int((long long)n * (k) / kChunks))
  Body<T>* tp = tmp;
#pragma omp taskloop grainsize(1) shared(cnt) firstprivate(bodies, b, n, c)
  for (int k = 0; k < kChunks; k++) {
    std::array<int, 8> h{};
    for (int i = XFMM_CHUNK_LO(k); i < XFMM_CHUNK_LO(k + 1); i++) h[octant_of(bodies[i].X, c)]++;
    cnt[k] = h;
  }
  std::array<int, 8> tot{};
  for (int k = 0; k < kChunks; k++)
    for (int o = 0; o < 8; o++) tot[o] += cnt[k][o];
  for (int o = 0; o < 8; o++) off[o + 1] = off[o] + tot[o];
  // start of chunk k's bodies of octant o
  std::vector<std::array<int, 8>> start(kChunks);
  for (int o = 0; o < 8; o++) {
    int p = off[o];
    for (int k = 0; k < kChunks; k++) {
      start[k][o] = p;
      p += cnt[k][o];
    }
  }
#pragma omp taskloop grainsize(1) shared(start) firstprivate(bodies, tp, b, n, c)
  for (int k = 0; k < kChunks; k++) {
    std::array<int, 8> pos = start[k];
    for (int i = XFMM_CHUNK_LO(k); i < XFMM_CHUNK_LO(k + 1); i++)
      std::memcpy(tp + pos[octant_of(bodies[i].X, c)]++, bodies + i, sizeof(Body<T>));
  }
#pragma omp taskloop grainsize(1) firstprivate(bodies, tp, b, n)
  for (int k = 0; k < kChunks; k++)
    std::memcpy(bodies + XFMM_CHUNK_LO(k), tp + XFMM_CHUNK_LO(k),
                sizeof(Body<T>) * size_t(XFMM_CHUNK_LO(k + 1) - XFMM_CHUNK_LO(k)));
#undef XFMM_CHUNK_LO
  return off;
}

// Shape of a subtree, computed in parallel before any Node is written: body
// ranges after the octant sort and the size of the block of descendants.
struct Shape {
  int sb = 0, se = 0, tb = 0, te = 0, level = 0, octant = 0;
  vec3 x;
  real_t r = 0;
  uint64_t key = 0;
  bool leaf = true;
  size_t desc = 0;  // descendants (children blocks of the whole subtree)
  std::unique_ptr<Shape> ch[NCHILD];
};

// Ranges with more bodies than this are split in tasks of their own.
constexpr int kTaskBodies = 20000;

template <typename T>
struct TreeBuilder {
  Bodies<T>& src;
  Bodies<T>& trg;
  // sort scratch: raw storage (bodies are trivially copyable and only ever
  // memcpy'd in and out)
  Body<T>* stmp;
  Body<T>* ttmp;
  Nodes<T>& nodes;
  const Keys& leafkeys;
  FmmBase<T>& fmm;

  TreeBuilder(Bodies<T>& s, Bodies<T>& t, Nodes<T>& n, const Keys& lk, FmmBase<T>& f)
      : src(s), trg(t), stmp(static_cast<Body<T>*>(std::malloc(sizeof(Body<T>) * (s.size() + 1)))),
        ttmp(static_cast<Body<T>*>(std::malloc(sizeof(Body<T>) * (t.size() + 1)))), nodes(n), leafkeys(lk), fmm(f) {
    if (!stmp || !ttmp) throw std::bad_alloc();
  }
  ~TreeBuilder() {
    std::free(stmp);
    std::free(ttmp);
  }
  TreeBuilder(const TreeBuilder&) = delete;
  TreeBuilder& operator=(const TreeBuilder&) = delete;

  // Pass 1: leaf test, octant sort of the node's bodies, children.
  void shape(Shape* sh) {
    sh->key = getKey(get3DIndex(sh->x, sh->level, fmm.x0, fmm.r0), sh->level);
    bool leaf_key = true;
    if (!leafkeys.empty())
      leaf_key = sh->level < int(leafkeys.size()) && leafkeys[sh->level].count(sh->key) > 0;
    const int ns = sh->se - sh->sb, nt = sh->te - sh->tb;
    if (ns <= fmm.ncrit && nt <= fmm.ncrit && leaf_key) {
      sh->leaf = true;
      sh->desc = 0;
      return;
    }
    sh->leaf = false;
    std::vector<int> so = split_octants(src.data(), stmp, sh->sb, sh->se, sh->x);
    std::vector<int> to = split_octants(trg.data(), ttmp, sh->tb, sh->te, sh->x);
    for (int o = 0; o < NCHILD; o++) {
      Shape* c = new Shape;
      sh->ch[o].reset(c);
      c->r = sh->r / 2;
      c->x = sh->x;
      for (int d = 0; d < 3; d++) c->x[d] += c->r * (((o >> d) & 1) * 2 - 1);
      c->level = sh->level + 1;
      c->octant = o;
      c->sb = so[o];
      c->se = so[o + 1];
      c->tb = to[o];
      c->te = to[o + 1];
    }
    for (int o = 0; o < NCHILD; o++) {
      Shape* c = sh->ch[o].get();
      if ((c->se - c->sb) + (c->te - c->tb) > kTaskBodies) {
#pragma omp task firstprivate(c)
        shape(c);
      } else {
        shape(c);
      }
    }
#pragma omp taskwait
    sh->desc = NCHILD;
    for (int o = 0; o < NCHILD; o++) sh->desc += sh->ch[o]->desc;
  }

  // Pass 2: node `idx` and its descendants from `base` on: the 8 children of
  // a split node are one block, followed by the descendant blocks of child
  // 0, 1, ..., 7 (depth-first storage, build_tree.h:147-165).
  void fill(const Shape* sh, size_t idx, size_t base, Node<T>* parent) {
    Node<T>& nd = nodes[idx];
    nd.idx = idx;
    nd.x = sh->x;
    nd.r = sh->r;
    nd.level = sh->level;
    nd.octant = sh->octant;
    nd.parent = parent;
    nd.key = sh->key;
    nd.nsrcs = sh->se - sh->sb;
    nd.ntrgs = sh->te - sh->tb;
    nd.is_leaf = sh->leaf;
    if (sh->leaf) {
      nd.trg_value.assign(size_t(nd.ntrgs) * 4, T(0));
      nd.src_coord.resize(size_t(nd.nsrcs) * 3);
      nd.isrcs.resize(nd.nsrcs);
      nd.src_value.resize(nd.nsrcs);
      for (int i = 0; i < nd.nsrcs; i++) {
        const Body<T>& b = src[sh->sb + i];
        for (int d = 0; d < 3; d++) nd.src_coord[3 * size_t(i) + d] = b.X[d];
        nd.isrcs[i] = b.ibody;
        nd.src_value[i] = b.q;
      }
      nd.trg_coord.resize(size_t(nd.ntrgs) * 3);
      nd.itrgs.resize(nd.ntrgs);
      for (int i = 0; i < nd.ntrgs; i++) {
        const Body<T>& b = trg[sh->tb + i];
        for (int d = 0; d < 3; d++) nd.trg_coord[3 * size_t(i) + d] = b.X[d];
        nd.itrgs[i] = b.ibody;
      }
      return;
    }
    nd.children.resize(NCHILD);
    for (int o = 0; o < NCHILD; o++) nd.children[o] = &nodes[base + o];
    size_t next = base + NCHILD;
    for (int o = 0; o < NCHILD; o++) {
      const Shape* c = sh->ch[o].get();
      const size_t cb = next;
      next += c->desc;
      Node<T>* self = &nd;
      if ((c->se - c->sb) + (c->te - c->tb) > kTaskBodies) {
#pragma omp task firstprivate(c, cb, self, o)
        fill(c, base + o, cb, self);
      } else {
        fill(c, base + o, cb, self);
      }
    }
#pragma omp taskwait
  }
};

}  // namespace detail

#if !EXAFMM_T_NON_ADAPTIVE
// Build the (unbalanced, or with `leafkeys` the balanced) tree.  The bodies are
// reordered in place by octant path, as in the reference.
template <typename T>
Nodes<T> build_tree(Bodies<T>& sources, Bodies<T>& targets, NodePtrs<T>& leafs, NodePtrs<T>& nonleafs,
                    FmmBase<T>& fmm, const Keys& leafkeys = Keys()) {
  detail::Shape root;
  root.sb = 0;
  root.se = int(sources.size());
  root.tb = 0;
  root.te = int(targets.size());
  root.x = fmm.x0;
  root.r = fmm.r0;
  root.level = 0;
  Nodes<T> nodes;
  detail::TreeBuilder<T> tb(sources, targets, nodes, leafkeys, fmm);
#pragma omp parallel
#pragma omp single
  tb.shape(&root);
  nodes.resize(1 + root.desc);
#pragma omp parallel
#pragma omp single
  tb.fill(&root, 0, 1, nullptr);
  leafs.clear();
  nonleafs.clear();
  // leafs/nonleafs in creation (depth-first) order, as the reference pushes them
  std::vector<Node<T>*> stack{&nodes[0]};
  while (!stack.empty()) {
    Node<T>* n = stack.back();
    stack.pop_back();
    if (n->is_leaf) {
      if (n->nsrcs || n->ntrgs) leafs.push_back(n);
    } else {
      nonleafs.push_back(n);
      for (int o = NCHILD - 1; o >= 0; o--) stack.push_back(n->children[o]);
    }
  }
  return nodes;
}

// Per-level key sets of every node, breadth-first (build_tree.h:209-235).
template <typename T>
Keys breadth_first_traversal(Node<T>* root, std::unordered_map<uint64_t, size_t>& key2id) {
  Keys keys;
  std::queue<Node<T>*> q;
  q.push(root);
  while (!q.empty()) {
    Node<T>* n = q.front();
    q.pop();
    if (int(keys.size()) <= n->level) keys.resize(n->level + 1);
    keys[n->level].insert(n->key);
    key2id[n->key] = n->idx;
    if (!n->is_leaf)
      for (int o = 0; o < NCHILD; o++) q.push(n->children[o]);
  }
  return keys;
}

// Balanced key sets (build_tree.h:247-299).
template <typename T>
Keys balance_tree(const Keys& keys, const std::unordered_map<uint64_t, size_t>& key2id, const Nodes<T>& nodes) {
  const int nlevels = int(keys.size());
  const int maxlevel = nlevels - 1;
  Keys bkeys(nlevels);
  std::set<uint64_t> refine_below;  // S: level-l keys that must be refined
  for (int l = maxlevel; l > 0; --l) {
    std::set<uint64_t> refined = refine_below;  // N = S + nonleafs(l)
    for (uint64_t k : keys[l])
      if (!nodes[key2id.at(k)].is_leaf) refined.insert(k);
    refine_below.clear();
    const int nside = 1 << l;
    for (uint64_t k : refined) {
      const ivec3 iX = get3DIndex(k);
      for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            if (!dx && !dy && !dz) continue;
            ivec3 c;
            c[0] = iX[0] + dz;  // the neighbour set is symmetric in all three axes
            c[1] = iX[1] + dy;
            c[2] = iX[2] + dx;
            if (c[0] < 0 || c[1] < 0 || c[2] < 0 || c[0] >= nside || c[1] >= nside || c[2] >= nside) continue;
            refine_below.insert(getParent(getKey(c, l)));
          }
    }
    if (l != maxlevel) {
      for (uint64_t k : refined) {
        const uint64_t c0 = getChild(k);
        for (int o = 0; o < NCHILD; o++) bkeys[l + 1].insert(c0 + o);
      }
    }
  }
  bkeys[0].insert(0);
  for (uint64_t k = 1; k < 9; k++) bkeys[1].insert(k);
  return bkeys;
}

// Keys without children in the balanced set (build_tree.h:308-323).
inline Keys find_leaf_keys(const Keys& keys) {
  Keys leafkeys(keys.size());
  for (int l = int(keys.size()) - 1; l >= 1; --l) {
    std::set<uint64_t> parents = keys[l - 1];
    for (uint64_t k : keys[l]) parents.erase(getParent(k));
    leafkeys[l - 1] = parents;
  }
  leafkeys.back() = keys.back();
  return leafkeys;
}

// Rebuild the tree with the 2:1 constraint and set fmm.depth
// (build_tree.h:336-359).
template <typename T>
void balance_tree(Nodes<T>& nodes, Bodies<T>& sources, Bodies<T>& targets, NodePtrs<T>& leafs,
                  NodePtrs<T>& nonleafs, FmmBase<T>& fmm) {
  std::unordered_map<uint64_t, size_t> key2id;
  Keys keys = breadth_first_traversal(&nodes[0], key2id);
  fmm.depth = int(keys.size()) - 1;
  if (nodes.size() == 1) return;
  Keys bkeys = balance_tree(keys, key2id, nodes);
  // Already balanced (the balanced key set is the tree's own): the rebuild
  // would reproduce this tree node for node, so it is kept as it is.
  if (bkeys == keys) return;
  Keys leaf_keys = find_leaf_keys(bkeys);
  nodes.clear();
  leafs.clear();
  nonleafs.clear();
  nodes = build_tree(sources, targets, leafs, nonleafs, fmm, leaf_keys);
  // leafs/nonleafs hold pointers into the new vector (build_tree returned it by
  // move, so the element storage did not change)
}
#endif  // !EXAFMM_T_NON_ADAPTIVE

}  // namespace exafmm_t
