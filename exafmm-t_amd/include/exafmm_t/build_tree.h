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
//    node, level by level from the bottom (build_tree.h:247-299).
#pragma once
#include <unordered_map>
#include <queue>
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
  for (const auto& b : sources) { lo = min(lo, b.X); hi = max(hi, b.X); }
  for (const auto& b : targets) { lo = min(lo, b.X); hi = max(hi, b.X); }
  x0 = (hi + lo) / 2;
  r0 = std::max(max(x0 - lo), max(hi - x0));
  r0 *= 1.00001;
}

namespace detail {

inline int octant_of(const vec3& x, const vec3& c) {
  return int(x[0] > c[0]) + (int(x[1] > c[1]) << 1) + (int(x[2] > c[2]) << 2);
}

// Stable counting sort of bodies [b,e) by octant of centre c; returns the 9
// octant boundaries.
template <typename T>
std::vector<int> split_octants(Body<T>* bodies, std::vector<Body<T>>& tmp, int b, int e, const vec3& c) {
  std::vector<int> cnt(8, 0), off(9, b);
  for (int i = b; i < e; i++) cnt[octant_of(bodies[i].X, c)]++;
  for (int o = 0; o < 8; o++) off[o + 1] = off[o] + cnt[o];
  std::vector<int> pos(off.begin(), off.begin() + 8);
  for (int i = b; i < e; i++) tmp[pos[octant_of(bodies[i].X, c)]++] = bodies[i];
  for (int i = b; i < e; i++) bodies[i] = tmp[i];
  return off;
}

template <typename T>
struct TreeBuilder {
  Bodies<T>& src;
  Bodies<T>& trg;
  std::vector<Body<T>> stmp, ttmp;
  Nodes<T>& nodes;
  const Keys& leafkeys;
  FmmBase<T>& fmm;
  std::vector<std::vector<int>> children;  // per node index: 8 child indices

  TreeBuilder(Bodies<T>& s, Bodies<T>& t, Nodes<T>& n, const Keys& lk, FmmBase<T>& f)
      : src(s), trg(t), stmp(s.size()), ttmp(t.size()), nodes(n), leafkeys(lk), fmm(f) {}

  void build(size_t ni, int sb, int se, int tb, int te) {
    Node<T>& nd = nodes[ni];
    nd.idx = ni;
    nd.nsrcs = se - sb;
    nd.ntrgs = te - tb;
    nd.up_equiv.assign(fmm.nsurf, T(0));
    nd.dn_equiv.assign(fmm.nsurf, T(0));
    nd.key = getKey(get3DIndex(nd.x, nd.level, fmm.x0, fmm.r0), nd.level);
    bool leaf_key = true;
    if (!leafkeys.empty()) {
      leaf_key = nd.level < int(leafkeys.size()) && leafkeys[nd.level].count(nd.key) > 0;
    }
    if (nd.nsrcs <= fmm.ncrit && nd.ntrgs <= fmm.ncrit && leaf_key) {
      nd.is_leaf = true;
      nd.trg_value.assign(size_t(nd.ntrgs) * 4, T(0));
      if (!leafkeys.empty()) {
        for (int i = sb; i < se; i++) {
          for (int d = 0; d < 3; d++) nd.src_coord.push_back(src[i].X[d]);
          nd.isrcs.push_back(src[i].ibody);
          nd.src_value.push_back(src[i].q);
        }
        for (int i = tb; i < te; i++) {
          for (int d = 0; d < 3; d++) nd.trg_coord.push_back(trg[i].X[d]);
          nd.itrgs.push_back(trg[i].ibody);
        }
      }
      return;
    }
    nd.is_leaf = false;
    const vec3 c = nd.x;
    std::vector<int> so = split_octants(src.data(), stmp, sb, se, c);
    std::vector<int> to = split_octants(trg.data(), ttmp, tb, te, c);
    if (nodes.size() + NCHILD > nodes.capacity()) throw std::runtime_error("build_tree: node capacity exceeded");
    const size_t first = nodes.size();
    nodes.resize(first + NCHILD);
    Node<T>& parent = nodes[ni];  // re-fetch: resize does not reallocate (capacity checked)
    for (int o = 0; o < NCHILD; o++) {
      Node<T>& ch = nodes[first + o];
      ch.r = parent.r / 2;
      ch.x = parent.x;
      for (int d = 0; d < 3; d++) ch.x[d] += ch.r * (((o >> d) & 1) * 2 - 1);
      ch.parent = &parent;
      ch.octant = o;
      ch.level = parent.level + 1;
    }
    parent.children.resize(NCHILD);
    for (int o = 0; o < NCHILD; o++) parent.children[o] = &nodes[first + o];
    for (int o = 0; o < NCHILD; o++) build(first + o, so[o], so[o + 1], to[o], to[o + 1]);
  }
};

}  // namespace detail

#if !EXAFMM_T_NON_ADAPTIVE
// Build the (unbalanced, or with `leafkeys` the balanced) tree.  The bodies are
// reordered in place by octant path, as in the reference.
template <typename T>
Nodes<T> build_tree(Bodies<T>& sources, Bodies<T>& targets, NodePtrs<T>& leafs, NodePtrs<T>& nonleafs,
                    FmmBase<T>& fmm, const Keys& leafkeys = Keys()) {
  Nodes<T> nodes(1);
  nodes.reserve((sources.size() + targets.size()) * (32 / std::max(fmm.ncrit, 1) + 1) + 64);
  nodes[0].parent = nullptr;
  nodes[0].octant = 0;
  nodes[0].x = fmm.x0;
  nodes[0].r = fmm.r0;
  nodes[0].level = 0;
  detail::TreeBuilder<T> tb(sources, targets, nodes, leafkeys, fmm);
  tb.build(0, 0, int(sources.size()), 0, int(targets.size()));
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
  Keys leaf_keys = (nodes.size() == 1) ? keys : find_leaf_keys(balance_tree(keys, key2id, nodes));
  nodes.clear();
  leafs.clear();
  nonleafs.clear();
  nodes = build_tree(sources, targets, leafs, nonleafs, fmm, leaf_keys);
  // leafs/nonleafs hold pointers into the new vector (build_tree returned it by
  // move, so the element storage did not change)
  fmm.depth = int(keys.size()) - 1;
}
#endif  // !EXAFMM_T_NON_ADAPTIVE

}  // namespace exafmm_t
