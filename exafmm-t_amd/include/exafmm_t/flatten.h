// Flatten a host octree into the device layout of include/exafmm_t_amd.h.
//
// Device order: level-major; level l+1 is the concatenation, over the
// non-leaves of level l sorted along a 3-D Hilbert curve, of their 8 children
// in octant order.  Particles are grouped by leaf, leaves in Hilbert order.
// The Hilbert index is Skilling's transpose construction (J. Skilling,
// "Programming the Hilbert curve", AIP Conf. Proc. 707, 2004), which
// exafmm-t also carries (disabled) in include/hilbert.h:80-97.
#pragma once
#include <algorithm>
#include <numeric>
#include "exafmm_t.h"
#include "../../../include/exafmm_t_amd.h"

namespace exafmm_t {

// Hilbert index of integer point (x,y,z) with `bits` bits per axis.
inline uint64_t hilbert_index(uint32_t x, uint32_t y, uint32_t z, int bits) {
  if (bits <= 0) return 0;
  uint32_t X[3] = {x, y, z};
  const uint32_t M = 1u << (bits - 1);
  for (uint32_t Q = M; Q > 1; Q >>= 1) {  // inverse undo excess work
    const uint32_t P = Q - 1;
    for (int i = 0; i < 3; i++) {
      if (X[i] & Q) {
        X[0] ^= P;
      } else {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  for (int i = 1; i < 3; i++) X[i] ^= X[i - 1];  // Gray encode
  uint32_t t = 0;
  for (uint32_t Q = M; Q > 1; Q >>= 1)
    if (X[2] & Q) t ^= Q - 1;
  for (int i = 0; i < 3; i++) X[i] ^= t;
  uint64_t h = 0;  // interleave the transposed bits, most significant first
  for (int b = bits - 1; b >= 0; b--)
    for (int i = 0; i < 3; i++) h = (h << 1) | ((X[i] >> b) & 1u);
  return h;
}

template <typename T>
struct FlatTree {
  int depth = 0;
  real_t r0 = 0;
  vec3 x0;
  std::vector<int> level_offset, child0, src_begin, src_count, trg_begin, trg_count;
  std::vector<double> center, src_coord, trg_coord;
  std::vector<int> p2p_offset, p2p_source, m2p_offset, m2p_source, p2l_offset, p2l_source, m2l_source;
  std::vector<Node<T>*> dev2node;   // device index -> host node
  std::vector<int> node2dev;        // host Node::idx -> device index
  std::vector<Node<T>*> leaf_order; // leaves in particle order
  std::vector<int> src_ibody, trg_ibody;

  int nnodes() const { return int(dev2node.size()); }
  xfmm_tree view() const {
    xfmm_tree t;
    t.nnodes = nnodes();
    t.depth = depth;
    t.r0 = r0;
    for (int d = 0; d < 3; d++) t.x0[d] = x0[d];
    t.level_offset = level_offset.data();
    t.center = center.data();
    t.child0 = child0.data();
    t.src_begin = src_begin.data();
    t.src_count = src_count.data();
    t.trg_begin = trg_begin.data();
    t.trg_count = trg_count.data();
    t.nsrcs = int(src_ibody.size());
    t.ntrgs = int(trg_ibody.size());
    t.src_coord = src_coord.data();
    t.trg_coord = trg_coord.data();
    t.p2p_offset = p2p_offset.data();
    t.p2p_source = p2p_source.data();
    t.m2p_offset = m2p_offset.data();
    t.m2p_source = m2p_source.data();
    t.p2l_offset = p2l_offset.data();
    t.p2l_source = p2l_source.data();
    t.m2l_source = m2l_source.data();
    return t;
  }
};

template <typename T>
FlatTree<T> flatten(Nodes<T>& nodes, int depth, real_t r0, const vec3& x0) {
  FlatTree<T> f;
  f.depth = depth;
  f.r0 = r0;
  f.x0 = x0;
  const int nn = int(nodes.size());
  const int D = std::max(depth, 0);
  // Hilbert key of each node's lower corner at the finest level
  std::vector<uint64_t> hk(nn);
  for (int i = 0; i < nn; i++) {
    const Node<T>& n = nodes[i];
    const real_t dx = 2 * r0 / real_t(uint64_t(1) << n.level);
    uint32_t ix[3];
    for (int d = 0; d < 3; d++) {
      long v = long(std::floor((n.x[d] - (x0[d] - r0)) / dx));
      v = std::max(0L, std::min(v, (1L << n.level) - 1));
      ix[d] = uint32_t(v) << (D - n.level);
    }
    hk[i] = hilbert_index(ix[0], ix[1], ix[2], D);
  }
  std::vector<std::vector<Node<T>*>> lev(1, std::vector<Node<T>*>{&nodes[0]});
  for (int l = 0;; l++) {
    std::vector<Node<T>*> nl;
    for (Node<T>* n : lev[l])
      if (!n->is_leaf) nl.push_back(n);
    if (nl.empty()) break;
    std::stable_sort(nl.begin(), nl.end(), [&](Node<T>* a, Node<T>* b) { return hk[a->idx] < hk[b->idx]; });
    std::vector<Node<T>*> next;
    next.reserve(nl.size() * NCHILD);
    for (Node<T>* n : nl)
      for (int o = 0; o < NCHILD; o++) next.push_back(n->children[o]);
    lev.push_back(std::move(next));
  }
  if (int(lev.size()) - 1 != D && !(D == 0 && lev.size() == 1))
    throw std::runtime_error("flatten: tree depth does not match the FMM depth");
  f.level_offset.assign(lev.size() + 1, 0);
  for (size_t l = 0; l < lev.size(); l++) {
    f.level_offset[l + 1] = f.level_offset[l] + int(lev[l].size());
    for (Node<T>* n : lev[l]) f.dev2node.push_back(n);
  }
  if (int(f.dev2node.size()) != nn) throw std::runtime_error("flatten: unreachable nodes in the tree");
  f.node2dev.assign(nn, -1);
  for (int i = 0; i < nn; i++) f.node2dev[f.dev2node[i]->idx] = i;
  f.center.resize(size_t(nn) * 3);
  f.child0.assign(nn, -1);
  for (int i = 0; i < nn; i++) {
    Node<T>* n = f.dev2node[i];
    for (int d = 0; d < 3; d++) f.center[3 * size_t(i) + d] = n->x[d];
    if (!n->is_leaf) {
      f.child0[i] = f.node2dev[n->children[0]->idx];
      for (int o = 1; o < NCHILD; o++)
        if (f.node2dev[n->children[o]->idx] != f.child0[i] + o) throw std::runtime_error("flatten: child block broken");
    }
  }
  // particles: leaves in Hilbert order
  std::vector<int> leaves;
  for (int i = 0; i < nn; i++)
    if (f.dev2node[i]->is_leaf) leaves.push_back(i);
  std::stable_sort(leaves.begin(), leaves.end(),
                   [&](int a, int b) { return hk[f.dev2node[a]->idx] < hk[f.dev2node[b]->idx]; });
  f.src_begin.assign(nn, 0);
  f.src_count.assign(nn, 0);
  f.trg_begin.assign(nn, 0);
  f.trg_count.assign(nn, 0);
  const int nl = int(leaves.size());
  f.leaf_order.resize(nl);
  int ns_tot = 0, nt_tot = 0;
  for (int k = 0; k < nl; k++) {
    const int i = leaves[k];
    Node<T>* n = f.dev2node[i];
    f.leaf_order[k] = n;
    f.src_begin[i] = ns_tot;
    f.src_count[i] = int(n->src_coord.size() / 3);
    f.trg_begin[i] = nt_tot;
    f.trg_count[i] = int(n->trg_coord.size() / 3);
    ns_tot += f.src_count[i];
    nt_tot += f.trg_count[i];
  }
  f.src_coord.resize(3 * size_t(ns_tot));
  f.trg_coord.resize(3 * size_t(nt_tot));
  f.src_ibody.resize(ns_tot);
  f.trg_ibody.resize(nt_tot);
  // leaves own disjoint ranges: copy them in parallel
#pragma omp parallel for schedule(dynamic, 64)
  for (int k = 0; k < nl; k++) {
    const int i = leaves[k];
    Node<T>* n = f.dev2node[i];
    const int ns = f.src_count[i], nt = f.trg_count[i], sb = f.src_begin[i], tb = f.trg_begin[i];
    std::copy(n->src_coord.begin(), n->src_coord.end(), f.src_coord.begin() + 3 * size_t(sb));
    std::copy(n->trg_coord.begin(), n->trg_coord.end(), f.trg_coord.begin() + 3 * size_t(tb));
    for (int j = 0; j < ns; j++) f.src_ibody[sb + j] = j < int(n->isrcs.size()) ? n->isrcs[j] : -1;
    for (int j = 0; j < nt; j++) f.trg_ibody[tb + j] = j < int(n->itrgs.size()) ? n->itrgs[j] : -1;
  }
  // lists
  auto csr = [&](auto getter, std::vector<int>& off, std::vector<int>& lst) {
    off.assign(nn + 1, 0);
    lst.clear();
    for (int i = 0; i < nn; i++) {
      for (Node<T>* s : getter(f.dev2node[i])) lst.push_back(f.node2dev[s->idx]);
      off[i + 1] = int(lst.size());
    }
  };
  csr([](Node<T>* n) -> std::vector<Node<T>*>& { return n->P2P_list; }, f.p2p_offset, f.p2p_source);
  csr([](Node<T>* n) -> std::vector<Node<T>*>& { return n->M2P_list; }, f.m2p_offset, f.m2p_source);
  csr([](Node<T>* n) -> std::vector<Node<T>*>& { return n->P2L_list; }, f.p2l_offset, f.p2l_source);
  f.m2l_source.assign(size_t(nn) * 26, -1);
  for (int i = 0; i < nn; i++) {
    Node<T>* n = f.dev2node[i];
    for (size_t r = 0; r < n->M2L_list.size() && r < 26; r++)
      if (n->M2L_list[r]) f.m2l_source[size_t(i) * 26 + r] = f.node2dev[n->M2L_list[r]->idx];
  }
  return f;
}

}  // namespace exafmm_t
