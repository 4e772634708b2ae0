// Non-adaptive (uniform) octree: the -DNON_ADAPTIVE build variant of exafmm-t
// (include/build_non_adaptive_tree.h:17-194).
//
//  * every node holding a source or a target is split down to fmm.depth
//    (set by the caller, e.g. from the reference tests' maxlevel,
//    tests/fmm_laplace.cpp:25-27); the nodes at fmm.depth are the leaves
//    (build_non_adaptive_tree.h:110-132);
//  * a node with neither sources nor targets is a leaf wherever it is (a
//    ghost: no bodies, not in `leafs`; build_non_adaptive_tree.h:104-108);
//  * bodies are counting-sorted by octant at every split and the 8 children
//    of a split node are appended as one block, as in the adaptive builder.
// No 2:1 balance is needed (all non-empty leaves share one level).  Lists are
// built with build_list(nodes, fmm, true).
#pragma once
#include "build_tree.h"

namespace exafmm_t {
namespace detail {

template <typename T>
struct UniformTreeBuilder {
  Bodies<T>& src;
  Bodies<T>& trg;
  std::vector<Body<T>> stmp, ttmp;
  Nodes<T>& nodes;
  FmmBase<T>& fmm;

  UniformTreeBuilder(Bodies<T>& s, Bodies<T>& t, Nodes<T>& n, FmmBase<T>& f)
      : src(s), trg(t), stmp(s.size()), ttmp(t.size()), nodes(n), fmm(f) {}

  void build(size_t ni, int sb, int se, int tb, int te) {
    Node<T>& nd = nodes[ni];
    nd.idx = ni;
    nd.nsrcs = se - sb;
    nd.ntrgs = te - tb;
    nd.key = getKey(get3DIndex(nd.x, nd.level, fmm.x0, fmm.r0), nd.level);
    if (nd.nsrcs == 0 && nd.ntrgs == 0) {  // ghost
      nd.is_leaf = true;
      return;
    }
    if (nd.level == fmm.depth) {
      nd.is_leaf = true;
      nd.trg_value.assign(size_t(nd.ntrgs) * 4, T(0));
      for (int i = sb; i < se; i++) {
        for (int d = 0; d < 3; d++) nd.src_coord.push_back(src[i].X[d]);
        nd.isrcs.push_back(src[i].ibody);
        nd.src_value.push_back(src[i].q);
      }
      for (int i = tb; i < te; i++) {
        for (int d = 0; d < 3; d++) nd.trg_coord.push_back(trg[i].X[d]);
        nd.itrgs.push_back(trg[i].ibody);
      }
      return;
    }
    nd.is_leaf = false;
    const vec3 c = nd.x;
    std::vector<int> so = split_octants(src.data(), stmp.data(), sb, se, c);
    std::vector<int> to = split_octants(trg.data(), ttmp.data(), tb, te, c);
    if (nodes.size() + NCHILD > nodes.capacity()) throw std::runtime_error("build_uniform_tree: node capacity exceeded");
    const size_t first = nodes.size();
    nodes.resize(first + NCHILD);
    Node<T>& parent = nodes[ni];  // capacity checked: no reallocation
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

// Uniform tree to fmm.depth over the bounding box fmm.x0, fmm.r0 (get_bounds).
// The bodies are reordered in place by octant path; leafs = the non-empty
// leaves, nonleafs = the split nodes, both in depth-first creation order.
template <typename T>
Nodes<T> build_uniform_tree(Bodies<T>& sources, Bodies<T>& targets, NodePtrs<T>& leafs, NodePtrs<T>& nonleafs,
                            FmmBase<T>& fmm) {
  if (fmm.depth < 0 || fmm.depth > 20) throw std::runtime_error("build_uniform_tree: fmm.depth must be in [0, 20]");
  // every split node holds a body, so at most depth * (ns + nt) + 1 splits;
  // never more than the full tree
  const double full = (std::pow(8.0, fmm.depth + 1) - 1.0) / 7.0;
  const double bound = 1.0 + 8.0 * (double(fmm.depth) * double(sources.size() + targets.size()) + 1.0);
  Nodes<T> nodes(1);
  nodes.reserve(size_t(std::min(full, bound)));
  nodes[0].parent = nullptr;
  nodes[0].octant = 0;
  nodes[0].x = fmm.x0;
  nodes[0].r = fmm.r0;
  nodes[0].level = 0;
  detail::UniformTreeBuilder<T> tb(sources, targets, nodes, fmm);
  tb.build(0, 0, int(sources.size()), 0, int(targets.size()));
  leafs.clear();
  nonleafs.clear();
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

#if EXAFMM_T_NON_ADAPTIVE
// The reference's name for the builder in a NON_ADAPTIVE build.
template <typename T>
Nodes<T> build_tree(Bodies<T>& sources, Bodies<T>& targets, NodePtrs<T>& leafs, NodePtrs<T>& nonleafs,
                    FmmBase<T>& fmm) {
  return build_uniform_tree(sources, targets, leafs, nonleafs, fmm);
}
#endif

}  // namespace exafmm_t
