// Colleagues and interaction lists (host side).
//
// Same list semantics as exafmm-t include/build_list.h:16-197:
//  * colleagues: the 27 same-level neighbours, slot (dx+1) + 3(dy+1) + 9(dz+1)
//    (build_list.h:146-186);
//  * parent level (P2P0 / P2L): leaf colleagues of the parent; P2L switches to
//    P2P when the target is a leaf with ntrgs <= nsurf (build_list.h:16-43);
//  * current level (P2P1 / M2L): colleagues, M2L_list slot = M2L relpos index
//    (build_list.h:52-74);
//  * child level (P2P2 / M2P): children of non-leaf colleagues; M2P switches to
//    P2P when the source is a leaf with nsrcs <= nsurf (build_list.h:84-114).
#pragma once
#include "geometry.h"
#include "fmm_base.h"

namespace exafmm_t {

namespace detail {

// The colleagues of one node from those of its parent (which must be set).
template <typename T>
void colleagues_of(Node<T>* node) {
  node->colleagues.assign(27, nullptr);
  if (node->level == 0) {
    node->colleagues[13] = node;
    return;
  }
  const int o = node->octant;
  for (int i = 0; i < 27; i++) {
    Node<T>* pc = node->parent->colleagues[i];
    if (!pc || pc->is_leaf) continue;
    for (int j = 0; j < NCHILD; j++) {
      Node<T>* ch = pc->children[j];
      if (!ch) continue;
      int slot = 0, stride = 1;
      bool ok = true;
      for (int d = 0; d < 3; d++) {
        const int pd = (i / stride) % 3 - 1;                       // parent colleague offset
        const int diff = 2 * pd + ((j >> d) & 1) - ((o >> d) & 1);  // child offset
        if (diff < -1 || diff > 1) ok = false;
        slot += (diff + 1) * stride;
        stride *= 3;
      }
      if (ok) node->colleagues[slot] = ch;
    }
  }
}

}  // namespace detail

// Colleagues of node and of its subtree, top-down.
template <typename T>
void set_colleagues(Node<T>* node) {
  detail::colleagues_of(node);
  if (!node->is_leaf)
    for (int c = 0; c < NCHILD; c++)
      if (node->children[c]) set_colleagues(node->children[c]);
}

// Colleagues of every node, level by level (a node reads only its parent's
// colleagues, so the nodes of one level are independent).
template <typename T>
void set_colleagues(Nodes<T>& nodes) {
  int maxl = 0;
  for (auto& n : nodes) maxl = std::max(maxl, n.level);
  std::vector<std::vector<Node<T>*>> lev(maxl + 1);
  for (auto& n : nodes) lev[n.level].push_back(&n);
  for (auto& l : lev) {
#pragma omp parallel for schedule(dynamic, 256)
    for (long i = 0; i < long(l.size()); i++) detail::colleagues_of(l[i]);
  }
}

namespace detail {

template <typename T>
void list_parent_level(Node<T>* n, int nsurf) {
  if (!n->parent) return;
  const int o = n->octant;
  for (int i = 0; i < 27; i++) {
    Node<T>* pc = n->parent->colleagues[i];
    if (!pc || !pc->is_leaf) continue;
    ivec3 rel;
    rel[0] = (i % 3) * 4 - 4 - ((o & 1) ? 2 : 0) + 1;
    rel[1] = ((i / 3) % 3) * 4 - 4 - ((o & 2) ? 2 : 0) + 1;
    rel[2] = ((i / 9) % 3) * 4 - 4 - ((o & 4) ? 2 : 0) + 1;
    const int h = hash(rel);
    if (n->is_leaf && HASH_LUT[P2P0_Type][h] >= 0) n->P2P_list.push_back(pc);
    if (HASH_LUT[P2L_Type][h] >= 0) {
      if (n->is_leaf && n->ntrgs <= nsurf) n->P2P_list.push_back(pc);
      else n->P2L_list.push_back(pc);
    }
  }
}

template <typename T>
void list_current_level(Node<T>* n) {
  for (int i = 0; i < 27; i++) {
    Node<T>* col = n->colleagues[i];
    if (!col) continue;
    ivec3 rel;
    rel[0] = i % 3 - 1;
    rel[1] = (i / 3) % 3 - 1;
    rel[2] = (i / 9) % 3 - 1;
    const int h = hash(rel);
    if (col->is_leaf && n->is_leaf) {
      if (HASH_LUT[P2P1_Type][h] >= 0) n->P2P_list.push_back(col);
    } else if (!col->is_leaf && !n->is_leaf) {
      const int k = HASH_LUT[M2L_Type][h];
      if (k >= 0) n->M2L_list[k] = col;
    }
  }
}

template <typename T>
bool list_child_level(Node<T>* n, int nsurf) {
  if (!n->is_leaf) return true;
  for (int i = 0; i < 27; i++) {
    Node<T>* col = n->colleagues[i];
    if (!col || col->is_leaf) continue;
    for (int j = 0; j < NCHILD; j++) {
      Node<T>* cc = col->children[j];
      ivec3 rel;
      rel[0] = (i % 3) * 4 - 4 + ((j & 1) ? 2 : 0) - 1;
      rel[1] = ((i / 3) % 3) * 4 - 4 + ((j & 2) ? 2 : 0) - 1;
      rel[2] = ((i / 9) % 3) * 4 - 4 + ((j & 4) ? 2 : 0) - 1;
      const int h = hash(rel);
      if (HASH_LUT[P2P2_Type][h] >= 0) {
        if (!cc->is_leaf) return false;  // tree is not 2:1 balanced
        n->P2P_list.push_back(cc);
      }
      if (HASH_LUT[M2P_Type][h] >= 0) {
        if (cc->is_leaf && cc->nsrcs <= nsurf) n->P2P_list.push_back(cc);
        else n->M2P_list.push_back(cc);
      }
    }
  }
  return true;
}

}  // namespace detail

// non_adaptive: child-level lists (P2P2, M2P) only for nodes with targets, as
// the reference's NON_ADAPTIVE build does (build_list.h:131-136); in a
// non-adaptive tree the empty (ghost) leaves above the leaf level would
// otherwise see unrefined neighbours.
template <typename T>
void build_list(Nodes<T>& nodes, const FmmBase<T>& fmm, bool non_adaptive = EXAFMM_T_NON_ADAPTIVE) {
  init_rel_coord();
  const int npos = int(REL_COORD[M2L_Type].size());
  int unbalanced = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(|:unbalanced)
  for (long i = 0; i < long(nodes.size()); i++) {
    Node<T>* n = &nodes[i];
    n->P2P_list.clear();
    n->P2L_list.clear();
    n->M2P_list.clear();
    n->M2L_list.assign(npos, nullptr);
    detail::list_parent_level(n, fmm.nsurf);
    detail::list_current_level(n);
    if ((!non_adaptive || n->ntrgs > 0) && !detail::list_child_level(n, fmm.nsurf)) unbalanced = 1;
  }
  if (unbalanced) throw std::runtime_error("build_list: tree is not 2:1 balanced");
}

}  // namespace exafmm_t
