// pybind11 module `exafmm` (built as exafmm._exafmm): the Python API of
// exafmm-t (reference python/exafmm.cpp:35-537) on top of the gfx950 engine.
//
// Submodules laplace / helmholtz / modified_helmholtz with Body, Node, Tree,
// *Fmm (verify, wavek, plus `precision`), init_sources, init_targets, setup,
// evaluate, update_charges, clear_values.  Bodies/Nodes are exposed as opaque
// vectors instead of Python lists (same element access, no per-body copies).
// Extra, for tests and benchmarks: Fmm.flat_tree() / Fmm.operators() return
// the exact arrays handed to the C ABI, and Fmm.sorted_charges().
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>
#include <pybind11/stl_bind.h>
#include <pybind11/complex.h>

#include "../include/exafmm_t/build_list.h"
#include "../include/exafmm_t/build_tree.h"
#include "../include/exafmm_t/build_non_adaptive_tree.h"
#include "../include/exafmm_t/helmholtz.h"
#include "../include/exafmm_t/laplace.h"
#include "../include/exafmm_t/modified_helmholtz.h"

namespace py = pybind11;
using namespace exafmm_t;

PYBIND11_MAKE_OPAQUE(Bodies<real_t>);
PYBIND11_MAKE_OPAQUE(Bodies<complex_t>);
PYBIND11_MAKE_OPAQUE(Nodes<real_t>);
PYBIND11_MAKE_OPAQUE(Nodes<complex_t>);
PYBIND11_MAKE_OPAQUE(NodePtrs<real_t>);
PYBIND11_MAKE_OPAQUE(NodePtrs<complex_t>);

template <typename T>
struct Tree {
  Nodes<T> nodes;
  NodePtrs<T> leafs;
  NodePtrs<T> nonleafs;
};

template <typename T>
Bodies<T> init_sources(py::array_t<real_t, py::array::c_style | py::array::forcecast> coords,
                       py::array_t<T, py::array::c_style | py::array::forcecast> charges) {
  if (coords.ndim() != 2 || charges.ndim() != 1 || coords.shape(1) != 3)
    throw std::runtime_error("coords should have a shape of (n, 3), charges should have a shape of (n,)");
  if (coords.shape(0) != charges.shape(0)) throw std::runtime_error("Number of coords and charges must match");
  const ssize_t n = charges.shape(0);
  Bodies<T> b(n);
  auto c = coords.template unchecked<2>();
  auto q = charges.template unchecked<1>();
#pragma omp parallel for schedule(static)
  for (ssize_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) b[i].X[d] = c(i, d);
    b[i].q = q(i);
    b[i].ibody = int(i);
  }
  return b;
}

template <typename T>
Bodies<T> init_targets(py::array_t<real_t, py::array::c_style | py::array::forcecast> coords) {
  if (coords.ndim() != 2 || coords.shape(1) != 3) throw std::runtime_error("coords should have a shape of (n, 3)");
  const ssize_t n = coords.shape(0);
  Bodies<T> b(n);
  auto c = coords.template unchecked<2>();
#pragma omp parallel for schedule(static)
  for (ssize_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) b[i].X[d] = c(i, d);
    b[i].ibody = int(i);
  }
  return b;
}

// setup (python/exafmm.cpp:155-180): tree + 2:1 balance, lists, M2L_setup,
// precompute.  The bodies are reordered in place (as in the reference).
// precompute = false (an addition) skips the operators, for callers that
// share one operator set between many trees (exafmm.batch).
template <typename T, typename F>
Tree<T> setup(Bodies<T>& sources, Bodies<T>& targets, F& fmm, bool precompute) {
  if (sources.empty() || targets.empty()) throw std::runtime_error("setup: sources and targets must be non-empty");
  Tree<T> tree;
  get_bounds<T>(sources, targets, fmm.x0, fmm.r0);
  tree.nodes = build_tree<T>(sources, targets, tree.leafs, tree.nonleafs, fmm);
  balance_tree<T>(tree.nodes, sources, targets, tree.leafs, tree.nonleafs, fmm);
  init_rel_coord();
  set_colleagues<T>(tree.nodes);
  build_list<T>(tree.nodes, fmm);
  fmm.M2L_setup(tree.nonleafs);
  if (precompute) fmm.precompute();
  return tree;
}

// setup of the reference's NON_ADAPTIVE build (python/exafmm.cpp:133-134 with
// build_non_adaptive_tree.h; depth as the C++ tests set it,
// tests/fmm_laplace.cpp:25-27): uniform tree to `depth`, lists with
// target-only child lists (build_list.h:131-136), M2L_setup, precompute.
template <typename T, typename F>
Tree<T> setup_non_adaptive(Bodies<T>& sources, Bodies<T>& targets, F& fmm, int depth) {
  if (sources.empty() || targets.empty()) throw std::runtime_error("setup: sources and targets must be non-empty");
  if (depth < 0 || depth > 20) throw std::runtime_error("setup_non_adaptive: depth must be in [0, 20]");
  Tree<T> tree;
  get_bounds<T>(sources, targets, fmm.x0, fmm.r0);
  fmm.depth = depth;
  tree.nodes = build_uniform_tree<T>(sources, targets, tree.leafs, tree.nonleafs, fmm);
  init_rel_coord();
  set_colleagues<T>(tree.nodes);
  build_list<T>(tree.nodes, fmm, true);
  fmm.M2L_setup(tree.nonleafs);
  fmm.precompute();
  return tree;
}

// evaluate (python/exafmm.cpp:190-267): upward + downward pass, results in
// the original target order as an (ntrgs, 4) array.
template <typename T, typename F>
py::array_t<T> evaluate(Tree<T>& tree, F& fmm, bool verbose) {
  {
    py::gil_scoped_release nogil;
    fmm.upward_pass(tree.nodes, tree.leafs, verbose);
    fmm.downward_pass(tree.nodes, tree.leafs, verbose);
  }
  const ssize_t nt = tree.nodes[0].ntrgs;
  py::array_t<T> out({nt, ssize_t(4)});
  auto r = out.template mutable_unchecked<2>();
  for (Node<T>* leaf : tree.leafs)
    for (size_t j = 0; j < leaf->itrgs.size(); j++)
      for (int d = 0; d < 4; d++) r(leaf->itrgs[j], d) = leaf->trg_value[4 * j + d];
  return out;
}

template <typename T>
void update_charges(Tree<T>& tree, py::array_t<T, py::array::c_style | py::array::forcecast> charges) {
  auto q = charges.template unchecked<1>();
  for (Node<T>* leaf : tree.leafs)
    for (size_t j = 0; j < leaf->isrcs.size(); j++) leaf->src_value[j] = q(leaf->isrcs[j]);
}

template <typename T>
void clear_values(Tree<T>& tree) {
  for (auto& n : tree.nodes) {
    std::fill(n.up_equiv.begin(), n.up_equiv.end(), T(0));
    std::fill(n.dn_equiv.begin(), n.dn_equiv.end(), T(0));
    if (n.is_leaf) std::fill(n.trg_value.begin(), n.trg_value.end(), T(0));
  }
}

template <typename V>
py::array_t<V> as_array(const std::vector<V>& v) {
  py::array_t<V> a(ssize_t(v.size()));
  if (!v.empty()) std::memcpy(a.mutable_data(), v.data(), v.size() * sizeof(V));
  return a;
}

template <typename T, typename F>
py::dict flat_tree(F& fmm, Tree<T>& tree) {
  fmm.ensure_flat(tree.nodes);
  const FlatTree<T>& f = fmm.flat_tree();
  py::dict d;
  d["nnodes"] = f.nnodes();
  d["depth"] = f.depth;
  d["r0"] = f.r0;
  d["x0"] = std::vector<double>{f.x0[0], f.x0[1], f.x0[2]};
  d["level_offset"] = as_array(f.level_offset);
  d["center"] = as_array(f.center);
  d["child0"] = as_array(f.child0);
  d["src_begin"] = as_array(f.src_begin);
  d["src_count"] = as_array(f.src_count);
  d["trg_begin"] = as_array(f.trg_begin);
  d["trg_count"] = as_array(f.trg_count);
  d["src_coord"] = as_array(f.src_coord);
  d["trg_coord"] = as_array(f.trg_coord);
  d["p2p_offset"] = as_array(f.p2p_offset);
  d["p2p_source"] = as_array(f.p2p_source);
  d["m2p_offset"] = as_array(f.m2p_offset);
  d["m2p_source"] = as_array(f.m2p_source);
  d["p2l_offset"] = as_array(f.p2l_offset);
  d["p2l_source"] = as_array(f.p2l_source);
  d["m2l_source"] = as_array(f.m2l_source);
  d["src_ibody"] = as_array(f.src_ibody);
  d["trg_ibody"] = as_array(f.trg_ibody);
  std::vector<int> lvl(f.nnodes()), idx(f.nnodes());
  for (int i = 0; i < f.nnodes(); i++) {
    lvl[i] = f.dev2node[i]->level;
    idx[i] = int(f.dev2node[i]->idx);
  }
  d["level"] = as_array(lvl);
  d["node_idx"] = as_array(idx);
  return d;
}

template <typename T, typename F>
py::dict operators(F& fmm) {
  py::dict d;
  auto cat = [](const std::vector<std::vector<T>>& v) {
    std::vector<T> o;
    for (auto& x : v) o.insert(o.end(), x.begin(), x.end());
    return as_array(o);
  };
  auto cat3 = [](const std::vector<std::vector<std::vector<T>>>& v) {
    std::vector<T> o;
    for (auto& l : v)
      for (auto& x : l) o.insert(o.end(), x.begin(), x.end());
    return as_array(o);
  };
  d["nlevels"] = fmm.op_levels();
  d["nm2l"] = fmm.m2l_slots();
  d["uc2e_u"] = cat(fmm.matrix_UC2E_U);
  d["uc2e_v"] = cat(fmm.matrix_UC2E_V);
  d["dc2e_u"] = cat(fmm.matrix_DC2E_U);
  d["dc2e_v"] = cat(fmm.matrix_DC2E_V);
  d["m2m"] = cat3(fmm.matrix_M2M);
  d["l2l"] = cat3(fmm.matrix_L2L);
  std::vector<double> m;
  for (auto& x : fmm.matrix_M2L) m.insert(m.end(), x.begin(), x.end());
  d["m2l"] = as_array(m);
  return d;
}

template <typename T, typename F>
py::array_t<T> sorted_charges(F& fmm, Tree<T>& tree) {
  fmm.ensure_flat(tree.nodes);
  std::vector<T> q;
  for (Node<T>* n : fmm.flat_tree().leaf_order) q.insert(q.end(), n->src_value.begin(), n->src_value.end());
  return as_array(q);
}

template <typename T>
void bind_common(py::module& m, const char* suffix) {
  (void)suffix;
  py::class_<Body<T>>(m, "Body")
      .def(py::init<>())
      .def_readwrite("q", &Body<T>::q)
      .def_readwrite("p", &Body<T>::p)
      .def_property("X", [](const Body<T>& b) { return std::vector<double>{b.X[0], b.X[1], b.X[2]}; },
                    [](Body<T>& b, std::vector<double> x) { for (int d = 0; d < 3; d++) b.X[d] = x.at(d); })
      .def_property("F", [](const Body<T>& b) { return std::vector<T>{b.F[0], b.F[1], b.F[2]}; },
                    [](Body<T>& b, std::vector<T> x) { for (int d = 0; d < 3; d++) b.F[d] = x.at(d); })
      .def_readwrite("ibody", &Body<T>::ibody);
  py::bind_vector<Bodies<T>>(m, "Bodies");
  py::class_<Node<T>>(m, "Node")
      .def(py::init<>())
      .def_readwrite("isrcs", &Node<T>::isrcs)
      .def_readwrite("itrgs", &Node<T>::itrgs)
      .def_readwrite("trg_value", &Node<T>::trg_value)
      .def_readwrite("key", &Node<T>::key)
      .def_readonly("parent", &Node<T>::parent, py::return_value_policy::reference)
      .def_property_readonly("x", [](const Node<T>& n) { return std::vector<double>{n.x[0], n.x[1], n.x[2]}; })
      .def_readwrite("r", &Node<T>::r)
      .def_readwrite("nsrcs", &Node<T>::nsrcs)
      .def_readwrite("ntrgs", &Node<T>::ntrgs)
      .def_readwrite("level", &Node<T>::level)
      .def_readwrite("is_leaf", &Node<T>::is_leaf)
      .def_readwrite("up_equiv", &Node<T>::up_equiv)
      .def_readwrite("dn_equiv", &Node<T>::dn_equiv);
  py::bind_vector<Nodes<T>>(m, "Nodes");
  py::bind_vector<NodePtrs<T>>(m, "NodePtrs");
  py::class_<Tree<T>>(m, "Tree")
      .def(py::init<>())
      .def_readonly("nodes", &Tree<T>::nodes)
      .def_readonly("leafs", &Tree<T>::leafs)
      .def_readonly("nonleafs", &Tree<T>::nonleafs);
  m.def("init_sources", &init_sources<T>, "initialize sources", py::arg("coords"), py::arg("charges"));
  m.def("init_targets", &init_targets<T>, "initialize targets", py::arg("coords"));
  m.def("update_charges", &update_charges<T>, "update charges of sources");
  m.def("clear_values", &clear_values<T>, "clear target potentials, equivalent charges and check potentials");
}

template <typename T, typename F>
void bind_fmm(py::module& m, py::class_<F>& c) {
  c.def("verify", &F::verify, py::arg("leafs"), py::arg("sample") = true)
      .def_readwrite("p", &F::p)
      .def_readwrite("ncrit", &F::ncrit)
      .def_readonly("nsurf", &F::nsurf)
      .def_readonly("nfreq", &F::nfreq)
      .def_readonly("depth", &F::depth)
      .def_readonly("r0", &F::r0)
      .def_property_readonly("x0", [](const F& f) { return std::vector<double>{f.x0[0], f.x0[1], f.x0[2]}; })
      .def(
          "set_bounds",
          [](F& f, std::vector<double> x0, double r0) {
            if (x0.size() != 3 || !(r0 > 0)) throw std::runtime_error("set_bounds: x0 needs 3 values, r0 > 0");
            for (int d = 0; d < 3; d++) f.x0[d] = x0[d];
            f.r0 = r0;
          },
          "set the root box (centre x0, half side r0) that precompute() uses, without a tree",
          py::arg("x0"), py::arg("r0"))
      .def_readwrite("filename", &F::filename)
      .def_readwrite("precision", &F::precision,
                     "0: FP64 device path; 1: mixed (FP32 P2P and M2L, FP64 pseudo-inverse solves and "
                     "translations); 2: FLOAT (all evaluation stages FP32).  1 and 2 truncate the pseudo-inverses "
                     "at 4e-8 max(S); set before setup or call precompute()")
      .def_readwrite("device", &F::device)
      .def("flat_tree", [](F& f, Tree<T>& t) { return flat_tree<T, F>(f, t); })
      .def("operators", [](F& f) { return operators<T, F>(f); })
      .def("precompute", &F::precompute,
           "recompute the operators for the current precision (mixed and FLOAT truncate the pseudo-inverses "
           "at 4e-8 max(S))")
      .def("sorted_charges", [](F& f, Tree<T>& t) { return sorted_charges<T, F>(f, t); })
      .def("sync_equiv_to_host", [](F& f, Tree<T>& t) { f.sync_equiv_to_host(t.nodes); })
      .def("release_device", &F::release_device);
  m.def("setup", &setup<T, F>, "setup FMM: tree, 2:1 balance, lists, M2L setup and precomputation",
        py::arg("sources"), py::arg("targets"), py::arg("fmm"), py::arg("precompute") = true);
  m.def("setup_non_adaptive", &setup_non_adaptive<T, F>,
        "setup FMM on the non-adaptive (uniform) tree of the given depth: tree, lists, M2L setup, precomputation",
        py::arg("sources"), py::arg("targets"), py::arg("fmm"), py::arg("depth"));
  m.def("evaluate", &evaluate<T, F>, "evaluate potential and gradient", py::arg("tree"), py::arg("fmm"),
        py::arg("verbose") = false);
}

PYBIND11_MODULE(_exafmm, m) {
  m.doc() = "MI355X (gfx950) evaluator with the exafmm-t Python API";
  py::module m0 = m.def_submodule("laplace", "Laplace kernel");
  py::module m1 = m.def_submodule("helmholtz", "Helmholtz kernel");
  py::module m2 = m.def_submodule("modified_helmholtz", "Modified Helmholtz kernel");

  bind_common<real_t>(m0, "");
  bind_common<complex_t>(m1, "");
  for (const char* a : {"Body", "Bodies", "Node", "Nodes", "NodePtrs", "Tree", "init_sources", "init_targets",
                        "update_charges", "clear_values"})
    m2.attr(a) = m0.attr(a);

  py::class_<LaplaceFmm> lc(m0, "LaplaceFmm");
  lc.def(py::init<>()).def(py::init<int, int, std::string>(), py::arg("p"), py::arg("ncrit"),
                           py::arg("filename") = std::string());
  bind_fmm<real_t, LaplaceFmm>(m0, lc);

  py::class_<HelmholtzFmm> hc(m1, "HelmholtzFmm");
  hc.def(py::init<>())
      .def(py::init<int, int, complex_t, std::string>(), py::arg("p"), py::arg("ncrit"), py::arg("wavek"),
           py::arg("filename") = std::string())
      .def_readwrite("wavek", &HelmholtzFmm::wavek);
  bind_fmm<complex_t, HelmholtzFmm>(m1, hc);

  py::class_<ModifiedHelmholtzFmm> yc(m2, "ModifiedHelmholtzFmm");
  yc.def(py::init<>())
      .def(py::init<int, int, real_t, std::string>(), py::arg("p"), py::arg("ncrit"), py::arg("wavek"),
           py::arg("filename") = std::string())
      .def_readwrite("wavek", &ModifiedHelmholtzFmm::wavek);
  bind_fmm<real_t, ModifiedHelmholtzFmm>(m2, yc);
}
