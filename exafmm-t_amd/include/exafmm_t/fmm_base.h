// FMM object: host precomputation + device (gfx950) evaluation.
//
// Public surface follows exafmm-t's FmmBase<T> (include/fmm_base.h:11-268) and
// its derived FmmScaleInvariant<T> / Fmm<T> (precompute, M2L_setup,
// upward_pass, downward_pass, verify, potential_P2P, gradient_P2P).  The
// evaluation runs through the C ABI of include/exafmm_t_amd.h; this header
// only prepares its inputs (operators, flattened tree, sorted charges) and
// moves results back into Node::trg_value.
#pragma once
#include <cstdio>
#include <fstream>
#include <memory>
#include "exafmm_t.h"
#include "flatten.h"
#include "geometry.h"
#include "linalg.h"
#include "../../../include/exafmm_t_amd.h"

namespace exafmm_t {

inline void xfmm_check(int rc, const xfmm_context* ctx, const char* what) {
  if (rc != XFMM_OK)
    throw std::runtime_error(std::string(what) + " failed (" + std::to_string(rc) + "): " + xfmm_last_error(ctx));
}

template <typename T>
class FmmBase {
 public:
  int p = 0;
  int nsurf = 0;
  int nconv = 0;
  int nfreq = 0;
  int ncrit = 0;
  int depth = 0;
  real_t r0 = 0;
  vec3 x0;
  bool is_precomputed = false;
  bool is_real = std::is_same<T, real_t>::value;
  std::string filename;
  int precision = XFMM_FP64;  // XFMM_FP64, XFMM_MIXED or XFMM_FLOAT (device arithmetic of the stages)
  int device = 0;             // HIP device ordinal

  // Precomputed operators, FP64, indexed by level (a single level for the
  // scale-invariant Laplace kernel): reference fmm_scale_invariant.h:21-27,
  // fmm.h:15-20.  matrix_M2L[slot] is [26][nfreq][8*8][2] (slot = target level).
  std::vector<std::vector<T>> matrix_UC2E_U, matrix_UC2E_V, matrix_DC2E_U, matrix_DC2E_V;
  std::vector<std::vector<std::vector<T>>> matrix_M2M, matrix_L2L;
  std::vector<RealVec> matrix_M2L;

  FmmBase() {}
  FmmBase(int p_, int ncrit_, std::string filename_ = std::string()) : p(p_), ncrit(ncrit_), filename(filename_) {
    set_order(p_);
  }
  FmmBase(const FmmBase&) = delete;
  FmmBase& operator=(const FmmBase&) = delete;
  virtual ~FmmBase() { release_device(); }

  void set_order(int p_) {
    p = p_;
    nsurf = 6 * (p - 1) * (p - 1) + 2;
    const int n1 = 2 * p;
    nconv = n1 * n1 * n1;
    nfreq = is_real ? n1 * n1 * (n1 / 2 + 1) : nconv;
  }

  // kernel identity (XFMM_LAPLACE / XFMM_HELMHOLTZ / XFMM_MODIFIED_HELMHOLTZ) and wave number
  virtual int kernel_kind() const = 0;
  virtual complex_t kernel_wavek() const { return complex_t(0, 0); }
  // true for the scale-invariant Laplace kernel: one operator set for all levels
  bool scale_invariant() const { return kernel_kind() == XFMM_LAPLACE; }

  // Potential of a unit charge at distance vector d (host, exact FP64).
  T kernel_value(real_t dx, real_t dy, real_t dz) const {
    const real_t r2 = dx * dx + dy * dy + dz * dz;
    if (r2 == 0) return T(0);
    const real_t r = std::sqrt(r2);
    const real_t inv4pi = 1.0 / (4 * PI);
    const int kind = kernel_kind();
    if constexpr (std::is_same<T, complex_t>::value) {
      const complex_t k = kernel_wavek();
      return std::exp(complex_t(0, 1) * k * r) / r * inv4pi;
    } else {
      if (kind == XFMM_MODIFIED_HELMHOLTZ) return std::exp(-kernel_wavek().real() * r) / r * inv4pi;
      return inv4pi / r;
    }
  }

  // matrix[i * ntrgs + j] = K(src_i, trg_j) (reference fmm_base.h:65-76).
  void kernel_matrix(const RealVec& src_coord, const RealVec& trg_coord, std::vector<T>& matrix) const {
    const int ns = int(src_coord.size() / 3), nt = int(trg_coord.size() / 3);
    matrix.resize(size_t(ns) * nt);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ns; i++)
      for (int j = 0; j < nt; j++)
        matrix[size_t(i) * nt + j] = kernel_value(trg_coord[3 * j] - src_coord[3 * i],
                                                  trg_coord[3 * j + 1] - src_coord[3 * i + 1],
                                                  trg_coord[3 * j + 2] - src_coord[3 * i + 2]);
  }

  // ------------------------------------------------------------ precompute
  int op_levels() const { return scale_invariant() ? 1 : depth + 1; }
  int m2l_slots() const { return scale_invariant() ? 1 : std::max(depth, 1); }

  // Pseudo-inverse factors of the check-to-equivalent matrix at `level`
  // (fmm_scale_invariant.h:446-514, fmm.h:451-521): A = U S V^H,
  // UC2E_U = U^H, UC2E_V = V S^+, DC2E_U = V^T (= VH for real), DC2E_V = conj(U) S^+,
  // singular values below 4 eps max(S) dropped (eps = pinv_eps()).
  void precompute_check2equiv(int level) {
    real_t c[3] = {0, 0, 0};
    RealVec up_check = surface(p, r0, level, c, 2.95);
    RealVec up_equiv = surface(p, r0, level, c, 1.05);
    std::vector<T> A;
    kernel_matrix(up_check, up_equiv, A);
    std::vector<T> U, VH;
    RealVec S;
    svd(nsurf, A, U, S, VH);
    const real_t smax = S.empty() ? 0 : S[0];
    RealVec Sinv(nsurf);
    const real_t eps = pinv_eps();
    for (int i = 0; i < nsurf; i++) Sinv[i] = S[i] > eps * smax * 4 ? 1.0 / S[i] : 0.0;
    std::vector<T> UH = conjugate_transpose(U, nsurf, nsurf);
    std::vector<T> V = conjugate_transpose(VH, nsurf, nsurf);
    std::vector<T>& uu = matrix_UC2E_U[level];
    std::vector<T>& uv = matrix_UC2E_V[level];
    std::vector<T>& du = matrix_DC2E_U[level];
    std::vector<T>& dv = matrix_DC2E_V[level];
    uu = UH;
    uv.assign(size_t(nsurf) * nsurf, T(0));
    dv.assign(size_t(nsurf) * nsurf, T(0));
    du = transpose(V, nsurf, nsurf);
    for (int i = 0; i < nsurf; i++)
      for (int j = 0; j < nsurf; j++) {
        uv[size_t(i) * nsurf + j] = V[size_t(i) * nsurf + j] * Sinv[j];
        dv[size_t(i) * nsurf + j] = conj_if(U[size_t(i) * nsurf + j]) * Sinv[j];
      }
  }

  // M2M / L2L for parent level `level` (fmm_scale_invariant.h:51-77, fmm.h:47-74).
  void precompute_M2M(int level) {
    real_t pc[3] = {0, 0, 0};
    RealVec parent_up_check = surface(p, r0, level, pc, 2.95);
    const real_t s = r0 * std::pow(0.5, level + 1);
    for (int i = 0; i < 8; i++) {
      const ivec3& rel = REL_COORD[M2M_Type][i];
      real_t cc[3] = {pc[0] + rel[0] * s, pc[1] + rel[1] * s, pc[2] + rel[2] * s};
      RealVec child_up_equiv = surface(p, r0, level + 1, cc, 1.05);
      std::vector<T> K, buf(size_t(nsurf) * nsurf);
      kernel_matrix(parent_up_check, child_up_equiv, K);
      gemm(nsurf, nsurf, nsurf, matrix_UC2E_U[level].data(), K.data(), buf.data());
      matrix_M2M[level][i].assign(size_t(nsurf) * nsurf, T(0));
      gemm(nsurf, nsurf, nsurf, matrix_UC2E_V[level].data(), buf.data(), matrix_M2M[level][i].data());
      std::vector<T> Kt = transpose(K, nsurf, nsurf);
      gemm(nsurf, nsurf, nsurf, Kt.data(), matrix_DC2E_V[level].data(), buf.data());
      matrix_L2L[level][i].assign(size_t(nsurf) * nsurf, T(0));
      gemm(nsurf, nsurf, nsurf, buf.data(), matrix_DC2E_U[level].data(), matrix_L2L[level][i].data());
    }
  }

  // Frequency-domain M2L operators for child level `child_level`
  // (fmm_scale_invariant.h:517-558 with child_level = 0 and relative
  // coordinates x 2 r0; fmm.h:524-629 with child_level = l, coordinates x r0 2^(1-l)).
  void precompute_M2L(int slot, int child_level) {
    const int n1 = 2 * p;
    const int nhelper = int(REL_COORD[M2L_Helper_Type].size());
    std::vector<std::vector<complex_t>> helper(nhelper);
    const real_t unit = scale_invariant() ? r0 / 0.5 : r0 * std::pow(0.5, child_level - 1);
    RealVec origin(3, 0.0);
#pragma omp parallel for schedule(dynamic)
    for (int i = 0; i < nhelper; i++) {
      real_t coord[3];
      for (int d = 0; d < 3; d++) coord[d] = REL_COORD[M2L_Helper_Type][i][d] * unit;
      RealVec grid = convolution_grid(p, r0, child_level, coord);
      std::vector<T> val;
      kernel_matrix(grid, origin, val);
      if constexpr (std::is_same<T, complex_t>::value) helper[i] = dft3_c2c(val, n1);
      else helper[i] = dft3_r2c(val, n1);
    }
    RealVec& M = matrix_M2L[slot];
    M.assign(size_t(26) * nfreq * 128, 0.0);
    for (int i = 0; i < 26; i++)
      for (int j = 0; j < 64; j++) {
        const int h = M2L_INDEX_MAP[i][j];
        if (h < 0) continue;
        for (int k = 0; k < nfreq; k++) {
          const size_t o = (size_t(i) * nfreq + k) * 128 + 2 * j;
          M[o] = helper[h][k].real() / nconv;
          M[o + 1] = helper[h][k].imag() / nconv;
        }
      }
  }

  void initialize_matrix() {
    const int L = op_levels();
    matrix_UC2E_U.assign(L, {});
    matrix_UC2E_V.assign(L, {});
    matrix_DC2E_U.assign(L, {});
    matrix_DC2E_V.assign(L, {});
    matrix_M2M.assign(L, std::vector<std::vector<T>>(8));
    matrix_L2L.assign(L, std::vector<std::vector<T>>(8));
    matrix_M2L.assign(m2l_slots(), {});
  }

  // Binary cache in the reference layout (fmm_scale_invariant.h:86-148 and
  // fmm.h:77-139): r0 | per level {UC2E_U, UC2E_V, DC2E_U, DC2E_V, M2M[8], L2L[8]} | M2L.
  size_t matrix_file_size() const {
    return size_t(op_levels()) * (2 * 8 + 4) * nsurf * nsurf * sizeof(T) +
           size_t(m2l_slots()) * 26 * nfreq * 128 * sizeof(real_t) + sizeof(real_t);
  }
  void save_matrix() const {
    if (filename.empty()) return;
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    if (!f) return;
    f.write(reinterpret_cast<const char*>(&r0), sizeof(real_t));
    const size_t n2 = size_t(nsurf) * nsurf * sizeof(T);
    for (int l = 0; l < op_levels(); l++) {
      f.write(reinterpret_cast<const char*>(matrix_UC2E_U[l].data()), n2);
      f.write(reinterpret_cast<const char*>(matrix_UC2E_V[l].data()), n2);
      f.write(reinterpret_cast<const char*>(matrix_DC2E_U[l].data()), n2);
      f.write(reinterpret_cast<const char*>(matrix_DC2E_V[l].data()), n2);
      for (auto& m : matrix_M2M[l]) f.write(reinterpret_cast<const char*>(m.data()), n2);
      for (auto& m : matrix_L2L[l]) f.write(reinterpret_cast<const char*>(m.data()), n2);
    }
    for (auto& m : matrix_M2L) f.write(reinterpret_cast<const char*>(m.data()), m.size() * sizeof(real_t));
  }
  bool load_matrix() {
    if (filename.empty()) return false;
    std::ifstream f(filename, std::ios::binary);
    if (!f.good()) return false;
    f.seekg(0, std::ios::end);
    if (size_t(f.tellg()) != matrix_file_size()) return false;
    f.seekg(0, std::ios::beg);
    real_t r;
    f.read(reinterpret_cast<char*>(&r), sizeof(real_t));
    if (r != r0) return false;
    const size_t n2 = size_t(nsurf) * nsurf;
    for (int l = 0; l < op_levels(); l++) {
      for (auto* m : {&matrix_UC2E_U[l], &matrix_UC2E_V[l], &matrix_DC2E_U[l], &matrix_DC2E_V[l]}) {
        m->resize(n2);
        f.read(reinterpret_cast<char*>(m->data()), n2 * sizeof(T));
      }
      for (auto& m : matrix_M2M[l]) { m.resize(n2); f.read(reinterpret_cast<char*>(m.data()), n2 * sizeof(T)); }
      for (auto& m : matrix_L2L[l]) { m.resize(n2); f.read(reinterpret_cast<char*>(m.data()), n2 * sizeof(T)); }
    }
    for (auto& m : matrix_M2L) {
      m.resize(size_t(26) * nfreq * 128);
      f.read(reinterpret_cast<char*>(m.data()), m.size() * sizeof(real_t));
    }
    if (!f) return false;
    // The reference layout records neither the wave number nor the
    // truncation (fmm.h:103-138 accepts any file of the right size): check
    // that the loaded pseudo-inverses invert this kernel's check matrix at the
    // coarsest and finest operator level, else recompute.
    return pinv_consistent(0) && pinv_consistent(op_levels() - 1);
  }

  // A A+ A = A on a fixed probe vector for the upward check matrix A of
  // `level` (relative residual 1e-6; an operator set of another wave number
  // or kernel misses it by orders of magnitude).
  bool pinv_consistent(int level) const {
    real_t c[3] = {0, 0, 0};
    RealVec up_check = surface(p, r0, level, c, 2.95);
    RealVec up_equiv = surface(p, r0, level, c, 1.05);
    std::vector<T> A;
    kernel_matrix(up_check, up_equiv, A);
    std::vector<T> x(nsurf), y(nsurf, T(0)), t(nsurf, T(0)), z(nsurf, T(0)), w(nsurf, T(0));
    for (int i = 0; i < nsurf; i++) x[i] = T(std::cos(0.7 * i + 0.3));
    auto matvec = [&](const std::vector<T>& M, const std::vector<T>& v, std::vector<T>& out) {
      for (int i = 0; i < nsurf; i++) {
        T s = T(0);
        for (int j = 0; j < nsurf; j++) s += M[size_t(i) * nsurf + j] * v[j];
        out[i] = s;
      }
    };
    matvec(A, x, y);
    matvec(matrix_UC2E_U[level], y, t);
    matvec(matrix_UC2E_V[level], t, z);
    matvec(A, z, w);
    double num = 0, den = 0;
    for (int i = 0; i < nsurf; i++) {
      num += std::norm(w[i] - y[i]);
      den += std::norm(y[i]);
    }
    return den > 0 && std::sqrt(num / den) < 1e-6;
  }

  // Truncation of the check-to-equivalent pseudo-inverses: EPS (1e-16) for
  // the FP64 path, 1e-8 (the EPS of the reference's -DFLOAT build,
  // exafmm_t.h:23) for the modes whose M2L runs in FP32.  Those modes feed
  // FP32-rounded check potentials into the downward solve; the singular
  // directions below 4e-8 max(S) only amplify that rounding (at p = 10 they
  // carry no signal the FP32 data resolves), so dropping them is a
  // regularisation of the FP64 solve, not a lower arithmetic precision.
  real_t pinv_eps() const { return precision == XFMM_FP64 ? EPS : real_t(1e-8); }

  // precompute(): needs r0 (get_bounds) and, for level-dependent kernels, depth.
  // The cache file holds operators of the default truncation only (the FLOAT
  // mode's operators are computed, not cached).
  void precompute() {
    init_rel_coord();
    initialize_matrix();
    const bool cacheable = pinv_eps() == EPS;
    is_precomputed = cacheable && load_matrix();
    if (!is_precomputed) {
      for (int l = 0; l < op_levels(); l++) precompute_check2equiv(l);
      for (int l = 0; l < op_levels(); l++) precompute_M2M(l);
      if (scale_invariant()) precompute_M2L(0, 0);
      else
        for (int l = 1; l <= depth; l++) precompute_M2L(l - 1, l);
      if (cacheable) save_matrix();
      is_precomputed = true;
    }
    ops_uploaded = false;
  }

  // M2L_setup (fmm_scale_invariant.h:277-341): the M2L pairs are taken from
  // Node::M2L_list when the tree is flattened; this records the call only.
  void M2L_setup(NodePtrs<T>& nonleafs) {
    (void)nonleafs;
    tree_sig = {};
    flat_sig = {};
  }

  // ------------------------------------------------------------ device glue
  xfmm_context* context() {
    if (!ctx) {
      const complex_t k = kernel_wavek();
      xfmm_check(xfmm_create(kernel_kind(), p, k.real(), k.imag(), precision, device, &ctx), nullptr, "xfmm_create");
      ctx_precision = precision;
      ctx_wavek = k;
    }
    return ctx;
  }
  void release_device() {
    if (ctx) xfmm_destroy(ctx);
    ctx = nullptr;
    ops_uploaded = false;
    tree_sig = {};
  }
  void upload_operators() {
    if (!is_precomputed) throw std::runtime_error("FMM: precompute() must run before evaluation");
    const int L = op_levels(), S = m2l_slots();
    const size_t n2 = size_t(nsurf) * nsurf;
    std::vector<T> uu, uv, du, dv, mm, ll;
    RealVec m2l;
    for (int l = 0; l < L; l++) {
      uu.insert(uu.end(), matrix_UC2E_U[l].begin(), matrix_UC2E_U[l].end());
      uv.insert(uv.end(), matrix_UC2E_V[l].begin(), matrix_UC2E_V[l].end());
      du.insert(du.end(), matrix_DC2E_U[l].begin(), matrix_DC2E_U[l].end());
      dv.insert(dv.end(), matrix_DC2E_V[l].begin(), matrix_DC2E_V[l].end());
      for (int o = 0; o < 8; o++) mm.insert(mm.end(), matrix_M2M[l][o].begin(), matrix_M2M[l][o].end());
      for (int o = 0; o < 8; o++) ll.insert(ll.end(), matrix_L2L[l][o].begin(), matrix_L2L[l][o].end());
    }
    (void)n2;
    for (int s = 0; s < S; s++) m2l.insert(m2l.end(), matrix_M2L[s].begin(), matrix_M2L[s].end());
    xfmm_check(xfmm_set_operators(context(), L, dbl(uu), dbl(uv), dbl(du), dbl(dv), dbl(mm), dbl(ll), S, m2l.data()),
               ctx, "xfmm_set_operators");
    ops_uploaded = true;
  }

  // Flatten the host tree (host only) when it changed.
  void ensure_flat(Nodes<T>& nodes) {
    TreeSig sig{nodes.data(), nodes.size(), r0, depth};
    if (sig == flat_sig) return;
    flat = flatten(nodes, depth, r0, x0);
    flat_sig = sig;
  }

  // Flatten + upload the tree when it changed since the last evaluation.
  void sync_tree(Nodes<T>& nodes) {
    // the device context holds the precision and the wave number it was
    // created with (the reference reads wavek on every call)
    if (ctx && (ctx_precision != precision || ctx_wavek != kernel_wavek())) release_device();
    if (!ops_uploaded) upload_operators();
    ensure_flat(nodes);
    if (flat_sig == tree_sig) return;
    xfmm_tree v = flat.view();
    xfmm_check(xfmm_set_tree(context(), &v), ctx, "xfmm_set_tree");
    tree_sig = flat_sig;
  }

  void upload_charges() {
    std::vector<T> q;
    q.reserve(flat.src_ibody.size());
    for (Node<T>* n : flat.leaf_order) q.insert(q.end(), n->src_value.begin(), n->src_value.end());
    if (q.size() != flat.src_ibody.size()) throw std::runtime_error("FMM: leaf charges do not match the tree");
    xfmm_check(xfmm_set_charges_host(ctx, dbl(q)), ctx, "xfmm_set_charges");
  }

  // FmmBase::upward_pass (fmm_base.h:162-171): P2M + M2M on the device.
  void upward_pass(Nodes<T>& nodes, NodePtrs<T>& leafs, bool verbose = true) {
    (void)leafs;
    sync_tree(nodes);
    upload_charges();
    xfmm_check(xfmm_upward_pass(ctx), ctx, "xfmm_upward_pass");
    if (verbose) print_timings(true);
  }

  // FmmBase::downward_pass (fmm_base.h:179-200): P2L, M2P, P2P, M2L, L2L, L2P
  // on the device; results are added to Node::trg_value (+=, as the reference).
  void downward_pass(Nodes<T>& nodes, NodePtrs<T>& leafs, bool verbose = true) {
    (void)leafs;
    sync_tree(nodes);
    xfmm_check(xfmm_downward_pass(ctx), ctx, "xfmm_downward_pass");
    std::vector<T> v(flat.trg_ibody.size() * 4);
    xfmm_check(xfmm_get_values(ctx, dbl(v)), ctx, "xfmm_get_values");
    size_t o = 0;
    for (Node<T>* n : flat.leaf_order) {
      for (size_t j = 0; j < n->trg_value.size(); j++) n->trg_value[j] += v[o + j];
      o += n->trg_value.size();
    }
    if (verbose) print_timings(false);
  }

  // Copy the device up/dn vectors into Node::up_equiv / dn_equiv.
  void sync_equiv_to_host(Nodes<T>& nodes) {
    if (!ctx) return;
    const size_t nn = flat.dev2node.size();
    std::vector<T> up(nn * nsurf), dn(nn * nsurf);
    xfmm_check(xfmm_get_equiv(ctx, dbl(up), dbl(dn)), ctx, "xfmm_get_equiv");
    for (size_t i = 0; i < nn; i++) {
      Node<T>* n = flat.dev2node[i];
      n->up_equiv.assign(up.begin() + i * nsurf, up.begin() + (i + 1) * nsurf);
      n->dn_equiv.assign(dn.begin() + i * nsurf, dn.begin() + (i + 1) * nsurf);
    }
    (void)nodes;
  }

  void print_timings(bool upward) const {
    double ms[11];
    const int n = xfmm_timings(ctx, ms, 11);
    static const char* names[11] = {"P2M", "M2M", "P2L", "M2P", "P2P", "M2L", "L2L", "L2P", "  FFT", "  Hadamard", "  IFFT"};
    for (int i = 0; i < n; i++) {
      if (upward != (i < 2)) continue;
      std::printf("%-20s : %.6e s\n", names[i], ms[i] * 1e-3);
    }
  }

  // Direct sums on the device (laplace.h:32-164 and the Helmholtz /
  // modified-Helmholtz counterparts): trg_value += sum over sources.
  void potential_P2P(RealVec& src_coord, std::vector<T>& src_value, RealVec& trg_coord, std::vector<T>& trg_value) {
    direct(src_coord, src_value, trg_coord, trg_value, 0);
  }
  void gradient_P2P(RealVec& src_coord, std::vector<T>& src_value, RealVec& trg_coord, std::vector<T>& trg_value) {
    direct(src_coord, src_value, trg_coord, trg_value, 1);
  }
  void direct(const RealVec& src_coord, const std::vector<T>& src_value, const RealVec& trg_coord,
              std::vector<T>& trg_value, int gradient) {
    const int ns = int(src_coord.size() / 3), nt = int(trg_coord.size() / 3);
    if (int(trg_value.size()) < nt * (gradient ? 4 : 1)) throw std::runtime_error("P2P: trg_value too small");
    const complex_t k = kernel_wavek();
    xfmm_check(xfmm_direct(kernel_kind(), precision, k.real(), k.imag(), src_coord.data(), dbl(src_value), ns,
                           trg_coord.data(), nt, dbl(trg_value), gradient),
               nullptr, "xfmm_direct");
  }

  // verify (fmm_base.h:208-267): FP64 direct sum over all sources for the
  // targets of 10 strided leaves (or all leaves), relative L2 errors of the
  // potential and the gradient against Node::trg_value.
  RealVec verify(NodePtrs<T>& leafs, bool sample = false) {
    std::vector<Node<T>*> targets;
    if (sample) {
      const int ns = 10;
      const size_t stride = leafs.size() / ns;
      for (int i = 0; i < ns && i * stride < leafs.size(); i++) targets.push_back(leafs[i * stride]);
    } else {
      targets.assign(leafs.begin(), leafs.end());
    }
    RealVec src;
    std::vector<T> q;
    for (Node<T>* l : leafs) {
      src.insert(src.end(), l->src_coord.begin(), l->src_coord.end());
      q.insert(q.end(), l->src_value.begin(), l->src_value.end());
    }
    RealVec trg;
    for (Node<T>* t : targets) trg.insert(trg.end(), t->trg_coord.begin(), t->trg_coord.end());
    std::vector<T> ref(trg.size() / 3 * 4, T(0));
    const int save = precision;
    precision = XFMM_FP64;
    direct(src, q, trg, ref, 1);
    precision = save;
    double pd = 0, pn = 0, gd = 0, gn = 0;
    size_t o = 0;
    for (Node<T>* t : targets)
      for (int j = 0; j < int(t->trg_coord.size() / 3); j++, o++) {
        pn += std::norm(ref[4 * o]);
        pd += std::norm(ref[4 * o] - t->trg_value[4 * j]);
        for (int d = 1; d < 4; d++) {
          gn += std::norm(ref[4 * o + d]);
          gd += std::norm(ref[4 * o + d] - t->trg_value[4 * j + d]);
        }
      }
    return RealVec{std::sqrt(pd / pn), std::sqrt(gd / gn)};
  }

  const FlatTree<T>& flat_tree() const { return flat; }

 protected:
  struct TreeSig {
    const void* data = nullptr;
    size_t n = 0;
    real_t r0 = 0;
    int depth = -1;
    bool operator==(const TreeSig& o) const { return data == o.data && n == o.n && r0 == o.r0 && depth == o.depth; }
  };
  static double* dbl(std::vector<T>& v) { return reinterpret_cast<double*>(v.data()); }
  static const double* dbl(const std::vector<T>& v) { return reinterpret_cast<const double*>(v.data()); }

  xfmm_context* ctx = nullptr;
  int ctx_precision = XFMM_FP64;
  complex_t ctx_wavek = complex_t(0, 0);
  bool ops_uploaded = false;
  TreeSig tree_sig, flat_sig;
  FlatTree<T> flat;
};

}  // namespace exafmm_t
