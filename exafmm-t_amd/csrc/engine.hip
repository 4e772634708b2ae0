// C-ABI implementation of the gfx950 FMM evaluation engine
// (include/exafmm_t_amd.h).  Owns the device copy of the flattened tree, the
// operators and all evaluation buffers of one FMM object, and drives the
// kernels of p2p.hip / gemm.hip / m2l.hip on one HIP stream.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <exception>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/exafmm_t_amd.h"
#include "kernels.h"

using namespace xfmm;

namespace {

thread_local std::string g_thread_error;

struct HipError {
  std::string msg;
};

#define HIPCHK(expr)                                                                 \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess) throw HipError{std::string(#expr) + ": " + hipGetErrorString(e_)}; \
  } while (0)

struct ArgError {
  std::string msg;
};

template <typename T>
struct DevBuf {
  T* ptr = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    hipError_t e = hipMalloc(&ptr, count * sizeof(T));
    if (e != hipSuccess) {
      ptr = nullptr;
      throw HipError{std::string("hipMalloc(") + std::to_string(count * sizeof(T)) + " bytes): " + hipGetErrorString(e)};
    }
    n = count;
  }
  void upload(const T* host, size_t count, hipStream_t s) {
    alloc(count);
    if (count) HIPCHK(hipMemcpyAsync(ptr, host, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

}  // namespace

struct xfmm_context {
  int kind = 0, p = 0, prec = 0, device = 0;
  // stages that run in FP32 (mixed path only): near-field P2P, M2L
  // (transforms + Hadamard), and the surface evaluations P2M/L2P/M2P/P2L
  bool f32_p2p = false, f32_m2l = false, f32_surf = false, f32_m2l_fft = false, f32_far = false;
  bool f16_had = false;  // split-FP16 Hadamard (needs f32_m2l)
  // FP32 MFMA GEMMs: the pseudo-inverse solves of P2M / L2P (FLOAT mode only)
  // and the M2M / L2L translations (both need the float-truncated operators)
  bool f32_pinv = false, f32_trans = false;
  // mixed path: the current charges' net sum exceeds a tenth of their
  // absolute sum (one-signed charges).  Their potentials are then large
  // against their gradients, and the FP32 rounding of M2M / L2L and of the
  // transforms' arithmetic, relative to the potential, is not small against
  // the gradient: those stages run in FP64 while this holds.
  bool net_heavy = false;
  bool trans32() const { return f32_trans && !net_heavy; }
  bool fft32() const { return f32_m2l_fft && !net_heavy; }
  // dn and trg_value are zero (cleared by the last upward pass or clear_values
  // and not written since), so the downward pass need not clear them itself
  bool down_clean = false;
  double wk_re = 0, wk_im = 0;
  int nsurf = 0, nconv = 0, nfreq = 0, n1 = 0, vw = 1;
  bool cplx = false;
  hipStream_t stream = nullptr;
  // Second stream: the near-field P2P of an evaluation runs there, into
  // `near`, concurrently with the upward pass and M2L (it needs only the
  // charges), and is added to the target values at the end of the downward
  // pass.  Every upward pass starts a new one; the next downward pass
  // consumes it.
  hipStream_t stream2 = nullptr;
  bool near_pending = false;   // `near` is being filled with P2P of the current charges
  bool near_launched = false;  // a P2P was ever queued on stream2 (ev[20] valid)
  bool near_overlapped = false;  // the last downward pass took P2P from stream2
  // graph of one evaluation (xfmm_graph_capture); timing events are not
  // recorded while capturing
  bool capturing = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  bool graph_stale = false;  // captured with the kernels of the other charge class (net_heavy)
  std::string err;

  // ---- operators
  bool have_ops = false;
  int nlevels = 0, nm2l = 0;
  DevBuf<double> uc2e_uT, uc2e_vT, dc2e_uT, dc2e_vT, m2mcat, l2lcat;
  // numerical rank of the upward / downward pseudo-inverse per operator level
  // (uc2e_uT / dc2e_uT hold only the first r_up / r_dn columns, row stride r)
  std::vector<int> r_up, r_dn;
  DevBuf<float> uc2e_uT32, uc2e_vT32, dc2e_uT32, dc2e_vT32, m2mcat32, l2lcat32;  // f32_pinv / f32_trans
  DevBuf<double> m2l64;  // M2L operators in the expanded MFMA form, [L][nfreq][26][64][4]
  DevBuf<float> m2l32;   // FP32 form, or the split-FP16 form of the same size when f16_had
  DevBuf<float> m2l_hscale;  // f16_had: operator scale per [L][nfreq]
  DevBuf<double> m2l_amax;   // f16_had: largest |spectrum value| of the evaluation (float bits)
  DevBuf<double> unit_surf;
  DevBuf<int> map_up, map_dn, inv_dn;

  // ---- tree
  bool have_tree = false;
  int nnodes = 0, depth = 0, nsrcs = 0, ntrgs = 0;
  double r0 = 0;
  std::vector<int> h_level_offset, h_level, h_child0;
  DevBuf<double> center, src, trg, q;
  DevBuf<int> level, child0, src_begin, src_count, trg_begin, trg_count;
  DevBuf<int> p2p_off, p2p_src, m2p_off, m2p_src, p2l_off, p2l_src;
  // work lists
  DevBuf<int> w_p2p, w_m2p, w_p2l, w_l2p_eval, w_leaf;  // w_leaf: P2M/L2P conversion leaves, sorted by level
  int n_p2p = 0, n_m2p = 0, n_p2l = 0, n_l2p_eval = 0, n_leaf = 0;
  DevBuf<int> t_p2p, t_l2p;  // (leaf, 64-target tile) pairs of w_p2p / w_l2p_eval
  int nt_p2p = 0, nt_l2p = 0;
  std::vector<int> leaf_level_off;                        // [depth+2] ranges in w_leaf
  DevBuf<long> leaf_slot_rows, leaf_node_rows;            // i*nsurf, node*nsurf
  std::vector<int> nl_level_off;                          // [depth+2] ranges of the non-leaf list
  DevBuf<long> nl_rows, nl_child_rows;                    // node*nsurf, child0*nsurf
  // M2L
  int n_m2l_src = 0, n_m2l_trg = 0, n_m2l_blk = 0;
  DevBuf<int> m2l_src_node, m2l_trg_node, m2l_blk_trg, m2l_blk_uni, m2l_blk_nu;
  DevBuf<short> m2l_blk_loc;
  DevBuf<double> m2l_scale;
  std::vector<int> blk_level_off;                         // [depth+2] block ranges per target level
  DevBuf<char> fft_in, fft_out;
  // buffers
  DevBuf<double> up, dn, trg_value, check, tmp, gemm_scratch, near;
  // near-coincident P2P pairs ([4] ints each) corrected in FP64 after an
  // FP32 near field
  DevBuf<int> close_pairs, close_off;
  DevBuf<double> close_coef;
  DevBuf<double> charge_sums;  // partial sums of the charge balance rule (device charges)
  int nclose = 0;  // targets with near-coincident pairs
  DevBuf<double> src_rel, src_pack, surf_tab;
  DevBuf<float> surf_tab32;  // P2P sources relative to their leaf centre; packed (x, y, z, q)

  // per-target output factors (phi, gradient) applied at the end of each
  // downward pass (xfmm_set_value_scale); empty = none
  DevBuf<double> vscale;
  // tuning (xfmm_set_tuning / xfmm_autotune)
  int tune_kch = HAD_KCH;  // Hadamard frequencies per workgroup
  int tune_groups = 2;     // split-FP16 Hadamard: 64-target blocks per workgroup
  int tune_tiles = 1;      // split-FP16 Hadamard: 16-target tiles per wave
  int tune_split = 32;     // real-kernel P2P / L2P: largest tail tile whose sources are split over lanes

  // ---- timing
  hipEvent_t ev[24] = {};
  bool ev_ok = false;
  double t_ms[11] = {0};

  // scale of the P2M / L2P conversion of a level-l leaf: 0.5^l for the
  // scale-invariant Laplace operators (fmm_scale_invariant.h:176-193, 211-219), else 1
  double leaf_scale(int l) const { return kind == XFMM_LAPLACE ? std::ldexp(1.0, -l) : 1.0; }
  DevTree dev_tree() const {
    DevTree t;
    t.nnodes = nnodes;
    t.depth = depth;
    t.p = p;
    t.nsurf = nsurf;
    t.vw = vw;
    t.r0 = r0;
    t.center = center.ptr;
    t.level = level.ptr;
    t.child0 = child0.ptr;
    t.src_begin = src_begin.ptr;
    t.src_count = src_count.ptr;
    t.trg_begin = trg_begin.ptr;
    t.trg_count = trg_count.ptr;
    t.src = src.ptr;
    t.q = q.ptr;
    t.trg = trg.ptr;
    t.p2p_off = p2p_off.ptr;
    t.p2p_src = p2p_src.ptr;
    t.m2p_off = m2p_off.ptr;
    t.m2p_src = m2p_src.ptr;
    t.p2l_off = p2l_off.ptr;
    t.p2l_src = p2l_src.ptr;
    t.unit_surf = unit_surf.ptr;
    t.nsrcs = nsrcs;
    t.src_rel = src_rel.ptr;
    t.surf_tab = surf_tab.ptr;
    t.surf_tab32 = surf_tab32.ptr;
    t.src_pack = src_pack.ptr;
    t.up = up.ptr;
    t.dn = dn.ptr;
    t.trg_value = trg_value.ptr;
    t.p2p_split = tune_split;
    t.close = close_pairs.ptr;
    t.close_off = close_off.ptr;
    t.close_coef = close_coef.ptr;
    t.nclose = nclose;
    return t;
  }
  M2LDev m2l_dev() const {
    M2LDev m;
    m.p = p;
    m.n1 = n1;
    m.nfreq = nfreq;
    m.nsurf = nsurf;
    m.cplx = cplx;
    m.mixed = f32_m2l;
    m.mixed_fft = fft32();
    m.map_up = map_up.ptr;
    m.map_dn = map_dn.ptr;
    m.inv_dn = inv_dn.ptr;
    m.fft_in = fft_in.ptr;
    m.fft_out = fft_out.ptr;
    m.amax = f16_had ? reinterpret_cast<unsigned*>(m2l_amax.ptr) : nullptr;
    return m;
  }
  Wavek wavek() const { return Wavek{wk_re, wk_im}; }
  int op_level(int l) const { return nlevels == 1 ? 0 : l; }
};

namespace {

// Unit-cube surface in the reference point order (geometry.h:21-63).
std::vector<double> unit_surface(int p) {
  const int n = 6 * (p - 1) * (p - 1) + 2;
  std::vector<double> u(3 * n);
  auto co = [&](int m) { return (2.0 * m - p + 1) / (p - 1); };
  int c = 0;
  auto put = [&](double a, double b, double d) { u[3 * c] = a; u[3 * c + 1] = b; u[3 * c + 2] = d; c++; };
  put(-1, -1, -1);
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(-1, co(i + 1), co(j));
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(co(i), -1, co(j + 1));
  for (int i = 0; i < p - 1; i++)
    for (int j = 0; j < p - 1; j++) put(co(i + 1), co(j), -1);
  for (int i = 0; i < 3 * c; i++) u[3 * c + i] = -u[i];
  return u;
}

// surf2conv maps (geometry.h:101-134) as local indices of the p^3 corner cube
// that holds the data: upward (p-1-x, ...) in [0,p), downward (2p-1-x, ...)
// in [p,2p) shifted by -p.
std::vector<int> local_map(int p, bool up) {
  const double half = 0.5 * (p - 1);
  std::vector<double> u = unit_surface(p);
  const int n = int(u.size() / 3);
  std::vector<int> m(n);
  for (int i = 0; i < n; i++) {
    int idx[3];
    for (int d = 0; d < 3; d++) {
      const double s = u[3 * i + d] * (double(p - 1) * 0.5) + half;  // lattice coordinate in [0, p-1]
      idx[d] = up ? int((p - 1) - s) : int((2 * p - 1) - s) - p;
    }
    m[i] = idx[0] + p * idx[1] + p * p * idx[2];
  }
  return m;
}

int fail(xfmm_context* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  else g_thread_error = msg;
  return code;
}

template <typename F>
int guarded(xfmm_context* ctx, F&& f) {
  if (!ctx) return fail(nullptr, XFMM_ERR_ARG, "null context");
  try {
    HIPCHK(hipSetDevice(ctx->device));
    f();
    return XFMM_OK;
  } catch (const ArgError& e) {
    return fail(ctx, XFMM_ERR_ARG, e.msg);
  } catch (const HipError& e) {
    const bool oom = e.msg.find("hipMalloc") != std::string::npos;
    return fail(ctx, oom ? XFMM_ERR_NOMEM : XFMM_ERR_HIP, e.msg);
  } catch (const std::bad_alloc&) {
    return fail(ctx, XFMM_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(ctx, XFMM_ERR_ARG, e.what());
  }
}

void require(bool c, const std::string& msg) {
  if (!c) throw ArgError{msg};
}

// A captured graph holds the device addresses of the buffers it was captured
// with: drop it before anything is reallocated.
void drop_graph(xfmm_context* c) {
  if (!c->graph_exec && !c->graph) return;
  HIPCHK(hipStreamSynchronize(c->stream));
  if (c->graph_exec) HIPCHK(hipGraphExecDestroy(c->graph_exec));
  if (c->graph) HIPCHK(hipGraphDestroy(c->graph));
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_stale = false;
}

// ---------------------------------------------------------------- operators

void op_P2M(xfmm_context* c);
void op_M2M(xfmm_context* c);
void op_P2L(xfmm_context* c);
void op_M2P(xfmm_context* c);
void op_P2P(xfmm_context* c);
void op_L2L(xfmm_context* c);
void op_L2P(xfmm_context* c, const double* near_add = nullptr, bool wait_near = false);

void check_ready(xfmm_context* c) {
  if (!c->have_tree) throw ArgError{"xfmm: tree not set (call xfmm_set_tree)"};
  if (!c->have_ops) throw ArgError{"xfmm: operators not set (call xfmm_set_operators)"};
  if (c->nlevels != 1 && c->nlevels < c->depth + 1)
    throw ArgError{"xfmm: level-dependent operators cover fewer levels than the tree depth"};
  if (c->nm2l != 1 && c->nm2l < c->depth) throw ArgError{"xfmm: M2L operators cover fewer levels than the tree depth"};
}

// B operand of a translation GEMM: the FP32 copy for FP32 GEMMs, else the
// FP64 operator; off = element offset.
const void* gemm_op(bool f32, const xfmm_context* c, const DevBuf<double>& d, const DevBuf<float>& f, size_t off) {
  (void)c;
  if (f32) return f.ptr + off;
  return d.ptr + off;
}

// Doubles per nsurf x nsurf operator in GEMM form: complex operators are held
// as real 2 x 2 blocks [[Re, Im], [-Im, Re]] per entry, so that a complex
// row-GEMM on interleaved (re, im) rows is one real GEMM with K and N doubled
// (on the FP64 matrix cores like the real kernels).
size_t gemm_op_size(const xfmm_context* c) { return size_t(c->nsurf) * c->nsurf * c->vw * c->vw; }

void op_P2M(xfmm_context* c) {
  // check potentials of every leaf, then up = alpha * UC2E_V (UC2E_U check)
  DevTree t = c->dev_tree();
  launch_eval_surface(c->kind, c->f32_surf, MODE_P2M, t, c->wavek(), c->w_leaf.ptr, c->n_leaf, c->check.ptr,
                      c->stream);
  const size_t ns2 = gemm_op_size(c);
  const int vw = c->vw;
  for (int l = 0; l <= c->depth; l++) {
    const int b = c->leaf_level_off[l], e = c->leaf_level_off[l + 1];
    if (e <= b) continue;
    const int L = c->op_level(l);
    // only the first r singular directions carry a nonzero S^+ entry: the
    // first product computes those r coefficients, the second sums over them
    const int r = c->r_up[L];
    launch_gemm_rows(c->f32_pinv, e - b, r * vw, c->nsurf * vw, c->check.ptr, c->leaf_slot_rows.ptr + b,
                     gemm_op(c->f32_pinv, c, c->uc2e_uT, c->uc2e_uT32, L * ns2), c->tmp.ptr, c->leaf_slot_rows.ptr + b,
                     1.0, 0, c->stream);
    launch_gemm_rows(c->f32_pinv, e - b, c->nsurf * vw, r * vw, c->tmp.ptr, c->leaf_slot_rows.ptr + b,
                     gemm_op(c->f32_pinv, c, c->uc2e_vT, c->uc2e_vT32, L * ns2), c->up.ptr, c->leaf_node_rows.ptr + b,
                     c->leaf_scale(l), 0, c->stream);
  }
}

void op_M2M(xfmm_context* c) {
  const size_t blk = 8 * gemm_op_size(c);
  const int vw = c->vw;
  for (int l = c->depth - 1; l >= 0; l--) {
    const int b = c->nl_level_off[l], e = c->nl_level_off[l + 1];
    if (e <= b) continue;
    const int L = c->op_level(l);
    // parent row = [child_0 .. child_7] (contiguous, K = 8 nsurf) x M2Mcat
    launch_gemm_rows(c->trans32(), e - b, c->nsurf * vw, 8 * c->nsurf * vw, c->up.ptr, c->nl_child_rows.ptr + b,
                     gemm_op(c->trans32(), c, c->m2mcat, c->m2mcat32, L * blk), c->up.ptr, c->nl_rows.ptr + b, 1.0, 1,
                     c->stream, c->gemm_scratch.ptr, c->gemm_scratch.n);
  }
}

void op_L2L(xfmm_context* c) {
  const size_t blk = 8 * gemm_op_size(c);
  const int vw = c->vw;
  for (int l = 0; l < c->depth; l++) {
    const int b = c->nl_level_off[l], e = c->nl_level_off[l + 1];
    if (e <= b) continue;
    const int L = c->op_level(l);
    launch_gemm_rows(c->trans32(), e - b, 8 * c->nsurf * vw, c->nsurf * vw, c->dn.ptr, c->nl_rows.ptr + b,
                     gemm_op(c->trans32(), c, c->l2lcat, c->l2lcat32, L * blk), c->dn.ptr, c->nl_child_rows.ptr + b,
                     1.0, 1, c->stream, c->gemm_scratch.ptr, c->gemm_scratch.n);
  }
}

// near_add (may be null): per-target values added to the L2P result (the
// near field of a full evaluation); wait_near: they come from the second
// stream, joined before the L2P kernel.
void op_L2P(xfmm_context* c, const double* near_add, bool wait_near) {
  const size_t ns2 = gemm_op_size(c);
  const int vw = c->vw;
  for (int l = 0; l <= c->depth; l++) {
    const int b = c->leaf_level_off[l], e = c->leaf_level_off[l + 1];
    if (e <= b) continue;
    const int L = c->op_level(l);
    const int r = c->r_dn[L];
    launch_gemm_rows(c->f32_pinv, e - b, r * vw, c->nsurf * vw, c->dn.ptr, c->leaf_node_rows.ptr + b,
                     gemm_op(c->f32_pinv, c, c->dc2e_uT, c->dc2e_uT32, L * ns2), c->tmp.ptr, c->leaf_slot_rows.ptr + b,
                     1.0, 0, c->stream);
    launch_gemm_rows(c->f32_pinv, e - b, c->nsurf * vw, r * vw, c->tmp.ptr, c->leaf_slot_rows.ptr + b,
                     gemm_op(c->f32_pinv, c, c->dc2e_vT, c->dc2e_vT32, L * ns2), c->dn.ptr, c->leaf_node_rows.ptr + b,
                     c->leaf_scale(l), 0, c->stream);
  }
  DevTree t = c->dev_tree();
  if (wait_near) HIPCHK(hipStreamWaitEvent(c->stream, c->ev[20], 0));
  t.near_add = near_add;
  launch_eval_targets(c->kind, c->f32_far, MODE_L2P, t, c->wavek(), c->w_l2p_eval.ptr, c->n_l2p_eval, c->t_l2p.ptr,
                      c->nt_l2p, c->stream);
}

void op_P2L(xfmm_context* c) {
  DevTree t = c->dev_tree();
  launch_eval_surface(c->kind, c->f32_surf, MODE_P2L, t, c->wavek(), c->w_p2l.ptr, c->n_p2l, nullptr,
                      c->stream);
}

void op_M2P(xfmm_context* c) {
  DevTree t = c->dev_tree();
  launch_eval_targets(c->kind, c->f32_far, MODE_M2P, t, c->wavek(), c->w_m2p.ptr, c->n_m2p, nullptr, 0, c->stream);
}

// The near-coincident pairs of the P2P lists (FP32 near field only), found
// once per tree: a first pass with room for ntrgs / 8 pairs, a second with
// the exact count when that was too small; then sorted by (target, source)
// and grouped per target on the host, so that the correction is applied in a
// fixed order, and each pair's correction per unit charge computed once.
void find_close_pairs(xfmm_context* c) {
  c->close_pairs.release();
  c->close_off.release();
  c->close_coef.release();
  c->nclose = 0;
  if (!c->f32_p2p || c->n_p2p <= 0) return;
  DevBuf<int> count, found;
  count.alloc(1);
  int cap = std::max(1024, c->ntrgs / 8), n = 0;
  for (int pass = 0; pass < 2; pass++) {
    found.alloc(size_t(cap) * 4);
    HIPCHK(hipMemsetAsync(count.ptr, 0, sizeof(int), c->stream));
    launch_close_find(c->dev_tree(), c->w_p2p.ptr, c->n_p2p, found.ptr, cap, count.ptr, c->stream);
    HIPCHK(hipMemcpyAsync(&n, count.ptr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (n <= cap) break;
    cap = n;
  }
  n = std::min(n, cap);
  if (n == 0) return;
  std::vector<int> h(size_t(n) * 4);
  HIPCHK(hipMemcpy(h.data(), found.ptr, h.size() * sizeof(int), hipMemcpyDeviceToHost));
  std::vector<int> order(n);
  for (int i = 0; i < n; i++) order[i] = i;
  std::sort(order.begin(), order.end(), [&](int a, int b) {
    return h[4 * size_t(a)] != h[4 * size_t(b)] ? h[4 * size_t(a)] < h[4 * size_t(b)]
                                                : h[4 * size_t(a) + 1] < h[4 * size_t(b) + 1];
  });
  std::vector<int> sorted(size_t(n) * 4), off;
  for (int i = 0; i < n; i++) {
    for (int d = 0; d < 4; d++) sorted[4 * size_t(i) + d] = h[4 * size_t(order[i]) + d];
    if (i == 0 || sorted[4 * size_t(i)] != sorted[4 * size_t(i - 1)]) off.push_back(i);
  }
  off.push_back(n);
  c->close_pairs.upload(sorted, c->stream);
  c->close_off.upload(off, c->stream);
  c->close_coef.alloc(size_t(n) * 4 * c->vw);
  c->nclose = int(off.size()) - 1;
  launch_close_coef(c->kind, c->dev_tree(), c->wavek(), n, c->close_coef.ptr, c->stream);
  HIPCHK(hipStreamSynchronize(c->stream));
}

void op_P2P(xfmm_context* c) {
  DevTree t = c->dev_tree();
  launch_eval_targets(c->kind, c->f32_p2p, MODE_P2P, t, c->wavek(), c->w_p2p.ptr, c->n_p2p, c->t_p2p.ptr,
                      c->nt_p2p, c->stream);
}

void record(xfmm_context* c, int i) {
  if (c->ev_ok && !c->capturing) HIPCHK(hipEventRecord(c->ev[i], c->stream));
}

// Queue the near-field P2P of the current charges on the second stream, after
// everything already queued on the main stream.
void launch_near(xfmm_context* c) {
  if (!c->ev_ok || c->near_pending) return;
  HIPCHK(hipEventRecord(c->ev[23], c->stream));
  HIPCHK(hipStreamWaitEvent(c->stream2, c->ev[23], 0));
  launch_zero(c->near.ptr, c->near.n, c->stream2);
  if (!c->capturing) HIPCHK(hipEventRecord(c->ev[21], c->stream2));
  DevTree t = c->dev_tree();
  t.trg_value = c->near.ptr;
  launch_eval_targets(c->kind, c->f32_p2p, MODE_P2P, t, c->wavek(), c->w_p2p.ptr, c->n_p2p, c->t_p2p.ptr,
                      c->nt_p2p, c->stream2);
  if (!c->capturing) HIPCHK(hipEventRecord(c->ev[22], c->stream2));
  HIPCHK(hipEventRecord(c->ev[20], c->stream2));
  c->near_pending = true;
  c->near_launched = true;
}

// hadamard_product of every operator level (after the forward transforms)
void run_hadamards(xfmm_context* c, const M2LDev& m) {
  const size_t opsz = size_t(NM2L) * c->nfreq * 64 * 4;
  // one operator set (Laplace): a single launch over the blocks of all levels
  const int nlaunch = (c->nm2l == 1) ? 1 : c->depth;
  for (int l = 0; l < nlaunch; l++) {
    const int b = (c->nm2l == 1) ? 0 : c->blk_level_off[l];
    const int e = (c->nm2l == 1) ? c->n_m2l_blk : c->blk_level_off[l + 1];
    if (e <= b) continue;
    const int L = (c->nm2l == 1) ? 0 : l;
    const void* ops = c->f32_m2l ? static_cast<const void*>(c->m2l32.ptr + L * opsz)
                                              : static_cast<const void*>(c->m2l64.ptr + L * opsz);
    HadBlocks hb;
    hb.nblk = e - b;
    hb.trg = c->m2l_blk_trg.ptr + size_t(b) * 64;
    hb.uni = c->m2l_blk_uni.ptr + size_t(b) * HAD_UCAP;
    hb.nunion = c->m2l_blk_nu.ptr + b;
    hb.loc = c->m2l_blk_loc.ptr + size_t(b) * NM2L * 64;
    launch_hadamard(m, ops, c->f16_had ? c->m2l_hscale.ptr + size_t(L) * c->nfreq : nullptr, hb, c->tune_kch,
                    c->tune_groups, c->tune_tiles, c->stream);
  }
}

void op_M2L(xfmm_context* c, bool timed) {
  if (c->n_m2l_trg == 0) return;
  M2LDev m = c->m2l_dev();
  DevTree t = c->dev_tree();
  if (timed) record(c, 16);
  if (m.amax) launch_zero(c->m2l_amax.ptr, 1, c->stream);
  launch_fft_up(m, t, c->m2l_src_node.ptr, c->n_m2l_src, c->stream);
  if (timed) record(c, 17);
  run_hadamards(c, m);
  if (timed) record(c, 18);
  launch_ifft_dn(m, t, c->m2l_trg_node.ptr, c->m2l_scale.ptr, c->n_m2l_trg, c->stream);
  if (timed) record(c, 19);
}

}  // namespace

// =================================================================== C ABI

// ---------------------------------------------------------- allocation-free direct sum
//
// Partial sums of xfmm_direct_noalloc live in this file-scope array, one copy
// per device, for the life of the process.  Host code never passes the name
// `scratch` to a kernel (in host code it is a host-side stand-in); it uses the
// device address that hipGetSymbolAddress returns for the current device.
static __device__ __align__(256) double scratch[XFMM_NOALLOC_SCRATCH_BYTES / sizeof(double)];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device (looked up once per device).
double* scratch_ptr(int dev) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) throw HipError{"hipGetSymbolAddress(scratch) returned a null address"};
    g_scratch_ptr[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_ptr[dev];
}
}  // namespace

extern "C" {

int xfmm_device_count(int* count) {
  if (!count) return fail(nullptr, XFMM_ERR_ARG, "xfmm_device_count: null output");
  *count = 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return XFMM_OK;
  *count = n;
  return XFMM_OK;
}

int xfmm_create(int kernel, int p, double wavek_re, double wavek_im, int precision, int device, xfmm_context** out) {
  if (!out) return fail(nullptr, XFMM_ERR_ARG, "xfmm_create: null output pointer");
  *out = nullptr;
  if (kernel < XFMM_LAPLACE || kernel > XFMM_MODIFIED_HELMHOLTZ)
    return fail(nullptr, XFMM_ERR_ARG, "xfmm_create: unknown kernel");
  if (p < 2 || p > 16) return fail(nullptr, XFMM_ERR_ARG, "xfmm_create: expansion order p must be in [2, 16]");
  if (precision != XFMM_FP64 && precision != XFMM_MIXED && precision != XFMM_FLOAT)
    return fail(nullptr, XFMM_ERR_ARG, "xfmm_create: unknown precision mode");
  xfmm_context* c = new (std::nothrow) xfmm_context;
  if (!c) return fail(nullptr, XFMM_ERR_NOMEM, "xfmm_create: host allocation failed");
  c->kind = kernel;
  c->p = p;
  c->prec = precision;
  if (precision == XFMM_MIXED) {
    // Mixed policy (the operators' pseudo-inverses truncated at 4e-8 max(S),
    // which the host layer precomputes for this mode): FP32 near field, FP32
    // M2L (transforms, spectra; the Hadamard on split FP16 = two FP16 numbers
    // per FP32 value, FP32 accumulation), FP32 check potentials (P2M, P2L),
    // FP32 L2P / M2P evaluation and FP32 M2M / L2L translations; the
    // check-to-equivalent pseudo-inverse solves of P2M and L2P stay FP64.
    // XFMM_MIXED_STAGES overrides the FP32 set with a comma list of p2p, m2l
    // (FP32 spectra + Hadamard), had16 (split-FP16 Hadamard, with m2l), fft
    // (FP32 transform arithmetic, else FP64), surf (check potentials), far
    // (L2P / M2P evaluation), trans (M2M / L2L) and pinv (the solves).
    const char* env = std::getenv("XFMM_MIXED_STAGES");
    const std::string st = env ? env : "p2p,m2l,had16,fft,surf,far,trans";
    c->f32_p2p = st.find("p2p") != std::string::npos;
    c->f32_m2l = st.find("m2l") != std::string::npos;
    c->f32_surf = st.find("surf") != std::string::npos;
    c->f32_far = st.find("far") != std::string::npos;
    c->f32_m2l_fft = c->f32_m2l && st.find("fft") != std::string::npos;
    c->f16_had = c->f32_m2l && st.find("had16") != std::string::npos;
    c->f32_trans = st.find("trans") != std::string::npos;
    c->f32_pinv = st.find("pinv") != std::string::npos;
  } else if (precision == XFMM_FLOAT) {
    c->f32_p2p = c->f32_m2l = c->f32_m2l_fft = c->f16_had = c->f32_surf = c->f32_far = true;
    c->f32_pinv = c->f32_trans = true;
  }
  c->device = device;
  c->wk_re = wavek_re;
  c->wk_im = wavek_im;
  c->cplx = kernel == XFMM_HELMHOLTZ;
  c->vw = c->cplx ? 2 : 1;
  c->n1 = 2 * p;
  c->nsurf = 6 * (p - 1) * (p - 1) + 2;
  c->nconv = c->n1 * c->n1 * c->n1;
  c->nfreq = c->cplx ? c->nconv : c->n1 * c->n1 * (c->n1 / 2 + 1);
  try {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw ArgError{"xfmm_create: device ordinal out of range"};
    HIPCHK(hipSetDevice(device));
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&c->stream2, hipStreamNonBlocking));
    bool ok = true;
    for (auto& e : c->ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    c->ev_ok = ok;
    std::vector<double> u = unit_surface(p);
    c->unit_surf.upload(u, c->stream);
    // the maps index the LDS cube of the transform kernels: must be a
    // one-to-one map into [0, p^3), onto the cube's surface points (the
    // forward transform stores only those, the inverse evaluates only those)
    for (int up = 0; up < 2; up++) {
      std::vector<int> m = local_map(p, up == 1);
      std::vector<char> seen(size_t(p) * p * p, 0);
      for (int v : m) {
        if (v < 0 || v >= p * p * p || seen[v]) throw ArgError{"xfmm_create: surface-to-grid map is not one-to-one"};
        const int x = v % p, y = (v / p) % p, z = v / (p * p);
        const bool face = x == 0 || x == p - 1 || y == 0 || y == p - 1 || z == 0 || z == p - 1;
        if (!face) throw ArgError{"xfmm_create: surface point inside the grid cube"};
        seen[v] = 1;
      }
      (up ? c->map_up : c->map_dn).upload(m, c->stream);
      if (!up) {
        std::vector<int> inv(size_t(p) * p * p, -1);
        for (size_t i = 0; i < m.size(); i++) inv[m[i]] = int(i);
        c->inv_dn.upload(inv, c->stream);
      }
    }
    M2LDev m = c->m2l_dev();
    const size_t lds = m2l_lds_bytes(m);
    if (lds > 160 * 1024) throw ArgError{"xfmm_create: p too large for the LDS-resident M2L transforms"};
    HIPCHK(hipStreamSynchronize(c->stream));
  } catch (const ArgError& e) {
    xfmm_destroy(c);
    return fail(nullptr, XFMM_ERR_ARG, e.msg);
  } catch (const HipError& e) {
    xfmm_destroy(c);
    return fail(nullptr, XFMM_ERR_HIP, e.msg);
  } catch (...) {
    xfmm_destroy(c);
    return fail(nullptr, XFMM_ERR_NOMEM, "xfmm_create: allocation failed");
  }
  *out = c;
  return XFMM_OK;
}

void xfmm_destroy(xfmm_context* c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->stream2) (void)hipStreamSynchronize(c->stream2);
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  if (c->stream2) (void)hipStreamDestroy(c->stream2);
  delete c;  // DevBuf destructors free device memory
}

const char* xfmm_last_error(const xfmm_context* c) { return c ? c->err.c_str() : g_thread_error.c_str(); }

int xfmm_sizes(const xfmm_context* c, int* nsurf, int* nconv, int* nfreq) {
  if (!c) return fail(nullptr, XFMM_ERR_ARG, "null context");
  if (nsurf) *nsurf = c->nsurf;
  if (nconv) *nconv = c->nconv;
  if (nfreq) *nfreq = c->nfreq;
  return XFMM_OK;
}

int xfmm_set_operators(xfmm_context* c, int nlevels, const double* uc2e_u, const double* uc2e_v,
                       const double* dc2e_u, const double* dc2e_v, const double* m2m, const double* l2l, int nm2l,
                       const double* m2l) {
  return guarded(c, [&] {
    require(nlevels >= 1 && nm2l >= 1, "xfmm_set_operators: nlevels and nm2l must be >= 1");
    drop_graph(c);
    require(uc2e_u && uc2e_v && dc2e_u && dc2e_v && m2m && l2l && m2l, "xfmm_set_operators: null operator");
    if (c->kind == XFMM_LAPLACE) require(nlevels == 1 && nm2l == 1, "xfmm_set_operators: Laplace operators are scale invariant (nlevels = nm2l = 1)");
    const int ns = c->nsurf, vw = c->vw;
    const size_t ns2 = size_t(ns) * ns;
    // transposes for row-GEMM: out_row = in_row x op^T
    auto transpose_set = [&](const double* src, std::vector<double>& dst) {
      dst.resize(size_t(nlevels) * ns2 * vw);
      for (int L = 0; L < nlevels; L++)
        for (int i = 0; i < ns; i++)
          for (int j = 0; j < ns; j++)
            for (int w = 0; w < vw; w++)
              dst[(L * ns2 + size_t(j) * ns + i) * vw + w] = src[(L * ns2 + size_t(i) * ns + j) * vw + w];
    };
    // Rank of C2E_V = V S^+ per level: 1 + its last column that is not
    // exactly zero (S^+ holds 0 for the truncated singular values, which
    // come last).  The products with the columns past it are exact zeros, so
    // the pseudo-inverse is applied as (first r rows of C2E_U) then (first r
    // columns of C2E_V) with the same result.
    auto rank_of = [&](const double* v, std::vector<int>& r) {
      r.assign(nlevels, 1);
      for (int L = 0; L < nlevels; L++)
        for (int i = 0; i < ns; i++)
          for (int j = ns - 1; j >= r[L]; j--) {
            bool nz = false;
            for (int w = 0; w < vw; w++) nz |= v[(L * ns2 + size_t(i) * ns + j) * vw + w] != 0.0;
            if (nz) {
              r[L] = j + 1;
              break;
            }
          }
    };
    rank_of(uc2e_v, c->r_up);
    rank_of(dc2e_v, c->r_dn);
    // keep the first r columns of each level's transposed C2E_U, row stride r
    auto compact = [&](std::vector<double>& d, const std::vector<int>& r) {
      for (int L = 0; L < nlevels; L++)
        for (int k = 0; k < ns; k++)
          for (int n = 0; n < r[L]; n++)
            for (int w = 0; w < vw; w++)
              d[(L * ns2 + size_t(k) * r[L] + n) * vw + w] = d[(L * ns2 + size_t(k) * ns + n) * vw + w];
    };
    // complex operators: each level's K x N matrix (K rows of N complex
    // entries, row stride N) -> real 2K x 2N with [[Re, Im], [-Im, Re]]
    // blocks, levels at a stride of 4 ns^2 doubles (see gemm_op_size)
    auto expand = [&](const std::vector<double>& d, int K, const std::vector<int>* rk, int N, std::vector<double>& out) {
      const size_t src_level = size_t(K) * N * 2, dst_level = size_t(K) * N * 4;
      out.assign(size_t(nlevels) * dst_level, 0.0);
      for (int L = 0; L < nlevels; L++) {
        const int n = rk ? (*rk)[L] : N;
        const double* src = d.data() + size_t(L) * src_level;
        double* dst = out.data() + size_t(L) * dst_level;
        for (int k = 0; k < K; k++)
          for (int j = 0; j < n; j++) {
            const double re = src[(size_t(k) * n + j) * 2], im = src[(size_t(k) * n + j) * 2 + 1];
            dst[size_t(2 * k) * (2 * n) + 2 * j] = re;
            dst[size_t(2 * k) * (2 * n) + 2 * j + 1] = im;
            dst[size_t(2 * k + 1) * (2 * n) + 2 * j] = -im;
            dst[size_t(2 * k + 1) * (2 * n) + 2 * j + 1] = re;
          }
      }
    };
    std::vector<double> h, hx;
    auto put = [&](DevBuf<double>& dst, const std::vector<int>* rk) {
      if (!c->cplx) {
        dst.upload(h, c->stream);
        return;
      }
      expand(h, ns, rk, ns, hx);
      dst.upload(hx, c->stream);
      HIPCHK(hipStreamSynchronize(c->stream));  // hx is reused
    };
    transpose_set(uc2e_u, h);
    compact(h, c->r_up);
    put(c->uc2e_uT, &c->r_up);
    transpose_set(uc2e_v, h);
    put(c->uc2e_vT, nullptr);
    transpose_set(dc2e_u, h);
    compact(h, c->r_dn);
    put(c->dc2e_uT, &c->r_dn);
    transpose_set(dc2e_v, h);
    put(c->dc2e_vT, nullptr);
    // M2Mcat[L]: (8 ns) x ns, row o*ns + k, col n = M2M[L][o][n][k]
    // L2Lcat[L]: ns x (8 ns), row k, col o*ns + n = L2L[L][o][n][k]
    std::vector<double> mm(size_t(nlevels) * 8 * ns2 * vw), ll(size_t(nlevels) * 8 * ns2 * vw);
    for (int L = 0; L < nlevels; L++)
      for (int o = 0; o < 8; o++)
        for (int n = 0; n < ns; n++)
          for (int k = 0; k < ns; k++)
            for (int w = 0; w < vw; w++) {
              const size_t s = ((size_t(L) * 8 + o) * ns2 + size_t(n) * ns + k) * vw + w;
              mm[((size_t(L) * 8 * ns2) + (size_t(o) * ns + k) * ns + n) * vw + w] = m2m[s];
              ll[((size_t(L) * 8 * ns2) + size_t(k) * 8 * ns + size_t(o) * ns + n) * vw + w] = l2l[s];
            }
    if (c->cplx) {
      // M2Mcat: (8 ns) x ns complex -> (16 ns) x (2 ns); L2Lcat: ns x (8 ns) -> (2 ns) x (16 ns)
      expand(mm, 8 * ns, nullptr, ns, hx);
      c->m2mcat.upload(hx, c->stream);
      HIPCHK(hipStreamSynchronize(c->stream));
      expand(ll, ns, nullptr, 8 * ns, hx);
      c->l2lcat.upload(hx, c->stream);
      HIPCHK(hipStreamSynchronize(c->stream));
    } else {
      c->m2mcat.upload(mm, c->stream);
      c->l2lcat.upload(ll, c->stream);
    }
    // FP32 copies of the GEMM operators for the FP32 translation GEMMs
    const std::pair<DevBuf<double>*, DevBuf<float>*> f32ops[] = {
        {&c->uc2e_uT, &c->uc2e_uT32}, {&c->uc2e_vT, &c->uc2e_vT32}, {&c->dc2e_uT, &c->dc2e_uT32},
        {&c->dc2e_vT, &c->dc2e_vT32}, {&c->m2mcat, &c->m2mcat32},   {&c->l2lcat, &c->l2lcat32}};
    for (int i = 0; i < 6; i++) {
      const auto& pr = f32ops[i];
      if (i < 4 ? c->f32_pinv : c->f32_trans) {
        pr.second->alloc(pr.first->n);
        launch_to_float(pr.first->ptr, pr.second->ptr, pr.first->n, c->stream);
      } else {
        pr.second->release();
      }
    }
    const size_t msz = size_t(nm2l) * NM2L * c->nfreq * 64 * 2;
    // reference layout uploaded, then expanded on the device into the per-lane
    // MFMA operand form (twice the size) used by the Hadamard kernel
    if (c->f32_m2l) {
      std::vector<float> f(msz);
      for (size_t i = 0; i < msz; i++) f[i] = float(m2l[i]);
      DevBuf<float> compact;
      compact.upload(f, c->stream);
      c->m2l32.alloc(2 * msz);
      if (c->f16_had) {
        c->m2l_hscale.alloc(size_t(nm2l) * c->nfreq);
        if (!c->m2l_amax.ptr) c->m2l_amax.alloc(1);
      }
      launch_m2l_expand(c->m2l_dev(), nm2l, compact.ptr, c->m2l32.ptr, c->f16_had ? c->m2l_hscale.ptr : nullptr,
                        c->stream);
      HIPCHK(hipStreamSynchronize(c->stream));
      c->m2l64.release();
    } else {
      DevBuf<double> compact;
      compact.upload(m2l, msz, c->stream);
      c->m2l64.alloc(2 * msz);
      launch_m2l_expand(c->m2l_dev(), nm2l, compact.ptr, c->m2l64.ptr, nullptr, c->stream);
      HIPCHK(hipStreamSynchronize(c->stream));
      c->m2l32.release();
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    c->nlevels = nlevels;
    c->nm2l = nm2l;
    c->have_ops = true;
  });
}

int xfmm_set_tree(xfmm_context* c, const xfmm_tree* t) {
  return guarded(c, [&] {
    require(t != nullptr, "xfmm_set_tree: null tree");
    const int nn = t->nnodes, D = t->depth;
    require(nn >= 1 && D >= 0 && D < 30, "xfmm_set_tree: bad node count or depth");
    require(t->level_offset && t->center && t->child0 && t->src_begin && t->src_count && t->trg_begin &&
                t->trg_count && t->p2p_offset && t->m2p_offset && t->p2l_offset && t->m2l_source,
            "xfmm_set_tree: null array");
    require(t->nsrcs >= 0 && t->ntrgs >= 0, "xfmm_set_tree: negative body count");
    require((t->nsrcs == 0 || t->src_coord) && (t->ntrgs == 0 || t->trg_coord), "xfmm_set_tree: null coordinates");
    std::vector<int> lo(t->level_offset, t->level_offset + D + 2);
    // level 0 holds the root, or the roots of a forest of independent trees
    require(lo[0] == 0 && lo[1] >= 1 && lo[D + 1] == nn, "xfmm_set_tree: level_offset must start 0,>=1 and end at nnodes");
    std::vector<int> lev(nn);
    for (int l = 0; l <= D; l++) {
      require(lo[l + 1] >= lo[l], "xfmm_set_tree: level_offset not monotone");
      for (int i = lo[l]; i < lo[l + 1]; i++) lev[i] = l;
    }
    std::vector<int> ch(t->child0, t->child0 + nn);
    for (int i = 0; i < nn; i++) {
      if (ch[i] < 0) continue;
      require(ch[i] + 8 <= nn && lev[i] < D && ch[i] >= lo[lev[i] + 1] && ch[i] + 8 <= lo[lev[i] + 2],
              "xfmm_set_tree: child block outside the next level");
    }
    for (int i = 0; i < nn; i++) {
      const bool leaf = ch[i] < 0;
      const int sb = t->src_begin[i], sc = t->src_count[i], tb = t->trg_begin[i], tc = t->trg_count[i];
      require(sc >= 0 && tc >= 0 && sb >= 0 && tb >= 0 && sb + sc <= t->nsrcs && tb + tc <= t->ntrgs,
              "xfmm_set_tree: body range out of bounds");
      require(leaf || sc == 0, "xfmm_set_tree: a non-leaf holds sources");
    }
    auto check_csr = [&](const int* off, const int* lst, const char* name, bool leaf_sources) {
      require(off[0] == 0, std::string("xfmm_set_tree: ") + name + " offsets must start at 0");
      for (int i = 0; i < nn; i++) require(off[i + 1] >= off[i], std::string("xfmm_set_tree: ") + name + " offsets not monotone");
      const int m = off[nn];
      require(m == 0 || lst, std::string("xfmm_set_tree: null ") + name + " list");
      for (int j = 0; j < m; j++) {
        require(lst[j] >= 0 && lst[j] < nn, std::string("xfmm_set_tree: ") + name + " entry out of range");
        if (leaf_sources) require(ch[lst[j]] < 0, std::string("xfmm_set_tree: ") + name + " source is not a leaf");
      }
    };
    check_csr(t->p2p_offset, t->p2p_source, "p2p", true);
    check_csr(t->m2p_offset, t->m2p_source, "m2p", false);
    check_csr(t->p2l_offset, t->p2l_source, "p2l", true);
    for (int i = 0; i < nn; i++)
      for (int r = 0; r < NM2L; r++) {
        const int s = t->m2l_source[size_t(i) * NM2L + r];
        if (s < 0) continue;
        require(s < nn && ch[s] >= 0 && ch[i] >= 0 && lev[s] == lev[i],
                "xfmm_set_tree: M2L pair must join two non-leaves of one level");
      }

    hipStream_t st = c->stream;
    drop_graph(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    c->near_pending = false;
    c->have_tree = false;
    c->vscale.release();
    c->nnodes = nn;
    c->depth = D;
    c->nsrcs = t->nsrcs;
    c->ntrgs = t->ntrgs;
    c->r0 = t->r0;
    c->h_level_offset = lo;
    c->h_level = lev;
    c->h_child0 = ch;
    c->center.upload(t->center, size_t(nn) * 3, st);
    c->level.upload(lev, st);
    c->child0.upload(ch, st);
    c->src_begin.upload(t->src_begin, nn, st);
    c->src_count.upload(t->src_count, nn, st);
    c->trg_begin.upload(t->trg_begin, nn, st);
    c->trg_count.upload(t->trg_count, nn, st);
    c->src.upload(t->src_coord, size_t(t->nsrcs) * 3, st);
    {
      std::vector<double> rel(size_t(t->nsrcs) * 3);
      for (int i = 0; i < nn; i++) {
        if (ch[i] >= 0) continue;
        for (int k = t->src_begin[i]; k < t->src_begin[i] + t->src_count[i]; k++)
          for (int d = 0; d < 3; d++) rel[3 * size_t(k) + d] = t->src_coord[3 * size_t(k) + d] - t->center[3 * i + d];
      }
      c->src_rel.upload(rel, st);
      c->src_pack.alloc(std::max<size_t>(1, size_t(t->nsrcs) * 4));
      // L2P equivalent surface per level, relative to the box centre
      const std::vector<double> us = unit_surface(c->p);
      std::vector<double> tab(size_t(D + 1) * c->nsurf * 4, 0.0);
      for (int l = 0; l <= D; l++) {
        const double b = 2.95 * t->r0 * std::ldexp(1.0, -l);
        for (int i = 0; i < c->nsurf; i++)
          for (int d = 0; d < 3; d++) tab[(size_t(l) * c->nsurf + i) * 4 + d] = us[3 * size_t(i) + d] * b;
      }
      c->surf_tab.upload(tab, st);
      c->surf_tab32.upload(std::vector<float>(tab.begin(), tab.end()), st);
      HIPCHK(hipStreamSynchronize(st));  // the FP32 copy is a temporary
    }
    c->trg.upload(t->trg_coord, size_t(t->ntrgs) * 3, st);
    c->p2p_off.upload(t->p2p_offset, nn + 1, st);
    c->p2p_src.upload(t->p2p_source, t->p2p_offset[nn], st);
    c->m2p_off.upload(t->m2p_offset, nn + 1, st);
    c->m2p_src.upload(t->m2p_source, t->m2p_offset[nn], st);
    c->p2l_off.upload(t->p2l_offset, nn + 1, st);
    c->p2l_src.upload(t->p2l_source, t->p2l_offset[nn], st);

    // work lists
    std::vector<int> wp2p, wm2p, wp2l, wl2p, leaf;
    for (int i = 0; i < nn; i++) {
      const bool isleaf = ch[i] < 0;
      if (isleaf && t->trg_count[i] > 0) {
        if (t->p2p_offset[i + 1] > t->p2p_offset[i]) wp2p.push_back(i);
        if (t->m2p_offset[i + 1] > t->m2p_offset[i]) wm2p.push_back(i);
        wl2p.push_back(i);
      }
      if (t->p2l_offset[i + 1] > t->p2l_offset[i]) wp2l.push_back(i);
    }
    // leaves for P2M/L2P conversions (reference: all leaves with bodies), by level
    c->leaf_level_off.assign(D + 2, 0);
    for (int l = 0; l <= D; l++) {
      c->leaf_level_off[l] = int(leaf.size());
      for (int i = lo[l]; i < lo[l + 1]; i++)
        if (ch[i] < 0 && (t->src_count[i] > 0 || t->trg_count[i] > 0)) leaf.push_back(i);
    }
    c->leaf_level_off[D + 1] = int(leaf.size());
    std::vector<long> lsr(leaf.size()), lnr(leaf.size());
    for (size_t i = 0; i < leaf.size(); i++) {
      lsr[i] = long(i) * c->nsurf * c->vw;  // row offsets of the GEMMs, in doubles
      lnr[i] = long(leaf[i]) * c->nsurf * c->vw;
    }
    auto tiles = [&](const std::vector<int>& w) {
      std::vector<int> v;
      for (int leaf : w)
        for (int k = 0; 64 * k < t->trg_count[leaf]; k++) {
          v.push_back(leaf);
          v.push_back(k);
        }
      return v;
    };
    const std::vector<int> tp2p = tiles(wp2p), tl2p = tiles(wl2p);
    c->t_p2p.upload(tp2p, st);
    c->t_l2p.upload(tl2p, st);
    c->nt_p2p = int(tp2p.size() / 2);
    c->nt_l2p = int(tl2p.size() / 2);
    c->w_p2p.upload(wp2p, st);
    c->w_m2p.upload(wm2p, st);
    c->w_p2l.upload(wp2l, st);
    c->w_l2p_eval.upload(wl2p, st);
    c->w_leaf.upload(leaf, st);
    c->leaf_slot_rows.upload(lsr, st);
    c->leaf_node_rows.upload(lnr, st);
    c->n_p2p = int(wp2p.size());
    c->n_m2p = int(wm2p.size());
    c->n_p2l = int(wp2l.size());
    c->n_l2p_eval = int(wl2p.size());
    c->n_leaf = int(leaf.size());
    // non-leaves by level (level-major order already)
    std::vector<long> nr, ncr;
    c->nl_level_off.assign(D + 2, 0);
    for (int l = 0; l <= D; l++) {
      c->nl_level_off[l] = int(nr.size());
      for (int i = lo[l]; i < lo[l + 1]; i++)
        if (ch[i] >= 0) {
          nr.push_back(long(i) * c->nsurf * c->vw);
          ncr.push_back(long(ch[i]) * c->nsurf * c->vw);
        }
    }
    c->nl_level_off[D + 1] = int(nr.size());
    c->nl_rows.upload(nr, st);
    c->nl_child_rows.upload(ncr, st);
    find_close_pairs(c);

    // M2L: source / target slots and 64-target blocks per level
    std::vector<int> src_slot(nn, -1), trg_slot(nn, -1), srcs, trgs;
    std::vector<double> scale;
    for (int i = 0; i < nn; i++) {
      bool any = false;
      for (int r = 0; r < NM2L; r++) {
        const int s = t->m2l_source[size_t(i) * NM2L + r];
        if (s >= 0) {
          any = true;
          if (src_slot[s] < 0) {
            src_slot[s] = 0;  // mark; numbered below in node order
          }
        }
      }
      if (any) {
        trg_slot[i] = int(trgs.size());
        trgs.push_back(i);
        scale.push_back(c->kind == XFMM_LAPLACE ? std::ldexp(1.0, lev[i] + 1) : 1.0);
      }
    }
    for (int i = 0; i < nn; i++)
      if (src_slot[i] == 0 || src_slot[i] > 0) {
        src_slot[i] = int(srcs.size());
        srcs.push_back(i);
      }
    // blocks: consecutive targets of one level (Hilbert order), at most 64
    // targets and HAD_UROWS union rows per colour each (slot and lane
    // placement rules at HAD_UCAP)
    auto lattice = [&](int node, int* ijk) {
      const double h = 2.0 * t->r0 / double(1 << lev[node]);
      for (int d = 0; d < 3; d++) {
        const double v = std::floor((t->center[3 * size_t(node) + d] - (t->x0[d] - t->r0)) / h);
        ijk[d] = int(std::max(0.0, v));
      }
    };
    auto colour = [&](int node) {
      int q[3];
      lattice(node, q);
      return (q[0] + 4 * q[1] + 8 * q[2]) & 15;
    };
    std::vector<int> src_colour(srcs.size());
    for (size_t k = 0; k < srcs.size(); k++) src_colour[k] = colour(srcs[k]);
    std::vector<int> btrg, buni, bnu;
    std::vector<short> bloc;
    c->blk_level_off.assign(D + 2, 0);
    int nblk = 0;
    std::vector<int> lidx(std::max<size_t>(1, srcs.size()), -1);  // src slot -> union slot in current block
    std::vector<int> cur_t, cur_u;
    int rows[16] = {0};
    auto flush = [&]() {
      if (cur_t.empty()) return;
      const size_t bt = btrg.size(), bu = buni.size(), bl = bloc.size();
      btrg.resize(bt + 64, -1);
      buni.resize(bu + HAD_UCAP, -1);
      bloc.resize(bl + size_t(NM2L) * 64, short(-1));
      int nu = 0;
      for (int u : cur_u) {
        buni[bu + lidx[u]] = u;
        nu = std::max(nu, lidx[u] + 1);
      }
      // lane placement: preferred position, else the first free one
      std::vector<int> pos(cur_t.size(), -1);
      bool used[64] = {false};
      for (size_t jt = 0; jt < cur_t.size(); jt++) {
        int q[3];
        lattice(cur_t[jt], q);
        const int want = 16 * (((q[1] >> 1) & 1) + 2 * ((q[2] >> 1) & 1)) + (q[0] & 3) + 4 * (q[1] & 1) + 8 * (q[2] & 1);
        if (!used[want]) { used[want] = true; pos[jt] = want; }
      }
      int nextfree = 0;
      for (size_t jt = 0; jt < cur_t.size(); jt++) {
        if (pos[jt] >= 0) continue;
        while (used[nextfree]) nextfree++;
        used[nextfree] = true;
        pos[jt] = nextfree;
      }
      for (size_t jt = 0; jt < cur_t.size(); jt++) {
        const int node = cur_t[jt];
        btrg[bt + pos[jt]] = trg_slot[node];
        for (int r = 0; r < NM2L; r++) {
          const int sn = t->m2l_source[size_t(node) * NM2L + r];
          if (sn >= 0) bloc[bl + size_t(r) * 64 + pos[jt]] = short(lidx[src_slot[sn]]);
        }
      }
      for (int u : cur_u) lidx[u] = -1;
      for (int& r : rows) r = 0;
      bnu.push_back(nu);
      nblk++;
      cur_t.clear();
      cur_u.clear();
    };
    for (int l = 0; l <= D; l++) {
      c->blk_level_off[l] = nblk;
      for (int i = lo[l]; i < lo[l + 1]; i++) {
        if (trg_slot[i] < 0) continue;
        int extra[16] = {0};
        bool over = false;
        for (int r = 0; r < NM2L; r++) {
          const int sn = t->m2l_source[size_t(i) * NM2L + r];
          if (sn < 0 || lidx[src_slot[sn]] >= 0) continue;
          // (a source repeated over r is counted twice: conservative)
          const int col = src_colour[src_slot[sn]];
          if (rows[col] + ++extra[col] > HAD_UROWS) over = true;
        }
        if (cur_t.size() == 64 || over) flush();
        cur_t.push_back(i);
        for (int r = 0; r < NM2L; r++) {
          const int sn = t->m2l_source[size_t(i) * NM2L + r];
          if (sn < 0) continue;
          const int ss = src_slot[sn];
          if (lidx[ss] < 0) {
            const int col = src_colour[ss];
            lidx[ss] = 16 * rows[col]++ + col;
            cur_u.push_back(ss);
          }
        }
      }
      flush();
    }
    c->blk_level_off[D + 1] = nblk;
    c->n_m2l_src = int(srcs.size());
    c->n_m2l_trg = int(trgs.size());
    c->n_m2l_blk = nblk;
    c->m2l_src_node.upload(srcs, st);
    c->m2l_trg_node.upload(trgs, st);
    c->m2l_scale.upload(scale, st);
    c->m2l_blk_trg.upload(btrg, st);
    c->m2l_blk_uni.upload(buni, st);
    c->m2l_blk_nu.upload(bnu, st);
    c->m2l_blk_loc.upload(bloc, st);
    const size_t rb = c->f32_m2l ? 4 : 8;
    const size_t per = size_t(c->nfreq) * NCHILD * 2 * rb;
    c->fft_in.alloc(per * std::max(1, c->n_m2l_src));
    c->fft_out.alloc(per * std::max(1, c->n_m2l_trg));

    // buffers
    const size_t nv = size_t(nn) * c->nsurf * c->vw;
    c->up.alloc(nv);
    c->dn.alloc(nv);
    c->trg_value.alloc(std::max<size_t>(1, size_t(t->ntrgs) * 4 * c->vw));
    c->check.alloc(std::max<size_t>(1, leaf.size() * c->nsurf * c->vw));
    c->tmp.alloc(std::max<size_t>(1, leaf.size() * c->nsurf * c->vw));
    {
      size_t need = 0;  // split-K scratch of the per-level M2M / L2L GEMMs
      for (int l = 0; l < D; l++) {
        const int rows = c->nl_level_off[l + 1] - c->nl_level_off[l];
        if (rows <= 0) continue;
        need = std::max(need, gemm_scratch_doubles(rows, c->nsurf * c->vw, 8 * c->nsurf * c->vw));
        need = std::max(need, gemm_scratch_doubles(rows, 8 * c->nsurf * c->vw, c->nsurf * c->vw));
      }
      c->gemm_scratch.alloc(need);
    }
    c->q.alloc(std::max<size_t>(1, size_t(t->nsrcs) * c->vw));
    c->near.alloc(c->trg_value.n);
    launch_zero(c->q.ptr, c->q.n, st);
    launch_zero(c->up.ptr, c->up.n, st);
    launch_zero(c->dn.ptr, c->dn.n, st);
    launch_zero(c->trg_value.ptr, c->trg_value.n, st);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    c->have_tree = true;
    c->down_clean = false;
    c->net_heavy = false;  // zero charges
  });
}

// Mixed path: classify the charges by |sum q| / sum |q| (the net_heavy rule
// above; sums = {Re sum, Im sum, sum |q|}).  A captured graph holds the
// kernels of the class it was captured with: a change of class marks it for
// recapture at its next launch.
static void classify_charges(xfmm_context* c, const double* sums) {
  if (c->prec != XFMM_MIXED) return;
  const bool heavy = std::hypot(sums[0], sums[1]) > 0.1 * sums[2];
  if (heavy != c->net_heavy && c->graph_exec) c->graph_stale = !c->graph_stale;
  c->net_heavy = heavy;
}

int xfmm_set_charges_host(xfmm_context* c, const double* q) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_set_charges: tree not set");
    require(q || c->nsrcs == 0, "xfmm_set_charges: null charges");
    // a near-field P2P still reading the old charges on the second stream
    if (c->near_launched) HIPCHK(hipStreamWaitEvent(c->stream, c->ev[20], 0));
    c->near_pending = false;
    if (c->nsrcs) HIPCHK(hipMemcpyAsync(c->q.ptr, q, size_t(c->nsrcs) * c->vw * sizeof(double), hipMemcpyHostToDevice, c->stream));
    double sums[3] = {0, 0, 0};
    for (size_t i = 0; i < size_t(c->nsrcs); i++) {
      const double re = q[i * c->vw], im = c->vw == 2 ? q[2 * i + 1] : 0.0;
      sums[0] += re;
      sums[1] += im;
      sums[2] += std::hypot(re, im);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    classify_charges(c, sums);
  });
}

int xfmm_set_charges_device(xfmm_context* c, const void* q) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_set_charges: tree not set");
    require(q || c->nsrcs == 0, "xfmm_set_charges: null charges");
    // a near-field P2P still reading the old charges on the second stream
    if (c->near_launched) HIPCHK(hipStreamWaitEvent(c->stream, c->ev[20], 0));
    c->near_pending = false;
    if (c->nsrcs) HIPCHK(hipMemcpyAsync(c->q.ptr, q, size_t(c->nsrcs) * c->vw * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    if (c->prec == XFMM_MIXED) {
      double sums[3] = {0, 0, 0};
      if (c->nsrcs) {
        c->charge_sums.alloc(3 * (CHARGE_SUM_BLOCKS + 1));
        launch_charge_sums(c->q.ptr, c->nsrcs, c->vw == 2, c->charge_sums.ptr, c->stream);
        HIPCHK(hipMemcpyAsync(sums, c->charge_sums.ptr + 3 * CHARGE_SUM_BLOCKS, sizeof(sums), hipMemcpyDeviceToHost,
                              c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
      }
      classify_charges(c, sums);
    }
  });
}

int xfmm_clear_values(xfmm_context* c) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_clear_values: tree not set");
    launch_zero(c->up.ptr, c->up.n, c->stream);
    launch_zero(c->dn.ptr, c->dn.n, c->stream);
    launch_zero(c->trg_value.ptr, c->trg_value.n, c->stream);
    c->down_clean = true;
    HIPCHK(hipGetLastError());
  });
}

int xfmm_upward_pass(xfmm_context* c) {
  return guarded(c, [&] {
    check_ready(c);
    // clear the accumulators of both passes (the near field of the previous
    // evaluation has been consumed by then)
    launch_zero(c->up.ptr, c->up.n, c->stream);
    launch_zero(c->dn.ptr, c->dn.n, c->stream);
    launch_zero(c->trg_value.ptr, c->trg_value.n, c->stream);
    c->down_clean = true;
    record(c, 0);
    op_P2M(c);
    record(c, 1);
    op_M2M(c);
    record(c, 2);
    // the near field goes on the second stream here, so that it overlaps
    // the downward pass (measured: starting it with the upward pass, where it
    // competes with the VALU-bound P2M check potentials, is slower)
    launch_near(c);
    HIPCHK(hipGetLastError());
  });
}

int xfmm_downward_pass(xfmm_context* c) {
  return guarded(c, [&] {
    check_ready(c);
    if (!c->down_clean) {
      launch_zero(c->dn.ptr, c->dn.n, c->stream);
      launch_zero(c->trg_value.ptr, c->trg_value.n, c->stream);
    }
    c->down_clean = false;
    record(c, 3);
    op_P2L(c);
    record(c, 4);
    op_M2P(c);
    record(c, 5);
    // near field: taken from the second stream when the upward pass started
    // it, else computed here; either way it is added last, so both give
    // identical values
    const bool pend = c->near_pending;
    if (!pend) {
      DevTree t = c->dev_tree();
      t.trg_value = c->near.ptr;
      launch_zero(c->near.ptr, c->near.n, c->stream);
      launch_eval_targets(c->kind, c->f32_p2p, MODE_P2P, t, c->wavek(), c->w_p2p.ptr, c->n_p2p, c->t_p2p.ptr,
                      c->nt_p2p, c->stream);
    }
    record(c, 6);
    op_M2L(c, true);
    record(c, 7);
    op_L2L(c);
    record(c, 8);
    // every leaf with targets has an L2P work item, so the L2P kernel adds
    // the near field to each target's result exactly once
    op_L2P(c, c->near.ptr, pend);
    record(c, 9);
    if (c->vscale.ptr) launch_scale_values(c->trg_value.ptr, c->vscale.ptr, c->ntrgs, c->vw, c->stream);
    c->near_pending = false;
    c->near_overlapped = pend;
    HIPCHK(hipGetLastError());
  });
}

int xfmm_graph_capture(xfmm_context* c) {
  return guarded(c, [&] {
    check_ready(c);
    drop_graph(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    c->near_pending = false;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = xfmm_upward_pass(c);
    if (rc == XFMM_OK) rc = xfmm_downward_pass(c);
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc != XFMM_OK) {
      if (g) (void)hipGraphDestroy(g);
      throw ArgError{"xfmm_graph_capture: " + c->err};
    }
    HIPCHK(e);
    c->graph = g;
    HIPCHK(hipGraphInstantiate(&c->graph_exec, c->graph, nullptr, nullptr, 0));
    c->near_pending = false;
    c->down_clean = false;
  });
}

int xfmm_graph_launch(xfmm_context* c) {
  return guarded(c, [&] {
    require(c->graph_exec != nullptr, "xfmm_graph_launch: no graph (call xfmm_graph_capture)");
    if (c->graph_stale && xfmm_graph_capture(c) != XFMM_OK) throw ArgError{c->err};
    // the graph zeroes and recomputes everything it produces; eager state
    // afterwards is that of a completed evaluation
    HIPCHK(hipGraphLaunch(c->graph_exec, c->stream));
    c->near_pending = false;
    c->down_clean = false;
  });
}

#define XFMM_OP(NAME)                                   \
  int xfmm_##NAME(xfmm_context* c) {                    \
    return guarded(c, [&] {                             \
      check_ready(c);                                   \
      c->down_clean = false;                            \
      op_##NAME(c);                                     \
      HIPCHK(hipGetLastError());                        \
    });                                                 \
  }
XFMM_OP(P2M)
XFMM_OP(M2M)
XFMM_OP(P2L)
XFMM_OP(M2P)
XFMM_OP(P2P)
XFMM_OP(L2L)
XFMM_OP(L2P)
int xfmm_M2L(xfmm_context* c) {
  return guarded(c, [&] {
    check_ready(c);
    c->down_clean = false;
    op_M2L(c, false);
    HIPCHK(hipGetLastError());
  });
}

int xfmm_get_values(xfmm_context* c, double* v) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_get_values: tree not set");
    require(v || c->ntrgs == 0, "xfmm_get_values: null output");
    if (c->ntrgs)
      HIPCHK(hipMemcpyAsync(v, c->trg_value.ptr, size_t(c->ntrgs) * 4 * c->vw * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int xfmm_get_equiv(xfmm_context* c, double* up, double* dn) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_get_equiv: tree not set");
    const size_t n = size_t(c->nnodes) * c->nsurf * c->vw * sizeof(double);
    if (up) HIPCHK(hipMemcpyAsync(up, c->up.ptr, n, hipMemcpyDeviceToHost, c->stream));
    if (dn) HIPCHK(hipMemcpyAsync(dn, c->dn.ptr, n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int xfmm_set_equiv(xfmm_context* c, const double* up, const double* dn) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_set_equiv: tree not set");
    const size_t n = size_t(c->nnodes) * c->nsurf * c->vw * sizeof(double);
    if (up) HIPCHK(hipMemcpyAsync(c->up.ptr, up, n, hipMemcpyHostToDevice, c->stream));
    if (dn) HIPCHK(hipMemcpyAsync(c->dn.ptr, dn, n, hipMemcpyHostToDevice, c->stream));
    if (dn) c->down_clean = false;
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int xfmm_values_device(xfmm_context* c, void** ptr) {
  return guarded(c, [&] {
    require(c->have_tree && ptr, "xfmm_values_device: tree not set or null output");
    *ptr = c->trg_value.ptr;
    c->down_clean = false;  // the caller may write through the pointer
  });
}

int xfmm_synchronize(xfmm_context* c) {
  return guarded(c, [&] {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
  });
}

int xfmm_timings(const xfmm_context* cc, double* ms, int cap) {
  xfmm_context* c = const_cast<xfmm_context*>(cc);
  if (!c || !ms) return 0;
  if (!c->ev_ok) return 0;
  if (hipStreamSynchronize(c->stream) != hipSuccess || hipStreamSynchronize(c->stream2) != hipSuccess) return 0;
  // P2M M2M P2L M2P P2P M2L L2L L2P FFT HAD IFFT; P2P from the second stream
  // when it ran there (overlapped with the other operators)
  const bool ov = c->near_overlapped;
  const int pairs[11][2] = {{0, 1}, {1, 2}, {3, 4}, {4, 5}, {ov ? 21 : 5, ov ? 22 : 6}, {6, 7}, {7, 8}, {8, 9},
                            {16, 17}, {17, 18}, {18, 19}};
  int n = 0;
  for (int i = 0; i < 11 && n < cap; i++) {
    float v = 0;
    if (hipEventElapsedTime(&v, c->ev[pairs[i][0]], c->ev[pairs[i][1]]) != hipSuccess) v = -1;
    ms[n++] = v;
  }
  return n;
}

int xfmm_set_value_scale(xfmm_context* c, const double* scale) {
  return guarded(c, [&] {
    require(c->have_tree, "xfmm_set_value_scale: tree not set");
    drop_graph(c);  // a captured evaluation holds the previous setting
    if (!scale) {
      c->vscale.release();
      return;
    }
    for (size_t i = 0; i < size_t(c->ntrgs) * 2; i++)
      require(std::isfinite(scale[i]), "xfmm_set_value_scale: non-finite factor");
    c->vscale.upload(scale, size_t(c->ntrgs) * 2, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int xfmm_set_tuning(xfmm_context* c, int param, int value) {
  return guarded(c, [&] {
    if (param == XFMM_TUNE_HAD_KCH) {
      require(value >= 1 && value <= 256, "xfmm_set_tuning: Hadamard frequency chunk must be in [1, 256]");
      c->tune_kch = value;
    } else if (param == XFMM_TUNE_P2P_SPLIT) {
      require(value >= 0 && value <= 32, "xfmm_set_tuning: P2P split threshold must be in [0, 32]");
      c->tune_split = value;
    } else if (param == XFMM_TUNE_HAD_GROUPS) {
      require(value == 1 || value == 2 || value == 4, "xfmm_set_tuning: Hadamard groups must be 1, 2 or 4");
      c->tune_groups = value;
    } else if (param == XFMM_TUNE_HAD_TILES) {
      require(value == 1 || value == 2, "xfmm_set_tuning: Hadamard tiles per wave must be 1 or 2");
      c->tune_tiles = value;
    } else {
      require(false, "xfmm_set_tuning: unknown parameter");
    }
    drop_graph(c);
  });
}

int xfmm_get_tuning(const xfmm_context* c, int param, int* value) {
  if (!c || !value) return fail(nullptr, XFMM_ERR_ARG, "xfmm_get_tuning: null argument");
  if (param == XFMM_TUNE_HAD_KCH) *value = c->tune_kch;
  else if (param == XFMM_TUNE_P2P_SPLIT) *value = c->tune_split;
  else if (param == XFMM_TUNE_HAD_GROUPS) *value = c->tune_groups;
  else if (param == XFMM_TUNE_HAD_TILES) *value = c->tune_tiles;
  else return fail(const_cast<xfmm_context*>(c), XFMM_ERR_ARG, "xfmm_get_tuning: unknown parameter");
  return XFMM_OK;
}

int xfmm_autotune(xfmm_context* c, int reps, double* best_ms) {
  return guarded(c, [&] {
    check_ready(c);
    require(reps >= 1 && reps <= 100, "xfmm_autotune: reps must be in [1, 100]");
    drop_graph(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipStreamSynchronize(c->stream2));
    c->near_pending = false;
    c->down_clean = false;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIPCHK(hipEventCreate(&e0));
    struct EvGuard {
      hipEvent_t& a;
      hipEvent_t& b;
      ~EvGuard() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
      }
    } eg{e0, e1};
    HIPCHK(hipEventCreate(&e1));
    // fastest of `reps` runs of f (one warm-up run first)
    auto time_it = [&](auto&& f) {
      f();
      float best = 1e30f;
      for (int r = 0; r < reps; r++) {
        HIPCHK(hipEventRecord(e0, c->stream));
        f();
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float v = 0;
        HIPCHK(hipEventElapsedTime(&v, e0, e1));
        best = std::min(best, v);
      }
      HIPCHK(hipGetLastError());
      return double(best);
    };
    double ms_split = 0, ms_kch = 0;
    // near field and L2P of the real kernels: largest tail tile whose
    // sources are split over lanes (results go to the scratch `near`)
    if (!c->cplx && c->nt_p2p + c->nt_l2p > 0) {
      const int cand[] = {0, 8, 16, 32};
      int best = c->tune_split;
      double bt = 1e30;
      for (int v : cand) {
        c->tune_split = v;
        const double ms = time_it([&] {
          DevTree t = c->dev_tree();
          t.trg_value = c->near.ptr;
          launch_eval_targets(c->kind, c->f32_p2p, MODE_P2P, t, c->wavek(), c->w_p2p.ptr, c->n_p2p, c->t_p2p.ptr,
                              c->nt_p2p, c->stream);
          launch_eval_targets(c->kind, c->f32_far, MODE_L2P, t, c->wavek(), c->w_l2p_eval.ptr, c->n_l2p_eval,
                              c->t_l2p.ptr, c->nt_l2p, c->stream);
        });
        if (ms < bt) {
          bt = ms;
          best = v;
        }
      }
      c->tune_split = best;
      ms_split = bt;
    }
    // Hadamard frequency chunk (spectra of the current equivalent densities)
    if (c->n_m2l_trg > 0) {
      M2LDev m = c->m2l_dev();
      DevTree t = c->dev_tree();
      if (m.amax) launch_zero(c->m2l_amax.ptr, 1, c->stream);
      launch_fft_up(m, t, c->m2l_src_node.ptr, c->n_m2l_src, c->stream);
      const int cand[] = {4, 8, 16, 32, 64};
      // groups and tiles per wave matter only for the split-FP16 kernel
      const int ngr = (c->f16_had && m.amax) ? 4 : 1;
      int best = c->tune_kch, bestg = c->tune_groups, bestt = c->tune_tiles;
      double bt = 1e30;
      for (int gi = 0; gi < ngr; gi++)
        for (int v : cand) {
          c->tune_kch = v;
          if (ngr > 1) {
            c->tune_groups = (gi & 1) ? 2 : 1;
            c->tune_tiles = (gi & 2) ? 2 : 1;
          }
          const double ms = time_it([&] { run_hadamards(c, m); });
          if (ms < bt) {
            bt = ms;
            best = v;
            bestg = c->tune_groups;
            bestt = c->tune_tiles;
          }
        }
      c->tune_kch = best;
      c->tune_groups = bestg;
      c->tune_tiles = bestt;
      ms_kch = bt;
    }
    launch_zero(c->near.ptr, c->near.n, c->stream);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (best_ms) {
      best_ms[0] = ms_split;
      best_ms[1] = ms_kch;
    }
  });
}

int xfmm_direct(int kernel, int precision, double wavek_re, double wavek_im, const double* src, const double* q,
                int ns, const double* trg, int nt, double* trg_value, int gradient) {
  if (ns < 0 || nt < 0) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct: negative size");
  if (kernel < XFMM_LAPLACE || kernel > XFMM_MODIFIED_HELMHOLTZ) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct: unknown kernel");
  if (nt == 0) return XFMM_OK;
  if ((ns && (!src || !q)) || !trg || !trg_value) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct: null pointer");
  const int vw = kernel == XFMM_HELMHOLTZ ? 2 : 1;
  const int na = (gradient ? 4 : 1) * vw;
  try {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard { hipStream_t s; ~StreamGuard() { (void)hipStreamDestroy(s); } } sg{s};
    DevBuf<double> dsrc, dq, dtrg, dout, scratch;
    dsrc.upload(src, size_t(ns) * 3, s);
    dq.upload(q, size_t(ns) * vw, s);
    dtrg.upload(trg, size_t(nt) * 3, s);
    dout.upload(trg_value, size_t(nt) * na, s);
    const int tblocks = (nt + 63) / 64;
    int nsplit = std::max(1, std::min(2048 / tblocks + 1, (ns + 255) / 256));
    scratch.alloc(size_t(nsplit) * nt * na);
    launch_direct(kernel, precision != XFMM_FP64, vw == 2, Wavek{wavek_re, wavek_im}, dsrc.ptr, dq.ptr, ns, dtrg.ptr, nt,
                  dout.ptr, gradient, scratch.ptr, nsplit, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(trg_value, dout.ptr, size_t(nt) * na * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
  } catch (const HipError& e) {
    return fail(nullptr, e.msg.find("hipMalloc") != std::string::npos ? XFMM_ERR_NOMEM : XFMM_ERR_HIP, e.msg);
  } catch (...) {
    return fail(nullptr, XFMM_ERR_NOMEM, "xfmm_direct: allocation failed");
  }
  return XFMM_OK;
}


int xfmm_direct_noalloc(int kernel, int precision, double wavek_re, double wavek_im, const double* src,
                        const double* q, int ns, const double* trg, int nt, double* trg_value, int gradient,
                        void* stream) {
  if (ns < 0 || nt < 0) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct_noalloc: negative size");
  if (kernel < XFMM_LAPLACE || kernel > XFMM_MODIFIED_HELMHOLTZ)
    return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct_noalloc: unknown kernel");
  if (nt == 0) return XFMM_OK;
  if ((ns && (!src || !q)) || !trg || !trg_value) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct_noalloc: null pointer");
  const int vw = kernel == XFMM_HELMHOLTZ ? 2 : 1;
  const int na = (gradient ? 4 : 1) * vw;
  constexpr size_t cap = XFMM_NOALLOC_SCRATCH_BYTES / sizeof(double);
  const size_t per_split = size_t(nt) * na;
  if (per_split > cap)
    return fail(nullptr, XFMM_ERR_ARG,
                "xfmm_direct_noalloc: " + std::to_string(nt) + " targets x " + std::to_string(na) +
                    " values exceed the " + std::to_string(XFMM_NOALLOC_SCRATCH_BYTES) + "-byte scratch buffer (at most " +
                    std::to_string(cap / na) + " targets for this kernel and output)");
  try {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, XFMM_ERR_ARG, "xfmm_direct_noalloc: device ordinal out of range");
    double* part = scratch_ptr(dev);
    // same source split as xfmm_direct, capped by the scratch capacity
    const int tblocks = (nt + 63) / 64;
    int nsplit = std::max(1, std::min(2048 / tblocks + 1, (ns + 255) / 256));
    nsplit = int(std::min<size_t>(size_t(nsplit), cap / per_split));
    hipStream_t s = static_cast<hipStream_t>(stream);
    launch_direct(kernel, precision != XFMM_FP64, vw == 2, Wavek{wavek_re, wavek_im}, src, q, ns, trg, nt, trg_value,
                  gradient, part, nsplit, s);
    HIPCHK(hipGetLastError());
  } catch (const HipError& e) {
    return fail(nullptr, XFMM_ERR_HIP, e.msg);
  } catch (...) {
    return fail(nullptr, XFMM_ERR_HIP, "xfmm_direct_noalloc: unexpected error");
  }
  return XFMM_OK;
}

}  // extern "C"
