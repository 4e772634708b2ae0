/*
 * exafmm_t_amd.h -- C ABI of the MI355X (gfx950) FMM evaluation engine.
 *
 * The engine replaces the evaluation phase of exafmm-t, i.e.
 * FmmBase<T>::upward_pass / downward_pass (reference include/fmm_base.h:162-200)
 * and the operators they call.  Everything that exafmm-t does before the
 * evaluation (tree, 2:1 balance, lists, precompute, M2L_setup) stays on the
 * host and is handed over once through xfmm_set_operators / xfmm_set_tree.
 *
 * Conventions
 *  - Plain C types only.  All host pointers are read during the call and may
 *    be freed afterwards.  "T" below means double for the real kernels
 *    (Laplace, modified Helmholtz) and interleaved (re, im) double pairs for
 *    the Helmholtz kernel.
 *  - Every function returns XFMM_OK (0) or a positive status code; nothing is
 *    thrown across the ABI.  xfmm_last_error() returns the message of the last
 *    failure of that context (or of the calling thread for context creation).
 *  - One context = two HIP streams on one device (the near-field P2P of an
 *    evaluation runs on the second, concurrently with the upward pass and
 *    M2L).  Contexts are independent; calls on one context must not overlap.
 *  - Sorted order: sources/targets in the order given by xfmm_tree (grouped by
 *    leaf); the engine never sees the caller's original body order.
 */
#ifndef EXAFMM_T_AMD_H
#define EXAFMM_T_AMD_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { XFMM_LAPLACE = 0, XFMM_HELMHOLTZ = 1, XFMM_MODIFIED_HELMHOLTZ = 2 };
enum { XFMM_FP64 = 0, XFMM_MIXED = 1, XFMM_FLOAT = 2 };
enum {
  XFMM_OK = 0,
  XFMM_ERR_ARG = 1,    /* invalid argument / inconsistent tree */
  XFMM_ERR_STATE = 2,  /* call order violated (e.g. evaluate before set_tree) */
  XFMM_ERR_HIP = 3,    /* HIP runtime error (message has the HIP error string) */
  XFMM_ERR_NOMEM = 4   /* device allocation failed */
};

typedef struct xfmm_context xfmm_context;

/* Number of HIP devices visible to the engine's HIP runtime (0 if none).
 * The engine links the system ROCm runtime; a process that has already
 * initialised another HIP runtime copy (for example one bundled with a
 * Python wheel) sees no device here. */
int xfmm_device_count(int* count);

/* Flattened octree.  Node order is level-major: nodes of level l are
 * [level_offset[l], level_offset[l+1]).  The 8 children of a non-leaf n are
 * the consecutive nodes child0[n] .. child0[n]+7 in octant order (x bit 0,
 * y bit 1, z bit 2).  Leaves have child0 = -1.  Lists are CSR arrays indexed
 * by target node: list entries of node n are list[off[n] .. off[n+1]).
 * Replaces the Node pointer graph of the reference (include/exafmm_t.h:102-129,
 * lists built by include/build_list.h:16-197, M2L_setup of
 * include/fmm_scale_invariant.h:277-341 / include/fmm.h:265-333). */
typedef struct {
  int nnodes;
  int depth;                 /* deepest level; FmmBase::depth */
  double r0;                 /* half side of the root box; FmmBase::r0 */
  double x0[3];              /* centre of the root box */
  const int* level_offset;   /* [depth + 2] */
  const double* center;      /* [nnodes][3] */
  const int* child0;         /* [nnodes] */
  const int* src_begin;      /* [nnodes] leaves: first sorted source; else 0 */
  const int* src_count;      /* [nnodes] leaves: number of sources; else 0 */
  const int* trg_begin;      /* [nnodes] */
  const int* trg_count;      /* [nnodes] */
  int nsrcs;
  int ntrgs;
  const double* src_coord;   /* [nsrcs][3], sorted by leaf */
  const double* trg_coord;   /* [ntrgs][3], sorted by leaf */
  const int* p2p_offset;     /* [nnodes + 1]; sources: leaf nodes (Node::P2P_list) */
  const int* p2p_source;
  const int* m2p_offset;     /* [nnodes + 1]; sources: any node (Node::M2P_list) */
  const int* m2p_source;
  const int* p2l_offset;     /* [nnodes + 1]; sources: leaf nodes (Node::P2L_list) */
  const int* p2l_source;
  const int* m2l_source;     /* [nnodes][26]; source node per M2L relpos or -1 (Node::M2L_list) */
} xfmm_tree;

/* Create a context (FmmBase ctor: p, kernel, wavek; reference
 * include/laplace.h:15, include/helmholtz.h:16, include/modified_helmholtz.h:17).
 * wavek_im is used by the Helmholtz kernel only.  device = HIP device ordinal.
 * precision: XFMM_FP64 computes everything in FP64.  XFMM_MIXED is the
 * mixed-precision path: FP32 near field (P2P, Kahan-combined tiles), FP32 M2L
 * (transforms; Hadamard products on the FP16 matrix cores with split-FP16
 * operands and FP32 accumulation), FP32 check potentials (P2M, P2L), FP32
 * L2P / M2P evaluation and FP32 M2M / L2L translations, while the
 * check-to-equivalent pseudo-inverse solves of P2M and L2P stay FP64 (FP64
 * operators and storage throughout).  It expects operators whose
 * pseudo-inverses drop singular values below 4e-8 max(S) (the reference's
 * FLOAT EPS, exafmm_t.h:23), which the C++ / Python host layer precomputes
 * for this mode: those directions only amplify FP32 rounding.  The environment
 * variable XFMM_MIXED_STAGES (comma list of p2p, m2l, had16, fft, surf, far,
 * trans, pinv) overrides the FP32 stage set of XFMM_MIXED when a context is
 * created.  Charges whose net sum exceeds a tenth of their absolute sum
 * (|sum q| > 0.1 sum |q|, e.g. charges of one sign) make the potential large
 * against its gradient; while such charges are set, XFMM_MIXED runs M2M / L2L
 * and the transforms' arithmetic in FP64 (their FP32 rounding, relative to
 * the potential, would otherwise reach ~1e-5 of the gradient).  XFMM_FLOAT is the counterpart of the reference's -DFLOAT build
 * (include/exafmm_t.h:21-37): every evaluation stage computes in FP32,
 * including the pseudo-inverse solves (FP32 MFMA); values stay FP64 in memory.
 * It expects the same truncated operators: with untruncated pseudo-inverses
 * FP32 arithmetic in P2M, L2P or the GEMMs loses the result. */
int xfmm_create(int kernel, int p, double wavek_re, double wavek_im, int precision, int device,
                xfmm_context** out);
void xfmm_destroy(xfmm_context* ctx);
const char* xfmm_last_error(const xfmm_context* ctx); /* ctx may be NULL */

/* Derived sizes (reference include/fmm_base.h:27-36). */
int xfmm_sizes(const xfmm_context* ctx, int* nsurf, int* nconv, int* nfreq);

/* Precomputed operators (outputs of precompute(); reference
 * include/fmm_scale_invariant.h:151-160, include/fmm.h:141-153).
 * nlevels = 1 for Laplace (scale invariant), depth + 1 otherwise.
 *   uc2e_u, uc2e_v, dc2e_u, dc2e_v : T[nlevels][nsurf][nsurf] (row-major)
 *   m2m, l2l                       : T[nlevels][8][nsurf][nsurf]
 *   m2l                            : double[nm2l][26][nfreq][8*8][2]
 *     (layout of matrix_M2L, fmm_scale_invariant.h:545-552); nm2l = 1 for
 *     Laplace, depth for the others (slot l serves target level l). */
int xfmm_set_operators(xfmm_context* ctx, int nlevels, const double* uc2e_u, const double* uc2e_v,
                       const double* dc2e_u, const double* dc2e_v, const double* m2m, const double* l2l,
                       int nm2l, const double* m2l);

/* Upload the tree once (validates the layout; returns XFMM_ERR_ARG on a
 * malformed tree).  Charges are reset to zero. */
int xfmm_set_tree(xfmm_context* ctx, const xfmm_tree* tree);

/* Sorted charges, T[nsrcs].  Host pointer (..._host) or device pointer
 * already resident in HBM (..._device, same layout; XFMM_MIXED sums them on
 * the device and waits for the two sums).  XFMM_MIXED classifies the charges
 * by |sum q| / sum |q| (see xfmm_create); after a change of class the next
 * xfmm_graph_launch captures the graph again. */
int xfmm_set_charges_host(xfmm_context* ctx, const double* q);
int xfmm_set_charges_device(xfmm_context* ctx, const void* q_device);

/* Zero target values and equivalent/check vectors (python clear_values,
 * reference python/exafmm.cpp:313-323). */
int xfmm_clear_values(xfmm_context* ctx);

/* FmmBase::upward_pass / downward_pass (include/fmm_base.h:162-200).  Each
 * pass computes from zero, so repeated evaluation does not need clear_values:
 * upward_pass zeroes up_equiv, dn_equiv and the target values (all before
 * the near field starts); downward_pass zeroes dn_equiv and the target values
 * again only if anything wrote them since.
 * upward_pass also starts the near-field P2P of the current charges on the
 * second stream; downward_pass adds it to the target values (setting new
 * charges in between discards it and downward_pass recomputes it). */
int xfmm_upward_pass(xfmm_context* ctx);
int xfmm_downward_pass(xfmm_context* ctx);

/* One evaluation (upward_pass + downward_pass, the same kernels on the same
 * two streams) captured into a HIP graph held by the context, and its replay
 * on the context's stream.  A replay reads the charges set at the time it
 * runs and leaves the same values as an eager evaluation.  For callers that
 * evaluate many times, or many small problems (one graph per context, all
 * replayed back to back); no counterpart in the reference, whose evaluation
 * is host-driven.  The graph is recaptured by calling capture again (and by
 * the next launch itself after charges of the other balance class were set,
 * XFMM_MIXED, see xfmm_create), and dropped by xfmm_set_tree /
 * xfmm_set_operators.  Per-operator timings are
 * not recorded by replays. */
int xfmm_graph_capture(xfmm_context* ctx);
int xfmm_graph_launch(xfmm_context* ctx);

/* Per-target output factors, double[ntrgs][2] in sorted target order: at the
 * end of every downward pass the potential of target i is multiplied by
 * scale[2 i] and its gradient by scale[2 i + 1] (NULL removes them).  No
 * counterpart in the reference; used to evaluate many independent Laplace
 * problems as one forest of normalised trees (a problem scaled by s into the
 * shared root box has potential s phi' and gradient s^2 grad'; see the
 * exafmm.batch module).  Reset by xfmm_set_tree; drops a captured graph. */
int xfmm_set_value_scale(xfmm_context* ctx, const double* scale);

/* Launch-shape parameters of one context (no counterpart in the reference,
 * whose loops have fixed blocking; none of them changes which pairs are
 * summed):
 *   XFMM_TUNE_HAD_KCH    frequencies per Hadamard workgroup, 1..256 (default 16)
 *   XFMM_TUNE_P2P_SPLIT  real kernels: P2P / L2P tail tiles of at most this many
 *                        targets (0..32) split each target's sources over
 *                        several lanes (default 32; the partial sums then add
 *                        in another order)
 *   XFMM_TUNE_HAD_GROUPS split-FP16 Hadamard (mixed / FLOAT): 64-target blocks
 *                        per workgroup, 1, 2 or 4, which share the operators
 *                        staged in LDS for a frequency (default 2)
 *   XFMM_TUNE_HAD_TILES  split-FP16 Hadamard: 16-target tiles per wave, 1 or 2;
 *                        a wave applies each operator operand it reads from
 *                        LDS to all of its tiles (default 1)
 * Setting one drops a captured graph. */
enum { XFMM_TUNE_HAD_KCH = 0, XFMM_TUNE_P2P_SPLIT = 1, XFMM_TUNE_HAD_GROUPS = 2, XFMM_TUNE_HAD_TILES = 3 };
int xfmm_set_tuning(xfmm_context* ctx, int param, int value);
int xfmm_get_tuning(const xfmm_context* ctx, int param, int* value);

/* Autotuner: on the uploaded tree and operators, times every candidate of
 * each tuning parameter (P2P split 0, 8, 16, 32 over the P2P + L2P kernels;
 * Hadamard chunk 4, 8, 16, 32, 64 times groups 1, 2 times tiles per wave 1, 2
 * over the Hadamard launches), fastest of `reps` runs each (1..100), and keeps the fastest.
 * best_ms (may be NULL) receives the two best times.  Leaves no evaluation result behind (call the
 * passes again); drops a captured graph, so capture after tuning. */
int xfmm_autotune(xfmm_context* ctx, int reps, double* best_ms);

/* Individual operators, same order and semantics as the reference
 * (accumulating into the device buffers; no zeroing):
 *   P2M  fmm_scale_invariant.h:163-195 / fmm.h:156-185
 *   M2M  fmm_scale_invariant.h:233-253 / fmm.h:221-241 (all levels, bottom-up)
 *   P2L  fmm_base.h:127-154     M2P  fmm_base.h:97-124    P2P  fmm_base.h:82-94
 *   M2L  fmm_scale_invariant.h:407-440 / fmm.h:395-446
 *   L2L  fmm_scale_invariant.h:256-275 / fmm.h:244-263 (all levels, top-down)
 *   L2P  fmm_scale_invariant.h:198-230 / fmm.h:188-218 */
int xfmm_P2M(xfmm_context* ctx);
int xfmm_M2M(xfmm_context* ctx);
int xfmm_P2L(xfmm_context* ctx);
int xfmm_M2P(xfmm_context* ctx);
int xfmm_P2P(xfmm_context* ctx);
int xfmm_M2L(xfmm_context* ctx);
int xfmm_L2L(xfmm_context* ctx);
int xfmm_L2P(xfmm_context* ctx);

/* Results: T[ntrgs][4] = (phi, dphi/dx, dphi/dy, dphi/dz) at each sorted
 * target (Node::trg_value).  Equivalent vectors: T[nnodes][nsurf] in tree node
 * order (Node::up_equiv / Node::dn_equiv; dn holds downward check potentials
 * until L2P converts the leaves). */
int xfmm_get_values(xfmm_context* ctx, double* trg_value);
int xfmm_get_equiv(xfmm_context* ctx, double* up_equiv, double* dn_equiv);
int xfmm_set_equiv(xfmm_context* ctx, const double* up_equiv, const double* dn_equiv);
/* Device pointer of the target values (valid until the next set_tree). */
int xfmm_values_device(xfmm_context* ctx, void** trg_value_device);
int xfmm_synchronize(xfmm_context* ctx);

/* Per-operator device time of the last passes in milliseconds, measured with
 * HIP events on the context stream.  names: P2M M2M P2L M2P P2P M2L L2L L2P,
 * plus the M2L stages FFT HAD IFFT.  Returns the number of entries written. */
int xfmm_timings(const xfmm_context* ctx, double* ms, int cap);

/* Direct summation on the device, no tree: LaplaceFmm/HelmholtzFmm/
 * ModifiedHelmholtzFmm::gradient_P2P (laplace.h:91-164, helmholtz.h:106-208,
 * modified_helmholtz.h:94-171) when gradient = 1, potential_P2P otherwise
 * (1 value per target).  Accumulates (+=) into trg_value like the reference.
 * src: [ns][3], q: T[ns], trg: [nt][3], trg_value: T[nt][gradient ? 4 : 1]. */
int xfmm_direct(int kernel, int precision, double wavek_re, double wavek_im, const double* src,
                const double* q, int ns, const double* trg, int nt, double* trg_value, int gradient);

/* Size of the device scratch array used by xfmm_direct_noalloc (32 MiB per
 * device, held from the first call on that device until the process ends). */
#define XFMM_NOALLOC_SCRATCH_BYTES (32u << 20)

/* Allocation-free variant of xfmm_direct.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * kernels, arguments and result (+= into trg_value) as xfmm_direct, except
 * that src, q, trg and trg_value are DEVICE pointers on the current device
 * and the work is only enqueued on `stream` (a hipStream_t; NULL = the null
 * stream): the call returns before the result is ready.  It calls no
 * hipMalloc / hipFree; the per-source-split partial sums go to a file-scope
 * device array of XFMM_NOALLOC_SCRATCH_BYTES, so nt * (gradient ? 4 : 1) *
 * (Helmholtz ? 2 : 1) doubles must fit in it (at most 1,048,576 targets for a
 * real kernel with gradient), else XFMM_ERR_ARG.  Every call on one device
 * shares that device's array: calls that overlap on one device (other
 * streams, other threads) race and must be serialised by the caller. */
int xfmm_direct_noalloc(int kernel, int precision, double wavek_re, double wavek_im, const double* src,
                        const double* q, int ns, const double* trg, int nt, double* trg_value, int gradient,
                        void* stream);

#ifdef __cplusplus
}
#endif
#endif
