// Host-callable launchers of the gfx950 kernels (internal interface of the
// engine library; the public interface is include/exafmm_t_amd.h).
#pragma once
#include <hip/hip_runtime.h>
#include "common.h"

namespace xfmm {

// Device view of the flattened tree and the evaluation buffers.  Values of
// kind T are doubles (real kernels) or interleaved (re, im) doubles
// (Helmholtz); `vw` is the number of doubles per value (1 or 2).
struct DevTree {
  int nnodes = 0, depth = 0, p = 0, nsurf = 0, vw = 1, nsrcs = 0;
  double r0 = 0;
  const double* center = nullptr;  // [nnodes][3]
  const int* level = nullptr;      // [nnodes]
  const int* child0 = nullptr;     // [nnodes]
  const int* src_begin = nullptr;
  const int* src_count = nullptr;
  const int* trg_begin = nullptr;
  const int* trg_count = nullptr;
  const double* src = nullptr;     // [nsrcs][3]
  const double* q = nullptr;       // T[nsrcs]
  const double* trg = nullptr;     // [ntrgs][3]
  const int* p2p_off = nullptr;
  const int* p2p_src = nullptr;
  const int* m2p_off = nullptr;
  const int* m2p_src = nullptr;
  const int* p2l_off = nullptr;
  const int* p2l_src = nullptr;
  const double* unit_surf = nullptr;  // [nsurf][3] unit-cube surface (reference point order)
  const double* src_rel = nullptr;    // [nsrcs][3] source positions relative to their leaf centre
  const double* surf_tab = nullptr;   // [depth + 1][nsurf][4] L2P equivalent surface (radius 2.95 r_l), (x, y, z, 0) relative to the centre
  const float* surf_tab32 = nullptr;  // the same in FP32 (FP32 L2P)
  void* src_pack = nullptr;           // [nsrcs] (x, y, z, q) relative to the leaf centre, in the P2P type
  double* up = nullptr;               // T[nnodes][nsurf]
  double* dn = nullptr;               // T[nnodes][nsurf]
  double* trg_value = nullptr;        // T[ntrgs][4]
  const double* near_add = nullptr;   // L2P only (may be null): [ntrgs][4] values that the L2P kernels add to
                                      // each target's result after their own sum (the evaluation's near field)
  int p2p_split = 32;                 // real-kernel P2P / L2P: tail tiles of <= p2p_split targets (<= 32) split their sources over lanes
  const int* close = nullptr;         // near-coincident P2P pairs, [4] ints (target, source, target leaf, source leaf), sorted by target
  const int* close_off = nullptr;     // [nclose + 1]: pairs of the k-th target with any are close[close_off[k] .. close_off[k + 1])
  const double* close_coef = nullptr; // [pairs][4 or 8]: FP64 minus FP32 value of each pair per unit charge
  int nclose = 0;                     // targets with near-coincident pairs
};

enum TargetMode { MODE_P2P = 0, MODE_M2P = 1, MODE_L2P = 2 };

// Near-coincident pairs of the P2P lists (distance below 1/16 of the target
// leaf's side, coinciding points excluded), as [4] ints (target, source,
// target leaf, source leaf) into out[0..cap); *count receives the number
// found (which may exceed cap: then call again with a larger buffer).
// After an FP32 near field, launch_eval_targets replaces the FP32
// contribution of each listed pair by its FP64 value (t.close, t.close_off,
// t.close_coef, t.nclose: the list sorted by target, grouped per target).
void launch_close_find(const DevTree& t, const int* work, int nwork, int* out, int cap, int* count, hipStream_t s);
// coef[k][4 or 8] of the sorted list t.close (npairs pairs): the pair's FP64
// minus FP32 value per unit charge, FP32 operands formed as the near-field
// kernel of `kind` forms them
void launch_close_coef(int kind, const DevTree& t, Wavek k, int npairs, double* coef, hipStream_t s);
enum SurfaceMode { MODE_P2M = 0, MODE_P2L = 1 };

// Particle targets of the leaves in work[0..nwork): P2P (sources: leaf
// particles of the P2P list), M2P (sources: upward equivalent surfaces of the
// M2P list, alpha 1.05), L2P (source: the leaf's own downward equivalent
// surface, alpha 2.95).  Adds grad/potential into trg_value.  tiles[2 i],
// tiles[2 i + 1] (i < ntiles) = (leaf, tile of 64 targets) of the same
// leaves: the P2P and L2P kernels of the real kernels take one tile per wave.
void launch_eval_targets(int kind, bool f32, int mode, const DevTree& t, Wavek k, const int* work, int nwork,
                         const int* tiles, int ntiles, hipStream_t s);

// Surface targets: P2M writes the upward check potential of each work leaf to
// check[slot] (slot = position in work); P2L adds the potential of the P2L list
// sources on the downward check surface (alpha 1.05) to dn[node].
void launch_eval_surface(int kind, bool f32, int mode, const DevTree& t, Wavek k, const int* work, int nwork,
                         double* check, hipStream_t s);

// Direct sum of ns sources onto nt targets (potential or potential+gradient);
// out has nt*(gradient?4:1) values of kind T and is accumulated into.
void launch_direct(int kind, bool mixed, bool cplx, Wavek k, const double* src, const double* q, int ns,
                   const double* trg, int nt, double* out, int gradient, double* scratch, int nsplit,
                   hipStream_t s);

// C[crow[i]] (beta=0: =, beta=1: +=) alpha * A[arow[i]] x B for i < m.
// A rows have K values, B is K x N row-major, C rows have N values (doubles;
// arow/crow are row offsets in doubles, not rows); alpha is a power of two.
// FP64 MFMA, or with f32 FP32 MFMA (operands rounded to FP32 while staged,
// FP32 accumulation, FP64 results; B float); when the output has few tiles
// the K range is split over workgroups, partial products go to `scratch` (at
// least gemm_scratch_doubles(m, n, k) doubles) and are reduced in a fixed
// order.  Complex GEMMs are called as real ones on the 2 x 2 block form.
void launch_gemm_rows(bool f32, int m, int n, int k, const double* A, const long* arow, const void* B, double* C,
                      const long* crow, double alpha, int beta, hipStream_t s, double* scratch = nullptr,
                      size_t scratch_doubles = 0);
// out[i] = float(in[i]) for i < n
void launch_to_float(const double* in, float* out, size_t n, hipStream_t s);
size_t gemm_scratch_doubles(int m, int n, int k);

// M2L stages.
struct M2LDev {
  int p = 0, n1 = 0, nfreq = 0, nsurf = 0;
  bool cplx = false;                 // Helmholtz: complex c2c transforms
  bool mixed = false;                // FP32 spectra and Hadamard
  bool mixed_fft = false;            // FP32 transform arithmetic too (else FP64 arithmetic, FP32 storage)
  const int* map_up = nullptr;       // [nsurf] local index x + p (y + p z) in the [0,p)^3 input cube
  const int* map_dn = nullptr;       // [nsurf] local index in the [p,2p)^3 output cube (offset p removed)
  const int* inv_dn = nullptr;       // [p^3] inverse of map_dn: surface index of a cube point, or -1
  void* fft_in = nullptr;            // [nsrc_slots][nfreq][8] complex
  void* fft_out = nullptr;           // [ntrg_slots][nfreq][8] complex
  unsigned* amax = nullptr;          // split-FP16 Hadamard: largest |spectrum value| (float bits), else null
};

// fft_up_equiv: one (source slot, child) per work item; src_node[slot] gives
// the node whose 8 children are transformed.
void launch_fft_up(const M2LDev& m, const DevTree& t, const int* src_node, int nslots, hipStream_t s);
// hadamard_product over blocks of <= 64 M2L targets of one operator level.
// Block b: trg[b][64] target slots (-1 = padding), uni[b][HAD_UCAP] source
// slots of the union of the block's sources (nunion[b] used), loc[b][26][64]
// local union index of the source of target t for relative position r (-1 =
// none).
// Union slots of a block: slot = 16 row + colour, colour = (x + 4 y + 8 z) mod 16
// of the source's lattice position at its level, at most HAD_UROWS rows.
// Targets sit at lane (x mod 4) + 4 (y mod 2) + 8 (z mod 2) of tile
// (y / 2 mod 2) + 2 (z / 2 mod 2) where that is free, so that the 16 sources a
// wave gathers for one relative position fall on 16 different colours (LDS
// banks).
constexpr int HAD_UROWS = 18;
constexpr int HAD_UCAP = 16 * HAD_UROWS;  // union slots per block (LDS staging)
constexpr int HAD_KCH = 16;    // default frequencies per workgroup (tunable per context)
struct HadBlocks {
  int nblk = 0;
  const int* trg = nullptr;
  const int* uni = nullptr;
  const int* nunion = nullptr;
  const short* loc = nullptr;
};
// m2l_opsx: operators of the level in the expanded MFMA form of launch_m2l_expand
// hscale != null (mixed, with m.amax set): split-FP16 operators of
// launch_m2l_expand with their per-frequency scales, spectra scaled by *m.amax
// kch: frequencies per workgroup (<= 0: HAD_KCH); groups: 64-target blocks
// per workgroup of the split-FP16 kernel (1, 2 or 4; they share the staged
// operators of a frequency); tiles: 16-target tiles per wave of that kernel
// (1 or 2; a wave's tiles share the operator operands it reads)
void launch_hadamard(const M2LDev& m, const void* m2l_opsx, const float* hscale, const HadBlocks& hb, int kch,
                     int groups, int tiles, hipStream_t s);
// Expand nlev operator levels from the reference layout [L][26][nfreq][64]
// complex into the per-lane MFMA A-operand form [L][nfreq][26][64 lanes][4]
// (float for the mixed path, else double), frequencies in the order of the
// transforms that m selects.
// With hscale != null (mixed only) the split-FP16 form of k_hadamard_h
// instead (same size), and hscale[L][k] receives the operator scales.
void launch_m2l_expand(const M2LDev& m, int nlev, const void* M, void* X, float* hscale, hipStream_t s);
// ifft_dn_check: one (target slot, child) per work item, adds
// scale[slot] * result to the child's dn vector.
void launch_ifft_dn(const M2LDev& m, const DevTree& t, const int* trg_node, const double* scale, int nslots,
                    hipStream_t s);

// Bytes of dynamic LDS the transform kernels need for this configuration.
size_t m2l_lds_bytes(const M2LDev& m);

void launch_zero(double* p, size_t n, hipStream_t s);
// {Re sum q, Im sum q, sum |q|} of n charges (complex: interleaved pairs),
// in a fixed order: per-block partial sums in out[0 .. 3 CHARGE_SUM_BLOCKS),
// their total in out[3 CHARGE_SUM_BLOCKS .. + 3)
constexpr int CHARGE_SUM_BLOCKS = 256;
void launch_charge_sums(const double* q, int n, bool cplx, double* out, hipStream_t s);
// values T[n][4]: value 0 (potential) times scale[i][0], values 1..3
// (gradient) times scale[i][1]; vw doubles per value
void launch_scale_values(double* val, const double* scale, int n, int vw, hipStream_t s);
void launch_scatter_values(const double* sorted_val, const int* perm, int n, int vw, double* out, hipStream_t s);

}  // namespace xfmm
