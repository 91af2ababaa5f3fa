// lsqt_launch.h -- host-side launchers of the device kernels (internal).
#pragma once

#include "lsqt_device.h"

namespace lsqt {

// FP64 drop-in path (kernels_f64.hip)
// tiles (moment partials per step) of the fused step for n sites and K
// slots, and its sites per thread
int hpsi_tiles(int n, int K);
int hpsi_spt(int K);
void launch_hpsi(const HamDev& h, const double2* psi1, const double2* psi0, double2* out,
                 const double2* L, double2* partial, bool recur, bool ip, hipStream_t st);
void launch_xop(const HamDev& h, const double2* in, double2* out, bool current, hipStream_t st);
// one expansion term of U (of [X, U]): s2 = 2 H~ s1 - s0 (and s2x) with
// s += bm (label phase) s2 (s2x); bm_prev / label_prev add the previous
// launch's term (s1, s1x) first, label 0 adds nothing, and with both 0 s is
// neither read nor written (terms added in pairs: deferred_term)
void launch_cheb2(const HamDev& h, const double2* s0, const double2* s1, double2* s2,
                  double2* s, double bm, int label, hipStream_t st, double bm_prev = 0.0,
                  int label_prev = 0);
void launch_cheb2x(const HamDev& h, const double2* s0, const double2* s0x, const double2* s1,
                   const double2* s1x, double2* s2, double2* s2x, double2* s, double bm,
                   int label, hipStream_t st, double bm_prev = 0.0, int label_prev = 0);
// the irregular-site mask of h's slot pattern (h.pat, h.pat_p set): the
// mask holds pattern_mask_words(h.n) words
int pattern_mask_words(int n);
void launch_pattern_mask(const HamDev& h, uint32_t* mask, hipStream_t st);
void launch_zero(double2* v, int n, hipStream_t st);
void launch_copy(double2* dst, const double2* src, int n, hipStream_t st);
void launch_add(double2* dst, const double2* src, int n, hipStream_t st);
void launch_sz(double2* dst, const double2* src, int n, hipStream_t st);
void launch_cheb01(const double2* s0, const double2* s1, double2* s, double b0, double b1,
                   int direction, int n, hipStream_t st);
void launch_cheb1x(const double2* s1x, double2* s, double b1, int n, hipStream_t st);
// partials of conj(L) psi over tiles of spt * 256 sites: spt 1 gives
// dot_tiles(n) tiles, hpsi_spt(K) the tiles of the fused step (the moment
// mu_0 of a moment set)
int dot_tiles(int n);
void launch_dot(const double2* psi, const double2* L, double2* partial, int n, hipStream_t st,
                int spt = 1);
void launch_reduce_partials(const double2* partial, size_t stride, int tiles, int count,
                            double2* mu, hipStream_t st);

// Two Chebyshev steps per launch on slab lattices (kernels_sweep.h): from
// psi2 = psi_{m-2} and psi1 = psi_{m-1}, out0 = psi_m and out1 = psi_{m+1}
// (four distinct buffers), with the tile partials of <L|psi_m> in part0 and
// of <L|psi_{m+1}> in part1 (sm.tiles each)
void launch_sweep2(const HamDev& h, const SweepMap& sm, const double2* psi2,
                   const double2* psi1, double2* out0, double2* out1, const double2* L,
                   double2* part0, double2* part1, hipStream_t st);
void launch_sweep2_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                      const float2* psi1, float2* out0, float2* out1, const float2* L,
                      double2* part0, double2* part1, hipStream_t st);
// NS = sm.ns steps per launch (kernels_sweepn.h): from psi2 = psi_{m-2} and
// psi1 = psi_{m-1}, out0 = psi_{m+NS-2} and out1 = psi_{m+NS-1} (four
// distinct buffers), the tile partials of <L|psi_{m+l}> in part + l *
// part_stride (sm.tiles each, l < NS); sweepn_lds: its static LDS bytes
void launch_sweepn(const HamDev& h, const SweepMap& sm, const double2* psi2,
                   const double2* psi1, double2* out0, double2* out1, const double2* L,
                   double2* part, size_t part_stride, hipStream_t st);
void launch_sweepn_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                      const float2* psi1, float2* out0, float2* out1, const float2* L,
                      double2* part, size_t part_stride, hipStream_t st);
size_t sweepn_lds(bool f32, int ns, int nb);
// sm.ns terms of U(t) per launch: from psi2 = phi_{k-2}, psi1 = phi_{k-1},
// out0 = phi_{k+ns-2}, out1 = phi_{k+ns-1}, term l added into the FP64 state
// S with terms.bm[l] / label[l], in order (the additions of one-term launches)
void launch_sweepn_evo(const HamDev& h, const SweepMap& sm, const double2* psi2,
                       const double2* psi1, double2* out0, double2* out1, double2* S,
                       const SweepTerms& terms, hipStream_t st);
void launch_sweepn_evo_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                          const float2* psi1, float2* out0, float2* out1, double2* S,
                          const SweepTerms& terms, hipStream_t st);
// the check of a sweep table against every packed word: one int per
// workgroup of sweep_check_blocks(n), nonzero where a word differs
int sweep_check_blocks(int n);
void launch_sweep_check(const HamDev& h, const SweepMap& sm, int* out, hipStream_t st);

// FP32-storage path (kernels_f32.hip)
int hpsi32_tiles(int n);
void launch_hpsi32(const HamDev& h, const float2* psi1, const float2* psi0, float2* out,
                   const float2* L, double2* partial, bool recur, bool ip, hipStream_t st);
void launch_comm32(const HamDev& h, const float2* in, float2* out, hipStream_t st);
// the same with FP32 recursion vectors (FP64 state s)
void launch_cheb2_32(const HamDev& h, const float2* s0, const float2* s1, float2* s2,
                     double2* s, double bm, int label, hipStream_t st, double bm_prev = 0.0,
                     int label_prev = 0);
void launch_cheb2x_32(const HamDev& h, const float2* s0, const float2* s0x, const float2* s1,
                      const float2* s1x, float2* s2, float2* s2x, double2* s, double bm,
                      int label, hipStream_t st, double bm_prev = 0.0, int label_prev = 0);
void launch_zero32(float2* v, int n, hipStream_t st);
// packed slot words of a lattice model (class m*K + i for hopping i of orbital m)
void launch_lattice_words(const LatticeDev& L, uint32_t* words, int stride, int idx_bits,
                          int Kmax, int n_pristine, hipStream_t st);
void launch_to_f32(const double2* src, float2* dst, int n, hipStream_t st);
void launch_to_f64(const float2* src, double2* dst, int n, hipStream_t st);
void launch_cheb01_32(const float2* s1, double2* s, double b0, double b1, int direction, int n,
                      hipStream_t st);
void launch_cheb1x_32(const float2* s1x, double2* s, double b1, int n, hipStream_t st);
void launch_dot32(const float2* psi, const float2* L, double2* partial, int n, hipStream_t st);
void launch_norm2(const double2* s, double2* partial, int n, hipStream_t st);
void launch_renormalise(double2* s, const double2* norms, int n, hipStream_t st);

// block path (kernels_block.hip)
// walk_pass > 0 asks for the column walk with passes of walk_pass sites
TileMap make_tile_map(int n, int rows, int slab, int span, int row, int walk_pass);
// Sliding-window schedule (block steps only): tiles of `by` rows of one chunk
// of kBlock / (R / W) sites; the other block kernels see natural tiles of
// by * chunk sites (the same tile count).
TileMap make_window_map(int n, int slab, int row, int by, int chunk);
// whether the window schedule applies to this lattice, tile height and shape
bool window_fits(int n, int slab, int row, int by, int R, int W);
Grid block_grid(const TileMap& t);
// Moment products of a block step (values of the kernels' BlockProducts):
// kStepLeft <L|out> into partial; kStepSelf <psi1|out> into partial and
// <out|out> into partial2 (L unused); kStepPair: the R columns are R/2
// pairs (2p from a left, 2p + 1 from a right recursion), <psi1_2p|out_2p+1>
// into partial and <out_2p|out_2p+1> into partial2, [tile][R/2] (W >= 2)
enum StepProducts { kStepLeft = 1, kStepSelf = 2, kStepPair = 3 };
// db: double-buffered pass loop (compile-time slot counts only)
void launch_block_step(const HamDev& h, int R, int W, bool db, const TileMap& tm,
                       const float2* psi1, const float2* psi0, float2* out, const float2* L,
                       double2* partial, double2* partial2, bool recur, int products,
                       hipStream_t st);
// partial[t][r] = tile sums of conj(L_r) psi_r; pairs: R columns as R/2
// pairs, partial[t][p] = conj(psi_2p) psi_2p+1 (L unused)
void launch_block_dot(const float2* psi, const float2* L, int R, const TileMap& tm,
                      double2* partial, hipStream_t st, bool pairs = false);
// dst ([n][2R]) = L and Rb ([n][R] each) as R pairs of columns (2r, 2r + 1)
void launch_block_interleave(const float2* L, const float2* Rb, int n, int R, float2* dst,
                             hipStream_t st);
// block time evolution (op: 0 H~, 1 Chebyshev term of U, 2 [X, H~], 3 V,
// 4 Chebyshev term of [X, U]); S is the FP64 accumulated state [n][R].
// Ops 1 and 4 add bm_prev * (label_prev phase of s1, or s1x: the previous
// launch's term) and then bm * (label phase of the new s2, or s2x) to S;
// label 0 adds nothing, and with both 0 S is neither read nor written.
void launch_block_evo(const HamDev& h, int R, int W, const TileMap& tm, int op,
                      const float2* s0, const float2* s1, float2* s2, const float2* s0x,
                      const float2* s1x, float2* s2x, double2* S, double bm, int label,
                      hipStream_t st, double bm_prev = 0.0, int label_prev = 0);
// partial[t][r] = (|S_r|^2 over tile t, 0)
void launch_block_norm64(const double2* S, int R, const TileMap& tm, double2* partial,
                         hipStream_t st);
// S_r *= sqrt(before[k][r] / after[k][r]) in segment k
void launch_block_renorm(double2* S, int R, const TileMap& tm, const double2* before,
                         const double2* after, hipStream_t st);
// vector-major host layout ([R][n] re and im) <-> site-major blocks ([n][R]);
// n * R < 2^31
void launch_site_major(const double* re, const double* im, int n, int R, double2* dst,
                       hipStream_t st);
void launch_site_major(const double* re, const double* im, int n, int R, float2* dst,
                       hipStream_t st);
void launch_vector_major(const float2* src, int n, int R, double* re, double* im,
                         hipStream_t st);
// dst = S_z src (sign of the odd sites flipped) for every vector of a block
void launch_block_sz(const float2* src, int n, int R, float2* dst, hipStream_t st);
// mu[(k*R + r)*M + m0 + i] for segment k, i < count (partials of moment
// m0 + i at slot i; slots stride_tiles * R apart, tm.tiles * R where 0)
void launch_block_reduce(const double2* partial, const TileMap& tm, int R, int M, int m0,
                         int count, double2* mu, hipStream_t st, int stride_tiles = 0);
// Two steps per launch on the block path (kernels_block_sweep.h): from p2 =
// psi_{m-2} and p1 = psi_{m-1} of `cols` columns, o0 = psi_m and o1 =
// psi_{m+1} (four distinct buffers) with the self products (pairs: the pair
// products of cols / 2 pairs) of both steps: part[0], part[1] of step 1,
// part[2], part[3] of step 2 (sm.tiles partials each).  W = 2 for cols <= 16,
// 4 for cols 32 and 64; cols / W lanes per site.
size_t block_sweep_lds(int W);
void launch_block_sweep2(const HamDev& h, const BlockSweepMap& sm, int cols, int W, bool pairs,
                         const float2* p2, const float2* p1, float2* o0, float2* o1,
                         double2* const (&part)[4], hipStream_t st);
// Four steps per launch on the block path (k_block_sweepn, cols <= 16):
// from p2 = psi_{m-2} and p1 = psi_{m-1}, o0 = psi_{m+2} and o1 = psi_{m+3};
// level l = 1 .. 4 (psi_{m+l-1}) adds its two products to part + (2 l - 2)
// part_stride and part + (2 l - 1) part_stride (sm.tiles partials each)
size_t block_sweepn_lds();
void launch_block_sweepn(const HamDev& h, const BlockSweepMap& sm, int cols, bool pairs,
                         const float2* p2, const float2* p1, float2* o0, float2* o1,
                         double2* part, size_t part_stride, hipStream_t st);
// four terms of U(t) per launch on the block path (cols <= 16): from p2 =
// phi_{k-2}, p1 = phi_{k-1}, o0 = phi_{k+2}, o1 = phi_{k+3}, term l added
// into the FP64 state blocks with terms.bm[l] / label[l], in order
void launch_block_sweepn_evo(const HamDev& h, const BlockSweepMap& sm, int cols,
                             const float2* p2, const float2* p1, float2* o0, float2* o1,
                             double2* state, const SweepTerms& terms, hipStream_t st);
// the check of one system's sweep table against its packed words: one int
// per workgroup of block_sweep_check_blocks(g), nonzero where a word differs
int block_sweep_check_blocks(const BlockSweepSeg& g);
void launch_block_sweep_check(const HamDev& h, const BlockSweepSeg& g, const int4* tab, int* out,
                              hipStream_t st);

}  // namespace lsqt
