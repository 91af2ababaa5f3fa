// kernels_f64.hip -- FP64 kernels of the drop-in path for gfx950.
//
// Every kernel reproduces the per-site arithmetic of the reference CPU
// build in the same order -- on-site term first, neighbours in slot order,
// then the division by energy_max, then the recursion / accumulation
// (src/hamiltonian.cu:188-221, 273-304, 354-383, 425-443, 531-595, 635-648,
// 760-852, 927-966) -- and this file is compiled with -ffp-contract=off, so
// the state vectors are bit-identical to the reference's.
//
// Data layout: complex vectors are interleaved double2 (one 16-byte load per
// neighbour gather); ELL slots are column-major so that slot m of 256
// consecutive sites is one coalesced load; a packed pair word carries the
// neighbour index and the index of its (re, im, xx) triple.  One workgroup of
// 256 threads owns a tile of 256*SPT consecutive sites, tiles are assigned so
// that each XCD sweeps one contiguous range (neighbour gathers then hit that
// XCD's L2), and the moment product of a Chebyshev step is fused into the
// step: per-tile partial sums, reduced by a fixed wave64 shuffle tree, are
// summed per moment in tile order at the end.  Nothing uses atomics, so
// every result is run-to-run deterministic.
#include "kernels_sweep.h"
#include "kernels_sweepn.h"
#include "lsqt_device.h"
#include "lsqt_launch.h"

namespace lsqt {

// --- class tables in LDS -------------------------------------------------
template <int LAYOUT, bool WITH_XX>
__device__ __forceinline__ void stage_table(const HamDev& h, double2* s_tab, double* s_xx)
{
  if constexpr (LAYOUT == kPackedLds) {
    for (int i = threadIdx.x; i < h.ncls; i += kBlock) {
      s_tab[i] = h.tab[i];
      if constexpr (WITH_XX)
        s_xx[i] = h.tab_xx[i];
    }
    __syncthreads();
  }
}

// --- H~-type sums over the slots of site n, in slot order -----------------
// hsum:  t += sum_m h_nm psi[j_m]                 (apply, kernel_polynomial)
// xsum:  t -= sum_m h_nm xx_nm psi[j_m]  (or += for the current operator)
// Every slot's loads are issued before any is used; the adds are predicated
// so the skipped slots leave the sums bit-identical to a loop that stops at
// neighbor_number[n] (src/hamiltonian.cu:206-215).
template <int LAYOUT, int KT, bool NT, bool PAT = false>
__device__ __forceinline__ void hsum(const HamDev& h, const double2* s_tab,
                                     const double2* __restrict__ psi, int n, double& tr,
                                     double& ti, const int2* s_pat = nullptr)
{
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  for (int m0 = 0; m0 < K; m0 += C) {
    int j[C];
    double a[C], b[C], x[C];
    bool ok[C];
    load_slots<LAYOUT, C, NT, false, PAT>(h, s_tab, nullptr, n, m0, K, j, a, b, x, ok, s_pat);
    double2 c[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      c[q] = psi[j[q]];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        tr += a[q] * c[q].x - b[q] * c[q].y;
        ti += a[q] * c[q].y + b[q] * c[q].x;
      }
    }
  }
}

// --- deterministic workgroup reduction of a complex partial -------------
__device__ __forceinline__ void block_sum_store(double re, double im, double2* out)
{
  __shared__ double s_re[kBlock / 64], s_im[kBlock / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    re += __shfl_xor(re, off, 64);
    im += __shfl_xor(im, off, 64);
  }
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    s_re[wave] = re;
    s_im[wave] = im;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r = s_re[0], i = s_im[0];
#pragma unroll
    for (int w = 1; w < kBlock / 64; ++w) {
      r += s_re[w];
      i += s_im[w];
    }
    *out = make_double2(r, i);
  }
}

// --- H~ psi1, optionally 2 H~ psi1 - psi0, optionally the moment product --
// RECUR: out = 2 H~ psi1 - psi0 (kernel_polynomial) else out = H~ psi1 (apply)
// IP:    partial[tile] = sum over the tile of conj(L) out
// out may alias psi0 (each site reads its own psi0 before writing out).
// With a slot pattern (packed layouts) the pattern table is staged in LDS and
// regular sites compute their slot words instead of loading them.
template <int LAYOUT, int KT, int SPT, bool RECUR, bool IP, bool NT>
__global__ __launch_bounds__(kBlock) void k_hpsi(HamDev h, const double2* __restrict__ psi1,
                                                 const double2* psi0, double2* out,
                                                 const double2* __restrict__ L,
                                                 double2* __restrict__ partial, int tiles,
                                                 int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ int2 s_pat[LAYOUT == kFull ? 1 : kPatternMaxEntries];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage_table<LAYOUT, false>(h, s_tab, nullptr);
  const bool pat = LAYOUT != kFull && h.pat != nullptr;  // the same for the whole launch
  if (pat) {
    for (int i = threadIdx.x; i < h.pat_p * h.K; i += kBlock)
      s_pat[i] = h.pat[i];
    __syncthreads();
  }
  double acc_re = 0.0, acc_im = 0.0;
  if constexpr (KT > 0) {
    // all loads of the thread's SPT sites first (out may alias psi0, so the
    // compiler would not move a later site's loads above an earlier store),
    // then the same per-site arithmetic as below
    constexpr int C = KT;
    double2 p1[SPT], p0[SPT], l[SPT], c[SPT][C];
    double v[SPT], a[SPT][C], b[SPT][C], x[SPT][C];
    int j[SPT][C];
    bool ok[SPT][C];
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int nk = (tile * SPT + k) * kBlock + threadIdx.x;
      const int n = nk < h.n ? nk : h.n - 1;  // an in-bounds site; its results are dropped
      p1[k] = psi1[n];
      v[k] = ld<NT>(h.pot + n);
      if constexpr (RECUR)
        p0[k] = ld2<NT>(psi0 + n);
      if constexpr (IP)
        l[k] = ld2<NT>(L + n);
      if (pat) {
        // the pattern's words for every site, so that the gathers issue
        // without waiting for the mask (indices clamped into range); the
        // irregular sites load their words and gather again below (measured:
        // 93.5 us per step against 99.9 us with the mask read first).  At an
        // irregular site n + offset may leave [0, 2^idx_bits) and carry into
        // the class bits, so the class index is clamped into the table too
        // (the value read there is replaced before it is used)
        const int r = pattern_residue(h, n);
#pragma unroll
        for (int q = 0; q < C; ++q) {
          const uint32_t w = pattern_word(h, s_pat[q * h.pat_p + r], n);
          ok[k][q] = w != kEmptySlot;
          const int jj = ok[k][q] ? (int)(w & h.idx_mask) : n;
          j[k][q] = (unsigned)jj < (unsigned)h.n ? jj : n;
          const uint32_t cw = ok[k][q] ? (w >> h.idx_bits) : 0u;
          const int cc = cw < (uint32_t)h.ncls ? (int)cw : 0;
          const double2 t = (LAYOUT == kPackedLds) ? s_tab[cc] : h.tab[cc];
          a[k][q] = t.x;
          b[k][q] = t.y;
        }
      } else
        load_slots<LAYOUT, C, NT, false>(h, s_tab, nullptr, n, 0, C, j[k], a[k], b[k], x[k],
                                         ok[k]);
    }
#pragma unroll
    for (int k = 0; k < SPT; ++k)
#pragma unroll
      for (int q = 0; q < C; ++q)
        c[k][q] = psi1[j[k][q]];
    if (pat) {
#pragma unroll
      for (int k = 0; k < SPT; ++k) {
        const int nk = (tile * SPT + k) * kBlock + threadIdx.x;
        const int n = nk < h.n ? nk : h.n - 1;
        if (pattern_irregular(h, n)) {
          load_slots<LAYOUT, C, NT, false>(h, s_tab, nullptr, n, 0, C, j[k], a[k], b[k], x[k],
                                           ok[k]);
#pragma unroll
          for (int q = 0; q < C; ++q)
            c[k][q] = psi1[j[k][q]];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < SPT; ++k) {
      const int n = (tile * SPT + k) * kBlock + threadIdx.x;
      if (n < h.n) {
        double tr = v[k] * p1[k].x;
        double ti = v[k] * p1[k].y;
#pragma unroll
        for (int q = 0; q < C; ++q) {
          if (ok[k][q]) {
            tr += a[k][q] * c[k][q].x - b[k][q] * c[k][q].y;
            ti += a[k][q] * c[k][q].y + b[k][q] * c[k][q].x;
          }
        }
        tr /= h.emax;
        ti /= h.emax;
        if constexpr (RECUR) {
          tr = 2.0 * tr - p0[k].x;
          ti = 2.0 * ti - p0[k].y;
        }
        st2<false>(out + n, make_double2(tr, ti));
        if constexpr (IP) {
          acc_re += tr * l[k].x + ti * l[k].y;
          acc_im += ti * l[k].x - tr * l[k].y;
        }
      }
    }
    if constexpr (IP)
      block_sum_store(acc_re, acc_im, partial + tile);
    return;
  }
#pragma unroll
  for (int k = 0; k < SPT; ++k) {
    int n = (tile * SPT + k) * kBlock + threadIdx.x;
    if (n < h.n) {
      double2 p1 = psi1[n];
      double v = ld<NT>(h.pot + n);
      double tr = v * p1.x;
      double ti = v * p1.y;
      if (pat)
        hsum<LAYOUT, KT, NT, true>(h, s_tab, psi1, n, tr, ti, s_pat);
      else
        hsum<LAYOUT, KT, NT>(h, s_tab, psi1, n, tr, ti);
      tr /= h.emax;
      ti /= h.emax;
      if constexpr (RECUR) {
        double2 p0 = ld2<NT>(psi0 + n);
        tr = 2.0 * tr - p0.x;
        ti = 2.0 * ti - p0.y;
      }
      st2<false>(out + n, make_double2(tr, ti));
      if constexpr (IP) {
        double2 l = ld2<NT>(L + n);
        acc_re += tr * l.x + ti * l.y;
        acc_im += ti * l.x - tr * l.y;
      }
    }
  }
  if constexpr (IP)
    block_sum_store(acc_re, acc_im, partial + tile);
}

// --- [X, H~] psi and V psi ----------------------------------------------
template <int LAYOUT, int KT, bool CURRENT, bool NT>
__global__ __launch_bounds__(kBlock) void k_xop(HamDev h, const double2* __restrict__ in,
                                                double2* __restrict__ out, int tiles,
                                                int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ double s_xx[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage_table<LAYOUT, true>(h, s_tab, s_xx);
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  int n = tile * kBlock + threadIdx.x;
  if (n >= h.n)
    return;
  double tr = 0.0, ti = 0.0;
  for (int m0 = 0; m0 < K; m0 += C) {
    int j[C];
    double a[C], b[C], x[C];
    bool ok[C];
    load_slots<LAYOUT, C, NT, true>(h, s_tab, s_xx, n, m0, K, j, a, b, x, ok);
    double2 c[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      c[q] = in[j[q]];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        if constexpr (CURRENT) {
          tr += (a[q] * c[q].x - b[q] * c[q].y) * x[q];
          ti += (a[q] * c[q].y + b[q] * c[q].x) * x[q];
        } else {
          tr -= (a[q] * c[q].x - b[q] * c[q].y) * x[q];
          ti -= (a[q] * c[q].y + b[q] * c[q].x) * x[q];
        }
      }
    }
  }
  if constexpr (CURRENT)
    st2<false>(out + n, make_double2(+ti, -tr));
  else
    st2<false>(out + n, make_double2(tr / h.emax, ti / h.emax));
}

// the four (-i)^m phase cases of src/hamiltonian.cu:504-525
__device__ __forceinline__ void accumulate_label(double2& s, double bm, double tr, double ti,
                                                 int label)
{
  switch (label) {
    case 1:
      s.x += bm * tr;
      s.y += bm * ti;
      break;
    case 2:
      s.x -= bm * tr;
      s.y -= bm * ti;
      break;
    case 3:
      s.x += bm * ti;
      s.y -= bm * tr;
      break;
    case 4:
      s.x -= bm * ti;
      s.y += bm * tr;
      break;
  }
}

// --- one term of U(t): s2 = 2 H~ s1 - s0, s += b (phase) s2 -------------
// bm_prev / label_prev: the previous launch's term, s1, added first where
// terms are added in pairs (label 0 adds nothing; with both 0 the state s is
// neither read nor written): the same additions, in the same order
template <int LAYOUT, int KT, bool NT>
__global__ __launch_bounds__(kBlock) void k_cheb2(HamDev h, const double2* s0,
                                                  const double2* __restrict__ s1, double2* s2,
                                                  double2* __restrict__ s, double bm, int label,
                                                  double bm_prev, int label_prev, int tiles,
                                                  int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage_table<LAYOUT, false>(h, s_tab, nullptr);
  int n = tile * kBlock + threadIdx.x;
  if (n >= h.n)
    return;
  double2 p1 = s1[n];
  double v = ld<NT>(h.pot + n);
  double tr = v * p1.x;
  double ti = v * p1.y;
  hsum<LAYOUT, KT, NT>(h, s_tab, s1, n, tr, ti);
  tr /= h.emax;
  ti /= h.emax;
  double2 p0 = ld2<NT>(s0 + n);
  tr = 2.0 * tr - p0.x;
  ti = 2.0 * ti - p0.y;
  if (label != 0 || label_prev != 0) {
    double2 st = ld2<NT>(s + n);
    accumulate_label(st, bm_prev, p1.x, p1.y, label_prev);
    accumulate_label(st, bm, tr, ti, label);
    st2<false>(s + n, st);
  }
  st2<false>(s2 + n, make_double2(tr, ti));
}

// --- one term of [X, U(t)]: coupled recursion of s and sx ----------------
template <int LAYOUT, int KT, bool NT>
__global__ __launch_bounds__(kBlock) void k_cheb2x(HamDev h, const double2* s0,
                                                   const double2* s0x,
                                                   const double2* __restrict__ s1,
                                                   const double2* __restrict__ s1x, double2* s2,
                                                   double2* s2x, double2* __restrict__ s,
                                                   double bm, int label, double bm_prev,
                                                   int label_prev, int tiles, int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ double s_xx[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage_table<LAYOUT, true>(h, s_tab, s_xx);
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  int n = tile * kBlock + threadIdx.x;
  if (n >= h.n)
    return;
  double v = ld<NT>(h.pot + n);
  double2 p1 = s1[n], p1x = s1x[n];
  double tr = v * p1.x, ti = v * p1.y;
  double txr = v * p1x.x, txi = v * p1x.y;
  for (int m0 = 0; m0 < K; m0 += C) {
    int j[C];
    double a[C], b[C], x[C];
    bool ok[C];
    load_slots<LAYOUT, C, NT, true>(h, s_tab, s_xx, n, m0, K, j, a, b, x, ok);
    double2 c[C], cx[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      c[q] = s1[j[q]];
      cx[q] = s1x[j[q]];
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        tr += a[q] * c[q].x - b[q] * c[q].y;
        ti += a[q] * c[q].y + b[q] * c[q].x;
        txr += a[q] * cx[q].x - b[q] * cx[q].y;
        txi += a[q] * cx[q].y + b[q] * cx[q].x;
        txr -= (a[q] * c[q].x - b[q] * c[q].y) * x[q];
        txi -= (a[q] * c[q].y + b[q] * c[q].x) * x[q];
      }
    }
  }
  tr /= h.emax;
  ti /= h.emax;
  double2 p0 = ld2<NT>(s0 + n);
  tr = 2.0 * tr - p0.x;
  ti = 2.0 * ti - p0.y;
  st2<false>(s2 + n, make_double2(tr, ti));
  txr /= h.emax;
  txi /= h.emax;
  double2 p0x = ld2<NT>(s0x + n);
  txr = 2.0 * txr - p0x.x;
  txi = 2.0 * txi - p0x.y;
  st2<false>(s2x + n, make_double2(txr, txi));
  if (label != 0 || label_prev != 0) {
    double2 st = ld2<NT>(s + n);
    accumulate_label(st, bm_prev, p1x.x, p1x.y, label_prev);
    accumulate_label(st, bm, txr, txi, label);
    st2<false>(s + n, st);
  }
}

// --- element-wise kernels -------------------------------------------------
__global__ __launch_bounds__(kBlock) void k_zero(double2* v, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    v[i] = make_double2(0.0, 0.0);
}

__global__ __launch_bounds__(kBlock) void k_copy(double2* __restrict__ dst,
                                                 const double2* __restrict__ src, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    dst[i] = src[i];
}

__global__ __launch_bounds__(kBlock) void k_add(double2* __restrict__ dst,
                                                const double2* __restrict__ src, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double2 d = dst[i], s = src[i];
    d.x += s.x;
    d.y += s.y;
    dst[i] = d;
  }
}

__global__ __launch_bounds__(kBlock) void k_sz(double2* __restrict__ dst,
                                               const double2* __restrict__ src, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double2 s = src[i];
    dst[i] = (i % 2 == 0) ? s : make_double2(-s.x, -s.y);
  }
}

// state = b0 s0 + b1 dir (Im s1, -Re s1)   (src/hamiltonian.cu:404-423)
__global__ __launch_bounds__(kBlock) void k_cheb01(const double2* __restrict__ s0,
                                                   const double2* __restrict__ s1,
                                                   double2* __restrict__ s, double b0,
                                                   double b1, int direction, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double bessel_1 = b1 * direction;
    double2 a = s0[i], c = s1[i];
    s[i] = make_double2(b0 * a.x + bessel_1 * c.y, b0 * a.y - bessel_1 * c.x);
  }
}

// state = b1 (Im s1x, -Re s1x)   (src/hamiltonian.cu:619-633)
__global__ __launch_bounds__(kBlock) void k_cheb1x(const double2* __restrict__ s1x,
                                                   double2* __restrict__ s, double b1, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double2 c = s1x[i];
    s[i] = make_double2(+b1 * c.y, -b1 * c.x);
  }
}

// partial[tile] = sum over the tile of conj(L) psi
// partial[tile] = sum over tile of conj(L) psi, tiles of spt * kBlock sites
// (the tiling of the fused step whose partials it joins)
__global__ __launch_bounds__(kBlock) void k_dot(const double2* __restrict__ psi,
                                                const double2* __restrict__ L,
                                                double2* __restrict__ partial, int n, int spt,
                                                int tiles, int blocks)
{
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  double re = 0.0, im = 0.0;
  for (int k = 0; k < spt; ++k) {
    const int i = (tile * spt + k) * kBlock + threadIdx.x;
    if (i < n) {
      double2 p = psi[i], l = L[i];
      re += p.x * l.x + p.y * l.y;
      im += p.y * l.x - p.x * l.y;
    }
  }
  block_sum_store(re, im, partial + tile);
}

// mu[m] = sum_t partial[m * stride + t], t = 0..tiles-1, in a fixed order
__global__ __launch_bounds__(kBlock) void k_reduce_partials(const double2* __restrict__ partial,
                                                            size_t stride, int tiles,
                                                            double2* __restrict__ mu)
{
  const double2* p = partial + (size_t)blockIdx.x * stride;
  double re = 0.0, im = 0.0;
  for (int t = threadIdx.x; t < tiles; t += kBlock) {
    double2 v = p[t];
    re += v.x;
    im += v.y;
  }
  block_sum_store(re, im, mu + blockIdx.x);
}

// --- the irregular-site mask of a slot pattern -----------------------------
// bit n of mask: site n's words differ from the pattern in some slot; one
// wave64 writes the two mask words of its 64 sites.  Sites past h.n (the
// padding that two-site kernels may touch) are marked irregular, so their
// words (all empty) are loaded, never computed.
__global__ __launch_bounds__(kBlock) void k_pattern_mask(HamDev h, uint32_t* __restrict__ mask)
{
  const int n = blockIdx.x * kBlock + threadIdx.x;
  bool irr = n >= h.n;
  if (n < h.n) {
    const int r = pattern_residue(h, n);
    for (int m = 0; m < h.K; ++m)
      irr |= pattern_word(h, h.pat[m * h.pat_p + r], n) != h.words[(size_t)m * h.stride + n];
  }
  const unsigned long long bits = __ballot(irr);
  if ((threadIdx.x & 63) == 0) {
    const int w = (blockIdx.x * kBlock + threadIdx.x) >> 5;
    mask[w] = (uint32_t)bits;
    mask[w + 1] = (uint32_t)(bits >> 32);
  }
}

// Check of the two-step sweep's periodic slot table (sm.tab): per workgroup,
// out[b] = 1 where the packed word of some slot of one of its sites differs
// from the word the table gives (neighbour at the periodic slab and
// position offsets, class, or an unused slot).
__global__ __launch_bounds__(kBlock) void k_sweep_check(HamDev h, SweepMap sm, int* __restrict__ out)
{
  __shared__ int s_bad[kBlock / 64];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int bad = 0;
  if (i < h.n) {
    const int x = div_by(i, sm.sdiv), pos = i - x * sm.S, o = pos % sm.p;
    for (int q = 0; q < h.K; ++q) {
      const uint32_t w = h.words[(size_t)q * h.stride + i];
      const int4 e = sm.tab[q * sm.p + o];
      if (!e.w) {
        bad |= w != kEmptySlot;
        continue;
      }
      int xs = x + e.x;
      xs = xs < 0 ? xs + sm.nslab : (xs >= sm.nslab ? xs - sm.nslab : xs);
      int ps = pos + e.y;
      ps = ps < 0 ? ps + sm.S : (ps >= sm.S ? ps - sm.S : ps);
      const uint32_t want = ((uint32_t)e.z << h.idx_bits) | (uint32_t)(xs * sm.S + ps);
      bad |= w != want;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
    bad |= __shfl_xor(bad, o, 64);
  if ((threadIdx.x & 63) == 0)
    s_bad[threadIdx.x >> 6] = bad;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / 64; ++w)
      bad |= s_bad[w];
    out[blockIdx.x] = bad;
  }
}

// ==========================================================================
// launchers
// ==========================================================================
namespace {

inline int blocks_1d(int n)
{
  return (n + kBlock - 1) / kBlock;
}

}  // namespace

#define LSQT_LAYOUT_SWITCH(layout, L, ...)        \
  switch (layout) {                               \
    case kFull: {                                 \
      constexpr int L = kFull;                    \
      __VA_ARGS__;                                \
    } break;                                      \
    case kPackedLds: {                            \
      constexpr int L = kPackedLds;               \
      __VA_ARGS__;                                \
    } break;                                      \
    default: {                                    \
      constexpr int L = kPackedGlobal;            \
      __VA_ARGS__;                                \
    } break;                                      \
  }

#define LSQT_K_SWITCH(kval, KT, ...) \
  switch (kval) {                    \
    case 2: {                        \
      constexpr int KT = 2;          \
      __VA_ARGS__;                   \
    } break;                         \
    case 3: {                        \
      constexpr int KT = 3;          \
      __VA_ARGS__;                   \
    } break;                         \
    case 4: {                        \
      constexpr int KT = 4;          \
      __VA_ARGS__;                   \
    } break;                         \
    case 6: {                        \
      constexpr int KT = 6;          \
      __VA_ARGS__;                   \
    } break;                         \
    default: {                       \
      constexpr int KT = 0;          \
      __VA_ARGS__;                   \
    } break;                         \
  }

#define LSQT_NT_SWITCH(flag, NT, ...) \
  if (flag) {                          \
    constexpr bool NT = true;          \
    __VA_ARGS__;                       \
  } else {                             \
    constexpr bool NT = false;         \
    __VA_ARGS__;                       \
  }

// sites per thread of the fused step, by slot count.  Measured on the
// 8.0M-site honeycomb (3 slots) in one process: 1 site 102.6 us per step, 2
// sites 102.2 us; with the slot pattern 105.2 and 100.4 us, and with the
// pattern's gathers issued before the mask read 2 sites 93.8 us, 3 sites
// 100.0 us, 4 sites 108.2 us.  On the 160^3 cube (6 slots) 2 sites were 9 %
// slower than 1.  (bench.py names the profiled kernel k_hpsi<1, 3, 2, ...>
// after this choice.)
constexpr int hpsi_spt_for(int K)
{
  return K == 3 ? 2 : 1;
}

int hpsi_spt(int K)
{
  return hpsi_spt_for(K);
}

void launch_hpsi(const HamDev& h, const double2* psi1, const double2* psi0, double2* out,
                 const double2* L, double2* partial, bool recur, bool ip, hipStream_t st)
{
  const bool nt = h.n >= kStreamingMinSites;
  LSQT_LAYOUT_SWITCH(h.layout, LAY, LSQT_K_SWITCH(h.K, KT, LSQT_NT_SWITCH(nt, NT, {
    constexpr int SPT = hpsi_spt_for(KT);
    const Grid g = make_grid(h.n, SPT);
    if (recur && ip)
      k_hpsi<LAY, KT, SPT, true, true, NT>
        <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
    else if (recur)
      k_hpsi<LAY, KT, SPT, true, false, NT>
        <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
    else if (ip)
      k_hpsi<LAY, KT, SPT, false, true, NT>
        <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
    else
      k_hpsi<LAY, KT, SPT, false, false, NT>
        <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
  })))
}

int hpsi_tiles(int n, int K)
{
  return make_grid(n, hpsi_spt_for(K)).tiles;
}

void launch_xop(const HamDev& h, const double2* in, double2* out, bool current, hipStream_t st)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT_LAYOUT_SWITCH(h.layout, LAY, LSQT_K_SWITCH(h.K, KT, LSQT_NT_SWITCH(nt, NT, {
    if (current)
      k_xop<LAY, KT, true, NT><<<g.blocks, kBlock, 0, st>>>(h, in, out, g.tiles, g.blocks);
    else
      k_xop<LAY, KT, false, NT><<<g.blocks, kBlock, 0, st>>>(h, in, out, g.tiles, g.blocks);
  })))
}

void launch_cheb2(const HamDev& h, const double2* s0, const double2* s1, double2* s2,
                  double2* s, double bm, int label, hipStream_t st, double bm_prev,
                  int label_prev)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT_LAYOUT_SWITCH(h.layout, LAY, LSQT_K_SWITCH(h.K, KT, LSQT_NT_SWITCH(nt, NT, {
    k_cheb2<LAY, KT, NT><<<g.blocks, kBlock, 0, st>>>(h, s0, s1, s2, s, bm, label, bm_prev,
                                                        label_prev, g.tiles, g.blocks);
  })))
}

void launch_cheb2x(const HamDev& h, const double2* s0, const double2* s0x, const double2* s1,
                   const double2* s1x, double2* s2, double2* s2x, double2* s, double bm,
                   int label, hipStream_t st, double bm_prev, int label_prev)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT_LAYOUT_SWITCH(h.layout, LAY, LSQT_K_SWITCH(h.K, KT, LSQT_NT_SWITCH(nt, NT, {
    k_cheb2x<LAY, KT, NT><<<g.blocks, kBlock, 0, st>>>(h, s0, s0x, s1, s1x, s2, s2x, s, bm,
                                                         label, bm_prev, label_prev, g.tiles,
                                                         g.blocks);
  })))
}

int pattern_mask_words(int n)
{
  return blocks_1d(n) * (kBlock / 32);
}

void launch_pattern_mask(const HamDev& h, uint32_t* mask, hipStream_t st)
{
  k_pattern_mask<<<blocks_1d(h.n), kBlock, 0, st>>>(h, mask);
}

void launch_zero(double2* v, int n, hipStream_t st)
{
  if (n > 0)
    k_zero<<<blocks_1d(n), kBlock, 0, st>>>(v, n);
}

void launch_copy(double2* dst, const double2* src, int n, hipStream_t st)
{
  if (n > 0)
    k_copy<<<blocks_1d(n), kBlock, 0, st>>>(dst, src, n);
}

void launch_add(double2* dst, const double2* src, int n, hipStream_t st)
{
  if (n > 0)
    k_add<<<blocks_1d(n), kBlock, 0, st>>>(dst, src, n);
}

void launch_sz(double2* dst, const double2* src, int n, hipStream_t st)
{
  if (n > 0)
    k_sz<<<blocks_1d(n), kBlock, 0, st>>>(dst, src, n);
}

void launch_cheb01(const double2* s0, const double2* s1, double2* s, double b0, double b1,
                   int direction, int n, hipStream_t st)
{
  if (n > 0)
    k_cheb01<<<blocks_1d(n), kBlock, 0, st>>>(s0, s1, s, b0, b1, direction, n);
}

void launch_cheb1x(const double2* s1x, double2* s, double b1, int n, hipStream_t st)
{
  if (n > 0)
    k_cheb1x<<<blocks_1d(n), kBlock, 0, st>>>(s1x, s, b1, n);
}

int sweep_check_blocks(int n)
{
  return blocks_1d(n);
}

void launch_sweep_check(const HamDev& h, const SweepMap& sm, int* out, hipStream_t st)
{
  k_sweep_check<<<blocks_1d(h.n), kBlock, 0, st>>>(h, sm, out);
}

void launch_sweepn(const HamDev& h, const SweepMap& sm, const double2* psi2,
                   const double2* psi1, double2* out0, double2* out1, const double2* L,
                   double2* part, size_t part_stride, hipStream_t st)
{
  // the shape sweep_shape_n makes: six steps, blocks of 1 slab
  const double2* left = L ? L : psi1;
  if (sm.ns == 6 && sm.nb == 1)
    sweepn_launch<double2, 6, 1>(h, sm, psi2, psi1, out0, out1, left, part, part_stride, st);
}

void launch_sweepn_evo(const HamDev& h, const SweepMap& sm, const double2* psi2,
                       const double2* psi1, double2* out0, double2* out1, double2* S,
                       const SweepTerms& terms, hipStream_t st)
{
  if (sm.ns == 6 && sm.nb == 1)
    sweepn_evo_launch<double2, 6, 1>(h, sm, psi2, psi1, out0, out1, S, terms, st);
}

size_t sweepn_lds(bool f32, int ns, int nb)
{
  const size_t v = f32 ? sizeof(float2) : sizeof(double2);
  return v * ns * (nb + 2) * kSweepThreads + sizeof(double) * (kSweepThreads / 64) * 2 * ns;
}

void launch_sweep2(const HamDev& h, const SweepMap& sm, const double2* psi2,
                   const double2* psi1, double2* out0, double2* out1, const double2* L,
                   double2* part0, double2* part1, hipStream_t st)
{
  sweep_launch<double2>(h, sm, psi2, psi1, out0, out1, L, part0, part1, st);
}

int dot_tiles(int n)
{
  return make_grid(n, 1).tiles;
}

void launch_dot(const double2* psi, const double2* L, double2* partial, int n, hipStream_t st,
                int spt)
{
  Grid g = make_grid(n, spt);
  k_dot<<<g.blocks, kBlock, 0, st>>>(psi, L, partial, n, spt, g.tiles, g.blocks);
}

void launch_reduce_partials(const double2* partial, size_t stride, int tiles, int count,
                            double2* mu, hipStream_t st)
{
  if (count > 0)
    k_reduce_partials<<<count, kBlock, 0, st>>>(partial, stride, tiles, mu);
}

}  // namespace lsqt
