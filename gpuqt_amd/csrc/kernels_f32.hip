// kernels_f32.hip -- the precision path: Chebyshev states stored in FP32.
//
// What is FP32 and what is not (measured on the 20,000-site honeycomb with
// 3000 moments: quantising the hopping / E_max values to FP32 costs 5e-6 of
// the DOS, storing the states in FP32 costs 9e-8):
//   - stored in FP32 (the bytes that move every step): the recursion states
//     psi_m, the left vector, the on-site potential;
//   - FP64: the hopping table, every product and sum inside a site update
//     (FP64 registers cost nothing on an HBM-bound step), the moment
//     inner products and their reduction, and the time-evolved state that
//     the Bessel series accumulates into.
// One fused step moves pot 4 + words 4z + psi_{m-1} 8 + psi_{m-2} 8 +
// psi_m 8 + left 8 bytes per site (48 at z = 3, against 84 in FP64).
//
// The block path keeps R vectors per site contiguous ([site][R] float2), so a
// neighbour gather reads R*8 contiguous bytes and the Hamiltonian is read
// once for all R vectors; VPW lanes share a site, each lane carries R/VPW
// adjacent vectors.
#include "kernels_sweep.h"
#include "kernels_sweepn.h"
#include "lsqt_device.h"
#include "lsqt_launch.h"

namespace lsqt {

template <bool NT>
__device__ __forceinline__ float2 ldf2(const float2* p)
{
  if constexpr (NT) {
    fvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const fvec2_t*>(p));
    return make_float2(v.x, v.y);
  } else {
    return *p;
  }
}

template <int LAYOUT>
__device__ __forceinline__ bool decode32(const HamDev& h, const double2* s_tab, uint32_t w,
                                         int& j, double& a, double& b)
{
  if (w == kEmptySlot)
    return false;
  j = (int)(w & h.idx_mask);
  int c = (int)(w >> h.idx_bits);
  double2 t = (LAYOUT == kPackedLds) ? s_tab[c] : h.tab[c];
  a = t.x;
  b = t.y;
  return true;
}

template <int LAYOUT, bool WITH_XX>
__device__ __forceinline__ void stage32(const HamDev& h, double2* s_tab, double* s_xx)
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

__device__ __forceinline__ void block_sum2(double re, double im, double2* out)
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

// H~ psi1 in FP64 registers from FP32 states
template <int LAYOUT, int KT, bool NT>
__device__ __forceinline__ void hpsi32_site(const HamDev& h, const double2* s_tab,
                                            const float2* __restrict__ psi1, int n, double& tr,
                                            double& ti)
{
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  float2 p1 = psi1[n];
  double v = (double)ld<NT>(h.pot32 + n);
  tr = v * (double)p1.x;
  ti = v * (double)p1.y;
  for (int m0 = 0; m0 < K; m0 += C) {
    int j[C];
    double a[C], b[C], x[C];
    bool ok[C];
    load_slots<LAYOUT, C, NT, false>(h, s_tab, nullptr, n, m0, K, j, a, b, x, ok);
    float2 c[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      c[q] = psi1[j[q]];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        double cx = c[q].x, cy = c[q].y;
        tr += a[q] * cx - b[q] * cy;
        ti += a[q] * cy + b[q] * cx;
      }
    }
  }
  tr /= h.emax;
  ti /= h.emax;
}

// out = H~ psi1 (RECUR false) or 2 H~ psi1 - psi0; optional moment partial.
// Each thread owns two consecutive sites and moves them with 16-byte vector
// loads (measured on MI355X: 80 us against 96 us per step for one site per
// thread on 8.0M sites).  The FP32 arrays are padded to the pair stride, so
// the second site of the last pair is always addressable; it is excluded
// from the moment when it lies beyond n.
template <int LAYOUT, int KT, bool RECUR, bool IP, bool NT>
__global__ __launch_bounds__(kBlock) void k_hpsi32(HamDev h, const float2* __restrict__ psi1,
                                                   const float2* psi0, float2* out,
                                                   const float2* __restrict__ L,
                                                   double2* __restrict__ partial, int tiles,
                                                   int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ int2 s_pat[kPatternMaxEntries];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage32<LAYOUT, false>(h, s_tab, nullptr);
  // a slot pattern: regular sites compute their words (the same words)
  const bool pat = h.pat != nullptr;  // the same for the whole launch
  if (pat) {
    for (int i = threadIdx.x; i < h.pat_p * h.K; i += kBlock)
      s_pat[i] = h.pat[i];
    __syncthreads();
  }
  const int K = KT > 0 ? KT : h.K;
  const int n = 2 * (tile * kBlock + threadIdx.x);
  double acc_re = 0.0, acc_im = 0.0;
  if (n < h.n) {
    int r0 = 0, r1 = 0;
    bool irr = false;
    if (pat) {
      r0 = pattern_residue(h, n);
      r1 = r0 + 1 == h.pat_p ? 0 : r0 + 1;
      // n is even: both bits sit in one mask word (padding past h.n is marked)
      irr = pattern_irregular(h, n) || pattern_irregular(h, n + 1);
    }
    fvec4_t p1 = *reinterpret_cast<const fvec4_t*>(psi1 + n);
    // slot words and potential with ordinary loads (non-temporal: 1 % slower
    // here), the once-read streams non-temporally and with the first loads
    // (their latency overlaps the gathers; out may be psi0: each site is
    // read before written)
    fvec2_t v = *reinterpret_cast<const fvec2_t*>(h.pot32 + n);
    fvec4_t p0, l;
    if constexpr (RECUR)
      p0 = ld<NT>(reinterpret_cast<const fvec4_t*>(psi0 + n));
    if constexpr (IP)
      l = ld<NT>(reinterpret_cast<const fvec4_t*>(L + n));
    double tr0 = (double)v.x * (double)p1.x, ti0 = (double)v.x * (double)p1.y;
    double tr1 = (double)v.y * (double)p1.z, ti1 = (double)v.y * (double)p1.w;
    constexpr int C = SlotBatch<KT>::C;
    for (int m0 = 0; m0 < K; m0 += C) {
      uvec2_t w[C];
      if (pat) {
#pragma unroll
        for (int q = 0; q < C; ++q)
          w[q] = (m0 + q < K)
                   ? uvec2_t{pattern_word(h, s_pat[(m0 + q) * h.pat_p + r0], n),
                             pattern_word(h, s_pat[(m0 + q) * h.pat_p + r1], n + 1)}
                   : uvec2_t{kEmptySlot, kEmptySlot};
        if (irr) {
#pragma unroll
          for (int q = 0; q < C; ++q)
            if (m0 + q < K)
              w[q] = ld<false>(
                reinterpret_cast<const uvec2_t*>(h.words + (size_t)(m0 + q) * h.stride + n));
        }
      } else {
#pragma unroll
        for (int q = 0; q < C; ++q)
          w[q] = (m0 + q < K)
                   ? ld<false>(reinterpret_cast<const uvec2_t*>(h.words + (size_t)(m0 + q) * h.stride + n))
                   : uvec2_t{kEmptySlot, kEmptySlot};
      }
      int j0[C], j1[C];
      double a0[C], b0[C], a1[C], b1[C];
      bool ok0[C], ok1[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        ok0[q] = decode32<LAYOUT>(h, s_tab, w[q].x, j0[q], a0[q], b0[q]);
        ok1[q] = decode32<LAYOUT>(h, s_tab, w[q].y, j1[q], a1[q], b1[q]);
        if (!ok0[q])
          j0[q] = n;
        if (!ok1[q])
          j1[q] = n;
      }
      float2 c0[C], c1[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c0[q] = psi1[j0[q]];
        c1[q] = psi1[j1[q]];
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
        if (ok0[q]) {
          double cx = c0[q].x, cy = c0[q].y;
          tr0 += a0[q] * cx - b0[q] * cy;
          ti0 += a0[q] * cy + b0[q] * cx;
        }
        if (ok1[q]) {
          double cx = c1[q].x, cy = c1[q].y;
          tr1 += a1[q] * cx - b1[q] * cy;
          ti1 += a1[q] * cy + b1[q] * cx;
        }
      }
    }
    tr0 /= h.emax;
    ti0 /= h.emax;
    tr1 /= h.emax;
    ti1 /= h.emax;
    if constexpr (RECUR) {
      tr0 = 2.0 * tr0 - (double)p0.x;
      ti0 = 2.0 * ti0 - (double)p0.y;
      tr1 = 2.0 * tr1 - (double)p0.z;
      ti1 = 2.0 * ti1 - (double)p0.w;
    }
    fvec4_t o;
    o.x = (float)tr0;
    o.y = (float)ti0;
    o.z = (float)tr1;
    o.w = (float)ti1;
    *reinterpret_cast<fvec4_t*>(out + n) = o;
    if constexpr (IP) {
      double ox = o.x, oy = o.y, lx = l.x, ly = l.y;
      acc_re = ox * lx + oy * ly;
      acc_im = oy * lx - ox * ly;
      if (n + 1 < h.n) {
        double ox1 = o.z, oy1 = o.w, lx1 = l.z, ly1 = l.w;
        acc_re += ox1 * lx1 + oy1 * ly1;
        acc_im += oy1 * lx1 - ox1 * ly1;
      }
    }
  }
  if constexpr (IP)
    block_sum2(acc_re, acc_im, partial + tile);
}

// the same for the full layout (separate index / value arrays), one site
// per thread
template <int KT, bool RECUR, bool IP, bool NT>
__global__ __launch_bounds__(kBlock) void k_hpsi32_full(HamDev h, const float2* __restrict__ psi1,
                                                        const float2* psi0, float2* out,
                                                        const float2* __restrict__ L,
                                                        double2* __restrict__ partial,
                                                        int tiles, int blocks)
{
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  double acc_re = 0.0, acc_im = 0.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    int n = (2 * tile + k) * kBlock + threadIdx.x;
    if (n < h.n) {
      double tr, ti;
      hpsi32_site<kFull, KT, NT>(h, nullptr, psi1, n, tr, ti);
      if constexpr (RECUR) {
        float2 p0 = ldf2<NT>(psi0 + n);
        tr = 2.0 * tr - (double)p0.x;
        ti = 2.0 * ti - (double)p0.y;
      }
      float2 o = make_float2((float)tr, (float)ti);
      out[n] = o;
      if constexpr (IP) {
        float2 l = ldf2<NT>(L + n);
        double lx = l.x, ly = l.y, ox = o.x, oy = o.y;
        acc_re += ox * lx + oy * ly;
        acc_im += oy * lx - ox * ly;
      }
    }
  }
  if constexpr (IP)
    block_sum2(acc_re, acc_im, partial + tile);
}

// [X, H~] psi in FP32 storage
template <int LAYOUT, int KT, bool NT>
__global__ __launch_bounds__(kBlock) void k_comm32(HamDev h, const float2* __restrict__ in,
                                                   float2* __restrict__ out, int tiles,
                                                   int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ double s_xx[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage32<LAYOUT, true>(h, s_tab, s_xx);
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
    float2 c[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      c[q] = in[j[q]];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        double cx = c[q].x, cy = c[q].y;
        tr -= (a[q] * cx - b[q] * cy) * x[q];
        ti -= (a[q] * cy + b[q] * cx) * x[q];
      }
    }
  }
  out[n] = make_float2((float)(tr / h.emax), (float)(ti / h.emax));
}

__device__ __forceinline__ void accumulate_label64(double2& s, double bm, double tr, double ti,
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

// one term of U(t): FP32 recursion, FP64 accumulation into state.
// bm_prev / label_prev: the previous launch's term, s1, added first where
// terms are added in pairs (label 0 adds nothing; with both 0 the state is
// neither read nor written): the same additions, in the same order
template <int LAYOUT, int KT, bool NT>
__global__ __launch_bounds__(kBlock) void k_cheb2_32(HamDev h, const float2* s0,
                                                     const float2* __restrict__ s1, float2* s2,
                                                     double2* __restrict__ s, double bm,
                                                     int label, double bm_prev, int label_prev,
                                                     int tiles, int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage32<LAYOUT, false>(h, s_tab, nullptr);
  int n = tile * kBlock + threadIdx.x;
  if (n >= h.n)
    return;
  double tr, ti;
  hpsi32_site<LAYOUT, KT, NT>(h, s_tab, s1, n, tr, ti);
  float2 p0 = ldf2<NT>(s0 + n);
  tr = 2.0 * tr - (double)p0.x;
  ti = 2.0 * ti - (double)p0.y;
  float2 o = make_float2((float)tr, (float)ti);
  s2[n] = o;
  if (label != 0 || label_prev != 0) {
    double2 st = ld2<NT>(s + n);
    if (label_prev != 0) {
      const float2 p1 = s1[n];  // a cache hit: the on-site term read it
      accumulate_label64(st, bm_prev, (double)p1.x, (double)p1.y, label_prev);
    }
    accumulate_label64(st, bm, (double)o.x, (double)o.y, label);
    s[n] = st;
  }
}

// one term of [X, U(t)]: coupled FP32 recursions, FP64 accumulation
template <int LAYOUT, int KT, bool NT>
__global__ __launch_bounds__(kBlock) void k_cheb2x_32(HamDev h, const float2* s0,
                                                      const float2* s0x,
                                                      const float2* __restrict__ s1,
                                                      const float2* __restrict__ s1x,
                                                      float2* s2, float2* s2x,
                                                      double2* __restrict__ s, double bm,
                                                      int label, double bm_prev,
                                                      int label_prev, int tiles, int blocks)
{
  __shared__ double2 s_tab[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  __shared__ double s_xx[LAYOUT == kPackedLds ? kLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  stage32<LAYOUT, true>(h, s_tab, s_xx);
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  int n = tile * kBlock + threadIdx.x;
  if (n >= h.n)
    return;
  double v = (double)ld<NT>(h.pot32 + n);
  float2 p1 = s1[n], p1x = s1x[n];
  double tr = v * (double)p1.x, ti = v * (double)p1.y;
  double txr = v * (double)p1x.x, txi = v * (double)p1x.y;
  for (int m0 = 0; m0 < K; m0 += C) {
    int j[C];
    double a[C], b[C], x[C];
    bool ok[C];
    load_slots<LAYOUT, C, NT, true>(h, s_tab, s_xx, n, m0, K, j, a, b, x, ok);
    float2 c[C], cxv[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      c[q] = s1[j[q]];
      cxv[q] = s1x[j[q]];
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (ok[q]) {
        double cx = c[q].x, cy = c[q].y, dx = cxv[q].x, dy = cxv[q].y;
        double hr = a[q] * cx - b[q] * cy, hi = a[q] * cy + b[q] * cx;
        tr += hr;
        ti += hi;
        txr += a[q] * dx - b[q] * dy;
        txi += a[q] * dy + b[q] * dx;
        txr -= hr * x[q];
        txi -= hi * x[q];
      }
    }
  }
  float2 p0 = ldf2<NT>(s0 + n), p0x = ldf2<NT>(s0x + n);
  tr = 2.0 * (tr / h.emax) - (double)p0.x;
  ti = 2.0 * (ti / h.emax) - (double)p0.y;
  txr = 2.0 * (txr / h.emax) - (double)p0x.x;
  txi = 2.0 * (txi / h.emax) - (double)p0x.y;
  s2[n] = make_float2((float)tr, (float)ti);
  float2 ox = make_float2((float)txr, (float)txi);
  s2x[n] = ox;
  if (label != 0 || label_prev != 0) {
    double2 st = ld2<NT>(s + n);
    accumulate_label64(st, bm_prev, (double)p1x.x, (double)p1x.y, label_prev);
    accumulate_label64(st, bm, (double)ox.x, (double)ox.y, label);
    s[n] = st;
  }
}

// element-wise: conversions and the first terms of the series
__global__ __launch_bounds__(kBlock) void k_to_f32(const double2* __restrict__ src,
                                                   float2* __restrict__ dst, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

__global__ __launch_bounds__(kBlock) void k_to_f64(const float2* __restrict__ src,
                                                   double2* __restrict__ dst, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    float2 v = src[i];
    dst[i] = make_double2((double)v.x, (double)v.y);
  }
}

__global__ __launch_bounds__(kBlock) void k_zero32(float2* v, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n)
    v[i] = make_float2(0.0f, 0.0f);
}

// state = b0 s0 + b1 dir (Im s1, -Re s1), s0 = state (FP64), s1 FP32
__global__ __launch_bounds__(kBlock) void k_cheb01_32(const float2* __restrict__ s1,
                                                      double2* __restrict__ s, double b0,
                                                      double b1, int direction, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double bessel_1 = b1 * direction;
    double2 a = s[i];
    float2 c = s1[i];
    s[i] = make_double2(b0 * a.x + bessel_1 * (double)c.y, b0 * a.y - bessel_1 * (double)c.x);
  }
}

__global__ __launch_bounds__(kBlock) void k_cheb1x_32(const float2* __restrict__ s1x,
                                                      double2* __restrict__ s, double b1, int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    float2 c = s1x[i];
    s[i] = make_double2(+b1 * (double)c.y, -b1 * (double)c.x);
  }
}

__global__ __launch_bounds__(kBlock) void k_dot32(const float2* __restrict__ psi,
                                                  const float2* __restrict__ L,
                                                  double2* __restrict__ partial, int n,
                                                  int tiles, int blocks)
{
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int k = 0; k < 2; ++k) {
    int i = 2 * (tile * kBlock + threadIdx.x) + k;
    if (i < n) {
      float2 p = psi[i], l = L[i];
      double px = p.x, py = p.y, lx = l.x, ly = l.y;
      re += px * lx + py * ly;
      im += py * lx - px * ly;
    }
  }
  block_sum2(re, im, partial + tile);
}

// per-tile partial of |s|^2 for the FP64 state (norm before / after evolve)
__global__ __launch_bounds__(kBlock) void k_norm2(const double2* __restrict__ s,
                                                  double2* __restrict__ partial, int n,
                                                  int tiles, int blocks)
{
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  double acc = 0.0;
  int i = tile * kBlock + threadIdx.x;
  if (i < n) {
    double2 v = s[i];
    acc = v.x * v.x + v.y * v.y;
  }
  block_sum2(acc, 0.0, partial + tile);
}

// s *= sqrt(target / current), both norms taken from the reduced partials
__global__ __launch_bounds__(kBlock) void k_renormalise(double2* __restrict__ s,
                                                        const double2* __restrict__ norms,
                                                        int n)
{
  int i = blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    double target = norms[0].x, current = norms[1].x;
    double f = current > 0.0 ? sqrt(target / current) : 1.0;
    double2 v = s[i];
    s[i] = make_double2(v.x * f, v.y * f);
  }
}

// ==========================================================================
// launchers
// ==========================================================================
#define LSQT32_LAYOUT_SWITCH(layout, L, ...) \
  switch (layout) {                          \
    case kFull: {                            \
      constexpr int L = kFull;               \
      __VA_ARGS__;                           \
    } break;                                 \
    case kPackedLds: {                       \
      constexpr int L = kPackedLds;          \
      __VA_ARGS__;                           \
    } break;                                 \
    default: {                               \
      constexpr int L = kPackedGlobal;       \
      __VA_ARGS__;                           \
    } break;                                 \
  }

#define LSQT32_K_SWITCH(kval, KT, ...) \
  switch (kval) {                      \
    case 3: {                          \
      constexpr int KT = 3;            \
      __VA_ARGS__;                     \
    } break;                           \
    case 4: {                          \
      constexpr int KT = 4;            \
      __VA_ARGS__;                     \
    } break;                           \
    case 6: {                          \
      constexpr int KT = 6;            \
      __VA_ARGS__;                     \
    } break;                           \
    default: {                         \
      constexpr int KT = 0;            \
      __VA_ARGS__;                     \
    } break;                           \
  }

#define LSQT32_NT_SWITCH(flag, NT, ...) \
  if (flag) {                           \
    constexpr bool NT = true;           \
    __VA_ARGS__;                        \
  } else {                              \
    constexpr bool NT = false;          \
    __VA_ARGS__;                        \
  }

static inline int blocks_1d32(int n)
{
  return (n + kBlock - 1) / kBlock;
}

void launch_sweepn_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                      const float2* psi1, float2* out0, float2* out1, const float2* L,
                      double2* part, size_t part_stride, hipStream_t st)
{
  // the shape sweep_shape_n makes for FP32 states: six steps, blocks of 1 slab
  const float2* left = L ? L : psi1;
  if (sm.ns == 6 && sm.nb == 1)
    sweepn_launch<float2, 6, 1>(h, sm, psi2, psi1, out0, out1, left, part, part_stride, st);
}

void launch_sweepn_evo_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                          const float2* psi1, float2* out0, float2* out1, double2* S,
                          const SweepTerms& terms, hipStream_t st)
{
  if (sm.ns == 6 && sm.nb == 1)
    sweepn_evo_launch<float2, 6, 1>(h, sm, psi2, psi1, out0, out1, S, terms, st);
}

void launch_sweep2_32(const HamDev& h, const SweepMap& sm, const float2* psi2,
                      const float2* psi1, float2* out0, float2* out1, const float2* L,
                      double2* part0, double2* part1, hipStream_t st)
{
  sweep_launch<float2>(h, sm, psi2, psi1, out0, out1, L, part0, part1, st);
}

int hpsi32_tiles(int n)
{
  return make_grid(n, 2).tiles;
}

void launch_hpsi32(const HamDev& h, const float2* psi1, const float2* psi0, float2* out,
                   const float2* L, double2* partial, bool recur, bool ip, hipStream_t st)
{
  Grid g = make_grid(h.n, 2);
  const bool nt = h.n >= kStreamingMinSites;
  if (h.layout == kFull) {
    LSQT32_K_SWITCH(h.K, KT, LSQT32_NT_SWITCH(nt, NT, {
      if (recur && ip)
        k_hpsi32_full<KT, true, true, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else if (recur)
        k_hpsi32_full<KT, true, false, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else if (ip)
        k_hpsi32_full<KT, false, true, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else
        k_hpsi32_full<KT, false, false, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
    }))
    return;
  }
  LSQT32_LAYOUT_SWITCH(h.layout, LAY, LSQT32_K_SWITCH(h.K, KT, LSQT32_NT_SWITCH(nt, NT, {
    if constexpr (LAY != kFull) {
      if (recur && ip)
        k_hpsi32<LAY, KT, true, true, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else if (recur)
        k_hpsi32<LAY, KT, true, false, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else if (ip)
        k_hpsi32<LAY, KT, false, true, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
      else
        k_hpsi32<LAY, KT, false, false, NT>
          <<<g.blocks, kBlock, 0, st>>>(h, psi1, psi0, out, L, partial, g.tiles, g.blocks);
    }
  })))
}

void launch_comm32(const HamDev& h, const float2* in, float2* out, hipStream_t st)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT32_LAYOUT_SWITCH(h.layout, LAY, LSQT32_K_SWITCH(h.K, KT, LSQT32_NT_SWITCH(nt, NT, {
    k_comm32<LAY, KT, NT><<<g.blocks, kBlock, 0, st>>>(h, in, out, g.tiles, g.blocks);
  })))
}

void launch_cheb2_32(const HamDev& h, const float2* s0, const float2* s1, float2* s2,
                     double2* s, double bm, int label, hipStream_t st, double bm_prev,
                     int label_prev)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT32_LAYOUT_SWITCH(h.layout, LAY, LSQT32_K_SWITCH(h.K, KT, LSQT32_NT_SWITCH(nt, NT, {
    k_cheb2_32<LAY, KT, NT><<<g.blocks, kBlock, 0, st>>>(h, s0, s1, s2, s, bm, label, bm_prev,
                                                          label_prev, g.tiles, g.blocks);
  })))
}

void launch_cheb2x_32(const HamDev& h, const float2* s0, const float2* s0x, const float2* s1,
                      const float2* s1x, float2* s2, float2* s2x, double2* s, double bm,
                      int label, hipStream_t st, double bm_prev, int label_prev)
{
  Grid g = make_grid(h.n, 1);
  const bool nt = h.n >= kStreamingMinSites;
  LSQT32_LAYOUT_SWITCH(h.layout, LAY, LSQT32_K_SWITCH(h.K, KT, LSQT32_NT_SWITCH(nt, NT, {
    k_cheb2x_32<LAY, KT, NT><<<g.blocks, kBlock, 0, st>>>(h, s0, s0x, s1, s1x, s2, s2x, s, bm,
                                                           label, bm_prev, label_prev, g.tiles,
                                                           g.blocks);
  })))
}

void launch_to_f32(const double2* src, float2* dst, int n, hipStream_t st)
{
  if (n > 0)
    k_to_f32<<<blocks_1d32(n), kBlock, 0, st>>>(src, dst, n);
}

void launch_to_f64(const float2* src, double2* dst, int n, hipStream_t st)
{
  if (n > 0)
    k_to_f64<<<blocks_1d32(n), kBlock, 0, st>>>(src, dst, n);
}

void launch_zero32(float2* v, int n, hipStream_t st)
{
  if (n > 0)
    k_zero32<<<blocks_1d32(n), kBlock, 0, st>>>(v, n);
}

void launch_cheb01_32(const float2* s1, double2* s, double b0, double b1, int direction, int n,
                      hipStream_t st)
{
  if (n > 0)
    k_cheb01_32<<<blocks_1d32(n), kBlock, 0, st>>>(s1, s, b0, b1, direction, n);
}

void launch_cheb1x_32(const float2* s1x, double2* s, double b1, int n, hipStream_t st)
{
  if (n > 0)
    k_cheb1x_32<<<blocks_1d32(n), kBlock, 0, st>>>(s1x, s, b1, n);
}

void launch_dot32(const float2* psi, const float2* L, double2* partial, int n, hipStream_t st)
{
  Grid g = make_grid(n, 2);
  k_dot32<<<g.blocks, kBlock, 0, st>>>(psi, L, partial, n, g.tiles, g.blocks);
}

void launch_norm2(const double2* s, double2* partial, int n, hipStream_t st)
{
  Grid g = make_grid(n, 1);
  k_norm2<<<g.blocks, kBlock, 0, st>>>(s, partial, n, g.tiles, g.blocks);
}

void launch_renormalise(double2* s, const double2* norms, int n, hipStream_t st)
{
  if (n > 0)
    k_renormalise<<<blocks_1d32(n), kBlock, 0, st>>>(s, norms, n);
}

}  // namespace lsqt
