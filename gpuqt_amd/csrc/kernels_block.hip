// kernels_block.hip -- the block path: R random vectors per sparse sweep.
//
// A block holds R complex vectors stored site-major with the R values of a
// site contiguous ([site][R] float2), so that
//   - the Hamiltonian (potential, pair words) is read once per site for all R
//     vectors: (4 + 4z) / R bytes per site-update instead of 4 + 4z;
//   - a neighbour gather reads R*8 contiguous bytes (256 B at R = 32).
// VPW lanes share a site ("vectors per wavefront" = 64/VPW sites per wave),
// each lane carrying W = R/VPW adjacent vectors (8W-byte vector loads); a
// workgroup sweeps a tile of ROWS consecutive sites ("rows per block").  Both
// are picked per (R, z, layout, size) by the autotuner (capi_block.hip).
//
// Precision: FP32 states, left vectors, on-site energies and hopping values,
// FP32 arithmetic for the sparse product and the three-term recursion; the
// moment products <L_r|psi_r> are formed and summed in FP64 (a product of two
// FP32 numbers is exact in FP64), reduced per vector by a fixed shuffle / LDS
// tree into per-tile partials [M][tiles][R] and then in tile order.
//
// Kernels: k_block_step (one recursion step with its moment products: <L|.>,
// the self products of the product identities, or pairs of a left and a
// right recursion), k_block_win (the same step on a lattice with rows
// y-1..y+1 of the recursion block staged in LDS), k_block_evo (the terms of
// U(t), [X, U(t)] and the operators, FP64 accumulated state), reductions,
// norms and layout changes.
#include <algorithm>

#include "lsqt_device.h"
#include "lsqt_launch.h"

namespace lsqt {

constexpr int kBlockLdsTableMax = 1024;

template <int W>
struct Vec {
  float2 v[W];
};

template <int W, bool NT>
__device__ __forceinline__ Vec<W> ldv(const float2* p)
{
  Vec<W> r;
  if constexpr (W == 1) {
    if constexpr (NT) {
      fvec2_t t = __builtin_nontemporal_load(reinterpret_cast<const fvec2_t*>(p));
      r.v[0] = make_float2(t.x, t.y);
    } else {
      r.v[0] = *p;
    }
  } else {
#pragma unroll
    for (int i = 0; i < W; i += 2) {
      fvec4_t t;
      if constexpr (NT)
        t = __builtin_nontemporal_load(reinterpret_cast<const fvec4_t*>(p + i));
      else
        t = *reinterpret_cast<const fvec4_t*>(p + i);
      r.v[i] = make_float2(t.x, t.y);
      r.v[i + 1] = make_float2(t.z, t.w);
    }
  }
  return r;
}

template <int W, bool NT = false>
__device__ __forceinline__ void stv(float2* p, const Vec<W>& r)
{
  if constexpr (W == 1) {
    if constexpr (NT) {
      fvec2_t t;
      t.x = r.v[0].x;
      t.y = r.v[0].y;
      __builtin_nontemporal_store(t, reinterpret_cast<fvec2_t*>(p));
    } else {
      *p = r.v[0];
    }
  } else {
#pragma unroll
    for (int i = 0; i < W; i += 2) {
      fvec4_t t;
      t.x = r.v[i].x;
      t.y = r.v[i].y;
      t.z = r.v[i + 1].x;
      t.w = r.v[i + 1].y;
      if constexpr (NT)
        __builtin_nontemporal_store(t, reinterpret_cast<fvec4_t*>(p + i));
      else
        *reinterpret_cast<fvec4_t*>(p + i) = t;
    }
  }
}

__device__ __forceinline__ void tile_range(const TileMap& t, int tile, int& begin, int& end)
{
  if (t.table) {
    const int2 e = t.table[tile];
    begin = e.x;
    end = e.y;
  } else if (t.slab == 0) {
    begin = tile * t.rows;
    end = min(begin + t.rows, t.n);
  } else {
    // blocked schedule: tile = (row block, slab) with slabs fastest, so the
    // tiles that run together are the same rows of neighbouring slabs
    const int rb = tile / t.slabs, sl = tile % t.slabs;
    begin = sl * t.slab + rb * t.span;
    end = begin + t.span;
  }
}

// Packed-layout slot words of one site (all K = C slots), issued together.
template <int C, bool NT>
__device__ __forceinline__ void load_words(const HamDev& h, int site, uint32_t (&w)[C])
{
#pragma unroll
  for (int q = 0; q < C; ++q)
    w[q] = ld<NT>(h.words + (size_t)q * h.stride + site);
}

// Decodes slot words into neighbour index and FP32 hopping value.  Unused
// slots get the site itself as index (an in-bounds address) and a zero
// hopping, so that every slot is gathered and summed without a branch (the
// gathered values are finite, and adding +0 leaves a sum unchanged).
template <int LAYOUT, int C>
__device__ __forceinline__ void decode_words32(const HamDev& h, const float2* s_tab, int site,
                                               const uint32_t (&w)[C], int (&j)[C],
                                               float (&a)[C], float (&b)[C])
{
#pragma unroll
  for (int q = 0; q < C; ++q) {
    const bool ok = w[q] != kEmptySlot;
    j[q] = ok ? (int)(w[q] & h.idx_mask) : site;
    const int c = ok ? (int)(w[q] >> h.idx_bits) : 0;
    float2 t;
    if constexpr (LAYOUT == kPackedLds) {
      t = s_tab[c];
    } else {
      const double2 d = h.tab[c];
      t = make_float2((float)d.x, (float)d.y);
    }
    a[q] = ok ? t.x : 0.0f;
    b[q] = ok ? t.y : 0.0f;
  }
}

// Slot batch [m0, m0 + C) of a site whose slot count is known only at run
// time: all words of the batch are issued before any is decoded.
template <int LAYOUT, int C, bool NT>
__device__ __forceinline__ void load_slots32(const HamDev& h, const float2* s_tab, int site,
                                             int m0, int count, int (&j)[C], float (&a)[C],
                                             float (&b)[C])
{
  uint32_t w[C];
#pragma unroll
  for (int q = 0; q < C; ++q)
    w[q] = (m0 + q < count) ? ld<NT>(h.words + (size_t)(m0 + q) * h.stride + site) : kEmptySlot;
  decode_words32<LAYOUT, C>(h, s_tab, site, w, j, a, b);
}

// Per-vector tile sums of the moment products (FP64), by a fixed tree: lanes
// holding the same vectors differ in the site bits of the lane id (shuffles),
// then the workgroup's waves in order (LDS).  Ends with a barrier, so that
// s_red can be reused.
template <int W>
__device__ __forceinline__ void tile_sums(double (&acc_re)[W], double (&acc_im)[W], int vpw, int R,
                                          int r0, double (*s_red)[kBlock / 64][64 * 4],
                                          double2* __restrict__ dst)
{
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int w = 0; w < W; ++w) {
    for (int off = vpw; off < 64; off <<= 1) {
      acc_re[w] += __shfl_xor(acc_re[w], off, 64);
      acc_im[w] += __shfl_xor(acc_im[w], off, 64);
    }
  }
  if (lane < vpw) {
#pragma unroll
    for (int w = 0; w < W; ++w) {
      s_red[0][wave][r0 + w] = acc_re[w];
      s_red[1][wave][r0 + w] = acc_im[w];
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < R; r += kBlock) {
    double re = s_red[0][0][r], im = s_red[1][0][r];
#pragma unroll
    for (int wv = 1; wv < kBlock / 64; ++wv) {
      re += s_red[0][wv][r];
      im += s_red[1][wv][r];
    }
    dst[r] = make_double2(re, im);
  }
  __syncthreads();
}

// Moment products fused into a block step.
enum BlockProducts {
  kLeft = 1,  // <L_r|out_r> into partial
  kSelf = 2,  // <psi1_r|out_r> into partial, <out_r|out_r> into partial2
  // a block of R/2 pairs of recursions, columns 2p (from L_p) and 2p + 1
  // (from R_p): <psi1_2p|out_2p+1> into partial, <out_2p|out_2p+1> into
  // partial2, per pair p (W even)
  kPair = 3,
};

// Per-lane FP64 accumulators of a tile's moment products.
template <int W>
struct Products {
  double re[W], im[W], re2[W], im2[W];
};

template <int W>
__device__ __forceinline__ void clear_products(Products<W>& a)
{
#pragma unroll
  for (int w = 0; w < W; ++w)
    a.re[w] = a.im[w] = a.re2[w] = a.im2[w] = 0.0;
}

// adds one site's products: o = the new values, p1 = psi1 at the site, l =
// the left vector (kLeft only); conj(x) y in FP64 (exact products of FP32
// values)
template <int IPM, int W>
__device__ __forceinline__ void add_products(Products<W>& a, const Vec<W>& o, const Vec<W>& p1,
                                             const Vec<W>& l)
{
  if constexpr (IPM == kPair) {
#pragma unroll
    for (int w = 0; w + 1 < W; w += 2) {
      const double lx = p1.v[w].x, ly = p1.v[w].y;  // psi_{m-1} of the L recursion
      const double ax = o.v[w].x, ay = o.v[w].y;    // psi_m of the L recursion
      const double rx = o.v[w + 1].x, ry = o.v[w + 1].y;  // psi_m of the R recursion
      a.re[w / 2] += lx * rx + ly * ry;
      a.im[w / 2] += lx * ry - ly * rx;
      a.re2[w / 2] += ax * rx + ay * ry;
      a.im2[w / 2] += ax * ry - ay * rx;
    }
  } else {
    const Vec<W>& m = IPM == kSelf ? p1 : l;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const double ox = o.v[w].x, oy = o.v[w].y, lx = m.v[w].x, ly = m.v[w].y;
      a.re[w] += ox * lx + oy * ly;
      a.im[w] += oy * lx - ox * ly;
      if constexpr (IPM == kSelf)
        a.re2[w] += ox * ox + oy * oy;
    }
  }
}

// the tile's per-vector (or per-pair) sums into partial / partial2 [tile][.]
template <int IPM, int W>
__device__ __forceinline__ void store_products(Products<W>& a, int vpw, int R, int r0, int tile,
                                               double (*s_red)[kBlock / 64][64 * 4],
                                               double2* partial, double2* partial2)
{
  if constexpr (IPM == kPair) {
    constexpr int WP = W >= 2 ? W / 2 : 1;
    double r1[WP], i1[WP], r2[WP], i2[WP];
#pragma unroll
    for (int w = 0; w < WP; ++w) {
      r1[w] = a.re[w];
      i1[w] = a.im[w];
      r2[w] = a.re2[w];
      i2[w] = a.im2[w];
    }
    const int P = R / 2;
    tile_sums<WP>(r1, i1, vpw, P, r0 / 2, s_red, partial + (size_t)tile * P);
    tile_sums<WP>(r2, i2, vpw, P, r0 / 2, s_red, partial2 + (size_t)tile * P);
  } else {
    tile_sums<W>(a.re, a.im, vpw, R, r0, s_red, partial + (size_t)tile * R);
    if constexpr (IPM == kSelf)
      tile_sums<W>(a.re2, a.im2, vpw, R, r0, s_red, partial2 + (size_t)tile * R);
  }
}

// One Chebyshev step for R vectors (RECUR: out = 2 H~ psi1 - psi0, else
// out = H~ psi1) in FP32 arithmetic, fused with the per-vector moment
// products of mode IPM (FP64 products and sums, per-tile partials).  The
// streams read or written once per step (psi0, L, out) bypass the caches
// (nontemporal); psi1, whose rows are gathered again by neighbouring sites,
// and the slot words (NT) use ordinary loads.  On the 160^3 cube at R = 32
// the nontemporal streams raised the step rate by about 4 %.
template <int LAYOUT, int KT, int W, bool RECUR, int IPM, bool NT, bool DB>
__global__ __launch_bounds__(kBlock) void k_block_step(
  HamDev h, int R, TileMap tm, const float2* __restrict__ psi1, const float2* psi0, float2* out,
  const float2* __restrict__ L, double2* __restrict__ partial, double2* __restrict__ partial2,
  int tiles, int blocks)
{
  __shared__ float2 s_tab[LAYOUT == kPackedLds ? kBlockLdsTableMax : 1];
  __shared__ double s_red[2][kBlock / 64][64 * 4];  // [re/im][wave][vector] (R <= 256)
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  if constexpr (LAYOUT == kPackedLds) {
    for (int i = threadIdx.x; i < h.ncls; i += kBlock) {
      const double2 d = h.tab[i];
      s_tab[i] = make_float2((float)d.x, (float)d.y);
    }
    __syncthreads();
  }
  const int K = KT > 0 ? KT : h.K;
  const int vpw = R / W;                       // lanes per site
  const int per_pass = kBlock / vpw;           // sites per workgroup pass
  const int lane_site = threadIdx.x / vpw;
  const int r0 = (threadIdx.x % vpw) * W;      // first vector of this lane
  int begin, end;
  tile_range(tm, tile, begin, end);
  const float inv_scale = (float)(1.0 / h.emax);
  Products<W> acc;
  clear_products(acc);

  // With a compile-time slot count the loop is software-pipelined: the words
  // of the next site are loaded while this site's gathers are in flight, so a
  // pass costs one memory round trip instead of two.  The prefetch is issued
  // unconditionally (the last pass re-reads its own words) so that the
  // compiler can count the loads in flight and wait only for the ones used.
  //
  // Pass order: consecutive sites (s += per_pass), or with the column walk
  // (blocked tiles of whole lattice rows) the tile's rows are swept side by
  // side, one chunk of per_pass sites of every row before the next chunk, so
  // that neighbours one row apart are read one pass apart.
  constexpr int C = SlotBatch<KT>::C;
  constexpr bool kPipe = KT > 0;
  const int by = tm.walk ? tm.span / tm.row : 1;  // rows per tile (column walk)
  int s = begin + lane_site;
  int wy = 0;  // row of the column walk
  const int np = tm.walk ? by * (tm.row / per_pass) : (end - s + per_pass - 1) / per_pass;
  auto advance = [&](int& site, int& y) {
    if (tm.walk) {
      if (++y == by) {
        y = 0;
        site += per_pass - (by - 1) * tm.row;
      } else {
        site += tm.row;
      }
    } else {
      site += per_pass;
    }
  };
  if constexpr (DB && kPipe) {
    // Double buffering: the loads of pass p+1 (own rows, psi0, left rows,
    // gathers) are issued before pass p is computed, the slot words two
    // passes ahead.  Past the last pass the first site of the tile stands in
    // (its loads are in bounds and never used).
    struct Pass {
      Vec<W> p1, p0, l;
      Vec<W> c[C];
      float v;
      float a[C], b[C];
    };
    const int first = s;
    auto issue = [&](int site, const uint32_t (&wd)[C], Pass& P) {
      int j[C];
      decode_words32<LAYOUT, C>(h, s_tab, site, wd, j, P.a, P.b);
      const size_t base = (size_t)site * R + r0;
      P.p1 = ldv<W, false>(psi1 + base);
      P.v = ld<NT>(h.pot32 + site);
      if constexpr (RECUR)
        P.p0 = ldv<W, true>(psi0 + base);
      if constexpr (IPM == kLeft)
        P.l = ldv<W, true>(L + base);
#pragma unroll
      for (int q = 0; q < C; ++q)
        P.c[q] = ldv<W, false>(psi1 + (size_t)j[q] * R + r0);
    };
    if (np > 0) {
      int s1 = s, y1 = wy;
      advance(s1, y1);
      int s2 = s1, y2 = y1;
      advance(s2, y2);
      uint32_t wd[C];
      load_words<C, NT>(h, s, wd);
      Pass cur;
      issue(s, wd, cur);
      load_words<C, NT>(h, np > 1 ? s1 : first, wd);
      for (int p = 0; p < np; ++p) {
        Pass nxt;
        issue(p + 1 < np ? s1 : first, wd, nxt);
        load_words<C, NT>(h, p + 2 < np ? s2 : first, wd);
        const size_t base = (size_t)s * R + r0;
        float tr[W], ti[W];
#pragma unroll
        for (int w = 0; w < W; ++w) {
          tr[w] = cur.v * cur.p1.v[w].x;
          ti[w] = cur.v * cur.p1.v[w].y;
        }
#pragma unroll
        for (int q = 0; q < C; ++q) {
#pragma unroll
          for (int w = 0; w < W; ++w) {
            const float cx = cur.c[q].v[w].x, cy = cur.c[q].v[w].y;
            tr[w] = __builtin_fmaf(cur.a[q], cx, tr[w]);
            tr[w] = __builtin_fmaf(-cur.b[q], cy, tr[w]);
            ti[w] = __builtin_fmaf(cur.a[q], cy, ti[w]);
            ti[w] = __builtin_fmaf(cur.b[q], cx, ti[w]);
          }
        }
        Vec<W> o;
#pragma unroll
        for (int w = 0; w < W; ++w) {
          if constexpr (RECUR) {
            o.v[w].x = __builtin_fmaf(2.0f, tr[w] * inv_scale, -cur.p0.v[w].x);
            o.v[w].y = __builtin_fmaf(2.0f, ti[w] * inv_scale, -cur.p0.v[w].y);
          } else {
            o.v[w].x = tr[w] * inv_scale;
            o.v[w].y = ti[w] * inv_scale;
          }
        }
        stv<W, true>(out + base, o);
        add_products<IPM, W>(acc, o, cur.p1, cur.l);
        cur = nxt;
        s = s1;
        wy = y1;
        s1 = s2;
        y1 = y2;
        advance(s2, y2);
      }
    }
  } else {
  uint32_t wnext[C];
  if constexpr (kPipe) {
    if (np > 0)
      load_words<C, NT>(h, s, wnext);
  }
  for (int p = 0; p < np; ++p) {
    int sn = s, yn = wy;
    advance(sn, yn);
    const size_t base = (size_t)s * R + r0;
    Vec<W> p1 = ldv<W, false>(psi1 + base);
    const float v = ld<NT>(h.pot32 + s);
    Vec<W> p0, l;
    if constexpr (RECUR)
      p0 = ldv<W, true>(psi0 + base);
    if constexpr (IPM == kLeft)
      l = ldv<W, true>(L + base);
    float tr[W], ti[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      tr[w] = v * p1.v[w].x;
      ti[w] = v * p1.v[w].y;
    }
    for (int m0 = 0; m0 < K; m0 += C) {
      int j[C];
      float a[C], b[C];
      if constexpr (kPipe) {
        uint32_t wcur[C];
#pragma unroll
        for (int q = 0; q < C; ++q)
          wcur[q] = wnext[q];
        load_words<C, NT>(h, p + 1 < np ? sn : s, wnext);
        decode_words32<LAYOUT, C>(h, s_tab, s, wcur, j, a, b);
      } else {
        load_slots32<LAYOUT, C, NT>(h, s_tab, s, m0, K, j, a, b);
      }
      Vec<W> c[C];
#pragma unroll
      for (int q = 0; q < C; ++q)
        c[q] = ldv<W, false>(psi1 + (size_t)j[q] * R + r0);
#pragma unroll
      for (int q = 0; q < C; ++q) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          // explicit FMAs: the same rounding in every kernel shape
          const float cx = c[q].v[w].x, cy = c[q].v[w].y;
          tr[w] = __builtin_fmaf(a[q], cx, tr[w]);
          tr[w] = __builtin_fmaf(-b[q], cy, tr[w]);
          ti[w] = __builtin_fmaf(a[q], cy, ti[w]);
          ti[w] = __builtin_fmaf(b[q], cx, ti[w]);
        }
      }
    }
    Vec<W> o;
    if constexpr (RECUR) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        o.v[w].x = __builtin_fmaf(2.0f, tr[w] * inv_scale, -p0.v[w].x);
        o.v[w].y = __builtin_fmaf(2.0f, ti[w] * inv_scale, -p0.v[w].y);
      }
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        o.v[w].x = tr[w] * inv_scale;
        o.v[w].y = ti[w] * inv_scale;
      }
    }
    stv<W, true>(out + base, o);
    // kSelf / kPair use psi1, loaded for the on-site term
    add_products<IPM, W>(acc, o, p1, l);
    s = sn;
    wy = yn;
  }
  }

  store_products<IPM, W>(acc, vpw, R, r0, tile, s_red, partial, partial2);
}

// One tile of the sliding window of WindowMap: the tile is rows y0 .. y0 +
// by - 1 of one chunk (z0 .. z0 + P - 1) of one slab; rows y - 1, y, y + 1
// of a source block (the chunk plus one site either side, periodic in y and
// z) sit in a ring of four LDS slots of (P + 2) * R float2.  Tile order:
// chunk fastest, then slab, then row block, so the tiles in flight on one
// XCD cover whole lattice rows (contiguous memory) of a run of neighbouring
// slabs (their x-neighbour gathers hit that XCD's L2).
template <int W>
struct WindowTile {
  int slab, row, ny, by, R, P, span, z0, y0, xbase;
  Divisor row_div;
  // a window row is (P + 2) * vpw items of W vectors: item i = (position
  // i / vpw, lane group i % vpw); thread t loads item t and, for t < 2 vpw,
  // item kBlock + t
  int d_a, g_a, d_b, g_b;
  bool has_b;

  __device__ WindowTile(const WindowMap& m, int R_, int tile)
  {
    slab = m.slab;
    row = m.row;
    ny = m.ny;
    by = m.by;
    row_div = m.row_div;
    R = R_;
    P = m.chunk;
    span = (P + 2) * R;
    z0 = (tile % m.chunks) * P;
    const int u = tile / m.chunks;
    xbase = (u % m.slabs) * slab;
    y0 = (u / m.slabs) * by;
    const int vpw = R / W;
    d_a = threadIdx.x / vpw;
    g_a = threadIdx.x % vpw;
    has_b = threadIdx.x < 2 * vpw;
    d_b = (kBlock + threadIdx.x) / vpw;
    g_b = (kBlock + threadIdx.x) % vpw;
  }
  // site at position d (0 .. P + 1) of window row y (any integer row)
  __device__ int site(int y, int d) const
  {
    int yy = y;
    if (yy < 0)
      yy += ny;
    else if (yy >= ny)
      yy -= ny;
    int z = z0 - 1 + d;
    if (z < 0)
      z += row;
    else if (z >= row)
      z -= row;
    return xbase + yy * row + z;
  }
  __device__ void fetch(const float2* src, int y, Vec<W>& va, Vec<W>& vb) const
  {
    va = ldv<W, false>(src + (size_t)site(y, d_a) * R + g_a * W);
    if (has_b)
      vb = ldv<W, false>(src + (size_t)site(y, d_b) * R + g_b * W);
  }
  __device__ void put(float2* slots, int slot, const Vec<W>& va, const Vec<W>& vb) const
  {
    stv<W>(slots + slot * span + d_a * R + g_a * W, va);
    if (has_b)
      stv<W>(slots + slot * span + d_b * R + g_b * W, vb);
  }
  // rows y0 - 1 .. y0 + 1 into slots 0 .. 2 (callers then synchronise)
  __device__ void prologue(const float2* src, float2* slots) const
  {
    Vec<W> a0, b0, a1, b1, a2, b2;
    fetch(src, y0 - 1, a0, b0);
    fetch(src, y0, a1, b1);
    fetch(src, y0 + 1, a2, b2);
    put(slots, 0, a0, b0);
    put(slots, 1, a1, b1);
    put(slots, 2, a2, b2);
  }
  // slot offset (float2) of neighbour j's values for vectors r0.. in the
  // window while row y = y0 + p is swept, or -1 where j lies outside it
  // (other slab, |dy| > 1, or z outside the chunk +- 1)
  __device__ int pos(int j, int p, int r0) const
  {
    const int rr = j - xbase;
    if ((unsigned)rr >= (unsigned)slab)
      return -1;
    const int yj = div_by(rr, row_div);
    int dy = yj - (y0 + p);
    if (dy == ny - 1)
      dy = -1;
    else if (dy == 1 - ny)
      dy = 1;
    int dz = rr - yj * row - z0 + 1;  // -row < dz <= row
    if (dz < 0)
      dz += row;
    else if (dz >= row)
      dz -= row;
    if (dy < -1 || dy > 1 || dz >= P + 2)
      return -1;
    return ((p + 1 + dy) & 3) * span + dz * R + r0;
  }
  static __device__ Vec<W> read(const float2* slots, int pos)
  {
    Vec<W> v;
#pragma unroll
    for (int w = 0; w < W; ++w)
      v.v[w] = slots[pos + w];
    return v;
  }
};

// The block step on a lattice with the sliding window (WindowTile): the
// workgroup sweeps the rows of its tile; row y + 2 of psi1 is loaded while
// row y is computed.  Neighbours inside the window are read from LDS, all
// others (the neighbouring slabs and anything the lattice does not predict)
// from global memory, so any Hamiltonian gives the same values as
// k_block_step; only how often a row of psi1 crosses the L1 / L2 path
// changes (once per row instead of once per neighbour).  Same arithmetic and
// products as k_block_step; per-tile sums over the tile's by x chunk sites.
template <int LAYOUT, int KT, int W, bool RECUR, int IPM>
__global__ __launch_bounds__(kBlock) void k_block_win(
  HamDev h, int R, TileMap tm, const float2* __restrict__ psi1, const float2* psi0, float2* out,
  const float2* __restrict__ L, double2* __restrict__ partial, double2* __restrict__ partial2,
  int tiles, int blocks)
{
  static_assert(KT > 0, "the window kernel needs a compile-time slot count");
  __shared__ float2 s_tab[LAYOUT == kPackedLds ? kBlockLdsTableMax : 1];
  // the window ring during the sweep, the tile sums after it
  extern __shared__ __align__(16) unsigned char s_dyn[];
  float2* s_win = reinterpret_cast<float2*>(s_dyn);
  auto s_red = reinterpret_cast<double (*)[kBlock / 64][64 * 4]>(s_dyn);
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  if constexpr (LAYOUT == kPackedLds) {
    for (int i = threadIdx.x; i < h.ncls; i += kBlock) {
      const double2 d = h.tab[i];
      s_tab[i] = make_float2((float)d.x, (float)d.y);
    }
  }
  const WindowTile<W> wt(tm.win, R, tile);
  const int vpw = R / W;  // lanes per site
  const int lane_site = threadIdx.x / vpw;
  const int r0 = (threadIdx.x % vpw) * W;
  wt.prologue(psi1, s_win);
  __syncthreads();

  const float inv_scale = (float)(1.0 / h.emax);
  Products<W> acc;
  clear_products(acc);
  constexpr int C = KT;
  int s = wt.xbase + wt.y0 * wt.row + wt.z0 + lane_site;
  uint32_t wnext[C];
  load_words<C, false>(h, s, wnext);
  for (int p = 0; p < wt.by; ++p) {
    const bool more = p + 1 < wt.by;
    Vec<W> na, nb;
    if (more)
      wt.fetch(psi1, wt.y0 + p + 2, na, nb);
    const int sn = more ? s + wt.row : s;
    const size_t base = (size_t)s * R + r0;
    const float v = h.pot32[s];
    Vec<W> p0, l;
    if constexpr (RECUR)
      p0 = ldv<W, true>(psi0 + base);
    if constexpr (IPM == kLeft)
      l = ldv<W, true>(L + base);
    int j[C];
    float a[C], b[C];
    {
      uint32_t wcur[C];
#pragma unroll
      for (int q = 0; q < C; ++q)
        wcur[q] = wnext[q];
      load_words<C, false>(h, sn, wnext);
      decode_words32<LAYOUT, C>(h, s_tab, s, wcur, j, a, b);
    }
    int pos[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      pos[q] = wt.pos(j[q], p, r0);
    Vec<W> c[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (pos[q] < 0)
        c[q] = ldv<W, false>(psi1 + (size_t)j[q] * R + r0);
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (pos[q] >= 0)
        c[q] = WindowTile<W>::read(s_win, pos[q]);
    const Vec<W> p1 =
      WindowTile<W>::read(s_win, ((p + 1) & 3) * wt.span + (lane_site + 1) * R + r0);
    float tr[W], ti[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      tr[w] = v * p1.v[w].x;
      ti[w] = v * p1.v[w].y;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const float cx = c[q].v[w].x, cy = c[q].v[w].y;
        tr[w] = __builtin_fmaf(a[q], cx, tr[w]);
        tr[w] = __builtin_fmaf(-b[q], cy, tr[w]);
        ti[w] = __builtin_fmaf(a[q], cy, ti[w]);
        ti[w] = __builtin_fmaf(b[q], cx, ti[w]);
      }
    }
    Vec<W> o;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      if constexpr (RECUR) {
        o.v[w].x = __builtin_fmaf(2.0f, tr[w] * inv_scale, -p0.v[w].x);
        o.v[w].y = __builtin_fmaf(2.0f, ti[w] * inv_scale, -p0.v[w].y);
      } else {
        o.v[w].x = tr[w] * inv_scale;
        o.v[w].y = ti[w] * inv_scale;
      }
    }
    stv<W, true>(out + base, o);
    add_products<IPM, W>(acc, o, p1, l);
    if (more)
      wt.put(s_win, (p + 3) & 3, na, nb);
    __syncthreads();
    s = sn;
  }

  store_products<IPM, W>(acc, vpw, R, r0, tile, s_red, partial, partial2);
}

// mu[(k*R + r)*M + m0 + i] = sum over the tiles t of segment k of
// partial[i][t][r] (slots `stride` tiles apart), in tile order, i <
// gridDim.x, k = blockIdx.z (one segment of all tiles without a segment
// table)
__global__ __launch_bounds__(kBlock) void k_block_reduce(const double2* __restrict__ partial,
                                                         int tiles, int stride, int R, int M,
                                                         int m0,
                                                         const int* __restrict__ seg_tiles,
                                                         double2* __restrict__ mu)
{
  __shared__ double s_re[kBlock], s_im[kBlock];
  const int m = m0 + blockIdx.x;
  const int k = blockIdx.z;
  const int t0 = seg_tiles ? seg_tiles[k] : 0;
  const int t1 = seg_tiles ? seg_tiles[k + 1] : tiles;
  const double2* p = partial + (size_t)blockIdx.x * stride * R;
  mu += (size_t)k * R * M;
  for (int r = blockIdx.y; r < R; r += gridDim.y) {
    double re = 0.0, im = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock) {
      double2 v = p[(size_t)t * R + r];
      re += v.x;
      im += v.y;
    }
    s_re[threadIdx.x] = re;
    s_im[threadIdx.x] = im;
    __syncthreads();
    for (int half = kBlock / 2; half > 0; half >>= 1) {
      if (threadIdx.x < half) {
        s_re[threadIdx.x] += s_re[threadIdx.x + half];
        s_im[threadIdx.x] += s_im[threadIdx.x + half];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0)
      mu[(size_t)r * M + m] = make_double2(s_re[0], s_im[0]);
    __syncthreads();
  }
}

// partial[t][r] = sum over the tile of conj(L) psi for each vector (moment 0)
// PAIRS: the R columns are R/2 pairs, partial[t][p] = conj(psi_2p) psi_2p+1
template <bool PAIRS>
__global__ __launch_bounds__(kBlock) void k_block_dot(const float2* __restrict__ psi,
                                                      const float2* __restrict__ L, int R,
                                                      TileMap tm, double2* __restrict__ partial,
                                                      int tiles, int blocks)
{
  __shared__ double s_re[kBlock], s_im[kBlock];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  int b0, e0;
  tile_range(tm, tile, b0, e0);
  const int V = PAIRS ? R / 2 : R;  // products per site
  const size_t begin = (size_t)b0 * V;
  const size_t end = (size_t)e0 * V;
  // thread t accumulates product t % V over the tile, for V dividing kBlock
  double re = 0.0, im = 0.0;
  for (size_t i = begin + threadIdx.x; i < end; i += kBlock) {
    float2 p, l;
    if constexpr (PAIRS) {
      l = psi[2 * i];
      p = psi[2 * i + 1];
    } else {
      p = psi[i];
      l = L[i];
    }
    double px = p.x, py = p.y, lx = l.x, ly = l.y;
    re += px * lx + py * ly;
    im += py * lx - px * ly;
  }
  s_re[threadIdx.x] = re;
  s_im[threadIdx.x] = im;
  __syncthreads();
  for (int r = threadIdx.x; r < V; r += kBlock) {
    double a = 0.0, b = 0.0;
    for (int k = r; k < kBlock; k += V) {
      a += s_re[k];
      b += s_im[k];
    }
    partial[(size_t)tile * V + r] = make_double2(a, b);
  }
}

// ==========================================================================
// time evolution of blocks: FP32 states and recursion, FP64 accumulated state
// ==========================================================================
enum BlockEvo {
  kApplyH = 0,   // out = H~ in
  kCheb2 = 1,    // s2 = 2 H~ s1 - s0; S += bm * phase(s2)
  kComm = 2,     // out = [X, H~] in = -sum t xx in_j / E_max
  kCurrent = 3,  // out = V in = -i sum t xx in_j
  kCheb2x = 4,   // s2 = 2 H~ s1 - s0; s2x = 2 (H~ s1x + [X, H~] s1) - s0x; S += bm * phase(s2x)
};

// the four (-i)^m phase cases of src/hamiltonian.cu:504-525 in FP64
__device__ __forceinline__ void accumulate_phase(double2& s, double bm, double tr, double ti,
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

}  // namespace lsqt

#include "kernels_block_sweep.h"

namespace lsqt {

// Expansion terms into S (kCheb2, kCheb2x): label / bm for this launch's
// term (the new s2, or s2x), label_prev / bm_prev for the term of the
// launch before, whose value at each site is this launch's source s1 (s1x)
// there; label 0 adds nothing, and a launch with both 0 leaves S alone.
// Callers defer every other term to the next launch, so S is read and
// written every second launch; the additions are the same, in the same
// order, as one term per launch.
template <int LAYOUT, int KT, int W, int OP, bool NT>
__global__ __launch_bounds__(kBlock) void k_block_evo(
  HamDev h, int R, TileMap tm, const float2* s0, const float2* __restrict__ s1, float2* s2,
  const float2* s0x, const float2* __restrict__ s1x, float2* s2x, double2* __restrict__ S,
  double bm, int label, double bm_prev, int label_prev, int tiles, int blocks)
{
  constexpr bool kXX = OP == kComm || OP == kCurrent || OP == kCheb2x;
  constexpr bool kOnsite = OP == kApplyH || OP == kCheb2 || OP == kCheb2x;
  __shared__ float2 s_tab[LAYOUT == kPackedLds ? kBlockLdsTableMax : 1];
  __shared__ float s_xx[LAYOUT == kPackedLds && kXX ? kBlockLdsTableMax : 1];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  if constexpr (LAYOUT == kPackedLds) {
    for (int i = threadIdx.x; i < h.ncls; i += kBlock) {
      const double2 d = h.tab[i];
      s_tab[i] = make_float2((float)d.x, (float)d.y);
      if constexpr (kXX)
        s_xx[i] = (float)h.tab_xx[i];
    }
    __syncthreads();
  }
  constexpr int C = SlotBatch<KT>::C;
  const int K = KT > 0 ? KT : h.K;
  const int vpw = R / W;
  const int per_pass = kBlock / vpw;
  const int lane_site = threadIdx.x / vpw;
  const int r0 = (threadIdx.x % vpw) * W;
  int begin, end;
  tile_range(tm, tile, begin, end);
  const float inv_scale = (float)(1.0 / h.emax);
  const bool touch = label != 0 || label_prev != 0;  // the same for the whole launch
  for (int s = begin + lane_site; s < end; s += per_pass) {
    const size_t base = (size_t)s * R + r0;
    float tr[W], ti[W], xr[W], xi[W];
    // the streams read once per pass (psi_{m-2}, the accumulated state) are
    // loaded with the first loads of the pass, so that their latency
    // overlaps the gathers (s0 may be s2: each site is read before written);
    // slot words and potential take ordinary loads (non-temporal: 2 % slower)
    constexpr bool kRecur = OP == kCheb2 || OP == kCheb2x;
    Vec<W> p0, p0x, p1, p1x;
    double2 Sv[W];
    if constexpr (kRecur) {
      p0 = ldv<W, NT>(s0 + base);
      if constexpr (OP == kCheb2x)
        p0x = ldv<W, NT>(s0x + base);
      if (touch) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          Sv[w] = S[base + w];
      }
    }
    if constexpr (kOnsite) {
      const float v = ld<false>(h.pot32 + s);
      p1 = ldv<W, false>(s1 + base);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        tr[w] = v * p1.v[w].x;
        ti[w] = v * p1.v[w].y;
      }
      if constexpr (OP == kCheb2x) {
        p1x = ldv<W, false>(s1x + base);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          xr[w] = v * p1x.v[w].x;
          xi[w] = v * p1x.v[w].y;
        }
      }
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w)
        tr[w] = ti[w] = 0.0f;
    }
    for (int m0 = 0; m0 < K; m0 += C) {
      uint32_t wd[C];
#pragma unroll
      for (int q = 0; q < C; ++q)
        wd[q] = (m0 + q < K) ? ld<false>(h.words + (size_t)(m0 + q) * h.stride + s) : kEmptySlot;
      int j[C];
      float a[C], b[C], x[C];
      decode_words32<LAYOUT, C>(h, s_tab, s, wd, j, a, b);
      if constexpr (kXX) {
#pragma unroll
        for (int q = 0; q < C; ++q) {
          const int c = wd[q] != kEmptySlot ? (int)(wd[q] >> h.idx_bits) : 0;
          x[q] = (LAYOUT == kPackedLds) ? s_xx[c] : (float)h.tab_xx[c];
        }
      }
      Vec<W> c[C], cx[C];
#pragma unroll
      for (int q = 0; q < C; ++q) {
        c[q] = ldv<W, false>(s1 + (size_t)j[q] * R + r0);
        if constexpr (OP == kCheb2x)
          cx[q] = ldv<W, false>(s1x + (size_t)j[q] * R + r0);
      }
#pragma unroll
      for (int q = 0; q < C; ++q) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          const float cr = c[q].v[w].x, ci = c[q].v[w].y;
          // t * in_j (a zero hopping for an absent slot adds nothing)
          const float hr = __builtin_fmaf(a[q], cr, -b[q] * ci);
          const float hi = __builtin_fmaf(a[q], ci, b[q] * cr);
          if constexpr (OP == kApplyH || OP == kCheb2) {
            tr[w] += hr;
            ti[w] += hi;
          } else if constexpr (OP == kComm) {
            tr[w] = __builtin_fmaf(-hr, x[q], tr[w]);
            ti[w] = __builtin_fmaf(-hi, x[q], ti[w]);
          } else if constexpr (OP == kCurrent) {
            tr[w] = __builtin_fmaf(hr, x[q], tr[w]);
            ti[w] = __builtin_fmaf(hi, x[q], ti[w]);
          } else {  // kCheb2x
            tr[w] += hr;
            ti[w] += hi;
            const float dr = cx[q].v[w].x, di = cx[q].v[w].y;
            xr[w] += __builtin_fmaf(a[q], dr, -b[q] * di);
            xi[w] += __builtin_fmaf(a[q], di, b[q] * dr);
            xr[w] = __builtin_fmaf(-hr, x[q], xr[w]);
            xi[w] = __builtin_fmaf(-hi, x[q], xi[w]);
          }
        }
      }
    }
    Vec<W> o;
    if constexpr (OP == kApplyH || OP == kComm) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(tr[w] * inv_scale, ti[w] * inv_scale);
      stv<W>(s2 + base, o);
    } else if constexpr (OP == kCurrent) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(ti[w], -tr[w]);
      stv<W>(s2 + base, o);
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(__builtin_fmaf(2.0f, tr[w] * inv_scale, -p0.v[w].x),
                             __builtin_fmaf(2.0f, ti[w] * inv_scale, -p0.v[w].y));
      stv<W>(s2 + base, o);
      const Vec<W>* acc = &o;
      const Vec<W>* prev = &p1;
      Vec<W> ox;
      if constexpr (OP == kCheb2x) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          ox.v[w] = make_float2(__builtin_fmaf(2.0f, xr[w] * inv_scale, -p0x.v[w].x),
                                __builtin_fmaf(2.0f, xi[w] * inv_scale, -p0x.v[w].y));
        stv<W>(s2x + base, ox);
        acc = &ox;
        prev = &p1x;
      }
      if (touch) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          accumulate_phase(Sv[w], bm_prev, (double)prev->v[w].x, (double)prev->v[w].y,
                           label_prev);
          accumulate_phase(Sv[w], bm, (double)acc->v[w].x, (double)acc->v[w].y, label);
          S[base + w] = Sv[w];
        }
      }
    }
  }
}

// k_block_evo on a lattice with the sliding window (WindowTile): the source
// block s1 (and s1x for the [X, U] term) is staged row by row in LDS and
// neighbours inside the window are read from there; the same arithmetic, in
// the same order, as k_block_evo.
template <int LAYOUT, int KT, int W, int OP>
__global__ __launch_bounds__(kBlock) void k_block_evo_win(
  HamDev h, int R, TileMap tm, const float2* s0, const float2* __restrict__ s1, float2* s2,
  const float2* s0x, const float2* __restrict__ s1x, float2* s2x, double2* __restrict__ S,
  double bm, int label, double bm_prev, int label_prev, int tiles, int blocks)
{
  static_assert(KT > 0, "the window kernel needs a compile-time slot count");
  constexpr bool kXX = OP == kComm || OP == kCurrent || OP == kCheb2x;
  constexpr bool kOnsite = OP == kApplyH || OP == kCheb2 || OP == kCheb2x;
  constexpr bool kRecur = OP == kCheb2 || OP == kCheb2x;
  constexpr bool kTwo = OP == kCheb2x;  // windows of s1 and s1x
  __shared__ float2 s_tab[LAYOUT == kPackedLds ? kBlockLdsTableMax : 1];
  __shared__ float s_xx[LAYOUT == kPackedLds && kXX ? kBlockLdsTableMax : 1];
  extern __shared__ __align__(16) unsigned char s_dyn[];
  float2* s_win = reinterpret_cast<float2*>(s_dyn);
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  if constexpr (LAYOUT == kPackedLds) {
    for (int i = threadIdx.x; i < h.ncls; i += kBlock) {
      const double2 d = h.tab[i];
      s_tab[i] = make_float2((float)d.x, (float)d.y);
      if constexpr (kXX)
        s_xx[i] = (float)h.tab_xx[i];
    }
  }
  const WindowTile<W> wt(tm.win, R, tile);
  float2* s_winx = s_win + 4 * wt.span;
  const int vpw = R / W;
  const int lane_site = threadIdx.x / vpw;
  const int r0 = (threadIdx.x % vpw) * W;
  wt.prologue(s1, s_win);
  if constexpr (kTwo)
    wt.prologue(s1x, s_winx);
  __syncthreads();

  const float inv_scale = (float)(1.0 / h.emax);
  const bool touch = label != 0 || label_prev != 0;  // the same for the whole launch
  constexpr int C = KT;
  int s = wt.xbase + wt.y0 * wt.row + wt.z0 + lane_site;
  uint32_t wnext[C];
  load_words<C, false>(h, s, wnext);
  for (int p = 0; p < wt.by; ++p) {
    const bool more = p + 1 < wt.by;
    Vec<W> na, nb, nax, nbx;
    if (more) {
      wt.fetch(s1, wt.y0 + p + 2, na, nb);
      if constexpr (kTwo)
        wt.fetch(s1x, wt.y0 + p + 2, nax, nbx);
    }
    const int sn = more ? s + wt.row : s;
    const size_t base = (size_t)s * R + r0;
    Vec<W> p0, p0x, p1, p1x;
    double2 Sv[W];
    if constexpr (kRecur) {
      p0 = ldv<W, true>(s0 + base);
      if constexpr (kTwo)
        p0x = ldv<W, true>(s0x + base);
      if (touch) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          Sv[w] = S[base + w];
      }
    }
    const int own = ((p + 1) & 3) * wt.span + (lane_site + 1) * R + r0;
    float tr[W], ti[W], xr[W], xi[W];
    if constexpr (kOnsite) {
      const float v = h.pot32[s];
      p1 = WindowTile<W>::read(s_win, own);
#pragma unroll
      for (int w = 0; w < W; ++w) {
        tr[w] = v * p1.v[w].x;
        ti[w] = v * p1.v[w].y;
      }
      if constexpr (kTwo) {
        p1x = WindowTile<W>::read(s_winx, own);
#pragma unroll
        for (int w = 0; w < W; ++w) {
          xr[w] = v * p1x.v[w].x;
          xi[w] = v * p1x.v[w].y;
        }
      }
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w)
        tr[w] = ti[w] = 0.0f;
    }
    uint32_t wd[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      wd[q] = wnext[q];
    load_words<C, false>(h, sn, wnext);
    int j[C];
    float a[C], b[C], x[C];
    decode_words32<LAYOUT, C>(h, s_tab, s, wd, j, a, b);
    if constexpr (kXX) {
#pragma unroll
      for (int q = 0; q < C; ++q) {
        const int c = wd[q] != kEmptySlot ? (int)(wd[q] >> h.idx_bits) : 0;
        x[q] = (LAYOUT == kPackedLds) ? s_xx[c] : (float)h.tab_xx[c];
      }
    }
    int pos[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      pos[q] = wt.pos(j[q], p, r0);
    Vec<W> c[C], cx[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (pos[q] < 0) {
        c[q] = ldv<W, false>(s1 + (size_t)j[q] * R + r0);
        if constexpr (kTwo)
          cx[q] = ldv<W, false>(s1x + (size_t)j[q] * R + r0);
      }
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (pos[q] >= 0) {
        c[q] = WindowTile<W>::read(s_win, pos[q]);
        if constexpr (kTwo)
          cx[q] = WindowTile<W>::read(s_winx, pos[q]);
      }
#pragma unroll
    for (int q = 0; q < C; ++q) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const float cr = c[q].v[w].x, ci = c[q].v[w].y;
        const float hr = __builtin_fmaf(a[q], cr, -b[q] * ci);
        const float hi = __builtin_fmaf(a[q], ci, b[q] * cr);
        if constexpr (OP == kApplyH || OP == kCheb2) {
          tr[w] += hr;
          ti[w] += hi;
        } else if constexpr (OP == kComm) {
          tr[w] = __builtin_fmaf(-hr, x[q], tr[w]);
          ti[w] = __builtin_fmaf(-hi, x[q], ti[w]);
        } else if constexpr (OP == kCurrent) {
          tr[w] = __builtin_fmaf(hr, x[q], tr[w]);
          ti[w] = __builtin_fmaf(hi, x[q], ti[w]);
        } else {  // kCheb2x
          tr[w] += hr;
          ti[w] += hi;
          const float dr = cx[q].v[w].x, di = cx[q].v[w].y;
          xr[w] += __builtin_fmaf(a[q], dr, -b[q] * di);
          xi[w] += __builtin_fmaf(a[q], di, b[q] * dr);
          xr[w] = __builtin_fmaf(-hr, x[q], xr[w]);
          xi[w] = __builtin_fmaf(-hi, x[q], xi[w]);
        }
      }
    }
    Vec<W> o;
    if constexpr (OP == kApplyH || OP == kComm) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(tr[w] * inv_scale, ti[w] * inv_scale);
      stv<W>(s2 + base, o);
    } else if constexpr (OP == kCurrent) {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(ti[w], -tr[w]);
      stv<W>(s2 + base, o);
    } else {
#pragma unroll
      for (int w = 0; w < W; ++w)
        o.v[w] = make_float2(__builtin_fmaf(2.0f, tr[w] * inv_scale, -p0.v[w].x),
                             __builtin_fmaf(2.0f, ti[w] * inv_scale, -p0.v[w].y));
      stv<W>(s2 + base, o);
      const Vec<W>* acc = &o;
      const Vec<W>* prev = &p1;
      Vec<W> ox;
      if constexpr (OP == kCheb2x) {
#pragma unroll
        for (int w = 0; w < W; ++w)
          ox.v[w] = make_float2(__builtin_fmaf(2.0f, xr[w] * inv_scale, -p0x.v[w].x),
                                __builtin_fmaf(2.0f, xi[w] * inv_scale, -p0x.v[w].y));
        stv<W>(s2x + base, ox);
        acc = &ox;
        prev = &p1x;
      }
      if (touch) {
#pragma unroll
        for (int w = 0; w < W; ++w) {
          accumulate_phase(Sv[w], bm_prev, (double)prev->v[w].x, (double)prev->v[w].y,
                           label_prev);
          accumulate_phase(Sv[w], bm, (double)acc->v[w].x, (double)acc->v[w].y, label);
          S[base + w] = Sv[w];
        }
      }
    }
    if (more) {
      wt.put(s_win, (p + 3) & 3, na, nb);
      if constexpr (kTwo)
        wt.put(s_winx, (p + 3) & 3, nax, nbx);
    }
    __syncthreads();
    s = sn;
  }
}

// partial[t][r] = (sum over the tile of |S_r|^2, 0), FP64 state
__global__ __launch_bounds__(kBlock) void k_block_norm64(const double2* __restrict__ S, int R,
                                                         TileMap tm, double2* __restrict__ partial,
                                                         int tiles, int blocks)
{
  __shared__ double s_re[kBlock];
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  int b0, e0;
  tile_range(tm, tile, b0, e0);
  double acc = 0.0;
  for (size_t i = (size_t)b0 * R + threadIdx.x; i < (size_t)e0 * R; i += kBlock) {
    const double2 v = S[i];
    acc += v.x * v.x + v.y * v.y;
  }
  s_re[threadIdx.x] = acc;
  __syncthreads();
  for (int r = threadIdx.x; r < R; r += kBlock) {
    double a = 0.0;
    for (int k = r; k < kBlock; k += R)
      a += s_re[k];
    partial[(size_t)tile * R + r] = make_double2(a, 0.0);
  }
}

// S_r *= sqrt(before_r / after_r) per segment and vector (U is unitary: the
// norm before an evolution step is the exact norm after it)
__global__ __launch_bounds__(kBlock) void k_block_renorm(double2* __restrict__ S, int R, TileMap tm,
                                                         const double2* __restrict__ before,
                                                         const double2* __restrict__ after,
                                                         int tiles, int blocks)
{
  int tile = xcd_tile(blockIdx.x, blocks);
  if (tile >= tiles)
    return;
  int b0, e0;
  tile_range(tm, tile, b0, e0);
  const int seg = tm.tile_seg ? tm.tile_seg[tile] : 0;
  for (size_t i = (size_t)b0 * R + threadIdx.x; i < (size_t)e0 * R; i += kBlock) {
    const int r = (int)(i % (size_t)R);
    const double a = after[(size_t)seg * R + r].x, b = before[(size_t)seg * R + r].x;
    const double f = a > 0.0 ? sqrt(b / a) : 1.0;
    double2 v = S[i];
    v.x *= f;
    v.y *= f;
    S[i] = v;
  }
}

// Layout changes between the host's vector-major arrays (vector r at
// [r*n, (r+1)*n)) and site-major blocks ([n][R]): one thread per value,
// coalesced on the vector-major side.
template <class T2>
__global__ __launch_bounds__(kBlock) void k_site_major(const double* __restrict__ re,
                                                      const double* __restrict__ im, int n,
                                                      int R, T2* __restrict__ dst)
{
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n * R)
    return;
  const int r = idx / n, i = idx - r * n;
  T2 v;
  v.x = re[idx];
  v.y = im[idx];
  dst[(size_t)i * R + r] = v;
}

__global__ __launch_bounds__(kBlock) void k_vector_major(const float2* __restrict__ src, int n,
                                                        int R, double* __restrict__ re,
                                                        double* __restrict__ im)
{
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n * R)
    return;
  const int r = idx / n, i = idx - r * n;
  const float2 v = src[(size_t)i * R + r];
  re[idx] = v.x;
  im[idx] = v.y;
}

// dst = S_z src for every vector: site i keeps its sign when i is even and
// changes it when odd (src/vector.cu:159-172)
__global__ __launch_bounds__(kBlock) void k_block_sz(const float2* __restrict__ src, int n, int R,
                                                    float2* __restrict__ dst)
{
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= n * R)
    return;
  const float2 v = src[idx];
  dst[idx] = ((idx / R) % 2 == 0) ? v : make_float2(-v.x, -v.y);
}

// dst[s][2r] = L[s][r], dst[s][2r + 1] = Rb[s][r]: two R-vector blocks as one
// block of R pairs (columns 2r, 2r + 1); n * R < 2^31
__global__ __launch_bounds__(kBlock) void k_block_interleave(const float2* __restrict__ L,
                                                            const float2* __restrict__ Rb, int nR,
                                                            float2* __restrict__ dst)
{
  const int idx = blockIdx.x * kBlock + threadIdx.x;
  if (idx >= nR)
    return;
  fvec4_t t;
  const float2 a = L[idx], b = Rb[idx];
  t.x = a.x;
  t.y = a.y;
  t.z = b.x;
  t.w = b.y;
  __builtin_nontemporal_store(t, reinterpret_cast<fvec4_t*>(dst) + idx);
}

// ==========================================================================
// launchers
// ==========================================================================
// the block kernel stages class tables of up to kBlockLdsTableMax entries
// in LDS (one table per tile of `rows` sites costs little next to R vectors)
#define LSQTB_LAYOUT_SWITCH(h, L, ...)                                   \
  if (h.layout == kPackedLds ||                                          \
      (h.layout == kPackedGlobal && h.ncls <= kBlockLdsTableMax)) {      \
    constexpr int L = kPackedLds;           \
    __VA_ARGS__;                            \
  } else {                                  \
    constexpr int L = kPackedGlobal;        \
    __VA_ARGS__;                            \
  }

#define LSQTB_K_SWITCH(kval, KT, ...) \
  switch (kval) {                     \
    case 3: {                         \
      constexpr int KT = 3;           \
      __VA_ARGS__;                    \
    } break;                          \
    case 4: {                         \
      constexpr int KT = 4;           \
      __VA_ARGS__;                    \
    } break;                          \
    case 6: {                         \
      constexpr int KT = 6;           \
      __VA_ARGS__;                    \
    } break;                          \
    default: {                        \
      constexpr int KT = 0;           \
      __VA_ARGS__;                    \
    } break;                          \
  }

#define LSQTB_W_SWITCH(wval, W, ...) \
  switch (wval) {                    \
    case 1: {                        \
      constexpr int W = 1;           \
      __VA_ARGS__;                   \
    } break;                         \
    case 2: {                        \
      constexpr int W = 2;           \
      __VA_ARGS__;                   \
    } break;                         \
    case 4: {                        \
      constexpr int W = 4;           \
      __VA_ARGS__;                   \
    } break;                         \
    default: {                       \
      constexpr int W = 8;           \
      __VA_ARGS__;                   \
    } break;                         \
  }

#define LSQTB_WIN_W_SWITCH(wval, W, ...) \
  switch (wval) {                        \
    case 1: {                            \
      constexpr int W = 1;               \
      __VA_ARGS__;                       \
    } break;                             \
    case 2: {                            \
      constexpr int W = 2;               \
      __VA_ARGS__;                       \
    } break;                             \
    default: {                           \
      constexpr int W = 4;               \
      __VA_ARGS__;                       \
    } break;                             \
  }

TileMap make_tile_map(int n, int rows, int slab, int span, int row, int walk_pass)
{
  TileMap t{};
  t.n = n;
  t.rows = rows;
  if (slab > 0 && span > 0 && slab % span == 0 && n % slab == 0) {
    t.slab = slab;
    t.span = span;
    t.slabs = n / slab;
    t.tiles = (slab / span) * t.slabs;
    // the column walk needs tiles of whole rows cut into whole passes; any
    // other request keeps the consecutive pass order (same sites, same sums)
    if (walk_pass > 0 && row > 0 && span % row == 0 && row % walk_pass == 0) {
      t.row = row;
      t.walk = 1;
    }
  } else {
    t.tiles = (n + rows - 1) / rows;
  }
  return t;
}

TileMap make_window_map(int n, int slab, int row, int by, int chunk)
{
  TileMap t = make_tile_map(n, by * chunk, 0, 0, 0, 0);
  WindowMap& w = t.win;
  w.slab = slab;
  w.row = row;
  w.slabs = n / slab;
  w.chunks = row / chunk;
  w.ny = slab / row;
  w.by = by;
  w.chunk = chunk;
  w.row_div = make_divisor(row);
  return t;
}

bool window_fits(int n, int slab, int row, int by, int R, int W)
{
  if (W > 4 || slab <= 0 || row < 2 || by < 1)
    return false;
  const int chunk = kBlock / (R / W);
  return n % slab == 0 && slab % row == 0 && slab / row >= 3 && (slab / row) % by == 0 &&
         row % chunk == 0;
}

Grid block_grid(const TileMap& t)
{
  int blocks = (t.tiles + kXcds - 1) / kXcds * kXcds;
  return Grid{t.tiles, blocks};
}

void launch_block_step(const HamDev& h, int R, int W, bool db, const TileMap& tm,
                       const float2* psi1, const float2* psi0, float2* out, const float2* L,
                       double2* partial, double2* partial2, bool recur, int products,
                       hipStream_t st)
{
  static_assert((int)kStepLeft == (int)kLeft && (int)kStepSelf == (int)kSelf &&
                  (int)kStepPair == (int)kPair,
                "product modes");
  Grid g = block_grid(tm);
  if (tm.win.by > 0 && W <= 4 && (h.K == 3 || h.K == 4 || h.K == 6)) {
    // ring of four window rows; the tile sums reuse the same memory
    const size_t lds = std::max(sizeof(float2) * 4 * (size_t)(tm.win.chunk + 2) * R,
                                sizeof(double) * 2 * (kBlock / 64) * 64 * 4);
#define LSQTB_WIN(RC, MODE)                                                                 \
  k_block_win<LAY, KT, WW, RC, MODE><<<g.blocks, kBlock, lds, st>>>(h, R, tm, psi1, psi0, out, \
                                                                   L, partial, partial2,       \
                                                                   g.tiles, g.blocks)
#define LSQTB_WIN_MODE                                                  \
  if (products == kPair) {                                              \
    if constexpr (WW >= 2)                                              \
      recur ? LSQTB_WIN(true, kPair) : LSQTB_WIN(false, kPair);         \
  } else if (products == kSelf) {                                       \
    recur ? LSQTB_WIN(true, kSelf) : LSQTB_WIN(false, kSelf);           \
  } else {                                                              \
    recur ? LSQTB_WIN(true, kLeft) : LSQTB_WIN(false, kLeft);           \
  }
    LSQTB_LAYOUT_SWITCH(h, LAY, {
      switch (h.K) {
        case 3: {
          constexpr int KT = 3;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_WIN_MODE)
        } break;
        case 4: {
          constexpr int KT = 4;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_WIN_MODE)
        } break;
        default: {
          constexpr int KT = 6;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_WIN_MODE)
        } break;
      }
    })
#undef LSQTB_WIN_MODE
#undef LSQTB_WIN
    return;
  }
#define LSQTB_STEP(RC, MODE, DBV)                                                 \
  k_block_step<LAY, KT, WW, RC, MODE, false, DBV><<<g.blocks, kBlock, 0, st>>>(   \
    h, R, tm, psi1, psi0, out, L, partial, partial2, g.tiles, g.blocks)
#define LSQTB_STEP_DB(DBV)                                                  \
  if (products == kPair) {                                                  \
    if constexpr (WW >= 2)                                                  \
      recur ? LSQTB_STEP(true, kPair, DBV) : LSQTB_STEP(false, kPair, DBV); \
  } else if (products == kSelf) {                                           \
    recur ? LSQTB_STEP(true, kSelf, DBV) : LSQTB_STEP(false, kSelf, DBV);   \
  } else {                                                                  \
    recur ? LSQTB_STEP(true, kLeft, DBV) : LSQTB_STEP(false, kLeft, DBV);   \
  }
  LSQTB_LAYOUT_SWITCH(h, LAY, LSQTB_K_SWITCH(h.K, KT, LSQTB_W_SWITCH(W, WW, {
    if (db && KT > 0) {
      LSQTB_STEP_DB(true)
    } else {
      LSQTB_STEP_DB(false)
    }
  })))
#undef LSQTB_STEP_DB
#undef LSQTB_STEP
}

size_t block_sweep_lds(int W)
{
  return W == 2 ? block_sweep_lds_bytes<2, 4>() : block_sweep_lds_bytes<4, 1>();
}

// W = 2 (blocks of 4 slabs) or 4 (one slab at a time: the rings of wider
// entries in the same 49 KB of LDS), slot counts 2, 3, 4 and 6
void launch_block_sweep2(const HamDev& h, const BlockSweepMap& sm, int cols, int W, bool pairs,
                         const float2* p2, const float2* p1, float2* o0, float2* o1,
                         double2* const (&part)[4], hipStream_t st)
{
  const size_t lds = block_sweep_lds(W);
#define LSQTB_SW(KK, WW, BB)                                                                \
  do {                                                                                      \
    if (pairs)                                                                              \
      k_block_sweep2<KK, WW, kPair, BB><<<sm.blocks, kBlock, lds, st>>>(                    \
        h, sm, cols, p2, p1, o0, o1, part[0], part[1], part[2], part[3]);                   \
    else                                                                                    \
      k_block_sweep2<KK, WW, kSelf, BB><<<sm.blocks, kBlock, lds, st>>>(                    \
        h, sm, cols, p2, p1, o0, o1, part[0], part[1], part[2], part[3]);                   \
  } while (0)
#define LSQTB_SW_K(KK)        \
  do {                        \
    if (W == 2)               \
      LSQTB_SW(KK, 2, 4);     \
    else                      \
      LSQTB_SW(KK, 4, 1);     \
  } while (0)
  switch (h.K) {
    case 2:
      LSQTB_SW_K(2);
      break;
    case 3:
      LSQTB_SW_K(3);
      break;
    case 4:
      LSQTB_SW_K(4);
      break;
    case 6:
      LSQTB_SW_K(6);
      break;
    default:
      break;
  }
#undef LSQTB_SW_K
#undef LSQTB_SW
}

size_t block_sweepn_lds()
{
  return block_sweepn_lds_bytes<2, kBlockSweepSteps, 1>();
}

// four steps per launch, one slab per loop iteration, 2 columns per lane
// (cols <= 16): the rings of four levels in 48 KB of LDS
void launch_block_sweepn(const HamDev& h, const BlockSweepMap& sm, int cols, bool pairs,
                         const float2* p2, const float2* p1, float2* o0, float2* o1,
                         double2* part, size_t part_stride, hipStream_t st)
{
  const size_t lds = block_sweepn_lds();
#define LSQTB_SWN(KK)                                                                       \
  do {                                                                                      \
    if (pairs)                                                                              \
      k_block_sweepn<KK, 2, kPair, kBlockSweepSteps, 1><<<sm.blocks, kBlock, lds, st>>>(                   \
        h, sm, cols, p2, p1, o0, o1, part, part_stride);                                    \
    else                                                                                    \
      k_block_sweepn<KK, 2, kSelf, kBlockSweepSteps, 1><<<sm.blocks, kBlock, lds, st>>>(                   \
        h, sm, cols, p2, p1, o0, o1, part, part_stride);                                    \
  } while (0)
  switch (h.K) {
    case 2:
      LSQTB_SWN(2);
      break;
    case 3:
      LSQTB_SWN(3);
      break;
    case 4:
      LSQTB_SWN(4);
      break;
    case 6:
      LSQTB_SWN(6);
      break;
    default:
      break;
  }
#undef LSQTB_SWN
}

// four terms of U(t) per launch on blocks of cols <= 16 columns
void launch_block_sweepn_evo(const HamDev& h, const BlockSweepMap& sm, int cols,
                             const float2* p2, const float2* p1, float2* o0, float2* o1,
                             double2* state, const SweepTerms& terms, hipStream_t st)
{
  const size_t lds = block_sweepn_lds();
#define LSQTB_SWNE(KK)                                                                     \
  k_block_sweepn<KK, 2, kSelf, kBlockSweepSteps, 1, true><<<sm.blocks, kBlock, lds, st>>>(                \
    h, sm, cols, p2, p1, o0, o1, nullptr, 0, state, terms)
  switch (h.K) {
    case 2:
      LSQTB_SWNE(2);
      break;
    case 3:
      LSQTB_SWNE(3);
      break;
    case 4:
      LSQTB_SWNE(4);
      break;
    case 6:
      LSQTB_SWNE(6);
      break;
    default:
      break;
  }
#undef LSQTB_SWNE
}

int block_sweep_check_blocks(const BlockSweepSeg& g)
{
  return (int)(((long long)g.S * g.nslab + kBlock - 1) / kBlock);
}

void launch_block_sweep_check(const HamDev& h, const BlockSweepSeg& g, const int4* tab, int* out,
                              hipStream_t st)
{
  k_block_sweep_check<<<block_sweep_check_blocks(g), kBlock, 0, st>>>(h, g, tab, out);
}

void launch_block_dot(const float2* psi, const float2* L, int R, const TileMap& tm,
                      double2* partial, hipStream_t st, bool pairs)
{
  Grid g = block_grid(tm);
  if (pairs)
    k_block_dot<true><<<g.blocks, kBlock, 0, st>>>(psi, psi, R, tm, partial, g.tiles, g.blocks);
  else
    k_block_dot<false><<<g.blocks, kBlock, 0, st>>>(psi, L, R, tm, partial, g.tiles, g.blocks);
}

void launch_block_reduce(const double2* partial, const TileMap& tm, int R, int M, int m0,
                         int count, double2* mu, hipStream_t st, int stride_tiles)
{
  if (count <= 0)
    return;
  const int segs = tm.seg_tiles ? tm.segs : 1;
  dim3 grid(count, R < 8 ? R : 8, segs);
  k_block_reduce<<<grid, kBlock, 0, st>>>(partial, tm.tiles,
                                          stride_tiles > 0 ? stride_tiles : tm.tiles, R, M, m0,
                                          tm.seg_tiles, mu);
}

void launch_block_evo(const HamDev& h, int R, int W, const TileMap& tm, int op,
                      const float2* s0, const float2* s1, float2* s2, const float2* s0x,
                      const float2* s1x, float2* s2x, double2* S, double bm, int label,
                      hipStream_t st, double bm_prev, int label_prev)
{
  Grid g = block_grid(tm);
  // one ring of four window rows per source block (two for kCheb2x), beside
  // the class tables (12 KB); shapes whose rings exceed the 64 KB of LDS a
  // workgroup may hold take the kernel without the window
  const size_t lds = sizeof(float2) * 4 * (size_t)(tm.win.chunk + 2) * R * (op == kCheb2x ? 2 : 1);
  if (tm.win.by > 0 && W <= 4 && (h.K == 3 || h.K == 4 || h.K == 6) &&
      lds + sizeof(float2) * kBlockLdsTableMax + sizeof(float) * kBlockLdsTableMax <= 65536) {
#define LSQTB_EVOW(OPV)                                                                    \
  k_block_evo_win<LAY, KT, WW, OPV><<<g.blocks, kBlock, lds, st>>>(h, R, tm, s0, s1, s2, s0x, \
                                                                  s1x, s2x, S, bm, label,     \
                                                                  bm_prev, label_prev, g.tiles, \
                                                                  g.blocks)
#define LSQTB_EVOW_OP                   \
  switch (op) {                         \
    case kApplyH:                       \
      LSQTB_EVOW(kApplyH);              \
      break;                            \
    case kCheb2:                        \
      LSQTB_EVOW(kCheb2);               \
      break;                            \
    case kComm:                         \
      LSQTB_EVOW(kComm);                \
      break;                            \
    case kCurrent:                      \
      LSQTB_EVOW(kCurrent);             \
      break;                            \
    default:                            \
      LSQTB_EVOW(kCheb2x);              \
      break;                            \
  }
    LSQTB_LAYOUT_SWITCH(h, LAY, {
      switch (h.K) {
        case 3: {
          constexpr int KT = 3;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_EVOW_OP)
        } break;
        case 4: {
          constexpr int KT = 4;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_EVOW_OP)
        } break;
        default: {
          constexpr int KT = 6;
          LSQTB_WIN_W_SWITCH(W, WW, LSQTB_EVOW_OP)
        } break;
      }
    })
#undef LSQTB_EVOW_OP
#undef LSQTB_EVOW
    return;
  }
  const bool nt = (size_t)h.n * R >= (size_t)kStreamingMinSites;
#define LSQTB_EVO(OPV, NTV)                                                              \
  k_block_evo<LAY, KT, WW, OPV, NTV><<<g.blocks, kBlock, 0, st>>>(                       \
    h, R, tm, s0, s1, s2, s0x, s1x, s2x, S, bm, label, bm_prev, label_prev, g.tiles, g.blocks)
#define LSQTB_EVO_OP(NTV)                  \
  switch (op) {                            \
    case kApplyH:                          \
      LSQTB_EVO(kApplyH, NTV);             \
      break;                               \
    case kCheb2:                           \
      LSQTB_EVO(kCheb2, NTV);              \
      break;                               \
    case kComm:                            \
      LSQTB_EVO(kComm, NTV);               \
      break;                               \
    case kCurrent:                         \
      LSQTB_EVO(kCurrent, NTV);            \
      break;                               \
    default:                               \
      LSQTB_EVO(kCheb2x, NTV);             \
      break;                               \
  }
  LSQTB_LAYOUT_SWITCH(h, LAY, LSQTB_K_SWITCH(h.K, KT, LSQTB_W_SWITCH(W, WW, {
    if (nt) {
      LSQTB_EVO_OP(true)
    } else {
      LSQTB_EVO_OP(false)
    }
  })))
#undef LSQTB_EVO_OP
#undef LSQTB_EVO
}

void launch_block_norm64(const double2* S, int R, const TileMap& tm, double2* partial,
                         hipStream_t st)
{
  Grid g = block_grid(tm);
  k_block_norm64<<<g.blocks, kBlock, 0, st>>>(S, R, tm, partial, g.tiles, g.blocks);
}

void launch_block_renorm(double2* S, int R, const TileMap& tm, const double2* before,
                         const double2* after, hipStream_t st)
{
  Grid g = block_grid(tm);
  k_block_renorm<<<g.blocks, kBlock, 0, st>>>(S, R, tm, before, after, g.tiles, g.blocks);
}

void launch_site_major(const double* re, const double* im, int n, int R, double2* dst,
                       hipStream_t st)
{
  const int total = n * R;
  k_site_major<double2><<<(total + kBlock - 1) / kBlock, kBlock, 0, st>>>(re, im, n, R, dst);
}

void launch_site_major(const double* re, const double* im, int n, int R, float2* dst,
                       hipStream_t st)
{
  const int total = n * R;
  k_site_major<float2><<<(total + kBlock - 1) / kBlock, kBlock, 0, st>>>(re, im, n, R, dst);
}

void launch_vector_major(const float2* src, int n, int R, double* re, double* im,
                         hipStream_t st)
{
  const int total = n * R;
  k_vector_major<<<(total + kBlock - 1) / kBlock, kBlock, 0, st>>>(src, n, R, re, im);
}

void launch_block_interleave(const float2* L, const float2* Rb, int n, int R, float2* dst,
                             hipStream_t st)
{
  const int total = n * R;
  k_block_interleave<<<(total + kBlock - 1) / kBlock, kBlock, 0, st>>>(L, Rb, total, dst);
}

void launch_block_sz(const float2* src, int n, int R, float2* dst, hipStream_t st)
{
  const int total = n * R;
  k_block_sz<<<(total + kBlock - 1) / kBlock, kBlock, 0, st>>>(src, n, R, dst);
}

}  // namespace lsqt
