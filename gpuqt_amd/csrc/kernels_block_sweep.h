// kernels_block_sweep.h -- two Chebyshev steps per launch for blocks of
// vectors on periodic slab lattices (included by kernels_block.hip, after
// its vector, product and tile-sum helpers).
//
// The block form of k_sweep2 (kernels_sweep.h): a workgroup owns a chunk of
// positions of a run of slabs of one system and walks the run slab by slab,
// holding in LDS B + 2 slabs of psi_{m-1} (the chunk and 2h positions either
// side) and B + 2 slabs of psi_m it computed (the chunk and h either side),
// and advances B slabs at a time: psi_m at slabs k .. k + B - 1, then
// psi_{m+1} at slabs k - 1 .. k + B - 2.  An entry of a slab row is one site
// with all of its columns: VPW = cols / W lanes per site, each holding W
// adjacent columns (the block step's lane layout), so a 256-thread row holds
// 256 / VPW positions.
//
// It serves the product-identity moment sets (self products of R vectors,
// pairs of a left and a right recursion in 2R columns), whose products need
// only the step's own input and output at a site: step 1 adds S_{2m-1} and
// S_{2m} of psi_{m-1} and psi_m, step 2 S_{2m+1} and S_{2m+2} of psi_m and
// psi_{m+1}, four ring slots per launch.  Per site the arithmetic is the
// block step's (FP32, the on-site term, then every slot in slot order with
// explicit FMAs, unused slots as zero hoppings on the site's own value, the
// recursion as one FMA), so the states are those of one-step launches bit
// for bit; only the moment partials group the sites by sweep tile.
//
// Per site and launch: psi_{m-2} and psi_{m-1} read, psi_m and psi_{m+1}
// written (8 cols bytes each), the FP32 potential read: no slot words (the
// per-orbital table gives the neighbours), against 3 vectors and 4 + 4K
// bytes per site for each one-step launch.
//
// A segmented Hamiltonian (many systems of different size) runs as one
// launch: dispatch slot -> (tile, system, chunk, run) from a table that
// puts the longest runs first on every XCD.
#pragma once

#include <type_traits>

namespace lsqt {

template <int W>
__device__ __forceinline__ Vec<W> lds_vec(const float2* p)
{
  Vec<W> r;
#pragma unroll
  for (int i = 0; i < W; i += 2) {
    const float4 t = *reinterpret_cast<const float4*>(p + i);
    r.v[i] = make_float2(t.x, t.y);
    r.v[i + 1] = make_float2(t.z, t.w);
  }
  return r;
}

template <int W>
__device__ __forceinline__ void lds_put(float2* p, const Vec<W>& a)
{
#pragma unroll
  for (int i = 0; i < W; i += 2)
    *reinterpret_cast<float4*>(p + i) = make_float4(a.v[i].x, a.v[i].y, a.v[i + 1].x, a.v[i + 1].y);
}

// dynamic LDS of k_block_sweep2<., W, ., B>: two rings of B + 2 rows of
// kBlock entries of W columns, then one FP32 potential per thread (the last
// slab of the previous block, used by step 2)
template <int W, int B>
constexpr size_t block_sweep_lds_bytes()
{
  return sizeof(float2) * 2 * (B + 2) * kBlock * W + sizeof(float) * kBlock;
}

template <int K, int W, int IPM, int B>
__global__ __launch_bounds__(kBlock) void k_block_sweep2(
  HamDev h, BlockSweepMap sm, int cols, const float2* __restrict__ p2,
  const float2* __restrict__ p1, float2* __restrict__ o0, float2* __restrict__ o1,
  double2* __restrict__ s0, double2* __restrict__ s1, double2* __restrict__ s2,
  double2* __restrict__ s3)
{
  static_assert(W % 2 == 0, "W even: 16-byte LDS entries");
  static_assert(block_sweep_lds_bytes<W, B>() >= sizeof(double) * 2 * (kBlock / 64) * 64 * 4,
                "the tile sums reuse the rings");
  constexpr int RR = B + 2;        // ring rows
  constexpr int RW = kBlock * W;   // ring row length in float2
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int4 od = sm.order[xcd_tile(blockIdx.x, sm.blocks)];
  if (od.x < 0)
    return;
  const BlockSweepSeg g = sm.seg[od.y];
  float2* s_w = reinterpret_cast<float2*>(s_dyn);  // psi_{m-1}: sweep index q in row (q + 1) mod RR
  float2* s_m = s_w + RR * RW;                      // psi_m: sweep index q in row q mod RR
  float* s_cv = reinterpret_cast<float*>(s_m + RR * RW);

  const int t = threadIdx.x;
  const int vpw = cols / W;
  const int tpos = t / vpw;                  // held entry of this thread's site
  const int r0 = (t - tpos * vpw) * W;       // its first column
  const int S = g.S, nslab = g.nslab, hh = g.hh;
  const int c0 = od.z * g.cs;
  const int clen = min(g.cs, S - c0);
  const int wlen = clen + 4 * hh;  // psi_{m-1} positions held per slab
  const int len1 = clen + 2 * hh;  // psi_m positions computed per slab
  int ws = c0 - 2 * hh;            // first position held (periodic)
  if (ws < 0)
    ws += S;
  const int run = od.w;
  const int x_lo = (int)((long long)run * nslab / g.nruns);
  const int x_hi = (int)((long long)(run + 1) * nslab / g.nruns);
  const int bx = x_hi - x_lo;
  const int dir = (run & 1) ? -1 : 1;
  const int start = dir > 0 ? x_lo - 1 : x_hi;  // slab of sweep index 0
  auto slab_of = [&](int q) {  // clamped past the run's end (values unused)
    q = q > bx + 2 ? bx + 2 : q;
    int x = start + dir * q;
    x = x < 0 ? x + nslab : x;
    return x >= nslab ? x - nslab : x;
  };
  auto wpos = [&](int w) {
    const int p = ws + w;
    return p >= S ? p - S : p;
  };
  auto ring = [](int q) {  // q mod RR for q >= -RR
    return (q + RR) % RR;
  };
  // step 1 computes held entry tpos + hh (tpos < len1), step 2 the same
  // site for the interior entries
  const bool act1 = tpos < len1;
  const bool inner = tpos >= hh && tpos < hh + clen;
  const int e1 = act1 ? tpos + hh : 0;
  const int pos1 = wpos(e1 < wlen ? e1 : 0);
  const int pos_a = wpos(tpos < wlen ? tpos : 0);
  const int o_self = e1 * vpw * W + r0;  // ring-row offset of the site's own entry
  const int o_mid = hh * vpw * W;        // psi_m rows hold entry e at e - hh
  const int r = pos1 % g.p;
  int eo[K], dq[K];
  float ha[K], hb[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int4 e = sm.tab[g.tab + q * g.p + r];
    const bool ok = e.w != 0;
    int b = e1 + (ok ? e.y : 0);
    b = b < 0 ? b + S : (b >= S ? b - S : b);  // the whole slab in one chunk
    b = b < 0 ? 0 : (b < wlen ? b : wlen - 1);
    eo[q] = b * vpw * W + r0;
    dq[q] = ok ? e.x * dir : 0;
    const double2 hv = h.tab[ok ? e.z : 0];
    ha[q] = ok ? (float)hv.x : 0.0f;
    hb[q] = ok ? (float)hv.y : 0.0f;
  }
  const size_t base = (size_t)g.base;
  auto gidx = [&](int slab, int pos) {
    return (base + (size_t)slab * S + pos) * (size_t)cols + r0;
  };
  auto fetch = [&](int q) { return ldv<W, false>(p1 + gidx(slab_of(q), pos_a)); };
  auto put = [&](int q, const Vec<W>& a) { lds_put<W>(s_w + ring(q + 1) * RW + t * W, a); };
  {
    const Vec<W> a0 = fetch(-1), a1 = fetch(0);
    put(-1, a0);
    put(0, a1);
  }
  const float inv_scale = (float)(1.0 / h.emax);
  auto site = [&](const Vec<W>& self, float v, const Vec<W> (&c)[K], const Vec<W>& prev) {
    float tr[W], ti[W];
#pragma unroll
    for (int w = 0; w < W; ++w) {
      tr[w] = v * self.v[w].x;
      ti[w] = v * self.v[w].y;
    }
#pragma unroll
    for (int q = 0; q < K; ++q) {
#pragma unroll
      for (int w = 0; w < W; ++w) {
        const float cx = c[q].v[w].x, cy = c[q].v[w].y;
        tr[w] = __builtin_fmaf(ha[q], cx, tr[w]);
        tr[w] = __builtin_fmaf(-hb[q], cy, tr[w]);
        ti[w] = __builtin_fmaf(ha[q], cy, ti[w]);
        ti[w] = __builtin_fmaf(hb[q], cx, ti[w]);
      }
    }
    Vec<W> o;
#pragma unroll
    for (int w = 0; w < W; ++w) {
      o.v[w].x = __builtin_fmaf(2.0f, tr[w] * inv_scale, -prev.v[w].x);
      o.v[w].y = __builtin_fmaf(2.0f, ti[w] * inv_scale, -prev.v[w].y);
    }
    return o;
  };

  Products<W> acc0, acc1;
  clear_products(acc0);
  clear_products(acc1);
  for (int k0 = 0; k0 <= bx + 1; k0 += B) {
    // every global load of the block is issued here and used below
    Vec<W> w[B], ip2[B];
    float iv[B];
#pragma unroll
    for (int i = 0; i < B; ++i)
      w[i] = fetch(k0 + 1 + i);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      const size_t at = base + (size_t)slab_of(k > bx + 1 ? bx + 1 : k) * S + pos1;
      ip2[i] = ldv<W, true>(p2 + at * cols + r0);
      iv[i] = h.pot32[at];
    }
    const float cv_prev = s_cv[t];
    s_cv[t] = iv[B - 1];
#pragma unroll
    for (int i = 0; i < B; ++i)
      put(k0 + 1 + i, w[i]);
    __syncthreads();
    // step 1: psi_m at slabs k0 + i (stored after all B are computed)
    Vec<W> out[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      const Vec<W> self = lds_vec<W>(s_w + ring(k + 1) * RW + o_self);
      Vec<W> c[K];
#pragma unroll
      for (int q = 0; q < K; ++q)
        c[q] = lds_vec<W>(s_w + ring(k + dq[q] + 1) * RW + eo[q]);
      const Vec<W> o = site(self, iv[i], c, ip2[i]);
      lds_put<W>(s_m + ring(k) * RW + t * W, o);
      out[i] = o;
      if (inner && k >= 1 && k <= bx)
        add_products<IPM, W>(acc0, o, self, self);
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      if (inner && k >= 1 && k <= bx)
        stv<W, true>(o0 + gidx(slab_of(k), pos1), out[i]);
    }
    __syncthreads();
    // step 2: psi_{m+1} at slabs k0 - 1 + i on the interior entries
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int j = k0 - 1 + i;
      if (j >= 1 && j <= bx && inner) {
        const float vj = i == 0 ? cv_prev : iv[i - 1];
        const Vec<W> self = lds_vec<W>(s_m + ring(j) * RW + t * W);
        Vec<W> c[K];
#pragma unroll
        for (int q = 0; q < K; ++q)
          c[q] = lds_vec<W>(s_m + ring(j + dq[q]) * RW + eo[q] - o_mid);
        const Vec<W> prev = lds_vec<W>(s_w + ring(j + 1) * RW + o_self);  // psi_{m-1}
        const Vec<W> o = site(self, vj, c, prev);
        stv<W, true>(o1 + gidx(slab_of(j), pos1), o);
        add_products<IPM, W>(acc1, o, self, self);
      }
    }
    __syncthreads();
  }
  // the tile's sums (the rings are free after the loop's last barrier)
  auto s_red = reinterpret_cast<double(*)[kBlock / 64][64 * 4]>(s_dyn);
  store_products<IPM, W>(acc0, vpw, cols, r0, od.x, s_red, s0, s1);
  store_products<IPM, W>(acc1, vpw, cols, r0, od.x, s_red, s2, s3);
}

// NS steps per launch (NS >= 3) for the same moment sets: levels 1 .. NS
// (psi_m .. psi_{m+NS-1}; level 0 is psi_{m-1}), each one slab behind the
// level before it and kept in an LDS ring of B + 2 slab rows (levels
// 0 .. NS - 1), as k_sweepn does for single vectors; the potential of the
// last B + NS - 1 slabs is kept per thread in registers.  Only the last two
// levels are written.  Level l adds its two products to part[2 (l - 1)] and
// part[2 (l - 1) + 1] (moments 2 (m + l - 1) - 1 and 2 (m + l - 1) of the
// product identities).  A chunk holds the row's positions less 2 NS h.
// As in k_sweepn: each iteration's global loads are issued one iteration
// ahead and waited for at the end of the iteration before, ahead of its
// stores; the loop is unrolled by the period of the ring rows, so a
// neighbour's LDS offset is a per-thread constant of its slot and row.
template <int W, int NS, int B>
constexpr size_t block_sweepn_lds_bytes()
{
  return sizeof(float2) * NS * (B + 2) * kBlock * W;
}

// EVO: NS terms of the Chebyshev series of U(t) instead of moment products,
// with the per-site arithmetic of the one-term block kernel (k_block_evo):
// level l adds terms.bm / label[l - 1] times its value into the FP64 state
// blocks (cols double2 per site), each site's state held in registers from
// its slab's level 1 to level NS and written once; the additions of one
// term per launch, in the same order.
template <int K, int W, int IPM, int NS, int B, bool EVO = false>
__global__ __launch_bounds__(kBlock) void k_block_sweepn(
  HamDev h, BlockSweepMap sm, int cols, const float2* __restrict__ p2,
  const float2* __restrict__ p1, float2* __restrict__ o0, float2* __restrict__ o1,
  double2* __restrict__ part, size_t part_stride, double2* __restrict__ state = nullptr,
  SweepTerms terms = SweepTerms{})
{
  static_assert(W % 2 == 0 && NS >= 3, "W even, NS >= 3");
  static_assert(block_sweepn_lds_bytes<W, NS, B>() >= sizeof(double) * 2 * (kBlock / 64) * 64 * 4,
                "the tile sums reuse the rings");
  constexpr int RR = B + 2;        // rows of a level ring
  constexpr int RB = B + NS - 1;   // slabs of potential held per thread
  constexpr int RW = kBlock * W;   // ring row length in float2
  constexpr int RL = RR * RW;      // a level's ring in float2
  constexpr int U = RR / sweep_gcd(B, RR);  // iterations until the rows repeat
  static_assert(U >= 1 && U <= 3, "unrolled by one, two or three iterations");
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int4 od = sm.order[xcd_tile(blockIdx.x, sm.blocks)];
  if (od.x < 0)
    return;
  const BlockSweepSeg g = sm.seg[od.y];
  float2* s_lv = reinterpret_cast<float2*>(s_dyn);  // [level][row][entry][W]
  auto lv_at = [&](int l, int row) { return s_lv + ((size_t)l * RR + row) * RW; };

  const int t = threadIdx.x;
  const int vpw = cols / W;
  const int tpos = t / vpw;             // held entry (window position) of this thread
  const int r0 = (t - tpos * vpw) * W;  // its first column
  const int S = g.S, nslab = g.nslab, hh = g.hh;
  const int c0 = od.z * g.cs;
  const int clen = min(g.cs, S - c0);
  const int wlen = clen + 2 * NS * hh;
  int ws = c0 - NS * hh;
  while (ws < 0)
    ws += S;
  const int run = od.w;
  const int x_lo = (int)((long long)run * nslab / g.nruns);
  const int x_hi = (int)((long long)(run + 1) * nslab / g.nruns);
  const int bx = x_hi - x_lo;
  const int dir = (run & 1) ? -1 : 1;
  // sweep index q of slab start + dir q: the run's slabs are q = NS - 1 ..
  // NS - 2 + bx, level 0 is needed at q = -1 .. bx + 2 NS - 2
  const int start = dir > 0 ? x_lo - (NS - 1) : x_hi + NS - 2;
  const int q_last = bx + 2 * NS - 2;
  auto slab_of = [&](int q) {
    q = q < -1 ? -1 : (q > q_last ? q_last : q);
    int x = start + dir * q;
    while (x < 0)
      x += nslab;
    while (x >= nslab)
      x -= nslab;
    return x;
  };
  auto ring = [](int q) {  // q mod RR for q >= -8 RR
    return (q + 8 * RR) % RR;
  };
  auto wrap_pos = [&](int e) {
    const int p = ws + e;
    return p >= S ? p - S : p;
  };
  const int te = tpos < wlen ? tpos : 0;
  const int pos = wrap_pos(te);
  // psi_{m-2} and the potential are used on level 1's entries only, the
  // state (EVO) on the interior
  const int pos_p = wrap_pos(min(max(tpos, hh), max(wlen - hh - 1, 0)));
  const int pos_l = wrap_pos(min(max(tpos, NS * hh), NS * hh + clen - 1));
  const bool interior = tpos >= NS * hh && tpos < NS * hh + clen;
  const int o_self = te * vpw * W + r0;
  const int r = pos % g.p;
  int eo[K], dq[K];
  float ha[K], hb[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int4 e = sm.tab[g.tab + q * g.p + r];
    const bool ok = e.w != 0;
    int b = te + (ok ? e.y : 0);
    b = b < 0 ? b + S : (b >= S ? b - S : b);  // the whole slab in one chunk
    b = b < 0 ? 0 : (b < wlen ? b : wlen - 1);
    eo[q] = b * vpw * W + r0;
    dq[q] = ok ? e.x * dir : 0;
    const double2 hv = h.tab[ok ? e.z : 0];
    ha[q] = ok ? (float)hv.x : 0.0f;
    hb[q] = ok ? (float)hv.y : 0.0f;
  }
  const size_t base = (size_t)g.base;
  auto gidx = [&](int slab, int p) {
    return (base + (size_t)slab * S + p) * (size_t)cols + r0;
  };
  {
    const Vec<W> a0 = ldv<W, false>(p1 + gidx(slab_of(-1), pos));
    const Vec<W> a1 = ldv<W, false>(p1 + gidx(slab_of(0), pos));
    lds_put<W>(lv_at(0, ring(-1)) + t * W, a0);
    lds_put<W>(lv_at(0, ring(0)) + t * W, a1);
  }
  const float inv_scale = (float)(1.0 / h.emax);
  Products<W> acc[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l)
    clear_products(acc[l]);
  float pv[RB];  // potential (EVO: and state) of slabs k0 - NS + 1 + j
  double2 sv[EVO ? RB : 1][W];
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    pv[j] = 0.0f;
    if constexpr (EVO) {
#pragma unroll
      for (int c = 0; c < W; ++c)
        sv[j][c] = make_double2(0.0, 0.0);
    }
  }

  // neighbour s's entry (float2 offset in a level's ring) when the site's
  // slab sits in row c
  int nbo[K][RR];
#pragma unroll
  for (int q = 0; q < K; ++q) {
#pragma unroll
    for (int c = 0; c < RR; ++c)
      nbo[q][c] = (((c + dq[q]) % RR + RR) % RR) * RW + eo[q];
  }
  // the slabs of iteration k: level 0 at k + 1 .. k + B; psi_{m-2},
  // potential (EVO: state) at level 1's slabs k .. k + B - 1 (past the last
  // iteration: slabs already read)
  const int k_last = bx + 2 * NS - 3;
  Vec<W> nw[B], np2[B];
  float npv[B];
  double2 nsv[EVO ? B : 1][W];
  auto load_slabs = [&](int k) {
#pragma unroll
    for (int i = 0; i < B; ++i)
      nw[i] = ldv<W, false>(p1 + gidx(slab_of(k + 1 + i), pos));
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int sl = slab_of(min(k + i, k_last));
      const size_t at = base + (size_t)sl * S + pos_p;
      np2[i] = ldv<W, true>(p2 + at * cols + r0);
      npv[i] = h.pot32[at];
      if constexpr (EVO) {
        const size_t as = (base + (size_t)sl * S + pos_l) * cols + r0;
#pragma unroll
        for (int c = 0; c < W; ++c)
          nsv[i][c] = state[as + c];
      }
    }
  };
  auto hold_slabs = [&]() {  // the loads of load_slabs landed
#pragma unroll
    for (int i = 0; i < B; ++i) {
#pragma unroll
      for (int c = 0; c < W; ++c) {
        sweep_hold(nw[i].v[c]);
        sweep_hold(np2[i].v[c]);
        if constexpr (EVO)
          sweep_hold(nsv[i][c]);
      }
      sweep_hold(npv[i]);
    }
  };
  load_slabs(0);
  hold_slabs();

  // one iteration, k0 = c B modulo RR: the ring row of slab k0 + d is
  // row(d), a constant once the loops over levels and slabs are unrolled
  auto iterate = [&](auto phase, int k0) {
    constexpr int c = decltype(phase)::value;
    auto row = [](int d) { return ((c * B + d) % RR + RR) % RR; };
    Vec<W> w[B], ip2[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      w[i] = nw[i];
      ip2[i] = np2[i];
      pv[NS - 1 + i] = npv[i];
      if constexpr (EVO) {
#pragma unroll
        for (int cc = 0; cc < W; ++cc)
          sv[NS - 1 + i][cc] = nsv[i][cc];
      }
    }
    load_slabs(k0 + B);
#pragma unroll
    for (int i = 0; i < B; ++i)
      lds_put<W>(s_lv + row(1 + i) * RW + t * W, w[i]);
    __syncthreads();
    Vec<W> out[2][B];  // levels NS - 1 and NS, stored at the iteration's end
#pragma unroll
    for (int l = 1; l <= NS; ++l) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int q = k0 - (l - 1) + i;
        const int rq = row(i - (l - 1));
        const float2* lvl = s_lv + (l - 1) * RL;
        const Vec<W> self = lds_vec<W>(lvl + rq * RW + o_self);
        const float v = pv[NS - l + i];
        float tr[W], ti[W];
#pragma unroll
        for (int cc = 0; cc < W; ++cc) {
          tr[cc] = v * self.v[cc].x;
          ti[cc] = v * self.v[cc].y;
        }
#pragma unroll
        for (int s = 0; s < K; ++s) {
          const Vec<W> nb = lds_vec<W>(lvl + nbo[s][rq]);
#pragma unroll
          for (int cc = 0; cc < W; ++cc) {
            const float cx = nb.v[cc].x, cy = nb.v[cc].y;
            if constexpr (EVO) {  // k_block_evo's order
              tr[cc] += __builtin_fmaf(ha[s], cx, -hb[s] * cy);
              ti[cc] += __builtin_fmaf(ha[s], cy, hb[s] * cx);
            } else {
              tr[cc] = __builtin_fmaf(ha[s], cx, tr[cc]);
              tr[cc] = __builtin_fmaf(-hb[s], cy, tr[cc]);
              ti[cc] = __builtin_fmaf(ha[s], cy, ti[cc]);
              ti[cc] = __builtin_fmaf(hb[s], cx, ti[cc]);
            }
          }
        }
        const Vec<W> prev =
          l == 1 ? ip2[i] : lds_vec<W>(s_lv + (l - 2) * RL + rq * RW + o_self);
        Vec<W> o;
#pragma unroll
        for (int cc = 0; cc < W; ++cc) {
          o.v[cc].x = __builtin_fmaf(2.0f, tr[cc] * inv_scale, -prev.v[cc].x);
          o.v[cc].y = __builtin_fmaf(2.0f, ti[cc] * inv_scale, -prev.v[cc].y);
        }
        if (l >= NS - 1)
          out[l - (NS - 1)][i] = o;
        if (l < NS)
          lds_put<W>(s_lv + l * RL + rq * RW + t * W, o);
        if constexpr (EVO) {
#pragma unroll
          for (int cc = 0; cc < W; ++cc)
            accumulate_phase(sv[NS - l + i][cc], terms.bm[l - 1], (double)o.v[cc].x,
                             (double)o.v[cc].y, terms.label[l - 1]);
        } else if (interior && q >= NS - 1 && q <= NS - 2 + bx) {
          add_products<IPM, W>(acc[l - 1], o, self, self);
        }
      }
      __syncthreads();
    }
    hold_slabs();  // the next iteration's loads, before this iteration's stores
#pragma unroll
    for (int j = 0; j < 2; ++j) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int q = k0 - (NS - 2 + j) + i;  // level NS - 1 + j
        if (interior && q >= NS - 1 && q <= NS - 2 + bx) {
          stv<W, true>((j ? o1 : o0) + gidx(slab_of(q), pos), out[j][i]);
          if constexpr (EVO) {
            if (j) {  // every term of the launch added
              const size_t as = gidx(slab_of(q), pos);
#pragma unroll
              for (int cc = 0; cc < W; ++cc)
                state[as + cc] = sv[i][cc];
            }
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j + B < RB; ++j) {
      pv[j] = pv[j + B];
      if constexpr (EVO) {
#pragma unroll
        for (int cc = 0; cc < W; ++cc)
          sv[j][cc] = sv[j + B][cc];
      }
    }
  };

  for (int k0 = 0; k0 <= k_last; k0 += U * B) {
    iterate(std::integral_constant<int, 0>{}, k0);
    if constexpr (U > 1) {
      if (k0 + B > k_last)
        break;
      iterate(std::integral_constant<int, 1>{}, k0 + B);
    }
    if constexpr (U > 2) {
      if (k0 + 2 * B > k_last)
        break;
      iterate(std::integral_constant<int, 2>{}, k0 + 2 * B);
    }
  }
  if constexpr (EVO)
    return;
  auto s_red = reinterpret_cast<double(*)[kBlock / 64][64 * 4]>(s_dyn);
#pragma unroll
  for (int l = 0; l < NS; ++l)
    store_products<IPM, W>(acc[l], vpw, cols, r0, od.x, s_red, part + 2 * l * part_stride,
                           part + (2 * l + 1) * part_stride);
}

// Check of one system's slot table against its packed words: out[b] = 1
// where some word of workgroup b's sites differs from the table's.
__global__ __launch_bounds__(kBlock) void k_block_sweep_check(HamDev h, BlockSweepSeg g,
                                                              const int4* __restrict__ tab,
                                                              int* __restrict__ out)
{
  __shared__ int s_bad[kBlock / 64];
  const long long count = (long long)g.S * g.nslab;
  const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
  int bad = 0;
  if (i < count) {
    const int x = (int)(i / g.S), pos = (int)(i - (long long)x * g.S), o = pos % g.p;
    const int site = g.base + (int)i;
    for (int q = 0; q < h.K; ++q) {
      const uint32_t w = h.words[(size_t)q * h.stride + site];
      const int4 e = tab[g.tab + q * g.p + o];
      if (!e.w) {
        bad |= w != kEmptySlot;
        continue;
      }
      int xs = x + e.x;
      xs = xs < 0 ? xs + g.nslab : (xs >= g.nslab ? xs - g.nslab : xs);
      int ps = pos + e.y;
      ps = ps < 0 ? ps + g.S : (ps >= g.S ? ps - g.S : ps);
      const uint32_t want = ((uint32_t)e.z << h.idx_bits) | (uint32_t)(g.base + xs * g.S + ps);
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

}  // namespace lsqt
