// kernels_sweep.h -- two Chebyshev steps per launch on periodic slab lattices.
//
// A lattice numbered x-slowest (src/model_lattice.cu:148-162) comes in slabs
// of S consecutive sites (one x cell).  On a 2D lattice with periodic
// boundaries every neighbour of a site sits at a fixed slab offset d (-1, 0
// or 1) and position offset dp within the slab that depend only on the
// site's orbital (its index modulo the pattern period p, which divides S),
// both taken periodically -- the wrapped neighbours at the lattice's edges
// included.  The Hamiltonian's creation finds that per-orbital table from
// the packed words and verifies every word of every site against it on the
// device; where it holds, the recursion psi_m = 2 H~ psi_{m-1} - psi_{m-2}
// advances two steps per launch: a workgroup owns a chunk of positions of a
// run of slabs and walks the run slab by slab, holding in LDS
//   - B + 2 slabs of psi_{m-1} (the chunk plus 2h positions either side, h
//     the largest |dp|), and
//   - B + 2 slabs of psi_m it computed (the chunk plus h either side: the
//     halo is computed again by the neighbouring chunks' tiles),
// and advances B slabs at a time: psi_m at slabs k .. k + B - 1, then
// psi_{m+1} at slabs k - 1 .. k + B - 2, all from LDS.
// psi_{m-1} and psi_{m-2} are read from HBM once, psi_m and psi_{m+1} written
// once, the left vector read once for both moment products: 5 vector streams
// per two site-updates instead of 8.  Runs alternate their direction, so the
// two tiles that share a run boundary read its halo slabs at the same time
// (both at their start or both at their end).
//
// Each thread keeps one position of its chunk, so its orbital, and with it
// its neighbours' LDS offsets and hopping values, are fixed for the whole
// sweep: no slot words are read.  A block's global loads (B new slabs of
// psi_{m-1}, B slabs of psi_{m-2}, potential and left vector) are all issued
// at its start and used within it, so no load is in flight across a loop
// iteration (whose register copies would make the compiler wait for every
// load); the memory latency is met once per B slabs and hidden by the other
// workgroups on the CU.
//
// Per site, the arithmetic is that of the one-step kernels in the same order
// (on-site term, neighbours in slot order, division by energy_max, the
// recursion), in FP64: the FP64 instantiation (kernels_f64.hip, compiled
// without contraction) gives the bits of k_hpsi and of the reference; the
// FP32-storage one (kernels_f32.hip) the values of k_hpsi32.  Only the moment
// partial sums group the sites differently (one partial per tile and moment,
// summed in tile order), so the moments agree with the one-step kernels'
// to rounding of the sums.
#pragma once

#include "lsqt_device.h"

namespace lsqt {

template <typename V>
struct SweepIO;

template <>
struct SweepIO<double2> {
  static __device__ __forceinline__ double pot(const HamDev& h, int i) { return h.pot[i]; }
  static __device__ __forceinline__ double2 wide(double2 v) { return v; }
  static __device__ __forceinline__ double2 narrow(double x, double y)
  {
    return make_double2(x, y);
  }
};

template <>
struct SweepIO<float2> {
  static __device__ __forceinline__ double pot(const HamDev& h, int i)
  {
    return (double)h.pot32[i];
  }
  static __device__ __forceinline__ double2 wide(float2 v)
  {
    return make_double2((double)v.x, (double)v.y);
  }
  static __device__ __forceinline__ float2 narrow(double x, double y)
  {
    return make_float2((float)x, (float)y);
  }
};

// Dynamic LDS of k_sweep2 with blocks of B slabs: psi_{m-1} and psi_m in
// B + 2 rows of kSweepThreads entries each (the held entries, wlen <=
// kSweepThreads, then padding), then each thread's potential and left
// vector entry of the last slab of a block (used by the next block).
template <typename V, int B>
inline size_t sweep_lds_bytes(const SweepMap&)
{
  return (sizeof(V) * 2 * (B + 2) + sizeof(double) + sizeof(V)) * kSweepThreads;
}

// FAST (FP32 storage only, where no bit pattern is owed): the division by
// energy_max becomes a multiplication by its FP64 reciprocal (one FP64
// rounding more), and real hoppings (sm.real) skip their imaginary terms.
template <typename V, int K, int B, bool FAST>
__global__ __launch_bounds__(kSweepThreads) void k_sweep2(HamDev h, SweepMap sm,
                                                          const V* __restrict__ p2,
                                                          const V* __restrict__ p1,
                                                          V* __restrict__ o0, V* __restrict__ o1,
                                                          const V* __restrict__ L,
                                                          double2* __restrict__ part0,
                                                          double2* __restrict__ part1)
{
  using IO = SweepIO<V>;
  constexpr int R = B + 2;  // ring rows
  extern __shared__ __align__(16) unsigned char s_dyn[];
  const int tile = xcd_tile(blockIdx.x, sm.blocks);
  if (tile >= sm.tiles)
    return;
  const int t = threadIdx.x;
  constexpr int W = kSweepThreads;        // LDS row length
  V* s_w = reinterpret_cast<V*>(s_dyn);  // psi_{m-1}: sweep index q in ring row (q + 1) mod R
  V* s_m = s_w + R * W;                  // psi_m: sweep index q in ring row q mod R
  // potential and left vector of the previous block's last slab, per thread
  // (kept in LDS: a register carried across the loop would be copied, and
  // the copy would wait for the block's stores)
  V* s_cl = s_m + R * W;
  double* s_cv = reinterpret_cast<double*>(s_cl + kSweepThreads);

  // tile geometry
  const int S = sm.S, nslab = sm.nslab, hh = sm.hh;
  // chunks fastest: the tiles in flight together cover whole slabs (measured
  // on the 8.0M-site honeycomb: 73.7 against 75.3 us per FP64 step with runs
  // fastest)
  const int chunk = tile % sm.nch;
  const int run = tile / sm.nch;
  const int c0 = chunk * sm.cs;
  const int clen = min(sm.cs, S - c0);
  const int wlen = clen + 4 * hh;  // psi_{m-1} positions held per slab (<= kSweepThreads)
  const int len1 = clen + 2 * hh;  // psi_m positions computed per slab
  int ws = c0 - 2 * hh;            // first position held (periodic)
  if (ws < 0)
    ws += S;
  const int x_lo = (int)((long long)run * nslab / sm.nruns);
  const int x_hi = (int)((long long)(run + 1) * nslab / sm.nruns);
  const int bx = x_hi - x_lo;
  const int dir = (run & 1) ? -1 : 1;
  const int start = dir > 0 ? x_lo - 1 : x_hi;  // slab of sweep index 0
  // slab of sweep index q, clamped to -1 .. bx + 2 (loads past the run's end
  // read its last slab again; their values are not used)
  auto slab_of = [&](int q) {
    q = q > bx + 2 ? bx + 2 : q;
    int x = start + dir * q;
    x = x < 0 ? x + nslab : x;
    return x >= nslab ? x - nslab : x;
  };
  auto wpos = [&](int w) {  // position of held entry w
    const int p = ws + w;
    return p >= S ? p - S : p;
  };
  auto ring = [](int q) {  // q mod R for q >= -R
    return (q + R) % R;
  };

  // this thread: held entry t of a psi_{m-1} slab, psi_m entry t (step 1,
  // t < len1; held entry t + hh), the interior entry t (step 2)
  const bool act1 = t < len1;
  const bool inner = t >= hh && t < hh + clen;
  const int pos1 = wpos(act1 ? t + hh : hh);
  const int pos_a = wpos(t < wlen ? t : 0);
  // its slots: held-entry offset of the neighbour, slab offset along the
  // sweep, hopping value (zero for unused slots)
  const int r = pos1 % sm.p;
  int bq[K], dq[K];
  double ha[K], hb[K];
  bool okq[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int4 e = sm.tab[q * sm.p + r];
    okq[q] = e.w != 0;
    dq[q] = e.x * dir;
    int b = t + hh + e.y;
    b = b < 0 ? b + S : (b >= S ? b - S : b);  // one chunk of the whole slab
    bq[q] = b < 0 ? 0 : (b < wlen ? b : wlen - 1);
    const double2 hv = okq[q] ? h.tab[e.z] : make_double2(0.0, 0.0);
    ha[q] = hv.x;
    hb[q] = hv.y;
  }
  auto fetch = [&](int q) { return p1[(size_t)slab_of(q) * S + pos_a]; };
  auto put = [&](int q, const V& a) {  // entries past wlen are padding
    s_w[ring(q + 1) * W + t] = a;
  };
  {
    const V a0 = fetch(-1), a1 = fetch(0);
    put(-1, a0);
    put(0, a1);
  }

  const double emax = h.emax;
  const double inv_emax = 1.0 / h.emax;
  auto site_sum = [&](double& tr, double& ti, int q, double2 c) {
    if constexpr (FAST) {
      tr += ha[q] * c.x;
      ti += ha[q] * c.y;
    } else {
      tr += ha[q] * c.x - hb[q] * c.y;
      ti += ha[q] * c.y + hb[q] * c.x;
    }
  };
  auto scale = [&](double& tr, double& ti) {
    if constexpr (FAST) {
      tr *= inv_emax;
      ti *= inv_emax;
    } else {
      tr /= emax;
      ti /= emax;
    }
  };
  double a0r = 0.0, a0i = 0.0, a1r = 0.0, a1i = 0.0;
  // block of step-1 slabs k0 .. k0 + B - 1 and step-2 slabs k0 - 1 .. k0 + B - 2:
  // every load is issued at the block's start and used inside the block
  for (int k0 = 0; k0 <= bx + 1; k0 += B) {
    V w[B], ip2[B], il[B];
    double iv[B];
#pragma unroll
    for (int i = 0; i < B; ++i)
      w[i] = fetch(k0 + 1 + i);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      const int site = slab_of(k > bx + 1 ? bx + 1 : k) * S + pos1;
      ip2[i] = p2[site];
      iv[i] = IO::pot(h, site);
      il[i] = L[site];
    }
    // every load of the block is used here, before any store is issued:
    // the previous block's carried entries out of LDS, this block's in
    const double cv_prev = s_cv[t];
    const V cl_prev = s_cl[t];
    s_cv[t] = iv[B - 1];
    s_cl[t] = il[B - 1];
#pragma unroll
    for (int i = 0; i < B; ++i)
      put(k0 + 1 + i, w[i]);
    __syncthreads();
    // step 1: psi_m at slabs k0 + i on the held positions hh .. hh + len1 - 1
    // (stored after all B are computed, so that no store is issued before
    // the last load is used)
    V out[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      const double2 self = IO::wide(s_w[ring(k + 1) * W + (act1 ? t + hh : 0)]);
      double tr = iv[i] * self.x;
      double ti = iv[i] * self.y;
#pragma unroll
      for (int q = 0; q < K; ++q) {
        const double2 c = IO::wide(s_w[ring(k + dq[q] + 1) * W + bq[q]]);
        if (okq[q])
          site_sum(tr, ti, q, c);
      }
      scale(tr, ti);
      const double2 q2 = IO::wide(ip2[i]);
      tr = 2.0 * tr - q2.x;
      ti = 2.0 * ti - q2.y;
      const V o = IO::narrow(tr, ti);
      s_m[ring(k) * W + t] = o;  // entries past len1 are padding
      out[i] = o;
      if (inner && k >= 1 && k <= bx) {
        const double2 ow = IO::wide(o), lw = IO::wide(il[i]);
        a0r += ow.x * lw.x + ow.y * lw.y;
        a0i += ow.y * lw.x - ow.x * lw.y;
      }
    }
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int k = k0 + i;
      if (inner && k >= 1 && k <= bx)
        o0[(size_t)slab_of(k) * S + pos1] = out[i];
    }
    __syncthreads();
    // step 2: psi_{m+1} at slabs k0 - 1 + i on the interior positions
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const int j = k0 - 1 + i;
      if (j >= 1 && j <= bx && inner) {
        const double vj = i == 0 ? cv_prev : iv[i - 1];
        const V lj = i == 0 ? cl_prev : il[i - 1];
        const double2 self = IO::wide(s_m[ring(j) * W + t]);
        double tr = vj * self.x;
        double ti = vj * self.y;
#pragma unroll
        for (int q = 0; q < K; ++q) {
          const double2 c = IO::wide(s_m[ring(j + dq[q]) * W + bq[q] - hh]);
          if (okq[q])
            site_sum(tr, ti, q, c);
        }
        scale(tr, ti);
        const double2 q1 = IO::wide(s_w[ring(j + 1) * W + t + hh]);  // psi_{m-1} at slab j
        tr = 2.0 * tr - q1.x;
        ti = 2.0 * ti - q1.y;
        const V o = IO::narrow(tr, ti);
        o1[(size_t)slab_of(j) * S + pos1] = o;
        const double2 ow = IO::wide(o), lw = IO::wide(lj);
        a1r += ow.x * lw.x + ow.y * lw.y;
        a1i += ow.y * lw.x - ow.x * lw.y;
      }
    }
    __syncthreads();
  }

  // the tile's two moment partials: wave64 shuffle tree, then the waves in order
  __shared__ double s_red[kSweepThreads / 64][4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a0r += __shfl_xor(a0r, o, 64);
    a0i += __shfl_xor(a0i, o, 64);
    a1r += __shfl_xor(a1r, o, 64);
    a1i += __shfl_xor(a1i, o, 64);
  }
  if ((t & 63) == 0) {
    s_red[t >> 6][0] = a0r;
    s_red[t >> 6][1] = a0i;
    s_red[t >> 6][2] = a1r;
    s_red[t >> 6][3] = a1i;
  }
  __syncthreads();
  if (t == 0 && part0) {
    double r0 = s_red[0][0], i0 = s_red[0][1], r1 = s_red[0][2], i1 = s_red[0][3];
#pragma unroll
    for (int wv = 1; wv < kSweepThreads / 64; ++wv) {
      r0 += s_red[wv][0];
      i0 += s_red[wv][1];
      r1 += s_red[wv][2];
      i1 += s_red[wv][3];
    }
    part0[tile] = make_double2(r0, i0);
    part1[tile] = make_double2(r1, i1);
  }
}

// Host launcher of k_sweep2 for storage type V and compile-time slot counts
// 2, 3, 4 and 6 (the shapes that detect_sweep accepts), in blocks of
// kSweepBlock slabs.  L may be null (no moment products: psi_{m-1} stands in
// for the left vector, and part0 / part1 must be null too).
template <typename V, int K, int B>
void sweep_launch_b(const HamDev& h, const SweepMap& sm, const V* p2, const V* p1, V* o0, V* o1,
                    const V* L, double2* part0, double2* part1, hipStream_t st)
{
  const size_t lds = sweep_lds_bytes<V, B>(sm);
  if (sizeof(V) == sizeof(float2) && sm.real)
    k_sweep2<V, K, B, sizeof(V) == sizeof(float2)>
      <<<sm.blocks, kSweepThreads, lds, st>>>(h, sm, p2, p1, o0, o1, L, part0, part1);
  else
    k_sweep2<V, K, B, false>
      <<<sm.blocks, kSweepThreads, lds, st>>>(h, sm, p2, p1, o0, o1, L, part0, part1);
}

template <typename V>
void sweep_launch(const HamDev& h, const SweepMap& sm, const V* p2, const V* p1, V* o0, V* o1,
                  const V* L, double2* part0, double2* part1, hipStream_t st)
{
  const V* left = L ? L : p1;
  switch (h.K) {
    case 2:
      sweep_launch_b<V, 2, kSweepBlock>(h, sm, p2, p1, o0, o1, left, part0, part1, st);
      break;
    case 3:
      sweep_launch_b<V, 3, kSweepBlock>(h, sm, p2, p1, o0, o1, left, part0, part1, st);
      break;
    case 4:
      sweep_launch_b<V, 4, kSweepBlock>(h, sm, p2, p1, o0, o1, left, part0, part1, st);
      break;
    case 6:
      sweep_launch_b<V, 6, kSweepBlock>(h, sm, p2, p1, o0, o1, left, part0, part1, st);
      break;
    default:
      break;
  }
}

}  // namespace lsqt
