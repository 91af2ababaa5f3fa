// kernels_sweepn.h -- NS Chebyshev steps per launch on periodic slab
// lattices (included by kernels_f64.hip and kernels_f32.hip after
// kernels_sweep.h, whose SweepIO it uses).
//
// The deeper form of k_sweep2: from psi_{m-2} and psi_{m-1} a launch makes
// psi_m .. psi_{m+NS-1} ("levels" 1 .. NS; level 0 is psi_{m-1}) and writes
// only the last two, the ones the recursion continues from, with the moment
// products <L|psi> of every level fused.  Per site and launch: psi_{m-2},
// psi_{m-1}, L and the potential read, two states written -- the bytes of one
// two-step launch for NS site-updates.
//
// A workgroup owns a chunk of positions of a run of slabs (as k_sweep2),
// one position per thread: level l is computed on the chunk and
// (NS - l) h positions either side (h the largest |position offset|; the
// halos are computed again by the neighbouring chunks' tiles), and it lags
// level l - 1 by one slab, since l at slab q needs l - 1 at q - 1 .. q + 1
// (and l - 2 at q for the recursion).  Each level 0 .. NS - 1 keeps the last
// B + 2 slabs it made in an LDS ring; the potential and left vector of the
// last B + NS - 1 slabs are kept per thread in registers (a level reaches
// slab q up to NS - 1 slabs after level 1 did), shifted by B slabs per
// iteration.  A loop iteration takes B new slabs of level 0 (and psi_{m-2},
// potential, left vector for level 1's B slabs) from registers loaded one
// iteration earlier, issues the loads of the next iteration's slabs, then
// computes each level's B slabs in turn, one barrier between levels: the
// loads are in flight while an iteration computes.  At the iteration's end
// the loads are waited for, then the two written levels go out through
// buffer stores that every lane issues (a lane with nothing to write gets an
// offset the hardware drops).  Loads and stores share one counter: a wait
// for a load issued before a store waits for the store too, so no load is
// waited for after the stores that follow it -- the stores have the next
// iteration to complete.
//
// The loop is unrolled by the period of the ring rows (RR / gcd(B, RR)
// iterations), so every level's ring row is a constant and a neighbour's LDS
// address is a per-thread offset of its slot and row, computed once.
//
// Per site the arithmetic is k_sweep2's (the one-step kernels', in the same
// order): the FP64 instantiation gives the bits of k_hpsi and of the
// reference at every level; only the moment partial sums group the sites by
// sweep tile.  A slot that an orbital does not use adds nothing: FULL (every
// orbital uses every slot) adds all of them, FAST (FP32, real hoppings) adds
// the unused ones' zero hoppings, the exact FP64 form keeps the sum of the
// used ones by a select.
//
// EVO: NS terms of the Chebyshev series of U(t) instead of moments: level l
// adds its Bessel-weighted value to the FP64 state S (terms.bm / label of
// level l: the phase cases of the one-term kernels; label 0 adds nothing),
// the state of a site held in registers from its slab's level 1 to its
// level NS and written once.  The additions are those of one term per
// launch, in the same order, so the state's bits are k_cheb2's.
#pragma once

#include <type_traits>

#include "kernels_sweep.h"

namespace lsqt {

__device__ __forceinline__ void sweep_add_term(double2& s, double bm, double tr, double ti,
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

// static LDS of k_sweepn<V, K, NS, B>: the level rings and the tile sums
// store of val at base[idx] through a buffer resource over elems elements
// (elems * sizeof(V) < 2^31): with ok false the offset lies past the range
// and the hardware drops the store
template <typename V>
__device__ __forceinline__ void store_or_drop(V* base, int elems, int idx, bool ok, V val)
{
  using u32x4 = unsigned __attribute__((ext_vector_type(4)));
  using u32x2 = unsigned __attribute__((ext_vector_type(2)));
  const __amdgpu_buffer_rsrc_t r =
    __builtin_amdgcn_make_buffer_rsrc(base, 0, elems * (int)sizeof(V), 0x00020000);
  const int off = ok ? idx * (int)sizeof(V) : (int)0x80000000u;
  if constexpr (sizeof(V) == 16)
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, val), r, off, 0, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, val), r, off, 0, 0);
}

template <typename V, int NS, int B>
constexpr size_t sweepn_lds_bytes()
{
  return sizeof(V) * NS * (B + 2) * kSweepThreads + sizeof(double) * (kSweepThreads / 64) * 2 * NS;
}

template <typename V, int K, int NS, int B, bool FAST, bool EVO = false, bool FULL = false>
__global__ __launch_bounds__(kSweepThreads) void k_sweepn(HamDev h, SweepMap sm,
                                                          const V* __restrict__ p2,
                                                          const V* __restrict__ p1,
                                                          V* __restrict__ o0, V* __restrict__ o1,
                                                          const V* __restrict__ L,
                                                          double2* __restrict__ part,
                                                          size_t part_stride,
                                                          double2* __restrict__ state = nullptr,
                                                          SweepTerms terms = SweepTerms{})
{
  static_assert(NS >= 3 && NS <= kSweepMaxSteps, "two steps per launch: k_sweep2");
  using IO = SweepIO<V>;
  constexpr int T = kSweepThreads;
  constexpr int RR = B + 2;       // rows of a level ring
  constexpr int RB = B + NS - 1;  // slabs of potential and left vector held
  constexpr int U = RR / sweep_gcd(B, RR);  // iterations until the rows repeat
  static_assert(U >= 1 && U <= 3, "unrolled by one, two or three iterations");
  constexpr int kRow = T * (int)sizeof(V);  // bytes of a ring row
  constexpr int kLevel = RR * kRow;         // bytes of a level's ring
  __shared__ V s_lv[NS][RR][T];             // levels 0 .. NS - 1
  __shared__ double s_red[T / 64][2 * NS];
  const int tile = xcd_tile(blockIdx.x, sm.blocks);
  if (tile >= sm.tiles)
    return;
  const int t = threadIdx.x;
  const int S = sm.S, nslab = sm.nslab, hh = sm.hh;
  const int chunk = tile % sm.nch;
  const int run = tile / sm.nch;
  const int c0 = chunk * sm.cs;
  const int clen = min(sm.cs, S - c0);
  const int wlen = clen + 2 * NS * hh;  // positions held per slab (<= T)
  int ws = c0 - NS * hh;                // first position held (periodic)
  while (ws < 0)
    ws += S;
  const int x_lo = (int)((long long)run * nslab / sm.nruns);
  const int x_hi = (int)((long long)(run + 1) * nslab / sm.nruns);
  const int bx = x_hi - x_lo;
  const int dir = (run & 1) ? -1 : 1;
  // sweep index q of slab start + dir q; the run's slabs are q = NS - 1 ..
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
  // this thread: window entry t (position pos), its neighbours' entries
  const int te = t < wlen ? t : 0;
  int pos = ws + te;
  pos = pos >= S ? pos - S : pos;
  const bool interior = t >= NS * hh && t < NS * hh + clen;
  const int r = pos % sm.p;
  int bq[K], dq[K];
  double ha[K], hb[K];
  bool okq[K];
#pragma unroll
  for (int q = 0; q < K; ++q) {
    const int4 e = sm.tab[q * sm.p + r];
    okq[q] = e.w != 0;
    dq[q] = e.x * dir;
    int b = te + e.y;
    b = b < 0 ? b + S : (b >= S ? b - S : b);  // the whole slab in one chunk
    bq[q] = b < 0 ? 0 : (b < wlen ? b : wlen - 1);
    const double2 hv = okq[q] ? h.tab[e.z] : make_double2(0.0, 0.0);
    ha[q] = hv.x;
    hb[q] = hv.y;
  }
  // byte offsets in a level's ring: entry te (the site), entry t (the one
  // written), slot s's neighbour when the site's slab sits in row c
  const int so = te * (int)sizeof(V), to = t * (int)sizeof(V);
  int nbo[K][RR];
#pragma unroll
  for (int q = 0; q < K; ++q) {
#pragma unroll
    for (int c = 0; c < RR; ++c)
      nbo[q][c] = (((c + dq[q]) % RR + RR) % RR) * kRow + bq[q] * (int)sizeof(V);
  }
  char* const lds = reinterpret_cast<char*>(&s_lv[0][0][0]);
  auto at = [&](int off) -> V& { return *reinterpret_cast<V*>(lds + off); };
  const double emax = h.emax;
  const double inv_emax = 1.0 / h.emax;
  auto gsite = [&](int q) { return (size_t)slab_of(q) * S + pos; };
  // psi_{m-2} and the potential are used on level 1's entries, the left
  // vector on the interior: the other threads load their nearest such
  // entry's (a cache line the wave reads anyway) instead of their own
  auto wrap_pos = [&](int e) {
    const int p = ws + e;
    return p >= S ? p - S : p;
  };
  const int pos_p = wrap_pos(min(max(t, hh), max(wlen - hh - 1, 0)));
  const int pos_l = wrap_pos(min(max(t, NS * hh), NS * hh + clen - 1));

  // FP64: this thread's own entries of the level rings also in registers
  // (same rows), read there as the site's value and psi_{m-2}; a thread past
  // the window (te = 0) holds its own values, which nothing reads
  constexpr bool kOwn = sizeof(V) == sizeof(double2);
  V own[kOwn ? NS : 1][kOwn ? RR : 1];
  // level 0 at q = -1, 0
  {
    const V a0 = p1[gsite(-1)], a1 = p1[gsite(0)];
    s_lv[0][ring(-1)][t] = a0;
    s_lv[0][ring(0)][t] = a1;
    if constexpr (kOwn) {
#pragma unroll
      for (int c = 0; c < RR; ++c)
        own[0][c] = c == ring(0) ? a1 : a0;
    }
  }
  double acc_r[NS], acc_i[NS];
#pragma unroll
  for (int l = 0; l < NS; ++l)
    acc_r[l] = acc_i[l] = 0.0;
  // potential and left vector (EVO: the state) of slabs k0 - NS + 1 + j,
  // j < RB
  double pv[RB];
  V lv[EVO ? 1 : RB];
  double2 sv[EVO ? RB : 1];
#pragma unroll
  for (int j = 0; j < RB; ++j) {
    pv[j] = 0.0;
    if constexpr (EVO)
      sv[j] = make_double2(0.0, 0.0);
    else
      lv[j] = IO::narrow(0.0, 0.0);
  }

  // the slabs of iteration k: level 0 at k + 1 .. k + B; psi_{m-2},
  // potential and left vector (EVO: state) at level 1's slabs k .. k + B - 1
  // (past the last iteration: slabs already read, so cache hits)
  const int k_last = bx + 2 * NS - 3;
  V nw[B], np2[B], nl[EVO ? 1 : B];
  double npv[B];
  double2 ns[EVO ? B : 1];
  auto load_slabs = [&](int k) {
#pragma unroll
    for (int i = 0; i < B; ++i)
      nw[i] = p1[gsite(k + 1 + i)];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      const size_t x = (size_t)slab_of(min(k + i, k_last)) * S;
      np2[i] = p2[x + pos_p];
      npv[i] = IO::pot(h, (int)(x + pos_p));
      if constexpr (EVO)
        ns[i] = state[x + pos_l];
      else
        nl[i] = L[x + pos_l];
    }
  };
  load_slabs(0);
  auto hold_slabs = [&]() {  // the loads of load_slabs landed
#pragma unroll
    for (int i = 0; i < B; ++i) {
      sweep_hold(nw[i]);
      sweep_hold(np2[i]);
      sweep_hold(npv[i]);
      if constexpr (EVO)
        sweep_hold(ns[i]);
      else
        sweep_hold(nl[i]);
    }
  };
  hold_slabs();

  // one iteration, k0 = c B modulo RR: the ring row of slab k0 + d is
  // row(d), a constant once the loops over levels and slabs are unrolled
  auto iterate = [&](auto phase, int k0) {
    constexpr int c = decltype(phase)::value;
    auto row = [](int d) { return ((c * B + d) % RR + RR) % RR; };
    V w[B], ip2[B];
#pragma unroll
    for (int i = 0; i < B; ++i) {
      w[i] = nw[i];
      ip2[i] = np2[i];
      pv[NS - 1 + i] = npv[i];
      if constexpr (EVO)
        sv[NS - 1 + i] = ns[i];
      else
        lv[NS - 1 + i] = nl[i];
    }
    load_slabs(k0 + B);
#pragma unroll
    for (int i = 0; i < B; ++i) {
      at(row(1 + i) * kRow + to) = w[i];
      if constexpr (kOwn)
        own[0][row(1 + i)] = w[i];
    }
    __syncthreads();
    V out[2][B];  // levels NS - 1 and NS, stored at the iteration's end
#pragma unroll
    for (int l = 1; l <= NS; ++l) {
#pragma unroll
      for (int i = 0; i < B; ++i) {
        const int q = k0 - (l - 1) + i;
        const int rq = row(i - (l - 1));
        const double2 self =
          IO::wide(kOwn ? own[kOwn ? l - 1 : 0][kOwn ? rq : 0] : at((l - 1) * kLevel + rq * kRow + so));
        const double v = pv[NS - l + i];  // slab q = k0 - NS + 1 + (NS - l + i)
        double tr = v * self.x;
        double ti = v * self.y;
        // every neighbour read first (a branch per slot would wait for each
        // read in turn)
        double2 cn[K];
#pragma unroll
        for (int s = 0; s < K; ++s)
          cn[s] = IO::wide(at((l - 1) * kLevel + nbo[s][rq]));
#pragma unroll
        for (int s = 0; s < K; ++s) {
          if constexpr (FAST) {  // unused slots: zero hoppings
            tr += ha[s] * cn[s].x;
            ti += ha[s] * cn[s].y;
          } else if constexpr (FULL) {
            tr += ha[s] * cn[s].x - hb[s] * cn[s].y;
            ti += ha[s] * cn[s].y + hb[s] * cn[s].x;
          } else {  // the sum of the used slots, in order
            const double nr = tr + (ha[s] * cn[s].x - hb[s] * cn[s].y);
            const double ni = ti + (ha[s] * cn[s].y + hb[s] * cn[s].x);
            tr = okq[s] ? nr : tr;
            ti = okq[s] ? ni : ti;
          }
        }
        if constexpr (FAST) {
          tr *= inv_emax;
          ti *= inv_emax;
        } else {
          tr /= emax;
          ti /= emax;
        }
        const double2 prev =
          l == 1 ? IO::wide(ip2[i])
                 : IO::wide(kOwn ? own[kOwn ? l - 2 : 0][kOwn ? rq : 0]
                                 : at((l - 2) * kLevel + rq * kRow + so));
        tr = 2.0 * tr - prev.x;
        ti = 2.0 * ti - prev.y;
        const V o = IO::narrow(tr, ti);
        if (l >= NS - 1)
          out[l - (NS - 1)][i] = o;
        if (l < NS) {
          at(l * kLevel + rq * kRow + to) = o;
          if constexpr (kOwn)
            own[l][rq] = o;
        }
        if constexpr (EVO) {
          // the term of this level in the FP64 state, from the value stored
          const double2 ow = IO::wide(o);
          sweep_add_term(sv[NS - l + i], terms.bm[l - 1], ow.x, ow.y, terms.label[l - 1]);
        } else if (interior && q >= NS - 1 && q <= NS - 2 + bx) {
          const double2 ow = IO::wide(o), lw = IO::wide(lv[NS - l + i]);
          acc_r[l - 1] += ow.x * lw.x + ow.y * lw.y;
          acc_i[l - 1] += ow.y * lw.x - ow.x * lw.y;
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
        const bool ok = interior && q >= NS - 1 && q <= NS - 2 + bx;
        const size_t x = (size_t)slab_of(q) * S;
        store_or_drop(j ? o1 + x : o0 + x, S, pos, ok, out[j][i]);
        if constexpr (EVO) {
          if (j)  // every term of the launch added
            store_or_drop(state + x, S, pos, ok, sv[i]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j + B < RB; ++j) {
      pv[j] = pv[j + B];
      if constexpr (EVO)
        sv[j] = sv[j + B];
      else
        lv[j] = lv[j + B];
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

  // the tile's NS moment partials: wave64 shuffle tree, then the waves in order
#pragma unroll
  for (int l = 0; l < NS; ++l) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      acc_r[l] += __shfl_xor(acc_r[l], o, 64);
      acc_i[l] += __shfl_xor(acc_i[l], o, 64);
    }
  }
  if ((t & 63) == 0) {
#pragma unroll
    for (int l = 0; l < NS; ++l) {
      s_red[t >> 6][2 * l] = acc_r[l];
      s_red[t >> 6][2 * l + 1] = acc_i[l];
    }
  }
  __syncthreads();
  if (t < NS && part) {
    double re = s_red[0][2 * t], im = s_red[0][2 * t + 1];
#pragma unroll
    for (int wv = 1; wv < T / 64; ++wv) {
      re += s_red[wv][2 * t];
      im += s_red[wv][2 * t + 1];
    }
    part[(size_t)t * part_stride + tile] = make_double2(re, im);
  }
}

// NS terms of U(t) per launch (EVO): from p2 = phi_{k-2}, p1 = phi_{k-1},
// o0 = phi_{k+NS-2}, o1 = phi_{k+NS-1}, level l's term added into S with
// terms.bm[l - 1] / label[l - 1]
template <typename V, int NS, int B>
void sweepn_evo_launch(const HamDev& h, const SweepMap& sm, const V* p2, const V* p1, V* o0,
                       V* o1, double2* S, const SweepTerms& terms, hipStream_t st)
{
  const bool fast = sizeof(V) == sizeof(float2) && sm.real;
#define LSQT_SWNE(KK)                                                                         \
  do {                                                                                        \
    if (fast)                                                                                 \
      k_sweepn<V, KK, NS, B, sizeof(V) == sizeof(float2), true>                               \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, p1, nullptr, 0, S, terms); \
    else if (sm.full)                                                                         \
      k_sweepn<V, KK, NS, B, false, true, true>                                               \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, p1, nullptr, 0, S, terms); \
    else                                                                                      \
      k_sweepn<V, KK, NS, B, false, true>                                                     \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, p1, nullptr, 0, S, terms); \
  } while (0)
  switch (h.K) {
    case 2:
      LSQT_SWNE(2);
      break;
    case 3:
      LSQT_SWNE(3);
      break;
    case 4:
      LSQT_SWNE(4);
      break;
    case 6:
      LSQT_SWNE(6);
      break;
    default:
      break;
  }
#undef LSQT_SWNE
}

// Host launcher of k_sweepn: NS steps per launch in blocks of B slabs for
// slot counts 2, 3, 4 and 6 (part: NS rows of partials part_stride apart)
template <typename V, int NS, int B>
void sweepn_launch(const HamDev& h, const SweepMap& sm, const V* p2, const V* p1, V* o0, V* o1,
                   const V* L, double2* part, size_t part_stride, hipStream_t st)
{
  const bool fast = sizeof(V) == sizeof(float2) && sm.real;
#define LSQT_SWN(KK)                                                                          \
  do {                                                                                        \
    if (fast)                                                                                 \
      k_sweepn<V, KK, NS, B, sizeof(V) == sizeof(float2)>                                     \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, L, part, part_stride);   \
    else if (sm.full)                                                                         \
      k_sweepn<V, KK, NS, B, false, false, true>                                              \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, L, part, part_stride);   \
    else                                                                                      \
      k_sweepn<V, KK, NS, B, false>                                                           \
        <<<sm.blocks, kSweepThreads, 0, st>>>(h, sm, p2, p1, o0, o1, L, part, part_stride);   \
  } while (0)
  switch (h.K) {
    case 2:
      LSQT_SWN(2);
      break;
    case 3:
      LSQT_SWN(3);
      break;
    case 4:
      LSQT_SWN(4);
      break;
    case 6:
      LSQT_SWN(6);
      break;
    default:
      break;
  }
#undef LSQT_SWN
}

}  // namespace lsqt
