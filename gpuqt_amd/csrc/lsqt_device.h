// lsqt_device.h -- internal types shared by the device sources of
// liblsqt_hip.so.  Not part of the ABI (that is include/lsqt_hip.h).
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cstddef>
#include <cstdint>

namespace lsqt {

constexpr int kBlock = 256;          // threads per workgroup (4 wave64)
constexpr int kXcds = 8;             // MI355X: 8 XCDs, workgroups dispatched round-robin
constexpr uint32_t kEmptySlot = 0xFFFFFFFFu;  // packed word of an unused ELL slot
constexpr int kLdsTableMax = 512;    // packed tables up to this size are staged in LDS

enum Layout { kFull = 0, kPackedLds = 1, kPackedGlobal = 2 };

// Slot patterns of packed words (periodic lattices): the longest period and
// the most entries (period x slots) a pattern may have; the entry of a slot
// that the pattern leaves unused
constexpr int kPatternMaxPeriod = 64;
constexpr int kPatternMaxEntries = 512;
constexpr int kPatternEmpty = INT_MIN;

// Division of 0 <= v < 2^31 by an invariant divisor 2 <= d < 2^31 with one
// high multiply and a shift (Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994, section 4): with
// l = ceil(log2 d) and m = ceil(2^(31+l) / d) < 2^32, v / d equals
// umulhi(v, m) >> (l - 1), since m d - 2^(31+l) < d keeps the error of
// v m / 2^(31+l) below 1/d.
struct Divisor {
  uint32_t m;
  int shift;
};

inline Divisor make_divisor(int d)
{
  int l = 0;
  while ((1ll << l) < d)
    ++l;
  const unsigned long long p = 1ull << (31 + l);
  return Divisor{(uint32_t)((p + (unsigned long long)d - 1) / (unsigned long long)d), l - 1};
}

__device__ __forceinline__ int div_by(int v, const Divisor& q)
{
  return (int)(__umulhi((uint32_t)v, q.m) >> q.shift);
}

// Device view of a Hamiltonian.  ELL slots are column-major: slot m of site
// n is at m*n_sites + n.  Unused slots are marked (kEmptySlot / index -1)
// and always come after the used ones, as in the reference's row-major
// arrays (src/model_lattice.cu:356-383).
struct HamDev {
  int n;
  int K;
  int stride;  // column stride of the pair arrays (n rounded up to 64)
  double emax;
  const double* pot;
  // packed layout
  const uint32_t* words;
  uint32_t idx_mask;
  int idx_bits;
  const double2* tab;     // (re, im) per class
  const double* tab_xx;   // xx per class
  int ncls;
  // full layout
  const int* nl;
  const double* hr;
  const double* hi;
  const double* xx;
  int layout;
  // FP32 copy of the potential (the precision path; null until first used)
  const float* pot32;
  // periodic slot pattern of the packed words (pat null: none).  Site n
  // whose bit in `irregular` (32 sites per word) is clear holds in slot m
  // exactly pattern_word(pat[m * pat_p + n mod pat_p], n); the moment
  // kernels compute those words instead of loading them
  const int2* pat;
  const uint32_t* irregular;
  int pat_p;
  Divisor pat_div;  // for pat_p >= 2
};

// the packed word of a slot that follows pattern entry e at site n
__device__ __forceinline__ uint32_t pattern_word(const HamDev& h, int2 e, int n)
{
  return e.x == kPatternEmpty ? kEmptySlot
                              : (((uint32_t)e.y << h.idx_bits) | (uint32_t)(n + e.x));
}

__device__ __forceinline__ int pattern_residue(const HamDev& h, int n)
{
  return h.pat_p == 1 ? 0 : n - div_by(n, h.pat_div) * h.pat_p;
}

// whether site n's words differ from the pattern (they are then loaded)
__device__ __forceinline__ bool pattern_irregular(const HamDev& h, int n)
{
  return (h.irregular[n >> 5] >> (n & 31)) & 1u;
}

constexpr int kPairStrideAlign = 64;

inline int pair_stride(int n)
{
  return (n + kPairStrideAlign - 1) / kPairStrideAlign * kPairStrideAlign;
}

// Streaming policy: arrays read once per step (potential, pair words,
// psi_{m-2}, the left vector) are loaded non-temporally on large systems, so
// that the L2 keeps the gathered psi_{m-1}; stores stay normal (measured on
// MI355X: non-temporal stores cost ~15% on the fused step, non-temporal
// loads gain ~18%).  Below this many sites everything fits in the caches and
// normal loads are used.
constexpr int kStreamingMinSites = 1 << 20;

typedef double dvec2_t __attribute__((ext_vector_type(2)));
typedef float fvec2_t __attribute__((ext_vector_type(2)));
typedef float fvec4_t __attribute__((ext_vector_type(4)));
typedef uint32_t uvec2_t __attribute__((ext_vector_type(2)));

template <bool NT, typename T>
__device__ __forceinline__ T ld(const T* p)
{
  if constexpr (NT)
    return __builtin_nontemporal_load(p);
  else
    return *p;
}

template <bool NT>
__device__ __forceinline__ double2 ld2(const double2* p)
{
  if constexpr (NT) {
    dvec2_t v = __builtin_nontemporal_load(reinterpret_cast<const dvec2_t*>(p));
    return make_double2(v.x, v.y);
  } else {
    return *p;
  }
}

template <bool NT>
__device__ __forceinline__ void st2(double2* p, double2 v)
{
  if constexpr (NT) {
    dvec2_t w;
    w.x = v.x;
    w.y = v.y;
    __builtin_nontemporal_store(w, reinterpret_cast<dvec2_t*>(p));
  } else {
    *p = v;
  }
}

// Loads slots [m0, m0 + C) of site n, every load issued before any is used
// (independent loads overlap their latencies; a slot-by-slot loop serialises
// three dependent round trips per slot).  Slots at or beyond `count` and
// unused slots get j = n (an in-bounds address) and ok = false; callers skip
// their terms.  The pair arrays are padded to `stride`, so every slot load
// below `count` is in bounds.
//
// PAT (packed layouts with a slot pattern, s_pat = the pattern table): the
// words of a regular site are computed from the pattern and only those of
// irregular sites are loaded; both give the same words.
template <int LAYOUT, int C, bool NT, bool WITH_XX, bool PAT = false>
__device__ __forceinline__ void load_slots(const HamDev& h, const double2* s_tab,
                                           const double* s_xx, int n, int m0, int count,
                                           int (&j)[C], double (&a)[C], double (&b)[C],
                                           double (&x)[C], bool (&ok)[C],
                                           const int2* s_pat = nullptr)
{
  if constexpr (LAYOUT == 0) {  // full layout: index, re, im (and xx) per slot
    int jj[C];
#pragma unroll
    for (int q = 0; q < C; ++q)
      jj[q] = (m0 + q < count) ? ld<NT>(h.nl + (size_t)(m0 + q) * h.stride + n) : -1;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const bool in = m0 + q < count;
      a[q] = in ? ld<NT>(h.hr + (size_t)(m0 + q) * h.stride + n) : 0.0;
      b[q] = in ? ld<NT>(h.hi + (size_t)(m0 + q) * h.stride + n) : 0.0;
      if constexpr (WITH_XX)
        x[q] = in ? ld<NT>(h.xx + (size_t)(m0 + q) * h.stride + n) : 0.0;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      ok[q] = jj[q] >= 0;
      j[q] = ok[q] ? jj[q] : n;
    }
  } else {  // packed: one word per slot, (re, im, xx) from the class table
    uint32_t w[C];
    if constexpr (PAT) {
      const int r = pattern_residue(h, n);
#pragma unroll
      for (int q = 0; q < C; ++q)
        w[q] = (m0 + q < count) ? pattern_word(h, s_pat[(m0 + q) * h.pat_p + r], n) : kEmptySlot;
      if (pattern_irregular(h, n)) {
#pragma unroll
        for (int q = 0; q < C; ++q)
          if (m0 + q < count)
            w[q] = ld<NT>(h.words + (size_t)(m0 + q) * h.stride + n);
      }
    } else {
#pragma unroll
      for (int q = 0; q < C; ++q)
        w[q] = (m0 + q < count) ? ld<NT>(h.words + (size_t)(m0 + q) * h.stride + n) : kEmptySlot;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      ok[q] = w[q] != kEmptySlot;
      j[q] = ok[q] ? (int)(w[q] & h.idx_mask) : n;
      const int c = ok[q] ? (int)(w[q] >> h.idx_bits) : 0;
      const double2 t = (LAYOUT == 1) ? s_tab[c] : h.tab[c];
      a[q] = t.x;
      b[q] = t.y;
      if constexpr (WITH_XX)
        x[q] = (LAYOUT == 1) ? s_xx[c] : h.tab_xx[c];
    }
  }
}

// Slots per load batch: all K slots when K is a compile-time constant,
// otherwise batches of 4 (in slot order).
template <int KT>
struct SlotBatch {
  static constexpr int C = KT > 0 ? KT : 4;
};

// Grid shape of a one-tile-per-workgroup launch: `tiles` tiles of
// kBlock*spt sites, padded to a multiple of the XCD count so that the
// round-robin dispatch gives every XCD one contiguous range of tiles.
struct Grid {
  int tiles;
  int blocks;
};

// Block-path tile schedule: natural (tile t = sites [t*rows, (t+1)*rows))
// or blocked (slab > 0: tile t = the span of sites at offset (t / slabs)*span
// in slab t % slabs); walk != 0 sweeps a blocked tile made of whole rows of
// `row` sites chunk by chunk across its rows (the column walk).  With a tile
// table (segmented Hamiltonians: independent systems side by side) tile t is
// sites [table[t].x, table[t].y) of segment tile_seg[t], and the tiles of
// segment k are seg_tiles[k] .. seg_tiles[k+1]-1.
// Sliding-window geometry of a lattice (sites x * slab + y * row + z): a tile
// is `by` consecutive rows y of one slab x, restricted to one chunk of
// `chunk` consecutive z, swept row by row with rows y-1, y, y+1 of psi_{m-1}
// (chunk plus one site either side, periodic in y and z) held in LDS.
// by == 0: no window.
struct WindowMap {
  int slab, row, slabs, chunks, ny, by, chunk;
  Divisor row_div;
};

// Two Chebyshev steps per launch on periodic slab lattices (k_sweep2):
// slabs of S sites (n = S * nslab); slot q of a site at position pos of
// slab x holds, for its orbital o = pos mod p (p divides S), the neighbour
// at slab x + tab[q p + o].x (mod nslab) and position pos + tab[..].y
// (mod S) with class tab[..].z, or nothing where tab[..].w is 0.  hh is the
// largest |position offset|.  A tile is a chunk of cs positions (the last
// chunk of a slab may be shorter; one chunk of the whole slab with hh = 0
// when it fits a workgroup) of a run of slabs; tiles are numbered chunks
// fastest (run * nch + chunk), runs split the slabs evenly.
constexpr int kSweepThreads = 256;
// slabs per loop iteration of k_sweep2 (measured on the 8.0M-site honeycomb:
// 2, 4 and 8 gave 73.0-75.4, 73.7 and 102.8 us per FP64 step, 38.9-40.6,
// 38.8 and 53.7 us per FP32 step)
constexpr int kSweepBlock = 4;
constexpr int kSweepMaxHalo = 8;
constexpr int kSweepMinSites = 1 << 14;
constexpr int kComputeUnits = 256;     // MI355X: 8 XCDs x 32 CUs
constexpr size_t kLdsPerCu = 160 * 1024;
struct SweepMap {
  int S;       // sites per slab (0: no two-step sweep)
  int nslab;
  int hh;      // in-slab halo
  int cs;      // positions per chunk
  int nch;     // chunks per slab
  int wmax;    // psi_{m-1} entries held per slab: cs + 4 hh
  int nruns;   // runs of slabs per chunk (set per launch shape)
  int tiles;   // nch * nruns
  int blocks;  // tiles padded to a multiple of kXcds
  Divisor sdiv;
  int real;          // every used hopping has a zero imaginary part
  int p;             // orbitals per pattern period
  const int4* tab;   // [slot][orbital]: slab offset, position offset, class, used
  int halo;          // largest |position offset| of the table
  int ns;            // steps per launch of the shape (k_sweepn: 3 or 4)
  int nb;            // slabs per loop iteration (k_sweepn)
  int full;          // every orbital uses every slot
};

// the series coefficients of the levels of a multi-step launch of the
// evolution (k_sweepn): 2 J_m(t) and the phase case of each term
constexpr int kSweepMaxSteps = 8;
struct SweepTerms {
  double bm[kSweepMaxSteps];
  int label[kSweepMaxSteps];
};

// Two steps per launch for blocks of vectors (k_block_sweep2): the slab
// geometry of one system of a (possibly segmented) Hamiltonian, and its
// chunks and runs for one launch shape.
struct BlockSweepSeg {
  int base;   // first site
  int S;      // sites per slab
  int nslab;  // slabs (periodic)
  int halo;   // largest |position offset| of its table
  int p;      // table period (orbitals)
  int tab;    // offset of its [slot][orbital] table in BlockSweepMap::tab
  // launch shape
  int hh;     // in-slab halo held per chunk (0: the whole slab is one chunk)
  int cs;     // positions per chunk
  int nch;    // chunks per slab
  int nruns;  // runs of slabs per chunk
  int tile0;  // first tile (moment partial) of the system
};

struct BlockSweepMap {
  int tiles;             // moment partials per launch (tiles of all systems)
  int blocks;            // workgroups: dispatch slots, a multiple of kXcds
  int segs;
  const BlockSweepSeg* seg;
  const int4* order;     // per dispatch slot: tile, system, chunk, run (tile < 0: idle)
  const int4* tab;       // the systems' slot tables, as SweepMap::tab
  const int* seg_tiles;  // first tile of each system, then the tile count
};

struct TileMap {
  int n;
  int rows;
  int slab;
  int span;
  int slabs;
  int tiles;
  int row;
  int walk;
  const int2* table;
  const int* tile_seg;
  const int* seg_tiles;
  int segs;
  WindowMap win;
};

// A lattice model for device-side construction: cells[3] x orbitals
// pristine sites, K hopping slots per orbital in the tables; hop[m*K + i] =
// (dx, dy, dz, target orbital) of hopping i of orbital m (hop_count[m] of
// them), cls[m*K + i] its class in the Hamiltonian's table.  site_index
// (null: the identity) maps a pristine site to its index after vacancies
// were removed, -1 for a vacancy; offset is the lattice's first site in the
// Hamiltonian (segmented Hamiltonians hold several lattices side by side).
struct LatticeDev {
  int cells[3];
  int pbc[3];
  int orbitals;
  int K;
  const int* hop_count;
  const int4* hop;
  const int* cls;
  const int* site_index;
  int offset;
};

inline Grid make_grid(int n, int spt)
{
  int per_tile = kBlock * spt;
  int tiles = (n + per_tile - 1) / per_tile;
  int blocks = (tiles + kXcds - 1) / kXcds * kXcds;
  return Grid{tiles, blocks};
}

// steps (U(t) terms) per launch of the block path's deep sweep, k_block_sweepn
// (uneven batch, 16 and 8 columns: 4.54e11 site-updates/s with four, 4.37e11
// with six, whose 72 KB of rings leave two workgroups per CU)
constexpr int kBlockSweepSteps = 4;

// the multi-step sweeps unroll their loop by the period of their ring rows
constexpr int sweep_gcd(int a, int b)
{
  return b ? sweep_gcd(b, a % b) : a;
}

// an empty asm reading the registers of v: the compiler places the wait for
// the load that wrote them here
__device__ __forceinline__ void sweep_hold(float v)
{
  asm volatile("" : : "v"(v));
}
__device__ __forceinline__ void sweep_hold(double v)
{
  asm volatile("" : : "v"(v));
}
__device__ __forceinline__ void sweep_hold(float2 v)
{
  asm volatile("" : : "v"(v.x), "v"(v.y));
}
__device__ __forceinline__ void sweep_hold(double2 v)
{
  asm volatile("" : : "v"(v.x), "v"(v.y));
}

// workgroup b runs on XCD b % 8; map it to tile (b % 8) * (blocks/8) + b / 8
__device__ __forceinline__ int xcd_tile(int b, int blocks)
{
  return (b % kXcds) * (blocks / kXcds) + b / kXcds;
}

}  // namespace lsqt
