// capi_block.hip -- the block path of the C ABI: blocks of R vectors, the
// block moment recursion (direct, self products and pairs of recursions by
// the product identities), the block time-correlation drivers, segmented
// Hamiltonians, hipGraph replay of moment sets, and the autotuner that picks
// the kernel shape.
//
// Autotuner: for a Hamiltonian and R it times a few fused steps of every
// candidate (W = vectors per lane in {1, 2, 4, 8}, i.e. vectors-per-wavefront
// VPW = R / W lanes per site; natural tiles of 256, 1024 or 4096 sites; on a
// detected lattice also blocked tiles, the column walk and the sliding
// window; one or two passes in flight) on the caller's data, keeps the
// fastest, and remembers it per (R, z, table placement, size class, lattice
// slab, pair mode) in memory and in a text cache file ($LSQT_TUNE_CACHE,
// default $HOME/.cache/lsqt_hip_tune.txt).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "bessel.h"
#include "lsqt_capi_internal.h"
#include "lsqt_hip.h"
#include "lsqt_launch.h"

using namespace lsqt;

struct lsqt_block {
  lsqt_ctx* ctx = nullptr;
  int n = 0;
  int R = 0;
  float2* d = nullptr;  // [n][R]
};

namespace {
struct BlockScratch {
  float2* A = nullptr;  // moment recursion ping-pong, evolution temporaries
  float2* B = nullptr;
  size_t elems = 0;
  float2* C = nullptr;  // the two outputs of two-step launches
  float2* D = nullptr;
  size_t cd_elems = 0;
  float2* F[3] = {nullptr, nullptr, nullptr};  // left vector; [X, U] temporaries
  size_t f_elems = 0;
  double2* S[3] = {nullptr, nullptr, nullptr};  // FP64 states of the time drivers
  size_t s_elems = 0;
  double2* norms = nullptr;  // [2][segments][R]
  size_t norms_cap = 0;
  double2* ring = nullptr;
  size_t ring_cap = 0;
  double2* mu = nullptr;
  size_t mu_cap = 0;
  double2* mu_host = nullptr;
  size_t mu_host_cap = 0;
};
}  // namespace

// A kernel shape: W vectors per lane (R/W lanes per site), tile size and
// tile schedule of the block step.  blocked: 0 natural tiles of `rows`
// sites, 1 blocked tiles (rows at one offset of every lattice slab), 2
// blocked with the column walk, 3 the sliding window (tiles of `rows`
// lattice rows of one chunk; see make_window_map).
struct BlockShape {
  int W = 0;
  int rows = 0;
  int blocked = 0;
  int depth = 1;  // passes in flight per wave: 1, or 2 (double-buffered loop)
};

// tiles of a segmented Hamiltonian for one tile size: `rows` sites per tile,
// never across a segment boundary
struct SegTable {
  int tiles = 0;
  int2* table = nullptr;
  int* tile_seg = nullptr;
  int* seg_tiles = nullptr;
};

// a captured moment set: every argument its kernels were recorded with
struct GraphKey {
  int self, M, R, W, rows, blocked, depth, tiles;
  const void *left, *A, *B, *ring, *mu, *C, *D, *sweep, *deep;
  bool operator<(const GraphKey& o) const
  {
    return std::tie(self, M, R, W, rows, blocked, depth, tiles, left, A, B, ring, mu, C, D,
                    sweep, deep) < std::tie(o.self, o.M, o.R, o.W, o.rows, o.blocked,
                                            o.depth, o.tiles, o.left, o.A, o.B, o.ring, o.mu,
                                            o.C, o.D, o.sweep, o.deep);
  }
};

// The launch shape of the block path's two-step sweep for one column count:
// its tile map (on the device) and the tile map that reduces its partials
// per system.
struct BlockSweepShape {
  bool ok = false;
  int W = 0;
  BlockSweepMap map{};
  TileMap red{};
  std::vector<void*> mem;
};

struct lsqt_block_state {
  BlockScratch s;
  std::map<int, BlockShape> forced;  // R -> shape set by the caller
  std::map<int, SegTable> seg;       // rows -> tile table
  std::map<int, BlockSweepShape> bsw;  // 8 columns + steps -> sweep shape
  // moment sets are captured once and replayed as hipGraphs (LSQT_GRAPHS=0
  // or lsqt_block_set_graphs(h, 0) launches them one by one instead)
  bool graphs = true;
  std::map<GraphKey, hipGraphExec_t> graph;
  lsqt_block_state()
  {
    const char* e = getenv("LSQT_GRAPHS");
    graphs = !(e && e[0] == '0');
  }
};

void lsqt::block_release(lsqt_ham* h)
{
  if (!h->blk)
    return;
  // graphs first: nothing of theirs may run or refer to memory freed below
  (void)hipStreamSynchronize(h->ctx->stream);
  for (auto& kv : h->blk->graph)
    (void)hipGraphExecDestroy(kv.second);
  h->blk->graph.clear();
  BlockScratch& s = h->blk->s;
  for (void* p : {(void*)s.A, (void*)s.B, (void*)s.C, (void*)s.D, (void*)s.ring, (void*)s.mu,
                  (void*)s.F[0], (void*)s.F[1], (void*)s.F[2], (void*)s.S[0], (void*)s.S[1],
                  (void*)s.S[2], (void*)s.norms})
    if (p)
      (void)hipFree(p);
  for (auto& kv : h->blk->bsw)
    for (void* p : kv.second.mem)
      (void)hipFree(p);
  for (auto& kv : h->blk->seg)
    for (void* p : {(void*)kv.second.table, (void*)kv.second.tile_seg,
                    (void*)kv.second.seg_tiles})
      if (p)
        (void)hipFree(p);
  if (s.mu_host)
    (void)hipHostFree(s.mu_host);
  delete h->blk;
  h->blk = nullptr;
}

namespace {

constexpr int kRing = 64;  // moment partials are reduced every kRing steps

// How candidate shapes are timed (and cached apart): direct steps
// <left|psi_m> (also the shape of self-product steps: timing those on their
// own picked shapes up to 4 % slower over a whole moment set on the 160^3
// cube), or pair moment sets (R columns of R/2 pairs in `right`, left
// unused, shapes with W >= 2).
enum TuneMode { kTuneDirect = 0, kTunePairs = 1 };

// pair: the TuneMode of the cached shape
struct TuneKey {
  int R, K, lds, size_class, slab, pair = 0;
  bool operator<(const TuneKey& o) const
  {
    return std::tie(R, K, lds, size_class, slab, pair) <
           std::tie(o.R, o.K, o.lds, o.size_class, o.slab, o.pair);
  }
};

struct TuneVal {
  BlockShape shape;
  double us;
};

std::mutex g_tune_mutex;
std::map<TuneKey, TuneVal> g_tune;
bool g_tune_loaded = false;

std::string tune_path()
{
  if (const char* p = getenv("LSQT_TUNE_CACHE"))
    return p;
  const char* home = getenv("HOME");
  return std::string(home ? home : "/tmp") + "/.cache/lsqt_hip_tune.txt";
}

void tune_load()
{
  if (g_tune_loaded)
    return;
  g_tune_loaded = true;
  FILE* f = fopen(tune_path().c_str(), "r");
  if (!f)
    return;
  char arch[32];
  TuneKey k;
  TuneVal v;
  while (fscanf(f, "%31s %d %d %d %d %d %d %d %d %d %lf", arch, &k.R, &k.K, &k.lds,
                &k.size_class, &k.slab, &v.shape.W, &v.shape.rows, &v.shape.blocked,
                &v.shape.depth, &v.us) == 11) {
    k.pair = !std::strcmp(arch, "gfx950-pair") ? kTunePairs : kTuneDirect;
    if (k.pair != kTuneDirect || std::strcmp(arch, "gfx950") == 0)
      g_tune[k] = v;
  }
  fclose(f);
}

void tune_store(const TuneKey& k, const TuneVal& v)
{
  g_tune[k] = v;
  const std::string path = tune_path();
  std::error_code ec;
  std::filesystem::create_directories(std::filesystem::path(path).parent_path(), ec);
  FILE* f = fopen(path.c_str(), "a");
  if (!f)
    return;  // the cache file is an optimisation only
  static const char* kTag[2] = {"gfx950", "gfx950-pair"};
  fprintf(f, "%s %d %d %d %d %d %d %d %d %d %.3f\n", kTag[k.pair], k.R,
          k.K, k.lds, k.size_class, k.slab, v.shape.W, v.shape.rows, v.shape.blocked,
          v.shape.depth, v.us);
  fclose(f);
}

int size_class(int n)
{
  int c = 0;
  while ((1 << (c + 1)) <= n)
    ++c;
  return c;
}

bool valid_R(int R)
{
  return R >= 1 && R <= 256 && (R & (R - 1)) == 0;
}

bool valid_W(int R, int W)
{
  return (W == 1 || W == 2 || W == 4 || W == 8) && R % W == 0 && R / W <= 64;
}

// whether a block-step shape can run on this Hamiltonian (the conditions
// lsqt_block_set_config checks)
bool shape_usable(const lsqt_ham* h, int R, const BlockShape& c)
{
  if (!valid_W(R, c.W) || (c.depth != 1 && c.depth != 2))
    return false;
  if (c.blocked == 3)
    return c.depth == 1 && (h->d.K == 3 || h->d.K == 4 || h->d.K == 6) &&
           window_fits(h->d.n, h->slab, h->row, c.rows, R, c.W);
  return c.blocked >= 0 && c.blocked <= 2 && c.rows >= 64 && c.rows % 64 == 0;
}

int block_alloc(lsqt_ctx* ctx, size_t count, float2** out)
{
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(float2) * (count ? count : 1));
  if (e != hipSuccess)
    return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu float2): %s", count, hipGetErrorString(e));
  *out = (float2*)p;
  (void)ctx;
  return LSQT_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
// blocks
// ---------------------------------------------------------------------------
extern "C" int lsqt_block_create(lsqt_ctx* ctx, int n, int R, lsqt_block** out)
{
  if (!ctx || !out)
    return set_error(LSQT_ERR_INVALID, "null context or output handle");
  *out = nullptr;
  if (n <= 0 || !valid_R(R))
    return set_error(LSQT_ERR_INVALID, "block of %d sites x %d vectors: R must be a power of 2 "
                     "in 1..256", n, R);
  if ((size_t)n * R > (size_t)INT_MAX)
    return set_error(LSQT_ERR_INVALID, "block of %d sites x %d vectors exceeds 2^31 values", n,
                     R);
  LSQT_TRY(ctx_activate(ctx));
  lsqt_block* b = new lsqt_block();
  b->ctx = ctx;
  ++ctx->live;
  b->n = n;
  b->R = R;
  int rc = block_alloc(ctx, (size_t)n * R, &b->d);
  if (rc != LSQT_OK) {
    --ctx->live;
    delete b;
    return rc;
  }
  launch_zero32(b->d, (int)((size_t)n * R), ctx->stream);
  *out = b;
  return LSQT_OK;
}

extern "C" int lsqt_block_destroy(lsqt_block* b)
{
  if (!b)
    return LSQT_OK;
  (void)hipSetDevice(b->ctx->device);
  (void)hipStreamSynchronize(b->ctx->stream);
  (void)hipFree(b->d);
  --b->ctx->live;
  delete b;
  return LSQT_OK;
}

namespace {

// a device buffer of `bytes` for one layout change, freed by the destructor
struct Staging {
  void* p = nullptr;
  ~Staging()
  {
    if (p)
      (void)hipFree(p);
  }
  int alloc(size_t bytes)
  {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess)
      return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    return LSQT_OK;
  }
};

}  // namespace

extern "C" int lsqt_block_from_host(lsqt_block* b, const double* re, const double* im)
{
  if (!b || !re || !im)
    return set_error(LSQT_ERR_INVALID, "null argument");
  LSQT_TRY(ctx_activate(b->ctx));
  const size_t count = (size_t)b->n * b->R;
  Staging tmp;
  LSQT_TRY(tmp.alloc(2 * sizeof(double) * count));
  double* d_re = (double*)tmp.p;
  double* d_im = d_re + count;
  hipStream_t st = b->ctx->stream;
  LSQT_HIP(hipStreamSynchronize(st));
  LSQT_HIP(hipMemcpy(d_re, re, sizeof(double) * count, hipMemcpyHostToDevice));
  LSQT_HIP(hipMemcpy(d_im, im, sizeof(double) * count, hipMemcpyHostToDevice));
  launch_site_major(d_re, d_im, b->n, b->R, b->d, st);
  LSQT_HIP(hipGetLastError());
  LSQT_HIP(hipStreamSynchronize(st));
  return LSQT_OK;
}

extern "C" int lsqt_block_to_host(const lsqt_block* b, double* re, double* im)
{
  if (!b || !re || !im)
    return set_error(LSQT_ERR_INVALID, "null argument");
  LSQT_TRY(ctx_activate(b->ctx));
  const size_t count = (size_t)b->n * b->R;
  Staging tmp;
  LSQT_TRY(tmp.alloc(2 * sizeof(double) * count));
  double* d_re = (double*)tmp.p;
  double* d_im = d_re + count;
  hipStream_t st = b->ctx->stream;
  launch_vector_major(b->d, b->n, b->R, d_re, d_im, st);
  LSQT_HIP(hipGetLastError());
  LSQT_HIP(hipStreamSynchronize(st));
  LSQT_HIP(hipMemcpy(re, d_re, sizeof(double) * count, hipMemcpyDeviceToHost));
  LSQT_HIP(hipMemcpy(im, d_im, sizeof(double) * count, hipMemcpyDeviceToHost));
  return LSQT_OK;
}

// ---------------------------------------------------------------------------
// block moments
// ---------------------------------------------------------------------------
namespace {

int segments(const lsqt_ham* h)
{
  return h->seg_starts.empty() ? 1 : (int)h->seg_starts.size() - 1;
}

int scratch_for(lsqt_ham* h, int R, int M, int tiles, BlockScratch** out)
{
  if (!h->blk)
    h->blk = new lsqt_block_state();
  BlockScratch& s = h->blk->s;
  const size_t elems = (size_t)h->d.n * R;
  if (s.elems < elems) {
    if (s.A)
      (void)hipFree(s.A);
    if (s.B)
      (void)hipFree(s.B);
    s.A = s.B = nullptr;
    s.elems = 0;
    LSQT_TRY(block_alloc(h->ctx, elems, &s.A));
    LSQT_TRY(block_alloc(h->ctx, elems, &s.B));
    s.elems = elems;
  }
  LSQT_TRY(ensure_device((void**)&s.ring, &s.ring_cap,
                         sizeof(double2) * (size_t)kRing * tiles * R));
  const size_t outs = (size_t)M * R * segments(h);
  LSQT_TRY(ensure_device((void**)&s.mu, &s.mu_cap, sizeof(double2) * outs));
  LSQT_TRY(ensure_pinned((void**)&s.mu_host, &s.mu_host_cap, sizeof(double2) * outs));
  *out = &s;
  return LSQT_OK;
}

// enqueues steps [from, to) of the recursion; step m's partials go to ring
// slot m % kRing, and full rings are reduced into mu
// the tile table of a segmented Hamiltonian for tiles of `rows` sites
int seg_table(lsqt_ham* h, int rows, SegTable** out)
{
  if (!h->blk)
    h->blk = new lsqt_block_state();
  auto it = h->blk->seg.find(rows);
  if (it == h->blk->seg.end()) {
    std::vector<int2> table;
    std::vector<int> tile_seg, seg_tiles;
    const int S = segments(h);
    for (int k = 0; k < S; ++k) {
      seg_tiles.push_back((int)table.size());
      for (int b = h->seg_starts[k]; b < h->seg_starts[k + 1]; b += rows) {
        table.push_back(make_int2(b, std::min(b + rows, h->seg_starts[k + 1])));
        tile_seg.push_back(k);
      }
    }
    seg_tiles.push_back((int)table.size());
    SegTable t;
    t.tiles = (int)table.size();
    size_t cap = 0;
    LSQT_TRY(ensure_device((void**)&t.table, &cap, sizeof(int2) * table.size()));
    cap = 0;
    LSQT_TRY(ensure_device((void**)&t.tile_seg, &cap, sizeof(int) * tile_seg.size()));
    cap = 0;
    LSQT_TRY(ensure_device((void**)&t.seg_tiles, &cap, sizeof(int) * seg_tiles.size()));
    LSQT_HIP(hipMemcpy(t.table, table.data(), sizeof(int2) * table.size(),
                       hipMemcpyHostToDevice));
    LSQT_HIP(hipMemcpy(t.tile_seg, tile_seg.data(), sizeof(int) * tile_seg.size(),
                       hipMemcpyHostToDevice));
    LSQT_HIP(hipMemcpy(t.seg_tiles, seg_tiles.data(), sizeof(int) * seg_tiles.size(),
                       hipMemcpyHostToDevice));
    it = h->blk->seg.emplace(rows, t).first;
  }
  *out = &it->second;
  return LSQT_OK;
}

// blocked: 0 natural, 1 blocked, 2 blocked with the column walk; segmented
// Hamiltonians use their tile table (natural tiles inside each segment)
int tile_map(lsqt_ham* h, int R, const BlockShape& c, TileMap* tm)
{
  if (!h->seg_starts.empty()) {
    SegTable* t;
    LSQT_TRY(seg_table(h, c.rows, &t));
    TileMap m{};
    m.n = h->d.n;
    m.rows = c.rows;
    m.tiles = t->tiles;
    m.table = t->table;
    m.tile_seg = t->tile_seg;
    m.seg_tiles = t->seg_tiles;
    m.segs = segments(h);
    *tm = m;
    return LSQT_OK;
  }
  const int per_pass = kBlock / (R / c.W);
  if (c.blocked == 3) {
    if (!window_fits(h->d.n, h->slab, h->row, c.rows, R, c.W))
      return set_error(LSQT_ERR_INVALID, "the sliding window does not fit this lattice "
                       "(slab %d, row %d, %d rows per tile, R=%d, W=%d)", h->slab, h->row,
                       c.rows, R, c.W);
    *tm = make_window_map(h->d.n, h->slab, h->row, c.rows, per_pass);
    return LSQT_OK;
  }
  *tm = make_tile_map(h->d.n, c.rows, c.blocked ? h->slab : 0, c.blocked ? c.rows : 0, h->row,
                      c.blocked == 2 ? per_pass : 0);
  return LSQT_OK;
}

// The launch shape of the block sweep with `steps` (2 or kBlockSweepSteps)
// steps per launch
// for blocks of `cols` columns (*out null where the Hamiltonian has no sweep
// geometry for each of its systems or the columns do not fit): W = 2
// columns per lane up to 16 columns, else 4 (two steps only), so cols / W
// lanes per site and kBlock W / cols positions per slab row; a chunk holds
// them less 2 steps halo positions (the whole slab where it fits).
// Runs of slabs per chunk: about twice as many tiles as resident
// workgroups, and the dispatch order puts the longest runs first, in groups
// of 8 neighbouring chunks dealt to the XCDs in turn, so that the systems of
// an uneven batch finish together.
int block_sweep_shape(lsqt_ham* h, int cols, int steps, BlockSweepShape** out)
{
  *out = nullptr;
  if (!h->blk)
    h->blk = new lsqt_block_state();
  if (h->bsweep.empty() || !h->bsweep_tab || cols < 2 || cols > 64 || !valid_R(cols) ||
      (int)h->bsweep.size() != segments(h))
    return LSQT_OK;
  if (!(h->d.K == 2 || h->d.K == 3 || h->d.K == 4 || h->d.K == 6))
    return LSQT_OK;
  if (steps == kBlockSweepSteps && (cols > 16 || h->bsweep_steps < kBlockSweepSteps))
    return LSQT_OK;  // the deep sweep's levels of wider entries leave too little LDS
  auto it = h->blk->bsw.find(cols * 8 + steps);
  if (it != h->blk->bsw.end()) {
    *out = it->second.ok ? &it->second : nullptr;
    return LSQT_OK;
  }
  BlockSweepShape& sh = h->blk->bsw[cols * 8 + steps];
  const int W = cols <= 16 ? 2 : 4;
  const int P = kBlock * W / cols;
  std::vector<BlockSweepSeg> segs = h->bsweep;
  long long units = 0;
  for (size_t k = 0; k < segs.size(); ++k) {
    BlockSweepSeg& g = segs[k];
    const long long lo = h->seg_starts.empty() ? 0 : h->seg_starts[k];
    const long long hi = h->seg_starts.empty() ? h->d.n : h->seg_starts[k + 1];
    if (g.base != lo || (long long)g.S * g.nslab != hi - lo || g.nslab < 4)
      return LSQT_OK;
    if (g.S <= P) {
      g.hh = 0;
      g.cs = g.S;
      g.nch = 1;
    } else if (4 * steps * g.halo <= P) {
      // a chunk holds the row's positions less the halo of every step
      g.hh = g.halo;
      g.cs = P - 2 * steps * g.hh;
      g.nch = (g.S + g.cs - 1) / g.cs;
    } else {
      return LSQT_OK;
    }
    units += (long long)g.nch * g.nslab;
  }
  const size_t lds = steps == kBlockSweepSteps ? block_sweepn_lds() : block_sweep_lds(W);
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, kLdsPerCu / lds));
  const long long run_len = std::max(16ll, units / (2ll * kComputeUnits * per_cu));
  int tiles = 0;
  struct Entry {
    int tile, seg, chunk, run, len;
  };
  std::vector<Entry> list;
  for (size_t k = 0; k < segs.size(); ++k) {
    BlockSweepSeg& g = segs[k];
    long long runs = (g.nslab + run_len / 2) / run_len;
    g.nruns = (int)std::max(1ll, std::min(runs, (long long)g.nslab / 2));
    g.tile0 = tiles;
    if ((long long)tiles + (long long)g.nch * g.nruns > INT_MAX / 8)
      return LSQT_OK;
    tiles += g.nch * g.nruns;
    for (int r = 0; r < g.nruns; ++r) {
      const int len = (int)((long long)(r + 1) * g.nslab / g.nruns) -
                      (int)((long long)r * g.nslab / g.nruns);
      for (int c = 0; c < g.nch; ++c)
        list.push_back(Entry{g.tile0 + r * g.nch + c, (int)k, c, r, len});
    }
  }
  std::stable_sort(list.begin(), list.end(),
                   [](const Entry& a, const Entry& b) { return a.len > b.len; });
  constexpr size_t kGroup = 8;
  std::vector<std::vector<int4>> per_xcd(kXcds);
  for (size_t i = 0; i < list.size(); ++i)
    per_xcd[(i / kGroup) % kXcds].push_back(
      make_int4(list[i].tile, list[i].seg, list[i].chunk, list[i].run));
  size_t per = 0;
  for (const auto& v : per_xcd)
    per = std::max(per, v.size());
  std::vector<int4> order(per * kXcds, make_int4(-1, 0, 0, 0));
  for (int x = 0; x < kXcds; ++x)
    std::copy(per_xcd[x].begin(), per_xcd[x].end(), order.begin() + (size_t)x * per);
  std::vector<int> seg_tiles;
  for (const BlockSweepSeg& g : segs)
    seg_tiles.push_back(g.tile0);
  seg_tiles.push_back(tiles);
  auto up = [&](const void* src, size_t bytes, void** dst) -> int {
    LSQT_HIP(hipMalloc(dst, bytes));
    sh.mem.push_back(*dst);
    LSQT_HIP(hipMemcpy(*dst, src, bytes, hipMemcpyHostToDevice));
    return LSQT_OK;
  };
  void *d_seg, *d_order, *d_tiles;
  LSQT_TRY(up(segs.data(), sizeof(BlockSweepSeg) * segs.size(), &d_seg));
  LSQT_TRY(up(order.data(), sizeof(int4) * order.size(), &d_order));
  LSQT_TRY(up(seg_tiles.data(), sizeof(int) * seg_tiles.size(), &d_tiles));
  sh.W = W;
  sh.map.tiles = tiles;
  sh.map.blocks = (int)per * kXcds;
  sh.map.segs = (int)segs.size();
  sh.map.seg = (const BlockSweepSeg*)d_seg;
  sh.map.order = (const int4*)d_order;
  sh.map.tab = h->bsweep_tab;
  sh.map.seg_tiles = (const int*)d_tiles;
  sh.red = TileMap{};
  sh.red.n = h->d.n;
  sh.red.tiles = tiles;
  sh.red.seg_tiles = (const int*)d_tiles;
  sh.red.segs = (int)segs.size();
  sh.ok = true;
  *out = &sh;
  return LSQT_OK;
}

// scratch of a two-step moment set of `cols` columns: the outputs C and D,
// and a ring of kRing slots of stride_tiles partials per column
int block_sweep_scratch(lsqt_ham* h, int cols, int stride_tiles, BlockScratch* s)
{
  const size_t elems = (size_t)h->d.n * cols;
  if (s->cd_elems < elems) {
    if (s->C)
      (void)hipFree(s->C);
    if (s->D)
      (void)hipFree(s->D);
    s->C = s->D = nullptr;
    s->cd_elems = 0;
    LSQT_TRY(block_alloc(h->ctx, elems, &s->C));
    LSQT_TRY(block_alloc(h->ctx, elems, &s->D));
    s->cd_elems = elems;
  }
  LSQT_TRY(ensure_device((void**)&s->ring, &s->ring_cap,
                         sizeof(double2) * (size_t)kRing * stride_tiles * cols));
  return LSQT_OK;
}

void enqueue_steps(lsqt_ham* h, int R, const BlockShape& c, const TileMap& tm, const float2* L,
                   float2*& A, float2*& B, double2* ring, double2* mu, int M, int from, int to,
                   hipStream_t st)
{
  const int tiles = tm.tiles;
  const size_t slot = (size_t)tiles * R;
  for (int m = from; m < to; ++m) {
    double2* part = ring + (size_t)(m % kRing) * slot;
    if (m == 0) {
      launch_block_dot(A, L, R, tm, part, st);
    } else if (m == 1) {
      launch_block_step(h->d, R, c.W, c.depth == 2, tm, A, nullptr, B, L, part, nullptr, false,
                        kStepLeft, st);
    } else {
      launch_block_step(h->d, R, c.W, c.depth == 2, tm, B, A, A, L, part, nullptr, true,
                        kStepLeft, st);
      float2* t = A;
      A = B;
      B = t;
    }
    if (mu && (m % kRing == kRing - 1 || m == M - 1)) {
      int m0 = m - m % kRing;
      launch_block_reduce(ring, tm, R, M, m0, m - m0 + 1, mu, st);
    }
  }
}

// The product-identity moments (A is overwritten).  R vectors (pairs
// false): S_0 = <v|v>; step m >= 1 makes psi_m (psi_0 = v) and adds
// S_{2m-1} = <psi_{m-1}|psi_m>, S_{2m} = <psi_m|psi_m>.  Pairs: A holds R
// pairs of columns (a left vector L_r, a right vector R_r), both recursions
// advance together and S_0 = <L|R>, S_{2m-1} = <T_{m-1} L|T_m R>, S_{2m} =
// <T_m L|T_m R>.  The moments follow from T_{2m} = 2 T_m^2 - T_0 and
// T_{2m+1} = 2 T_m T_{m+1} - T_1 with T_m Hermitian: mu_0 = S_0, mu_1 = S_1,
// mu_k = 2 S_k - mu_{k mod 2}.  M/2 steps give moments 0..M-1.
// With sweep shapes (sw: two steps, deep: kBlockSweepSteps; outputs C and
// D), steps 2 .. are taken kBlockSweepSteps per launch, then two, and the
// rest (and step 1) by the one-step kernels.
// Moments k >= 1 fill ring slots [0, used) from moment k_first, slots
// stride tiles apart; a ring is reduced when the next launch's slots would
// not fit or the launches change their tile map, and at the end.
void enqueue_self_steps(lsqt_ham* h, int R, const BlockShape& c, const TileMap& tm, float2*& A,
                        float2*& B, double2* ring, double2* mu, int M, hipStream_t st,
                        bool pairs = false, const BlockSweepShape* sw = nullptr,
                        float2* C = nullptr, float2* D = nullptr,
                        const BlockSweepShape* deep = nullptr)
{
  const int cols = pairs ? 2 * R : R;
  const int mode = pairs ? kStepPair : kStepSelf;
  const int stride = std::max(tm.tiles, std::max(sw ? sw->map.tiles : 0,
                                                 deep ? deep->map.tiles : 0));
  const size_t slot = (size_t)stride * R;
  launch_block_dot(A, A, cols, tm, ring, st, pairs);
  launch_block_reduce(ring, tm, R, M, 0, 1, mu, st, stride);
  int used = 0, k_first = 1;
  const TileMap* cur = &tm;
  auto flush = [&]() {
    launch_block_reduce(ring, *cur, R, M, k_first, std::min(used, M - k_first), mu, st, stride);
    k_first += used;
    used = 0;
  };
  auto take = [&](int count, const TileMap* map) {
    if (used > 0 && (map != cur || used + count > kRing))
      flush();
    cur = map;
    double2* p = ring + (size_t)used * slot;
    used += count;
    return p;
  };
  const int steps = M / 2;
  for (int m = 1; m <= steps;) {
    if (deep && m >= 2 && m + kBlockSweepSteps - 1 <= steps) {
      // steps m .. m + kBlockSweepSteps - 1: two moments each from 2m - 1
      double2* p = take(2 * kBlockSweepSteps, &deep->red);
      launch_block_sweepn(h->d, deep->map, cols, pairs, A, B, C, D, p, slot, st);
      std::swap(A, C);
      std::swap(B, D);
      m += kBlockSweepSteps;
      continue;
    }
    if (sw && m >= 2 && m + 1 <= steps) {
      double2* p = take(4, &sw->red);
      double2* const part[4] = {p, p + slot, p + 2 * slot, p + 3 * slot};
      launch_block_sweep2(h->d, sw->map, cols, sw->W, pairs, A, B, C, D, part, st);
      std::swap(A, C);
      std::swap(B, D);
      m += 2;
      continue;
    }
    double2* p = take(2, &tm);
    if (m == 1) {
      launch_block_step(h->d, cols, c.W, c.depth == 2, tm, A, nullptr, B, nullptr, p, p + slot,
                        false, mode, st);
    } else {
      launch_block_step(h->d, cols, c.W, c.depth == 2, tm, B, A, A, nullptr, p, p + slot, true,
                        mode, st);
      float2* t = A;
      A = B;
      B = t;
    }
    ++m;
  }
  if (used > 0)
    flush();
}

int pick_config(lsqt_ham* h, int R, const float2* left, const float2* right, BlockShape* out,
                double* us_out, bool force, int mode = kTuneDirect);

}  // namespace

static int check_block(lsqt_ham* h, const lsqt_block* a, const lsqt_block* b)
{
  if (!h || !a || !b)
    return set_error(LSQT_ERR_INVALID, "null argument");
  if (a->n != h->d.n || b->n != h->d.n || a->R != b->R || a->ctx != h->ctx || b->ctx != h->ctx)
    return set_error(LSQT_ERR_INVALID, "block shapes do not match the Hamiltonian");
  if (h->d.layout == kFull)
    return set_error(LSQT_ERR_UNSUPPORTED, "the block path needs the packed layout");
  return ctx_activate(h->ctx);
}

extern "C" int lsqt_block_set_config(lsqt_ham* h, int R, int W, int rows, int blocked,
                                     int depth)
{
  if (!h || !valid_R(R))
    return set_error(LSQT_ERR_INVALID, "bad Hamiltonian or R");
  if (!h->blk)
    h->blk = new lsqt_block_state();
  if (W == 0) {
    h->blk->forced.erase(R);
    return LSQT_OK;
  }
  if (blocked < 0 || blocked > 3)
    return set_error(LSQT_ERR_INVALID, "tile schedule %d is not 0, 1, 2 or 3", blocked);
  if (depth != 1 && depth != 2)
    return set_error(LSQT_ERR_INVALID, "depth %d is not 1 or 2", depth);
  if (blocked == 3) {
    if (!valid_W(R, W) || depth != 1 || h->d.layout == kFull ||
        !(h->d.K == 3 || h->d.K == 4 || h->d.K == 6) ||
        !window_fits(h->d.n, h->slab, h->row, rows, R, W))
      return set_error(LSQT_ERR_INVALID, "the sliding window needs a lattice of >= 3 rows per "
                       "slab that %d rows per tile divide, rows of whole chunks, W <= 4, 3, 4 "
                       "or 6 slots and depth 1 (slab %d, row %d, W=%d, R=%d)", rows, h->slab,
                       h->row, W, R);
    BlockShape c;
    c.W = W;
    c.rows = rows;
    c.blocked = 3;
    h->blk->forced[R] = c;
    return LSQT_OK;
  }
  if (!valid_W(R, W) || rows < 64 || rows % 64 != 0)
    return set_error(LSQT_ERR_INVALID, "W=%d rows=%d invalid for R=%d", W, rows, R);
  if (blocked && (h->slab == 0 || h->slab % rows != 0))
    return set_error(LSQT_ERR_INVALID, "blocked tiles of %d sites need a lattice slab they "
                     "divide (slab %d)", rows, h->slab);
  if (blocked == 2 && (rows % h->row != 0 || h->row % (kBlock / (R / W)) != 0))
    return set_error(LSQT_ERR_INVALID, "the column walk needs tiles of whole rows (%d sites) "
                     "that passes of %d sites divide", h->row, kBlock / (R / W));
  BlockShape c;
  c.W = W;
  c.rows = rows;
  c.blocked = blocked;
  c.depth = depth;
  h->blk->forced[R] = c;
  return LSQT_OK;
}

namespace {

// How a moment set is formed: <left|T_m|right> directly, the self products
// of `right`, or the products of a left and a right recursion run side by
// side (both by the product identities, M/2 steps).
enum MomentMode { kDirect = 0, kSelfProducts = 1, kPairProducts = 2 };

// One moment set for R vectors of every segment: <left_r|T_m|right_r>,
// downloaded into mu[(k*R + r)*M + m].  A and B of the scratch are
// overwritten.  kPairProducts: c and tm are the shape and tile map of 2R
// columns, A and B hold 2R columns, and left / right differ from A.
int moment_set(lsqt_ham* h, int R, const BlockShape& c, const TileMap& tm, BlockScratch* s,
               const float2* left, const float2* right, int mode, int M, double* mu_re,
               double* mu_im)
{
  hipStream_t st = h->ctx->stream;
  float2* A = s->A;
  float2* B = s->B;
  if (mode == kPairProducts) {
    if (left == A || right == A)
      return set_error(LSQT_ERR_INVALID, "pair moment set inputs overlap its scratch");
    launch_block_interleave(left, right, h->d.n, R, A, st);
  } else if (A != right) {
    LSQT_HIP(hipMemcpyAsync(A, right, sizeof(float2) * (size_t)h->d.n * R,
                            hipMemcpyDeviceToDevice, st));
  }
  // two steps per launch where the Hamiltonian's systems allow (product
  // identities only: their moments need no left vector)
  BlockSweepShape* sw = nullptr;
  BlockSweepShape* deep = nullptr;
  if (mode != kDirect && M / 2 >= 3) {
    const int cols = mode == kPairProducts ? 2 * R : R;
    LSQT_TRY(block_sweep_shape(h, cols, 2, &sw));
    if (sw && M / 2 >= kBlockSweepSteps + 1)
      LSQT_TRY(block_sweep_shape(h, cols, kBlockSweepSteps, &deep));
    if (sw)
      LSQT_TRY(block_sweep_scratch(
        h, cols, std::max(tm.tiles, std::max(sw->map.tiles, deep ? deep->map.tiles : 0)), s));
  }
  auto enqueue = [&]() {
    float2* a = A;
    float2* b = B;
    if (mode == kDirect)
      enqueue_steps(h, R, c, tm, left, a, b, s->ring, s->mu, M, 0, M, st);
    else
      enqueue_self_steps(h, R, c, tm, a, b, s->ring, s->mu, M, st, mode == kPairProducts, sw,
                         s->C, s->D, deep);
  };
  if (h->blk->graphs) {
    const GraphKey key{mode, M, R, c.W, c.rows, c.blocked, c.depth, tm.tiles,
                       mode == kDirect ? left : nullptr, A, B, s->ring, s->mu,
                       sw ? s->C : nullptr, sw ? s->D : nullptr, sw, deep};
    auto it = h->blk->graph.find(key);
    if (it == h->blk->graph.end()) {
      hipGraph_t g = nullptr;
      hipGraphExec_t exec = nullptr;
      LSQT_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      enqueue();
      hipError_t e = hipStreamEndCapture(st, &g);
      if (e != hipSuccess)
        return set_error(LSQT_ERR_HIP, "capturing a moment set: %s", hipGetErrorString(e));
      e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (e != hipSuccess)
        return set_error(LSQT_ERR_HIP, "instantiating a moment set: %s", hipGetErrorString(e));
      it = h->blk->graph.emplace(key, exec).first;
    }
    LSQT_HIP(hipGraphLaunch(it->second, st));
  } else {
    enqueue();
  }
  LSQT_HIP(hipGetLastError());
  const size_t count = (size_t)M * R * segments(h);
  LSQT_HIP(hipMemcpyAsync(s->mu_host, s->mu, sizeof(double2) * count, hipMemcpyDeviceToHost, st));
  LSQT_HIP(hipStreamSynchronize(st));
  for (size_t v = 0; v < count / M; ++v) {
    const double2* S = s->mu_host + v * M;
    double* re = mu_re + v * M;
    double* im = mu_im + v * M;
    if (mode == kDirect) {
      for (int k = 0; k < M; ++k) {
        re[k] = S[k].x;
        im[k] = S[k].y;
      }
      continue;
    }
    re[0] = S[0].x;
    im[0] = S[0].y;
    re[1] = S[1].x;
    im[1] = S[1].y;
    for (int k = 2; k < M; ++k) {
      re[k] = 2.0 * S[k].x - re[k % 2];
      im[k] = 2.0 * S[k].y - im[k % 2];
    }
  }
  return LSQT_OK;
}

// The shape and tile map of pair moment sets for R vectors (2R columns):
// autotuned on the pair block of (left, right), or, where the caller fixed
// the R-vector shape c, c with twice the vectors per lane (the same lanes
// per site, tiles and schedule; else W = 2 on natural tiles of 256 sites);
// scratch A / B of 2R columns.  *ok false (no error) where 2R columns exceed
// the block limit.
int pair_setup(lsqt_ham* h, int R, const BlockShape& c, const float2* left,
               const float2* right, int M, BlockShape* c2, TileMap* tm2, BlockScratch** s,
               bool* ok)
{
  *ok = false;
  const int R2 = 2 * R;
  if (!valid_R(R2) || (size_t)h->d.n * R2 > (size_t)INT_MAX)
    return LSQT_OK;
  BlockShape p;
  if (h->blk && h->blk->forced.count(R)) {
    // a shape set for R: the same lanes per site, tiles and schedule
    p = c;
    p.W = 2 * c.W;
    if (!shape_usable(h, R2, p)) {
      p = BlockShape{};
      p.W = 2;
      p.rows = 256;
      if (!valid_W(R2, p.W))
        p.W = 4;
      if (!valid_W(R2, p.W))
        return LSQT_OK;
    }
  } else {
    // tuned on the pair block of (left, right), staged in B
    BlockScratch* t;
    LSQT_TRY(scratch_for(h, R2, 2, 1, &t));
    launch_block_interleave(left, right, h->d.n, R, t->B, h->ctx->stream);
    double us;
    LSQT_TRY(pick_config(h, R2, nullptr, t->B, &p, &us, false, kTunePairs));
  }
  LSQT_TRY(tile_map(h, R2, p, tm2));
  LSQT_TRY(scratch_for(h, R2, M, tm2->tiles, s));
  *c2 = p;
  *ok = true;
  return LSQT_OK;
}

// shape, tile map and scratch (M moments) for R vectors
int block_setup(lsqt_ham* h, int R, const float2* left, const float2* right, int M,
                BlockShape* c, TileMap* tm, BlockScratch** s)
{
  LSQT_TRY(f32_potential(h));
  double us;
  LSQT_TRY(pick_config(h, R, left, right, c, &us, false));
  LSQT_TRY(tile_map(h, R, *c, tm));
  LSQT_TRY(scratch_for(h, R, M, tm->tiles, s));
  return LSQT_OK;
}

}  // namespace

extern "C" int lsqt_moments_block(lsqt_ham* h, const lsqt_block* left, const lsqt_block* right,
                                  int M, double* mu_re, double* mu_im)
{
  LSQT_TRY(check_block(h, left, right));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  BlockShape c;
  TileMap tm;
  BlockScratch* s;
  LSQT_TRY(block_setup(h, left->R, left->d, right->d, M, &c, &tm, &s));
  return moment_set(h, left->R, c, tm, s, left->d, right->d, kDirect, M, mu_re, mu_im);
}

extern "C" int lsqt_moments_block_pair(lsqt_ham* h, const lsqt_block* left,
                                       const lsqt_block* right, int M, double* mu_re,
                                       double* mu_im)
{
  LSQT_TRY(check_block(h, left, right));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  const int R = left->R;
  BlockShape c, c2;
  TileMap tm, tm2;
  BlockScratch* s;
  LSQT_TRY(block_setup(h, R, left->d, right->d, M, &c, &tm, &s));
  bool pairs = false;
  LSQT_TRY(pair_setup(h, R, c, left->d, right->d, M, &c2, &tm2, &s, &pairs));
  if (!pairs)
    return set_error(LSQT_ERR_UNSUPPORTED, "R = %d vectors: pairs need 2R <= 256", R);
  return moment_set(h, R, c2, tm2, s, left->d, right->d, kPairProducts, M, mu_re, mu_im);
}

extern "C" int lsqt_moments_block_self(lsqt_ham* h, const lsqt_block* v, int M, double* mu_re,
                                       double* mu_im)
{
  LSQT_TRY(check_block(h, v, v));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  BlockShape c;
  TileMap tm;
  BlockScratch* s;
  LSQT_TRY(block_setup(h, v->R, v->d, v->d, M, &c, &tm, &s));
  return moment_set(h, v->R, c, tm, s, nullptr, v->d, kSelfProducts, M, mu_re, mu_im);
}

// ---------------------------------------------------------------------------
// time evolution of blocks and the VAC / MSD drivers
// ---------------------------------------------------------------------------
namespace {

int evo_scratch(lsqt_ham* h, int R, BlockScratch* s)
{
  const size_t elems = (size_t)h->d.n * R;
  if (s->f_elems < elems) {
    for (float2*& p : s->F) {
      if (p)
        (void)hipFree(p);
      p = nullptr;
    }
    s->f_elems = 0;
    for (float2*& p : s->F)
      LSQT_TRY(block_alloc(h->ctx, elems, &p));
    s->f_elems = elems;
  }
  if (s->s_elems < elems) {
    for (double2*& p : s->S) {
      if (p)
        (void)hipFree(p);
      p = nullptr;
    }
    s->s_elems = 0;
    for (double2*& p : s->S) {
      void* q = nullptr;
      hipError_t e = hipMalloc(&q, sizeof(double2) * elems);
      if (e != hipSuccess)
        return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu double2): %s", elems,
                         hipGetErrorString(e));
      p = (double2*)q;
    }
    s->s_elems = elems;
  }
  LSQT_TRY(ensure_device((void**)&s->norms, &s->norms_cap,
                         sizeof(double2) * 2 * (size_t)segments(h) * R));
  return LSQT_OK;
}

// |S_r|^2 per segment into norms slot `slot` (ring slot 0 as partials)
void block_norms(lsqt_ham* h, int R, const TileMap& tm, BlockScratch* s, const double2* S,
                 int slot)
{
  hipStream_t st = h->ctx->stream;
  launch_block_norm64(S, R, tm, s->ring, st);
  launch_block_reduce(s->ring, tm, R, 1, 0, 1, s->norms + (size_t)slot * segments(h) * R, st);
}

// S <- U(direction t) S for every vector (src/sigma.cu:116-151 with FP32
// recursion states and FP64 accumulation), re-normalised to the norm before
// the step when `renormalise`
void block_evolve(lsqt_ham* h, int R, const BlockShape& c, const TileMap& tm, BlockScratch* s,
                  int direction, double t, double2* S, bool renormalise)
{
  hipStream_t st = h->ctx->stream;
  const int nR = (int)((size_t)h->d.n * R);
  // kBlockSweepSteps terms per launch where the lattice allows (its launch
  // shape and the two more recursion buffers are made on first use; without
  // them the terms take one launch each)
  BlockSweepShape* deep = nullptr;
  if (block_sweep_shape(h, R, kBlockSweepSteps, &deep) != LSQT_OK ||
      (deep && block_sweep_scratch(h, R, tm.tiles, s) != LSQT_OK))
    deep = nullptr;
  float2* A = s->A;
  float2* B = s->B;
  if (renormalise)
    block_norms(h, R, tm, s, S, 0);
  BesselTerms bt = evolve_terms(direction, t);
  launch_to_f32(S, A, nR, st);
  launch_block_evo(h->d, R, c.W, tm, 0, nullptr, A, B, nullptr, nullptr, nullptr, nullptr, 0.0,
                   0, st);
  launch_cheb01_32(B, S, bt.b0, bt.b1, direction, nR, st);
  size_t k = 0;
  float2* C = s->C;
  float2* D = s->D;
  if (deep) {
    // kBlockSweepSteps terms per launch on periodic lattices (the same
    // additions in the same order), the rest below, from an even term
    for (; k + kBlockSweepSteps <= bt.bm.size(); k += kBlockSweepSteps) {
      SweepTerms terms{};
      for (int l = 0; l < kBlockSweepSteps; ++l) {
        terms.bm[l] = bt.bm[k + l];
        terms.label[l] = bt.label[k + l];
      }
      launch_block_sweepn_evo(h->d, deep->map, R, A, B, C, D, S, terms, st);
      std::swap(A, C);
      std::swap(B, D);
    }
  }
  for (; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_block_evo(h->d, R, c.W, tm, 1, A, B, A, nullptr, nullptr, nullptr, S, d.bm, d.label,
                     st, d.bm_prev, d.label_prev);
    std::swap(A, B);
  }
  if (renormalise) {
    block_norms(h, R, tm, s, S, 1);
    launch_block_renorm(S, R, tm, s->norms, s->norms + (size_t)segments(h) * R, st);
  }
}

// S <- [X, U(t)] S (src/sigma.cu:157-200)
void block_evolvex(lsqt_ham* h, int R, const BlockShape& c, const TileMap& tm, BlockScratch* s,
                   int direction, double t, double2* S)
{
  hipStream_t st = h->ctx->stream;
  const int nR = (int)((size_t)h->d.n * R);
  float2 *A = s->A, *B = s->B, *Ax = s->F[1], *Bx = s->F[2];
  BesselTerms bt = evolvex_terms(direction, t);
  launch_to_f32(S, A, nR, st);
  launch_zero32(Ax, nR, st);
  launch_block_evo(h->d, R, c.W, tm, 0, nullptr, A, B, nullptr, nullptr, nullptr, nullptr, 0.0,
                   0, st);
  launch_block_evo(h->d, R, c.W, tm, 2, nullptr, A, Bx, nullptr, nullptr, nullptr, nullptr, 0.0,
                   0, st);
  launch_cheb1x_32(Bx, S, bt.b1, nR, st);
  for (size_t k = 0; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_block_evo(h->d, R, c.W, tm, 4, A, B, A, Ax, Bx, Ax, S, d.bm, d.label, st, d.bm_prev,
                     d.label_prev);
    std::swap(A, B);
    std::swap(Ax, Bx);
  }
}

// host vectors (vector-major [R][n]) into a site-major FP64 block, through
// `staging` (n*R double2 of device memory, overwritten)
int upload_states(lsqt_ham* h, int R, const double* re, const double* im, double2* S,
                  double2* staging)
{
  const size_t count = (size_t)h->d.n * R;
  double* d_re = (double*)staging;
  double* d_im = d_re + count;
  hipStream_t st = h->ctx->stream;
  LSQT_HIP(hipStreamSynchronize(st));
  LSQT_HIP(hipMemcpy(d_re, re, sizeof(double) * count, hipMemcpyHostToDevice));
  LSQT_HIP(hipMemcpy(d_im, im, sizeof(double) * count, hipMemcpyHostToDevice));
  launch_site_major(d_re, d_im, h->d.n, R, S, st);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

int check_driver(lsqt_ham* h, int R, const double* re, const double* im, int M, int n_steps,
                 const double* steps, const double* mu_re, const double* mu_im)
{
  if (!h || !re || !im || !steps || !mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null argument");
  if (!valid_R(R))
    return set_error(LSQT_ERR_INVALID, "R = %d must be a power of 2 in 1..256", R);
  if (M < 2 || n_steps < 1)
    return set_error(LSQT_ERR_INVALID, "M = %d (>= 2) and %d time steps (>= 1)", M, n_steps);
  if ((size_t)h->d.n * R > (size_t)INT_MAX)
    return set_error(LSQT_ERR_INVALID, "%d sites x %d vectors exceeds 2^31 values", h->d.n, R);
  if (h->d.layout == kFull)
    return set_error(LSQT_ERR_UNSUPPORTED, "the block path needs the packed layout");
  return ctx_activate(h->ctx);
}

}  // namespace

extern "C" int lsqt_block_vac(lsqt_ham* h, int R, const double* re, const double* im, int M,
                              int n_steps, const double* time_step_scaled, double* mu_re,
                              double* mu_im)
{
  LSQT_TRY(check_driver(h, R, re, im, M, n_steps, time_step_scaled, mu_re, mu_im));
  BlockScratch* s;
  LSQT_TRY(scratch_for(h, R, 2, 1, &s));
  LSQT_TRY(evo_scratch(h, R, s));
  double2* left = s->S[0];   // U(-t) phi
  double2* right = s->S[1];  // U(-t) V phi
  LSQT_TRY(upload_states(h, R, re, im, left, s->S[2]));
  hipStream_t st = h->ctx->stream;
  const int nR = (int)((size_t)h->d.n * R);
  float2* L = s->F[0];
  // the shape is tuned on the random vectors themselves
  launch_to_f32(left, L, nR, st);
  BlockShape c;
  TileMap tm;
  LSQT_TRY(block_setup(h, R, L, L, M, &c, &tm, &s));
  // the moment sets <state_right|T_m|state_left_copy> as pairs of
  // recursions by the product identities where 2R columns fit
  BlockShape c2;
  TileMap tm2;
  bool pairs = false;
  LSQT_TRY(pair_setup(h, R, c, L, L, M, &c2, &tm2, &s, &pairs));
  LSQT_TRY(evo_scratch(h, R, s));
  // state_right = V state_left (src/sigma.cu:293)
  launch_to_f32(left, s->B, nR, st);
  launch_block_evo(h->d, R, c.W, tm, 3, nullptr, s->B, s->A, nullptr, nullptr, nullptr, nullptr,
                   0.0, 0, st);
  launch_to_f64(s->A, right, nR, st);
  const size_t per_step = (size_t)M * R * segments(h);
  for (int m = 0; m < n_steps; ++m) {
    // state_left_copy = V state_left; moments <state_right|T|state_left_copy>
    float2* copy = pairs ? s->F[1] : s->A;
    launch_to_f32(left, s->B, nR, st);
    launch_block_evo(h->d, R, c.W, tm, 3, nullptr, s->B, copy, nullptr, nullptr, nullptr,
                     nullptr, 0.0, 0, st);
    launch_to_f32(right, L, nR, st);
    if (pairs)
      LSQT_TRY(moment_set(h, R, c2, tm2, s, L, copy, kPairProducts, M, mu_re + m * per_step,
                          mu_im + m * per_step));
    else
      LSQT_TRY(moment_set(h, R, c, tm, s, L, copy, kDirect, M, mu_re + m * per_step,
                          mu_im + m * per_step));
    if (m < n_steps - 1) {
      block_evolve(h, R, c, tm, s, -1, time_step_scaled[m], left, true);
      block_evolve(h, R, c, tm, s, -1, time_step_scaled[m], right, true);
    }
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_block_msd(lsqt_ham* h, int R, const double* re, const double* im, int M,
                              int n_steps, const double* time_step_scaled, double* mu_re,
                              double* mu_im)
{
  LSQT_TRY(check_driver(h, R, re, im, M, n_steps, time_step_scaled, mu_re, mu_im));
  BlockScratch* s;
  LSQT_TRY(scratch_for(h, R, 2, 1, &s));
  LSQT_TRY(evo_scratch(h, R, s));
  double2* state = s->S[0];
  double2* state_x = s->S[1];
  double2* state_copy = s->S[2];
  LSQT_TRY(upload_states(h, R, re, im, state, s->S[2]));
  hipStream_t st = h->ctx->stream;
  const int nR = (int)((size_t)h->d.n * R);
  launch_to_f32(state, s->F[0], nR, st);
  BlockShape c;
  TileMap tm;
  LSQT_TRY(block_setup(h, R, s->F[0], s->F[0], M, &c, &tm, &s));
  LSQT_TRY(evo_scratch(h, R, s));
  launch_copy(state_x, state, nR, st);
  block_evolve(h, R, c, tm, s, 1, time_step_scaled[0], state, true);
  block_evolvex(h, R, c, tm, s, 1, time_step_scaled[0], state_x);
  const size_t per_step = (size_t)M * R * segments(h);
  for (int m = 0; m < n_steps; ++m) {
    launch_to_f32(state_x, s->A, nR, st);
    LSQT_TRY(moment_set(h, R, c, tm, s, nullptr, s->A, kSelfProducts, M, mu_re + m * per_step,
                        mu_im + m * per_step));
    if (m < n_steps - 1) {
      const double t = time_step_scaled[m + 1];
      launch_copy(state_copy, state, nR, st);
      block_evolvex(h, R, c, tm, s, 1, t, state_copy);
      block_evolve(h, R, c, tm, s, 1, t, state_x, true);
      launch_add(state_x, state_copy, nR, st);
      block_evolve(h, R, c, tm, s, 1, t, state, true);
    }
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_block_spin(lsqt_ham* h, int R, const double* re, const double* im, int M,
                               int n_steps, const double* time_step_scaled, double* mu_re,
                               double* mu_im)
{
  LSQT_TRY(check_driver(h, R, re, im, M, n_steps, time_step_scaled, mu_re, mu_im));
  if (!h->seg_starts.empty())
    return set_error(LSQT_ERR_UNSUPPORTED, "the spin driver takes one system (S_z alternates "
                     "over the sites of a system)");
  BlockScratch* s;
  LSQT_TRY(scratch_for(h, R, 2, 1, &s));
  LSQT_TRY(evo_scratch(h, R, s));
  double2* state = s->S[0];
  LSQT_TRY(upload_states(h, R, re, im, state, s->S[2]));
  hipStream_t st = h->ctx->stream;
  const int nR = (int)((size_t)h->d.n * R);
  float2* L = s->F[0];
  launch_to_f32(state, L, nR, st);
  BlockShape c;
  TileMap tm;
  LSQT_TRY(block_setup(h, R, L, L, M, &c, &tm, &s));
  BlockShape c2;
  TileMap tm2;
  bool pairs = false;
  LSQT_TRY(pair_setup(h, R, c, L, L, M, &c2, &tm2, &s, &pairs));
  LSQT_TRY(evo_scratch(h, R, s));
  const size_t per_step = (size_t)M * R;
  for (int m = 0; m < n_steps; ++m) {
    // <state|T_m|S_z state> (src/sigma.cu:395-442)
    float2* sz = pairs ? s->F[1] : s->A;
    launch_to_f32(state, L, nR, st);
    launch_block_sz(L, h->d.n, R, sz, st);
    if (pairs)
      LSQT_TRY(moment_set(h, R, c2, tm2, s, L, sz, kPairProducts, M, mu_re + m * per_step,
                          mu_im + m * per_step));
    else
      LSQT_TRY(moment_set(h, R, c, tm, s, L, sz, kDirect, M, mu_re + m * per_step,
                          mu_im + m * per_step));
    if (m < n_steps - 1)
      block_evolve(h, R, c, tm, s, 1, time_step_scaled[m], state, true);
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_block_apply(lsqt_ham* h, int op, const lsqt_block* in, lsqt_block* out)
{
  LSQT_TRY(check_block(h, in, out));
  if (op < 0 || op > 2)
    return set_error(LSQT_ERR_INVALID, "operator %d is not 0 (H), 1 ([X,H]) or 2 (V)", op);
  if (in == out)
    return set_error(LSQT_ERR_INVALID, "input and output blocks must differ");
  BlockShape c;
  TileMap tm;
  BlockScratch* s;
  LSQT_TRY(block_setup(h, in->R, in->d, in->d, 2, &c, &tm, &s));
  static const int kOp[3] = {0, 2, 3};  // k_block_evo operators: H~, [X, H~], V
  launch_block_evo(h->d, in->R, c.W, tm, kOp[op], nullptr, in->d, out->d, nullptr, nullptr,
                   nullptr, nullptr, 0.0, 0, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_block_set_graphs(lsqt_ham* h, int enable)
{
  if (!h)
    return set_error(LSQT_ERR_INVALID, "null Hamiltonian");
  if (!h->blk)
    h->blk = new lsqt_block_state();
  h->blk->graphs = enable != 0;
  return LSQT_OK;
}

extern "C" int lsqt_block_tune(lsqt_ham* h, const lsqt_block* left, const lsqt_block* right,
                               int force, int* W, int* rows, int* blocked, int* depth,
                               double* us_per_step)
{
  LSQT_TRY(check_block(h, left, right));
  LSQT_TRY(f32_potential(h));
  BlockShape c;
  double us;
  LSQT_TRY(pick_config(h, left->R, left->d, right->d, &c, &us, force != 0));
  if (W)
    *W = c.W;
  if (rows)
    *rows = c.rows;
  if (blocked)
    *blocked = c.blocked;
  if (depth)
    *depth = c.depth;
  if (us_per_step)
    *us_per_step = us;
  return LSQT_OK;
}

namespace {

int pick_config(lsqt_ham* h, int R, const float2* left, const float2* right, BlockShape* out,
                double* us_out, bool force, int mode)
{
  const bool pairs = mode == kTunePairs;
  if (h->blk && !pairs) {
    auto fit = h->blk->forced.find(R);
    if (fit != h->blk->forced.end()) {
      *out = fit->second;
      *us_out = 0.0;
      return LSQT_OK;
    }
  }
  std::lock_guard<std::mutex> lock(g_tune_mutex);
  tune_load();
  // the slab of a detected lattice enters the key: blocked shapes exist only
  // for lattices, and their best tile size depends on the slab
  TuneKey key{R, h->d.K, h->d.layout == kPackedLds ? 1 : 0, size_class(h->d.n), h->slab, mode};
  auto it = g_tune.find(key);
  // a cached shape is taken only if it is valid here (the cache file is
  // plain text, and the key holds the slab, not the row length that a
  // window shape also depends on)
  if (it != g_tune.end() && !force && shape_usable(h, R, it->second.shape) &&
      (!pairs || it->second.shape.W >= 2)) {
    *out = it->second.shape;
    *us_out = it->second.us;
    return LSQT_OK;
  }
  // candidates: natural tiles of 256..4096 sites; on a detected lattice also
  // blocked tiles of a whole number of rows that divide the slab.  Tiles of
  // 64 and 128 sites (one to four passes) are tried only where the ring of
  // moment partials they need stays within 4 GB.
  std::vector<BlockShape> cands;
  for (int W : {1, 2, 4, 8}) {
    if (!valid_W(R, W) || (pairs && W < 2))
      continue;
    for (int rows : {64, 128, 256, 1024, 4096}) {
      if (rows < kBlock / (R / W))
        continue;
      const double ring_bytes = (double)kRing * ((double)h->d.n / rows) * R * sizeof(double2);
      if (rows < 256 && ring_bytes > 4.0e9)
        continue;
      BlockShape c;
      c.W = W;
      c.rows = rows;
      cands.push_back(c);
    }
    if (h->slab > 0) {
      for (int span = h->row; span <= 8192 && span <= h->slab; span += h->row) {
        if (h->slab % span != 0 || span % 64 != 0 || span < 512 || span < kBlock / (R / W))
          continue;
        BlockShape c;
        c.W = W;
        c.rows = span;
        c.blocked = 1;
        cands.push_back(c);
        if (h->row % (kBlock / (R / W)) == 0) {
          c.blocked = 2;
          cands.push_back(c);
        }
      }
    }
  }
  // every shape also with the double-buffered loop (compile-time slot counts)
  if (h->d.K == 3 || h->d.K == 4 || h->d.K == 6) {
    const size_t single = cands.size();
    for (size_t i = 0; i < single; ++i) {
      BlockShape c = cands[i];
      c.depth = 2;
      cands.push_back(c);
    }
    // the sliding window: tiles of about 8, 16 and 32 lattice rows
    if (h->slab > 0 && h->row > 0 && h->slab % h->row == 0) {
      const int ny = h->slab / h->row;
      for (int W : {1, 2, 4}) {
        if (!valid_W(R, W) || (pairs && W < 2))
          continue;
        int last = 0;
        for (int want : {8, 16, 32}) {
          int by = 0;  // the divisor of ny closest to `want` from below
          for (int d = want; d >= 4 && by == 0; --d)
            if (ny % d == 0)
              by = d;
          if (by == 0 || by == last || !window_fits(h->d.n, h->slab, h->row, by, R, W))
            continue;
          last = by;
          BlockShape c;
          c.W = W;
          c.rows = by;
          c.blocked = 3;
          cands.push_back(c);
        }
      }
    }
  }
  const int steps = 6;
  int max_tiles = 0;
  for (const BlockShape& c : cands) {
    TileMap tm;
    LSQT_TRY(tile_map(h, R, c, &tm));
    max_tiles = std::max(max_tiles, tm.tiles);
  }
  BlockScratch* s;
  // moment outputs for the 2 * steps moments of a pair set
  LSQT_TRY(scratch_for(h, R, 2 * steps, max_tiles, &s));
  hipStream_t st = h->ctx->stream;
  hipEvent_t e0, e1;
  LSQT_HIP(hipEventCreate(&e0));
  LSQT_HIP(hipEventCreate(&e1));
  TuneVal best{BlockShape{}, 1e300};
  for (const BlockShape& c : cands) {
    // time a few fused steps on the caller's blocks, twice (the better of
    // the two counts), after an untimed warm-up: the first launches of a
    // kernel shape also load its code
    TileMap tm;
    LSQT_TRY(tile_map(h, R, c, &tm));
    double us = 1e300;
    for (int rep = 0; rep < 3; ++rep) {
      float2* A = s->A;
      float2* B = s->B;
      LSQT_HIP(hipMemcpyAsync(A, right, sizeof(float2) * (size_t)h->d.n * R,
                              hipMemcpyDeviceToDevice, st));
      const int n_steps = rep == 0 ? 2 : steps;
      if (pairs) {
        // a pair moment set of 2 n_steps moments
        LSQT_HIP(hipEventRecord(e0, st));
        enqueue_self_steps(h, R / 2, c, tm, A, B, s->ring, s->mu, 2 * n_steps, st, true);
        LSQT_HIP(hipEventRecord(e1, st));
      } else {
        enqueue_steps(h, R, c, tm, left, A, B, s->ring, nullptr, n_steps, 0, 2, st);
        LSQT_HIP(hipEventRecord(e0, st));
        enqueue_steps(h, R, c, tm, left, A, B, s->ring, nullptr, n_steps, 2, n_steps, st);
        LSQT_HIP(hipEventRecord(e1, st));
      }
      LSQT_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      LSQT_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (rep > 0)
        us = std::min(us, 1e3 * ms / (pairs ? steps : steps - 2));
    }
    if (us < best.us)
      best = TuneVal{c, us};
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  LSQT_HIP(hipGetLastError());
  tune_store(key, best);
  *out = best.shape;
  *us_out = best.us;
  return LSQT_OK;
}

}  // namespace
