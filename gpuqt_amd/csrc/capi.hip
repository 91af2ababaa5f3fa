// capi.hip -- the C ABI of liblsqt_hip.so (include/lsqt_hip.h): contexts,
// Hamiltonian upload and layout selection, vectors, and the fused moment /
// evolution drivers that keep the whole recursion on the device.
#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <unordered_map>
#include <vector>

#include "bessel.h"
#include "lsqt_capi_internal.h"
#include "lsqt_hip.h"
#include "lsqt_launch.h"

using namespace lsqt;

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[1024];

int lsqt::set_error(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char* lsqt_last_error(void)
{
  return g_err;
}

extern "C" const char* lsqt_strerror(int status)
{
  switch (status) {
    case LSQT_OK:
      return "ok";
    case LSQT_ERR_INVALID:
      return "invalid argument";
    case LSQT_ERR_HIP:
      return "HIP runtime error";
    case LSQT_ERR_NOMEM:
      return "out of memory";
    case LSQT_ERR_NODEVICE:
      return "no usable GPU";
    case LSQT_ERR_UNSUPPORTED:
      return "unsupported";
    default:
      return "unknown status";
  }
}

extern "C" const char* lsqt_version(void)
{
  return "lsqt_hip 0.1 gfx950";
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
int lsqt::ctx_activate(lsqt_ctx* ctx)
{
  LSQT_HIP(hipSetDevice(ctx->device));
  return LSQT_OK;
}

int lsqt::ensure_device(void** p, size_t* cap, size_t bytes)
{
  if (*cap >= bytes)
    return LSQT_OK;
  if (*p)
    LSQT_HIP(hipFree(*p));
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  }
  *cap = bytes;
  return LSQT_OK;
}

int lsqt::ensure_pinned(void** p, size_t* cap, size_t bytes)
{
  if (*cap >= bytes)
    return LSQT_OK;
  if (*p)
    LSQT_HIP(hipHostFree(*p));
  *p = nullptr;
  *cap = 0;
  hipError_t e = hipHostMalloc(p, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    *p = nullptr;
    return set_error(LSQT_ERR_NOMEM, "hipHostMalloc(%zu bytes): %s", bytes,
                     hipGetErrorString(e));
  }
  *cap = bytes;
  return LSQT_OK;
}

extern "C" int lsqt_device_count(int* count)
{
  if (!count)
    return set_error(LSQT_ERR_INVALID, "null count");
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) {
    *count = 0;
    return set_error(LSQT_ERR_NODEVICE, "hipGetDeviceCount: %s", hipGetErrorString(e));
  }
  return LSQT_OK;
}

extern "C" int lsqt_ctx_create(int device, lsqt_ctx** out)
{
  if (!out)
    return set_error(LSQT_ERR_INVALID, "null output handle");
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
    return set_error(LSQT_ERR_NODEVICE, "no HIP device visible");
  if (device < 0 || device >= count)
    return set_error(LSQT_ERR_INVALID, "device %d out of range (%d devices)", device, count);
  hipDeviceProp_t prop;
  LSQT_HIP(hipGetDeviceProperties(&prop, device));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    return set_error(LSQT_ERR_NODEVICE, "device %d is %s; this library is built for gfx950",
                     device, prop.gcnArchName);
  LSQT_HIP(hipSetDevice(device));
  noalloc_setup(device);
  lsqt_ctx* c = new lsqt_ctx();
  c->device = device;
  c->cu_count = prop.multiProcessorCount;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return set_error(LSQT_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  *out = c;
  return LSQT_OK;
}

extern "C" int lsqt_ctx_destroy(lsqt_ctx* ctx)
{
  if (!ctx)
    return LSQT_OK;
  if (ctx->live > 0)
    return set_error(LSQT_ERR_INVALID, "%d Hamiltonians, vectors or blocks still use this "
                     "context; destroy them first", ctx->live);
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  if (ctx->partial)
    (void)hipFree(ctx->partial);
  if (ctx->mu_dev)
    (void)hipFree(ctx->mu_dev);
  if (ctx->mu_host)
    (void)hipHostFree(ctx->mu_host);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LSQT_OK;
}

extern "C" int lsqt_ctx_sync(lsqt_ctx* ctx)
{
  if (!ctx)
    return set_error(LSQT_ERR_INVALID, "null context");
  LSQT_TRY(ctx_activate(ctx));
  LSQT_HIP(hipStreamSynchronize(ctx->stream));
  return LSQT_OK;
}

extern "C" void* lsqt_ctx_stream(lsqt_ctx* ctx)
{
  return ctx ? (void*)ctx->stream : nullptr;
}

// ---------------------------------------------------------------------------
// Hamiltonian
// ---------------------------------------------------------------------------
namespace {

struct TripleKey {
  uint64_t a, b, c;
  bool operator==(const TripleKey& o) const { return a == o.a && b == o.b && c == o.c; }
};

struct TripleHash {
  size_t operator()(const TripleKey& k) const
  {
    uint64_t h = k.a * 0x9E3779B97F4A7C15ull;
    h ^= (k.b + 0x632BE59BD9B4E019ull) + (h << 6) + (h >> 2);
    h ^= (k.c + 0x85EBCA77C2B2AE63ull) + (h << 6) + (h >> 2);
    return (size_t)h;
  }
};

uint64_t bits_of(double x)
{
  uint64_t u;
  std::memcpy(&u, &x, 8);
  return u;
}

int index_bits_for(int n)
{
  int bits = 1;
  while (bits < 31 && (1u << bits) < (uint32_t)n)
    ++bits;
  return bits;
}

template <typename T>
int upload(lsqt_ham* h, const T** dst, const T* src, size_t count)
{
  void* p = nullptr;
  size_t bytes = sizeof(T) * (count ? count : 1);
  hipError_t e = hipMalloc(&p, bytes);
  if (e != hipSuccess)
    return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
  h->allocations.push_back(p);
  if (count) {
    e = hipMemcpy(p, src, sizeof(T) * count, hipMemcpyHostToDevice);
    if (e != hipSuccess)
      return set_error(LSQT_ERR_HIP, "hipMemcpy H2D: %s", hipGetErrorString(e));
  }
  *dst = (const T*)p;
  return LSQT_OK;
}

// Finds a 3-level natural order (slabs of `slab` sites made of rows of `row`
// sites) from the neighbour offsets |j - i|: the slab is the largest offset
// carried by a regular fraction of the pairs (wrap-around offsets above n/2
// excluded) that divides n, the row the largest smaller regular offset that
// divides the slab.  Lattices numbered x-slowest (src/model_lattice.cu:162)
// and the simple-cubic general inputs have this form; anything else gets 0.
// Returns the regular offsets (ascending), the candidates of the two-step
// sweep's slab.
std::vector<long long> detect_lattice_offsets(lsqt_ham* h, int n, int K, const int* nn,
                                              const int* nl)
{
  std::unordered_map<long long, long long> count;
  long long pairs = 0;
  for (int i = 0; i < n; ++i)
    for (int m = 0; m < nn[i]; ++m) {
      long long d = std::llabs((long long)nl[(size_t)i * K + m] - i);
      if (d == 0 || 2 * d > n)
        continue;
      ++count[d];
      ++pairs;
      if (count.size() > 4096)
        return {};  // irregular structure
    }
  if (pairs == 0)
    return {};
  std::vector<long long> regular;
  for (auto& kv : count)
    if (kv.second * 8 * K >= pairs)
      regular.push_back(kv.first);
  std::sort(regular.begin(), regular.end());
  for (int a = (int)regular.size() - 1; a >= 0; --a) {
    long long D = regular[a];
    if (n % D != 0 || D < 64)
      continue;
    for (int b = a - 1; b >= 0; --b) {
      long long d = regular[b];
      if (D % d == 0 && d >= 8 && D / d >= 4) {
        h->slab = (int)D;
        h->row = (int)d;
        return regular;
      }
    }
  }
  return regular;
}

// Launch shape of the two-step sweep for one storage precision: runs of
// slabs per chunk so that the tiles fill the device once at the occupancy
// the LDS rows allow (LSQT_SWEEP_TILES=<count> in the environment sets the
// tile count instead), at least two slabs per run.  The shape depends on the
// geometry only, not on the class table or slot pattern, so a Hamiltonian
// built with and without its pattern sums the same tiles.
SweepMap sweep_shape(SweepMap g, bool f32)
{
  const size_t vb = f32 ? sizeof(float2) : sizeof(double2);
  const size_t lds = (vb * 2 * (kSweepBlock + 2) + 8 + vb) * kSweepThreads;
  const int per_cu = (int)std::max<size_t>(1, std::min<size_t>(8, kLdsPerCu / lds));
  long long target = (long long)kComputeUnits * per_cu;
  if (const char* env = getenv("LSQT_SWEEP_TILES"))
    if (atoll(env) > 0)
      target = atoll(env);
  // at most one tile per resident workgroup: a second, partial wave of
  // tiles would take as long as the first
  long long runs = target / g.nch;
  runs = std::max(1ll, std::min(runs, (long long)(g.nslab / 2)));
  g.nruns = (int)runs;
  g.tiles = g.nch * g.nruns;
  g.blocks = (g.tiles + kXcds - 1) / kXcds * kXcds;
  return g;
}

// The deeper sweep's shape for the same geometry: ns steps per launch in
// blocks of nb slabs; a chunk holds 2 ns h fewer positions than a workgroup
// (none where that leaves fewer than half), runs as sweep_shape.
SweepMap sweep_shape_n(SweepMap g, bool f32, int ns, int nb)
{
  // the written levels go through buffer resources over one slab
  if ((long long)g.S * sizeof(double2) >= (1ll << 31)) {
    g.S = 0;
    return g;
  }
  const bool whole = g.hh == 0 && g.nch == 1 && g.cs == g.S;
  if (!whole) {
    g.cs = kSweepThreads - 2 * ns * g.hh;
    if (g.cs < kSweepThreads / 2) {
      g.S = 0;
      return g;
    }
    g.nch = (g.S + g.cs - 1) / g.cs;
  }
  g.wmax = g.cs + 2 * ns * g.hh;
  g.ns = ns;
  g.nb = nb;
  const int per_cu =
    (int)std::max<size_t>(1, std::min<size_t>(8, kLdsPerCu / sweepn_lds(f32, ns, nb)));
  long long target = (long long)kComputeUnits * per_cu;
  if (const char* env = getenv("LSQT_SWEEP_TILES"))
    if (atoll(env) > 0)
      target = atoll(env);
  // at most one tile per resident workgroup: a second, partial wave of
  // tiles would take as long as the first
  long long runs = target / g.nch;
  runs = std::max(1ll, std::min(runs, (long long)(g.nslab / 2)));
  g.nruns = (int)runs;
  g.tiles = g.nch * g.nruns;
  g.blocks = (g.tiles + kXcds - 1) / kXcds * kXcds;
  return g;
}

bool find_slot_pattern(const HamDev& d, int* period_out, std::vector<int2>* table_out);

// The two-step sweep (k_sweep2) of a packed Hamiltonian with compile-time
// slot count: from the per-residue slot pattern (period p), for slab sizes S
// near each candidate offset that are multiples of p and divide n into at
// least 4 slabs, every pattern offset is split into a slab offset (-1, 0 or
// 1) and a position offset (at most kSweepMaxHalo, unless a whole slab fits
// one workgroup), and a device pass checks every packed word against the
// periodic table so obtained; the first S that passes is kept.
// LSQT_SWEEP=0 in the environment turns the sweep off.  Any failure leaves
// the Hamiltonian without it (it is an optimisation only).
void detect_sweep(lsqt_ham* h, const std::vector<long long>& candidates)
{
  h->sweep64 = SweepMap{};
  h->sweep32 = SweepMap{};
  h->sweepn64 = SweepMap{};
  h->sweepn32 = SweepMap{};
  h->bsweep.clear();
  h->bsweep_tab = nullptr;
  const HamDev& d = h->d;
  const char* env = getenv("LSQT_SWEEP");
  if ((env && env[0] == '0') || d.layout == kFull || d.n < kSweepMinSites)
    return;
  if (d.K != 2 && d.K != 3 && d.K != 4 && d.K != 6)
    return;
  int period = 0;
  std::vector<int2> pattern;
  if (!find_slot_pattern(d, &period, &pattern))
    return;
  const int n = d.n;
  const int blocks = sweep_check_blocks(n);
  int* dev_bad = nullptr;
  int4* dev_tab = nullptr;
  if (hipMalloc(&dev_bad, sizeof(int) * (size_t)blocks) != hipSuccess)
    return;
  if (hipMalloc(&dev_tab, sizeof(int4) * pattern.size()) != hipSuccess) {
    (void)hipFree(dev_bad);
    return;
  }
  std::vector<int> bad(blocks);
  std::vector<double2> classes(d.ncls);
  if (hipMemcpy(classes.data(), d.tab, sizeof(double2) * classes.size(),
                hipMemcpyDeviceToHost) != hipSuccess)
    classes.clear();
  std::vector<long long> tried;
  bool found = false;
  for (auto it = candidates.rbegin(); it != candidates.rend() && !found; ++it) {
    for (long long S = *it - kSweepMaxHalo - 1; S <= *it + kSweepMaxHalo + 1 && !found; ++S) {
      if (S < 128 || S % period != 0 || n % S != 0 || n / S < 4)
        continue;
      if (std::find(tried.begin(), tried.end(), S) != tried.end())
        continue;
      tried.push_back(S);
      const bool whole = S <= kSweepThreads;
      std::vector<int4> tab(pattern.size());
      int halo = 0;
      bool ok = true;
      for (size_t e = 0; e < pattern.size() && ok; ++e) {
        const int off = pattern[e].x;
        if (off == kPatternEmpty) {
          tab[e] = make_int4(0, 0, 0, 0);
          continue;
        }
        // off = ds * S + dp with |dp| <= S / 2
        long long ds = off >= 0 ? (off + S / 2) / S : -((-off + S / 2) / S);
        const long long dp = off - ds * S;
        if (ds < -1 || ds > 1 || (!whole && std::llabs(dp) > kSweepMaxHalo))
          ok = false;
        halo = std::max(halo, (int)std::llabs(dp));
        tab[e] = make_int4((int)ds, (int)dp, pattern[e].y, 1);
      }
      if (!ok)
        continue;
      SweepMap g{};
      g.S = (int)S;
      g.nslab = (int)(n / S);
      g.sdiv = make_divisor((int)S);
      g.p = period;
      g.tab = dev_tab;
      g.halo = halo;
      if (whole) {
        g.hh = 0;
        g.nch = 1;
        g.cs = (int)S;
      } else {
        g.hh = halo;
        g.cs = kSweepThreads - 4 * halo;  // the held positions fit one workgroup
        g.nch = (int)((S + g.cs - 1) / g.cs);
      }
      g.wmax = g.cs + 4 * g.hh;
      g.real = 1;
      g.full = 1;
      for (size_t e = 0; e < tab.size(); ++e) {
        if (tab[e].w && (tab[e].z >= (int)classes.size() || classes[tab[e].z].y != 0.0))
          g.real = 0;
        if (!tab[e].w)
          g.full = 0;
      }
      if (hipMemcpy(dev_tab, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice) !=
          hipSuccess)
        break;
      launch_sweep_check(d, g, dev_bad, h->ctx->stream);
      if (hipMemcpyAsync(bad.data(), dev_bad, sizeof(int) * bad.size(), hipMemcpyDeviceToHost,
                         h->ctx->stream) != hipSuccess ||
          hipStreamSynchronize(h->ctx->stream) != hipSuccess)
        break;
      if (std::any_of(bad.begin(), bad.end(), [](int b) { return b != 0; }))
        continue;
      h->allocations.push_back(dev_tab);  // kept with the Hamiltonian
      h->sweep64 = sweep_shape(g, false);
      h->sweep32 = sweep_shape(g, true);
      // more steps per launch on the same geometry (LSQT_SWEEP_STEPS=2 in
      // the environment keeps two): six steps, one slab per loop iteration,
      // FP64 and FP32 states (measured on the 8.0M-site honeycomb: 32.7 and
      // 18.3 us per step against 74.4 and 38.3 with two steps per launch;
      // FP64 four, five, seven and eight steps 39.8, 34.3, 40.3, 40.2; FP32
      // six steps in blocks of two and four slabs 18.4 and 19.9, seven and
      // eight steps 18.9)
      const char* steps = getenv("LSQT_SWEEP_STEPS");
      if (!(steps && atoi(steps) == 2)) {
        h->sweepn64 = sweep_shape_n(h->sweep64, false, 6, 1);
        h->sweepn32 = sweep_shape_n(h->sweep32, true, 6, 1);
      }
      // the same geometry serves the block path's sweeps
      h->bsweep_steps = (steps && atoi(steps) == 2) ? 2 : kBlockSweepSteps;
      BlockSweepSeg b{};
      b.S = g.S;
      b.nslab = g.nslab;
      b.halo = halo;
      b.p = period;
      h->bsweep.assign(1, b);
      h->bsweep_tab = dev_tab;
      found = true;
    }
  }
  (void)hipGetLastError();
  (void)hipFree(dev_bad);
  if (!found)
    (void)hipFree(dev_tab);
}

// The block path's two-step sweep of several pristine lattices side by side
// (a segmented Hamiltonian built by lsqt_ham_create_lattices): lattice k's
// slabs are its x cells (S = Ny Nz orbitals sites), its table period the
// orbital count, and hopping i of orbital m goes to slab offset dx and
// position offset (dy Nz + dz) orbitals + m' - m, taken periodically within
// the slab.  A device pass checks every packed word of every lattice against
// its table; the sweep is kept only if all lattices pass (LSQT_SWEEP=0 in
// the environment turns it off).  Any failure leaves the Hamiltonian without
// it.
void detect_block_sweep(lsqt_ham* h, int segments, const lsqt_lattice* lats,
                        const int* const* site_index, const std::vector<int>& starts,
                        const std::vector<std::vector<int>>& cls)
{
  h->bsweep.clear();
  h->bsweep_tab = nullptr;
  const HamDev& d = h->d;
  const char* env = getenv("LSQT_SWEEP");
  if ((env && env[0] == '0') || d.layout == kFull)
    return;
  const int K = d.K;
  std::vector<BlockSweepSeg> segs;
  std::vector<int4> tab;
  for (int s = 0; s < segments; ++s) {
    const lsqt_lattice* L = &lats[s];
    if (site_index && site_index[s])
      return;  // vacancies: no periodic table
    const long long S = (long long)L->cells[1] * L->cells[2] * L->orbitals;
    const int No = L->orbitals, Ks = L->max_neighbor, Nz = L->cells[2];
    if (L->cells[0] < 4 || S < 8 || S > INT_MAX / 4)
      return;
    BlockSweepSeg g{};
    g.base = starts[s];
    g.S = (int)S;
    g.nslab = L->cells[0];
    g.p = No;
    g.tab = (int)tab.size();
    tab.resize(tab.size() + (size_t)K * No, make_int4(0, 0, 0, 0));
    for (int m = 0; m < No; ++m)
      for (int i = 0; i < L->hop_count[m]; ++i) {
        const int* c = L->hop_cell + 3 * (m * Ks + i);
        if (c[0] < -1 || c[0] > 1)
          return;
        long long dp = ((long long)c[1] * Nz + c[2]) * No + L->hop_orbital[m * Ks + i] - m;
        dp %= S;
        if (dp < 0)
          dp += S;
        if (2 * dp > S)
          dp -= S;
        g.halo = std::max(g.halo, (int)std::llabs(dp));
        tab[(size_t)g.tab + (size_t)i * No + m] = make_int4(c[0], (int)dp, cls[s][m * Ks + i], 1);
      }
    segs.push_back(g);
  }
  int4* dev_tab = nullptr;
  int* dev_bad = nullptr;
  if (hipMalloc(&dev_tab, sizeof(int4) * tab.size()) != hipSuccess)
    return;
  bool ok = hipMemcpy(dev_tab, tab.data(), sizeof(int4) * tab.size(), hipMemcpyHostToDevice) ==
            hipSuccess;
  int most = 0;
  for (const BlockSweepSeg& g : segs)
    most = std::max(most, block_sweep_check_blocks(g));
  ok = ok && hipMalloc(&dev_bad, sizeof(int) * (size_t)most * segs.size()) == hipSuccess;
  std::vector<int> bad((size_t)most * segs.size(), 0);
  if (ok) {
    for (size_t s = 0; s < segs.size(); ++s)
      launch_block_sweep_check(d, segs[s], dev_tab, dev_bad + s * most, h->ctx->stream);
    for (size_t s = 0; s < segs.size(); ++s)
      ok = ok && hipMemcpyAsync(bad.data() + s * most, dev_bad + s * most,
                                sizeof(int) * (size_t)block_sweep_check_blocks(segs[s]),
                                hipMemcpyDeviceToHost, h->ctx->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->ctx->stream) == hipSuccess;
    ok = ok && std::none_of(bad.begin(), bad.end(), [](int b) { return b != 0; });
  }
  (void)hipGetLastError();
  if (dev_bad)
    (void)hipFree(dev_bad);
  if (!ok) {
    (void)hipFree(dev_tab);
    return;
  }
  h->allocations.push_back(dev_tab);
  h->bsweep = segs;
  h->bsweep_tab = dev_tab;
  const char* steps = getenv("LSQT_SWEEP_STEPS");
  h->bsweep_steps = (steps && atoi(steps) == 2) ? 2 : kBlockSweepSteps;
}

// The slot pattern of a packed Hamiltonian.  On a periodic lattice whose
// orbital index runs fastest, slot m of site n holds neighbour n + off and
// class c that depend only on n mod p, except at the boundaries.  The period
// is the smallest p <= 64 (with p x slots <= 512 table entries) whose most
// common entries per residue describe at least 90 % of 4096 sites from the
// middle of the system; a device pass then marks every site whose words
// differ from the pattern in any slot, and the pattern is kept where at most
// 10 % of the sites are so marked.  The kernels that use it compute the words
// of the unmarked sites, which equal the stored ones bit for bit, and load
// the marked sites' words.  LSQT_PATTERN=0 in the environment turns it off.
// Any failure leaves the Hamiltonian without a pattern (it is an
// optimisation only).
// The per-residue slot pattern of a packed Hamiltonian's words, sampled on
// 4096 sites from the middle of the system: the smallest period p <= 64
// (p x slots <= 512 entries) whose most common (offset, class) entries per
// residue describe at least 90 % of the sample; table[m * p + r] = (offset,
// class) of slot m at residue r (offset kPatternEmpty: unused).  False where
// there is none.
bool find_slot_pattern(const HamDev& d, int* period_out, std::vector<int2>* table_out)
{
  if (d.layout == kFull || d.n < 4096)
    return false;
  const int K = d.K;
  constexpr int kSample = 4096;
  const int s0 = (d.n / 2 - kSample / 2) / 64 * 64;
  std::vector<uint32_t> sample((size_t)K * kSample);
  for (int m = 0; m < K; ++m)
    if (hipMemcpy(sample.data() + (size_t)m * kSample, d.words + (size_t)m * d.stride + s0,
                  sizeof(uint32_t) * kSample, hipMemcpyDeviceToHost) != hipSuccess)
      return false;
  // (offset, class) of sample site i in slot m, as two 32-bit halves
  auto entry = [&](int i, int m) -> unsigned long long {
    const uint32_t w = sample[(size_t)m * kSample + i];
    const int off = w == kEmptySlot ? kPatternEmpty : (int)(w & d.idx_mask) - (s0 + i);
    const uint32_t cls = w == kEmptySlot ? 0u : w >> d.idx_bits;
    return ((unsigned long long)(uint32_t)off << 32) | cls;
  };
  int period = 0;
  std::vector<int2> table;
  for (int p = 1; p <= kPatternMaxPeriod && p * K <= kPatternMaxEntries && !period; ++p) {
    std::vector<int2> tab((size_t)p * K);
    long long matched = 0;
    for (int r = 0; r < p; ++r) {
      std::map<std::vector<unsigned long long>, int> counts;
      for (int i = (r - s0 % p + p) % p; i < kSample; i += p) {
        std::vector<unsigned long long> key(K);
        for (int m = 0; m < K; ++m)
          key[m] = entry(i, m);
        ++counts[key];
      }
      auto best = counts.begin();
      for (auto it = counts.begin(); it != counts.end(); ++it)
        if (it->second > best->second)
          best = it;
      if (best == counts.end())
        continue;
      matched += best->second;
      for (int m = 0; m < K; ++m)
        tab[(size_t)m * p + r] = make_int2((int)(uint32_t)(best->first[m] >> 32),
                                           (int)(uint32_t)best->first[m]);
    }
    if (matched * 10 >= 9ll * kSample) {
      period = p;
      table.swap(tab);
    }
  }
  if (!period)
    return false;
  *period_out = period;
  table_out->swap(table);
  return true;
}

void detect_slot_pattern(lsqt_ham* h)
{
  HamDev& d = h->d;
  const char* env = getenv("LSQT_PATTERN");
  if (env && env[0] == '0')
    return;
  int period = 0;
  std::vector<int2> table;
  if (!find_slot_pattern(d, &period, &table))
    return;
  const int2* pat = nullptr;
  if (upload(h, &pat, table.data(), table.size()) != LSQT_OK)
    return;
  const int words = pattern_mask_words(d.n);
  void* mp = nullptr;
  if (hipMalloc(&mp, sizeof(uint32_t) * words) != hipSuccess)
    return;
  h->allocations.push_back(mp);
  uint32_t* mask = (uint32_t*)mp;
  d.pat = pat;
  d.pat_p = period;
  d.pat_div = period >= 2 ? make_divisor(period) : Divisor{0u, 0};
  launch_pattern_mask(d, mask, h->ctx->stream);
  std::vector<uint32_t> bits(words);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(h->ctx->stream) != hipSuccess ||
      hipMemcpy(bits.data(), mask, sizeof(uint32_t) * words, hipMemcpyDeviceToHost) !=
        hipSuccess) {
    d.pat = nullptr;
    d.pat_p = 0;
    return;
  }
  long long irregular = 0;
  for (int i = 0; i < d.n; ++i)
    irregular += (bits[i >> 5] >> (i & 31)) & 1u;
  if (irregular * 10 > d.n) {
    d.pat = nullptr;
    d.pat_p = 0;
    return;
  }
  d.irregular = mask;
  h->pat_irregular = irregular;
}

}  // namespace

int lsqt::ham_scratch(lsqt_ham* h, int count)
{
  for (int i = 0; i < count; ++i) {
    if (!h->scratch[i]) {
      void* p = nullptr;
      size_t bytes = sizeof(double2) * (size_t)(h->d.n > 0 ? h->d.n : 1);
      hipError_t e = hipMalloc(&p, bytes);
      if (e != hipSuccess)
        return set_error(LSQT_ERR_NOMEM, "hipMalloc(%zu bytes): %s", bytes,
                         hipGetErrorString(e));
      h->scratch[i] = (double2*)p;
    }
  }
  return LSQT_OK;
}

extern "C" int lsqt_ham_create(lsqt_ctx* ctx, int n, int K, const int* nn, const int* nl,
                               const double* pot, const double* hr, const double* hi,
                               const double* xx, double energy_max, int layout_hint,
                               lsqt_ham** out)
{
  if (!ctx || !out)
    return set_error(LSQT_ERR_INVALID, "null context or output handle");
  *out = nullptr;
  if (n <= 0 || K <= 0)
    return set_error(LSQT_ERR_INVALID, "n=%d and max_neighbor=%d must be positive", n, K);
  if (!nn || !nl || !pot || !hr || !hi || !xx)
    return set_error(LSQT_ERR_INVALID, "null Hamiltonian array");
  if (!(energy_max > 0.0))
    return set_error(LSQT_ERR_INVALID, "energy_max must be > 0");
  if (layout_hint < LSQT_LAYOUT_AUTO || layout_hint > LSQT_LAYOUT_PACKED)
    return set_error(LSQT_ERR_INVALID, "unknown layout %d", layout_hint);
  // validate the structure before anything touches the device
  for (int i = 0; i < n; ++i) {
    if (nn[i] < 0 || nn[i] > K)
      return set_error(LSQT_ERR_INVALID, "site %d has %d neighbours (max %d)", i, nn[i], K);
    for (int m = 0; m < nn[i]; ++m) {
      int j = nl[(size_t)i * K + m];
      if (j < 0 || j >= n)
        return set_error(LSQT_ERR_INVALID, "site %d slot %d: neighbour %d out of range", i, m,
                         j);
    }
  }
  LSQT_TRY(ctx_activate(ctx));

  const int stride = pair_stride(n);
  const size_t P = (size_t)stride * K;  // column-major pair slots, padded columns
  lsqt_ham* h = new lsqt_ham();
  h->ctx = ctx;
  ++ctx->live;
  HamDev& d = h->d;
  std::memset(&d, 0, sizeof(d));
  d.n = n;
  d.K = K;
  d.stride = stride;
  d.emax = energy_max;

  // distinct (re, im, xx) triples, numbered by first appearance
  const int ib = index_bits_for(n);
  const int cb = 32 - ib;
  const long long max_classes =
    std::min<long long>((1ll << cb) - 1, (long long)kPackedMaxClasses);
  std::vector<uint32_t> words;
  std::vector<double2> tab;
  std::vector<double> tab_xx;
  bool packed = layout_hint != LSQT_LAYOUT_FULL;
  if (packed) {
    std::unordered_map<TripleKey, int, TripleHash> ids;
    words.assign(P, kEmptySlot);
    TripleKey last{~0ull, ~0ull, ~0ull};
    int last_id = -1;
    for (int i = 0; i < n && packed; ++i) {
      for (int m = 0; m < nn[i]; ++m) {
        size_t q = (size_t)i * K + m;
        TripleKey key{bits_of(hr[q]), bits_of(hi[q]), bits_of(xx[q])};
        int id;
        if (key == last) {
          id = last_id;
        } else {
          auto it = ids.find(key);
          if (it == ids.end()) {
            if ((long long)tab.size() >= max_classes) {
              packed = false;
              break;
            }
            id = (int)tab.size();
            ids.emplace(key, id);
            tab.push_back(make_double2(hr[q], hi[q]));
            tab_xx.push_back(xx[q]);
          } else {
            id = it->second;
          }
          last = key;
          last_id = id;
        }
        words[(size_t)m * stride + i] = ((uint32_t)id << ib) | (uint32_t)nl[q];
      }
    }
    if (!packed && layout_hint == LSQT_LAYOUT_PACKED) {
      --ctx->live;
      delete h;
      return set_error(LSQT_ERR_UNSUPPORTED,
                       "packed layout needs at most %lld distinct hopping triples", max_classes);
    }
  }

  int rc = upload(h, &d.pot, pot, (size_t)n);
  if (rc == LSQT_OK && packed) {
    d.layout = tab.size() <= (size_t)kLdsTableMax ? kPackedLds : kPackedGlobal;
    d.idx_bits = ib;
    d.idx_mask = (ib >= 32) ? 0xFFFFFFFFu : ((1u << ib) - 1u);
    d.ncls = (int)tab.size();
    rc = upload(h, &d.words, words.data(), P);
    if (rc == LSQT_OK)
      rc = upload(h, &d.tab, tab.data(), tab.size());
    if (rc == LSQT_OK)
      rc = upload(h, &d.tab_xx, tab_xx.data(), tab_xx.size());
    words.clear();
    words.shrink_to_fit();
  } else if (rc == LSQT_OK) {
    d.layout = kFull;
    std::vector<int> nl_cm(P, -1);
    std::vector<double> hr_cm(P, 0.0), hi_cm(P, 0.0), xx_cm(P, 0.0);
    for (int i = 0; i < n; ++i)
      for (int m = 0; m < nn[i]; ++m) {
        size_t q = (size_t)i * K + m, c = (size_t)m * stride + i;
        nl_cm[c] = nl[q];
        hr_cm[c] = hr[q];
        hi_cm[c] = hi[q];
        xx_cm[c] = xx[q];
      }
    rc = upload(h, &d.nl, nl_cm.data(), P);
    if (rc == LSQT_OK)
      rc = upload(h, &d.hr, hr_cm.data(), P);
    if (rc == LSQT_OK)
      rc = upload(h, &d.hi, hi_cm.data(), P);
    if (rc == LSQT_OK)
      rc = upload(h, &d.xx, xx_cm.data(), P);
  }
  if (rc != LSQT_OK) {
    lsqt_ham_destroy(h);
    return rc;
  }
  const std::vector<long long> offsets = detect_lattice_offsets(h, n, K, nn, nl);
  detect_slot_pattern(h);
  detect_sweep(h, offsets);
  // host copies for the FP32 / block paths, built on demand
  h->host_pot.assign(pot, pot + n);
  h->host_tab = tab;
  h->host_tab_xx = tab_xx;
  *out = h;
  return LSQT_OK;
}

namespace {

// checks one lattice description and its vacancy map; n_pristine and
// n_sites (after vacancies) on success
int check_lattice(const lsqt_lattice* L, const int* site_index, int* n_pristine, int* n_sites)
{
  if (!L->orbital_x || !L->hop_count || !L->hop_cell || !L->hop_orbital || !L->hop_re ||
      !L->hop_im)
    return set_error(LSQT_ERR_INVALID, "null lattice table");
  const int No = L->orbitals, K = L->max_neighbor;
  if (No <= 0 || K <= 0 || L->transport < 0 || L->transport > 2)
    return set_error(LSQT_ERR_INVALID, "orbitals, max_neighbor or transport direction invalid");
  long long nll = No;
  for (int d = 0; d < 3; ++d) {
    if (L->cells[d] <= 0 || (L->pbc[d] != 0 && L->pbc[d] != 1))
      return set_error(LSQT_ERR_INVALID, "cells and boundaries must be > 0 and 0/1");
    nll *= L->cells[d];
  }
  if (nll > INT_MAX)
    return set_error(LSQT_ERR_INVALID, "%lld sites exceed the int32 site index", nll);
  for (int m = 0; m < No; ++m) {
    if (L->hop_count[m] < 0 || L->hop_count[m] > K)
      return set_error(LSQT_ERR_INVALID, "orbital %d has %d hoppings (max %d)", m,
                       L->hop_count[m], K);
    for (int i = 0; i < L->hop_count[m]; ++i) {
      const int k = m * K + i;
      if (L->hop_orbital[k] < 0 || L->hop_orbital[k] >= No)
        return set_error(LSQT_ERR_INVALID, "hopping %d of orbital %d: orbital %d out of range",
                         i, m, L->hop_orbital[k]);
      // a hopping that some site keeps (within the cell count of every open
      // direction) may wrap once across a periodic one
      bool kept = true;
      int beyond = -1;
      for (int d = 0; d < 3; ++d) {
        const int c = std::abs(L->hop_cell[3 * k + d]);
        if (!L->pbc[d] && c >= L->cells[d])
          kept = false;
        if (L->pbc[d] && c > L->cells[d])
          beyond = d;
      }
      if (kept && beyond >= 0)
        return set_error(LSQT_ERR_INVALID, "hopping %d of orbital %d reaches beyond one "
                         "period in direction %d", i, m, beyond);
    }
  }
  const int np = (int)nll;
  int kept = np;
  if (site_index) {
    // a map onto 0 .. kept - 1, each index once, -1 for the vacancies
    std::vector<char> seen((size_t)np, 0);
    kept = 0;
    for (int p = 0; p < np; ++p) {
      const int v = site_index[p];
      if (v < 0)
        continue;
      if (v >= np || seen[(size_t)v])
        return set_error(LSQT_ERR_INVALID, "site index %d of pristine site %d is out of range "
                         "or repeated", v, p);
      seen[(size_t)v] = 1;
      ++kept;
    }
    for (int v = 0; v < kept; ++v)
      if (!seen[(size_t)v])
        return set_error(LSQT_ERR_INVALID, "the site index leaves site %d unused", v);
  }
  if (kept == 0)
    return set_error(LSQT_ERR_INVALID, "a lattice without sites");
  *n_pristine = np;
  *n_sites = kept;
  return LSQT_OK;
}

}  // namespace

extern "C" int lsqt_ham_create_lattices(lsqt_ctx* ctx, int segments, const lsqt_lattice* lats,
                                        const int* const* site_index, const double* potential,
                                        double energy_max, lsqt_ham** out)
{
  if (!ctx || !lats || !potential || !out || segments < 1)
    return set_error(LSQT_ERR_INVALID, "null argument or no lattice");
  *out = nullptr;
  if (!(energy_max > 0.0))
    return set_error(LSQT_ERR_INVALID, "energy_max must be > 0");
  std::vector<int> n_pristine((size_t)segments), starts((size_t)segments + 1, 0);
  int K = 0;
  long long total = 0;
  for (int k = 0; k < segments; ++k) {
    int np = 0, nk = 0;
    LSQT_TRY(check_lattice(&lats[k], site_index ? site_index[k] : nullptr, &np, &nk));
    n_pristine[k] = np;
    starts[k] = (int)total;
    total += nk;
    if (total > INT_MAX)
      return set_error(LSQT_ERR_INVALID, "%lld sites exceed the int32 site index", total);
    K = std::max(K, lats[k].max_neighbor);
  }
  const int n = (int)total;
  starts[segments] = n;
  const int ib = index_bits_for(n);
  // classes: the distinct (re, im, xx) of every (lattice, orbital, hopping),
  // xx = a[transport] * offset[transport] + x[orbital'] - x[m] in the Model's
  // operation order (src/model_lattice.cu:373-375)
  std::unordered_map<TripleKey, int, TripleHash> ids;
  std::vector<double2> tab;
  std::vector<double> tab_xx;
  std::vector<std::vector<int>> cls((size_t)segments);
  std::vector<std::vector<int4>> hop((size_t)segments);
  const long long max_classes =
    std::min<long long>(ib >= 32 ? 0 : (1ll << (32 - ib)) - 1, (long long)kPackedMaxClasses);
  for (int s = 0; s < segments; ++s) {
    const lsqt_lattice* L = &lats[s];
    const int No = L->orbitals, Ks = L->max_neighbor, t = L->transport;
    cls[s].assign((size_t)No * Ks, 0);
    hop[s].assign((size_t)No * Ks, make_int4(0, 0, 0, 0));
    for (int m = 0; m < No; ++m)
      for (int i = 0; i < L->hop_count[m]; ++i) {
        const int k = m * Ks + i;
        double x12 = L->lattice_constant[t] * L->hop_cell[3 * k + t];
        x12 += L->orbital_x[L->hop_orbital[k]] - L->orbital_x[m];
        TripleKey key{bits_of(L->hop_re[k]), bits_of(L->hop_im[k]), bits_of(x12)};
        auto it = ids.find(key);
        if (it == ids.end()) {
          if ((long long)tab.size() >= max_classes)
            return set_error(LSQT_ERR_UNSUPPORTED, "%zu hopping classes do not fit the packed "
                             "words of %d sites", tab.size() + 1, n);
          it = ids.emplace(key, (int)tab.size()).first;
          tab.push_back(make_double2(L->hop_re[k], L->hop_im[k]));
          tab_xx.push_back(x12);
        }
        cls[s][k] = it->second;
        hop[s][k] = make_int4(L->hop_cell[3 * k], L->hop_cell[3 * k + 1], L->hop_cell[3 * k + 2],
                              L->hop_orbital[k]);
      }
  }
  if (tab.empty()) {  // no hoppings at all: a table of one unused class
    tab.push_back(make_double2(0.0, 0.0));
    tab_xx.push_back(0.0);
  }
  LSQT_TRY(ctx_activate(ctx));
  const int stride = pair_stride(n);
  lsqt_ham* h = new lsqt_ham();
  h->ctx = ctx;
  ++ctx->live;
  HamDev& d = h->d;
  std::memset(&d, 0, sizeof(d));
  d.n = n;
  d.K = K;
  d.stride = stride;
  d.emax = energy_max;
  d.layout = tab.size() <= (size_t)kLdsTableMax ? kPackedLds : kPackedGlobal;
  d.idx_bits = ib;
  d.idx_mask = (1u << ib) - 1u;
  d.ncls = (int)tab.size();
  int rc = upload(h, &d.pot, potential, (size_t)n);
  if (rc == LSQT_OK)
    rc = upload(h, &d.tab, tab.data(), tab.size());
  if (rc == LSQT_OK)
    rc = upload(h, &d.tab_xx, tab_xx.data(), tab_xx.size());
  uint32_t* words = nullptr;
  if (rc == LSQT_OK) {
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, sizeof(uint32_t) * (size_t)stride * K);
    if (e != hipSuccess) {
      rc = set_error(LSQT_ERR_NOMEM, "hipMalloc(words): %s", hipGetErrorString(e));
    } else {
      h->allocations.push_back(p);
      words = (uint32_t*)p;
    }
  }
  if (rc == LSQT_OK) {
    // the padding columns past n stay marked empty
    hipError_t e = hipMemsetD32Async((hipDeviceptr_t)words, (int)kEmptySlot,
                                     (size_t)stride * K, ctx->stream);
    if (e != hipSuccess)
      rc = set_error(LSQT_ERR_HIP, "clearing the lattice words: %s", hipGetErrorString(e));
  }
  for (int s = 0; s < segments && rc == LSQT_OK; ++s) {
    const lsqt_lattice* L = &lats[s];
    const int* dev_count = nullptr;
    const int4* dev_hop = nullptr;
    const int* dev_cls = nullptr;
    const int* dev_index = nullptr;
    rc = upload(h, &dev_count, L->hop_count, (size_t)L->orbitals);
    if (rc == LSQT_OK)
      rc = upload(h, &dev_hop, hop[s].data(), hop[s].size());
    if (rc == LSQT_OK)
      rc = upload(h, &dev_cls, cls[s].data(), cls[s].size());
    if (rc == LSQT_OK && site_index && site_index[s])
      rc = upload(h, &dev_index, site_index[s], (size_t)n_pristine[s]);
    if (rc != LSQT_OK)
      break;
    LatticeDev ld{};
    for (int q = 0; q < 3; ++q) {
      ld.cells[q] = L->cells[q];
      ld.pbc[q] = L->pbc[q];
    }
    ld.orbitals = L->orbitals;
    ld.K = L->max_neighbor;
    ld.hop_count = dev_count;
    ld.hop = dev_hop;
    ld.cls = dev_cls;
    ld.site_index = dev_index;
    ld.offset = starts[s];
    launch_lattice_words(ld, words, stride, ib, K, n_pristine[s], ctx->stream);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess)
      rc = set_error(LSQT_ERR_HIP, "building the lattice words: %s", hipGetErrorString(e));
  }
  if (rc == LSQT_OK) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess)
      rc = set_error(LSQT_ERR_HIP, "building the lattice words: %s", hipGetErrorString(e));
    d.words = words;
  }
  if (rc != LSQT_OK) {
    lsqt_ham_destroy(h);
    return rc;
  }
  if (segments > 1) {
    h->seg_starts = starts;  // lattice tile schedules do not cross segments
  } else if (!site_index || !site_index[0]) {
    // block tile schedules of a pristine lattice: slabs of one x cell, rows
    // of one y cell
    const lsqt_lattice* L = &lats[0];
    const long long slab = (long long)L->cells[1] * L->cells[2] * L->orbitals;
    const long long row = (long long)L->cells[2] * L->orbitals;
    if (L->cells[0] > 1 && slab >= 64 && row >= 8 && slab / row >= 4) {
      h->slab = (int)slab;
      h->row = (int)row;
    }
  }
  detect_slot_pattern(h);
  if (segments == 1 && (!site_index || !site_index[0]))
    detect_sweep(h, {(long long)lats[0].cells[1] * lats[0].cells[2] * lats[0].orbitals});
  else if (segments > 1)
    detect_block_sweep(h, segments, lats, site_index, starts, cls);
  h->host_pot.assign(potential, potential + n);
  h->host_tab = tab;
  h->host_tab_xx = tab_xx;
  *out = h;
  return LSQT_OK;
}

extern "C" int lsqt_ham_create_lattice(lsqt_ctx* ctx, const lsqt_lattice* L,
                                       const double* potential, double energy_max,
                                       lsqt_ham** out)
{
  if (!ctx || !L || !potential || !out)
    return set_error(LSQT_ERR_INVALID, "null argument");
  return lsqt_ham_create_lattices(ctx, 1, L, nullptr, potential, energy_max, out);
}

extern "C" int lsqt_ham_create_segmented(lsqt_ctx* ctx, int segments, const int* starts, int n,
                                         int K, const int* nn, const int* nl, const double* pot,
                                         const double* hr, const double* hi, const double* xx,
                                         double energy_max, lsqt_ham** out)
{
  if (!out || !starts || !nn || !nl)
    return set_error(LSQT_ERR_INVALID, "null argument");
  *out = nullptr;
  if (segments < 1 || starts[0] != 0 || starts[segments] != n)
    return set_error(LSQT_ERR_INVALID, "segment starts must run from 0 to n = %d", n);
  for (int k = 0; k < segments; ++k) {
    if (starts[k + 1] <= starts[k])
      return set_error(LSQT_ERR_INVALID, "segment %d is empty or reversed", k);
    for (int i = starts[k]; i < starts[k + 1]; ++i)
      for (int m = 0; m < nn[i] && m < K; ++m) {
        const int j = nl[(size_t)i * K + m];
        if (j < starts[k] || j >= starts[k + 1])
          return set_error(LSQT_ERR_INVALID, "site %d of segment %d has neighbour %d outside "
                           "its segment", i, k, j);
      }
  }
  lsqt_ham* h = nullptr;
  LSQT_TRY(lsqt_ham_create(ctx, n, K, nn, nl, pot, hr, hi, xx, energy_max,
                           LSQT_LAYOUT_PACKED, &h));
  h->seg_starts.assign(starts, starts + segments + 1);
  h->slab = h->row = 0;  // lattice tile schedules do not cross segments
  if (segments > 1)
    h->bsweep.clear();  // a geometry per system is known only for lattices
  *out = h;
  return LSQT_OK;
}

extern "C" int lsqt_ham_destroy(lsqt_ham* h)
{
  if (!h)
    return LSQT_OK;
  (void)hipSetDevice(h->ctx->device);
  (void)hipStreamSynchronize(h->ctx->stream);
  for (void* p : h->allocations)
    (void)hipFree(p);
  for (double2* p : h->scratch)
    if (p)
      (void)hipFree(p);
  ham_release_extra(h);
  --h->ctx->live;
  delete h;
  return LSQT_OK;
}

extern "C" int lsqt_ham_get_info(const lsqt_ham* h, lsqt_ham_info* info)
{
  if (!h || !info)
    return set_error(LSQT_ERR_INVALID, "null argument");
  info->n = h->d.n;
  info->max_neighbor = h->d.K;
  info->layout = h->d.layout == kFull ? LSQT_LAYOUT_FULL : LSQT_LAYOUT_PACKED;
  info->classes = h->d.ncls;
  info->index_bits = h->d.idx_bits;
  info->energy_max = h->d.emax;
  // potential 8 + per slot (word 4 | index 4 + re 8 + im 8) + psi_{m-1} 16 +
  // psi_{m-2} 16 + psi_m 16 + left vector 16; with a slot pattern the words
  // of the irregular sites only, and the mask (1 bit per site)
  double per_slot = h->d.layout == kFull ? 20.0 : 4.0;
  info->pattern_period = h->d.pat ? h->d.pat_p : 0;
  info->pattern_irregular = h->d.pat ? (double)h->pat_irregular / h->d.n : 0.0;
  if (h->d.pat)
    info->bytes_per_site_f64 =
      8.0 + per_slot * h->d.K * info->pattern_irregular + 1.0 / 8.0 + 64.0;
  else
    info->bytes_per_site_f64 = 8.0 + per_slot * h->d.K + 64.0;
  info->slab = h->slab;
  info->row = h->row;
  info->sweep_slab = h->sweep64.S;
  info->sweep_halo = h->sweep64.hh;
  info->sweep_tiles = h->sweep64.tiles;
  info->sweep_tiles_f32 = h->sweep32.tiles;
  // one two-step launch per site: psi_{m-2} and psi_{m-1} read, psi_m and
  // psi_{m+1} written, the left vector and the potential read (the slots
  // come from the per-orbital table: no words)
  if (h->sweep64.S) {
    info->bytes_per_site_sweep_f64 = 5 * 16.0 + 8.0;
    info->bytes_per_site_sweep_f32 = 5 * 8.0 + 4.0;
  } else {
    info->bytes_per_site_sweep_f64 = 0.0;
    info->bytes_per_site_sweep_f32 = 0.0;
  }
  info->block_sweep_systems = (int)h->bsweep.size();
  info->sweep_steps = h->sweepn64.S ? h->sweepn64.ns : (h->sweep64.S ? 2 : 0);
  info->sweep_steps_f32 = h->sweepn32.S ? h->sweepn32.ns : (h->sweep32.S ? 2 : 0);
  info->sweep_tiles_deep = h->sweepn64.S ? h->sweepn64.tiles : 0;
  info->block_sweep_steps = h->bsweep.empty() ? 0 : h->bsweep_steps;
  return LSQT_OK;
}

// ---------------------------------------------------------------------------
// vectors
// ---------------------------------------------------------------------------
extern "C" int lsqt_vec_create(lsqt_ctx* ctx, int n, lsqt_vec** out)
{
  if (!ctx || !out)
    return set_error(LSQT_ERR_INVALID, "null context or output handle");
  *out = nullptr;
  if (n <= 0)
    return set_error(LSQT_ERR_INVALID, "vector length %d must be positive", n);
  LSQT_TRY(ctx_activate(ctx));
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(double2) * (size_t)n);
  if (e != hipSuccess)
    return set_error(LSQT_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
  launch_zero((double2*)p, n, ctx->stream);
  e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(p);
    return set_error(LSQT_ERR_HIP, "zeroing a vector: %s", hipGetErrorString(e));
  }
  lsqt_vec* v = new lsqt_vec();
  v->ctx = ctx;
  ++ctx->live;
  v->n = n;
  v->d = (double2*)p;
  *out = v;
  return LSQT_OK;
}

extern "C" int lsqt_vec_clone(const lsqt_vec* src, lsqt_vec** out)
{
  if (!src || !out)
    return set_error(LSQT_ERR_INVALID, "null argument");
  LSQT_TRY(lsqt_vec_create(src->ctx, src->n, out));
  launch_copy((*out)->d, src->d, src->n, src->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_destroy(lsqt_vec* v)
{
  if (!v)
    return LSQT_OK;
  (void)hipSetDevice(v->ctx->device);
  (void)hipStreamSynchronize(v->ctx->stream);
  (void)hipFree(v->d);
  --v->ctx->live;
  delete v;
  return LSQT_OK;
}

extern "C" int lsqt_vec_size(const lsqt_vec* v, int* n)
{
  if (!v || !n)
    return set_error(LSQT_ERR_INVALID, "null argument");
  *n = v->n;
  return LSQT_OK;
}

extern "C" int lsqt_vec_from_host(lsqt_vec* v, const double* re, const double* im)
{
  if (!v || !re || !im)
    return set_error(LSQT_ERR_INVALID, "null argument");
  LSQT_TRY(ctx_activate(v->ctx));
  std::vector<double2> tmp((size_t)v->n);
  for (int i = 0; i < v->n; ++i)
    tmp[i] = make_double2(re[i], im[i]);
  LSQT_HIP(hipStreamSynchronize(v->ctx->stream));
  LSQT_HIP(hipMemcpy(v->d, tmp.data(), sizeof(double2) * (size_t)v->n, hipMemcpyHostToDevice));
  return LSQT_OK;
}

extern "C" int lsqt_vec_to_host(const lsqt_vec* v, double* re, double* im)
{
  if (!v || !re || !im)
    return set_error(LSQT_ERR_INVALID, "null argument");
  LSQT_TRY(ctx_activate(v->ctx));
  std::vector<double2> tmp((size_t)v->n);
  LSQT_HIP(hipStreamSynchronize(v->ctx->stream));
  LSQT_HIP(hipMemcpy(tmp.data(), v->d, sizeof(double2) * (size_t)v->n, hipMemcpyDeviceToHost));
  for (int i = 0; i < v->n; ++i) {
    re[i] = tmp[i].x;
    im[i] = tmp[i].y;
  }
  return LSQT_OK;
}

static int same_shape(const lsqt_vec* a, const lsqt_vec* b)
{
  if (!a || !b)
    return set_error(LSQT_ERR_INVALID, "null vector");
  if (a->n != b->n || a->ctx != b->ctx)
    return set_error(LSQT_ERR_INVALID, "vector sizes/contexts differ (%d vs %d)", a->n, b->n);
  return LSQT_OK;
}

extern "C" int lsqt_vec_copy(lsqt_vec* dst, const lsqt_vec* src)
{
  LSQT_TRY(same_shape(dst, src));
  LSQT_TRY(ctx_activate(dst->ctx));
  if (dst->d != src->d)
    launch_copy(dst->d, src->d, dst->n, dst->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_add(lsqt_vec* dst, const lsqt_vec* src)
{
  LSQT_TRY(same_shape(dst, src));
  if (dst->d == src->d)
    return set_error(LSQT_ERR_INVALID, "add: source and destination are the same vector");
  LSQT_TRY(ctx_activate(dst->ctx));
  launch_add(dst->d, src->d, dst->n, dst->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_apply_sz(lsqt_vec* dst, const lsqt_vec* src)
{
  LSQT_TRY(same_shape(dst, src));
  if (dst->d == src->d)
    return set_error(LSQT_ERR_INVALID, "apply_sz: source and destination are the same vector");
  LSQT_TRY(ctx_activate(dst->ctx));
  launch_sz(dst->d, src->d, dst->n, dst->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_swap(lsqt_vec* a, lsqt_vec* b)
{
  LSQT_TRY(same_shape(a, b));
  double2* t = a->d;
  a->d = b->d;
  b->d = t;
  return LSQT_OK;
}

int lsqt::ctx_reserve_moments(lsqt_ctx* ctx, int M, int tiles)
{
  LSQT_TRY(ensure_device((void**)&ctx->partial, &ctx->partial_cap,
                         sizeof(double2) * (size_t)M * (size_t)tiles));
  LSQT_TRY(ensure_device((void**)&ctx->mu_dev, &ctx->mu_cap, sizeof(double2) * (size_t)M));
  LSQT_TRY(ensure_pinned((void**)&ctx->mu_host, &ctx->mu_host_cap,
                         sizeof(double2) * (size_t)M));
  return LSQT_OK;
}

int lsqt::ctx_download_moments(lsqt_ctx* ctx, int M, double* mu_re, double* mu_im)
{
  LSQT_HIP(hipMemcpyAsync(ctx->mu_host, ctx->mu_dev, sizeof(double2) * (size_t)M,
                          hipMemcpyDeviceToHost, ctx->stream));
  LSQT_HIP(hipStreamSynchronize(ctx->stream));
  for (int m = 0; m < M; ++m) {
    mu_re[m] = ctx->mu_host[m].x;
    mu_im[m] = ctx->mu_host[m].y;
  }
  return LSQT_OK;
}

extern "C" int lsqt_vec_dot(const lsqt_vec* left, const lsqt_vec* psi, double* re, double* im)
{
  LSQT_TRY(same_shape(left, psi));
  if (!re || !im)
    return set_error(LSQT_ERR_INVALID, "null output");
  lsqt_ctx* ctx = left->ctx;
  LSQT_TRY(ctx_activate(ctx));
  int tiles = dot_tiles(left->n);
  LSQT_TRY(ctx_reserve_moments(ctx, 1, tiles));
  launch_dot(psi->d, left->d, ctx->partial, left->n, ctx->stream);
  launch_reduce_partials(ctx->partial, (size_t)tiles, tiles, 1, ctx->mu_dev, ctx->stream);
  LSQT_HIP(hipGetLastError());
  return ctx_download_moments(ctx, 1, re, im);
}

// sites per partial of Vector::inner_product_1: the reference's BLOCK_SIZE
// (src/common.h), one k_dot tile of two sites per thread
constexpr int kInnerProductBlock = 512;

extern "C" int lsqt_vec_inner_product_1(const lsqt_vec* v, int number_of_atoms,
                                        const lsqt_vec* other, lsqt_vec* target, int offset)
{
  LSQT_TRY(same_shape(v, other));
  if (!target || target->ctx != v->ctx)
    return set_error(LSQT_ERR_INVALID, "inner_product_1: null target or another context");
  if (number_of_atoms <= 0 || number_of_atoms > v->n)
    return set_error(LSQT_ERR_INVALID, "inner_product_1: %d atoms for vectors of length %d",
                     number_of_atoms, v->n);
  const long long blocks = (number_of_atoms - 1) / kInnerProductBlock + 1;
  if (offset < 0 || offset + blocks > target->n)
    return set_error(LSQT_ERR_INVALID, "inner_product_1: %lld partials at offset %d do not fit "
                     "a target of length %d", blocks, offset, target->n);
  if (target->d == v->d || target->d == other->d)
    return set_error(LSQT_ERR_INVALID, "inner_product_1: the target must be another vector");
  LSQT_TRY(ctx_activate(v->ctx));
  launch_dot(v->d, other->d, target->d + offset, number_of_atoms, v->ctx->stream,
             kInnerProductBlock / 256);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_inner_product_2(const lsqt_vec* v, int number_of_atoms,
                                        int number_of_moments, lsqt_vec* target)
{
  if (!v || !target || target->ctx != v->ctx)
    return set_error(LSQT_ERR_INVALID, "inner_product_2: null vector or another context");
  if (number_of_atoms <= 0 || number_of_moments <= 0)
    return set_error(LSQT_ERR_INVALID, "inner_product_2: %d atoms, %d moments",
                     number_of_atoms, number_of_moments);
  const long long blocks = (number_of_atoms - 1) / kInnerProductBlock + 1;
  if (blocks * number_of_moments > v->n || number_of_moments > target->n)
    return set_error(LSQT_ERR_INVALID, "inner_product_2: %lld x %d partials in a vector of "
                     "length %d, %d moments into a target of length %d", blocks,
                     number_of_moments, v->n, number_of_moments, target->n);
  if (target->d == v->d)
    return set_error(LSQT_ERR_INVALID, "inner_product_2: the target must be another vector");
  LSQT_TRY(ctx_activate(v->ctx));
  launch_reduce_partials(v->d, (size_t)blocks, (int)blocks, number_of_moments, target->d,
                         v->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_vec_device_data(const lsqt_vec* v, void** data)
{
  if (!v || !data)
    return set_error(LSQT_ERR_INVALID, "null vector or output");
  *data = (void*)v->d;
  return LSQT_OK;
}

// ---------------------------------------------------------------------------
// operators
// ---------------------------------------------------------------------------
static int check_ham_vecs(lsqt_ham* h, std::initializer_list<const lsqt_vec*> vs)
{
  if (!h)
    return set_error(LSQT_ERR_INVALID, "null Hamiltonian");
  for (const lsqt_vec* v : vs) {
    if (!v)
      return set_error(LSQT_ERR_INVALID, "null vector");
    if (v->n != h->d.n || v->ctx != h->ctx)
      return set_error(LSQT_ERR_INVALID, "vector of length %d does not match the %d-site "
                       "Hamiltonian", v->n, h->d.n);
  }
  return ctx_activate(h->ctx);
}

extern "C" int lsqt_apply(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out)
{
  LSQT_TRY(check_ham_vecs(h, {in, out}));
  if (in->d == out->d)
    return set_error(LSQT_ERR_INVALID, "apply: input and output must differ");
  launch_hpsi(h->d, in->d, nullptr, out->d, nullptr, nullptr, false, false, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_apply_commutator(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out)
{
  LSQT_TRY(check_ham_vecs(h, {in, out}));
  if (in->d == out->d)
    return set_error(LSQT_ERR_INVALID, "apply_commutator: input and output must differ");
  launch_xop(h->d, in->d, out->d, false, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_apply_current(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out)
{
  LSQT_TRY(check_ham_vecs(h, {in, out}));
  if (in->d == out->d)
    return set_error(LSQT_ERR_INVALID, "apply_current: input and output must differ");
  launch_xop(h->d, in->d, out->d, true, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_kernel_polynomial(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1,
                                      lsqt_vec* s2)
{
  LSQT_TRY(check_ham_vecs(h, {s0, s1, s2}));
  if (s1->d == s2->d || s1->d == s0->d)
    return set_error(LSQT_ERR_INVALID, "kernel_polynomial: s1 must differ from s0 and s2");
  launch_hpsi(h->d, s1->d, s0->d, s2->d, nullptr, nullptr, true, false, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_kernel_polynomial_pair(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1,
                                           lsqt_vec* s2, lsqt_vec* s3)
{
  LSQT_TRY(check_ham_vecs(h, {s0, s1, s2, s3}));
  if (h->sweep64.S == 0)
    return set_error(LSQT_ERR_UNSUPPORTED, "kernel_polynomial_pair: the Hamiltonian has no "
                     "two-step sweep (lsqt_ham_info.sweep_slab is 0)");
  const double2* p[4] = {s0->d, s1->d, s2->d, s3->d};
  for (int a = 0; a < 4; ++a)
    for (int b = a + 1; b < 4; ++b)
      if (p[a] == p[b])
        return set_error(LSQT_ERR_INVALID, "kernel_polynomial_pair: s0, s1, s2 and s3 must "
                         "be distinct");
  launch_sweep2(h->d, h->sweep64, s0->d, s1->d, s2->d, s3->d, nullptr, nullptr, nullptr,
                h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_chebyshev_01(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1,
                                 lsqt_vec* s, double b0, double b1, int direction)
{
  LSQT_TRY(check_ham_vecs(h, {s0, s1, s}));
  launch_cheb01(s0->d, s1->d, s->d, b0, b1, direction, h->d.n, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_chebyshev_2(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1,
                                lsqt_vec* s2, lsqt_vec* s, double bm, int label)
{
  LSQT_TRY(check_ham_vecs(h, {s0, s1, s2, s}));
  if (label < 1 || label > 4)
    return set_error(LSQT_ERR_INVALID, "label %d must be 1..4", label);
  if (s1->d == s2->d || s1->d == s->d || s2->d == s->d)
    return set_error(LSQT_ERR_INVALID, "chebyshev_2: s1, s2 and s must be distinct");
  launch_cheb2(h->d, s0->d, s1->d, s2->d, s->d, bm, label, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_chebyshev_1x(lsqt_ham* h, const lsqt_vec* s1x, lsqt_vec* s, double b1)
{
  LSQT_TRY(check_ham_vecs(h, {s1x, s}));
  launch_cheb1x(s1x->d, s->d, b1, h->d.n, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_chebyshev_2x(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s0x,
                                 const lsqt_vec* s1, const lsqt_vec* s1x, lsqt_vec* s2,
                                 lsqt_vec* s2x, lsqt_vec* s, double bm, int label)
{
  LSQT_TRY(check_ham_vecs(h, {s0, s0x, s1, s1x, s2, s2x, s}));
  if (label < 1 || label > 4)
    return set_error(LSQT_ERR_INVALID, "label %d must be 1..4", label);
  launch_cheb2x(h->d, s0->d, s0x->d, s1->d, s1x->d, s2->d, s2x->d, s->d, bm, label,
                h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

// ---------------------------------------------------------------------------
// fused algorithms
// ---------------------------------------------------------------------------
int lsqt::moments_enqueue(lsqt_ham* h, const double2* left, const double2* right, int M)
{
  lsqt_ctx* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  const int n = h->d.n;
  const int tiles = hpsi_tiles(n, h->d.K);
  // two steps per launch where the lattice allows (the same per-site values;
  // one partial per sweep tile)
  const SweepMap& sw = h->sweep64;
  const SweepMap& sn = h->sweepn64;
  const bool sweep = sw.S > 0 && M >= 4;
  const bool deep = sweep && sn.S > 0 && M >= 4 + sn.ns;
  const int stride_tiles =
    std::max(tiles, std::max(sweep ? sw.tiles : 0, deep ? sn.tiles : 0));
  LSQT_TRY(ctx_reserve_moments(ctx, M, stride_tiles));
  LSQT_TRY(ham_scratch(h, sweep ? 4 : 2));
  double2* A = h->scratch[0];
  double2* B = h->scratch[1];
  const size_t stride = (size_t)stride_tiles;
  // psi_0 = right; mu_0 = <left|psi_0>
  launch_copy(A, right, n, st);
  launch_dot(A, left, ctx->partial, n, st, hpsi_spt(h->d.K));
  // psi_1 = H~ psi_0; mu_1
  launch_hpsi(h->d, A, nullptr, B, left, ctx->partial + stride, false, true, st);
  // several steps per launch where the lattice allows (the same per-site
  // values; one partial per sweep tile): moments 2, 3 by a two-step launch,
  // then NS at a time from 4, then two at a time, the rest one at a time
  int m = 2, m_deep0 = 2, m_deep = 2;
  if (sweep) {
    double2* C = h->scratch[2];
    double2* D = h->scratch[3];
    auto two = [&]() {
      launch_sweep2(h->d, sw, A, B, C, D, left, ctx->partial + (size_t)m * stride,
                    ctx->partial + (size_t)(m + 1) * stride, st);
      std::swap(A, C);
      std::swap(B, D);
      m += 2;
    };
    if (deep) {
      two();
      m_deep0 = m;
      for (; m + sn.ns - 1 < M; m += sn.ns) {
        launch_sweepn(h->d, sn, A, B, C, D, left, ctx->partial + (size_t)m * stride, stride,
                      st);
        std::swap(A, C);
        std::swap(B, D);
      }
    }
    m_deep = m;
    while (m + 1 < M)
      two();
  }
  const int m_sweep = m;
  // psi_m = 2 H~ psi_{m-1} - psi_{m-2}, written over psi_{m-2}
  for (; m < M; ++m) {
    launch_hpsi(h->d, B, A, A, left, ctx->partial + (size_t)m * stride, true, true, st);
    std::swap(A, B);
  }
  launch_reduce_partials(ctx->partial, stride, tiles, 2, ctx->mu_dev, st);
  launch_reduce_partials(ctx->partial + 2 * stride, stride, sw.tiles, m_deep0 - 2,
                         ctx->mu_dev + 2, st);
  launch_reduce_partials(ctx->partial + (size_t)m_deep0 * stride, stride, sn.tiles,
                         m_deep - m_deep0, ctx->mu_dev + m_deep0, st);
  launch_reduce_partials(ctx->partial + (size_t)m_deep * stride, stride, sw.tiles,
                         m_sweep - m_deep, ctx->mu_dev + m_deep, st);
  launch_reduce_partials(ctx->partial + (size_t)m_sweep * stride, stride, tiles, M - m_sweep,
                         ctx->mu_dev + m_sweep, st);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_moments(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                            double* mu_re, double* mu_im)
{
  LSQT_TRY(check_ham_vecs(h, {left, right}));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  LSQT_TRY(moments_enqueue(h, left->d, right->d, M));
  return ctx_download_moments(h->ctx, M, mu_re, mu_im);
}

extern "C" int lsqt_moments_to_vec(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right,
                                   int M, lsqt_vec* out)
{
  LSQT_TRY(check_ham_vecs(h, {left, right}));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!out || out->ctx != h->ctx || out->n < M)
    return set_error(LSQT_ERR_INVALID, "moment output: a vector of at least %d entries on the "
                     "Hamiltonian's context", M);
  LSQT_TRY(moments_enqueue(h, left->d, right->d, M));
  launch_copy(out->d, h->ctx->mu_dev, M, h->ctx->stream);
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

int lsqt::evolve_enqueue(lsqt_ham* h, int direction, double t, double2* state)
{
  hipStream_t st = h->ctx->stream;
  const int n = h->d.n;
  const SweepMap& sn = h->sweepn64;
  LSQT_TRY(ham_scratch(h, sn.S ? 4 : 2));
  double2* A = h->scratch[0];  // psi_{m-2}
  double2* B = h->scratch[1];  // psi_{m-1}
  BesselTerms bt = evolve_terms(direction, t);
  launch_copy(A, state, n, st);
  launch_hpsi(h->d, state, nullptr, B, nullptr, nullptr, false, false, st);
  launch_cheb01(A, B, state, bt.b0, bt.b1, direction, n, st);
  size_t k = 0;
  if (sn.S > 0) {
    // sn.ns terms per launch on a periodic lattice (the same additions in
    // the same order), the rest below, from an even term
    double2* C = h->scratch[2];
    double2* D = h->scratch[3];
    for (; k + sn.ns <= bt.bm.size(); k += sn.ns) {
      SweepTerms terms{};
      for (int l = 0; l < sn.ns; ++l) {
        terms.bm[l] = bt.bm[k + l];
        terms.label[l] = bt.label[k + l];
      }
      launch_sweepn_evo(h->d, sn, A, B, C, D, state, terms, st);
      std::swap(A, C);
      std::swap(B, D);
    }
  }
  for (; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_cheb2(h->d, A, B, A, state, d.bm, d.label, st, d.bm_prev, d.label_prev);
    double2* tmp = A;
    A = B;
    B = tmp;
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

int lsqt::evolvex_enqueue(lsqt_ham* h, int direction, double t, double2* state)
{
  hipStream_t st = h->ctx->stream;
  const int n = h->d.n;
  LSQT_TRY(ham_scratch(h, 4));
  double2 *A = h->scratch[0], *Ax = h->scratch[1];  // psi_{m-2}, [X, .] part
  double2 *B = h->scratch[2], *Bx = h->scratch[3];  // psi_{m-1}, [X, .] part
  BesselTerms bt = evolvex_terms(direction, t);
  launch_copy(A, state, n, st);
  launch_zero(Ax, n, st);
  launch_hpsi(h->d, state, nullptr, B, nullptr, nullptr, false, false, st);
  launch_xop(h->d, state, Bx, false, st);
  launch_cheb1x(Bx, state, bt.b1, n, st);
  for (size_t k = 0; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_cheb2x(h->d, A, Ax, B, Bx, A, Ax, state, d.bm, d.label, st, d.bm_prev,
                  d.label_prev);
    double2* tmp = A;
    A = B;
    B = tmp;
    tmp = Ax;
    Ax = Bx;
    Bx = tmp;
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_evolve(lsqt_ham* h, int direction, double t, lsqt_vec* state)
{
  LSQT_TRY(check_ham_vecs(h, {state}));
  if (direction != 1 && direction != -1)
    return set_error(LSQT_ERR_INVALID, "direction %d must be +1 or -1", direction);
  return evolve_enqueue(h, direction, t, state->d);
}

extern "C" int lsqt_evolvex(lsqt_ham* h, int direction, double t, lsqt_vec* state)
{
  LSQT_TRY(check_ham_vecs(h, {state}));
  if (direction != 1 && direction != -1)
    return set_error(LSQT_ERR_INVALID, "direction %d must be +1 or -1", direction);
  return evolvex_enqueue(h, direction, t, state->d);
}

extern "C" int lsqt_bessel_terms(double t, int* last_m)
{
  if (!last_m)
    return set_error(LSQT_ERR_INVALID, "null output");
  BesselTerms bt = evolve_terms(1, t);
  *last_m = 1 + (int)bt.bm.size();
  return LSQT_OK;
}
