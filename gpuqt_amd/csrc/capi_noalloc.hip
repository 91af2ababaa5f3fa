// capi_noalloc.hip -- lsqt_moments_noalloc: the FP64 moment set of
// lsqt_moments without any device allocation, for callers that cannot
// allocate (code under graph capture, tight serving loops).
//
// lsqt_moments keeps its recursion vectors on the Hamiltonian and its moment
// partials on the context, both allocated on first use.  Here every buffer
// beyond the caller's inputs and outputs lives in `scratch`, one fixed array
// per GPU: the recursion vectors (four where lsqt_moments advances several
// steps per launch), a ring of kRing moments' partials (reduced whenever the
// next launch's moments would not fit), the moments, and their split real /
// imaginary parts for the download.  The
// kernels and their order are those of lsqt_moments, so the moments are
// bit-identical to it.  Nothing in the scratch must start at zero: every
// partial a reduction reads is written by the step before it.
#include <algorithm>
#include <mutex>
#include <utility>

#include "lsqt_capi_internal.h"
#include "lsqt_launch.h"

namespace lsqt {
namespace {

constexpr int kNoAllocMaxSites = LSQT_NOALLOC_MAX_SITES;
constexpr int kNoAllocMaxMoments = LSQT_NOALLOC_MAX_MOMENTS;
constexpr int kRing = 64;                                     // moments per partial ring
static_assert(kRing >= kSweepMaxSteps, "a multi-step launch's moments fit the ring");
constexpr size_t kMaxTiles = (kNoAllocMaxSites + kBlock - 1) / kBlock;  // at 1 site / thread
// offsets in double2 units; every region starts on a 256-byte boundary
constexpr size_t kVecA = 0;
constexpr size_t kVecB = kVecA + (size_t)kNoAllocMaxSites;
constexpr size_t kVecC = kVecB + (size_t)kNoAllocMaxSites;
constexpr size_t kVecD = kVecC + (size_t)kNoAllocMaxSites;
constexpr size_t kRingAt = kVecD + (size_t)kNoAllocMaxSites;
constexpr size_t kMuAt = kRingAt + (size_t)kRing * kMaxTiles;
constexpr size_t kSplitAt = kMuAt + (size_t)kNoAllocMaxMoments;  // re then im, as doubles
constexpr size_t kScratchLen = kSplitAt + (size_t)kNoAllocMaxMoments;
static_assert(kVecB % 16 == 0 && kVecC % 16 == 0 && kVecD % 16 == 0 && kRingAt % 16 == 0 &&
                kMuAt % 16 == 0 && kSplitAt % 16 == 0,
              "scratch regions must stay 256-byte aligned");

// The scratch of every GPU the library runs on (the loader gives each device
// its own copy).  Host code must not use this name as an address: it is a
// host-side stand-in; the device address comes from hipGetSymbolAddress.
static __device__ __align__(256) double2 scratch[kScratchLen];

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
bool g_resolved[kMaxDevices] = {};
double2* g_scratch[kMaxDevices] = {};
hipError_t g_status[kMaxDevices] = {};

__global__ __launch_bounds__(kBlock) void k_split_moments(const double2* __restrict__ mu,
                                                          double* __restrict__ re,
                                                          double* __restrict__ im, int M)
{
  const int m = blockIdx.x * kBlock + threadIdx.x;
  if (m < M) {
    re[m] = mu[m].x;
    im[m] = mu[m].y;
  }
}

}  // namespace

// Resolves the scratch address of `device` (the current device) once per
// process; called by lsqt_ctx_create.  A failure is kept and reported by the
// entry point; contexts work without it.
void noalloc_setup(int device)
{
  if (device < 0 || device >= kMaxDevices)
    return;
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_resolved[device])
    return;
  void* p = nullptr;
  g_status[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  g_scratch[device] = g_status[device] == hipSuccess ? static_cast<double2*>(p) : nullptr;
  g_resolved[device] = true;
}

}  // namespace lsqt

using namespace lsqt;

extern "C" int lsqt_moments_noalloc(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right,
                                    int M, double* mu_re, double* mu_im)
{
  if (!h || !left || !right)
    return set_error(LSQT_ERR_INVALID, "null Hamiltonian or vector");
  if (left->n != h->d.n || right->n != h->d.n || left->ctx != h->ctx || right->ctx != h->ctx)
    return set_error(LSQT_ERR_INVALID, "vectors do not match the %d-site Hamiltonian", h->d.n);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  if (M < 2 || M > kNoAllocMaxMoments)
    return set_error(LSQT_ERR_INVALID, "lsqt_moments_noalloc: %d moments; the fixed scratch "
                     "holds 2 to %d (use lsqt_moments)", M, kNoAllocMaxMoments);
  if (h->d.n > kNoAllocMaxSites)
    return set_error(LSQT_ERR_INVALID, "lsqt_moments_noalloc: %d sites; the fixed scratch "
                     "holds at most %d (use lsqt_moments)", h->d.n, kNoAllocMaxSites);
  lsqt_ctx* ctx = h->ctx;
  const int dev = ctx->device;
  double2* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_mutex);
    if (dev >= 0 && dev < kMaxDevices && g_resolved[dev]) {
      if (g_status[dev] != hipSuccess)
        return set_error(LSQT_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev,
                         hipGetErrorString(g_status[dev]));
      base = g_scratch[dev];
    }
  }
  if (!base)
    return set_error(LSQT_ERR_INVALID, "no scratch resolved for device %d", dev);
  LSQT_TRY(ctx_activate(ctx));
  hipStream_t st = ctx->stream;
  const int n = h->d.n;
  const int tiles = hpsi_tiles(n, h->d.K);
  const SweepMap& sw = h->sweep64;
  const SweepMap& sn = h->sweepn64;
  const bool sweep = sw.S > 0 && M >= 4;  // as lsqt_moments
  const bool deep = sweep && sn.S > 0 && M >= 4 + sn.ns;
  const int stride_tiles =
    std::max(tiles, std::max(sweep ? sw.tiles : 0, deep ? sn.tiles : 0));
  if ((size_t)stride_tiles > kMaxTiles)
    return set_error(LSQT_ERR_INVALID, "%d moment partials per step exceed the ring",
                     stride_tiles);
  double2* A = base + kVecA;
  double2* B = base + kVecB;
  double2* C = base + kVecC;
  double2* D = base + kVecD;
  double2* ring = base + kRingAt;
  double2* mu = base + kMuAt;
  double* split = reinterpret_cast<double*>(base + kSplitAt);
  const size_t stride = (size_t)stride_tiles;
  // the launch schedule of lsqt_moments: moments 2, 3 by a two-step launch,
  // [4, m_deep) by sn.ns-step launches, [m_deep, m_sweep) by two-step ones,
  // the rest one step per launch (a part is empty where its shape is absent)
  const int m_deep0 = deep ? 4 : 2;
  const int m_deep = deep ? 4 + (M - 4) / sn.ns * sn.ns : 2;
  const int m_sweep = sweep ? m_deep + (M - m_deep) / 2 * 2 : 2;
  auto tiles_of = [&](int m) {
    return m < 2 ? tiles
                 : (m < m_deep0 ? sw.tiles : (m < m_deep ? sn.tiles : (m < m_sweep ? sw.tiles : tiles)));
  };
  // the ring holds the partials of moments rb .. rb + kRing - 1 (moment m in
  // slot m - rb); before a launch whose moments would not fit, the ring's
  // moments are reduced, in pieces of one partial count, and it starts over
  // at the launch's first moment
  int rb = 0;
  auto partial = [&](int m) { return ring + (size_t)(m - rb) * stride; };
  auto reduce = [&](int m_end) {
    int cuts[6] = {rb, 2, m_deep0, m_deep, m_sweep, m_end};
    for (int c = 1; c < 5; ++c)
      cuts[c] = std::min(std::max(cuts[c], rb), m_end);
    for (int p = 0; p < 5; ++p) {
      const int a = cuts[p], b = cuts[p + 1];
      if (b > a)
        launch_reduce_partials(ring + (size_t)(a - rb) * stride, stride, tiles_of(a), b - a,
                               mu + a, st);
    }
    rb = m_end;
  };
  launch_copy(A, right->d, n, st);
  launch_dot(A, left->d, partial(0), n, st, hpsi_spt(h->d.K));
  launch_hpsi(h->d, A, nullptr, B, left->d, partial(1), false, true, st);
  int m = 2;
  auto take = [&](int count) {  // room in the ring for moments m .. m + count - 1
    if (m + count - rb > kRing)
      reduce(m);
  };
  auto two = [&]() {
    take(2);
    launch_sweep2(h->d, sw, A, B, C, D, left->d, partial(m), partial(m + 1), st);
    std::swap(A, C);
    std::swap(B, D);
    m += 2;
  };
  if (deep) {
    two();
    for (; m < m_deep; m += sn.ns) {
      take(sn.ns);
      launch_sweepn(h->d, sn, A, B, C, D, left->d, partial(m), stride, st);
      std::swap(A, C);
      std::swap(B, D);
    }
  }
  while (m < m_sweep)
    two();
  for (; m < M; ++m) {
    take(1);
    launch_hpsi(h->d, B, A, A, left->d, partial(m), true, true, st);
    std::swap(A, B);
  }
  reduce(M);
  k_split_moments<<<(M + kBlock - 1) / kBlock, kBlock, 0, st>>>(mu, split, split + M, M);
  LSQT_HIP(hipGetLastError());
  LSQT_HIP(hipMemcpyAsync(mu_re, split, sizeof(double) * (size_t)M, hipMemcpyDeviceToHost, st));
  LSQT_HIP(hipMemcpyAsync(mu_im, split + M, sizeof(double) * (size_t)M, hipMemcpyDeviceToHost,
                          st));
  LSQT_HIP(hipStreamSynchronize(st));
  return LSQT_OK;
}
