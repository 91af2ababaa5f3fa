// capi_f32.hip -- the precision path of the C ABI (include/lsqt_hip.h):
// moments and time evolution with FP32-stored Chebyshev states, FP64
// arithmetic, FP64 moment accumulation and an FP64 evolved state.
#include <hip/hip_runtime.h>

#include <cstring>
#include <utility>
#include <vector>

#include "bessel.h"
#include "lsqt_capi_internal.h"
#include "lsqt_hip.h"
#include "lsqt_launch.h"

using namespace lsqt;

struct lsqt_f32_state {
  float* pot32 = nullptr;
  float2* v[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // A, B, Ax, Bx, left
  double2* norms = nullptr;                                      // [target, current]
};

void lsqt::ham_release_extra(lsqt_ham* h)
{
  block_release(h);
  if (!h->f32)
    return;
  if (h->f32->pot32)
    (void)hipFree(h->f32->pot32);
  for (float2* p : h->f32->v)
    if (p)
      (void)hipFree(p);
  if (h->f32->norms)
    (void)hipFree(h->f32->norms);
  delete h->f32;
  h->f32 = nullptr;
  h->d.pot32 = nullptr;
}

static int f32_prepare(lsqt_ham* h, int nvec)
{
  if (!h->f32)
    h->f32 = new lsqt_f32_state();
  lsqt_f32_state* f = h->f32;
  // FP32 arrays are padded to the pair stride (a multiple of 64 sites) and the
  // padding is zero, so paired kernels may always touch site n | 1
  const size_t n = (size_t)h->d.n;
  const size_t padded = (size_t)h->d.stride;
  if (!f->pot32) {
    std::vector<float> p(padded, 0.0f);
    for (size_t i = 0; i < n; ++i)
      p[i] = (float)h->host_pot[i];
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float) * padded);
    if (e != hipSuccess)
      return set_error(LSQT_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
    f->pot32 = (float*)d;
    LSQT_HIP(hipMemcpy(f->pot32, p.data(), sizeof(float) * padded, hipMemcpyHostToDevice));
    h->d.pot32 = f->pot32;
  }
  for (int i = 0; i < nvec; ++i) {
    if (!f->v[i]) {
      void* d = nullptr;
      hipError_t e = hipMalloc(&d, sizeof(float2) * padded);
      if (e != hipSuccess)
        return set_error(LSQT_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
      f->v[i] = (float2*)d;
      launch_zero32(f->v[i], (int)padded, h->ctx->stream);
    }
  }
  if (!f->norms) {
    void* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(double2) * 2);
    if (e != hipSuccess)
      return set_error(LSQT_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
    f->norms = (double2*)d;
  }
  return LSQT_OK;
}

int lsqt::f32_potential(lsqt_ham* h)
{
  return f32_prepare(h, 0);
}

static int check_f32(lsqt_ham* h, std::initializer_list<const lsqt_vec*> vs)
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

int lsqt::moments_f32_enqueue(lsqt_ham* h, const double2* left, const double2* right, int M)
{
  lsqt_ctx* ctx = h->ctx;
  hipStream_t st = ctx->stream;
  const int n = h->d.n;
  const int tiles = hpsi32_tiles(n);
  // several steps per launch where the lattice allows (the schedule of
  // moments_enqueue)
  const SweepMap& sw = h->sweep32;
  const SweepMap& sn = h->sweepn32;
  const bool sweep = sw.S > 0 && M >= 4;
  const bool deep = sweep && sn.S > 0 && M >= 4 + sn.ns;
  const int stride_tiles =
    std::max(tiles, std::max(sweep ? sw.tiles : 0, deep ? sn.tiles : 0));
  LSQT_TRY(ctx_reserve_moments(ctx, M, stride_tiles));
  LSQT_TRY(f32_prepare(h, 5));
  float2* A = h->f32->v[0];
  float2* B = h->f32->v[1];
  float2* L = h->f32->v[4];
  const size_t stride = (size_t)stride_tiles;
  launch_to_f32(left, L, n, st);
  launch_to_f32(right, A, n, st);
  launch_dot32(A, L, ctx->partial, n, st);
  launch_hpsi32(h->d, A, nullptr, B, L, ctx->partial + stride, false, true, st);
  int m = 2, m_deep0 = 2, m_deep = 2;
  if (sweep) {
    // the outputs go to the buffers of the [X, U] recursion (free during a
    // moment set)
    float2* C = h->f32->v[2];
    float2* D = h->f32->v[3];
    auto two = [&]() {
      launch_sweep2_32(h->d, sw, A, B, C, D, L, ctx->partial + (size_t)m * stride,
                       ctx->partial + (size_t)(m + 1) * stride, st);
      std::swap(A, C);
      std::swap(B, D);
      m += 2;
    };
    if (deep) {
      two();
      m_deep0 = m;
      for (; m + sn.ns - 1 < M; m += sn.ns) {
        launch_sweepn_32(h->d, sn, A, B, C, D, L, ctx->partial + (size_t)m * stride, stride,
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
  for (; m < M; ++m) {
    launch_hpsi32(h->d, B, A, A, L, ctx->partial + (size_t)m * stride, true, true, st);
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

extern "C" int lsqt_moments_f32(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                                double* mu_re, double* mu_im)
{
  LSQT_TRY(check_f32(h, {left, right}));
  if (M < 2)
    return set_error(LSQT_ERR_INVALID, "number of moments %d must be >= 2", M);
  if (!mu_re || !mu_im)
    return set_error(LSQT_ERR_INVALID, "null moment output");
  LSQT_TRY(moments_f32_enqueue(h, left->d, right->d, M));
  return ctx_download_moments(h->ctx, M, mu_re, mu_im);
}

// norms[slot] = |state|^2, reduced on the device
static int norm_into(lsqt_ham* h, const double2* state, int slot)
{
  lsqt_ctx* ctx = h->ctx;
  const int n = h->d.n;
  const int tiles = make_grid(n, 1).tiles;
  LSQT_TRY(ctx_reserve_moments(ctx, 1, tiles));
  launch_norm2(state, ctx->partial, n, ctx->stream);
  launch_reduce_partials(ctx->partial, (size_t)tiles, tiles, 1, h->f32->norms + slot,
                         ctx->stream);
  return LSQT_OK;
}

int lsqt::evolve_f32_enqueue(lsqt_ham* h, int direction, double t, double2* state,
                             int renormalise)
{
  hipStream_t st = h->ctx->stream;
  const int n = h->d.n;
  LSQT_TRY(f32_prepare(h, 5));
  float2* A = h->f32->v[0];
  float2* B = h->f32->v[1];
  if (renormalise)
    LSQT_TRY(norm_into(h, state, 0));
  BesselTerms bt = evolve_terms(direction, t);
  launch_to_f32(state, A, n, st);
  launch_hpsi32(h->d, A, nullptr, B, nullptr, nullptr, false, false, st);
  launch_cheb01_32(B, state, bt.b0, bt.b1, direction, n, st);
  size_t k = 0;
  const SweepMap& sn = h->sweepn32;
  if (sn.S > 0) {
    // sn.ns terms per launch on a periodic lattice, into the buffers of the
    // [X, U] recursion (free here); the rest below, from an even term
    float2* C = h->f32->v[2];
    float2* D = h->f32->v[3];
    for (; k + sn.ns <= bt.bm.size(); k += sn.ns) {
      SweepTerms terms{};
      for (int l = 0; l < sn.ns; ++l) {
        terms.bm[l] = bt.bm[k + l];
        terms.label[l] = bt.label[k + l];
      }
      launch_sweepn_evo_32(h->d, sn, A, B, C, D, state, terms, st);
      std::swap(A, C);
      std::swap(B, D);
    }
  }
  for (; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_cheb2_32(h->d, A, B, A, state, d.bm, d.label, st, d.bm_prev, d.label_prev);
    float2* tmp = A;
    A = B;
    B = tmp;
  }
  if (renormalise) {
    LSQT_TRY(norm_into(h, state, 1));
    launch_renormalise(state, h->f32->norms, n, st);
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

int lsqt::evolvex_f32_enqueue(lsqt_ham* h, int direction, double t, double2* state)
{
  hipStream_t st = h->ctx->stream;
  const int n = h->d.n;
  LSQT_TRY(f32_prepare(h, 5));
  float2 *A = h->f32->v[0], *B = h->f32->v[1], *Ax = h->f32->v[2], *Bx = h->f32->v[3];
  BesselTerms bt = evolvex_terms(direction, t);
  launch_to_f32(state, A, n, st);
  launch_zero32(Ax, n, st);
  launch_hpsi32(h->d, A, nullptr, B, nullptr, nullptr, false, false, st);
  launch_comm32(h->d, A, Bx, st);
  launch_cheb1x_32(Bx, state, bt.b1, n, st);
  for (size_t k = 0; k < bt.bm.size(); ++k) {
    const DeferredTerm d = deferred_term(bt, k);
    launch_cheb2x_32(h->d, A, Ax, B, Bx, A, Ax, state, d.bm, d.label, st, d.bm_prev,
                     d.label_prev);
    float2* tmp = A;
    A = B;
    B = tmp;
    tmp = Ax;
    Ax = Bx;
    Bx = tmp;
  }
  LSQT_HIP(hipGetLastError());
  return LSQT_OK;
}

extern "C" int lsqt_evolve_f32(lsqt_ham* h, int direction, double t, lsqt_vec* state,
                               int renormalise)
{
  LSQT_TRY(check_f32(h, {state}));
  if (direction != 1 && direction != -1)
    return set_error(LSQT_ERR_INVALID, "direction %d must be +1 or -1", direction);
  return evolve_f32_enqueue(h, direction, t, state->d, renormalise);
}

extern "C" int lsqt_evolvex_f32(lsqt_ham* h, int direction, double t, lsqt_vec* state)
{
  LSQT_TRY(check_f32(h, {state}));
  if (direction != 1 && direction != -1)
    return set_error(LSQT_ERR_INVALID, "direction %d must be +1 or -1", direction);
  return evolvex_f32_enqueue(h, direction, t, state->d);
}
