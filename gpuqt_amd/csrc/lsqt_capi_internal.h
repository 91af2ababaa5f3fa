// lsqt_capi_internal.h -- handle structs and helpers behind the C ABI
// (internal to liblsqt_hip.so).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "lsqt_device.h"
#include "lsqt_hip.h"

struct lsqt_ctx {
  int device = 0;
  int cu_count = 0;
  hipStream_t stream = nullptr;
  // moment partial sums [M][tiles], moments [M] on device, pinned host copy
  double2* partial = nullptr;
  size_t partial_cap = 0;
  double2* mu_dev = nullptr;
  size_t mu_cap = 0;
  double2* mu_host = nullptr;
  size_t mu_host_cap = 0;
  // Hamiltonians, vectors and blocks that use this context's device and
  // stream; the context cannot be destroyed while any is alive
  int live = 0;
};

struct lsqt_f32_state;    // FP32-path device data (capi_f32.hip)
struct lsqt_block_state;  // block-path scratch and choices (capi_block.hip)

struct lsqt_ham {
  lsqt_ctx* ctx = nullptr;
  lsqt::HamDev d{};
  std::vector<void*> allocations;
  double2* scratch[4] = {nullptr, nullptr, nullptr, nullptr};
  // host copies kept for building the FP32 data on demand
  std::vector<double> host_pot;
  std::vector<double2> host_tab;
  std::vector<double> host_tab_xx;
  lsqt_f32_state* f32 = nullptr;
  lsqt_block_state* blk = nullptr;
  // neighbour-offset structure of a lattice in natural order (0 when none):
  // sites come in slabs of `slab` consecutive sites whose neighbours are at
  // most one slab away, and slabs in rows of `row` sites (slab % row == 0)
  int slab = 0;
  int row = 0;
  // sites whose slot words differ from the slot pattern (d.pat set)
  long long pat_irregular = 0;
  // two Chebyshev steps per launch (k_sweep2) with FP64 and with FP32
  // storage: the slab geometry and the launch shape of each (S = 0: none)
  lsqt::SweepMap sweep64{};
  lsqt::SweepMap sweep32{};
  // the same geometry with more steps per launch (k_sweepn; S = 0: none)
  lsqt::SweepMap sweepn64{};
  lsqt::SweepMap sweepn32{};
  // two steps per launch on the block path (k_block_sweep2): the slab
  // geometry of every system (empty: none) and their slot tables on the
  // device (owned by `allocations`); launch shapes are made per block width
  std::vector<lsqt::BlockSweepSeg> bsweep;
  const int4* bsweep_tab = nullptr;
  int bsweep_steps = lsqt::kBlockSweepSteps;  // most steps per launch (LSQT_SWEEP_STEPS=2: 2)
  // segmented Hamiltonian: independent systems on sites [seg_starts[k],
  // seg_starts[k+1]); empty for one system
  std::vector<int> seg_starts;
};

struct lsqt_vec {
  lsqt_ctx* ctx = nullptr;
  int n = 0;
  double2* d = nullptr;
};

namespace lsqt {

constexpr long long kPackedMaxClasses = 4096;

int set_error(int code, const char* fmt, ...);
int ctx_activate(lsqt_ctx* ctx);
int ensure_device(void** p, size_t* cap, size_t bytes);
int ensure_pinned(void** p, size_t* cap, size_t bytes);
int ctx_reserve_moments(lsqt_ctx* ctx, int M, int tiles);
int ctx_download_moments(lsqt_ctx* ctx, int M, double* mu_re, double* mu_im);
int ham_scratch(lsqt_ham* h, int count);
void ham_release_extra(lsqt_ham* h);
void block_release(lsqt_ham* h);
int f32_potential(lsqt_ham* h);
int moments_enqueue(lsqt_ham* h, const double2* left, const double2* right, int M);
int evolve_enqueue(lsqt_ham* h, int direction, double t, double2* state);
int evolvex_enqueue(lsqt_ham* h, int direction, double t, double2* state);
int moments_f32_enqueue(lsqt_ham* h, const double2* left, const double2* right, int M);
int evolve_f32_enqueue(lsqt_ham* h, int direction, double t, double2* state, int renormalise);
int evolvex_f32_enqueue(lsqt_ham* h, int direction, double t, double2* state);
// resolves the fixed scratch of lsqt_moments_noalloc on `device` (current)
void noalloc_setup(int device);

}  // namespace lsqt

#define LSQT_HIP(call)                                                                   \
  do {                                                                                   \
    hipError_t _e = (call);                                                              \
    if (_e != hipSuccess)                                                                \
      return lsqt::set_error(LSQT_ERR_HIP, "%s failed: %s (%s:%d)", #call,               \
                             hipGetErrorString(_e), __FILE__, __LINE__);                 \
  } while (0)

#define LSQT_TRY(call)       \
  do {                       \
    int _rc = (call);        \
    if (_rc != LSQT_OK)      \
      return _rc;            \
  } while (0)
