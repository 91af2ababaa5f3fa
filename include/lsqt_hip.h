/*
 * lsqt_hip.h -- C ABI of the MI355X (gfx950) LSQT device library,
 * liblsqt_hip.so.
 *
 * The reference has no FFI: its boundary is the C++ classes Hamiltonian
 * (src/hamiltonian.h:25-54), Vector (src/vector.h:23-47) and the algorithm
 * functions of src/sigma.cu.  Each entry point below replaces one member or
 * function of that interface (cited per entry point); the C++ host classes
 * in gpuqt_amd/host/ keep the reference's names on top of it.
 *
 * Conventions
 *   - Every call returns an int status (LSQT_OK == 0).  The library never
 *     exits the process; lsqt_last_error() gives a message for the calling
 *     thread.  (The reference prints and exit(1)s on any CUDA error,
 *     src/common.h:31-42; the host wrapper keeps that behaviour.)
 *   - Plain pointers and sizes only.  Host arrays are caller-owned and never
 *     freed by the library; device buffers live in opaque handles released
 *     by the matching *_destroy.
 *   - Complex vectors cross the ABI as split real/imag host arrays, as the
 *     reference's Vector::copy_from_host/copy_to_host do; on the device they
 *     are stored interleaved (re, im) so a neighbour gather is one load.
 *   - One HIP stream per context; calls on a context are ordered and not
 *     thread-safe.  Calls that return host data (vector download, moments)
 *     synchronise the stream; everything else is asynchronous.
 *   - FP64 entry points reproduce the reference CPU build's per-site
 *     arithmetic in the same operation order without contraction, so state
 *     vectors are bit-identical to it; only moment sums differ, by their
 *     reduction order.
 */
#ifndef LSQT_HIP_H
#define LSQT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
  LSQT_OK = 0,
  LSQT_ERR_INVALID = 1,  /* bad argument, size mismatch, null handle */
  LSQT_ERR_HIP = 2,      /* a HIP runtime call failed */
  LSQT_ERR_NOMEM = 3,    /* device or host allocation failed */
  LSQT_ERR_NODEVICE = 4, /* no usable gfx950 device */
  LSQT_ERR_UNSUPPORTED = 5
};

/* Hamiltonian storage layouts on the device (lsqt_ham_create layout_hint) */
enum {
  LSQT_LAYOUT_AUTO = 0,   /* packed when the pair values have few distinct
                             triples, full otherwise */
  LSQT_LAYOUT_FULL = 1,   /* column-major int32 index + FP64 re/im/xx per pair */
  LSQT_LAYOUT_PACKED = 2  /* column-major 32-bit word per pair: neighbour
                             index in the low bits, index of the pair's
                             (re, im, xx) triple in a small table above it */
};

/* Operators for lsqt_apply_op */
enum { LSQT_OP_H = 0, LSQT_OP_COMMUTATOR = 1, LSQT_OP_CURRENT = 2 };

typedef struct lsqt_ctx lsqt_ctx;
typedef struct lsqt_ham lsqt_ham;
typedef struct lsqt_vec lsqt_vec;

typedef struct {
  int n;              /* sites */
  int max_neighbor;   /* slots per site */
  int layout;         /* LSQT_LAYOUT_FULL or LSQT_LAYOUT_PACKED */
  int classes;        /* distinct (re, im, xx) triples (packed layout) */
  int index_bits;     /* neighbour-index bits of a packed word */
  double energy_max;
  double bytes_per_site_f64; /* compulsory HBM bytes of one FP64 Chebyshev
                                site-update with this layout (fused moment
                                product included) */
  int slab;           /* detected lattice slab in sites (0: none), see the */
  int row;            /* block tile schedules below; row: the next level    */
  int pattern_period; /* slot-pattern period (0: none): packed words that
                         depend only on the site index modulo the period
                         (up to a shift of the neighbour index) are computed
                         by the moment kernels instead of loaded, except on
                         the irregular sites; LSQT_PATTERN=0 turns it off */
  double pattern_irregular; /* fraction of the sites with irregular words */
  int sweep_slab;     /* slab in sites of the two-step moment sweep (0: the
                         moment sets take one Chebyshev step per launch):
                         every neighbour lies within one slab and within
                         sweep_halo positions of its site, periodically, so
                         two steps advance per launch from LDS; LSQT_SWEEP=0
                         turns it off */
  int sweep_halo;
  int sweep_tiles;    /* workgroups (moment partials) per two-step launch,
                         FP64 states */
  int sweep_tiles_f32;  /* the same with FP32 states */
  double bytes_per_site_sweep_f64; /* compulsory HBM bytes per site of one
                                      two-step launch (two site-updates) */
  double bytes_per_site_sweep_f32;
  int block_sweep_systems; /* systems (segments) whose block moment sets of
                              self and pair products take two steps per
                              launch (0: none; either every system of the
                              Hamiltonian or none) */
  int sweep_steps;         /* Chebyshev steps per launch of the FP64 moment
                              sets' deepest sweep: 4 (k_sweepn), 2 (k_sweep2
                              only: LSQT_SWEEP_STEPS=2 in the environment at
                              creation) or 0 (no sweep); a launch reads and
                              writes bytes_per_site_sweep_f64 per site either
                              way */
  int sweep_steps_f32;     /* the same for FP32 states: 6, 2 or 0 */
  int sweep_tiles_deep;    /* workgroups per four-step launch, FP64 states */
  int block_sweep_steps;   /* most steps per launch of the block path's
                              sweeps (4 up to 16 columns, else 2; 2 with
                              LSQT_SWEEP_STEPS=2; 0: none) */
} lsqt_ham_info;

const char* lsqt_strerror(int status);
const char* lsqt_last_error(void);
/* library version string (build identity) */
const char* lsqt_version(void);

/* ---- context: device + stream --------------------------------------------
 * (no counterpart: the reference uses the CUDA default device and stream) */
int lsqt_ctx_create(int device, lsqt_ctx** out);
int lsqt_ctx_destroy(lsqt_ctx* ctx);
int lsqt_ctx_sync(lsqt_ctx* ctx);
/* the hipStream_t the context launches on, for events recorded by callers */
void* lsqt_ctx_stream(lsqt_ctx* ctx);
int lsqt_device_count(int* count);

/* ---- Hamiltonian ----------------------------------------------------------
 * Replaces Hamiltonian::Hamiltonian(Model&) / initialize_gpu
 * (src/hamiltonian.cu:27-89,124-131): the arrays are the Model's row-major
 * ELL arrays (pair (n, m) at n*max_neighbor + m, slots m >= neighbor_number[n]
 * ignored), uploaded column-major.  Unlike the reference, the caller's
 * arrays are not freed. */
int lsqt_ham_create(lsqt_ctx* ctx, int n, int max_neighbor, const int* neighbor_number,
                    const int* neighbor_list, const double* potential,
                    const double* hopping_real, const double* hopping_imag, const double* xx,
                    double energy_max, int layout_hint, lsqt_ham** out);
/* Hamiltonian::~Hamiltonian (src/hamiltonian.cu:133-150) */
int lsqt_ham_destroy(lsqt_ham* h);

/* A lattice model (the contents of lattice.in, src/model_lattice.cu:165-341)
 * for construction on the device.  Sites n = ((nx*Ny + ny)*Nz + nz)*orbitals
 * + m.  Hopping i of orbital m (i < hop_count[m] <= max_neighbor) goes to
 * cell offset hop_cell[3*(m*max_neighbor + i) + 0..2], orbital
 * hop_orbital[m*max_neighbor + i], with value hop_re + i hop_im at the same
 * index. */
typedef struct {
  int cells[3];
  int pbc[3];
  int transport;               /* 0, 1, 2 */
  double lattice_constant[3];
  int orbitals;
  const double* orbital_x;     /* [orbitals]: x of each orbital in its cell */
  int max_neighbor;
  const int* hop_count;        /* [orbitals] */
  const int* hop_cell;         /* [orbitals][max_neighbor][3] */
  const int* hop_orbital;      /* [orbitals][max_neighbor] */
  const double* hop_re;        /* [orbitals][max_neighbor] */
  const double* hop_im;
} lsqt_lattice;
/* Builds the Hamiltonian of a pristine lattice (no vacancies) on the device
 * (the neighbour expansion of src/model_lattice.cu:342-386): the same
 * neighbour order, open-boundary skips, single periodic wrap and transport
 * offsets x12 = a[transport] * offset[transport] + x[orbital'] - x[m] as the
 * Model's arrays, so every result equals that of lsqt_ham_create on those
 * arrays bit for bit.  potential[n] comes from the host (the disorder
 * draws of the Model).  Packed layout, one class per (orbital, hopping). */
int lsqt_ham_create_lattice(lsqt_ctx* ctx, const lsqt_lattice* lattice, const double* potential,
                            double energy_max, lsqt_ham** out);
/* The same for lattice models with vacancies and for several lattices side
 * by side: the neighbour expansion and the vacancy re-indexing of the Model
 * (src/model_lattice.cu:58-146, 342-386) on the device.  site_index (NULL,
 * or an array with an entry per lattice that may itself be NULL) maps each
 * pristine site ((nx*Ny + ny)*Nz + nz)*orbitals + m of lattice k to its
 * index after the vacancies were removed, -1 for a vacancy
 * (Model::find_new_atom_index): vacancies get no site, neighbours that are
 * vacancies are dropped and the later slots move up, as in
 * Model::add_vacancies.  Lattice k then holds sites [s_k, s_k + n_k) with
 * s_0 = 0 and n_k its sites after vacancies; potential holds all of them in
 * that order (the Model's disorder draws).  segments > 1 makes a segmented
 * Hamiltonian (as lsqt_ham_create_segmented).  Packed layout; the classes
 * are the distinct (re, im, x12) of all (lattice, orbital, hopping) entries.
 * Every result equals that of lsqt_ham_create / lsqt_ham_create_segmented on
 * the Model's arrays bit for bit. */
int lsqt_ham_create_lattices(lsqt_ctx* ctx, int segments, const lsqt_lattice* lattices,
                             const int* const* site_index, const double* potential,
                             double energy_max, lsqt_ham** out);
/* (no counterpart in the reference, which runs one system per input
 * directory)  A block-diagonal Hamiltonian of `segments` independent systems
 * side by side: system k on sites [starts[k], starts[k+1]), starts[0] = 0,
 * starts[segments] = n, every neighbour inside its own system (checked), one
 * energy_max for all.  Packed layout.  The block entry points then return
 * moments per system: S = segments times as many, segment-major.  Used for
 * many disorder realisations of uneven size in one launch per step. */
int lsqt_ham_create_segmented(lsqt_ctx* ctx, int segments, const int* starts, int n,
                              int max_neighbor, const int* neighbor_number,
                              const int* neighbor_list, const double* potential,
                              const double* hopping_real, const double* hopping_imag,
                              const double* xx, double energy_max, lsqt_ham** out);
int lsqt_ham_get_info(const lsqt_ham* h, lsqt_ham_info* info);

/* ---- Vector ---------------------------------------------------------------
 * Vector(int n): zero-filled (src/vector.cu:61-71) */
int lsqt_vec_create(lsqt_ctx* ctx, int n, lsqt_vec** out);
/* Vector(Vector&): deep copy (src/vector.cu:92-105) */
int lsqt_vec_clone(const lsqt_vec* src, lsqt_vec** out);
/* ~Vector (src/vector.cu:107-116) */
int lsqt_vec_destroy(lsqt_vec* v);
int lsqt_vec_size(const lsqt_vec* v, int* n);
/* Vector::copy_from_host (src/vector.cu:199-208) */
int lsqt_vec_from_host(lsqt_vec* v, const double* re, const double* im);
/* Vector::copy_to_host (src/vector.cu:210-219); synchronises */
int lsqt_vec_to_host(const lsqt_vec* v, double* re, double* im);
/* Vector::copy: dst = src (src/vector.cu:148-157) */
int lsqt_vec_copy(lsqt_vec* dst, const lsqt_vec* src);
/* Vector::add: dst += src (src/vector.cu:137-146) */
int lsqt_vec_add(lsqt_vec* dst, const lsqt_vec* src);
/* Vector::apply_sz: dst = (-1)^n src (src/vector.cu:188-197) */
int lsqt_vec_apply_sz(lsqt_vec* dst, const lsqt_vec* src);
/* Vector::swap: exchanges storage (src/vector.cu:221-228) */
int lsqt_vec_swap(lsqt_vec* a, lsqt_vec* b);
/* sum_n conj(left_n) psi_n (Vector::inner_product_1 + inner_product_2,
 * src/vector.cu:331-344,425-436, for one vector pair); synchronises */
int lsqt_vec_dot(const lsqt_vec* left, const lsqt_vec* psi, double* re, double* im);

/* Vector::inner_product_1 (src/vector.cu:331-344): for every block b of 512
 * sites (n in [512 b, 512 b + 512), n < number_of_atoms),
 * target[offset + b] = sum_n v_n conj(other_n).  target must hold offset +
 * ceil(number_of_atoms / 512) entries.  Asynchronous. */
int lsqt_vec_inner_product_1(const lsqt_vec* v, int number_of_atoms, const lsqt_vec* other,
                             lsqt_vec* target, int offset);
/* Vector::inner_product_2 (src/vector.cu:425-436): target[m] = sum_b
 * v[m G + b] over the G = ceil(number_of_atoms / 512) partials of moment m,
 * m < number_of_moments.  Asynchronous. */
int lsqt_vec_inner_product_2(const lsqt_vec* v, int number_of_atoms, int number_of_moments,
                             lsqt_vec* target);
/* The device address of the vector's storage: n interleaved (re, im) FP64
 * pairs.  (The reference exposes its split device arrays as the public
 * members real_part / imag_part, src/vector.h:39-40; this build stores the
 * two parts interleaved, so one address replaces both.)  Valid until the
 * vector is destroyed or swapped. */
int lsqt_vec_device_data(const lsqt_vec* v, void** data);

/* ---- operators ------------------------------------------------------------
 * Hamiltonian::apply: out = H~ in, H~ = H / energy_max (src/hamiltonian.cu:225-237) */
int lsqt_apply(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out);
/* Hamiltonian::apply_commutator: out = [X, H~] in (src/hamiltonian.cu:308-320) */
int lsqt_apply_commutator(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out);
/* Hamiltonian::apply_current: out = V in, not scaled (src/hamiltonian.cu:387-399) */
int lsqt_apply_current(lsqt_ham* h, const lsqt_vec* in, lsqt_vec* out);
/* Hamiltonian::kernel_polynomial: s2 = 2 H~ s1 - s0 (src/hamiltonian.cu:970-984);
 * s2 may be s0 */
int lsqt_kernel_polynomial(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1, lsqt_vec* s2);
/* (no counterpart in the reference: two calls of Hamiltonian::kernel_polynomial,
 * src/hamiltonian.cu:970-984, in one launch)  s2 = 2 H~ s1 - s0, then
 * s3 = 2 H~ s2 - s1, bit-identical to the two calls, on a Hamiltonian with a
 * two-step sweep (lsqt_ham_info.sweep_slab > 0; LSQT_ERR_UNSUPPORTED
 * otherwise); s0, s1, s2 and s3 must be distinct.  The moment sets use the
 * same launch. */
int lsqt_kernel_polynomial_pair(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1,
                                lsqt_vec* s2, lsqt_vec* s3);
/* Hamiltonian::chebyshev_01 (src/hamiltonian.cu:447-460) */
int lsqt_chebyshev_01(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1, lsqt_vec* s,
                      double bessel_0, double bessel_1, int direction);
/* Hamiltonian::chebyshev_2 (src/hamiltonian.cu:599-614); s2 may be s0 */
int lsqt_chebyshev_2(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s1, lsqt_vec* s2,
                     lsqt_vec* s, double bessel_m, int label);
/* Hamiltonian::chebyshev_1x (src/hamiltonian.cu:652-662) */
int lsqt_chebyshev_1x(lsqt_ham* h, const lsqt_vec* s1x, lsqt_vec* s, double bessel_1);
/* Hamiltonian::chebyshev_2x (src/hamiltonian.cu:856-882); s2 may be s0 and
 * s2x may be s0x */
int lsqt_chebyshev_2x(lsqt_ham* h, const lsqt_vec* s0, const lsqt_vec* s0x, const lsqt_vec* s1,
                      const lsqt_vec* s1x, lsqt_vec* s2, lsqt_vec* s2x, lsqt_vec* s,
                      double bessel_m, int label);

/* ---- fused algorithms (src/sigma.cu) --------------------------------------
 * find_moments_chebyshev (src/sigma.cu:33-61): mu_m = <left|T_m(H~)|right>,
 * m = 0..M-1, M >= 2.  The whole recursion runs on the device, each step
 * fused with its moment product; one download of M complex values at the
 * end.  mu_re/mu_im are host arrays of M doubles. */
int lsqt_moments(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                 double* mu_re, double* mu_im);
/* find_moments_chebyshev with its Vector output (src/sigma.cu:33-61):
 * out[m] = mu_m, m < M, left on the device (out holds at least M entries);
 * asynchronous */
int lsqt_moments_to_vec(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                        lsqt_vec* out);
/* (no counterpart in the reference)  lsqt_moments for callers that cannot
 * allocate device memory: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  The same kernels in the same order, so
 * the same moments bit for bit, but every buffer beyond the inputs and
 * outputs (up to four recursion vectors, the moment partials, the moments)
 * lives in one fixed scratch array per GPU, held from the library's first
 * context on that GPU until the process ends: 70 MiB, for up to LSQT_NOALLOC_MAX_SITES
 * sites and LSQT_NOALLOC_MAX_MOMENTS moments; larger inputs are rejected
 * with LSQT_ERR_INVALID.  No hipMalloc / hipFree or other allocating call.
 * All calls on one GPU share its scratch: calls that overlap on one GPU (two
 * contexts on different streams, or threads) race and must be serialised by
 * the caller.  Synchronises the context's stream, like lsqt_moments. */
#define LSQT_NOALLOC_MAX_SITES (1 << 20)
#define LSQT_NOALLOC_MAX_MOMENTS (1 << 16)
int lsqt_moments_noalloc(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                         double* mu_re, double* mu_im);
/* evolve (src/sigma.cu:116-151): state <- U(direction * t) state, where
 * time_step_scaled = dt * energy_max.  Bessel coefficients from the host's
 * libm j0/j1/jn, stopping at the first |J_m| < 1e-15 (m >= 2). */
int lsqt_evolve(lsqt_ham* h, int direction, double time_step_scaled, lsqt_vec* state);
/* evolvex (src/sigma.cu:157-200): state <- [X, U(t)] state (direction +1)
 * or [U(-t), X] state (direction -1) */
int lsqt_evolvex(lsqt_ham* h, int direction, double time_step_scaled, lsqt_vec* state);
/* ---- precision path --------------------------------------------------------
 * (no counterpart in the reference: its only precision switch is the
 * compile-time -DUSE_SP that makes everything FP32, src/common.h:22-26)
 * Same results as lsqt_moments / lsqt_evolve / lsqt_evolvex to FP32
 * accuracy: the Chebyshev states are stored in FP32 (half the bytes per
 * step), every site update is computed in FP64 registers with the FP64
 * hopping values, moments are accumulated in FP64, and the evolved state
 * stays FP64.  Vectors cross the ABI in FP64 as above. */
int lsqt_moments_f32(lsqt_ham* h, const lsqt_vec* left, const lsqt_vec* right, int M,
                     double* mu_re, double* mu_im);
/* renormalise != 0 rescales the state to its norm before the step (U is
 * unitary; this removes the FP32 recursion's drift of the norm) */
int lsqt_evolve_f32(lsqt_ham* h, int direction, double time_step_scaled, lsqt_vec* state,
                    int renormalise);
int lsqt_evolvex_f32(lsqt_ham* h, int direction, double time_step_scaled, lsqt_vec* state);

/* ---- block path --------------------------------------------------------------
 * (no counterpart in the reference, which runs one random vector at a time,
 * src/lsqt.cu:140-150)  A block holds R vectors of n sites (R a power of 2,
 * at most 256) in FP32, site-major; one sparse sweep updates all R.
 * Needs the packed layout.  Host arrays are vector-major: vector r at
 * re[r*n .. r*n + n). */
typedef struct lsqt_block lsqt_block;
int lsqt_block_create(lsqt_ctx* ctx, int n, int R, lsqt_block** out);  /* zero-filled */
int lsqt_block_destroy(lsqt_block* b);
int lsqt_block_from_host(lsqt_block* b, const double* re, const double* im);
int lsqt_block_to_host(const lsqt_block* b, double* re, double* im);
/* mu[r*M + m] = <left_r|T_m(H~)|right_r>, r < R, m < M (M >= 2); on a
 * segmented Hamiltonian mu[(k*R + r)*M + m] for segment k */
int lsqt_moments_block(lsqt_ham* h, const lsqt_block* left, const lsqt_block* right, int M,
                       double* mu_re, double* mu_im);
/* mu[r*M + m] = <v_r|T_m(H~)|v_r> (the DOS moments) in M/2 steps, from the
 * product identities T_{2m} = 2 T_m^2 - T_0 and T_{2m+1} = 2 T_m T_{m+1} - T_1
 * (H Hermitian): step m adds <psi_{m-1}|psi_m> and <psi_m|psi_m>, and no left
 * block is read.  The same numbers as lsqt_moments_block(h, v, v, ...) up to
 * FP32 rounding of the states. */
int lsqt_moments_block_self(lsqt_ham* h, const lsqt_block* v, int M, double* mu_re,
                            double* mu_im);
/* mu[r*M + m] = <left_r|T_m(H~)|right_r> (as lsqt_moments_block) in M/2
 * steps: the recursions of left and right run side by side as R pairs of
 * columns of one 2R-vector block (2R <= 256), and <L|T_{2m}|R> =
 * 2 <T_m L|T_m R> - <L|R>, <L|T_{2m+1}|R> = 2 <T_m L|T_{m+1} R> - <L|T_1|R>
 * (T_m Hermitian): per vector, site and moment 24 bytes move instead of 32
 * (two recursions of 24 bytes per step, two moments per step, no left
 * stream).  The VAC and spin drivers below take this route whenever
 * 2R <= 256. */
int lsqt_moments_block_pair(lsqt_ham* h, const lsqt_block* left, const lsqt_block* right,
                            int M, double* mu_re, double* mu_im);
/* out = op(in) for every vector of a block: op 0 = H~ (apply), 1 = [X, H~]
 * (apply_commutator), 2 = V (apply_current); FP32 arithmetic
 * (src/hamiltonian.cu:225-237, 308-320, 387-399) */
int lsqt_block_apply(lsqt_ham* h, int op, const lsqt_block* in, lsqt_block* out);
/* Time-correlation drivers for R vectors at once (FP32 recursion states,
 * FP64 accumulated and re-normalised evolved states, FP64 moments):
 * re/im hold the random vectors vector-major ([R][n], FP64, as
 * Model::initialize_state makes them), time_step_scaled[t] = dt_t *
 * energy_max, and mu[((t*S + k)*R + r)*M + m] receives moment m of time step
 * t, segment k (S = 1 unless the Hamiltonian is segmented), vector r.
 *   lsqt_block_vac: find_vac (src/sigma.cu:278-329): <U(-t) V phi|T_m|V U(-t) phi>,
 *                   as pair moment sets (lsqt_moments_block_pair) when 2R <= 256
 *   lsqt_block_msd: find_msd (src/sigma.cu:333-393): <[X,U(t)] phi|T_m|[X,U(t)] phi>,
 *                   by the product identities (M/2 recursion steps) */
int lsqt_block_vac(lsqt_ham* h, int R, const double* re, const double* im, int M, int n_steps,
                   const double* time_step_scaled, double* mu_re, double* mu_im);
int lsqt_block_msd(lsqt_ham* h, int R, const double* re, const double* im, int M, int n_steps,
                   const double* time_step_scaled, double* mu_re, double* mu_im);
/*   lsqt_block_spin: find_spin_polarization (src/sigma.cu:395-442):
 *                   <U(t) phi|T_m|S_z U(t) phi> (one system only), as pair
 *                   moment sets when 2R <= 256 */
int lsqt_block_spin(lsqt_ham* h, int R, const double* re, const double* im, int M, int n_steps,
                    const double* time_step_scaled, double* mu_re, double* mu_im);
/* Block moment sets are captured once per shape and buffer set and replayed
 * as hipGraphs (default; also LSQT_GRAPHS=0 in the environment turns it
 * off): enable = 0 launches their kernels one by one. */
int lsqt_block_set_graphs(lsqt_ham* h, int enable);
/* The autotuner: times every kernel shape (W vectors per lane, i.e. R/W lanes
 * per site; rows = sites per workgroup tile; blocked = tile schedule) on these
 * blocks and keeps the fastest for (R, max_neighbor, table placement, size
 * class, lattice slab), also in the cache file $LSQT_TUNE_CACHE (default
 * ~/.cache/lsqt_hip_tune.txt).  force re-tunes; the choice and its time per
 * step are returned.
 * Shapes also differ in depth (1 or 2 passes in flight per wave).
 * Tile schedules (blocked = 0, 1, 2): 0 natural (tile t = sites [t*rows,
 * (t+1)*rows)); 1 blocked, for Hamiltonians whose neighbour offsets show a
 * lattice of slabs (sites [k*slab, (k+1)*slab) for the slowest lattice index
 * k) made of rows: tile (b, k) = rows sites at offset b*rows of slab k, with k
 * fastest, so that the tiles in flight together are the same rows of
 * neighbouring slabs and share their neighbours in cache; 2 blocked with the
 * column walk: a tile of whole rows is swept one chunk of every row at a time,
 * so that neighbours one row apart are also read close together; 3 the
 * sliding window, for lattices of at least 3 rows per slab whose rows are
 * whole chunks of 256*W/R sites (W <= 4, max_neighbor 3, 4 or 6): rows =
 * lattice rows per tile (a divisor of the rows per slab); a tile is one
 * chunk of those rows in one slab, swept row by row with rows y-1 .. y+1 of
 * the recursion block in LDS (neighbours there are read from LDS, all
 * others from global memory). */
int lsqt_block_tune(lsqt_ham* h, const lsqt_block* left, const lsqt_block* right, int force,
                    int* W, int* rows, int* blocked, int* depth, double* us_per_step);
/* fixes the shape for R on this Hamiltonian (W = 0 returns to the tuner);
 * blocked = 1 or 2 needs a detected slab that rows divides, blocked = 2 also
 * tiles of whole rows that passes of 256*W/R sites divide, blocked = 3 the
 * conditions above with rows = lattice rows per tile and depth 1; depth =
 * passes in flight per wave (2: the loads of the next pass are issued before
 * the current one is computed; used for max_neighbor 3, 4, 6) */
int lsqt_block_set_config(lsqt_ham* h, int R, int W, int rows, int blocked, int depth);

/* the number of Chebyshev terms evolve/evolvex use for time_step_scaled */
int lsqt_bessel_terms(double time_step_scaled, int* last_m);

#ifdef __cplusplus
}
#endif
#endif
