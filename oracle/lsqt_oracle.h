/*
 * lsqt_oracle.h -- CPU restatement of GPUQT's linear-scaling quantum
 * transport path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library supplies the expected results: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only to check (or time beside) the HIP path.  No product code links it.
 *
 * It restates, in plain C, the CPU build of the reference (`lsqt_cpu`,
 * src/makefile.cpu:1-2) step by step, in the same floating-point operation
 * order, so that on the same inputs it reproduces the reference's results
 * bit for bit:
 *   - model construction and RNG draw order   src/model.cu, src/model_lattice.cu,
 *                                              src/model_general.cu, src/anderson.cu,
 *                                              src/charge.cu
 *   - operator kernels (CPU branches)          src/hamiltonian.cu:188-221,273-304,
 *                                              354-383,425-443,531-595,635-648,
 *                                              760-852,927-966
 *   - vectors and the chunked inner product    src/vector.cu:300-328,403-422
 *   - moments, damping, summation, evolution   src/sigma.cu:33-200
 *   - quantity drivers and .out rows           src/sigma.cu:204-442, src/lsqt.cu:135-152
 *
 * Pinning (what checks each part): the reference's own fixed-seed outputs
 * (the .out files of examples/fan2018cpc/lattice/diffusive and .../localized), bit-exact hex
 * dumps from the reference compiled from source (oracle/_ref, built by
 * oracle/Makefile), and closed forms (diagonal H, mu_0 = <r|r>, unitarity,
 * dense eigendecomposition on tiny systems).  The tests under tests/ name
 * the check for every function below.
 *
 * Site loops are parallelised with OpenMP where each site is independent;
 * reductions keep the reference's sequential order (512-site chunks, then
 * chunks in order), so the thread count never changes a bit.
 */
#ifndef LSQT_ORACLE_H
#define LSQT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- random numbers: std::mt19937 + libstdc++ 11 distributions ---------- */
typedef struct {
  uint32_t mt[624];
  int mti;
} orc_mt19937;

void orc_mt_seed(orc_mt19937* g, uint32_t seed);
uint32_t orc_mt_next(orc_mt19937* g);
/* std::uniform_real_distribution<double>(a, b)(g) */
double orc_uniform_real(orc_mt19937* g, double a, double b);
/* std::uniform_int_distribution<int>(0, INT_MAX)(g) */
int orc_uniform_int_0_intmax(orc_mt19937* g);

/* ---- model --------------------------------------------------------------- */
typedef struct {
  /* para.in (src/model.cu:243-310, defaults src/model.h:34-49) */
  int use_lattice_model;
  int calculate_vac0, calculate_vac, calculate_msd, calculate_spin, calculate_ldos;
  int has_anderson;
  double anderson_W;
  int has_charge;
  int charge_Ni;
  double charge_W, charge_xi;
  int has_vacancy;
  int number_of_vacancies;
  int number_of_random_vectors;
  int number_of_moments;
  double energy_max;
  int requires_time;

  /* Hamiltonian data, row-major ELL: pair (n, m) at n*max_neighbor + m */
  int number_of_atoms, max_neighbor, number_of_pairs;
  int* neighbor_number;
  int* neighbor_list;
  double* potential;
  double* hopping_real;
  double* hopping_imag;
  double* xx;
  double volume;

  int number_of_energy_points;
  double* energy;
  int number_of_steps_correlation;
  double* time_step;
  int number_of_local_orbitals;
  int* local_orbitals;

  char input_dir[4096];
  orc_mt19937 generator;
  char error[1024];
} orc_model;

/* Builds the model exactly as Model::Model (src/model.cu:30-65), including
 * all RNG draws of the model phase.  Returns 0, or 1 with model->error set. */
int orc_model_load(const char* dir, orc_model* model);
void orc_model_free(orc_model* model);

/* Model::initialize_state (src/model.cu:77-109): orbital < 0 draws random
 * phases (consuming the model's generator); orbital >= 0 gives sqrt(N) at
 * that orbital. */
void orc_initialize_state(orc_model* model, int orbital, double* re, double* im);

/* ---- operators (row-major ELL, as the reference CPU build) -------------- */
typedef struct {
  int n, max_neighbor;
  double energy_max;
  const int* neighbor_number;
  const int* neighbor_list;
  const double* potential;
  const double* hopping_real;
  const double* hopping_imag;
  const double* xx;
} orc_ham;

void orc_ham_from_model(const orc_model* model, orc_ham* h);

void orc_apply_hamiltonian(const orc_ham* h, const double* ir, const double* ii,
                           double* outr, double* outi);
void orc_apply_commutator(const orc_ham* h, const double* ir, const double* ii,
                          double* outr, double* outi);
void orc_apply_current(const orc_ham* h, const double* ir, const double* ii,
                       double* outr, double* outi);
void orc_kernel_polynomial(const orc_ham* h, const double* s0r, const double* s0i,
                           const double* s1r, const double* s1i, double* s2r, double* s2i);
void orc_chebyshev_01(int n, const double* s0r, const double* s0i, const double* s1r,
                      const double* s1i, double* sr, double* si, double b0, double b1,
                      int direction);
void orc_chebyshev_2(const orc_ham* h, const double* s0r, const double* s0i,
                     const double* s1r, const double* s1i, double* s2r, double* s2i,
                     double* sr, double* si, double bessel_m, int label);
void orc_chebyshev_1x(int n, const double* s1xr, const double* s1xi, double* sr, double* si,
                      double b1);
void orc_chebyshev_2x(const orc_ham* h, const double* s0r, const double* s0i,
                      const double* s0xr, const double* s0xi, const double* s1r,
                      const double* s1i, const double* s1xr, const double* s1xi, double* s2r,
                      double* s2i, double* s2xr, double* s2xi, double* sr, double* si,
                      double bessel_m, int label);

/* sum_n conj(L_n) psi_n in the CPU build's order (src/vector.cu:300-328,403-422) */
void orc_inner_product(int n, const double* psir, const double* psii, const double* Lr,
                       const double* Li, double* out_re, double* out_im);

/* ---- algorithms (src/sigma.cu) ------------------------------------------ */
/* mu_m = <L|T_m(H~)|R>, m = 0..M-1 (src/sigma.cu:33-61) */
void orc_find_moments(const orc_ham* h, int M, const double* Lr, const double* Li,
                      const double* Rr, const double* Ri, double* mu_re, double* mu_im);
/* Jackson damping (src/sigma.cu:76-86) */
void orc_apply_damping(int M, double* mu_re, double* mu_im);
/* Lorentz damping, compile-time LORENTZ variant (src/sigma.cu:64-74) */
void orc_apply_damping_lorentz(int M, double* mu_re, double* mu_im);
/* src/sigma.cu:89-110 */
void orc_chebyshev_summation(int M, const double* mu_re, int n_energy, const double* energy,
                             double energy_max, double volume, double* out);
/* Number of Bessel terms the evolution uses: the first m >= 2 with
 * |J_m(t)| < 1e-15 (src/sigma.cu:129-133). */
int orc_bessel_stop(double time_step_scaled);
/* src/sigma.cu:116-151 and 157-200; state is updated in place */
void orc_evolve(const orc_ham* h, int direction, double time_step_scaled, double* sr,
                double* si);
void orc_evolvex(const orc_ham* h, int direction, double time_step_scaled, double* sr,
                 double* si);

/* ---- quantity drivers ----------------------------------------------------
 * Each returns the damped-and-summed rows (n_rows x N_E, row-major) and,
 * when mu_out is non-NULL, the undamped moments of every row (n_rows x M
 * complex, re then im per row: mu_out[row*2*M + m], mu_out[row*2*M + M + m]).
 * The random_state r is not modified.                                   */
void orc_find_dos(orc_model* model, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu_out);
void orc_find_vac0(orc_model* model, const orc_ham* h, const double* rr, const double* ri,
                   double* rows, double* mu_out);
void orc_find_vac(orc_model* model, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu_out);
void orc_find_msd(orc_model* model, const orc_ham* h, const double* rr, const double* ri,
                  double* rows, double* mu_out);
void orc_find_spin(orc_model* model, const orc_ham* h, const double* rr, const double* ri,
                   double* rows, double* mu_out);

/* The whole lsqt(dir) driver (src/lsqt.cu:135-152): builds the model, runs
 * every random vector and the LDOS loop, and appends rows to the .out files
 * in dir with the reference's text format.  Returns 0 or 1 (err set). */
int orc_lsqt(const char* dir, char* err, int err_len);

/* Writes one .out row exactly as `os << value << " "` ... `<< std::endl`
 * with default stream state (src/sigma.cu:230-233). */
int orc_format_row(const double* values, int n, char* buf, int buf_len);

/* OpenMP thread count actually used by the site loops (1 without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
