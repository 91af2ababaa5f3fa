/*
 * lsqt_host.h -- C entry points of liblsqt_host.so, the host side of the
 * drop-in (C++ classes Model / Hamiltonian / Vector and the src/sigma.cu
 * drivers, in gpuqt_amd/host/, running on liblsqt_hip.so).
 *
 * Every call returns 0 on success.  Input errors return 1 and copy the
 * reference's message into err (the reference prints it and exits,
 * e.g. src/model.cu:128-310); device errors print a report and exit(1) as
 * the reference's CHECK does (src/common.h:31-42).
 */
#ifndef LSQT_HOST_H
#define LSQT_HOST_H

#include "lsqt_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* main() of the program (src/main.cu:34-56): argv[1] is the driver file,
 * one input directory per non-empty line.  Returns the exit status. */
int lsqt_host_main(int argc, char** argv);

/* lsqt(std::string) (src/lsqt.h:24, src/lsqt.cu:135-152) for one directory */
int lsqt_host_run(const char* input_dir, char* err, int err_len);

/* apply_damping + perform_chebyshev_summation (src/sigma.cu:76-110) on one
 * moment set: mu_re (and mu_im, if not NULL) are damped in place; out gets
 * n_energy values of the correlation function */
int lsqt_host_damped_sum(int M, double* mu_re, double* mu_im, int n_energy,
                         const double* energy, double energy_max, double volume, double* out);

/* Conductivity from the correlation rows (the post-processing of the
 * reference's example scripts, examples/fan2018cpc/lattice/diffusive/
 * plot_results.m:16-44): dos_rows rows of n_energy DOS values, vac_rows /
 * msd_rows rows (whole blocks of n_time rows, one block per random vector;
 * NULL or 0 skips), time_step[n_time].  Outputs [n_time][n_energy] (NULL
 * skips): sigma_vac(k) = 2 pi * trapezoid integral of the averaged VAC over
 * t = 0, dt0, dt0 + dt1, ...; sigma_msd(k) = pi (MSD(k) - MSD(k-1)) / dt(k),
 * MSD(-1) = 0 (units e^2/h); vac_true / msd_true = averages / averaged DOS
 * (need dos_rows > 0). */
int lsqt_host_conductivity(int n_energy, int n_time, const double* time_step, const double* dos,
                           int dos_rows, const double* vac, int vac_rows, const double* msd,
                           int msd_rows, double* sigma_vac, double* sigma_msd, double* vac_true,
                           double* msd_true, char* err, int err_len);
/* the same from input_dir's energy.in, time_step.in, dos.out, vac.out and
 * msd.out, written to input_dir/sigma_vac.out and sigma_msd.out (n_time rows
 * of n_energy values in the .out format, replacing earlier files); also
 * `lsqt_gpu --sigma <driver-file>` for every listed directory */
int lsqt_host_postprocess(const char* input_dir, char* err, int err_len);

/* Writes the general-model input files (read by src/model_general.cu:32-185)
 * into dir: neighbor.in, position.in (box, volume, x[n]), potential.in
 * (skipped if potential is NULL), hopping.in ("complex", or "real" if
 * hopping_imag is NULL), numbers in the shortest form that reads back to the
 * same bits.  The arrays are the Model's row-major ELL arrays. */
int lsqt_host_write_general(const char* dir, int n, int max_neighbor,
                            const int* neighbor_number, const int* neighbor_list,
                            const double* x, double box, double volume, const double* potential,
                            const double* hopping_real, const double* hopping_imag, char* err,
                            int err_len);

/* significant digits of the .out rows (default 6, the reference's) */
void lsqt_host_set_output_precision(int digits);

/* Model (src/model.h:27-104), for inspection: construction with every
 * model-phase random draw, the row-major ELL arrays, the random vectors */
typedef struct lsqt_host_model lsqt_host_model;
int lsqt_host_model_load(const char* input_dir, lsqt_host_model** out, char* err, int err_len);
/* the same with flags: LSQT_HOST_LATTICE_ON_DEVICE keeps a lattice model's
 * description instead of expanding its neighbour arrays (what the drop-in
 * does before building the Hamiltonian on the GPU); lsqt_host_model_arrays
 * then fails, lsqt_host_model_lattice and lsqt_host_model_potential give
 * what lsqt_ham_create_lattices takes */
#define LSQT_HOST_LATTICE_ON_DEVICE 1
int lsqt_host_model_load_ex(const char* input_dir, int flags, lsqt_host_model** out, char* err,
                            int err_len);
void lsqt_host_model_free(lsqt_host_model* m);
/* ints: number_of_atoms, max_neighbor, number_of_moments,
 *       number_of_random_vectors, number_of_energy_points,
 *       number_of_steps_correlation, number_of_local_orbitals,
 *       flags (bit 0 vac0, 1 vac, 2 msd, 3 spin, 4 ldos)
 * reals: energy_max, volume */
int lsqt_host_model_info(const lsqt_host_model* m, int* ints8, double* reals2);
/* copies out the arrays (sizes from lsqt_host_model_info) */
int lsqt_host_model_arrays(const lsqt_host_model* m, int* neighbor_number, int* neighbor_list,
                           double* potential, double* hopping_real, double* hopping_imag,
                           double* xx);
int lsqt_host_model_inputs(const lsqt_host_model* m, double* energy, double* time_step,
                           int* local_orbitals);
/* the description of a model loaded with LSQT_HOST_LATTICE_ON_DEVICE (its
 * arrays stay owned by the model): the lattice, the index of every pristine
 * site after the vacancies (NULL: none) and the pristine site count; 1 when
 * the model has no such description (a general model, or arrays expanded) */
int lsqt_host_model_lattice(const lsqt_host_model* m, lsqt_lattice* lattice,
                            const int** site_index, int* n_pristine);
/* expands the neighbour arrays of such a model on the host after all (the
 * arrays lsqt_host_model_load gives, bit for bit) */
int lsqt_host_model_expand_lattice(lsqt_host_model* m);
/* the on-site energies (number_of_atoms values) */
int lsqt_host_model_potential(const lsqt_host_model* m, double* potential);
/* Model::initialize_state into host arrays (consumes random draws) */
int lsqt_host_model_random_state(lsqt_host_model* m, int orbital, double* re, double* im);

/* The model's Hamiltonian (built on the first of these calls; like
 * Hamiltonian(Model&), src/hamiltonian.cu:27-89, that releases the model's
 * arrays, so lsqt_host_model_arrays fails afterwards) applied by the host
 * classes:
 *   lsqt_host_model_evolve: kind 0 evolve (src/sigma.cu:116-151), kind 1
 *     evolvex (src/sigma.cu:157-200) of the state re/im, replaced by the
 *     result; by_terms != 0 sequences it as the reference does, one
 *     Hamiltonian::chebyshev_2 / chebyshev_2x call per Bessel term with the
 *     states rotated by Vector::swap, by_terms == 0 with the fused entry
 *     points (the same bits);
 *   lsqt_host_model_moments: find_moments_chebyshev (src/sigma.cu:33-61) of
 *     the model's number_of_moments into a device Vector, downloaded to
 *     mu_re / mu_im; by_terms != 0 through Hamiltonian::kernel_polynomial and
 *     Vector::inner_product_1 / inner_product_2, by_terms == 0 fused. */
int lsqt_host_model_evolve(lsqt_host_model* m, int kind, int direction, double time_step_scaled,
                           int by_terms, double* re, double* im);
int lsqt_host_model_moments(lsqt_host_model* m, const double* left_re, const double* left_im,
                            const double* right_re, const double* right_im, int by_terms,
                            double* mu_re, double* mu_im);

#ifdef __cplusplus
}
#endif
#endif
