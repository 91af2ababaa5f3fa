// sigma.cpp -- find_dos / find_vac0 / find_vac / find_msd /
// find_spin_polarization (src/sigma.cu:204-442) on top of the fused device
// entry points: each moment set is one lsqt_moments call (recursion and
// moment products on the GPU, one download of M complex moments), each
// time step one lsqt_evolve / lsqt_evolvex call.  Damping and the Chebyshev
// summation stay on the host as in the reference.
#include "sigma.h"

#include <chrono>
#include <algorithm>
#include <cmath>
#include <fstream>
#include <math.h>
#include <iostream>
#include <string>
#include <vector>

#include "hamiltonian.h"
#include "lsqt_hip.h"
#include "model.h"
#include "vector.h"

#define PI 3.141592653589793

namespace lsqt_host {
static int g_precision = 6;
void set_output_precision(int digits)
{
  g_precision = digits > 0 ? digits : 6;
}
int output_precision()
{
  return g_precision;
}

static Path g_path = Path::f64;
Path path()
{
  return g_path;
}
void set_path(Path p)
{
  g_path = p;
}
Path path_from_environment()
{
  const char* e = getenv("LSQT_PATH");
  if (!e || !*e || std::string(e) == "f64")
    return Path::f64;
  if (std::string(e) == "f32")
    return Path::f32;
  if (std::string(e) == "block")
    return Path::block;
  if (std::string(e) == "batch")
    return Path::batch;
  throw InputError(std::string("Error: LSQT_PATH must be f64, f32, block or batch, not ") + e);
}
}  // namespace lsqt_host

void find_moments_chebyshev(Model& model, Hamiltonian& H, Vector& left, Vector& right,
                            real* mu_re, real* mu_im)
{
  if (lsqt_host::path() == lsqt_host::Path::f64)
    LSQT_CHECK(lsqt_moments(H.handle(), left.handle(), right.handle(), model.number_of_moments,
                            mu_re, mu_im));
  else
    LSQT_CHECK(lsqt_moments_f32(H.handle(), left.handle(), right.handle(),
                                model.number_of_moments, mu_re, mu_im));
}

void find_moments_chebyshev(Model& model, Hamiltonian& H, Vector& left, Vector& right,
                            Vector& output)
{
  if (lsqt_host::path() == lsqt_host::Path::f64) {
    LSQT_CHECK(lsqt_moments_to_vec(H.handle(), left.handle(), right.handle(),
                                   model.number_of_moments, output.handle()));
    return;
  }
  const int M = model.number_of_moments;
  std::vector<real> re((size_t)output.size(), 0.0), im((size_t)output.size(), 0.0);
  LSQT_CHECK(lsqt_moments_f32(H.handle(), left.handle(), right.handle(), M, re.data(),
                              im.data()));
  output.copy_from_host(re.data(), im.data());
}

namespace lsqt_host {

namespace {
// (-i)^m-type phase case of Bessel term m (m >= 2) of U(direction t) and of
// [X, U(t)] / [U(-t), X] (src/sigma.cu:136-146, 180-190)
int evolve_label(int m, int direction)
{
  switch (m % 4) {
    case 0: return 1;
    case 2: return 2;
    case 1: return direction == 1 ? 3 : 4;
    default: return direction == 1 ? 4 : 3;
  }
}

int evolvex_label(int m, int direction)
{
  switch (m % 4) {
    case 1: return 3;
    case 3: return 4;
    case 0: return direction == 1 ? 1 : 2;
    default: return direction == 1 ? 2 : 1;
  }
}

// 2 J_m(t) while |J_m(t)| >= 1e-15 (src/sigma.cu:126-130, 171-175)
bool bessel_term(int m, real t, real& coefficient)
{
  const real b = jn(m, t);
  if (b < 1.0e-15 && b > -1.0e-15)
    return false;
  coefficient = 2.0 * b;
  return true;
}
}  // namespace

void find_moments_by_terms(Model& model, Hamiltonian& H, Vector& left, Vector& right,
                           Vector& output)
{
  const int n = model.number_of_atoms, M = model.number_of_moments;
  const int blocks = (n - 1) / 512 + 1;
  Vector partials(blocks * M);
  Vector psi_prev(right), psi(n), psi_next(n);
  psi_prev.inner_product_1(n, left, partials, 0);
  H.apply(psi_prev, psi);
  psi.inner_product_1(n, left, partials, blocks);
  for (int m = 2; m < M; ++m) {
    H.kernel_polynomial(psi_prev, psi, psi_next);
    psi_next.inner_product_1(n, left, partials, m * blocks);
    psi_prev.swap(psi);
    psi.swap(psi_next);
  }
  partials.inner_product_2(n, M, output);
}

void evolve_by_terms(Model& model, int direction, real t, Hamiltonian& H, Vector& state)
{
  const int n = model.number_of_atoms;
  Vector t_prev(state), t_cur(n), t_next(n);
  H.apply(state, t_cur);
  H.chebyshev_01(t_prev, t_cur, state, j0(t), 2.0 * j1(t), direction);
  real bm;
  for (int m = 2; bessel_term(m, t, bm); ++m) {
    H.chebyshev_2(t_prev, t_cur, t_next, state, bm, evolve_label(m, direction));
    t_prev.swap(t_cur);
    t_cur.swap(t_next);
  }
}

void evolvex_by_terms(Model& model, int direction, real t, Hamiltonian& H, Vector& state)
{
  const int n = model.number_of_atoms;
  Vector t_prev(state), x_prev(n), t_cur(n), x_cur(n), t_next(n), x_next(n);
  H.apply(state, t_cur);
  H.apply_commutator(state, x_cur);
  H.chebyshev_1x(x_cur, state, 2.0 * j1(t));
  real bm;
  for (int m = 2; bessel_term(m, t, bm); ++m) {
    H.chebyshev_2x(t_prev, x_prev, t_cur, x_cur, t_next, x_next, state, bm,
                   evolvex_label(m, direction));
    t_prev.swap(t_cur);
    t_cur.swap(t_next);
    x_prev.swap(x_cur);
    x_cur.swap(x_next);
  }
}

// Jackson damping, alpha = 1/(M+1) (doc/gpuqt.tex:266-267), then
// rho(E) = 2/(pi Omega) (mu_0 + 2 sum_k mu_k T_k(E~)) / sqrt(1 - E~^2) / E_max
// from the real parts
void damped_chebyshev_sum(int M, real* mu_re, real* mu_im, int n_energy, const real* energy,
                          real energy_max, real volume, real* out)
{
  for (int k = 0; k < M; ++k) {
    real factor = 1.0 / (M + 1);
    real damping =
      (1 - k * factor) * cos(k * PI * factor) + sin(k * PI * factor) * factor / tan(PI * factor);
    mu_re[k] *= damping;
    if (mu_im)
      mu_im[k] *= damping;
  }
  for (int e = 0; e < n_energy; ++e) {
    real x = energy[e] / energy_max;
    real t0 = 1.0, t1 = x, t2;
    real sum = mu_re[1] * t1;
    for (int k = 2; k < M; ++k) {
      t2 = 2.0 * x * t1 - t0;
      t0 = t1;
      t1 = t2;
      sum += mu_re[k] * t2;
    }
    sum *= 2.0;
    sum += mu_re[0];
    sum *= 2.0 / (PI * volume);
    sum /= sqrt(1.0 - x * x);
    out[e] = sum / energy_max;
  }
}

}  // namespace lsqt_host

void apply_damping(Model& model, real* inner_product_real, real* inner_product_imag)
{
  lsqt_host::damped_chebyshev_sum(model.number_of_moments, inner_product_real,
                                  inner_product_imag, 0, nullptr, model.energy_max,
                                  model.volume, nullptr);
}

void perform_chebyshev_summation(Model& model, real* inner_product_real, real*,
                                 real* correlation_function)
{
  // summation only: damping was applied by apply_damping
  const int M = model.number_of_moments;
  for (int e = 0; e < model.number_of_energy_points; ++e) {
    real x = model.energy[e] / model.energy_max;
    real t0 = 1.0, t1 = x, t2;
    real sum = inner_product_real[1] * t1;
    for (int k = 2; k < M; ++k) {
      t2 = 2.0 * x * t1 - t0;
      t0 = t1;
      t1 = t2;
      sum += inner_product_real[k] * t2;
    }
    sum *= 2.0;
    sum += inner_product_real[0];
    sum *= 2.0 / (PI * model.volume);
    sum /= sqrt(1.0 - x * x);
    correlation_function[e] = sum / model.energy_max;
  }
}

void evolve(Model&, int direction, real time_step_scaled, Hamiltonian& H, Vector& state)
{
  if (lsqt_host::path() == lsqt_host::Path::f64)
    LSQT_CHECK(lsqt_evolve(H.handle(), direction, time_step_scaled, state.handle()));
  else
    LSQT_CHECK(lsqt_evolve_f32(H.handle(), direction, time_step_scaled, state.handle(), 1));
}

void evolvex(Model&, int direction, real time_step_scaled, Hamiltonian& H, Vector& state)
{
  if (lsqt_host::path() == lsqt_host::Path::f64)
    LSQT_CHECK(lsqt_evolvex(H.handle(), direction, time_step_scaled, state.handle()));
  else
    LSQT_CHECK(lsqt_evolvex_f32(H.handle(), direction, time_step_scaled, state.handle()));
}

namespace {

// one moment set -> one damped, summed row
std::vector<real> correlation_row(Model& model, Hamiltonian& H, Vector& left, Vector& right)
{
  std::vector<real> mu_re((size_t)model.number_of_moments), mu_im((size_t)model.number_of_moments);
  find_moments_chebyshev(model, H, left, right, mu_re.data(), mu_im.data());
  std::vector<real> row((size_t)model.number_of_energy_points);
  lsqt_host::damped_chebyshev_sum(model.number_of_moments, mu_re.data(), mu_im.data(),
                                  model.number_of_energy_points, model.energy,
                                  model.energy_max, model.volume, row.data());
  return row;
}

std::ofstream open_output(Model& model, const std::string& name)
{
  std::ofstream output(model.input_dir + name, std::ios::app);
  if (!output.is_open()) {
    std::cout << "Error: cannot open " + model.input_dir + name << std::endl;
    exit(1);
  }
  output.precision(lsqt_host::output_precision());
  return output;
}

void write_row(std::ofstream& output, const std::vector<real>& row)
{
  for (real v : row)
    output << v << " ";
  output << std::endl;
}

}  // namespace

void find_dos(Model& model, Hamiltonian& H, Vector& random_state, int flag)
{
  std::vector<real> row = correlation_row(model, H, random_state, random_state);
  std::ofstream output = open_output(model, flag == 0 ? "/dos.out" : "/ldos.out");
  write_row(output, row);
}

void find_vac0(Model& model, Hamiltonian& H, Vector& random_state)
{
  Vector state(model.number_of_atoms);
  H.apply_current(random_state, state);
  std::vector<real> row = correlation_row(model, H, state, state);
  std::ofstream output = open_output(model, "/vac0.out");
  write_row(output, row);
}

void find_vac(Model& model, Hamiltonian& H, Vector& random_state)
{
  Vector state_left(random_state);
  Vector state_left_copy(model.number_of_atoms);
  Vector state_right(random_state);
  H.apply_current(state_left, state_right);
  std::ofstream output = open_output(model, "/vac.out");
  for (int m = 0; m < model.number_of_steps_correlation; ++m) {
    std::cout << "- calculating VAC step " << m << std::endl;
    H.apply_current(state_left, state_left_copy);
    write_row(output, correlation_row(model, H, state_right, state_left_copy));
    if (m < model.number_of_steps_correlation - 1) {
      real time_step_scaled = model.time_step[m] * model.energy_max;
      evolve(model, -1, time_step_scaled, H, state_left);
      evolve(model, -1, time_step_scaled, H, state_right);
    }
  }
}

void find_msd(Model& model, Hamiltonian& H, Vector& random_state)
{
  Vector state(random_state);
  Vector state_x(random_state);
  Vector state_copy(model.number_of_atoms);
  real time_step_scaled = model.time_step[0] * model.energy_max;
  evolve(model, 1, time_step_scaled, H, state);
  evolvex(model, 1, time_step_scaled, H, state_x);
  std::ofstream output = open_output(model, "/msd.out");
  for (int m = 0; m < model.number_of_steps_correlation; ++m) {
    std::cout << "- calculating MSD step " << m << std::endl;
    write_row(output, correlation_row(model, H, state_x, state_x));
    if (m < model.number_of_steps_correlation - 1) {
      time_step_scaled = model.time_step[m + 1] * model.energy_max;
      // [X, U^m] |phi> -> [X, U^(m+1)] |phi>
      state_copy.copy(state);
      evolvex(model, 1, time_step_scaled, H, state_copy);
      evolve(model, 1, time_step_scaled, H, state_x);
      state_x.add(state_copy);
      // U^m |phi> -> U^(m+1) |phi>
      evolve(model, 1, time_step_scaled, H, state);
    }
  }
}

void find_spin_polarization(Model& model, Hamiltonian& H, Vector& random_state)
{
  Vector state(random_state);
  Vector state_sz(model.number_of_atoms);
  std::ofstream output = open_output(model, "/S.out");
  for (int m = 0; m < model.number_of_steps_correlation; ++m) {
    std::cout << "- calculating spin step " << m << std::endl;
    state_sz.apply_sz(state);
    write_row(output, correlation_row(model, H, state, state_sz));
    if (m < model.number_of_steps_correlation - 1) {
      real time_step_scaled = model.time_step[m] * model.energy_max;
      evolve(model, 1, time_step_scaled, H, state);
    }
  }
}

namespace lsqt_host {

namespace {

// the rows of R moment sets (mu[r*M + m]) for the first `used` vectors
std::vector<std::vector<real>> rows_of(Model& model, std::vector<real>& mu_re,
                                       std::vector<real>& mu_im, int used)
{
  const int M = model.number_of_moments;
  std::vector<std::vector<real>> rows((size_t)used,
                                      std::vector<real>((size_t)model.number_of_energy_points));
  for (int r = 0; r < used; ++r)
    damped_chebyshev_sum(M, mu_re.data() + (size_t)r * M, mu_im.data() + (size_t)r * M,
                         model.number_of_energy_points, model.energy, model.energy_max,
                         model.volume, rows[r].data());
  return rows;
}

struct BlockHandle {
  lsqt_block* b = nullptr;
  ~BlockHandle() { lsqt_block_destroy(b); }
};

}  // namespace

void find_block_quantities(std::vector<BlockMember>& members, lsqt_ham* h, int R,
                           const std::vector<real>& re, const std::vector<real>& im)
{
  Model& first = *members[0].model;
  const int M = first.number_of_moments;
  const int Nt = first.number_of_steps_correlation;
  const int S = (int)members.size();
  lsqt_ham_info info;
  LSQT_CHECK(lsqt_ham_get_info(h, &info));
  const int n = info.n;
  int used = 0;
  for (auto& mb : members)
    used += mb.used;
  auto t0 = std::chrono::steady_clock::now();
  auto lap = [&](const char* what) {
    auto t1 = std::chrono::steady_clock::now();
    std::cout << "- Time used for finding " << what << " of " << used << " random vectors in "
              << S << (S == 1 ? " system" : " systems") << " = "
              << std::chrono::duration<double>(t1 - t0).count() << " s" << std::endl;
    t0 = t1;
  };
  // the rows of member k from moment sets mu[(k*R + r)*M + m]
  auto write_rows = [&](const char* name, std::vector<real>& mu_re, std::vector<real>& mu_im) {
    for (int k = 0; k < S; ++k) {
      Model& model = *members[k].model;
      std::vector<real> sr(mu_re.begin() + (size_t)k * R * M,
                           mu_re.begin() + (size_t)(k + 1) * R * M);
      std::vector<real> si(mu_im.begin() + (size_t)k * R * M,
                           mu_im.begin() + (size_t)(k + 1) * R * M);
      auto rows = rows_of(model, sr, si, members[k].used);
      std::ofstream output = open_output(model, name);
      for (auto& row : rows)
        write_row(output, row);
    }
  };
  std::vector<real> mu_re((size_t)S * R * M), mu_im((size_t)S * R * M);
  BlockHandle phi, tmp;
  LSQT_CHECK(lsqt_block_create(default_context(), n, R, &phi.b));
  LSQT_CHECK(lsqt_block_from_host(phi.b, re.data(), im.data()));
  // DOS: self products (src/sigma.cu:204-239)
  LSQT_CHECK(lsqt_moments_block_self(h, phi.b, M, mu_re.data(), mu_im.data()));
  write_rows("/dos.out", mu_re, mu_im);
  lap("DOS");
  if (first.calculate_vac0) {  // <V phi|T_m|V phi> (src/sigma.cu:243-274)
    LSQT_CHECK(lsqt_block_create(default_context(), n, R, &tmp.b));
    LSQT_CHECK(lsqt_block_apply(h, 2, phi.b, tmp.b));
    LSQT_CHECK(lsqt_moments_block_self(h, tmp.b, M, mu_re.data(), mu_im.data()));
    write_rows("/vac0.out", mu_re, mu_im);
    lap("VAC0");
  }
  if ((first.calculate_vac || first.calculate_msd || (first.calculate_spin && S == 1)) &&
      Nt > 0) {
    std::vector<real> steps((size_t)Nt);
    for (int t = 0; t < Nt; ++t)
      steps[t] = first.time_step[t] * first.energy_max;
    const size_t per_step = (size_t)S * R * M;
    std::vector<real> all_re((size_t)Nt * per_step), all_im((size_t)Nt * per_step);
    for (int q = 0; q < 3; ++q) {
      const bool wanted[3] = {first.calculate_vac, first.calculate_msd,
                              first.calculate_spin && S == 1};
      if (!wanted[q])
        continue;
      static const char* kFile[3] = {"/vac.out", "/msd.out", "/S.out"};
      static const char* kWhat[3] = {"VAC", "MSD", "spin polarization"};
      auto driver = q == 0 ? lsqt_block_vac : q == 1 ? lsqt_block_msd : lsqt_block_spin;
      LSQT_CHECK(driver(h, R, re.data(), im.data(), M, Nt, steps.data(), all_re.data(),
                        all_im.data()));
      for (int k = 0; k < S; ++k) {
        Model& model = *members[k].model;
        // the reference writes vector by vector, each with its Nt rows
        std::vector<std::vector<std::vector<real>>> by_step;
        for (int t = 0; t < Nt; ++t) {
          const size_t at = (size_t)t * per_step + (size_t)k * R * M;
          std::vector<real> sr(all_re.begin() + at, all_re.begin() + at + (size_t)R * M);
          std::vector<real> si(all_im.begin() + at, all_im.begin() + at + (size_t)R * M);
          by_step.push_back(rows_of(model, sr, si, members[k].used));
        }
        std::ofstream output = open_output(model, kFile[q]);
        for (int r = 0; r < members[k].used; ++r)
          for (int t = 0; t < Nt; ++t)
            write_row(output, by_step[t][r]);
      }
      lap(kWhat[q]);
    }
  }
}

void find_block_ldos(Model& model, lsqt_ham* h)
{
  const int M = model.number_of_moments;
  const int n = model.number_of_atoms;
  const int count = model.number_of_local_orbitals;
  std::ofstream output = open_output(model, "/ldos.out");
  for (int first = 0; first < count; first += 256) {
    const int used = std::min(256, count - first);
    int R = 1;
    while (R < used)
      R *= 2;
    // the reference's LDOS vectors: sqrt(N) on one orbital (src/model.cu:79-109)
    std::vector<real> re((size_t)n * R, 0.0), im((size_t)n * R, 0.0);
    for (int r = 0; r < used; ++r)
      model.initialize_state_host(re.data() + (size_t)r * n, im.data() + (size_t)r * n,
                                  model.local_orbitals[first + r]);
    std::vector<real> mu_re((size_t)R * M), mu_im((size_t)R * M);
    BlockHandle phi;
    LSQT_CHECK(lsqt_block_create(default_context(), n, R, &phi.b));
    LSQT_CHECK(lsqt_block_from_host(phi.b, re.data(), im.data()));
    LSQT_CHECK(lsqt_moments_block_self(h, phi.b, M, mu_re.data(), mu_im.data()));
    for (auto& row : rows_of(model, mu_re, mu_im, used))
      write_row(output, row);
  }
}

}  // namespace lsqt_host
