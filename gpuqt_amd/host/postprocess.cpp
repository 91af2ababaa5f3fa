// postprocess.cpp -- conductivity from the correlation rows (see
// postprocess.h for the formulas and their source).
#include "postprocess.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <sstream>

#include "common.h"
#include "sigma.h"

#define PI 3.141592653589793

namespace lsqt_host {
namespace {

// "count, then values" files (energy.in, time_step.in)
std::vector<double> read_counted(const std::string& filename)
{
  std::ifstream input(filename);
  if (!input.is_open())
    throw InputError("Error: cannot open " + filename);
  long long count = -1;
  input >> count;
  if (!input || count < 0)
    throw InputError("Error: bad count in " + filename);
  std::vector<double> v((size_t)count);
  for (auto& x : v)
    if (!(input >> x))
      throw InputError("Error: " + filename + " holds fewer values than its count");
  return v;
}

// rows of whitespace-separated numbers; blank lines are skipped.  Returns
// false if the file does not exist.
bool read_rows(const std::string& filename, int Ne, std::vector<double>& data, int& rows)
{
  std::ifstream input(filename);
  rows = 0;
  data.clear();
  if (!input.is_open())
    return false;
  std::string line;
  while (std::getline(input, line)) {
    const char* p = line.c_str();
    char* end = nullptr;
    int count = 0;
    for (;;) {
      errno = 0;
      double v = std::strtod(p, &end);
      if (end == p)
        break;
      data.push_back(v);
      ++count;
      p = end;
    }
    while (*p == ' ' || *p == '\t' || *p == '\r')
      ++p;
    if (*p != '\0')
      throw InputError("Error: unreadable value in " + filename + " row " +
                       std::to_string(rows + 1));
    if (count == 0)
      continue;
    if (count != Ne)
      throw InputError("Error: " + filename + " row " + std::to_string(rows + 1) + " has " +
                       std::to_string(count) + " values for " + std::to_string(Ne) +
                       " energies");
    ++rows;
  }
  return true;
}

void write_table(const std::string& filename, const std::vector<double>& v, int Nt, int Ne)
{
  std::ofstream output(filename, std::ios::trunc);
  if (!output.is_open())
    throw InputError("Error: cannot open " + filename);
  output.precision(output_precision());
  for (int t = 0; t < Nt; ++t) {
    for (int e = 0; e < Ne; ++e)
      output << v[(size_t)t * Ne + e] << " ";
    output << std::endl;
  }
}

// average of `blocks` blocks of Nt rows of Ne values
std::vector<double> block_average(const double* rows, int blocks, int Nt, int Ne)
{
  std::vector<double> avg((size_t)Nt * Ne, 0.0);
  for (int b = 0; b < blocks; ++b)
    for (size_t i = 0; i < avg.size(); ++i)
      avg[i] += rows[(size_t)b * Nt * Ne + i];
  for (double& x : avg)
    x /= blocks;
  return avg;
}

}  // namespace

Conductivity conductivity(int Ne, int Nt, const double* time_step, const double* dos, int n_dos,
                          const double* vac, int n_vac, const double* msd, int n_msd)
{
  if (Ne <= 0 || Nt <= 0 || !time_step)
    throw InputError("Error: the conductivity needs energies and time steps");
  if (!vac)
    n_vac = 0;
  if (!msd)
    n_msd = 0;
  if (n_vac == 0 && n_msd == 0)
    throw InputError("Error: the conductivity needs VAC or MSD rows");
  for (int rows : {n_vac, n_msd})
    if (rows % Nt != 0)
      throw InputError("Error: " + std::to_string(rows) +
                       " correlation rows are not whole blocks of " + std::to_string(Nt) +
                       " time steps");
  Conductivity c;
  c.Ne = Ne;
  c.Nt = Nt;
  c.t_vac.assign((size_t)Nt, 0.0);
  c.t_msd.assign((size_t)Nt, 0.0);
  double t = 0.0;
  for (int k = 0; k < Nt; ++k) {
    c.t_vac[k] = t;
    t += time_step[k];
    c.t_msd[k] = t - 0.5 * time_step[k];
  }
  std::vector<double> dos_avg;
  if (dos && n_dos > 0) {
    dos_avg.assign((size_t)Ne, 0.0);
    for (int r = 0; r < n_dos; ++r)
      for (int e = 0; e < Ne; ++e)
        dos_avg[e] += dos[(size_t)r * Ne + e];
    for (double& x : dos_avg)
      x /= n_dos;
  }
  if (n_vac > 0) {
    std::vector<double> avg = block_average(vac, n_vac / Nt, Nt, Ne);
    c.sigma_vac.assign((size_t)Nt * Ne, 0.0);
    for (int k = 1; k < Nt; ++k) {
      const double dt = c.t_vac[k] - c.t_vac[k - 1];
      for (int e = 0; e < Ne; ++e)
        c.sigma_vac[(size_t)k * Ne + e] =
          c.sigma_vac[(size_t)(k - 1) * Ne + e] +
          2.0 * PI * dt * 0.5 * (avg[(size_t)(k - 1) * Ne + e] + avg[(size_t)k * Ne + e]);
    }
    if (!dos_avg.empty()) {
      c.vac = avg;
      for (int k = 0; k < Nt; ++k)
        for (int e = 0; e < Ne; ++e)
          c.vac[(size_t)k * Ne + e] /= dos_avg[e];
    }
  }
  if (n_msd > 0) {
    std::vector<double> avg = block_average(msd, n_msd / Nt, Nt, Ne);
    c.sigma_msd.assign((size_t)Nt * Ne, 0.0);
    for (int k = 0; k < Nt; ++k)
      for (int e = 0; e < Ne; ++e) {
        const double prev = k > 0 ? avg[(size_t)(k - 1) * Ne + e] : 0.0;
        c.sigma_msd[(size_t)k * Ne + e] = PI * (avg[(size_t)k * Ne + e] - prev) / time_step[k];
      }
    if (!dos_avg.empty()) {
      c.msd = avg;
      for (int k = 0; k < Nt; ++k)
        for (int e = 0; e < Ne; ++e)
          c.msd[(size_t)k * Ne + e] /= dos_avg[e];
    }
  }
  return c;
}

Conductivity postprocess_directory(const std::string& dir)
{
  const std::vector<double> energy = read_counted(dir + "/energy.in");
  const std::vector<double> time_step = read_counted(dir + "/time_step.in");
  const int Ne = (int)energy.size();
  const int Nt = (int)time_step.size();
  std::vector<double> dos, vac, msd;
  int n_dos = 0, n_vac = 0, n_msd = 0;
  read_rows(dir + "/dos.out", Ne, dos, n_dos);
  const bool has_vac = read_rows(dir + "/vac.out", Ne, vac, n_vac);
  const bool has_msd = read_rows(dir + "/msd.out", Ne, msd, n_msd);
  if (!has_vac && !has_msd)
    throw InputError("Error: neither " + dir + "/vac.out nor " + dir + "/msd.out exists");
  Conductivity c = conductivity(Ne, Nt, time_step.data(), dos.data(), n_dos, vac.data(), n_vac,
                                msd.data(), n_msd);
  if (!c.sigma_vac.empty())
    write_table(dir + "/sigma_vac.out", c.sigma_vac, Nt, Ne);
  if (!c.sigma_msd.empty())
    write_table(dir + "/sigma_msd.out", c.sigma_msd, Nt, Ne);
  return c;
}

}  // namespace lsqt_host
