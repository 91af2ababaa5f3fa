// postprocess.h -- conductivity from the appended correlation rows: the
// post-processing of the reference's example scripts
// (examples/fan2018cpc/lattice/diffusive/plot_results.m:16-44,
// examples/review/square/diffusive/plot_results.m:33-44), in C++.
//
// Inputs are the rows the drivers append: dos.out (one row of Ne values per
// random vector), vac.out / msd.out (Nt rows per random vector), and the
// time steps of time_step.in.  Averages run over the random vectors; then
//   t_vac(0) = 0, t_vac(k) = t_vac(k-1) + dt(k-1)
//   sigma_vac(k) = 2 pi * trapezoid integral of VAC over t_vac(0..k)
//   sigma_msd(k) = pi * (MSD(k) - MSD(k-1)) / dt(k),   MSD(-1) = 0
// in units of e^2/h (the correlation rows carry the DOS factor), and the
// "true" VAC and MSD are the averages divided by the averaged DOS.
#pragma once

#include <string>
#include <vector>

namespace lsqt_host {

struct Conductivity {
  int Ne = 0;  // energies
  int Nt = 0;  // time steps
  std::vector<double> t_vac;      // [Nt] start time of each VAC row
  std::vector<double> t_msd;      // [Nt] mid-point of each MSD difference
  std::vector<double> sigma_vac;  // [Nt][Ne], empty without VAC rows
  std::vector<double> sigma_msd;  // [Nt][Ne], empty without MSD rows
  std::vector<double> vac;        // [Nt][Ne] averaged VAC / DOS (empty without DOS)
  std::vector<double> msd;        // [Nt][Ne] averaged MSD / DOS (empty without DOS)
};

// Pure computation.  dos: n_dos rows of Ne (may be null), vac / msd: n_vac /
// n_msd rows of Ne (null or 0 rows: skipped), each a whole number of blocks of
// Nt rows.  Throws InputError on inconsistent sizes.
Conductivity conductivity(int Ne, int Nt, const double* time_step, const double* dos, int n_dos,
                          const double* vac, int n_vac, const double* msd, int n_msd);

// Reads dir/energy.in, dir/time_step.in, dir/dos.out and whichever of
// dir/vac.out and dir/msd.out exist, and writes dir/sigma_vac.out and
// dir/sigma_msd.out (Nt rows of Ne values in the .out row format, replacing
// earlier files).  Throws InputError.
Conductivity postprocess_directory(const std::string& dir);

}  // namespace lsqt_host
