// general_io.h -- writer of the general-model input files (the format that
// src/model_general.cu:32-185 reads): neighbor.in, position.in,
// potential.in, hopping.in, in the shortest decimal form that reads back to
// the same bits.
#pragma once

#include <string>

namespace lsqt_host {

// n sites, K slots per site; nl row-major [n][K] (first nn[i] used); x[n]
// the transport coordinate; pot (may be null: no potential.in); hi null
// writes "real" hoppings.  Throws InputError if a file cannot be written.
void write_general(const std::string& dir, int n, int K, const int* nn, const int* nl,
                   const double* x, double box, double volume, const double* pot,
                   const double* hr, const double* hi);

}  // namespace lsqt_host
