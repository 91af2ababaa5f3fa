// lsqt.h -- the per-directory driver (src/lsqt.h:24).
#pragma once

#include <string>
#include <vector>

// Runs every calculation para.in asks for in input_directory and appends
// the rows to its .out files.  Throws lsqt_host::InputError on bad input.
void lsqt(std::string input_directory);

// The same for all the directories at once on the batch path: directories
// with the same moments, time steps, energy_max, flags and vector count run
// together as independent systems of one segmented Hamiltonian (one launch
// per recursion step for all of them); each directory gets its own rows.
void lsqt_batch(const std::vector<std::string>& input_directories);
