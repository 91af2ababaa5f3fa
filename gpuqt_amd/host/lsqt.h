// lsqt.h -- the per-directory driver (src/lsqt.h:24).
#pragma once

#include <string>

// Runs every calculation para.in asks for in input_directory and appends
// the rows to its .out files.  Throws lsqt_host::InputError on bad input.
void lsqt(std::string input_directory);
