// Input file of the `mif` driver (include/InputParser.h,
// src/InputParser.cpp:8-77 of the reference): lines "key : value" with the
// keys Nx, Ny, Nz (global node counts), dt, Nt (time steps), Py, Pz
// (processors per axis) and test_case_2 ("true" selects the second
// lid-driven case).  Lines without ':' are skipped; an unknown or repeated
// key, a missing key, an unparsable number or a file that cannot be opened
// throw std::runtime_error.
#ifndef MIF_SHIM_INPUT_PARSER_H
#define MIF_SHIM_INPUT_PARSER_H

#include <cstddef>
#include <string>

#include "Real.h"

namespace mif {

void parse_input_file(const std::string &filename, size_t &Nx_global, size_t &Ny_global, size_t &Nz_global, Real &dt,
                      unsigned int &num_time_steps, int &Py, int &Pz, bool &test_case_2);

}  // namespace mif

#endif  // MIF_SHIM_INPUT_PARSER_H
