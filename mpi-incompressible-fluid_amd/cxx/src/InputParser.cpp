// Parser of the `mif` driver's input file (src/InputParser.cpp:8-77 of the
// reference): the same keys, the same trimming (trailing blanks of the key,
// leading blanks of the value), first-occurrence-only keys and the same
// error messages.
#include "InputParser.h"

#include <fstream>
#include <map>
#include <stdexcept>

namespace mif {

namespace {

const char *kBlanks = " \t\n\r\f\v";

}  // namespace

void parse_input_file(const std::string &filename, size_t &Nx_global, size_t &Ny_global, size_t &Nz_global, Real &dt,
                      unsigned int &num_time_steps, int &Py, int &Pz, bool &test_case_2) {
  std::ifstream in(filename);
  if (!in.is_open()) throw std::runtime_error("Error opening input file: " + filename);

  // the keys in the reference's order, and whether each was seen
  std::map<std::string, bool> seen{{"Nx", false}, {"Ny", false}, {"Nz", false}, {"dt", false},
                                   {"Nt", false}, {"Py", false}, {"Pz", false}, {"test_case_2", false}};
  std::string line;
  while (std::getline(in, line)) {
    const size_t colon = line.find(':');
    if (colon == std::string::npos) continue;
    std::string key = line.substr(0, colon);
    std::string value = line.substr(colon + 1);
    key.erase(key.find_last_not_of(kBlanks) + 1);
    const size_t first = value.find_first_not_of(kBlanks);
    value.erase(0, first == std::string::npos ? value.size() : first);

    auto it = seen.find(key);
    if (it == seen.end() || it->second) throw std::runtime_error("Unknown or duplicate key: " + key);
    it->second = true;
    // std::stoul / std::stod / std::stoi throw std::invalid_argument or
    // std::out_of_range (both std::exception) on a bad number, as there
    if (key == "Nx")
      Nx_global = std::stoul(value);
    else if (key == "Ny")
      Ny_global = std::stoul(value);
    else if (key == "Nz")
      Nz_global = std::stoul(value);
    else if (key == "dt")
      dt = std::stod(value);
    else if (key == "Nt")
      num_time_steps = static_cast<unsigned int>(std::stoul(value));
    else if (key == "Py")
      Py = std::stoi(value);
    else if (key == "Pz")
      Pz = std::stoi(value);
    else
      test_case_2 = value == "true";
  }
  for (const auto &kv : seen)
    if (!kv.second)
      throw std::runtime_error("Missing one or more inputs. Required inputs: Nx, Ny, Nz, dt, Nt, Py, Pz, test_case_2.");
}

}  // namespace mif
