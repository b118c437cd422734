// The reference's .info input language (io.f90:360-831, io_block_parse
// :1765-1935, io_convert_all :324-358) and its command-line help
// (io_cl_parser / io_help / io_search / io_list_params, io.f90:994-1673).
#pragma once
#include <array>
#include <map>
#include <string>
#include <vector>

#include "common.hpp"

namespace derek {
namespace info {

enum class Kind { Bool, Int, Real, Str, Int3, Real3, Block };

struct KeyDef {
  const char* name;
  Kind kind;
  const char* def;      // default value as written in the reference's type (io.f90:60-112)
  const char* allowed;  // allowed values (for the help text)
  const char* help;     // one-line description
};

// the reference's 45 keys in its spell-check order (io.f90:1290-1334),
// followed by this build's extension keys
const std::vector<KeyDef>& keys();
int n_reference_keys();

// Levenshtein distance (io_levenshtein, io.f90:2540-2590)
int levenshtein(const std::string& a, const std::string& b);

struct Well {
  double c[3], height, width;
};

struct Params {
  std::map<std::string, std::vector<double>> num;  // Bool / Int / Real / triplets, atomic units
  std::map<std::string, std::string> str;           // Str keys
  std::vector<std::string> present;                 // keys set by the file (incl. external_pot of a block)
  double cell[9] = {0};                             // cell(row, col) column-major, bohr
  bool cell_declared = false;
  std::vector<Well> wells;                          // gaussian_pot block, atomic units
  std::string log;                                  // spell-check table and notes
};

// Reads <path>; safety "Safe" | "Normal" | "Risky" (io.f90:397-407).
// Throws ArgError with the reference's message on any input error.
Params read(const std::string& path, const std::string& safety);

// unit name -> factor to atomic units (units.f90:36-128, 223-360)
double to_atomic(double q, const std::string& unit);

// External potentials the reference builds from keys (pot.f90:130-163) on
// the fine grid (flat, z fastest; fractional coordinates i/n, i = 1..n):
// "jelly", "finite_barrier", "periodic_pot".  The scalar potential is the
// diagonal of the reference's 2x2 potential.
void preset_potential(const Params& P, const int fine[3], double* v);

// help texts for the command line: one key, a search over names and
// descriptions, the full list
std::string help_key(const std::string& key);
std::string search(const std::string& term);
std::string list();

}  // namespace info
}  // namespace derek
