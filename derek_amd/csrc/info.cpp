// .info reader: the reference's key table, line rules, spell check,
// blocks, validation messages, output_level presets and unit conversion.
#include "info.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <sstream>

namespace derek {
namespace info {

namespace {

// ---- units (units.f90:36-128) from the CODATA 2018 constants (constants.f90:61-106)
constexpr double kPi = 3.141592653589793238462643383279502884197;
constexpr double kTwoPi = 2.0 * kPi;
constexpr double c_si = 299792458.0, h_si = 6.62607015e-34, e_si = 1.602176634e-19, me_si = 9.1093837015e-31;
constexpr double na_si = 6.02214076e23, r_si = 8.314462618;
constexpr double hbar_si = h_si / kTwoPi;
constexpr double mu0_si = (4.0 * kPi) * 1.0E-7;
constexpr double eps0_si = 1.0 / (mu0_si * (c_si * c_si));
constexpr double alpha_si = (e_si * e_si) / (4.0 * kPi * eps0_si * hbar_si * c_si);
constexpr double kb_si = r_si / na_si;
constexpr double metre = me_si * c_si * alpha_si / hbar_si;
constexpr double electron_volt = e_si / ((alpha_si * alpha_si) * me_si * (c_si * c_si));
constexpr double joule = 1.0 / ((alpha_si * alpha_si) * me_si * (c_si * c_si));
constexpr double kelvin = kb_si * joule;
constexpr double angstrom = metre * 1e-10;
constexpr double invmetre = 1.0 / metre;

struct Unit {
  const char* name;
  double v;
};
const Unit kUnits[] = {
    {"bohr", 1.0}, {"m", metre}, {"cm", metre * 1e-2}, {"mm", metre * 1e-3}, {"mum", metre * 1e-6},
    {"nm", metre * 1e-9}, {"A", angstrom}, {"pm", metre * 1e-12},
    {"Ha", 1.0}, {"mHa", 1e-3}, {"eV", electron_volt}, {"meV", electron_volt * 1e-3}, {"Ry", 0.5},
    {"mRy", 0.5 * 1e-3}, {"J", joule}, {"K", kelvin},
    {"Ha/Bohr/e", 1.0}, {"eV/A/e", electron_volt / angstrom}, {"N/C", joule / (metre / e_si)},
    {"T", e_si * (invmetre * invmetre) * kTwoPi / h_si}, {"G", 1.0e-4 * e_si * (invmetre * invmetre) * kTwoPi / h_si},
    {"agr", 1.0},
};

std::string lower(std::string s) {
  for (char& ch : s) ch = (char)std::tolower((unsigned char)ch);
  return s;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// list-directed input items: blank- or comma-separated, quotes allowed
std::vector<std::string> items(const std::string& s) {
  std::vector<std::string> out;
  size_t i = 0;
  while (i < s.size()) {
    while (i < s.size() && (std::isspace((unsigned char)s[i]) || s[i] == ',')) ++i;
    if (i >= s.size()) break;
    if (s[i] == '"' || s[i] == '\'') {
      const char q = s[i++];
      const size_t j = s.find(q, i);
      out.push_back(s.substr(i, j == std::string::npos ? std::string::npos : j - i));
      i = j == std::string::npos ? s.size() : j + 1;
    } else {
      size_t j = i;
      while (j < s.size() && !std::isspace((unsigned char)s[j]) && s[j] != ',') ++j;
      out.push_back(s.substr(i, j - i));
      i = j;
    }
  }
  return out;
}

bool parse_logical(const std::string& t, double& v) {
  std::string u = t;
  if (!u.empty() && u[0] == '.') u = u.substr(1);
  if (u.empty()) return false;
  const char c = (char)std::tolower((unsigned char)u[0]);
  if (c == 't') v = 1.0;
  else if (c == 'f') v = 0.0;
  else return false;
  return true;
}

bool parse_int(const std::string& t, double& v) {
  if (t.empty()) return false;
  char* end = nullptr;
  const long long x = std::strtoll(t.c_str(), &end, 10);
  if (*end != '\0') return false;
  v = (double)x;
  return true;
}

bool parse_real(const std::string& t, double& v) {
  if (t.empty()) return false;
  std::string u = t;
  for (char& ch : u)
    if (ch == 'd' || ch == 'D') ch = 'e';
  char* end = nullptr;
  v = std::strtod(u.c_str(), &end);
  return *end == '\0';
}

const std::vector<KeyDef> kKeys = {
    {"check", Kind::Bool, "F", "Boolean", "Check the input and stop before the calculation."},
    {"profilling", Kind::Bool, "F", "Boolean", "Write the routine timing profile."},
    {"n_electrons", Kind::Int, "1", "Integer > 0", "Number of electrons in the cell."},
    {"energy_tol", Kind::Real, "1.0E-7", "Real > 0 (unit_energy)", "SCF stops when successive total energies differ by less."},
    {"xc_functional", Kind::Str, "LDA", "LDA", "Exchange-correlation functional."},
    {"cut_off_energy", Kind::Real, "200.0", "Real > 0 (unit_energy)", "Plane-wave kinetic energy cut-off."},
    {"g_fine_scale", Kind::Real, "2.0", "Real > 1.0", "Fine (density) grid size relative to the wavefunction sphere."},
    {"write_wvfn", Kind::Bool, "F", "Boolean", "Write the wavefunctions to <seed>.wvfn."},
    {"task", Kind::Str, "singlepoint", "singlepoint, bandstructure, dos", "Kind of calculation."},
    {"soc", Kind::Bool, "F", "Boolean", "Spin-orbit coupling."},
    {"max_scf", Kind::Int, "40", "Integer > 0", "Largest number of SCF steps."},
    {"scf_method", Kind::Str, "EDFT", "DM, EDFT", "SCF minimisation scheme."},
    {"electronic_temp", Kind::Real, "0.0", "Real >= 0 (Kelvin)", "Electronic (smearing) temperature, in Kelvin."},
    {"conduction_bands", Kind::Int, "20", "Integer >= 0", "Empty bands above the electrons."},
    {"write_density", Kind::Bool, "F", "Boolean", "Write the density to <seed>.den."},
    {"write_potential", Kind::Bool, "F", "Boolean", "Write the total potential to <seed>.pot."},
    {"write_state", Kind::Bool, "T", "Boolean", "Write the state file."},
    {"external_pot", Kind::Str, "jelly", "jelly, periodic_pot, finite_barrier, <file>.potex",
     "External potential (gaussian_pot is given as a data block)."},
    {"kpt_mp_grid", Kind::Int3, "1 1 1", "Three integers > 0", "Monkhorst-Pack mesh of the SCF."},
    {"write_spec", Kind::Bool, "T", "Boolean", "Write the spectral file."},
    {"write_memory", Kind::Bool, "F", "Boolean", "Write the memory report."},
    {"print_level", Kind::Int, "1", "1, 2, 3", "Verbosity of the output."},
    {"finite_barrier_height", Kind::Real, "10.0", "Real (unit_energy)", "Height of the finite-barrier potential."},
    {"finite_barrier_width", Kind::Real3, "0.5 0.5 0.5", "Three reals (unit_length)", "Widths of the finite barrier."},
    {"periodic_pot_grid", Kind::Int3, "1 1 1", "Three integers > 0", "Periods of the periodic potential."},
    {"periodic_pot_amp", Kind::Real, "10.0", "Real (unit_energy)", "Amplitude of the periodic potential."},
    {"random_seed", Kind::Int, "0", "Any integer", "Seed of the random initial wavefunctions."},
    {"write_formatted_potential", Kind::Bool, "F", "Boolean", "Write the total potential as text."},
    {"write_potex", Kind::Bool, "F", "Boolean", "Write the external potential."},
    {"write_formatted_density", Kind::Bool, "F", "Boolean", "Write the density as text."},
    {"write_formatted_potex", Kind::Bool, "F", "Boolean", "Write the external potential as text."},
    {"output_level", Kind::Str, "minimal", "none, minimal, all", "Preset of the files written."},
    {"ext_efield", Kind::Real3, "0.0 0.0 0.0", "Three reals (unit_efield)", "External electric field."},
    {"ext_bfield", Kind::Real3, "0.0 0.0 0.0", "Three reals (unit_bfield)", "External magnetic field."},
    {"unit_energy", Kind::Str, "eV", "Ha, mHa, eV, meV, Ry, mRy, J", "Unit of energies in the input."},
    {"unit_length", Kind::Str, "A", "m, mm, cm, nm, mum, pm, A, bohr", "Unit of lengths in the input."},
    {"unit_efield", Kind::Str, "eV/A/e", "eV/A/e, Ha/Bohr/e, N/C", "Unit of the electric field in the input."},
    {"unit_bfield", Kind::Str, "T", "T, G, agr", "Unit of the magnetic field in the input."},
    {"unit_energy_out", Kind::Str, "eV", "Ha, mHa, eV, meV, Ry, mRy, J", "Unit of energies in the output."},
    {"unit_length_out", Kind::Str, "A", "m, mm, cm, nm, mum, pm, A, bohr", "Unit of lengths in the output."},
    {"unit_efield_out", Kind::Str, "eV/A/e", "eV/A/e, Ha/Bohr/e, N/C", "Unit of the electric field in the output."},
    {"unit_bfield_out", Kind::Str, "G", "T, G, agr", "Unit of the magnetic field in the output."},
    {"lattice", Kind::Block, "", "$DATA lattice block: 3 rows of 3 reals (unit_length)", "Lattice vectors (rows)."},
    {"dos_kpt_mp_grid", Kind::Int3, "11 11 11", "Three integers > 0", "Monkhorst-Pack mesh of task : dos."},
    {"write_xsf", Kind::Bool, "F", "Boolean", "Write the potentials as XSF files."},
    // extension keys of this build
    {"nbands", Kind::Int, "0", "Integer >= 0", "Scalar (spin-degenerate) bands; 0 = n_electrons/2 + conduction_bands."},
    {"cos_pot_amp", Kind::Real, "0.0", "Real (unit_energy)", "Adds amp*cos(2 pi x / L) to the external potential."},
    {"wfn_seed", Kind::Int, "1", "Integer", "Seed of this build's initial wavefunctions."},
    {"precision", Kind::Str, "fp64", "fp64, mixed", "Arithmetic of the GPU path."},
};

const KeyDef* find_key(const std::string& name) {
  for (const KeyDef& k : kKeys)
    if (name == k.name) return &k;
  return nullptr;
}

// message prefix of a failed value read, as the reference words it per key
std::string parse_msg(const std::string& key) {
  static const char* late[] = {"ext_efield", "ext_bfield", "unit_energy", "unit_length", "unit_efield",
                               "unit_bfield", "unit_energy_out", "unit_length_out", "unit_efield_out",
                               "unit_bfield_out", "dos_kpt_mp_grid", "write_xsf"};
  if (key == "output_level") return "Error parsing value: ";
  for (const char* k : late)
    if (key == k) return "Error in I/O: Error parsing value: ";
  return " Error parsing value: ";
}

bool in(const std::string& v, std::initializer_list<const char*> list) {
  for (const char* x : list)
    if (v == x) return true;
  return false;
}

}  // namespace

const std::vector<KeyDef>& keys() { return kKeys; }
int n_reference_keys() { return 45; }

double to_atomic(double q, const std::string& unit) {
  for (const Unit& u : kUnits)
    if (unit == u.name) return q * u.v;
  throw ArgError("Unknown unit: " + unit);
}

int levenshtein(const std::string& a, const std::string& b) {
  std::vector<std::vector<int>> d(a.size() + 1, std::vector<int>(b.size() + 1, 0));
  for (size_t i = 0; i <= a.size(); ++i) d[i][0] = (int)i;
  for (size_t j = 0; j <= b.size(); ++j) d[0][j] = (int)j;
  for (size_t i = 1; i <= a.size(); ++i)
    for (size_t j = 1; j <= b.size(); ++j)
      d[i][j] = std::min({d[i - 1][j] + 1, d[i][j - 1] + 1, d[i - 1][j - 1] + (a[i - 1] == b[j - 1] ? 0 : 1)});
  return d[a.size()][b.size()];
}

namespace {

void set_default(Params& P, const KeyDef& k) {
  if (k.kind == Kind::Block) return;
  if (k.kind == Kind::Str) {
    P.str[k.name] = k.def;
    return;
  }
  std::vector<double> v;
  for (const std::string& t : items(k.def)) {
    double x = 0.0;
    if (k.kind == Kind::Bool) parse_logical(t, x);
    else parse_real(t, x);
    v.push_back(x);
  }
  P.num[k.name] = v;
}

void set_value(Params& P, const KeyDef& k, const std::string& param) {
  const std::vector<std::string> t = items(param);
  const std::string err = parse_msg(k.name) + param;
  if (k.kind == Kind::Str) {
    if (t.empty()) throw ArgError(err);
    // the reference stores text in character(len=30); file names (external_pot)
    // are kept whole here
    P.str[k.name] = std::string(k.name) == "external_pot" ? t[0] : t[0].substr(0, 30);
    return;
  }
  const size_t n = (k.kind == Kind::Int3 || k.kind == Kind::Real3) ? 3 : 1;
  if (t.size() < n) throw ArgError(err);
  std::vector<double> v(n);
  for (size_t i = 0; i < n; ++i) {
    bool ok = false;
    if (k.kind == Kind::Bool) ok = parse_logical(t[i], v[i]);
    else if (k.kind == Kind::Int || k.kind == Kind::Int3) ok = parse_int(t[i], v[i]);
    else ok = parse_real(t[i], v[i]);
    if (!ok) throw ArgError(err);
  }
  P.num[k.name] = v;
}

// key-specific checks of io_read_param (io.f90:518-745)
void validate(const std::string& key, const std::string& param, const Params& P) {
  const std::string v = P.str.count(key) ? P.str.at(key) : std::string();
  if (key == "task" && !in(v, {"singlepoint", "bandstructure", "dos"})) throw ArgError("Unknown task: " + param);
  if (key == "print_level") {
    const double x = P.num.at(key)[0];
    if (x < 1 || x > 3) throw ArgError("print_level  must be between 1 and 3");
  }
  if (key == "output_level" && !in(v, {"none", "minimal", "all"}))
    throw ArgError("Invalid output: " + param + " for key: " + key);
  if (key == "unit_energy" && !in(v, {"Ha", "mHa", "eV", "meV", "Ry", "mRy", "J"})) throw ArgError("Invalid energy unit");
  if (key == "unit_length" && !in(v, {"m", "mm", "cm", "nm", "mum", "pm", "A", "bohr"}))
    throw ArgError("Invalid length unit");
  if (key == "unit_efield" && !in(v, {"Ha/Bohr/e", "eV/A/e", "N/C"})) throw ArgError("Invalid electric field unit");
  if (key == "unit_bfield" && !in(v, {"T", "G", "agr"})) throw ArgError("Invalid magnetic field unit");
  if (key == "unit_energy_out" && !in(v, {"Ha", "mHa", "eV", "meV", "Ry", "mRy", "J"}))
    throw ArgError("Invalid output energy unit");
  if (key == "unit_length_out" && !in(v, {"m", "mm", "cm", "nm", "mum", "pm", "A", "bohr"}))
    throw ArgError("Invalid output length unit");
  if (key == "unit_efield_out" && !in(v, {"Ha/Bohr/e", "eV/A/e", "N/C"}))
    throw ArgError("Invalid output electric field unit");
  if (key == "unit_bfield_out" && !in(v, {"T", "G", "agr"})) throw ArgError("Invalid out magnetic field unit");
  if (key == "precision" && !in(v, {"fp64", "mixed"})) throw ArgError("Invalid precision: " + param);
  if (key == "external_pot") {
    if (in(v, {"jelly", "finite_barrier", "periodic_pot"})) return;
    if (in(v, {"gaussian_pot", "lorentzian_pot"}))
      throw ArgError("External pot type " + v + " should only be set in a data block.");
    std::ifstream f(v);
    if (!f.good()) throw ArgError("Potential file does not exist: " + v);
    if (v.find("potex") == std::string::npos) throw ArgError("Unsupported external potential file: " + v);
  }
}

// $DATA <type> ... $ENDDATA <type> (io.f90:1765-1935); `lines` holds the
// file, `i` the header line, advanced past the closing line.  Data lines
// whose first character is '#', '!' or a blank are comments there.
void block(Params& P, const std::vector<std::string>& lines, size_t& i) {
  const std::vector<std::string> h = items(lower(trim(lines[i])));
  if (h.size() < 2) throw ArgError(" problem parsing block");
  const std::string type = h[1];
  std::vector<std::string> data;
  size_t j = i + 1;
  for (;; ++j) {
    if (j >= lines.size()) throw ArgError("Block " + type + " not closed");
    const std::string b = lower(trim(lines[j]));
    if (b.find("$enddata") != std::string::npos) {
      const std::vector<std::string> e = items(b);
      if (e.size() < 2) throw ArgError("Problem parsing end block");
      if (e[1] != type) throw ArgError("Block " + type + " not closed");
      break;
    }
    const std::string& raw = lines[j];
    if (raw.empty() || raw[0] == '#' || raw[0] == '!' || raw[0] == ' ' || raw[0] == '\t') continue;
    data.push_back(raw);
  }
  if (data.empty()) throw ArgError("No data in block");
  if (type == "lattice") {
    if (data.size() != 3) throw ArgError("LATTICE must have 3 lines");
    for (int r = 0; r < 3; ++r) {
      const std::vector<std::string> t = items(data[r]);
      double x[3];
      if (t.size() < 3 || !parse_real(t[0], x[0]) || !parse_real(t[1], x[1]) || !parse_real(t[2], x[2]))
        throw ArgError("Error parsing lattice matrix");
      for (int c = 0; c < 3; ++c) P.cell[r + 3 * c] = to_atomic(x[c], P.str.at("unit_length"));
    }
    P.cell_declared = true;
  } else if (type == "gaussian_pot" || type == "lorentzian_pot") {
    P.str["external_pot"] = type;
    P.wells.clear();
    for (const std::string& d : data) {
      const std::vector<std::string> t = items(d);
      double x[5];
      if (t.size() < 5) throw ArgError("Error parsing " + type + " block");
      for (int q = 0; q < 5; ++q)
        if (!parse_real(t[q], x[q])) throw ArgError("Error parsing " + type + " block");
      Well w;
      w.c[0] = x[0];
      w.c[1] = x[1];
      w.c[2] = x[2];
      w.height = to_atomic(x[3], P.str.at("unit_energy"));
      w.width = to_atomic(x[4], P.str.at("unit_length"));
      P.wells.push_back(w);
    }
  } else {
    throw ArgError(" Unknown block type " + type);
  }
  i = j;
}

}  // namespace

Params read(const std::string& path, const std::string& safety) {
  int max_lev = 3;
  if (safety == "Safe") max_lev = 0;
  else if (safety == "Normal") max_lev = 3;
  else if (safety == "Risky") max_lev = 5;
  else throw ArgError("Unknown maximum spell check safety");
  std::ifstream f(path);
  if (!f.good()) throw ArgError(" file '" + path + "' does not exist.");
  std::vector<std::string> lines;
  for (std::string l; std::getline(f, l);) {
    if (!l.empty() && l.back() == '\r') l.pop_back();
    lines.push_back(l);
  }
  Params P;
  for (const KeyDef& k : kKeys) set_default(P, k);
  bool spelling = false;
  for (size_t i = 0; i < lines.size(); ++i) {
    const std::string& line = lines[i];
    if (trim(line).empty()) continue;
    if (line.find('!') != std::string::npos) continue;
    if (lower(line).find("$data") != std::string::npos) {
      const std::string before = P.str["external_pot"];
      block(P, lines, i);
      if (P.str["external_pot"] != before) P.present.push_back("external_pot");
      continue;
    }
    if (line[0] == '#') continue;
    const size_t d = line.find_first_of(":=");
    if (d == std::string::npos) throw ArgError(" Error parsing line:  " + trim(line));
    std::string key = lower(trim(line.substr(0, d)));
    std::string param = trim(line.substr(d + 1));
    if (key.find("unit") == std::string::npos) param = lower(param);
    // nearest key (first in table order on ties)
    int best = 1 << 30;
    std::string match;
    for (const KeyDef& k : kKeys) {
      const int ld = levenshtein(key, k.name);
      if (ld < best) {
        best = ld;
        match = k.name;
      }
    }
    if (best > 0 && best < max_lev) {
      if (!spelling) {
        P.log += "POSSIBLE TYPOS DETECTED IN INFO FILE, ATTEMPTING AUTOCORRECT\n";
        P.log += "  UNKNOWN PARAMETER        AUTOCORRECT       LEVENSHTEIN DISTANCE\n";
        spelling = true;
      }
      char buf[160];
      std::snprintf(buf, sizeof buf, "  %-24s %-24s %3d\n", key.c_str(), match.c_str(), best);
      P.log += buf;
      key = match;
    } else if (best >= max_lev) {
      throw ArgError("Unknown parameter - " + key);
    }
    const KeyDef* k = find_key(key);
    if (!k || k->kind == Kind::Block) throw ArgError(" Error parsing keyword: " + key);
    set_value(P, *k, param);
    validate(key, param, P);
    P.present.push_back(key);
  }
  for (size_t a = 0; a < P.present.size(); ++a)
    for (size_t b = a + 1; b < P.present.size(); ++b)
      if (P.present[a] == P.present[b]) throw ArgError(" Duplicate parameter found: " + P.present[a]);
  // output_level presets (io.f90:778-812)
  if (std::find(P.present.begin(), P.present.end(), "output_level") != P.present.end()) {
    const std::string lvl = P.str["output_level"];
    const bool all = lvl == "all", none = lvl == "none";
    P.num["write_wvfn"] = {all ? 1.0 : 0.0};
    P.num["write_spec"] = {all ? 1.0 : 0.0};
    P.num["write_state"] = {none ? 0.0 : 1.0};
    P.num["write_potex"] = {none ? 0.0 : 1.0};
    P.num["write_formatted_potex"] = {all ? 1.0 : 0.0};
    P.num["write_density"] = {none ? 0.0 : 1.0};
    P.num["write_formatted_density"] = {all ? 1.0 : 0.0};
    P.num["write_potential"] = {all ? 1.0 : 0.0};
    P.num["write_formatted_potential"] = {all ? 1.0 : 0.0};
  }
  if (P.str["task"] == "bandstructure" || P.str["task"] == "dos") P.num["write_spec"] = {1.0};
  // io_convert_all (io.f90:324-358)
  const std::string ue = P.str["unit_energy"], ul = P.str["unit_length"];
  for (const char* k : {"energy_tol", "cut_off_energy", "finite_barrier_height", "periodic_pot_amp", "cos_pot_amp"})
    P.num[k][0] = to_atomic(P.num[k][0], ue);
  P.num["electronic_temp"][0] = to_atomic(P.num["electronic_temp"][0], "K");
  for (double& w : P.num["finite_barrier_width"]) w = to_atomic(w, ul);
  for (double& x : P.num["ext_efield"]) x = to_atomic(x, P.str["unit_efield"]);
  for (double& x : P.num["ext_bfield"]) x = to_atomic(x, P.str["unit_bfield"]);
  return P;
}

void preset_potential(const Params& P, const int fine[3], double* v) {
  const std::string type = P.str.at("external_pot");
  const int64_t nx = fine[0], ny = fine[1], nz = fine[2];
  const int64_t n = nx * ny * nz;
  auto frac = [](int64_t i, int64_t m) { return ((double)(i + 1) * 1.0) / (double)m; };
  if (type == "jelly") {
    // 0.5 * n_electrons / num_fine_grid_points, the count formed as
    // fine_ngx * fine_ngy * fine_ngy (basis.f90:109-111)
    const double val = 0.5 * P.num.at("n_electrons")[0] / (double)(nx * ny * ny);
    for (int64_t p = 0; p < n; ++p) v[p] = val;
  } else if (type == "finite_barrier") {
    const std::vector<double>& w = P.num.at("finite_barrier_width");
    const double b[3] = {w[0] / 2.0, w[1] / 2.0, w[2] / 2.0};
    const double hgt = 0.5 * P.num.at("finite_barrier_height")[0];
    for (int64_t x = 0; x < nx; ++x)
      for (int64_t y = 0; y < ny; ++y)
        for (int64_t z = 0; z < nz; ++z) {
          const double t[3] = {frac(x, nx) - 0.5, frac(y, ny) - 0.5, frac(z, nz) - 0.5};
          const bool inside = std::fabs(t[0]) < b[0] && std::fabs(t[1]) < b[1] && std::fabs(t[2]) < b[2];
          v[(x * ny + y) * nz + z] = inside ? hgt : 0.0;
        }
  } else if (type == "periodic_pot") {
    // every factor uses the x coordinate, as pot.f90:156-163 does
    const std::vector<double>& g = P.num.at("periodic_pot_grid");
    const double amp = P.num.at("periodic_pot_amp")[0];
    for (int64_t x = 0; x < nx; ++x) {
      const double fx = frac(x, nx);
      const double val = 0.5 * amp * std::sin(kTwoPi * g[0] * fx) * std::sin(kTwoPi * g[1] * fx) *
                         std::sin(kTwoPi * g[2] * fx);
      for (int64_t q = 0; q < ny * nz; ++q) v[x * ny * nz + q] = val;
    }
  } else {
    throw ArgError("no preset potential named " + type);
  }
}

std::string help_key(const std::string& key) {
  const KeyDef* k = find_key(lower(key));
  if (!k) return "\n ************** NO MATCHING PARAMETERS **************\n\n";
  std::ostringstream o;
  std::string up = k->name;
  for (char& ch : up) ch = (char)std::toupper((unsigned char)ch);
  o << "\n " << up << "\n " << k->help << "\n\n Allowed Values: " << k->allowed << "\n Default:        "
    << k->def << "\n\n";
  return o.str();
}

std::string search(const std::string& term) {
  std::ostringstream o;
  const std::string t = lower(term);
  bool found = false;
  for (const KeyDef& k : kKeys) {
    if (std::string(k.name).find(t) != std::string::npos || lower(k.help).find(t) != std::string::npos) {
      std::string up = k.name;
      for (char& ch : up) ch = (char)std::toupper((unsigned char)ch);
      char buf[256];
      std::snprintf(buf, sizeof buf, " %-33s %s\n", up.c_str(), k.help);
      o << buf;
      found = true;
    }
  }
  if (!found) o << "\n ************** NO MATCHING PARAMETERS **************\n\n";
  return o.str();
}

std::string list() {
  std::ostringstream o;
  for (const KeyDef& k : kKeys) {
    char buf[320];
    std::snprintf(buf, sizeof buf, " %-27s %-14s %s\n", k.name, k.def, k.help);
    o << buf;
  }
  return o.str();
}

}  // namespace info
}  // namespace derek
