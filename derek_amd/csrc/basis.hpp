// Host restatement of the reference discretisation (Source/basis.f90,
// Source/io.f90 lattice and k-point routines) and of the benchmark external
// potential.  Built with -ffp-contract=off so every map is bit-identical to
// the reference's operation order.
#pragma once
#include <cstdint>
#include <vector>

namespace derek {
namespace basis {

int prime_fact(int n);
void grid_dims(const double cell[9], double ecut, double g_fine_scale, int std_dims[3], int fine_dims[3]);
void inv_cell(const double cell[9], double inv[9]);  // column-major 3x3, orthorhombic only
double volume(const double cell[9]);
void g_map(const double inv[9], const int dims[3], double* g, int32_t* m);        // N x 3, column-major
void real_map(const double cell[9], const int dims[3], double* r, double* frac);  // N x 3, column-major
int kpoints(const int mp[3], double* kfrac);                                      // nk x 3, column-major
// Plane waves of one k-point in map order; returns the count.  Outputs may
// be null (count only); cap bounds the outputs.
int64_t sphere(const double inv[9], const int dims[3], const double kfrac[3], double ecut, int64_t cap,
               int32_t* fft_index, double* kin, int32_t* m);
void gsq(const double inv[9], const int dims[3], double* out);
void gaussian_pot(const double cell[9], const int dims[3], int nwells, const double* centres,
                  const double* heights, const double* widths, double cos_amp, double* v);

// Position on the standard grid (basis order: x outer, z inner, m from
// -n/2 to n-n/2-1, basis.f90:124-135) of each fine-grid FFT bin; -1 where
// the Miller index lies outside the standard grid.
void std_index(const int fine[3], const int std_dims[3], int64_t n, const int32_t* fft_index, int64_t* out);
}  // namespace basis
}  // namespace derek
