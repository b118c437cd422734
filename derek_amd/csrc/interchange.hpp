// DEReK's binary interchange files: Fortran unformatted sequential records
// as gfortran writes them with -fconvert=big-endian (the reference's build
// flag, Makefile:86), holding the derived types that the reference writes
// with its defined unformatted I/O procedures:
//   .pot / .potex  potential  (pot.f90:194-224 pot_write, :303-340 pot_read)
//   .den           elec_den   (density.f90:133-163 density_write)
//   .wvfn          wavefunction (wave.f90:896-926 wave_write)
// A child data transfer statement of a defined I/O procedure continues the
// record of its parent statement (Fortran 2008, 9.6.4.8.3), so each file is
// one record: the three header integers, the arrays and the logicals.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "common.hpp"

namespace derek {

struct IoError : Error {  // file or record format failure -> DEREK_ERR_IO
  using Error::Error;
};

namespace fio {

// gfortran splits a record longer than this many bytes into subrecords
// (its manual, "Unformatted sequential file format"); DEREK_IO_SUBRECORD
// (bytes) lowers it, for tests of the split.
int64_t max_subrecord();

// One record written in order; the total length is known up front so the
// markers of every subrecord can be written as the bytes stream through.
class RecordWriter {
 public:
  RecordWriter(const std::string& path, int64_t total_bytes);
  ~RecordWriter();
  void put_i32(int32_t v);
  void put_logical(bool v);
  // n complex128 values (re, im interleaved doubles), byte-swapped on the way
  void put_c128(const double* re_im, int64_t n);
  void put_zero_c128(int64_t n);
  void finish();  // checks that exactly total_bytes were written

 private:
  void put_raw(const unsigned char* be, int64_t n);  // already big-endian
  void open_sub();
  void close_sub();
  FILE* f_ = nullptr;
  std::string path_;
  int64_t total_, done_ = 0, sub_len_ = 0, sub_left_ = 0, max_;
  bool first_ = true;
};

// One record read in order across its subrecords.
class RecordReader {
 public:
  explicit RecordReader(const std::string& path);
  ~RecordReader();
  int64_t length() const { return total_; }  // payload bytes of the record
  int32_t get_i32();
  bool get_logical();
  void get_c128(double* re_im, int64_t n);
  void skip(int64_t bytes);
  void finish();  // all bytes consumed

 private:
  void get_raw(unsigned char* dst, int64_t n);
  FILE* f_ = nullptr;
  std::string path_;
  int64_t total_ = 0, pos_ = 0;
  struct Sub {
    int64_t offset, len;  // file offset of the payload, payload bytes
  };
  std::vector<Sub> subs_;
  size_t cur_ = 0;
  int64_t cur_left_ = 0;
};

}  // namespace fio

// The reference's record layouts (sizes in bytes):
//   potential:    3 x int32 (std grid dims) | nc_pot(nfine, 2, 2) complex128 | logical
//   elec_den:     3 x int32 | charge(nfine) complex128 | spin(nfine, 3) complex128 | logical
//   wavefunction: 3 x int32 | coeff(nstd, nk, nb, 2) complex128 | nbands int32 | kpts int32 | logical
int64_t pot_record_bytes(int64_t nfine);
int64_t den_record_bytes(int64_t nfine);
int64_t wvfn_record_bytes(int64_t nstd, int64_t nk, int64_t nb);

void pot_write(const std::string& path, const int32_t std_dims[3], int64_t nfine, const double* nc_pot,
               bool allocated);
bool pot_read(const std::string& path, const int32_t std_dims[3], int64_t nfine, double* nc_pot);
void den_write(const std::string& path, const int32_t std_dims[3], int64_t nfine, const double* charge,
               const double* spin, bool allocated);
bool den_read(const std::string& path, const int32_t std_dims[3], int64_t nfine, double* charge, double* spin);
// Sphere coefficients (per k: npw[k] x nb, column-major, k-points one after
// the other) placed at std_index on the standard grid (spinor 1), zeros
// elsewhere and in spinor 2.
void wvfn_write(const std::string& path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
                const int64_t* std_index, const double* psi, bool allocated);
bool wvfn_read(const std::string& path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
               const int64_t* std_index, double* psi);

}  // namespace derek
