// Fortran unformatted sequential records (gfortran layout, big-endian) and
// the reference's potential / density / wavefunction records.
#include "interchange.hpp"

#include <algorithm>
#include <cerrno>
#include <cstdlib>
#include <cstring>

namespace derek {
namespace fio {

namespace {

constexpr int64_t kGfortranMaxSub = 2147483639;  // 2^31 - 9 bytes

inline void be32(int32_t v, unsigned char* out) {
  const uint32_t u = __builtin_bswap32((uint32_t)v);
  std::memcpy(out, &u, 4);
}
inline int32_t from_be32(const unsigned char* in) {
  uint32_t u;
  std::memcpy(&u, in, 4);
  return (int32_t)__builtin_bswap32(u);
}

std::string sys_msg(const std::string& what, const std::string& path) {
  return what + " " + path + ": " + std::strerror(errno);
}

}  // namespace

int64_t max_subrecord() {
  if (const char* s = std::getenv("DEREK_IO_SUBRECORD")) {
    const long long v = std::atoll(s);
    if (v >= 8 && v <= kGfortranMaxSub) return (int64_t)v;
  }
  return kGfortranMaxSub;
}

// ---------------------------------------------------------------- writer
RecordWriter::RecordWriter(const std::string& path, int64_t total_bytes)
    : path_(path), total_(total_bytes), max_(max_subrecord()) {
  DEREK_REQUIRE(total_bytes >= 0, "negative record length");
  f_ = std::fopen(path.c_str(), "wb");
  if (!f_) throw IoError(sys_msg("unable to open", path));
  open_sub();
}

RecordWriter::~RecordWriter() {
  if (f_) std::fclose(f_);
}

// leading marker: -length when more subrecords of the record follow
void RecordWriter::open_sub() {
  sub_len_ = std::min(max_, total_ - done_);
  sub_left_ = sub_len_;
  const bool more = done_ + sub_len_ < total_;
  unsigned char m[4];
  be32((int32_t)(more ? -sub_len_ : sub_len_), m);
  if (std::fwrite(m, 1, 4, f_) != 4) throw IoError(sys_msg("unable to write to", path_));
}

// trailing marker: -length when the subrecord continues an earlier one
void RecordWriter::close_sub() {
  unsigned char m[4];
  be32((int32_t)(first_ ? sub_len_ : -sub_len_), m);
  if (std::fwrite(m, 1, 4, f_) != 4) throw IoError(sys_msg("unable to write to", path_));
  first_ = false;
}

void RecordWriter::put_raw(const unsigned char* be, int64_t n) {
  while (n > 0) {
    if (sub_left_ == 0) {
      if (done_ >= total_) throw ArgError("record longer than declared: " + path_);
      close_sub();
      open_sub();
    }
    const int64_t take = std::min(n, sub_left_);
    if ((int64_t)std::fwrite(be, 1, (size_t)take, f_) != take) throw IoError(sys_msg("unable to write to", path_));
    be += take;
    n -= take;
    sub_left_ -= take;
    done_ += take;
  }
}

void RecordWriter::put_i32(int32_t v) {
  unsigned char b[4];
  be32(v, b);
  put_raw(b, 4);
}

// gfortran's default LOGICAL is 4 bytes, .true. = 1
void RecordWriter::put_logical(bool v) { put_i32(v ? 1 : 0); }

void RecordWriter::put_c128(const double* re_im, int64_t n) {
  constexpr int64_t kChunk = 1 << 15;  // doubles per conversion buffer
  std::vector<uint64_t> buf((size_t)std::min<int64_t>(kChunk, std::max<int64_t>(2 * n, 1)));
  int64_t left = 2 * n;
  while (left > 0) {
    const int64_t c = std::min<int64_t>(left, (int64_t)buf.size());
    std::memcpy(buf.data(), re_im, (size_t)c * 8);
    for (int64_t i = 0; i < c; ++i) buf[i] = __builtin_bswap64(buf[i]);
    put_raw(reinterpret_cast<const unsigned char*>(buf.data()), c * 8);
    re_im += c;
    left -= c;
  }
}

void RecordWriter::put_zero_c128(int64_t n) {
  static const unsigned char zeros[1 << 16] = {};
  int64_t left = 16 * n;
  while (left > 0) {
    const int64_t c = std::min<int64_t>(left, (int64_t)sizeof(zeros));
    put_raw(zeros, c);
    left -= c;
  }
}

void RecordWriter::finish() {
  if (done_ != total_) throw ArgError("record shorter than declared: " + path_);
  close_sub();
  const int rc = std::fclose(f_);
  f_ = nullptr;
  if (rc != 0) throw IoError(sys_msg("unable to write to", path_));
}

// ---------------------------------------------------------------- reader
RecordReader::RecordReader(const std::string& path) : path_(path) {
  f_ = std::fopen(path.c_str(), "rb");
  if (!f_) throw IoError(sys_msg("unable to open", path));
  bool first = true;
  while (true) {
    unsigned char m[4];
    if (std::fread(m, 1, 4, f_) != 4) throw IoError("unable to read record marker of " + path);
    const int32_t head = from_be32(m);
    if (head == INT32_MIN) throw IoError("bad record marker in " + path);
    const int64_t len = head < 0 ? -(int64_t)head : head;
    const long off = std::ftell(f_);
    if (off < 0 || std::fseek(f_, (long)len, SEEK_CUR) != 0) throw IoError("truncated record in " + path);
    if (std::fread(m, 1, 4, f_) != 4) throw IoError("truncated record in " + path);
    const int32_t tail = from_be32(m);
    const int64_t tlen = tail < 0 ? -(int64_t)tail : tail;
    if (tlen != len || (tail < 0) == first) throw IoError("record markers do not match in " + path);
    subs_.push_back({(int64_t)off, len});
    total_ += len;
    first = false;
    if (head >= 0) break;
  }
  cur_ = 0;
  cur_left_ = subs_[0].len;
  if (std::fseek(f_, (long)subs_[0].offset, SEEK_SET) != 0) throw IoError("unable to read " + path);
}

RecordReader::~RecordReader() {
  if (f_) std::fclose(f_);
}

void RecordReader::get_raw(unsigned char* dst, int64_t n) {
  if (pos_ + n > total_) throw IoError("record too short in " + path_);
  while (n > 0) {
    if (cur_left_ == 0) {
      ++cur_;
      cur_left_ = subs_[cur_].len;
      if (std::fseek(f_, (long)subs_[cur_].offset, SEEK_SET) != 0) throw IoError("unable to read " + path_);
    }
    const int64_t take = std::min(n, cur_left_);
    if (dst) {
      if ((int64_t)std::fread(dst, 1, (size_t)take, f_) != take) throw IoError("unable to read " + path_);
      dst += take;
    } else if (std::fseek(f_, (long)take, SEEK_CUR) != 0) {
      throw IoError("unable to read " + path_);
    }
    n -= take;
    cur_left_ -= take;
    pos_ += take;
  }
}

int32_t RecordReader::get_i32() {
  unsigned char b[4];
  get_raw(b, 4);
  return from_be32(b);
}

bool RecordReader::get_logical() { return get_i32() != 0; }

void RecordReader::get_c128(double* re_im, int64_t n) {
  get_raw(reinterpret_cast<unsigned char*>(re_im), 16 * n);
  uint64_t* u = reinterpret_cast<uint64_t*>(re_im);
  for (int64_t i = 0; i < 2 * n; ++i) u[i] = __builtin_bswap64(u[i]);
}

void RecordReader::skip(int64_t bytes) { get_raw(nullptr, bytes); }

void RecordReader::finish() {
  if (pos_ != total_) throw IoError("unread data left in the record of " + path_);
}

}  // namespace fio

// ---------------------------------------------------------------- record layouts
int64_t pot_record_bytes(int64_t nfine) { return 12 + 16 * 4 * nfine + 4; }
int64_t den_record_bytes(int64_t nfine) { return 12 + 16 * 4 * nfine + 4; }
int64_t wvfn_record_bytes(int64_t nstd, int64_t nk, int64_t nb) { return 12 + 16 * 2 * nstd * nk * nb + 12; }

namespace {

void put_dims(fio::RecordWriter& w, const int32_t d[3]) {
  for (int i = 0; i < 3; ++i) w.put_i32(d[i]);
}

// header check of pot_read (pot.f90:326-332): the stored standard-grid
// dimensions must be the current ones
void check_dims(fio::RecordReader& r, const int32_t d[3], const std::string& path) {
  int32_t g[3];
  for (int i = 0; i < 3; ++i) g[i] = r.get_i32();
  if (g[0] != d[0] || g[1] != d[1] || g[2] != d[2])
    throw ArgError("fine grid mismatch: " + path + " holds " + std::to_string(g[0]) + "x" + std::to_string(g[1]) +
                   "x" + std::to_string(g[2]));
}

void check_std(const int32_t d[3]) {
  DEREK_REQUIRE(d && d[0] > 0 && d[1] > 0 && d[2] > 0, "grid dimensions must be positive");
}

}  // namespace

void pot_write(const std::string& path, const int32_t std_dims[3], int64_t nfine, const double* nc_pot,
               bool allocated) {
  check_std(std_dims);
  DEREK_REQUIRE(nfine > 0 && nc_pot, "empty potential");
  fio::RecordWriter w(path, pot_record_bytes(nfine));
  put_dims(w, std_dims);
  w.put_c128(nc_pot, 4 * nfine);
  w.put_logical(allocated);
  w.finish();
}

bool pot_read(const std::string& path, const int32_t std_dims[3], int64_t nfine, double* nc_pot) {
  check_std(std_dims);
  DEREK_REQUIRE(nfine > 0 && nc_pot, "empty potential");
  fio::RecordReader r(path);
  check_dims(r, std_dims, path);
  if (r.length() < pot_record_bytes(nfine)) throw IoError("unable to read .pot file " + path + ": record too short");
  r.get_c128(nc_pot, 4 * nfine);
  return r.get_logical();
}

void den_write(const std::string& path, const int32_t std_dims[3], int64_t nfine, const double* charge,
               const double* spin, bool allocated) {
  check_std(std_dims);
  DEREK_REQUIRE(nfine > 0 && charge, "empty density");
  fio::RecordWriter w(path, den_record_bytes(nfine));
  put_dims(w, std_dims);
  w.put_c128(charge, nfine);
  if (spin)
    w.put_c128(spin, 3 * nfine);
  else
    w.put_zero_c128(3 * nfine);
  w.put_logical(allocated);
  w.finish();
}

bool den_read(const std::string& path, const int32_t std_dims[3], int64_t nfine, double* charge, double* spin) {
  check_std(std_dims);
  DEREK_REQUIRE(nfine > 0 && charge, "empty density");
  fio::RecordReader r(path);
  check_dims(r, std_dims, path);
  if (r.length() < den_record_bytes(nfine)) throw IoError("unable to read .den file " + path + ": record too short");
  r.get_c128(charge, nfine);
  if (spin)
    r.get_c128(spin, 3 * nfine);
  else
    r.skip(16 * 3 * nfine);
  return r.get_logical();
}

namespace {

struct SphereLayout {
  int64_t nstd = 0;
  std::vector<int64_t> kofs;  // first coefficient of k in psi (per band 0) / std_index
};

SphereLayout sphere_layout(const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
                           const int64_t* std_index) {
  check_std(std_dims);
  DEREK_REQUIRE(nk > 0 && nb > 0 && npw && std_index, "empty wavefunction");
  SphereLayout L;
  L.nstd = (int64_t)std_dims[0] * std_dims[1] * std_dims[2];
  L.kofs.resize((size_t)nk + 1, 0);
  for (int32_t k = 0; k < nk; ++k) {
    DEREK_REQUIRE(npw[k] >= 0 && npw[k] <= L.nstd, "plane-wave count outside the standard grid");
    L.kofs[k + 1] = L.kofs[k] + npw[k];
  }
  for (int64_t i = 0; i < L.kofs[nk]; ++i)
    DEREK_REQUIRE(std_index[i] >= 0 && std_index[i] < L.nstd, "plane wave outside the standard grid");
  return L;
}

}  // namespace

void wvfn_write(const std::string& path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
                const int64_t* std_index, const double* psi, bool allocated) {
  const SphereLayout L = sphere_layout(std_dims, nk, nb, npw, std_index);
  DEREK_REQUIRE(psi, "null wavefunction");
  fio::RecordWriter w(path, wvfn_record_bytes(L.nstd, nk, nb));
  put_dims(w, std_dims);
  std::vector<double> col((size_t)(2 * L.nstd));
  // coeff(gvec, kpt, band, spinor): spinor 1 holds the scalar wavefunction
  for (int32_t b = 0; b < nb; ++b) {
    for (int32_t k = 0; k < nk; ++k) {
      std::fill(col.begin(), col.end(), 0.0);
      const int64_t n = npw[k];
      const double* src = psi + 2 * (nb * L.kofs[k] + (int64_t)b * n);
      const int64_t* idx = std_index + L.kofs[k];
      for (int64_t p = 0; p < n; ++p) {
        col[2 * idx[p]] = src[2 * p];
        col[2 * idx[p] + 1] = src[2 * p + 1];
      }
      w.put_c128(col.data(), L.nstd);
    }
  }
  w.put_zero_c128(L.nstd * nk * nb);  // spinor 2
  w.put_i32(nb);
  w.put_i32(nk);
  w.put_logical(allocated);
  w.finish();
}

bool wvfn_read(const std::string& path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
               const int64_t* std_index, double* psi) {
  const SphereLayout L = sphere_layout(std_dims, nk, nb, npw, std_index);
  DEREK_REQUIRE(psi, "null wavefunction");
  fio::RecordReader r(path);
  check_dims(r, std_dims, path);
  if (r.length() != wvfn_record_bytes(L.nstd, nk, nb))
    throw ArgError("wavefunction record of " + path + " does not hold " + std::to_string(nk) + " k-points x " +
                   std::to_string(nb) + " bands on this grid");
  std::vector<double> col((size_t)(2 * L.nstd));
  for (int32_t b = 0; b < nb; ++b) {
    for (int32_t k = 0; k < nk; ++k) {
      r.get_c128(col.data(), L.nstd);
      const int64_t n = npw[k];
      double* dst = psi + 2 * (nb * L.kofs[k] + (int64_t)b * n);
      const int64_t* idx = std_index + L.kofs[k];
      for (int64_t p = 0; p < n; ++p) {
        dst[2 * p] = col[2 * idx[p]];
        dst[2 * p + 1] = col[2 * idx[p] + 1];
      }
    }
  }
  r.skip(16 * L.nstd * nk * nb);  // spinor 2
  const int32_t rb = r.get_i32(), rk = r.get_i32();
  if (rb != nb || rk != nk)
    throw ArgError("wavefunction record of " + path + " holds " + std::to_string(rk) + " k-points x " +
                   std::to_string(rb) + " bands");
  const bool alloc = r.get_logical();
  r.finish();
  return alloc;
}

}  // namespace derek
