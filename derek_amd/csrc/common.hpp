// Shared definitions of the gfx950 library: complex helpers, error type.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace derek {

struct Error : std::runtime_error {  // HIP/runtime failure -> DEREK_ERR_HIP
  using std::runtime_error::runtime_error;
};
struct ArgError : Error {  // invalid argument -> DEREK_ERR_ARG
  using Error::Error;
};
struct NumericError : Error {  // -> DEREK_ERR_NUMERIC
  using Error::Error;
};
struct StateError : Error {  // -> DEREK_ERR_STATE
  using Error::Error;
};

#define DEREK_HIP(x)                                                              \
  do {                                                                            \
    hipError_t e_ = (x);                                                          \
    if (e_ != hipSuccess)                                                         \
      throw ::derek::Error(std::string("HIP error '") + hipGetErrorString(e_) +   \
                           "' at " + __FILE__ + ":" + std::to_string(__LINE__) + \
                           " (" #x ")");                                          \
  } while (0)

#define DEREK_REQUIRE(cond, msg)                        \
  do {                                                  \
    if (!(cond)) throw ::derek::ArgError(std::string(msg)); \
  } while (0)

// Complex numbers are HIP's 2-vectors: x = real part, y = imaginary part.
template <typename R> struct Cx;
template <> struct Cx<double> { using T = double2; };
template <> struct Cx<float> { using T = float2; };
template <typename R> using cx_t = typename Cx<R>::T;

template <typename C>
__device__ __forceinline__ C cmk(decltype(C::x) a, decltype(C::x) b) {
  C r;
  r.x = a;
  r.y = b;
  return r;
}
// Element at a 32-bit byte offset from a wave-uniform base pointer: the
// access compiles to a scalar base plus a 32-bit vector offset, with no
// 64-bit address arithmetic per lane.
template <typename T> __device__ __forceinline__ T ld_at(const T* base, uint32_t boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T> __device__ __forceinline__ void st_at(T* base, uint32_t boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}
// Raw buffer access (CDNA buffer resources): a wave-uniform base and a byte
// size; a load at an offset outside [0, size) returns zero and a store there
// is dropped by the hardware, so masked-out rows take the offset kOob
// instead of a branch.  Sizes stay below kOob.
constexpr uint32_t kOob = 0x7FFF0000u;
#ifdef __HIPCC__
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), (short)0, (int)bytes, 0x00020000);
}
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2_t __attribute__((ext_vector_type(2)));
// CP is the cache policy: 0 for ordinary accesses, kNt (the non-temporal bit)
// for streaming data that a kernel reads or writes once.
constexpr int kNt = 2;
template <typename C, int CP = 0> __device__ __forceinline__ C buf_ld(__amdgpu_buffer_rsrc_t r, uint32_t off) {
  if constexpr (sizeof(C) == 16) {
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, CP));
  } else {
    static_assert(sizeof(C) == 8, "8- or 16-byte elements");
    return __builtin_bit_cast(C, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, CP));
  }
}
template <typename C, int CP = 0> __device__ __forceinline__ void buf_st(__amdgpu_buffer_rsrc_t r, uint32_t off, C v) {
  if constexpr (sizeof(C) == 16) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4_t, v), r, (int)off, 0, CP);
  } else {
    static_assert(sizeof(C) == 8, "8- or 16-byte elements");
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2_t, v), r, (int)off, 0, CP);
  }
}
#endif
template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return cmk<C>(a.x + b.x, a.y + b.y); }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return cmk<C>(a.x - b.x, a.y - b.y); }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return cmk<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename C> __device__ __forceinline__ C cscale(C a, decltype(C::x) s) { return cmk<C>(a.x * s, a.y * s); }
// multiply by sign*i
template <int S, typename C> __device__ __forceinline__ C cmul_si(C a) {
  return S > 0 ? cmk<C>(-a.y, a.x) : cmk<C>(a.y, -a.x);
}
template <typename C> __device__ __forceinline__ C cconj(C a) { return cmk<C>(a.x, -a.y); }

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace derek
