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

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace derek
