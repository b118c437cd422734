// Timing of the library's tall-skinny complex GEMMs (subspace overlaps and
// rotations) at the block shapes of the SCF workloads: 161214 plane waves x
// 8 k-points with 300 / 600 columns (144^3 grid, 300 bands) and 62669 plane
// waves with 80 / 160 columns (108^3 grid, 80 bands).  Prints, per call,
// the best of five HIP-event timings and the rate in TFLOP/s counted as
// 8 m n npw per complex product (the four-real-product count; Hermitian
// overlaps compute only the upper tiles, so their rate counts the full
// product as well).  complex64 runs on the FP32 matrix cores ("fp32") and
// on the BF16 ones with 1, 2 and 3 bf16 terms per value ("bf16x1" ...).
//
//   derek_amd/gemm_bench [bf16]   (built by make -C derek_amd/csrc)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "launch.hpp"

using namespace derek;

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
    std::exit(1);
  }
}

template <typename C> __global__ void k_fill(C* p, int64_t n) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const uint32_t h = (uint32_t)(i * 2654435761u);
    p[i].x = (float)((int)(h & 1023) - 512) / 1024.0f;
    p[i].y = (float)((int)((h >> 10) & 1023) - 512) / 1024.0f;
  }
}

template <typename R> void run(int64_t npw, int nb, int batch, int terms = 0, int tile = 64) {
  using C = cx_t<R>;
  const int64_t ld = ceil_div(npw, 64) * 64;
  const int m2 = 2 * nb;
  const int64_t ks = ld * m2;
  C *X = nullptr, *Z = nullptr;
  double2 *S = nullptr, *U = nullptr, *work = nullptr;
  check(hipMalloc(&X, sizeof(C) * ks * batch), "hipMalloc");
  check(hipMalloc(&Z, sizeof(C) * ks * batch), "hipMalloc");
  check(hipMalloc(&S, sizeof(double2) * m2 * m2 * batch), "hipMalloc");
  check(hipMalloc(&U, sizeof(double2) * m2 * m2 * batch), "hipMalloc");
  hipLaunchKernelGGL(k_fill<C>, dim3(4096), dim3(256), 0, 0, X, ks * batch);
  hipLaunchKernelGGL(k_fill<double2>, dim3(4096), dim3(256), 0, 0, U, (int64_t)m2 * m2 * batch);
  check(hipGetLastError(), "fill");
  const int64_t cap = std::max(overlap_work_elems(m2, m2, npw, batch), overlap_work_elems(nb, nb, npw, batch));
  check(hipMalloc(&work, sizeof(double2) * cap), "hipMalloc");
  hipEvent_t a, b;
  check(hipEventCreate(&a), "hipEventCreate");
  check(hipEventCreate(&b), "hipEventCreate");
  auto timeit = [&](const char* what, double flops, auto f) {
    f();
    check(hipDeviceSynchronize(), "sync");
    float best = 1e30f;
    for (int r = 0; r < 5; ++r) {
      check(hipEventRecord(a, 0), "hipEventRecord");
      f();
      check(hipEventRecord(b, 0), "hipEventRecord");
      check(hipEventSynchronize(b), "hipEventSynchronize");
      float t = 0.f;
      check(hipEventElapsedTime(&t, a, b), "hipEventElapsedTime");
      best = std::min(best, t);
    }
    char tag[16];
    std::snprintf(tag, sizeof tag, "%s", sizeof(R) == 8 ? "fp64" : "fp32");
    if (terms) std::snprintf(tag, sizeof tag, "bf16x%d%s", terms, tile == 128 ? "/128" : "");
    std::printf("%-6s npw %6lld b%-2d %-24s %9.3f ms %7.1f TF/s\n", tag, (long long)npw, batch, what, best,
                flops / best / 1e9);
    std::fflush(stdout);
  };
  char w[64];
  for (int m : {nb, m2}) {
    std::snprintf(w, sizeof w, "overlap herm %dx%d", m, m);
    timeit(w, 8.0 * m * m * npw * batch, [&] {
      overlap<R>(X, ld, ks, X, ld, ks, m, m, npw, nullptr, S, m, (int64_t)m * m, batch, work, cap, true, 0, terms, 0,
                 2, tile);
    });
    std::snprintf(w, sizeof w, "overlap full %dx%d", m, m);
    timeit(w, 8.0 * m * m * npw * batch, [&] {
      overlap<R>(X, ld, ks, X, ld, ks, m, m, npw, nullptr, S, m, (int64_t)m * m, batch, work, cap, false, 0, terms, 0,
                 2, tile);
    });
    std::snprintf(w, sizeof w, "rotate %d->%d", m, nb);
    timeit(w, 8.0 * m * nb * npw * batch, [&] {
      rotate<R>(X, ld, ks, U, m2, (int64_t)m2 * m2, m, nb, npw, nullptr, Z, ld, ks, batch, 0, terms);
    });
  }
  check(hipFree(X), "hipFree");
  check(hipFree(Z), "hipFree");
  check(hipFree(S), "hipFree");
  check(hipFree(U), "hipFree");
  check(hipFree(work), "hipFree");
}

int main(int argc, char** argv) {
  // argument "bf16": the complex64 runs only, FP32 then BF16 with 1..3 terms,
  // then BF16 overlaps with 1 and 2 terms on 128-wide tiles
  const bool bf_only = argc > 1 && std::string(argv[1]) == "bf16";
  if (!bf_only) run<double>(161214, 300, 8);
  for (int t = 0; t <= (bf_only ? 3 : 0); ++t) run<float>(161214, 300, 8, t);
  // BF16 overlaps on 128-wide tiles (1 and 2 terms)
  if (bf_only)
    for (int t = 1; t <= 2; ++t) run<float>(161214, 300, 8, t, 128);
  if (!bf_only) run<double>(62669, 80, 1);
  for (int t = 0; t <= (bf_only ? 3 : 0); ++t) run<float>(62669, 80, 1, t);
  return 0;
}
