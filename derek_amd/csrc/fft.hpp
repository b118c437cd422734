// Stockham FFT staged through LDS, for gfx950.
//
// A workgroup owns LINES one-dimensional transforms of length N that sit in
// LDS as two ping-pong buffers.  Each radix-R stage reads R elements spaced
// N/R apart, applies twiddles exp(S*2*pi*i*k*r/(Ns*R)) (k = j mod Ns), does
// an R-point DFT in registers and writes them Ns apart from
// (j/Ns)*Ns*R + k (Stockham autosort: natural order in, natural order out).
//
// Two LDS layouts:
//   CONTIG  element (line, idx) at line*(N+1) + idx - lines contiguous in
//           global memory (the z axis of the flat z-fastest grid);
//   TILE    element (line, idx) at idx*LINES + line - LINES neighbouring
//           z columns of a strided axis (y or x), so global loads are rows
//           of LINES consecutive elements.
// The transform length is a template constant for the hot sizes (the
// compiler then folds all index arithmetic) and a runtime plan otherwise.
#pragma once
#include "common.hpp"

namespace derek {

constexpr int kMaxStages = 16;

struct FftPlan {
  int n;
  int ns;  // number of stages, -1 if n has a prime factor outside {2,3,5,7,11,13}
  int r[kMaxStages];
};

// Radices with a register DFT; composite ones (6, 9, 12, 16) are built from
// the prime-power kernels with internal twiddles.
constexpr int kRadices[12] = {16, 12, 9, 8, 6, 5, 4, 3, 2, 7, 11, 13};

// Depth-first search for the factorisation with the fewest stages, ties
// broken by the largest smallest factor (balanced stages).
constexpr void plan_search(int m, int depth, int* cur, int* best, int& best_len, int& best_min) {
  if (m == 1) {
    int mn = 1 << 30;
    for (int i = 0; i < depth; ++i) mn = cur[i] < mn ? cur[i] : mn;
    if (depth < best_len || (depth == best_len && mn > best_min)) {
      best_len = depth;
      best_min = mn;
      for (int i = 0; i < depth; ++i) best[i] = cur[i];
    }
    return;
  }
  if (depth + 1 > best_len || depth >= kMaxStages) return;
  for (int q = 0; q < 12; ++q) {
    const int r = kRadices[q];
    if (m % r == 0) {
      cur[depth] = r;
      plan_search(m / r, depth + 1, cur, best, best_len, best_min);
    }
  }
}

constexpr FftPlan make_plan(int n) {
  FftPlan p{n, 0, {}};
  if (n < 1) {
    p.ns = -1;
    return p;
  }
  int cur[kMaxStages] = {};
  int best[kMaxStages] = {};
  int best_len = kMaxStages + 1, best_min = 0;
  plan_search(n, 0, cur, best, best_len, best_min);
  if (best_len > kMaxStages) {
    p.ns = -1;
    return p;
  }
  p.ns = best_len;
  // larger radix first
  for (int i = 0; i < best_len; ++i) p.r[i] = best[i];
  for (int i = 0; i < best_len; ++i)
    for (int j = i + 1; j < best_len; ++j)
      if (p.r[j] > p.r[i]) {
        const int t = p.r[i];
        p.r[i] = p.r[j];
        p.r[j] = t;
      }
  if (n == 1) p.ns = 0;
  return p;
}

// ---------------------------------------------------------------- R-point DFTs
// X_k = sum_j v_j exp(S*2*pi*i*j*k/R), in place.
template <int S, typename C> __device__ __forceinline__ void dft2(C* v) {
  C a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int S, typename C> __device__ __forceinline__ void dft3(C* v) {
  using R = decltype(C::x);
  const R h = (R)0.86602540378443864676372317075294;  // sin(2pi/3)
  C t = cadd(v[1], v[2]);
  C d = csub(v[1], v[2]);
  C m = csub(v[0], cscale(t, (R)0.5));
  C s = cmul_si<S>(cscale(d, h));
  v[0] = cadd(v[0], t);
  v[1] = cadd(m, s);
  v[2] = csub(m, s);
}

template <int S, typename C> __device__ __forceinline__ void dft4(C* v) {
  C a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  C c = cadd(v[1], v[3]), d = cmul_si<S>(csub(v[1], v[3]));
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

template <int S, typename C> __device__ __forceinline__ void dft5(C* v) {
  using R = decltype(C::x);
  const R c1 = (R)0.30901699437494742410229341718282;   // cos(2pi/5)
  const R c2 = (R)-0.80901699437494742410229341718282;  // cos(4pi/5)
  const R s1 = (R)0.95105651629515357211643933337938;   // sin(2pi/5)
  const R s2 = (R)0.58778525229247312916870595463907;   // sin(4pi/5)
  C p1 = cadd(v[1], v[4]), m1 = csub(v[1], v[4]);
  C p2 = cadd(v[2], v[3]), m2 = csub(v[2], v[3]);
  C a1 = cadd(v[0], cadd(cscale(p1, c1), cscale(p2, c2)));
  C a2 = cadd(v[0], cadd(cscale(p1, c2), cscale(p2, c1)));
  C b1 = cmul_si<S>(cadd(cscale(m1, s1), cscale(m2, s2)));
  C b2 = cmul_si<S>(csub(cscale(m1, s2), cscale(m2, s1)));
  v[0] = cadd(v[0], cadd(p1, p2));
  v[1] = cadd(a1, b1);
  v[4] = csub(a1, b1);
  v[2] = cadd(a2, b2);
  v[3] = csub(a2, b2);
}

template <int S, typename C> __device__ __forceinline__ void dft8(C* v) {
  using R = decltype(C::x);
  const R h = (R)0.70710678118654752440084436210485;
  C e[4] = {v[0], v[2], v[4], v[6]};
  C o[4] = {v[1], v[3], v[5], v[7]};
  dft4<S>(e);
  dft4<S>(o);
  // w^1 = (1 + S i)/sqrt2, w^2 = S i, w^3 = (-1 + S i)/sqrt2
  C o1 = cscale(cadd(o[1], cmul_si<S>(o[1])), h);
  C o2 = cmul_si<S>(o[2]);
  C o3 = cscale(csub(cmul_si<S>(o[3]), o[3]), h);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// Twiddle table entry exp(S*2*pi*i*t/N); the table holds the S = -1 values.
template <int S, typename C> __device__ __forceinline__ C twid(const C* tw, int t) {
  C w = tw[t];
  return S > 0 ? cconj(w) : w;
}

// Odd primes 7, 11, 13: direct O(R^2) sum with table roots of unity.
template <int R, int S, typename C> __device__ __forceinline__ void dft_direct(C* v, const C* tw, int N) {
  C out[R];
  const int step = N / R;
#pragma unroll
  for (int k = 0; k < R; ++k) {
    C acc = v[0];
#pragma unroll
    for (int j = 1; j < R; ++j) acc = cadd(acc, cmul(v[j], twid<S>(tw, ((j * k) % R) * step)));
    out[k] = acc;
  }
#pragma unroll
  for (int k = 0; k < R; ++k) v[k] = out[k];
}

template <int R, int S, typename C> __device__ __forceinline__ void dftR(C* v, const C* tw, int N);

// cos and sin of 2*pi*m/R for the composite radices 6, 9, 12, 16 (m is a
// compile-time constant once the radix loops are unrolled, so these fold
// into immediates instead of table loads held in registers).
template <int R> __device__ __forceinline__ void root_cs(int m, double& c, double& s) {
  constexpr double h3 = 0.86602540378443864676;  // sin(pi/3)
  constexpr double h2 = 0.70710678118654752440;  // sin(pi/4)
  constexpr double c8 = 0.92387953251128675613, s8 = 0.38268343236508977173;  // pi/8
  constexpr double c40 = 0.76604444311897803520, s40 = 0.64278760968653932632;
  constexpr double c80 = 0.17364817766693034885, s80 = 0.98480775301220805936;
  constexpr double c20 = 0.93969262078590838405, s20 = 0.34202014332566873304;  // 20 degrees
  const int q = ((m % R) + R) % R;
  if constexpr (R == 6) {
    constexpr double cc[6] = {1.0, 0.5, -0.5, -1.0, -0.5, 0.5};
    constexpr double ss[6] = {0.0, h3, h3, 0.0, -h3, -h3};
    c = cc[q];
    s = ss[q];
  } else if constexpr (R == 9) {
    constexpr double cc[9] = {1.0, c40, c80, -0.5, -c20, -c20, -0.5, c80, c40};
    constexpr double ss[9] = {0.0, s40, s80, h3, s20, -s20, -h3, -s80, -s40};
    c = cc[q];
    s = ss[q];
  } else if constexpr (R == 12) {
    constexpr double cc[12] = {1.0, h3, 0.5, 0.0, -0.5, -h3, -1.0, -h3, -0.5, 0.0, 0.5, h3};
    constexpr double ss[12] = {0.0, 0.5, h3, 1.0, h3, 0.5, 0.0, -0.5, -h3, -1.0, -h3, -0.5};
    c = cc[q];
    s = ss[q];
  } else {
    static_assert(R == 16, "composite radix without a root table");
    constexpr double cc[16] = {1.0, c8, h2, s8, 0.0, -s8, -h2, -c8, -1.0, -c8, -h2, -s8, 0.0, s8, h2, c8};
    constexpr double ss[16] = {0.0, s8, h2, c8, 1.0, c8, h2, s8, 0.0, -s8, -h2, -c8, -1.0, -c8, -h2, -s8};
    c = cc[q];
    s = ss[q];
  }
}

// exp(S*2*pi*i*m/R)
template <int R, int S, typename C> __device__ __forceinline__ C root(int m) {
  using Rl = decltype(C::x);
  double c, s;
  root_cs<R>(m, c, s);
  return cmk<C>((Rl)c, (Rl)(S > 0 ? s : -s));
}

// t * exp(S*2*pi*i*m/R) for a compile-time m; quarter turns are sign flips
// and swaps, which fold into the neighbouring adds.
template <int R, int S, typename C> __device__ __forceinline__ C mul_root(C t, int m) {
  const int q = m % R;
  if (q == 0) return t;
  if ((4 * q) % R == 0) {
    const int h = 4 * q / R;
    return h == 1 ? cmul_si<S>(t) : (h == 2 ? cmk<C>(-t.x, -t.y) : cmul_si<-S>(t));
  }
  return cmul(t, root<R, S, C>(q));
}

// Composite R = A*B in registers: n = B*a + b, k = ka + A*kb;
// B DFTs of length A, twiddles exp(S*2*pi*i*b*ka/R), A DFTs of length B.
template <int A, int B, int S, typename C> __device__ __forceinline__ void dft_comp(C* v, const C* tw, int N) {
  constexpr int R = A * B;
  C y[B][A];
#pragma unroll
  for (int b = 0; b < B; ++b) {
    C t[A];
#pragma unroll
    for (int a = 0; a < A; ++a) t[a] = v[B * a + b];
    dftR<A, S>(t, tw, N);
#pragma unroll
    for (int ka = 0; ka < A; ++ka) y[b][ka] = mul_root<R, S>(t[ka], b * ka);
  }
#pragma unroll
  for (int ka = 0; ka < A; ++ka) {
    C t[B];
#pragma unroll
    for (int b = 0; b < B; ++b) t[b] = y[b][ka];
    dftR<B, S>(t, tw, N);
#pragma unroll
    for (int kb = 0; kb < B; ++kb) v[ka + A * kb] = t[kb];
  }
}

template <int R, int S, typename C> __device__ __forceinline__ void dftR(C* v, const C* tw, int N) {
  if constexpr (R == 2) dft2<S>(v);
  else if constexpr (R == 3) dft3<S>(v);
  else if constexpr (R == 4) dft4<S>(v);
  else if constexpr (R == 5) dft5<S>(v);
  else if constexpr (R == 8) dft8<S>(v);
  else if constexpr (R == 6) dft_comp<3, 2, S>(v, tw, N);
  else if constexpr (R == 9) dft_comp<3, 3, S>(v, tw, N);
  else if constexpr (R == 12) dft_comp<4, 3, S>(v, tw, N);
  else if constexpr (R == 16) dft_comp<4, 4, S>(v, tw, N);
  else dft_direct<R, S>(v, tw, N);
}

// ---------------------------------------------------------------- one stage
// CONTIG lines are padded by one element so that the lines a stage touches
// at once start in different LDS banks.
__host__ __device__ constexpr int contig_stride(int N) { return N + 1; }

template <bool TILE, int LINES>
__device__ __forceinline__ int lds_addr(int line, int idx, int N) {
  return TILE ? idx * LINES + line : line * contig_stride(N) + idx;
}

template <int R, int S, bool TILE, int LINES, int NT, typename C>
__device__ __forceinline__ void fft_stage(const C* __restrict__ src, C* __restrict__ dst, int Ns, int N,
                                          int lines, const C* __restrict__ tw) {
  const int NR = N / R;
  const int total = lines * NR;
  const int tstride = N / (Ns * R);
  for (int t = threadIdx.x; t < total; t += NT) {
    int line, j;
    if (TILE) { line = t % LINES; j = t / LINES; }
    else { line = t / NR; j = t - line * NR; }
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[lds_addr<TILE, LINES>(line, j + r * NR, N)];
    const int k = j % Ns;
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], twid<S>(tw, k * r * tstride));
    }
    dftR<R, S>(v, tw, N);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[lds_addr<TILE, LINES>(line, base + r * Ns, N)] = v[r];
  }
}

// Compile-time plan: stages unrolled by recursion.
template <int S, bool TILE, int LINES, int NT, int N, int STAGE, int NS, typename C>
__device__ __forceinline__ C* fft_stages_ct(C* src, C* dst, const C* tw, int lines) {
  constexpr FftPlan p = make_plan(N);
  if constexpr (STAGE >= p.ns) {
    return src;
  } else {
    constexpr int R = p.r[STAGE];
    fft_stage<R, S, TILE, LINES, NT>(src, dst, NS, N, lines, tw);
    __syncthreads();
    return fft_stages_ct<S, TILE, LINES, NT, N, STAGE + 1, NS * R>(dst, src, tw, lines);
  }
}

// Runs the transform on buf0 (buf1 is scratch); returns the buffer holding
// the result.  NC > 0: compile-time length NC; NC == 0: runtime plan.
template <int S, bool TILE, int LINES, int NT, int NC, typename C>
__device__ __forceinline__ C* fft_lds(C* buf0, C* buf1, const C* tw, const FftPlan& plan, int lines) {
  if constexpr (NC > 0) {
    return fft_stages_ct<S, TILE, LINES, NT, NC, 0, 1>(buf0, buf1, tw, lines);
  } else {
    C* src = buf0;
    C* dst = buf1;
    int Ns = 1;
    const int N = plan.n;
    for (int s = 0; s < plan.ns; ++s) {
      const int R = plan.r[s];
      switch (R) {
        case 2: fft_stage<2, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 3: fft_stage<3, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 4: fft_stage<4, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 5: fft_stage<5, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 6: fft_stage<6, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 7: fft_stage<7, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 8: fft_stage<8, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 9: fft_stage<9, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 11: fft_stage<11, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 12: fft_stage<12, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        case 16: fft_stage<16, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
        default: fft_stage<13, S, TILE, LINES, NT>(src, dst, Ns, N, lines, tw); break;
      }
      __syncthreads();
      C* t = src;
      src = dst;
      dst = t;
      Ns *= R;
    }
    return src;
  }
}

}  // namespace derek
