// Register-resident FFTs for gfx950, one transform per quad (4 lanes).
//
// A length N = 4*M complex transform is held by the 4 lanes of a quad, M
// points per lane.  Two layouts are used:
//   strided: lane l holds elements 4*r + l, r = 0..M-1          (x[r])
//   blocked: lane l holds elements s + M*brev2(l), s = 0..M-1   (x[s])
// where brev2 reverses the two bits of the lane index.
//
//   quad_dit: strided in  -> blocked out   (decimation in time)
//   quad_dif: blocked in  -> strided out   (decimation in frequency)
//
// Both compute X[k] = sum_n x[n] e(SIGN * n k / N), e(t) = exp(2 pi i t),
// unnormalised.  The in-lane M-point DFT is a compile-time mixed-radix
// Cooley-Tukey recursion (fully unrolled, twiddles are compile-time
// constants); the 4-point stage across the quad is two radix-2 butterflies
// over DPP quad permutations, so no LDS is touched inside a transform.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace flct {

// ---------------------------------------------------------------- utilities

template <int... Is, class F>
__host__ __device__ __forceinline__ void static_for_impl(std::integer_sequence<int, Is...>, F&& f) {
  (f(std::integral_constant<int, Is>{}), ...);
}
// calls f(std::integral_constant<int, i>) for i = 0..N-1
template <int N, class F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  static_for_impl(std::make_integer_sequence<int, N>{}, f);
}

// compile-time cos(2 pi a / b), exact rational reduction to [0, pi/4]
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double taylor_cos(double x) {
  double term = 1.0, sum = 1.0;
  for (int n = 1; n < 16; ++n) {
    term *= -x * x / double((2 * n - 1) * (2 * n));
    sum += term;
  }
  return sum;
}
constexpr double taylor_sin(double x) {
  double term = x, sum = x;
  for (int n = 1; n < 16; ++n) {
    term *= -x * x / double((2 * n) * (2 * n + 1));
    sum += term;
  }
  return sum;
}
constexpr double cos2pi(long long a, long long b) {
  // angle 2 pi a / b
  if (b < 0) { a = -a; b = -b; }
  a %= b;
  if (a < 0) a += b;
  if (2 * a > b) a = b - a;              // cos(2pi t) = cos(2pi (1 - t)), t in [0, 1/2]
  if (4 * a > b) return -cos2pi(b - 2 * a, 2 * b);  // t in (1/4, 1/2]: -cos(2pi(1/2 - t))
  if (8 * a > b) return taylor_sin(2.0 * kPi * double(b - 4 * a) / double(4 * b));  // sin(2pi(1/4 - t))
  return taylor_cos(2.0 * kPi * double(a) / double(b));
}
constexpr double sin2pi(long long a, long long b) {
  // sin(2 pi a/b) = cos(2 pi (b - 4a) / (4b))
  return cos2pi(b - 4 * a, 4 * b);
}

// ---------------------------------------------------------------- complex

template <class T>
struct cplx {
  T re, im;
};

template <class T>
__device__ __forceinline__ cplx<T> cadd(cplx<T> a, cplx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <class T>
__device__ __forceinline__ cplx<T> csub(cplx<T> a, cplx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <class T>
__device__ __forceinline__ cplx<T> cmul(cplx<T> a, cplx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <class T>
__device__ __forceinline__ cplx<T> cconj(cplx<T> a) { return {a.re, -a.im}; }
// multiply by SIGN*i
template <int SIGN, class T>
__device__ __forceinline__ cplx<T> mul_si(cplx<T> a) {
  return SIGN > 0 ? cplx<T>{-a.im, a.re} : cplx<T>{a.im, -a.re};
}
// multiply by the compile-time root e(SIGN * A / B)
template <int SIGN, long long A, long long B, class T>
__device__ __forceinline__ cplx<T> mul_root(cplx<T> a) {
  constexpr long long r = ((A % B) + B) % B;
  if constexpr (r == 0) {
    return a;
  } else if constexpr (4 * r == B) {
    return mul_si<SIGN>(a);
  } else if constexpr (2 * r == B) {
    return {-a.re, -a.im};
  } else if constexpr (4 * r == 3 * B) {
    return mul_si<-SIGN>(a);
  } else {
    constexpr T c = T(cos2pi(r, B));
    constexpr T s = T(SIGN * sin2pi(r, B));
    return {a.re * c - a.im * s, a.re * s + a.im * c};
  }
}

// ---------------------------------------------------------------- in-lane DFT

constexpr int smallest_odd_factor(int m, int f) {
  return (f >= m) ? m : (m % f == 0) ? f : smallest_odd_factor(m, f + 2);
}
template <int M>
struct radix_of {
  static constexpr int value = (M == 1) ? 1 : (M % 4 == 0) ? 4 : (M % 2 == 0) ? 2 : smallest_odd_factor(M, 3);
};

template <int M, int SIGN, class T>
__device__ __forceinline__ void dft(cplx<T> (&x)[M]);

// prime (or 4 / 2) point DFT, direct form with compile-time roots
template <int R, int SIGN, class T>
__device__ __forceinline__ void dft_base(cplx<T> (&x)[R]) {
  if constexpr (R == 1) {
    return;
  } else if constexpr (R == 2) {
    cplx<T> a = x[0], b = x[1];
    x[0] = cadd(a, b);
    x[1] = csub(a, b);
  } else if constexpr (R == 4) {
    cplx<T> s0 = cadd(x[0], x[2]), d0 = csub(x[0], x[2]);
    cplx<T> s1 = cadd(x[1], x[3]), d1 = mul_si<SIGN>(csub(x[1], x[3]));
    x[0] = cadd(s0, s1);
    x[2] = csub(s0, s1);
    x[1] = cadd(d0, d1);
    x[3] = csub(d0, d1);
  } else {
    // odd R: pair n and R-n
    constexpr int H = (R - 1) / 2;
    cplx<T> a[H + 1], b[H + 1];
    static_for<H>([&](auto I) {
      constexpr int n = I + 1;
      a[n] = cadd(x[n], x[R - n]);
      b[n] = csub(x[n], x[R - n]);
    });
    cplx<T> x0 = x[0];
    cplx<T> sum = x0;
    static_for<H>([&](auto I) { sum = cadd(sum, a[I + 1]); });
    x[0] = sum;
    static_for<H>([&](auto K) {
      constexpr int k = K + 1;
      cplx<T> A = x0, Bv = {T(0), T(0)};
      static_for<H>([&](auto Nn) {
        constexpr int n = Nn + 1;
        constexpr T c = T(cos2pi((long long)n * k, R));
        constexpr T s = T(sin2pi((long long)n * k, R));
        A.re += a[n].re * c;
        A.im += a[n].im * c;
        Bv.re += b[n].re * s;
        Bv.im += b[n].im * s;
      });
      cplx<T> jb = mul_si<SIGN>(Bv);
      x[k] = cadd(A, jb);
      x[R - k] = csub(A, jb);
    });
  }
}

// M = R * S, decimation in time: x[R n2 + n1] -> X[k1 + S k2]
template <int M, int SIGN, class T>
__device__ __forceinline__ void dft(cplx<T> (&x)[M]) {
  constexpr int R = radix_of<M>::value;
  if constexpr (M == R) {
    dft_base<R, SIGN>(x);
  } else {
    constexpr int S = M / R;
    cplx<T> y[M];
    static_for<R>([&](auto N1) {
      constexpr int n1 = N1;
      cplx<T> sub[S];
      static_for<S>([&](auto N2) { sub[N2] = x[R * N2 + n1]; });
      dft<S, SIGN>(sub);
      static_for<S>([&](auto K1) {
        constexpr int k1 = K1;
        y[n1 * S + k1] = mul_root<SIGN, (long long)n1 * k1, M>(sub[k1]);
      });
    });
    static_for<S>([&](auto K1) {
      constexpr int k1 = K1;
      cplx<T> t[R];
      static_for<R>([&](auto N1) { t[N1] = y[N1 * S + k1]; });
      dft_base<R, SIGN>(t);
      static_for<R>([&](auto K2) { x[k1 + S * K2] = t[K2]; });
    });
  }
}

// ---------------------------------------------------------------- quad exchange

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
  hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xF, 0xF, true));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) { return dpp_f64<CTRL>(v); }
template <int CTRL>
__device__ __forceinline__ float dpp(float v) { return dpp_f32<CTRL>(v); }

// quad_perm encodings
constexpr int kQuadXor1 = 1 | (0 << 2) | (3 << 4) | (2 << 6);  // [1,0,3,2]
constexpr int kQuadXor2 = 2 | (3 << 2) | (0 << 4) | (1 << 6);  // [2,3,0,1]
constexpr int kQuadXor3 = 3 | (2 << 2) | (1 << 4) | (0 << 6);  // [3,2,1,0]
constexpr int kQuadSwap23 = 0 | (1 << 2) | (3 << 4) | (2 << 6); // [0,1,3,2]
template <int Q>
constexpr int kQuadBcast = Q | (Q << 2) | (Q << 4) | (Q << 6);

template <int CTRL, class T>
__device__ __forceinline__ cplx<T> dpp(cplx<T> v) {
  return {dpp<CTRL>(v.re), dpp<CTRL>(v.im)};
}

// radix-2 butterfly across the quad: partner = lane ^ mask; the lane whose
// bit is clear keeps lo + hi, the other lo - hi
template <int CTRL, class T>
__device__ __forceinline__ cplx<T> quad_bfly(cplx<T> v, bool bit) {
  cplx<T> p = dpp<CTRL>(v);
  T sg = bit ? T(-1) : T(1);
  return {p.re + sg * v.re, p.im + sg * v.im};
}

template <int SIGN, class T>
__device__ __forceinline__ cplx<T> lane3_rot(cplx<T> v, int l) {
  cplx<T> r = mul_si<SIGN>(v);
  bool on = (l == 3);
  return {on ? r.re : v.re, on ? r.im : v.im};
}

// e(SIGN * l / N) for l = 0..3
template <int N, int SIGN, class T>
__device__ __forceinline__ cplx<T> lane_root(int l) {
  constexpr T c1 = T(cos2pi(1, N)), s1 = T(SIGN * sin2pi(1, N));
  constexpr T c2 = T(cos2pi(2, N)), s2 = T(SIGN * sin2pi(2, N));
  constexpr T c3 = T(cos2pi(3, N)), s3 = T(SIGN * sin2pi(3, N));
  T c = (l == 0) ? T(1) : (l == 1) ? c1 : (l == 2) ? c2 : c3;
  T s = (l == 0) ? T(0) : (l == 1) ? s1 : (l == 2) ? s2 : s3;
  return {c, s};
}

// x[s] *= e(SIGN * l * s / N), s = 0..M-1, by recurrence from e(SIGN l / N)
template <int M, int SIGN, class T>
__device__ __forceinline__ void lane_twiddle(cplx<T> (&x)[M], int l) {
  constexpr int N = 4 * M;
  cplx<T> w1 = lane_root<N, SIGN, T>(l);
  cplx<T> w = w1;
  static_for<M>([&](auto S) {
    if constexpr (S > 0) {
      x[S] = cmul(x[S], w);
      if constexpr (S + 1 < M) w = cmul(w, w1);
    }
  });
}

// strided in (lane l: elements 4r + l) -> blocked out (lane l: s + M brev2(l))
template <int M, int SIGN, class T>
__device__ __forceinline__ void quad_dit(cplx<T> (&x)[M], int l) {
  dft<M, SIGN>(x);
  lane_twiddle<M, SIGN>(x, l);
  const bool b0 = l & 1, b1 = (l >> 1) & 1;
  static_for<M>([&](auto S) { x[S] = quad_bfly<kQuadXor2>(x[S], b1); });
  static_for<M>([&](auto S) { x[S] = lane3_rot<SIGN>(x[S], l); });
  static_for<M>([&](auto S) { x[S] = quad_bfly<kQuadXor1>(x[S], b0); });
}

// blocked in (lane l: s + M brev2(l)) -> strided out (lane l: 4r + l)
template <int M, int SIGN, class T>
__device__ __forceinline__ void quad_dif(cplx<T> (&x)[M], int l) {
  const bool b0 = l & 1, b1 = (l >> 1) & 1;
  static_for<M>([&](auto S) { x[S] = quad_bfly<kQuadXor1>(x[S], b0); });
  static_for<M>([&](auto S) { x[S] = lane3_rot<SIGN>(x[S], l); });
  static_for<M>([&](auto S) { x[S] = quad_bfly<kQuadXor2>(x[S], b1); });
  lane_twiddle<M, SIGN>(x, l);
  dft<M, SIGN>(x);
}

__device__ __forceinline__ int brev2(int l) { return ((l & 1) << 1) | ((l >> 1) & 1); }

}  // namespace flct
