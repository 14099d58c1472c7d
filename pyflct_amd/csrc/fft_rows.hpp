// Batched 1-D DFTs of rows of any length n <= 4096 in LDS, for the warp's
// spectrum, shift_frac2d and the sigma = 0 correlation (the reference's
// fftw_plan_dft_2d / r2c / c2r calls, warp.c:772-806, flct.c:1380-1436).
//
// out[r][k] = sum_t in[r][t] e^{sign 2 pi i t k / n}, k < nout, unnormalised.
// One block per row.  A power-of-two n is one radix-2 transform of length n.
// Any other n uses Bluestein's chirp: with w_m = e^{sign pi i m^2 / n},
// X_k = w_k sum_t (x_t w_t) conj(w_{k-t}), a circular convolution of length
// M = 2^ceil(log2(2n - 1)) done as a forward transform (decimation in
// frequency, natural order in, bit-reversed out), a product with the
// transform of the conjugate chirp (stored bit-reversed, scaled by 1/M) and
// an inverse transform (decimation in time, bit-reversed in, natural out),
// so no permutation pass is needed.
#pragma once

#include <hip/hip_runtime.h>

namespace flct {

constexpr int kRowDftMaxLog2 = 13;  // transforms up to 8192 points (128 KB of LDS)
constexpr int kRowDftThreads = 512;

struct RowDftArgs {
  const double* in;
  long long irs, ies;  // element (r, t) at in[r*irs + t*ies] (complex units unless real_in)
  int real_in;
  int n, nout, nrows;
  double* out;  // element (r, k) at out[2*(r*ors + k*oes)]
  long long ors, oes;
  int sign;
  int logm;             // transform length M = 2^logm
  int bluestein;        // 0: M = n
  const double2* tw;    // e^{+2 pi i e / M}, e < M/2
  const double2* chirp;  // w_t, t < n (Bluestein)
  const double2* bhat;  // transform of conj(w) over the circle, bit-reversed, / M (Bluestein)
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// decimation in frequency, in place: natural order in, bit-reversed out
__device__ void lds_fft_dif(double2* x, int logn, const double2* __restrict__ tw, int sign) {
  const int L = 1 << logn;
  for (int ls = logn - 1; ls >= 0; --ls) {
    const int s = 1 << ls;
    for (int t = threadIdx.x; t < L / 2; t += blockDim.x) {
      const int j = t & (s - 1);
      const int i0 = ((t >> ls) << (ls + 1)) + j, i1 = i0 + s;
      double2 w = tw[j << (logn - 1 - ls)];
      if (sign < 0) w.y = -w.y;
      const double2 u = x[i0], v = x[i1];
      x[i0] = make_double2(u.x + v.x, u.y + v.y);
      x[i1] = cmul(make_double2(u.x - v.x, u.y - v.y), w);
    }
    __syncthreads();
  }
}

// decimation in time, in place: bit-reversed order in, natural out
__device__ void lds_fft_dit(double2* x, int logn, const double2* __restrict__ tw, int sign) {
  const int L = 1 << logn;
  for (int ls = 0; ls < logn; ++ls) {
    const int s = 1 << ls;
    for (int t = threadIdx.x; t < L / 2; t += blockDim.x) {
      const int j = t & (s - 1);
      const int i0 = ((t >> ls) << (ls + 1)) + j, i1 = i0 + s;
      double2 w = tw[j << (logn - 1 - ls)];
      if (sign < 0) w.y = -w.y;
      const double2 u = x[i0], v = cmul(x[i1], w);
      x[i0] = make_double2(u.x + v.x, u.y + v.y);
      x[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

__global__ void __launch_bounds__(kRowDftThreads) k_row_dft(RowDftArgs q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* x = reinterpret_cast<double2*>(lds_raw);
  const int r = blockIdx.x;
  const int M = 1 << q.logm;
  for (int t = threadIdx.x; t < M; t += blockDim.x) {
    double2 v = make_double2(0., 0.);
    if (t < q.n) {
      const long long e = (long long)r * q.irs + (long long)t * q.ies;
      v = q.real_in ? make_double2(q.in[e], 0.) : make_double2(q.in[2 * e], q.in[2 * e + 1]);
      if (q.bluestein) v = cmul(v, q.chirp[t]);
    }
    x[t] = v;
  }
  __syncthreads();
  const int logn = q.logm;
  if (!q.bluestein) {
    lds_fft_dif(x, logn, q.tw, q.sign);
    for (int k = threadIdx.x; k < q.nout; k += blockDim.x) {
      const double2 v = x[__builtin_bitreverse32((unsigned)k) >> (32 - logn)];
      const long long o = 2 * ((long long)r * q.ors + (long long)k * q.oes);
      q.out[o] = v.x;
      q.out[o + 1] = v.y;
    }
    return;
  }
  lds_fft_dif(x, logn, q.tw, -1);
  for (int t = threadIdx.x; t < M; t += blockDim.x) x[t] = cmul(x[t], q.bhat[t]);
  __syncthreads();
  lds_fft_dit(x, logn, q.tw, +1);
  for (int k = threadIdx.x; k < q.nout; k += blockDim.x) {
    const double2 v = cmul(x[k], q.chirp[k]);
    const long long o = 2 * ((long long)r * q.ors + (long long)k * q.oes);
    q.out[o] = v.x;
    q.out[o + 1] = v.y;
  }
}

}  // namespace flct
