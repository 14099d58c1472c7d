// warp_frac2d (warp.c:883-1011) evaluated as a type-2 non-uniform FFT.
//
// The reference sums the image's whole Fourier series at every displaced
// pixel: out(p) = Re sum_k F_k e^{i k.x_p} / (nx ny), O(nx ny) work per
// pixel.  Here the series is spread once onto an oversampled periodic grid
// (Mx x My, powers of two >= 2 nx, 2 ny) and each pixel interpolates that
// grid with a separable kernel of w = 16 taps per axis:
//
//   g(m) = sum_k F_k / (phihat_x(kx) phihat_y(ky)) e^{2 pi i k.m / M}
//   f(x) = sum_{m near x} phi(x / h - m) g(m),   h = 2 pi / M,
//
// phi(z) = exp(beta (sqrt(1 - (2 z / w)^2) - 1)) ("exponential of
// semicircle", Barnett, Magland and af Klinteberg, SIAM J. Sci. Comput. 41
// (2019) C479), beta = 2.30 w, and phihat its Fourier transform (host
// quadrature).  phi is real and even, so only Re g is interpolated.  With
// w = 16 the result differs from the direct sum by about 1e-13 of max|image|,
// the size of the direct sum's own rounding.
//
// Kernels: k_nufft_rows builds each spectrum row a (kx = make_freq(a)) on
// the My grid, deconvolved and zero padded, and transforms it in LDS; the
// rows go to T (nx x My).  k_nufft_cols gathers column c of T onto the Mx
// grid, transforms it in LDS and stores Re into row c of Rt (My x Mx), so
// both passes write whole rows.  k_nufft_gather interpolates Rt at every
// pixel's displaced position.
#pragma once

#include <hip/hip_runtime.h>

#include "warp_kernels.hpp"

namespace flct {

constexpr int kNuW = 16;             // kernel taps per axis
constexpr double kNuBetaPerW = 2.30;  // beta / w for grid oversampling 2
constexpr int kNuMaxLog2 = 13;       // grid lengths up to 8192 (128 KB of LDS)
constexpr int kNuThreads = 512;

__device__ __forceinline__ int brev_n(int i, int logn) {
  return (int)(__builtin_bitreverse32((unsigned)i) >> (32 - logn));
}

// in-place radix-2 decimation-in-time transform of the L = 2^logn complex
// values in x (held in bit-reversed order), sign +1: y_m = sum_k x_k
// e^{+2 pi i k m / L}.  tw[e] = e^{+2 pi i e / L}, e < L/2.
__device__ void lds_fft_pos(double2* x, int logn, const double2* __restrict__ tw) {
  const int L = 1 << logn;
  for (int ls = 0; ls < logn; ++ls) {
    const int s = 1 << ls;
    for (int t = threadIdx.x; t < L / 2; t += blockDim.x) {
      const int j = t & (s - 1);
      const int i0 = ((t >> ls) << (ls + 1)) + j, i1 = i0 + s;
      const double2 w = tw[j << (logn - 1 - ls)];
      const double2 u = x[i0], v0 = x[i1];
      const double2 v = make_double2(v0.x * w.x - v0.y * w.y, v0.x * w.y + v0.y * w.x);
      x[i0] = make_double2(u.x + v.x, u.y + v.y);
      x[i1] = make_double2(u.x - v.x, u.y - v.y);
    }
    __syncthreads();
  }
}

struct NufftArgs {
  const double* F;      // half spectrum nx x nb (complex), nb = ny/2 + 1
  int nx, ny, nb;
  int lmx, lmy;         // log2 of the grid lengths Mx, My
  const double* dcx;    // 1 / phihat_x(kx_a), a < nx
  const double* dcy;    // 1 / phihat_y(ky_b), b < ny
  const double2* twx;   // e^{+2 pi i e / Mx}, e < Mx/2
  const double2* twy;   // e^{+2 pi i e / My}, e < My/2
  double2* T;           // nx x My
  double* Rt;           // My x Mx
  const double* delx;
  const double* dely;
  int transp;
  double* out;
};

// full-spectrum element (a, b) of a real image from its half spectrum
__device__ __forceinline__ double2 spec_at(const NufftArgs& q, int a, int b) {
  if (b < q.nb) {
    const long long o = 2 * ((long long)a * q.nb + b);
    return make_double2(q.F[o], q.F[o + 1]);
  }
  const int am = (a == 0) ? 0 : q.nx - a;
  const long long o = 2 * ((long long)am * q.nb + (q.ny - b));
  return make_double2(q.F[o], -q.F[o + 1]);
}

// one block per spectrum row a
__global__ void __launch_bounds__(kNuThreads) k_nufft_rows(NufftArgs q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* x = reinterpret_cast<double2*>(lds_raw);
  const int a = blockIdx.x;
  const int My = 1 << q.lmy;
  for (int t = threadIdx.x; t < My; t += blockDim.x) x[t] = make_double2(0., 0.);
  __syncthreads();
  const double da = q.dcx[a];
  for (int b = threadIdx.x; b < q.ny; b += blockDim.x) {
    const double2 f = spec_at(q, a, b);
    const double d = da * q.dcy[b];
    const int m = ((int)freq_of(b, q.ny)) & (My - 1);
    x[brev_n(m, q.lmy)] = make_double2(f.x * d, f.y * d);
  }
  __syncthreads();
  lds_fft_pos(x, q.lmy, q.twy);
  double2* row = q.T + (long long)a * My;
  for (int t = threadIdx.x; t < My; t += blockDim.x) row[t] = x[t];
}

// one block per grid column c
__global__ void __launch_bounds__(kNuThreads) k_nufft_cols(NufftArgs q) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  double2* x = reinterpret_cast<double2*>(lds_raw);
  const int c = blockIdx.x;
  const int Mx = 1 << q.lmx, My = 1 << q.lmy;
  for (int t = threadIdx.x; t < Mx; t += blockDim.x) x[t] = make_double2(0., 0.);
  __syncthreads();
  for (int a = threadIdx.x; a < q.nx; a += blockDim.x) {
    const int m = ((int)freq_of(a, q.nx)) & (Mx - 1);
    x[brev_n(m, q.lmx)] = q.T[(long long)a * My + c];
  }
  __syncthreads();
  lds_fft_pos(x, q.lmx, q.twx);
  double* row = q.Rt + (long long)c * Mx;
  for (int t = threadIdx.x; t < Mx; t += blockDim.x) row[t] = x[t].x;
}

__device__ __forceinline__ double es_phi(double z, double beta) {
  const double u = 2. * z * (1. / kNuW);
  const double r = 1. - u * u;
  return (r > 0.) ? exp(beta * (sqrt(r) - 1.)) : 0.;
}

// One thread per output pixel; a block takes a tile of kNuTile x kNuTile
// pixels, i fastest within a wave.  Grid coordinates follow i and j (about
// 2 cells per pixel), so for one tap the lanes of a wave read a few grid rows
// at every other column, and a block's taps stay inside a small window of
// the grid.
constexpr int kNuTile = 16;

__global__ void __launch_bounds__(kNuTile * kNuTile) k_nufft_gather(NufftArgs q) {
  const int tiles_i = (q.nx + kNuTile - 1) / kNuTile;
  const int ti = (int)(blockIdx.x % tiles_i), tj = (int)(blockIdx.x / tiles_i);
  const int i = ti * kNuTile + (int)(threadIdx.x % kNuTile);
  const int j = tj * kNuTile + (int)(threadIdx.x / kNuTile);
  if (i >= q.nx || j >= q.ny) return;
  const long long p = (long long)i * q.ny + j;
  double shiftx, shifty;
  if (q.transp) {
    shiftx = q.dely[p];
    shifty = q.delx[p];
  } else {
    shiftx = q.delx[p];
    shifty = q.dely[p];
  }
  const double nxinv = 1. / ((double)q.nx), nyinv = 1. / ((double)q.ny);
  // warp.c:958-959
  const double xarg = 2. * kWarpPi * (((double)i) - shiftx) * nxinv;
  const double yarg = 2. * kWarpPi * (((double)j) - shifty) * nyinv;
  const int Mx = 1 << q.lmx, My = 1 << q.lmy;
  const double inv2pi = 0.15915494309189533576888376337251;  // 1 / (2 pi)
  double tx = xarg * ((double)Mx * inv2pi), ty = yarg * ((double)My * inv2pi);
  // the grid is periodic: bring far displacements back to one period so the
  // integer tap origin below stays small
  tx -= (double)Mx * floor(tx * (1. / (double)Mx));
  ty -= (double)My * floor(ty * (1. / (double)My));
  const double fx = floor(tx - 0.5 * kNuW), fy = floor(ty - 0.5 * kNuW);
  const int mx0 = (int)fx + 1, my0 = (int)fy + 1;
  const double zx0 = tx - (fx + 1.), zy0 = ty - (fy + 1.);
  const double beta = kNuBetaPerW * kNuW;
  double wx[kNuW];
#pragma unroll
  for (int u = 0; u < kNuW; ++u) wx[u] = es_phi(zx0 - (double)u, beta);
  double acc = 0.;
#pragma unroll 2
  for (int v = 0; v < kNuW; ++v) {
    const double wy = es_phi(zy0 - (double)v, beta);
    const double* row = q.Rt + (long long)((my0 + v) & (My - 1)) * Mx;
    double s = 0.;
#pragma unroll
    for (int u = 0; u < kNuW; ++u) s += wx[u] * row[(mx0 + u) & (Mx - 1)];
    acc += wy * s;
  }
  q.out[p] = acc * (1. / ((double)q.nx * q.ny));
}

}  // namespace flct
