// Fourier shifter kernels for gfx950: warp_frac2d (warp.c:883-1011) and
// shift_frac2d (warp.c:757-881).
//
// Spectra come from direct DFT passes along rows and columns (one output
// element per lane, roots from a host-built table indexed by t*k mod n).
// warp_frac2d's per-pixel Fourier sum
//     out(i,j) = 1/(nx ny) sum_a sum_b Re(F[a][b] X_a Y_b),
//     X_a = exp(i kx[a] xarg), Y_b = exp(i ky[b] yarg),
// is folded onto the half spectrum b = 0..ny/2 using F[-a][-b] = conj(F[a][b])
// (the image is real), keeping the reference's Nyquist convention kx = +n/2
// (make_freq, warp.c:316-360): the negative-b columns pair with the positive
// ones through XX_a = X_a + X~_a, where X~_a = X_a except X~_{nx/2} =
// conj(X_{nx/2}).  The sum over b = 1..(ny-1)/2 is a complex contraction
// Z[a][p] = sum_b F[a][b] Y_b(p), blocked 64 rows x 64 pixels x 16 columns in
// LDS; the contraction over a with XX_a(p) happens in the epilogue.
#pragma once

#include <hip/hip_runtime.h>

namespace flct {

constexpr double kWarpPi = 3.1415926535898;  // warp.c:893

// out[r][k] = sum_t in[r][t] e(sign t k / n), k = 0..nout-1.  Element (r, t)
// of the input is at in[r*irs + t*ies] (complex units; real input when
// REAL_IN, then imaginary part 0), output (r, k) at out[r*ors + k*oes].
// tw[2*e] = cos(2 pi e / n), tw[2*e+1] = sin(2 pi e / n).
template <bool REAL_IN>
__global__ void k_dft(const double* __restrict__ in, long long irs, long long ies, int n,
                      int nout, int nrows, double* __restrict__ out, long long ors,
                      long long oes, int sign, const double* __restrict__ tw) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nrows * nout) return;
  const int r = (int)(g / nout), k = (int)(g - (long long)r * nout);
  double sr = 0., si = 0.;
  int e = 0;
  for (int t = 0; t < n; ++t) {
    const double c = tw[2 * e], s = sign * tw[2 * e + 1];
    if (REAL_IN) {
      const double x = in[r * irs + t * ies];
      sr += x * c;
      si += x * s;
    } else {
      const double xr = in[2 * (r * irs + t * ies)], xi = in[2 * (r * irs + t * ies) + 1];
      sr += xr * c - xi * s;
      si += xr * s + xi * c;
    }
    e += k;
    if (e >= n) e -= n;
  }
  out[2 * (r * ors + k * oes)] = sr;
  out[2 * (r * ors + k * oes) + 1] = si;
}

// make_freq value of index i for size n (flct.c:1651-1692)
__device__ __forceinline__ double freq_of(int i, int n) {
  const int n21 = n / 2 - 1;
  if (i <= n21) return (double)i;
  if ((n / 2) * 2 != n) {
    if (i == n21 + 1) return (double)(n / 2);
    if (i == n21 + 2) return -(double)(n / 2);
    return -(double)(n21 - (i - (n21 + 3)));
  }
  if (i == n21 + 1) return (double)(n / 2);
  return -(double)(n21 - (i - (n21 + 2)));
}

// phase ramp of shift_frac2d (warp.c:822-852), spectrum in place
__global__ void k_phase_ramp(double* __restrict__ a, int nx, int ny, double dxarg, double dyarg) {
#pragma clang fp contract(off)
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nx * ny) return;
  const int i = (int)(g / ny), j = (int)(g - (long long)i * ny);
  const double pi = kWarpPi;
  const double kx = freq_of(i, nx), ky = freq_of(j, ny);
  const double crx = cos(2. * pi * kx * dxarg), cix = sin(2. * pi * kx * dxarg);
  const double cry = cos(2. * pi * ky * dyarg), ciy = sin(2. * pi * ky * dyarg);
  const double er = crx * cry - cix * ciy;
  const double ei = crx * ciy + cix * cry;
  const double orr = a[2 * g], oi = a[2 * g + 1];
  a[2 * g] = (orr * er - oi * ei);
  a[2 * g + 1] = (orr * ei + oi * er);
}

__global__ void k_real_scale(const double* __restrict__ a, double* __restrict__ out, long long n,
                             double normfac) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  out[g] = (double)(a[2 * g]) * normfac;
}

struct WarpArgs {
  const double* F;    // half spectrum, nx rows x nb columns (complex), nb = ny/2 + 1
  int nx, ny, nb;
  int kpos;           // number of paired columns b = 1..kpos
  const double* delx;
  const double* dely;
  int transp;
  double* out;
};

constexpr int kWP = 64;   // pixels per block
constexpr int kWA = 64;   // spectrum rows per block step
constexpr int kWK = 16;   // spectrum columns per step

__global__ void __launch_bounds__(256) k_warp(WarpArgs w) {
  __shared__ double Fs[kWA][kWK][2];
  __shared__ double Ys[kWK][kWP][2];
  __shared__ double px[kWP], py[kWP], pec[kWP], pes[kWP], pnc[kWP], pns[kWP];
  __shared__ double red[16][kWP];
  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;
  const long long npix = (long long)w.nx * w.ny;
  const long long p0 = (long long)blockIdx.x * kWP;
  const double nxinv = 1. / ((double)w.nx), nyinv = 1. / ((double)w.ny);
  if (tid < kWP) {
    const long long p = p0 + tid;
    double xarg = 0., yarg = 0.;
    if (p < npix) {
      const int i = (int)(p / w.ny), j = (int)(p - (long long)i * w.ny);
      double shiftx, shifty;
      if (w.transp) {
        shiftx = w.dely[p];
        shifty = w.delx[p];
      } else {
        shiftx = w.delx[p];
        shifty = w.dely[p];
      }
      // warp.c:958-959
      xarg = 2. * kWarpPi * (((double)i) - shiftx) * nxinv;
      yarg = 2. * kWarpPi * (((double)j) - shifty) * nyinv;
    }
    px[tid] = xarg;
    py[tid] = yarg;
    double s, c;
    sincos(yarg, &s, &c);
    pec[tid] = c;
    pes[tid] = s;
    if ((w.ny & 1) == 0) {
      sincos((double)(w.ny / 2) * yarg, &s, &c);
    } else {
      s = 0.;
      c = 0.;
    }
    pnc[tid] = c;
    pns[tid] = s;
  }
  __syncthreads();
  double part[4] = {0., 0., 0., 0.};
  const int ya_b = tid >> 6;     // Y generation: 4 column groups of 4
  const int ya_p = tid & 63;     // pixel
  for (int a0 = 0; a0 < w.nx; a0 += kWA) {
    double acc[4][4][2];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int v = 0; v < 4; ++v) acc[u][v][0] = acc[u][v][1] = 0.;
    for (int b0 = 1; b0 <= w.kpos; b0 += kWK) {
      // stage F rows a0..a0+63, columns b0..b0+15
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int e = tid + 256 * q;
        const int r = e >> 4, c = e & 15;
        const int a = a0 + r, b = b0 + c;
        double fr = 0., fi = 0.;
        if (a < w.nx && b <= w.kpos) {
          const long long o = 2 * ((long long)a * w.nb + b);
          fr = w.F[o];
          fi = w.F[o + 1];
        }
        Fs[r][c][0] = fr;
        Fs[r][c][1] = fi;
      }
      // Y_b(p) = exp(i b yarg): seeded exactly, then 3 steps of exp(i yarg)
      {
        const int bb = b0 + 4 * ya_b;
        double s, c;
        sincos((double)bb * py[ya_p], &s, &c);
        const double ec = pec[ya_p], es = pes[ya_p];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          Ys[4 * ya_b + t][ya_p][0] = c;
          Ys[4 * ya_b + t][ya_p][1] = s;
          const double nc = c * ec - s * es, ns = c * es + s * ec;
          c = nc;
          s = ns;
        }
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < kWK; ++kk) {
        double fr[4], fi[4], yr[4], yi[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          fr[u] = Fs[ty + 16 * u][kk][0];
          fi[u] = Fs[ty + 16 * u][kk][1];
        }
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          yr[v] = Ys[kk][tx + 16 * v][0];
          yi[v] = Ys[kk][tx + 16 * v][1];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            acc[u][v][0] += fr[u] * yr[v] - fi[u] * yi[v];
            acc[u][v][1] += fr[u] * yi[v] + fi[u] * yr[v];
          }
      }
      __syncthreads();
    }
    // epilogue: contract over this block's rows
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int a = a0 + ty + 16 * u;
      if (a >= w.nx) continue;
      const double kx = freq_of(a, w.nx);
      const bool nyq = ((w.nx & 1) == 0) && (a == w.nx / 2);
      const double f0r = w.F[2 * ((long long)a * w.nb)], f0i = w.F[2 * ((long long)a * w.nb) + 1];
      double fnr = 0., fni = 0.;
      if ((w.ny & 1) == 0) {
        fnr = w.F[2 * ((long long)a * w.nb + w.ny / 2)];
        fni = w.F[2 * ((long long)a * w.nb + w.ny / 2) + 1];
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = tx + 16 * v;
        double s, c;
        sincos(kx * px[q], &s, &c);
        // single terms: b = 0 and b = ny/2
        const double zr = f0r + (fnr * pnc[q] - fni * pns[q]);
        const double zi = f0i + (fnr * pns[q] + fni * pnc[q]);
        double t = c * zr - s * zi;
        // paired columns
        if (nyq) {
          t += 2. * c * acc[u][v][0];
        } else {
          t += 2. * (c * acc[u][v][0] - s * acc[u][v][1]);
        }
        part[v] += t;
      }
    }
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) red[ty][tx + 16 * v] = part[v];
  __syncthreads();
  if (tid < kWP) {
    const long long p = p0 + tid;
    double s = 0.;
    for (int r = 0; r < 16; ++r) s += red[r][tid];
    if (p < npix) w.out[p] = s * (1. / ((double)w.nx * w.ny));
  }
}

}  // namespace flct
