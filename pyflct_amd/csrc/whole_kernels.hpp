// Whole-image correlation for sigma = 0 (flct.c:567-586, 666-675, 802-808,
// 920): one cross_cor of the two full nx x ny images with window 1 and the
// whole-image means, any nx, ny.  The transforms are the direct DFT passes of
// warp_kernels.hpp; these kernels do the means, the packing, the cross power
// over the full spectrum, |.| with the fftshift, and the first-strict-maximum
// search.
#pragma once

#include <hip/hip_runtime.h>

#include "warp_kernels.hpp"

namespace flct {

constexpr int kSumBlocks = 256;

// partial sums of f over a fixed grid-stride partition (deterministic)
__global__ void k_sum_partial(const double* __restrict__ f, long long n, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0.;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    acc += f[t];
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

// mean = (sum of the partials in order) / n, one thread
__global__ void k_sum_final(const double* __restrict__ part, int nparts, long long n,
                            double* __restrict__ mean) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double s = 0.;
    for (int k = 0; k < nparts; ++k) s += part[k];
    *mean = s / (double)n;
  }
}

// z = (f1 - mean1) + i (f2 - mean2): both real transforms in one complex one
__global__ void k_pack_centered(const double* __restrict__ f1, const double* __restrict__ f2,
                                const double* __restrict__ means, long long n,
                                double* __restrict__ z) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  z[2 * t] = f1[t] - means[0];
  z[2 * t + 1] = f2[t] - means[1];
}

// C = conj(A) B normfac^2 over the full spectrum, A = (Z + conj Z(-k,-l))/2,
// B = (Z - conj Z(-k,-l))/(2i) (flct.c:1421-1477); with the filter both
// spectra are multiplied by exp(-(kx/(kr kx[nx/2]))^2 - (ky/(kr ky[ny/2]))^2)
// (flct.c:1445-1457, 1694-1715)
__global__ void k_cross_full(const double* __restrict__ z, int nx, int ny, int filter, double kr,
                             double* __restrict__ c) {
#pragma clang fp contract(off)
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= (long long)nx * ny) return;
  const int i = (int)(g / ny), j = (int)(g - (long long)i * ny);
  const int im = (nx - i) % nx, jm = (ny - j) % ny;
  const double zr = z[2 * g], zi = z[2 * g + 1];
  const long long gm = (long long)im * ny + jm;
  const double mr = z[2 * gm], mi = -z[2 * gm + 1];
  double ar = 0.5 * (zr + mr), ai = 0.5 * (zi + mi);
  double br = 0.5 * (zi - mi), bi = -0.5 * (zr - mr);
  if (filter) {
    const double kxroll = kr * freq_of(nx / 2, nx), kyroll = kr * freq_of(ny / 2, ny);
    const double argx = freq_of(i, nx) * (1. / kxroll), argy = freq_of(j, ny) * (1. / kyroll);
    const double fv = exp(-(argx * argx + argy * argy));
    ar = ar * fv;
    ai = ai * fv;
    br = br * fv;
    bi = bi * fv;
  }
  double normfac = (1. / ((double)nx * ny));
  normfac *= normfac;
  c[2 * g] = (ar * br + ai * bi) * normfac;
  c[2 * g + 1] = (ar * bi - ai * br) * normfac;
}

// |Re c| moved to the fftshift position (flct.c:1485-1510, 2111-2140) and the
// maximum value (atomic max of the bits of a non-negative double)
__global__ void k_abs_shift(const double* __restrict__ c, int nx, int ny, double* __restrict__ s,
                            unsigned long long* __restrict__ vmax) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.;
  if (g < (long long)nx * ny) {
    const int i = (int)(g / ny), j = (int)(g - (long long)i * ny);
    const int ii = (i + nx / 2) % nx, jj = (j + ny / 2) % ny;
    v = fabs(c[2 * g]);
    s[(long long)ii * ny + jj] = v;
  }
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) atomicMax(vmax, (unsigned long long)__double_as_longlong(v));
}

__global__ void k_first_max_init(unsigned long long* loc) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *loc = ~0ull;
}

// smallest row-major index of the maximum (the first strict maximum,
// flct.c:2142-2161)
__global__ void k_first_max(const double* __restrict__ s, long long n,
                            const unsigned long long* __restrict__ vmax,
                            unsigned long long* __restrict__ loc) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long m = ~0ull;
  if (g < n && (unsigned long long)__double_as_longlong(s[g]) == *vmax) m = (unsigned long long)g;
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long q = __shfl_xor(m, o);
    m = q < m ? q : m;
  }
  if ((threadIdx.x & 63) == 0 && m != ~0ull) atomicMin(loc, m);
}

}  // namespace flct
