// Small per-pixel kernels around the tracker: relative-threshold maximum,
// tracked-pixel selection, skip fill by cubic convolution and the final
// xmiss reset.  Integer and comparison work here is bit-exact with the
// reference; floating-point expressions keep the reference's operation
// order with contraction off.
#pragma once

#include <hip/hip_runtime.h>

#include "track_kernels.hpp"

namespace flct {

constexpr double kXmiss = -999999.;  // flct.c:287

// max |f| over n values, one partial per block (max is order independent)
__global__ void k_absmax(const double* __restrict__ f, long long n,
                         unsigned long long* __restrict__ out) {
  double m = 0.;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (long long)gridDim.x * blockDim.x)
    m = fmax(m, fabs(f[t]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o));
  if ((threadIdx.x & 63) == 0) atomicMax(out, (unsigned long long)__double_as_longlong(m));
}

// Box sums for the local means of flct.c:815-827.  The clamped box of any
// pixel starts at (imin, jmin) with 0 <= imin <= nx-N, 0 <= jmin <= ny-N.
// Pass 1: hs[r][s] = sum_{c<N} f[r][s+c]; pass 2: mean[imin][jmin] =
// (sum_{r<N} hs[imin+r][jmin]) / (N*N), rows summed column-first like the
// reference's row-major loop.
__global__ void k_rowsum(const double* __restrict__ f, int nx, int ny, int N,
                         double* __restrict__ hs) {
  const int ns = ny - N + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nx * ns) return;
  const int r = (int)(t / ns), sidx = (int)(t - (long long)r * ns);
  const double* p = f + (long long)r * ny + sidx;
  double acc = 0.;
  for (int c = 0; c < N; ++c) acc = acc + p[c];
  hs[t] = acc;
}

__global__ void k_boxmean(const double* __restrict__ hs, int nx, int ny, int N,
                          double* __restrict__ mean) {
  const int ns = ny - N + 1, nr = nx - N + 1;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nr * ns) return;
  const int r = (int)(t / ns), sidx = (int)(t - (long long)r * ns);
  double acc = 0.;
  for (int q = 0; q < N; ++q) acc = acc + hs[(long long)(r + q) * ns + sidx];
  mean[t] = acc / ((double)N * N);
}

struct SelectArgs {
  const double* f1;
  const double* f2;
  int nx, ny;
  double thresh;
  // relative threshold: thresh is multiplied by max(max |f1|, max |f2|) read
  // here (flct.c:594-629), or nullptr for an absolute threshold
  const unsigned long long* absmax;
  int skip;       // 0: no skip
  int xoffset;    // q offset, applies to i (transp = 1)
  int yoffset;    // p offset, applies to j
  int* list;
  unsigned int* count;
  double* vx;
  double* vy;
  double* vm;
};

// clears n counters (a kernel rather than a memset, so that graphs that
// capture it replay correctly)
__global__ void k_zero(unsigned int* __restrict__ p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0u;
}

// threshold (flct.c:721-722) and skip (flct.c:734-757) predicates; pixels
// that are not tracked get vx = vy = xmiss, vm = 0 (flct.c:918-930)
__global__ void k_select(SelectArgs a) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool tracked = false;
  if (t < n) {
    const int i = (int)(t / a.ny), j = (int)(t - (long long)i * a.ny);
    const double fabsbar = 0.5 * (fabs(a.f1[t] + a.f2[t]));
    const bool below = (fabsbar < resolve_thresh(a.thresh, a.absmax));
    bool noskip = true;
    if (a.skip) noskip = (!((i - a.xoffset) % a.skip)) && (!((j - a.yoffset) % a.skip));
    tracked = (!below) && noskip;
    if (!tracked) {
      a.vx[t] = kXmiss;
      a.vy[t] = kXmiss;
      a.vm[t] = 0.;
    }
  }
  // wave-aggregated append
  const unsigned long long mask = __ballot(tracked);
  const int lane = threadIdx.x & 63;
  const int cnt = __popcll(mask);
  unsigned int base = 0;
  if (lane == 0 && cnt) base = atomicAdd(a.count, (unsigned)cnt);
  base = __shfl(base, 0);
  if (tracked) {
    const int before = __popcll(mask & ((1ull << lane) - 1ull));
    a.list[base + before] = (int)t;
  }
}

// Keys cubic convolution weight, a = -0.5 (flct.c:2391-2395)
__device__ __forceinline__ double keys_w(double t) {
#pragma clang fp contract(off)
  return (t >= (double)1.0)
             ? (((((double)(-0.5)) * t + ((double)2.5)) * t) - (double)4.) * t + (double)2.
             : (((double)(1.5)) * t - ((double)(2.5))) * t * t + (double)1.;
}

struct FillArgs {
  int nx, ny;
  int skip, xoffset, yoffset;
  int ixcount, iycount;  // lattice size
  double* vx;
  double* vy;
  double* vm;
};

// lattice value (ic, jc) of field v: the computed pixel it came from
__device__ __forceinline__ double lat(const double* v, const FillArgs& a, int ic, int jc) {
  return v[(long long)(a.xoffset + ic * a.skip) * a.ny + (a.yoffset + jc * a.skip)];
}

// cdata of interpcc2d (flct.c:2266-2336): lattice padded by one cell with
// the reference's edge and corner extrapolation, index (ci, cj) in
// [0, ixcount+1] x [0, iycount+1]
__device__ double cdata_at(const double* v, const FillArgs& a, int ci, int cj) {
#pragma clang fp contract(off)
  const int nxl = a.ixcount, nyl = a.iycount;
  const bool ie = (ci == 0 || ci == nxl + 1), je = (cj == 0 || cj == nyl + 1);
  if (!ie && !je) return lat(v, a, ci - 1, cj - 1);
  if (ie && !je) {
    const int jc = cj - 1;
    if (ci == 0)
      return lat(v, a, 2, jc) - 3. * (lat(v, a, 1, jc)) + 3. * (lat(v, a, 0, jc));
    return lat(v, a, nxl - 3, jc) - 3. * (lat(v, a, nxl - 2, jc)) + 3. * (lat(v, a, nxl - 1, jc));
  }
  if (!ie && je) {
    const int ic = ci - 1;
    if (cj == 0)
      return lat(v, a, ic, 2) - 3. * (lat(v, a, ic, 1)) + 3. * (lat(v, a, ic, 0));
    return lat(v, a, ic, nyl - 3) - 3. * (lat(v, a, ic, nyl - 2)) + 3. * (lat(v, a, ic, nyl - 1));
  }
  // corners, from the edge values in column 0 / nyl+1 (rows) or row 0 (flct.c:2313-2336)
  auto edge_col0 = [&](int c) {  // cdata[c][0], 1 <= c <= nxl
    const int ic = c - 1;
    return lat(v, a, ic, 2) - 3. * (lat(v, a, ic, 1)) + 3. * (lat(v, a, ic, 0));
  };
  auto edge_colN = [&](int c) {  // cdata[c][nyl+1]
    const int ic = c - 1;
    return lat(v, a, ic, nyl - 3) - 3. * (lat(v, a, ic, nyl - 2)) + 3. * (lat(v, a, ic, nyl - 1));
  };
  auto edge_row0 = [&](int c) {  // cdata[0][c], 1 <= c <= nyl
    const int jc = c - 1;
    return lat(v, a, 2, jc) - 3. * (lat(v, a, 1, jc)) + 3. * (lat(v, a, 0, jc));
  };
  if (ci == 0 && cj == 0) return 3. * (edge_col0(1)) - 3. * (edge_col0(2)) + edge_col0(3);
  if (ci == nxl + 1 && cj == 0)
    return 3. * (edge_col0(nxl)) - 3. * (edge_col0(nxl - 1)) + edge_col0(nxl - 2);
  if (ci == 0 && cj == nyl + 1)
    return 3. * (edge_row0(nyl)) - 3. * (edge_row0(nyl - 1)) + edge_row0(nyl - 2);
  return 3. * (edge_colN(nxl)) - 3. * (edge_colN(nxl - 1)) + edge_colN(nxl - 2);
}

__device__ double interp_one(const double* v, const FillArgs& a, double xint, double yint) {
#pragma clang fp contract(off)
  const int nxl = a.ixcount, nyl = a.iycount;
  if ((yint < 0.) || (yint > (double)(nyl - 1)) || ((xint < 0) || (xint > (double)(nxl - 1))))
    return kXmiss;
  const int itemp = ((int)xint > 0) ? (int)xint : 0;
  const int izero = (itemp < (nxl - 2)) ? itemp : nxl - 2;
  const int jtemp = ((int)yint > 0) ? (int)yint : 0;
  const int jzero = (jtemp < (nyl - 2)) ? jtemp : nyl - 2;
  double ftmp = 0.;
  bool databad = false;
  for (int ii = -1; ii < 3; ii++) {
    const double tx = fabs(xint - (double)(izero + ii));
    const double rx = keys_w(tx);
    for (int jj = -1; jj < 3; jj++) {
      const double ty = fabs(yint - (double)(jzero + jj));
      const double ry = keys_w(ty);
      const double c = cdata_at(v, a, izero + 1 + ii, jzero + 1 + jj);
      ftmp = ftmp + c * rx * ry;
      if (c == kXmiss) databad = true;
    }
  }
  return databad ? kXmiss : ftmp;
}

// skipped pixels get the cubic-convolution fill and vm = 0.5
// (flct.c:1080-1107).  Lattice pixels are only read and skipped pixels only
// written, so the update is in place without races.
__global__ void k_fill(FillArgs a) {
#pragma clang fp contract(off)
  const long long n = (long long)a.nx * a.ny;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = (int)(t / a.ny), j = (int)(t - (long long)i * a.ny);
  const bool noskip = (!((i - a.xoffset) % a.skip)) && (!((j - a.yoffset) % a.skip));
  if (noskip) return;
  const double xint = ((double)(i - a.xoffset)) / (double)(a.skip);
  const double yint = ((double)(j - a.yoffset)) / (double)(a.skip);
  a.vx[t] = interp_one(a.vx, a, xint, yint);
  a.vy[t] = interp_one(a.vy, a, xint, yint);
  a.vm[t] = 0.5;
}

// flct.c:1125-1133
__global__ void k_finish(double* __restrict__ vx, double* __restrict__ vy,
                         double* __restrict__ vm, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double x = vx[t], y = vy[t];
  if (x == kXmiss) vm[t] = 0.;
  vx[t] = (x == kXmiss) ? 0. : x;
  vy[t] = (y == kXmiss) ? 0. : y;
}

}  // namespace flct
