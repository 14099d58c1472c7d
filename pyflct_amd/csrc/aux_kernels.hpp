// Small per-pixel kernels around the tracker: relative-threshold maximum,
// tracked-pixel selection, skip fill by cubic convolution and the final
// xmiss reset.  Integer and comparison work here is bit-exact with the
// reference; floating-point expressions keep the reference's operation
// order with contraction off.
#pragma once

#include <hip/hip_runtime.h>

#include "interp_cc.hpp"
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

// interpcc2d over the computed lattice (flct.c:2226-2463)
__device__ double interp_one(const double* v, const FillArgs& a, double xint, double yint) {
  auto at = [&](int ic, int jc) { return lat(v, a, ic, jc); };
  return cc_interp(at, a.ixcount, a.iycount, xint, yint, kXmiss);
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
