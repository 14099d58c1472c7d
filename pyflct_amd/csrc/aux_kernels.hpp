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
// Each output is summed in the same order as a plain loop over c (pass 1)
// and r (pass 2), so the means do not depend on the launch shape.
// Pass 1 stages kMeanTile + N - 1 values of one image row in LDS (dynamic,
// (kMeanTile + N - 1) doubles) and each thread sums its N from there; pass 2 gives each thread kMeanRows
// consecutive box origins of one column and adds every loaded row sum to the
// origins whose box holds it, so each value is loaded once per thread
// instead of once per output.
constexpr int kMeanTile = 256;     // pass-1 outputs per block (= threads)
constexpr int kMeanRows = 8;       // pass-2 origins per thread

__global__ void __launch_bounds__(kMeanTile) k_rowsum(const double* __restrict__ f, int nx, int ny, int N,
                                                      double* __restrict__ hs) {
  extern __shared__ double tile[];
  const int ns = ny - N + 1;
  const int nbx = (ns + kMeanTile - 1) / kMeanTile;
  const int r = (int)(blockIdx.x / nbx);
  const int s0 = (int)(blockIdx.x - (long long)r * nbx) * kMeanTile;
  const double* p = f + (long long)r * ny;
  for (int i = threadIdx.x; i < kMeanTile + N - 1; i += kMeanTile) {
    const int c = s0 + i;
    tile[i] = (c < ny) ? p[c] : 0.;
  }
  __syncthreads();
  const int sidx = s0 + (int)threadIdx.x;
  if (sidx >= ns) return;
  double acc = 0.;
  for (int c = 0; c < N; ++c) acc = acc + tile[threadIdx.x + c];
  hs[(long long)r * ns + sidx] = acc;
}

__global__ void __launch_bounds__(256) k_boxmean(const double* __restrict__ hs, int nx, int ny, int N,
                                                 double* __restrict__ mean) {
  const int ns = ny - N + 1, nr = nx - N + 1;
  const int nbx = (ns + 255) / 256;
  const int rb = (int)(blockIdx.x / nbx);
  const int sidx = (int)(blockIdx.x - (long long)rb * nbx) * 256 + (int)threadIdx.x;
  if (sidx >= ns) return;
  const int r0 = rb * kMeanRows;
  double acc[kMeanRows];
#pragma unroll
  for (int k = 0; k < kMeanRows; ++k) acc[k] = 0.;
  const int last = ((r0 + kMeanRows < nr) ? r0 + kMeanRows : nr) - 1 + N - 1;  // last row read
  // unrolled so that several row loads are in flight per thread
#pragma unroll 8
  for (int q = 0; r0 + q <= last; ++q) {
    const double v = hs[(long long)(r0 + q) * ns + sidx];
#pragma unroll
    for (int k = 0; k < kMeanRows; ++k) {
      // origin r0 + k sums rows r0 + k .. r0 + k + N - 1 in increasing order
      if (q >= k && q - k < N) acc[k] = acc[k] + v;
    }
  }
#pragma unroll
  for (int k = 0; k < kMeanRows; ++k)
    if (r0 + k < nr) mean[(long long)(r0 + k) * ns + sidx] = acc[k] / ((double)N * N);
}

// Transposed column strip: out[t * nx + i] = f[i * ny + c0 + t] for the N
// columns c0 .. c0 + N - 1 of an nx x ny image (writes coalesced along i).
__global__ void __launch_bounds__(256) k_strip(const double* __restrict__ f, int nx, int ny, int c0,
                                               int N, double* __restrict__ out) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)N * nx) return;
  const int t = (int)(e / nx), i = (int)(e - (long long)t * nx);
  out[e] = f[(long long)i * ny + c0 + t];
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
