// FLCT per-pixel tracker kernel for gfx950.
//
// One "slot" (N/2 quads = 2N lanes) tracks one pixel at a time; a block holds
// P slots and loops persistently over the tracked-pixel list.  For the box
// size N = 4*M the three phases per pixel are:
//
//  1. column transforms along the box rows ii: each lane loads its box
//     values, windows them g = G*(f - fbar) (flct.c:856-870, fbar from the
//     box-mean kernel, flct.c:815-827), both images packed as g1 + i g2; the
//     transform runs in registers and Z[k][c] goes to LDS;
//  2. row transforms along jj of the unpacked spectra A = r2c(g1),
//     B = r2c(g2) for k = 0..N/2 (rows k = 0 and k = N/2 packed into one
//     transform), the optional k-space filter on both (flct.c:1445-1457),
//     the cross power conj(A) B normfac (flct.c:1461-1477) and the inverse
//     transform along jj, written back in place as Q[k][jj];
//  3. column c2r along k for two columns per transform, |.|, the fftshift
//     by (N/2, N/2) as index arithmetic, the first-strict-maximum argmax
//     (flct.c:1512-1534, 2142-2161), the 3x3 Taylor fit (flct.c:1536-1556,
//     1639-1640) and vx, vy (flct.c:901-905).  With -h (HIRES, double) the
//     fit is replaced by the hires refinement of cross_cor: the surface is
//     interpolated by cubic convolution on a 21 x 21 grid of 0.1 px about
//     the integer peak, and the grid maximum gets the 3 x 3 fit
//     (flct.c:1564-1645).
//
// With bias correction on, the pixel is tracked a second time with the
// Gaussian windows of image 1 and image 2 centred at -d/2 and +d/2 about the
// pixel, d being the first estimate (window-offset correction).
#pragma once

#include "fft_quad.hpp"
#include "interp_cc.hpp"

namespace flct {

struct TrackArgs {
  const double* f1;
  const double* f2;
  int nx, ny;
  int nsize;            // box is 2*nsize
  double sigminv;       // 1/sigma
  double deltinv;       // 1/deltat
  double deltas;
  const double* gw;     // gw[d + N - 1] = exp(-(sigminv d)^2), d = -(N-1)..N-1
  const double* mean1;  // box means by box origin, (nx-N+1) x (ny-N+1)
  const double* mean2;
  const double* filt;   // (N/2+1) x N filter values, or nullptr
  const int* list;      // tracked pixel list, or nullptr for all pixels in order
  long long nwork;      // list length (or pixel count without a list)
  const unsigned int* nwork_dev;  // list length in device memory (overrides nwork), or nullptr
  const unsigned long long* absmax;  // max |f1|, max |f2| as double bits, or nullptr
  int* flag_list;                    // pixels to re-evaluate in FP64, or nullptr
  unsigned int* flag_count;
  int bias;
  int hires;  // -h: run the HIRES instantiation (cubic-convolution peak on a 0.1 px grid)
  double* vx;
  double* vy;
  double* vm;
  int slots;  // pixels tracked together per block, 1..the compiled count (0: compiled count)
};

// slots per block of a launch: the tuned count when it is 1..P, else P
__device__ __host__ __forceinline__ int run_slots(int slots, int P) {
  return (slots > 0 && slots < P) ? slots : P;
}

// The FP32 form lists the pixels its surface cannot settle for re-evaluation
// in FP64, like the split-BF16 form: another cell within kTieTol32 of the
// maximum, or a Taylor fit whose first-order response to relative errors
// kRelErr32 of the nine values exceeds kFitTol32 pixels.
constexpr double kTieTol32 = 1e-5;
constexpr double kRelErr32 = 1e-6;
constexpr double kFitTol32 = 1e-2;

// Slots (pixels tracked together) per block, from a sweep of 1, 2, 3, 4, 6
// and 8 slots over every box size at 2048^2 (every pixel tracked,
// one library per count built with TUNE_FLAGS=-DFLCT_TRACK_SLOTS=n), where the best count beat the default
// rule below by 4 % or more; 0: default rule.
constexpr int tuned_slots(int M, bool fp32) {
  if (fp32) {
    switch (M) {
      case 4: return 8;
      case 5: return 6;
      case 7: return 1;
      case 9: return 3;
      case 10: return 3;
      case 12: return 8;
      case 13: return 6;
      case 15: return 2;
      case 17: return 3;
      case 18: return 3;
      case 19: return 3;
      default: return 0;
    }
  }
  switch (M) {
    case 4: return 4;
    case 5: return 3;
    case 6: return 2;
    case 10: return 4;
    case 12: return 3;
    case 13: return 3;
    default: return 0;
  }
}

// slots per block for box size 4*M, real type T
template <int M, class T>
struct TrackCfg {
  static constexpr int N = 4 * M;
  static constexpr int kSlotThreads = 2 * N;
  // LDS row stride (complex elements), = 4 mod 8: the four columns a quad
  // reads from one row and the next quad's four from the next row fall in
  // different bank groups
  static constexpr int kRow = (N % 8 == 4) ? N + 8 : N + 4;
  static constexpr int kZBytes = (N * kRow + 8) * int(sizeof(cplx<T>));
  static constexpr int kSlotBytes = ((kZBytes + 4 * N * 8 + 9 * 8 + 48) + 15) / 16 * 16;
  static constexpr int waste(int p) {
    int t = p * kSlotThreads;
    int r = ((t + 63) / 64) * 64;
    return (r - t) * 1000 / r;
  }
  static constexpr int pick() {
    int best = 1, bw = waste(1);
    for (int p = 2; p <= 8; ++p) {
      if (p * kSlotBytes > 80 * 1024 || p * kSlotThreads > 512) break;
      int w = waste(p);
      if (w < bw) { best = p; bw = w; }
    }
    return best;
  }
  // the swept count, reduced until the block fits a CU (160 KB of LDS,
  // 1024 threads)
  static constexpr int fit(int p) {
    while (p > 1 && (p * kSlotBytes + 2 * N * 8 > 160 * 1024 || p * kSlotThreads > 1024)) --p;
    return p;
  }
#ifdef FLCT_TRACK_SLOTS
  static constexpr int P = fit(FLCT_TRACK_SLOTS);  // tuning builds: one count for every size
#else
  static constexpr int P = tuned_slots(M, sizeof(T) == 4) > 0 ? fit(tuned_slots(M, sizeof(T) == 4))
                                                             : pick();
#endif
  static constexpr int kThreads = ((P * kSlotThreads + 63) / 64) * 64;
  static constexpr int kLdsBytes = P * kSlotBytes + (2 * N) * 8;
  // minimum waves per SIMD asked of the register allocator (1: no bound),
  // where the registers it spills cost less than the extra wave gains
  // (measured at 2048^2): three for FP64 box 24, two for FP64 boxes 40..52
  // and FP32 boxes 72..96.  Elsewhere the bound was slower: FP64 boxes 28
  // and 32 at three waves (1.5x), FP64 56..64 at two (1.6x), FP32 32..36 at
  // three.
  static constexpr int kMinWaves =
      (sizeof(T) == 8 && M == 6) ? 3
      : ((sizeof(T) == 4 && M >= 18 && M <= 24) || (sizeof(T) == 8 && M >= 10 && M <= 13)) ? 2
                                                                                         : 1;
};

// LDS index of spectrum element (k, c).  Rows of output block p = k / M
// are offset by delta(p) elements so that the four lanes of a quad, which
// write rows M*p + s for p = 0..3 of one column, hit four bank groups.
template <int M>
__device__ __forceinline__ int zidx(int k, int c) {
  constexpr int N = 4 * M;
  constexpr int R = (N % 8 == 4) ? N + 8 : N + 4;
  const int p = k / M;
  const int delta = (M % 2 == 0) ? 2 * p : ((6 * p) & 7);
  return k * R + c + delta;
}

// column at which a quad parks blocked element l (lane brev2(l / M), slot
// l % M) of a private row: lane-strided, so the quad's stores are contiguous
template <int M>
__device__ __forceinline__ int park_col(int l) {
  return 4 * (l % M) + brev2(l / M);
}

// absolute threshold: thresh, or thresh * max(f1max, f2max) for a relative one
// (flct.c:594-629); absmax holds the two maxima as double bits
__device__ __forceinline__ double resolve_thresh(double thresh, const unsigned long long* absmax) {
  if (!absmax) return thresh;
  const double f1max = __longlong_as_double((long long)absmax[0]);
  const double f2max = __longlong_as_double((long long)absmax[1]);
  const double fmax = (f1max > f2max) ? f1max : f2max;
  return thresh * fmax;
}

__device__ __forceinline__ unsigned long long dbits(double v) {
  return (unsigned long long)__double_as_longlong(v);
}

// keeps the compiler from moving LDS accesses across this point; LDS
// requests of one wave are served in issue order, so lanes of a quad see
// each other's earlier writes
__device__ __forceinline__ void wave_lds_order() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
  asm volatile("" ::: "memory");
}

// an opaque copy of a lane value: address arithmetic derived from it stays
// inside the phase instead of being hoisted out of the persistent loop,
// where every distinct address would hold a register for the whole loop
__device__ __forceinline__ int fresh(int x) {
  asm volatile("" : "+v"(x));
  return x;
}

template <class T>
__device__ __forceinline__ cplx<T> cross_power(cplx<T> A, cplx<T> B, T norm) {
  return {(A.re * B.re + A.im * B.im) * norm, (A.re * B.im - A.im * B.re) * norm};
}

// HIRES (double only): the -h peak step replaces the 3 x 3 fit; a separate
// instantiation, so the plain kernels keep their register allocation
template <int M, class T, bool HIRES = false>
__global__ void __launch_bounds__((TrackCfg<M, T>::kThreads), (TrackCfg<M, T>::kMinWaves))
    k_track(TrackArgs a) {
  static_assert(!HIRES || sizeof(T) == 8, "the hires step runs in the double kernels");
  using Cfg = TrackCfg<M, T>;
  constexpr int N = 4 * M;
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const int slot = tid / Cfg::kSlotThreads;
  const int stid = tid - slot * Cfg::kSlotThreads;
  const int SP = run_slots(a.slots, Cfg::P);  // slots of this launch
  const bool in_slot = slot < SP;
  const int lane = stid & 3, qd = stid >> 2;
  double* gw = reinterpret_cast<double*>(lds);
  for (int t = tid; t < 2 * N - 1; t += blockDim.x) gw[t] = a.gw[t];
  unsigned char* base = lds + 2 * N * 8 + (in_slot ? slot : 0) * Cfg::kSlotBytes;
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(base);
  double* wx1 = reinterpret_cast<double*>(base + Cfg::kZBytes);
  double* wy1 = wx1 + N;
  double* wx2 = wy1 + N;
  double* wy2 = wx2 + N;
  double* nb = wy2 + N;
  double* res = nb + 9;
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(res + 2);
  unsigned long long* gmx = amax + 2;  // max |g| of image 1 and 2 (bits)
  unsigned int* aidx = reinterpret_cast<unsigned int*>(amax + 1);
  __syncthreads();

  const int nx = a.nx, ny = a.ny, nsize = a.nsize;
  const int ns = ny - N + 1;
  const T norm = T((1. / ((double)N * N)) * (1. / ((double)N * N)));
  const int npass = a.bias ? 2 : 1;
  // Work distribution.  List mode: slots stride through the list.  All
  // pixels: blocks are dealt to the 8 XCDs round robin, so XCD x = block % 8
  // takes rows [x nx/8, (x+1) nx/8) and walks them in strips of kStrip
  // columns, row by row inside a strip; the box rows it reads then stay in
  // its L2.
  constexpr int kXcd = 8, kStrip = 512;
  const bool dense = (a.list == nullptr);
  long long first, stride, count;
  int r0 = 0, rows_b = 0;
  if (dense) {
    const int nband = (int)gridDim.x < kXcd ? (int)gridDim.x : kXcd;
    const int x = blockIdx.x % nband;
    const long long nbx = (gridDim.x - x + nband - 1) / nband;
    r0 = (int)((long long)nx * x / nband);
    rows_b = (int)((long long)nx * (x + 1) / nband) - r0;
    first = (long long)(blockIdx.x / nband) * SP + slot;
    stride = nbx * SP;
    count = (long long)rows_b * ny;
  } else {
    first = (long long)blockIdx.x * SP + slot;
    stride = (long long)gridDim.x * SP;
    count = a.nwork_dev ? (long long)*a.nwork_dev : a.nwork;
  }
  const long long iters = (count + stride - 1) / stride;
  const int full_strips = ny / kStrip;

  for (long long it = 0; it < iters; ++it) {
    const long long w = first + it * stride;
    const bool active = in_slot && w < count;
    int i = 0, j = 0, imin = 0, jmin = 0;
    double f1bar = 0., f2bar = 0.;
    if (active) {
      if (dense) {
        const long long sw = (long long)rows_b * kStrip;
        const int sidx = (int)(w / sw);
        int width = kStrip;
        long long t = w - (long long)sidx * sw;
        if (sidx >= full_strips) {
          width = ny - full_strips * kStrip;
          t = w - (long long)full_strips * sw;
        }
        const int strip = sidx < full_strips ? sidx : full_strips;
        i = r0 + (int)(t / width);
        j = strip * kStrip + (int)(t % width);
      } else {
        const long long p = (long long)a.list[w];
        i = (int)(p / ny);
        j = (int)(p - (long long)i * ny);
      }
      // box clamp, flct.c:773-786
      const int imin0 = (0 > (i - (nsize - 1))) ? 0 : i - (nsize - 1);
      const int imax0 = ((nx - 1) < (i + nsize)) ? nx - 1 : i + nsize;
      imin = (imax0 == nx - 1) ? nx - 1 - (2 * nsize - 1) : imin0;
      const int jmin0 = (0 > (j - (nsize - 1))) ? 0 : j - (nsize - 1);
      const int jmax0 = ((ny - 1) < (j + nsize)) ? ny - 1 : j + nsize;
      jmin = (jmax0 == ny - 1) ? ny - 1 - (2 * nsize - 1) : jmin0;
      f1bar = a.mean1[(long long)imin * ns + jmin];
      f2bar = a.mean2[(long long)imin * ns + jmin];
    }
    double shiftx = 0., shifty = 0.;
    bool flagged = false;  // FP32: re-evaluate in FP64 (read by thread stid == 0 only)
    for (int pass = 0; pass < npass; ++pass) {
      // ------------- windows: separable Gaussian centred on the pixel, or
      // moved by -d/2 (image 1) and +d/2 (image 2) in the bias pass
      if (active) {
        const double cx = 0.5 * shiftx, cy = 0.5 * shifty;
        if (stid < N) {
          const int d = stid + imin - i;
          if (pass == 0) {
            const double g = gw[d + N - 1];
            wx1[stid] = g;
            wx2[stid] = g;
          } else {
            const double a1 = a.sigminv * ((double)d + cx), a2 = a.sigminv * ((double)d - cx);
            wx1[stid] = exp(-a1 * a1);
            wx2[stid] = exp(-a2 * a2);
          }
        } else {
          const int c = stid - N;
          const int d = c + jmin - j;
          if (pass == 0) {
            const double g = gw[d + N - 1];
            wy1[c] = g;
            wy2[c] = g;
          } else {
            const double b1 = a.sigminv * ((double)d + cy), b2 = a.sigminv * ((double)d - cy);
            wy1[c] = exp(-b1 * b1);
            wy2[c] = exp(-b2 * b2);
          }
        }
        if (stid == 0) {
          *amax = 0ull;
          *aidx = 0xffffffffu;
          gmx[0] = 0ull;
          gmx[1] = 0ull;
        }
      }
      __syncthreads();
      // ------------- phase 1: column transforms of the packed windowed
      // images g1 + i g2.  In FP32 each image is first scaled by a power of
      // two that brings its largest |g| to [1, 2): exact in floating point
      // (every later value carries the same factor and the argmax and the
      // fit are scale-invariant), and it keeps the two packed images at
      // comparable magnitude so that a nearly empty box (e.g. one that only
      // grazes the limb of a disk) is not lost in the other image's rounding.
      // FP64 has the margin to pack unscaled.
      if constexpr (sizeof(T) == 8) {
        if (active) {
          const int l = fresh(lane), quad = fresh(qd);
          const double* p1 = a.f1 + (long long)imin * ny + jmin;
          const double* p2 = a.f2 + (long long)imin * ny + jmin;
          const int kb = M * brev2(l);
#pragma unroll 1
          for (int t = 0; t < 2; ++t) {
            const int c = quad + t * (N / 2);
            const double w1c = wy1[c], w2c = wy2[c];
            cplx<T> z[M];
            static_for<M>([&](auto Rr) {
              const int ii = 4 * Rr + l;
              const long long off = (long long)ii * ny + c;
              z[Rr] = {T((wx1[ii] * w1c) * (p1[off] - f1bar)),
                       T((wx2[ii] * w2c) * (p2[off] - f2bar))};
            });
            quad_dit<M, -1>(z, l);
            static_for<M>([&](auto S) { Z[zidx<M>(kb + S, c)] = z[S]; });
          }
        }
      } else {
        // Boxes 44..96: two passes over the box, the maxima first, then the
        // scaled values recomputed with the same operations (nothing held
        // across the barrier: two waves per SIMD instead of one; boxes 72..96
        // also bound the allocator to two waves, spilling a few registers).
        // Otherwise the values stay in registers between the passes (no
        // second read of the box).
        if constexpr (M >= 11 && M <= 24) {
          if (active) {
            const int l = fresh(lane), quad = fresh(qd);
            const double* p1 = a.f1 + (long long)imin * ny + jmin;
            const double* p2 = a.f2 + (long long)imin * ny + jmin;
            double m1 = 0., m2 = 0.;
            static_for<2>([&](auto Tc) {
              const int c = quad + Tc * (N / 2);
              const double w1c = wy1[c], w2c = wy2[c];
              static_for<M>([&](auto Rr) {
                const int ii = 4 * Rr + l;
                const long long off = (long long)ii * ny + c;
                m1 = fmax(m1, fabs((wx1[ii] * w1c) * (p1[off] - f1bar)));
                m2 = fmax(m2, fabs((wx2[ii] * w2c) * (p2[off] - f2bar)));
              });
            });
            m1 = fmax(m1, dpp<kQuadXor1>(m1));
            m2 = fmax(m2, dpp<kQuadXor1>(m2));
            m1 = fmax(m1, dpp<kQuadXor2>(m1));
            m2 = fmax(m2, dpp<kQuadXor2>(m2));
            if (l == 0) {
              atomicMax(&gmx[0], dbits(m1));
              atomicMax(&gmx[1], dbits(m2));
            }
          }
          __syncthreads();
          if (active) {
            const int l = fresh(lane), quad = fresh(qd);
            const double* p1 = a.f1 + (long long)imin * ny + jmin;
            const double* p2 = a.f2 + (long long)imin * ny + jmin;
            const double g1max = __longlong_as_double((long long)gmx[0]);
            const double g2max = __longlong_as_double((long long)gmx[1]);
            const double s1 = (g1max > 0.) ? ldexp(1.0, -ilogb(g1max)) : 1.0;
            const double s2 = (g2max > 0.) ? ldexp(1.0, -ilogb(g2max)) : 1.0;
            const int kb = M * brev2(l);
#pragma unroll 1
            for (int Tc = 0; Tc < 2; ++Tc) {
              const int c = quad + Tc * (N / 2);
              const double w1c = wy1[c], w2c = wy2[c];
              cplx<T> z[M];
              static_for<M>([&](auto Rr) {
                const int ii = 4 * Rr + l;
                const long long off = (long long)ii * ny + c;
                const double g1 = (wx1[ii] * w1c) * (p1[off] - f1bar);
                const double g2 = (wx2[ii] * w2c) * (p2[off] - f2bar);
                z[Rr] = {T(g1 * s1), T(g2 * s2)};
              });
              quad_dit<M, -1>(z, l);
              static_for<M>([&](auto S) { Z[zidx<M>(kb + S, c)] = z[S]; });
            }
          }
        } else {
          double ga[2][M], gb[2][M];
          if (active) {
            const int l = fresh(lane), quad = fresh(qd);
            const double* p1 = a.f1 + (long long)imin * ny + jmin;
            const double* p2 = a.f2 + (long long)imin * ny + jmin;
            double m1 = 0., m2 = 0.;
            static_for<2>([&](auto Tc) {
              const int c = quad + Tc * (N / 2);
              const double w1c = wy1[c], w2c = wy2[c];
              static_for<M>([&](auto Rr) {
                const int ii = 4 * Rr + l;
                const long long off = (long long)ii * ny + c;
                ga[Tc][Rr] = (wx1[ii] * w1c) * (p1[off] - f1bar);
                gb[Tc][Rr] = (wx2[ii] * w2c) * (p2[off] - f2bar);
                m1 = fmax(m1, fabs(ga[Tc][Rr]));
                m2 = fmax(m2, fabs(gb[Tc][Rr]));
              });
            });
            m1 = fmax(m1, dpp<kQuadXor1>(m1));
            m2 = fmax(m2, dpp<kQuadXor1>(m2));
            m1 = fmax(m1, dpp<kQuadXor2>(m1));
            m2 = fmax(m2, dpp<kQuadXor2>(m2));
            if (l == 0) {
              atomicMax(&gmx[0], dbits(m1));
              atomicMax(&gmx[1], dbits(m2));
            }
          }
          __syncthreads();
          if (active) {
            const int l = fresh(lane), quad = fresh(qd);
            const double g1max = __longlong_as_double((long long)gmx[0]);
            const double g2max = __longlong_as_double((long long)gmx[1]);
            const double s1 = (g1max > 0.) ? ldexp(1.0, -ilogb(g1max)) : 1.0;
            const double s2 = (g2max > 0.) ? ldexp(1.0, -ilogb(g2max)) : 1.0;
            const int kb = M * brev2(l);
            static_for<2>([&](auto Tc) {
              const int c = quad + Tc * (N / 2);
              cplx<T> z[M];
              static_for<M>([&](auto Rr) { z[Rr] = {T(ga[Tc][Rr] * s1), T(gb[Tc][Rr] * s2)}; });
              quad_dit<M, -1>(z, l);
              static_for<M>([&](auto S) { Z[zidx<M>(kb + S, c)] = z[S]; });
            });
          }
        }
      }
      __syncthreads();
      // ------------- phase 2: row transforms, cross power, inverse along jj
      // (FP32: gmx[0], read in phase 1, now counts near-maximum cells)
      if constexpr (sizeof(T) == 4) {
        if (active && stid == 0) gmx[0] = 0ull;
      }
      if (active) {
        const int l = fresh(lane), quad = fresh(qd);
        const int t = quad;
        const int lb = M * brev2(l);
        // Row task t >= 1: rows k = t and N - t, A = (Zk + conj(Zm))/2 and
        // B = (Zk - conj(Zm))/(2i), A's spectrum parked in the mirror row,
        // which only this quad reads.  Task 0: rows 0 and N/2 of A and B are
        // real along jj, so each image is one packed transform (row 0 + i row
        // N/2), parked in rows N/2 (A) and 0 (B) and unpacked through them.
        // The transforms are shared code for every task (only the input, the
        // unpack and the stores depend on it), so the wave holding task 0
        // does not run them twice.
        const bool tz = (t == 0);
        const int k = tz ? 0 : t, km = tz ? N / 2 : N - t;
        cplx<T> u1[M], u2[M];
        static_for<M>([&](auto Rr) {
          const int jj = 4 * Rr + l;
          const cplx<T> zk = Z[zidx<M>(k, jj)], zm = Z[zidx<M>(km, jj)];
          const cplx<T> g1 = {T(0.5) * (zk.re + zm.re), T(0.5) * (zk.im - zm.im)};
          const cplx<T> g2 = {T(0.5) * (zk.im + zm.im), T(-0.5) * (zk.re - zm.re)};
          const cplx<T> p1 = {zk.re, zm.re}, p2 = {zk.im, zm.im};
          u1[Rr] = tz ? p1 : g1;
          u2[Rr] = tz ? p2 : g2;
        });
        quad_dit<M, -1>(u1, l);
        wave_lds_order();
        static_for<M>([&](auto S) { Z[zidx<M>(km, 4 * S + l)] = u1[S]; });
        wave_lds_order();
        quad_dit<M, -1>(u2, l);
        wave_lds_order();
        if (tz) {
          static_for<M>([&](auto S) { Z[zidx<M>(0, 4 * S + l)] = u2[S]; });
          wave_lds_order();
          static_for<M>([&](auto S) {
            const int lx = lb + S;
            const int lm = (N - lx) % N;
            const cplx<T> U = Z[zidx<M>(N / 2, park_col<M>(lx))];
            const cplx<T> Um = Z[zidx<M>(N / 2, park_col<M>(lm))];
            const cplx<T> V = Z[zidx<M>(0, park_col<M>(lx))];
            const cplx<T> Vm = Z[zidx<M>(0, park_col<M>(lm))];
            // row 0 = (U + conj(Um))/2, row N/2 = (U - conj(Um))/(2i)
            cplx<T> A0 = {T(0.5) * (U.re + Um.re), T(0.5) * (U.im - Um.im)};
            cplx<T> Ah = {T(0.5) * (U.im + Um.im), T(-0.5) * (U.re - Um.re)};
            cplx<T> B0 = {T(0.5) * (V.re + Vm.re), T(0.5) * (V.im - Vm.im)};
            cplx<T> Bh = {T(0.5) * (V.im + Vm.im), T(-0.5) * (V.re - Vm.re)};
            if (a.filt) {
              const T f0 = T(a.filt[0 * N + lx]), fh = T(a.filt[(N / 2) * N + lx]);
              A0 = {A0.re * f0, A0.im * f0};
              B0 = {B0.re * f0, B0.im * f0};
              Ah = {Ah.re * fh, Ah.im * fh};
              Bh = {Bh.re * fh, Bh.im * fh};
            }
            const cplx<T> P0 = cross_power(A0, B0, norm), Ph = cross_power(Ah, Bh, norm);
            u2[S] = {P0.re - Ph.im, P0.im + Ph.re};
          });
        } else {
          static_for<M>([&](auto S) {
            cplx<T> A = Z[zidx<M>(km, 4 * S + l)], B = u2[S];
            if (a.filt) {
              const T f = T(a.filt[k * N + lb + S]);
              A = {A.re * f, A.im * f};
              B = {B.re * f, B.im * f};
            }
            u2[S] = cross_power(A, B, norm);
          });
        }
        quad_dif<M, +1>(u2, l);
        if (tz) {
          wave_lds_order();
          static_for<M>([&](auto Rr) {
            const int jj = 4 * Rr + l;
            Z[zidx<M>(0, jj)] = {u2[Rr].re, T(0)};
            Z[zidx<M>(N / 2, jj)] = {u2[Rr].im, T(0)};
          });
        } else {
          static_for<M>([&](auto Rr) { Z[zidx<M>(k, 4 * Rr + l)] = u2[Rr]; });
        }
      }
      __syncthreads();
      // ------------- phase 3: c2r along k, two columns per transform
      T va[M], vb[M];
      const int l = fresh(lane), quad = fresh(qd);
      const int ca = quad, cb = quad + N / 2;
      const int ib = M * brev2(l);
      const int sa = (ca + N / 2) % N, sb = (cb + N / 2) % N;
      if (active) {
        cplx<T> z[M];
        static_for<M>([&](auto Rr) {
          const int k = 4 * Rr + l;
          cplx<T> qa, qb;
          if (k <= N / 2) {
            qa = Z[zidx<M>(k, ca)];
            qb = Z[zidx<M>(k, cb)];
          } else {
            qa = cconj(Z[zidx<M>(N - k, ca)]);
            qb = cconj(Z[zidx<M>(N - k, cb)]);
          }
          z[Rr] = {qa.re - qb.im, qa.im + qb.re};
        });
        quad_dit<M, +1>(z, l);
        double lmax = 0.;
        static_for<M>([&](auto S) {
          va[S] = fabs(z[S].re);
          vb[S] = fabs(z[S].im);
          lmax = fmax(lmax, (double)va[S]);
          lmax = fmax(lmax, (double)vb[S]);
        });
        lmax = fmax(lmax, dpp<kQuadXor1>(lmax));
        lmax = fmax(lmax, dpp<kQuadXor2>(lmax));
        if (l == 0) atomicMax(amax, dbits(lmax));
      }
      __syncthreads();
      const double gmax = __longlong_as_double((long long)*amax);
      if (active) {
        unsigned int best = 0xffffffffu;
        static_for<M>([&](auto S) {
          const int si = ((ib + S + N / 2) % N) * N;
          if ((double)va[S] == gmax) best = min(best, (unsigned)(si + sa));
          if ((double)vb[S] == gmax) best = min(best, (unsigned)(si + sb));
        });
        best = min(best, (unsigned)__builtin_amdgcn_mov_dpp((int)best, kQuadXor1, 0xF, 0xF, true));
        best = min(best, (unsigned)__builtin_amdgcn_mov_dpp((int)best, kQuadXor2, 0xF, 0xF, true));
        if (l == 0 && best != 0xffffffffu) atomicMin(aidx, best);
        if constexpr (sizeof(T) == 4) {
          if (a.flag_list && gmax != 0.) {
            const double thr = gmax * (1. - kTieTol32);
            unsigned cnt = 0;
            static_for<M>([&](auto S) {
              cnt += ((double)va[S] >= thr) ? 1u : 0u;
              cnt += ((double)vb[S] >= thr) ? 1u : 0u;
            });
            if (cnt) atomicAdd(reinterpret_cast<unsigned int*>(gmx), cnt);
          }
        }
      }
      __syncthreads();
      const unsigned loc = *aidx;
      const bool nonzero = (gmax != 0.);
      int ixmax = N / 2, iymax = N / 2;
      if (nonzero) {
        ixmax = (int)(loc / N);
        iymax = (int)(loc % N);
      }
      bool fit = nonzero && ixmax > 0 && ixmax < N - 1 && iymax > 0 && iymax < N - 1;
      int hq = -1;  // hires grid maximum
      if constexpr (HIRES) {
        {
          fit = false;
          // 6 x 6 window of the surface about the integer peak and the 6 x 6
          // cdata block the grid reads, in Z (whose last reads were before
          // the barriers above)
          double* hw = reinterpret_cast<double*>(Z);
          double* hc = hw + kHiresWin * kHiresWin;
          unsigned long long* hkey = reinterpret_cast<unsigned long long*>(hc + kHiresWin * kHiresWin);
          unsigned int* hloc = reinterpret_cast<unsigned int*>(hkey + 1);
          if (active) {
            static_for<M>([&](auto S) {
              const int di = (ib + S + N / 2) % N - ixmax;
              if (di >= -2 && di <= 3) {
                const int da = sa - iymax, db = sb - iymax;
                if (da >= -2 && da <= 3) hw[(di + 2) * kHiresWin + (da + 2)] = (double)va[S];
                if (db >= -2 && db <= 3) hw[(di + 2) * kHiresWin + (db + 2)] = (double)vb[S];
              }
            });
            if (stid == 0) {
              *hkey = 0ull;
              *hloc = 0xffffffffu;
            }
          }
          __syncthreads();
          // a slot has 2N threads, fewer than the 36 entries for boxes up to 16
          if (active)
            for (int t = stid; t < kHiresWin * kHiresWin; t += Cfg::kSlotThreads)
              hires_cdata_entry(HiresWindow{hw, ixmax, iymax}, N, N, t, hc);
          __syncthreads();
          unsigned long long kbest = 0ull;
          int qbest = 0x7fffffff;
          if (active) {
#pragma unroll 1
            for (int q = stid; q < kHiresN * kHiresN; q += Cfg::kSlotThreads) {
              const unsigned long long k = hires_key(hires_value_cp(hc, N, N, ixmax, iymax, q));
              if (k > kbest) {
                kbest = k;
                qbest = q;
              }
            }
            atomicMax(hkey, kbest);
          }
          __syncthreads();
          if (active && kbest == *hkey) atomicMin(hloc, (unsigned)qbest);
          __syncthreads();
          hq = (int)*hloc;
          if (active) {
            // 3 x 3 neighbourhood of the grid maximum, where it is interior
            // (9 entries; box 4 has 8 threads per slot)
            for (int t = stid; t < 9; t += Cfg::kSlotThreads) {
              const int ix = hq / kHiresN + t / 3 - 1, iy = hq % kHiresN + t % 3 - 1;
              if (ix >= 0 && ix < kHiresN && iy >= 0 && iy < kHiresN)
                nb[t] = hires_value_cp(hc, N, N, ixmax, iymax, ix * kHiresN + iy);
            }
          }
        }
      }
      if (active && fit) {
        static_for<M>([&](auto S) {
          const int si = (ib + S + N / 2) % N;
          const int di = si - ixmax;
          if (di >= -1 && di <= 1) {
            const int da = sa - iymax, db = sb - iymax;
            if (da >= -1 && da <= 1) nb[(di + 1) * 3 + (da + 1)] = (double)va[S];
            if (db >= -1 && db <= 1) nb[(di + 1) * 3 + (db + 1)] = (double)vb[S];
          }
        });
      }
      __syncthreads();
      if (active && stid == 0) {
#pragma clang fp contract(off)
        double shiftsubx = 0., shiftsuby = 0.;
        if (hq >= 0) {
          double shx, shy;
          hires_shift(hq, nb, ixmax, iymax, &shx, &shy);
          res[0] = shx - (double)(N / 2);
          res[1] = shy - (double)(N / 2);
        } else if (fit) {
          // nb[(dx+1)*3 + (dy+1)] = surface at (ixmax+dx, iymax+dy)
          const double c00 = nb[4];
          const double fx = 0.5 * (nb[7] - nb[1]);
          const double fy = 0.5 * (nb[5] - nb[3]);
          const double fxx = (nb[7] + nb[1] - 2. * (c00));
          const double fyy = (nb[5] + nb[3] - 2. * (c00));
          const double fxy = 0.25 * (nb[8] + nb[0] - nb[6] - nb[2]);
          shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
          shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
          if constexpr (sizeof(T) == 4) {
            const double det = fxx * fyy - fxy * fxy;
            const double hn = fabs(fxx) + fabs(fyy) + 2. * fabs(fxy);
            const double bound =
                kRelErr32 * gmax * hn * (1. + 4. * (fabs(shiftsubx) + fabs(shiftsuby)));
            if (!(bound <= kFitTol32 * fabs(det))) flagged = true;
          }
        }
        if constexpr (sizeof(T) == 4) {
          if (nonzero && *reinterpret_cast<const unsigned int*>(gmx) > 1u) flagged = true;
        }
        if (hq < 0) {
          res[0] = ((double)ixmax + shiftsubx) - (double)(N / 2);
          res[1] = ((double)iymax + shiftsuby) - (double)(N / 2);
        }
      }
      __syncthreads();
      shiftx = res[0];
      shifty = res[1];
    }
    if (active && stid == 0) {
      const long long p = (long long)i * ny + j;
      a.vx[p] = shifty * a.deltinv * a.deltas;
      a.vy[p] = shiftx * a.deltinv * a.deltas;
      a.vm[p] = 1.;
      if constexpr (sizeof(T) == 4) {
        if (flagged && a.flag_list) a.flag_list[atomicAdd(a.flag_count, 1u)] = (int)p;
      }
    }
  }
}

}  // namespace flct
