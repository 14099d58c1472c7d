// FP64 tracker, row-shared form, boxes N = 4*M up to 96 (path "fp64-rowseg").
//
// The windowed sub-image of pixel (i, j) is g(x, c) = Wr(x) Wc(c) (f(imin + x,
// jmin + c) - fbar) (flct.c:815-870).  All pixels of image row i share imin
// and the row window Wr, so the transform along the box rows is shared by a
// segment of up to 64 neighbouring pixels of the row: for every image column
// u the segment's boxes touch,
//
//   Phi(k, u) = sum_x Wr(x) (f(imin + x, u) - mref) e^{-2 pi i k x / N},
//
// both images packed as f1 + i f2, mref the box mean of the segment's first
// pixel, and Wr^(k), the transform of Wr, from a column of ones.  Pixel j then
// starts from
//
//   Z(k, c) = Wc(c) (Phi(k, jmin + c) - (fbar - mref) Wr^(k)),
//
// which is the column transform of g1 + i g2 exactly in real arithmetic
// (the small correction keeps the cancellation of a large mean out of the
// shared transform), so the N per-pixel column transforms of k_track's phase
// 1 become loads of Phi.  Phases 2 and 3 are k_track's: row transforms of the
// unpacked spectra A = r2c(g1), B = r2c(g2), the optional filter on both, the
// cross power conj(A) B (1/N^2)^2 (flct.c:1440-1477), the inverse along jj,
// then c2r along k, |.|, the shift by N/2 (flct.c:1485-1510, 2111-2140), the
// first strict maximum (flct.c:2142-2161) and the 3 x 3 fit (flct.c:1536-
// 1556).  Only rows k = 0..N/2 are kept in LDS between the phases, half of
// k_track's plane, so four 64-box slots fit a CU and the kernel runs two
// waves per SIMD instead of one.  Phi lives in block-private device memory
// (N x (64 + N) complex per block).
//
// Degenerate boxes.  Where either image's box is constant and equal to its
// mean, the reference's windowed sub-image is exactly zero and so is its
// surface (the max == 0 branch, v = 0); the shared transform would leave
// rounding noise there instead (mref differs from that box's mean), so the
// exact box minimum and maximum, kept per column during the shared
// transform, classify each pixel: such pixels get the zero surface's result
// directly, and pixels whose box is constant at an inexact mean, or whose
// structure (max - min) is below 1e-4 of its distance from mref, are listed
// for re-evaluation by k_track, like the unsettled pixels of the precision
// forms.
//
// The threshold (flct.c:721-722) is evaluated per segment, like the
// split-BF16 row-segment kernel; skip lattices, -b and -h run k_track.
#include "track_rs.hpp"

namespace flct {

// box structure (max - min) below this fraction of the box's distance from
// the segment's reference mean: re-evaluated by k_track
constexpr double kRsRange = 1e-4;

// inverse of hires_key: the double of an order key
__device__ __forceinline__ double key_value(unsigned long long k) {
  const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __builtin_bit_cast(double, b);
}

// FP32 row tasks: pixels whose surface the FP32 arithmetic cannot settle
// are listed for FP64 re-evaluation, as in the FP32 tracker (a second cell
// within kTieTol32 of the maximum, or an ill-conditioned fit)
template <int M, class T = double>
struct RsCfg {
  static constexpr int N = 4 * M;
  static constexpr int kSlotThreads = 2 * N;
  static constexpr int kRow = TrackCfg<M, T>::kRow;
  static constexpr int kZBytes = ((N / 2 + 1) * kRow + 8) * int(sizeof(cplx<T>));
  // plane, row window, 9 neighbours, result, argmax words, box extrema,
  // near-maximum count
  static constexpr int kSlotBytes = ((kZBytes + N * 8 + 9 * 8 + 2 * 8 + 16 + 32 + 16) + 15) / 16 * 16;
  static constexpr int kColStat = kRsSeg + N;  // columns of a segment's boxes, at most
  static constexpr int kShared = 2 * N * 8 + N * 16 + kRsSeg * 4 + 16 + 4 * kColStat * 8;
  // FP64: boxes up to 36 fit two waves per SIMD in registers (at most 255
  // VGPRs); 40..48 run two waves with spills, which measured 27-32 % faster
  // than one wave (2048^2: 48.7, 72.2, 79.2 against 66.3, 105.6, 108.9 ms);
  // larger boxes need up to ~500 registers per lane and run one wave per
  // SIMD in blocks of at most 256 threads, so that the allocator may use the
  // accumulation registers as well (two waves were slower from box 52 up:
  // box 64 227 against 143 ms).  FP32 row tasks always run
  // two waves per SIMD: up to box 40 without spills, above it spilling 38
  // to ~120 registers, which measured 1.5-2.1x faster than one wave with
  // none (4096^2 box 64: 454 against 685 ms; 2048^2 box 52: 96 against
  // 148 ms, box 80: 242 against 497 ms)
  static constexpr bool kTwo = (M <= 12) || (sizeof(T) == 4);
  static constexpr int kMinWaves = kTwo ? 2 : 1;
  static constexpr int kMaxThreads = kTwo ? 512 : 256;
  static constexpr int fit() {
    int p = 1;
    while (p < 8 && (p + 1) * kSlotBytes + kShared <= 160 * 1024 &&
           (p + 1) * kSlotThreads <= kMaxThreads)
      ++p;
    return p;
  }
  static constexpr int P = fit();
  static constexpr int kThreads = ((P * kSlotThreads + 63) / 64) * 64;
  static constexpr int kLdsBytes = P * kSlotBytes + kShared;
};

template <int M, class T>
__global__ void __launch_bounds__((RsCfg<M, T>::kThreads), (RsCfg<M, T>::kMinWaves)) k_track_rs(RsArgs A) {
  using Cfg = RsCfg<M, T>;
  constexpr int N = 4 * M;
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const int slot = tid / Cfg::kSlotThreads;
  const int stid = tid - slot * Cfg::kSlotThreads;
  const int SP = run_slots(A.t.slots, Cfg::P);  // slots of this launch
  const bool in_slot = slot < SP;
  const int lane = stid & 3, qd = stid >> 2;
  double* gw = reinterpret_cast<double*>(lds);                     // 2N - 1
  cplx<double>* wrh = reinterpret_cast<cplx<double>*>(gw + 2 * N);  // N: Wr^(k)
  int* slist = reinterpret_cast<int*>(wrh + N);                    // kRsSeg
  int* scount = slist + kRsSeg;
  double* colst = reinterpret_cast<double*>(scount + 4);  // [4][kColStat]: min f1, max f1, min f2, max f2
  unsigned char* base = reinterpret_cast<unsigned char*>(colst + 4 * Cfg::kColStat) +
                        (in_slot ? slot : 0) * Cfg::kSlotBytes;
  cplx<T>* Z = reinterpret_cast<cplx<T>*>(base);
  // N doubles: with whole-wave slots the per-wave box extrema words
  double* wy = reinterpret_cast<double*>(base + Cfg::kZBytes);
  double* nb = wy + N;
  double* res = nb + 9;
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(res + 2);
  unsigned int* aidx = reinterpret_cast<unsigned int*>(amax + 1);
  unsigned long long* bst = amax + 2;  // box min f1, max f1, min f2, max f2 (order keys)
  unsigned int* ntie = reinterpret_cast<unsigned int*>(bst + 4);  // FP32: near-maximum cells
  const TrackArgs& a = A.t;
  for (int t = tid; t < 2 * N - 1; t += blockDim.x) gw[t] = a.gw[t];

  const int nx = a.nx, ny = a.ny, nsize = a.nsize;
  const int ns = ny - N + 1;
  const T norm = T((1. / ((double)N * N)) * (1. / ((double)N * N)));
  const int us = A.ustride;
  double2* phi = A.phi + (size_t)blockIdx.x * N * us;  // [k][u]
  // Segment order: blocks are dealt to the 8 XCDs round robin; XCD x owns
  // image rows [x nx/8, (x+1) nx/8) and its blocks split that range's
  // segments, band by band, into contiguous runs of consecutive rows, whose
  // box rows overlap and stay in the XCD's L2.
  const int G = gridDim.x;
  const int nxcd = G < 8 ? G : 8;
  const int xc = blockIdx.x % nxcd, kb = blockIdx.x / nxcd;
  const int nbx = (G - xc + nxcd - 1) / nxcd;
  const int r0 = (int)((long long)nx * xc / nxcd), r1 = (int)((long long)nx * (xc + 1) / nxcd);
  const long long nsx = (long long)(r1 - r0) * A.nband;
  const long long s0 = nsx * kb / nbx, s1 = nsx * (kb + 1) / nbx;
  auto jmin_of = [&](int j) {  // box clamp, flct.c:780-786
    const int jmin0 = (0 > (j - (nsize - 1))) ? 0 : j - (nsize - 1);
    const int jmax0 = ((ny - 1) < (j + nsize)) ? ny - 1 : j + nsize;
    return (jmax0 == ny - 1) ? ny - 1 - (2 * nsize - 1) : jmin0;
  };
  __syncthreads();

  for (long long sidx = s0; sidx < s1; ++sidx) {
    const int band = (int)(sidx / (r1 - r0));
    const int i = r0 + (int)(sidx - (long long)band * (r1 - r0));
    const int jb0 = (int)((long long)band * ny / A.nband);
    const int jb1 = (int)((long long)(band + 1) * ny / A.nband);
    const int P = jb1 - jb0;
    // box rows, flct.c:773-779
    const int imin0 = (0 > (i - (nsize - 1))) ? 0 : i - (nsize - 1);
    const int imax0 = ((nx - 1) < (i + nsize)) ? nx - 1 : i + nsize;
    const int imin = (imax0 == nx - 1) ? nx - 1 - (2 * nsize - 1) : imin0;
    const int cb = jmin_of(jb0);
    const int U = jmin_of(jb1 - 1) - cb + N;  // box columns of the segment

    __syncthreads();  // the previous segment is done with Phi, Wr^ and slist
    // tracked pixels of the segment: threshold rule (flct.c:721-722)
    if (tid < 64) {
      bool tr = false;
      if (tid < P) {
#pragma clang fp contract(off)
        const long long t = (long long)i * ny + (jb0 + tid);
        const double fabsbar = 0.5 * (fabs(a.f1[t] + a.f2[t]));
        tr = !(fabsbar < resolve_thresh(A.thresh, A.thresh_rel ? A.absmax : nullptr));
      }
      const unsigned long long m = __ballot(tr);
      if (tr) slist[__popcll(m & ((1ull << tid) - 1ull))] = tid;
      if (tid == 0) *scount = __popcll(m);
    }
    __syncthreads();
    const int cnt = *scount;
    if (cnt == 0) continue;
    const double mref1 = a.mean1[(long long)imin * ns + cb];
    const double mref2 = a.mean2[(long long)imin * ns + cb];

    // ------------- shared transforms along the box rows: Phi(k, u) for
    // u < U, and Wr^(k) from a column of ones (u = U)
    if (in_slot) {
      const int l = fresh(lane);
      const int nq = SP * (Cfg::kSlotThreads / 4);
      const int kbl = M * brev2(l);
      const double* p1 = a.f1 + (long long)imin * ny + cb;
      const double* p2 = a.f2 + (long long)imin * ny + cb;
#pragma unroll 1
      for (int u = tid >> 2; u <= U; u += nq) {
        cplx<double> z[M];
        double mn1 = 0., mx1 = 0., mn2 = 0., mx2 = 0.;
        static_for<M>([&](auto Rr) {
          const int x = 4 * Rr + l;
          const double w = gw[x + imin - i + N - 1];
          if (u < U) {
            const long long off = (long long)x * ny + u;
            const double v1 = p1[off], v2 = p2[off];
            if constexpr (Rr == 0) {
              mn1 = mx1 = v1;
              mn2 = mx2 = v2;
            } else {
              mn1 = fmin(mn1, v1);
              mx1 = fmax(mx1, v1);
              mn2 = fmin(mn2, v2);
              mx2 = fmax(mx2, v2);
            }
            z[Rr] = {w * (v1 - mref1), w * (v2 - mref2)};
          } else {
            z[Rr] = {w, 0.};
          }
        });
        if (u < U) {
          // column minimum and maximum over the box rows
          mn1 = fmin(mn1, dpp<kQuadXor1>(mn1));
          mx1 = fmax(mx1, dpp<kQuadXor1>(mx1));
          mn2 = fmin(mn2, dpp<kQuadXor1>(mn2));
          mx2 = fmax(mx2, dpp<kQuadXor1>(mx2));
          mn1 = fmin(mn1, dpp<kQuadXor2>(mn1));
          mx1 = fmax(mx1, dpp<kQuadXor2>(mx1));
          mn2 = fmin(mn2, dpp<kQuadXor2>(mn2));
          mx2 = fmax(mx2, dpp<kQuadXor2>(mx2));
          if (l == 0) {
            colst[0 * Cfg::kColStat + u] = mn1;
            colst[1 * Cfg::kColStat + u] = mx1;
            colst[2 * Cfg::kColStat + u] = mn2;
            colst[3 * Cfg::kColStat + u] = mx2;
          }
        }
        quad_dit<M, -1>(z, l);
        if (u < U) {
          static_for<M>([&](auto S) { phi[(size_t)(kbl + S) * us + u] = make_double2(z[S].re, z[S].im); });
        } else {
          static_for<M>([&](auto S) { wrh[kbl + S] = z[S]; });
        }
      }
    }
    __threadfence_block();
    __syncthreads();

    // ------------- the segment's tracked pixels, one per slot at a time
    // Where a slot is one or two whole waves (boxes 32 and 64), each wave
    // reduces the box extrema of its slot itself, into its own four LDS words
    // (in the otherwise unused wy area), so neither the set-up nor the extrema
    // need a block barrier; the column window is read from gw directly.
    constexpr bool kWaveExt = (Cfg::kSlotThreads % 64 == 0) && N <= 64;
    unsigned long long* bstw = reinterpret_cast<unsigned long long*>(wy) + 4 * (stid >> 6);
    // The 3 x 3 fit and the outputs of a slot's pixel are done by one lane
    // of the slot (lane 64 where the slot has more than one wave: its first
    // wave carries the rows 0 / N/2 task), at the start of the slot's next
    // pixel or after its last one: the lane
    // keeps the peak data in registers, nb is not rewritten before the next
    // pixel's gather, and no block barrier waits on the single-lane FP64 work.
    constexpr int kFitLane = (Cfg::kSlotThreads >= 128) ? 64 : 0;
    constexpr bool kBranchTasks = sizeof(T) == 8 && (M == 20 || M == 21);
    bool pend = false, pexact = false, pfit = false, pflag = false;
    int pj = 0, pix = N / 2, piy = N / 2;
    double pgmax = 0.;
    auto finish = [&]() {
#pragma clang fp contract(off)
      double shiftsubx = 0., shiftsuby = 0.;
      bool flagged = pflag;
      if (pfit) {
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
          // first-order response of the fit to relative errors kRelErr32
          // of the nine values
          const double det = fxx * fyy - fxy * fxy;
          const double hn = fabs(fxx) + fabs(fyy) + 2. * fabs(fxy);
          const double bound =
              kRelErr32 * pgmax * hn * (1. + 4. * (fabs(shiftsubx) + fabs(shiftsuby)));
          if (!(bound <= kFitTol32 * fabs(det))) flagged = true;
        }
      }
      const double shiftx = ((double)pix + shiftsubx) - (double)(N / 2);
      const double shifty = ((double)piy + shiftsuby) - (double)(N / 2);
      const long long p = (long long)i * ny + pj;
      if (pexact) {
        a.flag_list[atomicAdd(a.flag_count, 1u)] = (int)p;  // k_track writes this pixel
      } else {
        a.vx[p] = shifty * a.deltinv * a.deltas;  // transp = 1 (flct.c:901-905)
        a.vy[p] = shiftx * a.deltinv * a.deltas;
        a.vm[p] = 1.;
        if (flagged) a.flag_list[atomicAdd(a.flag_count, 1u)] = (int)p;  // FP32: re-evaluated
      }
    };
    for (int rp = 0; rp < cnt; rp += SP) {
      if (pend) {
        finish();
        pend = false;
      }
      const bool active = in_slot && (rp + slot) < cnt;
      int j = 0, jmin = 0;
      double d1 = 0., d2 = 0., f1b = 0., f2b = 0.;
      if (active) {
        j = jb0 + slist[rp + slot];
        jmin = jmin_of(j);
        f1b = a.mean1[(long long)imin * ns + jmin];
        f2b = a.mean2[(long long)imin * ns + jmin];
        d1 = f1b - mref1;
        d2 = f2b - mref2;
        if constexpr (kWaveExt) {
          if ((stid & 63) == 0) {
            bstw[0] = ~0ull;
            bstw[1] = 0ull;
            bstw[2] = ~0ull;
            bstw[3] = 0ull;
          }
        }
        if (stid == 0) {
          *amax = 0ull;
          *aidx = 0xffffffffu;
          bst[0] = ~0ull;
          bst[1] = 0ull;
          bst[2] = ~0ull;
          bst[3] = 0ull;
          *ntie = 0u;
        }
      }
      if constexpr (kWaveExt) {
        wave_lds_order();
      } else {
        __syncthreads();
      }
      // exact box minimum and maximum of both images from the column values
      // (with kWaveExt in every wave of the slot, else in its first N lanes)
      unsigned long long* bs = kWaveExt ? bstw : bst;
      if (active && (kWaveExt ? (stid & 63) : stid) < N) {
        const int u = jmin - cb + (kWaveExt ? (stid & 63) : stid);
        double mn1 = colst[0 * Cfg::kColStat + u], mx1 = colst[1 * Cfg::kColStat + u];
        double mn2 = colst[2 * Cfg::kColStat + u], mx2 = colst[3 * Cfg::kColStat + u];
        mn1 = fmin(mn1, dpp<kQuadXor1>(mn1));
        mx1 = fmax(mx1, dpp<kQuadXor1>(mx1));
        mn2 = fmin(mn2, dpp<kQuadXor1>(mn2));
        mx2 = fmax(mx2, dpp<kQuadXor1>(mx2));
        mn1 = fmin(mn1, dpp<kQuadXor2>(mn1));
        mx1 = fmax(mx1, dpp<kQuadXor2>(mx1));
        mn2 = fmin(mn2, dpp<kQuadXor2>(mn2));
        mx2 = fmax(mx2, dpp<kQuadXor2>(mx2));
        if ((stid & 3) == 0) {
          atomicMin(&bs[0], hires_key(mn1));
          atomicMax(&bs[1], hires_key(mx1));
          atomicMin(&bs[2], hires_key(mn2));
          atomicMax(&bs[3], hires_key(mx2));
        }
      }
      if constexpr (kWaveExt) {
        wave_lds_order();
      } else {
        __syncthreads();
      }
      bool zero = false, exact = false;  // zero surface; re-evaluate in k_track
      double s1 = 1., s2 = 1.;             // FP32: image scales
      if (active) {
        const double mn1 = key_value(bs[0]), mx1 = key_value(bs[1]);
        const double mn2 = key_value(bs[2]), mx2 = key_value(bs[3]);
        zero = (mn1 == mx1 && mx1 == f1b) || (mn2 == mx2 && mx2 == f2b);
        const double o1 = fmax(fabs(mx1 - mref1), fabs(mn1 - mref1));
        const double o2 = fmax(fabs(mx2 - mref2), fabs(mn2 - mref2));
        exact = !zero && (!(mx1 - mn1 > kRsRange * o1) || !(mx2 - mn2 > kRsRange * o2));
        if constexpr (sizeof(T) == 4) {
          // powers of two bringing each image's largest |f - fbar| (a bound
          // of its windowed values) to [1, 2): exact, the argmax and the fit
          // are scale-invariant, and the row tasks round each image to FP32
          // at its own magnitude
          const double g1 = fmax(fabs(mx1 - f1b), fabs(mn1 - f1b));
          const double g2 = fmax(fabs(mx2 - f2b), fabs(mn2 - f2b));
          if (g1 > 0.) s1 = ldexp(1.0, -ilogb(g1));
          if (g2 > 0.) s2 = ldexp(1.0, -ilogb(g2));
        }
      }
      // A double of image X (1, 2) as a row-task value: FP64 unchanged, FP32
      // scaled and rounded once
      auto tv = [&](double x, double sc) -> T {
        if constexpr (sizeof(T) == 8) {
          (void)sc;
          return x;
        } else {
          return (T)(x * sc);
        }
      };
      const bool compute = active && !zero && !exact;
      // ------------- phase 2: row transforms of A and B, cross power,
      // inverse along jj; rows k = 0..N/2 of the result stay in LDS
      if (compute) {
        const int l = fresh(lane), quad = fresh(qd);
        const int t = quad;
        const int lb = M * brev2(l);
        const int c0 = jmin - cb;
        // Z(k, c) = Wc(c) (Phi(k, c0 + c) - (d1 + i d2) Wr^(k)), in FP64
        auto zval = [&](int k, int c) -> cplx<double> {
          const double2 f = phi[(size_t)k * us + c0 + c];
          const cplx<double> w = wrh[k];
          const double cr = d1 * w.re - d2 * w.im, ci = d1 * w.im + d2 * w.re;
          const double wc = gw[c + jmin - j + N - 1];
          return {wc * (f.x - cr), wc * (f.y - ci)};
        };
        // FP64 boxes 80 and 84 (one wave per SIMD, 200-226 accumulation
        // registers): the shared-transform form below allocated worse there
        // (404 and 493 ms at 2048^2 against 370 and 417 with a branch per
        // task kind), so these two keep the branch per task kind.
        if constexpr (kBranchTasks) {
          if (t >= 1) {
            const int k = t, km = N - t;
            // A = (Zk + conj(Zm))/2 and B = (Zk - conj(Zm))/(2i) from one read
            // of Phi rows k and N - k; both spectra stay in registers
            cplx<T> u[M], v[M];
            static_for<M>([&](auto Rr) {
              const int jj = 4 * Rr + l;
              const cplx<double> zk = zval(k, jj), zm = zval(km, jj);
              u[Rr] = {tv(0.5 * (zk.re + zm.re), s1), tv(0.5 * (zk.im - zm.im), s1)};
              v[Rr] = {tv(0.5 * (zk.im + zm.im), s2), tv(-0.5 * (zk.re - zm.re), s2)};
            });
            quad_dit<M, -1>(u, l);
            quad_dit<M, -1>(v, l);
            static_for<M>([&](auto S) {
              cplx<T> Aa = u[S], B = v[S];
              if (a.filt) {
                const T f = T(a.filt[k * N + lb + S]);
                Aa = {Aa.re * f, Aa.im * f};
                B = {B.re * f, B.im * f};
              }
              u[S] = cross_power(Aa, B, norm);
            });
            quad_dif<M, +1>(u, l);
            static_for<M>([&](auto Rr) { Z[zidx<M>(k, 4 * Rr + l)] = u[Rr]; });
          } else {
            // rows k = 0 and k = N/2 of A and B are real along jj: one packed
            // transform per image, parked in rows N/2 (A) and 0 (B)
            cplx<T> u[M], v[M];
            static_for<M>([&](auto Rr) {
              const int jj = 4 * Rr + l;
              const cplx<double> z0 = zval(0, jj), zh = zval(N / 2, jj);
              u[Rr] = {tv(z0.re, s1), tv(zh.re, s1)};
              v[Rr] = {tv(z0.im, s2), tv(zh.im, s2)};
            });
            quad_dit<M, -1>(u, l);
            wave_lds_order();
            static_for<M>([&](auto S) { Z[zidx<M>(N / 2, 4 * S + l)] = u[S]; });
            wave_lds_order();
            quad_dit<M, -1>(v, l);
            wave_lds_order();
            static_for<M>([&](auto S) { Z[zidx<M>(0, 4 * S + l)] = v[S]; });
            wave_lds_order();
            static_for<M>([&](auto S) {
              const int lx = lb + S;
              const int lm = (N - lx) % N;
              const cplx<T> Uu = Z[zidx<M>(N / 2, park_col<M>(lx))];
              const cplx<T> Um = Z[zidx<M>(N / 2, park_col<M>(lm))];
              const cplx<T> V = Z[zidx<M>(0, park_col<M>(lx))];
              const cplx<T> Vm = Z[zidx<M>(0, park_col<M>(lm))];
              // row 0 = (U + conj(Um))/2, row N/2 = (U - conj(Um))/(2i)
              cplx<T> A0 = {T(0.5) * (Uu.re + Um.re), T(0.5) * (Uu.im - Um.im)};
              cplx<T> Ah = {T(0.5) * (Uu.im + Um.im), T(-0.5) * (Uu.re - Um.re)};
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
              u[S] = {P0.re - Ph.im, P0.im + Ph.re};
            });
            quad_dif<M, +1>(u, l);
            wave_lds_order();
            static_for<M>([&](auto Rr) {
              const int jj = 4 * Rr + l;
              Z[zidx<M>(0, jj)] = {u[Rr].re, T(0)};
              Z[zidx<M>(N / 2, jj)] = {u[Rr].im, T(0)};
            });
          }
        } else {
          // Row task t >= 1: rows k = t and N - t; A = (Zk + conj(Zm))/2 and
          // B = (Zk - conj(Zm))/(2i) from one read of the two Phi rows, both
          // spectra in registers.  Task 0: rows 0 and N/2 of A and B are real
          // along jj, so each image is one packed transform (row 0 + i row
          // N/2), unpacked after the forward transforms through LDS.  The
          // transforms are shared code for every task: only the input, the
          // unpack and the stores depend on it, so the wave holding task 0
          // does not run the transforms twice.
          const bool tz = (t == 0);
          const int k = tz ? 0 : t, km = tz ? N / 2 : N - t;
          cplx<T> u[M], v[M];
          static_for<M>([&](auto Rr) {
            const int jj = 4 * Rr + l;
            const cplx<double> zk = zval(k, jj), zm = zval(km, jj);
            const cplx<T> ug = {tv(0.5 * (zk.re + zm.re), s1), tv(0.5 * (zk.im - zm.im), s1)};
            const cplx<T> vg = {tv(0.5 * (zk.im + zm.im), s2), tv(-0.5 * (zk.re - zm.re), s2)};
            const cplx<T> up = {tv(zk.re, s1), tv(zm.re, s1)};
            const cplx<T> vp = {tv(zk.im, s2), tv(zm.im, s2)};
            u[Rr] = tz ? up : ug;
            v[Rr] = tz ? vp : vg;
          });
          quad_dit<M, -1>(u, l);
          quad_dit<M, -1>(v, l);
          if (tz) {
            // the packed transforms parked in rows N/2 (A) and 0 (B)
            wave_lds_order();
            static_for<M>([&](auto S) { Z[zidx<M>(N / 2, 4 * S + l)] = u[S]; });
            static_for<M>([&](auto S) { Z[zidx<M>(0, 4 * S + l)] = v[S]; });
            wave_lds_order();
            static_for<M>([&](auto S) {
              const int lx = lb + S;
              const int lm = (N - lx) % N;
              const cplx<T> Uu = Z[zidx<M>(N / 2, park_col<M>(lx))];
              const cplx<T> Um = Z[zidx<M>(N / 2, park_col<M>(lm))];
              const cplx<T> V = Z[zidx<M>(0, park_col<M>(lx))];
              const cplx<T> Vm = Z[zidx<M>(0, park_col<M>(lm))];
              // row 0 = (U + conj(Um))/2, row N/2 = (U - conj(Um))/(2i)
              cplx<T> A0 = {T(0.5) * (Uu.re + Um.re), T(0.5) * (Uu.im - Um.im)};
              cplx<T> Ah = {T(0.5) * (Uu.im + Um.im), T(-0.5) * (Uu.re - Um.re)};
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
              u[S] = {P0.re - Ph.im, P0.im + Ph.re};
            });
          } else {
            static_for<M>([&](auto S) {
              cplx<T> Aa = u[S], B = v[S];
              if (a.filt) {
                const T f = T(a.filt[k * N + lb + S]);
                Aa = {Aa.re * f, Aa.im * f};
                B = {B.re * f, B.im * f};
              }
              u[S] = cross_power(Aa, B, norm);
            });
          }
          quad_dif<M, +1>(u, l);
          if (tz) {
            wave_lds_order();
            static_for<M>([&](auto Rr) {
              const int jj = 4 * Rr + l;
              Z[zidx<M>(0, jj)] = {u[Rr].re, T(0)};
              Z[zidx<M>(N / 2, jj)] = {u[Rr].im, T(0)};
            });
          } else {
            static_for<M>([&](auto Rr) { Z[zidx<M>(k, 4 * Rr + l)] = u[Rr]; });
          }
        }
      }
      __syncthreads();
      // ------------- phase 3: c2r along k, two columns per transform
      T va[M], vb[M];
      const int l = fresh(lane), quad = fresh(qd);
      const int ca = quad, cbq = quad + N / 2;
      const int ib = M * brev2(l);
      const int sa = (ca + N / 2) % N, sb = (cbq + N / 2) % N;
      if (compute) {
        cplx<T> z[M];
        static_for<M>([&](auto Rr) {
          const int k = 4 * Rr + l;
          cplx<T> qa, qb;
          if (k <= N / 2) {
            qa = Z[zidx<M>(k, ca)];
            qb = Z[zidx<M>(k, cbq)];
          } else {
            qa = cconj(Z[zidx<M>(N - k, ca)]);
            qb = cconj(Z[zidx<M>(N - k, cbq)]);
          }
          z[Rr] = {qa.re - qb.im, qa.im + qb.re};
        });
        quad_dit<M, +1>(z, l);
        double lmax = 0.;
        static_for<M>([&](auto S) {
          va[S] = fabs(z[S].re);
          vb[S] = fabs(z[S].im);
          lmax = fmax(lmax, va[S]);
          lmax = fmax(lmax, vb[S]);
        });
        lmax = fmax(lmax, dpp<kQuadXor1>(lmax));
        lmax = fmax(lmax, dpp<kQuadXor2>(lmax));
        if (l == 0) atomicMax(amax, dbits(lmax));
      }
      __syncthreads();
      const double gmax = __longlong_as_double((long long)*amax);
      if (compute) {
        unsigned int best = 0xffffffffu;
        static_for<M>([&](auto S) {
          const int si = ((ib + S + N / 2) % N) * N;
          if (va[S] == gmax) best = min(best, (unsigned)(si + sa));
          if (vb[S] == gmax) best = min(best, (unsigned)(si + sb));
        });
        best = min(best, (unsigned)__builtin_amdgcn_mov_dpp((int)best, kQuadXor1, 0xF, 0xF, true));
        best = min(best, (unsigned)__builtin_amdgcn_mov_dpp((int)best, kQuadXor2, 0xF, 0xF, true));
        if (l == 0 && best != 0xffffffffu) atomicMin(aidx, best);
        if constexpr (sizeof(T) == 4) {
          // cells within kTieTol32 of the maximum (one is the maximum itself)
          if (gmax != 0.) {
            const double thr = gmax * (1. - kTieTol32);
            unsigned cnt = 0;
            static_for<M>([&](auto S) {
              cnt += ((double)va[S] >= thr) ? 1u : 0u;
              cnt += ((double)vb[S] >= thr) ? 1u : 0u;
            });
            if (cnt) atomicAdd(ntie, cnt);
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
      const bool fit = nonzero && ixmax > 0 && ixmax < N - 1 && iymax > 0 && iymax < N - 1;
      if (compute && fit) {
        static_for<M>([&](auto S) {
          const int si = (ib + S + N / 2) % N;
          const int di = si - ixmax;
          if (di >= -1 && di <= 1) {
            const int da = sa - iymax, db = sb - iymax;
            if (da >= -1 && da <= 1) nb[(di + 1) * 3 + (da + 1)] = va[S];
            if (db >= -1 && db <= 1) nb[(di + 1) * 3 + (db + 1)] = vb[S];
          }
        });
      }
      if (active && stid == kFitLane) {
        pend = true;
        pj = j;
        pexact = exact;
        pfit = fit;
        pix = ixmax;
        piy = iymax;
        pgmax = gmax;
        if constexpr (sizeof(T) == 4) pflag = nonzero && *ntie > 1u;
      }
      __syncthreads();  // Z, nb and the maxima are rewritten by the next pixel
    }
    if (pend) {
      finish();
      pend = false;
    }
  }
}

// LDS bytes and threads of a block of sp slots (1..P)
template <int M, class T>
static void rs_block(int sp, size_t* lds, int* threads) {
  using Cfg = RsCfg<M, T>;
  *lds = (size_t)sp * Cfg::kSlotBytes + Cfg::kShared;
  *threads = (sp * Cfg::kSlotThreads + 63) / 64 * 64;
}

template <int M, class T>
static int launch_rs(const LaunchCtx& c, const RsArgs& a, long long grid) {
  size_t lds = 0;
  int threads = 0;
  rs_block<M, T>(run_slots(a.t.slots, RsCfg<M, T>::P), &lds, &threads);
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track_rs<M, T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  hipLaunchKernelGGL((k_track_rs<M, T>), dim3((unsigned)grid), dim3(threads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

template <class T>
static long long rs_grid_t(int M, int num_cu, long long nseg, int slots) {
  size_t lds = 0;
  int threads = 0;
  static_for<kTrackMaxM>([&](auto I) {
    constexpr int m = I + 1;
    if (m == M) rs_block<m, T>(run_slots(slots, RsCfg<m, T>::P), &lds, &threads);
  });
  if (lds == 0) return 0;
  long long per_cu = (160LL * 1024) / (long long)lds;
  if (per_cu > 2048 / threads) per_cu = 2048 / threads;
  if (per_cu < 1) per_cu = 1;
  if (per_cu > (sizeof(T) == 4 ? 8 : 4)) per_cu = (sizeof(T) == 4 ? 8 : 4);
  long long g = (long long)num_cu * per_cu;
  if (g > nseg) g = nseg;
  return g < 1 ? 1 : g;
}

long long rs_grid(int M, int num_cu, long long nseg, int slots, bool fp32) {
  return fp32 ? rs_grid_t<float>(M, num_cu, nseg, slots) : rs_grid_t<double>(M, num_cu, nseg, slots);
}

int rs_slots(int M, bool fp32) {
  int p = 0;
  static_for<kTrackMaxM>([&](auto I) {
    constexpr int m = I + 1;
    if (m == M) p = fp32 ? RsCfg<m, float>::P : RsCfg<m, double>::P;
  });
  return p;
}

int dispatch_track_rs(int M, const LaunchCtx& c, const RsArgs& a, long long grid, bool fp32) {
  int rc = FLCT_HIP_E_UNSUPPORTED;
  static_for<kTrackMaxM>([&](auto I) {
    constexpr int m = I + 1;
    if (m == M) rc = fp32 ? launch_rs<m, float>(c, a, grid) : launch_rs<m, double>(c, a, grid);
  });
  return rc;
}

}  // namespace flct
