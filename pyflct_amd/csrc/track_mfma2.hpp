// Split-BF16 MFMA tracker, radix-2 register-chained form, for boxes
// N = 36..64 with N % 4 == 0 (Nh = N/2 <= 32) on gfx950.
//
// One pixel is handled by two waves; wave kt owns the spectral columns
// k = 16 kt + (lane % 16) (k is the frequency along the box rows x, half
// spectrum).  With D_h[a][b] = exp(-2 pi i a b / Nh):
//
//   stage 1 (x -> k, radix-2 DIT on x):  E = G_even^T D_h, O = G_odd^T D_h,
//            T[c][k] = E + w^k O  (w = exp(-2 pi i / N));  the real DC and
//            Nyquist columns are packed as T[c][0] = T_dc + i T_nyq;
//   stage 2 (c -> l, radix-2 DIT on c):  S[l] = E' + w^l O',
//            S[l + Nh] = E' - w^l O';
//   cross power P = conj(S1) S2 normfac (flct.c:1461-1477), times filter^2
//            (flct.c:1445-1457); the packed column is unpacked through the
//            l <-> -l partner and re-packed as P_dc + i P_nyq;
//   inverse 1 (l -> jj, radix-2 DIF on l): U[2m+q][k] from Q+ = P[l] +
//            P[l+Nh] and Q- = (P[l] - P[l+Nh]) conj(w^l);
//   inverse 2 (k -> ii, Hermitian half sum): c[ii][jj] = Ce + Co,
//            c[ii + Nh][jj] = Ce - Co over even and odd k, with the real and
//            imaginary parts of U both laid along the contraction.
//
// Every MFMA output tile (4 consecutive rows per lane, fixed column) is the
// next stage's B operand without leaving registers: the contraction index is
// permuted to the rows the lane holds, and the matching A operand table is
// stored in that permuted order (TA[rho][kappa] = D_h[sig(rho)][pi1(kappa)],
// one table serving stage 2 and, conjugated, inverse 1).  Only G (built from
// the image) and U (transposed for the last contraction) pass through LDS.
// Products are hi*hi + hi*lo + lo*hi of BF16 splits, accumulated in FP32
// (v_mfma_f32_16x16x32_bf16).  Window, box means, argmax and the 3x3 Taylor
// fit (flct.c:1512-1556) follow the plain form; images are scaled by powers
// of two (exact, and the peak location is scale invariant).
#pragma once

#include "track_kernels.hpp"
#include "track_mfma.hpp"

namespace flct {

typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

template <int PIX>
struct R2Cfg {
  // bf16 row stride of 32-wide rows: 40 (80 B).  48 makes the fragment loads
  // free of bank conflicts (as in the row-segment kernel) but measured 4 %
  // slower here: 4.81 against 4.62 ms over the edge columns of a 4096^2
  // pair with 64 x 64 boxes
  static constexpr int RS = 40;
  static constexpr int kTabPlane = 32 * RS;  // one plane of a 32-row table
  static constexpr int kTabEl = 4 * kTabPlane;
  static constexpr int kPixPlane = 64 * RS;  // one plane of G or U
  static constexpr int kPixEl = 4 * kPixPlane;
  // per pixel: G/U planes, row windows (2 x 32 float), column-0 exchange
  // (2 x 64 float2), packed P (64 float2), 9 neighbours, argmax words, fit
  static constexpr int kPixBytes = kPixEl * 2 + 64 * 4 + 128 * 8 + 64 * 8 + 9 * 8 + 16 + 16;
  static constexpr int kPixStride = (kPixBytes + 15) / 16 * 16;
  static constexpr int kThreads = PIX * 128;
  static constexpr int kLdsBytes = 2 * kTabEl * 2 + 32 * 8 + PIX * kPixStride;
  static_assert(kLdsBytes <= 80 * 1024, "two blocks of k_track_r2 per CU");
};

// row permutation of the stage-2 / inverse-1 outputs: r = 16a + 4g + b
// holds l = 16a + 8(b >> 1) + 2g + (b & 1)
__device__ __forceinline__ int r2_sig(int r) {
  return (r & 16) | ((r & 2) << 2) | ((r >> 1) & 6) | (r & 1);
}
// contraction permutation of stage 2: kappa = 8g + j holds c' = 8(j >> 1) + 2g + (j & 1)
__device__ __forceinline__ int r2_pi1(int k) {
  return 8 * ((k & 7) >> 1) + 2 * (k >> 3) + (k & 1);
}

struct Frag2 {
  bf16x8 h, l;
};
struct CFrag {
  Frag2 re, im;
};
struct CF4 {
  f32x4 re, im;
};

// two floats -> packed BF16 hi pair and lo pair (hi = RNE(x), lo = RNE(x - hi)).
// The residuals come from one dot2 each: x + (-1) hi_x + 0 hi_y, whose exact
// value x - hi_x is representable in FP32 (hi_x is x rounded to 8 bits), so
// any rounding of the dot product returns it unchanged; 4 VALU instructions
// per pair instead of unpacking hi (2) and subtracting (2).  The (-1, 0)
// and (0, -1) operands are passed in scalar registers: as constants the
// compiler encodes (-1, 0) as the inline constant -1.0, which v_dot2c_f32_bf16
// on gfx950 reads as the FP32 bit pattern, i.e. (0, -1) (measured: wrong lo
// parts for 32763 of 32768 random pairs; the register form matches the
// unpack-and-subtract split on all of them)
__device__ __forceinline__ void split2k(float a, float b, unsigned& hi, unsigned& lo, unsigned k0,
                                        unsigned k1) {
  const bf16x2 h = __builtin_convertvector((f32x2v){a, b}, bf16x2);
  const bf16x2 m0 = __builtin_bit_cast(bf16x2, k0), m1 = __builtin_bit_cast(bf16x2, k1);
  const float ra = __builtin_amdgcn_fdot2_f32_bf16(h, m0, a, false);
  const float rb = __builtin_amdgcn_fdot2_f32_bf16(h, m1, b, false);
  const bf16x2 l = __builtin_convertvector((f32x2v){ra, rb}, bf16x2);
  hi = __builtin_bit_cast(unsigned, h);
  lo = __builtin_bit_cast(unsigned, l);
}
// the two dot2 operands, opaque to the compiler (kept in scalar registers)
struct SplitK {
  unsigned k0, k1;
};
__device__ __forceinline__ SplitK split_consts() {
  SplitK k{0x0000bf80u, 0xbf800000u};  // bf16 pairs (-1, 0), (0, -1)
  asm volatile("" : "+s"(k.k0), "+s"(k.k1));
  return k;
}
__device__ __forceinline__ void split2(float a, float b, unsigned& hi, unsigned& lo) {
  const SplitK k = split_consts();
  split2k(a, b, hi, lo, k.k0, k.k1);
}
__device__ __forceinline__ void split2(float a, float b, unsigned& hi, unsigned& lo, SplitK k) {
  split2k(a, b, hi, lo, k.k0, k.k1);
}

__device__ __forceinline__ bf16x8 ld8(const __bf16* p) {
  return *reinterpret_cast<const bf16x8*>(p);
}

__device__ __forceinline__ void split_store(__bf16* hi, __bf16* lo, int idx, double v) {
  const float f = (float)v;
  const __bf16 h = (__bf16)f;
  hi[idx] = h;
  lo[idx] = (__bf16)(f - (float)h);
}

// an opaque copy of a wave-uniform value (kept in a scalar register)
__device__ __forceinline__ int fresh_s(int x) {
  asm volatile("" : "+s"(x));
  return x;
}

__device__ __forceinline__ void wave_lds_fence() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

template <int PIX>
__global__ void __launch_bounds__(R2Cfg<PIX>::kThreads, 2) k_track_r2(TrackArgs a) {
  using C = R2Cfg<PIX>;
  constexpr int RS = C::RS, PT = C::kTabPlane, PL = C::kPixPlane;
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int ps = wave >> 1, kt = wave & 1;
  const int ptid = tid - ps * 128;
  const int r0l = lane & 15, g0 = lane >> 4;
  const int N0 = 2 * a.nsize, Nh0 = a.nsize;

  __bf16* TA0 = reinterpret_cast<__bf16*>(lds);
  __bf16* WT0 = TA0 + C::kTabEl;
  float2* tw0 = reinterpret_cast<float2*>(WT0 + C::kTabEl);
  unsigned char* pb = reinterpret_cast<unsigned char*>(tw0 + 32) + ps * C::kPixStride;
  __bf16* GB0 = reinterpret_cast<__bf16*>(pb);
  float* wxs = reinterpret_cast<float*>(pb + C::kPixEl * 2);    // [2 waves][32] row windows
  float2* zx = reinterpret_cast<float2*>(wxs + 64);             // [2][64]
  float2* zp = zx + 128;                                        // [64]
  double* nb = reinterpret_cast<double*>(zp + 64);
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(nb + 9);
  unsigned int* asec = reinterpret_cast<unsigned int*>(amax + 1);
  double* res = reinterpret_cast<double*>(amax + 2);

  // ---------------- constant tables
  for (int e = tid; e < 32 * 32; e += blockDim.x) {
    const int rho = e >> 5, kap = e & 31;
    // TA[rho][kappa] = D_h[sig(rho)][pi1(kappa)]
    {
      const int sr = r2_sig(rho), pk = r2_pi1(kap);
      double dr = 0., di = 0.;
      if (sr < Nh0 && pk < Nh0) {
        const int m = (sr * pk) % Nh0;
        double sn, cs;
        sincospi(2.0 * (double)m / (double)Nh0, &sn, &cs);
        dr = cs;
        di = -sn;
      }
      split_store(TA0, TA0 + PT, rho * RS + kap, dr);
      split_store(TA0 + 2 * PT, TA0 + 3 * PT, rho * RS + kap, di);
    }
    // WT[set][ii][kappa]: kappa = 2m' + part, k = 2m' + set; weights of
    // Re U (part 0) and Im U (part 1) in c[ii] = sum_k a_k Re(e^{2 pi i k ii/N0} U[k])
    {
      const int ii = rho, part = kap & 1;
      static_for<2>([&](auto s) {
        const int k = (kap & ~1) + s;
        double v = 0.;
        if (ii < Nh0) {
          if (s == 0 && k == 0) {
            v = (part == 0) ? 1. : ((ii & 1) ? -1. : 1.);  // DC, and Nyquist packed in Im
          } else if (k < Nh0) {
            const int m = (k * ii) % N0;
            double sn, cs;
            sincospi(2.0 * (double)m / (double)N0, &sn, &cs);
            v = (part == 0) ? 2. * cs : -2. * sn;
          }
        }
        split_store(WT0 + (2 * s) * PT, WT0 + (2 * s + 1) * PT, ii * RS + kap, v);
      });
    }
  }
  for (int l = tid; l < 32; l += blockDim.x) {
    float2 w = make_float2(0.f, 0.f);
    if (l < Nh0) {
      double sn, cs;
      sincospi(2.0 * (double)l / (double)N0, &sn, &cs);
      w = make_float2((float)cs, (float)-sn);
    }
    tw0[l] = w;
  }
  // stage-1 B operand in registers: D_h[x' = 8g + j][k = 16 kt + r]
  CFrag bs1;
  const int kcol = 16 * kt + r0l;
  {
    u32x4 rh, rl, ih, il;
    static_for<4>([&](auto d) {
      float v[2][2];
      static_for<2>([&](auto hf) {
        const int xp = 8 * g0 + 2 * d + hf;
        double dr = 0., di = 0.;
        if (xp < Nh0 && kcol < Nh0) {
          const int m = (xp * kcol) % Nh0;
          double sn, cs;
          sincospi(2.0 * (double)m / (double)Nh0, &sn, &cs);
          dr = cs;
          di = -sn;
        }
        v[0][hf] = (float)dr;
        v[1][hf] = (float)di;
      });
      unsigned h0, l0, h1, l1;
      split2(v[0][0], v[0][1], h0, l0);
      split2(v[1][0], v[1][1], h1, l1);
      rh[int(d)] = h0;
      rl[int(d)] = l0;
      ih[int(d)] = h1;
      il[int(d)] = l1;
    });
    bs1.re.h = __builtin_bit_cast(bf16x8, rh);
    bs1.re.l = __builtin_bit_cast(bf16x8, rl);
    bs1.im.h = __builtin_bit_cast(bf16x8, ih);
    bs1.im.l = __builtin_bit_cast(bf16x8, il);
  }
  float w1r = 1.f, w1i = 0.f;  // w^k for the stage-1 combine
  if (kcol < Nh0) {
    double sn, cs;
    sincospi(2.0 * (double)kcol / (double)N0, &sn, &cs);
    w1r = (float)cs;
    w1i = (float)-sn;
  }
  const bool k0 = (kcol == 0);
  __syncthreads();

  // image scales: powers of two bringing max |f| to [1, 2)
  double sc[2] = {1., 1.};
  if (a.absmax) {
    static_for<2>([&](auto X) {
      const double m = __longlong_as_double((long long)a.absmax[X]);
      if (m > 0. && m < 1e300) sc[X] = ldexp(1., -ilogb(m));
    });
  }

  const int nx = a.nx, ny = a.ny, nsize = a.nsize;
  const int ns = ny - N0 + 1;
  const float normf = (float)((1. / ((double)N0 * N0)) * (1. / ((double)N0 * N0)));
  const int npass = a.bias ? 2 : 1;

  // work distribution (as the plain form): XCD bands of rows, strips of
  // kStrip columns; list mode strides through the list
  constexpr int kXcd = 8, kStrip = 512;
  const bool dense = (a.list == nullptr);
  // 32-bit work indices: the host limits images to fewer than 2^31 pixels
  unsigned first, stride, count;
  int r0 = 0, rows_b = 0;
  if (dense) {
    const int nband = (int)gridDim.x < kXcd ? (int)gridDim.x : kXcd;
    const int xb = blockIdx.x % nband;
    const unsigned nbx = (gridDim.x - xb + nband - 1) / nband;
    r0 = (int)((long long)nx * xb / nband);
    rows_b = (int)((long long)nx * (xb + 1) / nband) - r0;
    first = (blockIdx.x / nband) * PIX + ps;
    stride = nbx * PIX;
    count = (unsigned)rows_b * (unsigned)ny;
  } else {
    first = blockIdx.x * PIX + ps;
    stride = gridDim.x * PIX;
    count = a.nwork_dev ? *a.nwork_dev : (unsigned)a.nwork;
  }
  const unsigned iters = (count + stride - 1) / stride;
  const int full_strips = ny / kStrip;
  const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};

  for (unsigned it = 0; it < iters; ++it) {
    const unsigned w = first + it * stride;
    const bool active = w < count;
    int i = 0, j = 0, imin = 0, jmin = 0;
    double f1bar = 0., f2bar = 0.;
    if (active) {
      if (dense) {
        const unsigned sw = (unsigned)rows_b * kStrip;
        const unsigned sidx = w / sw;
        unsigned width = kStrip;
        unsigned t = w - sidx * sw;
        if ((int)sidx >= full_strips) {
          width = (unsigned)(ny - full_strips * kStrip);
          t = w - (unsigned)full_strips * sw;
        }
        const int strip = (int)sidx < full_strips ? (int)sidx : full_strips;
        const unsigned row = t / width;
        i = r0 + (int)row;
        j = strip * kStrip + (int)(t - row * width);
      } else {
        const int p = a.list[w];
        i = p / ny;
        j = p - i * ny;
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
    bool flagged = false;  // meaningful in thread ptid == 0
    for (int pass = 0; pass < npass; ++pass) {
      // per-pass copies of the LDS bases and lane indices: keeps the
      // compiler from hoisting table loads and index math out of the loop
      const int zo = fresh(0);
      const int g = fresh(g0), r = fresh(r0l);
      const int N = fresh_s(N0), Nh = fresh_s(Nh0);
      const __bf16* TA = TA0 + zo;
      const __bf16* WT = WT0 + zo;
      const float2* tw = tw0 + zo;
      __bf16* GB = GB0 + zo;
      // ---------------- windows: lane q holds w(x = q) and w(c = q)
      float wxl[2] = {0.f, 0.f}, wyl[2] = {0.f, 0.f};
      if (active && lane < N) {
        const int dx = lane + imin - i, dy = lane + jmin - j;
        double x1, x2, y1, y2;
        if (pass == 0) {
          x1 = x2 = a.gw[dx + N - 1];
          y1 = y2 = a.gw[dy + N - 1];
        } else {
          const double cx = 0.5 * shiftx, cy = 0.5 * shifty;
          const double a1 = a.sigminv * ((double)dx + cx), a2 = a.sigminv * ((double)dx - cx);
          const double b1 = a.sigminv * ((double)dy + cy), b2 = a.sigminv * ((double)dy - cy);
          x1 = exp(-a1 * a1);
          x2 = exp(-a2 * a2);
          y1 = exp(-b1 * b1);
          y2 = exp(-b2 * b2);
        }
        wxl[0] = (float)x1;
        wxl[1] = (float)x2;
        wyl[0] = (float)(y1 * sc[0]);
        wyl[1] = (float)(y2 * sc[1]);
      }
      if (ptid == 0) {
        *amax = 0ull;
        *asec = 0u;
      }

      CF4 S1p[2], S1m[2], S2p[2], S2m[2];
      // ---------------- G: wave kt covers the rows x = 2x' + kt, lane = column c.
      // The raw rows of image 2 are loaded while image 1 is in stage 1.
      double fpre[32];
      auto load_rows = [&](auto X) {
        if (active) {
          const double* f = (X == 0) ? a.f1 : a.f2;
          // rows are wave-uniform: one scalar base plus a 32-bit byte offset
          // that advances by two rows per element (columns clamped into the
          // box, rows beyond it read row 0; padding is zeroed when built)
          const char* fb = reinterpret_cast<const char*>(f + ((long long)imin * ny + jmin));
          const int cl = lane < N ? lane : N - 1;
          // opaque per pixel, so the per-row offsets are not hoisted out of
          // the pixel loop as 32 live registers
          const unsigned step = (unsigned)fresh_s(16 * ny);
          unsigned boff = (unsigned)fresh(8 * (kt * ny + cl));
          static_for<32>([&](auto e) {
            const int x = 2 * e + kt;
            const unsigned bo = (x < N) ? boff : 8u * (unsigned)cl;
            fpre[int(e)] = *reinterpret_cast<const double*>(fb + bo);
            boff += step;
          });
        }
      };
      auto build_g = [&](auto X) {
        if (active) {
          const double fbar = (X == 0) ? f1bar : f2bar;
          // this wave's row windows w(x), x = 16 it4 + 2t + kt, as 4 x 8
          // floats in LDS (written and read by this wave only)
          float* wxt = wxs + kt * 32;
          if ((lane & 1) == kt) wxt[lane >> 1] = wxl[X];
          wave_lds_fence();
          const bool cval = lane < N;
          const float wy = wyl[X];
          static_for<4>([&](auto it4) {
            const float4 wa = *reinterpret_cast<const float4*>(wxt + 8 * it4);
            const float4 wb = *reinterpret_cast<const float4*>(wxt + 8 * it4 + 4);
            const float wxv[8] = {wa.x, wa.y, wa.z, wa.w, wb.x, wb.y, wb.z, wb.w};
            float v[8];
            static_for<8>([&](auto t) {
              const int x = 16 * it4 + 2 * t + kt;
              const float gv = (float)(fpre[8 * it4 + t] - fbar) * (wy * wxv[int(t)]);
              v[int(t)] = (x < N && cval) ? gv : 0.f;
            });
            u32x4 hv, lv;
            static_for<4>([&](auto d) {
              unsigned hh, ll;
              split2(v[2 * d], v[2 * d + 1], hh, ll);
              hv[int(d)] = hh;
              lv[int(d)] = ll;
            });
            *reinterpret_cast<u32x4*>(GB + (2 * kt) * PL + lane * RS + 8 * it4) = hv;
            *reinterpret_cast<u32x4*>(GB + (2 * kt + 1) * PL + lane * RS + 8 * it4) = lv;
          });
        }
      };
      // stages 1 and 2 at raised wave priority, so that the blocks sharing a
      // CU drift out of phase and one's MFMAs overlap another's VALU work
      // (as in k_track_seg)
      __builtin_amdgcn_s_setprio(1);
      static_for<2>([&](auto X) {
        if constexpr (X == 0) {
          load_rows(X);
          build_g(X);
        }
        __syncthreads();  // G of image X is complete
        if constexpr (X == 0) load_rows(std::integral_constant<int, 1>{});
        // ---------------- stage 1: T[c][k] for c = 0..63, this wave's k
        CF4 T[4];
        static_for<4>([&](auto MT) {
          const int row = (16 * MT + r) * RS + 8 * g;
          const bf16x8 eh = ld8(GB + 0 * PL + row), el = ld8(GB + 1 * PL + row);
          const bf16x8 oh = ld8(GB + 2 * PL + row), ol = ld8(GB + 3 * PL + row);
          const f32x4 Er = mfma3(eh, el, bs1.re.h, bs1.re.l, z4);
          const f32x4 Ei = mfma3(eh, el, bs1.im.h, bs1.im.l, z4);
          const f32x4 Or = mfma3(oh, ol, bs1.re.h, bs1.re.l, z4);
          const f32x4 Oi = mfma3(oh, ol, bs1.im.h, bs1.im.l, z4);
          static_for<4>([&](auto t) {
            const float tr = Er[int(t)] + (w1r * Or[int(t)] - w1i * Oi[int(t)]);
            const float ti = Ei[int(t)] + (w1r * Oi[int(t)] + w1i * Or[int(t)]);
            T[MT].re[int(t)] = tr;
            T[MT].im[int(t)] = k0 ? (Er[int(t)] - Or[int(t)]) : ti;
          });
        });
        // both waves have read G of image 1; build image 2 (its rows were
        // loaded during stage 1) while image 1 goes on to stage 2
        if constexpr (X == 0) {
          __syncthreads();
          build_g(std::integral_constant<int, 1>{});
        }
        // B operands of stage 2: set q holds the rows c = 16 MT + 4g + q (+2)
        CFrag B[2];
        static_for<2>([&](auto q) {
          u32x4 rh, rl, ih, il;
          static_for<4>([&](auto d) {
            unsigned h0, l0, h1, l1;
            split2(T[d].re[int(q)], T[d].re[int(q) + 2], h0, l0);
            split2(T[d].im[int(q)], T[d].im[int(q) + 2], h1, l1);
            rh[int(d)] = h0;
            rl[int(d)] = l0;
            ih[int(d)] = h1;
            il[int(d)] = l1;
          });
          B[q].re.h = __builtin_bit_cast(bf16x8, rh);
          B[q].re.l = __builtin_bit_cast(bf16x8, rl);
          B[q].im.h = __builtin_bit_cast(bf16x8, ih);
          B[q].im.l = __builtin_bit_cast(bf16x8, il);
        });
        // ---------------- stage 2: S[l][k], l = sig(16 LT + 4g + t) (+ Nh)
        static_for<2>([&](auto LT) {
          const int row = (16 * LT + r) * RS + 8 * g;
          const bf16x8 rh = ld8(TA + 0 * PT + row), rl = ld8(TA + 1 * PT + row);
          const bf16x8 ih = ld8(TA + 2 * PT + row), il = ld8(TA + 3 * PT + row);
          const bf16x8 nih = neg8(ih), nil = neg8(il);
          f32x4 Er = mfma3(rh, rl, B[0].re.h, B[0].re.l, z4);
          Er = mfma3(nih, nil, B[0].im.h, B[0].im.l, Er);
          f32x4 Ei = mfma3(rh, rl, B[0].im.h, B[0].im.l, z4);
          Ei = mfma3(ih, il, B[0].re.h, B[0].re.l, Ei);
          f32x4 Or = mfma3(rh, rl, B[1].re.h, B[1].re.l, z4);
          Or = mfma3(nih, nil, B[1].im.h, B[1].im.l, Or);
          f32x4 Oi = mfma3(rh, rl, B[1].im.h, B[1].im.l, z4);
          Oi = mfma3(ih, il, B[1].re.h, B[1].re.l, Oi);
          CF4 sp, sm;
          static_for<4>([&](auto t) {
            const float2 wl = tw[r2_sig(16 * LT + 4 * g + t)];
            const float wor = wl.x * Or[int(t)] - wl.y * Oi[int(t)];
            const float woi = wl.x * Oi[int(t)] + wl.y * Or[int(t)];
            sp.re[int(t)] = Er[int(t)] + wor;
            sp.im[int(t)] = Ei[int(t)] + woi;
            sm.re[int(t)] = Er[int(t)] - wor;
            sm.im[int(t)] = Ei[int(t)] - woi;
          });
          if constexpr (X == 0) {
            S1p[LT] = sp;
            S1m[LT] = sm;
          } else {
            S2p[LT] = sp;
            S2m[LT] = sm;
          }
        });
      });

      __builtin_amdgcn_s_setprio(0);
      // ---------------- cross power
      CF4 Pp[2], Pm[2];
      static_for<2>([&](auto LT) {
        static_for<4>([&](auto t) {
          const int l = r2_sig(16 * LT + 4 * g + t);
          float nfp = normf, nfm = normf;
          if (a.filt && active && l < Nh && kcol < Nh) {
            const float fp = (float)a.filt[kcol * N + l], fm = (float)a.filt[kcol * N + l + Nh];
            nfp *= fp * fp;
            nfm *= fm * fm;
          }
          {
            const float ar = S1p[LT].re[int(t)], ai = S1p[LT].im[int(t)];
            const float br = S2p[LT].re[int(t)], bi = S2p[LT].im[int(t)];
            Pp[LT].re[int(t)] = (ar * br + ai * bi) * nfp;
            Pp[LT].im[int(t)] = (ar * bi - ai * br) * nfp;
          }
          {
            const float ar = S1m[LT].re[int(t)], ai = S1m[LT].im[int(t)];
            const float br = S2m[LT].re[int(t)], bi = S2m[LT].im[int(t)];
            Pm[LT].re[int(t)] = (ar * br + ai * bi) * nfm;
            Pm[LT].im[int(t)] = (ar * bi - ai * br) * nfm;
          }
        });
      });
      // packed DC / Nyquist column (k = 0): unpack with the partner -l
      if (kt == 0) {
        if (r == 0) {
          static_for<2>([&](auto LT) {
            static_for<4>([&](auto t) {
              const int l = r2_sig(16 * LT + 4 * g + t);
              if (l < Nh) {
                zx[l] = make_float2(S1p[LT].re[int(t)], S1p[LT].im[int(t)]);
                zx[l + Nh] = make_float2(S1m[LT].re[int(t)], S1m[LT].im[int(t)]);
                zx[64 + l] = make_float2(S2p[LT].re[int(t)], S2p[LT].im[int(t)]);
                zx[64 + l + Nh] = make_float2(S2m[LT].re[int(t)], S2m[LT].im[int(t)]);
              }
            });
          });
        }
        wave_lds_fence();
        if (lane < N) {
          const int lp = (lane == 0) ? 0 : N - lane;
          const float2 z1 = zx[lane], q1 = zx[lp], z2 = zx[64 + lane], q2 = zx[64 + lp];
          // A0 = (z + conj zp)/2, A32 = (z - conj zp)/(2i)
          const float a0r1 = 0.5f * (z1.x + q1.x), a0i1 = 0.5f * (z1.y - q1.y);
          const float anr1 = 0.5f * (z1.y + q1.y), ani1 = -0.5f * (z1.x - q1.x);
          const float a0r2 = 0.5f * (z2.x + q2.x), a0i2 = 0.5f * (z2.y - q2.y);
          const float anr2 = 0.5f * (z2.y + q2.y), ani2 = -0.5f * (z2.x - q2.x);
          float n0 = normf, nn = normf;
          if (a.filt && active) {
            const float f0 = (float)a.filt[lane], fn = (float)a.filt[Nh * N + lane];
            n0 *= f0 * f0;
            nn *= fn * fn;
          }
          const float p0r = (a0r1 * a0r2 + a0i1 * a0i2) * n0;
          const float p0i = (a0r1 * a0i2 - a0i1 * a0r2) * n0;
          const float pnr = (anr1 * anr2 + ani1 * ani2) * nn;
          const float pni = (anr1 * ani2 - ani1 * anr2) * nn;
          zp[lane] = make_float2(p0r - pni, p0i + pnr);
        }
        wave_lds_fence();
        if (r == 0) {
          static_for<2>([&](auto LT) {
            static_for<4>([&](auto t) {
              const int l = r2_sig(16 * LT + 4 * g + t);
              if (l < Nh) {
                const float2 pp = zp[l], pm = zp[l + Nh];
                Pp[LT].re[int(t)] = pp.x;
                Pp[LT].im[int(t)] = pp.y;
                Pm[LT].re[int(t)] = pm.x;
                Pm[LT].im[int(t)] = pm.y;
              }
            });
          });
        }
      }

      // ---------------- inverse 1 (DIF on l): Q+ = P+ + P-, Q- = (P+ - P-) conj(w^l)
      CFrag Q[2];
      {
        u32x4 qh[2][2], ql[2][2];  // [q][re/im]
        static_for<2>([&](auto LT) {
          float qv[2][2][4];  // [q][re/im][t]
          static_for<4>([&](auto t) {
            const float2 wl = tw[r2_sig(16 * LT + 4 * g + t)];
            const float pr = Pp[LT].re[int(t)], pi = Pp[LT].im[int(t)];
            const float mr = Pm[LT].re[int(t)], mi = Pm[LT].im[int(t)];
            qv[0][0][int(t)] = pr + mr;
            qv[0][1][int(t)] = pi + mi;
            const float dr = pr - mr, di = pi - mi;
            qv[1][0][int(t)] = dr * wl.x + di * wl.y;
            qv[1][1][int(t)] = di * wl.x - dr * wl.y;
          });
          static_for<2>([&](auto q) {
            static_for<2>([&](auto c) {
              static_for<2>([&](auto hp) {
                unsigned hh, ll;
                split2(qv[q][c][2 * hp], qv[q][c][2 * hp + 1], hh, ll);
                qh[q][c][2 * LT + hp] = hh;
                ql[q][c][2 * LT + hp] = ll;
              });
            });
          });
        });
        static_for<2>([&](auto q) {
          Q[q].re.h = __builtin_bit_cast(bf16x8, qh[q][0]);
          Q[q].re.l = __builtin_bit_cast(bf16x8, ql[q][0]);
          Q[q].im.h = __builtin_bit_cast(bf16x8, qh[q][1]);
          Q[q].im.l = __builtin_bit_cast(bf16x8, ql[q][1]);
        });
      }
      CF4 U[2][2];  // [MT][q]
      static_for<2>([&](auto MT) {
        const int row = (16 * MT + r) * RS + 8 * g;
        const bf16x8 rh = ld8(TA + 0 * PT + row), rl = ld8(TA + 1 * PT + row);
        const bf16x8 ih = ld8(TA + 2 * PT + row), il = ld8(TA + 3 * PT + row);
        const bf16x8 nih = neg8(ih), nil = neg8(il);
        static_for<2>([&](auto q) {
          f32x4 ur = mfma3(rh, rl, Q[q].re.h, Q[q].re.l, z4);
          ur = mfma3(ih, il, Q[q].im.h, Q[q].im.l, ur);
          f32x4 ui = mfma3(rh, rl, Q[q].im.h, Q[q].im.l, z4);
          ui = mfma3(nih, nil, Q[q].re.h, Q[q].re.l, ui);
          U[MT][q].re = ur;
          U[MT][q].im = ui;
        });
      });
      __syncthreads();  // all G reads are done; U reuses the buffer
      {
        const int s = kcol & 1, kf = kcol & ~1;
        unsigned* uh = reinterpret_cast<unsigned*>(GB + (2 * s) * PL);
        unsigned* ul = reinterpret_cast<unsigned*>(GB + (2 * s + 1) * PL);
        static_for<2>([&](auto MT) {
          static_for<2>([&](auto q) {
            static_for<4>([&](auto t) {
              const int jj = 2 * r2_sig(16 * MT + 4 * g + t) + q;
              unsigned hh, ll;
              split2(U[MT][q].re[int(t)], U[MT][q].im[int(t)], hh, ll);
              uh[(jj * RS + kf) >> 1] = hh;
              ul[(jj * RS + kf) >> 1] = ll;
            });
          });
        });
      }
      __syncthreads();
      // ---------------- inverse 2: c^T[jj][ii] over the even and odd k sets
      f32x4 cp[2][2], cm[2][2];  // [jl][IT]
      {
        Frag2 wv[2][2];  // [IT][set]
        static_for<2>([&](auto IT) {
          const int row = (16 * IT + r) * RS + 8 * g;
          static_for<2>([&](auto s) {
            wv[IT][s].h = ld8(WT + (2 * s) * PT + row);
            wv[IT][s].l = ld8(WT + (2 * s + 1) * PT + row);
          });
        });
        static_for<2>([&](auto jl) {
          const int row = (16 * (2 * kt + jl) + r) * RS + 8 * g;
          const bf16x8 eh = ld8(GB + 0 * PL + row), el = ld8(GB + 1 * PL + row);
          const bf16x8 oh = ld8(GB + 2 * PL + row), ol = ld8(GB + 3 * PL + row);
          static_for<2>([&](auto IT) {
            const f32x4 ce = mfma3(eh, el, wv[IT][0].h, wv[IT][0].l, z4);
            const f32x4 co = mfma3(oh, ol, wv[IT][1].h, wv[IT][1].l, z4);
            cp[jl][IT] = ce + co;
            cm[jl][IT] = ce - co;
          });
        });
      }
      // ---------------- |c|, argmax with the first-strict-maximum rule, fit
      // element t of (jl, IT): jj = 16(2kt + jl) + 4g + t, ii = 16 IT + r
      // (cp) and ii + Nh (cm).  In fftshift order (flct.c:1490-1510) the
      // cell (ii, jj) sits at row ii + Nh (cp) or ii (cm) and column
      // jj + Nh (mod N); the scan index is row * N + column.  A key packs
      // |c| above the complemented scan index, so the largest key is the
      // first strict maximum.
      unsigned long long k1 = 0ull;  // best key of this lane
      float m2 = 0.f;                // largest other value of this lane
      if (active) {
        static_for<2>([&](auto jl) {
          static_for<4>([&](auto t) {
            const int jj = 16 * (2 * kt + jl) + 4 * g + t;
            const int sy = (jj < Nh) ? jj + Nh : jj - Nh;
            const bool jv = jj < N;
            static_for<2>([&](auto IT) {
              const int ii = 16 * IT + r;
              const bool v = jv && ii < Nh;
              static_for<2>([&](auto pm) {
                const float c = fabsf(pm == 0 ? cp[jl][IT][int(t)] : cm[jl][IT][int(t)]);
                const int sx = (pm == 0) ? ii + Nh : ii;
                const unsigned L = (unsigned)(sx * N + sy);
                const unsigned long long kk =
                    v ? (((unsigned long long)__float_as_uint(c) << 32) | (unsigned long long)(~L)) : 0ull;
                const bool nbest = kk > k1;
                const float prev = __uint_as_float((unsigned)(k1 >> 32));
                m2 = fmaxf(m2, nbest ? prev : (v ? c : 0.f));
                k1 = nbest ? kk : k1;
              });
            });
          });
        });
        unsigned long long kw = k1;
        for (int o = 32; o > 0; o >>= 1) {
          const unsigned long long ko = __shfl_xor(kw, o);
          kw = ko > kw ? ko : kw;
        }
        if (lane == 0) atomicMax(amax, kw);
      }
      __syncthreads();
      const unsigned long long kwin = *amax;
      const double gmax = (double)__uint_as_float((unsigned)(kwin >> 32));
      const unsigned loc = ~(unsigned)kwin;
      const bool nonzero = (gmax != 0.);
      int ixmax = Nh, iymax = Nh;
      if (nonzero) {
        ixmax = (int)(loc / (unsigned)N);
        iymax = (int)(loc - (unsigned)ixmax * (unsigned)N);
      }
      const bool fit = nonzero && ixmax > 0 && ixmax < N - 1 && iymax > 0 && iymax < N - 1;
      if (active) {
        // largest value outside the winning cell (for the near-tie test)
        float sec = (k1 == kwin) ? m2 : __uint_as_float((unsigned)(k1 >> 32));
        for (int o = 32; o > 0; o >>= 1) sec = fmaxf(sec, __shfl_xor(sec, o));
        if (lane == 0) atomicMax(asec, __float_as_uint(sec));
        if (fit) {
          static_for<2>([&](auto IT) {
            static_for<2>([&](auto pm) {
              const int ii = 16 * IT + r;
              const int dx = ((pm == 0) ? ii + Nh : ii) - ixmax;
              if (ii < Nh && dx >= -1 && dx <= 1) {
                static_for<2>([&](auto jl) {
                  static_for<4>([&](auto t) {
                    const int jj = 16 * (2 * kt + jl) + 4 * g + t;
                    const int dy = ((jj < Nh) ? jj + Nh : jj - Nh) - iymax;
                    if (jj < N && dy >= -1 && dy <= 1)
                      nb[(dx + 1) * 3 + (dy + 1)] =
                          (double)fabsf(pm == 0 ? cp[jl][IT][int(t)] : cm[jl][IT][int(t)]);
                  });
                });
              }
            });
          });
        }
      }
      __syncthreads();
      if (active && ptid == 0) {
#pragma clang fp contract(off)
        double shiftsubx = 0., shiftsuby = 0.;
        if (fit) {
          const double c00 = nb[4];
          const double fx = 0.5 * (nb[7] - nb[1]);
          const double fy = 0.5 * (nb[5] - nb[3]);
          const double fxx = (nb[7] + nb[1] - 2. * (c00));
          const double fyy = (nb[5] + nb[3] - 2. * (c00));
          const double fxy = 0.25 * (nb[8] + nb[0] - nb[6] - nb[2]);
          shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
          shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
          // first-order bound of the fit's response to errors of
          // kRelErr * gmax in the nine values
          const double det = fxx * fyy - fxy * fxy;
          const double hn = fabs(fxx) + fabs(fyy) + 2. * fabs(fxy);
          const double bound = kRelErr * gmax * hn * (1. + 4. * (fabs(shiftsubx) + fabs(shiftsuby)));
          if (!(bound <= kFitTol * fabs(det))) flagged = true;
        }
        // another cell within the split-BF16 error of the maximum: the
        // peak cell itself is uncertain
        if (nonzero && (double)__uint_as_float(*asec) >= gmax * (1. - kTieTol)) flagged = true;
        res[0] = ((double)ixmax + shiftsubx) - (double)Nh;
        res[1] = ((double)iymax + shiftsuby) - (double)Nh;
      }
      __syncthreads();
      shiftx = res[0];
      shifty = res[1];
    }
    if (active && ptid == 0) {
      const long long p = (long long)i * ny + j;
      a.vx[p] = shifty * a.deltinv * a.deltas;
      a.vy[p] = shiftx * a.deltinv * a.deltas;
      a.vm[p] = 1.;
      if (flagged && a.flag_list) a.flag_list[atomicAdd(a.flag_count, 1u)] = (int)p;
    }
  }
}

}  // namespace flct
