// Split-BF16 MFMA form of the tracker for gfx950.
//
// Each sub-image DFT is a pair of dense DFT-matrix contractions on the
// matrix cores: T = D G along the box rows, then S = T D along the columns,
// with D[r][c] = exp(-2 pi i r c / N) (symmetric).  Operands are split into
// BF16 hi + lo parts and each real product is accumulated in FP32 as
// hi*hi + hi*lo + lo*hi (v_mfma_f32_16x16x32_bf16).  Only the half spectrum
// k = 0..N/2 is formed (the images are real).  The cross power
// conj(S1) S2 normfac (flct.c:1461-1477, optional filter^2, flct.c:1445-1457)
// is elementwise in registers.  The inverse is the Hermitian sum
//     c(ii,jj) = sum_k a_k Re(exp(+2 pi i k ii/N) U(k,jj)),
//     U = P conj(D),  a_0 = a_{N/2} = 1, else 2,
// again as two contractions.  Window, box means (FP64, shared with the FFT
// form), the argmax and the 3x3 Taylor fit (FP64, flct.c:1512-1556) are as
// in the plain form.
//
// Every contraction is oriented so that the FP32 C-tile of one stage (four
// consecutive rows per lane) is stored as 8-byte BF16 runs that are exactly
// the 16-byte k-runs the next stage reads as its A or B fragment; D being
// symmetric, one row-major copy of D in LDS serves as either operand.
//
// Block: PIX pixels, two waves per pixel; the waves of a pixel split every
// contraction's output tiles in halves.
#pragma once

#include "track_kernels.hpp"

namespace flct {

// Pixels whose result the split-BF16 surface cannot settle are listed for
// re-evaluation in FP64: another cell within kTieTol of the maximum, or a
// Taylor fit whose first-order response to relative errors kRelErr of the
// nine values exceeds kFitTol pixels.  kRelErr is a few times the observed
// relative error of the surface values.
constexpr double kTieTol = 1e-4;
constexpr double kRelErr = 1e-5;
constexpr double kFitTol = 1e-2;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

template <int NP>
struct MfmaCfg {
  static constexpr int TP = NP / 16;                   // 16-wide tiles over NP
  static constexpr int KSN = NP / 32;                  // 32-deep k steps over NP
  static constexpr int KH = ((NP / 2 + 1 + 15) / 16) * 16;  // half-spectrum rows
  static constexpr int KHT = KH / 16;
  static constexpr int KP = ((KH + 31) / 32) * 32;     // k contraction length of the last stage
  static constexpr int KSP = KP / 32;
  // LDS row strides (bf16 elements): an odd multiple of 8, so the 16 rows
  // of a fragment read start in different 16-byte bank groups
  static constexpr int RS = NP + 8;                    // rows of NP columns
  static constexpr int RSU = KP + 8;                   // rows of KP columns (U^T)
  static constexpr int PIX = (NP <= 64) ? 2 : 1;       // pixels per block
  static constexpr int kThreads = PIX * 128;
  static constexpr int cmax(int x, int y) { return x > y ? x : y; }
  // per-pixel bf16 work buffer, reused by the stages: G^T (2 x NP x RS),
  // T (4 x KH x RS), P (4 x KH x RS), U^T (4 x NP x RSU)
  static constexpr int kBufEl = cmax(cmax(2 * NP * RS, 4 * KH * RS), 4 * NP * RSU);
  static constexpr int kDEl = 4 * NP * RS;             // D: r_hi, r_lo, i_hi, i_lo
  static constexpr int kPixBytes = ((kBufEl * 2 + 4 * NP * 8 + 9 * 8 + 64) + 15) / 16 * 16;
  static constexpr int kLdsBytes = kDEl * 2 + PIX * kPixBytes + (2 * NP) * 8;
};

__device__ __forceinline__ f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hi*hi + hi*lo + lo*hi, small terms first
__device__ __forceinline__ f32x4 mfma3(bf16x8 ah, bf16x8 al, bf16x8 bh, bf16x8 bl, f32x4 c) {
  c = mfma16(al, bh, c);
  c = mfma16(ah, bl, c);
  return mfma16(ah, bh, c);
}

__device__ __forceinline__ bf16x8 neg8(bf16x8 v) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  u32x4 u = __builtin_bit_cast(u32x4, v);
  u ^= 0x80008000u;
  return __builtin_bit_cast(bf16x8, u);
}

// 8 consecutive bf16 at base[row * stride + k0 + 8 * (lane / 16)]
__device__ __forceinline__ bf16x8 ldfrag(const __bf16* base, int row, int k0, int stride,
                                         int fg) {
  return *reinterpret_cast<const bf16x8*>(base + row * stride + k0 + 8 * fg);
}

__device__ __forceinline__ void split_bf16(float x, __bf16& hi, __bf16& lo) {
  hi = (__bf16)x;
  lo = (__bf16)(x - (float)hi);
}

// store 4 consecutive floats as hi and lo bf16 runs
__device__ __forceinline__ void st4(__bf16* hi, __bf16* lo, f32x4 v) {
  bf16x4 h, o;
  static_for<4>([&](auto t) {
    __bf16 a, b;
    split_bf16(v[int(t)], a, b);
    h[int(t)] = a;
    o[int(t)] = b;
  });
  *reinterpret_cast<bf16x4*>(hi) = h;
  *reinterpret_cast<bf16x4*>(lo) = o;
}

template <int NP>
__global__ void __launch_bounds__((MfmaCfg<NP>::kThreads)) k_track_mfma(TrackArgs a) {
  using C = MfmaCfg<NP>;
  constexpr int TP = C::TP, KHT = C::KHT, RS = C::RS;
  constexpr int HT = TP / 2;  // tiles per wave along a split dimension
  extern __shared__ __align__(16) unsigned char lds[];
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int ps = wave >> 1, h = wave & 1;  // pixel slot, half
  const int ptid = tid - ps * 128;         // thread index within the pixel
  const int fr = lane & 15, fg = lane >> 4;
  const int N = 2 * a.nsize;

  __bf16* Dsh = reinterpret_cast<__bf16*>(lds);
  double* gw = reinterpret_cast<double*>(lds + C::kDEl * 2);
  unsigned char* pbase = lds + C::kDEl * 2 + (2 * NP) * 8 + ps * C::kPixBytes;
  __bf16* Bf = reinterpret_cast<__bf16*>(pbase);
  double* wx1 = reinterpret_cast<double*>(pbase + C::kBufEl * 2);
  double* wy1 = wx1 + NP;
  double* wx2 = wy1 + NP;
  double* wy2 = wx2 + NP;
  double* nb = wy2 + NP;
  double* res = nb + 9;
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(res + 2);
  unsigned int* aidx = reinterpret_cast<unsigned int*>(amax + 1);
  unsigned int* aeq = aidx + 1;   // cells equal to the maximum
  unsigned int* asec = aidx + 2;  // largest value below the maximum (float bits)

  // D in LDS: rows r, columns c < NP, zero outside the N x N DFT
  for (int e = tid; e < NP * NP; e += blockDim.x) {
    const int r = e / NP, c = e - r * NP;
    double dr = 0., di = 0.;
    if (r < N && c < N) {
      const int m = (int)(((long long)r * c) % N);
      double s, co;
      sincospi(2.0 * (double)m / (double)N, &s, &co);
      dr = co;
      di = -s;
    }
    const __bf16 h0 = (__bf16)(float)dr, l0 = (__bf16)(float)(dr - (double)(float)h0);
    const __bf16 h1 = (__bf16)(float)di, l1 = (__bf16)(float)(di - (double)(float)h1);
    Dsh[0 * NP * RS + r * RS + c] = h0;
    Dsh[1 * NP * RS + r * RS + c] = l0;
    Dsh[2 * NP * RS + r * RS + c] = h1;
    Dsh[3 * NP * RS + r * RS + c] = l1;
  }
  for (int t = tid; t < 2 * N - 1; t += blockDim.x) gw[int(t)] = a.gw[int(t)];
  __syncthreads();
  const __bf16* Drh = Dsh;
  const __bf16* Drl = Dsh + 1 * NP * RS;
  const __bf16* Dih = Dsh + 2 * NP * RS;
  const __bf16* Dil = Dsh + 3 * NP * RS;

  const int nx = a.nx, ny = a.ny, nsize = a.nsize;
  const int ns = ny - N + 1;
  const float norm = (float)((1. / ((double)N * N)) * (1. / ((double)N * N)));
  const long long nslots = (long long)gridDim.x * C::PIX;
  const long long first = (long long)blockIdx.x * C::PIX + ps;
  const long long nwork = (a.list && a.nwork_dev) ? (long long)*a.nwork_dev : a.nwork;
  const long long iters = (nwork + nslots - 1) / nslots;
  const int npass = a.bias ? 2 : 1;

  for (long long it = 0; it < iters; ++it) {
    const long long w = first + it * nslots;
    const bool active = w < nwork;
    int i = 0, j = 0, imin = 0, jmin = 0;
    double f1bar = 0., f2bar = 0.;
    if (active) {
      const long long p = a.list ? (long long)a.list[w] : w;
      i = (int)(p / ny);
      j = (int)(p - (long long)i * ny);
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
      // ---------------- windows
      if (active) {
        const double cx = 0.5 * shiftx, cy = 0.5 * shifty;
        for (int t = ptid; t < 2 * NP; t += 128) {
          const bool isx = t < NP;
          const int q = isx ? t : t - NP;
          double w1 = 0., w2 = 0.;
          if (q < N) {
            const int d = q + (isx ? imin - i : jmin - j);
            if (pass == 0) {
              w1 = w2 = gw[d + N - 1];
            } else {
              const double off = isx ? cx : cy;
              const double a1 = a.sigminv * ((double)d + off), a2 = a.sigminv * ((double)d - off);
              w1 = exp(-a1 * a1);
              w2 = exp(-a2 * a2);
            }
          }
          if (isx) {
            wx1[q] = w1;
            wx2[q] = w2;
          } else {
            wy1[q] = w1;
            wy2[q] = w2;
          }
        }
        if (ptid == 0) {
          *amax = 0ull;
          *aidx = 0xffffffffu;
          *aeq = 0u;
          *asec = 0u;
        }
      }
      __syncthreads();

      // ---------------- forward transforms, one image at a time
      f32x4 S1r[HT][KHT], S1i[HT][KHT];  // image 1 spectrum tiles (l-tile, k-tile)
      f32x4 S2r[HT][KHT], S2i[HT][KHT];
      static_for<2>([&](auto X) {
        // G^T[c][x] (x contiguous), hi and lo
        __bf16* Gh = Bf;
        __bf16* Gl = Bf + NP * RS;
        if (active) {
          const double* f = (X == 0) ? a.f1 : a.f2;
          const double fbar = (X == 0) ? f1bar : f2bar;
          const double* wxx = (X == 0) ? wx1 : wx2;
          const double* wyy = (X == 0) ? wy1 : wy2;
          for (int e = ptid; e < NP * (NP / 8); e += 128) {
            const int c = e % NP, x0 = (e / NP) * 8;
            bf16x8 vh, vl;
            static_for<8>([&](auto t) {
              const int x = x0 + t;
              float g = 0.f;
              if (x < N && c < N)
                g = (float)((wxx[x] * wyy[c]) * (f[(long long)(imin + x) * ny + jmin + c] - fbar));
              __bf16 hh, ll;
              split_bf16(g, hh, ll);
              vh[int(t)] = hh;
              vl[int(t)] = ll;
            });
            *reinterpret_cast<bf16x8*>(Gh + c * RS + x0) = vh;
            *reinterpret_cast<bf16x8*>(Gl + c * RS + x0) = vl;
          }
        }
        __syncthreads();
        // stage 1: T^T(c, k) = sum_x G^T(c, x) D(k, x); this wave: c-tiles h*HT..
        f32x4 Tr[HT][KHT], Ti[HT][KHT];
        static_for<HT>([&](auto MT) {
          static_for<KHT>([&](auto NT) {
            Tr[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
            Ti[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
          });
        });
        if (active) {
          static_for<C::KSN>([&](auto KS) {
            static_for<HT>([&](auto MT) {
              const int crow = (h * HT + MT) * 16 + fr;
              const bf16x8 gh = ldfrag(Gh, crow, KS * 32, RS, fg);
              const bf16x8 gl = ldfrag(Gl, crow, KS * 32, RS, fg);
              static_for<KHT>([&](auto NT) {
                const int krow = NT * 16 + fr;
                const bf16x8 drh = ldfrag(Drh, krow, KS * 32, RS, fg);
                const bf16x8 drl = ldfrag(Drl, krow, KS * 32, RS, fg);
                const bf16x8 dih = ldfrag(Dih, krow, KS * 32, RS, fg);
                const bf16x8 dil = ldfrag(Dil, krow, KS * 32, RS, fg);
                Tr[MT][NT] = mfma3(gh, gl, drh, drl, Tr[MT][NT]);
                Ti[MT][NT] = mfma3(gh, gl, dih, dil, Ti[MT][NT]);
              });
            });
          });
        }
        __syncthreads();
        // T[k][c] row-major, hi/lo of r and i
        __bf16* Trh = Bf;
        __bf16* Trl = Bf + 1 * C::KH * RS;
        __bf16* Tih = Bf + 2 * C::KH * RS;
        __bf16* Til = Bf + 3 * C::KH * RS;
        if (active) {
          static_for<HT>([&](auto MT) {
            static_for<KHT>([&](auto NT) {
              const int k = NT * 16 + fr, c0 = (h * HT + MT) * 16 + 4 * fg;
              st4(Trh + k * RS + c0, Trl + k * RS + c0, Tr[MT][NT]);
              st4(Tih + k * RS + c0, Til + k * RS + c0, Ti[MT][NT]);
            });
          });
        }
        __syncthreads();
        // stage 2: S^T(l, k) = sum_c D(l, c) T(k, c); this wave: l-tiles h*HT..
        f32x4 Sr[HT][KHT], Si[HT][KHT];
        static_for<HT>([&](auto MT) {
          static_for<KHT>([&](auto NT) {
            Sr[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
            Si[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
          });
        });
        if (active) {
          static_for<C::KSN>([&](auto KS) {
            static_for<HT>([&](auto MT) {
              const int lrow = (h * HT + MT) * 16 + fr;
              const bf16x8 drh = ldfrag(Drh, lrow, KS * 32, RS, fg);
              const bf16x8 drl = ldfrag(Drl, lrow, KS * 32, RS, fg);
              const bf16x8 dih = ldfrag(Dih, lrow, KS * 32, RS, fg);
              const bf16x8 dil = ldfrag(Dil, lrow, KS * 32, RS, fg);
              const bf16x8 ndih = neg8(dih), ndil = neg8(dil);
              static_for<KHT>([&](auto NT) {
                const int krow = NT * 16 + fr;
                const bf16x8 trh = ldfrag(Trh, krow, KS * 32, RS, fg);
                const bf16x8 trl = ldfrag(Trl, krow, KS * 32, RS, fg);
                const bf16x8 tih = ldfrag(Tih, krow, KS * 32, RS, fg);
                const bf16x8 til = ldfrag(Til, krow, KS * 32, RS, fg);
                // Sr = Dr Tr - Di Ti, Si = Dr Ti + Di Tr
                Sr[MT][NT] = mfma3(drh, drl, trh, trl, Sr[MT][NT]);
                Sr[MT][NT] = mfma3(ndih, ndil, tih, til, Sr[MT][NT]);
                Si[MT][NT] = mfma3(drh, drl, tih, til, Si[MT][NT]);
                Si[MT][NT] = mfma3(dih, dil, trh, trl, Si[MT][NT]);
              });
            });
          });
        }
        static_for<HT>([&](auto MT) {
          static_for<KHT>([&](auto NT) {
            if constexpr (X == 0) {
              S1r[MT][NT] = Sr[MT][NT];
              S1i[MT][NT] = Si[MT][NT];
            } else {
              S2r[MT][NT] = Sr[MT][NT];
              S2i[MT][NT] = Si[MT][NT];
            }
          });
        });
        __syncthreads();
      });

      // ---------------- cross power P(k, l) in the S^T tile layout:
      // lane element t of tile (MT, NT) is k = NT*16 + fr, l = (h*HT+MT)*16 + 4 fg + t
      __bf16* Prh = Bf;
      __bf16* Prl = Bf + 1 * C::KH * RS;
      __bf16* Pih = Bf + 2 * C::KH * RS;
      __bf16* Pil = Bf + 3 * C::KH * RS;
      if (active) {
        static_for<HT>([&](auto MT) {
          static_for<KHT>([&](auto NT) {
            const int k = NT * 16 + fr, l0 = (h * HT + MT) * 16 + 4 * fg;
            f32x4 pr, pi;
            static_for<4>([&](auto t) {
              const int l = l0 + t;
              float ar = S1r[MT][NT][int(t)], ai = S1i[MT][NT][int(t)];
              float br = S2r[MT][NT][int(t)], bi = S2i[MT][NT][int(t)];
              if (a.filt && k <= N / 2 && l < N) {
                const float fl = (float)a.filt[k * N + l];
                ar *= fl;
                ai *= fl;
                br *= fl;
                bi *= fl;
              }
              float r_ = (ar * br + ai * bi) * norm;
              float i_ = (ar * bi - ai * br) * norm;
              if (k > N / 2 || l >= N) r_ = i_ = 0.f;
              pr[int(t)] = r_;
              pi[int(t)] = i_;
            });
            st4(Prh + k * RS + l0, Prl + k * RS + l0, pr);
            st4(Pih + k * RS + l0, Pil + k * RS + l0, pi);
          });
        });
      }
      __syncthreads();
      // ---------------- inverse stage 1: U(k, jj) = sum_l P(k, l) conj(D)(l, jj)
      //   Ur = Pr Dr + Pi Di, Ui = Pi Dr - Pr Di; this wave: jj-tiles h*HT..
      f32x4 Ur[KHT][HT], Ui[KHT][HT];
      static_for<KHT>([&](auto MT) {
        static_for<HT>([&](auto NT) {
          Ur[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
          Ui[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f};
        });
      });
      if (active) {
        static_for<C::KSN>([&](auto KS) {
          static_for<HT>([&](auto NT) {
            const int jrow = (h * HT + NT) * 16 + fr;
            const bf16x8 drh = ldfrag(Drh, jrow, KS * 32, RS, fg);
            const bf16x8 drl = ldfrag(Drl, jrow, KS * 32, RS, fg);
            const bf16x8 dih = ldfrag(Dih, jrow, KS * 32, RS, fg);
            const bf16x8 dil = ldfrag(Dil, jrow, KS * 32, RS, fg);
            const bf16x8 ndih = neg8(dih), ndil = neg8(dil);
            static_for<KHT>([&](auto MT) {
              const int krow = MT * 16 + fr;
              const bf16x8 prh = ldfrag(Prh, krow, KS * 32, RS, fg);
              const bf16x8 prl = ldfrag(Prl, krow, KS * 32, RS, fg);
              const bf16x8 pih = ldfrag(Pih, krow, KS * 32, RS, fg);
              const bf16x8 pil = ldfrag(Pil, krow, KS * 32, RS, fg);
              Ur[MT][NT] = mfma3(prh, prl, drh, drl, Ur[MT][NT]);
              Ur[MT][NT] = mfma3(pih, pil, dih, dil, Ur[MT][NT]);
              Ui[MT][NT] = mfma3(pih, pil, drh, drl, Ui[MT][NT]);
              Ui[MT][NT] = mfma3(prh, prl, ndih, ndil, Ui[MT][NT]);
            });
          });
        });
      }
      __syncthreads();
      // U^T[jj][k] (k contiguous, KP wide), rows weighted by a_k (0 beyond N/2)
      constexpr int RSU = C::RSU;
      __bf16* Urh = Bf;
      __bf16* Url = Bf + 1 * NP * RSU;
      __bf16* Uih = Bf + 2 * NP * RSU;
      __bf16* Uil = Bf + 3 * NP * RSU;
      if (active) {
        static_for<KHT>([&](auto MT) {
          static_for<HT>([&](auto NT) {
            const int jj = (h * HT + NT) * 16 + fr, k0 = MT * 16 + 4 * fg;
            f32x4 ur = Ur[MT][NT], ui = Ui[MT][NT];
            static_for<4>([&](auto t) {
              const int k = k0 + t;
              const float al = (k == 0 || k == N / 2) ? 1.f : ((k < N / 2) ? 2.f : 0.f);
              ur[int(t)] *= al;
              ui[int(t)] *= al;
            });
            st4(Urh + jj * RSU + k0, Url + jj * RSU + k0, ur);
            st4(Uih + jj * RSU + k0, Uil + jj * RSU + k0, ui);
          });
        });
        // zero the contraction padding k = KH..KP-1
        if constexpr (C::KP > C::KH) {
          for (int e = h * 64 + lane; e < NP * (C::KP - C::KH); e += 128) {
            const int jj = e / (C::KP - C::KH), k = C::KH + e % (C::KP - C::KH);
            Urh[jj * RSU + k] = Url[jj * RSU + k] = Uih[jj * RSU + k] = Uil[jj * RSU + k] =
                (__bf16)0.f;
          }
        }
      }
      __syncthreads();
      // ---------------- inverse stage 2: c(ii, jj) = sum_k Dr(ii,k) U'r(k,jj) + Di(ii,k) U'i(k,jj)
      f32x4 cc[HT][TP];
      static_for<HT>([&](auto MT) {
        static_for<TP>([&](auto NT) { cc[MT][NT] = f32x4{0.f, 0.f, 0.f, 0.f}; });
      });
      if (active) {
        static_for<C::KSP>([&](auto KS) {
          static_for<HT>([&](auto MT) {
            const int irow = (h * HT + MT) * 16 + fr;
            const bf16x8 drh = ldfrag(Drh, irow, KS * 32, RS, fg);
            const bf16x8 drl = ldfrag(Drl, irow, KS * 32, RS, fg);
            const bf16x8 dih = ldfrag(Dih, irow, KS * 32, RS, fg);
            const bf16x8 dil = ldfrag(Dil, irow, KS * 32, RS, fg);
            static_for<TP>([&](auto NT) {
              const int jrow = NT * 16 + fr;
              const bf16x8 urh = ldfrag(Urh, jrow, KS * 32, RSU, fg);
              const bf16x8 url = ldfrag(Url, jrow, KS * 32, RSU, fg);
              const bf16x8 uih = ldfrag(Uih, jrow, KS * 32, RSU, fg);
              const bf16x8 uil = ldfrag(Uil, jrow, KS * 32, RSU, fg);
              cc[MT][NT] = mfma3(drh, drl, urh, url, cc[MT][NT]);
              cc[MT][NT] = mfma3(dih, dil, uih, uil, cc[MT][NT]);
            });
          });
        });
      }
      // ---------------- |c|, argmax with the first-strict-maximum rule, fit
      // lane element t of tile (MT, NT): ii = (h*HT+MT)*16 + 4 fg + t, jj = NT*16 + fr
      if (active) {
        float lmax = 0.f;
        static_for<HT>([&](auto MT) {
          static_for<TP>([&](auto NT) {
            static_for<4>([&](auto t) {
              const int ii = (h * HT + MT) * 16 + 4 * fg + t, jj = NT * 16 + fr;
              const float v = fabsf(cc[MT][NT][int(t)]);
              cc[MT][NT][int(t)] = v;
              if (ii < N && jj < N) lmax = fmaxf(lmax, v);
            });
          });
        });
        for (int o = 32; o > 0; o >>= 1) lmax = fmaxf(lmax, __shfl_xor(lmax, o));
        if (lane == 0) atomicMax(amax, dbits((double)lmax));
      }
      __syncthreads();
      const double gmax = __longlong_as_double((long long)*amax);
      if (active) {
        unsigned best = 0xffffffffu;
        unsigned neq = 0u;
        float sec = 0.f;
        static_for<HT>([&](auto MT) {
          static_for<TP>([&](auto NT) {
            static_for<4>([&](auto t) {
              const int ii = (h * HT + MT) * 16 + 4 * fg + t, jj = NT * 16 + fr;
              if (ii < N && jj < N) {
                const float v = cc[MT][NT][int(t)];
                if ((double)v == gmax) {
                  const unsigned L = (unsigned)(((ii + N / 2) % N) * N + (jj + N / 2) % N);
                  best = min(best, L);
                  ++neq;
                } else {
                  sec = fmaxf(sec, v);
                }
              }
            });
          });
        });
        for (int o = 32; o > 0; o >>= 1) {
          best = min(best, (unsigned)__shfl_xor((int)best, o));
          neq += (unsigned)__shfl_xor((int)neq, o);
          sec = fmaxf(sec, __shfl_xor(sec, o));
        }
        if (lane == 0) {
          if (best != 0xffffffffu) atomicMin(aidx, best);
          atomicAdd(aeq, neq);
          atomicMax(asec, __float_as_uint(sec));
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
      if (active && fit) {
        static_for<HT>([&](auto MT) {
          static_for<TP>([&](auto NT) {
            static_for<4>([&](auto t) {
              const int ii = (h * HT + MT) * 16 + 4 * fg + t, jj = NT * 16 + fr;
              if (ii < N && jj < N) {
                const int di = (ii + N / 2) % N - ixmax, dj = (jj + N / 2) % N - iymax;
                if (di >= -1 && di <= 1 && dj >= -1 && dj <= 1)
                  nb[(di + 1) * 3 + (dj + 1)] = (double)cc[MT][NT][int(t)];
              }
            });
          });
        });
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
          // first-order bound of the fit's response to relative errors
          // kRelErr of the nine values
          const double det = fxx * fyy - fxy * fxy;
          const double hn = fabs(fxx) + fabs(fyy) + 2. * fabs(fxy);
          const double bound = kRelErr * gmax * hn * (1. + 4. * (fabs(shiftsubx) + fabs(shiftsuby)));
          if (!(bound <= kFitTol * fabs(det))) flagged = true;
        }
        // another cell within the split-BF16 error of the maximum
        if (nonzero && (*aeq > 1u || (double)__uint_as_float(*asec) >= gmax * (1. - kTieTol)))
          flagged = true;
        res[0] = ((double)ixmax + shiftsubx) - (double)(N / 2);
        res[1] = ((double)iymax + shiftsuby) - (double)(N / 2);
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
