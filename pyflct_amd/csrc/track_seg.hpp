// Split-BF16 MFMA tracker, row-segment form, for boxes N = 8..64 with
// N % 4 == 0 and columns whose box is not clamped at the left or right edge.
//
// The windowed sub-image of pixel (i, j) is g(x, c) = Wr(x) Wc(c) (f(imin + x,
// jmin + c) - fbar) (flct.c:815-870).  Every pixel of one image row shares
// imin and the row window Wr, and for interior columns jmin = j - (nsize - 1)
// and the column window Wc are the same up to a shift of the box.  So the
// transform along the box rows is done once per segment of P <= 64
// neighbouring pixels of a row, for every image column the segment's boxes
// touch:
//
//   stage 1 (shared, x -> k):  T(u, k) = sum_x Wr(x) (f(imin + x, cb + u) - cm(u)) D(x, k),
//            cm(u) the column's Wr-weighted mean over the box rows (so the
//            windowed DC of every column is zero), D(x, k) = exp(-2 pi i k x / N)
//            for k = 1 .. N/2-1; the real DC and Nyquist rows are packed into
//            k = 0 as T_dc + i T_nyq by the table (D(x, 0) = 1 + i (-1)^x).  A
//            column of ones (u = 127) gives sc Wr^(k).
//   stage 2 (per pixel, c -> l, radix-2 on c, window folded into the tables):
//            E(l) = sum_c' Wc(2c') T(p + 2c', k) e^{-2 pi i l c' / Nh},
//            O(l) = sum_c' Wc(2c'+1) T(p + 2c' + 1, k) e^{-2 pi i l (2c'+1) / N},
//            S(l) = E + O, S(l + Nh) = E - O, plus the mean correction
//            Wr^(k) V(l), V the same transform of the column vector
//            cm(p + c) - fbar (exact in real arithmetic; every term is of the
//            size of the box's own variation, and a constant box gives S = 0
//            exactly, the reference's max == 0 case).  V of all the segment's
//            pixels is one MFMA product per segment (B column = pixel), kept
//            in a block-private slice of device scratch.
//   cross power P = conj(S1) S2 (flct.c:1461-1477; normfac dropped: the
//            argmax and the fit are scale invariant), times filter^2
//            (flct.c:1445-1457); the packed column is unpacked through l <-> -l.
//   inverse 1 (l -> jj, radix-2 DIF): U(k, 2m'+q) from Q0 = P(l) + P(l+Nh) and
//            Q1 = P(l) - P(l+Nh), the twiddle folded into the q = 1 table.
//   inverse 2 (k -> ii, Hermitian half sum, radix-2 on ii):
//            c(ii, jj) = Ce + Co, c(ii + Nh, jj) = Ce - Co over even / odd k.
//
// TL is the number of 16-wide tiles along k, l and the outputs: 2 for boxes
// 36..64, 1 for boxes up to 32 (Nh <= 16), where the second tiles would be
// padding.
//
// One wave tracks one pixel at a time.  Stage-2 outputs are rows l for the
// lane's column k, which is exactly the A operand (row k, 8 contraction
// values l) of inverse 1; inverse 1's outputs are rows k for the lane's column
// jj, the A operand of inverse 2.  So after stage 1 nothing of a pixel passes
// through LDS except the packed-column exchange; the 3x3 neighbourhood of the
// peak is read from its lanes with v_readlane.
// Products are hi*hi + hi*lo + lo*hi of BF16 splits accumulated in FP32
// (v_mfma_f32_16x16x32_bf16); the argmax uses the values only, the 3x3 Taylor
// fit (flct.c:1536-1556) is FP64.  Pixels the BF16 surface cannot settle (a
// second cell within kTieTol of the maximum, an ill-conditioned fit) are
// listed for FP64 re-evaluation as in the per-pixel kernels.
#pragma once

#include "track_mfma2.hpp"

namespace flct {

struct SegArgs {
  TrackArgs t;           // images, box size, means, outputs, filter, flags, scales
  const __bf16* tab;     // split tables, SegCfg layout (global)
  double thresh;         // threshold (flct.c:721-722), 0: off
  int thresh_rel;        // thresh is relative: times max(max |f1|, max |f2|) from t.absmax
  int skip, xoffset, yoffset;  // flct.c:734-757 (transp = 1)
  int jlo;               // first interior column, nsize - 1
  int width;             // interior columns, ny - N + 1
  int nband;             // column bands of at most 64 pixels
  int nseg;              // nband * nx segments
  int xchunk;            // > 0: XCD-grouped order with this many row chunks (see k_track_seg)
  float* vscr;           // per block SegCfg::kVFloats: the segment's mean-correction transforms
  // rows run: [rlo, rlo + nrows) (the whole image: 0, nx)
  int rlo, nrows;
  // strip mode: the images are a transposed edge strip, strip row t = image
  // column cb0 + t and strip column = image row, so that the strip's
  // interior columns are the image's clamped edge columns (flct.c:780-786)
  // outside its edge rows.  Outputs go to image pixel (strip column) * nyo +
  // cb0 + (strip row), with vx and vy exchanged; the skip lattice is tested
  // in image coordinates.
  int strip, cb0, nyo;
};

struct SegCfg {
  static constexpr int P = 64;            // pixels per segment, at most
  // waves per block: one stage-1 column tile each (4- and 12-wave blocks
  // measured slower)
  static constexpr int kWaves = 8;
  static constexpr int kThreads = 64 * kWaves;
  // table row stride (bf16): 48 = 24 dwords.  A fragment load is a
  // ds_read_b128 of row n = lane & 15 at dword 4 (lane >> 4); measured on
  // gfx950 (SQ_LDS_BANK_CONFLICT), row strides of 24 or 40 dwords serve it
  // without bank conflicts, 20 dwords (the padded minimum) costs 4 conflict
  // cycles per load
  static constexpr int RS = 48;
  static constexpr int kPlane = 32 * RS;
  static constexpr int RS1 = 72;          // stage-1 table row stride, 64 + 8
  static constexpr int kD1Plane = 32 * RS1;
  // global table: D1 (4 planes, stays in global), then the LDS part
  static constexpr int kLdsTabOff = 4 * kD1Plane;
  // LDS tables (bf16 planes of 32 rows): De, Do (re_h, re_l, im_h, im_l,
  // -im_h, -im_l), BI0, BI1 (re_h, re_l, im_h, im_l), WT (set 0 h, l, set 1 h, l)
  static constexpr int kDe = 0, kDo = 6 * kPlane, kBI0 = 12 * kPlane, kBI1 = 16 * kPlane,
                       kWT = 20 * kPlane;
  static constexpr int kTabEl = 24 * kPlane;
  static constexpr int kGlobalEl = kLdsTabOff + kTabEl;
  // T planes [X][par][k][comp][uh], comp = re_h, re_l, im_h, im_l.  A k row
  // is 4 runs of RW half columns plus padding: 138 dwords, so the 16 rows a
  // fragment read touches start 10 banks apart and a lane's dword pair never
  // shares a bank with another lane's
  static constexpr int RW = 68;           // >= 32 + 32 + 1 (odd reads), even
  static constexpr int RK = 4 * RW + 4;   // k row stride
  static constexpr int kTPlane = 32 * RK;
  static constexpr int kTEl = 4 * kTPlane;
  static constexpr int kScratch = 128 * 8 + 64 * 8;  // zx, zp
  static constexpr int kGw = 128;  // window table gw[d + N - 1], 2N - 1 <= 127 doubles
  // mean-correction transforms of one segment (global, per block):
  // [pixel P][l-tile 2][image 2][column parity 2][re, im][l 16] floats
  static constexpr int kVFloats = P * 2 * 2 * 2 * 2 * 16;
  static constexpr int kLdsBytes =
      kTabEl * 2 + 256 * 8 + kTEl * 2 + 68 * 4 + kWaves * kScratch + kGw * 8;
  static_assert(kLdsBytes <= 160 * 1024, "k_track_seg needs more LDS than a CU has");
};

__device__ __forceinline__ bf16x8 u4_to_bf8(u32x4 v) { return __builtin_bit_cast(bf16x8, v); }

// 8 consecutive bf16 of an LDS row from any element (odd starts uniform over
// the wave).  Branch-free: five dwords from the one at or below the start,
// q, shifted by sh = 2 bytes when the start is odd (by 0, the low dword
// itself, when it is even), so the loads of one tile can be scheduled
// across the MFMAs of the previous one.  Callers form q from one address
// per run, so that every read is an immediate offset from it.
__device__ __forceinline__ bf16x8 ld8_sh(const unsigned* q, unsigned sh) {
  const unsigned d0 = q[0], d1 = q[1], d2 = q[2], d3 = q[3], d4 = q[4];
  u32x4 r;
  r[0] = __builtin_amdgcn_alignbyte(d1, d0, sh);
  r[1] = __builtin_amdgcn_alignbyte(d2, d1, sh);
  r[2] = __builtin_amdgcn_alignbyte(d3, d2, sh);
  r[3] = __builtin_amdgcn_alignbyte(d4, d3, sh);
  return u4_to_bf8(r);
}

__device__ __forceinline__ float wave_max(float v) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

typedef const __attribute__((address_space(4))) SegArgs* SegArgsPtr;

// The kernel's argument block, read through an opaque pointer: loads of the
// fields are issued where they are used instead of being kept in scalar
// registers across the segment and pixel loops.
__device__ __forceinline__ SegArgsPtr seg_args() {
  SegArgsPtr p = __builtin_bit_cast(SegArgsPtr, __builtin_amdgcn_kernarg_segment_ptr());
  asm volatile("" : "+s"(p));
  return p;
}

template <int TL>
__global__ void __launch_bounds__(SegCfg::kThreads, 1) k_track_seg(SegArgs A0) {
  using C = SegCfg;
  constexpr int RS = C::RS, PL = C::kPlane, RW = C::RW, RK = C::RK, TP = C::kTPlane;
  extern __shared__ __align__(16) unsigned char lds[];
  (void)A0;
  SegArgsPtr Ak = seg_args();
  const auto& A = *Ak;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int n0 = lane & 15, g0 = lane >> 4;

  __bf16* TB = reinterpret_cast<__bf16*>(lds);
  double* cmd = reinterpret_cast<double*>(TB + C::kTabEl);  // [X][u] cm(u)
  __bf16* TT = reinterpret_cast<__bf16*>(cmd + 256);
  int* slist = reinterpret_cast<int*>(TT + C::kTEl);
  int* scount = slist + 64;
  double* gwl = reinterpret_cast<double*>(reinterpret_cast<unsigned char*>(slist + 68) +
                                          C::kWaves * C::kScratch);  // window table
  unsigned char* wsc = reinterpret_cast<unsigned char*>(slist + 68) + wave * C::kScratch;
  // Packed DC / Nyquist column per wave, planar: zx [image][re, im][64] the
  // spectra, zp [re, im][64] their cross power.  Row l < Nh sits at l and
  // row l + Nh at 32 + l, so a group of four rows is one aligned f32x4 for
  // every box size.
  float* zx = reinterpret_cast<float*>(wsc);
  float* zp = zx + 256;

  // ---------------- tables to LDS, T planes zeroed (padding stays zero)
  {
    const u32x4* src = reinterpret_cast<const u32x4*>(A.tab + C::kLdsTabOff);
    u32x4* dst = reinterpret_cast<u32x4*>(TB);
    for (int e = tid; e < C::kTabEl / 8; e += blockDim.x) dst[e] = src[e];
    u32x4* tz = reinterpret_cast<u32x4*>(TT);
    const u32x4 z = {0u, 0u, 0u, 0u};
    for (int e = tid; e < C::kTEl / 8; e += blockDim.x) tz[e] = z;
    for (int e = tid; e < 2 * A.t.nsize * 2 - 1; e += blockDim.x) gwl[e] = A.t.gw[e];
  }

  const f32x4 z4 = {0.f, 0.f, 0.f, 0.f};
  // Segment order.  Default: block b takes segments [b nseg / G, (b+1) nseg
  // / G) of the band-major order, consecutive rows of one band.  XCD order
  // (xchunk > 0): blocks are dealt to the 8 XCDs round robin, XCD x = b % 8
  // owns the bands [x nband / 8, (x+1) nband / 8), and its blocks take one
  // band and one of xchunk row chunks each, adjacent bands side by side; the
  // blocks of an XCD then read overlapping box columns of the same rows, and
  // their working set stays in the XCD's L2.
  const int G = gridDim.x;
  int s0, s1, sband = 0;
  if (A.xchunk > 0) {
    const int x = blockIdx.x & 7, k = blockIdx.x >> 3;
    const int nbg = A.nband >> 3;
    sband = x * nbg + k % nbg;
    const int ch = k / nbg;
    s0 = (int)((long long)ch * A.t.nx / A.xchunk);
    s1 = (int)((long long)(ch + 1) * A.t.nx / A.xchunk);
  } else {
    s0 = (int)((long long)blockIdx.x * A.nseg / G);
    s1 = (int)((long long)(blockIdx.x + 1) * A.nseg / G);
  }

  for (int sidx = s0; sidx < s1; ++sidx) {
    SegArgsPtr As = seg_args();
    const auto& A = *As;
    const auto& a = A.t;
    const int nx = a.nx, ny = a.ny, nsize = a.nsize;
    const int N = 2 * nsize, Nh = nsize;
    const int ns = ny - N + 1;
    // image scales: powers of two bringing max |f| to [1, 2)
    double sc[2] = {1., 1.};
    if (a.absmax) {
      static_for<2>([&](auto X) {
        const double m = __longlong_as_double((long long)a.absmax[X]);
        if (m > 0. && m < 1e300) sc[X] = ldexp(1., -ilogb(m));
      });
    }
    const int band = (A.xchunk > 0) ? sband : sidx / A.nrows;
    const int i = (A.xchunk > 0) ? sidx : A.rlo + (sidx - band * A.nrows);
    const int jb0 = A.jlo + (int)((long long)band * A.width / A.nband);
    const int jb1 = A.jlo + (int)((long long)(band + 1) * A.width / A.nband);
    const int P = jb1 - jb0;
    // box rows (flct.c:773-779)
    const int imin0 = (0 > (i - (nsize - 1))) ? 0 : i - (nsize - 1);
    const int imax0 = ((nx - 1) < (i + nsize)) ? nx - 1 : i + nsize;
    const int imin = (imax0 == nx - 1) ? nx - 1 - (2 * nsize - 1) : imin0;

    // Stage-1 loads of this segment, issued before its barriers so that
    // their latency overlaps the wait for the other waves (107.5 -> 105.7 ms
    // per 4096^2 step): the box column values of both images of the wave's
    // column tile ct = wave (rows x = 32 s + 8 g + e of this lane) and the
    // stage-1 table fragments.
    static_assert(C::kWaves == 8, "one stage-1 column tile per wave");
    const int cb = jb0 - (nsize - 1);  // image column of T column 0
    const int ucols = P + N - 1;       // T columns the boxes touch
    // per-segment copies of the lane indices: keeps the table loads below
    // from being hoisted out of the segment loop
    const int n1 = fresh(n0), g1 = fresh(g0);
    const int u = 16 * wave + n1;
    const bool colok = u < ucols;
    const bool ones = (u == 127);
    const int ucl = colok ? u : 0;  // loads stay inside the image
    double fva[2][16];
    static_for<2>([&](auto X) {
      const double* f = (X == 0) ? a.f1 : a.f2;
      static_for<16>([&](auto q) {
        const int x = 32 * (q >> 3) + 8 * g1 + (q & 7);
        const int xr = (x < N) ? x : 0;
        fva[X][int(q)] = f[(long long)(imin + xr) * ny + (cb + ucl)];
      });
    });
    bf16x8 d1f[2][2][4];  // [kt][s][re_h, re_l, im_h, im_l]
    static_for<TL>([&](auto kt) {
      static_for<2>([&](auto s) {
        const int off = (16 * kt + n1) * C::RS1 + 32 * s + 8 * g1;
        static_for<4>([&](auto pl) { d1f[kt][s][pl] = ld8(A.tab + pl * C::kD1Plane + off); });
      });
    });

    // the pixel values of the threshold test (wave 0), read only with a
    // threshold: with thresh = 0 the test fabsbar < thresh fails for every
    // pixel (flct.c:721-722)
    double f1t = 0., f2t = 0.;
    if (tid < 64 && tid < P && A.thresh != 0.) {
      const long long t = (long long)i * ny + (jb0 + tid);
      f1t = a.f1[t];
      f2t = a.f2[t];
    }
    // the box means of this wave's two columns of the mean-correction
    // product below (images 0 and 1, pixel 16 (wave & 3) + n)
    const int vpp = 16 * (wave & 3) + n1;
    const bool vpin = vpp < P;
    double fbv[2];
    {
      const long long vo = (long long)imin * ns + (jb0 + (vpin ? vpp : 0) - (nsize - 1));
      fbv[0] = a.mean1[vo];
      fbv[1] = a.mean2[vo];
    }

    __syncthreads();  // the previous segment's pixels are done with T and slist
    // tracked pixels of the segment: threshold and skip rules (flct.c:721-757)
    if (tid < 64) {
      bool tr = false;
      if (tid < P) {
#pragma clang fp contract(off)
        const int j = jb0 + tid;
        const double fabsbar = 0.5 * (fabs(f1t + f2t));
        const bool below = (fabsbar < resolve_thresh(A.thresh, A.thresh_rel ? a.absmax : nullptr));
        bool noskip = true;
        // image coordinates of the pixel (transposed in strip mode)
        const int io = A.strip ? j : i, jo = A.strip ? A.cb0 + i : j;
        if (A.skip) noskip = (!((io - A.xoffset) % A.skip)) && (!((jo - A.yoffset) % A.skip));
        tr = (!below) && noskip;
      }
      const unsigned long long m = __ballot(tr);
      if (tr) slist[__popcll(m & ((1ull << tid) - 1ull))] = tid;
      if (tid == 0) *scount = __popcll(m);
    }
    __syncthreads();
    const int cnt = __builtin_amdgcn_readfirstlane(*scount);
    if (cnt == 0) continue;

    // ---------------- stage 1: column tile ct = wave (T columns 16ct .. 16ct+15)
    {
      const int ct = wave;
      const SplitK sk = split_consts();
      float wr[16];
      static_for<16>([&](auto q) {
        const int x = 32 * (q >> 3) + 8 * g1 + (q & 7);
        wr[int(q)] = (x < N) ? (float)gwl[(x < N ? x : 0) + imin - i + N - 1] : 0.f;
      });
      static_for<2>([&](auto X) {
        const float scf = (float)sc[X];
        bf16x8 Ah[2], Al[2];
        const double* fv = fva[X];
        double csum = 0., wsum = 0.;
        static_for<16>([&](auto q) {
          csum += (double)wr[int(q)] * fv[int(q)];  // wr is zero for x >= N
          wsum += (double)wr[int(q)];
        });
        // window-weighted column mean over the box rows (the four lanes of a
        // column hold 16 rows each): every column's windowed DC term is zero
        csum += __shfl_xor(csum, 16);
        csum += __shfl_xor(csum, 32);
        wsum += __shfl_xor(wsum, 16);
        wsum += __shfl_xor(wsum, 32);
        const double cm = csum / wsum;
        // the four lanes of a column hold the same value: no branch
        cmd[128 * X + u] = colok ? cm : 0.;
        static_for<2>([&](auto s) {
          float v[8];
          static_for<8>([&](auto e) {
            const float ws = wr[8 * s + e] * scf;  // zero for x >= N
            const float gv = (float)(fv[8 * s + e] - cm) * ws;
            v[int(e)] = ones ? ws : (colok ? gv : 0.f);
          });
          u32x4 hv, lv;
          static_for<4>([&](auto d) {
            unsigned hh, ll;
            split2(v[2 * d], v[2 * d + 1], hh, ll, sk);
            hv[int(d)] = hh;
            lv[int(d)] = ll;
          });
          Ah[s] = u4_to_bf8(hv);
          Al[s] = u4_to_bf8(lv);
        });
        static_for<TL>([&](auto kt) {
          const int k = 16 * kt + n1;
          f32x4 tre = z4, tim = z4;
          static_for<2>([&](auto s) {
            tre = mfma3(Ah[s], Al[s], d1f[kt][s][0], d1f[kt][s][1], tre);
            tim = mfma3(Ah[s], Al[s], d1f[kt][s][2], d1f[kt][s][3], tim);
          });
          // lane holds T(u = 16w + 4g + t, k = 16kt + n): parity t & 1,
          // half column uh = 8w + 2g + (t >> 1)
          const int uh = 8 * ct + 2 * g1;
          static_for<2>([&](auto par) {
            __bf16* pl = TT + (X * 2 + par) * TP + k * RK + uh;
            unsigned h0, l0, h1, l1;
            split2(tre[int(par)], tre[int(par) + 2], h0, l0, sk);
            split2(tim[int(par)], tim[int(par) + 2], h1, l1, sk);
            *reinterpret_cast<unsigned*>(pl + 0 * RW) = h0;
            *reinterpret_cast<unsigned*>(pl + 1 * RW) = l0;
            *reinterpret_cast<unsigned*>(pl + 2 * RW) = h1;
            *reinterpret_cast<unsigned*>(pl + 3 * RW) = l1;
          });
        });
      });
    }
    __syncthreads();  // T complete

    // sc_X Wr^(k) for the lane's k (T column 127: parity 1, uh 63)
    float2 wrh[2][2];  // [X][kt]
    static_for<2>([&](auto X) {
      static_for<TL>([&](auto kt) {
        const __bf16* pl = TT + (X * 2 + 1) * TP + (16 * kt + n0) * RK + 63;
        const float re = (float)pl[0 * RW] + (float)pl[1 * RW];
        const float im = (float)pl[2 * RW] + (float)pl[3 * RW];
        wrh[X][kt] = make_float2(re, im);
      });
    });

    // ---------------- mean-correction transforms of the segment's pixels
    // V_X(l; p) = sum_c' D_s(l, c') (cm(p + 2c' + s) - fbar_p), the stage-2
    // transform of pixel p's real column vector over the even (s = 0) or odd
    // (s = 1) columns, exact in real arithmetic beside the box-row spectra.
    // One MFMA product per segment serves all its pixels: B column n is
    // pixel 16 PT + n, so each output lane holds V for its column's pixel at
    // the rows l = 16 LT + 4g + t that the pixel loop's accumulators take.
    // The vector is formed in FP64 and rounded once (the column means and
    // the box mean can be large beside their difference).
    float* vb = A.vscr + (long long)blockIdx.x * C::kVFloats;
    static_for<2>([&](auto X) {  // (X, s, PT) = (X, (wave >> 2) & 1, wave & 3)
      const int g1 = fresh(g0);
      const SplitK sk = split_consts();
      const int sp = (wave >> 2) & 1;
      const int pp = vpp;
      const bool pin = vpin;
      const double fb = fbv[X];
      const double* cr = cmd + 128 * X + pp + sp + 16 * g1;  // index <= 126
      float v[8];
      static_for<8>([&](auto e) { v[int(e)] = pin ? (float)(cr[2 * e] - fb) : 0.f; });
      u32x4 hv, lv;
      static_for<4>([&](auto d) {
        unsigned hh, ll;
        split2(v[2 * d], v[2 * d + 1], hh, ll, sk);
        hv[int(d)] = hh;
        lv[int(d)] = ll;
      });
      const bf16x8 vh = u4_to_bf8(hv), vl = u4_to_bf8(lv);
      static_for<TL>([&](auto LT) {
        const __bf16* D = TB + (n1 * RS + 8 * g1) + (sp ? C::kDo : C::kDe) + 16 * LT * RS;
        const f32x4 rr = mfma3(ld8(D + 0 * PL), ld8(D + 1 * PL), vh, vl, z4);
        const f32x4 ri = mfma3(ld8(D + 2 * PL), ld8(D + 3 * PL), vh, vl, z4);
        float* o = vb + ((((pp * TL + LT) * 2 + X) * 2 + sp) * 2) * 16 + 4 * g1;
        *reinterpret_cast<f32x4*>(o) = rr;
        *reinterpret_cast<f32x4*>(o + 16) = ri;
      });
    });
    __syncthreads();  // V of every pixel stored

    // ---------------- pixels of the segment, one per wave at a time
    for (int qi = wave; qi < cnt; qi += C::kWaves) {
      SegArgsPtr Ap = seg_args();
      const auto& a = Ap->t;
      const int p = __builtin_amdgcn_readfirstlane(slist[qi]);
      const int j = jb0 + p;
      // per-pixel copies of lane indices: keeps index math out of registers
      // across the pixel loop
      const int n = fresh(n0), g = fresh(g0);
      const SplitK sk = split_consts();
      // lane bases: table row n, T row k = n, 8 contraction values from 8g;
      // every operand below is an immediate offset from one of them
      const __bf16* TBl = TB + (n * RS + 8 * g);
      const __bf16* TTl = TT + (n * RK + 8 * g);
      const int pe = p & 1, be = p >> 1;             // even columns c = 2c'
      const int po = pe ^ 1, bo = (p + 1) >> 1;      // odd columns
      const bool oe = (be & 1) != 0, oo = (bo & 1) != 0;

      // A operands of stage 2 (window-folded half-DFT tables) for l-tile LT;
      // the row index is refreshed per use so that the loads are not kept
      // live across the loops below
      struct DFr {
        bf16x8 rh, rl, ih, il, nh, nl;
      };
      auto dtab = [&](int base, auto LT) {
        const __bf16* D = TBl + base + 16 * LT * RS;
        DFr d;
        d.rh = ld8(D + 0 * PL);
        d.rl = ld8(D + 1 * PL);
        d.ih = ld8(D + 2 * PL);
        d.il = ld8(D + 3 * PL);
        d.nh = ld8(D + 4 * PL);
        d.nl = ld8(D + 5 * PL);
        return d;
      };
      // complex (D T) over the 32 half columns: re = Dr Tr - Di Ti, im = Dr Ti + Di Tr
      // re, im enter as the accumulators' starting values
      auto cmul_mfma = [&](const DFr& d, bf16x8 trh, bf16x8 trl, bf16x8 tih, bf16x8 til,
                           f32x4& re, f32x4& im) {
        re = mfma3(d.rh, d.rl, trh, trl, re);
        re = mfma3(d.nh, d.nl, tih, til, re);
        im = mfma3(d.rh, d.rl, tih, til, im);
        im = mfma3(d.ih, d.il, trh, trl, im);
      };

      CF4 Pp[2][2], Pm[2][2];  // [kt][LT] cross power
      // the pixel's mean-correction transforms for every l-tile, loaded at
      // once (the same for every lane n)
      f32x4 vld[TL][2][2][2];  // [LT][X][s][re, im]
      {
        const float* vp = vb + p * (TL * 128) + 4 * g;
        static_for<TL>([&](auto LT) {
          static_for<2>([&](auto X) {
            static_for<2>([&](auto sp) {
              static_for<2>([&](auto ri) {
                vld[LT][X][sp][ri] = *reinterpret_cast<const f32x4*>(
                    vp + (((LT * 2 + X) * 2 + sp) * 2 + ri) * 16);
              });
            });
          });
        });
      }
      // Stage 2 runs at raised wave priority.  The two waves of a SIMD leave
      // the segment barrier together and would otherwise keep step, their
      // MFMA-dense stage 2 and VALU/SALU-dense tails coinciding; the wave in
      // stage 2 now issues first and the pair drifts out of phase, so one
      // wave's MFMAs overlap the other's VALU work (4096^2 box 64: 123.4 ->
      // 114.0 ms per step, identical output; priority 3 the same, raising
      // it also over the inverse stages less)
      __builtin_amdgcn_s_setprio(1);
      // Loop order k-tile, image, l-tile: the box-row spectra of a k-tile and
      // image (read at the pixel's 2-byte offset and shifted into place, 4
      // VALU per fragment) serve both l-tiles, and image 0's S of both
      // l-tiles waits in registers for image 1's.  The l-tile-outer order
      // read them twice: 111.2 -> 107.0 ms per 4096^2 step, identical
      // output, 213 VGPRs.
      static_for<TL>([&](auto kt) {
        const int k = 16 * kt + n;
        CF4 s1[TL][2];  // image 0's S(l), S(l + Nh) of every l-tile
        static_for<2>([&](auto X) {
          // the even and odd column runs of this k-tile and image, each from
          // one address (the element offset, less one where it is odd, held
          // opaque in a scalar register): the 20 dword reads of a run are
          // immediate offsets from it, where a constant folded into the
          // address took an add per read pair beyond the 1 KB reach of the
          // paired-read offsets
          const int ofe = fresh_s((X * 2 + pe) * TP + 16 * kt * RK + be - (oe ? 1 : 0));
          const int ofo = fresh_s((X * 2 + po) * TP + 16 * kt * RK + bo - (oo ? 1 : 0));
          const unsigned* tre = reinterpret_cast<const unsigned*>(TTl + ofe);
          const unsigned* tro = reinterpret_cast<const unsigned*>(TTl + ofo);
          const unsigned she = oe ? 2u : 0u, sho = oo ? 2u : 0u;
          const bf16x8 eh = ld8_sh(tre + 0 * RW / 2, she), el = ld8_sh(tre + 1 * RW / 2, she);
          const bf16x8 euh = ld8_sh(tre + 2 * RW / 2, she), eul = ld8_sh(tre + 3 * RW / 2, she);
          const bf16x8 oh = ld8_sh(tro + 0 * RW / 2, sho), ol = ld8_sh(tro + 1 * RW / 2, sho);
          const bf16x8 ouh = ld8_sh(tro + 2 * RW / 2, sho), oul = ld8_sh(tro + 3 * RW / 2, sho);
          static_for<TL>([&](auto LT) {
            // the table fragments of this l-tile, reloaded per use rather
            // than held across the loops
            asm volatile("" ::: "memory");
            const DFr dE = dtab(C::kDe, LT), dO = dtab(C::kDo, LT);
            // stage-2 tile of image X: S(l), S(l + Nh) for the lane's rows.
            // The mean correction w V enters as the accumulators' start:
            // E' = E + w Ve, O' = O + w Vo, so E' + O' = E + O + w (Ve + Vo)
            // and E' - O' = E - O + w (Ve - Vo)
            const float2 w = wrh[X][kt];
            f32x4 er, ei, orr, oi;
            static_for<4>([&](auto t) {
              const float a = vld[LT][X][0][0][int(t)], b = vld[LT][X][0][1][int(t)];
              const float c = vld[LT][X][1][0][int(t)], d = vld[LT][X][1][1][int(t)];
              er[int(t)] = w.x * a - w.y * b;
              ei[int(t)] = w.x * b + w.y * a;
              orr[int(t)] = w.x * c - w.y * d;
              oi[int(t)] = w.x * d + w.y * c;
            });
            cmul_mfma(dE, eh, el, euh, eul, er, ei);
            cmul_mfma(dO, oh, ol, ouh, oul, orr, oi);
            CF4 sp, sm;
            static_for<4>([&](auto t) {
              sp.re[int(t)] = er[int(t)] + orr[int(t)];
              sp.im[int(t)] = ei[int(t)] + oi[int(t)];
              sm.re[int(t)] = er[int(t)] - orr[int(t)];
              sm.im[int(t)] = ei[int(t)] - oi[int(t)];
            });
            if constexpr (kt == 0) {
              // the packed DC / Nyquist column to the exchange buffer: a
              // group of four rows below Nh as one vector store per part, a
              // group across Nh row by row
              if (n == 0) {
                const int l0 = 16 * LT + 4 * g;
                float* zr = zx + 128 * X;
                if (l0 + 3 < Nh) {
                  *reinterpret_cast<f32x4*>(zr + l0) = sp.re;
                  *reinterpret_cast<f32x4*>(zr + 64 + l0) = sp.im;
                  *reinterpret_cast<f32x4*>(zr + 32 + l0) = sm.re;
                  *reinterpret_cast<f32x4*>(zr + 96 + l0) = sm.im;
                } else {
                  static_for<4>([&](auto t) {
                    const int l = l0 + t;
                    if (l < Nh) {
                      zr[l] = sp.re[int(t)];
                      zr[64 + l] = sp.im[int(t)];
                      zr[32 + l] = sm.re[int(t)];
                      zr[96 + l] = sm.im[int(t)];
                    }
                  });
                }
              }
            }
            if constexpr (X == 0) {
              s1[LT][0] = sp;
              s1[LT][1] = sm;
            } else {
              // cross power conj(S1) S2
              static_for<4>([&](auto t) {
                {
                  const float ar = s1[LT][0].re[int(t)], ai = s1[LT][0].im[int(t)];
                  const float br = sp.re[int(t)], bi = sp.im[int(t)];
                  Pp[kt][LT].re[int(t)] = ar * br + ai * bi;
                  Pp[kt][LT].im[int(t)] = ar * bi - ai * br;
                }
                {
                  const float ar = s1[LT][1].re[int(t)], ai = s1[LT][1].im[int(t)];
                  const float br = sm.re[int(t)], bi = sm.im[int(t)];
                  Pm[kt][LT].re[int(t)] = ar * br + ai * bi;
                  Pm[kt][LT].im[int(t)] = ar * bi - ai * br;
                }
              });
              if (a.filt) {
                // filter^2 on the cross power (flct.c:1445-1457)
                static_for<4>([&](auto t) {
                  const int l = 16 * LT + 4 * g + t;
                  if (k < Nh && l < Nh) {
                    const float fp = (float)a.filt[k * N + l], fm = (float)a.filt[k * N + l + Nh];
                    Pp[kt][LT].re[int(t)] *= fp * fp;
                    Pp[kt][LT].im[int(t)] *= fp * fp;
                    Pm[kt][LT].re[int(t)] *= fm * fm;
                    Pm[kt][LT].im[int(t)] *= fm * fm;
                  }
                });
              }
            }
          });
        });
      });

      __builtin_amdgcn_s_setprio(0);
      // packed columns of both images, unpacked through the partner -l:
      // A0 = (z + conj z(-l)) / 2, A_nyq = (z - conj z(-l)) / (2i)
      wave_lds_fence();
      if (lane < N) {
        const int lp = (lane == 0) ? 0 : N - lane;
        // rows l and -l in the planar layout (l >= Nh at 32 + l - Nh)
        const int il = (lane < Nh) ? lane : lane - Nh + 32;
        const int ip = (lp < Nh) ? lp : lp - Nh + 32;
        const float2 z1 = make_float2(zx[il], zx[64 + il]), q1 = make_float2(zx[ip], zx[64 + ip]);
        const float2 z2 = make_float2(zx[128 + il], zx[192 + il]);
        const float2 q2 = make_float2(zx[128 + ip], zx[192 + ip]);
        const float a0r1 = 0.5f * (z1.x + q1.x), a0i1 = 0.5f * (z1.y - q1.y);
        const float anr1 = 0.5f * (z1.y + q1.y), ani1 = -0.5f * (z1.x - q1.x);
        const float a0r2 = 0.5f * (z2.x + q2.x), a0i2 = 0.5f * (z2.y - q2.y);
        const float anr2 = 0.5f * (z2.y + q2.y), ani2 = -0.5f * (z2.x - q2.x);
        float n0f = 1.f, nnf = 1.f;
        if (a.filt) {
          const float f0 = (float)a.filt[lane], fn = (float)a.filt[Nh * N + lane];
          n0f = f0 * f0;
          nnf = fn * fn;
        }
        const float p0r = (a0r1 * a0r2 + a0i1 * a0i2) * n0f;
        const float p0i = (a0r1 * a0i2 - a0i1 * a0r2) * n0f;
        const float pnr = (anr1 * anr2 + ani1 * ani2) * nnf;
        const float pni = (anr1 * ani2 - ani1 * anr2) * nnf;
        zp[il] = p0r - pni;
        zp[64 + il] = p0i + pnr;
      }
      wave_lds_fence();
      if (n == 0) {
        static_for<TL>([&](auto LT) {
          const int l0 = 16 * LT + 4 * g;
          if (l0 + 3 < Nh) {
            // a whole group of four rows: one vector load per part
            Pp[0][LT].re = *reinterpret_cast<const f32x4*>(zp + l0);
            Pp[0][LT].im = *reinterpret_cast<const f32x4*>(zp + 64 + l0);
            Pm[0][LT].re = *reinterpret_cast<const f32x4*>(zp + 32 + l0);
            Pm[0][LT].im = *reinterpret_cast<const f32x4*>(zp + 96 + l0);
          } else {
            static_for<4>([&](auto t) {
              const int l = l0 + t;
              if (l < Nh) {
                Pp[0][LT].re[int(t)] = zp[l];
                Pp[0][LT].im[int(t)] = zp[64 + l];
                Pm[0][LT].re[int(t)] = zp[32 + l];
                Pm[0][LT].im[int(t)] = zp[96 + l];
              }
            });
          }
        });
      }
      wave_lds_fence();

      // ---------------- inverse 1: A = Q_q (row k, contraction over l)
      unsigned Uh[2][2][2][4], Ul[2][2][2][4];  // [kt][NT][q][t']
      static_for<TL>([&](auto kt) {
        static_for<2>([&](auto q) {
          u32x4 rh, rl, ih, il;
          static_for<2>([&](auto LT) {
            static_for<2>([&](auto th) {
              float vr[2], vi[2];
              static_for<2>([&](auto b) {
                constexpr int t = 2 * decltype(th)::value + decltype(b)::value;
                if constexpr (decltype(LT)::value < TL) {
                  const float pr = Pp[kt][LT].re[t], pi = Pp[kt][LT].im[t];
                  const float mr = Pm[kt][LT].re[t], mi = Pm[kt][LT].im[t];
                  vr[int(b)] = (q == 0) ? pr + mr : pr - mr;
                  vi[int(b)] = (q == 0) ? pi + mi : pi - mi;
                } else {
                  vr[int(b)] = 0.f;  // rows l >= 16 do not exist for boxes up to 32
                  vi[int(b)] = 0.f;
                }
              });
              unsigned h0, l0, h1, l1;
              split2(vr[0], vr[1], h0, l0, sk);
              split2(vi[0], vi[1], h1, l1, sk);
              rh[2 * LT + th] = h0;
              rl[2 * LT + th] = l0;
              ih[2 * LT + th] = h1;
              il[2 * LT + th] = l1;
            });
          });
          const bf16x8 qrh = u4_to_bf8(rh), qrl = u4_to_bf8(rl);
          const bf16x8 qih = u4_to_bf8(ih), qil = u4_to_bf8(il);
          const bf16x8 nqih = neg8(qih), nqil = neg8(qil);
          static_for<TL>([&](auto NT) {
            const __bf16* B = TBl + ((q == 0) ? C::kBI0 : C::kBI1) + 16 * NT * RS;
            const bf16x8 brh = ld8(B + 0 * PL), brl = ld8(B + 1 * PL);
            const bf16x8 bih = ld8(B + 2 * PL), bil = ld8(B + 3 * PL);
            f32x4 ur = mfma3(qrh, qrl, brh, brl, z4);
            ur = mfma3(nqih, nqil, bih, bil, ur);
            f32x4 ui = mfma3(qrh, qrl, bih, bil, z4);
            ui = mfma3(qih, qil, brh, brl, ui);
            static_for<4>([&](auto t) {
              unsigned hh, ll;
              split2(ur[int(t)], ui[int(t)], hh, ll, sk);
              Uh[kt][NT][q][int(t)] = hh;
              Ul[kt][NT][q][int(t)] = ll;
            });
          });
        });
      });

      // ---------------- inverse 2: c over (NT, q) column tiles of jj
      f32x4 cv[2][2][2][2];  // [NT][q][IT][pm]
      static_for<TL>([&](auto NT) {
        static_for<2>([&](auto q) {
          bf16x8 ah[2], al[2];
          static_for<2>([&](auto s) {
            u32x4 vh, vl;
            vh[0] = Uh[0][NT][q][s];
            vh[1] = Uh[0][NT][q][2 + s];
            vl[0] = Ul[0][NT][q][s];
            vl[1] = Ul[0][NT][q][2 + s];
            if constexpr (TL > 1) {
              vh[2] = Uh[1][NT][q][s];
              vh[3] = Uh[1][NT][q][2 + s];
              vl[2] = Ul[1][NT][q][s];
              vl[3] = Ul[1][NT][q][2 + s];
            } else {
              vh[2] = vh[3] = vl[2] = vl[3] = 0u;  // no k >= 16 for boxes up to 32
            }
            ah[s] = u4_to_bf8(vh);
            al[s] = u4_to_bf8(vl);
          });
          static_for<TL>([&](auto IT) {
            const __bf16* W = TBl + C::kWT + 16 * IT * RS;
            const f32x4 ce = mfma3(ah[0], al[0], ld8(W + 0 * PL), ld8(W + 1 * PL), z4);
            const f32x4 co = mfma3(ah[1], al[1], ld8(W + 2 * PL), ld8(W + 3 * PL), z4);
            cv[NT][q][IT][0] = ce + co;
            cv[NT][q][IT][1] = ce - co;
          });
        });
      });

      // ---------------- argmax on |c|, first strict maximum in the scan order
      // cell (NT, q, IT, t, pm) of lane (n, g): ii = 16 IT + n, x = ii + pm Nh,
      // jj = 32 NT + 8 g + 2 t + q; in fftshift order (flct.c:1490-1510) it
      // sits at row sx = (x + Nh) mod N, column sy = (jj + Nh) mod N.
      float lmax = 0.f;
      static_for<TL>([&](auto NT) {
        static_for<2>([&](auto q) {
          static_for<TL>([&](auto IT) {
            static_for<2>([&](auto pm) {
              const f32x4 v = cv[NT][q][IT][pm];
              lmax = fmaxf(lmax, fmaxf(fmaxf(fabsf(v[0]), fabsf(v[1])), fmaxf(fabsf(v[2]), fabsf(v[3]))));
            });
          });
        });
      });
      // the same in every lane: held as a wave-uniform (scalar) value, so the
      // peak location and the fit below are scalar
      const float gmaxf = __builtin_bit_cast(
          float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, wave_max(lmax))));
      const float thr = gmaxf * (float)(1. - kTieTol);
      // cells within kTieTol of the maximum: exactly one (the maximum) settles
      // the peak, and its register slot and lane are read off the sums below
      // (for a single cell ssel is its slot and msel its lane bit); with more
      // the pixel is re-evaluated in FP64, which rewrites all its outputs, so
      // the cell these sums then name (masked into range) only has to be
      // deterministic.  A zero surface is the max == 0 case.  Register slot
      // s = (((NT 2 + q) TL + IT) 2 + pm) 4 + t, in the loop order below.
      // held as vectors of at most 32 registers, which the compiler indexes
      // in place for a wave-uniform slot (larger ones go through scratch)
      constexpr int NS = TL * 2 * TL * 2 * 4;
      constexpr int NH = (NS > 32) ? 32 : NS;
      typedef float slotv __attribute__((ext_vector_type(NH)));
      slotv cva[NS / NH];
      static_for<TL>([&](auto NT) {
        static_for<2>([&](auto q) {
          static_for<TL>([&](auto IT) {
            static_for<2>([&](auto pm) {
              static_for<4>([&](auto t) {
                constexpr int sl = (((NT * 2 + q) * TL + IT) * 2 + pm) * 4 + t;
                cva[sl / NH][sl % NH] = cv[NT][q][IT][pm][int(t)];
              });
            });
          });
        });
      });
      int nge = 0, ssel = 0;
      unsigned long long msel = 0ull;
      if (gmaxf > 0.f) {
        static_for<NS>([&](auto sl) {
          const unsigned long long m = __ballot(fabsf(cva[sl / NH][sl % NH]) >= thr);
          // branch-free: a zero mask leaves the sums as they are
          const int pc = __popcll(m);
          nge += pc;
          ssel += pc * int(sl);
          msel |= m;
          // summed here: left to the compiler, the sums sink to the end and
          // every mask stays live in scalar registers
          asm volatile("" : "+s"(nge), "+s"(ssel), "+s"(msel));
        });
      }
      const double gmax = (double)gmaxf;
      const bool nonzero = (gmaxf != 0.f);
      int ixmax = Nh, iymax = Nh;
      if (nonzero) {
        // cell (NT, q, IT, pm, t) of lane (n, g): ii = 16 IT + n, x = ii + pm
        // Nh, jj = 32 NT + 8 g + 2 t + q; in fftshift order (flct.c:1490-1510)
        // it sits at row sx = (x + Nh) mod N, column sy = (jj + Nh) mod N
        const int ln = __builtin_ctzll(msel | (1ull << 63));
        const int sl = ssel & (NS - 1);
        const int t = sl & 3, pm = (sl >> 2) & 1, IT = (sl >> 3) % TL;
        const int q = (sl / (8 * TL)) & 1, NT = sl / (16 * TL);
        const int ii = 16 * IT + (ln & 15);
        ixmax = (pm == 0) ? ii + Nh : ii;
        const int jj = 32 * NT + 8 * (ln >> 4) + 2 * t + q;
        iymax = (jj < Nh) ? jj + Nh : jj - Nh;
      }
      const bool fit = nonzero && ixmax > 0 && ixmax < N - 1 && iymax > 0 && iymax < N - 1;
      // the 3x3 neighbourhood, read from its lanes (wave-uniform cells)
      double nbv[9];
      static_for<9>([&](auto c) {
        nbv[int(c)] = 0.;
        if (fit) {
          const int sx = ixmax + int(c) / 3 - 1, sy = iymax + int(c) % 3 - 1;
          const int pm = (sx >= Nh) ? 0 : 1;
          const int ii = (sx >= Nh) ? sx - Nh : sx;
          const int jj = (sy < Nh) ? sy + Nh : sy - Nh;
          const int sl = ((((jj >> 5) * 2 + (jj & 1)) * TL + (ii >> 4)) * 2 + pm) * 4 + ((jj >> 1) & 3);
          const int src = (ii & 15) + 16 * ((jj >> 3) & 3);
          // a copy first: a bit cast of a vector element reads element 0
          // with this compiler, whatever the index
          float v = cva[0][sl % NH];
          if constexpr (NS > NH) {
            if (sl >= NH) v = cva[1][sl % NH];
          }
          const float r = __builtin_bit_cast(
              float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), src));
          nbv[int(c)] = (double)fabsf(r);
        }
      });
      bool flagged = nonzero && nge >= 2;
      double shiftsubx = 0., shiftsuby = 0.;
      if (fit) {
#pragma clang fp contract(off)
        const double* nb = nbv;
        const double c00 = nb[4];
        const double fx = 0.5 * (nb[7] - nb[1]);
        const double fy = 0.5 * (nb[5] - nb[3]);
        const double fxx = (nb[7] + nb[1] - 2. * (c00));
        const double fyy = (nb[5] + nb[3] - 2. * (c00));
        const double fxy = 0.25 * (nb[8] + nb[0] - nb[6] - nb[2]);
        shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
        shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
        const double det = fxx * fyy - fxy * fxy;
        const double hn = fabs(fxx) + fabs(fyy) + 2. * fabs(fxy);
        const double bound = kRelErr * gmax * hn * (1. + 4. * (fabs(shiftsubx) + fabs(shiftsuby)));
        if (!(bound <= kFitTol * fabs(det))) flagged = true;
      }
      if (lane == 0) {
        const double shiftx = ((double)ixmax + shiftsubx) - (double)Nh;
        const double shifty = ((double)iymax + shiftsuby) - (double)Nh;
        const long long pix =
            A.strip ? (long long)j * A.nyo + (A.cb0 + i) : (long long)i * ny + j;
        // in strip mode the surface's rows are image columns: shiftx is the
        // image's column shift
        a.vx[pix] = (A.strip ? shiftx : shifty) * a.deltinv * a.deltas;
        a.vy[pix] = (A.strip ? shifty : shiftx) * a.deltinv * a.deltas;
        a.vm[pix] = 1.;
        if (flagged && a.flag_list) a.flag_list[atomicAdd(a.flag_count, 1u)] = (int)pix;
      }
    }
  }
}

// Edge pixels for the per-pixel kernel: tracked pixels whose box is clamped
// at the left or right image edge (columns outside [jlo, jhi]) and whose row
// is outside [ilo, ihi] (rows inside it run through the transposed edge
// strips; an empty range, ilo > ihi, keeps every row).  With a tracked-pixel
// list (threshold or skip; its length in *nlist) the list is filtered into
// out / *count by a grid-stride loop; otherwise the edge pixels are
// enumerated.
__global__ void k_edge_list(const int* __restrict__ list, const unsigned int* __restrict__ nlist,
                            int nx, int ny, int jlo, int jhi, int ilo, int ihi, int* __restrict__ out,
                            unsigned int* __restrict__ count) {
  const int nskip = (ihi >= ilo) ? ihi - ilo + 1 : 0;  // rows left to the strips
  if (list) {
    const long long n = (long long)*nlist;
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long t0 = (long long)blockIdx.x * blockDim.x; t0 < n; t0 += stride) {
      const long long t = t0 + threadIdx.x;
      bool e = false;
      int p = 0;
      if (t < n) {
        p = list[t];
        const int r = p / ny;
        const int j = p - r * ny;
        e = ((j < jlo) || (j > jhi)) && !(nskip > 0 && r >= ilo && r <= ihi);
      }
      const unsigned long long mask = __ballot(e);
      const int ln = threadIdx.x & 63;
      unsigned int base = 0;
      if (ln == 0 && mask) base = atomicAdd(count, (unsigned)__popcll(mask));
      base = __shfl(base, 0);
      if (e) out[base + __popcll(mask & ((1ull << ln) - 1ull))] = p;
    }
  } else {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int we = jlo + (ny - 1 - jhi);
    if (t < (long long)(nx - nskip) * we) {
      const int rr = (int)(t / we), c = (int)(t - (long long)rr * we);
      const int r = (nskip > 0 && rr >= ilo) ? rr + nskip : rr;
      const int j = (c < jlo) ? c : jhi + 1 + (c - jlo);
      out[t] = r * ny + j;
    }
  }
}

}  // namespace flct
