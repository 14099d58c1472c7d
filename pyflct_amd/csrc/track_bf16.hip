// Split-BF16 MFMA tracker launches.  Boxes N = 8..64 with N % 4 == 0 use the
// row-segment kernel for interior columns (no bias pass) and a per-pixel
// kernel for the edge columns: the radix-2 kernel for 36..64, the dense DFT
// kernel padded to NP = 32, 64 or 96 otherwise.
#include <cmath>
#include <cstring>

#include "track_launch.hpp"
#include "track_mfma.hpp"
#include "track_mfma2.hpp"
#include "track_seg.hpp"

namespace flct {

static int launch_grid(const LaunchCtx& c, size_t lds, int pix, long long nwork, long long max_per_cu,
                       long long& grid) {
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  long long per_cu = (long long)(160 * 1024) / (long long)lds;
  if (per_cu < 1) per_cu = 1;
  if (per_cu > max_per_cu) per_cu = max_per_cu;
  grid = (long long)c.num_cu * per_cu;
  const long long need = (nwork + pix - 1) / pix;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  return FLCT_HIP_OK;
}

template <int NP>
static int launch_mfma(const LaunchCtx& c, const TrackArgs& a) {
  using Cfg = MfmaCfg<NP>;
  const size_t lds = Cfg::kLdsBytes;
  long long grid = 0;
  if (int rc = launch_grid(c, lds, Cfg::PIX, a.nwork, 4, grid)) return rc;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track_mfma<NP>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_track_mfma<NP>), dim3((unsigned)grid), dim3(Cfg::kThreads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

template <int PIX>
static int launch_r2(const LaunchCtx& c, const TrackArgs& a) {
  using Cfg = R2Cfg<PIX>;
  const size_t lds = Cfg::kLdsBytes;
  long long grid = 0;
  if (int rc = launch_grid(c, lds, PIX, a.nwork, 4, grid)) return rc;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track_r2<PIX>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_track_r2<PIX>), dim3((unsigned)grid), dim3(Cfg::kThreads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

int dispatch_track_bf16(int N, const LaunchCtx& c, const TrackArgs& a) {
  if (N >= 36 && N <= 64 && N % 4 == 0) return launch_r2<2>(c, a);
  if (N <= 32) return launch_mfma<32>(c, a);
  if (N <= 64) return launch_mfma<64>(c, a);
  if (N <= 96) return launch_mfma<96>(c, a);
  return FLCT_HIP_E_UNSUPPORTED;
}

bool seg_supported(int N, int nx, int ny, int bias) {
  (void)nx;
  return !bias && N >= 8 && N <= 64 && N % 4 == 0 && ny - N + 1 >= 1;
}

// round to nearest even BF16
static unsigned short to_bf16(float f) {
  unsigned u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

static double from_bf16(unsigned short h) {
  const unsigned u = (unsigned)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return (double)f;
}

// hi = bf16(v), lo = bf16(v - hi)
static void put_split(std::vector<unsigned short>& t, int hi_idx, int lo_idx, double v) {
  const unsigned short h = to_bf16((float)v);
  t[hi_idx] = h;
  t[lo_idx] = to_bf16((float)(v - from_bf16(h)));
}

// e^{-2 pi i m / n} for integer m (reduced mod n before the angle is formed)
static void root(long long m, int n, double& re, double& im) {
  m %= n;
  if (m < 0) m += n;
  const double ang = 2.0 * M_PI * (double)m / (double)n;
  re = std::cos(ang);
  im = -std::sin(ang);
}

void seg_tables(int N, double sigma, std::vector<unsigned short>& tab) {
  using C = SegCfg;
  const int Nh = N / 2, nsize = N / 2;
  tab.assign(C::kGlobalEl, 0);
  const double sigminv = 1. / sigma;
  // column window of an interior box, as the window table of the host call
  std::vector<double> wc(N);
  for (int c = 0; c < N; ++c) {
    const double ax = sigminv * (double)(c - (nsize - 1));
    wc[c] = std::exp(-ax * ax);
  }
  // stage 1: D1[k][x], k = 0 packs DC (re) and Nyquist (im)
  for (int k = 0; k < 32; ++k)
    for (int x = 0; x < N; ++x) {
      double re = 0., im = 0.;
      if (k == 0) {
        re = 1.;
        im = (x & 1) ? -1. : 1.;
      } else if (k < Nh) {
        root((long long)k * x, N, re, im);
      } else {
        continue;
      }
      const int o = k * C::RS1 + x;
      put_split(tab, 0 * C::kD1Plane + o, 1 * C::kD1Plane + o, re);
      put_split(tab, 2 * C::kD1Plane + o, 3 * C::kD1Plane + o, im);
    }
  const int L = C::kLdsTabOff, PL = C::kPlane, RS = C::RS;
  // stage 2: De[l][c'] = Wc(2c') e^{-2 pi i l c'/Nh}, Do[l][c'] = Wc(2c'+1) e^{-2 pi i l (2c'+1)/N}
  for (int l = 0; l < Nh; ++l)
    for (int cp = 0; cp < Nh; ++cp) {
      double er, ei, orr, oi;
      root((long long)l * cp, Nh, er, ei);
      root((long long)l * (2 * cp + 1), N, orr, oi);
      const double we = wc[2 * cp], wo = wc[2 * cp + 1];
      const int o = l * RS + cp;
      const double v[2][3] = {{we * er, we * ei, -we * ei}, {wo * orr, wo * oi, -wo * oi}};
      for (int s = 0; s < 2; ++s) {
        const int base = L + (s == 0 ? C::kDe : C::kDo);
        for (int comp = 0; comp < 3; ++comp)
          put_split(tab, base + (2 * comp) * PL + o, base + (2 * comp + 1) * PL + o, v[s][comp]);
      }
    }
  // inverse 1: BI_q[m'][kk] = e^{+2 pi i l (2m' + q) / N}, l = l(kk)
  for (int mp = 0; mp < Nh; ++mp)
    for (int kk = 0; kk < 32; ++kk) {
      const int l = 16 * ((kk >> 2) & 1) + 4 * (kk >> 3) + (kk & 3);
      if (l >= Nh) continue;
      for (int q = 0; q < 2; ++q) {
        double re, im;
        root(-(long long)l * (2 * mp + q), N, re, im);
        const int base = L + (q == 0 ? C::kBI0 : C::kBI1), o = mp * RS + kk;
        put_split(tab, base + 0 * PL + o, base + 1 * PL + o, re);
        put_split(tab, base + 2 * PL + o, base + 3 * PL + o, im);
      }
    }
  // inverse 2: WT_s[ii][kk], kk = 8g + 4kt + 2th + part, k = 16kt + 4g + 2th + s
  for (int s = 0; s < 2; ++s)
    for (int ii = 0; ii < Nh; ++ii)
      for (int kk = 0; kk < 32; ++kk) {
        const int g = kk >> 3, kt = (kk >> 2) & 1, th = (kk >> 1) & 1, part = kk & 1;
        const int k = 16 * kt + 4 * g + 2 * th + s;
        double v = 0.;
        if (k == 0) {
          v = (part == 0) ? 1. : ((ii & 1) ? -1. : 1.);
        } else if (k < Nh) {
          double re, im;
          root(-(long long)k * ii, N, re, im);  // e^{+2 pi i k ii / N}
          v = (part == 0) ? 2. * re : -2. * im;
        }
        const int o = L + C::kWT + ii * RS + kk;
        put_split(tab, o + (2 * s) * PL, o + (2 * s + 1) * PL, v);
      }
}

int launch_edge_list(const LaunchCtx& c, const int* list, const unsigned int* nlist, int nx, int ny,
                     int nsize, int* out, unsigned int* count, long long* nedge, bool corners_only) {
  const int jlo = nsize - 1, jhi = ny - nsize - 1;
  // rows whose boxes are not clamped: their edge pixels run through the strips
  const int ilo = corners_only ? nsize - 1 : 1, ihi = corners_only ? nx - nsize - 1 : 0;
  const int nskip = (ihi >= ilo) ? ihi - ilo + 1 : 0;
  if (list) {
    // count must be zero on entry; the per-pixel kernel reads it on the device
    hipLaunchKernelGGL(k_edge_list, dim3(1024), dim3(256), 0, c.stream, list, nlist, nx, ny, jlo, jhi,
                       ilo, ihi, out, count);
    *nedge = -1;
  } else {
    const long long ne = (long long)(nx - nskip) * (jlo + (ny - 1 - jhi));
    if (ne > 0)
      hipLaunchKernelGGL(k_edge_list, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, c.stream,
                         (const int*)nullptr, (const unsigned int*)nullptr, nx, ny, jlo, jhi, ilo, ihi,
                         out, count);
    *nedge = ne;
  }
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

static_assert(SegCfg::kVFloats * 4LL == kSegVBytes, "k_track_seg scratch size");

int launch_track_seg(const LaunchCtx& c, const TrackArgs& a, const void* tab, double thresh,
                     int thresh_rel, int skip, int xoffset, int yoffset, int segpx, float* vscr,
                     const SegStrip* st) {
  using C = SegCfg;
  const int N = 2 * a.nsize;
  SegArgs s;
  s.t = a;
  s.tab = reinterpret_cast<const __bf16*>(tab);
  s.thresh = thresh;
  s.thresh_rel = thresh_rel;
  s.skip = skip;
  s.xoffset = xoffset;
  s.yoffset = yoffset;
  s.vscr = vscr;  // c.num_cu blocks of SegCfg::kVFloats
  if (!vscr) return FLCT_HIP_E_ARG;
  s.strip = st ? 1 : 0;
  s.rlo = st ? st->rlo : 0;
  s.nrows = st ? st->nrows : a.nx;
  s.cb0 = st ? st->cb0 : 0;
  s.nyo = st ? st->nyo : a.ny;
  if (s.nrows < 1) return FLCT_HIP_OK;
  s.jlo = a.nsize - 1;
  s.width = a.ny - N + 1;
  const int w = (segpx >= 8 && segpx < C::P) ? segpx : C::P;
  s.nband = (s.width + w - 1) / w;
  const long long nseg = (long long)s.nband * s.nrows;
  if (nseg > 0x7fffffffLL) return FLCT_HIP_E_UNSUPPORTED;
  s.nseg = (int)nseg;
  const size_t lds = C::kLdsBytes;
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  const void* kern = (N <= 32) ? (const void*)k_track_seg<1> : (const void*)k_track_seg<2>;
  (void)hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  long long grid = c.num_cu;
  if (grid > nseg) grid = nseg;
  if (grid < 1) return FLCT_HIP_OK;
  // XCD order when the bands split evenly over the 8 XCDs and each XCD's
  // blocks evenly over its bands (k_track_seg)
  s.xchunk = 0;
  if (!st && grid % 8 == 0 && s.nband % 8 == 0 && (grid / 8) % (s.nband / 8) == 0)
    s.xchunk = (int)((grid / 8) / (s.nband / 8));
  if (N <= 32)
    hipLaunchKernelGGL(k_track_seg<1>, dim3((unsigned)grid), dim3(C::kThreads), lds, c.stream, s);
  else
    hipLaunchKernelGGL(k_track_seg<2>, dim3((unsigned)grid), dim3(C::kThreads), lds, c.stream, s);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

}  // namespace flct
