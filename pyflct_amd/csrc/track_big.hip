// FP64 tracker for boxes N > 96 (sigma >= 24): the reference bounds the box
// only by the image size (flct.c:567-580), and above 96 neither the quad-FFT
// kernels (registers) nor an LDS-resident N x N spectrum (160 KB) fit.
//
// One block tracks one pixel at a time and keeps the packed spectrum of that
// pixel in a block-private plane P (N x N complex) in device memory; the
// transforms are batches of R rows (or columns) staged through LDS, radix-2
// for power-of-two N and Bluestein's chirp otherwise (the same scheme and
// tables as the row transforms of fft_rows.hpp).  Per pixel and pass:
//   A  rows of z = g1 + i g2 (window, box mean, flct.c:815-870), forward;
//   B  columns, forward (P = FFT2(z));
//   C  column tasks closed under l -> -l: the spectra of the two images are
//      A(k) = (Z(k) + conj Z(-k)) / 2, B(k) = (Z(k) - conj Z(-k)) / 2i; the
//      cross power conj(A) B (1/N^2)^2 with the optional filter on both
//      spectra (flct.c:1440-1477) is Hermitian, so one element of each
//      orbit {k, -k} is formed and its mirror is its conjugate; inverse
//      column transforms;
//   D  inverse row transforms; |.| moved by N/2 in both axes (flct.c:1485-
//      1510, 2111-2140) into the real plane Q, with the first strict maximum
//      in row-major order of the moved array (flct.c:2142-2161) as an LDS
//      atomicMax of the value bits and atomicMin of the index;
// then the 3 x 3 Taylor fit (flct.c:1536-1556) or the -h grid step
// (flct.c:1564-1645), and the -b second pass with re-centred windows.
#include <hip/hip_runtime.h>

#include "../../include/flct_hip.h"
#include "interp_cc.hpp"
#include "track_big.hpp"

namespace flct {

namespace {

constexpr int kBigThreads = 256;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// R rows of length M = 2^logm at x[r * M + t], in place
__device__ void rows_fft_dif(double2* x, int R, int logm, const double2* __restrict__ tw, int sign) {
  const int L = 1 << logm, H = L >> 1;
  for (int ls = logm - 1; ls >= 0; --ls) {
    const int s = 1 << ls;
    for (int u = threadIdx.x; u < R * H; u += blockDim.x) {
      const int r = u >> (logm - 1), t = u & (H - 1);
      const int j = t & (s - 1);
      const int i0 = r * L + ((t >> ls) << (ls + 1)) + j, i1 = i0 + s;
      double2 w = tw[j << (logm - 1 - ls)];
      if (sign < 0) w.y = -w.y;
      const double2 a = x[i0], b = x[i1];
      x[i0] = make_double2(a.x + b.x, a.y + b.y);
      x[i1] = cmul(make_double2(a.x - b.x, a.y - b.y), w);
    }
    __syncthreads();
  }
}

__device__ void rows_fft_dit(double2* x, int R, int logm, const double2* __restrict__ tw, int sign) {
  const int L = 1 << logm, H = L >> 1;
  for (int ls = 0; ls < logm; ++ls) {
    const int s = 1 << ls;
    for (int u = threadIdx.x; u < R * H; u += blockDim.x) {
      const int r = u >> (logm - 1), t = u & (H - 1);
      const int j = t & (s - 1);
      const int i0 = r * L + ((t >> ls) << (ls + 1)) + j, i1 = i0 + s;
      double2 w = tw[j << (logm - 1 - ls)];
      if (sign < 0) w.y = -w.y;
      const double2 a = x[i0], b = cmul(x[i1], w);
      x[i0] = make_double2(a.x + b.x, a.y + b.y);
      x[i1] = make_double2(a.x - b.x, a.y - b.y);
    }
    __syncthreads();
  }
}

// Transform of R staged rows (inputs already multiplied by the chirp for
// Bluestein, zero beyond n).  Output element k of row r is then
// x[r * M + out_index(k)], times chirp[k] for Bluestein.
__device__ void rows_dft(double2* x, int R, const BigTab& t) {
  if (!t.bluestein) {
    rows_fft_dif(x, R, t.logm, t.tw, t.sign);
    return;
  }
  const int M = 1 << t.logm;
  rows_fft_dif(x, R, t.logm, t.tw, -1);
  for (int u = threadIdx.x; u < R * M; u += blockDim.x) x[u] = cmul(x[u], t.bhat[u & (M - 1)]);
  __syncthreads();
  rows_fft_dit(x, R, t.logm, t.tw, +1);
}

__device__ __forceinline__ double2 row_out(const double2* x, int r, int k, const BigTab& t) {
  const int M = 1 << t.logm;
  if (!t.bluestein) return x[r * M + (int)(__builtin_bitreverse32((unsigned)k) >> (32 - t.logm))];
  return cmul(x[r * M + k], t.chirp[k]);
}

__device__ __forceinline__ double2 row_in(double2 v, int k, const BigTab& t) {
  return t.bluestein ? cmul(v, t.chirp[k]) : v;
}

}  // namespace

__global__ void __launch_bounds__(kBigThreads) k_track_big(BigArgs a) {
  extern __shared__ __align__(16) unsigned char lds[];
  const int N = a.N, R = a.R;
  const int M = 1 << a.fwd.logm;
  double2* X = reinterpret_cast<double2*>(lds);
  double* wx1 = reinterpret_cast<double*>(X + (size_t)R * M);
  double* wy1 = wx1 + N;
  double* wx2 = wy1 + N;
  double* wy2 = wx2 + N;
  double* nb = wy2 + N;                 // 9
  double* res = nb + 9;                 // 2
  double* hw = res + 2;                 // 36 (-h window)
  double* hc = hw + kHiresWin * kHiresWin;  // 36 (-h cdata block)
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(hc + kHiresWin * kHiresWin);
  unsigned long long* hkey = amax + 1;
  unsigned int* aidx = reinterpret_cast<unsigned int*>(hkey + 1);
  unsigned int* hloc = aidx + 1;

  double2* P = a.plane + (size_t)blockIdx.x * N * N;
  double* Q = a.surf + (size_t)blockIdx.x * N * N;
  const int nx = a.nx, ny = a.ny, nsize = a.nsize, ns = ny - N + 1;
  const double norm = (1. / ((double)N * N)) * (1. / ((double)N * N));
  const long long count = a.nwork_dev ? (long long)*a.nwork_dev : a.nwork;
  const int tid = threadIdx.x, nt = blockDim.x;

  for (long long w = blockIdx.x; w < count; w += gridDim.x) {
    const long long p = a.list ? (long long)a.list[w] : w;
    const int i = (int)(p / ny), j = (int)(p - (p / ny) * ny);
    // box clamp, flct.c:773-786
    const int imin0 = (0 > (i - (nsize - 1))) ? 0 : i - (nsize - 1);
    const int imax0 = ((nx - 1) < (i + nsize)) ? nx - 1 : i + nsize;
    const int imin = (imax0 == nx - 1) ? nx - 1 - (2 * nsize - 1) : imin0;
    const int jmin0 = (0 > (j - (nsize - 1))) ? 0 : j - (nsize - 1);
    const int jmax0 = ((ny - 1) < (j + nsize)) ? ny - 1 : j + nsize;
    const int jmin = (jmax0 == ny - 1) ? ny - 1 - (2 * nsize - 1) : jmin0;
    const double f1bar = a.mean1[(long long)imin * ns + jmin];
    const double f2bar = a.mean2[(long long)imin * ns + jmin];
    double shiftx = 0., shifty = 0.;
    const int npass = a.bias ? 2 : 1;
    for (int pass = 0; pass < npass; ++pass) {
      // windows: the Gaussian centred on the pixel (flct.c:861-862), or for
      // the -b pass centred at pixel -+ d/2 (image 1 / image 2)
      const double cx = 0.5 * shiftx, cy = 0.5 * shifty;
      for (int x = tid; x < N; x += nt) {
        if (pass == 0) {
          wx1[x] = wx2[x] = a.gw[imin + x - i + N - 1];
          wy1[x] = wy2[x] = a.gw[jmin + x - j + N - 1];
        } else {
          const double di = (double)(imin + x - i), dj = (double)(jmin + x - j);
          const double a1 = a.sigminv * (di + cx), a2 = a.sigminv * (di - cx);
          const double b1 = a.sigminv * (dj + cy), b2 = a.sigminv * (dj - cy);
          wx1[x] = exp(-a1 * a1);
          wx2[x] = exp(-a2 * a2);
          wy1[x] = exp(-b1 * b1);
          wy2[x] = exp(-b2 * b2);
        }
      }
      __syncthreads();
      // A: rows of z, forward
      for (int r0 = 0; r0 < N; r0 += R) {
        const int nr = (N - r0 < R) ? N - r0 : R;
        for (int u = tid; u < nr * M; u += nt) {
          const int r = u / M, t = u - r * M;
          double2 v = make_double2(0., 0.);
          if (t < N) {
            const long long e = (long long)(imin + r0 + r) * ny + (jmin + t);
            const double g1 = (wx1[r0 + r] * wy1[t]) * (a.f1[e] - f1bar);
            const double g2 = (wx2[r0 + r] * wy2[t]) * (a.f2[e] - f2bar);
            v = row_in(make_double2(g1, g2), t, a.fwd);
          }
          X[u] = v;
        }
        __syncthreads();
        rows_dft(X, nr, a.fwd);
        for (int u = tid; u < nr * N; u += nt) {
          const int r = u / N, k = u - r * N;
          P[(size_t)(r0 + r) * N + k] = row_out(X, r, k, a.fwd);
        }
        __syncthreads();
      }
      __threadfence_block();
      __syncthreads();
      // B: columns, forward; staged with the column index fastest so that
      // the R adjacent columns of one plane row are read together
      for (int c0 = 0; c0 < N; c0 += R) {
        const int nc = (N - c0 < R) ? N - c0 : R;
        for (int u = tid; u < nc * M; u += nt) {
          const int t = u / nc, c = u - t * nc;
          if (t < M) {
            double2 v = make_double2(0., 0.);
            if (t < N) v = row_in(P[(size_t)t * N + c0 + c], t, a.fwd);
            X[c * M + t] = v;
          }
        }
        __syncthreads();
        rows_dft(X, nc, a.fwd);
        for (int u = tid; u < nc * N; u += nt) {
          const int k = u / nc, c = u - k * nc;
          P[(size_t)k * N + c0 + c] = row_out(X, c, k, a.fwd);
        }
        __syncthreads();
      }
      __threadfence_block();
      __syncthreads();
      // C: column tasks {0, N/2} and pairs (l, N - l), cross power, inverse
      const int half = R / 2;
      const int ntask = 1 + (N / 2 - 1 + half - 1) / half;
      for (int task = 0; task < ntask; ++task) {
        // columns of the task: list slot q holds column col(q); its mirror
        // N - col (mod N) sits in slot mq(q)
        int nc, l0 = 0, h = 0;
        if (task == 0) {
          nc = 2;
        } else {
          l0 = 1 + (task - 1) * half;
          h = (N / 2 - l0 < half) ? N / 2 - l0 : half;
          nc = 2 * h;
        }
        auto col = [&](int q) {
          if (task == 0) return q == 0 ? 0 : N / 2;
          return q < h ? l0 + q : N - l0 - (q - h);
        };
        auto mq = [&](int q) {
          if (task == 0) return q;
          return q < h ? q + h : q - h;
        };
        for (int u = tid; u < nc * N; u += nt) {
          const int t = u / nc, q = u - t * nc;
          X[q * M + t] = P[(size_t)t * N + col(q)];
        }
        __syncthreads();
        // one thread per orbit {(q, k), (mq, -k)}
        for (int u = tid; u < nc * N; u += nt) {
          const int q = u / N, k = u - q * N;
          const int m = mq(q), km = (N - k) % N;
          if (m * N + km < q * N + k) continue;
          const double2 z = X[q * M + k], zm = X[m * M + km];
          // A = (z + conj zm) / 2, B = (z - conj zm) / 2i
          double ar = 0.5 * (z.x + zm.x), ai = 0.5 * (z.y - zm.y);
          double br = 0.5 * (z.y + zm.y), bi = -0.5 * (z.x - zm.x);
          if (a.filt) {
            const int kr = (k <= N / 2) ? k : N - k;
            const double fv = a.filt[kr * N + col(q)];
            ar *= fv;
            ai *= fv;
            br *= fv;
            bi *= fv;
          }
          const double cr = (ar * br + ai * bi) * norm, ci = (ar * bi - ai * br) * norm;
          X[q * M + k] = make_double2(cr, ci);
          X[m * M + km] = make_double2(cr, -ci);
        }
        __syncthreads();
        for (int u = tid; u < nc * M; u += nt) {
          const int q = u / M, t = u - q * M;
          X[u] = (t < N) ? row_in(X[u], t, a.inv) : make_double2(0., 0.);
        }
        __syncthreads();
        rows_dft(X, nc, a.inv);
        for (int u = tid; u < nc * N; u += nt) {
          const int k = u / nc, q = u - k * nc;
          P[(size_t)k * N + col(q)] = row_out(X, q, k, a.inv);
        }
        __syncthreads();
      }
      __threadfence_block();
      // D: rows, inverse; |.| moved by N/2 into Q; first strict maximum
      if (tid == 0) {
        *amax = 0ull;
        *aidx = 0xffffffffu;
      }
      __syncthreads();
      // per thread: its largest value and the smallest moved index holding it
      double lmax = 0.;
      unsigned int lidx = 0xffffffffu;
      for (int r0 = 0; r0 < N; r0 += R) {
        const int nr = (N - r0 < R) ? N - r0 : R;
        for (int u = tid; u < nr * M; u += nt) {
          const int r = u / M, t = u - r * M;
          X[u] = (t < N) ? row_in(P[(size_t)(r0 + r) * N + t], t, a.inv) : make_double2(0., 0.);
        }
        __syncthreads();
        rows_dft(X, nr, a.inv);
        for (int u = tid; u < nr * N; u += nt) {
          const int r = u / N, k = u - r * N;
          const double v = fabs(row_out(X, r, k, a.inv).x);
          const int si = (r0 + r + N / 2) % N, sj = (k + N / 2) % N;
          const unsigned int idx = (unsigned)(si * N + sj);
          Q[idx] = v;
          if (v > lmax) {
            lmax = v;
            lidx = idx;
          } else if (v == lmax && idx < lidx) {
            lidx = idx;
          }
        }
        __syncthreads();
      }
      atomicMax(amax, dbits(lmax));
      __threadfence_block();
      __syncthreads();
      const double gmax = __longlong_as_double((long long)*amax);
      if (lmax == gmax && lidx != 0xffffffffu) atomicMin(aidx, lidx);
      __syncthreads();
      const bool nonzero = (gmax != 0.);
      int ixmax = N / 2, iymax = N / 2;
      if (nonzero) {
        ixmax = (int)(*aidx / (unsigned)N);
        iymax = (int)(*aidx % (unsigned)N);
      }
      if (a.hires) {
        // 6 x 6 window of the surface about the integer peak, its cdata
        // block, the 21 x 21 grid at 0.1 px and its first strict maximum
        for (int t = tid; t < kHiresWin * kHiresWin; t += nt) {
          const int r = ixmax - 2 + t / kHiresWin, c = iymax - 2 + t % kHiresWin;
          hw[t] = (r >= 0 && r < N && c >= 0 && c < N) ? Q[(size_t)r * N + c] : 0.;
        }
        if (tid == 0) {
          *hkey = 0ull;
          *hloc = 0xffffffffu;
        }
        __syncthreads();
        for (int t = tid; t < kHiresWin * kHiresWin; t += nt)
          hires_cdata_entry(HiresWindow{hw, ixmax, iymax}, N, N, t, hc);
        __syncthreads();
        unsigned long long kbest = 0ull;
        int qbest = 0x7fffffff;
        for (int q = tid; q < kHiresN * kHiresN; q += nt) {
          const unsigned long long k = hires_key(hires_value_cp(hc, N, N, ixmax, iymax, q));
          if (k > kbest) {
            kbest = k;
            qbest = q;
          }
        }
        atomicMax(hkey, kbest);
        __syncthreads();
        if (kbest == *hkey) atomicMin(hloc, (unsigned)qbest);
        __syncthreads();
        const int hq = (int)*hloc;
        for (int t = tid; t < 9; t += nt) {
          const int ix = hq / kHiresN + t / 3 - 1, iy = hq % kHiresN + t % 3 - 1;
          nb[t] = (ix >= 0 && ix < kHiresN && iy >= 0 && iy < kHiresN)
                      ? hires_value_cp(hc, N, N, ixmax, iymax, ix * kHiresN + iy)
                      : 0.;
        }
        __syncthreads();
        if (tid == 0) {
          double shx, shy;
          hires_shift(hq, nb, ixmax, iymax, &shx, &shy);
          res[0] = shx - (double)(N / 2);
          res[1] = shy - (double)(N / 2);
        }
      } else {
        const bool fit = nonzero && ixmax > 0 && ixmax < N - 1 && iymax > 0 && iymax < N - 1;
        if (tid == 0) {
#pragma clang fp contract(off)
          double shiftsubx = 0., shiftsuby = 0.;
          if (fit) {
            for (int t = 0; t < 9; ++t)
              nb[t] = Q[(size_t)(ixmax + t / 3 - 1) * N + (iymax + t % 3 - 1)];
            const double c00 = nb[4];
            const double fx = 0.5 * (nb[7] - nb[1]);
            const double fy = 0.5 * (nb[5] - nb[3]);
            const double fxx = (nb[7] + nb[1] - 2. * (c00));
            const double fyy = (nb[5] + nb[3] - 2. * (c00));
            const double fxy = 0.25 * (nb[8] + nb[0] - nb[6] - nb[2]);
            shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
            shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
          }
          res[0] = ((double)ixmax + shiftsubx) - (double)(N / 2);
          res[1] = ((double)iymax + shiftsuby) - (double)(N / 2);
        }
      }
      __syncthreads();
      shiftx = res[0];
      shifty = res[1];
      __syncthreads();
    }
    if (tid == 0) {
      a.vx[p] = shifty * a.deltinv * a.deltas;  // transp = 1 (flct.c:901-905)
      a.vy[p] = shiftx * a.deltinv * a.deltas;
      a.vm[p] = 1.;
    }
  }
}

size_t big_lds_bytes(int N, int R, int logm) {
  return (size_t)R * ((size_t)1 << logm) * 16 + (size_t)(4 * N + 9 + 2 + 2 * kHiresWin * kHiresWin) * 8 +
         2 * 8 + 2 * 4;
}

int big_rows(int logm) {
  const int M = 1 << logm;
  int R = 65536 / (16 * M);
  if (R < 2) R = 2;
  if (R > 16) R = 16;
  return R & ~1;
}

long long big_grid(int num_cu, int N, long long nwork) {
  // two blocks per CU, at most 4 GB of planes (24 N^2 bytes per block)
  long long g = 2LL * num_cu;
  const long long cap = (4LL << 30) / (24LL * N * N);
  if (g > cap) g = cap;
  if (g < num_cu && cap >= num_cu) g = num_cu;
  if (g < 1) g = 1;
  if (nwork > 0 && g > nwork) g = nwork;
  return g;
}

int launch_track_big(const LaunchCtx& c, const BigArgs& a, long long grid) {
  const size_t lds = big_lds_bytes(a.N, a.R, a.fwd.logm);
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track_big, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
  hipLaunchKernelGGL(k_track_big, dim3((unsigned)grid), dim3(kBigThreads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

}  // namespace flct
