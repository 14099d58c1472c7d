// FP64 tracker for boxes N > 96: arguments and launch (track_big.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

#include "track_launch.hpp"

namespace flct {

// boxes above the quad-FFT kernels' 96, up to kBigMaxN (Bluestein length
// 2048: two staged rows of 32 KB each)
constexpr int kBigMaxN = 1024;

// Tables of one length-N transform direction (fft_rows.hpp layout):
// twiddles e^{+2 pi i e / M} (M/2), chirp (N) and its transform (M,
// bit-reversed, / M) for Bluestein; M = N for power-of-two N.
struct BigTab {
  int logm = 0, bluestein = 0, sign = -1;
  const double2* tw = nullptr;
  const double2* chirp = nullptr;
  const double2* bhat = nullptr;
};

struct BigArgs {
  const double* f1 = nullptr;
  const double* f2 = nullptr;
  int nx = 0, ny = 0, nsize = 0, N = 0;
  int R = 2;  // rows (columns) staged in LDS per batch
  double sigminv = 0., deltinv = 0., deltas = 0.;
  const double* gw = nullptr;     // gw[d + N - 1] = exp(-(sigminv d)^2)
  const double* mean1 = nullptr;  // box means by origin, (nx-N+1) x (ny-N+1)
  const double* mean2 = nullptr;
  const double* filt = nullptr;   // (N/2+1) x N filter, or null
  const int* list = nullptr;      // tracked pixels, or null for all
  long long nwork = 0;
  const unsigned int* nwork_dev = nullptr;
  int bias = 0, hires = 0;
  double* vx = nullptr;
  double* vy = nullptr;
  double* vm = nullptr;
  double2* plane = nullptr;  // grid x N x N complex
  double* surf = nullptr;    // grid x N x N real (the moved |c|)
  BigTab fwd, inv;
};

size_t big_lds_bytes(int N, int R, int logm);
int big_rows(int logm);
// blocks of the persistent grid; the planes take 24 N^2 bytes per block
long long big_grid(int num_cu, int N, long long nwork);
int launch_track_big(const LaunchCtx& c, const BigArgs& a, long long grid);

}  // namespace flct
