// Launch of the tracker kernel for one box size, shared by the translation
// units that instantiate the box-size ranges (built in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/flct_hip.h"
#include "track_kernels.hpp"

namespace flct {

constexpr int kTrackMaxM = 24;  // box sizes N = 4..96

struct LaunchCtx {
  hipStream_t stream;
  int num_cu;
  size_t lds_per_block;
};

// Blocks hold a.slots pixels (1..P, 0: the compiled count P); LDS and
// threads per block follow the slot count, and a CU takes as many blocks as
// its 160 KB of LDS and 2048 threads hold (at most 8).
template <int M, class T, bool HIRES = false>
int launch_track(const LaunchCtx& c, const TrackArgs& a) {
  using Cfg = TrackCfg<M, T>;
  const int sp = run_slots(a.slots, Cfg::P);
  const size_t lds = (size_t)sp * Cfg::kSlotBytes + 2 * Cfg::N * 8;
  const int threads = (sp * Cfg::kSlotThreads + 63) / 64 * 64;
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track<M, T, HIRES>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  long long per_cu = (long long)(160 * 1024) / (long long)lds;
  if (per_cu > 2048 / threads) per_cu = 2048 / threads;
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 8) per_cu = 8;
  long long grid = (long long)c.num_cu * per_cu;
  const long long need = (a.nwork + sp - 1) / sp;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_track<M, T, HIRES>), dim3((unsigned)grid), dim3(threads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

// compiled slots per block of k_track for box size 4*M (0: no such size)
template <class T>
int track_slots(int M) {
  int p = 0;
  static_for<kTrackMaxM>([&](auto I) {
    constexpr int m = I + 1;
    if (m == M) p = TrackCfg<m, T>::P;
  });
  return p;
}

// dispatchers, one per translation unit: box sizes M = 1..12 and 13..24
int dispatch_track_f64_lo(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f64_hi(int M, const LaunchCtx& c, const TrackArgs& a);
// double with the -h peak step (a.hires), M = 1..12 and 13..24
int dispatch_track_f64h_lo(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f64h_hi(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f32_lo(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f32_hi(int M, const LaunchCtx& c, const TrackArgs& a);
// split-BF16 MFMA form, box size N = 4..96
int dispatch_track_bf16(int N, const LaunchCtx& c, const TrackArgs& a);

// split-BF16 row-segment form: boxes N = 8..64 with N % 4 == 0, interior
// columns, no bias pass.  seg_tables fills the split DFT tables for box N and
// window width sigma (bf16 bits, copied to the device by the caller).
// launch_edge_list lists the tracked pixels of the clamped edge columns: from
// a tracked-pixel list of device length *nlist into out / *count (zero on
// entry; *nedge = -1, the count stays on the device), or every edge pixel
// when list is null (*nedge = their number).  corners_only: only those in
// clamped edge rows too (the others run through the transposed strips).
bool seg_supported(int N, int nx, int ny, int bias);
void seg_tables(int N, double sigma, std::vector<unsigned short>& tab);
int launch_edge_list(const LaunchCtx& c, const int* list, const unsigned int* nlist, int nx, int ny,
                     int nsize, int* out, unsigned int* count, long long* nedge, bool corners_only);
// segpx: pixels per row segment, 8..64 (0: 64)
// k_track_seg scratch per block: the mean-correction transforms of one
// segment, [64 pixels][2 l-tiles][2 images][2 column parities][re, im][16] floats
constexpr long long kSegVBytes = 64LL * 2 * 2 * 2 * 2 * 16 * 4;
// Strip mode of k_track_seg (SegArgs::strip): rows [rlo, rlo + nrows) of a
// transposed edge strip whose row t is image column cb0 + t; nyo is the
// image's ny.
struct SegStrip {
  int rlo, nrows, cb0, nyo;
};
int launch_track_seg(const LaunchCtx& c, const TrackArgs& a, const void* tab, double thresh,
                     int thresh_rel, int skip, int xoffset, int yoffset, int segpx, float* vscr,
                     const SegStrip* st = nullptr);

}  // namespace flct
