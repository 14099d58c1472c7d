// Launch of the tracker kernel for one box size, shared by the translation
// units that instantiate the box-size ranges (built in parallel).
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/flct_hip.h"
#include "track_kernels.hpp"

namespace flct {

constexpr int kTrackMaxM = 24;  // box sizes N = 4..96

struct LaunchCtx {
  hipStream_t stream;
  int num_cu;
  size_t lds_per_block;
};

template <int M, class T>
int launch_track(const LaunchCtx& c, const TrackArgs& a) {
  using Cfg = TrackCfg<M, T>;
  const size_t lds = Cfg::kLdsBytes;
  if (lds > c.lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  if (lds > 65536)
    (void)hipFuncSetAttribute((const void*)k_track<M, T>,
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  long long per_cu = (long long)(160 * 1024) / (long long)lds;
  if (per_cu < 1) per_cu = 1;
  if (per_cu > 8) per_cu = 8;
  long long grid = (long long)c.num_cu * per_cu;
  const long long need = (a.nwork + Cfg::P - 1) / Cfg::P;
  if (grid > need) grid = need;
  if (grid < 1) grid = 1;
  hipLaunchKernelGGL((k_track<M, T>), dim3((unsigned)grid), dim3(Cfg::kThreads), lds, c.stream, a);
  return hipGetLastError() == hipSuccess ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
}

// dispatchers, one per translation unit: box sizes M = 1..12 and 13..24
int dispatch_track_f64_lo(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f64_hi(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f32_lo(int M, const LaunchCtx& c, const TrackArgs& a);
int dispatch_track_f32_hi(int M, const LaunchCtx& c, const TrackArgs& a);
// split-BF16 MFMA form, box size N = 4..96
int dispatch_track_bf16(int N, const LaunchCtx& c, const TrackArgs& a);

}  // namespace flct
