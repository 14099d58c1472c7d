// FP64 tracker, row-shared form (track_rs.hip): arguments and launch.
#pragma once

#include <hip/hip_runtime.h>

#include "track_launch.hpp"

namespace flct {

constexpr int kRsSeg = 64;  // pixels per row segment, at most

struct RsArgs {
  TrackArgs t;  // images, box size, means, windows, filter, outputs
  double thresh = 0.;  // threshold (flct.c:721-722), 0: off
  int thresh_rel = 0;  // thresh is relative: times max(max |f1|, max |f2|) from absmax
  const unsigned long long* absmax = nullptr;
  int nband = 1;       // column bands of at most kRsSeg pixels
  int ustride = 0;     // row stride of Phi (complex), kRsSeg + N
  double2* phi = nullptr;  // per block: N x ustride complex
};

// blocks of the persistent grid for box size 4*M with `slots` pixels per
// block (0: the compiled count; 0 blocks: no such box); fp32: the row tasks
// in FP32 (path fp32-rowseg), else FP64
long long rs_grid(int M, int num_cu, long long nseg, int slots, bool fp32);
// compiled slots per block of k_track_rs for box size 4*M
int rs_slots(int M, bool fp32);
int dispatch_track_rs(int M, const LaunchCtx& c, const RsArgs& a, long long grid, bool fp32);

}  // namespace flct
