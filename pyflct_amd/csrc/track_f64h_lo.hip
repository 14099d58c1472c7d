// Tracker instantiations with the -h peak step: double, box sizes N = 4*M for M = 1..12.
#include "track_launch.hpp"

namespace flct {

int dispatch_track_f64h_lo(int M, const LaunchCtx& c, const TrackArgs& a) {
  int rc = FLCT_HIP_E_UNSUPPORTED;
  static_for<12>([&](auto I) {
    constexpr int m = I + 1;
    if (m == M) rc = launch_track<m, double, true>(c, a);
  });
  return rc;
}

}  // namespace flct
