// Split-BF16 MFMA tracker instantiations.  Boxes N = 36..64 with N % 4 == 0
// use the radix-2 register-chained kernel; other sizes use the dense DFT
// kernel padded to NP = 32, 64 or 96.
#include "track_launch.hpp"
#include "track_mfma.hpp"
#include "track_mfma2.hpp"

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

}  // namespace flct
