// C ABI of the gfx950 FLCT tracker and Fourier shifter (include/flct_hip.h).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/flct_hip.h"
#include "aux_kernels.hpp"
#include "track_kernels.hpp"
#include "warp_kernels.hpp"
#include "track_launch.hpp"

using namespace flct;

namespace {


struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct flct_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  size_t lds_per_block = 0;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  DevBuf f1, f2, vx, vy, vm, list, count, gw, filt, absmax, wtmp, wF, wtw1, wtw2, w3, w4, w5,
      hs1, hs2, mean1, mean2, flags, fcount, segtab, elist, ecount;
  int seg_n = 0;          // box size of the split tables in segtab
  double seg_sigma = 0.;  // window width of the split tables
  int bf16_kernel = 0;    // FLCT_HIP_BF16_KERNEL: 0 automatic, 1 per-pixel kernels only
  long long last_tracked = 0;
  long long refined = 0;  // pixels re-evaluated in FP64 by the last split-BF16 call
  double last_kernel_ms = 0.;
  double last_total_ms = 0.;
  double last_main_ms = 0.;    // the tracking kernel alone
  double last_refine_ms = 0.;  // FP64 re-evaluation of flagged pixels (split-BF16 form)
};

#define HIPCHK(x)                              \
  do {                                         \
    hipError_t e_ = (x);                       \
    if (e_ != hipSuccess) return FLCT_HIP_E_HIP; \
  } while (0)

static int ensure(DevBuf& b, size_t bytes) {
  if (b.bytes >= bytes && b.p) return FLCT_HIP_OK;
  if (b.p) (void)hipFree(b.p);
  b.p = nullptr;
  b.bytes = 0;
  if (hipMalloc(&b.p, bytes < 256 ? 256 : bytes) != hipSuccess) return FLCT_HIP_E_ALLOC;
  b.bytes = bytes < 256 ? 256 : bytes;
  return FLCT_HIP_OK;
}

extern "C" void flct_hip_opts_default(flct_hip_opts* o) {
  std::memset(o, 0, sizeof(*o));
  o->deltat = 1.;
  o->deltas = 1.;
  o->sigma = 0.;
  o->precision = FLCT_HIP_FP64;
}

extern "C" const char* flct_hip_strerror(int s) {
  switch (s) {
    case FLCT_HIP_OK: return "ok";
    case FLCT_HIP_E_ARG: return "bad argument";
    case FLCT_HIP_E_SKIP_NONPOS: return "skip is zero or negative";
    case FLCT_HIP_E_PQ_RANGE: return "abs(p) or abs(q) >= skip";
    case FLCT_HIP_E_SKIP_TOO_BIG: return "skip is too big compared to nx or ny";
    case FLCT_HIP_E_BOX: return "illegal box size";
    case FLCT_HIP_E_UNSUPPORTED: return "unsupported request";
    case FLCT_HIP_E_NODEVICE: return "no usable HIP device";
    case FLCT_HIP_E_ALLOC: return "device allocation failed";
    case FLCT_HIP_E_HIP: return "HIP runtime error";
    case FLCT_HIP_E_SHAPE: return "bad shift arrays";
    default: return "unknown status";
  }
}

extern "C" int flct_hip_create(flct_hip_ctx** out, int device) {
  if (!out) return FLCT_HIP_E_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return FLCT_HIP_E_NODEVICE;
  if (hipSetDevice(device) != hipSuccess) return FLCT_HIP_E_NODEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return FLCT_HIP_E_NODEVICE;
  auto* c = new flct_hip_ctx();
  c->device = device;
  c->num_cu = prop.multiProcessorCount;
  c->lds_per_block = prop.sharedMemPerBlock;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c;
    return FLCT_HIP_E_HIP;
  }
  for (auto& e : c->ev) {
    if (hipEventCreate(&e) != hipSuccess) {
      delete c;
      return FLCT_HIP_E_HIP;
    }
  }
  // kernel choice of the split-BF16 form: "pixel" keeps every pixel on the
  // per-pixel kernels (for comparisons); default is automatic
  if (const char* k = std::getenv("FLCT_HIP_BF16_KERNEL"))
    c->bf16_kernel = (std::strcmp(k, "pixel") == 0) ? 1 : 0;
  *out = c;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_destroy(flct_hip_ctx* c) {
  if (!c) return FLCT_HIP_OK;
  (void)hipSetDevice(c->device);
  DevBuf* bufs[] = {&c->f1, &c->f2, &c->vx, &c->vy, &c->vm, &c->list, &c->count, &c->gw,
                    &c->filt, &c->absmax, &c->wtmp, &c->wF, &c->wtw1, &c->wtw2, &c->w3,
                    &c->w4, &c->w5, &c->hs1, &c->hs2, &c->mean1, &c->mean2, &c->flags,
                    &c->fcount, &c->segtab, &c->elist, &c->ecount};
  for (DevBuf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_last_stats(flct_hip_ctx* c, long long* tracked, double* kernel_ms,
                                   double* total_ms) {
  if (!c) return FLCT_HIP_E_ARG;
  if (tracked) *tracked = c->last_tracked;
  if (kernel_ms) *kernel_ms = c->last_kernel_ms;
  if (total_ms) *total_ms = c->last_total_ms;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_last_refined(flct_hip_ctx* c, long long* refined) {
  if (!c || !refined) return FLCT_HIP_E_ARG;
  *refined = c->refined;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_last_kernel_ms(flct_hip_ctx* c, double* main_ms, double* refine_ms) {
  if (!c) return FLCT_HIP_E_ARG;
  if (main_ms) *main_ms = c->last_main_ms;
  if (refine_ms) *refine_ms = c->last_refine_ms;
  return FLCT_HIP_OK;
}

// ------------------------------------------------------------------ tracker

namespace {

struct Resolved {
  int nsize = 0, N = 0;
  int poff = 0, qoff = 0, skipon = 0;
};

// argument rules of the reference CLI (flct.c:459-472, 537-538, 559-563,
// 567-580)
int resolve(const flct_hip_opts* o, int nx, int ny, Resolved* r) {
  if (!(o->sigma > 0.)) return FLCT_HIP_E_UNSUPPORTED;
  int p = o->poff, q = o->qoff;
  r->skipon = 0;
  if (o->skip != 0 || p != 0 || q != 0) {
    if (o->skip <= 0) return FLCT_HIP_E_SKIP_NONPOS;
    if ((std::abs(p) >= o->skip) || (std::abs(q) >= o->skip)) return FLCT_HIP_E_PQ_RANGE;
    r->skipon = o->skip + std::abs(q) + std::abs(p);
  }
  if (p < 0) p = o->skip - std::abs(p);
  if (q < 0) q = o->skip - std::abs(q);
  if ((o->skip >= nx) || (o->skip >= ny)) return FLCT_HIP_E_SKIP_TOO_BIG;
  const double tol = 1e-2;
  int nt = (int)o->sigma * std::sqrt(std::log(1. / tol));
  int ndmin = (nx < ny) ? (((nx / 3) / 2)) * 2 : (((ny / 3) / 2)) * 2;
  int nsize = ((nt / 2) * 2 < ndmin) ? (nt / 2) * 2 : ndmin;
  if (nsize <= 0) return FLCT_HIP_E_BOX;
  r->nsize = nsize;
  r->N = 2 * nsize;
  r->poff = p;
  r->qoff = q;
  return FLCT_HIP_OK;
}

}  // namespace

extern "C" int flct_hip_check_opts(int nx, int ny, const flct_hip_opts* o, int* nbox) {
  if (!o || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  Resolved r;
  const int rc = resolve(o, nx, ny, &r);
  if (rc) return rc;
  if (nbox) *nbox = r.N;
  if (r.N % 4 != 0 || r.N / 4 > kTrackMaxM) return FLCT_HIP_E_UNSUPPORTED;
  if (o->precision < FLCT_HIP_FP64 || o->precision > FLCT_HIP_SPLIT_BF16) return FLCT_HIP_E_ARG;
  if (o->deltat == 0.) return FLCT_HIP_E_ARG;
  return FLCT_HIP_OK;
}

// Split-BF16 form through the row-segment kernel: interior columns by
// segments of one row, clamped edge columns by the per-pixel kernel.
static int track_bf16_seg(flct_hip_ctx* c, const LaunchCtx& lc, const TrackArgs& ta, double sigma,
                          double thresh, int skip, int xoffset, int yoffset) {
  const int N = 2 * ta.nsize;
  int rc;
  hipStream_t s = c->stream;
  if (c->seg_n != N || c->seg_sigma != sigma) {
    std::vector<unsigned short> tab;
    seg_tables(N, sigma, tab);
    if ((rc = ensure(c->segtab, tab.size() * 2))) return rc;
    HIPCHK(hipMemcpyAsync(c->segtab.p, tab.data(), tab.size() * 2, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    c->seg_n = N;
    c->seg_sigma = sigma;
  }
  const long long nedge_max = (ta.list ? ta.nwork : (long long)ta.nx * (N - 1));
  if ((rc = ensure(c->elist, (size_t)(nedge_max > 0 ? nedge_max : 1) * sizeof(int))) ||
      (rc = ensure(c->ecount, sizeof(unsigned int))))
    return rc;
  long long nedge = 0;
  if ((rc = launch_edge_list(lc, ta.list, ta.nwork, ta.nx, ta.ny, ta.nsize, (int*)c->elist.p,
                             (unsigned int*)c->ecount.p, &nedge)))
    return rc;
  if (nedge > 0) {
    TrackArgs te = ta;
    te.list = (const int*)c->elist.p;
    te.nwork = nedge;
    if ((rc = dispatch_track_bf16(N, lc, te))) return rc;
  }
  return launch_track_seg(lc, ta, c->segtab.p, thresh, skip, xoffset, yoffset);
}

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.;
  return (double)ms;
}

extern "C" int flct_hip_track_device(flct_hip_ctx* c, const double* f1, const double* f2, int nx,
                                     int ny, const flct_hip_opts* o, double* vx, double* vy,
                                     double* vm) {
  if (!c || !f1 || !f2 || !o || !vx || !vy || !vm || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  if (o->deltat == 0.) return FLCT_HIP_E_ARG;
  Resolved r;
  int rc = resolve(o, nx, ny, &r);
  if (rc) return rc;
  if (o->precision != FLCT_HIP_FP64 && o->precision != FLCT_HIP_FP32 &&
      o->precision != FLCT_HIP_SPLIT_BF16)
    return FLCT_HIP_E_UNSUPPORTED;
  const int N = r.N, M = N / 4;
  if (N % 4 != 0 || M < 1 || M > kTrackMaxM) return FLCT_HIP_E_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->device));
  const long long npix = (long long)nx * ny;
  if (npix > 0x7fffffffLL) return FLCT_HIP_E_UNSUPPORTED;
  hipStream_t s = c->stream;
  HIPCHK(hipEventRecord(c->ev[0], s));

  // window table gw[d + N - 1] = exp(-(sigminv d)^2) (flct.c:652-663, separable)
  const double sigminv = 1. / o->sigma;
  std::vector<double> gw(2 * N - 1);
  for (int d = -(N - 1); d <= N - 1; ++d) {
    const double ax = sigminv * (double)d;
    gw[d + N - 1] = std::exp(-ax * ax);
  }
  if ((rc = ensure(c->gw, gw.size() * sizeof(double)))) return rc;
  HIPCHK(hipMemcpyAsync(c->gw.p, gw.data(), gw.size() * sizeof(double), hipMemcpyHostToDevice, s));
  // filter table over (k, l), k = 0..N/2, l = 0..N-1 (flct.c:1694-1715)
  const double* d_filt = nullptr;
  if (o->filter) {
    std::vector<double> kx(N), filt((N / 2 + 1) * N);
    {
      int n21 = (N / 2) - 1, inext;
      for (int i = 0; i < n21 + 1; i++) kx[i] = (double)i;
      inext = n21 + 1;
      kx[inext] = (double)(N / 2);
      inext++;
      for (int i = inext; i < N; i++) kx[i] = -(double)(n21 - (i - inext));
    }
    const double kxroll = o->kr * kx[N / 2];
    const double smxinv = 1. / kxroll;
    for (int k = 0; k <= N / 2; ++k) {
      const double argx = kx[k] * smxinv;
      for (int l = 0; l < N; ++l) {
        const double argy = kx[l] * smxinv;
        filt[k * N + l] = std::exp(-(argx * argx + argy * argy));
      }
    }
    if ((rc = ensure(c->filt, filt.size() * sizeof(double)))) return rc;
    HIPCHK(hipMemcpyAsync(c->filt.p, filt.data(), filt.size() * sizeof(double),
                          hipMemcpyHostToDevice, s));
    d_filt = (const double*)c->filt.p;
  }

  // threshold: relative -> absolute (flct.c:594-629)
  double thresh = o->thresh;
  if ((thresh > 0.) && (thresh < 1.) && (!o->thresh_abs)) {
    if ((rc = ensure(c->absmax, 2 * sizeof(unsigned long long)))) return rc;
    HIPCHK(hipMemsetAsync(c->absmax.p, 0, 2 * sizeof(unsigned long long), s));
    auto* am = (unsigned long long*)c->absmax.p;
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f1, npix, am);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f2, npix, am + 1);
    HIPCHK(hipGetLastError());
    unsigned long long h[2];
    HIPCHK(hipMemcpyAsync(h, am, sizeof(h), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    double f1max, f2max;
    std::memcpy(&f1max, &h[0], 8);
    std::memcpy(&f2max, &h[1], 8);
    const double fmax = (f1max > f2max) ? f1max : f2max;
    thresh *= fmax;
  }

  // tracked pixel selection
  const bool sparse = (thresh > 0.) || r.skipon;
  long long nwork = npix;
  const int* d_list = nullptr;
  const int xoffset = r.qoff, yoffset = r.poff;  // transp = 1 (flct.c:736-740)
  if (sparse) {
    if ((rc = ensure(c->list, npix * sizeof(int)))) return rc;
    if ((rc = ensure(c->count, sizeof(unsigned int)))) return rc;
    HIPCHK(hipMemsetAsync(c->count.p, 0, sizeof(unsigned int), s));
    SelectArgs sa;
    sa.f1 = f1;
    sa.f2 = f2;
    sa.nx = nx;
    sa.ny = ny;
    sa.thresh = thresh;
    sa.skip = r.skipon ? o->skip : 0;
    sa.xoffset = xoffset;
    sa.yoffset = yoffset;
    sa.list = (int*)c->list.p;
    sa.count = (unsigned int*)c->count.p;
    sa.vx = vx;
    sa.vy = vy;
    sa.vm = vm;
    hipLaunchKernelGGL(k_select, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, sa);
    HIPCHK(hipGetLastError());
    unsigned int cnt = 0;
    HIPCHK(hipMemcpyAsync(&cnt, c->count.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    nwork = cnt;
    d_list = (const int*)c->list.p;
  }

  // box means by box origin (flct.c:815-827)
  const int nsx = ny - N + 1, nrx = nx - N + 1;
  if ((rc = ensure(c->hs1, (size_t)nx * nsx * 8)) || (rc = ensure(c->hs2, (size_t)nx * nsx * 8)) ||
      (rc = ensure(c->mean1, (size_t)nrx * nsx * 8)) || (rc = ensure(c->mean2, (size_t)nrx * nsx * 8)))
    return rc;
  {
    const long long nh = (long long)nx * nsx, nm = (long long)nrx * nsx;
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, f1, nx, ny, N,
                       (double*)c->hs1.p);
    hipLaunchKernelGGL(k_rowsum, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, s, f2, nx, ny, N,
                       (double*)c->hs2.p);
    hipLaunchKernelGGL(k_boxmean, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s,
                       (const double*)c->hs1.p, nx, ny, N, (double*)c->mean1.p);
    hipLaunchKernelGGL(k_boxmean, dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, s,
                       (const double*)c->hs2.p, nx, ny, N, (double*)c->mean2.p);
    HIPCHK(hipGetLastError());
  }

  TrackArgs ta;
  ta.f1 = f1;
  ta.f2 = f2;
  ta.nx = nx;
  ta.ny = ny;
  ta.nsize = r.nsize;
  ta.sigminv = sigminv;
  ta.deltinv = 1. / o->deltat;
  ta.deltas = o->deltas;
  ta.gw = (const double*)c->gw.p;
  ta.mean1 = (const double*)c->mean1.p;
  ta.mean2 = (const double*)c->mean2.p;
  ta.filt = d_filt;
  ta.list = d_list;
  ta.nwork = nwork;
  ta.bias = o->bias ? 1 : 0;
  ta.vx = vx;
  ta.vy = vy;
  ta.vm = vm;
  ta.absmax = nullptr;
  ta.flag_list = nullptr;
  ta.flag_count = nullptr;
  c->refined = 0;
  if (o->precision == FLCT_HIP_SPLIT_BF16 && nwork > 0) {
    // pixels the BF16 surface cannot settle are listed for FP64 re-evaluation
    if ((rc = ensure(c->flags, npix * sizeof(int))) || (rc = ensure(c->fcount, sizeof(unsigned int))))
      return rc;
    HIPCHK(hipMemsetAsync(c->fcount.p, 0, sizeof(unsigned int), s));
    ta.flag_list = (int*)c->flags.p;
    ta.flag_count = (unsigned int*)c->fcount.p;
    // image maxima for the power-of-two input scaling of the BF16 form
    if ((rc = ensure(c->absmax, 2 * sizeof(unsigned long long)))) return rc;
    auto* am = (unsigned long long*)c->absmax.p;
    HIPCHK(hipMemsetAsync(am, 0, 2 * sizeof(unsigned long long), s));
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f1, npix, am);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f2, npix, am + 1);
    HIPCHK(hipGetLastError());
    ta.absmax = am;
  }
  HIPCHK(hipEventRecord(c->ev[1], s));
  if (nwork > 0) {
    LaunchCtx lc{c->stream, c->num_cu, c->lds_per_block};
    if (o->precision == FLCT_HIP_SPLIT_BF16) {
      if (!c->bf16_kernel && seg_supported(N, nx, ny, o->bias))
        rc = track_bf16_seg(c, lc, ta, o->sigma, thresh, r.skipon ? o->skip : 0, xoffset, yoffset);
      else
        rc = dispatch_track_bf16(N, lc, ta);
      if (!rc) {
        HIPCHK(hipEventRecord(c->ev[4], s));
        unsigned int nflag = 0;
        HIPCHK(hipMemcpyAsync(&nflag, c->fcount.p, sizeof(nflag), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        c->refined = nflag;
        if (nflag > 0) {
          TrackArgs tb = ta;
          tb.list = (const int*)c->flags.p;
          tb.nwork = nflag;
          tb.absmax = nullptr;
          tb.flag_list = nullptr;
          tb.flag_count = nullptr;
          rc = (M <= 12) ? dispatch_track_f64_lo(M, lc, tb) : dispatch_track_f64_hi(M, lc, tb);
        }
      }
    }
    else if (o->precision == FLCT_HIP_FP32)
      rc = (M <= 12) ? dispatch_track_f32_lo(M, lc, ta) : dispatch_track_f32_hi(M, lc, ta);
    else
      rc = (M <= 12) ? dispatch_track_f64_lo(M, lc, ta) : dispatch_track_f64_hi(M, lc, ta);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(c->ev[2], s));

  // skip fill (flct.c:958-1120) and final reset (flct.c:1125-1133)
  const unsigned pblocks = (unsigned)((npix + 255) / 256);
  if (r.skipon && o->interp) {
    FillArgs fa;
    fa.nx = nx;
    fa.ny = ny;
    fa.skip = o->skip;
    fa.xoffset = xoffset;
    fa.yoffset = yoffset;
    int ixc = 0, iyc = 0;
    for (int i = 0; i < nx; i++)
      if (!((i - xoffset) % o->skip)) ixc++;
    for (int j = 0; j < ny; j++)
      if (!((j - yoffset) % o->skip)) iyc++;
    // interpcc2d reads three lattice points on each side (flct.c:2294-2336)
    if (ixc < 3 || iyc < 3) return FLCT_HIP_E_UNSUPPORTED;
    fa.ixcount = ixc;
    fa.iycount = iyc;
    fa.vx = vx;
    fa.vy = vy;
    fa.vm = vm;
    hipLaunchKernelGGL(k_fill, dim3(pblocks), dim3(256), 0, s, fa);
    HIPCHK(hipGetLastError());
  }
  hipLaunchKernelGGL(k_finish, dim3(pblocks), dim3(256), 0, s, vx, vy, vm, npix);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[3], s));
  HIPCHK(hipStreamSynchronize(s));
  c->last_tracked = nwork;
  c->last_kernel_ms = ev_ms(c->ev[1], c->ev[2]);
  c->last_total_ms = ev_ms(c->ev[0], c->ev[3]);
  if (o->precision == FLCT_HIP_SPLIT_BF16 && nwork > 0) {
    c->last_main_ms = ev_ms(c->ev[1], c->ev[4]);
    c->last_refine_ms = ev_ms(c->ev[4], c->ev[2]);
  } else {
    c->last_main_ms = c->last_kernel_ms;
    c->last_refine_ms = 0.;
  }
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_track(flct_hip_ctx* c, const double* f1, const double* f2, int nx, int ny,
                              const flct_hip_opts* o, double* vx, double* vy, double* vm) {
  if (!c || !f1 || !f2 || !o || !vx || !vy || !vm || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)nx * ny * sizeof(double);
  int rc;
  if ((rc = ensure(c->f1, bytes)) || (rc = ensure(c->f2, bytes)) || (rc = ensure(c->vx, bytes)) ||
      (rc = ensure(c->vy, bytes)) || (rc = ensure(c->vm, bytes)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->f1.p, f1, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->f2.p, f2, bytes, hipMemcpyHostToDevice, c->stream));
  rc = flct_hip_track_device(c, (const double*)c->f1.p, (const double*)c->f2.p, nx, ny, o,
                             (double*)c->vx.p, (double*)c->vy.p, (double*)c->vm.p);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(vx, c->vx.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(vy, c->vy.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(vm, c->vm.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FLCT_HIP_OK;
}

// ------------------------------------------------------------------ warp

static void root_table(std::vector<double>& t, int n) {
  t.resize(2 * (size_t)n);
  for (int e = 0; e < n; ++e) {
    const double a = 2.0 * M_PI * (double)e / (double)n;
    t[2 * e] = std::cos(a);
    t[2 * e + 1] = std::sin(a);
  }
}

static unsigned nblk(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

extern "C" int flct_hip_warp_device(flct_hip_ctx* c, const double* arr, const double* delx,
                                    const double* dely, double* out, int nx, int ny, int transp) {
  if (!c || !arr || !delx || !dely || !out || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int nb = ny / 2 + 1;
  const int kpos = (ny % 2) ? (ny - 1) / 2 : ny / 2 - 1;
  int rc;
  std::vector<double> t1, t2;
  root_table(t1, ny);
  root_table(t2, nx);
  if ((rc = ensure(c->wtw1, t1.size() * 8)) || (rc = ensure(c->wtw2, t2.size() * 8)) ||
      (rc = ensure(c->wtmp, (size_t)nx * nb * 16)) || (rc = ensure(c->wF, (size_t)nx * nb * 16)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->wtw1.p, t1.data(), t1.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->wtw2.p, t2.data(), t2.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipEventRecord(c->ev[0], s));
  // rows: Frow[i][b] = sum_j arr[i][j] e(-j b / ny), b = 0..ny/2
  hipLaunchKernelGGL(k_dft<true>, dim3(nblk((long long)nx * nb, 256)), dim3(256), 0, s, arr,
                     (long long)ny, 1LL, ny, nb, nx, (double*)c->wtmp.p, (long long)nb, 1LL, -1,
                     (const double*)c->wtw1.p);
  // columns: F[a][b] = sum_i Frow[i][b] e(-i a / nx)
  hipLaunchKernelGGL(k_dft<false>, dim3(nblk((long long)nb * nx, 256)), dim3(256), 0, s,
                     (const double*)c->wtmp.p, 1LL, (long long)nb, nx, nx, nb, (double*)c->wF.p,
                     1LL, (long long)nb, -1, (const double*)c->wtw2.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[1], s));
  WarpArgs w;
  w.F = (const double*)c->wF.p;
  w.nx = nx;
  w.ny = ny;
  w.nb = nb;
  w.kpos = kpos;
  w.delx = delx;
  w.dely = dely;
  w.transp = transp;
  w.out = out;
  hipLaunchKernelGGL(k_warp, dim3(nblk((long long)nx * ny, kWP)), dim3(256), 0, s, w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[2], s));
  HIPCHK(hipStreamSynchronize(s));
  c->last_kernel_ms = ev_ms(c->ev[1], c->ev[2]);
  c->last_total_ms = ev_ms(c->ev[0], c->ev[2]);
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_warp(flct_hip_ctx* c, const double* arr, const double* delx,
                             const double* dely, double* outarr, int nx, int ny, int transp,
                             int quiet) {
  (void)quiet;
  if (!c || !arr || !delx || !dely || !outarr || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)nx * ny * 8;
  int rc;
  if ((rc = ensure(c->f1, bytes)) || (rc = ensure(c->vx, bytes)) || (rc = ensure(c->vy, bytes)) ||
      (rc = ensure(c->vm, bytes)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->f1.p, arr, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->vx.p, delx, bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->vy.p, dely, bytes, hipMemcpyHostToDevice, c->stream));
  rc = flct_hip_warp_device(c, (const double*)c->f1.p, (const double*)c->vx.p,
                            (const double*)c->vy.p, (double*)c->vm.p, nx, ny, transp);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(outarr, c->vm.p, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_shift(flct_hip_ctx* c, const double* arr, double delx, double dely,
                              double* outarr, int nx, int ny, int transp, int quiet) {
  (void)quiet;
  if (!c || !arr || !outarr || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const long long n = (long long)nx * ny;
  int rc;
  std::vector<double> t1, t2;
  root_table(t1, ny);
  root_table(t2, nx);
  if ((rc = ensure(c->f1, n * 8)) || (rc = ensure(c->w3, n * 16)) || (rc = ensure(c->w4, n * 16)) ||
      (rc = ensure(c->wtw1, t1.size() * 8)) || (rc = ensure(c->wtw2, t2.size() * 8)) ||
      (rc = ensure(c->vx, n * 8)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->f1.p, arr, n * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->wtw1.p, t1.data(), t1.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->wtw2.p, t2.data(), t2.size() * 8, hipMemcpyHostToDevice, s));
  double* A = (double*)c->w3.p;
  double* B = (double*)c->w4.p;
  // forward 2D transform (warp.c:806)
  hipLaunchKernelGGL(k_dft<true>, dim3(nblk(n, 256)), dim3(256), 0, s, (const double*)c->f1.p,
                     (long long)ny, 1LL, ny, ny, nx, A, (long long)ny, 1LL, -1,
                     (const double*)c->wtw1.p);
  hipLaunchKernelGGL(k_dft<false>, dim3(nblk(n, 256)), dim3(256), 0, s, (const double*)A, 1LL,
                     (long long)ny, nx, nx, ny, B, 1LL, (long long)ny, -1, (const double*)c->wtw2.p);
  // phase ramp (warp.c:811-852)
  double dxarg, dyarg;
  if (transp) {
    dxarg = -dely / ((double)nx);
    dyarg = -delx / ((double)ny);
  } else {
    dxarg = -delx / ((double)nx);
    dyarg = -dely / ((double)ny);
  }
  hipLaunchKernelGGL(k_phase_ramp, dim3(nblk(n, 256)), dim3(256), 0, s, B, nx, ny, dxarg, dyarg);
  // backward 2D transform (warp.c:855)
  hipLaunchKernelGGL(k_dft<false>, dim3(nblk(n, 256)), dim3(256), 0, s, (const double*)B,
                     (long long)ny, 1LL, ny, ny, nx, A, (long long)ny, 1LL, +1,
                     (const double*)c->wtw1.p);
  hipLaunchKernelGGL(k_dft<false>, dim3(nblk(n, 256)), dim3(256), 0, s, (const double*)A, 1LL,
                     (long long)ny, nx, nx, ny, B, 1LL, (long long)ny, +1, (const double*)c->wtw2.p);
  hipLaunchKernelGGL(k_real_scale, dim3(nblk(n, 256)), dim3(256), 0, s, (const double*)B,
                     (double*)c->vx.p, n, (1. / ((double)nx * ny)));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(outarr, c->vx.p, n * 8, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  return FLCT_HIP_OK;
}

// ------------------------------------------------------------------ memory helpers

extern "C" int flct_hip_malloc(flct_hip_ctx* c, void** dptr, long long bytes) {
  if (!c || !dptr || bytes <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  if (hipMalloc(dptr, (size_t)bytes) != hipSuccess) return FLCT_HIP_E_ALLOC;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_free(flct_hip_ctx* c, void* dptr) {
  if (!c) return FLCT_HIP_E_ARG;
  if (!dptr) return FLCT_HIP_OK;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipFree(dptr));
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_upload(flct_hip_ctx* c, void* dptr, const void* src, long long bytes) {
  if (!c || !dptr || !src || bytes < 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dptr, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_download(flct_hip_ctx* c, void* dst, const void* dptr, long long bytes) {
  if (!c || !dptr || !dst || bytes < 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpyAsync(dst, dptr, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return FLCT_HIP_OK;
}
