// C ABI of the gfx950 FLCT tracker and Fourier shifter (include/flct_hip.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/flct_hip.h"
#include "aux_kernels.hpp"
#include "track_kernels.hpp"
#include "warp_kernels.hpp"
#include "interp_cc.hpp"
#include "whole_kernels.hpp"
#include "nufft_kernels.hpp"
#include "fft_rows.hpp"
#include "track_launch.hpp"
#include "track_big.hpp"
#include "track_rs.hpp"

using namespace flct;


// Device memory of flct_hip_track_device_noalloc: every table and scratch
// buffer of one such call is carved from this array, which each device
// holds once for the life of the process.  In host code the name only
// stands for the symbol; its address on a device comes from
// hipGetSymbolAddress (flct_hip_create) and is the only form passed to
// kernels and copies.
constexpr size_t kNoAllocBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {


struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// A kernel configuration: the path and its runtime shape.  slots: pixels
// tracked together per block by k_track / k_track_rs (1 up to the compiled
// count; 0: the compiled count).  segpx: pixels per row segment of
// k_track_seg / k_track_rs (8..64; 0: 64).
struct Tuning {
  int path = -1;
  int slots = 0;
  int segpx = 0;
};

}  // namespace

struct flct_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cu = 0;
  size_t lds_per_block = 0;
  hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  DevBuf f1, f2, vx, vy, vm, list, count, gw, filt, absmax, wtmp, wF, wtw1, wtw2, w3, w4, w5,
      hs1, hs2, mean1, mean2, flags, fcount, segtab, elist, ecount, nuT, nuR, nuTab, bigp;
  // tables of the row transforms (fft_rows.hpp) by (length, sign)
  struct DftTab {
    int n, sign, logm, bluestein;
    DevBuf buf;  // twiddles (M/2) | chirp (n) | chirp transform (M), as double2
  };
  std::vector<DftTab> dft_tabs;
  // batch: the pairs run on up to kBranches concurrent branches of one
  // graph, each with its own stream (during capture) and scratch; the
  // table arena and the per-pair counters are shared; the cached graph
  static constexpr int kBranches = 4;
  DevBuf bscr[kBranches][8], barena, bcount;
  hipStream_t bstream[kBranches] = {};
  hipEvent_t bev[kBranches + 1] = {};  // fork, then one join per branch
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  std::vector<unsigned char> gkey;  // pairs the cached graph was captured for
  int seg_n = 0;          // box size of the split tables in segtab
  double seg_sigma = 0.;  // window width of the split tables
  int bf16_kernel = 0;    // FLCT_HIP_BF16_KERNEL: 0 automatic, 1 per-pixel kernels only
  // kernel configuration by (precision, box size, bias), chosen by
  // flct_hip_autotune; -h always runs the FP64 kernels and is never stored
  std::map<std::tuple<int, int, int>, Tuning> tuned;
  int tune_gen = 0;       // bumped by every tuning (part of the batch graph key)
  Tuning force;           // configuration under test while tuning (path -1: none)
  int last_path = 0;
  int last_slots = 0, last_segpx = 0;  // runtime shape of the last call's main kernel
  long long last_tracked = 0;
  long long refined = 0;  // pixels re-evaluated in FP64 by the last split-BF16 or FP32 call
  double last_kernel_ms = 0.;
  double last_total_ms = 0.;
  double last_main_ms = 0.;    // the tracking kernel alone
  double last_refine_ms = 0.;  // FP64 re-evaluation of flagged pixels (split-BF16, FP32)
  int nu_nx = 0, nu_ny = 0;  // image size of the NUFFT tables in nuTab
  unsigned char* noalloc = nullptr;  // this device's address of `scratch`
  int warp_method = FLCT_HIP_WARP_AUTO;
  int last_warp_method = 0;
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
    case FLCT_HIP_E_NOSCRATCH: return "problem needs more than the 256 MiB no-allocation scratch";
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
  // the no-allocation entry point's array on this device (set-up, once per
  // context, with the context's device current)
  void* sym = nullptr;
  if (hipGetSymbolAddress(&sym, HIP_SYMBOL(scratch)) == hipSuccess && sym) {
    c->noalloc = static_cast<unsigned char*>(sym);
  } else {
    c->noalloc = nullptr;  // flct_hip_track_device_noalloc then returns FLCT_HIP_E_HIP
    (void)hipGetLastError();
  }
  // kernel choice of the split-BF16 form: "pixel" keeps every pixel on the
  // per-pixel kernels where they meet the form's contract (boxes up to
  // kBf16PixelMaxBox; for comparisons); default is automatic
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
                    &c->fcount, &c->segtab, &c->elist, &c->ecount, &c->nuT, &c->nuR,
                    &c->nuTab, &c->bigp};
  for (DevBuf* b : bufs)
    if (b->p) (void)hipFree(b->p);
  for (auto& set : c->bscr)
    for (DevBuf& b : set)
      if (b.p) (void)hipFree(b.p);
  for (auto& st : c->bstream)
    if (st) (void)hipStreamDestroy(st);
  for (auto& e : c->bev)
    if (e) (void)hipEventDestroy(e);
  for (auto& t : c->dft_tabs)
    if (t.buf.p) (void)hipFree(t.buf.p);
  if (c->barena.p) (void)hipFree(c->barena.p);
  if (c->bcount.p) (void)hipFree(c->bcount.p);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
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
  int whole = 0;  // sigma = 0: one correlation of the whole image
};

// argument rules of the reference CLI (flct.c:459-472, 537-538, 559-563,
// 567-580)
int resolve(const flct_hip_opts* o, int nx, int ny, Resolved* r) {
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
  r->poff = p;
  r->qoff = q;
  // any sigma that is not > 0 (zero, negative, NaN) correlates the whole
  // image once: flct.c:357-365 sets sigmaeq0 and skips the box rules
  // (flct.c:567-586)
  if (!(o->sigma > 0.)) {
    r->whole = 1;
    return FLCT_HIP_OK;
  }
  const double tol = 1e-2;
  int nt = (int)o->sigma * std::sqrt(std::log(1. / tol));
  int ndmin = (nx < ny) ? (((nx / 3) / 2)) * 2 : (((ny / 3) / 2)) * 2;
  int nsize = ((nt / 2) * 2 < ndmin) ? (nt / 2) * 2 : ndmin;
  if (nsize <= 0) return FLCT_HIP_E_BOX;
  r->nsize = nsize;
  r->N = 2 * nsize;
  return FLCT_HIP_OK;
}

}  // namespace

extern "C" int flct_hip_check_opts(int nx, int ny, const flct_hip_opts* o, int* nbox) {
  if (!o || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  Resolved r;
  const int rc = resolve(o, nx, ny, &r);
  if (rc) return rc;
  if (nbox) *nbox = r.N;
  if (!r.whole && (r.N % 4 != 0 || r.N > kBigMaxN)) return FLCT_HIP_E_UNSUPPORTED;
  if (o->precision < FLCT_HIP_FP64 || o->precision > FLCT_HIP_SPLIT_BF16) return FLCT_HIP_E_ARG;
  if (o->deltat == 0.) return FLCT_HIP_E_ARG;
  return FLCT_HIP_OK;
}

// Kernel paths of the tracker.  A precision form may run on any path whose
// error is within the form's contract: FP64 serves every form.
// PATH_FP64_PLANE: boxes above 96, every form (k_track_big, FP64)
enum Path {
  PATH_FP64 = 0,
  PATH_FP32 = 1,
  PATH_BF16_SEG = 2,
  PATH_BF16_PIXEL = 3,
  PATH_FP64_PLANE = 4,
  PATH_FP64_ROWSEG = 5,  // FP64, shared row transforms (k_track_rs)
  PATH_FP32_ROWSEG = 6   // shared row transforms in FP64, row tasks in FP32 (k_track_rs)
};
constexpr int kQuadMaxN = 4 * kTrackMaxM;  // largest box of the quad-FFT kernels

static const char* path_name(int p) {
  switch (p) {
    case PATH_FP64: return "fp64-fft";
    case PATH_FP32: return "fp32-fft";
    case PATH_BF16_SEG: return "bf16-segment";
    case PATH_BF16_PIXEL: return "bf16-pixel";
    case PATH_FP64_PLANE: return "fp64-plane";
    case PATH_FP64_ROWSEG: return "fp64-rowseg";
    case PATH_FP32_ROWSEG: return "fp32-rowseg";
    default: return "none";
  }
}

// Largest box on which the per-pixel split-BF16 kernels keep the form's
// contract (RMS <= 1e-4 px against FP64).  On smooth fields with broad
// correlation peaks (random pairs about three boxes wide, smoothed over
// box / 16 px, four seeds) the dense-DFT kernel's RMS error grows with the
// box: at most 6.5e-5 px up to box 72, 8.8e-5 at 76 and 80, 1.2e-4 to
// 1.5e-4 at 84, 92 and 96.
constexpr int kBf16PixelMaxBox = 72;

// candidate paths of a form and box size (the first is the untuned default
// unless default_path says otherwise)
// The row-shared FP64 kernel serves calls without a skip lattice, -b or -h.
static bool rowseg_ok(int N, int bias, int hires, int skipon) {
  return N <= kQuadMaxN && !bias && !hires && !skipon;
}

static std::vector<int> path_candidates(int precision, int N, int nx, int ny, int bias,
                                        int hires, int skipon = 0) {
  if (N > kQuadMaxN) return {PATH_FP64_PLANE};  // boxes above 96: FP64 in every form
  if (hires) return {PATH_FP64};  // the hires refinement runs in the FP64 kernels
  std::vector<int> v;
  if (precision == FLCT_HIP_FP32) {
    v = {PATH_FP32, PATH_FP64};
  } else if (precision == FLCT_HIP_SPLIT_BF16) {
    if (seg_supported(N, nx, ny, bias)) v.push_back(PATH_BF16_SEG);
    if (N <= kBf16PixelMaxBox) v.push_back(PATH_BF16_PIXEL);
    v.push_back(PATH_FP32);
    v.push_back(PATH_FP64);
  } else {
    v = {PATH_FP64};
  }
  if (rowseg_ok(N, bias, hires, skipon)) {
    if (precision != FLCT_HIP_FP64) v.push_back(PATH_FP32_ROWSEG);
    v.push_back(PATH_FP64_ROWSEG);
  }
  return v;
}

// Untuned choice: the split-BF16 request on boxes up to 16 runs the FP32
// kernels (2048^2 full-disk pair, whole-call ms, fp32-fft / fp32-rowseg /
// bf16-segment: 8 3.0/2.0/11.4, 12 4.7/4.1/11.5, 16 7.3/5.9/11.6, 20
// 12.5/18.1/11.7, 24 18.8/20.5/11.8, 32 30.5/25.4/12.0: the row-segment
// kernel's cost is flat below box 32), the row-segment kernel from 20 (box
// 20 with -b: FP32).  On boxes above 64 it runs the FP32 kernels (with their FP64 re-evaluation of
// unsettled pixels), which are within the split-BF16 error bound and were
// measured 1.9-3.1x faster than the FP64 kernels there (2048^2, boxes 68..84:
// 133-161 ms against 257-493 ms) and 2.9-3.7x faster than the dense-DFT
// BF16 kernel even before their shared row transforms.  Among the FP32
// kernels the row-shared one is taken where it was faster (rowseg32_default).
// Box sizes where a row-shared kernel was measured faster than the quad-FFT
// kernel of its precision by 4 % or more (2048^2 full-disk pair, every pixel
// tracked, whole-call ms, quad FFT / row-shared).
// FP64: 12 5.3/4.4, 16 6.8/6.1, 24 20.3/18.4, 28 28.8/22.2, 32 27.1/22.2,
// 40 77.9/39.4, 44 106/55.4, 48 118/80.2, 52 173/101, 60 180/113,
// 64 195/114, 68 389/257, 72 390/261, 76 431/309, 80 428/371, 84 474/418,
// 88 588/476, 96 616/561; 8 2.8/2.7 (kept); level or slower at 4
// (1.1/1.7), 20 (13.6/16.2) and 92 (609/635).
static bool rowseg_default(int N) {
  switch (N) {
    case 8: case 12: case 16: case 24: case 28: case 32: case 36: case 40: case 44: case 48: case 52:
    case 56: case 60: case 64: case 68: case 72: case 76: case 80: case 84: case 88: case 96:
      return true;
    default:
      return false;
  }
}
// FP32 row tasks (fp32-fft / fp32-rowseg): 8 2.7/1.7, 12 4.3/3.6,
// 16 6.8/5.1, 28 25.0/21.1, 32 27.7/21.0, 36 53.5/30.9, 40 53.9/37.2,
// 44 100/51.0, 48 104/50.0, 52 148/77.7, 60 134/74.9, 64 155/88.0,
// 68 243/133, 72 258/132, 76 291/157, 80 316/144, 84 352/161, 88 387/211,
// 92 411/246, 96 427/228; level or slower at 4 (1.1/1.1), 20 (11.6/15.1)
// and 24 (16.7/16.8).  Used by the FP32 form and by the split-BF16 form on
// the boxes it runs in FP32.
static bool rowseg32_default(int N) {
  switch (N) {
    case 8: case 12: case 16: case 28: case 32: case 36: case 40: case 44: case 48: case 52:
    case 56: case 60: case 64: case 68: case 72: case 76: case 80: case 84: case 88: case 92: case 96:
      return true;
    default:
      return false;
  }
}

static int default_path(int precision, int N, int nx, int ny, int bias, int bf16_kernel) {
  if (N > kQuadMaxN) return PATH_FP64_PLANE;
  if (precision == FLCT_HIP_FP32) return PATH_FP32;
  if (precision == FLCT_HIP_SPLIT_BF16) {
    if (bf16_kernel) return (N <= kBf16PixelMaxBox) ? PATH_BF16_PIXEL : PATH_FP32;
    if (N > 64 || N <= 16) return PATH_FP32;
    if (seg_supported(N, nx, ny, bias)) return PATH_BF16_SEG;
    return (N <= 20) ? PATH_FP32 : PATH_BF16_PIXEL;
  }
  return PATH_FP64;
}

namespace {

// Host plan of one track call: the validated options, the paths taken and
// the host-built tables.  No device work.
struct Plan {
  int nx = 0, ny = 0;
  flct_hip_opts o;
  Resolved r;
  int N = 0, M = 0;
  bool relative = false;  // -t below 1 without 'a': times the image maxima (flct.c:594-629)
  bool sparse = false;    // threshold or skip: a tracked-pixel list
  bool seg = false;       // split-BF16 row-segment kernel for interior columns
  int path = PATH_FP64;   // kernel path
  int ixc = 0, iyc = 0;   // skip-fill lattice size
  long long big_grid = 0; // blocks of k_track_big (boxes above 96)
  int rs_nband = 0;       // k_track_rs: column bands of at most segpx pixels
  long long rs_grid = 0;  // k_track_rs: blocks
  int slots = 0;          // k_track / k_track_rs pixels per block (0: compiled count)
  int segpx = 0;          // k_track_seg / k_track_rs pixels per segment (0: 64)
  long long seg_grid = 0; // k_track_seg: blocks at most (one scratch slice each)
  bool strip = false;     // k_track_seg also runs the transposed edge-column strips
  std::vector<double> gw, filt;
};

int make_plan(int nx, int ny, const flct_hip_opts* o, const flct_hip_ctx* c, Plan* pl,
              const Tuning* force = nullptr) {
  if (!o || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  if (o->deltat == 0.) return FLCT_HIP_E_ARG;
  int rc = resolve(o, nx, ny, &pl->r);
  if (rc) return rc;
  if (pl->r.whole) return FLCT_HIP_E_UNSUPPORTED;  // sigma = 0 runs track_whole
  if (o->precision != FLCT_HIP_FP64 && o->precision != FLCT_HIP_FP32 &&
      o->precision != FLCT_HIP_SPLIT_BF16)
    return FLCT_HIP_E_UNSUPPORTED;
  pl->nx = nx;
  pl->ny = ny;
  pl->o = *o;
  pl->N = pl->r.N;
  pl->M = pl->N / 4;
  if (pl->N % 4 != 0 || pl->M < 1 || pl->N > kBigMaxN) return FLCT_HIP_E_UNSUPPORTED;
  if ((long long)nx * ny > 0x7fffffffLL) return FLCT_HIP_E_UNSUPPORTED;
  const int N = pl->N;
  // window table gw[d + N - 1] = exp(-(sigminv d)^2) (flct.c:652-663, separable)
  const double sigminv = 1. / o->sigma;
  pl->gw.resize(2 * N - 1);
  for (int d = -(N - 1); d <= N - 1; ++d) {
    const double ax = sigminv * (double)d;
    pl->gw[d + N - 1] = std::exp(-ax * ax);
  }
  // filter table over (k, l), k = 0..N/2, l = 0..N-1 (flct.c:1694-1715)
  pl->filt.clear();
  if (o->filter) {
    std::vector<double> kx(N);
    int n21 = (N / 2) - 1, inext;
    for (int i = 0; i < n21 + 1; i++) kx[i] = (double)i;
    inext = n21 + 1;
    kx[inext] = (double)(N / 2);
    inext++;
    for (int i = inext; i < N; i++) kx[i] = -(double)(n21 - (i - inext));
    const double kxroll = o->kr * kx[N / 2];
    const double smxinv = 1. / kxroll;
    pl->filt.resize((N / 2 + 1) * N);
    for (int k = 0; k <= N / 2; ++k) {
      const double argx = kx[k] * smxinv;
      for (int l = 0; l < N; ++l) {
        const double argy = kx[l] * smxinv;
        pl->filt[k * N + l] = std::exp(-(argx * argx + argy * argy));
      }
    }
  }
  pl->relative = (o->thresh > 0.) && (o->thresh < 1.) && (!o->thresh_abs);
  pl->sparse = (o->thresh > 0.) || pl->r.skipon;
  if (force && force->path >= 0) {
    pl->path = force->path;
    pl->slots = force->slots;
    pl->segpx = force->segpx;
  } else {
    const auto it = c->tuned.find(std::make_tuple(o->precision, N, o->bias ? 1 : 0));
    if (it != c->tuned.end()) {
      pl->path = it->second.path;
      pl->slots = it->second.slots;
      pl->segpx = it->second.segpx;
    } else {
      pl->path = default_path(o->precision, N, nx, ny, o->bias, c->bf16_kernel);
      // without a threshold: the row-shared kernels where they were faster
      // (with one, a tracked-pixel list serves sparse selections better)
      if (o->thresh == 0. && rowseg_ok(N, o->bias, o->hires, pl->r.skipon)) {
        if (pl->path == PATH_FP64 && rowseg_default(N)) pl->path = PATH_FP64_ROWSEG;
        if (pl->path == PATH_FP32 && rowseg32_default(N)) pl->path = PATH_FP32_ROWSEG;
      }
    }
  }
  if (pl->path == PATH_BF16_SEG && !seg_supported(N, nx, ny, o->bias)) pl->path = PATH_BF16_PIXEL;
  if (pl->path == PATH_FP64_ROWSEG || pl->path == PATH_FP32_ROWSEG) {
    if (rowseg_ok(N, o->bias, o->hires, pl->r.skipon)) {
      const int w = (pl->segpx >= 8 && pl->segpx < kRsSeg) ? pl->segpx : kRsSeg;
      pl->rs_nband = (ny + w - 1) / w;
      pl->rs_grid = rs_grid(N / 4, c->num_cu, (long long)nx * pl->rs_nband, pl->slots,
                            pl->path == PATH_FP32_ROWSEG);
    } else {
      pl->path = (pl->path == PATH_FP32_ROWSEG) ? PATH_FP32 : PATH_FP64;
    }
  }
  if (o->hires) pl->path = PATH_FP64;  // -h: the FP64 kernels carry the hires refinement
  if (N > kQuadMaxN) {
    // boxes above 96 run k_track_big in every form, -h and -b included
    pl->path = PATH_FP64_PLANE;
    pl->big_grid = big_grid(c->num_cu, N, (long long)nx * ny);
  }
  pl->seg = (pl->path == PATH_BF16_SEG);
  pl->seg_grid = pl->seg ? c->num_cu : 0;
  // the clamped edge columns outside the clamped edge rows run through the
  // row-segment kernel on transposed column strips; the four corners stay
  // on the per-pixel kernel
  pl->strip = pl->seg && nx >= N;
  if (pl->r.skipon && o->interp) {
    int ixc = 0, iyc = 0;
    for (int i = 0; i < nx; i++)
      if (!((i - pl->r.qoff) % o->skip)) ixc++;
    for (int j = 0; j < ny; j++)
      if (!((j - pl->r.poff) % o->skip)) iyc++;
    // interpcc2d reads three lattice points on each side (flct.c:2294-2336)
    if (ixc < 3 || iyc < 3) return FLCT_HIP_E_UNSUPPORTED;
    pl->ixc = ixc;
    pl->iyc = iyc;
  }
  return FLCT_HIP_OK;
}

// Device tables of one call.
struct Tables {
  const double* gw = nullptr;
  const double* filt = nullptr;
  const void* segtab = nullptr;
  BigTab big_fwd, big_inv;  // transform tables of boxes above 96
};

// Device scratch of one call, reused by the pairs of a batch in stream order.
struct Scratch {
  int* list = nullptr;   // tracked pixels (threshold or skip)
  double* hs1 = nullptr;
  double* hs2 = nullptr;
  double* mean1 = nullptr;
  double* mean2 = nullptr;
  int* flags = nullptr;  // split-BF16 and FP32 pixels for FP64 re-evaluation
  int* elist = nullptr;  // edge-column pixels of the row-segment path
  void* planes = nullptr;  // per block: k_track_big N x N planes, k_track_rs spectra, k_track_seg V
};

constexpr int kNumScratch = 8;

// The transposed edge strips of the split-BF16 form (after the row-segment
// kernel's slices in the planes buffer): per side (left, right) and image the
// strip (N x nx), its row sums (N x (nx - N + 1)) and its box means
// (nx - N + 1), as doubles.
size_t strip_bytes(const Plan& pl) {
  if (!pl.strip) return 0;
  const size_t nss = (size_t)(pl.nx - pl.N + 1);
  return 4 * ((size_t)pl.N * pl.nx + (size_t)pl.N * nss + nss) * sizeof(double);
}

// Scratch bytes of one plan: list, hs1, hs2, mean1, mean2, flags, elist, planes
void scratch_sizes(const Plan& pl, size_t sz[kNumScratch]) {
  const long long npix = (long long)pl.nx * pl.ny;
  const int N = pl.N, nsx = pl.ny - N + 1, nrx = pl.nx - N + 1;
  sz[0] = pl.sparse ? npix * sizeof(int) : 0;
  sz[1] = sz[2] = (size_t)pl.nx * nsx * 8;
  sz[3] = sz[4] = (size_t)nrx * nsx * 8;
  sz[5] = (pl.path != PATH_FP64) ? npix * sizeof(int) : 0;
  sz[6] = pl.seg ? (pl.sparse ? npix : (long long)pl.nx * (N - 1)) * sizeof(int) : 0;
  sz[7] = (pl.path == PATH_FP64_PLANE)    ? (size_t)pl.big_grid * 24 * (size_t)N * N
          : (pl.path == PATH_FP64_ROWSEG || pl.path == PATH_FP32_ROWSEG)
              ? (size_t)pl.rs_grid * 16 * N * (size_t)(kRsSeg + N)
          : pl.seg ? (size_t)pl.seg_grid * (size_t)kSegVBytes + strip_bytes(pl)
                                          : 0;
}

// Enqueues the tracking pipeline of one pair on lc.stream: no host
// synchronisation and no memset, so a graph can capture it.
// counts[0]: tracked-list length (threshold or skip); counts[1]: pixels
// listed for FP64 re-evaluation; counts[2]: edge pixels; absmax: the two
// image maxima.  ev (optional): events around the tracking kernels.
int enqueue_track(const LaunchCtx& lc, const Plan& pl, const Tables& tb, const Scratch& sc,
                  unsigned int* counts, unsigned long long* absmax, const double* f1,
                  const double* f2, double* vx, double* vy, double* vm, hipEvent_t* ev) {
  hipStream_t s = lc.stream;
  const flct_hip_opts* o = &pl.o;
  const int nx = pl.nx, ny = pl.ny, N = pl.N, M = pl.M;
  const long long npix = (long long)nx * ny;
  const bool bf16 = (pl.path == PATH_BF16_SEG || pl.path == PATH_BF16_PIXEL);
  const bool big = (pl.path == PATH_FP64_PLANE);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, counts, 4);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned int*>(absmax), 4);
  // image maxima: relative threshold (flct.c:594-629) and the split-BF16 scaling
  if (pl.relative || bf16) {
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f1, npix, absmax);
    hipLaunchKernelGGL(k_absmax, dim3(1024), dim3(256), 0, s, f2, npix, absmax + 1);
  }
  const int xoffset = pl.r.qoff, yoffset = pl.r.poff;  // transp = 1 (flct.c:736-740)
  if (pl.sparse) {
    SelectArgs sa;
    sa.f1 = f1;
    sa.f2 = f2;
    sa.nx = nx;
    sa.ny = ny;
    sa.thresh = o->thresh;
    sa.absmax = pl.relative ? absmax : nullptr;
    sa.skip = pl.r.skipon ? o->skip : 0;
    sa.xoffset = xoffset;
    sa.yoffset = yoffset;
    sa.list = sc.list;
    sa.count = counts;
    sa.vx = vx;
    sa.vy = vy;
    sa.vm = vm;
    hipLaunchKernelGGL(k_select, dim3((unsigned)((npix + 255) / 256)), dim3(256), 0, s, sa);
  }
  // box means by box origin (flct.c:815-827)
  const long long nbx = (ny - N + 1 + kMeanTile - 1) / kMeanTile;
  const long long nrb = (nx - N + 1 + kMeanRows - 1) / kMeanRows, nby = (ny - N + 1 + 255) / 256;
  const size_t rs_lds = (size_t)(kMeanTile + N - 1) * sizeof(double);
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)(nx * nbx)), dim3(kMeanTile), rs_lds, s, f1, nx, ny, N, sc.hs1);
  hipLaunchKernelGGL(k_rowsum, dim3((unsigned)(nx * nbx)), dim3(kMeanTile), rs_lds, s, f2, nx, ny, N, sc.hs2);
  hipLaunchKernelGGL(k_boxmean, dim3((unsigned)(nrb * nby)), dim3(256), 0, s,
                     (const double*)sc.hs1, nx, ny, N, sc.mean1);
  hipLaunchKernelGGL(k_boxmean, dim3((unsigned)(nrb * nby)), dim3(256), 0, s,
                     (const double*)sc.hs2, nx, ny, N, sc.mean2);
  HIPCHK(hipGetLastError());

  TrackArgs ta;
  ta.f1 = f1;
  ta.f2 = f2;
  ta.nx = nx;
  ta.ny = ny;
  ta.nsize = pl.r.nsize;
  ta.sigminv = 1. / o->sigma;
  ta.deltinv = 1. / o->deltat;
  ta.deltas = o->deltas;
  ta.gw = tb.gw;
  ta.mean1 = sc.mean1;
  ta.mean2 = sc.mean2;
  ta.filt = tb.filt;
  ta.list = pl.sparse ? sc.list : nullptr;
  ta.nwork = npix;  // upper bound of the list length, for the grid
  ta.nwork_dev = pl.sparse ? counts : nullptr;
  ta.bias = o->bias ? 1 : 0;
  ta.hires = o->hires ? 1 : 0;
  ta.vx = vx;
  ta.vy = vy;
  ta.vm = vm;
  ta.absmax = bf16 ? absmax : nullptr;
  ta.slots = pl.slots;
  // the split-BF16 and FP32 kernels list the pixels their surface cannot
  // settle; those are tracked again by the FP64 kernels
  const bool refine = bf16 || pl.path == PATH_FP32 || pl.path == PATH_FP64_ROWSEG ||
                      pl.path == PATH_FP32_ROWSEG;
  ta.flag_list = refine ? sc.flags : nullptr;
  ta.flag_count = refine ? counts + 1 : nullptr;
  if (ev) HIPCHK(hipEventRecord(ev[1], s));
  int rc = FLCT_HIP_OK;
  if (big) {
    BigArgs ba;
    ba.f1 = f1;
    ba.f2 = f2;
    ba.nx = nx;
    ba.ny = ny;
    ba.nsize = pl.r.nsize;
    ba.N = N;
    ba.R = big_rows(tb.big_fwd.logm);
    ba.sigminv = ta.sigminv;
    ba.deltinv = ta.deltinv;
    ba.deltas = ta.deltas;
    ba.gw = tb.gw;
    ba.mean1 = sc.mean1;
    ba.mean2 = sc.mean2;
    ba.filt = tb.filt;
    ba.list = ta.list;
    ba.nwork = npix;
    ba.nwork_dev = ta.nwork_dev;
    ba.bias = ta.bias;
    ba.hires = ta.hires;
    ba.vx = vx;
    ba.vy = vy;
    ba.vm = vm;
    ba.plane = (double2*)sc.planes;
    ba.surf = (double*)((char*)sc.planes + (size_t)pl.big_grid * 16 * (size_t)N * N);
    ba.fwd = tb.big_fwd;
    ba.inv = tb.big_inv;
    if ((rc = launch_track_big(lc, ba, pl.big_grid))) return rc;
  } else if (pl.path == PATH_FP64_ROWSEG || pl.path == PATH_FP32_ROWSEG) {
    RsArgs ra;
    ra.t = ta;
    ra.thresh = o->thresh;
    ra.thresh_rel = pl.relative ? 1 : 0;
    ra.absmax = absmax;
    ra.nband = pl.rs_nband;
    ra.ustride = kRsSeg + N;
    ra.phi = (double2*)sc.planes;
    if ((rc = dispatch_track_rs(M, lc, ra, pl.rs_grid, pl.path == PATH_FP32_ROWSEG))) return rc;
  } else if (bf16) {
    if (pl.seg) {
      long long nedge = 0;
      if ((rc = launch_edge_list(lc, ta.list, ta.nwork_dev, nx, ny, ta.nsize, sc.elist, counts + 2,
                                 &nedge, pl.strip)))
        return rc;
      if (nedge != 0) {
        TrackArgs te = ta;
        te.slots = 0;
        te.list = sc.elist;
        te.nwork = (nedge > 0) ? nedge : npix;
        te.nwork_dev = (nedge > 0) ? nullptr : counts + 2;
        if ((rc = dispatch_track_bf16(N, lc, te))) return rc;
      }
      if ((rc = launch_track_seg(lc, ta, tb.segtab, o->thresh, pl.relative ? 1 : 0,
                                 pl.r.skipon ? o->skip : 0, xoffset, yoffset, pl.segpx,
                                 (float*)sc.planes)))
        return rc;
      if (pl.strip) {
        // Edge columns: the left and right N-column strips of both images,
        // transposed, with their box means; in a strip the image's clamped
        // edge columns are rows whose pixels (the image rows whose boxes are
        // not clamped) are interior columns, which the row-segment kernel
        // tracks with the same windows and boxes (flct.c:773-870)
        const int nss = nx - N + 1;
        const size_t simg = (size_t)N * nx, shs = (size_t)N * nss;
        double* sb = (double*)((char*)sc.planes + (size_t)pl.seg_grid * (size_t)kSegVBytes);
        const long long nbx_s = (nss + kMeanTile - 1) / kMeanTile, nby_s = (nss + 255) / 256;
        for (int side = 0; side < 2; ++side) {
          const int c0 = side ? ny - N : 0;
          double* img[2];
          double* mn[2];
          for (int X = 0; X < 2; ++X) {
            const int k = 2 * side + X;
            img[X] = sb + k * simg;
            double* hs = sb + 4 * simg + k * shs;
            mn[X] = sb + 4 * simg + 4 * shs + (size_t)k * nss;
            hipLaunchKernelGGL(k_strip, dim3((unsigned)((simg + 255) / 256)), dim3(256), 0, s,
                               X ? f2 : f1, nx, ny, c0, N, img[X]);
            hipLaunchKernelGGL(k_rowsum, dim3((unsigned)(N * nbx_s)), dim3(kMeanTile), rs_lds, s,
                               (const double*)img[X], N, nx, N, hs);
            hipLaunchKernelGGL(k_boxmean, dim3((unsigned)nby_s), dim3(256), 0, s, (const double*)hs,
                               N, nx, N, mn[X]);
          }
          HIPCHK(hipGetLastError());
          TrackArgs ts = ta;
          ts.f1 = img[0];
          ts.f2 = img[1];
          ts.nx = N;
          ts.ny = nx;
          ts.mean1 = mn[0];
          ts.mean2 = mn[1];
          ts.list = nullptr;
          const SegStrip st{side ? ta.nsize : 0, side ? ta.nsize : ta.nsize - 1, c0, ny};
          if ((rc = launch_track_seg(lc, ts, tb.segtab, o->thresh, pl.relative ? 1 : 0,
                                     pl.r.skipon ? o->skip : 0, xoffset, yoffset, pl.segpx,
                                     (float*)sc.planes, &st)))
            return rc;
        }
      }
    } else if ((rc = dispatch_track_bf16(N, lc, ta))) {
      return rc;
    }
  } else if (pl.path == PATH_FP32) {
    if ((rc = (M <= 12) ? dispatch_track_f32_lo(M, lc, ta) : dispatch_track_f32_hi(M, lc, ta)))
      return rc;
  }
  if (big) {
    // tracked above
  } else if (refine) {
    if (ev) HIPCHK(hipEventRecord(ev[4], s));
    // FP64 re-evaluation of the flagged pixels (count read on the device)
    TrackArgs tr = ta;
    tr.slots = 0;  // the re-evaluation runs the compiled shape
    tr.list = sc.flags;
    tr.nwork = npix;
    tr.nwork_dev = counts + 1;
    tr.absmax = nullptr;
    tr.flag_list = nullptr;
    tr.flag_count = nullptr;
    rc = (M <= 12) ? dispatch_track_f64_lo(M, lc, tr) : dispatch_track_f64_hi(M, lc, tr);
  } else {
    if (ta.hires)
      rc = (M <= 12) ? dispatch_track_f64h_lo(M, lc, ta) : dispatch_track_f64h_hi(M, lc, ta);
    else
      rc = (M <= 12) ? dispatch_track_f64_lo(M, lc, ta) : dispatch_track_f64_hi(M, lc, ta);
  }
  if (rc) return rc;
  if (ev) HIPCHK(hipEventRecord(ev[2], s));

  // skip fill (flct.c:958-1120) and final reset (flct.c:1125-1133)
  const unsigned pblocks = (unsigned)((npix + 255) / 256);
  if (pl.r.skipon && o->interp) {
    FillArgs fa;
    fa.nx = nx;
    fa.ny = ny;
    fa.skip = o->skip;
    fa.xoffset = xoffset;
    fa.yoffset = yoffset;
    fa.ixcount = pl.ixc;
    fa.iycount = pl.iyc;
    fa.vx = vx;
    fa.vy = vy;
    fa.vm = vm;
    hipLaunchKernelGGL(k_fill, dim3(pblocks), dim3(256), 0, s, fa);
  }
  hipLaunchKernelGGL(k_finish, dim3(pblocks), dim3(256), 0, s, vx, vy, vm, npix);
  HIPCHK(hipGetLastError());
  return FLCT_HIP_OK;
}

}  // namespace

static double ev_ms(hipEvent_t a, hipEvent_t b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.;
  return (double)ms;
}

// Scratch buffers of the context sized for a plan
static int ensure_scratch(flct_hip_ctx* c, const Plan& pl, Scratch* sc) {
  size_t sz[kNumScratch];
  scratch_sizes(pl, sz);
  DevBuf* b[kNumScratch] = {&c->list, &c->hs1,   &c->hs2,   &c->mean1,
                            &c->mean2, &c->flags, &c->elist, &c->bigp};
  for (int k = 0; k < kNumScratch; ++k)
    if (int rc = ensure(*b[k], sz[k] ? sz[k] : 256)) return rc;
  sc->list = (int*)c->list.p;
  sc->hs1 = (double*)c->hs1.p;
  sc->hs2 = (double*)c->hs2.p;
  sc->mean1 = (double*)c->mean1.p;
  sc->mean2 = (double*)c->mean2.p;
  sc->flags = (int*)c->flags.p;
  sc->elist = (int*)c->elist.p;
  sc->planes = c->bigp.p;
  return FLCT_HIP_OK;
}

static void root_table(std::vector<double>& t, int n);
static int dft_tables(flct_hip_ctx* c, int n, int sign, const flct_hip_ctx::DftTab** out);

// transform tables of a box above 96 in both directions (cached in the
// context; the fft_rows.hpp layout: twiddles, chirp, chirp transform)
static int big_tables(flct_hip_ctx* c, int N, Tables* tb) {
  for (int sign : {-1, +1}) {
    const flct_hip_ctx::DftTab* t = nullptr;
    if (int rc = dft_tables(c, N, sign, &t)) return rc;
    BigTab& b = (sign < 0) ? tb->big_fwd : tb->big_inv;
    const int M = 1 << t->logm;
    const double2* base = (const double2*)t->buf.p;
    b.logm = t->logm;
    b.bluestein = t->bluestein;
    b.sign = sign;
    b.tw = base;
    b.chirp = base + M / 2;
    b.bhat = base + M / 2 + N;
  }
  return FLCT_HIP_OK;
}
static int row_dft(flct_hip_ctx* c, const double* in, long long irs, long long ies, bool real_in,
                   int n, int nout, int nrows, double* out, long long ors, long long oes, int sign,
                   const double* tw_direct);

// sigma = 0 (flct.c:567-586, 666-675, 802-808, 920-947): one cross_cor of
// the whole nx x ny images with window 1 and whole-image means; threshold and
// skip are ignored and the velocity is a single value.  Device inputs; the
// result (vx, vy, vm) goes to host memory.
static int track_whole(flct_hip_ctx* c, const double* f1, const double* f2, int nx, int ny,
                       const flct_hip_opts* o, double out[3]) {
  const long long n = (long long)nx * ny;
  if (n > 0x7fffffffLL) return FLCT_HIP_E_UNSUPPORTED;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  int rc;
  std::vector<double> t1, t2;
  root_table(t1, ny);
  root_table(t2, nx);
  if ((rc = ensure(c->w3, n * 16)) || (rc = ensure(c->w4, n * 16)) ||
      (rc = ensure(c->wtw1, t1.size() * 8)) || (rc = ensure(c->wtw2, t2.size() * 8)) ||
      (rc = ensure(c->w5, (2 * kSumBlocks + 8) * 8)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->wtw1.p, t1.data(), t1.size() * 8, hipMemcpyHostToDevice, s));
  HIPCHK(hipMemcpyAsync(c->wtw2.p, t2.data(), t2.size() * 8, hipMemcpyHostToDevice, s));
  double* part = (double*)c->w5.p;
  double* means = part + 2 * kSumBlocks;
  unsigned long long* red = (unsigned long long*)(means + 2);  // vmax, loc
  double* z = (double*)c->w3.p;
  double* t = (double*)c->w4.p;
  const double* tw1 = (const double*)c->wtw1.p;
  const double* tw2 = (const double*)c->wtw2.p;
  const unsigned nb = (unsigned)((n + 255) / 256);
  HIPCHK(hipEventRecord(c->ev[0], s));
  hipLaunchKernelGGL(k_sum_partial, dim3(kSumBlocks), dim3(256), 0, s, f1, n, part);
  hipLaunchKernelGGL(k_sum_partial, dim3(kSumBlocks), dim3(256), 0, s, f2, n, part + kSumBlocks);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, s, (const double*)part, kSumBlocks, n, means);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, s, (const double*)(part + kSumBlocks),
                     kSumBlocks, n, means + 1);
  hipLaunchKernelGGL(k_pack_centered, dim3(nb), dim3(256), 0, s, f1, f2, (const double*)means, n, z);
  HIPCHK(hipEventRecord(c->ev[1], s));
  // forward 2-D transform of z (rows, then columns), e^{-2 pi i}
  if ((rc = row_dft(c, z, ny, 1, false, ny, ny, nx, t, ny, 1, -1, tw1)) ||
      (rc = row_dft(c, t, 1, ny, false, nx, nx, ny, z, 1, ny, -1, tw2)))
    return rc;
  hipLaunchKernelGGL(k_cross_full, dim3(nb), dim3(256), 0, s, (const double*)z, nx, ny,
                     o->filter ? 1 : 0, o->kr, t);
  // inverse (c2r: real part of the backward transform)
  if ((rc = row_dft(c, t, ny, 1, false, ny, ny, nx, z, ny, 1, +1, tw1)) ||
      (rc = row_dft(c, z, 1, ny, false, nx, nx, ny, t, 1, ny, +1, tw2)))
    return rc;
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned int*>(red), 4);
  hipLaunchKernelGGL(k_first_max_init, dim3(1), dim3(64), 0, s, red + 1);
  double* sh = z;  // shifted |c|, reusing the spectrum buffer
  hipLaunchKernelGGL(k_abs_shift, dim3(nb), dim3(256), 0, s, (const double*)t, nx, ny, sh, red);
  hipLaunchKernelGGL(k_first_max, dim3(nb), dim3(256), 0, s, (const double*)sh, n,
                     (const unsigned long long*)red, red + 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[2], s));
  unsigned long long hr[2];
  HIPCHK(hipMemcpyAsync(hr, red, sizeof(hr), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  double vmaxd;
  std::memcpy(&vmaxd, &hr[0], 8);
  // peak and 3x3 fit (flct.c:1512-1556, 1639-1640), or the hires refinement
  // (flct.c:1564-1645)
  int ixmax, iymax;
  const bool absccmax = (vmaxd != 0.);
  if (absccmax) {
    ixmax = (int)(hr[1] / (unsigned long long)ny);
    iymax = (int)(hr[1] - (unsigned long long)ixmax * ny);
  } else {
    ixmax = nx / 2;
    iymax = ny / 2;
  }
  double shx = (double)ixmax, shy = (double)iymax;  // peak in surface index units
  if (o->hires) {
    // the 6 x 6 window about the integer peak holds every value it reads
    double win[kHiresWin * kHiresWin] = {0.};
    const int c0 = std::max(0, iymax - 2), c1 = std::min(ny - 1, iymax + 3);
    for (int r = std::max(0, ixmax - 2); r <= std::min(nx - 1, ixmax + 3); ++r)
      HIPCHK(hipMemcpy(&win[(r - ixmax + 2) * kHiresWin + (c0 - iymax + 2)],
                       sh + (long long)r * ny + c0, (size_t)(c1 - c0 + 1) * 8,
                       hipMemcpyDeviceToHost));
    const HiresWindow hw{win, ixmax, iymax};
    int qmax = 0;
    double vmx = hires_value(hw, nx, ny, 0);
    for (int q = 1; q < kHiresN * kHiresN; ++q) {  // maxloc (flct.c:2142-2161)
      const double v = hires_value(hw, nx, ny, q);
      if (v > vmx) {
        vmx = v;
        qmax = q;
      }
    }
    double nb[9] = {0.};
    for (int t = 0; t < 9; ++t) {
      const int ix = qmax / kHiresN + t / 3 - 1, iy = qmax % kHiresN + t % 3 - 1;
      if (ix >= 0 && ix < kHiresN && iy >= 0 && iy < kHiresN)
        nb[t] = hires_value(hw, nx, ny, ix * kHiresN + iy);
    }
    hires_shift(qmax, nb, ixmax, iymax, &shx, &shy);
  } else if (ixmax > 0 && ixmax < nx - 1 && iymax > 0 && iymax < ny - 1 && absccmax) {
    double a[3][3];
    for (int r = 0; r < 3; ++r)
      HIPCHK(hipMemcpy(a[r], sh + (long long)(ixmax - 1 + r) * ny + (iymax - 1), 3 * 8,
                       hipMemcpyDeviceToHost));
    {
#pragma clang fp contract(off)
      const double fx = 0.5 * (a[2][1] - a[0][1]);
      const double fy = 0.5 * (a[1][2] - a[1][0]);
      const double fxx = (a[2][1] + a[0][1] - 2. * (a[1][1]));
      const double fyy = (a[1][2] + a[1][0] - 2. * (a[1][1]));
      const double fxy = 0.25 * (a[2][2] + a[0][0] - a[2][0] - a[0][2]);
      const double shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
      const double shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
      shx = (double)ixmax + shiftsubx;
      shy = (double)iymax + shiftsuby;
    }
  }
  const double shiftx = shx - (double)(nx / 2);
  const double shifty = shy - (double)(ny / 2);
  const double deltinv = 1. / o->deltat;
  out[0] = shifty * deltinv * o->deltas;  // transp = 1 (flct.c:900-905)
  out[1] = shiftx * deltinv * o->deltas;
  out[2] = 1.;
  c->last_tracked = 1;
  c->refined = 0;
  c->last_path = PATH_FP64;
  c->last_kernel_ms = c->last_main_ms = ev_ms(c->ev[1], c->ev[2]);
  c->last_refine_ms = 0.;
  c->last_total_ms = ev_ms(c->ev[0], c->ev[2]);
  return FLCT_HIP_OK;
}

static int end_call(flct_hip_ctx* c, const Plan& pl, const unsigned int* counts);

extern "C" int flct_hip_track_device(flct_hip_ctx* c, const double* f1, const double* f2, int nx,
                                     int ny, const flct_hip_opts* o, double* vx, double* vy,
                                     double* vm) {
  if (!c || !f1 || !f2 || !o || !vx || !vy || !vm || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  {
    Resolved r0;
    const int rc0 = resolve(o, nx, ny, &r0);
    if (rc0) return rc0;
    if (r0.whole) {
      if (o->deltat == 0.) return FLCT_HIP_E_ARG;
      double out[3];
      const int rc = track_whole(c, f1, f2, nx, ny, o, out);
      if (rc) return rc;
      HIPCHK(hipMemcpy(vx, &out[0], 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(vy, &out[1], 8, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(vm, &out[2], 8, hipMemcpyHostToDevice));
      return FLCT_HIP_OK;
    }
  }
  Plan pl;
  int rc = make_plan(nx, ny, o, c, &pl, &c->force);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  HIPCHK(hipEventRecord(c->ev[0], s));
  // tables
  Tables tb;
  if ((rc = ensure(c->gw, pl.gw.size() * 8))) return rc;
  HIPCHK(hipMemcpyAsync(c->gw.p, pl.gw.data(), pl.gw.size() * 8, hipMemcpyHostToDevice, s));
  tb.gw = (const double*)c->gw.p;
  if (!pl.filt.empty()) {
    if ((rc = ensure(c->filt, pl.filt.size() * 8))) return rc;
    HIPCHK(hipMemcpyAsync(c->filt.p, pl.filt.data(), pl.filt.size() * 8, hipMemcpyHostToDevice, s));
    tb.filt = (const double*)c->filt.p;
  }
  if (pl.seg) {
    if (c->seg_n != pl.N || c->seg_sigma != o->sigma) {
      std::vector<unsigned short> tab;
      seg_tables(pl.N, o->sigma, tab);
      if ((rc = ensure(c->segtab, tab.size() * 2))) return rc;
      HIPCHK(hipMemcpyAsync(c->segtab.p, tab.data(), tab.size() * 2, hipMemcpyHostToDevice, s));
      c->seg_n = pl.N;
      c->seg_sigma = o->sigma;
    }
    tb.segtab = c->segtab.p;
  }
  if (pl.path == PATH_FP64_PLANE && (rc = big_tables(c, pl.N, &tb))) return rc;
  Scratch sc;
  if ((rc = ensure_scratch(c, pl, &sc))) return rc;
  if ((rc = ensure(c->count, 4 * sizeof(unsigned int))) ||
      (rc = ensure(c->absmax, 2 * sizeof(unsigned long long))))
    return rc;
  LaunchCtx lc{s, c->num_cu, c->lds_per_block};
  unsigned int* counts = (unsigned int*)c->count.p;
  if ((rc = enqueue_track(lc, pl, tb, sc, counts, (unsigned long long*)c->absmax.p, f1, f2, vx, vy,
                          vm, c->ev)))
    return rc;
  return end_call(c, pl, counts);
}

// End of a single-pair call: waits for the stream, reads the counters and
// keeps the call's statistics in the context.
static int end_call(flct_hip_ctx* c, const Plan& pl, const unsigned int* counts) {
  hipStream_t s = c->stream;
  const int nx = pl.nx, ny = pl.ny;
  HIPCHK(hipEventRecord(c->ev[3], s));
  unsigned int hc[4];
  HIPCHK(hipMemcpyAsync(hc, counts, sizeof(hc), hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  const bool refine = (pl.path == PATH_FP32 || pl.path == PATH_BF16_SEG ||
                       pl.path == PATH_BF16_PIXEL || pl.path == PATH_FP64_ROWSEG ||
                       pl.path == PATH_FP32_ROWSEG);
  c->last_path = pl.path;
  c->last_slots = pl.slots;
  c->last_segpx = pl.segpx;
  c->last_tracked = pl.sparse ? (long long)hc[0] : (long long)nx * ny;
  c->refined = refine ? (long long)hc[1] : 0;
  c->last_kernel_ms = ev_ms(c->ev[1], c->ev[2]);
  c->last_total_ms = ev_ms(c->ev[0], c->ev[3]);
  if (refine) {
    c->last_main_ms = ev_ms(c->ev[1], c->ev[4]);
    c->last_refine_ms = ev_ms(c->ev[4], c->ev[2]);
  } else {
    c->last_main_ms = c->last_kernel_ms;
    c->last_refine_ms = 0.;
  }
  return FLCT_HIP_OK;
}

static int dft_table_host(int n, int sign, std::vector<double>& h, int* logm, int* bluestein);

// flct_hip_track_device with every table and scratch buffer carved from the
// static `scratch` array of the device (its address from flct_hip_create):
// no device allocation.  The tables are copied from host memory on the
// stream; the counters are cleared by kernels (enqueue_track).  Boxes above
// 96 run k_track_big with as many blocks as their planes leave room for.
extern "C" int flct_hip_track_device_noalloc(flct_hip_ctx* c, const double* f1, const double* f2,
                                             int nx, int ny, const flct_hip_opts* o, double* vx,
                                             double* vy, double* vm) {
  if (!c || !f1 || !f2 || !o || !vx || !vy || !vm || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  {
    Resolved r0;
    const int rc0 = resolve(o, nx, ny, &r0);
    if (rc0) return rc0;
    if (r0.whole) return FLCT_HIP_E_UNSUPPORTED;  // sigma <= 0: flct_hip_track_device
  }
  if (!c->noalloc) return FLCT_HIP_E_HIP;
  Plan pl;
  int rc = make_plan(nx, ny, o, c, &pl, &c->force);
  if (rc) return rc;
  const int N = pl.N;
  const long long npix = (long long)nx * ny;
  // host tables
  std::vector<unsigned short> segt;
  if (pl.seg) seg_tables(N, o->sigma, segt);
  std::vector<double> bt[2];
  int blog[2] = {0, 0}, bblue[2] = {0, 0};
  const bool big = (pl.path == PATH_FP64_PLANE);
  if (big)
    for (int k = 0; k < 2; ++k)
      if ((rc = dft_table_host(N, k ? +1 : -1, bt[k], &blog[k], &bblue[k]))) return rc;
  // layout of the call in the array, 256-byte aligned pieces
  size_t off = 0;
  auto take = [&off](size_t bytes) {
    const size_t at = off;
    off += (bytes + 255) / 256 * 256;
    return at;
  };
  const size_t o_cnt = take(4 * sizeof(unsigned int)), o_max = take(2 * sizeof(unsigned long long));
  const size_t o_gw = take(pl.gw.size() * 8);
  const size_t o_filt = pl.filt.empty() ? 0 : take(pl.filt.size() * 8);
  const size_t o_seg = segt.empty() ? 0 : take(segt.size() * 2);
  const size_t o_bt0 = big ? take(bt[0].size() * 8) : 0, o_bt1 = big ? take(bt[1].size() * 8) : 0;
  size_t sz[kNumScratch];
  if (big) {
    // the planes take what is left: fewer blocks when they do not all fit
    pl.big_grid = 0;
    scratch_sizes(pl, sz);
    size_t fixed = off;
    for (int k = 0; k < kNumScratch; ++k) fixed += (sz[k] + 255) / 256 * 256 + 256;
    const size_t plane = 24 * (size_t)N * N;
    const long long fit = (fixed < kNoAllocBytes) ? (long long)((kNoAllocBytes - fixed) / plane) : 0;
    if (fit < 1) return FLCT_HIP_E_NOSCRATCH;
    pl.big_grid = std::min(big_grid(c->num_cu, N, npix), fit);
  }
  scratch_sizes(pl, sz);
  size_t o_sc[kNumScratch];
  for (int k = 0; k < kNumScratch; ++k) o_sc[k] = take(sz[k] ? sz[k] : 256);
  if (off > kNoAllocBytes) return FLCT_HIP_E_NOSCRATCH;

  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  unsigned char* base = c->noalloc;
  HIPCHK(hipEventRecord(c->ev[0], s));
  Tables tb;
  HIPCHK(hipMemcpyAsync(base + o_gw, pl.gw.data(), pl.gw.size() * 8, hipMemcpyHostToDevice, s));
  tb.gw = (const double*)(base + o_gw);
  if (!pl.filt.empty()) {
    HIPCHK(hipMemcpyAsync(base + o_filt, pl.filt.data(), pl.filt.size() * 8, hipMemcpyHostToDevice, s));
    tb.filt = (const double*)(base + o_filt);
  }
  if (!segt.empty()) {
    HIPCHK(hipMemcpyAsync(base + o_seg, segt.data(), segt.size() * 2, hipMemcpyHostToDevice, s));
    tb.segtab = base + o_seg;
  }
  if (big) {
    const size_t ob[2] = {o_bt0, o_bt1};
    for (int k = 0; k < 2; ++k) {
      HIPCHK(hipMemcpyAsync(base + ob[k], bt[k].data(), bt[k].size() * 8, hipMemcpyHostToDevice, s));
      BigTab& b = k ? tb.big_inv : tb.big_fwd;
      const int M = 1 << blog[k];
      const double2* t = (const double2*)(base + ob[k]);
      b.logm = blog[k];
      b.bluestein = bblue[k];
      b.sign = k ? +1 : -1;
      b.tw = t;
      b.chirp = t + M / 2;
      b.bhat = t + M / 2 + N;
    }
  }
  Scratch sc;
  sc.list = (int*)(base + o_sc[0]);
  sc.hs1 = (double*)(base + o_sc[1]);
  sc.hs2 = (double*)(base + o_sc[2]);
  sc.mean1 = (double*)(base + o_sc[3]);
  sc.mean2 = (double*)(base + o_sc[4]);
  sc.flags = (int*)(base + o_sc[5]);
  sc.elist = (int*)(base + o_sc[6]);
  sc.planes = base + o_sc[7];
  unsigned int* counts = (unsigned int*)(base + o_cnt);
  LaunchCtx lc{s, c->num_cu, c->lds_per_block};
  if ((rc = enqueue_track(lc, pl, tb, sc, counts, (unsigned long long*)(base + o_max), f1, f2, vx,
                          vy, vm, c->ev)))
    return rc;
  return end_call(c, pl, counts);
}

extern "C" int flct_hip_track(flct_hip_ctx* c, const double* f1, const double* f2, int nx, int ny,
                              const flct_hip_opts* o, double* vx, double* vy, double* vm) {
  if (!c || !f1 || !f2 || !o || !vx || !vy || !vm || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)nx * ny * sizeof(double);
  int rc;
  if (!(o->sigma > 0.)) {  // one velocity for the whole image: vx[0], vy[0], vm[0]
    Resolved r0;
    if ((rc = resolve(o, nx, ny, &r0))) return rc;
    if (o->deltat == 0.) return FLCT_HIP_E_ARG;
    if ((rc = ensure(c->f1, bytes)) || (rc = ensure(c->f2, bytes))) return rc;
    HIPCHK(hipMemcpyAsync(c->f1.p, f1, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->f2.p, f2, bytes, hipMemcpyHostToDevice, c->stream));
    double out[3];
    if ((rc = track_whole(c, (const double*)c->f1.p, (const double*)c->f2.p, nx, ny, o, out)))
      return rc;
    vx[0] = out[0];
    vy[0] = out[1];
    vm[0] = out[2];
    return FLCT_HIP_OK;
  }
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

extern "C" const char* flct_hip_last_path(flct_hip_ctx* c) {
  return c ? path_name(c->last_path) : "none";
}

extern "C" int flct_hip_last_shape(flct_hip_ctx* c, int* slots, int* segpx) {
  if (!c) return FLCT_HIP_E_ARG;
  if (slots) *slots = c->last_slots;
  if (segpx) *segpx = c->last_segpx;
  return FLCT_HIP_OK;
}

// compiled slots per block of a path's main kernel at box N (0: the path
// has no slot axis)
static int path_slots(int path, int N) {
  const int M = N / 4;
  if (N % 4 != 0 || M < 1 || M > kTrackMaxM) return 0;
  if (path == PATH_FP64) return track_slots<double>(M);
  if (path == PATH_FP32) return track_slots<float>(M);
  if (path == PATH_FP64_ROWSEG) return rs_slots(M, false);
  if (path == PATH_FP32_ROWSEG) return rs_slots(M, true);
  return 0;
}

static bool path_has_segpx(int path) {
  return path == PATH_BF16_SEG || path == PATH_FP64_ROWSEG || path == PATH_FP32_ROWSEG;
}

// a runtime shape the path accepts: slots 0 or 1..its compiled count,
// segpx 0 or 8..64 on the row-segment paths
static bool shape_ok(int path, int N, int slots, int segpx) {
  if (slots != 0 && (slots < 1 || slots > path_slots(path, N))) return false;
  if (segpx != 0 && (!path_has_segpx(path) || segpx < 8 || segpx > 64)) return false;
  return true;
}

extern "C" int flct_hip_set_config(flct_hip_ctx* c, int precision, int nbox, const char* path,
                                   int slots, int segpx) {
  if (!c || nbox < 4 || nbox % 4 != 0 || nbox > kBigMaxN) return FLCT_HIP_E_ARG;
  if (precision < FLCT_HIP_FP64 || precision > FLCT_HIP_SPLIT_BF16) return FLCT_HIP_E_ARG;
  // one choice for calls with and without -b (a -b call the row-segment
  // kernel cannot serve runs the per-pixel one)
  const auto key = std::make_tuple(precision, nbox, 0), keyb = std::make_tuple(precision, nbox, 1);
  if (!path || !*path) {  // back to the untuned default
    c->tuned.erase(key);
    c->tuned.erase(keyb);
    ++c->tune_gen;
    return FLCT_HIP_OK;
  }
  int p = -1;
  for (int q = PATH_FP64; q <= PATH_FP32_ROWSEG; ++q)
    if (std::strcmp(path, path_name(q)) == 0) p = q;
  if (p < 0) return FLCT_HIP_E_ARG;
  // the path must meet the form's error contract; the row-segment kernel
  // also needs a box of 8..64 (calls it cannot serve run the per-pixel one)
  bool ok = false;
  for (int q : path_candidates(precision, nbox, 1 << 20, 1 << 20, 0, 0)) ok = ok || (q == p);
  if (!ok || !shape_ok(p, nbox, slots, segpx)) return FLCT_HIP_E_ARG;
  Tuning t;
  t.path = p;
  t.slots = slots;
  t.segpx = segpx;
  c->tuned[key] = t;
  c->tuned[keyb] = t;
  ++c->tune_gen;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_set_path(flct_hip_ctx* c, int precision, int nbox, const char* path) {
  return flct_hip_set_config(c, precision, nbox, path, 0, 0);
}

// Autotuner.  Stage 1 times every candidate path at its compiled shape;
// stage 2 times the winner's runtime shapes (slots per block, segment
// width) around the compiled one.  Each configuration runs three times on
// the caller's inputs and the last two are summed (the first warms up).
extern "C" int flct_hip_autotune(flct_hip_ctx* c, const double* f1, const double* f2, int nx,
                                 int ny, const flct_hip_opts* o) {
  if (!c || !f1 || !f2 || !o || nx <= 0 || ny <= 0) return FLCT_HIP_E_ARG;
  {
    Resolved r0;
    const int rc0 = resolve(o, nx, ny, &r0);
    if (rc0) return rc0;
    if (r0.whole) return FLCT_HIP_OK;  // one path for sigma = 0
  }
  Plan pl;
  int rc = make_plan(nx, ny, o, c, &pl);
  if (rc) return rc;
  HIPCHK(hipSetDevice(c->device));
  const size_t bytes = (size_t)nx * ny * 8;
  if ((rc = ensure(c->vx, bytes)) || (rc = ensure(c->vy, bytes)) || (rc = ensure(c->vm, bytes)))
    return rc;
  auto time_cfg = [&](const Tuning& t, double* ms) {
    c->force = t;
    double sum = 0.;
    int r = FLCT_HIP_OK;
    for (int rep = 0; rep < 3 && !r; ++rep) {
      r = flct_hip_track_device(c, f1, f2, nx, ny, o, (double*)c->vx.p, (double*)c->vy.p,
                                (double*)c->vm.p);
      if (rep > 0) sum += c->last_total_ms;
    }
    c->force = Tuning();
    *ms = sum;
    return r;
  };
  Tuning best;
  double best_ms = 0.;
  for (int p : path_candidates(o->precision, pl.N, nx, ny, o->bias, o->hires, pl.r.skipon)) {
    Tuning t;
    t.path = p;
    double ms = 0.;
    if ((rc = time_cfg(t, &ms))) return rc;
    if (best.path < 0 || ms < best_ms) {
      best = t;
      best_ms = ms;
    }
  }
  // stage 2: the winner's shapes (fewer slots per block give more blocks
  // on small images; narrower segments more segments per row)
  std::vector<Tuning> shapes;
  const int P = path_slots(best.path, pl.N);
  for (int sl : {P / 2, P / 4, 1})
    if (sl >= 1 && sl < P) {
      Tuning t = best;
      t.slots = sl;
      bool dup = false;
      for (const Tuning& u : shapes) dup = dup || (u.slots == sl && u.segpx == 0);
      if (!dup) shapes.push_back(t);
    }
  if (path_has_segpx(best.path))
    for (int w : {32, 16}) {
      Tuning t = best;
      t.segpx = w;
      shapes.push_back(t);
    }
  for (const Tuning& t : shapes) {
    double ms = 0.;
    if ((rc = time_cfg(t, &ms))) return rc;
    if (ms < best_ms) {
      best = t;
      best_ms = ms;
    }
  }
  // -h forces the FP64 kernels whatever is stored: nothing to keep
  if (!o->hires) {
    c->tuned[std::make_tuple(o->precision, pl.N, o->bias ? 1 : 0)] = best;
    c->tune_gen++;
  }
  c->last_path = best.path;
  c->last_slots = best.slots;
  c->last_segpx = best.segpx;
  return FLCT_HIP_OK;
}

// ------------------------------------------------------------------ batch

static void drop_graph(flct_hip_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  c->gkey.clear();
}

// ensure() for buffers a cached graph may reference: a reallocation drops it
static int ensure_b(flct_hip_ctx* c, DevBuf& b, size_t bytes) {
  if (b.p && b.bytes >= bytes) return FLCT_HIP_OK;
  drop_graph(c);
  return ensure(b, bytes);
}

extern "C" int flct_hip_track_batch(flct_hip_ctx* c, flct_hip_pair* pairs, int npairs) {
  if (!c || npairs < 0 || (npairs > 0 && !pairs)) return FLCT_HIP_E_ARG;
  if (npairs == 0) return FLCT_HIP_OK;
  HIPCHK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  // plans of every pair; nothing runs when one is invalid
  std::vector<Plan> pl(npairs);
  int first_err = FLCT_HIP_OK;
  for (int k = 0; k < npairs; ++k) {
    flct_hip_pair& q = pairs[k];
    int rc = (!q.f1 || !q.f2 || !q.vx || !q.vy || !q.vm) ? FLCT_HIP_E_ARG
                                                          : make_plan(q.nx, q.ny, &q.opts, c, &pl[k]);
    q.status = rc;
    q.tracked = 0;
    q.refined = 0;
    if (rc && !first_err) first_err = rc;
  }
  if (first_err) return first_err;
  // table arena: gw, filt, split tables (one copy per distinct box size and sigma)
  std::vector<unsigned char> arena;
  auto put = [&](const void* src, size_t bytes) {
    const size_t off = (arena.size() + 255) / 256 * 256;
    arena.resize(off + bytes);
    std::memcpy(arena.data() + off, src, bytes);
    return off;
  };
  std::vector<size_t> o_gw(npairs), o_filt(npairs, (size_t)-1), o_seg(npairs, (size_t)-1);
  std::vector<std::pair<std::pair<int, double>, size_t>> segs;
  for (int k = 0; k < npairs; ++k) {
    o_gw[k] = put(pl[k].gw.data(), pl[k].gw.size() * 8);
    if (!pl[k].filt.empty()) o_filt[k] = put(pl[k].filt.data(), pl[k].filt.size() * 8);
    if (pl[k].seg) {
      const std::pair<int, double> key(pl[k].N, pl[k].o.sigma);
      for (auto& e : segs)
        if (e.first == key) o_seg[k] = e.second;
      if (o_seg[k] == (size_t)-1) {
        std::vector<unsigned short> tab;
        seg_tables(pl[k].N, pl[k].o.sigma, tab);
        o_seg[k] = put(tab.data(), tab.size() * 2);
        segs.push_back({key, o_seg[k]});
      }
    }
  }
  // Branches: pairs in decreasing order of estimated work (pixels times the
  // FFT-form flops per pixel of their box), each to the branch with the
  // least work so far (longest processing time first).  The branches run
  // concurrently; inside a branch the pairs run in that order and share the
  // branch's scratch.
  const int nbr = std::min(npairs, flct_hip_ctx::kBranches);
  std::vector<int> order(npairs), branch(npairs, 0);
  std::vector<double> cost(npairs), load(nbr, 0.);
  for (int k = 0; k < npairs; ++k) {
    const double n = (double)pl[k].N;
    cost[k] = (double)pl[k].nx * pl[k].ny * (15. * n * n * std::log2(n) + 14. * n * n);
    order[k] = k;
  }
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return cost[a] > cost[b]; });
  for (int k : order) {
    int b = 0;
    for (int q = 1; q < nbr; ++q)
      if (load[q] < load[b]) b = q;
    branch[k] = b;
    load[b] += cost[k];
  }
  // scratch of each branch for its largest pair, counters per pair
  int rc;
  for (int b = 0; b < nbr; ++b) {
    size_t need[kNumScratch];
    for (int i = 0; i < kNumScratch; ++i) need[i] = 256;
    for (int k = 0; k < npairs; ++k) {
      if (branch[k] != b) continue;
      size_t sz[kNumScratch];
      scratch_sizes(pl[k], sz);
      for (int i = 0; i < kNumScratch; ++i) need[i] = std::max(need[i], sz[i]);
    }
    for (int i = 0; i < kNumScratch; ++i)
      if ((rc = ensure_b(c, c->bscr[b][i], need[i]))) return rc;
    if (!c->bstream[b] && hipStreamCreateWithFlags(&c->bstream[b], hipStreamNonBlocking) != hipSuccess)
      return FLCT_HIP_E_HIP;
  }
  for (auto& e : c->bev)
    if (!e && hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return FLCT_HIP_E_HIP;
  // transform tables of boxes above 96 (cached in the context before capture)
  std::vector<Tables> btab(npairs);
  for (int k = 0; k < npairs; ++k)
    if (pl[k].path == PATH_FP64_PLANE && (rc = big_tables(c, pl[k].N, &btab[k]))) return rc;
  if ((rc = ensure_b(c, c->barena, arena.size())) ||
      (rc = ensure_b(c, c->bcount, (size_t)npairs * 32)))
    return rc;
  HIPCHK(hipMemcpyAsync(c->barena.p, arena.data(), arena.size(), hipMemcpyHostToDevice, s));
  // the graph is reused while the pairs are the same
  std::vector<unsigned char> key((size_t)npairs * sizeof(flct_hip_pair) + sizeof(int));
  std::memcpy(key.data() + (size_t)npairs * sizeof(flct_hip_pair), &c->tune_gen, sizeof(int));
  for (int k = 0; k < npairs; ++k) {
    flct_hip_pair q = pairs[k];
    q.tracked = 0;
    q.refined = 0;
    q.status = 0;
    std::memcpy(key.data() + (size_t)k * sizeof(q), &q, sizeof(q));
  }
  if (!c->gexec || key != c->gkey) {
    drop_graph(c);
    unsigned char* base = (unsigned char*)c->barena.p;
    HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
    // fork: every branch stream joins the capture through an event of s
    int erc = (hipEventRecord(c->bev[0], s) == hipSuccess) ? FLCT_HIP_OK : FLCT_HIP_E_HIP;
    for (int b = 0; b < nbr && !erc; ++b)
      if (hipStreamWaitEvent(c->bstream[b], c->bev[0], 0) != hipSuccess) erc = FLCT_HIP_E_HIP;
    for (int k : order) {
      if (erc) break;
      const int b = branch[k];
      Scratch sc;
      sc.list = (int*)c->bscr[b][0].p;
      sc.hs1 = (double*)c->bscr[b][1].p;
      sc.hs2 = (double*)c->bscr[b][2].p;
      sc.mean1 = (double*)c->bscr[b][3].p;
      sc.mean2 = (double*)c->bscr[b][4].p;
      sc.flags = (int*)c->bscr[b][5].p;
      sc.elist = (int*)c->bscr[b][6].p;
      sc.planes = c->bscr[b][7].p;
      LaunchCtx lc{c->bstream[b], c->num_cu, c->lds_per_block};
      Tables tb = btab[k];
      tb.gw = (const double*)(base + o_gw[k]);
      tb.filt = (o_filt[k] != (size_t)-1) ? (const double*)(base + o_filt[k]) : nullptr;
      tb.segtab = (o_seg[k] != (size_t)-1) ? (const void*)(base + o_seg[k]) : nullptr;
      unsigned char* cnt = (unsigned char*)c->bcount.p + (size_t)k * 32;
      erc = enqueue_track(lc, pl[k], tb, sc, (unsigned int*)cnt, (unsigned long long*)(cnt + 16),
                          pairs[k].f1, pairs[k].f2, pairs[k].vx, pairs[k].vy, pairs[k].vm, nullptr);
    }
    // join: s waits for every branch before the capture ends
    for (int b = 0; b < nbr; ++b) {
      if (hipEventRecord(c->bev[1 + b], c->bstream[b]) != hipSuccess ||
          hipStreamWaitEvent(s, c->bev[1 + b], 0) != hipSuccess)
        if (!erc) erc = FLCT_HIP_E_HIP;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s, &g);
    if (erc) {
      if (g) (void)hipGraphDestroy(g);
      return erc;
    }
    if (ec != hipSuccess || !g) return FLCT_HIP_E_HIP;
    c->graph = g;
    if (hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0) != hipSuccess) {
      drop_graph(c);
      return FLCT_HIP_E_HIP;
    }
    c->gkey = key;
  }
  HIPCHK(hipEventRecord(c->ev[0], s));
  HIPCHK(hipGraphLaunch(c->gexec, s));
  HIPCHK(hipEventRecord(c->ev[3], s));
  std::vector<unsigned int> hc((size_t)npairs * 8);
  HIPCHK(hipMemcpyAsync(hc.data(), c->bcount.p, hc.size() * 4, hipMemcpyDeviceToHost, s));
  HIPCHK(hipStreamSynchronize(s));
  long long tracked = 0, refined = 0;
  for (int k = 0; k < npairs; ++k) {
    const unsigned int* ck = hc.data() + (size_t)k * 8;
    pairs[k].tracked = pl[k].sparse ? (long long)ck[0] : (long long)pl[k].nx * pl[k].ny;
    const int pk = pl[k].path;
    pairs[k].refined =
        (pk == PATH_FP32 || pk == PATH_BF16_SEG || pk == PATH_BF16_PIXEL || pk == PATH_FP64_ROWSEG ||
         pk == PATH_FP32_ROWSEG)
            ? (long long)ck[1]
            : 0;
    tracked += pairs[k].tracked;
    refined += pairs[k].refined;
  }
  c->last_tracked = tracked;
  c->refined = refined;
  c->last_kernel_ms = c->last_total_ms = c->last_main_ms = ev_ms(c->ev[0], c->ev[3]);
  c->last_refine_ms = 0.;
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

// in-place radix-2 transform on the host, sign -1 (tables of the row DFTs)
static void host_fft(std::vector<double>& re, std::vector<double>& im, int logm) {
  const int M = 1 << logm;
  for (int i = 0; i < M; ++i) {
    const int j = (int)(__builtin_bitreverse32((unsigned)i) >> (32 - logm));
    if (j > i) {
      std::swap(re[i], re[j]);
      std::swap(im[i], im[j]);
    }
  }
  for (int len = 2; len <= M; len <<= 1) {
    for (int j = 0; j < len / 2; ++j) {
      const double a = -2. * M_PI * (double)j / (double)len;
      const double wr = std::cos(a), wi = std::sin(a);
      for (int i = 0; i < M; i += len) {
        const int i0 = i + j, i1 = i0 + len / 2;
        const double vr = re[i1] * wr - im[i1] * wi, vi = re[i1] * wi + im[i1] * wr;
        re[i1] = re[i0] - vr;
        im[i1] = im[i0] - vi;
        re[i0] += vr;
        im[i0] += vi;
      }
    }
  }
}

// host tables of the length-n row transform with the given sign (the
// fft_rows.hpp layout: twiddles, then chirp and chirp transform for
// Bluestein), as double2 entries
static int dft_table_host(int n, int sign, std::vector<double>& h, int* logm, int* bluestein) {
  const int blue = (n & (n - 1)) != 0;
  int l = 0;
  while ((1 << l) < (blue ? 2 * n - 1 : n)) ++l;
  const int M = 1 << l;
  if (l > kRowDftMaxLog2) return FLCT_HIP_E_UNSUPPORTED;
  *logm = l;
  *bluestein = blue;
  h.clear();
  h.reserve(2 * (M / 2 + n + M));
  for (int e = 0; e < M / 2; ++e) {
    h.push_back(std::cos(2. * M_PI * e / M));
    h.push_back(std::sin(2. * M_PI * e / M));
  }
  if (blue) {
    // w_m = e^{sign pi i m^2 / n}, the exponent reduced mod 2n exactly
    std::vector<double> wr(n), wi(n);
    for (int m = 0; m < n; ++m) {
      const long long r = ((long long)m * m) % (2LL * n);
      const double a = (double)sign * M_PI * (double)r / (double)n;
      wr[m] = std::cos(a);
      wi[m] = std::sin(a);
      h.push_back(wr[m]);
      h.push_back(wi[m]);
    }
    std::vector<double> br(M, 0.), bi(M, 0.);
    for (int m = 0; m < n; ++m) {
      br[m] = wr[m];
      bi[m] = -wi[m];
      if (m) {
        br[M - m] = wr[m];
        bi[M - m] = -wi[m];
      }
    }
    host_fft(br, bi, l);
    for (int q = 0; q < M; ++q) {
      const int k = (int)(__builtin_bitreverse32((unsigned)q) >> (32 - l));
      h.push_back(br[k] / M);
      h.push_back(bi[k] / M);
    }
  }
  return FLCT_HIP_OK;
}

// tables of the length-n row transform with the given sign (cached)
static int dft_tables(flct_hip_ctx* c, int n, int sign, const flct_hip_ctx::DftTab** out) {
  for (const auto& t : c->dft_tabs)
    if (t.n == n && t.sign == sign) {
      *out = &t;
      return FLCT_HIP_OK;
    }
  flct_hip_ctx::DftTab t;
  t.n = n;
  t.sign = sign;
  std::vector<double> h;
  int rc;
  if ((rc = dft_table_host(n, sign, h, &t.logm, &t.bluestein))) return rc;
  if ((rc = ensure(t.buf, h.size() * 8))) return rc;
  HIPCHK(hipMemcpy(t.buf.p, h.data(), h.size() * 8, hipMemcpyHostToDevice));
  c->dft_tabs.push_back(t);
  *out = &c->dft_tabs.back();
  return FLCT_HIP_OK;
}

// out[r][k] = sum_t in[r][t] e^{sign 2 pi i t k / n}, k < nout, on the
// context's stream: k_row_dft for n <= 4096, else the direct k_dft with the
// root table tw_direct (root_table(n))
static int row_dft(flct_hip_ctx* c, const double* in, long long irs, long long ies, bool real_in,
                   int n, int nout, int nrows, double* out, long long ors, long long oes, int sign,
                   const double* tw_direct) {
  hipStream_t s = c->stream;
  const flct_hip_ctx::DftTab* t = nullptr;
  if (n <= (1 << (kRowDftMaxLog2 - 1)) && dft_tables(c, n, sign, &t) == FLCT_HIP_OK) {
    const int M = 1 << t->logm;
    const double2* base = (const double2*)t->buf.p;
    RowDftArgs q;
    q.in = in;
    q.irs = irs;
    q.ies = ies;
    q.real_in = real_in ? 1 : 0;
    q.n = n;
    q.nout = nout;
    q.nrows = nrows;
    q.out = out;
    q.ors = ors;
    q.oes = oes;
    q.sign = sign;
    q.logm = t->logm;
    q.bluestein = t->bluestein;
    q.tw = base;
    q.chirp = base + M / 2;
    q.bhat = base + M / 2 + n;
    const size_t lds = (size_t)M * 16;
    if (lds > c->lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
    (void)hipFuncSetAttribute((const void*)k_row_dft, hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)lds);
    hipLaunchKernelGGL(k_row_dft, dim3(nrows), dim3(kRowDftThreads), lds, s, q);
  } else if (real_in) {
    hipLaunchKernelGGL(k_dft<true>, dim3(nblk((long long)nrows * nout, 256)), dim3(256), 0, s, in,
                       irs, ies, n, nout, nrows, out, ors, oes, sign, tw_direct);
  } else {
    hipLaunchKernelGGL(k_dft<false>, dim3(nblk((long long)nrows * nout, 256)), dim3(256), 0, s, in,
                       irs, ies, n, nout, nrows, out, ors, oes, sign, tw_direct);
  }
  HIPCHK(hipGetLastError());
  return FLCT_HIP_OK;
}

// Gauss-Legendre nodes and weights on [-1, 1] (Newton on P_n)
static void gauss_legendre(int n, std::vector<double>& x, std::vector<double>& w) {
  x.resize(n);
  w.resize(n);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1., p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2. * k - 1.) * z * p1 - (k - 1.) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2. / ((1. - z * z) * dp * dp);
  }
}

// 1 / phihat(2 pi k / M) of the warp's spreading kernel for the modes
// k = make_freq(a, n), a < n: phihat(xi) = 2 int_0^{w/2} phi(z) cos(xi z) dz
static void nufft_deconv(int n, int M, std::vector<double>& d) {
  static std::vector<double> gx, gw;
  if (gx.empty()) gauss_legendre(128, gx, gw);
  const double half = 0.5 * kNuW, beta = kNuBetaPerW * kNuW;
  std::vector<double> z(gx.size()), pw(gx.size());
  for (size_t q = 0; q < gx.size(); ++q) {
    z[q] = 0.5 * half * (gx[q] + 1.);
    const double u = 2. * z[q] / kNuW;
    pw[q] = std::exp(beta * (std::sqrt(std::max(0., 1. - u * u)) - 1.)) * gw[q] * 0.5 * half;
  }
  d.resize(n);
  const int n21 = n / 2 - 1;
  for (int a = 0; a < n; ++a) {
    double k;  // make_freq (warp.c:316-360)
    if (a <= n21) k = a;
    else if (n % 2) k = (a == n21 + 1) ? n / 2 : (a == n21 + 2) ? -(n / 2) : -(n21 - (a - (n21 + 3)));
    else k = (a == n21 + 1) ? n / 2 : -(n21 - (a - (n21 + 2)));
    const double xi = 2. * M_PI * k / (double)M;
    double acc = 0.;
    for (size_t q = 0; q < z.size(); ++q) acc += pw[q] * std::cos(xi * z[q]);
    d[a] = 1. / (2. * acc);
  }
}

static int log2_grid(int n) {  // smallest power of two >= 2 n
  int l = 0;
  while ((1 << l) < 2 * n) ++l;
  return l;
}

static bool nufft_fits(int nx, int ny) {
  return nx >= 16 && ny >= 16 && log2_grid(nx) <= kNuMaxLog2 && log2_grid(ny) <= kNuMaxLog2;
}

// NUFFT tables of an nx x ny warp (dcx | dcy | twx | twy as doubles), built
// and uploaded once per image size; synchronous, before any stream work
static int nufft_prepare(flct_hip_ctx* c, int nx, int ny) {
  if (c->nu_nx == nx && c->nu_ny == ny && c->nuTab.p) return FLCT_HIP_OK;
  const int Mx = 1 << log2_grid(nx), My = 1 << log2_grid(ny);
  std::vector<double> dcx, dcy;
  nufft_deconv(nx, Mx, dcx);
  nufft_deconv(ny, My, dcy);
  std::vector<double> tab;
  tab.reserve(nx + ny + Mx + My);
  tab.insert(tab.end(), dcx.begin(), dcx.end());
  tab.insert(tab.end(), dcy.begin(), dcy.end());
  for (int e = 0; e < Mx / 2; ++e) {
    tab.push_back(std::cos(2. * M_PI * e / Mx));
    tab.push_back(std::sin(2. * M_PI * e / Mx));
  }
  for (int e = 0; e < My / 2; ++e) {
    tab.push_back(std::cos(2. * M_PI * e / My));
    tab.push_back(std::sin(2. * M_PI * e / My));
  }
  int rc;
  if ((rc = ensure(c->nuTab, tab.size() * 8))) return rc;
  HIPCHK(hipMemcpy(c->nuTab.p, tab.data(), tab.size() * 8, hipMemcpyHostToDevice));
  c->nu_nx = nx;
  c->nu_ny = ny;
  return FLCT_HIP_OK;
}

// warp_frac2d as a type-2 NUFFT on the half spectrum c->wF (already formed;
// tables from nufft_prepare)
static int warp_nufft(flct_hip_ctx* c, const double* delx, const double* dely, double* out, int nx,
                      int ny, int transp) {
  hipStream_t s = c->stream;
  const int lmx = log2_grid(nx), lmy = log2_grid(ny);
  const int Mx = 1 << lmx, My = 1 << lmy;
  const size_t otx = (size_t)nx + ny, oty = otx + Mx;
  int rc;
  if ((rc = ensure(c->nuT, (size_t)nx * My * 16)) || (rc = ensure(c->nuR, (size_t)My * Mx * 8)))
    return rc;
  const double* tb = (const double*)c->nuTab.p;
  NufftArgs q;
  q.F = (const double*)c->wF.p;
  q.nx = nx;
  q.ny = ny;
  q.nb = ny / 2 + 1;
  q.lmx = lmx;
  q.lmy = lmy;
  q.dcx = tb;
  q.dcy = tb + nx;
  q.twx = reinterpret_cast<const double2*>(tb + otx);
  q.twy = reinterpret_cast<const double2*>(tb + oty);
  q.T = (double2*)c->nuT.p;
  q.Rt = (double*)c->nuR.p;
  q.delx = delx;
  q.dely = dely;
  q.transp = transp;
  q.out = out;
  const size_t ldsr = (size_t)My * 16, ldsc = (size_t)Mx * 16;
  if (ldsr > c->lds_per_block || ldsc > c->lds_per_block) return FLCT_HIP_E_UNSUPPORTED;
  (void)hipFuncSetAttribute((const void*)k_nufft_rows, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ldsr);
  (void)hipFuncSetAttribute((const void*)k_nufft_cols, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)ldsc);
  hipLaunchKernelGGL(k_nufft_rows, dim3(nx), dim3(kNuThreads), ldsr, s, q);
  hipLaunchKernelGGL(k_nufft_cols, dim3(My), dim3(kNuThreads), ldsc, s, q);
  const long long tiles = (long long)((nx + kNuTile - 1) / kNuTile) * ((ny + kNuTile - 1) / kNuTile);
  hipLaunchKernelGGL(k_nufft_gather, dim3((unsigned)tiles), dim3(kNuTile * kNuTile), 0, s, q);
  HIPCHK(hipGetLastError());
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_set_warp_method(flct_hip_ctx* c, int method) {
  if (!c || method < FLCT_HIP_WARP_AUTO || method > FLCT_HIP_WARP_NUFFT) return FLCT_HIP_E_ARG;
  c->warp_method = method;
  return FLCT_HIP_OK;
}

extern "C" int flct_hip_last_warp_method(flct_hip_ctx* c) { return c ? c->last_warp_method : 0; }

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
  const bool nufft = (c->warp_method == FLCT_HIP_WARP_NUFFT) ||
                     (c->warp_method == FLCT_HIP_WARP_AUTO && nufft_fits(nx, ny));
  if (nufft) {
    if (!nufft_fits(nx, ny)) return FLCT_HIP_E_UNSUPPORTED;
    if ((rc = nufft_prepare(c, nx, ny))) return rc;
  }
  // row-transform tables (cached), before the timed region
  {
    const flct_hip_ctx::DftTab* t;
    if (ny <= (1 << (kRowDftMaxLog2 - 1))) (void)dft_tables(c, ny, -1, &t);
    if (nx <= (1 << (kRowDftMaxLog2 - 1))) (void)dft_tables(c, nx, -1, &t);
  }
  HIPCHK(hipEventRecord(c->ev[0], s));
  // rows: Frow[i][b] = sum_j arr[i][j] e(-j b / ny), b = 0..ny/2
  if ((rc = row_dft(c, arr, ny, 1, true, ny, nb, nx, (double*)c->wtmp.p, nb, 1, -1,
                    (const double*)c->wtw1.p)))
    return rc;
  // columns: F[a][b] = sum_i Frow[i][b] e(-i a / nx)
  if ((rc = row_dft(c, (const double*)c->wtmp.p, 1, nb, false, nx, nx, nb, (double*)c->wF.p, 1, nb,
                    -1, (const double*)c->wtw2.p)))
    return rc;
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(c->ev[1], s));
  if (nufft) {
    if ((rc = warp_nufft(c, delx, dely, out, nx, ny, transp))) return rc;
    HIPCHK(hipEventRecord(c->ev[2], s));
    HIPCHK(hipStreamSynchronize(s));
    c->last_warp_method = FLCT_HIP_WARP_NUFFT;
    c->last_kernel_ms = ev_ms(c->ev[1], c->ev[2]);
    c->last_total_ms = ev_ms(c->ev[0], c->ev[2]);
    return FLCT_HIP_OK;
  }
  c->last_warp_method = FLCT_HIP_WARP_DIRECT;
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
  if ((rc = row_dft(c, (const double*)c->f1.p, ny, 1, true, ny, ny, nx, A, ny, 1, -1,
                    (const double*)c->wtw1.p)) ||
      (rc = row_dft(c, A, 1, ny, false, nx, nx, ny, B, 1, ny, -1, (const double*)c->wtw2.p)))
    return rc;
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
  if ((rc = row_dft(c, B, ny, 1, false, ny, ny, nx, A, ny, 1, +1, (const double*)c->wtw1.p)) ||
      (rc = row_dft(c, A, 1, ny, false, nx, nx, ny, B, 1, ny, +1, (const double*)c->wtw2.p)))
    return rc;
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
