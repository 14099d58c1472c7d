/*
 * flct_hip.h -- C ABI of the MI355X (gfx950) FLCT tracker and Fourier warp.
 *
 * Arrays are C-space, row-major: element (i, j) of an nx x ny image is at
 * i*ny + j, where nx is the IDL y size and ny the IDL x size (the reference
 * runs with transp = 1, flct.c:291; the vcimage header is read swapped,
 * flct.c:1192-1196).  No torch or HIP types appear in these signatures.
 *
 * Every entry point returns a status code (FLCT_HIP_OK = 0).  Each code that
 * corresponds to an exit(1) site of the reference CLI is named after it.
 * A context is bound to one GPU and one HIP stream; calls on one context are
 * serialised by the caller (the reference is single-threaded).
 */
#ifndef FLCT_HIP_H
#define FLCT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define FLCT_HIP_OK 0
#define FLCT_HIP_E_ARG 1           /* bad argument (null pointer, size <= 0) */
#define FLCT_HIP_E_SKIP_NONPOS 2   /* flct.c:459-464  "skip =%d is zero or negative" */
#define FLCT_HIP_E_PQ_RANGE 3      /* flct.c:466-472  "abs(p) or abs(q) >= skip" */
#define FLCT_HIP_E_SKIP_TOO_BIG 4  /* flct.c:559-563  "skip = %d is too big" */
#define FLCT_HIP_E_BOX 5           /* flct.c:575-579  "illegal box size" */
#define FLCT_HIP_E_UNSUPPORTED 6   /* a request the library does not handle */
#define FLCT_HIP_E_NODEVICE 7      /* no usable gfx950 device */
#define FLCT_HIP_E_ALLOC 8         /* device allocation failed */
#define FLCT_HIP_E_HIP 9           /* a HIP runtime call failed */
#define FLCT_HIP_E_SHAPE 10        /* warp.c:182-188 "bad shift arrays" */
#define FLCT_HIP_E_NOSCRATCH 11    /* flct_hip_track_device_noalloc: the call needs
                                      more than its 256 MiB device array */

/* arithmetic forms of the tracker */
#define FLCT_HIP_FP64 0        /* plain form: FP64 throughout (drop-in) */
#define FLCT_HIP_FP32 1        /* FP32 transforms, FP64 means, fit and outputs;
                                  unsettled pixels re-evaluated in FP64 */
#define FLCT_HIP_SPLIT_BF16 2  /* DFTs as split-BF16 MFMA contractions, FP32 accumulate */

typedef struct flct_hip_ctx flct_hip_ctx;

/* Options of one flct run: the reference command line
 * "flct ifile ofile deltat deltas sigma [-t thr[a]] [-k kr] [-s N[pP][qQ][i]] [-h] [-b] [-q]"
 * (flct.c:304, flct.c:370-487). */
typedef struct flct_hip_opts {
  double deltat;     /* time between images */
  double deltas;     /* length unit of one pixel */
  double sigma;      /* Gaussian window width in pixels; any value not > 0
                        (0, negative) = one whole-image correlation
                        (flct.c:357-365, 567-586) */
  double thresh;     /* -t value (0: off) */
  int thresh_abs;    /* -t value carried an 'a' suffix: absolute units */
  int skip;          /* -s N (0: off) */
  int poff;          /* -s ...pP: offset along IDL x (C-space j) */
  int qoff;          /* -s ...qQ: offset along IDL y (C-space i) */
  int interp;        /* -s ...i: fill skipped pixels by cubic convolution */
  int filter;        /* -k given */
  double kr;         /* -k value */
  int bias;          /* -b: window-offset bias correction (not in the reference) */
  int precision;     /* FLCT_HIP_FP64, FLCT_HIP_FP32, FLCT_HIP_SPLIT_BF16 */
  int quiet;         /* -q: no progress messages */
  int hires;         /* -h: peak by cubic-convolution interpolation of the
                        correlation on a 0.1 px grid (flct.c:383, 1357-1368,
                        1564-1645); runs the FP64 kernels for every precision */
} flct_hip_opts;

/* Fill *opts with the reference defaults (flct.c:370-375): thresh 0, no
 * skip, no filter, FP64. */
void flct_hip_opts_default(flct_hip_opts *opts);

/* Human-readable text for a status code. */
const char *flct_hip_strerror(int status);

/* Validate opts for an nx x ny image on the host (no device needed): the
 * reference's option and box-size checks (flct.c:459-472, 559-563, 575-579).
 * On success *nbox (if non-null) receives the sliding-box size 2*nsize. */
int flct_hip_check_opts(int nx, int ny, const flct_hip_opts *opts, int *nbox);

/* Create a context on HIP device `device`.  Allocates nothing large. */
int flct_hip_create(flct_hip_ctx **ctx, int device);
int flct_hip_destroy(flct_hip_ctx *ctx);

/*
 * flct_hip_track replaces the per-pixel body of flct's main program,
 * flct.c:567-1133: box size, relative threshold, the i,j loop with its
 * threshold and skip rules, the windowed FFT cross-correlation of
 * cross_cor (flct.c:1325-1649, with or without -h), the skip fill by interpcc2d
 * (flct.c:958-1120, 2226-2463) and the final xmiss reset.
 * f1, f2: host arrays (nx*ny doubles, C-space).  vx, vy, vm: caller-allocated
 * host arrays of nx*ny doubles.  sigma = 0 correlates the two whole images
 * once with window 1 and the whole-image means, ignoring the threshold and
 * skip (flct.c:567-586, 666-675, 802-808, 920-947): the single velocity goes
 * to vx[0], vy[0], vm[0] (= 1) and the other elements are not written.
 */
int flct_hip_track(flct_hip_ctx *ctx, const double *f1, const double *f2, int nx, int ny,
                   const flct_hip_opts *opts, double *vx, double *vy, double *vm);

/* Same as flct_hip_track with every array already in device memory (HIP
 * device pointers); the call is ordered on the context's stream and returns
 * after the outputs are complete. */
int flct_hip_track_device(flct_hip_ctx *ctx, const double *d_f1, const double *d_f2, int nx,
                          int ny, const flct_hip_opts *opts, double *d_vx, double *d_vy,
                          double *d_vm);

/* flct_hip_track_device without any device allocation.  No counterpart in
 * the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Every table and scratch buffer of the call (box means and their row sums,
 * pixel lists, counters, transform tables, the planes of boxes above 96,
 * the row-segment kernels' per-block spectra) lives in one static 256 MiB
 * device array that each device holds from library load until the process
 * ends; the caller passes only the inputs and outputs.  A call whose
 * scratch exceeds the array returns FLCT_HIP_E_NOSCRATCH (every pixel
 * tracked fits up to about 2900 x 2900 pixels in FP64, 2600 x 2600 in
 * split-BF16); sigma <= 0 returns FLCT_HIP_E_UNSUPPORTED.  All
 * calls on one device share the array: calls that overlap on one device,
 * from any context or thread, race.  Results equal flct_hip_track_device's
 * for the same inputs and options. */
int flct_hip_track_device_noalloc(flct_hip_ctx *ctx, const double *d_f1, const double *d_f2,
                                  int nx, int ny, const flct_hip_opts *opts, double *d_vx,
                                  double *d_vy, double *d_vm);

/* One image pair of a batch: device arrays (C-space, nx*ny doubles) and the
 * options of its own flct run; tracked, refined and status are outputs. */
typedef struct flct_hip_pair {
  const double *f1, *f2;
  double *vx, *vy, *vm;
  int nx, ny;
  flct_hip_opts opts;
  long long tracked;  /* out: pixels whose cross-correlation was computed */
  long long refined;  /* out: split-BF16 pixels re-evaluated in FP64 */
  int status;         /* out: FLCT_HIP_OK or the pair's option error */
} flct_hip_pair;

/*
 * flct_hip_track_batch runs flct_hip_track_device for every pair: the
 * equivalent of npairs invocations of the reference flct (flct.c:261-1154,
 * one pair per process), each with its own size and options.  All pairs'
 * kernels are captured into one HIP graph on the context's stream and
 * launched once.  Inside the graph the pairs run on up to four concurrent
 * branches (largest estimated work first, each pair to the least loaded
 * branch), each branch with its own scratch; so pairs run in no fixed
 * order, and a pair's output arrays must not overlap any other pair's input
 * or output arrays.  The graph is kept and relaunched while the next call
 * brings the same pairs (arrays, sizes, options).  Every pair is validated
 * first: if any is invalid, nothing runs, each pair's status is set and the
 * first error is returned.  flct_hip_last_stats reports the sum of the
 * tracked pixels and the device time of the graph.  A pair with sigma = 0
 * is rejected with FLCT_HIP_E_UNSUPPORTED (call flct_hip_track_device).
 */
int flct_hip_track_batch(flct_hip_ctx *ctx, flct_hip_pair *pairs, int npairs);

/* Number of pixels whose cross-correlation the last track call computed
 * (not thresholded, not skipped), and the device time of its tracking
 * kernel in milliseconds (HIP events on the context's stream). */
int flct_hip_last_stats(flct_hip_ctx *ctx, long long *tracked, double *kernel_ms,
                        double *total_ms);

/* Number of pixels the last FLCT_HIP_SPLIT_BF16 or FLCT_HIP_FP32 call
 * re-evaluated in FP64 because its correlation surface could not settle them
 * (a second cell within 1e-4 (split-BF16) or 1e-5 (FP32) of the maximum, or
 * an ill-conditioned peak fit).  The FP64 row-shared path ("fp64-rowseg")
 * counts the pixels it hands to the quad-FFT kernels: boxes constant at an
 * inexact mean, or whose structure is below 1e-4 of their distance from the
 * segment's reference mean; the quad-FFT FP64 path reports 0.  No
 * counterpart in the reference, which has one precision. */
int flct_hip_last_refined(flct_hip_ctx *ctx, long long *refined);

/* Device time in milliseconds of the last track call's tracking kernel alone
 * (main_ms) and of the FP64 re-evaluation of flagged pixels (refine_ms, 0
 * for FLCT_HIP_FP64); HIP events on the context's stream.  No counterpart in
 * the reference. */
int flct_hip_last_kernel_ms(flct_hip_ctx *ctx, double *main_ms, double *refine_ms);

/* Kernel path of the last track call: "fp64-fft" and "fp32-fft" (quad-FFT
 * kernels, one pixel per 4-lane group), "fp64-rowseg" and "fp32-rowseg"
 * (row-shared kernels: the box-row transforms of a row segment computed
 * once, FP64 in both), "bf16-segment" and "bf16-pixel" (split-BF16 MFMA
 * kernels) or "fp64-plane" (boxes above 96).  A precision form may run on
 * any path whose error is within the form's contract; FP64 serves every
 * form.  No counterpart in the reference. */
const char *flct_hip_last_path(flct_hip_ctx *ctx);

/* Runtime shape of the last call's main tracking kernel: *slots, the
 * pixels tracked together per block by the FFT kernels (fp64-fft, fp32-fft,
 * fp64-rowseg, fp32-rowseg), and *segpx, the pixels per row segment of the
 * row-segment kernels (bf16-segment, fp64-rowseg, fp32-rowseg); 0 stands
 * for the compiled shape (its slot count, 64 pixels per segment).  No
 * counterpart in the reference. */
int flct_hip_last_shape(flct_hip_ctx *ctx, int *slots, int *segpx);

/* Autotuner: runs every kernel path that meets the precision form of opts
 * on the given device inputs, times the whole call with HIP events, keeps
 * the fastest, then times that path's runtime shapes (half, a quarter and
 * one of its compiled slots per block; segments of 32 and 16 pixels) and
 * makes the fastest configuration the context's for that form and box size
 * (until the context is destroyed or tuned again).  Untuned, the split-BF16 form
 * runs the FP32 kernels for boxes up to 20 and above 64 and the row-segment
 * kernel from 24 to 64 (measured fastest there); its per-pixel MFMA kernels
 * are eligible up to box 72, above which their error can exceed the form's
 * contract.  Untuned calls without a threshold, skip lattice, -b or -h run
 * the row-shared kernel of their FP64 or FP32 precision on the box sizes
 * where it was measured faster than the quad-FFT one.  sigma = 0 has one
 * path: nothing to tune.  No counterpart in the reference. */
int flct_hip_autotune(flct_hip_ctx *ctx, const double *d_f1, const double *d_f2, int nx, int ny,
                      const flct_hip_opts *opts);

/* Pins the kernel path of one precision form and box size in the context,
 * as flct_hip_autotune would: path is one of the names flct_hip_last_path
 * returns and must meet the form's error contract (FLCT_HIP_E_ARG
 * otherwise); NULL or "" restores the untuned default.  No counterpart in
 * the reference. */
int flct_hip_set_path(flct_hip_ctx *ctx, int precision, int nbox, const char *path);

/* flct_hip_set_path with a runtime shape: slots 0 (compiled count) or 1 up
 * to the path's compiled slots per block, segpx 0 (64) or 8..64 on the
 * row-segment paths; FLCT_HIP_E_ARG for a shape the path does not have.
 * No counterpart in the reference. */
int flct_hip_set_config(flct_hip_ctx *ctx, int precision, int nbox, const char *path, int slots,
                        int segpx);

/*
 * flct_hip_warp replaces warp_frac2d (warp.c:883-1011): out(i,j) is the
 * image's Fourier series evaluated at (i - shiftx, j - shifty), where with
 * transp != 0 shiftx = dely[i*ny+j], shifty = delx[i*ny+j] (warp.c:948-957).
 * Host arrays of nx*ny doubles; out caller-allocated.
 */
int flct_hip_warp(flct_hip_ctx *ctx, const double *arr, const double *delx,
                  const double *dely, double *outarr, int nx, int ny, int transp, int quiet);

/* Evaluation method of flct_hip_warp: the reference's direct sum of the
 * Fourier series at every pixel (O(nx ny) work per pixel, warp.c:960-991),
 * or the same series as a type-2 non-uniform FFT (a 2x oversampled grid,
 * 16 x 16 interpolation taps per pixel), which differs from the direct sum
 * by about 1e-13 of max|image|.  AUTO (the default) takes the NUFFT for
 * 16 <= nx, ny <= 4096 and the direct sum otherwise.  No counterpart in the
 * reference. */
#define FLCT_HIP_WARP_AUTO 0
#define FLCT_HIP_WARP_DIRECT 1
#define FLCT_HIP_WARP_NUFFT 2
int flct_hip_set_warp_method(flct_hip_ctx *ctx, int method);
/* Method the last warp call ran (FLCT_HIP_WARP_DIRECT or _NUFFT; 0 before any). */
int flct_hip_last_warp_method(flct_hip_ctx *ctx);

/* Device-pointer form of flct_hip_warp. */
int flct_hip_warp_device(flct_hip_ctx *ctx, const double *d_arr, const double *d_delx,
                         const double *d_dely, double *d_out, int nx, int ny, int transp);

/*
 * flct_hip_shift replaces shift_frac2d (warp.c:757-881): a uniform
 * sub-pixel shift by a Fourier phase ramp.  Host arrays.
 */
int flct_hip_shift(flct_hip_ctx *ctx, const double *arr, double delx, double dely,
                   double *outarr, int nx, int ny, int transp, int quiet);

/* Device memory helpers for callers of the *_device entry points (no
 * counterpart in the reference, which has no device).  Allocations belong to
 * the context's device; copies are ordered on the context's stream and
 * complete before return. */
int flct_hip_malloc(flct_hip_ctx *ctx, void **dptr, long long bytes);
int flct_hip_free(flct_hip_ctx *ctx, void *dptr);
int flct_hip_upload(flct_hip_ctx *ctx, void *dptr, const void *src, long long bytes);
int flct_hip_download(flct_hip_ctx *ctx, void *dst, const void *dptr, long long bytes);

#ifdef __cplusplus
}
#endif

#endif /* FLCT_HIP_H */
