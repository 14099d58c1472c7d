/*
 * flct -- drop-in command line for the MI355X tracker.
 *
 *   flct ifile ofile deltat deltas sigma [-t thr[a]] [-k kr] [-s N[pP][qQ][i]] [-h] [-b] [-q]
 *
 * Same argument grammar as the reference program (flct.c:298-487): options
 * in any order after the five required arguments, matched on their first two
 * characters; -t with a trailing 'a' is absolute; -s is parsed right to left
 * (i, then q, then p, then N).  The input file holds two vcimage arrays, the
 * output file vx, vy, vm (flct.c:1138).  Errors print a message and exit(1)
 * like the reference.  -b (bias correction) is an addition.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/flct_hip.h"
#include "vcimage_io.h"

static void usage(const char *prog) {
  printf("flct (MI355X build): local correlation tracking between two images\n\n");
  printf("usage: %s ifile ofile deltat deltas sigma [-t thr[a]] [-k kr] [-s N[pP][qQ][i]] [-h] [-b] [-q]\n\n",
         prog);
  printf("  ifile   two images (vcimage format)\n");
  printf("  ofile   vx, vy, vm (vm: 1 computed, 0 not computed, 0.5 interpolated)\n");
  printf("  deltat  time between the images; deltas  length of one pixel\n");
  printf("  sigma   width in pixels of the Gaussian window; 0 (or less) correlates the whole images once\n");
  printf("  -t thr  skip pixels whose mean |image| is below thr; 0 < thr < 1 is\n");
  printf("          relative to the largest |pixel| unless followed by 'a'\n");
  printf("  -k kr   Gaussian low-pass of the sub-images, kr in units of the Nyquist wavenumber\n");
  printf("  -s NpPqQ  track every N-th pixel (offsets P in x, Q in y); 'i' interpolates the rest\n");
  printf("  -h      peak by cubic-convolution interpolation on a 0.1 pixel grid\n");
  printf("  -b      window-offset bias correction\n");
  printf("  -q      quiet\n");
}

int main(int argc, char *argv[]) {
  char buf[100];
  flct_hip_opts o;
  int iarg, quiet = 0, hires = 0, rc, nx = 0, ny = 0, nbox = 0, aflag = 0;
  double *in[2] = {NULL, NULL};
  double *vx, *vy, *vm;
  flct_hip_ctx *ctx = NULL;
  if ((argc < 6) || (argc > 15)) {
    usage(argv[0]);
    exit(1);
  }
  flct_hip_opts_default(&o);
  o.deltat = atof(argv[3]);
  o.deltas = atof(argv[4]);
  o.sigma = atof(argv[5]);
  for (iarg = 6; iarg < argc; iarg++) {
    if (!strncmp("-h", argv[iarg], 2)) hires = 1;
    if (!strncmp("-q", argv[iarg], 2)) quiet = 1;
    if (!strncmp("-b", argv[iarg], 2)) o.bias = 1;
    if (!strncmp("-t", argv[iarg], 2) && (iarg + 1) < argc) {
      char *a;
      strncpy(buf, argv[iarg + 1], 99);
      buf[99] = '\0';
      a = strchr(buf, 'a');
      if (a) *a = '\0';
      o.thresh = atof(buf);
      o.thresh_abs = (a != NULL);
      aflag = (a != NULL);
    }
    if (!strncmp("-s", argv[iarg], 2) && (iarg + 1) < argc) {
      char *p;
      strncpy(buf, argv[iarg + 1], 99);
      buf[99] = '\0';
      o.skip = o.poff = o.qoff = o.interp = 0;
      if ((p = strchr(buf, 'i')) != NULL) {
        o.interp = 1;
        *p = '\0';
      }
      if ((p = strchr(buf, 'q')) != NULL) {
        o.qoff = atoi(p + 1);
        *p = '\0';
      }
      if ((p = strchr(buf, 'p')) != NULL) {
        o.poff = atoi(p + 1);
        *p = '\0';
      }
      o.skip = atoi(buf);
      if (o.skip <= 0) {
        printf("flct: skip =%d is zero or negative, fatal\n", o.skip);
        exit(1);
      }
      if (abs(o.poff) >= o.skip || abs(o.qoff) >= o.skip) {
        printf("flct: p=%d,q=%d, abs(p) or abs(q) >= skip=%d\n", o.poff, o.qoff, o.skip);
        exit(1);
      }
    }
    if (!strncmp("-k", argv[iarg], 2) && (iarg + 1) < argc) {
      o.kr = atof(argv[iarg + 1]);
      o.filter = 1;
    }
  }
  o.quiet = quiet;
  o.hires = hires;
  if (!quiet) {
    printf("flct: MI355X build\n");
    printf("flct: infile = %s\nflct: outfile = %s\n", argv[1], argv[2]);
    printf("flct: deltat = %g\nflct: deltas = %g\nflct: sigma = %g\n", o.deltat, o.deltas, o.sigma);
    printf("flct: threshold image value for LCT is %g\n", o.thresh);
    if (aflag) printf("flct: threshold forced to be in abs. units\n");
    if (o.skip) printf("flct: skip = %d pixels with p=%d, q=%d\n", o.skip, o.poff, o.qoff);
    if (o.interp) printf("flct: skipped pixels interpolated with cubic convolution\n");
    if (o.filter) printf("flct: filter rolloff value for input images is %g\n", o.kr);
    if (o.hires) printf("flct: hires option turned on\n");
    if (o.bias) printf("flct: bias correction on\n");
  }
  if (vc_read("read2images", argv[1], 2, &nx, &ny, in)) exit(1);
  if (!quiet) printf("flct: from input file, nx = %d, ny = %d\n", nx, ny);
  rc = flct_hip_check_opts(nx, ny, &o, &nbox);
  if (rc == FLCT_HIP_E_SKIP_TOO_BIG) {
    printf("flct: skip = %d is too big compared to nx or ny, fatal\n", o.skip);
    exit(1);
  }
  if (rc == FLCT_HIP_E_BOX) {
    printf("flct: error - illegal box size, exiting\n");
    exit(1);
  }
  if (rc) {
    printf("flct: %s\n", flct_hip_strerror(rc));
    exit(1);
  }
  if (!quiet && o.sigma > 0.) printf("flct: nominal sliding box size = %d\n", nbox);
  vx = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  vy = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  vm = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  if (!vx || !vy || !vm) {
    printf("flct: memory request failed\n");
    exit(1);
  }
  rc = flct_hip_create(&ctx, 0);
  if (rc) {
    printf("flct: %s\n", flct_hip_strerror(rc));
    exit(1);
  }
  rc = flct_hip_track(ctx, in[0], in[1], nx, ny, &o, vx, vy, vm);
  if (rc) {
    printf("flct: %s\n", flct_hip_strerror(rc));
    exit(1);
  }
  if (!quiet) {
    long long tracked = 0;
    double kms = 0., tms = 0.;
    flct_hip_last_stats(ctx, &tracked, &kms, &tms);
    printf("flct: tracked %lld pixels in %.3f ms on the GPU\n", tracked, tms);
  }
  if (!(o.sigma > 0.)) {
    /* one velocity for the whole image, written as 1 x 1 arrays (flct.c:581-586) */
    if (!quiet) printf("\nflct: vx = %g vy = %g \n", vx[0], vy[0]);
    nx = 1;
    ny = 1;
  }
  {
    const double *outs[3] = {vx, vy, vm};
    if (vc_write("write3images", argv[2], 3, nx, ny, outs)) exit(1);
  }
  flct_hip_destroy(ctx);
  free(in[0]);
  free(in[1]);
  free(vx);
  free(vy);
  free(vm);
  if (!quiet) printf("flct: finished\n");
  return 0;
}
