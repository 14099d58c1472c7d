/*
 * warp -- drop-in command line for the MI355X Fourier shifter.
 *
 *   warp ifile sfile ofile [-q]
 *
 * ifile holds one image; sfile two arrays (delx, dely).  Arrays of the
 * image's size warp it pixel by pixel (warp_frac2d, warp.c:164-172); 1 x 1
 * arrays shift it uniformly (shift_frac2d, warp.c:173-181); anything else is
 * an error (warp.c:182-188).  Like the reference, -q is honoured only as the
 * fourth argument (warp.c:115-123).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/flct_hip.h"
#include "vcimage_io.h"

int main(int argc, char *argv[]) {
  int quiet = 0, nx, ny, nxd, nyd, rc;
  double *img[1], *del[2], *out;
  flct_hip_ctx *ctx = NULL;
  if ((argc < 4) || (argc > 6)) {
    printf("warp (MI355X build): shift or warp an image by Fourier interpolation\n\n");
    printf("usage: %s ifile sfile ofile [-q]\n", argv[0]);
    printf("  ifile  one image; sfile  delx, dely (image-sized arrays, or 1 x 1 for a uniform shift)\n");
    printf("  ofile  the shifted image; -q  quiet\n");
    exit(1);
  }
  if (argc == 5 && !strncmp("-q", argv[argc - 1], 2)) quiet = 1;
  if (vc_read("readimage", argv[1], 1, &nx, &ny, img)) exit(1);
  if (!quiet) printf("warp: from input file, nx = %d, ny = %d\n", nx, ny);
  if (vc_read("read2images", argv[2], 2, &nxd, &nyd, del)) exit(1);
  out = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  rc = flct_hip_create(&ctx, 0);
  if (rc) {
    printf("warp: %s\n", flct_hip_strerror(rc));
    exit(1);
  }
  if (nxd == nx && nyd == ny) {
    rc = flct_hip_warp(ctx, img[0], del[0], del[1], out, nx, ny, 1, quiet);
  } else if (nxd == 1 && nyd == 1) {
    rc = flct_hip_shift(ctx, img[0], del[0][0], del[1][0], out, nx, ny, 1, quiet);
  } else {
    printf("warp: bad shift arrays: nx=%d, nxdel=%d, ny=%d, nydel=%d\n", nx, nxd, ny, nyd);
    exit(1);
  }
  if (rc) {
    printf("warp: %s\n", flct_hip_strerror(rc));
    exit(1);
  }
  {
    const double *outs[1] = {out};
    if (vc_write("writeimage", argv[3], 1, nx, ny, outs)) exit(1);
  }
  flct_hip_destroy(ctx);
  if (!quiet) printf("warp: finished\n");
  return 0;
}
