/*
 * flct_oracle.c -- CPU restatement of the FLCT 1.01 per-pixel local
 * correlation tracker and of warp's Fourier shifter, in FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark load it.
 * The product path (libflct_hip.so) never links or calls it.
 *
 * Every function follows a named piece of the reference (file:line in
 * Cadair/pyflct, source/flct.c and source/warp.c).  The reference links
 * FFTW3, which is absent here; the DFTs below are this file's own
 * mixed-radix Cooley-Tukey transforms (unnormalised, e^{-2 pi i} forward),
 * which is the contract the reference's fftw_plan_dft_* calls rely on.
 *
 * Parity pins (checked in tests/test_oracle.py):
 *   - oracle_flct on tests/golden/hashgauss.dat with "1 1 5 -k 0.5"
 *     reproduces tests/golden/testgaussvel.dat at float32;
 *   - oracle_warp_frac2d(hashgauss f1; deltaxygauss) reproduces hashgauss f2;
 *   - the FFTs against numpy.fft and a brute-force DFT;
 *   - the peak fit on sampled quadratics, interpcc2d on quadratics and
 *     hand-computed border cases, box-size and make_freq closed forms;
 *   - invariants (identical images, negated images, swapped images,
 *     integer-shift warp == roll, uniform warp == shift_frac2d).
 *   - the -h peak (oracle_peak_hires) on sampled quadratics, whose cubic
 *     convolution is exact, and its all-zero and edge rules;
 *   - the -b window-offset bias correction (no reference counterpart) by a
 *     direct-sum restatement of its definition on boxes 8 and 12, by exact
 *     cases (d = 0 is the plain pass bit for bit; a zero-sum integer blob
 *     moved by an integer s, given d = s, returns s; swapping the images
 *     negates the velocity) and by four planted mistakes.
 *
 * Build: gcc -O2 -ffp-contract=off -fPIC -shared (oracle/Makefile).
 *
 * PLANT_* macros (never set in the normal build) compile deliberately wrong
 * variants; the test suite builds them to show that the pins above fail
 * on a plausible mistake.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define XMISS (-999999.)
/* pi literal used by warp.c:893 and warp.c:770 */
#define WARP_PI 3.1415926535898

/* ------------------------------------------------------------------ */
/* FFT: recursive mixed-radix decimation in time with a master table. */
/* ------------------------------------------------------------------ */

typedef struct {
  int n;
  int nf;
  int fac[40];
  double *cs; /* cos(2 pi e / n), e = 0..n-1 */
  double *sn; /* sin(2 pi e / n) */
} fft_plan_t;

static void plan_init(fft_plan_t *p, int n) {
  int m = n, i;
  p->n = n;
  p->nf = 0;
  while (m > 1) {
    int r;
    if (m % 4 == 0) r = 4;
    else if (m % 2 == 0) r = 2;
    else if (m % 3 == 0) r = 3;
    else if (m % 5 == 0) r = 5;
    else {
      r = 7;
      while (m % r) r += 2;
    }
    p->fac[p->nf++] = r;
    m /= r;
  }
  p->cs = (double *)malloc(sizeof(double) * (size_t)n);
  p->sn = (double *)malloc(sizeof(double) * (size_t)n);
  for (i = 0; i < n; i++) {
    /* exact symmetric evaluation of the roots of unity */
    double a = 2.0 * M_PI * (double)i / (double)n;
    p->cs[i] = cos(a);
    p->sn[i] = sin(a);
  }
}

static void plan_free(fft_plan_t *p) {
  free(p->cs);
  free(p->sn);
  p->cs = p->sn = NULL;
}

/* out[k] = sum_t x[t*stride] * exp(sign*2*pi*i*t*k/n), n = product of fac[] */
static void fft_rec(const fft_plan_t *P, int n, const double *x, int stride,
                    double *out, const int *fac, int sign) {
  int r, m, q, k, pp;
  int N = P->n;
  if (n == 1) {
    out[0] = x[0];
    out[1] = x[1];
    return;
  }
  r = fac[0];
  m = n / r;
  for (q = 0; q < r; q++)
    fft_rec(P, m, x + 2 * q * stride, stride * r, out + 2 * q * m, fac + 1,
            sign);
  {
    double tr[r], ti[r], yr[r], yi[r]; /* r <= n: VLAs sized by the radix */
    int scale = N / n;
    for (k = 0; k < m; k++) {
      for (q = 0; q < r; q++) {
        double ar = out[2 * (q * m + k)], ai = out[2 * (q * m + k) + 1];
        int e = ((q * k) % n) * scale;
        double c = P->cs[e], s = sign * P->sn[e];
        tr[q] = ar * c - ai * s;
        ti[q] = ar * s + ai * c;
      }
      if (r == 2) {
        yr[0] = tr[0] + tr[1]; yi[0] = ti[0] + ti[1];
        yr[1] = tr[0] - tr[1]; yi[1] = ti[0] - ti[1];
      } else if (r == 4) {
        double s0r = tr[0] + tr[2], s0i = ti[0] + ti[2];
        double d0r = tr[0] - tr[2], d0i = ti[0] - ti[2];
        double s1r = tr[1] + tr[3], s1i = ti[1] + ti[3];
        double d1r = tr[1] - tr[3], d1i = ti[1] - ti[3];
        /* multiply d1 by (sign*i): forward sign=-1 -> -i */
        double jr = -sign * d1i, ji = sign * d1r;
        yr[0] = s0r + s1r; yi[0] = s0i + s1i;
        yr[2] = s0r - s1r; yi[2] = s0i - s1i;
        yr[1] = d0r + jr;  yi[1] = d0i + ji;
        yr[3] = d0r - jr;  yi[3] = d0i - ji;
      } else {
        for (pp = 0; pp < r; pp++) {
          double sr = 0., si = 0.;
          for (q = 0; q < r; q++) {
            int e = ((q * pp) % r) * (N / r);
            double c = P->cs[e], s = sign * P->sn[e];
            sr += tr[q] * c - ti[q] * s;
            si += tr[q] * s + ti[q] * c;
          }
          yr[pp] = sr;
          yi[pp] = si;
        }
      }
      for (pp = 0; pp < r; pp++) {
        out[2 * (k + m * pp)] = yr[pp];
        out[2 * (k + m * pp) + 1] = yi[pp];
      }
    }
  }
}

/* in-place 1D complex transform of n points with element stride (in complex
 * units); tmp must hold 4n doubles */
static void fft1(const fft_plan_t *P, double *a, int estride, int sign,
                 double *tmp) {
  int n = P->n, t;
  double *x = tmp, *y = tmp + 2 * n;
  for (t = 0; t < n; t++) {
    x[2 * t] = a[2 * t * estride];
    x[2 * t + 1] = a[2 * t * estride + 1];
  }
  fft_rec(P, n, x, 1, y, P->fac, sign);
  for (t = 0; t < n; t++) {
    a[2 * t * estride] = y[2 * t];
    a[2 * t * estride + 1] = y[2 * t + 1];
  }
}

/* in-place 2D complex transform of an nx x ny row-major array */
static void fft2(const fft_plan_t *Px, const fft_plan_t *Py, double *a, int sign,
                 double *tmp) {
  int nx = Px->n, ny = Py->n, i, j;
  for (i = 0; i < nx; i++) fft1(Py, a + 2 * i * ny, 1, sign, tmp);
  for (j = 0; j < ny; j++) fft1(Px, a + 2 * j, ny, sign, tmp);
}

/* exported for tests: 2D complex DFT, sign -1 forward, +1 backward, unnormalised */
int oracle_fft2(double *a, int nx, int ny, int sign) {
  fft_plan_t px, py;
  double *tmp = (double *)malloc(sizeof(double) * 4 * (size_t)(nx > ny ? nx : ny));
  plan_init(&px, nx);
  plan_init(&py, ny);
  fft2(&px, &py, a, sign, tmp);
  plan_free(&px);
  plan_free(&py);
  free(tmp);
  return 0;
}

/* ------------------------------------------------------------------ */
/* helpers from flct.c                                                 */
/* ------------------------------------------------------------------ */

/* make_freq: flct.c:1651-1692, warp.c:316-360 (Nyquist = +n/2) */
int oracle_make_freq(double *k, int ndim) {
  int n21 = (ndim / 2) - 1, i, inext;
  for (i = 0; i < n21 + 1; i++) k[i] = (double)i;
  inext = n21 + 1;
#ifdef PLANT_NYQUIST_NEG
  if (1) {
    k[inext] = -(double)(ndim / 2);
    inext++;
    if ((ndim / 2) * 2 != ndim) {
      k[inext] = -(double)(ndim / 2);
      inext++;
    }
  } else
#endif
  if ((ndim / 2) * 2 != ndim) {
    k[inext] = (double)(ndim / 2);
    inext++;
    k[inext] = -(double)(ndim / 2);
    inext++;
  } else {
    k[inext] = (double)(ndim / 2);
    inext++;
  }
  for (i = inext; i < ndim; i++) k[i] = -(double)(n21 - (i - inext));
  return 0;
}

/* gaussfilt: flct.c:1694-1715 */
static void gaussfilt(double *filter, const double *kx, const double *ky, int nx,
                      int ny, double kr) {
  double kxmax = kx[nx / 2], kymax = ky[ny / 2];
  double kxroll = kr * kxmax, kyroll = kr * kymax;
  double smxinv = 1. / kxroll, smyinv = 1. / kyroll;
  int i, j;
  for (i = 0; i < nx; i++) {
    double argx = kx[i] * smxinv;
    for (j = 0; j < ny; j++) {
      double argy = ky[j] * smyinv;
      filter[i * ny + j] = exp(-(argx * argx + argy * argy));
    }
  }
}

/* maxloc: flct.c:2142-2161, first strict maximum */
static int maxloc(const double *arr, int n) {
  double amax = arr[0];
  int loc = 0, i;
  for (i = 1; i < n; i++)
    if (arr[i] > amax) {
      amax = arr[i];
      loc = i;
    }
  return loc;
}

/* box size: flct.c:567-580.  Returns nsize (box is 2*nsize), or -1 when the
 * reference would exit with "illegal box size". */
int oracle_box_nsize(double sigma, int nx, int ny) {
  double tol = 1e-2;
  int nt, ndmin, nsize;
  nt = (int)sigma * sqrt(log(1. / tol));
  ndmin = (nx < ny) ? (((nx / 3) / 2)) * 2 : (((ny / 3) / 2)) * 2;
  nsize = ((nt / 2) * 2 < ndmin) ? (nt / 2) * 2 : ndmin;
  if (nsize <= 0) return -1;
  return nsize;
}

/* ------------------------------------------------------------------ */
/* cross_cor: flct.c:1325-1649 (hires = -1, expand = 1 branch)         */
/* ------------------------------------------------------------------ */

typedef struct {
  int nx, ny;
  fft_plan_t px, py;
  double *z;       /* 2 nx ny: packed complex input / spectrum */
  double *c;       /* 2 nx ny: cross-power spectrum, full plane */
  double *absc;    /* nx ny */
  double *shifted; /* nx ny */
  double *filter;  /* nx ny (flct.c:1396-1403) */
  double *tmp;
  int filterflag;
} xcor_ws_t;

static void ws_init(xcor_ws_t *w, int nx, int ny, int filterflag, double kr) {
  int nmax = nx > ny ? nx : ny;
  w->nx = nx;
  w->ny = ny;
  plan_init(&w->px, nx);
  plan_init(&w->py, ny);
  w->z = (double *)malloc(sizeof(double) * 2 * (size_t)nx * ny);
  w->c = (double *)malloc(sizeof(double) * 2 * (size_t)nx * ny);
  w->absc = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  w->shifted = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  w->filter = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  w->tmp = (double *)malloc(sizeof(double) * 4 * (size_t)nmax);
  w->filterflag = filterflag;
  if (filterflag) {
    double *kx = (double *)malloc(sizeof(double) * nx);
    double *ky = (double *)malloc(sizeof(double) * ny);
    oracle_make_freq(kx, nx);
    oracle_make_freq(ky, ny);
    gaussfilt(w->filter, kx, ky, nx, ny, kr);
    free(kx);
    free(ky);
  }
}

static void ws_free(xcor_ws_t *w) {
  plan_free(&w->px);
  plan_free(&w->py);
  free(w->z);
  free(w->c);
  free(w->absc);
  free(w->shifted);
  free(w->filter);
  free(w->tmp);
}

/* Correlation surface |c2r(conj(A) B normfac)| shifted by (nx/2, ny/2), where
 * A = r2c(arr), B = r2c(barr) (flct.c:1421-1510).  Result in w->shifted. */
static void xcor_surface(xcor_ws_t *w, const double *arr, const double *barr) {
  int nx = w->nx, ny = w->ny, i, j, nh = ny / 2 + 1;
  double normfac;
  double *z = w->z, *c = w->c;
  /* both real inputs packed into one complex transform */
  for (i = 0; i < nx * ny; i++) {
    z[2 * i] = arr[i];
    z[2 * i + 1] = barr[i];
  }
  fft2(&w->px, &w->py, z, -1, w->tmp);
  normfac = (1. / ((double)nx * ny));
  normfac *= normfac;
  /* half spectrum l = 0..ny/2 (r2c layout), then Hermitian extension */
  for (i = 0; i < nx; i++) {
    int im = (nx - i) % nx;
    for (j = 0; j < nh; j++) {
      int jm = (ny - j) % ny;
      double zr = z[2 * (i * ny + j)], zi = z[2 * (i * ny + j) + 1];
      double mr = z[2 * (im * ny + jm)], mi = -z[2 * (im * ny + jm) + 1];
      /* A = (Z + conj(Zm))/2, B = (Z - conj(Zm))/(2i) */
      double ar = 0.5 * (zr + mr), ai = 0.5 * (zi + mi);
      double br = 0.5 * (zi - mi), bi = -0.5 * (zr - mr);
      double cr, ci;
      if (w->filterflag) {
        /* flct.c:1445-1457: both spectra multiplied by filter[i*ny+j] */
        double f = w->filter[i * ny + j];
        ar = ar * f; ai = ai * f;
#ifndef PLANT_FILTER_ONCE
        br = br * f; bi = bi * f;
#endif
      }
      /* flct.c:1461-1477 */
      cr = (ar * br + ai * bi) * normfac;
      ci = (ar * bi - ai * br) * normfac;
      c[2 * (i * ny + j)] = cr;
      c[2 * (i * ny + j) + 1] = ci;
    }
  }
  /* c2r: columns l > ny/2 from Hermitian symmetry C[k][l] = conj(C[-k][-l]) */
  for (i = 0; i < nx; i++) {
    int im = (nx - i) % nx;
    for (j = nh; j < ny; j++) {
      int jm = ny - j;
      c[2 * (i * ny + j)] = c[2 * (im * ny + jm)];
      c[2 * (i * ny + j) + 1] = -c[2 * (im * ny + jm) + 1];
    }
  }
  fft2(&w->px, &w->py, c, +1, w->tmp);
  /* flct.c:1485-1488: absolute value of the real c2r output */
  for (i = 0; i < nx * ny; i++) w->absc[i] = fabs(c[2 * i]);
  /* shift2d: flct.c:2111-2140 */
  for (i = 0; i < nx; i++) {
    int ii = (i + nx / 2) % nx;
    for (j = 0; j < ny; j++) {
      int jj = (j + ny / 2) % ny;
      w->shifted[ii * ny + jj] = w->absc[i * ny + j];
    }
  }
}

/* peak + 3x3 Taylor fit: flct.c:1512-1556, 1639-1640 */
static void xcor_peak(const double *a, int nx, int ny, double *shiftx,
                      double *shifty) {
  int maxind, ixmax, iymax, absccmax = 1;
  double shiftx0, shifty0, shiftsubx = 0., shiftsuby = 0.;
#ifdef PLANT_TIE_LAST
  {
    int q;
    maxind = 0;
    for (q = 1; q < nx * ny; q++)
      if (a[q] >= a[maxind]) maxind = q;
  }
#else
  maxind = maxloc(a, nx * ny);
#endif
  if (a[maxind] == (double)0.) absccmax = 0;
  if (absccmax == 1) {
    ixmax = maxind / ny;
    iymax = maxind % ny;
  } else {
    ixmax = nx / 2;
    iymax = ny / 2;
  }
  shiftx0 = ixmax;
  shifty0 = iymax;
  if ((ixmax > 0) && (ixmax < (nx - 1)) && (iymax > 0) && (iymax < (ny - 1)) &&
      (absccmax == 1)) {
    double fx, fy, fxx, fyy, fxy;
    fx = 0.5 * (a[(ixmax + 1) * ny + iymax] - a[(ixmax - 1) * ny + iymax]);
    fy = 0.5 * (a[ixmax * ny + iymax + 1] - a[ixmax * ny + iymax - 1]);
    fxx = (a[(ixmax + 1) * ny + iymax] + a[(ixmax - 1) * ny + iymax] -
           2. * (a[ixmax * ny + iymax]));
    fyy = (a[ixmax * ny + iymax + 1] + a[ixmax * ny + iymax - 1] -
           2. * (a[ixmax * ny + iymax]));
    fxy = 0.25 * (a[(ixmax + 1) * ny + iymax + 1] + a[(ixmax - 1) * ny + iymax - 1] -
                  a[(ixmax + 1) * ny + iymax - 1] - a[(ixmax - 1) * ny + iymax + 1]);
#ifdef PLANT_FIT_SIGN
    fxy = -fxy;
#endif
    shiftsubx = (fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy);
    shiftsuby = (fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy);
  }
  *shiftx = (shiftx0 + shiftsubx) - (double)(nx / 2);
  *shifty = (shifty0 + shiftsuby) - (double)(ny / 2);
}

/* exported: one cross_cor call on two nx x ny arrays (hires=-1, expand=1).
 * absccor (optional) receives the shifted |ccor| surface. */
int oracle_cross_cor(const double *arr, const double *barr, int nx, int ny,
                     int filterflag, double kr, double *shiftx, double *shifty,
                     double *absccor) {
  xcor_ws_t w;
  ws_init(&w, nx, ny, filterflag, kr);
  xcor_surface(&w, arr, barr);
  xcor_peak(w.shifted, nx, ny, shiftx, shifty);
  if (absccor) memcpy(absccor, w.shifted, sizeof(double) * (size_t)nx * ny);
  ws_free(&w);
  return 0;
}

/* exported: the Taylor fit alone on an nx x ny surface (already shifted) */
int oracle_peak_fit(const double *a, int nx, int ny, double *shiftx,
                    double *shifty) {
  xcor_peak(a, nx, ny, shiftx, shifty);
  return 0;
}

/* ------------------------------------------------------------------ */
/* interpcc2d: flct.c:2226-2463                                        */
/* ------------------------------------------------------------------ */

static double keys(double t) {
  return (t >= (double)1.0)
             ? (((((double)(-0.5)) * t + ((double)2.5)) * t) - (double)4.) * t +
                   (double)2.
             : (((double)(1.5)) * t - ((double)(2.5))) * t * t + (double)1.;
}

int oracle_interpcc2d(const double *fdata, double xmiss, int nx, int ny,
                      const double *xwant, int nxinterp, const double *ywant,
                      int nyinterp, double *finterp) {
  double *cdata = (double *)malloc(sizeof(double) * (size_t)(nx + 2) * (ny + 2));
  int i, j, ii, jj;
  for (i = 0; i < nx; i++)
    for (j = 0; j < ny; j++) cdata[(i + 1) * (ny + 2) + (j + 1)] = fdata[i * ny + j];
  for (j = 0; j < ny; j++) {
    cdata[0 * (ny + 2) + (j + 1)] =
        fdata[2 * ny + j] - 3. * (fdata[1 * ny + j]) + 3. * (fdata[0 * ny + j]);
    cdata[(nx + 1) * (ny + 2) + (j + 1)] = fdata[(nx - 3) * ny + j] -
                                           3. * (fdata[(nx - 2) * ny + j]) +
                                           3. * (fdata[(nx - 1) * ny + j]);
  }
  for (i = 0; i < nx; i++) {
    cdata[(i + 1) * (ny + 2) + 0] =
        fdata[i * ny + 2] - 3. * (fdata[i * ny + 1]) + 3. * (fdata[i * ny + 0]);
    cdata[(i + 1) * (ny + 2) + ny + 1] = fdata[i * ny + ny - 3] -
                                         3. * (fdata[i * ny + ny - 2]) +
                                         3. * (fdata[i * ny + ny - 1]);
  }
  /* corners (flct.c:2313-2336); the first uses row stride nx+2 on index 0 */
  cdata[0] = 3. * (cdata[1 * (ny + 2) + 0]) - 3. * (cdata[2 * (ny + 2) + 0]) +
             cdata[3 * (ny + 2) + 0];
  cdata[(nx + 1) * (ny + 2) + 0] = 3. * (cdata[nx * (ny + 2) + 0]) -
                                   3. * (cdata[(nx - 1) * (ny + 2) + 0]) +
                                   cdata[(nx - 2) * (ny + 2) + 0];
  cdata[0 * (ny + 2) + ny + 1] = 3. * (cdata[0 * (ny + 2) + ny]) -
                                 3. * (cdata[0 * (ny + 2) + ny - 1]) +
                                 cdata[0 * (ny + 2) + ny - 2];
  cdata[(nx + 1) * (ny + 2) + ny + 1] =
      3. * (cdata[nx * (ny + 2) + ny + 1]) -
      3. * (cdata[(nx - 1) * (ny + 2) + ny + 1]) + cdata[(nx - 2) * (ny + 2) + ny + 1];

  for (i = 0; i < nxinterp; i++) {
    double xint = xwant[i];
    int itemp = ((int)xint > 0) ? (int)xint : 0;
    int izero = (itemp < (nx - 2)) ? itemp : nx - 2;
    for (j = 0; j < nyinterp; j++) {
      double yint = ywant[j];
      if ((yint < 0.) || (yint > (double)(ny - 1)) || ((xint < 0) || (xint > (double)(nx - 1)))) {
        finterp[i * nyinterp + j] = xmiss;
      } else {
        int jtemp = ((int)yint > 0) ? (int)yint : 0;
        int jzero = (jtemp < (ny - 2)) ? jtemp : ny - 2;
        double ftmp = 0.;
        int databad = 0;
        for (ii = -1; ii < 3; ii++) {
          double tx = fabs(xint - (double)(izero + ii));
          double rx = keys(tx);
          for (jj = -1; jj < 3; jj++) {
            double ty = fabs(yint - (double)(jzero + jj));
            double ry = keys(ty);
            double cv = cdata[(izero + 1 + ii) * (ny + 2) + jzero + 1 + jj];
            ftmp = ftmp + cv * rx * ry;
            if (cv == xmiss) databad = 1;
          }
        }
        finterp[i * nyinterp + j] = databad ? xmiss : ftmp;
      }
    }
  }
  free(cdata);
  return 0;
}

/* peak of cross_cor with -h (hires = 0): the integer peak as in xcor_peak
 * (flct.c:1512-1534), no fit there (flct.c:1536), then interpcc2d of the
 * whole surface with xmiss = 0 on a 21 x 21 grid of 0.1 px about it
 * (flct.c:1357-1368, 1564-1590), maxloc of that grid (flct.c:1598-1600) and
 * the 3 x 3 fit there scaled to pixels (flct.c:1605-1631, 1639-1640). */
static void xcor_peak_hires(const double *a, int nx, int ny, double *shiftx,
                            double *shifty) {
  int nxinterp = 21, nyinterp = 21, nfgppergp = 10;
  int maxind, ixmax, iymax, absccmax = 1, i, j, maxfine, ixx, iyy;
  double shiftx0, shifty0, rangex, rangey, shiftsubx = 0., shiftsuby = 0.;
  double xwant[21], ywant[21], peakarea[21 * 21];
  maxind = maxloc(a, nx * ny);
  if (a[maxind] == (double)0.) absccmax = 0;
  if (absccmax == 1) {
    ixmax = maxind / ny;
    iymax = maxind % ny;
  } else {
    ixmax = nx / 2;
    iymax = ny / 2;
  }
  shiftx0 = ixmax;
  shifty0 = iymax;
  rangex = (double)(nxinterp - 1) / nfgppergp;
  rangey = (double)(nyinterp - 1) / nfgppergp;
  for (i = 0; i < nxinterp; i++)
    xwant[i] = ((((double)i) * rangex) / ((double)(nxinterp - 1))) - 0.5 * rangex + shiftx0;
  for (j = 0; j < nyinterp; j++)
    ywant[j] = ((((double)j) * rangey) / ((double)(nyinterp - 1))) - 0.5 * rangey + shifty0;
  oracle_interpcc2d(a, 0., nx, ny, xwant, nxinterp, ywant, nyinterp, peakarea);
  maxfine = maxloc(peakarea, nxinterp * nyinterp);
  ixx = maxfine / nyinterp;
  iyy = maxfine % nyinterp;
  if ((ixx > 0) && (ixx < (nxinterp - 1)) && (iyy > 0) && (iyy < (nyinterp - 1))) {
    const double *p = peakarea;
    double fx, fy, fxx, fyy, fxy;
    fx = 0.5 * (p[(ixx + 1) * nyinterp + iyy] - p[(ixx - 1) * nyinterp + iyy]);
    fy = 0.5 * (p[ixx * nyinterp + iyy + 1] - p[ixx * nyinterp + iyy - 1]);
    fxx = (p[(ixx + 1) * nyinterp + iyy] + p[(ixx - 1) * nyinterp + iyy] -
           2. * (p[ixx * nyinterp + iyy]));
    fyy = (p[ixx * nyinterp + iyy + 1] + p[ixx * nyinterp + iyy - 1] -
           2. * (p[ixx * nyinterp + iyy]));
    fxy = 0.25 * (p[(ixx + 1) * nyinterp + iyy + 1] + p[(ixx - 1) * nyinterp + iyy - 1] -
                  p[(ixx + 1) * nyinterp + iyy - 1] - p[(ixx - 1) * nyinterp + iyy + 1]);
#ifdef PLANT_HIRES_SCALE
    shiftsubx = ((fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy));
    shiftsuby = ((fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy));
#else
    shiftsubx = ((fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy)) * rangex / ((double)(nxinterp - 1));
    shiftsuby = ((fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy)) * rangey / ((double)(nyinterp - 1));
#endif
  }
  *shiftx = (xwant[ixx] + shiftsubx) - (double)(nx / 2);
  *shifty = (ywant[iyy] + shiftsuby) - (double)(ny / 2);
}

/* exported: the hires peak alone on an nx x ny surface (already shifted) */
int oracle_peak_hires(const double *a, int nx, int ny, double *shiftx, double *shifty) {
  xcor_peak_hires(a, nx, ny, shiftx, shifty);
  return 0;
}

/* the peak step of cross_cor for the -h setting */
static void cc_peak(int hires, const double *a, int nx, int ny, double *shiftx,
                    double *shifty) {
  if (hires)
    xcor_peak_hires(a, nx, ny, shiftx, shifty);
  else
    xcor_peak(a, nx, ny, shiftx, shifty);
}

/* ------------------------------------------------------------------ */
/* per-pixel driver: flct.c:567-1133                                   */
/* ------------------------------------------------------------------ */

typedef struct {
  double deltat, deltas, sigma;
  double thresh;
  int thresh_abs;
  int skip, poff, qoff, interp;
  int filter;
  double kr;
  int bias;
  int hires;
} oracle_opts_t;

/* status codes shared with include/flct_hip.h */
#define OR_OK 0
#define OR_E_ARG 1
#define OR_E_SKIP_NONPOS 2
#define OR_E_PQ_RANGE 3
#define OR_E_SKIP_TOO_BIG 4
#define OR_E_BOX 5
#define OR_E_UNSUPPORTED 6

typedef struct {
  int nx, ny, nsize, nbox;
  double sigminv;
  double *gtab; /* Gaussian table over offsets, (2nbox-1)^2, identical values
                   to gaussdata (flct.c:652-663) at the used offsets */
  double *g1, *g2;
  xcor_ws_t ws;
} track_ctx_t;

/* window value exp(-argx^2 - argy^2) at integer offsets (di, dj) from the
 * pixel: argx = sigminv*(double)di as in flct.c:657-660 */
static double gauss_off(const track_ctx_t *t, int di, int dj) {
  int nb = t->nbox;
  return t->gtab[(di + nb - 1) * (2 * nb - 1) + (dj + nb - 1)];
}

/* box of pixel (i, j) and its means: clamp flct.c:773-808, means
 * flct.c:815-827 (summed row-major over the clamped box) */
static void pixel_box(const track_ctx_t *t, const double *f1, const double *f2, int i, int j,
                      int *imin_out, int *jmin_out, double *f1bar_out, double *f2bar_out) {
  int nx = t->nx, ny = t->ny, nsize = t->nsize, ii, jj;
  int imin0, imax0, imin, jmin0, jmax0, jmin, isize, jsize;
  double f1bar = 0., f2bar = 0.;
  imin0 = (0 > (i - (nsize - 1))) ? 0 : i - (nsize - 1);
  imax0 = ((nx - 1) < (i + nsize)) ? nx - 1 : i + nsize;
  imin = (imax0 == nx - 1) ? nx - 1 - (2 * nsize - 1) : imin0;
  jmin0 = (0 > (j - (nsize - 1))) ? 0 : j - (nsize - 1);
  jmax0 = ((ny - 1) < (j + nsize)) ? ny - 1 : j + nsize;
  jmin = (jmax0 == ny - 1) ? ny - 1 - (2 * nsize - 1) : jmin0;
  isize = jsize = 2 * nsize;
  for (ii = 0; ii < isize; ii++)
    for (jj = 0; jj < jsize; jj++) {
      f1bar = f1bar + f1[(ii + imin) * ny + (jj + jmin)];
      f2bar = f2bar + f2[(ii + imin) * ny + (jj + jmin)];
    }
  f1bar = f1bar / ((double)isize * jsize);
  f2bar = f2bar / ((double)isize * jsize);
#ifdef PLANT_NO_MEAN
  f1bar = f2bar = 0.;
#endif
  *imin_out = imin;
  *jmin_out = jmin;
  *f1bar_out = f1bar;
  *f2bar_out = f2bar;
}

/* -b second pass (no reference counterpart; window-offset correction,
 * central-difference form): with the first estimate d = (dsx, dsy) in
 * surface units (rows, columns), image 1 is windowed by the Gaussian
 * centred at pixel - d/2 and image 2 by the one centred at pixel + d/2,
 * over the same box and with the same box means, and cross_cor runs again.
 * d = 0 reproduces the plain pass bit for bit (same exp arguments). */
static void bias_pass(track_ctx_t *t, const double *f1, const double *f2, int i, int j,
                      int imin, int jmin, double f1bar, double f2bar, int hires, double dsx,
                      double dsy, double *shiftx, double *shifty) {
  int ny = t->ny, isize = t->nbox, jsize = t->nbox, ii, jj;
  double cx = 0.5 * dsx, cy = 0.5 * dsy;
#ifdef PLANT_BIAS_FULL_D
  cx = dsx;
  cy = dsy;
#endif
#ifdef PLANT_BIAS_REMEAN
  {
    double s1 = 0., s2 = 0., w1s = 0., w2s = 0.;
    for (ii = 0; ii < isize; ii++)
      for (jj = 0; jj < jsize; jj++) {
        double dx = (double)(ii + imin - i), dy = (double)(jj + jmin - j);
        double a1 = t->sigminv * (dx + cx), a2 = t->sigminv * (dx - cx);
        double b1 = t->sigminv * (dy + cy), b2 = t->sigminv * (dy - cy);
        double w1 = exp(-a1 * a1 - b1 * b1), w2 = exp(-a2 * a2 - b2 * b2);
        s1 += w1 * f1[(ii + imin) * ny + (jj + jmin)];
        s2 += w2 * f2[(ii + imin) * ny + (jj + jmin)];
        w1s += w1;
        w2s += w2;
      }
    f1bar = s1 / w1s;
    f2bar = s2 / w2s;
  }
#endif
  for (ii = 0; ii < isize; ii++) {
    double dx = (double)(ii + imin - i);
    double a1 = t->sigminv * (dx + cx), a2 = t->sigminv * (dx - cx);
#ifdef PLANT_BIAS_SIGN
    a1 = t->sigminv * (dx - cx);
    a2 = t->sigminv * (dx + cx);
#endif
#ifdef PLANT_BIAS_ONE_IMAGE
    a1 = t->sigminv * dx;
#endif
    for (jj = 0; jj < jsize; jj++) {
      double dy = (double)(jj + jmin - j);
      double b1 = t->sigminv * (dy + cy), b2 = t->sigminv * (dy - cy);
#ifdef PLANT_BIAS_SIGN
      b1 = t->sigminv * (dy - cy);
      b2 = t->sigminv * (dy + cy);
#endif
#ifdef PLANT_BIAS_ONE_IMAGE
      b1 = t->sigminv * dy;
#endif
      double w1 = exp(-a1 * a1 - b1 * b1), w2 = exp(-a2 * a2 - b2 * b2);
      t->g1[ii * jsize + jj] = w1 * (f1[(ii + imin) * ny + (jj + jmin)] - f1bar);
      t->g2[ii * jsize + jj] = w2 * (f2[(ii + imin) * ny + (jj + jmin)] - f2bar);
    }
  }
  xcor_surface(&t->ws, t->g1, t->g2);
  cc_peak(hires, t->ws.shifted, isize, jsize, shiftx, shifty);
}

/* one pixel: box clamp (773-808), means (815-827), window (856-870),
 * cross_cor (876), velocity (901-905).  With bias != 0 the plain estimate
 * d feeds bias_pass once and its result replaces d. */
static void track_pixel(track_ctx_t *t, const double *f1, const double *f2, int i,
                        int j, const oracle_opts_t *o, double *vxx, double *vyy) {
  int ny = t->ny, ii, jj, imin, jmin, isize = t->nbox, jsize = t->nbox;
  double f1bar, f2bar, shiftx, shifty, deltinv = 1. / o->deltat;
  pixel_box(t, f1, f2, i, j, &imin, &jmin, &f1bar, &f2bar);
  for (ii = 0; ii < isize; ii++)
    for (jj = 0; jj < jsize; jj++) {
      double g = gauss_off(t, ii + imin - i, jj + jmin - j);
      t->g1[ii * jsize + jj] = g * (f1[(ii + imin) * ny + (jj + jmin)] - f1bar);
      t->g2[ii * jsize + jj] = g * (f2[(ii + imin) * ny + (jj + jmin)] - f2bar);
    }
  xcor_surface(&t->ws, t->g1, t->g2);
  cc_peak(o->hires, t->ws.shifted, isize, jsize, &shiftx, &shifty);
  if (o->bias)
    bias_pass(t, f1, f2, i, j, imin, jmin, f1bar, f2bar, o->hires, shiftx, shifty, &shiftx,
              &shifty);
  /* transp = 1 (flct.c:901-905) */
#ifdef PLANT_NO_SWAP
  *vxx = shiftx * deltinv * o->deltas;
  *vyy = shifty * deltinv * o->deltas;
#else
  *vxx = shifty * deltinv * o->deltas;
  *vyy = shiftx * deltinv * o->deltas;
#endif
}

static int track_ctx_init(track_ctx_t *t, int nx, int ny, const oracle_opts_t *o) {
  int nsize = oracle_box_nsize(o->sigma, nx, ny), nb, a, b;
  if (nsize < 0) return OR_E_BOX;
  t->nx = nx;
  t->ny = ny;
  t->nsize = nsize;
  t->nbox = nb = 2 * nsize;
  t->sigminv = 1. / o->sigma;
  t->gtab = (double *)malloc(sizeof(double) * (size_t)(2 * nb - 1) * (2 * nb - 1));
  for (a = -(nb - 1); a <= nb - 1; a++) {
    double argx = t->sigminv * (double)a;
    for (b = -(nb - 1); b <= nb - 1; b++) {
      double argy = t->sigminv * (double)b;
      t->gtab[(a + nb - 1) * (2 * nb - 1) + (b + nb - 1)] =
          exp(-argx * argx - argy * argy);
    }
  }
  t->g1 = (double *)malloc(sizeof(double) * (size_t)nb * nb);
  t->g2 = (double *)malloc(sizeof(double) * (size_t)nb * nb);
  ws_init(&t->ws, nb, nb, o->filter, o->kr);
  return OR_OK;
}

static void track_ctx_free(track_ctx_t *t) {
  free(t->gtab);
  free(t->g1);
  free(t->g2);
  ws_free(&t->ws);
}

/* option checks done by the reference CLI (flct.c:459-472, 537-538,
 * 559-563); returns the resolved (poff, qoff) through pointers */
static int check_opts(const oracle_opts_t *o, int nx, int ny, int *poff, int *qoff,
                      int *skipon) {
  int p = o->poff, q = o->qoff;
  *skipon = 0;
  if (o->skip != 0 || p != 0 || q != 0) {
    if (o->skip <= 0) return OR_E_SKIP_NONPOS;
    if ((abs(p) >= o->skip) || (abs(q) >= o->skip)) return OR_E_PQ_RANGE;
    *skipon = o->skip + abs(q) + abs(p);
  }
  if (p < 0) p = o->skip - abs(p);
  if (q < 0) q = o->skip - abs(q);
  if ((o->skip >= nx) || (o->skip >= ny)) return OR_E_SKIP_TOO_BIG;
  *poff = p;
  *qoff = q;
  return OR_OK;
}

/* relative threshold: flct.c:594-629 */
static double resolve_thresh(const double *f1, const double *f2, int n,
                             const oracle_opts_t *o) {
  double thresh = o->thresh;
  if ((thresh > 0.) && (thresh < 1.) && (!o->thresh_abs)) {
    double f1max = fabs(f1[0]), f2max = fabs(f2[0]), fmax;
    int i;
    for (i = 1; i < n; i++) {
      if (fabs(f1[i]) > f1max) f1max = fabs(f1[i]);
      if (fabs(f2[i]) > f2max) f2max = fabs(f2[i]);
    }
    fmax = (f1max > f2max) ? f1max : f2max;
    thresh *= fmax;
  }
  return thresh;
}

int oracle_resolve_thresh(const double *f1, const double *f2, int n, double thresh,
                          int thresh_abs, double *out) {
  oracle_opts_t o;
  memset(&o, 0, sizeof(o));
  o.thresh = thresh;
  o.thresh_abs = thresh_abs;
  *out = resolve_thresh(f1, f2, n, &o);
  return 0;
}

/* sigma = 0 (flct.c:361-365, 567-586, 666-675, 802-808, 920-947): one
 * cross_cor of the whole images with window 1 and the whole-image means
 * (summed row-major, flct.c:815-827); threshold and skip are ignored; one
 * velocity, written to vx[0], vy[0], vm[0]. */
static int flct_whole(const double *f1, const double *f2, int nx, int ny,
                      const oracle_opts_t *o, double *vx, double *vy, double *vm) {
  double f1bar = 0., f2bar = 0., shiftx, shifty, deltinv = 1. / o->deltat;
  double *g1 = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  double *g2 = (double *)malloc(sizeof(double) * (size_t)nx * ny);
  int ii, jj;
  for (ii = 0; ii < nx; ii++)
    for (jj = 0; jj < ny; jj++) {
      f1bar = f1bar + f1[ii * ny + jj];
      f2bar = f2bar + f2[ii * ny + jj];
    }
  f1bar = f1bar / ((double)nx * ny);
  f2bar = f2bar / ((double)nx * ny);
  for (ii = 0; ii < nx * ny; ii++) {
    g1[ii] = (double)1. * (f1[ii] - f1bar);
    g2[ii] = (double)1. * (f2[ii] - f2bar);
  }
  {
    xcor_ws_t w;
    ws_init(&w, nx, ny, o->filter, o->kr);
    xcor_surface(&w, g1, g2);
    cc_peak(o->hires, w.shifted, nx, ny, &shiftx, &shifty);
    ws_free(&w);
  }
  vx[0] = shifty * deltinv * o->deltas;
  vy[0] = shiftx * deltinv * o->deltas;
  vm[0] = 1.;
  free(g1);
  free(g2);
  return OR_OK;
}

/* The whole per-pixel loop plus skip fill and final reset
 * (flct.c:567-1133).  Outputs are C-space nx x ny (1 x 1 for sigma = 0). */
int oracle_flct(const double *f1, const double *f2, int nx, int ny, double deltat,
                double deltas, double sigma, double thresh_in, int thresh_abs,
                int skip, int poff_in, int qoff_in, int interp, int filter,
                double kr, int bias, int hires, double *vx, double *vy, double *vm) {
  oracle_opts_t o;
  track_ctx_t t;
  int poff, qoff, skipon, rc, i, j, xoffset, yoffset;
  double thresh;
  o.deltat = deltat; o.deltas = deltas; o.sigma = sigma;
  o.thresh = thresh_in; o.thresh_abs = thresh_abs;
  o.skip = skip; o.poff = poff_in; o.qoff = qoff_in; o.interp = interp;
  o.filter = filter; o.kr = kr; o.bias = bias; o.hires = hires;
  rc = check_opts(&o, nx, ny, &poff, &qoff, &skipon);
  if (rc) return rc;
  /* sigmaeq0 for any sigma that is not > 0 (flct.c:357-365) */
  if (!(sigma > 0.)) return flct_whole(f1, f2, nx, ny, &o, vx, vy, vm);
  rc = track_ctx_init(&t, nx, ny, &o);
  if (rc) return rc;
  thresh = resolve_thresh(f1, f2, nx * ny, &o);
  /* transp = 1: p offsets j, q offsets i (flct.c:736-740) */
  xoffset = qoff;
  yoffset = poff;
  for (i = 0; i < nx; i++) {
    for (j = 0; j < ny; j++) {
      double fabsbar = 0.5 * (fabs(f1[i * ny + j] + f2[i * ny + j]));
      int belowthresh = (fabsbar < thresh);
      int noskipxy = 1;
      double vxx = 0., vyy = 0., vmask = 1.;
      if (skipon) noskipxy = (!((i - xoffset) % skip)) * (!((j - yoffset) % skip));
      if ((!belowthresh) && noskipxy) track_pixel(&t, f1, f2, i, j, &o, &vxx, &vyy);
      if (belowthresh || !noskipxy) {
        vxx = XMISS;
        vyy = XMISS;
        vmask = 0.;
      }
      vx[i * ny + j] = vxx;
      vy[i * ny + j] = vyy;
      vm[i * ny + j] = vmask;
    }
  }
  if (skipon && interp) {
    int ixcount = 0, iycount = 0, ic, jc;
    double *vxnoi, *vynoi, *xwant, *ywant, *vxint, *vyint;
    for (i = 0; i < nx; i++) if (!((i - xoffset) % skip)) ixcount++;
    for (j = 0; j < ny; j++) if (!((j - yoffset) % skip)) iycount++;
    vxnoi = (double *)malloc(sizeof(double) * (size_t)ixcount * iycount);
    vynoi = (double *)malloc(sizeof(double) * (size_t)ixcount * iycount);
    ic = -1;
    for (i = 0; i < nx; i++) {
      int nsx = !((i - xoffset) % skip);
      if (nsx) ic++;
      jc = -1;
      for (j = 0; j < ny; j++) {
        int nsy = !((j - yoffset) % skip);
        if (nsy) jc++;
        if (nsx && nsy) {
          vxnoi[ic * iycount + jc] = vx[i * ny + j];
          vynoi[ic * iycount + jc] = vy[i * ny + j];
        }
      }
    }
    xwant = (double *)malloc(sizeof(double) * nx);
    ywant = (double *)malloc(sizeof(double) * ny);
    for (i = 0; i < nx; i++) xwant[i] = ((double)(i - xoffset)) / (double)(skip);
    for (j = 0; j < ny; j++) ywant[j] = ((double)(j - yoffset)) / (double)(skip);
    vxint = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    vyint = (double *)malloc(sizeof(double) * (size_t)nx * ny);
    oracle_interpcc2d(vxnoi, XMISS, ixcount, iycount, xwant, nx, ywant, ny, vxint);
    oracle_interpcc2d(vynoi, XMISS, ixcount, iycount, xwant, nx, ywant, ny, vyint);
    for (i = 0; i < nx; i++) {
      int nsx = !((i - xoffset) % skip);
      for (j = 0; j < ny; j++) {
        int nsy = !((j - yoffset) % skip);
        if (!(nsx && nsy)) {
          vx[i * ny + j] = vxint[i * ny + j];
          vy[i * ny + j] = vyint[i * ny + j];
          vm[i * ny + j] = 0.5;
        }
      }
    }
    free(vxnoi); free(vynoi); free(xwant); free(ywant); free(vxint); free(vyint);
  }
  /* flct.c:1125-1133 */
  for (i = 0; i < nx * ny; i++) {
    if (vx[i] == XMISS) vm[i] = 0.;
    if (vx[i] == XMISS) vx[i] = 0.;
    if (vy[i] == XMISS) vy[i] = 0.;
  }
  track_ctx_free(&t);
  return OR_OK;
}

/* Track only the listed pixels (C-space linear indices), with the same
 * threshold and skip rules as oracle_flct but no skip fill and no final
 * xmiss reset: outputs hold vx, vy (or XMISS) and vm in {0, 1} per listed
 * pixel.  Used for sampled parity checks and the CPU baseline. */
int oracle_flct_pixels(const double *f1, const double *f2, int nx, int ny,
                       double deltat, double deltas, double sigma, double thresh_in,
                       int thresh_abs, int skip, int poff_in, int qoff_in,
                       int filter, double kr, int bias, int hires, const long long *pix,
                       long long npix, double *vx, double *vy, double *vm) {
  oracle_opts_t o;
  track_ctx_t t;
  int poff, qoff, skipon, rc;
  long long p;
  double thresh;
  o.deltat = deltat; o.deltas = deltas; o.sigma = sigma;
  o.thresh = thresh_in; o.thresh_abs = thresh_abs;
  o.skip = skip; o.poff = poff_in; o.qoff = qoff_in; o.interp = 0;
  o.filter = filter; o.kr = kr; o.bias = bias; o.hires = hires;
  rc = check_opts(&o, nx, ny, &poff, &qoff, &skipon);
  if (rc) return rc;
  if (!(sigma > 0.)) return OR_E_UNSUPPORTED;  /* one whole-image velocity: oracle_flct */
  rc = track_ctx_init(&t, nx, ny, &o);
  if (rc) return rc;
  thresh = resolve_thresh(f1, f2, nx * ny, &o);
  for (p = 0; p < npix; p++) {
    int i = (int)(pix[p] / ny), j = (int)(pix[p] % ny);
    double fabsbar = 0.5 * (fabs(f1[i * ny + j] + f2[i * ny + j]));
    int belowthresh = (fabsbar < thresh), noskipxy = 1;
    double vxx = XMISS, vyy = XMISS, vmask = 0.;
    if (skipon) noskipxy = (!((i - qoff) % skip)) * (!((j - poff) % skip));
    if ((!belowthresh) && noskipxy) {
      track_pixel(&t, f1, f2, i, j, &o, &vxx, &vyy);
      vmask = 1.;
    }
    vx[p] = vxx;
    vy[p] = vyy;
    vm[p] = vmask;
  }
  track_ctx_free(&t);
  return OR_OK;
}

/* Test hook of the -b pins: the second pass alone, for pixel (i, j), with a
 * given offset d = (dsx, dsy) in surface units (rows, columns) instead of
 * the plain estimate.  Returns the surface shifts (before the transp swap
 * of flct.c:901-905).  d = 0 is the plain pass. */
int oracle_bias_pass(const double *f1, const double *f2, int nx, int ny, double sigma,
                     int filter, double kr, int hires, int i, int j, double dsx, double dsy,
                     double *shiftx, double *shifty) {
  oracle_opts_t o;
  track_ctx_t t;
  int imin, jmin, rc;
  double f1bar, f2bar;
  memset(&o, 0, sizeof(o));
  o.deltat = 1.; o.deltas = 1.; o.sigma = sigma; o.filter = filter; o.kr = kr;
  o.hires = hires;
  if (!(sigma > 0.) || i < 0 || i >= nx || j < 0 || j >= ny) return OR_E_ARG;
  rc = track_ctx_init(&t, nx, ny, &o);
  if (rc) return rc;
  pixel_box(&t, f1, f2, i, j, &imin, &jmin, &f1bar, &f2bar);
  bias_pass(&t, f1, f2, i, j, imin, jmin, f1bar, f2bar, hires, dsx, dsy, shiftx, shifty);
  track_ctx_free(&t);
  return OR_OK;
}

/* ------------------------------------------------------------------ */
/* warp.c                                                              */
/* ------------------------------------------------------------------ */

/* shift_frac2d: warp.c:757-881 */
int oracle_shift_frac2d(const double *arr, double delx, double dely, double *outarr,
                        int nx, int ny, int transp) {
  double pi = WARP_PI, normfac, dxarg, dyarg;
  double *a = (double *)malloc(sizeof(double) * 2 * (size_t)nx * ny);
  double *kx = (double *)malloc(sizeof(double) * nx), *ky = (double *)malloc(sizeof(double) * ny);
  double *crx = (double *)malloc(sizeof(double) * nx), *cix = (double *)malloc(sizeof(double) * nx);
  double *cry = (double *)malloc(sizeof(double) * ny), *ciy = (double *)malloc(sizeof(double) * ny);
  int i, j;
  oracle_make_freq(kx, nx);
  oracle_make_freq(ky, ny);
  normfac = (1. / ((double)nx * ny));
  for (i = 0; i < nx * ny; i++) {
    a[2 * i] = arr[i];
    a[2 * i + 1] = 0.;
  }
  oracle_fft2(a, nx, ny, -1);
  if (transp) {
    dxarg = -dely / ((double)nx);
    dyarg = -delx / ((double)ny);
  } else {
    dxarg = -delx / ((double)nx);
    dyarg = -dely / ((double)ny);
  }
  for (i = 0; i < nx; i++) {
    crx[i] = cos(2. * pi * kx[i] * dxarg);
    cix[i] = sin(2. * pi * kx[i] * dxarg);
  }
  for (j = 0; j < ny; j++) {
    cry[j] = cos(2. * pi * ky[j] * dyarg);
    ciy[j] = sin(2. * pi * ky[j] * dyarg);
  }
  for (i = 0; i < nx; i++)
    for (j = 0; j < ny; j++) {
      double er = crx[i] * cry[j] - cix[i] * ciy[j];
      double ei = crx[i] * ciy[j] + cix[i] * cry[j];
      double orr = a[2 * (i * ny + j)], oi = a[2 * (i * ny + j) + 1];
      a[2 * (i * ny + j)] = (orr * er - oi * ei);
      a[2 * (i * ny + j) + 1] = (orr * ei + oi * er);
    }
  oracle_fft2(a, nx, ny, +1);
  for (i = 0; i < nx * ny; i++) outarr[i] = (double)(a[2 * i]) * normfac;
  free(a); free(kx); free(ky); free(crx); free(cix); free(cry); free(ciy);
  return 0;
}

/* warp_frac2d: warp.c:883-1011.  rows [r0, r1) only (full image when
 * r0 = 0, r1 = nx); the sum order over (ii, jj) is the reference's. */
int oracle_warp_frac2d_rows(const double *arr, const double *delx,
                            const double *dely, double *outarr, int nx, int ny,
                            int transp, int r0, int r1) {
  double pi = WARP_PI, normfac, nxinv, nyinv;
  double *a = (double *)malloc(sizeof(double) * 2 * (size_t)nx * ny);
  double *kx = (double *)malloc(sizeof(double) * nx), *ky = (double *)malloc(sizeof(double) * ny);
  double *cskx = (double *)malloc(sizeof(double) * nx), *snkx = (double *)malloc(sizeof(double) * nx);
  double *csky = (double *)malloc(sizeof(double) * ny), *snky = (double *)malloc(sizeof(double) * ny);
  int i, j, ii, jj;
  oracle_make_freq(kx, nx);
  oracle_make_freq(ky, ny);
  normfac = (1. / ((double)nx * ny));
  nxinv = 1. / ((double)nx);
  nyinv = 1. / ((double)ny);
  for (i = 0; i < nx * ny; i++) {
    a[2 * i] = arr[i];
    a[2 * i + 1] = 0.;
  }
  oracle_fft2(a, nx, ny, -1);
  for (i = r0; i < r1; i++) {
    for (j = 0; j < ny; j++) {
      double shiftx, shifty, xarg, yarg, total = 0.;
      if (transp) {
        shiftx = dely[i * ny + j];
        shifty = delx[i * ny + j];
      } else {
        shiftx = delx[i * ny + j];
        shifty = dely[i * ny + j];
      }
      xarg = 2. * pi * (((double)i) - shiftx) * nxinv;
      yarg = 2. * pi * (((double)j) - shifty) * nyinv;
      for (ii = 0; ii < nx; ii++) {
        cskx[ii] = cos(kx[ii] * xarg);
        snkx[ii] = sin(kx[ii] * xarg);
      }
      for (jj = 0; jj < ny; jj++) {
        csky[jj] = cos(ky[jj] * yarg);
        snky[jj] = sin(ky[jj] * yarg);
      }
      for (ii = 0; ii < nx; ii++)
        for (jj = 0; jj < ny; jj++) {
          double er = cskx[ii] * csky[jj] - snkx[ii] * snky[jj];
          double ei = cskx[ii] * snky[jj] + snkx[ii] * csky[jj];
          total += (a[2 * (ii * ny + jj)] * er - a[2 * (ii * ny + jj) + 1] * ei);
        }
      outarr[i * ny + j] = (double)total * normfac;
    }
  }
  free(a); free(kx); free(ky); free(cskx); free(snkx); free(csky); free(snky);
  return 0;
}

int oracle_warp_frac2d(const double *arr, const double *delx, const double *dely,
                       double *outarr, int nx, int ny, int transp) {
  return oracle_warp_frac2d_rows(arr, delx, dely, outarr, nx, ny, transp, 0, nx);
}
