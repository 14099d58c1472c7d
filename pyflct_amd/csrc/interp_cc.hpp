// Cubic-convolution interpolation of interpcc2d (flct.c:2226-2463) over an
// nxl x nyl array read through an accessor at(i, j), and the hires peak
// refinement of cross_cor built on it (flct.c:1357-1368, 1564-1645).
// Host and device: the skip fill, the per-pixel tracker and the whole-image
// path share one statement of the arithmetic.
#pragma once

#include <hip/hip_runtime.h>

namespace flct {

// Keys cubic convolution weight, a = -0.5 (flct.c:2391-2395)
__host__ __device__ __forceinline__ double keys_w(double t) {
#pragma clang fp contract(off)
  return (t >= (double)1.0)
             ? (((((double)(-0.5)) * t + ((double)2.5)) * t) - (double)4.) * t + (double)2.
             : (((double)(1.5)) * t - ((double)(2.5))) * t * t + (double)1.;
}

// cdata of interpcc2d (flct.c:2266-2336): the array padded by one cell with
// the reference's edge and corner extrapolation, index (ci, cj) in
// [0, nxl+1] x [0, nyl+1]
template <class At>
__host__ __device__ double cc_cdata(const At& at, int nxl, int nyl, int ci, int cj) {
#pragma clang fp contract(off)
  const bool ie = (ci == 0 || ci == nxl + 1), je = (cj == 0 || cj == nyl + 1);
  if (!ie && !je) return at(ci - 1, cj - 1);
  if (ie && !je) {
    const int jc = cj - 1;
    if (ci == 0) return at(2, jc) - 3. * (at(1, jc)) + 3. * (at(0, jc));
    return at(nxl - 3, jc) - 3. * (at(nxl - 2, jc)) + 3. * (at(nxl - 1, jc));
  }
  if (!ie && je) {
    const int ic = ci - 1;
    if (cj == 0) return at(ic, 2) - 3. * (at(ic, 1)) + 3. * (at(ic, 0));
    return at(ic, nyl - 3) - 3. * (at(ic, nyl - 2)) + 3. * (at(ic, nyl - 1));
  }
  // corners, from the edge values in column 0 / nyl+1 (rows) or row 0 (flct.c:2313-2336)
  auto edge_col0 = [&](int c) {  // cdata[c][0], 1 <= c <= nxl
    const int ic = c - 1;
    return at(ic, 2) - 3. * (at(ic, 1)) + 3. * (at(ic, 0));
  };
  auto edge_colN = [&](int c) {  // cdata[c][nyl+1]
    const int ic = c - 1;
    return at(ic, nyl - 3) - 3. * (at(ic, nyl - 2)) + 3. * (at(ic, nyl - 1));
  };
  auto edge_row0 = [&](int c) {  // cdata[0][c], 1 <= c <= nyl
    const int jc = c - 1;
    return at(2, jc) - 3. * (at(1, jc)) + 3. * (at(0, jc));
  };
  if (ci == 0 && cj == 0) return 3. * (edge_col0(1)) - 3. * (edge_col0(2)) + edge_col0(3);
  if (ci == nxl + 1 && cj == 0)
    return 3. * (edge_col0(nxl)) - 3. * (edge_col0(nxl - 1)) + edge_col0(nxl - 2);
  if (ci == 0 && cj == nyl + 1)
    return 3. * (edge_row0(nyl)) - 3. * (edge_row0(nyl - 1)) + edge_row0(nyl - 2);
  return 3. * (edge_colN(nxl)) - 3. * (edge_colN(nxl - 1)) + edge_colN(nxl - 2);
}

// one interpolated value at (xint, yint) (flct.c:2343-2440): xmiss outside
// [0, nxl-1] x [0, nyl-1] or when any of the 16 cdata taps equals xmiss
template <class At>
__host__ __device__ double cc_interp(const At& at, int nxl, int nyl, double xint, double yint,
                                     double xmiss) {
#pragma clang fp contract(off)
  if ((yint < 0.) || (yint > (double)(nyl - 1)) || ((xint < 0) || (xint > (double)(nxl - 1))))
    return xmiss;
  const int itemp = ((int)xint > 0) ? (int)xint : 0;
  const int izero = (itemp < (nxl - 2)) ? itemp : nxl - 2;
  const int jtemp = ((int)yint > 0) ? (int)yint : 0;
  const int jzero = (jtemp < (nyl - 2)) ? jtemp : nyl - 2;
  double ftmp = 0.;
  bool databad = false;
  for (int ii = -1; ii < 3; ii++) {
    const double tx = fabs(xint - (double)(izero + ii));
    const double rx = keys_w(tx);
    for (int jj = -1; jj < 3; jj++) {
      const double ty = fabs(yint - (double)(jzero + jj));
      const double ry = keys_w(ty);
      const double c = cc_cdata(at, nxl, nyl, izero + 1 + ii, jzero + 1 + jj);
      ftmp = ftmp + c * rx * ry;
      if (c == xmiss) databad = true;
    }
  }
  return databad ? xmiss : ftmp;
}

// hires grid of cross_cor (flct.c:1357-1368, hires = 0 branch): 21 x 21
// points at 0.1 pixel spacing centred on the integer peak
constexpr int kHiresN = 21;
constexpr int kHiresPerPixel = 10;

// grid coordinate i of the hires grid about integer peak index p
// (flct.c:1569-1583): ((i * range) / (n - 1)) - 0.5 * range + p
__host__ __device__ __forceinline__ double hires_want(int i, int p) {
#pragma clang fp contract(off)
  const double range = (double)(kHiresN - 1) / kHiresPerPixel;
  return ((((double)i) * range) / ((double)(kHiresN - 1))) - 0.5 * range + (double)p;
}

// The surface window a hires refinement reads: rows ixmax-2..ixmax+3 and
// columns iymax-2..iymax+3 of the nx x ny surface (those inside it).  Every
// cdata tap of a grid point within one pixel of the peak, the edge and
// corner extrapolations included, falls in this window.
constexpr int kHiresWin = 6;

// accessor of the window: w[(i - ixmax + 2) * 6 + (j - iymax + 2)]
struct HiresWindow {
  const double* w;
  int ixmax, iymax;
  __host__ __device__ double operator()(int i, int j) const {
    return w[(i - ixmax + 2) * kHiresWin + (j - iymax + 2)];
  }
};

// value of hires grid point q = ix * 21 + iy (interpcc2d with xmiss = 0,
// flct.c:1589-1590); -0 is returned as +0 so that bit order is value order
__host__ __device__ __forceinline__ double hires_value(const HiresWindow& win, int nx, int ny,
                                                       int q) {
  const int ix = q / kHiresN, iy = q - (q / kHiresN) * kHiresN;
  const double v = cc_interp(win, nx, ny, hires_want(ix, win.ixmax), hires_want(iy, win.iymax), 0.);
  return v + 0.;
}

// The padded-array (cdata) values a hires grid reads: cdata rows
// ixmax-1..ixmax+4 and columns iymax-1..iymax+4 (cdata index = surface index
// + 1), kept as cp[(ci - ixmax + 1) * 6 + (cj - iymax + 1)].  Entry t of the
// 6 x 6 block from the window; entries outside [0, n+1]^2 are not needed.
__host__ __device__ __forceinline__ void hires_cdata_entry(const HiresWindow& win, int nx, int ny,
                                                           int t, double* cp) {
  const int ci = win.ixmax - 1 + t / kHiresWin, cj = win.iymax - 1 + t % kHiresWin;
  if (ci >= 0 && ci <= nx + 1 && cj >= 0 && cj <= ny + 1) cp[t] = cc_cdata(win, nx, ny, ci, cj);
}

// hires grid point q from the cdata block: the 16-tap sum of cc_interp in
// the same order, so the value equals hires_value's bit for bit
__host__ __device__ __forceinline__ double hires_value_cp(const double* cp, int nx, int ny,
                                                          int ixmax, int iymax, int q) {
#pragma clang fp contract(off)
  const int ix = q / kHiresN, iy = q - (q / kHiresN) * kHiresN;
  const double xint = hires_want(ix, ixmax), yint = hires_want(iy, iymax);
  if ((yint < 0.) || (yint > (double)(ny - 1)) || ((xint < 0) || (xint > (double)(nx - 1))))
    return 0.;
  const int itemp = ((int)xint > 0) ? (int)xint : 0;
  const int izero = (itemp < (nx - 2)) ? itemp : nx - 2;
  const int jtemp = ((int)yint > 0) ? (int)yint : 0;
  const int jzero = (jtemp < (ny - 2)) ? jtemp : ny - 2;
  const double* row = cp + (izero - ixmax + 1) * kHiresWin + (jzero - iymax + 1);
  double ftmp = 0.;
  bool databad = false;
  for (int ii = -1; ii < 3; ii++) {
    const double tx = fabs(xint - (double)(izero + ii));
    const double rx = keys_w(tx);
    for (int jj = -1; jj < 3; jj++) {
      const double ty = fabs(yint - (double)(jzero + jj));
      const double ry = keys_w(ty);
      const double c = row[(ii + 1) * kHiresWin + (jj + 1)];
      ftmp = ftmp + c * rx * ry;
      if (c == 0.) databad = true;
    }
  }
  return databad ? 0. : ftmp + 0.;
}

// total order of doubles as unsigned keys (larger value, larger key)
__host__ __device__ __forceinline__ unsigned long long hires_key(double v) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}

// shift of the hires peak: grid maximum at q (first strict maximum,
// flct.c:1598-1600) and its 3 x 3 neighbourhood nb[(dx+1)*3 + (dy+1)]
// (flct.c:1605-1628); the result is in surface index units, before the
// - n/2 of flct.c:1639-1640
__host__ __device__ __forceinline__ void hires_shift(int q, const double* nb, int ixmax, int iymax,
                                                     double* shx, double* shy) {
#pragma clang fp contract(off)
  const int ixx = q / kHiresN, iyy = q - (q / kHiresN) * kHiresN;
  const double range = (double)(kHiresN - 1) / kHiresPerPixel;
  double shiftsubx = 0., shiftsuby = 0.;
  if ((ixx > 0) && (ixx < (kHiresN - 1)) && (iyy > 0) && (iyy < (kHiresN - 1))) {
    const double c00 = nb[4];
    const double fx = 0.5 * (nb[7] - nb[1]);
    const double fy = 0.5 * (nb[5] - nb[3]);
    const double fxx = (nb[7] + nb[1] - 2. * (c00));
    const double fyy = (nb[5] + nb[3] - 2. * (c00));
    const double fxy = 0.25 * (nb[8] + nb[0] - nb[6] - nb[2]);
    shiftsubx = ((fyy * fx - fy * fxy) / (fxy * fxy - fxx * fyy)) * range / ((double)(kHiresN - 1));
    shiftsuby = ((fxx * fy - fx * fxy) / (fxy * fxy - fxx * fyy)) * range / ((double)(kHiresN - 1));
  }
  *shx = hires_want(ixx, ixmax) + shiftsubx;
  *shy = hires_want(iyy, iymax) + shiftsuby;
}

}  // namespace flct
