/*
 * grid_oracle.c -- CPU restatement of the reference's grid generation from a
 * layer file (gd_curv_gen_layer, gd_t.c:987-1480, with gd_grid_z_interp
 * :1483-1562 and the cubic spline of gd_SPLine / gd_seval / gd_SPL
 * :1564-1700).  TEST INFRASTRUCTURE ONLY: the checker of the engine's
 * cgfdh_gen_grid_layer; nothing in the product links it.  gd_t.c needs
 * netCDF and is not compiled here, so this restatement is pinned by
 * mathematics (tests/test_grid_oracle.py): the spline against the published
 * Forsythe-Malcolm-Moler method in float64 and its exactness on cubics, the
 * vertical spacing by its closed forms and invariants, flat layers by the
 * Cartesian grid they give.  Every expression keeps the reference's
 * float/double mix and order; compiled without FMA contraction.
 */
#include <math.h>
#include <stdlib.h>

/* gd_SPLine with natural-free ends (end1 = end2 = 0): the spline
 * coefficients b, c, d of y(x); 0, or 1 (n < 2) / 2 (x not ascending) */
int orc_spline(int n, const float *x, const float *y, float *b, float *c, float *d)
{
  if (n < 2) return 1;
  for (int i = 1; i < n; ++i)
    if (x[i] <= x[i - 1]) return 2;
  const int nm1 = n - 1;
  if (n >= 3) {
    d[0] = x[1] - x[0];
    c[1] = (y[1] - y[0]) / d[0];
    for (int i = 1; i < nm1; ++i) {
      d[i] = x[i + 1] - x[i];
      b[i] = 2.0 * (d[i - 1] + d[i]);
      c[i + 1] = (y[i + 1] - y[i]) / d[i];
      c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[nm1] = -d[n - 2];
    c[0] = 0.0;
    c[nm1] = 0.0;
    if (n != 3) {
      c[0] = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0]);
      c[nm1] = c[n - 2] / (x[nm1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4]);
      c[0] = c[0] * d[0] * d[0] / (x[3] - x[0]);
      c[nm1] = -c[nm1] * d[n - 2] * d[n - 2] / (x[nm1] - x[n - 4]);
    }
    for (int i = 1; i < n; ++i) {
      float t = d[i - 1] / b[i - 1];
      b[i] = b[i] - t * d[i - 1];
      c[i] = c[i] - t * c[i - 1];
    }
    c[nm1] = c[nm1] / b[nm1];
    for (int ib = 0; ib < nm1; ++ib) {
      int i = n - ib - 2;
      c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    }
    b[nm1] = (y[nm1] - y[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2.0 * c[nm1]);
    for (int i = 0; i < nm1; ++i) {
      b[i] = (y[i + 1] - y[i]) / d[i] - d[i] * (c[i + 1] + 2.0 * c[i]);
      d[i] = (c[i + 1] - c[i]) / d[i];
      c[i] = 3.0 * c[i];
    }
    c[nm1] = 3.0 * c[nm1];
    d[nm1] = d[n - 2];
  } else {
    b[0] = (y[1] - y[0]) / (x[1] - x[0]);
    c[0] = 0.0;
    d[0] = 0.0;
    b[1] = b[0];
    c[1] = 0.0;
    d[1] = 0.0;
  }
  return 0;
}

/* gd_seval: the spline at u, starting the interval search from *last */
float orc_spline_eval(float u, int n, const float *x, const float *y, const float *b,
                      const float *c, const float *d, int *last)
{
  int i = *last;
  if (i >= n - 1) i = 0;
  if (i < 0) i = 0;
  if ((x[i] > u) || (x[i + 1] < u)) {
    int j = n;
    i = 0;
    do {
      int k = (i + j) / 2;
      if (u < x[k]) j = k;
      if (u >= x[k]) i = k;
    } while (j > i + 1);
  }
  *last = i;
  float w = u - x[i];
  w = y[i] + w * (b[i] + w * (c[i] + w * d[i]));
  return w;
}

/* gd_SPL: the spline through (x, y) at xi[0..ni), in order */
int orc_spline_at(int n, const float *x, const float *y, int ni, const float *xi, float *yi)
{
  float *b = malloc(3 * sizeof(float) * (n > 2 ? n : 2));
  float *c = b + (n > 2 ? n : 2), *d = c + (n > 2 ? n : 2);
  int rc = orc_spline(n, x, y, b, c, d), last = 0;
  if (!rc)
    for (int i = 0; i < ni; i++) yi[i] = orc_spline_eval(xi[i], n, x, y, b, c, d, &last);
  free(b);
  return rc;
}

/* gd_grid_z_interp: the nz_interp z values of one column from its interface
 * heights; equal spacing in a layer, or for a middle layer marked unequal
 * whose neighbours' spacings bracket its own, two runs of linearly changing
 * spacing that join its neighbours' */
void orc_grid_z(float *z, const float *zl, const int *ncell, const int *equal, int nlay)
{
  float dz[nlay > 0 ? nlay : 1];
  int nmax = 0, sum = 0;
  for (int i = 0; i < nlay; i++) nmax = ncell[i] > nmax ? ncell[i] : nmax;
  float range1[nmax + 2];
  for (int i = 0; i < nlay; i++) dz[i] = (zl[i + 1] - zl[i]) / ncell[i];
  for (int i = 0; i < nlay; i++) {
    z[sum] = zl[i];
    for (int ii = sum; ii < sum + ncell[i] - 1; ii++) z[ii + 1] = zl[i] + (ii - sum + 1) * dz[i];
    if (equal[i] < 1 && i > 0 && i < nlay - 1) {
      range1[0] = zl[i];
      if ((dz[i] - dz[i - 1]) * (dz[i + 1] - dz[i]) > 0) {
        int N1 = ((dz[i] - dz[i - 1]) / (dz[i + 1] - dz[i - 1])) * (ncell[i] + 2) + 1;
        if (N1 > ncell[i] - 1) N1 = ncell[i] - 1;
        int N2 = ncell[i] + 2 - N1;
        float dist = zl[i + 1] - zl[i] + dz[i - 1] + dz[i + 1];
        float mid = (2 * dist - dz[i - 1] * N1 - dz[i + 1] * (N2 + 1)) / (N1 + N2 - 1);
        float s1 = (mid - dz[i - 1]) / (N1 - 1);
        float s2 = (dz[i + 1] - mid) / (N2);
        /* the reference's range buffer holds ncell_max values; the last two
         * of this recursion are never read */
        for (int j = 1; j < N1 && j < nmax + 2; j++) range1[j] = range1[j - 1] + dz[i - 1] + j * s1;
        for (int j = N1; j < N1 + N2 && j < nmax + 2; j++)
          range1[j] = range1[j - 1] + dz[i + 1] - (N2 + N1 - j - 1) * s2;
        for (int ii = sum; ii < sum + ncell[i]; ii++) z[ii] = range1[ii - sum];
      }
    }
    sum += ncell[i];
  }
  z[sum] = zl[nlay];
}

#define IDX(i, j, k, nx, ny) ((size_t)(i) + (size_t)(j) * (nx) + (size_t)(k) * (nx) * (ny))

/* gd_curv_gen_layer after the file is read: n_if interfaces (ncell_in and
 * equal per layer), layer_xyz = x, y, z of the nx_l x ny_l samples of every
 * interface ([3][n_if][ny_l][nx_l]); resample and start as the par file's
 * refine factors and start indices; the block's x3d/y3d/z3d ([nz][ny][nx],
 * ghosts included).  0, or a negative code where the reference stops. */
int orc_gen_layer(int n_if, const int *ncell_in, const int *equal, int nx_l, int ny_l,
                  const float *layer_xyz, const int *resample, const int *start, int ntx,
                  int nty, int ntz, int ni, int nj, int nk, int gni1, int gnj1, int gnk1,
                  float *x3d, float *y3d, float *z3d)
{
  const int ng = 3, nlay = n_if - 1;
  const int nx = ni + 2 * ng, ny = nj + 2 * ng;
  if (nlay < 1) return -1;
  int ncell[nlay];
  const int fx = abs(resample[0]), fy = abs(resample[1]), fz = abs(resample[2]);
  int nxi, nyi, nzi = 1, nx_first, ny_first;
  if (resample[0] < 0) {
    nxi = (nx_l + 1) / fx;
    nx_first = (start[0] + 1) / fx;
  } else {
    nxi = (nx_l - 1) * fx + 1;
    nx_first = (start[0] - 1) * fx + 1;
  }
  if (resample[1] < 0) {
    nyi = (ny_l + 1) / fy;
    ny_first = (start[1] + 1) / fy;
  } else {
    nyi = (ny_l - 1) * fy + 1;
    ny_first = (start[1] - 1) * fy + 1;
  }
  for (int k = 0; k < nlay; k++) {
    ncell[k] = resample[2] < 0 ? ncell_in[k] / fz : ncell_in[k] * fz;
    nzi += ncell[k];
    /* a layer needs a cell, and an unequal middle layer two: with one the
     * reference's spacing recursion reads before its buffer */
    if (ncell[k] < 1 || (equal[k] < 1 && k > 0 && k < nlay - 1 && ncell[k] < 2)) return -6;
  }
  const int nz_first = nzi - start[2] - ntz;
  const int xg = nx_first + gni1 - ng, yg = ny_first + gnj1 - ng, zg = nz_first + gnk1 - ng;
  if (nx_first - ng < 0 || ny_first - ng < 0) return -2;
  if (nx_first + ntx + ng > nxi || ny_first + nty + ng > nyi) return -3;
  if (zg < 0) return -4;

  /* the interfaces resampled to nxi x nyi: every fx-th / fy-th sample, or a
   * spline along x then along y through the input samples */
  const size_t vin = (size_t)nx_l * ny_l * n_if, vit = (size_t)nxi * nyi * n_if;
  float *L = malloc(3 * vit * sizeof(float));
  int maxn = nx_l > ny_l ? nx_l : ny_l, maxi = nxi > nyi ? nxi : nyi;
  float *xn = malloc(maxn * sizeof(float)), *yn = malloc(maxn * sizeof(float));
  float *xi = malloc(maxi * sizeof(float)), *yi = malloc(maxi * sizeof(float));
  int rc = 0;
  if (resample[0] < 2 && resample[1] < 2) {
    for (int kk = 0; kk < n_if; kk++)
      for (int jj = 0; jj < nyi; jj++)
        for (int ii = 0; ii < nxi; ii++)
          for (int c = 0; c < 3; c++)
            L[c * vit + IDX(ii, jj, kk, nxi, nyi)] =
                layer_xyz[c * vin + IDX(ii * fx, jj * fy, kk, nx_l, ny_l)];
  } else if (resample[0] < 2) {          /* every fx-th in x, spline in y */
    for (int j = 0; j < nyi; j++) xi[j] = j;
    for (int j = 0; j < ny_l; j++) xn[j] = j * fy;
    for (int kk = 0; kk < n_if; kk++)
      for (int ii = 0; ii < nxi; ii++)
        for (int c = 0; c < 3; c++) {
          for (int j = 0; j < ny_l; j++) yn[j] = layer_xyz[c * vin + IDX(ii * fx, j, kk, nx_l, ny_l)];
          rc |= orc_spline_at(ny_l, xn, yn, nyi, xi, yi);
          for (int jj = 0; jj < nyi; jj++) L[c * vit + IDX(ii, jj, kk, nxi, nyi)] = yi[jj];
        }
  } else if (resample[1] < 2) {          /* spline in x, every fy-th in y */
    for (int i = 0; i < nxi; i++) xi[i] = i;
    for (int i = 0; i < nx_l; i++) xn[i] = i * fx;
    for (int kk = 0; kk < n_if; kk++)
      for (int jj = 0; jj < nyi; jj++)
        for (int c = 0; c < 3; c++) {
          for (int i = 0; i < nx_l; i++) yn[i] = layer_xyz[c * vin + IDX(i, jj * fy, kk, nx_l, ny_l)];
          rc |= orc_spline_at(nx_l, xn, yn, nxi, xi, yi);
          for (int ii = 0; ii < nxi; ii++) L[c * vit + IDX(ii, jj, kk, nxi, nyi)] = yi[ii];
        }
  } else {                               /* spline in x on the input rows, then in y */
    for (int i = 0; i < nxi; i++) xi[i] = i;
    for (int i = 0; i < nx_l; i++) xn[i] = i * fx;
    for (int kk = 0; kk < n_if; kk++)
      for (int jj = 0; jj < ny_l; jj++)
        for (int c = 0; c < 3; c++) {
          for (int i = 0; i < nx_l; i++) yn[i] = layer_xyz[c * vin + IDX(i, jj, kk, nx_l, ny_l)];
          rc |= orc_spline_at(nx_l, xn, yn, nxi, xi, yi);
          for (int ii = 0; ii < nxi; ii++) L[c * vit + IDX(ii, jj * fy, kk, nxi, nyi)] = yi[ii];
        }
    float *xi2 = malloc(nyi * sizeof(float)), *xn2 = malloc(ny_l * sizeof(float));
    float *yn2 = malloc(ny_l * sizeof(float)), *yi2 = malloc(nyi * sizeof(float));
    for (int j = 0; j < nyi; j++) xi2[j] = j;
    for (int j = 0; j < ny_l; j++) xn2[j] = j * fy;
    for (int kk = 0; kk < n_if; kk++)
      for (int ii = 0; ii < nxi; ii++)
        for (int c = 0; c < 3; c++) {
          for (int j = 0; j < ny_l; j++) yn2[j] = L[c * vit + IDX(ii, j * fy, kk, nxi, nyi)];
          rc |= orc_spline_at(ny_l, xn2, yn2, nyi, xi2, yi2);
          for (int jj = 0; jj < nyi; jj++) L[c * vit + IDX(ii, jj, kk, nxi, nyi)] = yi2[jj];
        }
    free(xi2); free(xn2); free(yn2); free(yi2);
  }
  free(xn); free(yn); free(xi); free(yi);
  if (rc) { free(L); return -5; }

  /* every column: z through the layers, x and y as splines of z, three ghost
   * points mirrored above the top */
  float *zp = malloc((nzi + ng) * sizeof(float)), *xp = malloc((nzi + ng) * sizeof(float));
  float *yp = malloc((nzi + ng) * sizeof(float));
  float xl[n_if], yl[n_if], zl[n_if];
  for (int i = 0; i < nx && !rc; i++)
    for (int j = 0; j < ny && !rc; j++) {
      for (int k = 0; k < n_if; k++) {
        size_t q = IDX(i + xg, j + yg, k, nxi, nyi);
        xl[k] = L[q]; yl[k] = L[vit + q]; zl[k] = L[2 * vit + q];
      }
      orc_grid_z(zp, zl, ncell, equal, nlay);
      rc |= orc_spline_at(n_if, zl, xl, nzi, zp, xp);
      rc |= orc_spline_at(n_if, zl, yl, nzi, zp, yp);
      for (int k = 1; k < ng + 1; k++) {
        int t = nzi - 1;
        xp[t + k] = xp[t] * 2 - xp[t - k];
        yp[t + k] = yp[t] * 2 - yp[t - k];
        zp[t + k] = zp[t] * 2 - zp[t - k];
      }
      for (int k = 0; k < nk + 2 * ng; k++) {
        size_t q = IDX(i, j, k, nx, ny);
        x3d[q] = xp[zg + k]; y3d[q] = yp[zg + k]; z3d[q] = zp[zg + k];
      }
    }
  free(zp); free(xp); free(yp); free(L);
  return rc ? -5 : 0;
}
