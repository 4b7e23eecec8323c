/*
 * host_setup.c -- host-side construction of the engine's small static tables:
 * CFS-PML profiles, source time-function tables, moment tensors and the
 * Gaussian source weights.  Each follows the reference function it replaces
 * with the same float/double mix, so the tables are bit-identical to the
 * reference's (built with -ffp-contract=off like the reference's x86-64 code).
 */
#include <math.h>
#include <stddef.h>
#include <string.h>

#include "cgfdh.h"

#ifndef PI
#define PI 3.14159265358979323846264338327950288419716939937510
#endif

/* region of one PML face with its own layer count, bdry_t.c:160-186 */
static void pml_region(int ni, int nj, int nk, int face, int L, int r[6])
{
  int ni1 = 3, ni2 = 3 + ni - 1, nj1 = 3, nj2 = 3 + nj - 1, nk1 = 3, nk2 = 3 + nk - 1;
  r[0] = ni1; r[1] = ni2; r[2] = nj1; r[3] = nj2; r[4] = nk1; r[5] = nk2;
  if (face == 0) r[1] = ni1 + L;
  if (face == 1) r[0] = ni2 - L;
  if (face == 2) r[3] = nj1 + L;
  if (face == 3) r[2] = nj2 - L;
  if (face == 4) r[5] = nk1 + L;
  if (face == 5) r[4] = nk2 - L;
}

/* average grid spacing along the face normal over the PML region,
 * bdry_cal_abl_len_dh for curvilinear grids (bdry_t.c:399-503) */
static float avg_dh(int ni, int nj, const float *xyz, size_t vol, const int r[6], int idim,
                    float *L0)
{
  size_t nx = (size_t)ni + 6, ny = (size_t)nj + 6, sl = nx, ss = nx * ny;
  const float *x3d = xyz, *y3d = xyz + vol, *z3d = xyz + 2 * vol;
  int o1, o2, m1, m2, c1, c2;
  size_t so, sm, sa;
  if (idim == 0) { o1 = r[4]; o2 = r[5]; so = ss; m1 = r[2]; m2 = r[3]; sm = sl; c1 = r[0]; c2 = r[1]; sa = 1; }
  else if (idim == 1) { o1 = r[4]; o2 = r[5]; so = ss; m1 = r[0]; m2 = r[1]; sm = 1; c1 = r[2]; c2 = r[3]; sa = sl; }
  else { o1 = r[2]; o2 = r[3]; so = sl; m1 = r[0]; m2 = r[1]; sm = 1; c1 = r[4]; c2 = r[5]; sa = ss; }
  double L = 0.0;
  int num = 0;
  for (int a = o1; a <= o2; a++)
    for (int b = m1; b <= m2; b++) {
      size_t p = (size_t)a * so + (size_t)b * sm + (size_t)c1 * sa;
      double x0 = x3d[p], y0 = y3d[p], z0 = z3d[p];
      for (int c = c1 + 1; c <= c2; c++) {
        size_t q = (size_t)a * so + (size_t)b * sm + (size_t)c * sa;
        double x1 = x3d[q], y1 = y3d[q], z1 = z3d[q];
        L += sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0) + (z1 - z0) * (z1 - z0));
        x0 = x1; y0 = y1; z0 = z1;
        num += 1;
      }
    }
  float dh = (float)(L / num);
  *L0 = dh * (c2 - c1);
  return dh;
}

int cgfdh_pml_face_coefs(int ni, int nj, int nk, const float *xyz, int face, int num_layers,
                         float velocity, float alpha_max, float beta_max,
                         float *A, float *B, float *D)
{
  if (!xyz || !A || !B || !D || face < 0 || face > 5 || num_layers < 1) return CGFDH_EINVAL;
  int r[6];
  pml_region(ni, nj, nk, face, num_layers, r);
  size_t vol = ((size_t)ni + 6) * ((size_t)nj + 6) * ((size_t)nk + 6);
  float L0;
  float dh = avg_dh(ni, nj, xyz, vol, r, face / 2, &L0);
  /* bdry_t.c:237-269 with bdry_pml_cal_R/dmax/d/a/b (bdry_t.c:333-366):
   * powers p_d = 2, p_a = 1, p_b = 2 */
  int npoints = num_layers + 1;
  float num_lay = npoints - 1;
  float Rpp = (float)(pow(10, -((log10((double)num_lay) - 1.0) / log10(2.0) + 4.0)));
  float dmax = (float)(-velocity / (2.0 * L0) * log(Rpp) * (2.0f + 1.0));
  for (int n = 0; n < npoints; n++) {
    float L = n * dh;
    int i = (face % 2 == 0) ? npoints - 1 - n : n;
    D[i] = (L < 0) ? 0.0f : (float)(dmax * pow(L / L0, 2.0f));
    A[i] = (L < 0) ? 0.0f : (float)(alpha_max * (1.0 - pow(L / L0, 1.0f)));
    B[i] = (L < 0) ? 1.0f : (float)(1.0 + (beta_max - 1.0) * pow(L / L0, 2.0f));
    D[i] /= B[i];
    A[i] += D[i];
    B[i] = 1.0 / B[i];
  }
  return CGFDH_OK;
}

/* fun_ricker, src_t.c:2159-2166 */
float cgfdh_ricker(float t, float fc, float t0)
{
  float u = (t - t0) * 2.0 * PI * fc;
  float v = (1 - u * u / 2.0) * exp(-u * u / 4.0);
  return v;
}

/* angle2moment and src_muDA_to_moment, src_t.c:2344-2409.
 * m: Mxx, Myy, Mzz, Myz, Mxz, Mxy (the .src file order) */
void cgfdh_moment_from_angles(float strike, float dip, float rake, float mu, float D,
                              float A, float *m)
{
  float dp = dip / 180.0 * PI, st = strike / 180.0 * PI, rk = rake / 180.0 * PI;
  float M11 = -(sin(dp) * cos(rk) * sin(2.0 * st) + sin(2.0 * dp) * sin(rk) * sin(st) * sin(st));
  float M22 = sin(dp) * cos(rk) * sin(2.0 * st) - sin(2.0 * dp) * sin(rk) * cos(st) * cos(st);
  float M33 = -(M11 + M22);
  float M12 = sin(dp) * cos(rk) * cos(2.0 * st) + 0.5 * sin(2.0 * dp) * sin(rk) * sin(2.0 * st);
  float M13 = -(cos(dp) * cos(rk) * cos(st) + cos(2.0 * dp) * sin(rk) * sin(st));
  float M23 = -(cos(dp) * cos(rk) * sin(st) - cos(2.0 * dp) * sin(rk) * cos(st));
  float t[6] = {M22, M11, M33, -M13, -M23, M12};
  float M0 = mu * D * A;
  for (int n = 0; n < 6; n++) m[n] = M0 * t[n];
}

/* source time-function table of one moment source given by wavelet name
 * "ricker" (src_read_locate_file src_t.c:550-555 for max_nt, and
 * src_put_into_struct src_t.c:1029-1091).  m6 in the .src order
 * Mxx,Myy,Mzz,Myz,Mxz,Mxy; out: 6 arrays [max_nt][4] in the engine order
 * Mxx,Myy,Mzz,Mxz,Myz,Mxy, each with stride out_stride floats. */
int cgfdh_ricker_moment_table(float t0, float dt, float stf_length, float wavelet_tstart,
                              float fc, float tpeak, const float *m6, int *it_begin,
                              int *it_end, int *max_nt, float *out, long out_stride)
{
  static const float rk_t[4] = {0.0f, 0.5f, 0.5f, 1.0f};
  int nt = (int)(stf_length / dt + 0.5);
  if (max_nt) *max_nt = nt;
  int ib = (int)((wavelet_tstart - t0) / dt);
  if (it_begin) *it_begin = ib;
  if (it_end) *it_end = ib + nt - 1;
  if (!out) return CGFDH_OK;
  float t_shift = wavelet_tstart - (ib * dt + t0);
  for (int it = 0; it < nt; it++)
    for (int s = 0; s < 4; s++) {
      float t = it * dt + rk_t[s] * dt - t_shift;
      float v = cgfdh_ricker(t, fc, tpeak);
      long q = (long)it * 4 + s;
      out[0 * out_stride + q] = v * m6[0];
      out[1 * out_stride + q] = v * m6[1];
      out[2 * out_stride + q] = v * m6[2];
      out[3 * out_stride + q] = v * m6[4];   /* Mxz */
      out[4 * out_stride + q] = v * m6[3];   /* Myz */
      out[5 * out_stride + q] = v * m6[5];
    }
  return CGFDH_OK;
}

/* normalised Gaussian weights, src_cal_norm_delt3d_z2fre (src_t.c:2111-2151)
 * with fun_gauss (src_t.c:2179-2184) */
static float fun_gauss(float t, float a, float t0)
{
  float f = exp(-(t - t0) * (t - t0) / (a * a)) / (sqrtf(PI) * a);
  return f;
}

int cgfdh_gauss_weights(float *w, float x0, float y0, float z0, float a, int half,
                        int half_below)
{
  if (!w || half < 0) return CGFDH_EINVAL;
  int n = 0;
  for (int k = -half; k <= half; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) {
        float D1 = fun_gauss(i - x0, a, 0.0);
        float D2 = fun_gauss(j - y0, a, 0.0);
        float D3 = fun_gauss(k - z0, a, 0.0);
        w[n++] = D1 * D2 * D3;
      }
  float sum = 0.0;
  n = 0;
  for (int k = -half; k <= half_below; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) sum += w[n++];
  if (sum < 1e-20) return CGFDH_EINVAL;
  int len = 2 * half + 1;
  for (int m = 0; m < len * len * len; m++) w[m] /= sum;
  return CGFDH_OK;
}

/* ------------------------------------------------------------------------ */
/* GMB visco-elastic coefficients: md_vis_GMB_cal_Y (md_t.c:643-792) with its */
/* least-squares fit md_visco_LS / md_visco_LS_mat_inv (md_t.c:795-914)       */
/* ------------------------------------------------------------------------ */

#define GMB_N CGFDH_MAXWELL_MAX
#define GMB_M (2 * CGFDH_MAXWELL_MAX)

/* inverse of the n x n normal matrix a (destroyed) by Gauss-Jordan, swapping
 * in a later row when a pivot is exactly zero */
static int gmb_invert(float a[GMB_N][GMB_N], float b[GMB_N][GMB_N], int n)
{
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) b[r][c] = r == c ? 1.0f : 0.0f;
  for (int k = 0; k < n; k++) {
    if (a[k][k] == 0) {
      int sw = k + 1;
      while (sw < n && a[sw][k] == 0) sw++;
      if (sw >= n) return -1;
      for (int c = 0; c < n; c++) {
        float t = a[k][c]; a[k][c] = a[sw][c]; a[sw][c] = t;
        t = b[k][c]; b[k][c] = b[sw][c]; b[sw][c] = t;
      }
    }
    float piv = a[k][k];
    for (int c = 0; c < n; c++) {
      a[k][c] = a[k][c] / piv;
      b[k][c] = b[k][c] / piv;
    }
    for (int r = 0; r < n; r++) {
      if (r == k) continue;
      float f = a[r][k];
      for (int c = 0; c < n; c++) {
        a[r][c] = a[r][c] - a[k][c] * f;
        b[r][c] = b[r][c] - b[k][c] * f;
      }
    }
  }
  return 0;
}

/* least squares y = (G^T G)^-1 G^T d for a right-hand side constant d */
static int gmb_fit(float G[GMB_M][GMB_N], float d, int m, int n, float *y)
{
  float gtg[GMB_N][GMB_N], inv[GMB_N][GMB_N], x[GMB_N][GMB_M];
  for (int r = 0; r < n; r++)
    for (int c = 0; c < n; c++) {
      float acc = 0;
      for (int k = 0; k < m; k++) acc = acc + G[k][r] * G[k][c];
      gtg[r][c] = acc;
    }
  if (gmb_invert(gtg, inv, n)) return -1;
  for (int r = 0; r < n; r++)
    for (int c = 0; c < m; c++) {
      float acc = 0;
      for (int k = 0; k < n; k++) acc = acc + inv[r][k] * G[c][k];
      x[r][c] = acc;
    }
  for (int r = 0; r < n; r++) {
    float acc = 0.0;
    for (int c = 0; c < m; c++) acc = acc + x[r][c] * d;
    y[r] = acc;
  }
  return 0;
}

/* phase terms of a GMB modulus at the reference frequency: theta1 = 1 -
 * sum Y/(1+(wr/wl)^2), theta2 = sum Y (wr/wl)/(1+(wr/wl)^2), R = |theta| */
static void gmb_theta(const float *Y, const float *wl, int n, float wr, float *t1, float *R)
{
  float acc = 0.0;
  for (int q = 0; q < n; q++) acc = acc + Y[q] / (1 + pow(wr / wl[q], 2));
  *t1 = 1 - acc;
  acc = 0.0;
  for (int q = 0; q < n; q++) acc = acc + Y[q] * wr / wl[q] / (1 + pow(wr / wl[q], 2));
  const float t2 = acc;
  *R = sqrt(pow(*t1, 2) + pow(t2, 2));
}

int cgfdh_visco_gmb(long npts, int nmaxwell, float freq, float fmin, float fmax,
                    float *lambda, float *mu, const float *Qp, const float *Qs,
                    float *ylam_ymu, float *wl)
{
  if (npts < 0 || nmaxwell < 1 || nmaxwell > CGFDH_MAXWELL_MAX || !lambda || !mu || !Qp ||
      !Qs || !ylam_ymu || !wl || !(fmin > 0) || !(fmax > fmin))
    return CGFDH_EINVAL;
  const int m = 2 * nmaxwell - 1;
  const float wr = 2 * PI * freq, w0 = 2 * PI * fmin, w1 = 2 * PI * fmax;
  const float ratio = w1 / w0;
  float wk[GMB_M];
  for (int k = 0; k < m; k++) wk[k] = w0 * pow(ratio, k / (float)(m - 1));
  for (int q = 0; q < nmaxwell; q++) wl[q] = wk[2 * q];
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (long p = 0; p < npts; p++) {
    float GP[GMB_M][GMB_N], GS[GMB_M][GMB_N], YP[GMB_N], YS[GMB_N];
    const float qp1 = 1.0 / round(Qp[p]), qs1 = 1.0 / round(Qs[p]);
    for (int k = 0; k < m; k++)
      for (int q = 0; q < nmaxwell; q++) {
        GP[k][q] = (wl[q] * wk[k] + pow(wl[q], 2) * qp1) / (pow(wl[q], 2) + pow(wk[k], 2));
        GS[k][q] = (wl[q] * wk[k] + pow(wl[q], 2) * qs1) / (pow(wl[q], 2) + pow(wk[k], 2));
      }
    if (gmb_fit(GP, qp1, m, nmaxwell, YP) || gmb_fit(GS, qs1, m, nmaxwell, YS)) {
      bad = 1;
      continue;
    }
    const float lam = lambda[p], muu = mu[p];
    /* P modulus lambda + 2 mu and S modulus mu scaled to unrelaxed values:
     * M_u = M (R + theta1) / (2 R^2) */
    float t1, R;
    gmb_theta(YP, wl, nmaxwell, wr, &t1, &R);
    const float kappa = (lam + 2 * muu) * (R + t1) / (2 * pow(R, 2));
    gmb_theta(YS, wl, nmaxwell, wr, &t1, &R);
    const float mu_u = muu * (R + t1) / (2 * pow(R, 2));
    mu[p] = mu_u;
    lambda[p] = kappa - 2 * mu_u;
    for (int q = 0; q < nmaxwell; q++) {
      ylam_ymu[(2 * q) * npts + p] = (1 + 2 * mu[p] / lambda[p]) * YP[q] - 2 * mu[p] / lambda[p] * YS[q];
      ylam_ymu[(2 * q + 1) * npts + p] = YS[q];
    }
  }
  return bad ? CGFDH_EINVAL : CGFDH_OK;
}
