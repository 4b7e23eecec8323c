/*
 * cgfd_oracle.c -- CPU restatement of the reference's elastic hot path
 * (isotropic, VTI and general anisotropic media).  TEST INFRASTRUCTURE ONLY
 * (see cgfd_oracle.h).
 *
 * Every function follows the cited reference loop expression by expression,
 * including where the reference evaluates in double and rounds to float, so
 * that the FP32 build reproduces the reference bit for bit (it is compiled
 * with -ffp-contract=off, as x86-64 -O3 code has no FMA).  The one deliberate
 * difference: the traction-image mirror on the surface row with the forward
 * z operator reads the real grid points below the window (the reference's
 * own alternative formula, sv_curv_col_el.c:151-157) where the reference
 * indexes veczt[-1] and veczt[-2] out of bounds (sv_curv_col_el.c:160-163).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "cgfd_oracle.h"

#ifndef PI
#define PI 3.14159265358979323846264338327950288419716939937510
#endif

/* ------------------------------------------------------------------------ */
/* scheme tables: fd_t.c:36-49 (RK4), :75-131 (one-sided ops), :185-236     */
/* (pair schedule)                                                          */
/* ------------------------------------------------------------------------ */

/* [layer][dir][n]; layer 3 is the interior DRP op, 0..2 the near-surface ones */
static const int op_indx[4][2][5] = {
  {{0, 1, 0, 0, 0}, {-1, 0, 0, 0, 0}},
  {{0, 1, 0, 0, 0}, {-1, 0, 0, 0, 0}},
  {{0, 1, 2, 0, 0}, {-2, -1, 0, 0, 0}},
  {{-1, 0, 1, 2, 3}, {-3, -2, -1, 0, 1}}};
static const double op_coef[4][2][5] = {
  {{-1.0, 1.0, 0, 0, 0}, {-1.0, 1.0, 0, 0, 0}},
  {{-1.0, 1.0, 0, 0, 0}, {-1.0, 1.0, 0, 0, 0}},
  {{-7.0 / 6.0, 8.0 / 6.0, -1.0 / 6.0, 0, 0}, {1.0 / 6.0, -8.0 / 6.0, 7.0 / 6.0, 0, 0}},
  {{-0.30874, -0.6326, 1.233, -0.3334, 0.04168},
   {-0.04168, 0.3334, -1.233, 0.6326, 0.30874}}};
static const int op_len[4] = {2, 2, 3, 5};
/* forward(0)/backward(1) flag per pair for x,y,z; per stage dir=(flag+s)%2 */
static const int pair_flags[8][3] = {
  {0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
  {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
static const double rk_a_tab[3] = {0.5, 0.5, 1.0};
static const double rk_b_tab[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
static const double rk_t_tab[4] = {0.0, 0.5, 0.5, 1.0};
/* centred 7-point op for the metric, fd_t.c:161-171 (row 3) */
static const double fdc_coef[7] = {-0.02084, 0.1667, -0.7709, 0.0, 0.7709, -0.1667, 0.02084};

int orc_fd_ops(int *indx, float *coef, int *len, float *rk_a, float *rk_b, float *rk_t)
{
  for (int p = 0; p < 8; p++)
    for (int s = 0; s < 4; s++)
      for (int slot = 0; slot < 6; slot++) {
        int d = slot < 2 ? slot : 2;
        int dir = (pair_flags[p][d] + s) % 2;
        int lay = slot < 2 ? 3 : slot - 2;
        int base = (p * 4 + s) * 6 + slot;
        len[base] = op_len[lay];
        for (int n = 0; n < 5; n++) {
          indx[base * 5 + n] = n < op_len[lay] ? op_indx[lay][dir][n] : 0;
          coef[base * 5 + n] = n < op_len[lay] ? (float)op_coef[lay][dir][n] : 0.0f;
        }
      }
  for (int s = 0; s < 4; s++) {
    rk_b[s] = (float)rk_b_tab[s];
    rk_t[s] = (float)rk_t_tab[s];
    if (s < 3) rk_a[s] = (float)rk_a_tab[s];
  }
  return 8;
}

/* ------------------------------------------------------------------------ */
/* setup: PML profile (bdry_t.c:237-269, 333-366), average spacing          */
/* (bdry_t.c:399-503), regions (bdry_t.c:160-186)                           */
/* ------------------------------------------------------------------------ */

void orc_pml_profile(int num_layers, float L0, float dh, float vel,
                     float amax, float bmax, int iside,
                     float *A, float *B, float *D)
{
  int npoints = num_layers + 1;
  float num_lay = npoints - 1;
  float Rpp = (float)(pow(10, -((log10((double)num_lay) - 1.0) / log10(2.0) + 4.0)));
  float dmax = (float)(-vel / (2.0 * L0) * log(Rpp) * (2.0f + 1.0));
  for (int n = 0; n < npoints; n++) {
    float L = n * dh;
    int i = iside == 0 ? npoints - 1 - n : n;
    D[i] = (L < 0) ? 0.0f : (float)(dmax * pow(L / L0, 2.0f));
    A[i] = (L < 0) ? 0.0f : (float)(amax * (1.0 - pow(L / L0, 1.0f)));
    B[i] = (L < 0) ? 1.0f : (float)(1.0 + (bmax - 1.0) * pow(L / L0, 2.0f));
    D[i] /= B[i];
    A[i] += D[i];
    B[i] = 1.0 / B[i];
  }
}

void orc_pml_region(int ni, int nj, int nk, const int pml_layers[6],
                    const int pml_on[6], int region[6][6])
{
  int ni1 = ORC_NGHOST, ni2 = ni1 + ni - 1;
  int nj1 = ORC_NGHOST, nj2 = nj1 + nj - 1;
  int nk1 = ORC_NGHOST, nk2 = nk1 + nk - 1;
  for (int f = 0; f < 6; f++) {
    int L = pml_on[f] ? pml_layers[f] : 0;
    int *r = region[f];
    r[0] = ni1; r[1] = ni2; r[2] = nj1; r[3] = nj2; r[4] = nk1; r[5] = nk2;
    if (f == 0) r[1] = ni1 + L;
    if (f == 1) r[0] = ni2 - L;
    if (f == 2) r[3] = nj1 + L;
    if (f == 3) r[2] = nj2 - L;
    if (f == 4) r[5] = nk1 + L;
    if (f == 5) r[4] = nk2 - L;
  }
}

float orc_pml_avg_dh(int ni, int nj, int nk, const float *x3d, const float *y3d,
                     const float *z3d, const int r[6], int idim, float *L0)
{
  size_t nx = ni + 6, ny = nj + 6;
  size_t sl = nx, ss = nx * ny;
  double L = 0.0;
  int num = 0;
  /* walk each grid line of the region along idim, summing segment lengths */
  int a1, a2, b1, b2, c1, c2;      /* outer, middle, along-line ranges */
  size_t so, sm, sa;
  if (idim == 0) { a1 = r[4]; a2 = r[5]; so = ss; b1 = r[2]; b2 = r[3]; sm = sl; c1 = r[0]; c2 = r[1]; sa = 1; }
  else if (idim == 1) { a1 = r[4]; a2 = r[5]; so = ss; b1 = r[0]; b2 = r[1]; sm = 1; c1 = r[2]; c2 = r[3]; sa = sl; }
  else { a1 = r[2]; a2 = r[3]; so = sl; b1 = r[0]; b2 = r[1]; sm = 1; c1 = r[4]; c2 = r[5]; sa = ss; }
  for (int a = a1; a <= a2; a++)
    for (int b = b1; b <= b2; b++) {
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

/* ------------------------------------------------------------------------ */
/* setup: free-surface matrices, sv_curv_col_el_iso.c:1259-1375 with        */
/* fdlib_math_invert3x3 / matmul3x3 (fdlib_math.c:7-49)                     */
/* ------------------------------------------------------------------------ */

static void inv3(float m[3][3])
{
  float inv[3][3], det;
  inv[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  inv[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  inv[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  inv[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  inv[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  inv[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  inv[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  inv[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  inv[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0f / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = inv[i][j] * det;
}

static void mul3(float A[3][3], float B[3][3], float C[3][3])
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      C[i][j] = 0.0;
      for (int k = 0; k < 3; k++) C[i][j] += A[i][k] * B[k][j];
    }
}

/* ------------------------------------------------------------------------ */
/* VTI and general anisotropic stress-velocity coupling                     */
/* ------------------------------------------------------------------------ */

/* Stress rows in Voigt order Txx,Tyy,Tzz,Tyz,Txz,Txy (wavefield slots 3..8).
 * For one direction with metric (e1,e2,e3), the rate of stress row t holds
 * the derivative of V_b times a coefficient K[t][b]:
 *   anisotropic: K = c[a0]*e1 + c[a1]*e2 + c[a2]*e3 with a = aniso_ix[t][b]
 *     into c11,c12,c13,c14,c15,c16,c22,c23,c24,c25,c26,c33,c34,c35,c36,c44,
 *     c45,c46,c55,c56,c66 (sv_curv_col_el_aniso.c:468-495);
 *   VTI: K = c[vti_c]*e[vti_e] into {c11, c12 = c11 - 2 c66, c13, c33, c55,
 *     c66}, no term where vti_c < 0 (sv_curv_col_el_vti.c:453-475). */
static const int aniso_ix[6][3][3] = {
  {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}},
  {{1, 10, 9}, {10, 6, 8}, {9, 8, 7}},
  {{2, 14, 13}, {14, 7, 12}, {13, 12, 11}},
  {{3, 17, 16}, {17, 8, 15}, {16, 15, 12}},
  {{4, 19, 18}, {19, 9, 16}, {18, 16, 13}},
  {{5, 20, 19}, {20, 10, 17}, {19, 17, 14}}};
static const int vti_c[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 2, 3}, {-1, 4, 4}, {4, -1, 4}, {5, 5, -1}};
static const int vti_e[6][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2}, {0, 2, 1}, {2, 0, 0}, {1, 0, 0}};
/* rows of the traction on a zeta surface: (Txx,Txy,Txz), (Txy,Tyy,Tyz),
 * (Txz,Tyz,Tzz) against zeta_x, zeta_y, zeta_z */
static const int trac_row[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

static inline int has_term(int mt, int t, int b) { return mt != ORC_MEDIUM_VTI || vti_c[t][b] >= 0; }

int orc_medium_ncmp(int mt)
{
  return mt == ORC_MEDIUM_VTI ? 6 : (mt == ORC_MEDIUM_ANISO ? 22 : 3);
}

#define DEFINE_COUPLING(NAME, T)                                                  \
  static void NAME(int mt, const T *c, const T e[3], T K[6][3])                   \
  {                                                                               \
    for (int t = 0; t < 6; t++)                                                   \
      for (int b = 0; b < 3; b++) {                                               \
        if (mt == ORC_MEDIUM_VTI) {                                               \
          K[t][b] = vti_c[t][b] < 0 ? (T)0 : c[vti_c[t][b]] * e[vti_e[t][b]];     \
        } else {                                                                  \
          const int *a = aniso_ix[t][b];                                          \
          K[t][b] = c[a[0]] * e[0] + c[a[1]] * e[1] + c[a[2]] * e[2];             \
        }                                                                         \
      }                                                                           \
  }                                                                               \
  /* point medium: VTI {c11,c12,c13,c33,c55,c66}, anisotropic the 21 Cij */       \
  static void NAME##_medium(int mt, const T *med, size_t V, size_t p, T *c)       \
  {                                                                               \
    if (mt == ORC_MEDIUM_VTI) {                                                   \
      T c11 = med[p], c13 = med[V + p], c33 = med[2 * V + p];                     \
      T c55 = med[3 * V + p], c66 = med[4 * V + p];                               \
      c[0] = c11; c[1] = c11 - 2.0 * c66; c[2] = c13;                             \
      c[3] = c33; c[4] = c55; c[5] = c66;                                         \
    } else {                                                                      \
      for (int q = 0; q < 21; q++) c[q] = med[q * V + p];                         \
    }                                                                             \
  }
DEFINE_COUPLING(coupling, real_t)
DEFINE_COUPLING(coupling_f, float)

/* the six stress rates: one left-to-right sum over directions d and velocity
 * components b of K_d[t][b] * D[d][b], the order the reference writes them */
static void stress_rates(int mt, const real_t *c, int ndir, const real_t (*e)[3],
                         const real_t (*D)[3], real_t out[6])
{
  real_t K[3][6][3];
  for (int d = 0; d < ndir; d++) coupling(mt, c, e[d], K[d]);
  for (int t = 0; t < 6; t++) {
    real_t acc = 0;
    int first = 1;
    for (int d = 0; d < ndir; d++)
      for (int b = 0; b < 3; b++) {
        if (!has_term(mt, t, b)) continue;
        real_t term = K[d][t][b] * D[d][b];
        acc = first ? term : acc + term;
        first = 0;
      }
    out[t] = acc;
  }
}

/* traction-free matrices for VTI / anisotropic media: A[i][j] sums
 * K_zeta[row][j] * zeta_m, B and C the same with K_xi and K_eta; stored as
 * -A^-1 B and -A^-1 C (sv_curv_col_el_vti.c:1159-1199,
 * sv_curv_col_el_aniso.c:1377-1413) */
void orc_dvh2dvz(int mt, int ni, int nj, int nk, const float *metric10,
                 const float *med, float *mats)
{
  if (mt != ORC_MEDIUM_VTI && mt != ORC_MEDIUM_ANISO) {
    orc_dvh2dvz_iso(ni, nj, nk, metric10, med, mats);
    return;
  }
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = nx * ny * nz, ss = nx * ny;
  float *mVx = mats, *mVy = mats + ss * 9;
  size_t k = ORC_NGHOST + nk - 1;
  for (size_t j = ORC_NGHOST; j < ORC_NGHOST + (size_t)nj; j++)
    for (size_t i = ORC_NGHOST; i < ORC_NGHOST + (size_t)ni; i++) {
      size_t p = i + j * nx + k * ss;
      float c[21], xi[3], et[3], zt[3];
      coupling_f_medium(mt, med, vol, p, c);
      for (int m = 0; m < 3; m++) {
        xi[m] = metric10[(1 + m) * vol + p];
        et[m] = metric10[(4 + m) * vol + p];
        zt[m] = metric10[(7 + m) * vol + p];
      }
      float Kx[6][3], Ky[6][3], Kz[6][3];
      coupling_f(mt, c, xi, Kx);
      coupling_f(mt, c, et, Ky);
      coupling_f(mt, c, zt, Kz);
      float A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
      for (int r = 0; r < 3; r++)
        for (int b = 0; b < 3; b++) {
          int first = 1;
          for (int m = 0; m < 3; m++) {
            int t = trac_row[r][m];
            if (!has_term(mt, t, b)) continue;
            float a = Kz[t][b] * zt[m], bb = Kx[t][b] * zt[m], cc = Ky[t][b] * zt[m];
            if (first) { A[r][b] = a; B[r][b] = bb; C[r][b] = cc; }
            else { A[r][b] += a; B[r][b] += bb; C[r][b] += cc; }
            first = 0;
          }
        }
      inv3(A);
      mul3(A, B, AB);
      mul3(A, C, AC);
      size_t ij = (j * nx + i) * 9;
      for (int r = 0; r < 3; r++)
        for (int b = 0; b < 3; b++) {
          mVx[ij + r * 3 + b] = -1.0f * AB[r][b];
          mVy[ij + r * 3 + b] = -1.0f * AC[r][b];
        }
    }
}

/* visco-elastic surface matrices, sv_curv_col_vis_iso.c:354-506 */
void orc_dvh2dvz_vis(int ni, int nj, int nk, const float *metric10,
                     const float *med, const float *xyz, float *mats4)
{
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = nx * ny * nz, ss = nx * ny;
  const float *x3d = xyz, *y3d = xyz + vol, *z3d = xyz + 2 * vol;
  float *mVx = mats4, *mVy = mats4 + ss * 9, *mD = mats4 + 3 * ss * 9;
  /* the isotropic matrices: the same expressions (:435-468) */
  float *tmp = (float *)calloc(3 * ss * 9, sizeof(float));
  orc_dvh2dvz_iso(ni, nj, nk, metric10, med, tmp);
  size_t k = ORC_NGHOST + nk - 1;
  float fc[7];
  for (int n = 0; n < 7; n++) fc[n] = (float)fdc_coef[n];
  for (size_t j = ORC_NGHOST; j < ORC_NGHOST + (size_t)nj; j++)
    for (size_t i = ORC_NGHOST; i < ORC_NGHOST + (size_t)ni; i++) {
      size_t p = i + j * nx + k * ss, ij = (j * nx + i) * 9;
      float ztx = metric10[7 * vol + p], zty = metric10[8 * vol + p], ztz = metric10[9 * vol + p];
      /* M_FD_SHIFT along eta with the centred op (fd->fdc, fd_t.c:161-171) */
      float xet = fc[0] * x3d[p - 3 * nx], yet = fc[0] * y3d[p - 3 * nx], zet = fc[0] * z3d[p - 3 * nx];
      for (int n = 1; n < 7; n++) {
        long q = (long)p + (long)(n - 3) * (long)nx;
        xet += fc[n] * x3d[q];
        yet += fc[n] * y3d[q];
        zet += fc[n] * z3d[q];
      }
      float e_n = 1.0 / sqrt(xet * xet + yet * yet + zet * zet);
      float e_m = 1.0 / sqrt(ztx * ztx + zty * zty + ztz * ztz);
      float e_nm = e_n * e_m;
      float D[9];
      D[0] = (yet * ztz - zet * zty) * e_nm;
      D[1] = (zet * ztx - xet * ztz) * e_nm;
      D[2] = (xet * zty - yet * ztx) * e_nm;
      D[3] = xet * e_n;
      D[4] = yet * e_n;
      D[5] = zet * e_n;
      D[6] = ztx * e_m;
      D[7] = zty * e_m;
      D[8] = ztz * e_m;
      for (int q = 0; q < 9; q++) {
        mVx[ij + q] = tmp[ij + q];
        mVy[ij + q] = tmp[ss * 9 + ij + q];
        mD[ij + q] = D[q];
      }
    }
  free(tmp);
}

/* md_visco_LS_mat_inv (md_t.c:857-914): Gauss-Jordan with row swap on a zero
 * pivot */
static void ls_mat_inv(float m[][ORC_MAXWELL_MAX], float inv[][ORC_MAXWELL_MAX], int n)
{
  float tmp;
  int j = 0;
  for (int k = 0; k < n; k++) {
    if (m[k][k] == 0) {
      for (int jj = k + 1; jj < n; jj++) {
        j = jj;
        if (m[j][k] != 0) break;
      }
      for (int i = 0; i < n; i++) {
        tmp = m[k][i]; m[k][i] = m[j][i]; m[j][i] = tmp;
        tmp = inv[k][i]; inv[k][i] = inv[j][i]; inv[j][i] = tmp;
      }
    }
    tmp = m[k][k];
    for (int jj = 0; jj < n; jj++) {
      m[k][jj] = m[k][jj] / tmp;
      inv[k][jj] = inv[k][jj] / tmp;
    }
    for (int i = 0; i < n; i++) {
      if (i == k) continue;
      tmp = m[i][k];
      for (int jj = 0; jj < n; jj++) {
        m[i][jj] = m[i][jj] - m[k][jj] * tmp;
        inv[i][jj] = inv[i][jj] - inv[k][jj] * tmp;
      }
    }
  }
}

/* md_visco_LS (md_t.c:795-852): y = (G^T G)^-1 G^T d for a constant d */
static void visco_ls(float G[][ORC_MAXWELL_MAX], float *y, float d, int m, int n)
{
  float tr[ORC_MAXWELL_MAX][2 * ORC_MAXWELL_MAX], mu[ORC_MAXWELL_MAX][ORC_MAXWELL_MAX];
  float inv[ORC_MAXWELL_MAX][ORC_MAXWELL_MAX], pr[ORC_MAXWELL_MAX][2 * ORC_MAXWELL_MAX];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) inv[i][j] = i == j ? 1 : 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < m; j++) tr[i][j] = G[j][i];
  for (int i = 0; i < n; i++)
    for (int j = 0; j < n; j++) {
      mu[i][j] = 0;
      for (int k = 0; k < m; k++) mu[i][j] = mu[i][j] + tr[i][k] * G[k][j];
    }
  ls_mat_inv(mu, inv, n);
  for (int i = 0; i < n; i++)
    for (int j = 0; j < m; j++) {
      pr[i][j] = 0;
      for (int k = 0; k < n; k++) pr[i][j] = pr[i][j] + inv[i][k] * tr[k][j];
    }
  for (int i = 0; i < n; i++) {
    float sum = 0.0;
    for (int j = 0; j < m; j++) sum = sum + pr[i][j] * d;
    y[i] = sum;
  }
}

void orc_visco_gmb(long npts, int nmaxwell, float freq, float fmin, float fmax,
                   float *lambda, float *mu, const float *Qp, const float *Qs,
                   float *Ylam, float *Ymu, float *wl)
{
  int kmax = 2 * nmaxwell - 1;
  float wr = 2 * PI * freq, wmin = 2 * PI * fmin, wmax = 2 * PI * fmax;
  float wratio = wmax / wmin;
  float wk[2 * ORC_MAXWELL_MAX];
  for (int k = 0; k < kmax; k++) wk[k] = wmin * pow(wratio, k / (float)(kmax - 1));
  for (int n = 0; n < nmaxwell; n++) wl[n] = wk[2 * n];
#pragma omp parallel for schedule(static)
  for (long p = 0; p < npts; p++) {
    float GP[2 * ORC_MAXWELL_MAX][ORC_MAXWELL_MAX], GS[2 * ORC_MAXWELL_MAX][ORC_MAXWELL_MAX];
    float YP[ORC_MAXWELL_MAX], YS[ORC_MAXWELL_MAX];
    float QP1 = 1.0 / round(Qp[p]), QS1 = 1.0 / round(Qs[p]);
    float lam = lambda[p], muu = mu[p];
    for (int m = 0; m < kmax; m++)
      for (int n = 0; n < nmaxwell; n++) {
        GP[m][n] = (wl[n] * wk[m] + pow(wl[n], 2) * QP1) / (pow(wl[n], 2) + pow(wk[m], 2));
        GS[m][n] = (wl[n] * wk[m] + pow(wl[n], 2) * QS1) / (pow(wl[n], 2) + pow(wk[m], 2));
      }
    visco_ls(GP, YP, QP1, kmax, nmaxwell);
    visco_ls(GS, YS, QS1, kmax, nmaxwell);
    float theta1, theta2, thetatmp = 0.0, R;
    for (int n = 0; n < nmaxwell; n++) thetatmp = thetatmp + YP[n] / (1 + pow(wr / wl[n], 2));
    theta1 = 1 - thetatmp;
    thetatmp = 0.0;
    for (int n = 0; n < nmaxwell; n++)
      thetatmp = thetatmp + YP[n] * wr / wl[n] / (1 + pow(wr / wl[n], 2));
    theta2 = thetatmp;
    R = sqrt(pow(theta1, 2) + pow(theta2, 2));
    float kappa = (lam + 2 * muu) * (R + theta1) / (2 * pow(R, 2));
    thetatmp = 0.0;
    for (int n = 0; n < nmaxwell; n++) thetatmp = thetatmp + YS[n] / (1 + pow(wr / wl[n], 2));
    theta1 = 1 - thetatmp;
    thetatmp = 0.0;
    for (int n = 0; n < nmaxwell; n++)
      thetatmp = thetatmp + YS[n] * wr / wl[n] / (1 + pow(wr / wl[n], 2));
    theta2 = thetatmp;
    R = sqrt(pow(theta1, 2) + pow(theta2, 2));
    float muunrelax = muu * (R + theta1) / (2 * pow(R, 2));
    mu[p] = muunrelax;
    lambda[p] = kappa - 2 * muunrelax;
    for (int n = 0; n < nmaxwell; n++) {
      Ylam[p * nmaxwell + n] = (1 + 2 * mu[p] / lambda[p]) * YP[n] - 2 * mu[p] / lambda[p] * YS[n];
      Ymu[p * nmaxwell + n] = YS[n];
    }
  }
}

void orc_dvh2dvz_iso(int ni, int nj, int nk, const float *metric10,
                     const float *med3, float *mats)
{
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = nx * ny * nz, ss = nx * ny;
  const float *xi_x = metric10 + vol, *xi_y = metric10 + 2 * vol, *xi_z = metric10 + 3 * vol;
  const float *et_x = metric10 + 4 * vol, *et_y = metric10 + 5 * vol, *et_z = metric10 + 6 * vol;
  const float *zt_x = metric10 + 7 * vol, *zt_y = metric10 + 8 * vol, *zt_z = metric10 + 9 * vol;
  const float *lam3d = med3, *mu3d = med3 + vol;
  float *mVx = mats, *mVy = mats + ss * 9, *mF = mats + 2 * ss * 9;
  size_t k = ORC_NGHOST + nk - 1;
  for (size_t j = ORC_NGHOST; j < ORC_NGHOST + (size_t)nj; j++)
    for (size_t i = ORC_NGHOST; i < ORC_NGHOST + (size_t)ni; i++) {
      size_t p = i + j * nx + k * ss;
      float e11 = xi_x[p], e12 = xi_y[p], e13 = xi_z[p];
      float e21 = et_x[p], e22 = et_y[p], e23 = et_z[p];
      float e31 = zt_x[p], e32 = zt_y[p], e33 = zt_z[p];
      float lam = lam3d[p], mu = mu3d[p];
      float lam2mu = lam + 2.0f * mu;
      float A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
      A[0][0] = lam2mu * e31 * e31 + mu * (e32 * e32 + e33 * e33);
      A[0][1] = lam * e31 * e32 + mu * e32 * e31;
      A[0][2] = lam * e31 * e33 + mu * e33 * e31;
      A[1][0] = lam * e32 * e31 + mu * e31 * e32;
      A[1][1] = lam2mu * e32 * e32 + mu * (e31 * e31 + e33 * e33);
      A[1][2] = lam * e32 * e33 + mu * e33 * e32;
      A[2][0] = lam * e33 * e31 + mu * e31 * e33;
      A[2][1] = lam * e33 * e32 + mu * e32 * e33;
      A[2][2] = lam2mu * e33 * e33 + mu * (e31 * e31 + e32 * e32);
      inv3(A);
      B[0][0] = -lam2mu * e31 * e11 - mu * (e32 * e12 + e33 * e13);
      B[0][1] = -lam * e31 * e12 - mu * e32 * e11;
      B[0][2] = -lam * e31 * e13 - mu * e33 * e11;
      B[1][0] = -lam * e32 * e11 - mu * e31 * e12;
      B[1][1] = -lam2mu * e32 * e12 - mu * (e31 * e11 + e33 * e13);
      B[1][2] = -lam * e32 * e13 - mu * e33 * e12;
      B[2][0] = -lam * e33 * e11 - mu * e31 * e13;
      B[2][1] = -lam * e33 * e12 - mu * e32 * e13;
      B[2][2] = -lam2mu * e33 * e13 - mu * (e31 * e11 + e32 * e12);
      C[0][0] = -lam2mu * e31 * e21 - mu * (e32 * e22 + e33 * e23);
      C[0][1] = -lam * e31 * e22 - mu * e32 * e21;
      C[0][2] = -lam * e31 * e23 - mu * e33 * e21;
      C[1][0] = -lam * e32 * e21 - mu * e31 * e22;
      C[1][1] = -lam2mu * e32 * e22 - mu * (e31 * e21 + e33 * e23);
      C[1][2] = -lam * e32 * e23 - mu * e33 * e22;
      C[2][0] = -lam * e33 * e21 - mu * e31 * e23;
      C[2][1] = -lam * e33 * e22 - mu * e32 * e23;
      C[2][2] = -lam2mu * e33 * e23 - mu * (e31 * e21 + e32 * e22);
      mul3(A, B, AB);
      mul3(A, C, AC);
      size_t ij = (j * nx + i) * 9;
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
          mVx[ij + r * 3 + c] = AB[r][c];
          mVy[ij + r * 3 + c] = AC[r][c];
          mF[ij + r * 3 + c] = A[r][c];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* Graves Qs attenuation, sv_curv_col_el_graves_Qs (sv_curv_col_el.c:638-667),*/
/* applied to w_end after the last stage's combine (drv_rk_curv_col.c:414-416):*/
/* every component at physical points times expf(coef / Qs), coef =          */
/* -PI * freq * dt evaluated in double (PI is a double literal) and stored   */
/* as float                                                                  */
/* ------------------------------------------------------------------------ */

void orc_graves_qs(int ni, int nj, int nk, int ncmp, float *w, const float *Qs, float freq,
                   float dt)
{
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6, vol = nx * ny * nz;
  float coef = (float)(-3.14159265358979323846 * (double)freq * (double)dt);
  for (int c = 0; c < ncmp; c++)
    for (int k = 3; k < 3 + nk; k++)
      for (int j = 3; j < 3 + nj; j++)
        for (int i = 3; i < 3 + ni; i++) {
          size_t p = i + j * nx + k * nx * ny;
          w[c * vol + p] *= expf(coef / Qs[p]);
        }
}

/* ------------------------------------------------------------------------ */
/* ablexp sponge, bdry_ablexp_apply (bdry_t.c:840-890): blocks in order,    */
/* every component times min(Ex[i], Ey[j], Ez[k]); blocks as {enable, ni1, */
/* ni2, nj1, nj2, nk1, nk2}                                                 */
/* ------------------------------------------------------------------------ */

void orc_ablexp_apply(int nx, int ny, int nz, int ncmp, float *w, const float *Ex,
                      const float *Ey, const float *Ez, const int *blocks)
{
  size_t vol = (size_t)nx * ny * nz;
  for (int c = 0; c < ncmp; c++)
    for (int n = 0; n < 6; n++) {
      const int *b = blocks + 7 * n;
      if (!b[0]) continue;
      for (int k = b[5]; k <= b[6]; k++)
        for (int j = b[3]; j <= b[4]; j++)
          for (int i = b[1]; i <= b[2]; i++) {
            float m = Ex[i] < Ey[j] ? Ex[i] : Ey[j];
            if (m > Ez[k]) m = Ez[k];
            w[c * vol + i + (size_t)j * nx + (size_t)k * nx * ny] *= m;
          }
    }
}

/* ------------------------------------------------------------------------ */
/* setup: stable time step, blk_dt_esti_curv (blk_t.c:2840-2930) with the   */
/* point-to-plane distance of fdlib_math.c:52-89; float throughout except   */
/* where the reference calls the double sqrt or writes double literals      */
/* ------------------------------------------------------------------------ */

float orc_dist_point2plane(const float x0[3], const float x1[3], const float x2[3],
                           const float x3[3])
{
  float a[3], b[3], p[3];
  for (int c = 0; c < 3; c++) {
    a[c] = x2[c] - x1[c];
    b[c] = x3[c] - x1[c];
  }
  p[0] = a[1] * b[2] - a[2] * b[1];
  p[1] = a[2] * b[0] - a[0] * b[2];
  p[2] = a[0] * b[1] - a[1] * b[0];
  float d = 0.0f, e = 0.0f, n = 0.0f;
  for (int c = 0; c < 3; c++) d += p[c] * x1[c];
  for (int c = 0; c < 3; c++) e += p[c] * x0[c];
  for (int c = 0; c < 3; c++) n += p[c] * p[c];
  float L = (float)fabs(e - d);
  return L / sqrtf(n);
}

/* vols by medium: iso and visco lambda, mu, rho (visco then Ylam_n, Ymu_n
 * interleaved per Maxwell body, the md_t order);
 * VTI c11, c33, rho; aniso c11, rho.  ijk = -1 and vp = L = 0 if no point
 * beats the reference's starting value 1e10. */
void orc_dt_estimate(int ni, int nj, int nk, const float *xyz, int medium_type,
                     const float *vols, int nmaxwell, float cfl, float *dtmax,
                     float *vpmax, float *lmin, int *ijk)
{
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = nx * ny * nz, sy = nx, sz = nx * ny;
  const float *x = xyz, *y = xyz + vol, *z = xyz + 2 * vol;
  float best = 1.0e10f;
  ijk[0] = ijk[1] = ijk[2] = -1;
  *vpmax = 0.0f;
  *lmin = 0.0f;
  for (int k = 3; k < 3 + nk; k++)
    for (int j = 3; j < 3 + nj; j++)
      for (int i = 3; i < 3 + ni; i++) {
        size_t p = i + j * sy + k * sz;
        float vp;
        if (medium_type == ORC_MEDIUM_VTI) {
          float vpv = (float)sqrt(vols[vol + p] / vols[2 * vol + p]);
          float vph = (float)sqrt(vols[p] / vols[2 * vol + p]);
          vp = vph > vpv ? vph : vpv;
        } else if (medium_type == ORC_MEDIUM_ANISO) {
          vp = (float)sqrt(vols[p] / vols[vol + p]);
        } else {
          vp = (float)sqrt(((double)vols[p] + 2.0 * (double)vols[vol + p]) /
                           (double)vols[2 * vol + p]);
          if (medium_type == ORC_MEDIUM_VISCO) {
            const float *yv = vols + 3 * vol;   /* Ylam_n, Ymu_n interleaved per body */
            float ymax = yv[p];
            for (int n = 0; n < nmaxwell; n++) {
              float yl = yv[2 * n * vol + p], ym = yv[(2 * n + 1) * vol + p];
              ymax = ymax > yl ? ymax : yl;
              ymax = ymax > ym ? ymax : ym;
            }
            float vvis = (float)((double)vp * (1.0 + sqrt((double)(nmaxwell * ymax))));
            vp = vvis > vp ? vvis : vp;
          }
        }
        float L = 1.0e20f;
        float p0[3] = {x[p], y[p], z[p]};
        for (int kk = -1; kk <= 1; kk += 2)
          for (int jj = -1; jj <= 1; jj += 2)
            for (int ii = -1; ii <= 1; ii += 2) {
              size_t q1 = p - ii, q2 = p - jj * sy, q3 = p - kk * sz;
              float p1[3] = {x[q1], y[q1], z[q1]};
              float p2[3] = {x[q2], y[q2], z[q2]};
              float p3[3] = {x[q3], y[q3], z[q3]};
              float l = orc_dist_point2plane(p0, p1, p2, p3);
              if (L > l) L = l;
            }
        float dt = cfl / vp * L;
        if (dt < best) {
          best = dt;
          ijk[0] = i; ijk[1] = j; ijk[2] = k;
          *vpmax = vp;
          *lmin = L;
        }
      }
  *dtmax = best;
}

/* ------------------------------------------------------------------------ */
/* setup: metric, gd_t.c:191-402 (7-point centred op, cross products,       */
/* ghost mirroring in x1,x2,y1,y2,z1,z2 order)                              */
/* ------------------------------------------------------------------------ */

void orc_metric(int ni, int nj, int nk, const float *xyz, float *metric10)
{
  size_t nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = nx * ny * nz, sl = nx, ss = nx * ny;
  const float *x3d = xyz, *y3d = xyz + vol, *z3d = xyz + 2 * vol;
  float c[7];
  for (int n = 0; n < 7; n++) c[n] = (float)fdc_coef[n];
  size_t sh[3] = {1, sl, ss};
  int ni1 = 3, ni2 = 3 + ni - 1, nj1 = 3, nj2 = 3 + nj - 1, nk1 = 3, nk2 = 3 + nk - 1;
  for (size_t k = nk1; k <= (size_t)nk2; k++)
    for (size_t j = nj1; j <= (size_t)nj2; j++)
      for (size_t i = ni1; i <= (size_t)ni2; i++) {
        size_t p = i + j * sl + k * ss;
        float d[3][3]; /* d[axis][coord]: x_xi,y_xi,z_xi / x_et.. / x_zt.. */
        for (int a = 0; a < 3; a++) {
          const float *v[3] = {x3d, y3d, z3d};
          for (int cc = 0; cc < 3; cc++) {
            float s = c[0] * v[cc][p + (long)(-3) * (long)sh[a]];
            for (int n = 1; n < 7; n++) s += c[n] * v[cc][p + (long)(n - 3) * (long)sh[a]];
            d[a][cc] = s;
          }
        }
        float *v1 = d[0], *v2 = d[1], *v3 = d[2], g[3];
        g[0] = v1[1] * v2[2] - v1[2] * v2[1];
        g[1] = v1[2] * v2[0] - v1[0] * v2[2];
        g[2] = v1[0] * v2[1] - v1[1] * v2[0];
        float jac = 0.0;
        for (int n = 0; n < 3; n++) jac += g[n] * v3[n];
        metric10[p] = jac;
        g[0] = v2[1] * v3[2] - v2[2] * v3[1];
        g[1] = v2[2] * v3[0] - v2[0] * v3[2];
        g[2] = v2[0] * v3[1] - v2[1] * v3[0];
        for (int n = 0; n < 3; n++) metric10[(1 + n) * vol + p] = g[n] / jac;
        g[0] = v3[1] * v1[2] - v3[2] * v1[1];
        g[1] = v3[2] * v1[0] - v3[0] * v1[2];
        g[2] = v3[0] * v1[1] - v3[1] * v1[0];
        for (int n = 0; n < 3; n++) metric10[(4 + n) * vol + p] = g[n] / jac;
        g[0] = v1[1] * v2[2] - v1[2] * v2[1];
        g[1] = v1[2] * v2[0] - v1[0] * v2[2];
        g[2] = v1[0] * v2[1] - v1[1] * v2[0];
        for (int n = 0; n < 3; n++) metric10[(7 + n) * vol + p] = g[n] / jac;
      }
  for (int m = 0; m < 10; m++) {
    float *v = metric10 + m * vol;
    for (size_t k = 0; k < nz; k++)
      for (size_t j = 0; j < ny; j++) {
        for (size_t i = 0; i < (size_t)ni1; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p + (ni1 - i) * 2 - 1]; }
        for (size_t i = ni2 + 1; i < nx; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p - (i - ni2) * 2 + 1]; }
      }
    for (size_t k = 0; k < nz; k++) {
      for (size_t j = 0; j < (size_t)nj1; j++)
        for (size_t i = 0; i < nx; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p + ((nj1 - j) * 2 - 1) * sl]; }
      for (size_t j = nj2 + 1; j < ny; j++)
        for (size_t i = 0; i < nx; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p - ((j - nj2) * 2 - 1) * sl]; }
    }
    for (size_t k = 0; k < (size_t)nk1; k++)
      for (size_t j = 0; j < ny; j++)
        for (size_t i = 0; i < nx; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p + ((nk1 - k) * 2 - 1) * ss]; }
    for (size_t k = nk2 + 1; k < nz; k++)
      for (size_t j = 0; j < ny; j++)
        for (size_t i = 0; i < nx; i++) { size_t p = i + j * sl + k * ss; v[p] = v[p - ((k - nk2) * 2 - 1) * ss]; }
  }
}

/* ------------------------------------------------------------------------ */
/* setup: source pieces, src_t.c:2111-2151 (Gaussian weights), :2159-2166   */
/* (Ricker), :2179-2184 (Gaussian), :2344-2383 (strike/dip/rake)            */
/* ------------------------------------------------------------------------ */

static float fgauss(float t, float a, float t0)
{
  float f = exp(-(t - t0) * (t - t0) / (a * a)) / (sqrtf(PI) * a);
  return f;
}

void orc_gauss_weights(float *w, float x0, float y0, float z0, float a,
                       int half, int half_below)
{
  int n = 0;
  for (int k = -half; k <= half; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) {
        float D1 = fgauss(i - x0, a, 0.0);
        float D2 = fgauss(j - y0, a, 0.0);
        float D3 = fgauss(k - z0, a, 0.0);
        w[n++] = D1 * D2 * D3;
      }
  float sum = 0.0;
  n = 0;
  for (int k = -half; k <= half_below; k++)
    for (int j = -half; j <= half; j++)
      for (int i = -half; i <= half; i++) sum += w[n++];
  int len = 2 * half + 1;
  for (int m = 0; m < len * len * len; m++) w[m] /= sum;
}

float orc_ricker(float t, float fc, float t0)
{
  float u = (t - t0) * 2.0 * PI * fc;
  float v = (1 - u * u / 2.0) * exp(-u * u / 4.0);
  return v;
}

float orc_ricker_deriv(float t, float fc, float t0)
{
  float u = (t - t0) * 2.0 * PI * fc;
  float v = u * (-3 + 1.0 / 2.0 * u * u) * exp(-u * u / 4.0) * PI * fc;
  return v;
}

void orc_angle2moment(float strike, float dip, float rake, float m[6])
{
  float dp = dip / 180.0 * PI, st = strike / 180.0 * PI, rk = rake / 180.0 * PI;
  float M11 = -(sin(dp) * cos(rk) * sin(2.0 * st) + sin(2.0 * dp) * sin(rk) * sin(st) * sin(st));
  float M22 = sin(dp) * cos(rk) * sin(2.0 * st) - sin(2.0 * dp) * sin(rk) * cos(st) * cos(st);
  float M33 = -(M11 + M22);
  float M12 = sin(dp) * cos(rk) * cos(2.0 * st) + 0.5 * sin(2.0 * dp) * sin(rk) * sin(2.0 * st);
  float M13 = -(cos(dp) * cos(rk) * cos(st) + cos(2.0 * dp) * sin(rk) * sin(st));
  float M23 = -(cos(dp) * cos(rk) * sin(st) - cos(2.0 * dp) * sin(rk) * cos(st));
  m[0] = M22; m[1] = M11; m[2] = M33; m[3] = -M13; m[4] = -M23; m[5] = M12;
}

/* ------------------------------------------------------------------------ */
/* solver state                                                             */
/* ------------------------------------------------------------------------ */

/* the visco-elastic medium runs the isotropic elastic RHS (lambda, mu) */
#define ISO_LIKE(o) ((o)->mt == ORC_MEDIUM_ISO || (o)->mt == ORC_MEDIUM_VISCO)

struct orc_s {
  orc_cfg cfg;
  int nx, ny, nz;
  size_t vol, sl, ss;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  real_t *metric;           /* 10 volumes */
  int mt, ncmp;             /* medium type, volumes per medium */
  int nw;                   /* wavefield components: 9, + 6 per Maxwell body */
  int nmax;                 /* visco-elastic: Maxwell bodies */
  real_t wl[ORC_MAXWELL_MAX];
  real_t *med;              /* ncmp volumes, slowness last */
  const real_t *slw;
  real_t *mats;             /* matVx2Vz2, matVy2Vz2, matF2Vz2, matD */
  int region[6][6];
  size_t aux_off[6], aux_n[6], aux_total;
  real_t *pA[6], *pB[6], *pD[6];
  real_t *srcm;             /* 6 x nsrc*max_nt*4, NULL without moment */
  real_t *srcf;             /* 3 x nsrc*max_nt*4, NULL without force */
  real_t *srcfr;            /* 3 x nsurf*max_nt*4: surface forces' rates */
  real_t *tsrc, *vsrc;      /* TxSrc..TzSrc, VxSrc..VzSrc: 3 x [ny][nx] each */
  real_t *pg, *dis;         /* peak ground motion maps, free surface only */
};

size_t orc_aux_len(const orc_t *o) { return o->aux_total; }
size_t orc_vol(const orc_t *o) { return o->vol; }
int orc_ncomp(const orc_t *o) { return o->nw; }

orc_t *orc_create(const orc_cfg *cfg, const float *metric10, const float *med)
{
  orc_t *o = (orc_t *)calloc(1, sizeof(orc_t));
  o->cfg = *cfg;
  o->mt = cfg->medium_type ? cfg->medium_type : ORC_MEDIUM_ISO;
  o->ncmp = orc_medium_ncmp(o->mt);
  o->nw = 9;
  if (o->mt == ORC_MEDIUM_VISCO) {
    o->nmax = cfg->nmaxwell;
    o->ncmp = 3 + 2 * o->nmax;
    o->nw = 9 + 6 * o->nmax;
    for (int n = 0; n < o->nmax; n++) o->wl[n] = cfg->visco_wl[n];
  }
  o->nx = cfg->ni + 6; o->ny = cfg->nj + 6; o->nz = cfg->nk + 6;
  o->sl = o->nx; o->ss = (size_t)o->nx * o->ny; o->vol = o->ss * o->nz;
  o->ni1 = 3; o->ni2 = 3 + cfg->ni - 1;
  o->nj1 = 3; o->nj2 = 3 + cfg->nj - 1;
  o->nk1 = 3; o->nk2 = 3 + cfg->nk - 1;
  o->metric = (real_t *)malloc(10 * o->vol * sizeof(real_t));
  o->med = (real_t *)malloc(o->ncmp * o->vol * sizeof(real_t));
  for (size_t p = 0; p < 10 * o->vol; p++) o->metric[p] = metric10[p];
  for (size_t p = 0; p < o->ncmp * o->vol; p++) o->med[p] = med[p];
  o->slw = o->med + (ISO_LIKE(o) ? 2 : o->ncmp - 1) * o->vol;
  /* free-surface matrices, computed in the oracle's own precision */
  size_t nm = o->ss * 9;
  o->mats = (real_t *)calloc(4 * nm, sizeof(real_t));
  if (cfg->free_top) {
    float *m = (float *)calloc(4 * nm, sizeof(float));
    if (o->mt == ORC_MEDIUM_VISCO)
      orc_dvh2dvz_vis(cfg->ni, cfg->nj, cfg->nk, metric10, med, cfg->xyz, m);
    else
      orc_dvh2dvz(o->mt, cfg->ni, cfg->nj, cfg->nk, metric10, med, m);
    for (size_t p = 0; p < 4 * nm; p++) o->mats[p] = m[p];
    free(m);
    o->pg = (real_t *)calloc(15 * o->ss, sizeof(real_t));
    o->dis = (real_t *)calloc(3 * o->ss, sizeof(real_t));
  }
  orc_pml_region(cfg->ni, cfg->nj, cfg->nk, cfg->pml_layers, cfg->pml_on, o->region);
  size_t off = 0;
  for (int f = 0; f < 6; f++) {
    int *r = o->region[f];
    o->aux_off[f] = off;
    o->aux_n[f] = 0;
    if (!cfg->pml_on[f]) continue;
    size_t n = (size_t)(r[1] - r[0] + 1) * (r[3] - r[2] + 1) * (r[5] - r[4] + 1);
    o->aux_n[f] = n;
    off += 9 * n;
    int np = cfg->pml_layers[f] + 1;
    o->pA[f] = (real_t *)malloc(np * sizeof(real_t));
    o->pB[f] = (real_t *)malloc(np * sizeof(real_t));
    o->pD[f] = (real_t *)malloc(np * sizeof(real_t));
    for (int q = 0; q < np; q++) {
      o->pA[f][q] = cfg->pml_A[f][q];
      o->pB[f][q] = cfg->pml_B[f][q];
      o->pD[f][q] = cfg->pml_D[f][q];
    }
  }
  o->aux_total = off;
  if (cfg->nsrc > 0 && cfg->src_mij) {
    size_t n = (size_t)cfg->nsrc * cfg->src_max_nt * 4;
    o->srcm = (real_t *)malloc(6 * n * sizeof(real_t));
    for (size_t p = 0; p < 6 * n; p++) o->srcm[p] = cfg->src_mij[p];
  }
  if (cfg->nsrc > 0 && cfg->src_fxyz) {
    size_t n = (size_t)cfg->nsrc * cfg->src_max_nt * 4;
    o->srcf = (real_t *)malloc(3 * n * sizeof(real_t));
    for (size_t p = 0; p < 3 * n; p++) o->srcf[p] = cfg->src_fxyz[p];
  }
  o->tsrc = (real_t *)calloc(3 * o->ss, sizeof(real_t));
  o->vsrc = (real_t *)calloc(3 * o->ss, sizeof(real_t));
  if (cfg->nsrc > 0 && cfg->src_fxyz && cfg->src_force_strict && cfg->nsurf > 0) {
    size_t n = (size_t)cfg->nsurf * cfg->src_max_nt * 4;
    o->srcfr = (real_t *)malloc(3 * n * sizeof(real_t));
    for (size_t p = 0; p < 3 * n; p++) o->srcfr[p] = cfg->src_frate[p];
  }
  return o;
}

void orc_destroy(orc_t *o)
{
  if (!o) return;
  free(o->metric); free(o->med); free(o->mats); free(o->srcm); free(o->srcf); free(o->pg);
  free(o->srcfr); free(o->tsrc); free(o->vsrc);
  free(o->dis);
  for (int f = 0; f < 6; f++) { free(o->pA[f]); free(o->pB[f]); free(o->pD[f]); }
  free(o);
}

/* wavefield component slots (wav_t.c:93-136) */
enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };


typedef struct { int len; long sh[5]; real_t c[5]; } op_t;

static void make_op(op_t *op, int lay, int dir, long stride)
{
  op->len = op_len[lay];
  for (int n = 0; n < op->len; n++) {
    op->sh[n] = op_indx[lay][dir][n] * stride;
    op->c[n] = (real_t)(float)op_coef[lay][dir][n];
  }
#if ORC_REAL_IS_DOUBLE
  for (int n = 0; n < op->len; n++) op->c[n] = op_coef[lay][dir][n];
#endif
}

/* M_FD_SHIFT order: c0*v0, then += c_n*v_n (fd_t.h:11-31) */
static inline real_t dop(const op_t *op, const real_t *v, size_t p)
{
  real_t d = op->c[0] * v[p + op->sh[0]];
  for (int n = 1; n < op->len; n++) d += op->c[n] * v[p + op->sh[n]];
  return d;
}

/* interior RHS, sv_curv_col_el_iso.c:209-441 */
static void rhs_inner(orc_t *o, const real_t *w, real_t *h,
                      const op_t *ox, const op_t *oy, const op_t *oz)
{
  size_t V = o->vol;
  const real_t *M = o->metric, *md = o->med;
#pragma omp parallel for schedule(static)
  for (int k = o->nk1; k <= o->nk2; k++)
    for (int j = o->nj1; j <= o->nj2; j++)
      for (int i = o->ni1; i <= o->ni2; i++) {
        size_t p = i + j * o->sl + k * o->ss;
        real_t D[3][9];
        for (int c = 0; c < 9; c++) {
          D[0][c] = dop(ox, w + c * V, p);
          D[1][c] = dop(oy, w + c * V, p);
          D[2][c] = dop(oz, w + c * V, p);
        }
        real_t xix = M[V + p], xiy = M[2 * V + p], xiz = M[3 * V + p];
        real_t etx = M[4 * V + p], ety = M[5 * V + p], etz = M[6 * V + p];
        real_t ztx = M[7 * V + p], zty = M[8 * V + p], ztz = M[9 * V + p];
        real_t slw = o->slw[p];
        real_t (*X) = D[0], (*Y) = D[1], (*Z) = D[2];
        if (!ISO_LIKE(o)) {
          /* sv_curv_col_el_vti.c:453-475, sv_curv_col_el_aniso.c:455-491 */
          real_t c[21], e[3][3], Dv[3][3], out[6];
          coupling_medium(o->mt, md, V, p, c);
          for (int d = 0; d < 3; d++)
            for (int m = 0; m < 3; m++) { e[d][m] = M[(1 + 3 * d + m) * V + p]; Dv[d][m] = D[d][m]; }
          stress_rates(o->mt, c, 3, (const real_t (*)[3])e, (const real_t (*)[3])Dv, out);
          for (int t = 0; t < 6; t++) h[(TXX + t) * V + p] = out[t];
        }
        h[VX * V + p] = slw * (xix * X[TXX] + xiy * X[TXY] + xiz * X[TXZ]
                              + etx * Y[TXX] + ety * Y[TXY] + etz * Y[TXZ]
                              + ztx * Z[TXX] + zty * Z[TXY] + ztz * Z[TXZ]);
        h[VY * V + p] = slw * (xix * X[TXY] + xiy * X[TYY] + xiz * X[TYZ]
                              + etx * Y[TXY] + ety * Y[TYY] + etz * Y[TYZ]
                              + ztx * Z[TXY] + zty * Z[TYY] + ztz * Z[TYZ]);
        h[VZ * V + p] = slw * (xix * X[TXZ] + xiy * X[TYZ] + xiz * X[TZZ]
                              + etx * Y[TXZ] + ety * Y[TYZ] + etz * Y[TZZ]
                              + ztx * Z[TXZ] + zty * Z[TYZ] + ztz * Z[TZZ]);
        if (!ISO_LIKE(o)) continue;
        real_t lam = md[p], mu = md[V + p];
        real_t lam2mu = lam + 2.0 * mu;
        h[TXX * V + p] = lam2mu * (xix * X[VX] + etx * Y[VX] + ztx * Z[VX])
                       + lam * (xiy * X[VY] + ety * Y[VY] + zty * Z[VY]
                                + xiz * X[VZ] + etz * Y[VZ] + ztz * Z[VZ]);
        h[TYY * V + p] = lam2mu * (xiy * X[VY] + ety * Y[VY] + zty * Z[VY])
                       + lam * (xix * X[VX] + etx * Y[VX] + ztx * Z[VX]
                                + xiz * X[VZ] + etz * Y[VZ] + ztz * Z[VZ]);
        h[TZZ * V + p] = lam2mu * (xiz * X[VZ] + etz * Y[VZ] + ztz * Z[VZ])
                       + lam * (xix * X[VX] + etx * Y[VX] + ztx * Z[VX]
                                + xiy * X[VY] + ety * Y[VY] + zty * Z[VY]);
        h[TXY * V + p] = mu * (xiy * X[VX] + xix * X[VY]
                               + ety * Y[VX] + etx * Y[VY]
                               + zty * Z[VX] + ztx * Z[VY]);
        h[TXZ * V + p] = mu * (xiz * X[VX] + xix * X[VZ]
                               + etz * Y[VX] + etx * Y[VZ]
                               + ztz * Z[VX] + ztx * Z[VZ]);
        h[TYZ * V + p] = mu * (xiz * X[VY] + xiy * X[VZ]
                               + etz * Y[VY] + ety * Y[VZ]
                               + ztz * Z[VY] + zty * Z[VZ]);
      }
}

/* traction image for hV on the top layers, sv_curv_col_el.c:31-305, with
 * the surface traction TxSrc..TzSrc of the strict surface forces (zero
 * without them) at the surface row and in the mirror */
static void rhs_timg(orc_t *o, const real_t *w, real_t *h,
                     const op_t *ox, const op_t *oy, const op_t *oz, const int *zidx)
{
  size_t V = o->vol;
  const real_t *M = o->metric;
  const real_t *jac = M;
  int zlen = oz->len;
  int k_min = o->nk2 - zidx[zlen - 1];
  /* traction rows: (Txx,Txy,Txz), (Txy,Tyy,Tyz), (Txz,Tyz,Tzz) */
  static const int trow[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
  for (int k = k_min; k <= o->nk2; k++) {
    int n_free = o->nk2 - k - zidx[0];
#pragma omp parallel for schedule(static)
    for (int j = o->nj1; j <= o->nj2; j++)
      for (int i = o->ni1; i <= o->ni2; i++) {
        size_t p = i + j * o->sl + k * o->ss;
        real_t slwjac = o->slw[p] / jac[p];
        for (int comp = 0; comp < 3; comp++) {
          const real_t *T0 = w + trow[comp][0] * V, *T1 = w + trow[comp][1] * V,
                       *T2 = w + trow[comp][2] * V;
          real_t vx[5], vy[5], vz[5];
          for (int n = 0; n < ox->len; n++) {
            size_t q = p + ox->sh[n];
            vx[n] = jac[q] * (M[V + q] * T0[q] + M[2 * V + q] * T1[q] + M[3 * V + q] * T2[q]);
          }
          for (int n = 0; n < oy->len; n++) {
            size_t q = p + oy->sh[n];
            vy[n] = jac[q] * (M[4 * V + q] * T0[q] + M[5 * V + q] * T1[q] + M[6 * V + q] * T2[q]);
          }
          for (int n = 0; n < n_free; n++) {
            size_t q = p + oz->sh[n];
            vz[n] = jac[q] * (M[7 * V + q] * T0[q] + M[8 * V + q] * T1[q] + M[9 * V + q] * T2[q]);
          }
          const real_t ts = o->tsrc[comp * o->ss + i + j * o->sl];
          vz[n_free] = ts;   /* surface traction */
          for (int n = n_free + 1; n < zlen; n++) {
            int m = n_free - (n - n_free);
            real_t below;
            if (m >= 0) {
              below = vz[m];
            } else {
              /* mirror point lies below the operator window: evaluate it at
               * its grid point k + zidx[n] - 2(n - n_free) */
              size_t q = p + (long)(zidx[n] - 2 * (n - n_free)) * (long)o->ss;
              below = jac[q] * (M[7 * V + q] * T0[q] + M[8 * V + q] * T1[q] + M[9 * V + q] * T2[q]);
            }
            vz[n] = 2.0 * ts - below;
          }
          real_t DxT = ox->c[0] * vx[0], DyT = oy->c[0] * vy[0], DzT = oz->c[0] * vz[0];
          for (int n = 1; n < ox->len; n++) DxT += ox->c[n] * vx[n];
          for (int n = 1; n < oy->len; n++) DyT += oy->c[n] * vy[n];
          for (int n = 1; n < zlen; n++) DzT += oz->c[n] * vz[n];
          h[comp * V + p] = (DxT + DyT + DzT) * slwjac;
        }
      }
  }
}

/* Hooke's law from 9 velocity derivatives (the expression of rhs_inner) */
static inline void hooke(real_t *h, size_t V, size_t p, const real_t *m9,
                         real_t lam, real_t lam2mu, real_t mu,
                         const real_t *X, const real_t *Y, const real_t *Z)
{
  real_t xix = m9[0], xiy = m9[1], xiz = m9[2];
  real_t etx = m9[3], ety = m9[4], etz = m9[5];
  real_t ztx = m9[6], zty = m9[7], ztz = m9[8];
  h[TXX * V + p] = lam2mu * (xix * X[0] + etx * Y[0] + ztx * Z[0])
                 + lam * (xiy * X[1] + ety * Y[1] + zty * Z[1]
                          + xiz * X[2] + etz * Y[2] + ztz * Z[2]);
  h[TYY * V + p] = lam2mu * (xiy * X[1] + ety * Y[1] + zty * Z[1])
                 + lam * (xix * X[0] + etx * Y[0] + ztx * Z[0]
                          + xiz * X[2] + etz * Y[2] + ztz * Z[2]);
  h[TZZ * V + p] = lam2mu * (xiz * X[2] + etz * Y[2] + ztz * Z[2])
                 + lam * (xix * X[0] + etx * Y[0] + ztx * Z[0]
                          + xiy * X[1] + ety * Y[1] + zty * Z[1]);
  h[TXY * V + p] = mu * (xiy * X[0] + xix * X[1] + ety * Y[0] + etx * Y[1] + zty * Z[0] + ztx * Z[1]);
  h[TXZ * V + p] = mu * (xiz * X[0] + xix * X[2] + etz * Y[0] + etx * Y[2] + ztz * Z[0] + ztx * Z[2]);
  h[TYZ * V + p] = mu * (xiz * X[1] + xiy * X[2] + etz * Y[1] + ety * Y[2] + ztz * Z[1] + zty * Z[2]);
}

/* hT on the top 3 layers, sv_curv_col_el_iso.c:452-634; at the surface the
 * isotropic forms add matF2Vz2 times the strict surface forces' velocity
 * source VxSrc..VzSrc (sv_curv_col_el_iso.c:583-592; the VTI and anisotropic
 * forms have no such term) */
static void rhs_vlow(orc_t *o, const real_t *w, real_t *h,
                     const op_t *ox, const op_t *oy, const op_t ozl[4])
{
  size_t V = o->vol;
  const real_t *M = o->metric, *md = o->med;
  const real_t *mVx = o->mats, *mVy = o->mats + o->ss * 9;
  for (int n = 0; n < 3; n++) {
    int k = o->nk2 - n;
#pragma omp parallel for schedule(static)
    for (int j = o->nj1; j <= o->nj2; j++)
      for (int i = o->ni1; i <= o->ni2; i++) {
        size_t p = i + j * o->sl + k * o->ss;
        real_t m9[9];
        for (int q = 0; q < 9; q++) m9[q] = M[(1 + q) * V + p];
        real_t X[3], Y[3], Z[3];
        for (int c = 0; c < 3; c++) {
          X[c] = dop(ox, w + c * V, p);
          Y[c] = dop(oy, w + c * V, p);
        }
        if (k == o->nk2) {
          size_t ij = (i + j * o->sl) * 9;
          for (int r = 0; r < 3; r++)
            Z[r] = mVx[ij + 3 * r + 0] * X[0] + mVx[ij + 3 * r + 1] * X[1] + mVx[ij + 3 * r + 2] * X[2]
                 + mVy[ij + 3 * r + 0] * Y[0] + mVy[ij + 3 * r + 1] * Y[1] + mVy[ij + 3 * r + 2] * Y[2];
          if (ISO_LIKE(o)) {
            const real_t *mF = o->mats + 2 * o->ss * 9;
            size_t q = i + j * o->sl;
            const real_t *vs = o->vsrc;
            for (int r = 0; r < 3; r++)
              Z[r] += mF[ij + 3 * r + 0] * vs[q] + mF[ij + 3 * r + 1] * vs[o->ss + q]
                    + mF[ij + 3 * r + 2] * vs[2 * o->ss + q];
          }
        } else {
          for (int c = 0; c < 3; c++) Z[c] = dop(&ozl[n], w + c * V, p);
        }
        if (ISO_LIKE(o)) {
          real_t lam = md[p], mu = md[V + p];
          real_t lam2mu = lam + 2.0 * mu;
          hooke(h, V, p, m9, lam, lam2mu, mu, X, Y, Z);
        } else {
          /* sv_curv_col_el_vti.c:566-586, sv_curv_col_el_aniso.c:676-702 */
          real_t c[21], Dv[3][3], out[6];
          coupling_medium(o->mt, md, V, p, c);
          for (int m = 0; m < 3; m++) { Dv[0][m] = X[m]; Dv[1][m] = Y[m]; Dv[2][m] = Z[m]; }
          stress_rates(o->mt, c, 3, (const real_t (*)[3])m9, (const real_t (*)[3])Dv, out);
          for (int t = 0; t < 6; t++) h[(TXX + t) * V + p] = out[t];
        }
      }
  }
}

/* CFS-PML ADE, sv_curv_col_el_iso.c:645-1146 */
static void rhs_pml(orc_t *o, const real_t *w, real_t *h, const real_t *aux,
                    real_t *aux_h, const op_t *ox, const op_t *oy, const op_t *oz)
{
  size_t V = o->vol;
  const real_t *M = o->metric, *md = o->med;
  const real_t *mVx = o->mats, *mVy = o->mats + o->ss * 9;
  for (int f = 0; f < 6; f++) {
    if (!o->cfg.pml_on[f]) continue;
    int idim = f / 2;
    const int *r = o->region[f];
    const op_t *op = idim == 0 ? ox : (idim == 1 ? oy : oz);
    const real_t *mt = M + (1 + 3 * idim) * V;      /* xi_*, eta_* or zeta_* */
    size_t na = o->aux_n[f];
    const real_t *a = aux + o->aux_off[f];
    real_t *ah = aux_h + o->aux_off[f];
    size_t fnx = r[1] - r[0] + 1, fny = r[3] - r[2] + 1;
#pragma omp parallel for schedule(static)
    for (int k = r[4]; k <= r[5]; k++)
      for (int j = r[2]; j <= r[3]; j++)
        for (int i = r[0]; i <= r[1]; i++) {
          size_t ia = (i - r[0]) + (j - r[2]) * fnx + (k - r[4]) * fnx * fny;
          size_t p = i + j * o->sl + k * o->ss;
          int ic = idim == 0 ? i - r[0] : (idim == 1 ? j - r[2] : k - r[4]);
          real_t cD = o->pD[f][ic], cA = o->pA[f][ic], cB = o->pB[f][ic];
          real_t cBm1 = cB - 1.0;
          real_t e1 = mt[p], e2 = mt[V + p], e3 = mt[2 * V + p];
          real_t slw = o->slw[p];
          real_t lam = 0, mu = 0, lam2mu = 0, cm[21];
          if (ISO_LIKE(o)) {
            lam = md[p]; mu = md[V + p];
            lam2mu = lam + 2.0 * mu;
          } else {
            coupling_medium(o->mt, md, V, p, cm);
          }
          real_t D[9];
          for (int c = 0; c < 9; c++) D[c] = dop(op, w + c * V, p);
          real_t rr[9];
          rr[VX] = slw * (e1 * D[TXX] + e2 * D[TXY] + e3 * D[TXZ]);
          rr[VY] = slw * (e1 * D[TXY] + e2 * D[TYY] + e3 * D[TYZ]);
          rr[VZ] = slw * (e1 * D[TXZ] + e2 * D[TYZ] + e3 * D[TZZ]);
          if (ISO_LIKE(o)) {
            rr[TXX] = lam2mu * e1 * D[VX] + lam * e2 * D[VY] + lam * e3 * D[VZ];
            rr[TYY] = lam * e1 * D[VX] + lam2mu * e2 * D[VY] + lam * e3 * D[VZ];
            rr[TZZ] = lam * e1 * D[VX] + lam * e2 * D[VY] + lam2mu * e3 * D[VZ];
            rr[TXY] = mu * (e2 * D[VX] + e1 * D[VY]);
            rr[TXZ] = mu * (e3 * D[VX] + e1 * D[VZ]);
            rr[TYZ] = mu * (e3 * D[VY] + e2 * D[VZ]);
          } else {
            /* one direction of the Hooke sums (sv_curv_col_el_vti.c:757-765) */
            real_t e[1][3] = {{e1, e2, e3}}, Dv[1][3] = {{D[VX], D[VY], D[VZ]}};
            stress_rates(o->mt, cm, 1, (const real_t (*)[3])e, (const real_t (*)[3])Dv, rr + TXX);
          }
          /* reference update order: Vx,Vy,Vz,Txx,Tyy,Tzz,Txz,Tyz,Txy */
          static const int ord[9] = {VX, VY, VZ, TXX, TYY, TZZ, TXZ, TYZ, TXY};
          for (int q = 0; q < 9; q++) {
            int c = ord[q];
            h[c * V + p] += cBm1 * rr[c] - cB * a[c * na + ia];
          }
          for (int q = 0; q < 9; q++) {
            int c = ord[q];
            ah[c * na + ia] = cD * rr[c] - cA * a[c * na + ia];
          }
          if (idim < 2 && o->cfg.free_top && k == o->nk2) {
            size_t ij = (i + j * o->sl) * 9;
            const real_t *mm = idim == 0 ? mVx : mVy;
            real_t Dz[3];
            for (int rr3 = 0; rr3 < 3; rr3++)
              Dz[rr3] = mm[ij + 3 * rr3 + 0] * D[VX] + mm[ij + 3 * rr3 + 1] * D[VY]
                      + mm[ij + 3 * rr3 + 2] * D[VZ];
            real_t ztx = M[7 * V + p], zty = M[8 * V + p], ztz = M[9 * V + p];
            real_t s[9];
            if (ISO_LIKE(o)) {
              s[TXX] = lam2mu * (ztx * Dz[0]) + lam * (zty * Dz[1] + ztz * Dz[2]);
              s[TYY] = lam2mu * (zty * Dz[1]) + lam * (ztx * Dz[0] + ztz * Dz[2]);
              s[TZZ] = lam2mu * (ztz * Dz[2]) + lam * (ztx * Dz[0] + zty * Dz[1]);
              s[TXY] = mu * (zty * Dz[0] + ztx * Dz[1]);
              s[TXZ] = mu * (ztz * Dz[0] + ztx * Dz[2]);
              s[TYZ] = mu * (ztz * Dz[1] + zty * Dz[2]);
            } else {
              /* sv_curv_col_el_vti.c:817-822 */
              real_t e[1][3] = {{ztx, zty, ztz}}, Dv[1][3] = {{Dz[0], Dz[1], Dz[2]}};
              stress_rates(o->mt, cm, 1, (const real_t (*)[3])e, (const real_t (*)[3])Dv, s + TXX);
            }
            static const int ordT[6] = {TXX, TYY, TZZ, TXZ, TYZ, TXY};
            for (int q = 0; q < 6; q++) {
              int c = ordT[q];
              h[c * V + p] += (cB - 1.0) * s[c];
            }
            for (int q = 0; q < 6; q++) {
              int c = ordT[q];
              ah[c * na + ia] += cD * s[c];
            }
          }
        }
  }
}

/* moment and force sources, sv_curv_col_el.c:312-479: a force adds
 * f * (coef) * slw / jac to the velocity rates (float division first), a
 * moment subtracts M * (coef) / jac from the stress rates */
static void add_force(orc_t *o, const real_t *F, size_t p, real_t coef, real_t *h)
{
  size_t V = o->vol;
  real_t Vf = coef * o->slw[p] / o->metric[p];
  h[VX * V + p] += F[0] * Vf;
  h[VY * V + p] += F[1] * Vf;
  h[VZ * V + p] += F[2] * Vf;
}

static void rhs_src(orc_t *o, int it, int istage, real_t *h)
{
  const orc_cfg *c = &o->cfg;
  size_t V = o->vol;
  const real_t *jac = o->metric;
  size_t nper = (size_t)c->nsrc * c->src_max_nt * 4;
  int half = c->src_ext_half;
  int len = 2 * half + 1;
  float *ext = (float *)malloc(sizeof(float) * len * len * len);
  for (int is = 0; is < c->nsrc; is++) {
    if (!(it >= c->src_it_begin[is] && it <= c->src_it_end[is])) continue;
    size_t ip = (size_t)is * c->src_max_nt * 4 + (size_t)(it - c->src_it_begin[is]) * 4 + istage;
    real_t Mv[6] = {0, 0, 0, 0, 0, 0}, Fv[3] = {0, 0, 0};
    if (o->srcm)
      for (int m = 0; m < 6; m++) Mv[m] = o->srcm[m * nper + ip];
    if (o->srcf)
      for (int m = 0; m < 3; m++) Fv[m] = o->srcf[m * nper + ip];
    int si = c->src_ijk[3 * is], sj = c->src_ijk[3 * is + 1], sk = c->src_ijk[3 * is + 2];
    /* M order Mxx,Myy,Mzz,Mxz,Myz,Mxy onto slots Txx,Tyy,Tzz,Txz,Tyz,Txy */
    static const int slot[6] = {TXX, TYY, TZZ, TXZ, TYZ, TXY};
    if (!c->src_gauss) {
      size_t p = si + sj * o->sl + sk * o->ss;
      if (o->srcf && (!c->src_force_strict || sk < o->nk2)) {
        real_t Vf = o->slw[p] / jac[p];
        h[VX * V + p] += Fv[0] * Vf;
        h[VY * V + p] += Fv[1] * Vf;
        h[VZ * V + p] += Fv[2] * Vf;
      }
      if (o->srcm) {
        real_t rjac = 1.0 / jac[p];
        for (int m = 0; m < 6; m++) h[slot[m] * V + p] -= Mv[m] * rjac;
      }
    } else {
      int kb = (sk + half < o->nk2) ? half : o->nk2 - sk;
      orc_gauss_weights(ext, c->src_inc[3 * is], c->src_inc[3 * is + 1], c->src_inc[3 * is + 2],
                        c->src_ext_coef, half, kb);
      int ie = 0;
      for (int ke = -half; ke <= kb; ke++)
        for (int je = -half; je <= half; je++)
          for (int iee = -half; iee <= half; iee++, ie++) {
            int i = si + iee, j = sj + je, k = sk + ke;
            if (i < o->ni1 || i > o->ni2 || j < o->nj1 || j > o->nj2 || k < o->nk1 || k > o->nk2)
              continue;
            size_t p = i + j * o->sl + k * o->ss;
            real_t coef = ext[ie];
            if (o->srcf && (!c->src_force_strict || k < o->nk2)) add_force(o, Fv, p, coef, h);
            if (o->srcm) {
              real_t rjac = coef / jac[p];
              for (int m = 0; m < 6; m++) h[slot[m] * V + p] -= Mv[m] * rjac;
            }
          }
    }
  }
  free(ext);
}

/* dd sources at grid points, sv_curv_col_el_rhs_srcdd (sv_curv_col_el.c:
 * 486-555): for step it < dd_max_nt, each source in order adds vi * (slw /
 * jac) to the velocity rates and subtracts mij * (1 / jac), the reciprocal
 * in double, from the stress rates (mij order Mxx,Myy,Mzz,Myz,Mxz,Mxy) */
static void rhs_srcdd(orc_t *o, int it, int istage, real_t *h)
{
  const orc_cfg *c = &o->cfg;
  if (c->ndd <= 0 || it >= c->dd_max_nt) return;
  size_t V = o->vol;
  const real_t *jac = o->metric;
  if (c->dd_vi) {
    const float *vi = c->dd_vi + ((size_t)it * 4 + istage) * c->ndd * 3;
    for (int is = 0; is < c->ndd; is++) {
      size_t p = (size_t)c->dd_indx[is];
      real_t Vf = o->slw[p] / jac[p];
      h[VX * V + p] += vi[3 * is + 0] * Vf;
      h[VY * V + p] += vi[3 * is + 1] * Vf;
      h[VZ * V + p] += vi[3 * is + 2] * Vf;
    }
  }
  if (c->dd_mij) {
    const float *m = c->dd_mij + ((size_t)it * 4 + istage) * c->ndd * 6;
    for (int is = 0; is < c->ndd; is++) {
      size_t p = (size_t)c->dd_indx[is];
      real_t rjac = 1.0 / jac[p];
      h[TXX * V + p] -= m[6 * is + 0] * rjac;
      h[TYY * V + p] -= m[6 * is + 1] * rjac;
      h[TZZ * V + p] -= m[6 * is + 2] * rjac;
      h[TYZ * V + p] -= m[6 * is + 3] * rjac;
      h[TXZ * V + p] -= m[6 * is + 4] * rjac;
      h[TXY * V + p] -= m[6 * is + 5] * rjac;
    }
  }
}

/* GMB attenuation after the PML terms, sv_curv_col_vis_iso.c:251-342: memory
 * variable rates and the memory terms subtracted from the stress rates.
 * J slots of body n: 9+6n + {Jxx, Jyy, Jzz, Jyz, Jxz, Jxy} */
static void rhs_atten(orc_t *o, const real_t *w, real_t *h)
{
  size_t V = o->vol;
  const real_t *md = o->med;
  int nm = o->nmax;
#pragma omp parallel for schedule(static)
  for (int k = o->nk1; k <= o->nk2; k++)
    for (int j = o->nj1; j <= o->nj2; j++)
      for (int i = o->ni1; i <= o->ni2; i++) {
        size_t p = i + j * o->sl + k * o->ss;
        real_t lam = md[p], mu = md[V + p];
        real_t hTxx = h[TXX * V + p], hTyy = h[TYY * V + p], hTzz = h[TZZ * V + p];
        real_t hTxy = h[TXY * V + p], hTxz = h[TXZ * V + p], hTyz = h[TYZ * V + p];
        real_t sum_hxyz = (hTxx + hTyy + hTzz) / (3 * lam + 2 * mu);
        real_t EVxx = ((2.0 * hTxx - hTyy - hTzz) / (2 * mu) + sum_hxyz) / 3;
        real_t EVyy = ((2.0 * hTyy - hTxx - hTzz) / (2 * mu) + sum_hxyz) / 3;
        real_t EVzz = ((2.0 * hTzz - hTxx - hTyy) / (2 * mu) + sum_hxyz) / 3;
        real_t EVxy = hTxy / mu * 0.5;
        real_t EVxz = hTxz / mu * 0.5;
        real_t EVyz = hTyz / mu * 0.5;
        real_t sum_Jxyz = 0.0, sum_Jxx = 0.0, sum_Jyy = 0.0, sum_Jzz = 0.0;
        real_t sum_Jxy = 0.0, sum_Jxz = 0.0, sum_Jyz = 0.0;
        for (int n = 0; n < nm; n++) {
          size_t b = (size_t)(9 + 6 * n) * V + p;
          real_t Jxx = w[b], Jyy = w[b + V], Jzz = w[b + 2 * V];
          real_t Jyz = w[b + 3 * V], Jxz = w[b + 4 * V], Jxy = w[b + 5 * V];
          real_t wl = o->wl[n];
          h[b] = wl * (EVxx - Jxx);
          h[b + V] = wl * (EVyy - Jyy);
          h[b + 2 * V] = wl * (EVzz - Jzz);
          h[b + 5 * V] = wl * (EVxy - Jxy);
          h[b + 4 * V] = wl * (EVxz - Jxz);
          h[b + 3 * V] = wl * (EVyz - Jyz);
        }
        for (int n = 0; n < nm; n++) {
          size_t b = (size_t)(9 + 6 * n) * V + p;
          real_t Jxx = w[b], Jyy = w[b + V], Jzz = w[b + 2 * V];
          real_t Jyz = w[b + 3 * V], Jxz = w[b + 4 * V], Jxy = w[b + 5 * V];
          real_t Yl = md[(3 + 2 * n) * V + p], Ym = md[(4 + 2 * n) * V + p];
          sum_Jxyz += Yl * (Jxx + Jyy + Jzz);
          sum_Jxx += Ym * Jxx;
          sum_Jyy += Ym * Jyy;
          sum_Jzz += Ym * Jzz;
          sum_Jxy += Ym * Jxy;
          sum_Jxz += Ym * Jxz;
          sum_Jyz += Ym * Jyz;
        }
        real_t mem_Txx = lam * sum_Jxyz + 2.0 * mu * sum_Jxx;
        real_t mem_Tyy = lam * sum_Jxyz + 2.0 * mu * sum_Jyy;
        real_t mem_Tzz = lam * sum_Jxyz + 2.0 * mu * sum_Jzz;
        real_t mem_Txy = 2.0 * mu * sum_Jxy;
        real_t mem_Txz = 2.0 * mu * sum_Jxz;
        real_t mem_Tyz = 2.0 * mu * sum_Jyz;
        h[TXX * V + p] = hTxx - mem_Txx;
        h[TYY * V + p] = hTyy - mem_Tyy;
        h[TZZ * V + p] = hTzz - mem_Tzz;
        h[TXY * V + p] = hTxy - mem_Txy;
        h[TXZ * V + p] = hTxz - mem_Txz;
        h[TYZ * V + p] = hTyz - mem_Tyz;
      }
}

/* rotation of the surface-row stresses after the last stage,
 * sv_curv_col_vis_iso_free (sv_curv_col_vis_iso.c:510-623) */
static void vis_free(orc_t *o, real_t *w)
{
  size_t V = o->vol;
  const real_t *mD = o->mats + 3 * o->ss * 9;
  real_t *Txx = w + TXX * V, *Tyy = w + TYY * V, *Tzz = w + TZZ * V;
  real_t *Txy = w + TXY * V, *Txz = w + TXZ * V, *Tyz = w + TYZ * V;
  for (int j = o->nj1; j <= o->nj2; j++)
    for (int i = o->ni1; i <= o->ni2; i++) {
      size_t p = i + j * o->sl + (size_t)o->nk2 * o->ss, ij = (i + j * o->sl) * 9;
      real_t D[3][3], DT[3][3], Tl[3][3], DTl[3][3], Tg[3][3];
      for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) { D[r][c] = mD[ij + 3 * r + c]; DT[c][r] = D[r][c]; }
      real_t d11 = D[0][0], d12 = D[0][1], d13 = D[0][2];
      real_t d21 = D[1][0], d22 = D[1][1], d23 = D[1][2];
      Tl[0][0] = d11 * d11 * Txx[p] + d12 * d12 * Tyy[p] + d13 * d13 * Tzz[p]
               + 2 * (d11 * d12 * Txy[p] + d11 * d13 * Txz[p] + d12 * d13 * Tyz[p]);
      Tl[0][1] = d11 * d21 * Txx[p] + d12 * d22 * Tyy[p] + d13 * d23 * Tzz[p]
               + (d11 * d22 + d12 * d21) * Txy[p] + (d11 * d23 + d21 * d13) * Txz[p]
               + (d12 * d23 + d22 * d13) * Tyz[p];
      Tl[1][1] = d21 * d21 * Txx[p] + d22 * d22 * Tyy[p] + d23 * d23 * Tzz[p]
               + 2 * (d21 * d22 * Txy[p] + d21 * d23 * Txz[p] + d22 * d23 * Tyz[p]);
      Tl[1][0] = Tl[0][1];
      Tl[0][2] = 0.0; Tl[1][2] = 0.0; Tl[2][0] = 0.0; Tl[2][1] = 0.0; Tl[2][2] = 0.0;
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          DTl[a][b] = 0.0;
          for (int c = 0; c < 3; c++) DTl[a][b] += DT[a][c] * Tl[c][b];
        }
      for (int a = 0; a < 3; a++)
        for (int b = 0; b < 3; b++) {
          Tg[a][b] = 0.0;
          for (int c = 0; c < 3; c++) Tg[a][b] += DTl[a][c] * D[c][b];
        }
      Txx[p] = Tg[0][0];
      Tyy[p] = Tg[1][1];
      Tzz[p] = Tg[2][2];
      Txy[p] = Tg[0][1];
      Txz[p] = Tg[0][2];
      Tyz[p] = Tg[1][2];
    }
}

/* surface traction and velocity source of the strict surface forces for
 * step it, stage istage, src_set_surface_layer_for_force (src_t.c:154-314):
 * the slices start at zero; each listed source active at it adds its force
 * (times its weight on the surface row, for the Gaussian) to TxSrc.. and its
 * force rate (times the weight) over the surface jacobian to VxSrc.. */
static void surface_layer(orc_t *o, int it, int istage)
{
  const orc_cfg *c = &o->cfg;
  for (size_t q = 0; q < 3 * o->ss; q++) { o->tsrc[q] = 0; o->vsrc[q] = 0; }
  if (!o->srcfr) return;
  const real_t *jac = o->metric;
  size_t nper = (size_t)c->nsrc * c->src_max_nt * 4;
  size_t nperr = (size_t)c->nsurf * c->src_max_nt * 4;
  int half = c->src_ext_half, len = 2 * half + 1;
  float *ext = (float *)malloc(sizeof(float) * len * len * len);
  for (int n = 0; n < c->nsurf; n++) {
    int is = c->src_surf_idx[n];
    if (it < c->src_it_begin[is] || it > c->src_it_end[is]) continue;
    int si = c->src_ijk[3 * is], sj = c->src_ijk[3 * is + 1], sk = c->src_ijk[3 * is + 2];
    size_t itr = (size_t)(it - c->src_it_begin[is]) * 4 + istage;
    size_t ip = (size_t)is * c->src_max_nt * 4 + itr;
    size_t ipr = (size_t)n * c->src_max_nt * 4 + itr;
    real_t fv[3], fr[3];
    for (int m = 0; m < 3; m++) {
      fv[m] = o->srcf[m * nper + ip];
      fr[m] = o->srcfr[m * nperr + ipr];
    }
    if (!c->src_gauss) {
      size_t p = si + sj * o->sl + sk * o->ss, q = si + sj * o->sl;
      for (int m = 0; m < 3; m++) {
        o->tsrc[m * o->ss + q] += fv[m];
        o->vsrc[m * o->ss + q] += fr[m] / jac[p];
      }
    } else {
      int ke = o->nk2 - sk;
      orc_gauss_weights(ext, c->src_inc[3 * is], c->src_inc[3 * is + 1], c->src_inc[3 * is + 2],
                        c->src_ext_coef, half, ke);
      int ie = (ke + half) * len * len;
      int k = sk + ke;
      for (int je = -half; je <= half; je++)
        for (int iee = -half; iee <= half; iee++, ie++) {
          int i = si + iee, j = sj + je;
          real_t coef = ext[ie];
          size_t p = i + j * o->sl + k * o->ss, q = i + j * o->sl;
          for (int m = 0; m < 3; m++) {
            o->tsrc[m * o->ss + q] += fv[m] * coef;
            o->vsrc[m * o->ss + q] += fr[m] * coef / jac[p];
          }
        }
    }
  }
  free(ext);
}

void orc_surface_force_layer(orc_t *o, int it, int istage, real_t *ts, real_t *vs)
{
  surface_layer(o, it, istage);
  memcpy(ts, o->tsrc, 3 * o->ss * sizeof(real_t));
  memcpy(vs, o->vsrc, 3 * o->ss * sizeof(real_t));
}

void orc_onestage(orc_t *o, int it, int ipair, int istage,
                  const real_t *w_cur, const real_t *aux_cur,
                  real_t *rhs, real_t *aux_rhs)
{
  /* the driver sets the surface-force slices before every stage
   * (drv_rk_curv_col.c:227-230) */
  surface_layer(o, it, istage);
  int dx = (pair_flags[ipair][0] + istage) % 2;
  int dy = (pair_flags[ipair][1] + istage) % 2;
  int dz = (pair_flags[ipair][2] + istage) % 2;
  op_t ox, oy, oz, ozl[4];
  make_op(&ox, 3, dx, 1);
  make_op(&oy, 3, dy, (long)o->sl);
  make_op(&oz, 3, dz, (long)o->ss);
  for (int l = 0; l < 4; l++) make_op(&ozl[l], l, dz, (long)o->ss);
  rhs_inner(o, w_cur, rhs, &ox, &oy, &oz);
  if (o->cfg.free_top) {
    rhs_timg(o, w_cur, rhs, &ox, &oy, &oz, op_indx[3][dz]);
    rhs_vlow(o, w_cur, rhs, &ox, &oy, ozl);
  }
  if (o->aux_total) rhs_pml(o, w_cur, rhs, aux_cur, aux_rhs, &ox, &oy, &oz);
  if (o->mt == ORC_MEDIUM_VISCO) rhs_atten(o, w_cur, rhs);
  if (o->cfg.nsrc > 0) rhs_src(o, it, istage, rhs);
  if (ISO_LIKE(o)) rhs_srcdd(o, it, istage, rhs);
}

/* peak ground motion on the free surface, PG_calcu (wav_t.c:380-455):
 * 15 maps PGV,PGVh,PGVx,PGVy,PGVz,PGA,PGAh,PGAx..z,PGD,PGDh,PGDx..z and the
 * trapezoid displacement accumulator, from w_end and the previous step */
static void pg_calcu(orc_t *o, const real_t *w1, const real_t *w0, float dt)
{
  size_t V = o->vol, S = o->ss;
  const real_t *Vx1 = w1, *Vy1 = w1 + V, *Vz1 = w1 + 2 * V;
  const real_t *Vx0 = w0, *Vy0 = w0 + V, *Vz0 = w0 + 2 * V;
  real_t *PG = o->pg, *Dx = o->dis, *Dy = o->dis + S, *Dz = o->dis + 2 * S;
  for (int j = o->nj1; j <= o->nj2; j++)
    for (int i = o->ni1; i <= o->ni2; i++) {
      size_t p = i + j * o->sl + (size_t)o->nk2 * S, q = i + j * o->sl;
      real_t Vv, Vh, D, Dh, A, Ah, Ax, Ay, Az;
      Ax = fabs((Vx1[p] - Vx0[p]) / dt);
      Ay = fabs((Vy1[p] - Vy0[p]) / dt);
      Az = fabs((Vz1[p] - Vz0[p]) / dt);
      Dx[q] += 0.5 * (Vx1[p] + Vx0[p]) * dt;
      Dy[q] += 0.5 * (Vy1[p] + Vy0[p]) * dt;
      Dz[q] += 0.5 * (Vz1[p] + Vz0[p]) * dt;
      Vv = sqrt(Vx1[p] * Vx1[p] + Vy1[p] * Vy1[p] + Vz1[p] * Vz1[p]);
      Vh = sqrt(Vx1[p] * Vx1[p] + Vy1[p] * Vy1[p]);
      A = sqrt(Ax * Ax + Ay * Ay + Az * Az);
      Ah = sqrt(Ax * Ax + Ay * Ay);
      D = sqrt(Dx[q] * Dx[q] + Dy[q] * Dy[q] + Dz[q] * Dz[q]);
      Dh = sqrt(Dx[q] * Dx[q] + Dy[q] * Dy[q]);
      real_t *m = PG;
      if (m[2 * S + q] < fabs(Vx1[p])) m[2 * S + q] = fabs(Vx1[p]);
      if (m[3 * S + q] < fabs(Vy1[p])) m[3 * S + q] = fabs(Vy1[p]);
      if (m[4 * S + q] < fabs(Vz1[p])) m[4 * S + q] = fabs(Vz1[p]);
      if (m[7 * S + q] < Ax) m[7 * S + q] = Ax;
      if (m[8 * S + q] < Ay) m[8 * S + q] = Ay;
      if (m[9 * S + q] < Az) m[9 * S + q] = Az;
      if (m[12 * S + q] < fabs(Dx[q])) m[12 * S + q] = fabs(Dx[q]);
      if (m[13 * S + q] < fabs(Dy[q])) m[13 * S + q] = fabs(Dy[q]);
      if (m[14 * S + q] < fabs(Dz[q])) m[14 * S + q] = fabs(Dz[q]);
      if (m[0 * S + q] < Vv) m[0 * S + q] = Vv;
      if (m[1 * S + q] < Vh) m[1 * S + q] = Vh;
      if (m[5 * S + q] < A) m[5 * S + q] = A;
      if (m[6 * S + q] < Ah) m[6 * S + q] = Ah;
      if (m[10 * S + q] < D) m[10 * S + q] = D;
      if (m[11 * S + q] < Dh) m[11 * S + q] = Dh;
    }
}

void orc_get_pg(const orc_t *o, real_t *pg, real_t *dis)
{
  if (!o->pg) return;
  memcpy(pg, o->pg, 15 * o->ss * sizeof(real_t));
  memcpy(dis, o->dis, 3 * o->ss * sizeof(real_t));
}

void orc_run(orc_t *o, float dt, int it0, int nt, real_t *w, real_t *aux,
             int nrecv, const int *recv_idx, const float *recv_w,
             int nt_total, real_t *seismo)
{
  size_t L = (size_t)o->nw * o->vol, LA = o->aux_total;
  real_t *tmp = (real_t *)calloc(L, sizeof(real_t));
  real_t *rhs = (real_t *)calloc(L, sizeof(real_t));
  real_t *end = (real_t *)calloc(L, sizeof(real_t));
  real_t *atmp = (real_t *)calloc(LA + 1, sizeof(real_t));
  real_t *arhs = (real_t *)calloc(LA + 1, sizeof(real_t));
  real_t *aend = (real_t *)calloc(LA + 1, sizeof(real_t));
  real_t *pre = w, *apre = aux;
  for (int n = 0; n < nt; n++) {
    int it = it0 + n;
    int ipair = it % 8;
    for (int s = 0; s < 4; s++) {
      const real_t *cur = s == 0 ? pre : tmp;
      const real_t *acur = s == 0 ? apre : atmp;
      orc_onestage(o, it, ipair, s, cur, acur, rhs, arhs);
      real_t ca = s < 3 ? (real_t)((real_t)(float)rk_a_tab[s] * (real_t)dt) : 0;
      real_t cb = (real_t)((real_t)(float)rk_b_tab[s] * (real_t)dt);
#if ORC_REAL_IS_DOUBLE
      ca = s < 3 ? rk_a_tab[s] * (double)dt : 0;
      cb = rk_b_tab[s] * (double)dt;
#endif
      /* drv_rk_curv_col.c:291-446, whole level */
#pragma omp parallel for schedule(static)
      for (long p = 0; p < (long)L; p++) {
        if (s < 3) tmp[p] = pre[p] + ca * rhs[p];
        if (s == 0) end[p] = pre[p] + cb * rhs[p];
        else end[p] += cb * rhs[p];
      }
#pragma omp parallel for schedule(static)
      for (long p = 0; p < (long)LA; p++) {
        if (s < 3) atmp[p] = apre[p] + ca * arhs[p];
        if (s == 0) aend[p] = apre[p] + cb * arhs[p];
        else aend[p] += cb * arhs[p];
      }
      /* drv_rk_curv_col.c:440-445, after the last stage's combine */
      if (s == 3 && o->mt == ORC_MEDIUM_VISCO && o->cfg.free_top) vis_free(o, end);
    }
    /* receivers, io_funcs.c:1585-1623 (indx1d[5] twice, indx1d[4] unused) */
    for (int rcv = 0; rcv < nrecv; rcv++) {
      const int *ix = recv_idx + 8 * rcv;
      real_t Lx2 = recv_w[3 * rcv], Ly2 = recv_w[3 * rcv + 1], Lz2 = recv_w[3 * rcv + 2];
      real_t Lx1 = 1.0 - Lx2, Ly1 = 1.0 - Ly2, Lz1 = 1.0 - Lz2;
      for (int c = 0; c < o->nw; c++) {
        const real_t *v = end + c * o->vol;
        seismo[((size_t)rcv * o->nw + c) * nt_total + it] =
            v[ix[0]] * Lx1 * Ly1 * Lz1 + v[ix[1]] * Lx2 * Ly1 * Lz1
          + v[ix[2]] * Lx1 * Ly2 * Lz1 + v[ix[3]] * Lx2 * Ly2 * Lz1
          + v[ix[5]] * Lx1 * Ly1 * Lz2 + v[ix[5]] * Lx2 * Ly1 * Lz2
          + v[ix[6]] * Lx1 * Ly2 * Lz2 + v[ix[7]] * Lx2 * Ly2 * Lz2;
      }
    }
    if (o->pg) pg_calcu(o, end, pre, dt);   /* drv_rk_curv_col.c:492-495 */
    memcpy(pre, end, L * sizeof(real_t));
    if (LA) memcpy(apre, aend, LA * sizeof(real_t));
  }
  free(tmp); free(rhs); free(end); free(atmp); free(arhs); free(aend);
}
