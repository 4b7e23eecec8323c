// setup_kernels.h -- one-off device kernels that build the solver's static
// inputs: the curvilinear metric from grid coordinates and the free-surface
// velocity-gradient matrices.  Same float arithmetic as the reference's host
// code, so the results are bit-identical to it.
#pragma once

namespace cgfdh {

// centred 7-point operator of the metric, fd_t.c:161-171 (row 3)
#define CGFDH_C0 ((float)-0.02084)
#define CGFDH_C1 ((float)0.1667)
#define CGFDH_C2 ((float)-0.7709)
#define CGFDH_C3 ((float)0.0)
#define CGFDH_C4 ((float)0.7709)
#define CGFDH_C5 ((float)-0.1667)
#define CGFDH_C6 ((float)0.02084)

__device__ __forceinline__ float dc7(const float* __restrict__ v, long p, long s)
{
  float d = CGFDH_C0 * v[p - 3 * s];
  d += CGFDH_C1 * v[p - 2 * s];
  d += CGFDH_C2 * v[p - s];
  d += CGFDH_C3 * v[p];
  d += CGFDH_C4 * v[p + s];
  d += CGFDH_C5 * v[p + 2 * s];
  d += CGFDH_C6 * v[p + 3 * s];
  return d;
}

// gd_curv_metric_cal, physical points (gd_t.c:191-289)
__global__ void metric_kernel(const float* __restrict__ xyz, float* __restrict__ met, int ni,
                              int nj, int nk, int nx, long ss, long vol)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  const int k = 3 + blockIdx.z;
  if (i > 3 + ni - 1) return;
  const long p = (long)i + (long)j * nx + (long)k * ss;
  const float* X = xyz;
  const float* Y = xyz + vol;
  const float* Z = xyz + 2 * vol;
  float v1[3] = {dc7(X, p, 1), dc7(Y, p, 1), dc7(Z, p, 1)};
  float v2[3] = {dc7(X, p, nx), dc7(Y, p, nx), dc7(Z, p, nx)};
  float v3[3] = {dc7(X, p, ss), dc7(Y, p, ss), dc7(Z, p, ss)};
  float g[3];
  g[0] = v1[1] * v2[2] - v1[2] * v2[1];
  g[1] = v1[2] * v2[0] - v1[0] * v2[2];
  g[2] = v1[0] * v2[1] - v1[1] * v2[0];
  float jac = 0.0f;
  jac += g[0] * v3[0];
  jac += g[1] * v3[1];
  jac += g[2] * v3[2];
  met[p] = jac;
  g[0] = v2[1] * v3[2] - v2[2] * v3[1];
  g[1] = v2[2] * v3[0] - v2[0] * v3[2];
  g[2] = v2[0] * v3[1] - v2[1] * v3[0];
  met[1 * vol + p] = g[0] / jac;
  met[2 * vol + p] = g[1] / jac;
  met[3 * vol + p] = g[2] / jac;
  g[0] = v3[1] * v1[2] - v3[2] * v1[1];
  g[1] = v3[2] * v1[0] - v3[0] * v1[2];
  g[2] = v3[0] * v1[1] - v3[1] * v1[0];
  met[4 * vol + p] = g[0] / jac;
  met[5 * vol + p] = g[1] / jac;
  met[6 * vol + p] = g[2] / jac;
  g[0] = v1[1] * v2[2] - v1[2] * v2[1];
  g[1] = v1[2] * v2[0] - v1[0] * v2[2];
  g[2] = v1[0] * v2[1] - v1[1] * v2[0];
  met[7 * vol + p] = g[0] / jac;
  met[8 * vol + p] = g[1] / jac;
  met[9 * vol + p] = g[2] / jac;
}

// ghost mirroring of the metric (gd_t.c:291-401): pass 0 mirrors x ghosts
// for every (j,k), pass 1 y ghosts for every (i,k), pass 2 z ghosts for every
// (i,j); run in that order as the reference does
__global__ void metric_mirror_kernel(float* __restrict__ met, int pass, int nx, int ny, int nz,
                                     int ni1, int ni2, int nj1, int nj2, int nk1, int nk2,
                                     long ss, long vol)
{
  const int a = blockIdx.x * blockDim.x + threadIdx.x;   // fastest free index
  const int b = blockIdx.y;                              // slower free index
  int g = blockIdx.z;                                    // ghost slot 0..5
  long dst, src;
  if (pass == 0) {
    if (a >= ny) return;
    const int i = g < 3 ? g : ni2 + 1 + (g - 3);
    const int is = g < 3 ? 2 * ni1 - i - 1 : 2 * ni2 - i + 1;
    dst = (long)i + (long)a * nx + (long)b * ss;
    src = (long)is + (long)a * nx + (long)b * ss;
  } else if (pass == 1) {
    if (a >= nx) return;
    const int j = g < 3 ? g : nj2 + 1 + (g - 3);
    const int js = g < 3 ? 2 * nj1 - j - 1 : 2 * nj2 - j + 1;
    dst = (long)a + (long)j * nx + (long)b * ss;
    src = (long)a + (long)js * nx + (long)b * ss;
  } else {
    if (a >= nx) return;
    const int k = g < 3 ? g : nk2 + 1 + (g - 3);
    const int ks = g < 3 ? 2 * nk1 - k - 1 : 2 * nk2 - k + 1;
    dst = (long)a + (long)b * nx + (long)k * ss;
    src = (long)a + (long)b * nx + (long)ks * ss;
  }
  for (int m = 0; m < 10; m++) met[m * vol + dst] = met[m * vol + src];
}

__device__ __forceinline__ void inv3(float m[3][3])
{
  float inv[3][3];
  inv[0][0] = m[1][1] * m[2][2] - m[2][1] * m[1][2];
  inv[0][1] = m[2][1] * m[0][2] - m[0][1] * m[2][2];
  inv[0][2] = m[0][1] * m[1][2] - m[0][2] * m[1][1];
  inv[1][0] = m[1][2] * m[2][0] - m[1][0] * m[2][2];
  inv[1][1] = m[0][0] * m[2][2] - m[2][0] * m[0][2];
  inv[1][2] = m[1][0] * m[0][2] - m[0][0] * m[1][2];
  inv[2][0] = m[1][0] * m[2][1] - m[1][1] * m[2][0];
  inv[2][1] = m[2][0] * m[0][1] - m[0][0] * m[2][1];
  inv[2][2] = m[0][0] * m[1][1] - m[0][1] * m[1][0];
  float det = inv[0][0] * m[0][0] + inv[0][1] * m[1][0] + inv[0][2] * m[2][0];
  det = 1.0f / det;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) m[i][j] = inv[i][j] * det;
}

__device__ __forceinline__ void mul3(const float A[3][3], const float B[3][3], float C[3][3])
{
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) {
      float c = 0.0f;
      for (int k = 0; k < 3; k++) c += A[i][k] * B[k][j];
      C[i][j] = c;
    }
}

// sv_curv_col_el_iso_dvh2dvz (sv_curv_col_el_iso.c:1259-1375) at k = nk2
// (the visco-elastic version, sv_curv_col_vis_iso.c:354-506, computes the same
// two matrices and leaves matF2Vz2 unset: write_f = 0)
__global__ void dvh2dvz_iso_kernel(const float* __restrict__ met, const float* __restrict__ lam3,
                                   const float* __restrict__ mu3, float* __restrict__ mats,
                                   int ni, int nj, int nx, int nk2, long ss, long vol,
                                   int write_f)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  if (i > 3 + ni - 1) return;
  const long p = (long)i + (long)j * nx + (long)nk2 * ss;
  const float e11 = met[1 * vol + p], e12 = met[2 * vol + p], e13 = met[3 * vol + p];
  const float e21 = met[4 * vol + p], e22 = met[5 * vol + p], e23 = met[6 * vol + p];
  const float e31 = met[7 * vol + p], e32 = met[8 * vol + p], e33 = met[9 * vol + p];
  const float lam = lam3[p], mu = mu3[p];
  const float lam2mu = lam + 2.0f * mu;
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
  const long ij = ((long)j * nx + i) * 9;
  const long n = 9 * ss;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      mats[ij + r * 3 + c] = AB[r][c];
      mats[n + ij + r * 3 + c] = AC[r][c];
      if (write_f) mats[2 * n + ij + r * 3 + c] = A[r][c];
    }
}

// visco-elastic surface rotation matrix matD (sv_curv_col_vis_iso.c:415-501):
// rows = the surface's tangent pair and normal, from the eta derivatives of
// the coordinates (centred 7-point operator) and the zeta metric
__global__ void vis_matd_kernel(const float* __restrict__ xyz, const float* __restrict__ met,
                                float* __restrict__ matD, int ni, int nj, int nx, int nk2,
                                long ss, long vol)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  if (i > 3 + ni - 1) return;
  const long p = (long)i + (long)j * nx + (long)nk2 * ss;
  const float xet = dc7(xyz, p, nx), yet = dc7(xyz + vol, p, nx), zet = dc7(xyz + 2 * vol, p, nx);
  const float ztx = met[7 * vol + p], zty = met[8 * vol + p], ztz = met[9 * vol + p];
  const float e_n = (float)(1.0 / sqrt((double)(xet * xet + yet * yet + zet * zet)));
  const float e_m = (float)(1.0 / sqrt((double)(ztx * ztx + zty * zty + ztz * ztz)));
  const float e_nm = e_n * e_m;
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
  const long ij = ((long)j * nx + i) * 9;
  for (int q = 0; q < 9; q++) matD[ij + q] = D[q];
}

// VTI / general anisotropic traction-free matrices at k = nk2
// (sv_curv_col_el_vti.c:1085-1203, sv_curv_col_el_aniso.c:1267-1417).
// K[t][b] is the coefficient of dV_b in stress row t (medium.h); rows of the
// traction on the zeta surface: (Txx,Txy,Txz), (Txy,Tyy,Tyz), (Txz,Tyz,Tzz)
// against zeta_x, zeta_y, zeta_z.  A sums K_zeta[row][b] * zeta_m, B and C
// the same with K_xi and K_eta; stored as -A^-1 B and -A^-1 C.  VTI has no
// term where kVtiHas is false; the third matrix is not written.
__device__ constexpr bool kVtiHas[6][3] = {{1, 1, 1}, {1, 1, 1}, {1, 1, 1},
                                           {0, 1, 1}, {1, 0, 1}, {1, 1, 0}};
// VTI term c * e: index into {c11, c12, c13, c33, c55, c66} and into (e1,e2,e3)
__device__ constexpr int kVtiC[6][3] = {{0, 1, 2}, {1, 0, 2}, {2, 2, 3},
                                        {0, 4, 4}, {4, 0, 4}, {5, 5, 0}};
__device__ constexpr int kVtiE[6][3] = {{0, 1, 2}, {0, 1, 2}, {0, 1, 2},
                                        {0, 2, 1}, {2, 0, 0}, {1, 0, 0}};
__device__ constexpr int kTracRow[3][3] = {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}};

template <int MT>
__device__ __forceinline__ float coupling_k(const float* cf, int t, int b, const float* e)
{
  if (MT == MEDIUM_VTI) return cf[kVtiC[t][b]] * e[kVtiE[t][b]];
  return cf[kAnisoIx[t][b][0]] * e[0] + cf[kAnisoIx[t][b][1]] * e[1] +
         cf[kAnisoIx[t][b][2]] * e[2];
}

template <int MT>
__global__ void dvh2dvz_aniso_kernel(const float* __restrict__ met, const float* __restrict__ med,
                                     float* __restrict__ mats, int ni, int nj, int nx, int nk2,
                                     long ss, long vol)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  if (i > 3 + ni - 1) return;
  const long p = (long)i + (long)j * nx + (long)nk2 * ss;
  float xi[3], et[3], zt[3];
#pragma unroll
  for (int m = 0; m < 3; m++) {
    xi[m] = met[(1 + m) * vol + p];
    et[m] = met[(4 + m) * vol + p];
    zt[m] = met[(7 + m) * vol + p];
  }
  float cf[21];
  if (MT == MEDIUM_VTI) {
    const float c11 = med[p], c66 = med[4 * vol + p];
    cf[0] = c11;
    cf[1] = (float)((double)c11 - 2.0 * (double)c66);
    cf[2] = med[vol + p];
    cf[3] = med[2 * vol + p];
    cf[4] = med[3 * vol + p];
    cf[5] = c66;
  } else {
#pragma unroll
    for (int q = 0; q < 21; q++) cf[q] = med[q * vol + p];
  }
  float A[3][3], B[3][3], C[3][3], AB[3][3], AC[3][3];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int b = 0; b < 3; b++) {
      bool first = true;
#pragma unroll
      for (int m = 0; m < 3; m++) {
        const int t = kTracRow[r][m];
        if (MT == MEDIUM_VTI && !kVtiHas[t][b]) continue;
        const float a = coupling_k<MT>(cf, t, b, zt) * zt[m];
        const float bb = coupling_k<MT>(cf, t, b, xi) * zt[m];
        const float cc = coupling_k<MT>(cf, t, b, et) * zt[m];
        if (first) { A[r][b] = a; B[r][b] = bb; C[r][b] = cc; }
        else { A[r][b] += a; B[r][b] += bb; C[r][b] += cc; }
        first = false;
      }
    }
  inv3(A);
  mul3(A, B, AB);
  mul3(A, C, AC);
  const long ij = ((long)j * nx + i) * 9;
  const long n = 9 * ss;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) {
      mats[ij + r * 3 + c] = -1.0f * AB[r][c];
      mats[n + ij + r * 3 + c] = -1.0f * AC[r][c];
    }
}

}  // namespace cgfdh
