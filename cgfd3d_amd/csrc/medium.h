// medium.h -- stress-velocity coupling of the three elastic media, one
// policy type per medium, used by every stage kernel.
//
//   Medium<2>  isotropic: lambda, mu          (sv_curv_col_el_iso.c)
//   Medium<3>  VTI: c11, c13, c33, c55, c66   (sv_curv_col_el_vti.c)
//   Medium<4>  general anisotropic: 21 Cij    (sv_curv_col_el_aniso.c)
//   Medium<5>  visco-elastic GMB: the isotropic coupling with unrelaxed
//              lambda, mu; its memory variables are handled by atten()
//              in stage_common.h               (sv_curv_col_vis_iso.c)
//
// Each medium lives in HBM as volumes (ncoef coefficient volumes first, read
// by load(), then the slowness): lambda, mu, slw / c11, c13, c33, c55, c66,
// slw / c11, c12, ..., c66, slw / lambda, mu, slw, Ylam_0, Ymu_0, Ylam_1, ...
// Every expression keeps the reference's written order; the kernels are
// compiled without FMA contraction, so results match the reference's FP32
// evaluation bit for bit.
#pragma once
#include <hip/hip_runtime.h>

namespace cgfdh {

// Streaming accesses: values a stage reads or writes once per point (metric,
// medium, the RK levels other than the stencil's input, PML auxiliaries).
// Plain loads: marking them non-temporal measured slower (it also keeps them
// out of the last-level cache that serves part of the stencil's re-reads).
__device__ __forceinline__ float ld_stream(const float* p) { return *p; }
__device__ __forceinline__ void st_stream(float* p, float v) { *p = v; }

enum { MEDIUM_ISO = 2, MEDIUM_VTI = 3, MEDIUM_ANISO = 4, MEDIUM_VISCO = 5 };

template <int MT> struct Medium;

// wavefield component slots (wav_t.c:93-136)
enum { VX = 0, VY, VZ, TXX, TYY, TZZ, TYZ, TXZ, TXY };

template <> struct Medium<MEDIUM_ISO> {
  static constexpr int ncoef = 2;
  float lam, mu, lam2mu;

  // from the coefficient volumes' values at one point (lambda, mu)
  __device__ __forceinline__ void from(const float* v)
  {
    lam = v[0];
    mu = v[1];
    lam2mu = (float)((double)lam + 2.0 * (double)mu);
  }

  __device__ __forceinline__ void load(const float* __restrict__ med, long V, long p)
  {
    const float v[2] = {ld_stream(med + p), ld_stream(med + V + p)};
    from(v);
  }

  // stress rates from the velocity derivatives along xi (X), eta (Y) and
  // zeta (Z), m = xi_x..zeta_z (sv_curv_col_el_iso.c:409-435)
  __device__ __forceinline__ void hooke(float* r, const float* m, const float* X,
                                        const float* Y, const float* Z) const
  {
    const float xix = m[0], xiy = m[1], xiz = m[2];
    const float etx = m[3], ety = m[4], etz = m[5];
    const float ztx = m[6], zty = m[7], ztz = m[8];
    r[TXX] = lam2mu * (xix * X[0] + etx * Y[0] + ztx * Z[0]) +
             lam * (xiy * X[1] + ety * Y[1] + zty * Z[1] + xiz * X[2] + etz * Y[2] + ztz * Z[2]);
    r[TYY] = lam2mu * (xiy * X[1] + ety * Y[1] + zty * Z[1]) +
             lam * (xix * X[0] + etx * Y[0] + ztx * Z[0] + xiz * X[2] + etz * Y[2] + ztz * Z[2]);
    r[TZZ] = lam2mu * (xiz * X[2] + etz * Y[2] + ztz * Z[2]) +
             lam * (xix * X[0] + etx * Y[0] + ztx * Z[0] + xiy * X[1] + ety * Y[1] + zty * Z[1]);
    r[TXY] = mu * (xiy * X[0] + xix * X[1] + ety * Y[0] + etx * Y[1] + zty * Z[0] + ztx * Z[1]);
    r[TXZ] = mu * (xiz * X[0] + xix * X[2] + etz * Y[0] + etx * Y[2] + ztz * Z[0] + ztx * Z[2]);
    r[TYZ] = mu * (xiz * X[1] + xiy * X[2] + etz * Y[1] + ety * Y[2] + ztz * Z[1] + zty * Z[2]);
  }

  // the one-direction stress rates of a PML face with metric (e1,e2,e3)
  // (sv_curv_col_el_iso.c:735-743)
  __device__ __forceinline__ void one_dir(float* r, float e1, float e2, float e3,
                                          const float* D) const
  {
    r[TXX] = lam2mu * e1 * D[0] + lam * e2 * D[1] + lam * e3 * D[2];
    r[TYY] = lam * e1 * D[0] + lam2mu * e2 * D[1] + lam * e3 * D[2];
    r[TZZ] = lam * e1 * D[0] + lam * e2 * D[1] + lam2mu * e3 * D[2];
    r[TXY] = mu * (e2 * D[0] + e1 * D[1]);
    r[TXZ] = mu * (e3 * D[0] + e1 * D[2]);
    r[TYZ] = mu * (e3 * D[1] + e2 * D[2]);
  }

  // zeta term of a PML face's free-surface coupling, Dz converted from the
  // face direction's derivatives (sv_curv_col_el_iso.c:800-806)
  __device__ __forceinline__ void surf(float* s, float ztx, float zty, float ztz,
                                       const float* Dz) const
  {
    s[TXX] = lam2mu * (ztx * Dz[0]) + lam * (zty * Dz[1] + ztz * Dz[2]);
    s[TYY] = lam2mu * (zty * Dz[1]) + lam * (ztx * Dz[0] + ztz * Dz[2]);
    s[TZZ] = lam2mu * (ztz * Dz[2]) + lam * (ztx * Dz[0] + zty * Dz[1]);
    s[TXY] = mu * (zty * Dz[0] + ztx * Dz[1]);
    s[TXZ] = mu * (ztz * Dz[0] + ztx * Dz[2]);
    s[TYZ] = mu * (ztz * Dz[1] + zty * Dz[2]);
  }
};

// VTI: every term is one coefficient product (c * e) times a derivative.
template <> struct Medium<MEDIUM_VTI> {
  static constexpr int ncoef = 5;
  float c11, c12, c13, c33, c55, c66;

  // from c11, c13, c33, c55, c66 at one point
  __device__ __forceinline__ void from(const float* v)
  {
    c11 = v[0];
    c13 = v[1];
    c33 = v[2];
    c55 = v[3];
    c66 = v[4];
    c12 = (float)((double)c11 - 2.0 * (double)c66);
  }

  __device__ __forceinline__ void load(const float* __restrict__ med, long V, long p)
  {
    float v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) v[q] = ld_stream(med + q * V + p);
    from(v);
  }

  // sv_curv_col_el_vti.c:453-475
  __device__ __forceinline__ void hooke(float* r, const float* m, const float* X,
                                        const float* Y, const float* Z) const
  {
    const float xix = m[0], xiy = m[1], xiz = m[2];
    const float etx = m[3], ety = m[4], etz = m[5];
    const float ztx = m[6], zty = m[7], ztz = m[8];
    r[TXX] = (c11 * xix) * X[0] + (c12 * xiy) * X[1] + (c13 * xiz) * X[2] +
             (c11 * etx) * Y[0] + (c12 * ety) * Y[1] + (c13 * etz) * Y[2] +
             (c11 * ztx) * Z[0] + (c12 * zty) * Z[1] + (c13 * ztz) * Z[2];
    r[TYY] = (c12 * xix) * X[0] + (c11 * xiy) * X[1] + (c13 * xiz) * X[2] +
             (c12 * etx) * Y[0] + (c11 * ety) * Y[1] + (c13 * etz) * Y[2] +
             (c12 * ztx) * Z[0] + (c11 * zty) * Z[1] + (c13 * ztz) * Z[2];
    r[TZZ] = (c13 * xix) * X[0] + (c13 * xiy) * X[1] + (c33 * xiz) * X[2] +
             (c13 * etx) * Y[0] + (c13 * ety) * Y[1] + (c33 * etz) * Y[2] +
             (c13 * ztx) * Z[0] + (c13 * zty) * Z[1] + (c33 * ztz) * Z[2];
    r[TYZ] = (c55 * xiz) * X[1] + (c55 * xiy) * X[2] + (c55 * etz) * Y[1] +
             (c55 * ety) * Y[2] + (c55 * ztz) * Z[1] + (c55 * zty) * Z[2];
    r[TXZ] = (c55 * xiz) * X[0] + (c55 * xix) * X[2] + (c55 * etz) * Y[0] +
             (c55 * etx) * Y[2] + (c55 * ztz) * Z[0] + (c55 * ztx) * Z[2];
    r[TXY] = (c66 * xiy) * X[0] + (c66 * xix) * X[1] + (c66 * ety) * Y[0] +
             (c66 * etx) * Y[1] + (c66 * zty) * Z[0] + (c66 * ztx) * Z[1];
  }

  // sv_curv_col_el_vti.c:760-765
  __device__ __forceinline__ void one_dir(float* r, float e1, float e2, float e3,
                                          const float* D) const
  {
    r[TXX] = (c11 * e1) * D[0] + (c12 * e2) * D[1] + (c13 * e3) * D[2];
    r[TYY] = (c12 * e1) * D[0] + (c11 * e2) * D[1] + (c13 * e3) * D[2];
    r[TZZ] = (c13 * e1) * D[0] + (c13 * e2) * D[1] + (c33 * e3) * D[2];
    r[TYZ] = (c55 * e3) * D[1] + (c55 * e2) * D[2];
    r[TXZ] = (c55 * e3) * D[0] + (c55 * e1) * D[2];
    r[TXY] = (c66 * e2) * D[0] + (c66 * e1) * D[1];
  }

  // sv_curv_col_el_vti.c:817-822: the same form with the zeta metric
  __device__ __forceinline__ void surf(float* s, float ztx, float zty, float ztz,
                                       const float* Dz) const
  {
    one_dir(s, ztx, zty, ztz, Dz);
  }
};

// General anisotropy: the coefficient of dV_b in stress row t (Voigt order
// Txx,Tyy,Tzz,Tyz,Txz,Txy) along a direction with metric (e1,e2,e3) is
// c[a0]*e1 + c[a1]*e2 + c[a2]*e3, a = kAnisoIx[t][b] into c11,c12,..,c66
// (sv_curv_col_el_aniso.c:468-491).
__device__ constexpr int kAnisoIx[6][3][3] = {
    {{0, 5, 4}, {5, 1, 3}, {4, 3, 2}},       {{1, 10, 9}, {10, 6, 8}, {9, 8, 7}},
    {{2, 14, 13}, {14, 7, 12}, {13, 12, 11}}, {{3, 17, 16}, {17, 8, 15}, {16, 15, 12}},
    {{4, 19, 18}, {19, 9, 16}, {18, 16, 13}}, {{5, 20, 19}, {20, 10, 17}, {19, 17, 14}}};

template <> struct Medium<MEDIUM_ANISO> {
  static constexpr int ncoef = 21;
  float c[21];

  __device__ __forceinline__ void from(const float* v)
  {
#pragma unroll
    for (int q = 0; q < 21; q++) c[q] = v[q];
  }

  __device__ __forceinline__ void load(const float* __restrict__ med, long V, long p)
  {
#pragma unroll
    for (int q = 0; q < 21; q++) c[q] = ld_stream(med + q * V + p);
  }

  __device__ __forceinline__ float K(int t, int b, float e1, float e2, float e3) const
  {
    return c[kAnisoIx[t][b][0]] * e1 + c[kAnisoIx[t][b][1]] * e2 + c[kAnisoIx[t][b][2]] * e3;
  }

  // sv_curv_col_el_aniso.c:468-491: one sum over xi, eta, zeta terms
  __device__ __forceinline__ void hooke(float* r, const float* m, const float* X,
                                        const float* Y, const float* Z) const
  {
#pragma unroll
    for (int t = 0; t < 6; t++) {
      float acc = K(t, 0, m[0], m[1], m[2]) * X[0];
      acc += K(t, 1, m[0], m[1], m[2]) * X[1];
      acc += K(t, 2, m[0], m[1], m[2]) * X[2];
      acc += K(t, 0, m[3], m[4], m[5]) * Y[0];
      acc += K(t, 1, m[3], m[4], m[5]) * Y[1];
      acc += K(t, 2, m[3], m[4], m[5]) * Y[2];
      acc += K(t, 0, m[6], m[7], m[8]) * Z[0];
      acc += K(t, 1, m[6], m[7], m[8]) * Z[1];
      acc += K(t, 2, m[6], m[7], m[8]) * Z[2];
      r[TXX + t] = acc;
    }
  }

  // sv_curv_col_el_aniso.c:912-917
  __device__ __forceinline__ void one_dir(float* r, float e1, float e2, float e3,
                                          const float* D) const
  {
#pragma unroll
    for (int t = 0; t < 6; t++)
      r[TXX + t] = K(t, 0, e1, e2, e3) * D[0] + K(t, 1, e1, e2, e3) * D[1] +
                   K(t, 2, e1, e2, e3) * D[2];
  }

  // sv_curv_col_el_aniso.c:969-974
  __device__ __forceinline__ void surf(float* s, float ztx, float zty, float ztz,
                                       const float* Dz) const
  {
    one_dir(s, ztx, zty, ztz, Dz);
  }
};

// visco-elastic GMB: the stress-velocity coupling is the isotropic one
template <> struct Medium<MEDIUM_VISCO> : Medium<MEDIUM_ISO> {};

}  // namespace cgfdh
