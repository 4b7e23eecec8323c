// stage_common.h -- device-side pieces shared by the stage kernels.
//
// Arithmetic follows the reference expression by expression and is compiled
// with -ffp-contract=off (no FMA contraction), so every value matches the
// reference's x86-64 FP32 evaluation.  Where the reference evaluates in
// double and rounds to float (lam + 2.0*mu, coef_B - 1.0, 1.0/jac, ...),
// the device code does the same.
#pragma once
#include <hip/hip_runtime.h>

#include "medium.h"

namespace cgfdh {

// DRP MacCormack operators as float, fd_t.c:109-112 (row 3);
// dir 0 = {-1..3} ("forward"), dir 1 = {-3..1}
#define CGFDH_F0 ((float)-0.30874)
#define CGFDH_F1 ((float)-0.6326)
#define CGFDH_F2 ((float)1.233)
#define CGFDH_F3 ((float)-0.3334)
#define CGFDH_F4 ((float)0.04168)

template <int DIR>
__device__ __forceinline__ float d5(const float* __restrict__ v, long p, long s)
{
  if (DIR == 0)
    return CGFDH_F0 * v[p - s] + CGFDH_F1 * v[p] + CGFDH_F2 * v[p + s] +
           CGFDH_F3 * v[p + 2 * s] + CGFDH_F4 * v[p + 3 * s];
  else
    return -CGFDH_F4 * v[p - 3 * s] + -CGFDH_F3 * v[p - 2 * s] + -CGFDH_F2 * v[p - s] +
           -CGFDH_F1 * v[p] + -CGFDH_F0 * v[p + s];
}

// the same operator applied to 5 values already gathered in window order
template <int DIR>
__device__ __forceinline__ float d5v(float a0, float a1, float a2, float a3, float a4)
{
  if (DIR == 0)
    return CGFDH_F0 * a0 + CGFDH_F1 * a1 + CGFDH_F2 * a2 + CGFDH_F3 * a3 + CGFDH_F4 * a4;
  else
    return -CGFDH_F4 * a0 + -CGFDH_F3 * a1 + -CGFDH_F2 * a2 + -CGFDH_F1 * a3 + -CGFDH_F0 * a4;
}

template <int DIR> struct Win {
  // offset of window entry 0 from the point
  static constexpr int lo = DIR == 0 ? -1 : -3;
};

struct PmlDev {
  int idim;                 // 0,1,2
  int i1, i2, j1, j2, k1, k2;
  int fnx, fny;
  long fn;                  // points per component
  const float* A;
  const float* B;
  const float* D;
  const float* aux_cur;     // 9*fn, component order of the wavefield
  const float* aux_pre;
  float* aux_tmp;           // stage output (may alias aux_cur)
  float* aux_end;
};

struct SrcDev {
  int si, sj, sk;
  int it_begin, it_end;
  int kb;                   // k extent below the point (src k2_ext_below_surface)
  int coef_off;             // offset of this source's weights in ext_coefs
};

// visco-elastic: at most this many Maxwell bodies, 6 memory variables each
constexpr int MAXWELL_MAX = 4;
constexpr int NW_MAX = 9 + 6 * MAXWELL_MAX;

struct StageParams {
  int ni1, ni2, nj1, nj2, nk1, nk2;
  int nx;
  long ss, vol;
  const float* met;         // 10 volumes
  const float* med;         // the medium's volumes (see medium.h)
  const float* slw;         // its last volume
  const float* wcur;        // 9 volumes each
  const float* wpre;
  float* wtmp;
  float* wend;
  float* rhs_out;           // KIND_RHS only
  float ca, cb;
  int free_top;
  const float* matVx;       // [ny][nx][3][3]
  const float* matVy;
  int npml;
  PmlDev pml[6];
  float* aux_rhs_out[6];    // KIND_RHS only, per enabled face
  int nsrc;
  const SrcDev* src;
  const float* mij;         // 6 arrays of mij_stride (NULL: no moment sources)
  const float* fxyz;        // 3 arrays of mij_stride (NULL: no force sources)
  long mij_stride;
  int src_max_nt;
  int src_spatial;          // 1 point, 2 gaussian
  int src_half;
  const float* src_coef;
  // strict surface force: forces on the surface row are left to the free
  // surface, which takes the listed sources' forces (fxyz) as surface
  // traction and their force rates (frate, 3 arrays of frate_stride) as a
  // velocity source
  int src_strict;
  int nsurf;
  const int* surf_idx;
  const float* frate;
  long frate_stride;
  const float* matF;        // matF2Vz2 [ny][nx][3][3]
  // dd sources at grid points: one bit per point marks them, dd_idx holds
  // their indices sorted (equal indices in input order) and dd_perm each
  // one's input position; step it reads row it % dd_cap of the value tables
  int ndd, dd_max_nt, dd_cap;
  const unsigned* dd_mask;
  const long* dd_idx;
  const int* dd_perm;
  const float* dd_vi;       // [dd_cap][4][ndd][3] (NULL: none)
  const float* dd_mij;      // [dd_cap][4][ndd][6] (NULL: none)
  int it, istage;
  const int* it_off;        // if set, a device step offset added to it (graph replays)
  // visco-elastic GMB: bodies, wavefield components (9 + 6 nmax), relaxation
  // frequencies, the Ylam_n / Ymu_n volumes (interleaved, from ylam)
  int nmax, nw;
  float wl[MAXWELL_MAX];
  const float* ylam;
  // tiling of the z-marching kernel
  int kchunk;
  int tiles_x, tiles_y;
};

enum { KIND_FIRST = 0, KIND_MID = 1, KIND_LAST = 2, KIND_RHS = 3 };

// momentum equation (sv_curv_col_el_iso.c:398-406); X,Y,Z hold the 9
// derivatives in component-slot order
__device__ __forceinline__ void momentum(float* r, const float* m, float slw,
                                         const float* X, const float* Y, const float* Z)
{
  const float xix = m[0], xiy = m[1], xiz = m[2];
  const float etx = m[3], ety = m[4], etz = m[5];
  const float ztx = m[6], zty = m[7], ztz = m[8];
  r[VX] = slw * (xix * X[TXX] + xiy * X[TXY] + xiz * X[TXZ] + etx * Y[TXX] + ety * Y[TXY] +
                 etz * Y[TXZ] + ztx * Z[TXX] + zty * Z[TXY] + ztz * Z[TXZ]);
  r[VY] = slw * (xix * X[TXY] + xiy * X[TYY] + xiz * X[TYZ] + etx * Y[TXY] + ety * Y[TYY] +
                 etz * Y[TYZ] + ztx * Z[TXY] + zty * Z[TYY] + ztz * Z[TYZ]);
  r[VZ] = slw * (xix * X[TXZ] + xiy * X[TYZ] + xiz * X[TZZ] + etx * Y[TXZ] + ety * Y[TYZ] +
                 etz * Y[TZZ] + ztx * Z[TXZ] + zty * Z[TYZ] + ztz * Z[TZZ]);
}

// conservative traction vector jac*(e1*T0 + e2*T1 + e3*T2) at q
__device__ __forceinline__ float tvec(const StageParams& P, long q, int e, const float* T0,
                                      const float* T1, const float* T2)
{
  const float* M = P.met;
  return M[q] * (M[e * P.vol + q] * T0[q] + M[(e + 1) * P.vol + q] * T1[q] +
                 M[(e + 2) * P.vol + q] * T2[q]);
}

// the surface-force slices at surface point (i, j) for the current step and
// stage, src_set_surface_layer_for_force (src_t.c:154-314): ts = TxSrc,
// TySrc, TzSrc (the listed sources' forces, times the Gaussian weight of
// the surface row) and vs = VxSrc, VySrc, VzSrc (their force rates times the
// weight, over the jacobian at the source's surface point), summed in list
// order from zero
__device__ __forceinline__ void surface_force(const StageParams& P, int i, int j, float* ts,
                                              float* vs)
{
  ts[0] = ts[1] = ts[2] = 0.0f;
  vs[0] = vs[1] = vs[2] = 0.0f;
  const int it = P.it + (P.it_off ? *P.it_off : 0);
  for (int n = 0; n < P.nsurf; n++) {
    const int is = P.surf_idx[n];
    const SrcDev S = P.src[is];
    if (it < S.it_begin || it > S.it_end) continue;
    const long itr = (long)(it - S.it_begin) * 4 + P.istage;
    const long ip = (long)is * P.src_max_nt * 4 + itr;
    const long ipr = (long)n * P.src_max_nt * 4 + itr;
    if (P.src_spatial == 1) {
      if (i != S.si || j != S.sj) continue;
      const float jac = P.met[(long)S.si + (long)S.sj * P.nx + (long)S.sk * P.ss];
#pragma unroll
      for (int m = 0; m < 3; m++) {
        ts[m] += P.fxyz[m * P.mij_stride + ip];
        vs[m] += P.frate[m * P.frate_stride + ipr] / jac;
      }
    } else {
      const int h = P.src_half;
      const int di = i - S.si, dj = j - S.sj;
      if (di < -h || di > h || dj < -h || dj > h) continue;
      const int L = 2 * h + 1;
      const int ke = P.nk2 - S.sk;
      const float coef = P.src_coef[S.coef_off + ((ke + h) * L + (dj + h)) * L + (di + h)];
      const float jac = P.met[(long)i + (long)j * P.nx + (long)P.nk2 * P.ss];
#pragma unroll
      for (int m = 0; m < 3; m++) {
        ts[m] += P.fxyz[m * P.mij_stride + ip] * coef;
        vs[m] += P.frate[m * P.frate_stride + ipr] * coef / jac;
      }
    }
  }
}

// traction-image hV on the top rows (sv_curv_col_el.c:31-305): the surface
// traction (zero, or the strict surface forces' TxSrc..TzSrc) at the
// surface row and mirrored above it.  Mirror values come from the grid
// points they stand for.
template <int DX, int DY, int DZ>
__device__ void timg_hv(const StageParams& P, int i, int j, int k, long p, float* r)
{
  const float* w = P.wcur;
  const long V = P.vol, sl = P.nx, ss = P.ss;
  const int zlo = Win<DZ>::lo;
  const int n_free = P.nk2 - k - zlo;
  const float slwjac = P.slw[p] / P.met[p];
  const int rows[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
  float ts[3] = {0.0f, 0.0f, 0.0f};
  if (P.nsurf > 0) {
    float vs[3];
    surface_force(P, i, j, ts, vs);
  }
#pragma unroll
  for (int comp = 0; comp < 3; comp++) {
    const float* T0 = w + rows[comp][0] * V;
    const float* T1 = w + rows[comp][1] * V;
    const float* T2 = w + rows[comp][2] * V;
    float vx[5], vy[5], vz[5];
#pragma unroll
    for (int n = 0; n < 5; n++) {
      vx[n] = tvec(P, p + (Win<DX>::lo + n), 1, T0, T1, T2);
      vy[n] = tvec(P, p + (long)(Win<DY>::lo + n) * sl, 4, T0, T1, T2);
    }
#pragma unroll
    for (int n = 0; n < 5; n++) {
      if (n < n_free) {
        vz[n] = tvec(P, p + (long)(zlo + n) * ss, 7, T0, T1, T2);
      } else if (n == n_free) {
        vz[n] = ts[comp];
      } else {
        const long off = (zlo + n) - 2 * (n - n_free);
        const float below = tvec(P, p + off * ss, 7, T0, T1, T2);
        vz[n] = (float)(2.0 * (double)ts[comp] - (double)below);
      }
    }
    const float DxT = d5v<DX>(vx[0], vx[1], vx[2], vx[3], vx[4]);
    const float DyT = d5v<DY>(vy[0], vy[1], vy[2], vy[3], vy[4]);
    const float DzT = d5v<DZ>(vz[0], vz[1], vz[2], vz[3], vz[4]);
    r[comp] = (DxT + DyT + DzT) * slwjac;
  }
}

// hT on the top 3 rows (sv_curv_col_el_iso.c:452-634 and the VTI / general
// anisotropic equivalents): matrices at the surface, the 2- and 3-point
// one-sided z operators below it.  The isotropic forms (elastic and
// visco-elastic) add matF2Vz2 times the strict surface forces' velocity
// source at the surface (sv_curv_col_el_iso.c:583-592); with no such force
// that term is zero and skipped.
template <int DZ, int MT>
__device__ void vlow_ht(const StageParams& P, int i, int j, int k, long p, const float* m,
                        const Medium<MT>& md, const float* X, const float* Y, float* r)
{
  float Z[3];
  const long V = P.vol, ss = P.ss;
  if (k == P.nk2) {
    const long ij = ((long)i + (long)j * P.nx) * 9;
#pragma unroll
    for (int q = 0; q < 3; q++)
      Z[q] = P.matVx[ij + 3 * q + 0] * X[0] + P.matVx[ij + 3 * q + 1] * X[1] +
             P.matVx[ij + 3 * q + 2] * X[2] + P.matVy[ij + 3 * q + 0] * Y[0] +
             P.matVy[ij + 3 * q + 1] * Y[1] + P.matVy[ij + 3 * q + 2] * Y[2];
    if ((MT == MEDIUM_ISO || MT == MEDIUM_VISCO) && P.nsurf > 0) {
      float ts[3], vs[3];
      surface_force(P, i, j, ts, vs);
#pragma unroll
      for (int q = 0; q < 3; q++)
        Z[q] += P.matF[ij + 3 * q + 0] * vs[0] + P.matF[ij + 3 * q + 1] * vs[1] +
                P.matF[ij + 3 * q + 2] * vs[2];
    }
  } else if (k == P.nk2 - 1) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float* v = P.wcur + c * V;
      float d;
      if (DZ == 0) { d = -1.0f * v[p]; d += 1.0f * v[p + ss]; }
      else { d = -1.0f * v[p - ss]; d += 1.0f * v[p]; }
      Z[c] = d;
    }
  } else {
    const float c0 = DZ == 0 ? (float)(-7.0 / 6.0) : (float)(1.0 / 6.0);
    const float c1 = DZ == 0 ? (float)(8.0 / 6.0) : (float)(-8.0 / 6.0);
    const float c2 = DZ == 0 ? (float)(-1.0 / 6.0) : (float)(7.0 / 6.0);
#pragma unroll
    for (int c = 0; c < 3; c++) {
      const float* v = P.wcur + c * V;
      float d;
      if (DZ == 0) { d = c0 * v[p]; d += c1 * v[p + ss]; d += c2 * v[p + 2 * ss]; }
      else { d = c0 * v[p - 2 * ss]; d += c1 * v[p - ss]; d += c2 * v[p]; }
      Z[c] = d;
    }
  }
  md.hooke(r, m, X, Y, Z);
}

// offset of point (i, j, k) in face F's memory-variable arrays
__device__ __forceinline__ long pml_aux_index(const PmlDev& F, int i, int j, int k)
{
  return (long)(i - F.i1) + (long)(j - F.j1) * F.fnx + (long)(k - F.k1) * F.fnx * F.fny;
}

// face F's memory variables at the point: the current level and the levels
// the RK combine of this stage kind reads
template <int KIND>
__device__ __forceinline__ void pml_aux_load(const PmlDev& F, long ia, float* a, float* apre,
                                             float* aend)
{
#pragma unroll
  for (int c = 0; c < 9; c++) {
    a[c] = ld_stream(F.aux_cur + c * F.fn + ia);
    if (KIND == KIND_FIRST || KIND == KIND_MID) apre[c] = ld_stream(F.aux_pre + c * F.fn + ia);
    if (KIND == KIND_MID || KIND == KIND_LAST) aend[c] = ld_stream(F.aux_end + c * F.fn + ia);
  }
}

// CFS-PML correction of one face (sv_curv_col_el_iso.c:645-1146) and the
// RK combine of its auxiliary variables, from memory variables already
// loaded; IDIM the face normal (0 x, 1 y, 2 z) as a template argument, so
// that no register array is indexed by it
template <int KIND, int MT, int IDIM>
__device__ void pml_face_apply(const StageParams& P, const PmlDev& F, int fslot, int i, int j,
                               int k, long ia, const float* Dd, const Medium<MT>& md, float slw,
                               const float* m, float* r, const float* a, const float* apre,
                               const float* aend)
{
  // m: the point's metric xi_x..zeta_z, already loaded by the caller
  const int ic = IDIM == 0 ? i - F.i1 : (IDIM == 1 ? j - F.j1 : k - F.k1);
  const float cD = F.D[ic], cA = F.A[ic], cB = F.B[ic];
  const float cBm1 = (float)((double)cB - 1.0);
  const float e1 = m[3 * IDIM], e2 = m[3 * IDIM + 1], e3 = m[3 * IDIM + 2];
  float rr[9];
  rr[VX] = slw * (e1 * Dd[TXX] + e2 * Dd[TXY] + e3 * Dd[TXZ]);
  rr[VY] = slw * (e1 * Dd[TXY] + e2 * Dd[TYY] + e3 * Dd[TYZ]);
  rr[VZ] = slw * (e1 * Dd[TXZ] + e2 * Dd[TYZ] + e3 * Dd[TZZ]);
  md.one_dir(rr, e1, e2, e3, Dd);
  float ah[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    r[c] += cBm1 * rr[c] - cB * a[c];
    ah[c] = cD * rr[c] - cA * a[c];
  }
  if (IDIM < 2 && P.free_top && k == P.nk2) {
    const long ij = ((long)i + (long)j * P.nx) * 9;
    const float* mm = IDIM == 0 ? P.matVx : P.matVy;
    float Dz[3];
#pragma unroll
    for (int q = 0; q < 3; q++)
      Dz[q] = mm[ij + 3 * q + 0] * Dd[VX] + mm[ij + 3 * q + 1] * Dd[VY] +
              mm[ij + 3 * q + 2] * Dd[VZ];
    const float ztx = m[6], zty = m[7], ztz = m[8];
    float s[9];
    md.surf(s, ztx, zty, ztz, Dz);
#pragma unroll
    for (int c = TXX; c <= TXY; c++) {
      r[c] = (float)((double)r[c] + ((double)cB - 1.0) * (double)s[c]);
      ah[c] += cD * s[c];
    }
  }
  if (KIND == KIND_RHS) {
#pragma unroll
    for (int c = 0; c < 9; c++) P.aux_rhs_out[fslot][c * F.fn + ia] = ah[c];
  } else if (KIND == KIND_FIRST) {
#pragma unroll
    for (int c = 0; c < 9; c++) {
      st_stream(F.aux_tmp + c * F.fn + ia, apre[c] + P.ca * ah[c]);
      st_stream(F.aux_end + c * F.fn + ia, apre[c] + P.cb * ah[c]);
    }
  } else if (KIND == KIND_MID) {
#pragma unroll
    for (int c = 0; c < 9; c++) {
      st_stream(F.aux_tmp + c * F.fn + ia, apre[c] + P.ca * ah[c]);
      st_stream(F.aux_end + c * F.fn + ia, aend[c] + P.cb * ah[c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < 9; c++) st_stream(F.aux_end + c * F.fn + ia, aend[c] + P.cb * ah[c]);
  }
}

// the same with the memory variables loaded here, ahead of the arithmetic
template <int KIND, int MT, int IDIM>
__device__ void pml_face_d(const StageParams& P, const PmlDev& F, int fslot, int i, int j,
                           int k, long p, const float* Dd, const Medium<MT>& md, float slw,
                           const float* m, float* r)
{
  const long ia = pml_aux_index(F, i, j, k);
  float a[9], apre[9], aend[9];
  pml_aux_load<KIND>(F, ia, a, apre, aend);
  pml_face_apply<KIND, MT, IDIM>(P, F, fslot, i, j, k, ia, Dd, md, slw, m, r, a, apre, aend);
}

// pml_face_d for face F, with the derivatives along its normal (X, Y or Z)
template <int KIND, int MT>
__device__ __forceinline__ void pml_face(const StageParams& P, const PmlDev& F, int fslot, int i,
                                         int j, int k, long p, const float* X, const float* Y,
                                         const float* Z, const Medium<MT>& md, float slw,
                                         const float* m, float* r)
{
  if (F.idim == 0)
    pml_face_d<KIND, MT, 0>(P, F, fslot, i, j, k, p, X, md, slw, m, r);
  else if (F.idim == 1)
    pml_face_d<KIND, MT, 1>(P, F, fslot, i, j, k, p, Y, md, slw, m, r);
  else
    pml_face_d<KIND, MT, 2>(P, F, fslot, i, j, k, p, Z, md, slw, m, r);
}

// moment and force sources (sv_curv_col_el.c:312-479): a moment subtracts
// M * coef / jac from the stress rates, a force adds f * (coef * slw / jac)
// to the velocity rates, except on the surface row when the surface force
// is strict (the free surface takes it there)
__device__ __forceinline__ void add_sources(const StageParams& P, int i, int j, int k, long p,
                                            float* r)
{
  const int it = P.it + (P.it_off ? *P.it_off : 0);
  for (int is = 0; is < P.nsrc; is++) {
    const SrcDev S = P.src[is];
    if (it < S.it_begin || it > S.it_end) continue;
    float coefv, forcev;
    bool hit;
    const bool force = P.fxyz != nullptr && (!P.src_strict || k < P.nk2);
    if (P.src_spatial == 1) {
      hit = (i == S.si && j == S.sj && k == S.sk);
      if (!hit) continue;
      coefv = (float)(1.0 / (double)P.met[p]);
      if (force) forcev = P.slw[p] / P.met[p];
    } else {
      const int h = P.src_half;
      const int di = i - S.si, dj = j - S.sj, dk = k - S.sk;
      hit = di >= -h && di <= h && dj >= -h && dj <= h && dk >= -h && dk <= S.kb;
      if (!hit) continue;
      const int L = 2 * h + 1;
      const float coef = P.src_coef[S.coef_off + ((dk + h) * L + (dj + h)) * L + (di + h)];
      coefv = coef / P.met[p];
      if (force) forcev = coef * P.slw[p] / P.met[p];
    }
    const long ip = (long)is * P.src_max_nt * 4 + (long)(it - S.it_begin) * 4 + P.istage;
    if (force) {
#pragma unroll
      for (int m = 0; m < 3; m++) r[VX + m] += P.fxyz[m * P.mij_stride + ip] * forcev;
    }
    if (P.mij != nullptr) {
      const int slot[6] = {TXX, TYY, TZZ, TXZ, TYZ, TXY};
#pragma unroll
      for (int m = 0; m < 6; m++) r[slot[m]] -= P.mij[m * P.mij_stride + ip] * coefv;
    }
  }
}

// dd sources at grid points (sv_curv_col_el.c:486-555): every source at
// this point, in input order, adds vi * (slw / jac) to the velocity rates
// and subtracts mij * (1 / jac) (Mxx,Myy,Mzz,Myz,Mxz,Mxy) from the stress
// rates; nothing from step dd_max_nt on
__device__ __forceinline__ void add_dd(const StageParams& P, long p, float* r)
{
  if (!((P.dd_mask[p >> 5] >> (p & 31)) & 1u)) return;
  const int it = P.it + (P.it_off ? *P.it_off : 0);
  if (it >= P.dd_max_nt) return;
  int lo = 0, hi = P.ndd;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (P.dd_idx[mid] < p) lo = mid + 1;
    else hi = mid;
  }
  const long row = ((long)(it % P.dd_cap) * 4 + P.istage) * P.ndd;
  const float V = P.slw[p] / P.met[p];
  const float rjac = (float)(1.0 / (double)P.met[p]);
  for (int q = lo; q < P.ndd && P.dd_idx[q] == p; q++) {
    const long e = row + P.dd_perm[q];
    if (P.dd_vi != nullptr) {
#pragma unroll
      for (int c = 0; c < 3; c++) r[VX + c] += P.dd_vi[3 * e + c] * V;
    }
    if (P.dd_mij != nullptr) {
      const int slot[6] = {TXX, TYY, TZZ, TYZ, TXZ, TXY};
#pragma unroll
      for (int c = 0; c < 6; c++) r[slot[c]] -= P.dd_mij[6 * e + c] * rjac;
    }
  }
}

// the reference adds dd sources in the isotropic and GMB right-hand sides
// only (sv_curv_col_el_iso.c:193-199, sv_curv_col_vis_iso.c:235-241)
template <int MT>
__device__ __forceinline__ void add_dd_for(const StageParams& P, long p, float* r)
{
  if constexpr (MT == MEDIUM_ISO || MT == MEDIUM_VISCO)
    if (P.ndd > 0) add_dd(P, p, r);
}

// RK combine of the wavefield (drv_rk_curv_col.c:291-446); NW > 9 when the
// visco-elastic memory variables follow (P.nw of them are live)
template <int KIND, int NW = 9>
__device__ __forceinline__ void combine(const StageParams& P, long p, const float* r)
{
  const long V = P.vol;
  if (KIND == KIND_RHS) {
#pragma unroll
    for (int c = 0; c < NW; c++)
      if (c < 9 || c < P.nw) P.rhs_out[c * V + p] = r[c];
  } else if (KIND == KIND_FIRST) {
#pragma unroll
    for (int c = 0; c < NW; c++) {
      if (c >= 9 && c >= P.nw) break;
      const float pre = ld_stream(P.wpre + c * V + p);
      st_stream(P.wtmp + c * V + p, pre + P.ca * r[c]);
      st_stream(P.wend + c * V + p, pre + P.cb * r[c]);
    }
  } else if (KIND == KIND_MID) {
#pragma unroll
    for (int c = 0; c < NW; c++) {
      if (c >= 9 && c >= P.nw) break;
      const float pre = ld_stream(P.wpre + c * V + p);
      const float end = ld_stream(P.wend + c * V + p);
      st_stream(P.wtmp + c * V + p, pre + P.ca * r[c]);
      st_stream(P.wend + c * V + p, end + P.cb * r[c]);
    }
  } else {
#pragma unroll
    for (int c = 0; c < NW; c++) {
      if (c >= 9 && c >= P.nw) break;
      const float end = ld_stream(P.wend + c * V + p);
      st_stream(P.wend + c * V + p, end + P.cb * r[c]);
    }
  }
}

// combine() of the 9 elastic fields split in two: the loads of the levels
// the stage kind reads, and the stores with the same arithmetic
template <int KIND>
__device__ __forceinline__ void combine_load(const StageParams& P, long p, float* pre, float* end)
{
  const long V = P.vol;
#pragma unroll
  for (int c = 0; c < 9; c++) {
    if (KIND == KIND_FIRST || KIND == KIND_MID) pre[c] = ld_stream(P.wpre + c * V + p);
    if (KIND == KIND_MID || KIND == KIND_LAST) end[c] = ld_stream(P.wend + c * V + p);
  }
}

template <int KIND>
__device__ __forceinline__ void combine_store(const StageParams& P, long p, const float* r,
                                              const float* pre, const float* end)
{
  const long V = P.vol;
#pragma unroll
  for (int c = 0; c < 9; c++) {
    if (KIND == KIND_RHS) {
      P.rhs_out[c * V + p] = r[c];
    } else if (KIND == KIND_FIRST) {
      st_stream(P.wtmp + c * V + p, pre[c] + P.ca * r[c]);
      st_stream(P.wend + c * V + p, pre[c] + P.cb * r[c]);
    } else if (KIND == KIND_MID) {
      st_stream(P.wtmp + c * V + p, pre[c] + P.ca * r[c]);
      st_stream(P.wend + c * V + p, end[c] + P.cb * r[c]);
    } else {
      st_stream(P.wend + c * V + p, end[c] + P.cb * r[c]);
    }
  }
}

// GMB attenuation after the PML terms (sv_curv_col_vis_iso.c:251-342): the
// memory-variable rates into r[9 + 6n + {xx,yy,zz,yz,xz,xy}] and the memory
// terms subtracted from the stress rates, with the reference's double
// intermediates
__device__ __forceinline__ void atten(const StageParams& P, long p, float lam, float mu, float* r)
{
  const long V = P.vol;
  const float hTxx = r[TXX], hTyy = r[TYY], hTzz = r[TZZ];
  const float hTxy = r[TXY], hTxz = r[TXZ], hTyz = r[TYZ];
  const float sum_hxyz = (hTxx + hTyy + hTzz) / (3 * lam + 2 * mu);
  const double mu2 = (double)(2 * mu);
  const float EVxx = (float)(((2.0 * hTxx - hTyy - hTzz) / mu2 + sum_hxyz) / 3);
  const float EVyy = (float)(((2.0 * hTyy - hTxx - hTzz) / mu2 + sum_hxyz) / 3);
  const float EVzz = (float)(((2.0 * hTzz - hTxx - hTyy) / mu2 + sum_hxyz) / 3);
  const float EVxy = (float)((double)(hTxy / mu) * 0.5);
  const float EVxz = (float)((double)(hTxz / mu) * 0.5);
  const float EVyz = (float)((double)(hTyz / mu) * 0.5);
  float sJxyz = 0.0f, sJxx = 0.0f, sJyy = 0.0f, sJzz = 0.0f, sJxy = 0.0f, sJxz = 0.0f,
        sJyz = 0.0f;
#pragma unroll
  for (int n = 0; n < MAXWELL_MAX; n++) {
    if (n >= P.nmax) break;
    const float* J = P.wcur + (9 + 6 * n) * V + p;
    const float Jxx = J[0], Jyy = J[V], Jzz = J[2 * V], Jyz = J[3 * V], Jxz = J[4 * V],
                Jxy = J[5 * V];
    const float wl = P.wl[n];
    float* h = r + 9 + 6 * n;
    h[0] = wl * (EVxx - Jxx);
    h[1] = wl * (EVyy - Jyy);
    h[2] = wl * (EVzz - Jzz);
    h[3] = wl * (EVyz - Jyz);
    h[4] = wl * (EVxz - Jxz);
    h[5] = wl * (EVxy - Jxy);
    const float Yl = ld_stream(P.ylam + (2 * n) * V + p);
    const float Ym = ld_stream(P.ylam + (2 * n + 1) * V + p);
    sJxyz += Yl * (Jxx + Jyy + Jzz);
    sJxx += Ym * Jxx;
    sJyy += Ym * Jyy;
    sJzz += Ym * Jzz;
    sJxy += Ym * Jxy;
    sJxz += Ym * Jxz;
    sJyz += Ym * Jyz;
  }
  r[TXX] = hTxx - (float)(lam * sJxyz + 2.0 * mu * sJxx);
  r[TYY] = hTyy - (float)(lam * sJxyz + 2.0 * mu * sJyy);
  r[TZZ] = hTzz - (float)(lam * sJxyz + 2.0 * mu * sJzz);
  r[TXY] = hTxy - (float)(2.0 * mu * sJxy);
  r[TXZ] = hTxz - (float)(2.0 * mu * sJxz);
  r[TYZ] = hTyz - (float)(2.0 * mu * sJyz);
}

// the interior form of everything after the derivatives: momentum, Hooke's
// law, sources and the RK combine, for points with no PML face and no
// free-surface row
template <int KIND, int MT>
__device__ __forceinline__ void interior_finish(const StageParams& P, int i, int j, int k,
                                                long p, const float* X, const float* Y,
                                                const float* Z)
{
  const long V = P.vol;
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = ld_stream(P.met + (1 + n) * V + p);
  Medium<MT> md;
  md.load(P.med, V, p);
  const float slw = ld_stream(P.slw + p);
  // the elastic media's RK combine inputs are loaded here, with the
  // coefficients, so their latency overlaps the arithmetic below instead of
  // following the source branch
  float cpre[9], cend[9];
  if constexpr (MT != MEDIUM_VISCO) combine_load<KIND>(P, p, cpre, cend);
  float r[MT == MEDIUM_VISCO ? NW_MAX : 9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  if constexpr (MT == MEDIUM_VISCO) atten(P, p, md.lam, md.mu, r);
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  if constexpr (MT != MEDIUM_VISCO)
    combine_store<KIND>(P, p, r, cpre, cend);
  else
    combine<KIND, NW_MAX>(P, p, r);
}

// which boundary treatments a kernel instance has to handle
enum { HAS_SURF = 1, HAS_PML = 2 };

// everything after the derivatives: inner formulas, surface overrides, PML,
// sources, RK combine.  X/Y/Z: inner-operator derivatives of the 9 fields.
// FLAGS prunes the surface and PML code from instances whose points never
// need them (the result is the same either way).
// finish_point with the point's metric m, medium md and slowness already
// loaded
template <int DX, int DY, int DZ, int KIND, int MT, int FLAGS = HAS_SURF | HAS_PML>
__device__ __forceinline__ void finish_point_loaded(const StageParams& P, int i, int j, int k,
                                                    long p, const float* X, const float* Y,
                                                    const float* Z, const float* m,
                                                    const Medium<MT>& md, float slw)
{
  // the elastic media's RK combine inputs are requested before the
  // arithmetic and the boundary branches
  float cpre[9], cend[9];
  if constexpr (MT != MEDIUM_VISCO) combine_load<KIND>(P, p, cpre, cend);
  float r[MT == MEDIUM_VISCO ? NW_MAX : 9];
  momentum(r, m, slw, X, Y, Z);
  if ((FLAGS & HAS_SURF) && P.free_top && k >= P.nk2 - (DZ == 0 ? 3 : 1)) {
    // hV: traction image on rows nk2-3..nk2 (forward z op) or nk2-1..nk2
    timg_hv<DX, DY, DZ>(P, i, j, k, p, r);
  }
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]};
  if ((FLAGS & HAS_SURF) && P.free_top && k >= P.nk2 - 2) {
    vlow_ht<DZ, MT>(P, i, j, k, p, m, md, Xv, Yv, r);
  } else {
    float Zv[3] = {Z[VX], Z[VY], Z[VZ]};
    md.hooke(r, m, Xv, Yv, Zv);
  }
  for (int f = 0; (FLAGS & HAS_PML) && f < P.npml; f++) {
    const PmlDev& F = P.pml[f];
    if (i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2) {
      pml_face<KIND, MT>(P, F, f, i, j, k, p, X, Y, Z, md, slw, m, r);
    }
  }
  if constexpr (MT == MEDIUM_VISCO) atten(P, p, md.lam, md.mu, r);
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  if constexpr (MT != MEDIUM_VISCO)
    combine_store<KIND>(P, p, r, cpre, cend);
  else
    combine<KIND, NW_MAX>(P, p, r);
}

template <int DX, int DY, int DZ, int KIND, int MT, int FLAGS = HAS_SURF | HAS_PML>
__device__ __forceinline__ void finish_point(const StageParams& P, int i, int j, int k, long p,
                                             const float* X, const float* Y, const float* Z)
{
  const long V = P.vol;
  float m[9];
#pragma unroll
  for (int q = 0; q < 9; q++) m[q] = ld_stream(P.met + (1 + q) * V + p);
  Medium<MT> md;
  md.load(P.med, V, p);
  const float slw = ld_stream(P.slw + p);
  finish_point_loaded<DX, DY, DZ, KIND, MT, FLAGS>(P, i, j, k, p, X, Y, Z, m, md, slw);
}

}  // namespace cgfdh
