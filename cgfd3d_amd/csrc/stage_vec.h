// stage_vec.h -- the x-vectorised interior stage kernel.
//
// One thread owns NQ (2 or 4) consecutive x points, so every stencil row,
// every metric / medium / RK-level access is one 8- or 16-byte load or store
// instead of NQ 4-byte ones: the kernel issues 1/NQ of the vector memory
// instructions of the one-point-per-thread form, whose address units (TA)
// otherwise saturate before HBM does.  The x window of the NQ points is
// NQ + 4 values (two loads), y and z windows five loads each.  Arithmetic per point follows the
// reference's expressions and order, so results are bit-identical to the
// other kernels.
#pragma once

namespace cgfdh {

// NQ consecutive floats as one load / store (rows are 4-byte aligned only)
template <int NQ> struct Vec;
template <> struct Vec<2> {
  typedef float T __attribute__((ext_vector_type(2)));
  typedef T U __attribute__((aligned(4)));
};
template <> struct Vec<4> {
  typedef float T __attribute__((ext_vector_type(4)));
  typedef T U __attribute__((aligned(4)));
};

template <int NQ>
__device__ __forceinline__ void ldq(const float* p, float* o)
{
  const typename Vec<NQ>::T t = *(const typename Vec<NQ>::U*)p;
#pragma unroll
  for (int q = 0; q < NQ; q++) o[q] = t[q];
}
template <int NQ>
__device__ __forceinline__ void ldq_stream(const float* p, float* o)
{
  ldq<NQ>(p, o);
}
template <int NQ>
__device__ __forceinline__ void stq_stream(float* p, const float* v)
{
  typename Vec<NQ>::T t;
#pragma unroll
  for (int q = 0; q < NQ; q++) t[q] = v[q];
  *(typename Vec<NQ>::U*)p = t;
}

// keeps the scheduler from hoisting the next phase's loads above this point,
// which bounds the registers live at once
#define CGFDH_PHASE() __builtin_amdgcn_sched_barrier(0)

// derivatives along x of one field at the NQ points starting at v: the
// window is NQ + 4 values
template <int DX, int NQ>
__device__ __forceinline__ void dxq(const float* v, float* d)
{
  constexpr int lo = Win<DX>::lo;
  float w[NQ + 4];
  ldq<4>(v + lo, w);
  ldq<NQ>(v + lo + 4, w + 4);
#pragma unroll
  for (int q = 0; q < NQ; q++) d[q] = d5v<DX>(w[q], w[q + 1], w[q + 2], w[q + 3], w[q + 4]);
}

// derivatives along a strided direction (y: s = nx, z: s = nx*ny): one
// NQ-wide load per window row
template <int DD, int NQ>
__device__ __forceinline__ void dsq(const float* v, long s, float* d)
{
  constexpr int lo = Win<DD>::lo;
  float w[5][NQ];
#pragma unroll
  for (int n = 0; n < 5; n++) ldq<NQ>(v + (long)(lo + n) * s, w[n]);
#pragma unroll
  for (int q = 0; q < NQ; q++) d[q] = d5v<DD>(w[0][q], w[1][q], w[2][q], w[3][q], w[4][q]);
}

// RK combine of component c for the NQ points (drv_rk_curv_col.c:291-446);
// r[q] is the rate at point q; the first stage's "pre" level is the stencil's
// input level
template <int KIND, int NQ>
__device__ __forceinline__ void combine_q(const StageParams& P, const float* wcur, long p, int c,
                                          const float* r)
{
  const long V = P.vol;
  float o1[NQ], o2[NQ];
  if (KIND == KIND_RHS) {
    stq_stream<NQ>(P.rhs_out + c * V + p, r);
  } else if (KIND == KIND_FIRST) {
    float pr[NQ];
    ldq<NQ>(wcur + c * V, pr);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      o1[q] = pr[q] + P.ca * r[q];
      o2[q] = pr[q] + P.cb * r[q];
    }
    stq_stream<NQ>(P.wtmp + c * V + p, o1);
    stq_stream<NQ>(P.wend + c * V + p, o2);
  } else if (KIND == KIND_MID) {
    float pr[NQ], en[NQ];
    ldq_stream<NQ>(P.wpre + c * V + p, pr);
    ldq_stream<NQ>(P.wend + c * V + p, en);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      o1[q] = pr[q] + P.ca * r[q];
      o2[q] = en[q] + P.cb * r[q];
    }
    stq_stream<NQ>(P.wtmp + c * V + p, o1);
    stq_stream<NQ>(P.wend + c * V + p, o2);
  } else {
    float en[NQ];
    ldq_stream<NQ>(P.wend + c * V + p, en);
#pragma unroll
    for (int q = 0; q < NQ; q++) o2[q] = en[q] + P.cb * r[q];
    stq_stream<NQ>(P.wend + c * V + p, o2);
  }
}

// block: 64 x BY threads over a 64*NQ x BY tile of one plane, grouped block
// order (grouped_block); the box width is a multiple of NQ.  Phases: Hooke's
// law from the velocity derivatives (field by field) and the stress
// combine; then the momentum sums one direction at a time, folded in the
// reference's order (xi terms, eta terms, zeta terms), and the velocity
// combine.  The sources' moment terms join the first phase, their force
// terms the second.
template <int DX, int DY, int DZ, int KIND, int MT, int BY, int NQ>
__global__ void __launch_bounds__(64 * BY, NQ == 2 ? 4 : 2)
stage_interior_vec_kernel(StageParams P, BoxParams B)
{
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  const int i = B.i0 + bx * 64 * NQ + NQ * (int)threadIdx.x;
  const int j = B.j0 + by * BY + (int)threadIdx.y;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const long V = P.vol;
  const float* w = P.wcur + p;
  constexpr int NM = Medium<MT>::ncoef;
  float m4[9][NQ];
#pragma unroll
  for (int n = 0; n < 9; n++) ldq_stream<NQ>(P.met + (1 + n) * V + p, m4[n]);
  // ---- stresses: Hooke's law, sources, combine
  {
    float X[3][NQ], Y[3][NQ], Z[3][NQ];
#pragma unroll
    for (int c = 0; c < 3; c++) {
      dxq<DX, NQ>(w + c * V, X[c]);
      dsq<DY, NQ>(w + c * V, P.nx, Y[c]);
      dsq<DZ, NQ>(w + c * V, P.ss, Z[c]);
      CGFDH_PHASE();
    }
    float mv[NM][NQ];
#pragma unroll
    for (int n = 0; n < NM; n++) ldq_stream<NQ>(P.med + n * V + p, mv[n]);
    float rT[6][NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      float m[9], cv[NM], r[9];
#pragma unroll
      for (int n = 0; n < 9; n++) m[n] = m4[n][q];
#pragma unroll
      for (int n = 0; n < NM; n++) cv[n] = mv[n][q];
      Medium<MT> md;
      md.from(cv);
      const float Xq[3] = {X[0][q], X[1][q], X[2][q]};
      const float Yq[3] = {Y[0][q], Y[1][q], Y[2][q]};
      const float Zq[3] = {Z[0][q], Z[1][q], Z[2][q]};
      md.hooke(r, m, Xq, Yq, Zq);
      // the moment terms of the sources here, the force terms in the
      // velocity pass below
      r[VX] = r[VY] = r[VZ] = 0.0f;
      if (P.nsrc > 0) add_sources(P, i + q, j, k, p + q, r);
      add_dd_for<MT>(P, p + q, r);
#pragma unroll
      for (int t = 0; t < 6; t++) rT[t][q] = r[TXX + t];
    }
#pragma unroll
    for (int t = 0; t < 6; t++) combine_q<KIND, NQ>(P, w, p, TXX + t, rT[t]);
  }
  CGFDH_PHASE();
  // ---- velocities: momentum, one direction at a time
  {
    constexpr int rows[3][3] = {{TXX, TXY, TXZ}, {TXY, TYY, TYZ}, {TXZ, TYZ, TZZ}};
    float acc[3][NQ];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      float D[9][NQ];
#pragma unroll
      for (int c = TXX; c <= TXY; c++) {
        if (d == 0) dxq<DX, NQ>(w + c * V, D[c]);
        else if (d == 1) dsq<DY, NQ>(w + c * V, P.nx, D[c]);
        else dsq<DZ, NQ>(w + c * V, P.ss, D[c]);
      }
      // sv_curv_col_el_iso.c:398-406: xix*DxTxx + xiy*DxTxy + xiz*DxTxz + etx*DyTxx ...
#pragma unroll
      for (int v = 0; v < 3; v++)
#pragma unroll
        for (int q = 0; q < NQ; q++) {
          float a = m4[3 * d][q] * D[rows[v][0]][q];
          if (d > 0) a = acc[v][q] + a;
          a = a + m4[3 * d + 1][q] * D[rows[v][1]][q];
          a = a + m4[3 * d + 2][q] * D[rows[v][2]][q];
          acc[v][q] = a;
        }
      CGFDH_PHASE();
    }
    float s4[NQ];
    ldq_stream<NQ>(P.slw + p, s4);
    float rv[3][NQ];
#pragma unroll
    for (int v = 0; v < 3; v++)
#pragma unroll
      for (int q = 0; q < NQ; q++) rv[v][q] = s4[q] * acc[v][q];
    if ((P.nsrc > 0 && P.fxyz != nullptr) || P.ndd > 0) {
#pragma unroll
      for (int q = 0; q < NQ; q++) {
        float r[9] = {rv[0][q], rv[1][q], rv[2][q], 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (P.nsrc > 0 && P.fxyz != nullptr) add_sources(P, i + q, j, k, p + q, r);
        add_dd_for<MT>(P, p + q, r);
#pragma unroll
        for (int v = 0; v < 3; v++) rv[v][q] = r[VX + v];
      }
    }
#pragma unroll
    for (int v = 0; v < 3; v++) combine_q<KIND, NQ>(P, w, p, VX + v, rv[v]);
  }
}

template <int KIND, int MT, int BY, int NQ>
static void launch_interior_vec(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  const int grid = grouped_grid(B, BY, 64 * NQ);
  dim3 blk(64, BY, 1);
#define CGFDH_IV(c, DX, DY, DZ)                                                            \
  case c:                                                                                  \
    hipLaunchKernelGGL((stage_interior_vec_kernel<DX, DY, DZ, KIND, MT, BY, NQ>), dim3(grid), \
                       blk, 0, st, P, B);                                                  \
    break;
  switch (code) {
    CGFDH_IV(0, 0, 0, 0) CGFDH_IV(1, 0, 0, 1) CGFDH_IV(2, 0, 1, 0) CGFDH_IV(3, 0, 1, 1)
    CGFDH_IV(4, 1, 0, 0) CGFDH_IV(5, 1, 0, 1) CGFDH_IV(6, 1, 1, 0) CGFDH_IV(7, 1, 1, 1)
  }
#undef CGFDH_IV
}

}  // namespace cgfdh
