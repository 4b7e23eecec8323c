// stage_launch.h -- launch of one fused RK stage for one medium type.
//
// The kernels are templated on the medium (medium.h); each medium's
// instances are compiled in a translation unit of their own
// (stage_iso.hip, stage_vti.hip, stage_aniso.hip), which defines
// launch_stage_<medium> from launch_stage_mt below.
//
// Variants: 0 the per-point kernel over the whole grid; 1 the z-marching
// tiled kernel over the whole grid; 2/3 the z-marching interior kernel
// (64x8 / 64x16 tiles) plus the per-point kernel on the boundary shell;
// 4/5 the lean per-point interior kernel (64x4 / 64x2 blocks) plus the
// shell; 6/7 the x-vectorised interior kernel (2 points per thread on 64x4
// threads / 4 points per thread on 64x2 threads) plus the shell; 8/9 the
// LDS-tiled interior kernel (one 64x8 / 64x16 plane tile per block, x/y
// operands from LDS) plus the shell; 10/11 the per-point interior kernel
// marching 2 / 4 planes per block plus the shell; 12 / 13 the two-phase
// per-point interior kernel (all loads, then the arithmetic; 13 loads the RK
// combine inputs after the arithmetic) plus the shell; 14 the LDS-tiled
// plane kernel (64x8) with the same phased loads plus the shell; 15 variant
// 12 compiled for 3 waves per SIMD (a few registers spill) plus the shell;
// 16 the two-phase kernel over whole rows with the PML faces folded in, the
// free-surface rows on the general kernel; 17 the two-phase interior kernel
// on the tiles clear of every PML face and the phased shell kernel on the
// PML slabs (x edges in whole 64-point tiles, y and z slabs in whole rows),
// free-surface rows on the general kernel.
// The VTI and anisotropic media build variants 0 and 4..17, the
// visco-elastic medium all but 1..3, 6, 7 and 12..17.
#pragma once
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "stage_common.h"
#include "stage_tiled.h"
#include "stage_interior.h"
#include "stage_vec.h"

namespace cgfdh {

struct LaunchCtx {
  hipStream_t stream;
  int variant;
  int ni, nj, nk;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  int free_top;
  int pml_on[6], pml_layers[6];
  int tgx, tgy;            // tile group of the grouped block order (0: default)
  int merge_shell;         // 1: the five PML slabs of the shell in one launch
  hipStream_t shell_stream;  // stream of the shell boxes (= stream, or one beside it)
  int slab_only;           // variant 16: launch only the free-surface rows (the batch's
                           // fused launch runs the whole rows of every grid)
};

// the whole-row box of variant 16 with its block grouping resolved; returns
// the launch's block count (a multiple of 8)
inline int rows_box(const LaunchCtx& L, BoxParams& b)
{
  const int kmain = L.free_top ? L.nk2 - 4 : L.nk2;
  memset(&b, 0, sizeof b);
  b.gx = L.tgx; b.gy = L.tgy;
  b.i0 = L.ni1; b.j0 = L.nj1; b.k0 = L.nk1;
  b.ni = L.ni2 - L.ni1 + 1; b.nj = L.nj2 - L.nj1 + 1; b.nk = kmain - L.nk1 + 1;
  const int tiles_x = (b.ni + 63) / 64;
  if (b.gx <= 0) b.gx = tiles_x;
  if (b.gy <= 0) b.gy = 2;
  return grouped_grid(b, 4);
}

// per-point kernel: one thread per grid point, every stencil operand read
// from global memory.  The plain reference form of the stage.
template <int DX, int DY, int DZ, int KIND, int MT>
__global__ void __launch_bounds__(256) stage_point_kernel(StageParams P)
{
  const int i = P.ni1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = P.nj1 + blockIdx.y * blockDim.y + threadIdx.y;
  const int k = P.nk1 + blockIdx.z;
  if (i > P.ni2 || j > P.nj2) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const long V = P.vol;
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* v = P.wcur + c * V;
    X[c] = d5<DX>(v, p, 1);
    Y[c] = d5<DY>(v, p, P.nx);
    Z[c] = d5<DZ>(v, p, P.ss);
  }
  finish_point<DX, DY, DZ, KIND, MT>(P, i, j, k, p, X, Y, Z);
}

template <int KIND, int MT>
static void launch_kind(const LaunchCtx& L, const StageParams& P, int dx, int dy, int dz)
{
  const int code = dx * 4 + dy * 2 + dz;
  hipStream_t st = L.stream;
  if (L.variant == 0) {
    dim3 blk(64, 4, 1);
    dim3 grd((L.ni + 63) / 64, (L.nj + 3) / 4, L.nk);
#define CGFDH_PT(c, DX, DY, DZ)                                                              \
  case c:                                                                                    \
    hipLaunchKernelGGL((stage_point_kernel<DX, DY, DZ, KIND, MT>), grd, blk, 0, st, P); \
    break;
    switch (code) {
      CGFDH_PT(0, 0, 0, 0) CGFDH_PT(1, 0, 0, 1) CGFDH_PT(2, 0, 1, 0) CGFDH_PT(3, 0, 1, 1)
      CGFDH_PT(4, 1, 0, 0) CGFDH_PT(5, 1, 0, 1) CGFDH_PT(6, 1, 1, 0) CGFDH_PT(7, 1, 1, 1)
    }
#undef CGFDH_PT
    return;
  }
  if constexpr (MT == MEDIUM_ISO) {
    if (L.variant == 1) {
      launch_tiled<KIND, MT>(st, P, L.ni, L.nj, L.nk, code);
      return;
    }
  }
  // interior box (no PML face, no free-surface row) on a lean kernel; the
  // shell around it on the general per-point kernel
  const int ilo = L.pml_on[0] ? L.ni1 + L.pml_layers[0] + 1 : L.ni1;
  const int ihi = L.pml_on[1] ? L.ni2 - L.pml_layers[1] - 1 : L.ni2;
  const int jlo = L.pml_on[2] ? L.nj1 + L.pml_layers[2] + 1 : L.nj1;
  const int jhi = L.pml_on[3] ? L.nj2 - L.pml_layers[3] - 1 : L.nj2;
  const int klo = L.pml_on[4] ? L.nk1 + L.pml_layers[4] + 1 : L.nk1;
  const int khi = L.free_top ? L.nk2 - 4 : (L.pml_on[5] ? L.nk2 - L.pml_layers[5] - 1 : L.nk2);
  auto box = [&L](int i0, int i1, int j0, int j1, int k0, int k1) {
    BoxParams b;
    memset(&b, 0, sizeof b);
    b.gx = L.tgx; b.gy = L.tgy;
    b.i0 = i0; b.j0 = j0; b.k0 = k0;
    b.ni = i1 - i0 + 1; b.nj = j1 - j0 + 1; b.nk = k1 - k0 + 1;
    return b;
  };
  if (L.variant == 16 || (MT != MEDIUM_VISCO && L.variant == 18)) {
    // whole rows, PML faces included, on the two-phase kernel (18: its
    // operands without a stencil loaded straight to LDS, elastic media);
    // the free-surface rows on the general kernel
    const int kmain = L.free_top ? L.nk2 - 4 : L.nk2;
    const BoxParams rows = box(L.ni1, L.ni2, L.nj1, L.nj2, L.nk1, kmain);
    bool done = L.slab_only != 0;
    if constexpr (KIND != KIND_RHS && MT != MEDIUM_VISCO) {
      if (!done && L.variant == 18) {
        launch_rows_lds<KIND, MT, 4>(st, P, rows, code);
        done = true;
      }
    }
    if (!done) launch_rows_phased<KIND, MT, 4>(st, P, rows, code);
    if (L.free_top)
      launch_box<KIND, MT, HAS_SURF | HAS_PML>(L.shell_stream, P,
                                               box(L.ni1, L.ni2, L.nj1, L.nj2, kmain + 1,
                                                   L.nk2),
                                               code);
    return;
  }
  if constexpr (MT != MEDIUM_VISCO) {
    if (L.variant == 17) {
      // one 64-point tile grid along x from the first physical column: the
      // interior box holds the tiles clear of the x faces' PML, the tiles
      // that meet them form the x edge boxes; the y and z PML slabs span
      // whole rows.  Interior on the two-phase kernel; the PML slabs on the
      // phased shell kernel (each box requests its own face's memory
      // variables with its other operands); free-surface rows on the
      // general kernel.
      int slot[6], ns = 0;
      for (int f = 0; f < 6; f++) slot[f] = L.pml_on[f] ? ns++ : -1;
      int xa = L.ni1, xb = L.ni2 + 1;   // interior x range [xa, xb)
      if (L.pml_on[0]) xa = std::min(L.ni2 + 1, L.ni1 + (L.pml_layers[0] + 1 + 63) / 64 * 64);
      if (L.pml_on[1]) xb = L.ni1 + (L.ni2 - L.pml_layers[1] - L.ni1) / 64 * 64;
      if (xb < xa) xb = xa;
      const BoxParams inner = box(xa, xb - 1, jlo, jhi, klo, khi);
      if (inner.ni > 0 && inner.nj > 0 && inner.nk > 0)
        launch_interior_phased<KIND, MT, 4, true>(st, P, inner, code);
      BoxParams sh[6];
      int nsh = 0;
      auto add = [&](BoxParams b, int face) {
        b.fslot = face >= 0 ? slot[face] : -1;
        if (b.ni > 0 && b.nj > 0 && b.nk > 0) sh[nsh++] = b;
      };
      if (L.pml_on[4]) add(box(L.ni1, L.ni2, L.nj1, L.nj2, L.nk1, klo - 1), 4);
      if (!L.free_top && L.pml_on[5]) add(box(L.ni1, L.ni2, L.nj1, L.nj2, khi + 1, L.nk2), 5);
      if (L.pml_on[2]) add(box(L.ni1, L.ni2, L.nj1, jlo - 1, klo, khi), 2);
      if (L.pml_on[3]) add(box(L.ni1, L.ni2, jhi + 1, L.nj2, klo, khi), 3);
      if (xa > L.ni1) add(box(L.ni1, xa - 1, jlo, jhi, klo, khi), L.pml_on[0] ? 0 : -1);
      if (xb <= L.ni2) add(box(xb, L.ni2, jlo, jhi, klo, khi), L.pml_on[1] ? 1 : -1);
      launch_shellp<KIND, MT>(L.shell_stream, P, sh, nsh, code);
      if (L.free_top)
        launch_box<KIND, MT, HAS_SURF | HAS_PML>(L.shell_stream, P,
                                                 box(L.ni1, L.ni2, L.nj1, L.nj2, khi + 1,
                                                     L.nk2),
                                                 code);
      return;
    }
  }
  const BoxParams in = box(ilo, ihi, jlo, jhi, klo, khi);
  if (MT != MEDIUM_VISCO && in.ni > 0 && in.nj > 0 && in.nk > 0 &&
      (L.variant == 6 || L.variant == 7)) {
    // whole point groups on the vectorised kernel, the last columns on the
    // per-point box kernel
    BoxParams vb = in;
    const int nq = L.variant == 7 ? 4 : 2;
    vb.ni = in.ni / nq * nq;
    if constexpr (MT != MEDIUM_VISCO) {
      if (vb.ni > 0) {
        if (L.variant == 7)
          launch_interior_vec<KIND, MT, 2, 4>(st, P, vb, code);
        else
          launch_interior_vec<KIND, MT, 4, 2>(st, P, vb, code);
      }
    }
    if (in.ni > vb.ni)
      launch_box<KIND, MT, HAS_PML>(st, P, box(ilo + vb.ni, ihi, jlo, jhi, klo, khi), code);
  } else if (in.ni > 0 && in.nj > 0 && in.nk > 0) {
    if (L.variant == 15) {
      if constexpr (MT != MEDIUM_VISCO) launch_interior_phased<KIND, MT, 4, true, 3>(st, P, in, code);
    } else if (L.variant == 14) {
      if constexpr (MT != MEDIUM_VISCO) launch_interior_ldsp<KIND, MT, 8>(st, P, in, code);
    } else if (L.variant == 12 || L.variant == 13) {
      if constexpr (MT != MEDIUM_VISCO) {
        if (L.variant == 12)
          launch_interior_phased<KIND, MT, 4, true>(st, P, in, code);
        else
          launch_interior_phased<KIND, MT, 4, false>(st, P, in, code);
      }
    } else if (L.variant == 10) {
      launch_interior_pointz<KIND, MT, 4, 2>(st, P, in, code);
    } else if (L.variant == 11) {
      launch_interior_pointz<KIND, MT, 4, 4>(st, P, in, code);
    } else if (L.variant == 8) {
      launch_interior_lds<KIND, MT, 8>(st, P, in, code);
    } else if (L.variant == 9) {
      launch_interior_lds<KIND, MT, 16>(st, P, in, code);
    } else if (L.variant == 5) {
      launch_interior_point<KIND, MT, 2>(st, P, in, code);
    } else if (L.variant == 4) {
      launch_interior_point<KIND, MT, 4>(st, P, in, code);
    } else {
      if constexpr (MT == MEDIUM_ISO) {
        if (L.variant == 3)
          launch_interior_ty<KIND, MT, 16>(st, P, in, code);
        else
          launch_interior_ty<KIND, MT, 8>(st, P, in, code);
      } else {
        launch_interior_point<KIND, MT, 4>(st, P, in, code);
      }
    }
  }
  // only the top slab holds free-surface rows (the others end at khi)
  constexpr int SURF = HAS_SURF | HAS_PML;
  st = L.shell_stream;
  launch_box<KIND, MT, SURF>(st, P, box(L.ni1, L.ni2, L.nj1, L.nj2, khi + 1, L.nk2), code);
  const BoxParams faces[5] = {box(L.ni1, L.ni2, L.nj1, L.nj2, L.nk1, klo - 1),
                              box(L.ni1, ilo - 1, L.nj1, L.nj2, klo, khi),
                              box(ihi + 1, L.ni2, L.nj1, L.nj2, klo, khi),
                              box(ilo, ihi, L.nj1, jlo - 1, klo, khi),
                              box(ilo, ihi, jhi + 1, L.nj2, klo, khi)};
  if (L.merge_shell) {
    // the five side and bottom slabs in one launch
    launch_shell<KIND, MT, HAS_PML>(st, P, faces, 5, code);
  } else {
    for (const BoxParams& f : faces) launch_box<KIND, MT, HAS_PML>(st, P, f, code);
  }
}

template <int MT>
void launch_stage_mt(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz)
{
  switch (kind) {
    case KIND_FIRST: launch_kind<KIND_FIRST, MT>(L, P, dx, dy, dz); break;
    case KIND_MID: launch_kind<KIND_MID, MT>(L, P, dx, dy, dz); break;
    case KIND_LAST: launch_kind<KIND_LAST, MT>(L, P, dx, dy, dz); break;
    default: launch_kind<KIND_RHS, MT>(L, P, dx, dy, dz); break;
  }
}

template <int MT>
void launch_rows_multi_mt(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                          int dy, int dz)
{
  const int code = dx * 4 + dy * 2 + dz;
  switch (kind) {
    case KIND_FIRST: launch_rows_multi<KIND_FIRST, MT, 4>(st, R, nblocks, code); break;
    case KIND_MID: launch_rows_multi<KIND_MID, MT, 4>(st, R, nblocks, code); break;
    default: launch_rows_multi<KIND_LAST, MT, 4>(st, R, nblocks, code); break;
  }
}

// defined in stage_iso.hip, stage_vti.hip and stage_aniso.hip
void launch_stage_iso(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz);
void launch_stage_vti(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy, int dz);
void launch_stage_aniso(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy,
                        int dz);
void launch_stage_visco(const LaunchCtx& L, const StageParams& P, int kind, int dx, int dy,
                        int dz);
void launch_rows_multi_iso(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                           int dy, int dz);
void launch_rows_multi_vti(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                           int dy, int dz);
void launch_rows_multi_visco(hipStream_t st, const RowsMulti& R, int nblocks, int kind, int dx,
                             int dy, int dz);

}  // namespace cgfdh
