// stage_tiled.h -- the z-marching stage kernel.
//
// A block owns a 64 x TY column of grid points and walks a chunk of k planes.
// Each thread keeps the five z-window values of its point for all 9 fields in
// registers (a queue shifted once per plane), so z derivatives cost one new
// load per field per plane.  The current plane of the block, plus the x and y
// halo the one-sided operators need, is staged in LDS for the x and y
// derivatives.  Everything after the derivatives is the shared finish_point
// (surface rows, PML, sources, RK combine), so the tiled and per-point kernels
// produce identical bits.
#pragma once

namespace cgfdh {

constexpr int TILE_X = 64;
constexpr int TILE_Y = 8;
constexpr int HALO = 4;                       // 1+3 or 3+1 along x and along y
constexpr int LDS_W = TILE_X + HALO;          // 68
constexpr int LDS_H = TILE_Y + HALO;          // 12
constexpr int LDS_PLANE = LDS_W * LDS_H;

template <int DX, int DY, int DZ, int KIND, int MT>
__global__ void __launch_bounds__(TILE_X * TILE_Y)
stage_tiled_kernel(StageParams P)
{
  __shared__ float sm[9 * LDS_PLANE];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int tid = ty * TILE_X + tx;
  int b = blockIdx.x;
  const int ntile = P.tiles_x * P.tiles_y;
  const int chunk = b / ntile;
  b -= chunk * ntile;
  const int tyi = b / P.tiles_x;
  const int txi = b - tyi * P.tiles_x;
  const int i0 = P.ni1 + txi * TILE_X;
  const int j0 = P.nj1 + tyi * TILE_Y;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = P.nk1 + chunk * P.kchunk;
  const int kend = min(P.nk2, kbeg + P.kchunk - 1);
  const int nx = P.nx;
  const int ny = (int)(P.ss / nx);
  const long V = P.vol, ss = P.ss;
  const bool inrow = (i < nx) && (j < ny);        // inside the allocation
  const bool valid = (i <= P.ni2) && (j <= P.nj2);
  const long pij = (long)i + (long)j * nx;
  constexpr int xl = -Win<DX>::lo;              // left halo along x: 1 or 3
  constexpr int yl = -Win<DY>::lo;              // left halo along y
  constexpr int zlo = Win<DZ>::lo;
  constexpr int zc = -zlo;                      // centre slot of the z queue

  // y-halo rows (tile-relative): DY=0 needs -1 and TY..TY+2; DY=1 needs -3..-1 and TY
  // x-halo cols: DX=0 needs -1 and 64..66; DX=1 needs -3..-1 and 64

  float q[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++) {
#pragma unroll
    for (int n = 1; n < 5; n++) {
      const long pz = pij + (long)(kbeg + zlo + n - 1) * ss;
      q[c][n] = inrow ? P.wcur[c * V + pz] : 0.0f;
    }
  }

  for (int k = kbeg; k <= kend; k++) {
    // advance the z queue: slot n holds plane k + zlo + n
#pragma unroll
    for (int c = 0; c < 9; c++) {
#pragma unroll
      for (int n = 0; n < 4; n++) q[c][n] = q[c][n + 1];
      const long pz = pij + (long)(k + zlo + 4) * ss;
      q[c][4] = inrow ? P.wcur[c * V + pz] : 0.0f;
    }
    __syncthreads();   // the previous plane's LDS reads are done
    // interior of the plane from registers
#pragma unroll
    for (int c = 0; c < 9; c++) sm[c * LDS_PLANE + (ty + yl) * LDS_W + (tx + xl)] = q[c][zc];
    // halos from global memory
    const long pk = (long)k * ss;
    if (ty < HALO) {
      // y halo: one row per thread row 0..3, 64 columns
      const int rrel = DY == 0 ? (ty == 0 ? -1 : TILE_Y + ty - 1) : (ty < 3 ? ty - 3 : TILE_Y);
      const int jj = j0 + rrel;
      const bool ok = (i < nx) && (jj >= 0) && (jj < ny);
      const long pp = (long)i + (long)jj * nx + pk;
#pragma unroll
      for (int c = 0; c < 9; c++)
        sm[c * LDS_PLANE + (rrel + yl) * LDS_W + (tx + xl)] = ok ? P.wcur[c * V + pp] : 0.0f;
    } else if (ty < HALO + 1 && tx < HALO * TILE_Y) {
      // x halo: 4 columns x TILE_Y rows handled by 32 threads of thread row 4
      const int row = tx / HALO, hc = tx % HALO;
      const int crel = DX == 0 ? (hc == 0 ? -1 : TILE_X + hc - 1) : (hc < 3 ? hc - 3 : TILE_X);
      const int ii = i0 + crel, jj = j0 + row;
      const bool ok = (ii >= 0) && (ii < nx) && (jj < ny);
      const long pp = (long)ii + (long)jj * nx + pk;
#pragma unroll
      for (int c = 0; c < 9; c++)
        sm[c * LDS_PLANE + (row + yl) * LDS_W + (crel + xl)] = ok ? P.wcur[c * V + pp] : 0.0f;
    }
    __syncthreads();
    if (valid) {
      float X[9], Y[9], Z[9];
#pragma unroll
      for (int c = 0; c < 9; c++) {
        const float* s = sm + c * LDS_PLANE;
        const int rowb = (ty + yl) * LDS_W + tx;           // x window starts at tx+xl-xl
        X[c] = d5v<DX>(s[rowb + 0], s[rowb + 1], s[rowb + 2], s[rowb + 3], s[rowb + 4]);
        const int colb = ty * LDS_W + tx + xl;             // y window starts at ty+yl-yl
        Y[c] = d5v<DY>(s[colb], s[colb + LDS_W], s[colb + 2 * LDS_W], s[colb + 3 * LDS_W],
                       s[colb + 4 * LDS_W]);
        Z[c] = d5v<DZ>(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4]);
      }
      finish_point<DX, DY, DZ, KIND, MT>(P, i, j, k, pij + pk, X, Y, Z);
    }
  }
}

template <int KIND, int MT>
static void launch_tiled(hipStream_t st, StageParams P, int ni, int nj, int nk, int code)
{
  P.tiles_x = (ni + TILE_X - 1) / TILE_X;
  P.tiles_y = (nj + TILE_Y - 1) / TILE_Y;
  // enough blocks to fill 256 CUs several times over
  const int ntile = P.tiles_x * P.tiles_y;
  int nchunk = (4096 + ntile - 1) / ntile;
  if (nchunk > nk / 8) nchunk = nk / 8 > 0 ? nk / 8 : 1;
  if (nchunk < 1) nchunk = 1;
  P.kchunk = (nk + nchunk - 1) / nchunk;
  nchunk = (nk + P.kchunk - 1) / P.kchunk;
  dim3 blk(TILE_X, TILE_Y, 1);
  dim3 grd(ntile * nchunk, 1, 1);
#define CGFDH_TL(c, DX, DY, DZ) \
  case c: hipLaunchKernelGGL((stage_tiled_kernel<DX, DY, DZ, KIND, MT>), grd, blk, 0, st, P); break;
  switch (code) {
    CGFDH_TL(0, 0, 0, 0) CGFDH_TL(1, 0, 0, 1) CGFDH_TL(2, 0, 1, 0) CGFDH_TL(3, 0, 1, 1)
    CGFDH_TL(4, 1, 0, 0) CGFDH_TL(5, 1, 0, 1) CGFDH_TL(6, 1, 1, 0) CGFDH_TL(7, 1, 1, 1)
  }
#undef CGFDH_TL
}

}  // namespace cgfdh
