// stage_interior.h -- the stage kernel of the interior box: every point
// whose stencil, Hooke's law and RK combine need no PML face, no
// free-surface row and no one-sided operator.  The boundary shell around it
// runs the general per-point kernel; together they cover each point once.
//
// Block = 64 x TY threads over a 64 x TY tile of one z chunk; the block walks
// its k planes keeping the 5-value z window of all 9 fields in registers and
// stages the current plane (plus its x/y halo) in a double-buffered LDS tile,
// so each plane costs one barrier.  Block ids are remapped so that each XCD
// owns a contiguous run of tiles (x fastest, then y, then chunk): tiles that
// share cache lines and halo rows meet in the same L2.
#pragma once

namespace cgfdh {

struct BoxParams {
  int i0, j0, k0;           // first point of the box
  int ni, nj, nk;           // extent
  int tiles_x, tiles_y, nchunk, kchunk;
  int nblocks;              // real blocks (grid is padded to a multiple of 8)
};

template <int TY> struct LdsTile {
  static constexpr int W = 64 + 4;
  static constexpr int H = TY + 4;
  static constexpr int PLANE = W * H;
};

template <int DX, int DY, int DZ, int KIND, int MT, int TY>
__global__ void __launch_bounds__(64 * TY)
stage_interior_kernel(StageParams P, BoxParams B)
{
  using T = LdsTile<TY>;
  __shared__ float sm[2][9 * T::PLANE];
  const int tx = threadIdx.x, ty = threadIdx.y;
  // XCD-aware block order: hardware dispatches block b to XCD b % 8
  const int nb8 = gridDim.x / 8;
  int b = (blockIdx.x % 8) * nb8 + blockIdx.x / 8;
  if (b >= B.nblocks) return;
  const int ntile = B.tiles_x * B.tiles_y;
  const int chunk = b / ntile;
  b -= chunk * ntile;
  const int tyi = b / B.tiles_x;
  const int txi = b - tyi * B.tiles_x;
  const int i0 = B.i0 + txi * 64;
  const int j0 = B.j0 + tyi * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = B.k0 + chunk * B.kchunk;
  const int kend = min(B.k0 + B.nk - 1, kbeg + B.kchunk - 1);
  const int nx = P.nx;
  const long V = P.vol, ss = P.ss;
  // tiles may run past the box (and past the allocation's rows): such
  // threads load zeros or in-bounds neighbours and compute nothing
  const bool valid = (i < B.i0 + B.ni) && (j < B.j0 + B.nj);
  const bool inrow = (i < nx) && (j < (int)(ss / nx));
  const long pij = (long)i + (long)j * nx;
  constexpr int xl = -Win<DX>::lo;
  constexpr int yl = -Win<DY>::lo;
  constexpr int zlo = Win<DZ>::lo;
  constexpr int zc = -zlo;

  float q[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++)
#pragma unroll
    for (int n = 1; n < 5; n++)
      q[c][n] = inrow ? P.wcur[c * V + pij + (long)(kbeg + zlo + n - 1) * ss] : 0.0f;

  int buf = 0;
  for (int k = kbeg; k <= kend; k++, buf ^= 1) {
#pragma unroll
    for (int c = 0; c < 9; c++) {
#pragma unroll
      for (int n = 0; n < 4; n++) q[c][n] = q[c][n + 1];
      q[c][4] = inrow ? P.wcur[c * V + pij + (long)(k + zlo + 4) * ss] : 0.0f;
    }
    float* s0 = sm[buf];
#pragma unroll
    for (int c = 0; c < 9; c++) s0[c * T::PLANE + (ty + yl) * T::W + (tx + xl)] = q[c][zc];
    const long pk = (long)k * ss;
    if (ty < 4) {
      const int rrel = DY == 0 ? (ty == 0 ? -1 : TY + ty - 1) : (ty < 3 ? ty - 3 : TY);
      const int jj = j0 + rrel;
      const long pp = (long)i + (long)jj * nx + pk;
      const bool ok = i < nx && jj < (int)(ss / nx);
#pragma unroll
      for (int c = 0; c < 9; c++)
        s0[c * T::PLANE + (rrel + yl) * T::W + (tx + xl)] = ok ? P.wcur[c * V + pp] : 0.0f;
    } else if (ty == 4 && tx < 4 * TY) {
      const int row = tx / 4, hc = tx % 4;
      const int crel = DX == 0 ? (hc == 0 ? -1 : 64 + hc - 1) : (hc < 3 ? hc - 3 : 64);
      const int ii = i0 + crel;
      const long pp = (long)ii + (long)(j0 + row) * nx + pk;
      const bool ok = ii < nx && j0 + row < (int)(ss / nx);
#pragma unroll
      for (int c = 0; c < 9; c++)
        s0[c * T::PLANE + (row + yl) * T::W + (crel + xl)] = ok ? P.wcur[c * V + pp] : 0.0f;
    }
    __syncthreads();
    if (valid) {
      const long p = pij + pk;
      float X[9], Y[9], Z[9];
#pragma unroll
      for (int c = 0; c < 9; c++) {
        const float* s = s0 + c * T::PLANE;
        const int rowb = (ty + yl) * T::W + tx;
        X[c] = d5v<DX>(s[rowb], s[rowb + 1], s[rowb + 2], s[rowb + 3], s[rowb + 4]);
        const int colb = ty * T::W + tx + xl;
        Y[c] = d5v<DY>(s[colb], s[colb + T::W], s[colb + 2 * T::W], s[colb + 3 * T::W],
                       s[colb + 4 * T::W]);
        Z[c] = d5v<DZ>(q[c][0], q[c][1], q[c][2], q[c][3], q[c][4]);
      }
      interior_finish<KIND, MT>(P, i, j, k, p, X, Y, Z);
    }
    // the other LDS buffer is written next; this one is reused two planes on,
    // after the next barrier, so one barrier per plane suffices
  }
}

template <int KIND, int MT, int TY>
static void launch_interior_ty(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  B.tiles_x = (B.ni + 63) / 64;
  B.tiles_y = (B.nj + TY - 1) / TY;
  const int ntile = B.tiles_x * B.tiles_y;
  int nchunk = (6144 + ntile - 1) / ntile;
  nchunk = std::max(1, std::min(nchunk, B.nk / 16));
  B.kchunk = (B.nk + nchunk - 1) / nchunk;
  B.nchunk = (B.nk + B.kchunk - 1) / B.kchunk;
  B.nblocks = ntile * B.nchunk;
  const int grid = (B.nblocks + 7) / 8 * 8;
  dim3 blk(64, TY, 1);
#define CGFDH_IN(c, DX, DY, DZ)                                                              \
  case c:                                                                                    \
    hipLaunchKernelGGL((stage_interior_kernel<DX, DY, DZ, KIND, MT, TY>), dim3(grid), blk, 0, st, \
                       P, B);                                                                \
    break;
  switch (code) {
    CGFDH_IN(0, 0, 0, 0) CGFDH_IN(1, 0, 0, 1) CGFDH_IN(2, 0, 1, 0) CGFDH_IN(3, 0, 1, 1)
    CGFDH_IN(4, 1, 0, 0) CGFDH_IN(5, 1, 0, 1) CGFDH_IN(6, 1, 1, 0) CGFDH_IN(7, 1, 1, 1)
  }
#undef CGFDH_IN
}

// lean per-point kernel for the interior box: one thread per point, stencil
// operands straight from global memory (L1/L2 serve the overlap), no LDS and
// no barrier; block ids remapped so each XCD owns a contiguous run of
// (x,y,z) blocks.  BY rows of 64 points per block.
template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__global__ void __launch_bounds__(64 * BY)
stage_interior_point_kernel(StageParams P, BoxParams B)
{
  const int nb8 = gridDim.x / 8;
  int b = (blockIdx.x % 8) * nb8 + blockIdx.x / 8;
  if (b >= B.nblocks) return;
  const int bz = b / (B.tiles_x * B.tiles_y);
  b -= bz * B.tiles_x * B.tiles_y;
  const int by = b / B.tiles_x;
  const int bx = b - by * B.tiles_x;
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
  const int k = B.k0 + bz;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
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
  interior_finish<KIND, MT>(P, i, j, k, p, X, Y, Z);
}

template <int KIND, int MT, int BY>
static void launch_interior_point(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  B.tiles_x = (B.ni + 63) / 64;
  B.tiles_y = (B.nj + BY - 1) / BY;
  B.nblocks = B.tiles_x * B.tiles_y * B.nk;
  const int grid = (B.nblocks + 7) / 8 * 8;
  dim3 blk(64, BY, 1);
#define CGFDH_IP(c, DX, DY, DZ)                                                          \
  case c:                                                                                \
    hipLaunchKernelGGL((stage_interior_point_kernel<DX, DY, DZ, KIND, MT, BY>), dim3(grid), blk, \
                       0, st, P, B);                                                     \
    break;
  switch (code) {
    CGFDH_IP(0, 0, 0, 0) CGFDH_IP(1, 0, 0, 1) CGFDH_IP(2, 0, 1, 0) CGFDH_IP(3, 0, 1, 1)
    CGFDH_IP(4, 1, 0, 0) CGFDH_IP(5, 1, 0, 1) CGFDH_IP(6, 1, 1, 0) CGFDH_IP(7, 1, 1, 1)
  }
#undef CGFDH_IP
}

// general per-point kernel restricted to a box (the boundary shell); FLAGS
// says whether the box touches the free-surface rows
template <int DX, int DY, int DZ, int KIND, int MT, int FLAGS>
__global__ void __launch_bounds__(256) stage_box_kernel(StageParams P, BoxParams B)
{
  const int i = B.i0 + blockIdx.x * 64 + threadIdx.x;
  const int j = B.j0 + blockIdx.y * 4 + threadIdx.y;
  const int k = B.k0 + blockIdx.z;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
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
  finish_point<DX, DY, DZ, KIND, MT, FLAGS>(P, i, j, k, p, X, Y, Z);
}

template <int KIND, int MT, int FLAGS>
static void launch_box(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  if (B.ni <= 0 || B.nj <= 0 || B.nk <= 0) return;
  dim3 blk(64, 4, 1);
  dim3 grd((B.ni + 63) / 64, (B.nj + 3) / 4, B.nk);
#define CGFDH_BX(c, DX, DY, DZ)                                                                \
  case c:                                                                                      \
    hipLaunchKernelGGL((stage_box_kernel<DX, DY, DZ, KIND, MT, FLAGS>), grd, blk, 0, st, P, B); \
    break;
  switch (code) {
    CGFDH_BX(0, 0, 0, 0) CGFDH_BX(1, 0, 0, 1) CGFDH_BX(2, 0, 1, 0) CGFDH_BX(3, 0, 1, 1)
    CGFDH_BX(4, 1, 0, 0) CGFDH_BX(5, 1, 0, 1) CGFDH_BX(6, 1, 1, 0) CGFDH_BX(7, 1, 1, 1)
  }
#undef CGFDH_BX
}

}  // namespace cgfdh
