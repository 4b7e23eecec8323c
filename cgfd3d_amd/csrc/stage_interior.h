// stage_interior.h -- the stage kernel of the interior box: every point
// whose stencil, Hooke's law and RK combine need no PML face, no
// free-surface row and no one-sided operator.  The boundary shell around it
// runs the general per-point kernel; together they cover each point once.
//
// Block = 64 x TY threads over a 64 x TY tile of one z chunk; the block walks
// its k planes keeping the 5-value z window of all 9 fields in registers and
// stages the current plane (plus its x/y halo) in a double-buffered LDS tile,
// so each plane costs one barrier.  The newest window entry and the halo of
// the next plane are loaded one plane ahead, so their HBM latency overlaps
// the current plane's arithmetic.  Block ids are remapped so that each XCD
// owns a contiguous run of tiles (x fastest, then y, then chunk): tiles that
// share cache lines and halo rows meet in the same L2.
#pragma once

namespace cgfdh {

struct BoxParams {
  int fslot;                // phased shell kernel: the PML face slot whose memory
                            // variables are requested with the other operands (-1 none)
  int i0, j0, k0;           // first point of the box
  int ni, nj, nk;           // extent
  int tiles_x, tiles_y, nchunk, kchunk;
  int nblocks;              // real blocks (grid is padded to a multiple of 8)
  int gx, gy;               // per-point kernels: tiles per block group in x, y
  int groups_x;             // groups along x
  int tw;                   // box kernel: tile width in x (tile = tw x 256/tw)
};

// Tile-group shape of the per-point interior kernel, "GXxGY" tiles, from
// CGFDH_TILE_GROUP when set (a tuning knob), else 9x2.
inline void tile_group(int& gx, int& gy)
{
  static int cgx = 0, cgy = 0;
  if (!cgx) {
    cgx = 9; cgy = 2;
    if (const char* e = getenv("CGFDH_TILE_GROUP")) {
      int a = 0, b = 0;
      if (sscanf(e, "%dx%d", &a, &b) == 2 && a > 0 && b > 0) { cgx = a; cgy = b; }
    }
  }
  gx = cgx; gy = cgy;
}

// block of a grouped, XCD-remapped 1-D grid over a box tiled 64 x by: the
// tile (bx, by) and plane k it covers, or false for padding blocks
__device__ __forceinline__ bool grouped_block(const BoxParams& B, int& bx, int& by, int& k)
{
  const int nb8 = gridDim.x / 8;
  const int b = (blockIdx.x % 8) * nb8 + blockIdx.x / 8;
  if (b >= B.nblocks) return false;
  const int gt = B.gx * B.gy;
  const int per_group = gt * B.nk;
  const int g = b / per_group;
  const int r = b - g * per_group;
  const int kk = r / gt;
  const int t = r - kk * gt;
  const int gyi = g / B.groups_x;
  bx = (g - gyi * B.groups_x) * B.gx + t % B.gx;
  by = gyi * B.gy + t / B.gx;
  k = B.k0 + kk;
  return bx < B.tiles_x && by < B.tiles_y;
}

// grouped_block for block bidx of a sub-grid of nblk blocks (nblk a
// multiple of 8 starting at a multiple of 8, so bidx % 8 is still the XCD)
__device__ __forceinline__ bool grouped_block_at(const BoxParams& B, int bidx, int nblk, int& bx,
                                                 int& by, int& k)
{
  const int nb8 = nblk / 8;
  const int b = (bidx % 8) * nb8 + bidx / 8;
  if (b >= B.nblocks) return false;
  const int gt = B.gx * B.gy;
  const int per_group = gt * B.nk;
  const int g = b / per_group;
  const int r = b - g * per_group;
  const int kk = r / gt;
  const int t = r - kk * gt;
  const int gyi = g / B.groups_x;
  bx = (g - gyi * B.groups_x) * B.gx + t % B.gx;
  by = gyi * B.gy + t / B.gx;
  k = B.k0 + kk;
  return bx < B.tiles_x && by < B.tiles_y;
}

// the grouped grid of a box tiled tw x by
inline int grouped_grid(BoxParams& B, int by, int tw = 64)
{
  B.tiles_x = (B.ni + tw - 1) / tw;
  B.tiles_y = (B.nj + by - 1) / by;
  if (B.gx <= 0 || B.gy <= 0) tile_group(B.gx, B.gy);
  B.gx = std::min(B.gx, B.tiles_x);
  B.gy = std::min(B.gy, B.tiles_y);
  B.groups_x = (B.tiles_x + B.gx - 1) / B.gx;
  const int groups_y = (B.tiles_y + B.gy - 1) / B.gy;
  B.nblocks = B.groups_x * groups_y * B.gx * B.gy * B.nk;
  return (B.nblocks + 7) / 8 * 8;
}

template <int TY> struct LdsTile {
  static constexpr int W = 64 + 4;
  static constexpr int H = TY + 4;
  static constexpr int PLANE = W * H;
};

template <int DX, int DY, int DZ, int KIND, int MT, int TY>
__global__ void __launch_bounds__(64 * TY, TY <= 8 ? 2 : 4)
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
  const int ny = (int)(P.ss / nx);
  const long V = P.vol, ss = P.ss;
  // tiles may run past the box (and past the allocation's rows): such
  // threads load zeros or in-bounds neighbours and compute nothing
  const bool valid = (i < B.i0 + B.ni) && (j < B.j0 + B.nj);
  const bool inrow = (i < nx) && (j < ny);
  const long pij = (long)i + (long)j * nx;
  constexpr int xl = -Win<DX>::lo;
  constexpr int yl = -Win<DY>::lo;
  constexpr int zlo = Win<DZ>::lo;
  constexpr int zc = -zlo;

  // halo duty: threads ty < 4 load one y-halo row each, threads of row 4
  // with tx < 4*TY one x-halo column entry; hidx is the LDS slot, hp the
  // point (minus the plane offset), hok whether it lies in the allocation
  int hslot = -1;
  long hp = 0;
  bool hok = false;
  if (ty < 4) {
    const int rrel = DY == 0 ? (ty == 0 ? -1 : TY + ty - 1) : (ty < 3 ? ty - 3 : TY);
    const int jj = j0 + rrel;
    hslot = (rrel + yl) * T::W + (tx + xl);
    hp = (long)i + (long)jj * nx;
    hok = i < nx && jj < ny;
  } else if (ty == 4 && tx < 4 * TY) {
    const int row = tx / 4, hc = tx % 4;
    const int crel = DX == 0 ? (hc == 0 ? -1 : 64 + hc - 1) : (hc < 3 ? hc - 3 : 64);
    const int ii = i0 + crel;
    hslot = (row + yl) * T::W + (crel + xl);
    hp = (long)ii + (long)(j0 + row) * nx;
    hok = ii < nx && j0 + row < ny;
  }

  // software pipeline: the z window of plane k sits in q, the newest window
  // entry and the halo of plane k+1 are loaded one iteration ahead
  float q[9][5], qn[9], hv[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
#pragma unroll
    for (int n = 0; n < 4; n++)
      q[c][n + 1] = inrow ? P.wcur[c * V + pij + (long)(kbeg + zlo + n) * ss] : 0.0f;
    qn[c] = inrow ? P.wcur[c * V + pij + (long)(kbeg + zlo + 4) * ss] : 0.0f;
    hv[c] = hok ? P.wcur[c * V + hp + (long)kbeg * ss] : 0.0f;
  }

  int buf = 0;
  for (int k = kbeg; k <= kend; k++, buf ^= 1) {
    float* s0 = sm[buf];
#pragma unroll
    for (int c = 0; c < 9; c++) {
#pragma unroll
      for (int n = 0; n < 4; n++) q[c][n] = q[c][n + 1];
      q[c][4] = qn[c];
      s0[c * T::PLANE + (ty + yl) * T::W + (tx + xl)] = q[c][zc];
      if (hslot >= 0) s0[c * T::PLANE + hslot] = hv[c];
    }
    if (k < kend) {
      const long pn = (long)(k + 1) * ss;
#pragma unroll
      for (int c = 0; c < 9; c++) {
        qn[c] = inrow ? P.wcur[c * V + pij + (long)(k + 1 + zlo + 4) * ss] : 0.0f;
        hv[c] = hok ? P.wcur[c * V + hp + pn] : 0.0f;
      }
    }
    __syncthreads();
    if (valid) {
      const long p = pij + (long)k * ss;
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
// no barrier.  BY rows of 64 points per block.  Block order: groups of
// gx x gy tiles, k planes inside a group, tiles fastest; block ids remapped
// so each XCD owns a contiguous run.  The blocks an XCD runs together then
// cover a compact brick of tiles and planes, so the z and y neighbours one
// block reads are mostly still in that XCD's L2 when its neighbours read them.
template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__global__ void __launch_bounds__(64 * BY)
stage_interior_point_kernel(StageParams P, BoxParams B)
{
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
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
  const int grid = grouped_grid(B, BY);
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

// the per-point kernel in two phases for the elastic media: every operand a
// point needs (the 13 stencil points of the 9 fields, metric, medium, the RK
// combine inputs) is requested first, then the arithmetic runs.  A schedule
// barrier keeps the compiler from interleaving loads and waits, so each
// wave has all its loads in flight at once (fewer waves, far more memory
// requests outstanding per wave).
template <int DX, int DY, int DZ, int KIND, int MT, int BY, bool PRE, int MINB = 1>
__global__ void __launch_bounds__(64 * BY, MINB)
stage_interior_phased_kernel(StageParams P, BoxParams B)
{
  static_assert(MT != MEDIUM_VISCO, "elastic media only");
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const long V = P.vol;
  constexpr int xl = Win<DX>::lo, yl = Win<DY>::lo, zl = Win<DZ>::lo;
  float ax[9][5], ay[9][5], az[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* v = P.wcur + c * V;
#pragma unroll
    for (int n = 0; n < 5; n++) {
      ax[c][n] = v[p + (xl + n)];
      ay[c][n] = v[p + (long)(yl + n) * P.nx];
      az[c][n] = v[p + (long)(zl + n) * P.ss];
    }
  }
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = ld_stream(P.met + (1 + n) * V + p);
  float mv[Medium<MT>::ncoef];
#pragma unroll
  for (int n = 0; n < Medium<MT>::ncoef; n++) mv[n] = ld_stream(P.med + n * V + p);
  const float slw = ld_stream(P.slw + p);
  // PRE: the RK combine inputs join the first phase too (more registers)
  float cpre[9], cend[9];
  if (PRE) combine_load<KIND>(P, p, cpre, cend);
  __builtin_amdgcn_sched_barrier(0);
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    X[c] = d5v<DX>(ax[c][0], ax[c][1], ax[c][2], ax[c][3], ax[c][4]);
    Y[c] = d5v<DY>(ay[c][0], ay[c][1], ay[c][2], ay[c][3], ay[c][4]);
    Z[c] = d5v<DZ>(az[c][0], az[c][1], az[c][2], az[c][3], az[c][4]);
  }
  Medium<MT> md;
  md.from(mv);
  float r[9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  if (!PRE) combine_load<KIND>(P, p, cpre, cend);
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  combine_store<KIND>(P, p, r, cpre, cend);
}

template <int KIND, int MT, int BY, bool PRE, int MINB = 1>
static void launch_interior_phased(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  const int grid = grouped_grid(B, BY);
  dim3 blk(64, BY, 1);
#define CGFDH_IF(c, DX, DY, DZ)                                                            \
  case c:                                                                                  \
    hipLaunchKernelGGL((stage_interior_phased_kernel<DX, DY, DZ, KIND, MT, BY, PRE, MINB>),  \
                       dim3(grid), blk, 0, st, P, B);                                      \
    break;
  switch (code) {
    CGFDH_IF(0, 0, 0, 0) CGFDH_IF(1, 0, 0, 1) CGFDH_IF(2, 0, 1, 0) CGFDH_IF(3, 0, 1, 1)
    CGFDH_IF(4, 1, 0, 0) CGFDH_IF(5, 1, 0, 1) CGFDH_IF(6, 1, 1, 0) CGFDH_IF(7, 1, 1, 1)
  }
#undef CGFDH_IF
}

// the two-phase per-point kernel over whole rows: every point of the grid
// except the free-surface rows, PML faces included.  Tiles are 64 points of
// one row segment starting at the first physical column, so rows are read
// in full and the x-face PML columns share their cache lines with the
// neighbouring interior points in the same launch.  A block whose tile meets
// no PML face runs the interior form; a tile that meets one applies the
// face corrections (and their memory-variable RK combine) to its PML lanes
// after the common arithmetic.  The grouped block order spans whole rows.
template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__device__ __forceinline__ void rows_phased_tile(const StageParams& P, const BoxParams& B, int bx,
                                                 int by, int k)
{
  // does this tile meet a PML face (block-uniform)?
  const int ti0 = B.i0 + bx * 64, ti1 = min(ti0 + 63, B.i0 + B.ni - 1);
  const int tj0 = B.j0 + by * BY, tj1 = min(tj0 + BY - 1, B.j0 + B.nj - 1);
  bool tile_pml = false;
  for (int f = 0; f < P.npml; f++) {
    const PmlDev& F = P.pml[f];
    tile_pml |= ti1 >= F.i1 && ti0 <= F.i2 && tj1 >= F.j1 && tj0 <= F.j2 && k >= F.k1 &&
                k <= F.k2;
  }
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const long V = P.vol;
  constexpr int xl = Win<DX>::lo, yl = Win<DY>::lo, zl = Win<DZ>::lo;
  float ax[9][5], ay[9][5], az[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* v = P.wcur + c * V;
#pragma unroll
    for (int n = 0; n < 5; n++) {
      ax[c][n] = v[p + (xl + n)];
      ay[c][n] = v[p + (long)(yl + n) * P.nx];
      az[c][n] = v[p + (long)(zl + n) * P.ss];
    }
  }
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = ld_stream(P.met + (1 + n) * V + p);
  float mv[Medium<MT>::ncoef];
#pragma unroll
  for (int n = 0; n < Medium<MT>::ncoef; n++) mv[n] = ld_stream(P.med + n * V + p);
  const float slw = ld_stream(P.slw + p);
  // the elastic media's RK combine inputs with the other operands; the
  // visco-elastic medium combines all its fields after the arithmetic
  float cpre[9], cend[9];
  if constexpr (MT != MEDIUM_VISCO) combine_load<KIND>(P, p, cpre, cend);
  __builtin_amdgcn_sched_barrier(0);
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    X[c] = d5v<DX>(ax[c][0], ax[c][1], ax[c][2], ax[c][3], ax[c][4]);
    Y[c] = d5v<DY>(ay[c][0], ay[c][1], ay[c][2], ay[c][3], ay[c][4]);
    Z[c] = d5v<DZ>(az[c][0], az[c][1], az[c][2], az[c][3], az[c][4]);
  }
  Medium<MT> md;
  md.from(mv);
  float r[MT == MEDIUM_VISCO ? NW_MAX : 9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  if (tile_pml) {
    for (int f = 0; f < P.npml; f++) {
      const PmlDev& F = P.pml[f];
      if (i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2) {
        pml_face<KIND, MT>(P, F, f, i, j, k, p, X, Y, Z, md, slw, m, r);
      }
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

template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__global__ void __launch_bounds__(64 * BY)
stage_rows_phased_kernel(StageParams P, BoxParams B)
{
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  rows_phased_tile<DX, DY, DZ, KIND, MT, BY>(P, B, bx, by, k);
}

// the whole-row stage of several grids in one launch (the batch's
// cross-grid tile queue): blocks start[q] .. start[q+1]-1 (multiples of 8)
// belong to grid q, whose parameters and box sit in device memory; each
// grid keeps its own XCD-grouped block order
struct RowsMulti {
  const StageParams* P;
  const BoxParams* B;
  const int* start;   // n + 1
  int n;
};

template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__global__ void __launch_bounds__(64 * BY)
__attribute__((amdgpu_waves_per_eu(KIND == KIND_MID ? 2 : 3))) stage_rows_multi_kernel(RowsMulti R)
{
  // the tables are read through the constant address space, as kernel
  // arguments are: the kernel never writes them, so their fields load into
  // scalar registers once instead of being re-read after every store; the
  // grid index is wave-uniform
  typedef const __attribute__((address_space(4))) StageParams CStageParams;
  typedef const __attribute__((address_space(4))) BoxParams CBoxParams;
  typedef const __attribute__((address_space(4))) int CInt;
  CInt* start = (CInt*)R.start;
  int lo = 0, hi = R.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= (int)blockIdx.x) lo = mid;
    else hi = mid - 1;
  }
  lo = __builtin_amdgcn_readfirstlane(lo);
  const BoxParams& B = *(const BoxParams*)((CBoxParams*)R.B + lo);
  const int s0 = start[lo], s1 = start[lo + 1];
  int bx, by, k;
  if (!grouped_block_at(B, blockIdx.x - s0, s1 - s0, bx, by, k)) return;
  rows_phased_tile<DX, DY, DZ, KIND, MT, BY>(*(const StageParams*)((CStageParams*)R.P + lo), B,
                                             bx, by, k);
}

template <int KIND, int MT, int BY>
static void launch_rows_multi(hipStream_t st, const RowsMulti& R, int nblocks, int code)
{
  dim3 blk(64, BY, 1);
#define CGFDH_RM(c, DX, DY, DZ)                                                            \
  case c:                                                                                  \
    hipLaunchKernelGGL((stage_rows_multi_kernel<DX, DY, DZ, KIND, MT, BY>), dim3(nblocks),  \
                       blk, 0, st, R);                                                     \
    break;
  switch (code) {
    CGFDH_RM(0, 0, 0, 0) CGFDH_RM(1, 0, 0, 1) CGFDH_RM(2, 0, 1, 0) CGFDH_RM(3, 0, 1, 1)
    CGFDH_RM(4, 1, 0, 0) CGFDH_RM(5, 1, 0, 1) CGFDH_RM(6, 1, 1, 0) CGFDH_RM(7, 1, 1, 1)
  }
#undef CGFDH_RM
}

template <int KIND, int MT, int BY>
static void launch_rows_phased(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  // groups of whole rows: every tile of a row in one group
  const int tiles_x = (B.ni + 63) / 64;
  if (B.gx <= 0) B.gx = tiles_x;
  if (B.gy <= 0) B.gy = 2;
  const int grid = grouped_grid(B, BY);
  dim3 blk(64, BY, 1);
#define CGFDH_RF(c, DX, DY, DZ)                                                            \
  case c:                                                                                  \
    hipLaunchKernelGGL((stage_rows_phased_kernel<DX, DY, DZ, KIND, MT, BY>), dim3(grid),   \
                       blk, 0, st, P, B);                                                  \
    break;
  switch (code) {
    CGFDH_RF(0, 0, 0, 0) CGFDH_RF(1, 0, 0, 1) CGFDH_RF(2, 0, 1, 0) CGFDH_RF(3, 0, 1, 1)
    CGFDH_RF(4, 1, 0, 0) CGFDH_RF(5, 1, 0, 1) CGFDH_RF(6, 1, 1, 0) CGFDH_RF(7, 1, 1, 1)
  }
#undef CGFDH_RF
}

// the whole-row kernel with the operands that have no stencil -- the
// metric, the medium, the slowness and the RK combine inputs -- brought in
// by direct-to-LDS loads (global_load_lds_dword): each wave's lanes land
// in its own LDS slots (slot s of wave w holds the 64 lanes' values at
// lds[w][s][lane]), so the loads stay in flight without holding VGPRs and
// only the 135 stencil values occupy registers during the load phase.  The
// values are read back by the same lanes after the loads have completed
// (s_waitcnt vmcnt(0)); no barrier is needed.  Same arithmetic as
// stage_rows_phased_kernel.
template <int MT, int KIND> struct RowsLdsSlots {
  static constexpr int ncomb = KIND == KIND_MID ? 18 : 9;
  static constexpr int met = 0, med = 9, slw = 9 + Medium<MT>::ncoef;
  static constexpr int comb = slw + 1;
  static constexpr int n = comb + ncomb;
};

__device__ __forceinline__ void lds_dma_dword(const float* src, float* lds_wave_slot)
{
  __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                   (__attribute__((address_space(3))) void*)lds_wave_slot, 4, 0,
                                   0);
}

template <int DX, int DY, int DZ, int KIND, int MT, int BY>
__global__ void __launch_bounds__(64 * BY, 4)
stage_rows_lds_kernel(StageParams P, BoxParams B)
{
  static_assert(MT != MEDIUM_VISCO, "elastic media only");
  static_assert(KIND != KIND_RHS, "time-stepping stages only");
  using S = RowsLdsSlots<MT, KIND>;
  __shared__ float lds[BY][S::n][64];
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  const int ti0 = B.i0 + bx * 64, ti1 = min(ti0 + 63, B.i0 + B.ni - 1);
  const int tj0 = B.j0 + by * BY, tj1 = min(tj0 + BY - 1, B.j0 + B.nj - 1);
  bool tile_pml = false;
  for (int f = 0; f < P.npml; f++) {
    const PmlDev& F = P.pml[f];
    tile_pml |= ti1 >= F.i1 && ti0 <= F.i2 && tj1 >= F.j1 && tj0 <= F.j2 && k >= F.k1 &&
                k <= F.k2;
  }
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const long V = P.vol;
  float (*L)[64] = lds[threadIdx.y];
  // the operands without a stencil go straight to LDS
#pragma unroll
  for (int n = 0; n < 9; n++) lds_dma_dword(P.met + (1 + n) * V + p, L[S::met + n]);
#pragma unroll
  for (int n = 0; n < Medium<MT>::ncoef; n++) lds_dma_dword(P.med + n * V + p, L[S::med + n]);
  lds_dma_dword(P.slw + p, L[S::slw]);
#pragma unroll
  for (int c = 0; c < 9; c++) {
    if (KIND == KIND_FIRST || KIND == KIND_MID) lds_dma_dword(P.wpre + c * V + p, L[S::comb + c]);
    if (KIND == KIND_MID) lds_dma_dword(P.wend + c * V + p, L[S::comb + 9 + c]);
    if (KIND == KIND_LAST) lds_dma_dword(P.wend + c * V + p, L[S::comb + c]);
  }
  constexpr int xl = Win<DX>::lo, yl = Win<DY>::lo, zl = Win<DZ>::lo;
  float ax[9][5], ay[9][5], az[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* v = P.wcur + c * V;
#pragma unroll
    for (int n = 0; n < 5; n++) {
      ax[c][n] = v[p + (xl + n)];
      ay[c][n] = v[p + (long)(yl + n) * P.nx];
      az[c][n] = v[p + (long)(zl + n) * P.ss];
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    X[c] = d5v<DX>(ax[c][0], ax[c][1], ax[c][2], ax[c][3], ax[c][4]);
    Y[c] = d5v<DY>(ay[c][0], ay[c][1], ay[c][2], ay[c][3], ay[c][4]);
    Z[c] = d5v<DZ>(az[c][0], az[c][1], az[c][2], az[c][3], az[c][4]);
  }
  // every load of this lane has landed (the stencil values are in use)
  __builtin_amdgcn_s_waitcnt(0);
  __builtin_amdgcn_sched_barrier(0);
  const int t = threadIdx.x;
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = L[S::met + n][t];
  float mv[Medium<MT>::ncoef];
#pragma unroll
  for (int n = 0; n < Medium<MT>::ncoef; n++) mv[n] = L[S::med + n][t];
  const float slw = L[S::slw][t];
  Medium<MT> md;
  md.from(mv);
  float r[9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  if (tile_pml) {
    for (int f = 0; f < P.npml; f++) {
      const PmlDev& F = P.pml[f];
      if (i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2) {
        pml_face<KIND, MT>(P, F, f, i, j, k, p, X, Y, Z, md, slw, m, r);
      }
    }
  }
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  float cpre[9], cend[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    cpre[c] = (KIND == KIND_FIRST || KIND == KIND_MID) ? L[S::comb + c][t] : 0.0f;
    cend[c] = KIND == KIND_MID ? L[S::comb + 9 + c][t] : (KIND == KIND_LAST ? L[S::comb + c][t]
                                                                              : 0.0f);
  }
  combine_store<KIND>(P, p, r, cpre, cend);
}

template <int KIND, int MT, int BY>
static void launch_rows_lds(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  const int tiles_x = (B.ni + 63) / 64;
  if (B.gx <= 0) B.gx = tiles_x;
  if (B.gy <= 0) B.gy = 2;
  const int grid = grouped_grid(B, BY);
  dim3 blk(64, BY, 1);
#define CGFDH_RL(c, DX, DY, DZ)                                                          \
  case c:                                                                                \
    hipLaunchKernelGGL((stage_rows_lds_kernel<DX, DY, DZ, KIND, MT, BY>), dim3(grid), blk, \
                       0, st, P, B);                                                     \
    break;
  switch (code) {
    CGFDH_RL(0, 0, 0, 0) CGFDH_RL(1, 0, 0, 1) CGFDH_RL(2, 0, 1, 0) CGFDH_RL(3, 0, 1, 1)
    CGFDH_RL(4, 1, 0, 0) CGFDH_RL(5, 1, 0, 1) CGFDH_RL(6, 1, 1, 0) CGFDH_RL(7, 1, 1, 1)
  }
#undef CGFDH_RL
}

// the LDS-tiled plane kernel in phases for the elastic media: the tile loads
// first, then every other operand of the point (z stencil, metric, medium,
// RK combine inputs); the LDS stores wait only for the tile loads, so the
// other requests stay in flight across the barrier.
template <int DX, int DY, int DZ, int KIND, int MT, int TY>
__global__ void __launch_bounds__(64 * TY)
stage_interior_ldsp_kernel(StageParams P, BoxParams B)
{
  static_assert(MT != MEDIUM_VISCO, "elastic media only");
  using T = LdsTile<TY>;
  constexpr int NE = (T::PLANE + 64 * TY - 1) / (64 * TY);   // tile entries per thread
  __shared__ float sm[9 * T::PLANE];
  int bx, by, k;
  // block-uniform exit: every thread of a block takes the same branch
  if (!grouped_block(B, bx, by, k)) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = B.i0 + bx * 64, j0 = B.j0 + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int nx = P.nx;
  const int ny = (int)(P.ss / nx);
  const long V = P.vol, ss = P.ss;
  const long pk = (long)k * ss;
  constexpr int xl = -Win<DX>::lo;
  constexpr int yl = -Win<DY>::lo;
  constexpr int zl = Win<DZ>::lo;
  const bool valid = i < B.i0 + B.ni && j < B.j0 + B.nj;
  // out-of-box threads read the box's first point instead (in bounds, unused)
  const long p = valid ? (long)i + (long)j * nx + pk : (long)B.i0 + (long)B.j0 * nx + pk;
  float tv[NE][9];
#pragma unroll
  for (int q = 0; q < NE; q++) {
    const int e = ty * 64 + tx + q * 64 * TY;
    const int r = e / T::W, col = e - r * T::W;
    const int gi = i0 - xl + col, gj = j0 - yl + r;
    const bool ok = e < T::PLANE && gi < nx && gj < ny;
    const long g = ok ? (long)gi + (long)gj * nx + pk : p;
#pragma unroll
    for (int c = 0; c < 9; c++) tv[q][c] = P.wcur[c * V + g];
    if (!ok) {
#pragma unroll
      for (int c = 0; c < 9; c++) tv[q][c] = 0.0f;
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  float az[9][5];
#pragma unroll
  for (int c = 0; c < 9; c++) {
#pragma unroll
    for (int n = 0; n < 5; n++) az[c][n] = P.wcur[c * V + p + (long)(zl + n) * ss];
  }
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = ld_stream(P.met + (1 + n) * V + p);
  float mv[Medium<MT>::ncoef];
#pragma unroll
  for (int n = 0; n < Medium<MT>::ncoef; n++) mv[n] = ld_stream(P.med + n * V + p);
  const float slw = ld_stream(P.slw + p);
  float cpre[9], cend[9];
  combine_load<KIND>(P, p, cpre, cend);
  __builtin_amdgcn_sched_barrier(0);
#pragma unroll
  for (int q = 0; q < NE; q++) {
    const int e = ty * 64 + tx + q * 64 * TY;
    if (e < T::PLANE) {
#pragma unroll
      for (int c = 0; c < 9; c++) sm[c * T::PLANE + e] = tv[q][c];
    }
  }
  __syncthreads();
  if (!valid) return;
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* s = sm + c * T::PLANE;
    const int rowb = (ty + yl) * T::W + tx;
    X[c] = d5v<DX>(s[rowb], s[rowb + 1], s[rowb + 2], s[rowb + 3], s[rowb + 4]);
    const int colb = ty * T::W + tx + xl;
    Y[c] = d5v<DY>(s[colb], s[colb + T::W], s[colb + 2 * T::W], s[colb + 3 * T::W],
                   s[colb + 4 * T::W]);
    Z[c] = d5v<DZ>(az[c][0], az[c][1], az[c][2], az[c][3], az[c][4]);
  }
  Medium<MT> md;
  md.from(mv);
  float r[9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  combine_store<KIND>(P, p, r, cpre, cend);
}

template <int KIND, int MT, int TY>
static void launch_interior_ldsp(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  const int grid = grouped_grid(B, TY);
  dim3 blk(64, TY, 1);
#define CGFDH_IP2(c, DX, DY, DZ)                                                            \
  case c:                                                                                   \
    hipLaunchKernelGGL((stage_interior_ldsp_kernel<DX, DY, DZ, KIND, MT, TY>), dim3(grid), blk, \
                       0, st, P, B);                                                        \
    break;
  switch (code) {
    CGFDH_IP2(0, 0, 0, 0) CGFDH_IP2(1, 0, 0, 1) CGFDH_IP2(2, 0, 1, 0) CGFDH_IP2(3, 0, 1, 1)
    CGFDH_IP2(4, 1, 0, 0) CGFDH_IP2(5, 1, 0, 1) CGFDH_IP2(6, 1, 1, 0) CGFDH_IP2(7, 1, 1, 1)
  }
#undef CGFDH_IP2
}

// the per-point kernel marching KZ consecutive planes per block, so that the
// z neighbours one plane reads are read again by the same block moments
// later (L1/L2 hits) instead of by blocks far apart in time.  The box's nk
// holds the number of KZ-plane chunks and nchunk the true plane count.
template <int DX, int DY, int DZ, int KIND, int MT, int BY, int KZ>
__global__ void __launch_bounds__(64 * BY)
stage_interior_pointz_kernel(StageParams P, BoxParams B)
{
  int bx, by, kc;
  if (!grouped_block(B, bx, by, kc)) return;
  const int i = B.i0 + bx * 64 + threadIdx.x;
  const int j = B.j0 + by * BY + threadIdx.y;
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const int kbeg = B.k0 + (kc - B.k0) * KZ;
  const int kend = min(B.k0 + B.nchunk, kbeg + KZ);
  const long V = P.vol;
  for (int k = kbeg; k < kend; k++) {
    const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
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
}

template <int KIND, int MT, int BY, int KZ>
static void launch_interior_pointz(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  B.nchunk = B.nk;
  B.nk = (B.nk + KZ - 1) / KZ;
  const int grid = grouped_grid(B, BY);
  dim3 blk(64, BY, 1);
#define CGFDH_IZ(c, DX, DY, DZ)                                                               \
  case c:                                                                                     \
    hipLaunchKernelGGL((stage_interior_pointz_kernel<DX, DY, DZ, KIND, MT, BY, KZ>), dim3(grid), \
                       blk, 0, st, P, B);                                                     \
    break;
  switch (code) {
    CGFDH_IZ(0, 0, 0, 0) CGFDH_IZ(1, 0, 0, 1) CGFDH_IZ(2, 0, 1, 0) CGFDH_IZ(3, 0, 1, 1)
    CGFDH_IZ(4, 1, 0, 0) CGFDH_IZ(5, 1, 0, 1) CGFDH_IZ(6, 1, 1, 0) CGFDH_IZ(7, 1, 1, 1)
  }
#undef CGFDH_IZ
}

// interior box, one plane of a 64 x TY tile per block: the x and y stencil
// operands of all 9 fields come from an LDS copy of the tile plus its 4-wide
// x/y halo (each value loaded once per block instead of once per reader),
// the z operands straight from global memory, issued before the barrier.
// Same grouped block order as the per-point kernel.
template <int DX, int DY, int DZ, int KIND, int MT, int TY>
__global__ void __launch_bounds__(64 * TY)
stage_interior_lds_kernel(StageParams P, BoxParams B)
{
  using T = LdsTile<TY>;
  __shared__ float sm[9 * T::PLANE];
  int bx, by, k;
  // block-uniform exit: every thread of a block takes the same branch
  if (!grouped_block(B, bx, by, k)) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int i0 = B.i0 + bx * 64, j0 = B.j0 + by * TY;
  const int i = i0 + tx, j = j0 + ty;
  const int nx = P.nx;
  const int ny = (int)(P.ss / nx);
  const long V = P.vol, ss = P.ss;
  const long pk = (long)k * ss;
  constexpr int xl = -Win<DX>::lo;
  constexpr int yl = -Win<DY>::lo;
  const bool valid = i < B.i0 + B.ni && j < B.j0 + B.nj;
  const long p = (long)i + (long)j * nx + pk;
  float Z[9];
  if (valid) {
#pragma unroll
    for (int c = 0; c < 9; c++) Z[c] = d5<DZ>(P.wcur + c * V, p, ss);
  }
  // tile entry e = row * W + col holds point (i0 - xl + col, j0 - yl + row);
  // the interior box lies >= 3 points inside the allocation, so only the
  // right/back ends of a partial tile can leave a row (those read zeros and
  // feed no valid point)
  for (int e = ty * 64 + tx; e < T::PLANE; e += 64 * TY) {
    const int r = e / T::W, col = e - r * T::W;
    const int gi = i0 - xl + col, gj = j0 - yl + r;
    const bool ok = gi < nx && gj < ny;
    const long g = (long)gi + (long)gj * nx + pk;
#pragma unroll
    for (int c = 0; c < 9; c++) sm[c * T::PLANE + e] = ok ? P.wcur[c * V + g] : 0.0f;
  }
  __syncthreads();
  if (!valid) return;
  float X[9], Y[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* s = sm + c * T::PLANE;
    const int rowb = (ty + yl) * T::W + tx;
    X[c] = d5v<DX>(s[rowb], s[rowb + 1], s[rowb + 2], s[rowb + 3], s[rowb + 4]);
    const int colb = ty * T::W + tx + xl;
    Y[c] = d5v<DY>(s[colb], s[colb + T::W], s[colb + 2 * T::W], s[colb + 3 * T::W],
                   s[colb + 4 * T::W]);
  }
  interior_finish<KIND, MT>(P, i, j, k, p, X, Y, Z);
}

template <int KIND, int MT, int TY>
static void launch_interior_lds(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  const int grid = grouped_grid(B, TY);
  dim3 blk(64, TY, 1);
#define CGFDH_IL(c, DX, DY, DZ)                                                          \
  case c:                                                                                \
    hipLaunchKernelGGL((stage_interior_lds_kernel<DX, DY, DZ, KIND, MT, TY>), dim3(grid), blk, \
                       0, st, P, B);                                                     \
    break;
  switch (code) {
    CGFDH_IL(0, 0, 0, 0) CGFDH_IL(1, 0, 0, 1) CGFDH_IL(2, 0, 1, 0) CGFDH_IL(3, 0, 1, 1)
    CGFDH_IL(4, 1, 0, 0) CGFDH_IL(5, 1, 0, 1) CGFDH_IL(6, 1, 1, 0) CGFDH_IL(7, 1, 1, 1)
  }
#undef CGFDH_IL
}

// general per-point kernel restricted to a box (the boundary shell); FLAGS
// says whether the box touches the free-surface rows.  256-thread blocks over
// tiles tw points wide and 256/tw rows deep (tw = the box width when that is
// under 64, so the thin PML slabs along x keep whole waves busy), in the
// grouped block order of the interior kernel.
template <int DX, int DY, int DZ, int KIND, int MT, int FLAGS>
__global__ void __launch_bounds__(256) stage_box_kernel(StageParams P, BoxParams B)
{
  int bx, by, k;
  if (!grouped_block(B, bx, by, k)) return;
  const int rows = 256 / B.tw;
  const int t = threadIdx.x;
  const int ty = t / B.tw;
  if (ty >= rows) return;
  const int i = B.i0 + bx * B.tw + (t - ty * B.tw);
  const int j = B.j0 + by * rows + ty;
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

// several shell boxes of one FLAGS class in one launch: block ranges
// start[q] .. start[q+1]-1 (multiples of 8) belong to box q
constexpr int SHELL_MAX = 6;
struct ShellBoxes {
  BoxParams b[SHELL_MAX];
  int start[SHELL_MAX + 1];
  int n;
};

template <int DX, int DY, int DZ, int KIND, int MT, int FLAGS>
__global__ void __launch_bounds__(256) stage_shell_kernel(StageParams P, ShellBoxes S)
{
  int q = 0;
  while (q + 1 < S.n && (int)blockIdx.x >= S.start[q + 1]) ++q;
  const BoxParams& B = S.b[q];
  int bx, by, k;
  if (!grouped_block_at(B, blockIdx.x - S.start[q], S.start[q + 1] - S.start[q], bx, by, k))
    return;
  const int rows = 256 / B.tw;
  const int t = threadIdx.x;
  const int ty = t / B.tw;
  if (ty >= rows) return;
  const int i = B.i0 + bx * B.tw + (t - ty * B.tw);
  const int j = B.j0 + by * rows + ty;
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

// one point of the phased shell kernel: for a point inside the box's own
// PML face (IDIM its normal, -1 none) that face's memory variables are
// requested first, so their HBM latency overlaps the stencil that follows;
// then the stencil, metric and medium (loads and arithmetic interleaved by
// the compiler), the PML corrections in face order (the box's own face from
// the values already loaded, any other face holding the point with its own
// loads), sources and the RK combine.  No free-surface rows.
template <int DX, int DY, int DZ, int KIND, int MT, int IDIM>
__device__ __forceinline__ void shellp_point(const StageParams& P, int fslot, int i, int j, int k,
                                             long p)
{
  const long V = P.vol;
  float a[9], apre[9], aend[9];
  long ia = 0;
  if constexpr (IDIM >= 0) {
    const PmlDev& F = P.pml[fslot];
    if (i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2) {
      ia = pml_aux_index(F, i, j, k);
      pml_aux_load<KIND>(F, ia, a, apre, aend);
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  float X[9], Y[9], Z[9];
#pragma unroll
  for (int c = 0; c < 9; c++) {
    const float* v = P.wcur + c * V;
    X[c] = d5<DX>(v, p, 1);
    Y[c] = d5<DY>(v, p, P.nx);
    Z[c] = d5<DZ>(v, p, P.ss);
  }
  float m[9];
#pragma unroll
  for (int n = 0; n < 9; n++) m[n] = ld_stream(P.met + (1 + n) * V + p);
  Medium<MT> md;
  md.load(P.med, V, p);
  const float slw = ld_stream(P.slw + p);
  float cpre[9], cend[9];
  combine_load<KIND>(P, p, cpre, cend);
  float r[9];
  momentum(r, m, slw, X, Y, Z);
  float Xv[3] = {X[VX], X[VY], X[VZ]}, Yv[3] = {Y[VX], Y[VY], Y[VZ]},
        Zv[3] = {Z[VX], Z[VY], Z[VZ]};
  md.hooke(r, m, Xv, Yv, Zv);
  for (int f = 0; f < P.npml; f++) {
    const PmlDev& F = P.pml[f];
    if (!(i >= F.i1 && i <= F.i2 && j >= F.j1 && j <= F.j2 && k >= F.k1 && k <= F.k2)) continue;
    if (IDIM >= 0 && f == fslot) {
      if constexpr (IDIM >= 0)
        pml_face_apply<KIND, MT, (IDIM < 0 ? 0 : IDIM)>(
            P, F, f, i, j, k, ia, IDIM == 0 ? X : (IDIM == 1 ? Y : Z), md, slw, m, r, a, apre,
            aend);
    } else {
      pml_face<KIND, MT>(P, F, f, i, j, k, p, X, Y, Z, md, slw, m, r);
    }
  }
  if (P.nsrc > 0) add_sources(P, i, j, k, p, r);
  add_dd_for<MT>(P, p, r);
  combine_store<KIND>(P, p, r, cpre, cend);
}

// the PML shell (no free-surface rows) of the elastic media in one launch:
// boxes of 64 x 4 tiles in the grouped block order, each naming its own PML
// face; block ranges start[q] .. start[q+1]-1 belong to box q
template <int DX, int DY, int DZ, int KIND, int MT>
__global__ void __launch_bounds__(256) stage_shellp_kernel(StageParams P, ShellBoxes S)
{
  int q = 0;
  while (q + 1 < S.n && (int)blockIdx.x >= S.start[q + 1]) ++q;
  const BoxParams& B = S.b[q];
  int bx, by, k;
  if (!grouped_block_at(B, blockIdx.x - S.start[q], S.start[q + 1] - S.start[q], bx, by, k))
    return;
  const int i = B.i0 + bx * 64 + (int)(threadIdx.x & 63);
  const int j = B.j0 + by * 4 + (int)(threadIdx.x >> 6);
  if (i >= B.i0 + B.ni || j >= B.j0 + B.nj) return;
  const long p = (long)i + (long)j * P.nx + (long)k * P.ss;
  const int fs = B.fslot;
  const int idim = fs >= 0 ? P.pml[fs].idim : -1;
  if (idim == 0)
    shellp_point<DX, DY, DZ, KIND, MT, 0>(P, fs, i, j, k, p);
  else if (idim == 1)
    shellp_point<DX, DY, DZ, KIND, MT, 1>(P, fs, i, j, k, p);
  else if (idim == 2)
    shellp_point<DX, DY, DZ, KIND, MT, 2>(P, fs, i, j, k, p);
  else
    shellp_point<DX, DY, DZ, KIND, MT, -1>(P, fs, i, j, k, p);
}

template <int KIND, int MT>
static void launch_shellp(hipStream_t st, const StageParams& P, const BoxParams* boxes, int n,
                          int code)
{
  ShellBoxes S;
  memset(&S, 0, sizeof S);
  int total = 0;
  for (int q = 0; q < n && S.n < SHELL_MAX; q++) {
    BoxParams B = boxes[q];
    if (B.ni <= 0 || B.nj <= 0 || B.nk <= 0) continue;
    // groups of whole box rows (at most 16 tiles wide), two tile rows deep
    const int tiles_x = (B.ni + 63) / 64;
    if (B.gx <= 0) B.gx = std::min(tiles_x, 16);
    if (B.gy <= 0) B.gy = 2;
    const int g = grouped_grid(B, 4, 64);
    S.b[S.n] = B;
    S.start[S.n] = total;
    total += g;
    S.n++;
  }
  if (!S.n) return;
  S.start[S.n] = total;
  dim3 blk(256, 1, 1);
#define CGFDH_SP(c, DX, DY, DZ)                                                                \
  case c:                                                                                      \
    hipLaunchKernelGGL((stage_shellp_kernel<DX, DY, DZ, KIND, MT>), dim3(total), blk, 0, st, P, \
                       S);                                                                     \
    break;
  switch (code) {
    CGFDH_SP(0, 0, 0, 0) CGFDH_SP(1, 0, 0, 1) CGFDH_SP(2, 0, 1, 0) CGFDH_SP(3, 0, 1, 1)
    CGFDH_SP(4, 1, 0, 0) CGFDH_SP(5, 1, 0, 1) CGFDH_SP(6, 1, 1, 0) CGFDH_SP(7, 1, 1, 1)
  }
#undef CGFDH_SP
}

template <int KIND, int MT, int FLAGS>
static void launch_shell(hipStream_t st, const StageParams& P, const BoxParams* boxes, int n,
                         int code)
{
  ShellBoxes S;
  memset(&S, 0, sizeof S);
  int total = 0;
  for (int q = 0; q < n && S.n < SHELL_MAX; q++) {
    BoxParams B = boxes[q];
    if (B.ni <= 0 || B.nj <= 0 || B.nk <= 0) continue;
    B.tw = std::min(64, B.ni);
    const int g = grouped_grid(B, 256 / B.tw, B.tw);
    S.b[S.n] = B;
    S.start[S.n] = total;
    total += g;
    S.n++;
  }
  if (!S.n) return;
  S.start[S.n] = total;
  dim3 blk(256, 1, 1);
#define CGFDH_SH(c, DX, DY, DZ)                                                                  \
  case c:                                                                                        \
    hipLaunchKernelGGL((stage_shell_kernel<DX, DY, DZ, KIND, MT, FLAGS>), dim3(total), blk, 0, st, \
                       P, S);                                                                    \
    break;
  switch (code) {
    CGFDH_SH(0, 0, 0, 0) CGFDH_SH(1, 0, 0, 1) CGFDH_SH(2, 0, 1, 0) CGFDH_SH(3, 0, 1, 1)
    CGFDH_SH(4, 1, 0, 0) CGFDH_SH(5, 1, 0, 1) CGFDH_SH(6, 1, 1, 0) CGFDH_SH(7, 1, 1, 1)
  }
#undef CGFDH_SH
}

template <int KIND, int MT, int FLAGS>
static void launch_box(hipStream_t st, const StageParams& P, BoxParams B, int code)
{
  if (B.ni <= 0 || B.nj <= 0 || B.nk <= 0) return;
  B.tw = std::min(64, B.ni);
  dim3 blk(256, 1, 1);
  dim3 grd(grouped_grid(B, 256 / B.tw, B.tw), 1, 1);
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
