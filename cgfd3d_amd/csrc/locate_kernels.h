// locate_kernels.h -- point location on the curvilinear grid on the device:
// the cell search of gd_curv_coord_to_local_indx (gd_t.c:1998-2173) with the
// point-in-cell test of isPointInHexahedron (media/media_geometry3d.cpp:
// 150-267), the sampled shift of gd_curv_coord2shift_sample (gd_t.c:
// 2304-2378) and the surface search of gd_curv_depth_to_axis (gd_t.c:
// 2179-2296).  The reference scans tiles and cells in order and stops at the
// first hit; here every cell of a candidate tile is tested in parallel and
// the hit with the smallest scan-order key wins (atomicMin), which is the
// same cell.  Same float/double arithmetic as the reference, so the located
// indices and shifts are bit-identical to it.
#pragma once

namespace cgfdh {

constexpr int LOC_TILES = 4;    // GD_TILE_NX / _NY / _NZ (gd_t.h:9-11)
constexpr int LOC_SAMPLES = 100;  // samples per cell edge (gd_t.c:2133-2137)

struct LocGrid {
  const float* x;
  const float* y;
  const float* z;
  int nx, ny;
  long ss;
  int ts[3][LOC_TILES], te[3][LOC_TILES];   // tile ranges per axis (i, j, k)
};

struct lv3 { float x, y, z; };

__device__ __forceinline__ lv3 lsub(lv3 a, lv3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ float ldot(lv3 a, lv3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ lv3 lcross(lv3 a, lv3 b)
{
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
// the length as the geometry code takes it: sqrt of the float dot product in
// double, rounded to float
__device__ __forceinline__ float llen(lv3 a) { return (float)sqrt((double)ldot(a, a)); }
__device__ __forceinline__ float larea(lv3 a, lv3 b, lv3 c)
{
  return llen(lcross(lsub(b, a), lsub(c, a)));
}

// the ten tetrahedra of the cell (two five-way cuts), four faces each as
// vertex triples whose edge cross products point inwards
__constant__ unsigned char kLocTet[10][4][3] = {
  {{0, 4, 2}, {2, 4, 1}, {1, 4, 0}, {0, 2, 1}},
  {{4, 6, 2}, {2, 6, 7}, {7, 6, 4}, {4, 2, 7}},
  {{2, 7, 3}, {3, 7, 1}, {1, 2, 3}, {1, 7, 2}},
  {{7, 4, 5}, {1, 7, 5}, {5, 4, 1}, {1, 4, 7}},
  {{2, 4, 7}, {7, 4, 1}, {2, 7, 1}, {4, 2, 1}},
  {{4, 5, 6}, {4, 6, 0}, {4, 0, 5}, {6, 5, 0}},
  {{6, 2, 0}, {6, 3, 2}, {0, 2, 3}, {3, 6, 0}},
  {{6, 5, 7}, {3, 6, 7}, {7, 5, 3}, {5, 6, 3}},
  {{5, 0, 1}, {3, 5, 1}, {0, 3, 1}, {0, 5, 3}},
  {{5, 6, 0}, {5, 0, 3}, {0, 6, 3}, {3, 6, 5}},
};

// a point is in the cell when one tetrahedron holds it: on a face triangle
// (its three sub-areas sum to the whole within 1e-6), or on the inner side
// of all four faces within -1e-6 of the normalised distance
__device__ bool point_in_hexa(lv3 p, const lv3* v)
{
  for (int t = 0; t < 10; t++) {
    bool in = true;
    for (int n = 0; n < 4 && in; n++) {
      const lv3 a = v[kLocTet[t][n][0]], b = v[kLocTet[t][n][1]], c = v[kLocTet[t][n][2]];
      const float s1 = larea(p, a, b), s2 = larea(p, b, c), s3 = larea(p, c, a);
      const float d = s1 + s2 + s3 - larea(a, b, c);
      if (fabs((double)d) < 1e-6) break;   // on this face: in the tetrahedron
      lv3 nrm = lcross(lsub(b, a), lsub(c, a));
      const float l = llen(nrm);
      nrm.x = nrm.x / l; nrm.y = nrm.y / l; nrm.z = nrm.z / l;
      const lv3 p2f = lsub(a, p);
      float s = ldot(p2f, nrm);
      s /= llen(p2f);
      if (s <= -1.0e-6f) in = false;
    }
    if (in) return true;
  }
  return false;
}

__device__ __forceinline__ void cell_corners(const LocGrid& G, int i, int j, int k, lv3* v)
{
#pragma unroll
  for (int n = 0; n < 8; n++) {
    const long q = (long)(i + (n & 1)) + (long)(j + ((n >> 1) & 1)) * G.nx +
                   (long)(k + (n >> 2)) * G.ss;
    v[n] = {G.x[q], G.y[q], G.z[q]};
  }
}

// tile boxes: the coordinate range of every point a tile's cells touch (its
// point ranges plus one), which is the union of its cells' boxes
// (gd_curv_set_minmax, gd_t.c:1808-1873).  One block per tile; box[t*6 ..]
// = xmin, xmax, ymin, ymax, zmin, zmax.
__global__ void __launch_bounds__(256) tile_box_kernel(LocGrid G, float* box)
{
  const int t = blockIdx.x;
  const int ti = t % LOC_TILES, tj = (t / LOC_TILES) % LOC_TILES;
  const int tk = t / (LOC_TILES * LOC_TILES);
  const int i0 = G.ts[0][ti], j0 = G.ts[1][tj], k0 = G.ts[2][tk];
  const int ni = G.te[0][ti] - i0 + 2, nj = G.te[1][tj] - j0 + 2, nk = G.te[2][tk] - k0 + 2;
  float lo[3] = {1.0e26f, 1.0e26f, 1.0e26f}, hi[3] = {-1.0e26f, -1.0e26f, -1.0e26f};
  const long n = (ni > 1 && nj > 1 && nk > 1) ? (long)ni * nj * nk : 0;
  for (long q = threadIdx.x; q < n; q += 256) {
    const int i = i0 + (int)(q % ni), j = j0 + (int)(q / ni % nj), k = k0 + (int)(q / ni / nj);
    const long p = (long)i + (long)j * G.nx + (long)k * G.ss;
    const float v[3] = {G.x[p], G.y[p], G.z[p]};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      lo[d] = fminf(lo[d], v[d]);
      hi[d] = fmaxf(hi[d], v[d]);
    }
  }
  __shared__ float red[6][256];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    red[2 * d][threadIdx.x] = lo[d];
    red[2 * d + 1][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
#pragma unroll
      for (int d = 0; d < 3; d++) {
        red[2 * d][threadIdx.x] = fminf(red[2 * d][threadIdx.x], red[2 * d][threadIdx.x + h]);
        red[2 * d + 1][threadIdx.x] =
            fmaxf(red[2 * d + 1][threadIdx.x], red[2 * d + 1][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) box[t * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// cell search over (point, tile) pairs: every cell of the tile whose box
// holds the point and which holds it by the tetrahedra test offers its
// scan-order key (tile number, then cell order within the tile)
__global__ void __launch_bounds__(256) cell_search_kernel(LocGrid G, const float* pts,
                                                         const int* pairs,
                                                         unsigned long long* key)
{
  const int q = pairs[2 * blockIdx.y], t = pairs[2 * blockIdx.y + 1];
  const int ti = t % LOC_TILES, tj = (t / LOC_TILES) % LOC_TILES, tk = t / (LOC_TILES * LOC_TILES);
  const int i0 = G.ts[0][ti], j0 = G.ts[1][tj], k0 = G.ts[2][tk];
  const int ni = G.te[0][ti] - i0 + 1, nj = G.te[1][tj] - j0 + 1, nk = G.te[2][tk] - k0 + 1;
  const long n = (long)ni * nj * nk;
  const lv3 p = {pts[3 * q], pts[3 * q + 1], pts[3 * q + 2]};
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long)gridDim.x * 256) {
    const int i = i0 + (int)(c % ni), j = j0 + (int)(c / ni % nj), k = k0 + (int)(c / ni / nj);
    lv3 v[8];
    cell_corners(G, i, j, k, v);
    // the cell's box (gd_curv_set_minmax) first, then the tetrahedra
    float lo[3] = {v[0].x, v[0].y, v[0].z}, hi[3] = {v[0].x, v[0].y, v[0].z};
#pragma unroll
    for (int m = 1; m < 8; m++) {
      lo[0] = fminf(lo[0], v[m].x); hi[0] = fmaxf(hi[0], v[m].x);
      lo[1] = fminf(lo[1], v[m].y); hi[1] = fmaxf(hi[1], v[m].y);
      lo[2] = fminf(lo[2], v[m].z); hi[2] = fmaxf(hi[2], v[m].z);
    }
    const bool out = p.x < lo[0] || p.x > hi[0] || p.y < lo[1] || p.y > hi[1] ||
                     p.z < lo[2] || p.z > hi[2];
    if (out || !point_in_hexa(p, v)) continue;
    atomicMin(&key[q], ((unsigned long long)t << 40) | (unsigned long long)c);
  }
}

// the nearest of the 101^3 trilinear samples of the located cell: key =
// distance bits, then sample index (n3, n2, n1 order), so the minimum is the
// reference's first minimum; its start, the distance to corner 0 with index
// 0, is sample 0's own value
__global__ void __launch_bounds__(256) sample_shift_kernel(LocGrid G, const float* pts,
                                                          const int* cell,
                                                          unsigned long long* key)
{
  const int q = blockIdx.x;
  const int i = cell[3 * q], j = cell[3 * q + 1], k = cell[3 * q + 2];
  if (i < 0) return;
  lv3 v[8];
  cell_corners(G, i, j, k, v);
  const float sx = pts[3 * q], sy = pts[3 * q + 1], sz = pts[3 * q + 2];
  constexpr int NS = LOC_SAMPLES + 1;
  unsigned long long best = ~0ull;
  for (int s = threadIdx.x; s < NS * NS * NS; s += 256) {
    const int n1 = s % NS, n2 = s / NS % NS, n3 = s / (NS * NS);
    float Lx[2], Ly[2], Lz[2];
    Lz[1] = (float)n3 / (float)LOC_SAMPLES;
    Lz[0] = (float)(1.0 - (double)Lz[1]);
    Ly[1] = (float)n2 / (float)LOC_SAMPLES;
    Ly[0] = (float)(1.0 - (double)Ly[1]);
    Lx[1] = (float)n1 / (float)LOC_SAMPLES;
    Lx[0] = (float)(1.0 - (double)Lx[1]);
    float x = 0.0f, y = 0.0f, z = 0.0f;
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const float w = Lx[c & 1] * Ly[(c >> 1) & 1] * Lz[c >> 2];
      x += w * v[c].x;
      y += w * v[c].y;
      z += w * v[c].z;
    }
    const float d = sqrtf((sx - x) * (sx - x) + (sy - y) * (sy - y) + (sz - z) * (sz - z));
    const unsigned long long kk =
        ((unsigned long long)__float_as_uint(d) << 32) | (unsigned long long)s;
    best = kk < best ? kk : best;
  }
  __shared__ unsigned long long red[256];
  red[threadIdx.x] = best;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h && red[threadIdx.x + h] < red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0) key[q] = red[0];
}

// fdlib_math_isPoint2InTri (fdlib_math.c:246-266): on the left of (or on)
// all three edges, each edge's cross product divided by its length in double
__device__ __forceinline__ bool in_tri2(float px, float py, const float* vx, const float* vy)
{
  for (int i = 0; i < 3; i++) {
    const float ex = i == 2 ? (vx[0] - vx[2]) : (vx[i + 1] - vx[i]);
    const float ey = i == 2 ? (vy[0] - vy[2]) : (vy[i + 1] - vy[i]);
    const float qx = px - vx[i], qy = py - vy[i];
    float s = ex * qy - ey * qx;
    s = (float)((double)s / sqrt((double)(ex * ex + ey * ey)));
    if (s < 0) return false;
  }
  return true;
}

// fdlib_math_computeZfromTri (fdlib_math.c:307-345)
__device__ __forceinline__ float z_on_tri(float x, float y, float x1, float y1, float z1,
                                          float x2, float y2, float z2, float x3, float y3,
                                          float z3)
{
  const float ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
  const float bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
  const float a = ay * bz - az * by;
  const float b = az * bx - ax * bz;
  const float c = ax * by - ay * bx;
  const float d = -(a * x1 + b * y1 + c * z1);
  return -(a * x + b * y + d) / c;
}

// which of the surface cell's two triangles holds (x, y): 2 for 3-2-0, else
// 1 for 0-1-3, else 0 (fdlib_math_Point2InWhichTri, fdlib_math.c:285-304);
// z of the surface there
__device__ int surface_tri(const LocGrid& G, int i, int j, int k, float x, float y, float* zt)
{
  float qx[4], qy[4], qz[4];
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const long q = (long)(i + (n & 1)) + (long)(j + (n >> 1)) * G.nx + (long)k * G.ss;
    qx[n] = G.x[q]; qy[n] = G.y[q]; qz[n] = G.z[q];
  }
  const float tx[5] = {qx[0], qx[1], qx[3], qx[2], qx[0]};
  const float ty[5] = {qy[0], qy[1], qy[3], qy[2], qy[0]};
  if (in_tri2(x, y, tx + 2, ty + 2)) {
    *zt = z_on_tri(x, y, qx[3], qy[3], qz[3], qx[2], qy[2], qz[2], qx[0], qy[0], qz[0]);
    return 2;
  }
  if (in_tri2(x, y, tx, ty)) {
    *zt = z_on_tri(x, y, qx[0], qy[0], qz[0], qx[1], qy[1], qz[1], qx[3], qy[3], qz[3]);
    return 1;
  }
  return 0;
}

// the surface search over (point, top tile) pairs: surface cells (k = nk2)
// whose x-y box holds the point and one of whose triangles does
__global__ void __launch_bounds__(256) surface_search_kernel(LocGrid G, const float* pts,
                                                            const int* pairs,
                                                            unsigned long long* key)
{
  const int q = pairs[2 * blockIdx.y], t = pairs[2 * blockIdx.y + 1];
  const int ti = t % LOC_TILES, tj = t / LOC_TILES;
  const int i0 = G.ts[0][ti], j0 = G.ts[1][tj];
  const int ni = G.te[0][ti] - i0 + 1, nj = G.te[1][tj] - j0 + 1;
  const int k = G.te[2][LOC_TILES - 1];
  const float px = pts[3 * q], py = pts[3 * q + 1];
  for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < (long)ni * nj;
       c += (long)gridDim.x * 256) {
    const int i = i0 + (int)(c % ni), j = j0 + (int)(c / ni);
    lv3 v[8];
    cell_corners(G, i, j, k, v);
    float lo[2] = {v[0].x, v[0].y}, hi[2] = {v[0].x, v[0].y};
#pragma unroll
    for (int m = 1; m < 8; m++) {
      lo[0] = fminf(lo[0], v[m].x); hi[0] = fmaxf(hi[0], v[m].x);
      lo[1] = fminf(lo[1], v[m].y); hi[1] = fmaxf(hi[1], v[m].y);
    }
    if (px < lo[0] || px > hi[0] || py < lo[1] || py > hi[1]) continue;
    float zt;
    if (!surface_tri(G, i, j, k, px, py, &zt)) continue;
    atomicMin(&key[q], ((unsigned long long)t << 40) | (unsigned long long)c);
  }
}

// z of the winning surface cell minus the depth; points with no cell keep
// their depth
__global__ void surface_finish_kernel(LocGrid G, const float* pts, int npts,
                                      const unsigned long long* key, float* z, int* found)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= npts) return;
  const unsigned long long kk = key[q];
  if (kk == ~0ull) {
    found[q] = 0;
    z[q] = pts[3 * q + 2];
    return;
  }
  const int t = (int)(kk >> 40);
  const long c = (long)(kk & ((1ull << 40) - 1));
  const int ti = t % LOC_TILES, tj = t / LOC_TILES;
  const int ni = G.te[0][ti] - G.ts[0][ti] + 1;
  const int i = G.ts[0][ti] + (int)(c % ni), j = G.ts[1][tj] + (int)(c / ni);
  float zt = 0.0f;
  surface_tri(G, i, j, G.te[2][LOC_TILES - 1], pts[3 * q], pts[3 * q + 1], &zt);
  z[q] = zt - pts[3 * q + 2];
  found[q] = 1;
}

}  // namespace cgfdh
