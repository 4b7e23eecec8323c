/*
 * locate_oracle.c -- CPU restatement of the reference's point location on the
 * curvilinear grid.  TEST INFRASTRUCTURE ONLY: the checker of the engine's
 * cgfdh_locate_points / cgfdh_depth_to_axis; nothing in the product links it.
 *
 *  - orc_point_in_hexa: isPointInHexahedron (media/media_geometry3d.cpp:
 *    150-267): the cell cut into ten tetrahedra (two five-way cuts), a point
 *    is inside when one tetrahedron holds it (isPointInPolyhedron :167-182:
 *    on a face triangle by the area test, else on the inner side of all four
 *    faces within -1e-6 of the normalised distance).  Pinned to the
 *    reference's own function compiled from media_geometry3d.cpp.
 *  - orc_sample_shift: gd_curv_coord2shift_sample (gd_t.c:2304-2378), the
 *    nearest of (n+1)^3 trilinear samples of the cell.
 *  - orc_locate_point: gd_curv_coord_to_local_indx (gd_t.c:1998-2173) over
 *    the tiles and cell boxes of gd_curv_set_minmax (gd_t.c:1708-1880).
 *  - orc_depth_to_axis: gd_curv_depth_to_axis (gd_t.c:2179-2296), the top
 *    surface split into two triangles per cell (fdlib_math.c:246-345).
 * Every expression keeps the reference's float/double mix and order;
 * compiled without FMA contraction.
 */
#include <math.h>
#include <stddef.h>

#define TILES 4   /* GD_TILE_NX / _NY / _NZ (gd_t.h:9-11) */

/* ---- vectors as the geometry code forms them (float components) ---- */
typedef struct { float x, y, z; } v3;

static v3 vsub(v3 a, v3 b) { v3 r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static float vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static v3 vcross(v3 a, v3 b)
{
  v3 r = {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
  return r;
}
/* sqrt of the float dot product, taken in double and rounded to float */
static float vlen(v3 a) { return (float)sqrt((double)vdot(a, a)); }
static float area2(v3 a, v3 b, v3 c) { return vlen(vcross(vsub(b, a), vsub(c, a))); }

/* a point on the triangle: the three sub-areas add up to the whole within
 * 1e-6 (dcmp, media_geometry3d.cpp:33-38) */
static int on_triangle(v3 p, v3 a, v3 b, v3 c)
{
  float s1 = area2(p, a, b), s2 = area2(p, b, c), s3 = area2(p, c, a);
  float d = s1 + s2 + s3 - area2(a, b, c);
  return fabs((double)d) < 1e-6;
}

/* the ten tetrahedra as four faces each, vertex triples in the order whose
 * normals (cross of the edges from the first vertex) point inwards */
static const unsigned char kTetFaces[10][4][3] = {
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

static int in_tetra(v3 p, const v3 *v, const unsigned char f[4][3])
{
  for (int n = 0; n < 4; n++) {
    v3 a = v[f[n][0]], b = v[f[n][1]], c = v[f[n][2]];
    if (on_triangle(p, a, b, c)) return 1;
    v3 nrm = vcross(vsub(b, a), vsub(c, a));
    float d = vlen(nrm);
    nrm.x = nrm.x / d; nrm.y = nrm.y / d; nrm.z = nrm.z / d;
    v3 p2f = vsub(a, p);
    float s = vdot(p2f, nrm);
    s /= vlen(p2f);
    if (s <= -1.0e-6f) return 0;
  }
  return 1;
}

/* vertex n = corner (n&1, (n>>1)&1, n>>2) of the cell in i, j, k */
int orc_point_in_hexa(float px, float py, float pz, const float *vx, const float *vy,
                      const float *vz)
{
  v3 p = {px, py, pz}, v[8];
  for (int n = 0; n < 8; n++) { v[n].x = vx[n]; v[n].y = vy[n]; v[n].z = vz[n]; }
  for (int t = 0; t < 10; t++)
    if (in_tetra(p, v, kTetFaces[t])) return 1;
  return 0;
}

/* the sample of the trilinear cell nearest the point; first minimum in
 * n3, n2, n1 order, starting from corner 0 */
void orc_sample_shift(float sx, float sy, float sz, const float *px, const float *py,
                      const float *pz, int nxs, int nys, int nzs, float *shift)
{
  float best = sqrtf((sx - px[0]) * (sx - px[0]) + (sy - py[0]) * (sy - py[0]) +
                     (sz - pz[0]) * (sz - pz[0]));
  int bi = 0, bj = 0, bk = 0;
  for (int n3 = 0; n3 <= nzs; n3++) {
    float Lz[2];
    Lz[1] = (float)n3 / (float)nzs;
    Lz[0] = 1.0 - Lz[1];
    for (int n2 = 0; n2 <= nys; n2++) {
      float Ly[2];
      Ly[1] = (float)n2 / (float)nys;
      Ly[0] = 1.0 - Ly[1];
      for (int n1 = 0; n1 <= nxs; n1++) {
        float Lx[2];
        Lx[1] = (float)n1 / (float)nxs;
        Lx[0] = 1.0 - Lx[1];
        float x = 0, y = 0, z = 0;
        for (int c = 0; c < 8; c++) {
          const int a = c & 1, b = (c >> 1) & 1, e = c >> 2;
          x += Lx[a] * Ly[b] * Lz[e] * px[c];
          y += Lx[a] * Ly[b] * Lz[e] * py[c];
          z += Lx[a] * Ly[b] * Lz[e] * pz[c];
        }
        float d = sqrtf((sx - x) * (sx - x) + (sy - y) * (sy - y) + (sz - z) * (sz - z));
        if (best > d) { best = d; bi = n1; bj = n2; bk = n3; }
      }
    }
  }
  shift[0] = (float)bi / (float)nxs;
  shift[1] = (float)bj / (float)nys;
  shift[2] = (float)bk / (float)nzs;
}

/* the tile ranges of gd_curv_set_minmax: physical points split evenly, the
 * first (n % 4) tiles one point longer */
static void tile_ranges(int n, int *s, int *e)
{
  int avg = n / TILES, left = n % TILES;
  for (int t = 0; t < TILES; t++) {
    s[t] = t == 0 ? 3 : e[t - 1] + 1;
    e[t] = s[t] + avg - 1 + (t < left ? 1 : 0);
  }
}

typedef struct {
  int nx, ny;
  size_t sl, ss;
  const float *x, *y, *z;
} grid3;

static void cell_box(const grid3 *g, int i, int j, int k, float *bmin, float *bmax)
{
  size_t p = i + j * g->sl + k * g->ss;
  bmin[0] = bmax[0] = g->x[p];
  bmin[1] = bmax[1] = g->y[p];
  bmin[2] = bmax[2] = g->z[p];
  for (int n = 0; n < 8; n++) {
    size_t q = p + (n >> 2) * g->ss + ((n >> 1) & 1) * g->sl + (n & 1);
    float v[3] = {g->x[q], g->y[q], g->z[q]};
    for (int d = 0; d < 3; d++) {
      bmin[d] = bmin[d] < v[d] ? bmin[d] : v[d];
      bmax[d] = bmax[d] > v[d] ? bmax[d] : v[d];
    }
  }
}

static void cell_corners(const grid3 *g, int i, int j, int k, float *vx, float *vy, float *vz)
{
  for (int n = 0; n < 8; n++) {
    size_t q = (i + (n & 1)) + (j + ((n >> 1) & 1)) * g->sl + (k + (n >> 2)) * g->ss;
    vx[n] = g->x[q]; vy[n] = g->y[q]; vz[n] = g->z[q];
  }
}

static int outside(const float *p, const float *bmin, const float *bmax, int nd)
{
  for (int d = 0; d < nd; d++)
    if (p[d] < bmin[d] || p[d] > bmax[d]) return 1;
  return 0;
}

/* the box of every cell of tile (ti, tj, tk) */
static void tile_box(const grid3 *g, const int *is, const int *ie, const int *js,
                     const int *je, const int *ks, const int *ke, int ti, int tj, int tk,
                     float *bmin, float *bmax)
{
  for (int d = 0; d < 3; d++) { bmin[d] = 1.0e26f; bmax[d] = -1.0e26f; }
  for (int k = ks[tk]; k <= ke[tk]; k++)
    for (int j = js[tj]; j <= je[tj]; j++)
      for (int i = is[ti]; i <= ie[ti]; i++) {
        float cmin[3], cmax[3];
        cell_box(g, i, j, k, cmin, cmax);
        for (int d = 0; d < 3; d++) {
          bmin[d] = bmin[d] < cmin[d] ? bmin[d] : cmin[d];
          bmax[d] = bmax[d] > cmax[d] ? bmax[d] : cmax[d];
        }
      }
}

static void grid_range(const grid3 *g, int nz, float *gmin, float *gmax)
{
  size_t n = (size_t)g->ss * nz;
  gmin[0] = gmax[0] = g->x[0]; gmin[1] = gmax[1] = g->y[0]; gmin[2] = gmax[2] = g->z[0];
  for (size_t p = 0; p < n; p++) {
    float v[3] = {g->x[p], g->y[p], g->z[p]};
    for (int d = 0; d < 3; d++) {
      gmin[d] = gmin[d] < v[d] ? gmin[d] : v[d];
      gmax[d] = gmax[d] > v[d] ? gmax[d] : v[d];
    }
  }
}

/* 1 and the index (with ghosts) and shift of the point, or 0; cell (may be
 * NULL) receives the cell that holds it, before the shift past one half
 * moves the index to the next point */
int orc_locate_point(int ni, int nj, int nk, const float *x, const float *y, const float *z,
                     float sx, float sy, float sz, int *ijk, float *inc, int *cell)
{
  grid3 g = {ni + 6, nj + 6, (size_t)(ni + 6), (size_t)(ni + 6) * (nj + 6), x, y, z};
  float p[3] = {sx, sy, sz}, gmin[3], gmax[3];
  grid_range(&g, nk + 6, gmin, gmax);
  if (outside(p, gmin, gmax, 3)) return 0;
  int is[TILES], ie[TILES], js[TILES], je[TILES], ks[TILES], ke[TILES];
  tile_ranges(ni, is, ie);
  tile_ranges(nj, js, je);
  tile_ranges(nk, ks, ke);
  for (int tk = 0; tk < TILES; tk++)
    for (int tj = 0; tj < TILES; tj++)
      for (int ti = 0; ti < TILES; ti++) {
        float tmin[3], tmax[3];
        tile_box(&g, is, ie, js, je, ks, ke, ti, tj, tk, tmin, tmax);
        if (outside(p, tmin, tmax, 3)) continue;
        for (int k = ks[tk]; k <= ke[tk]; k++)
          for (int j = js[tj]; j <= je[tj]; j++)
            for (int i = is[ti]; i <= ie[ti]; i++) {
              float cmin[3], cmax[3];
              cell_box(&g, i, j, k, cmin, cmax);
              if (outside(p, cmin, cmax, 3)) continue;
              float vx[8], vy[8], vz[8];
              cell_corners(&g, i, j, k, vx, vy, vz);
              if (!orc_point_in_hexa(sx, sy, sz, vx, vy, vz)) continue;
              orc_sample_shift(sx, sy, sz, vx, vy, vz, 100, 100, 100, inc);
              ijk[0] = i; ijk[1] = j; ijk[2] = k;
              if (cell) { cell[0] = i; cell[1] = j; cell[2] = k; }
              for (int d = 0; d < 3; d++)
                if (inc[d] > 0.5) { ijk[d] += 1; inc[d] -= 1.0; }
              return 1;
            }
      }
  return 0;
}

/* fdlib_math_isPoint2InTri: on the left of (or on) all three edges */
static int in_tri2(float px, float py, const float *vx, const float *vy)
{
  for (int i = 0; i < 3; i++) {
    float ex = i == 2 ? (vx[0] - vx[2]) : (vx[i + 1] - vx[i]);
    float ey = i == 2 ? (vy[0] - vy[2]) : (vy[i + 1] - vy[i]);
    float qx = px - vx[i], qy = py - vy[i];
    float s = ex * qy - ey * qx;
    s /= sqrt(ex * ex + ey * ey);
    if (s < 0) return 0;
  }
  return 1;
}

/* fdlib_math_computeZfromTri: the plane through three points at (x, y) */
static float z_on_tri(float x, float y, float x1, float y1, float z1, float x2, float y2,
                      float z2, float x3, float y3, float z3)
{
  float ax = x2 - x1, ay = y2 - y1, az = z2 - z1;
  float bx = x3 - x1, by = y3 - y1, bz = z3 - z1;
  float a = ay * bz - az * by;
  float b = az * bx - ax * bz;
  float c = ax * by - ay * bx;
  float d = -(a * x1 + b * y1 + c * z1);
  return -(a * x + b * y + d) / c;
}

/* *sz: depth below the top surface in, z coordinate out; unchanged when the
 * point is not over a surface cell (the return value says which) */
int orc_depth_to_axis(int ni, int nj, int nk, const float *x, const float *y, const float *z,
                      float sx, float sy, float *sz)
{
  grid3 g = {ni + 6, nj + 6, (size_t)(ni + 6), (size_t)(ni + 6) * (nj + 6), x, y, z};
  float p[2] = {sx, sy}, gmin[3], gmax[3];
  grid_range(&g, nk + 6, gmin, gmax);
  if (outside(p, gmin, gmax, 2)) return 0;
  int is[TILES], ie[TILES], js[TILES], je[TILES], ks[TILES], ke[TILES];
  tile_ranges(ni, is, ie);
  tile_ranges(nj, js, je);
  tile_ranges(nk, ks, ke);
  const int tk = TILES - 1, k = ke[tk];
  for (int tj = 0; tj < TILES; tj++)
    for (int ti = 0; ti < TILES; ti++) {
      float tmin[3], tmax[3];
      tile_box(&g, is, ie, js, je, ks, ke, ti, tj, tk, tmin, tmax);
      if (outside(p, tmin, tmax, 2)) continue;
      for (int j = js[tj]; j <= je[tj]; j++)
        for (int i = is[ti]; i <= ie[ti]; i++) {
          float cmin[3], cmax[3];
          cell_box(&g, i, j, k, cmin, cmax);
          if (outside(p, cmin, cmax, 2)) continue;
          float qx[4], qy[4], qz[4];
          for (int n = 0; n < 4; n++) {
            size_t q = (i + (n & 1)) + (j + (n >> 1)) * g.sl + k * g.ss;
            qx[n] = x[q]; qy[n] = y[q]; qz[n] = z[q];
          }
          /* fdlib_math_Point2InWhichTri: triangle 3-2-0 first, then 0-1-3 */
          float tx[5] = {qx[0], qx[1], qx[3], qx[2], qx[0]};
          float ty[5] = {qy[0], qy[1], qy[3], qy[2], qy[0]};
          float zt;
          if (in_tri2(sx, sy, tx + 2, ty + 2))
            zt = z_on_tri(sx, sy, qx[3], qy[3], qz[3], qx[2], qy[2], qz[2], qx[0], qy[0], qz[0]);
          else if (in_tri2(sx, sy, tx, ty))
            zt = z_on_tri(sx, sy, qx[0], qy[0], qz[0], qx[1], qy[1], qz[1], qx[3], qy[3], qz[3]);
          else
            continue;
          *sz = zt - (*sz);
          return 1;
        }
    }
  return 0;
}
