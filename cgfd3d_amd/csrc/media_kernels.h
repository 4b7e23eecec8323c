// media_kernels.h -- the layered model discretised onto the grid on the
// device (media/media_layer2model.cpp), for every media type of the .md3lay
// file and the three equivalent-medium methods:
//  * the point value (AssignLayerMediaPara2Point / CalPointValue_layer
//    :422-1238): each interface's elevation interpolated bilinearly at the
//    point (BilinearInterpolation, media_utility.cpp:173-210), the layer
//    holding it (findLastGreaterEqualIndex :88-113), value + dz^pow * grad
//    per property, the mean of the two layers' values on an interface (the
//    stiffness-tensor and TTI types keep the lower layer's dz for the upper
//    value, as their cases never recompute it);
//  * the conversion to coefficients: lambda, mu (the el_iso loop :1334-1385),
//    para2vti and para2tti with the Bond rotation (media_utility.cpp:835-1100);
//  * "loc": one thread per grid point;
//  * "har" / "ari" on a curvilinear grid (:1603-2830): the half grid
//    (GenerateHalfGrid), each half point's layer (MarkInterfaceNumber), the
//    points whose eight surrounding half points lie in more than one layer
//    (NumOfValues), and for each such point one block that evaluates the
//    9x9x9 sub-points of its half cell (GenerateHalfMesh, MeshSubdivide) in
//    parallel into LDS and then forms the reference's running sums in its
//    sub-point order, one coefficient per thread, so the float accumulation
//    rounds as the reference's does;
//  * the expansion media_layer2model_el_aniso applies to isotropic and VTI
//    files (:288-370);
//  * the grid model of media_grid2model (.md3grd): surfaces grouped into
//    layers, values interpolated linearly in z between the surfaces around
//    a point, the layer number of each half point among the layers' top
//    surfaces, sub-points on a layer boundary left out of the averages, no
//    averaging with one layer (media_grid2model.cpp:424-2800).
// Same float/double arithmetic as the reference (the power law, cos and sin
// in double); results agree with it to the last bit except where two math
// libraries' pow, cos or sin differ.
#pragma once

#include <cfloat>

namespace cgfdh {

constexpr int MEDIA_MAX_IF = 64;       // interfaces of a layered model
constexpr int MEDIA_MAX_SURF = 256;    // surfaces of a grid model
constexpr int MEDIA_MAX_LAYERS = 64;   // layers of a grid model
constexpr int MEDIA_NG = 8;
constexpr int MEDIA_NSG = (MEDIA_NG + 1) * (MEDIA_NG + 1) * (MEDIA_NG + 1);

// media type codes of the .md3lay header (media_utility.hpp:37-45)
enum : int {
  MD_ONE = 0, MD_AC = 10, MD_ISO = 21, MD_PREM = 22, MD_THOMSEN = 23, MD_VCIJ = 24,
  MD_TTHOMSEN = 25, MD_TBOND = 26, MD_ACIJ = 27
};

// (value, gradient, power) triples per sample of a media type
__host__ __device__ inline int media_nprop(int mt)
{
  switch (mt) {
  case MD_ONE: return 1;
  case MD_AC: return 2;
  case MD_ISO: return 3;
  case MD_PREM: case MD_THOMSEN: case MD_VCIJ: return 6;
  case MD_TTHOMSEN: case MD_TBOND: return 8;
  case MD_ACIJ: return 22;
  }
  return -1;
}

// coefficients of a point before any expansion: var | lambda mu rho |
// c11 c33 c55 c66 c13 rho | c11 c12 .. c66 (21) rho
__host__ __device__ inline int media_native(int mt)
{
  switch (mt) {
  case MD_ONE: return 1;
  case MD_ISO: return 3;
  case MD_PREM: case MD_THOMSEN: case MD_VCIJ: return 6;
  case MD_TTHOMSEN: case MD_TBOND: case MD_ACIJ: return 22;
  }
  return -1;
}

struct LayerModel {
  // layered model (grid 0): [1 + 3 nprop][ni][my][mx], elevation then
  // value/grad/pow per property; grid model (grid 1): [1 + nprop][ni][my][mx],
  // elevation then the values, ni surfaces in nl layers
  const float* f;
  int ni, mx, my, nprop, mt;
  float minx, miny, dx, dy;
  int grid, nl;
  int first[MEDIA_MAX_LAYERS];           // grid model: each layer's top surface
  unsigned bound[MEDIA_MAX_SURF / 32];   // grid model: surfaces that begin a layer (not the first)
  unsigned long long* bad;   // first failing point: code << 48 | index
};

// where the native coefficients go among the outputs ([nout][n])
struct LayerOut {
  float* out;
  long n;
  int nn;                // native coefficients
  int slot[22];
};

__device__ __forceinline__ float layer_bilinear(const LayerModel& M, const float* v, float xq,
                                                float yq)
{
  const float x0 = M.minx + M.dx * 0, x1 = M.minx + M.dx * 1;
  const float y0 = M.miny + M.dy * 0, y1 = M.miny + M.dy * 1;
  const int ix = (int)((xq - x0) / (x1 - x0));
  const int iy = (int)((yq - y0) / (y1 - y0));
  if (ix < 0 || iy < 0) return FLT_MAX;
  const long indx = (long)iy * M.mx + ix;
  if (ix == M.mx - 1 || iy == M.my - 1) return v[indx];
  const float xa = M.minx + M.dx * ix, xb = M.minx + M.dx * (ix + 1);
  const float ya = M.miny + M.dy * iy, yb = M.miny + M.dy * (iy + 1);
  const float area = (xb - xa) * (yb - ya);
  float vq = v[indx] * (xb - xq) * (yb - yq) + v[indx + 1] * (xq - xa) * (yb - yq) +
             v[indx + M.mx] * (xb - xq) * (yq - ya) + v[indx + 1 + M.mx] * (xq - xa) * (yq - ya);
  vq /= area;
  return vq;
}

__device__ __forceinline__ int layer_last_ge(float value, const float* x, int n)
{
  if (value > x[0]) return -1;
  float dist = FLT_MAX;
  int indx = -1;
  for (int i = 0; i < n; i++) {
    if (x[i] == -FLT_MAX) continue;
    const float d = x[i] - value;
    if (d < 0) break;
    if (d >= 0 && d <= dist) {
      dist = d;
      indx = i;
    }
  }
  return indx;
}

__device__ __forceinline__ int layer_first_ge(float value, const float* x, int n)
{
  float dist = FLT_MAX;
  int idx = -1;
  for (int i = 0; i < n; i++) {
    if (x[i] == -FLT_MAX) continue;
    const float d = x[i] - value;
    if (d >= 0 && d <= dist) {
      if (d <= 1e-6) return i;
      dist = d;
      idx = i;
    }
  }
  return idx;
}

// one property's value + dz^pow * grad at layer mi, in double, rounded
__device__ __forceinline__ float layer_law(const LayerModel& M, long vol, int q, int mi,
                                           float x, float y, float dz)
{
  const long sl = (long)M.mx * M.my;
  const float* v = M.f + (1 + 3 * q) * vol + mi * sl;
  const float v0 = layer_bilinear(M, v, x, y);
  const float vg = layer_bilinear(M, v + vol, x, y);
  const float vp = layer_bilinear(M, v + 2 * vol, x, y);
  return (float)((double)v0 + pow((double)dz, (double)vp) * (double)vg);
}

// the layer of a point (MarkInterfaceNumber's search): above the top
// interface, the top layer of non-zero thickness
__device__ __forceinline__ int layer_index(const LayerModel& M, float x, float y, float z)
{
  const long sl = (long)M.mx * M.my;
  float e[MEDIA_MAX_LAYERS > MEDIA_MAX_IF ? MEDIA_MAX_LAYERS : MEDIA_MAX_IF];
  // grid model: LayerNumberAtPoint (media_grid2model.cpp:1269-1310) over
  // the layers' top surfaces
  const int n = M.grid ? M.nl : M.ni;
  for (int q = 0; q < n; q++) e[q] = layer_bilinear(M, M.f + (M.grid ? M.first[q] : q) * sl, x, y);
  int mi = layer_last_ge(z, e, n);
  if (mi == -1) mi = layer_last_ge(e[0], e, n);
  return mi;
}

// the point's property values: 0, 1 outside the mesh, 2 an isotropic point
// with vp <= sqrt(2) vs (the reference stops on both)
__device__ int layer_values(const LayerModel& M, float x, float y, float z, float* var)
{
  const float maxx = M.minx + (M.mx - 1) * M.dx, maxy = M.miny + (M.my - 1) * M.dy;
  if (x < M.minx || x > maxx || y < M.miny || y > maxy) return 1;
  const long sl = (long)M.mx * M.my, vol = (long)M.ni * sl;
  float e[MEDIA_MAX_IF];
  for (int n = 0; n < M.ni; n++) e[n] = layer_bilinear(M, M.f + n * sl, x, y);
  int mi = layer_last_ge(z, e, M.ni);
  if (mi == -1) {
    mi = layer_last_ge(e[0], e, M.ni);
    for (int q = 0; q < M.nprop; q++)
      var[q] = layer_bilinear(M, M.f + (1 + 3 * q) * vol + mi * sl, x, y);
    return 0;
  }
  const float dz = e[mi] - z;
  for (int q = 0; q < M.nprop; q++) var[q] = layer_law(M, vol, q, mi, x, y, dz);
  if (M.mt == MD_ISO && (double)var[1] <= (double)var[2] * sqrt(2.0)) return 2;
  if (fabsf(dz - 0.0f) < FLT_EPSILON) {
    const int m2 = layer_first_ge(z, e, M.ni) - 1;
    if (m2 >= 0) {
      const bool stale = M.mt == MD_VCIJ || M.mt == MD_TTHOMSEN || M.mt == MD_TBOND ||
                         M.mt == MD_ACIJ || M.mt == MD_AC;
      const float dz2 = stale ? dz : e[m2] - z;
      for (int q = 0; q < M.nprop; q++) {
        const float top = layer_law(M, vol, q, m2, x, y, dz2);
        var[q] = (float)(((double)var[q] + (double)top) / 2.0);
      }
    }
  }
  return 0;
}

__device__ __forceinline__ bool grid_bound(const LayerModel& M, int s)
{
  return (M.bound[s >> 5] >> (s & 31)) & 1u;
}

// the grid model's point values (AssignGridMediaPara2Point /
// CalPointValue_grid, media_grid2model.cpp:424-958): linear in z between
// the surfaces around the point, the top surface's values above it, the
// last's below; on a boundary between layers the mean of the surfaces
// either side.  on_if: the point lies on a layer boundary.  Codes as
// layer_values.
__device__ int grid_values(const LayerModel& M, float x, float y, float z, float* var, int& on_if)
{
  on_if = 0;
  const float maxx = M.minx + (M.mx - 1) * M.dx, maxy = M.miny + (M.my - 1) * M.dy;
  if (x < M.minx || x > maxx || y < M.miny || y > maxy) return 1;
  const long sl = (long)M.mx * M.my, vol = (long)M.ni * sl;
  float e[MEDIA_MAX_SURF];
  for (int n = 0; n < M.ni; n++) e[n] = layer_bilinear(M, M.f + n * sl, x, y);
  int mi = layer_last_ge(z, e, M.ni);
  on_if = mi >= 0 && grid_bound(M, mi) && fabsf(z - e[mi]) < FLT_EPSILON;
  float dis_r = 1.0f / 2.0f;
  if (mi >= 0 && mi < M.ni - 1 && fabs((double)(e[mi + 1] - e[mi])) > 1e-6)
    dis_r = (z - e[mi]) / (e[mi + 1] - e[mi]);
  if (mi == -1 || mi == M.ni - 1) {
    if (mi == -1) mi = layer_last_ge(e[0], e, M.ni);
    for (int q = 0; q < M.nprop; q++)
      var[q] = layer_bilinear(M, M.f + (1 + q) * vol + mi * sl, x, y);
    return 0;
  }
  if (mi > 0 && fabsf(e[mi] - z) < FLT_EPSILON && grid_bound(M, mi)) {
    mi--;
    dis_r = 1.0f / 2.0f;
  }
  for (int q = 0; q < M.nprop; q++) {
    const float v0 = layer_bilinear(M, M.f + (1 + q) * vol + mi * sl, x, y);
    const float v1 = layer_bilinear(M, M.f + (1 + q) * vol + (mi + 1) * sl, x, y);
    var[q] = v0 + (v1 - v0) * dis_r;
  }
  if (M.mt == MD_ISO && (double)var[1] <= sqrt(2.0) * (double)var[2]) return 2;
  return 0;
}

__device__ __forceinline__ int model_values(const LayerModel& M, float x, float y, float z,
                                            float* var, int& on_if)
{
  if (M.grid) return grid_values(M, x, y, z, var, on_if);
  on_if = 0;
  return layer_values(M, x, y, z, var);
}

// BondTransform (dip theta, azimuth phi in degrees) of a transversely
// isotropic stiffness about z: C = R C0 R^T, products summed in float in
// row order; o receives c11 c12 .. c66
__device__ void layer_bond(float c11, float c13, float c33, float c44, float c66, float theta,
                           float phi, float* o)
{
  const double pi = acos(-1.0);
  theta = (float)((double)theta * (pi / 180.0));
  phi = (float)((double)phi * (pi / 180.0));
  const double ct = cos((double)theta), st = sin((double)theta);
  const double cp = cos((double)phi), sp = sin((double)phi);
  const float a11 = (float)(ct * sp), a12 = (float)cp, a13 = (float)(st * sp);
  const float a21 = (float)(-ct * cp), a22 = (float)sp, a23 = (float)(-st * cp);
  const float a31 = (float)(-st), a32 = 0.0f, a33 = (float)ct;
  const float R[36] = {
      a11 * a11, a12 * a12, a13 * a13, 2 * a12 * a13, 2 * a13 * a11, 2 * a11 * a12,
      a21 * a21, a22 * a22, a23 * a23, 2 * a22 * a23, 2 * a23 * a21, 2 * a21 * a22,
      a31 * a31, a32 * a32, a33 * a33, 2 * a32 * a33, 2 * a33 * a31, 2 * a31 * a32,
      a21 * a31, a22 * a32, a23 * a33, a22 * a33 + a23 * a32, a21 * a33 + a23 * a31,
      a22 * a31 + a21 * a32,
      a31 * a11, a32 * a12, a33 * a13, a12 * a33 + a13 * a32, a13 * a31 + a11 * a33,
      a11 * a32 + a12 * a31,
      a11 * a21, a12 * a22, a13 * a23, a12 * a23 + a13 * a22, a13 * a21 + a11 * a23,
      a11 * a22 + a12 * a21};
  const float c12 = c11 - 2 * c66;
  const float C0[36] = {c11, c12, c13, 0, 0, 0,   c12, c11, c13, 0, 0, 0,
                        c13, c13, c33, 0, 0, 0,   0, 0, 0, c44, 0, 0,
                        0, 0, 0, 0, c44, 0,       0, 0, 0, 0, 0, c66};
  float T[36];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      float s = 0;
      for (int k = 0; k < 6; k++) s += R[i * 6 + k] * C0[k * 6 + j];
      T[i * 6 + j] = s;
    }
  int n = 0;
  for (int i = 0; i < 6; i++)
    for (int j = i; j < 6; j++) {
      float s = 0;
      for (int k = 0; k < 6; k++) s += T[i * 6 + k] * R[j * 6 + k];
      o[n++] = s;
    }
}

// the point's coefficients from its property values
__device__ void layer_native(int mt, const float* var, float* o)
{
  switch (mt) {
  case MD_ONE:
    o[0] = var[0];
    break;
  case MD_ISO: {
    const float vp = var[1], vs = var[2], rho = var[0];
    o[1] = vs * vs * rho;
    o[0] = (float)((double)(vp * vp * rho) - 2.0 * (double)o[1]);
    o[2] = rho;
  } break;
  case MD_PREM: {
    const float vph = var[1], vpv = var[2], vsh = var[3], vsv = var[4], rho = var[0];
    const float eta = var[5];
    o[0] = vph * vph * rho;
    o[1] = vpv * vpv * rho;
    o[3] = vsh * vsh * rho;
    o[2] = vsv * vsv * rho;
    o[5] = rho;
    o[4] = (float)((double)eta * ((double)o[0] - 2.0 * (double)o[2]));
  } break;
  case MD_THOMSEN: {
    const float rho = var[0], vp0 = var[1], vs0 = var[2], epsil = var[3], delta = var[4];
    const float gamma = var[5];
    const float c33 = vp0 * vp0 * rho, c55 = vs0 * vs0 * rho;
    o[5] = rho;
    o[1] = c33;
    o[2] = c55;
    o[3] = (float)(2.0 * (double)gamma * (double)c55 + (double)c55);
    o[0] = (float)(2.0 * (double)epsil * (double)c33 + (double)c33);
    const float d = c33 - c55;
    o[4] = (float)(sqrt(2.0 * (double)delta * (double)c33 * (double)d + (double)(d * d)) -
                   (double)c55);
  } break;
  case MD_VCIJ:
    o[0] = var[1]; o[1] = var[2]; o[2] = var[3]; o[3] = var[4]; o[4] = var[5]; o[5] = var[0];
    break;
  case MD_TTHOMSEN: {
    const float rho = var[0], vp0 = var[1], vs0 = var[2], epsil = var[3], delta = var[4];
    const float gamma = var[5];
    const float c33 = vp0 * vp0 * rho;
    const float c44 = vs0 * vs0 * rho;
    const float c66 = (float)(2.0 * (double)gamma * (double)c44 + (double)c44);
    const float c11 = (float)(2.0 * (double)epsil * (double)c33 + (double)c33);
    const float d = c33 - c44;
    const float c13 = (float)(sqrt(2.0 * (double)delta * (double)c33 * (double)d +
                                   (double)(d * d)) - (double)c44);
    o[21] = rho;
    layer_bond(c11, c13, c33, c44, c66, var[7], var[6], o);
  } break;
  case MD_TBOND:
    o[21] = var[0];
    layer_bond(var[1], var[5], var[2], var[3], var[4], var[7], var[6], o);
    break;
  case MD_ACIJ:
    for (int q = 0; q < 21; q++) o[q] = var[1 + q];
    o[21] = var[0];
    break;
  }
}

// "loc": every point's coefficients into its output slots
__global__ void __launch_bounds__(256)
layer_loc_kernel(LayerModel M, const float* x, const float* y, const float* z, LayerOut O)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= O.n) return;
  float var[22], nat[22];
  int on_if;
  const int rc = model_values(M, x[p], y[p], z[p], var, on_if);
  if (rc) {
    atomicMin(M.bad, ((unsigned long long)rc << 48) | (unsigned long long)p);
    return;
  }
  layer_native(M.mt, var, nat);
  for (int q = 0; q < O.nn; q++) O.out[(long)O.slot[q] * O.n + p] = nat[q];
}

// GenerateHalfGrid on a curvilinear grid: the mean of each cell's eight
// corners (float sum in the reference's corner order, then / 8), the last
// plane of each direction the grid itself; all three coordinates
__global__ void __launch_bounds__(256)
half_grid_kernel(const float* g, float* h, int nx, int ny, int nz)
{
  const long n = (long)nx * ny * nz;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long)nx * ny));
  const long line = nx, slice = (long)nx * ny;
  for (int d = 0; d < 3; d++) {
    const float* G = g + d * n;
    float v = G[p];
    if (i < nx - 1 && j < ny - 1 && k < nz - 1)
      v = (float)((double)(G[p] + G[p + 1] + G[p + line + 1] + G[p + line] + G[p + slice] +
                           G[p + slice + 1] + G[p + slice + line + 1] + G[p + line + slice]) /
                  8.0);
    h[d * n + p] = v;
  }
}

// MarkInterfaceNumber: each half point's layer
__global__ void __launch_bounds__(256)
layer_mark_kernel(LayerModel M, const float* h, long n, int* mat)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  mat[p] = layer_index(M, h[p], h[n + p], h[2 * n + p]);
}

// the points (1..n-2 in each direction) whose eight surrounding half points
// lie in more than one layer and none above every interface (NumOfValues > 1)
__global__ void __launch_bounds__(256)
layer_mixed_kernel(const int* mat, int nx, int ny, int nz, int* list, unsigned* count)
{
  const long n = (long)nx * ny * nz;
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % nx), j = (int)((p / nx) % ny), k = (int)(p / ((long)nx * ny));
  if (i < 1 || j < 1 || k < 1 || i > nx - 2 || j > ny - 2 || k > nz - 2) return;
  const long line = nx, slice = (long)nx * ny;
  const int v[8] = {mat[p - 1 - line - slice], mat[p - line - slice], mat[p - slice],
                    mat[p - 1 - slice],        mat[p - 1 - line],     mat[p - line],
                    mat[p],                    mat[p - 1]};
  bool mixed = false;
  for (int q = 0; q < 8; q++) {
    if (v[q] == -1) return;
    mixed |= v[q] != v[0];
  }
  if (mixed) list[atomicAdd(count, 1u)] = (int)p;
}

// the averaged coefficients of one mixed point per block: phase 1 fills
// sub[q][s] with the native coefficients of the 729 sub-points of its half
// cell; phase 2 forms each coefficient's running sum in sub-point order
// (thread q, or thread 0 for the isotropic combination)
template <int NN>
__global__ void __launch_bounds__(256)
layer_average_kernel(LayerModel M, const float* h, const int* list, int harmonic, LayerOut O,
                     int nx, int ny)
{
  __shared__ float sub[NN][MEDIA_NSG];
  __shared__ unsigned char use[MEDIA_NSG];   // grid model: off the layer boundaries
  __shared__ int fail;
  const long n = O.n;
  const long p = list[blockIdx.x];
  const long line = nx, slice = (long)nx * ny;
  if (threadIdx.x == 0) fail = 0;
  // GenerateHalfMesh's vertex order
  const long hv[8] = {p - 1 - line - slice, p - 1 - slice, p - slice, p - line - slice,
                      p - 1 - line,         p - 1,         p,         p - line};
  __syncthreads();
  for (int s = threadIdx.x; s < MEDIA_NSG; s += 256) {
    const int ii = s % 9, jj = (s / 9) % 9, kk = s / 81;
    float P[3];
    const float ng = (float)MEDIA_NG;
    for (int d = 0; d < 3; d++) {
      const float* H = h + d * n;
      const float v0 = H[hv[0]], v1 = H[hv[1]], v2 = H[hv[2]], v3 = H[hv[3]];
      const float v4 = H[hv[4]], v5 = H[hv[5]], v6 = H[hv[6]], v7 = H[hv[7]];
      const float ly00 = v0 + (v1 - v0) / ng * (float)jj;
      const float ly10 = v3 + (v2 - v3) / ng * (float)jj;
      const float ly01 = v4 + (v5 - v4) / ng * (float)jj;
      const float ly11 = v7 + (v6 - v7) / ng * (float)jj;
      const float lxy0 = ly00 + (ly10 - ly00) / ng * (float)ii;
      const float lxy1 = ly01 + (ly11 - ly01) / ng * (float)ii;
      P[d] = lxy0 + (lxy1 - lxy0) / ng * (float)kk;
    }
    float var[22], nat[22];
    int on_if;
    const int rc = model_values(M, P[0], P[1], P[2], var, on_if);
    if (rc) {
      atomicMin(M.bad, ((unsigned long long)rc << 48) | (unsigned long long)p);
      fail = 1;
      continue;
    }
    layer_native(M.mt, var, nat);
    for (int q = 0; q < NN; q++) sub[q][s] = nat[q];
    use[s] = !on_if;
  }
  __syncthreads();
  if (fail) return;
  const int t = threadIdx.x;
  // the sums run over the used sub-points in order and divide by their
  // count (a float sum over the int count and over the count in double
  // round alike, so one expression serves both models)
  int num = 0;
  for (int k = 0; k < MEDIA_NSG; k++) num += use[k];
  if (M.mt == MD_ONE) {
    if (t == 0) {
      float s = 0.0f;
      if (harmonic) {
        for (int k = 0; k < MEDIA_NSG; k++)
          if (use[k]) s = (float)((double)s + 1.0 / (double)sub[0][k]);
        O.out[(long)O.slot[0] * n + p] = (float)(num * 1.0 / (double)s);
      } else {
        for (int k = 0; k < MEDIA_NSG; k++)
          if (use[k]) s += sub[0][k];
        O.out[(long)O.slot[0] * n + p] = (float)((double)s / (num * 1.0));
      }
    }
  } else if (M.mt == MD_ISO) {
    // native lambda, mu, rho; three running sums side by side
    __shared__ float acc[3];
    if (t < 3) {
      float s = 0.0f;
      for (int k = 0; k < MEDIA_NSG; k++) {
        if (!use[k]) continue;
        const float lam = sub[0][k], mu = sub[1 % NN][k], rho = sub[2 % NN][k];
        if (!harmonic) s += t == 0 ? lam : (t == 1 ? mu : rho);
        else if (t == 0) s += rho;
        else if (t == 1) s = (float)((double)s + 1.0 / ((double)lam + 2.0 / 3.0 * (double)mu));
        else s = (float)((double)s + 1.0 / (double)mu);
      }
      acc[t] = s;
    }
    __syncthreads();
    if (t == 0) {
      float lam, mu, rho;
      if (harmonic) {
        const float har_mu = (float)(num * 1.0 / (double)acc[2]);
        const float har_kappa = (float)(num * 1.0 / (double)acc[1]);
        rho = (float)((double)acc[0] / (num * 1.0));
        lam = (float)((double)har_kappa - 2.0 / 3.0 * (double)har_mu);
        mu = har_mu;
      } else {
        lam = (float)((double)acc[0] / (num * 1.0));
        mu = (float)((double)acc[1] / (num * 1.0));
        rho = (float)((double)acc[2] / (num * 1.0));
      }
      O.out[(long)O.slot[0] * n + p] = lam;
      O.out[(long)O.slot[1 % NN] * n + p] = mu;
      O.out[(long)O.slot[2 % NN] * n + p] = rho;
    }
  } else if (t < NN) {
    // stiffnesses harmonic or arithmetic (float sums, / the int count),
    // the density (last) arithmetic
    float s = 0.0f;
    if (harmonic && t < NN - 1) {
      for (int k = 0; k < MEDIA_NSG; k++)
        if (use[k]) s = (float)((double)s + 1.0 / (double)sub[t][k]);
      s = (float)(1.0 * num / (double)s);
    } else {
      for (int k = 0; k < MEDIA_NSG; k++)
        if (use[k]) s += sub[t][k];
      s = s / (float)num;
    }
    O.out[(long)O.slot[t] * n + p] = s;
  }
}

// media_layer2model_el_aniso's expansion of an isotropic (c13 = lambda,
// c44 = mu) or VTI (c11 c33 c55 c66 c13) result; outputs rho c11 c12 .. c66
__global__ void __launch_bounds__(256) layer_expand_kernel(int vti, float* out, long n)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  float* c[22];
  for (int q = 0; q < 22; q++) c[q] = out + q * n;
  if (!vti) {
    const float c11 = (float)((double)c[3][p] + 2.0 * (double)c[16][p]);
    c[1][p] = c11; c[7][p] = c11; c[12][p] = c11;
    c[2][p] = c[3][p]; c[8][p] = c[3][p];
    c[19][p] = c[16][p]; c[21][p] = c[16][p];
  } else {
    c[2][p] = (float)((double)c[1][p] - 2.0 * (double)c[21][p]);
    c[7][p] = c[1][p];
    c[8][p] = c[3][p];
    c[16][p] = c[19][p];
  }
  const int zeros[12] = {4, 9, 13, 5, 10, 14, 6, 11, 15, 17, 18, 20};
  for (int z = 0; z < 12; z++) c[zeros[z]][p] = 0.0f;
}

// ---- binary volumes (media_bin2model.cpp) --------------------------------
struct BinModel {
  const float* vol;          // [nvol][nz][ny][nx]
  const float *xv, *yv, *zv; // the windowed mesh vectors (z may decrease)
  int nx, ny, nz, nvol;
};

// TrilinearInterpolation (media_utility.cpp:304-401): the bracketing
// samples by floor / ceil of the offset over the spacing, clamped to the
// mesh, then linear in x, y, z in turn (a direction whose two samples
// coincide taken as is)
__device__ float bin_trilinear(const BinModel& B, const float* v, float xq, float yq, float zq)
{
  const float* x = B.xv;
  const float* y = B.yv;
  const float* z = B.zv;
  const int nx = B.nx, ny = B.ny, nz = B.nz;
  const float dx = x[1] - x[0], dy = y[1] - y[0], dz = z[1] - z[0];
  int ix0 = (int)floor((double)((xq - x[0]) / dx)), iy0 = (int)floor((double)((yq - y[0]) / dy));
  int ix2 = (int)ceil((double)((xq - x[0]) / dx)), iy2 = (int)ceil((double)((yq - y[0]) / dy));
  ix0 = ix0 < 0 ? 0 : ix0;
  ix2 = ix2 < 0 ? 0 : ix2;
  ix0 = ix0 > nx - 1 ? nx - 1 : ix0;
  ix2 = ix2 > nx - 1 ? nx - 1 : ix2;
  iy0 = iy0 < 0 ? 0 : iy0;
  iy2 = iy2 < 0 ? 0 : iy2;
  iy0 = iy0 > ny - 1 ? ny - 1 : iy0;
  iy2 = iy2 > ny - 1 ? ny - 1 : iy2;
  int iz0 = (int)floor((double)((zq - z[0]) / dz)), iz2 = (int)ceil((double)((zq - z[0]) / dz));
  iz0 = iz0 < 0 ? 0 : iz0;
  iz0 = iz0 > nz - 1 ? nz - 1 : iz0;
  iz2 = iz2 > nz - 1 ? nz - 1 : iz2;
  iz2 = iz2 < 0 ? 0 : iz2;
  const long line = nx, slice = (long)nx * ny;
  if (ix0 == ix2 && iy0 == iy2 && iz0 == iz2) return v[ix0 + iy0 * line + iz0 * slice];
  const float v000 = v[ix0 + iy0 * line + iz0 * slice], v002 = v[ix0 + iy0 * line + iz2 * slice];
  const float v200 = v[ix2 + iy0 * line + iz0 * slice], v202 = v[ix2 + iy0 * line + iz2 * slice];
  const float v020 = v[ix0 + iy2 * line + iz0 * slice], v022 = v[ix0 + iy2 * line + iz2 * slice];
  const float v220 = v[ix2 + iy2 * line + iz0 * slice], v222 = v[ix2 + iy2 * line + iz2 * slice];
  float v100, v102, v122, v120;
  if (ix0 == ix2) {
    v100 = v000; v102 = v002; v122 = v022; v120 = v020;
  } else {
    const float w = (xq - x[ix0]) / (x[ix2] - x[ix0]);
    v100 = v000 + w * (v200 - v000);
    v102 = v002 + w * (v202 - v002);
    v122 = v022 + w * (v222 - v022);
    v120 = v020 + w * (v220 - v020);
  }
  float v110, v112;
  if (iy0 == iy2) {
    v110 = v100; v112 = v102;
  } else {
    const float w = (yq - y[iy0]) / (y[iy2] - y[iy0]);
    v110 = v100 + w * (v120 - v100);
    v112 = v102 + w * (v122 - v102);
  }
  if (iz0 == iz2) return v110;
  return v110 + (zq - z[iz0]) / (z[iz2] - z[iz0]) * (v112 - v110);
}

// every point's samples: rho, vp, vs into rho, lambda, mu
// (parameterization_bin_el_iso_loc :185-234), or Qp, Qs as they are
__global__ void __launch_bounds__(256)
bin_loc_kernel(BinModel B, const float* x, const float* y, const float* z, long n, float* out)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= n) return;
  const long v3 = (long)B.nx * B.ny * B.nz;
  float s[3] = {0.0f, 0.0f, 0.0f};
  for (int q = 0; q < B.nvol; q++) s[q] = bin_trilinear(B, B.vol + q * v3, x[p], y[p], z[p]);
  if (B.nvol == 3) {
    const float rho = s[0], vp = s[1], vs = s[2];
    const float mu = vs * vs * rho;
    out[2 * n + p] = mu;
    out[n + p] = (float)((double)(vp * vp * rho) - 2.0 * (double)mu);
    out[p] = rho;
  } else {
    out[p] = s[0];
    out[n + p] = s[1];
  }
}

// the smallest and largest of n floats (ordered-integer keys, one
// atomicMin / atomicMax per block)
__device__ __forceinline__ unsigned float_key(float f)
{
  const unsigned b = __float_as_uint(f);
  return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

__global__ void __launch_bounds__(256) minmax_kernel(const float* v, long n, unsigned* mm)
{
  __shared__ unsigned lo[256], hi[256];
  unsigned a = 0xffffffffu, b = 0u;
  for (long p = (long)blockIdx.x * 256 + threadIdx.x; p < n; p += (long)gridDim.x * 256) {
    const unsigned k = float_key(v[p]);
    a = min(a, k);
    b = max(b, k);
  }
  lo[threadIdx.x] = a;
  hi[threadIdx.x] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      lo[threadIdx.x] = min(lo[threadIdx.x], lo[threadIdx.x + s]);
      hi[threadIdx.x] = max(hi[threadIdx.x], hi[threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomicMin(&mm[0], lo[0]);
    atomicMax(&mm[1], hi[0]);
  }
}

inline float float_unkey(unsigned k)
{
  const unsigned b = (k & 0x80000000u) ? (k & 0x7fffffffu) : ~k;
  float f;
  memcpy(&f, &b, sizeof f);
  return f;
}

}  // namespace cgfdh
