// media_kernels.h -- the layered model discretised onto the grid on the
// device with local values (media_layer2model_el_iso / _onecmp with the
// "loc" method, media/media_layer2model.cpp:47-190, 422-560, 1239-1383): one
// thread per grid point interpolates every interface's elevation at the
// point (BilinearInterpolation, media_utility.cpp:173-210), finds the layer
// holding it (findLastGreaterEqualIndex :88-113) and evaluates the layer's
// power law value + dz^pow * grad for each property, averaging the two
// layers' values on an interface.  Same float/double arithmetic as the
// reference (the power law in double); results agree with it to the last
// bit except where two math libraries' pow differ.
#pragma once

#include <cfloat>

namespace cgfdh {

constexpr int MEDIA_MAX_IF = 64;

struct LayerModel {
  const float* f;        // [nf][ni][my][mx]: elevation, then value/grad/pow per property
  int ni, mx, my, nprop;
  float minx, miny, dx, dy;
  const float* x;        // the grid, one point each
  const float* y;
  const float* z;
  long n;
  float* a;              // var, or lambda
  float* b;              // mu
  float* c;              // rho
  unsigned long long* bad;   // first failing point: code << 48 | index
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

__global__ void __launch_bounds__(256) layer2model_kernel(LayerModel M)
{
  const long p = (long)blockIdx.x * 256 + threadIdx.x;
  if (p >= M.n) return;
  const float x = M.x[p], y = M.y[p], z = M.z[p];
  const float maxx = M.minx + (M.mx - 1) * M.dx, maxy = M.miny + (M.my - 1) * M.dy;
  if (x < M.minx || x > maxx || y < M.miny || y > maxy) {
    atomicMin(M.bad, (1ull << 48) | (unsigned long long)p);
    return;
  }
  const long sl = (long)M.mx * M.my, vol = (long)M.ni * sl;
  float e[MEDIA_MAX_IF];
  for (int n = 0; n < M.ni; n++) e[n] = layer_bilinear(M, M.f + n * sl, x, y);
  float out[3];
  int mi = layer_last_ge(z, e, M.ni);
  if (mi == -1) {
    // above the top interface: the top layer of non-zero thickness
    mi = layer_last_ge(e[0], e, M.ni);
    for (int q = 0; q < M.nprop; q++)
      out[q] = layer_bilinear(M, M.f + (1 + 3 * q) * vol + mi * sl, x, y);
  } else {
    const float dz = e[mi] - z;
    for (int q = 0; q < M.nprop; q++) out[q] = layer_law(M, vol, q, mi, x, y, dz);
    if (M.nprop == 3 && (double)out[1] <= (double)out[2] * sqrt(2.0)) {
      atomicMin(M.bad, (2ull << 48) | (unsigned long long)p);
      return;
    }
    // on an interface: the mean of the two layers' values there
    if (fabsf(dz - 0.0f) < FLT_EPSILON) {
      const int m2 = layer_first_ge(z, e, M.ni) - 1;
      if (m2 >= 0) {
        const float dz2 = e[m2] - z;
        for (int q = 0; q < M.nprop; q++) {
          const float top = layer_law(M, vol, q, m2, x, y, dz2);
          out[q] = (float)(((double)out[q] + (double)top) / 2.0);
        }
      }
    }
  }
  if (M.nprop == 1) {
    M.a[p] = out[0];
  } else {
    const float vp = out[1], vs = out[2], rho = out[0];
    const float mu = vs * vs * rho;
    M.b[p] = mu;
    M.a[p] = (float)((double)(vp * vp * rho) - 2.0 * (double)mu);
    M.c[p] = rho;
  }
}

}  // namespace cgfdh
