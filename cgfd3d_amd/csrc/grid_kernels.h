// grid_kernels.h -- the curvilinear grid generated from a layer file on the
// device (gd_curv_gen_layer, gd_t.c:987-1480): the interfaces resampled in
// x and y (every |f|-th sample, or a cubic spline through the samples), then
// one thread per grid column builds its z values through the layers
// (gd_grid_z_interp, gd_t.c:1483-1562) and x, y as cubic splines of z, with
// three ghost points mirrored above the top.  The spline is the reference's
// (gd_SPLine / gd_seval, gd_t.c:1564-1700: the Forsythe-Malcolm-Moler
// method with its end conditions and interval search), in the same
// float/double arithmetic, so the grid is bit-identical to the reference's.
#pragma once

namespace cgfdh {

constexpr int GRID_MAX_IF = 64;   // interfaces a column handles

// the spline coefficients of y(x) at n knots (gd_SPLine, end1 = end2 = 0);
// false for fewer than 2 or unordered knots
__device__ bool spline_coef(int n, const float* x, const float* y, float* b, float* c, float* d)
{
  if (n < 2) return false;
  for (int i = 1; i < n; ++i)
    if (x[i] <= x[i - 1]) return false;
  const int nm1 = n - 1;
  if (n >= 3) {
    d[0] = x[1] - x[0];
    c[1] = (y[1] - y[0]) / d[0];
    for (int i = 1; i < nm1; ++i) {
      d[i] = x[i + 1] - x[i];
      b[i] = (float)(2.0 * (double)(d[i - 1] + d[i]));
      c[i + 1] = (y[i + 1] - y[i]) / d[i];
      c[i] = c[i + 1] - c[i];
    }
    b[0] = -d[0];
    b[nm1] = -d[n - 2];
    c[0] = 0.0f;
    c[nm1] = 0.0f;
    if (n != 3) {
      c[0] = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0]);
      c[nm1] = c[n - 2] / (x[nm1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4]);
      c[0] = c[0] * d[0] * d[0] / (x[3] - x[0]);
      c[nm1] = -c[nm1] * d[n - 2] * d[n - 2] / (x[nm1] - x[n - 4]);
    }
    for (int i = 1; i < n; ++i) {
      const float t = d[i - 1] / b[i - 1];
      b[i] = b[i] - t * d[i - 1];
      c[i] = c[i] - t * c[i - 1];
    }
    c[nm1] = c[nm1] / b[nm1];
    for (int ib = 0; ib < nm1; ++ib) {
      const int i = n - ib - 2;
      c[i] = (c[i] - d[i] * c[i + 1]) / b[i];
    }
    b[nm1] = (float)((double)((y[nm1] - y[n - 2]) / d[n - 2]) +
                     (double)d[n - 2] * ((double)c[n - 2] + 2.0 * (double)c[nm1]));
    for (int i = 0; i < nm1; ++i) {
      b[i] = (float)((double)((y[i + 1] - y[i]) / d[i]) -
                     (double)d[i] * ((double)c[i + 1] + 2.0 * (double)c[i]));
      d[i] = (c[i + 1] - c[i]) / d[i];
      c[i] = (float)(3.0 * (double)c[i]);
    }
    c[nm1] = (float)(3.0 * (double)c[nm1]);
    d[nm1] = d[n - 2];
  } else {
    b[0] = (y[1] - y[0]) / (x[1] - x[0]);
    c[0] = 0.0f;
    d[0] = 0.0f;
    b[1] = b[0];
    c[1] = 0.0f;
    d[1] = 0.0f;
  }
  return true;
}

// gd_seval: the interval kept from the last call unless u lies outside it
__device__ __forceinline__ float spline_eval(float u, int n, const float* x, const float* y,
                                             const float* b, const float* c, const float* d,
                                             int* last)
{
  int i = *last;
  if (i >= n - 1) i = 0;
  if (i < 0) i = 0;
  if ((x[i] > u) || (x[i + 1] < u)) {
    int j = n;
    i = 0;
    do {
      const int k = (i + j) / 2;
      if (u < x[k]) j = k;
      if (u >= x[k]) i = k;
    } while (j > i + 1);
  }
  *last = i;
  float w = u - x[i];
  w = y[i] + w * (b[i] + w * (c[i] + w * d[i]));
  return w;
}

// one line of samples respaced: knots at j * f (j < n) read with stride
// sin, the spline at 0, 1, .., ni-1 written with stride sout.  Line q's
// knots, values and coefficients live in work[q * 5n ..]; lines come as
// (base_in, base_out) pairs.
struct SplineLines {
  const float* src;
  float* dst;
  const long* base;      // 2 per line: first input, first output offset
  long sin, sout;
  int n, ni, f, nline;
  float* work;
  int* bad;
};

__global__ void spline_lines_kernel(SplineLines S)
{
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= S.nline) return;
  float* x = S.work + (long)q * 5 * S.n;
  float* y = x + S.n;
  float* b = y + S.n;
  float* c = b + S.n;
  float* d = c + S.n;
  const float* in = S.src + S.base[2 * q];
  for (int j = 0; j < S.n; j++) {
    x[j] = (float)(j * S.f);
    y[j] = in[j * S.sin];
  }
  if (!spline_coef(S.n, x, y, b, c, d)) {
    atomicExch(S.bad, 1);
    return;
  }
  float* out = S.dst + S.base[2 * q + 1];
  int last = 0;
  for (int i = 0; i < S.ni; i++) out[i * S.sout] = spline_eval((float)i, S.n, x, y, b, c, d, &last);
}

// every |f|-th sample of the interfaces: L[c][k][jj][ii] = In[c][k][jj fy][ii fx]
__global__ void sample_layers_kernel(const float* in, float* L, int nxl, int nyl, int nif,
                                     int nxi, int nyi, int fx, int fy)
{
  const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long per = (long)nxi * nyi * nif;
  if (q >= 3 * per) return;
  const int c = (int)(q / per);
  const long r = q % per;
  const int ii = (int)(r % nxi), jj = (int)(r / nxi % nyi), kk = (int)(r / nxi / nyi);
  L[q] = in[(long)c * nxl * nyl * nif + (long)ii * fx + (long)jj * fy * nxl + (long)kk * nxl * nyl];
}

struct GridColumns {
  const float* L;        // resampled interfaces [3][nif][nyi][nxi]
  int nxi, nyi, nif;
  int ncell[GRID_MAX_IF], equal[GRID_MAX_IF];
  int xg, yg, zg;        // first resampled x, y index and first z index of the block
  int nx, ny, nz;        // the block with ghosts
  float* x3d;
  float* y3d;
  float* z3d;
  int* bad;
};

// the z values of one column, in order, as gd_grid_z_interp leaves them
struct ZWalk {
  const float* zl;
  const int* ncell;
  const int* equal;
  float dz[GRID_MAX_IF];
  int nlay;
  int layer, m;          // current layer and index inside it
  bool graded;
  int N1, N2;
  float s1, s2, r;

  __device__ void start_layer()
  {
    const int i = layer;
    graded = false;
    if (i < nlay && equal[i] < 1 && i > 0 && i < nlay - 1 &&
        (dz[i] - dz[i - 1]) * (dz[i + 1] - dz[i]) > 0) {
      graded = true;
      N1 = ((dz[i] - dz[i - 1]) / (dz[i + 1] - dz[i - 1])) * (ncell[i] + 2) + 1;
      if (N1 > ncell[i] - 1) N1 = ncell[i] - 1;
      N2 = ncell[i] + 2 - N1;
      const float dist = zl[i + 1] - zl[i] + dz[i - 1] + dz[i + 1];
      const float mid = (2 * dist - dz[i - 1] * N1 - dz[i + 1] * (N2 + 1)) / (N1 + N2 - 1);
      s1 = (mid - dz[i - 1]) / (N1 - 1);
      s2 = (dz[i + 1] - mid) / (N2);
    }
    m = 0;
  }
  __device__ void init(const float* z, const int* nc, const int* eq, int n)
  {
    zl = z; ncell = nc; equal = eq; nlay = n;
    for (int i = 0; i < nlay; i++) dz[i] = (zl[i + 1] - zl[i]) / ncell[i];
    layer = 0;
    start_layer();
  }
  // the next value (nz_interp of them in all)
  __device__ float next()
  {
    if (layer == nlay) return zl[nlay];
    const int i = layer;
    float v;
    if (!graded) {
      v = m == 0 ? zl[i] : zl[i] + m * dz[i];
    } else {
      if (m == 0) r = zl[i];
      else if (m < N1) r = r + dz[i - 1] + m * s1;
      else r = r + dz[i + 1] - (N2 + N1 - m - 1) * s2;
      v = r;
    }
    if (++m == ncell[i]) {
      ++layer;
      start_layer();
    }
    return v;
  }
};

__global__ void __launch_bounds__(64) grid_columns_kernel(GridColumns G)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.nx * G.ny) return;
  const int i = t % G.nx, j = t / G.nx;
  float xl[GRID_MAX_IF], yl[GRID_MAX_IF], zl[GRID_MAX_IF];
  float bx[GRID_MAX_IF], cx[GRID_MAX_IF], dx[GRID_MAX_IF];
  float by[GRID_MAX_IF], cy[GRID_MAX_IF], dy[GRID_MAX_IF];
  const long vit = (long)G.nxi * G.nyi * G.nif;
  for (int k = 0; k < G.nif; k++) {
    const long q = (long)(i + G.xg) + (long)(j + G.yg) * G.nxi + (long)k * G.nxi * G.nyi;
    xl[k] = G.L[q];
    yl[k] = G.L[vit + q];
    zl[k] = G.L[2 * vit + q];
  }
  if (!spline_coef(G.nif, zl, xl, bx, cx, dx) || !spline_coef(G.nif, zl, yl, by, cy, dy)) {
    atomicExch(G.bad, 1);
    return;
  }
  ZWalk W;
  W.init(zl, G.ncell, G.equal, G.nif - 1);
  int nzi = 1;
  for (int k = 0; k < G.nif - 1; k++) nzi += G.ncell[k];
  int lx = 0, ly = 0;
  float hx[4], hy[4], hz[4];   // the last four values, for the mirrored ghosts
  const long ss = (long)G.nx * G.ny, p0 = (long)i + (long)j * G.nx;
  for (int n = 0; n < nzi; n++) {
    const float z = W.next();
    const float x = spline_eval(z, G.nif, zl, xl, bx, cx, dx, &lx);
    const float y = spline_eval(z, G.nif, zl, yl, by, cy, dy, &ly);
    const int k = n - G.zg;
    if (k >= 0 && k < G.nz) {
      G.x3d[p0 + k * ss] = x;
      G.y3d[p0 + k * ss] = y;
      G.z3d[p0 + k * ss] = z;
    }
    hx[n & 3] = x; hy[n & 3] = y; hz[n & 3] = z;
  }
  // ghosts above the top: v[top + g] = 2 v[top] - v[top - g]
  const int top = nzi - 1;
  for (int g = 1; g <= 3; g++) {
    const int k = top + g - G.zg;
    if (k < 0 || k >= G.nz) continue;
    const int a = top & 3, b = (top - g) & 3;
    G.x3d[p0 + k * ss] = hx[a] * 2 - hx[b];
    G.y3d[p0 + k * ss] = hy[a] * 2 - hy[b];
    G.z3d[p0 + k * ss] = hz[a] * 2 - hz[b];
  }
}

}  // namespace cgfdh
