/*
 * media_oracle.c -- CPU restatement of the reference's layered-model
 * discretisation (media/media_layer2model.cpp): the point values of every
 * media type (AssignLayerMediaPara2Point / CalPointValue_layer :422-1238),
 * their conversion to the solver's coefficients (para2vti / para2tti /
 * BondTransform, media_utility.cpp:835-1100), the local-value method "loc"
 * (:1239-1515) and the volume-averaging methods "har" and "ari" on a
 * curvilinear grid (:1515-2830 with GenerateHalfGrid, MarkInterfaceNumber,
 * GenerateHalfMesh, MeshSubdivide and NumOfValues, media_utility.cpp:
 * 600-830), plus the element-wise expansion media_layer2model_el_aniso
 * applies to isotropic and VTI types (:249-420).
 * TEST INFRASTRUCTURE ONLY: the checker of the engine's cgfdh_layer2model*;
 * nothing in the product links it.  Pinned to the reference's own functions,
 * compiled from the media library into oracle/_ref
 * (tests/test_media_oracle.py).  Every expression keeps the reference's
 * float/double mix and order (the power law, cos and sin in double, as the
 * C++ code's calls of two floats resolve to the double functions); compiled
 * without FMA contraction.
 */
#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

/* media type codes of the .md3lay header (media_utility.hpp:37-45) */
enum {
  MT_ONE = 0, MT_AC = 10, MT_ISO = 21, MT_PREM = 22, MT_THOMSEN = 23, MT_VCIJ = 24,
  MT_TTHOMSEN = 25, MT_TBOND = 26, MT_ACIJ = 27
};
/* targets: the reference entry point being restated */
enum { TG_ONECMP = 0, TG_ISO = 1, TG_VTI = 2, TG_ANISO = 3 };
/* methods */
enum { ME_LOC = 0, ME_HAR = 1, ME_ARI = 2 };

#define NG 8
#define NSG ((NG + 1) * (NG + 1) * (NG + 1))

/* the interface mesh: ni interfaces of mx x my samples from (minx, miny) */
typedef struct {
  int ni, mx, my;
  float minx, miny, dx, dy;
} mesh;

/* (value, gradient, power) triples per sample of each media type */
int orc_media_nprop(int mt)
{
  switch (mt) {
  case MT_ONE: return 1;
  case MT_AC: return 2;
  case MT_ISO: return 3;
  case MT_PREM: case MT_THOMSEN: case MT_VCIJ: return 6;
  case MT_TTHOMSEN: case MT_TBOND: return 8;
  case MT_ACIJ: return 22;
  }
  return -1;
}

/* coefficients a point carries before any target expansion: var; lambda,
 * mu, rho; c11 c33 c55 c66 c13 rho; c11..c66 (21) then rho */
static int native_count(int mt)
{
  switch (mt) {
  case MT_ONE: return 1;
  case MT_ISO: return 3;
  case MT_PREM: case MT_THOMSEN: case MT_VCIJ: return 6;
  case MT_TTHOMSEN: case MT_TBOND: case MT_ACIJ: return 22;
  }
  return -1;
}

/* BilinearInterpolation: cell from the uniform spacing, the last row or
 * column takes the sample itself; FLT_MAX left of or below the mesh */
static float bilinear(const mesh *m, const float *v, float xq, float yq)
{
  float x0 = m->minx + m->dx * 0, x1 = m->minx + m->dx * 1;
  float y0 = m->miny + m->dy * 0, y1 = m->miny + m->dy * 1;
  int ix = (xq - x0) / (x1 - x0);
  int iy = (yq - y0) / (y1 - y0);
  if (ix < 0 || iy < 0) return FLT_MAX;
  int indx = iy * m->mx + ix;
  if (ix == m->mx - 1 || iy == m->my - 1) return v[indx];
  float xa = m->minx + m->dx * ix, xb = m->minx + m->dx * (ix + 1);
  float ya = m->miny + m->dy * iy, yb = m->miny + m->dy * (iy + 1);
  float area = (xb - xa) * (yb - ya);
  float vq = v[indx] * (xb - xq) * (yb - yq) + v[indx + 1] * (xq - xa) * (yb - yq) +
             v[indx + m->mx] * (xb - xq) * (yq - ya) + v[indx + 1 + m->mx] * (xq - xa) * (yq - ya);
  vq /= area;
  return vq;
}

static int is_equal(float a, float b) { return fabs(a - b) < FLT_EPSILON; }

/* findLastGreaterEqualIndex over elevations ordered top to bottom */
static int last_ge(float value, const float *x, int n)
{
  if (value > x[0]) return -1;
  float dist = FLT_MAX;
  int indx = -1;
  for (int i = 0; i < n; i++) {
    if (x[i] == -FLT_MAX) continue;
    float d = x[i] - value;
    if (d < 0) break;
    if (d >= 0 && d <= dist) { dist = d; indx = i; }
  }
  return indx;
}

/* findFirstGreaterEqualIndex */
static int first_ge(float value, const float *x, int n)
{
  float dist = FLT_MAX;
  int idx = -1;
  for (int i = 0; i < n; i++) {
    if (x[i] == -FLT_MAX) continue;
    float d = x[i] - value;
    if (d >= 0 && d <= dist) {
      if (d <= 1e-6) return i;
      dist = d;
      idx = i;
    }
  }
  return idx;
}

/* value + dz^pow * grad of one property at layer mi, in double, rounded */
static float law(const mesh *m, const float *v, const float *g, const float *p, int mi, float x,
                 float y, float dz)
{
  size_t sl = (size_t)m->mx * m->my;
  float v0 = bilinear(m, v + mi * sl, x, y);
  float vg = bilinear(m, g + mi * sl, x, y);
  float vp = bilinear(m, p + mi * sl, x, y);
  return v0 + pow(dz, vp) * vg;
}

static int stale_dz(int mt)
{
  return mt == MT_VCIJ || mt == MT_TTHOMSEN || mt == MT_TBOND || mt == MT_ACIJ || mt == MT_AC;
}

/* the point's property values (AssignLayerMediaPara2Point with
 * CalPointValue_layer): 0; 1 when it lies outside the mesh; 2 when an
 * isotropic point has vp <= sqrt(2) vs (the reference stops on both) */
static int point_var(int mt, int nprop, const mesh *m, const float *f, float x, float y,
                     float z, float *out)
{
  float maxx = m->minx + (m->mx - 1) * m->dx, maxy = m->miny + (m->my - 1) * m->dy;
  if (x < m->minx || x > maxx || y < m->miny || y > maxy) return 1;
  const size_t vol = (size_t)m->ni * m->mx * m->my, sl = (size_t)m->mx * m->my;
  float e[m->ni];
  for (int n = 0; n < m->ni; n++) e[n] = bilinear(m, f + n * sl, x, y);
  int mi = last_ge(z, e, m->ni);
  if (mi == -1) {
    mi = last_ge(e[0], e, m->ni);
    for (int q = 0; q < nprop; q++) out[q] = bilinear(m, f + (1 + 3 * q) * vol + mi * sl, x, y);
    return 0;
  }
  float dz = e[mi] - z;
  for (int q = 0; q < nprop; q++) {
    const float *v = f + (1 + 3 * q) * vol;
    out[q] = law(m, v, v + vol, v + 2 * vol, mi, x, y, dz);
  }
  if (mt == MT_ISO && out[1] <= out[2] * sqrt(2)) return 2;
  if (is_equal(dz, 0.0)) {
    int m2 = first_ge(z, e, m->ni) - 1;
    if (m2 >= 0) {
      /* the stiffness-tensor and TTI types keep the lower layer's dz for
       * the upper layer's value (their cases never recompute it,
       * media_layer2model.cpp:708-1238) */
      float dz2 = stale_dz(mt) ? dz : e[m2] - z;
      for (int q = 0; q < nprop; q++) {
        const float *v = f + (1 + 3 * q) * vol;
        float top = law(m, v, v + vol, v + 2 * vol, m2, x, y, dz2);
        out[q] = (out[q] + top) / 2.0;
      }
    }
  }
  return 0;
}

/* BondTransform (theta = dip, phi = azimuth in degrees): C = R C0 R^T with
 * the Bond matrix of the rotation, products summed in float in row order */
static void bond(const float c0[36], float theta, float phi, float c[36])
{
  const double pi = acos(-1);
  theta *= (pi / 180.0);
  phi *= (pi / 180.0);
  float a11 = cos(theta) * sin(phi), a12 = cos(phi), a13 = sin(theta) * sin(phi);
  float a21 = -cos(theta) * cos(phi), a22 = sin(phi), a23 = -sin(theta) * cos(phi);
  float a31 = -sin(theta), a32 = 0, a33 = cos(theta);
  float R[36] = {
      a11 * a11, a12 * a12, a13 * a13, 2 * a12 * a13, 2 * a13 * a11, 2 * a11 * a12,
      a21 * a21, a22 * a22, a23 * a23, 2 * a22 * a23, 2 * a23 * a21, 2 * a21 * a22,
      a31 * a31, a32 * a32, a33 * a33, 2 * a32 * a33, 2 * a33 * a31, 2 * a31 * a32,
      a21 * a31, a22 * a32, a23 * a33, a22 * a33 + a23 * a32, a21 * a33 + a23 * a31,
      a22 * a31 + a21 * a32,
      a31 * a11, a32 * a12, a33 * a13, a12 * a33 + a13 * a32, a13 * a31 + a11 * a33,
      a11 * a32 + a12 * a31,
      a11 * a21, a12 * a22, a13 * a23, a12 * a23 + a13 * a22, a13 * a21 + a11 * a23,
      a11 * a22 + a12 * a21};
  float T[36];
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      float s = 0;
      for (int k = 0; k < 6; k++) s += R[i * 6 + k] * c0[k * 6 + j];
      T[i * 6 + j] = s;
    }
  for (int i = 0; i < 6; i++)
    for (int j = 0; j < 6; j++) {
      float s = 0;
      for (int k = 0; k < 6; k++) s += T[i * 6 + k] * R[j * 6 + k];
      c[i * 6 + j] = s;
    }
}

/* upper triangle of a symmetric 6x6 in c11 c12 .. c66 order */
static void upper21(const float c[36], float *o)
{
  int n = 0;
  for (int i = 0; i < 6; i++)
    for (int j = i; j < 6; j++) o[n++] = c[i * 6 + j];
}

/* a transversely isotropic medium about z (c11 c13 c33 c44 c66) rotated by
 * dip and azimuth (para2tti's call of BondTransform) */
static void tti_rotate(float c11, float c13, float c33, float c44, float c66, float dip,
                       float azi, float *o)
{
  float c0[36] = {c11, c11 - 2 * c66, c13, 0, 0, 0,
                  c11 - 2 * c66, c11, c13, 0, 0, 0,
                  c13, c13, c33, 0, 0, 0,
                  0, 0, 0, c44, 0, 0,
                  0, 0, 0, 0, c44, 0,
                  0, 0, 0, 0, 0, c66};
  float c[36];
  bond(c0, dip, azi, c);
  upper21(c, o);
}

/* a point's coefficients from its property values (the conversion of the
 * el_iso loc loop, para2vti, para2tti) */
static void to_native(int mt, const float *var, float *o)
{
  switch (mt) {
  case MT_ONE:
    o[0] = var[0];
    break;
  case MT_ISO: {
    float vp = var[1], vs = var[2], rho = var[0];
    o[1] = vs * vs * rho;
    o[0] = vp * vp * rho - 2.0 * o[1];
    o[2] = rho;
  } break;
  case MT_PREM: {
    float vph = var[1], vpv = var[2], vsh = var[3], vsv = var[4], rho = var[0], eta = var[5];
    o[0] = vph * vph * rho;
    o[1] = vpv * vpv * rho;
    o[3] = vsh * vsh * rho;
    o[2] = vsv * vsv * rho;
    o[5] = rho;
    o[4] = eta * (o[0] - 2.0 * o[2]);
  } break;
  case MT_THOMSEN: {
    float rho = var[0], vp0 = var[1], vs0 = var[2], epsil = var[3], delta = var[4];
    float gamma = var[5];
    float c33 = vp0 * vp0 * rho, c55 = vs0 * vs0 * rho;
    o[5] = rho;
    o[1] = c33;
    o[2] = c55;
    o[3] = 2.0 * gamma * c55 + c55;
    o[0] = 2.0 * epsil * c33 + c33;
    o[4] = sqrt(2.0 * delta * c33 * (c33 - c55) + (c33 - c55) * (c33 - c55)) - c55;
  } break;
  case MT_VCIJ:
    o[0] = var[1]; o[1] = var[2]; o[2] = var[3]; o[3] = var[4]; o[4] = var[5]; o[5] = var[0];
    break;
  case MT_TTHOMSEN: {
    float rho = var[0], vp0 = var[1], vs0 = var[2], epsil = var[3], delta = var[4];
    float gamma = var[5];
    float c33 = vp0 * vp0 * rho;
    float c44 = vs0 * vs0 * rho;
    float c66 = 2.0 * gamma * c44 + c44;
    float c11 = 2.0 * epsil * c33 + c33;
    float c13 = sqrt(2.0 * delta * c33 * (c33 - c44) + (c33 - c44) * (c33 - c44)) - c44;
    o[21] = rho;
    tti_rotate(c11, c13, c33, c44, c66, var[7], var[6], o);
  } break;
  case MT_TBOND:
    o[21] = var[0];
    tti_rotate(var[1], var[5], var[2], var[3], var[4], var[7], var[6], o);
    break;
  case MT_ACIJ:
    for (int q = 0; q < 21; q++) o[q] = var[1 + q];
    o[21] = var[0];
    break;
  }
}

/* the native coefficients' places among the target's outputs (the
 * reference functions' argument order):
 *   onecmp: var;  el_iso: lambda mu rho (the grid model: rho lambda mu);
 *   el_vti: rho c11 c33 c55 c66 c13;
 *   el_aniso: rho c11 c12 c13 c14 c15 c16 c22 c23 c24 c25 c26 c33 c34 c35
 *             c36 c44 c45 c46 c55 c56 c66 */
static int native_slots(int mt, int target, int grid, int *slot)
{
  /* media_grid2model_el_iso takes rho, lambda, mu */
  static const int iso_in_grid_iso[3] = {1, 2, 0};
  static const int vti_in_vti[6] = {1, 2, 3, 4, 5, 0};
  static const int iso_in_aniso[3] = {3, 16, 0};            /* lambda -> c13, mu -> c44 */
  static const int vti_in_aniso[6] = {1, 12, 19, 21, 3, 0};  /* c11 c33 c55 c66 c13 rho */
  int n = native_count(mt);
  if (target == TG_ONECMP && mt == MT_ONE) { slot[0] = 0; return 1; }
  if (target == TG_ISO && mt == MT_ISO) {
    for (int q = 0; q < 3; q++) slot[q] = grid ? iso_in_grid_iso[q] : q;
    return 3;
  }
  if (target == TG_VTI && (mt == MT_PREM || mt == MT_THOMSEN || mt == MT_VCIJ)) {
    for (int q = 0; q < 6; q++) slot[q] = vti_in_vti[q];
    return 6;
  }
  if (target == TG_ANISO) {
    if (mt == MT_ISO) { for (int q = 0; q < 3; q++) slot[q] = iso_in_aniso[q]; return 3; }
    if (mt == MT_PREM || mt == MT_THOMSEN || mt == MT_VCIJ) {
      for (int q = 0; q < 6; q++) slot[q] = vti_in_aniso[q];
      return 6;
    }
    if (mt == MT_TTHOMSEN || mt == MT_TBOND || mt == MT_ACIJ) {
      for (int q = 0; q < 21; q++) slot[q] = 1 + q;
      slot[21] = 0;
      return 22;
    }
  }
  (void)n;
  return -1;
}

/* a sub-cell point's averaged coefficients: the reference's running sums
 * (float accumulators, the harmonic terms in double) in sub-point order */
/* the averaged coefficients of the sub-points with use[p] set (all when
 * use is NULL; the grid model leaves out sub-points on a layer boundary):
 * the reference's running sums (float accumulators, the harmonic terms in
 * double) in sub-point order, divided by the count.  A float sum divided by
 * the int count and by the count in double round alike (both operands are
 * floats), so one expression serves the layered and the grid model. */
static void average(int mt, int method, const float *nat, const unsigned char *use, float *o)
{
  int num = 0;
  for (int p = 0; p < NSG; p++) num += use ? use[p] != 0 : 1;
  if (mt == MT_ONE) {
    float s = 0.0;
    for (int p = 0; p < NSG; p++) {
      if (use && !use[p]) continue;
      if (method == ME_HAR) s += (1.0 / nat[p]);
      else s += nat[p];
    }
    o[0] = method == ME_HAR ? num * 1.0 / s : s / (num * 1.0);
  } else if (mt == MT_ISO) {
    if (method == ME_HAR) {
      float ari_rho = 0.0, har_kappa = 0.0, har_mu = 0.0;
      for (int p = 0; p < NSG; p++) {
        if (use && !use[p]) continue;
        float lambda = nat[3 * p], mu = nat[3 * p + 1], rho = nat[3 * p + 2];
        ari_rho += rho;
        har_kappa += (1.0 / (lambda + 2.0 / 3.0 * mu));
        har_mu += (1.0 / mu);
      }
      har_mu = num * 1.0 / har_mu;
      har_kappa = num * 1.0 / har_kappa;
      ari_rho = ari_rho / (num * 1.0);
      o[0] = har_kappa - 2.0 / 3.0 * har_mu;
      o[1] = har_mu;
      o[2] = ari_rho;
    } else {
      float ari_rho = 0.0, ari_lam = 0.0, ari_mu = 0.0;
      for (int p = 0; p < NSG; p++) {
        if (use && !use[p]) continue;
        ari_rho += nat[3 * p + 2];
        ari_lam += nat[3 * p];
        ari_mu += nat[3 * p + 1];
      }
      o[0] = ari_lam / (num * 1.0);
      o[1] = ari_mu / (num * 1.0);
      o[2] = ari_rho / (num * 1.0);
    }
  } else {
    /* VTI and general anisotropic: each stiffness harmonic or arithmetic,
     * the density arithmetic */
    const int nn = native_count(mt), nc = nn - 1;
    for (int q = 0; q < nc; q++) {
      float s = 0.0;
      for (int p = 0; p < NSG; p++) {
        if (use && !use[p]) continue;
        if (method == ME_HAR) s += (1.0 / nat[nn * p + q]);
        else s += nat[nn * p + q];
      }
      o[q] = method == ME_HAR ? 1.0 * num / s : s / num;
    }
    float r = 0.0;
    for (int p = 0; p < NSG; p++)
      if (!use || use[p]) r += nat[nn * p + nc];
    o[nc] = r / num;
  }
}

/* the twelve zero stiffnesses and copies media_layer2model_el_aniso
 * writes for an isotropic or VTI file (slots as native_slots) */
static void expand_aniso(int mt, float *const *out, size_t n)
{
  static const int zeros[12] = {4, 9, 13, 5, 10, 14, 6, 11, 15, 17, 18, 20};
  for (size_t i = 0; i < n; i++) {
    if (mt == MT_ISO) {
      out[1][i] = out[3][i] + 2.0 * out[16][i];
      out[7][i] = out[1][i]; out[12][i] = out[1][i];
      out[2][i] = out[3][i]; out[8][i] = out[3][i];
      out[19][i] = out[16][i]; out[21][i] = out[16][i];
    } else {
      out[2][i] = out[1][i] - 2.0 * out[21][i];
      out[7][i] = out[1][i];
      out[8][i] = out[3][i];
      out[16][i] = out[19][i];
    }
    for (int z = 0; z < 12; z++) out[zeros[z]][i] = 0.0;
  }
}

/* one point as the one-component or isotropic entry sees it (kept for the
 * point-list tests): nprop 1 or 3 values (var, or rho, vp, vs) */
int orc_layer_point(int nprop, const int *dims, const float *geo, const float *f, float x,
                    float y, float z, float *out)
{
  mesh m = {dims[0], dims[1], dims[2], geo[0], geo[1], geo[2], geo[3]};
  return point_var(nprop == 1 ? MT_ONE : MT_ISO, nprop, &m, f, x, y, z, out);
}

/* points one by one with "loc": one component into a[], or lambda, mu, rho
 * into a[], b[], c[]; 0, or the first failing point's code */
int orc_layer2model(int nprop, const int *dims, const float *geo, const float *f, size_t n,
                    const float *x, const float *y, const float *z, float *a, float *b,
                    float *c)
{
  for (size_t p = 0; p < n; p++) {
    float v[3];
    int rc = orc_layer_point(nprop, dims, geo, f, x[p], y[p], z[p], v);
    if (rc) return rc;
    if (nprop == 1) {
      a[p] = v[0];
    } else {
      float vp = v[1], vs = v[2], rho = v[0];
      b[p] = vs * vs * rho;
      a[p] = vp * vp * rho - 2.0 * b[p];
      c[p] = rho;
    }
  }
  return 0;
}

/* the half grid of a curvilinear grid (GenerateHalfGrid, GRID_CURV): the
 * mean of each cell's eight corners, the last plane of each direction the
 * grid itself */
static void half_grid(const float *g, float *h, int nx, int ny, int nz)
{
  const size_t line = nx, slice = (size_t)nx * ny, vol = slice * nz;
  for (size_t i = 0; i < vol; i++) h[i] = g[i];
  for (int k = 0; k < nz - 1; k++)
    for (int j = 0; j < ny - 1; j++)
      for (int i = 0; i < nx - 1; i++) {
        size_t p = i + j * line + k * slice;
        h[p] = (g[p] + g[p + 1] + g[p + line + 1] + g[p + line] + g[p + slice] + g[p + slice + 1] +
                g[p + slice + line + 1] + g[p + line + slice]) / 8.0;
      }
}

/* MarkInterfaceNumber (GRID_CURV): the layer each half-grid point lies in,
 * the top layer of non-zero thickness above the surface */
static int material(const mesh *m, const float *f, float x, float y, float z)
{
  const size_t sl = (size_t)m->mx * m->my;
  float e[m->ni];
  for (int n = 0; n < m->ni; n++) e[n] = bilinear(m, f + n * sl, x, y);
  int mi = last_ge(z, e, m->ni);
  if (mi == -1) mi = last_ge(e[0], e, m->ni);
  return mi;
}

/* NumOfValues: distinct layers among the eight, 0 if one lies above all */
static int num_values(const int *v, int ni)
{
  int tab[ni];
  for (int i = 0; i < ni; i++) tab[i] = 0;
  int num = 0;
  for (int i = 0; i < 8; i++) {
    if (v[i] == -1) return 0;
    if (tab[v[i]] == 0) num++;
    tab[v[i]]++;
  }
  return num;
}

/* the grid-level entry: media_layer2model_onecmp / el_iso / el_vti /
 * el_aniso on a curvilinear grid (nx*ny*nz points, x fastest) with "loc",
 * "har" or "ari"; out holds the target's arrays in the reference's argument
 * order.  0, or: 1 a point (or sub-point) outside the mesh, 2 an isotropic
 * vp <= sqrt(2) vs, 3 a media type the target does not take, 4 out of
 * memory */
int orc_layer2model_grid(int mt, int target, int method, const int *dims, const float *geo,
                         const float *f, int nx, int ny, int nz, const float *x, const float *y,
                         const float *z, float *const *out)
{
  mesh m = {dims[0], dims[1], dims[2], geo[0], geo[1], geo[2], geo[3]};
  int slot[22];
  const int nn = native_slots(mt, target, 0, slot);
  if (nn < 0 || method < 0 || method > 2) return 3;
  const int nprop = orc_media_nprop(mt);
  const size_t line = nx, slice = (size_t)nx * ny, n = slice * nz;
  float var[22], nat[22];
  for (size_t p = 0; p < n; p++) {
    int rc = point_var(mt, nprop, &m, f, x[p], y[p], z[p], var);
    if (rc) return rc;
    to_native(mt, var, nat);
    for (int q = 0; q < nn; q++) out[slot[q]][p] = nat[q];
  }
  if (method != ME_LOC) {
    float *hx = malloc(n * sizeof(float)), *hy = malloc(n * sizeof(float));
    float *hz = malloc(n * sizeof(float)), *sub = malloc((size_t)NSG * 22 * sizeof(float));
    int *mat = malloc(n * sizeof(int));
    if (!hx || !hy || !hz || !sub || !mat) {
      free(hx); free(hy); free(hz); free(sub); free(mat);
      return 4;
    }
    half_grid(x, hx, nx, ny, nz);
    half_grid(y, hy, nx, ny, nz);
    half_grid(z, hz, nx, ny, nz);
    for (size_t p = 0; p < n; p++) mat[p] = material(&m, f, hx[p], hy[p], hz[p]);
    int rc = 0;
    for (int j = 1; j < ny - 1 && !rc; j++)
      for (int k = 1; k < nz - 1 && !rc; k++)
        for (int i = 1; i < nx - 1 && !rc; i++) {
          const size_t p = i + j * line + k * slice;
          const size_t c[8] = {p - 1 - line - slice, p - line - slice, p - slice, p - 1 - slice,
                               p - 1 - line, p - line, p, p - 1};
          int v[8];
          for (int q = 0; q < 8; q++) v[q] = mat[c[q]];
          if (num_values(v, m.ni) <= 1) continue;
          /* the half-grid cell around the point, vertices in GenerateHalfMesh's order */
          const size_t hv[8] = {p - 1 - line - slice, p - 1 - slice, p - slice, p - line - slice,
                                p - 1 - line, p - 1, p, p - line};
          float V[8][3];
          for (int q = 0; q < 8; q++) {
            V[q][0] = hx[hv[q]]; V[q][1] = hy[hv[q]]; V[q][2] = hz[hv[q]];
          }
          const float ng = NG;
          for (int kk = 0; kk <= NG && !rc; kk++)
            for (int jj = 0; jj <= NG && !rc; jj++)
              for (int ii = 0; ii <= NG && !rc; ii++) {
                float P[3];
                for (int d = 0; d < 3; d++) {
                  float ly00 = V[0][d] + (V[1][d] - V[0][d]) / ng * jj;
                  float ly10 = V[3][d] + (V[2][d] - V[3][d]) / ng * jj;
                  float ly01 = V[4][d] + (V[5][d] - V[4][d]) / ng * jj;
                  float ly11 = V[7][d] + (V[6][d] - V[7][d]) / ng * jj;
                  float lxy0 = ly00 + (ly10 - ly00) / ng * ii;
                  float lxy1 = ly01 + (ly11 - ly01) / ng * ii;
                  P[d] = lxy0 + (lxy1 - lxy0) / ng * kk;
                }
                rc = point_var(mt, nprop, &m, f, P[0], P[1], P[2], var);
                if (!rc) to_native(mt, var, sub + (size_t)nn * (ii + jj * 9 + kk * 81));
              }
          if (rc) break;
          average(mt, method, sub, NULL, nat);
          for (int q = 0; q < nn; q++) out[slot[q]][p] = nat[q];
        }
    free(hx); free(hy); free(hz); free(sub); free(mat);
    if (rc) return rc;
  }
  if (target == TG_ANISO && (mt == MT_ISO || nn == 6)) expand_aniso(mt, out, n);
  return 0;
}

/* ---- the grid model (.md3grd; media/media_grid2model.cpp) -----------------
 * NI surfaces grouped into NL layers of ngz[l] surfaces, each sample the
 * surface's elevation and the property values (no gradients or powers);
 * fields [1 + nprop][NI][my][mx].  A point's value is interpolated linearly
 * in z between the two surfaces around it. */

/* the point's values and whether it lies on a layer boundary
 * (AssignGridMediaPara2Point / CalPointValue_grid :424-958): above the top
 * the top surface's values, below the last the last's; on a boundary
 * between layers the mean of the surfaces either side; 0, 1 outside the
 * mesh, 2 an isotropic vp <= sqrt(2) vs */
static int gpoint_var(int mt, int nprop, const mesh *m, const unsigned char *bound,
                      const float *f, float x, float y, float z, float *var, int *on_if)
{
  float maxx = m->minx + (m->mx - 1) * m->dx, maxy = m->miny + (m->my - 1) * m->dy;
  if (x < m->minx || x > maxx || y < m->miny || y > maxy) return 1;
  const int ni = m->ni;
  if (ni < 1) return 1;
  const size_t sl = (size_t)m->mx * m->my, vol = (size_t)ni * sl;
  float e[ni];
  for (int n = 0; n < ni; n++) e[n] = bilinear(m, f + n * sl, x, y);
  int mi = last_ge(z, e, ni);
  *on_if = mi >= 0 && bound[mi] && is_equal(z, e[mi]);
  float dis_r = 1.0 / 2.0;
  if (mi >= 0 && mi < ni - 1 && fabs(e[mi + 1] - e[mi]) > 1e-6)
    dis_r = (z - e[mi]) / (e[mi + 1] - e[mi]);
  if (mi == -1 || mi == ni - 1) {
    if (mi == -1) mi = last_ge(e[0], e, ni);
    for (int q = 0; q < nprop; q++) var[q] = bilinear(m, f + (1 + q) * vol + mi * sl, x, y);
    return 0;
  }
  if (mi > 0 && is_equal(e[mi], z) && bound[mi]) {
    mi--;
    dis_r = 1.0 / 2.0;
  }
  for (int q = 0; q < nprop; q++) {
    float v0 = bilinear(m, f + (1 + q) * vol + mi * sl, x, y);
    float v1 = bilinear(m, f + (1 + q) * vol + (mi + 1) * sl, x, y);
    var[q] = v0 + (v1 - v0) * dis_r;
  }
  if (mt == MT_ISO && var[1] <= sqrt(2) * var[2]) return 2;
  return 0;
}

/* LayerNumberAtPoint: the layer (0..NL-1) among the layers' top surfaces */
static int glayer_number(const mesh *m, int nl, const int *first, const float *f, float x,
                         float y, float z)
{
  const size_t sl = (size_t)m->mx * m->my;
  float e[nl];
  for (int l = 0; l < nl; l++) e[l] = bilinear(m, f + first[l] * sl, x, y);
  int mi = last_ge(z, e, nl);
  if (mi == -1) mi = last_ge(e[0], e, nl);
  return mi;
}

/* the grid-level entry of the grid model: media_grid2model_onecmp / el_iso
 * / el_vti / el_aniso on a curvilinear grid with "loc", "har" or "ari"
 * (codes as orc_layer2model_grid); dims = NI surfaces, mx, my; ngz[nl]
 * surfaces per layer.  Averaging leaves out the sub-points on a layer
 * boundary and needs two layers (with one the local values stay). */
int orc_grid2model_grid(int mt, int target, int method, const int *dims, const float *geo,
                        int nl, const int *ngz, const float *f, int nx, int ny, int nz,
                        const float *x, const float *y, const float *z, float *const *out)
{
  mesh m = {dims[0], dims[1], dims[2], geo[0], geo[1], geo[2], geo[3]};
  int slot[22];
  const int nn = native_slots(mt, target, 1, slot);
  if (nn < 0 || method < 0 || method > 2 || nl < 1) return 3;
  int first[nl];
  unsigned char bound[m.ni];
  memset(bound, 0, sizeof bound);
  int acc = 0;
  for (int l = 0; l < nl; l++) {
    first[l] = acc;
    acc += ngz[l];
    if (l < nl - 1 && acc < m.ni) bound[acc] = 1;
  }
  if (acc != m.ni) return 3;
  const int nprop = orc_media_nprop(mt);
  const size_t line = nx, slice = (size_t)nx * ny, n = slice * nz;
  float var[22], nat[22];
  int on_if;
  for (size_t p = 0; p < n; p++) {
    int rc = gpoint_var(mt, nprop, &m, bound, f, x[p], y[p], z[p], var, &on_if);
    if (rc) return rc;
    to_native(mt, var, nat);
    for (int q = 0; q < nn; q++) out[slot[q]][p] = nat[q];
  }
  if (method != ME_LOC && nl >= 2) {
    float *hx = malloc(n * sizeof(float)), *hy = malloc(n * sizeof(float));
    float *hz = malloc(n * sizeof(float)), *sub = malloc((size_t)NSG * 22 * sizeof(float));
    int *mat = malloc(n * sizeof(int));
    unsigned char use[NSG];
    if (!hx || !hy || !hz || !sub || !mat) {
      free(hx); free(hy); free(hz); free(sub); free(mat);
      return 4;
    }
    half_grid(x, hx, nx, ny, nz);
    half_grid(y, hy, nx, ny, nz);
    half_grid(z, hz, nx, ny, nz);
    for (size_t p = 0; p < n; p++) mat[p] = glayer_number(&m, nl, first, f, hx[p], hy[p], hz[p]);
    int rc = 0;
    for (int j = 1; j < ny - 1 && !rc; j++)
      for (int k = 1; k < nz - 1 && !rc; k++)
        for (int i = 1; i < nx - 1 && !rc; i++) {
          const size_t p = i + j * line + k * slice;
          const size_t c[8] = {p - 1 - line - slice, p - line - slice, p - slice, p - 1 - slice,
                               p - 1 - line, p - line, p, p - 1};
          int v[8];
          for (int q = 0; q < 8; q++) v[q] = mat[c[q]];
          if (num_values(v, nl) <= 1) continue;
          const size_t hv[8] = {p - 1 - line - slice, p - 1 - slice, p - slice, p - line - slice,
                                p - 1 - line, p - 1, p, p - line};
          float V[8][3];
          for (int q = 0; q < 8; q++) {
            V[q][0] = hx[hv[q]]; V[q][1] = hy[hv[q]]; V[q][2] = hz[hv[q]];
          }
          const float ng = NG;
          for (int kk = 0; kk <= NG && !rc; kk++)
            for (int jj = 0; jj <= NG && !rc; jj++)
              for (int ii = 0; ii <= NG && !rc; ii++) {
                float P[3];
                for (int d = 0; d < 3; d++) {
                  float ly00 = V[0][d] + (V[1][d] - V[0][d]) / ng * jj;
                  float ly10 = V[3][d] + (V[2][d] - V[3][d]) / ng * jj;
                  float ly01 = V[4][d] + (V[5][d] - V[4][d]) / ng * jj;
                  float ly11 = V[7][d] + (V[6][d] - V[7][d]) / ng * jj;
                  float lxy0 = ly00 + (ly10 - ly00) / ng * ii;
                  float lxy1 = ly01 + (ly11 - ly01) / ng * ii;
                  P[d] = lxy0 + (lxy1 - lxy0) / ng * kk;
                }
                const int sp = ii + jj * 9 + kk * 81;
                rc = gpoint_var(mt, nprop, &m, bound, f, P[0], P[1], P[2], var, &on_if);
                if (!rc) {
                  to_native(mt, var, sub + (size_t)nn * sp);
                  use[sp] = !on_if;
                }
              }
          if (rc) break;
          average(mt, method, sub, use, nat);
          for (int q = 0; q < nn; q++) out[slot[q]][p] = nat[q];
        }
    free(hx); free(hy); free(hz); free(sub); free(mat);
    if (rc) return rc;
  }
  if (target == TG_ANISO && (mt == MT_ISO || nn == 6)) expand_aniso(mt, out, n);
  return 0;
}

/* ---- binary volumes (media/media_bin2model.cpp) ---------------------------
 * Three float32 volumes (rho, vp, vs; or Qp, Qs) on a regular mesh whose
 * axes are stored in bin_order (0 x, 1 y, 2 z; the first fastest), sampled
 * at each grid point by TrilinearInterpolation (media_utility.cpp:304-401).
 * The window read and the mesh vectors follow media_bin2model_el_iso
 * (:27-183): ix_start = (xmin - x0) / dx truncated (0 below the origin),
 * ix_end = ceil((xmax - x0) / dx) clipped to the last sample, y alike, every
 * z sample; the vectors x0 + (ix_start + i) dx, z0 + k dz. */

/* TrilinearInterpolation over vectors x[nx], y[ny], z[nz] (z may
 * decrease); v[nz][ny][nx] */
static float trilinear(const float *x, int nx, const float *y, int ny, const float *z, int nz,
                       const float *v, float xq, float yq, float zq)
{
  float dx = x[1] - x[0], dy = y[1] - y[0], dz = z[1] - z[0];
  int ix0 = floor((xq - x[0]) / dx), iy0 = floor((yq - y[0]) / dy);
  int ix2 = ceil((xq - x[0]) / dx), iy2 = ceil((yq - y[0]) / dy);
  ix0 = ix0 < 0 ? 0 : ix0;
  ix2 = ix2 < 0 ? 0 : ix2;
  ix0 = ix0 > nx - 1 ? nx - 1 : ix0;
  ix2 = ix2 > nx - 1 ? nx - 1 : ix2;
  iy0 = iy0 < 0 ? 0 : iy0;
  iy2 = iy2 < 0 ? 0 : iy2;
  iy0 = iy0 > ny - 1 ? ny - 1 : iy0;
  iy2 = iy2 > ny - 1 ? ny - 1 : iy2;
  int iz0 = floor((zq - z[0]) / dz), iz2 = ceil((zq - z[0]) / dz);
  iz0 = iz0 < 0 ? 0 : iz0;
  iz0 = iz0 > nz - 1 ? nz - 1 : iz0;
  iz2 = iz2 > nz - 1 ? nz - 1 : iz2;
  iz2 = iz2 < 0 ? 0 : iz2;
  const size_t line = nx, slice = (size_t)nx * ny;
  if (ix0 == ix2 && iy0 == iy2 && iz0 == iz2) return v[ix0 + iy0 * line + iz0 * slice];
  float v000 = v[ix0 + iy0 * line + iz0 * slice], v002 = v[ix0 + iy0 * line + iz2 * slice];
  float v200 = v[ix2 + iy0 * line + iz0 * slice], v202 = v[ix2 + iy0 * line + iz2 * slice];
  float v020 = v[ix0 + iy2 * line + iz0 * slice], v022 = v[ix0 + iy2 * line + iz2 * slice];
  float v220 = v[ix2 + iy2 * line + iz0 * slice], v222 = v[ix2 + iy2 * line + iz2 * slice];
  float v100, v102, v122, v120;
  if (ix0 == ix2) {
    v100 = v000; v102 = v002; v122 = v022; v120 = v020;
  } else {
    float w = (xq - x[ix0]) / (x[ix2] - x[ix0]);
    v100 = v000 + w * (v200 - v000);
    v102 = v002 + w * (v202 - v002);
    v122 = v022 + w * (v222 - v022);
    v120 = v020 + w * (v220 - v020);
  }
  float v110, v112;
  if (iy0 == iy2) {
    v110 = v100; v112 = v102;
  } else {
    float w = (yq - y[iy0]) / (y[iy2] - y[iy0]);
    v110 = v100 + w * (v120 - v100);
    v112 = v102 + w * (v122 - v102);
  }
  if (iz0 == iz2) return v110;
  return v110 + (zq - z[iz0]) / (z[iz2] - z[iz0]) * (v112 - v110);
}

/* the read window of a binary model over the grid's extent: start/end per
 * stored axis, the mesh vectors (x, y windowed; z whole) and their lengths;
 * 0, or 5 when the model misses the grid in x, y or z (the reference stops) */
int orc_bin_window(const int *order, const int *size, const float *spacing, const float *origin,
                   float xmin, float xmax, float ymin, float ymax, float zmin, float zmax,
                   int *start, int *end, float *xv, float *yv, float *zv, int *n3)
{
  int dim[3] = {-1, -1, -1};
  for (int i = 0; i < 3; i++) {
    if (order[i] < 0 || order[i] > 2) return 5;
    dim[order[i]] = i;
  }
  if (dim[0] < 0 || dim[1] < 0 || dim[2] < 0) return 5;
  const int dx_ = dim[0], dy_ = dim[1], dz_ = dim[2];
  float x0 = origin[dx_], y0 = origin[dy_], z0 = origin[dz_];
  float dx = spacing[dx_], dy = spacing[dy_], dz = spacing[dz_];
  int ix_start = (xmin < x0 ? 0 : (xmin - x0) / dx);
  if (ix_start >= size[dx_] || xmax < x0) return 5;
  int ix_end = ceil((xmax - x0) / dx);
  if (ix_end >= size[dx_]) ix_end = size[dx_] - 1;
  int iy_start = (ymin < y0 ? 0 : (ymin - y0) / dy);
  if (iy_start >= size[dy_] || ymax < y0) return 5;
  int iy_end = ceil((ymax - y0) / dy);
  if (iy_end >= size[dy_]) iy_end = size[dy_] - 1;
  float zn = z0 + dz * (size[dz_] - 1);
  if (zmin >= fmaxf(zn, z0) || zmax <= fminf(z0, zn)) return 5;
  n3[0] = ix_end - ix_start + 1;
  n3[1] = iy_end - iy_start + 1;
  n3[2] = size[dz_];
  if (xv) for (int i = 0; i < n3[0]; i++) xv[i] = x0 + (ix_start + i) * dx;
  if (yv) for (int i = 0; i < n3[1]; i++) yv[i] = y0 + (iy_start + i) * dy;
  if (zv) for (int i = 0; i < n3[2]; i++) zv[i] = z0 + i * dz;
  start[dx_] = ix_start; end[dx_] = ix_end;
  start[dy_] = iy_start; end[dy_] = iy_end;
  start[dz_] = 0; end[dz_] = n3[2] - 1;
  return 0;
}

/* media_bin2model_el_iso (nvol 3: rho, vp, vs into rho, lambda, mu) or
 * media_bin2model_vis_iso (nvol 2: Qp, Qs as they are) on a curvilinear
 * grid, from windowed volumes vol[nvol][nz][ny][nx] on the vectors */
int orc_bin2model_grid(int nvol, const int *n3, const float *xv, const float *yv,
                       const float *zv, const float *vol, size_t npts, const float *x,
                       const float *y, const float *z, float *const *out)
{
  const size_t v3 = (size_t)n3[0] * n3[1] * n3[2];
  for (size_t p = 0; p < npts; p++) {
    float s[3] = {0, 0, 0};
    for (int q = 0; q < nvol; q++)
      s[q] = trilinear(xv, n3[0], yv, n3[1], zv, n3[2], vol + q * v3, x[p], y[p], z[p]);
    if (nvol == 3) {
      float rho = s[0], vp = s[1], vs = s[2];
      out[2][p] = vs * vs * rho;
      out[1][p] = vp * vp * rho - 2.0 * out[2][p];
      out[0][p] = rho;
    } else {
      out[0][p] = s[0];
      out[1][p] = s[1];
    }
  }
  return 0;
}
