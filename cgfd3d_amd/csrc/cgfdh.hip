// cgfdh.hip -- gfx950 stage kernels and the C ABI declared in include/cgfdh.h.
//
// One RK4 stage is one fused launch: RHS (interior stencil, traction image,
// vlow, CFS-PML with its auxiliary RK combine, moment sources) and the RK
// combine of the wavefield, so the stage never writes an rhs level to HBM.
// Levels: pre, tmpA, tmpB, end; the stage reading tmpA writes tmpB and the
// other way round, because the stencil reads neighbours of the level it
// replaces.  PML aux needs only pre/tmp/end: it is read at its own point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "cgfdh.h"
#include "cgfdh_io.h"
#include "stage_launch.h"

using namespace cgfdh;

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------

// receivers: trilinear sample of w_end after a step (io_funcs.c:1585-1623),
// indx1d[5] used twice and indx1d[4] not, as the reference does
__global__ void recv_kernel(const float* __restrict__ w, long vol, int nrecv, int nw,
                            const int* __restrict__ idx, const float* __restrict__ dijk,
                            float* __restrict__ seismo, int nt_total, int it,
                            const int* __restrict__ it_off)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nrecv * nw) return;
  if (it_off) it += *it_off;
  const int r = t / nw, c = t % nw;
  const int* ix = idx + 8 * r;
  const float Lx2 = dijk[3 * r], Ly2 = dijk[3 * r + 1], Lz2 = dijk[3 * r + 2];
  const float Lx1 = (float)(1.0 - (double)Lx2);
  const float Ly1 = (float)(1.0 - (double)Ly2);
  const float Lz1 = (float)(1.0 - (double)Lz2);
  const float* v = w + (long)c * vol;
  const float s = v[ix[0]] * Lx1 * Ly1 * Lz1 + v[ix[1]] * Lx2 * Ly1 * Lz1 +
                  v[ix[2]] * Lx1 * Ly2 * Lz1 + v[ix[3]] * Lx2 * Ly2 * Lz1 +
                  v[ix[5]] * Lx1 * Ly1 * Lz2 + v[ix[5]] * Lx2 * Ly1 * Lz2 +
                  v[ix[6]] * Lx1 * Ly2 * Lz2 + v[ix[7]] * Lx2 * Ly2 * Lz2;
  seismo[((long)r * nw + c) * nt_total + it] = s;
}

// visco-elastic free surface after the last stage: the surface-row stresses
// rotated into the surface frame with the zeta-normal rows zeroed and back
// (sv_curv_col_vis_iso_free, sv_curv_col_vis_iso.c:510-623)
__global__ void vis_free_kernel(float* __restrict__ w, const float* __restrict__ matD, int ni,
                                int nj, int nx, long ss, long vol, int nk2)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  if (i > 3 + ni - 1) return;
  const long p = (long)i + (long)j * nx + (long)nk2 * ss;
  const long ij = ((long)i + (long)j * nx) * 9;
  float* Txx = w + TXX * vol;
  float* Tyy = w + TYY * vol;
  float* Tzz = w + TZZ * vol;
  float* Txy = w + TXY * vol;
  float* Txz = w + TXZ * vol;
  float* Tyz = w + TYZ * vol;
  float D[3][3], DT[3][3], Tl[3][3], DTl[3][3], Tg[3][3];
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      D[a][b] = matD[ij + 3 * a + b];
      DT[b][a] = D[a][b];
    }
  const float d11 = D[0][0], d12 = D[0][1], d13 = D[0][2];
  const float d21 = D[1][0], d22 = D[1][1], d23 = D[1][2];
  const float txx = Txx[p], tyy = Tyy[p], tzz = Tzz[p], txy = Txy[p], txz = Txz[p], tyz = Tyz[p];
  Tl[0][0] = d11 * d11 * txx + d12 * d12 * tyy + d13 * d13 * tzz +
             2 * (d11 * d12 * txy + d11 * d13 * txz + d12 * d13 * tyz);
  Tl[0][1] = d11 * d21 * txx + d12 * d22 * tyy + d13 * d23 * tzz + (d11 * d22 + d12 * d21) * txy +
             (d11 * d23 + d21 * d13) * txz + (d12 * d23 + d22 * d13) * tyz;
  Tl[1][1] = d21 * d21 * txx + d22 * d22 * tyy + d23 * d23 * tzz +
             2 * (d21 * d22 * txy + d21 * d23 * txz + d22 * d23 * tyz);
  Tl[1][0] = Tl[0][1];
  Tl[0][2] = 0.0f; Tl[1][2] = 0.0f; Tl[2][0] = 0.0f; Tl[2][1] = 0.0f; Tl[2][2] = 0.0f;
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      float c = 0.0f;
      for (int q = 0; q < 3; q++) c += DT[a][q] * Tl[q][b];
      DTl[a][b] = c;
    }
  for (int a = 0; a < 3; a++)
    for (int b = 0; b < 3; b++) {
      float c = 0.0f;
      for (int q = 0; q < 3; q++) c += DTl[a][q] * D[q][b];
      Tg[a][b] = c;
    }
  Txx[p] = Tg[0][0];
  Tyy[p] = Tg[1][1];
  Tzz[p] = Tg[2][2];
  Txy[p] = Tg[0][1];
  Txz[p] = Tg[0][2];
  Tyz[p] = Tg[1][2];
}

// peak ground motion on the free surface after a step (PG_calcu,
// wav_t.c:380-455): maps PGV,PGVh,PGVx,PGVy,PGVz, PGA.., PGD.. and the
// trapezoid displacement, from w_end and the state before the step
__global__ void pg_kernel(const float* __restrict__ w1, const float* __restrict__ w0,
                          float* __restrict__ PG, float* __restrict__ Dis, int ni, int nj,
                          int nx, long ss, long vol, int nk2, float dt)
{
  const int i = 3 + blockIdx.x * blockDim.x + threadIdx.x;
  const int j = 3 + blockIdx.y;
  if (i > 3 + ni - 1) return;
  const long q = (long)i + (long)j * nx;
  const long p = q + (long)nk2 * ss;
  const float vx1 = w1[p], vy1 = w1[vol + p], vz1 = w1[2 * vol + p];
  const float vx0 = w0[p], vy0 = w0[vol + p], vz0 = w0[2 * vol + p];
  const float Ax = fabsf((vx1 - vx0) / dt);
  const float Ay = fabsf((vy1 - vy0) / dt);
  const float Az = fabsf((vz1 - vz0) / dt);
  float* Dx = Dis;
  float* Dy = Dis + ss;
  float* Dz = Dis + 2 * ss;
  const float dx = (float)((double)Dx[q] + 0.5 * (double)(vx1 + vx0) * (double)dt);
  const float dy = (float)((double)Dy[q] + 0.5 * (double)(vy1 + vy0) * (double)dt);
  const float dz = (float)((double)Dz[q] + 0.5 * (double)(vz1 + vz0) * (double)dt);
  Dx[q] = dx; Dy[q] = dy; Dz[q] = dz;
  const float V = (float)sqrt((double)(vx1 * vx1 + vy1 * vy1 + vz1 * vz1));
  const float Vh = (float)sqrt((double)(vx1 * vx1 + vy1 * vy1));
  const float A = (float)sqrt((double)(Ax * Ax + Ay * Ay + Az * Az));
  const float Ah = (float)sqrt((double)(Ax * Ax + Ay * Ay));
  const float D = (float)sqrt((double)(dx * dx + dy * dy + dz * dz));
  const float Dh = (float)sqrt((double)(dx * dx + dy * dy));
  const float val[15] = {V, Vh, fabsf(vx1), fabsf(vy1), fabsf(vz1), A, Ah, Ax, Ay, Az,
                         D, Dh, fabsf(dx), fabsf(dy), fabsf(dz)};
#pragma unroll
  for (int m = 0; m < 15; m++)
    if (PG[m * ss + q] < val[m]) PG[m * ss + q] = val[m];
}

#include "setup_kernels.h"
#include "locate_kernels.h"
#include "grid_kernels.h"
#include "media_kernels.h"

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static thread_local char g_err[512];

static int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
  return code;
}

// error reporting for the other translation units of the library
int cgfdh_internal_fail(int code, const char* msg) { return fail(code, "%s", msg); }

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(CGFDH_EDEVICE, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                             \
  } while (0)

struct FaceHost {
  int on = 0;
  int region[6] = {0};
  int fnx = 0, fny = 0, fnz = 0;
  long fn = 0;
  float* A = nullptr;
  float* B = nullptr;
  float* D = nullptr;
  float* aux[3] = {nullptr, nullptr, nullptr};   // pre, tmp, end
  bool coefs_set = false;
};

struct cgfdh_solver {
  cgfdh_desc d;
  int nx, ny, nz;
  long ss, vol;
  int ni1, ni2, nj1, nj2, nk1, nk2;
  hipStream_t stream = nullptr;
  hipStream_t stream2 = nullptr; // shell boxes, when they run beside the interior
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  float* met = nullptr;
  int ncmp = 3;                  // medium volumes (medium.h)
  int nw = 9;                    // wavefield components (9 + 6 per Maxwell body)
  float wl[MAXWELL_MAX] = {0};   // visco-elastic relaxation frequencies
  float* med = nullptr;
  float* lvl[4] = {nullptr, nullptr, nullptr, nullptr};   // pre, tmpA, tmpB, end
  float* mats = nullptr;         // matVx2Vz2, matVy2Vz2, matF2Vz2, matD
  FaceHost face[6];
  int nsrc = 0;
  SrcDev* src = nullptr;
  float* mij = nullptr;          // moment tables (NULL: none)
  float* fxyz = nullptr;         // force tables (NULL: none)
  long mij_stride = 0;
  int src_max_nt = 0, src_spatial = 1, src_half = 0;
  float* src_coef = nullptr;
  int src_strict = 0;            // strict surface force
  int nsurf = 0;                 // surface forces: indices and force-rate tables
  int* surf_idx = nullptr;
  float* frate = nullptr;
  // dd sources at grid points: marks, sorted indices and their input
  // positions on the device, a dd_cap-row ring of value tables, and per row
  // the step whose values it holds (-1: none)
  int ndd = 0, dd_max_nt = 0, dd_cap = 0;
  unsigned* dd_mask = nullptr;
  long* dd_idx = nullptr;
  int* dd_perm = nullptr;
  float* dd_vi = nullptr;
  float* dd_mij = nullptr;
  std::vector<int> dd_row_it;
  int nrecv = 0;
  int* recv_idx = nullptr;
  float* recv_w = nullptr;
  float* seismo = nullptr;
  float* pg = nullptr;           // 15 peak maps + 3 displacement maps, [ny][nx] each
  float* outbuf = nullptr;       // staging for packed output sub-volumes (grown on demand)
  float* scratch = nullptr;      // device address of the static sub-volume buffer
  unsigned long long* nan_key = nullptr;   // result word of the NaN scan
  // asynchronous output: a copy stream and device staging slots, each
  // packed on the solver's stream and copied out on the copy stream
  hipStream_t copy_stream = nullptr;
  struct OutSlot {
    float* d = nullptr;
    size_t n = 0;
    hipEvent_t packed = nullptr, done = nullptr;
    unsigned long gen = 0;
  };
  std::vector<OutSlot> oslots;
  std::vector<std::pair<int, unsigned long>> tickets;   // (slot, generation)
  unsigned long out_gen = 0;
  float* qs = nullptr;           // Graves-Qs quality factor per point (NULL: off)
  float qs_coef = 0.0f;          // -pi * freq * dt
  float* abl = nullptr;          // ablexp masks Ex[nx], Ey[ny], Ez[nz]
  int abl_blk[6][6] = {};        // enabled ablexp blocks: i1, i2, j1, j2, k1, k2
  int abl_n = 0;
  size_t outbuf_n = 0;
  int variant = 12;              // set per medium in cgfdh_create
  int tgx = 0, tgy = 0;          // tile group (0: the default)
  bool metric_set = false, medium_set = false, mats_set = false;
  // configuration generation: every setter bumps it, so a batch graph
  // captured under an older configuration is recaptured before its replay
  unsigned long gen = 0;
  hipEvent_t ev_run[2];
  std::vector<hipEvent_t> ev_stage;
  double last_run_ms = 0, last_stage_ms = 0;
  int last_launches = 0;
};

// operator-pair flags, fd_t.c:185-195; stage s uses dir (flag + s) % 2
static const int kPairFlags[8][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}, {0, 0, 1},
                                     {0, 1, 0}, {1, 0, 0}, {1, 0, 1}, {0, 1, 1}};
static const double kRkA[3] = {0.5, 0.5, 1.0};
static const double kRkB[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

extern "C" const char* cgfdh_last_error(void) { return g_err; }

static void face_region(const cgfdh_solver* s, int f, int L, int r[6])
{
  r[0] = s->ni1; r[1] = s->ni2; r[2] = s->nj1; r[3] = s->nj2; r[4] = s->nk1; r[5] = s->nk2;
  if (f == 0) r[1] = s->ni1 + L;
  if (f == 1) r[0] = s->ni2 - L;
  if (f == 2) r[3] = s->nj1 + L;
  if (f == 3) r[2] = s->nj2 - L;
  if (f == 4) r[5] = s->nk1 + L;
  if (f == 5) r[4] = s->nk2 - L;
}

static void free_all(cgfdh_solver* s)
{
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(s->met); F(s->med); F(s->mats);
  for (auto& l : s->lvl) F(l);
  for (auto& f : s->face) {
    F(f.A); F(f.B); F(f.D);
    for (auto& a : f.aux) F(a);
  }
  F(s->src); F(s->mij); F(s->fxyz); F(s->src_coef); F(s->surf_idx); F(s->frate);
  F(s->recv_idx); F(s->recv_w); F(s->seismo); F(s->pg); F(s->outbuf); F(s->abl);
  F(s->qs);
  F(s->dd_mask); F(s->dd_idx); F(s->dd_perm); F(s->dd_vi); F(s->dd_mij);
  F(s->nan_key);
}

static int scratch_resolve(float** out);

extern "C" int cgfdh_create(const cgfdh_desc* d, cgfdh_solver** out)
{
  if (!d || !out) return fail(CGFDH_EINVAL, "cgfdh_create: null argument");
  if (d->ni < 8 || d->nj < 8 || d->nk < 8)
    return fail(CGFDH_EINVAL, "cgfdh_create: grid %dx%dx%d below the 8-point minimum", d->ni,
                d->nj, d->nk);
  if (d->medium_type != CGFDH_MEDIUM_ELASTIC_ISO && d->medium_type != CGFDH_MEDIUM_ELASTIC_VTI &&
      d->medium_type != CGFDH_MEDIUM_ELASTIC_ANISO && d->medium_type != CGFDH_MEDIUM_VISCO_ISO)
    return fail(CGFDH_EINVAL, "cgfdh_create: medium_type %d not supported", d->medium_type);
  if (d->medium_type == CGFDH_MEDIUM_VISCO_ISO && (d->nmaxwell < 1 || d->nmaxwell > MAXWELL_MAX))
    return fail(CGFDH_EINVAL, "cgfdh_create: nmaxwell %d outside 1..%d", d->nmaxwell,
                MAXWELL_MAX);
  if (d->nt_total < 1) return fail(CGFDH_EINVAL, "cgfdh_create: nt_total %d", d->nt_total);
  for (int f = 0; f < 6; f++) {
    int n = f < 2 ? d->ni : (f < 4 ? d->nj : d->nk);
    if (d->pml_on[f] && (d->pml_layers[f] < 1 || d->pml_layers[f] >= n))
      return fail(CGFDH_EINVAL, "cgfdh_create: face %d has %d PML layers for %d points", f,
                  d->pml_layers[f], n);
  }
  if (d->free_top && d->pml_on[5])
    return fail(CGFDH_EINVAL, "cgfdh_create: z2 cannot be both free surface and PML");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    return fail(CGFDH_EDEVICE, "cgfdh_create: no HIP device visible");
  cgfdh_solver* s = new cgfdh_solver();
  s->d = *d;
  s->nx = d->ni + 6; s->ny = d->nj + 6; s->nz = d->nk + 6;
  s->ss = (long)s->nx * s->ny;
  s->vol = s->ss * s->nz;
  s->ni1 = 3; s->ni2 = 3 + d->ni - 1;
  s->nj1 = 3; s->nj2 = 3 + d->nj - 1;
  s->nk1 = 3; s->nk2 = 3 + d->nk - 1;
  auto alloc = [&](float** p, size_t n) -> int {
    if (hipMalloc((void**)p, n * sizeof(float)) != hipSuccess) return -1;
    if (hipMemset(*p, 0, n * sizeof(float)) != hipSuccess) return -1;
    return 0;
  };
  int bad = 0;
  bad |= alloc(&s->met, 10 * s->vol);
  // coefficient volumes + slowness (+ Ylam_n, Ymu_n for the visco-elastic medium)
  s->ncmp = d->medium_type == CGFDH_MEDIUM_ELASTIC_VTI     ? Medium<MEDIUM_VTI>::ncoef + 1
            : d->medium_type == CGFDH_MEDIUM_ELASTIC_ANISO ? Medium<MEDIUM_ANISO>::ncoef + 1
            : d->medium_type == CGFDH_MEDIUM_VISCO_ISO     ? 3 + 2 * d->nmaxwell
                                                          : Medium<MEDIUM_ISO>::ncoef + 1;
  s->nw = d->medium_type == CGFDH_MEDIUM_VISCO_ISO ? 9 + 6 * d->nmaxwell : 9;
  bad |= alloc(&s->med, (size_t)s->ncmp * s->vol);
  for (auto& l : s->lvl) bad |= alloc(&l, (size_t)s->nw * s->vol);
  bad |= alloc(&s->mats, 4 * 9 * s->ss);
  if (d->free_top) bad |= alloc(&s->pg, 18 * s->ss);
  for (int f = 0; f < 6 && !bad; f++) {
    FaceHost& F = s->face[f];
    F.on = d->pml_on[f];
    if (!F.on) continue;
    face_region(s, f, d->pml_layers[f], F.region);
    F.fnx = F.region[1] - F.region[0] + 1;
    F.fny = F.region[3] - F.region[2] + 1;
    F.fnz = F.region[5] - F.region[4] + 1;
    F.fn = (long)F.fnx * F.fny * F.fnz;
    int np = d->pml_layers[f] + 1;
    bad |= alloc(&F.A, np) | alloc(&F.B, np) | alloc(&F.D, np);
    for (auto& a : F.aux) bad |= alloc(&a, 9 * F.fn);
  }
  // default stage kernels: the whole-row two-phase kernel with the PML faces
  // folded in for the elastic media, the per-point kernel for the
  // visco-elastic one
  s->variant = d->medium_type == CGFDH_MEDIUM_ELASTIC_ANISO ? 18 : 16;
  if (!bad && hipStreamCreate(&s->stream) != hipSuccess) bad = 1;
  if (!bad && (hipStreamCreate(&s->stream2) != hipSuccess ||
               hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess ||
               hipEventCreateWithFlags(&s->ev_join, hipEventDisableTiming) != hipSuccess))
    bad = 1;
  if (!bad && (hipEventCreate(&s->ev_run[0]) != hipSuccess ||
               hipEventCreate(&s->ev_run[1]) != hipSuccess))
    bad = 1;
  if (bad) {
    free_all(s);
    delete s;
    return fail(CGFDH_ENOMEM, "cgfdh_create: device allocation failed for %dx%dx%d", d->ni,
                d->nj, d->nk);
  }
  // the static sub-volume buffer's address on this device (its entry point
  // reports a failure here; the solver works without it)
  if (scratch_resolve(&s->scratch) != CGFDH_OK) s->scratch = nullptr;
  *out = s;
  return CGFDH_OK;
}

extern "C" int cgfdh_destroy(cgfdh_solver* s)
{
  if (!s) return CGFDH_OK;
  (void)hipStreamSynchronize(s->stream);
  if (s->copy_stream) {
    (void)hipStreamSynchronize(s->copy_stream);
    for (auto& o : s->oslots) {
      if (o.d) (void)hipFree(o.d);
      if (o.packed) (void)hipEventDestroy(o.packed);
      if (o.done) (void)hipEventDestroy(o.done);
    }
    (void)hipStreamDestroy(s->copy_stream);
  }
  free_all(s);
  for (auto e : s->ev_stage) (void)hipEventDestroy(e);
  (void)hipEventDestroy(s->ev_run[0]);
  (void)hipEventDestroy(s->ev_run[1]);
  (void)hipStreamSynchronize(s->stream2);
  if (s->ev_fork) (void)hipEventDestroy(s->ev_fork);
  if (s->ev_join) (void)hipEventDestroy(s->ev_join);
  (void)hipStreamDestroy(s->stream2);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return CGFDH_OK;
}

static int h2d(cgfdh_solver* s, float* dst, const float* src, size_t n)
{
  HIPCHK(hipMemcpyAsync(dst, src, n * sizeof(float), hipMemcpyHostToDevice, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

extern "C" int cgfdh_set_metric(cgfdh_solver* s, const float* m)
{
  if (s) s->gen++;
  if (!s || !m) return fail(CGFDH_EINVAL, "cgfdh_set_metric: null argument");
  int rc = h2d(s, s->met, m, 10 * s->vol);
  if (rc == 0) s->metric_set = true;
  return rc;
}

extern "C" int cgfdh_set_medium_iso(cgfdh_solver* s, const float* lam, const float* mu,
                                    const float* slw)
{
  if (s) s->gen++;
  if (!s || !lam || !mu || !slw) return fail(CGFDH_EINVAL, "cgfdh_set_medium_iso: null argument");
  if (s->d.medium_type != CGFDH_MEDIUM_ELASTIC_ISO)
    return fail(CGFDH_ESTATE, "cgfdh_set_medium_iso: solver medium_type is %d", s->d.medium_type);
  int rc = h2d(s, s->med, lam, s->vol);
  if (!rc) rc = h2d(s, s->med + s->vol, mu, s->vol);
  if (!rc) rc = h2d(s, s->med + 2 * s->vol, slw, s->vol);
  if (!rc) s->medium_set = true;
  return rc;
}

extern "C" int cgfdh_set_medium_vti(cgfdh_solver* s, const float* c11, const float* c13,
                                    const float* c33, const float* c55, const float* c66,
                                    const float* slw)
{
  if (s) s->gen++;
  if (!s || !c11 || !c13 || !c33 || !c55 || !c66 || !slw)
    return fail(CGFDH_EINVAL, "cgfdh_set_medium_vti: null argument");
  if (s->d.medium_type != CGFDH_MEDIUM_ELASTIC_VTI)
    return fail(CGFDH_ESTATE, "cgfdh_set_medium_vti: solver medium_type is %d", s->d.medium_type);
  const float* v[6] = {c11, c13, c33, c55, c66, slw};
  int rc = 0;
  for (int q = 0; q < 6 && !rc; q++) rc = h2d(s, s->med + q * s->vol, v[q], s->vol);
  if (!rc) s->medium_set = true;
  return rc;
}

extern "C" int cgfdh_set_medium_aniso(cgfdh_solver* s, const float* cij, const float* slw)
{
  if (s) s->gen++;
  if (!s || !cij || !slw) return fail(CGFDH_EINVAL, "cgfdh_set_medium_aniso: null argument");
  if (s->d.medium_type != CGFDH_MEDIUM_ELASTIC_ANISO)
    return fail(CGFDH_ESTATE, "cgfdh_set_medium_aniso: solver medium_type is %d",
                s->d.medium_type);
  int rc = h2d(s, s->med, cij, 21 * s->vol);
  if (!rc) rc = h2d(s, s->med + 21 * s->vol, slw, s->vol);
  if (!rc) s->medium_set = true;
  return rc;
}

extern "C" int cgfdh_set_medium_visco(cgfdh_solver* s, const float* lambda, const float* mu,
                                      const float* slw, const float* ylam_ymu, const float* wl)
{
  if (s) s->gen++;
  if (!s || !lambda || !mu || !slw || !ylam_ymu || !wl)
    return fail(CGFDH_EINVAL, "cgfdh_set_medium_visco: null argument");
  if (s->d.medium_type != CGFDH_MEDIUM_VISCO_ISO)
    return fail(CGFDH_ESTATE, "cgfdh_set_medium_visco: solver medium_type is %d",
                s->d.medium_type);
  int rc = h2d(s, s->med, lambda, s->vol);
  if (!rc) rc = h2d(s, s->med + s->vol, mu, s->vol);
  if (!rc) rc = h2d(s, s->med + 2 * s->vol, slw, s->vol);
  if (!rc) rc = h2d(s, s->med + 3 * s->vol, ylam_ymu, 2 * (size_t)s->d.nmaxwell * s->vol);
  for (int n = 0; n < s->d.nmaxwell; n++) s->wl[n] = wl[n];
  if (!rc) s->medium_set = true;
  return rc;
}

extern "C" int cgfdh_set_free_surface_matD(cgfdh_solver* s, const float* matD)
{
  if (s) s->gen++;
  if (!s || !matD) return fail(CGFDH_EINVAL, "cgfdh_set_free_surface_matD: null argument");
  return h2d(s, s->mats + 27 * s->ss, matD, 9 * (size_t)s->ss);
}

extern "C" int cgfdh_set_graves_qs(cgfdh_solver* s, const float* Qs, float freq)
{
  if (s) s->gen++;
  if (!s || !Qs) return fail(CGFDH_EINVAL, "cgfdh_set_graves_qs: null argument");
  // any elastic medium (md->visco_type is Graves-Qs, drv_rk_curv_col.c:414-416);
  // the GMB visco-elastic medium has its own attenuation
  if (s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO)
    return fail(CGFDH_ESTATE, "cgfdh_set_graves_qs: needs an elastic medium "
                "(medium_type %d)", s->d.medium_type);
  if (!(freq > 0.0f)) return fail(CGFDH_EINVAL, "cgfdh_set_graves_qs: freq %g", (double)freq);
  if (!s->qs) HIPCHK(hipMalloc(&s->qs, (size_t)s->vol * sizeof(float)));
  int rc = h2d(s, s->qs, Qs, s->vol);
  if (rc) return rc;
  s->qs_coef = (float)(-3.14159265358979323846 * (double)freq * (double)s->d.dt);
  return CGFDH_OK;
}

extern "C" int cgfdh_set_ablexp(cgfdh_solver* s, const float* Ex, const float* Ey,
                                const float* Ez, const int* blocks)
{
  if (s) s->gen++;
  if (!s || !Ex || !Ey || !Ez || !blocks)
    return fail(CGFDH_EINVAL, "cgfdh_set_ablexp: null argument");
  int blk[6][6] = {}, n = 0;
  const int hi[3] = {s->nx - 1, s->ny - 1, s->nz - 1};
  for (int q = 0; q < 6; q++) {
    const int* b = blocks + 7 * q;
    if (!b[0]) continue;
    for (int a = 0; a < 3; a++)
      if (b[1 + 2 * a] < 0 || b[2 + 2 * a] < b[1 + 2 * a] || b[2 + 2 * a] > hi[a])
        return fail(CGFDH_EINVAL, "cgfdh_set_ablexp: block %d axis %d range %d..%d outside 0..%d",
                    q, a, b[1 + 2 * a], b[2 + 2 * a], hi[a]);
    for (int c = 0; c < 6; c++) blk[n][c] = b[1 + c];
    n++;
  }
  const size_t len = (size_t)s->nx + s->ny + s->nz;
  if (!s->abl) HIPCHK(hipMalloc(&s->abl, len * sizeof(float)));
  int rc = h2d(s, s->abl, Ex, s->nx);
  if (!rc) rc = h2d(s, s->abl + s->nx, Ey, s->ny);
  if (!rc) rc = h2d(s, s->abl + s->nx + s->ny, Ez, s->nz);
  if (rc) return rc;
  memcpy(s->abl_blk, blk, sizeof blk);
  s->abl_n = n;
  return CGFDH_OK;
}

extern "C" int cgfdh_set_pml_face(cgfdh_solver* s, int f, const float* A, const float* B,
                                  const float* D)
{
  if (s) s->gen++;
  if (!s || f < 0 || f > 5 || !A || !B || !D)
    return fail(CGFDH_EINVAL, "cgfdh_set_pml_face: bad argument (face %d)", f);
  if (!s->face[f].on) return fail(CGFDH_EINVAL, "cgfdh_set_pml_face: face %d has no PML", f);
  int np = s->d.pml_layers[f] + 1;
  int rc = h2d(s, s->face[f].A, A, np);
  if (!rc) rc = h2d(s, s->face[f].B, B, np);
  if (!rc) rc = h2d(s, s->face[f].D, D, np);
  if (!rc) s->face[f].coefs_set = true;
  return rc;
}

extern "C" int cgfdh_set_free_surface(cgfdh_solver* s, const float* mx, const float* my,
                                      const float* mf)
{
  if (s) s->gen++;
  if (!s || !mx || !my || !mf) return fail(CGFDH_EINVAL, "cgfdh_set_free_surface: null argument");
  size_t n = 9 * (size_t)s->ss;
  int rc = h2d(s, s->mats, mx, n);
  if (!rc) rc = h2d(s, s->mats + n, my, n);
  if (!rc) rc = h2d(s, s->mats + 2 * n, mf, n);
  if (!rc) s->mats_set = true;
  return rc;
}

extern "C" int cgfdh_set_sources(cgfdh_solver* s, int nsrc, const int* ijk,
                                 const int* it_begin, const int* it_end, int max_nt,
                                 const float* mij, const float* fxyz, int spatial,
                                 int ext_half, const float* ext_coefs, int strict, int nsurf,
                                 const int* surf_idx, const float* frate)
{
  if (s) s->gen++;
  if (!s || nsrc < 0 || nsurf < 0) return fail(CGFDH_EINVAL, "cgfdh_set_sources: bad argument");
  auto F = [](auto*& p) { if (p) { (void)hipFree(p); p = nullptr; } };
  F(s->src); F(s->mij); F(s->fxyz); F(s->src_coef); F(s->surf_idx); F(s->frate);
  s->nsrc = 0;
  s->nsurf = 0;
  s->src_strict = 0;
  if (nsrc == 0) return CGFDH_OK;
  if (!ijk || !it_begin || !it_end || (!mij && !fxyz) || max_nt < 1)
    return fail(CGFDH_EINVAL, "cgfdh_set_sources: missing tables");
  if (spatial != CGFDH_SRC_POINT && spatial != CGFDH_SRC_GAUSSIAN)
    return fail(CGFDH_EINVAL, "cgfdh_set_sources: spatial type %d", spatial);
  if (spatial == CGFDH_SRC_GAUSSIAN && (!ext_coefs || ext_half < 0 || ext_half > 8))
    return fail(CGFDH_EINVAL, "cgfdh_set_sources: Gaussian needs weights, half %d", ext_half);
  const int half = spatial == CGFDH_SRC_GAUSSIAN ? ext_half : 0;
  if (nsurf > 0 && (!fxyz || !strict || !surf_idx || !frate))
    return fail(CGFDH_EINVAL,
                "cgfdh_set_sources: surface forces need force tables, strict mode, their "
                "indices and rate tables");
  if (nsurf > 0 && !s->d.free_top)
    return fail(CGFDH_EINVAL, "cgfdh_set_sources: surface forces need the free surface");
  std::vector<SrcDev> hs(nsrc);
  int L = 2 * ext_half + 1;
  for (int is = 0; is < nsrc; is++) {
    SrcDev& q = hs[is];
    q.si = ijk[3 * is]; q.sj = ijk[3 * is + 1]; q.sk = ijk[3 * is + 2];
    if (q.si < s->ni1 || q.si > s->ni2 || q.sj < s->nj1 || q.sj > s->nj2 || q.sk < s->nk1 ||
        q.sk > s->nk2)
      return fail(CGFDH_EINVAL, "cgfdh_set_sources: source %d at (%d,%d,%d) outside the grid",
                  is, q.si, q.sj, q.sk);
    q.it_begin = it_begin[is];
    q.it_end = it_end[is];
    if (q.it_end - q.it_begin + 1 > max_nt)
      return fail(CGFDH_EINVAL, "cgfdh_set_sources: source %d spans more than max_nt", is);
    q.kb = (q.sk + ext_half < s->nk2) ? ext_half : s->nk2 - q.sk;
    q.coef_off = is * L * L * L;
  }
  // a surface force reaches the surface (src_glob_ext_reach_free,
  // src_t.c:361-371): its point, or its Gaussian's top row, is on it
  for (int n = 0; n < nsurf; n++) {
    const int is = surf_idx[n];
    if (is < 0 || is >= nsrc)
      return fail(CGFDH_EINVAL, "cgfdh_set_sources: surface force %d names source %d", n, is);
    if (hs[is].sk + half < s->nk2)
      return fail(CGFDH_EINVAL,
                  "cgfdh_set_sources: surface force %d (source %d, k %d) does not reach the "
                  "surface",
                  n, is, hs[is].sk);
    if (spatial == CGFDH_SRC_GAUSSIAN &&
        (hs[is].si - half < 0 || hs[is].si + half >= s->nx || hs[is].sj - half < 0 ||
         hs[is].sj + half >= s->ny))
      return fail(CGFDH_EINVAL, "cgfdh_set_sources: surface force %d overhangs the array", n);
  }
  HIPCHK(hipMalloc((void**)&s->src, nsrc * sizeof(SrcDev)));
  HIPCHK(hipMemcpy(s->src, hs.data(), nsrc * sizeof(SrcDev), hipMemcpyHostToDevice));
  s->mij_stride = (long)nsrc * max_nt * 4;
  if (mij) {
    HIPCHK(hipMalloc((void**)&s->mij, 6 * s->mij_stride * sizeof(float)));
    HIPCHK(hipMemcpy(s->mij, mij, 6 * s->mij_stride * sizeof(float), hipMemcpyHostToDevice));
  }
  if (fxyz) {
    HIPCHK(hipMalloc((void**)&s->fxyz, 3 * s->mij_stride * sizeof(float)));
    HIPCHK(hipMemcpy(s->fxyz, fxyz, 3 * s->mij_stride * sizeof(float), hipMemcpyHostToDevice));
  }
  if (spatial == CGFDH_SRC_GAUSSIAN) {
    size_t n = (size_t)nsrc * L * L * L;
    HIPCHK(hipMalloc((void**)&s->src_coef, n * sizeof(float)));
    HIPCHK(hipMemcpy(s->src_coef, ext_coefs, n * sizeof(float), hipMemcpyHostToDevice));
  }
  if (nsurf > 0) {
    const size_t nr = 3 * (size_t)nsurf * max_nt * 4;
    HIPCHK(hipMalloc((void**)&s->surf_idx, nsurf * sizeof(int)));
    HIPCHK(hipMemcpy(s->surf_idx, surf_idx, nsurf * sizeof(int), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc((void**)&s->frate, nr * sizeof(float)));
    HIPCHK(hipMemcpy(s->frate, frate, nr * sizeof(float), hipMemcpyHostToDevice));
  }
  s->nsrc = nsrc;
  s->nsurf = nsurf;
  s->src_strict = fxyz && strict ? 1 : 0;
  s->src_max_nt = max_nt;
  s->src_spatial = spatial;
  s->src_half = half;
  return CGFDH_OK;
}

extern "C" int cgfdh_set_dd_sources(cgfdh_solver* s, int ndd, const size_t* indx, int max_nt,
                                    int nt_cap, int has_vi, int has_mij)
{
  if (s) s->gen++;
  if (!s || ndd < 0) return fail(CGFDH_EINVAL, "cgfdh_set_dd_sources: bad argument");
  auto F = [](auto*& p) { if (p) { (void)hipFree(p); p = nullptr; } };
  F(s->dd_mask); F(s->dd_idx); F(s->dd_perm); F(s->dd_vi); F(s->dd_mij);
  s->ndd = 0;
  s->dd_row_it.clear();
  if (ndd == 0) return CGFDH_OK;
  const int mt = s->d.medium_type;
  if (mt != CGFDH_MEDIUM_ELASTIC_ISO && mt != CGFDH_MEDIUM_VISCO_ISO)
    return fail(CGFDH_EINVAL,
                "cgfdh_set_dd_sources: dd sources need the isotropic or GMB medium (medium "
                "type %d)",
                mt);
  if (!indx || max_nt < 1 || nt_cap < 1 || (!has_vi && !has_mij))
    return fail(CGFDH_EINVAL, "cgfdh_set_dd_sources: missing points, steps or value kinds");
  // the points, sorted by index with equal indices kept in input order
  std::vector<int> perm(ndd);
  for (int q = 0; q < ndd; q++) {
    const long p = (long)indx[q];
    const long i = p % s->nx, j = p / s->nx % s->ny, k = p / s->ss;
    if (indx[q] >= (size_t)s->vol || i < s->ni1 || i > s->ni2 || j < s->nj1 || j > s->nj2 ||
        k < s->nk1 || k > s->nk2)
      return fail(CGFDH_EINVAL,
                  "cgfdh_set_dd_sources: source %d at index %zu is not a physical grid point", q,
                  indx[q]);
    perm[q] = q;
  }
  std::stable_sort(perm.begin(), perm.end(),
                   [&](int a, int b) { return indx[a] < indx[b]; });
  std::vector<long> sorted(ndd);
  std::vector<unsigned> mask(((size_t)s->vol + 31) / 32, 0u);
  for (int q = 0; q < ndd; q++) {
    sorted[q] = (long)indx[perm[q]];
    mask[sorted[q] >> 5] |= 1u << (sorted[q] & 31);
  }
  const int cap = nt_cap < max_nt ? nt_cap : max_nt;
  const size_t rows = (size_t)cap * 4 * ndd;
  HIPCHK(hipMalloc((void**)&s->dd_mask, mask.size() * sizeof(unsigned)));
  HIPCHK(hipMemcpy(s->dd_mask, mask.data(), mask.size() * sizeof(unsigned),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void**)&s->dd_idx, ndd * sizeof(long)));
  HIPCHK(hipMemcpy(s->dd_idx, sorted.data(), ndd * sizeof(long), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void**)&s->dd_perm, ndd * sizeof(int)));
  HIPCHK(hipMemcpy(s->dd_perm, perm.data(), ndd * sizeof(int), hipMemcpyHostToDevice));
  if (has_vi) HIPCHK(hipMalloc((void**)&s->dd_vi, 3 * rows * sizeof(float)));
  if (has_mij) HIPCHK(hipMalloc((void**)&s->dd_mij, 6 * rows * sizeof(float)));
  s->ndd = ndd;
  s->dd_max_nt = max_nt;
  s->dd_cap = cap;
  s->dd_row_it.assign(cap, -1);
  return CGFDH_OK;
}

extern "C" int cgfdh_put_dd_values(cgfdh_solver* s, int it0, int nt, const float* vi,
                                   const float* mij)
{
  if (!s || it0 < 0 || nt < 0) return fail(CGFDH_EINVAL, "cgfdh_put_dd_values: bad argument");
  if (s->ndd == 0) return fail(CGFDH_EINVAL, "cgfdh_put_dd_values: no dd sources set");
  if (nt > s->dd_cap || it0 + nt > s->dd_max_nt)
    return fail(CGFDH_EINVAL,
                "cgfdh_put_dd_values: steps %d..%d exceed the %d-row table or max_nt %d", it0,
                it0 + nt - 1, s->dd_cap, s->dd_max_nt);
  if ((s->dd_vi && !vi) || (s->dd_mij && !mij))
    return fail(CGFDH_EINVAL, "cgfdh_put_dd_values: missing value tables");
  // rows in flight on the device are read by enqueued stages: wait for them
  HIPCHK(hipStreamSynchronize(s->stream));
  const size_t per = (size_t)4 * s->ndd;
  for (int n = 0; n < nt; n++) {
    const int row = (it0 + n) % s->dd_cap;
    if (s->dd_vi)
      HIPCHK(hipMemcpy(s->dd_vi + 3 * per * row, vi + 3 * per * n, 3 * per * sizeof(float),
                       hipMemcpyHostToDevice));
    if (s->dd_mij)
      HIPCHK(hipMemcpy(s->dd_mij + 6 * per * row, mij + 6 * per * n, 6 * per * sizeof(float),
                       hipMemcpyHostToDevice));
    s->dd_row_it[row] = it0 + n;
  }
  return CGFDH_OK;
}

// the value rows of steps it0 .. it0 + nsteps - 1 that add dd sources are loaded
static int dd_check(const cgfdh_solver* s, int it0, int nsteps, const char* who)
{
  if (s->ndd == 0) return CGFDH_OK;
  for (int it = it0; it < it0 + nsteps && it < s->dd_max_nt; it++)
    if (s->dd_row_it[it % s->dd_cap] != it)
      return fail(CGFDH_EINVAL, "%s: dd source values of step %d are not loaded", who, it);
  return CGFDH_OK;
}

extern "C" int cgfdh_set_moment_sources(cgfdh_solver* s, int nsrc, const int* ijk,
                                        const int* it_begin, const int* it_end, int max_nt,
                                        const float* mij, int spatial, int ext_half,
                                        const float* ext_coefs)
{
  if (s) s->gen++;
  if (nsrc > 0 && !mij) return fail(CGFDH_EINVAL, "cgfdh_set_moment_sources: missing tables");
  return cgfdh_set_sources(s, nsrc, ijk, it_begin, it_end, max_nt, mij, nullptr, spatial,
                           ext_half, ext_coefs, 0, 0, nullptr, nullptr);
}

extern "C" int cgfdh_set_receivers(cgfdh_solver* s, int nrecv, const int* idx, const float* dijk)
{
  if (s) s->gen++;
  if (!s || nrecv < 0) return fail(CGFDH_EINVAL, "cgfdh_set_receivers: bad argument");
  if (s->recv_idx) { (void)hipFree(s->recv_idx); s->recv_idx = nullptr; }
  if (s->recv_w) { (void)hipFree(s->recv_w); s->recv_w = nullptr; }
  if (s->seismo) { (void)hipFree(s->seismo); s->seismo = nullptr; }
  s->nrecv = 0;
  if (nrecv == 0) return CGFDH_OK;
  if (!idx || !dijk) return fail(CGFDH_EINVAL, "cgfdh_set_receivers: missing tables");
  for (int r = 0; r < 8 * nrecv; r++)
    if (idx[r] < 0 || idx[r] >= s->vol)
      return fail(CGFDH_EINVAL, "cgfdh_set_receivers: corner index %d out of the volume", idx[r]);
  HIPCHK(hipMalloc((void**)&s->recv_idx, 8 * nrecv * sizeof(int)));
  HIPCHK(hipMemcpy(s->recv_idx, idx, 8 * nrecv * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc((void**)&s->recv_w, 3 * nrecv * sizeof(float)));
  HIPCHK(hipMemcpy(s->recv_w, dijk, 3 * nrecv * sizeof(float), hipMemcpyHostToDevice));
  size_t n = (size_t)nrecv * s->nw * s->d.nt_total;
  HIPCHK(hipMalloc((void**)&s->seismo, n * sizeof(float)));
  HIPCHK(hipMemset(s->seismo, 0, n * sizeof(float)));
  s->nrecv = nrecv;
  return CGFDH_OK;
}

extern "C" int cgfdh_set_wavefield(cgfdh_solver* s, const float* w)
{
  if (s) s->gen++;
  if (!s || !w) return fail(CGFDH_EINVAL, "cgfdh_set_wavefield: null argument");
  return h2d(s, s->lvl[0], w, (size_t)s->nw * s->vol);
}

extern "C" int cgfdh_get_wavefield(cgfdh_solver* s, float* w)
{
  if (!s || !w) return fail(CGFDH_EINVAL, "cgfdh_get_wavefield: null argument");
  HIPCHK(hipMemcpyAsync(w, s->lvl[0], (size_t)s->nw * s->vol * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

__global__ void set_u64_kernel(unsigned long long* p, unsigned long long v) { *p = v; }

// the first NaN of a level in wav_check_value's scan order (wav_t.c:253-273:
// component after component, array index within each): the smallest flat
// index icmp * vol + iptr that holds a NaN.  Grid-stride, so a thread's
// first hit is its smallest; block minimum, then one atomicMin per block.
__global__ void __launch_bounds__(256) nan_scan_kernel(const float* w, long n,
                                                       unsigned long long* key)
{
  __shared__ unsigned long long red[256];
  unsigned long long k = ~0ull;
  const long stride = (long)gridDim.x * 256;
  for (long q = (long)blockIdx.x * 256 + threadIdx.x; q < n; q += stride)
    if (w[q] != w[q]) {
      k = (unsigned long long)q;
      break;
    }
  red[threadIdx.x] = k;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h && red[threadIdx.x + h] < red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + h];
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != ~0ull) atomicMin(key, red[0]);
}

extern "C" int cgfdh_check_nan(cgfdh_solver* s, int* icmp, long* iptr)
{
  if (!s || !icmp || !iptr) return fail(CGFDH_EINVAL, "cgfdh_check_nan: null argument");
  if (!s->nan_key) HIPCHK(hipMalloc((void**)&s->nan_key, sizeof(unsigned long long)));
  const long n = (long)s->nw * s->vol;
  hipLaunchKernelGGL(set_u64_kernel, dim3(1), dim3(1), 0, s->stream, s->nan_key, ~0ull);
  const long blocks = std::min<long>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(nan_scan_kernel, dim3((unsigned)blocks), dim3(256), 0, s->stream, s->lvl[0],
                     n, s->nan_key);
  HIPCHK(hipGetLastError());
  unsigned long long k = 0;
  HIPCHK(hipMemcpyAsync(&k, s->nan_key, sizeof k, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (k == ~0ull) {
    *icmp = -1;
    *iptr = -1;
  } else {
    *icmp = (int)((long)k / s->vol);
    *iptr = (long)k % s->vol;
  }
  return CGFDH_OK;
}

extern "C" size_t cgfdh_pml_aux_len(cgfdh_solver* s)
{
  if (!s) return 0;
  size_t n = 0;
  for (auto& f : s->face)
    if (f.on) n += 9 * f.fn;
  return n;
}

static int aux_copy(cgfdh_solver* s, int level, float* host, bool to_dev)
{
  size_t off = 0;
  for (auto& f : s->face) {
    if (!f.on) continue;
    size_t n = 9 * f.fn * sizeof(float);
    if (to_dev)
      HIPCHK(hipMemcpyAsync(f.aux[level], host + off, n, hipMemcpyHostToDevice, s->stream));
    else
      HIPCHK(hipMemcpyAsync(host + off, f.aux[level], n, hipMemcpyDeviceToHost, s->stream));
    off += 9 * f.fn;
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

extern "C" int cgfdh_set_pml_aux(cgfdh_solver* s, const float* aux)
{
  if (s) s->gen++;
  if (!s || !aux) return fail(CGFDH_EINVAL, "cgfdh_set_pml_aux: null argument");
  return aux_copy(s, 0, (float*)aux, true);
}

extern "C" int cgfdh_get_pml_aux(cgfdh_solver* s, float* aux)
{
  if (!s || !aux) return fail(CGFDH_EINVAL, "cgfdh_get_pml_aux: null argument");
  return aux_copy(s, 0, aux, false);
}

extern "C" int cgfdh_get_seismograms(cgfdh_solver* s, float* out)
{
  if (!s || !out) return fail(CGFDH_EINVAL, "cgfdh_get_seismograms: null argument");
  if (s->nrecv == 0) return CGFDH_OK;
  HIPCHK(hipMemcpyAsync(out, s->seismo, (size_t)s->nrecv * s->nw * s->d.nt_total * sizeof(float),
                        hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

extern "C" int cgfdh_set_kernel_variant(cgfdh_solver* s, int v)
{
  if (s) s->gen++;
  if (!s || v < 0 || v > 18) return fail(CGFDH_EINVAL, "cgfdh_set_kernel_variant: %d", v);
  if (s->d.medium_type != CGFDH_MEDIUM_ELASTIC_ISO && v >= 1 && v <= 3)
    return fail(CGFDH_EINVAL, "cgfdh_set_kernel_variant: variant %d is built for the isotropic "
                "medium only (VTI / anisotropic: 0, 4..18; visco-elastic: 0, 4, 5, 8..11, 16)", v);
  if (s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO && (v == 6 || v == 7 || (v >= 12 && v != 16)))
    return fail(CGFDH_EINVAL, "cgfdh_set_kernel_variant: variant %d is not built for the "
                "visco-elastic medium (0, 4, 5, 8..11, 16)", v);
  s->variant = v;
  return CGFDH_OK;
}

extern "C" int cgfdh_set_tile_group(cgfdh_solver* s, int gx, int gy)
{
  if (s) s->gen++;
  if (!s || gx < 0 || gy < 0) return fail(CGFDH_EINVAL, "cgfdh_set_tile_group: bad argument");
  s->tgx = gx;
  s->tgy = gy;
  return CGFDH_OK;
}

static int ready(cgfdh_solver* s, const char* who)
{
  if (!s->metric_set || !s->medium_set)
    return fail(CGFDH_ESTATE, "%s: metric and medium must be set first", who);
  if (s->d.free_top && !s->mats_set)
    return fail(CGFDH_ESTATE, "%s: free surface enabled but its matrices are not set", who);
  for (int f = 0; f < 6; f++)
    if (s->face[f].on && !s->face[f].coefs_set)
      return fail(CGFDH_ESTATE, "%s: PML face %d coefficients not set", who, f);
  return CGFDH_OK;
}

// fill the launch parameters of one stage; wcur/wtmp and aux levels chosen
// by the caller
static StageParams make_params(cgfdh_solver* s, int it, int istage, const float* wcur,
                               float* wtmp, int aux_cur_level)
{
  StageParams P;
  memset(&P, 0, sizeof P);
  P.ni1 = s->ni1; P.ni2 = s->ni2; P.nj1 = s->nj1; P.nj2 = s->nj2; P.nk1 = s->nk1; P.nk2 = s->nk2;
  P.nx = s->nx; P.ss = s->ss; P.vol = s->vol;
  P.met = s->met;
  const bool iso_like = s->d.medium_type == CGFDH_MEDIUM_ELASTIC_ISO ||
                        s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO;
  P.med = s->med;
  P.slw = s->med + (iso_like ? 2 : s->ncmp - 1) * s->vol;
  P.nw = s->nw;
  P.nmax = s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO ? s->d.nmaxwell : 0;
  for (int n = 0; n < MAXWELL_MAX; n++) P.wl[n] = s->wl[n];
  P.ylam = s->med + 3 * s->vol;
  P.wcur = wcur; P.wpre = s->lvl[0]; P.wtmp = wtmp; P.wend = s->lvl[3];
  P.ca = istage < 3 ? (float)kRkA[istage] * s->d.dt : 0.0f;
  P.cb = (float)kRkB[istage] * s->d.dt;
  P.free_top = s->d.free_top;
  P.matVx = s->mats; P.matVy = s->mats + 9 * s->ss;
  int n = 0;
  for (int f = 0; f < 6; f++) {
    FaceHost& F = s->face[f];
    if (!F.on) continue;
    PmlDev& D = P.pml[n++];
    D.idim = f / 2;
    D.i1 = F.region[0]; D.i2 = F.region[1]; D.j1 = F.region[2]; D.j2 = F.region[3];
    D.k1 = F.region[4]; D.k2 = F.region[5];
    D.fnx = F.fnx; D.fny = F.fny; D.fn = F.fn;
    D.A = F.A; D.B = F.B; D.D = F.D;
    D.aux_cur = F.aux[aux_cur_level];
    D.aux_pre = F.aux[0]; D.aux_tmp = F.aux[1]; D.aux_end = F.aux[2];
  }
  P.npml = n;
  P.nsrc = s->nsrc; P.src = s->src; P.mij = s->mij; P.fxyz = s->fxyz;
  P.mij_stride = s->mij_stride;
  P.src_max_nt = s->src_max_nt; P.src_spatial = s->src_spatial; P.src_half = s->src_half;
  P.src_coef = s->src_coef;
  P.src_strict = s->src_strict;
  P.nsurf = s->nsurf; P.surf_idx = s->surf_idx; P.frate = s->frate;
  P.frate_stride = (long)s->nsurf * s->src_max_nt * 4;
  P.matF = s->mats + 18 * s->ss;
  P.ndd = s->ndd; P.dd_max_nt = s->dd_max_nt; P.dd_cap = s->dd_cap;
  P.dd_mask = s->dd_mask; P.dd_idx = s->dd_idx; P.dd_perm = s->dd_perm;
  P.dd_vi = s->dd_vi; P.dd_mij = s->dd_mij;
  P.it = it; P.istage = istage;
  return P;
}

// whether the PML slabs of the shell go in one launch: CGFDH_MERGE_SHELL=0/1
// (a tuning knob), default on
static int merge_shell()
{
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("CGFDH_MERGE_SHELL");
    v = (e && e[0] == '0') ? 0 : 1;
  }
  return v;
}

// whether the shell boxes run on a second stream beside the interior kernel
// (eager runs only; a captured batch keeps one stream per solver):
// CGFDH_SHELL_STREAM=0/1 (a tuning knob), default off
static int shell_stream()
{
  static int v = -1;
  if (v < 0) {
    const char* e = getenv("CGFDH_SHELL_STREAM");
    v = (e && e[0] == '1') ? 1 : 0;
  }
  return v;
}

static LaunchCtx make_ctx(const cgfdh_solver* s, hipStream_t stream)
{
  LaunchCtx L;
  L.stream = stream;
  L.variant = s->variant;
  L.ni = s->d.ni; L.nj = s->d.nj; L.nk = s->d.nk;
  L.ni1 = s->ni1; L.ni2 = s->ni2; L.nj1 = s->nj1; L.nj2 = s->nj2; L.nk1 = s->nk1; L.nk2 = s->nk2;
  L.free_top = s->d.free_top;
  L.tgx = s->tgx;
  L.tgy = s->tgy;
  L.merge_shell = merge_shell();
  for (int f = 0; f < 6; f++) {
    L.pml_on[f] = s->d.pml_on[f];
    L.pml_layers[f] = s->d.pml_layers[f];
  }
  L.shell_stream = stream;
  L.slab_only = 0;
  return L;
}

// slab_only: variant 16's free-surface rows alone (the batch's fused launch
// runs the whole rows of all its grids)
static void launch_stage(cgfdh_solver* s, hipStream_t stream, const StageParams& P, int kind,
                         int ipair, int istage, bool slab_only = false)
{
  const int dx = (kPairFlags[ipair][0] + istage) % 2;
  const int dy = (kPairFlags[ipair][1] + istage) % 2;
  const int dz = (kPairFlags[ipair][2] + istage) % 2;
  LaunchCtx L = make_ctx(s, stream);
  L.slab_only = slab_only ? 1 : 0;
  bool two = false;
  if (shell_stream() && s->variant >= 2) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    two = hipStreamIsCapturing(stream, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone;
  }
  if (two) {
    (void)hipEventRecord(s->ev_fork, stream);
    (void)hipStreamWaitEvent(s->stream2, s->ev_fork, 0);
    L.shell_stream = s->stream2;
  }
  if (s->d.medium_type == CGFDH_MEDIUM_ELASTIC_VTI)
    launch_stage_vti(L, P, kind, dx, dy, dz);
  else if (s->d.medium_type == CGFDH_MEDIUM_ELASTIC_ANISO)
    launch_stage_aniso(L, P, kind, dx, dy, dz);
  else if (s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO)
    launch_stage_visco(L, P, kind, dx, dy, dz);
  else
    launch_stage_iso(L, P, kind, dx, dy, dz);
  if (two) {
    (void)hipEventRecord(s->ev_join, s->stream2);
    (void)hipStreamWaitEvent(stream, s->ev_join, 0);
  }
}

// one RK4 step of solver s on stream st (drv_rk_curv_col.c:167-544): four fused
// stages, the visco-elastic surface rotation, peak ground motion, receivers,
// then the pre/end swap.  ev (optional) gets a record before and after each
// stage; it_off (optional) is a device step offset added to it, for graph
// replays of a captured cycle.
// ablexp sponge, bdry_ablexp_apply (bdry_t.c:840-890): every component of
// w_end at the points of one block times min(Ex[i], Ey[j], Ez[k]); the blocks
// are launched one after another in the reference's order, so a point that
// two blocks share is damped twice, as there
__global__ void ablexp_kernel(float* __restrict__ w, long vol, int nw, int nx, long ss,
                              const float* __restrict__ Ex, const float* __restrict__ Ey,
                              const float* __restrict__ Ez, int i1, int ni, int j1, int nj,
                              int k1, long npt)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npt) return;
  const long r = t / ni;
  const int i = i1 + (int)(t - r * ni);
  const int k = k1 + (int)(r / nj);
  const int j = j1 + (int)(r - (long)(k - k1) * nj);
  float mask = Ex[i] < Ey[j] ? Ex[i] : Ey[j];
  if (mask > Ez[k]) mask = Ez[k];
  const long p = (long)i + (long)j * nx + (long)k * ss;
  for (int c = 0; c < nw; c++) w[c * vol + p] *= mask;
}

// Graves-Qs attenuation, sv_curv_col_el_graves_Qs (sv_curv_col_el.c:638-667):
// every component at physical points times exp(coef / Qs), coef = -pi f dt.
// The exponential is evaluated in double and rounded to float (correctly
// rounded but for rare double-rounding cases); the reference's libm expf
// differs from it by one ulp on a few inputs in 10^5.
__global__ void graves_qs_kernel(float* __restrict__ w, long vol, int nw, int nx, long ss,
                                 const float* __restrict__ qs, float coef, int ni, int nj,
                                 long npt)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npt) return;
  const long r = t / ni;
  const int i = 3 + (int)(t - r * ni);
  const int k = 3 + (int)(r / nj);
  const int j = 3 + (int)(r - (long)(k - 3) * nj);
  const long p = (long)i + (long)j * nx + (long)k * ss;
  const float att = (float)exp((double)(coef / qs[p]));
  for (int c = 0; c < nw; c++) w[c * vol + p] *= att;
}

static void enqueue_ablexp(cgfdh_solver* s, hipStream_t st)
{
  const float* Ex = s->abl;
  const float* Ey = s->abl + s->nx;
  const float* Ez = s->abl + s->nx + s->ny;
  for (int n = 0; n < s->abl_n; n++) {
    const int* b = s->abl_blk[n];
    const int ni = b[1] - b[0] + 1, nj = b[3] - b[2] + 1, nk = b[5] - b[4] + 1;
    const long npt = (long)ni * nj * nk;
    hipLaunchKernelGGL(ablexp_kernel, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st,
                       s->lvl[3], s->vol, s->nw, s->nx, s->ss, Ex, Ey, Ez, b[0], ni, b[2], nj,
                       b[4], npt);
  }
}

// what follows the four stages of step it (drv_rk_curv_col.c:414-544): the
// visco-elastic surface rotation, Graves-Qs, the sponge, peak ground
// motion, receivers, then the pre/end swap
static int enqueue_post_step(cgfdh_solver* s, hipStream_t st, int it, const int* it_off)
{
  if (s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO && s->d.free_top)   // drv_rk_curv_col.c:440-445
    hipLaunchKernelGGL(vis_free_kernel, dim3((s->d.ni + 255) / 256, s->d.nj), dim3(256), 0, st,
                       s->lvl[3], s->mats + 27 * s->ss, s->d.ni, s->d.nj, s->nx, s->ss, s->vol,
                       s->nk2);
  if (s->qs) {   // after the last stage's combine, drv_rk_curv_col.c:414-416
    const long npt = (long)s->d.ni * s->d.nj * s->d.nk;
    hipLaunchKernelGGL(graves_qs_kernel, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, st,
                       s->lvl[3], s->vol, s->nw, s->nx, s->ss, s->qs, s->qs_coef, s->d.ni,
                       s->d.nj, npt);
  }
  if (s->abl_n) enqueue_ablexp(s, st);   // drv_rk_curv_col.c:482-486
  if (s->pg)   // drv_rk_curv_col.c:492-495, before the pre/end swap
    hipLaunchKernelGGL(pg_kernel, dim3((s->d.ni + 255) / 256, s->d.nj), dim3(256), 0, st,
                       s->lvl[3], s->lvl[0], s->pg, s->pg + 15 * s->ss, s->d.ni, s->d.nj, s->nx,
                       s->ss, s->vol, s->nk2, s->d.dt);
  if (s->nrecv > 0) {
    const int nt = s->nrecv * s->nw;
    hipLaunchKernelGGL(recv_kernel, dim3((nt + 255) / 256), dim3(256), 0, st, s->lvl[3], s->vol,
                       s->nrecv, s->nw, s->recv_idx, s->recv_w, s->seismo, s->d.nt_total, it,
                       it_off);
  }
  std::swap(s->lvl[0], s->lvl[3]);
  for (auto& f : s->face)
    if (f.on) std::swap(f.aux[0], f.aux[2]);
  return CGFDH_OK;
}

static int enqueue_step(cgfdh_solver* s, hipStream_t st, int it, const int* it_off,
                        hipEvent_t* ev)
{
  const int ipair = it % 8;
  float* tA = s->lvl[1];
  float* tB = s->lvl[2];
  for (int sg = 0; sg < 4; sg++) {
    const float* wcur = sg == 0 ? s->lvl[0] : (sg % 2 == 1 ? tA : tB);
    float* wtmp = sg % 2 == 0 ? tA : tB;
    StageParams P = make_params(s, it, sg, wcur, wtmp, sg == 0 ? 0 : 1);
    P.it_off = it_off;
    const int kind = sg == 0 ? KIND_FIRST : (sg == 3 ? KIND_LAST : KIND_MID);
    if (ev) HIPCHK(hipEventRecord(ev[2 * sg], st));
    launch_stage(s, st, P, kind, ipair, sg);
    if (ev) HIPCHK(hipEventRecord(ev[2 * sg + 1], st));
  }
  return enqueue_post_step(s, st, it, it_off);
}

extern "C" int cgfdh_run(cgfdh_solver* s, int it0, int nsteps)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_run: null solver");
  int rc = ready(s, "cgfdh_run");
  if (rc) return rc;
  if (it0 < 0 || nsteps < 0 || it0 + nsteps > s->d.nt_total)
    return fail(CGFDH_EINVAL, "cgfdh_run: steps %d..%d outside nt_total %d", it0, it0 + nsteps,
                s->d.nt_total);
  rc = dd_check(s, it0, nsteps, "cgfdh_run");
  if (rc) return rc;
  const int nev = 8 * nsteps;
  while ((int)s->ev_stage.size() < nev) {
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    s->ev_stage.push_back(e);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(s->ev_run[0], s->stream));
  for (int n = 0; n < nsteps; n++) {
    rc = enqueue_step(s, s->stream, it0 + n, nullptr, s->ev_stage.data() + 8 * n);
    if (rc) return rc;
  }
  HIPCHK(hipEventRecord(s->ev_run[1], s->stream));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, s->ev_run[0], s->ev_run[1]));
  s->last_run_ms = ms;
  double st_ms = 0;
  for (int e = 0; e < 4 * nsteps; e++) {
    float t = 0;
    HIPCHK(hipEventElapsedTime(&t, s->ev_stage[2 * e], s->ev_stage[2 * e + 1]));
    st_ms += t;
  }
  s->last_stage_ms = st_ms;
  s->last_launches = 4 * nsteps;
  return CGFDH_OK;
}

extern "C" int cgfdh_last_timing(cgfdh_solver* s, double* run_ms, double* stage_ms, int* n)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_last_timing: null solver");
  if (run_ms) *run_ms = s->last_run_ms;
  if (stage_ms) *stage_ms = s->last_stage_ms;
  if (n) *n = s->last_launches;
  return CGFDH_OK;
}

extern "C" int cgfdh_onestage(cgfdh_solver* s, int it, int ipair, int istage, const float* w_cur,
                              const float* aux_cur, float* rhs, float* aux_rhs)
{
  if (!s || !w_cur || !rhs) return fail(CGFDH_EINVAL, "cgfdh_onestage: null argument");
  if (ipair < 0 || ipair > 7 || istage < 0 || istage > 3)
    return fail(CGFDH_EINVAL, "cgfdh_onestage: pair %d stage %d", ipair, istage);
  int rc = ready(s, "cgfdh_onestage");
  if (!rc) rc = dd_check(s, it, 1, "cgfdh_onestage");
  if (rc) return rc;
  // inputs go to tmpA / aux tmp, the rhs to tmpB / aux end
  const size_t lv = (size_t)s->nw * s->vol;
  HIPCHK(hipMemcpyAsync(s->lvl[1], w_cur, lv * sizeof(float), hipMemcpyHostToDevice,
                        s->stream));
  HIPCHK(hipMemsetAsync(s->lvl[2], 0, lv * sizeof(float), s->stream));
  if (aux_cur) {
    rc = aux_copy(s, 1, (float*)aux_cur, true);
    if (rc) return rc;
  }
  StageParams P = make_params(s, it, istage, s->lvl[1], nullptr, 1);
  P.rhs_out = s->lvl[2];
  int n = 0;
  for (auto& f : s->face)
    if (f.on) P.aux_rhs_out[n++] = f.aux[2];
  launch_stage(s, s->stream, P, KIND_RHS, ipair, istage);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(rhs, s->lvl[2], lv * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  if (aux_rhs) return aux_copy(s, 2, aux_rhs, false);
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// setup on the device
// ---------------------------------------------------------------------------

extern "C" int cgfdh_compute_metric(cgfdh_solver* s, const float* xyz)
{
  if (s) s->gen++;
  if (!s || !xyz) return fail(CGFDH_EINVAL, "cgfdh_compute_metric: null argument");
  float* d = nullptr;
  HIPCHK(hipMalloc((void**)&d, 3 * s->vol * sizeof(float)));
  hipError_t e = hipMemcpyAsync(d, xyz, 3 * s->vol * sizeof(float), hipMemcpyHostToDevice,
                                s->stream);
  if (e == hipSuccess) {
    dim3 blk(256);
    hipLaunchKernelGGL(metric_kernel, dim3((s->d.ni + 255) / 256, s->d.nj, s->d.nk), blk, 0,
                       s->stream, d, s->met, s->d.ni, s->d.nj, s->d.nk, s->nx, s->ss, s->vol);
    hipLaunchKernelGGL(metric_mirror_kernel, dim3((s->ny + 255) / 256, s->nz, 6), blk, 0,
                       s->stream, s->met, 0, s->nx, s->ny, s->nz, s->ni1, s->ni2, s->nj1,
                       s->nj2, s->nk1, s->nk2, s->ss, s->vol);
    hipLaunchKernelGGL(metric_mirror_kernel, dim3((s->nx + 255) / 256, s->nz, 6), blk, 0,
                       s->stream, s->met, 1, s->nx, s->ny, s->nz, s->ni1, s->ni2, s->nj1,
                       s->nj2, s->nk1, s->nk2, s->ss, s->vol);
    hipLaunchKernelGGL(metric_mirror_kernel, dim3((s->nx + 255) / 256, s->ny, 6), blk, 0,
                       s->stream, s->met, 2, s->nx, s->ny, s->nz, s->ni1, s->ni2, s->nj1,
                       s->nj2, s->nk1, s->nk2, s->ss, s->vol);
    // the visco-elastic surface rotation needs the coordinates: build it now
    if (s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO && s->d.free_top)
      hipLaunchKernelGGL(vis_matd_kernel, dim3((s->d.ni + 255) / 256, s->d.nj), blk, 0,
                         s->stream, d, s->met, s->mats + 27 * s->ss, s->d.ni, s->d.nj, s->nx,
                         s->nk2, s->ss, s->vol);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  }
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(CGFDH_EDEVICE, "cgfdh_compute_metric: %s", hipGetErrorString(e));
  s->metric_set = true;
  return CGFDH_OK;
}

extern "C" int cgfdh_compute_free_surface(cgfdh_solver* s)
{
  if (s) s->gen++;
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_compute_free_surface: null solver");
  if (!s->metric_set || !s->medium_set)
    return fail(CGFDH_ESTATE, "cgfdh_compute_free_surface: metric and medium first");
  const dim3 grd((s->d.ni + 255) / 256, s->d.nj), blk(256);
  if (s->d.medium_type == CGFDH_MEDIUM_ELASTIC_VTI)
    hipLaunchKernelGGL(dvh2dvz_aniso_kernel<MEDIUM_VTI>, grd, blk, 0, s->stream, s->met, s->med,
                       s->mats, s->d.ni, s->d.nj, s->nx, s->nk2, s->ss, s->vol);
  else if (s->d.medium_type == CGFDH_MEDIUM_ELASTIC_ANISO)
    hipLaunchKernelGGL(dvh2dvz_aniso_kernel<MEDIUM_ANISO>, grd, blk, 0, s->stream, s->met,
                       s->med, s->mats, s->d.ni, s->d.nj, s->nx, s->nk2, s->ss, s->vol);
  else
    hipLaunchKernelGGL(dvh2dvz_iso_kernel, grd, blk, 0, s->stream, s->met, s->med,
                       s->med + s->vol, s->mats, s->d.ni, s->d.nj, s->nx, s->nk2, s->ss, s->vol,
                       s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO ? 0 : 1);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(s->stream));
  s->mats_set = true;
  return CGFDH_OK;
}

extern "C" int cgfdh_get_metric(cgfdh_solver* s, float* m)
{
  if (!s || !m) return fail(CGFDH_EINVAL, "cgfdh_get_metric: null argument");
  HIPCHK(hipMemcpy(m, s->met, 10 * s->vol * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

extern "C" int cgfdh_get_free_surface(cgfdh_solver* s, float* m)
{
  if (!s || !m) return fail(CGFDH_EINVAL, "cgfdh_get_free_surface: null argument");
  HIPCHK(hipMemcpy(m, s->mats, 27 * s->ss * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

extern "C" int cgfdh_get_free_surface_matD(cgfdh_solver* s, float* m)
{
  if (!s || !m) return fail(CGFDH_EINVAL, "cgfdh_get_free_surface_matD: null argument");
  HIPCHK(hipMemcpy(m, s->mats + 27 * s->ss, 9 * s->ss * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

extern "C" int cgfdh_get_peak_ground_motion(cgfdh_solver* s, float* pg, float* dis)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_get_peak_ground_motion: null solver");
  if (!s->pg) return fail(CGFDH_ESTATE, "cgfdh_get_peak_ground_motion: no free surface");
  if (pg) HIPCHK(hipMemcpy(pg, s->pg, 15 * s->ss * sizeof(float), hipMemcpyDeviceToHost));
  if (dis)
    HIPCHK(hipMemcpy(dis, s->pg + 15 * s->ss, 3 * s->ss * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// output sub-volumes: snapshots and slices packed on the device, so only the
// packed points cross PCIe.  box = {i1, ni, di, j1, nj, dj, k1, nk, dk} in
// array indices (ghosts included); the packed order is i fastest, then j,
// then k, as io_snap_pack_buff (io_funcs.c:1525-1556) writes.
// ---------------------------------------------------------------------------

struct SubBox {
  int i1, ni, di, j1, nj, dj, k1, nk, dk;
};
struct CompList {
  int c[NW_MAX];
};

__device__ inline long subbox_point(const SubBox& b, long t, int nx, long ss)
{
  const long r = t / b.ni;
  const int n1 = (int)(t - r * b.ni);
  const int n3 = (int)(r / b.nj);
  const int n2 = (int)(r - (long)n3 * b.nj);
  return (long)(b.i1 + n1 * b.di) + (long)(b.j1 + n2 * b.dj) * nx + (long)(b.k1 + n3 * b.dk) * ss;
}

// out[c][t] = w[comp c][point t of the box]; blockIdx.y is the component
__global__ void subvolume_kernel(const float* __restrict__ w, long vol, int nx, long ss, SubBox b,
                                 long npt, CompList cl, float* __restrict__ out)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npt) return;
  const int c = blockIdx.y;
  out[(long)c * npt + t] = w[(long)cl.c[c] * vol + subbox_point(b, t, nx, ss)];
}

// strains from the stresses and the isotropic lambda, mu, with the rounding
// of io_snap_stress_to_strain_eliso (io_funcs.c:1456-1522): the compliance
// factors in double rounded to float, the normal strains in float, the
// shear strains as 0.5 * E3 * T in double.  Output order Exx, Eyy, Ezz,
// Exz, Eyz, Exy: the order of the strain variables io_snap_nc_put writes
// (io_funcs.c:1226-1252).
__global__ void strain_kernel(const float* __restrict__ w, const float* __restrict__ med, long vol,
                              int nx, long ss, SubBox b, long npt, float* __restrict__ out)
{
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npt) return;
  const long p = subbox_point(b, t, nx, ss);
  const float lam = med[p], mu = med[vol + p];
  const float Txx = w[3 * vol + p], Tyy = w[4 * vol + p], Tzz = w[5 * vol + p];
  const float Tyz = w[6 * vol + p], Txz = w[7 * vol + p], Txy = w[8 * vol + p];
  const double den = 3.0 * (double)lam + 2.0 * (double)mu;
  const float E1 = (float)((double)(lam + mu) / ((double)mu * den));
  const float E2 = (float)((double)(-lam) / (2.0 * (double)mu * den));
  const float E3 = (float)(1.0 / (double)mu);
  const float E0 = E2 * (Txx + Tyy + Tzz);
  const float d = E2 - E1;
  out[t] = E0 - d * Txx;
  out[npt + t] = E0 - d * Tyy;
  out[2 * npt + t] = E0 - d * Tzz;
  out[3 * npt + t] = (float)(0.5 * (double)E3 * (double)Txz);
  out[4 * npt + t] = (float)(0.5 * (double)E3 * (double)Tyz);
  out[5 * npt + t] = (float)(0.5 * (double)E3 * (double)Txy);
}

static int check_box(cgfdh_solver* s, const int* box, const char* who, SubBox* b, long* npt)
{
  if (!box) return fail(CGFDH_EINVAL, "%s: null box", who);
  const int dims[3] = {s->nx, s->ny, s->nz};
  for (int a = 0; a < 3; a++) {
    const int o = box[3 * a], n = box[3 * a + 1], d = box[3 * a + 2];
    if (n < 1 || d < 1 || o < 0 || (long)o + (long)(n - 1) * d >= dims[a])
      return fail(CGFDH_EINVAL, "%s: box axis %d (start %d, count %d, stride %d) outside 0..%d",
                  who, a, o, n, d, dims[a] - 1);
  }
  *b = SubBox{box[0], box[1], box[2], box[3], box[4], box[5], box[6], box[7], box[8]};
  *npt = (long)b->ni * b->nj * b->nk;
  return CGFDH_OK;
}

static int outbuf_reserve(cgfdh_solver* s, size_t n)
{
  if (n <= s->outbuf_n) return CGFDH_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->outbuf) HIPCHK(hipFree(s->outbuf));
  s->outbuf = nullptr;
  s->outbuf_n = 0;
  HIPCHK(hipMalloc(&s->outbuf, n * sizeof(float)));
  s->outbuf_n = n;
  return CGFDH_OK;
}

static dim3 subvolume_grid(long npt, int ny)
{
  return dim3((unsigned)((npt + 255) / 256), (unsigned)ny, 1);
}

extern "C" int cgfdh_get_subvolume(cgfdh_solver* s, const int* box, int ncomp, const int* comps,
                                   float* out)
{
  if (!s || !comps || !out) return fail(CGFDH_EINVAL, "cgfdh_get_subvolume: null argument");
  if (ncomp < 1 || ncomp > s->nw)
    return fail(CGFDH_EINVAL, "cgfdh_get_subvolume: ncomp %d outside 1..%d", ncomp, s->nw);
  CompList cl{};
  for (int c = 0; c < ncomp; c++) {
    if (comps[c] < 0 || comps[c] >= s->nw)
      return fail(CGFDH_EINVAL, "cgfdh_get_subvolume: component %d outside 0..%d", comps[c],
                  s->nw - 1);
    cl.c[c] = comps[c];
  }
  SubBox b;
  long npt;
  int rc = check_box(s, box, "cgfdh_get_subvolume", &b, &npt);
  if (!rc) rc = outbuf_reserve(s, (size_t)ncomp * npt);
  if (rc) return rc;
  hipLaunchKernelGGL(subvolume_kernel, subvolume_grid(npt, ncomp), dim3(256), 0, s->stream,
                     s->lvl[0], s->vol, s->nx, s->ss, b, npt, cl, s->outbuf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->outbuf, (size_t)ncomp * npt * sizeof(float),
                        hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// asynchronous steps and output: the steps and the packing run in order on
// the solver's stream, each packed box is copied to the host on a second
// stream, so the copy (and the caller's use of earlier output) overlaps the
// steps enqueued after it
// ---------------------------------------------------------------------------

extern "C" int cgfdh_run_async(cgfdh_solver* s, int it0, int nsteps)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_run_async: null solver");
  int rc = ready(s, "cgfdh_run_async");
  if (rc) return rc;
  if (it0 < 0 || nsteps < 0 || it0 + nsteps > s->d.nt_total)
    return fail(CGFDH_EINVAL, "cgfdh_run_async: steps %d..%d outside nt_total %d", it0,
                it0 + nsteps, s->d.nt_total);
  rc = dd_check(s, it0, nsteps, "cgfdh_run_async");
  if (rc) return rc;
  for (int n = 0; n < nsteps; n++) {
    rc = enqueue_step(s, s->stream, it0 + n, nullptr, nullptr);
    if (rc) return rc;
  }
  HIPCHK(hipGetLastError());
  return CGFDH_OK;
}

extern "C" int cgfdh_sync(cgfdh_solver* s)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_sync: null solver");
  HIPCHK(hipStreamSynchronize(s->stream));
  if (s->copy_stream) HIPCHK(hipStreamSynchronize(s->copy_stream));
  return CGFDH_OK;
}

// a staging slot of at least n floats whose last copy has finished: a free
// one, a new one (at most kOutSlots), or the oldest after its copy
static int out_slot(cgfdh_solver* s, size_t n, int* out)
{
  constexpr int kOutSlots = 16;
  if (!s->copy_stream) HIPCHK(hipStreamCreate(&s->copy_stream));
  int pick = -1;
  for (int i = 0; i < (int)s->oslots.size(); i++) {
    hipError_t q = hipEventQuery(s->oslots[i].done);
    if (q == hipErrorNotReady) continue;
    if (q != hipSuccess) HIPCHK(q);
    if (pick < 0 || (s->oslots[i].n >= n && s->oslots[pick].n < n)) pick = i;
  }
  if (pick < 0 && (int)s->oslots.size() < kOutSlots) {
    cgfdh_solver::OutSlot o;
    HIPCHK(hipEventCreateWithFlags(&o.packed, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&o.done, hipEventDisableTiming));
    s->oslots.push_back(o);
    pick = (int)s->oslots.size() - 1;
  }
  if (pick < 0) {
    pick = 0;
    for (int i = 1; i < (int)s->oslots.size(); i++)
      if (s->oslots[i].gen < s->oslots[pick].gen) pick = i;
    HIPCHK(hipEventSynchronize(s->oslots[pick].done));
  }
  auto& o = s->oslots[pick];
  if (o.n < n) {
    if (o.d) HIPCHK(hipFree(o.d));
    o.d = nullptr;
    o.n = 0;
    HIPCHK(hipMalloc(&o.d, n * sizeof(float)));
    o.n = n;
  }
  *out = pick;
  return CGFDH_OK;
}

extern "C" int cgfdh_subvolume_async(cgfdh_solver* s, const int* box, int ncomp,
                                     const int* comps, float* out, int* ticket)
{
  const char* who = "cgfdh_subvolume_async";
  if (!s || !comps || !out || !ticket) return fail(CGFDH_EINVAL, "%s: null argument", who);
  if (ncomp < 1 || ncomp > s->nw)
    return fail(CGFDH_EINVAL, "%s: ncomp %d outside 1..%d", who, ncomp, s->nw);
  CompList cl{};
  for (int c = 0; c < ncomp; c++) {
    if (comps[c] < 0 || comps[c] >= s->nw)
      return fail(CGFDH_EINVAL, "%s: component %d outside 0..%d", who, comps[c], s->nw - 1);
    cl.c[c] = comps[c];
  }
  SubBox b;
  long npt;
  int rc = check_box(s, box, who, &b, &npt);
  int k = -1;
  if (!rc) rc = out_slot(s, (size_t)ncomp * npt, &k);
  if (rc) return rc;
  auto& o = s->oslots[k];
  hipLaunchKernelGGL(subvolume_kernel, subvolume_grid(npt, ncomp), dim3(256), 0, s->stream,
                     s->lvl[0], s->vol, s->nx, s->ss, b, npt, cl, o.d);
  HIPCHK(hipGetLastError());
  HIPCHK(hipEventRecord(o.packed, s->stream));
  HIPCHK(hipStreamWaitEvent(s->copy_stream, o.packed, 0));
  HIPCHK(hipMemcpyAsync(out, o.d, (size_t)ncomp * npt * sizeof(float), hipMemcpyDeviceToHost,
                        s->copy_stream));
  HIPCHK(hipEventRecord(o.done, s->copy_stream));
  o.gen = ++s->out_gen;
  s->tickets.push_back({k, o.gen});
  *ticket = (int)s->tickets.size() - 1;
  return CGFDH_OK;
}

extern "C" int cgfdh_wait(cgfdh_solver* s, int ticket)
{
  if (!s || ticket < 0 || ticket >= (int)s->tickets.size())
    return fail(CGFDH_EINVAL, "cgfdh_wait: no ticket %d", ticket);
  const auto t = s->tickets[ticket];
  // a slot is reused only after its copy has finished, so a newer generation
  // means this ticket's copy is done
  if (s->oslots[t.first].gen == t.second) HIPCHK(hipEventSynchronize(s->oslots[t.first].done));
  return CGFDH_OK;
}

extern "C" int cgfdh_host_alloc(size_t bytes, void** p)
{
  if (!p) return fail(CGFDH_EINVAL, "cgfdh_host_alloc: null argument");
  *p = nullptr;
  HIPCHK(hipHostMalloc(p, bytes ? bytes : 1, hipHostMallocDefault));
  return CGFDH_OK;
}

extern "C" int cgfdh_host_free(void* p)
{
  if (p) HIPCHK(hipHostFree(p));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// cgfdh_get_subvolume without any device allocation: the packed points go
// through a fixed file-scope device array instead of the solver's staging
// buffer.  Each GPU holds its own copy of the array from the code object's
// load until the process ends; host code reaches it only through the device
// address hipGetSymbolAddress gives for that device (kept per device at
// solver creation, with the device current), never through the array's name.
// ---------------------------------------------------------------------------
static __device__ __align__(256) float scratch[CGFDH_SUBVOLUME_STATIC_MAX];

constexpr int kScratchDevices = 64;
static float* g_scratch_ptr[kScratchDevices];
static std::mutex g_scratch_mu;

// the scratch array's device address on the current device, resolved once
static int scratch_resolve(float** out)
{
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices)
    return fail(CGFDH_EDEVICE, "scratch: device %d beyond %d", dev, kScratchDevices);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(CGFDH_EDEVICE, "scratch: hipGetSymbolAddress failed: %s",
                  hipGetErrorString(e));
    g_scratch_ptr[dev] = (float*)p;
  }
  *out = g_scratch_ptr[dev];
  return CGFDH_OK;
}

extern "C" int cgfdh_get_subvolume_noalloc(cgfdh_solver* s, const int* box, int ncomp,
                                           const int* comps, float* out)
{
  if (!s || !comps || !out) return fail(CGFDH_EINVAL, "cgfdh_get_subvolume_noalloc: null argument");
  if (ncomp < 1 || ncomp > s->nw)
    return fail(CGFDH_EINVAL, "cgfdh_get_subvolume_noalloc: ncomp %d outside 1..%d", ncomp, s->nw);
  CompList cl{};
  for (int c = 0; c < ncomp; c++) {
    if (comps[c] < 0 || comps[c] >= s->nw)
      return fail(CGFDH_EINVAL, "cgfdh_get_subvolume_noalloc: component %d outside 0..%d",
                  comps[c], s->nw - 1);
    cl.c[c] = comps[c];
  }
  SubBox b;
  long npt;
  int rc = check_box(s, box, "cgfdh_get_subvolume_noalloc", &b, &npt);
  if (rc) return rc;
  if ((size_t)ncomp * (size_t)npt > (size_t)CGFDH_SUBVOLUME_STATIC_MAX)
    return fail(CGFDH_EINVAL,
                "cgfdh_get_subvolume_noalloc: %d components x %ld points exceed the %d-value "
                "static buffer (use cgfdh_get_subvolume)",
                ncomp, npt, CGFDH_SUBVOLUME_STATIC_MAX);
  if (!s->scratch) return fail(CGFDH_ESTATE, "cgfdh_get_subvolume_noalloc: no scratch address");
  hipLaunchKernelGGL(subvolume_kernel, subvolume_grid(npt, ncomp), dim3(256), 0, s->stream,
                     s->lvl[0], s->vol, s->nx, s->ss, b, npt, cl, s->scratch);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->scratch, (size_t)ncomp * npt * sizeof(float),
                        hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

extern "C" int cgfdh_get_strain_subvolume(cgfdh_solver* s, const int* box, float* out)
{
  if (!s || !out) return fail(CGFDH_EINVAL, "cgfdh_get_strain_subvolume: null argument");
  if (s->d.medium_type != CGFDH_MEDIUM_ELASTIC_ISO && s->d.medium_type != CGFDH_MEDIUM_VISCO_ISO)
    return fail(CGFDH_ESTATE, "cgfdh_get_strain_subvolume: needs an isotropic lambda, mu "
                "(medium_type %d)", s->d.medium_type);
  if (!s->medium_set) return fail(CGFDH_ESTATE, "cgfdh_get_strain_subvolume: medium not set");
  SubBox b;
  long npt;
  int rc = check_box(s, box, "cgfdh_get_strain_subvolume", &b, &npt);
  if (!rc) rc = outbuf_reserve(s, 6 * (size_t)npt);
  if (rc) return rc;
  hipLaunchKernelGGL(strain_kernel, subvolume_grid(npt, 1), dim3(256), 0, s->stream, s->lvl[0],
                     s->med, s->vol, s->nx, s->ss, b, npt, s->outbuf);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, s->outbuf, 6 * (size_t)npt * sizeof(float), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// stable time step, blk_dt_esti_curv (blk_t.c:2840-2930): per physical point
// cfl / Vp * L, L the smallest distance from the point to the 8 planes through
// its diagonal neighbour triples (fdlib_math_dist_point2plane,
// fdlib_math.c:52-89), float arithmetic with the reference's double sqrt and
// literals.  The minimum and its point come from one 64-bit key per point,
// (dt bits << 32) | iptr: non-negative floats order like their bits, and the
// smaller iptr wins a tie, which is the point the reference's k-j-i loop
// keeps (it replaces only on a strictly smaller value).
// ---------------------------------------------------------------------------

__device__ inline float dist_point2plane(const float* p0, const float* p1, const float* p2,
                                         const float* p3)
{
  float a[3], b[3], n[3];
  for (int c = 0; c < 3; c++) {
    a[c] = p2[c] - p1[c];
    b[c] = p3[c] - p1[c];
  }
  n[0] = a[1] * b[2] - a[2] * b[1];
  n[1] = a[2] * b[0] - a[0] * b[2];
  n[2] = a[0] * b[1] - a[1] * b[0];
  float d = 0.0f, e = 0.0f, nn = 0.0f;
  for (int c = 0; c < 3; c++) d += n[c] * p1[c];
  for (int c = 0; c < 3; c++) e += n[c] * p0[c];
  for (int c = 0; c < 3; c++) nn += n[c] * n[c];
  const float L = fabsf(e - d);
  return L / (float)sqrt((double)nn);     // sqrtf, correctly rounded
}

struct DtGrid {
  const float *x, *y, *z, *vols;
  long vol, ss;
  int nx, ni, nj, nk, mt, nmaxwell;
  float cfl;
};

__device__ inline void dt_at(const DtGrid& G, long p, float* dt, float* vp_out, float* L_out)
{
  const float* v = G.vols;
  const long V = G.vol;
  float vp;
  if (G.mt == CGFDH_MEDIUM_ELASTIC_VTI) {
    const float vpv = (float)sqrt((double)(v[V + p] / v[2 * V + p]));
    const float vph = (float)sqrt((double)(v[p] / v[2 * V + p]));
    vp = vph > vpv ? vph : vpv;
  } else if (G.mt == CGFDH_MEDIUM_ELASTIC_ANISO) {
    vp = (float)sqrt((double)(v[p] / v[V + p]));
  } else {
    vp = (float)sqrt(((double)v[p] + 2.0 * (double)v[V + p]) / (double)v[2 * V + p]);
    if (G.mt == CGFDH_MEDIUM_VISCO_ISO) {
      const float* y = v + 3 * V;       // Ylam_n, Ymu_n interleaved per body
      float ymax = y[p];
      for (int n = 0; n < G.nmaxwell; n++) {
        const float yl = y[2 * n * V + p], ym = y[(2 * n + 1) * V + p];
        ymax = ymax > yl ? ymax : yl;
        ymax = ymax > ym ? ymax : ym;
      }
      const float vvis = (float)((double)vp * (1.0 + sqrt((double)((float)G.nmaxwell * ymax))));
      vp = vvis > vp ? vvis : vp;
    }
  }
  float L = 1.0e20f;
  const float p0[3] = {G.x[p], G.y[p], G.z[p]};
  for (int kk = -1; kk <= 1; kk += 2)
    for (int jj = -1; jj <= 1; jj += 2)
      for (int ii = -1; ii <= 1; ii += 2) {
        const long q1 = p - ii, q2 = p - (long)jj * G.nx, q3 = p - (long)kk * G.ss;
        const float p1[3] = {G.x[q1], G.y[q1], G.z[q1]};
        const float p2[3] = {G.x[q2], G.y[q2], G.z[q2]};
        const float p3[3] = {G.x[q3], G.y[q3], G.z[q3]};
        const float l = dist_point2plane(p0, p1, p2, p3);
        if (L > l) L = l;
      }
  *dt = G.cfl / vp * L;
  *vp_out = vp;
  *L_out = L;
}

__global__ void __launch_bounds__(256) dt_min_kernel(DtGrid G, unsigned long long* key)
{
  __shared__ unsigned long long red[256];
  const long npt = (long)G.ni * G.nj * G.nk;
  const long t = (long)blockIdx.x * 256 + threadIdx.x;
  unsigned long long k = ~0ull;
  if (t < npt) {
    const long r = t / G.ni;
    const int i = 3 + (int)(t - r * G.ni);
    const int kz = 3 + (int)(r / G.nj);
    const int j = 3 + (int)(r - (long)(kz - 3) * G.nj);
    const long p = (long)i + (long)j * G.nx + (long)kz * G.ss;
    float dt, vp, L;
    dt_at(G, p, &dt, &vp, &L);
    if (dt < 1.0e10f)   // the reference's starting value; NaN never wins
      k = ((unsigned long long)__float_as_uint(dt) << 32) | (unsigned long long)(unsigned)p;
  }
  red[threadIdx.x] = k;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const unsigned long long o = red[threadIdx.x + s];
      if (o < red[threadIdx.x]) red[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0 && red[0] != ~0ull) atomicMin(key, red[0]);
}

// res = {dt, vp, L} and ijk of the winning point; ijk = -1 if none
__global__ void dt_finish_kernel(DtGrid G, const unsigned long long* key, float* res, int* ijk)
{
  const unsigned long long k = *key;
  if (k == ~0ull) {
    res[0] = 1.0e10f; res[1] = 0.0f; res[2] = 0.0f;
    ijk[0] = ijk[1] = ijk[2] = -1;
    return;
  }
  const long p = (long)(k & 0xffffffffull);
  dt_at(G, p, &res[0], &res[1], &res[2]);
  const long kz = p / G.ss;
  const long rem = p - kz * G.ss;
  ijk[2] = (int)kz;
  ijk[1] = (int)(rem / G.nx);
  ijk[0] = (int)(rem - (long)ijk[1] * G.nx);
}

extern "C" int cgfdh_estimate_dt(int ni, int nj, int nk, const float* xyz, int medium_type,
                                 const float* vols, int nmaxwell, float cfl, float* dtmax,
                                 float* vpmax, float* lmin, int* ijk)
{
  if (!xyz || !vols || !dtmax || !vpmax || !lmin || !ijk)
    return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: null argument");
  if (ni < 1 || nj < 1 || nk < 1)
    return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: grid %dx%dx%d", ni, nj, nk);
  if (!(cfl > 0.0f)) return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: cfl %g", (double)cfl);
  int nv;
  switch (medium_type) {
    case CGFDH_MEDIUM_ELASTIC_ISO: nv = 3; break;
    case CGFDH_MEDIUM_ELASTIC_VTI: nv = 3; break;
    case CGFDH_MEDIUM_ELASTIC_ANISO: nv = 2; break;
    case CGFDH_MEDIUM_VISCO_ISO:
      if (nmaxwell < 1 || nmaxwell > MAXWELL_MAX)
        return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: nmaxwell %d outside 1..%d", nmaxwell,
                    MAXWELL_MAX);
      nv = 3 + 2 * nmaxwell;
      break;
    default: return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: medium_type %d", medium_type);
  }
  const int nx = ni + 6, ny = nj + 6, nz = nk + 6;
  const long vol = (long)nx * ny * nz;
  if (vol > 0xffffffffL)
    return fail(CGFDH_EINVAL, "cgfdh_estimate_dt: %ld points exceed the 32-bit point index", vol);
  float *dx = nullptr, *dv = nullptr, *dres = nullptr;
  unsigned long long* dkey = nullptr;
  int* dijk = nullptr;
  int rc = CGFDH_OK;
  const unsigned long long none = ~0ull;
  float res[3];
  DtGrid G;
  const long npt = (long)ni * nj * nk;
#define DTCHK(x)                                                                              \
  do {                                                                                       \
    hipError_t e_ = (x);                                                                     \
    if (e_ != hipSuccess) {                                                                  \
      rc = fail(CGFDH_EDEVICE, "cgfdh_estimate_dt: %s: %s", #x, hipGetErrorString(e_));          \
      goto done;                                                                             \
    }                                                                                        \
  } while (0)
  DTCHK(hipMalloc(&dx, 3 * (size_t)vol * sizeof(float)));
  DTCHK(hipMalloc(&dv, (size_t)nv * vol * sizeof(float)));
  DTCHK(hipMalloc(&dres, 3 * sizeof(float)));
  DTCHK(hipMalloc(&dkey, sizeof(unsigned long long)));
  DTCHK(hipMalloc(&dijk, 3 * sizeof(int)));
  DTCHK(hipMemcpy(dx, xyz, 3 * (size_t)vol * sizeof(float), hipMemcpyHostToDevice));
  DTCHK(hipMemcpy(dv, vols, (size_t)nv * vol * sizeof(float), hipMemcpyHostToDevice));
  DTCHK(hipMemcpy(dkey, &none, sizeof none, hipMemcpyHostToDevice));
  G = DtGrid{dx, dx + vol, dx + 2 * vol, dv, vol, (long)nx * ny, nx, ni, nj, nk, medium_type,
             medium_type == CGFDH_MEDIUM_VISCO_ISO ? nmaxwell : 0, cfl};
  hipLaunchKernelGGL(dt_min_kernel, dim3((unsigned)((npt + 255) / 256)), dim3(256), 0, 0, G,
                     dkey);
  DTCHK(hipGetLastError());
  hipLaunchKernelGGL(dt_finish_kernel, dim3(1), dim3(1), 0, 0, G, dkey, dres, dijk);
  DTCHK(hipGetLastError());
  DTCHK(hipMemcpy(res, dres, sizeof res, hipMemcpyDeviceToHost));
  DTCHK(hipMemcpy(ijk, dijk, 3 * sizeof(int), hipMemcpyDeviceToHost));
  *dtmax = res[0];
  *vpmax = res[1];
  *lmin = res[2];
#undef DTCHK
done:
  (void)hipFree(dx);
  (void)hipFree(dv);
  (void)hipFree(dres);
  (void)hipFree(dkey);
  (void)hipFree(dijk);
  return rc;
}

// ---------------------------------------------------------------------------
// point location (gd_curv_coord_to_local_indx, gd_curv_depth_to_axis)
// ---------------------------------------------------------------------------

namespace {
// a device buffer freed when it goes out of scope
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() { if (p) (void)hipFree(p); }
  template <class T> T* as() { return (T*)p; }
};

// the tile ranges of gd_curv_set_minmax (gd_t.c:1808-1850): the physical
// points of an axis split evenly, the first (n % 4) tiles one point longer
void loc_tile_ranges(int n, int* s, int* e)
{
  const int avg = n / LOC_TILES, left = n % LOC_TILES;
  for (int t = 0; t < LOC_TILES; t++) {
    s[t] = t == 0 ? 3 : e[t - 1] + 1;
    e[t] = s[t] + avg - 1 + (t < left ? 1 : 0);
  }
}

// coordinates on the device and the 64 tile boxes on the host
int loc_setup(const char* who, int ni, int nj, int nk, const float* xyz, int npts,
              const float* pts, DevBuf& dxyz, DevBuf& dpts, LocGrid& G, std::vector<float>& box)
{
  if (!xyz || (npts > 0 && !pts)) return fail(CGFDH_EINVAL, "%s: null argument", who);
  if (ni < 1 || nj < 1 || nk < 1 || npts < 0)
    return fail(CGFDH_EINVAL, "%s: grid %dx%dx%d, %d points", who, ni, nj, nk, npts);
  const long nx = ni + 6, ny = nj + 6, nz = nk + 6, vol = nx * ny * nz;
  HIPCHK(hipMalloc(&dxyz.p, 3 * (size_t)vol * sizeof(float)));
  HIPCHK(hipMemcpy(dxyz.p, xyz, 3 * (size_t)vol * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dpts.p, 3 * (size_t)std::max(npts, 1) * sizeof(float)));
  if (npts > 0)
    HIPCHK(hipMemcpy(dpts.p, pts, 3 * (size_t)npts * sizeof(float), hipMemcpyHostToDevice));
  G.x = dxyz.as<float>(); G.y = G.x + vol; G.z = G.x + 2 * vol;
  G.nx = (int)nx; G.ny = (int)ny; G.ss = nx * ny;
  loc_tile_ranges(ni, G.ts[0], G.te[0]);
  loc_tile_ranges(nj, G.ts[1], G.te[1]);
  loc_tile_ranges(nk, G.ts[2], G.te[2]);
  constexpr int NT = LOC_TILES * LOC_TILES * LOC_TILES;
  DevBuf dbox;
  HIPCHK(hipMalloc(&dbox.p, NT * 6 * sizeof(float)));
  hipLaunchKernelGGL(tile_box_kernel, dim3(NT), dim3(256), 0, 0, G, dbox.as<float>());
  HIPCHK(hipGetLastError());
  box.resize(NT * 6);
  HIPCHK(hipMemcpy(box.data(), dbox.p, NT * 6 * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

// one search kernel over (point, tile) pairs, in launches of at most 65535
template <class K>
int loc_search(K kernel, const LocGrid& G, const float* dpts, const std::vector<int>& pairs,
               long max_cells, unsigned long long* dkey)
{
  const long npair = (long)pairs.size() / 2;
  if (npair == 0) return CGFDH_OK;
  DevBuf dpairs;
  HIPCHK(hipMalloc(&dpairs.p, pairs.size() * sizeof(int)));
  HIPCHK(hipMemcpy(dpairs.p, pairs.data(), pairs.size() * sizeof(int), hipMemcpyHostToDevice));
  const unsigned bx = (unsigned)std::min<long>(std::max<long>((max_cells + 255) / 256, 1), 64);
  for (long p0 = 0; p0 < npair; p0 += 65535) {
    const unsigned ny = (unsigned)std::min<long>(65535, npair - p0);
    hipLaunchKernelGGL(kernel, dim3(bx, ny), dim3(256), 0, 0, G, dpts,
                       dpairs.as<int>() + 2 * p0, dkey);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipDeviceSynchronize());
  return CGFDH_OK;
}
}  // namespace

extern "C" int cgfdh_locate_points(int ni, int nj, int nk, const float* xyz, int npts,
                                   const float* pts, int* is_here, int* ijk, float* inc)
{
  const char* who = "cgfdh_locate_points";
  if (npts > 0 && (!is_here || !ijk || !inc)) return fail(CGFDH_EINVAL, "%s: null argument", who);
  DevBuf dxyz, dpts, dkey, dcell;
  LocGrid G;
  std::vector<float> box;
  int rc = loc_setup(who, ni, nj, nk, xyz, npts, pts, dxyz, dpts, G, box);
  if (rc || npts == 0) return rc;
  // candidate tiles: the tile's box holds the point (gd_t.c:2050-2078)
  std::vector<int> pairs;
  long max_cells = 0;
  for (int t = 0; t < LOC_TILES * LOC_TILES * LOC_TILES; t++) {
    const int ti = t % LOC_TILES, tj = t / LOC_TILES % LOC_TILES, tk = t / (LOC_TILES * LOC_TILES);
    const long n = (long)(G.te[0][ti] - G.ts[0][ti] + 1) * (G.te[1][tj] - G.ts[1][tj] + 1) *
                   (G.te[2][tk] - G.ts[2][tk] + 1);
    max_cells = std::max(max_cells, n);
  }
  for (int q = 0; q < npts; q++) {
    const float* p = pts + 3 * q;
    for (int t = 0; t < LOC_TILES * LOC_TILES * LOC_TILES; t++) {
      const float* b = box.data() + 6 * t;
      if (p[0] < b[0] || p[0] > b[1] || p[1] < b[2] || p[1] > b[3] || p[2] < b[4] || p[2] > b[5])
        continue;
      pairs.push_back(q);
      pairs.push_back(t);
    }
  }
  std::vector<unsigned long long> key(npts, ~0ull);
  HIPCHK(hipMalloc(&dkey.p, npts * sizeof(unsigned long long)));
  HIPCHK(hipMemcpy(dkey.p, key.data(), npts * sizeof(unsigned long long), hipMemcpyHostToDevice));
  rc = loc_search(cell_search_kernel, G, dpts.as<float>(), pairs, max_cells,
                  dkey.as<unsigned long long>());
  if (rc) return rc;
  HIPCHK(hipMemcpy(key.data(), dkey.p, npts * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  // the holding cells, then the nearest sample in each
  std::vector<int> cell(3 * (size_t)npts, -1);
  for (int q = 0; q < npts; q++) {
    if (key[q] == ~0ull) continue;
    const int t = (int)(key[q] >> 40);
    const long c = (long)(key[q] & ((1ull << 40) - 1));
    const int ti = t % LOC_TILES, tj = t / LOC_TILES % LOC_TILES, tk = t / (LOC_TILES * LOC_TILES);
    const int nti = G.te[0][ti] - G.ts[0][ti] + 1, ntj = G.te[1][tj] - G.ts[1][tj] + 1;
    cell[3 * q] = G.ts[0][ti] + (int)(c % nti);
    cell[3 * q + 1] = G.ts[1][tj] + (int)(c / nti % ntj);
    cell[3 * q + 2] = G.ts[2][tk] + (int)(c / nti / ntj);
  }
  HIPCHK(hipMalloc(&dcell.p, cell.size() * sizeof(int)));
  HIPCHK(hipMemcpy(dcell.p, cell.data(), cell.size() * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(sample_shift_kernel, dim3(npts), dim3(256), 0, 0, G, dpts.as<float>(),
                     dcell.as<int>(), dkey.as<unsigned long long>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(key.data(), dkey.p, npts * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  constexpr int NS = LOC_SAMPLES + 1;
  for (int q = 0; q < npts; q++) {
    is_here[q] = cell[3 * q] >= 0;
    for (int d = 0; d < 3; d++) {
      ijk[3 * q + d] = cell[3 * q + d];
      inc[3 * q + d] = 0.0f;
    }
    if (!is_here[q]) continue;
    const int s = (int)(key[q] & 0xffffffffull);
    const int n[3] = {s % NS, s / NS % NS, s / (NS * NS)};
    for (int d = 0; d < 3; d++) {
      float f = (float)n[d] / (float)LOC_SAMPLES;
      // past one half the nearest point is the next one (gd_t.c:2146-2157)
      if (f > 0.5) {
        ijk[3 * q + d] += 1;
        f = (float)((double)f - 1.0);
      }
      inc[3 * q + d] = f;
    }
  }
  return CGFDH_OK;
}

extern "C" int cgfdh_depth_to_axis(int ni, int nj, int nk, const float* xyz, int npts,
                                   const float* xyd, int* found, float* z)
{
  const char* who = "cgfdh_depth_to_axis";
  if (npts > 0 && (!found || !z)) return fail(CGFDH_EINVAL, "%s: null argument", who);
  DevBuf dxyz, dpts, dkey, dz, dfound;
  LocGrid G;
  std::vector<float> box;
  int rc = loc_setup(who, ni, nj, nk, xyz, npts, xyd, dxyz, dpts, G, box);
  if (rc || npts == 0) return rc;
  // candidate tiles of the top tile layer by their x-y box (gd_t.c:2205-2216)
  std::vector<int> pairs;
  long max_cells = 0;
  const int tk = LOC_TILES - 1;
  for (int t = 0; t < LOC_TILES * LOC_TILES; t++)
    max_cells = std::max<long>(max_cells, (long)(G.te[0][t % LOC_TILES] - G.ts[0][t % LOC_TILES] + 1) *
                                              (G.te[1][t / LOC_TILES] - G.ts[1][t / LOC_TILES] + 1));
  for (int q = 0; q < npts; q++) {
    const float* p = xyd + 3 * q;
    for (int t = 0; t < LOC_TILES * LOC_TILES; t++) {
      const float* b = box.data() + 6 * (tk * LOC_TILES * LOC_TILES + t);
      if (p[0] < b[0] || p[0] > b[1] || p[1] < b[2] || p[1] > b[3]) continue;
      pairs.push_back(q);
      pairs.push_back(t);
    }
  }
  std::vector<unsigned long long> key(npts, ~0ull);
  HIPCHK(hipMalloc(&dkey.p, npts * sizeof(unsigned long long)));
  HIPCHK(hipMemcpy(dkey.p, key.data(), npts * sizeof(unsigned long long), hipMemcpyHostToDevice));
  rc = loc_search(surface_search_kernel, G, dpts.as<float>(), pairs, max_cells,
                  dkey.as<unsigned long long>());
  if (rc) return rc;
  HIPCHK(hipMalloc(&dz.p, npts * sizeof(float)));
  HIPCHK(hipMalloc(&dfound.p, npts * sizeof(int)));
  hipLaunchKernelGGL(surface_finish_kernel, dim3((npts + 255) / 256), dim3(256), 0, 0, G,
                     dpts.as<float>(), npts, dkey.as<unsigned long long>(), dz.as<float>(),
                     dfound.as<int>());
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpy(z, dz.p, npts * sizeof(float), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(found, dfound.p, npts * sizeof(int), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// grid generation from a layer file (gd_curv_gen_layer)
// ---------------------------------------------------------------------------

namespace {
// the spline lines of one resampling pass: bases in (input, output) pairs
int spline_pass(const float* src, float* dst, const std::vector<long>& base, long sin, long sout,
                int n, int ni, int f, int* dbad)
{
  const int nline = (int)(base.size() / 2);
  if (nline == 0) return CGFDH_OK;
  DevBuf dbase, dwork;
  HIPCHK(hipMalloc(&dbase.p, base.size() * sizeof(long)));
  HIPCHK(hipMemcpy(dbase.p, base.data(), base.size() * sizeof(long), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dwork.p, (size_t)nline * 5 * n * sizeof(float)));
  SplineLines S{src, dst, dbase.as<long>(), sin, sout, n, ni, f, nline, dwork.as<float>(), dbad};
  hipLaunchKernelGGL(spline_lines_kernel, dim3((nline + 63) / 64), dim3(64), 0, 0, S);
  HIPCHK(hipGetLastError());
  HIPCHK(hipDeviceSynchronize());
  return CGFDH_OK;
}
}  // namespace

extern "C" int cgfdh_gen_grid_layer(int n_if, const int* ncell_in, const int* equal, int nx_l,
                                    int ny_l, const float* layer_xyz, const int* resample,
                                    const int* start, int ntx, int nty, int ntz, int ni, int nj,
                                    int nk, int gni1, int gnj1, int gnk1, float* xyz)
{
  const char* who = "cgfdh_gen_grid_layer";
  if (!ncell_in || !equal || !layer_xyz || !resample || !start || !xyz)
    return fail(CGFDH_EINVAL, "%s: null argument", who);
  const int ng = 3, nlay = n_if - 1;
  if (nlay < 1 || n_if > GRID_MAX_IF)
    return fail(CGFDH_EINVAL, "%s: %d interfaces outside 2..%d", who, n_if, GRID_MAX_IF);
  if (nx_l < 2 || ny_l < 2 || ni < 1 || nj < 1 || nk < 1)
    return fail(CGFDH_EINVAL, "%s: %dx%d samples, block %dx%dx%d", who, nx_l, ny_l, ni, nj, nk);
  const int fx = std::abs(resample[0]), fy = std::abs(resample[1]), fz = std::abs(resample[2]);
  if (fx < 1 || fy < 1 || fz < 1)
    return fail(CGFDH_EINVAL, "%s: resample factors %d %d %d", who, resample[0], resample[1],
                resample[2]);
  // the resampled sizes and the block's first indices (gd_t.c:1066-1112)
  const int nxi = resample[0] < 0 ? (nx_l + 1) / fx : (nx_l - 1) * fx + 1;
  const int nx_first = resample[0] < 0 ? (start[0] + 1) / fx : (start[0] - 1) * fx + 1;
  const int nyi = resample[1] < 0 ? (ny_l + 1) / fy : (ny_l - 1) * fy + 1;
  const int ny_first = resample[1] < 0 ? (start[1] + 1) / fy : (start[1] - 1) * fy + 1;
  GridColumns C{};
  int nzi = 1;
  for (int k = 0; k < nlay; k++) {
    C.ncell[k] = resample[2] < 0 ? ncell_in[k] / fz : ncell_in[k] * fz;
    C.equal[k] = equal[k];
    nzi += C.ncell[k];
    // an unequal middle layer with one cell would read before the
    // reference's spacing buffer
    if (C.ncell[k] < 1 || (equal[k] < 1 && k > 0 && k < nlay - 1 && C.ncell[k] < 2))
      return fail(CGFDH_EINVAL, "%s: layer %d has %d cells", who, k, C.ncell[k]);
  }
  if (nzi < 4) return fail(CGFDH_EINVAL, "%s: %d grid levels, 4 needed", who, nzi);
  const int xg = nx_first + gni1 - ng, yg = ny_first + gnj1 - ng;
  const int zg = nzi - start[2] - ntz + gnk1 - ng;
  // the reference's parameter checks (gd_t.c:1151-1177)
  if (nx_first - ng < 0 || ny_first - ng < 0)
    return fail(CGFDH_EINVAL, "%s: horizontal start index %d %d leaves no room for ghosts", who,
                start[0], start[1]);
  if (nx_first + ntx + ng > nxi || ny_first + nty + ng > nyi)
    return fail(CGFDH_EINVAL, "%s: %dx%d points from %d,%d exceed the %dx%d resampled layers",
                who, ntx, nty, nx_first, ny_first, nxi, nyi);
  if (zg < 0) return fail(CGFDH_EINVAL, "%s: vertical start index %d too deep", who, start[2]);
  if (xg + ni + 2 * ng > nxi || yg + nj + 2 * ng > nyi || zg + nk + 2 * ng > nzi + ng)
    return fail(CGFDH_EINVAL, "%s: the block reaches past the resampled layers", who);

  const long vin = (long)nx_l * ny_l * n_if, vit = (long)nxi * nyi * n_if;
  DevBuf din, dL, dout, dbad;
  HIPCHK(hipMalloc(&din.p, 3 * vin * sizeof(float)));
  HIPCHK(hipMemcpy(din.p, layer_xyz, 3 * vin * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dL.p, 3 * vit * sizeof(float)));
  HIPCHK(hipMalloc(&dbad.p, sizeof(int)));
  const int zero = 0;
  HIPCHK(hipMemcpy(dbad.p, &zero, sizeof(int), hipMemcpyHostToDevice));
  const float* In = din.as<float>();
  float* L = dL.as<float>();
  int* bad = dbad.as<int>();
  int rc = CGFDH_OK;
  std::vector<long> base;
  if (resample[0] < 2 && resample[1] < 2) {
    const long n = 3 * vit;
    hipLaunchKernelGGL(sample_layers_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                       In, L, nx_l, ny_l, n_if, nxi, nyi, fx, fy);
    HIPCHK(hipGetLastError());
  } else if (resample[0] < 2) {          // every fx-th in x, a spline along y
    for (int c = 0; c < 3; c++)
      for (int kk = 0; kk < n_if; kk++)
        for (int ii = 0; ii < nxi; ii++) {
          base.push_back(c * vin + (long)ii * fx + (long)kk * nx_l * ny_l);
          base.push_back(c * vit + ii + (long)kk * nxi * nyi);
        }
    rc = spline_pass(In, L, base, nx_l, nxi, ny_l, nyi, fy, bad);
  } else if (resample[1] < 2) {          // a spline along x, every fy-th in y
    for (int c = 0; c < 3; c++)
      for (int kk = 0; kk < n_if; kk++)
        for (int jj = 0; jj < nyi; jj++) {
          base.push_back(c * vin + (long)jj * fy * nx_l + (long)kk * nx_l * ny_l);
          base.push_back(c * vit + (long)jj * nxi + (long)kk * nxi * nyi);
        }
    rc = spline_pass(In, L, base, 1, 1, nx_l, nxi, fx, bad);
  } else {                               // along x on the input rows, then along y
    for (int c = 0; c < 3; c++)
      for (int kk = 0; kk < n_if; kk++)
        for (int jj = 0; jj < ny_l; jj++) {
          base.push_back(c * vin + (long)jj * nx_l + (long)kk * nx_l * ny_l);
          base.push_back(c * vit + (long)jj * fy * nxi + (long)kk * nxi * nyi);
        }
    rc = spline_pass(In, L, base, 1, 1, nx_l, nxi, fx, bad);
    base.clear();
    for (int c = 0; c < 3 && !rc; c++)
      for (int kk = 0; kk < n_if; kk++)
        for (int ii = 0; ii < nxi; ii++) {
          base.push_back(c * vit + ii + (long)kk * nxi * nyi);
          base.push_back(c * vit + ii + (long)kk * nxi * nyi);
        }
    if (!rc) rc = spline_pass(L, L, base, (long)fy * nxi, nxi, ny_l, nyi, fy, bad);
  }
  if (rc) return rc;
  const int nx = ni + 2 * ng, ny = nj + 2 * ng, nz = nk + 2 * ng;
  const long vol = (long)nx * ny * nz;
  HIPCHK(hipMalloc(&dout.p, 3 * vol * sizeof(float)));
  C.L = L; C.nxi = nxi; C.nyi = nyi; C.nif = n_if;
  C.xg = xg; C.yg = yg; C.zg = zg;
  C.nx = nx; C.ny = ny; C.nz = nz;
  C.x3d = dout.as<float>(); C.y3d = C.x3d + vol; C.z3d = C.x3d + 2 * vol;
  C.bad = bad;
  hipLaunchKernelGGL(grid_columns_kernel, dim3((nx * ny + 63) / 64), dim3(64), 0, 0, C);
  HIPCHK(hipGetLastError());
  int hbad = 0;
  HIPCHK(hipMemcpy(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (hbad)
    return fail(CGFDH_EINVAL, "%s: interface samples or heights are not strictly increasing",
                who);
  HIPCHK(hipMemcpy(xyz, dout.p, 3 * vol * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

// ---------------------------------------------------------------------------
// layered model onto the grid (media_layer2model, local values)
// ---------------------------------------------------------------------------

// the native coefficients' places among a target's outputs, in the
// reference functions' argument order: onecmp var | el_iso lambda mu rho |
// el_vti rho c11 c33 c55 c66 c13 | el_aniso rho c11 c12 .. c66; -1 when the
// target does not take the media type
static int layer_slots(int mt, int target, int gridmodel, int* slot)
{
  static const int iso_in_grid_iso[3] = {1, 2, 0};   // media_grid2model_el_iso: rho lambda mu
  static const int vti_in_vti[6] = {1, 2, 3, 4, 5, 0};
  static const int iso_in_aniso[3] = {3, 16, 0};
  static const int vti_in_aniso[6] = {1, 12, 19, 21, 3, 0};
  const bool vti = mt == MD_PREM || mt == MD_THOMSEN || mt == MD_VCIJ;
  const bool full = mt == MD_TTHOMSEN || mt == MD_TBOND || mt == MD_ACIJ;
  if (target == CGFDH_L2M_ONECMP && mt == MD_ONE) { slot[0] = 0; return 1; }
  if (target == CGFDH_L2M_EL_ISO && mt == MD_ISO) {
    for (int q = 0; q < 3; q++) slot[q] = gridmodel ? iso_in_grid_iso[q] : q;
    return 3;
  }
  if (target == CGFDH_L2M_EL_VTI && vti) {
    for (int q = 0; q < 6; q++) slot[q] = vti_in_vti[q];
    return 6;
  }
  if (target == CGFDH_L2M_EL_ANISO) {
    if (mt == MD_ISO) { for (int q = 0; q < 3; q++) slot[q] = iso_in_aniso[q]; return 3; }
    if (vti) { for (int q = 0; q < 6; q++) slot[q] = vti_in_aniso[q]; return 6; }
    if (full) {
      for (int q = 0; q < 21; q++) slot[q] = 1 + q;
      slot[21] = 0;
      return 22;
    }
  }
  return -1;
}

static int layer_nout(int target)
{
  return target == CGFDH_L2M_ONECMP ? 1 : target == CGFDH_L2M_EL_ISO ? 3
                                        : target == CGFDH_L2M_EL_VTI ? 6 : 22;
}

// the layered model (nl = 0) or the grid model (nl layers of ngz[l]
// surfaces, n_if in all) on points (method loc; grid = 0) or on a
// curvilinear grid of nx*ny*nz points (any method)
static int layer2model_run(const char* who, int mt, int target, int method, int n_if, int mx,
                           int my, const float* geo, const float* fields, int grid, int nx,
                           int ny, int nz, long npts, const float* x, const float* y,
                           const float* z, float* const* out, int nl = 0,
                           const int* ngz = nullptr)
{
  if (!geo || !fields || !x || !y || !z || !out) return fail(CGFDH_EINVAL, "%s: null argument", who);
  if (target < CGFDH_L2M_ONECMP || target > CGFDH_L2M_EL_ANISO)
    return fail(CGFDH_EINVAL, "%s: target %d", who, target);
  if (method < CGFDH_L2M_LOC || method > CGFDH_L2M_ARI || (!grid && method != CGFDH_L2M_LOC))
    return fail(CGFDH_EINVAL, "%s: method %d (loc 0, har 1, ari 2; averaging needs a grid)",
                who, method);
  LayerOut O;
  O.nn = layer_slots(mt, target, nl > 0, O.slot);
  if (O.nn < 0) return fail(CGFDH_EINVAL, "%s: media type %d is not supported by this target", who, mt);
  if (n_if < 1 || n_if > (nl > 0 ? MEDIA_MAX_SURF : MEDIA_MAX_IF) || mx < 2 || my < 2 || npts < 0)
    return fail(CGFDH_EINVAL, "%s: %d %s of %dx%d samples, %ld points", who, n_if,
                nl > 0 ? "surfaces" : "interfaces", mx, my, npts);
  LayerModel M;
  memset(&M, 0, sizeof M);
  if (nl > 0) {
    if (!ngz || nl > MEDIA_MAX_LAYERS) return fail(CGFDH_EINVAL, "%s: %d layers", who, nl);
    int acc = 0;
    for (int l = 0; l < nl; l++) {
      if (ngz[l] < 1) return fail(CGFDH_EINVAL, "%s: layer %d has %d surfaces", who, l, ngz[l]);
      M.first[l] = acc;
      acc += ngz[l];
      if (l < nl - 1 && acc < n_if) M.bound[acc >> 5] |= 1u << (acc & 31);
    }
    if (acc != n_if)
      return fail(CGFDH_EINVAL, "%s: the layers hold %d surfaces, not %d", who, acc, n_if);
    M.grid = 1;
    M.nl = nl;
  }
  if (grid && (nx < 1 || ny < 1 || nz < 1 || (long)nx * ny * nz != npts || npts >= INT_MAX))
    return fail(CGFDH_EINVAL, "%s: grid %dx%dx%d", who, nx, ny, nz);
  const int nout = layer_nout(target);
  for (int q = 0; q < nout; q++)
    if (!out[q]) return fail(CGFDH_EINVAL, "%s: null output %d", who, q);
  if (npts == 0) return CGFDH_OK;
  const int nprop = media_nprop(mt);
  const size_t nf = (nl > 0 ? 1 : 1 + 2 * (size_t)nprop) + nprop, vol = (size_t)n_if * mx * my;
  const size_t n = npts;
  DevBuf df, dx, dout, dbad;
  HIPCHK(hipMalloc(&df.p, nf * vol * sizeof(float)));
  HIPCHK(hipMemcpy(df.p, fields, nf * vol * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dx.p, 3 * n * sizeof(float)));
  float* d = dx.as<float>();
  HIPCHK(hipMemcpy(d, x, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d + n, y, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d + 2 * n, z, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dout.p, (size_t)nout * n * sizeof(float)));
  HIPCHK(hipMalloc(&dbad.p, sizeof(unsigned long long)));
  const unsigned long long none = ~0ull;
  HIPCHK(hipMemcpy(dbad.p, &none, sizeof none, hipMemcpyHostToDevice));
  M.f = df.as<float>();
  M.ni = n_if; M.mx = mx; M.my = my; M.nprop = nprop; M.mt = mt;
  M.minx = geo[0]; M.miny = geo[1]; M.dx = geo[2]; M.dy = geo[3];
  M.bad = dbad.as<unsigned long long>();
  O.out = dout.as<float>();
  O.n = npts;
  const dim3 g((unsigned)((n + 255) / 256));
  auto report = [&](unsigned long long bad, bool sub) {
    const long p = (long)(bad & ((1ull << 48) - 1));
    if ((bad >> 48) == 1)
      return fail(CGFDH_EINVAL, "%s: point %ld (%g, %g, %g)%s is outside the interface mesh", who,
                  p, (double)x[p], (double)y[p], (double)z[p],
                  sub ? ": a sub-point of its half cell" : "");
    return fail(CGFDH_EINVAL, "%s: point %ld%s has vp <= sqrt(2) vs", who, p,
                sub ? " (a sub-point of its half cell)" : "");
  };
  hipLaunchKernelGGL(layer_loc_kernel, g, dim3(256), 0, 0, M, d, d + n, d + 2 * n, O);
  HIPCHK(hipGetLastError());
  unsigned long long bad = 0;
  HIPCHK(hipMemcpy(&bad, dbad.p, sizeof bad, hipMemcpyDeviceToHost));
  if (bad != none) return report(bad, false);
  // the grid model averages nothing with one layer (the reference keeps
  // the local values then)
  if (method != CGFDH_L2M_LOC && nl != 1) {
    DevBuf dh, dmat, dlist, dcnt;
    HIPCHK(hipMalloc(&dh.p, 3 * n * sizeof(float)));
    HIPCHK(hipMalloc(&dmat.p, n * sizeof(int)));
    HIPCHK(hipMalloc(&dlist.p, n * sizeof(int)));
    HIPCHK(hipMalloc(&dcnt.p, sizeof(unsigned)));
    HIPCHK(hipMemset(dcnt.p, 0, sizeof(unsigned)));
    hipLaunchKernelGGL(half_grid_kernel, g, dim3(256), 0, 0, d, dh.as<float>(), nx, ny, nz);
    hipLaunchKernelGGL(layer_mark_kernel, g, dim3(256), 0, 0, M, dh.as<float>(), npts,
                       dmat.as<int>());
    hipLaunchKernelGGL(layer_mixed_kernel, g, dim3(256), 0, 0, dmat.as<int>(), nx, ny, nz,
                       dlist.as<int>(), dcnt.as<unsigned>());
    HIPCHK(hipGetLastError());
    unsigned cnt = 0;
    HIPCHK(hipMemcpy(&cnt, dcnt.p, sizeof cnt, hipMemcpyDeviceToHost));
    const int har = method == CGFDH_L2M_HAR;
    if (cnt > 0) {
#define CGFDH_AVG(NN)                                                                         \
  hipLaunchKernelGGL(layer_average_kernel<NN>, dim3(cnt), dim3(256), 0, 0, M, dh.as<float>(), \
                     dlist.as<int>(), har, O, nx, ny)
      switch (O.nn) {
      case 1: CGFDH_AVG(1); break;
      case 3: CGFDH_AVG(3); break;
      case 6: CGFDH_AVG(6); break;
      default: CGFDH_AVG(22); break;
      }
#undef CGFDH_AVG
      HIPCHK(hipGetLastError());
      HIPCHK(hipMemcpy(&bad, dbad.p, sizeof bad, hipMemcpyDeviceToHost));
      if (bad != none) return report(bad, true);
    }
  }
  if (target == CGFDH_L2M_EL_ANISO && O.nn < 22) {
    hipLaunchKernelGGL(layer_expand_kernel, g, dim3(256), 0, 0, O.nn == 6 ? 1 : 0, O.out, npts);
    HIPCHK(hipGetLastError());
  }
  for (int q = 0; q < nout; q++)
    HIPCHK(hipMemcpy(out[q], O.out + (size_t)q * n, n * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

extern "C" int cgfdh_layer2model(int nprop, int n_if, int mx, int my, const float* geo,
                                 const float* fields, long npts, const float* x, const float* y,
                                 const float* z, float* a, float* b, float* c)
{
  if (nprop != 1 && nprop != 3)
    return fail(CGFDH_EINVAL, "cgfdh_layer2model: nprop %d (1 or 3)", nprop);
  float* out[3] = {a, b, c};
  return layer2model_run("cgfdh_layer2model", nprop == 1 ? MD_ONE : MD_ISO,
                         nprop == 1 ? CGFDH_L2M_ONECMP : CGFDH_L2M_EL_ISO, CGFDH_L2M_LOC, n_if,
                         mx, my, geo, fields, 0, 0, 0, 0, npts, x, y, z, out);
}

extern "C" int cgfdh_layer2model_grid(int media_type, int target, int method, int n_if, int mx,
                                      int my, const float* geo, const float* fields, int nx,
                                      int ny, int nz, const float* x3d, const float* y3d,
                                      const float* z3d, float* const* out)
{
  return layer2model_run("cgfdh_layer2model_grid", media_type, target, method, n_if, mx, my,
                         geo, fields, 1, nx, ny, nz, (long)nx * ny * nz, x3d, y3d, z3d, out);
}

// the reference's file-level entry points: read the .md3lay file, check
// the grid type and method names, run on the device
static int layer2model_file(const char* who, int target, float* const* out, const float* x3d,
                            const float* y3d, const float* z3d, size_t nx, size_t ny, size_t nz,
                            int grid_type, const char* path, const char* method_name)
{
  if (!path || !method_name) return fail(CGFDH_EINVAL, "%s: null argument", who);
  if (grid_type != 3)
    return fail(CGFDH_EINVAL, "%s: grid type %d (only the curvilinear grid, 3, is provided)", who,
                grid_type);
  int method;
  if (strcmp(method_name, "loc") == 0) method = CGFDH_L2M_LOC;
  else if (strcmp(method_name, "har") == 0) method = CGFDH_L2M_HAR;
  else if (strcmp(method_name, "ari") == 0) method = CGFDH_L2M_ARI;
  else return fail(CGFDH_EINVAL, "%s: equivalent medium method %s is not provided (loc, har, ari)",
                   who, method_name);
  if (nx > (size_t)INT_MAX || ny > (size_t)INT_MAX || nz > (size_t)INT_MAX)
    return fail(CGFDH_EINVAL, "%s: grid too large", who);
  int mt, nprop, n_if, mx, my;
  float geo[4];
  if (cgfdh_io_md3lay_info(path, &mt, &nprop, &n_if, &mx, &my, geo))
    return fail(CGFDH_EINVAL, "%s: %s", who, cgfdh_io_last_error());
  int slot[22];
  if (layer_slots(mt, target, 0, slot) < 0)
    return fail(CGFDH_EINVAL, "%s: media type %d of %s is not supported by this function", who,
                mt, path);
  std::vector<float> fields((size_t)(1 + 3 * nprop) * n_if * mx * my);
  if (cgfdh_io_read_md3lay(path, fields.data()))
    return fail(CGFDH_EINVAL, "%s: %s", who, cgfdh_io_last_error());
  return layer2model_run(who, mt, target, method, n_if, mx, my, geo, fields.data(), 1, (int)nx,
                         (int)ny, (int)nz, (long)(nx * ny * nz), x3d, y3d, z3d, out);
}

extern "C" int cgfdh_media_layer2model_onecmp(float* var3d, const float* x3d, const float* y3d,
                                              const float* z3d, size_t nx, size_t ny, size_t nz,
                                              int grid_type, const char* in_var_file,
                                              const char* average_method, int myid)
{
  (void)myid;
  float* out[1] = {var3d};
  return layer2model_file("cgfdh_media_layer2model_onecmp", CGFDH_L2M_ONECMP, out, x3d, y3d, z3d,
                          nx, ny, nz, grid_type, in_var_file, average_method);
}

extern "C" int cgfdh_media_layer2model_el_iso(float* lam3d, float* mu3d, float* rho3d,
                                              const float* x3d, const float* y3d,
                                              const float* z3d, size_t nx, size_t ny, size_t nz,
                                              int grid_type, const char* in_3lay_file,
                                              const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[3] = {lam3d, mu3d, rho3d};
  return layer2model_file("cgfdh_media_layer2model_el_iso", CGFDH_L2M_EL_ISO, out, x3d, y3d, z3d,
                          nx, ny, nz, grid_type, in_3lay_file, equivalent_medium_method);
}

extern "C" int cgfdh_media_layer2model_el_vti(float* rho, float* c11, float* c33, float* c55,
                                              float* c66, float* c13, const float* x3d,
                                              const float* y3d, const float* z3d, size_t nx,
                                              size_t ny, size_t nz, int grid_type,
                                              const char* in_3lay_file,
                                              const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[6] = {rho, c11, c33, c55, c66, c13};
  return layer2model_file("cgfdh_media_layer2model_el_vti", CGFDH_L2M_EL_VTI, out, x3d, y3d, z3d,
                          nx, ny, nz, grid_type, in_3lay_file, equivalent_medium_method);
}

extern "C" int cgfdh_media_layer2model_el_aniso(
    float* rho, float* c11, float* c12, float* c13, float* c14, float* c15, float* c16,
    float* c22, float* c23, float* c24, float* c25, float* c26, float* c33, float* c34,
    float* c35, float* c36, float* c44, float* c45, float* c46, float* c55, float* c56,
    float* c66, const float* x3d, const float* y3d, const float* z3d, size_t nx, size_t ny,
    size_t nz, int grid_type, const char* in_3lay_file, const char* equivalent_medium_method,
    int myid)
{
  (void)myid;
  float* out[22] = {rho, c11, c12, c13, c14, c15, c16, c22, c23, c24, c25,
                    c26, c33, c34, c35, c36, c44, c45, c46, c55, c56, c66};
  return layer2model_file("cgfdh_media_layer2model_el_aniso", CGFDH_L2M_EL_ANISO, out, x3d, y3d,
                          z3d, nx, ny, nz, grid_type, in_3lay_file, equivalent_medium_method);
}

extern "C" int cgfdh_bin2model_grid(int nvol, const int* n3, const float* xv, const float* yv,
                                    const float* zv, const float* vols, long npts, const float* x,
                                    const float* y, const float* z, float* const* out)
{
  const char* who = "cgfdh_bin2model_grid";
  if (!n3 || !xv || !yv || !zv || !vols || !x || !y || !z || !out)
    return fail(CGFDH_EINVAL, "%s: null argument", who);
  if ((nvol != 2 && nvol != 3) || n3[0] < 2 || n3[1] < 2 || n3[2] < 2 || npts < 0)
    return fail(CGFDH_EINVAL, "%s: %d volumes of %dx%dx%d samples, %ld points (each axis needs "
                "two samples)", who, nvol, n3[0], n3[1], n3[2], npts);
  for (int q = 0; q < nvol; q++)
    if (!out[q]) return fail(CGFDH_EINVAL, "%s: null output %d", who, q);
  if (npts == 0) return CGFDH_OK;
  const size_t v3 = (size_t)n3[0] * n3[1] * n3[2], n = npts;
  DevBuf dvol, dvec, dx, dout;
  HIPCHK(hipMalloc(&dvol.p, nvol * v3 * sizeof(float)));
  HIPCHK(hipMemcpy(dvol.p, vols, nvol * v3 * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dvec.p, ((size_t)n3[0] + n3[1] + n3[2]) * sizeof(float)));
  float* vec = dvec.as<float>();
  HIPCHK(hipMemcpy(vec, xv, n3[0] * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(vec + n3[0], yv, n3[1] * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(vec + n3[0] + n3[1], zv, n3[2] * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dx.p, 3 * n * sizeof(float)));
  float* d = dx.as<float>();
  HIPCHK(hipMemcpy(d, x, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d + n, y, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(d + 2 * n, z, n * sizeof(float), hipMemcpyHostToDevice));
  HIPCHK(hipMalloc(&dout.p, nvol * n * sizeof(float)));
  BinModel B;
  B.vol = dvol.as<float>();
  B.xv = vec; B.yv = vec + n3[0]; B.zv = vec + n3[0] + n3[1];
  B.nx = n3[0]; B.ny = n3[1]; B.nz = n3[2]; B.nvol = nvol;
  hipLaunchKernelGGL(bin_loc_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, B, d,
                     d + n, d + 2 * n, npts, dout.as<float>());
  HIPCHK(hipGetLastError());
  for (int q = 0; q < nvol; q++)
    HIPCHK(hipMemcpy(out[q], dout.as<float>() + q * n, n * sizeof(float), hipMemcpyDeviceToHost));
  return CGFDH_OK;
}

// the window and mesh vectors of a binary model over the grid's extent
// (media_bin2model_el_iso :46-140): start / end per stored axis, vectors
// x0 + (ix_start + i) dx (x, y windowed, z whole); false when the model
// misses the grid (the reference stops)
static bool bin_window(const int* order, const int* size, const float* spacing,
                       const float* origin, float xmin, float xmax, float ymin, float ymax,
                       float zmin, float zmax, int* start, int* end, std::vector<float>& xv,
                       std::vector<float>& yv, std::vector<float>& zv, char* why, size_t nwhy)
{
  int dim[3] = {-1, -1, -1};
  for (int a = 0; a < 3; a++) {
    if (order[a] < 0 || order[a] > 2) {
      snprintf(why, nwhy, "bin_order[%d] = %d (0, 1 or 2)", a, order[a]);
      return false;
    }
    dim[order[a]] = a;
  }
  if (dim[0] < 0 || dim[1] < 0 || dim[2] < 0) {
    snprintf(why, nwhy, "bin_order %d,%d,%d misses a direction", order[0], order[1], order[2]);
    return false;
  }
  const float x0 = origin[dim[0]], y0 = origin[dim[1]], z0 = origin[dim[2]];
  const float dx = spacing[dim[0]], dy = spacing[dim[1]], dz = spacing[dim[2]];
  const int ix_start = (int)(xmin < x0 ? 0.0f : (xmin - x0) / dx);
  if (ix_start >= size[dim[0]] || xmax < x0) {
    snprintf(why, nwhy, "the model does not coincide with the grid in x");
    return false;
  }
  int ix_end = (int)ceil((double)((xmax - x0) / dx));
  if (ix_end >= size[dim[0]]) ix_end = size[dim[0]] - 1;
  const int iy_start = (int)(ymin < y0 ? 0.0f : (ymin - y0) / dy);
  if (iy_start >= size[dim[1]] || ymax < y0) {
    snprintf(why, nwhy, "the model does not coincide with the grid in y");
    return false;
  }
  int iy_end = (int)ceil((double)((ymax - y0) / dy));
  if (iy_end >= size[dim[1]]) iy_end = size[dim[1]] - 1;
  const float zn = z0 + dz * (size[dim[2]] - 1);
  if (zmin >= std::max(zn, z0) || zmax <= std::min(z0, zn)) {
    snprintf(why, nwhy, "the model does not coincide with the grid in z");
    return false;
  }
  xv.resize(ix_end - ix_start + 1);
  yv.resize(iy_end - iy_start + 1);
  zv.resize(size[dim[2]]);
  for (size_t i = 0; i < xv.size(); i++) xv[i] = x0 + (ix_start + (int)i) * dx;
  for (size_t i = 0; i < yv.size(); i++) yv[i] = y0 + (iy_start + (int)i) * dy;
  for (size_t i = 0; i < zv.size(); i++) zv[i] = z0 + (int)i * dz;
  start[dim[0]] = ix_start; end[dim[0]] = ix_end;
  start[dim[1]] = iy_start; end[dim[1]] = iy_end;
  start[dim[2]] = 0; end[dim[2]] = size[dim[2]] - 1;
  return true;
}

// media_bin2model_el_iso / _vis_iso: the z extent of the grid on the
// device, the window, the volumes read, then cgfdh_bin2model_grid
static int bin2model_file(const char* who, int nvol, float* const* out, const float* x3d,
                          const float* y3d, const float* z3d, size_t nx, size_t ny, size_t nz,
                          float xmin, float xmax, float ymin, float ymax, int grid_type,
                          const int* order, const int* size, const float* spacing,
                          const float* origin, const char* const* files)
{
  if (!x3d || !y3d || !z3d || !order || !size || !spacing || !origin)
    return fail(CGFDH_EINVAL, "%s: null argument", who);
  for (int q = 0; q < nvol; q++)
    if (!files[q]) return fail(CGFDH_EINVAL, "%s: null file name", who);
  if (grid_type != 3)
    return fail(CGFDH_EINVAL, "%s: grid type %d (only the curvilinear grid, 3, is provided)", who,
                grid_type);
  const size_t n = nx * ny * nz;
  if (n == 0) return CGFDH_OK;
  unsigned mm[2] = {0xffffffffu, 0u};
  {
    DevBuf dz, dmm;
    HIPCHK(hipMalloc(&dz.p, n * sizeof(float)));
    HIPCHK(hipMemcpy(dz.p, z3d, n * sizeof(float), hipMemcpyHostToDevice));
    HIPCHK(hipMalloc(&dmm.p, sizeof mm));
    HIPCHK(hipMemcpy(dmm.p, mm, sizeof mm, hipMemcpyHostToDevice));
    const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(minmax_kernel, dim3(blocks), dim3(256), 0, 0, dz.as<float>(), (long)n,
                       dmm.as<unsigned>());
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpy(mm, dmm.p, sizeof mm, hipMemcpyDeviceToHost));
  }
  int start[3], end[3];
  std::vector<float> xv, yv, zv;
  char why[160];
  if (!bin_window(order, size, spacing, origin, xmin, xmax, ymin, ymax, float_unkey(mm[0]),
                  float_unkey(mm[1]), start, end, xv, yv, zv, why, sizeof why))
    return fail(CGFDH_EINVAL, "%s: %s", who, why);
  const int n3[3] = {(int)xv.size(), (int)yv.size(), (int)zv.size()};
  const size_t v3 = (size_t)n3[0] * n3[1] * n3[2];
  std::vector<float> vols(nvol * v3);
  for (int q = 0; q < nvol; q++)
    if (cgfdh_io_read_bin_window(files[q], order, size, start, end, vols.data() + q * v3))
      return fail(CGFDH_EINVAL, "%s: %s", who, cgfdh_io_last_error());
  return cgfdh_bin2model_grid(nvol, n3, xv.data(), yv.data(), zv.data(), vols.data(), (long)n,
                              x3d, y3d, z3d, out);
}

extern "C" int cgfdh_media_bin2model_el_iso(float* rho3d, float* lam3d, float* mu3d,
                                            const float* x3d, const float* y3d, const float* z3d,
                                            size_t nx, size_t ny, size_t nz, float xmin,
                                            float xmax, float ymin, float ymax, int grid_type,
                                            int* bin_order, int* bin_size, float* bin_spacing,
                                            float* bin_origin, const char* bin_file_rho,
                                            const char* bin_file_vp, const char* bin_file_vs)
{
  float* out[3] = {rho3d, lam3d, mu3d};
  const char* files[3] = {bin_file_rho, bin_file_vp, bin_file_vs};
  return bin2model_file("cgfdh_media_bin2model_el_iso", 3, out, x3d, y3d, z3d, nx, ny, nz, xmin,
                        xmax, ymin, ymax, grid_type, bin_order, bin_size, bin_spacing,
                        bin_origin, files);
}

extern "C" int cgfdh_media_bin2model_vis_iso(float* Qp3d, float* Qs3d, const float* x3d,
                                             const float* y3d, const float* z3d, size_t nx,
                                             size_t ny, size_t nz, float xmin, float xmax,
                                             float ymin, float ymax, int grid_type,
                                             int* bin_order, int* bin_size, float* bin_spacing,
                                             float* bin_origin, const char* bin_file_Qp,
                                             const char* bin_file_Qs)
{
  float* out[2] = {Qp3d, Qs3d};
  const char* files[2] = {bin_file_Qp, bin_file_Qs};
  return bin2model_file("cgfdh_media_bin2model_vis_iso", 2, out, x3d, y3d, z3d, nx, ny, nz, xmin,
                        xmax, ymin, ymax, grid_type, bin_order, bin_size, bin_spacing,
                        bin_origin, files);
}

extern "C" int cgfdh_grid2model_grid(int media_type, int target, int method, int n_layers,
                                     const int* ngz, int mx, int my, const float* geo,
                                     const float* fields, int nx, int ny, int nz,
                                     const float* x3d, const float* y3d, const float* z3d,
                                     float* const* out)
{
  const char* who = "cgfdh_grid2model_grid";
  if (!ngz || n_layers < 1) return fail(CGFDH_EINVAL, "%s: %d layers", who, n_layers);
  long ns = 0;
  for (int l = 0; l < n_layers; l++) ns += ngz[l];
  if (ns < 1 || ns > MEDIA_MAX_SURF) return fail(CGFDH_EINVAL, "%s: %ld surfaces", who, ns);
  return layer2model_run(who, media_type, target, method, (int)ns, mx, my, geo, fields, 1, nx, ny,
                         nz, (long)nx * ny * nz, x3d, y3d, z3d, out, n_layers, ngz);
}

// the reference's grid-model file entry points: read the .md3grd window
// covering [Xmin, Xmax] x [Ymin, Ymax], check the grid type and method
// names, run on the device
static int grid2model_file(const char* who, int target, float* const* out, const float* x3d,
                           const float* y3d, const float* z3d, int nx, int ny, int nz,
                           float xmin, float xmax, float ymin, float ymax, int grid_type,
                           const char* path, const char* method_name)
{
  if (!path || !method_name) return fail(CGFDH_EINVAL, "%s: null argument", who);
  if (grid_type != 3)
    return fail(CGFDH_EINVAL, "%s: grid type %d (only the curvilinear grid, 3, is provided)", who,
                grid_type);
  int method;
  if (strcmp(method_name, "loc") == 0) method = CGFDH_L2M_LOC;
  else if (strcmp(method_name, "har") == 0) method = CGFDH_L2M_HAR;
  else if (strcmp(method_name, "ari") == 0) method = CGFDH_L2M_ARI;
  else return fail(CGFDH_EINVAL, "%s: equivalent medium method %s is not provided (loc, har, ari)",
                   who, method_name);
  int mt, nprop, nl, ns, mx, my;
  std::vector<int> ngz(MEDIA_MAX_LAYERS);
  float geo[4];
  if (cgfdh_io_md3grd_info(path, xmin, xmax, ymin, ymax, &mt, &nprop, &nl, ngz.data(),
                           MEDIA_MAX_LAYERS, &ns, &mx, &my, geo))
    return fail(CGFDH_EINVAL, "%s: %s", who, cgfdh_io_last_error());
  int slot[22];
  if (layer_slots(mt, target, 1, slot) < 0)
    return fail(CGFDH_EINVAL, "%s: media type %d of %s is not supported by this function", who,
                mt, path);
  std::vector<float> fields((size_t)(1 + nprop) * ns * mx * my);
  if (cgfdh_io_read_md3grd(path, xmin, xmax, ymin, ymax, fields.data()))
    return fail(CGFDH_EINVAL, "%s: %s", who, cgfdh_io_last_error());
  if (nx < 1 || ny < 1 || nz < 1) return fail(CGFDH_EINVAL, "%s: grid %dx%dx%d", who, nx, ny, nz);
  return layer2model_run(who, mt, target, method, ns, mx, my, geo, fields.data(), 1, nx, ny, nz,
                         (long)nx * ny * nz, x3d, y3d, z3d, out, nl, ngz.data());
}

extern "C" int cgfdh_media_grid2model_onecmp(float* var3d, const float* x3d, const float* y3d,
                                             const float* z3d, int nx, int ny, int nz, float Xmin,
                                             float Xmax, float Ymin, float Ymax, int grid_type,
                                             const char* in_media_file,
                                             const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[1] = {var3d};
  return grid2model_file("cgfdh_media_grid2model_onecmp", CGFDH_L2M_ONECMP, out, x3d, y3d, z3d,
                         nx, ny, nz, Xmin, Xmax, Ymin, Ymax, grid_type, in_media_file,
                         equivalent_medium_method);
}

extern "C" int cgfdh_media_grid2model_el_iso(float* rho3d, float* lam3d, float* mu3d,
                                             const float* x3d, const float* y3d,
                                             const float* z3d, int nx, int ny, int nz, float Xmin,
                                             float Xmax, float Ymin, float Ymax, int grid_type,
                                             const char* in_media_file,
                                             const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[3] = {rho3d, lam3d, mu3d};
  return grid2model_file("cgfdh_media_grid2model_el_iso", CGFDH_L2M_EL_ISO, out, x3d, y3d, z3d,
                         nx, ny, nz, Xmin, Xmax, Ymin, Ymax, grid_type, in_media_file,
                         equivalent_medium_method);
}

extern "C" int cgfdh_media_grid2model_el_vti(float* rho, float* c11, float* c33, float* c55,
                                             float* c66, float* c13, const float* x3d,
                                             const float* y3d, const float* z3d, int nx, int ny,
                                             int nz, float Xmin, float Xmax, float Ymin,
                                             float Ymax, int grid_type, const char* in_media_file,
                                             const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[6] = {rho, c11, c33, c55, c66, c13};
  return grid2model_file("cgfdh_media_grid2model_el_vti", CGFDH_L2M_EL_VTI, out, x3d, y3d, z3d,
                         nx, ny, nz, Xmin, Xmax, Ymin, Ymax, grid_type, in_media_file,
                         equivalent_medium_method);
}

extern "C" int cgfdh_media_grid2model_el_aniso(
    float* rho, float* c11, float* c12, float* c13, float* c14, float* c15, float* c16,
    float* c22, float* c23, float* c24, float* c25, float* c26, float* c33, float* c34,
    float* c35, float* c36, float* c44, float* c45, float* c46, float* c55, float* c56,
    float* c66, const float* x3d, const float* y3d, const float* z3d, int nx, int ny, int nz,
    float Xmin, float Xmax, float Ymin, float Ymax, int grid_type, const char* in_media_file,
    const char* equivalent_medium_method, int myid)
{
  (void)myid;
  float* out[22] = {rho, c11, c12, c13, c14, c15, c16, c22, c23, c24, c25,
                    c26, c33, c34, c35, c36, c44, c45, c46, c55, c56, c66};
  return grid2model_file("cgfdh_media_grid2model_el_aniso", CGFDH_L2M_EL_ANISO, out, x3d, y3d,
                         z3d, nx, ny, nz, Xmin, Xmax, Ymin, Ymax, grid_type, in_media_file,
                         equivalent_medium_method);
}

// ---------------------------------------------------------------------------
// batched runs: independent solvers advanced together.  A cycle of 8 steps
// (one full operator-pair schedule, an even number of level swaps, so every
// pointer returns to where it started) of all solvers is captured once into
// a hipGraph, each solver on its own stream so the graph's branches run
// concurrently, and replayed; a device step offset shifts the captured step
// numbers of sources and receivers by 8 per replay.
// ---------------------------------------------------------------------------

__global__ void set_int_kernel(int* p, int v) { *p = v; }

// what a captured cycle of one member depends on beyond device memory
// contents: its configuration generation and the level pointers it started
// from (eager steps in between swap them)
struct CaptureKey {
  unsigned long gen;
  const float* lvl0;
  const float* aux0[6];
  bool operator==(const CaptureKey& o) const
  {
    if (gen != o.gen || lvl0 != o.lvl0) return false;
    for (int f = 0; f < 6; f++)
      if (aux0[f] != o.aux0[f]) return false;
    return true;
  }
};

static CaptureKey capture_key(const cgfdh_solver* s)
{
  CaptureKey k;
  k.gen = s->gen;
  k.lvl0 = s->lvl[0];
  for (int f = 0; f < 6; f++) k.aux0[f] = s->face[f].aux[0];
  return k;
}

struct cgfdh_batch {
  std::vector<cgfdh_solver*> sv;
  std::vector<CaptureKey> keys;     // members' keys when the graph was captured
  hipStream_t stream = nullptr;
  int* it_off = nullptr;
  hipGraphExec_t exec = nullptr;
  int cap_it0 = -1;                 // step the captured cycle starts at
  std::vector<hipEvent_t> ev;       // fork / join
  hipEvent_t t[2];
  double last_ms = 0;
  // fused cycle: the whole-row stage of every grid in one launch; device
  // copies of the 32 stages' parameters of every member, their row boxes
  // and the block offsets of each member in the launch
  bool fused = false;
  StageParams* fz_p = nullptr;
  BoxParams* fz_b = nullptr;
  int* fz_start = nullptr;
};

static void batch_drop_graph(cgfdh_batch* b)
{
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  b->exec = nullptr;
  b->cap_it0 = -1;
}

extern "C" int cgfdh_batch_create(int n, cgfdh_solver* const* solvers, cgfdh_batch** out)
{
  if (n < 1 || !solvers || !out) return fail(CGFDH_EINVAL, "cgfdh_batch_create: bad argument");
  for (int i = 0; i < n; i++)
    if (!solvers[i]) return fail(CGFDH_EINVAL, "cgfdh_batch_create: solver %d is null", i);
  cgfdh_batch* b = new cgfdh_batch();
  b->sv.assign(solvers, solvers + n);
  bool bad = hipStreamCreate(&b->stream) != hipSuccess ||
             hipMalloc((void**)&b->it_off, sizeof(int)) != hipSuccess ||
             hipEventCreate(&b->t[0]) != hipSuccess || hipEventCreate(&b->t[1]) != hipSuccess;
  for (int i = 0; i < 2 * n + 1 && !bad; i++) {
    hipEvent_t e;
    bad = hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess;
    if (!bad) b->ev.push_back(e);
  }
  if (bad) {
    delete b;
    return fail(CGFDH_EDEVICE, "cgfdh_batch_create: stream / event / allocation failed");
  }
  *out = b;
  return CGFDH_OK;
}

extern "C" int cgfdh_batch_destroy(cgfdh_batch* b)
{
  if (!b) return CGFDH_OK;
  (void)hipStreamSynchronize(b->stream);
  batch_drop_graph(b);
  for (auto e : b->ev) (void)hipEventDestroy(e);
  (void)hipEventDestroy(b->t[0]);
  (void)hipEventDestroy(b->t[1]);
  (void)hipFree(b->it_off);
  (void)hipFree(b->fz_p);
  (void)hipFree(b->fz_b);
  (void)hipFree(b->fz_start);
  (void)hipStreamDestroy(b->stream);
  delete b;
  return CGFDH_OK;
}

// capture steps it0..it0+7 of every solver into one graph
static int batch_capture(cgfdh_batch* b, int it0)
{
  batch_drop_graph(b);
  b->keys.clear();
  for (auto s : b->sv) b->keys.push_back(capture_key(s));
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
  int rc = CGFDH_OK;
  const int n = (int)b->sv.size();
  hipError_t e = hipEventRecord(b->ev[2 * n], b->stream);
  for (int i = 0; i < n && e == hipSuccess && !rc; i++) {
    cgfdh_solver* s = b->sv[i];
    e = hipStreamWaitEvent(s->stream, b->ev[2 * n], 0);
    for (int q = 0; q < 8 && e == hipSuccess && !rc; q++)
      rc = enqueue_step(s, s->stream, it0 + q, b->it_off, nullptr);
    if (e == hipSuccess) e = hipEventRecord(b->ev[i], s->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev[i], 0);
  }
  hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (e == hipSuccess) e = ee;
  if (e == hipSuccess && !rc) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (rc) return rc;
  if (e != hipSuccess) {
    b->exec = nullptr;
    return fail(CGFDH_EDEVICE, "cgfdh_batch: graph capture failed: %s", hipGetErrorString(e));
  }
  b->cap_it0 = it0;
  b->fused = false;
  return CGFDH_OK;
}

// the fused cycle (opt-in: CGFDH_BATCH_FUSED=1) serves batches whose
// members all run variant 16 on one isotropic, VTI or visco-elastic medium
static bool batch_fusable(const cgfdh_batch* b)
{
  const char* e = getenv("CGFDH_BATCH_FUSED");
  if (!e || e[0] != '1') return false;
  const int mt = b->sv[0]->d.medium_type;
  if (mt != CGFDH_MEDIUM_ELASTIC_ISO && mt != CGFDH_MEDIUM_ELASTIC_VTI &&
      mt != CGFDH_MEDIUM_VISCO_ISO)
    return false;
  for (auto s : b->sv)
    if (s->variant != 16 || s->d.medium_type != mt) return false;
  return true;
}

// capture steps it0..it0+7 of every solver into one graph whose stages are
// single launches over all grids' whole rows (a cross-grid tile queue:
// stage_rows_multi_kernel); each member's free-surface rows and its
// post-step work run on its own stream beside them
static int batch_capture_fused(cgfdh_batch* b, int it0)
{
  batch_drop_graph(b);
  const int n = (int)b->sv.size();
  std::vector<BoxParams> boxes(n);
  std::vector<int> start(n + 1, 0);
  for (int i = 0; i < n; i++) {
    const long nb = start[i] + (long)rows_box(make_ctx(b->sv[i], b->stream), boxes[i]);
    if (nb > INT_MAX) return fail(CGFDH_EINVAL, "cgfdh_batch: fused launch too large");
    start[i + 1] = (int)nb;
  }
  (void)hipFree(b->fz_p);
  (void)hipFree(b->fz_b);
  (void)hipFree(b->fz_start);
  b->fz_p = nullptr; b->fz_b = nullptr; b->fz_start = nullptr;
  HIPCHK(hipMalloc((void**)&b->fz_p, (size_t)32 * n * sizeof(StageParams)));
  HIPCHK(hipMalloc((void**)&b->fz_b, (size_t)n * sizeof(BoxParams)));
  HIPCHK(hipMalloc((void**)&b->fz_start, (size_t)(n + 1) * sizeof(int)));
  HIPCHK(hipMemcpy(b->fz_b, boxes.data(), (size_t)n * sizeof(BoxParams), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->fz_start, start.data(), (size_t)(n + 1) * sizeof(int),
                   hipMemcpyHostToDevice));
  b->keys.clear();
  for (auto s : b->sv) b->keys.push_back(capture_key(s));
  // the member's level pointers as they stand before the cycle: capture
  // swaps them step by step, as replays will
  std::vector<StageParams> ph((size_t)32 * n);
  const int mt = b->sv[0]->d.medium_type;
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeGlobal));
  int rc = CGFDH_OK;
  hipError_t e = hipSuccess;
  auto fork = [&]() {
    if (e == hipSuccess) e = hipEventRecord(b->ev[2 * n], b->stream);
  };
  auto join = [&](int i) {
    if (e == hipSuccess) e = hipEventRecord(b->ev[i], b->sv[i]->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->ev[i], 0);
  };
  for (int q = 0; q < 8 && e == hipSuccess && !rc; q++) {
    const int it = it0 + q;
    const int ipair = it % 8;
    for (int sg = 0; sg < 4 && e == hipSuccess; sg++) {
      const int kind = sg == 0 ? KIND_FIRST : (sg == 3 ? KIND_LAST : KIND_MID);
      StageParams* pq = ph.data() + (size_t)(q * 4 + sg) * n;
      fork();
      for (int i = 0; i < n && e == hipSuccess; i++) {
        cgfdh_solver* s = b->sv[i];
        const float* wcur = sg == 0 ? s->lvl[0] : (sg % 2 == 1 ? s->lvl[1] : s->lvl[2]);
        float* wtmp = sg % 2 == 0 ? s->lvl[1] : s->lvl[2];
        pq[i] = make_params(s, it, sg, wcur, wtmp, sg == 0 ? 0 : 1);
        pq[i].it_off = b->it_off;
        if (!s->d.free_top) continue;
        e = hipStreamWaitEvent(s->stream, b->ev[2 * n], 0);
        if (e == hipSuccess) launch_stage(s, s->stream, pq[i], kind, ipair, sg, true);
        join(i);
      }
      if (e != hipSuccess) break;
      RowsMulti R;
      R.P = b->fz_p + (size_t)(q * 4 + sg) * n;
      R.B = b->fz_b;
      R.start = b->fz_start;
      R.n = n;
      const int dx = (kPairFlags[ipair][0] + sg) % 2;
      const int dy = (kPairFlags[ipair][1] + sg) % 2;
      const int dz = (kPairFlags[ipair][2] + sg) % 2;
      if (mt == CGFDH_MEDIUM_ELASTIC_VTI)
        launch_rows_multi_vti(b->stream, R, start[n], kind, dx, dy, dz);
      else if (mt == CGFDH_MEDIUM_VISCO_ISO)
        launch_rows_multi_visco(b->stream, R, start[n], kind, dx, dy, dz);
      else
        launch_rows_multi_iso(b->stream, R, start[n], kind, dx, dy, dz);
    }
    fork();
    for (int i = 0; i < n && e == hipSuccess && !rc; i++) {
      cgfdh_solver* s = b->sv[i];
      e = hipStreamWaitEvent(s->stream, b->ev[2 * n], 0);
      if (e == hipSuccess) rc = enqueue_post_step(s, s->stream, it, b->it_off);
      join(i);
    }
  }
  hipError_t ee = hipStreamEndCapture(b->stream, &g);
  if (e == hipSuccess) e = ee;
  if (e == hipSuccess && !rc) e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (g) (void)hipGraphDestroy(g);
  if (e == hipSuccess && !rc)
    e = hipMemcpy(b->fz_p, ph.data(), ph.size() * sizeof(StageParams), hipMemcpyHostToDevice);
  if (rc || e != hipSuccess) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
    // capture swapped the members' level pointers as the steps went; the
    // cycle never ran, so put back the pointers the keys describe
    for (int i = 0; i < n; i++) {
      cgfdh_solver* s = b->sv[i];
      if (s->lvl[0] != b->keys[i].lvl0) std::swap(s->lvl[0], s->lvl[3]);
      for (int f = 0; f < 6; f++)
        if (s->face[f].on && s->face[f].aux[0] != b->keys[i].aux0[f])
          std::swap(s->face[f].aux[0], s->face[f].aux[2]);
    }
    b->keys.clear();
    return rc ? rc
              : fail(CGFDH_EDEVICE, "cgfdh_batch: fused graph capture failed: %s",
                     hipGetErrorString(e));
  }
  b->cap_it0 = it0;
  b->fused = true;
  return CGFDH_OK;
}

extern "C" int cgfdh_batch_run(cgfdh_batch* b, int it0, int nsteps)
{
  if (!b || it0 < 0 || nsteps < 0) return fail(CGFDH_EINVAL, "cgfdh_batch_run: bad argument");
  for (size_t i = 0; i < b->sv.size(); i++) {
    cgfdh_solver* s = b->sv[i];
    int rc = ready(s, "cgfdh_batch_run");
    if (!rc) rc = dd_check(s, it0, nsteps, "cgfdh_batch_run");
    if (rc) return rc;
    if (it0 + nsteps > s->d.nt_total)
      return fail(CGFDH_EINVAL, "cgfdh_batch_run: steps %d..%d outside solver %zu's nt_total %d",
                  it0, it0 + nsteps, i, s->d.nt_total);
  }
  for (auto s : b->sv) HIPCHK(hipStreamSynchronize(s->stream));
  HIPCHK(hipEventRecord(b->t[0], b->stream));
  int it = it0;
  const int cycles = nsteps / 8;
  if (cycles > 0) {
    // a captured cycle serves every start step congruent to it mod 8, while
    // no member's configuration or level pointers have changed since
    const bool fuse = batch_fusable(b);
    bool stale = b->fused != fuse;
    for (size_t i = 0; i < b->sv.size() && b->exec && !stale; i++)
      stale = !(capture_key(b->sv[i]) == b->keys[i]);
    if (!b->exec || stale || (it - b->cap_it0) % 8 != 0 || it < b->cap_it0) {
      int rc = fuse ? batch_capture_fused(b, it) : batch_capture(b, it);
      if (rc) return rc;
    }
    for (int c = 0; c < cycles; c++, it += 8) {
      hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, b->stream, b->it_off,
                         it - b->cap_it0);
      HIPCHK(hipGraphLaunch(b->exec, b->stream));
    }
  }
  // the remaining 0..7 steps, eagerly, solver after solver on the batch stream
  for (; it < it0 + nsteps; it++)
    for (auto s : b->sv) {
      int rc = enqueue_step(s, b->stream, it, nullptr, nullptr);
      if (rc) return rc;
    }
  HIPCHK(hipEventRecord(b->t[1], b->stream));
  HIPCHK(hipStreamSynchronize(b->stream));
  HIPCHK(hipGetLastError());
  float ms = 0;
  HIPCHK(hipEventElapsedTime(&ms, b->t[0], b->t[1]));
  b->last_ms = ms;
  return CGFDH_OK;
}

extern "C" int cgfdh_batch_last_timing(cgfdh_batch* b, double* ms)
{
  if (!b || !ms) return fail(CGFDH_EINVAL, "cgfdh_batch_last_timing: bad argument");
  *ms = b->last_ms;
  return CGFDH_OK;
}

// autotune: per solver, time one step of each candidate kernel variant and
// tile group from the current state (restored after every trial) and keep
// the fastest.  Restores the state, so the caller's run starts unchanged.
extern "C" int cgfdh_batch_autotune(cgfdh_batch* b, int it0)
{
  if (!b) return fail(CGFDH_EINVAL, "cgfdh_batch_autotune: null batch");
  batch_drop_graph(b);
  for (size_t i = 0; i < b->sv.size(); i++) {
    cgfdh_solver* s = b->sv[i];
    int rc = ready(s, "cgfdh_batch_autotune");
    if (rc) return rc;
    if (it0 + 1 > s->d.nt_total) return fail(CGFDH_EINVAL, "cgfdh_batch_autotune: it0 %d", it0);
    // saved state: the pre level, the PML aux pre levels, peak ground motion
    const size_t lv = (size_t)s->nw * s->vol;
    size_t naux = 0;
    for (auto& f : s->face)
      if (f.on) naux += 9 * f.fn;
    const size_t npg = s->pg ? 18 * (size_t)s->ss : 0;
    float* save = nullptr;
    HIPCHK(hipMalloc((void**)&save, (lv + naux + npg + 1) * sizeof(float)));
    auto copy_state = [&](bool to_save) -> hipError_t {
      hipError_t e = to_save ? hipMemcpyAsync(save, s->lvl[0], lv * 4, hipMemcpyDeviceToDevice,
                                              s->stream)
                             : hipMemcpyAsync(s->lvl[0], save, lv * 4, hipMemcpyDeviceToDevice,
                                              s->stream);
      size_t off = lv;
      for (auto& f : s->face) {
        if (!f.on || e != hipSuccess) continue;
        e = to_save ? hipMemcpyAsync(save + off, f.aux[0], 9 * f.fn * 4,
                                     hipMemcpyDeviceToDevice, s->stream)
                    : hipMemcpyAsync(f.aux[0], save + off, 9 * f.fn * 4,
                                     hipMemcpyDeviceToDevice, s->stream);
        off += 9 * f.fn;
      }
      if (npg && e == hipSuccess)
        e = to_save ? hipMemcpyAsync(save + off, s->pg, npg * 4, hipMemcpyDeviceToDevice, s->stream)
                    : hipMemcpyAsync(s->pg, save + off, npg * 4, hipMemcpyDeviceToDevice, s->stream);
      return e;
    };
    hipError_t e = copy_state(true);
    const bool iso = s->d.medium_type == CGFDH_MEDIUM_ELASTIC_ISO;
    const bool visco = s->d.medium_type == CGFDH_MEDIUM_VISCO_ISO;
    std::vector<int> variants = visco ? std::vector<int>{4, 5, 8, 9, 16}
                                      : std::vector<int>{4, 6, 7, 8, 12, 14, 16, 18};
    if (iso) variants.push_back(3);
    const int groups[][2] = {{9, 2}, {3, 4}, {9, 8}};
    float best = 1e30f;
    int bv = s->variant, bgx = s->tgx, bgy = s->tgy;
    hipEvent_t t0 = b->t[0], t1 = b->t[1];
    for (int v : variants)
      for (auto& g : groups) {
        if (e != hipSuccess) break;
        s->variant = v;
        s->tgx = g[0];
        s->tgy = g[1];
        float ms = 0;
        for (int rep = 0; rep < 2 && e == hipSuccess; rep++) {   // first run warms up
          e = copy_state(false);
          if (e == hipSuccess) e = hipEventRecord(t0, s->stream);
          if (e == hipSuccess && enqueue_step(s, s->stream, it0, nullptr, nullptr)) e = hipErrorUnknown;
          if (e == hipSuccess) e = hipEventRecord(t1, s->stream);
          if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
          if (e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
          std::swap(s->lvl[0], s->lvl[3]);      // undo the step's swaps
          for (auto& f : s->face)
            if (f.on) std::swap(f.aux[0], f.aux[2]);
        }
        if (e == hipSuccess && ms < best) { best = ms; bv = v; bgx = g[0]; bgy = g[1]; }
      }
    if (e == hipSuccess) e = copy_state(false);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(save);
    s->variant = bv;
    s->tgx = bgx;
    s->tgy = bgy;
    if (e != hipSuccess)
      return fail(CGFDH_EDEVICE, "cgfdh_batch_autotune: %s", hipGetErrorString(e));
  }
  return CGFDH_OK;
}

extern "C" int cgfdh_get_kernel_choice(cgfdh_solver* s, int* variant, int* gx, int* gy)
{
  if (!s) return fail(CGFDH_EINVAL, "cgfdh_get_kernel_choice: null solver");
  if (variant) *variant = s->variant;
  if (gx) *gx = s->tgx;
  if (gy) *gy = s->tgy;
  return CGFDH_OK;
}
