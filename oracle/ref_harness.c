/*
 * ref_harness.c -- test-only driver around the reference's own hot-path
 * sources (fd_t.c, sv_curv_col_el.c, sv_curv_col_el_iso.c,
 * sv_curv_col_el_vti.c, sv_curv_col_el_aniso.c, sv_curv_col_vis_iso.c, bdry_t.c, wav_t.c,
 * blk_t.c, fdlib_math.c, fdlib_mem.c), compiled unchanged from /root/reference by
 * oracle/Makefile into oracle/_ref/libcgfd3d_ref.so.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product links or loads this.
 *
 * The reference files that need netCDF (gd_t.c, md_t.c, src_t.c, io_funcs.c)
 * cannot be compiled in this image, so this harness fills the reference
 * structs directly and re-implements only the glue the reference's time loop
 * performs around the RHS (drv_rk_curv_col.c:167-542): the RK4 combine over
 * whole levels (including ghosts), the PML auxiliary combine, the pre/end swap,
 * the trilinear receiver sampling (io_funcs.c:1585-1623, with its duplicated
 * indx1d[5] term) and the ghost zeroing of the rhs level (wav_t.c:276-378 is
 * called as the reference function itself).  Sources are point moment
 * sources only: the Gaussian-smoothed path calls src_cal_norm_delt3d_z2fre and
 * gd_lindx_is_inner, which live in the netCDF-dependent files.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <math.h>

#include "constants.h"
#include "fdlib_mem.h"
#include "fdlib_math.h"
#include "fd_t.h"
#include "gd_t.h"
#include "md_t.h"
#include "wav_t.h"
#include "src_t.h"
#include "bdry_t.h"
#include "sv_curv_col_el.h"
#include "sv_curv_col_el_iso.h"
#include "sv_curv_col_el_vti.h"
#include "sv_curv_col_el_aniso.h"
#include "sv_curv_col_vis_iso.h"
#include "blk_t.h"

typedef struct {
  int ni, nj, nk;
  int free_top;
  int pml_on[6];          /* face order idim*2+iside: x1,x2,y1,y2,z1,z2 */
  int pml_layers[6];
  float pml_alpha[6];
  float pml_beta[6];
  float pml_vel[6];
  /* point moment sources, reference layout M??[is*max_nt*4 + it*4 + istage] */
  int nsrc;
  int src_max_nt;
  const int *src_ijk;     /* 3*nsrc, local indices including ghosts */
  const int *src_it_begin;
  const int *src_it_end;
  const float *src_mij;   /* 6 arrays back to back: Mxx,Myy,Mzz,Mxz,Myz,Mxy */
  /* CONST_MEDIUM_ELASTIC_ISO (or 0), _VTI or _ANISO; the medium array holds
   * lambda,mu,slw / c11,c13,c33,c55,c66,slw / the 21 Cij (c11,c12,..,c66) and
   * slw, one volume each */
  int medium_type;
  /* CONST_MEDIUM_VISCOELASTIC_ISO: lambda, mu, slw, then Ylam_n, Ymu_n per
   * Maxwell body; wl the bodies' relaxation frequencies */
  int nmaxwell;
  const float *visco_wl;
  /* point force sources: Fx, Fy, Fz back to back, same layout as src_mij;
   * NULL for none (src_mij may then be NULL too) */
  const float *src_fxyz;
  /* strict surface force: src->is_surface_force_strict, and the surface
   * traction / velocity-source slices TxSrc,TySrc,TzSrc,VxSrc,VySrc,VzSrc
   * ([6][ny][nx]) of every step and stage, [surf_nt][4][6][ny][nx], as
   * src_set_surface_layer_for_force (src_t.c:154-314, in a file that needs
   * netCDF) leaves them before each stage (drv_rk_curv_col.c:227-230).
   * refh_onestage reads entry 0.  NULL: all zero. */
  int src_force_strict;
  const float *surf_slices;
  int surf_nt;
  /* dd sources at points (src_t dd_* members): dd_indx the 1-D indices,
   * dd_vi [dd_max_nt][4][ndd][3] and dd_mij [dd_max_nt][4][ndd][6] the whole
   * value tables (one read of dd_max_nt steps), NULL for a kind not given */
  int ndd;
  int dd_max_nt;
  const size_t *dd_indx;
  const float *dd_vi;
  const float *dd_mij;
} refh_cfg;

typedef struct {
  fd_t fd;
  gd_t gd;
  gdcurv_metric_t metric;
  md_t md;
  wav_t wav;
  bdry_t bdry;
  src_t src;
  float *T0;   /* surface traction / velocity-source slices, 6 x [ny][nx] */
  const float *surf;   /* the configured slices per step and stage, or NULL */
  int surf_nt;
  float *Ylam[16], *Ymu[16], wl[16];
} refh_ctx;

static void fill_gd(gd_t *gd, int ni, int nj, int nk,
                    float *x3d, float *y3d, float *z3d)
{
  memset(gd, 0, sizeof(*gd));
  gd->type = GD_TYPE_CURV;
  gd->ni = ni; gd->nj = nj; gd->nk = nk;
  gd->nx = ni + 6; gd->ny = nj + 6; gd->nz = nk + 6;
  gd->ni1 = 3; gd->ni2 = 3 + ni - 1;
  gd->nj1 = 3; gd->nj2 = 3 + nj - 1;
  gd->nk1 = 3; gd->nk2 = 3 + nk - 1;
  gd->npoint_ghosts = 3;
  gd->fdx_nghosts = gd->fdy_nghosts = gd->fdz_nghosts = 3;
  gd->siz_iy = gd->siz_line = gd->nx;
  gd->siz_iz = gd->siz_slice = (size_t)gd->nx * gd->ny;
  gd->siz_icmp = gd->siz_volume = (size_t)gd->nx * gd->ny * gd->nz;
  gd->x3d = x3d; gd->y3d = y3d; gd->z3d = z3d;
}

static int medium_of(const refh_cfg *cfg)
{
  return cfg->medium_type ? cfg->medium_type : CONST_MEDIUM_ELASTIC_ISO;
}

static void setup(refh_ctx *c, const refh_cfg *cfg, float *coords,
                  float *metric10, float *med)
{
  memset(c, 0, sizeof(*c));
  fd_set_macdrp(&c->fd);
  size_t vol = (size_t)(cfg->ni + 6) * (cfg->nj + 6) * (cfg->nk + 6);
  fill_gd(&c->gd, cfg->ni, cfg->nj, cfg->nk,
          coords, coords ? coords + vol : NULL, coords ? coords + 2 * vol : NULL);
  gd_t *gd = &c->gd;

  gdcurv_metric_t *m = &c->metric;
  m->nx = gd->nx; m->ny = gd->ny; m->nz = gd->nz; m->ncmp = 10;
  m->siz_iy = gd->siz_iy; m->siz_iz = gd->siz_iz; m->siz_icmp = gd->siz_icmp;
  m->v4d = metric10;
  m->jac = metric10;
  m->xi_x = metric10 + 1 * vol; m->xi_y = metric10 + 2 * vol; m->xi_z = metric10 + 3 * vol;
  m->eta_x = metric10 + 4 * vol; m->eta_y = metric10 + 5 * vol; m->eta_z = metric10 + 6 * vol;
  m->zeta_x = metric10 + 7 * vol; m->zeta_y = metric10 + 8 * vol; m->zeta_z = metric10 + 9 * vol;

  md_t *md = &c->md;
  md->nx = gd->nx; md->ny = gd->ny; md->nz = gd->nz;
  md->siz_iy = gd->siz_iy; md->siz_iz = gd->siz_iz; md->siz_icmp = gd->siz_icmp;
  md->medium_type = medium_of(cfg);
  if (md->medium_type == CONST_MEDIUM_ELASTIC_VTI) {
    float **f[5] = {&md->c11, &md->c13, &md->c33, &md->c55, &md->c66};
    for (int q = 0; q < 5; q++) *f[q] = med + q * vol;
    md->rho = med + 5 * vol;
  } else if (md->medium_type == CONST_MEDIUM_ELASTIC_ANISO) {
    float **f[21] = {&md->c11, &md->c12, &md->c13, &md->c14, &md->c15, &md->c16,
                     &md->c22, &md->c23, &md->c24, &md->c25, &md->c26,
                     &md->c33, &md->c34, &md->c35, &md->c36,
                     &md->c44, &md->c45, &md->c46, &md->c55, &md->c56, &md->c66};
    for (int q = 0; q < 21; q++) *f[q] = med + q * vol;
    md->rho = med + 21 * vol;
  } else {
    md->lambda = med; md->mu = med + vol; md->rho = med + 2 * vol;
  }
  int visco = md->medium_type == CONST_MEDIUM_VISCOELASTIC_ISO;
  if (visco) {
    md->nmaxwell = cfg->nmaxwell;
    md->visco_type = CONST_VISCO_GMB;
    for (int n = 0; n < cfg->nmaxwell; n++) {
      c->Ylam[n] = med + (3 + 2 * n) * vol;
      c->Ymu[n] = med + (4 + 2 * n) * vol;
      c->wl[n] = cfg->visco_wl[n];
    }
    md->Ylam = c->Ylam;
    md->Ymu = c->Ymu;
    md->wl = c->wl;
  }

  wav_init(gd, &c->wav, 4, visco ? CONST_VISCO_GMB : 0, visco ? cfg->nmaxwell : 0);

  int neighid[6] = {MPI_PROC_NULL, MPI_PROC_NULL, MPI_PROC_NULL,
                    MPI_PROC_NULL, MPI_PROC_NULL, MPI_PROC_NULL};
  int is_free[3][2] = {{0, 0}, {0, 0}, {0, cfg->free_top ? 1 : 0}};
  int is_pml[3][2], nlay[3][2];
  float amax[3][2], bmax[3][2], vel[3][2];
  for (int f = 0; f < 6; f++) {
    is_pml[f / 2][f % 2] = cfg->pml_on[f];
    nlay[f / 2][f % 2] = cfg->pml_on[f] ? cfg->pml_layers[f] : 0;
    amax[f / 2][f % 2] = cfg->pml_alpha[f];
    bmax[f / 2][f % 2] = cfg->pml_beta[f];
    vel[f / 2][f % 2] = cfg->pml_vel[f];
  }
  bdry_init(&c->bdry, gd->nx, gd->ny, gd->nz);
  bdry_free_set(gd, &c->bdry, neighid, is_free, 0);
  bdry_pml_set(gd, &c->wav, &c->bdry, neighid, is_pml, nlay, amax, bmax, vel, 0);

  src_t *s = &c->src;
  c->T0 = (float *)calloc(6 * gd->siz_iz, sizeof(float));
  s->TxSrc = c->T0; s->TySrc = c->T0 + gd->siz_iz; s->TzSrc = c->T0 + 2 * gd->siz_iz;
  s->VxSrc = c->T0 + 3 * gd->siz_iz; s->VySrc = c->T0 + 4 * gd->siz_iz;
  s->VzSrc = c->T0 + 5 * gd->siz_iz;
  c->surf = cfg->surf_slices;
  c->surf_nt = cfg->surf_nt;
  s->total_number = cfg->nsrc;
  s->max_stage = 4;
  s->dd_is_valid = cfg->ndd > 0;
  if (cfg->ndd > 0) {
    s->dd_is_add_at_point = 1;
    s->dd_total_number = cfg->ndd;
    s->dd_max_nt = cfg->dd_max_nt;
    s->dd_nt_per_read = cfg->dd_max_nt;
    s->dd_it_here = 0;
    s->dd_indx = (size_t *)cfg->dd_indx;
    s->dd_vi = (float *)cfg->dd_vi;
    s->dd_mij = (float *)cfg->dd_mij;
    s->dd_vi_actived = cfg->dd_vi != NULL;
    s->dd_mij_actived = cfg->dd_mij != NULL;
  }
  if (cfg->nsrc > 0) {
    s->max_nt = cfg->src_max_nt;
    s->max_stage = 4;
    s->itype_spatial_ext = CONST_SRC_SPATIAL_POINT;
    s->force_actived = cfg->src_fxyz != NULL;
    s->moment_actived = cfg->src_mij != NULL;
    s->is_surface_force_strict = cfg->src_force_strict;
    s->si = (int *)malloc(cfg->nsrc * sizeof(int));
    s->sj = (int *)malloc(cfg->nsrc * sizeof(int));
    s->sk = (int *)malloc(cfg->nsrc * sizeof(int));
    s->it_begin = (int *)malloc(cfg->nsrc * sizeof(int));
    s->it_end = (int *)malloc(cfg->nsrc * sizeof(int));
    for (int is = 0; is < cfg->nsrc; is++) {
      s->si[is] = cfg->src_ijk[3 * is + 0];
      s->sj[is] = cfg->src_ijk[3 * is + 1];
      s->sk[is] = cfg->src_ijk[3 * is + 2];
      s->it_begin[is] = cfg->src_it_begin[is];
      s->it_end[is] = cfg->src_it_end[is];
    }
    size_t n = (size_t)cfg->nsrc * cfg->src_max_nt * 4;
    if (cfg->src_mij) {
      float *mij = (float *)cfg->src_mij;
      s->Mxx = mij + 0 * n; s->Myy = mij + 1 * n; s->Mzz = mij + 2 * n;
      s->Mxz = mij + 3 * n; s->Myz = mij + 4 * n; s->Mxy = mij + 5 * n;
    }
    if (cfg->src_fxyz) {
      float *f = (float *)cfg->src_fxyz;
      s->Fx = f; s->Fy = f + n; s->Fz = f + 2 * n;
    }
  }
}

static void teardown(refh_ctx *c)
{
  free(c->wav.v5d);
  for (int d = 0; d < 3; d++)
    for (int s = 0; s < 2; s++) {
      free(c->bdry.auxvar[d][s].var);
      free(c->bdry.A[d][s]); free(c->bdry.B[d][s]); free(c->bdry.D[d][s]);
    }
  free(c->bdry.matVx2Vz2); free(c->bdry.matVy2Vz2);
  free(c->bdry.matF2Vz2); free(c->bdry.matD);
  free(c->wav.cmp_pos);
  if (c->wav.Jxx_pos) {
    free(c->wav.Jxx_pos); free(c->wav.Jyy_pos); free(c->wav.Jzz_pos);
    free(c->wav.Jxy_pos); free(c->wav.Jxz_pos); free(c->wav.Jyz_pos);
    free(c->wav.Jxx_seq); free(c->wav.Jyy_seq); free(c->wav.Jzz_seq);
    free(c->wav.Jxy_seq); free(c->wav.Jxz_seq); free(c->wav.Jyz_seq);
  }
  free(c->T0);
  free(c->src.si); free(c->src.sj); free(c->src.sk);
  free(c->src.it_begin); free(c->src.it_end);
}


/* ---- the RHS sequence of sv_curv_col_el_iso_onestage --------------------
 * sv_curv_col_el_iso.c:23-202 calls inner -> timg -> vlow -> cfspml -> src.
 * The harness makes the same calls to the reference functions, and (when
 * patch != 0) re-evaluates the one row the reference leaves undefined: with
 * the forward z operator {-1..3} the traction-image mirror at k = nk2 reads
 * veczt[-1] and veczt[-2] (sv_curv_col_el.c:160-163, out of bounds; host
 * ASan reports a dynamic-stack-buffer-overflow).  The re-evaluation takes the
 * mirror values from the grid points they stand for, k-2 and k-3, as the
 * reference's commented-out formula does (sv_curv_col_el.c:151-157), and runs
 * before vlow/cfspml/src so their updates land on top as in the reference. */
static void patch_timg_surface_row(refh_ctx *c, float *w_cur, float *rhs, fd_op_t *fx,
                                   fd_op_t *fy, fd_op_t *fz)
{
  gd_t *gd = &c->gd;
  size_t V = gd->siz_icmp, sl = gd->siz_iy, ss = gd->siz_iz;
  gdcurv_metric_t *m = &c->metric;
  int k = gd->nk2;
  int n_free = -fz->indx[0];           /* = 1 for the {-1..3} operator */
  const size_t rows[3][3] = {{c->wav.Txx_pos, c->wav.Txy_pos, c->wav.Txz_pos},
                             {c->wav.Txy_pos, c->wav.Tyy_pos, c->wav.Tyz_pos},
                             {c->wav.Txz_pos, c->wav.Tyz_pos, c->wav.Tzz_pos}};
  const size_t hpos[3] = {c->wav.Vx_pos, c->wav.Vy_pos, c->wav.Vz_pos};
  for (int j = gd->nj1; j <= gd->nj2; j++)
    for (int i = gd->ni1; i <= gd->ni2; i++) {
      size_t p = i + j * sl + k * ss;
      float slwjac = c->md.rho[p] / m->jac[p];
      for (int comp = 0; comp < 3; comp++) {
        float *T0 = w_cur + rows[comp][0], *T1 = w_cur + rows[comp][1], *T2 = w_cur + rows[comp][2];
        float vx[5], vy[5], vz[5];
        for (int n = 0; n < fx->total_len; n++) {
          size_t q = p + fx->indx[n];
          vx[n] = m->jac[q] * (m->xi_x[q] * T0[q] + m->xi_y[q] * T1[q] + m->xi_z[q] * T2[q]);
        }
        for (int n = 0; n < fy->total_len; n++) {
          size_t q = p + fy->indx[n] * sl;
          vy[n] = m->jac[q] * (m->eta_x[q] * T0[q] + m->eta_y[q] * T1[q] + m->eta_z[q] * T2[q]);
        }
        for (int n = 0; n < n_free; n++) {
          size_t q = p + fz->indx[n] * ss;
          vz[n] = m->jac[q] * (m->zeta_x[q] * T0[q] + m->zeta_y[q] * T1[q] + m->zeta_z[q] * T2[q]);
        }
        const float *TS = comp == 0 ? c->src.TxSrc : (comp == 1 ? c->src.TySrc : c->src.TzSrc);
        const float ts = TS[i + j * sl];
        vz[n_free] = ts;
        for (int n = n_free + 1; n < fz->total_len; n++) {
          long off = fz->indx[n] - 2 * (n - n_free);
          size_t q = p + off * (long)ss;
          float below = m->jac[q] * (m->zeta_x[q] * T0[q] + m->zeta_y[q] * T1[q] + m->zeta_z[q] * T2[q]);
          vz[n] = 2.0 * ts - below;
        }
        float DxT = fx->coef[0] * vx[0], DyT = fy->coef[0] * vy[0], DzT = fz->coef[0] * vz[0];
        for (int n = 1; n < fx->total_len; n++) DxT += fx->coef[n] * vx[n];
        for (int n = 1; n < fy->total_len; n++) DyT += fy->coef[n] * vy[n];
        for (int n = 1; n < fz->total_len; n++) DzT += fz->coef[n] * vz[n];
        rhs[hpos[comp] + p] = (DxT + DyT + DzT) * slwjac;
      }
    }
}

/* the surface-force slices of step it, stage istage into src_t */
static void load_surface_slices(refh_ctx *c, int it, int istage)
{
  size_t n = 6 * c->gd.siz_iz;
  if (!c->surf || it >= c->surf_nt) {
    memset(c->T0, 0, n * sizeof(float));
    return;
  }
  memcpy(c->T0, c->surf + ((size_t)it * 4 + istage) * n, n * sizeof(float));
}

#define METRIC9(m) (m)->xi_x, (m)->xi_y, (m)->xi_z, (m)->eta_x, (m)->eta_y, (m)->eta_z, \
                   (m)->zeta_x, (m)->zeta_y, (m)->zeta_z
#define VTI_C(md) (md)->c11, (md)->c13, (md)->c33, (md)->c55, (md)->c66
#define ANISO_C(md) (md)->c11, (md)->c12, (md)->c13, (md)->c14, (md)->c15, (md)->c16, \
                    (md)->c22, (md)->c23, (md)->c24, (md)->c25, (md)->c26, (md)->c33, \
                    (md)->c34, (md)->c35, (md)->c36, (md)->c44, (md)->c45, (md)->c46, \
                    (md)->c55, (md)->c56, (md)->c66

/* sv_curv_col_el_{iso,vti,aniso}_onestage: inner -> timg -> vlow -> cfspml -> src */
static void harness_onestage(refh_ctx *c, float *w_cur, float *rhs, int ipair, int istage,
                             int patch)
{
  wav_t *wav = &c->wav;
  gd_t *gd = &c->gd;
  gdcurv_metric_t *m = &c->metric;
  md_t *md = &c->md;
  fd_op_t *fx = c->fd.pair_fdx_op[ipair][istage] + (c->fd.num_of_fdx_op - 1);
  fd_op_t *fy = c->fd.pair_fdy_op[ipair][istage] + (c->fd.num_of_fdy_op - 1);
  fd_op_t *fzl = c->fd.pair_fdz_op[ipair][istage];
  fd_op_t *fz = fzl + (c->fd.num_of_fdz_op - 1);
  float *Vx = w_cur + wav->Vx_pos, *Vy = w_cur + wav->Vy_pos, *Vz = w_cur + wav->Vz_pos;
  float *Txx = w_cur + wav->Txx_pos, *Tyy = w_cur + wav->Tyy_pos, *Tzz = w_cur + wav->Tzz_pos;
  float *Txz = w_cur + wav->Txz_pos, *Tyz = w_cur + wav->Tyz_pos, *Txy = w_cur + wav->Txy_pos;
  float *hVx = rhs + wav->Vx_pos, *hVy = rhs + wav->Vy_pos, *hVz = rhs + wav->Vz_pos;
  float *hTxx = rhs + wav->Txx_pos, *hTyy = rhs + wav->Tyy_pos, *hTzz = rhs + wav->Tzz_pos;
  float *hTxz = rhs + wav->Txz_pos, *hTyz = rhs + wav->Tyz_pos, *hTxy = rhs + wav->Txy_pos;
  size_t sl = gd->siz_line, ss = gd->siz_slice;
  int mt = md->medium_type;
  if (mt == CONST_MEDIUM_ELASTIC_VTI)
    sv_curv_col_el_vti_rhs_inner(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                 hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                 VTI_C(md), md->rho,
                                 gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                 fx->total_len, fx->indx, fx->coef,
                                 fy->total_len, fy->indx, fy->coef,
                                 fz->total_len, fz->indx, fz->coef, 0, 0);
  else if (mt == CONST_MEDIUM_ELASTIC_ANISO)
    sv_curv_col_el_aniso_rhs_inner(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                   hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                   ANISO_C(md), md->rho,
                                   gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                   fx->total_len, fx->indx, fx->coef,
                                   fy->total_len, fy->indx, fy->coef,
                                   fz->total_len, fz->indx, fz->coef, 0, 0);
  else
    sv_curv_col_el_iso_rhs_inner(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                 hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                                 m->xi_x, m->xi_y, m->xi_z, m->eta_x, m->eta_y, m->eta_z,
                                 m->zeta_x, m->zeta_y, m->zeta_z, md->lambda, md->mu, md->rho,
                                 gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                 fx->total_len, fx->indx, fx->coef,
                                 fy->total_len, fy->indx, fy->coef,
                                 fz->total_len, fz->indx, fz->coef, 0, 0);
  if (c->bdry.is_sides_free[2][1] == 1) {
    sv_curv_col_el_rhs_timg_z2(Txx, Tyy, Tzz, Txz, Tyz, Txy, hVx, hVy, hVz,
                               m->xi_x, m->xi_y, m->xi_z, m->eta_x, m->eta_y, m->eta_z,
                               m->zeta_x, m->zeta_y, m->zeta_z, m->jac, md->rho,
                               c->src.TxSrc, c->src.TySrc, c->src.TzSrc,
                               gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                               fx->total_len, fx->indx, fx->coef,
                               fy->total_len, fy->indx, fy->coef,
                               fz->total_len, fz->indx, fz->coef, 0, 0);
    if (patch && fz->indx[0] == -1) patch_timg_surface_row(c, w_cur, rhs, fx, fy, fz);
    if (mt == CONST_MEDIUM_ELASTIC_VTI)
      sv_curv_col_el_vti_rhs_vlow_z2(Vx, Vy, Vz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                     VTI_C(md), md->rho, c->bdry.matVx2Vz2, c->bdry.matVy2Vz2,
                                     gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                     fx->total_len, fx->indx, fx->coef,
                                     fy->total_len, fy->indx, fy->coef,
                                     c->fd.num_of_fdz_op, fzl, c->fd.fdz_max_len, 0, 0);
    else if (mt == CONST_MEDIUM_ELASTIC_ANISO)
      sv_curv_col_el_aniso_rhs_vlow_z2(Vx, Vy, Vz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                       ANISO_C(md), md->rho, c->bdry.matVx2Vz2, c->bdry.matVy2Vz2,
                                       gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                       fx->total_len, fx->indx, fx->coef,
                                       fy->total_len, fy->indx, fy->coef,
                                       c->fd.num_of_fdz_op, fzl, c->fd.fdz_max_len, 0, 0);
    else
      sv_curv_col_el_iso_rhs_vlow_z2(Vx, Vy, Vz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                                     m->xi_x, m->xi_y, m->xi_z, m->eta_x, m->eta_y, m->eta_z,
                                     m->zeta_x, m->zeta_y, m->zeta_z, md->lambda, md->mu, md->rho,
                                     c->bdry.matVx2Vz2, c->bdry.matVy2Vz2, c->bdry.matF2Vz2,
                                     c->src.VxSrc, c->src.VySrc, c->src.VzSrc,
                                     gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                                     fx->total_len, fx->indx, fx->coef,
                                     fy->total_len, fy->indx, fy->coef,
                                     c->fd.num_of_fdz_op, fzl, c->fd.fdz_max_len, 0, 0);
  }
  if (c->bdry.is_enable_pml == 1 && mt == CONST_MEDIUM_ELASTIC_VTI)
    sv_curv_col_el_vti_rhs_cfspml(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                  hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                  VTI_C(md), md->rho, gd->nk2, sl, ss,
                                  fx->total_len, fx->indx, fx->coef,
                                  fy->total_len, fy->indx, fy->coef,
                                  fz->total_len, fz->indx, fz->coef, &c->bdry, 0, 0);
  else if (c->bdry.is_enable_pml == 1 && mt == CONST_MEDIUM_ELASTIC_ANISO)
    sv_curv_col_el_aniso_rhs_cfspml(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                    hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy, METRIC9(m),
                                    ANISO_C(md), md->rho, gd->nk2, sl, ss,
                                    fx->total_len, fx->indx, fx->coef,
                                    fy->total_len, fy->indx, fy->coef,
                                    fz->total_len, fz->indx, fz->coef, &c->bdry, 0, 0);
  else
  if (c->bdry.is_enable_pml == 1)
    sv_curv_col_el_iso_rhs_cfspml(Vx, Vy, Vz, Txx, Tyy, Tzz, Txz, Tyz, Txy,
                                  hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                                  m->xi_x, m->xi_y, m->xi_z, m->eta_x, m->eta_y, m->eta_z,
                                  m->zeta_x, m->zeta_y, m->zeta_z, md->lambda, md->mu, md->rho,
                                  gd->nk2, sl, ss,
                                  fx->total_len, fx->indx, fx->coef,
                                  fy->total_len, fy->indx, fy->coef,
                                  fz->total_len, fz->indx, fz->coef, &c->bdry, 0, 0);
  if (mt == CONST_MEDIUM_VISCOELASTIC_ISO) {
    /* sv_curv_col_vis_iso.c:207-215 */
    int nm = md->nmaxwell;
    float *J[6][16], *hJ[6][16];
    for (int n = 0; n < nm; n++) {
      const size_t pos[6] = {wav->Jxx_pos[n], wav->Jyy_pos[n], wav->Jzz_pos[n],
                             wav->Jxy_pos[n], wav->Jxz_pos[n], wav->Jyz_pos[n]};
      for (int q = 0; q < 6; q++) { J[q][n] = w_cur + pos[q]; hJ[q][n] = rhs + pos[q]; }
    }
    sv_curv_col_vis_iso_atten(hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                              J[0], J[1], J[2], J[3], J[4], J[5],
                              hJ[0], hJ[1], hJ[2], hJ[3], hJ[4], hJ[5],
                              md->lambda, md->mu, md->rho, md->wl, md->Ylam, md->Ymu,
                              gd->ni1, gd->ni2, gd->nj1, gd->nj2, gd->nk1, gd->nk2, sl, ss,
                              nm, 0, 0);
  }
  if (c->src.total_number > 0)
    sv_curv_col_el_rhs_src(hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                           m->jac, md->rho, gd, &c->src, 0, 0);
  /* sv_curv_col_el_iso.c:193-199, sv_curv_col_vis_iso.c:235-241 */
  if (c->src.dd_is_valid == 1 &&
      (mt == CONST_MEDIUM_ELASTIC_ISO || mt == CONST_MEDIUM_VISCOELASTIC_ISO))
    sv_curv_col_el_rhs_srcdd(hVx, hVy, hVz, hTxx, hTyy, hTzz, hTxz, hTyz, hTxy,
                             m->jac, md->rho, &c->src, 0, 0);
}

/* ---- operator tables: [pair 8][stage 4][dim 3 (z has 4 layers)] -------- */
/* indx_out[((p*4+s)*6 + slot)*5 + n], slot 0=x, 1=y, 2..5 = z layer 0..3 */
int refh_fd_ops(int *indx_out, float *coef_out, int *len_out,
                float *rk_a, float *rk_b, float *rk_t)
{
  fd_t fd;
  memset(&fd, 0, sizeof(fd));
  fd_set_macdrp(&fd);
  for (int p = 0; p < 8; p++)
    for (int s = 0; s < 4; s++)
      for (int slot = 0; slot < 6; slot++) {
        fd_op_t *op = slot == 0 ? fd.pair_fdx_op[p][s]
                    : slot == 1 ? fd.pair_fdy_op[p][s]
                                : fd.pair_fdz_op[p][s] + (slot - 2);
        int base = (p * 4 + s) * 6 + slot;
        len_out[base] = op->total_len;
        for (int n = 0; n < 5; n++) {
          indx_out[base * 5 + n] = n < op->total_len ? op->indx[n] : 0;
          coef_out[base * 5 + n] = n < op->total_len ? op->coef[n] : 0.0f;
        }
      }
  for (int s = 0; s < 4; s++) {
    rk_b[s] = fd.rk_b[s];
    rk_t[s] = fd.rk_rhs_time[s];
    if (s < 3) rk_a[s] = fd.rk_a[s];
  }
  return fd.num_of_pairs;
}

/* ---- PML profiles and regions (bdry_pml_set) ---------------------------- */
/* A/B/D_out: 6 faces x maxlen; region_out: 6 faces x (ni1,ni2,nj1,nj2,nk1,nk2) */
int refh_pml(const refh_cfg *cfg, float *coords, int maxlen,
             float *A_out, float *B_out, float *D_out, int *region_out)
{
  size_t vol = (size_t)(cfg->ni + 6) * (cfg->nj + 6) * (cfg->nk + 6);
  float *metric = (float *)calloc(10 * vol, sizeof(float));
  float *med = (float *)calloc(3 * vol, sizeof(float));
  refh_ctx c;
  setup(&c, cfg, coords, metric, med);
  for (int f = 0; f < 6; f++) {
    int d = f / 2, s = f % 2;
    region_out[f * 6 + 0] = c.bdry.ni1[d][s]; region_out[f * 6 + 1] = c.bdry.ni2[d][s];
    region_out[f * 6 + 2] = c.bdry.nj1[d][s]; region_out[f * 6 + 3] = c.bdry.nj2[d][s];
    region_out[f * 6 + 4] = c.bdry.nk1[d][s]; region_out[f * 6 + 5] = c.bdry.nk2[d][s];
    if (!c.bdry.is_sides_pml[d][s]) continue;
    int np = c.bdry.num_of_layers[d][s] + 1;
    for (int n = 0; n < np && n < maxlen; n++) {
      A_out[f * maxlen + n] = c.bdry.A[d][s][n];
      B_out[f * maxlen + n] = c.bdry.B[d][s][n];
      D_out[f * maxlen + n] = c.bdry.D[d][s][n];
    }
  }
  teardown(&c);
  free(metric); free(med);
  return 0;
}

static void ref_dvh2dvz(refh_ctx *c)
{
  int mt = c->md.medium_type;
  if (mt == CONST_MEDIUM_ELASTIC_VTI)
    sv_curv_col_el_vti_dvh2dvz(&c->gd, &c->metric, &c->md, &c->bdry, 0);
  else if (mt == CONST_MEDIUM_ELASTIC_ANISO)
    sv_curv_col_el_aniso_dvh2dvz(&c->gd, &c->metric, &c->md, &c->bdry, 0);
  else if (mt == CONST_MEDIUM_VISCOELASTIC_ISO)
    sv_curv_col_vis_iso_dvh2dvz(&c->gd, &c->metric, &c->md, &c->bdry, c->fd.fdc_len,
                                c->fd.fdc_indx, c->fd.fdc_coef, 0);
  else
    sv_curv_col_el_iso_dvh2dvz(&c->gd, &c->metric, &c->md, &c->bdry, 0);
}

/* ---- free-surface matrices (sv_curv_col_el_{iso,vti,aniso}_dvh2dvz) ----- */
/* mats_out: 3 arrays of nx*ny*9 (matVx2Vz2, matVy2Vz2, matF2Vz2) */
int refh_dvh2dvz(const refh_cfg *cfg, float *metric10, float *med3, float *mats_out)
{
  refh_ctx c;
  refh_cfg nopml = *cfg;   /* no coordinates here: the PML set-up needs them */
  memset(nopml.pml_on, 0, sizeof nopml.pml_on);
  setup(&c, &nopml, NULL, metric10, med3);
  ref_dvh2dvz(&c);
  size_t n = c.gd.siz_iz * 9;
  memcpy(mats_out, c.bdry.matVx2Vz2, n * sizeof(float));
  memcpy(mats_out + n, c.bdry.matVy2Vz2, n * sizeof(float));
  memcpy(mats_out + 2 * n, c.bdry.matF2Vz2, n * sizeof(float));
  teardown(&c);
  return 0;
}

/* visco-elastic matrices (sv_curv_col_vis_iso_dvh2dvz): mats_out holds 4
 * arrays of nx*ny*9, matVx2Vz2, matVy2Vz2, matF2Vz2, matD */
int refh_dvh2dvz_vis(const refh_cfg *cfg, float *coords, float *metric10, float *med,
                     float *mats_out)
{
  refh_ctx c;
  refh_cfg nopml = *cfg;
  memset(nopml.pml_on, 0, sizeof nopml.pml_on);
  setup(&c, &nopml, coords, metric10, med);
  ref_dvh2dvz(&c);
  size_t n = c.gd.siz_iz * 9;
  memcpy(mats_out, c.bdry.matVx2Vz2, n * sizeof(float));
  memcpy(mats_out + n, c.bdry.matVy2Vz2, n * sizeof(float));
  memcpy(mats_out + 2 * n, c.bdry.matF2Vz2, n * sizeof(float));
  memcpy(mats_out + 3 * n, c.bdry.matD, n * sizeof(float));
  teardown(&c);
  return 0;
}

static size_t aux_total(refh_ctx *c)
{
  size_t t = 0;
  for (int f = 0; f < 6; f++)
    if (c->bdry.is_sides_pml[f / 2][f % 2]) t += 9 * c->bdry.auxvar[f / 2][f % 2].siz_icmp;
  return t;
}

/* ---- one RHS evaluation (sv_curv_col_el_*_onestage) --------------------- */
/* aux_cur / aux_rhs_out: the enabled PML faces' 9-component aux arrays, face
 * after face in x1,x2,y1,y2,z1,z2 order.  Returns the aux length in floats. */
long refh_onestage(const refh_cfg *cfg, float *coords, float *metric10, float *med3,
                   int it, int ipair, int istage, int patch,
                   const float *w_cur, const float *aux_cur,
                   float *rhs_out, float *aux_rhs_out)
{
  refh_ctx c;
  setup(&c, cfg, coords, metric10, med3);
  if (cfg->free_top) ref_dvh2dvz(&c);
  size_t L = c.wav.siz_ilevel;
  float *wcur = c.wav.v5d;          /* level 0 */
  float *rhs = c.wav.v5d + 2 * L;   /* level 2, zero */
  memcpy(wcur, w_cur, L * sizeof(float));
  size_t off = 0;
  for (int f = 0; f < 6; f++) {
    bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
    if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
    a->cur = a->var;
    a->rhs = a->var + 2 * a->siz_ilevel;
    /* the 9 elastic components; a visco-elastic wavefield's memory variables
     * have aux slots too (bdry_t.c:300), which stay zero */
    if (aux_cur) memcpy(a->cur, aux_cur + off, 9 * a->siz_icmp * sizeof(float));
    off += 9 * a->siz_icmp;
  }
  c.src.it = it;
  c.src.istage = istage;
  /* one read holds every step: row it, none from dd_max_nt on */
  if (cfg->ndd > 0) {
    c.src.dd_it_here = it;
    c.src.dd_is_valid = it < cfg->dd_max_nt;
  }
  load_surface_slices(&c, 0, 0);
  harness_onestage(&c, wcur, rhs, ipair, istage, patch);
  memcpy(rhs_out, rhs, L * sizeof(float));
  off = 0;
  for (int f = 0; f < 6; f++) {
    bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
    if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
    if (aux_rhs_out) memcpy(aux_rhs_out + off, a->rhs, 9 * a->siz_icmp * sizeof(float));
    off += 9 * a->siz_icmp;
  }
  teardown(&c);
  return (long)off;
}

/* ---- the time loop around the reference RHS ----------------------------- */
/* recv_idx: 8 corner 1-D indices per receiver (io_funcs.c:382-389 order);
 * recv_w: di,dj,dk per receiver.  seismo_out: [nrecv][9][nt]. */
int refh_run(const refh_cfg *cfg, float *coords, float *metric10, float *med3,
              float dt, int nt, const float *w_init,
              int nrecv, const int *recv_idx, const float *recv_w,
              float *w_final, float *seismo_out, float *pg_out, float *dis_out)
{
  refh_ctx c;
  setup(&c, cfg, coords, metric10, med3);
  if (cfg->free_top) ref_dvh2dvz(&c);
  size_t L = c.wav.siz_ilevel, V = c.wav.siz_icmp;
  float *w_pre = c.wav.v5d, *w_tmp = c.wav.v5d + L, *w_rhs = c.wav.v5d + 2 * L,
        *w_end = c.wav.v5d + 3 * L, *w_cur;
  if (w_init) memcpy(w_pre, w_init, L * sizeof(float));
  float *pg = NULL, *dis = NULL;
  if (cfg->free_top && pg_out) {
    pg = (float *)calloc(15 * c.gd.siz_iz, sizeof(float));
    dis = (float *)calloc(3 * c.gd.siz_iz, sizeof(float));
  }
  for (int f = 0; f < 6; f++) {
    bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
    if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
    a->pre = a->var; a->tmp = a->var + a->siz_ilevel;
    a->rhs = a->var + 2 * a->siz_ilevel; a->end = a->var + 3 * a->siz_ilevel;
  }
  for (int it = 0; it < nt; it++) {
    int ipair = it % c.fd.num_of_pairs;
    /* the step bookkeeping of src_dd_accit_loadstf (src_t.c:1929-1952, in a
     * file that needs netCDF) for one read of every step: row it, none from
     * dd_max_nt on */
    if (cfg->ndd > 0) {
      c.src.dd_it_here = it;
      c.src.dd_is_valid = it < cfg->dd_max_nt;
    }
    for (int istage = 0; istage < 4; istage++) {
      w_cur = istage == 0 ? w_pre : w_tmp;
      for (int f = 0; f < 6; f++) {
        bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
        if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
        a->cur = istage == 0 ? a->pre : a->tmp;
      }
      c.src.it = it; c.src.istage = istage;
      load_surface_slices(&c, it, istage);
      harness_onestage(&c, w_cur, w_rhs, ipair, istage, 1);
      float ca = istage < 3 ? c.fd.rk_a[istage] * dt : 0.0f;
      float cb = c.fd.rk_b[istage] * dt;
      if (istage < 3)
        for (size_t p = 0; p < L; p++) w_tmp[p] = w_pre[p] + ca * w_rhs[p];
      if (istage == 0)
        for (size_t p = 0; p < L; p++) w_end[p] = w_pre[p] + cb * w_rhs[p];
      else
        for (size_t p = 0; p < L; p++) w_end[p] += cb * w_rhs[p];
      for (int f = 0; f < 6; f++) {
        bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
        if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
        size_t n = a->siz_ilevel;
        if (istage < 3)
          for (size_t p = 0; p < n; p++) a->tmp[p] = a->pre[p] + ca * a->rhs[p];
        if (istage == 0)
          for (size_t p = 0; p < n; p++) a->end[p] = a->pre[p] + cb * a->rhs[p];
        else
          for (size_t p = 0; p < n; p++) a->end[p] += cb * a->rhs[p];
      }
      /* drv_rk_curv_col.c:440-445 */
      if (istage == 3 && c.md.medium_type == CONST_MEDIUM_VISCOELASTIC_ISO && cfg->free_top)
        sv_curv_col_vis_iso_free(w_end, &c.wav, &c.gd, &c.metric, &c.md, &c.bdry, 0, 0);
    }
    for (int r = 0; r < nrecv; r++) {
      const int *ix = recv_idx + 8 * r;
      float Lx2 = recv_w[3 * r], Ly2 = recv_w[3 * r + 1], Lz2 = recv_w[3 * r + 2];
      float Lx1 = 1.0 - Lx2, Ly1 = 1.0 - Ly2, Lz1 = 1.0 - Lz2;
      for (int icmp = 0; icmp < c.wav.ncmp; icmp++) {
        float *var = w_end + icmp * V;
        /* same expression as the reference receiver sampler, including the
         * repeated indx1d[5] term and the absent indx1d[4] term */
        float v = var[ix[0]] * Lx1 * Ly1 * Lz1
                + var[ix[1]] * Lx2 * Ly1 * Lz1
                + var[ix[2]] * Lx1 * Ly2 * Lz1
                + var[ix[3]] * Lx2 * Ly2 * Lz1
                + var[ix[5]] * Lx1 * Ly1 * Lz2
                + var[ix[5]] * Lx2 * Ly1 * Lz2
                + var[ix[6]] * Lx1 * Ly2 * Lz2
                + var[ix[7]] * Lx2 * Ly2 * Lz2;
        seismo_out[((size_t)r * c.wav.ncmp + icmp) * nt + it] = v;
      }
    }
    /* peak ground motion from w_end and w_pre before the swap
     * (drv_rk_curv_col.c:492-495), the reference's own PG_calcu */
    if (cfg->free_top && pg) PG_calcu(w_end, w_pre, &c.gd, pg, dis, dt);
    wav_zero_edge(&c.gd, &c.wav, w_rhs);
    float *t = w_pre; w_pre = w_end; w_end = t;
    for (int f = 0; f < 6; f++) {
      bdrypml_auxvar_t *a = &c.bdry.auxvar[f / 2][f % 2];
      if (!c.bdry.is_sides_pml[f / 2][f % 2]) continue;
      float *q = a->pre; a->pre = a->end; a->end = q;
    }
  }
  if (w_final) memcpy(w_final, w_pre, L * sizeof(float));
  if (pg) {
    memcpy(pg_out, pg, 15 * c.gd.siz_iz * sizeof(float));
    if (dis_out) memcpy(dis_out, dis, 3 * c.gd.siz_iz * sizeof(float));
    free(pg); free(dis);
  }
  teardown(&c);
  return 0;
}

/* blk_dt_esti_curv (blk_t.c:2840-2930) itself on filled gd_t / md_t structs.
 * vols: lambda, mu, rho (iso; visco then Ylam_n, Ymu_n per body), c11, c33,
 * rho (VTI), c11, rho (aniso).  out3 = dtmax, Vp, L; ijk stays -1 when the
 * reference keeps no point. */
void refh_dt_esti(int ni, int nj, int nk, float *xyz, int medium_type, float *vols,
                  int nmaxwell, float cfl, float *out3, int *ijk)
{
  size_t vol = (size_t)(ni + 6) * (nj + 6) * (nk + 6);
  gd_t gd;
  fill_gd(&gd, ni, nj, nk, xyz, xyz + vol, xyz + 2 * vol);
  md_t md;
  memset(&md, 0, sizeof md);
  md.medium_type = medium_type;
  md.nmaxwell = nmaxwell;
  float *Ylam[16], *Ymu[16];
  if (medium_type == CONST_MEDIUM_ELASTIC_VTI) {
    md.c11 = vols; md.c33 = vols + vol; md.rho = vols + 2 * vol;
  } else if (medium_type == CONST_MEDIUM_ELASTIC_ANISO) {
    md.c11 = vols; md.rho = vols + vol;
  } else {
    md.lambda = vols; md.mu = vols + vol; md.rho = vols + 2 * vol;
    for (int n = 0; n < nmaxwell && n < 16; n++) {
      Ylam[n] = vols + (3 + 2 * (size_t)n) * vol;
      Ymu[n] = vols + (4 + 2 * (size_t)n) * vol;
    }
    md.Ylam = Ylam; md.Ymu = Ymu;
  }
  out3[0] = out3[1] = out3[2] = 0.0f;
  ijk[0] = ijk[1] = ijk[2] = -1;
  blk_dt_esti_curv(&gd, &md, cfl, &out3[0], &out3[1], &out3[2], &ijk[0], &ijk[1], &ijk[2]);
}

/* bdry_ablexp_apply (bdry_t.c:840-890) itself on a filled bdry_t: masks
 * Ex[nx], Ey[ny], Ez[nz], blocks as {enable, ni1, ni2, nj1, nj2, nk1, nk2}. */
void refh_ablexp_apply(int nx, int ny, int nz, int ncmp, float *w, float *Ex, float *Ey,
                       float *Ez, const int *blocks)
{
  bdry_t bdry;
  memset(&bdry, 0, sizeof bdry);
  bdry.nx = nx; bdry.ny = ny; bdry.nz = nz;
  bdry.ablexp_Ex = Ex; bdry.ablexp_Ey = Ey; bdry.ablexp_Ez = Ez;
  for (int n = 0; n < CONST_NDIM_2; n++) {
    const int *b = blocks + 7 * n;
    bdry_block_t *D = bdry.bdry_blk + n;
    D->enable = b[0];
    D->ni1 = b[1]; D->ni2 = b[2]; D->ni = b[2] - b[1] + 1;
    D->nj1 = b[3]; D->nj2 = b[4]; D->nj = b[4] - b[3] + 1;
    D->nk1 = b[5]; D->nk2 = b[6]; D->nk = b[6] - b[5] + 1;
  }
  bdry_ablexp_apply(&bdry, w, ncmp, (size_t)nx * ny * nz);
}

/* sv_curv_col_el_graves_Qs (sv_curv_col_el.c:638-667) itself on filled gd_t /
 * md_t structs. */
void refh_graves_qs(int ni, int nj, int nk, int ncmp, float *w, float *Qs, float freq, float dt)
{
  gd_t gd;
  fill_gd(&gd, ni, nj, nk, NULL, NULL, NULL);
  md_t md;
  memset(&md, 0, sizeof md);
  md.Qs = Qs;
  md.visco_Qs_freq = freq;
  sv_curv_col_el_graves_Qs(w, ncmp, dt, &gd, &md);
}
