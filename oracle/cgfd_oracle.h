/*
 * cgfd_oracle.h -- CPU restatement of CGFD3D's isotropic elastic
 * time-stepping hot path.  TEST INFRASTRUCTURE ONLY: the product never links
 * this; tests/, the smoke check and the CPU leg of the benchmark call it as
 * the checker.
 *
 * Built twice by oracle/Makefile: liboracle.so with real_t = float (the
 * reference's arithmetic, expression by expression, no FMA contraction) and
 * liboracle64.so with real_t = double (the same scheme in FP64, used to state
 * every FP32 result's error).
 *
 * Array layout is the reference's (gd_t.c:2890-2897, wav_t.c:93-136):
 * x fastest, nx = ni+6, ny = nj+6, nz = nk+6, iptr = i + j*nx + k*nx*ny;
 * wavefield components Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy, one volume each;
 * metric jac,xi_x,xi_y,xi_z,eta_x,eta_y,eta_z,zeta_x,zeta_y,zeta_z;
 * medium volumes (md_t.c:28-260, density replaced by slowness 1/rho as
 * md_rho_to_slow does, md_t.c:602-623):
 *   isotropic (2):  lambda, mu, slw
 *   VTI (3):        c11, c13, c33, c55, c66, slw
 *   anisotropic (4): c11,c12,c13,c14,c15,c16,c22,c23,c24,c25,c26,c33,c34,c35,
 *                    c36,c44,c45,c46,c55,c56,c66, slw
 *   visco-elastic GMB (5): lambda, mu (unrelaxed), slw, then Ylam_n, Ymu_n for
 *                    each of the nmaxwell bodies (md_t.c:222-249); the
 *                    wavefield then carries 6 memory variables per body after
 *                    the 9 fields, Jxx,Jyy,Jzz,Jyz,Jxz,Jxy (wav_t.c:138-171)
 */
#ifndef CGFD_ORACLE_H
#define CGFD_ORACLE_H

#include <stddef.h>

#ifndef ORC_REAL
#define ORC_REAL float
#endif
typedef ORC_REAL real_t;

#define ORC_NGHOST 3

/* medium types, constants.h:22-26 */
#define ORC_MEDIUM_ISO   2
#define ORC_MEDIUM_VTI   3
#define ORC_MEDIUM_ANISO 4
#define ORC_MEDIUM_VISCO 5
#define ORC_MAXWELL_MAX 8

/* face order matches the reference's idim*2+iside: x1,x2,y1,y2,z1,z2 */
typedef struct {
  int ni, nj, nk;
  int free_top;            /* traction-image free surface on z2 */
  int pml_on[6];
  int pml_layers[6];
  const float *pml_A[6];   /* per face, num_layers+1 values (bdry_t.c:246-269) */
  const float *pml_B[6];
  const float *pml_D[6];
  /* moment sources (sv_curv_col_el.c:312-479) */
  int nsrc;
  int src_gauss;           /* 0: point, 1: Gaussian 7^3 smoothing */
  int src_ext_half;        /* 3 for the Gaussian */
  float src_ext_coef;      /* 1.5 (= ext_half/2, src_t.c:41-46) */
  const int *src_ijk;      /* local indices with ghosts, 3 per source */
  const float *src_inc;    /* si_inc,sj_inc,sk_inc, 3 per source */
  const int *src_it_begin;
  const int *src_it_end;
  int src_max_nt;
  const float *src_mij;    /* Mxx,Myy,Mzz,Mxz,Myz,Mxy, each [nsrc][max_nt][4] */
  int medium_type;         /* ORC_MEDIUM_*; 0 is read as isotropic */
  /* visco-elastic only */
  int nmaxwell;            /* memory-variable bodies */
  const float *visco_wl;   /* relaxation frequencies, nmaxwell values */
  const float *xyz;        /* coordinates (3 volumes), for the surface matD */
  /* force sources (sv_curv_col_el.c:356-364, 440-452), not the strict surface
   * force: Fx,Fy,Fz, each [nsrc][max_nt][4]; NULL for none.  src_mij may be
   * NULL for force-only sources. */
  const float *src_fxyz;
  /* strict surface force (par_t.c:821-825, src_t.c:154-314, 1020-1170):
   * with src_force_strict the force terms at k == nk2 are left out of the
   * RHS (sv_curv_col_el.c:388, 447), and the nsurf sources listed in
   * src_surf_idx (src->force_rate_indx: the force sources whose extent
   * reaches the surface) set the surface traction TxSrc..TzSrc of the
   * traction image and the velocity source VxSrc..VzSrc of vlow every stage;
   * src_frate holds Fx_rate, Fy_rate, Fz_rate, each [nsurf][max_nt][4] */
  int src_force_strict;
  int nsurf;
  const int *src_surf_idx;
  const float *src_frate;
  /* distributed (dd) sources added at grid points (sv_curv_col_el_rhs_srcdd,
   * sv_curv_col_el.c:486-555; called by the isotropic and visco-elastic
   * RHS only, sv_curv_col_el_iso.c:194-199, sv_curv_col_vis_iso.c:236-241):
   * dd_indx the points' array indices; dd_vi [dd_max_nt][4][ndd][3] and
   * dd_mij [dd_max_nt][4][ndd][6] (Mxx,Myy,Mzz,Myz,Mxz,Mxy) the values of
   * step it and stage (NULL for none); nothing is added from step
   * dd_max_nt on (src_dd_accit_loadstf, src_t.c:1929-1944) */
  int ndd;
  int dd_max_nt;
  const long *dd_indx;
  const float *dd_vi;
  const float *dd_mij;
} orc_cfg;

typedef struct orc_s orc_t;

/* ---- scheme tables (fd_t.c:18-304) ---- */
/* slot 0=x, 1=y, 2..5 = z near-surface layers 0..3; 5 entries per slot */
int orc_fd_ops(int *indx, float *coef, int *len,
               float *rk_a, float *rk_b, float *rk_t);

/* ---- setup pieces ---- */
void orc_pml_profile(int num_layers, float L0, float dh, float vel,
                     float amax, float bmax, int iside,
                     float *A, float *B, float *D);
float orc_pml_avg_dh(int ni, int nj, int nk, const float *x3d, const float *y3d,
                     const float *z3d, const int region[6], int idim, float *L0);
void orc_pml_region(int ni, int nj, int nk, const int pml_layers[6],
                    const int pml_on[6], int region[6][6]);
void orc_dvh2dvz_iso(int ni, int nj, int nk, const float *metric10,
                     const float *med3, float *mats /* 3 x nx*ny*9 */);
/* VTI and anisotropic matrices (sv_curv_col_el_vti.c:1085-1203,
 * sv_curv_col_el_aniso.c:1267-1417); the third matrix is left untouched */
void orc_dvh2dvz(int medium_type, int ni, int nj, int nk, const float *metric10,
                 const float *med, float *mats);
int orc_medium_ncmp(int medium_type);
/* visco-elastic free-surface matrices, sv_curv_col_vis_iso_dvh2dvz
 * (sv_curv_col_vis_iso.c:354-506): matVx2Vz2, matVy2Vz2 as the isotropic
 * ones, matD (the 4th [ny][nx][3][3] array) from the eta derivatives of the
 * coordinates with the centred 7-point operator; matF2Vz2 left untouched */
void orc_dvh2dvz_vis(int ni, int nj, int nk, const float *metric10,
                     const float *med, const float *xyz, float *mats4);
/* GMB memory-variable coefficients, md_vis_GMB_cal_Y (md_t.c:643-792) with
 * md_visco_LS (:795-852): from relaxed lambda, mu and Qp, Qs at npts points,
 * in place to unrelaxed lambda, mu, plus Ylam/Ymu [npts][nmaxwell] and wl */
void orc_visco_gmb(long npts, int nmaxwell, float freq, float fmin, float fmax,
                   float *lambda, float *mu, const float *Qp, const float *Qs,
                   float *Ylam, float *Ymu, float *wl);
void orc_metric(int ni, int nj, int nk, const float *xyz /* 3 volumes */,
                float *metric10);
void orc_graves_qs(int ni, int nj, int nk, int ncmp, float *w, const float *Qs, float freq,
                   float dt);
void orc_ablexp_apply(int nx, int ny, int nz, int ncmp, float *w, const float *Ex,
                      const float *Ey, const float *Ez, const int *blocks);
float orc_dist_point2plane(const float x0[3], const float x1[3], const float x2[3],
                           const float x3[3]);
void orc_dt_estimate(int ni, int nj, int nk, const float *xyz, int medium_type,
                     const float *vols, int nmaxwell, float cfl, float *dtmax,
                     float *vpmax, float *lmin, int *ijk);
void orc_gauss_weights(float *w, float x0, float y0, float z0, float a,
                       int half, int half_below);
float orc_ricker(float t, float fc, float t0);
void orc_angle2moment(float strike, float dip, float rake, float m[6]);
/* the derivative of fun_ricker, fun_ricker_deriv (src_t.c:2168-2176) */
float orc_ricker_deriv(float t, float fc, float t0);

/* ---- solver ---- */
orc_t *orc_create(const orc_cfg *cfg, const float *metric10, const float *med);
void   orc_destroy(orc_t *o);
size_t orc_aux_len(const orc_t *o);           /* floats of all PML aux, 1 level */
size_t orc_vol(const orc_t *o);
int    orc_ncomp(const orc_t *o);             /* wavefield components */

/* one RHS evaluation: rhs must be zero outside the physical region */
void orc_onestage(orc_t *o, int it, int ipair, int istage,
                  const real_t *w_cur, const real_t *aux_cur,
                  real_t *rhs, real_t *aux_rhs);

/* nt RK4 steps from level w (in/out) and aux (in/out, may be NULL if no PML);
 * seismo[(r*9+icmp)*nt_total + it0+n] sampled after each step when nrecv>0. */
void orc_get_pg(const orc_t *o, real_t *pg, real_t *dis);
/* the surface traction and velocity-source slices of step it, stage istage,
 * src_set_surface_layer_for_force (src_t.c:154-314): ts = TxSrc,TySrc,TzSrc
 * and vs = VxSrc,VySrc,VzSrc, [3][ny][nx] each (zero without surface forces) */
void orc_surface_force_layer(orc_t *o, int it, int istage, real_t *ts, real_t *vs);
void orc_run(orc_t *o, float dt, int it0, int nt, real_t *w, real_t *aux,
             int nrecv, const int *recv_idx, const float *recv_w,
             int nt_total, real_t *seismo);

#endif
