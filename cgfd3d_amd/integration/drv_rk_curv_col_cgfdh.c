/*
 * drv_rk_curv_col_cgfdh.c -- drop-in replacement of drv_rk_curv_col_allstep
 * (forward/drv_rk_curv_col.h:17-37) for the reference build: the same C
 * signature, the reference's structs unpacked into the engine's C ABI
 * (include/cgfdh.h).  Compile it into the reference's main program in place
 * of forward/drv_rk_curv_col.c and link cgfd3d_amd/libcgfdh.so.
 *
 * Covered: isotropic, VTI and general anisotropic elastic media and the GMB
 * visco-elastic medium, CFS-PML faces, traction-image free surface, point
 * and Gaussian moment and force sources (the strict surface force
 * included), dd sources at points (their value chunks streamed to the
 * device as src_dd_accit_loadstf reads them), receivers and receiver lines,
 * slices, snapshots (velocity, stress, strain), the peak-ground-motion maps,
 * the ablexp sponge and Graves-Qs attenuation with any elastic medium, and
 * the NaN check every qc_check_nan_num_of_step steps (wav_check_value; the
 * reference's driver has that call commented out, the shim makes it).
 * Slices and snapshots are packed on the GPU and written by the library's
 * own netCDF writer (include/cgfdh_io.h) in the reference's file layouts;
 * the output_all debug dump copies the whole wavefield back each step.
 * Anything else the par file asks for (acoustic medium, smoothed dd
 * sources, several MPI ranks) stops the run with a message, the reference's
 * fatal-error convention.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>

#include "drv_rk_curv_col.h"
#include "cgfdh.h"
#include "cgfdh_io.h"

static void cgfdh_fatal(const char *what, int rc)
{
  fprintf(stderr, "ERROR: %s failed (%d): %s\n", what, rc, cgfdh_last_error());
  MPI_Abort(MPI_COMM_WORLD, 1);
}

#define CHK(call) do { int rc_ = (call); if (rc_) cgfdh_fatal(#call, rc_); } while (0)

static void cgfdh_io_fatal(const char *what)
{
  fprintf(stderr, "ERROR: %s failed: %s\n", what, cgfdh_io_last_error());
  MPI_Abort(MPI_COMM_WORLD, 1);
}

#define CHKIO(call) do { if ((call) != 0) cgfdh_io_fatal(#call); } while (0)

/* the condition under which io_snap_nc_put writes snapshot n at step it
 * (io_funcs.c:1148-1154) */
static int snap_due(const iosnap_t *sn, int n, int it, int nt_total)
{
  int it1 = sn->it1[n], dit = sn->dit[n];
  return it >= it1 && (it - it1) % dit == 0 && (it - it1) / dit <= (nt_total - it1) / dit;
}

void
drv_rk_curv_col_allstep(
  fd_t            *fd,
  gd_t        *gd,
  gdcurv_metric_t *metric,
  md_t      *md,
  src_t     *src,
  bdry_t    *bdry,
  wav_t     *wav,
  mympi_t    *mympi,
  iorecv_t   *iorecv,
  ioline_t   *ioline,
  ioslice_t  *ioslice,
  iosnap_t   *iosnap,
  float dt, int nt_total, float t0,
  char *output_fname_part,
  char *output_dir,
  int qc_check_nan_num_of_step,
  const int output_all,
  const int verbose)
{
  for (int n = 0; n < 4; n++)
    if (mympi->neighid[n] != MPI_PROC_NULL) {
      fprintf(stderr, "ERROR: the GPU time loop runs one block (1 MPI rank)\n");
      MPI_Abort(MPI_COMM_WORLD, 1);
    }
  int mt = md->medium_type;
  int visco = mt == CONST_MEDIUM_VISCOELASTIC_ISO;
  /* the isotropic and GMB right-hand sides add dd sources
   * (sv_curv_col_el_iso.c:193-199, sv_curv_col_vis_iso.c:235-241), at points
   * only (sv_curv_col_el.c:572-577 stops on smoothed ones) */
  int dd_on = src->dd_is_valid == 1 && (mt == CONST_MEDIUM_ELASTIC_ISO || visco);
  if (dd_on && src->dd_is_add_at_point != 1) {
    fprintf(stderr, "ERROR: ddsrc with smoooth not implemented\n");
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  if ((mt != CONST_MEDIUM_ELASTIC_ISO && mt != CONST_MEDIUM_ELASTIC_VTI &&
       mt != CONST_MEDIUM_ELASTIC_ANISO && !visco) ||
      (visco && md->visco_type != CONST_VISCO_GMB)) {
    fprintf(stderr, "ERROR: par file asks for a feature the GPU time loop lacks\n");
    MPI_Abort(MPI_COMM_WORLD, 1);
  }
  for (int n = 0; n < iosnap->num_of_snap; n++)
    if (iosnap->dit[n] <= 0) {   /* io_snap_nc_put divides by it (io_funcs.c:1148-1150) */
      fprintf(stderr, "ERROR: snapshot %d has time increment %d\n", n, iosnap->dit[n]);
      MPI_Abort(MPI_COMM_WORLD, 1);
    }

  cgfdh_desc d;
  memset(&d, 0, sizeof d);
  d.ni = gd->ni; d.nj = gd->nj; d.nk = gd->nk;
  d.medium_type = mt;              /* the engine's codes are the reference's (constants.h) */
  d.nmaxwell = visco ? md->nmaxwell : 0;
  d.free_top = bdry->is_sides_free[2][1];
  for (int f = 0; f < 6; f++) {
    d.pml_on[f] = bdry->is_sides_pml[f / 2][f % 2];
    d.pml_layers[f] = d.pml_on[f] ? bdry->num_of_layers[f / 2][f % 2] : 0;
  }
  d.dt = dt;
  d.nt_total = nt_total;

  cgfdh_solver *s = NULL;
  CHK(cgfdh_create(&d, &s));
  size_t vol = (size_t)gd->nx * gd->ny * gd->nz;
  if (visco && d.free_top)         /* device metric build also forms matD from the coords */
    CHK(cgfdh_compute_metric(s, gd->x3d));
  CHK(cgfdh_set_metric(s, metric->v4d));   /* the reference's own metric */
  /* md->rho holds 1/rho here (md_rho_to_slow, main_curv_col_el_3d.c:843) */
  if (mt == CONST_MEDIUM_ELASTIC_ISO) {
    CHK(cgfdh_set_medium_iso(s, md->lambda, md->mu, md->rho));
  } else if (mt == CONST_MEDIUM_ELASTIC_VTI) {
    CHK(cgfdh_set_medium_vti(s, md->c11, md->c13, md->c33, md->c55, md->c66, md->rho));
  } else if (mt == CONST_MEDIUM_ELASTIC_ANISO) {
    const float *c[21] = {md->c11, md->c12, md->c13, md->c14, md->c15, md->c16, md->c22,
                          md->c23, md->c24, md->c25, md->c26, md->c33, md->c34, md->c35,
                          md->c36, md->c44, md->c45, md->c46, md->c55, md->c56, md->c66};
    float *cij = malloc(21 * vol * sizeof(float));
    for (int n = 0; n < 21; n++) memcpy(cij + n * vol, c[n], vol * sizeof(float));
    CHK(cgfdh_set_medium_aniso(s, cij, md->rho));
    free(cij);
  } else {
    float *y = malloc(2 * (size_t)md->nmaxwell * vol * sizeof(float));
    for (int n = 0; n < md->nmaxwell; n++) {
      memcpy(y + (2 * n) * vol, md->Ylam[n], vol * sizeof(float));
      memcpy(y + (2 * n + 1) * vol, md->Ymu[n], vol * sizeof(float));
    }
    CHK(cgfdh_set_medium_visco(s, md->lambda, md->mu, md->rho, y, md->wl));
    free(y);
  }
  for (int f = 0; f < 6; f++)
    if (d.pml_on[f])
      CHK(cgfdh_set_pml_face(s, f, bdry->A[f / 2][f % 2], bdry->B[f / 2][f % 2],
                             bdry->D[f / 2][f % 2]));
  if (d.free_top) CHK(cgfdh_compute_free_surface(s));   /* drv_rk_curv_col.c:132-159 */

  int ns = src->total_number;
  if (ns > 0 && (src->moment_actived || src->force_actived)) {
    size_t per = (size_t)ns * src->max_nt * src->max_stage;
    int gauss = src->itype_spatial_ext != CONST_SRC_SPATIAL_POINT;
    int half = src->ext_half_npoint, len = 2 * half + 1;
    int *ijk = malloc(3 * ns * sizeof(int));
    float *mij = src->moment_actived ? malloc(6 * per * sizeof(float)) : NULL;
    float *fxyz = src->force_actived ? malloc(3 * per * sizeof(float)) : NULL;
    float *w = gauss ? malloc((size_t)ns * len * len * len * sizeof(float)) : NULL;
    if (mij) {
      memcpy(mij + 0 * per, src->Mxx, per * sizeof(float));
      memcpy(mij + 1 * per, src->Myy, per * sizeof(float));
      memcpy(mij + 2 * per, src->Mzz, per * sizeof(float));
      memcpy(mij + 3 * per, src->Mxz, per * sizeof(float));
      memcpy(mij + 4 * per, src->Myz, per * sizeof(float));
      memcpy(mij + 5 * per, src->Mxy, per * sizeof(float));
    }
    if (fxyz) {
      memcpy(fxyz + 0 * per, src->Fx, per * sizeof(float));
      memcpy(fxyz + 1 * per, src->Fy, per * sizeof(float));
      memcpy(fxyz + 2 * per, src->Fz, per * sizeof(float));
    }
    /* strict surface forces: indices and force-rate tables (src_t.c:76-95,
     * 1020-1028, 1080-1090); the engine forms TxSrc.. / VxSrc.. from them
     * before every stage as src_set_surface_layer_for_force does */
    int nsurf = src->force_actived ? src->total_number_surface_force : 0;
    float *frate = NULL;
    if (nsurf > 0) {
      size_t perr = (size_t)nsurf * src->max_nt * src->max_stage;
      frate = malloc(3 * perr * sizeof(float));
      memcpy(frate + 0 * perr, src->Fx_rate, perr * sizeof(float));
      memcpy(frate + 1 * perr, src->Fy_rate, perr * sizeof(float));
      memcpy(frate + 2 * perr, src->Fz_rate, perr * sizeof(float));
    }
    for (int is = 0; is < ns; is++) {
      ijk[3 * is] = src->si[is]; ijk[3 * is + 1] = src->sj[is]; ijk[3 * is + 2] = src->sk[is];
      if (gauss) {   /* the weights sv_curv_col_el_rhs_src computes per stage */
        int kb = src->sk[is] + half < gd->nk2 ? half : gd->nk2 - src->sk[is];
        CHK(cgfdh_gauss_weights(w + (size_t)is * len * len * len, src->si_inc[is],
                                src->sj_inc[is], src->sk_inc[is], src->ext_func_coef, half, kb));
      }
    }
    CHK(cgfdh_set_sources(s, ns, ijk, src->it_begin, src->it_end, src->max_nt, mij, fxyz,
                          gauss ? CGFDH_SRC_GAUSSIAN : CGFDH_SRC_POINT, half, w,
                          src->is_surface_force_strict, nsurf, src->force_rate_indx, frate));
    free(ijk); free(mij); free(fxyz); free(frate); free(w);
  }

  /* dd sources: the points and the chunk src_dd_read2local read first
   * (dd_nt_per_read steps, src_t.c:1838-1900) */
  const float *dd_vi = NULL, *dd_mij = NULL;
  if (dd_on) {
    dd_vi = src->dd_vi_actived ? src->dd_vi : NULL;
    dd_mij = src->dd_mij_actived ? src->dd_mij : NULL;
    CHK(cgfdh_set_dd_sources(s, src->dd_total_number, src->dd_indx, src->dd_max_nt,
                             src->dd_nt_per_read, dd_vi != NULL, dd_mij != NULL));
    CHK(cgfdh_put_dd_values(s, 0, src->dd_nt_per_read, dd_vi, dd_mij));
  }

  /* receivers, then the points of the receiver lines as receivers with zero
   * interpolation weights on their own grid point: the engine's trilinear
   * sum (io_funcs.c:1598-1612) then returns the point's value, which is what
   * io_line_keep records (io_funcs.c:1626-1645) */
  int nr = iorecv->total_number;
  int nl = 0;
  for (int n = 0; n < ioline->num_of_lines; n++) nl += ioline->line_nr[n];
  if (nr + nl > 0) {
    int *idx = malloc(8 * (size_t)(nr + nl) * sizeof(int));
    float *dijk = malloc(3 * (size_t)(nr + nl) * sizeof(float));
    for (int r = 0; r < nr; r++) {
      iorecv_one_t *o = iorecv->recvone + r;
      memcpy(idx + 8 * r, o->indx1d, 8 * sizeof(int));
      dijk[3 * r] = o->di; dijk[3 * r + 1] = o->dj; dijk[3 * r + 2] = o->dk;
    }
    int q = nr;
    for (int n = 0; n < ioline->num_of_lines; n++)
      for (int ir = 0; ir < ioline->line_nr[n]; ir++, q++) {
        for (int c = 0; c < 8; c++) idx[8 * q + c] = ioline->recv_iptr[n][ir];
        dijk[3 * q] = dijk[3 * q + 1] = dijk[3 * q + 2] = 0.0f;
      }
    CHK(cgfdh_set_receivers(s, nr + nl, idx, dijk));
    free(idx); free(dijk);
  }

  if (!visco && md->visco_type == CONST_VISCO_GRAVES_QS)   /* drv_rk_curv_col.c:414-416 */
    CHK(cgfdh_set_graves_qs(s, md->Qs, md->visco_Qs_freq));

  if (bdry->is_enable_ablexp) {   /* masks and blocks of bdry_ablexp_set (bdry_t.c:513-820) */
    int blk[6 * 7];
    for (int n = 0; n < 6; n++) {
      const bdry_block_t *D = bdry->bdry_blk + n;
      int *b = blk + 7 * n;
      b[0] = D->enable;
      b[1] = D->ni1; b[2] = D->ni2;
      b[3] = D->nj1; b[4] = D->nj2;
      b[5] = D->nk1; b[6] = D->nk2;
    }
    CHK(cgfdh_set_ablexp(s, bdry->ablexp_Ex, bdry->ablexp_Ey, bdry->ablexp_Ez, blk));
  }

  /* slices every step, snapshots at their own steps, the debug dump of
   * output_all (drv_rk_curv_col.c:503-525): the engine runs up to each output
   * step, packs the output points on the GPU, and the library's own netCDF
   * writer puts them into files laid out as io_slice_nc_create /
   * io_snap_nc_create / io_var3d_export_nc lay them out (cgfdh_io.h) */
  int *topoid = mympi->topoid;
  int nvs = visco ? 9 : wav->ncmp;   /* io_slice_nc_create: GMB slices hold 9 */
  const char *const *vnames = (const char *const *)wav->cmp_name;
  int nslice[3] = {ioslice->num_of_slice_x, ioslice->num_of_slice_y, ioslice->num_of_slice_z};
  int nsl = nslice[0] + nslice[1] + nslice[2];
  cgfdh_nc **slf = nsl ? calloc(nsl, sizeof(cgfdh_nc *)) : NULL;
  int q = 0;
  for (int n = 0; n < nslice[0]; n++, q++)
    CHKIO(cgfdh_io_slice_create(ioslice->slice_x_fname[n], 'x', gd->nk, gd->nj,
                                ioslice->slice_x_indx[n], topoid, nvs, vnames, &slf[q]));
  for (int n = 0; n < nslice[1]; n++, q++)
    CHKIO(cgfdh_io_slice_create(ioslice->slice_y_fname[n], 'y', gd->nk, gd->ni,
                                ioslice->slice_y_indx[n], topoid, nvs, vnames, &slf[q]));
  for (int n = 0; n < nslice[2]; n++, q++)
    CHKIO(cgfdh_io_slice_create(ioslice->slice_z_fname[n], 'z', gd->nj, gd->ni,
                                ioslice->slice_z_indx[n], topoid, nvs, vnames, &slf[q]));
  int nsn = iosnap->num_of_snap;
  cgfdh_nc **snf = nsn ? calloc(nsn, sizeof(cgfdh_nc *)) : NULL;
  int *snrec = nsn ? calloc(nsn, sizeof(int)) : NULL;
  for (int n = 0; n < nsn; n++) {
    if (iosnap->out_strain[n] && mt != CONST_MEDIUM_ELASTIC_ISO && !visco) {
      fprintf(stderr, "ERROR: strain snapshots need an isotropic medium\n");
      MPI_Abort(MPI_COMM_WORLD, 1);
    }
    const int g_start[3] = {iosnap->i1_to_glob[n], iosnap->j1_to_glob[n], iosnap->k1_to_glob[n]};
    const int l_start[3] = {iosnap->i1[n], iosnap->j1[n], iosnap->k1[n]};
    const int stride[3] = {iosnap->di[n], iosnap->dj[n], iosnap->dk[n]};
    CHKIO(cgfdh_io_snap_create(iosnap->fname[n], iosnap->nk[n], iosnap->nj[n], iosnap->ni[n],
                               iosnap->out_vel[n], iosnap->out_stress[n], iosnap->out_strain[n],
                               g_start, l_start, stride, topoid, &snf[n]));
  }
  /* output jobs (the slices, then the snapshots), each with two pinned host
   * buffers used in turn: a due step's boxes are packed behind its step and
   * copied out on the engine's copy stream while the next steps run, and
   * written to their files at the next due step (or the end), in order */
  int njob = nsl + nsn;
  size_t *jn = njob ? calloc(njob, sizeof(size_t)) : NULL;
  for (q = 0; q < nsl; q++) {
    int a = q < nslice[0] ? 0 : (q < nslice[0] + nslice[1] ? 1 : 2);
    size_t n1 = a == 0 ? gd->nk : (a == 1 ? gd->nk : gd->nj);
    size_t n2 = a == 0 ? gd->nj : gd->ni;
    jn[q] = (size_t)nvs * n1 * n2;
  }
  for (int n = 0; n < nsn; n++)
    jn[nsl + n] = 15 * (size_t)iosnap->ni[n] * iosnap->nj[n] * iosnap->nk[n];
  float **hb[2] = {njob ? calloc(njob, sizeof(float *)) : NULL,
                   njob ? calloc(njob, sizeof(float *)) : NULL};
  for (int h = 0; h < 2; h++)
    for (q = 0; q < njob; q++) CHK(cgfdh_host_alloc(jn[q] * sizeof(float), (void **)&hb[h][q]));
  typedef struct {
    int live, it;
    float t_end;
    int *tk;     /* two tickets per job: slice / snapshot velocities, stresses */
    int *rec;    /* per snapshot: its record, or -1 when not due */
  } out_set;
  out_set os[2];
  for (int h = 0; h < 2; h++) {
    os[h].live = 0;
    os[h].tk = calloc(2 * (size_t)(njob ? njob : 1), sizeof(int));
    os[h].rec = calloc(nsn ? nsn : 1, sizeof(int));
  }
  float *w4d = wav->v5d;
  int cmps[CGFDH_MAXWELL_MAX * 6 + 9];
  for (int c = 0; c < nvs; c++) cmps[c] = c;
  /* stresses in the snapshot file's order Txx Tyy Tzz Txz Tyz Txy (wavefield
   * slots 3 4 5 7 6 8) */
  const int stress_cmps[6] = {3, 4, 5, 7, 6, 8};
  int it_run = 0, cur = 0;
  for (int it = 0; it <= nt_total; it++) {
    int last = it == nt_total;
    int nan_due = 0, due = last;
    if (!last) {
      /* drv_rk_curv_col.c:179-181; after a new chunk is read the steps
       * before it run on the old one, then the new one goes to the device */
      if (src->dd_is_valid == 1) {
        src_dd_accit_loadstf(src, it, mympi->myid);
        if (dd_on && it > 0 && src->dd_is_valid == 1 && src->dd_it_here == 0) {
          if (it_run < it) CHK(cgfdh_run_async(s, it_run, it - it_run));
          it_run = it;
          CHK(cgfdh_put_dd_values(s, it, src->dd_nt_this_read, dd_vi, dd_mij));
        }
      }
      /* the value check of check_nan_every_nummber_of_steps
       * (drv_rk_curv_col.c:476-479), wav_check_value's message and exit */
      nan_due = qc_check_nan_num_of_step > 0 && it % qc_check_nan_num_of_step == 0;
      due = nsl > 0 || output_all || nan_due;
      for (int n = 0; n < nsn && !due; n++) due = snap_due(iosnap, n, it, nt_total);
    }
    if (!due) continue;
    if (it_run < (last ? nt_total : it + 1)) {
      CHK(cgfdh_run_async(s, it_run, (last ? nt_total : it + 1) - it_run));
      it_run = last ? nt_total : it + 1;
    }
    out_set *o = &os[cur];
    if (!last) {
      if (nan_due) {
        int nan_c;
        long nan_p;
        CHK(cgfdh_check_nan(s, &nan_c, &nan_p));
        if (nan_c >= 0) {
          fprintf(stderr, "ERROR: NaN occurs at iptr=%ld icmp=%d\n", nan_p, nan_c);
          fflush(stderr);
          exit(-1);
        }
      }
      /* this step's boxes, packed behind it (a set is reused two due steps
       * later, after its files were written) */
      o->live = 1;
      o->it = it;
      o->t_end = it * dt + t0 + dt;
      q = 0;
      for (int n = 0; n < nslice[0]; n++, q++) {
        int box[9] = {ioslice->slice_x_indx[n], 1, 1, gd->nj1, gd->nj, 1, gd->nk1, gd->nk, 1};
        CHK(cgfdh_subvolume_async(s, box, nvs, cmps, hb[cur][q], &o->tk[2 * q]));
      }
      for (int n = 0; n < nslice[1]; n++, q++) {
        int box[9] = {gd->ni1, gd->ni, 1, ioslice->slice_y_indx[n], 1, 1, gd->nk1, gd->nk, 1};
        CHK(cgfdh_subvolume_async(s, box, nvs, cmps, hb[cur][q], &o->tk[2 * q]));
      }
      for (int n = 0; n < nslice[2]; n++, q++) {
        int box[9] = {gd->ni1, gd->ni, 1, gd->nj1, gd->nj, 1, ioslice->slice_z_indx[n], 1, 1};
        CHK(cgfdh_subvolume_async(s, box, nvs, cmps, hb[cur][q], &o->tk[2 * q]));
      }
      for (int n = 0; n < nsn; n++) {
        int j = nsl + n;
        o->rec[n] = -1;
        o->tk[2 * j] = o->tk[2 * j + 1] = -1;
        if (!snap_due(iosnap, n, it, nt_total)) continue;
        int box[9] = {iosnap->i1[n], iosnap->ni[n], iosnap->di[n], iosnap->j1[n], iosnap->nj[n],
                      iosnap->dj[n], iosnap->k1[n], iosnap->nk[n], iosnap->dk[n]};
        size_t npt = (size_t)box[1] * box[4] * box[7];
        if (iosnap->out_vel[n])
          CHK(cgfdh_subvolume_async(s, box, 3, cmps, hb[cur][j], &o->tk[2 * j]));
        if (iosnap->out_stress[n])
          CHK(cgfdh_subvolume_async(s, box, 6, stress_cmps, hb[cur][j] + 3 * npt,
                                    &o->tk[2 * j + 1]));
        if (iosnap->out_strain[n]) CHK(cgfdh_get_strain_subvolume(s, box, hb[cur][j] + 9 * npt));
        o->rec[n] = snrec[n]++;
      }
      if (output_all) {   /* debug dump of the whole w_end (drv_rk_curv_col.c:513-525) */
        CHK(cgfdh_get_wavefield(s, w4d));
        CHKIO(cgfdh_io_var3d_export(output_dir, "w3d", topoid, it, w4d, wav->cmp_pos,
                                    vnames, wav->ncmp, (const char *const *)gd->index_name,
                                    gd->nx, gd->ny, gd->nz));
      }
    }
    /* the previous due step's files, while this step runs on the GPU; at the
     * end both sets, older first */
    for (int pass = 0; pass < (last ? 2 : 1); pass++) {
      int h = last ? (cur + 1 + pass) % 2 : 1 - cur;
      out_set *p = &os[h];
      if (!p->live) continue;
      for (q = 0; q < nsl; q++) {
        CHK(cgfdh_wait(s, p->tk[2 * q]));
        CHKIO(cgfdh_io_slice_put(slf[q], p->it, p->t_end, nvs, hb[h][q]));
      }
      for (int n = 0; n < nsn; n++) {
        int j = nsl + n;
        if (p->rec[n] < 0) continue;
        size_t npt = (size_t)iosnap->ni[n] * iosnap->nj[n] * iosnap->nk[n];
        if (p->tk[2 * j] >= 0) CHK(cgfdh_wait(s, p->tk[2 * j]));
        if (p->tk[2 * j + 1] >= 0) CHK(cgfdh_wait(s, p->tk[2 * j + 1]));
        CHKIO(cgfdh_io_snap_put(snf[n], p->rec[n], p->t_end,
                                iosnap->out_vel[n] ? hb[h][j] : NULL,
                                iosnap->out_stress[n] ? hb[h][j] + 3 * npt : NULL,
                                iosnap->out_strain[n] ? hb[h][j] + 9 * npt : NULL));
      }
      p->live = 0;
    }
    cur = 1 - cur;
  }
  CHK(cgfdh_sync(s));
  for (int h = 0; h < 2; h++) {
    for (q = 0; q < njob; q++) CHK(cgfdh_host_free(hb[h][q]));
    free(hb[h]);
    free(os[h].tk);
    free(os[h].rec);
  }
  free(jn);
  for (q = 0; q < nsl; q++) CHKIO(cgfdh_nc_close(slf[q]));
  for (int n = 0; n < nsn; n++) CHKIO(cgfdh_nc_close(snf[n]));
  free(slf); free(snf); free(snrec);

  if (d.free_top) {   /* PG_slice_output (drv_rk_curv_col.c:546-550): 15 maps [ny][nx] */
    float *pg = malloc(15 * (size_t)gd->nx * gd->ny * sizeof(float));
    CHK(cgfdh_get_peak_ground_motion(s, pg, NULL));
    CHKIO(cgfdh_io_pg_output(output_dir, output_fname_part, gd->nx, gd->ny, gd->ni, gd->nj,
                             gd->gni1, gd->gnj1, topoid, pg));
    free(pg);
  }

  if (nr + nl > 0) {   /* recvone[r].seismo and line seismo, [icmp*max_nt + it] */
    int nc = wav->ncmp;            /* 9, or 9 + 6 nmaxwell (wav_t.c:26-33) */
    float *seis = malloc((size_t)(nr + nl) * nc * nt_total * sizeof(float));
    CHK(cgfdh_get_seismograms(s, seis));
    for (int r = 0; r < nr; r++)
      for (int c = 0; c < nc; c++)
        memcpy(iorecv->recvone[r].seismo + (size_t)c * iorecv->max_nt,
               seis + ((size_t)r * nc + c) * nt_total, (size_t)nt_total * sizeof(float));
    int q = nr;
    for (int n = 0; n < ioline->num_of_lines; n++)
      for (int ir = 0; ir < ioline->line_nr[n]; ir++, q++)
        for (int c = 0; c < nc; c++)
          memcpy(ioline->recv_seismo[n] + ((size_t)ir * nc + c) * ioline->max_nt,
                 seis + ((size_t)q * nc + c) * nt_total, (size_t)nt_total * sizeof(float));
    free(seis);
  }
  CHK(cgfdh_destroy(s));
}
