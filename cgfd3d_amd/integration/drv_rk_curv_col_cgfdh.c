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
 * included), receivers and receiver lines, slices, snapshots (velocity,
 * stress, strain), the peak-ground-motion maps, the ablexp sponge and
 * Graves-Qs attenuation with any elastic medium.  Slices
 * and snapshots are packed on the GPU; only their points are copied into the host wavefield that the reference's own writers
 * (io_slice_nc_put, io_snap_nc_put) then read, so the files are the
 * reference's; the output_all debug dump copies the whole wavefield back each
 * step.  Anything else the par file asks for (acoustic medium, dd sources,
 * several MPI ranks) stops the run with a message, the reference's
 * fatal-error convention.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mpi.h>

#include "drv_rk_curv_col.h"
#include "cgfdh.h"

static void cgfdh_fatal(const char *what, int rc)
{
  fprintf(stderr, "ERROR: %s failed (%d): %s\n", what, rc, cgfdh_last_error());
  MPI_Abort(MPI_COMM_WORLD, 1);
}

#define CHK(call) do { int rc_ = (call); if (rc_) cgfdh_fatal(#call, rc_); } while (0)

/* copy the packed box {i1,ni,di, j1,nj,dj, k1,nk,dk} of components comps
 * from the engine into the host wavefield w4d at the same points */
static void fill_box(cgfdh_solver *s, const gd_t *gd, size_t siz_icmp, float *w4d,
                     const int box[9], int ncomp, const int *comps, float *pack)
{
  CHK(cgfdh_get_subvolume(s, box, ncomp, comps, pack));
  size_t t = 0;
  for (int c = 0; c < ncomp; c++) {
    float *dst = w4d + (size_t)comps[c] * siz_icmp;
    for (int n3 = 0; n3 < box[7]; n3++)
      for (int n2 = 0; n2 < box[4]; n2++)
        for (int n1 = 0; n1 < box[1]; n1++, t++) {
          size_t i = box[0] + (size_t)n1 * box[2];
          size_t j = box[3] + (size_t)n2 * box[5];
          size_t k = box[6] + (size_t)n3 * box[8];
          dst[i + j * gd->siz_iy + k * gd->siz_iz] = pack[t];
        }
  }
}

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
  if ((mt != CONST_MEDIUM_ELASTIC_ISO && mt != CONST_MEDIUM_ELASTIC_VTI &&
       mt != CONST_MEDIUM_ELASTIC_ANISO && !visco) ||
      (visco && md->visco_type != CONST_VISCO_GMB) ||
      src->dd_is_valid) {
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

  /* slices every step, snapshots at their own steps (drv_rk_curv_col.c:503-509):
   * the engine runs up to each output step, then the packed points go into
   * level 0 of the reference's host wavefield (wav->v5d), level 1 serving as
   * the writers' buffer as w_rhs does in the reference */
  int *topoid = mympi->topoid;
  ioslice_nc_t ioslice_nc;
  io_slice_nc_create(ioslice, wav->ncmp, wav->visco_type, wav->cmp_name, gd->ni, gd->nj,
                     gd->nk, topoid, &ioslice_nc);
  iosnap_nc_t iosnap_nc;
  io_snap_nc_create(iosnap, &iosnap_nc, topoid);
  int nslice[3] = {ioslice->num_of_slice_x, ioslice->num_of_slice_y, ioslice->num_of_slice_z};
  int has_slice = nslice[0] + nslice[1] + nslice[2] > 0;
  size_t pack_n = 9 * (size_t)gd->nx * gd->ny;
  if (9 * (size_t)gd->nx * gd->nz > pack_n) pack_n = 9 * (size_t)gd->nx * gd->nz;
  if (9 * (size_t)gd->ny * gd->nz > pack_n) pack_n = 9 * (size_t)gd->ny * gd->nz;
  for (int n = 0; n < iosnap->num_of_snap; n++) {
    size_t m = 6 * (size_t)iosnap->ni[n] * iosnap->nj[n] * iosnap->nk[n];
    if (m > pack_n) pack_n = m;
  }
  float *pack = (has_slice || iosnap->num_of_snap > 0) ? malloc(pack_n * sizeof(float)) : NULL;
  float *w4d = wav->v5d;
  float *wbuf = wav->v5d + wav->siz_ilevel;
  const int cmp9[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  int it_run = 0;
  char ou_file[CONST_MAX_STRLEN];
  for (int it = 0; it < nt_total; it++) {
    int due = has_slice || output_all;
    for (int n = 0; n < iosnap->num_of_snap && !due; n++) due = snap_due(iosnap, n, it, nt_total);
    if (!due) continue;
    CHK(cgfdh_run(s, it_run, it + 1 - it_run));
    it_run = it + 1;
    float t_cur = it * dt + t0;
    float t_end = t_cur + dt;
    /* slices: components 0..8 (all of them elastic, the first 9 for GMB) */
    for (int n = 0; n < nslice[0]; n++) {
      int box[9] = {ioslice->slice_x_indx[n], 1, 1, gd->nj1, gd->nj, 1, gd->nk1, gd->nk, 1};
      fill_box(s, gd, wav->siz_icmp, w4d, box, 9, cmp9, pack);
    }
    for (int n = 0; n < nslice[1]; n++) {
      int box[9] = {gd->ni1, gd->ni, 1, ioslice->slice_y_indx[n], 1, 1, gd->nk1, gd->nk, 1};
      fill_box(s, gd, wav->siz_icmp, w4d, box, 9, cmp9, pack);
    }
    for (int n = 0; n < nslice[2]; n++) {
      int box[9] = {gd->ni1, gd->ni, 1, gd->nj1, gd->nj, 1, ioslice->slice_z_indx[n], 1, 1};
      fill_box(s, gd, wav->siz_icmp, w4d, box, 9, cmp9, pack);
    }
    /* snapshots: velocities, and the stresses the stress and strain outputs read */
    for (int n = 0; n < iosnap->num_of_snap; n++) {
      if (!snap_due(iosnap, n, it, nt_total)) continue;
      int box[9] = {iosnap->i1[n], iosnap->ni[n], iosnap->di[n], iosnap->j1[n], iosnap->nj[n],
                    iosnap->dj[n], iosnap->k1[n], iosnap->nk[n], iosnap->dk[n]};
      if (iosnap->out_vel[n]) fill_box(s, gd, wav->siz_icmp, w4d, box, 3, cmp9, pack);
      if (iosnap->out_stress[n] || iosnap->out_strain[n])
        fill_box(s, gd, wav->siz_icmp, w4d, box, 6, cmp9 + 3, pack);
    }
    io_slice_nc_put(ioslice, &ioslice_nc, gd, w4d, wbuf, it, t_end, 0, wav->ncmp - 1,
                    wav->visco_type);
    io_snap_nc_put(iosnap, &iosnap_nc, gd, md, wav, w4d, wbuf, nt_total, it, t_end, 1, 1, 1);
    if (output_all) {   /* debug dump of the whole w_end (drv_rk_curv_col.c:513-525) */
      CHK(cgfdh_get_wavefield(s, w4d));
      io_build_fname_time(output_dir, "w3d", ".nc", topoid, it, ou_file);
      io_var3d_export_nc(ou_file, w4d, wav->cmp_pos, wav->cmp_name, wav->ncmp, gd->index_name,
                         gd->nx, gd->ny, gd->nz);
    }
  }
  if (it_run < nt_total) CHK(cgfdh_run(s, it_run, nt_total - it_run));
  free(pack);
  io_slice_nc_close(&ioslice_nc);
  io_snap_nc_close(&iosnap_nc);

  if (d.free_top) {   /* PG_slice_output (drv_rk_curv_col.c:546-550): 15 maps [ny][nx] */
    float *pg = malloc(15 * (size_t)gd->nx * gd->ny * sizeof(float));
    CHK(cgfdh_get_peak_ground_motion(s, pg, NULL));
    PG_slice_output(pg, gd, output_dir, output_fname_part, topoid);
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
