/*
 * abi_replay.c -- a C caller of libcgfdh.so built only from include/cgfdh.h.
 * It replays the call order of the drop-in drv_rk_curv_col_allstep
 * (cgfd3d_amd/integration/drv_rk_curv_col_cgfdh.c) on inputs a test writes:
 *
 *   create -> compute_metric (the metric of the coordinates, on the device)
 *   -> set_medium_iso -> pml_face_coefs + set_pml_face
 *   per face -> compute_free_surface -> set_sources (moment tables, force
 *   tables, strict surface forces) -> set_receivers (receivers, then line
 *   points as zero-weight receivers) -> set_graves_qs -> chunked cgfdh_run
 *   up to every output step with get_subvolume after each chunk ->
 *   get_seismograms, get_peak_ground_motion, get_wavefield -> destroy.
 *
 * Usage: abi_replay <input.bin> <output.bin>
 * Input (native-endian int32 / float32, in this order):
 *   int   ni, nj, nk, nt_total, npml_layers, nsrc, max_nt, nsurf, nrecv,
 *         nline, nout, strict, use_graves
 *   float dt, pml_vel, pml_alpha, pml_beta, qs_freq
 *   float xyz[3][nz][ny][nx], lam[vol], mu[vol], slw[vol]
 *   int   ijk[3 nsrc], it_begin[nsrc], it_end[nsrc]
 *   float mij[6][nsrc][max_nt][4], fxyz[3][nsrc][max_nt][4]
 *   int   surf_idx[nsurf]
 *   float frate[3][nsurf][max_nt][4]
 *   int   recv_idx[nrecv][8]
 *   float recv_w[nrecv][3]
 *   int   line_ptr[nline]
 *   float qs[vol]                                  (if use_graves)
 *   int   out_step[nout], box[9]                   (steps after which the box is read)
 * Output: float seismo[(nrecv+nline)][9][nt_total], pg[15][ny][nx],
 *         w[9][vol], then per output step the box's 9 components.
 * Exit status 0 on success; a failing call prints cgfdh_last_error().
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cgfdh.h"

#define CHK(call)                                                                 \
  do {                                                                            \
    int rc_ = (call);                                                             \
    if (rc_) {                                                                    \
      fprintf(stderr, "abi_replay: %s failed (%d): %s\n", #call, rc_,             \
              cgfdh_last_error());                                                \
      exit(2);                                                                    \
    }                                                                             \
  } while (0)

static FILE *fin;

static void *rd(size_t n, size_t sz)
{
  void *p = malloc(n * sz + 1);
  if (!p || fread(p, sz, n, fin) != n) {
    fprintf(stderr, "abi_replay: short input\n");
    exit(3);
  }
  return p;
}

int main(int argc, char **argv)
{
  if (argc != 3) {
    fprintf(stderr, "usage: abi_replay <input.bin> <output.bin>\n");
    return 1;
  }
  fin = fopen(argv[1], "rb");
  if (!fin) {
    perror(argv[1]);
    return 1;
  }
  int *h = rd(13, sizeof(int));
  int ni = h[0], nj = h[1], nk = h[2], nt_total = h[3], nlay = h[4], nsrc = h[5];
  int max_nt = h[6], nsurf = h[7], nrecv = h[8], nline = h[9], nout = h[10];
  int strict = h[11], use_graves = h[12];
  float *hf = rd(5, sizeof(float));
  float dt = hf[0], pvel = hf[1], palpha = hf[2], pbeta = hf[3], qs_freq = hf[4];
  int nx = ni + 6, ny = nj + 6, nz = nk + 6;
  size_t vol = (size_t)nx * ny * nz;
  float *xyz = rd(3 * vol, sizeof(float));
  float *lam = rd(vol, sizeof(float)), *mu = rd(vol, sizeof(float));
  float *slw = rd(vol, sizeof(float));
  int *ijk = rd(3 * (size_t)nsrc, sizeof(int));
  int *itb = rd(nsrc, sizeof(int)), *ite = rd(nsrc, sizeof(int));
  size_t per = (size_t)nsrc * max_nt * 4;
  float *mij = rd(6 * per, sizeof(float)), *fxyz = rd(3 * per, sizeof(float));
  int *surf_idx = rd(nsurf, sizeof(int));
  float *frate = rd(3 * (size_t)nsurf * max_nt * 4, sizeof(float));
  int *ridx = rd(8 * (size_t)nrecv, sizeof(int));
  float *rw = rd(3 * (size_t)nrecv, sizeof(float));
  int *lptr = rd(nline, sizeof(int));
  float *qs = use_graves ? rd(vol, sizeof(float)) : NULL;
  int *out_step = rd(nout, sizeof(int));
  int *box = rd(9, sizeof(int));
  fclose(fin);

  cgfdh_desc d;
  memset(&d, 0, sizeof d);
  d.ni = ni; d.nj = nj; d.nk = nk;
  d.medium_type = CGFDH_MEDIUM_ELASTIC_ISO;
  d.free_top = 1;
  for (int f = 0; f < 5; f++) {
    d.pml_on[f] = 1;
    d.pml_layers[f] = nlay;
  }
  d.dt = dt;
  d.nt_total = nt_total;
  cgfdh_solver *s = NULL;
  CHK(cgfdh_create(&d, &s));
  CHK(cgfdh_compute_metric(s, xyz));
  CHK(cgfdh_set_medium_iso(s, lam, mu, slw));
  float *A = malloc((nlay + 1) * sizeof(float)), *B = malloc((nlay + 1) * sizeof(float));
  float *D = malloc((nlay + 1) * sizeof(float));
  for (int f = 0; f < 5; f++) {
    CHK(cgfdh_pml_face_coefs(ni, nj, nk, xyz, f, nlay, pvel, palpha, pbeta, A, B, D));
    CHK(cgfdh_set_pml_face(s, f, A, B, D));
  }
  CHK(cgfdh_compute_free_surface(s));
  CHK(cgfdh_set_sources(s, nsrc, ijk, itb, ite, max_nt, mij, fxyz, CGFDH_SRC_POINT, 0, NULL,
                        strict, nsurf, nsurf ? surf_idx : NULL, nsurf ? frate : NULL));
  /* receivers, then the line points as receivers whose 8 corners are the
   * point, with zero weights */
  int nall = nrecv + nline;
  int *idx = malloc(8 * (size_t)nall * sizeof(int));
  float *dijk = malloc(3 * (size_t)nall * sizeof(float));
  memcpy(idx, ridx, 8 * (size_t)nrecv * sizeof(int));
  memcpy(dijk, rw, 3 * (size_t)nrecv * sizeof(float));
  for (int q = 0; q < nline; q++) {
    for (int c = 0; c < 8; c++) idx[8 * (nrecv + q) + c] = lptr[q];
    dijk[3 * (nrecv + q)] = dijk[3 * (nrecv + q) + 1] = dijk[3 * (nrecv + q) + 2] = 0.0f;
  }
  CHK(cgfdh_set_receivers(s, nall, idx, dijk));
  if (use_graves) CHK(cgfdh_set_graves_qs(s, qs, qs_freq));

  FILE *fo = fopen(argv[2], "wb");
  if (!fo) {
    perror(argv[2]);
    return 1;
  }
  size_t nbox = (size_t)box[1] * box[4] * box[7];
  float *pack = malloc(9 * nbox * sizeof(float));
  const int comps[9] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
  float **boxes = malloc(nout * sizeof(float *));
  int it_run = 0;
  for (int n = 0; n < nout; n++) {
    /* run up to and including the output step, as the shim does */
    CHK(cgfdh_run(s, it_run, out_step[n] + 1 - it_run));
    it_run = out_step[n] + 1;
    CHK(cgfdh_get_subvolume(s, box, 9, comps, pack));
    boxes[n] = malloc(9 * nbox * sizeof(float));
    memcpy(boxes[n], pack, 9 * nbox * sizeof(float));
  }
  if (it_run < nt_total) CHK(cgfdh_run(s, it_run, nt_total - it_run));
  float *seis = malloc((size_t)nall * 9 * nt_total * sizeof(float));
  float *pg = malloc(15 * (size_t)nx * ny * sizeof(float));
  float *w = malloc(9 * vol * sizeof(float));
  CHK(cgfdh_get_seismograms(s, seis));
  CHK(cgfdh_get_peak_ground_motion(s, pg, NULL));
  CHK(cgfdh_get_wavefield(s, w));
  CHK(cgfdh_destroy(s));
  int ok = fwrite(seis, sizeof(float), (size_t)nall * 9 * nt_total, fo) ==
               (size_t)nall * 9 * nt_total &&
           fwrite(pg, sizeof(float), 15 * (size_t)nx * ny, fo) == 15 * (size_t)nx * ny &&
           fwrite(w, sizeof(float), 9 * vol, fo) == 9 * vol;
  for (int n = 0; n < nout && ok; n++)
    ok = fwrite(boxes[n], sizeof(float), 9 * nbox, fo) == 9 * nbox;
  fclose(fo);
  if (!ok) {
    fprintf(stderr, "abi_replay: short write\n");
    return 4;
  }
  printf("abi_replay: %dx%dx%d, %d steps, %d receivers + %d line points, %d outputs\n", ni, nj,
         nk, nt_total, nrecv, nline, nout);
  return 0;
}
