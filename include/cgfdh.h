/*
 * cgfdh.h -- C ABI of the MI355X (gfx950) time-stepping engine for the
 * curvilinear collocated-grid elastic solver.
 *
 * The reference's drop-in boundary for this path is the C function
 *   drv_rk_curv_col_allstep(fd_t*, gd_t*, gdcurv_metric_t*, md_t*, src_t*,
 *                           bdry_t*, wav_t*, mympi_t*, iorecv_t*, ioline_t*,
 *                           ioslice_t*, iosnap_t*, float dt, int nt_total,
 *                           float t0, char*, char*, int, int, int)
 *   (forward/drv_rk_curv_col.h:17-37, called from main_curv_col_el_3d.c:849-858)
 * together with the one-stage RHS it dispatches to,
 *   sv_curv_col_el_iso_onestage(...)    (forward/sv_curv_col_el_iso.h:17-32),
 *   sv_curv_col_el_vti_onestage(...)    (forward/sv_curv_col_el_vti.h:17-31),
 *   sv_curv_col_el_aniso_onestage(...)  (forward/sv_curv_col_el_aniso.h:17-32),
 *   sv_curv_col_vis_iso_onestage(...)   (forward/sv_curv_col_vis_iso.h:17-33).
 * This ABI carries the same data with plain pointers and sizes, so a C shim
 * that unpacks those structs into these calls replaces the reference's time
 * loop (cgfd3d_amd/integration/drv_rk_curv_col_cgfdh.c is that shim).  No
 * torch or HIP types appear.
 *
 * Conventions (all follow the reference):
 *  - grids carry 3 ghost points per side: nx=ni+6, ny=nj+6, nz=nk+6;
 *    iptr = i + j*nx + k*nx*ny, x fastest (gd_t.c:2890-2897);
 *  - wavefield components Vx,Vy,Vz,Txx,Tyy,Tzz,Tyz,Txz,Txy (wav_t.c:93-136),
 *    followed for the visco-elastic medium by Jxx,Jyy,Jzz,Jyz,Jxz,Jxy of
 *    each Maxwell body (wav_t.c:138-171): ncomp = 9 + 6*nmaxwell components
 *    in every wavefield array and seismogram of this ABI;
 *  - metric components jac,xi_x,xi_y,xi_z,eta_x,eta_y,eta_z,zeta_x,zeta_y,
 *    zeta_z (gd_t.c:101-180);
 *  - the medium's density is passed as slowness 1/rho (md_t.c:602-623);
 *  - PML faces are numbered idim*2+iside: x1,x2,y1,y2,z1,z2 (bdry_t.h);
 *  - time step it uses the operator pair it % 8 (drv_rk_curv_col.c:175).
 * Every function returns 0 on success or a negative CGFDH_E* code; the
 * message of the last failure on the calling thread is cgfdh_last_error().
 * The library is single-caller per solver, like the reference's driver.
 */
#ifndef CGFDH_H
#define CGFDH_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CGFDH_OK 0
#define CGFDH_EINVAL (-1)
#define CGFDH_ENOMEM (-2)
#define CGFDH_EDEVICE (-3)
#define CGFDH_ESTATE (-4)

/* medium codes are the reference's (constants.h) */
#define CGFDH_MEDIUM_ELASTIC_ISO 2
#define CGFDH_MEDIUM_ELASTIC_VTI 3
#define CGFDH_MEDIUM_ELASTIC_ANISO 4
#define CGFDH_MEDIUM_VISCO_ISO 5       /* visco-elastic, GMB memory variables */
#define CGFDH_MAXWELL_MAX 4

#define CGFDH_SRC_POINT 1     /* CONST_SRC_SPATIAL_POINT */
#define CGFDH_SRC_GAUSSIAN 2  /* CONST_SRC_SPATIAL_GAUSSIAN */

typedef struct cgfdh_solver cgfdh_solver;

typedef struct {
  int ni, nj, nk;       /* physical grid points of the block */
  int medium_type;      /* CGFDH_MEDIUM_ELASTIC_ISO, _VTI, _ANISO or _VISCO_ISO */
  int free_top;         /* 1: traction-image free surface on z2 */
  int pml_on[6];        /* CFS-PML per face */
  int pml_layers[6];    /* number_of_layers per face */
  float dt;             /* size_of_time_step */
  int nt_total;         /* number_of_time_steps + 1 (main_curv_col_el_3d.c:615) */
  int nmaxwell;         /* visco-elastic: Maxwell bodies, 1..CGFDH_MAXWELL_MAX */
} cgfdh_desc;

/* Replaces the allocations of wav_init / bdry_pml_auxvar_init / bdry_free_set
 * for the device (wav_t.c:15-179, bdry_t.c:291-327, bdry_t.c:49-112). */
int cgfdh_create(const cgfdh_desc *desc, cgfdh_solver **out);
int cgfdh_destroy(cgfdh_solver *s);
const char *cgfdh_last_error(void);

/* gdcurv_metric_t v4d, 10 volumes in the order above (gd_t.c:98-185). */
int cgfdh_set_metric(cgfdh_solver *s, const float *metric10);

/* md_t lambda, mu and rho-as-slowness (md_t.c:90-110, 602-623). */
int cgfdh_set_medium_iso(cgfdh_solver *s, const float *lambda, const float *mu,
                         const float *slowness);

/* md_t c11, c13, c33, c55, c66 of a VTI medium (md_t.c:112-134) and
 * rho-as-slowness; c12 = c11 - 2 c66 is formed per point as
 * sv_curv_col_el_vti.c:351 does. */
int cgfdh_set_medium_vti(cgfdh_solver *s, const float *c11, const float *c13,
                         const float *c33, const float *c55, const float *c66,
                         const float *slowness);

/* md_t c11,c12,c13,c14,c15,c16,c22,c23,c24,c25,c26,c33,c34,c35,c36,c44,c45,
 * c46,c55,c56,c66 of a general anisotropic medium (md_t.c:136-221), 21
 * volumes back to back, and rho-as-slowness. */
int cgfdh_set_medium_aniso(cgfdh_solver *s, const float *cij21,
                           const float *slowness);

/* md_t of the GMB visco-elastic medium after md_vis_GMB_cal_Y (md_t.c:643-792;
 * cgfdh_visco_gmb builds it): unrelaxed lambda, mu, rho-as-slowness, then
 * Ylam_n, Ymu_n interleaved per Maxwell body (2*nmaxwell volumes, the md_t
 * order, md_t.c:240-249), and the bodies' relaxation frequencies wl. */
int cgfdh_set_medium_visco(cgfdh_solver *s, const float *lambda, const float *mu,
                           const float *slowness, const float *ylam_ymu,
                           const float *wl);

/* bdry_t A/B/D of one face, num_layers+1 values each, already converted as
 * bdry_pml_set leaves them (D/=B, A+=D, B=1/B; bdry_t.c:259-268). */
int cgfdh_set_pml_face(cgfdh_solver *s, int face, const float *A, const float *B,
                       const float *D);

/* bdry_t matVx2Vz2 / matVy2Vz2 / matF2Vz2, [ny][nx][3][3] each
 * (sv_curv_col_el_iso_dvh2dvz, sv_curv_col_el_iso.c:1259-1375; the VTI and
 * anisotropic versions, sv_curv_col_el_vti.c:1085-1203 and
 * sv_curv_col_el_aniso.c:1267-1417, leave matF2Vz2 zero). */
int cgfdh_set_free_surface(cgfdh_solver *s, const float *matVx2Vz2,
                           const float *matVy2Vz2, const float *matF2Vz2);

/* bdry_t matD [ny][nx][3][3] of the visco-elastic free surface
 * (sv_curv_col_vis_iso.c:485-501), the rotation applied to the surface-row
 * stresses after every step (sv_curv_col_vis_iso_free, drv_rk_curv_col.c:440-445);
 * cgfdh_compute_metric builds it on the device for such a solver. */
int cgfdh_set_free_surface_matD(cgfdh_solver *s, const float *matD);
int cgfdh_get_free_surface_matD(cgfdh_solver *s, float *matD);

/* Graves-Qs attenuation (md->visco_type CONST_VISCO_GRAVES_QS, any elastic
 * medium: isotropic, VTI or anisotropic): md->Qs, one volume, and
 * md->visco_Qs_freq.  After the
 * last stage of every step each physical point, all components, is
 * multiplied by exp(-pi * freq * dt / Qs) (sv_curv_col_el_graves_Qs,
 * sv_curv_col_el.c:638-667, called at drv_rk_curv_col.c:414-416); the
 * exponential is the double one rounded to float, within one ulp of the
 * reference's expf. */
int cgfdh_set_graves_qs(cgfdh_solver *s, const float *Qs, float freq);

/* The ablexp sponge as bdry_ablexp_set leaves it (bdry_t.c:513-820): masks
 * Ex[nx], Ey[ny], Ez[nz] (bdry->ablexp_Ex/Ey/Ez) and the 6 blocks of
 * bdry->bdry_blk as {enable, ni1, ni2, nj1, nj2, nk1, nk2} each (array
 * indices, inclusive).  After every step each enabled block's points, all
 * ncomp components, are multiplied by min(Ex[i], Ey[j], Ez[k]) as
 * bdry_ablexp_apply does (bdry_t.c:840-890, drv_rk_curv_col.c:482-486),
 * before peak ground motion and receivers.  No enabled block turns it off. */
int cgfdh_set_ablexp(cgfdh_solver *s, const float *Ex, const float *Ey, const float *Ez,
                     const int *blocks);

/* Moment sources as src_t holds them (src_t.h; src_t.c:985-1174):
 *  ijk      local indices with ghosts (si,sj,sk), 3 per source;
 *  it_begin/it_end per source; max_nt;
 *  mij      Mxx,Myy,Mzz,Mxz,Myz,Mxy, each [nsrc][max_nt][4 stages];
 *  spatial  CGFDH_SRC_POINT or CGFDH_SRC_GAUSSIAN;
 *  ext_half half width of the Gaussian (src->ext_half_npoint);
 *  ext_coefs the normalised (2*ext_half+1)^3 weights per source, as
 *           src_cal_norm_delt3d_z2fre leaves them (NULL for point sources). */
int cgfdh_set_moment_sources(cgfdh_solver *s, int nsrc, const int *ijk,
                             const int *it_begin, const int *it_end, int max_nt,
                             const float *mij, int spatial, int ext_half,
                             const float *ext_coefs);

/* Force and moment sources together, as src_t holds them (src_t.h;
 * src_t.c:26-100, 985-1174), replacing cgfdh_set_moment_sources' set:
 *  ijk, it_begin, it_end, max_nt, spatial, ext_half, ext_coefs as there;
 *  mij      the 6 moment tables, or NULL (src->moment_actived 0);
 *  fxyz     Fx, Fy, Fz, each [nsrc][max_nt][4 stages], or NULL
 *           (src->force_actived 0); a force adds f * coef * slw / jac to
 *           the velocity rates (sv_curv_col_el.c:356-364, 440-452);
 *  strict   src->is_surface_force_strict: forces on the surface row k = nk2
 *           are then left to the free surface (sv_curv_col_el.c:388, 447);
 *  nsurf    src->total_number_surface_force: the strict forces that reach
 *           the surface (src_glob_ext_reach_free, src_t.c:361-371: the
 *           point on k = nk2, or the Gaussian's top row on it), named by
 *  surf_idx src->force_rate_indx (source numbers), with
 *  frate    Fx_rate, Fy_rate, Fz_rate, each [nsurf][max_nt][4]
 *           (src_t.c:1080-1090).  Before every stage the engine forms the
 *           reference's surface-force slices from them
 *           (src_set_surface_layer_for_force, src_t.c:154-314): TxSrc..TzSrc
 *           become the surface traction of the traction image
 *           (sv_curv_col_el.c:147-163) and VxSrc..VzSrc enter the surface
 *           velocity derivatives through matF2Vz2 (isotropic media,
 *           sv_curv_col_el_iso.c:583-592).
 * nsurf > 0 needs fxyz, strict and the free surface. */
int cgfdh_set_sources(cgfdh_solver *s, int nsrc, const int *ijk, const int *it_begin,
                      const int *it_end, int max_nt, const float *mij, const float *fxyz,
                      int spatial, int ext_half, const float *ext_coefs, int strict,
                      int nsurf, const int *surf_idx, const float *frate);

/* Distributed (dd) sources added at grid points, the src_t dd_* members that
 * src_dd_read2local fills (src_t.c:1180-1900, dd_is_add_at_point = 1) and
 * sv_curv_col_el_rhs_srcdd adds (sv_curv_col_el.c:486-555, called from the
 * isotropic and GMB right-hand sides only): ndd points at the 1-D array
 * indices indx (src->dd_indx, physical points, duplicates allowed and added
 * in input order), steps 0 .. max_nt-1 (src->dd_max_nt), a device table of
 * nt_cap steps (src->dd_nt_per_read), velocity values (has_vi,
 * src->dd_vi_actived) and/or moment values (has_mij, src->dd_mij_actived).
 * ndd = 0 removes them.  Isotropic and GMB media only. */
int cgfdh_set_dd_sources(cgfdh_solver *s, int ndd, const size_t *indx, int max_nt, int nt_cap,
                         int has_vi, int has_mij);
/* Values of steps it0 .. it0+nt-1 (nt <= nt_cap), as the reference's chunk
 * holds them after src_dd_accit_loadstf (src_t.c:1929-2005):
 * vi[nt][4][ndd][3] and mij[nt][4][ndd][6] (Mxx,Myy,Mzz,Myz,Mxz,Mxy), each
 * [stage][source] row the value of that RK stage.  Step it occupies table
 * row it % nt_cap, replacing what it held.  A step below max_nt whose values
 * are not in the table makes cgfdh_run / cgfdh_onestage / cgfdh_batch_run
 * fail; from step max_nt on nothing is added. */
int cgfdh_put_dd_values(cgfdh_solver *s, int it0, int nt, const float *vi, const float *mij);

/* Receivers as iorecv_one_t keeps them: indx1d[8] and di,dj,dk
 * (io_funcs.c:358-389); samples follow io_recv_keep (io_funcs.c:1585-1623). */
int cgfdh_set_receivers(cgfdh_solver *s, int nrecv, const int *indx1d,
                        const float *dijk);

/* Whole-state access (level "pre": the state at the start of step it0),
 * ncomp volumes. */
int cgfdh_set_wavefield(cgfdh_solver *s, const float *w9);
int cgfdh_get_wavefield(cgfdh_solver *s, float *w9);
/* The value check of wav_check_value (wav_t.c:253-273) on the current state,
 * as the par file's check_nan_every_nummber_of_steps asks for it
 * (drv_rk_curv_col.c:476-479): *icmp and *iptr of the first NaN in its scan
 * order (component by component, array index within each), or -1 and -1
 * when there is none. */
int cgfdh_check_nan(cgfdh_solver *s, int *icmp, long *iptr);
size_t cgfdh_pml_aux_len(cgfdh_solver *s);   /* floats, all enabled faces */
int cgfdh_set_pml_aux(cgfdh_solver *s, const float *aux);
int cgfdh_get_pml_aux(cgfdh_solver *s, float *aux);

/* Advance steps it0 .. it0+nsteps-1 (drv_rk_curv_col.c:167-544): four fused
 * stage launches per step, receivers sampled after every step. */
int cgfdh_run(cgfdh_solver *s, int it0, int nsteps);

/* seismo[(r*ncomp + icmp)*nt_total + it], the layout of
 * iorecv->recvone[r].seismo (all ncomp components, io_funcs.c:1598-1612) */
int cgfdh_get_seismograms(cgfdh_solver *s, float *seismo);

/* Peak ground motion on the free surface, updated after every step as
 * PG_calcu does (wav_t.c:380-455, called at drv_rk_curv_col.c:492-495):
 * pg = 15 maps [ny][nx] in the order PGV,PGVh,PGVx,PGVy,PGVz,PGA,PGAh,PGAx,
 * PGAy,PGAz,PGD,PGDh,PGDx,PGDy,PGDz (the PG buffer PG_slice_output writes);
 * dis = the 3 accumulated displacement maps.  Either pointer may be NULL. */
int cgfdh_get_peak_ground_motion(cgfdh_solver *s, float *pg, float *dis);

/* Output sub-volumes of the wavefield after the last step, packed on the
 * device so that only the packed points are copied to the host.
 * box = {i1, ni, di, j1, nj, dj, k1, nk, dk}: start, count and stride per
 * axis in array indices (ghosts included), every point inside the array.
 *
 * cgfdh_get_subvolume: out[c][n3][n2][n1] = component comps[c] at
 * (i1 + n1*di, j1 + n2*dj, k1 + n3*dk), i fastest: io_snap_pack_buff
 * (io_funcs.c:1525-1556) for each velocity/stress snapshot component
 * (io_funcs.c:1162-1222), and the x/y/z slices of io_slice_nc_put
 * (io_funcs.c:1022-1108) as boxes one point thick with unit strides.
 *
 * cgfdh_get_strain_subvolume: out [6][nk][nj][ni] = Exx, Eyy, Ezz, Exz,
 * Eyz, Exy from the stresses and lambda, mu, with the rounding of
 * io_snap_stress_to_strain_eliso (io_funcs.c:1456-1522), in the order
 * io_snap_nc_put writes the strain variables (io_funcs.c:1226-1252).
 * Isotropic elastic and visco-elastic media only. */
int cgfdh_get_subvolume(cgfdh_solver *s, const int *box, int ncomp, const int *comps,
                        float *out);

/* cgfdh_get_subvolume for callers that cannot allocate (code under stream
 * capture, tight serving loops, harnesses that forbid hipMalloc); it has no
 * counterpart in the reference.  The same packed output, through a fixed
 * device buffer of CGFDH_SUBVOLUME_STATIC_MAX floats (32 MiB per GPU, held
 * from the library's load until the process ends) instead of the solver's
 * growing staging buffer: no hipMalloc, hipFree or other allocating call.
 * ncomp times the box's point count above the limit is rejected with
 * CGFDH_EINVAL.  Every call on one GPU shares that buffer, which keeps the
 * last call's packed values: calls that overlap on one GPU (several host
 * threads, or solvers on different streams of one device at once) race. */
#define CGFDH_SUBVOLUME_STATIC_MAX (8 * 1024 * 1024)
int cgfdh_get_subvolume_noalloc(cgfdh_solver *s, const int *box, int ncomp, const int *comps,
                                float *out);
int cgfdh_get_strain_subvolume(cgfdh_solver *s, const int *box, float *out);

/* Asynchronous steps and output, so that packing and the copy to the host
 * overlap the steps after them (the reference's loop writes its slices and
 * snapshots between steps, drv_rk_curv_col.c:497-525).
 * cgfdh_run_async enqueues steps it0 .. it0+nsteps-1 and returns;
 * cgfdh_subvolume_async packs a box of the state left by every step
 * enqueued so far (cgfdh_get_subvolume's layout) on the solver's stream and
 * copies it to out on a copy stream, returning a ticket; cgfdh_wait(ticket)
 * returns once out holds the values, cgfdh_sync once everything enqueued
 * has finished.  out must stay valid until then; it overlaps with the steps
 * when it is pinned memory from cgfdh_host_alloc.  Up to 16 boxes can be in
 * flight per solver (a 17th request waits for the oldest copy). */
int cgfdh_run_async(cgfdh_solver *s, int it0, int nsteps);
int cgfdh_subvolume_async(cgfdh_solver *s, const int *box, int ncomp, const int *comps,
                          float *out, int *ticket);
int cgfdh_wait(cgfdh_solver *s, int ticket);
int cgfdh_sync(cgfdh_solver *s);
int cgfdh_host_alloc(size_t bytes, void **p);
int cgfdh_host_free(void *p);

/* Point location on a curvilinear grid, computed on the GPU, for npts
 * points at once: gd_curv_coord_to_local_indx (gd_t.c:1998-2173) with its
 * cell boxes and 4x4x4 tiles (gd_curv_set_minmax, gd_t.c:1708-1880), the
 * point-in-cell test isPointInHexahedron (media_geometry3d.cpp:198-270) and
 * the 101^3-sample shift of gd_curv_coord2shift_sample (gd_t.c:2304-2378).
 * xyz = x3d, y3d, z3d with ghosts as gd_t holds them; pts[n][3] the x, y, z
 * of each point.  Per point: is_here (1 when a physical cell holds it, the
 * first in the reference's tile and cell order), ijk (index with ghosts, the
 * cell's corner or the next point where the shift passed one half) and inc
 * (the shift, in (-0.5, 0.5]); ijk = -1 and inc = 0 when not here.  The
 * global index of a one-block grid is ijk - 3 (gd_curv_coord_to_glob_indx). */
int cgfdh_locate_points(int ni, int nj, int nk, const float *xyz, int npts, const float *pts,
                        int *is_here, int *ijk, float *inc);

/* gd_curv_depth_to_axis (gd_t.c:2179-2296) for npts points at once: xyd[n]
 * = x, y, depth below the top surface; z[n] = the surface's z at (x, y) from
 * the two triangles of its surface cell (fdlib_math_Point2InWhichTri,
 * fdlib_math_computeZfromTri) minus the depth, found[n] = 1; where no
 * surface cell lies over (x, y), found = 0 and z keeps the depth, as the
 * reference leaves it. */
int cgfdh_depth_to_axis(int ni, int nj, int nk, const float *xyz, int npts, const float *xyd,
                        int *found, float *z);

/* The curvilinear grid of one block generated from a layer file on the GPU,
 * gd_curv_gen_layer (gd_t.c:987-1480) after the file is read
 * (cgfdh_io_read_grid_layer reads it): n_if interfaces with ncell cells and
 * an equal-spacing flag per layer (n_if - 1 each), layer_xyz the x, y, z of
 * every interface sample ([3][n_if][ny_layers][nx_layers], bottom interface
 * first); resample and start the par file's refine factors and start
 * indices (x, y, z); nt* the total physical points and g*1 the block's first
 * global index (0 for one block).  The interfaces are resampled in x and y
 * (every |f|-th sample, or the reference's cubic spline), every column gets
 * its z values through the layers (gd_grid_z_interp, graded spacing in
 * unequal middle layers) and x, y as cubic splines of z, and three ghost
 * points are mirrored above the top.  xyz receives x3d, y3d, z3d of the
 * block with ghosts.  The reference's parameter checks (gd_t.c:1151-1177)
 * and unordered samples or heights are errors. */
int cgfdh_gen_grid_layer(int n_if, const int *ncell, const int *equal, int nx_layers,
                         int ny_layers, const float *layer_xyz, const int *resample,
                         const int *start, int ntx, int nty, int ntz, int ni, int nj, int nk,
                         int gni1, int gnj1, int gnk1, float *xyz);

/* A layered model discretised onto grid points on the GPU with local values:
 * media_layer2model_onecmp (nprop 1) and media_layer2model_el_iso (nprop 3)
 * with equivalent_medium_method "loc" (media_layer2model.cpp:47-190,
 * 422-560, 1239-1383).  The interface mesh has n_if interfaces of mx x my
 * samples at geo = minx, miny, dx, dy; fields as cgfdh_io_read_md3lay gives
 * them.  Per point (x, y, z of npts points, e.g. x3d, y3d, z3d with ghosts):
 * each interface's elevation interpolated bilinearly, the layer holding the
 * point, value + dz^pow * grad of each property (the mean of both layers'
 * values on an interface, the top layer's values above the surface).
 * Outputs var into a, or lambda, mu, rho into a, b, c (mu = vs^2 rho,
 * lambda = vp^2 rho - 2 mu).  A point outside the mesh, or with
 * vp <= sqrt(2) vs, is an error (the reference stops there too).  Points
 * need no grid here; the averaging methods and the other media types are
 * cgfdh_layer2model_grid's. */
int cgfdh_layer2model(int nprop, int n_if, int mx, int my, const float *geo,
                      const float *fields, long npts, const float *x, const float *y,
                      const float *z, float *a, float *b, float *c);

/* targets of cgfdh_layer2model_grid: the reference entry point restated */
#define CGFDH_L2M_ONECMP 0   /* media_layer2model_onecmp: var */
#define CGFDH_L2M_EL_ISO 1   /* media_layer2model_el_iso: lambda, mu, rho */
#define CGFDH_L2M_EL_VTI 2   /* media_layer2model_el_vti: rho, c11, c33, c55, c66, c13 */
#define CGFDH_L2M_EL_ANISO 3 /* media_layer2model_el_aniso: rho, c11, c12, .. c66 (22) */
/* equivalent medium methods */
#define CGFDH_L2M_LOC 0
#define CGFDH_L2M_HAR 1
#define CGFDH_L2M_ARI 2

/* A layered model discretised onto a curvilinear grid on the GPU, every
 * media type and method of media_layer2model_{onecmp,el_iso,el_vti,el_aniso}
 * (media_layer2model.cpp:47-420, with grid type GRID_CURV).  media_type is
 * the file's type code (cgfdh_io_md3lay_info), fields its samples; the grid
 * has nx*ny*nz points (x fastest; x3d, y3d, z3d with ghosts).  out holds
 * the target's arrays in the reference function's argument order (1, 3, 6
 * or 22 host arrays of nx*ny*nz floats).
 *   loc: each point's own value (the point value of its layer, the mean of
 *        both layers' on an interface, the top layer's above the surface),
 *        converted to lambda/mu, VTI or rotated TTI stiffnesses (para2vti,
 *        para2tti with BondTransform);
 *   har, ari: a point (1..n-2 in each direction) whose surrounding half-grid
 *        cell spans layers takes the harmonic (stiffness; kappa and mu for
 *        the isotropic medium) or arithmetic average over the cell's 9^3
 *        sub-points, density always arithmetic (the reference's NG = 8).
 * el_aniso also takes isotropic and VTI files and fills the other
 * stiffnesses as the reference does.  Errors (the reference exits): a
 * media type the target does not take, a point or sub-point outside the
 * interface mesh, an isotropic vp <= sqrt(2) vs.  The "tti" method is not
 * provided (only the isotropic type has one in the reference). */
int cgfdh_layer2model_grid(int media_type, int target, int method, int n_if, int mx, int my,
                           const float *geo, const float *fields, int nx, int ny, int nz,
                           const float *x3d, const float *y3d, const float *z3d,
                           float *const *out);

/* Drop-in replacements of the reference's media_layer2model_onecmp,
 * _el_iso, _el_vti and _el_aniso (media_layer2model.cpp:47-420, same
 * arguments): read the .md3lay file, then cgfdh_layer2model_grid.  grid_type
 * must be 3 (GRID_CURV, what main_curv_col_el_3d.c:328-380 passes); the
 * method is "loc", "har" or "ari"; myid is unused.  CGFDH_OK or an error
 * code with the message in cgfdh_last_error (the reference exits instead). */
int cgfdh_media_layer2model_onecmp(float *var3d, const float *x3d, const float *y3d,
                                   const float *z3d, size_t nx, size_t ny, size_t nz,
                                   int grid_type, const char *in_var_file,
                                   const char *average_method, int myid);
int cgfdh_media_layer2model_el_iso(float *lam3d, float *mu3d, float *rho3d, const float *x3d,
                                   const float *y3d, const float *z3d, size_t nx, size_t ny,
                                   size_t nz, int grid_type, const char *in_3lay_file,
                                   const char *equivalent_medium_method, int myid);
int cgfdh_media_layer2model_el_vti(float *rho, float *c11, float *c33, float *c55, float *c66,
                                   float *c13, const float *x3d, const float *y3d,
                                   const float *z3d, size_t nx, size_t ny, size_t nz,
                                   int grid_type, const char *in_3lay_file,
                                   const char *equivalent_medium_method, int myid);
int cgfdh_media_layer2model_el_aniso(float *rho, float *c11, float *c12, float *c13, float *c14,
                                     float *c15, float *c16, float *c22, float *c23, float *c24,
                                     float *c25, float *c26, float *c33, float *c34, float *c35,
                                     float *c36, float *c44, float *c45, float *c46, float *c55,
                                     float *c56, float *c66, const float *x3d, const float *y3d,
                                     const float *z3d, size_t nx, size_t ny, size_t nz,
                                     int grid_type, const char *in_3lay_file,
                                     const char *equivalent_medium_method, int myid);

/* The grid model of media_grid2model (.md3grd) on a curvilinear grid on
 * the GPU, every media type and method of media_grid2model_{onecmp,el_iso,
 * el_vti,el_aniso} (media_grid2model.cpp:29-2800, GRID_CURV): n_layers
 * layers of ngz[l] surfaces (at most 64 layers, 256 surfaces), fields
 * [1 + nprop][surfaces][my][mx] and geo of the cropped mesh as
 * cgfdh_io_read_md3grd gives them.  A point's values are interpolated
 * linearly in z between the surfaces around it (the top surface's above,
 * the last's below, the mean of both sides on a boundary between layers);
 * har / ari as cgfdh_layer2model_grid except that sub-points on a layer
 * boundary are left out and a single layer keeps the local values.
 * Outputs in the reference functions' order; note el_iso's is rho, lambda,
 * mu here (media_grid2model_el_iso:133-150).  Errors as
 * cgfdh_layer2model_grid. */
int cgfdh_grid2model_grid(int media_type, int target, int method, int n_layers, const int *ngz,
                          int mx, int my, const float *geo, const float *fields, int nx, int ny,
                          int nz, const float *x3d, const float *y3d, const float *z3d,
                          float *const *out);

/* Binary volumes sampled at the grid points on the GPU
 * (parameterization_bin_el_iso_loc / _vis_iso_loc, media_bin2model.cpp:
 * 185-432): nvol 3 volumes rho, vp, vs into rho, lambda, mu (out in that
 * order), or nvol 2 volumes Qp, Qs as they are; each volume n3[0] x n3[1] x
 * n3[2] samples (x fastest) on the mesh vectors xv, yv, zv (z may
 * decrease), as cgfdh_io_read_bin_window cuts them; npts points x, y, z.
 * TrilinearInterpolation (media_utility.cpp:304-401) per volume, points
 * outside the mesh clamped to its edge. */
int cgfdh_bin2model_grid(int nvol, const int *n3, const float *xv, const float *yv,
                         const float *zv, const float *vols, long npts, const float *x,
                         const float *y, const float *z, float *const *out);

/* Drop-in replacements of media_bin2model_el_iso and media_bin2model_vis_iso
 * (media_bin2model.cpp:27-183, 236-386, same arguments; main_curv_col_el_3d.c:
 * 466-515): the window of the binary volumes over the grid's extent (x, y
 * from xmin..ymax, z from z3d), read, sampled on the GPU.  grid_type must be
 * 3.  CGFDH_OK or an error code (the reference exits) when bin_order is not
 * a permutation of 0, 1, 2, the model misses the grid, or a file is short. */
int cgfdh_media_bin2model_el_iso(float *rho3d, float *lam3d, float *mu3d, const float *x3d,
                                 const float *y3d, const float *z3d, size_t nx, size_t ny,
                                 size_t nz, float xmin, float xmax, float ymin, float ymax,
                                 int grid_type, int *bin_order, int *bin_size, float *bin_spacing,
                                 float *bin_origin, const char *bin_file_rho,
                                 const char *bin_file_vp, const char *bin_file_vs);
int cgfdh_media_bin2model_vis_iso(float *Qp3d, float *Qs3d, const float *x3d, const float *y3d,
                                  const float *z3d, size_t nx, size_t ny, size_t nz, float xmin,
                                  float xmax, float ymin, float ymax, int grid_type,
                                  int *bin_order, int *bin_size, float *bin_spacing,
                                  float *bin_origin, const char *bin_file_Qp,
                                  const char *bin_file_Qs);

/* Drop-in replacements of media_grid2model_onecmp, _el_iso, _el_vti and
 * _el_aniso (media_grid2model.cpp:29-420, same arguments; Xmin..Ymax the
 * grid's extent, main_curv_col_el_3d.c:388-460): read the window of the
 * .md3grd file, then cgfdh_grid2model_grid.  grid_type must be 3; the
 * method is "loc", "har" or "ari"; myid is unused.  CGFDH_OK or an error
 * code with the message in cgfdh_last_error (the reference exits). */
int cgfdh_media_grid2model_onecmp(float *var3d, const float *x3d, const float *y3d,
                                  const float *z3d, int nx, int ny, int nz, float Xmin, float Xmax,
                                  float Ymin, float Ymax, int grid_type, const char *in_media_file,
                                  const char *equivalent_medium_method, int myid);
int cgfdh_media_grid2model_el_iso(float *rho3d, float *lam3d, float *mu3d, const float *x3d,
                                  const float *y3d, const float *z3d, int nx, int ny, int nz,
                                  float Xmin, float Xmax, float Ymin, float Ymax, int grid_type,
                                  const char *in_media_file, const char *equivalent_medium_method,
                                  int myid);
int cgfdh_media_grid2model_el_vti(float *rho, float *c11, float *c33, float *c55, float *c66,
                                  float *c13, const float *x3d, const float *y3d,
                                  const float *z3d, int nx, int ny, int nz, float Xmin,
                                  float Xmax, float Ymin, float Ymax, int grid_type,
                                  const char *in_media_file, const char *equivalent_medium_method,
                                  int myid);
int cgfdh_media_grid2model_el_aniso(float *rho, float *c11, float *c12, float *c13, float *c14,
                                    float *c15, float *c16, float *c22, float *c23, float *c24,
                                    float *c25, float *c26, float *c33, float *c34, float *c35,
                                    float *c36, float *c44, float *c45, float *c46, float *c55,
                                    float *c56, float *c66, const float *x3d, const float *y3d,
                                    const float *z3d, int nx, int ny, int nz, float Xmin,
                                    float Xmax, float Ymin, float Ymax, int grid_type,
                                    const char *in_media_file,
                                    const char *equivalent_medium_method, int myid);

/* Stable time step of a curvilinear grid, blk_dt_esti_curv
 * (blk_t.c:2840-2930), computed on the GPU: the minimum over physical points
 * of cfl / Vp * L, L the smallest distance from the point to the 8 planes
 * through its diagonal neighbour triples (fdlib_math_dist_point2plane,
 * fdlib_math.c:71-89), Vp per medium as the reference takes it.  Host
 * arrays, ghosts included: xyz = x3d, y3d, z3d; vols = lambda, mu, rho
 * (isotropic), c11, c33, rho (VTI), c11, rho (anisotropic), or lambda, mu,
 * rho, then Ylam_n, Ymu_n interleaved per Maxwell body as for
 * cgfdh_set_medium_visco (visco-elastic, n = nmaxwell).  Outputs the
 * reference's dtmax, dtmaxVp, dtmaxL and dtmaxi/j/k (the first point in its
 * k-j-i loop order on a tie); if no point gives a step below 1e10 (the
 * reference's starting value), dtmax = 1e10, Vp = L = 0 and ijk = -1. */
int cgfdh_estimate_dt(int ni, int nj, int nk, const float *xyz, int medium_type,
                      const float *vols, int nmaxwell, float cfl, float *dtmax,
                      float *vpmax, float *lmin, int *ijk);

/* One RHS evaluation without the RK combine (sv_curv_col_el_iso_onestage):
 * rhs and aux_rhs receive what the reference leaves in w_rhs and the PML
 * aux rhs level.  Host buffers; for parity checks and drop-in use. */
int cgfdh_onestage(cgfdh_solver *s, int it, int ipair, int istage,
                   const float *w_cur, const float *aux_cur, float *rhs,
                   float *aux_rhs);

/* Device time of the last cgfdh_run, and the summed duration of its stage
 * launches, in milliseconds (HIP events on the solver's stream). */
int cgfdh_last_timing(cgfdh_solver *s, double *run_ms, double *stage_ms,
                      int *stage_launches);

/* Selects the stage kernels: 0 = per-point kernel everywhere, 1 = general
 * tiled z-marching kernel everywhere, 2 / 3 = lean z-marching kernel on the
 * interior box (64x8 / 64x16 tiles), 4 / 5 = lean per-point kernel on the
 * interior box (64x4 / 64x2 blocks), 6 / 7 = x-vectorised kernel on the
 * interior box (2 / 4 points per thread), 8 / 9 = LDS-tiled plane kernel on
 * the interior box (64x8 / 64x16 tiles), 10 / 11 = per-point kernel
 * marching 2 / 4 planes per block, 12 / 13 = two-phase per-point kernel
 * (all loads, then the arithmetic; 13 loads the RK combine inputs last),
 * 14 = LDS-tiled plane kernel with the same phased loads, 15 = variant 12
 * compiled for 3 waves per SIMD; 2..15 run the per-point kernel on the PML
 * and free-surface shell; 16 = the two-phase kernel over whole rows with the
 * PML faces folded in (free-surface rows on the per-point kernel); 17 = the
 * two-phase interior kernel on the tiles clear of every PML face plus a
 * shell kernel that requests each PML slab's memory variables first; 18 =
 * variant 16 with the metric, medium, slowness and RK combine inputs loaded
 * straight to LDS (global_load_lds), so the stencil values alone hold
 * registers and four waves fit per SIMD.
 * Default 16 for the isotropic, VTI and visco-elastic media, 18 for the
 * anisotropic medium (whose 21 coefficients otherwise hold registers).
 * Variants 1..3 exist for the isotropic medium only; the visco-elastic
 * medium has 0, 4, 5, 8..11 and 16.
 * All variants produce identical bits. */
int cgfdh_set_kernel_variant(cgfdh_solver *s, int variant);

/* Tile group of the grouped block order (gx x gy tiles per group, 0 0 = the
 * default 9x2) and the solver's current kernel choice. */
int cgfdh_set_tile_group(cgfdh_solver *s, int gx, int gy);
int cgfdh_get_kernel_choice(cgfdh_solver *s, int *variant, int *gx, int *gy);

/* Batched runs of independent solvers (no reference counterpart: the
 * reference runs one grid per process; this serves many-source studies on
 * one GPU).  Every solver is set up with the calls above; cgfdh_batch_run
 * advances them all over the same steps, capturing 8-step cycles of all
 * solvers once into a hipGraph (one stream per solver, concurrent branches)
 * and replaying it; results equal cgfdh_run of each solver bit for bit.
 * With the environment variable CGFDH_BATCH_FUSED=1, when every solver runs
 * variant 16 on one isotropic, VTI or visco-elastic medium, each RK stage of
 * the cycle is instead one launch over the whole rows of all grids (a
 * cross-grid tile queue; each grid's free-surface rows and post-step work on
 * its own stream beside it); same results.
 * cgfdh_batch_autotune times each solver's kernel variants and tile groups
 * for one step from the current state (restored afterwards) and keeps the
 * fastest. */
typedef struct cgfdh_batch cgfdh_batch;
int cgfdh_batch_create(int n, cgfdh_solver *const *solvers, cgfdh_batch **out);
int cgfdh_batch_destroy(cgfdh_batch *b);
int cgfdh_batch_autotune(cgfdh_batch *b, int it0);
int cgfdh_batch_run(cgfdh_batch *b, int it0, int nsteps);
int cgfdh_batch_last_timing(cgfdh_batch *b, double *ms);

/* ---- setup on the device ------------------------------------------------ */

/* Metric from grid coordinates x3d,y3d,z3d (3 volumes with ghosts), as
 * gd_curv_metric_cal computes it, ghost mirroring included (gd_t.c:191-402). */
int cgfdh_compute_metric(cgfdh_solver *s, const float *xyz);

/* Free-surface matrices from the device metric and medium, as
 * sv_curv_col_el_{iso,vti,aniso}_dvh2dvz (sv_curv_col_el_iso.c:1259-1375,
 * sv_curv_col_el_vti.c:1085-1203, sv_curv_col_el_aniso.c:1267-1417), which
 * the reference runs at the start of its time loop (drv_rk_curv_col.c:132-159). */
int cgfdh_compute_free_surface(cgfdh_solver *s);

int cgfdh_get_metric(cgfdh_solver *s, float *metric10);
int cgfdh_get_free_surface(cgfdh_solver *s, float *mats /* 3 x [ny][nx][3][3] */);

/* ---- host table builders ------------------------------------------------ */

/* A/B/D of one PML face from the grid coordinates: bdry_pml_set with
 * bdry_cal_abl_len_dh and bdry_pml_cal_{R,dmax,d,a,b} (bdry_t.c:115-506). */
int cgfdh_pml_face_coefs(int ni, int nj, int nk, const float *xyz, int face, int num_layers,
                         float velocity, float alpha_max, float beta_max,
                         float *A, float *B, float *D);

/* fun_ricker (src_t.c:2159-2166). */
float cgfdh_ricker(float t, float fc, float t0);

/* angle2moment + src_muDA_to_moment (src_t.c:2344-2409); m in the .src order
 * Mxx,Myy,Mzz,Myz,Mxz,Mxy. */
void cgfdh_moment_from_angles(float strike, float dip, float rake, float mu, float D,
                              float A, float *m);

/* Ricker moment-rate table of one source (src_t.c:550-555, 1029-1091):
 * fills it_begin/it_end/max_nt and, if out != NULL, 6 arrays (engine order
 * Mxx,Myy,Mzz,Mxz,Myz,Mxy) of [max_nt][4] at stride out_stride floats. */
int cgfdh_ricker_moment_table(float t0, float dt, float stf_length, float wavelet_tstart,
                              float fc, float tpeak, const float *m6, int *it_begin,
                              int *it_end, int *max_nt, float *out, long out_stride);

/* GMB visco-elastic coefficients (md_vis_GMB_cal_Y, md_t.c:643-792, with the
 * least-squares fit md_visco_LS, md_t.c:795-914): relaxation frequencies wl
 * log-spaced over [fmin, fmax], Y fitted so 1/Q is flat at 1/round(Qp|Qs);
 * lambda, mu (relaxed, npts values) are replaced in place by the unrelaxed
 * moduli at the reference frequency freq; ylam_ymu gets the 2*nmaxwell
 * volumes Ylam_n, Ymu_n of npts values in cgfdh_set_medium_visco's order. */
int cgfdh_visco_gmb(long npts, int nmaxwell, float freq, float fmin, float fmax,
                    float *lambda, float *mu, const float *Qp, const float *Qs,
                    float *ylam_ymu, float *wl);

/* Normalised Gaussian source weights (src_cal_norm_delt3d_z2fre,
 * src_t.c:2111-2151): (2*half+1)^3 values, normalised over k <= half_below. */
int cgfdh_gauss_weights(float *w, float x0, float y0, float z0, float a, int half,
                        int half_below);

#ifdef __cplusplus
}
#endif

#endif
