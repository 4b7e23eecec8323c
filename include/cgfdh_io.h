/*
 * cgfdh_io.h -- output files of the time loop in the reference's formats,
 * written by the library's own code (no netCDF library needed):
 *
 *  - a netCDF-3 classic (CDF-1) writer: dimensions (one unlimited),
 *    float variables, int attributes, hyperslab puts;
 *  - the reference's output files on top of it, same names, dimensions,
 *    attributes and variable order: slices (io_slice_nc_create / _put,
 *    io_funcs.c:780-883, 1006-1108), snapshots (io_snap_nc_create / _put,
 *    io_funcs.c:885-995, 1114-1268), peak ground motion (PG_slice_output,
 *    io_funcs.c:1944-2003) and the whole-wavefield debug dump
 *    (io_var3d_export_nc + io_build_fname_time, io_funcs.c:2006-2066,
 *    1913-1941);
 *  - SAC traces as sacExport1C1R writes them (sacLib.c:8-84, with the
 *    header defaults of sacHeaderInit, sacLib.c:86-279) and the receiver
 *    files of io_recv_output_sac (io_funcs.c:1648-1697).
 *
 * Every function returns 0 on success or a negative code; the message of the
 * last failure is cgfdh_io_last_error().  Host code only.
 */
#ifndef CGFDH_IO_H
#define CGFDH_IO_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

const char *cgfdh_io_last_error(void);

/* ---- netCDF-3 classic writer --------------------------------------------- */
typedef struct cgfdh_nc cgfdh_nc;

#define CGFDH_NC_GLOBAL (-1)
#define CGFDH_NC_UNLIMITED 0

/* nc_create(path, NC_CLOBBER): a new classic file in define mode */
int cgfdh_nc_create(const char *path, cgfdh_nc **out);
/* nc_def_dim; len CGFDH_NC_UNLIMITED is the record dimension (at most one,
 * and first in any variable that uses it); *dimid receives its id */
int cgfdh_nc_def_dim(cgfdh_nc *nc, const char *name, size_t len, int *dimid);
/* nc_def_var of an NC_FLOAT variable */
int cgfdh_nc_def_var_float(cgfdh_nc *nc, const char *name, int ndims, const int *dimids,
                           int *varid);
/* nc_put_att_int (NC_INT values) on a variable or CGFDH_NC_GLOBAL */
int cgfdh_nc_put_att_int(cgfdh_nc *nc, int varid, const char *name, int n, const int *vals);
/* nc_enddef: the header is laid out and written */
int cgfdh_nc_enddef(cgfdh_nc *nc);
/* nc_put_vara_float: start/count per dimension of the variable; new records
 * are filled with the classic float fill value first, as the netCDF
 * library's default fill mode does */
int cgfdh_nc_put_vara_float(cgfdh_nc *nc, int varid, const size_t *start, const size_t *count,
                            const float *data);
int cgfdh_nc_close(cgfdh_nc *nc);

/* ---- the reference's output files ---------------------------------------- */

/* A slice file of io_slice_nc_create: axis 'x' (dims time, k, j), 'y'
 * (time, k, i) or 'z' (time, j, i); global attributes
 * "<a>_index_with_ghosts_in_this_thread" (the slice's array index) and
 * "coords_of_mpi_topo"; a "time" variable and one float variable per name. */
int cgfdh_io_slice_create(const char *path, char axis, int n1, int n2, int index_with_ghosts,
                          const int *topoid, int nvars, const char *const *names,
                          cgfdh_nc **out);
/* io_slice_nc_put for one slice file: record it, its time value, then every
 * variable's [n1][n2] plane (vals holds nvars planes back to back) */
int cgfdh_io_slice_put(cgfdh_nc *nc, int it, float time, int nvars, const float *vals);

/* A snapshot file of io_snap_nc_create: dims time, k, j, i (nk, nj, ni);
 * variables time, then Vx Vy Vz (out_vel), Txx Tyy Tzz Txz Tyz Txy
 * (out_stress), Exx Eyy Ezz Exz Eyz Exy (out_strain); global attributes
 * first_index_to_snapshot_output (g_start, 3), first_index_in_this_thread_
 * with_ghosts (l_start, 3), index_stride_in_this_thread (stride, 3) and
 * coords_of_mpi_topo (2). */
int cgfdh_io_snap_create(const char *path, int nk, int nj, int ni, int out_vel,
                         int out_stress, int out_strain, const int *g_start,
                         const int *l_start, const int *stride, const int *topoid,
                         cgfdh_nc **out);
/* io_snap_nc_put for one snapshot record: time, then the present variable
 * groups in file order, each [nk][nj][ni]: vel (3), stress (6), strain (6)
 * as given (NULL for absent groups) */
int cgfdh_io_snap_put(cgfdh_nc *nc, int record, float time, const float *vel,
                      const float *stress, const float *strain);

/* PG_slice_output: <output_dir>/PG_V_A_D_<frame_coords>.nc with dims j (ny),
 * i (nx), the 15 maps PGV..PGDz [ny][nx] and the global attributes
 * global_index_of_first_physical_points (gni1, gnj1),
 * count_index_of_physical_points (ni, nj) and coords_of_mpi_topo. */
int cgfdh_io_pg_output(const char *output_dir, const char *frame_coords, int nx, int ny,
                       int ni, int nj, int gni1, int gnj1, const int *topoid, const float *pg);

/* io_build_fname_time + io_var3d_export_nc: <output_dir>/<prefix>_<t0>_<t1>
 * _it<it>.nc (t = topoid) holding the components of w4d at the offsets
 * cmp_pos (floats) over the whole [nz][ny][nx] array; the dimensions are
 * defined in the order dim_names[0] (nx), dim_names[1] (ny), dim_names[2]
 * (nz) and every variable is [nz][ny][nx] */
int cgfdh_io_var3d_export(const char *output_dir, const char *prefix, const int *topoid,
                          int it, const float *w4d, const size_t *cmp_pos,
                          const char *const *cmp_names, int ncmp,
                          const char *const *dim_names, int nx, int ny, int nz);

/* sacExport1C1R: one SAC trace (version 6 header, native byte order) with
 * the event and station fields given, b = 0, e = delta * (npts - 1),
 * every other header field undefined (-12345). */
int cgfdh_sac_export_1c1r(const char *path, const float *seismo, float evla, float evlo,
                          float evel, float evdp, float stla, float stlo, float stel,
                          float stdp, float delta, int npts);

/* io_recv_output_sac: <output_dir>/<evtnm>.<name>.<cmp>.sac for every
 * receiver and each of the first ncomp components (9 for the GMB medium,
 * whose memory variables are not written); seismo[(r * ncomp_all + c) *
 * nt + it] as cgfdh_get_seismograms gives it, xyz the receivers'
 * coordinates (3 per receiver); stdp and delta both take dt, as the
 * reference passes them. */
int cgfdh_io_recv_output_sac(const char *output_dir, const char *evtnm, int nrecv,
                             const char *const *names, const float *xyz, const float *seismo,
                             int ncomp_all, int ncomp, const char *const *cmp_names, int nt,
                             float dt);

/* The grid layer file gd_curv_gen_layer reads (in_grid_layer_file,
 * gd_t.c:1031-1060): first its sizes, then its contents into caller arrays:
 * ncell and equal (n_interfaces - 1 each) and layer_xyz = x, then y, then z
 * of every sample ([3][n_interfaces][ny_layers][nx_layers]), the input of
 * cgfdh_gen_grid_layer. */
int cgfdh_io_grid_layer_dims(const char *path, int *n_interfaces, int *nx_layers,
                             int *ny_layers);
int cgfdh_io_read_grid_layer(const char *path, int *ncell, int *equal, float *layer_xyz);

/* The layered-model file of media_layer2model (.md3lay, read_interface_file,
 * media_read_file.cpp:22-620), every media type: first its header (the type
 * code of media_utility.hpp:37-45 -- one_component 0, acoustic_isotropic 10,
 * elastic_isotropic 21, elastic_vti_prem 22, _thomsen 23, _cij 24,
 * elastic_tti_thomsen 25, elastic_tti_bond 26, elastic_aniso_cij 27 --, the
 * properties per sample nprop (1, 2, 3, 6, 6, 6, 8, 8, 22), interfaces,
 * samples nx x ny, geo = minx, miny, dx, dy), then its samples into fields
 * ([1 + 3 nprop][n_interfaces][ny][nx]: elevation, then value, gradient and
 * power of each property in the file's order), the input of
 * cgfdh_layer2model_grid.  _dims is _info without the type. */
int cgfdh_io_md3lay_info(const char *path, int *media_type, int *nprop, int *n_interfaces,
                         int *nx, int *ny, float *geo);
int cgfdh_io_md3lay_dims(const char *path, int *nprop, int *n_interfaces, int *nx, int *ny,
                         float *geo);
int cgfdh_io_read_md3lay(const char *path, float *fields);

/* The grid-model file of media_grid2model (.md3grd, read_grid_file,
 * media_read_file.cpp:670-1320), every media type: the header (type code
 * and nprop as for .md3lay, n_layers and the surfaces of each into ngz --
 * at most ngz_cap --, n_surfaces their sum) and the samples of the window of
 * the mesh that covers [xmin, xmax] x [ymin, ymax] (the grid's extent, as
 * the reference takes it): nx x ny samples from geo = minx, miny, dx, dy,
 * the reference's cropped mesh; fields [1 + nprop][n_surfaces][ny][nx]
 * (elevation, then each property's value), the input of
 * cgfdh_grid2model_grid.  A mesh that does not cover the window is an
 * error. */
int cgfdh_io_md3grd_info(const char *path, float xmin, float xmax, float ymin, float ymax,
                         int *media_type, int *nprop, int *n_layers, int *ngz, int ngz_cap,
                         int *n_surfaces, int *nx, int *ny, float *geo);
int cgfdh_io_read_md3grd(const char *path, float xmin, float xmax, float ymin, float ymax,
                         float *fields);

/* One float32 volume of media_bin2model (read_bin_file,
 * media_read_file.cpp:1274-1320): size[0] x size[1] x size[2] samples, the
 * first axis fastest, axis a holding direction order[a] (0 x, 1 y, 2 z);
 * the samples start[a]..end[a] of each axis are kept, x fastest then y then
 * z (the layout cgfdh_bin2model_grid takes). */
int cgfdh_io_read_bin_window(const char *path, const int *order, const int *size,
                             const int *start, const int *end, float *out);

#ifdef __cplusplus
}
#endif

#endif
