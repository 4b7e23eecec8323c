/*
 * host_io.c -- the time loop's output files in the reference's formats,
 * written without the netCDF library: a netCDF-3 classic (CDF-1) writer,
 * the reference's slice / snapshot / peak-ground-motion / debug-dump files
 * on top of it, and SAC traces.  See include/cgfdh_io.h.
 *
 * netCDF-3 classic layout (the published netCDF classic format
 * specification): header = "CDF\1", numrecs, dimension list, global
 * attribute list, variable list (name, dimension ids, attributes, type,
 * vsize, begin); all integers and values big-endian, names and values padded
 * to 4 bytes; the non-record variables' data follow the header in variable
 * order, then the records, each holding one slab of every record variable
 * in variable order.
 */
#include <errno.h>
#include <math.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "cgfdh_io.h"

static char g_io_err[512];

static int io_fail(const char *fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_io_err, sizeof g_io_err, fmt, ap);
  va_end(ap);
  return -1;
}

const char *cgfdh_io_last_error(void) { return g_io_err; }

/* ---- netCDF-3 classic writer ----------------------------------------------- */

#define NC_MAXDIMS 16
#define NC_MAXVARS 64
#define NC_MAXATTS 16
#define NC_MAXNAME 128
#define NC_MAXVDIMS 8

enum { NCT_INT = 4, NCT_FLOAT = 5 };
enum { TAG_DIMENSION = 0x0A, TAG_VARIABLE = 0x0B, TAG_ATTRIBUTE = 0x0C };
/* the classic format's default float fill value, 9.9692099683868690e+36 */
#define NC_FILL_FLOAT_BITS 0x7cf00000u

typedef struct {
  char name[NC_MAXNAME];
  int n;
  int *vals;
} nc_att;

typedef struct {
  char name[NC_MAXNAME];
  size_t len;
} nc_dim;

typedef struct {
  char name[NC_MAXNAME];
  int ndims;
  int dimids[NC_MAXVDIMS];
  int natts;
  nc_att atts[NC_MAXATTS];
  int is_rec;
  size_t vsize;          /* bytes of the variable (of one record's slab for a record variable) */
  uint64_t begin;
} nc_var;

struct cgfdh_nc {
  FILE *fp;
  char path[1024];
  int defmode;
  int ndims, nvars, ngatts;
  nc_dim dims[NC_MAXDIMS];
  nc_var vars[NC_MAXVARS];
  nc_att gatts[NC_MAXATTS];
  int rec_dim;           /* -1: none */
  size_t numrecs;
  uint64_t recsize;
};

static void put_be32(unsigned char *b, uint32_t v)
{
  b[0] = (unsigned char)(v >> 24);
  b[1] = (unsigned char)(v >> 16);
  b[2] = (unsigned char)(v >> 8);
  b[3] = (unsigned char)v;
}

/* header bytes, growable */
typedef struct {
  unsigned char *p;
  size_t n, cap;
} hbuf;

static int hb_put(hbuf *h, const void *src, size_t n)
{
  if (h->n + n > h->cap) {
    size_t cap = h->cap ? 2 * h->cap : 1024;
    while (cap < h->n + n) cap *= 2;
    unsigned char *q = (unsigned char *)realloc(h->p, cap);
    if (!q) return -1;
    h->p = q;
    h->cap = cap;
  }
  memcpy(h->p + h->n, src, n);
  h->n += n;
  return 0;
}

static int hb_u32(hbuf *h, uint32_t v)
{
  unsigned char b[4];
  put_be32(b, v);
  return hb_put(h, b, 4);
}

static int hb_name(hbuf *h, const char *s)
{
  const size_t n = strlen(s);
  static const unsigned char zero[4] = {0, 0, 0, 0};
  if (hb_u32(h, (uint32_t)n) || hb_put(h, s, n)) return -1;
  return hb_put(h, zero, (4 - n % 4) % 4);
}

static int hb_atts(hbuf *h, const nc_att *atts, int natts)
{
  if (natts == 0) return hb_u32(h, 0) || hb_u32(h, 0);   /* ABSENT */
  if (hb_u32(h, TAG_ATTRIBUTE) || hb_u32(h, (uint32_t)natts)) return -1;
  for (int a = 0; a < natts; a++) {
    if (hb_name(h, atts[a].name) || hb_u32(h, NCT_INT) || hb_u32(h, (uint32_t)atts[a].n))
      return -1;
    for (int q = 0; q < atts[a].n; q++)
      if (hb_u32(h, (uint32_t)atts[a].vals[q])) return -1;
  }
  return 0;
}

/* the header with the variables' current begin offsets */
static int build_header(const cgfdh_nc *nc, hbuf *h)
{
  static const unsigned char magic[4] = {'C', 'D', 'F', 1};
  h->n = 0;
  if (hb_put(h, magic, 4) || hb_u32(h, (uint32_t)nc->numrecs)) return -1;
  if (nc->ndims == 0) {
    if (hb_u32(h, 0) || hb_u32(h, 0)) return -1;
  } else {
    if (hb_u32(h, TAG_DIMENSION) || hb_u32(h, (uint32_t)nc->ndims)) return -1;
    for (int d = 0; d < nc->ndims; d++)
      if (hb_name(h, nc->dims[d].name) || hb_u32(h, (uint32_t)nc->dims[d].len)) return -1;
  }
  if (hb_atts(h, nc->gatts, nc->ngatts)) return -1;
  if (nc->nvars == 0) return hb_u32(h, 0) || hb_u32(h, 0);
  if (hb_u32(h, TAG_VARIABLE) || hb_u32(h, (uint32_t)nc->nvars)) return -1;
  for (int v = 0; v < nc->nvars; v++) {
    const nc_var *V = nc->vars + v;
    if (hb_name(h, V->name) || hb_u32(h, (uint32_t)V->ndims)) return -1;
    for (int d = 0; d < V->ndims; d++)
      if (hb_u32(h, (uint32_t)V->dimids[d])) return -1;
    if (hb_atts(h, V->atts, V->natts) || hb_u32(h, NCT_FLOAT) ||
        hb_u32(h, (uint32_t)V->vsize) || hb_u32(h, (uint32_t)V->begin))
      return -1;
  }
  return 0;
}

static void free_atts(nc_att *a, int n)
{
  for (int q = 0; q < n; q++) free(a[q].vals);
}

static void nc_free(cgfdh_nc *nc)
{
  if (!nc) return;
  if (nc->fp) fclose(nc->fp);
  free_atts(nc->gatts, nc->ngatts);
  for (int v = 0; v < nc->nvars; v++) free_atts(nc->vars[v].atts, nc->vars[v].natts);
  free(nc);
}

int cgfdh_nc_create(const char *path, cgfdh_nc **out)
{
  if (!path || !out) return io_fail("cgfdh_nc_create: null argument");
  cgfdh_nc *nc = (cgfdh_nc *)calloc(1, sizeof *nc);
  if (!nc) return io_fail("cgfdh_nc_create: out of memory");
  nc->fp = fopen(path, "w+b");
  if (!nc->fp) {
    free(nc);
    return io_fail("cgfdh_nc_create: %s: %s", path, strerror(errno));
  }
  snprintf(nc->path, sizeof nc->path, "%s", path);
  nc->defmode = 1;
  nc->rec_dim = -1;
  *out = nc;
  return 0;
}

int cgfdh_nc_def_dim(cgfdh_nc *nc, const char *name, size_t len, int *dimid)
{
  if (!nc || !name || !nc->defmode) return io_fail("cgfdh_nc_def_dim: not in define mode");
  if (nc->ndims >= NC_MAXDIMS || strlen(name) >= NC_MAXNAME)
    return io_fail("cgfdh_nc_def_dim: too many dimensions or name too long (%s)", name);
  if (len == CGFDH_NC_UNLIMITED) {
    if (nc->rec_dim >= 0) return io_fail("cgfdh_nc_def_dim: a second unlimited dimension");
    nc->rec_dim = nc->ndims;
  } else if (len > 0x7fffffffu) {
    return io_fail("cgfdh_nc_def_dim: %s length %zu beyond the classic format", name, len);
  }
  snprintf(nc->dims[nc->ndims].name, NC_MAXNAME, "%s", name);
  nc->dims[nc->ndims].len = len;
  if (dimid) *dimid = nc->ndims;
  nc->ndims++;
  return 0;
}

int cgfdh_nc_def_var_float(cgfdh_nc *nc, const char *name, int ndims, const int *dimids,
                           int *varid)
{
  if (!nc || !name || !nc->defmode) return io_fail("cgfdh_nc_def_var_float: not in define mode");
  if (nc->nvars >= NC_MAXVARS || ndims < 0 || ndims > NC_MAXVDIMS || strlen(name) >= NC_MAXNAME)
    return io_fail("cgfdh_nc_def_var_float: bad variable %s", name);
  nc_var *V = nc->vars + nc->nvars;
  memset(V, 0, sizeof *V);
  snprintf(V->name, NC_MAXNAME, "%s", name);
  V->ndims = ndims;
  for (int d = 0; d < ndims; d++) {
    if (dimids[d] < 0 || dimids[d] >= nc->ndims)
      return io_fail("cgfdh_nc_def_var_float: %s: bad dimension id %d", name, dimids[d]);
    if (dimids[d] == nc->rec_dim && d != 0)
      return io_fail("cgfdh_nc_def_var_float: %s: the unlimited dimension must come first", name);
    V->dimids[d] = dimids[d];
  }
  V->is_rec = ndims > 0 && dimids[0] == nc->rec_dim;
  uint64_t n = 1;
  for (int d = V->is_rec ? 1 : 0; d < ndims; d++) n *= nc->dims[dimids[d]].len;
  if (4 * n > 0xffffffffull) return io_fail("cgfdh_nc_def_var_float: %s too large", name);
  V->vsize = (size_t)(4 * n);
  if (varid) *varid = nc->nvars;
  nc->nvars++;
  return 0;
}

int cgfdh_nc_put_att_int(cgfdh_nc *nc, int varid, const char *name, int n, const int *vals)
{
  if (!nc || !name || n < 0 || (n > 0 && !vals) || !nc->defmode)
    return io_fail("cgfdh_nc_put_att_int: bad argument");
  nc_att *list;
  int *cnt;
  if (varid == CGFDH_NC_GLOBAL) {
    list = nc->gatts;
    cnt = &nc->ngatts;
  } else {
    if (varid < 0 || varid >= nc->nvars) return io_fail("cgfdh_nc_put_att_int: varid %d", varid);
    list = nc->vars[varid].atts;
    cnt = &nc->vars[varid].natts;
  }
  if (*cnt >= NC_MAXATTS || strlen(name) >= NC_MAXNAME)
    return io_fail("cgfdh_nc_put_att_int: too many attributes (%s)", name);
  nc_att *a = list + *cnt;
  snprintf(a->name, NC_MAXNAME, "%s", name);
  a->n = n;
  a->vals = (int *)malloc((n > 0 ? n : 1) * sizeof(int));
  if (!a->vals) return io_fail("cgfdh_nc_put_att_int: out of memory");
  memcpy(a->vals, vals, n * sizeof(int));
  (*cnt)++;
  return 0;
}

static int write_header(cgfdh_nc *nc)
{
  hbuf h = {0, 0, 0};
  int rc = build_header(nc, &h);
  if (!rc && (fseek(nc->fp, 0, SEEK_SET) || fwrite(h.p, 1, h.n, nc->fp) != h.n)) rc = -1;
  free(h.p);
  return rc ? io_fail("%s: header write failed", nc->path) : 0;
}

int cgfdh_nc_enddef(cgfdh_nc *nc)
{
  if (!nc || !nc->defmode) return io_fail("cgfdh_nc_enddef: not in define mode");
  hbuf h = {0, 0, 0};
  if (build_header(nc, &h)) {
    free(h.p);
    return io_fail("cgfdh_nc_enddef: out of memory");
  }
  /* the data start right after the header (4-byte aligned); non-record
   * variables first, then the record section */
  uint64_t off = (h.n + 3) / 4 * 4;
  free(h.p);
  for (int v = 0; v < nc->nvars; v++)
    if (!nc->vars[v].is_rec) {
      nc->vars[v].begin = off;
      off += nc->vars[v].vsize;
    }
  nc->recsize = 0;
  for (int v = 0; v < nc->nvars; v++)
    if (nc->vars[v].is_rec) {
      nc->vars[v].begin = off + nc->recsize;
      nc->recsize += nc->vars[v].vsize;
    }
  if (off > 0xffffffffull) return io_fail("%s: beyond the classic format's offsets", nc->path);
  nc->defmode = 0;
  if (write_header(nc)) return -1;
  /* fixed-size variables hold the fill value until written */
  for (int v = 0; v < nc->nvars; v++) {
    const nc_var *V = nc->vars + v;
    if (V->is_rec) continue;
    unsigned char fill[4096];
    for (int q = 0; q < 4096; q += 4) put_be32(fill + q, NC_FILL_FLOAT_BITS);
    if (fseek(nc->fp, (long)V->begin, SEEK_SET)) return io_fail("%s: seek failed", nc->path);
    for (size_t done = 0; done < V->vsize;) {
      size_t n = V->vsize - done < 4096 ? V->vsize - done : 4096;
      if (fwrite(fill, 1, n, nc->fp) != n) return io_fail("%s: write failed", nc->path);
      done += n;
    }
  }
  return 0;
}

/* fill records numrecs .. upto-1 of every record variable */
static int grow_records(cgfdh_nc *nc, size_t upto)
{
  unsigned char fill[4096];
  for (int q = 0; q < 4096; q += 4) put_be32(fill + q, NC_FILL_FLOAT_BITS);
  for (size_t r = nc->numrecs; r < upto; r++)
    for (int v = 0; v < nc->nvars; v++) {
      const nc_var *V = nc->vars + v;
      if (!V->is_rec) continue;
      if (fseek(nc->fp, (long)(V->begin + r * nc->recsize), SEEK_SET))
        return io_fail("%s: seek failed", nc->path);
      for (size_t done = 0; done < V->vsize;) {
        size_t n = V->vsize - done < 4096 ? V->vsize - done : 4096;
        if (fwrite(fill, 1, n, nc->fp) != n) return io_fail("%s: write failed", nc->path);
        done += n;
      }
    }
  if (upto > nc->numrecs) nc->numrecs = upto;
  return 0;
}

int cgfdh_nc_put_vara_float(cgfdh_nc *nc, int varid, const size_t *start, const size_t *count,
                            const float *data)
{
  if (!nc || nc->defmode || varid < 0 || varid >= nc->nvars || !data)
    return io_fail("cgfdh_nc_put_vara_float: bad argument");
  const nc_var *V = nc->vars + varid;
  const int nd = V->ndims;
  size_t shape[NC_MAXVDIMS];
  size_t total = 1;
  for (int d = 0; d < nd; d++) {
    shape[d] = nc->dims[V->dimids[d]].len;
    const int rec = V->is_rec && d == 0;
    if (!rec && start[d] + count[d] > shape[d])
      return io_fail("cgfdh_nc_put_vara_float: %s: dimension %d range %zu+%zu beyond %zu",
                     V->name, d, start[d], count[d], shape[d]);
    total *= count[d];
  }
  if (total == 0) return 0;
  if (V->is_rec && start[0] + count[0] > nc->numrecs && grow_records(nc, start[0] + count[0]))
    return -1;
  /* element strides of the variable's own layout (a record's slab for the
   * record dimension) */
  uint64_t stride[NC_MAXVDIMS];
  uint64_t s = 1;
  for (int d = nd - 1; d >= 0; d--) {
    stride[d] = s;
    if (!(V->is_rec && d == 0)) s *= shape[d];
  }
  const size_t run = nd > 0 ? count[nd - 1] : 1;
  unsigned char *buf = (unsigned char *)malloc(4 * run);
  if (!buf) return io_fail("cgfdh_nc_put_vara_float: out of memory");
  size_t idx[NC_MAXVDIMS] = {0};
  const size_t nruns = total / run;
  const float *src = data;
  for (size_t q = 0; q < nruns; q++) {
    uint64_t off = V->begin;
    for (int d = 0; d < nd; d++) {
      const uint64_t i = start[d] + idx[d];
      if (V->is_rec && d == 0)
        off += i * nc->recsize;
      else
        off += 4 * i * stride[d];
    }
    for (size_t e = 0; e < run; e++) {
      uint32_t u;
      memcpy(&u, src + e, 4);
      put_be32(buf + 4 * e, u);
    }
    if (fseek(nc->fp, (long)off, SEEK_SET) || fwrite(buf, 4, run, nc->fp) != run) {
      free(buf);
      return io_fail("%s: write failed", nc->path);
    }
    src += run;
    for (int d = nd - 2; d >= 0; d--) {   /* next run: odometer over the outer dimensions */
      if (++idx[d] < count[d]) break;
      idx[d] = 0;
    }
  }
  free(buf);
  return 0;
}

int cgfdh_nc_close(cgfdh_nc *nc)
{
  if (!nc) return 0;
  int rc = 0;
  if (nc->defmode) rc = cgfdh_nc_enddef(nc);
  if (!rc) rc = write_header(nc);   /* the final record count */
  if (!rc && fflush(nc->fp)) rc = io_fail("%s: flush failed", nc->path);
  nc_free(nc);
  return rc;
}

/* ---- the reference's output files ------------------------------------------ */

int cgfdh_io_slice_create(const char *path, char axis, int n1, int n2, int index_with_ghosts,
                          const int *topoid, int nvars, const char *const *names,
                          cgfdh_nc **out)
{
  /* io_slice_nc_create (io_funcs.c:780-883) */
  const char *d1, *d2, *att;
  if (axis == 'x') { d1 = "k"; d2 = "j"; att = "i_index_with_ghosts_in_this_thread"; }
  else if (axis == 'y') { d1 = "k"; d2 = "i"; att = "j_index_with_ghosts_in_this_thread"; }
  else if (axis == 'z') { d1 = "j"; d2 = "i"; att = "k_index_with_ghosts_in_this_thread"; }
  else return io_fail("cgfdh_io_slice_create: axis %c", axis);
  cgfdh_nc *nc = NULL;
  int dimid[3], varid;
  int rc = cgfdh_nc_create(path, &nc);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "time", CGFDH_NC_UNLIMITED, &dimid[0]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, d1, (size_t)n1, &dimid[1]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, d2, (size_t)n2, &dimid[2]);
  if (!rc) rc = cgfdh_nc_def_var_float(nc, "time", 1, dimid, &varid);
  for (int v = 0; v < nvars && !rc; v++) rc = cgfdh_nc_def_var_float(nc, names[v], 3, dimid, &varid);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, att, 1, &index_with_ghosts);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "coords_of_mpi_topo", 2, topoid);
  if (!rc) rc = cgfdh_nc_enddef(nc);
  if (rc) {
    nc_free(nc);
    return rc;
  }
  *out = nc;
  return 0;
}

int cgfdh_io_slice_put(cgfdh_nc *nc, int it, float time, int nvars, const float *vals)
{
  /* io_slice_nc_put (io_funcs.c:1006-1108): the time value, then each
   * variable's plane at record it */
  if (!nc || nvars != nc->nvars - 1) return io_fail("cgfdh_io_slice_put: %d variables", nvars);
  const size_t n1 = nc->dims[1].len, n2 = nc->dims[2].len;
  size_t t0 = (size_t)it, t1 = 1;
  int rc = cgfdh_nc_put_vara_float(nc, 0, &t0, &t1, &time);
  const size_t start[3] = {(size_t)it, 0, 0}, count[3] = {1, n1, n2};
  for (int v = 0; v < nvars && !rc; v++)
    rc = cgfdh_nc_put_vara_float(nc, 1 + v, start, count, vals + (size_t)v * n1 * n2);
  return rc;
}

int cgfdh_io_snap_create(const char *path, int nk, int nj, int ni, int out_vel,
                         int out_stress, int out_strain, const int *g_start,
                         const int *l_start, const int *stride, const int *topoid,
                         cgfdh_nc **out)
{
  /* io_snap_nc_create (io_funcs.c:885-995) */
  static const char *vn[3] = {"Vx", "Vy", "Vz"};
  static const char *tn[6] = {"Txx", "Tyy", "Tzz", "Txz", "Tyz", "Txy"};
  static const char *en[6] = {"Exx", "Eyy", "Ezz", "Exz", "Eyz", "Exy"};
  cgfdh_nc *nc = NULL;
  int dimid[4], varid;
  int rc = cgfdh_nc_create(path, &nc);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "time", CGFDH_NC_UNLIMITED, &dimid[0]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "k", (size_t)nk, &dimid[1]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "j", (size_t)nj, &dimid[2]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "i", (size_t)ni, &dimid[3]);
  if (!rc) rc = cgfdh_nc_def_var_float(nc, "time", 1, dimid, &varid);
  for (int q = 0; q < 3 && out_vel && !rc; q++) rc = cgfdh_nc_def_var_float(nc, vn[q], 4, dimid, &varid);
  for (int q = 0; q < 6 && out_stress && !rc; q++)
    rc = cgfdh_nc_def_var_float(nc, tn[q], 4, dimid, &varid);
  for (int q = 0; q < 6 && out_strain && !rc; q++)
    rc = cgfdh_nc_def_var_float(nc, en[q], 4, dimid, &varid);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "first_index_to_snapshot_output", 3,
                                     g_start);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL,
                                     "first_index_in_this_thread_with_ghosts", 3, l_start);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "index_stride_in_this_thread", 3,
                                     stride);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "coords_of_mpi_topo", 2, topoid);
  if (!rc) rc = cgfdh_nc_enddef(nc);
  if (rc) {
    nc_free(nc);
    return rc;
  }
  *out = nc;
  return 0;
}

int cgfdh_io_snap_put(cgfdh_nc *nc, int record, float time, const float *vel,
                      const float *stress, const float *strain)
{
  /* io_snap_nc_put (io_funcs.c:1114-1268) for one record: variables in
   * file order after "time", each group given or skipped */
  if (!nc) return io_fail("cgfdh_io_snap_put: null file");
  const size_t nk = nc->dims[1].len, nj = nc->dims[2].len, ni = nc->dims[3].len;
  const size_t npt = nk * nj * ni;
  size_t t0 = (size_t)record, t1 = 1;
  int rc = cgfdh_nc_put_vara_float(nc, 0, &t0, &t1, &time);
  const size_t start[4] = {(size_t)record, 0, 0, 0}, count[4] = {1, nk, nj, ni};
  int v = 1;
  const float *groups[3] = {vel, stress, strain};
  for (int g = 0; g < 3 && !rc; g++) {
    /* a group is in the file if its first variable's name starts with V/T/E */
    const char lead = g == 0 ? 'V' : (g == 1 ? 'T' : 'E');
    if (v >= nc->nvars || nc->vars[v].name[0] != lead) continue;
    const int nvg = g == 0 ? 3 : 6;
    if (!groups[g]) {
      v += nvg;
      continue;
    }
    for (int q = 0; q < nvg && !rc; q++, v++)
      rc = cgfdh_nc_put_vara_float(nc, v, start, count, groups[g] + (size_t)q * npt);
  }
  return rc;
}

int cgfdh_io_pg_output(const char *output_dir, const char *frame_coords, int nx, int ny,
                       int ni, int nj, int gni1, int gnj1, const int *topoid, const float *pg)
{
  /* PG_slice_output (io_funcs.c:1944-2003) */
  static const char *names[15] = {"PGV", "PGVh", "PGVx", "PGVy", "PGVz", "PGA", "PGAh", "PGAx",
                                  "PGAy", "PGAz", "PGD", "PGDh", "PGDx", "PGDy", "PGDz"};
  char path[2048];
  snprintf(path, sizeof path, "%s/%s_%s.nc", output_dir, "PG_V_A_D", frame_coords);
  cgfdh_nc *nc = NULL;
  int dimid[2], varid[15];
  int rc = cgfdh_nc_create(path, &nc);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "j", (size_t)ny, &dimid[0]);
  if (!rc) rc = cgfdh_nc_def_dim(nc, "i", (size_t)nx, &dimid[1]);
  for (int q = 0; q < 15 && !rc; q++) rc = cgfdh_nc_def_var_float(nc, names[q], 2, dimid, &varid[q]);
  const int g_start[2] = {gni1, gnj1}, phy[2] = {ni, nj};
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "global_index_of_first_physical_points",
                                     2, g_start);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "count_index_of_physical_points", 2,
                                     phy);
  if (!rc) rc = cgfdh_nc_put_att_int(nc, CGFDH_NC_GLOBAL, "coords_of_mpi_topo", 2, topoid);
  if (!rc) rc = cgfdh_nc_enddef(nc);
  const size_t start[2] = {0, 0}, count[2] = {(size_t)ny, (size_t)nx};
  for (int q = 0; q < 15 && !rc; q++)
    rc = cgfdh_nc_put_vara_float(nc, varid[q], start, count, pg + (size_t)q * nx * ny);
  if (rc) {
    nc_free(nc);
    return rc;
  }
  return cgfdh_nc_close(nc);
}

int cgfdh_io_var3d_export(const char *output_dir, const char *prefix, const int *topoid,
                          int it, const float *w4d, const size_t *cmp_pos,
                          const char *const *cmp_names, int ncmp,
                          const char *const *dim_names, int nx, int ny, int nz)
{
  /* io_build_fname_time (io_funcs.c:46-55) and io_var3d_export_nc
   * (io_funcs.c:121-163): dimensions defined x, y, z; variables [z][y][x] */
  char path[2048];
  snprintf(path, sizeof path, "%s/%s_%d_%d_it%d%s", output_dir, prefix, topoid[0], topoid[1], it,
           ".nc");
  cgfdh_nc *nc = NULL;
  int dx, dy, dz, varid;
  int rc = cgfdh_nc_create(path, &nc);
  if (!rc) rc = cgfdh_nc_def_dim(nc, dim_names[0], (size_t)nx, &dx);
  if (!rc) rc = cgfdh_nc_def_dim(nc, dim_names[1], (size_t)ny, &dy);
  if (!rc) rc = cgfdh_nc_def_dim(nc, dim_names[2], (size_t)nz, &dz);
  const int dimid[3] = {dz, dy, dx};
  for (int v = 0; v < ncmp && !rc; v++) rc = cgfdh_nc_def_var_float(nc, cmp_names[v], 3, dimid, &varid);
  if (!rc) rc = cgfdh_nc_enddef(nc);
  const size_t start[3] = {0, 0, 0}, count[3] = {(size_t)nz, (size_t)ny, (size_t)nx};
  for (int v = 0; v < ncmp && !rc; v++)
    rc = cgfdh_nc_put_vara_float(nc, v, start, count, w4d + cmp_pos[v]);
  if (rc) {
    nc_free(nc);
    return rc;
  }
  return cgfdh_nc_close(nc);
}

/* ---- SAC ------------------------------------------------------------------- */

/* The SAC version 6 header: 70 floats, 40 ints, 192 characters (8-character
 * fields, kevnm 16) in the format's field order.  Indices of the fields set
 * here. */
enum {
  SAC_DELTA = 0, SAC_B = 5, SAC_E = 6, SAC_STLA = 31, SAC_STLO = 32, SAC_STEL = 33,
  SAC_STDP = 34, SAC_EVLA = 35, SAC_EVLO = 36, SAC_EVEL = 37, SAC_EVDP = 38,
  SAC_CMPAZ = 57, SAC_CMPINC = 58
};
enum { SAC_NVHDR = 6, SAC_NPTS = 9, SAC_IFTYPE = 15, SAC_LEVEN = 35 };

int cgfdh_sac_export_1c1r(const char *path, const float *seismo, float evla, float evlo,
                          float evel, float evdp, float stla, float stlo, float stel,
                          float stdp, float delta, int npts)
{
  /* sacExport1C1R (sacLib.c:8-84): header defaults of sacHeaderInit, then
   * delta, b = 0, e = delta * (npts - 1), nvhdr 6, npts, iftype ITIME (1),
   * leven 1, the event and station fields, cmpaz = cmpinc = 0, kstnm and
   * kcmpnm "-12345" (NUL-padded) */
  if (!path || !seismo || npts < 0) return io_fail("cgfdh_sac_export_1c1r: bad argument");
  float hf[70];
  int32_t hi[40];
  char hc[192];
  for (int q = 0; q < 70; q++) hf[q] = -12345.0f;
  for (int q = 0; q < 40; q++) hi[q] = -12345;
  for (int q = 0; q < 24; q++) memcpy(hc + 8 * q, "-12345  ", 8);
  memset(hc + 16, 0, 8);   /* kevnm is 16 wide: the 8 undefined characters, then NULs */
  hf[SAC_DELTA] = delta;
  hf[SAC_B] = 0.0f;
  hf[SAC_E] = delta * (float)(npts - 1);
  hf[SAC_STLA] = stla; hf[SAC_STLO] = stlo; hf[SAC_STEL] = stel; hf[SAC_STDP] = stdp;
  hf[SAC_EVLA] = evla; hf[SAC_EVLO] = evlo; hf[SAC_EVEL] = evel; hf[SAC_EVDP] = evdp;
  hf[SAC_CMPAZ] = 0.0f;
  hf[SAC_CMPINC] = 0.0f;
  hi[SAC_NVHDR] = 6;
  hi[SAC_NPTS] = npts;
  hi[SAC_IFTYPE] = 1;
  hi[SAC_LEVEN] = 1;
  /* kstnm (the first 8 characters) and kcmpnm (the 21st 8-character slot
   * counting kevnm as two): "-12345" then NULs */
  memset(hc, 0, 8);
  memcpy(hc, "-12345", 6);
  memset(hc + 160, 0, 8);
  memcpy(hc + 160, "-12345", 6);
  FILE *fp = fopen(path, "wb");
  if (!fp) return io_fail("cgfdh_sac_export_1c1r: %s: %s", path, strerror(errno));
  int ok = fwrite(hf, 4, 70, fp) == 70 && fwrite(hi, 4, 40, fp) == 40 &&
           fwrite(hc, 1, 192, fp) == 192 && fwrite(seismo, 4, (size_t)npts, fp) == (size_t)npts;
  ok = (fclose(fp) == 0) && ok;
  return ok ? 0 : io_fail("cgfdh_sac_export_1c1r: %s: write failed", path);
}

int cgfdh_io_recv_output_sac(const char *output_dir, const char *evtnm, int nrecv,
                             const char *const *names, const float *xyz, const float *seismo,
                             int ncomp_all, int ncomp, const char *const *cmp_names, int nt,
                             float dt)
{
  /* io_recv_output_sac (io_funcs.c:1648-1697): event fields 0 */
  if (ncomp > ncomp_all) return io_fail("cgfdh_io_recv_output_sac: ncomp %d > %d", ncomp, ncomp_all);
  char path[2048];
  for (int r = 0; r < nrecv; r++)
    for (int c = 0; c < ncomp; c++) {
      snprintf(path, sizeof path, "%s/%s.%s.%s.sac", output_dir, evtnm, names[r], cmp_names[c]);
      int rc = cgfdh_sac_export_1c1r(path, seismo + ((size_t)r * ncomp_all + c) * nt, 0.0f, 0.0f,
                                     0.0f, 0.0f, xyz[3 * r], xyz[3 * r + 1], xyz[3 * r + 2], dt,
                                     dt, nt);
      if (rc) return rc;
    }
  return 0;
}

/* ---- the grid layer file of gd_curv_gen_layer (gd_t.c:1031-1060) ---------
 * whitespace-separated text: the number of interfaces n; n-1 cell counts;
 * n-1 equal-spacing flags; nx and ny of the samples; then x y z of every
 * sample, x fastest, then y, then interface (bottom first). */
static int layer_head(FILE *fp, const char *path, int *n_if, int *ncell, int *equal, int *nx,
                      int *ny)
{
  if (fscanf(fp, "%d", n_if) != 1 || *n_if < 2)
    return io_fail("grid layer file %s: bad number of interfaces", path);
  for (int i = 0; i < *n_if - 1; i++) {
    int v;
    if (fscanf(fp, "%d", &v) != 1) return io_fail("grid layer file %s: cell counts", path);
    if (ncell) ncell[i] = v;
  }
  for (int i = 0; i < *n_if - 1; i++) {
    int v;
    if (fscanf(fp, "%d", &v) != 1) return io_fail("grid layer file %s: spacing flags", path);
    if (equal) equal[i] = v;
  }
  if (fscanf(fp, "%d %d", nx, ny) != 2 || *nx < 1 || *ny < 1)
    return io_fail("grid layer file %s: sample counts", path);
  return 0;
}

int cgfdh_io_grid_layer_dims(const char *path, int *n_interfaces, int *nx_layers,
                             int *ny_layers)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return io_fail("cgfdh_io_grid_layer_dims: %s: %s", path, strerror(errno));
  int rc = layer_head(fp, path, n_interfaces, NULL, NULL, nx_layers, ny_layers);
  fclose(fp);
  return rc;
}

int cgfdh_io_read_grid_layer(const char *path, int *ncell, int *equal, float *layer_xyz)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return io_fail("cgfdh_io_read_grid_layer: %s: %s", path, strerror(errno));
  int n_if, nx, ny;
  int rc = layer_head(fp, path, &n_if, ncell, equal, &nx, &ny);
  const size_t n = (size_t)nx * ny * n_if;
  for (size_t i = 0; i < n && !rc; i++)
    if (fscanf(fp, "%f %f %f", &layer_xyz[i], &layer_xyz[i + n], &layer_xyz[i + 2 * n]) != 3)
      rc = io_fail("grid layer file %s: sample %zu of %zu", path, i, n);
  fclose(fp);
  return rc;
}

/* ---- the layered-model file of media_layer2model (.md3lay) ---------------
 * read_interface_file (media_read_file.cpp:22-420): lines starting with '#'
 * or empty are skipped; then the media type, the number of interfaces, the
 * mesh (nx ny minx miny dx dy) and per interface, per sample (x fastest) the
 * elevation followed by value, gradient and power of each property of the
 * type (media_read_file.cpp:110-620; the type names and codes of
 * media_utility.hpp:37-45). */
typedef struct {
  char *text;
  char *pos;
} md_tokens;

static int md_load(const char *path, md_tokens *t)
{
  FILE *fp = fopen(path, "r");
  if (!fp) return io_fail("md3lay file %s: %s", path, strerror(errno));
  size_t cap = 1 << 16, n = 0;
  char *buf = malloc(cap), line[8192];
  if (!buf) { fclose(fp); return io_fail("md3lay file %s: out of memory", path); }
  while (fgets(line, sizeof line, fp)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    size_t l = strlen(line);
    if (n + l + 2 > cap) {
      while (n + l + 2 > cap) cap *= 2;
      char *nb = realloc(buf, cap);
      if (!nb) { free(buf); fclose(fp); return io_fail("md3lay file %s: out of memory", path); }
      buf = nb;
    }
    memcpy(buf + n, line, l);
    n += l;
    buf[n++] = ' ';
  }
  buf[n] = 0;
  fclose(fp);
  t->text = t->pos = buf;
  return 0;
}

static int md_next(md_tokens *t, char *tok, size_t cap)
{
  while (*t->pos && (*t->pos == ' ' || *t->pos == '\t' || *t->pos == '\n' || *t->pos == '\r'))
    t->pos++;
  if (!*t->pos) return 0;
  size_t k = 0;
  while (*t->pos && !(*t->pos == ' ' || *t->pos == '\t' || *t->pos == '\n' || *t->pos == '\r')) {
    if (k + 1 < cap) tok[k++] = *t->pos;
    t->pos++;
  }
  tok[k] = 0;
  return 1;
}

static const struct {
  const char *name;
  int code, nprop;
} md_types[] = {{"one_component", 0, 1},        {"acoustic_isotropic", 10, 2},
                {"elastic_isotropic", 21, 3},    {"elastic_vti_prem", 22, 6},
                {"elastic_vti_thomsen", 23, 6},  {"elastic_vti_cij", 24, 6},
                {"elastic_tti_thomsen", 25, 8},  {"elastic_tti_bond", 26, 8},
                {"elastic_aniso_cij", 27, 22}};

static int md_head(const char *path, md_tokens *t, int *mtype, int *nprop, int *n_if, int *mx,
                   int *my, float *geo)
{
  char tok[256];
  if (!md_next(t, tok, sizeof tok)) return io_fail("md3lay file %s: no media type", path);
  *nprop = 0;
  for (size_t q = 0; q < sizeof md_types / sizeof md_types[0]; q++)
    if (strcmp(tok, md_types[q].name) == 0) {
      *mtype = md_types[q].code;
      *nprop = md_types[q].nprop;
    }
  if (!*nprop) return io_fail("md3lay file %s: media type %s is not supported", path, tok);
  int iv[3];
  for (int k = 0; k < 3; k++) {
    char *e;
    if (!md_next(t, tok, sizeof tok)) return io_fail("md3lay file %s: short header", path);
    iv[k] = (int)strtol(tok, &e, 10);
  }
  *n_if = iv[0]; *mx = iv[1]; *my = iv[2];
  for (int k = 0; k < 4; k++) {
    if (!md_next(t, tok, sizeof tok)) return io_fail("md3lay file %s: short header", path);
    geo[k] = strtof(tok, NULL);
  }
  if (*n_if < 1 || *mx < 2 || *my < 2)
    return io_fail("md3lay file %s: %d interfaces of %d x %d samples", path, *n_if, *mx, *my);
  return 0;
}

int cgfdh_io_md3lay_info(const char *path, int *media_type, int *nprop, int *n_interfaces,
                         int *nx, int *ny, float *geo)
{
  md_tokens t;
  int rc = md_load(path, &t);
  if (rc) return rc;
  rc = md_head(path, &t, media_type, nprop, n_interfaces, nx, ny, geo);
  free(t.text);
  return rc;
}

int cgfdh_io_md3lay_dims(const char *path, int *nprop, int *n_interfaces, int *nx, int *ny,
                         float *geo)
{
  int mt;
  return cgfdh_io_md3lay_info(path, &mt, nprop, n_interfaces, nx, ny, geo);
}

int cgfdh_io_read_md3lay(const char *path, float *fields)
{
  md_tokens t;
  int rc = md_load(path, &t);
  if (rc) return rc;
  int mt, nprop, n_if, mx, my;
  float geo[4];
  rc = md_head(path, &t, &mt, &nprop, &n_if, &mx, &my, geo);
  const int nf = 1 + 3 * nprop;
  const size_t vol = (size_t)n_if * mx * my;
  char tok[256];
  for (size_t p = 0; p < vol && !rc; p++)
    for (int f = 0; f < nf && !rc; f++) {
      if (!md_next(&t, tok, sizeof tok))
        rc = io_fail("md3lay file %s: insufficient data at sample %zu", path, p);
      else
        fields[f * vol + p] = strtof(tok, NULL);
    }
  free(t.text);
  return rc;
}

/* ---- the grid-model file of media_grid2model (.md3grd) -------------------
 * read_grid_file (media_read_file.cpp:670-1320): the media type, the number
 * of layers NL and the surfaces of each, the mesh (NX NY MINX MINY DX DY),
 * then per surface, per sample (x fastest) the elevation followed by the
 * property values.  Only the samples of the window covering [xmin, xmax] x
 * [ymin, ymax] are kept: ix0 = (xmin - MINX) / DX truncated, ix1 =
 * ceil((xmax - MINX) / DX), the kept mesh starting at ix0 * DX + MINX (the
 * reference's float arithmetic, so interpolation on the kept mesh rounds
 * as the reference's does).  A window reaching past the last sample is an
 * error (the reference would read samples it never filled). */
#define MG_MAX_LAYERS 4096

static int mg_head(const char *path, md_tokens *t, float xmin, float xmax, float ymin, float ymax,
                   int *mtype, int *nprop, int *nl, int *ngz, int ngz_cap, int *ni, int *full,
                   int *win, float *geo)
{
  char tok[256];
  if (!md_next(t, tok, sizeof tok)) return io_fail("md3grd file %s: no media type", path);
  *nprop = 0;
  for (size_t q = 0; q < sizeof md_types / sizeof md_types[0]; q++)
    if (strcmp(tok, md_types[q].name) == 0) {
      *mtype = md_types[q].code;
      *nprop = md_types[q].nprop;
    }
  if (!*nprop) return io_fail("md3grd file %s: media type %s is not supported", path, tok);
  if (!md_next(t, tok, sizeof tok)) return io_fail("md3grd file %s: short header", path);
  *nl = (int)strtol(tok, NULL, 10);
  if (*nl < 1) return io_fail("md3grd file %s: %d layers", path, *nl);
  if (ngz && *nl > ngz_cap)
    return io_fail("md3grd file %s: %d layers, room for %d", path, *nl, ngz_cap);
  *ni = 0;
  for (int l = 0; l < *nl; l++) {
    if (!md_next(t, tok, sizeof tok)) return io_fail("md3grd file %s: short header", path);
    const int g = (int)strtol(tok, NULL, 10);
    if (g < 1) return io_fail("md3grd file %s: layer %d has %d surfaces", path, l, g);
    if (ngz) ngz[l] = g;
    *ni += g;
  }
  int iv[2];
  float fv[4];
  for (int k = 0; k < 2; k++) {
    if (!md_next(t, tok, sizeof tok)) return io_fail("md3grd file %s: short header", path);
    iv[k] = (int)strtol(tok, NULL, 10);
  }
  for (int k = 0; k < 4; k++) {
    if (!md_next(t, tok, sizeof tok)) return io_fail("md3grd file %s: short header", path);
    fv[k] = strtof(tok, NULL);
  }
  const int NX = iv[0], NY = iv[1];
  const float MINX = fv[0], MINY = fv[1], DX = fv[2], DY = fv[3];
  if (NX < 2 || NY < 2) return io_fail("md3grd file %s: %d x %d samples", path, NX, NY);
  const float fx0 = (xmin - MINX) / DX, fy0 = (ymin - MINY) / DY;
  const double cx1 = ceil((xmax - MINX) / DX), cy1 = ceil((ymax - MINY) / DY);
  if (fx0 < 0 || cx1 > NX - 1)
    return io_fail("md3grd file %s: the media mesh does not cover the grid in x", path);
  if (fy0 < 0 || cy1 > NY - 1)
    return io_fail("md3grd file %s: the media mesh does not cover the grid in y", path);
  const size_t ix0 = (size_t)fx0, iy0 = (size_t)fy0, ix1 = (size_t)cx1, iy1 = (size_t)cy1;
  full[0] = NX;
  full[1] = NY;
  win[0] = (int)ix0; win[1] = (int)ix1; win[2] = (int)iy0; win[3] = (int)iy1;
  geo[0] = ix0 * DX + MINX;
  geo[1] = iy0 * DY + MINY;
  geo[2] = DX;
  geo[3] = DY;
  return 0;
}

int cgfdh_io_md3grd_info(const char *path, float xmin, float xmax, float ymin, float ymax,
                         int *media_type, int *nprop, int *n_layers, int *ngz, int ngz_cap,
                         int *n_surfaces, int *nx, int *ny, float *geo)
{
  md_tokens t;
  int rc = md_load(path, &t);
  if (rc) return rc;
  int full[2], win[4];
  rc = mg_head(path, &t, xmin, xmax, ymin, ymax, media_type, nprop, n_layers, ngz, ngz_cap,
               n_surfaces, full, win, geo);
  if (!rc) {
    *nx = win[1] - win[0] + 1;
    *ny = win[3] - win[2] + 1;
  }
  free(t.text);
  return rc;
}

int cgfdh_io_read_md3grd(const char *path, float xmin, float xmax, float ymin, float ymax,
                         float *fields)
{
  md_tokens t;
  int rc = md_load(path, &t);
  if (rc) return rc;
  int mt, nprop, nl, ni, full[2], win[4];
  int ngz[MG_MAX_LAYERS];
  float geo[4];
  rc = mg_head(path, &t, xmin, xmax, ymin, ymax, &mt, &nprop, &nl, ngz, MG_MAX_LAYERS, &ni, full,
               win, geo);
  const int nf = 1 + nprop, nx = win[1] - win[0] + 1, ny = win[3] - win[2] + 1;
  const size_t vol = (size_t)ni * nx * ny;
  char tok[256];
  for (int s = 0; s < ni && !rc; s++)
    for (int iy = 0; iy < full[1] && !rc; iy++)
      for (int ix = 0; ix < full[0] && !rc; ix++) {
        const int keep = ix >= win[0] && ix <= win[1] && iy >= win[2] && iy <= win[3];
        const size_t p = (size_t)s * nx * ny + (size_t)(iy - win[2]) * nx + (ix - win[0]);
        for (int f = 0; f < nf && !rc; f++) {
          if (!md_next(&t, tok, sizeof tok))
            rc = io_fail("md3grd file %s: insufficient data at surface %d sample (%d, %d)", path,
                         s, ix, iy);
          else if (keep)
            fields[f * vol + p] = strtof(tok, NULL);
        }
      }
  free(t.text);
  /* checkGridData (media_read_file.cpp:1242-1265): each layer's last
   * surface lies on the next layer's first */
  if (!rc && nl > 1) {
    const size_t sl = (size_t)nx * ny;
    int s = 0;
    for (int l = 0; l < nl - 1 && !rc; l++) {
      s += ngz[l];
      for (size_t q = 0; q < sl && !rc; q++)
        if (fields[(size_t)(s - 1) * sl + q] != fields[(size_t)s * sl + q])
          rc = io_fail("md3grd file %s: the last surface of layer %d is not the first of layer %d",
                       path, l, l + 1);
    }
  }
  return rc;
}

/* ---- binary volumes of media_bin2model ------------------------------------
 * read_bin_file (media_read_file.cpp:1274-1320): size[0] x size[1] x size[2]
 * float32 samples, the first axis fastest; axis a holds the grid direction
 * order[a] (0 x, 1 y, 2 z).  The samples with start[a] <= i[a] <= end[a]
 * are kept, laid out x fastest, then y, then z: out[(ix - sx) + (iy - sy) nx
 * + (iz - sz) nx ny] (the reference always keeps every z sample, sz = 0). */
int cgfdh_io_read_bin_window(const char *path, const int *order, const int *size,
                             const int *start, const int *end, float *out)
{
  int dim[3] = {-1, -1, -1};
  for (int a = 0; a < 3; a++) {
    if (order[a] < 0 || order[a] > 2)
      return io_fail("binary model %s: bin_order[%d] = %d (0, 1 or 2)", path, a, order[a]);
    dim[order[a]] = a;
  }
  if (dim[0] < 0 || dim[1] < 0 || dim[2] < 0)
    return io_fail("binary model %s: bin_order misses a direction", path);
  for (int a = 0; a < 3; a++)
    if (size[a] < 1 || start[a] < 0 || end[a] < start[a] || end[a] >= size[a])
      return io_fail("binary model %s: window %d..%d of axis %d (%d samples)", path, start[a],
                     end[a], a, size[a]);
  FILE *fp = fopen(path, "rb");
  if (!fp) return io_fail("binary model %s: %s", path, strerror(errno));
  const size_t nx = end[dim[0]] - start[dim[0]] + 1, ny = end[dim[1]] - start[dim[1]] + 1;
  const size_t line = nx, slice = nx * ny;
  float *row = malloc((size_t)size[0] * sizeof(float));
  if (!row) { fclose(fp); return io_fail("binary model %s: out of memory", path); }
  int rc = 0;
  int i[3];
  for (i[2] = 0; i[2] < size[2] && !rc; i[2]++)
    for (i[1] = 0; i[1] < size[1] && !rc; i[1]++) {
      if (fread(row, sizeof(float), (size_t)size[0], fp) != (size_t)size[0]) {
        rc = io_fail("binary model %s: insufficient data", path);
        break;
      }
      if (i[1] < start[1] || i[1] > end[1] || i[2] < start[2] || i[2] > end[2]) continue;
      for (i[0] = start[0]; i[0] <= end[0]; i[0]++) {
        const size_t p = (size_t)(i[dim[0]] - start[dim[0]]) +
                         (size_t)(i[dim[1]] - start[dim[1]]) * line +
                         (size_t)(i[dim[2]] - start[dim[2]]) * slice;
        out[p] = row[i[0]];
      }
    }
  free(row);
  fclose(fp);
  return rc;
}
