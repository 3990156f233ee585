/* vfio.c -- field output of the host layer: the reference's PETSc binary and
 * VTK writers (FieldsBinaryWrite / FieldsVTKWrite, VFCommon.c:1353-1460) for
 * the fields of this path.
 *
 * PETSc binary: VecView with a binary viewer writes, per vector, the class id
 * VEC_FILE_CLASSID = 1211214 and the length as 32-bit integers, then the
 * values as doubles, all big-endian, in the DMDA natural order -- the order
 * the arrays here already have.  A sequence of vectors is a sequence of such
 * records, read back in the same order (VecLoad, Utils/vtkexport.c:94-110).
 *
 * VTK: an XML structured grid (.vts, ASCII), one array per component named
 * as VecViewVTKDof names them (VFCommon.c:1383-1406: "<name>_<c>" for
 * vectors with several components). */
#include <errno.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/vfhost.h"

static void put_be32(FILE *f, int32_t v) {
  const uint32_t u = (uint32_t)v;
  const unsigned char b[4] = {(unsigned char)(u >> 24), (unsigned char)(u >> 16), (unsigned char)(u >> 8),
                              (unsigned char)u};
  fwrite(b, 1, 4, f);
}
static void put_be64d(FILE *f, double d) {
  uint64_t u;
  memcpy(&u, &d, 8);
  unsigned char b[8];
  for (int i = 0; i < 8; i++) b[i] = (unsigned char)(u >> (56 - 8 * i));
  fwrite(b, 1, 8, f);
}

int vfh_petsc_vec_write(const char *file, int nvec, const double *const *vec, const int64_t *len, int append) {
  if (!file || nvec < 0 || (nvec && (!vec || !len))) return VFG_E_ARG;
  for (int v = 0; v < nvec; v++)
    if (len[v] < 0 || len[v] > INT32_MAX) return VFG_E_ARG;
  FILE *f = fopen(file, append ? "ab" : "wb");
  if (!f) return VFG_E_ARG;
  for (int v = 0; v < nvec; v++) {
    put_be32(f, 1211214);  /* VEC_FILE_CLASSID */
    put_be32(f, (int32_t)len[v]);
    for (int64_t i = 0; i < len[v]; i++) put_be64d(f, vec[v] ? vec[v][i] : 0.);
  }
  const int bad = ferror(f);
  if (fclose(f) || bad) return VFG_E_ARG;
  return VFG_OK;
}

static void vts_array(FILE *f, const char *name, const double *a, int64_t n) {
  fprintf(f, "        <DataArray type=\"Float64\" Name=\"%s\" format=\"ascii\">\n", name);
  for (int64_t i = 0; i < n; i++) fprintf(f, "%.17g%c", a ? a[i] : 0., (i % 8 == 7 || i == n - 1) ? '\n' : ' ');
  fprintf(f, "        </DataArray>\n");
}

int vfh_vts_write(const char *file, int nx, int ny, int nz, double lx, double ly, double lz, int npt,
                  const char *const *pnames, const double *const *pdata, int ncell, const char *const *cnames,
                  const double *const *cdata) {
  if (!file || nx < 2 || ny < 2 || nz < 2 || npt < 0 || ncell < 0) return VFG_E_ARG;
  FILE *f = fopen(file, "w");
  if (!f) return VFG_E_ARG;
  const int64_t nn = (int64_t)nx * ny * nz, ne = (int64_t)(nx - 1) * (ny - 1) * (nz - 1);
  fprintf(f, "<?xml version=\"1.0\"?>\n<VTKFile type=\"StructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n");
  fprintf(f, "  <StructuredGrid WholeExtent=\"0 %d 0 %d 0 %d\">\n    <Piece Extent=\"0 %d 0 %d 0 %d\">\n", nx - 1,
          ny - 1, nz - 1, nx - 1, ny - 1, nz - 1);
  /* node coordinates X[i] = i l / (n - 1) (VFCommon.c:393-397) */
  fprintf(f, "      <Points>\n        <DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n");
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++)
        fprintf(f, "%.17g %.17g %.17g\n", i == 0 ? 0. : i * lx / (nx - 1.), j == 0 ? 0. : j * ly / (ny - 1.),
                k == 0 ? 0. : k * lz / (nz - 1.));
  fprintf(f, "        </DataArray>\n      </Points>\n      <PointData>\n");
  for (int q = 0; q < npt; q++) vts_array(f, pnames[q], pdata[q], nn);
  fprintf(f, "      </PointData>\n      <CellData>\n");
  for (int q = 0; q < ncell; q++) vts_array(f, cnames[q], cdata[q], ne);
  fprintf(f, "      </CellData>\n    </Piece>\n  </StructuredGrid>\n</VTKFile>\n");
  const int bad = ferror(f);
  if (fclose(f) || bad) return VFG_E_ARG;
  return VFG_OK;
}

/* host copies of U, V and the nodal crack opening of a context */
static int fetch(vfh_ctx *c, int64_t nn, double **U, double **V, double **open) {
  *U = malloc(sizeof(double) * 3 * nn);
  *V = malloc(sizeof(double) * nn);
  *open = malloc(sizeof(double) * nn);
  if (!*U || !*V || !*open) return VFG_E_NOMEM;
  int rc = vfh_get_state(c, *U, *V, NULL);
  if (!rc) rc = vfh_crack_opening_field(c, *open);
  return rc;
}

static void grid_of(vfh_ctx *c, vfg_grid *g) { *g = vfh_options_of(c)->grid; }

int vfh_fields_binary_write(vfh_ctx *c, const char *prefix, int step, const double *theta, const double *pressure) {
  if (!c || !prefix) return VFG_E_ARG;
  vfg_grid g;
  grid_of(c, &g);
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz, ne = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
  double *U = NULL, *V = NULL, *open = NULL;
  int rc = fetch(c, nn, &U, &V, &open);
  if (!rc) {
    char file[4096];
    snprintf(file, sizeof file, "%s.%.5i.bin", prefix, step);
    /* U, velocity, V, pmult, theta, pressure, VolCrackOpening, VolLeakOffRate
     * (VFCommon.c:1365-1372); the flow fields hold the values the reference
     * initialises them to (VecSet(.,0.), VFCommon.c:675, 687, 695) */
    const double *vec[8] = {U, NULL, V, NULL, theta, pressure, open, NULL};
    const int64_t len[8] = {3 * nn, 3 * nn, nn, ne, nn, nn, nn, nn};
    rc = vfh_petsc_vec_write(file, 8, vec, len, 0);
  }
  free(U);
  free(V);
  free(open);
  return rc;
}

int vfh_coordinates_write(vfh_ctx *c, const char *file) {
  if (!c || !file) return VFG_E_ARG;
  vfg_grid g;
  grid_of(c, &g);
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz;
  double *X = malloc(sizeof(double) * 3 * nn);
  if (!X) return VFG_E_NOMEM;
  int64_t q = 0;
  for (int k = 0; k < g.nz; k++)
    for (int j = 0; j < g.ny; j++)
      for (int i = 0; i < g.nx; i++) {
        X[q++] = i == 0 ? 0. : i * g.lx / (g.nx - 1.);
        X[q++] = j == 0 ? 0. : j * g.ly / (g.ny - 1.);
        X[q++] = k == 0 ? 0. : k * g.lz / (g.nz - 1.);
      }
  const double *vec[1] = {X};
  const int64_t len[1] = {3 * nn};
  int rc = vfh_petsc_vec_write(file, 1, vec, len, 0);
  free(X);
  return rc;
}

int vfh_fields_vts_write(vfh_ctx *c, const char *prefix, int step, const double *theta, const double *pressure) {
  if (!c || !prefix) return VFG_E_ARG;
  vfg_grid g;
  grid_of(c, &g);
  const int64_t nn = (int64_t)g.nx * g.ny * g.nz, ne = (int64_t)(g.nx - 1) * (g.ny - 1) * (g.nz - 1);
  double *U = NULL, *V = NULL, *open = NULL, *Uc = NULL, *E3 = NULL, *Ec = NULL;
  int rc = fetch(c, nn, &U, &V, &open);
  if (!rc) {
    Uc = malloc(sizeof(double) * 3 * nn);
    E3 = malloc(sizeof(double) * 3 * ne);
    Ec = malloc(sizeof(double) * 3 * ne);
    if (!Uc || !E3 || !Ec) rc = VFG_E_NOMEM;
  }
  if (!rc) rc = vfh_energy_fields(c, E3);
  if (!rc) {
    for (int comp = 0; comp < 3; comp++)
      for (int64_t n = 0; n < nn; n++) Uc[comp * nn + n] = U[3 * n + comp];
    for (int comp = 0; comp < 3; comp++)
      for (int64_t e = 0; e < ne; e++) Ec[comp * ne + e] = E3[3 * e + comp];
    char file[4096];
    /* names of VFCommon.c:646, 654, 662, 670, 690 */
    const char *pn[7] = {"Displacement_0", "Displacement_1", "Displacement_2", "Fracture", "Temperature",
                         "Pressure", "Volumetric Crack Opening"};
    const double *pd[7] = {Uc, Uc + nn, Uc + 2 * nn, V, theta, pressure, open};
    snprintf(file, sizeof file, "%s_nodal.%.5i.vts", prefix, step);
    rc = vfh_vts_write(file, g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, 7, pn, pd, 0, NULL, NULL);
    if (!rc) {
      const char *cn[3] = {"Elastic Energy", "Pressure Work", "Surface Energy"};
      const double *cd[3] = {Ec, Ec + ne, Ec + 2 * ne};
      snprintf(file, sizeof file, "%s_cell.%.5i.vts", prefix, step);
      rc = vfh_vts_write(file, g.nx, g.ny, g.nz, g.lx, g.ly, g.lz, 0, NULL, NULL, 3, cn, cd);
    }
  }
  free(U);
  free(V);
  free(open);
  free(Uc);
  free(E3);
  free(Ec);
  return rc;
}
