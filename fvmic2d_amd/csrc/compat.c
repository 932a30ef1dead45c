/*
 * compat.c -- the reference's marker-path functions with their own signatures
 * (markers.h, markerProjection.h, subgridTemperatureChanges.h,
 * subgridStressChanges.h, updateMarkerStress.h, updateMarkerStrainPressure.h,
 * residual.h), running on the GPU through fvm_abi.h.
 *
 * Each wrapper uploads the host state it reads (the AoS MarkerSet, the nodal
 * Vecs, the grid), calls the C ABI and downloads what the reference function
 * writes, so the caller's host data changes exactly as with the reference.
 * One process-wide context and a small pool of device buffers are kept between
 * calls.  Where the reference aborts, the wrappers print the reference's
 * message and return a nonzero status instead.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <stddef.h>

#include "fvm_abi.h"
#include "fvm_compat.h"

#define NPOOL 24

static fvm_ctx *g_ctx = NULL;
static fvm_markers g_mk;
static int g_mk_alloc = 0;
static double *g_gridbuf = NULL;   /* 4 padded coordinate arrays */
static size_t g_grid_cap = 0;      /* doubles */
static fvm_grid g_grid;
static double *g_pool[NPOOL];
static size_t g_pool_cap = 0;      /* doubles per pool buffer */
static double *g_scalar = NULL;    /* dense per-marker array */
static size_t g_scalar_cap = 0;
static int g_last_status = 0;
static double *g_S = NULL; /* device copy of the interleaved solution vector */
static size_t g_S_cap = 0;

#define NRES 100 /* residual.c:9 */
static double g_residuals[NRES];

static int ensure_ctx(void) {
  if (g_ctx) return FVM_OK;
  return fvm_ctx_create(&g_ctx, NULL);
}

/* A Vec's host array is checked out only around the copy that uses it.  With
 * PETSc, reads go through VecGetArrayRead and writes hold VecGetArray until the
 * device-to-host copy has completed (fvm_memcpy_dtoh synchronises); the
 * VecRestoreArray that follows advances the Vec's object state, so PETSc drops
 * its cached norms and extrema of the old values.  One rank: the array holds
 * every node in natural order. */
static int vec_to_device(double *dev, Vec v, size_t n) {
#ifdef FVM_WITH_PETSC
  const PetscScalar *a = NULL;
  if (VecGetArrayRead(v, &a)) return FVM_ERR_ARG;
  int rc = fvm_memcpy_htod(g_ctx, dev, a, n * sizeof(double));
  VecRestoreArrayRead(v, &a);
  return rc;
#else
  if (!v || !v->a || (size_t)v->n < n) return FVM_ERR_ARG;
  return fvm_memcpy_htod(g_ctx, dev, v->a, n * sizeof(double));
#endif
}

static int vec_from_device(Vec v, const double *dev, size_t n) {
#ifdef FVM_WITH_PETSC
  PetscScalar *a = NULL;
  if (VecGetArray(v, &a)) return FVM_ERR_ARG;
  int rc = fvm_memcpy_dtoh(g_ctx, a, dev, n * sizeof(double));
  VecRestoreArray(v, &a);
  return rc;
#else
  if (!v || !v->a || (size_t)v->n < n) return FVM_ERR_ARG;
  return fvm_memcpy_dtoh(g_ctx, v->a, dev, n * sizeof(double));
#endif
}

static int upload_grid(const GridData *g) {
  const size_t need = (size_t)(g->NX + 4) * 2 + (size_t)(g->NY + 4) * 2;
  int rc;
  if (need > g_grid_cap) {
    if (g_gridbuf) fvm_free(g_gridbuf);
    g_gridbuf = NULL;
    if ((rc = fvm_malloc((void **)&g_gridbuf, need * sizeof(double))) != FVM_OK) return rc;
    g_grid_cap = need;
  }
  double *dx = g_gridbuf, *dxc = dx + g->NX + 4, *dy = dxc + g->NX + 4, *dyc = dy + g->NY + 4;
  /* the reference's arrays are offset by +2 into NX+4 / NY+4 doubles */
  if ((rc = fvm_memcpy_htod(g_ctx, dx, g->x - 2, (size_t)(g->NX + 4) * sizeof(double)))) return rc;
  if ((rc = fvm_memcpy_htod(g_ctx, dxc, g->xc - 2, (size_t)(g->NX + 4) * sizeof(double)))) return rc;
  if ((rc = fvm_memcpy_htod(g_ctx, dy, g->y - 2, (size_t)(g->NY + 4) * sizeof(double)))) return rc;
  if ((rc = fvm_memcpy_htod(g_ctx, dyc, g->yc - 2, (size_t)(g->NY + 4) * sizeof(double)))) return rc;
  g_grid.NX = g->NX;
  g_grid.NY = g->NY;
  g_grid.LX = g->LX;
  g_grid.LY = g->LY;
  g_grid.xperiodic = g->xperiodic;
  g_grid.x = dx + 2;
  g_grid.xc = dxc + 2;
  g_grid.y = dy + 2;
  g_grid.yc = dyc + 2;
  return FVM_OK;
}

static int upload_markers(const MarkerSet *ms) {
  int rc;
  if (!g_mk_alloc || g_mk.maxMark < ms->nMark) {
    if (g_mk_alloc) fvm_markers_free(&g_mk);
    g_mk_alloc = 0;
    if ((rc = fvm_markers_alloc(g_ctx, ms->nMark > 0 ? ms->nMark : 1, &g_mk)) != FVM_OK) return rc;
    g_mk_alloc = 1;
  }
  return fvm_markers_upload_aos(g_ctx, &g_mk, ms->markers, ms->nMark);
}

static int download_markers(MarkerSet *ms) { return fvm_markers_download_aos(g_ctx, &g_mk, ms->markers); }

static int pool(const GridData *g, int i, double **out) {
  const size_t need = (size_t)g->NX * g->NY;
  if (need > g_pool_cap) {
    for (int k = 0; k < NPOOL; k++) {
      if (g_pool[k]) fvm_free(g_pool[k]);
      g_pool[k] = NULL;
    }
    g_pool_cap = 0;
    for (int k = 0; k < NPOOL; k++) {
      int rc = fvm_malloc((void **)&g_pool[k], need * sizeof(double));
      if (rc) return rc;
    }
    g_pool_cap = need;
  }
  *out = g_pool[i];
  return FVM_OK;
}

static int put(const GridData *g, int i, Vec v, double **dev) {
  int rc = pool(g, i, dev);
  if (rc) return rc;
  return vec_to_device(*dev, v, (size_t)g->NX * g->NY);
}

static int get(const GridData *g, const double *dev, Vec v) {
  return vec_from_device(v, dev, (size_t)g->NX * g->NY);
}

/* a double field of Marker, named by its address in markers[0] (markerProjection.c:99-102) */
static double *column_of(const MarkerSet *ms, const double *fieldptr) {
  const ptrdiff_t off = (const char *)fieldptr - (const char *)ms->markers;
  switch (off) {
    case offsetof(Marker, X): return g_mk.X;
    case offsetof(Marker, Y): return g_mk.Y;
    case offsetof(Marker, VX): return g_mk.VX;
    case offsetof(Marker, VY): return g_mk.VY;
    case offsetof(Marker, T): return g_mk.T;
    case offsetof(Marker, Tdot): return g_mk.Tdot;
    case offsetof(Marker, s.T11): return g_mk.sxx;
    case offsetof(Marker, s.T22): return g_mk.syy;
    case offsetof(Marker, s.T12): return g_mk.sxy;
    case offsetof(Marker, eta): return g_mk.eta;
    case offsetof(Marker, mu): return g_mk.mu;
    case offsetof(Marker, rho): return g_mk.rho;
    case offsetof(Marker, rhodot): return g_mk.rhodot;
    case offsetof(Marker, p): return g_mk.p;
    case offsetof(Marker, e.T11): return g_mk.exx;
    case offsetof(Marker, e.T22): return g_mk.eyy;
    case offsetof(Marker, e.T12): return g_mk.exy;
    case offsetof(Marker, E.T11): return g_mk.Exx;
    case offsetof(Marker, E.T22): return g_mk.Eyy;
    case offsetof(Marker, E.T12): return g_mk.Exy;
    case offsetof(Marker, Eii): return g_mk.Eii;
    case offsetof(Marker, wxy): return g_mk.wxy;
    case offsetof(Marker, strainRateResidual): return g_mk.strainRateResidual;
    default: return NULL;
  }
}

#define TRY(x)             \
  do {                     \
    int r_ = (x);          \
    if (r_ != FVM_OK) {    \
      report(r_);          \
      return r_;           \
    }                      \
  } while (0)

static void report(int rc) {
  g_last_status = rc;
  if (rc) fprintf(stderr, "fvm_compat: status %d: %s\n", rc, g_ctx ? fvm_last_error(g_ctx) : "no context");
}

static int begin(const GridData *g, const MarkerSet *ms) {
  int rc;
  if ((rc = ensure_ctx())) return rc;
  if ((rc = upload_grid(g))) return rc;
  if (ms && (rc = upload_markers(ms))) return rc;
  return FVM_OK;
}

/* ---- markers.c:85-106 -------------------------------------------------------- */
void findMarkerCells(MarkerSet *markerset, GridData *grid) {
  int32_t nout = 0;
  int rc = begin(grid, markerset);
  if (!rc) rc = fvm_find_marker_cells(g_ctx, &g_grid, &g_mk, 0, &nout);
  if (rc == FVM_ERR_CELL_OUT_OF_BOUNDS) printf("cell out of bounds marker\n"); /* markers.c:75 */
  if (!rc || rc == FVM_ERR_CELL_OUT_OF_BOUNDS) {
    int r2 = download_markers(markerset);
    if (!rc) rc = r2;
  }
  markerset->nOut = nout;
  report(rc);
}

/* ---- markerProjection.c:11-68 ------------------------------------------------- */
int projectMarkersNodesAll2(MarkerSet *markerset, GridData *grid, NodalFields *nf, Materials *materials,
                            Options *options) {
  (void)options;
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  Vec outs[12] = {nf->rho, nf->rhodot, nf->etaN, nf->etaS, nf->kThermal, nf->lastT,
                  nf->Cp,  nf->soxx,   nf->soyy, nf->soxy, nf->muS,      nf->muN};
  double *dev[12];
  for (int i = 0; i < 12; i++) TRY(pool(grid, i, &dev[i]));
  d.rho = dev[0], d.rhodot = dev[1], d.etaN = dev[2], d.etaS = dev[3], d.kThermal = dev[4];
  d.lastT = dev[5], d.Cp = dev[6], d.soxx = dev[7], d.soyy = dev[8], d.soxy = dev[9];
  d.muS = dev[10], d.muN = dev[11];
  TRY(fvm_project_markers_nodes_all2(g_ctx, &g_grid, &g_mk, &d, materials));
  for (int i = 0; i < 12; i++) TRY(get(grid, dev[i], outs[i]));
  return 0;
}

/* ---- markerProjection.c:71-250 ------------------------------------------------ */
int projectMarkerFieldToNodes(GridData *grid, MarkerSet *markerset, double *materialProperty,
                              double *markerfield, Vec nodalField, const int32_t stagx, const int32_t stagy,
                              const PROJECTION_METHOD pm) {
  TRY(begin(grid, markerset));
  double *out;
  TRY(pool(grid, 0, &out));
  fvm_projection p;
  memset(&p, 0, sizeof(p));
  if (materialProperty) {
    p.materialProperty = materialProperty;
  } else {
    p.markerField = column_of(markerset, markerfield);
    if (!p.markerField) TRY(FVM_ERR_ARG);
  }
  p.nodalField = out;
  p.stagx = stagx;
  p.stagy = stagy;
  p.method = pm;
  TRY(fvm_project_marker_fields(g_ctx, &g_grid, &g_mk, 1, &p));
  TRY(get(grid, out, nodalField));
  return 0;
}

/* ---- markers.c:1507-1594 -------------------------------------------------------- */
int projectMarkersNodesFromScalar(MarkerSet *markerset, GridData *grid, double *markerField, Vec nodalField) {
  TRY(begin(grid, markerset));
  const size_t n = (size_t)(markerset->nMark > 0 ? markerset->nMark : 1);
  if (n > g_scalar_cap) {
    if (g_scalar) fvm_free(g_scalar);
    g_scalar = NULL;
    TRY(fvm_malloc((void **)&g_scalar, n * sizeof(double)));
    g_scalar_cap = n;
  }
  if (markerset->nMark > 0)
    TRY(fvm_memcpy_htod(g_ctx, g_scalar, markerField, (size_t)markerset->nMark * sizeof(double)));
  double *out;
  TRY(pool(grid, 0, &out));
  TRY(fvm_project_markers_nodes_from_scalar(g_ctx, &g_grid, &g_mk, g_scalar, out));
  TRY(get(grid, out, nodalField));
  return 0;
}

/* ---- markers.c:1670-1706 -------------------------------------------------------- */
int projectNodalFieldToMarkersS(NodalFields *nodalFields, Vec nodalField, MarkerSet *markerset, double *fieldptr,
                                GridData *grid) {
  (void)nodalFields;
  TRY(begin(grid, markerset));
  double *f;
  TRY(put(grid, 0, nodalField, &f));
  double *col = column_of(markerset, fieldptr);
  if (!col) TRY(FVM_ERR_ARG);
  TRY(fvm_project_nodal_field_to_markers(g_ctx, &g_grid, f, &g_mk, col));
  TRY(download_markers(markerset));
  return 0;
}

/* ---- markers.c:1865-2250 -------------------------------------------------------- */
int advectMarkersRK(MarkerSet *markerset, NodalFields *nf, GridData *grid, Options *options, BoundaryValues *bv,
                    double dt) {
  (void)options;
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->vx, &d.vx));
  TRY(put(grid, 1, nf->vy, &d.vy));
  TRY(put(grid, 2, nf->wxy, &d.wxy));
#ifdef JAUMANN
  const int jaumann = 1;
#else
  const int jaumann = 0;
#endif
  TRY(fvm_advect_markers_rk(g_ctx, &g_mk, &d, &g_grid, bv, dt, jaumann));
  TRY(download_markers(markerset));
  return 0;
}

/* ---- markers.c:1381-1505 -------------------------------------------------------- */
int checkPlasticYielding(GridData *grid, MarkerSet *markerset, Materials *materials, double dt, int32_t *isYielding,
                         Options *options, double elapsedTime) {
  TRY(begin(grid, markerset));
  const int rc = fvm_check_plastic_yielding(g_ctx, &g_grid, &g_mk, materials, dt, isYielding, options, elapsedTime);
  if (rc == FVM_ERR_VISCOSITY_LAW) printf("Error! Viscosity law undefined\n"); /* viscosity.c:61 */
  TRY(rc);
  TRY(download_markers(markerset));
  return 0;
}

/* ---- viscosity.c:90-96 ------------------------------------------------------------ */
void limitViscosity(MarkerSet *markerset, Options *options, Materials *materials) {
  (void)materials;
  int rc = ensure_ctx();
  if (!rc && (!g_mk_alloc || g_mk.maxMark < markerset->nMark)) {
    if (g_mk_alloc) fvm_markers_free(&g_mk);
    g_mk_alloc = 0;
    rc = fvm_markers_alloc(g_ctx, markerset->nMark > 0 ? markerset->nMark : 1, &g_mk);
    if (!rc) g_mk_alloc = 1;
  }
  if (!rc) rc = fvm_markers_upload_aos(g_ctx, &g_mk, markerset->markers, markerset->nMark);
  if (!rc) rc = fvm_limit_viscosity(g_ctx, &g_mk, options);
  if (!rc) rc = download_markers(markerset);
  report(rc);
}

/* ---- subgridTemperatureChanges.c:5-124 ------------------------------------------- */
int subgridTemperatureChanges(Vec thisT, GridData *grid, NodalFields *nf, MarkerSet *markerset,
                              Materials *materials, double dt, Options *options) {
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  double *t;
  TRY(put(grid, 0, thisT, &t));
  TRY(put(grid, 1, nf->lastT, &d.lastT));
  TRY(put(grid, 2, nf->rho, &d.rho));
  TRY(put(grid, 3, nf->Cp, &d.Cp));
  TRY(put(grid, 4, nf->kThermal, &d.kThermal));
  TRY(fvm_subgrid_temperature_changes(g_ctx, t, &g_grid, &d, &g_mk, materials, dt,
                                      options->subgridTemperatureDiffusivity));
  TRY(download_markers(markerset));
  return 0;
}

/* ---- subgridStressChanges.c:5-318 ------------------------------------------------- */
int subgridStressChanges(GridData *grid, NodalFields *nf, MarkerSet *markerset, Materials *materials, double dt,
                         double dsubgrid) {
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->soxx, &d.soxx));
  TRY(put(grid, 1, nf->soyy, &d.soyy));
  TRY(put(grid, 2, nf->soxy, &d.soxy));
  TRY(put(grid, 3, nf->dsxx, &d.dsxx));
  TRY(put(grid, 4, nf->dsyy, &d.dsyy));
  TRY(put(grid, 5, nf->dsxy, &d.dsxy));
  TRY(fvm_subgrid_stress_changes(g_ctx, &g_grid, &d, &g_mk, materials, dt, dsubgrid));
  TRY(download_markers(markerset));
  TRY(get(grid, d.dsxx, nf->dsxx));
  TRY(get(grid, d.dsyy, nf->dsyy));
  TRY(get(grid, d.dsxy, nf->dsxy));
  return 0;
}

/* ---- updateMarkerStress.c:5-163 ---------------------------------------------------- */
int updateMarkerStress(GridData *grid, NodalFields *nf, MarkerSet *markerset, Materials *materials) {
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->dsxx, &d.dsxx));
  TRY(put(grid, 1, nf->dsyy, &d.dsyy));
  TRY(put(grid, 2, nf->dsxy, &d.dsxy));
  TRY(fvm_update_marker_stress(g_ctx, &g_grid, &d, &g_mk, materials));
  TRY(download_markers(markerset));
  return 0;
}

/* ---- updateMarkerStrainPressure.c:9-399 -------------------------------------------- */
int updateMarkerStrainPressure(GridData *grid, NodalFields *nf, MarkerSet *markerset, Materials *materials,
                               Options *options, double dt) {
  TRY(begin(grid, markerset));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->dsxx, &d.dsxx));
  TRY(put(grid, 1, nf->dsyy, &d.dsyy));
  TRY(put(grid, 2, nf->dsxy, &d.dsxy));
  TRY(put(grid, 3, nf->soxx, &d.soxx));
  TRY(put(grid, 4, nf->soyy, &d.soyy));
  TRY(put(grid, 5, nf->soxy, &d.soxy));
  TRY(put(grid, 6, nf->edotxx, &d.edotxx));
  TRY(put(grid, 7, nf->edotyy, &d.edotyy));
  TRY(put(grid, 8, nf->edotxy, &d.edotxy));
  TRY(put(grid, 9, nf->wxy, &d.wxy));
  TRY(put(grid, 10, nf->p, &d.p));
  const int rc = fvm_update_marker_strain_pressure(g_ctx, &g_grid, &d, &g_mk, materials, options, dt);
  if (rc == FVM_ERR_NAN_RESIDUAL) printf("marker has nan strain rate residual\n"); /* :303 */
  if (rc == FVM_OK || rc == FVM_ERR_NAN_RESIDUAL) {
    int r2 = download_markers(markerset);
    if (r2) TRY(r2);
  }
  TRY(rc);
  return 0;
}

/* ---- residual.c:12-77 ---------------------------------------------------------------- */
int updateGlobalStrainRateResidual(GridData *grid, MarkerSet *markerset, int32_t iter) {
  TRY(begin(grid, markerset));
  double r = 0.0;
  TRY(fvm_update_global_strain_rate_residual(g_ctx, &g_grid, &g_mk, &r));
  g_residuals[iter % NRES] = r;
  return 0;
}

double getGlobalStrainRateResidual(int32_t iter) { return g_residuals[iter % NRES]; }

int32_t checkConvergence(double atol, double rtol, int32_t iter) {
  if (getGlobalStrainRateResidual(iter) < atol) return 1;
  if (iter > 0) {
    const double thisres = getGlobalStrainRateResidual(iter);
    const double lastres = getGlobalStrainRateResidual(iter - 1);
    double d = (thisres - lastres) / lastres;
    if (d < 0) d = -d;
    if (d < rtol) return 2;
  }
  return 0;
}

/* ---- retrieveSolutions.c:4-82 ---------------------------------------------------- */
int retrieveSolutions(GridData *grid, NodalFields *nf, Vec S, double Kcont, int32_t *foundnan) {
  TRY(begin(grid, NULL));
  const size_t nn = (size_t)grid->NX * grid->NY;
  if (3 * nn > g_S_cap) {
    if (g_S) fvm_free(g_S);
    g_S = NULL;
    g_S_cap = 0;
    TRY(fvm_malloc((void **)&g_S, 3 * nn * sizeof(double)));
    g_S_cap = 3 * nn;
  }
  TRY(vec_to_device(g_S, S, 3 * nn));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(pool(grid, 0, &d.vx));
  TRY(pool(grid, 1, &d.vy));
  TRY(pool(grid, 2, &d.p));
  TRY(fvm_retrieve_solutions(g_ctx, &g_grid, &d, g_S, Kcont, foundnan));
  TRY(vec_from_device(S, g_S, 3 * nn)); /* pressure scaled in place */
  TRY(get(grid, d.vx, nf->vx));
  TRY(get(grid, d.vy, nf->vy));
  TRY(get(grid, d.p, nf->p));
  return 0;
}

/* ---- initialPressureGuess.c:6-184 ----------------------------------------------------- */
/* The reference solves the pressure Poisson problem with the energy_ KSP (a direct LU);
 * here BiCGStab runs to a relative residual of 1e-12.  Only the pressure entries of S
 * are written. */
int initialPressureGuess(GridData *grid, NodalFields *nf, Options *options, Vec S) {
  TRY(begin(grid, NULL));
  const size_t nn = (size_t)grid->NX * grid->NY;
  if (3 * nn > g_S_cap) {
    if (g_S) fvm_free(g_S);
    g_S = NULL;
    g_S_cap = 0;
    TRY(fvm_malloc((void **)&g_S, 3 * nn * sizeof(double)));
    g_S_cap = 3 * nn;
  }
  double *rho;
  TRY(put(grid, 0, nf->rho, &rho));
  TRY(vec_to_device(g_S, S, 3 * nn));
  int32_t its = 0;
  double rel = 0.0;
  TRY(fvm_initial_pressure_guess(g_ctx, &g_grid, rho, options->gy, 0.0, 1e-12, 1000000, g_S, NULL, &its, &rel));
  TRY(vec_from_device(S, g_S, 3 * nn));
  return 0;
}

/* ---- nodalStressStrain.c:6-485 ------------------------------------------------------ */
int nodalStressStrain(GridData *grid, NodalFields *nf, Options *options, BoundaryValues *bv, double dt,
                      Vec nodalHeating, double gy) {
  (void)options;
  (void)gy;
  TRY(begin(grid, NULL));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->vx, (double **)&d.vx));
  TRY(put(grid, 1, nf->vy, (double **)&d.vy));
  TRY(put(grid, 2, nf->etaN, &d.etaN));
  TRY(put(grid, 3, nf->muN, &d.muN));
  TRY(put(grid, 4, nf->etaS, &d.etaS));
  TRY(put(grid, 5, nf->muS, &d.muS));
  TRY(put(grid, 6, nf->soxx, &d.soxx));
  TRY(put(grid, 7, nf->soyy, &d.soyy));
  TRY(put(grid, 8, nf->soxy, &d.soxy));
  TRY(pool(grid, 9, &d.edotxx));
  TRY(pool(grid, 10, &d.edotyy));
  TRY(pool(grid, 11, &d.edotxy));
  TRY(pool(grid, 12, &d.wxy));
  TRY(pool(grid, 13, &d.dsxx));
  TRY(pool(grid, 14, &d.dsyy));
  TRY(pool(grid, 15, &d.dsxy));
  TRY(pool(grid, 16, &d.eii));
  TRY(pool(grid, 17, &d.sii));
  double *heat;
  TRY(pool(grid, 18, &heat));
  TRY(fvm_nodal_stress_strain(g_ctx, &g_grid, &d, bv, dt, heat));
  TRY(get(grid, d.edotxx, nf->edotxx));
  TRY(get(grid, d.edotyy, nf->edotyy));
  TRY(get(grid, d.edotxy, nf->edotxy));
  TRY(get(grid, d.wxy, nf->wxy));
  TRY(get(grid, d.soxx, nf->soxx));
  TRY(get(grid, d.soyy, nf->soyy));
  TRY(get(grid, d.soxy, nf->soxy));
  TRY(get(grid, d.dsxx, nf->dsxx));
  TRY(get(grid, d.dsyy, nf->dsyy));
  TRY(get(grid, d.dsxy, nf->dsxy));
  TRY(get(grid, d.eii, nf->eii));
  TRY(get(grid, d.sii, nf->sii));
  TRY(get(grid, heat, nodalHeating));
  return 0;
}

/* ---- limitTimestep.c:6-133 ---------------------------------------------------------- */
int limitDisplacementTimestep(GridData *grid, NodalFields *nf, double *displacementdt, Options *options) {
  TRY(begin(grid, NULL));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  TRY(put(grid, 0, nf->vx, &d.vx));
  TRY(put(grid, 1, nf->vy, &d.vy));
  TRY(put(grid, 2, nf->rho, &d.rho));
  TRY(put(grid, 3, nf->rhodot, &d.rhodot));
  TRY(fvm_limit_displacement_timestep(g_ctx, &g_grid, &d, displacementdt, options));
  return 0;
}

/* ---- io.c:299-402 / restart.c:8-125 (one rank; files output/Markers.<iMonte>.<iTime>) -- */
int saveMarkersBinary(MarkerSet *markerset, int32_t iMonte, int32_t iTime, double elapsedTime) {
  char fn[80];
  snprintf(fn, sizeof(fn), "output/Markers.%d.%d", iMonte, iTime);
  TRY(ensure_ctx());
  TRY(upload_markers(markerset));
  TRY(fvm_save_markers_binary(g_ctx, &g_mk, fn, elapsedTime, NULL));
  return 0;
}

/* io.c:150-296 with format OUTPUT_NODAL_FIELDS_BINARY (the only one written
 * here; the ASCII/VTK viewers are PETSc's own): output/loadNodalFields_<iMonte>_<iTime>.petscbin */
int saveNodalFields(NodalFields *nf, GridData *grid, int32_t iMonte, int32_t iTime, double dt, double elapsedTime,
                    int32_t format) {
  (void)dt;
  (void)elapsedTime;
  if (format != FVM_OUTPUT_NODAL_FIELDS_BINARY) {
    fprintf(stderr, "fvm_compat: saveNodalFields writes the binary format only\n");
    g_last_status = FVM_ERR_ARG;
    return FVM_ERR_ARG;
  }
  char fn[80];
  snprintf(fn, sizeof(fn), "output/loadNodalFields_%d_%d.petscbin", iMonte, iTime);
  TRY(begin(grid, NULL));
  NodalArrays d;
  memset(&d, 0, sizeof(d));
  Vec src[20] = {nf->vx,   nf->vy,   nf->p,    nf->rho,    nf->rhodot, nf->etaN,   nf->etaS,
                 nf->kThermal, nf->lastT, nf->Cp, nf->muN, nf->muS,    nf->soxx,   nf->soyy,
                 nf->soxy, nf->edotxx, nf->edotyy, nf->edotxy, nf->wxy, nf->ha};
  double **dst[19] = {&d.vx,   &d.vy,   &d.p,    &d.rho,    &d.rhodot, &d.etaN,   &d.etaS,
                      &d.kThermal, &d.lastT, &d.Cp, &d.muN, &d.muS,    &d.soxx,   &d.soyy,
                      &d.soxy, &d.edotxx, &d.edotyy, &d.edotxy, &d.wxy};
  for (int k = 0; k < 19; k++) TRY(put(grid, k, src[k], dst[k]));
  double *ha = NULL;
  if (src[19]) TRY(put(grid, 19, src[19], &ha));
  TRY(fvm_save_nodal_fields(g_ctx, &g_grid, &d, ha, fn));
  return 0;
}

/* Allocates markerset->markers (release it with free()) with room for
 * nMark * options->maxMarkFactor markers, as restart.c:45 does. */
int restartFromMarkers(MarkerSet *markerset, GridData *grid, Materials *materials, Options *options, int32_t iMonte,
                       int32_t iTime, double *elapsedTime) {
  (void)grid;
  (void)materials;
  char fn[80];
  snprintf(fn, sizeof(fn), "output/Markers.%d.%d", iMonte, iTime);
  TRY(ensure_ctx());
  FILE *f = fopen(fn, "rb");
  int32_t n = -1;
  if (f) {
    if (fread(&n, sizeof(int32_t), 1, f) != 1) n = -1;
    fclose(f);
  }
  if (n < 0) {
    fprintf(stderr, "fvm_compat: cannot read %s\n", fn);
    g_last_status = FVM_ERR_IO;
    return FVM_ERR_IO;
  }
  int32_t cap = (int32_t)((double)n * options->maxMarkFactor);
  if (cap < n) cap = n;
  if (!g_mk_alloc || g_mk.maxMark < cap) {
    if (g_mk_alloc) fvm_markers_free(&g_mk);
    g_mk_alloc = 0;
    TRY(fvm_markers_alloc(g_ctx, cap > 0 ? cap : 1, &g_mk));
    g_mk_alloc = 1;
  }
  TRY(fvm_restart_from_markers(g_ctx, &g_mk, fn, elapsedTime));
  Marker *m = (Marker *)calloc(cap > 0 ? (size_t)cap : 1, sizeof(Marker));
  if (!m) return FVM_ERR_NOMEM;
  const int rc = fvm_markers_download_aos(g_ctx, &g_mk, m);
  if (rc != FVM_OK) {
    free(m);
    report(rc);
    return rc;
  }
  for (int32_t i = n; i < cap; i++) { /* resetMarker (markers.c:250-330) */
    m[i].X = -1.0;
    m[i].Y = -1.0;
    m[i].Mat = -1;
  }
  markerset->markers = m;
  markerset->nMark = n;
  markerset->maxMark = cap;
  markerset->nOut = 0;
  return 0;
}

int fvm_compat_last_status(void) { return g_last_status; }

void fvm_compat_finalize(void) {
  if (g_mk_alloc) fvm_markers_free(&g_mk);
  g_mk_alloc = 0;
  for (int k = 0; k < NPOOL; k++) {
    if (g_pool[k]) fvm_free(g_pool[k]);
    g_pool[k] = NULL;
  }
  g_pool_cap = 0;
  if (g_gridbuf) fvm_free(g_gridbuf);
  g_gridbuf = NULL;
  g_grid_cap = 0;
  if (g_scalar) fvm_free(g_scalar);
  g_scalar = NULL;
  g_scalar_cap = 0;
  if (g_S) fvm_free(g_S);
  g_S = NULL;
  g_S_cap = 0;
  if (g_ctx) fvm_ctx_destroy(g_ctx);
  g_ctx = NULL;
}
