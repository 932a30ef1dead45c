/*
 * oracle.c -- CPU restatement of fvmic2d's marker hot path.  TEST INFRASTRUCTURE
 * ONLY (see oracle.h).  Build with -ffp-contract=off: the reference was built by
 * gcc -O3 without -march (makefile:14), i.e. no fused multiply-add, and the
 * position / cell chain must reproduce its roundings bit for bit.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>

#include "oracle.h"

/* ---- ghosted local arrays (DMDA stencil width 2, nodalFields.c:11-20) ----
 * x is DM_BOUNDARY_PERIODIC when grid->xperiodic, else DM_BOUNDARY_GHOSTED;
 * y is always DM_BOUNDARY_GHOSTED. */
typedef struct {
  int NX, NY, per;
  double *a; /* (NY+4) x (NX+4), zero-initialised */
} Local;

static int local_new(Local *l, int NX, int NY, int per) {
  l->NX = NX;
  l->NY = NY;
  l->per = per != 0;
  l->a = (double *)calloc((size_t)(NX + 4) * (NY + 4), sizeof(double));
  return l->a ? ORC_OK : ORC_ERR_NOMEM;
}
static void local_free(Local *l) { free(l->a); l->a = NULL; }
static double *local_at(Local *l, int jy, int ix) {
  return &l->a[(size_t)(jy + 2) * (l->NX + 4) + (ix + 2)];
}
/* DMGlobalToLocal: interior from the global vector; on a periodic x axis the
 * two ghost columns on each side hold the wrapped columns (ix -> ix mod NX),
 * every other ghost stays zero. */
static void local_fill(Local *l, const double *global) {
  const int NX = l->NX;
  for (int jy = 0; jy < l->NY; jy++) {
    for (int ix = 0; ix < NX; ix++) *local_at(l, jy, ix) = global[(size_t)jy * NX + ix];
    if (l->per)
      for (int ix = -2; ix < NX + 2; ix += (ix == -1 ? NX + 1 : 1))
        *local_at(l, jy, ix) = global[(size_t)jy * NX + (ix + NX) % NX];
  }
}
/* DMLocalToGlobal ADD_VALUES: ghosts of a ghosted axis are dropped; on a
 * periodic x axis a ghost column adds into the node it wraps to.  Every
 * wrapped node then receives exactly two addends (NX >= 4), so the sum is
 * independent of the order in which PETSc's scatter applies them. */
static void local_add_to_global(Local *l, double *global) {
  const int NX = l->NX;
  for (int jy = 0; jy < l->NY; jy++) {
    for (int ix = 0; ix < NX; ix++) {
      size_t i = (size_t)jy * NX + ix;
      global[i] = global[i] + *local_at(l, jy, ix);
    }
    if (l->per)
      for (int ix = -2; ix < NX + 2; ix += (ix == -1 ? NX + 1 : 1)) {
        size_t i = (size_t)jy * NX + (ix + NX) % NX;
        global[i] = global[i] + *local_at(l, jy, ix);
      }
  }
}
#define LA(l, jy, ix) (*local_at(&(l), (jy), (ix)))

static double field_of(const Marker *mk, size_t off) {
  double v;
  memcpy(&v, (const char *)mk + off, sizeof(double));
  return v;
}
static void set_field_of(Marker *mk, size_t off, double v) { memcpy((char *)mk + off, &v, sizeof(double)); }

/* ---- markers.c:41-106 ---------------------------------------------------- */
static int find_marker_cell(Marker *mk, const GridData *g) {
  const int NX = g->NX, NY = g->NY;
  if (mk->X >= 0.0 && mk->Y >= 0.0 && mk->X <= g->LX && mk->Y <= g->LY) {
    /* a periodic x axis has NX cells: x[NX] = LX closes the last one */
    int lo = 0, hi = g->xperiodic ? NX : NX - 1;
    while (hi - lo > 1) {
      int mid = (hi + lo) / 2;
      if (mid == lo) mid++;
      else if (mid == hi) mid--;
      if (mk->X >= g->x[mid]) lo = mid;
      else hi = mid;
    }
    mk->cellX = lo;
    lo = 0;
    hi = NY - 1;
    while (hi - lo > 1) {
      int mid = (hi + lo) / 2;
      if (mid == lo) mid++;
      else if (mid == hi) mid--;
      if (mk->Y >= g->y[mid]) lo = mid;
      else hi = mid;
    }
    mk->cellY = lo;
    if (mk->cellX < 0 || mk->cellX > (g->xperiodic ? NX - 1 : NX - 2) || mk->cellY < 0 || mk->cellY > NY - 2)
      return ORC_ERR_CELL_OUT_OF_BOUNDS; /* markers.c:74-77 aborts */
  } else {
    mk->cellX = -1;
    mk->cellY = -1;
  }
  return ORC_OK;
}

int orc_find_marker_cells(MarkerSet *ms, const GridData *g, int rank) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  int nout = 0, err = ORC_OK;
  for (int m = 0; m < ms->nMark; m++) {
    int e = find_marker_cell(&ms->markers[m], g);
    if (e && !err) err = e;
    ms->markers[m].cpu = rank;
    if (ms->markers[m].cellX == -1) nout++;
  }
  ms->nOut = nout;
  return err;
}

/* ---- markerProjection.c:252-269 ----------------------------------------- */
static void find_cell_stag(const Marker *mk, const GridData *g, int stagx, int stagy, int *cx, int *cy) {
  *cx = mk->cellX;
  *cy = mk->cellY;
  if (stagx == 0 && stagy == 0) return;
  if (stagx == 1) {
    if (mk->X < g->xc[mk->cellX + 1]) (*cx)--;
  } else if (stagx == -1) {
    if (mk->X >= g->xc[mk->cellX + 1]) (*cx)++;
  }
  if (stagy == 1) {
    if (mk->Y < g->yc[mk->cellY + 1]) (*cy)--;
  } else if (stagy == -1) {
    if (mk->Y >= g->yc[mk->cellY + 1]) (*cy)++;
  }
}

/* ---- markerProjection.c:71-250 ----------------------------------------- */
int orc_project_marker_field(const GridData *g, const MarkerSet *ms, const double *materialProperty,
                             size_t field_offset, double *nodalField, int stagx, int stagy,
                             int method) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  if (stagx < -1 || stagx > 1 || stagy < -1 || stagy > 1) return ORC_ERR_BAD_STAG;
  if (method != ARITHMETIC && method != ARITHMETIC_LOCAL && method != HARMONIC) return ORC_ERR_BAD_METHOD;
  const int NX = g->NX, NY = g->NY;
  const size_t nn = (size_t)NX * NY;
  double *wtg = (double *)calloc(nn, sizeof(double));
  Local wt, nf;
  if (!wtg || local_new(&wt, NX, NY, g->xperiodic) || local_new(&nf, NX, NY, g->xperiodic)) return ORC_ERR_NOMEM;
  memset(nodalField, 0, nn * sizeof(double));
  const Marker *mk = ms->markers;
  double *mf = (double *)malloc((size_t)(ms->nMark > 0 ? ms->nMark : 1) * sizeof(double));
  if (!mf) return ORC_ERR_NOMEM;
  for (int m = 0; m < ms->nMark; m++) {
    if (materialProperty == NULL) mf[m] = field_of(&mk[m], field_offset);
    else mf[m] = (mk[m].cellX != -1) ? materialProperty[mk[m].Mat] : 0.0; /* only active markers are used */
  }
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    int cellX, cellY;
    find_cell_stag(&mk[m], g, stagx, stagy, &cellX, &cellY);
    double mdx, mdy;
    if (stagx == -1) mdx = (mk[m].X - g->xc[cellX]) / (g->xc[cellX + 1] - g->xc[cellX]);
    else if (stagx == 0) mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    else mdx = (mk[m].X - g->xc[cellX + 1]) / (g->xc[cellX + 2] - g->xc[cellX + 1]);
    if (stagy == -1) mdy = (mk[m].Y - g->yc[cellY]) / (g->yc[cellY + 1] - g->yc[cellY]);
    else if (stagy == 0) mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    else mdy = (mk[m].Y - g->yc[cellY + 1]) / (g->yc[cellY + 2] - g->yc[cellY + 1]);
    if (method == ARITHMETIC) {
      const double w1 = (1.0 - mdx) * (1.0 - mdy);
      LA(nf, cellY, cellX) += w1 * mf[m];
      LA(wt, cellY, cellX) += w1;
      const double w2 = mdx * (1.0 - mdy);
      LA(nf, cellY, cellX + 1) += w2 * mf[m];
      LA(wt, cellY, cellX + 1) += w2;
      const double w3 = (1.0 - mdx) * (mdy);
      LA(nf, cellY + 1, cellX) += w3 * mf[m];
      LA(wt, cellY + 1, cellX) += w3;
      const double w4 = mdx * mdy;
      LA(nf, cellY + 1, cellX + 1) += w4 * mf[m];
      LA(wt, cellY + 1, cellX + 1) += w4;
    } else if (method == ARITHMETIC_LOCAL) {
      double w = 1.0;
      if (mdx >= 0.5) {
        cellX++;
        w *= mdx;
      } else {
        w *= (1.0 - mdx);
      }
      if (mdy >= 0.5) {
        cellY++;
        w *= mdy;
      } else {
        w *= (1.0 - mdy);
      }
      LA(nf, cellY, cellX) += w * mf[m];
      LA(wt, cellY, cellX) += w;
    } else { /* HARMONIC */
      if (mdx >= 0.5) cellX++;
      if (mdy >= 0.5) cellY++;
      LA(nf, cellY, cellX) += 1.0 / mf[m];
      LA(wt, cellY, cellX) += 1.0;
    }
  }
  local_add_to_global(&nf, nodalField);
  local_add_to_global(&wt, wtg);
  for (size_t i = 0; i < nn; i++) {
    if (method == HARMONIC) nodalField[i] = wtg[i] / nodalField[i];
    else nodalField[i] = nodalField[i] / wtg[i];
  }
  free(mf);
  free(wtg);
  local_free(&wt);
  local_free(&nf);
  return ORC_OK;
}

/* ---- markerProjection.c:11-68 ------------------------------------------- */
int orc_project_markers_nodes_all2(const MarkerSet *ms, const GridData *g, NodalArrays *nf,
                                   const Materials *mat) {
  const size_t nn = (size_t)g->NX * g->NY;
  int e = 0;
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, rho), nf->rho, 0, 0, ARITHMETIC);
  memset(nf->rhodot, 0, nn * sizeof(double));
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, eta), nf->etaN, -1, -1, ARITHMETIC);
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, eta), nf->etaS, 0, 0, ARITHMETIC);
  for (size_t i = 0; i < nn; i++) {
    nf->muS[i] = 1e99;
    nf->muN[i] = 1e99;
  }
  e |= orc_project_marker_field(g, ms, mat->materialkThermal, 0, nf->kThermal, 0, 0, ARITHMETIC);
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, T), nf->lastT, 0, 0, ARITHMETIC);
  e |= orc_project_marker_field(g, ms, mat->materialCp, 0, nf->Cp, 0, 0, ARITHMETIC);
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, s.T11), nf->soxx, -1, -1, ARITHMETIC_LOCAL);
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, s.T22), nf->soyy, -1, -1, ARITHMETIC_LOCAL);
  e |= orc_project_marker_field(g, ms, NULL, offsetof(Marker, s.T12), nf->soxy, 0, 0, ARITHMETIC_LOCAL);
  return e;
}

/* ---- markers.c:1507-1594 ------------------------------------------------ */
int orc_project_markers_nodes_from_scalar(const MarkerSet *ms, const GridData *g,
                                          const double *markerField, double *nodalField) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  const size_t nn = (size_t)NX * NY;
  double *wtng = (double *)calloc(nn, sizeof(double));
  Local wtn, nf;
  if (!wtng || local_new(&wtn, NX, NY, g->xperiodic) || local_new(&nf, NX, NY, g->xperiodic)) return ORC_ERR_NOMEM;
  memset(nodalField, 0, nn * sizeof(double));
  const Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    const int cellX = mk[m].cellX, cellY = mk[m].cellY;
    const double dx = g->x[cellX + 1] - g->x[cellX];
    const double dy = g->y[cellY + 1] - g->y[cellY];
    const double mdx = (mk[m].X - g->x[cellX]) / dx;
    const double mdy = (mk[m].Y - g->y[cellY]) / dy;
    double w = (1.0 - mdx) * (1.0 - mdy);
    LA(wtn, cellY, cellX) += w;
    LA(nf, cellY, cellX) += w * markerField[m];
    w = mdx * (1.0 - mdy);
    LA(wtn, cellY, cellX + 1) += w;
    LA(nf, cellY, cellX + 1) += w * markerField[m];
    w = (1.0 - mdx) * mdy;
    LA(wtn, cellY + 1, cellX) += w;
    LA(nf, cellY + 1, cellX) += w * markerField[m];
    w = mdx * mdy;
    LA(wtn, cellY + 1, cellX + 1) += w;
    LA(nf, cellY + 1, cellX + 1) += w * markerField[m];
  }
  local_add_to_global(&wtn, wtng);
  local_add_to_global(&nf, nodalField);
  for (size_t i = 0; i < nn; i++) nodalField[i] = nodalField[i] / wtng[i];
  free(wtng);
  local_free(&wtn);
  local_free(&nf);
  return ORC_OK;
}

/* ---- markers.c:1670-1706 ------------------------------------------------ */
int orc_project_nodal_field_to_markers(const double *nodalField, MarkerSet *ms, size_t field_offset,
                                       const GridData *g) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  Local f;
  if (local_new(&f, g->NX, g->NY, g->xperiodic)) return ORC_ERR_NOMEM;
  local_fill(&f, nodalField);
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    const int cellX = mk[m].cellX, cellY = mk[m].cellY;
    const double mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    const double mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    const double val = LA(f, cellY, cellX) * (1 - mdx) * (1 - mdy) + LA(f, cellY + 1, cellX) * (1 - mdx) * mdy +
                       LA(f, cellY, cellX + 1) * (mdx) * (1 - mdy) + LA(f, cellY + 1, cellX + 1) * mdx * mdy;
    set_field_of(&mk[m], field_offset, val);
  }
  local_free(&f);
  return ORC_OK;
}

/* ---- markers.c:1865-2250 ------------------------------------------------ */
int orc_advect_markers_rk(MarkerSet *ms, const NodalArrays *nfa, const GridData *g,
                          const BoundaryValues *bv, double dt, int jaumann) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  Local vx, vy, wxy;
  if (local_new(&vx, NX, NY, g->xperiodic) || local_new(&vy, NX, NY, g->xperiodic) || local_new(&wxy, NX, NY, g->xperiodic)) return ORC_ERR_NOMEM;
  local_fill(&vx, nfa->vx);
  local_fill(&vy, nfa->vy);
  local_fill(&wxy, nfa->wxy);
  const int per = g->xperiodic != 0;
  /* markers.c:1908-1915: a periodic walk may leave [0, NX-1] by one cell */
  const int XCMAX = per ? NX - 1 : NX - 2, XCMIN = per ? -1 : 0;
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    double xcur = mk[m].X, ycur = mk[m].Y;
    double mvx[4], mvy[4], mwxy[4];
    for (int irk = 0; irk < 4; irk++) {
      int cellX = mk[m].cellX, cellY = mk[m].cellY;
      if (xcur > g->x[cellX + 1]) {
        while (xcur > g->x[cellX + 1] && cellX < XCMAX) cellX++;
      } else if (xcur < g->x[cellX]) {
        while (xcur < g->x[cellX] && cellX > XCMIN) cellX--;
      }
      if (ycur > g->y[cellY + 1]) {
        while (ycur > g->y[cellY + 1] && cellY < NY - 2) cellY++;
      } else if (ycur < g->y[cellY]) {
        while (ycur < g->y[cellY] && cellY > 0) cellY--;
      }
      int vxX = cellX, vxY = cellY;
      if (vxY > -1 && ycur < g->yc[vxY + 1]) vxY--;
      int vyX = cellX;
      if (vyX > -1 && xcur < g->xc[vyX + 1]) vyX--;
      int vyY = cellY;
      if (!per) { /* markers.c:1965-1971: no x clamp on a periodic grid */
        if (vxX < 0) vxX = 0;
        else if (vxX > NX - 2) vxX = NX - 2;
        if (vyX < -1) vyX = -1;
        else if (vyX > NX - 2) vyX = NX - 2;
      }
      if (vxY < -1) vxY = -1;
      else if (vxY > NY - 2) vxY = NY - 2;
      if (vyY < 0) vyY = 0;
      else if (vyY > NY - 2) vyY = NY - 2;

      double deltaxm = xcur - g->x[vxX];
      double deltaym = ycur - g->yc[vxY + 1];
      double dx = g->x[vxX + 1] - g->x[vxX];
      double dy = g->yc[vxY + 2] - g->yc[vxY + 1];
      double vxul, vxur;
      if (vxY == -1 && bv->mechBCTop.type[0] == 0) {
        vxul = 2.0 * bv->mechBCTop.value[0] - LA(vx, vxY + 1, vxX);
        vxur = 2.0 * bv->mechBCTop.value[0] - LA(vx, vxY + 1, vxX + 1);
      } else if (vxY == -1 && bv->mechBCTop.type[0] == 1) {
        vxul = LA(vx, vxY + 1, vxX);
        vxur = LA(vx, vxY + 1, vxX + 1);
      } else {
        vxul = LA(vx, vxY, vxX);
        vxur = LA(vx, vxY, vxX + 1);
      }
      const double vxll = LA(vx, vxY + 1, vxX);
      const double vxlr = LA(vx, vxY + 1, vxX + 1);
      mvx[irk] = vxul * (1.0 - deltaxm / dx) * (1.0 - deltaym / dy);
      mvx[irk] += vxur * (deltaxm / dx) * (1.0 - deltaym / dy);
      mvx[irk] += vxll * (1.0 - deltaxm / dx) * (deltaym / dy);
      mvx[irk] += vxlr * (deltaxm * deltaym) / dx / dy;

      double vyul;
      const double vyur = LA(vy, vyY, vyX + 1);
      double vyll;
      const double vylr = LA(vy, vyY + 1, vyX + 1);
      if (!per && vyX == -1 && bv->mechBCLeft.type[1] == 0) {
        vyul = 2.0 * bv->mechBCLeft.value[1] - LA(vy, vyY, vyX + 1);
        vyll = 2.0 * bv->mechBCLeft.value[1] - LA(vy, vyY + 1, vyX + 1);
      } else if (!per && vyX == -1 && bv->mechBCLeft.type[1] == 1) {
        vyul = LA(vy, vyY, vyX + 1);
        vyll = LA(vy, vyY + 1, vyX + 1);
      } else {
        vyul = LA(vy, vyY, vyX);
        vyll = LA(vy, vyY + 1, vyX);
      }
      deltaxm = xcur - g->xc[vyX + 1];
      deltaym = ycur - g->y[vyY];
      dx = g->xc[vyX + 2] - g->xc[vyX + 1];
      dy = g->y[vyY + 1] - g->y[vyY];
      mvy[irk] = vyul * (1.0 - deltaxm / dx) * (1.0 - deltaym / dy);
      mvy[irk] += vyur * (deltaxm / dx) * (1.0 - deltaym / dy);
      mvy[irk] += vyll * (1.0 - deltaxm / dx) * (deltaym / dy);
      mvy[irk] += vylr * (deltaxm * deltaym) / dx / dy;

      const double mdx = (xcur - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
      const double mdy = (ycur - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
      mwxy[irk] = (1.0 - mdx) * (1.0 - mdy) * LA(wxy, cellY, cellX);
      mwxy[irk] += (mdx) * (1.0 - mdy) * LA(wxy, cellY, cellX + 1);
      mwxy[irk] += (1.0 - mdx) * (mdy) * LA(wxy, cellY + 1, cellX);
      mwxy[irk] += (mdx) * (mdy) * LA(wxy, cellY + 1, cellX + 1);

      if (irk < 3) {
        if (irk < 2) {
          xcur = mk[m].X + dt / 2 * mvx[irk];
          ycur = mk[m].Y + dt / 2 * mvy[irk];
        } else {
          xcur = mk[m].X + dt * mvx[irk];
          ycur = mk[m].Y + dt * mvy[irk];
        }
      }
    }
    mk[m].VX = (mvx[0] + 2 * mvx[1] + 2 * mvx[2] + mvx[3]) / 6;
    mk[m].VY = (mvy[0] + 2 * mvy[1] + 2 * mvy[2] + mvy[3]) / 6;
    mk[m].wxy = (mwxy[0] + 2 * mwxy[1] + 2 * mwxy[2] + mwxy[3]) / 6.0;
    mk[m].X += dt * mk[m].VX;
    mk[m].Y += dt * mk[m].VY;

    const double msxx0 = mk[m].s.T11, msyy0 = mk[m].s.T22, msxy0 = mk[m].s.T12;
    if (jaumann) {
      const double wxy1 = mk[m].wxy * dt;
      mk[m].s.T11 += 2.0 * msxy0 * -wxy1;
      mk[m].s.T22 += 2.0 * msxy0 * wxy1;
      mk[m].s.T12 += (msxx0 - msyy0) * wxy1;
    } else {
      const double thetaz = -mk[m].wxy * dt;
      mk[m].s.T11 = msxx0 * pow(cos(thetaz), 2.0) + msyy0 * pow(sin(thetaz), 2.0) - msxy0 * sin(2.0 * thetaz);
      mk[m].s.T22 = msxx0 * pow(sin(thetaz), 2.0) + msyy0 * pow(cos(thetaz), 2.0) + msxy0 * sin(2.0 * thetaz);
      mk[m].s.T12 = 0.5 * (msxx0 - msyy0) * sin(2.0 * thetaz) + msxy0 * cos(2.0 * thetaz);
    }
    mk[m].Eii += dt * sqrt(0.5 * (mk[m].e.T11 * mk[m].e.T11 + mk[m].e.T22 * mk[m].e.T22) + mk[m].e.T12 * mk[m].e.T12);
    if (per) { /* markers.c:2225-2231 */
      if (mk[m].X > g->LX) mk[m].X -= g->LX;
      else if (mk[m].X < 0) mk[m].X += g->LX;
    }
  }
  local_free(&vx);
  local_free(&vy);
  local_free(&wxy);
  return ORC_OK;
}

/* ---- subgridTemperatureChanges.c:5-124 --------------------------------- */
int orc_subgrid_temperature_changes(const double *thisT, const GridData *g, const NodalArrays *nfa,
                                    MarkerSet *ms, const Materials *mat, double dt, double dsubgrid) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  const size_t nn = (size_t)NX * NY;
  Local T, rho, Cp, kT;
  if (local_new(&T, NX, NY, g->xperiodic) || local_new(&rho, NX, NY, g->xperiodic) || local_new(&Cp, NX, NY, g->xperiodic) || local_new(&kT, NX, NY, g->xperiodic))
    return ORC_ERR_NOMEM;
  local_fill(&T, nfa->lastT);
  local_fill(&rho, nfa->rho);
  local_fill(&Cp, nfa->Cp);
  local_fill(&kT, nfa->kThermal);
  double *dTsub = (double *)malloc((size_t)(ms->nMark > 0 ? ms->nMark : 1) * sizeof(double));
  double *deltaT = (double *)malloc(nn * sizeof(double));
  if (!dTsub || !deltaT) return ORC_ERR_NOMEM;
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    dTsub[m] = 0.0;
    if (mk[m].cellX == -1) continue;
    const int cellY = mk[m].cellY, cellX = mk[m].cellX;
    const double dx = g->x[cellX + 1] - g->x[cellX];
    const double dy = g->y[cellY + 1] - g->y[cellY];
    const double mdx = (mk[m].X - g->x[cellX]) / dx;
    const double mdy = (mk[m].Y - g->y[cellY]) / dy;
    const double Tm = LA(T, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) + LA(T, cellY, cellX + 1) * mdx * (1.0 - mdy) +
                      LA(T, cellY + 1, cellX) * (1.0 - mdx) * (mdy) + LA(T, cellY + 1, cellX + 1) * mdx * mdy;
    const double rhom = mat->materialRho[0] + LA(rho, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) +
                        LA(rho, cellY, cellX + 1) * (mdx) * (1.0 - mdy) + LA(rho, cellY + 1, cellX) * (1.0 - mdx) * (mdy) +
                        LA(rho, cellY + 1, cellX + 1) * mdx * mdy;
    const double Cpm = LA(Cp, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) + LA(Cp, cellY, cellX + 1) * (mdx) * (1.0 - mdy) +
                       LA(Cp, cellY + 1, cellX) * (1.0 - mdx) * (mdy) + LA(Cp, cellY + 1, cellX + 1) * mdx * mdy;
    const double km = LA(kT, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) + LA(kT, cellY, cellX + 1) * (mdx) * (1.0 - mdy) +
                      LA(kT, cellY + 1, cellX) * (1.0 - mdx) * (mdy) + LA(kT, cellY + 1, cellX + 1) * mdx * mdy;
    const double dTdiff = Cpm * rhom / km / (2.0 / (dx * dx) + 2.0 / (dy * dy));
    dTsub[m] = (Tm - mk[m].T) * (1.0 - exp(-dsubgrid * dt / dTdiff));
  }
  int e = orc_project_markers_nodes_from_scalar(ms, g, dTsub, deltaT);
  if (e) return e;
  /* VecAXPBYPCZ(deltaT, 1, -1, -1, thisT, lastT): z = x + (-1)y + (-1)z */
  for (size_t i = 0; i < nn; i++) deltaT[i] = thisT[i] + -1.0 * nfa->lastT[i] + -1.0 * deltaT[i];
  Local dT;
  if (local_new(&dT, NX, NY, g->xperiodic)) return ORC_ERR_NOMEM;
  local_fill(&dT, deltaT);
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    const int cellY = mk[m].cellY, cellX = mk[m].cellX;
    const double dx = g->x[cellX + 1] - g->x[cellX];
    const double dy = g->y[cellY + 1] - g->y[cellY];
    const double mdx = (mk[m].X - g->x[cellX]) / dx;
    const double mdy = (mk[m].Y - g->y[cellY]) / dy;
    const double dtm = (LA(dT, cellY, cellX) * (1.0 - mdy) * (1.0 - mdx) + LA(dT, cellY + 1, cellX) * mdy * (1.0 - mdx) +
                        LA(dT, cellY, cellX + 1) * (1.0 - mdy) * mdx + LA(dT, cellY + 1, cellX + 1) * mdy * mdx) +
                       dTsub[m];
    mk[m].T += dtm;
    mk[m].Tdot = dtm / dt;
  }
  local_free(&dT);
  free(dTsub);
  free(deltaT);
  local_free(&T);
  local_free(&rho);
  local_free(&Cp);
  local_free(&kT);
  return ORC_OK;
}

/* ---- subgridStressChanges.c:5-318 -------------------------------------- */
int orc_subgrid_stress_changes(const GridData *g, NodalArrays *nfa, MarkerSet *ms, const Materials *mat,
                               double dt, double dsubgrid) {
  (void)mat;
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  const size_t nn = (size_t)NX * NY;
  Local soxx, soyy, soxy, dsxxn, dsyyn, dsxyn, wtetan, wtetas;
  if (local_new(&soxx, NX, NY, g->xperiodic) || local_new(&soyy, NX, NY, g->xperiodic) || local_new(&soxy, NX, NY, g->xperiodic) ||
      local_new(&dsxxn, NX, NY, g->xperiodic) || local_new(&dsyyn, NX, NY, g->xperiodic) || local_new(&dsxyn, NX, NY, g->xperiodic) ||
      local_new(&wtetan, NX, NY, g->xperiodic) || local_new(&wtetas, NX, NY, g->xperiodic))
    return ORC_ERR_NOMEM;
  local_fill(&soxx, nfa->soxx);
  local_fill(&soyy, nfa->soyy);
  local_fill(&soxy, nfa->soxy);
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    const double sdm = mk[m].eta / mk[m].mu;
    double sdif = -dsubgrid * dt / sdm;
    if (sdif < -30.0) sdif = -30.0;
    sdif = 1.0 - exp(sdif);

    int cellY = mk[m].cellY, cellX = mk[m].cellX;
    if (cellX < 0) cellX = 0;
    if (cellX > NX - 2) cellX = NX - 2;
    if (cellY < 0) cellY = 0;
    if (cellY > NY - 2) cellY = NY - 2;
    double mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    double mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    double sxym = 0.0;
    sxym += (1.0 - mdx) * (1.0 - mdy) * LA(soxy, cellY, cellX);
    sxym += (1.0 - mdx) * mdy * LA(soxy, cellY + 1, cellX);
    sxym += mdx * (1.0 - mdy) * LA(soxy, cellY, cellX + 1);
    sxym += mdx * mdy * LA(soxy, cellY + 1, cellX + 1);
    double dsxym = sxym - mk[m].s.T12;
    dsxym *= sdif;
    mk[m].s.T12 += dsxym;

    cellY = mk[m].cellY;
    cellX = mk[m].cellX;
    mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    if (mdx <= 0.5 && mdy <= 0.5) {
      LA(dsxyn, cellY, cellX) = LA(dsxyn, cellY, cellX) + (1.0 - mdx) * (1.0 - mdy) * dsxym;
      LA(wtetas, cellY, cellX) = LA(wtetas, cellY, cellX) + (1.0 - mdx) * (1.0 - mdy);
    }
    if (mdx <= 0.5 && mdy >= 0.5) {
      LA(dsxyn, cellY + 1, cellX) = LA(dsxyn, cellY + 1, cellX) + (1.0 - mdx) * (mdy) * dsxym;
      LA(wtetas, cellY + 1, cellX) = LA(wtetas, cellY + 1, cellX) + (1.0 - mdx) * (mdy);
    }
    if (mdx >= 0.5 && mdy <= 0.5) {
      LA(dsxyn, cellY, cellX + 1) = LA(dsxyn, cellY, cellX + 1) + (mdx) * (1.0 - mdy) * dsxym;
      LA(wtetas, cellY, cellX + 1) = LA(wtetas, cellY, cellX + 1) + (mdx) * (1.0 - mdy);
    }
    if (mdx >= 0.5 && mdy >= 0.5) {
      LA(dsxyn, cellY + 1, cellX + 1) = LA(dsxyn, cellY + 1, cellX + 1) + (mdx) * (mdy) * dsxym;
      LA(wtetas, cellY + 1, cellX + 1) = LA(wtetas, cellY + 1, cellX + 1) + (mdx) * (mdy);
    }

    cellX = mk[m].cellX;
    cellY = mk[m].cellY;
    if (cellX < NX - 1 && mk[m].X >= g->xc[cellX + 1]) cellX++;
    if (cellY < NY - 1 && mk[m].Y >= g->yc[cellY + 1]) cellY++;
    if (cellY < 1) cellY = 1;
    else if (cellY > NY - 2) cellY = NY - 2;
    if (cellX < 1) cellX = 1;
    else if (cellX > NX - 2) cellX = NX - 2;
    mdx = (mk[m].X - g->xc[cellX]) / (g->xc[cellX + 1] - g->xc[cellX]);
    mdy = (mk[m].Y - g->yc[cellY]) / (g->yc[cellY + 1] - g->yc[cellY]);
    double sxxm = 0.0;
    sxxm += (1.0 - mdx) * (1.0 - mdy) * LA(soxx, cellY, cellX);
    sxxm += (1.0 - mdx) * mdy * LA(soxx, cellY + 1, cellX);
    sxxm += mdx * (1.0 - mdy) * LA(soxx, cellY, cellX + 1);
    sxxm += mdx * mdy * LA(soxx, cellY + 1, cellX + 1);
    double dsxxm = sxxm - mk[m].s.T11;
    dsxxm *= sdif;
    mk[m].s.T11 += dsxxm;
    double syym = 0.0;
    syym += (1.0 - mdx) * (1.0 - mdy) * LA(soyy, cellY, cellX);
    syym += (1.0 - mdx) * mdy * LA(soyy, cellY + 1, cellX);
    syym += mdx * (1.0 - mdy) * LA(soyy, cellY, cellX + 1);
    syym += mdx * mdy * LA(soyy, cellY + 1, cellX + 1);
    double dsyym = syym - mk[m].s.T22;
    dsyym *= sdif;
    mk[m].s.T22 += dsyym;

    cellX = mk[m].cellX;
    cellY = mk[m].cellY;
    mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    LA(dsxxn, cellY + 1, cellX + 1) += (1.0 - fabs(0.5 - mdx)) * (1.0 - fabs(0.5 - mdy)) * dsxxm;
    LA(dsyyn, cellY + 1, cellX + 1) += (1.0 - fabs(0.5 - mdx)) * (1.0 - fabs(0.5 - mdy)) * dsyym;
    LA(wtetan, cellY + 1, cellX + 1) += (1.0 - fabs(0.5 - mdx)) * (1.0 - fabs(0.5 - mdy));
  }
  double *wn = (double *)calloc(nn, sizeof(double)), *ws = (double *)calloc(nn, sizeof(double));
  double *gx = (double *)calloc(nn, sizeof(double)), *gy = (double *)calloc(nn, sizeof(double));
  double *gxy = (double *)calloc(nn, sizeof(double));
  if (!wn || !ws || !gx || !gy || !gxy) return ORC_ERR_NOMEM;
  local_add_to_global(&wtetan, wn);
  local_add_to_global(&wtetas, ws);
  local_add_to_global(&dsxxn, gx);
  local_add_to_global(&dsyyn, gy);
  local_add_to_global(&dsxyn, gxy);
  for (size_t i = 0; i < nn; i++) {
    gxy[i] = gxy[i] / ws[i];
    gx[i] = gx[i] / wn[i];
    gy[i] = gy[i] / wn[i];
    nfa->dsxx[i] = nfa->dsxx[i] + -1.0 * gx[i];
    nfa->dsyy[i] = nfa->dsyy[i] + -1.0 * gy[i];
    nfa->dsxy[i] = nfa->dsxy[i] + -1.0 * gxy[i];
  }
  free(wn);
  free(ws);
  free(gx);
  free(gy);
  free(gxy);
  local_free(&soxx);
  local_free(&soyy);
  local_free(&soxy);
  local_free(&dsxxn);
  local_free(&dsyyn);
  local_free(&dsxyn);
  local_free(&wtetan);
  local_free(&wtetas);
  return ORC_OK;
}

/* ---- updateMarkerStress.c:5-163 ---------------------------------------- */
int orc_update_marker_stress(const GridData *g, const NodalArrays *nfa, MarkerSet *ms) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  Local dsxx, dsyy, dsxy;
  if (local_new(&dsxx, NX, NY, g->xperiodic) || local_new(&dsyy, NX, NY, g->xperiodic) || local_new(&dsxy, NX, NY, g->xperiodic)) return ORC_ERR_NOMEM;
  local_fill(&dsxx, nfa->dsxx);
  local_fill(&dsyy, nfa->dsyy);
  local_fill(&dsxy, nfa->dsxy);
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    int cellX = mk[m].cellX, cellY = mk[m].cellY;
    int cellXMax = g->xperiodic ? NX : NX - 2; /* updateMarkerStress.c:66 */
    if (cellX < cellXMax && mk[m].X >= g->xc[cellX + 1]) cellX++;
    if (cellY < NY - 1 && mk[m].Y >= g->yc[cellY + 1]) cellY++;
    if (cellY < 1) cellY = 1;
    else if (cellY > NY - 2) cellY = NY - 2;
    cellXMax = g->xperiodic ? NX - 1 : NX - 2;
    if (cellX < 1) cellX = 1;
    else if (cellX > cellXMax) cellX = cellXMax;
    double mdx = (mk[m].X - g->xc[cellX]) / (g->xc[cellX + 1] - g->xc[cellX]);
    double mdy = (mk[m].Y - g->yc[cellY]) / (g->yc[cellY + 1] - g->yc[cellY]);
    double mdsxx = 0.0;
    mdsxx += (1.0 - mdx) * (1.0 - mdy) * LA(dsxx, cellY, cellX);
    mdsxx += (1.0 - mdx) * mdy * LA(dsxx, cellY + 1, cellX);
    mdsxx += mdx * (1.0 - mdy) * LA(dsxx, cellY, cellX + 1);
    mdsxx += mdx * mdy * LA(dsxx, cellY + 1, cellX + 1);
    double mdsyy = 0.0;
    mdsyy += (1.0 - mdx) * (1.0 - mdy) * LA(dsyy, cellY, cellX);
    mdsyy += (1.0 - mdx) * mdy * LA(dsyy, cellY + 1, cellX);
    mdsyy += mdx * (1.0 - mdy) * LA(dsyy, cellY, cellX + 1);
    mdsyy += mdx * mdy * LA(dsyy, cellY + 1, cellX + 1);
    mk[m].s.T11 += mdsxx;
    mk[m].s.T22 += mdsyy;
    cellX = mk[m].cellX;
    cellY = mk[m].cellY;
    cellXMax = g->xperiodic ? NX - 1 : NX - 2;
    if (cellY < 0) cellY = 0;
    else if (cellY > NY - 2) cellY = NY - 2;
    if (cellX < 0) cellX = 0;
    else if (cellX > cellXMax) cellX = cellXMax;
    mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    double mdsxy = 0.0;
    mdsxy += (1.0 - mdx) * (1.0 - mdy) * LA(dsxy, cellY, cellX);
    mdsxy += (1.0 - mdx) * mdy * LA(dsxy, cellY + 1, cellX);
    mdsxy += mdx * (1.0 - mdy) * LA(dsxy, cellY, cellX + 1);
    mdsxy += mdx * mdy * LA(dsxy, cellY + 1, cellX + 1);
    mk[m].s.T12 += mdsxy;
  }
  local_free(&dsxx);
  local_free(&dsyy);
  local_free(&dsxy);
  return ORC_OK;
}

/* ---- updateMarkerStrainPressure.c:9-399 -------------------------------- */
static double bil_c(Local *f, int cellY, int cellX, double mdx, double mdy) {
  double v = 0.0;
  v += (1.0 - mdx) * (1.0 - mdy) * LA(*f, cellY, cellX);
  v += (1.0 - mdx) * mdy * LA(*f, cellY + 1, cellX);
  v += mdx * (1.0 - mdy) * LA(*f, cellY, cellX + 1);
  v += mdx * mdy * LA(*f, cellY + 1, cellX + 1);
  return v;
}

int orc_update_marker_strain_pressure(const GridData *g, const NodalArrays *nfa, MarkerSet *ms, double dt) {
  if (g->xperiodic && g->NX < 4) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  Local dsxx, dsyy, dsxy, soxx, soyy, soxy, exx, eyy, exy, wxy, p;
  if (local_new(&dsxx, NX, NY, g->xperiodic) || local_new(&dsyy, NX, NY, g->xperiodic) || local_new(&dsxy, NX, NY, g->xperiodic) ||
      local_new(&soxx, NX, NY, g->xperiodic) || local_new(&soyy, NX, NY, g->xperiodic) || local_new(&soxy, NX, NY, g->xperiodic) ||
      local_new(&exx, NX, NY, g->xperiodic) || local_new(&eyy, NX, NY, g->xperiodic) || local_new(&exy, NX, NY, g->xperiodic) ||
      local_new(&wxy, NX, NY, g->xperiodic) || local_new(&p, NX, NY, g->xperiodic))
    return ORC_ERR_NOMEM;
  local_fill(&dsxx, nfa->dsxx);
  local_fill(&dsyy, nfa->dsyy);
  local_fill(&dsxy, nfa->dsxy);
  local_fill(&soxx, nfa->soxx);
  local_fill(&soyy, nfa->soyy);
  local_fill(&soxy, nfa->soxy);
  local_fill(&exx, nfa->edotxx);
  local_fill(&eyy, nfa->edotyy);
  local_fill(&exy, nfa->edotxy);
  local_fill(&wxy, nfa->wxy);
  local_fill(&p, nfa->p);
  int err = ORC_OK;
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    int cellX = mk[m].cellX, cellY = mk[m].cellY;
    int cellXMax = g->xperiodic ? NX : NX - 2; /* updateMarkerStrainPressure.c:137 */
    if (cellX < cellXMax && mk[m].X >= g->xc[cellX + 1]) cellX++;
    if (cellY < NY - 1 && mk[m].Y >= g->yc[cellY + 1]) cellY++;
    if (cellY < 1) cellY = 1;
    else if (cellY > NY - 2) cellY = NY - 2;
    cellXMax = g->xperiodic ? NX - 1 : NX - 2;
    if (cellX < 1) cellX = 1;
    else if (cellX > cellXMax) cellX = cellXMax;
    double mdx = (mk[m].X - g->xc[cellX]) / (g->xc[cellX + 1] - g->xc[cellX]);
    double mdy = (mk[m].Y - g->yc[cellY]) / (g->yc[cellY + 1] - g->yc[cellY]);
    mk[m].p = bil_c(&p, cellY, cellX, mdx, mdy);
    const double mexx = bil_c(&exx, cellY, cellX, mdx, mdy);
    (void)bil_c(&eyy, cellY, cellX, mdx, mdy); /* meyy: computed, unused (:166-170) */
    const double msxx = bil_c(&soxx, cellY, cellX, mdx, mdy);
    const double msyy = bil_c(&soyy, cellY, cellX, mdx, mdy);
    const double mdsxx = bil_c(&dsxx, cellY, cellX, mdx, mdy);
    const double mdsyy = bil_c(&dsyy, cellY, cellX, mdx, mdy);

    cellX = mk[m].cellX;
    cellY = mk[m].cellY;
    cellXMax = g->xperiodic ? NX - 1 : NX - 2;
    if (cellY < 0) cellY = 0;
    else if (cellY > NY - 2) cellY = NY - 2;
    if (cellX < 0) cellX = 0;
    else if (cellX > cellXMax) cellX = cellXMax;
    mdx = (mk[m].X - g->x[cellX]) / (g->x[cellX + 1] - g->x[cellX]);
    mdy = (mk[m].Y - g->y[cellY]) / (g->y[cellY + 1] - g->y[cellY]);
    mk[m].wxy = bil_c(&wxy, cellY, cellX, mdx, mdy);
    const double mexy = bil_c(&exy, cellY, cellX, mdx, mdy);
    const double msxy = bil_c(&soxy, cellY, cellX, mdx, mdy);
    const double mdsxy = bil_c(&dsxy, cellY, cellX, mdx, mdy);

    mk[m].e.T11 = msxx / 2.0 / mk[m].eta + mdsxx / 2.0 / dt / mk[m].mu;
    mk[m].e.T22 = msyy / 2.0 / mk[m].eta + mdsyy / 2.0 / dt / mk[m].mu;
    mk[m].e.T12 = msxy / 2.0 / mk[m].eta + mdsxy / 2.0 / dt / mk[m].mu;
    mk[m].strainRateResidual = sqrt(pow(mk[m].e.T11 - mexx, 2.0) + pow(mk[m].e.T12 - mexy, 2.0));
    if (isnan(mk[m].strainRateResidual) && !err) err = ORC_ERR_NAN_RESIDUAL; /* :302-306 aborts */
  }
  local_free(&dsxx);
  local_free(&dsyy);
  local_free(&dsxy);
  local_free(&soxx);
  local_free(&soyy);
  local_free(&soxy);
  local_free(&exx);
  local_free(&eyy);
  local_free(&exy);
  local_free(&wxy);
  local_free(&p);
  return err;
}

/* ---- viscosity.c:12-67 ------------------------------------------------- */
int orc_compute_rheology(const Marker *mk, const Options *o, const Materials *mat, double *eta, double *mu) {
  const int mm = mk->Mat;
  double thiseta;
  const double thismu = mat->materialMu[mm];
  const int law = mat->hasEtaT[mm];
  if (law == 1) {
    thiseta = mat->materialEta[mm] * exp(mat->QonR[mm] * (1.0 / mk->T - 1.0 / mat->Tref[mm]));
  } else if (law == 2) {
    const double T0 = mat->Tref[mm];
    thiseta = mat->materialEta[mm] * exp(mat->QonR[mm] / T0 * (1.0 - (mk->T - T0) / T0));
  } else if (law == 3) {
    thiseta = mat->materialEta[mm] * mk->T;
  } else if (law == 5) {
    thiseta = mat->materialEta[mm] * exp(-mat->QonR[mm] * mk->T);
  } else if (law == 0) {
    thiseta = mat->materialEta[mm];
  } else if (law == 6) {
    const double depth = mk->Y, distance = mk->X;
    thiseta = mat->materialEta[mm];
    if (distance > depth / tan(o->slabAngle)) {
      if (distance < (depth / tan(o->slabAngle) + o->decoupleThickness)) {
        if (depth > o->minDecoupleDepth && depth < o->maxDecoupleDepth) thiseta *= o->decoupleEtaScaler;
      }
    }
  } else {
    return ORC_ERR_VISCOSITY_LAW; /* viscosity.c:60-62 aborts */
  }
  *eta = thiseta;
  *mu = thismu;
  return ORC_OK;
}

/* ---- markers.c:1381-1505 ----------------------------------------------- */
int orc_check_plastic_yielding(const GridData *g, MarkerSet *ms, const Materials *mat, double dt,
                               int *isYielding, const Options *o) {
  (void)g;
  const double etamin = o->etamin, etamax = o->etamax, fractionalEtamin = o->fractionalEtamin;
  int err = ORC_OK;
  Marker *mk = ms->markers;
  for (int m = 0; m < ms->nMark; m++) {
    if (mk[m].cellX == -1) continue;
    double etanew = mk[m].eta;
    const double X = mk[m].eta / (mk[m].mu * dt + mk[m].eta);
    const double sxxnew = mk[m].s.T11 * X + 2 * mk[m].eta * mk[m].e.T11 * (1.0 - X);
    const double syynew = mk[m].s.T22 * X + 2 * mk[m].eta * mk[m].e.T22 * (1.0 - X);
    const double sxynew = mk[m].s.T12 * X + 2 * mk[m].eta * mk[m].e.T12 * (1.0 - X);
    const double siinew = sqrt(0.5 * (sxxnew * sxxnew + syynew * syynew) + sxynew * sxynew);
    {
      double eta, mu;
      int e = orc_compute_rheology(&mk[m], o, mat, &eta, &mu);
      if (e) {
        if (!err) err = e;
        continue;
      }
      mk[m].eta = eta;
      mk[m].mu = mu;
    }
    if (mat->hasPlasticity[mk[m].Mat]) {
      const int thismat = mk[m].Mat;
      double siiyield;
      const int plast = mat->hasPlasticity[thismat];
      if (plast == 0) {
        siiyield = DBL_MAX;
      } else if (plast == 1) {
        siiyield = mat->materialCohesion[thismat] + mat->materialFriction[thismat] * mk[m].p;
      } else if (plast == 2) {
        const double mEii = mk[m].Eii;
        double c, f;
        if (mEii < mat->gamma[thismat][0]) {
          c = mat->C[thismat][0];
          f = sin(mat->F[thismat][0] / 180.0 * M_PI);
        } else if (mEii >= mat->gamma[thismat][0] && mEii <= mat->gamma[thismat][1]) {
          const double dg = (mEii - mat->gamma[thismat][0]) / (mat->gamma[thismat][1] - mat->gamma[thismat][0]);
          c = (mat->C[thismat][1] - mat->C[thismat][0]) * dg + mat->C[thismat][0];
          f = (sin(mat->F[thismat][1] / 180.0 * M_PI) - sin(mat->F[thismat][0] / 180.0 * M_PI)) * dg +
              sin(mat->F[thismat][0] / 180.0 * M_PI);
        } else {
          c = mat->C[thismat][1];
          f = sin(mat->F[thismat][1] / 180.0 * M_PI);
        }
        siiyield = c + f * mk[m].p;
      } else {
        siiyield = 0.0;
      }
      if (siiyield < 0.0) siiyield = 0.0;
      if (siinew > siiyield) {
        const double exx = mk[m].e.T11, eyy = mk[m].e.T22, exy = mk[m].e.T12;
        const double eiiold = sqrt(0.5 * (exx * exx + eyy * eyy) + exy * exy);
        etanew = siiyield / 2.0 / eiiold;
        if (etanew / mk[m].eta < fractionalEtamin) etanew = fractionalEtamin * mk[m].eta;
        if (etanew < etamin) etanew = etamin;
        else if (etanew > etamax) etanew = etamax;
        const double siiold =
            sqrt(0.5 * (mk[m].s.T11 * mk[m].s.T11 + mk[m].s.T22 * mk[m].s.T22) + mk[m].s.T12 * mk[m].s.T12);
        if (siiold != 0.0) {
          mk[m].s.T11 *= siiyield / siiold;
          mk[m].s.T22 *= siiyield / siiold;
          mk[m].s.T12 *= siiyield / siiold;
        }
        isYielding[0] = 1;
      }
      mk[m].eta = etanew;
    }
  }
  return err;
}

/* ---- viscosity.c:90-104 (every marker m < nMark, active or not) ---------- */
void orc_limit_viscosity(MarkerSet *ms, const Options *o) {
  for (int m = 0; m < ms->nMark; m++) {
    Marker *mk = &ms->markers[m];
    if (mk->eta > o->etamax) mk->eta = o->etamax;
    else if (mk->eta < o->etamin) mk->eta = o->etamin;
  }
}

/* ---- residual.c:12-51 --------------------------------------------------- */
int orc_update_global_strain_rate_residual(const GridData *g, const MarkerSet *ms, double *residual) {
  const int NX = g->NX, NY = g->NY;
  double *srr = (double *)malloc((size_t)NX * NY * sizeof(double));
  if (!srr) return ORC_ERR_NOMEM;
  int e = orc_project_marker_field(g, ms, NULL, offsetof(Marker, strainRateResidual), srr, -1, -1, ARITHMETIC);
  if (e) {
    free(srr);
    return e;
  }
  for (int ix = 0; ix < NX; ix++)
    for (int jy = 0; jy < NY; jy++) {
      double *v = &srr[(size_t)jy * NX + ix];
      if (ix > 0 && jy > 0) {
        const double dx = g->x[ix] - g->x[ix - 1];
        const double dy = g->y[jy] - g->y[jy - 1];
        *v *= dx * dy;
      } else {
        *v = 0.0;
      }
    }
  double sum = 0.0;
  for (size_t i = 0; i < (size_t)NX * NY; i++) sum += srr[i];
  sum /= (g->LX * g->LY);
  *residual = sum;
  free(srr);
  return ORC_OK;
}

/* ---- main_subduction.c glue loops --------------------------------------- */
void orc_update_marker_density(MarkerSet *ms, const Materials *mat) {
  for (int m = 0; m < ms->nMark; m++) {
    Marker *mk = &ms->markers[m];
    if (mk->cellX == -1) continue;
    mk->rho = -mat->materialRho[mk->Mat] * (mat->materialAlpha[mk->Mat] * mk->T);
    mk->rhodot = 0.0;
  }
}

double orc_max_marker_T(const MarkerSet *ms) {
  double maxT = 0.0;
  for (int m = 0; m < ms->nMark; m++)
    if (ms->markers[m].cellX != -1 && ms->markers[m].T > maxT) maxT = ms->markers[m].T;
  return maxT;
}

double orc_max_abs_sxx(const MarkerSet *ms) {
  double v = 0.0;
  for (int m = 0; m < ms->nMark; m++)
    if (ms->markers[m].cellX != -1 && fabs(ms->markers[m].s.T11) > v) v = fabs(ms->markers[m].s.T11);
  return v;
}

/* ---- nodal step between the Stokes solve and the marker updates ----------- */
/* nodalStressStrain.c:6-485.  Serial restatement over the whole grid (one rank:
 * every node owned; the ghosted local arrays read only owned nodes here).
 * PETSc vector operations are written out in their element order: VecAXPY
 * y = y + a*x, VecPointwiseDivide/Mult, VecScale, VecShift, VecAYPX(y,-1,x)
 * y = x - y.  Nodes that the interior loop (:370-394) does not visit get
 * eii = sii = 0 (the freshly duplicated local vectors are zero) and the heating
 * vector is zeroed first (:355), so it ends as the shear heating alone. */
static int vep_node_update(double eta, double mu, double dt, double *so, double e, double *ds) {
  double X = eta + dt * mu;           /* :93-94 VecCopy, VecAXPY */
  X = eta / X;                        /* :95 VecPointwiseDivide */
  double X1 = X * -2.0;               /* :97-98 VecCopy, VecScale */
  X1 = X1 + 2.0;                      /* :99 VecShift */
  X1 = eta * X1;                      /* :100 VecPointwiseMult */
  const double et = X1 * e;           /* :108 */
  double st = X * (*so);              /* :109 */
  st = st + 1.0 * et;                 /* :110 VecAXPY(stemp, 1.0, etemp) */
  *ds = st - *so;                     /* :112-113 VecCopy, VecAYPX(ds, -1, stemp) */
  *so = st;                           /* :114 */
  return 0;
}

int orc_nodal_stress_strain(const GridData *g, NodalArrays *nf, const BoundaryValues *bv, double dt,
                            double *nodalHeating) {
  if (g->xperiodic) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  /* the ghost velocities of :162-180 are defined for prescribed (0) and free-slip (1) sides only */
  if ((bv->mechBCLeft.type[1] != 0 && bv->mechBCLeft.type[1] != 1) ||
      (bv->mechBCTop.type[0] != 0 && bv->mechBCTop.type[0] != 1))
    return ORC_ERR_BC;
#define N2(a, jy, ix) ((a)[(size_t)(jy) * NX + (ix)])
  /* cell-centre deviatoric strain rates (:54-75) */
  for (int ix = 0; ix < NX; ix++)
    for (int jy = 0; jy < NY; jy++) {
      if (ix == 0 || jy == 0) {
        N2(nf->edotxx, jy, ix) = 0.0;
        N2(nf->edotyy, jy, ix) = 0.0;
      } else {
        const double dx = g->x[ix] - g->x[ix - 1];
        const double dy = g->y[jy] - g->y[jy - 1];
        const double divv = (N2(nf->vx, jy - 1, ix) - N2(nf->vx, jy - 1, ix - 1)) / dx +
                            (N2(nf->vy, jy, ix - 1) - N2(nf->vy, jy - 1, ix - 1)) / dy;
        N2(nf->edotxx, jy, ix) = (N2(nf->vx, jy - 1, ix) - N2(nf->vx, jy - 1, ix - 1)) / dx - divv / 3.0;
        N2(nf->edotyy, jy, ix) = (N2(nf->vy, jy, ix - 1) - N2(nf->vy, jy - 1, ix - 1)) / dy - divv / 3.0;
      }
    }
  /* normal stresses on the cell-centre nodes (:88-121) */
  for (size_t i = 0; i < (size_t)NX * NY; i++) {
    vep_node_update(nf->etaN[i], nf->muN[i], dt, &nf->soxx[i], nf->edotxx[i], &nf->dsxx[i]);
    vep_node_update(nf->etaN[i], nf->muN[i], dt, &nf->soyy[i], nf->edotyy[i], &nf->dsyy[i]);
  }
  /* vorticity and shear strain rate on the basic nodes (:157-236) */
  for (int jy = 0; jy < NY; jy++)
    for (int ix = 0; ix < NX; ix++) {
      double vym, vxm;
      if (ix == 0)
        vym = bv->mechBCLeft.type[1] == 0 ? 2.0 * bv->mechBCLeft.value[1] - N2(nf->vy, jy, ix) : N2(nf->vy, jy, ix);
      else
        vym = N2(nf->vy, jy, ix - 1);
      const double vyp = N2(nf->vy, jy, ix);
      if (jy == 0)
        vxm = bv->mechBCTop.type[0] == 0 ? 2.0 * bv->mechBCTop.value[0] - N2(nf->vx, jy, ix) : N2(nf->vx, jy, ix);
      else
        vxm = N2(nf->vx, jy - 1, ix);
      const double vxp = N2(nf->vx, jy, ix);
      const double dx = g->xc[ix + 1] - g->xc[ix];
      const double dy = g->yc[jy + 1] - g->yc[jy];
      N2(nf->wxy, jy, ix) = 0.5 * ((vxp - vxm) / dy - (vyp - vym) / dx);
      N2(nf->edotxy, jy, ix) = 0.5 * ((vxp - vxm) / dy + (vyp - vym) / dx);
    }
  /* shear stress on the basic nodes (:257-271) */
  for (size_t i = 0; i < (size_t)NX * NY; i++)
    vep_node_update(nf->etaS[i], nf->muS[i], dt, &nf->soxy[i], nf->edotxy[i], &nf->dsxy[i]);
  /* shear heating, second invariants (:355-394) */
  for (size_t i = 0; i < (size_t)NX * NY; i++) {
    nodalHeating[i] = 0.0;
    nf->eii[i] = 0.0;
    nf->sii[i] = 0.0;
  }
  for (int jy = 1; jy < NY - 1; jy++)
    for (int ix = 1; ix < NX - 1; ix++) {
      const double *sxx = nf->soxx, *syy = nf->soyy, *eN = nf->etaN, *exx = nf->edotxx, *eyy = nf->edotyy;
      double hs = (N2(sxx, jy, ix) * N2(sxx, jy, ix) / N2(eN, jy, ix) +
                   N2(sxx, jy + 1, ix) * N2(sxx, jy + 1, ix) / N2(eN, jy + 1, ix) +
                   N2(sxx, jy, ix + 1) * N2(sxx, jy, ix + 1) / N2(eN, jy, ix + 1) +
                   N2(sxx, jy + 1, ix + 1) * N2(sxx, jy + 1, ix + 1) / N2(eN, jy + 1, ix + 1)) /
                  8.0;
      hs = hs + (N2(syy, jy, ix) * N2(syy, jy, ix) / N2(eN, jy, ix) +
                 N2(syy, jy + 1, ix) * N2(syy, jy + 1, ix) / N2(eN, jy + 1, ix) +
                 N2(syy, jy, ix + 1) * N2(syy, jy, ix + 1) / N2(eN, jy, ix + 1) +
                 N2(syy, jy + 1, ix + 1) * N2(syy, jy + 1, ix + 1) / N2(eN, jy + 1, ix + 1)) /
                    8.0;
      hs = hs + N2(nf->soxy, jy, ix) * N2(nf->soxy, jy, ix) / N2(nf->etaS, jy, ix);
      N2(nodalHeating, jy, ix) = N2(nodalHeating, jy, ix) + hs;
      N2(nf->eii, jy, ix) =
          sqrt((N2(exx, jy, ix) * N2(exx, jy, ix) + N2(exx, jy + 1, ix) * N2(exx, jy + 1, ix) +
                N2(exx, jy, ix + 1) * N2(exx, jy, ix + 1) + N2(exx, jy + 1, ix + 1) * N2(exx, jy + 1, ix + 1)) /
                   8.0 +
               (N2(eyy, jy, ix) * N2(eyy, jy, ix) + N2(eyy, jy + 1, ix) * N2(eyy, jy + 1, ix) +
                N2(eyy, jy, ix + 1) * N2(eyy, jy, ix + 1) + N2(eyy, jy + 1, ix + 1) * N2(eyy, jy + 1, ix + 1)) /
                   8.0 +
               N2(nf->edotxy, jy, ix) * N2(nf->edotxy, jy, ix));
      N2(nf->sii, jy, ix) =
          sqrt((N2(sxx, jy, ix) * N2(sxx, jy, ix) + N2(sxx, jy + 1, ix) * N2(sxx, jy + 1, ix) +
                N2(sxx, jy, ix + 1) * N2(sxx, jy, ix + 1) + N2(sxx, jy + 1, ix + 1) * N2(sxx, jy + 1, ix + 1)) /
                   8.0 +
               (N2(syy, jy, ix) * N2(syy, jy, ix) + N2(syy, jy + 1, ix) * N2(syy, jy + 1, ix) +
                N2(syy, jy, ix + 1) * N2(syy, jy, ix + 1) + N2(syy, jy + 1, ix + 1) * N2(syy, jy + 1, ix + 1)) /
                   8.0 +
               N2(nf->soxy, jy, ix) * N2(nf->soxy, jy, ix));
    }
  return ORC_OK;
}

/* retrieveSolutions.c:4-82 (without TEXTURE: foundnan stays 0).  S is the
 * interleaved Stokes vector [jy][ix][vx, vy, p] (fdcode.h:242-245); its
 * pressure entries are scaled in place by Kcont (VecStrideScale, :10), then the
 * row-1 pressure mean weighted by dx/LX, summed in index order (:49-61), is
 * removed from every node with jy > 0 and ix > 0 (:65-75). */
int orc_retrieve_solutions(const GridData *g, NodalArrays *nf, double *S, double Kcont, int *foundnan) {
  if (g->xperiodic) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  const size_t nn = (size_t)NX * NY;
  for (size_t i = 0; i < nn; i++) S[3 * i + 2] = S[3 * i + 2] * Kcont;
  for (size_t i = 0; i < nn; i++) {
    nf->p[i] = S[3 * i + 2];
    nf->vx[i] = S[3 * i + 0];
    nf->vy[i] = S[3 * i + 1];
  }
  double p_mean = 0.0;
  if (NY > 1)
    for (int i = 1; i < NX; i++) {
      const double dx = g->x[i] - g->x[i - 1];
      p_mean += N2(nf->p, 1, i) * dx / g->LX;
    }
  for (int j = 1; j < NY; j++)
    for (int i = 1; i < NX; i++) N2(nf->p, j, i) -= p_mean;
  *foundnan = 0;
  return ORC_OK;
}

/* limitTimestep.c:6-133 limitDisplacementTimestep.  *dtdisp is the running
 * minimum over the nodes (vx, vy, density-change candidates; a zero velocity
 * or density rate offers no candidate), then the 1% density-change limit from
 * the extreme nodal rhodot (VecMax/VecMin, :112-121). */
int orc_limit_displacement_timestep(const GridData *g, const NodalArrays *nf, double *dtdisp,
                                    double dispFactor) {
  if (g->xperiodic) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  double d = *dtdisp;
  for (int ix = 0; ix < NX; ix++)
    for (int jy = 0; jy < NY; jy++) {
      double dx, dy, dt1;
      if (ix == 0) {
        dx = g->x[ix + 1] - g->x[ix];
      } else if (ix == NX - 1) {
        dx = g->x[ix] - g->x[ix - 1];
      } else {
        const double dx1 = g->x[ix + 1] - g->x[ix], dx2 = g->x[ix] - g->x[ix - 1];
        dx = dx1 <= dx2 ? dx1 : dx2;
      }
      dt1 = N2(nf->vx, jy, ix) != 0.0 ? dx / (fabs(N2(nf->vx, jy, ix))) * dispFactor : d;
      if (d > dt1) d = dt1;
      if (jy == 0) {
        dy = g->y[jy + 1] - g->y[jy];
      } else if (jy == NY - 1) {
        dy = g->y[jy] - g->y[jy - 1];
      } else {
        const double dy1 = g->y[jy + 1] - g->y[jy], dy2 = g->y[jy] - g->y[jy - 1];
        dy = dy1 <= dy2 ? dy1 : dy2;
      }
      dt1 = N2(nf->vy, jy, ix) != 0.0 ? dy / (fabs(N2(nf->vy, jy, ix))) * dispFactor : d;
      if (d > dt1) d = dt1;
      if (N2(nf->rhodot, jy, ix) != 0.0) dt1 = fabs(0.01 * N2(nf->rho, jy, ix) / N2(nf->rhodot, jy, ix));
      if (d > dt1) d = dt1;
    }
  /* VecMax / VecMin: the first entry, then strictly larger / smaller ones */
  const size_t nn = (size_t)NX * NY;
  double rmax = nf->rhodot[0], rmin = nf->rhodot[0];
  for (size_t i = 1; i < nn; i++) {
    if (nf->rhodot[i] > rmax) rmax = nf->rhodot[i];
    if (nf->rhodot[i] < rmin) rmin = nf->rhodot[i];
  }
  rmax = fabs(rmax);
  rmin = fabs(rmin);
  if (rmin > rmax) rmax = rmin;
  if (rmax * d > 0.01) d = 0.01 / rmax;
  *dtdisp = d;
  return ORC_OK;
}
#undef N2

/* ---- marker files (io.c:299-402, restart.c:8-125) ------------------------- */
/* One rank.  The file: int32 count of active markers (cellX != -1), double
 * elapsed time, then one column per field over the active markers in marker
 * order; the field order and types are those of saveMarkersBinary and of
 * post_processing/getBinaryMarkers.m. */
typedef struct {
  size_t off;
  int is_int;
} FileField;
static const FileField kFileFields[21] = {
    {offsetof(Marker, cpu), 1},   {offsetof(Marker, cellX), 1}, {offsetof(Marker, cellY), 1},
    {offsetof(Marker, X), 0},     {offsetof(Marker, Y), 0},     {offsetof(Marker, VX), 0},
    {offsetof(Marker, VY), 0},    {offsetof(Marker, Mat), 1},   {offsetof(Marker, T), 0},
    {offsetof(Marker, Tdot), 0},  {offsetof(Marker, eta), 0},   {offsetof(Marker, mu), 0},
    {offsetof(Marker, e.T11), 0}, {offsetof(Marker, e.T12), 0}, {offsetof(Marker, Eii), 0},
    {offsetof(Marker, s.T11), 0}, {offsetof(Marker, s.T22), 0}, {offsetof(Marker, s.T12), 0},
    {offsetof(Marker, p), 0},     {offsetof(Marker, rho), 0},   {offsetof(Marker, rhodot), 0}};

int orc_save_markers_binary(const MarkerSet *ms, const char *path, double elapsedTime) {
  FILE *f = fopen(path, "wb");
  if (!f) return ORC_ERR_IO;
  int32_t n = 0;
  for (int m = 0; m < ms->nMark; m++) n += ms->markers[m].cellX != -1;
  int ok = fwrite(&n, sizeof(int32_t), 1, f) == 1 && fwrite(&elapsedTime, sizeof(double), 1, f) == 1;
  for (int k = 0; k < 21 && ok; k++)
    for (int m = 0; m < ms->nMark && ok; m++) {
      if (ms->markers[m].cellX == -1) continue;
      const char *src = (const char *)&ms->markers[m] + kFileFields[k].off;
      ok = fwrite(src, kFileFields[k].is_int ? 4 : 8, 1, f) == 1;
    }
  return (fclose(f) == 0 && ok) ? ORC_OK : ORC_ERR_IO;
}

/* resetMarker (markers.c:250-330) on every slot of the capacity, then the file's
 * fields on the first n; strainRateResidual, which resetMarker leaves as it
 * was, is zeroed. */
int orc_restart_from_markers(MarkerSet *ms, const char *path, double *elapsedTime) {
  FILE *f = fopen(path, "rb");
  if (!f) return ORC_ERR_IO;
  int32_t n;
  if (fread(&n, sizeof(int32_t), 1, f) != 1 || fread(elapsedTime, sizeof(double), 1, f) != 1 || n < 0 ||
      n > ms->maxMark) {
    fclose(f);
    return ORC_ERR_IO;
  }
  for (int m = 0; m < ms->maxMark; m++) {
    Marker *mk = &ms->markers[m];
    memset(mk, 0, sizeof(*mk));
    mk->X = -1;
    mk->Y = -1;
    mk->Mat = -1;
  }
  int ok = 1;
  for (int k = 0; k < 21 && ok; k++)
    for (int m = 0; m < n && ok; m++) {
      char *dst = (char *)&ms->markers[m] + kFileFields[k].off;
      ok = fread(dst, kFileFields[k].is_int ? 4 : 8, 1, f) == 1;
    }
  if (ok && fgetc(f) != EOF) ok = 0; /* longer than the layout says */
  fclose(f);
  ms->nMark = n;
  ms->nOut = 0;
  return ok ? ORC_OK : ORC_ERR_IO;
}

/* ---- markers.c:455-781 checkMarkerDensity (one rank) ---------------------
 * The reference draws its random numbers from drand48 (which markers leave an
 * over-full quarter cell) and PetscRandom (where added markers go), neither of
 * which can be reproduced here; both are replaced by the counter-based stream
 * orc_pc_u01 below, which the GPU path uses as well:
 *   u(seed, stream, ctr) = (mix(mix(seed + G*(stream+1)) ^ ctr) >> 11) * 2^-53,
 *   mix = the SplitMix64 output function, G = 0x9E3779B97F4A7C15;
 *   stream 1, ctr (q << 32 | k): removal key of the k-th member of quarter q;
 *   stream 2 / 3, ctr iMark: x / y draw of the iMark-th added marker.
 * Everything else follows the reference: quarter cells in its order (j over
 * 2(NY-1) rows outside, i over 2(NX-1) columns inside), its target count
 * ceil(NMX*(NX-1)/2.0*dx/LX)*ceil(...) per under-full quarter, removal of the
 * lowest keys (first index on ties) with geometric membership, the out list in
 * index order followed by the removed markers, recycling from its end, the
 * slab and inflow markers of benchmarkInitialConditions.c:52-92, and the
 * nearest-neighbour copy over the original markers m1 < nMark with the
 * one-cell-width window (recycled slots already filled count as candidates).
 * Two departures, both where the reference has no defined result:
 *   - a quarter with fewer geometric members than markers to remove (members
 *     on x = LX or y = LY are counted but not members) loses only its members;
 *     the reference would re-append its first member or read an unset index;
 *   - cpu and strainRateResidual of a slab or inflow marker are 0 (the
 *     reference copies an uninitialised local).
 * The reference aborts where this returns ORC_ERR_NEGATIVE_ADD (the target of
 * an under-full quarter does not exceed its count) or ORC_ERR_CAPACITY. */
static uint64_t pc_mix(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

double orc_pc_u01(uint64_t seed, uint64_t stream, uint64_t ctr) {
  const uint64_t key = pc_mix(seed + 0x9E3779B97F4A7C15ull * (stream + 1));
  return (double)(pc_mix(key ^ ctr) >> 11) * (1.0 / 9007199254740992.0);
}

/* benchmarkInitialConditions.c:80-92 plate_depth */
static double pc_plate_depth(double x, const Options *o) {
  if (fabs(x - o->rootCenter) < o->rootWidth)
    return o->overridingPlateThickness +
           o->rootThickness * pow(cos(M_PI * (x - o->rootCenter) / o->rootWidth / 2.0), 2.0);
  return o->overridingPlateThickness;
}

/* markers.c:255-330 resetMarker */
static void pc_reset_marker(Marker *m) {
  memset(m, 0, sizeof(*m)); /* fields resetMarker leaves unset: 0 here */
  m->X = -1;
  m->Y = -1;
  m->Mat = -1;
}

/* benchmarkInitialConditions.c:52-72 new_slab_marker / new_inflow_marker */
static int pc_new_marker(Marker *m, double x, double y, int slab, const Options *o, const Materials *mat) {
  pc_reset_marker(m);
  m->X = x;
  m->Y = y;
  m->Mat = 0;
  if (slab) {
    const double Ts = 273, T0 = 1623, kappa = 0.7272e-6, s_in_year = 3.15576e7;
    const double slab_age_s = o->slabAge * 1.0e6 * s_in_year;
    m->T = Ts + (T0 - Ts) * erf((y - x * tan(o->slabAngle)) / (2.0 * sqrt(kappa * slab_age_s)));
  } else {
    m->T = 1573.0;
  }
  return orc_compute_rheology(m, o, mat, &m->eta, &m->mu);
}

int orc_check_marker_density(MarkerSet *ms, const GridData *g, const Options *o, const Materials *mat,
                             uint64_t seed, int *nAdded, int *nRemoved) {
  if (g->xperiodic) return ORC_ERR_PERIODIC;
  const int NX = g->NX, NY = g->NY;
  const int m = NX - 1, n = NY - 1, W = 2 * m;
  const int nq = 4 * m * n;
  const int nMark = ms->nMark;
  Marker *mk = ms->markers;
  int *cnt = (int *)calloc((size_t)nq, sizeof(int));
  int *isout = (int *)malloc((size_t)(nMark > 0 ? nMark : 1) * sizeof(int));
  if (!cnt || !isout) return ORC_ERR_NOMEM;
  int nOut = 0;
  for (int m1 = 0; m1 < nMark; m1++) {
    if (mk[m1].cellX != -1) {
      const int cx = 2 * mk[m1].cellX + (mk[m1].X >= g->xc[mk[m1].cellX + 1] ? 1 : 0);
      const int cy = 2 * mk[m1].cellY + (mk[m1].Y >= g->yc[mk[m1].cellY + 1] ? 1 : 0);
      cnt[cx + cy * W] += 1;
    } else {
      isout[nOut++] = m1;
    }
  }
  int nAdd = 0, nRemove = 0, err = ORC_OK;
  for (int j = 0; j < 2 * n; j++) {
    for (int i = 0; i < W; i++) {
      int *c = &cnt[i + W * j];
      if (*c < o->minMarkers) {
        double dy = g->y[j / 2 + 1] - g->y[j / 2];
        double dx = g->x[i / 2 + 1] - g->x[i / 2];
        dx /= 2.0;
        dy /= 2.0;
        const int mynmx = (int)ceil(o->NMX * (g->NX - 1) / 2.0 * dx / g->LX);
        const int mynmy = (int)ceil(o->NMY * (g->NY - 1) / 2.0 * dy / g->LY);
        const int mynadd = mynmx * mynmy - *c;
        if (mynadd <= 0) err = ORC_ERR_NEGATIVE_ADD;
        nAdd += mynadd;
        *c = mynadd;
      } else if (*c > o->maxMarkers) {
        const int r = *c - o->maxMarkers;
        nRemove += r;
        *c = -r;
      } else {
        *c = 0;
      }
    }
  }
  if (err) {
    free(cnt);
    free(isout);
    return err;
  }
  /* removals */
  const int nOutBeforeRemoval = nOut;
  isout = (int *)realloc(isout, (size_t)(nOut + nRemove + 1) * sizeof(int));
  if (!isout) return ORC_ERR_NOMEM;
  if (nRemove > 0) {
    int *idx = (int *)malloc((size_t)(nMark > 0 ? nMark : 1) * sizeof(int));
    double *key = (double *)malloc((size_t)(nMark > 0 ? nMark : 1) * sizeof(double));
    if (!idx || !key) return ORC_ERR_NOMEM;
    for (int j = 0; j < 2 * n; j++) {
      const double ymin = !(j % 2) ? g->y[j / 2] : g->yc[(j + 1) / 2];
      const double ymax = !(j % 2) ? g->yc[j / 2 + 1] : g->y[(j + 1) / 2];
      for (int i = 0; i < W; i++) {
        const double xmin = !(i % 2) ? g->x[i / 2] : g->xc[(i + 1) / 2];
        const double xmax = !(i % 2) ? g->xc[i / 2 + 1] : g->x[(i + 1) / 2];
        const int q = i + W * j;
        if (cnt[q] >= 0) continue;
        const int nToRemove = -cnt[q];
        int nCell = 0;
        for (int im = 0; im < nMark; im++)
          if (mk[im].X >= xmin && mk[im].X < xmax && mk[im].Y >= ymin && mk[im].Y < ymax) idx[nCell++] = im;
        for (int k = 0; k < nCell; k++) key[k] = orc_pc_u01(seed, 1, ((uint64_t)q << 32) | (uint64_t)k);
        for (int irem = 0; irem < nToRemove && irem < nCell; irem++) {
          double minv = 2.0;
          int minidx = -1;
          for (int k = 0; k < nCell; k++)
            if (key[k] < minv) {
              minv = key[k];
              minidx = k;
            }
          key[minidx] = 2.0;
          Marker *r = &mk[idx[minidx]];
          r->X = -1.0;
          r->Y = -1.0;
          r->cellX = -1;
          r->cellY = -1;
          isout[nOut++] = idx[minidx];
        }
      }
    }
    free(idx);
    free(key);
  }
  const int removedActual = nOut - nOutBeforeRemoval;
  int added = 0;
  if (nAdd > 0 && (nAdd + ms->nMark - nOut) < ms->maxMark) {
    double *newX = (double *)malloc((size_t)nAdd * sizeof(double));
    double *newY = (double *)malloc((size_t)nAdd * sizeof(double));
    if (!newX || !newY) return ORC_ERR_NOMEM;
    int iMark = 0;
    for (int j = 0; j < 2 * n; j++) {
      for (int i = 0; i < W; i++) {
        const int q = i + W * j;
        for (int k = 0; k < cnt[q]; k++) {
          double lo, hi;
          if (!(i % 2)) lo = g->x[i / 2], hi = g->xc[i / 2 + 1];
          else lo = g->xc[(i + 1) / 2], hi = g->x[(i + 1) / 2];
          newX[iMark] = (hi - lo) * orc_pc_u01(seed, 2, (uint64_t)iMark) + lo;
          if (!(j % 2)) lo = g->y[j / 2], hi = g->yc[j / 2 + 1];
          else lo = g->yc[(j + 1) / 2], hi = g->y[(j + 1) / 2];
          newY[iMark] = (hi - lo) * orc_pc_u01(seed, 3, (uint64_t)iMark) + lo;
          iMark++;
        }
      }
    }
    const double pd = pc_plate_depth(g->LX, o);
    for (int i = 0; i < nAdd; i++) {
      int id;
      if (nOut > 0) {
        id = isout[nOut - 1];
        nOut--;
      } else {
        id = ms->nMark;
        ms->nMark++;
      }
      if (newX[i] < g->x[1] || (newX[i] >= g->x[NX - 2] && newY[i] >= pd)) {
        const int e = pc_new_marker(&mk[id], newX[i], newY[i], newX[i] < g->x[1], o, mat);
        if (e && !err) err = e;
        find_marker_cell(&mk[id], g);
      } else {
        double mindist = g->LX * g->LY;
        int minidx = -1;
        for (int m1 = 0; m1 < nMark; m1++) {
          if (m1 != id && mk[m1].cellX != -1) {
            if (fabs(mk[m1].X - newX[i]) <= (g->x[mk[m1].cellX + 1] - g->x[mk[m1].cellX]) &&
                fabs(mk[m1].Y - newY[i]) <= (g->y[mk[m1].cellY + 1] - g->y[mk[m1].cellY])) {
              const double x1 = (mk[m1].X - newX[i]);
              const double y1 = (mk[m1].Y - newY[i]);
              const double r1 = sqrt(x1 * x1 + y1 * y1);
              if (r1 < mindist) {
                mindist = r1;
                minidx = m1;
              }
            }
          }
        }
        if (minidx != -1) {
          mk[id] = mk[minidx];
          mk[id].X = newX[i];
          mk[id].Y = newY[i];
        }
      }
    }
    added = nAdd;
    free(newX);
    free(newY);
  } else if ((nAdd + ms->nMark - nOut) > ms->maxMark) {
    err = ORC_ERR_CAPACITY;
  }
  free(cnt);
  free(isout);
  if (nAdded) *nAdded = added;
  if (nRemoved) *nRemoved = removedActual;
  return err;
}
