// Per-marker kernels of the hot path: RK4 advection, subgrid temperature and
// stress diffusion, marker stress / strain-rate / pressure updates, plastic
// yielding, viscosity limits, the strain-rate residual and the glue loops.
// Each kernel keeps the reference's expression order (see the file:line
// citations).  Periodic x grids follow the reference's xperiodic branches.
#include <float.h>

#include "marker_common.h"
#include "rheology.h"
#include "proj_common.h"

int fvm_fill(fvm_ctx *ctx, double *a, int64_t n, double v);  // project.hip

// ---- A6: advectMarkersRK (markers.c:1865-2250), body in marker_common.h ----------
// With LDS the grid coordinates x, y, xc, yc (2 NX + 2 NY + 2 doubles) are staged
// in LDS once per block, and each block walks `tiles` tiles of 256 markers so
// that the staging is amortised (2 for small grids, one more per ADV_TILE_BYTES
// of staged grid beyond that, at most 8); the per-marker arithmetic is unchanged.
// k_rotate_eii then rotates the stresses and updates Eii in a streaming pass.
#define ADV_TILE_BYTES 4096
// A periodic grid (PER) reads the grid coordinates at ghost indices -1 .. NX+1,
// so it takes the path without staging.
template <bool LDS, bool PER>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) k_advect_rk4(DGrid g, int32_t n, double *__restrict__ MX, double *__restrict__ MY,
                                                    double *__restrict__ MVX, double *__restrict__ MVY,
                                                    double *__restrict__ Mwxy,
                                                    const int32_t *__restrict__ cellXa, const int32_t *__restrict__ cellYa,
                                                    const double *__restrict__ vx, const double *__restrict__ vy,
                                                    const double *__restrict__ wxy, BoundaryValues bv, double dt,
                                                    double *__restrict__ MT, double *__restrict__ MTdot,
                                                    const double *__restrict__ dTn, const double *__restrict__ dTsub,
                                                    double dt_thermal, int tiles) {
  if (LDS) {
    extern __shared__ double sgrid[];
    for (int i = threadIdx.x; i < g.NX; i += blockDim.x) sgrid[i] = g.x[i];
    for (int i = threadIdx.x; i < g.NY; i += blockDim.x) sgrid[g.NX + i] = g.y[i];
    for (int i = threadIdx.x; i <= g.NX; i += blockDim.x) sgrid[g.NX + g.NY + i] = g.xc[i];
    for (int i = threadIdx.x; i <= g.NY; i += blockDim.x) sgrid[2 * g.NX + g.NY + 1 + i] = g.yc[i];
    __syncthreads();
    DGrid gl = g;
    gl.x = sgrid;
    gl.y = sgrid + g.NX;
    gl.xc = sgrid + g.NX + g.NY;
    gl.yc = sgrid + 2 * g.NX + g.NY + 1;
    const int64_t m0 = (int64_t)blockIdx.x * (256 * tiles) + threadIdx.x;
#pragma unroll 1
    for (int t = 0; t < tiles; t++) {
      const int64_t m = m0 + 256 * t;
      if (m < n)
        advect_marker<PER>(m, gl, MX, MY, MVX, MVY, Mwxy, cellXa, cellYa, vx, vy, wxy, bv, dt, MT, MTdot, dTn, dTsub,
                      dt_thermal);
    }
  } else {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    advect_marker<PER>(m, g, MX, MY, MVX, MVY, Mwxy, cellXa, cellYa, vx, vy, wxy, bv, dt, MT, MTdot, dTn, dTsub,
                  dt_thermal);
  }
}

__global__ void __launch_bounds__(256) k_rotate_eii(int32_t n, int ncx, int NY, const int32_t *__restrict__ cellXa,
                                                    const int32_t *__restrict__ cellYa,
                                                    const double *__restrict__ Mwxy, double *__restrict__ sxx,
                                                    double *__restrict__ syy, double *__restrict__ sxy,
                                                    const double *__restrict__ exx, const double *__restrict__ eyy,
                                                    const double *__restrict__ exy, double *__restrict__ Eii,
                                                    double dt, int jaumann) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  rotate_eii_marker(m, ncx, NY, cellXa, cellYa, Mwxy, sxx, syy, sxy, exx, eyy, exy, Eii, dt, jaumann);
}

static int launch_advect(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                         const BoundaryValues *bv, double dt, int jaumann, const double *deltaT, const double *dTsub,
                         double dt_thermal) {
  const int32_t n = mk->nMark;
  const size_t lds = (size_t)(2 * g->NX + 2 * g->NY + 2) * sizeof(double);
  double *T = deltaT ? mk->T : nullptr, *Tdot = deltaT ? mk->Tdot : nullptr;
  if (g->xperiodic) {
    k_advect_rk4<false, true><<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(
        dgrid(g), n, mk->X, mk->Y, mk->VX, mk->VY, mk->wxy, mk->cellX, mk->cellY, nf->vx, nf->vy, nf->wxy, *bv, dt, T,
        Tdot, deltaT, dTsub, dt_thermal, 1);
  } else if (lds <= 32 * 1024) {
    const int tiles = std::max(2, std::min(8, (int)(lds / ADV_TILE_BYTES)));
    k_advect_rk4<true, false><<<fvm_blocks(n, 256 * tiles), 256, lds, ctx->stream>>>(
        dgrid(g), n, mk->X, mk->Y, mk->VX, mk->VY, mk->wxy, mk->cellX, mk->cellY, nf->vx, nf->vy, nf->wxy, *bv, dt, T,
        Tdot, deltaT, dTsub, dt_thermal,
        tiles);
  } else {
    k_advect_rk4<false, false><<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(
        dgrid(g), n, mk->X, mk->Y, mk->VX, mk->VY, mk->wxy, mk->cellX, mk->cellY, nf->vx, nf->vy, nf->wxy, *bv, dt, T,
        Tdot, deltaT, dTsub, dt_thermal,
        1);
  }
  FVM_LAUNCH_CHECK(ctx);
  k_rotate_eii<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, fvm_ncx(g), g->NY, mk->cellX, mk->cellY, mk->wxy, mk->sxx,
                                                            mk->syy, mk->sxy, mk->exx, mk->eyy, mk->exy, mk->Eii, dt,
                                                            jaumann);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_advect_markers_rk(fvm_ctx *ctx, fvm_markers *mk, const NodalArrays *nf, const fvm_grid *g,
                                     const BoundaryValues *bv, double dt, int32_t jaumann) {
  if (!ctx || !mk || !nf || !bv || !nf->vx || !nf->vy || !nf->wxy) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (mk->nMark == 0) return FVM_OK;
  {
    FVM_PHASE(ctx, FVM_PHASE_ADVECT);
    FVM_TRY(launch_advect(ctx, g, mk, nf, bv, dt, jaumann, nullptr, nullptr, 0.0));
  }
  fvm_markers_moved(ctx);
  return FVM_OK;
}

// ---- A7: subgridTemperatureChanges (subgridTemperatureChanges.c:5-124) -------
__global__ void __launch_bounds__(256) k_subgrid_t2(DGrid g, int32_t n, const double *__restrict__ MX,
                                                    const double *__restrict__ MY, double *__restrict__ MT,
                                                    double *__restrict__ MTdot, const int32_t *__restrict__ cX,
                                                    const int32_t *__restrict__ cY, const double *__restrict__ dT,
                                                    const double *__restrict__ dTsub, double dt) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  subgrid_t2_marker(m, g, MX, MY, MT, MTdot, cX, cY, dT, dTsub, dt);
}

// Pass 1 and the dTsub projection of subgridTemperatureChanges; leaves dTsub
// (per marker) and deltaT (nodal) in the context scratch for pass 2.
int fvm_subgrid_t_pass1(fvm_ctx *ctx, const double *thisT, const fvm_grid *g, const NodalArrays *nf,
                        fvm_markers *mk, const Materials *mat, double dt, double dsubgrid, double **dTsub_out,
                        double **deltaT_out) {
  const int32_t n = mk->nMark;
  const int64_t nn = (int64_t)g->NX * g->NY;
  FVM_TRY(fvm_ws(ctx, ctx->scratch, ((size_t)n + (size_t)nn + 64) * sizeof(double)));
  double *dTsub = (double *)ctx->scratch.ptr;
  double *deltaT = dTsub + ((n + 31) / 32) * 32;
  ProjGroup pg;
  memset(&pg, 0, sizeof(pg));
  pg.stagx = pg.stagy = 0;
  pg.method = PM_ARITH;
  pg.nf = 1;
  pg.src[0].col = dTsub;
  pg.out[0] = deltaT;
  pg.epilogue = EPI_SUBGRID_T;
  pg.ea = thisT;
  pg.eb = nf->lastT;
  // pass 1 runs inside the projection's partials pass, one lattice cell (= the
  // markers' basic cell) per lane group; markers without a valid cell are in
  // no bucket and get no dTsub (pass 2 skips them too)
  SubT sb;
  sb.T = nf->lastT;
  sb.rho = nf->rho;
  sb.Cp = nf->Cp;
  sb.k = nf->kThermal;
  sb.MT = mk->T;
  sb.dTsub = dTsub;
  sb.rho0 = mat->materialRho[0];
  sb.dt = dt;
  sb.dsub = dsubgrid;
  FVM_TRY(fvm_project_groups(ctx, g, mk, 1, &pg, &sb));
  *dTsub_out = dTsub;
  *deltaT_out = deltaT;
  return FVM_OK;
}

extern "C" int fvm_subgrid_temperature_changes(fvm_ctx *ctx, const double *thisT, const fvm_grid *g,
                                               const NodalArrays *nf, fvm_markers *mk, const Materials *mat,
                                               double dt, double dsubgrid) {
  if (!ctx || !mk || !nf || !mat || !thisT || !nf->lastT || !nf->rho || !nf->Cp || !nf->kThermal)
    return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  double *dTsub = nullptr, *deltaT = nullptr;
  FVM_TRY(fvm_subgrid_t_pass1(ctx, thisT, g, nf, mk, mat, dt, dsubgrid, &dTsub, &deltaT));
  const int32_t n = mk->nMark;
  if (n > 0) {
    FVM_PHASE(ctx, FVM_PHASE_SUBGRID_T2);
    k_subgrid_t2<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(dgrid(g), n, mk->X, mk->Y, mk->T, mk->Tdot, mk->cellX,
                                                               mk->cellY, deltaT, dTsub, dt);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

// advectMarkersRK with subgridTemperatureChanges pass 2 fused in front of it
// (both start from the stored cell; T is not read by the advection).
int fvm_advect_with_subgrid_t2(fvm_ctx *ctx, fvm_markers *mk, const NodalArrays *nf, const fvm_grid *g,
                               const BoundaryValues *bv, double dt, const double *deltaT, const double *dTsub,
                               double dt_thermal) {
  if (mk->nMark == 0) return FVM_OK;
  {
    FVM_PHASE(ctx, FVM_PHASE_ADVECT);
    FVM_TRY(launch_advect(ctx, g, mk, nf, bv, dt, ctx->jaumann, deltaT, dTsub, dt_thermal));
  }
  fvm_markers_moved(ctx);
  return FVM_OK;
}

// ---- A8: subgridStressChanges (subgridStressChanges.c:5-318) ------------------
__global__ void __launch_bounds__(256) k_subgrid_s1(DGrid g, int32_t n, const double *__restrict__ MX,
                                                    const double *__restrict__ MY, const double *__restrict__ eta,
                                                    const double *__restrict__ mu, double *__restrict__ sxx,
                                                    double *__restrict__ syy, double *__restrict__ sxy,
                                                    const int32_t *__restrict__ cXa, const int32_t *__restrict__ cYa,
                                                    const double *__restrict__ soxx, const double *__restrict__ soyy,
                                                    const double *__restrict__ soxy, double dt, double dsubgrid,
                                                    double *__restrict__ dxy, double *__restrict__ dxx,
                                                    double *__restrict__ dyy) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int c0x = cXa[m], c0y = cYa[m];
  if (!cell_ok(c0x, c0y, g.ncx(), g.NY)) {
    dxy[m] = 0.0;
    dxx[m] = 0.0;
    dyy[m] = 0.0;
    return;
  }
  // no xperiodic branch in subgridStressChanges.c: the clamps keep both
  // gathers inside [0, NX-1] on a periodic grid too
  const int NX = g.NX, NY = g.NY, P = g.per;
  const double X = MX[m], Y = MY[m];
  const double sdm = eta[m] / mu[m];
  double sdif = -dsubgrid * dt / sdm;
  if (sdif < -30.0) sdif = -30.0;
  sdif = 1.0 - exp(sdif);

  int cellY = c0y, cellX = c0x;
  if (cellX < 0) cellX = 0;
  if (cellX > NX - 2) cellX = NX - 2;
  if (cellY < 0) cellY = 0;
  if (cellY > NY - 2) cellY = NY - 2;
  double mdx = (X - g.x[cellX]) / (g.x[cellX + 1] - g.x[cellX]);
  double mdy = (Y - g.y[cellY]) / (g.y[cellY + 1] - g.y[cellY]);
  double sxym = 0.0;
  sxym += (1.0 - mdx) * (1.0 - mdy) * gat(soxy, NX, NY, P, cellY, cellX);
  sxym += (1.0 - mdx) * mdy * gat(soxy, NX, NY, P, cellY + 1, cellX);
  sxym += mdx * (1.0 - mdy) * gat(soxy, NX, NY, P, cellY, cellX + 1);
  sxym += mdx * mdy * gat(soxy, NX, NY, P, cellY + 1, cellX + 1);
  double sxy0 = sxy[m];
  double dsxym = sxym - sxy0;
  dsxym *= sdif;
  sxy[m] = sxy0 + dsxym;
  dxy[m] = dsxym;

  cellX = c0x;
  cellY = c0y;
  if (cellX < NX - 1 && X >= g.xc[cellX + 1]) cellX++;
  if (cellY < NY - 1 && Y >= g.yc[cellY + 1]) cellY++;
  if (cellY < 1) cellY = 1;
  else if (cellY > NY - 2) cellY = NY - 2;
  if (cellX < 1) cellX = 1;
  else if (cellX > NX - 2) cellX = NX - 2;
  mdx = (X - g.xc[cellX]) / (g.xc[cellX + 1] - g.xc[cellX]);
  mdy = (Y - g.yc[cellY]) / (g.yc[cellY + 1] - g.yc[cellY]);
  const double sxxm = bil_acc(soxx, NX, NY, P, cellY, cellX, mdx, mdy);
  const double sxx0 = sxx[m];
  double dsxxm = sxxm - sxx0;
  dsxxm *= sdif;
  sxx[m] = sxx0 + dsxxm;
  const double syym = bil_acc(soyy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double syy0 = syy[m];
  double dsyym = syym - syy0;
  dsyym *= sdif;
  syy[m] = syy0 + dsyym;
  dxx[m] = dsxxm;
  dyy[m] = dsyym;
}

extern "C" int fvm_subgrid_stress_changes(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, fvm_markers *mk,
                                          const Materials *mat, double dt, double dsubgrid) {
  (void)mat;
  if (!ctx || !mk || !nf || !nf->soxx || !nf->soyy || !nf->soxy || !nf->dsxx || !nf->dsyy || !nf->dsxy)
    return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int32_t n = mk->nMark;
  const size_t cs = ((size_t)n + 31) / 32 * 32;
  FVM_TRY(fvm_ws(ctx, ctx->scratch, (3 * cs + 64) * sizeof(double)));
  double *dxy = (double *)ctx->scratch.ptr, *dxx = dxy + cs, *dyy = dxx + cs;
  if (n > 0) {
    {
      FVM_PHASE(ctx, FVM_PHASE_SUBGRID_S1);
      k_subgrid_s1<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(dgrid(g), n, mk->X, mk->Y, mk->eta, mk->mu, mk->sxx,
                                                                 mk->syy, mk->sxy, mk->cellX, mk->cellY, nf->soxx,
                                                                 nf->soyy, nf->soxy, dt, dsubgrid, dxy, dxx, dyy);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  // subgridStressChanges.c:134-175 (sxy, basic nodes, <=/>= 0.5 ties) and
  // :223-233 (sxx, syy, one cell-centre node); two groups, one node pass
  ProjGroup pg[2];
  memset(pg, 0, sizeof(pg));
  pg[0].stagx = pg[0].stagy = 0;
  pg[0].method = PM_TIES;
  pg[0].nf = 1;
  pg[0].src[0].col = dxy;
  pg[0].out[0] = nf->dsxy;
  pg[0].epilogue = EPI_SUB_FROM;
  pg[1].stagx = pg[1].stagy = 0;
  pg[1].method = PM_CENTRE;
  pg[1].nf = 2;
  pg[1].src[0].col = dxx;
  pg[1].src[1].col = dyy;
  pg[1].out[0] = nf->dsxx;
  pg[1].out[1] = nf->dsyy;
  pg[1].epilogue = EPI_SUB_FROM;
  FVM_TRY(fvm_project_groups(ctx, g, mk, 2, pg));
  return FVM_OK;
}

// ---- A9: updateMarkerStress (updateMarkerStress.c:5-163) ----------------------
__global__ void __launch_bounds__(256) k_update_stress(DGrid g, int32_t n, const double *__restrict__ MX,
                                                       const double *__restrict__ MY, double *__restrict__ sxx,
                                                       double *__restrict__ syy, double *__restrict__ sxy,
                                                       const int32_t *__restrict__ cXa, const int32_t *__restrict__ cYa,
                                                       const double *__restrict__ dsxx, const double *__restrict__ dsyy,
                                                       const double *__restrict__ dsxy) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int c0x = cXa[m], c0y = cYa[m];
  if (!cell_ok(c0x, c0y, g.ncx(), g.NY)) return;
  const int NX = g.NX, NY = g.NY, P = g.per;
  const double X = MX[m], Y = MY[m];
  int cellX = c0x, cellY = c0y;
  // periodic: the centre cell may step to NX and is then clamped to NX-1
  // (cellXMax NX, then NX-1); node column NX is column 0
  if (cellX < (P ? NX : NX - 2) && X >= g.xc[cellX + 1]) cellX++;
  if (cellY < NY - 1 && Y >= g.yc[cellY + 1]) cellY++;
  if (cellY < 1) cellY = 1;
  else if (cellY > NY - 2) cellY = NY - 2;
  if (cellX < 1) cellX = 1;
  else if (cellX > (P ? NX - 1 : NX - 2)) cellX = P ? NX - 1 : NX - 2;
  double mdx = (X - g.xc[cellX]) / (g.xc[cellX + 1] - g.xc[cellX]);
  double mdy = (Y - g.yc[cellY]) / (g.yc[cellY + 1] - g.yc[cellY]);
  const double mdsxx = bil_acc(dsxx, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mdsyy = bil_acc(dsyy, NX, NY, P, cellY, cellX, mdx, mdy);
  sxx[m] = sxx[m] + mdsxx;
  syy[m] = syy[m] + mdsyy;
  cellX = c0x;
  cellY = c0y;
  if (cellY < 0) cellY = 0;
  else if (cellY > NY - 2) cellY = NY - 2;
  if (cellX < 0) cellX = 0;
  else if (cellX > (P ? NX - 1 : NX - 2)) cellX = P ? NX - 1 : NX - 2;
  mdx = (X - g.x[cellX]) / (g.x[cellX + 1] - g.x[cellX]);
  mdy = (Y - g.y[cellY]) / (g.y[cellY + 1] - g.y[cellY]);
  const double mdsxy = bil_acc(dsxy, NX, NY, P, cellY, cellX, mdx, mdy);
  sxy[m] = sxy[m] + mdsxy;
}

extern "C" int fvm_update_marker_stress(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, fvm_markers *mk,
                                        const Materials *mat) {
  (void)mat;
  if (!ctx || !mk || !nf || !nf->dsxx || !nf->dsyy || !nf->dsxy) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (mk->nMark == 0) return FVM_OK;
  {
    FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
    k_update_stress<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(dgrid(g), mk->nMark, mk->X, mk->Y, mk->sxx,
                                                                           mk->syy, mk->sxy, mk->cellX, mk->cellY,
                                                                           nf->dsxx, nf->dsyy, nf->dsxy);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

// ---- A10: updateMarkerStrainPressure (updateMarkerStrainPressure.c:9-399) -----
struct StrainIn {
  const double *dsxx, *dsyy, *dsxy, *soxx, *soyy, *soxy, *exx, *eyy, *exy, *wxy, *p;
};
__global__ void __launch_bounds__(256) k_strain_pressure(DGrid g, int32_t n, const double *__restrict__ MX,
                                                         const double *__restrict__ MY, const double *__restrict__ eta,
                                                         const double *__restrict__ mu, double *__restrict__ Mp,
                                                         double *__restrict__ Mwxy, double *__restrict__ e11,
                                                         double *__restrict__ e22, double *__restrict__ e12,
                                                         double *__restrict__ srr, const int32_t *__restrict__ cXa,
                                                         const int32_t *__restrict__ cYa, StrainIn F, double dt,
                                                         int32_t *__restrict__ flags) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int c0x = cXa[m], c0y = cYa[m];
  if (!cell_ok(c0x, c0y, g.ncx(), g.NY)) return;
  const int NX = g.NX, NY = g.NY, P = g.per;
  const double X = MX[m], Y = MY[m];
  int cellX = c0x, cellY = c0y;
  // periodic: the centre cell may step to NX and is then clamped to NX-1
  // (cellXMax NX, then NX-1); node column NX is column 0
  if (cellX < (P ? NX : NX - 2) && X >= g.xc[cellX + 1]) cellX++;
  if (cellY < NY - 1 && Y >= g.yc[cellY + 1]) cellY++;
  if (cellY < 1) cellY = 1;
  else if (cellY > NY - 2) cellY = NY - 2;
  if (cellX < 1) cellX = 1;
  else if (cellX > (P ? NX - 1 : NX - 2)) cellX = P ? NX - 1 : NX - 2;
  double mdx = (X - g.xc[cellX]) / (g.xc[cellX + 1] - g.xc[cellX]);
  double mdy = (Y - g.yc[cellY]) / (g.yc[cellY + 1] - g.yc[cellY]);
  Mp[m] = bil_acc(F.p, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mexx = bil_acc(F.exx, NX, NY, P, cellY, cellX, mdx, mdy);
  const double msxx = bil_acc(F.soxx, NX, NY, P, cellY, cellX, mdx, mdy);
  const double msyy = bil_acc(F.soyy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mdsxx = bil_acc(F.dsxx, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mdsyy = bil_acc(F.dsyy, NX, NY, P, cellY, cellX, mdx, mdy);
  cellX = c0x;
  cellY = c0y;
  if (cellY < 0) cellY = 0;
  else if (cellY > NY - 2) cellY = NY - 2;
  if (cellX < 0) cellX = 0;
  else if (cellX > (P ? NX - 1 : NX - 2)) cellX = P ? NX - 1 : NX - 2;
  mdx = (X - g.x[cellX]) / (g.x[cellX + 1] - g.x[cellX]);
  mdy = (Y - g.y[cellY]) / (g.y[cellY + 1] - g.y[cellY]);
  Mwxy[m] = bil_acc(F.wxy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mexy = bil_acc(F.exy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double msxy = bil_acc(F.soxy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double mdsxy = bil_acc(F.dsxy, NX, NY, P, cellY, cellX, mdx, mdy);
  const double em = eta[m], um = mu[m];
  const double t11 = msxx / 2.0 / em + mdsxx / 2.0 / dt / um;
  const double t22 = msyy / 2.0 / em + mdsyy / 2.0 / dt / um;
  const double t12 = msxy / 2.0 / em + mdsxy / 2.0 / dt / um;
  e11[m] = t11;
  e22[m] = t22;
  e12[m] = t12;
  // pow(v, 2.0) folded to v*v as in the reference's gcc -O3 build
  const double d11 = t11 - mexx, d12 = t12 - mexy;
  const double r = sqrt(d11 * d11 + d12 * d12);
  srr[m] = r;
  // the flag only goes from 0 to 1: a plain store, never an atomic per marker
  if (r != r) __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

extern "C" int fvm_update_marker_strain_pressure(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                                                 fvm_markers *mk, const Materials *mat, const Options *opt,
                                                 double dt) {
  (void)mat;
  (void)opt;
  if (!ctx || !mk || !nf) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  StrainIn F = {nf->dsxx, nf->dsyy, nf->dsxy, nf->soxx, nf->soyy, nf->soxy,
                nf->edotxx, nf->edotyy, nf->edotxy, nf->wxy, nf->p};
  if (!F.dsxx || !F.dsyy || !F.dsxy || !F.soxx || !F.soyy || !F.soxy || !F.exx || !F.exy || !F.wxy || !F.p)
    return FVM_ERR_ARG;
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096));
  int32_t *flags = (int32_t *)ctx->red.ptr + 8;
  FVM_HIP(ctx, hipMemsetAsync(flags, 0, sizeof(int32_t), ctx->stream));
  if (mk->nMark > 0) {
    {
      FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
      k_strain_pressure<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(
          dgrid(g), mk->nMark, mk->X, mk->Y, mk->eta, mk->mu, mk->p, mk->wxy, mk->exx, mk->eyy, mk->exy,
          mk->strainRateResidual, mk->cellX, mk->cellY, F, dt, flags);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  int32_t *h = (int32_t *)ctx->pinned.ptr;
  FVM_HIP(ctx, hipMemcpyAsync(h, flags, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (h[0]) return fvm_set_err(ctx, FVM_ERR_NAN_RESIDUAL, "NaN strain rate residual on an active marker");
  return FVM_OK;
}

// ---- A11: checkPlasticYielding + computeRheology (markers.c:1381-1505,
//           viscosity.c:12-75) ------------------------------------------------
// checkPlasticYielding's per-marker body (markers.c:1381-1505 with the viscosity
// law of viscosity.c:12-104): 0, 1 when the marker yields, 2 when its material
// or viscosity law is undefined.
__device__ __forceinline__ int yield_marker(int64_t m, const YieldParams &P, const double *__restrict__ MX,
                                            const double *__restrict__ MY, const double *__restrict__ MT,
                                            const int32_t *__restrict__ Mat, double *__restrict__ eta,
                                            double *__restrict__ mu, double *__restrict__ sxx,
                                            double *__restrict__ syy, double *__restrict__ sxy,
                                            const double *__restrict__ exx, const double *__restrict__ eyy,
                                            const double *__restrict__ exy, const double *__restrict__ Mp,
                                            const double *__restrict__ MEii, const int32_t *__restrict__ cX) {
  if (cX[m] == -1) return 0;
  const int mm = Mat[m];
  if ((unsigned)mm >= FVM_MAXMAT) return 2;
  const double eta0 = eta[m], mu0 = mu[m];
  double s11 = sxx[m], s22 = syy[m], s12 = sxy[m];
  const double e11 = exx[m], e22 = eyy[m], e12 = exy[m];
  double etanew = eta0;
  int yielded = 0;
  const double X = eta0 / (mu0 * P.dt + eta0);
  const double sxxnew = s11 * X + 2 * eta0 * e11 * (1.0 - X);
  const double syynew = s22 * X + 2 * eta0 * e22 * (1.0 - X);
  const double sxynew = s12 * X + 2 * eta0 * e12 * (1.0 - X);
  const double siinew = sqrt(0.5 * (sxxnew * sxxnew + syynew * syynew) + sxynew * sxynew);
  double etaL, muL;
  if (rheology(P, mm, MT[m], MX[m], MY[m], &etaL, &muL)) return 2;
  mu[m] = muL;
  double etaOut = etaL;
  const Materials &mat = P.mat;
  if (mat.hasPlasticity[mm]) {
    double siiyield;
    const int plast = mat.hasPlasticity[mm];
    const double pm = Mp[m];
    if (plast == 0) {
      siiyield = DBL_MAX;
    } else if (plast == 1) {
      siiyield = mat.materialCohesion[mm] + mat.materialFriction[mm] * pm;
    } else if (plast == 2) {
      const double mEii = MEii[m];
      double c, f;
      if (mEii < mat.gamma[mm][0]) {
        c = mat.C[mm][0];
        f = P.sinF[mm][0];
      } else if (mEii >= mat.gamma[mm][0] && mEii <= mat.gamma[mm][1]) {
        const double dg = (mEii - mat.gamma[mm][0]) / (mat.gamma[mm][1] - mat.gamma[mm][0]);
        c = (mat.C[mm][1] - mat.C[mm][0]) * dg + mat.C[mm][0];
        f = (P.sinF[mm][1] - P.sinF[mm][0]) * dg + P.sinF[mm][0];
      } else {
        c = mat.C[mm][1];
        f = P.sinF[mm][1];
      }
      siiyield = c + f * pm;
    } else {
      siiyield = 0.0;
    }
    if (siiyield < 0.0) siiyield = 0.0;
    if (siinew > siiyield) {
      const double eiiold = sqrt(0.5 * (e11 * e11 + e22 * e22) + e12 * e12);
      etanew = siiyield / 2.0 / eiiold;
      // markers[m].eta already holds the law's value here (markers.c:1410)
      if (etanew / etaL < P.fractionalEtamin) etanew = P.fractionalEtamin * etaL;
      if (etanew < P.etamin) etanew = P.etamin;
      else if (etanew > P.etamax) etanew = P.etamax;
      const double siiold = sqrt(0.5 * (s11 * s11 + s22 * s22) + s12 * s12);
      if (siiold != 0.0) {
        s11 *= siiyield / siiold;
        s22 *= siiyield / siiold;
        s12 *= siiyield / siiold;
        sxx[m] = s11;
        syy[m] = s22;
        sxy[m] = s12;
      }
      yielded = 1;
    }
    etaOut = etanew;  // markers.c:1493 (the pre-call eta when not yielding)
  }
  eta[m] = etaOut;
  return yielded;
}

__global__ void __launch_bounds__(256) k_yield(int32_t n, YieldParams P, const double *__restrict__ MX,
                                               const double *__restrict__ MY, const double *__restrict__ MT,
                                               const int32_t *__restrict__ Mat, double *__restrict__ eta,
                                               double *__restrict__ mu, double *__restrict__ sxx,
                                               double *__restrict__ syy, double *__restrict__ sxy,
                                               const double *__restrict__ exx, const double *__restrict__ eyy,
                                               const double *__restrict__ exy, const double *__restrict__ Mp,
                                               const double *__restrict__ MEii, const int32_t *__restrict__ cX,
                                               int32_t *__restrict__ flags) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int r = yield_marker(m, P, MX, MY, MT, Mat, eta, mu, sxx, syy, sxy, exx, eyy, exy, Mp, MEii, cX);
  // The flags only ever go from 0 to 1, so a wavefront that has any raises them
  // with a plain store: every writer stores the same value and the host reads
  // after the kernel.  (A device-scope atomic per wavefront on one word is
  // serialised where it executes: 125k of them cost ~0.7 ms at 8M markers.)
  const unsigned long long yb = __ballot(r == 1), bb = __ballot(r == 2);
  if ((yb | bb) && __lane_id() == __ffsll((unsigned long long)__ballot(1)) - 1) {
    if (yb) __hip_atomic_store(&flags[0], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (bb) __hip_atomic_store(&flags[1], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  }
}

extern "C" int fvm_check_plastic_yielding(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const Materials *mat,
                                          double dt, int32_t *isYielding, const Options *opt, double elapsedTime) {
  (void)g;
  (void)elapsedTime;
  if (!ctx || !mk || !mat || !opt) return FVM_ERR_ARG;
  const YieldParams P = make_yield_params(mat, opt, dt);
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096));
  int32_t *flags = (int32_t *)ctx->red.ptr + 16;
  FVM_HIP(ctx, hipMemsetAsync(flags, 0, 2 * sizeof(int32_t), ctx->stream));
  if (mk->nMark > 0) {
    {
      FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
      k_yield<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(mk->nMark, P, mk->X, mk->Y, mk->T, mk->Mat, mk->eta,
                                                                     mk->mu, mk->sxx, mk->syy, mk->sxy, mk->exx,
                                                                     mk->eyy, mk->exy, mk->p, mk->Eii, mk->cellX, flags);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  int32_t *h = (int32_t *)ctx->pinned.ptr;
  FVM_HIP(ctx, hipMemcpyAsync(h, flags, 2 * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (isYielding && h[0]) *isYielding = 1;
  if (h[1]) return fvm_set_err(ctx, FVM_ERR_VISCOSITY_LAW, "viscosity law undefined for a marker material");
  return FVM_OK;
}

__global__ void k_limit_viscosity(int32_t n, double *__restrict__ eta, double etamin, double etamax) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const double e = eta[m];
  if (e > etamax) eta[m] = etamax;
  else if (e < etamin) eta[m] = etamin;
}

extern "C" int fvm_limit_viscosity(fvm_ctx *ctx, fvm_markers *mk, const Options *opt) {
  if (!ctx || !mk || !opt) return FVM_ERR_ARG;
  if (mk->nMark == 0) return FVM_OK;
  {
    FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
    k_limit_viscosity<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(mk->nMark, mk->eta, opt->etamin,
                                                                             opt->etamax);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

// ---- deterministic sums ----------------------------------------------------------
#define RED_BS 256
__global__ void __launch_bounds__(RED_BS) k_block_sum(const double *__restrict__ a, int64_t n, int64_t per,
                                                      double *__restrict__ out) {
  __shared__ double s[RED_BS];
  const int64_t b0 = (int64_t)blockIdx.x * per;
  const int64_t b1 = b0 + per < n ? b0 + per : n;
  double acc = 0.0;
  for (int64_t i = b0 + threadIdx.x; i < b1; i += RED_BS) acc += a[i];
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int off = RED_BS / 2; off > 0; off >>= 1) {
    if (threadIdx.x < (unsigned)off) s[threadIdx.x] += s[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = s[0];
}

int fvm_reduce_sum(fvm_ctx *ctx, const double *d, int64_t n, double *out) {
  const int64_t nb = n < 1024 * RED_BS ? (n + RED_BS - 1) / RED_BS : 1024;
  const int64_t per = nb > 0 ? (n + nb - 1) / nb : 0;
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096 + (size_t)(nb + 1) * sizeof(double)));
  double *parts = (double *)((char *)ctx->red.ptr + 4096);
  if (nb > 0) {
    {
      FVM_PHASE(ctx, FVM_PHASE_OTHER);
      k_block_sum<<<(unsigned)nb, RED_BS, 0, ctx->stream>>>(d, n, per, parts);
      FVM_LAUNCH_CHECK(ctx);
    }
    {
      FVM_PHASE(ctx, FVM_PHASE_OTHER);
      k_block_sum<<<1, RED_BS, 0, ctx->stream>>>(parts, nb, nb, parts + nb);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  double *h = (double *)ctx->pinned.ptr + 8;
  if (nb > 0) {
    FVM_HIP(ctx, hipMemcpyAsync(h, parts + nb, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    *out = *h;
  } else {
    *out = 0.0;
  }
  return FVM_OK;
}

// ---- A12: updateGlobalStrainRateResidual (residual.c:12-51) --------------------
__global__ void k_residual_area(DGrid g, double *__restrict__ srr) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)g.NX * g.NY) return;
  const int ix = (int)(i % g.NX), jy = (int)(i / g.NX);
  if (ix > 0 && jy > 0) {
    const double dx = g.x[ix] - g.x[ix - 1];
    const double dy = g.y[jy] - g.y[jy - 1];
    srr[i] = srr[i] * (dx * dy);
  } else {
    srr[i] = 0.0;
  }
}

extern "C" int fvm_update_global_strain_rate_residual(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                                      double *residual) {
  if (!ctx || !mk || !residual) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int64_t nn = (int64_t)g->NX * g->NY;
  FVM_TRY(fvm_ws(ctx, ctx->scratch, ((size_t)nn + 64) * sizeof(double)));
  double *srr = (double *)ctx->scratch.ptr;
  ProjGroup pg;
  memset(&pg, 0, sizeof(pg));
  pg.stagx = pg.stagy = -1;
  pg.method = PM_ARITH;
  pg.nf = 1;
  pg.src[0].col = mk->strainRateResidual;
  pg.out[0] = srr;
  pg.epilogue = EPI_DIVIDE;
  FVM_TRY(fvm_project_group(ctx, g, mk, pg));
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_residual_area<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(dgrid(g), srr);
    FVM_LAUNCH_CHECK(ctx);
  }
  double sum = 0.0;
  FVM_TRY(fvm_reduce_sum(ctx, srr, nn, &sum));
  sum /= (g->LX * g->LY);
  *residual = sum;
  return FVM_OK;
}

// ---- A13: projectNodalFieldToMarkersS (markers.c:1670-1706) --------------------
__global__ void k_nodal_to_markers(DGrid g, int32_t n, const double *__restrict__ MX, const double *__restrict__ MY,
                                   const int32_t *__restrict__ cX, const int32_t *__restrict__ cY,
                                   const double *__restrict__ f, double *__restrict__ out) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const int cellX = cX[m], cellY = cY[m];
  if (!cell_ok(cellX, cellY, g.ncx(), g.NY)) return;
  const int NX = g.NX, NY = g.NY, P = g.per;
  const double mdx = (MX[m] - g.x[cellX]) / (g.x[cellX + 1] - g.x[cellX]);
  const double mdy = (MY[m] - g.y[cellY]) / (g.y[cellY + 1] - g.y[cellY]);
  out[m] = gat(f, NX, NY, P, cellY, cellX) * (1 - mdx) * (1 - mdy) +
           gat(f, NX, NY, P, cellY + 1, cellX) * (1 - mdx) * mdy +
           gat(f, NX, NY, P, cellY, cellX + 1) * (mdx) * (1 - mdy) +
           gat(f, NX, NY, P, cellY + 1, cellX + 1) * mdx * mdy;
}

extern "C" int fvm_project_nodal_field_to_markers(fvm_ctx *ctx, const fvm_grid *g, const double *nodalField,
                                                  fvm_markers *mk, double *markerField) {
  if (!ctx || !mk || !nodalField || !markerField) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (mk->nMark == 0) return FVM_OK;
  {
    FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
    k_nodal_to_markers<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(dgrid(g), mk->nMark, mk->X, mk->Y,
                                                                              mk->cellX, mk->cellY, nodalField,
                                                                              markerField);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

// ---- A14: glue loops of main_subduction.c ---------------------------------------
struct MatTab {
  double rho[FVM_MAXMAT], alpha[FVM_MAXMAT];
};
__global__ void k_density(int32_t n, const int32_t *__restrict__ cX, const int32_t *__restrict__ Mat,
                          const double *__restrict__ T, double *__restrict__ rho, double *__restrict__ rhodot,
                          MatTab t) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  if (cX[m] == -1) return;
  const int mm = Mat[m] & (FVM_MAXMAT - 1);
  rho[m] = -t.rho[mm] * (t.alpha[mm] * T[m]);
  rhodot[m] = 0.0;
}

extern "C" int fvm_update_marker_density(fvm_ctx *ctx, fvm_markers *mk, const Materials *mat) {
  if (!ctx || !mk || !mat) return FVM_ERR_ARG;
  if (mk->nMark == 0) return FVM_OK;
  MatTab t;
  for (int i = 0; i < FVM_MAXMAT; i++) {
    t.rho[i] = mat->materialRho[i];
    t.alpha[i] = mat->materialAlpha[i];
  }
  {
    FVM_PHASE(ctx, FVM_PHASE_MARKER_UPDATE);
    k_density<<<fvm_blocks(mk->nMark, 256), 256, 0, ctx->stream>>>(mk->nMark, mk->cellX, mk->Mat, mk->T, mk->rho,
                                                                     mk->rhodot, t);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

// max of non-negative doubles via their ordered bit patterns (order-free, exact).
// A bounded grid strides over the markers and each block issues one atomic, so
// the number of atomics on the result word stays at the block count.
#define MAXPOS_BLOCKS 1024
__global__ void __launch_bounds__(256) k_max_pos(int32_t n, const int32_t *__restrict__ cX,
                                                 const double *__restrict__ v, int absval,
                                                 unsigned long long *__restrict__ out) {
  __shared__ unsigned long long s_w[4];
  unsigned long long b = 0;
  for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n; m += (int64_t)gridDim.x * blockDim.x) {
    if (cX[m] == -1) continue;
    double x = absval ? fabs(v[m]) : v[m];
    if (!(x > 0.0)) x = 0.0;
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    b = u > b ? u : b;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(b, off);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = b;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) b = s_w[k] > b ? s_w[k] : b;
    if (b) atomicMax(out, b);
  }
}

static int max_pos(fvm_ctx *ctx, const fvm_markers *mk, const double *col, int absval, double *res) {
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096));
  unsigned long long *d = (unsigned long long *)((char *)ctx->red.ptr + 256);
  FVM_HIP(ctx, hipMemsetAsync(d, 0, 8, ctx->stream));
  if (mk->nMark > 0) {
    {
      FVM_PHASE(ctx, FVM_PHASE_OTHER);
      const unsigned nb = std::min(fvm_blocks(mk->nMark, 256), (unsigned)MAXPOS_BLOCKS);
      k_max_pos<<<nb, 256, 0, ctx->stream>>>(mk->nMark, mk->cellX, col, absval, d);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  unsigned long long *h = (unsigned long long *)ctx->pinned.ptr + 16;
  FVM_HIP(ctx, hipMemcpyAsync(h, d, 8, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  double r;
  memcpy(&r, h, 8);
  *res = r;
  return FVM_OK;
}

extern "C" int fvm_max_marker_T(fvm_ctx *ctx, const fvm_markers *mk, double *maxT) {
  if (!ctx || !mk || !maxT) return FVM_ERR_ARG;
  return max_pos(ctx, mk, mk->T, 0, maxT);
}
extern "C" int fvm_max_abs_sxx(fvm_ctx *ctx, const fvm_markers *mk, double *sxxmax) {
  if (!ctx || !mk || !sxxmax) return FVM_ERR_ARG;
  return max_pos(ctx, mk, mk->sxx, 1, sxxmax);
}
