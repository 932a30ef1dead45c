// The energy equation's system (formThermalSystem.c:6-249) assembled on the GPU
// and solved there (the reference's "energy_" KSP): one thread per basic node
// writes its row as five coefficients -- diagonal, east (ix+1), west (ix-1),
// south (jy+1), north (jy-1), 0 where the row has no such entry -- and its
// right-hand side, in the reference's branch order: left wall, right wall
// (plate or inflow), right outflow, top, bottom, interior.  The boundary
// temperatures are the subduction set-up's (benchmarkInitialConditions.c:
// 75-135).  Reference expression order, compiled without contraction.
#include <math.h>

#include "fvm_internal.h"

namespace {

struct ThermalParams {
  int NX, NY;
  const double *x, *y;
  int topType;
  double topValue;
  double slabAngle, slabAge;
  double plateThickness, rootThickness, rootCenter, rootWidth;
  double rho0, dt;
  int doSteady;
};

// benchmarkInitialConditions.c:80-93 (pow(c, 2.0) is folded to c*c by the
// reference's gcc -O3 build)
__device__ __forceinline__ double plate_depth(double x, const ThermalParams &P) {
  if (fabs(x - P.rootCenter) < P.rootWidth) {
    const double c = cos(M_PI * (x - P.rootCenter) / P.rootWidth / 2.0);
    return P.plateThickness + P.rootThickness * (c * c);
  }
  return P.plateThickness;
}
// :94-97, mantle temperature 1573 (:75-78)
__device__ __forceinline__ double plate_geotherm(double x, double y, const ThermalParams &P) {
  return y / plate_depth(x, P) * (1573.0 - 273.0) + 273.0;
}
// :128-135
__device__ __forceinline__ double slab_inflow_temperature(double x, double y, double angle, double age) {
  const double Ts = 273, T0 = 1623, kappa = 0.7272e-6, s_in_year = 3.15576e7;
  const double slab_age_s = age * 1.0e6 * s_in_year;
  return Ts + (T0 - Ts) * erf((y - x * tan(angle)) / (2.0 * sqrt(kappa * slab_age_s)));
}

__global__ void __launch_bounds__(256) k_thermal_rows(ThermalParams P, const double *__restrict__ k,
                                                      const double *__restrict__ rho, const double *__restrict__ Cp,
                                                      const double *__restrict__ lastT,
                                                      const double *__restrict__ vx, double *__restrict__ cd,
                                                      double *__restrict__ ce, double *__restrict__ cw,
                                                      double *__restrict__ cs, double *__restrict__ cn,
                                                      double *__restrict__ rhs) {
  const int NX = P.NX, NY = P.NY;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  double d = 0.0, e = 0.0, w = 0.0, s = 0.0, n = 0.0, b = rhs[i];
  const double *gx = P.x, *gy = P.y;
  if (ix == 0) {  // left wall: slab inflow temperature (:80-81, :160-162)
    d = 1.0;
    b = slab_inflow_temperature(gx[ix], gy[jy], P.slabAngle, P.slabAge);
  } else if (ix == NX - 1 &&
             (gy[jy] <= plate_depth(gx[NX - 1], P) || (jy > 0 && (vx[i] + vx[i - NX]) / 2.0 < 0.0))) {
    d = 1.0;  // right wall, plate or inflow (:82-83, :163-171)
    b = gy[jy] <= plate_depth(gx[ix], P) ? plate_geotherm(gx[ix], gy[jy], P) : 1573.0;
  } else if (ix == NX - 1) {  // outflow (:84-85, :172-174)
    d = 1.0;
    b = lastT[i];
  } else if (jy == 0) {  // top (:86-94, :175-181)
    if (P.topType == 0) {
      d = 1.0;
      b = P.topValue;
    } else {
      const double dyt1 = gy[jy + 1] - gy[jy];
      d = -1.0 / dyt1;
      s = 1.0 / dyt1;
      b = 0.0;
    }
  } else if (jy == NY - 1) {  // bottom (:95-97, :182-184)
    d = 1.0;
    b = lastT[i];
  } else {  // interior (:98-131)
    const double ka = (k[i] + k[i - 1]) / 2.0;
    const double kb = (k[i] + k[i + 1]) / 2.0;
    const double kc = (k[i] + k[i - NX]) / 2.0;
    const double kd = (k[i] + k[i + NX]) / 2.0;
    const double dx2 = gx[ix + 1] - gx[ix];
    const double dx1 = gx[ix] - gx[ix - 1];
    const double dy1 = gy[jy] - gy[jy - 1];
    const double dy2 = gy[jy + 1] - gy[jy];
    const double lap = 2.0 / (dx2 + dx1) * (kb / dx2 + ka / dx1) + 2.0 / (dy1 + dy2) * (kd / dy2 + kc / dy1);
    if (!P.doSteady) {
      b = b + (P.rho0 + rho[i]) * Cp[i] / P.dt * lastT[i];
      d = (P.rho0 + rho[i]) * Cp[i] / P.dt + lap;
    } else {
      d = lap;
    }
    e = -2.0 / (dx1 + dx2) * kb / dx2;
    w = -2.0 / (dx1 + dx2) * ka / dx1;
    n = -2.0 / (dy1 + dy2) * kc / dy1;
    s = -2.0 / (dy1 + dy2) * kd / dy2;
  }
  cd[i] = d;
  ce[i] = e;
  cw[i] = w;
  cs[i] = s;
  cn[i] = n;
  rhs[i] = b;
}

// rows divided by their diagonal: the solve's residual then weighs the
// Dirichlet rows (scale 1) and the diffusion rows (scale k/h^2) alike
__global__ void __launch_bounds__(256) k_row_scale(int64_t nn, const double *__restrict__ coef,
                                                   const double *__restrict__ rhs, double *__restrict__ sc,
                                                   double *__restrict__ srhs) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double d = coef[i];
  sc[i] = 1.0;
  for (int r = 1; r < 5; r++) sc[r * nn + i] = coef[r * nn + i] / d;
  srhs[i] = rhs[i] / d;
}

}  // namespace

extern "C" int fvm_form_thermal_system(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, const Materials *mat,
                                       const Options *opt, double dt, int32_t doSteady, double *coef, double *rhs) {
  if (!ctx || !nf || !mat || !opt || !coef || !rhs || !nf->kThermal || !nf->rho || !nf->Cp || !nf->lastT ||
      !nf->vx)
    return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (g->xperiodic) return fvm_set_err(ctx, FVM_ERR_PERIODIC, "formThermalSystem: periodic grids are not supported");
  const int tt = opt->thermalBCTop.type[0];
  if (tt != 0 && tt != 1)
    return fvm_set_err(ctx, FVM_ERR_BC, "formThermalSystem: top thermal boundary type %d sets no row", tt);
  ThermalParams P;
  P.NX = g->NX;
  P.NY = g->NY;
  P.x = g->x;
  P.y = g->y;
  P.topType = tt;
  P.topValue = opt->thermalBCTop.value[0];
  P.slabAngle = opt->slabAngle;
  P.slabAge = opt->slabAge;
  P.plateThickness = opt->overridingPlateThickness;
  P.rootThickness = opt->rootThickness;
  P.rootCenter = opt->rootCenter;
  P.rootWidth = opt->rootWidth;
  P.rho0 = mat->materialRho[0];
  P.dt = dt;
  P.doSteady = doSteady ? 1 : 0;
  const int64_t nn = (int64_t)g->NX * g->NY;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_thermal_rows<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(P, nf->kThermal, nf->rho, nf->Cp, nf->lastT, nf->vx,
                                                                 coef, coef + nn, coef + 2 * nn, coef + 3 * nn,
                                                                 coef + 4 * nn, rhs);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_solve_thermal_system(fvm_ctx *ctx, const fvm_grid *g, const double *coef, const double *rhs,
                                        double *T, double rtol, int32_t maxit, int32_t *iters, double *relres) {
  if (!ctx || !coef || !rhs || !T || !(rtol > 0.0) || maxit < 1) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "thermal solve"));
  const int64_t nn = (int64_t)g->NX * g->NY;
  FVM_TRY(fvm_ws(ctx, ctx->partials, (size_t)6 * nn * sizeof(double)));
  double *sc = (double *)ctx->partials.ptr, *srhs = sc + 5 * nn;
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_row_scale<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(nn, coef, rhs, sc, srhs);
    FVM_LAUNCH_CHECK(ctx);
  }
  return fvm_bicgstab_5pt(ctx, g->NX, g->NY, sc, sc + nn, sc + 2 * nn, sc + 3 * nn, sc + 4 * nn, srhs, T, rtol,
                          maxit, iters, relres);
}
