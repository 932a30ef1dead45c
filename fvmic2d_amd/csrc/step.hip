// Whole marker steps in the reference's call order (main_subduction.c:227-463),
// enqueued without host round trips except where a scalar must come back.
#include "fvm_internal.h"

extern "C" int fvm_step_thermal(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                                const Materials *mat, const Options *opt, const BoundaryValues *bv, double dt,
                                int32_t nproj, const fvm_projection *list) {
  if (!ctx || !mk || !nf || !mat || !opt || !bv) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  FVM_TRY(fvm_locate_async(ctx, g, mk, 0, mat, true));         // :243 + density :255-264
  if (list)                                                     // :268
    FVM_TRY(fvm_project_marker_fields(ctx, g, mk, nproj, list));
  else
    FVM_TRY(fvm_project_markers_nodes_all2(ctx, g, mk, nf, mat));
  if (!nf->thisT || !nf->lastT || !nf->rho || !nf->Cp || !nf->kThermal || !nf->vx || !nf->vy || !nf->wxy)
    return fvm_set_err(ctx, FVM_ERR_ARG, "step: missing nodal field");
  double *dTsub = nullptr, *deltaT = nullptr;                   // :415 subgridTemperatureChanges
  // marker T is unchanged since the projection above, so its bucket-ordered
  // copy can be reused by the fused dTsub pass
  ctx->reuse_payload = true;
  const int rc = fvm_subgrid_t_pass1(ctx, nf->thisT, g, nf, mk, mat, dt, opt->subgridTemperatureDiffusivity, &dTsub,
                                     &deltaT);
  ctx->reuse_payload = false;
  FVM_TRY(rc);
  // its pass 2 runs inside the advection kernel (:435 advectMarkersRK)
  FVM_TRY(fvm_advect_with_subgrid_t2(ctx, mk, nf, g, bv, dt, deltaT, dTsub, dt));
  return FVM_OK;
}

extern "C" int fvm_step_vep(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                            const NodalArrays *host, const Materials *mat, const Options *opt,
                            const BoundaryValues *bv, double dt, int32_t *isYielding, double *residual) {
  if (!ctx || !mk || !nf || !host || !mat || !opt || !bv) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  FVM_TRY(fvm_locate_async(ctx, g, mk, 0, mat, true));         // :243 + density :255-264
  int32_t y = 0;
  FVM_TRY(fvm_check_plastic_yielding(ctx, g, mk, mat, dt, &y, opt, 0.0));  // :299
  FVM_TRY(fvm_limit_viscosity(ctx, mk, opt));                   // :300
  FVM_TRY(fvm_project_markers_nodes_all2(ctx, g, mk, nf, mat)); // :303
  // host: formVEPSystem, KSP, retrieveSolutions, nodalStressStrain (:309-335)
  FVM_TRY(fvm_update_marker_strain_pressure(ctx, g, host, mk, mat, opt, dt));  // :341
  double r = 0.0;
  FVM_TRY(fvm_update_global_strain_rate_residual(ctx, g, mk, &r));            // :345
  if (opt->subgridStressDiffusivity > 0.0)                                      // :365
    FVM_TRY(fvm_subgrid_stress_changes(ctx, g, host, mk, mat, dt, opt->subgridStressDiffusivity));
  FVM_TRY(fvm_update_marker_stress(ctx, g, host, mk, mat));                     // :368
  NodalArrays th = *nf;  // subgrid T reads the projected lastT/rho/Cp/k
  FVM_TRY(fvm_subgrid_temperature_changes(ctx, host->thisT, g, &th, mk, mat, dt,  // :415
                                          opt->subgridTemperatureDiffusivity));
  NodalArrays adv = *host;
  FVM_TRY(fvm_advect_markers_rk(ctx, mk, &adv, g, bv, dt, ctx->jaumann));  // :435
  if (isYielding) *isYielding = y;
  if (residual) *residual = r;
  return FVM_OK;
}
