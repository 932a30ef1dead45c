// Whole marker steps in the reference's call order (main_subduction.c:227-463),
// enqueued without host round trips except where a scalar must come back.
#include "fvm_internal.h"

extern "C" int fvm_step_thermal(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                                const Materials *mat, const Options *opt, const BoundaryValues *bv, double dt,
                                int32_t nproj, const fvm_projection *list) {
  if (!ctx || !mk || !nf || !mat || !opt || !bv) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  FVM_TRY(fvm_locate_async(ctx, g, mk, 0));                     // :243
  FVM_TRY(fvm_update_marker_density(ctx, mk, mat));             // :255-264
  if (list)                                                     // :268
    FVM_TRY(fvm_project_marker_fields(ctx, g, mk, nproj, list));
  else
    FVM_TRY(fvm_project_markers_nodes_all2(ctx, g, mk, nf, mat));
  FVM_TRY(fvm_subgrid_temperature_changes(ctx, nf->thisT, g, nf, mk, mat, dt,  // :415
                                          opt->subgridTemperatureDiffusivity));
  FVM_TRY(fvm_advect_markers_rk(ctx, mk, nf, g, bv, dt, 0));   // :435
  return FVM_OK;
}

extern "C" int fvm_step_vep(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                            const NodalArrays *host, const Materials *mat, const Options *opt,
                            const BoundaryValues *bv, double dt, int32_t *isYielding, double *residual) {
  if (!ctx || !mk || !nf || !host || !mat || !opt || !bv) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  FVM_TRY(fvm_locate_async(ctx, g, mk, 0));                     // :243
  FVM_TRY(fvm_update_marker_density(ctx, mk, mat));             // :255-264
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
  FVM_TRY(fvm_advect_markers_rk(ctx, mk, &adv, g, bv, dt, 0));  // :435
  if (isYielding) *isYielding = y;
  if (residual) *residual = r;
  return FVM_OK;
}
