/*
 * oracle.h -- CPU restatement of fvmic2d's marker hot path.  TEST INFRASTRUCTURE
 * ONLY: it is the checker the HIP path is compared against (tests/, the smoke
 * test and bench.py's CPU baseline leg); nothing in the product links it.
 *
 * Every function follows one reference function line by line over the
 * reference's own AoS Marker layout, single rank.  The PETSc DMDA semantics the
 * reference relies on are restated as:
 *   - scatter (DMLocalToGlobal ADD_VALUES, DM_BOUNDARY_GHOSTED): contributions
 *     to node indices outside [0,NX-1]x[0,NY-1] are dropped;
 *   - gather from a local vector outside the domain reads 0.0 (fresh, zeroed
 *     local vector);
 *   - periodic x (grid->xperiodic, DM_BOUNDARY_PERIODIC, nodalFields.c:11-20):
 *     x node index ix in [-2, NX+1] stands for ix mod NX, both ways (gathers
 *     read the wrapped column, scatters add into it); y stays ghosted;
 *   - VecPointwiseDivide is IEEE division (0/0 = NaN at zero-weight nodes).
 * The marker functions take periodic grids (NX >= 4, ghost coordinates as
 * makePeriodic, gridSpacing.c:159-168, sets them); the nodal functions
 * (nodal stress/strain, solution retrieval, displacement time step, marker
 * population control) return ORC_ERR_PERIODIC for them.
 */
#ifndef FVM_ORACLE_H
#define FVM_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_OK = 0,
  ORC_ERR_PERIODIC = 1,
  ORC_ERR_CELL_OUT_OF_BOUNDS = 2,
  ORC_ERR_BAD_STAG = 3,
  ORC_ERR_BAD_METHOD = 4,
  ORC_ERR_NAN_RESIDUAL = 5,
  ORC_ERR_VISCOSITY_LAW = 6,
  ORC_ERR_NOMEM = 7,
  ORC_ERR_BC = 8,
  ORC_ERR_IO = 9,
  ORC_ERR_NEGATIVE_ADD = 10,
  ORC_ERR_CAPACITY = 11
};

/* markers.c:85-106 findMarkerCells (+ findMarkerCell 41-83) */
int orc_find_marker_cells(MarkerSet *ms, const GridData *g, int rank);

/* markerProjection.c:71-250 projectMarkerFieldToNodes (+ findCellStag2D 252-269).
 * field_offset: byte offset of the projected double inside Marker, used when
 * materialProperty is NULL (the reference's &markers[0].field pointer). */
int orc_project_marker_field(const GridData *g, const MarkerSet *ms,
                             const double *materialProperty, size_t field_offset,
                             double *nodalField, int stagx, int stagy, int method);

/* markerProjection.c:11-68 projectMarkersNodesAll2 */
int orc_project_markers_nodes_all2(const MarkerSet *ms, const GridData *g, NodalArrays *nf,
                                   const Materials *mat);

/* markers.c:1507-1594 projectMarkersNodesFromScalar (dense marker array) */
int orc_project_markers_nodes_from_scalar(const MarkerSet *ms, const GridData *g,
                                          const double *markerField, double *nodalField);

/* markers.c:1670-1706 projectNodalFieldToMarkersS */
int orc_project_nodal_field_to_markers(const double *nodalField, MarkerSet *ms,
                                       size_t field_offset, const GridData *g);

/* markers.c:1865-2250 advectMarkersRK.  jaumann != 0 selects the -DJAUMANN
 * stress rotation (markers.c:2147-2156); 0 the default exact rotation. */
int orc_advect_markers_rk(MarkerSet *ms, const NodalArrays *nf, const GridData *g,
                          const BoundaryValues *bv, double dt, int jaumann);

/* subgridTemperatureChanges.c:5-124 */
int orc_subgrid_temperature_changes(const double *thisT, const GridData *g, const NodalArrays *nf,
                                    MarkerSet *ms, const Materials *mat, double dt,
                                    double dsubgrid);

/* subgridStressChanges.c:5-318 (updates nf->dsxx, dsyy, dsxy in place) */
int orc_subgrid_stress_changes(const GridData *g, NodalArrays *nf, MarkerSet *ms,
                               const Materials *mat, double dt, double dsubgrid);

/* updateMarkerStress.c:5-163 */
int orc_update_marker_stress(const GridData *g, const NodalArrays *nf, MarkerSet *ms);

/* updateMarkerStrainPressure.c:9-399 (non-TEXTURE branch) */
int orc_update_marker_strain_pressure(const GridData *g, const NodalArrays *nf, MarkerSet *ms,
                                      double dt);

/* viscosity.c:12-67 computeRheology */
int orc_compute_rheology(const Marker *mk, const Options *o, const Materials *mat,
                         double *eta, double *mu);

/* markers.c:1381-1505 checkPlasticYielding (ORs into *isYielding) */
int orc_check_plastic_yielding(const GridData *g, MarkerSet *ms, const Materials *mat,
                               double dt, int *isYielding, const Options *o);

/* viscosity.c:90-104 limitViscosity */
void orc_limit_viscosity(MarkerSet *ms, const Options *o);

/* residual.c:12-51 updateGlobalStrainRateResidual; returns the residual */
int orc_update_global_strain_rate_residual(const GridData *g, const MarkerSet *ms,
                                           double *residual);

/* main_subduction.c:255-264 marker density perturbation loop */
void orc_update_marker_density(MarkerSet *ms, const Materials *mat);
/* main_subduction.c:271-289 maximum active-marker T (0 if none above 0) */
double orc_max_marker_T(const MarkerSet *ms);
/* main_subduction.c:437-448 maximum |sxx| over active markers */
double orc_max_abs_sxx(const MarkerSet *ms);

/* nodalStressStrain.c:6-485 nodalStressStrain (strain rates, VEP nodal stress
 * update, shear heating, second invariants).  ORC_ERR_BC for a left/top side
 * whose ghost velocity the reference leaves undefined (type not 0 or 1). */
int orc_nodal_stress_strain(const GridData *g, NodalArrays *nf, const BoundaryValues *bv, double dt,
                            double *nodalHeating);
/* retrieveSolutions.c:4-82 retrieveSolutions (S: interleaved [vx, vy, p], p scaled in place) */
int orc_retrieve_solutions(const GridData *g, NodalArrays *nf, double *S, double Kcont, int *foundnan);
/* limitTimestep.c:6-133 limitDisplacementTimestep (dispFactor = options.displacementStepLimit) */
int orc_limit_displacement_timestep(const GridData *g, const NodalArrays *nf, double *dtdisp,
                                    double dispFactor);

/* io.c:299-402 saveMarkersBinary / restart.c:8-125 restartFromMarkers (one rank;
 * the path is the caller's: the reference names it output/Markers.<iMonte>.<iTime>) */
int orc_save_markers_binary(const MarkerSet *ms, const char *path, double elapsedTime);
int orc_restart_from_markers(MarkerSet *ms, const char *path, double *elapsedTime);

/* markers.c:455-781 checkMarkerDensity on one rank, with the counter-based
 * random stream orc_pc_u01 in place of drand48 / PetscRandom (see oracle.c).
 * Reads the cells that findMarkerCells left; may raise ms->nMark up to
 * ms->maxMark.  *nAdded = markers added (0 when the reference would neither add
 * nor abort), *nRemoved = markers removed. */
int orc_check_marker_density(MarkerSet *ms, const GridData *g, const Options *o, const Materials *mat,
                             uint64_t seed, int *nAdded, int *nRemoved);
double orc_pc_u01(uint64_t seed, uint64_t stream, uint64_t ctr);

#ifdef __cplusplus
}
#endif
#endif
