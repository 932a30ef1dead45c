/*
 * fvm_compat.h -- the reference's own function signatures for the marker hot
 * path, implemented on the GPU through the C ABI (fvm_abi.h).
 *
 * A driver written against the reference (main_subduction.c) calls these with
 * its host-side MarkerSet (AoS, 216-byte Marker), GridData and NodalFields; each
 * call copies what it reads to the GPU, runs the kernels and copies back what
 * it writes, so the host data ends up exactly as the reference's function
 * would leave it.  The copies cross PCIe on every call; a driver that keeps its
 * state on the GPU uses fvm_abi.h directly instead.
 *
 * Vec: without PETSc (FVM_WITH_PETSC undefined) a Vec is the small host
 * stand-in below (a dense NX*NY array, single rank).  With PETSc, define
 * FVM_WITH_PETSC before including this header and Vec is PETSc's own; the
 * wrappers then read and write it through VecGetArray/VecRestoreArray (one
 * rank: the global vector holds every node in natural order).
 */
#ifndef FVM_COMPAT_H
#define FVM_COMPAT_H

#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

#ifdef FVM_WITH_PETSC
#include <petscvec.h>
#else
typedef struct fvm_host_vec {
  int32_t n;  /* number of doubles (NX*NY) */
  double *a;  /* row-major [jy][ix] */
} *Vec;
#endif

/* fdcode.h:40-71 with Vec members (field order kept) */
typedef struct {
  Vec rho, rhodot, etaS, etaN;
  Vec thisT, lastT, kThermal, Cp;
  double boundaryT;
  Vec edotxx, edotxy;
  Vec edotyy, edotzz, edotxz, edotyz;
  Vec soxx, soxy;
  Vec soyy;
  Vec dsxx, dsyy, dsxy;
  Vec muS, muN;
  Vec vx, vy, p;
  Vec eii, sii, wxy;
  Vec ha;
  Vec nodalHeating;
  Vec VPTensorC;
  Vec VPTensorB;
  Vec strainRateResidual;
} NodalFields;

/* markers.h:14 */
void findMarkerCells(MarkerSet *markerset, GridData *grid);
/* markerProjection.h:11 */
int projectMarkersNodesAll2(MarkerSet *markerset, GridData *grid, NodalFields *nodalFields,
                            Materials *materials, Options *options);
/* markerProjection.h:10 */
int projectMarkerFieldToNodes(GridData *grid, MarkerSet *markerset, double *materialProperty,
                              double *markerfield, Vec nodalField, const int32_t stagx,
                              const int32_t stagy, const PROJECTION_METHOD pm);
/* markers.h:10 */
int projectMarkersNodesFromScalar(MarkerSet *markerset, GridData *grid, double *markerField,
                                  Vec nodalField);
/* markers.h:21 */
int projectNodalFieldToMarkersS(NodalFields *nodalFields, Vec nodalField, MarkerSet *markerset,
                                double *fieldptr, GridData *grid);
/* markers.h:20 */
int advectMarkersRK(MarkerSet *markerset, NodalFields *nodalFields, GridData *grid, Options *options,
                    BoundaryValues *bv, double dt);
/* markers.h:13 */
int checkPlasticYielding(GridData *grid, MarkerSet *markerset, Materials *materials, double dt,
                         int32_t *isYielding, Options *options, double elapsedTime);
/* viscosity.h: limitViscosity (viscosity.c:90-96) */
void limitViscosity(MarkerSet *markerset, Options *options, Materials *materials);
/* subgridTemperatureChanges.h:2 */
int subgridTemperatureChanges(Vec thisT, GridData *grid, NodalFields *nodalFields, MarkerSet *markerset,
                              Materials *materials, double dt, Options *options);
/* subgridStressChanges.h:1 */
int subgridStressChanges(GridData *grid, NodalFields *nodalFields, MarkerSet *markerset,
                         Materials *materials, double dt, double dsubgrid);
/* updateMarkerStress.h:1 */
int updateMarkerStress(GridData *grid, NodalFields *nodalFields, MarkerSet *markerset,
                       Materials *materials);
/* updateMarkerStrainPressure.h:1 */
int updateMarkerStrainPressure(GridData *grid, NodalFields *nodalFields, MarkerSet *markerset,
                               Materials *materials, Options *options, double dt);
/* residual.h: updateGlobalStrainRateResidual / getGlobalStrainRateResidual (residual.c:12-77) */
int updateGlobalStrainRateResidual(GridData *grid, MarkerSet *markerset, int32_t iter);
double getGlobalStrainRateResidual(int32_t iter);
int32_t checkConvergence(double atol, double rtol, int32_t iter);

/* initialPressureGuess.h:2: writes the pressure entries of S (a Vec of NX*NY*3) */
int initialPressureGuess(GridData *grid, NodalFields *nodalFields, Options *options, Vec S);
/* retrieveSolutions.h: S is the interleaved Stokes solution (a Vec of NX*NY*3) */
int retrieveSolutions(GridData *grid, NodalFields *nodalFields, Vec S, double Kcont, int32_t *foundnan);
/* nodalStressStrain.h (nodalStressStrain.c:6-485) */
int nodalStressStrain(GridData *grid, NodalFields *nodalFields, Options *options, BoundaryValues *bv,
                      double dt, Vec nodalHeating, double gy);
/* limitTimestep.h (limitTimestep.c:6-133) */
int limitDisplacementTimestep(GridData *grid, NodalFields *nodalFields, double *displacementdt,
                              Options *options);
/* io.h saveMarkersBinary (io.c:299-402): writes output/Markers.<iMonte>.<iTime> */
int saveMarkersBinary(MarkerSet *markerset, int32_t iMonte, int32_t iTime, double elapsedTime);
/* restart.h restartFromMarkers (restart.c:8-125): allocates markerset->markers
 * (nMark * options->maxMarkFactor slots; release with free()) */
int restartFromMarkers(MarkerSet *markerset, GridData *grid, Materials *materials, Options *options,
                       int32_t iMonte, int32_t iTime, double *elapsedTime);

/* io.h saveNodalFields (io.c:150-296): the binary format (fdcode.h's
 * OUTPUT_NODAL_FIELDS_BINARY) only; writes output/loadNodalFields_<iMonte>_<iTime>.petscbin */
#define FVM_OUTPUT_NODAL_FIELDS_BINARY 0 /* fdcode.h:249 */
int saveNodalFields(NodalFields *nodalFields, GridData *grid, int32_t iMonte, int32_t iTime, double dt,
                    double elapsedTime, int32_t format);

/* Release the GPU state the wrappers keep (context, device copies). */
void fvm_compat_finalize(void);
/* Status of the last wrapper call that has no status return (findMarkerCells). */
int fvm_compat_last_status(void);

#ifdef __cplusplus
}
#endif
#endif
