/*
 * fvm_abi.h -- C ABI of the MI355X (gfx950) marker hot path.
 *
 * Every entry point takes plain pointers and sizes (no PETSc, no torch), returns
 * an int status (FVM_OK = 0) and never aborts.  Marker state lives on the GPU as
 * structure-of-arrays columns (fvm_markers), nodal fields as dense NX*NY device
 * arrays in the reference's DMDA natural order ([jy][ix], ix fastest).  Grid
 * coordinate arrays are device copies of the reference's padded arrays: each
 * pointer addresses element 0 of an array that is valid for indices -2..N+1
 * (gridGenerator.c:123-130).
 *
 * Work is enqueued on the context's HIP stream; a call returns once its kernels
 * are enqueued unless it has to hand a scalar back (nOut, isYielding, residual),
 * in which case it synchronises the stream.
 *
 * Semantics kept from the reference's PETSc layer: scatters to node indices
 * outside the grid are dropped (DMLocalToGlobal ADD on DM_BOUNDARY_GHOSTED),
 * gathers outside the grid read 0, normalisation is IEEE division (nodes with
 * zero weight become NaN).
 *
 * Periodic x (xperiodic != 0, DM_BOUNDARY_PERIODIC, nodalFields.c:11-20): the
 * grid has NX cells along x, the last closed by x[NX] = LX, and the ghost
 * coordinates are those makePeriodic sets (gridSpacing.c:159-168).  Node column
 * ix outside [0, NX-1] stands for ix mod NX in scatters and gathers, and the
 * marker functions follow the reference's xperiodic branches (markers.c:49,73,
 * 1908-1915, 1965-1971, 2014-2023, 2225-2231; updateMarkerStress.c:66,76,110;
 * updateMarkerStrainPressure.c:137,146,211).  It needs NX >= 4.  The nodal
 * functions (nodal stress/strain, solution retrieval, displacement time step,
 * thermal system, initial pressure guess), marker population control and the
 * batched ensemble return FVM_ERR_PERIODIC for periodic grids.
 *
 * Projections are deterministic: markers are bucketed by (basic cell, quadrant)
 * into a sort order that is a pure function of the marker positions, and every
 * nodal sum is formed in a fixed order, so repeated calls give bitwise
 * identical nodal fields.
 */
#ifndef FVM_ABI_H
#define FVM_ABI_H

#include <stddef.h>
#include <stdint.h>
#include "structures.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------------- */
enum {
  FVM_OK = 0,
  FVM_ERR_PERIODIC = 1,          /* a periodic grid where it is not supported, or with NX < 4 */
  FVM_ERR_CELL_OUT_OF_BOUNDS = 2,/* markers.c:74-77 (reference aborts) */
  FVM_ERR_BAD_STAG = 3,          /* markerProjection.c:137-153 (reference aborts) */
  FVM_ERR_BAD_METHOD = 4,        /* markerProjection.c:224-225, 241-243 (aborts) */
  FVM_ERR_NAN_RESIDUAL = 5,      /* updateMarkerStrainPressure.c:302-306 (aborts) */
  FVM_ERR_VISCOSITY_LAW = 6,     /* viscosity.c:60-62 (aborts) */
  FVM_ERR_NOMEM = 7,
  FVM_ERR_HIP = 8,               /* a HIP runtime call failed; see fvm_last_error */
  FVM_ERR_ARG = 9,               /* invalid argument (NULL pointer, size) */
  FVM_ERR_TOO_MANY_FIELDS = 10,  /* more projected fields in one group than supported */
  FVM_ERR_NOT_CONVERGED = 11,    /* an iterative solve missed its tolerance */
  FVM_ERR_BC = 12,               /* a boundary type whose ghost value the reference leaves undefined */
  FVM_ERR_IO = 13,               /* a marker file could not be opened, read or written, or has another layout */
  FVM_ERR_CAPACITY = 14,         /* markers.c:772-775: more markers to add than free slots (reference aborts) */
  FVM_ERR_NEGATIVE_ADD = 15,     /* markers.c:548-551: an under-full quarter's target is not above its count (aborts) */
  FVM_ERR_STALLED = 16           /* a bounded wait of the ensemble's persistent step ran out; that step is incomplete */
};

/* ---- device marker store: one device column per Marker field --------------- */
typedef struct {
  int32_t nMark;   /* markers visited: m < nMark (MarkerSet.nMark) */
  int32_t maxMark; /* capacity of every column */
  double *X, *Y, *VX, *VY, *T, *Tdot;
  double *sxx, *syy, *sxy;            /* Marker.s.T11, T22, T12 */
  double *eta, *mu, *rho, *rhodot, *p;
  double *exx, *eyy, *exy;            /* Marker.e */
  double *Exx, *Eyy, *Exy;            /* Marker.E */
  double *Eii, *wxy, *strainRateResidual;
  int32_t *Mat, *cellX, *cellY, *cpu, *isYielding;
} fvm_markers;

/* ---- device grid -------------------------------------------------------- */
typedef struct {
  int32_t NX, NY;
  double LX, LY;
  int32_t xperiodic;
  const double *x, *y, *xc, *yc; /* device, valid for indices -2..N+1 */
} fvm_grid;

/* ---- context ------------------------------------------------------------ */
typedef struct fvm_ctx fvm_ctx;

/* Create a context on the current device.  stream: a hipStream_t to enqueue on,
 * or NULL for a stream owned by the context. */
int fvm_ctx_create(fvm_ctx **ctx, void *stream);
int fvm_ctx_destroy(fvm_ctx *ctx);
const char *fvm_last_error(const fvm_ctx *ctx);
void *fvm_ctx_stream(fvm_ctx *ctx);
/* Waits for the context's stream (every call before it has completed). */
int fvm_ctx_sync(fvm_ctx *ctx);
/* Bytes of device workspace the context currently holds (sort buffers, partials). */
size_t fvm_ctx_workspace_bytes(const fvm_ctx *ctx);
/* Lanes per lattice cell in the projection passes: 8 (default), 4, 16 or 32 (the
 * fused dTsub pass of subgridTemperatureChanges takes 8 when this is 4).  The
 * lane count fixes the summation tree, so results are bitwise reproducible for
 * a given setting and agree across settings to the projection tolerance. */
int fvm_ctx_set_lanes(fvm_ctx *ctx, int32_t lanes);
/* Stress rotation of the whole-step entry points (fvm_step_thermal,
 * fvm_step_vep): 0 (default) the exact rotation of advectMarkersRK, 1 the
 * -DJAUMANN form (markers.c:2147-2156), as the reference selects it at compile
 * time.  fvm_advect_markers_rk takes the choice as its own argument. */
int fvm_ctx_set_jaumann(fvm_ctx *ctx, int32_t jaumann);

/* ---- device memory helpers (so a host program needs no other GPU API) ---- */
int fvm_device_count(int *n);
int fvm_malloc(void **ptr, size_t bytes);
int fvm_free(void *ptr);
int fvm_memcpy_htod(fvm_ctx *ctx, void *dst, const void *src, size_t bytes);
int fvm_memcpy_dtoh(fvm_ctx *ctx, void *dst, const void *src, size_t bytes);
int fvm_memcpy_dtod(fvm_ctx *ctx, void *dst, const void *src, size_t bytes);
int fvm_memset(fvm_ctx *ctx, void *dst, int value, size_t bytes);
/* Pinned (page-locked) host memory, for the packed exchange below. */
int fvm_host_alloc(void **ptr, size_t bytes);
int fvm_host_free(void *ptr);
/* Per-step exchange with host-side solves (no counterpart in the reference, whose
 * nodal Vecs live on the host): na device arrays of n doubles each, packed back to
 * back in `host` (array k at host[k*n]).  One host<->device copy and one kernel per
 * call instead of one copy per array; pass pinned memory from fvm_host_alloc for
 * full bandwidth.  At most 32 arrays; both synchronise. */
int fvm_nodal_upload(fvm_ctx *ctx, int64_t n, int32_t na, double *const *dev, const double *host);
int fvm_nodal_download(fvm_ctx *ctx, int64_t n, int32_t na, const double *const *dev, double *host);

/* Allocate every column of a marker store in one device slab (freed by
 * fvm_markers_free).  Columns are zero-filled, cellX/cellY = -1. */
int fvm_markers_alloc(fvm_ctx *ctx, int32_t maxMark, fvm_markers *mk);
int fvm_markers_free(fvm_markers *mk);
/* Host AoS block (216-byte Marker, fdcode.h:97-129) <-> device columns.
 * Upload sets mk->nMark = nMark. */
int fvm_markers_upload_aos(fvm_ctx *ctx, fvm_markers *mk, const Marker *aos, int32_t nMark);
int fvm_markers_download_aos(fvm_ctx *ctx, const fvm_markers *mk, Marker *aos);

/* ---- A2: findMarkerCells (markers.c:85-106, findMarkerCell 41-83) ---------- */
/* Writes cellX, cellY, cpu = rank; *nOut (may be NULL) receives the number of
 * markers outside the domain. */
int fvm_find_marker_cells(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t rank,
                          int32_t *nOut);

/* Store reorder.  No counterpart in the reference: a layout call for callers
 * that do not need a marker to keep its index between steps (the reference's
 * own marker exchange moves markers between slots: outgoing markers are
 * flagged -1 in place and incoming ones take the first flagged slot or are
 * appended, markers.c:160-170,218-226).  Advection scrambles the index order
 * of the markers in space over many steps; every marker kernel then gathers
 * from scattered lines (a randomly ordered store runs the thermal step about
 * twice as long).  This call moves the markers back into cell order: located
 * markers first, by (basic cell, quadrant) bucket -- the projections' bucket
 * order, (cellY*ncx + cellX, quadrant), ascending index inside a bucket --
 * then the unlocated ones (cellX == -1) by index.  Every column moves; nMark
 * is unchanged.  order (device, nMark int32, may be NULL) receives the old
 * index of the marker now in each slot.  The cells must be current (a
 * fvm_find_marker_cells or whole step since the markers last moved), as for a
 * projection.  Per-marker results are unchanged (each marker's arithmetic is
 * its own); nodal sums stay within rounding of the unordered store's (their
 * in-bucket order follows the new indices); population control, which picks
 * by index, then picks among the new indices.  Costs about 0.8 KB of HBM
 * traffic per marker. */
int fvm_markers_reorder(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, int32_t *order);

/* ---- A3/A4/A5: projection -------------------------------------------------- */
/* One projection of markerProjection.c:71 projectMarkerFieldToNodes: a marker
 * column (markerField, device) or a material property table (materialProperty,
 * host array of FVM_MAXMAT, indexed by Mat) onto the node family (stagx,stagy)
 * with the given PROJECTION_METHOD, result in nodalField (device, NX*NY). */
typedef struct {
  const double *markerField;      /* device column, or NULL */
  const double *materialProperty; /* host table, or NULL */
  double *nodalField;             /* device output */
  int32_t stagx, stagy;
  int32_t method;
} fvm_projection;

/* markerProjection.c:71-250 projectMarkerFieldToNodes: a list of projections
 * sharing one marker sort.  Entries with the same (stagx, stagy, method) share
 * one pass over the markers (up to 8 fields per group). */
int fvm_project_marker_fields(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                              int32_t n, const fvm_projection *list);

/* markerProjection.c:11-68 projectMarkersNodesAll2: rho, etaN, etaS, kThermal,
 * lastT, Cp, soxx, soyy, soxy; sets rhodot = 0 and muS = muN = 1e99. */
int fvm_project_markers_nodes_all2(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                   const NodalArrays *nf, const Materials *mat);

/* markers.c:1507-1594 projectMarkersNodesFromScalar (dense device marker array). */
int fvm_project_markers_nodes_from_scalar(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                          const double *markerField, double *nodalField);

/* ---- A13: markers.c:1670-1706 projectNodalFieldToMarkersS ------------------ */
int fvm_project_nodal_field_to_markers(fvm_ctx *ctx, const fvm_grid *g, const double *nodalField,
                                       fvm_markers *mk, double *markerField);

/* ---- A6: markers.c:1865-2250 advectMarkersRK -------------------------------- */
/* Reads nf->vx, vy, wxy.  jaumann = 0: default exact stress rotation
 * (markers.c:2157-2164); 1: the -DJAUMANN variant (markers.c:2147-2156). */
int fvm_advect_markers_rk(fvm_ctx *ctx, fvm_markers *mk, const NodalArrays *nf, const fvm_grid *g,
                          const BoundaryValues *bv, double dt, int32_t jaumann);

/* ---- A7: subgridTemperatureChanges.c:5-124 --------------------------------- */
/* Reads nf->lastT, rho, Cp, kThermal; thisT is the host thermal solution. */
int fvm_subgrid_temperature_changes(fvm_ctx *ctx, const double *thisT, const fvm_grid *g,
                                    const NodalArrays *nf, fvm_markers *mk, const Materials *mat,
                                    double dt, double dsubgrid);

/* ---- A8: subgridStressChanges.c:5-318 --------------------------------------- */
/* Reads nf->soxx, soyy, soxy; updates nf->dsxx, dsyy, dsxy in place. */
int fvm_subgrid_stress_changes(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                               fvm_markers *mk, const Materials *mat, double dt, double dsubgrid);

/* ---- A9: updateMarkerStress.c:5-163 ----------------------------------------- */
int fvm_update_marker_stress(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                             fvm_markers *mk, const Materials *mat);

/* ---- A10: updateMarkerStrainPressure.c:9-399 -------------------------------- */
/* Returns FVM_ERR_NAN_RESIDUAL where the reference aborts (a NaN strain-rate
 * residual on any active marker); marker fields are written either way. */
int fvm_update_marker_strain_pressure(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                                      fvm_markers *mk, const Materials *mat, const Options *opt,
                                      double dt);

/* ---- A11: markers.c:1381-1505 checkPlasticYielding + viscosity.c:12-75 ------- */
/* ORs 1 into *isYielding when any marker yields. */
int fvm_check_plastic_yielding(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk,
                               const Materials *mat, double dt, int32_t *isYielding,
                               const Options *opt, double elapsedTime);
/* viscosity.c:90-104 limitViscosity (every marker m < nMark). */
int fvm_limit_viscosity(fvm_ctx *ctx, fvm_markers *mk, const Options *opt);

/* ---- A12: residual.c:12-51 updateGlobalStrainRateResidual ------------------- */
int fvm_update_global_strain_rate_residual(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                           double *residual);

/* ---- A14: step glue loops of main_subduction.c ------------------------------ */
/* :255-264 rho = -rho0(Mat) alpha(Mat) T, rhodot = 0 on active markers. */
int fvm_update_marker_density(fvm_ctx *ctx, fvm_markers *mk, const Materials *mat);
/* :271-289 max T over active markers (0 if none is positive). */
int fvm_max_marker_T(fvm_ctx *ctx, const fvm_markers *mk, double *maxT);
/* :437-448 max |sxx| over active markers. */
int fvm_max_abs_sxx(fvm_ctx *ctx, const fvm_markers *mk, double *sxxmax);

/* ---- nodal step between the Stokes solve and the marker updates ---------- */
/* retrieveSolutions.c:4-82 retrieveSolutions(grid, nodalFields, S, Kcont, &foundnan):
 * S (device, NX*NY*3 doubles) is the interleaved Stokes solution [jy][ix][vx, vy, p]
 * (fdcode.h:242-245); its pressure entries are scaled by Kcont in place, vx, vy, p
 * are written to nf, and the Kcont-scaled row-1 pressure mean (weights dx/LX,
 * summed in index order) is removed from every node with jy > 0, ix > 0.
 * *foundnan (optional) is set to 0: the reference checks only when built with
 * TEXTURE.  Called at main_subduction.c:320. */
int fvm_retrieve_solutions(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, double *S,
                           double Kcont, int32_t *foundnan);
/* nodalStressStrain.c:6-485 nodalStressStrain(grid, nodalFields, options, bv, dt,
 * nodalHeating, gy): from nf->vx, vy, etaN, muN, etaS, muS and the old stresses
 * soxx, soyy, soxy, writes edotxx, edotyy (cell centres, 0 on row/column 0),
 * edotxy and wxy (basic nodes; left/top ghost velocities by mechBCLeft.type[1] /
 * mechBCTop.type[0], which must be 0 or 1, else FVM_ERR_BC), the new stresses
 * soxx, soyy, soxy and their changes dsxx, dsyy, dsxy, and on the interior basic
 * nodes the shear heating (nodalHeating, zero elsewhere) and the second
 * invariants eii, sii (zero elsewhere).  options and gy are not read by the
 * reference either.  Called at main_subduction.c:335. */
int fvm_nodal_stress_strain(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                            const BoundaryValues *bv, double dt, double *nodalHeating);
/* limitTimestep.c:6-133 limitDisplacementTimestep(grid, nodalFields, &dt, options):
 * *displacementdt (host, in/out) becomes the minimum of itself and, over every
 * node, dx/|vx|*displacementStepLimit, dy/|vy|*displacementStepLimit (nonzero
 * velocities; dx, dy the smaller adjacent spacing) and |0.01 rho/rhodot| (nonzero
 * rhodot), then is capped at 0.01/max|rhodot|.  Exact (a minimum is order-free);
 * synchronises.  Called at main_subduction.c:349. */
int fvm_limit_displacement_timestep(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                                    double *displacementdt, const Options *opt);
/* No counterpart in the reference: fvm_limit_displacement_timestep for callers
 * that cannot allocate (code under HIP graph capture, tight loops, harnesses
 * that forbid hipMalloc).  It takes no context and allocates nothing: its
 * per-block partials live in a static device array of 24 KiB per GPU (1024
 * partials; any grid size fits), held until the process ends.  *dt_device is
 * device memory: in, the step to limit; out, the limited step, bitwise the
 * value fvm_limit_displacement_timestep returns.  Enqueued on `stream` (a
 * hipStream_t, NULL for the default stream) without synchronising.  Calls that
 * overlap on one GPU share the array and race. */
int fvm_limit_displacement_timestep_noalloc(void *stream, const fvm_grid *g, const NodalArrays *nf,
                                            const Options *opt, double *dt_device);

/* ---- energy equation ------------------------------------------------------ */
/* formThermalSystem.c:6-249 formThermalSystem(problem, grid, nodalFields, LHS,
 * RHS, dt, options, doSteady) for the subduction set-up: the 5-point system on
 * the basic nodes as five device arrays in coef (5*NX*NY doubles: diagonal, east
 * (ix+1), west (ix-1), south (jy+1), north (jy-1); 0 where a row has no such
 * entry), rows in node order.  rhs (device NX*NY, in/out) holds the heating
 * terms on entry, as the reference's RHS does; interior rows gain
 * (rho0 + rho) Cp / dt * lastT (doSteady == 0), boundary rows get their values:
 * left wall the slab inflow temperature, right wall the plate geotherm or 1573
 * where y <= plate depth or the mean vx of the node's two vx values is negative,
 * else lastT (outflow), top thermalBCTop (type 0 value, type 1 zero gradient;
 * other types FVM_ERR_BC), bottom lastT.  Reads nf->kThermal, rho, Cp, lastT, vx
 * and materialRho[0]. */
int fvm_form_thermal_system(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, const Materials *mat,
                            const Options *opt, double dt, int32_t doSteady, double *coef, double *rhs);
/* The "energy_" solve of that system (the reference uses a direct LU through
 * its options file): BiCGStab from T = 0 on the system with every row divided by
 * its diagonal, to rtol relative residual of that scaled system.
 * Synchronises; FVM_ERR_NOT_CONVERGED when rtol is not met (T is written). */
int fvm_solve_thermal_system(fvm_ctx *ctx, const fvm_grid *g, const double *coef, const double *rhs, double *T,
                             double rtol, int32_t maxit, int32_t *iters, double *relres);

/* ---- marker files (checkpoint / restart) ------------------------------------ */
/* io.c:299-402 saveMarkersBinary(markerset, iMonte, iTime, elapsedTime): writes
 * `path` (the reference names it output/Markers.<iMonte>.<iTime>) in the
 * reference's layout -- int32 count of active markers (cellX != -1), double
 * elapsedTime, then per field the active markers in marker order: cpu, cellX,
 * cellY (int32), X, Y, VX, VY (double), Mat (int32), T, Tdot, eta, mu, exx, exy,
 * Eii, sxx, syy, sxy, p, rho, rhodot (double); the layout
 * post_processing/getBinaryMarkers.m reads.  *nSaved (optional): markers
 * written.  Synchronises; FVM_ERR_IO when the file cannot be written. */
int fvm_save_markers_binary(fvm_ctx *ctx, const fvm_markers *mk, const char *path, double elapsedTime,
                            int32_t *nSaved);
/* restart.c:8-125 restartFromMarkers: reads such a file into mk (at most
 * mk->maxMark markers, FVM_ERR_ARG beyond), sets mk->nMark and *elapsedTime.
 * Every slot of the capacity is first reset as resetMarker does
 * (markers.c:250-330: X = Y = -1, Mat = -1, the rest 0); the file's fields are
 * then written to the first nMark slots.  FVM_ERR_IO for a missing, short or
 * differently sized file.  Synchronises. */
int fvm_restart_from_markers(fvm_ctx *ctx, fvm_markers *mk, const char *path, double *elapsedTime);

/* io.c:150-296 saveNodalFields(..., OUTPUT_NODAL_FIELDS_BINARY): writes `path`
 * (the reference names it output/loadNodalFields_<iMonte>_<iTime>.petscbin) as 20
 * PETSc binary Vec records in the reference's order -- vx, vy, p, rho, rhodot,
 * etaN, etaS, kThermal, lastT, Cp, muN, muS, soxx, soyy, soxy, edotxx, edotyy,
 * edotxy, wxy, ha -- each the big-endian int32 class id 1211214, the big-endian
 * int32 length NX*NY and the values as big-endian doubles in natural order
 * ([jy][ix]).  ha (adiabatic heating, device NX*NY) may be NULL: zeros are
 * written.  PETSc's companion .info file is not written.  Synchronises. */
int fvm_save_nodal_fields(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, const double *ha,
                          const char *path);

/* ---- marker population control ------------------------------------------ */
/* markers.c:455-781 checkMarkerDensity(problem, markerset, grid, options, r), one
 * rank; called at main_subduction.c:250 after findMarkerCells, whose cells it
 * reads.  Quarter cells with fewer than opt->minMarkers markers are refilled to
 * ceil(NMX*(NX-1)/2.0*dx/LX) * ceil(NMY*(NY-1)/2.0*dy/LY) (dx, dy the quarter's
 * size), those with more than opt->maxMarkers lose the excess; added markers
 * recycle the slots of markers without a cell (last first) and then grow
 * mk->nMark up to mk->maxMark.  An added marker left of x[1] is a slab-inflow
 * marker, one right of x[NX-2] below plate_depth(LX) a mantle-inflow marker
 * (benchmarkInitialConditions.c:52-92, viscosity from material 0's law);
 * any other copies every field of the nearest marker within one cell width
 * and height of that marker's cell (slots refilled earlier in the call
 * included), keeping its own position, or stays as it is when there is none.
 * Random numbers: the reference's drand48 / PetscRandom are replaced by the
 * counter stream u(seed, s, c) = (mix(mix(seed + G (s + 1)) ^ c) >> 11) / 2^53,
 * mix the SplitMix64 output function, G = 0x9E3779B97F4A7C15: stream 1 with
 * c = q << 32 | k keys the k-th member (in index order) of over-full quarter q
 * (q = i + 2(NX-1) j, the reference's order) and the lowest keys leave; streams
 * 2 and 3 with c = the added marker's number give its x and y,
 * (hi - lo) * u + lo within its quarter.  Synchronises.  *nAdded is 0 when
 * nAdd + nMark - nOut equals maxMark (the reference then neither adds nor
 * aborts); FVM_ERR_CAPACITY when it exceeds it, FVM_ERR_NEGATIVE_ADD when a
 * target does not exceed its count (both abort in the reference). */
int fvm_check_marker_density(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const Materials *mat,
                             const Options *opt, uint64_t seed, int32_t *nAdded, int32_t *nRemoved);

/* ---- whole marker step, in the reference's call order (main_subduction.c) -- */
/* Thermal-convection step: findMarkerCells (:243), density loop (:255-264),
 * the projection list (or projectMarkersNodesAll2 when list == NULL, :268),
 * subgridTemperatureChanges (:415, thisT = nf->thisT), advectMarkersRK (:435).
 * The host solves between them are outside this call: their outputs
 * (vx, vy, wxy, thisT) are read from nf. */
int fvm_step_thermal(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                     const Materials *mat, const Options *opt, const BoundaryValues *bv,
                     double dt, int32_t nproj, const fvm_projection *list);

/* Visco-elasto-plastic step: findMarkerCells, density loop, checkPlasticYielding
 * + limitViscosity (:299-300), projectMarkersNodesAll2 into nf (:303); then, with
 * the host Stokes/nodalStressStrain outputs in host (vx, vy, p, edotxx, edotyy,
 * edotxy, wxy, soxx, soyy, soxy, dsxx, dsyy, dsxy, thisT):
 * updateMarkerStrainPressure (:341), updateGlobalStrainRateResidual (:345),
 * subgridStressChanges (:365), updateMarkerStress (:368),
 * subgridTemperatureChanges (:415), advectMarkersRK (:435).
 * Optional outputs (NULL to skip): *isYielding, *residual. */
int fvm_step_vep(fvm_ctx *ctx, const fvm_grid *g, fvm_markers *mk, const NodalArrays *nf,
                 const NodalArrays *host, const Materials *mat, const Options *opt,
                 const BoundaryValues *bv, double dt, int32_t *isYielding, double *residual);

/* ---- timing on the context's stream (HIP events) --------------------------- */
/* Per-phase kernel timing: while enabled, every kernel launch of the library is
 * bracketed by a pair of HIP events on the context's stream.  Phases:
 * FVM_PHASE_* below.  fvm_profile_read synchronises and returns, per phase,
 * the summed milliseconds and the number of launches. */
enum {
  FVM_PHASE_LOCATE = 0,    /* findMarkerCells */
  FVM_PHASE_SORT = 1,      /* bucket keys, scan, scatter, in-bucket order */
  FVM_PHASE_PARTIALS = 2,  /* projection pass 1 (per lattice cell) */
  FVM_PHASE_NODES = 3,     /* projection pass 2 (per node) */
  FVM_PHASE_SUBGRID_T1 = 4,
  FVM_PHASE_SUBGRID_T2 = 5,
  FVM_PHASE_ADVECT = 6,
  FVM_PHASE_SUBGRID_S1 = 7,
  FVM_PHASE_MARKER_UPDATE = 8, /* stress, strain/pressure, yield, viscosity, density */
  FVM_PHASE_OTHER = 9,
  FVM_PHASE_PERSISTENT = 10,   /* the batched ensemble's one-launch step (every phase) */
  FVM_NUM_PHASES = 11
};
int fvm_profile_enable(fvm_ctx *ctx, int32_t on);
int fvm_profile_read(fvm_ctx *ctx, double *ms, int32_t *launches);

int fvm_timer_start(fvm_ctx *ctx);
/* Milliseconds between fvm_timer_start and now (synchronises). */
int fvm_timer_stop(fvm_ctx *ctx, float *ms);

/* ---- A15: initialPressureGuess.c:6-184 ----------------------------------------
 * The plain initial pressure guess: the 5-point variable-spacing Poisson
 * problem for the cell-centre pressure with the density-gradient source,
 * assembled as in initialPressureGuess.c:59-135 (dummy rows on ix == 0 and
 * jy == 0, pressure pinned in cell (3, 3), zero-flux sides, dp/dy = rho g on
 * top and bottom), solved on the GPU (matrix-free operator, Jacobi BiCGStab,
 * relative residual <= rtol, at most maxit iterations).  rho: device NX*NY
 * nodal density.  The pressure goes to S[3*(jy*NX+ix) + 2] (the DOF_P slot of
 * the interleaved [vx, vy, p] Stokes vector, fdcode.h:242-245; velocities are
 * left untouched) and/or p_out (NX*NY); either may be NULL but not both.
 * Returns FVM_ERR_NOT_CONVERGED when rtol is not met (p is still written). */
int fvm_initial_pressure_guess(fvm_ctx *ctx, const fvm_grid *g, const double *rho, double gy,
                               double gx, double rtol, int32_t maxit, double *S, double *p_out,
                               int32_t *iters, double *relres);
/* The same operator applied to x (y = A x) and, if b != NULL, its right-hand
 * side -- for callers that run their own solver. */
int fvm_pressure_guess_apply(fvm_ctx *ctx, const fvm_grid *g, const double *rho, double gy,
                             double gx, const double *x, double *y, double *b);

/* ---- batched ensemble of independent problems ------------------------------
 * No counterpart in the reference, which runs a parameter sweep as separate
 * processes (main_subduction.c:167-172 loops over Monte-Carlo realisations).
 * An ensemble steps many independent thermal-marker problems (own grid, marker
 * set, nodal fields, materials, dt; sizes may differ) with one launch per
 * phase over all of them, optionally replayed as one captured HIP graph.
 * Each instance's step equals fvm_step_thermal with the same projection list
 * (findMarkerCells + density loop, projections, subgridTemperatureChanges,
 * advectMarkersRK), and with the context's lane count (fvm_ctx_set_lanes,
 * default 8) it is bitwise identical to it.
 * Every instance must use the same projection pattern (stag/method/material
 * kind per entry; the fields and property tables may differ); methods
 * ARITHMETIC and ARITHMETIC_LOCAL.  Device arrays stay owned by the caller and
 * must outlive the ensemble. */
typedef struct {
  fvm_grid grid;
  fvm_markers markers;   /* nMark is fixed at creation */
  NodalArrays nodal;     /* vx, vy, wxy, thisT read; lastT, rho, Cp, kThermal + list outputs written */
  int32_t nproj;
  const fvm_projection *proj;  /* copied at creation */
  Materials materials;
  BoundaryValues bv;
  double dt;
  double dsubgrid;       /* Options.subgridTemperatureDiffusivity */
  int32_t jaumann;       /* stress rotation: 0 exact, 1 -DJAUMANN (as fvm_ctx_set_jaumann) */
} fvm_instance;

/* Launch shape: threads per block {128, 256, 512}; lanes per lattice cell in
 * the projection pass {4, 8, 16, 32} (4: per-phase launches only); persistent
 * 0: one launch per phase,
 * 1: the whole step as one launch of resident workgroups that take work items
 * (blocks of every phase of every instance) from a ticket, an item waiting
 * only for the previous phase of its own instance (needs block 256 and a
 * projection list whose groups are all ARITHMETIC with the same stored field
 * count, at most 5). */
typedef struct {
  int32_t block;
  int32_t lanes;
  int32_t persistent;
} fvm_tuning;

typedef struct fvm_ensemble fvm_ensemble;

int fvm_ensemble_create(fvm_ctx *ctx, int32_t n, const fvm_instance *inst, fvm_ensemble **out);
int fvm_ensemble_destroy(fvm_ensemble *e);
/* One step of every instance, launched on the context's stream. */
int fvm_ensemble_step(fvm_ctx *ctx, fvm_ensemble *e);
int fvm_ensemble_set_tuning(fvm_ctx *ctx, fvm_ensemble *e, const fvm_tuning *t);
int fvm_ensemble_get_tuning(const fvm_ensemble *e, fvm_tuning *t);
int64_t fvm_ensemble_markers(const fvm_ensemble *e);
/* Capture one step as a HIP graph (replaces a previous capture). */
int fvm_ensemble_capture(fvm_ctx *ctx, fvm_ensemble *e);
/* Replay the captured step (captures first if needed). */
int fvm_ensemble_launch(fvm_ctx *ctx, fvm_ensemble *e);
/* Time every launch shape on `reps` graph replays and keep the fastest; each
 * trial advances the ensemble's state. */
int fvm_ensemble_autotune(fvm_ctx *ctx, fvm_ensemble *e, int32_t reps, fvm_tuning *best,
                          float *ms_per_step);
/* Markers outside the domain / with no valid cell, summed over instances, in
 * the last step (synchronises).  Returns FVM_ERR_STALLED if a persistent step
 * since the ensemble's creation gave up waiting (its results are incomplete). */
int fvm_ensemble_counters(fvm_ctx *ctx, fvm_ensemble *e, int32_t *nOut, int32_t *nBad);

#ifdef __cplusplus
}
#endif
#endif /* FVM_ABI_H */
