/*
 * structures.h -- data types of the fvmic2d marker hot path.
 *
 * The reference ships an empty structures.h; its types live in fdcode.h:13-238.
 * The types below keep that memory layout field for field (same order, same
 * widths: PetscScalar = double, PetscInt = int32), so an AoS marker block written
 * by the reference (216 bytes per Marker) can be handed to this library as-is.
 * PETSc handles (DM, Vec) are replaced by plain pointers: DM is an opaque
 * pointer, and a nodal field is a dense row-major [jy][ix] array of NX*NY doubles.
 *
 * tests/test_abi_cpu.py compiles this header with gcc and checks Marker's size
 * (216 B) and every field offset against the reference's layout (X 0 ...
 * isYielding 208), and the other structs against the Python ctypes mirror.
 */
#ifndef FVM_STRUCTURES_H
#define FVM_STRUCTURES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FVM_MAXMAT 8 /* fdcode.h:11 MAXMAT */

/* fdcode.h:27-29 */
typedef struct {
  double T11, T22, T12;
} Tensor33s;

/* fdcode.h:13-24.  x/y/xc/yc point two elements into arrays of NX+4 / NY+4
 * doubles (gridGenerator.c:123-130), so indices -2..N+1 are addressable. */
typedef struct {
  void *da, *vda;            /* opaque: PETSc DMDA handles in the reference */
  double LX, LY;
  int32_t NX, NY;
  int32_t mp, np;
  double *x, *y, *xc, *yc;
  int32_t xperiodic, yperiodic;
} GridData;

/* fdcode.h:74-95 */
typedef struct {
  int32_t nMaterials;
  double materialEta[FVM_MAXMAT];
  double materialRho[FVM_MAXMAT];
  double materialkThermal[FVM_MAXMAT];
  double materialCp[FVM_MAXMAT];
  double materialAlpha[FVM_MAXMAT];
  double materialMu[FVM_MAXMAT];
  double materialFriction[FVM_MAXMAT];
  double materialCohesion[FVM_MAXMAT];
  int32_t hasPlasticity[FVM_MAXMAT];
  double gamma[FVM_MAXMAT][2];
  double C[FVM_MAXMAT][2];
  double F[FVM_MAXMAT][2];
  int32_t hasEtaT[FVM_MAXMAT];
  double QonR[FVM_MAXMAT];
  double Tref[FVM_MAXMAT];
} Materials;

/* fdcode.h:97-129 (no TEXTURE): 216 bytes. A marker is inactive iff cellX == -1. */
typedef struct {
  double X;                  /*   0 */
  double Y;                  /*   8 */
  double VX;                 /*  16 */
  double VY;                 /*  24 */
  int32_t Mat;               /*  32 */
  double T, Tdot;            /*  40, 48 */
  Tensor33s s;               /*  56: deviatoric stress T11, T22, T12 */
  double eta;                /*  80 */
  double mu;                 /*  88 */
  double rho;                /*  96 */
  double rhodot;             /* 104 */
  double p;                  /* 112 */
  Tensor33s e;               /* 120: strain rate */
  Tensor33s E;               /* 144: total strain */
  double Eii;                /* 168 */
  double wxy;                /* 176 */
  int32_t cellX, cellY, cpu; /* 184, 188, 192 */
  double strainRateResidual; /* 200 */
  int32_t isYielding;        /* 208 */
} Marker;

/* fdcode.h:131-138 */
typedef struct {
  int32_t nMark;
  int32_t maxMark;
  int32_t nOut;
  Marker *markers;
} MarkerSet;

/* fdcode.h:141-152.  type: 0 prescribed velocity, 1 free slip, 2 periodic */
typedef struct {
  int32_t type[3];
  double value[3];
} BCInfo;

typedef struct {
  BCInfo mechBCLeft;
  BCInfo mechBCRight;
  BCInfo mechBCTop;
  BCInfo mechBCBottom;
} BoundaryValues;

/* fdcode.h:155-211 */
typedef struct {
  int32_t NX, NY;
  double LX, LY;
  int32_t NMX;
  int32_t NMY;
  double gridRatio;
  int32_t nTime;
  int32_t restartStep;
  double totalTime;
  int32_t maxNumPlasticity;
  int32_t plasticDelay;
  double Tperturb;
  int32_t shearHeating, adiabaticHeating;
  double gy;
  double gx;
  double gz;
  BCInfo mechBCLeft, mechBCRight, mechBCTop, mechBCBottom;
  BCInfo thermalBCLeft, thermalBCRight, thermalBCTop, thermalBCBottom;
  double fractionalEtamin;
  double etamin;
  double etamax;
  double dtMax;
  double displacementStepLimit;
  double maxTempChange;
  int32_t saveInterval;
  int32_t markerSaveInterval;
  double subgridStressDiffusivity;
  double subgridTemperatureDiffusivity;
  double maxMarkFactor;
  int32_t minMarkers, maxMarkers;
  double grainSize;
  double overridingPlateThickness;
  double slabAge;
  double slabAngle;
  double slabVelocity;
  double rootThickness;
  double rootCenter;
  double rootWidth;
  int32_t staticVelocity;
  double minDecoupleDepth;
  double maxDecoupleDepth;
  double decoupleThickness;
  double decoupleEtaScaler;
} Options;

/* markerProjection.h:3-8 */
typedef enum {
  ARITHMETIC = 0,
  ARITHMETIC_LOCAL = 1,
  HARMONIC = 2,
  GEOMETRIC = 3
} PROJECTION_METHOD;

/* The nodal fields the marker path reads or writes (fdcode.h:40-71), each a
 * dense NX*NY row-major array ([jy][ix], ix fastest).  Host pointers for the
 * oracle and the compatibility layer, device pointers for the C ABI. */
typedef struct {
  double *rho, *rhodot, *etaS, *etaN;
  double *thisT, *lastT, *kThermal, *Cp;
  double *edotxx, *edotxy, *edotyy;
  double *soxx, *soxy, *soyy;
  double *dsxx, *dsyy, *dsxy;
  double *muS, *muN;
  double *vx, *vy, *p;
  double *wxy;
  double *strainRateResidual;
  double *eii, *sii; /* second invariants (nodalStressStrain.c:375-389) */
} NodalArrays;

#ifdef __cplusplus
}
#endif
#endif /* FVM_STRUCTURES_H */
