// computeRheology (viscosity.c:12-67) on the device, shared by the yield pass
// (ops.hip) and the markers that population control creates (population.hip).
#pragma once

#include <math.h>

#include "fvm_internal.h"

struct YieldParams {
  Materials mat;
  double etamin, etamax, fractionalEtamin;
  double slabAngle, decoupleThickness, minDecoupleDepth, maxDecoupleDepth, decoupleEtaScaler;
  double dt;
  // sin(F[m][k] / 180 * pi) and tan(slabAngle), evaluated once on the host with
  // the C library the reference calls (markers.c:1444-1462, viscosity.c:55)
  double sinF[FVM_MAXMAT][2];
  double tanSlab;
};

__device__ __forceinline__ int rheology(const YieldParams &P, int mm, double T, double X, double Y, double *eta,
                                        double *mu) {
  const Materials &mat = P.mat;
  const double thismu = mat.materialMu[mm];
  const int law = mat.hasEtaT[mm];
  double thiseta;
  if (law == 1) {
    thiseta = mat.materialEta[mm] * exp(mat.QonR[mm] * (1.0 / T - 1.0 / mat.Tref[mm]));
  } else if (law == 2) {
    const double T0 = mat.Tref[mm];
    thiseta = mat.materialEta[mm] * exp(mat.QonR[mm] / T0 * (1.0 - (T - T0) / T0));
  } else if (law == 3) {
    thiseta = mat.materialEta[mm] * T;
  } else if (law == 5) {
    thiseta = mat.materialEta[mm] * exp(-mat.QonR[mm] * T);
  } else if (law == 0) {
    thiseta = mat.materialEta[mm];
  } else if (law == 6) {
    const double depth = Y, distance = X;
    thiseta = mat.materialEta[mm];
    if (distance > depth / P.tanSlab) {
      if (distance < (depth / P.tanSlab + P.decoupleThickness)) {
        if (depth > P.minDecoupleDepth && depth < P.maxDecoupleDepth) thiseta *= P.decoupleEtaScaler;
      }
    }
  } else {
    return 1;
  }
  *eta = thiseta;
  *mu = thismu;
  return 0;
}

// The parameters from the reference's Materials and Options; the two
// transcendental constants come from the host C library, as the reference
// evaluates them.
static inline YieldParams make_yield_params(const Materials *mat, const Options *opt, double dt) {
  YieldParams P;
  P.mat = *mat;
  P.etamin = opt->etamin;
  P.etamax = opt->etamax;
  P.fractionalEtamin = opt->fractionalEtamin;
  P.slabAngle = opt->slabAngle;
  P.decoupleThickness = opt->decoupleThickness;
  P.minDecoupleDepth = opt->minDecoupleDepth;
  P.maxDecoupleDepth = opt->maxDecoupleDepth;
  P.decoupleEtaScaler = opt->decoupleEtaScaler;
  P.dt = dt;
  for (int i = 0; i < FVM_MAXMAT; i++)
    for (int k = 0; k < 2; k++) P.sinF[i][k] = sin(mat->F[i][k] / 180.0 * M_PI);
  P.tanSlab = tan(opt->slabAngle);
  return P;
}
