// Device bodies of the per-marker passes that both the single-instance
// kernels (ops.hip) and the batched ensemble kernels (ensemble.hip) run:
// advectMarkersRK (with subgridTemperatureChanges pass 2 in front of it) and
// subgridTemperatureChanges passes 1 and 2.  Reference expression order kept;
// compiled without contraction.
#pragma once
#include "fvm_internal.h"

// ---- A6: advectMarkersRK (markers.c:1865-2250) --------------------------------
__device__ __forceinline__ void advect_marker(int64_t m, const DGrid &g,
    double *__restrict__ MX,
    double *__restrict__ MY,
    double *__restrict__ MVX,
    double *__restrict__ MVY,
    double *__restrict__ Mwxy,
    double *__restrict__ sxx,
    double *__restrict__ syy,
    double *__restrict__ sxy,
    const double *__restrict__ exx,
    const double *__restrict__ eyy,
    const double *__restrict__ exy,
    double *__restrict__ Eii,
    const int32_t *__restrict__ cellXa,
    const int32_t *__restrict__ cellYa,
    const double *__restrict__ vx,
    const double *__restrict__ vy,
    const double *__restrict__ wxy,
    BoundaryValues bv,
    double dt,
    int jaumann,
    double *__restrict__ MT,
    double *__restrict__ MTdot,
    const double *__restrict__ dTn,
    const double *__restrict__ dTsub,
    double dt_thermal) {
  const int c0x = cellXa[m], c0y = cellYa[m];
  if (!cell_ok(c0x, c0y, g.NX, g.NY)) return;
  const int NX = g.NX, NY = g.NY;
  const int XCMAX = NX - 2, XCMIN = 0;
  const double X0 = MX[m], Y0 = MY[m];
  if (MT) {
    // subgridTemperatureChanges.c:89-103 (pass 2) on the same stored cell,
    // fused here because advectMarkersRK starts from that cell too
    const double dx = g.x[c0x + 1] - g.x[c0x];
    const double dy = g.y[c0y + 1] - g.y[c0y];
    const double mdx = (X0 - g.x[c0x]) / dx;
    const double mdy = (Y0 - g.y[c0y]) / dy;
    const double dtm = (gat(dTn, NX, NY, c0y, c0x) * (1.0 - mdy) * (1.0 - mdx) +
                        gat(dTn, NX, NY, c0y + 1, c0x) * mdy * (1.0 - mdx) +
                        gat(dTn, NX, NY, c0y, c0x + 1) * (1.0 - mdy) * mdx +
                        gat(dTn, NX, NY, c0y + 1, c0x + 1) * mdy * mdx) +
                       dTsub[m];
    MT[m] = MT[m] + dtm;
    MTdot[m] = dtm / dt_thermal;
  }
  double xcur = X0, ycur = Y0;
  double mvx[4], mvy[4], mwxy[4];
#pragma unroll
  for (int irk = 0; irk < 4; irk++) {
    int cellX = c0x, cellY = c0y;
    if (xcur > g.x[cellX + 1]) {
      while (xcur > g.x[cellX + 1] && cellX < XCMAX) cellX++;
    } else if (xcur < g.x[cellX]) {
      while (xcur < g.x[cellX] && cellX > XCMIN) cellX--;
    }
    if (ycur > g.y[cellY + 1]) {
      while (ycur > g.y[cellY + 1] && cellY < NY - 2) cellY++;
    } else if (ycur < g.y[cellY]) {
      while (ycur < g.y[cellY] && cellY > 0) cellY--;
    }
    int vxX = cellX, vxY = cellY;
    if (vxY > -1 && ycur < g.yc[vxY + 1]) vxY--;
    int vyX = cellX;
    if (vyX > -1 && xcur < g.xc[vyX + 1]) vyX--;
    int vyY = cellY;
    if (vxX < 0) vxX = 0;
    else if (vxX > NX - 2) vxX = NX - 2;
    if (vyX < -1) vyX = -1;
    else if (vyX > NX - 2) vyX = NX - 2;
    if (vxY < -1) vxY = -1;
    else if (vxY > NY - 2) vxY = NY - 2;
    if (vyY < 0) vyY = 0;
    else if (vyY > NY - 2) vyY = NY - 2;

    double deltaxm = xcur - g.x[vxX];
    double deltaym = ycur - g.yc[vxY + 1];
    double dx = g.x[vxX + 1] - g.x[vxX];
    double dy = g.yc[vxY + 2] - g.yc[vxY + 1];
    // vxX == cellX (cellX stays in [0, NX-2]), so this quotient is also the wxy mdx below
    const double qx = deltaxm / dx;
    double vxul, vxur;
    if (vxY == -1 && bv.mechBCTop.type[0] == 0) {
      vxul = 2.0 * bv.mechBCTop.value[0] - gat(vx, NX, NY, vxY + 1, vxX);
      vxur = 2.0 * bv.mechBCTop.value[0] - gat(vx, NX, NY, vxY + 1, vxX + 1);
    } else if (vxY == -1 && bv.mechBCTop.type[0] == 1) {
      vxul = gat(vx, NX, NY, vxY + 1, vxX);
      vxur = gat(vx, NX, NY, vxY + 1, vxX + 1);
    } else {
      vxul = gat(vx, NX, NY, vxY, vxX);
      vxur = gat(vx, NX, NY, vxY, vxX + 1);
    }
    const double vxll = gat(vx, NX, NY, vxY + 1, vxX);
    const double vxlr = gat(vx, NX, NY, vxY + 1, vxX + 1);
    const double qyv = deltaym / dy;
    double a = vxul * (1.0 - qx) * (1.0 - qyv);
    a += vxur * (qx) * (1.0 - qyv);
    a += vxll * (1.0 - qx) * (qyv);
    a += vxlr * (deltaxm * deltaym) / dx / dy;
    mvx[irk] = a;

    double vyul;
    const double vyur = gat(vy, NX, NY, vyY, vyX + 1);
    double vyll;
    const double vylr = gat(vy, NX, NY, vyY + 1, vyX + 1);
    if (vyX == -1 && bv.mechBCLeft.type[1] == 0) {
      vyul = 2.0 * bv.mechBCLeft.value[1] - gat(vy, NX, NY, vyY, vyX + 1);
      vyll = 2.0 * bv.mechBCLeft.value[1] - gat(vy, NX, NY, vyY + 1, vyX + 1);
    } else if (vyX == -1 && bv.mechBCLeft.type[1] == 1) {
      vyul = gat(vy, NX, NY, vyY, vyX + 1);
      vyll = gat(vy, NX, NY, vyY + 1, vyX + 1);
    } else {
      vyul = gat(vy, NX, NY, vyY, vyX);
      vyll = gat(vy, NX, NY, vyY + 1, vyX);
    }
    deltaxm = xcur - g.xc[vyX + 1];
    deltaym = ycur - g.y[vyY];
    dx = g.xc[vyX + 2] - g.xc[vyX + 1];
    dy = g.y[vyY + 1] - g.y[vyY];
    // vyY == cellY (cellY stays in [0, NY-2]), so this quotient is also the wxy mdy below
    const double qy = deltaym / dy;
    const double qxv = deltaxm / dx;
    double b = vyul * (1.0 - qxv) * (1.0 - qy);
    b += vyur * (qxv) * (1.0 - qy);
    b += vyll * (1.0 - qxv) * (qy);
    b += vylr * (deltaxm * deltaym) / dx / dy;
    mvy[irk] = b;

    // markers.c:2089-2090: (xcur - x[cellX]) / (x[cellX+1] - x[cellX]), the same
    // operands as qx and qy above, hence the same values
    const double mdx = qx;
    const double mdy = qy;
    double w = (1.0 - mdx) * (1.0 - mdy) * gat(wxy, NX, NY, cellY, cellX);
    w += (mdx) * (1.0 - mdy) * gat(wxy, NX, NY, cellY, cellX + 1);
    w += (1.0 - mdx) * (mdy) * gat(wxy, NX, NY, cellY + 1, cellX);
    w += (mdx) * (mdy) * gat(wxy, NX, NY, cellY + 1, cellX + 1);
    mwxy[irk] = w;

    if (irk < 3) {
      if (irk < 2) {
        xcur = X0 + dt / 2 * mvx[irk];
        ycur = Y0 + dt / 2 * mvy[irk];
      } else {
        xcur = X0 + dt * mvx[irk];
        ycur = Y0 + dt * mvy[irk];
      }
    }
  }
  const double VX = (mvx[0] + 2 * mvx[1] + 2 * mvx[2] + mvx[3]) / 6;
  const double VY = (mvy[0] + 2 * mvy[1] + 2 * mvy[2] + mvy[3]) / 6;
  const double wm = (mwxy[0] + 2 * mwxy[1] + 2 * mwxy[2] + mwxy[3]) / 6.0;
  MVX[m] = VX;
  MVY[m] = VY;
  Mwxy[m] = wm;
  MX[m] = X0 + dt * VX;
  MY[m] = Y0 + dt * VY;

  const double msxx0 = sxx[m], msyy0 = syy[m], msxy0 = sxy[m];
  if (jaumann) {
    const double wxy1 = wm * dt;
    sxx[m] = msxx0 + 2.0 * msxy0 * -wxy1;
    syy[m] = msyy0 + 2.0 * msxy0 * wxy1;
    sxy[m] = msxy0 + (msxx0 - msyy0) * wxy1;
  } else {
    const double thetaz = -wm * dt;
    const double c = cos(thetaz), s = sin(thetaz), s2 = sin(2.0 * thetaz), c2 = cos(2.0 * thetaz);
    // pow(v, 2.0) of markers.c:2162-2163: gcc -O3 (makefile:14) folds it to v*v,
    // which is what the reference binary executes
    sxx[m] = msxx0 * (c * c) + msyy0 * (s * s) - msxy0 * s2;
    syy[m] = msxx0 * (s * s) + msyy0 * (c * c) + msxy0 * s2;
    sxy[m] = 0.5 * (msxx0 - msyy0) * s2 + msxy0 * c2;
  }
  const double e11 = exx[m], e22 = eyy[m], e12 = exy[m];
  Eii[m] = Eii[m] + dt * sqrt(0.5 * (e11 * e11 + e22 * e22) + e12 * e12);
}

// ---- A7: subgridTemperatureChanges pass 1 (subgridTemperatureChanges.c:46-70) ----
__device__ __forceinline__ void subgrid_t1_marker(int64_t m, const DGrid &g,
    const double *__restrict__ MX,
    const double *__restrict__ MY,
    const double *__restrict__ MT,
    const int32_t *__restrict__ cX,
    const int32_t *__restrict__ cY,
    const double *__restrict__ T,
    const double *__restrict__ rho,
    const double *__restrict__ Cp,
    const double *__restrict__ kT,
    double rho0,
    double dt,
    double dsubgrid,
    double *__restrict__ dTsub) {
  const int cellX = cX[m], cellY = cY[m];
  if (!cell_ok(cellX, cellY, g.NX, g.NY)) {
    dTsub[m] = 0.0;
    return;
  }
  const int NX = g.NX, NY = g.NY;
  const double dx = g.x[cellX + 1] - g.x[cellX];
  const double dy = g.y[cellY + 1] - g.y[cellY];
  const double mdx = (MX[m] - g.x[cellX]) / dx;
  const double mdy = (MY[m] - g.y[cellY]) / dy;
  const double Tm = gat(T, NX, NY, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) +
                    gat(T, NX, NY, cellY, cellX + 1) * mdx * (1.0 - mdy) +
                    gat(T, NX, NY, cellY + 1, cellX) * (1.0 - mdx) * (mdy) +
                    gat(T, NX, NY, cellY + 1, cellX + 1) * mdx * mdy;
  const double rhom = rho0 + gat(rho, NX, NY, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) +
                      gat(rho, NX, NY, cellY, cellX + 1) * (mdx) * (1.0 - mdy) +
                      gat(rho, NX, NY, cellY + 1, cellX) * (1.0 - mdx) * (mdy) +
                      gat(rho, NX, NY, cellY + 1, cellX + 1) * mdx * mdy;
  const double Cpm = gat(Cp, NX, NY, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) +
                     gat(Cp, NX, NY, cellY, cellX + 1) * (mdx) * (1.0 - mdy) +
                     gat(Cp, NX, NY, cellY + 1, cellX) * (1.0 - mdx) * (mdy) +
                     gat(Cp, NX, NY, cellY + 1, cellX + 1) * mdx * mdy;
  const double km = gat(kT, NX, NY, cellY, cellX) * (1.0 - mdx) * (1.0 - mdy) +
                    gat(kT, NX, NY, cellY, cellX + 1) * (mdx) * (1.0 - mdy) +
                    gat(kT, NX, NY, cellY + 1, cellX) * (1.0 - mdx) * (mdy) +
                    gat(kT, NX, NY, cellY + 1, cellX + 1) * mdx * mdy;
  const double dTdiff = Cpm * rhom / km / (2.0 / (dx * dx) + 2.0 / (dy * dy));
  dTsub[m] = (Tm - MT[m]) * (1.0 - exp(-dsubgrid * dt / dTdiff));
}

// ---- A7: subgridTemperatureChanges pass 2 (subgridTemperatureChanges.c:89-103) ---
__device__ __forceinline__ void subgrid_t2_marker(int64_t m, const DGrid &g,
    const double *__restrict__ MX,
    const double *__restrict__ MY,
    double *__restrict__ MT,
    double *__restrict__ MTdot,
    const int32_t *__restrict__ cX,
    const int32_t *__restrict__ cY,
    const double *__restrict__ dT,
    const double *__restrict__ dTsub,
    double dt) {
  const int cellX = cX[m], cellY = cY[m];
  if (!cell_ok(cellX, cellY, g.NX, g.NY)) return;
  const int NX = g.NX, NY = g.NY;
  const double dx = g.x[cellX + 1] - g.x[cellX];
  const double dy = g.y[cellY + 1] - g.y[cellY];
  const double mdx = (MX[m] - g.x[cellX]) / dx;
  const double mdy = (MY[m] - g.y[cellY]) / dy;
  const double dtm = (gat(dT, NX, NY, cellY, cellX) * (1.0 - mdy) * (1.0 - mdx) +
                      gat(dT, NX, NY, cellY + 1, cellX) * mdy * (1.0 - mdx) +
                      gat(dT, NX, NY, cellY, cellX + 1) * (1.0 - mdy) * mdx +
                      gat(dT, NX, NY, cellY + 1, cellX + 1) * mdy * mdx) +
                     dTsub[m];
  MT[m] = MT[m] + dtm;
  MTdot[m] = dtm / dt;
}
