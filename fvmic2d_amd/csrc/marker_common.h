// Device bodies of the per-marker passes that both the single-instance
// kernels (ops.hip) and the batched ensemble kernels (ensemble.hip) run:
// advectMarkersRK (with subgridTemperatureChanges pass 2 in front of it) and
// subgridTemperatureChanges passes 1 and 2.  Reference expression order kept;
// compiled without contraction.
#pragma once
#include "fvm_internal.h"

// ---- A6: advectMarkersRK (markers.c:1865-2250) --------------------------------
// PER: x periodic (g.per == 1).  The walk may then leave [0, NX-1] by one cell
// (XCMIN = -1, XCMAX = NX-1, markers.c:1908-1915), the staggered x indices are
// not clamped (:1965-1971), node columns wrap, the left-wall ghost rule is off
// (:2014-2023) and the new X is wrapped into [0, LX] (:2225-2231).  The grid
// coordinates are then read at indices -1 .. NX+1 (set by makePeriodic).
template <bool PER>
__device__ __forceinline__ void advect_marker(int64_t m, const DGrid &g,
    double *__restrict__ MX,
    double *__restrict__ MY,
    double *__restrict__ MVX,
    double *__restrict__ MVY,
    double *__restrict__ Mwxy,
    const int32_t *__restrict__ cellXa,
    const int32_t *__restrict__ cellYa,
    const double *__restrict__ vx,
    const double *__restrict__ vy,
    const double *__restrict__ wxy,
    BoundaryValues bv,
    double dt,
    double *__restrict__ MT,
    double *__restrict__ MTdot,
    const double *__restrict__ dTn,
    const double *__restrict__ dTsub,
    double dt_thermal) {
  const int c0x = cellXa[m], c0y = cellYa[m];
  if (!cell_ok(c0x, c0y, g.ncx(), g.NY)) return;
  const int NX = g.NX, NY = g.NY;
  const int XCMAX = PER ? NX - 1 : NX - 2, XCMIN = PER ? -1 : 0;
  // column offset of node ix + 1 from node ix (wraps past the last column)
  auto right = [&](int ix) -> int { return (PER && ix + 1 >= NX) ? 1 - NX : 1; };
  const double X0 = MX[m], Y0 = MY[m];
  if (MT) {
    // subgridTemperatureChanges.c:89-103 (pass 2) on the same stored cell,
    // fused here because advectMarkersRK starts from that cell too
    const double dx = g.x[c0x + 1] - g.x[c0x];
    const double dy = g.y[c0y + 1] - g.y[c0y];
    const double mdx = (X0 - g.x[c0x]) / dx;
    const double mdy = (Y0 - g.y[c0y]) / dy;
    // the stored cell is inside the grid: all four corners are owned nodes
    const double *d0 = dTn + (int64_t)c0y * NX + c0x, *d1 = d0 + NX;
    const int r0 = right(c0x);
    const double dtm = (d0[0] * (1.0 - mdy) * (1.0 - mdx) + d1[0] * mdy * (1.0 - mdx) + d0[r0] * (1.0 - mdy) * mdx +
                        d1[r0] * mdy * mdx) +
                       dTsub[m];
    MT[m] = MT[m] + dtm;
    MTdot[m] = dtm / dt_thermal;
  }
  double xcur = X0, ycur = Y0;
  // RK4 weights (k1 + 2 k2 + 2 k3 + k4) / 6 (markers.c:2118-2120), summed in the
  // expression's left-to-right order as the stages complete
  double svx = 0.0, svy = 0.0, swxy = 0.0;
  // the stages stay a loop: unrolled, they held 116-126 VGPRs (four waves per
  // SIMD); rolled, 90-94 (five), and the issue-bound stages then hide more of
  // their gather latency (ensemble advection 3.16 ms instead of 3.31)
#pragma unroll 1
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
    if (!PER) {
      if (vxX < 0) vxX = 0;
      else if (vxX > NX - 2) vxX = NX - 2;
      if (vyX < -1) vyX = -1;
      else if (vyX > NX - 2) vyX = NX - 2;
    }
    if (vxY < -1) vxY = -1;
    else if (vxY > NY - 2) vxY = NY - 2;
    if (vyY < 0) vyY = 0;
    else if (vyY > NY - 2) vyY = NY - 2;

    double deltaxm = xcur - g.x[vxX];
    double deltaym = ycur - g.yc[vxY + 1];
    double dx = g.x[vxX + 1] - g.x[vxX];
    double dy = g.yc[vxY + 2] - g.yc[vxY + 1];
    // vxX == cellX (cellX stays in [0, NX-2], or [-1, NX-1] when PER), so this
    // quotient is also the wxy mdx below
    const double qx = deltaxm / dx;
    // vx nodes: rows vxY (-1 is the ghost row, which reads 0) and vxY + 1,
    // columns vxX and vxX + 1 (both owned: vxX is in [0, NX-2]; when PER the
    // columns wrap).  Loads are unconditional from a valid address; the ghost
    // row is selected away.
    const bool vrow = vxY >= 0;
    const int vxc = PER && vxX < 0 ? vxX + NX : vxX;
    const int vxr = right(vxc);
    const double *vr1 = vx + (int64_t)(vxY + 1) * NX + vxc;
    const double *vr0 = vrow ? vr1 - NX : vr1;
    const double vxll = vr1[0], vxlr = vr1[vxr];
    const double vx0 = vr0[0], vx1 = vr0[vxr];
    double vxul, vxur;
    if (vxY == -1 && bv.mechBCTop.type[0] == 0) {
      vxul = 2.0 * bv.mechBCTop.value[0] - vxll;
      vxur = 2.0 * bv.mechBCTop.value[0] - vxlr;
    } else if (vxY == -1 && bv.mechBCTop.type[0] == 1) {
      vxul = vxll;
      vxur = vxlr;
    } else {
      vxul = vrow ? vx0 : 0.0;
      vxur = vrow ? vx1 : 0.0;
    }
    const double qyv = deltaym / dy;
    double a = vxul * (1.0 - qx) * (1.0 - qyv);
    a += vxur * (qx) * (1.0 - qyv);
    a += vxll * (1.0 - qx) * (qyv);
    a += vxlr * (deltaxm * deltaym) / dx / dy;

    // vy nodes: columns vyX (-1 is the ghost column, which reads 0; when PER
    // it is column NX-1) and vyX + 1, rows vyY and vyY + 1 (both owned: vyY
    // is in [0, NY-2])
    const bool vcol = PER || vyX >= 0;
    const int vyc = PER && vyX + 1 >= NX ? vyX + 1 - NX : vyX + 1;  // column of vyX + 1
    const int vyl = (PER && vyc == 0) ? NX - 1 : -1;                // offset of column vyX
    const double *q0 = vy + (int64_t)vyY * NX + vyc, *q1 = q0 + NX;
    const double vyur = q0[0], vylr = q1[0];
    const double vy0 = (vcol ? q0 + vyl : q0)[0], vy1 = (vcol ? q1 + vyl : q1)[0];
    double vyul, vyll;
    if (!PER && vyX == -1 && bv.mechBCLeft.type[1] == 0) {
      vyul = 2.0 * bv.mechBCLeft.value[1] - vyur;
      vyll = 2.0 * bv.mechBCLeft.value[1] - vylr;
    } else if (!PER && vyX == -1 && bv.mechBCLeft.type[1] == 1) {
      vyul = vyur;
      vyll = vylr;
    } else {
      vyul = vcol ? vy0 : 0.0;
      vyll = vcol ? vy1 : 0.0;
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

    // markers.c:2089-2090: (xcur - x[cellX]) / (x[cellX+1] - x[cellX]), the same
    // operands as qx and qy above, hence the same values
    const double mdx = qx;
    const double mdy = qy;
    // cellX in [0, NX-2] ([-1, NX-1] wrapped when PER), cellY in [0, NY-2]:
    // four owned nodes
    const int wc = PER && cellX < 0 ? cellX + NX : cellX;
    const int wr = right(wc);
    const double *w0 = wxy + (int64_t)cellY * NX + wc, *w1 = w0 + NX;
    double w = (1.0 - mdx) * (1.0 - mdy) * w0[0];
    w += (mdx) * (1.0 - mdy) * w0[wr];
    w += (1.0 - mdx) * (mdy) * w1[0];
    w += (mdx) * (mdy) * w1[wr];
    if (irk == 0) {
      svx = a;
      svy = b;
      swxy = w;
    } else if (irk < 3) {
      svx = svx + 2 * a;
      svy = svy + 2 * b;
      swxy = swxy + 2 * w;
    } else {
      svx = svx + a;
      svy = svy + b;
      swxy = swxy + w;
    }
    if (irk < 3) {
      if (irk < 2) {
        xcur = X0 + dt / 2 * a;
        ycur = Y0 + dt / 2 * b;
      } else {
        xcur = X0 + dt * a;
        ycur = Y0 + dt * b;
      }
    }
  }
  const double VX = svx / 6;
  const double VY = svy / 6;
  const double wm = swxy / 6.0;
  MVX[m] = VX;
  MVY[m] = VY;
  Mwxy[m] = wm;
  double Xn = X0 + dt * VX;
  if (PER) {
    if (Xn > g.LX) Xn -= g.LX;
    else if (Xn < 0) Xn += g.LX;
  }
  MX[m] = Xn;
  MY[m] = Y0 + dt * VY;
}

// ---- A6, second pass: the stress rotation by the marker spin and the strain
// invariant update (markers.c:2141-2164, 2223), after advect_marker has stored the spin
// wxy = (k1 + 2 k2 + 2 k3 + k4) / 6.  A pass of its own: the RK4 pass then holds
// fewer live registers and hides more of its gather latency, and this one streams.
// ncx: basic cells along x (NX - 1, or NX on a periodic grid)
__device__ __forceinline__ void rotate_eii_marker(int64_t m, int ncx, int NY,
    const int32_t *__restrict__ cellXa,
    const int32_t *__restrict__ cellYa,
    const double *__restrict__ Mwxy,
    double *__restrict__ sxx,
    double *__restrict__ syy,
    double *__restrict__ sxy,
    const double *__restrict__ exx,
    const double *__restrict__ eyy,
    const double *__restrict__ exy,
    double *__restrict__ Eii,
    double dt,
    int jaumann) {
  if (!cell_ok(cellXa[m], cellYa[m], ncx, NY)) return;
  const double wm = Mwxy[m];
  const double msxx0 = sxx[m], msyy0 = syy[m], msxy0 = sxy[m];
  if (jaumann) {
    const double wxy1 = wm * dt;
    sxx[m] = msxx0 + 2.0 * msxy0 * -wxy1;
    syy[m] = msyy0 + 2.0 * msxy0 * wxy1;
    sxy[m] = msxy0 + (msxx0 - msyy0) * wxy1;
  } else {
    const double thetaz = -wm * dt;
    // one shared range reduction per angle (sincos); the values are those of sin and cos
    double c, s, s2, c2;
    sincos(thetaz, &s, &c);
    sincos(2.0 * thetaz, &s2, &c2);
    // pow(v, 2.0) of markers.c:2162-2163: gcc -O3 (makefile:14) folds it to v*v,
    // which is what the reference binary executes
    sxx[m] = msxx0 * (c * c) + msyy0 * (s * s) - msxy0 * s2;
    syy[m] = msxx0 * (s * s) + msyy0 * (c * c) + msxy0 * s2;
    sxy[m] = 0.5 * (msxx0 - msyy0) * s2 + msxy0 * c2;
  }
  const double e11 = exx[m], e22 = eyy[m], e12 = exy[m];
  Eii[m] = Eii[m] + dt * sqrt(0.5 * (e11 * e11 + e22 * e22) + e12 * e12);
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
  if (!cell_ok(cellX, cellY, g.ncx(), g.NY)) return;
  const int NX = g.NX, NY = g.NY, P = g.per;
  const double dx = g.x[cellX + 1] - g.x[cellX];
  const double dy = g.y[cellY + 1] - g.y[cellY];
  const double mdx = (MX[m] - g.x[cellX]) / dx;
  const double mdy = (MY[m] - g.y[cellY]) / dy;
  const double dtm = (gat(dT, NX, NY, P, cellY, cellX) * (1.0 - mdy) * (1.0 - mdx) +
                      gat(dT, NX, NY, P, cellY + 1, cellX) * mdy * (1.0 - mdx) +
                      gat(dT, NX, NY, P, cellY, cellX + 1) * (1.0 - mdy) * mdx +
                      gat(dT, NX, NY, P, cellY + 1, cellX + 1) * mdy * mdx) +
                     dTsub[m];
  MT[m] = MT[m] + dtm;
  MTdot[m] = dtm / dt;
}
