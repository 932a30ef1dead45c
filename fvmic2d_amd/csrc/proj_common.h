// Device bodies of the deterministic projection (markerProjection.c:71-250 and
// the scatter parts of markers.c:1507-1594, subgridStressChanges.c:128-233),
// shared by the single-instance kernels (project.hip) and the batched ensemble
// kernels (ensemble.hip).
//
// Node family (stagx, stagy) has its own lattice of cells; a marker belongs to
// lattice cell l = cellX + {0 (stag 0), bx (stag -1), bx-1 (stag +1)} along x
// (findCellStag2D, markerProjection.c:252-269), likewise along y, and it
// contributes to the lattice cell's four corner nodes.
#pragma once
#include "fvm_internal.h"

struct Lat {
  int lo, n;  // lattice cell index range [lo, lo+n)
};
__host__ __device__ __forceinline__ Lat lattice(int stag, int N) {
  Lat L;
  if (stag == 0) {
    L.lo = 0;
    L.n = N - 1;
  } else if (stag == -1) {
    L.lo = 0;
    L.n = N;
  } else {
    L.lo = -1;
    L.n = N;
  }
  return L;
}

// fractional position inside lattice cell l (markerProjection.c:127-154)
__device__ __forceinline__ double frac(int stag, const double *__restrict__ n, const double *__restrict__ c, int l,
                                       double X) {
  if (stag == -1) return (X - c[l]) / (c[l + 1] - c[l]);
  if (stag == 0) return (X - n[l]) / (n[l + 1] - n[l]);
  return (X - c[l + 1]) / (c[l + 2] - c[l + 1]);
}

// Accumulate one marker's contribution into the four corner sums.
template <int METHOD, int NF>
__device__ __forceinline__ void accumulate(double mdx, double mdy, const double (&f)[NF], double (&sw)[4],
                                           double (&sf)[4][NF]) {
  if (METHOD == PM_ARITH) {
    const double w0 = (1.0 - mdx) * (1.0 - mdy), w1 = mdx * (1.0 - mdy), w2 = (1.0 - mdx) * (mdy), w3 = mdx * mdy;
    sw[0] += w0;
    sw[1] += w1;
    sw[2] += w2;
    sw[3] += w3;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sf[0][k] += w0 * f[k];
      sf[1][k] += w1 * f[k];
      sf[2][k] += w2 * f[k];
      sf[3][k] += w3 * f[k];
    }
  } else if (METHOD == PM_LOCAL || METHOD == PM_HARM) {
    // one corner: nearest in x (mdx >= 0.5 -> right) and y; the other corners
    // receive +0.0, which leaves their sums unchanged (and keeps a NaN or Inf
    // value confined to the chosen node, as in the reference)
    const bool hx = mdx >= 0.5, hy = mdy >= 0.5;
    double w = 1.0;
    w *= hx ? mdx : (1.0 - mdx);
    w *= hy ? mdy : (1.0 - mdy);
    const int c = (hx ? 1 : 0) | (hy ? 2 : 0);
    double v[NF];
#pragma unroll
    for (int k = 0; k < NF; k++) v[k] = METHOD == PM_LOCAL ? w * f[k] : 1.0 / f[k];
    const double wc = METHOD == PM_LOCAL ? w : 1.0;
#pragma unroll
    for (int cc = 0; cc < 4; cc++) {
      const bool on = cc == c;
      sw[cc] += on ? wc : 0.0;
#pragma unroll
      for (int k = 0; k < NF; k++) sf[cc][k] += on ? v[k] : 0.0;
    }
  } else if (METHOD == PM_TIES) {
    const bool xl = mdx <= 0.5, xh = mdx >= 0.5, yl = mdy <= 0.5, yh = mdy >= 0.5;
    if (xl && yl) {
      const double w = (1.0 - mdx) * (1.0 - mdy);
      sw[0] += w;
#pragma unroll
      for (int k = 0; k < NF; k++) sf[0][k] += w * f[k];
    }
    if (xl && yh) {
      const double w = (1.0 - mdx) * (mdy);
      sw[2] += w;
#pragma unroll
      for (int k = 0; k < NF; k++) sf[2][k] += w * f[k];
    }
    if (xh && yl) {
      const double w = (mdx) * (1.0 - mdy);
      sw[1] += w;
#pragma unroll
      for (int k = 0; k < NF; k++) sf[1][k] += w * f[k];
    }
    if (xh && yh) {
      const double w = (mdx) * (mdy);
      sw[3] += w;
#pragma unroll
      for (int k = 0; k < NF; k++) sf[3][k] += w * f[k];
    }
  } else {  // PM_CENTRE: node [cellY+1][cellX+1]
    const double w = (1.0 - fabs(0.5 - mdx)) * (1.0 - fabs(0.5 - mdy));
    sw[3] += w;
#pragma unroll
    for (int k = 0; k < NF; k++) sf[3][k] += w * f[k];
  }
}

// Pass 1 for one lattice cell (lx, ly) = flat index t, handled by a group of G
// lanes (lane j; dead groups, live == false, still join the shuffles).  Writes
// the corner partial sums to part[v * nL + t], v = corner * (NF + 1) + 0 for
// the weight, + 1 + k for field k.  `start` addresses this grid's bucket
// starts; `perm` holds indices into this instance's marker columns.
template <int METHOD, int NF, int G>
__device__ __forceinline__ void partials_cell(const DGrid &g, const ProjGroup &pg, const int32_t *__restrict__ start,
                                              const int32_t *__restrict__ perm, const double *__restrict__ MX,
                                              const double *__restrict__ MY, const int32_t *__restrict__ Mat,
                                              double *__restrict__ part, int64_t nL, int64_t t, int lx, int ly, int j,
                                              bool live) {
  const int sx = pg.stagx, sy = pg.stagy;
  bool any_mat = false;
#pragma unroll
  for (int k = 0; k < NF; k++) any_mat |= pg.src[k].use_mat != 0;

  // Bucket ranges making up this lattice cell (at most four), in key order.
  // Along each axis: stag 0 -> basic cell l, both halves; stag -1 -> (l-1,
  // upper half) then (l, lower half); stag +1 -> (l, upper half) then (l+1,
  // lower half).  Range r is computed with r a compile-time constant, so the
  // four (start, length) pairs live in registers.
  int rs[4], rl[4];
#pragma unroll
  for (int r = 0; r < 4; r++) {
    int cx = lx, cy = ly, qlo = 0, qhi = 4;
    bool ok;
    if (sx == 0 && sy == 0) {
      ok = r == 0;
    } else if (sx == 0) {  // two cells along y, each one half (two quadrants)
      ok = r < 2;
      cy = sy == -1 ? ly - 1 + r : ly + r;
      qlo = 2 * (1 - r);
      qhi = qlo + 2;
    } else if (sy == 0) {  // two cells along x, each one half in x (quadrants q, q+2)
      const int jx = r >> 1;
      ok = true;
      cx = sx == -1 ? lx - 1 + jx : lx + jx;
      qlo = (1 - jx) + 2 * (r & 1);
      qhi = qlo + 1;
    } else {  // four cells, one quadrant each
      const int jy = r >> 1, jx = r & 1;
      ok = true;
      cy = sy == -1 ? ly - 1 + jy : ly + jy;
      cx = sx == -1 ? lx - 1 + jx : lx + jx;
      qlo = (1 - jx) + 2 * (1 - jy);
      qhi = qlo + 1;
    }
    ok = ok && live && cx >= 0 && cx <= g.NX - 2 && cy >= 0 && cy <= g.NY - 2;
    const int kb = (cy * (g.NX - 1) + cx) << 2;
    const int a0 = ok ? start[kb + qlo] : 0;
    const int a1 = ok ? start[kb + qhi] : 0;
    rs[r] = a0;
    rl[r] = a1 - a0;
  }
  const int p1 = rl[0], p2 = p1 + rl[1], p3 = p2 + rl[2], tot = p3 + rl[3];

  double sw[4], sf[4][NF];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    sw[c] = 0.0;
#pragma unroll
    for (int k = 0; k < NF; k++) sf[c][k] = 0.0;
  }
  // lane j takes entries j, j+16, ... of the concatenated ranges, in order
  for (int e = j; e < tot; e += G) {
    const int idx = e < p1 ? rs[0] + e : (e < p2 ? rs[1] + (e - p1) : (e < p3 ? rs[2] + (e - p2) : rs[3] + (e - p3)));
    const int m = perm[idx];
    const double mdx = frac(sx, g.x, g.xc, lx, MX[m]);
    const double mdy = frac(sy, g.y, g.yc, ly, MY[m]);
    double f[NF];
    const int mat = any_mat ? Mat[m] : 0;
#pragma unroll
    for (int k = 0; k < NF; k++)
      f[k] = pg.src[k].use_mat ? (((unsigned)mat < FVM_MAXMAT) ? pg.src[k].mat[mat] : __builtin_nan(""))
                               : pg.src[k].col[m];
    accumulate<METHOD, NF>(mdx, mdy, f, sw, sf);
  }
  // fixed butterfly over the 16 lanes of the group (a + b == b + a exactly,
  // so every lane ends with the same, order-fixed total)
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      sw[c] += __shfl_xor(sw[c], off, G);
#pragma unroll
      for (int k = 0; k < NF; k++) sf[c][k] += __shfl_xor(sf[c][k], off, G);
    }
  }
  if (live && j == 0) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      part[(int64_t)(c * (NF + 1)) * nL + t] = sw[c];
#pragma unroll
      for (int k = 0; k < NF; k++) part[(int64_t)(c * (NF + 1) + 1 + k) * nL + t] = sf[c][k];
    }
  }
}

// ---- pass 2: one thread per node, every group of the call -------------------------
#define MAX_NODE_GROUPS 6
struct NodeGroup {
  const double *part;
  int nf, harm, nLx, nLy, lox, loy, epilogue;
  double *out[FVM_MAXF];
  const double *ea, *eb;
};
struct NodeGroups {
  int ng;
  NodeGroup q[MAX_NODE_GROUPS];
};

// Pass 2 for node (ix, jy) (flat index i) of one group.
__device__ __forceinline__ void node_group(const NodeGroup &q, int ix, int jy, int64_t i) {
  const int64_t nL = (int64_t)q.nLx * q.nLy;
  const int st = q.nf + 1;
  // (lattice cell, corner) pairs touching node (ix, jy), in a fixed order:
  // (ix-1, jy-1) corner 3, (ix, jy-1) corner 2, (ix-1, jy) corner 1, (ix, jy) corner 0
  const int ax0 = ix - 1 - q.lox, ax1 = ix - q.lox, ay0 = jy - 1 - q.loy, ay1 = jy - q.loy;
  const bool vx0 = ax0 >= 0 && ax0 < q.nLx, vx1 = ax1 >= 0 && ax1 < q.nLx;
  const bool vy0 = ay0 >= 0 && ay0 < q.nLy, vy1 = ay1 >= 0 && ay1 < q.nLy;
  const int64_t L3 = (int64_t)ay0 * q.nLx + ax0, L2 = (int64_t)ay0 * q.nLx + ax1;
  const int64_t L1 = (int64_t)ay1 * q.nLx + ax0, L0 = (int64_t)ay1 * q.nLx + ax1;
  auto gather = [&](int v) {  // v = 0: weight, 1 + k: field k
    double s = 0.0;
    if (vx0 && vy0) s += q.part[(int64_t)(3 * st + v) * nL + L3];
    if (vx1 && vy0) s += q.part[(int64_t)(2 * st + v) * nL + L2];
    if (vx0 && vy1) s += q.part[(int64_t)(1 * st + v) * nL + L1];
    if (vx1 && vy1) s += q.part[(int64_t)(0 * st + v) * nL + L0];
    return s;
  };
  // global = 0 + local (DMLocalToGlobal ADD into a zeroed vector), then divide
  const double den = 0.0 + gather(0);
  for (int k = 0; k < q.nf; k++) {
    const double num = 0.0 + gather(1 + k);
    const double r = q.harm ? den / num : num / den;
    double *o = q.out[k];
    if (q.epilogue == EPI_DIVIDE) {
      o[i] = r;
    } else if (q.epilogue == EPI_SUBGRID_T) {
      // VecAXPBYPCZ(z, 1, -1, -1, x=a, y=b): z = x + (-1)*y + (-1)*z
      o[i] = q.ea[i] + -1.0 * q.eb[i] + -1.0 * r;
    } else {
      // VecAXPY(out, -1, r)
      o[i] = o[i] + -1.0 * r;
    }
  }
}
