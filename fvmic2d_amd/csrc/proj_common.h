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
// N - 1 basic cells along the axis: N = NX (NX + 1 along a periodic x) or NY
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

// The fractional position inside lattice cell l (markerProjection.c:127-154) is
// (X - lo) / (hi - lo) with lo, hi = xc[l], xc[l+1] (stag -1), x[l], x[l+1]
// (stag 0) or xc[l+1], xc[l+2] (stag +1); partials_cell_impl loads lo and hi
// once per lattice cell.

// ---- lane-group sums -------------------------------------------------------------
// A DPP move of a double between lanes of the same row (both 32-bit halves).
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Butterfly step s of a G-lane group sum: lanes pair with lane ^ 1, then
// lane ^ 2 (quad permutes), then the mirror lane inside 8 and inside 16
// (DPP row_half_mirror / row_mirror), then lane ^ 16 (cross-row shuffle).
// After step s every lane of a 2^(s+1)-lane block holds the same bits (IEEE
// addition commutes), so each step adds two equal-pattern partial sums and
// the group ends with one order-fixed total in every lane.
template <int S>
__device__ __forceinline__ double group_step(double v) {
  if constexpr (S == 0) return v + dpp_f64<0xB1>(v);       // quad_perm [1,0,3,2]
  else if constexpr (S == 1) return v + dpp_f64<0x4E>(v);  // quad_perm [2,3,0,1]
  else if constexpr (S == 2) return v + dpp_f64<0x141>(v); // row_half_mirror
  else if constexpr (S == 3) return v + dpp_f64<0x140>(v); // row_mirror
  else return v + __shfl_xor(v, 16, 32);
}

// Accumulate one marker's contribution into the four corner sums.
template <int METHOD, int NF>
__device__ __forceinline__ void accumulate(double mdx, double mdy, const double (&f)[NF], double (&sw)[4],
                                           double (&sf)[4][NF]) {
  if (METHOD == PM_ARITH) {
    // the weights keep the reference's products; the sums use fused
    // multiply-adds (the projection is compared to 1e-12, not bitwise)
    const double w0 = (1.0 - mdx) * (1.0 - mdy), w1 = mdx * (1.0 - mdy), w2 = (1.0 - mdx) * (mdy), w3 = mdx * mdy;
    sw[0] += w0;
    sw[1] += w1;
    sw[2] += w2;
    sw[3] += w3;
#pragma unroll
    for (int k = 0; k < NF; k++) {
      sf[0][k] = fma(w0, f[k], sf[0][k]);
      sf[1][k] = fma(w1, f[k], sf[1][k]);
      sf[2][k] = fma(w2, f[k], sf[2][k]);
      sf[3][k] = fma(w3, f[k], sf[3][k]);
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
// subgridTemperatureChanges pass 1 (subgridTemperatureChanges.c:46-70) evaluated
// inside the projection of its result onto the basic nodes (:74): the lattice
// cell of family (0,0) is the marker's own basic cell, so the quotients mdx, mdy
// of the projection are the ones pass 1 interpolates with, and the cell's four
// corner values of lastT, rho, Cp and k are loaded once per lattice cell.
struct SubT {
  const double *T, *rho, *Cp, *k;  // nodal lastT, rho, Cp, kThermal
  const double *MT;                // marker T
  double *dTsub;                   // per-marker result (read again by pass 2)
  double rho0, dt, dsub;           // materialRho[0], dt, subgrid diffusivity
};

// SORTED: MX, MY, Mat and the source columns are payload arrays in bucket
// order (indexed by the position in perm) instead of marker columns.  mtab:
// the group's material tables staged in LDS (stage_mat_tables; unused by SUBT).
template <int METHOD, int NF, int G, bool SUBT, bool SORTED>
__device__ __forceinline__ void partials_cell_impl(const DGrid &g, const ProjGroup &pg,
                                                   const int32_t *__restrict__ start, const int32_t *__restrict__ perm,
                                                   const double *__restrict__ MX, const double *__restrict__ MY,
                                                   const int32_t *__restrict__ Mat, const double *__restrict__ mtab,
                                                   double *__restrict__ part, int64_t nL, int64_t t, int lx, int ly,
                                                   int j, bool live, const SubT &sb) {
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
    ok = ok && live && cx >= 0 && cx < g.ncx() && cy >= 0 && cy <= g.NY - 2;
    const int kb = (cy * g.ncx() + cx) << 2;
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
  // pass-1 inputs at the corners of basic cell (lx, ly) (SUBT only; a cell with
  // markers has lx in [0, ncx-1], ly in [0, NY-2]; the right column lx + 1 is
  // node 0 for the last cell of a periodic grid)
  double T00 = 0, T10 = 0, T01 = 0, T11 = 0, r00 = 0, r10 = 0, r01 = 0, r11 = 0;
  double C00 = 0, C10 = 0, C01 = 0, C11 = 0, k00 = 0, k10 = 0, k01 = 0, k11 = 0, sden = 1.0;
  if (SUBT && tot > 0) {
    const int64_t i0 = (int64_t)ly * g.NX + lx, i1 = i0 + g.NX;
    const int64_t e = (lx + 1 == g.NX) ? 1 - g.NX : 1;  // offset of the right column
    T00 = sb.T[i0], T10 = sb.T[i0 + e], T01 = sb.T[i1], T11 = sb.T[i1 + e];
    r00 = sb.rho[i0], r10 = sb.rho[i0 + e], r01 = sb.rho[i1], r11 = sb.rho[i1 + e];
    C00 = sb.Cp[i0], C10 = sb.Cp[i0 + e], C01 = sb.Cp[i1], C11 = sb.Cp[i1 + e];
    k00 = sb.k[i0], k10 = sb.k[i0 + e], k01 = sb.k[i1], k11 = sb.k[i1 + e];
    const double dx = g.x[lx + 1] - g.x[lx];
    const double dy = g.y[ly + 1] - g.y[ly];
    sden = 2.0 / (dx * dx) + 2.0 / (dy * dy);
  }
  // lattice-cell bounds of frac(), loaded once: mdx = (X - x0) / (x1 - x0)
  double x0 = 0.0, x1 = 1.0, y0 = 0.0, y1 = 1.0;
  if (tot > 0) {
    x0 = sx == -1 ? g.xc[lx] : (sx == 0 ? g.x[lx] : g.xc[lx + 1]);
    x1 = sx == -1 ? g.xc[lx + 1] : (sx == 0 ? g.x[lx + 1] : g.xc[lx + 2]);
    y0 = sy == -1 ? g.yc[ly] : (sy == 0 ? g.y[ly] : g.yc[ly + 1]);
    y1 = sy == -1 ? g.yc[ly + 1] : (sy == 0 ? g.y[ly + 1] : g.yc[ly + 2]);
  }
  const double xden = x1 - x0, yden = y1 - y0;
  // ARITHMETIC weights are continuous in the fractional position, so they take
  // it from one reciprocal per lattice cell (within an ulp of the quotient: the
  // sums stay within rounding); the nearest-corner and tie rules compare it
  // with 0.5 and the dTsub pass follows subgridTemperatureChanges' expression,
  // so those keep the division
  const double xinv = 1.0 / xden, yinv = 1.0 / yden;
  // Lane j takes entries j, j+G, j+2G, ... of the concatenated ranges and adds
  // them in that order.  U entries are fetched together (their loads in
  // flight at once) before they are added; with more than three fields one at
  // a time (98 instead of 112 VGPRs; the 64-instance ensemble's partials pass
  // 2.38 ms instead of 2.44, the 201 x 201 step's 60.5 us instead of 61.8), and
  // two at a time in the dTsub pass (122 instead of 130 VGPRs, four waves per
  // SIMD instead of three: 0.76 ms instead of 0.83 in the ensemble).
  constexpr int U = SUBT ? 2 : (NF > 3 ? 1 : 4);
  constexpr int NV = SUBT ? 1 : NF;  // loaded values per marker
  for (int e0 = j; e0 < tot; e0 += U * G) {
    int mm[U];  // marker index (SORTED: bucket position) of entry e0 + u*G, -1 past the end
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int e = e0 + u * G;
      const int idx =
          e < p1 ? rs[0] + e : (e < p2 ? rs[1] + (e - p1) : (e < p3 ? rs[2] + (e - p2) : rs[3] + (e - p3)));
      mm[u] = e < tot ? (SORTED ? idx : perm[idx]) : -1;
    }
    double xs[U], ys[U], vs[U][NV];
    int mats[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int m = mm[u] >= 0 ? mm[u] : mm[0];  // mm[0] >= 0: e0 < tot
      if constexpr (SUBT) {
        xs[u] = MX[m];
        ys[u] = MY[m];
        vs[u][0] = sb.MT[m];
        mats[u] = 0;
      } else {
        // the material index first: the table reads wait for it alone
        mats[u] = any_mat ? Mat[m] : 0;
#pragma unroll
        for (int k = 0; k < NF; k++) vs[u][k] = pg.src[k].use_mat ? 0.0 : global_ptr(pg.src[k].col)[m];
        xs[u] = MX[m];
        ys[u] = MY[m];
      }
    }
    // every load of the round is issued before the first use of any of them
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int u = 0; u < U; u++) {
      if (u > 0 && mm[u] < 0) break;  // mm[0] >= 0: e0 < tot
      const int m = mm[u];
      const double mdx = METHOD == PM_ARITH && !SUBT ? (xs[u] - x0) * xinv : (xs[u] - x0) / xden;
      const double mdy = METHOD == PM_ARITH && !SUBT ? (ys[u] - y0) * yinv : (ys[u] - y0) / yden;
      double f[NF];
      if constexpr (SUBT) {
        // subgridTemperatureChanges.c:53-67, same operations and order
        const double Tm = T00 * (1.0 - mdx) * (1.0 - mdy) + T10 * mdx * (1.0 - mdy) + T01 * (1.0 - mdx) * (mdy) +
                          T11 * mdx * mdy;
        const double rhom = sb.rho0 + r00 * (1.0 - mdx) * (1.0 - mdy) + r10 * (mdx) * (1.0 - mdy) +
                            r01 * (1.0 - mdx) * (mdy) + r11 * mdx * mdy;
        const double Cpm = C00 * (1.0 - mdx) * (1.0 - mdy) + C10 * (mdx) * (1.0 - mdy) +
                           C01 * (1.0 - mdx) * (mdy) + C11 * mdx * mdy;
        const double km = k00 * (1.0 - mdx) * (1.0 - mdy) + k10 * (mdx) * (1.0 - mdy) + k01 * (1.0 - mdx) * (mdy) +
                          k11 * mdx * mdy;
        const double dTdiff = Cpm * rhom / km / sden;
        const double d = (Tm - vs[u][0]) * (1.0 - exp(-sb.dsub * sb.dt / dTdiff));
        sb.dTsub[SORTED ? perm[m] : m] = d;
#pragma unroll
        for (int k = 0; k < NF; k++) f[k] = d;
      } else {
        // the table read is unconditional (index masked into the table); a
        // material index outside it gives NaN
        const int mat = mats[u];
        const bool mok = (unsigned)mat < FVM_MAXMAT;
#pragma unroll
        for (int k = 0; k < NF; k++) {
          const double tv = mtab[k * FVM_MAXMAT + (mat & (FVM_MAXMAT - 1))];
          f[k] = pg.src[k].use_mat ? (mok ? tv : __builtin_nan("")) : vs[u][k];
        }
      }
      accumulate<METHOD, NF>(mdx, mdy, f, sw, sf);
    }
  }
  // fixed butterfly over the G lanes of the group (group_step)
  static_assert(G == 4 || G == 8 || G == 16 || G == 32, "lane groups of 4, 8, 16 or 32");
#pragma unroll
  for (int c = 0; c < 4; c++) {
    sw[c] = group_step<1>(group_step<0>(sw[c]));
#pragma unroll
    for (int k = 0; k < NF; k++) sf[c][k] = group_step<1>(group_step<0>(sf[c][k]));
    if constexpr (G >= 8) {
      sw[c] = group_step<2>(sw[c]);
#pragma unroll
      for (int k = 0; k < NF; k++) sf[c][k] = group_step<2>(sf[c][k]);
    }
    if constexpr (G >= 16) {
      sw[c] = group_step<3>(sw[c]);
#pragma unroll
      for (int k = 0; k < NF; k++) sf[c][k] = group_step<3>(sf[c][k]);
    }
    if constexpr (G == 32) {
      sw[c] = group_step<4>(sw[c]);
#pragma unroll
      for (int k = 0; k < NF; k++) sf[c][k] = group_step<4>(sf[c][k]);
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

// The group's material tables in LDS, smat[k * FVM_MAXMAT + mat] for its NF
// fields: a marker's lookup is then an LDS read, which does not wait behind
// the marker's own global loads.  Every thread of the block calls it with the
// same group; the first barrier lets a previous table go out of use.
template <int NF>
__device__ __forceinline__ const double *stage_mat_tables(const ProjGroup &pg, double *smat) {
  __syncthreads();
  for (int i = threadIdx.x; i < NF * FVM_MAXMAT; i += blockDim.x)
    smat[i] = pg.src[i / FVM_MAXMAT].mat[i % FVM_MAXMAT];
  __syncthreads();
  return smat;
}

// Pass 1 over the bucket-ordered payload (xs, ys, ms and the columns named in pg
// or sb.MT hold position-ordered copies; perm maps a position to its marker).
template <int METHOD, int NF, int G>
__device__ __forceinline__ void partials_cell_sorted(const DGrid &g, const ProjGroup &pg,
                                                     const int32_t *__restrict__ start,
                                                     const double *__restrict__ xs, const double *__restrict__ ys,
                                                     const int32_t *__restrict__ ms, const double *__restrict__ mtab,
                                                     double *__restrict__ part, int64_t nL, int64_t t, int lx, int ly,
                                                     int j, bool live) {
  const SubT none = {};
  partials_cell_impl<METHOD, NF, G, false, true>(g, pg, start, nullptr, xs, ys, ms, mtab, part, nL, t, lx, ly, j,
                                                 live, none);
}

template <int G>
__device__ __forceinline__ void partials_cell_subt_sorted(const DGrid &g, const ProjGroup &pg,
                                                          const int32_t *__restrict__ start,
                                                          const int32_t *__restrict__ perm,
                                                          const double *__restrict__ xs,
                                                          const double *__restrict__ ys, double *__restrict__ part,
                                                          int64_t nL, int64_t t, int lx, int ly, int j, bool live,
                                                          const SubT &sb) {
  partials_cell_impl<PM_ARITH, 1, G, true, true>(g, pg, start, perm, xs, ys, nullptr, nullptr, part, nL, t, lx, ly,
                                                 j, live, sb);
}

// ---- pass 2: one thread per (node, field) ------------------------------------------
#define MAX_NODE_GROUPS 6
struct NodeGroup {
  const double *part;
  int nf, harm, nLx, nLy, lox, loy, epilogue;
  int pst;  // partials per corner: stored fields + 1 (>= nf + 1)
  double *out[FVM_MAXF];
  const double *ea, *eb;
};
struct NodeGroups {
  int ng, npair;
  NodeGroup q[MAX_NODE_GROUPS];
  // pair p = (group pg[p], field pk[p]): one launch row (blockIdx.y) each
  unsigned char pg[MAX_NODE_GROUPS * FVM_MAXF], pk[MAX_NODE_GROUPS * FVM_MAXF];
};

// The result of one field at node i from its summed weight and value.
__device__ __forceinline__ void node_store(const NodeGroup &q, int k, int64_t i, double den, double num, double ea,
                                           double eb) {
  const double r = q.harm ? den / num : num / den;
  double *o = q.out[k];
  if (q.epilogue == EPI_DIVIDE) o[i] = r;
  // VecAXPBYPCZ(z, 1, -1, -1, x=a, y=b): z = x + (-1)*y + (-1)*z
  else if (q.epilogue == EPI_SUBGRID_T) o[i] = ea + -1.0 * eb + -1.0 * r;
  // VecAXPY(out, -1, r)
  else o[i] = o[i] + -1.0 * r;
}

// Pass 2 for node (ix, jy) (flat index i), every field of one group: the
// (lattice cell, corner) pairs touching the node are (ix-1, jy-1) corner 3,
// (ix, jy-1) corner 2, (ix-1, jy) corner 1, (ix, jy) corner 0, added in that
// order (global = 0 + local: DMLocalToGlobal ADD into a zeroed vector), then
// normalised.  The corner weights are read once.
__device__ __forceinline__ void node_group(const NodeGroup &q, int ix, int jy, int64_t i) {
  const int64_t nL = (int64_t)q.nLx * q.nLy;
  const int st = q.pst;
  const int ax0 = ix - 1 - q.lox, ax1 = ix - q.lox, ay0 = jy - 1 - q.loy, ay1 = jy - q.loy;
  const bool vx0 = ax0 >= 0 && ax0 < q.nLx, vx1 = ax1 >= 0 && ax1 < q.nLx;
  const bool vy0 = ay0 >= 0 && ay0 < q.nLy, vy1 = ay1 >= 0 && ay1 < q.nLy;
  const bool v3 = vx0 && vy0, v2 = vx1 && vy0, v1 = vx0 && vy1, v0 = vx1 && vy1;
  // absent corners read element 0 (a valid address) and are dropped below
  const int64_t L3 = v3 ? (int64_t)ay0 * q.nLx + ax0 : 0, L2 = v2 ? (int64_t)ay0 * q.nLx + ax1 : 0;
  const int64_t L1 = v1 ? (int64_t)ay1 * q.nLx + ax0 : 0, L0 = v0 ? (int64_t)ay1 * q.nLx + ax1 : 0;
  const double *P = q.part;
  const double w3 = P[(int64_t)(3 * st) * nL + L3], w2 = P[(int64_t)(2 * st) * nL + L2];
  const double w1 = P[(int64_t)(1 * st) * nL + L1], w0 = P[L0];
  double den = 0.0;
  if (v3) den += w3;
  if (v2) den += w2;
  if (v1) den += w1;
  if (v0) den += w0;
  den = 0.0 + den;
  double ea = 0.0, eb = 0.0;
  if (q.epilogue == EPI_SUBGRID_T) ea = q.ea[i], eb = q.eb[i];
  for (int k = 0; k < q.nf; k++) {
    const double f3 = P[(int64_t)(3 * st + 1 + k) * nL + L3], f2 = P[(int64_t)(2 * st + 1 + k) * nL + L2];
    const double f1 = P[(int64_t)(1 * st + 1 + k) * nL + L1], f0 = P[(int64_t)(1 + k) * nL + L0];
    double num = 0.0;
    if (v3) num += f3;
    if (v2) num += f2;
    if (v1) num += f1;
    if (v0) num += f0;
    num = 0.0 + num;
    node_store(q, k, i, den, num, ea, eb);
  }
}

// The same on a periodic x axis: node ix also collects the corners that land
// on its ghost images ix + NX and ix - NX (DM_BOUNDARY_PERIODIC adds a ghost
// column into the node it wraps to).  Images in the order ix, ix + NX, ix - NX;
// corners in the order above within each.
__device__ __forceinline__ void node_group_periodic(const NodeGroup &q, int ix, int jy, int64_t i, int NX) {
  const int64_t nL = (int64_t)q.nLx * q.nLy;
  const int st = q.pst;
  const int ay0 = jy - 1 - q.loy, ay1 = jy - q.loy;
  const bool vy0 = ay0 >= 0 && ay0 < q.nLy, vy1 = ay1 >= 0 && ay1 < q.nLy;
  int64_t L[12];
  bool v[12];
  int c[12];
#pragma unroll
  for (int s = 0; s < 3; s++) {
    const int xi = ix + (s == 0 ? 0 : (s == 1 ? NX : -NX));
    const int ax0 = xi - 1 - q.lox, ax1 = xi - q.lox;
    const bool vx0 = ax0 >= 0 && ax0 < q.nLx, vx1 = ax1 >= 0 && ax1 < q.nLx;
    v[4 * s + 0] = vx0 && vy0, c[4 * s + 0] = 3, L[4 * s + 0] = (int64_t)ay0 * q.nLx + ax0;
    v[4 * s + 1] = vx1 && vy0, c[4 * s + 1] = 2, L[4 * s + 1] = (int64_t)ay0 * q.nLx + ax1;
    v[4 * s + 2] = vx0 && vy1, c[4 * s + 2] = 1, L[4 * s + 2] = (int64_t)ay1 * q.nLx + ax0;
    v[4 * s + 3] = vx1 && vy1, c[4 * s + 3] = 0, L[4 * s + 3] = (int64_t)ay1 * q.nLx + ax1;
  }
  const double *P = q.part;
  double den = 0.0;
#pragma unroll
  for (int e = 0; e < 12; e++)
    if (v[e]) den += P[(int64_t)(c[e] * st) * nL + L[e]];
  den = 0.0 + den;
  double ea = 0.0, eb = 0.0;
  if (q.epilogue == EPI_SUBGRID_T) ea = q.ea[i], eb = q.eb[i];
  for (int k = 0; k < q.nf; k++) {
    double num = 0.0;
#pragma unroll
    for (int e = 0; e < 12; e++)
      if (v[e]) num += P[(int64_t)(c[e] * st + 1 + k) * nL + L[e]];
    num = 0.0 + num;
    node_store(q, k, i, den, num, ea, eb);
  }
}
