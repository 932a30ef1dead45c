// Deterministic marker -> node projection (markerProjection.c:71-250 and the
// scatter parts of markers.c:1507-1594, subgridStressChanges.c:128-233).
//
// Node family (stagx, stagy) has its own lattice of cells; a marker belongs to
// lattice cell l = cellX + {0 (stag 0), bx (stag -1), bx-1 (stag +1)} along x
// (findCellStag2D, markerProjection.c:252-269), likewise along y, and it
// contributes to the lattice cell's four corner nodes.
//
// Pass 1 (k_partials): one thread per lattice cell walks the (at most four)
//   buckets that make up the cell, in bucket order and ascending marker index
//   within a bucket, and accumulates per corner the weight sum and the
//   weighted field sums.  No atomics: each partial is owned by one thread.
// Pass 2 (k_nodes): one thread per node adds the corner partials of the (at
//   most four) lattice cells around it in a fixed order and normalises.
// Both passes are pure functions of the bucket order, so the result is
// bitwise reproducible; it differs from the reference's sequential sum only
// by the association of the additions.
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

// ---- pass 1: 16 lanes per lattice cell ------------------------------------------
#define LPG 16

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

template <int METHOD, int NF>
__global__ void __launch_bounds__(256) k_partials(DGrid g, ProjGroup pg, const int32_t *__restrict__ start,
                                                  const int32_t *__restrict__ perm, const double *__restrict__ MX,
                                                  const double *__restrict__ MY, const int32_t *__restrict__ Mat,
                                                  double *__restrict__ part, int nLx, int nLy, int lox, int loy) {
  const int64_t t = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / LPG;
  const int j = threadIdx.x % LPG;
  const int64_t nL = (int64_t)nLx * nLy;
  const bool live = t < nL;  // dead groups still take part in the shuffles
  const int lx = live ? (int)(t % nLx) + lox : 0;
  const int ly = live ? (int)(t / nLx) + loy : 0;
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
  for (int e = j; e < tot; e += LPG) {
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
  for (int off = LPG / 2; off > 0; off >>= 1) {
#pragma unroll
    for (int c = 0; c < 4; c++) {
      sw[c] += __shfl_xor(sw[c], off, LPG);
#pragma unroll
      for (int k = 0; k < NF; k++) sf[c][k] += __shfl_xor(sf[c][k], off, LPG);
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

__global__ void __launch_bounds__(256) k_nodes(int NX, int NY, NodeGroups G) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  for (int gi = 0; gi < G.ng; gi++) {
    const NodeGroup &q = G.q[gi];
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
}

template <int METHOD, int NF>
static int launch_partials(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg,
                           double *part) {
  const Lat Lx = lattice(pg.stagx, g->NX), Ly = lattice(pg.stagy, g->NY);
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  FVM_PHASE(ctx, FVM_PHASE_PARTIALS);
  k_partials<METHOD, NF><<<fvm_blocks(nL * LPG, 256), 256, 0, ctx->stream>>>(
      dgrid(g), pg, (const int32_t *)ctx->start.ptr, (const int32_t *)ctx->perm.ptr, mk->X, mk->Y, mk->Mat, part,
      Lx.n, Ly.n, Lx.lo, Ly.lo);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

template <int METHOD>
static int dispatch_nf(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg, double *part) {
  switch (pg.nf) {
    case 1: return launch_partials<METHOD, 1>(ctx, g, mk, pg, part);
    case 2: return launch_partials<METHOD, 2>(ctx, g, mk, pg, part);
    case 3: return launch_partials<METHOD, 3>(ctx, g, mk, pg, part);
    case 4: return launch_partials<METHOD, 4>(ctx, g, mk, pg, part);
    case 5: return launch_partials<METHOD, 5>(ctx, g, mk, pg, part);
    case 6: return launch_partials<METHOD, 6>(ctx, g, mk, pg, part);
    case 7: return launch_partials<METHOD, 7>(ctx, g, mk, pg, part);
    case 8: return launch_partials<METHOD, 8>(ctx, g, mk, pg, part);
    default: return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "projection group with %d fields", pg.nf);
  }
}

static int dispatch_method(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg,
                           double *part) {
  switch (pg.method) {
    case PM_ARITH: return dispatch_nf<PM_ARITH>(ctx, g, mk, pg, part);
    case PM_LOCAL: return dispatch_nf<PM_LOCAL>(ctx, g, mk, pg, part);
    case PM_HARM: return dispatch_nf<PM_HARM>(ctx, g, mk, pg, part);
    case PM_TIES: return dispatch_nf<PM_TIES>(ctx, g, mk, pg, part);
    case PM_CENTRE: return dispatch_nf<PM_CENTRE>(ctx, g, mk, pg, part);
    default: return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "projection method %d not implemented", pg.method);
  }
}

// Project several groups over one bucket order: one partials launch per group,
// then one node launch for (up to MAX_NODE_GROUPS of) them.
int fvm_project_groups(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, int ng, const ProjGroup *pgs) {
  size_t total = 0;
  size_t off[64];
  if (ng > 64) return fvm_set_err(ctx, FVM_ERR_ARG, "too many projection groups");
  for (int i = 0; i < ng; i++) {
    const ProjGroup &pg = pgs[i];
    if (pg.stagx < -1 || pg.stagx > 1 || pg.stagy < -1 || pg.stagy > 1)
      return fvm_set_err(ctx, FVM_ERR_BAD_STAG, "invalid stagger (%d,%d)", pg.stagx, pg.stagy);
    if (pg.nf < 1 || pg.nf > FVM_MAXF)
      return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "projection group with %d fields", pg.nf);
    const Lat Lx = lattice(pg.stagx, g->NX), Ly = lattice(pg.stagy, g->NY);
    off[i] = total;
    total += (size_t)Lx.n * Ly.n * 4 * (pg.nf + 1);
  }
  FVM_TRY(fvm_sort_markers(ctx, g, mk));
  FVM_TRY(fvm_ws(ctx, ctx->partials, total * sizeof(double)));
  double *part = (double *)ctx->partials.ptr;
  for (int i = 0; i < ng; i++) FVM_TRY(dispatch_method(ctx, g, mk, pgs[i], part + off[i]));
  const int64_t nn = (int64_t)g->NX * g->NY;
  for (int i0 = 0; i0 < ng; i0 += MAX_NODE_GROUPS) {
    NodeGroups G;
    memset(&G, 0, sizeof(G));
    for (int i = i0; i < ng && G.ng < MAX_NODE_GROUPS; i++) {
      const ProjGroup &pg = pgs[i];
      const Lat Lx = lattice(pg.stagx, g->NX), Ly = lattice(pg.stagy, g->NY);
      NodeGroup &q = G.q[G.ng++];
      q.part = part + off[i];
      q.nf = pg.nf;
      q.harm = pg.method == PM_HARM;
      q.nLx = Lx.n;
      q.nLy = Ly.n;
      q.lox = Lx.lo;
      q.loy = Ly.lo;
      q.epilogue = pg.epilogue;
      for (int k = 0; k < pg.nf; k++) q.out[k] = pg.out[k];
      q.ea = pg.ea;
      q.eb = pg.eb;
    }
    FVM_PHASE(ctx, FVM_PHASE_NODES);
    k_nodes<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(g->NX, g->NY, G);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

int fvm_project_group(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg) {
  return fvm_project_groups(ctx, g, mk, 1, &pg);
}

// ---- public projection entry points ------------------------------------------
extern "C" int fvm_project_marker_fields(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, int32_t n,
                                         const fvm_projection *list) {
  if (!ctx || !mk || (n > 0 && !list)) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  // group entries by (stagx, stagy, method) keeping first-appearance order
  bool done[256] = {false};
  if (n > 256) return fvm_set_err(ctx, FVM_ERR_ARG, "at most 256 projections per call");
  for (int i = 0; i < n; i++) {
    const fvm_projection &e = list[i];
    if (e.method != ARITHMETIC && e.method != ARITHMETIC_LOCAL && e.method != HARMONIC)
      return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "projection method %d not implemented", e.method);
    if (!e.nodalField || (!e.markerField && !e.materialProperty))
      return fvm_set_err(ctx, FVM_ERR_ARG, "projection %d: missing field or output", i);
  }
  ProjGroup groups[64];
  int ng = 0;
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    const fvm_projection &a = list[i];
    ProjGroup &pg = groups[ng++];
    memset(&pg, 0, sizeof(pg));
    pg.stagx = a.stagx;
    pg.stagy = a.stagy;
    pg.method = a.method == ARITHMETIC ? PM_ARITH : (a.method == ARITHMETIC_LOCAL ? PM_LOCAL : PM_HARM);
    pg.epilogue = EPI_DIVIDE;
    for (int j = i; j < n; j++) {
      const fvm_projection &b = list[j];
      if (done[j] || b.stagx != a.stagx || b.stagy != a.stagy || b.method != a.method) continue;
      if (pg.nf == FVM_MAXF) break;
      ProjSrc &s = pg.src[pg.nf];
      if (b.materialProperty) {
        s.use_mat = 1;
        for (int k = 0; k < FVM_MAXMAT; k++) s.mat[k] = b.materialProperty[k];
      } else {
        s.col = b.markerField;
      }
      pg.out[pg.nf] = b.nodalField;
      pg.nf++;
      done[j] = true;
    }
    if (ng == 64) break;
  }
  return fvm_project_groups(ctx, g, mk, ng, groups);
}

extern "C" int fvm_project_markers_nodes_from_scalar(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                                     const double *markerField, double *nodalField) {
  if (!ctx || !mk || !markerField || !nodalField) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  ProjGroup pg;
  memset(&pg, 0, sizeof(pg));
  pg.stagx = pg.stagy = 0;
  pg.method = PM_ARITH;
  pg.nf = 1;
  pg.src[0].col = markerField;
  pg.out[0] = nodalField;
  pg.epilogue = EPI_DIVIDE;
  return fvm_project_group(ctx, g, mk, pg);
}

__global__ void k_fill(double *__restrict__ a, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

int fvm_fill(fvm_ctx *ctx, double *a, int64_t n, double v) {
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_fill<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(a, n, v);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

extern "C" int fvm_project_markers_nodes_all2(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk,
                                              const NodalArrays *nf, const Materials *mat) {
  if (!ctx || !mk || !nf || !mat) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int64_t nn = (int64_t)g->NX * g->NY;
  // markerProjection.c:19-58, grouped by node family and method
  fvm_projection list[9] = {
      {mk->rho, nullptr, nf->rho, 0, 0, ARITHMETIC},
      {mk->eta, nullptr, nf->etaN, -1, -1, ARITHMETIC},
      {mk->eta, nullptr, nf->etaS, 0, 0, ARITHMETIC},
      {nullptr, mat->materialkThermal, nf->kThermal, 0, 0, ARITHMETIC},
      {mk->T, nullptr, nf->lastT, 0, 0, ARITHMETIC},
      {nullptr, mat->materialCp, nf->Cp, 0, 0, ARITHMETIC},
      {mk->sxx, nullptr, nf->soxx, -1, -1, ARITHMETIC_LOCAL},
      {mk->syy, nullptr, nf->soyy, -1, -1, ARITHMETIC_LOCAL},
      {mk->sxy, nullptr, nf->soxy, 0, 0, ARITHMETIC_LOCAL},
  };
  FVM_TRY(fvm_project_marker_fields(ctx, g, mk, 9, list));
  FVM_HIP(ctx, hipMemsetAsync(nf->rhodot, 0, (size_t)nn * sizeof(double), ctx->stream));  // :21
  FVM_TRY(fvm_fill(ctx, nf->muS, nn, 1e99));                                               // :36
  FVM_TRY(fvm_fill(ctx, nf->muN, nn, 1e99));                                               // :37
  return FVM_OK;
}
