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

template <int METHOD, int NF>
__global__ void __launch_bounds__(256) k_partials(DGrid g, ProjGroup pg, const int32_t *__restrict__ start,
                                                  const int32_t *__restrict__ perm, const double *__restrict__ MX,
                                                  const double *__restrict__ MY, const int32_t *__restrict__ Mat,
                                                  double *__restrict__ part, int nLx, int nLy, int lox, int loy) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nL = (int64_t)nLx * nLy;
  if (t >= nL) return;
  const int lx = (int)(t % nLx) + lox;
  const int ly = (int)(t / nLx) + loy;
  const int sx = pg.stagx, sy = pg.stagy;
  bool any_mat = false;
#pragma unroll
  for (int k = 0; k < NF; k++) any_mat |= pg.src[k].use_mat != 0;

  double sw[4], sf[4][NF];
#pragma unroll
  for (int c = 0; c < 4; c++) {
    sw[c] = 0.0;
#pragma unroll
    for (int k = 0; k < NF; k++) sf[c][k] = 0.0;
  }

  // Buckets making up this lattice cell.  Along each axis: stag 0 -> basic
  // cell l, both halves; stag -1 -> (l-1, upper half), (l, lower half);
  // stag +1 -> (l, upper half), (l+1, lower half).  j = 0, 1 enumerates them
  // in ascending cell order; half = -1 means both.
  const int nby = sy == 0 ? 1 : 2, nbx = sx == 0 ? 1 : 2;
  for (int jy = 0; jy < nby; jy++) {
    const int cy = sy == 0 ? ly : (sy == -1 ? ly - 1 + jy : ly + jy);
    const int hy = sy == 0 ? -1 : 1 - jy;
    if (cy < 0 || cy > g.NY - 2) continue;
    for (int jx = 0; jx < nbx; jx++) {
      const int cx = sx == 0 ? lx : (sx == -1 ? lx - 1 + jx : lx + jx);
      const int hx = sx == 0 ? -1 : 1 - jx;
      if (cx < 0 || cx > g.NX - 2) continue;
      const int kb = (cy * (g.NX - 1) + cx) << 2;
      // quadrants q = bx | by<<1 in ascending order
      for (int q = 0; q < 4; q++) {
        if (hx >= 0 && hx != (q & 1)) continue;
        if (hy >= 0 && hy != (q >> 1)) continue;
        const int s0 = start[kb | q], s1 = start[(kb | q) + 1];
        for (int s = s0; s < s1; s++) {
          const int m = perm[s];
          const double mdx = frac(sx, g.x, g.xc, lx, MX[m]);
          const double mdy = frac(sy, g.y, g.yc, ly, MY[m]);
          double f[NF];
          const int mat = any_mat ? Mat[m] : 0;
#pragma unroll
          for (int k = 0; k < NF; k++) {
            if (pg.src[k].use_mat) {
              f[k] = ((unsigned)mat < FVM_MAXMAT) ? pg.src[k].mat[mat] : __builtin_nan("");
            } else {
              f[k] = pg.src[k].col[m];
            }
          }
          if (METHOD == PM_ARITH) {
            const double w[4] = {(1.0 - mdx) * (1.0 - mdy), mdx * (1.0 - mdy), (1.0 - mdx) * (mdy), mdx * mdy};
#pragma unroll
            for (int c = 0; c < 4; c++) {
              sw[c] += w[c];
#pragma unroll
              for (int k = 0; k < NF; k++) sf[c][k] += w[c] * f[k];
            }
          } else if (METHOD == PM_LOCAL || METHOD == PM_HARM) {
            int c = 0;
            double w = 1.0;
            if (mdx >= 0.5) {
              c |= 1;
              w *= mdx;
            } else {
              w *= (1.0 - mdx);
            }
            if (mdy >= 0.5) {
              c |= 2;
              w *= mdy;
            } else {
              w *= (1.0 - mdy);
            }
            // unrolled corner select keeps the accumulators in registers
#pragma unroll
            for (int cc = 0; cc < 4; cc++) {
              if (cc == c) {
                if (METHOD == PM_LOCAL) {
                  sw[cc] += w;
#pragma unroll
                  for (int k = 0; k < NF; k++) sf[cc][k] += w * f[k];
                } else {
                  sw[cc] += 1.0;
#pragma unroll
                  for (int k = 0; k < NF; k++) sf[cc][k] += 1.0 / f[k];
                }
              }
            }
          } else if (METHOD == PM_TIES) {
            const bool xl = mdx <= 0.5, xh = mdx >= 0.5, yl = mdy <= 0.5, yh = mdy >= 0.5;
            if (xl && yl) {
              const double w = (1.0 - mdx) * (1.0 - mdy);
              sw[0] = sw[0] + w;
#pragma unroll
              for (int k = 0; k < NF; k++) sf[0][k] = sf[0][k] + w * f[k];
            }
            if (xl && yh) {
              const double w = (1.0 - mdx) * (mdy);
              sw[2] = sw[2] + w;
#pragma unroll
              for (int k = 0; k < NF; k++) sf[2][k] = sf[2][k] + w * f[k];
            }
            if (xh && yl) {
              const double w = (mdx) * (1.0 - mdy);
              sw[1] = sw[1] + w;
#pragma unroll
              for (int k = 0; k < NF; k++) sf[1][k] = sf[1][k] + w * f[k];
            }
            if (xh && yh) {
              const double w = (mdx) * (mdy);
              sw[3] = sw[3] + w;
#pragma unroll
              for (int k = 0; k < NF; k++) sf[3][k] = sf[3][k] + w * f[k];
            }
          } else {  // PM_CENTRE: node [cellY+1][cellX+1]
            const double w = (1.0 - fabs(0.5 - mdx)) * (1.0 - fabs(0.5 - mdy));
            sw[3] += w;
#pragma unroll
            for (int k = 0; k < NF; k++) sf[3][k] += w * f[k];
          }
        }
      }
    }
  }
#pragma unroll
  for (int c = 0; c < 4; c++) {
    part[(int64_t)(c * (NF + 1)) * nL + t] = sw[c];
#pragma unroll
    for (int k = 0; k < NF; k++) part[(int64_t)(c * (NF + 1) + 1 + k) * nL + t] = sf[c][k];
  }
}

template <int METHOD, int NF>
__global__ void __launch_bounds__(256) k_nodes(int NX, int NY, ProjGroup pg, const double *__restrict__ part,
                                               int nLx, int nLy, int lox, int loy) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  const int64_t nL = (int64_t)nLx * nLy;
  double sw = 0.0, sf[NF];
#pragma unroll
  for (int k = 0; k < NF; k++) sf[k] = 0.0;
  // (lattice cell, corner) pairs touching node (ix, jy), fixed order
  const int lxs[4] = {ix - 1, ix, ix - 1, ix};
  const int lys[4] = {jy - 1, jy - 1, jy, jy};
  const int cs[4] = {3, 2, 1, 0};
#pragma unroll
  for (int p = 0; p < 4; p++) {
    const int ax = lxs[p] - lox, ay = lys[p] - loy;
    if (ax < 0 || ax >= nLx || ay < 0 || ay >= nLy) continue;
    const int64_t L = (int64_t)ay * nLx + ax;
    const int c = cs[p];
    sw += part[(int64_t)(c * (NF + 1)) * nL + L];
#pragma unroll
    for (int k = 0; k < NF; k++) sf[k] += part[(int64_t)(c * (NF + 1) + 1 + k) * nL + L];
  }
#pragma unroll
  for (int k = 0; k < NF; k++) {
    // global = 0 + local (DMLocalToGlobal ADD into a zeroed vector), then divide
    const double num = 0.0 + sf[k], den = 0.0 + sw;
    const double q = (METHOD == PM_HARM) ? den / num : num / den;
    double *o = pg.out[k];
    if (pg.epilogue == EPI_DIVIDE) {
      o[i] = q;
    } else if (pg.epilogue == EPI_SUBGRID_T) {
      // VecAXPBYPCZ(z, 1, -1, -1, x=a, y=b): z = x + (-1)*y + (-1)*z
      o[i] = pg.ea[i] + -1.0 * pg.eb[i] + -1.0 * q;
    } else {
      // VecAXPY(out, -1, q)
      o[i] = o[i] + -1.0 * q;
    }
  }
}

template <int METHOD, int NF>
static int run_group(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg) {
  const Lat Lx = lattice(pg.stagx, g->NX), Ly = lattice(pg.stagy, g->NY);
  const int64_t nL = (int64_t)Lx.n * Ly.n;
  FVM_TRY(fvm_ws(ctx, ctx->partials, (size_t)nL * 4 * (NF + 1) * sizeof(double)));
  double *part = (double *)ctx->partials.ptr;
  const DGrid dg = dgrid(g);
  {
    FVM_PHASE(ctx, FVM_PHASE_PARTIALS);
    k_partials<METHOD, NF><<<fvm_blocks(nL, 256), 256, 0, ctx->stream>>>(
        dg, pg, (const int32_t *)ctx->start.ptr, (const int32_t *)ctx->perm.ptr, mk->X, mk->Y, mk->Mat, part, Lx.n,
        Ly.n, Lx.lo, Ly.lo);
    FVM_LAUNCH_CHECK(ctx);
  }
  const int64_t nn = (int64_t)g->NX * g->NY;
  {
    FVM_PHASE(ctx, FVM_PHASE_NODES);
    k_nodes<METHOD, NF><<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(g->NX, g->NY, pg, part, Lx.n, Ly.n, Lx.lo,
                                                                       Ly.lo);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

template <int METHOD>
static int dispatch_nf(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg) {
  switch (pg.nf) {
    case 1: return run_group<METHOD, 1>(ctx, g, mk, pg);
    case 2: return run_group<METHOD, 2>(ctx, g, mk, pg);
    case 3: return run_group<METHOD, 3>(ctx, g, mk, pg);
    case 4: return run_group<METHOD, 4>(ctx, g, mk, pg);
    case 5: return run_group<METHOD, 5>(ctx, g, mk, pg);
    case 6: return run_group<METHOD, 6>(ctx, g, mk, pg);
    case 7: return run_group<METHOD, 7>(ctx, g, mk, pg);
    case 8: return run_group<METHOD, 8>(ctx, g, mk, pg);
    default: return fvm_set_err(ctx, FVM_ERR_TOO_MANY_FIELDS, "projection group with %d fields", pg.nf);
  }
}

int fvm_project_group(fvm_ctx *ctx, const fvm_grid *g, const fvm_markers *mk, const ProjGroup &pg) {
  if (pg.stagx < -1 || pg.stagx > 1 || pg.stagy < -1 || pg.stagy > 1)
    return fvm_set_err(ctx, FVM_ERR_BAD_STAG, "invalid stagger (%d,%d)", pg.stagx, pg.stagy);
  FVM_TRY(fvm_sort_markers(ctx, g, mk));
  switch (pg.method) {
    case PM_ARITH: return dispatch_nf<PM_ARITH>(ctx, g, mk, pg);
    case PM_LOCAL: return dispatch_nf<PM_LOCAL>(ctx, g, mk, pg);
    case PM_HARM: return dispatch_nf<PM_HARM>(ctx, g, mk, pg);
    case PM_TIES: return dispatch_nf<PM_TIES>(ctx, g, mk, pg);
    case PM_CENTRE: return dispatch_nf<PM_CENTRE>(ctx, g, mk, pg);
    default: return fvm_set_err(ctx, FVM_ERR_BAD_METHOD, "projection method %d not implemented", pg.method);
  }
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
  for (int i = 0; i < n; i++) {
    if (done[i]) continue;
    const fvm_projection &a = list[i];
    ProjGroup pg;
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
    FVM_TRY(fvm_project_group(ctx, g, mk, pg));
  }
  return FVM_OK;
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
