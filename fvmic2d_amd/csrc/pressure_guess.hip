// The plain initial pressure guess (initialPressureGuess.c:6-184) on the GPU.
//
// The reference assembles a 5-point variable-spacing Poisson problem for the
// cell-centre pressure, div grad p = g . grad rho, and hands it to the
// "energy_" KSP.  Here the operator is applied matrix-free, row for row as
// assembled there (dummy rows for ix == 0 or jy == 0, the pinned cell (3, 3),
// zero-flux sides, dp/dy = rho g on top and bottom), and solved with
// Jacobi-preconditioned BiCGStab whose scalars stay on the device.  Dot
// products reduce in a fixed order (per-block trees, then one ordered pass),
// so the iteration is bitwise reproducible.
#include "fvm_internal.h"

namespace {

struct PGrid {
  int NX, NY;
  const double *x, *y, *xc, *yc;  // x[0..NX-1], xc[-2..NX+1] (padded)
};

PGrid pgrid(const fvm_grid *g) { return PGrid{g->NX, g->NY, g->x, g->y, g->xc, g->yc}; }

// Row (ix, jy) of the operator: diagonal d and neighbour coefficients
// (right, left, down, up); a missing neighbour has coefficient 0.
struct Row {
  double d, r, l, dn, up;
};

__device__ __forceinline__ Row pg_row(const PGrid &g, int ix, int jy) {
  Row w = {0.0, 0.0, 0.0, 0.0, 0.0};
  const double *x = g.x, *y = g.y, *xc = g.xc, *yc = g.yc;
  if (ix == 0 || jy == 0) {  // :65-71
    const double dx = x[1] - x[0], dy = y[1] - y[0];
    w.d = 4.0 / (dx * dx + dy * dy);
  } else if (ix == 3 && jy == 3) {  // :72-77
    const double dx = x[ix] - x[ix - 1], dy = y[jy] - y[jy - 1];
    w.d = 4.0 / (dx * dx + dy * dy);
  } else {  // :79-103
    const double hx = x[ix] - x[ix - 1], hy = y[jy] - y[jy - 1];
    if (ix != g.NX - 1) w.r = 1.0 / hx / (xc[ix + 1] - xc[ix]);
    w.d = 1.0 / hx * (-1.0 / (xc[ix + 1] - xc[ix]) - 1.0 / (xc[ix] - xc[ix - 1])) +
          1.0 / hy * (-1.0 / (yc[jy + 1] - yc[jy]) - 1.0 / (yc[jy] - yc[jy - 1]));
    if (ix != 1) w.l = 1.0 / hx / (xc[ix] - xc[ix - 1]);
    if (jy != g.NY - 1) w.dn = 1.0 / hy / (yc[jy + 1] - yc[jy]);
    if (jy != 1) w.up = 1.0 / hy / (yc[jy] - yc[jy - 1]);
  }
  // boundary terms added to the diagonal (:110-132)
  if (ix == 1) w.d += 1.0 / (x[ix] - x[ix - 1]) / (xc[ix] - xc[ix - 1]);
  else if (ix == g.NX - 1) w.d += 1.0 / (x[ix] - x[ix - 1]) / (xc[ix + 1] - xc[ix]);
  if (jy == 1) w.d += 1.0 / (y[jy] - y[jy - 1]) / (yc[jy] - yc[jy - 1]);
  else if (jy == g.NY - 1) w.d += 1.0 / (y[jy] - y[jy - 1]) / (yc[jy + 1] - yc[jy]);
  return w;
}

__global__ void __launch_bounds__(256) k_pg_rhs(PGrid g, const double *__restrict__ rho, double gy, double gx,
                                                double *__restrict__ b, double *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)g.NX * g.NY) return;
  const int ix = (int)(i % g.NX), jy = (int)(i / g.NX);
  const int NX = g.NX;
  double v = 0.0;
  if (!(ix == 0 || jy == 0) && !(ix == 3 && jy == 3)) {  // :104-106
    const double r00 = rho[(int64_t)jy * NX + ix], r01 = rho[(int64_t)jy * NX + ix - 1];
    const double r10 = rho[(int64_t)(jy - 1) * NX + ix], r11 = rho[(int64_t)(jy - 1) * NX + ix - 1];
    v = gy * ((r00 + r01) / 2.0 - (r10 + r11) / 2.0) / (g.y[jy] - g.y[jy - 1]) +
        gx * ((r00 + r10) / 2.0 - (r01 + r11) / 2.0) / (g.x[ix] - g.x[ix - 1]);
  }
  if (jy == 1 || jy == g.NY - 1) {
    // :119-132, applied to every row of these lines (also ix == 0); density
    // outside the grid reads 0 (a fresh local vector's ghost)
    auto R = [&](int j, int i2) { return i2 >= 0 ? rho[(int64_t)j * NX + i2] : 0.0; };
    const double rhoc = (R(jy, ix) + R(jy, ix - 1) + R(jy - 1, ix) + R(jy - 1, ix - 1)) / 4.0;
    const double dy = g.yc[jy] - g.yc[jy - 1];
    if (jy == 1) {
      const double c = 1.0 / (g.y[jy] - g.y[jy - 1]) / (g.yc[jy] - g.yc[jy - 1]);
      v += c * rhoc * gy * dy;
    } else {
      const double c = 1.0 / (g.y[jy] - g.y[jy - 1]) / (g.yc[jy + 1] - g.yc[jy]);
      v += -c * rhoc * gy * dy;
    }
  }
  b[i] = v;
  dinv[i] = 1.0 / pg_row(g, ix, jy).d;
}

// y = A x (y may not alias x)
__global__ void __launch_bounds__(256) k_pg_apply(PGrid g, const double *__restrict__ x, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)g.NX * g.NY) return;
  const int ix = (int)(i % g.NX), jy = (int)(i / g.NX);
  const Row w = pg_row(g, ix, jy);
  double s = w.d * x[i];
  if (w.r != 0.0) s += w.r * x[i + 1];
  if (w.l != 0.0) s += w.l * x[i - 1];
  if (w.dn != 0.0) s += w.dn * x[i + g.NX];
  if (w.up != 0.0) s += w.up * x[i - g.NX];
  y[i] = s;
}

// ---- deterministic dot products ---------------------------------------------
#define DOT_BS 256
#define DOT_ITEMS 8

struct DotArgs {
  int nd;
  const double *a[2], *b[2];
};

// partial[blk * nd + k] = sum over block blk of a_k . b_k (k < nd)
__global__ void __launch_bounds__(DOT_BS) k_dots(int64_t n, DotArgs d, double *__restrict__ partial) {
  __shared__ double sh[DOT_BS];
  for (int k = 0; k < d.nd; k++) {
    const double *a = d.a[k], *b = d.b[k];
    double s = 0.0;
    const int64_t base = (int64_t)blockIdx.x * DOT_BS * DOT_ITEMS + threadIdx.x;
#pragma unroll
    for (int it = 0; it < DOT_ITEMS; it++) {
      const int64_t i = base + (int64_t)it * DOT_BS;
      if (i < n) s = fma(a[i], b[i], s);
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = DOT_BS / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) partial[(int64_t)blockIdx.x * d.nd + k] = sh[0];
    __syncthreads();
  }
}

// out[k] = sum over blocks: strided per lane in block order, then a fixed tree
__global__ void __launch_bounds__(DOT_BS) k_dots_final(int nblk, int nd, const double *__restrict__ partial,
                                                       double *__restrict__ out) {
  __shared__ double sh[DOT_BS];
  for (int k = 0; k < nd; k++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += DOT_BS) s += partial[(int64_t)b * nd + k];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int off = DOT_BS / 2; off > 0; off >>= 1) {
      if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = sh[0];
    __syncthreads();
  }
}

// ---- BiCGStab with the scalars on the device ---------------------------------
enum { S_RHO = 0, S_RR, S_RV, S_TS, S_TT, S_BB, S_RHO_OLD, S_ALPHA, S_OMEGA, S_BETA, S_NSC };

__global__ void k_bcg_init(double *sc) {
  if (threadIdx.x == 0) {
    sc[S_RHO_OLD] = 1.0;
    sc[S_ALPHA] = 1.0;
    sc[S_OMEGA] = 1.0;
    sc[S_BETA] = 0.0;
  }
}

__global__ void k_bcg_beta(double *sc, int first) {
  if (threadIdx.x == 0)
    sc[S_BETA] = first ? 0.0 : (sc[S_RHO] / sc[S_RHO_OLD]) * (sc[S_ALPHA] / sc[S_OMEGA]);
}

__global__ void __launch_bounds__(256) k_bcg_p(int64_t n, const double *__restrict__ sc, const double *__restrict__ r,
                                               const double *__restrict__ v, double *__restrict__ p,
                                               const double *__restrict__ dinv, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double beta = sc[S_BETA], omega = sc[S_OMEGA];
  const double pi = r[i] + beta * (p[i] - omega * v[i]);
  p[i] = pi;
  y[i] = dinv[i] * pi;
}

__global__ void k_bcg_alpha(double *sc) {
  if (threadIdx.x == 0) sc[S_ALPHA] = sc[S_RV] != 0.0 ? sc[S_RHO] / sc[S_RV] : 0.0;
}

__global__ void __launch_bounds__(256) k_bcg_s(int64_t n, const double *__restrict__ sc, const double *__restrict__ r,
                                               const double *__restrict__ v, const double *__restrict__ y,
                                               double *__restrict__ x, double *__restrict__ s,
                                               const double *__restrict__ dinv, double *__restrict__ z) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double alpha = sc[S_ALPHA];
  x[i] = x[i] + alpha * y[i];
  const double si = r[i] - alpha * v[i];
  s[i] = si;
  z[i] = dinv[i] * si;
}

__global__ void k_bcg_omega(double *sc) {
  if (threadIdx.x == 0) {
    sc[S_OMEGA] = sc[S_TT] != 0.0 ? sc[S_TS] / sc[S_TT] : 0.0;
    sc[S_RHO_OLD] = sc[S_RHO];
  }
}

__global__ void __launch_bounds__(256) k_bcg_x(int64_t n, const double *__restrict__ sc, const double *__restrict__ z,
                                               const double *__restrict__ s, const double *__restrict__ t,
                                               double *__restrict__ x, double *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double omega = sc[S_OMEGA];
  x[i] = x[i] + omega * z[i];
  r[i] = s[i] - omega * t[i];
}

__global__ void __launch_bounds__(256) k_init_vecs(int64_t n, const double *__restrict__ b, double *__restrict__ x,
                                                   double *__restrict__ r, double *__restrict__ rh,
                                                   double *__restrict__ p, double *__restrict__ v) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  x[i] = 0.0;  // x0 = 0 (initialPressureGuess.c:148-149), so r0 = b
  r[i] = b[i];
  rh[i] = b[i];
  p[i] = 0.0;
  v[i] = 0.0;
}

__global__ void __launch_bounds__(256) k_store_p(int64_t n, const double *__restrict__ p, double *__restrict__ S,
                                                 double *__restrict__ pout) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (S) S[3 * i + 2] = p[i];  // S[jy][ix][DOF_P], DOF_P = 2 (fdcode.h:245)
  if (pout) pout[i] = p[i];
}

}  // namespace

// Apply the pressure-guess operator (for tests and callers that assemble their
// own right-hand sides): y = A x, b (optional) = the right-hand side.
extern "C" int fvm_pressure_guess_apply(fvm_ctx *ctx, const fvm_grid *g, const double *rho, double gy, double gx,
                                        const double *x, double *y, double *b) {
  if (!ctx || !rho || !x || !y) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (g->NX < 5 || g->NY < 5) return fvm_set_err(ctx, FVM_ERR_ARG, "pressure guess needs NX, NY >= 5");
  const int64_t n = (int64_t)g->NX * g->NY;
  const PGrid pg = pgrid(g);
  const unsigned nb = fvm_blocks(n, 256);
  FVM_TRY(fvm_ws(ctx, ctx->staging, (size_t)(2 * n + 64) * sizeof(double)));
  double *bb = (double *)ctx->staging.ptr, *dinv = bb + n;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_pg_rhs<<<nb, 256, 0, ctx->stream>>>(pg, rho, gy, gx, b ? b : bb, dinv);
  k_pg_apply<<<nb, 256, 0, ctx->stream>>>(pg, x, y);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_initial_pressure_guess(fvm_ctx *ctx, const fvm_grid *g, const double *rho, double gy, double gx,
                                          double rtol, int32_t maxit, double *S, double *p_out, int32_t *iters,
                                          double *relres) {
  if (!ctx || !rho || (!S && !p_out) || !(rtol > 0.0) || maxit < 1) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  if (g->NX < 5 || g->NY < 5) return fvm_set_err(ctx, FVM_ERR_ARG, "pressure guess needs NX, NY >= 5 (pinned cell (3,3))");
  const int64_t n = (int64_t)g->NX * g->NY;
  const int64_t np = (n + 63) / 64 * 64;
  const int nblk = (int)((n + DOT_BS * DOT_ITEMS - 1) / (DOT_BS * DOT_ITEMS));
  // b, dinv, x, r, rh, p, v, s, t, y, z; per-block partials (2 per block); scalars
  FVM_TRY(fvm_ws(ctx, ctx->scratch, (11 * np + 2 * (size_t)nblk + 64) * sizeof(double)));
  double *w = (double *)ctx->scratch.ptr;
  double *b = w, *dinv = w + np, *x = w + 2 * np, *r = w + 3 * np, *rh = w + 4 * np, *p = w + 5 * np;
  double *v = w + 6 * np, *s = w + 7 * np, *t = w + 8 * np, *y = w + 9 * np, *z = w + 10 * np;
  double *partial = w + 11 * np;
  double *sc = partial + 2 * (size_t)nblk;
  const PGrid pg = pgrid(g);
  const unsigned nb = fvm_blocks(n, 256);
  auto dots = [&](int nd, const double *a0, const double *b0, const double *a1, const double *b1,
                  double *out) -> int {
    DotArgs d;
    d.nd = nd;
    d.a[0] = a0;
    d.b[0] = b0;
    d.a[1] = a1;
    d.b[1] = b1;
    k_dots<<<nblk, DOT_BS, 0, ctx->stream>>>(n, d, partial);
    FVM_LAUNCH_CHECK(ctx);
    k_dots_final<<<1, DOT_BS, 0, ctx->stream>>>(nblk, nd, partial, out);
    FVM_LAUNCH_CHECK(ctx);
    return FVM_OK;
  };
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_pg_rhs<<<nb, 256, 0, ctx->stream>>>(pg, rho, gy, gx, b, dinv);
  k_init_vecs<<<nb, 256, 0, ctx->stream>>>(n, b, x, r, rh, p, v);
  k_bcg_init<<<1, 64, 0, ctx->stream>>>(sc);
  FVM_LAUNCH_CHECK(ctx);
  FVM_TRY(dots(1, b, b, nullptr, nullptr, sc + S_BB));
  double hs[S_NSC];
  FVM_HIP(ctx, hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double bnorm = sqrt(hs[S_BB]);
  int it = 0;
  double rel = 0.0;
  if (bnorm > 0.0) {
    rel = 1.0;
    for (;;) {
      // (rhat, r) for this iteration and (r, r) of the current residual
      FVM_TRY(dots(2, rh, r, r, r, sc + S_RHO));
      if (it % 8 == 0 || it == maxit) {
        FVM_HIP(ctx, hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
        FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
        rel = sqrt(hs[S_RR]) / bnorm;
        if (!(rel > rtol) || it == maxit) break;  // also stops on NaN
      }
      k_bcg_beta<<<1, 64, 0, ctx->stream>>>(sc, it == 0 ? 1 : 0);
      k_bcg_p<<<nb, 256, 0, ctx->stream>>>(n, sc, r, v, p, dinv, y);
      k_pg_apply<<<nb, 256, 0, ctx->stream>>>(pg, y, v);
      FVM_LAUNCH_CHECK(ctx);
      FVM_TRY(dots(1, rh, v, nullptr, nullptr, sc + S_RV));
      k_bcg_alpha<<<1, 64, 0, ctx->stream>>>(sc);
      k_bcg_s<<<nb, 256, 0, ctx->stream>>>(n, sc, r, v, y, x, s, dinv, z);
      k_pg_apply<<<nb, 256, 0, ctx->stream>>>(pg, z, t);
      FVM_LAUNCH_CHECK(ctx);
      FVM_TRY(dots(2, t, s, t, t, sc + S_TS));
      k_bcg_omega<<<1, 64, 0, ctx->stream>>>(sc);
      k_bcg_x<<<nb, 256, 0, ctx->stream>>>(n, sc, z, s, t, x, r);
      FVM_LAUNCH_CHECK(ctx);
      it++;
    }
  }
  k_store_p<<<nb, 256, 0, ctx->stream>>>(n, x, S, p_out);
  FVM_LAUNCH_CHECK(ctx);
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (!(rel <= rtol)) return fvm_set_err(ctx, FVM_ERR_NOT_CONVERGED, "pressure guess: relative residual %.3e after %d iterations", rel, it);
  return FVM_OK;
}
