// The plain initial pressure guess (initialPressureGuess.c:6-184) on the GPU.
//
// The reference assembles a 5-point variable-spacing Poisson problem for the
// cell-centre pressure, div grad p = g . grad rho, and hands it to the
// "energy_" KSP.  Here the operator is applied matrix-free, row for row as
// assembled there (dummy rows for ix == 0 or jy == 0, the pinned cell (3, 3),
// zero-flux sides, dp/dy = rho g on top and bottom), and solved with
// Jacobi-preconditioned BiCGStab in five kernels per iteration whose scalars
// stay on the device.  Dot products reduce in a fixed order (per-block trees,
// then the same ordered sum in every block of the next kernel), so the
// iteration is bitwise reproducible.
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

// ---- BiCGStab, five kernels per iteration -----------------------------------------
// Every vector kernel runs one thread per row in blocks of 256 (block k holds
// rows 256k ..).  Dot products are per-block partial sums (fixed trees) that
// the next kernel reduces: every block of it sums the same partials in the same
// order, so every thread holds the same scalar and the iteration is bitwise
// reproducible.  The stencil coefficients are tabulated once.
#define BCG_BS 256

enum { S_RHO = 0, S_RHO_OLD, S_ALPHA, S_OMEGA, S_NSC };

struct Coef {
  const double *d, *r, *l, *dn, *up;
};
struct BcgVecs {
  double *b, *dinv, *x, *r, *rh, *p, *v, *s, *t, *y, *z;
  double *cd, *cr, *cl, *cdn, *cup;          // stencil coefficients
  double *prv, *pts, *ptt, *prho, *prr;      // per-block partial sums
  double *sc;                                // scalars (S_*)
};

// sum over the block's threads with a fixed tree; every thread gets the total
__device__ __forceinline__ double block_total(double v) {
  __shared__ double sh[BCG_BS / 64];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();  // sh may still be read by a previous call
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int k = 0; k < BCG_BS / 64; k++) t += sh[k];
  return t;
}

// sum of nblk per-block partials, identical in every block
__device__ __forceinline__ double partial_total(const double *__restrict__ part, int nblk) {
  double s = 0.0;
  for (int k = threadIdx.x; k < nblk; k += BCG_BS) s += part[k];
  return block_total(s);
}

__device__ __forceinline__ double coef_apply(const Coef &c, const PGrid &g, const double *__restrict__ x, int64_t i) {
  double s = c.d[i] * x[i];
  const double r = c.r[i], l = c.l[i], dn = c.dn[i], up = c.up[i];
  if (r != 0.0) s += r * x[i + 1];
  if (l != 0.0) s += l * x[i - 1];
  if (dn != 0.0) s += dn * x[i + g.NX];
  if (up != 0.0) s += up * x[i - g.NX];
  return s;
}

// right-hand side, Jacobi scale, coefficients; x = 0, r = rhat = b, p = v = 0;
// partials of (rhat, r) and (r, r)
__global__ void __launch_bounds__(BCG_BS) k_bcg_setup(PGrid g, const double *__restrict__ rho, double gy, double gx,
                                                      BcgVecs V) {
  const int64_t n = (int64_t)g.NX * g.NY;
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  double bb = 0.0;
  if (i < n) {
    const int ix = (int)(i % g.NX), jy = (int)(i / g.NX);
    const Row w = pg_row(g, ix, jy);
    V.cd[i] = w.d, V.cr[i] = w.r, V.cl[i] = w.l, V.cdn[i] = w.dn, V.cup[i] = w.up;
    const double bi = V.b[i];
    V.x[i] = 0.0;  // x0 = 0 (initialPressureGuess.c:148-149), so r0 = b
    V.r[i] = bi;
    V.rh[i] = bi;
    V.p[i] = 0.0;
    V.v[i] = 0.0;
    bb = bi * bi;
  }
  const double tot = block_total(bb);
  if (threadIdx.x == 0) {
    V.prho[blockIdx.x] = tot;
    V.prr[blockIdx.x] = tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    V.sc[S_RHO_OLD] = 1.0;
    V.sc[S_ALPHA] = 1.0;
    V.sc[S_OMEGA] = 1.0;
  }
}

// the same start for a system given as tabulated coefficients: b from the
// caller, Jacobi scale 1/diagonal
__global__ void __launch_bounds__(BCG_BS) k_bcg_setup_sys(int64_t n, const double *__restrict__ b, Coef c,
                                                          BcgVecs V) {
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  double bb = 0.0;
  if (i < n) {
    const double bi = b[i];
    V.b[i] = bi;
    V.dinv[i] = 1.0 / c.d[i];
    V.x[i] = 0.0;
    V.r[i] = bi;
    V.rh[i] = bi;
    V.p[i] = 0.0;
    V.v[i] = 0.0;
    bb = bi * bi;
  }
  const double tot = block_total(bb);
  if (threadIdx.x == 0) {
    V.prho[blockIdx.x] = tot;
    V.prr[blockIdx.x] = tot;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    V.sc[S_RHO_OLD] = 1.0;
    V.sc[S_ALPHA] = 1.0;
    V.sc[S_OMEGA] = 1.0;
  }
}

// rho = (rhat, r); beta; p = r + beta (p - omega v); y = Dinv p
__global__ void __launch_bounds__(BCG_BS) k_bcg_p(int64_t n, int nblk, int first, BcgVecs V) {
  const double rho = partial_total(V.prho, nblk);
  const double omega = V.sc[S_OMEGA];
  // after an exact solve (r = 0) rho_old and omega are 0: restart with
  // beta = 0 instead of forming 0/0, so x stays the solution
  const double rho_old = V.sc[S_RHO_OLD];
  const double beta = (first || rho_old == 0.0 || omega == 0.0) ? 0.0 : (rho / rho_old) * (V.sc[S_ALPHA] / omega);
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  if (i < n) {
    const double pi = V.r[i] + beta * (V.p[i] - omega * V.v[i]);
    V.p[i] = pi;
    V.y[i] = V.dinv[i] * pi;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) V.sc[S_RHO] = rho;
}

// v = A y; partials of (rhat, v)
__global__ void __launch_bounds__(BCG_BS) k_bcg_v(PGrid g, Coef c, BcgVecs V) {
  const int64_t n = (int64_t)g.NX * g.NY;
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  double d = 0.0;
  if (i < n) {
    const double vi = coef_apply(c, g, V.y, i);
    V.v[i] = vi;
    d = V.rh[i] * vi;
  }
  const double tot = block_total(d);
  if (threadIdx.x == 0) V.prv[blockIdx.x] = tot;
}

// alpha = rho / (rhat, v); x += alpha y; s = r - alpha v; z = Dinv s
__global__ void __launch_bounds__(BCG_BS) k_bcg_s(int64_t n, int nblk, BcgVecs V) {
  const double rv = partial_total(V.prv, nblk);
  const double alpha = rv != 0.0 ? V.sc[S_RHO] / rv : 0.0;
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  if (i < n) {
    V.x[i] = V.x[i] + alpha * V.y[i];
    const double si = V.r[i] - alpha * V.v[i];
    V.s[i] = si;
    V.z[i] = V.dinv[i] * si;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) V.sc[S_ALPHA] = alpha;
}

// t = A z; partials of (t, s) and (t, t)
__global__ void __launch_bounds__(BCG_BS) k_bcg_t(PGrid g, Coef c, BcgVecs V) {
  const int64_t n = (int64_t)g.NX * g.NY;
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  double ts = 0.0, tt = 0.0;
  if (i < n) {
    const double ti = coef_apply(c, g, V.z, i);
    V.t[i] = ti;
    ts = ti * V.s[i];
    tt = ti * ti;
  }
  const double a = block_total(ts);
  const double b = block_total(tt);
  if (threadIdx.x == 0) {
    V.pts[blockIdx.x] = a;
    V.ptt[blockIdx.x] = b;
  }
}

// omega = (t, s) / (t, t); x += omega z; r = s - omega t; partials of (rhat, r), (r, r)
__global__ void __launch_bounds__(BCG_BS) k_bcg_x(int64_t n, int nblk, BcgVecs V) {
  const double ts = partial_total(V.pts, nblk);
  const double tt = partial_total(V.ptt, nblk);
  const double omega = tt != 0.0 ? ts / tt : 0.0;
  const int64_t i = (int64_t)blockIdx.x * BCG_BS + threadIdx.x;
  double rr0 = 0.0, rr = 0.0;
  if (i < n) {
    V.x[i] = V.x[i] + omega * V.z[i];
    const double ri = V.s[i] - omega * V.t[i];
    V.r[i] = ri;
    rr0 = V.rh[i] * ri;
    rr = ri * ri;
  }
  const double a = block_total(rr0);
  const double b = block_total(rr);
  if (threadIdx.x == 0) {
    V.prho[blockIdx.x] = a;
    V.prr[blockIdx.x] = b;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    V.sc[S_OMEGA] = omega;
    V.sc[S_RHO_OLD] = V.sc[S_RHO];
  }
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
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "initialPressureGuess"));
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

// workspace: 11 vectors, 5 coefficient arrays, 5 partial arrays, scalars
static int bcg_vecs(fvm_ctx *ctx, int64_t n, BcgVecs *V, int *nblk_out) {
  const int64_t np = (n + 63) / 64 * 64;
  const int nblk = (int)((n + BCG_BS - 1) / BCG_BS);
  const int64_t npart = ((int64_t)nblk + 63) / 64 * 64;
  FVM_TRY(fvm_ws(ctx, ctx->scratch, (16 * np + 5 * npart + 64) * sizeof(double)));
  double *w = (double *)ctx->scratch.ptr;
  double **vecs[] = {&V->b, &V->dinv, &V->x, &V->r, &V->rh, &V->p, &V->v, &V->s, &V->t, &V->y, &V->z,
                     &V->cd, &V->cr, &V->cl, &V->cdn, &V->cup};
  for (int k = 0; k < 16; k++) *vecs[k] = w + (int64_t)k * np;
  double *pw = w + 16 * np;
  V->prv = pw, V->pts = pw + npart, V->ptt = pw + 2 * npart, V->prho = pw + 3 * npart, V->prr = pw + 4 * npart;
  V->sc = pw + 5 * npart;
  *nblk_out = nblk;
  return FVM_OK;
}

// BiCGStab iterations after a setup kernel: the host reads ||r|| from the (r, r)
// partials (summed in block order) every eight iterations
static int bcg_loop(fvm_ctx *ctx, const PGrid &pg, const Coef &c, const BcgVecs &V, int64_t n, int nblk,
                    double rtol, int32_t maxit, int32_t *it_out, double *rel_out) {
  const unsigned nb = (unsigned)nblk;
  std::vector<double> hp((size_t)nblk);
  auto rnorm = [&](double *out) -> int {
    FVM_HIP(ctx, hipMemcpyAsync(hp.data(), V.prr, sizeof(double) * nblk, hipMemcpyDeviceToHost, ctx->stream));
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    double t = 0.0;
    for (int k = 0; k < nblk; k++) t += hp[(size_t)k];
    *out = sqrt(t);
    return FVM_OK;
  };
  double bnorm = 0.0;
  FVM_TRY(rnorm(&bnorm));
  auto iteration = [&](int first) -> int {
    k_bcg_p<<<nb, BCG_BS, 0, ctx->stream>>>(n, nblk, first, V);
    k_bcg_v<<<nb, BCG_BS, 0, ctx->stream>>>(pg, c, V);
    k_bcg_s<<<nb, BCG_BS, 0, ctx->stream>>>(n, nblk, V);
    k_bcg_t<<<nb, BCG_BS, 0, ctx->stream>>>(pg, c, V);
    k_bcg_x<<<nb, BCG_BS, 0, ctx->stream>>>(n, nblk, V);
    FVM_LAUNCH_CHECK(ctx);
    return FVM_OK;
  };
  int it = 0;
  double rel = 0.0;
  if (bnorm > 0.0) {
    rel = 1.0;
    for (;;) {
      if (it % 8 == 0 || it == maxit) {
        double rn = 0.0;
        FVM_TRY(rnorm(&rn));
        rel = rn / bnorm;
        if (!(rel > rtol) || it == maxit) break;  // also stops on NaN
      }
      FVM_TRY(iteration(it == 0 ? 1 : 0));
      it++;
    }
  }
  *it_out = it;
  *rel_out = rel;
  return FVM_OK;
}

extern "C" int fvm_initial_pressure_guess(fvm_ctx *ctx, const fvm_grid *g, const double *rho, double gy, double gx,
                                          double rtol, int32_t maxit, double *S, double *p_out, int32_t *iters,
                                          double *relres) {
  if (!ctx || !rho || (!S && !p_out) || !(rtol > 0.0) || maxit < 1) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "initialPressureGuess"));
  if (g->NX < 5 || g->NY < 5) return fvm_set_err(ctx, FVM_ERR_ARG, "pressure guess needs NX, NY >= 5 (pinned cell (3,3))");
  const int64_t n = (int64_t)g->NX * g->NY;
  BcgVecs V;
  int nblk = 0;
  FVM_TRY(bcg_vecs(ctx, n, &V, &nblk));
  const Coef c = {V.cd, V.cr, V.cl, V.cdn, V.cup};
  const PGrid pg = pgrid(g);
  const unsigned nb = (unsigned)nblk;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_pg_rhs<<<nb, BCG_BS, 0, ctx->stream>>>(pg, rho, gy, gx, V.b, V.dinv);
  k_bcg_setup<<<nb, BCG_BS, 0, ctx->stream>>>(pg, rho, gy, gx, V);
  FVM_LAUNCH_CHECK(ctx);
  int32_t it = 0;
  double rel = 0.0;
  FVM_TRY(bcg_loop(ctx, pg, c, V, n, nblk, rtol, maxit, &it, &rel));
  k_store_p<<<nb, BCG_BS, 0, ctx->stream>>>(n, V.x, S, p_out);
  FVM_LAUNCH_CHECK(ctx);
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (!(rel <= rtol)) return fvm_set_err(ctx, FVM_ERR_NOT_CONVERGED, "pressure guess: relative residual %.3e after %d iterations", rel, it);
  return FVM_OK;
}

// Jacobi BiCGStab on a 5-point system over the NX*NY nodes (fvm_internal.h)
int fvm_bicgstab_5pt(fvm_ctx *ctx, int NX, int NY, const double *d, const double *e, const double *w,
                     const double *sth, const double *nth, const double *b, double *x, double rtol, int32_t maxit,
                     int32_t *iters, double *relres) {
  const int64_t n = (int64_t)NX * NY;
  BcgVecs V;
  int nblk = 0;
  FVM_TRY(bcg_vecs(ctx, n, &V, &nblk));
  const Coef c = {d, e, w, sth, nth};
  const PGrid pg = PGrid{NX, NY, nullptr, nullptr, nullptr, nullptr};
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_bcg_setup_sys<<<(unsigned)nblk, BCG_BS, 0, ctx->stream>>>(n, b, c, V);
  FVM_LAUNCH_CHECK(ctx);
  int32_t it = 0;
  double rel = 0.0;
  FVM_TRY(bcg_loop(ctx, pg, c, V, n, nblk, rtol, maxit, &it, &rel));
  FVM_HIP(ctx, hipMemcpyAsync(x, V.x, sizeof(double) * n, hipMemcpyDeviceToDevice, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (!(rel <= rtol)) return fvm_set_err(ctx, FVM_ERR_NOT_CONVERGED, "5-point solve: relative residual %.3e after %d iterations", rel, it);
  return FVM_OK;
}
