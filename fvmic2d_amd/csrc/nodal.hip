// The nodal step between the host Stokes solve and the marker updates, on the
// device: retrieveSolutions (the solution vector to vx, vy, p), nodalStressStrain
// (strain rates, visco-elastic nodal stress update, shear heating, second
// invariants) and limitDisplacementTimestep (the displacement time step).
// Every nodal array is a dense NX*NY row-major array ([jy][ix], ix fastest);
// one thread per node, consecutive threads along ix.  Reference expression order,
// compiled without contraction.
#include <float.h>
#include <math.h>

#include <mutex>

#include "fvm_internal.h"

namespace {

#define N2(a, jy, ix) ((a)[(int64_t)(jy) * NX + (ix)])

// VecCopy/VecAXPY/VecPointwiseDivide/VecScale/VecShift/VecPointwiseMult/VecAYPX
// sequence of nodalStressStrain.c:93-114 (and :257-271 for the shear nodes),
// element by element
__device__ __forceinline__ void vep_node(double eta, double mu, double dt, double &so, double e, double &ds) {
  double X = eta + dt * mu;
  X = eta / X;
  double X1 = X * -2.0;
  X1 = X1 + 2.0;
  X1 = eta * X1;
  const double et = X1 * e;
  double st = X * so;
  st = st + 1.0 * et;
  ds = st - so;
  so = st;
}

struct NodalPtrs {
  const double *vx, *vy, *etaN, *muN, *etaS, *muS;
  double *exx, *eyy, *exy, *wxy, *soxx, *soyy, *soxy, *dsxx, *dsyy, *dsxy;
};

// nodalStressStrain.c:54-121 (cell-centre strain rates and normal stresses) and
// :157-271 (vorticity, shear strain rate, shear stress) for one node
__global__ void __launch_bounds__(256) k_nodal_stress(DGrid g, NodalPtrs F, double dt, int leftType,
                                                      double leftValue, int topType, double topValue) {
  const int NX = g.NX, NY = g.NY;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  double exx = 0.0, eyy = 0.0;
  if (ix > 0 && jy > 0) {
    const double dx = g.x[ix] - g.x[ix - 1];
    const double dy = g.y[jy] - g.y[jy - 1];
    const double dvx = N2(F.vx, jy - 1, ix) - N2(F.vx, jy - 1, ix - 1);
    const double dvy = N2(F.vy, jy, ix - 1) - N2(F.vy, jy - 1, ix - 1);
    const double divv = dvx / dx + dvy / dy;
    exx = dvx / dx - divv / 3.0;
    eyy = dvy / dy - divv / 3.0;
  }
  F.exx[i] = exx;
  F.eyy[i] = eyy;
  const double etaN = F.etaN[i], muN = F.muN[i];
  double s = F.soxx[i], ds;
  vep_node(etaN, muN, dt, s, exx, ds);
  F.soxx[i] = s;
  F.dsxx[i] = ds;
  s = F.soyy[i];
  vep_node(etaN, muN, dt, s, eyy, ds);
  F.soyy[i] = s;
  F.dsyy[i] = ds;

  const double vyp = F.vy[i], vxp = F.vx[i];
  double vym, vxm;
  if (ix == 0)
    vym = leftType == 0 ? 2.0 * leftValue - vyp : vyp;
  else
    vym = F.vy[i - 1];
  if (jy == 0)
    vxm = topType == 0 ? 2.0 * topValue - vxp : vxp;
  else
    vxm = F.vx[i - NX];
  const double dx = g.xc[ix + 1] - g.xc[ix];
  const double dy = g.yc[jy + 1] - g.yc[jy];
  F.wxy[i] = 0.5 * ((vxp - vxm) / dy - (vyp - vym) / dx);
  const double exy = 0.5 * ((vxp - vxm) / dy + (vyp - vym) / dx);
  F.exy[i] = exy;
  s = F.soxy[i];
  vep_node(F.etaS[i], F.muS[i], dt, s, exy, ds);
  F.soxy[i] = s;
  F.dsxy[i] = ds;
}

// sum over the 2x2 cell-centre nodes (jy..jy+1, ix..ix+1) of a*a/b, in the
// reference's term order (nodalStressStrain.c:373-374)
__device__ __forceinline__ double quad_ratio(const double *a, const double *b, int NX, int jy, int ix) {
  return N2(a, jy, ix) * N2(a, jy, ix) / N2(b, jy, ix) + N2(a, jy + 1, ix) * N2(a, jy + 1, ix) / N2(b, jy + 1, ix) +
         N2(a, jy, ix + 1) * N2(a, jy, ix + 1) / N2(b, jy, ix + 1) +
         N2(a, jy + 1, ix + 1) * N2(a, jy + 1, ix + 1) / N2(b, jy + 1, ix + 1);
}
__device__ __forceinline__ double quad_sq(const double *a, int NX, int jy, int ix) {
  return N2(a, jy, ix) * N2(a, jy, ix) + N2(a, jy + 1, ix) * N2(a, jy + 1, ix) +
         N2(a, jy, ix + 1) * N2(a, jy, ix + 1) + N2(a, jy + 1, ix + 1) * N2(a, jy + 1, ix + 1);
}

// nodalStressStrain.c:355-394: shear heating and the second invariants on the
// interior basic nodes; 0 on the boundary nodes
__global__ void __launch_bounds__(256) k_nodal_heat(int NX, int NY, const double *__restrict__ soxx,
                                                    const double *__restrict__ soyy, const double *__restrict__ soxy,
                                                    const double *__restrict__ etaN, const double *__restrict__ etaS,
                                                    const double *__restrict__ exx, const double *__restrict__ eyy,
                                                    const double *__restrict__ exy, double *__restrict__ heat,
                                                    double *__restrict__ eii, double *__restrict__ sii) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  double h = 0.0, e = 0.0, s = 0.0;
  if (ix > 0 && ix < NX - 1 && jy > 0 && jy < NY - 1) {
    double hs = quad_ratio(soxx, etaN, NX, jy, ix) / 8.0;
    hs = hs + quad_ratio(soyy, etaN, NX, jy, ix) / 8.0;
    hs = hs + soxy[i] * soxy[i] / etaS[i];
    h = 0.0 + hs;
    e = sqrt(quad_sq(exx, NX, jy, ix) / 8.0 + quad_sq(eyy, NX, jy, ix) / 8.0 + exy[i] * exy[i]);
    s = sqrt(quad_sq(soxx, NX, jy, ix) / 8.0 + quad_sq(soyy, NX, jy, ix) / 8.0 + soxy[i] * soxy[i]);
  }
  heat[i] = h;
  eii[i] = e;
  sii[i] = s;
}

// retrieveSolutions.c:10-36: scale the pressure entries of S in place, copy the
// three components out
__global__ void __launch_bounds__(256) k_retrieve(int64_t nn, double *__restrict__ S, double Kcont,
                                                  double *__restrict__ vx, double *__restrict__ vy,
                                                  double *__restrict__ p) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const double sp = S[3 * i + 2] * Kcont;
  S[3 * i + 2] = sp;
  p[i] = sp;
  vx[i] = S[3 * i + 0];
  vy[i] = S[3 * i + 1];
}

// retrieveSolutions.c:49-61: p_mean = sum over ix = 1..NX-1 of p[1][ix]*dx/LX,
// added in index order by one thread (the terms are formed in parallel)
#define PMEAN_CHUNK 2048
__global__ void __launch_bounds__(256) k_pmean(DGrid g, const double *__restrict__ p, double *__restrict__ out) {
  __shared__ double t[PMEAN_CHUNK];
  const int NX = g.NX;
  double acc = 0.0;
  for (int c0 = 1; c0 < NX; c0 += PMEAN_CHUNK) {
    const int cn = min(PMEAN_CHUNK, NX - c0);
    for (int k = threadIdx.x; k < cn; k += blockDim.x) {
      const int ix = c0 + k;
      const double dx = g.x[ix] - g.x[ix - 1];
      t[k] = N2(p, 1, ix) * dx / g.LX;
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int k = 0; k < cn; k++) acc += t[k];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = acc;
}

// retrieveSolutions.c:65-75: remove the mean from every node with jy > 0, ix > 0
__global__ void __launch_bounds__(256) k_pshift(int NX, int NY, double *__restrict__ p,
                                                const double *__restrict__ pm) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)NX * NY) return;
  const int ix = (int)(i % NX), jy = (int)(i / NX);
  if (ix > 0 && jy > 0) p[i] = p[i] - *pm;
}

// limitTimestep.c:39-107: every node offers up to three candidate steps (vx,
// vy, 1% density change); NaN candidates are never taken (the reference's
// `if (dt > dt1)` is false for them), so the running minimum of the reference
// equals min(initial step, every non-NaN candidate), which is order-free and
// exact.  Each block also keeps the extreme non-NaN rhodot.
#define DT_BLOCKS 1024
struct DtPart {
  double cmin, rmax, rmin;
};
__device__ __forceinline__ void take_min(double &m, double c) {
  if (m > c) m = c;
}
__global__ void __launch_bounds__(256) k_dt_partial(DGrid g, const double *__restrict__ vx,
                                                    const double *__restrict__ vy, const double *__restrict__ rho,
                                                    const double *__restrict__ rhodot, double dispFactor,
                                                    DtPart *__restrict__ part) {
  __shared__ DtPart s_p[4];
  const int NX = g.NX, NY = g.NY;
  double cm = INFINITY, rmax = -INFINITY, rmin = INFINITY;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (int64_t)NX * NY;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int ix = (int)(i % NX), jy = (int)(i / NX);
    double dx, dy;
    if (ix == 0) {
      dx = g.x[ix + 1] - g.x[ix];
    } else if (ix == NX - 1) {
      dx = g.x[ix] - g.x[ix - 1];
    } else {
      const double dx1 = g.x[ix + 1] - g.x[ix], dx2 = g.x[ix] - g.x[ix - 1];
      dx = dx1 <= dx2 ? dx1 : dx2;
    }
    if (jy == 0) {
      dy = g.y[jy + 1] - g.y[jy];
    } else if (jy == NY - 1) {
      dy = g.y[jy] - g.y[jy - 1];
    } else {
      const double dy1 = g.y[jy + 1] - g.y[jy], dy2 = g.y[jy] - g.y[jy - 1];
      dy = dy1 <= dy2 ? dy1 : dy2;
    }
    const double u = vx[i], v = vy[i], rd = rhodot[i];
    if (u != 0.0) take_min(cm, dx / (fabs(u)) * dispFactor);
    if (v != 0.0) take_min(cm, dy / (fabs(v)) * dispFactor);
    if (rd != 0.0) take_min(cm, fabs(0.01 * rho[i] / rd));
    if (rd > rmax) rmax = rd;
    if (rd < rmin) rmin = rd;
  }
  for (int off = 32; off > 0; off >>= 1) {
    take_min(cm, __shfl_xor(cm, off));
    const double a = __shfl_xor(rmax, off), b = __shfl_xor(rmin, off);
    if (a > rmax) rmax = a;
    if (b < rmin) rmin = b;
  }
  if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = DtPart{cm, rmax, rmin};
  __syncthreads();
  if (threadIdx.x == 0) {
    DtPart r = s_p[0];
    for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
      take_min(r.cmin, s_p[k].cmin);
      if (s_p[k].rmax > r.rmax) r.rmax = s_p[k].rmax;
      if (s_p[k].rmin < r.rmin) r.rmin = s_p[k].rmin;
    }
    part[blockIdx.x] = r;
  }
}

// limitTimestep.c:112-121 on the combined partials.  VecMax/VecMin start from
// the first entry and take strictly larger/smaller ones: a NaN first entry
// stays, later NaNs are skipped.
// The starting step is d0, or *d0p when d0p is given (then out may equal d0p:
// one thread reads it before it writes).
__global__ void __launch_bounds__(256) k_dt_final(int nb, const DtPart *__restrict__ part,
                                                  const double *__restrict__ rhodot, double d0,
                                                  const double *d0p, double *out) {
  // min and max are exact in any order: the block combines the partials, then
  // one thread applies the reference's final steps
  __shared__ DtPart s_p[4];
  DtPart r{INFINITY, -INFINITY, INFINITY};
  for (int k = threadIdx.x; k < nb; k += blockDim.x) {
    take_min(r.cmin, part[k].cmin);
    if (part[k].rmax > r.rmax) r.rmax = part[k].rmax;
    if (part[k].rmin < r.rmin) r.rmin = part[k].rmin;
  }
  for (int off = 32; off > 0; off >>= 1) {
    take_min(r.cmin, __shfl_xor(r.cmin, off));
    const double a = __shfl_xor(r.rmax, off), b = __shfl_xor(r.rmin, off);
    if (a > r.rmax) r.rmax = a;
    if (b < r.rmin) r.rmin = b;
  }
  if ((threadIdx.x & 63) == 0) s_p[threadIdx.x >> 6] = r;
  __syncthreads();
  if (threadIdx.x != 0) return;
  for (int k = 1; k < (int)(blockDim.x >> 6); k++) {
    take_min(r.cmin, s_p[k].cmin);
    if (s_p[k].rmax > r.rmax) r.rmax = s_p[k].rmax;
    if (s_p[k].rmin < r.rmin) r.rmin = s_p[k].rmin;
  }
  double d = d0p ? *d0p : d0;
  take_min(d, r.cmin);
  const double r0 = rhodot[0];
  double rmax = r0 != r0 ? r0 : r.rmax, rmin = r0 != r0 ? r0 : r.rmin;
  rmax = fabs(rmax);
  rmin = fabs(rmin);
  if (rmin > rmax) rmax = rmin;
  if (rmax * d > 0.01) d = 0.01 / rmax;
  *out = d;
}

#undef N2

}  // namespace

extern "C" int fvm_nodal_stress_strain(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                                       const BoundaryValues *bv, double dt, double *nodalHeating) {
  if (!ctx || !nf || !bv || !nodalHeating) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "nodalStressStrain"));
  const double *need[] = {nf->vx, nf->vy, nf->etaN, nf->muN, nf->etaS, nf->muS};
  double *outs[] = {nf->edotxx, nf->edotyy, nf->edotxy, nf->wxy,  nf->soxx, nf->soyy, nf->soxy,
                    nf->dsxx,   nf->dsyy,   nf->dsxy,   nf->eii, nf->sii};
  for (const double *q : need)
    if (!q) return fvm_set_err(ctx, FVM_ERR_ARG, "nodalStressStrain: missing nodal input");
  for (double *q : outs)
    if (!q) return fvm_set_err(ctx, FVM_ERR_ARG, "nodalStressStrain: missing nodal output");
  const int lt = bv->mechBCLeft.type[1], tt = bv->mechBCTop.type[0];
  if ((lt != 0 && lt != 1) || (tt != 0 && tt != 1))
    return fvm_set_err(ctx, FVM_ERR_BC, "nodalStressStrain: left/top velocity ghost needs type 0 or 1");
  const int64_t nn = (int64_t)g->NX * g->NY;
  NodalPtrs F = {nf->vx,     nf->vy,     nf->etaN,   nf->muN, nf->etaS, nf->muS,  nf->edotxx, nf->edotyy,
                 nf->edotxy, nf->wxy,    nf->soxx,   nf->soyy, nf->soxy, nf->dsxx, nf->dsyy,   nf->dsxy};
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_nodal_stress<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(dgrid(g), F, dt, lt, bv->mechBCLeft.value[1], tt,
                                                                 bv->mechBCTop.value[0]);
  FVM_LAUNCH_CHECK(ctx);
  k_nodal_heat<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(g->NX, g->NY, nf->soxx, nf->soyy, nf->soxy, nf->etaN,
                                                               nf->etaS, nf->edotxx, nf->edotyy, nf->edotxy,
                                                               nodalHeating, nf->eii, nf->sii);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_retrieve_solutions(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, double *S,
                                      double Kcont, int32_t *foundnan) {
  if (!ctx || !nf || !S || !nf->vx || !nf->vy || !nf->p) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "retrieveSolutions"));
  const int64_t nn = (int64_t)g->NX * g->NY;
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096));
  double *pm = (double *)((char *)ctx->red.ptr + 512);
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_retrieve<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(nn, S, Kcont, nf->vx, nf->vy, nf->p);
  FVM_LAUNCH_CHECK(ctx);
  if (g->NY > 1) {
    k_pmean<<<1, 256, 0, ctx->stream>>>(dgrid(g), nf->p, pm);
    FVM_LAUNCH_CHECK(ctx);
    k_pshift<<<fvm_blocks(nn, 256), 256, 0, ctx->stream>>>(g->NX, g->NY, nf->p, pm);
    FVM_LAUNCH_CHECK(ctx);
  }
  if (foundnan) *foundnan = 0;  // the reference checks for NaN only when built with TEXTURE
  return FVM_OK;
}

extern "C" int fvm_limit_displacement_timestep(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf,
                                               double *displacementdt, const Options *opt) {
  if (!ctx || !nf || !displacementdt || !opt || !nf->vx || !nf->vy || !nf->rho || !nf->rhodot)
    return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check_nonperiodic(ctx, g, "limitDisplacementTimestep"));
  const int64_t nn = (int64_t)g->NX * g->NY;
  const int nb = (int)std::min<int64_t>(fvm_blocks(nn, 256), DT_BLOCKS);
  FVM_TRY(fvm_ws(ctx, ctx->red, 4096 + (size_t)DT_BLOCKS * sizeof(DtPart)));
  double *dout = (double *)((char *)ctx->red.ptr + 520);
  DtPart *part = (DtPart *)((char *)ctx->red.ptr + 4096);
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_dt_partial<<<nb, 256, 0, ctx->stream>>>(dgrid(g), nf->vx, nf->vy, nf->rho, nf->rhodot,
                                              opt->displacementStepLimit, part);
    FVM_LAUNCH_CHECK(ctx);
    k_dt_final<<<1, 256, 0, ctx->stream>>>(nb, part, nf->rhodot, *displacementdt, nullptr, dout);
    FVM_LAUNCH_CHECK(ctx);
  }
  double *h = (double *)ctx->pinned.ptr + 24;
  FVM_HIP(ctx, hipMemcpyAsync(h, dout, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  *displacementdt = *h;
  return FVM_OK;
}

// ---- limitDisplacementTimestep without a context and without allocating --------------
// The per-block partials of k_dt_partial live in `scratch`, a file-scope device
// array of DT_BLOCKS entries (24 KiB; one copy on every GPU the library runs
// on, held until the process ends).  Host code never passes the array's name:
// its address on each device is looked up once, with that device current, by
// hipGetSymbolAddress and kept in g_scratch_addr.  Any grid size fits, since
// the partial pass runs at most DT_BLOCKS blocks over a grid-stride loop.
// Every call on one GPU shares the array, so calls that overlap on one GPU
// race; each call writes every partial it reads (no clearing is needed).
static __device__ __align__(256) DtPart scratch[DT_BLOCKS];

#define FVM_NOALLOC_MAX_DEVICES 64
static void *g_scratch_addr[FVM_NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int scratch_on_current_device(DtPart **out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return FVM_ERR_HIP;
  if (dev < 0 || dev >= FVM_NOALLOC_MAX_DEVICES) return FVM_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_addr[dev]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FVM_ERR_HIP;
    g_scratch_addr[dev] = p;
  }
  *out = (DtPart *)g_scratch_addr[dev];
  return FVM_OK;
}

extern "C" int fvm_limit_displacement_timestep_noalloc(void *stream, const fvm_grid *g, const NodalArrays *nf,
                                                       const Options *opt, double *dt_device) {
  if (!g || !nf || !opt || !dt_device || !nf->vx || !nf->vy || !nf->rho || !nf->rhodot) return FVM_ERR_ARG;
  if (g->xperiodic) return FVM_ERR_PERIODIC;
  if (g->NX < 2 || g->NY < 2 || !g->x || !g->y) return FVM_ERR_ARG;
  DtPart *part = nullptr;
  FVM_TRY(scratch_on_current_device(&part));
  const hipStream_t s = (hipStream_t)stream;
  const int64_t nn = (int64_t)g->NX * g->NY;
  const int nb = (int)std::min<int64_t>(fvm_blocks(nn, 256), DT_BLOCKS);
  k_dt_partial<<<nb, 256, 0, s>>>(dgrid(g), nf->vx, nf->vy, nf->rho, nf->rhodot, opt->displacementStepLimit, part);
  if (hipGetLastError() != hipSuccess) return FVM_ERR_HIP;
  k_dt_final<<<1, 256, 0, s>>>(nb, part, nf->rhodot, 0.0, dt_device, dt_device);
  if (hipGetLastError() != hipSuccess) return FVM_ERR_HIP;
  return FVM_OK;
}
