// Context, device memory helpers, AoS <-> SoA marker transposes, timers.
#include <stdarg.h>
#include <stdlib.h>

#include "fvm_internal.h"

int fvm_set_err(fvm_ctx *ctx, int code, const char *fmt, ...) {
  if (ctx) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(ctx->err, sizeof(ctx->err), fmt, ap);
    va_end(ap);
  }
  return code;
}

int fvm_ws(fvm_ctx *ctx, fvm_buf &b, size_t bytes) {
  if (bytes <= b.bytes) return FVM_OK;
  if (b.ptr) {
    // the buffer may still be read by enqueued work
    FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
    FVM_HIP(ctx, hipFree(b.ptr));
    b.ptr = nullptr;
    b.bytes = 0;
  }
  size_t want = bytes + bytes / 8 + 4096;
  hipError_t e = hipMalloc(&b.ptr, want);
  if (e != hipSuccess) {
    b.ptr = nullptr;
    return fvm_set_err(ctx, FVM_ERR_NOMEM, "workspace hipMalloc(%zu): %s", want, hipGetErrorString(e));
  }
  b.bytes = want;
  return FVM_OK;
}

extern "C" int fvm_ctx_create(fvm_ctx **out, void *stream) {
  if (!out) return FVM_ERR_ARG;
  fvm_ctx *ctx = new fvm_ctx();
  hipError_t e = hipGetDevice(&ctx->device);
  if (e != hipSuccess) {
    delete ctx;
    return FVM_ERR_HIP;
  }
  if (stream) {
    ctx->stream = (hipStream_t)stream;
  } else {
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
      delete ctx;
      return FVM_ERR_HIP;
    }
    ctx->own_stream = true;
  }
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess) {
    delete ctx;
    return FVM_ERR_HIP;
  }
  for (int i = 0; i < 2 * 1024; i++) {
    if (hipEventCreate(&ctx->prof_ev[i]) != hipSuccess) {
      delete ctx;
      return FVM_ERR_HIP;
    }
  }
  if (hipHostMalloc(&ctx->pinned.ptr, 4096, hipHostMallocDefault) != hipSuccess) {
    delete ctx;
    return FVM_ERR_HIP;
  }
  ctx->pinned.bytes = 4096;
  // reduction scratch: counters and flags at the front, block partials from +4 KiB
  if (fvm_ws(ctx, ctx->red, 64 * 1024) != FVM_OK || hipMemset(ctx->red.ptr, 0, 64 * 1024) != hipSuccess) {
    delete ctx;
    return FVM_ERR_NOMEM;
  }
  *out = ctx;
  return FVM_OK;
}

extern "C" int fvm_ctx_destroy(fvm_ctx *ctx) {
  if (!ctx) return FVM_OK;
  (void)hipStreamSynchronize(ctx->stream);
  fvm_buf *bufs[] = {&ctx->key,      &ctx->slot,     &ctx->perm,    &ctx->count,   &ctx->start,
                     &ctx->fill,     &ctx->blocksums, &ctx->partials, &ctx->scratch, &ctx->staging,
                     &ctx->red,      &ctx->payload,   &ctx->payload_xy, &ctx->xfer,    &ctx->popc,
                     &ctx->popc2,    &ctx->reord};
  for (fvm_buf *b : bufs)
    if (b->ptr) (void)hipFree(b->ptr);
  if (ctx->pinned.ptr) (void)hipHostFree(ctx->pinned.ptr);
  for (int i = 0; i < 2 * 1024; i++)
    if (ctx->prof_ev[i]) (void)hipEventDestroy(ctx->prof_ev[i]);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return FVM_OK;
}

extern "C" const char *fvm_last_error(const fvm_ctx *ctx) { return ctx ? ctx->err : "no context"; }
extern "C" void *fvm_ctx_stream(fvm_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }
extern "C" int fvm_ctx_sync(fvm_ctx *ctx) {
  if (!ctx) return FVM_ERR_ARG;
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}
extern "C" size_t fvm_ctx_workspace_bytes(const fvm_ctx *ctx) {
  if (!ctx) return 0;
  const fvm_buf *bufs[] = {&ctx->key,  &ctx->slot,    &ctx->perm,     &ctx->count,   &ctx->start,
                           &ctx->fill, &ctx->blocksums, &ctx->partials, &ctx->scratch, &ctx->staging,
                           &ctx->red,  &ctx->payload,   &ctx->payload_xy, &ctx->popc, &ctx->popc2,
                           &ctx->reord};
  size_t s = 0;
  for (const fvm_buf *b : bufs) s += b->bytes;
  return s;
}

extern "C" int fvm_ctx_set_jaumann(fvm_ctx *ctx, int32_t jaumann) {
  if (!ctx) return FVM_ERR_ARG;
  if (jaumann != 0 && jaumann != 1) return fvm_set_err(ctx, FVM_ERR_ARG, "jaumann must be 0 or 1");
  ctx->jaumann = jaumann;
  return FVM_OK;
}

extern "C" int fvm_ctx_set_lanes(fvm_ctx *ctx, int32_t lanes) {
  if (!ctx) return FVM_ERR_ARG;
  if (lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32)
    return fvm_set_err(ctx, FVM_ERR_ARG, "lanes must be 4, 8, 16 or 32");
  ctx->lanes = lanes;
  return FVM_OK;
}

extern "C" int fvm_device_count(int *n) {
  if (!n) return FVM_ERR_ARG;
  return hipGetDeviceCount(n) == hipSuccess ? FVM_OK : FVM_ERR_HIP;
}
extern "C" int fvm_malloc(void **p, size_t bytes) {
  if (!p) return FVM_ERR_ARG;
  return hipMalloc(p, bytes ? bytes : 1) == hipSuccess ? FVM_OK : FVM_ERR_NOMEM;
}
extern "C" int fvm_free(void *p) { return hipFree(p) == hipSuccess ? FVM_OK : FVM_ERR_HIP; }
extern "C" int fvm_host_alloc(void **ptr, size_t bytes) {
  if (!ptr) return FVM_ERR_ARG;
  return hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) == hipSuccess ? FVM_OK : FVM_ERR_NOMEM;
}
extern "C" int fvm_host_free(void *ptr) {
  if (ptr) (void)hipHostFree(ptr);
  return FVM_OK;
}

// ---- packed nodal exchange with host-side solves ------------------------------
#define XFER_MAX 32
struct XferTable {
  double *p[XFER_MAX];
};
// array k of `n` doubles <-> words [k*n, (k+1)*n) of the packed buffer
__global__ void __launch_bounds__(256) k_unpack_nodal(int64_t n, int32_t na, const double *__restrict__ packed,
                                                      XferTable t) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < na; k++) t.p[k][i] = packed[k * n + i];
}
__global__ void __launch_bounds__(256) k_pack_nodal(int64_t n, int32_t na, XferTable t, double *__restrict__ packed) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int k = 0; k < na; k++) packed[k * n + i] = t.p[k][i];
}

extern "C" int fvm_nodal_upload(fvm_ctx *ctx, int64_t n, int32_t na, double *const *dev, const double *host) {
  if (!ctx || n < 0 || na < 0 || na > XFER_MAX || (na > 0 && (!dev || !host))) return FVM_ERR_ARG;
  if (n == 0 || na == 0) return FVM_OK;
  XferTable t;
  for (int k = 0; k < na; k++) {
    if (!dev[k]) return fvm_set_err(ctx, FVM_ERR_ARG, "fvm_nodal_upload: array %d is NULL", k);
    t.p[k] = dev[k];
  }
  FVM_TRY(fvm_ws(ctx, ctx->xfer, (size_t)na * n * sizeof(double)));
  FVM_HIP(ctx, hipMemcpyAsync(ctx->xfer.ptr, host, (size_t)na * n * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  k_unpack_nodal<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, na, (const double *)ctx->xfer.ptr, t);
  FVM_LAUNCH_CHECK(ctx);
  // the host buffer may be reused as soon as this returns
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

extern "C" int fvm_nodal_download(fvm_ctx *ctx, int64_t n, int32_t na, const double *const *dev, double *host) {
  if (!ctx || n < 0 || na < 0 || na > XFER_MAX || (na > 0 && (!dev || !host))) return FVM_ERR_ARG;
  if (n == 0 || na == 0) return FVM_OK;
  XferTable t;
  for (int k = 0; k < na; k++) {
    if (!dev[k]) return fvm_set_err(ctx, FVM_ERR_ARG, "fvm_nodal_download: array %d is NULL", k);
    t.p[k] = const_cast<double *>(dev[k]);
  }
  FVM_TRY(fvm_ws(ctx, ctx->xfer, (size_t)na * n * sizeof(double)));
  k_pack_nodal<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>(n, na, t, (double *)ctx->xfer.ptr);
  FVM_LAUNCH_CHECK(ctx);
  FVM_HIP(ctx, hipMemcpyAsync(host, ctx->xfer.ptr, (size_t)na * n * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

extern "C" int fvm_memcpy_htod(fvm_ctx *ctx, void *dst, const void *src, size_t bytes) {
  FVM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}
extern "C" int fvm_memcpy_dtoh(fvm_ctx *ctx, void *dst, const void *src, size_t bytes) {
  FVM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}
extern "C" int fvm_memcpy_dtod(fvm_ctx *ctx, void *dst, const void *src, size_t bytes) {
  FVM_HIP(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  return FVM_OK;
}
extern "C" int fvm_memset(fvm_ctx *ctx, void *dst, int value, size_t bytes) {
  FVM_HIP(ctx, hipMemsetAsync(dst, value, bytes, ctx->stream));
  return FVM_OK;
}

// ---- marker store -----------------------------------------------------------
static const int kNumD = 23, kNumI = 5;

static double **dcol(fvm_markers *m, int i) {
  double **c[kNumD] = {&m->X,   &m->Y,   &m->VX,  &m->VY,     &m->T,   &m->Tdot, &m->sxx, &m->syy,
                       &m->sxy, &m->eta, &m->mu,  &m->rho,    &m->rhodot, &m->p, &m->exx, &m->eyy,
                       &m->exy, &m->Exx, &m->Eyy, &m->Exy,    &m->Eii, &m->wxy,  &m->strainRateResidual};
  return c[i];
}
static int32_t **icol(fvm_markers *m, int i) {
  int32_t **c[kNumI] = {&m->Mat, &m->cellX, &m->cellY, &m->cpu, &m->isYielding};
  return c[i];
}
// byte offset of each column's field inside the 216-byte Marker (structures.h)
__host__ __device__ constexpr int dcol_offset(int c) { return c < 4 ? c * 8 : (c < 22 ? 40 + (c - 4) * 8 : 200); }
__host__ __device__ constexpr int icol_offset(int c) { return c == 0 ? 32 : (c < 4 ? 184 + (c - 1) * 4 : 208); }
static_assert(dcol_offset(4) == 40 && dcol_offset(21) == 176 && dcol_offset(22) == 200, "T, wxy, residual");
static_assert(icol_offset(1) == 184 && icol_offset(3) == 192 && icol_offset(4) == 208, "cellX, cpu, isYielding");

static_assert(kNumD == FVM_NUM_DCOLS && kNumI == FVM_NUM_ICOLS, "column counts");
void fvm_marker_columns(fvm_markers *mk, double **d, int32_t **i) {
  for (int c = 0; c < kNumD; c++) d[c] = *dcol(mk, c);
  for (int c = 0; c < kNumI; c++) i[c] = *icol(mk, c);
}

static size_t col_stride(int32_t n) { return ((size_t)n * 8 + 255) / 256 * 256; }

extern "C" int fvm_markers_alloc(fvm_ctx *ctx, int32_t maxMark, fvm_markers *mk) {
  if (!mk || maxMark < 0) return FVM_ERR_ARG;
  memset(mk, 0, sizeof(*mk));
  const size_t s = col_stride(maxMark > 0 ? maxMark : 1);
  char *base = nullptr;
  hipError_t e = hipMalloc((void **)&base, s * (kNumD + kNumI));
  if (e != hipSuccess)
    return fvm_set_err(ctx, FVM_ERR_NOMEM, "marker store hipMalloc: %s", hipGetErrorString(e));
  for (int i = 0; i < kNumD; i++) *dcol(mk, i) = (double *)(base + s * i);
  for (int i = 0; i < kNumI; i++) *icol(mk, i) = (int32_t *)(base + s * (kNumD + i));
  mk->maxMark = maxMark;
  mk->nMark = 0;
  FVM_HIP(ctx, hipMemsetAsync(base, 0, s * (kNumD + kNumI), ctx->stream));
  FVM_HIP(ctx, hipMemsetAsync(mk->cellX, 0xff, s, ctx->stream));
  FVM_HIP(ctx, hipMemsetAsync(mk->cellY, 0xff, s, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

extern "C" int fvm_markers_free(fvm_markers *mk) {
  if (!mk) return FVM_OK;
  if (mk->X) (void)hipFree(mk->X);  // X is the slab base
  memset(mk, 0, sizeof(*mk));
  return FVM_OK;
}

struct ColTable {
  double *d[kNumD];
  int32_t *i[kNumI];
};

__global__ void k_aos_to_soa(const char *__restrict__ aos, int32_t n, ColTable t) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  const char *r = aos + m * 216;
#pragma unroll
  for (int c = 0; c < kNumD; c++) {
    double v;
    memcpy(&v, r + dcol_offset(c), 8);
    t.d[c][m] = v;
  }
#pragma unroll
  for (int c = 0; c < kNumI; c++) {
    int32_t v;
    memcpy(&v, r + icol_offset(c), 4);
    t.i[c][m] = v;
  }
}

__global__ void k_soa_to_aos(char *__restrict__ aos, int32_t n, ColTable t) {
  int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n) return;
  char *r = aos + m * 216;
  memset(r, 0, 216);
#pragma unroll
  for (int c = 0; c < kNumD; c++) {
    double v = t.d[c][m];
    memcpy(r + dcol_offset(c), &v, 8);
  }
#pragma unroll
  for (int c = 0; c < kNumI; c++) {
    int32_t v = t.i[c][m];
    memcpy(r + icol_offset(c), &v, 4);
  }
}

static ColTable cols(const fvm_markers *mk) {
  ColTable t;
  fvm_markers *m = const_cast<fvm_markers *>(mk);
  for (int i = 0; i < kNumD; i++) t.d[i] = *dcol(m, i);
  for (int i = 0; i < kNumI; i++) t.i[i] = *icol(m, i);
  return t;
}

extern "C" int fvm_markers_upload_aos(fvm_ctx *ctx, fvm_markers *mk, const Marker *aos, int32_t n) {
  if (!mk || (!aos && n > 0) || n < 0 || n > mk->maxMark)
    return fvm_set_err(ctx, FVM_ERR_ARG, "upload: bad arguments (n=%d, maxMark=%d)", n, mk ? mk->maxMark : -1);
  static_assert(sizeof(Marker) == 216, "Marker layout");
  mk->nMark = n;
  fvm_markers_moved(ctx);
  if (n == 0) return FVM_OK;
  FVM_TRY(fvm_ws(ctx, ctx->staging, (size_t)n * 216));
  FVM_HIP(ctx, hipMemcpyAsync(ctx->staging.ptr, aos, (size_t)n * 216, hipMemcpyHostToDevice, ctx->stream));
  k_aos_to_soa<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>((const char *)ctx->staging.ptr, n, cols(mk));
  FVM_LAUNCH_CHECK(ctx);
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

extern "C" int fvm_markers_download_aos(fvm_ctx *ctx, const fvm_markers *mk, Marker *aos) {
  if (!mk || (!aos && mk->nMark > 0)) return FVM_ERR_ARG;
  const int32_t n = mk->nMark;
  if (n == 0) return FVM_OK;
  FVM_TRY(fvm_ws(ctx, ctx->staging, (size_t)n * 216));
  k_soa_to_aos<<<fvm_blocks(n, 256), 256, 0, ctx->stream>>>((char *)ctx->staging.ptr, n, cols(mk));
  FVM_LAUNCH_CHECK(ctx);
  FVM_HIP(ctx, hipMemcpyAsync(aos, ctx->staging.ptr, (size_t)n * 216, hipMemcpyDeviceToHost, ctx->stream));
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return FVM_OK;
}

// ---- per-phase profiling ------------------------------------------------------
extern "C" int fvm_profile_enable(fvm_ctx *ctx, int32_t on) {
  if (!ctx) return FVM_ERR_ARG;
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  ctx->profiling = on ? 1 : 0;
  ctx->nprof = 0;
  return FVM_OK;
}

extern "C" int fvm_profile_read(fvm_ctx *ctx, double *ms, int32_t *launches) {
  if (!ctx || !ms || !launches) return FVM_ERR_ARG;
  FVM_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (int p = 0; p < FVM_NUM_PHASES; p++) {
    ms[p] = 0.0;
    launches[p] = 0;
  }
  for (int i = 0; i < ctx->nprof; i++) {
    float t = 0.f;
    FVM_HIP(ctx, hipEventElapsedTime(&t, ctx->prof_ev[2 * i], ctx->prof_ev[2 * i + 1]));
    const int p = ctx->prof_phase[i];
    if (p >= 0 && p < FVM_NUM_PHASES) {
      ms[p] += t;
      launches[p]++;
    }
  }
  ctx->nprof = 0;
  return FVM_OK;
}

// ---- timers ---------------------------------------------------------------------
extern "C" int fvm_timer_start(fvm_ctx *ctx) {
  FVM_HIP(ctx, hipEventRecord(ctx->ev0, ctx->stream));
  return FVM_OK;
}
extern "C" int fvm_timer_stop(fvm_ctx *ctx, float *ms) {
  FVM_HIP(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  FVM_HIP(ctx, hipEventSynchronize(ctx->ev1));
  float t = 0.f;
  FVM_HIP(ctx, hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
  if (ms) *ms = t;
  return FVM_OK;
}
