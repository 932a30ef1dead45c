// Marker files in the reference's binary layout, written by saveMarkersBinary
// (io.c:299-402) and read by restartFromMarkers (restart.c:8-125) and by
// post_processing/getBinaryMarkers.m: int32 count of active markers
// (cellX != -1), double elapsed time, then one column per field over the
// active markers in marker order -- cpu, cellX, cellY (int32), X, Y, VX, VY
// (double), Mat (int32), T, Tdot, eta, mu, exx, exy, Eii, sxx, syy, sxy, p,
// rho, rhodot (double).  152 bytes per marker.
//
// The file image is built on the device from the SoA columns: a scan of the
// active flags gives every active marker its row, one pass writes all 21
// columns, and the host issues one copy and one fwrite.  Reading reverses it.
// The image is a sequence of 32-bit words: a column of doubles starts at a
// 4-byte boundary only, so each double travels as two words.
#include <stdio.h>
#include <string.h>

#include "fvm_internal.h"

namespace {

#define NFILE 21
#define IO_BLOCK 256
// true for the int32 fields, in file order
constexpr bool kInt[NFILE] = {true,  true,  true,  false, false, false, false, true,  false, false, false,
                              false, false, false, false, false, false, false, false, false, false};

struct FileCols {
  const void *src[NFILE];
  void *dst[NFILE];
  int64_t word[NFILE];  // first word of each column in the body
};

static void file_cols(const fvm_markers *mk, int64_t n, FileCols *fc) {
  const void *cols[NFILE] = {mk->cpu, mk->cellX, mk->cellY, mk->X,   mk->Y,   mk->VX,  mk->VY,
                             mk->Mat, mk->T,     mk->Tdot,  mk->eta, mk->mu,  mk->exx, mk->exy,
                             mk->Eii, mk->sxx,   mk->syy,   mk->sxy, mk->p,   mk->rho, mk->rhodot};
  int64_t w = 0;
  for (int k = 0; k < NFILE; k++) {
    fc->src[k] = cols[k];
    fc->dst[k] = const_cast<void *>(cols[k]);
    fc->word[k] = w;
    w += kInt[k] ? n : 2 * n;
  }
}

__global__ void __launch_bounds__(IO_BLOCK) k_active(int32_t n, const int32_t *__restrict__ cellX,
                                                     int32_t *__restrict__ flag) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m < n) flag[m] = cellX[m] != -1 ? 1 : 0;
}

__global__ void __launch_bounds__(IO_BLOCK) k_pack(int32_t n, const int32_t *__restrict__ cellX,
                                                   const int32_t *__restrict__ row, FileCols fc,
                                                   uint32_t *__restrict__ body) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= n || cellX[m] == -1) return;
  const int64_t r = row[m];
#pragma unroll
  for (int k = 0; k < NFILE; k++) {
    if (kInt[k]) {
      body[fc.word[k] + r] = (uint32_t)((const int32_t *)fc.src[k])[m];
    } else {
      const unsigned long long b = (unsigned long long)__double_as_longlong(((const double *)fc.src[k])[m]);
      body[fc.word[k] + 2 * r] = (uint32_t)b;
      body[fc.word[k] + 2 * r + 1] = (uint32_t)(b >> 32);
    }
  }
}

// resetMarker (markers.c:250-330) on every slot, then the file's fields on the
// first n slots; strainRateResidual (left as it was by resetMarker) is zeroed
__global__ void __launch_bounds__(IO_BLOCK) k_unpack(int32_t n, int32_t cap, const uint32_t *__restrict__ body,
                                                     FileCols fc, fvm_markers mk) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= cap) return;
  double *zero_d[] = {mk.VX, mk.VY,  mk.T,   mk.Tdot, mk.sxx, mk.syy, mk.sxy, mk.eta, mk.mu,  mk.rho,
                      mk.rhodot, mk.p, mk.exx, mk.eyy, mk.exy, mk.Exx, mk.Eyy, mk.Exy, mk.Eii, mk.wxy,
                      mk.strainRateResidual};
  for (double *c : zero_d) c[m] = 0.0;
  mk.X[m] = -1.0;
  mk.Y[m] = -1.0;
  mk.Mat[m] = -1;
  mk.cellX[m] = 0;
  mk.cellY[m] = 0;
  mk.cpu[m] = 0;
  mk.isYielding[m] = 0;
  if (m >= n) return;
#pragma unroll
  for (int k = 0; k < NFILE; k++) {
    if (kInt[k]) {
      ((int32_t *)fc.dst[k])[m] = (int32_t)body[fc.word[k] + m];
    } else {
      const unsigned long long b =
          (unsigned long long)body[fc.word[k] + 2 * m] | ((unsigned long long)body[fc.word[k] + 2 * m + 1] << 32);
      ((double *)fc.dst[k])[m] = __longlong_as_double((long long)b);
    }
  }
}

// ---- nodal fields in PETSc's binary Vec format (io.c:150-296, -output_binary) ----
// 20 VecView records back to back, in saveNodalFields' order; each record is the
// big-endian int32 class id 1211214, the big-endian int32 length and the values
// as big-endian doubles in the DMDA's natural order ([jy][ix], ix fastest).
#define NNODAL 20
#define PETSC_VEC_CLASSID 1211214
struct NodalCols {
  const double *src[NNODAL];  // NULL writes zeros
};

__global__ void __launch_bounds__(IO_BLOCK) k_nodal_image(int64_t nn, NodalCols nc, uint32_t *__restrict__ img) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t rec = 2 + 2 * nn;  // words per record
#pragma unroll 4
  for (int k = 0; k < NNODAL; k++) {
    uint32_t *r = img + k * rec;
    if (i == 0) {
      r[0] = __builtin_bswap32((uint32_t)PETSC_VEC_CLASSID);
      r[1] = __builtin_bswap32((uint32_t)nn);
    }
    const unsigned long long u =
        nc.src[k] ? (unsigned long long)__double_as_longlong(nc.src[k][i]) : 0ull;
    r[2 + 2 * i] = __builtin_bswap32((uint32_t)(u >> 32));
    r[3 + 2 * i] = __builtin_bswap32((uint32_t)u);
  }
}

}  // namespace

extern "C" int fvm_save_nodal_fields(fvm_ctx *ctx, const fvm_grid *g, const NodalArrays *nf, const double *ha,
                                     const char *path) {
  if (!ctx || !nf || !path) return FVM_ERR_ARG;
  FVM_TRY(fvm_grid_check(ctx, g));
  const int64_t nn = (int64_t)g->NX * g->NY;
  if (nn > 0x7fffffff) return fvm_set_err(ctx, FVM_ERR_ARG, "PETSc binary Vec length limited to int32");
  // saveNodalFields' order (io.c:212-235)
  const double *cols[NNODAL] = {nf->vx,   nf->vy,     nf->p,      nf->rho,    nf->rhodot, nf->etaN,   nf->etaS,
                                nf->kThermal, nf->lastT, nf->Cp,   nf->muN,    nf->muS,    nf->soxx,   nf->soyy,
                                nf->soxy, nf->edotxx, nf->edotyy, nf->edotxy, nf->wxy,    ha};
  NodalCols nc;
  for (int k = 0; k < NNODAL; k++) {
    if (!cols[k] && k < NNODAL - 1)
      return fvm_set_err(ctx, FVM_ERR_ARG, "saveNodalFields: nodal field %d missing", k);
    nc.src[k] = cols[k];
  }
  const size_t bytes = (size_t)NNODAL * (8 + 8 * (size_t)nn);
  FVM_TRY(fvm_ws(ctx, ctx->staging, bytes));
  {
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_nodal_image<<<fvm_blocks(nn, IO_BLOCK), IO_BLOCK, 0, ctx->stream>>>(nn, nc, (uint32_t *)ctx->staging.ptr);
    FVM_LAUNCH_CHECK(ctx);
  }
  char *host = nullptr;
  FVM_HIP(ctx, hipHostMalloc((void **)&host, bytes, hipHostMallocDefault));
  int rc = FVM_OK;
  hipError_t e = hipMemcpyAsync(host, ctx->staging.ptr, bytes, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) rc = fvm_set_err(ctx, FVM_ERR_HIP, "nodal file copy: %s", hipGetErrorString(e));
  if (rc == FVM_OK) {
    FILE *f = fopen(path, "wb");
    if (!f) {
      rc = fvm_set_err(ctx, FVM_ERR_IO, "cannot open %s for writing", path);
    } else {
      const bool ok = fwrite(host, 1, bytes, f) == bytes;
      if (fclose(f) != 0 || !ok) rc = fvm_set_err(ctx, FVM_ERR_IO, "short write to %s", path);
    }
  }
  (void)hipHostFree(host);
  return rc;
}

extern "C" int fvm_save_markers_binary(fvm_ctx *ctx, const fvm_markers *mk, const char *path, double elapsedTime,
                                       int32_t *nSaved) {
  if (!ctx || !mk || !path) return FVM_ERR_ARG;
  const int32_t n = mk->nMark;
  // staging: flags (n), rows (n + 1), two spare counters, scan state, body
  const int64_t ntile = fvm_scan_tiles(n > 0 ? n : 1);
  const size_t off_rows = ((size_t)n * 4 + 255) / 256 * 256;
  const size_t off_cnt = off_rows + ((size_t)(n + 1) * 4 + 255) / 256 * 256;
  const size_t off_state = off_cnt + 256;
  const size_t off_body = off_state + ((size_t)(ntile + 2) * 8 + 255) / 256 * 256;
  const size_t body_bytes = (size_t)n * 152;
  FVM_TRY(fvm_ws(ctx, ctx->staging, off_body + (body_bytes > 0 ? body_bytes : 4)));
  char *base = (char *)ctx->staging.ptr;
  int32_t *flag = (int32_t *)base, *rows = (int32_t *)(base + off_rows), *cnt = (int32_t *)(base + off_cnt);
  unsigned long long *state = (unsigned long long *)(base + off_state);
  uint32_t *body = (uint32_t *)(base + off_body);
  int32_t nact = 0;
  if (n > 0) {
    FVM_TRY(fvm_launch_zero_sort(ctx, nullptr, 0, cnt, state, ntile + 2));
    {
      FVM_PHASE(ctx, FVM_PHASE_OTHER);
      k_active<<<fvm_blocks(n, IO_BLOCK), IO_BLOCK, 0, ctx->stream>>>(n, mk->cellX, flag);
      FVM_LAUNCH_CHECK(ctx);
    }
    FVM_TRY(fvm_launch_scan_onepass(ctx, flag, rows, state, n));
    int32_t *h = (int32_t *)ctx->pinned.ptr + 64;
    FVM_HIP(ctx, hipMemcpyAsync(h, rows + n, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
    FVM_TRY(fvm_ctx_sync(ctx));
    nact = *h;
    FileCols fc;
    file_cols(mk, nact, &fc);
    {
      FVM_PHASE(ctx, FVM_PHASE_OTHER);
      k_pack<<<fvm_blocks(n, IO_BLOCK), IO_BLOCK, 0, ctx->stream>>>(n, mk->cellX, rows, fc, body);
      FVM_LAUNCH_CHECK(ctx);
    }
  }
  const size_t bytes = 12 + (size_t)nact * 152;
  char *host = nullptr;
  FVM_HIP(ctx, hipHostMalloc((void **)&host, bytes, hipHostMallocDefault));
  memcpy(host, &nact, 4);
  memcpy(host + 4, &elapsedTime, 8);
  int rc = FVM_OK;
  if (nact > 0) {
    hipError_t e = hipMemcpyAsync(host + 12, body, (size_t)nact * 152, hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = fvm_set_err(ctx, FVM_ERR_HIP, "marker file copy: %s", hipGetErrorString(e));
  }
  if (rc == FVM_OK) {
    FILE *f = fopen(path, "wb");
    if (!f) {
      rc = fvm_set_err(ctx, FVM_ERR_IO, "cannot open %s for writing", path);
    } else {
      const bool ok = fwrite(host, 1, bytes, f) == bytes;
      if (fclose(f) != 0 || !ok) rc = fvm_set_err(ctx, FVM_ERR_IO, "short write to %s", path);
    }
  }
  (void)hipHostFree(host);
  if (rc == FVM_OK && nSaved) *nSaved = nact;
  return rc;
}

extern "C" int fvm_restart_from_markers(fvm_ctx *ctx, fvm_markers *mk, const char *path, double *elapsedTime) {
  if (!ctx || !mk || !path) return FVM_ERR_ARG;
  FILE *f = fopen(path, "rb");
  if (!f) return fvm_set_err(ctx, FVM_ERR_IO, "cannot open %s", path);
  int32_t n = -1;
  double t = 0.0;
  long size = -1;
  if (fread(&n, 4, 1, f) == 1 && fread(&t, 8, 1, f) == 1 && fseek(f, 0, SEEK_END) == 0) size = ftell(f);
  if (n < 0 || size != 12 + (long)n * 152) {
    fclose(f);
    return fvm_set_err(ctx, FVM_ERR_IO, "%s: not a marker file of this layout (count %d, %ld bytes)", path, n, size);
  }
  if (n > mk->maxMark) {
    fclose(f);
    return fvm_set_err(ctx, FVM_ERR_ARG, "%s holds %d markers, capacity is %d", path, n, mk->maxMark);
  }
  const size_t body_bytes = (size_t)n * 152;
  char *host = nullptr;
  if (hipHostMalloc((void **)&host, body_bytes > 0 ? body_bytes : 4, hipHostMallocDefault) != hipSuccess) {
    fclose(f);
    return fvm_set_err(ctx, FVM_ERR_NOMEM, "pinned buffer of %zu bytes", body_bytes);
  }
  const bool ok = fseek(f, 12, SEEK_SET) == 0 && fread(host, 1, body_bytes, f) == body_bytes;
  fclose(f);
  int rc = FVM_OK;
  if (!ok) rc = fvm_set_err(ctx, FVM_ERR_IO, "short read from %s", path);
  if (rc == FVM_OK) rc = fvm_ws(ctx, ctx->staging, body_bytes > 0 ? body_bytes : 4);
  if (rc == FVM_OK && n > 0) {
    hipError_t e = hipMemcpyAsync(ctx->staging.ptr, host, body_bytes, hipMemcpyHostToDevice, ctx->stream);
    if (e != hipSuccess) rc = fvm_set_err(ctx, FVM_ERR_HIP, "marker file copy: %s", hipGetErrorString(e));
  }
  if (rc == FVM_OK && mk->maxMark > 0) {
    FileCols fc;
    file_cols(mk, n, &fc);
    FVM_PHASE(ctx, FVM_PHASE_OTHER);
    k_unpack<<<fvm_blocks(mk->maxMark, IO_BLOCK), IO_BLOCK, 0, ctx->stream>>>(
        n, mk->maxMark, (const uint32_t *)ctx->staging.ptr, fc, *mk);
    if (hipGetLastError() != hipSuccess) rc = fvm_set_err(ctx, FVM_ERR_HIP, "k_unpack launch failed");
  }
  // the pinned buffer may be freed only after the copy has read it
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && rc == FVM_OK)
    rc = fvm_set_err(ctx, FVM_ERR_HIP, "marker file upload failed");
  (void)hipHostFree(host);
  if (rc != FVM_OK) return rc;
  mk->nMark = n;
  fvm_markers_moved(ctx);
  if (elapsedTime) *elapsedTime = t;
  return FVM_OK;
}
