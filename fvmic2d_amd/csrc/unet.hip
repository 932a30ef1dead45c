// Learned initial guess for the Stokes solve: the U-Net's layers as hand-written
// gfx950 kernels.  Activations are NHWC bfloat16 ([B][H][W][C], C fastest),
// weights bfloat16 [tap][Cout][Cin] (tap = ky*K + kx) with fp32 master copies
// for the optimizer; every product accumulates in fp32 on gfx950's double-K
// bf16 matrix instruction (v_mfma_f32_16x16x32_bf16: K = 32 per issue, twice
// the CDNA3 16x16x16 shape).  There is no reference counterpart: the
// reference's only initial guess is initialPressureGuess.c.
//
// MFMA 16x16x32 bf16 operand layout (one wavefront): lane l supplies
//   A: row m = l % 16, k = 8*(l/16) .. 8*(l/16)+7   (eight bf16, one 16-byte load)
//   B: col n = l % 16, k = 8*(l/16) .. 8*(l/16)+7
// and receives D[m = 4*(l/16) + i][n = l % 16], i = 0..3 (four fp32).
// fvm_nn_mfma_selftest checks this layout against a host product.
#include <hip/hip_bf16.h>

#include "fvm_internal.h"

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// round to nearest even (finite values; NaN stays NaN)
__device__ __forceinline__ uint16_t f2bf(float f) {
  const uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  const uint32_t r = u + 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(r >> 16);
}

__device__ __forceinline__ f4 mfma32(bf8 a, bf8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// 8 consecutive bf16 from a 16-byte aligned address, or zeros
__device__ __forceinline__ bf8 ld8(const uint16_t *p, bool ok) {
  const uint4 u = ok ? *(const uint4 *)p : make_uint4(0u, 0u, 0u, 0u);
  return __builtin_bit_cast(bf8, u);
}

// eight bf16 gathered one by one (the weight gradient's pixel-strided operands)
struct Bf8Builder {
  uint16_t h[8];
  __device__ __forceinline__ bf8 get() const {
    uint4 u;
    u.x = (uint32_t)h[0] | ((uint32_t)h[1] << 16);
    u.y = (uint32_t)h[2] | ((uint32_t)h[3] << 16);
    u.z = (uint32_t)h[4] | ((uint32_t)h[5] << 16);
    u.w = (uint32_t)h[6] | ((uint32_t)h[7] << 16);
    return __builtin_bit_cast(bf8, u);
  }
};

// ---- MFMA layout self-test: D(16x16) = A(16x32) B(32x16) ---------------------
__global__ void __launch_bounds__(64) k_mfma_selftest(const uint16_t *__restrict__ A, const uint16_t *__restrict__ B,
                                                      float *__restrict__ D) {
  const int l = threadIdx.x;
  const int m = l % 16, kg = l / 16;
  Bf8Builder a, b;
  for (int i = 0; i < 8; i++) {
    a.h[i] = A[m * 32 + 8 * kg + i];      // A row-major [m][k], 16 x 32
    b.h[i] = B[(8 * kg + i) * 16 + m];    // B row-major [k][n], 32 x 16, n = l % 16
  }
  f4 c = {0.f, 0.f, 0.f, 0.f};
  c = mfma32(a.get(), b.get(), c);
  for (int i = 0; i < 4; i++) D[(4 * kg + i) * 16 + m] = c[i];
}

// ---- convolution (implicit GEMM) ------------------------------------------------
// Y[p][co] = act(bias[co] + sum_{tap, ci} X[p + off(tap)][ci] * W[tap][co][ci])
// off(tap) = (ky - K/2, kx - K/2); X outside the image reads 0.
// Block: 256 threads = 4 wavefronts in a 2x2 arrangement, tile 64 pixels x 64
// output channels; each wavefront 32 x 32 = 2 x 2 MFMA tiles.
struct ConvArgs {
  const uint16_t *X;
  const uint16_t *W;
  const float *bias;  // may be null
  uint16_t *Y;        // bf16 output (or null)
  float *Yf;          // fp32 output (or null)
  int B, H, Wd, Cin, Cout, K, relu;
  int cout_real;  // channels >= cout_real are written as 0 (padding)
};

__global__ void __launch_bounds__(256) k_conv(ConvArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t npix = (int64_t)a.B * a.H * a.Wd;
  const int64_t p0 = (int64_t)blockIdx.x * 64 + (wv & 1) * 32;
  const int c0 = blockIdx.y * 64 + (wv >> 1) * 32;
  const int ml = lane % 16, kg = lane / 16;
  // the two pixels (one per M tile) this lane loads A rows for
  int64_t pp[2];
  int pb[2], ph[2], pw[2];
  bool pv[2];
  for (int t = 0; t < 2; t++) {
    const int64_t p = p0 + t * 16 + ml;
    pv[t] = p < npix;
    const int64_t q = pv[t] ? p : 0;
    pw[t] = (int)(q % a.Wd);
    ph[t] = (int)((q / a.Wd) % a.H);
    pb[t] = (int)(q / ((int64_t)a.Wd * a.H));
    pp[t] = q;
  }
  f4 acc[2][2];
  for (int i = 0; i < 2; i++)
    for (int j = 0; j < 2; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
  const int half = a.K / 2;
  for (int tap = 0; tap < a.K * a.K; tap++) {
    const int dy = tap / a.K - half, dx = tap % a.K - half;
    const uint16_t *xs[2];
    bool xv[2];
    for (int t = 0; t < 2; t++) {
      const int hh = ph[t] + dy, ww = pw[t] + dx;
      xv[t] = pv[t] && hh >= 0 && hh < a.H && ww >= 0 && ww < a.Wd;
      xs[t] = a.X + ((((int64_t)pb[t] * a.H + (xv[t] ? hh : 0)) * a.Wd + (xv[t] ? ww : 0)) * a.Cin);
    }
    const uint16_t *wt = a.W + (int64_t)tap * a.Cout * a.Cin;
    // Cin is a multiple of 16: a lane's eight channels lie wholly inside or
    // wholly past it (a 16-channel layer leaves half of K at zero)
    for (int ci0 = 0; ci0 < a.Cin; ci0 += 32) {
      const int ck = ci0 + 8 * kg;
      const bool kin = ck < a.Cin;
      bf8 fa[2], fb[2];
      for (int t = 0; t < 2; t++) fa[t] = ld8(xs[t] + ck, xv[t] && kin);
      for (int u = 0; u < 2; u++) {
        const int co = c0 + u * 16 + ml;
        fb[u] = ld8(wt + (int64_t)co * a.Cin + ck, co < a.Cout && kin);
      }
      for (int t = 0; t < 2; t++)
        for (int u = 0; u < 2; u++) acc[t][u] = mfma32(fa[t], fb[u], acc[t][u]);
    }
  }
  // epilogue: D[m = 4*kg + i][n = ml] of tile (t, u)
  for (int t = 0; t < 2; t++)
    for (int u = 0; u < 2; u++) {
      const int co = c0 + u * 16 + ml;
      if (co >= a.Cout) continue;
      const float bv = a.bias ? a.bias[co] : 0.f;
      for (int i = 0; i < 4; i++) {
        const int64_t p = p0 + t * 16 + 4 * kg + i;
        if (p >= npix) continue;
        float v = acc[t][u][i] + bv;
        if (a.relu) v = v > 0.f ? v : 0.f;
        if (co >= a.cout_real) v = 0.f;
        if (a.Y) a.Y[p * a.Cout + co] = f2bf(v);
        if (a.Yf) a.Yf[p * a.Cout + co] = v;
      }
    }
}

// ---- weight gradient --------------------------------------------------------------
// dW[tap][co][ci] = sum_p dY[p][co] * X[p + off(tap)][ci]: per tap a GEMM with
// M = Cout, N = Cin and the pixels as K, split into chunks of KCH pixels whose
// fp32 partials are summed in chunk order by k_wgrad_reduce.
#define KCH 256
struct WgradArgs {
  const uint16_t *X, *dY;
  float *part;  // [chunk][tap][Cout][Cin]
  int B, H, Wd, Cin, Cout, K;
  int64_t nchunk;
};

__global__ void __launch_bounds__(256) k_wgrad(WgradArgs a) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int tap = blockIdx.y;
  const int64_t chunk = blockIdx.x;
  const int ntn = (a.Cin + 31) / 32;  // 32-wide ci tiles per wavefront
  const int ntm = (a.Cout + 31) / 32;
  const int half = a.K / 2;
  const int dy = tap / a.K - half, dx = tap % a.K - half;
  const int64_t npix = (int64_t)a.B * a.H * a.Wd;
  const int ml = lane % 16, kg = lane / 16;
  for (int tile = blockIdx.z * 4 + wv; tile < ntm * ntn; tile += gridDim.z * 4) {
    const int co0 = (tile / ntn) * 32, ci0 = (tile % ntn) * 32;
    f4 acc[2][2];
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) acc[i][j] = (f4){0.f, 0.f, 0.f, 0.f};
    for (int k0 = 0; k0 < KCH; k0 += 32) {
      Bf8Builder fa[2], fb[2];
      for (int i = 0; i < 8; i++) {
        const int64_t p = chunk * KCH + k0 + 8 * kg + i;
        bool ok = p < npix;
        const int64_t q = ok ? p : 0;
        const int w = (int)(q % a.Wd), h = (int)((q / a.Wd) % a.H);
        const int64_t bq = q / ((int64_t)a.Wd * a.H);
        const int hh = h + dy, ww = w + dx;
        const bool xok = ok && hh >= 0 && hh < a.H && ww >= 0 && ww < a.Wd;
        const int64_t qx = ((bq * a.H + (xok ? hh : 0)) * a.Wd + (xok ? ww : 0));
        for (int t = 0; t < 2; t++) {
          const int co = co0 + t * 16 + ml;
          fa[t].h[i] = (ok && co < a.Cout) ? a.dY[q * a.Cout + co] : (uint16_t)0;
        }
        for (int u = 0; u < 2; u++) {
          const int ci = ci0 + u * 16 + ml;
          fb[u].h[i] = (xok && ci < a.Cin) ? a.X[qx * a.Cin + ci] : (uint16_t)0;
        }
      }
      for (int t = 0; t < 2; t++)
        for (int u = 0; u < 2; u++) acc[t][u] = mfma32(fa[t].get(), fb[u].get(), acc[t][u]);
    }
    float *out = a.part + ((chunk * (a.K * a.K) + tap) * a.Cout) * (int64_t)a.Cin;
    for (int t = 0; t < 2; t++)
      for (int u = 0; u < 2; u++) {
        const int ci = ci0 + u * 16 + ml;
        if (ci >= a.Cin) continue;
        for (int i = 0; i < 4; i++) {
          const int co = co0 + t * 16 + 4 * kg + i;
          if (co < a.Cout) out[(int64_t)co * a.Cin + ci] = acc[t][u][i];
        }
      }
  }
}

__global__ void __launch_bounds__(256) k_wgrad_reduce(const float *__restrict__ part, int64_t nchunk, int64_t nw,
                                                      float *__restrict__ dW) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nw) return;
  float s = 0.f;
  for (int64_t c = 0; c < nchunk; c++) s += part[c * nw + i];
  dW[i] = s;
}

// db[co] = sum_p dY[p][co], per block partial sums then a chunk-ordered pass
__global__ void __launch_bounds__(256) k_bias_grad(const uint16_t *__restrict__ dY, int64_t npix, int C,
                                                   float *__restrict__ part) {
  // block b sums pixels [b*KCH, (b+1)*KCH) for every channel (thread per channel)
  for (int c = threadIdx.x; c < C; c += blockDim.x) {
    float s = 0.f;
    for (int k = 0; k < KCH; k++) {
      const int64_t p = (int64_t)blockIdx.x * KCH + k;
      if (p < npix) s += bf2f(dY[p * C + c]);
    }
    part[(int64_t)blockIdx.x * C + c] = s;
  }
}

// dWd[tap'][ci][co] = W[K*K-1-tap'][co][ci]: weights of the data-gradient conv
__global__ void __launch_bounds__(256) k_flip_weights(const uint16_t *__restrict__ W, int K, int Cout, int Cin,
                                                      uint16_t *__restrict__ Wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)K * K * Cout * Cin;
  if (i >= n) return;
  const int ci = (int)(i % Cin);
  const int co = (int)((i / Cin) % Cout);
  const int tap = (int)(i / ((int64_t)Cin * Cout));
  Wd[((int64_t)(K * K - 1 - tap) * Cin + ci) * Cout + co] = W[i];
}

// ---- elementwise layers ------------------------------------------------------------
__global__ void __launch_bounds__(256) k_relu_bwd(int64_t n, const uint16_t *__restrict__ y,
                                                  const uint16_t *__restrict__ dy, uint16_t *__restrict__ dz) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dz[i] = bf2f(y[i]) > 0.f ? dy[i] : (uint16_t)0;
}

// 2x2 average pooling (H, W even)
__global__ void __launch_bounds__(256) k_pool_fwd(const uint16_t *__restrict__ X, int B, int H, int W, int C,
                                                  uint16_t *__restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int Ho = H / 2, Wo = W / 2;
  const int64_t n = (int64_t)B * Ho * Wo * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % Wo), h = (int)((q / Wo) % Ho);
  const int64_t b = q / ((int64_t)Wo * Ho);
  const int64_t r0 = ((b * H + 2 * h) * W + 2 * w) * C + c, r1 = r0 + (int64_t)W * C;
  const float s = (bf2f(X[r0]) + bf2f(X[r0 + C])) + (bf2f(X[r1]) + bf2f(X[r1 + C]));
  Y[i] = f2bf(0.25f * s);
}

__global__ void __launch_bounds__(256) k_pool_bwd(const uint16_t *__restrict__ dY, int B, int H, int W, int C,
                                                  uint16_t *__restrict__ dX) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)B * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % W), h = (int)((q / W) % H);
  const int64_t b = q / ((int64_t)W * H);
  const int Ho = H / 2, Wo = W / 2;
  dX[i] = f2bf(0.25f * bf2f(dY[((b * Ho + h / 2) * Wo + w / 2) * C + c]));
}

// nearest-neighbour 2x upsampling written into channels [c0, c0+C) of a Ctot-channel tensor
__global__ void __launch_bounds__(256) k_up_fwd(const uint16_t *__restrict__ X, int B, int H, int W, int C,
                                                uint16_t *__restrict__ Y, int Ctot, int c0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int H2 = 2 * H, W2 = 2 * W;
  const int64_t n = (int64_t)B * H2 * W2 * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % W2), h = (int)((q / W2) % H2);
  const int64_t b = q / ((int64_t)W2 * H2);
  Y[q * Ctot + c0 + c] = X[((b * H + h / 2) * W + w / 2) * C + c];
}

__global__ void __launch_bounds__(256) k_up_bwd(const uint16_t *__restrict__ dY, int B, int H, int W, int C,
                                                int Ctot, int c0, uint16_t *__restrict__ dX) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t n = (int64_t)B * H * W * C;
  if (i >= n) return;
  const int c = (int)(i % C);
  const int64_t q = i / C;
  const int w = (int)(q % W), h = (int)((q / W) % H);
  const int64_t b = q / ((int64_t)W * H);
  const int W2 = 2 * W;
  const int64_t r0 = ((b * 2 * H + 2 * h) * W2 + 2 * w) * Ctot + c0 + c, r1 = r0 + (int64_t)W2 * Ctot;
  const float s = (bf2f(dY[r0]) + bf2f(dY[r0 + Ctot])) + (bf2f(dY[r1]) + bf2f(dY[r1 + Ctot]));
  dX[i] = f2bf(s);
}

// copy channels: Y[p][c0 + c] = X[p][c] (C channels of a Cx-channel X into a Ctot-channel Y)
__global__ void __launch_bounds__(256) k_copy_channels(const uint16_t *__restrict__ X, int64_t npix, int C, int Cx,
                                                       int xc0, uint16_t *__restrict__ Y, int Ctot, int c0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * C) return;
  const int c = (int)(i % C);
  const int64_t p = i / C;
  Y[p * Ctot + c0 + c] = X[p * Cx + xc0 + c];
}

// add channels: Y[p][c] += X[p][xc0 + c]  (bf16 in, bf16 out, fp32 sum)
__global__ void __launch_bounds__(256) k_add_channels(const uint16_t *__restrict__ X, int64_t npix, int C, int Cx,
                                                      int xc0, uint16_t *__restrict__ Y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npix * C) return;
  const int c = (int)(i % C);
  const int64_t p = i / C;
  Y[i] = f2bf(bf2f(Y[i]) + bf2f(X[p * Cx + xc0 + c]));
}

// masked mean squared error over the first Creal of C fp32 output channels:
// loss partials per block; dOut (bf16, C channels) = 2 (out - tgt) mask / count
__global__ void __launch_bounds__(256) k_mse(const float *__restrict__ out, const float *__restrict__ tgt,
                                             const float *__restrict__ mask, int64_t npix, int C, int Creal,
                                             float inv_count, uint16_t *__restrict__ dOut, float *__restrict__ part) {
  __shared__ float sh[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float l = 0.f;
  if (i < npix * C) {
    const int c = (int)(i % C);
    const int64_t p = i / C;
    float g = 0.f;
    if (c < Creal) {
      const float m = mask ? mask[p * Creal + c] : 1.f;
      const float d = (out[i] - tgt[p * Creal + c]) * m;
      l = d * d * inv_count;
      g = 2.f * d * inv_count;
    }
    dOut[i] = f2bf(g);
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(256) k_sum_f(const float *__restrict__ part, int64_t n, float *__restrict__ out) {
  __shared__ float sh[256];
  float s = 0.f;
  for (int64_t i = threadIdx.x; i < n; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = sh[0];
}

// Adam on fp32 master weights; the bf16 working copy is refreshed
__global__ void __launch_bounds__(256) k_adam(int64_t n, float *__restrict__ w, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, uint16_t *__restrict__ wb,
                                              float lr, float b1, float b2, float eps, float c1, float c2, float wd) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i] + wd * w[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = f2bf(wi);
}

__global__ void __launch_bounds__(256) k_f2bf(int64_t n, const float *__restrict__ x, uint16_t *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = f2bf(x[i]);
}

unsigned nblk(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---- C ABI (sizes in elements; all pointers device) ---------------------------------
extern "C" int fvm_nn_mfma_selftest(fvm_ctx *ctx, const uint16_t *A, const uint16_t *B, float *D) {
  if (!ctx || !A || !B || !D) return FVM_ERR_ARG;
  k_mfma_selftest<<<1, 64, 0, ctx->stream>>>(A, B, D);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

static int conv_check(fvm_ctx *ctx, int B, int H, int W, int Cin, int Cout, int K) {
  if (B < 1 || H < 1 || W < 1 || Cin < 16 || Cin % 16 || Cout < 16 || Cout % 16 || (K != 1 && K != 3))
    return fvm_set_err(ctx, FVM_ERR_ARG, "conv: B,H,W >= 1, Cin and Cout multiples of 16, K in {1,3}");
  return FVM_OK;
}

extern "C" int fvm_nn_conv_fwd(fvm_ctx *ctx, const uint16_t *X, const uint16_t *W, const float *bias, int32_t B,
                               int32_t H, int32_t Wd, int32_t Cin, int32_t Cout, int32_t K, int32_t relu,
                               int32_t cout_real, uint16_t *Y, float *Yf) {
  if (!ctx || !X || !W || (!Y && !Yf)) return FVM_ERR_ARG;
  FVM_TRY(conv_check(ctx, B, H, Wd, Cin, Cout, K));
  ConvArgs a{X, W, bias, Y, Yf, B, H, Wd, Cin, Cout, K, relu, cout_real > 0 ? cout_real : Cout};
  const int64_t npix = (int64_t)B * H * Wd;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_conv<<<dim3((unsigned)((npix + 63) / 64), (unsigned)((Cout + 63) / 64)), 256, 0, ctx->stream>>>(a);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

// dX = conv(dY, flipped transposed W); Wd_scratch: K*K*Cout*Cin bf16
extern "C" int fvm_nn_conv_dgrad(fvm_ctx *ctx, const uint16_t *dY, const uint16_t *W, uint16_t *Wd_scratch, int32_t B,
                                 int32_t H, int32_t Wd, int32_t Cin, int32_t Cout, int32_t K, uint16_t *dX) {
  if (!ctx || !dY || !W || !Wd_scratch || !dX) return FVM_ERR_ARG;
  FVM_TRY(conv_check(ctx, B, H, Wd, Cin, Cout, K));
  const int64_t nw = (int64_t)K * K * Cout * Cin;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_flip_weights<<<nblk(nw), 256, 0, ctx->stream>>>(W, K, Cout, Cin, Wd_scratch);
  FVM_LAUNCH_CHECK(ctx);
  ConvArgs a{dY, Wd_scratch, nullptr, dX, nullptr, B, H, Wd, Cout, Cin, K, 0, Cin};
  const int64_t npix = (int64_t)B * H * Wd;
  k_conv<<<dim3((unsigned)((npix + 63) / 64), (unsigned)((Cin + 63) / 64)), 256, 0, ctx->stream>>>(a);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

// bytes of fp32 workspace fvm_nn_conv_wgrad needs
extern "C" size_t fvm_nn_wgrad_workspace(int32_t B, int32_t H, int32_t Wd, int32_t Cin, int32_t Cout, int32_t K) {
  const int64_t npix = (int64_t)B * H * Wd;
  const int64_t nchunk = (npix + KCH - 1) / KCH;
  return (size_t)nchunk * ((size_t)K * K * Cout * Cin + (size_t)Cout) * sizeof(float);
}

// dW (fp32 [tap][Cout][Cin]) and db (fp32 [Cout], may be null)
extern "C" int fvm_nn_conv_wgrad(fvm_ctx *ctx, const uint16_t *X, const uint16_t *dY, int32_t B, int32_t H,
                                 int32_t Wd, int32_t Cin, int32_t Cout, int32_t K, float *work, float *dW, float *db) {
  if (!ctx || !X || !dY || !work || !dW) return FVM_ERR_ARG;
  FVM_TRY(conv_check(ctx, B, H, Wd, Cin, Cout, K));
  const int64_t npix = (int64_t)B * H * Wd;
  const int64_t nchunk = (npix + KCH - 1) / KCH;
  const int64_t nw = (int64_t)K * K * Cout * Cin;
  WgradArgs a{X, dY, work, B, H, Wd, Cin, Cout, K, nchunk};
  const int tiles = ((Cout + 31) / 32) * ((Cin + 31) / 32);
  const int gz = (tiles + 3) / 4;
  FVM_PHASE(ctx, FVM_PHASE_OTHER);
  k_wgrad<<<dim3((unsigned)nchunk, K * K, gz), 256, 0, ctx->stream>>>(a);
  FVM_LAUNCH_CHECK(ctx);
  k_wgrad_reduce<<<nblk(nw), 256, 0, ctx->stream>>>(work, nchunk, nw, dW);
  FVM_LAUNCH_CHECK(ctx);
  if (db) {
    float *bp = work + nchunk * nw;
    k_bias_grad<<<(unsigned)nchunk, 256, 0, ctx->stream>>>(dY, npix, Cout, bp);
    k_wgrad_reduce<<<nblk(Cout), 256, 0, ctx->stream>>>(bp, nchunk, Cout, db);
    FVM_LAUNCH_CHECK(ctx);
  }
  return FVM_OK;
}

extern "C" int fvm_nn_relu_bwd(fvm_ctx *ctx, int64_t n, const uint16_t *y, const uint16_t *dy, uint16_t *dz) {
  if (!ctx || !y || !dy || !dz) return FVM_ERR_ARG;
  k_relu_bwd<<<nblk(n), 256, 0, ctx->stream>>>(n, y, dy, dz);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_pool_fwd(fvm_ctx *ctx, const uint16_t *X, int32_t B, int32_t H, int32_t W, int32_t C,
                               uint16_t *Y) {
  if (!ctx || !X || !Y || H % 2 || W % 2) return FVM_ERR_ARG;
  k_pool_fwd<<<nblk((int64_t)B * (H / 2) * (W / 2) * C), 256, 0, ctx->stream>>>(X, B, H, W, C, Y);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_pool_bwd(fvm_ctx *ctx, const uint16_t *dY, int32_t B, int32_t H, int32_t W, int32_t C,
                               uint16_t *dX) {
  if (!ctx || !dY || !dX || H % 2 || W % 2) return FVM_ERR_ARG;
  k_pool_bwd<<<nblk((int64_t)B * H * W * C), 256, 0, ctx->stream>>>(dY, B, H, W, C, dX);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_up_fwd(fvm_ctx *ctx, const uint16_t *X, int32_t B, int32_t H, int32_t W, int32_t C, uint16_t *Y,
                             int32_t Ctot, int32_t c0) {
  if (!ctx || !X || !Y || c0 < 0 || c0 + C > Ctot) return FVM_ERR_ARG;
  k_up_fwd<<<nblk((int64_t)B * 4 * H * W * C), 256, 0, ctx->stream>>>(X, B, H, W, C, Y, Ctot, c0);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_up_bwd(fvm_ctx *ctx, const uint16_t *dY, int32_t B, int32_t H, int32_t W, int32_t C,
                             int32_t Ctot, int32_t c0, uint16_t *dX) {
  if (!ctx || !dY || !dX || c0 < 0 || c0 + C > Ctot) return FVM_ERR_ARG;
  k_up_bwd<<<nblk((int64_t)B * H * W * C), 256, 0, ctx->stream>>>(dY, B, H, W, C, Ctot, c0, dX);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_copy_channels(fvm_ctx *ctx, const uint16_t *X, int64_t npix, int32_t C, int32_t Cx, int32_t xc0,
                                    uint16_t *Y, int32_t Ctot, int32_t c0) {
  if (!ctx || !X || !Y || xc0 < 0 || xc0 + C > Cx || c0 < 0 || c0 + C > Ctot) return FVM_ERR_ARG;
  k_copy_channels<<<nblk(npix * C), 256, 0, ctx->stream>>>(X, npix, C, Cx, xc0, Y, Ctot, c0);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_add_channels(fvm_ctx *ctx, const uint16_t *X, int64_t npix, int32_t C, int32_t Cx, int32_t xc0,
                                   uint16_t *Y) {
  if (!ctx || !X || !Y || xc0 < 0 || xc0 + C > Cx) return FVM_ERR_ARG;
  k_add_channels<<<nblk(npix * C), 256, 0, ctx->stream>>>(X, npix, C, Cx, xc0, Y);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

// loss (fp32 scalar, device) = sum over pixels and real channels of ((out - tgt) mask)^2 * inv_count
extern "C" int fvm_nn_mse(fvm_ctx *ctx, const float *out, const float *tgt, const float *mask, int64_t npix, int32_t C,
                          int32_t Creal, float inv_count, uint16_t *dOut, float *work, float *loss) {
  if (!ctx || !out || !tgt || !dOut || !work || !loss || Creal > C) return FVM_ERR_ARG;
  const int64_t n = npix * C;
  const unsigned nb = nblk(n);
  k_mse<<<nb, 256, 0, ctx->stream>>>(out, tgt, mask, npix, C, Creal, inv_count, dOut, work);
  k_sum_f<<<1, 256, 0, ctx->stream>>>(work, nb, loss);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_adam(fvm_ctx *ctx, int64_t n, float *w, const float *g, float *m, float *v, uint16_t *wb,
                           float lr, float b1, float b2, float eps, int32_t step, float wd) {
  if (!ctx || !w || !g || !m || !v || step < 1) return FVM_ERR_ARG;
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  k_adam<<<nblk(n), 256, 0, ctx->stream>>>(n, w, g, m, v, wb, lr, b1, b2, eps, c1, c2, wd);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}

extern "C" int fvm_nn_f32_to_bf16(fvm_ctx *ctx, int64_t n, const float *x, uint16_t *y) {
  if (!ctx || !x || !y) return FVM_ERR_ARG;
  k_f2bf<<<nblk(n), 256, 0, ctx->stream>>>(n, x, y);
  FVM_LAUNCH_CHECK(ctx);
  return FVM_OK;
}
