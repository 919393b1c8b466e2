// Kernels of the warm-start network: 3x3x3 convolutions as implicit GEMMs
// on the gfx950 matrix cores (v_mfma_f32_16x16x16_bf16: BF16 operands,
// FP32 accumulation), average pooling, nearest upsampling, skip additions,
// ReLU, the MSE loss and Adam.
//
// Tensors are voxel-major with channels fastest: T[v*C + c], v = x + X*(y +
// Y*z) -- the same order as the solver's packed vectors (r, theta, phi).
// Activations are BF16 (raw bits in uint16); gradients are FP32; weights
// are kept in FP32 (Adam state) with a BF16 copy [tap][c_out][c_in] for the
// matrix cores.  The z axis (phi) is periodic, x and y are zero-padded.
//
// MFMA 16x16x16 operand layout (lane l of a wave64):
//   A (16 x 16, M x K): lane holds A[l%16][4*(l/16) .. 4*(l/16)+3]
//   B (16 x 16, K x N): lane holds B[4*(l/16) .. 4*(l/16)+3][l%16]
//   D (16 x 16, M x N): lane holds D[4*(l/16) .. 4*(l/16)+3][l%16]
// (checked on the device against a host matmul by p3d_nn_mfma_check).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "p3d_nn.hpp"

namespace p3d {
namespace nn {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)(u >> 16);  // inf/nan
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

struct Dims {
  int X, Y, Z;
  int V;
};

// neighbour of (x,y,z) at tap t = (dx+1) + 3*(dy+1) + 9*(dz+1), or -1
__device__ __forceinline__ int neighbour(const Dims& d, int x, int y, int z, int t, int sign) {
  const int dx = (t % 3 - 1) * sign, dy = ((t / 3) % 3 - 1) * sign, dz = (t / 9 - 1) * sign;
  const int xn = x + dx, yn = y + dy;
  if (xn < 0 || xn >= d.X || yn < 0 || yn >= d.Y) return -1;
  int zn = z + dz;
  if (zn < 0) zn += d.Z;
  if (zn >= d.Z) zn -= d.Z;
  return xn + d.X * (yn + d.Y * zn);
}

__device__ __forceinline__ void coords(const Dims& d, int v, int& x, int& y, int& z) {
  x = v % d.X;
  const int r = v / d.X;
  y = r % d.Y;
  z = r / d.Y;
}

// ---------------------------------------------------------------------------
// y[v][co] = act( sum_t sum_ci x[nb(v,t)][ci] w[t][co][ci] + bias[co] )
// grid (ceil(V/64), Cout/16), block 256: each wave computes 16 voxels x 16
// output channels.
__global__ void __launch_bounds__(256)
    k_conv_fwd(Dims d, int Cin, int Cout, const uint16_t* __restrict__ x,
               const uint16_t* __restrict__ w, const float* __restrict__ bias,
               uint16_t* __restrict__ y, float* __restrict__ yf, int relu) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int v0 = (blockIdx.x * 4 + wave) * 16;
  const int co0 = blockIdx.y * 16;
  const int m = lane & 15, kb = lane >> 4;
  const int v = v0 + m;
  int vx = 0, vy = 0, vz = 0;
  const bool valid = v < d.V;
  if (valid) coords(d, v, vx, vy, vz);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int ncb = Cin / 16;
  for (int t = 0; t < 27; t++) {
    const int vn = valid ? neighbour(d, vx, vy, vz, t, 1) : -1;
    const uint16_t* wt = w + ((size_t)t * Cout + co0 + m) * Cin + 4 * kb;
    for (int cb = 0; cb < ncb; cb++) {
      bf16x4 a = {0, 0, 0, 0};
      if (vn >= 0) a = *(const bf16x4*)(x + (size_t)vn * Cin + cb * 16 + 4 * kb);
      const bf16x4 b = *(const bf16x4*)(wt + cb * 16);
      acc = mfma16(a, b, acc);
    }
  }
  const int co = co0 + m;
  const float bv = bias ? bias[co] : 0.f;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int vo = v0 + 4 * kb + i;
    if (vo < d.V) {
      float val = acc[i] + bv;
      if (relu && val < 0.f) val = 0.f;
      if (y) y[(size_t)vo * Cout + co] = f2bf(val);
      if (yf) yf[(size_t)vo * Cout + co] = val;
    }
  }
}

// dx[v][ci] = sum_t sum_co dy[nb(v,-t)][co] w[t][co][ci]   (FP32 out)
__global__ void __launch_bounds__(256)
    k_conv_bwd_data(Dims d, int Cin, int Cout, const float* __restrict__ dy,
                    const uint16_t* __restrict__ w, float* __restrict__ dx) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int v0 = (blockIdx.x * 4 + wave) * 16;
  const int ci0 = blockIdx.y * 16;
  const int m = lane & 15, kb = lane >> 4;
  const int v = v0 + m;
  int vx = 0, vy = 0, vz = 0;
  const bool valid = v < d.V;
  if (valid) coords(d, v, vx, vy, vz);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  const int ncb = Cout / 16;
  for (int t = 0; t < 27; t++) {
    const int vn = valid ? neighbour(d, vx, vy, vz, t, -1) : -1;
    for (int cb = 0; cb < ncb; cb++) {
      bf16x4 a = {0, 0, 0, 0};
      if (vn >= 0) {
        const float* src = dy + (size_t)vn * Cout + cb * 16 + 4 * kb;
#pragma unroll
        for (int q = 0; q < 4; q++) a[q] = (short)f2bf(src[q]);
      }
      bf16x4 b;
#pragma unroll
      for (int q = 0; q < 4; q++)
        b[q] = (short)w[((size_t)t * Cout + cb * 16 + 4 * kb + q) * Cin + ci0 + m];
      acc = mfma16(a, b, acc);
    }
  }
  const int ci = ci0 + m;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const int vo = v0 + 4 * kb + i;
    if (vo < d.V) dx[(size_t)vo * Cin + ci] = acc[i];
  }
}

// Partial weight gradients: part[chunk][t][co][ci] = sum over the chunk's
// voxels of dy[v][co] x[nb(v,t)][ci].  grid (nchunk, 27, (Cout/16)*(Cin/16)),
// block 64 (one wave per 16x16 tile).
__global__ void __launch_bounds__(64)
    k_conv_bwd_weight(Dims d, int Cin, int Cout, int chunk, const float* __restrict__ dy,
                      const uint16_t* __restrict__ x, float* __restrict__ part) {
  const int lane = threadIdx.x;
  const int t = blockIdx.y;
  const int nci = Cin / 16;
  const int co0 = (blockIdx.z / nci) * 16, ci0 = (blockIdx.z % nci) * 16;
  const int m = lane & 15, kb = lane >> 4;
  const int vbeg = blockIdx.x * chunk, vend = min(d.V, vbeg + chunk);
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int vb = vbeg; vb < vend; vb += 16) {
    bf16x4 a, b;
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int v = vb + 4 * kb + q;
      float av = 0.f;
      short bv = 0;
      if (v < vend) {
        av = dy[(size_t)v * Cout + co0 + m];
        int vx, vy, vz;
        coords(d, v, vx, vy, vz);
        const int vn = neighbour(d, vx, vy, vz, t, 1);
        if (vn >= 0) bv = (short)x[(size_t)vn * Cin + ci0 + m];
      }
      a[q] = (short)f2bf(av);
      b[q] = bv;
    }
    acc = mfma16(a, b, acc);
  }
  // D[co = 4*kb+i][ci = m]
  float* out = part + ((size_t)blockIdx.x * 27 + t) * (size_t)Cout * Cin;
#pragma unroll
  for (int i = 0; i < 4; i++) out[(size_t)(co0 + 4 * kb + i) * Cin + ci0 + m] = acc[i];
}

// gw[e] = sum over chunks (fixed order) of part[chunk][e]
__global__ void k_reduce_chunks(int nchunk, size_t n, const float* __restrict__ part,
                                float* __restrict__ gw) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  float s = 0.f;
  for (int c = 0; c < nchunk; c++) s += part[(size_t)c * n + e];
  gw[e] = s;
}

// gb[co] = sum_v dy[v][co] (one block per channel, fixed order)
__global__ void __launch_bounds__(256)
    k_bias_grad(int V, int C, const float* __restrict__ dy, float* __restrict__ gb) {
  __shared__ float sh[256];
  const int co = blockIdx.x;
  float s = 0.f;
  for (int v = threadIdx.x; v < V; v += 256) s += dy[(size_t)v * C + co];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) gb[co] = sh[0];
}

// ---------------------------------------------------------------------------
// 2x2x2 average pooling (children beyond the edge are skipped) and its
// backward; nearest upsampling and its backward (sum over children).
__device__ __forceinline__ int child_count(const Dims& f, int px, int py, int pz) {
  const int nx = min(2, f.X - 2 * px), ny = min(2, f.Y - 2 * py), nz = min(2, f.Z - 2 * pz);
  return nx * ny * nz;
}

__global__ void k_pool_fwd(Dims f, Dims c, int C, const uint16_t* __restrict__ x,
                           uint16_t* __restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)c.V * C) return;
  const int ch = (int)(e % C), v = (int)(e / C);
  int px, py, pz;
  coords(c, v, px, py, pz);
  float s = 0.f;
  int n = 0;
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const int X = 2 * px + dx, Y = 2 * py + dy, Z = 2 * pz + dz;
        if (X < f.X && Y < f.Y && Z < f.Z) {
          s += bf2f(x[((size_t)X + f.X * (Y + (size_t)f.Y * Z)) * C + ch]);
          n++;
        }
      }
  y[e] = f2bf(s / (float)n);
}

__global__ void k_pool_bwd(Dims f, Dims c, int C, const float* __restrict__ dy,
                           float* __restrict__ dx) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)f.V * C) return;
  const int ch = (int)(e % C), v = (int)(e / C);
  int X, Y, Z;
  coords(f, v, X, Y, Z);
  const int px = X / 2, py = Y / 2, pz = Z / 2;
  dx[e] = dy[((size_t)px + c.X * (py + (size_t)c.Y * pz)) * C + ch] /
          (float)child_count(f, px, py, pz);
}

__global__ void k_up_fwd(Dims c, Dims f, int C, const uint16_t* __restrict__ x,
                         uint16_t* __restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)f.V * C) return;
  const int ch = (int)(e % C), v = (int)(e / C);
  int X, Y, Z;
  coords(f, v, X, Y, Z);
  y[e] = x[((size_t)(X / 2) + c.X * ((Y / 2) + (size_t)c.Y * (Z / 2))) * C + ch];
}

__global__ void k_up_bwd(Dims c, Dims f, int C, const float* __restrict__ dy,
                         float* __restrict__ dx) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (size_t)c.V * C) return;
  const int ch = (int)(e % C), v = (int)(e / C);
  int px, py, pz;
  coords(c, v, px, py, pz);
  float s = 0.f;
  for (int dz = 0; dz < 2; dz++)
    for (int dy_ = 0; dy_ < 2; dy_++)
      for (int dx_ = 0; dx_ < 2; dx_++) {
        const int X = 2 * px + dx_, Y = 2 * py + dy_, Z = 2 * pz + dz;
        if (X < f.X && Y < f.Y && Z < f.Z)
          s += dy[((size_t)X + f.X * (Y + (size_t)f.Y * Z)) * C + ch];
      }
  dx[e] = s;
}

// y = a + b (BF16), the skip connection
__global__ void k_add_bf16(size_t n, const uint16_t* __restrict__ a,
                           const uint16_t* __restrict__ b, uint16_t* __restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(bf2f(a[e]) + bf2f(b[e]));
}

// ReLU backward through the post-activation y, summing two incoming
// gradients: out = (g + g2) * (y > 0)   (g2 may be null; out may alias g)
__global__ void k_relu_bwd(size_t n, const uint16_t* __restrict__ y, const float* g,
                           const float* g2, float* out) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float v = g[e] + (g2 ? g2[e] : 0.f);
  out[e] = bf2f(y[e]) > 0.f ? v : 0.f;
}

// loss partials and gradient of L = mean_v (out[v][0] - t[v])^2 (out FP32,
// C channels; only channel 0 is the prediction)
__global__ void __launch_bounds__(256)
    k_mse(int V, int C, const float* __restrict__ out, const float* __restrict__ tgt,
          float* __restrict__ g, float* __restrict__ part) {
  __shared__ float sh[256];
  const int v = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (v < V) {
    const float dlt = out[(size_t)v * C] - tgt[v];
    l = dlt * dlt;
    g[(size_t)v * C] = 2.f * dlt / (float)V;
    for (int c = 1; c < C; c++) g[(size_t)v * C + c] = 0.f;
  }
  sh[threadIdx.x] = l;
  __syncthreads();
  for (int off = 128; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) sh[threadIdx.x] += sh[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_sum(int n, const float* __restrict__ part, float* __restrict__ out, float scale) {
  if (blockIdx.x || threadIdx.x) return;
  float s = 0.f;
  for (int i = 0; i < n; i++) s += part[i];
  *out = s * scale;
}

// Adam (bias-corrected) on FP32 weights; refreshes the BF16 copy when wb.
__global__ void k_adam(size_t n, float* __restrict__ w, const float* __restrict__ g,
                       float* __restrict__ m1, float* __restrict__ m2, float lr, float b1,
                       float b2, float eps, float c1, float c2, uint16_t* __restrict__ wb) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const float gg = g[e];
  const float a = b1 * m1[e] + (1.f - b1) * gg;
  const float b = b2 * m2[e] + (1.f - b2) * gg * gg;
  m1[e] = a;
  m2[e] = b;
  const float upd = lr * (a / c1) / (sqrtf(b / c2) + eps);
  const float wn = w[e] - upd;
  w[e] = wn;
  if (wb) wb[e] = f2bf(wn);
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) y[e] = f2bf(x[e]);
}

__global__ void k_fill_u16(size_t n, uint16_t* p, uint16_t v) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) p[e] = v;
}

// Input features on the interior grid (a,b,c) from br0(nt,np) (Fortran
// order, interior at j+1,k+1) and a per-shell factor table:
//   ch0 = br0 * (r0/r)^2 / bscale,  ch1 = br0 / bscale,  ch2 = (r-r0)/(r1-r0)
__global__ void k_features(Dims d, int C, int nt, const double* __restrict__ br0,
                           const float* __restrict__ decay, const float* __restrict__ rpos,
                           float inv_bscale, int roll, int flip, uint16_t* __restrict__ x0) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= d.V) return;
  int i, j, k;
  coords(d, v, i, j, k);
  // the sample seen rotated by `roll` interior phi points and, with flip,
  // reflected in theta (j -> Y-1-j); both 0 outside training
  const int js = flip ? d.Y - 1 - j : j;
  const int ks = k + roll < d.Z ? k + roll : k + roll - d.Z;
  const float b = (float)br0[(size_t)(js + 1) + (size_t)nt * (ks + 1)] * inv_bscale;
  uint16_t* o = x0 + (size_t)v * C;
  o[0] = f2bf(b * decay[i]);
  o[1] = f2bf(b);
  o[2] = f2bf(rpos[i]);
  for (int c = 3; c < C; c++) o[c] = 0;
}

// Training target of a rotated/reflected sample: dst(i,j,k) = sign *
// src(i, flip ? Y-1-j : j, (k + roll) mod Z), the same map as k_features.
__global__ void k_gather_target(Dims d, int roll, int flip, float sign,
                                const float* __restrict__ src, float* __restrict__ dst) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= d.V) return;
  int i, j, k;
  coords(d, v, i, j, k);
  const int js = flip ? d.Y - 1 - j : j;
  const int ks = k + roll < d.Z ? k + roll : k + roll - d.Z;
  dst[v] = sign * src[i + d.X * (js + d.Y * ks)];
}

// x (FP64, packed) = scale * out[v][0]
__global__ void k_head_to_x(int V, int C, const float* __restrict__ out, double scale,
                            double* __restrict__ xo) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < V) xo[v] = scale * (double)out[(size_t)v * C];
}

// Trilinear prolongation of a padded coarse field phi_c(nrc,ntc,npc) to the
// fine interior points: per-axis index/weight tables (fine interior -> lower
// coarse padded index and weight of the upper neighbour), phi periodic.
__global__ void k_prolong(int af, int bf, int cf, int nrc, int ntc, int npc,
                          const double* __restrict__ phic, const int* __restrict__ ir,
                          const double* __restrict__ wr, const int* __restrict__ it,
                          const double* __restrict__ wt, const int* __restrict__ ip,
                          const double* __restrict__ wp, double* __restrict__ xf) {
  const size_t n = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t N = (size_t)af * bf * cf;
  if (n >= N) return;
  const int i = (int)(n % af);
  const int j = (int)((n / af) % bf);
  const int k = (int)(n / ((size_t)af * bf));
  const int i0 = ir[i], j0 = it[j], k0 = ip[k];
  const double a1 = wr[i], b1 = wt[j], c1 = wp[k];
  double s = 0.0;
  for (int dk = 0; dk < 2; dk++)
    for (int dj = 0; dj < 2; dj++)
      for (int di = 0; di < 2; di++) {
        const double w = (di ? a1 : 1.0 - a1) * (dj ? b1 : 1.0 - b1) * (dk ? c1 : 1.0 - c1);
        const size_t q = (size_t)(i0 + di) + (size_t)nrc * ((size_t)(j0 + dj) +
                                                             (size_t)ntc * (size_t)(k0 + dk));
        s += w * phic[q];
      }
  xf[n] = s;
}

// A 16x16x16 MFMA on given matrices (A row-major MxK, B row-major KxN, BF16
// bits) -> D (FP32, row-major): used to check the operand layout.
__global__ void k_mfma_check(const uint16_t* __restrict__ A, const uint16_t* __restrict__ B,
                             float* __restrict__ D) {
  const int l = threadIdx.x;
  bf16x4 a, b;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    a[q] = (short)A[(l % 16) * 16 + 4 * (l / 16) + q];
    b[q] = (short)B[(4 * (l / 16) + q) * 16 + (l % 16)];
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = mfma16(a, b, acc);
#pragma unroll
  for (int i = 0; i < 4; i++) D[(4 * (l / 16) + i) * 16 + (l % 16)] = acc[i];
}

// ---------------------------------------------------------------------------
// host launchers
static inline unsigned nblk(size_t n, int b) { return (unsigned)((n + b - 1) / b); }

void conv_fwd(const Dims3& d3, int Cin, int Cout, const uint16_t* x, const uint16_t* w,
              const float* bias, uint16_t* y, float* yf, bool relu, hipStream_t s) {
  Dims d{d3.X, d3.Y, d3.Z, d3.X * d3.Y * d3.Z};
  dim3 grid(nblk(d.V, 64), Cout / 16);
  k_conv_fwd<<<grid, 256, 0, s>>>(d, Cin, Cout, x, w, bias, y, yf, relu ? 1 : 0);
}

void conv_bwd(const Dims3& d3, int Cin, int Cout, const float* dy, const uint16_t* x,
              const uint16_t* w, float* dx, float* gw, float* gb, float* part, int nchunk,
              hipStream_t s) {
  Dims d{d3.X, d3.Y, d3.Z, d3.X * d3.Y * d3.Z};
  if (dx) {
    dim3 grid(nblk(d.V, 64), Cin / 16);
    k_conv_bwd_data<<<grid, 256, 0, s>>>(d, Cin, Cout, dy, w, dx);
  }
  int chunk = (d.V + nchunk - 1) / nchunk;
  chunk = (chunk + 15) / 16 * 16;
  const int nc = (d.V + chunk - 1) / chunk;
  dim3 gridw(nc, 27, (Cout / 16) * (Cin / 16));
  k_conv_bwd_weight<<<gridw, 64, 0, s>>>(d, Cin, Cout, chunk, dy, x, part);
  const size_t nw = (size_t)27 * Cout * Cin;
  k_reduce_chunks<<<nblk(nw, 256), 256, 0, s>>>(nc, nw, part, gw);
  k_bias_grad<<<Cout, 256, 0, s>>>(d.V, Cout, dy, gb);
}

void pool_fwd(const Dims3& f3, const Dims3& c3, int C, const uint16_t* x, uint16_t* y,
              hipStream_t s) {
  Dims f{f3.X, f3.Y, f3.Z, f3.X * f3.Y * f3.Z}, c{c3.X, c3.Y, c3.Z, c3.X * c3.Y * c3.Z};
  k_pool_fwd<<<nblk((size_t)c.V * C, 256), 256, 0, s>>>(f, c, C, x, y);
}
void pool_bwd(const Dims3& f3, const Dims3& c3, int C, const float* dy, float* dx,
              hipStream_t s) {
  Dims f{f3.X, f3.Y, f3.Z, f3.X * f3.Y * f3.Z}, c{c3.X, c3.Y, c3.Z, c3.X * c3.Y * c3.Z};
  k_pool_bwd<<<nblk((size_t)f.V * C, 256), 256, 0, s>>>(f, c, C, dy, dx);
}
void up_fwd(const Dims3& c3, const Dims3& f3, int C, const uint16_t* x, uint16_t* y,
            hipStream_t s) {
  Dims f{f3.X, f3.Y, f3.Z, f3.X * f3.Y * f3.Z}, c{c3.X, c3.Y, c3.Z, c3.X * c3.Y * c3.Z};
  k_up_fwd<<<nblk((size_t)f.V * C, 256), 256, 0, s>>>(c, f, C, x, y);
}
void up_bwd(const Dims3& c3, const Dims3& f3, int C, const float* dy, float* dx,
            hipStream_t s) {
  Dims f{f3.X, f3.Y, f3.Z, f3.X * f3.Y * f3.Z}, c{c3.X, c3.Y, c3.Z, c3.X * c3.Y * c3.Z};
  k_up_bwd<<<nblk((size_t)c.V * C, 256), 256, 0, s>>>(c, f, C, dy, dx);
}
void add_bf16(size_t n, const uint16_t* a, const uint16_t* b, uint16_t* y, hipStream_t s) {
  k_add_bf16<<<nblk(n, 256), 256, 0, s>>>(n, a, b, y);
}
void relu_bwd(size_t n, const uint16_t* y, const float* g, const float* g2, float* out,
              hipStream_t s) {
  k_relu_bwd<<<nblk(n, 256), 256, 0, s>>>(n, y, g, g2, out);
}
void mse(int V, int C, const float* out, const float* tgt, float* g, float* part, float* loss,
         hipStream_t s) {
  const int nb = (int)nblk(V, 256);
  k_mse<<<nb, 256, 0, s>>>(V, C, out, tgt, g, part);
  k_sum<<<1, 1, 0, s>>>(nb, part, loss, 1.0f / (float)V);
}
void adam(size_t n, float* w, const float* g, float* m1, float* m2, float lr, float b1,
          float b2, float eps, int step, uint16_t* wb, hipStream_t s) {
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  k_adam<<<nblk(n, 256), 256, 0, s>>>(n, w, g, m1, m2, lr, b1, b2, eps, c1, c2, wb);
}
void to_bf16(size_t n, const float* x, uint16_t* y, hipStream_t s) {
  k_to_bf16<<<nblk(n, 256), 256, 0, s>>>(n, x, y);
}
void fill_u16(size_t n, uint16_t* p, uint16_t v, hipStream_t s) {
  k_fill_u16<<<nblk(n, 256), 256, 0, s>>>(n, p, v);
}
void features(const Dims3& d3, int C, int nt, const double* br0, const float* decay,
              const float* rpos, float inv_bscale, int roll, int flip, uint16_t* x0,
              hipStream_t s) {
  Dims d{d3.X, d3.Y, d3.Z, d3.X * d3.Y * d3.Z};
  k_features<<<nblk(d.V, 256), 256, 0, s>>>(d, C, nt, br0, decay, rpos, inv_bscale, roll, flip,
                                             x0);
}
void gather_target(const Dims3& d3, int roll, int flip, float sign, const float* src, float* dst,
                   hipStream_t s) {
  Dims d{d3.X, d3.Y, d3.Z, d3.X * d3.Y * d3.Z};
  k_gather_target<<<nblk(d.V, 256), 256, 0, s>>>(d, roll, flip, sign, src, dst);
}
void head_to_x(int V, int C, const float* out, double scale, double* xo, hipStream_t s) {
  k_head_to_x<<<nblk(V, 256), 256, 0, s>>>(V, C, out, scale, xo);
}
void prolong(int af, int bf, int cf, int nrc, int ntc, int npc, const double* phic,
             const int* ir, const double* wr, const int* it, const double* wt, const int* ip,
             const double* wp, double* xf, hipStream_t s) {
  const size_t N = (size_t)af * bf * cf;
  k_prolong<<<nblk(N, 256), 256, 0, s>>>(af, bf, cf, nrc, ntc, npc, phic, ir, wr, it, wt, ip,
                                          wp, xf);
}
void mfma_check(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s) {
  k_mfma_check<<<1, 64, 0, s>>>(A, B, D);
}

}  // namespace nn
}  // namespace p3d
