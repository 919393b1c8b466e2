// gfx950 kernels of the PCG solve (matrix-free 7-point operator, fused CG
// vector updates with deterministic FP64 reductions, D-ILU preconditioner,
// boundary/pole/seam handling, energy diagnostic).
//
// Reference routines restated (predsci/POT3D src/pot3d.F90):
//   delsq :5337-5374 + set_boundary_points :5376-5487 + seam_hhh :5689-5760
//     -> k_stencil (ghost values folded into the neighbour reads)
//   cgsolve :4231-4369 and ernorm :4371-4485 -> k_cg_update, k_p_update,
//     epilogues of the fused reductions (cg_epilogue)
//   cgdot/global_sum :5571-5644 -> block_finish (FP64, fixed order)
//   prec_inv :5152-5225 (ifprec=1 diagonal, ifprec=2 ILU0)
//   RHS trick :3941-3950 -> k_rhs (closed form)
//   post-solve phi :3966-3969 -> k_phi_*
//   getb + magnetic_energy :6187-6321 -> k_energy, k_getb
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "p3d_kernels.hpp"

namespace p3d {

Launch plan_stream(uint32_t N, int max_blocks) {
  int G = (int)((N + kBlock - 1) / kBlock);
  if (G > max_blocks) G = max_blocks;
  if (G < 8) G = 8;
  G = (G + 7) / 8 * 8;
  uint32_t ch = (N + (uint32_t)G - 1) / (uint32_t)G;
  ch = (ch + kBlock - 1) / kBlock * kBlock;
  Launch L;
  L.G = G;
  L.CH = (int)ch;
  return L;
}

// Workgroups are placed round-robin over the 8 XCDs; map them so that each
// XCD (and its L2) streams one contiguous eighth of the domain, keeping the
// +-a and +-ab neighbour planes of a chunk in the same L2.
__device__ __forceinline__ int logical_block(int G) {
  const int b = blockIdx.x;
  return (b & 7) * (G >> 3) + (b >> 3);
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Block reduction of NV doubles, in two levels so that no single counter
// takes an atomic from every workgroup (thousands of same-address atomics
// serialise at the L2 and cost hundreds of microseconds per reduction).
// Every block stores its partial and counts in at its group's counter
// (kRedGroup consecutive blocks); the last block of a group sums the
// group's partials in order, stores the group partial and counts in at the
// slot's global counter; the last group sums the group partials in order,
// runs the epilogue and re-arms the counters.  The result is deterministic
// for a given launch shape.
//   counter: the slot's kRedCounters counters ([0] global, [1 + group]);
//   partial: NV*G block partials, then NV*ngroups group partials.
// Counts a block in at a reduction counter after its partial sums (agent-
// scope atomic stores, coherent by themselves) have completed.  A release
// ordering on the counter would make the compiler write the whole L2 back
// (buffer_wbl2) in every block, including the kernel's own output lines;
// waiting for the block's stores to complete is all the protocol needs.
__device__ __forceinline__ unsigned arrive(unsigned* counter) {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0);
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  return __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int BS, int NV>
__device__ void block_finish_bs(double (&v)[NV], double* partial, unsigned* counter, int G, int Lb,
                             CgState* st, Epi epi, double* out) {
  __shared__ double sh[NV][BS / 64];
  __shared__ int amlast;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int grp = Lb / kRedGroup, g0 = grp * kRedGroup;
  const int gn = min(kRedGroup, G - g0), ngrp = (G + kRedGroup - 1) / kRedGroup;
  double* gpart = partial + (size_t)NV * G;
#pragma unroll
  for (int q = 0; q < NV; q++) {
    const double t = wave_sum(v[q]);
    if (lane == 0) sh[q][w] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = 0.0;
      for (int ww = 0; ww < BS / 64; ww++) s += sh[q][ww];
      __hip_atomic_store(&partial[q * G + Lb], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    const unsigned prev = arrive(counter + 1 + grp);
    amlast = (prev == (unsigned)gn - 1);
  }
  __syncthreads();
  if (!amlast) return;
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  // last block of the group: the group partial
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0.0;
    for (int m = threadIdx.x; m < gn; m += BS)
      s += __hip_atomic_load(&partial[q * G + g0 + m], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    if (lane == 0) sh[q][w] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      double s = 0.0;
      for (int ww = 0; ww < BS / 64; ww++) s += sh[q][ww];
      __hip_atomic_store(&gpart[q * ngrp + grp], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __hip_atomic_store(counter + 1 + grp, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const unsigned prev = arrive(counter);
    amlast = (prev == (unsigned)ngrp - 1);
  }
  __syncthreads();
  if (!amlast || threadIdx.x != 0) return;
  __atomic_thread_fence(__ATOMIC_ACQUIRE);
  // last group: the total, in group order
  double tot[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) {
    double s = 0.0;
    for (int gi = 0; gi < ngrp; gi++)
      s += __hip_atomic_load(&gpart[q * ngrp + gi], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    tot[q] = s;
  }
  if (NV == 3) {
    if (out) {
      out[0] = tot[0];
      out[1] = tot[NV > 1 ? 1 : 0];
      out[2] = tot[NV > 2 ? 2 : 0];
    }
  } else {
    cg_epilogue(st, epi, tot[0], NV > 1 ? tot[NV > 1 ? 1 : 0] : 0.0, out);
  }
  __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int NV>
__device__ __forceinline__ void block_finish(double (&v)[NV], double* partial, unsigned* counter,
                                             int G, int Lb, CgState* st, Epi epi, double* out) {
  block_finish_bs<kBlock, NV>(v, partial, counter, G, Lb, st, epi, out);
}

// ---------------------------------------------------------------------------
// Pole averages: sum_k x(i,pole,k) * dph(k)/(2 pi)  (set_boundary_points
// :5442-5465), as kPoleSegs ordered partial sums per (pole, i).
template <typename T>
__global__ void __launch_bounds__(64) k_pole_sums(Metric m, Grid3 g, const T* __restrict__ x,
                                                 double* __restrict__ part) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= g.a) return;
  const int pole = blockIdx.y, seg = blockIdx.z;
  const int kper = (g.c + kPoleSegs - 1) / kPoleSegs;
  const int kl = seg * kper;
  const int kh = min(g.c, kl + kper);
  const size_t base = (size_t)i + (pole ? (size_t)(g.b - 1) * g.a : 0);
  double acc = 0.0;
  // the planes' loads in batches of 16, issued before the batch's sums
  // (which stay in plane order): one memory round trip per batch
  constexpr int kB = 16;
  for (int k = kl; k < kh; k += kB) {
    T v[kB];
    double w[kB];
#pragma unroll
    for (int u = 0; u < kB; u++) {
      const int kk = min(k + u, kh - 1);
      v[u] = x[base + (size_t)g.ab * kk];
      w[u] = m.PW[kk];
    }
#pragma unroll
    for (int u = 0; u < kB; u++)
      if (k + u < kh) acc += (double)v[u] * w[u];
  }
  part[(pole * kPoleSegs + seg) * g.a + i] = acc;
}

template <typename T>
void launch_pole_sums(const Metric& m, const Grid3& g, const T* x, double* part, hipStream_t s) {
  dim3 grid((g.a + 63) / 64, 2, kPoleSegs);
  k_pole_sums<T><<<grid, 64, 0, s>>>(m, g, x, part);
}

__device__ __forceinline__ double pole_sum(const double* part, int pole, int i, int a) {
  double s = 0.0;
#pragma unroll
  for (int q = 0; q < kPoleSegs; q++) s += part[(pole * kPoleSegs + q) * a + i];
  return s;
}

// ---------------------------------------------------------------------------
// Operator: y = A x (mode 0) or y = b - A x (mode 1), homogeneous ghosts:
//   r0 ghost  x(1)  = x(2)                    (vmask=0, :5414-5417)
//   r1 ghost  x(nr) = pm_r1 x(nrm1)           (:5424-5427)
//   poles     x(i,1,k) = 2<x(i,2,.)> - x(i,2,k) (and nt; :5475-5483)
//   phi seam  periodic wrap                   (:5746-5758)
// The coefficients are recomputed from the 1-D factor tables (no 56 B per
// unknown coefficient stream).
//
// Coefficients a1..a7 of delsq (:5337-5374) at one point from its column
// factors and its plane's factors, a4 = -(sum of the others).  Contraction
// is off: every operator kernel forms them with the same roundings.
__device__ __forceinline__ void op_coefs(double DRH, double CRL, double CRH, double SD,
                                         double CTL, double CTH, double P1, double dph,
                                         double dpl, double dpu, double (&a)[7]) {
#pragma clang fp contract(off)
  const double a1 = -(P1 * dpl);
  const double a7 = -(P1 * dpu);
  const double p2 = DRH * dph;
  const double a2 = -(p2 * CTL);
  const double a6 = -(p2 * CTH);
  const double p3 = SD * dph;
  const double a3 = -(p3 * CRL);
  const double a5 = -(p3 * CRH);
  a[0] = a1;
  a[1] = a2;
  a[2] = a3;
  a[3] = -(a1 + a2 + a3 + a5 + a6 + a7);
  a[4] = a5;
  a[5] = a6;
  a[6] = a7;
}

// The operator's value at one point from x and its six neighbours (xw/xe in
// i, xs/xn in j, xkm/xkp in k; the ghost values replace the boundary
// neighbours here), y or, with sub_b, b - y.  The sum is an explicit FMA
// chain, so the result is the same bits in every operator kernel.
//   FP64: the reference's delsq form, sum a_nb x_nb + a4 x.
//   FP32 storage: the same operator with neighbour differences, formed and
//   summed in double, only the result rounded to float.  For a smooth x,
//   sum a_nb x_nb + a4 x cancels to far below its terms (about 1/h^2 for a
//   Laplacian), and so does the sum of the differences (about 1/h); in
//   float either loses that factor times eps32, which stalls the FP32
//   Krylov recurrence on smooth (open-boundary, monopole-dominated)
//   solutions.
template <typename T>
__device__ __forceinline__ T op_point(const double (&a)[7], T xkm, T xs, T xw, T xk, T xe, T xn,
                                      T xkp, bool im, bool ip, bool jm, bool jp, double pm_r1,
                                      double s0d, double s1d, bool sub_b, T bval) {
#pragma clang fp contract(off)
  if constexpr (sizeof(T) == 8) {
    const T xim = im ? xw : xk;
    const T xip = ip ? xe : (T)pm_r1 * xk;
    const T xjm = jm ? xs : (T)s0d - xk;
    const T xjp = jp ? xn : (T)s1d - xk;
    T y = a[0] * xkm;
    y = fma(a[1], xjm, y);
    y = fma(a[2], xim, y);
    y = fma(a[3], xk, y);
    y = fma(a[4], xip, y);
    y = fma(a[5], xjp, y);
    y = fma(a[6], xkp, y);
    return sub_b ? bval - y : y;
  } else {
    const double c = (double)xk;
    const double dim = im ? (double)xw - c : 0.0;
    const double dip = ip ? (double)xe - c : (pm_r1 - 1.0) * c;
    const double djm = jm ? (double)xs - c : s0d - 2.0 * c;
    const double djp = jp ? (double)xn - c : s1d - 2.0 * c;
    double yd = a[0] * ((double)xkm - c);
    yd = fma(a[1], djm, yd);
    yd = fma(a[2], dim, yd);
    yd = fma(a[4], dip, yd);
    yd = fma(a[5], djp, yd);
    yd = fma(a[6], (double)xkp - c, yd);
    if (sub_b) yd = (double)bval - yd;
    return (T)yd;
  }
}

// 2.5-D form of the same operator: one thread per (i,j) column marching
// through a segment of phi planes.  x(k-1), x(k), x(k+1) stay in registers,
// the column's metric factors are loaded once, and the per-plane factors
// DPH/DPL/DPU are wave-uniform, so each point costs one HBM read of x, four
// cached neighbour reads and one write.  Partial sums of the fused dot are
// reduced per block with the same last-block epilogue.
constexpr int kSegPlanes = 16;

// Column blocks of the operator march: kBlock consecutive columns, or, with
// a tile-major output, kBlock/64 patches of 8x8 columns.
__host__ __device__ inline int stencil_col_blocks(const Grid3& g, const TileMap& tm) {
  if (tm.on) return (tm.TNI * tm.TNJ + kBlock / 64 - 1) / (kBlock / 64);
  return (int)((g.ab + kBlock - 1) / kBlock);
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_stencil_march(Metric m, Grid3 g, int nseg, const T* __restrict__ x, T* __restrict__ y,
                    const T* __restrict__ b, int mode, const double* __restrict__ pole_part,
                    CgState* st, double* partial, unsigned* counter, Epi epi, double* out,
                    int guard, int want_dot, TileMap tm) {
  if (guard && st_done(st)) return;
  const int G = gridDim.x;
  const int Lb = logical_block(G);
  const int ncol_blocks = stencil_col_blocks(g, tm);
  const int cb = Lb % ncol_blocks, seg = Lb / ncol_blocks;
  const int kper = (g.c + nseg - 1) / nseg;
  const int k0 = seg * kper, k1 = min(g.c, k0 + kper);
  // the columns of a wave: 64 consecutive mcol = i + a j, or, with the
  // tile-major output, one 8x8 (i, j) patch, so that a wave writes whole
  // 256-byte tile planes
  int i, j;
  bool col_ok;
  if (tm.on) {
    const int t = cb * (kBlock / 64) + (int)(threadIdx.x >> 6);
    const int lane = (int)(threadIdx.x & 63);
    i = (t % tm.TNI) * 8 + (lane & 7);
    j = (t / tm.TNI) * 8 + (lane >> 3);
    col_ok = t < tm.TNI * tm.TNJ && i < g.a && j < g.b;
  } else {
    const uint32_t mc = (uint32_t)cb * kBlock + threadIdx.x;
    const uint32_t jq = fdiv(mc, g.da);
    i = (int)(mc - jq * (uint32_t)g.a);
    j = (int)jq;
    col_ok = mc < g.ab;
  }
  double acc = 0.0;
  if (col_ok && k0 < k1) {
    const uint32_t mcol = (uint32_t)i + (uint32_t)g.a * (uint32_t)j;
    const double DRH = m.DRH[i], CRL = m.CRL[i], CRH = m.CRH[i];
    const double SD = m.SD[j], CTL = m.CTL[j], CTH = m.CTH[j];
    const double P1 = DRH * m.CP[j];
    const bool im = i > 0, ip = i < g.a - 1, jm = j > 0, jp = j < g.b - 1;
    const double s0d = jm ? 0.0 : 2.0 * pole_sum(pole_part, 0, i, g.a);
    const double s1d = jp ? 0.0 : 2.0 * pole_sum(pole_part, 1, i, g.a);
    const uint32_t ab = g.ab, a = (uint32_t)g.a;
    const T* xc = x + mcol;
    T xkm = xc[(size_t)ab * (k0 > 0 ? k0 - 1 : g.c - 1)];
    T xk = xc[(size_t)ab * k0];
    // blocks of U planes: every load of the block is issued before the
    // arithmetic, so each thread keeps U HBM reads in flight
    constexpr int U = sizeof(T) == 8 ? 4 : 8;
    for (int kb = k0; kb < k1; kb += U) {
      T xn[U], xw[U], xe[U], xs[U], xnn[U], bv[U];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = kb + u;
        xn[u] = xw[u] = xe[u] = xs[u] = xnn[u] = bv[u] = (T)0;
        if (k < k1) {
          const uint32_t n = mcol + ab * (uint32_t)k;
          xn[u] = xc[(size_t)ab * (k < g.c - 1 ? k + 1 : 0)];
          if (im) xw[u] = x[n - 1];
          if (ip) xe[u] = x[n + 1];
          if (jm) xs[u] = x[n - a];
          if (jp) xnn[u] = x[n + a];
          if (mode == 1) bv[u] = b[n];
        }
      }
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int k = kb + u;
        if (k < k1) {
          const uint32_t n = mcol + ab * (uint32_t)k;
          const T xkp = xn[u];
          double a[7];
          op_coefs(DRH, CRL, CRH, SD, CTL, CTH, P1, m.DPH[k], m.DPL[k], m.DPU[k], a);
          const T yv = op_point<T>(a, xkm, xs[u], xw[u], xk, xe[u], xnn[u], xkp, im, ip, jm, jp,
                                   m.pm_r1, s0d, s1d, mode == 1, bv[u]);
          y[tmap(tm, g, n)] = yv;
          acc += mode == 0 ? (double)xk * (double)yv : (double)yv * (double)yv;
          xkm = xk;
          xk = xkp;
        }
      }
    }
  }
  if (!want_dot) return;
  double v[1] = {acc};
  block_finish<1>(v, partial, counter, G, Lb, st, epi, out);
}

// ---------------------------------------------------------------------------
// Row-blocked operator march (Launch::stencil = 1): the operator and the
// per-point arithmetic of k_stencil_march, with the work laid out so that
// a point costs one coalesced row load of x instead of five neighbour loads.
// Each wave owns 64 consecutive i (one per lane) by kRows theta rows and
// marches a segment of phi planes.  Lanes 1..62 produce output; lanes 0 and
// 63 only load the i-1 / i+1 neighbours of the wave's edge columns.  Per
// plane a thread loads its kRows rows plus the halo rows j0-1 and j0+kRows
// (coalesced 64-lane rows); the i neighbours come from the adjacent lanes
// by DPP wave shifts, the j neighbours from the thread's own rows and the k
// neighbours from the march, which keeps planes k-1..k+1 in registers and
// the loads of plane k+2 in flight.  A workgroup's four waves take four
// consecutive row blocks, so their halo rows are each other's rows.
constexpr int kRows = 8;
constexpr int kRowLanes = 62;
constexpr int kWaveShr1 = 0x138;  // DPP: lane l reads lane l-1
constexpr int kWaveShl1 = 0x130;  // DPP: lane l reads lane l+1

template <int CTRL>
__device__ __forceinline__ float lane_shift(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xf, 0xf, true));
}
template <int CTRL>
__device__ __forceinline__ double lane_shift(double v) {
  const long long u = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(u & 0xffffffffLL), CTRL, 0xf, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(u >> 32), CTRL, 0xf, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__host__ __device__ inline void stencil_rows_shape(const Grid3& g, int& nwi, int& nwg) {
  nwi = (g.a + kRowLanes - 1) / kRowLanes;
  nwg = ((g.b + kRows - 1) / kRows + 3) / 4;
}

// The kernel the automatic choice takes: the row-blocked march for the
// FP32 operator when its 62-lane waves cover the r extent to at least 90%
// (measured faster there: 0.61 vs 0.71 ms at 299 x 599 x 1199; slower at
// 149 x 299 x 599, where a third of the third wave idles), else the column
// march (faster for the FP64 operator at both sizes).
inline bool stencil_rows_chosen(const Grid3& g, size_t elem_bytes) {
  if (elem_bytes != 4) return false;
  const int nwi = (g.a + kRowLanes - 1) / kRowLanes;
  return 10 * g.a >= 9 * nwi * kRowLanes;
}

template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_stencil_rows(Metric m, Grid3 g, int nseg, const T* __restrict__ x, T* __restrict__ y,
                   const T* __restrict__ b, int mode, const double* __restrict__ pole_part,
                   CgState* st, double* partial, unsigned* counter, Epi epi, double* out,
                   int guard, int want_dot, TileMap tm) {
  if (guard && st_done(st)) return;
  const int G = gridDim.x;
  const int Lb = logical_block(G);
  int nwi, nwg;
  stencil_rows_shape(g, nwi, nwg);
  // row groups fastest (neighbouring groups share halo rows in L2), then
  // i blocks, then phi segments
  const int seg = Lb / (nwi * nwg), rem = Lb - seg * (nwi * nwg);
  const int wi = rem / nwg, wg = rem - wi * nwg;
  const int kper = (g.c + nseg - 1) / nseg;
  const int k0 = seg * kper, k1 = min(g.c, k0 + kper);
  // the wave index as a wave-uniform value, so that the row factors and
  // per-plane factors below are scalar loads
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  const int lane = (int)(threadIdx.x & 63);
  const int j0 = (wg * 4 + wave) * kRows;
  double acc = 0.0;
  // wave-uniform: every lane of a working wave runs the loads and the DPP
  if (seg < nseg && j0 < g.b && k0 < k1) {
    const int i = wi * kRowLanes - 1 + lane;
    const bool out_lane = lane >= 1 && lane <= kRowLanes && i < g.a;
    const int il = min(max(i, 0), g.a - 1);
    const double DRH = m.DRH[il], CRL = m.CRL[il], CRH = m.CRH[il];
    const bool im = i > 0, ip = i < g.a - 1;
    const double s0d = j0 == 0 ? 2.0 * pole_sum(pole_part, 0, il, g.a) : 0.0;
    const double s1d = j0 + kRows >= g.b ? 2.0 * pole_sum(pole_part, 1, il, g.a) : 0.0;
    const uint32_t ab = g.ab, a_ = (uint32_t)g.a;
    // x offsets of the thread's rows j0-1 .. j0+kRows (clamped into the grid;
    // clamped rows are never used)
    uint32_t ro[kRows + 2];
#pragma unroll
    for (int r = 0; r < kRows + 2; r++) {
      const int j = min(max(j0 - 1 + r, 0), g.b - 1);
      ro[r] = (uint32_t)il + a_ * (uint32_t)j;
    }
    auto load = [&](T (&buf)[kRows + 2], int k) {
      const int kw = ((k % g.c) + g.c) % g.c;  // phi periodic (lookahead may pass c)
      const T* xp = x + (size_t)ab * (uint32_t)kw;
#pragma unroll
      for (int r = 0; r < kRows + 2; r++) buf[r] = xp[ro[r]];
    };
    auto plane = [&](const T (&pv)[kRows + 2], const T (&cv)[kRows + 2],
                     const T (&nv)[kRows + 2], int k) {
      const double dph = m.DPH[k], dpl = m.DPL[k], dpu = m.DPU[k];
#pragma unroll
      for (int r = 0; r < kRows; r++) {
        const int j = j0 + r;
        if (j < g.b) {
          const T xk = cv[r + 1];
          const T xl = lane_shift<kWaveShr1>(xk);
          const T xr = lane_shift<kWaveShl1>(xk);
          const bool jm = j > 0, jp = j < g.b - 1;
          const double SD = m.SD[j], CTL = m.CTL[j], CTH = m.CTH[j];
          const double P1 = DRH * m.CP[j];
          const T xkm = pv[r + 1], xkp = nv[r + 1];
          double a[7];
          op_coefs(DRH, CRL, CRH, SD, CTL, CTH, P1, dph, dpl, dpu, a);
          const uint32_t n = (uint32_t)il + a_ * (uint32_t)j + ab * (uint32_t)k;
          const T bval = (mode == 1 && out_lane) ? b[n] : (T)0;
          const T yv = op_point<T>(a, xkm, cv[r], xl, xk, xr, cv[r + 2], xkp, im, ip, jm, jp,
                                   m.pm_r1, s0d, s1d, mode == 1, bval);
          if (out_lane) {
            y[tmap(tm, g, n)] = yv;
            acc += mode == 0 ? (double)xk * (double)yv : (double)yv * (double)yv;
          }
        }
      }
    };
    // four plane buffers in rotation: planes k-1, k, k+1 and k+2 (loading)
    T B0[kRows + 2], B1[kRows + 2], B2[kRows + 2], B3[kRows + 2];
    load(B0, k0 - 1);
    load(B1, k0);
    load(B2, k0 + 1);
    int k = k0;
    for (; k + 4 <= k1; k += 4) {
      load(B3, k + 2);
      plane(B0, B1, B2, k);
      load(B0, k + 3);
      plane(B1, B2, B3, k + 1);
      load(B1, k + 4);
      plane(B2, B3, B0, k + 2);
      load(B2, k + 5);
      plane(B3, B0, B1, k + 3);
    }
    // the remaining 0..3 planes (buffers B0, B1, B2 hold planes k-1, k, k+1)
    if (k < k1) {
      load(B3, k + 2);
      plane(B0, B1, B2, k);
    }
    if (k + 1 < k1) {
      load(B0, k + 3);
      plane(B1, B2, B3, k + 1);
    }
    if (k + 2 < k1) {
      load(B1, k + 4);
      plane(B2, B3, B0, k + 2);
    }
  }
  if (!want_dot) return;
  double v[1] = {acc};
  block_finish<1>(v, partial, counter, G, Lb, st, epi, out);
}

// ---------------------------------------------------------------------------
// LDS-staged operator march (Launch::stencil = -1, 2, 3, 4; y = A x): the operator and
// the per-point arithmetic of k_stencil_march, with every value of x loaded
// from memory once per workgroup.  A workgroup of four waves owns a patch of
// kLdsW = 64 consecutive i by 4 WJ theta rows and marches a segment of phi
// planes.  Per plane, the patch and its one-point halo ((4 WJ + 2) rows of
// 66 values) are loaded D planes ahead into registers by coalesced loads,
// then written to one of D LDS slots; wave w reads rows w WJ .. w WJ + WJ - 1
// of the slot at column i0 + lane and their neighbours, and keeps the
// centres of planes k-1 and k in registers.  One barrier per plane orders
// the slot writes of plane k+1 before their reads; with D >= 3 slots the
// slot written at plane k was last read at plane k-2, which every wave
// finished before that barrier.  The barrier waits for LDS only, so the
// loads of the next planes stay in flight across it.  Patches are ordered r
// fastest: neighbouring patches share the cache lines at their edges (their
// halo columns, and the lines of y that both write part of), and they run
// together on one L2.
//
// WJ: theta rows per wave; D: planes loaded ahead (= LDS slots = unroll of
// the march); WPE: waves per SIMD the register allocation is held to.  The
// tall patches (WJ = 8) load fewer halo rows per point, the short ones
// (WJ = 4) make twice the workgroups for the smaller grids.
template <int WJ_, int D_, int WPE_>
struct LdsShape {
  static constexpr int WJ = WJ_, D = D_, WPE = WPE_;
};
template <typename T, bool kTall>
struct LdsCfg;
template <> struct LdsCfg<double, true> : LdsShape<8, 3, 2> {};
template <> struct LdsCfg<double, false> : LdsShape<4, 3, 1> {};
template <> struct LdsCfg<float, true> : LdsShape<8, 4, 2> {};
template <> struct LdsCfg<float, false> : LdsShape<4, 3, 4> {};
constexpr int kLdsW = 64;
constexpr int kLdsRow = kLdsW + 2;
constexpr int kLdsMaxPlanes = 96;  // phi planes of a segment (per-plane factors in LDS)

template <typename C>
__host__ __device__ inline void stencil_lds_shape(const Grid3& g, int& nwi, int& nwj) {
  nwi = (g.a + kLdsW - 1) / kLdsW;
  nwj = (g.b + 4 * C::WJ - 1) / (4 * C::WJ);
}

// Raw buffer stores of one value: a lane whose byte offset is kDropStore
// writes nothing (out of the buffer's range), so a store needs no branch.
constexpr uint32_t kDropStore = 0xffffffffu;
__device__ __forceinline__ void store_or_drop(__amdgpu_buffer_rsrc_t rs, uint32_t off,
                                              double v) {
  const unsigned long long u = (unsigned long long)__double_as_longlong(v);
  typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
  u32x2 d;
  d.x = (unsigned)(u & 0xffffffffull);
  d.y = (unsigned)(u >> 32);
  __builtin_amdgcn_raw_buffer_store_b64(d, rs, (int)off, 0, 0);
}
__device__ __forceinline__ void store_or_drop(__amdgpu_buffer_rsrc_t rs, uint32_t off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32((unsigned)__float_as_uint(v), rs, (int)off, 0, 0);
}

template <typename T, typename C>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(C::WPE)))
    k_stencil_lds(Metric m, Grid3 g, int nseg, const T* __restrict__ x, T* __restrict__ y,
                  const double* __restrict__ pole_part, CgState* st, double* partial,
                  unsigned* counter, Epi epi, double* out, int guard, int want_dot, TileMap tm) {
  constexpr int WJ = C::WJ, PJ = 4 * WJ, D = C::D;
  constexpr int PE = (PJ + 2) * kLdsRow;          // values of a plane slot
  constexpr int NL = (PE + kBlock - 1) / kBlock;  // loads per thread per plane
  // The march is straight-line code: no branch around a load or a store.  The
  // compiler's wait for a plane's loads counts the memory operations issued
  // after them on every path, so a skipped store or load would make it wait
  // for the planes behind as well.  Hence the slots are padded to NL kBlock
  // values (every thread loads and stages NL values; the padding receives
  // copies of the last one), the lookahead past the segment reloads its last
  // plane, and out-of-grid points are stored to an out-of-range offset.
  __shared__ T S[D][NL * kBlock];
  __shared__ double F[3][kLdsMaxPlanes];
  if (guard && st_done(st)) return;
  const int G = gridDim.x;
  const int Lb = logical_block(G);
  int nwi, nwj;
  stencil_lds_shape<C>(g, nwi, nwj);
  const int seg = Lb / (nwi * nwj), rem = Lb - seg * (nwi * nwj);
  const int wj = rem / nwi, wi = rem - wj * nwi;
  const int kper = (g.c + nseg - 1) / nseg;
  const int k0 = seg * kper, k1 = min(g.c, k0 + kper);
  const int tid = (int)threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int lane = tid & 63;
  const int i0 = wi * kLdsW, j0 = wj * PJ;
  double acc = 0.0;
  // workgroup-uniform: every thread takes part in the loads and barriers
  if (seg < nseg && k0 < k1) {
    const int n = k1 - k0;  // planes k0 .. k1-1; x planes k0-1 .. k1 are read
    const uint32_t ab = g.ab, a_ = (uint32_t)g.a;
    const int jw = j0 + wave * WJ;  // the wave's first row (wave-uniform)
    // factors of the wave's rows, loaded before any store or barrier (so as
    // scalar loads), and the segment's per-plane factors staged in LDS: the
    // march issues no vector load but the x planes
    double SDr[WJ], CTLr[WJ], CTHr[WJ], CPr[WJ];
#pragma unroll
    for (int r = 0; r < WJ; r++) {
      const int j = min(jw + r, g.b - 1);
      SDr[r] = m.SD[j];
      CTLr[r] = m.CTL[j];
      CTHr[r] = m.CTH[j];
      CPr[r] = m.CP[j];
    }
    for (int q = tid; q < n; q += kBlock) {
      F[0][q] = m.DPH[k0 + q];
      F[1][q] = m.DPL[k0 + q];
      F[2][q] = m.DPU[k0 + q];
    }
    // the thread's share of a slot: values e = tid + kBlock q, row e / (kLdsW
    // + 2) of the patch (theta j0-1+row), column e % (kLdsW + 2) (r i0-1+col),
    // clamped into the grid (clamped values are never used)
    uint32_t off[NL];
#pragma unroll
    for (int q = 0; q < NL; q++) {
      const int e = min(tid + kBlock * q, PE - 1);
      const int r = e / kLdsRow, c = e - r * kLdsRow;
      const int ii = min(max(i0 - 1 + c, 0), g.a - 1);
      const int jj = min(max(j0 - 1 + r, 0), g.b - 1);
      off[q] = (uint32_t)ii + a_ * (uint32_t)jj;
    }
    const int i = i0 + lane;
    const bool col_ok = i < g.a;
    const int il = min(i, g.a - 1);
    const double DRH = m.DRH[il], CRL = m.CRL[il], CRH = m.CRH[il];
    const bool im = i > 0, ip = i < g.a - 1;
    const double s0d = jw == 0 ? 2.0 * pole_sum(pole_part, 0, il, g.a) : 0.0;
    const double s1d = jw + WJ >= g.b ? 2.0 * pole_sum(pole_part, 1, il, g.a) : 0.0;
    const __amdgpu_buffer_rsrc_t ys =
        __builtin_amdgcn_make_buffer_rsrc(y, 0, (int)kDropStore, 0x00020000);
    // output offsets (in values) of the wave's rows at plane 0, natural or
    // tile-major (8x8 (i, j) tile planes of 64 values, 2^tk_log2 planes per
    // tile), and the per-plane step; out-of-grid points are dropped
    uint32_t ybase[WJ];
    bool yok[WJ];
#pragma unroll
    for (int r = 0; r < WJ; r++) {
      const int j = min(jw + r, g.b - 1);
      yok[r] = col_ok && jw + r < g.b;
      ybase[r] = tm.on ? ((((uint32_t)(il >> 3) + (uint32_t)tm.TNI * (uint32_t)(j >> 3))
                           << (tm.tk_log2 + 6)) +
                          (uint32_t)(((j & 7) << 3) | (il & 7)))
                       : (uint32_t)il + a_ * (uint32_t)j;
    }
    const uint32_t tplane = (uint32_t)tm.TNI * (uint32_t)tm.TNJ;
    auto yplane = [&](int k) -> uint32_t {
      return tm.on ? (((tplane * ((uint32_t)k >> tm.tk_log2)) << (tm.tk_log2 + 6)) +
                      (((uint32_t)k & ((1u << tm.tk_log2) - 1u)) << 6))
                   : ab * (uint32_t)k;
    };
    // the wave's rows in a slot: row 1 + wave WJ + r, column lane + 1
    const int rbase = (1 + wave * WJ) * kLdsRow + lane + 1;
    T B[D][NL];
    T prv[WJ], cur[WJ];
    // x plane p of the segment (p = 0 .. n+1 is phi plane k0-1+p, periodic);
    // the lookahead past plane n+1 reloads plane n+1 (cached, never staged)
    auto load = [&](T (&R)[NL], int p) {
      int kw = k0 - 1 + min(p, n + 1);
      kw = kw < 0 ? kw + g.c : (kw >= g.c ? kw - g.c : kw);
      const T* xp = x + (size_t)ab * (uint32_t)kw;
#pragma unroll
      for (int q = 0; q < NL; q++) R[q] = xp[off[q]];
      // keep the loads where they are in the sequence of memory operations
      __builtin_amdgcn_sched_barrier(0);
    };
    // step p: load plane p+D into buffer p%D, stage plane p+1 (buffer and
    // slot (p+1)%D), barrier; for p >= 1 compute plane p (slot p%D) with
    // the centres of plane p+1
    auto step = [&](auto pm_c, auto mode_c, int p) {
      constexpr int PM = decltype(pm_c)::value, PN = (PM + 1) % D;
      constexpr int MODE = decltype(mode_c)::value;  // 0 prologue, 1 first centres, 2 plane
      load(B[PM], p + D);
#pragma unroll
      for (int q = 0; q < NL; q++) S[PN][tid + kBlock * q] = B[PN][q];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
      __builtin_amdgcn_s_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
      const T* Sc = S[PM];
      const T* Sn = S[PN];
      if constexpr (MODE != 2) {
        // the prologue steps issue the plane's WJ stores too (dropped), so
        // the memory operations between a plane's loads and its staging are
        // the same on entry to the march as in it, and so are the waits
#pragma unroll
        for (int r = 0; r < WJ; r++) {
          uint32_t yo = kDropStore;
          asm volatile("" : "+v"(yo));
          store_or_drop(ys, yo, (T)0);
        }
      }
      if constexpr (MODE == 1) {
#pragma unroll
        for (int r = 0; r < WJ; r++) {
          prv[r] = Sc[rbase + r * kLdsRow];
          cur[r] = Sn[rbase + r * kLdsRow];
        }
      } else if constexpr (MODE == 2) {
        T nxt[WJ], xw[WJ], xe[WJ];
#pragma unroll
        for (int r = 0; r < WJ; r++) {
          nxt[r] = Sn[rbase + r * kLdsRow];
          xw[r] = Sc[rbase + r * kLdsRow - 1];
          xe[r] = Sc[rbase + r * kLdsRow + 1];
        }
        const T xup = Sc[rbase - kLdsRow];
        const T xdn = Sc[rbase + WJ * kLdsRow];
        const int k = k0 - 1 + p;
        const double dph = F[0][p - 1], dpl = F[1][p - 1], dpu = F[2][p - 1];
        const uint32_t yk = yplane(k);
#pragma unroll
        for (int r = 0; r < WJ; r++) {
          const int j = jw + r;
          const bool jm = j > 0, jp = j < g.b - 1;
          const double P1 = DRH * CPr[r];
          double co[7];
          op_coefs(DRH, CRL, CRH, SDr[r], CTLr[r], CTHr[r], P1, dph, dpl, dpu, co);
          const T xs = r == 0 ? xup : cur[r - 1];
          const T xn = r == WJ - 1 ? xdn : cur[r + 1];
          const T yv = op_point<T>(co, prv[r], xs, xw[r], cur[r], xe[r], xn, nxt[r], im, ip, jm,
                                   jp, m.pm_r1, s0d, s1d, false, (T)0);
          uint32_t yo = yok[r] ? (ybase[r] + yk) * (uint32_t)sizeof(T) : kDropStore;
          // an opaque offset: the store stays one unconditional instruction
          asm volatile("" : "+v"(yo));
          store_or_drop(ys, yo, yv);
          acc += yok[r] ? (double)cur[r] * (double)yv : 0.0;
        }
#pragma unroll
        for (int r = 0; r < WJ; r++) {
          prv[r] = cur[r];
          cur[r] = nxt[r];
        }
      }
    };
    using C0 = std::integral_constant<int, 0>;
    using C1 = std::integral_constant<int, 1>;
    using C2 = std::integral_constant<int, 2>;
    using C3 = std::integral_constant<int, 3>;
    using Cm = std::integral_constant<int, D - 1>;
    using Cp2 = std::integral_constant<int, 2 % D>;
    using Cp3 = std::integral_constant<int, 3 % D>;
    // prologue as steps -1 and 0: planes 0 and 1 staged, the centres of
    // planes 0 and 1 read, planes 2 .. D in flight
#pragma unroll
    for (int q = 0; q + 1 < D; q++) load(B[q], q);
    step(Cm{}, C0{}, -1);
    step(C0{}, C1{}, 0);
    // whole rounds of D planes, then the remaining ones: a loop with one
    // entry and one exit, entered in the state it repeats
    const int nfull = n / D, nrem = n - nfull * D;
    for (int t = 0; t < nfull; t++) {
      const int p = 1 + D * t;
      step(C1{}, C2{}, p);
      step(Cp2{}, C2{}, p + 1);
      step(Cp3{}, C2{}, p + 2);
      if constexpr (D == 4) step(C0{}, C2{}, p + 3);
    }
    const int pr = 1 + D * nfull;
    if (nrem >= 1) step(C1{}, C2{}, pr);
    if (nrem >= 2) step(Cp2{}, C2{}, pr + 1);
    if constexpr (D == 4)
      if (nrem >= 3) step(C3{}, C2{}, pr + 2);
  }
  if (!want_dot) return;
  double v[1] = {acc};
  block_finish<1>(v, partial, counter, G, Lb, st, epi, out);
}

// Workgroups of k_stencil_lds<T, C> the device holds at once (its
// occupancy on every CU; the same on every device of the process).
template <typename T, typename C>
static int stencil_lds_resident() {
  static const int R = [] {
    int dev = 0, ncu = 0, nb = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        ncu < 1)
      ncu = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_stencil_lds<T, C>, kBlock, 0) !=
            hipSuccess ||
        nb < 1)
      nb = 2;
    return nb * ncu;
  }();
  return R;
}

// Segments per patch column: whole rounds of the workgroups the device
// holds at once (about 4096 workgroups; a last round partly filled would
// run nearly alone), segments of >= kSegPlanes and <= kLdsMaxPlanes planes.
template <typename T, typename C>
static int stencil_lds_segments(const Grid3& g) {
  int nwi, nwj;
  stencil_lds_shape<C>(g, nwi, nwj);
  const int P = nwi * nwj, R = stencil_lds_resident<T, C>();
  const int rounds = std::max(1, (4096 + R / 2) / R);
  int nseg = rounds * R / P;
  const int maxseg = (g.c + kSegPlanes - 1) / kSegPlanes;
  if (nseg > maxseg) {
    const int fit = maxseg * P / R;  // whole rounds that fit
    nseg = fit >= 1 ? fit * R / P : maxseg;
  }
  nseg = std::max(nseg, (g.c + kLdsMaxPlanes - 1) / kLdsMaxPlanes);
  if (nseg < 1) nseg = 1;
  return nseg;
}

template <typename T, typename C>
static int stencil_lds_blocks(const Grid3& g) {
  int nwi, nwj;
  stencil_lds_shape<C>(g, nwi, nwj);
  const int G = nwi * nwj * stencil_lds_segments<T, C>(g);
  return (G + 7) / 8 * 8;
}

static int stencil_lds_blocks_max(const Grid3& g) {
  return std::max({stencil_lds_blocks<double, LdsCfg<double, true>>(g),
                   stencil_lds_blocks<double, LdsCfg<double, false>>(g),
                   stencil_lds_blocks<float, LdsCfg<float, true>>(g),
                   stencil_lds_blocks<float, LdsCfg<float, false>>(g)});
}

// The tall patches when they still make about 8 workgroups per CU.
template <typename T>
static bool stencil_lds_tall(const Grid3& g) {
  return stencil_lds_blocks<T, LdsCfg<T, true>>(g) >= 8 * 256;
}

template <typename T, typename C>
static void launch_stencil_lds(const Metric& m, const Grid3& g, const T* x, T* y,
                               const double* pole_part, CgState* st, Reduce red, int slot,
                               Epi epi, double* out, bool guard, int want, hipStream_t s,
                               const TileMap& tm) {
  const int nseg = stencil_lds_segments<T, C>(g);
  const int G = stencil_lds_blocks<T, C>(g);
  k_stencil_lds<T, C><<<G, kBlock, 0, s>>>(m, g, nseg, x, y, pole_part, st, red.partial,
                                            red.counter + (size_t)slot * kRedCounters, epi, out,
                                            guard ? 1 : 0, want, tm);
}

static int stencil_rows_segments(const Grid3& g) {
  // about 16 workgroups (64 waves) per CU in total, >= kSegPlanes planes each
  int nwi, nwg;
  stencil_rows_shape(g, nwi, nwg);
  int nseg = (16 * 256 + nwi * nwg - 1) / (nwi * nwg);
  const int maxseg = (g.c + kSegPlanes - 1) / kSegPlanes;
  if (nseg > maxseg) nseg = maxseg;
  if (nseg < 1) nseg = 1;
  return nseg;
}

static int stencil_rows_blocks(const Grid3& g) {
  int nwi, nwg;
  stencil_rows_shape(g, nwi, nwg);
  const int G = nwi * nwg * stencil_rows_segments(g);
  return (G + 7) / 8 * 8;
}

static int stencil_segments(const Grid3& g, const TileMap& tm) {
  // enough blocks to fill the chip several times, >= kSegPlanes planes each
  const int ncol_blocks = stencil_col_blocks(g, tm);
  int nseg = (8 * 256 * 4 + ncol_blocks - 1) / ncol_blocks;
  const int maxseg = (g.c + kSegPlanes - 1) / kSegPlanes;
  if (nseg > maxseg) nseg = maxseg;
  if (nseg < 1) nseg = 1;
  return nseg;
}

int stencil_blocks(const Grid3& g, const TileMap& tm) {
  const int G = stencil_col_blocks(g, tm) * stencil_segments(g, tm);
  return std::max({(G + 7) / 8 * 8, stencil_rows_blocks(g), stencil_lds_blocks_max(g)});
}

template <typename T>
void launch_stencil(const Metric& m, const Grid3& g, const Launch& L, const T* x, T* y,
                    const T* b, int mode, const double* pole_part, CgState* st, Reduce red,
                    int slot, Epi epi, double* out, bool guard, hipStream_t s,
                    const TileMap& tm) {
  const int want = (epi != EPI_NONE);
  // the LDS-staged march (y = A x only) stores through 32-bit buffer
  // offsets: every output offset, tile padding included, below kDropStore
  const double nout =
      tm.on ? (double)tm.TNI * tm.TNJ * 64.0 *
                  (double)(((g.c + (1 << tm.tk_log2) - 1) >> tm.tk_log2) << tm.tk_log2)
            : (double)g.N;
  const bool lds_ok = mode == 0 && nout * sizeof(T) < (double)kDropStore;
  if (lds_ok && (L.stencil < 0 || L.stencil >= 2)) {
    const bool tall = L.stencil == 3 || ((L.stencil < 0 || L.stencil == 2) &&
                                         stencil_lds_tall<T>(g));
    if (tall)
      launch_stencil_lds<T, LdsCfg<T, true>>(m, g, x, y, pole_part, st, red, slot, epi, out,
                                             guard, want, s, tm);
    else
      launch_stencil_lds<T, LdsCfg<T, false>>(m, g, x, y, pole_part, st, red, slot, epi, out,
                                              guard, want, s, tm);
    return;
  }
  if (L.stencil == 1 || (L.stencil < 0 && stencil_rows_chosen(g, sizeof(T)))) {
    const int nseg = stencil_rows_segments(g);
    const int G = stencil_rows_blocks(g);
    k_stencil_rows<T><<<G, kBlock, 0, s>>>(m, g, nseg, x, y, b, mode, pole_part, st,
                                            red.partial,
                                            red.counter + (size_t)slot * kRedCounters, epi, out,
                                            guard ? 1 : 0, want, tm);
    return;
  }
  const int nseg = stencil_segments(g, tm);
  const int G0 = stencil_col_blocks(g, tm) * nseg;
  const int G = (G0 + 7) / 8 * 8;
  k_stencil_march<T><<<G, kBlock, 0, s>>>(m, g, nseg, x, y, b, mode, pole_part, st,
                                           red.partial, red.counter + (size_t)slot * kRedCounters,
                                           epi, out, guard ? 1 : 0, want, tm);
}

// ---------------------------------------------------------------------------
// Diagonal preconditioner value a_i = real(1/a4, REAL32) (:4649-4661).
__device__ __forceinline__ float diag_inv(const Metric& m, int i, int j, int k) {
  const Coef cf = coef_at(m, i, j, k);
  return (float)(1.0 / cf.a4);
}

template <typename T>
__device__ __forceinline__ T apply_point_prec(const Metric& m, const Grid3& g, int prec,
                                              uint32_t n, T r) {
  if (prec == PREC_DIAG) {
    int i, j, k;
    unflatten(g, n, i, j, k);
    const float ai = diag_inv(m, i, j, k);
    return (T)((T)ai * r);
  }
  return r;
}

// x += alpha p ; r -= alpha ap ; (DIAG/NONE) z = M^-1 r with s0=r.z, s1=r.r
// (cgsolve :4340-4351).  For ILU only s1=r.r is reduced here.  With xs != nullptr (FP32 recurrence, FP64 solution) the
// update goes to the double vector xs instead of x: xs += (alpha scale) p,
// scale = *xscale being the factor of the scaled FP32 system.  Accumulating
// the solution in FP32 would leave its true residual at about
// eps32*cond(A) of the start, whatever the recurrence says.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_cg_update(Metric m, Grid3 g, int G, int CH, int prec, T* __restrict__ x,
                T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ ap, CgState* st,
                double* partial, unsigned* counter, double* __restrict__ xs,
                const double* __restrict__ xscale, TileMap tm) {
  if (st_done(st)) return;
  const int Lb = logical_block(G);
  const uint32_t n0 = (uint32_t)Lb * (uint32_t)CH;
  const uint32_t n1 = min(g.N, n0 + (uint32_t)CH);
  const double alpha_d = st_load(&st->alpha);
  const T alpha = (T)alpha_d;
  const double alpha_x = xs ? alpha_d * *xscale : 0.0;
  double s0 = 0.0, s1 = 0.0;
  // blocks of kUnroll points per thread, all loads issued first; the sums
  // keep the point order of the plain loop (one chunk per workgroup:
  // measured 5% faster here than the grid-stride of the FP32 streams)
  for (uint32_t nb = n0 + threadIdx.x; nb < n1; nb += kUnroll * kBlock) {
    T xv[kUnroll], pv[kUnroll], rv[kUnroll], av[kUnroll];
    double xd[kUnroll];
    uint32_t tr[kUnroll];  // position of point n in r
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint32_t n = nb + u * kBlock;
      xv[u] = pv[u] = rv[u] = av[u] = (T)0;
      xd[u] = 0.0;
      tr[u] = 0;
      if (n < n1) {
        tr[u] = tmap(tm, g, n);
        if (xs)
          xd[u] = xs[n];
        else
          xv[u] = x[n];
        pv[u] = p[n];
        rv[u] = r[tr[u]];
        av[u] = ap[n];
      }
    }
#pragma unroll
    for (int u = 0; u < kUnroll; u++) {
      const uint32_t n = nb + u * kBlock;
      if (n < n1) {
        if (xs)
          xs[n] = xd[u] + alpha_x * (double)pv[u];
        else
          x[n] = xv[u] + alpha * pv[u];
        const T rn = rv[u] - alpha * av[u];
        r[tr[u]] = rn;
        if (prec != PREC_ILU) {
          const T z = apply_point_prec<T>(m, g, prec, n, rn);
          s0 += (double)rn * (double)z;
        }
        s1 += (double)rn * (double)rn;
      }
    }
  }
  double v[2] = {s0, s1};
  block_finish<2>(v, partial, counter, G, Lb, st, prec == PREC_ILU ? EPI_RR : EPI_RITER,
                  nullptr);
}

template <typename T>
void launch_cg_update(const Metric& m, const Grid3& g, const Launch& L, int prec, T* x, T* r,
                      const T* p, const T* ap, CgState* st, Reduce red, int slot,
                      hipStream_t s, double* xs, const double* xscale, const TileMap& tm) {
  int G = L.G, CH = L.CH;
  if (sizeof(T) == 8 && L.Gcg > 0) {
    G = L.Gcg;
    CH = L.CHcg;
  }
  k_cg_update<T><<<G, kBlock, 0, s>>>(m, g, G, CH, prec, x, r, p, ap, st, red.partial,
                                       red.counter + (size_t)slot * kRedCounters, xs, xscale, tm);
}

// The residual update of the FP32 recurrence with the FP64 solution, one
// contiguous pass over the tile-major r and ap: r -= alpha ap (entries
// outside the grid are zero on both; r.r goes with r.z after the
// preconditioner).  The register-tile forward sweep folds the same update
// into its loads when it can (dilu_tiles_fuse_rupd): the same fmaf, so the
// same bits.  The solution update goes with the p update.
__global__ void __launch_bounds__(kBlock)
    k_r_update(uint32_t n_all, int G, float* __restrict__ r, const float* __restrict__ ap,
               const CgState* st) {
  if (st_done(st)) return;
  const uint32_t n1 = n_all;
  const float alpha = (float)st_load(&st->alpha);
  constexpr int kU = 16;  // points per thread in flight (measured: 8 -> 16, 4% faster)
  // grid-stride over runs of kU kBlock points in dispatch order: the
  // workgroups in flight stream one window of the vectors together
  for (uint32_t nb = blockIdx.x * (kU * kBlock) + threadIdx.x; nb < n1;
       nb += (uint32_t)G * (kU * kBlock)) {
    float rv[kU], av[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      rv[u] = av[u] = 0.f;
      if (n < n1) {
        rv[u] = r[n];
        av[u] = ap[n];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      if (n < n1) r[n] = __builtin_fmaf(-alpha, av[u], rv[u]);
    }
  }
}

void launch_r_update(uint32_t n_all, const Launch& Lt, float* r, const float* ap,
                     const CgState* st, hipStream_t s) {
  const int G = Lt.Gru > 0 ? Lt.Gru : Lt.G;
  k_r_update<<<G, kBlock, 0, s>>>(n_all, G, r, ap, st);
}

// p = beta p + z  (:4361-4363)
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_p_update(Metric m, Grid3 g, int G, int CH, int prec, T* __restrict__ p,
               const T* __restrict__ r, const T* __restrict__ z, CgState* st, TileMap tm,
               double* __restrict__ xs, const double* __restrict__ xscale,
               unsigned* counter) {
  // With xs (FP32 recurrence, FP64 solution): first the solution update of
  // the iteration whose alpha is current, xs += (alpha scale) p_old, once per
  // iteration, including the iteration that set done (CgState::x_ncg counts
  // the applied updates; the last workgroup advances it, after every
  // workgroup has read it); then p = z + beta p_old unless done.
  const int done = st_done(st);
  const int ncg = xs ? st_load_int(&st->ncg) : 0;
  const bool do_x = xs && st_load_int(&st->x_ncg) < ncg;
  if (done && !do_x) return;
  const int Lb = logical_block(G);
  const uint32_t n1 = g.N;
  const T beta = (T)st_load(&st->beta);
  const double alpha_x = do_x ? st_load(&st->alpha) * *xscale : 0.0;
  // 16 points per thread in flight, and the p and solution streams read
  // (and the solution written) non-temporally: neither is reused before it
  // has left the caches, and keeping them out leaves room for the gathered
  // z (measured: 1.60 -> 1.27 ms at 214.7 M points)
  constexpr int kU = 16;
  // grid-stride over runs of kU kBlock points in dispatch order: the
  // workgroups in flight stream one window of the vectors together
  for (uint32_t nb = blockIdx.x * (kU * kBlock) + threadIdx.x; nb < n1;
       nb += (uint32_t)G * (kU * kBlock)) {
    T pv[kU], zv[kU];
    double xv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      pv[u] = zv[u] = (T)0;
      xv[u] = 0.0;
      if (n < n1) {
        pv[u] = __builtin_nontemporal_load(p + n);
        if (do_x) xv[u] = __builtin_nontemporal_load(xs + n);
        if (!done) zv[u] = (prec == PREC_ILU) ? z[tmap(tm, g, n)] : r[n];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      if (n < n1) {
        if (do_x) __builtin_nontemporal_store(xv[u] + alpha_x * (double)pv[u], xs + n);
        if (!done) {
          const T zn = (prec == PREC_ILU) ? zv[u] : apply_point_prec<T>(m, g, prec, n, zv[u]);
          p[n] = beta * pv[u] + zn;
        }
      }
    }
  }
  if (do_x) {
    __shared__ int amlast;
    __syncthreads();
    // only the reads of x_ncg must precede its update: a plain count-in
    if (threadIdx.x == 0)
      amlast = (__hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                (unsigned)G - 1);
    __syncthreads();
    if (amlast && threadIdx.x == 0) {
      __hip_atomic_store(&st->x_ncg, ncg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

// The p update of the FP32 recurrence with the FP64 solution, the
// solution updated once per group of kXBatch iterations: iteration t of a
// group reads its direction p_t from pb.p[t] and writes p_(t+1) = z + beta
// p_t to pb.p[(t+1) % kXBatch]; at the group's last iteration, or when the
// recurrence has stopped, the solution takes the group's pending updates
// x += (alpha_u scale) p_u in iteration order -- the arithmetic, in the
// order, of one update per iteration (explicit fused multiply-adds), with
// the solution read and written once instead of kXBatch times.  The
// directions of earlier iterations are still in their buffers: p_(t+1)
// overwrites p_(t+1-kXBatch), read first for the flush.
template <int T_>
__global__ void __launch_bounds__(kBlock)
    k_p_update_x(Grid3 g, int G, int CH, PBuf pb, const float* __restrict__ z, CgState* st,
                 TileMap tm, double* __restrict__ xs, const double* __restrict__ xscale,
                 unsigned* counter) {
  const int done = st_done(st);
  const int ncg = st_load_int(&st->ncg);
  const int pending = ncg - st_load_int(&st->x_ncg);  // iterations not in x yet
  const bool flush = pending > 0 && (T_ == kXBatch - 1 || done);
  if (pending > 0 && blockIdx.x == 0 && threadIdx.x == 0) {
    // this iteration's alpha for its group's flush, and where p now is
    __hip_atomic_store(&st->alpha_hist[T_], st_load(&st->alpha), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&st->p_cur, done ? T_ : (T_ + 1) % kXBatch, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  if (done && !flush) return;
  const uint32_t n1 = g.N;
  const float beta = (float)st_load(&st->beta);
  const double sc = *xscale;
  double ax[kXBatch];
  bool use[kXBatch];
#pragma unroll
  for (int u = 0; u < kXBatch; u++) {
    use[u] = flush && u <= T_ && u >= T_ + 1 - pending;
    ax[u] = u == T_ ? st_load(&st->alpha) * sc : (use[u] ? st_load(&st->alpha_hist[u]) * sc : 0.0);
  }
  const float* __restrict__ pin = pb.p[T_];
  float* __restrict__ pout = pb.p[(T_ + 1) % kXBatch];
  constexpr int kU = 8;  // (16 measured 4% slower: half the resident waves)
  // grid-stride over runs of kU kBlock points in dispatch order: the
  // workgroups in flight stream one window of the vectors together
  for (uint32_t nb = blockIdx.x * (kU * kBlock) + threadIdx.x; nb < n1;
       nb += (uint32_t)G * (kU * kBlock)) {
    float pv[kU], zv[kU], pold[kXBatch][kU];
    double xv[kU];
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      pv[u] = zv[u] = 0.f;
      xv[u] = 0.0;
#pragma unroll
      for (int q = 0; q < kXBatch; q++) pold[q][u] = 0.f;
      if (n < n1) {
        pv[u] = __builtin_nontemporal_load(pin + n);
        if (flush) {
          xv[u] = __builtin_nontemporal_load(xs + n);
#pragma unroll
          for (int q = 0; q < T_; q++)
            if (use[q]) pold[q][u] = __builtin_nontemporal_load(pb.p[q] + n);
        }
        if (!done) zv[u] = z[tmap(tm, g, n)];
      }
    }
#pragma unroll
    for (int u = 0; u < kU; u++) {
      const uint32_t n = nb + u * kBlock;
      if (n < n1) {
        if (flush) {
          double x = xv[u];
#pragma unroll
          for (int q = 0; q < T_; q++)
            if (use[q]) x = __builtin_fma(ax[q], (double)pold[q][u], x);
          x = __builtin_fma(ax[T_], (double)pv[u], x);
          __builtin_nontemporal_store(x, xs + n);
        }
        if (!done) pout[n] = beta * pv[u] + zv[u];
      }
    }
  }
  if (flush) {
    __shared__ int amlast;
    __syncthreads();
    // only the reads of x_ncg must precede its update: a plain count-in
    if (threadIdx.x == 0)
      amlast = (__hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                (unsigned)G - 1);
    __syncthreads();
    if (amlast && threadIdx.x == 0) {
      __hip_atomic_store(&st->x_ncg, ncg, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
}

void launch_p_update_x(const Grid3& g, const Launch& L, int t, const PBuf& pb, const float* z,
                       CgState* st, hipStream_t s, const TileMap& tm, double* xs,
                       const double* xscale, unsigned* counter) {
  const int G = L.Gpx[t & 3] > 0 ? L.Gpx[t & 3] : L.G;
  switch (t) {
    case 0:
      k_p_update_x<0><<<G, kBlock, 0, s>>>(g, G, L.CH, pb, z, st, tm, xs, xscale, counter);
      break;
    case 1:
      k_p_update_x<1><<<G, kBlock, 0, s>>>(g, G, L.CH, pb, z, st, tm, xs, xscale, counter);
      break;
    case 2:
      k_p_update_x<2><<<G, kBlock, 0, s>>>(g, G, L.CH, pb, z, st, tm, xs, xscale, counter);
      break;
    default:
      k_p_update_x<3><<<G, kBlock, 0, s>>>(g, G, L.CH, pb, z, st, tm, xs, xscale, counter);
      break;
  }
}


template <typename T>
void launch_p_update(const Metric& m, const Grid3& g, const Launch& L, int prec, T* p,
                     const T* r, const T* z, CgState* st, hipStream_t s, const TileMap& tm,
                     double* xs, const double* xscale, unsigned* counter) {
  const int G = sizeof(T) == 8 && L.Gp > 0 ? L.Gp : L.G;
  k_p_update<T><<<G, kBlock, 0, s>>>(m, g, G, L.CH, prec, p, r, z, st, tm, xs, xscale, counter);
}

// s0 = x.y, s1 = u.v
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_dot2(Grid3 g, int G, int CH, const T* __restrict__ x, const T* __restrict__ y,
           const T* __restrict__ u, const T* __restrict__ v, CgState* st, double* partial,
           unsigned* counter, Epi epi, double* out, int guard, TileMap tm) {
  if (guard && st_done(st)) return;
  const int Lb = logical_block(G);
  const uint32_t n1 = g.N;
  double s0 = 0.0, s1 = 0.0;
  // 8 points per thread in flight; x and y read non-temporally (streams);
  // grid-stride over runs of kU kBlock points in dispatch order: the
  // workgroups in flight stream one window of the vectors together
  constexpr int kU = 8;
  for (uint32_t nb = blockIdx.x * (kU * kBlock) + threadIdx.x; nb < n1;
       nb += (uint32_t)G * (kU * kBlock)) {
    T xv[kU], yv[kU], uv[kU], vv[kU];
#pragma unroll
    for (int q = 0; q < kU; q++) {
      const uint32_t n = nb + q * kBlock;
      xv[q] = yv[q] = uv[q] = vv[q] = (T)0;
      if (n < n1) {
        const uint32_t t = tmap(tm, g, n);
        xv[q] = __builtin_nontemporal_load(x + t);
        yv[q] = __builtin_nontemporal_load(y + t);
        if (u) {
          uv[q] = u[t];
          vv[q] = v[t];
        }
      }
    }
#pragma unroll
    for (int q = 0; q < kU; q++) {
      const uint32_t n = nb + q * kBlock;
      if (n < n1) {
        s0 += (double)xv[q] * (double)yv[q];
        if (u) s1 += (double)uv[q] * (double)vv[q];
      }
    }
  }
  double sums[2] = {s0, s1};
  block_finish<2>(sums, partial, counter, G, Lb, st, epi, out);
}

template <typename T>
void launch_dot2(const Grid3& g, const Launch& L, const T* x, const T* y, const T* u,
                 const T* v, CgState* st, Reduce red, int slot, Epi epi, double* out,
                 bool guard, hipStream_t s, const TileMap& tm) {
  const int G = sizeof(T) == 8 && L.Gd2 > 0 ? L.Gd2 : L.G;
  k_dot2<T><<<G, kBlock, 0, s>>>(g, G, L.CH, x, y, u, v, st, red.partial,
                                  red.counter + (size_t)slot * kRedCounters, epi, out,
                                  guard ? 1 : 0, tm);
}

// z = M^-1 r (DIAG/NONE), reductions s0=r.z, s1=r.r.
template <typename T>
__global__ void __launch_bounds__(kBlock)
    k_point_prec(Metric m, Grid3 g, int G, int CH, int prec, const T* __restrict__ r,
                 T* __restrict__ z, CgState* st, double* partial, unsigned* counter, Epi epi,
                 double* out) {
  const int Lb = logical_block(G);
  const uint32_t n0 = (uint32_t)Lb * (uint32_t)CH;
  const uint32_t n1 = min(g.N, n0 + (uint32_t)CH);
  double s0 = 0.0, s1 = 0.0;
  for (uint32_t n = n0 + threadIdx.x; n < n1; n += kBlock) {
    const T rn = r[n];
    const T zn = apply_point_prec<T>(m, g, prec, n, rn);
    z[n] = zn;
    s0 += (double)rn * (double)zn;
    s1 += (double)rn * (double)rn;
  }
  if (epi == EPI_NONE) return;
  double sums[2] = {s0, s1};
  block_finish<2>(sums, partial, counter, G, Lb, st, epi, out);
}

template <typename T>
void launch_point_prec(const Metric& m, const Grid3& g, const Launch& L, int prec, const T* r,
                       T* z, CgState* st, Reduce red, int slot, Epi epi, double* out,
                       hipStream_t s) {
  k_point_prec<T><<<L.G, kBlock, 0, s>>>(m, g, L.G, L.CH, prec, r, z, st, red.partial,
                                          red.counter + (size_t)slot * kRedCounters, epi, out);
}

template <typename T>
__global__ void k_sub(Grid3 g, T* __restrict__ r, const T* __restrict__ ap, TileMap tm) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.N) return;
  const uint32_t t = tmap(tm, g, n);
  r[t] = r[t] - ap[n];
}

template <typename T>
void launch_sub(const Grid3& g, const Launch&, T* r, const T* ap, hipStream_t s,
                const TileMap& tm) {
  k_sub<T><<<(g.N + kBlock - 1) / kBlock, kBlock, 0, s>>>(g, r, ap, tm);
}

// dst[n] = src[n], or with a tile map: gather dst[n] = src[tmap(n)]
// (scatter == 0) or scatter dst[tmap(n)] = src[n] (scatter == 1).
template <typename T>
__global__ void k_copy(Grid3 g, T* __restrict__ d, const T* __restrict__ src, TileMap tm,
                       int scatter) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.N) return;
  const uint32_t t = tmap(tm, g, n);
  if (scatter)
    d[t] = src[n];
  else
    d[n] = src[t];
}

template <typename T>
void launch_copy(const Grid3& g, const Launch&, T* dst, const T* src, hipStream_t s,
                 const TileMap& tm, int scatter) {
  k_copy<T><<<(g.N + kBlock - 1) / kBlock, kBlock, 0, s>>>(g, dst, src, tm, scatter);
}

// op 0: dst[tmap(n)] = (TD)(src[n] * num/den)      op 1: dst[n] += (TD)src[n] * den/num
template <typename TD, typename TS>
__global__ void k_cast_scale(Grid3 g, TD* __restrict__ d, const TS* __restrict__ src,
                             const double* num, const double* den, int op, TileMap tm) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.N) return;
  double sc = 1.0;
  if (num && den) sc = (op == 0) ? (*num / *den) : (*den / *num);
  if (op == 0)
    d[tmap(tm, g, n)] = (TD)((double)src[n] * sc);
  else
    d[n] = (TD)((double)d[n] + (double)src[n] * sc);
}

template <typename TD, typename TS>
void launch_cast_scale(const Grid3& g, TD* dst, const TS* src, const double* num,
                       const double* den, int op, hipStream_t s, const TileMap& tm) {
  k_cast_scale<TD, TS><<<(g.N + kBlock - 1) / kBlock, kBlock, 0, s>>>(g, dst, src, num, den, op,
                                                                      tm);
}

__global__ void k_fill(double* p, double v, size_t n) {
  const size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill(double* p, double v, size_t n, hipStream_t s) {
  if (n) k_fill<<<(unsigned)((n + kBlock - 1) / kBlock), kBlock, 0, s>>>(p, v, n);
}

// ---------------------------------------------------------------------------
// RHS (boundary trick, :3941-3950): x_ax = 0 except the r ghosts
//   x(1)  = -br0*dr1            -> b(i=2)    = -(a3 * x(1))
//   x(nr) = +br1*drn (Neumann)  -> b(i=nrm1) = -(a5 * x(nr))
// (pole and seam ghosts of a zero interior are zero).
__global__ void k_rhs(Metric m, Grid3 g, const double* __restrict__ br0,
                      const double* __restrict__ br1, double dr1, double drn, int nt,
                      double* __restrict__ rhs) {
  const uint32_t n = blockIdx.x * kBlock + threadIdx.x;
  if (n >= g.N) return;
  int i, j, k;
  unflatten(g, n, i, j, k);
  const Coef cf = coef_at(m, i, j, k);
  double v = 0.0;
  // br0(j,k) Fortran (nt,np): interior j -> j+1 (0-based row), k -> k+1
  const size_t bidx = (size_t)(j + 1) + (size_t)nt * (size_t)(k + 1);
  if (i == 0) {
    const double x1 = 0.0 - 1.0 * br0[bidx] * dr1;
    v = v + cf.a3 * x1;
  }
  if (i == g.a - 1 && br1) {
    const double xn = 0.0 + 1.0 * br1[bidx] * drn;
    v = v + cf.a5 * xn;
  }
  rhs[n] = -v;
}

void launch_rhs(const Metric& m, const Grid3& g, const double* br0, const double* br1,
                double dr1, double drn, int nt, double* rhs, hipStream_t s) {
  k_rhs<<<(g.N + kBlock - 1) / kBlock, kBlock, 0, s>>>(m, g, br0, br1, dr1, drn, nt, rhs);
}

// ---------------------------------------------------------------------------
// D-ILU.  On the reference's ILU pattern (boundary r/theta couplings dropped,
// phi wrap kept) ILU(0) = D-ILU: L and U keep A's off-diagonals, only the
// pivots change:  d_p = a4_p - sum_{q<p} (a_pq/d_q) a_qp   (ilu0 :4717-4813
// on the diacsr matrix :4870-5018).  Points on hyperplane i+j+k=L depend
// only on hyperplane L-1 (the wrap partner (i,j,0) of (i,j,c-1) lies on an
// earlier hyperplane), so each hyperplane is one parallel step.

// Coefficient providers: matrix-free (from metric factors; symmetric) or
// stored (from the caller's CSR values, legacy ABI).
struct MFCoefs {
  Metric m;
  __device__ __forceinline__ double diag(int i, int j, int k, uint32_t) const {
    return coef_at(m, i, j, k).a4;
  }
  // lower couplings of p and the matching upper couplings of q
  __device__ __forceinline__ void lower(int i, int j, int k, uint32_t, double& ak, double& aj,
                                        double& ai, double& aw) const {
    const Coef c = coef_at(m, i, j, k);
    ak = c.a1;
    aj = c.a2;
    ai = c.a3;
    aw = c.a7;  // at k=c-1: coupling to (i,j,0)
  }
  __device__ __forceinline__ void upper(int i, int j, int k, uint32_t, double& bi, double& bj,
                                        double& bk, double& bw) const {
    const Coef c = coef_at(m, i, j, k);
    bi = c.a5;
    bj = c.a6;
    bk = c.a7;
    bw = c.a1;  // at k=0: coupling to (i,j,c-1)
  }
  // a_qp for the lower neighbours q of p (symmetric operator: a_qp = a_pq)
  __device__ __forceinline__ void lower_transpose(int i, int j, int k, uint32_t n, const Grid3&,
                                                  double& tk, double& tj, double& ti,
                                                  double& tw) const {
    lower(i, j, k, n, tk, tj, ti, tw);
  }
};

// Couplings of the level-scheduled triangular solves: the FP32 sweep
// couplings (sweep_coup_*), as doubles for the providers' interface (the
// factorisation keeps MFCoefs' FP64 coefficients).
struct MFSweep {
  Metric m;
  __device__ __forceinline__ void lower(int i, int j, int k, uint32_t, double& ak, double& aj,
                                        double& ai, double& aw) const {
    const SweepCoup c = sweep_coup_lower(m, i, j);
    ak = (double)(c.k * m.FDPL[k]);
    aj = (double)(c.j * m.FDPH[k]);
    ai = (double)(c.i * m.FDPH[k]);
    aw = (double)(c.k * m.FDPU[m.c - 1]);  // at k=c-1: coupling to (i,j,0)
  }
  __device__ __forceinline__ void upper(int i, int j, int k, uint32_t, double& bi, double& bj,
                                        double& bk, double& bw) const {
    const SweepCoup c = sweep_coup_upper(m, i, j);
    bi = (double)(c.i * m.FDPH[k]);
    bj = (double)(c.j * m.FDPH[k]);
    bk = (double)(c.k * m.FDPU[k]);
    bw = (double)(c.k * m.FDPL[0]);  // at k=0: coupling to (i,j,c-1)
  }
};

struct StoredCoefs {
  CoefArrays ca;
  const float* dg;
  int a, b, c;
  uint32_t ab;
  __device__ __forceinline__ double diag(int, int, int, uint32_t n) const { return (double)dg[n]; }
  __device__ __forceinline__ void lower(int i, int j, int k, uint32_t n, double& ak, double& aj,
                                        double& ai, double& aw) const {
    ak = (double)ca.lo_k[n];
    aj = (double)ca.lo_j[n];
    ai = (double)ca.lo_i[n];
    aw = (k == c - 1) ? (double)ca.wrap_lo[(uint32_t)i + (uint32_t)a * (uint32_t)j] : 0.0;
  }
  __device__ __forceinline__ void upper(int i, int j, int k, uint32_t n, double& bi, double& bj,
                                        double& bk, double& bw) const {
    bi = (double)ca.up_i[n];
    bj = (double)ca.up_j[n];
    bk = (double)ca.up_k[n];
    bw = (k == 0) ? (double)ca.wrap_up[(uint32_t)i + (uint32_t)a * (uint32_t)j] : 0.0;
  }
  __device__ __forceinline__ void lower_transpose(int i, int j, int k, uint32_t n, const Grid3& g,
                                                  double& tk, double& tj, double& ti,
                                                  double& tw) const {
    tk = (k > 0) ? (double)ca.up_k[n - g.ab] : 0.0;
    tj = (j > 0) ? (double)ca.up_j[n - (uint32_t)g.a] : 0.0;
    ti = (i > 0) ? (double)ca.up_i[n - 1] : 0.0;
    tw = (k == c - 1) ? (double)ca.wrap_up[(uint32_t)i + (uint32_t)a * (uint32_t)j] : 0.0;
  }
};

__device__ __forceinline__ bool level_point(const Grid3& g, int L, int kmin, int& i, int& j,
                                            int& k) {
  j = blockIdx.x * blockDim.x + threadIdx.x;
  k = kmin + (int)blockIdx.y;
  if (j >= g.b || k >= g.c) return false;
  i = L - j - k;
  return i >= 0 && i < g.a;
}

// bad (may be null): the smallest point index whose pivot is zero, not
// finite, or whose FP32 inverse overflows (left at 0xffffffff if none).
template <typename CP>
__global__ void k_dilu_factor_level(CP cp, Grid3 g, int L, int kmin, int wrap, IluBlocks bl,
                                    double* __restrict__ d, float* __restrict__ dinv,
                                    unsigned* __restrict__ bad) {
  int i, j, k;
  if (!level_point(g, L, kmin, i, j, k)) return;
  const uint32_t n = (uint32_t)i + (uint32_t)g.a * ((uint32_t)j + (uint32_t)g.b * (uint32_t)k);
  double ak, aj, ai, aw, tk, tj, ti, tw;
  cp.lower(i, j, k, n, ak, aj, ai, aw);
  cp.lower_transpose(i, j, k, n, g, tk, tj, ti, tw);
  double dp = cp.diag(i, j, k, n);
  if (wrap && k == g.c - 1 && g.c > 1) {
    const uint32_t q = (uint32_t)i + (uint32_t)g.a * (uint32_t)j;
    dp = dp - aw / d[q] * tw;
  }
  if (bl.lo_k(k)) dp = dp - ak / d[n - g.ab] * tk;
  if (bl.lo_j(j)) dp = dp - aj / d[n - (uint32_t)g.a] * tj;
  if (bl.lo_i(i)) dp = dp - ai / d[n - 1] * ti;
  d[n] = dp;
  const float di = (float)(1.0 / dp);
  dinv[n] = di;
  if (bad && !(dp != 0.0 && isfinite(dp) && isfinite(di))) atomicMin(bad, n);
}

static void level_grid(const Grid3& g, int L, dim3& grid, int& kmin) {
  kmin = L - (g.a - 1) - (g.b - 1);
  if (kmin < 0) kmin = 0;
  int kmax = L < g.c - 1 ? L : g.c - 1;
  grid = dim3((g.b + 63) / 64, (unsigned)(kmax - kmin + 1), 1);
}

static IluBlocks no_blocks(const Grid3& g, bool wrap) {
  return IluBlocks{g.a, g.b, g.c, wrap ? 1 : 0};
}

template <typename CP>
static void dilu_factor_levels(const CP& cp, const Grid3& g, const IluBlocks& bl, double* d,
                               float* dinv, unsigned* bad, hipStream_t s) {
  const int nlev = g.a + g.b + g.c - 2;
  for (int L = 0; L < nlev; L++) {
    dim3 grid;
    int kmin;
    level_grid(g, L, grid, kmin);
    k_dilu_factor_level<CP><<<grid, 64, 0, s>>>(cp, g, L, kmin, bl.wrap, bl, d, dinv, bad);
  }
}

void launch_dilu_factor(const Metric& m, const Grid3& g, const IluBlocks& bl, double* d,
                        float* dinv, unsigned* bad, hipStream_t s) {
  MFCoefs cp;
  cp.m = m;
  dilu_factor_levels(cp, g, bl, d, dinv, bad, s);
}

void launch_dilu_factor_stored(const CoefArrays& ca, const float* dg, const Grid3& g, bool wrap,
                               double* d, float* dinv, hipStream_t s) {
  StoredCoefs cp;
  cp.ca = ca;
  cp.dg = dg;
  cp.a = g.a;
  cp.b = g.b;
  cp.c = g.c;
  cp.ab = g.ab;
  dilu_factor_levels(cp, g, no_blocks(g, wrap), d, dinv, nullptr, s);
}

// Forward sweep on one hyperplane:
//   y_p = r_p - sum_{q<p} a_pq (dinv_q y_q)    (order: wrap, k-1, j-1, i-1)
template <typename T, typename CP>
__global__ void k_fwd_level(CP cp, Grid3 g, int L, int kmin, int wrap, IluBlocks bl,
                            const float* __restrict__ dinv, const T* __restrict__ r,
                            float* __restrict__ yv, const CgState* guard) {
  if (guard && st_done(guard)) return;
  int i, j, k;
  if (!level_point(g, L, kmin, i, j, k)) return;
  const uint32_t n = (uint32_t)i + (uint32_t)g.a * ((uint32_t)j + (uint32_t)g.b * (uint32_t)k);
  double ak, aj, ai, aw;
  cp.lower(i, j, k, n, ak, aj, ai, aw);
  float y = (float)r[n];
  if (wrap && k == g.c - 1 && g.c > 1) {
    const uint32_t q = (uint32_t)i + (uint32_t)g.a * (uint32_t)j;
    y = sub_prod(y, (float)aw, (dinv[q] * yv[q]));
  }
  if (bl.lo_k(k)) y = sub_prod(y, (float)ak, (dinv[n - g.ab] * yv[n - g.ab]));
  if (bl.lo_j(j)) y = sub_prod(y, (float)aj, (dinv[n - g.a] * yv[n - g.a]));
  if (bl.lo_i(i)) y = sub_prod(y, (float)ai, (dinv[n - 1] * yv[n - 1]));
  yv[n] = y;
}

// Backward sweep on one hyperplane (in place over yv):
//   z_p = (y_p - sum_{q>p} a_pq z_q) dinv_p   (order: i+1, j+1, k+1, wrap)
template <typename T, typename CP>
__global__ void k_bwd_level(CP cp, Grid3 g, int L, int kmin, int wrap, IluBlocks bl,
                            const float* __restrict__ dinv, float* __restrict__ yv,
                            T* __restrict__ z, const CgState* guard) {
  if (guard && st_done(guard)) return;
  int i, j, k;
  if (!level_point(g, L, kmin, i, j, k)) return;
  const uint32_t n = (uint32_t)i + (uint32_t)g.a * ((uint32_t)j + (uint32_t)g.b * (uint32_t)k);
  double bi, bj, bk, bw;
  cp.upper(i, j, k, n, bi, bj, bk, bw);
  float s = yv[n];
  if (bl.hi_i(i, g.a)) s = sub_prod(s, (float)bi, yv[n + 1]);
  if (bl.hi_j(j, g.b)) s = sub_prod(s, (float)bj, yv[n + g.a]);
  if (bl.hi_k(k, g.c)) s = sub_prod(s, (float)bk, yv[n + g.ab]);
  if (wrap && k == 0 && g.c > 1) s = sub_prod(s, (float)bw, yv[n + (uint32_t)(g.c - 1) * g.ab]);
  const float zz = s * dinv[n];
  yv[n] = zz;
  z[n] = (T)zz;
}

template <typename T, typename CP>
static void dilu_solve_levels(const CP& cp, const Grid3& g, const IluBlocks& bl,
                              const float* dinv, const T* r, float* yv, T* z,
                              const CgState* guard, hipStream_t s) {
  const int nlev = g.a + g.b + g.c - 2;
  for (int L = 0; L < nlev; L++) {
    dim3 grid;
    int kmin;
    level_grid(g, L, grid, kmin);
    k_fwd_level<T, CP><<<grid, 64, 0, s>>>(cp, g, L, kmin, bl.wrap, bl, dinv, r, yv, guard);
  }
  for (int L = nlev - 1; L >= 0; L--) {
    dim3 grid;
    int kmin;
    level_grid(g, L, grid, kmin);
    k_bwd_level<T, CP><<<grid, 64, 0, s>>>(cp, g, L, kmin, bl.wrap, bl, dinv, yv, z, guard);
  }
}

template <typename T>
void launch_dilu_solve_levels(const Metric& m, const Grid3& g, const float* dinv, const T* r,
                              float* yv, T* z, const CoefArrays* ca, bool wrap,
                              const IluBlocks& bl, const CgState* guard, hipStream_t s) {
  if (ca) {
    StoredCoefs cp;
    cp.ca = *ca;
    cp.dg = nullptr;
    cp.a = g.a;
    cp.b = g.b;
    cp.c = g.c;
    cp.ab = g.ab;
    dilu_solve_levels<T>(cp, g, no_blocks(g, wrap), dinv, r, yv, z, guard, s);
  } else {
    MFSweep cp;
    cp.m = m;
    dilu_solve_levels<T>(cp, g, bl, dinv, r, yv, z, guard, s);
  }
}

// ---------------------------------------------------------------------------
// Post-solve phi (:3966-3969): unpack + r ghosts (vmask=1), pole averages
// over all nr shells, pole ghosts, phi seam.  Padded layout (nr,nt,np).
__global__ void k_phi_interior(Metric m, int nr, int nt, int np, const double* __restrict__ x,
                               const double* __restrict__ br0, const double* __restrict__ br1,
                               double dr1, double drn, double* __restrict__ phi) {
  // one thread per (J,K) interior column, J in [1,nt-2], K in [1,np-2]
  const int J = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int K = 1 + blockIdx.y;
  if (J > nt - 2 || K > np - 2) return;
  const int a = nr - 2, b = nt - 2;
  const size_t col = (size_t)nr * ((size_t)J + (size_t)nt * K);
  const size_t xcol = (size_t)a * ((size_t)(J - 1) + (size_t)b * (K - 1));
  for (int I = 1; I <= nr - 2; I++) phi[col + I] = x[xcol + (I - 1)];
  const double b0 = br0[(size_t)J + (size_t)nt * K];
  phi[col + 0] = phi[col + 1] - 1.0 * b0 * dr1;
  if (br1)
    phi[col + nr - 1] = phi[col + nr - 2] + 1.0 * br1[(size_t)J + (size_t)nt * K] * drn;
  else
    phi[col + nr - 1] = m.pm_r1 * phi[col + nr - 2];
}

__global__ void k_phi_polesum(int nr, int nt, int np, const double* __restrict__ phi,
                              const double* __restrict__ dph, double pl_i,
                              double* __restrict__ sums) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nr) return;
  double s0 = 0.0, s1 = 0.0;
  for (int K = 1; K <= np - 2; K++) {
    s0 = s0 + phi[(size_t)I + (size_t)nr * (1 + (size_t)nt * K)] * dph[K] * pl_i;
    s1 = s1 + phi[(size_t)I + (size_t)nr * ((nt - 2) + (size_t)nt * K)] * dph[K] * pl_i;
  }
  sums[I] = s0;
  sums[nr + I] = s1;
}

__global__ void k_phi_poleset(int nr, int nt, int np, const double* __restrict__ sums,
                              double* __restrict__ phi) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int K = 1 + blockIdx.y;
  if (I >= nr || K > np - 2) return;
  const size_t pl = (size_t)nr * (size_t)nt * K;
  phi[pl + I] = 2.0 * sums[I] - phi[pl + (size_t)nr * 1 + I];
  phi[pl + (size_t)nr * (nt - 1) + I] = 2.0 * sums[nr + I] - phi[pl + (size_t)nr * (nt - 2) + I];
}

__global__ void k_phi_seam(int nr, int nt, int np, double* __restrict__ phi) {
  const size_t plane = (size_t)nr * nt;
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= plane) return;
  phi[e] = phi[(size_t)(np - 2) * plane + e];
  phi[(size_t)(np - 1) * plane + e] = phi[plane + e];
}

void launch_fill_phi(const Metric& m, const Grid3& g, int nr, int nt, int np, const double* x,
                     const double* br0, const double* br1, double dr1, double drn,
                     const double* dph_full, double* phi, double* sums, hipStream_t s) {
  (void)g;
  dim3 g1((nt - 2 + 63) / 64, np - 2);
  k_phi_interior<<<g1, 64, 0, s>>>(m, nr, nt, np, x, br0, br1, dr1, drn, phi);
  const double pl_i = 1.0 / (2.0 * 3.1415926535897932 - 0.0);
  k_phi_polesum<<<(nr + 63) / 64, 64, 0, s>>>(nr, nt, np, phi, dph_full, pl_i, sums);
  dim3 g3((nr + 63) / 64, np - 2);
  k_phi_poleset<<<g3, 64, 0, s>>>(nr, nt, np, sums, phi);
  const size_t plane = (size_t)nr * nt;
  k_phi_seam<<<(unsigned)((plane + 255) / 256), 256, 0, s>>>(nr, nt, np, phi);
}

// ---------------------------------------------------------------------------
// getb + magnetic_energy (:6217-6231, :6276-6292): cell-averaged B^2/2 dV
// over the interior, B from staggered differences of phi, fused (no B
// arrays).  Reduces (Br^2, Bt^2, Bp^2) in FP64 with a fixed order.
__global__ void __launch_bounds__(kBlock)
    k_energy(EnergyMesh em, int nr, int nt, int np, const double* __restrict__ phi,
             double* partial, unsigned* counter, double* out3, int G, uint32_t Nint,
             FastDiv da, FastDiv dab) {
  const int Lb = logical_block(G);
  const uint32_t CH = (Nint + (uint32_t)G - 1) / (uint32_t)G;
  const uint32_t n0 = (uint32_t)Lb * CH, n1 = min(Nint, n0 + CH);
  const int a = nr - 2, b = nt - 2;
  const size_t sJ = (size_t)nr, sK = (size_t)nr * nt;
  double wr = 0.0, wt = 0.0, wp = 0.0;
  for (uint32_t n = n0 + threadIdx.x; n < n1; n += kBlock) {
    const uint32_t kk = fdiv(n, dab);
    const uint32_t rem = n - kk * (uint32_t)(a * b);
    const uint32_t jj = fdiv(rem, da);
    const int I = 1 + (int)(rem - jj * (uint32_t)a), J = 1 + (int)jj, K = 1 + (int)kk;
    const size_t c0 = (size_t)I + sJ * J + sK * K;
    const double pc = phi[c0];
    const double rh = em.rh[I];
    const double dv = rh * rh * em.drh[I] * em.dth[J] * em.sth[J] * em.dph[K];
    const double br_i = (phi[c0 + 1] - pc) / em.dr[I];
    const double br_im = (pc - phi[c0 - 1]) / em.dr[I - 1];
    const double bt_j = (phi[c0 + sJ] - pc) / (rh * em.dt[J]);
    const double bt_jm = (pc - phi[c0 - sJ]) / (rh * em.dt[J - 1]);
    const double bp_k = (phi[c0 + sK] - pc) / (rh * em.sth[J] * em.dp[K]);
    const double bp_km = (pc - phi[c0 - sK]) / (rh * em.sth[J] * em.dp[K - 1]);
    const double brav = 0.5 * (br_i + br_im);
    const double btav = 0.5 * (bt_j + bt_jm);
    const double bpav = 0.5 * (bp_k + bp_km);
    wr = wr + 0.5 * (brav * brav) * dv;
    wt = wt + 0.5 * (btav * btav) * dv;
    wp = wp + 0.5 * (bpav * bpav) * dv;
  }
  double v[3] = {wr, wt, wp};
  block_finish<3>(v, partial, counter, G, Lb, nullptr, EPI_STORE, out3);
}

void launch_energy(const EnergyMesh& em, int nr, int nt, int np, const double* phi,
                   double* partial, unsigned* counter, double* out3, int G, hipStream_t s) {
  const uint32_t a = nr - 2, b = nt - 2, c = np - 2;
  k_energy<<<G, kBlock, 0, s>>>(em, nr, nt, np, phi, partial, counter, out3, G, a * b * c,
                                make_fastdiv(a), make_fastdiv(a * b));
}

__global__ void k_getb(EnergyMesh em, int nr, int nt, int np, const double* __restrict__ phi,
                       double* __restrict__ br, double* __restrict__ bt, double* __restrict__ bp) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  const int J = blockIdx.y, K = blockIdx.z;
  if (I >= nr) return;
  const size_t sJ = (size_t)nr, sK = (size_t)nr * nt;
  const size_t c0 = (size_t)I + sJ * J + sK * K;
  if (I < nr - 1)
    br[(size_t)I + (size_t)(nr - 1) * ((size_t)J + (size_t)nt * K)] =
        (phi[c0 + 1] - phi[c0]) / em.dr[I];
  if (J < nt - 1)
    bt[(size_t)I + (size_t)nr * ((size_t)J + (size_t)(nt - 1) * K)] =
        (phi[c0 + sJ] - phi[c0]) / (em.rh[I] * em.dt[J]);
  if (K < np - 1)
    bp[(size_t)I + (size_t)nr * ((size_t)J + (size_t)nt * K)] =
        (phi[c0 + sK] - phi[c0]) / (em.rh[I] * em.sth[J] * em.dp[K]);
}

void launch_getb(const EnergyMesh& em, int nr, int nt, int np, const double* phi, double* br,
                 double* bt, double* bp, hipStream_t s) {
  dim3 grid((nr + 63) / 64, nt, np);
  k_getb<<<grid, 64, 0, s>>>(em, nr, nt, np, phi, br, bt, bp);
}

// ---------------------------------------------------------------------------
// explicit instantiations
#define P3D_INST(T)                                                                             \
  template void launch_pole_sums<T>(const Metric&, const Grid3&, const T*, double*,             \
                                    hipStream_t);                                               \
  template void launch_stencil<T>(const Metric&, const Grid3&, const Launch&, const T*, T*,     \
                                  const T*, int, const double*, CgState*, Reduce, int, Epi,     \
                                  double*, bool, hipStream_t, const TileMap&);                  \
  template void launch_cg_update<T>(const Metric&, const Grid3&, const Launch&, int, T*, T*,    \
                                    const T*, const T*, CgState*, Reduce, int, hipStream_t,     \
                                    double*, const double*, const TileMap&);                    \
  template void launch_p_update<T>(const Metric&, const Grid3&, const Launch&, int, T*,         \
                                   const T*, const T*, CgState*, hipStream_t, const TileMap&,   \
                                   double*, const double*, unsigned*);                          \
  template void launch_dot2<T>(const Grid3&, const Launch&, const T*, const T*, const T*,       \
                               const T*, CgState*, Reduce, int, Epi, double*, bool,             \
                               hipStream_t, const TileMap&);                                    \
  template void launch_point_prec<T>(const Metric&, const Grid3&, const Launch&, int, const T*, \
                                     T*, CgState*, Reduce, int, Epi, double*, hipStream_t);     \
  template void launch_sub<T>(const Grid3&, const Launch&, T*, const T*, hipStream_t,           \
                              const TileMap&);                                                  \
  template void launch_copy<T>(const Grid3&, const Launch&, T*, const T*, hipStream_t,          \
                               const TileMap&, int);                                            \
  template void launch_dilu_solve_levels<T>(const Metric&, const Grid3&, const float*,          \
                                            const T*, float*, T*, const CoefArrays*, bool,      \
                                            const IluBlocks&, const CgState*, hipStream_t);
P3D_INST(double)
P3D_INST(float)
template void launch_cast_scale<float, double>(const Grid3&, float*, const double*,
                                               const double*, const double*, int, hipStream_t,
                                               const TileMap&);
template void launch_cast_scale<double, float>(const Grid3&, double*, const float*,
                                               const double*, const double*, int, hipStream_t,
                                               const TileMap&);
template void launch_cast_scale<double, double>(const Grid3&, double*, const double*,
                                                const double*, const double*, int, hipStream_t,
                                                const TileMap&);

void size_stream_grids(Launch& L, int ncu, uint32_t N) {
  // a multiple of 8 (the blocks' XCD mapping), below L.G, or 0 (L.G)
  auto fit = [&](auto kern) {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, kBlock, 0) != hipSuccess ||
        nb < 1)
      return 0;
    const int G = nb * ncu / 8 * 8;
    return G >= 8 && G < L.G ? G : 0;
  };
  L.Gcg = fit(k_cg_update<double>);
  if (L.Gcg > 0) {
    const uint32_t ch = (N + (uint32_t)L.Gcg - 1) / (uint32_t)L.Gcg;
    L.CHcg = (int)((ch + kBlock - 1) / kBlock * kBlock);
  }
  L.Gp = fit(k_p_update<double>);
  L.Gd2 = fit(k_dot2<double>);
  L.Gpx[0] = fit(k_p_update_x<0>);
  L.Gpx[1] = fit(k_p_update_x<1>);
  L.Gpx[2] = fit(k_p_update_x<2>);
  L.Gpx[3] = fit(k_p_update_x<3>);
  L.Gru = fit(k_r_update);
}

}  // namespace p3d
