// Internal definitions shared by the gfx950 kernels and the host side of
// libpot3d_hip.so.  Layout conventions (identical to the reference's packed
// solver vectors, src/pot3d.F90:3907-3917, 5323):
//   interior grid a=nr-2 (r), b=nt-2 (theta), c=np-2 (phi), N=a*b*c
//   packed index n = i + a*(j + b*k), 0-based i<a, j<b, k<c (r fastest)
// All 1-D metric factors live in one device table (MetricTables).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace p3d {

// Separable stencil factors (FP64).  For interior point (i,j,k):
//   a1 = -(DRH[i]*CP[j]) *DPL[k]   coupling to (i,j,k-1)
//   a7 = -(DRH[i]*CP[j]) *DPU[k]   coupling to (i,j,k+1)
//   a2 = -(DRH[i]*DPH[k])*CTL[j]   coupling to (i,j-1,k)
//   a6 = -(DRH[i]*DPH[k])*CTH[j]   coupling to (i,j+1,k)
//   a3 = -(SD[j]*DPH[k]) *CRL[i]   coupling to (i-1,j,k)
//   a5 = -(SD[j]*DPH[k]) *CRH[i]   coupling to (i+1,j,k)
//   a4 = -(a1+a2+a3+a5+a6+a7)
// with DRH=drh(i), CRL=r(i-1)^2/dr(i-1), CRH=r(i)^2/dr(i), SD=sth*dth,
// CP=dth/sth, CTL=st(j-1)/dt(j-1), CTH=st(j)/dt(j), DPH=dph(k),
// DPL=1/dp(k-1), DPU=1/dp(k)  -- the factors of load_matrix_pot3d_solve
// (src/pot3d.F90:4590-4611).  Because CRH[i]==CRL[i+1], CTH[j]==CTL[j+1]
// and DPU[k]==DPL[k+1], a5(i)==a3(i+1), a6(j)==a2(j+1), a7(k)==a1(k+1)
// bitwise: the operator is exactly symmetric, as the reference's is.
struct Metric {
  const double* DRH;  // a
  const double* CRL;  // a
  const double* CRH;  // a
  const double* SD;   // b
  const double* CP;   // b
  const double* CTL;  // b
  const double* CTH;  // b
  const double* DPH;  // c
  const double* DPL;  // c
  const double* DPU;  // c
  const double* PW;   // c: dph(k)/(2 pi), pole-average weights
  // DPH, DPL, DPU rounded to float: the per-plane factors of the FP32
  // couplings of the ILU triangular solves (sweep_coup_*)
  const float* FDPH;  // c
  const float* FDPL;  // c
  const float* FDPU;  // c
  int a, b, c;
  double pm_r1;       // r1 ghost factor (+1 Neumann, -1 Dirichlet)
};

// FP32 couplings of the D-ILU triangular solves.  A's off-diagonals split
// into a per-column factor and a per-plane factor; each is rounded to float
// once and the coupling is their float product:
//   forward (lower)   k-1: K*FDPL[k]   j-1: J*FDPH[k]   i-1: I*FDPH[k]
//                     wrap (k = c-1 to plane 0): K*FDPU[c-1]
//   backward (upper)  k+1: K*FDPU[k]   j+1: J*FDPH[k]   i+1: I*FDPH[k]
//                     wrap (k = 0 to plane c-1): K*FDPL[0]
// with K = -(DRH CP), J = -(DRH CTL) or -(DRH CTH), I = -(SD CRL) or
// -(SD CRH) in FP64 before the rounding.  A march step then forms a
// coupling with one float multiply.  The factorisation (the pivots) uses
// A's FP64 coefficients; every sweep schedule uses these couplings, so the
// schedules agree bit for bit.
struct SweepCoup {
  float k, j, i;
};
__device__ __forceinline__ SweepCoup sweep_coup_lower(const Metric& m, int i, int j) {
  const double drh = m.DRH[i];
  return SweepCoup{(float)(-(drh * m.CP[j])), (float)(-(drh * m.CTL[j])),
                   (float)(-(m.SD[j] * m.CRL[i]))};
}
__device__ __forceinline__ SweepCoup sweep_coup_upper(const Metric& m, int i, int j) {
  const double drh = m.DRH[i];
  return SweepCoup{(float)(-(drh * m.CP[j])), (float)(-(drh * m.CTH[j])),
                   (float)(-(m.SD[j] * m.CRH[i]))};
}

struct Coef {
  double a1, a2, a3, a4, a5, a6, a7;
};

__device__ __forceinline__ Coef coef_at(const Metric& m, int i, int j, int k) {
  Coef c;
  const double drh = m.DRH[i];
  const double p1 = drh * m.CP[j];
  const double dph = m.DPH[k];
  const double p2 = drh * dph;
  const double p3 = m.SD[j] * dph;
  c.a1 = -(p1 * m.DPL[k]);
  c.a7 = -(p1 * m.DPU[k]);
  c.a2 = -(p2 * m.CTL[j]);
  c.a6 = -(p2 * m.CTH[j]);
  c.a3 = -(p3 * m.CRL[i]);
  c.a5 = -(p3 * m.CRH[i]);
  c.a4 = -(c.a1 + c.a2 + c.a3 + c.a5 + c.a6 + c.a7);
  return c;
}

// n -> (i,j,k) with 32-bit magic-number division, exact for n < 2^31
// (shift = ceil(log2 d), mul = floor(2^32 (2^shift - d)/d) + 1).
struct FastDiv {
  uint32_t d, mul, shift;
};

__host__ inline FastDiv make_fastdiv(uint32_t d) {
  FastDiv f;
  f.d = d;
  uint32_t s = 0;
  while ((1ull << s) < d) s++;
  f.shift = s;
  uint64_t one = 1;
  f.mul = (uint32_t)(((one << 32) * ((one << s) - d)) / d + 1);
  return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
  return (__umulhi(n, f.mul) + n) >> f.shift;
}

// Block-Jacobi cut of the ILU matrix: blocks of bi x bj x bk points with
// block starts at multiples of the extents; every coupling between two
// blocks is dropped, and the phi wrap is kept only with a single phi block.
// This is the ILU0 the reference builds on a decomposed grid: each rank's
// diacsr drops the couplings at its local r and theta boundaries and, when
// nproc_p > 1, the phi ones (src/pot3d.F90:4940-4967).  Extents >= (a,b,c):
// no cut.
// y - a*b rounded once: the update every D-ILU triangular-solve schedule
// applies to a point, written as an explicit fused multiply-add so that the
// rounding does not depend on where the compiler places the product (all
// schedules must agree bit for bit).
__device__ __forceinline__ float sub_prod(float y, float a, float b) {
  return __builtin_fmaf(-a, b, y);
}

struct IluBlocks {
  int bi, bj, bk;
  int wrap;  // phi wrap kept
  __device__ __forceinline__ bool lo_i(int i) const { return i > 0 && i % bi != 0; }
  __device__ __forceinline__ bool lo_j(int j) const { return j > 0 && j % bj != 0; }
  __device__ __forceinline__ bool lo_k(int k) const { return k > 0 && k % bk != 0; }
  __device__ __forceinline__ bool hi_i(int i, int a) const { return i < a - 1 && (i + 1) % bi != 0; }
  __device__ __forceinline__ bool hi_j(int j, int b) const { return j < b - 1 && (j + 1) % bj != 0; }
  __device__ __forceinline__ bool hi_k(int k, int c) const { return k < c - 1 && (k + 1) % bk != 0; }
};

struct Grid3 {
  int a, b, c;
  uint32_t N, ab;
  FastDiv da, dab;
};

__device__ __forceinline__ void unflatten(const Grid3& g, uint32_t n, int& i, int& j, int& k) {
  uint32_t kk = fdiv(n, g.dab);
  uint32_t rem = n - kk * g.ab;
  uint32_t jj = fdiv(rem, g.da);
  i = (int)(rem - jj * (uint32_t)g.a);
  j = (int)jj;
  k = (int)kk;
}

// Placement of the vectors the single-wave ILU tile sweeps read and write
// (the CG residual r and preconditioned residual z): with on != 0, point
// (i, j, k) sits at ((tile << tk_log2) + q) * 64 + 8 jj + ii, tile = I + TNI
// (J + TNJ Kc), for I = i/8, ii = i%8, J = j/8, jj = j%8, Kc = k/TK,
// q = k%TK (TK = 1 << tk_log2): every 8x8xTK tile of the sweeps is one
// contiguous block of whole 64-point planes.  Entries of the block that are
// outside the grid stay zero.  on == 0: natural order.
struct TileMap {
  int on = 0;
  int TNI = 0, TNJ = 0;
  int tk_log2 = 0;
};

__device__ __forceinline__ uint32_t tmap(const TileMap& tm, const Grid3& g, uint32_t n) {
  if (!tm.on) return n;
  int i, j, k;
  unflatten(g, n, i, j, k);
  const uint32_t tile =
      (uint32_t)(i >> 3) +
      (uint32_t)tm.TNI * ((uint32_t)(j >> 3) + (uint32_t)tm.TNJ * ((uint32_t)k >> tm.tk_log2));
  const uint32_t q = (uint32_t)k & ((1u << tm.tk_log2) - 1u);
  return (((tile << tm.tk_log2) + q) << 6) + (uint32_t)(((j & 7) << 3) | (i & 7));
}

// Device-side solver state (one per solve, in device memory).
struct CgState {
  double bdotb;      // b . M^-1 b
  double rdotr;      // current r . z
  double rdotr_old;
  double pdotap;
  double alpha;
  double beta;
  double epsn;       // sign(rdotr/bdotb)*sqrt(|rdotr/bdotb|)
  double rr;         // r . r (two-norm residual squared)
  double bb;         // b . b
  double relres;     // sqrt(rr/bb)
  double epscg;
  double* hist;      // epsn per iteration (index ncg), may be null
  int hist_cap;
  int ncg;
  int ncgmax;
  int done;          // 0 running, 1 converged, 2 ncgmax reached,
                     // 3 non-finite residual, 4 zero right-hand side
  int stop_mode;     // 0: epsn (reference), 1: ||r||2/||b||2
  int x_ncg;         // iterations whose FP64 solution update is applied (FP32 recurrence)
  // FP32 recurrence with the FP64 solution updated every kXBatch
  // iterations: alpha of each iteration of the current group, and the
  // buffer holding the current search direction
  double alpha_hist[4];
  int p_cur;
};

// Iterations per FP64 solution update of the FP32 recurrence (the p
// update's group; directions of a group live in kXBatch rotating buffers).
constexpr int kXBatch = 4;
struct PBuf {
  float* p[kXBatch];
};

// Control-state reads in kernels go through agent-scope atomic loads, so
// every workgroup on every XCD sees the value the previous kernel's
// epilogue stored (no stale L2 line can make part of a grid skip work and
// break the last-block / last-wave protocols).
__device__ __forceinline__ int st_done(const CgState* st) {
  return __hip_atomic_load(&st->done, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ int st_load_int(const int* f) {
  return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double st_load(const double* f) {
  return __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Reduction slots used by the "last block finishes" pattern.
struct Reduce {
  double* partial;   // [3 * max_blocks + 3 * 256]: block, then group partials
  unsigned int* counter;  // 256 counters per reduction site (two-level reduction)
  int max_blocks;
};

}  // namespace p3d
