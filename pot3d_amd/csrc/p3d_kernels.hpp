// Host-side launchers of the gfx950 kernels (p3d_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "p3d_internal.hpp"

namespace p3d {

enum Prec { PREC_NONE = 0, PREC_DIAG = 1, PREC_ILU = 2 };

// Epilogues run by the last block of a fused reduction (one thread).
enum Epi {
  EPI_NONE = 0,
  EPI_BDOTB = 1,   // s0=b.z, s1=b.b        -> bdotb, bb (zero rhs -> done=4)
  EPI_RINIT = 2,   // s0=r.z, s1=r.r        -> rdotr, epsn, convergence at ncg=0
  EPI_PAP = 3,     // s0=p.ap               -> ncg++, pdotap, alpha
  EPI_RITER = 4,   // s0=r.z, s1=r.r        -> rdotr, epsn, convergence, beta
  EPI_RR = 5,      // s1=r.r                -> rr (ILU path; r.z comes later)
  EPI_RZITER = 6,  // s0=r.z (rr from EPI_RR) -> as EPI_RITER
  EPI_STORE = 7,   // s0,s1 -> out[0], out[1] (plain dot products)
  EPI_REPLACE = 8, // s0=r.z, s1=r.r        -> rdotr, rr, beta (residual replaced)
};

// ---------------------------------------------------------------------------
// CG scalar logic (cgsolve/ernorm), run by one thread after a reduction.
__device__ inline void cg_check(CgState* st) {
  const double epssq = st->rdotr / st->bdotb;
  const double epsn = copysign(1.0, epssq) * sqrt(fabs(epssq));
  st->epsn = epsn;
  st->relres = sqrt(st->rr / st->bb);
  if (st->hist && st->ncg < st->hist_cap) st->hist[st->ncg] = epsn;
  const double crit = st->stop_mode == 0 ? epsn : st->relres;
  if (!isfinite(epsn) || !isfinite(st->relres)) {
    st->done = 3;
  } else if (crit < st->epscg) {
    st->done = 1;
  } else if (st->ncg >= st->ncgmax) {
    st->done = 2;
  }
}

__device__ inline void cg_epilogue(CgState* st, Epi epi, double s0, double s1, double* out) {
  switch (epi) {
    case EPI_BDOTB:
      st->bdotb = s0;
      st->bb = s1;
      if (s0 == 0.0) st->done = 4;
      break;
    case EPI_RINIT:
      st->rdotr = s0;
      st->rr = s1;
      cg_check(st);
      break;
    case EPI_PAP:
      st->ncg += 1;
      st->pdotap = s0;
      st->alpha = st->rdotr / s0;
      break;
    case EPI_RR:
      st->rr = s1;
      break;
    case EPI_RITER:
    case EPI_RZITER:
      st->rdotr_old = st->rdotr;
      st->rdotr = s0;
      if (epi == EPI_RITER) st->rr = s1;
      cg_check(st);
      if (!st->done) st->beta = st->rdotr / st->rdotr_old;
      break;
    case EPI_REPLACE:
      // the recurrence stopped with rdotr_old = r.z of its last direction;
      // the replaced residual continues the same Krylov recurrence
      st->rdotr = s0;
      st->rr = s1;
      st->beta = s0 / st->rdotr_old;
      break;
    case EPI_STORE:
      if (out) {
        out[0] = s0;
        out[1] = s1;
      }
      break;
    default:
      break;
  }
}


struct Launch {
  int G;        // blocks of a streaming kernel (multiple of 8)
  int CH;       // points per block (multiple of 256)
  // operator kernel: 0 column march, 1 row-blocked march, 2 LDS-staged
  // patches (3 tall, 4 short), -1 the LDS-staged patches where they apply
  int stencil = -1;
  // grid-stride kernels whose registers allow fewer than G resident
  // workgroups (the grouped p update, t = 0..3, and the FP32 r update):
  // as many workgroups as the device holds at once (0: G), so that no
  // second wave of workgroups streams its share alone at the end
  int Gpx[4] = {0, 0, 0, 0};
  int Gru = 0;
  // the same for the FP64 CG update (chunks of CHcg points), p update and
  // dot products
  int Gcg = 0, CHcg = 0, Gp = 0, Gd2 = 0;
};

constexpr int kBlock = 256;
// two-level reductions (block_finish): blocks per group, counters per slot
// ([0] global + one per group), so at most kRedGroup*(kRedCounters-1) blocks
constexpr int kRedGroup = 128;
constexpr int kRedCounters = 256;
constexpr int kRedSlots = 16;
constexpr int kUnroll = 4;  // points per thread in flight in the streaming kernels
constexpr int kPoleSegs = 16;

Launch plan_stream(uint32_t N, int max_blocks);
// Fills L.Gpx, L.Gru, L.Gcg/CHcg, L.Gp and L.Gd2 from the kernels'
// occupancy on ncu CUs, for passes over N points.
void size_stream_grids(Launch& L, int ncu, uint32_t N);

// Pole averages (m=0 component) of x at j=0 and j=b-1, per interior i, as
// kPoleSegs partial sums each: part[(pole*kPoleSegs+seg)*a + i].
template <typename T>
void launch_pole_sums(const Metric& m, const Grid3& g, const T* x, double* part, hipStream_t s);

// Workgroups of the operator launch (its reductions need that many slots);
// tm: the placement of its output (natural by default).
int stencil_blocks(const Grid3& g, const TileMap& tm = TileMap{});

// y = A x (homogeneous boundary values; mode 0) or y = b - A x (mode 1).
// Reduces s0 = x.y (mode 0) or s0 = y.y (mode 1) and runs epilogue `epi`.
// tm places y (scatter; natural by default).
template <typename T>
void launch_stencil(const Metric& m, const Grid3& g, const Launch& L, const T* x, T* y,
                    const T* b, int mode, const double* pole_part, CgState* st, Reduce red,
                    int slot, Epi epi, double* out, bool guard, hipStream_t s,
                    const TileMap& tm = TileMap{});

// Residual update of the FP32 recurrence: r -= alpha ap over n_all
// tile-major entries (r.r is reduced with r.z after the preconditioner); Lt
// is the launch shape of n_all.
void launch_r_update(uint32_t n_all, const Launch& Lt, float* r, const float* ap,
                     const CgState* st, hipStream_t s);

// CG update x += alpha p, r -= alpha ap, and for PREC_DIAG/NONE the reduction
// of r.z and r.r with epilogue EPI_RITER; for PREC_ILU only r.r (EPI_RR).
// With xs non-null the solution update goes to the FP64 vector xs instead:
// xs += (alpha * *xscale) p (x is not touched).
template <typename T>
void launch_cg_update(const Metric& m, const Grid3& g, const Launch& L, int prec, T* x, T* r,
                      const T* p, const T* ap, CgState* st, Reduce red, int slot,
                      hipStream_t s, double* xs = nullptr, const double* xscale = nullptr,
                      const TileMap& tm = TileMap{});

// p update of the FP32 recurrence at iteration t of a group of kXBatch,
// with the group's FP64 solution updates at its last iteration (or when the
// recurrence has stopped): k_p_update_x in p3d_kernels.hip.
void launch_p_update_x(const Grid3& g, const Launch& L, int t, const PBuf& pb, const float* z,
                       CgState* st, hipStream_t s, const TileMap& tm, double* xs,
                       const double* xscale, unsigned* counter);

// p = z + beta p (z = M^-1 r computed in place for DIAG/NONE, read for ILU).
// tm: placement of r (cg_update) and z (p_update) (natural by default).
// With xs (FP32 recurrence, FP64 solution) it first applies the pending
// solution update xs += (alpha * *xscale) p, once per iteration
// (CgState::x_ncg; counter: a reduction slot's counter).
template <typename T>
void launch_p_update(const Metric& m, const Grid3& g, const Launch& L, int prec, T* p,
                     const T* r, const T* z, CgState* st, hipStream_t s,
                     const TileMap& tm = TileMap{}, double* xs = nullptr,
                     const double* xscale = nullptr, unsigned* counter = nullptr);

// Generic fused dot: s0 = x.y, s1 = u.v (v/u may be null), epilogue epi.
// All four vectors are placed by tm; the sums run in natural point order.
template <typename T>
void launch_dot2(const Grid3& g, const Launch& L, const T* x, const T* y, const T* u,
                 const T* v, CgState* st, Reduce red, int slot, Epi epi, double* out,
                 bool guard, hipStream_t s, const TileMap& tm = TileMap{});

// z = M^-1 r for DIAG/NONE (r,z may alias), with optional s0=r.z, s1=r.r.
template <typename T>
void launch_point_prec(const Metric& m, const Grid3& g, const Launch& L, int prec, const T* r,
                       T* z, CgState* st, Reduce red, int slot, Epi epi, double* out,
                       hipStream_t s);

// Residual init: r = r - ap (r placed by tm), and p = z later.
template <typename T>
void launch_sub(const Grid3& g, const Launch& L, T* r, const T* ap, hipStream_t s,
                const TileMap& tm = TileMap{});

// dst = src over the N points; with a tile map, gather (dst natural, src
// placed by tm) or scatter (scatter = 1: dst placed by tm, src natural).
template <typename T>
void launch_copy(const Grid3& g, const Launch& L, T* dst, const T* src, hipStream_t s,
                 const TileMap& tm = TileMap{}, int scatter = 0);

// op 0: dst = src * num/den, dst placed by tm; op 1: dst += src * den/num.
template <typename TD, typename TS>
void launch_cast_scale(const Grid3& g, TD* dst, const TS* src, const double* scale_num,
                       const double* scale_den, int op, hipStream_t s,
                       const TileMap& tm = TileMap{});

void launch_fill(double* p, double v, size_t n, hipStream_t s);

// RHS from the boundary data (closed form of the boundary trick).
void launch_rhs(const Metric& m, const Grid3& g, const double* br0, const double* br1,
                double dr1, double drn, int nt, double* rhs, hipStream_t s);

// D-ILU pivots (FP64, hyperplane level launches) -> dinv (FP32), with the
// block-Jacobi cut bl.  bad (may be null, preset to 0xffffffff) receives the
// smallest point index whose pivot is zero or not finite or whose FP32
// inverse overflows.
void launch_dilu_factor(const Metric& m, const Grid3& g, const IluBlocks& bl, double* d,
                        float* dinv, unsigned* bad, hipStream_t s);

// Stored couplings of the legacy CSR input (FP32, N each unless noted):
// lo_k/lo_j/lo_i couple p to (i,j,k-1)/(i,j-1,k)/(i-1,j,k), up_i/up_j/up_k
// to (i+1,j,k)/(i,j+1,k)/(i,j,k+1); zero where the reference drops the
// coupling.  wrap_lo[i+a*j]: (i,j,c-1)->(i,j,0); wrap_up[i+a*j]: (i,j,0)->
// (i,j,c-1).
struct CoefArrays {
  const float* lo_k;
  const float* lo_j;
  const float* lo_i;
  const float* up_i;
  const float* up_j;
  const float* up_k;
  const float* wrap_lo;
  const float* wrap_up;
};

// Level-scheduled D-ILU sweeps: r (T) -> z (T); yv is FP32 scratch of N.
// bl: block-Jacobi cut (matrix-free coefficients only; ca != null ignores it).
template <typename T>
void launch_dilu_solve_levels(const Metric& m, const Grid3& g, const float* dinv, const T* r,
                              float* yv, T* z, const CoefArrays* ca, bool wrap,
                              const IluBlocks& bl, const CgState* guard, hipStream_t s);

void launch_dilu_factor_stored(const CoefArrays& ca, const float* dg, const Grid3& g, bool wrap,
                               double* d, float* dinv, hipStream_t s);

// Post-solve phi (nr,nt,np) from packed x: unpack, boundary values with
// vmask=1, pole averages, phi seam.  sums: device scratch of 2*nr doubles.
void launch_fill_phi(const Metric& m, const Grid3& g, int nr, int nt, int np, const double* x,
                     const double* br0, const double* br1, double dr1, double drn,
                     const double* dph_full, double* phi, double* sums, hipStream_t s);

// Magnetic energy of phi (getb fused with magnetic_energy).
struct EnergyMesh {
  const double *dr, *rh, *drh, *dt, *dth, *sth, *dp, *dph;
};
void launch_energy(const EnergyMesh& em, int nr, int nt, int np, const double* phi,
                   double* partial, unsigned* counter, double* out3, int G, hipStream_t s);
void launch_getb(const EnergyMesh& em, int nr, int nt, int np, const double* phi, double* br,
                 double* bt, double* bp, hipStream_t s);

}  // namespace p3d
