// D-ILU triangular solves on the tile-diagonal schedule (see p3d_sweep.hpp).
//
// Arithmetic per point is identical to the hyperplane version in
// p3d_kernels.hip (same coefficients, same operation order):
//   forward  y_p = r_p - aw t_w - ak t_(k-1) - aj t_(j-1) - ai t_(i-1),
//            t_q = dinv_q y_q
//   backward z_p = (y_p - bi z_(i+1) - bj z_(j+1) - bk z_(k+1) - bw z_w) dinv_p
// so both schedules give bit-identical results.
#include <algorithm>
#include <type_traits>
#include <vector>

#include "p3d_kernels.hpp"
#include "p3d_sweep.hpp"

namespace p3d {

namespace {

constexpr int TI = 8, TJ = 8;
constexpr int SKEW = TI + TJ - 2;  // steps between the first and last lane of a tile
// planes a wave loads per round when it stages a tile (rounds of 32 for
// 32-plane tiles measured no faster)
constexpr int kTileRound = 16;
constexpr int U = 16;              // steps per prefetch block
}  // namespace
constexpr int kTileK = 32;         // default planes per tile of the tile-diagonal sweep
namespace {

// Neighbour exchange inside one wave holding an 8x8 block of columns, lane
// l = ii + 8 jj.  lanes_from_lower gives lane l the values of lanes l-1 and
// l-8, lanes_from_upper those of l+1 and l+8 (meaningful where those lanes
// are in the block; the edge lanes take a face value instead).  The i step
// is a DPP row shift (VALU).  The j step crosses 16-lane rows: one
// ds_bpermute through the LDS crossbar, no VALU work.  A VALU-only version
// (DPP rotate by 8, then the gfx950 permlane16/32 swaps and two selects)
// measured 1-2% slower in the tile sweeps, whose SIMDs are kept busy by the
// resident waves' VALU work.
#define P3D_DPP(v, ctrl) __builtin_amdgcn_mov_dpp((v), (ctrl), 0xf, 0xf, true)
constexpr int kDppRowShl1 = 0x101, kDppRowShr1 = 0x111;

__device__ __forceinline__ void lanes_from_lower(float t, float& from_i, float& from_j) {
  const int v = __float_as_int(t);
  const int lane = threadIdx.x & 63;
  from_i = __int_as_float(P3D_DPP(v, kDppRowShr1));
  from_j = __int_as_float(__builtin_amdgcn_ds_bpermute(((lane - 8) & 63) << 2, v));
}

__device__ __forceinline__ void lanes_from_upper(float t, float& from_i, float& from_j) {
  const int v = __float_as_int(t);
  const int lane = threadIdx.x & 63;
  from_i = __int_as_float(P3D_DPP(v, kDppRowShl1));
  from_j = __int_as_float(__builtin_amdgcn_ds_bpermute(((lane + 8) & 63) << 2, v));
}

// ---------------------------------------------------------------------------
// Tile-diagonal sweeps: the (i,j,k) grid is cut into 8x8xTK tiles and one
// launch processes every tile on one anti-diagonal I+J+K (reversed for the
// backward sweep).  A tile's inputs from its lower neighbours were finished
// by earlier launches, so there is no waiting between workgroups at all;
// inside a tile the lanes march through the TK planes: lane (ii,jj) handles
// plane k = s - ii - jj at step s, so its (i-1) and (j-1) neighbours hand
// their values over one step earlier by cross-lane shuffle.
// Coefficient sources of the tile sweeps: matrix-free (metric factors,
// the solver) or stored FP32 couplings (the legacy CSR input).
struct TileMF {
  Metric m;
  SweepCoup lo, up;
  __device__ __forceinline__ void init(int ic, int jc, uint32_t) {
    lo = sweep_coup_lower(m, ic, jc);
    up = sweep_coup_upper(m, ic, jc);
  }
  __device__ __forceinline__ float aw(uint32_t) const { return lo.k * m.FDPU[m.c - 1]; }
  __device__ __forceinline__ float bw(uint32_t) const { return up.k * m.FDPL[0]; }
  __device__ __forceinline__ void lower(int k, uint32_t, float& ak, float& aj, float& ai) const {
    ak = lo.k * m.FDPL[k];
    aj = lo.j * m.FDPH[k];
    ai = lo.i * m.FDPH[k];
  }
  __device__ __forceinline__ void upper(int k, uint32_t, float& bk, float& bj, float& bi) const {
    bk = up.k * m.FDPU[k];
    bj = up.j * m.FDPH[k];
    bi = up.i * m.FDPH[k];
  }
};

struct TileStored {
  CoefArrays ca;
  __device__ __forceinline__ void init(int, int, uint32_t) {}
  __device__ __forceinline__ float aw(uint32_t col) const { return ca.wrap_lo[col]; }
  __device__ __forceinline__ float bw(uint32_t col) const { return ca.wrap_up[col]; }
  __device__ __forceinline__ void lower(int, uint32_t p, float& ak, float& aj, float& ai) const {
    ak = ca.lo_k[p];
    aj = ca.lo_j[p];
    ai = ca.lo_i[p];
  }
  __device__ __forceinline__ void upper(int, uint32_t p, float& bk, float& bj, float& bi) const {
    bk = ca.up_k[p];
    bj = ca.up_j[p];
    bi = ca.up_i[p];
  }
};

template <typename T, typename CS>
__device__ void fwd_tile(CS cs, const Grid3& g, int I, int J, int k0, int k1,
                         const float* __restrict__ dinv, const T* __restrict__ r,
                         float* __restrict__ yv) {
  const int lane = threadIdx.x;
  const int ii = lane & (TI - 1), jj = lane / TI;
  const int i = I * TI + ii, j = J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  const uint32_t col = (uint32_t)ic + (uint32_t)g.a * (uint32_t)jc;
  cs.init(ic, jc, col);
  const float aw = (k1 == g.c) ? cs.aw(col) : 0.f;
  float t_own = 0.f, t_w = 0.f, t_pub = 0.f;
  if (act && k0 > 0) {
    const uint32_t q = col + g.ab * (uint32_t)(k0 - 1);
    t_own = dinv[q] * yv[q];
  }
  if (act && k0 > 0 && k1 == g.c) t_w = dinv[col] * yv[col];
  const int S = (k1 - k0) + SKEW;
  for (int s0 = 0; s0 < S; s0 += U) {
    float rr[U], dd[U], fl[U], fd[U], ak[U], aj[U], ai[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + s0 + u - ii - jj;
      rr[u] = dd[u] = fl[u] = fd[u] = ak[u] = aj[u] = ai[u] = 0.f;
      if (act && k >= k0 && k < k1) {
        const uint32_t p = col + g.ab * (uint32_t)k;
        rr[u] = (float)r[p];
        dd[u] = dinv[p];
        cs.lower(k, p, ak[u], aj[u], ai[u]);
        if (ii == 0 && i > 0) fl[u] = dinv[p - 1] * yv[p - 1];
        if (jj == 0 && j > 0) fd[u] = dinv[p - g.a] * yv[p - g.a];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = k0 + s0 + u - ii - jj;
      float from_i, from_j;
      lanes_from_lower(t_pub, from_i, from_j);
      if (act && k >= k0 && k < k1) {
        float y = rr[u];
        if (k == g.c - 1) y = sub_prod(y, aw, t_w);
        if (k > 0) y = sub_prod(y, ak[u], t_own);
        if (j > 0) y = sub_prod(y, aj[u], (jj == 0 ? fd[u] : from_j));
        if (i > 0) y = sub_prod(y, ai[u], (ii == 0 ? fl[u] : from_i));
        yv[col + g.ab * (uint32_t)k] = y;
        const float t = dd[u] * y;
        if (k == 0) t_w = t;
        t_own = t;
        t_pub = t;
      }
    }
  }
}

template <typename T, typename CS>
__device__ void bwd_tile(CS cs, const Grid3& g, int I, int J, int k0, int k1,
                         const float* __restrict__ dinv, float* __restrict__ yv,
                         T* __restrict__ z) {
  const int lane = threadIdx.x;
  const int ii = lane & (TI - 1), jj = lane / TI;
  const int rii = TI - 1 - ii, rjj = TJ - 1 - jj;
  const int i = I * TI + ii, j = J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  const uint32_t col = (uint32_t)ic + (uint32_t)g.a * (uint32_t)jc;
  cs.init(ic, jc, col);
  const float bw = (k0 == 0) ? cs.bw(col) : 0.f;
  float t_own = 0.f, t_w = 0.f, t_pub = 0.f;
  if (act && k1 < g.c) t_own = yv[col + g.ab * (uint32_t)k1];
  if (act && k1 < g.c && k0 == 0) t_w = yv[col + g.ab * (uint32_t)(g.c - 1)];
  const int S = (k1 - k0) + SKEW;
  for (int s0 = 0; s0 < S; s0 += U) {
    float yy[U], dd[U], fr[U], fu[U], bk[U], bj[U], bi[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = (k1 - 1) - (s0 + u - rii - rjj);
      yy[u] = dd[u] = fr[u] = fu[u] = bk[u] = bj[u] = bi[u] = 0.f;
      if (act && k >= k0 && k < k1) {
        const uint32_t p = col + g.ab * (uint32_t)k;
        yy[u] = yv[p];
        dd[u] = dinv[p];
        cs.upper(k, p, bk[u], bj[u], bi[u]);
        if (ii == TI - 1 && i < g.a - 1) fr[u] = yv[p + 1];
        if (jj == TJ - 1 && j < g.b - 1) fu[u] = yv[p + g.a];
      }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int k = (k1 - 1) - (s0 + u - rii - rjj);
      float from_i, from_j;
      lanes_from_upper(t_pub, from_i, from_j);
      if (act && k >= k0 && k < k1) {
        float sv = yy[u];
        if (i < g.a - 1) sv = sub_prod(sv, bi[u], (ii == TI - 1 ? fr[u] : from_i));
        if (j < g.b - 1) sv = sub_prod(sv, bj[u], (jj == TJ - 1 ? fu[u] : from_j));
        if (k < g.c - 1) sv = sub_prod(sv, bk[u], t_own);
        if (k == 0) sv = sub_prod(sv, bw, t_w);
        const float zz = sv * dd[u];
        const uint32_t p = col + g.ab * (uint32_t)k;
        yv[p] = zz;
        z[p] = (T)zz;
        if (k == g.c - 1) t_w = zz;
        t_own = zz;
        t_pub = zz;
      }
    }
  }
}

struct TileArgs {
  int NI, NJ, TK;
  const int* tiles;  // tiles of this diagonal: I + NI*(J + NJ*Kc)
};

template <typename T, typename CS>
__global__ void __launch_bounds__(64)
    k_tile_fwd(CS cs, Grid3 g, TileArgs ta, const float* __restrict__ dinv,
               const T* __restrict__ r, float* __restrict__ yv, const CgState* guard) {
  if (guard && st_done(guard)) return;
  const int t = ta.tiles[blockIdx.x];
  const int I = t % ta.NI, J = (t / ta.NI) % ta.NJ, Kc = t / (ta.NI * ta.NJ);
  const int k0 = Kc * ta.TK, k1 = min(g.c, k0 + ta.TK);
  fwd_tile<T, CS>(cs, g, I, J, k0, k1, dinv, r, yv);
}

template <typename T, typename CS>
__global__ void __launch_bounds__(64)
    k_tile_bwd(CS cs, Grid3 g, TileArgs ta, const float* __restrict__ dinv,
               float* __restrict__ yv, T* __restrict__ z, const CgState* guard) {
  if (guard && st_done(guard)) return;
  const int t = ta.tiles[blockIdx.x];
  const int I = t % ta.NI, J = (t / ta.NI) % ta.NJ, Kc = t / (ta.NI * ta.NJ);
  const int k0 = Kc * ta.TK, k1 = min(g.c, k0 + ta.TK);
  bwd_tile<T, CS>(cs, g, I, J, k0, k1, dinv, yv, z);
}

// LDS-staged tile sweeps for the matrix-free factors.  The skewed march
// touches a different plane in every lane, so reading r/dinv/y straight from
// HBM inside the march makes each load instruction gather 64 cache lines and
// leaves the march waiting on memory at every block of steps.  Here the tile
// is staged first: plane by plane (8 rows of 8 consecutive points per
// instruction) into LDS, together with the two incoming faces and the
// per-plane phi factors; the march then reads only LDS and registers and
// writes its result back into the staged plane, and the tile is stored
// plane by plane at the end.  Per-point arithmetic is that of fwd_tile /
// bwd_tile, so the results are bit-identical to the other schedules.
// Tiles of anti-diagonal L = I+J+Kc in the plan's order (Kc outer, J
// inner): the block's tile is found by a short uniform loop over Kc instead
// of a table read, so the first memory round trip of a launch already
// carries the tile's data.
struct DiagArgs {
  int L, NI, NJ, NK;
  const int* tiles;  // this diagonal's tiles (block-Jacobi plans), else null
  IluBlocks bl;      // block-Jacobi cut (extents >= the grid: none)
  float* face_i;     // per tile [TK][8]: its last (forward) / first (backward)
  float* face_j;     //   column (face_i) and row (face_j) values, for the next tile
  const float* dinv_t;  // dinv, tile-major (tile, plane, lane); zero outside the grid
  float* yv_t;          // forward result y in FP32, tile-major
  float* face_k;        // per tile [64]: backward z of its first plane (for the tile below)
  float* face_kt;       // per column of tiles [64]: backward z of plane c-1 (phi seam)
  int count;            // tiles on this diagonal
};

__device__ __forceinline__ void diag_tile(const DiagArgs& da, int b, int& I, int& J, int& Kc) {
  if (da.tiles) {
    const int t = da.tiles[b];
    I = t % da.NI;
    J = (t / da.NI) % da.NJ;
    Kc = t / (da.NI * da.NJ);
    return;
  }
  const int kc_lo = max(0, da.L - (da.NI - 1) - (da.NJ - 1));
  const int kc_hi = min(da.NK - 1, da.L);
  I = J = Kc = 0;
  for (int kc = kc_lo; kc <= kc_hi; kc++) {
    const int rest = da.L - kc;
    const int jlo = max(0, rest - (da.NI - 1)), jhi = min(da.NJ - 1, rest);
    const int cnt = jhi - jlo + 1;
    if (b < cnt) {
      Kc = kc;
      J = jlo + b;
      I = rest - J;
      return;
    }
    b -= cnt;
  }
}

static DiagArgs diag_args(const SweepPlan& plan, int L) {
  return DiagArgs{L,       plan.TNI, plan.TNJ,    plan.NK,
                  plan.blocked ? plan.tiles + plan.diag_off[L] : nullptr,
                  plan.bl, plan.face_i, plan.face_j, plan.dinv_t, plan.yv_t,
                  plan.face_k, plan.face_kt, plan.diag_off[L + 1] - plan.diag_off[L]};
}

__device__ __forceinline__ size_t tile_id(const DiagArgs& da, int I, int J, int Kc) {
  return (size_t)I + (size_t)da.NI * ((size_t)J + (size_t)da.NJ * (size_t)Kc);
}

// LDS of one 8x8xTK tile.  With couplings staged (STAGED), the staging waves
// form the float couplings of every plane and lane; otherwise the marching
// wave forms them from the per-plane factors (less LDS per tile, so more
// tiles resident per CU).
//
// Plane q of the tile lives in row q+1; a march step reads and writes row
// clamp(q, -1, nk) + 1, so the steps outside the tile's planes land in row
// 0 or row nk+1 and one index serves every access of a step.  The row the
// steps before a lane's first plane use (row 0 forward, row nk+1 backward)
// holds the incoming k value of the tile (the plane below / above) as its
// right-hand side, pivot 1 and zero couplings and faces: those steps then
// reproduce exactly that value, which is what the lane's first plane takes
// as its k neighbour.  The other end row is scratch.
template <int TK, bool STAGED>
struct TileLds {
  float v[TK + 2][64];  // r (forward) or y (backward), overwritten by the result
  float d[TK + 2][64];  // dinv
  // couplings in k, j, i: forward A(p, lower), backward A(p, upper)
  float ck[STAGED ? TK + 2 : 1][STAGED ? 64 : 1];
  float cj[STAGED ? TK + 2 : 1][STAGED ? 64 : 1];
  float ci[STAGED ? TK + 2 : 1][STAGED ? 64 : 1];
  float2 pf[STAGED ? 1 : TK + 2];  // (FDPH, FDPL) forward, (FDPH, FDPU) backward
  float fi[TK + 2][TJ];  // i face: forward t(i-1) at ii==0, backward z(i+1) at ii==TI-1
  float fj[TK + 2][TI];  // j face: forward t(j-1) at jj==0, backward z(j+1) at jj==TJ-1
};

// The end rows of a tile (rows 0 and nk+1, see TileLds): in the row the
// march enters from, right-hand side t_in, pivot 1; the other row zero;
// zero faces, couplings and plane factors in both.  Wave 0 (the marching
// wave) writes them, lane by lane.
template <int TK, bool STAGED>
__device__ __forceinline__ void tile_end_rows(TileLds<TK, STAGED>& sh, int entry, int other,
                                              int lane, float t_in) {
  sh.v[entry][lane] = t_in;
  sh.d[entry][lane] = 1.f;
  sh.v[other][lane] = 0.f;
  sh.d[other][lane] = 0.f;
  if constexpr (STAGED) {
    sh.ck[entry][lane] = sh.cj[entry][lane] = sh.ci[entry][lane] = 0.f;
    sh.ck[other][lane] = sh.cj[other][lane] = sh.ci[other][lane] = 0.f;
  } else {
    if (lane == 0) {
      sh.pf[entry] = make_float2(0.f, 0.f);
      sh.pf[other] = make_float2(0.f, 0.f);
    }
  }
  if (lane < 8) {
    sh.fi[entry][lane] = sh.fj[entry][lane] = 0.f;
    sh.fi[other][lane] = sh.fj[other][lane] = 0.f;
  }
}

// r, z, dinv and the FP32 sweep vector are all tile-major (TileMap): a
// tile reads and writes whole 64-point planes.
// W waves per tile: with W = TK/8 every wave stages 8 planes (one round of
// loads in flight per wave) and their float couplings, wave 0 marches and
// all waves store; with W = 1 the marching wave stages the whole tile
// itself, in rounds of up to 16 planes, and forms the couplings during the
// march (a workgroup is one wave and about 20 KB of LDS, so several tiles
// march on one CU at once).  The faces a tile needs from its lower
// (forward) or upper (backward) neighbours in i and j are not read from
// the scattered neighbour points: every tile leaves its boundary column and
// row in compact per-tile buffers (face_i, face_j; 8 values per plane),
// which the next tile reads as one contiguous block per face.  The march
// step is branch-free: lane data moves by DPP / permlane (no LDS round trip
// on the dependency chain), and its operands for 8 steps are read from LDS
// in one batch.
template <typename T, int TK, int W>
__global__ void __launch_bounds__(64 * W)
    k_tile_fwd_lds(Metric m, Grid3 g, DiagArgs da, const T* __restrict__ r,
                   const CgState* guard) {
  constexpr bool kStaged = W > 1;
  constexpr int kQPW = TK / W;                    // planes staged per wave
  constexpr int kRound = kQPW < kTileRound ? kQPW : kTileRound;  // planes per round of loads
  __shared__ TileLds<TK, kStaged> sh;
  int I, J, Kc;
  diag_tile(da, blockIdx.x, I, J, Kc);
  const int k0 = Kc * TK, k1 = min(g.c, k0 + TK), nk = k1 - k0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ii = lane & (TI - 1), jj = lane / TI;
  const int i = I * TI + ii, j = J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  // The march applies every coupling: where the block-Jacobi cut (tile-
  // aligned: inside a tile only the first plane can start a phi block) or
  // the domain boundary drops one, the value it multiplies is zero (the
  // incoming face, or the plane below the tile, is zeroed after loading).
  const bool lo_k0 = da.bl.lo_k(k0);
  const bool wrap = da.bl.wrap != 0;
  const SweepCoup cl = sweep_coup_lower(m, ic, jc);
  // this tile's block of the tile-major FP32 arrays (plane q at tb + 64 q)
  constexpr size_t kTile = (size_t)TK * 64;
  const size_t tb = tile_id(da, I, J, Kc) * kTile + lane;
  // Every global load of the tile is issued in one batch before any of
  // its values is used: the loads are unconditional, at clamped in-bounds
  // addresses (a tile's arrays hold TK planes, its faces TK rows), and the
  // values not wanted are dropped afterwards.  The incoming faces, the
  // per-plane factors and the planes below and at the seam go first, the
  // tile's own planes next, the done flag last; a tile then waits for
  // memory once before its march.
  constexpr int kFPT = (TK * 8 + 64 * W - 1) / (64 * W);  // face entries per thread
  const bool face_i = I > 0 && da.bl.lo_i(I * TI), face_j = J > 0 && da.bl.lo_j(J * TJ);
  float fiv[kFPT], fjv[kFPT];
  {
    const size_t fbi = tile_id(da, max(I - 1, 0), J, Kc) * TK * 8;
    const size_t fbj = tile_id(da, I, max(J - 1, 0), Kc) * TK * 8;
#pragma unroll
    for (int e = 0; e < kFPT; e++) {
      const int f = min((int)threadIdx.x + e * 64 * W, TK * 8 - 1);
      fiv[e] = da.face_i[fbi + f];
      fjv[e] = da.face_j[fbj + f];
    }
  }
  float dphv = 0.f, dpxv = 0.f;
  if constexpr (!kStaged) {
    const int kq = min(k0 + (int)threadIdx.x, g.c - 1);
    dphv = m.FDPH[kq];
    dpxv = m.FDPL[kq];
  }
  // plane k0-1 (the last plane of the tile below) and plane 0 (seam)
  const size_t qo = tile_id(da, I, J, max(Kc - 1, 0)) * kTile + (TK - 1) * 64 + lane;
  const size_t qw = tile_id(da, I, J, 0) * kTile + lane;
  const float od = da.dinv_t[qo], oy = da.yv_t[qo];
  const float wd = da.dinv_t[qw], wy = da.yv_t[qw];
  // stage: planes wave*kQPW .. +kQPW-1, rounds of kRound (planes past nk
  // are the zero padding of the tile-major arrays; the march does not read
  // their rows)
#pragma unroll 1
  for (int q0 = wave * kQPW; q0 < (wave + 1) * kQPW; q0 += kRound) {
    float rv[kRound], dv[kRound];
#pragma unroll
    for (int u = 0; u < kRound; u++) {
      const size_t o = tb + (size_t)(q0 + u) * 64;
      dv[u] = da.dinv_t[o];
      rv[u] = (float)r[o];
    }
    // the done flag is read after the tile's loads and tested before any store
    if (guard && st_done(guard)) return;  // uniform over the block
#pragma unroll
    for (int u = 0; u < kRound; u++) {
      const int q = q0 + u;
      if (q < nk) {  // rows past nk are not read (row nk+1 is an end row)
        sh.v[q + 1][lane] = rv[u];
        sh.d[q + 1][lane] = dv[u];
      }
      if constexpr (kStaged) {
        if (q < nk) {
          const float dph = m.FDPH[k0 + q];
          sh.ck[q + 1][lane] = cl.k * m.FDPL[k0 + q];
          sh.cj[q + 1][lane] = cl.j * dph;
          sh.ci[q + 1][lane] = cl.i * dph;
        }
      }
    }
  }
  const float t_own = (wave == 0 && act && lo_k0) ? od * oy : 0.f;
  const float t_w = (wave == 0 && act && wrap && k0 > 0 && k1 == g.c) ? wd * wy : 0.f;
#pragma unroll
  for (int e = 0; e < kFPT; e++) {
    const int f = (int)threadIdx.x + e * 64 * W;
    if (f < nk * 8) {
      sh.fi[(f >> 3) + 1][f & 7] = face_i ? fiv[e] : 0.f;
      sh.fj[(f >> 3) + 1][f & 7] = face_j ? fjv[e] : 0.f;
    }
  }
  if constexpr (!kStaged) {
    if ((int)threadIdx.x < nk) sh.pf[threadIdx.x + 1] = make_float2(dphv, dpxv);
  }
  // the march enters every lane through row 0 with the plane below
  if (wave == 0) tile_end_rows(sh, 0, nk + 1, lane, t_own);
  __syncthreads();
  if (wave == 0) {
    // Lane (ii, jj) handles plane q = s - ii - jj at step s.  t_k carries
    // the lane's latest t = dinv y (the k-1 coupling of its next plane and
    // what its i+1 / j+1 neighbours read one step later); it is refreshed
    // on every step, valid or not, because a neighbour only reads it on
    // steps where both planes are inside the tile.  The plane below the
    // tile enters through row 0 (the steps before a lane's plane 0 reproduce
    // it).  Only tiles holding plane c-1 of a wrapped phi range take the
    // seam coupling (to plane 0).
    const bool wrap_tile = wrap && k1 == g.c;
    const float aw = wrap_tile ? cl.k * m.FDPU[m.c - 1] : 0.f;
    const int S = nk + SKEW;
    auto march = [&](auto seam) {
      constexpr bool kSeam = decltype(seam)::value;
      float t_k = t_own, tw = t_w;
      for (int s0 = 0; s0 < S; s0 += 8) {
        float rr[8], dd[8], fl[8], fd[8], ak[8], aj[8], ai[8];
        int rw[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int qs = min(max(s0 + u - ii - jj, -1), nk) + 1;
          rw[u] = qs;
          rr[u] = sh.v[qs][lane];
          dd[u] = sh.d[qs][lane];
          if constexpr (kStaged) {
            ak[u] = sh.ck[qs][lane];
            aj[u] = sh.cj[qs][lane];
            ai[u] = sh.ci[qs][lane];
          } else {
            const float2 pf = sh.pf[qs];
            ak[u] = cl.k * pf.y;
            aj[u] = cl.j * pf.x;
            ai[u] = cl.i * pf.x;
          }
          fl[u] = sh.fi[qs][jj];
          fd[u] = sh.fj[qs][ii];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int q = s0 + u - ii - jj;
          float from_i, from_j;
          lanes_from_lower(t_k, from_i, from_j);
          float y = rr[u];
          if constexpr (kSeam) {
            if (q == nk - 1) y = sub_prod(y, aw, tw);
          }
          // the k neighbour: the lane's previous step (the end row's value
          // before its first plane)
          y = sub_prod(y, ak[u], t_k);
          y = sub_prod(y, aj[u], (jj == 0 ? fd[u] : from_j));
          y = sub_prod(y, ai[u], (ii == 0 ? fl[u] : from_i));
          sh.v[rw[u]][lane] = y;
          const float tt = dd[u] * y;
          if constexpr (kSeam) {
            if (k0 == 0 && q == 0) tw = tt;
          }
          t_k = tt;
        }
      }
    };
    if (wrap_tile)
      march(std::true_type{});
    else
      march(std::false_type{});
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int u = 0; u < kQPW; u++) {
      const int q = wave * kQPW + u;
      if (q < nk) da.yv_t[tb + (size_t)q * 64] = sh.v[q + 1][lane];
    }
  }
  // outgoing faces t = dinv*y of the last column (ii = 7) and row (jj = 7)
  for (int f = (int)threadIdx.x; f < nk * 8; f += 64 * W) {
    const int fq = f >> 3, fe = f & 7;
    const size_t tf = (tile_id(da, I, J, Kc) * TK + fq) * 8 + fe;
    da.face_i[tf] = sh.d[fq + 1][TI - 1 + TI * fe] * sh.v[fq + 1][TI - 1 + TI * fe];
    da.face_j[tf] = sh.d[fq + 1][fe + TI * (TJ - 1)] * sh.v[fq + 1][fe + TI * (TJ - 1)];
  }
}

template <typename T, int TK, int W>
__global__ void __launch_bounds__(64 * W)
    k_tile_bwd_lds(Metric m, Grid3 g, DiagArgs da, T* __restrict__ z, const CgState* guard) {
  constexpr bool kStaged = W > 1;
  constexpr int kQPW = TK / W;
  constexpr int kRound = kQPW < kTileRound ? kQPW : kTileRound;
  __shared__ TileLds<TK, kStaged> sh;
  int I, J, Kc;
  diag_tile(da, blockIdx.x, I, J, Kc);
  const int k0 = Kc * TK, k1 = min(g.c, k0 + TK), nk = k1 - k0;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ii = lane & (TI - 1), jj = lane / TI;
  const int rii = TI - 1 - ii, rjj = TJ - 1 - jj;
  const int i = I * TI + ii, j = J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  // every coupling applied, as in the forward sweep; the upper neighbours
  // outside the grid are padding lanes, whose dinv (hence z) is zero
  const bool hi_k1 = da.bl.hi_k(k1 - 1, g.c), wrap = da.bl.wrap != 0;
  const SweepCoup cu = sweep_coup_upper(m, ic, jc);
  constexpr size_t kTile = (size_t)TK * 64;
  const size_t tb = tile_id(da, I, J, Kc) * kTile + lane;
  // one batch of loads, as in the forward sweep: incoming faces (z of the
  // first column / row of the upper neighbours), per-plane factors, the
  // compact k faces (plane k1, the first plane of the tile above, and
  // plane c-1 at the seam), the tile's planes, the done flag
  constexpr int kFPT = (TK * 8 + 64 * W - 1) / (64 * W);  // face entries per thread
  const bool face_i = I + 1 < da.NI && da.bl.hi_i(I * TI + TI - 1, g.a);
  const bool face_j = J + 1 < da.NJ && da.bl.hi_j(J * TJ + TJ - 1, g.b);
  float fiv[kFPT], fjv[kFPT];
  {
    const size_t fbi = tile_id(da, min(I + 1, da.NI - 1), J, Kc) * TK * 8;
    const size_t fbj = tile_id(da, I, min(J + 1, da.NJ - 1), Kc) * TK * 8;
#pragma unroll
    for (int e = 0; e < kFPT; e++) {
      const int f = min((int)threadIdx.x + e * 64 * W, TK * 8 - 1);
      fiv[e] = da.face_i[fbi + f];
      fjv[e] = da.face_j[fbj + f];
    }
  }
  float dphv = 0.f, dpxv = 0.f;
  if constexpr (!kStaged) {
    const int kq = min(k0 + (int)threadIdx.x, g.c - 1);
    dphv = m.FDPH[kq];
    dpxv = m.FDPU[kq];
  }
  const float fk = da.face_k[tile_id(da, I, J, min(Kc + 1, da.NK - 1)) * 64 + lane];
  const float fkt = da.face_kt[((size_t)I + (size_t)da.NI * J) * 64 + lane];
#pragma unroll 1
  for (int q0 = wave * kQPW; q0 < (wave + 1) * kQPW; q0 += kRound) {
    float yv8[kRound], dv[kRound];
#pragma unroll
    for (int u = 0; u < kRound; u++) {
      const size_t o = tb + (size_t)(q0 + u) * 64;
      yv8[u] = da.yv_t[o];
      dv[u] = da.dinv_t[o];
    }
    if (guard && st_done(guard)) return;  // uniform over the block
#pragma unroll
    for (int u = 0; u < kRound; u++) {
      const int q = q0 + u;
      // y outside the grid is never written: zero it (rows past nk are not
      // read; row nk+1 is an end row)
      if (q < nk) {
        sh.v[q + 1][lane] = act ? yv8[u] : 0.f;
        sh.d[q + 1][lane] = act ? dv[u] : 0.f;
      }
      if constexpr (kStaged) {
        if (q < nk) {
          const float dph = m.FDPH[k0 + q];
          sh.ck[q + 1][lane] = cu.k * m.FDPU[k0 + q];
          sh.cj[q + 1][lane] = cu.j * dph;
          sh.ci[q + 1][lane] = cu.i * dph;
        }
      }
    }
  }
  const float t_own = (wave == 0 && act && hi_k1) ? fk : 0.f;
  const float t_w = (wave == 0 && act && wrap && k1 < g.c && k0 == 0) ? fkt : 0.f;
#pragma unroll
  for (int e = 0; e < kFPT; e++) {
    const int f = (int)threadIdx.x + e * 64 * W;
    if (f < nk * 8) {
      sh.fi[(f >> 3) + 1][f & 7] = face_i ? fiv[e] : 0.f;
      sh.fj[(f >> 3) + 1][f & 7] = face_j ? fjv[e] : 0.f;
    }
  }
  if constexpr (!kStaged) {
    if ((int)threadIdx.x < nk) sh.pf[threadIdx.x + 1] = make_float2(dphv, dpxv);
  }
  // the march enters every lane through row nk+1 with the plane above
  if (wave == 0) tile_end_rows(sh, nk + 1, 0, lane, t_own);
  __syncthreads();
  if (wave == 0) {
    // as in the forward sweep, mirrored: q = (nk-1) - (s - rii - rjj), the
    // plane above the tile enters through row nk+1, and only tiles holding
    // plane 0 of a wrapped phi range take the seam coupling (to c-1)
    const bool wrap_tile = wrap && k0 == 0;
    const float bw = wrap_tile ? cu.k * m.FDPL[0] : 0.f;
    const int S = nk + SKEW;
    auto march = [&](auto seam) {
      constexpr bool kSeam = decltype(seam)::value;
      float t_k = t_own, tw = t_w;
      for (int s0 = 0; s0 < S; s0 += 8) {
        float yy[8], dd[8], fr[8], fu[8], bk[8], bj[8], bi[8];
        int rw[8];
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int qs = min(max((nk - 1) - (s0 + u - rii - rjj), -1), nk) + 1;
          rw[u] = qs;
          yy[u] = sh.v[qs][lane];
          dd[u] = sh.d[qs][lane];
          if constexpr (kStaged) {
            bk[u] = sh.ck[qs][lane];
            bj[u] = sh.cj[qs][lane];
            bi[u] = sh.ci[qs][lane];
          } else {
            const float2 pf = sh.pf[qs];
            bk[u] = cu.k * pf.y;
            bj[u] = cu.j * pf.x;
            bi[u] = cu.i * pf.x;
          }
          fr[u] = sh.fi[qs][jj];
          fu[u] = sh.fj[qs][ii];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const int q = (nk - 1) - (s0 + u - rii - rjj);
          float from_i, from_j;
          lanes_from_upper(t_k, from_i, from_j);
          float sv = yy[u];
          sv = sub_prod(sv, bi[u], (ii == TI - 1 ? fr[u] : from_i));
          sv = sub_prod(sv, bj[u], (jj == TJ - 1 ? fu[u] : from_j));
          // the k neighbour: the lane's previous step (the end row's value
          // before its first plane)
          sv = sub_prod(sv, bk[u], t_k);
          if constexpr (kSeam) {
            if (q == 0) sv = sub_prod(sv, bw, tw);
          }
          const float zz = sv * dd[u];
          sh.v[rw[u]][lane] = zz;
          if constexpr (kSeam) {
            if (k1 == g.c && q == nk - 1) tw = zz;
          }
          t_k = zz;
        }
      }
    };
    if (wrap_tile)
      march(std::true_type{});
    else
      march(std::false_type{});
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int u = 0; u < kQPW; u++) {
      const int q = wave * kQPW + u;
      if (q < nk) z[tb + (size_t)q * 64] = (T)sh.v[q + 1][lane];
    }
    if (wave == 0) {  // k faces: the first plane, and plane c-1 for the seam
      da.face_k[tile_id(da, I, J, Kc) * 64 + lane] = sh.v[1][lane];
      if (k1 == g.c) da.face_kt[((size_t)I + (size_t)da.NI * J) * 64 + lane] = sh.v[nk][lane];
    }
  }
  // outgoing faces: z of the first column (ii = 0) and row (jj = 0)
  for (int f = (int)threadIdx.x; f < nk * 8; f += 64 * W) {
    const int fq = f >> 3, fe = f & 7;
    const size_t tf = (tile_id(da, I, J, Kc) * TK + fq) * 8 + fe;
    da.face_i[tf] = sh.v[fq + 1][TI * fe];
    da.face_j[tf] = sh.v[fq + 1][fe];
  }
}

// ---------------------------------------------------------------------------
// Register tiles (SweepPlan::stage_one == 2): the single-wave tile sweeps
// with the tile held in VGPRs instead of LDS.  The sweeps' own arrays (dinv
// and the FP32 y, tile-major) are stored skewed: lane l = ii + 8 jj keeps
// plane q of its column in row (q + ii + jj) mod TK of the tile, so at march
// step s every lane needs row s mod TK (forward) or (SKEW - 1 - s) mod TK
// (backward), a compile-time register.  r (forward input) and z (backward
// output) keep the solver's unskewed tile-major placement; each lane
// rotates its own column by ii + jj planes in registers (four conditional
// stages).  LDS holds only the per-plane factors and the incoming faces
// (under 1.5 KB per tile), so the number of resident tiles per CU is set by
// the VGPRs.  Point arithmetic and operation order are those of the other
// schedules.
//
// Steps where a lane is outside its tile's planes leave its registers and
// its k value unchanged (a select on s against ii + jj); the last tile of a
// phi range is marched as a full TK-plane tile whose planes past the grid
// are the zero padding of the tile-major arrays (r, y, dinv and the
// per-plane factors all zero there, so they stay zero and feed zero).
//
// Faces: face row c of a tile holds, for each of its 8 entries, the value
// the consuming edge lane needs at its march step c (mod TK), so a lane
// reads row s mod TK at step s.  A producer's lane computes that value at
// its own step c + 7, i.e. from its register (c + 7) mod TK (forward) or
// (6 - c) mod TK (backward).  The k faces are the lane's k value after the
// march (t of plane TK-1 forward, z of plane 0 backward) in face_k, and the
// phi-seam values (t of plane 0 forward, z of plane c-1 backward) in
// face_kt.
template <int TK>
__device__ __forceinline__ void rotate_col(float (&x)[TK], int m, int dir) {
  // x[rho] <- x[(rho - dir*m) mod TK], m in [0, 15], in four conditional
  // stages.  Both operands of each select pass through an empty asm, so that
  // it stays a select of two values: a select of two array loads may be
  // turned into one load from a selected address, which would keep x in
  // memory (scratch) instead of registers.
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int sh = 1 << b;
    const bool on = (m >> b) & 1;
    float y[TK];
#pragma unroll
    for (int rr = 0; rr < TK; rr++) {
      const int src = dir > 0 ? (rr - sh + TK) % TK : (rr + sh) % TK;
      float moved = x[src], kept = x[rr];
      asm("" : "+v"(moved), "+v"(kept));
      y[rr] = on ? moved : kept;
    }
#pragma unroll
    for (int rr = 0; rr < TK; rr++) x[rr] = y[rr];
  }
}

// Per-plane factors of the register tiles: row u holds plane u - SKEW,
// zero outside the tile's planes (TK + 2 SKEW <= 64 rows).
template <int TK>
struct TileRegLds {
  float2 pf[TK + 2 * SKEW];
  float fi[TK][TJ];  // incoming i face, consumer-step rows
  float fj[TK][TI];  // incoming j face
};

// One register tile's operands, this lane's share, loaded in one batch.
template <int TK>
struct RegIn {
  float v[TK];                  // forward: r of plane u (row u); backward: y, skewed
  float d[TK];                  // dinv, skewed
  float fi[TK / 8], fj[TK / 8];  // incoming faces: entries lane + 64 e
  float2 pf;                    // per-plane factors of row `lane`
  float fk;                     // k face of the neighbour tile in k
  float fkt;                    // phi-seam value (phi-wrapped plans)
  float ck, cj, ci;             // couplings of this lane's column
};

// Tile code on a diagonal: I + NI (J + NJ Kc).
__device__ __forceinline__ int reg_tile_code(const DiagArgs& da, int b) {
  if (da.tiles) return da.tiles[b];
  int I, J, Kc;
  diag_tile(da, b, I, J, Kc);
  return I + da.NI * (J + da.NJ * Kc);
}

// Where a tile sits and what it couples to; one per tile, from its code.
template <int TK>
struct RegTile {
  int I, J, Kc, k0, k1, nk;
  size_t tid, col, tb;
  __device__ __forceinline__ RegTile(const DiagArgs& da, const Grid3& g, int code, int lane) {
    I = code % da.NI;
    J = (code / da.NI) % da.NJ;
    Kc = code / (da.NI * da.NJ);
    k0 = Kc * TK;
    k1 = min(g.c, k0 + TK);
    nk = k1 - k0;
    tid = (size_t)code;
    col = (size_t)I + (size_t)da.NI * J;
    tb = tid * (size_t)TK * 64 + lane;
  }
};

// Stages a tile's faces and per-plane factors in the wave's LDS.  Only this
// wave touches its LDS, and a wavefront's LDS operations are performed in
// order, so wavefront-scope fences (compiler barriers, no instructions)
// are all that orders these writes after the previous tile's reads and
// before this tile's.
template <int TK>
__device__ __forceinline__ void reg_stage(TileRegLds<TK>& sh, const RegIn<TK>& in, int lane,
                                          int nk, bool face_i, bool face_j) {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int e = 0; e < TK / 8; e++) {
    (&sh.fi[0][0])[lane + 64 * e] = face_i ? in.fi[e] : 0.f;
    (&sh.fj[0][0])[lane + 64 * e] = face_j ? in.fj[e] : 0.f;
  }
  const int qf = lane - SKEW;
  if (lane < TK + 2 * SKEW) sh.pf[lane] = (qf >= 0 && qf < nk) ? in.pf : make_float2(0.f, 0.f);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
}

// kUpd: the residual update of the FP32 recurrence on the way in,
// r = r - alpha Ap (the tile's r is stored back by the caller).  kCoup:
// the column's couplings from the per-column table (the block-persistent
// sweeps, whose registers are the tighter), else formed from the metric's
// 1-D factors as the table was (the same values; the per-diagonal
// launches, where the table's 2.2 MB are not kept in the L2s).
template <typename T, int TK, bool kWrap, bool kUpd = false, bool kCoup = false>
__device__ __forceinline__ void fwd_reg_load(RegIn<TK>& in, const Metric& m, const Grid3& g,
                                             const DiagArgs& da, const float* coup,
                                             const T* __restrict__ r, int code, int lane,
                                             const float* __restrict__ ap = nullptr,
                                             float alpha = 0.f) {
  const RegTile<TK> t(da, g, code, lane);
  const size_t fbi = tile_id(da, max(t.I - 1, 0), t.J, t.Kc) * TK * 8;
  const size_t fbj = tile_id(da, t.I, max(t.J - 1, 0), t.Kc) * TK * 8;
#pragma unroll
  for (int e = 0; e < TK / 8; e++) {
    in.fi[e] = da.face_i[fbi + lane + 64 * e];
    in.fj[e] = da.face_j[fbj + lane + 64 * e];
  }
  const int qf = lane - SKEW;
  const int kq = (qf >= 0 && qf < t.nk) ? t.k0 + qf : 0;
  in.pf = make_float2(m.FDPH[kq], m.FDPL[kq]);
  in.fk = da.face_k[tile_id(da, t.I, t.J, max(t.Kc - 1, 0)) * 64 + lane];
  in.fkt = kWrap ? da.face_kt[t.col * 64 + lane] : 0.f;
  if constexpr (kCoup) {
    in.ck = coup[(t.col * 3 + 0) * 64 + lane];
    in.cj = coup[(t.col * 3 + 1) * 64 + lane];
    in.ci = coup[(t.col * 3 + 2) * 64 + lane];
  } else {
    const int i = t.I * TI + (lane & 7), j = t.J * TJ + (lane >> 3);
    const bool act = i < g.a && j < g.b;
    const SweepCoup c = sweep_coup_lower(m, act ? i : 0, act ? j : 0);
    in.ck = c.k;
    in.cj = c.j;
    in.ci = c.i;
  }
#pragma unroll
  for (int u = 0; u < TK; u++) {
    in.d[u] = da.dinv_t[t.tb + (size_t)u * 64];
    if constexpr (kUpd)
      in.v[u] = __builtin_fmaf(-alpha, ap[t.tb + (size_t)u * 64], (float)r[t.tb + (size_t)u * 64]);
    else
      in.v[u] = (float)r[t.tb + (size_t)u * 64];
  }
}

template <typename T, int TK, bool kWrap, bool kCoup = false>
__device__ __forceinline__ void bwd_reg_load(RegIn<TK>& in, const Metric& m, const Grid3& g,
                                             const DiagArgs& da, const float* coup, int code,
                                             int lane) {
  const RegTile<TK> t(da, g, code, lane);
  const size_t fbi = tile_id(da, min(t.I + 1, da.NI - 1), t.J, t.Kc) * TK * 8;
  const size_t fbj = tile_id(da, t.I, min(t.J + 1, da.NJ - 1), t.Kc) * TK * 8;
#pragma unroll
  for (int e = 0; e < TK / 8; e++) {
    in.fi[e] = da.face_i[fbi + lane + 64 * e];
    in.fj[e] = da.face_j[fbj + lane + 64 * e];
  }
  const int qf = lane - SKEW;
  const int kq = (qf >= 0 && qf < t.nk) ? t.k0 + qf : 0;
  in.pf = make_float2(m.FDPH[kq], m.FDPU[kq]);
  in.fk = da.face_k[tile_id(da, t.I, t.J, min(t.Kc + 1, da.NK - 1)) * 64 + lane];
  in.fkt = kWrap ? da.face_kt[t.col * 64 + lane] : 0.f;
  if constexpr (kCoup) {
    in.ck = coup[(t.col * 3 + 0) * 64 + lane];
    in.cj = coup[(t.col * 3 + 1) * 64 + lane];
    in.ci = coup[(t.col * 3 + 2) * 64 + lane];
  } else {
    const int i = t.I * TI + (lane & 7), j = t.J * TJ + (lane >> 3);
    const bool act = i < g.a && j < g.b;
    const SweepCoup c = sweep_coup_upper(m, act ? i : 0, act ? j : 0);
    in.ck = c.k;
    in.cj = c.j;
    in.ci = c.i;
  }
#pragma unroll
  for (int u = 0; u < TK; u++) {
    in.v[u] = da.yv_t[t.tb + (size_t)u * 64];
    in.d[u] = da.dinv_t[t.tb + (size_t)u * 64];
  }
}

template <int TK, bool kWrap>
__device__ __forceinline__ void fwd_reg_tile(RegIn<TK>& in, TileRegLds<TK>& sh, const Metric& m,
                                             const Grid3& g, const DiagArgs& da, int code,
                                             int lane) {
  const RegTile<TK> t(da, g, code, lane);
  const int k0 = t.k0, k1 = t.k1, nk = t.nk;
  const int ii = lane & (TI - 1), jj = lane / TI, dg = ii + jj;
  const int i = t.I * TI + ii, j = t.J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const bool lo_k0 = da.bl.lo_k(k0);
  const bool wrap = kWrap && da.bl.wrap != 0;
  const bool face_i = t.I > 0 && da.bl.lo_i(t.I * TI);
  const bool face_j = t.J > 0 && da.bl.lo_j(t.J * TJ);
  reg_stage(sh, in, lane, nk, face_i, face_j);
  float (&rv)[TK] = in.v;
  float (&dv)[TK] = in.d;
  // r of plane q to register (q + ii + jj) mod TK
  rotate_col<TK>(rv, dg, 1);
  const float t_own = (act && lo_k0) ? in.fk : 0.f;
  const float t_w = (act && wrap && k0 > 0 && k1 == g.c) ? in.fkt : 0.f;
  const bool wrap_tile = wrap && k1 == g.c;  // seam coupling at plane nk-1
  const bool cap0 = wrap && k0 == 0;         // records t of plane 0
  const float aw = wrap_tile ? in.ck * m.FDPU[m.c - 1] : 0.f;
  const float2* pfl = sh.pf + (SKEW - dg);  // pfl[s]: factors of plane s - dg
  const float clk = in.ck, clj = in.cj, cli = in.ci;
  float t_k = t_own, tw = t_w;
  auto march = [&](auto seam) {
    constexpr bool kSeam = decltype(seam)::value;
#pragma unroll
    for (int s = 0; s < TK + SKEW; s++) {
      const int rho = s % TK;
      const float2 f = pfl[s];
      const float fl = sh.fi[rho][jj], fd = sh.fj[rho][ii];
      float from_i, from_j;
      lanes_from_lower(t_k, from_i, from_j);
      const float ak = clk * f.y, aj = clj * f.x, ai = cli * f.x;
      float y = rv[rho];
      if constexpr (kSeam) {
        if (wrap_tile && s - dg == nk - 1) y = sub_prod(y, aw, tw);
      }
      y = sub_prod(y, ak, t_k);
      y = sub_prod(y, aj, (jj == 0 ? fd : from_j));
      y = sub_prod(y, ai, (ii == 0 ? fl : from_i));
      float tt = dv[rho] * y;
      // every lane computes every step (the operands are read in batches,
      // ahead of the dependency chain); only the results are selected
      asm volatile("" : "+v"(y), "+v"(tt));
      if constexpr (kSeam) {
        if (cap0 && s == dg) tw = tt;
      }
      if (s < SKEW || s >= TK) {
        const bool v = s < SKEW ? dg <= s : dg > s - TK;
        rv[rho] = v ? y : rv[rho];
        t_k = v ? tt : t_k;
      } else {
        rv[rho] = y;
        t_k = tt;
      }
    }
  };
  // the seam steps are compiled only into the kernel for phi-wrapped plans
  if constexpr (kWrap) {
    if (wrap && (k0 == 0 || k1 == g.c))
      march(std::true_type{});
    else
      march(std::false_type{});
  } else {
    march(std::false_type{});
  }
  // y outside the grid is stored as zero, so that the backward sweep's
  // padding lanes (dinv zero) produce z = 0
#pragma unroll
  for (int u = 0; u < TK; u++) da.yv_t[t.tb + (size_t)u * 64] = act ? rv[u] : 0.f;
  // outgoing faces t = dinv*y of the last column (ii = 7) and row (jj = 7)
  const size_t fo = t.tid * TK * 8;
  if (ii == TI - 1) {
#pragma unroll
    for (int c = 0; c < TK; c++)
      da.face_i[fo + c * 8 + jj] = dv[(c + 7) % TK] * rv[(c + 7) % TK];
  }
  if (jj == TJ - 1) {
#pragma unroll
    for (int c = 0; c < TK; c++)
      da.face_j[fo + c * 8 + ii] = dv[(c + 7) % TK] * rv[(c + 7) % TK];
  }
  da.face_k[t.tid * 64 + lane] = t_k;  // t of plane TK-1, for the tile above
  if (cap0 && k1 < g.c) da.face_kt[t.col * 64 + lane] = tw;
}

template <typename T, int TK, bool kWrap>
__device__ __forceinline__ void bwd_reg_tile(RegIn<TK>& in, TileRegLds<TK>& sh, const Metric& m,
                                             const Grid3& g, const DiagArgs& da,
                                             T* __restrict__ z, int code, int lane) {
  const RegTile<TK> t(da, g, code, lane);
  const int k0 = t.k0, k1 = t.k1, nk = t.nk;
  const int ii = lane & (TI - 1), jj = lane / TI, dg = ii + jj, rdg = SKEW - dg;
  const int i = t.I * TI + ii, j = t.J * TJ + jj;
  const bool act = (i < g.a) && (j < g.b);
  const bool hi_k1 = da.bl.hi_k(k1 - 1, g.c), wrap = kWrap && da.bl.wrap != 0;
  const bool face_i = t.I + 1 < da.NI && da.bl.hi_i(t.I * TI + TI - 1, g.a);
  const bool face_j = t.J + 1 < da.NJ && da.bl.hi_j(t.J * TJ + TJ - 1, g.b);
  reg_stage(sh, in, lane, nk, face_i, face_j);
  float (&yv)[TK] = in.v;
  float (&dv)[TK] = in.d;
  const float t_own = (act && hi_k1) ? in.fk : 0.f;
  const float t_w = (act && wrap && k1 < g.c && k0 == 0) ? in.fkt : 0.f;
  const bool wrap_tile = wrap && k0 == 0;    // seam coupling at plane 0
  const bool cap_last = wrap && k1 == g.c;   // records z of plane nk-1
  const float bw = wrap_tile ? in.ck * m.FDPL[0] : 0.f;
  // step s handles plane q = TK-1 - s + rdg, factors in row q + SKEW
  const float2* pfb = sh.pf + rdg;
  const float cuk = in.ck, cuj = in.cj, cui = in.ci;
  float t_k = t_own, tw = t_w;
  auto march = [&](auto seam) {
    constexpr bool kSeam = decltype(seam)::value;
#pragma unroll
    for (int s = 0; s < TK + SKEW; s++) {
      const int rho = ((SKEW - 1 - s) % TK + TK) % TK;
      const int crow = s % TK;
      const float2 f = pfb[TK - 1 + SKEW - s];
      const float fr = sh.fi[crow][jj], fu = sh.fj[crow][ii];
      float from_i, from_j;
      lanes_from_upper(t_k, from_i, from_j);
      const float bk = cuk * f.y, bj = cuj * f.x, bi = cui * f.x;
      float sv = yv[rho];
      sv = sub_prod(sv, bi, (ii == TI - 1 ? fr : from_i));
      sv = sub_prod(sv, bj, (jj == TJ - 1 ? fu : from_j));
      sv = sub_prod(sv, bk, t_k);
      const int q = TK - 1 - s + rdg;
      if constexpr (kSeam) {
        if (wrap_tile && q == 0) sv = sub_prod(sv, bw, tw);
      }
      float zz = sv * dv[rho];
      asm volatile("" : "+v"(zz));  // computed on every step, as in the forward march
      if constexpr (kSeam) {
        if (cap_last && q == nk - 1) tw = zz;
      }
      if (s < SKEW || s >= TK) {
        const bool v = s < SKEW ? rdg <= s : rdg > s - TK;
        yv[rho] = v ? zz : yv[rho];
        t_k = v ? zz : t_k;
      } else {
        yv[rho] = zz;
        t_k = zz;
      }
    }
  };
  if constexpr (kWrap) {
    if (wrap && (k0 == 0 || k1 == g.c))
      march(std::true_type{});
    else
      march(std::false_type{});
  } else {
    march(std::false_type{});
  }
  // outgoing faces: z of the first column (ii = 0) and row (jj = 0)
  const size_t fo = t.tid * TK * 8;
  if (ii == 0) {
#pragma unroll
    for (int c = 0; c < TK; c++) da.face_i[fo + c * 8 + jj] = yv[((6 - c) % TK + TK) % TK];
  }
  if (jj == 0) {
#pragma unroll
    for (int c = 0; c < TK; c++) da.face_j[fo + c * 8 + ii] = yv[((6 - c) % TK + TK) % TK];
  }
  da.face_k[t.tid * 64 + lane] = t_k;  // z of plane 0, for the tile below
  if (cap_last && k0 > 0) da.face_kt[t.col * 64 + lane] = tw;
  // z of plane q from register (q + ii + jj) mod TK back to row q
  rotate_col<TK>(yv, dg, -1);
  if (act) {
#pragma unroll
    for (int u = 0; u < TK; u++) z[t.tb + (size_t)u * 64] = (T)yv[u];
  }
}

// Waves per SIMD the register tiles are built for: 7 forward and 8
// backward (at most 72 and 64 VGPRs) for the FP32 16-plane tiles without
// the phi seam, which then fit without spilling (6 for the forward sweep
// that also updates the residual: its Ap values are in flight with r); the
// others take what they need.
template <typename T, int TK, bool kWrap, bool kFwd, bool kUpd = false>
constexpr int kRegWaves =
    (TK == 16 && !kWrap && sizeof(T) == 4) ? (kFwd ? (kUpd ? 6 : 7) : 8) : 1;
// Waves per workgroup: each marches a tile of its own (no data shared
// between the waves), two per workgroup so that the resident tiles per CU
// are not capped by the number of workgroups a CU holds.  (Looping a wave
// over several tiles with the next tile's loads in flight measured slower:
// the second operand set halves the resident waves.)
constexpr int kRegTilesPerGroup = 2;

// kUpd (FP32, with guard): the residual update r -= alpha Ap folded into
// the forward sweep, each tile storing its updated r before its march; the
// same fmaf as k_r_update, so r and the sweep's values are the same bits.
template <typename T, int TK, bool kWrap, bool kUpd = false>
__global__ void __launch_bounds__(64 * kRegTilesPerGroup)
    __attribute__((amdgpu_waves_per_eu(kRegWaves<T, TK, kWrap, true, kUpd>, 8)))
    k_tile_fwd_reg(Metric m, Grid3 g, DiagArgs da, const float* __restrict__ coup,
                   T* __restrict__ r, const CgState* guard,
                   const float* __restrict__ ap = nullptr) {
  static_assert(TK >= SKEW && TK + 2 * SKEW <= 64, "register tiles: 16 or 32 planes");
  __shared__ TileRegLds<TK> shw[kRegTilesPerGroup];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = (int)blockIdx.x * kRegTilesPerGroup + wave;
  if (b >= da.count) return;  // no workgroup barrier follows
  const int code = reg_tile_code(da, b);
  RegIn<TK> in;
  float alpha = 0.f;
  if constexpr (kUpd) alpha = (float)st_load(&guard->alpha);
  fwd_reg_load<T, TK, kWrap, kUpd>(in, m, g, da, coup, r, code, lane, ap, alpha);
  // the done flag is read after the tile's loads, before any store
  if (guard && st_done(guard)) return;
  if constexpr (kUpd) {
    const RegTile<TK> t(da, g, code, lane);
#pragma unroll
    for (int u = 0; u < TK; u++) r[t.tb + (size_t)u * 64] = in.v[u];
  }
  fwd_reg_tile<TK, kWrap>(in, shw[wave], m, g, da, code, lane);
}

template <typename T, int TK, bool kWrap>
__global__ void __launch_bounds__(64 * kRegTilesPerGroup)
    __attribute__((amdgpu_waves_per_eu(kRegWaves<T, TK, kWrap, false>, 8)))
    k_tile_bwd_reg(Metric m, Grid3 g, DiagArgs da, const float* __restrict__ coup,
                   T* __restrict__ z, const CgState* guard) {
  static_assert(TK >= SKEW && TK + 2 * SKEW <= 64, "register tiles: 16 or 32 planes");
  __shared__ TileRegLds<TK> shw[kRegTilesPerGroup];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int b = (int)blockIdx.x * kRegTilesPerGroup + wave;
  if (b >= da.count) return;  // no workgroup barrier follows
  const int code = reg_tile_code(da, b);
  RegIn<TK> in;
  bwd_reg_load<T, TK, kWrap>(in, m, g, da, coup, code, lane);
  if (guard && st_done(guard)) return;
  bwd_reg_tile<T, TK, kWrap>(in, shw[wave], m, g, da, z, code, lane);
}

// Block-persistent register-tile sweeps (block-Jacobi plans): one
// workgroup of kBlkWaves waves per block of the cut marches that block's
// tile diagonals in order (backward: in reverse), its waves taking the
// diagonal's tiles in turn, with a workgroup barrier between diagonals.
// The blocks are independent, so no workgroup ever waits for another: a
// sweep is one launch instead of one per diagonal, and a block moves on to
// its next diagonal as soon as its own tiles are done.  A tile's faces come
// from tiles of the same block on the previous diagonal, written by waves
// of the same workgroup before the barrier (its release and acquire order
// them for the workgroup); the per-tile work is that of k_tile_fwd_reg /
// k_tile_bwd_reg, so the results are the same bits.
constexpr int kBlkWaves = 8;
// Waves per SIMD of the block-persistent sweeps for the FP32 16-plane tiles
// without the seam: with kRoomy (at most 2 blocks per CU) 4, i.e. two
// workgroups per CU and registers enough for the loop over diagonals
// without spilling; else 6, three workgroups per CU (the loop then spills
// a few registers).
template <typename T, int TK, bool kWrap, bool kRoomy>
constexpr int kBlkRegWaves = (TK == 16 && !kWrap && sizeof(T) == 4) ? (kRoomy ? 4 : 6) : 1;

struct BlkArgs {
  const int* tiles;  // tile codes, by block, then diagonal
  const int* off;    // nblk * ndiag + 1 offsets into tiles
  int ndiag;
  int L0, L1;  // the diagonals this launch marches: L0 .. L1-1
};

template <typename T, int TK, bool kWrap, bool kUpd, bool kRoomy>
__global__ void __launch_bounds__(64 * kBlkWaves)
    __attribute__((amdgpu_waves_per_eu(kBlkRegWaves<T, TK, kWrap, kRoomy>, 8)))
    k_tile_fwd_blk(Metric m, Grid3 g, DiagArgs da, BlkArgs ba, const float* __restrict__ coup,
                   T* __restrict__ r, const CgState* guard,
                   const float* __restrict__ ap = nullptr) {
  static_assert(TK >= SKEW && TK + 2 * SKEW <= 64, "register tiles: 16 or 32 planes");
  __shared__ TileRegLds<TK> shw[kBlkWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  // the flag is written by none of this launch's workgroups: every thread
  // reads the same value and the whole workgroup returns or none does
  if (guard && st_done(guard)) return;
  float alpha = 0.f;
  if constexpr (kUpd) alpha = (float)st_load(&guard->alpha);
  const int* off = ba.off + (size_t)blockIdx.x * ba.ndiag;
  for (int L = ba.L0; L < ba.L1; L++) {
    const int o1 = off[L + 1];
    for (int b = off[L] + wave; b < o1; b += kBlkWaves) {
      const int code = ba.tiles[b];
      RegIn<TK> in;
      fwd_reg_load<T, TK, kWrap, kUpd, true>(in, m, g, da, coup, r, code, lane, ap, alpha);
      if constexpr (kUpd) {
        const RegTile<TK> t(da, g, code, lane);
#pragma unroll
        for (int u = 0; u < TK; u++) r[t.tb + (size_t)u * 64] = in.v[u];
      }
      fwd_reg_tile<TK, kWrap>(in, shw[wave], m, g, da, code, lane);
    }
    __syncthreads();
  }
}

template <typename T, int TK, bool kWrap, bool kRoomy>
__global__ void __launch_bounds__(64 * kBlkWaves)
    __attribute__((amdgpu_waves_per_eu(kBlkRegWaves<T, TK, kWrap, kRoomy>, 8)))
    k_tile_bwd_blk(Metric m, Grid3 g, DiagArgs da, BlkArgs ba, const float* __restrict__ coup,
                   T* __restrict__ z, const CgState* guard) {
  static_assert(TK >= SKEW && TK + 2 * SKEW <= 64, "register tiles: 16 or 32 planes");
  __shared__ TileRegLds<TK> shw[kBlkWaves];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (guard && st_done(guard)) return;  // as in the forward sweep
  const int* off = ba.off + (size_t)blockIdx.x * ba.ndiag;
  for (int L = ba.L1 - 1; L >= ba.L0; L--) {
    const int o1 = off[L + 1];
    for (int b = off[L] + wave; b < o1; b += kBlkWaves) {
      const int code = ba.tiles[b];
      RegIn<TK> in;
      bwd_reg_load<T, TK, kWrap, true>(in, m, g, da, coup, code, lane);
      bwd_reg_tile<T, TK, kWrap>(in, shw[wave], m, g, da, z, code, lane);
    }
    __syncthreads();
  }
}

// Couplings of every column of tiles for the register tiles: lower (k, j,
// i) at coup_lo[(col * 3 + c) * 64 + lane], upper likewise, col = I + TNI J;
// columns outside the grid take those of column (0, 0), as in the marches.
__global__ void k_reg_coup(Metric m, Grid3 g, int TNI, int TNJ, float* __restrict__ lo,
                           float* __restrict__ up) {
  const int col = blockIdx.x, lane = threadIdx.x;
  const int I = col % TNI, J = col / TNI;
  const int i = I * TI + (lane & 7), j = J * TJ + (lane >> 3);
  const bool act = i < g.a && j < g.b;
  const SweepCoup cl = sweep_coup_lower(m, act ? i : 0, act ? j : 0);
  const SweepCoup cu = sweep_coup_upper(m, act ? i : 0, act ? j : 0);
  float* l = lo + (size_t)col * 192 + lane;
  float* u = up + (size_t)col * 192 + lane;
  l[0] = cl.k;
  l[64] = cl.j;
  l[128] = cl.i;
  u[0] = cu.k;
  u[64] = cu.j;
  u[128] = cu.i;
}

// One sweep's launches of the register tiles.  Diagonals [0, plan.ndiag)
// in order (forward) or reverse (backward); with persist 1 one
// block-persistent launch marches all of them, with persist 2 one marches
// the leading diagonals [0, blk_lo) and one the trailing [blk_hi, ndiag)
// (the small ones, latency-bound as launches of their own), the others one
// launch each.
template <typename T, int TK, bool kFwd>
void tiles_sweep_reg(const SweepPlan& plan, const Metric& m, const Grid3& g, T* r, T* z,
                     const CgState* guard, hipStream_t s, const float* ap) {
  const bool wrap = plan.bl.wrap != 0;
  constexpr int kThreads = 64 * kRegTilesPerGroup;
  const bool pers = plan.persist != 0 && plan.blk_tiles;
  const int lo = !pers ? 0 : (plan.persist == 1 ? plan.ndiag : plan.blk_lo);
  const int hi = !pers ? plan.ndiag : (plan.persist == 1 ? plan.ndiag : plan.blk_hi);
  auto blk = [&](int L0, int L1) {
    if (L0 >= L1) return;
    DiagArgs da = diag_args(plan, 0);
    da.tiles = nullptr;
    da.count = 0;
    const BlkArgs ba{plan.blk_tiles, plan.blk_off, plan.ndiag, L0, L1};
    const int nb = plan.nblk;
    constexpr int kT = 64 * kBlkWaves;
    // the register budget only differs for the FP32 16-plane tiles
    // without the seam
    constexpr bool kTuned = TK == 16 && sizeof(T) == 4;
    const bool roomy = kTuned && plan.blk_roomy;
    if constexpr (kFwd) {
      bool upd = false;
      if constexpr (sizeof(T) == 4) {
        if (ap) {
          upd = true;
          if (wrap)
            k_tile_fwd_blk<T, TK, true, true, false>
                <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard, ap);
          else if (roomy)
            k_tile_fwd_blk<T, TK, false, true, kTuned>
                <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard, ap);
          else
            k_tile_fwd_blk<T, TK, false, true, false>
                <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard, ap);
        }
      }
      if (!upd) {
        if (wrap)
          k_tile_fwd_blk<T, TK, true, false, false>
              <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard);
        else if (roomy)
          k_tile_fwd_blk<T, TK, false, false, kTuned>
              <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard);
        else
          k_tile_fwd_blk<T, TK, false, false, false>
              <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_lo, r, guard);
      }
    } else {
      if (wrap)
        k_tile_bwd_blk<T, TK, true, false><<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_up, z, guard);
      else if (roomy)
        k_tile_bwd_blk<T, TK, false, kTuned>
            <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_up, z, guard);
      else
        k_tile_bwd_blk<T, TK, false, false>
            <<<nb, kT, 0, s>>>(m, g, da, ba, plan.coup_up, z, guard);
    }
  };
  auto diag = [&](int L) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    const DiagArgs da = diag_args(plan, L);
    const int nb = (cnt + kRegTilesPerGroup - 1) / kRegTilesPerGroup;
    if constexpr (kFwd) {
      if constexpr (sizeof(T) == 4) {
        if (ap) {
          if (wrap)
            k_tile_fwd_reg<T, TK, true, true>
                <<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_lo, r, guard, ap);
          else
            k_tile_fwd_reg<T, TK, false, true>
                <<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_lo, r, guard, ap);
          return;
        }
      }
      if (wrap)
        k_tile_fwd_reg<T, TK, true><<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_lo, r, guard);
      else
        k_tile_fwd_reg<T, TK, false><<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_lo, r, guard);
    } else {
      if (wrap)
        k_tile_bwd_reg<T, TK, true><<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_up, z, guard);
      else
        k_tile_bwd_reg<T, TK, false><<<nb, kThreads, 0, s>>>(m, g, da, plan.coup_up, z, guard);
    }
  };
  if constexpr (kFwd) {
    blk(0, lo);
    for (int L = lo; L < hi; L++) diag(L);
    blk(max(hi, lo), plan.ndiag);
  } else {
    blk(max(hi, lo), plan.ndiag);
    for (int L = hi - 1; L >= lo; L--) diag(L);
    blk(0, lo);
  }
}

// ap non-null (FP32, with guard): the forward sweep also applies the
// residual update r -= alpha Ap (k_tile_fwd_reg<.., kUpd>).
template <typename T, int TK>
void tiles_launch_reg(const SweepPlan& plan, const Metric& m, const Grid3& g, const T* r, T* z,
                      const CgState* guard, hipStream_t s, const float* ap = nullptr) {
  T* rw = const_cast<T*>(r);  // written only by the kUpd sweep
  tiles_sweep_reg<T, TK, true>(plan, m, g, rw, z, guard, s, ap);
  tiles_sweep_reg<T, TK, false>(plan, m, g, rw, z, guard, s, ap);
}

template <typename T, int TK, int W>
void tiles_launch_lds_w(const SweepPlan& plan, const Metric& m, const Grid3& g, const T* r, T* z,
                        const CgState* guard, hipStream_t s) {
  for (int L = 0; L < plan.ndiag; L++) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    k_tile_fwd_lds<T, TK, W><<<cnt, 64 * W, 0, s>>>(m, g, diag_args(plan, L), r, guard);
  }
  for (int L = plan.ndiag - 1; L >= 0; L--) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    k_tile_bwd_lds<T, TK, W><<<cnt, 64 * W, 0, s>>>(m, g, diag_args(plan, L), z, guard);
  }
}

template <typename T, int TK>
void tiles_launch_lds_tk(const SweepPlan& plan, const Metric& m, const Grid3& g,
                         const float* dinv, const T* r, float* yv, T* z, const CgState* guard,
                         hipStream_t s) {
  (void)dinv;
  (void)yv;  // the single-wave sweeps use the plan's tile-major copies
  if (plan.stage_one == 2) {
    if constexpr (TK + 2 * SKEW <= 64) tiles_launch_reg<T, TK>(plan, m, g, r, z, guard, s);
  } else if (plan.stage_one) {
    tiles_launch_lds_w<T, TK, 1>(plan, m, g, r, z, guard, s);
  } else {
    tiles_launch_lds_w<T, TK, TK / 8>(plan, m, g, r, z, guard, s);
  }
}

// ---------------------------------------------------------------------------
// Multi-wave tiles: a workgroup of SI x SJ waves owns a (8 SI) x (8 SJ) x TK
// tile, each wave an 8x8 column stack.  Wave (si,sj) runs the same skewed
// march 8(si+sj) steps after wave (0,0); the face values its edge lanes need
// from the wave below in i or j were produced one step earlier and are
// passed through LDS, with one barrier per step.  Bigger tiles mean fewer
// anti-diagonals, i.e. fewer launches per sweep: every launch boundary costs
// a dispatch and an L2 write-back on a multi-XCD part.
template <int SI, int SJ, int TK>
struct TileLdsW {
  float v[TK][64 * SI * SJ];   // staged r (forward) / y (backward), then the result
  float d[TK][64 * SI * SJ];   // dinv
  float fi[TK][8 * SJ];        // incoming i face of the tile (previous launch)
  float fj[TK][8 * SI];        // incoming j face of the tile
  float xi[SI > 1 ? SI - 1 : 1][SJ][TK][8];  // i faces between waves
  float xj[SI][SJ > 1 ? SJ - 1 : 1][TK][8];  // j faces between waves
  float dph[TK], dpx[TK];  // FDPH and FDPL / FDPU
};

template <typename T, int SI, int SJ, int TK>
__global__ void __launch_bounds__(64 * SI * SJ)
    k_tilew_fwd(Metric m, Grid3 g, DiagArgs da, const float* __restrict__ dinv,
                const T* __restrict__ r, float* __restrict__ yv, const CgState* guard) {
  const int done = guard ? st_done(guard) : 0;
  __shared__ TileLdsW<SI, SJ, TK> sh;
  int I, J, Kc;
  diag_tile(da, blockIdx.x, I, J, Kc);
  const int k0 = Kc * TK, k1 = min(g.c, k0 + TK), nk = k1 - k0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int si = wave % SI, sj = wave / SI;
  const int ii = lane & 7, jj = lane >> 3;
  const int i = (I * SI + si) * 8 + ii, j = (J * SJ + sj) * 8 + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  const uint32_t col = (uint32_t)ic + (uint32_t)g.a * (uint32_t)jc;
  const bool lo_i = da.bl.lo_i(i), lo_j = da.bl.lo_j(j), lo_k0 = da.bl.lo_k(k0);
  const bool wrap = da.bl.wrap != 0;
  const SweepCoup cl = sweep_coup_lower(m, ic, jc);
  float t_own = 0.f, t_w = 0.f, t_pub = 0.f;
  if (act && lo_k0) {
    const uint32_t q = col + g.ab * (uint32_t)(k0 - 1);
    t_own = dinv[q] * yv[q];
  }
  if (act && wrap && k0 > 0 && k1 == g.c) t_w = dinv[col] * yv[col];
  float dph_s = 0.f, dpx_s = 0.f;
  if (tid < nk) {
    dph_s = m.FDPH[k0 + tid];
    dpx_s = m.FDPL[k0 + tid];
  }
  const bool face_i = (si == 0 && ii == 0 && lo_i), face_j = (sj == 0 && jj == 0 && lo_j);
  for (int q0 = 0; q0 < TK; q0 += 16) {
    float rv[16], dv[16], fiv[16], fjv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = q0 + u;
      rv[u] = dv[u] = fiv[u] = fjv[u] = 0.f;
      if (act && q < nk) {
        const uint32_t p = col + g.ab * (uint32_t)(k0 + q);
        rv[u] = (float)r[p];
        dv[u] = dinv[p];
        if (face_i) fiv[u] = dinv[p - 1] * yv[p - 1];
        if (face_j) fjv[u] = dinv[p - g.a] * yv[p - g.a];
      }
    }
    if (done) return;  // uniform over the block
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = q0 + u;
      if (q < nk) {
        sh.v[q][tid] = rv[u];
        sh.d[q][tid] = dv[u];
        if (si == 0 && ii == 0) sh.fi[q][8 * sj + jj] = fiv[u];
        if (sj == 0 && jj == 0) sh.fj[q][8 * si + ii] = fjv[u];
      }
    }
  }
  if (tid < nk) {
    sh.dph[tid] = dph_s;
    sh.dpx[tid] = dpx_s;
  }
  __syncthreads();
  const float aw = (wrap && k1 == g.c) ? cl.k * m.FDPU[m.c - 1] : 0.f;
  const int off = 8 * (si + sj);
  const int S = nk + SKEW + 8 * (SI - 1) + 8 * (SJ - 1);
  for (int T0 = 0; T0 < S; T0 += 8) {
    float rr[8], dd[8], ak[8], aj[8], ai[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int q = T0 + u - off - ii - jj;
      rr[u] = dd[u] = ak[u] = aj[u] = ai[u] = 0.f;
      if (act && q >= 0 && q < nk) {
        rr[u] = sh.v[q][tid];
        dd[u] = sh.d[q][tid];
        const float dph = sh.dph[q];
        ak[u] = cl.k * sh.dpx[q];
        aj[u] = cl.j * dph;
        ai[u] = cl.i * dph;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int q = T0 + u - off - ii - jj;
      const int k = k0 + q;
      float from_i, from_j;
      lanes_from_lower(t_pub, from_i, from_j);
      if (act && q >= 0 && q < nk) {
        float fl = from_i, fd = from_j;
        if (ii == 0) fl = (si == 0) ? sh.fi[q][8 * sj + jj] : sh.xi[si > 0 ? si - 1 : 0][sj][q][jj];
        if (jj == 0) fd = (sj == 0) ? sh.fj[q][8 * si + ii] : sh.xj[si][sj > 0 ? sj - 1 : 0][q][ii];
        float y = rr[u];
        if (wrap && k == g.c - 1) y = sub_prod(y, aw, t_w);
        if (q > 0 || lo_k0) y = sub_prod(y, ak[u], t_own);
        if (lo_j) y = sub_prod(y, aj[u], fd);
        if (lo_i) y = sub_prod(y, ai[u], fl);
        sh.v[q][tid] = y;
        const float tt = dd[u] * y;
        if (ii == 7 && si < SI - 1) sh.xi[si][sj][q][jj] = tt;
        if (jj == 7 && sj < SJ - 1) sh.xj[si][sj][q][ii] = tt;
        if (k == 0) t_w = tt;
        t_own = tt;
        t_pub = tt;
      }
      __syncthreads();
    }
  }
  if (act)
    for (int q = 0; q < nk; q++) yv[col + g.ab * (uint32_t)(k0 + q)] = sh.v[q][tid];
}

template <typename T, int SI, int SJ, int TK>
__global__ void __launch_bounds__(64 * SI * SJ)
    k_tilew_bwd(Metric m, Grid3 g, DiagArgs da, const float* __restrict__ dinv,
                float* __restrict__ yv, T* __restrict__ z, const CgState* guard) {
  const int done = guard ? st_done(guard) : 0;
  __shared__ TileLdsW<SI, SJ, TK> sh;
  int I, J, Kc;
  diag_tile(da, blockIdx.x, I, J, Kc);
  const int k0 = Kc * TK, k1 = min(g.c, k0 + TK), nk = k1 - k0;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int si = wave % SI, sj = wave / SI;
  const int ii = lane & 7, jj = lane >> 3;
  const int rii = 7 - ii, rjj = 7 - jj;
  const int i = (I * SI + si) * 8 + ii, j = (J * SJ + sj) * 8 + jj;
  const bool act = (i < g.a) && (j < g.b);
  const int ic = act ? i : 0, jc = act ? j : 0;
  const uint32_t col = (uint32_t)ic + (uint32_t)g.a * (uint32_t)jc;
  const bool hi_i = da.bl.hi_i(i, g.a), hi_j = da.bl.hi_j(j, g.b);
  const bool hi_k1 = da.bl.hi_k(k1 - 1, g.c), wrap = da.bl.wrap != 0;
  const SweepCoup cu = sweep_coup_upper(m, ic, jc);
  float t_own = 0.f, t_w = 0.f, t_pub = 0.f;
  if (act && hi_k1) t_own = yv[col + g.ab * (uint32_t)k1];
  if (act && wrap && k1 < g.c && k0 == 0) t_w = yv[col + g.ab * (uint32_t)(g.c - 1)];
  float dph_s = 0.f, dpx_s = 0.f;
  if (tid < nk) {
    dph_s = m.FDPH[k0 + tid];
    dpx_s = m.FDPU[k0 + tid];
  }
  const bool face_i = (si == SI - 1 && ii == 7 && hi_i);
  const bool face_j = (sj == SJ - 1 && jj == 7 && hi_j);
  for (int q0 = 0; q0 < TK; q0 += 16) {
    float yv16[16], dv[16], fiv[16], fjv[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = q0 + u;
      yv16[u] = dv[u] = fiv[u] = fjv[u] = 0.f;
      if (act && q < nk) {
        const uint32_t p = col + g.ab * (uint32_t)(k0 + q);
        yv16[u] = yv[p];
        dv[u] = dinv[p];
        if (face_i) fiv[u] = yv[p + 1];
        if (face_j) fjv[u] = yv[p + g.a];
      }
    }
    if (done) return;  // uniform over the block
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int q = q0 + u;
      if (q < nk) {
        sh.v[q][tid] = yv16[u];
        sh.d[q][tid] = dv[u];
        if (si == SI - 1 && ii == 7) sh.fi[q][8 * sj + jj] = fiv[u];
        if (sj == SJ - 1 && jj == 7) sh.fj[q][8 * si + ii] = fjv[u];
      }
    }
  }
  if (tid < nk) {
    sh.dph[tid] = dph_s;
    sh.dpx[tid] = dpx_s;
  }
  __syncthreads();
  const float bw = (wrap && k0 == 0) ? cu.k * m.FDPL[0] : 0.f;
  const int off = 8 * ((SI - 1 - si) + (SJ - 1 - sj));
  const int S = nk + SKEW + 8 * (SI - 1) + 8 * (SJ - 1);
  for (int T0 = 0; T0 < S; T0 += 8) {
    float yy[8], dd[8], bk[8], bj[8], bi[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int q = (nk - 1) - (T0 + u - off - rii - rjj);
      yy[u] = dd[u] = bk[u] = bj[u] = bi[u] = 0.f;
      if (act && q >= 0 && q < nk) {
        yy[u] = sh.v[q][tid];
        dd[u] = sh.d[q][tid];
        const float dph = sh.dph[q];
        bk[u] = cu.k * sh.dpx[q];
        bj[u] = cu.j * dph;
        bi[u] = cu.i * dph;
      }
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int q = (nk - 1) - (T0 + u - off - rii - rjj);
      const int k = k0 + q;
      float from_i, from_j;
      lanes_from_upper(t_pub, from_i, from_j);
      if (act && q >= 0 && q < nk) {
        float fr = from_i, fu = from_j;
        if (ii == 7) fr = (si == SI - 1) ? sh.fi[q][8 * sj + jj] : sh.xi[si < SI - 1 ? si : 0][sj][q][jj];
        if (jj == 7) fu = (sj == SJ - 1) ? sh.fj[q][8 * si + ii] : sh.xj[si][sj < SJ - 1 ? sj : 0][q][ii];
        float sv = yy[u];
        if (hi_i) sv = sub_prod(sv, bi[u], fr);
        if (hi_j) sv = sub_prod(sv, bj[u], fu);
        if (q < nk - 1 || hi_k1) sv = sub_prod(sv, bk[u], t_own);
        if (wrap && k == 0) sv = sub_prod(sv, bw, t_w);
        const float zz = sv * dd[u];
        sh.v[q][tid] = zz;
        if (ii == 0 && si > 0) sh.xi[si - 1][sj][q][jj] = zz;
        if (jj == 0 && sj > 0) sh.xj[si][sj - 1][q][ii] = zz;
        if (k == g.c - 1) t_w = zz;
        t_own = zz;
        t_pub = zz;
      }
      __syncthreads();
    }
  }
  if (act)
    for (int q = 0; q < nk; q++) {
      const uint32_t p = col + g.ab * (uint32_t)(k0 + q);
      const float zz = sh.v[q][tid];
      yv[p] = zz;
      z[p] = (T)zz;
    }
}

template <typename T, int SI, int SJ, int TK>
void tiles_launch_w(const SweepPlan& plan, const Metric& m, const Grid3& g, const float* dinv,
                    const T* r, float* yv, T* z, const CgState* guard, hipStream_t s) {
  for (int L = 0; L < plan.ndiag; L++) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    k_tilew_fwd<T, SI, SJ, TK><<<cnt, 64 * SI * SJ, 0, s>>>(
        m, g, diag_args(plan, L), dinv, r, yv, guard);
  }
  for (int L = plan.ndiag - 1; L >= 0; L--) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    k_tilew_bwd<T, SI, SJ, TK><<<cnt, 64 * SI * SJ, 0, s>>>(
        m, g, diag_args(plan, L), dinv, yv, z, guard);
  }
}

template <typename T>
void tiles_launch_lds(const SweepPlan& plan, const Metric& m, const Grid3& g, const float* dinv,
                      const T* r, float* yv, T* z, const CgState* guard, hipStream_t s) {
  const int w = plan.SI * plan.SJ;
  if (w == 4) {
    if (plan.TK == 16)
      tiles_launch_w<T, 2, 2, 16>(plan, m, g, dinv, r, yv, z, guard, s);
    else if (plan.TK == 64)
      tiles_launch_w<T, 2, 2, 64>(plan, m, g, dinv, r, yv, z, guard, s);
    else
      tiles_launch_w<T, 2, 2, 32>(plan, m, g, dinv, r, yv, z, guard, s);
  } else if (w == 8) {
    if (plan.TK == 16)
      tiles_launch_w<T, 2, 4, 16>(plan, m, g, dinv, r, yv, z, guard, s);
    else
      tiles_launch_w<T, 2, 4, 32>(plan, m, g, dinv, r, yv, z, guard, s);
  } else if (plan.TK == 16) {
    tiles_launch_lds_tk<T, 16>(plan, m, g, dinv, r, yv, z, guard, s);
  } else if (plan.TK == 64) {
    tiles_launch_lds_tk<T, 64>(plan, m, g, dinv, r, yv, z, guard, s);
  } else {
    tiles_launch_lds_tk<T, 32>(plan, m, g, dinv, r, yv, z, guard, s);
  }
}

template <typename T, typename CS>
void tiles_launch(const SweepPlan& plan, const CS& cs, const Grid3& g, const float* dinv,
                  const T* r, float* yv, T* z, const CgState* guard, hipStream_t s) {
  for (int L = 0; L < plan.ndiag; L++) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    TileArgs ta{plan.TNI, plan.TNJ, plan.TK, plan.tiles + plan.diag_off[L]};
    k_tile_fwd<T, CS><<<cnt, 64, 0, s>>>(cs, g, ta, dinv, r, yv, guard);
  }
  for (int L = plan.ndiag - 1; L >= 0; L--) {
    const int cnt = plan.diag_off[L + 1] - plan.diag_off[L];
    TileArgs ta{plan.TNI, plan.TNJ, plan.TK, plan.tiles + plan.diag_off[L]};
    k_tile_bwd<T, CS><<<cnt, 64, 0, s>>>(cs, g, ta, dinv, yv, z, guard);
  }
}

}  // namespace

// dinv (natural order) -> tile-major copy: element t of tile (I, J, Kc),
// plane q, lane ii + 8 jj holds point (8I + ii, 8J + jj, TK Kc + q); zero
// where that point is outside the grid.
// dst[tile][row][lane] = src at plane q of the tile, q = row, or with skew
// q = (row - ii - jj) mod TK (the register tiles' placement).
__global__ void k_tile_gather(Grid3 g, int TNI, int TNJ, int TK, int skew, size_t n,
                              const float* __restrict__ src, float* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (t >= n) return;
  const int lane = (int)(t & 63);
  const size_t pq = t >> 6;  // tile * TK + row
  const int row = (int)(pq % (size_t)TK);
  const int q = skew ? (row - (lane & 7) - (lane >> 3) + 2 * TK) % TK : row;
  const size_t tile = pq / (size_t)TK;
  const int I = (int)(tile % (size_t)TNI);
  const int J = (int)((tile / (size_t)TNI) % (size_t)TNJ);
  const int Kc = (int)(tile / ((size_t)TNI * TNJ));
  const int i = I * TI + (lane & 7), j = J * TJ + (lane >> 3), k = Kc * TK + q;
  dst[t] = (i < g.a && j < g.b && k < g.c)
               ? src[(size_t)i + (size_t)g.a * j + (size_t)g.ab * k]
               : 0.f;
}

int build_sweep_plan(SweepPlan& plan, const Grid3& g, int num_cus, int tile_k, int tile_waves,
                     const IluBlocks* blk, const float* dinv, int stage_one, const Metric* m) {
  free_sweep_plan(plan);
  if (g.c < 2) return 1;
  // tile diagonals
  plan.stage_one = (stage_one == 1 || stage_one == 2) ? stage_one : 0;
  plan.TK = (tile_k == 16 || tile_k == 64) ? tile_k : kTileK;
  // register tiles: 8x8 columns, at most 32 planes (TK + 2 SKEW <= 64)
  if (plan.stage_one == 2 && (tile_waves != 1 || plan.TK + 2 * SKEW > 64 || !m)) {
    free_sweep_plan(plan);
    return 3;
  }
  plan.SI = plan.SJ = 1;
  if (tile_waves == 4) plan.SI = plan.SJ = 2;
  if (tile_waves == 8) {
    plan.SI = 2;
    plan.SJ = 4;
    if (plan.TK == 64) plan.TK = kTileK;  // the LDS holds at most 16x32x32
  }
  plan.TNI = (g.a + TI * plan.SI - 1) / (TI * plan.SI);
  plan.TNJ = (g.b + TJ * plan.SJ - 1) / (TJ * plan.SJ);
  plan.NK = (g.c + plan.TK - 1) / plan.TK;
  plan.bl = IluBlocks{g.a, g.b, g.c, 1};
  plan.blocked = 0;
  if (blk && (blk->bi < g.a || blk->bj < g.b || blk->bk < g.c)) {
    const int ti = TI * plan.SI, tj = TJ * plan.SJ;
    if ((blk->bi < g.a && blk->bi % ti) || (blk->bj < g.b && blk->bj % tj) ||
        (blk->bk < g.c && blk->bk % plan.TK)) {
      free_sweep_plan(plan);
      return 2;
    }
    plan.bl = *blk;
    plan.blocked = 1;
  }
  // tile extents of one block (the whole grid without a cut)
  const int bIt = plan.blocked && plan.bl.bi < g.a ? plan.bl.bi / (TI * plan.SI) : plan.TNI;
  const int bJt = plan.blocked && plan.bl.bj < g.b ? plan.bl.bj / (TJ * plan.SJ) : plan.TNJ;
  const int bKt = plan.blocked && plan.bl.bk < g.c ? plan.bl.bk / plan.TK : plan.NK;
  plan.ndiag = bIt + bJt + bKt - 2;
  std::vector<std::vector<int>> by_diag(plan.ndiag);
  for (int Kc = 0; Kc < plan.NK; Kc++)
    for (int J = 0; J < plan.TNJ; J++)
      for (int I = 0; I < plan.TNI; I++) {
        const int L = (I % bIt) + (J % bJt) + (Kc % bKt);
        by_diag[L].push_back(I + plan.TNI * (J + plan.TNJ * Kc));
      }
  std::vector<int> tl;
  plan.diag_off.assign(plan.ndiag + 1, 0);
  for (int L = 0; L < plan.ndiag; L++) {
    plan.diag_off[L] = (int)tl.size();
    tl.insert(tl.end(), by_diag[L].begin(), by_diag[L].end());
  }
  plan.diag_off[plan.ndiag] = (int)tl.size();
  const size_t nface = (size_t)plan.TNI * plan.TNJ * plan.NK * plan.TK * 8;
  if (hipMalloc((void**)&plan.face_i, nface * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&plan.face_j, nface * sizeof(float)) != hipSuccess ||
      hipMalloc((void**)&plan.tiles, tl.size() * sizeof(int)) != hipSuccess) {
    free_sweep_plan(plan);
    return 1;
  }
  hipMemcpy(plan.tiles, tl.data(), tl.size() * sizeof(int), hipMemcpyHostToDevice);
  if (plan.blocked && plan.stage_one == 2) {
    // the block-persistent register-tile sweeps: every block's tiles by
    // diagonal, each diagonal in the order of `tiles` (a counting sort on
    // block-major, diagonal-minor keys)
    const int nbI = (plan.TNI + bIt - 1) / bIt, nbJ = (plan.TNJ + bJt - 1) / bJt;
    const int nbK = (plan.NK + bKt - 1) / bKt;
    plan.nblk = nbI * nbJ * nbK;
    auto key = [&](int t, int L) {
      const int I = t % plan.TNI, J = (t / plan.TNI) % plan.TNJ, Kc = t / (plan.TNI * plan.TNJ);
      return ((size_t)(I / bIt) + (size_t)nbI * ((size_t)(J / bJt) + (size_t)nbJ * (Kc / bKt))) *
                 plan.ndiag + L;
    };
    std::vector<int> bo((size_t)plan.nblk * plan.ndiag + 1, 0);
    for (int L = 0; L < plan.ndiag; L++)
      for (int t : by_diag[L]) bo[key(t, L) + 1]++;
    for (size_t q = 1; q < bo.size(); q++) bo[q] += bo[q - 1];
    std::vector<int> at(bo.begin(), bo.end() - 1), bt(tl.size());
    for (int L = 0; L < plan.ndiag; L++)
      for (int t : by_diag[L]) bt[at[key(t, L)]++] = t;
    if (hipMalloc((void**)&plan.blk_tiles, bt.size() * sizeof(int)) != hipSuccess ||
        hipMalloc((void**)&plan.blk_off, bo.size() * sizeof(int)) != hipSuccess) {
      free_sweep_plan(plan);
      return 1;
    }
    hipMemcpy(plan.blk_tiles, bt.data(), bt.size() * sizeof(int), hipMemcpyHostToDevice);
    hipMemcpy(plan.blk_off, bo.data(), bo.size() * sizeof(int), hipMemcpyHostToDevice);
    // the small diagonals at either end: fewer tiles than the device holds
    // at once (24 waves per CU)
    const int small = 24 * (num_cus > 0 ? num_cus : 256);
    auto cnt = [&](int L) { return plan.diag_off[L + 1] - plan.diag_off[L]; };
    plan.blk_lo = 0;
    while (plan.blk_lo < plan.ndiag && cnt(plan.blk_lo) < small) plan.blk_lo++;
    plan.blk_hi = plan.ndiag;
    while (plan.blk_hi > plan.blk_lo && cnt(plan.blk_hi - 1) < small) plan.blk_hi--;
    // two 8-wave workgroups per CU hold every block at once
    plan.blk_roomy = plan.nblk <= 2 * (num_cus > 0 ? num_cus : 256);
  }
  if (plan.SI == 1 && plan.SJ == 1) {
    // tile-major dinv and y of the single-wave sweeps: every tile reads and
    // writes whole 256-byte planes instead of 32-byte row pieces
    plan.tile_elems = (size_t)plan.TNI * plan.TNJ * plan.NK * plan.TK * 64;
    const size_t ntile = (size_t)plan.TNI * plan.TNJ * plan.NK;
    if (hipMalloc((void**)&plan.dinv_t, plan.tile_elems * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&plan.yv_t, plan.tile_elems * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&plan.face_k, ntile * 64 * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&plan.face_kt, (size_t)plan.TNI * plan.TNJ * 64 * sizeof(float)) !=
            hipSuccess) {
      free_sweep_plan(plan);
      return 1;
    }
    if (dinv) {
      k_tile_gather<<<(unsigned)((plan.tile_elems + 255) / 256), 256>>>(
          g, plan.TNI, plan.TNJ, plan.TK, plan.stage_one == 2, plan.tile_elems, dinv,
          plan.dinv_t);
    } else {
      hipMemset(plan.dinv_t, 0, plan.tile_elems * sizeof(float));
    }
    if (plan.stage_one == 2) {
      const size_t ncol = (size_t)plan.TNI * plan.TNJ;
      if (hipMalloc((void**)&plan.coup_lo, ncol * 192 * sizeof(float)) != hipSuccess ||
          hipMalloc((void**)&plan.coup_up, ncol * 192 * sizeof(float)) != hipSuccess) {
        free_sweep_plan(plan);
        return 1;
      }
      k_reg_coup<<<(unsigned)ncol, 64>>>(*m, g, plan.TNI, plan.TNJ, plan.coup_lo, plan.coup_up);
    }
    if (hipDeviceSynchronize() != hipSuccess) {
      free_sweep_plan(plan);
      return 1;
    }
  }
  plan.ready = 1;
  return 0;
}

void free_sweep_plan(SweepPlan& plan) {
  if (plan.tiles) hipFree(plan.tiles);
  if (plan.face_i) hipFree(plan.face_i);
  if (plan.face_j) hipFree(plan.face_j);
  if (plan.dinv_t) hipFree(plan.dinv_t);
  if (plan.yv_t) hipFree(plan.yv_t);
  if (plan.face_k) hipFree(plan.face_k);
  if (plan.face_kt) hipFree(plan.face_kt);
  if (plan.coup_lo) hipFree(plan.coup_lo);
  if (plan.coup_up) hipFree(plan.coup_up);
  if (plan.blk_tiles) hipFree(plan.blk_tiles);
  if (plan.blk_off) hipFree(plan.blk_off);
  plan = SweepPlan();
}

template <typename T>
void launch_dilu_tiles(const SweepPlan& plan, const Metric& m, const Grid3& g, const float* dinv,
                       const T* r, float* yv, T* z, const CgState* guard, hipStream_t s) {
  tiles_launch_lds<T>(plan, m, g, dinv, r, yv, z, guard, s);
}

bool dilu_tiles_fuse_rupd(const SweepPlan& plan) {
  const int w = plan.SI * plan.SJ;
  return plan.stage_one == 2 && w != 4 && w != 8 && (plan.TK == 16 || plan.TK == 32);
}

void launch_dilu_tiles_rupd(const SweepPlan& plan, const Metric& m, const Grid3& g, float* r,
                            const float* ap, float* z, const CgState* guard, hipStream_t s) {
  if (plan.TK == 16)
    tiles_launch_reg<float, 16>(plan, m, g, r, z, guard, s, ap);
  else
    tiles_launch_reg<float, 32>(plan, m, g, r, z, guard, s, ap);
}

void launch_dilu_tiles_stored(const SweepPlan& plan, const CoefArrays& ca, const Grid3& g,
                              const float* dinv, const double* r, float* yv, double* z,
                              hipStream_t s) {
  TileStored cs;
  cs.ca = ca;
  tiles_launch<double, TileStored>(plan, cs, g, dinv, r, yv, z, nullptr, s);
}

template void launch_dilu_tiles<double>(const SweepPlan&, const Metric&, const Grid3&,
                                        const float*, const double*, float*, double*,
                                        const CgState*, hipStream_t);
template void launch_dilu_tiles<float>(const SweepPlan&, const Metric&, const Grid3&,
                                       const float*, const float*, float*, float*,
                                       const CgState*, hipStream_t);

}  // namespace p3d
