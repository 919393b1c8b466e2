// Tile-diagonal D-ILU triangular solves.
//
// The (i,j,k) grid is cut into 8x8xTK tiles (or (8 SI)x(8 SJ)xTK with
// several waves per tile); one launch processes every tile of one
// anti-diagonal I+J+Kc, so a tile's faces come from tiles finished by
// earlier launches and no workgroup ever waits for another.  Inside a tile
// one wave owns an 8x8 column stack (lane = ii + 8*jj) and marches along k
// (phi): lane (ii,jj) handles plane k = s - ii - jj at step s, so its (i-1)
// and (j-1) neighbours finished the same plane one step earlier and hand
// their values over by cross-lane shuffle; (k-1) and the phi wrap partner
// stay in the lane's registers.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "p3d_kernels.hpp"

namespace p3d {

struct SweepPlan {
  int ready = 0;
  // tile-diagonal schedule ((8 SI) x (8 SJ) x TK tiles, one launch per
  // anti-diagonal; SI x SJ waves per tile)
  int TK = 0, NK = 0, ndiag = 0;
  int SI = 1, SJ = 1, TNI = 0, TNJ = 0;
  // block-Jacobi cut: tiles are scheduled by their diagonal inside their
  // block (the blocks are independent), tile indices come from `tiles`
  int blocked = 0;
  IluBlocks bl{};
  float* face_i = nullptr;  // compact tile faces of the single-wave tile sweeps
  float* face_j = nullptr;
  int* tiles = nullptr;            // tiles ordered by diagonal
  std::vector<int> diag_off;       // ndiag+1 offsets into tiles
  // single-wave sweeps: dinv and the FP32 sweep vector in tile-major order
  // (tile I + TNI (J + TNJ Kc), then plane, then lane ii + 8 jj)
  size_t tile_elems = 0;
  float* dinv_t = nullptr;
  float* yv_t = nullptr;
  float* face_k = nullptr;   // per tile: backward z of its first plane
  float* face_kt = nullptr;  // per column of tiles: backward z of plane c-1
  // single-wave tiles: 1 when the marching wave stages its tile alone (one
  // wave and about 20 KB of LDS per tile), 0 when TK/8 waves stage it, 2
  // when the marching wave holds its tile in registers (one wave, about
  // 2.4 KB of LDS; dinv_t and yv_t then hold plane q of lane ii + 8 jj in
  // row (q + ii + jj) mod TK)
  int stage_one = 0;
  // register tiles: per-column couplings (lower, upper), 3 x 64 floats per
  // column of tiles
  float* coup_lo = nullptr;
  float* coup_up = nullptr;
  // block-Jacobi register tiles: the tiles of each of the nblk blocks by
  // diagonal (blk_off: nblk * ndiag + 1 offsets into blk_tiles); persist 1
  // runs each sweep as one launch, a workgroup per block marching its
  // diagonals, 2 so only the small diagonals [0, blk_lo) and [blk_hi,
  // ndiag) at either end (one launch each) and the others one launch per
  // diagonal, 0 one launch per diagonal (all the same bits)
  int nblk = 0;
  int* blk_tiles = nullptr;
  int* blk_off = nullptr;
  int blk_lo = 0, blk_hi = 0;
  int blk_roomy = 0;  // nblk <= 2 per CU: the block kernels' larger register budget
  int persist = 0;
};

// Returns 0 when the tile sweeps can run on this grid.
// tile_k: planes per tile of the tile-diagonal schedule (16, 32 or 64).
// tile_waves: waves per tile, 1 (8x8), 4 (16x16) or 8 (16x32; tile_k <= 32).
// bl: block-Jacobi cut of the ILU (null: none); its extents must be
// multiples of the tile extents (returns nonzero otherwise).
// dinv: the D-ILU's inverse pivots (natural order), copied tile-major for
// the single-wave tile sweeps (build the plan again after refactoring).
// stage_one: see SweepPlan::stage_one; 2 needs tile_waves 1, tile_k 16 or
// 32 and the operator's metric m (returns 3 otherwise).
int build_sweep_plan(SweepPlan& plan, const Grid3& g, int num_cus, int tile_k = 32,
                     int tile_waves = 1, const IluBlocks* bl = nullptr,
                     const float* dinv = nullptr, int stage_one = 0,
                     const Metric* m = nullptr);
void free_sweep_plan(SweepPlan& plan);
// Tile-diagonal sweeps: z = (LU)^-1 r with 2*ndiag launches and no waiting
// between workgroups (the schedule used inside the CG iteration).
template <typename T>
void launch_dilu_tiles(const SweepPlan& plan, const Metric& m, const Grid3& g, const float* dinv,
                       const T* r, float* yv, T* z, const CgState* guard, hipStream_t s);

// The FP32 register-tile sweeps with the residual update of the recurrence
// folded into the forward sweep: r -= alpha Ap (alpha from guard, which must
// be non-null), then z = M^-1 r; available when dilu_tiles_fuse_rupd(plan).
bool dilu_tiles_fuse_rupd(const SweepPlan& plan);
void launch_dilu_tiles_rupd(const SweepPlan& plan, const Metric& m, const Grid3& g, float* r,
                            const float* ap, float* z, const CgState* guard, hipStream_t s);

// Tile-diagonal sweeps with the stored FP32 couplings of a legacy CSR input.
void launch_dilu_tiles_stored(const SweepPlan& plan, const CoefArrays& ca, const Grid3& g,
                              const float* dinv, const double* r, float* yv, double* z,
                              hipStream_t s);

}  // namespace p3d
