// Host side of the whole-solve C-ABI (include/pot3d_hip.h, group 2).
//
// One solver object owns its device buffers and one HIP stream (default
// priority).  The PCG of cgsolve (src/pot3d.F90:4231-4369) runs with all
// scalars on the device: every reduction's last workgroup computes alpha,
// beta and the ernorm test (:4371-4485) and raises a device "done" flag;
// iterations are replayed as captured hipGraphs of K iterations whose
// kernels return at once after the flag is set; the host reads the state
// once per graph.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <vector>

#include "../../include/pot3d_hip.h"
#include "p3d_kernels.hpp"
#include "p3d_sweep.hpp"

using namespace p3d;

static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

namespace p3d {
// Error message shared by every entry point of the library (p3d_last_error).
int set_error(int code, const char* msg) {
  g_err = msg;
  return code;
}
}  // namespace p3d

#define HIPCHK(expr)                                                                \
  do {                                                                              \
    hipError_t e_ = (expr);                                                         \
    if (e_ != hipSuccess)                                                           \
      return fail(P3D_EHIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                  __FILE__, __LINE__);                                              \
  } while (0)

template <typename T>
struct Vecs {
  T *x, *r, *p, *ap, *z;
  double* xs = nullptr;             // FP64 solution updated instead of x (FP32 recurrence)
  const double* xscale = nullptr;  // scale of the FP32 system (device)
  // FP32 recurrence with the FP64 solution on tile-major r, ap, z: the
  // kXBatch rotating direction buffers (pb.p[0] = p), x updated per group
  PBuf pb{};
};

struct p3d_solver {
  int nr, nt, np;
  int ifprec;
  double pm_r1, dr1, drn;
  Grid3 g;
  Metric m;
  Launch L;
  hipStream_t stream;
  // device memory
  double* tables = nullptr;
  float* ftables = nullptr;  // FDPH | FDPL | FDPU
  double* mesh = nullptr;  // dr, rh, drh, dt, dth, sth, dp, dph (full length)
  EnergyMesh em;
  double* br0 = nullptr;
  double* br1 = nullptr;
  bool have_br0 = false, have_br1 = false;
  double *x = nullptr, *r = nullptr, *p = nullptr, *ap = nullptr, *z = nullptr, *b = nullptr;
  float* yv = nullptr;
  float* dinv = nullptr;
  float *x32 = nullptr, *r32 = nullptr, *p32 = nullptr, *ap32 = nullptr, *z32 = nullptr;
  float* p32b[kXBatch - 1] = {};  // the other direction buffers of the grouped x updates
  // r, z (and r32, z32) hold nvec >= N entries: with the single-wave tile
  // sweeps they are tile-major (tm.on), every other vector is natural
  size_t nvec = 0;
  TileMap tm;
  Launch Lt;  // launch shape of a pass over the tile-major entries (tm.on)
  Grid3 gt;   // g with N = the tile-major entries (contiguous FP32 passes)
  double* pole_part = nullptr;
  CgState* st = nullptr;
  CgState* st_host = nullptr;  // pinned
  double* hist = nullptr;
  int hist_cap = 0;
  Reduce red;
  double* out = nullptr;   // device scratch (8 doubles)
  double* out_host = nullptr;  // pinned
  double* phi = nullptr;
  double* sums = nullptr;
  double* e_partial = nullptr;
  int eG = 0;
  SweepPlan sweep;
  bool sweep_ready = false;
  // options
  int stop_mode = 0;
  int mixed = 0;
  double inner_tol = 1e-5;
  int chunk = 0;
  int use_graph = 1;
  int sweep_mode = 0;
  int verbose = 0;
  int guess_check = 0;
  int tile_k = 32, tile_waves = 1, tile_stage = 0;
  int apply_fp32 = 0;  // p3d_apply_A with the FP32 operator
  int fuse_rupd = -1;  // P3D_OPT_FUSE_RESIDUAL
  int blk_sweep = -1;  // P3D_OPT_BLOCK_SWEEP
  IluBlocks blocks{0, 0, 0, 1};  // block-Jacobi cut of the ILU (none by default)
  int nblk[3] = {1, 1, 1};       // blocks per axis (r, theta, phi)
  // graphs: [0] FP64, [1] FP32, [2] FP32 recurrence with the FP64 solution
  hipGraphExec_t gexec[3] = {nullptr, nullptr, nullptr};
  int gK[3] = {0, 0, 0};
  // stats
  p3d_stats stats;
  int last_hist = 0;
};

// ---------------------------------------------------------------------------
static int dalloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) {
    *p = nullptr;
    return fail(P3D_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  return P3D_OK;
}

template <typename A, typename B>
__global__ void k_convert(uint32_t n, const A* __restrict__ src, B* __restrict__ dst) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) dst[q] = (B)src[q];
}

static int ensure_mixed(p3d_solver* s) {
  if (s->x32) return P3D_OK;
  const size_t nb = (size_t)s->g.N * sizeof(float), nv = s->nvec * sizeof(float);
  int rc;
  if ((rc = dalloc((void**)&s->x32, nb)) || (rc = dalloc((void**)&s->r32, nv)) ||
      (rc = dalloc((void**)&s->p32, nb)) || (rc = dalloc((void**)&s->ap32, nv)) ||
      (rc = dalloc((void**)&s->z32, nv)))
    return rc;
  HIPCHK(hipMemsetAsync(s->r32, 0, nv, s->stream));
  HIPCHK(hipMemsetAsync(s->z32, 0, nv, s->stream));
  HIPCHK(hipMemsetAsync(s->ap32, 0, nv, s->stream));
  return P3D_OK;
}

// The other direction buffers of the FP32 recurrence's grouped solution
// updates (natural order, N floats each).
static int ensure_pbuf(p3d_solver* s) {
  int rc;
  for (int q = 0; q < kXBatch - 1; q++)
    if (!s->p32b[q] && (rc = dalloc((void**)&s->p32b[q], (size_t)s->g.N * sizeof(float))))
      return rc;
  return P3D_OK;
}

// Placement of r and z: tile-major exactly when the single-wave tile
// sweeps apply the ILU.  A change of placement drops the captured graphs
// and clears r and z (their entries outside the grid stay zero).
static int update_tilemap(p3d_solver* s) {
  TileMap t;
  if (s->ifprec == PREC_ILU && s->sweep_mode == 0 && s->sweep_ready && s->sweep.SI == 1 &&
      s->sweep.SJ == 1) {
    t.on = 1;
    t.TNI = s->sweep.TNI;
    t.TNJ = s->sweep.TNJ;
    t.tk_log2 = s->sweep.TK == 16 ? 4 : (s->sweep.TK == 64 ? 6 : 5);
  }
  if (t.on == s->tm.on && t.TNI == s->tm.TNI && t.TNJ == s->tm.TNJ && t.tk_log2 == s->tm.tk_log2)
    return P3D_OK;
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int q = 0; q < 3; q++) {
    if (s->gexec[q]) hipGraphExecDestroy(s->gexec[q]);
    s->gexec[q] = nullptr;
  }
  s->tm = t;
  if (t.on) {
    int dev = 0, ncu = 0;
    HIPCHK(hipGetDevice(&dev));
    HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    s->gt = s->g;
    s->gt.N = (uint32_t)s->sweep.tile_elems;
    // as many blocks as the reduction buffers hold (a multiple of 8)
    s->Lt = plan_stream(s->gt.N, ncu * 8 < s->red.max_blocks ? ncu * 8 : s->red.max_blocks);
    size_stream_grids(s->Lt, ncu, s->gt.N);
  }
  HIPCHK(hipMemsetAsync(s->r, 0, s->nvec * sizeof(double), s->stream));
  HIPCHK(hipMemsetAsync(s->z, 0, s->nvec * sizeof(double), s->stream));
  if (s->r32) HIPCHK(hipMemsetAsync(s->r32, 0, s->nvec * sizeof(float), s->stream));
  if (s->z32) HIPCHK(hipMemsetAsync(s->z32, 0, s->nvec * sizeof(float), s->stream));
  if (s->ap32) HIPCHK(hipMemsetAsync(s->ap32, 0, s->nvec * sizeof(float), s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

static int ensure_hist(p3d_solver* s, int ncgmax) {
  int want = ncgmax + 1;
  if (want > (1 << 22)) want = 1 << 22;
  if (want < 1) want = 1;
  if (s->hist_cap >= want) return P3D_OK;
  if (s->hist) hipFree(s->hist);
  s->hist = nullptr;
  s->hist_cap = 0;
  int rc = dalloc((void**)&s->hist, sizeof(double) * want);
  if (rc) return rc;
  s->hist_cap = want;
  return P3D_OK;
}

static int read_state(p3d_solver* s) {
  HIPCHK(hipMemcpyAsync(s->st_host, s->st, sizeof(CgState), hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

// ---------------------------------------------------------------------------
// Preconditioner application r -> z (+ fused dots with epilogue).
// ap_upd (FP32 register tiles, guard set): the residual update r -= alpha ap
// is folded into the forward sweep (r is written).
template <typename T>
static void prec_and_dot(p3d_solver* s, const T* r, T* z, Epi epi, double* out, bool guard,
                         const float* ap_upd = nullptr) {
  if (s->ifprec == PREC_ILU) {
    const CgState* gd = guard ? s->st : nullptr;
    if (s->sweep_mode == 0 && s->sweep_ready) {
      if constexpr (sizeof(T) == 4) {
        if (ap_upd)
          launch_dilu_tiles_rupd(s->sweep, s->m, s->g, const_cast<float*>(r), ap_upd, z, gd,
                                 s->stream);
        else
          launch_dilu_tiles<T>(s->sweep, s->m, s->g, s->dinv, r, s->yv, z, gd, s->stream);
      } else {
        launch_dilu_tiles<T>(s->sweep, s->m, s->g, s->dinv, r, s->yv, z, gd, s->stream);
      }
    } else {
      launch_dilu_solve_levels<T>(s->m, s->g, s->dinv, r, s->yv, z, nullptr, true, s->blocks, gd,
                                  s->stream);
    }
    if (epi != EPI_NONE) {
      const T* u = (epi == EPI_RZITER) ? nullptr : r;
      if (sizeof(T) == 4 && s->tm.on)  // FP32: contiguous, in tile order
        launch_dot2<T>(s->gt, s->Lt, r, z, u, u, s->st, s->red, 1, epi, out, guard, s->stream);
      else  // FP64: natural point order, as every other schedule
        launch_dot2<T>(s->g, s->L, r, z, u, u, s->st, s->red, 1, epi, out, guard, s->stream,
                       s->tm);
    }
  } else {
    launch_point_prec<T>(s->m, s->g, s->L, s->ifprec, r, z, s->st, s->red, 1, epi, out,
                         s->stream);
  }
}

// One CG iteration; t is its place in the group of kXBatch iterations of
// the FP32 recurrence's grouped solution updates (unused otherwise).
template <typename T>
static void cg_iteration(p3d_solver* s, Vecs<T> v, int t) {
  if constexpr (sizeof(T) == 4) {
    if (v.xs && s->tm.on && v.pb.p[0]) {
      // FP32 recurrence with the FP64 solution and tile-major r, ap, z:
      // every pass contiguous except the p update's read of z; direction
      // p_t in buffer t of the group, the FP64 x updated once per group
      float* p = v.pb.p[t];
      launch_pole_sums<float>(s->m, s->g, p, s->pole_part, s->stream);
      launch_stencil<float>(s->m, s->g, s->L, p, v.ap, nullptr, 0, s->pole_part, s->st, s->red,
                            0, EPI_PAP, nullptr, true, s->stream, s->tm);
      // r -= alpha ap, in the forward sweep of the register tiles where they
      // run, else on its own; then z = M^-1 r and r.z, r.r together
      const bool fused = (s->fuse_rupd == 1 || (s->fuse_rupd < 0 && s->g.N >= (1u << 26))) &&
                         s->sweep_mode == 0 && s->sweep_ready && dilu_tiles_fuse_rupd(s->sweep);
      if (!fused) launch_r_update(s->gt.N, s->Lt, v.r, v.ap, s->st, s->stream);
      prec_and_dot<float>(s, v.r, v.z, EPI_RITER, nullptr, true, fused ? v.ap : nullptr);
      launch_p_update_x(s->g, s->L, t, v.pb, v.z, s->st, s->stream, s->tm, v.xs, v.xscale,
                        s->red.counter + 5 * kRedCounters);
      return;
    }
  }
  launch_pole_sums<T>(s->m, s->g, v.p, s->pole_part, s->stream);
  launch_stencil<T>(s->m, s->g, s->L, v.p, v.ap, nullptr, 0, s->pole_part, s->st, s->red, 0,
                    EPI_PAP, nullptr, true, s->stream);
  launch_cg_update<T>(s->m, s->g, s->L, s->ifprec, v.x, v.r, v.p, v.ap, s->st, s->red, 2,
                      s->stream, v.xs, v.xscale, s->tm);
  if (s->ifprec == PREC_ILU) prec_and_dot<T>(s, v.r, v.z, EPI_RZITER, nullptr, true);
  launch_p_update<T>(s->m, s->g, s->L, s->ifprec, v.p, v.r, v.z, s->st, s->stream, s->tm);
}

template <typename T>
static int chunk_size(p3d_solver* s) {
  if (s->chunk > 0) return s->chunk;
  if (s->ifprec == PREC_ILU) {
    const int per = (s->sweep_mode == 0 && s->sweep_ready) ? 2 * s->sweep.ndiag + 6
                                                           : 2 * (s->g.a + s->g.b + s->g.c);
    int k = 4096 / per;
    if (k > 32) k = 32;
    // whole groups of the FP32 recurrence's grouped solution updates
    if (sizeof(T) == 4) k = k / kXBatch * kXBatch;
    return k < kXBatch ? kXBatch : k;
  }
  return 32;
}

template <typename T>
static int run_iterations(p3d_solver* s, Vecs<T> v, int K) {
  const int ti = sizeof(T) == 8 ? 0 : (v.xs ? 2 : 1);
  if (s->use_graph) {
    if (!s->gexec[ti] || s->gK[ti] != K) {
      if (s->gexec[ti]) hipGraphExecDestroy(s->gexec[ti]);
      s->gexec[ti] = nullptr;
      hipGraph_t graph;
      HIPCHK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
      for (int it = 0; it < K; it++) cg_iteration<T>(s, v, it % kXBatch);
      HIPCHK(hipStreamEndCapture(s->stream, &graph));
      hipError_t e = hipGraphInstantiate(&s->gexec[ti], graph, nullptr, nullptr, 0);
      hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(P3D_EHIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      s->gK[ti] = K;
    }
    HIPCHK(hipGraphLaunch(s->gexec[ti], s->stream));
  } else {
    for (int it = 0; it < K; it++) cg_iteration<T>(s, v, it % kXBatch);
    HIPCHK(hipGetLastError());
  }
  return P3D_OK;
}

// Replays chunks of CG iterations until the device's done flag is set.  The
// loop is also bounded on the host: every replay must advance the iteration
// count, and no more replays than the remaining ncgmax needs are allowed.
template <typename T>
static int run_until_done(p3d_solver* s, Vecs<T> v) {
  int rc;
  const int K = chunk_size<T>(s);
  const long max_launches = (long)(s->st_host->ncgmax - s->st_host->ncg) / K + 4;
  long launches = 0;
  while (!s->st_host->done) {
    const int ncg_before = s->st_host->ncg;
    if ((rc = run_iterations<T>(s, v, K))) return rc;
    if ((rc = read_state(s))) return rc;
    launches++;
    if (!s->st_host->done && (s->st_host->ncg <= ncg_before || launches > max_launches))
      return fail(P3D_ESTATE,
                  "CG state stopped advancing: ncg=%d (was %d) done=%d epsn=%g after %ld "
                  "graph replays of %d iterations",
                  s->st_host->ncg, ncg_before, s->st_host->done, s->st_host->epsn, launches, K);
  }
  return P3D_OK;
}

// PCG on vectors v (r holds the rhs on entry, x the initial guess).
// Returns P3D_OK with s->st_host holding the final state.
template <typename T>
static int pcg(p3d_solver* s, Vecs<T> v, double epscg, int ncgmax, int stop_mode) {
  int rc;
  if ((rc = ensure_hist(s, ncgmax))) return rc;
  CgState init;
  memset(&init, 0, sizeof init);
  init.epscg = epscg;
  init.ncgmax = ncgmax;
  init.stop_mode = stop_mode;
  init.hist = s->hist;
  init.hist_cap = s->hist_cap;
  *s->st_host = init;
  HIPCHK(hipMemcpyAsync(s->st, s->st_host, sizeof(CgState), hipMemcpyHostToDevice, s->stream));
  // bdotb = b . M^-1 b  (:4293-4298)
  prec_and_dot<T>(s, v.r, v.z, EPI_BDOTB, nullptr, false);
  if ((rc = read_state(s))) return rc;
  if (s->st_host->done == 4) {  // zero rhs: x = 0 (:4302-4311)
    HIPCHK(hipMemsetAsync(v.x, 0, sizeof(T) * s->g.N, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->st_host->epsn = 0.0;
    return P3D_OK;
  }
  // r = b - A x ; z = M^-1 r ; rdotr ; p = z  (:4317-4327)
  launch_pole_sums<T>(s->m, s->g, v.x, s->pole_part, s->stream);
  launch_stencil<T>(s->m, s->g, s->L, v.x, v.ap, nullptr, 0, s->pole_part, nullptr, s->red, 0,
                    EPI_NONE, nullptr, false, s->stream);
  launch_sub<T>(s->g, s->L, v.r, v.ap, s->stream, s->tm);
  prec_and_dot<T>(s, v.r, v.z, EPI_RINIT, nullptr, false);
  launch_copy<T>(s->g, s->L, v.p, v.z, s->stream, s->tm, 0);
  if ((rc = read_state(s))) return rc;
  return run_until_done<T>(s, v);
}

// ---------------------------------------------------------------------------
static void print_history(p3d_solver* s, int ncghist) {
  if (!s->verbose || ncghist <= 0) return;
  const int n = s->st_host->ncg + 1 < s->hist_cap ? s->st_host->ncg + 1 : s->hist_cap;
  std::vector<double> h(n);
  hipMemcpy(h.data(), s->hist, sizeof(double) * n, hipMemcpyDeviceToHost);
  for (int i = 0; i < n; i++)
    if (i % ncghist == 0) printf(" Iteration: %8d   Residual: %23.15E\n", i, h[i]);
  if (s->st_host->done == 1) printf(" Iteration: %8d   Residual: %23.15E\n", n - 1, h[n - 1]);
  fflush(stdout);
}

static int true_residual(p3d_solver* s, double* rr, double* bb) {
  // ap = b - A x (FP64), reductions into out[0] (r.r), out[2] (b.b)
  launch_pole_sums<double>(s->m, s->g, s->x, s->pole_part, s->stream);
  launch_stencil<double>(s->m, s->g, s->L, s->x, s->ap, s->b, 1, s->pole_part, nullptr, s->red,
                         3, EPI_STORE, s->out, false, s->stream);
  launch_dot2<double>(s->g, s->L, s->b, s->b, nullptr, nullptr, nullptr, s->red, 4, EPI_STORE,
                      s->out + 2, false, s->stream);
  HIPCHK(hipMemcpyAsync(s->out_host, s->out, 4 * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  *rr = s->out_host[0];
  *bb = s->out_host[2];
  return P3D_OK;
}

static int status_rc(int done) {
  if (done == 1 || done == 4) return P3D_OK;
  if (done == 2) return fail(P3D_ENOCONV, "the field solver did not converge (ncgmax reached)");
  if (done == 3) return fail(P3D_ENONFINITE, "non-finite residual in the CG iteration");
  return fail(P3D_ESTATE, "solver stopped in state %d", done);
}

// FP32 inner PCG inside FP64 defect correction: x += A^-1 (b - A x) until
// ||b - A x||2 / ||b||2 <= epscg.
static int solve_mixed(p3d_solver* s, double epscg, int ncgmax) {
  int rc;
  if ((rc = ensure_mixed(s))) return rc;
  double rr, bb;
  if ((rc = true_residual(s, &rr, &bb))) return rc;  // ap = b - A x
  int total = 0, outer = 0;
  const int max_outer = 60;
  s->stats.status = 1;
  if (bb == 0.0) {
    HIPCHK(hipMemsetAsync(s->x, 0, sizeof(double) * s->g.N, s->stream));
    s->stats.status = 4;
    s->stats.relres = s->stats.true_relres = 0.0;
    return P3D_OK;
  }
  Vecs<float> v{s->x32, s->r32, s->p32, s->ap32, s->z32};
  double relres = sqrt(rr / bb);
  while (relres > epscg) {
    if (!isfinite(relres)) {
      s->stats.status = 3;
      return fail(P3D_ENONFINITE, "non-finite residual in the defect correction");
    }
    if (outer >= max_outer || total >= ncgmax) {
      s->stats.status = 2;
      return fail(P3D_ENOCONV, "mixed-precision solve did not converge");
    }
    const double scale = sqrt(rr);
    // r32 = r64/scale ; x32 = 0
    s->out_host[4] = scale;
    s->out_host[5] = 1.0;
    HIPCHK(hipMemcpyAsync(s->out + 4, s->out_host + 4, 2 * sizeof(double), hipMemcpyHostToDevice,
                          s->stream));
    launch_cast_scale<float, double>(s->g, s->r32, s->ap, s->out + 5, s->out + 4, 0, s->stream,
                                     s->tm);
    HIPCHK(hipMemsetAsync(s->x32, 0, sizeof(float) * s->g.N, s->stream));
    double tol = s->inner_tol;
    const double need = 0.5 * epscg / relres;
    if (need > tol) tol = need;
    if ((rc = pcg<float>(s, v, tol, ncgmax - total, 1))) return rc;
    total += s->st_host->ncg;
    if (s->st_host->done == 3) {
      s->stats.status = 3;
      return fail(P3D_ENONFINITE, "non-finite residual in the inner solve");
    }
    // x += x32 * scale
    launch_cast_scale<double, float>(s->g, s->x, s->x32, s->out + 5, s->out + 4, 1, s->stream);
    if ((rc = true_residual(s, &rr, &bb))) return rc;
    relres = sqrt(rr / bb);
    outer++;
    s->stats.ncg = total;
    s->stats.n_outer = outer;
    if (s->verbose) {
      fprintf(stderr, "[p3d] mixed: correction %d at ncg=%d, ||b-Ax||/||b|| = %.3e\n", outer,
              total, relres);
      fflush(stderr);
    }
  }
  s->stats.ncg = total;
  s->stats.n_outer = outer;
  s->stats.relres = relres;
  s->stats.true_relres = relres;
  s->stats.epsn = s->st_host->epsn;
  return P3D_OK;
}

// FP32 PCG recurrence with an FP64 solution and reliable updates (residual
// replacement).  The Krylov recurrence (r, p, ap, z) runs in FP32 on vectors
// scaled by 1/||b - A x0||; every solution update goes straight to the FP64
// x (x += alpha scale p), so x does not carry FP32 rounding, whose effect on
// b - A x grows with cond(A).  Whenever the recurrence residual has fallen
// by the factor kReliableDelta since the last replacement, it is replaced by
// the FP64 b - A x and the recurrence continues (beta from the replaced
// residual, p = z + beta p): the Krylov space is not restarted as in
// solve_mixed.  Ends when ||b - A x||2/||b||2 <= epscg (true residual).
static constexpr double kReliableDelta = 1e-3;

static int solve_mixed_ru(p3d_solver* s, double epscg, int ncgmax) {
  int rc;
  if ((rc = ensure_mixed(s)) || (rc = ensure_hist(s, ncgmax))) return rc;
  double rr, bb;
  if ((rc = true_residual(s, &rr, &bb))) return rc;  // ap = b - A x
  s->stats.status = 1;
  if (bb == 0.0) {
    launch_fill(s->x, 0.0, (size_t)s->g.N, s->stream);
    s->stats.status = 4;
    return P3D_OK;
  }
  double relres = sqrt(rr / bb);
  if (!isfinite(relres)) {
    s->stats.status = 3;
    return fail(P3D_ENONFINITE, "non-finite residual of the initial guess");
  }
  s->stats.relres = s->stats.true_relres = relres;
  if (relres <= epscg) return P3D_OK;
  const double scale = sqrt(rr);
  s->out_host[4] = scale;
  s->out_host[5] = 1.0;
  HIPCHK(hipMemcpyAsync(s->out + 4, s->out_host + 4, 2 * sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  // the recurrence's solution updates go straight to the FP64 x, grouped
  // kXBatch iterations at a time with the tile-major placement
  Vecs<float> v{s->x32, s->r32, s->p32, s->ap32, s->z32, s->x, s->out + 4};
  if (s->tm.on) {
    if ((rc = ensure_pbuf(s))) return rc;
    v.pb.p[0] = s->p32;
    for (int q = 1; q < kXBatch; q++) v.pb.p[q] = s->p32b[q - 1];
  }
  // the residual update runs over every entry of the tile-major r and ap:
  // entries outside the grid must be zero in both (other users of the
  // buffers may have written them)
  HIPCHK(hipMemsetAsync(s->ap32, 0, s->nvec * sizeof(float), s->stream));
  HIPCHK(hipMemsetAsync(s->r32, 0, s->nvec * sizeof(float), s->stream));
  launch_cast_scale<float, double>(s->g, v.r, s->ap, s->out + 5, s->out + 4, 0, s->stream,
                                   s->tm);
  // the inner relres = ||r32|| scale / ||b|| estimates the true one
  const double final_tgt = 0.8 * epscg;
  CgState init;
  memset(&init, 0, sizeof init);
  init.epscg = fmax(kReliableDelta * relres, final_tgt);
  init.ncgmax = ncgmax;
  init.stop_mode = 1;
  init.hist = s->hist;
  init.hist_cap = s->hist_cap;
  init.bdotb = 1.0;
  init.bb = bb / rr;
  *s->st_host = init;
  HIPCHK(hipMemcpyAsync(s->st, s->st_host, sizeof(CgState), hipMemcpyHostToDevice, s->stream));
  prec_and_dot<float>(s, v.r, v.z, EPI_RINIT, nullptr, false);
  launch_copy<float>(s->g, s->L, v.p, v.z, s->stream, s->tm, 0);
  if ((rc = read_state(s))) return rc;
  int replacements = 0;
  for (;;) {
    if ((rc = run_until_done<float>(s, v))) return rc;
    const int done = s->st_host->done;
    if (done == 3) {
      s->stats.status = 3;
      return fail(P3D_ENONFINITE, "non-finite residual in the inner solve");
    }
    // ap = b - A x
    if ((rc = true_residual(s, &rr, &bb))) return rc;
    relres = sqrt(rr / bb);
    s->stats.ncg = s->st_host->ncg;
    s->stats.n_outer = ++replacements;
    if (s->verbose) {
      fprintf(stderr, "[p3d] mixed: replacement %d at ncg=%d, ||b-Ax||/||b|| = %.3e\n",
              replacements, s->st_host->ncg, relres);
      fflush(stderr);
    }
    if (!isfinite(relres)) {
      s->stats.status = 3;
      return fail(P3D_ENONFINITE, "non-finite residual in the defect correction");
    }
    if (relres <= epscg) break;
    if (done == 2 || s->st_host->ncg >= ncgmax) {
      s->stats.status = 2;
      s->stats.relres = s->stats.true_relres = relres;
      return fail(P3D_ENOCONV, "mixed-precision solve did not converge");
    }
    // replace the residual and continue the recurrence: r32 = ap/scale ;
    // z = M^-1 r32 ; beta = r.z / (r.z of the last direction) ; p = z + beta p
    launch_cast_scale<float, double>(s->g, v.r, s->ap, s->out + 5, s->out + 4, 0, s->stream,
                                   s->tm);
    s->st_host->epscg = fmax(kReliableDelta * relres, final_tgt);
    s->st_host->done = 0;
    HIPCHK(hipMemcpyAsync(s->st, s->st_host, sizeof(CgState), hipMemcpyHostToDevice,
                          s->stream));
    prec_and_dot<float>(s, v.r, v.z, EPI_REPLACE, nullptr, false);
    // the recurrence's direction to buffer 0, where the next group starts
    const int pc = s->st_host->p_cur;
    if (v.pb.p[0] && pc > 0 && pc < kXBatch) {
      launch_copy<float>(s->g, s->L, v.p, v.pb.p[pc], s->stream);
      s->st_host->p_cur = 0;
      HIPCHK(hipMemcpyAsync(&s->st->p_cur, &s->st_host->p_cur, sizeof(int),
                            hipMemcpyHostToDevice, s->stream));
    }
    launch_p_update<float>(s->m, s->g, s->L, s->ifprec, v.p, v.r, v.z, s->st, s->stream, s->tm);
    if ((rc = read_state(s))) return rc;
  }
  s->stats.relres = s->stats.true_relres = relres;
  s->stats.epsn = s->st_host->epsn;
  return P3D_OK;
}

// Core solve on the solver's own buffers: s->x guess, s->b rhs.
static int solve_core(p3d_solver* s, double epscg, int ncgmax, int ncghist) {
  int rc;
  auto t0 = std::chrono::steady_clock::now();
  memset(&s->stats, 0, sizeof s->stats);
  if (s->guess_check) {
    // a proposed guess that is worse than the reference's zero guess is
    // replaced by zero: mode 1 compares ||b - A x0|| with ||b||; mode 2
    // compares the A-norm errors, ||x* - x0||_A < ||x*||_A <=> x0.A x0 -
    // 2 b.x0 < 0 (the norm CG minimises)
    bool keep;
    if (s->guess_check == 2) {
      launch_pole_sums<double>(s->m, s->g, s->x, s->pole_part, s->stream);
      launch_stencil<double>(s->m, s->g, s->L, s->x, s->ap, nullptr, 0, s->pole_part, nullptr,
                             s->red, 3, EPI_STORE, s->out, false, s->stream);
      launch_dot2<double>(s->g, s->L, s->b, s->x, nullptr, nullptr, nullptr, s->red, 4,
                          EPI_STORE, s->out + 2, false, s->stream);
      HIPCHK(hipMemcpyAsync(s->out_host, s->out, 4 * sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
      HIPCHK(hipStreamSynchronize(s->stream));
      keep = s->out_host[0] - 2.0 * s->out_host[2] < 0.0;
    } else {
      double rr, bb;
      if ((rc = true_residual(s, &rr, &bb))) return rc;
      keep = rr < bb;
    }
    if (keep) {
      s->stats.guess = 1;
    } else {
      launch_fill(s->x, 0.0, (size_t)s->g.N, s->stream);
      s->stats.guess = 2;
    }
  }
  if (s->mixed) {
    rc = s->mixed == 2 ? solve_mixed_ru(s, epscg, ncgmax) : solve_mixed(s, epscg, ncgmax);
    HIPCHK(hipStreamSynchronize(s->stream));
  } else {
    launch_copy<double>(s->g, s->L, s->r, s->b, s->stream, s->tm, 1);
    Vecs<double> v{s->x, s->r, s->p, s->ap, s->z};
    if ((rc = pcg<double>(s, v, epscg, ncgmax, s->stop_mode))) return rc;
    print_history(s, ncghist);
    s->stats.ncg = s->st_host->ncg;
    s->stats.status = s->st_host->done;
    s->stats.epsn = s->st_host->epsn;
    s->stats.relres = s->st_host->relres;
    double rr, bb;
    if ((rc = true_residual(s, &rr, &bb))) return rc;
    s->stats.true_relres = bb > 0 ? sqrt(rr / bb) : 0.0;
    rc = status_rc(s->st_host->done);
  }
  s->last_hist = s->st_host->ncg + 1;
  auto t1 = std::chrono::steady_clock::now();
  s->stats.t_solve = std::chrono::duration<double>(t1 - t0).count();
  return rc;
}

// D-ILU factorisation (FP64 pivots, FP32 inverses) of the current
// block-Jacobi cut.  A zero or non-finite pivot, or one whose FP32 inverse
// overflows, fails with P3D_ENONFINITE: the reference's host ilu0 prints a
// zero pivot and goes on (src/pot3d.F90:4695-4697), which would leave the
// iteration running on inf/NaN.
static int factor_ilu(p3d_solver* s) {
  double* dpiv = nullptr;
  unsigned* bad = nullptr;
  int rc = dalloc((void**)&dpiv, sizeof(double) * s->g.N);
  if (rc) return rc;
  if ((rc = dalloc((void**)&bad, sizeof(unsigned)))) {
    hipFree(dpiv);
    return rc;
  }
  unsigned first = 0xffffffffu;
  double dv = 0.0;
  hipError_t e = hipMemcpyAsync(bad, &first, sizeof first, hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    launch_dilu_factor(s->m, s->g, s->blocks, dpiv, s->dinv, bad, s->stream);
    e = hipStreamSynchronize(s->stream);
  }
  if (e == hipSuccess) e = hipMemcpy(&first, bad, sizeof first, hipMemcpyDeviceToHost);
  if (e == hipSuccess && first != 0xffffffffu)
    e = hipMemcpy(&dv, dpiv + first, sizeof dv, hipMemcpyDeviceToHost);
  hipFree(dpiv);
  hipFree(bad);
  if (e != hipSuccess)
    return fail(P3D_EHIP, "D-ILU factorisation failed: %s", hipGetErrorString(e));
  if (first != 0xffffffffu) {
    const int a = s->g.a, b = s->g.b;
    const int i = (int)(first % (unsigned)a), j = (int)((first / (unsigned)a) % (unsigned)b);
    const int k = (int)(first / ((unsigned)a * (unsigned)b));
    return fail(P3D_ENONFINITE,
                "ILU0 pivot %g at row %u (padded point (%d,%d,%d)): zero, not finite or "
                "beyond FP32 range",
                dv, first + 1, i + 2, j + 2, k + 2);
  }
  return P3D_OK;
}

// ---------------------------------------------------------------------------
// C-ABI
extern "C" {

const char* p3d_last_error(void) { return g_err.c_str(); }

int p3d_destroy(p3d_solver* s) {
  if (!s) return P3D_OK;
  if (s->stream) hipStreamSynchronize(s->stream);
  for (int t = 0; t < 3; t++)
    if (s->gexec[t]) hipGraphExecDestroy(s->gexec[t]);
  void* bufs[] = {s->tables, s->ftables, s->mesh, s->br0, s->br1, s->x, s->r, s->p, s->ap, s->z, s->b,
                  s->yv, s->dinv, s->x32, s->r32, s->p32, s->ap32, s->z32, s->pole_part,
                  s->st, s->hist, s->red.partial, s->red.counter, s->out, s->phi, s->sums,
                  s->e_partial};
  for (void* b : bufs)
    if (b) hipFree(b);
  for (float* b : s->p32b)
    if (b) hipFree(b);
  free_sweep_plan(s->sweep);
  if (s->st_host) hipHostFree(s->st_host);
  if (s->out_host) hipHostFree(s->out_host);
  if (s->stream) hipStreamDestroy(s->stream);
  delete s;
  return P3D_OK;
}

static int block_sweep_on(const p3d_solver* s);

int p3d_create(p3d_solver** out, int nr, int nt, int np, const double* r, const double* dr,
               const double* rh, const double* drh, const double* st, const double* dt,
               const double* sth, const double* dth, const double* dp, const double* dph,
               double pm_r1, int ifprec) {
  if (!out) return fail(P3D_EINVAL, "null output handle");
  *out = nullptr;
  if (nr < 3 || nt < 3 || np < 4)
    return fail(P3D_EINVAL, "grid %dx%dx%d too small (need nr>=3, nt>=3, np>=4)", nr, nt, np);
  if (!r || !dr || !rh || !drh || !st || !dt || !sth || !dth || !dp || !dph)
    return fail(P3D_EINVAL, "null mesh array");
  if (ifprec < 0 || ifprec > 2) return fail(P3D_EINVAL, "ifprec=%d not in 0..2", ifprec);
  if (!(pm_r1 == 1.0 || pm_r1 == -1.0)) return fail(P3D_EINVAL, "pm_r1 must be +1 or -1");
  const uint64_t a = nr - 2, b = nt - 2, c = np - 2;
  const uint64_t N = a * b * c;
  if ((uint64_t)nr * nt * np >= (1ull << 31))
    return fail(P3D_EINVAL, "grid too large for 32-bit point indices (%llu points)",
                (unsigned long long)((uint64_t)nr * nt * np));
  if (ifprec == 2 && np < 8)
    return fail(P3D_EINVAL, "ILU0 needs np >= 8 (phi wrap must not create fill)");
  p3d_solver* s = new p3d_solver();
  memset(&s->stats, 0, sizeof s->stats);
  s->nr = nr;
  s->nt = nt;
  s->np = np;
  s->ifprec = ifprec;
  s->pm_r1 = pm_r1;
  s->dr1 = dr[0];
  s->drn = dr[nr - 2];
  s->g.a = (int)a;
  s->g.b = (int)b;
  s->g.c = (int)c;
  s->g.N = (uint32_t)N;
  s->g.ab = (uint32_t)(a * b);
  s->g.da = make_fastdiv((uint32_t)a);
  s->g.dab = make_fastdiv((uint32_t)(a * b));
  int dev = 0;
  hipDeviceProp_t prop;
  hipGetDevice(&dev);
  int ncu = 256;
  if (hipGetDeviceProperties(&prop, dev) == hipSuccess && prop.multiProcessorCount > 0)
    ncu = prop.multiProcessorCount;
  s->L = plan_stream((uint32_t)N, ncu * 8);
  size_stream_grids(s->L, ncu, (uint32_t)N);
  int rc = P3D_OK;
  auto bail = [&](int code) {
    p3d_destroy(s);
    return code;
  };
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
    return bail(fail(P3D_EHIP, "hipStreamCreate failed"));
  // metric factor tables
  const double PI = 3.1415926535897932;
  const double pl_i = 1.0 / (2.0 * PI - 0.0);
  std::vector<double> T(3 * a + 4 * b + 4 * c);
  double* DRH = T.data();
  double* CRL = DRH + a;
  double* CRH = CRL + a;
  double* SD = CRH + a;
  double* CP = SD + b;
  double* CTL = CP + b;
  double* CTH = CTL + b;
  double* DPH = CTH + b;
  double* DPL = DPH + c;
  double* DPU = DPL + c;
  double* PW = DPU + c;
  for (uint64_t i = 0; i < a; i++) {
    DRH[i] = drh[i + 1];
    CRL[i] = (r[i] * r[i]) * (1.0 / dr[i]);
    CRH[i] = (r[i + 1] * r[i + 1]) * (1.0 / dr[i + 1]);
  }
  for (uint64_t j = 0; j < b; j++) {
    SD[j] = sth[j + 1] * dth[j + 1];
    CP[j] = dth[j + 1] * (1.0 / sth[j + 1]);
    CTL[j] = st[j] * (1.0 / dt[j]);
    CTH[j] = st[j + 1] * (1.0 / dt[j + 1]);
  }
  for (uint64_t k = 0; k < c; k++) {
    DPH[k] = dph[k + 1];
    DPL[k] = 1.0 / dp[k];
    DPU[k] = 1.0 / dp[k + 1];
    PW[k] = dph[k + 1] * pl_i;
  }
  if ((rc = dalloc((void**)&s->tables, T.size() * sizeof(double)))) return bail(rc);
  if (hipMemcpy(s->tables, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice) !=
      hipSuccess)
    return bail(fail(P3D_EHIP, "copy of mesh tables failed"));
  double* dT = s->tables;
  s->m.DRH = dT;
  s->m.CRL = dT + a;
  s->m.CRH = dT + 2 * a;
  s->m.SD = dT + 3 * a;
  s->m.CP = dT + 3 * a + b;
  s->m.CTL = dT + 3 * a + 2 * b;
  s->m.CTH = dT + 3 * a + 3 * b;
  s->m.DPH = dT + 3 * a + 4 * b;
  s->m.DPL = dT + 3 * a + 4 * b + c;
  s->m.DPU = dT + 3 * a + 4 * b + 2 * c;
  s->m.PW = dT + 3 * a + 4 * b + 3 * c;
  {
    // float copies of the per-plane factors for the ILU sweeps' couplings
    std::vector<float> F(3 * c);
    for (uint64_t k = 0; k < c; k++) {
      F[k] = (float)DPH[k];
      F[c + k] = (float)DPL[k];
      F[2 * c + k] = (float)DPU[k];
    }
    if ((rc = dalloc((void**)&s->ftables, F.size() * sizeof(float)))) return bail(rc);
    if (hipMemcpy(s->ftables, F.data(), F.size() * sizeof(float), hipMemcpyHostToDevice) !=
        hipSuccess)
      return bail(fail(P3D_EHIP, "copy of mesh tables failed"));
    s->m.FDPH = s->ftables;
    s->m.FDPL = s->ftables + c;
    s->m.FDPU = s->ftables + 2 * c;
  }
  s->m.a = (int)a;
  s->m.b = (int)b;
  s->m.c = (int)c;
  s->m.pm_r1 = pm_r1;
  // full mesh arrays for phi / energy
  std::vector<double> M;
  size_t off[8];
  auto put = [&](int idx, const double* src, int n) {
    off[idx] = M.size();
    M.insert(M.end(), src, src + n);
  };
  put(0, dr, nr - 1);
  put(1, rh, nr);
  put(2, drh, nr);
  put(3, dt, nt - 1);
  put(4, dth, nt);
  put(5, sth, nt);
  put(6, dp, np);
  put(7, dph, np);
  if ((rc = dalloc((void**)&s->mesh, M.size() * sizeof(double)))) return bail(rc);
  hipMemcpy(s->mesh, M.data(), M.size() * sizeof(double), hipMemcpyHostToDevice);
  s->em.dr = s->mesh + off[0];
  s->em.rh = s->mesh + off[1];
  s->em.drh = s->mesh + off[2];
  s->em.dt = s->mesh + off[3];
  s->em.dth = s->mesh + off[4];
  s->em.sth = s->mesh + off[5];
  s->em.dp = s->mesh + off[6];
  s->em.dph = s->mesh + off[7];
  // vectors
  const size_t nb = (size_t)N * sizeof(double);
  // r and z also hold the tile-major placement: 8x8 columns x 64 planes
  {
    const size_t tiled = (size_t)((a + 7) / 8 * 8) * (size_t)((b + 7) / 8 * 8) *
                         (size_t)((c + 63) / 64 * 64);
    s->nvec = tiled > (size_t)N ? tiled : (size_t)N;
  }
  const size_t nbv = s->nvec * sizeof(double);
  if ((rc = dalloc((void**)&s->x, nb)) || (rc = dalloc((void**)&s->r, nbv)) ||
      (rc = dalloc((void**)&s->p, nb)) || (rc = dalloc((void**)&s->ap, nb)) ||
      (rc = dalloc((void**)&s->z, nbv)) || (rc = dalloc((void**)&s->b, nb)))
    return bail(rc);
  hipMemset(s->r, 0, nbv);
  hipMemset(s->z, 0, nbv);
  if ((rc = dalloc((void**)&s->br0, sizeof(double) * nt * np))) return bail(rc);
  if ((rc = dalloc((void**)&s->pole_part, sizeof(double) * 2 * kPoleSegs * a))) return bail(rc);
  if ((rc = dalloc((void**)&s->st, sizeof(CgState)))) return bail(rc);
  if (hipHostMalloc((void**)&s->st_host, sizeof(CgState), 0) != hipSuccess)
    return bail(fail(P3D_ENOMEM, "hipHostMalloc failed"));
  if (hipHostMalloc((void**)&s->out_host, 8 * sizeof(double), 0) != hipSuccess)
    return bail(fail(P3D_ENOMEM, "hipHostMalloc failed"));
  if ((rc = dalloc((void**)&s->out, 8 * sizeof(double)))) return bail(rc);
  {
    // the operator's launch with a natural or a tile-major (8x8 columns) output
    TileMap tsz;
    tsz.on = 1;
    tsz.TNI = (a + 7) / 8;
    tsz.TNJ = (b + 7) / 8;
    const int sb = std::max(stencil_blocks(s->g), stencil_blocks(s->g, tsz));
    s->red.max_blocks = s->L.G > sb ? s->L.G : sb;
  }
  if (s->red.max_blocks > kRedGroup * (kRedCounters - 1))
    return bail(fail(P3D_EINVAL, "grid too large for the reduction counters"));
  if ((rc = dalloc((void**)&s->red.partial,
                   sizeof(double) * 3 * (s->red.max_blocks + kRedCounters))))
    return bail(rc);
  if ((rc = dalloc((void**)&s->red.counter, sizeof(unsigned) * kRedSlots * kRedCounters)))
    return bail(rc);
  hipMemset(s->red.counter, 0, sizeof(unsigned) * kRedSlots * kRedCounters);
  hipMemset(s->x, 0, nb);
  hipMemset(s->b, 0, nb);
  hipMemset(s->pole_part, 0, sizeof(double) * 2 * kPoleSegs * a);
  if (ifprec == 2) {
    auto t0 = std::chrono::steady_clock::now();
    if ((rc = dalloc((void**)&s->yv, N * sizeof(float))) ||
        (rc = dalloc((void**)&s->dinv, N * sizeof(float))))
      return bail(rc);
    s->blocks = IluBlocks{s->g.a, s->g.b, s->g.c, 1};
    if ((rc = factor_ilu(s))) return bail(rc);
    s->sweep_ready = (build_sweep_plan(s->sweep, s->g, ncu, s->tile_k, s->tile_waves,
                                       &s->blocks, s->dinv, s->tile_stage, &s->m) == 0);
    s->sweep.persist = block_sweep_on(s);
    if ((rc = update_tilemap(s))) return bail(rc);
    s->stats.t_setup =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  if (hipDeviceSynchronize() != hipSuccess) return bail(fail(P3D_EHIP, "set-up failed"));
  *out = s;
  return P3D_OK;
}

// Block-persistent register-tile sweeps (P3D_OPT_BLOCK_SWEEP; -1: all
// diagonals below 2^26 unknowns, where one launch per diagonal leaves the
// device idle at the launch boundaries; above, one launch per diagonal:
// there the blocks' uneven share of the CUs costs more than the launches,
// and the forward sweep that also updates the residual runs slower
// block-persistent, also for the small diagonals alone).
static int block_sweep_on(const p3d_solver* s) {
  if (s->blk_sweep >= 0) return s->blk_sweep;
  return s->g.N < (1u << 26) ? 1 : 0;
}

// Block extents of the block-Jacobi cut for nblk blocks per axis, rounded
// up to multiples of 16 (r), 32 (theta) and 64 (phi) points so that every
// tile shape of the sweeps is aligned with them.
static IluBlocks ilu_blocks(const p3d_solver* s) {
  auto ext = [](int n, int nb, int q) {
    if (nb <= 1) return n;
    int e = (n + nb - 1) / nb;
    e = (e + q - 1) / q * q;
    return e >= n ? n : e;
  };
  IluBlocks bl;
  bl.bi = ext(s->g.a, s->nblk[0], 16);
  bl.bj = ext(s->g.b, s->nblk[1], 32);
  bl.bk = ext(s->g.c, s->nblk[2], 64);
  bl.wrap = bl.bk >= s->g.c ? 1 : 0;
  return bl;
}

// (Re)builds the tile plan and, with refactor, the D-ILU pivots for the
// current tile shape and block-Jacobi cut.
static int rebuild_ilu(p3d_solver* s, bool refactor) {
  HIPCHK(hipStreamSynchronize(s->stream));
  for (int t = 0; t < 3; t++) {
    if (s->gexec[t]) hipGraphExecDestroy(s->gexec[t]);
    s->gexec[t] = nullptr;
  }
  if (refactor) {
    s->blocks = ilu_blocks(s);
    const int rc = factor_ilu(s);
    if (rc) return rc;
  }
  int dev = 0, ncu = 0;
  HIPCHK(hipGetDevice(&dev));
  HIPCHK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int prc =
      build_sweep_plan(s->sweep, s->g, ncu, s->tile_k, s->tile_waves, &s->blocks, s->dinv,
                       s->tile_stage, &s->m);
  s->sweep_ready = prc == 0;
  s->sweep.persist = block_sweep_on(s);
  int rc = update_tilemap(s);
  if (prc == 2) return fail(P3D_EINVAL, "ILU blocks not aligned with the tiles");
  if (prc == 3)
    return fail(P3D_EINVAL, "P3D_OPT_TILE_STAGE 2 needs one wave per tile and 16 or 32 planes");
  if (!s->sweep_ready) return fail(P3D_ENOMEM, "sweep plan allocation failed");
  return rc;
}

int p3d_set_option(p3d_solver* s, int key, double value) {
  if (!s) return fail(P3D_EINVAL, "null solver");
  switch (key) {
    case P3D_OPT_STOP:
      if (value != 0 && value != 1) return fail(P3D_EINVAL, "P3D_OPT_STOP must be 0 or 1");
      s->stop_mode = (int)value;
      break;
    case P3D_OPT_MIXED:
      if (value != 0 && value != 1 && value != 2)
        return fail(P3D_EINVAL, "P3D_OPT_MIXED must be 0, 1 or 2");
      s->mixed = (int)value;
      break;
    case P3D_OPT_INNER_TOL:
      if (!(value > 0 && value < 1)) return fail(P3D_EINVAL, "inner tolerance must be in (0,1)");
      s->inner_tol = value;
      break;
    case P3D_OPT_CHUNK:
      s->chunk = (int)value;
      break;
    case P3D_OPT_GRAPH:
      s->use_graph = value != 0;
      break;
    case P3D_OPT_SWEEP:
      if (value != 0 && value != 1)
        return fail(P3D_EINVAL, "P3D_OPT_SWEEP must be 0 or 1");
      s->sweep_mode = (int)value;
      for (int t = 0; t < 3; t++) {
        if (s->gexec[t]) hipGraphExecDestroy(s->gexec[t]);
        s->gexec[t] = nullptr;
      }
      return update_tilemap(s);
    case P3D_OPT_VERBOSE:
      s->verbose = value != 0;
      break;
    case P3D_OPT_STENCIL: {
      const int v = (int)value;
      if (v < -1 || v > 4)
        return fail(P3D_EINVAL, "P3D_OPT_STENCIL must be -1 .. 4");
      if (v != s->L.stencil) {
        // captured iterations hold the other kernel
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int q = 0; q < 3; q++) {
          if (s->gexec[q]) hipGraphExecDestroy(s->gexec[q]);
          s->gexec[q] = nullptr;
        }
        s->L.stencil = v;
      }
      break;
    }
    case P3D_OPT_APPLY_FP32:
      s->apply_fp32 = value != 0;
      break;
    case P3D_OPT_FUSE_RESIDUAL: {
      const int v = (int)value;
      if (v < -1 || v > 1) return fail(P3D_EINVAL, "P3D_OPT_FUSE_RESIDUAL must be -1, 0 or 1");
      if (v != s->fuse_rupd) {
        // captured iterations hold the other kernel sequence
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int q = 0; q < 3; q++) {
          if (s->gexec[q]) hipGraphExecDestroy(s->gexec[q]);
          s->gexec[q] = nullptr;
        }
        s->fuse_rupd = v;
      }
      break;
    }
    case P3D_OPT_BLOCK_SWEEP: {
      const int v = (int)value;
      if (v < -1 || v > 2) return fail(P3D_EINVAL, "P3D_OPT_BLOCK_SWEEP must be -1, 0, 1 or 2");
      if (v != s->blk_sweep) {
        // captured iterations hold the other launch sequence
        HIPCHK(hipStreamSynchronize(s->stream));
        for (int q = 0; q < 3; q++) {
          if (s->gexec[q]) hipGraphExecDestroy(s->gexec[q]);
          s->gexec[q] = nullptr;
        }
        s->blk_sweep = v;
        s->sweep.persist = block_sweep_on(s);
      }
      break;
    }
    case P3D_OPT_TILE_K:
    case P3D_OPT_TILE_WAVES:
    case P3D_OPT_TILE_STAGE: {
      const int v = (int)value;
      if (key == P3D_OPT_TILE_K && v != 16 && v != 32 && v != 64)
        return fail(P3D_EINVAL, "P3D_OPT_TILE_K must be 16, 32 or 64");
      if (key == P3D_OPT_TILE_WAVES && v != 1 && v != 4 && v != 8)
        return fail(P3D_EINVAL, "P3D_OPT_TILE_WAVES must be 1, 4 or 8");
      if (key == P3D_OPT_TILE_STAGE && v != 0 && v != 1 && v != 2)
        return fail(P3D_EINVAL, "P3D_OPT_TILE_STAGE must be 0, 1 or 2");
      if (key == P3D_OPT_TILE_K) s->tile_k = v;
      else if (key == P3D_OPT_TILE_WAVES) s->tile_waves = v;
      else s->tile_stage = v;
      if (s->ifprec != PREC_ILU) break;
      return rebuild_ilu(s, false);
    }
    case P3D_OPT_ILU_BLOCKS_R:
    case P3D_OPT_ILU_BLOCKS_T:
    case P3D_OPT_ILU_BLOCKS_P: {
      const int v = (int)value;
      if (v < 1 || v > 1024) return fail(P3D_EINVAL, "block counts must be in 1..1024");
      s->nblk[key - P3D_OPT_ILU_BLOCKS_R] = v;
      if (s->ifprec != PREC_ILU) break;
      return rebuild_ilu(s, true);
    }
    case P3D_OPT_GUESS_CHECK:
      if (value != 0 && value != 1 && value != 2)
        return fail(P3D_EINVAL, "P3D_OPT_GUESS_CHECK must be 0, 1 or 2");
      s->guess_check = (int)value;
      break;
    default:
      return fail(P3D_EINVAL, "unknown option %d", key);
  }
  return P3D_OK;
}

int p3d_set_br0(p3d_solver* s, const double* br0) {
  if (!s || !br0) return fail(P3D_EINVAL, "null argument");
  HIPCHK(hipMemcpy(s->br0, br0, sizeof(double) * s->nt * s->np, hipMemcpyHostToDevice));
  s->have_br0 = true;
  return P3D_OK;
}

int p3d_set_br1(p3d_solver* s, const double* br1) {
  if (!s) return fail(P3D_EINVAL, "null solver");
  if (!br1) {
    s->have_br1 = false;
    return P3D_OK;
  }
  if (!s->br1) {
    int rc = dalloc((void**)&s->br1, sizeof(double) * s->nt * s->np);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(s->br1, br1, sizeof(double) * s->nt * s->np, hipMemcpyHostToDevice));
  s->have_br1 = true;
  return P3D_OK;
}

static int make_rhs(p3d_solver* s, double* dst) {
  if (!s->have_br0) return fail(P3D_ESTATE, "p3d_set_br0 must be called first");
  launch_rhs(s->m, s->g, s->br0, s->have_br1 ? s->br1 : nullptr, s->dr1, s->drn, s->nt, dst,
             s->stream);
  HIPCHK(hipGetLastError());
  return P3D_OK;
}

int p3d_rhs(p3d_solver* s, double* rhs) {
  if (!s || !rhs) return fail(P3D_EINVAL, "null argument");
  int rc = make_rhs(s, s->b);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(rhs, s->b, sizeof(double) * s->g.N, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

int p3d_solve_dev(p3d_solver* s, double* x_dev, double* rhs_dev, double epscg, int ncgmax,
                  int* ncg, double* epsn) {
  if (!s || !x_dev) return fail(P3D_EINVAL, "null argument");
  if (!(epscg > 0)) return fail(P3D_EINVAL, "epscg must be positive");
  const size_t nb = sizeof(double) * s->g.N;
  int rc;
  if (rhs_dev) {
    HIPCHK(hipMemcpyAsync(s->b, rhs_dev, nb, hipMemcpyDeviceToDevice, s->stream));
  } else if ((rc = make_rhs(s, s->b))) {
    return rc;
  }
  HIPCHK(hipMemcpyAsync(s->x, x_dev, nb, hipMemcpyDeviceToDevice, s->stream));
  rc = solve_core(s, epscg, ncgmax, 0);
  HIPCHK(hipMemcpyAsync(x_dev, s->x, nb, hipMemcpyDeviceToDevice, s->stream));
  if (rhs_dev) {
    // the final residual: FP64 recurrence r (natural order via ap), or the
    // true residual the mixed solve left in ap
    if (!s->mixed) launch_copy<double>(s->g, s->L, s->ap, s->r, s->stream, s->tm, 0);
    HIPCHK(hipMemcpyAsync(rhs_dev, s->ap, nb, hipMemcpyDeviceToDevice, s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  if (ncg) *ncg = s->stats.ncg;
  if (epsn) *epsn = s->stats.epsn;
  return rc;
}

int p3d_solve(p3d_solver* s, double* x, double* rhs, double epscg, int ncgmax, int ncghist,
              int* ncg, double* epsn) {
  if (!s || !x) return fail(P3D_EINVAL, "null argument");
  if (!(epscg > 0)) return fail(P3D_EINVAL, "epscg must be positive");
  const size_t nb = sizeof(double) * s->g.N;
  int rc;
  if (rhs) {
    HIPCHK(hipMemcpyAsync(s->b, rhs, nb, hipMemcpyHostToDevice, s->stream));
  } else if ((rc = make_rhs(s, s->b))) {
    return rc;
  }
  HIPCHK(hipMemcpyAsync(s->x, x, nb, hipMemcpyHostToDevice, s->stream));
  rc = solve_core(s, epscg, ncgmax, ncghist);
  HIPCHK(hipMemcpyAsync(x, s->x, nb, hipMemcpyDeviceToHost, s->stream));
  if (rhs) {
    if (!s->mixed) launch_copy<double>(s->g, s->L, s->ap, s->r, s->stream, s->tm, 0);
    HIPCHK(hipMemcpyAsync(rhs, s->ap, nb, hipMemcpyDeviceToHost, s->stream));
  }
  HIPCHK(hipStreamSynchronize(s->stream));
  if (ncg) *ncg = s->stats.ncg;
  if (epsn) *epsn = s->stats.epsn;
  return rc;
}

int p3d_get_stats(p3d_solver* s, p3d_stats* st) {
  if (!s || !st) return fail(P3D_EINVAL, "null argument");
  *st = s->stats;
  return P3D_OK;
}

int p3d_history(p3d_solver* s, double* hist, int n) {
  if (!s || !hist) return fail(P3D_EINVAL, "null argument");
  int m = s->last_hist;
  if (m > s->hist_cap) m = s->hist_cap;
  if (m > n) m = n;
  if (m > 0) HIPCHK(hipMemcpy(hist, s->hist, sizeof(double) * m, hipMemcpyDeviceToHost));
  return m;
}

static int ensure_phi(p3d_solver* s) {
  if (s->phi) return P3D_OK;
  int rc;
  if ((rc = dalloc((void**)&s->phi, sizeof(double) * s->nr * s->nt * s->np))) return rc;
  if ((rc = dalloc((void**)&s->sums, sizeof(double) * 2 * s->nr))) return rc;
  return P3D_OK;
}

int p3d_fill_phi(p3d_solver* s, const double* x, double* phi) {
  if (!s) return fail(P3D_EINVAL, "null solver");
  if (!s->have_br0) return fail(P3D_ESTATE, "p3d_set_br0 must be called first");
  int rc;
  if ((rc = ensure_phi(s))) return rc;
  const double* src = s->x;
  if (x) {
    HIPCHK(hipMemcpyAsync(s->z, x, sizeof(double) * s->g.N, hipMemcpyHostToDevice, s->stream));
    src = s->z;
  }
  launch_fill_phi(s->m, s->g, s->nr, s->nt, s->np, src, s->br0, s->have_br1 ? s->br1 : nullptr,
                  s->dr1, s->drn, s->em.dph, s->phi, s->sums, s->stream);
  HIPCHK(hipGetLastError());
  if (phi)
    HIPCHK(hipMemcpyAsync(phi, s->phi, sizeof(double) * s->nr * s->nt * s->np,
                          hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

int p3d_energy(p3d_solver* s, const double* phi, double* w, double* br, double* bt, double* bp) {
  if (!s || !w) return fail(P3D_EINVAL, "null argument");
  int rc;
  if ((rc = ensure_phi(s))) return rc;
  const int nr = s->nr, nt = s->nt, np = s->np;
  if (phi)
    HIPCHK(hipMemcpyAsync(s->phi, phi, sizeof(double) * nr * nt * np, hipMemcpyHostToDevice,
                          s->stream));
  if (!s->e_partial) {
    s->eG = s->L.G;
    if ((rc = dalloc((void**)&s->e_partial, sizeof(double) * 3 * (s->eG + kRedCounters))))
      return rc;
  }
  launch_energy(s->em, nr, nt, np, s->phi, s->e_partial, s->red.counter + 8 * kRedCounters,
                s->out, s->eG,
                s->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(s->out_host, s->out, 3 * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  w[0] = s->out_host[0];
  w[1] = s->out_host[1];
  w[2] = s->out_host[2];
  if (br || bt || bp) {
    double *dbr, *dbt, *dbp;
    const size_t nbr = (size_t)(nr - 1) * nt * np, nbt = (size_t)nr * (nt - 1) * np,
                 nbp = (size_t)nr * nt * (np - 1);
    if ((rc = dalloc((void**)&dbr, sizeof(double) * (nbr + nbt + nbp)))) return rc;
    dbt = dbr + nbr;
    dbp = dbt + nbt;
    launch_getb(s->em, nr, nt, np, s->phi, dbr, dbt, dbp, s->stream);
    if (br) hipMemcpyAsync(br, dbr, sizeof(double) * nbr, hipMemcpyDeviceToHost, s->stream);
    if (bt) hipMemcpyAsync(bt, dbt, sizeof(double) * nbt, hipMemcpyDeviceToHost, s->stream);
    if (bp) hipMemcpyAsync(bp, dbp, sizeof(double) * nbp, hipMemcpyDeviceToHost, s->stream);
    hipError_t e = hipStreamSynchronize(s->stream);
    hipFree(dbr);
    if (e != hipSuccess) return fail(P3D_EHIP, "getb failed: %s", hipGetErrorString(e));
  }
  return P3D_OK;
}

int p3d_apply_A(p3d_solver* s, const double* x, double* y) {
  if (!s || !x || !y) return fail(P3D_EINVAL, "null argument");
  const size_t nb = sizeof(double) * s->g.N;
  HIPCHK(hipMemcpyAsync(s->p, x, nb, hipMemcpyHostToDevice, s->stream));
  if (s->apply_fp32) {
    // the FP32 operator of the mixed recurrence, natural-order output
    float *xf = nullptr, *yf = nullptr;
    int rc;
    if ((rc = dalloc((void**)&xf, s->g.N * sizeof(float))) ||
        (rc = dalloc((void**)&yf, s->g.N * sizeof(float)))) {
      hipFree(xf);
      return rc;
    }
    const int nblk = (int)((s->g.N + kBlock - 1) / kBlock);
    k_convert<double, float><<<nblk, kBlock, 0, s->stream>>>(s->g.N, s->p, xf);
    launch_pole_sums<float>(s->m, s->g, xf, s->pole_part, s->stream);
    launch_stencil<float>(s->m, s->g, s->L, xf, yf, nullptr, 0, s->pole_part, nullptr, s->red, 0,
                          EPI_NONE, nullptr, false, s->stream);
    k_convert<float, double><<<nblk, kBlock, 0, s->stream>>>(s->g.N, yf, s->ap);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    hipFree(xf);
    hipFree(yf);
    if (e != hipSuccess) return fail(P3D_EHIP, "FP32 operator: %s", hipGetErrorString(e));
  } else {
    launch_pole_sums<double>(s->m, s->g, s->p, s->pole_part, s->stream);
    launch_stencil<double>(s->m, s->g, s->L, s->p, s->ap, nullptr, 0, s->pole_part, nullptr,
                           s->red, 0, EPI_NONE, nullptr, false, s->stream);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, s->ap, nb, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

int p3d_apply_prec(p3d_solver* s, double* x) {
  if (!s || !x) return fail(P3D_EINVAL, "null argument");
  const size_t nb = sizeof(double) * s->g.N;
  // x -> r and z -> x through the natural-order staging vector ap
  HIPCHK(hipMemcpyAsync(s->ap, x, nb, hipMemcpyHostToDevice, s->stream));
  launch_copy<double>(s->g, s->L, s->r, s->ap, s->stream, s->tm, 1);
  prec_and_dot<double>(s, s->r, s->z, EPI_NONE, nullptr, false);
  launch_copy<double>(s->g, s->L, s->ap, s->z, s->stream, s->tm, 0);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(x, s->ap, nb, hipMemcpyDeviceToHost, s->stream));
  HIPCHK(hipStreamSynchronize(s->stream));
  return P3D_OK;
}

int p3d_ilu_dinv(p3d_solver* s, float* dinv) {
  if (!s || !dinv) return fail(P3D_EINVAL, "null argument");
  if (s->ifprec != 2) return fail(P3D_ESTATE, "solver was not created with ifprec=2");
  HIPCHK(hipMemcpy(dinv, s->dinv, sizeof(float) * s->g.N, hipMemcpyDeviceToHost));
  return P3D_OK;
}

int p3d_time_kernel(p3d_solver* s, int which, int reps, double* avg_ms) {
  if (!s || !avg_ms || reps <= 0) return fail(P3D_EINVAL, "bad argument");
  if (which == 2 && s->ifprec != 2) return fail(P3D_ESTATE, "no ILU preconditioner");
  if ((which == 7 || which == 11) && s->ifprec != 2)
    return fail(P3D_ESTATE, "no ILU preconditioner");
  if ((which == 8 || which == 9 || which == 10 || which == 11) && !s->tm.on)
    return fail(P3D_ESTATE, "no tile-major placement (single-wave ILU tile sweeps)");
  if (which == 11 && !(s->sweep_mode == 0 && s->sweep_ready && dilu_tiles_fuse_rupd(s->sweep)))
    return fail(P3D_ESTATE, "no register-tile ILU sweeps (P3D_OPT_TILE_STAGE 2)");
  if (which == 1 || which == 3 || which >= 7) {
    int rc = ensure_mixed(s);
    if (!rc && which == 8) rc = ensure_pbuf(s);
    if (rc) return rc;
  }
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // a benign state: done=0 so guarded kernels run; alpha=beta=0 keeps x,r,p fixed
  CgState init;
  memset(&init, 0, sizeof init);
  init.bdotb = 1.0;
  init.bb = 1.0;
  init.epscg = 0.0;
  init.ncgmax = 1 << 30;
  init.ncg = which == 8 ? kXBatch : 1;
  *s->st_host = init;
  s->out_host[4] = 1.0;
  HIPCHK(hipMemcpyAsync(s->out + 4, s->out_host + 4, sizeof(double), hipMemcpyHostToDevice,
                        s->stream));
  HIPCHK(hipMemcpyAsync(s->st, s->st_host, sizeof(CgState), hipMemcpyHostToDevice, s->stream));
  Vecs<double> v{s->x, s->r, s->p, s->ap, s->z};
  if (which == 3) {
    // one genuine PCG run (r0 = p0 = 1, x0 = 0, never converging) replayed
    // through the captured graph: milliseconds per iteration
    launch_fill(s->x, 0.0, (size_t)s->g.N, s->stream);
    launch_fill(s->r, 1.0, (size_t)s->g.N, s->stream);
    launch_fill(s->p, 1.0, (size_t)s->g.N, s->stream);
    s->st_host->rdotr = 1.0;
    HIPCHK(hipMemcpyAsync(s->st, s->st_host, sizeof(CgState), hipMemcpyHostToDevice,
                          s->stream));
    const int K = chunk_size<double>(s);
    int rc = run_iterations<double>(s, v, K);  // warm-up (and graph capture)
    if (rc) return rc;
    HIPCHK(hipEventRecord(e0, s->stream));
    for (int i = 0; i < reps && !rc; i++) rc = run_iterations<double>(s, v, K);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e1, s->stream));
    HIPCHK(hipEventSynchronize(e1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, e0, e1));
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    if ((rc = read_state(s))) return rc;
    if (s->st_host->done) return fail(P3D_ESTATE, "timing run stopped (state %d)", s->st_host->done);
    *avg_ms = (double)ms / ((double)reps * K);
    return P3D_OK;
  }
  auto body = [&]() {
    switch (which) {
      case 0:
        launch_stencil<double>(s->m, s->g, s->L, s->p, s->ap, nullptr, 0, s->pole_part, nullptr,
                               s->red, 0, EPI_NONE, nullptr, false, s->stream);
        break;
      case 1:  // as in the FP32 recurrence: ap placed by the tile map
        launch_stencil<float>(s->m, s->g, s->L, s->p32, s->ap32, nullptr, 0, s->pole_part,
                              nullptr, s->red, 0, EPI_NONE, nullptr, false, s->stream, s->tm);
        break;
      case 2:
        prec_and_dot<double>(s, s->r, s->z, EPI_NONE, nullptr, false);
        break;
      case 4:
        launch_cg_update<double>(s->m, s->g, s->L, s->ifprec, s->x, s->r, s->p, s->ap, s->st,
                                 s->red, 2, s->stream, nullptr, nullptr, s->tm);
        break;
      case 5:
        launch_p_update<double>(s->m, s->g, s->L, s->ifprec, s->p, s->r, s->z, s->st,
                                s->stream, s->tm);
        break;
      case 6:  // plain copy of N doubles: the bandwidth reference
        launch_copy<double>(s->g, s->L, s->ap, s->p, s->stream);
        break;
      case 7:  // FP32 ILU apply (the mixed-precision inner preconditioner)
        if (s->ifprec == PREC_ILU) prec_and_dot<float>(s, s->r32, s->z32, EPI_NONE, nullptr, false);
        break;
      case 8: {  // one group of the FP32 recurrence's p updates (kXBatch launches, the
        // last with the FP64 solution update of the whole group; x_ncg cleared
        // so that it applies kXBatch pending updates; alpha = 0 leaves x);
        // timed per iteration
        PBuf pb{};
        pb.p[0] = s->p32;
        for (int q = 1; q < kXBatch; q++) pb.p[q] = s->p32b[q - 1];
        (void)hipMemsetAsync(&s->st->x_ncg, 0, sizeof(int), s->stream);  // checked below
        for (int t = 0; t < kXBatch; t++)
          launch_p_update_x(s->g, s->L, t, pb, s->z32, s->st, s->stream, s->tm, s->x,
                            s->out + 4, s->red.counter + 5 * kRedCounters);
        break;
      }
      case 9:  // residual update of the FP32 recurrence (tile-major r, ap)
        launch_r_update(s->gt.N, s->Lt, s->r32, s->ap32, s->st, s->stream);
        break;
      case 10:  // r.z and r.r of the FP32 recurrence, contiguous in tile order
        launch_dot2<float>(s->gt, s->Lt, s->r32, s->z32, s->r32, s->r32, s->st, s->red, 1,
                           EPI_NONE, nullptr, false, s->stream);
        break;
      case 11:  // FP32 ILU apply with the residual update in its forward sweep
        launch_dilu_tiles_rupd(s->sweep, s->m, s->g, s->r32, s->ap32, s->z32, s->st, s->stream);
        break;
      default:
        break;
    }
  };
  body();  // warm-up
  HIPCHK(hipEventRecord(e0, s->stream));
  for (int i = 0; i < reps; i++) body();
  HIPCHK(hipEventRecord(e1, s->stream));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *avg_ms = (double)ms / reps / (which == 8 ? kXBatch : 1);
  if (which > 11 || which < 0) return fail(P3D_EINVAL, "unknown kernel id %d", which);
  return P3D_OK;
}

}  // extern "C"
