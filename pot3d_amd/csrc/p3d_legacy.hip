// Legacy external-LU-solver ABI (include/pot3d_hip.h, group 1): the three
// symbols POT3D binds under -DHIPSPARSE (src/pot3d.F90:581-609), replacing
// src/lusol_hipsparse.c.
//
// The caller's CSR is the diacsr matrix (src/pot3d.F90:4870-5018): the
// 7-point stencil with r/theta boundary couplings dropped and, at one rank
// in phi, the periodic wrap kept.  The structure is recognised from IA/JA
// (grid a x b x c, wrap on/off); the couplings are split into stored
// planes and factorised as D-ILU (FP64 pivots), which equals ILU(0) on this
// pattern.  Anything else is rejected with exit(1), as the reference shim
// does on any sparse-library error.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include <hsa/hsa.h>
#include <hsa/hsa_ext_amd.h>

#include "../../include/pot3d_hip.h"
#include "p3d_kernels.hpp"
#include "p3d_sweep.hpp"

using namespace p3d;

// OpenMP mapping lookup, resolved only when the host program uses OpenMP
// offload (the -DHIPSPARSE convention passes host addresses of mapped
// arrays, src/lusol_hipsparse.c:75-79).
extern "C" void* omp_get_mapped_ptr(const void* ptr, int device_num) __attribute__((weak));
extern "C" int omp_get_default_device(void) __attribute__((weak));

namespace {

struct Legacy {
  bool loaded = false;
  int N = 0, a = 0, b = 0, c = 0;
  bool wrap = true;
  Grid3 g;
  hipStream_t stream = nullptr;
  float* planes = nullptr;  // 6*N couplings + 2*a*b wrap
  CoefArrays ca;
  float* dinv = nullptr;
  float* yv = nullptr;
  double* xbuf = nullptr;   // staging for host x
  SweepPlan plan;           // tile-diagonal schedule of the trisolves
  bool tiles = false;
};

Legacy G;

[[noreturn]] void die(const char* msg) {
  fprintf(stderr, " ERROR! external_lu_solver: %s\n", msg);
  fflush(stderr);
  exit(1);
}

void check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    fprintf(stderr, " ERROR! external_lu_solver: %s: %s\n", what, hipGetErrorString(e));
    fflush(stderr);
    exit(1);
  }
}

// Where a caller address lives.
//   Hip:  device (or managed) memory that HIP tracks: hipMemcpy and kernels.
//   Hsa:  an allocation of the HSA runtime that HIP does not track, such as
//         the device arrays of an OpenMP-offload program (libomptarget
//         allocates from the HSA memory pools; the -DCUSPARSE calling
//         convention passes them through use_device_addr, src/pot3d.F90:
//         4675-4683): reachable by kernels of this process only, so copies
//         go through a copy kernel.
//   Host: plain, pinned or registered host memory: memcpy.
enum Where { W_HOST = 0, W_HIP = 1, W_HSA = 2 };

Where classify(const void* p) {
  hipPointerAttribute_t at;
  if (hipPointerGetAttributes(&at, p) == hipSuccess) {
    if (at.type == hipMemoryTypeDevice || at.type == hipMemoryTypeManaged) return W_HIP;
    return W_HOST;
  }
  (void)hipGetLastError();
  hsa_amd_pointer_info_t info;
  memset(&info, 0, sizeof info);
  info.size = sizeof info;
  if (hsa_amd_pointer_info(p, &info, nullptr, nullptr, nullptr) == HSA_STATUS_SUCCESS &&
      (info.type == HSA_EXT_POINTER_TYPE_HSA || info.type == HSA_EXT_POINTER_TYPE_IPC ||
       info.type == HSA_EXT_POINTER_TYPE_HSA_VMEM))
    return W_HSA;
  return W_HOST;
}

struct View {
  void* p;
  Where w;
};

// Resolves a caller pointer: an OpenMP-mapped host address (the
// -DHIPSPARSE convention, src/lusol_hipsparse.c:75-79) -> its device
// address; anything else is classified as it is.
View view(void* p) {
  if (omp_get_mapped_ptr && omp_get_default_device) {
    void* d = omp_get_mapped_ptr(p, omp_get_default_device());
    if (d) return View{d, classify(d)};
  }
  return View{p, classify(p)};
}

__global__ void k_copy_words(const uint32_t* __restrict__ src, uint32_t* __restrict__ dst,
                             size_t n) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// Copies nbytes (a multiple of 4) between a caller buffer v and host memory
// h, in the direction to_host.
void move(View v, void* h, size_t nbytes, bool to_host) {
  if (v.w == W_HOST) {
    if (to_host)
      memcpy(h, v.p, nbytes);
    else
      memcpy(v.p, h, nbytes);
    return;
  }
  const hipMemcpyKind kind = to_host ? hipMemcpyDeviceToHost : hipMemcpyHostToDevice;
  if (v.w == W_HIP) {
    check(to_host ? hipMemcpy(h, v.p, nbytes, kind) : hipMemcpy(v.p, h, nbytes, kind),
          "copy CSR");
    return;
  }
  void* stage = nullptr;
  const size_t nw = nbytes / 4;
  const unsigned blocks = (unsigned)((nw + 255) / 256);
  check(hipMalloc(&stage, nbytes ? nbytes : 4), "allocate staging");
  if (to_host) {
    if (nw) k_copy_words<<<blocks, 256, 0, G.stream>>>((const uint32_t*)v.p, (uint32_t*)stage, nw);
    check(hipStreamSynchronize(G.stream), "copy from HSA memory");
    check(hipMemcpy(h, stage, nbytes, hipMemcpyDeviceToHost), "copy CSR");
  } else {
    check(hipMemcpy(stage, h, nbytes, hipMemcpyHostToDevice), "copy CSR");
    if (nw) k_copy_words<<<blocks, 256, 0, G.stream>>>((const uint32_t*)stage, (uint32_t*)v.p, nw);
    check(hipStreamSynchronize(G.stream), "copy to HSA memory");
  }
  hipFree(stage);
}

template <typename V>
void fetch(std::vector<V>& dst, const V* src, size_t n) {
  dst.resize(n);
  move(view((void*)src), dst.data(), n * sizeof(V), true);
}

// Expected columns (1-based, in diacsr order) of row p=(i,j,k).
int expected_row(int i, int j, int k, int a, int b, int c, bool wrap, long long* cols) {
  const long long N = (long long)a * b * c, ab = (long long)a * b;
  const long long p = (long long)i + (long long)a * j + ab * k + 1;
  const long long off[7] = {-ab, -a, -1, 0, 1, a, ab};
  bool ok[7] = {true, true, true, true, true, true, true};
  if (i == 0) ok[2] = false;
  if (i == a - 1) ok[4] = false;
  if (j == 0) ok[1] = false;
  if (j == b - 1) ok[5] = false;
  if (!wrap) {
    if (k == 0) ok[0] = false;
    if (k == c - 1) ok[6] = false;
  }
  int n = 0;
  for (int q = 0; q < 7; q++)
    if (ok[q] && p + off[q] > N) cols[n++] = p + off[q] - N;
  for (int q = 0; q < 7; q++)
    if (ok[q] && p + off[q] >= 1 && p + off[q] <= N) cols[n++] = p + off[q];
  for (int q = 0; q < 7; q++)
    if (ok[q] && p + off[q] < 1) cols[n++] = p + off[q] + N;
  return n;
}

bool match_pattern(const std::vector<int>& IA, const std::vector<int>& JA, int a, int b, int c,
                   bool wrap) {
  const long long N = (long long)a * b * c;
  if (IA[0] != 1) return false;
  long long cols[7];
  long long p = 0;
  for (int k = 0; k < c; k++)
    for (int j = 0; j < b; j++)
      for (int i = 0; i < a; i++, p++) {
        const int n = expected_row(i, j, k, a, b, c, wrap, cols);
        if (IA[p + 1] - IA[p] != n) return false;
        for (int q = 0; q < n; q++)
          if (JA[IA[p] - 1 + q] != cols[q]) return false;
      }
  return p == N;
}

void unload() {
  if (!G.loaded) return;
  if (G.stream) hipStreamSynchronize(G.stream);
  if (G.planes) hipFree(G.planes);
  if (G.dinv) hipFree(G.dinv);
  if (G.yv) hipFree(G.yv);
  if (G.xbuf) hipFree(G.xbuf);
  free_sweep_plan(G.plan);
  if (G.stream) hipStreamDestroy(G.stream);
  G = Legacy();
}

}  // namespace

extern "C" {

void load_external_lu_solver(float* csr_lu, int* csr_i, int* csr_j, int N, int M) {
  if (G.loaded) unload();
  if (N < 1 || M < 1) die("invalid matrix size");
  std::vector<int> IA, JA;
  std::vector<float> AV;
  fetch(IA, csr_i, (size_t)N + 1);
  const long long nnz = (long long)IA[N] - 1;
  if (nnz < N || nnz > (long long)M) die("row pointer array inconsistent with M");
  fetch(JA, csr_j, (size_t)nnz);
  fetch(AV, csr_lu, (size_t)nnz);
  // grid from row 1: off-diagonal columns 2, 1+a, 1+ab [, 1+(c-1)ab]
  std::vector<long long> offs;
  for (int e = IA[0] - 1; e < IA[1] - 1; e++)
    if (JA[e] != 1) offs.push_back(JA[e] - 1);
  bool found = false;
  int a = 0, b = 0, c = 0;
  bool wrap = true;
  if (offs.size() >= 3 && offs[0] == 1) {
    const long long A = offs[1], AB = offs[2];
    if (A >= 2 && AB % A == 0 && (long long)N % AB == 0) {
      a = (int)A;
      b = (int)(AB / A);
      c = (int)(N / AB);
      for (int w = 1; w >= 0 && !found; w--) {
        if (c >= 8 && match_pattern(IA, JA, a, b, c, w == 1)) {
          found = true;
          wrap = (w == 1);
        }
      }
    }
  }
  if (!found) die("matrix is not the POT3D 7-point pattern (unsupported by this solver)");
  G.N = N;
  G.a = a;
  G.b = b;
  G.c = c;
  G.wrap = wrap;
  G.g.a = a;
  G.g.b = b;
  G.g.c = c;
  G.g.N = (uint32_t)N;
  G.g.ab = (uint32_t)(a * b);
  G.g.da = make_fastdiv((uint32_t)a);
  G.g.dab = make_fastdiv((uint32_t)(a * b));
  // split into planes
  const size_t n = (size_t)N, ab = (size_t)a * b;
  std::vector<float> P(7 * n + 2 * ab, 0.0f);  // 6 couplings + diag + 2 wrap planes
  float *lo_k = P.data(), *lo_j = lo_k + n, *lo_i = lo_j + n, *up_i = lo_i + n,
        *up_j = up_i + n, *up_k = up_j + n, *dg = up_k + n, *w_lo = dg + n, *w_up = w_lo + ab;
  std::vector<int> role(nnz);  // 0..6 as planes, 7 wrap_lo, 8 wrap_up
  for (size_t p = 0; p < n; p++) {
    const int k = (int)(p / ab);
    const size_t ij = p - (size_t)k * ab;
    for (int e = IA[p] - 1; e < IA[p + 1] - 1; e++) {
      const long long q = (long long)JA[e] - 1;
      const long long d = q - (long long)p;
      const float v = AV[e];
      int rl;
      if (d == 0) {
        dg[p] = v;
        rl = 6;
      } else if (d == -1) {
        lo_i[p] = v;
        rl = 2;
      } else if (d == 1) {
        up_i[p] = v;
        rl = 3;
      } else if (d == -(long long)a) {
        lo_j[p] = v;
        rl = 1;
      } else if (d == (long long)a) {
        up_j[p] = v;
        rl = 4;
      } else if (k == c - 1 && q == (long long)ij) {
        w_lo[ij] = v;  // wrap to (i,j,0), lower (checked before the k-1 test)
        rl = 7;
      } else if (k == 0 && q == (long long)(ij + (size_t)(c - 1) * ab)) {
        w_up[ij] = v;
        rl = 8;
      } else if (d == -(long long)ab) {
        lo_k[p] = v;
        rl = 0;
      } else {
        up_k[p] = v;
        rl = 5;
      }
      role[e] = rl;
    }
  }
  check(hipStreamCreateWithFlags(&G.stream, hipStreamNonBlocking), "stream");
  check(hipMalloc((void**)&G.planes, P.size() * sizeof(float)), "allocate couplings");
  check(hipMemcpy(G.planes, P.data(), P.size() * sizeof(float), hipMemcpyHostToDevice),
        "copy couplings");
  float* dP = G.planes;
  G.ca.lo_k = dP;
  G.ca.lo_j = dP + n;
  G.ca.lo_i = dP + 2 * n;
  G.ca.up_i = dP + 3 * n;
  G.ca.up_j = dP + 4 * n;
  G.ca.up_k = dP + 5 * n;
  const float* d_dg = dP + 6 * n;
  G.ca.wrap_lo = dP + 7 * n;
  G.ca.wrap_up = dP + 7 * n + ab;
  check(hipMalloc((void**)&G.dinv, n * sizeof(float)), "allocate pivots");
  check(hipMalloc((void**)&G.yv, n * sizeof(float)), "allocate work");
  double* dpiv = nullptr;
  check(hipMalloc((void**)&dpiv, n * sizeof(double)), "allocate pivots");
  launch_dilu_factor_stored(G.ca, d_dg, G.g, wrap, dpiv, G.dinv, G.stream);
  check(hipStreamSynchronize(G.stream), "D-ILU factorisation");
  // zero-pivot check and in-place LU write-back (L: a_pq/d_q, U: a_pq,
  // diagonal: d_p), as csrilu02 leaves the caller's values.
  std::vector<double> d(n);
  check(hipMemcpy(d.data(), dpiv, n * sizeof(double), hipMemcpyDeviceToHost), "copy pivots");
  hipFree(dpiv);
  for (size_t p = 0; p < n; p++)
    if (d[p] == 0.0 || !isfinite(d[p])) {
      fprintf(stderr, " ERROR! ILU0 zero or non-finite pivot %g at row %zu\n", d[p], p + 1);
      fflush(stderr);
      exit(1);
    }
  for (size_t p = 0; p < n; p++) {
    for (int e = IA[p] - 1; e < IA[p + 1] - 1; e++) {
      const long long q = (long long)JA[e] - 1;
      if (role[e] == 6)
        AV[e] = (float)d[p];
      else if (q < (long long)p)
        AV[e] = (float)((double)AV[e] / d[q]);
    }
  }
  move(view(csr_lu), AV.data(), nnz * sizeof(float), false);
  int dev = 0, ncu = 256;
  hipDeviceProp_t prop;
  if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess)
    ncu = prop.multiProcessorCount;
  G.tiles = (build_sweep_plan(G.plan, G.g, ncu) == 0);
  G.loaded = true;
}

void external_lu_solver(double* x) {
  if (!G.loaded) die("external_lu_solver called before load_external_lu_solver");
  const View xv = view(x);
  double* dx = xv.w == W_HOST ? nullptr : (double*)xv.p;  // kernels work in place
  const size_t nb = (size_t)G.N * sizeof(double);
  if (!dx) {
    if (!G.xbuf) check(hipMalloc((void**)&G.xbuf, nb), "allocate x");
    check(hipMemcpyAsync(G.xbuf, x, nb, hipMemcpyHostToDevice, G.stream), "copy x");
    dx = G.xbuf;
  }
  if (G.tiles) {
    launch_dilu_tiles_stored(G.plan, G.ca, G.g, G.dinv, dx, G.yv, dx, G.stream);
  } else {
    Metric m;
    memset(&m, 0, sizeof m);
    launch_dilu_solve_levels<double>(m, G.g, G.dinv, dx, G.yv, dx, &G.ca, G.wrap,
                                     IluBlocks{G.g.a, G.g.b, G.g.c, G.wrap ? 1 : 0}, nullptr,
                                     G.stream);
  }
  check(hipGetLastError(), "trisolve launch");
  if (dx == G.xbuf) check(hipMemcpyAsync(x, G.xbuf, nb, hipMemcpyDeviceToHost, G.stream), "copy x");
  check(hipStreamSynchronize(G.stream), "trisolve");
}

void unload_external_lu_solver(void) { unload(); }

}  // extern "C"
