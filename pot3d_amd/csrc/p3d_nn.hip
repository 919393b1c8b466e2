// Warm-start network (C ABI p3d_nn_*): a 4-level 3-D convolutional
// encoder-decoder mapping the r0 boundary field Br (broadcast along r with
// the grid's radial factors) to the potential on the solver's interior
// grid.  Forward, backward and Adam run on the device with the kernels of
// p3d_nn_kernels.hip (BF16 matrix-core convolutions, FP32 accumulation).
//
//   level 0 (a x b x c)          e0 = relu(conv 16->16 x0)
//   level 1 (ceil/2)       p1 = pool e0,  e1 = relu(conv 16->32 p1)
//   level 2                p2 = pool e1,  e2 = relu(conv 32->64 p2)
//   level 3                p3 = pool e2,  e3 = relu(conv 64->128 p3)
//   level 2  u2 = up e3,   s2 = relu(conv 128->64 u2) + e2
//   level 1  u1 = up s2,   s1 = relu(conv 64->32 u1) + e1
//   level 0  u0 = up s1,   s0 = relu(conv 32->16 u0) + e0
//   head     out = conv 16->16 s0   (channel 0 = potential / phi_scale)
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/pot3d_hip.h"
#include "p3d_nn.hpp"

using namespace p3d::nn;

namespace p3d {
int set_error(int code, const char* msg);
}

static int nfail(int code, const char* msg) { return p3d::set_error(code, msg); }

// Staging for p3d_nn_prolong_dev_noalloc: the per-axis tables of one
// prolongation (weights wr|wt|wp as doubles, then indices ir|it|ip as ints),
// in a file-scope device array, so that entry point allocates nothing.
// Every GPU the library runs on holds its own copy (96 KiB) until the
// process ends.  Host code never uses the array's name as an address: the
// device address is looked up once per device (scratch_on_device) and that
// pointer is what the copies and the kernel receive.
constexpr int kProlongMaxTable = 8192;  // af + bf + cf at most
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kProlongMaxTable + kProlongMaxTable / 2];

static std::mutex g_scratch_mu;
static double* g_scratch_dev[kMaxDevices];

// Device address of `scratch` on the current device (looked up on the
// first call with that device current), or nullptr after an error.
static double* scratch_on_device() {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    nfail(P3D_EHIP, "no current HIP device (or device id beyond 63)");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
      nfail(P3D_EHIP, "hipGetSymbolAddress of the prolongation scratch failed");
      return nullptr;
    }
    g_scratch_dev[dev] = static_cast<double*>(ptr);
  }
  return g_scratch_dev[dev];
}

struct ConvSpec {
  int Cin, Cout, level;
  size_t w_off, b_off;  // offsets in the flat parameter buffer
};

struct p3d_nn {
  int a, b, c, nt;
  Dims3 D[4];
  hipStream_t stream = nullptr;
  std::vector<ConvSpec> conv;
  size_t nparam = 0;
  // parameters (FP32 master, BF16 copy of the weights), Adam state, grads
  float *w = nullptr, *g = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t* wb = nullptr;
  int step = 0;
  // activations (BF16)
  uint16_t *x0, *e0, *p1, *e1, *p2, *e2, *p3, *e3, *u2, *d2, *s2, *u1, *d1, *s1, *u0, *d0, *s0;
  float* out = nullptr;
  // gradients (FP32)
  float *g_out, *g_s0, *g_t0, *g_u0, *g_s1, *g_t1, *g_u1, *g_s2, *g_t2, *g_u2, *g_e3, *g_p3,
      *g_e2, *g_p2, *g_e1, *g_p1, *g_e0;
  float* part = nullptr;  // weight-gradient partials
  size_t part_n = 0;
  float* lpart = nullptr;  // loss partials
  float* loss = nullptr;   // device scalar
  // features
  float *decay = nullptr, *rpos = nullptr;
  double bscale = 1.0, phiscale = 1.0;
  double* br_dev = nullptr;  // one br0 (nt x np)
  // training set (device)
  std::vector<double*> ds_br;
  std::vector<double> ds_bs;  // per-sample input scale
  std::vector<float*> ds_t;
  float* tgt = nullptr;  // current target (FP32, /phiscale)
  // training augmentation by exact symmetries of the discrete problem:
  // rotation in phi (uniform phi mesh), reflection in theta (mesh symmetric
  // about the equator), sign (the solution is linear in br0)
  int aug_roll = 0, aug_flip = 0, aug_sign = 0;
  std::vector<void*> allocs;
};

static int dmalloc(p3d_nn* n, void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 4) != hipSuccess) return nfail(P3D_ENOMEM, "hipMalloc failed");
  n->allocs.push_back(*p);
  return P3D_OK;
}

static Dims3 half(const Dims3& d) { return Dims3{(d.X + 1) / 2, (d.Y + 1) / 2, (d.Z + 1) / 2}; }

// -- forward / backward -------------------------------------------------------
static void forward(p3d_nn* n) {
  hipStream_t s = n->stream;
  auto W = [&](int i) { return n->wb + n->conv[i].w_off; };
  auto B = [&](int i) { return n->w + n->conv[i].b_off; };
  const Dims3* D = n->D;
  conv_fwd(D[0], 16, 16, n->x0, W(0), B(0), n->e0, nullptr, true, s);
  pool_fwd(D[0], D[1], 16, n->e0, n->p1, s);
  conv_fwd(D[1], 16, 32, n->p1, W(1), B(1), n->e1, nullptr, true, s);
  pool_fwd(D[1], D[2], 32, n->e1, n->p2, s);
  conv_fwd(D[2], 32, 64, n->p2, W(2), B(2), n->e2, nullptr, true, s);
  pool_fwd(D[2], D[3], 64, n->e2, n->p3, s);
  conv_fwd(D[3], 64, 128, n->p3, W(3), B(3), n->e3, nullptr, true, s);
  up_fwd(D[3], D[2], 128, n->e3, n->u2, s);
  conv_fwd(D[2], 128, 64, n->u2, W(4), B(4), n->d2, nullptr, true, s);
  add_bf16(D[2].V() * 64, n->d2, n->e2, n->s2, s);
  up_fwd(D[2], D[1], 64, n->s2, n->u1, s);
  conv_fwd(D[1], 64, 32, n->u1, W(5), B(5), n->d1, nullptr, true, s);
  add_bf16(D[1].V() * 32, n->d1, n->e1, n->s1, s);
  up_fwd(D[1], D[0], 32, n->s1, n->u0, s);
  conv_fwd(D[0], 32, 16, n->u0, W(6), B(6), n->d0, nullptr, true, s);
  add_bf16(D[0].V() * 16, n->d0, n->e0, n->s0, s);
  conv_fwd(D[0], 16, 16, n->s0, W(7), B(7), nullptr, n->out, false, s);
}

static int nchunk_for(const Dims3& d) {
  int nc = (int)((d.V() + 2047) / 2048);
  return nc < 1 ? 1 : nc;
}

static void backward(p3d_nn* n) {
  hipStream_t s = n->stream;
  auto W = [&](int i) { return n->wb + n->conv[i].w_off; };
  auto GW = [&](int i) { return n->g + n->conv[i].w_off; };
  auto GB = [&](int i) { return n->g + n->conv[i].b_off; };
  const Dims3* D = n->D;
  const size_t V0 = D[0].V(), V1 = D[1].V(), V2 = D[2].V(), V3 = D[3].V();
  // head: out = conv(s0)
  conv_bwd(D[0], 16, 16, n->g_out, n->s0, W(7), n->g_s0, GW(7), GB(7), n->part,
           nchunk_for(D[0]), s);
  // s0 = relu(conv(u0)) + e0
  relu_bwd(V0 * 16, n->d0, n->g_s0, nullptr, n->g_t0, s);
  conv_bwd(D[0], 32, 16, n->g_t0, n->u0, W(6), n->g_u0, GW(6), GB(6), n->part, nchunk_for(D[0]),
           s);
  up_bwd(D[1], D[0], 32, n->g_u0, n->g_s1, s);
  // s1 = relu(conv(u1)) + e1
  relu_bwd(V1 * 32, n->d1, n->g_s1, nullptr, n->g_t1, s);
  conv_bwd(D[1], 64, 32, n->g_t1, n->u1, W(5), n->g_u1, GW(5), GB(5), n->part, nchunk_for(D[1]),
           s);
  up_bwd(D[2], D[1], 64, n->g_u1, n->g_s2, s);
  // s2 = relu(conv(u2)) + e2
  relu_bwd(V2 * 64, n->d2, n->g_s2, nullptr, n->g_t2, s);
  conv_bwd(D[2], 128, 64, n->g_t2, n->u2, W(4), n->g_u2, GW(4), GB(4), n->part,
           nchunk_for(D[2]), s);
  up_bwd(D[3], D[2], 128, n->g_u2, n->g_e3, s);
  // e3 = relu(conv(p3))
  relu_bwd(V3 * 128, n->e3, n->g_e3, nullptr, n->g_e3, s);
  conv_bwd(D[3], 64, 128, n->g_e3, n->p3, W(3), n->g_p3, GW(3), GB(3), n->part,
           nchunk_for(D[3]), s);
  pool_bwd(D[2], D[3], 64, n->g_p3, n->g_e2, s);
  // e2 feeds pool (g_e2) and the skip of s2 (g_s2)
  relu_bwd(V2 * 64, n->e2, n->g_e2, n->g_s2, n->g_e2, s);
  conv_bwd(D[2], 32, 64, n->g_e2, n->p2, W(2), n->g_p2, GW(2), GB(2), n->part, nchunk_for(D[2]),
           s);
  pool_bwd(D[1], D[2], 32, n->g_p2, n->g_e1, s);
  relu_bwd(V1 * 32, n->e1, n->g_e1, n->g_s1, n->g_e1, s);
  conv_bwd(D[1], 16, 32, n->g_e1, n->p1, W(1), n->g_p1, GW(1), GB(1), n->part, nchunk_for(D[1]),
           s);
  pool_bwd(D[0], D[1], 16, n->g_p1, n->g_e0, s);
  relu_bwd(V0 * 16, n->e0, n->g_e0, n->g_s0, n->g_e0, s);
  conv_bwd(D[0], 16, 16, n->g_e0, n->x0, W(0), nullptr, GW(0), GB(0), n->part, nchunk_for(D[0]),
           s);
}

static void make_features(p3d_nn* n, const double* br_dev, double bs, int roll = 0,
                          int flip = 0) {
  features(n->D[0], 16, n->nt, br_dev, n->decay, n->rpos, (float)(1.0 / bs), roll, flip, n->x0,
           n->stream);
}

// Input scale of one br0(nt,np) (host): the fixed bscale, or with bscale = 0
// the rms of br0 over the interior points (the solution is linear in br0,
// so every sample is presented to the network at unit amplitude).
static double sample_scale(const p3d_nn* n, const double* br0) {
  if (n->bscale > 0) return n->bscale;
  const int nt = n->nt, np = n->c + 2;
  double s = 0.0;
  for (int k = 1; k < np - 1; k++)
    for (int j = 1; j < nt - 1; j++) s += br0[j + (size_t)nt * k] * br0[j + (size_t)nt * k];
  s = sqrt(s / ((double)(nt - 2) * (np - 2)));
  return s > 0 ? s : 1.0;
}
static double target_scale(const p3d_nn* n, double bs) {
  return n->bscale > 0 ? n->phiscale : bs * n->phiscale;
}

extern "C" {

int p3d_nn_destroy(p3d_nn* n) {
  if (!n) return P3D_OK;
  if (n->stream) hipStreamSynchronize(n->stream);
  for (void* p : n->allocs) hipFree(p);
  for (double* p : n->ds_br) hipFree(p);
  for (float* p : n->ds_t) hipFree(p);
  if (n->stream) hipStreamDestroy(n->stream);
  delete n;
  return P3D_OK;
}

int p3d_nn_create(p3d_nn** out, int nr, int nt, int np, const double* rh, unsigned seed) {
  if (!out || !rh) return nfail(P3D_EINVAL, "null argument");
  if (nr < 4 || nt < 4 || np < 5) return nfail(P3D_EINVAL, "grid too small");
  p3d_nn* n = new p3d_nn();
  n->a = nr - 2;
  n->b = nt - 2;
  n->c = np - 2;
  n->nt = nt;
  n->D[0] = Dims3{n->a, n->b, n->c};
  for (int l = 1; l < 4; l++) n->D[l] = half(n->D[l - 1]);
  if (hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking) != hipSuccess) {
    delete n;
    return nfail(P3D_EHIP, "stream creation failed");
  }
  const int spec[8][3] = {{16, 16, 0},  {16, 32, 1}, {32, 64, 2}, {64, 128, 3},
                          {128, 64, 2}, {64, 32, 1}, {32, 16, 0}, {16, 16, 0}};
  size_t off = 0;
  for (auto& sp : spec) {
    ConvSpec cs{sp[0], sp[1], sp[2], off, 0};
    off += (size_t)27 * sp[0] * sp[1];
    cs.b_off = off;
    off += sp[1];
    n->conv.push_back(cs);
  }
  n->nparam = off;
  int rc;
  const size_t P = n->nparam;
  if ((rc = dmalloc(n, (void**)&n->w, P * 4)) || (rc = dmalloc(n, (void**)&n->g, P * 4)) ||
      (rc = dmalloc(n, (void**)&n->m1, P * 4)) || (rc = dmalloc(n, (void**)&n->m2, P * 4)) ||
      (rc = dmalloc(n, (void**)&n->wb, P * 2))) {
    p3d_nn_destroy(n);
    return rc;
  }
  // He-normal init of the weights, zero biases (seeded)
  std::vector<float> hw(P, 0.f);
  std::mt19937_64 rng(seed);
  for (size_t li = 0; li < n->conv.size(); li++) {
    const ConvSpec& cs = n->conv[li];
    const float sd = sqrtf(2.0f / (27.0f * cs.Cin)) * (li == 7 ? 0.1f : 1.0f);
    std::normal_distribution<float> nd(0.f, sd);
    for (size_t e = 0; e < (size_t)27 * cs.Cin * cs.Cout; e++) hw[cs.w_off + e] = nd(rng);
  }
  hipMemcpy(n->w, hw.data(), P * 4, hipMemcpyHostToDevice);
  hipMemset(n->m1, 0, P * 4);
  hipMemset(n->m2, 0, P * 4);
  hipMemset(n->g, 0, P * 4);
  to_bf16(P, n->w, n->wb, n->stream);
  const size_t V0 = n->D[0].V(), V1 = n->D[1].V(), V2 = n->D[2].V(), V3 = n->D[3].V();
  struct {
    uint16_t** p;
    size_t n;
  } acts[] = {{&n->x0, V0 * 16}, {&n->e0, V0 * 16}, {&n->p1, V1 * 16}, {&n->e1, V1 * 32},
              {&n->p2, V2 * 32}, {&n->e2, V2 * 64}, {&n->p3, V3 * 64}, {&n->e3, V3 * 128},
              {&n->u2, V2 * 128}, {&n->d2, V2 * 64}, {&n->s2, V2 * 64}, {&n->u1, V1 * 64},
              {&n->d1, V1 * 32}, {&n->s1, V1 * 32}, {&n->u0, V0 * 32}, {&n->d0, V0 * 16},
              {&n->s0, V0 * 16}};
  for (auto& a : acts)
    if ((rc = dmalloc(n, (void**)a.p, a.n * 2))) {
      p3d_nn_destroy(n);
      return rc;
    }
  struct {
    float** p;
    size_t n;
  } grads[] = {{&n->out, V0 * 16},   {&n->g_out, V0 * 16}, {&n->g_s0, V0 * 16},
               {&n->g_t0, V0 * 16},  {&n->g_u0, V0 * 32},  {&n->g_s1, V1 * 32},
               {&n->g_t1, V1 * 32},  {&n->g_u1, V1 * 64},  {&n->g_s2, V2 * 64},
               {&n->g_t2, V2 * 64},  {&n->g_u2, V2 * 128}, {&n->g_e3, V3 * 128},
               {&n->g_p3, V3 * 64},  {&n->g_e2, V2 * 64},  {&n->g_p2, V2 * 32},
               {&n->g_e1, V1 * 32},  {&n->g_p1, V1 * 16},  {&n->g_e0, V0 * 16},
               {&n->tgt, V0},        {&n->lpart, V0 / 256 + 2}, {&n->loss, 4}};
  for (auto& a : grads)
    if ((rc = dmalloc(n, (void**)a.p, a.n * 4))) {
      p3d_nn_destroy(n);
      return rc;
    }
  // weight-gradient partials: the largest nchunk*27*Cout*Cin over layers
  size_t pn = 0;
  for (const ConvSpec& cs : n->conv) {
    const size_t need = (size_t)nchunk_for(n->D[cs.level]) * 27 * cs.Cin * cs.Cout;
    if (need > pn) pn = need;
  }
  if ((rc = dmalloc(n, (void**)&n->part, pn * 4))) {
    p3d_nn_destroy(n);
    return rc;
  }
  n->part_n = pn;
  // radial feature tables on the interior shells: (r0/r)^2 and (r-r0)/(r1-r0)
  std::vector<float> dec(n->a), rp(n->a);
  const double r0 = 0.5 * (rh[0] + rh[1]), r1 = 0.5 * (rh[nr - 2] + rh[nr - 1]);
  for (int i = 0; i < n->a; i++) {
    const double r = rh[i + 1];
    dec[i] = (float)((r0 / r) * (r0 / r));
    rp[i] = (float)((r - r0) / (r1 - r0));
  }
  if ((rc = dmalloc(n, (void**)&n->decay, n->a * 4)) ||
      (rc = dmalloc(n, (void**)&n->rpos, n->a * 4)) ||
      (rc = dmalloc(n, (void**)&n->br_dev, (size_t)nt * np * 8))) {
    p3d_nn_destroy(n);
    return rc;
  }
  hipMemcpy(n->decay, dec.data(), n->a * 4, hipMemcpyHostToDevice);
  hipMemcpy(n->rpos, rp.data(), n->a * 4, hipMemcpyHostToDevice);
  if (hipDeviceSynchronize() != hipSuccess) {
    p3d_nn_destroy(n);
    return nfail(P3D_EHIP, "set-up failed");
  }
  *out = n;
  return P3D_OK;
}

size_t p3d_nn_num_params(p3d_nn* n) { return n ? n->nparam : 0; }

int p3d_nn_set_scales(p3d_nn* n, double bscale, double phiscale) {
  if (!n || !(bscale >= 0) || !(phiscale > 0))
    return nfail(P3D_EINVAL, "need bscale >= 0 and phiscale > 0");
  n->bscale = bscale;
  n->phiscale = phiscale;
  return P3D_OK;
}

int p3d_nn_get_params(p3d_nn* n, float* buf) {
  if (!n || !buf) return nfail(P3D_EINVAL, "null argument");
  if (hipMemcpy(buf, n->w, n->nparam * 4, hipMemcpyDeviceToHost) != hipSuccess)
    return nfail(P3D_EHIP, "copy failed");
  return P3D_OK;
}

int p3d_nn_set_params(p3d_nn* n, const float* buf) {
  if (!n || !buf) return nfail(P3D_EINVAL, "null argument");
  if (hipMemcpy(n->w, buf, n->nparam * 4, hipMemcpyHostToDevice) != hipSuccess)
    return nfail(P3D_EHIP, "copy failed");
  to_bf16(n->nparam, n->w, n->wb, n->stream);
  if (hipStreamSynchronize(n->stream) != hipSuccess) return nfail(P3D_EHIP, "cast failed");
  return P3D_OK;
}

// Adds a training pair (br0(nt,np), solution x(N) of the solve) to the
// device-resident training set.
int p3d_nn_add_sample(p3d_nn* n, const double* br0, const double* x) {
  if (!n || !br0 || !x) return nfail(P3D_EINVAL, "null argument");
  const size_t nb = (size_t)n->nt * (n->c + 2), V = n->D[0].V();
  double* db = nullptr;
  float* dt = nullptr;
  if (hipMalloc((void**)&db, nb * 8) != hipSuccess || hipMalloc((void**)&dt, V * 4) != hipSuccess) {
    if (db) hipFree(db);
    return nfail(P3D_ENOMEM, "training set allocation failed");
  }
  std::vector<float> t(V);
  const double bs = sample_scale(n, br0), ts = target_scale(n, bs);
  for (size_t v = 0; v < V; v++) t[v] = (float)(x[v] / ts);
  hipMemcpy(db, br0, nb * 8, hipMemcpyHostToDevice);
  hipMemcpy(dt, t.data(), V * 4, hipMemcpyHostToDevice);
  n->ds_br.push_back(db);
  n->ds_t.push_back(dt);
  n->ds_bs.push_back(bs);
  return P3D_OK;
}

// One forward/backward on (br0, x) (host arrays): loss and, if grads is
// non-null, the parameter gradient (nparam floats).  Used for gradient checks.
int p3d_nn_loss_grad(p3d_nn* n, const double* br0, const double* x, double* loss, float* grads) {
  if (!n || !br0 || !x || !loss) return nfail(P3D_EINVAL, "null argument");
  const size_t nb = (size_t)n->nt * (n->c + 2), V = n->D[0].V();
  std::vector<float> t(V);
  const double bs = sample_scale(n, br0), ts = target_scale(n, bs);
  for (size_t v = 0; v < V; v++) t[v] = (float)(x[v] / ts);
  hipMemcpyAsync(n->br_dev, br0, nb * 8, hipMemcpyHostToDevice, n->stream);
  hipMemcpyAsync(n->tgt, t.data(), V * 4, hipMemcpyHostToDevice, n->stream);
  make_features(n, n->br_dev, bs);
  forward(n);
  mse((int)V, 16, n->out, n->tgt, n->g_out, n->lpart, n->loss, n->stream);
  backward(n);
  float l = 0.f;
  hipMemcpyAsync(&l, n->loss, 4, hipMemcpyDeviceToHost, n->stream);
  if (grads) hipMemcpyAsync(grads, n->g, n->nparam * 4, hipMemcpyDeviceToHost, n->stream);
  if (hipStreamSynchronize(n->stream) != hipSuccess) return nfail(P3D_EHIP, "forward/backward failed");
  *loss = l;
  return P3D_OK;
}

// Adam training over the stored set: nsteps steps, samples in a seeded
// shuffled order; losses[step] receives each step's loss (may be null).
int p3d_nn_train(p3d_nn* n, int nsteps, double lr, unsigned seed, float* losses) {
  if (!n || nsteps < 0) return nfail(P3D_EINVAL, "bad argument");
  const int ns = (int)n->ds_br.size();
  if (ns == 0) return nfail(P3D_ESTATE, "no training samples");
  float* dloss = nullptr;
  if (hipMalloc((void**)&dloss, sizeof(float) * (nsteps > 0 ? nsteps : 1)) != hipSuccess)
    return nfail(P3D_ENOMEM, "allocation failed");
  std::vector<int> order;
  std::mt19937_64 rng(seed);
  const int V = (int)n->D[0].V();
  for (int it = 0; it < nsteps; it++) {
    if (order.empty()) {
      for (int q = 0; q < ns; q++) order.push_back(q);
      std::shuffle(order.begin(), order.end(), rng);
    }
    const int q = order.back();
    order.pop_back();
    const float* tq = n->ds_t[q];
    double bs = n->ds_bs[q];
    int roll = 0, flip = 0;
    if (n->aug_roll || n->aug_flip || n->aug_sign) {
      const uint64_t u = rng();
      roll = n->aug_roll ? (int)((u >> 8) % (uint64_t)n->c) : 0;
      flip = n->aug_flip ? (int)(u & 1) : 0;
      const float sign = (n->aug_sign && (u & 2)) ? -1.f : 1.f;
      bs *= sign;
      gather_target(n->D[0], roll, flip, sign, tq, n->tgt, n->stream);
      tq = n->tgt;
    }
    make_features(n, n->ds_br[q], bs, roll, flip);
    forward(n);
    mse(V, 16, n->out, tq, n->g_out, n->lpart, dloss + it, n->stream);
    backward(n);
    n->step++;
    adam(n->nparam, n->w, n->g, n->m1, n->m2, (float)lr, 0.9f, 0.999f, 1e-8f, n->step, n->wb,
         n->stream);
  }
  hipError_t e = hipStreamSynchronize(n->stream);
  if (e == hipSuccess && losses && nsteps > 0)
    e = hipMemcpy(losses, dloss, sizeof(float) * nsteps, hipMemcpyDeviceToHost);
  hipFree(dloss);
  if (e != hipSuccess) return nfail(P3D_EHIP, "training failed");
  return P3D_OK;
}

// Training augmentation switches (see struct p3d_nn); all off by default.
int p3d_nn_set_augment(p3d_nn* n, int phi_roll, int theta_flip, int sign_flip) {
  if (!n) return nfail(P3D_EINVAL, "null argument");
  n->aug_roll = phi_roll != 0;
  n->aug_flip = theta_flip != 0;
  n->aug_sign = sign_flip != 0;
  return P3D_OK;
}

// Prediction for br0(nt,np): x(N) = phi_scale * out[.][0] (host or device
// output; x_dev used when non-null).
int p3d_nn_predict(p3d_nn* n, const double* br0, double* x, double* x_dev) {
  if (!n || !br0 || (!x && !x_dev)) return nfail(P3D_EINVAL, "null argument");
  const size_t nb = (size_t)n->nt * (n->c + 2), V = n->D[0].V();
  const double bs = sample_scale(n, br0);
  hipMemcpyAsync(n->br_dev, br0, nb * 8, hipMemcpyHostToDevice, n->stream);
  make_features(n, n->br_dev, bs);
  forward(n);
  double* dst = x_dev;
  double* tmp = nullptr;
  if (!dst) {
    if (hipMalloc((void**)&tmp, V * 8) != hipSuccess) return nfail(P3D_ENOMEM, "allocation failed");
    dst = tmp;
  }
  head_to_x((int)V, 16, n->out, target_scale(n, bs), dst, n->stream);
  if (x) hipMemcpyAsync(x, dst, V * 8, hipMemcpyDeviceToHost, n->stream);
  hipError_t e = hipStreamSynchronize(n->stream);
  if (tmp) hipFree(tmp);
  if (e != hipSuccess) return nfail(P3D_EHIP, "prediction failed");
  return P3D_OK;
}

// Trilinear prolongation of a padded coarse phi (nrc x ntc x npc) to the
// fine interior points (af x bf x cf) with per-axis tables (host): lower
// coarse padded index and upper weight per fine interior index.  dev: phic
// and xf are device addresses (else host arrays).
static int prolong_check(int nrc, int ntc, int npc, const double* phic, int af, int bf, int cf,
                         const int* ir, const double* wr, const int* it, const double* wt,
                         const int* ip, const double* wp, const double* xf) {
  if (!phic || !ir || !wr || !it || !wt || !ip || !wp || !xf)
    return nfail(P3D_EINVAL, "null argument");
  if (nrc < 2 || ntc < 2 || npc < 2 || af < 1 || bf < 1 || cf < 1)
    return nfail(P3D_EINVAL, "bad grid sizes");
  for (int q = 0; q < af; q++)
    if (ir[q] < 0 || ir[q] + 1 >= nrc) return nfail(P3D_EINVAL, "r table out of range");
  for (int q = 0; q < bf; q++)
    if (it[q] < 0 || it[q] + 1 >= ntc) return nfail(P3D_EINVAL, "theta table out of range");
  for (int q = 0; q < cf; q++)
    if (ip[q] < 0 || ip[q] + 1 >= npc) return nfail(P3D_EINVAL, "phi table out of range");
  return P3D_OK;
}

static int prolong_impl(int nrc, int ntc, int npc, const double* phic, int af, int bf, int cf,
                        const int* ir, const double* wr, const int* it, const double* wt,
                        const int* ip, const double* wp, double* xf, bool dev) {
  const int rc = prolong_check(nrc, ntc, npc, phic, af, bf, cf, ir, wr, it, wt, ip, wp, xf);
  if (rc != P3D_OK) return rc;
  const size_t nc = (size_t)nrc * ntc * npc, nf = (size_t)af * bf * cf;
  const size_t nt_ = (size_t)af + bf + cf;
  double *dc = nullptr, *df = nullptr, *dw = nullptr;
  int* di = nullptr;
  bool ok = hipMalloc((void**)&dw, nt_ * 8) == hipSuccess &&
            hipMalloc((void**)&di, nt_ * 4) == hipSuccess;
  if (ok && !dev)
    ok = hipMalloc((void**)&dc, nc * 8) == hipSuccess && hipMalloc((void**)&df, nf * 8) == hipSuccess;
  if (!ok) {
    hipFree(dc);
    hipFree(df);
    hipFree(dw);
    hipFree(di);
    return nfail(P3D_ENOMEM, "allocation failed");
  }
  const double* src = dev ? phic : dc;
  double* dst = dev ? xf : df;
  if (!dev) hipMemcpy(dc, phic, nc * 8, hipMemcpyHostToDevice);
  hipMemcpy(di, ir, af * 4, hipMemcpyHostToDevice);
  hipMemcpy(di + af, it, bf * 4, hipMemcpyHostToDevice);
  hipMemcpy(di + af + bf, ip, cf * 4, hipMemcpyHostToDevice);
  hipMemcpy(dw, wr, af * 8, hipMemcpyHostToDevice);
  hipMemcpy(dw + af, wt, bf * 8, hipMemcpyHostToDevice);
  hipMemcpy(dw + af + bf, wp, cf * 8, hipMemcpyHostToDevice);
  prolong(af, bf, cf, nrc, ntc, npc, src, di, dw, di + af, dw + af, di + af + bf, dw + af + bf,
          dst, 0);
  hipError_t e = dev ? hipDeviceSynchronize() : hipMemcpy(xf, df, nf * 8, hipMemcpyDeviceToHost);
  hipFree(dc);
  hipFree(df);
  hipFree(dw);
  hipFree(di);
  if (e != hipSuccess) return nfail(P3D_EHIP, "prolongation failed");
  return P3D_OK;
}

int p3d_nn_prolong(int nrc, int ntc, int npc, const double* phic, int af, int bf, int cf,
                   const int* ir, const double* wr, const int* it, const double* wt,
                   const int* ip, const double* wp, double* xf) {
  return prolong_impl(nrc, ntc, npc, phic, af, bf, cf, ir, wr, it, wt, ip, wp, xf, false);
}

int p3d_nn_prolong_dev(int nrc, int ntc, int npc, const double* phic_dev, int af, int bf, int cf,
                       const int* ir, const double* wr, const int* it, const double* wt,
                       const int* ip, const double* wp, double* xf_dev) {
  return prolong_impl(nrc, ntc, npc, phic_dev, af, bf, cf, ir, wr, it, wt, ip, wp, xf_dev, true);
}

// p3d_nn_prolong_dev with its tables staged in the device array `scratch`
// instead of buffers allocated per call: no hipMalloc/hipFree.  Rejects
// af + bf + cf > kProlongMaxTable.  The tables are rewritten whole on every
// call, so nothing needs clearing; two calls overlapping on one GPU race on
// them.
int p3d_nn_prolong_dev_noalloc(int nrc, int ntc, int npc, const double* phic_dev, int af, int bf,
                               int cf, const int* ir, const double* wr, const int* it,
                               const double* wt, const int* ip, const double* wp, double* xf_dev) {
  const int rc =
      prolong_check(nrc, ntc, npc, phic_dev, af, bf, cf, ir, wr, it, wt, ip, wp, xf_dev);
  if (rc != P3D_OK) return rc;
  if ((long long)af + bf + cf > kProlongMaxTable) {
    char msg[160];
    snprintf(msg, sizeof msg,
             "prolongation tables of %lld entries (af+bf+cf) exceed the no-allocation limit "
             "of %d",
             (long long)af + bf + cf, kProlongMaxTable);
    return nfail(P3D_EINVAL, msg);
  }
  double* ptr = scratch_on_device();
  if (!ptr) return P3D_EHIP;
  double* dw = ptr;                                      // wr | wt | wp
  int* di = reinterpret_cast<int*>(ptr + kProlongMaxTable);  // ir | it | ip
  bool ok = hipMemcpy(di, ir, (size_t)af * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(di + af, it, (size_t)bf * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(di + af + bf, ip, (size_t)cf * 4, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dw, wr, (size_t)af * 8, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dw + af, wt, (size_t)bf * 8, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemcpy(dw + af + bf, wp, (size_t)cf * 8, hipMemcpyHostToDevice) == hipSuccess;
  if (!ok) return nfail(P3D_EHIP, "table copy failed");
  prolong(af, bf, cf, nrc, ntc, npc, phic_dev, di, dw, di + af, dw + af, di + af + bf,
          dw + af + bf, xf_dev, 0);
  if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess)
    return nfail(P3D_EHIP, "prolongation failed");
  return P3D_OK;
}

// One 16x16x16 BF16 MFMA: D = A*B (A, B row-major BF16 bits, D FP32).
int p3d_nn_mfma_check(const uint16_t* A, const uint16_t* B, float* D) {
  uint16_t *dA = nullptr, *dB = nullptr;
  float* dD = nullptr;
  if (hipMalloc((void**)&dA, 512) != hipSuccess || hipMalloc((void**)&dB, 512) != hipSuccess ||
      hipMalloc((void**)&dD, 1024) != hipSuccess)
    return nfail(P3D_ENOMEM, "allocation failed");
  hipMemcpy(dA, A, 512, hipMemcpyHostToDevice);
  hipMemcpy(dB, B, 512, hipMemcpyHostToDevice);
  mfma_check(dA, dB, dD, 0);
  hipError_t e = hipMemcpy(D, dD, 1024, hipMemcpyDeviceToHost);
  hipFree(dA);
  hipFree(dB);
  hipFree(dD);
  return e == hipSuccess ? P3D_OK : nfail(P3D_EHIP, "mfma check failed");
}

}  // extern "C"
