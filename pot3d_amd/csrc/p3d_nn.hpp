// Warm-start network: kernel launchers (p3d_nn_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace p3d {
namespace nn {

struct Dims3 {
  int X, Y, Z;
  size_t V() const { return (size_t)X * Y * Z; }
};

void conv_fwd(const Dims3& d, int Cin, int Cout, const uint16_t* x, const uint16_t* w,
              const float* bias, uint16_t* y, float* yf, bool relu, hipStream_t s);
// dx may be null; gw (27*Cout*Cin), gb (Cout) FP32; part: scratch of
// nchunk*27*Cout*Cin floats.
void conv_bwd(const Dims3& d, int Cin, int Cout, const float* dy, const uint16_t* x,
              const uint16_t* w, float* dx, float* gw, float* gb, float* part, int nchunk,
              hipStream_t s);
void pool_fwd(const Dims3& f, const Dims3& c, int C, const uint16_t* x, uint16_t* y,
              hipStream_t s);
void pool_bwd(const Dims3& f, const Dims3& c, int C, const float* dy, float* dx, hipStream_t s);
void up_fwd(const Dims3& c, const Dims3& f, int C, const uint16_t* x, uint16_t* y,
            hipStream_t s);
void up_bwd(const Dims3& c, const Dims3& f, int C, const float* dy, float* dx, hipStream_t s);
void add_bf16(size_t n, const uint16_t* a, const uint16_t* b, uint16_t* y, hipStream_t s);
void relu_bwd(size_t n, const uint16_t* y, const float* g, const float* g2, float* out,
              hipStream_t s);
void mse(int V, int C, const float* out, const float* tgt, float* g, float* part, float* loss,
         hipStream_t s);
void adam(size_t n, float* w, const float* g, float* m1, float* m2, float lr, float b1,
          float b2, float eps, int step, uint16_t* wb, hipStream_t s);
void to_bf16(size_t n, const float* x, uint16_t* y, hipStream_t s);
void fill_u16(size_t n, uint16_t* p, uint16_t v, hipStream_t s);
// roll/flip: the sample rotated by roll interior phi points and reflected in
// theta (training augmentation; 0, 0 otherwise)
void features(const Dims3& d, int C, int nt, const double* br0, const float* decay,
              const float* rpos, float inv_bscale, int roll, int flip, uint16_t* x0,
              hipStream_t s);
void gather_target(const Dims3& d, int roll, int flip, float sign, const float* src, float* dst,
                   hipStream_t s);
void head_to_x(int V, int C, const float* out, double scale, double* xo, hipStream_t s);
void prolong(int af, int bf, int cf, int nrc, int ntc, int npc, const double* phic,
             const int* ir, const double* wr, const int* it, const double* wt, const int* ip,
             const double* wp, double* xf, hipStream_t s);
void mfma_check(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s);

}  // namespace nn
}  // namespace p3d
