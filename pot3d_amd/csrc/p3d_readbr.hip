// readbr's arithmetic on the GPU: the photospheric Br map interpolated onto
// the solver's theta/phi half mesh, made periodic in phi, given its m=0 pole
// ghosts, flux balanced, and (option 'open') replaced by |Br|.
//
//   intrp2d + interp   src/pot3d.F90:6830-6910, 7028-7072 (bilinear)
//   readbr             src/pot3d.F90:6751-6829 (seam, poles, flux balance)
//   set_flux           src/pot3d.F90:3806-3833 (|Br| for option 'open')
//
// The driver keeps the HDF5 read (rdhdf_2d) and hands over the map and its
// scales.  Interpolation, seam and pole values follow the reference's
// expression order with FMA contraction off, so they are bit-identical to a
// plain FP64 evaluation; the three flux sums are FP64 in a fixed parallel
// order (one workgroup), so br00err can differ from a sequential sum in its
// last bits.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>

#include "../../include/pot3d_hip.h"

namespace p3d {
int set_error(int code, const char* msg);
}

namespace {

constexpr int kFluxThreads = 1024;

// interp (src/pot3d.F90:7028-7072): the first interval i with
// x[i] <= v <= x[i+1], alpha = (v - x[i]) / (x[i+1] - x[i]).  With x strictly
// increasing and x[0] <= v <= x[n-1] (checked on the host) that interval is
// the smallest i with x[i+1] >= v, found here by bisection.
__global__ void k_axis_table(int n, const double* __restrict__ x, int m,
                             const double* __restrict__ v, int* __restrict__ idx,
                             double* __restrict__ alpha) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m) return;
  const double xv = v[q];
  int lo = 0, hi = n - 2;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (x[mid + 1] >= xv)
      hi = mid;
    else
      lo = mid + 1;
  }
  idx[q] = lo;
  alpha[q] = (xv - x[lo]) / (x[lo + 1] - x[lo]);
}

// intrp2d's bilinear formula (src/pot3d.F90:6891-6892) on the interior
// (j = 2..nt-1, k = 2..np-1 in the reference's 1-based indices); br0 is
// (nt, np) in Fortran order, bn (ntn, npn).
__global__ void k_intrp2d(int nt, int np, int ntn, const double* __restrict__ bn,
                          const int* __restrict__ it, const double* __restrict__ at,
                          const int* __restrict__ ip, const double* __restrict__ ap,
                          double* __restrict__ br0) {
#pragma clang fp contract(off)
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = nt - 2, c = np - 2;
  if (q >= b * c) return;
  const int j = q % b, k = q / b;
  const int ii = it[j], jj = ip[k];
  const double ax = at[j], ay = ap[k];
  const double f00 = bn[ii + (size_t)ntn * jj], f01 = bn[ii + (size_t)ntn * (jj + 1)];
  const double f10 = bn[ii + 1 + (size_t)ntn * jj], f11 = bn[ii + 1 + (size_t)ntn * (jj + 1)];
  br0[(j + 1) + (size_t)nt * (k + 1)] =
      (1.0 - ax) * ((1.0 - ay) * f00 + ay * f01) + ax * ((1.0 - ay) * f10 + ay * f11);
}

// phi seam of the interior rows (br0(:,1) = br0(:,np-1), br0(:,np) = br0(:,2),
// src/pot3d.F90:6777-6778); the pole rows are set afterwards.
__global__ void k_seam(int nt, int np, double* __restrict__ br0) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < 1 || j >= nt - 1) return;
  br0[j] = br0[j + (size_t)nt * (np - 2)];
  br0[j + (size_t)nt * (np - 1)] = br0[j + (size_t)nt];
}

// m=0 pole ghosts (src/pot3d.F90:6784-6788): sum0/sum1 = sum over k=2..np-1
// of br0(2,k)*dph(k) (resp. row nt-1), in k order, times pl_i; then
// br0(1,:) = 2 sum0 - br0(2,:), br0(nt,:) = 2 sum1 - br0(nt-1,:).  One
// workgroup: two lanes form the ordered sums, all lanes write the rows.
__global__ void k_poles(int nt, int np, const double* __restrict__ dph, double pl_i,
                        double* __restrict__ br0) {
#pragma clang fp contract(off)
  __shared__ double sums[2];
  const int t = threadIdx.x;
  if (t < 2) {
    const int row = t == 0 ? 1 : nt - 2;
    double s = 0.0;
    for (int k = 1; k < np - 1; k++) s = s + br0[row + (size_t)nt * k] * dph[k];
    sums[t] = s * pl_i;
  }
  __syncthreads();
  for (int k = t; k < np; k += blockDim.x) {
    br0[(size_t)nt * k] = 2.0 * sums[0] - br0[1 + (size_t)nt * k];
    br0[nt - 1 + (size_t)nt * k] = 2.0 * sums[1] - br0[nt - 2 + (size_t)nt * k];
  }
}

// Positive flux, negative flux and area over the interior (src/pot3d.F90:
// 6792-6804), da = sth(j)*dth(j)*dph(k): one workgroup, each lane summing a
// fixed strided subset in order, then a fixed tree; out[0..2] = fluxp,
// fluxm, area.
__global__ void __launch_bounds__(kFluxThreads)
    k_flux(int nt, int np, const double* __restrict__ sth, const double* __restrict__ dth,
           const double* __restrict__ dph, const double* __restrict__ br0,
           double* __restrict__ out) {
#pragma clang fp contract(off)
  __shared__ double sp[kFluxThreads], sm[kFluxThreads], sa[kFluxThreads];
  const int t = threadIdx.x;
  const int b = nt - 2, c = np - 2;
  double fp = 0.0, fm = 0.0, ar = 0.0;
  for (int q = t; q < b * c; q += kFluxThreads) {
    const int j = 1 + q % b, k = 1 + q / b;
    const double da = sth[j] * dth[j] * dph[k];
    const double v = br0[j + (size_t)nt * k];
    if (v > 0.0)
      fp = fp + v * da;
    else
      fm = fm + v * da;
    ar = ar + da;
  }
  sp[t] = fp;
  sm[t] = fm;
  sa[t] = ar;
  __syncthreads();
  for (int w = kFluxThreads / 2; w > 0; w >>= 1) {
    if (t < w) {
      sp[t] = sp[t] + sp[t + w];
      sm[t] = sm[t] + sm[t + w];
      sa[t] = sa[t] + sa[t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    out[0] = sp[0];
    out[1] = sm[0];
    out[2] = sa[0];
  }
}

// br0 -= br00err with br00err = (fluxp + fluxm)/area when balancing
// (src/pot3d.F90:6812-6823); then |br0| for option 'open' (set_flux,
// :3832).  out[3] receives br00err (0 when not balanced).
__global__ void k_balance(int n, int balance, int abs_br, double* __restrict__ out,
                          double* __restrict__ br0) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  const double e = balance ? (out[0] + out[1]) / out[2] : 0.0;
  if (q == 0) out[3] = e;
  if (q >= n) return;
  double v = br0[q];
  if (balance) v = v - e;
  br0[q] = abs_br ? fabs(v) : v;
}

inline int nblk(long long n, int t) { return (int)((n + t - 1) / t); }

int fail(int code, const char* msg) { return p3d::set_error(code, msg); }

}  // namespace

extern "C" int p3d_readbr(int nt, int np, const double* th, const double* ph,
                          const double* sth, const double* dth, const double* dph, double pl,
                          int ntn, int npn, const double* tn, const double* pn, const double* bn,
                          int balance_flux, int abs_br, double* br0, double* flux) {
  if (!th || !ph || !sth || !dth || !dph || !tn || !pn || !bn || !br0)
    return fail(P3D_EINVAL, "null argument");
  if (nt < 3 || np < 3 || ntn < 2 || npn < 2 || !(pl > 0.0))
    return fail(P3D_EINVAL, "bad sizes (need nt, np >= 3, map scales of >= 2 points, pl > 0)");
  if ((long long)nt * np >= (1LL << 31) || (long long)ntn * npn >= (1LL << 31))
    return fail(P3D_EINVAL, "arrays too large for 32-bit indices");
  // intrp2d's checks (src/pot3d.F90:6870-6889): increasing scales, and every
  // mesh point inside them
  for (int i = 0; i + 1 < ntn; i++)
    if (!(tn[i + 1] > tn[i]))
      return fail(P3D_EINVAL, "theta scale of the Br map is not monotonically increasing");
  for (int i = 0; i + 1 < npn; i++)
    if (!(pn[i + 1] > pn[i]))
      return fail(P3D_EINVAL, "phi scale of the Br map is not monotonically increasing");
  for (int j = 1; j < nt - 1; j++)
    if (!(th[j] >= tn[0] && th[j] <= tn[ntn - 1]))
      return fail(P3D_EINVAL, "theta mesh point outside the range of the Br map's scale");
  for (int k = 1; k < np - 1; k++)
    if (!(ph[k] >= pn[0] && ph[k] <= pn[npn - 1]))
      return fail(P3D_EINVAL, "phi mesh point outside the range of the Br map's scale");

  const size_t nb = (size_t)nt * np, nm = (size_t)ntn * npn;
  // one device block: bn | tn | pn | th | ph | sth | dth | dph | alpha_t |
  // alpha_p | br0 | out[4]; then the int tables
  const size_t nd = nm + ntn + npn + 2 * (size_t)nt + 2 * (size_t)np + 2 * (size_t)nt + np + nb + 4;
  double* d = nullptr;
  int* di = nullptr;
  if (hipMalloc((void**)&d, nd * 8) != hipSuccess) return fail(P3D_ENOMEM, "allocation failed");
  if (hipMalloc((void**)&di, ((size_t)nt + np) * 4) != hipSuccess) {
    hipFree(d);
    return fail(P3D_ENOMEM, "allocation failed");
  }
  double* d_bn = d;
  double* d_tn = d_bn + nm;
  double* d_pn = d_tn + ntn;
  double* d_th = d_pn + npn;
  double* d_ph = d_th + nt;
  double* d_sth = d_ph + np;
  double* d_dth = d_sth + nt;
  double* d_dph = d_dth + nt;
  double* d_at = d_dph + np;
  double* d_ap = d_at + nt;
  double* d_br = d_ap + np;
  double* d_out = d_br + nb;
  int* d_it = di;
  int* d_ip = di + nt;
  hipStream_t st = nullptr;
  bool ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
  struct Up {
    double* dst;
    const double* src;
    size_t n;
  } ups[] = {{d_bn, bn, nm},  {d_tn, tn, (size_t)ntn}, {d_pn, pn, (size_t)npn},
             {d_th, th, (size_t)nt}, {d_ph, ph, (size_t)np}, {d_sth, sth, (size_t)nt},
             {d_dth, dth, (size_t)nt}, {d_dph, dph, (size_t)np}};
  for (const Up& u : ups)
    ok = ok && hipMemcpyAsync(u.dst, u.src, u.n * 8, hipMemcpyHostToDevice, st) == hipSuccess;
  if (ok) {
    // tables for the interior points th(2:nt-1), ph(2:np-1)
    k_axis_table<<<nblk(nt - 2, 256), 256, 0, st>>>(ntn, d_tn, nt - 2, d_th + 1, d_it, d_at);
    k_axis_table<<<nblk(np - 2, 256), 256, 0, st>>>(npn, d_pn, np - 2, d_ph + 1, d_ip, d_ap);
    k_intrp2d<<<nblk((long long)(nt - 2) * (np - 2), 256), 256, 0, st>>>(nt, np, ntn, d_bn, d_it,
                                                                       d_at, d_ip, d_ap, d_br);
    k_seam<<<nblk(nt, 256), 256, 0, st>>>(nt, np, d_br);
    k_poles<<<1, 256, 0, st>>>(nt, np, d_dph, 1.0 / pl, d_br);
    k_flux<<<1, kFluxThreads, 0, st>>>(nt, np, d_sth, d_dth, d_dph, d_br, d_out);
    k_balance<<<nblk((long long)nb, 256), 256, 0, st>>>((int)nb, balance_flux != 0, abs_br != 0,
                                                        d_out, d_br);
    ok = hipGetLastError() == hipSuccess;
  }
  double out[4] = {0, 0, 0, 0};
  ok = ok && hipMemcpyAsync(br0, d_br, nb * 8, hipMemcpyDeviceToHost, st) == hipSuccess;
  ok = ok && hipMemcpyAsync(out, d_out, 32, hipMemcpyDeviceToHost, st) == hipSuccess;
  ok = ok && hipStreamSynchronize(st) == hipSuccess;
  if (st) hipStreamDestroy(st);
  hipFree(d);
  hipFree(di);
  if (!ok) return fail(P3D_EHIP, "readbr kernels failed");
  if (flux)
    for (int q = 0; q < 4; q++) flux[q] = out[q];
  return P3D_OK;
}
