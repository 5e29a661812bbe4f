// Shell collocation (+gradient) on gfx950.
//
// Semantics: GauXC's collocation (gau2grid_collocation.cxx:25-145 on the
// host, collocation_device_masked_combined_kernel_deriv1 on the device):
//   B_mu(r) = R(|r-A|) Y_mu(r-A),  R(r) = sum_p c_p exp(-a_p r^2),
//   d/dx B = (dY/dx) R + Y x dR/dr / r = (dY/dx) R - 2 x Y sum_p a_p c_p e^{-a_p r^2},
// Cartesian functions in CCA order (x^l, x^(l-1)y, ..., z^l), spherical
// functions as real solid harmonics in CCA order m = -l..l, unnormalised as
// gau2grid's GG_SPHERICAL_CCA.
//
// Layout: task t owns [comp][nbe][npts_pad] in the chunk scratch
// (comp = B, dB/dx, dB/dy, dB/dz), point index fastest, padded points 0.
// Launch: one 256-thread block per (task, 64-point block); lane = point,
// the four waves stride over the task's shells.  Stores of one wave are
// 64 consecutive points of one basis function: fully coalesced.
#include "internal.hpp"

#include <cmath>

namespace xcq {

__constant__ double c_sph[kMaxL + 1][2 * kMaxL + 1][(kMaxL + 1) * (kMaxL + 2) / 2];

static double h_fact(int n) {
  double r = 1.;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
static double h_binom(int n, int k) {
  return (k < 0 || k > n) ? 0. : h_fact(n) / (h_fact(k) * h_fact(n - k));
}
static int cart_index(int l, int a, int b) {  // CCA position of x^a y^b z^(l-a-b)
  int idx = 0;
  for (int aa = l; aa >= 0; --aa)
    for (int bb = l - aa; bb >= 0; --bb, ++idx)
      if (aa == a && bb == b) return idx;
  return -1;
}

// Real solid harmonics from the closed form of Helgaker, Jorgensen & Olsen,
// "Molecular Electronic-Structure Theory", eq. 6.4.47-6.4.50:
//   S_lm = N_lm sum_{t,u,v} C^{lm}_{tuv} x^{2t+|m|-2(u+v)} y^{2(u+v)} z^{l-2t-|m|},
//   C = (-1)^{t+v-v_m} (1/4)^t binom(l,t) binom(l-t,|m|+t) binom(t,u) binom(|m|,2v),
//   N_lm = sqrt(2 (l+|m|)! (l-|m|)! / 2^{delta_m0}) / (2^|m| l!),
// v_m = 0 (m >= 0) or 1/2 (m < 0), v running in steps of 1 from v_m.
int upload_sph_table(hipStream_t s) {
  static double tbl[kMaxL + 1][2 * kMaxL + 1][(kMaxL + 1) * (kMaxL + 2) / 2];
  for (auto& a : tbl)
    for (auto& b : a)
      for (auto& c : b) c = 0.;
  for (int l = 0; l <= kMaxL; ++l)
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      const double N = std::sqrt(2. * h_fact(l + am) * h_fact(l - am) / (m == 0 ? 2. : 1.)) /
                       (std::ldexp(1.0, am) * h_fact(l));
      const int v2s = m < 0 ? 1 : 0;  // 2v starts at 0 (m >= 0) or 1 (m < 0)
      for (int t = 0; t <= (l - am) / 2; ++t)
        for (int u = 0; u <= t; ++u)
          for (int v2 = v2s; v2 <= am; v2 += 2) {
            const int se = t + (v2 - v2s) / 2;
            const double C = ((se & 1) ? -1. : 1.) * std::pow(0.25, t) * h_binom(l, t) *
                             h_binom(l - t, am + t) * h_binom(t, u) * h_binom(am, v2);
            const int ax = 2 * t + am - 2 * u - v2, ay = 2 * u + v2;
            tbl[l][m + l][cart_index(l, ax, ay)] += N * C;
          }
    }
  XCQ_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_sph), tbl, sizeof(tbl), 0,
                                 hipMemcpyHostToDevice, s));
  return XCQ_SUCCESS;
}

// bfloat16 (round to nearest even) of the part of v that its FP32 rounding
// drops: with it the stored value plane carries ~32 significant bits.
__device__ __forceinline__ unsigned short resid_bf16(double v) {
  const float hi = (float)v;
  return __builtin_bit_cast(unsigned short, (__bf16)(float)(v - (double)hi));
}

// Values (and gradients) of one shell at one point, written to the scratch.
// T is the storage type (double or float); arithmetic is in double for the
// FP64 kernel and float for the FP32 kernel (type C).  With FP32 storage the
// value plane also gets its bfloat16 residual (lo, same indexing as out).
template <int L, bool GRAD, typename C, typename T>
__device__ __forceinline__ void shell_eval(C x, C y, C z, C R, C dR, bool pure, bool valid,
                                           T* out, unsigned short* lo, int64_t ld,
                                           int64_t plane) {
  constexpr int NC = (L + 1) * (L + 2) / 2;
  C xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = C(1);
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  C cv[NC], cx[NC], cy[NC], cz[NC];
  const C Rx = dR * x, Ry = dR * y, Rz = dR * z;
  int ic = 0;
#pragma unroll
  for (int a = L; a >= 0; --a) {
#pragma unroll
    for (int b = L - a; b >= 0; --b) {
      const int c = L - a - b;
      const C mono = xp[a] * yp[b] * zp[c];
      cv[ic] = mono * R;
      if (GRAD) {
        const C dxm = a ? C(a) * xp[a > 0 ? a - 1 : 0] * yp[b] * zp[c] : C(0);
        const C dym = b ? C(b) * xp[a] * yp[b > 0 ? b - 1 : 0] * zp[c] : C(0);
        const C dzm = c ? C(c) * xp[a] * yp[b] * zp[c > 0 ? c - 1 : 0] : C(0);
        cx[ic] = dxm * R + mono * Rx;
        cy[ic] = dym * R + mono * Ry;
        cz[ic] = dzm * R + mono * Rz;
      }
      ++ic;
    }
  }
  if (pure) {
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      C v = 0, gx = 0, gy = 0, gz = 0;
#pragma unroll
      for (int k = 0; k < NC; ++k) {
        const C cf = C(c_sph[L][m][k]);
        v += cf * cv[k];
        if (GRAD) {
          gx += cf * cx[k];
          gy += cf * cy[k];
          gz += cf * cz[k];
        }
      }
      out[m * ld] = valid ? T(v) : T(0);
      if constexpr (sizeof(T) == 4) lo[m * ld] = valid ? resid_bf16(v) : (unsigned short)0;
      if (GRAD) {
        out[plane + m * ld] = valid ? T(gx) : T(0);
        out[2 * plane + m * ld] = valid ? T(gy) : T(0);
        out[3 * plane + m * ld] = valid ? T(gz) : T(0);
      }
    }
  } else {
    // Cartesian functions in CCA order.
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      out[k * ld] = valid ? T(cv[k]) : T(0);
      if constexpr (sizeof(T) == 4) lo[k * ld] = valid ? resid_bf16(cv[k]) : (unsigned short)0;
      if (GRAD) {
        out[plane + k * ld] = valid ? T(cx[k]) : T(0);
        out[2 * plane + k * ld] = valid ? T(cy[k]) : T(0);
        out[3 * plane + k * ld] = valid ? T(cz[k]) : T(0);
      }
    }
  }
}

template <bool GRAD, typename C, typename T>
__global__ __launch_bounds__(256) void k_collocation(const WorkItem* __restrict__ blocks,
                                                     const DevTask* __restrict__ tasks,
                                                     const DevShell* __restrict__ shells,
                                                     const int32_t* __restrict__ shell_list,
                                                     const int32_t* __restrict__ lsh_first,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     T* __restrict__ scratch) {
  const WorkItem wi = blocks[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = wi.a * kPtBlock + lane;
  const bool valid = p < t.npts;
  double gx = 0., gy = 0., gz = 0.;
  if (valid) {
    gx = px[t.pt_off + p];
    gy = py[t.pt_off + p];
    gz = pz[t.pt_off + p];
  }
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  const int64_t boff = sizeof(T) == 4 ? t.b_off_mx : t.b_off;
  T* base = scratch + boff + p;
  // FP32 storage: bfloat16 residual plane of the values after the five
  // FP32 planes (B, dB/dx, dB/dy, dB/dz, Z).
  unsigned short* lo_base = reinterpret_cast<unsigned short*>(scratch + boff + 5 * plane) + p;
  for (int s = wave; s < t.nshells; s += 4) {
    const int ish = shell_list[t.sh_off + s];
    const int bf0 = lsh_first[t.lsh_off + s];
    const DevShell* sh = shells + ish;
    const int l = sh->l;
    const int nprim = sh->nprim;
    const bool pure = sh->pure != 0;
    // Relative coordinates in double, then evaluated in C.
    const C x = C(gx - sh->O[0]), y = C(gy - sh->O[1]), z = C(gz - sh->O[2]);
    const C r2 = x * x + y * y + z * z;
    C R = 0, dR = 0;
    for (int k = 0; k < nprim; ++k) {
      C ex;
      if constexpr (sizeof(C) == 4) ex = expf(-C(sh->alpha[k]) * r2);
      else ex = exp(-C(sh->alpha[k]) * r2);
      const C e = C(sh->coeff[k]) * ex;
      R += e;
      dR += C(-2) * C(sh->alpha[k]) * e;
    }
    T* out = base + (int64_t)bf0 * ld;
    unsigned short* lo = lo_base + (int64_t)bf0 * ld;
    switch (l) {
      case 0: shell_eval<0, GRAD, C, T>(x, y, z, R, dR, pure, valid, out, lo, ld, plane); break;
      case 1: shell_eval<1, GRAD, C, T>(x, y, z, R, dR, pure, valid, out, lo, ld, plane); break;
      case 2: shell_eval<2, GRAD, C, T>(x, y, z, R, dR, pure, valid, out, lo, ld, plane); break;
      case 3: shell_eval<3, GRAD, C, T>(x, y, z, R, dR, pure, valid, out, lo, ld, plane); break;
      default: shell_eval<4, GRAD, C, T>(x, y, z, R, dR, pure, valid, out, lo, ld, plane); break;
    }
  }
}

int launch_collocation_fp64(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz,
                            double* scratch, int deriv) {
  if (n == 0) return XCQ_SUCCESS;
  if (deriv)
    hipLaunchKernelGGL((k_collocation<true, double, double>), dim3((unsigned)n), dim3(256), 0, s,
                       items, tasks, shells, shell_list, lsh_first, px, py, pz, scratch);
  else
    hipLaunchKernelGGL((k_collocation<false, double, double>), dim3((unsigned)n), dim3(256), 0,
                       s, items, tasks, shells, shell_list, lsh_first, px, py, pz, scratch);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Mixed precision: values evaluated in FP64, stored as FP32 (half the HBM
// traffic of the B planes; storage rounding is the only FP32 step).
int launch_collocation_fp32(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz,
                            float* scratch, int deriv) {
  if (n == 0) return XCQ_SUCCESS;
  if (deriv)
    hipLaunchKernelGGL((k_collocation<true, double, float>), dim3((unsigned)n), dim3(256), 0, s,
                       items, tasks, shells, shell_list, lsh_first, px, py, pz, scratch);
  else
    hipLaunchKernelGGL((k_collocation<false, double, float>), dim3((unsigned)n), dim3(256), 0,
                       s, items, tasks, shells, shell_list, lsh_first, px, py, pz, scratch);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
