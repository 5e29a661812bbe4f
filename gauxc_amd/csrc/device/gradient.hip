// EXC nuclear gradient, RKS and UKS, FP64 (GauXC eval_exc_grad with
// include_weight_derivatives: reference_replicated_xc_host_integrator_
// exc_grad.hpp:106-600, weights.cxx:806-990; device counterparts
// increment_exc_grad.cu and cuda_ssf_1d.cu).
//
// GauXC contracts, per task, point and function mu on an atom other than the
// task's parent,
//   g_mu = w vrho X_mu grad B_mu
//        + 2 w vgamma [ X_mu (H_mu grad rho) + grad B_mu (grad rho . Xg_mu) ]
// with X = 2 P_sub B, Xg_j = 2 P_sub dB/dj and H_mu the Hessian of B_mu,
// and adds -2 g to the function's atom and +2 g to the parent (the grid
// moves with the parent).  Here the first and third terms are one product:
//   w vrho X + 2 w vgamma grad rho . Xg = 2 P_sub Za,
//   Za = w vrho B + 2 w vgamma grad rho . grad B,
// so a task needs two plane products (X = 2 P_sub B for the Hessian term of
// GGAs, Y = 2 P_sub Za) instead of four, and
//   g_mu = grad B_mu . Y_mu + X_mu (H_mu c),  c = 2 w vgamma grad rho.
// grad B and H c are evaluated where they are used (k_grad_contract), so the
// chunk scratch's five planes suffice: B and grad B (collocation), X stored
// by the density stage itself into plane 4 (GGA), Za over grad_z B in
// plane 3 once the coefficients exist, then Y = 2 P_sub Za into plane 2.
// The grid-weight part contracts the SSF weight derivatives with w eps rho
// per point (k_ssf_grad).
//
// UKS (GauXC eval_exc_grad(Ps, Pz), the same driver with spin_dim 2): with
// X_N = P_s B, X_Z = P_z B (factor 1), rho_+- = (n +- m)/2 and the polarised
// functional,
//   g_mu = grad B_mu . Y_mu + X_N,mu (H_mu c_s) + X_Z,mu (H_mu c_z),
//   Y = P_s Za_s + P_z Za_z,
//   Za_s = w (vrho_a + vrho_b)/2 B + c_s . grad B,
//   Za_z = w (vrho_a - vrho_b)/2 B + c_z . grad B,
//   c_s = g1 grad n + g2 grad m,  c_z = g2 grad n + g3 grad m,
// g1 = (v_aa + v_ab + v_bb)/2, g2 = (v_aa - v_bb)/2, g3 = (v_aa - v_ab + v_bb)/2
// (w vsigma folded in), and f = w eps (rho_a + rho_b) for the weight part.
// Planes: Za_s -> 4 and Za_z -> 3 (k_grad_zpair), Y -> 1, then X_N -> 2 and
// X_Z -> 3.
#include "internal.hpp"
#include "functional.hpp"
#include "shell_eval.hpp"

namespace xcq {

// ---- per-point coefficients ------------------------------------------------
// From the raw sums of the P pass (factor 2: rho = sum B X, grad rho =
// 2 sum dB X): coef[block][4][64] = (w vrho, 2 w vsigma grad rho) -- the Za
// coefficients, whose gradient part is also the Hessian term's c -- and
// wf[block][64] = w eps rho; the block's (w eps rho, w rho) sums to red.
__global__ __launch_bounds__(64) void k_grad_points(const WorkItem* __restrict__ items,
                                                    const DevTask* __restrict__ tasks,
                                                    const double* __restrict__ weights,
                                                    int func, int gga,
                                                    const double* __restrict__ dp,
                                                    double* __restrict__ red,
                                                    double* __restrict__ coef,
                                                    double* __restrict__ wf) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x;
  const int64_t b = t.red_off + wi.a;
  const int p = wi.a * kPtBlock + tid;
  double s[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) s[c] = dp[b * 256 + c * 64 + tid];
  double g[3], sig = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g[d] = 2.0 * s[1 + d];
    sig += g[d] * g[d];
  }
  double e = 0., n = 0., a = 0., cv[3] = {0., 0., 0.};
  if (p < t.npts) {
    const double w = weights[t.pt_off + p];
    const XCOut o = dev_eval_functional(func, s[0], sig);
    e = w * o.e;
    n = w * s[0];
    a = w * o.vr;
    if (gga) {
#pragma unroll
      for (int d = 0; d < 3; ++d) cv[d] = 2.0 * w * o.vs * g[d];
    }
  }
  coef[b * 256 + tid] = a;
#pragma unroll
  for (int d = 0; d < 3; ++d) coef[b * 256 + (1 + d) * 64 + tid] = cv[d];
  wf[b * 64 + tid] = e;
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * b] = e;
    red[2 * b + 1] = n;
  }
}

// UKS: from the raw sums of the P_s and P_z passes (factor 1) at block b,
// coef_s[b] = (w (vrho_a + vrho_b)/2, c_s), coef_z[b] = (w (vrho_a -
// vrho_b)/2, c_z) and wf[b] = w eps (rho_a + rho_b).  coef_s and coef_z may
// be the sum buffers themselves: each thread reads its own entries first.
__global__ __launch_bounds__(64) void k_grad_points_uks(const WorkItem* __restrict__ items,
                                                        const DevTask* __restrict__ tasks,
                                                        const double* __restrict__ weights,
                                                        int func, int gga, double* ds,
                                                        double* dz, double* __restrict__ red,
                                                        double* __restrict__ wf) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x;
  const int64_t b = t.red_off + wi.a;
  const int p = wi.a * kPtBlock + tid;
  double sv[4], zv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    sv[c] = ds[b * 256 + c * 64 + tid];
    zv[c] = dz[b * 256 + c * 64 + tid];
  }
  const double ra = 0.5 * (sv[0] + zv[0]), rb = 0.5 * (sv[0] - zv[0]);
  double gn[3], gm[3], nn = 0., mm = 0., nm = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    gn[d] = 2.0 * sv[1 + d];
    gm[d] = 2.0 * zv[1 + d];
    nn += gn[d] * gn[d];
    mm += gm[d] * gm[d];
    nm += gn[d] * gm[d];
  }
  double e = 0., n = 0., as = 0., az = 0., cs[3] = {0., 0., 0.}, cz[3] = {0., 0., 0.};
  if (p < t.npts) {
    const double w = weights[t.pt_off + p];
    const double saa = gga ? 0.25 * (nn + mm) + 0.5 * nm : 0.;
    const double sab = gga ? 0.25 * (nn - mm) : 0.;
    const double sbb = gga ? 0.25 * (nn + mm) - 0.5 * nm : 0.;
    const XCOutPol o = dev_eval_functional_pol(func, ra, rb, saa, sab, sbb);
    e = w * o.e;
    n = w * (ra + rb);
    as = 0.5 * w * (o.vra + o.vrb);
    az = 0.5 * w * (o.vra - o.vrb);
    if (gga) {
      const double vpp = w * o.vsaa, vpm = w * o.vsab, vmm = w * o.vsbb;
      const double g1 = 0.5 * (vpp + vpm + vmm), g2 = 0.5 * (vpp - vmm),
                   g3 = 0.5 * (vpp - vpm + vmm);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        cs[d] = g1 * gn[d] + g2 * gm[d];
        cz[d] = g2 * gn[d] + g3 * gm[d];
      }
    }
  }
  ds[b * 256 + tid] = as;
  dz[b * 256 + tid] = az;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    ds[b * 256 + (1 + d) * 64 + tid] = cs[d];
    dz[b * 256 + (1 + d) * 64 + tid] = cz[d];
  }
  wf[b * 64 + tid] = e;
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * b] = e;
    red[2 * b + 1] = n;
  }
}

// Za_s = a_s B + c_s . grad B into plane pl_s and, when coef_z is given,
// Za_z = a_z B + c_z . grad B into plane pl_z (planes 3 or 4: each thread
// reads its element's B and grad B before it writes, so plane 3 may be
// overwritten in place) for one (task, 64-point block): lane = point, the
// four waves take rows w, w + 4, ...
template <bool GGA>
__global__ __launch_bounds__(256) void k_grad_zpair(const WorkItem* __restrict__ items,
                                                    const DevTask* __restrict__ tasks,
                                                    double* __restrict__ scratch,
                                                    const double* __restrict__ coef_s,
                                                    const double* __restrict__ coef_z, int pl_s,
                                                    int pl_z) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t b = t.red_off + wi.a;
  const bool two = coef_z != nullptr;
  double cs[4], cz[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    cs[c] = coef_s[b * 256 + c * 64 + lane];
    cz[c] = two ? coef_z[b * 256 + c * 64 + lane] : 0.;
  }
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  double* B = scratch + t.b_off + (int64_t)wi.a * kPtBlock + lane;
  for (int j = wave; j < t.nbe; j += 4) {
    const int64_t o = (int64_t)j * ld;
    const double bv = B[o];
    double zs = cs[0] * bv, zz = cz[0] * bv;
    if constexpr (GGA) {
      const double x = B[plane + o], y = B[2 * plane + o], z = B[3 * plane + o];
      zs += cs[1] * x + cs[2] * y + cs[3] * z;
      zz += cz[1] * x + cz[2] * y + cz[3] * z;
    }
    B[pl_s * plane + o] = zs;
    if (two) B[pl_z * plane + o] = zz;
  }
}

// ---- plane products ----------------------------------------------------------
// out[f][p] = fac sum_g P_sub[f][g] in[g][p] (+ out[f][p] with accumulate)
// for one (task, 64-point block), every 64-function tile of the task, on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64; lane l holds A[l%16][l/16] and
// B[l/16][l%16], accumulator register r of lane l is D[4r + l/16][l%16]).
// 256 threads = four waves in a 2 x 2 arrangement, each a 32 x 32 sub-tile
// (2 x 2 MFMA blocks); K steps of 16 staged in LDS.  P_sub is gathered
// through the basis map (symmetric, so the storage order of P does not
// matter).  Rows past nbe are neither read nor written.  The LDS stage is
// double-buffered: the next K step's operands are loaded into registers
// before the current step's matrix instructions and stored after them, one
// barrier per step.
using v4d = double __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_plane_product(const WorkItem* __restrict__ items,
                                                       const DevTask* __restrict__ tasks,
                                                       const int32_t* __restrict__ bfmap,
                                                       const double* __restrict__ P,
                                                       double* __restrict__ scratch, int in_pl,
                                                       int out_pl, double fac, int accumulate) {
  __shared__ double Ps[2][16][68];  // [buffer][g][f]
  __shared__ double Is[2][16][68];  // [buffer][g][p]
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  const int64_t m0 = (int64_t)wi.a * kPtBlock;
  const double* in = scratch + t.b_off + in_pl * plane + m0;
  double* out = scratch + t.b_off + out_pl * plane + m0;
  const int32_t* bf = bfmap + t.bf_off;
  const double* Pm = P + t.p_off;
  const int lr = lane & 15, lk = lane >> 4;
  // Thread tid moves elements (k, i) = ((tid + 256 q) >> 6, (tid + 256 q) & 63).
  const int ci = tid & 63, ck = tid >> 6;
  for (int f0 = 0; f0 < t.nbe; f0 += 64) {
    v4d acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b) acc[a][b] = v4d{0., 0., 0., 0.};
    const bool fok = f0 + ci < t.nbe;
    const int64_t pcol = fok ? bf[f0 + ci] : 0;
    double rp[4], ri[4];
    auto fetch = [&](int g0) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int g = g0 + ck + 4 * q;
        rp[q] = (fok && g < t.nbe) ? Pm[pcol + (int64_t)bf[g] * t.ldp] : 0.;
        ri[q] = g < t.nbe ? in[(int64_t)g * ld + ci] : 0.;
      }
    };
    auto put = [&](int buf) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        Ps[buf][ck + 4 * q][ci] = rp[q];
        Is[buf][ck + 4 * q][ci] = ri[q];
      }
    };
    __syncthreads();  // the previous tile's reads of both buffers are done
    fetch(0);
    put(0);
    int buf = 0;
    for (int g0 = 0; g0 < t.nbe; g0 += 16, buf ^= 1) {
      __syncthreads();  // stage buf is written; stage buf ^ 1 is no longer read
      const bool more = g0 + 16 < t.nbe;
      if (more) fetch(g0 + 16);
#pragma unroll
      for (int kk = 0; kk < 4; ++kk) {
        const int k = 4 * kk + lk;
        double a[2], b[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          a[q] = Ps[buf][k][wm * 32 + q * 16 + lr];
          b[q] = Is[buf][k][wn * 32 + q * 16 + lr];
        }
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
            acc[bi][bj] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b[bj], acc[bi][bj], 0, 0, 0);
      }
      if (more) put(buf ^ 1);
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int f = f0 + wm * 32 + bi * 16 + 4 * r + lk;
        if (f < t.nbe) {
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) {
            double* o = out + (int64_t)f * ld + wn * 32 + bj * 16 + lr;
            *o = accumulate ? fma(fac, acc[bi][bj][r], *o) : fac * acc[bi][bj][r];
          }
        }
      }
  }
}

// ---- gradient and Hessian-vector product of a shell's functions ------------
// For B = R(r^2) m(r) with R1 = -2 sum a c e^{-a r^2} and R2 = 4 sum a^2 c
// e^{-a r^2} (so d_i R = x_i R1, d_i d_j R = delta_ij R1 + x_i x_j R2):
//   d_i B = d_i m R + m x_i R1,
//   (H c)_i = R d_i(c . grad m) + R1 (d_i m (r . c) + x_i (c . grad m))
//           + m (c_i R1 + x_i (r . c) R2).
// Calls sink(f, gx, gy, gz, hx, hy, hz) per function (CCA order; spherical
// outputs one at a time from their nonzero Cartesian terms).
template <int L, typename Sink>
__device__ __forceinline__ void shell_grad_hc(double x, double y, double z, double R, double R1,
                                              double R2, double cx, double cy, double cz,
                                              bool pure, Sink&& sink) {
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.;
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  const double rc = x * cx + y * cy + z * cz;
  // x^n and its first and second derivatives (0 where the power is negative).
  auto pw = [](const double* t, int n) { return n < 0 ? 0. : t[n < 0 ? 0 : n]; };
  auto term = [&](int a, int b, double* o) {
    const int c = L - a - b;
    const double m = xp[a] * yp[b] * zp[c];
    const double mx = a * pw(xp, a - 1) * yp[b] * zp[c];
    const double my = b * xp[a] * pw(yp, b - 1) * zp[c];
    const double mz = c * xp[a] * yp[b] * pw(zp, c - 1);
    const double mxx = a * (a - 1) * pw(xp, a - 2) * yp[b] * zp[c];
    const double myy = b * (b - 1) * xp[a] * pw(yp, b - 2) * zp[c];
    const double mzz = c * (c - 1) * xp[a] * yp[b] * pw(zp, c - 2);
    const double mxy = a * b * pw(xp, a - 1) * pw(yp, b - 1) * zp[c];
    const double mxz = a * c * pw(xp, a - 1) * yp[b] * pw(zp, c - 1);
    const double myz = b * c * xp[a] * pw(yp, b - 1) * pw(zp, c - 1);
    const double cm = cx * mx + cy * my + cz * mz;
    o[0] = mx * R + m * x * R1;
    o[1] = my * R + m * y * R1;
    o[2] = mz * R + m * z * R1;
    o[3] = R * (cx * mxx + cy * mxy + cz * mxz) + R1 * (mx * rc + x * cm) + m * (cx * R1 + x * rc * R2);
    o[4] = R * (cx * mxy + cy * myy + cz * myz) + R1 * (my * rc + y * cm) + m * (cy * R1 + y * rc * R2);
    o[5] = R * (cx * mxz + cy * myz + cz * mzz) + R1 * (mz * rc + z * cm) + m * (cz * R1 + z * rc * R2);
  };
  if (pure) {
#pragma unroll
    for (int mo = 0; mo < 2 * L + 1; ++mo) {
      double acc[6] = {0., 0., 0., 0., 0., 0.};
      int ic = 0;
#pragma unroll
      for (int a = L; a >= 0; --a) {
#pragma unroll
        for (int b = L - a; b >= 0; --b) {
          const double cf = kSph.c[L][mo][ic];
          if (cf != 0.) {
            double o[6];
            term(a, b, o);
#pragma unroll
            for (int q = 0; q < 6; ++q) acc[q] += cf * o[q];
          }
          ++ic;
        }
      }
      sink(mo, acc[0], acc[1], acc[2], acc[3], acc[4], acc[5]);
    }
  } else {
    int ic = 0;
#pragma unroll
    for (int a = L; a >= 0; --a) {
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        double o[6];
        term(a, b, o);
        sink(ic, o[0], o[1], o[2], o[3], o[4], o[5]);
        ++ic;
      }
    }
  }
}

// Gradient and full Hessian of a shell's functions:
//   d_i d_j B = R d_i d_j m + R1 (d_i m x_j + d_j m x_i) + m (delta_ij R1 + x_i x_j R2).
// Calls sink(f, gx, gy, gz, h) with h = (xx, xy, xz, yy, yz, zz).
template <int L, typename Sink>
__device__ __forceinline__ void shell_grad_hess(double x, double y, double z, double R,
                                                double R1, double R2, bool pure, Sink&& sink) {
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.;
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  auto pw = [](const double* t, int n) { return n < 0 ? 0. : t[n < 0 ? 0 : n]; };
  auto term = [&](int a, int b, double* o) {
    const int c = L - a - b;
    const double m = xp[a] * yp[b] * zp[c];
    const double mx = a * pw(xp, a - 1) * yp[b] * zp[c];
    const double my = b * xp[a] * pw(yp, b - 1) * zp[c];
    const double mz = c * xp[a] * yp[b] * pw(zp, c - 1);
    const double mxx = a * (a - 1) * pw(xp, a - 2) * yp[b] * zp[c];
    const double myy = b * (b - 1) * xp[a] * pw(yp, b - 2) * zp[c];
    const double mzz = c * (c - 1) * xp[a] * yp[b] * pw(zp, c - 2);
    const double mxy = a * b * pw(xp, a - 1) * pw(yp, b - 1) * zp[c];
    const double mxz = a * c * pw(xp, a - 1) * yp[b] * pw(zp, c - 1);
    const double myz = b * c * xp[a] * pw(yp, b - 1) * pw(zp, c - 1);
    o[0] = mx * R + m * x * R1;
    o[1] = my * R + m * y * R1;
    o[2] = mz * R + m * z * R1;
    o[3] = R * mxx + 2. * R1 * mx * x + m * (R1 + x * x * R2);
    o[4] = R * mxy + R1 * (mx * y + my * x) + m * x * y * R2;
    o[5] = R * mxz + R1 * (mx * z + mz * x) + m * x * z * R2;
    o[6] = R * myy + 2. * R1 * my * y + m * (R1 + y * y * R2);
    o[7] = R * myz + R1 * (my * z + mz * y) + m * y * z * R2;
    o[8] = R * mzz + 2. * R1 * mz * z + m * (R1 + z * z * R2);
  };
  if (pure) {
#pragma unroll
    for (int mo = 0; mo < 2 * L + 1; ++mo) {
      double acc[9] = {0., 0., 0., 0., 0., 0., 0., 0., 0.};
      int ic = 0;
#pragma unroll
      for (int a = L; a >= 0; --a) {
#pragma unroll
        for (int b = L - a; b >= 0; --b) {
          const double cf = kSph.c[L][mo][ic];
          if (cf != 0.) {
            double o[9];
            term(a, b, o);
#pragma unroll
            for (int q = 0; q < 9; ++q) acc[q] += cf * o[q];
          }
          ++ic;
        }
      }
      sink(mo, acc[0], acc[1], acc[2], acc + 3);
    }
  } else {
    int ic = 0;
#pragma unroll
    for (int a = L; a >= 0; --a) {
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        double o[9];
        term(a, b, o);
        sink(ic, o[0], o[1], o[2], o + 3);
        ++ic;
      }
    }
  }
}

// The highest angular momentum a contraction kernel instantiates: the
// kernels are compiled for LMAX = 3 (the register budget of f functions)
// and LMAX = 4, and the launchers pick by the basis.
template <int LMAX>
constexpr int kTopL = LMAX > 3 ? 4 : 3;

// ---- basis-function part of the gradient -----------------------------------
// One (task, 64-point block) per 256-thread workgroup; lane = point, wave w
// takes the task's shells w, w + 4, ...  Per shell (with the weight
// derivatives: on an atom other than the parent): sum over its functions and the block's points of
// grad B . Y + X (H c) (X only for GGAs), reduced over the wave; runs of
// shells on one atom are added to that atom with one set of atomics, the
// parent's share (weight derivatives only) once per wave at the end.
template <bool GGA, int LMAX>
__global__ __launch_bounds__(256, 2) void k_grad_contract(
    const WorkItem* __restrict__ items, const DevTask* __restrict__ tasks,
    const DevShell* __restrict__ shells, const int32_t* __restrict__ shell_list,
    const int32_t* __restrict__ lsh_first, const int32_t* __restrict__ shell_atom,
    const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ pz, const double* __restrict__ scratch,
    const double* __restrict__ coef, int weight_derivs, double* __restrict__ grad, int x_pl,
    int y_pl) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p = wi.a * kPtBlock + lane;
  const bool valid = p < t.npts;
  const int pc = min(p, t.npts - 1);
  const double gx = px[t.pt_off + pc], gy = py[t.pt_off + pc], gz = pz[t.pt_off + pc];
  const int64_t b = t.red_off + wi.a;
  double cx = 0., cy = 0., cz = 0.;
  if (GGA && valid) {
    cx = coef[b * 256 + 64 + lane];
    cy = coef[b * 256 + 128 + lane];
    cz = coef[b * 256 + 192 + lane];
  }
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  const int64_t m0 = (int64_t)wi.a * kPtBlock + lane;
  const double* Xp = scratch + t.b_off + x_pl * plane + m0;
  const double* Yp = scratch + t.b_off + y_pl * plane + m0;
  int cur = -1;                            // atom of the running sum (uniform)
  double run[3] = {0., 0., 0.}, par[3] = {0., 0., 0.};
  auto flush = [&]() {
    if (cur >= 0 && lane == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * cur + d, -2.0 * run[d]);
    }
  };
  for (int k = wave; k < t.nshells; k += 4) {
    const int ish = __builtin_amdgcn_readfirstlane(shell_list[t.sh_off + k]);
    const int at = __builtin_amdgcn_readfirstlane(shell_atom[ish]);
    if (at < 0 || (weight_derivs && at == t.iparent)) continue;
    const int bf0 = __builtin_amdgcn_readfirstlane(lsh_first[t.lsh_off + k]);
    const DevShell* sh = shells + ish;
    const int l = sh->l, nprim = sh->nprim;
    const bool pure = sh->pure != 0;
    const double x = gx - sh->O[0], y = gy - sh->O[1], z = gz - sh->O[2];
    const double r2 = x * x + y * y + z * z;
    double R = 0., R1 = 0., R2 = 0.;
    for (int q = 0; q < nprim; ++q) {
      const double a = sh->alpha[q], e = sh->coeff[q] * exp(-a * r2);
      R += e;
      R1 -= 2. * a * e;
      R2 += 4. * a * a * e;
    }
    double s[3] = {0., 0., 0.};
    auto put = [&](int f, double bx, double by, double bz, double hx, double hy, double hz) {
      const int64_t r = (int64_t)(bf0 + f) * ld;
      const double Y = valid ? Yp[r] : 0.;
      s[0] = fma(bx, Y, s[0]);
      s[1] = fma(by, Y, s[1]);
      s[2] = fma(bz, Y, s[2]);
      if constexpr (GGA) {
        const double X = valid ? Xp[r] : 0.;
        s[0] = fma(X, hx, s[0]);
        s[1] = fma(X, hy, s[1]);
        s[2] = fma(X, hz, s[2]);
      }
    };
    if (l == 0) shell_grad_hc<0>(x, y, z, R, R1, R2, cx, cy, cz, pure, put);
    else if (l == 1) shell_grad_hc<1>(x, y, z, R, R1, R2, cx, cy, cz, pure, put);
    else if (l == 2) shell_grad_hc<2>(x, y, z, R, R1, R2, cx, cy, cz, pure, put);
    else if (l == 3) shell_grad_hc<3>(x, y, z, R, R1, R2, cx, cy, cz, pure, put);
    else shell_grad_hc<kTopL<LMAX>>(x, y, z, R, R1, R2, cx, cy, cz, pure, put);
#pragma unroll
    for (int d = 0; d < 3; ++d)
      for (int off = 32; off > 0; off >>= 1) s[d] += __shfl_xor(s[d], off);
    if (at != cur) {
      flush();
      cur = at;
#pragma unroll
      for (int d = 0; d < 3; ++d) run[d] = 0.;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      run[d] += s[d];
      par[d] += s[d];
    }
  }
  flush();
  if (weight_derivs && lane == 0 && cur >= 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * t.iparent + d, 2.0 * par[d]);
  }
}

// UKS basis-function part: as k_grad_contract with Y in plane 1, X_N in
// plane 2 and X_Z in plane 3; per function grad B . Y + H (X_N c_s + X_Z c_z)
// (the Hessian term for GGAs only).
template <bool GGA, int LMAX>
__global__ __launch_bounds__(256, 2) void k_grad_contract_uks(
    const WorkItem* __restrict__ items, const DevTask* __restrict__ tasks,
    const DevShell* __restrict__ shells, const int32_t* __restrict__ shell_list,
    const int32_t* __restrict__ lsh_first, const int32_t* __restrict__ shell_atom,
    const double* __restrict__ px, const double* __restrict__ py,
    const double* __restrict__ pz, const double* __restrict__ scratch,
    const double* __restrict__ coef_s, const double* __restrict__ coef_z, int weight_derivs,
    double* __restrict__ grad) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int p = wi.a * kPtBlock + lane;
  const bool valid = p < t.npts;
  const int pc = min(p, t.npts - 1);
  const double gx = px[t.pt_off + pc], gy = py[t.pt_off + pc], gz = pz[t.pt_off + pc];
  const int64_t b = t.red_off + wi.a;
  double cs[3] = {0., 0., 0.}, cz[3] = {0., 0., 0.};
  if (GGA && valid) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      cs[d] = coef_s[b * 256 + (1 + d) * 64 + lane];
      cz[d] = coef_z[b * 256 + (1 + d) * 64 + lane];
    }
  }
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  const int64_t m0 = (int64_t)wi.a * kPtBlock + lane;
  const double* Yp = scratch + t.b_off + plane + m0;
  const double* XNp = scratch + t.b_off + 2 * plane + m0;
  const double* XZp = scratch + t.b_off + 3 * plane + m0;
  int cur = -1;
  double run[3] = {0., 0., 0.}, par[3] = {0., 0., 0.};
  auto flush = [&]() {
    if (cur >= 0 && lane == 0) {
#pragma unroll
      for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * cur + d, -2.0 * run[d]);
    }
  };
  for (int k = wave; k < t.nshells; k += 4) {
    const int ish = __builtin_amdgcn_readfirstlane(shell_list[t.sh_off + k]);
    const int at = __builtin_amdgcn_readfirstlane(shell_atom[ish]);
    if (at < 0 || (weight_derivs && at == t.iparent)) continue;
    const int bf0 = __builtin_amdgcn_readfirstlane(lsh_first[t.lsh_off + k]);
    const DevShell* sh = shells + ish;
    const int l = sh->l, nprim = sh->nprim;
    const bool pure = sh->pure != 0;
    const double x = gx - sh->O[0], y = gy - sh->O[1], z = gz - sh->O[2];
    const double r2 = x * x + y * y + z * z;
    double R = 0., R1 = 0., R2 = 0.;
    for (int q = 0; q < nprim; ++q) {
      const double a = sh->alpha[q], e = sh->coeff[q] * exp(-a * r2);
      R += e;
      R1 -= 2. * a * e;
      R2 += 4. * a * a * e;
    }
    double s[3] = {0., 0., 0.};
    auto put = [&](int f, double bx, double by, double bz, const double* h) {
      const int64_t r = (int64_t)(bf0 + f) * ld;
      const double Y = valid ? Yp[r] : 0.;
      s[0] = fma(bx, Y, s[0]);
      s[1] = fma(by, Y, s[1]);
      s[2] = fma(bz, Y, s[2]);
      if constexpr (GGA) {
        const double xn = valid ? XNp[r] : 0., xz = valid ? XZp[r] : 0.;
        const double vx = xn * cs[0] + xz * cz[0], vy = xn * cs[1] + xz * cz[1],
                     vz = xn * cs[2] + xz * cz[2];
        s[0] += h[0] * vx + h[1] * vy + h[2] * vz;
        s[1] += h[1] * vx + h[3] * vy + h[4] * vz;
        s[2] += h[2] * vx + h[4] * vy + h[5] * vz;
      }
    };
    if constexpr (GGA) {
      if (l == 0) shell_grad_hess<0>(x, y, z, R, R1, R2, pure, put);
      else if (l == 1) shell_grad_hess<1>(x, y, z, R, R1, R2, pure, put);
      else if (l == 2) shell_grad_hess<2>(x, y, z, R, R1, R2, pure, put);
      else if (l == 3) shell_grad_hess<3>(x, y, z, R, R1, R2, pure, put);
      else shell_grad_hess<kTopL<LMAX>>(x, y, z, R, R1, R2, pure, put);
    } else {
      // LDA needs only grad B: the Hessian-vector form with c = 0.
      auto put3 = [&](int f, double bx, double by, double bz, double, double, double) {
        const double h0[6] = {0., 0., 0., 0., 0., 0.};
        put(f, bx, by, bz, h0);
      };
      if (l == 0) shell_grad_hc<0>(x, y, z, R, R1, R2, 0., 0., 0., pure, put3);
      else if (l == 1) shell_grad_hc<1>(x, y, z, R, R1, R2, 0., 0., 0., pure, put3);
      else if (l == 2) shell_grad_hc<2>(x, y, z, R, R1, R2, 0., 0., 0., pure, put3);
      else if (l == 3) shell_grad_hc<3>(x, y, z, R, R1, R2, 0., 0., 0., pure, put3);
      else shell_grad_hc<kTopL<LMAX>>(x, y, z, R, R1, R2, 0., 0., 0., pure, put3);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d)
      for (int off = 32; off > 0; off >>= 1) s[d] += __shfl_xor(s[d], off);
    if (at != cur) {
      flush();
      cur = at;
#pragma unroll
      for (int d = 0; d < 3; ++d) run[d] = 0.;
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      run[d] += s[d];
      par[d] += s[d];
    }
  }
  flush();
  if (weight_derivs && lane == 0 && cur >= 0) {
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * t.iparent + d, 2.0 * par[d]);
  }
}

// ---- grid-weight part ---------------------------------------------------------
// Derivatives of the SSF weights w_i = w0_i P_A(r_i) / sum_C P_C(r_i)
// contracted with f_i = w_i eps_i rho_i, the points moving with their parent
// atom A (weights.cxx:806-990).  P_B = prod_{C != B} s(mu_BC), s(mu) = 1
// (mu <= -a), 0 (mu >= a), (1 - g(mu / a)) / 2 otherwise (g the Frisch
// polynomial, a = 0.64), formed here as that product; its logarithmic
// derivative t(mu) = s'(mu) / s(mu) in the factored form
//   1 - g(v) = (1 - v)^4 (5 v^3 + 20 v^2 + 29 v + 16) / 16,  g'(v) = 35 (1 - v^2)^3 / 16,
//   t(mu) = 35 (1 + v)^3 / ((mu - a)(5 v^3 + 20 v^2 + 29 v + 16)),  v = mu / a,
// which stays accurate up to the |mu| < a - 1e-4 bound the contraction uses.
// One thread per point (lane of a 64-point block); distances are recomputed
// where needed.  Without neighbour lists every loop runs over all atoms,
// O(natoms^2) per point.  With them (nbr_idx / nbr_r: row A holds every atom
// in order of increasing R_AB, A itself first) the loops stop where the
// triangle inequality makes every remaining term vanish, as in the weights
// kernel: a factor s(mu_BC) differs from 1, and |mu_BC| can fall below a,
// only for R_BC < 2 d_B / (1 - a); P_B > 0 only for d_B <= (1 + a)/(1 - a)
// d_min; so the parent's list is read up to max(2 d_A / (1 - a),
// d_A + that bound) and each B's list up to 2 d_B / (1 - a), all widened by
// a relative 1e-10.  The terms skipped are exactly zero (factors exactly 1).
constexpr double kSsfA = 0.64;
constexpr double kSsfTol = 1e-13;

__device__ __forceinline__ double ssf_s(double mu) {
  if (mu <= -kSsfA) return 1.;
  if (mu >= kSsfA) return 0.;
  const double v = mu / kSsfA, v2 = v * v;
  const double g = v * (35. + v2 * (-35. + v2 * (21. - 5. * v2))) / 16.;
  return 0.5 * (1. - g);
}
__device__ __forceinline__ double ssf_t(double mu) {
  const double v = mu / kSsfA;
  return 35. * (1. + v) * (1. + v) * (1. + v) /
         ((mu - kSsfA) * (((5. * v + 20.) * v + 29.) * v + 16.));
}

__global__ __launch_bounds__(64) void k_ssf_grad(const WorkItem* __restrict__ items,
                                                 const DevTask* __restrict__ tasks,
                                                 const double* __restrict__ px,
                                                 const double* __restrict__ py,
                                                 const double* __restrict__ pz, int natoms,
                                                 const double* __restrict__ xyz,
                                                 const double* __restrict__ dnear,
                                                 const double* __restrict__ wf,
                                                 double* __restrict__ grad,
                                                 const int32_t* __restrict__ nbr_idx,
                                                 const double* __restrict__ nbr_r) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int lane = threadIdx.x;
  const int p = wi.a * kPtBlock + lane;
  if (p >= t.npts) return;
  const double f = wf[(t.red_off + wi.a) * 64 + lane];
  if (fabs(f) < 1e-12) return;
  const double r[3] = {px[t.pt_off + p], py[t.pt_off + p], pz[t.pt_off + p]};
  auto X = [&](int at, int d) { return xyz[3 * at + d]; };
  auto dist = [&](int at) {
    const double dx = r[0] - X(at, 0), dy = r[1] - X(at, 1), dz = r[2] - X(at, 2);
    return sqrt(dx * dx + dy * dy + dz * dz);
  };
  auto rab = [&](int a, int b) {
    const double dx = X(a, 0) - X(b, 0), dy = X(a, 1) - X(b, 1), dz = X(a, 2) - X(b, 2);
    return sqrt(dx * dx + dy * dy + dz * dz);
  };
  const int A = t.iparent;
  const double dA = dist(A);
  if (dA < 0.5 * (1. - kSsfA) * dnear[A]) return;
  const double bound = kSsfA - 1e-4;
  if (nbr_idx) {
    constexpr double kWiden = 1.0 + 1e-10;
    const int32_t* an = nbr_idx + (size_t)A * natoms;
    const double* ar = nbr_r + (size_t)A * natoms;
    // Nearest atom (R_AB* <= 2 d_A) and the candidate bound.
    double dmin = dA;
    for (int k = 1; k < natoms && ar[k] <= 2.0 * dA * kWiden; ++k) dmin = fmin(dmin, dist(an[k]));
    const double cbound = (1.0 + kSsfA) / (1.0 - kSsfA) * dmin * kWiden + 1e-12;
    const double rmax = fmax(2.0 * dA / (1.0 - kSsfA), dA + cbound) * kWiden;
    // P_B over B's list (factors beyond 2 d_B / (1 - a) are exactly 1).
    auto part_l = [&](int B, double dB) {
      if (dB > cbound) return 0.;
      const int32_t* bn = nbr_idx + (size_t)B * natoms;
      const double* br = nbr_r + (size_t)B * natoms;
      const double rf = 2.0 * dB / (1.0 - kSsfA) * kWiden;
      double v = 1.;
      for (int q = 1; q < natoms && br[q] < rf && v > 0.; ++q)
        v *= ssf_s((dB - dist(bn[q])) / br[q]);
      return v;
    };
    double sum = 0.;
    for (int k = 0; k < natoms && ar[k] <= rmax; ++k) {
      const int B = an[k];
      sum += part_l(B, k == 0 ? dA : dist(B));
    }
    const double pA = part_l(A, dA);
    double gA[3] = {0., 0., 0.};
    for (int k = 1; k < natoms && ar[k] <= rmax; ++k) {
      const int B = an[k];
      const double dB = dist(B);
      double gb[3] = {0., 0., 0.};
      const double rAB = ar[k];
      const double muAB = (dA - dB) / rAB;
      if (fabs(muAB) < bound) {
        const double c1 = ssf_t(muAB) / rAB * (pA - sum) / sum * f / dB;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          gb[d] = c1 * ((X(B, d) - r[d]) + muAB * (X(B, d) - X(A, d)) / rAB * dB);
      }
      const double pB = part_l(B, dB);
      if (pB > kSsfTol) {
        const int32_t* bn = nbr_idx + (size_t)B * natoms;
        const double* br = nbr_r + (size_t)B * natoms;
        const double rf = 2.0 * dB / (1.0 - kSsfA) * kWiden;
        for (int q = 1; q < natoms && br[q] < rf; ++q) {
          const int C = bn[q];
          const double rBC = br[q], dC = dist(C);
          const double muBC = (dB - dC) / rBC;
          if (fabs(muBC) >= bound) continue;
          const double c = pB * ssf_t(muBC) / rBC / sum * f;
#pragma unroll
          for (int d = 0; d < 3; ++d)
            gb[d] -= c * ((X(B, d) - r[d]) / dB - muBC * (X(B, d) - X(C, d)) / rBC);
          if (C != A) {
#pragma unroll
            for (int d = 0; d < 3; ++d) {
              const double gc = c * ((X(C, d) - r[d]) / dC + muBC * (X(C, d) - X(B, d)) / rBC);
              atomicAdd(grad + 3 * C + d, gc);
              gA[d] -= gc;
            }
          }
        }
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        if (gb[d] != 0.) atomicAdd(grad + 3 * B + d, gb[d]);
        gA[d] -= gb[d];
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * A + d, gA[d]);
    return;
  }
  auto part = [&](int B, double dB) {
    double v = 1.;
    for (int C = 0; C < natoms && v > 0.; ++C)
      if (C != B) v *= ssf_s((dB - dist(C)) / rab(B, C));
    return v;
  };
  double sum = 0.;
  for (int B = 0; B < natoms; ++B) sum += part(B, B == A ? dA : dist(B));
  const double pA = part(A, dA);
  double gA[3] = {0., 0., 0.};
  for (int B = 0; B < natoms; ++B) {
    if (B == A) continue;
    const double dB = dist(B);
    double gb[3] = {0., 0., 0.};
    const double rAB = rab(A, B);
    const double muAB = (dA - dB) / rAB;
    if (fabs(muAB) < bound) {
      const double c1 = ssf_t(muAB) / rAB * (pA - sum) / sum * f / dB;
#pragma unroll
      for (int d = 0; d < 3; ++d)
        gb[d] = c1 * ((X(B, d) - r[d]) + muAB * (X(B, d) - X(A, d)) / rAB * dB);
    }
    const double pB = part(B, dB);
    if (pB > kSsfTol) {
      for (int C = 0; C < natoms; ++C) {
        if (C == B) continue;
        const double rBC = rab(B, C), dC = dist(C);
        const double muBC = (dB - dC) / rBC;
        if (fabs(muBC) >= bound) continue;
        const double c = pB * ssf_t(muBC) / rBC / sum * f;
#pragma unroll
        for (int d = 0; d < 3; ++d)
          gb[d] -= c * ((X(B, d) - r[d]) / dB - muBC * (X(B, d) - X(C, d)) / rBC);
        if (C != A) {
#pragma unroll
          for (int d = 0; d < 3; ++d) {
            const double gc = c * ((X(C, d) - r[d]) / dC + muBC * (X(C, d) - X(B, d)) / rBC);
            atomicAdd(grad + 3 * C + d, gc);
            gA[d] -= gc;
          }
        }
      }
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      if (gb[d] != 0.) atomicAdd(grad + 3 * B + d, gb[d]);
      gA[d] -= gb[d];
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) atomicAdd(grad + 3 * A + d, gA[d]);
}

// ---- launchers ------------------------------------------------------------------
int launch_grad_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* dp,
                       double* coef, double* wf) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_grad_points, dim3((unsigned)r.n), dim3(64), 0, v.stream, v.work + r.off,
                     v.tasks, v.w, func, gga, dp, v.red, coef, wf);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_grad_points_uks(const CtxView& v, const WorkRange& r, int func, int gga, double* ds,
                           double* dz, double* wf) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_grad_points_uks, dim3((unsigned)r.n), dim3(64), 0, v.stream,
                     v.work + r.off, v.tasks, v.w, func, gga, ds, dz, v.red, wf);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_grad_zpair(const CtxView& v, const WorkRange& r, int gga, const double* coef_s,
                      const double* coef_z, int pl_s, int pl_z) {
  if (r.n == 0) return XCQ_SUCCESS;
  XCQ_REQUIRE((pl_s == 3 || pl_s == 4) && (!coef_z || ((pl_z == 3 || pl_z == 4) && pl_z != pl_s)),
              XCQ_ERROR_INVALID_ARGUMENT, "Za planes must be 3 or 4");
  double* scr = static_cast<double*>(v.scratch);
  if (gga)
    hipLaunchKernelGGL(k_grad_zpair<true>, dim3((unsigned)r.n), dim3(256), 0, v.stream,
                       v.work + r.off, v.tasks, scr, coef_s, coef_z, pl_s, pl_z);
  else
    hipLaunchKernelGGL(k_grad_zpair<false>, dim3((unsigned)r.n), dim3(256), 0, v.stream,
                       v.work + r.off, v.tasks, scr, coef_s, coef_z, pl_s, pl_z);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_plane_product(const CtxView& v, const WorkRange& r, int in_pl, int out_pl,
                         const double* P, double fac, int accumulate) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_plane_product, dim3((unsigned)r.n), dim3(256), 0, v.stream,
                     v.work + r.off, v.tasks, v.bfmap, P ? P : v.P,
                     static_cast<double*>(v.scratch), in_pl, out_pl, fac, accumulate);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_grad_contract(const CtxView& v, const WorkRange& r, int gga,
                         const int32_t* shell_atom, const double* coef, int weight_derivs,
                         double* grad, int x_pl, int y_pl, int lmax) {
  if (r.n == 0) return XCQ_SUCCESS;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)r.n), dim3(256), 0, v.stream, v.work + r.off,
                       v.tasks, v.shells, v.shell_list, v.lsh, shell_atom, v.px, v.py, v.pz,
                       static_cast<const double*>(v.scratch), coef, weight_derivs, grad, x_pl,
                       y_pl);
  };
  if (lmax <= 3) gga ? go(k_grad_contract<true, 3>) : go(k_grad_contract<false, 3>);
  else gga ? go(k_grad_contract<true, 4>) : go(k_grad_contract<false, 4>);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_grad_contract_uks(const CtxView& v, const WorkRange& r, int gga,
                             const int32_t* shell_atom, const double* coef_s,
                             const double* coef_z, int weight_derivs, double* grad, int lmax) {
  if (r.n == 0) return XCQ_SUCCESS;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)r.n), dim3(256), 0, v.stream, v.work + r.off,
                       v.tasks, v.shells, v.shell_list, v.lsh, shell_atom, v.px, v.py, v.pz,
                       static_cast<const double*>(v.scratch), coef_s, coef_z, weight_derivs,
                       grad);
  };
  if (lmax <= 3) gga ? go(k_grad_contract_uks<true, 3>) : go(k_grad_contract_uks<false, 3>);
  else gga ? go(k_grad_contract_uks<true, 4>) : go(k_grad_contract_uks<false, 4>);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_ssf_grad(const CtxView& v, const WorkRange& r, int natoms, const double* xyz,
                    const double* dnear, const double* wf, double* grad,
                    const int32_t* nbr_idx, const double* nbr_r) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_ssf_grad, dim3((unsigned)r.n), dim3(64), 0, v.stream, v.work + r.off,
                     v.tasks, v.px, v.py, v.pz, natoms, xyz, dnear, wf, grad, nbr_idx, nbr_r);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
