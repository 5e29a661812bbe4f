// FP64 contractions of the XC path on gfx950 matrix cores
// (v_mfma_f64_16x16x4_f64).
//
//  k_xmat_den : X = fac B P_sub (fac = 2 for RKS EXC/VXC, 1 for the
//               integrated density) for a 64-point x 64-function tile, with the
//               density reduction fused into the epilogue: X never reaches
//               HBM.  Per point it emits the partial sums
//                 rho_p += sum_j B_pj X_pj,  g_p += sum_j dB_pj X_pj
//               of its 64 columns into a [slot][4][npts_pad] buffer (two
//               slots per column tile, one per wave column), which the
//               functional kernel sums in a fixed order.  GauXC performs
//               the same contraction as one DGEMM per task
//               (scheme1_base.cxx:1604-1664, reference host
//               reference_local_host_work_driver.cxx:123-146) followed by
//               the uvvar kernels (uvvars_gga.hpp:22-157).
//  k_zmat     : Z = a B + c . grad B (eval_zmat_gga_vxc_rks,
//               reference_local_host_work_driver.cxx:678-713).
//  k_syr2k    : V = B^T Z + Z^T B for lower-triangle 64x64 tiles with the
//               scatter-add into VXC fused into the epilogue (GauXC: DSYR2K
//               per task + sym_inc_by_submat, scheme1_base.cxx:1712-1766).
//
// Workgroup: 256 threads = 4 waves in a 2x2 arrangement, each wave owning a
// 32x32 sub-tile = 2x2 MFMA 16x16 blocks.  K advances 16 per LDS stage.
// MFMA operand layout (16x16x4 f64, measured on gfx950): lane l holds
// A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane l is
// D[4*r + l/16][l%16] (rows interleave across the four 16-lane groups,
// unlike the FP32-output MFMAs whose register r holds row 4*(l/16) + r).
#include "internal.hpp"

namespace xcq {

using v4d = double __attribute__((ext_vector_type(4)));

constexpr int kLdsPad = 68;  // row length of the [16][64] LDS stages (doubles)

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_xmat_den(const WorkItem* __restrict__ tiles,
                                                  const DevTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ bfmap,
                                                  const double* __restrict__ P, int64_t ldp,
                                                  const double* __restrict__ scratch,
                                                  double* __restrict__ part, int gga,
                                                  double fac) {
  __shared__ double As[16][kLdsPad];
  __shared__ double Bs[16][kLdsPad];
  const WorkItem wi = tiles[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = wi.a * kTile, n0 = wi.b * kTile;
  const int64_t ld = t.npts_pad;
  const int nbe = t.nbe;
  const double* Bm = scratch + t.b_off;
  const int32_t* bf = bfmap + t.bf_off;

  // Loader coordinates (fixed over the K loop).
  const int a_m = tid & 63, a_k = tid >> 6;           // A: k = a_k + 4i
  const int b_k = tid & 15, b_n = tid >> 4;           // B: n = b_n + 16i
  int64_t colP[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int n = n0 + b_n + 16 * i;
    colP[i] = n < nbe ? (int64_t)bf[n] * ldp : -1;
  }

  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0., 0., 0., 0.};

  for (int k0 = 0; k0 < nbe; k0 += 16) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int k = a_k + 4 * i;
      As[k][a_m] = (k0 + k < nbe) ? Bm[(int64_t)(k0 + k) * ld + m0 + a_m] : 0.;
    }
    const int kk = k0 + b_k;
    const int64_t rowP = kk < nbe ? (int64_t)bf[kk] : -1;
#pragma unroll
    for (int i = 0; i < 4; ++i)
      Bs[b_k][b_n + 16 * i] = (rowP >= 0 && colP[i] >= 0) ? fac * P[rowP + colP[i]] : 0.;
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      const double a0 = As[kr][wm * 32 + (lane & 15)];
      const double a1 = As[kr][wm * 32 + 16 + (lane & 15)];
      const double b0 = Bs[kr][wn * 32 + (lane & 15)];
      const double b1 = Bs[kr][wn * 32 + 16 + (lane & 15)];
      acc[0][0] = mfma_f64(a0, b0, acc[0][0]);
      acc[0][1] = mfma_f64(a0, b1, acc[0][1]);
      acc[1][0] = mfma_f64(a1, b0, acc[1][0]);
      acc[1][1] = mfma_f64(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }

  // Epilogue: per-point partial densities of this tile's columns.
  const int64_t plane = (int64_t)nbe * ld;
  double* pslot = part + t.part_off + (int64_t)(2 * wi.b + wn) * 4 * ld;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + wm * 32 + bi * 16 + 4 * r + (lane >> 4);
      double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) {
        const int n = n0 + wn * 32 + bj * 16 + (lane & 15);
        if (n < nbe) {
          const double x = acc[bi][bj][r];
          const int64_t o = (int64_t)n * ld + m;
          s0 += Bm[o] * x;
          if (gga) {
            s1 += Bm[plane + o] * x;
            s2 += Bm[2 * plane + o] * x;
            s3 += Bm[3 * plane + o] * x;
          }
        }
      }
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) {
        s0 += __shfl_xor(s0, off);
        if (gga) {
          s1 += __shfl_xor(s1, off);
          s2 += __shfl_xor(s2, off);
          s3 += __shfl_xor(s3, off);
        }
      }
      if ((lane & 15) == 0) {
        pslot[m] = s0;
        if (gga) {
          pslot[ld + m] = s1;
          pslot[2 * ld + m] = s2;
          pslot[3 * ld + m] = s3;
        }
      }
    }
  }
}

int launch_xmat_den_fp64(hipStream_t s, const Chunk& c, const WorkItem* g1, const DevTask* tasks,
                         const int32_t* bfmap, const double* P, int64_t ldp,
                         const double* scratch, double* part, int gga, double fac) {
  if (c.ng1 == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_xmat_den, dim3((unsigned)c.ng1), dim3(256), 0, s, g1 + c.g1_off, tasks,
                     bfmap, P, ldp, scratch, part, gga, fac);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_zmat(const WorkItem* __restrict__ pblk,
                                              const DevTask* __restrict__ tasks,
                                              const double* __restrict__ zcoef, int64_t nq,
                                              double* __restrict__ scratch, int gga) {
  const WorkItem wi = pblk[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int p = wi.a * kPtBlock + lane;
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  const double* zc = zcoef + t.q_off + p;
  const double a = zc[0];
  const double cx = gga ? zc[nq] : 0., cy = gga ? zc[2 * nq] : 0., cz = gga ? zc[3 * nq] : 0.;
  double* B = scratch + t.b_off + p;
  double* Z = B + 4 * plane;
  for (int j = wave; j < t.nbe; j += 4) {
    const int64_t o = (int64_t)j * ld;
    double z = a * B[o];
    if (gga) z += cx * B[plane + o] + cy * B[2 * plane + o] + cz * B[3 * plane + o];
    Z[o] = z;
  }
}

int launch_zmat_fp64(hipStream_t s, const Chunk& c, const WorkItem* pblk, const DevTask* tasks,
                     const double* zcoef, double* scratch, int gga) {
  if (c.npblk == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_zmat, dim3((unsigned)c.npblk), dim3(256), 0, s, pblk + c.pblk_off, tasks,
                     zcoef, c.nq, scratch, gga);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_syr2k(const WorkItem* __restrict__ tiles,
                                               const DevTask* __restrict__ tasks,
                                               const int32_t* __restrict__ bfmap,
                                               const double* __restrict__ scratch,
                                               double* __restrict__ vxc, int64_t ldv) {
  __shared__ double Bi[16][kLdsPad];
  __shared__ double Zi[16][kLdsPad];
  __shared__ double Bj[16][kLdsPad];
  __shared__ double Zj[16][kLdsPad];
  const WorkItem wi = tiles[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int i0 = wi.a * kTile, j0 = wi.b * kTile;
  const bool diag = wi.a == wi.b;
  const int64_t ld = t.npts_pad;
  const int nbe = t.nbe;
  const int64_t plane = (int64_t)nbe * ld;
  const double* Bm = scratch + t.b_off;
  const double* Zm = Bm + 4 * plane;

  const int l_p = tid & 15, l_f = tid >> 4;  // loader: point, function (+16q)
  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0., 0., 0., 0.};

  for (int k0 = 0; k0 < ld; k0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int f = l_f + 16 * q;
      const int fi = i0 + f, fj = j0 + f;
      const int64_t oi = (int64_t)fi * ld + k0 + l_p;
      const int64_t oj = (int64_t)fj * ld + k0 + l_p;
      Bi[l_p][f] = fi < nbe ? Bm[oi] : 0.;
      Zi[l_p][f] = fi < nbe ? Zm[oi] : 0.;
      if (!diag) {
        Bj[l_p][f] = fj < nbe ? Bm[oj] : 0.;
        Zj[l_p][f] = fj < nbe ? Zm[oj] : 0.;
      }
    }
    __syncthreads();
    const double(*BJ)[kLdsPad] = diag ? Bi : Bj;
    const double(*ZJ)[kLdsPad] = diag ? Zi : Zj;
#pragma unroll
    for (int ks = 0; ks < 4; ++ks) {
      const int kr = ks * 4 + (lane >> 4);
      const int ci = wm * 32 + (lane & 15), cj = wn * 32 + (lane & 15);
      const double bi0 = Bi[kr][ci], bi1 = Bi[kr][ci + 16];
      const double zi0 = Zi[kr][ci], zi1 = Zi[kr][ci + 16];
      const double bj0 = BJ[kr][cj], bj1 = BJ[kr][cj + 16];
      const double zj0 = ZJ[kr][cj], zj1 = ZJ[kr][cj + 16];
      acc[0][0] = mfma_f64(bi0, zj0, acc[0][0]);
      acc[0][0] = mfma_f64(zi0, bj0, acc[0][0]);
      acc[0][1] = mfma_f64(bi0, zj1, acc[0][1]);
      acc[0][1] = mfma_f64(zi0, bj1, acc[0][1]);
      acc[1][0] = mfma_f64(bi1, zj0, acc[1][0]);
      acc[1][0] = mfma_f64(zi1, bj0, acc[1][0]);
      acc[1][1] = mfma_f64(bi1, zj1, acc[1][1]);
      acc[1][1] = mfma_f64(zi1, bj1, acc[1][1]);
    }
    __syncthreads();
  }

  const int32_t* bf = bfmap + t.bf_off;
#pragma unroll
  for (int bi = 0; bi < 2; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj) {
      const int j = j0 + wn * 32 + bj * 16 + (lane & 15);
      if (j >= nbe) continue;
      const int64_t cj = (int64_t)bf[j] * ldv;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wm * 32 + bi * 16 + 4 * r + (lane >> 4);
        if (i < nbe && (!diag || i >= j)) atomicAdd(vxc + bf[i] + cj, acc[bi][bj][r]);
      }
    }
}

int launch_syr2k_scatter_fp64(hipStream_t s, const Chunk& c, const WorkItem* g2,
                              const DevTask* tasks, const int32_t* bfmap, const double* scratch,
                              double* vxc, int64_t ldv) {
  if (c.ng2 == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_syr2k, dim3((unsigned)c.ng2), dim3(256), 0, s, g2 + c.g2_off, tasks,
                     bfmap, scratch, vxc, ldv);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
