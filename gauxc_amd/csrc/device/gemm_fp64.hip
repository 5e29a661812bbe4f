// FP64 contractions of the XC path on gfx950 matrix cores
// (v_mfma_f64_16x16x4_f64).
//
//  (X = fac B P_sub, the density reduction and Z live in density.hip.)
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
// Loader: thread (p = tid & 15, f = tid >> 4) moves point k0 + p of rows
// f, f+16, f+32, f+48 of each panel; 16 lanes cover 16 consecutive points
// (128 B) of one row.  Rows past nbe are read at row nbe-1 and zeroed by a
// select.  The next K step is loaded into registers while the current one
// runs on the matrix cores.
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
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int i0 = wi.a * kTile, j0 = wi.b * kTile;
  const bool diag = wi.a == wi.b;
  const int ld = t.npts_pad;
  const int nbe = t.nbe;
  const int plane = nbe * ld;
  const double* Bm = scratch + t.b_off;
  const double* Zm = Bm + 4 * plane;

  const int l_p = tid & 15, l_f = tid >> 4;
  int oi[4], oj[4];
  bool ki[4], kj[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int fi = i0 + l_f + 16 * q, fj = j0 + l_f + 16 * q;
    ki[q] = fi < nbe;
    kj[q] = fj < nbe;
    oi[q] = min(fi, nbe - 1) * ld + l_p;
    oj[q] = min(fj, nbe - 1) * ld + l_p;
  }
  double rbi[4], rzi[4], rbj[4], rzj[4];
  auto load = [&](int k0) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const double b = Bm[oi[q] + k0], z = Zm[oi[q] + k0];
      rbi[q] = ki[q] ? b : 0.;
      rzi[q] = ki[q] ? z : 0.;
    }
    if (!diag) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double b = Bm[oj[q] + k0], z = Zm[oj[q] + k0];
        rbj[q] = kj[q] ? b : 0.;
        rzj[q] = kj[q] ? z : 0.;
      }
    }
  };

  v4d acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0., 0., 0., 0.};

  load(0);
  for (int k0 = 0; k0 < ld; k0 += 16) {
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      Bi[l_p][l_f + 16 * q] = rbi[q];
      Zi[l_p][l_f + 16 * q] = rzi[q];
      if (!diag) {
        Bj[l_p][l_f + 16 * q] = rbj[q];
        Zj[l_p][l_f + 16 * q] = rzj[q];
      }
    }
    __syncthreads();
    if (k0 + 16 < ld) load(k0 + 16);
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

  // Scatter-add into VXC.  Element (i, j) of the tile (i >= j locally) is
  // added at (bf[j], bf[i]) of the column-major VXC: the 16 lanes of a lane
  // group hold 16 consecutive j, so each atomic instruction touches a few
  // 128-byte column segments instead of 16 columns.  The symmetrisation that
  // follows sums the two triangles, so which triangle an element lands in
  // does not matter.
  const int32_t* bf = bfmap + t.bf_off;
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int j = j0 + wn * 32 + bj * 16 + (lane & 15);
    if (j >= nbe) continue;
    const int rj = bf[j];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int i = i0 + wm * 32 + bi * 16 + 4 * r + (lane >> 4);
        if (i < nbe && (!diag || i >= j))
          atomicAdd(vxc + (int64_t)bf[i] * ldv + rj, acc[bi][bj][r]);
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
