// Density stage of the XC path: X = fac B P_sub, rho / grad rho, the
// functional with weight folding, and the EXC/NEL block sums, in one launch.
//
// One 256-thread workgroup owns one (task, 64-point block) and walks every
// 64-function column tile of the task.  Per tile it forms the transposed
// product X^T = fac P_sub B^T (64 functions x 64 points) on the matrix
// cores, so that an accumulator register holds one function of 16
// consecutive points across a lane group: the epilogue then reads B and
// grad B rows with coalesced 16-point segments (the planes are stored
// point-contiguous) and folds
//   rho_p = sum_j B_pj X_pj,   grad rho_p = 2 sum_j dB_pj X_pj
// (reference_local_host_work_driver.cxx:242-268) into per-point partials
// kept in LDS.  After the last tile the block's 64 points are complete: the
// functional runs in the same workgroup (host_integrator_exc_vxc.hpp:445-502)
// and writes the Z coefficients (a = w vrho / 2, c = 2 w vsigma grad rho)
// and the block's (w eps rho, w rho) sums.  X never reaches HBM.
#include "internal.hpp"
#include "functional.hpp"

namespace xcq {

using v4d = double __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

namespace {

constexpr int kPad64 = 68;  // FP64 stage row (doubles)
constexpr int kRowMx = 40;  // mixed stage row (bf16): 32 + 8 pad

__device__ __forceinline__ unsigned short bfbits(__bf16 v) {
  return __builtin_bit_cast(unsigned short, v);
}

__device__ __forceinline__ void split3(float x, unsigned short& a, unsigned short& b,
                                       unsigned short& c) {
  const __bf16 h1 = (__bf16)x;
  const float r1 = x - (float)h1;
  const __bf16 h2 = (__bf16)r1;
  const float r2 = r1 - (float)h2;
  const __bf16 h3 = (__bf16)r2;
  a = bfbits(h1);
  b = bfbits(h2);
  c = bfbits(h3);
}

__device__ __forceinline__ bf16x8 frag(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const us8*>(p));
}

__device__ __forceinline__ constexpr int sl_a(int p) {
  return p == 2 ? 1 : p == 4 ? 1 : p == 5 ? 2 : 0;
}
__device__ __forceinline__ constexpr int sl_b(int p) {
  return p == 1 ? 1 : p == 3 ? 2 : p == 4 ? 1 : 0;
}

// Folds one X^T tile (rows = functions, columns = points) into the
// per-point partials in LDS.  frow(bi, r) is the tile row (function) of
// accumulator register r of row block bi for this lane; the column (point)
// of lane l in column block bj is wn*32 + bj*16 + (l & 15).
template <typename T, typename FRow>
__device__ __forceinline__ void fold_tile(const double (&x)[2][2][4], FRow frow, int m0, int n0,
                                          int wm, int wn, int lane, int nbe, int64_t ld,
                                          int64_t plane, const T* Bm, int gga,
                                          double (*dens)[64][4]) {
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int pl = wn * 32 + bj * 16 + (lane & 15);
    const int64_t m = m0 + pl;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = n0 + wm * 32 + bi * 16 + frow(r);
        if (j < nbe) {
          const double xv = x[bi][bj][r];
          const int64_t o = (int64_t)j * ld + m;
          s0 += (double)Bm[o] * xv;
          if (gga) {
            s1 += (double)Bm[plane + o] * xv;
            s2 += (double)Bm[2 * plane + o] * xv;
            s3 += (double)Bm[3 * plane + o] * xv;
          }
        }
      }
    // Sum over the four 16-lane groups (different functions, same point).
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
      s0 += __shfl_xor(s0, off);
      if (gga) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
        s3 += __shfl_xor(s3, off);
      }
    }
    if (lane < 16) {
      dens[wm][pl][0] += s0;
      if (gga) {
        dens[wm][pl][1] += s1;
        dens[wm][pl][2] += s2;
        dens[wm][pl][3] += s3;
      }
    }
  }
}

// Functional on the block's 64 points (threads 0..63 = wave 0).
__device__ __forceinline__ void finish_points(const DevTask& t, int pb, int tid,
                                              double (*dens)[64][4],
                                              const double* __restrict__ weights, int func,
                                              int gga, int64_t nq, double* __restrict__ zcoef,
                                              double* __restrict__ red) {
  if (tid >= 64) return;
  const int p = pb * kPtBlock + tid;
  const double rho = dens[0][tid][0] + dens[1][tid][0];
  double gx = 0., gy = 0., gz = 0.;
  if (gga) {
    gx = 2.0 * (dens[0][tid][1] + dens[1][tid][1]);
    gy = 2.0 * (dens[0][tid][2] + dens[1][tid][2]);
    gz = 2.0 * (dens[0][tid][3] + dens[1][tid][3]);
  }
  double e = 0., n = 0., za = 0., zx = 0., zy = 0., zz = 0.;
  if (p < t.npts) {
    const double w = weights[t.pt_off + p];
    n = w * rho;
    if (func != 0) {
      const XCOut o = dev_eval_functional(func, rho, gx * gx + gy * gy + gz * gz);
      e = w * o.e;
      za = 0.5 * w * o.vr;
      const double c = 2.0 * w * o.vs;
      zx = c * gx;
      zy = c * gy;
      zz = c * gz;
    }
  }
  if (zcoef) {
    double* zc = zcoef + t.q_off + p;
    zc[0] = za;
    if (gga) {
      zc[nq] = zx;
      zc[2 * nq] = zy;
      zc[3 * nq] = zz;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * (t.red_off + pb)] = e;
    red[2 * (t.red_off + pb) + 1] = n;
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// FP64 (v_mfma_f64_16x16x4_f64; accumulator register r of lane l holds row
// 4r + l/16).  Waves: wm = function half, wn = point half of the tile.
__global__ __launch_bounds__(256) void k_density_f64(
    const WorkItem* __restrict__ pblk, const DevTask* __restrict__ tasks,
    const int32_t* __restrict__ bfmap, const double* __restrict__ P, int64_t ldp,
    const double* __restrict__ scratch, const double* __restrict__ weights, int func, int gga,
    double fac, int64_t nq, double* __restrict__ zcoef, double* __restrict__ red) {
  __shared__ double Pt[16][kPad64];  // [k][function]: A operand (P_sub)
  __shared__ double Bt[16][kPad64];  // [k][point]:    B operand (B^T)
  __shared__ double dens[2][64][4];
  const WorkItem wi = pblk[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = wi.a * kPtBlock;
  const int64_t ld = t.npts_pad;
  const int nbe = t.nbe;
  const int64_t plane = (int64_t)nbe * ld;
  const double* Bm = scratch + t.b_off;
  const int32_t* bf = bfmap + t.bf_off;
  for (int i = tid; i < 2 * 64 * 4; i += 256) (&dens[0][0][0])[i] = 0.;
  const int a_m = tid & 63, a_k = tid >> 6;  // plane loader: point, k (+4i)
  const int b_k = tid & 15, b_n = tid >> 4;  // P loader: k, function (+16i)
  auto frow = [&](int r) { return 4 * r + (lane >> 4); };

  for (int n0 = 0; n0 < nbe; n0 += 64) {
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
        Bt[k][a_m] = (k0 + k < nbe) ? Bm[(int64_t)(k0 + k) * ld + m0 + a_m] : 0.;
      }
      const int kk = k0 + b_k;
      const int64_t rowP = kk < nbe ? (int64_t)bf[kk] : -1;
#pragma unroll
      for (int i = 0; i < 4; ++i)
        Pt[b_k][b_n + 16 * i] = (rowP >= 0 && colP[i] >= 0) ? fac * P[rowP + colP[i]] : 0.;
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        const double a0 = Pt[kr][wm * 32 + (lane & 15)];
        const double a1 = Pt[kr][wm * 32 + 16 + (lane & 15)];
        const double b0 = Bt[kr][wn * 32 + (lane & 15)];
        const double b1 = Bt[kr][wn * 32 + 16 + (lane & 15)];
        acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, acc[0][0], 0, 0, 0);
        acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, acc[0][1], 0, 0, 0);
        acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, acc[1][0], 0, 0, 0);
        acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, acc[1][1], 0, 0, 0);
      }
      __syncthreads();
    }
    double x[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) x[i][j][r] = acc[i][j][r];
    fold_tile<double>(x, frow, m0, n0, wm, wn, lane, nbe, ld, plane, Bm, gga, dens);
  }
  __syncthreads();
  finish_points(t, wi.a, tid, dens, weights, func, gga, nq, zcoef, red);
}

// ---------------------------------------------------------------------------
// Mixed precision (v_mfma_f32_16x16x32_bf16, three-slice operands, FP32
// inside one K block, FP64 across; accumulator register r of lane l holds
// row 4(l/16) + r).
__global__ __launch_bounds__(256) void k_density_mx(
    const WorkItem* __restrict__ pblk, const DevTask* __restrict__ tasks,
    const int32_t* __restrict__ bfmap, const uint2* __restrict__ Ps, int64_t ldp,
    const float* __restrict__ scratch, const double* __restrict__ weights, int func, int gga,
    double fac, int64_t nq, double* __restrict__ zcoef, double* __restrict__ red) {
  __shared__ __attribute__((aligned(16))) unsigned short Pt[3][64][kRowMx];  // [s][fn][k]
  __shared__ __attribute__((aligned(16))) unsigned short Bt[3][64][kRowMx];  // [s][pt][k]
  __shared__ double dens[2][64][4];
  const WorkItem wi = pblk[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave & 1, wn = wave >> 1;
  const int m0 = wi.a * kPtBlock;
  const int64_t ld = t.npts_pad;
  const int nbe = t.nbe;
  const int64_t plane = (int64_t)nbe * ld;
  const float* Bm = scratch + t.b_off;
  const int32_t* bf = bfmap + t.bf_off;
  for (int i = tid; i < 2 * 64 * 4; i += 256) (&dens[0][0][0])[i] = 0.;
  const int r64 = tid & 63, kg = tid >> 6;
  const int kk = 8 * (lane >> 4);
  auto frow = [&](int r) { return 4 * (lane >> 4) + r; };

  for (int n0 = 0; n0 < nbe; n0 += 64) {
    // P is symmetric: read P(k, n) as row bf[k], column bf[n] so that
    // neighbouring threads (neighbouring n) touch neighbouring columns.
    const int ncol = n0 + r64;
    const int64_t colP = ncol < nbe ? (int64_t)bf[ncol] : -1;
    double acc[2][2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
      for (int b = 0; b < 2; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.;
    for (int k0 = 0; k0 < nbe; k0 += 32) {
      {  // B^T operand: 8 consecutive functions of one point per thread
        us8 s1, s2, s3;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + kg * 8 + e;
          const float xv = k < nbe ? Bm[(int64_t)k * ld + m0 + r64] : 0.f;
          unsigned short a, b, c;
          split3(xv, a, b, c);
          s1[e] = a;
          s2[e] = b;
          s3[e] = c;
        }
        *reinterpret_cast<us8*>(&Bt[0][r64][kg * 8]) = s1;
        *reinterpret_cast<us8*>(&Bt[1][r64][kg * 8]) = s2;
        *reinterpret_cast<us8*>(&Bt[2][r64][kg * 8]) = s3;
      }
      {  // P_sub operand: 8 consecutive functions of one column per thread
        us8 s1, s2, s3;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + kg * 8 + e;
          uint2 v = make_uint2(0u, 0u);
          if (k < nbe && colP >= 0) v = Ps[(int64_t)bf[k] * ldp + colP];
          s1[e] = (unsigned short)(v.x & 0xffffu);
          s2[e] = (unsigned short)(v.x >> 16);
          s3[e] = (unsigned short)(v.y & 0xffffu);
        }
        *reinterpret_cast<us8*>(&Pt[0][r64][kg * 8]) = s1;
        *reinterpret_cast<us8*>(&Pt[1][r64][kg * 8]) = s2;
        *reinterpret_cast<us8*>(&Pt[2][r64][kg * 8]) = s3;
      }
      __syncthreads();
      bf16x8 fa[2][3], fb[2][3];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int s = 0; s < 3; ++s) {
          fa[q][s] = frag(&Pt[s][wm * 32 + q * 16 + (lane & 15)][kk]);
          fb[q][s] = frag(&Bt[s][wn * 32 + q * 16 + (lane & 15)][kk]);
        }
      __syncthreads();
      v4f c[2][2];
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) c[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int pr = 0; pr < 6; ++pr) {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
            c[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[bi][sl_b(pr)],
                                                                 fb[bj][sl_a(pr)], c[bi][bj], 0,
                                                                 0, 0);
      }
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[bi][bj][r] += (double)c[bi][bj][r];
    }
    double x[2][2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) x[i][j][r] = fac * acc[i][j][r];
    fold_tile<float>(x, frow, m0, n0, wm, wn, lane, nbe, ld, plane, Bm, gga, dens);
  }
  __syncthreads();
  finish_points(t, wi.a, tid, dens, weights, func, gga, nq, zcoef, red);
}

int launch_density_fp64(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                        const DevTask* tasks, const int32_t* bfmap, const double* P,
                        int64_t ldp, const double* scratch, const double* weights, int func,
                        int gga, double fac, double* zcoef, double* red) {
  if (c.npblk == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_density_f64, dim3((unsigned)c.npblk), dim3(256), 0, s,
                     pblk + c.pblk_off, tasks, bfmap, P, ldp, scratch, weights, func, gga, fac,
                     c.nq, zcoef, red);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_density_mixed(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                         const DevTask* tasks, const int32_t* bfmap, const uint2* Ps,
                         int64_t ldp, const float* scratch, const double* weights, int func,
                         int gga, double fac, double* zcoef, double* red) {
  if (c.npblk == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_density_mx, dim3((unsigned)c.npblk), dim3(256), 0, s, pblk + c.pblk_off,
                     tasks, bfmap, Ps, ldp, scratch, weights, func, gga, fac, c.nq, zcoef, red);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
