// Mixed-precision contractions of the XC path on gfx950
// (v_mfma_f32_16x16x32_bf16).
//
// Numerics (chosen from a CPU emulation against the FP64 path):
//  * B, dB and Z are stored in FP32; P is held as three bfloat16 slices
//    P ~ p1 + p2 + p3 (24 significant bits).  Each FP32 operand element is
//    split on the fly into three bfloat16 slices the same way (exact: 3 x 8
//    bits cover the FP32 significand).
//  * A product A*B is formed from the six slice products a_s * b_t with
//    s + t <= 2 (dropped terms are ~2^-24 relative), accumulated in FP32
//    inside one 32-deep K block (one MFMA per slice product) and added to
//    FP64 accumulators across K blocks, so FP32 rounding does not grow with
//    the contraction length (npts can reach 2e4 in the B^T Z product).
//  * Densities, the functional, EXC/NEL and VXC accumulate in FP64.
//
// Kernels mirror the FP64 ones (gemm_fp64.hip; the density stage is in
// density.hip): k_syr2k_mx fuses the scatter-add into VXC into the
// B^T Z + Z^T B epilogue.  Operand layout of the MFMA (measured on gfx950): lane l holds
// A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. +7][l%16]; accumulator
// register r of lane l is D[4(l/16)+r][l%16].
#include "internal.hpp"

namespace xcq {

int launch_collocation_fp32(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz,
                            float* scratch, int deriv);

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int kKB = 32;    // K block = MFMA depth
constexpr int kRow = 40;   // LDS row length in bf16 (32 + 8 pad, 80 B: 16-B aligned)

__device__ __forceinline__ unsigned short bf_bits(__bf16 v) {
  return __builtin_bit_cast(unsigned short, v);
}

__device__ __forceinline__ void split3f(float x, unsigned short& a, unsigned short& b,
                                        unsigned short& c) {
  const __bf16 h1 = (__bf16)x;
  const float r1 = x - (float)h1;
  const __bf16 h2 = (__bf16)r1;
  const float r2 = r1 - (float)h2;
  const __bf16 h3 = (__bf16)r2;
  a = bf_bits(h1);
  b = bf_bits(h2);
  c = bf_bits(h3);
}

__device__ __forceinline__ bf16x8 ldfrag(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const us8*>(p));
}

__device__ __forceinline__ v4f mfma_v(bf16x8 a, bf16x8 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Slice pairs (s, t) with s + t <= 2, largest terms first:
// (0,0) (0,1) (1,0) (0,2) (1,1) (2,0).
__device__ __forceinline__ constexpr int pair_a(int p) {
  return p == 2 ? 1 : p == 4 ? 1 : p == 5 ? 2 : 0;
}
__device__ __forceinline__ constexpr int pair_b(int p) {
  return p == 1 ? 1 : p == 3 ? 2 : p == 4 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// P (FP64, nbf x nbf) -> packed (p1, p2, p3, 0) bfloat16 slices.
__global__ __launch_bounds__(256) void k_split_P(int64_t n, const double* __restrict__ P,
                                                 uint2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = P[i];
  const __bf16 h1 = (__bf16)(float)x;
  const double r1 = x - (double)(float)h1;
  const __bf16 h2 = (__bf16)(float)r1;
  const double r2 = r1 - (double)(float)h2;
  const __bf16 h3 = (__bf16)(float)r2;
  uint2 v;
  v.x = (unsigned)bf_bits(h1) | ((unsigned)bf_bits(h2) << 16);
  v.y = (unsigned)bf_bits(h3);
  out[i] = v;
}

// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_syr2k_mx(const WorkItem* __restrict__ tiles,
                                                  const DevTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ bfmap,
                                                  const float* __restrict__ scratch,
                                                  double* __restrict__ vxc, int64_t ldv) {
  __shared__ __attribute__((aligned(16))) unsigned short Bi[3][64][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Zi[3][64][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Bj[3][64][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Zj[3][64][kRow];
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
  const float* Bm = scratch + t.b_off;
  const float* Zm = Bm + 4 * plane;
  // Loader: thread (g = tid & 3, f = tid >> 2) moves points k0+8g..k0+8g+7
  // of row f of each panel (four lanes cover one row's 128 B); rows past
  // nbe are read at row nbe-1 and zeroed by a select.
  const int pg = tid & 3, f = tid >> 2;
  const int fi = i0 + f, fj = j0 + f;
  const bool ki = fi < nbe, kj = fj < nbe;
  const int oi = min(fi, nbe - 1) * ld + 8 * pg, oj = min(fj, nbe - 1) * ld + 8 * pg;

  double acc[2][2][4];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.;

  float4 pre[4][2];  // (Bi, Zi, Bj, Zj) x two float4
  auto load = [&](int k0) {
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const float4 b = *reinterpret_cast<const float4*>(Bm + oi + k0 + 4 * h);
      const float4 z = *reinterpret_cast<const float4*>(Zm + oi + k0 + 4 * h);
      pre[0][h] = ki ? b : z4;
      pre[1][h] = ki ? z : z4;
    }
    if (!diag) {
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const float4 b = *reinterpret_cast<const float4*>(Bm + oj + k0 + 4 * h);
        const float4 z = *reinterpret_cast<const float4*>(Zm + oj + k0 + 4 * h);
        pre[2][h] = kj ? b : z4;
        pre[3][h] = kj ? z : z4;
      }
    }
  };
  auto stage = [&](const float4 (&v)[2], unsigned short (*dst)[64][kRow]) {
    const float* pv = reinterpret_cast<const float*>(&v[0]);  // 8 consecutive points
    us8 s1, s2, s3;
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const float x = pv[e];
      unsigned short a, b, c;
      split3f(x, a, b, c);
      s1[e] = a;
      s2[e] = b;
      s3[e] = c;
    }
    *reinterpret_cast<us8*>(&dst[0][f][pg * 8]) = s1;
    *reinterpret_cast<us8*>(&dst[1][f][pg * 8]) = s2;
    *reinterpret_cast<us8*>(&dst[2][f][pg * 8]) = s3;
  };

  load(0);
  for (int k0 = 0; k0 < ld; k0 += kKB) {
    stage(pre[0], Bi);
    stage(pre[1], Zi);
    if (!diag) {
      stage(pre[2], Bj);
      stage(pre[3], Zj);
    }
    __syncthreads();
    if (k0 + kKB < ld) load(k0 + kKB);
    unsigned short(*BJ)[64][kRow] = diag ? Bi : Bj;
    unsigned short(*ZJ)[64][kRow] = diag ? Zi : Zj;
    const int kk = 8 * (lane >> 4);
    bf16x8 fbi[2][3], fzi[2][3], fbj[2][3], fzj[2][3];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const int ra = wm * 32 + q * 16 + (lane & 15);
        const int rb = wn * 32 + q * 16 + (lane & 15);
        fbi[q][s] = ldfrag(&Bi[s][ra][kk]);
        fzi[q][s] = ldfrag(&Zi[s][ra][kk]);
        fbj[q][s] = ldfrag(&BJ[s][rb][kk]);
        fzj[q][s] = ldfrag(&ZJ[s][rb][kk]);
      }
    __syncthreads();
    v4f c[2][2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj) c[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int pr = 0; pr < 6; ++pr) {
      const int sa = pair_a(pr), sb = pair_b(pr);
#pragma unroll
      for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int bj = 0; bj < 2; ++bj) {
          c[bi][bj] = mfma_v(fbi[bi][sa], fzj[bj][sb], c[bi][bj]);  // B_i^T Z_j
          c[bi][bj] = mfma_v(fzi[bi][sa], fbj[bj][sb], c[bi][bj]);  // Z_i^T B_j
        }
    }
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int bj = 0; bj < 2; ++bj)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[bi][bj][r] += (double)c[bi][bj][r];
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
        const int i = i0 + wm * 32 + bi * 16 + 4 * (lane >> 4) + r;
        if (i < nbe && (!diag || i >= j))
          atomicAdd(vxc + (int64_t)bf[i] * ldv + rj, acc[bi][bj][r]);
      }
  }
}

// ---------------------------------------------------------------------------
int launch_mixed_pipeline(struct ::xcq_context_s* ctx, const Chunk& c, int func, int gga,
                          int want_vxc, hipEvent_t* ev) {
  const CtxView v = view_of(ctx);
  hipStream_t s = v.stream;
  const int64_t nel = (int64_t)v.nbf * v.nbf;
  if (*v.Psplit == nullptr) {
    XCQ_HIP(hipMalloc(v.Psplit, (size_t)nel * sizeof(uint2)));
    hipLaunchKernelGGL(k_split_P, dim3((unsigned)((nel + 255) / 256)), dim3(256), 0, s, nel,
                       (const double*)v.P, (uint2*)*v.Psplit);
    XCQ_HIP(hipGetLastError());
  }
  float* scr = static_cast<float*>(v.scratch);
  XCQ_HIP(hipEventRecord(ev[0], s));
  int rc = launch_collocation_fp32(s, c, v.work, v.tasks, v.shells, v.shell_list, v.lsh, v.px,
                                   v.py, v.pz, scr, gga);
  if (rc) return rc;
  XCQ_HIP(hipEventRecord(ev[1], s));
  XCQ_HIP(hipEventRecord(ev[2], s));
  rc = launch_density_mixed(s, c, v.work, v.tasks, v.bfmap, (const uint2*)*v.Psplit,
                            (int64_t)v.nbf, scr, v.w, func, gga, func == 0 ? 1.0 : 2.0, want_vxc,
                            v.red);
  if (rc) return rc;
  XCQ_HIP(hipEventRecord(ev[3], s));
  XCQ_HIP(hipEventRecord(ev[4], s));
  XCQ_HIP(hipEventRecord(ev[5], s));
  if (!want_vxc) return XCQ_SUCCESS;
  XCQ_HIP(hipEventRecord(ev[6], s));
  XCQ_HIP(hipEventRecord(ev[7], s));
  XCQ_HIP(hipEventRecord(ev[8], s));
  if (c.ng2) {
    hipLaunchKernelGGL(k_syr2k_mx, dim3((unsigned)c.ng2), dim3(256), 0, s, v.work + c.g2_off,
                       v.tasks, v.bfmap, scr, v.vxc, (int64_t)v.nbf);
    XCQ_HIP(hipGetLastError());
  }
  XCQ_HIP(hipEventRecord(ev[9], s));
  return XCQ_SUCCESS;
}

}  // namespace xcq
