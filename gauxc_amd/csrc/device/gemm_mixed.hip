// Mixed-precision contractions of the XC path on gfx950
// (v_mfma_f32_16x16x32_bf16).
//
// Numerics (chosen from a CPU emulation against the FP64 path):
//  * dB/dx, dB/dy, dB/dz and Z are stored in FP32.  The value plane B is
//    stored as FP32 plus a bfloat16 residual (plane 5 of the task scratch),
//    about 32 significant bits: FP32 rounding of B was the largest error
//    term of the EXC/VXC budget.
//  * P is held as three bfloat16 slices p0 + p1 + p2 (24 significant bits).
//    Each FP32 operand element is split on the fly into three bfloat16
//    slices (exact: 3 x 8 bits cover the FP32 significand); B's residual is
//    its fourth slice b3.
//  * A product is formed from the leading slice product (one MFMA, its own
//    FP32 accumulator, so the large partial sum is rounded once per K block)
//    and six correction products (b0 z1, b1 z0, b0 z2, b1 z1, b2 z0, b3 z0;
//    dropped terms are ~2^-24 relative) in a second FP32 accumulator, both
//    added to FP64 accumulators after every 32-deep K block, so FP32
//    rounding does not grow with the contraction length (npts can reach
//    2e4 in the B^T Z product).
//  * Densities, the functional, EXC/NEL and VXC accumulate in FP64.
//
// Kernels mirror the FP64 ones (gemm_fp64.hip; the density stage is in
// density.hip): k_syr2k_mx fuses the scatter-add into VXC into the
// B^T Z + Z^T B epilogue.  Operand layout of the MFMA (measured on gfx950): lane l holds
// A[l%16][8(l/16) .. 8(l/16)+7] and B[8(l/16) .. +7][l%16]; accumulator
// register r of lane l is D[4(l/16)+r][l%16].
#include "internal.hpp"

#include <type_traits>

namespace xcq {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float v4f __attribute__((ext_vector_type(4)));
typedef unsigned short us8 __attribute__((ext_vector_type(8)));

constexpr int kKB = 32;    // K block = MFMA depth
constexpr int kRow = 40;   // LDS row length in bf16 (32 + 8 pad, 80 B: 16-B aligned)

__device__ __forceinline__ unsigned short bf_bits(__bf16 v) {
  return __builtin_bit_cast(unsigned short, v);
}

// Splits two FP32 values into three bfloat16 slices each (round to nearest
// even at every step, so x = a + b + c exactly); the slices of the pair come
// packed, one v_cvt_pk_bf16_f32 per slice.
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split3f_pair(f2v x, unsigned& a, unsigned& b, unsigned& c) {
  const bf2v h1 = __builtin_convertvector(x, bf2v);
  const f2v r1 = x - __builtin_convertvector(h1, f2v);
  const bf2v h2 = __builtin_convertvector(r1, bf2v);
  const f2v r2 = r1 - __builtin_convertvector(h2, f2v);
  const bf2v h3 = __builtin_convertvector(r2, bf2v);
  a = __builtin_bit_cast(unsigned, h1);
  b = __builtin_bit_cast(unsigned, h2);
  c = __builtin_bit_cast(unsigned, h3);
}

__device__ __forceinline__ bf16x8 ldfrag(const unsigned short* p) {
  return __builtin_bit_cast(bf16x8, *reinterpret_cast<const us8*>(p));
}

__device__ __forceinline__ v4f mfma_v(bf16x8 a, bf16x8 b, v4f c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Correction products (B slice, Z slice) after the leading b0 z0:
// (0,1) (1,0) (0,2) (1,1) (2,0) (3,0); slice 3 of B is the stored bfloat16
// residual of the value plane.
__device__ __forceinline__ constexpr int corr_b(int k) {
  return k == 1 ? 1 : k == 3 ? 1 : k == 4 ? 2 : k == 5 ? 3 : 0;
}
__device__ __forceinline__ constexpr int corr_z(int k) {
  return k == 0 ? 1 : k == 2 ? 2 : k == 3 ? 1 : 0;
}

// ---------------------------------------------------------------------------
// P (FP64, nbf x nbf) -> FP32 copy (the mixed density stage's P_sub source).
__global__ __launch_bounds__(256) void k_P_to_f32(int64_t n, const double* __restrict__ P,
                                                  float* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (float)P[i];
}

// ---------------------------------------------------------------------------
// B^T Z + Z^T B with the VXC scatter-add (mirrors k_syr2k of the FP64 path),
// templated on the function tile NT.  Loader: thread (g = tid & 3,
// f = tid >> 2) moves points k0+8g..k0+8g+7 of one row (four lanes cover a
// row's 128 B): for NT = 64 row f of the i and j panels, for NT = 32 row
// f & 31 of the i panels (f < 32) or of the j panels (f >= 32).  Rows past
// nbe are read at row nbe-1 and zeroed by a select.
template <int NT, int PF = 1>
__global__ __launch_bounds__(256, NT == 32 ? 4 : 2) void k_syr2k_mx(QueueArgs qa,
                                                  const DevTask* __restrict__ tasks,
                                                  const int32_t* __restrict__ bfmap,
                                                  const float* __restrict__ scratch,
                                                  double* __restrict__ vxc_all, double drop,
    const unsigned* __restrict__ tmax) {
  constexpr int BI = NT / 32;
  __shared__ __attribute__((aligned(16))) unsigned short Bi[4][NT][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Zi[3][NT][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Bj[4][NT][kRow];
  __shared__ __attribute__((aligned(16))) unsigned short Zj[3][NT][kRow];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int pg = tid & 3, f = tid >> 2;
  const int frow = NT == 64 ? f : (f & 31);
  const int kk = 8 * (lane >> 4);

  XCQ_FOR_EACH_ITEM(qa, it) {
    const int64_t t0 = qa.tstamp ? wall_clock64() : 0;
    const WorkItem wi = qa.items[it];
    const DevTask t = tasks[wi.task];
    const int i0 = wi.a * NT, j0 = wi.b * NT;
    const bool diag = wi.a == wi.b;
    // Tile pairs whose every element would be dropped are skipped whole
    // (uniform across the workgroup).
    if (tile_pair_negligible(t, tmax, i0, j0, NT, drop)) {
      stamp_item(qa, it, t0);
      continue;
    }
    const int ld = t.npts_pad;
    const int nbe = t.nbe;
    const int plane = nbe * ld;
    const float* Bm = scratch + t.b_off_mx;
    const float* Zm = Bm + 4 * plane;
    // bfloat16 residual of the value plane (plane 5, 2 bytes per element)
    const unsigned short* Lm = reinterpret_cast<const unsigned short*>(Bm + 5 * (int64_t)plane);
    const int64_t ldv = t.ldp;
    double* vxc = vxc_all + t.p_off;
    const int fi = i0 + frow, fj = j0 + frow;
    const bool ki = fi < nbe, kj = fj < nbe;
    const int oi = min(fi, nbe - 1) * ld + 8 * pg, oj = min(fj, nbe - 1) * ld + 8 * pg;

    double acc[BI][BI][4];
#pragma unroll
    for (int a = 0; a < BI; ++a)
#pragma unroll
      for (int b = 0; b < BI; ++b)
#pragma unroll
        for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.;

    // Raw prefetch registers, a ring of PF K blocks (PF = 3 for 32-wide
    // tiles: two blocks stay in flight while one is staged and multiplied;
    // 64-wide tiles have no register room for more than one).  The row
    // masks are applied when staged, so the loads stay in flight across the
    // matrix instructions.
    // Panel sides this thread loads: both (i then j) for 64-wide tiles, one
    // (i for f < 32, j otherwise) for 32-wide ones; the j side is skipped on
    // diagonal tiles (it equals the i side).
    constexpr int SIDES = NT == 64 ? 2 : 1;
    auto side_id = [&](int sd) { return NT == 64 ? sd : (f < 32 ? 0 : 1); };
    auto side_on = [&](int sd) { return side_id(sd) == 0 || !diag; };
    float4 pre[PF][SIDES][2][2];  // [slot][side][B, Z][two float4]
    uint4 prel[PF][SIDES];        // residuals of B (8 bfloat16)
    auto load = [&](int k0, auto S) {
      constexpr int sl = decltype(S)::value;
#pragma unroll
      for (int sd = 0; sd < SIDES; ++sd) {
        if (!side_on(sd)) continue;
        const int o = (side_id(sd) == 0 ? oi : oj) + k0;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
          pre[sl][sd][0][h] = *reinterpret_cast<const float4*>(Bm + o + 4 * h);
          pre[sl][sd][1][h] = *reinterpret_cast<const float4*>(Zm + o + 4 * h);
        }
        prel[sl][sd] = *reinterpret_cast<const uint4*>(Lm + o);
      }
    };
    auto stage = [&](const float4 (&v)[2], bool ok, unsigned short (*dst)[NT][kRow]) {
      const float* pv = reinterpret_cast<const float*>(&v[0]);  // 8 consecutive points
      uint4 s1, s2, s3;
      unsigned* w1 = reinterpret_cast<unsigned*>(&s1);
      unsigned* w2 = reinterpret_cast<unsigned*>(&s2);
      unsigned* w3 = reinterpret_cast<unsigned*>(&s3);
#pragma unroll
      for (int e = 0; e < 4; ++e)
        split3f_pair(f2v{ok ? pv[2 * e] : 0.f, ok ? pv[2 * e + 1] : 0.f}, w1[e], w2[e], w3[e]);
      *reinterpret_cast<uint4*>(&dst[0][frow][pg * 8]) = s1;
      *reinterpret_cast<uint4*>(&dst[1][frow][pg * 8]) = s2;
      *reinterpret_cast<uint4*>(&dst[2][frow][pg * 8]) = s3;
    };

    auto stage_lo = [&](const uint4& l, bool ok, unsigned short (*dst)[NT][kRow]) {
      *reinterpret_cast<uint4*>(&dst[3][frow][pg * 8]) = ok ? l : make_uint4(0u, 0u, 0u, 0u);
    };

    // One K block: stage ring slot S, refill it with block k0 + PF * kKB,
    // multiply.
    auto step = [&](int k0, auto S) {
      constexpr int sl = decltype(S)::value;
#pragma unroll
      for (int sd = 0; sd < SIDES; ++sd) {
        if (!side_on(sd)) continue;
        const bool i_side = side_id(sd) == 0;
        const bool ok = i_side ? ki : kj;
        unsigned short(*db)[NT][kRow] = i_side ? Bi : Bj;
        unsigned short(*dz)[NT][kRow] = i_side ? Zi : Zj;
        stage(pre[sl][sd][0], ok, db);
        stage_lo(prel[sl][sd], ok, db);
        stage(pre[sl][sd][1], ok, dz);
      }
      __syncthreads();
      if (k0 + PF * kKB < ld) load(k0 + PF * kKB, S);
      unsigned short(*BJ)[NT][kRow] = diag ? Bi : Bj;
      unsigned short(*ZJ)[NT][kRow] = diag ? Zi : Zj;
      bf16x8 fbi[BI][4], fzi[BI][3], fbj[BI][4], fzj[BI][3];
#pragma unroll
      for (int q = 0; q < BI; ++q)
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int ra = wm * 16 * BI + q * 16 + (lane & 15);
          const int rb = wn * 16 * BI + q * 16 + (lane & 15);
          fbi[q][s] = ldfrag(&Bi[s][ra][kk]);
          fbj[q][s] = ldfrag(&BJ[s][rb][kk]);
          if (s < 3) {
            fzi[q][s] = ldfrag(&Zi[s][ra][kk]);
            fzj[q][s] = ldfrag(&ZJ[s][rb][kk]);
          }
        }
      __syncthreads();
      // Leading products b0 z0 in one FP32 accumulator, the corrections
      // (b0 z1, b1 z0, b0 z2, b1 z1, b2 z0, b3 z0; b3 = stored residual) in
      // another, for B_i^T Z_j and Z_i^T B_j alike.
      v4f cm[BI][BI], cc[BI][BI];
#pragma unroll
      for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < BI; ++bj) {
          cm[bi][bj] = mfma_v(fbi[bi][0], fzj[bj][0], v4f{0.f, 0.f, 0.f, 0.f});
          cm[bi][bj] = mfma_v(fzi[bi][0], fbj[bj][0], cm[bi][bj]);
          cc[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};
        }
#pragma unroll
      for (int pr = 0; pr < 6; ++pr) {
        const int sb = corr_b(pr), sz = corr_z(pr);
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int bj = 0; bj < BI; ++bj) {
            cc[bi][bj] = mfma_v(fbi[bi][sb], fzj[bj][sz], cc[bi][bj]);  // B_i^T Z_j
            cc[bi][bj] = mfma_v(fzi[bi][sz], fbj[bj][sb], cc[bi][bj]);  // Z_i^T B_j
          }
      }
#pragma unroll
      for (int bi = 0; bi < BI; ++bi)
#pragma unroll
        for (int bj = 0; bj < BI; ++bj)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            acc[bi][bj][r] += (double)cm[bi][bj][r] + (double)cc[bi][bj][r];
    };

    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, (PF > 1 ? 1 : 0)>;
    using I2 = std::integral_constant<int, (PF > 2 ? 2 : 0)>;
    load(0, I0{});
    if (PF > 1 && kKB < ld) load(kKB, I1{});
    if (PF > 2 && 2 * kKB < ld) load(2 * kKB, I2{});
    for (int k0 = 0; k0 < ld; k0 += PF * kKB) {
      step(k0, I0{});
      if (PF > 1 && k0 + kKB < ld) step(k0 + kKB, I1{});
      if (PF > 2 && k0 + 2 * kKB < ld) step(k0 + 2 * kKB, I2{});
    }

    // Scatter-add into VXC in transposed order (see k_syr2k).
    const int32_t* bf = bfmap + t.bf_off;
#pragma unroll
    for (int bj = 0; bj < BI; ++bj) {
      const int j = j0 + wn * 16 * BI + bj * 16 + (lane & 15);
      if (j < nbe) {
        const int rj = bf[j];
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int i = i0 + wm * 16 * BI + bi * 16 + 4 * (lane >> 4) + r;
            // Negligible contributions are not added (NaN still is).
            if (i < nbe && (!diag || i >= j) && !(fabs(acc[bi][bj][r]) < drop))
              atomicAdd(vxc + (int64_t)bf[i] * ldv + rj, acc[bi][bj][r]);
          }
      }
    }
    stamp_item(qa, it, t0);
  }
}

int launch_syr2k_mixed(const CtxView& v, const WorkRange& r, int nt) {
  if (r.n == 0) return XCQ_SUCCESS;
  const QueueArgs qa = queue_args(v, r);
  const float* scr = static_cast<const float*>(v.scratch);
  if (nt == 32)
    hipLaunchKernelGGL(k_syr2k_mx<32>, dim3(grid_of(r)), dim3(256), 0, v.stream, qa, v.tasks,
                       v.bfmap, scr, v.vxc, v.vxc_drop, v.vxc_drop > 0.0 ? v.tmax : nullptr);
  else
    hipLaunchKernelGGL(k_syr2k_mx<64>, dim3(grid_of(r)), dim3(256), 0, v.stream, qa, v.tasks,
                       v.bfmap, scr, v.vxc, v.vxc_drop, v.vxc_drop > 0.0 ? v.tmax : nullptr);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_P_to_f32(hipStream_t s, int64_t n, const double* P, float* out) {
  if (n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_P_to_f32, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, n, P, out);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
