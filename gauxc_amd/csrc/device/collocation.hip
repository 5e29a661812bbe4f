// Shell collocation (+gradient) on gfx950.
//
// Semantics and the per-shell evaluation: shell_eval.hpp.
//
// Layout: task t owns [comp][nbe][npts_pad] in the chunk scratch
// (comp = B, dB/dx, dB/dy, dB/dz), point index fastest, padded points 0.
// Launch: one 256-thread block per (task, 64-point block); lane = point,
// the four waves stride over the task's shells.  Stores of one wave are
// 64 consecutive points of one basis function: fully coalesced.
#include "internal.hpp"
#include "shell_eval.hpp"

#include <cmath>
#include <type_traits>

namespace xcq {

int upload_sph_table(hipStream_t) { return XCQ_SUCCESS; }

// bfloat16 (round to nearest even) of the part of v that its FP32 rounding
// drops: with it the stored value plane carries ~32 significant bits.
__device__ __forceinline__ unsigned short resid_bf16(double v) {
  const float hi = (float)v;
  return __builtin_bit_cast(unsigned short, (__bf16)(float)(v - (double)hi));
}

// Values (and gradients) of one shell at one point, written to the scratch
// in double arithmetic.  T is the storage type (double or float); with FP32
// storage the value plane also gets its bfloat16 residual (lo, same
// indexing as out).  Stores go through a buffer resource over the task's scratch: the byte
// offset of a row is a scalar (row0 = offset of the shell's first function
// at the block's first point), the lane offset one shared register.
typedef unsigned U32x2 __attribute__((ext_vector_type(2)));

struct TaskStore {
  __amdgpu_buffer_rsrc_t rs;
  int vo4, vo2;       // lane byte offsets for 4-/8-byte and 2-byte elements
  unsigned lo_base;   // byte offset of the residual plane
};

template <typename T>
__device__ __forceinline__ void store_elem(const TaskStore& ts, int64_t idx, T v) {
  const int so = (int)(unsigned)(idx * (int64_t)sizeof(T));
  if constexpr (sizeof(T) == 4)
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(int, v), ts.rs, ts.vo4, so, 0);
  else
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(U32x2, v), ts.rs, ts.vo4 * 2, so,
                                          0);
}

template <int L, bool GRAD, typename T>
__device__ __forceinline__ void shell_eval(double x, double y, double z, double R, double dR,
                                           bool pure, bool valid, const TaskStore& ts,
                                           int64_t row0, int64_t ld, int64_t plane) {
  shell_functions<L, GRAD>(x, y, z, R, dR, pure,
                           [&](int f, double v, double gx, double gy, double gz) {
    const int64_t r = row0 + f * ld;
    store_elem<T>(ts, r, valid ? T(v) : T(0));
    if constexpr (sizeof(T) == 4)
      __builtin_amdgcn_raw_buffer_store_b16(
          (short)(valid ? resid_bf16(v) : (unsigned short)0), ts.rs, ts.vo2,
          (int)((unsigned)ts.lo_base + (unsigned)(2 * r)), 0);
    if constexpr (GRAD) {
      store_elem<T>(ts, plane + r, valid ? T(gx) : T(0));
      store_elem<T>(ts, 2 * plane + r, valid ? T(gy) : T(0));
      store_elem<T>(ts, 3 * plane + r, valid ? T(gz) : T(0));
    }
  });
}

constexpr int kCollShells = 256;  // shells staged per pass
constexpr int kCollPrims = 4;     // primitives per shell staged in LDS

// Occupancy target (waves per SIMD) by the highest angular momentum a
// launch evaluates: spherical f (LMAX 3) with gradients fits 168 registers,
// g (LMAX 4) 256.  Values alone up to f: six waves for FP64 planes (no
// spills), five for FP32 planes (six spill and measured slower: 24 waters,
// 1.23 ms at five, 1.29 ms at six, 1.32 ms at four).
constexpr int coll_waves(int lmax, bool grad, bool f64) {
  return lmax <= 2 ? 4 : lmax == 3 ? (grad ? 3 : f64 ? 6 : 5) : 2;
}

// Shell parameters of a task in LDS, one pass of up to kCollShells shells.
struct ShellStage {
  double O[3][kCollShells];
  double a[kCollPrims][kCollShells], c[kCollPrims][kCollShells];
  int info[kCollShells];  // l | pure << 8 | nprim << 16
  int bf0[kCollShells], ish[kCollShells];
};

// Walks the task's shells for the lane's point (gx, gy, gz): the shell
// parameters are staged in LDS once per pass of up to kCollShells shells by
// one thread each, so the shell loop reads them with LDS broadcasts instead
// of a chain of dependent scalar loads per shell; primitives past the first
// kCollPrims are read from the shell record.  Wave w takes shells w, w + 4,
// ... (ascending) and calls body(l, pure, first local function, x, y, z,
// R, dR) with the displacement from the centre, R = sum c e^{-a r^2} and
// dR = -2 sum a c e^{-a r^2}.
template <typename Body>
__device__ __forceinline__ void for_block_shells(const DevTask& t,
                                                 const DevShell* __restrict__ shells,
                                                 const int32_t* __restrict__ shell_list,
                                                 const int32_t* __restrict__ lsh_first,
                                                 double gx, double gy, double gz, ShellStage& st,
                                                 Body&& body) {
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  for (int pass = 0; pass < t.nshells; pass += kCollShells) {
    const int nsh = min(kCollShells, t.nshells - pass);
    __syncthreads();  // the previous pass is done with the staged shells
    for (int q = tid; q < nsh; q += 256) {
      const int ish = shell_list[t.sh_off + pass + q];
      const DevShell* sh = shells + ish;
      st.ish[q] = ish;
      st.bf0[q] = lsh_first[t.lsh_off + pass + q];
      st.info[q] = sh->l | (sh->pure ? 1 << 8 : 0) | (sh->nprim << 16);
      for (int d = 0; d < 3; ++d) st.O[d][q] = sh->O[d];
      const int np = min(sh->nprim, kCollPrims);
      for (int k = 0; k < np; ++k) {
        st.a[k][q] = sh->alpha[k];
        st.c[k][q] = sh->coeff[k];
      }
    }
    __syncthreads();
    for (int s = wave; s < nsh; s += 4) {
      const int info = __builtin_amdgcn_readfirstlane(st.info[s]);
      const int bf0 = __builtin_amdgcn_readfirstlane(st.bf0[s]);
      const int l = info & 0xff;
      const bool pure = (info >> 8) & 1;
      const int nprim = info >> 16;
      const double x = gx - st.O[0][s], y = gy - st.O[1][s], z = gz - st.O[2][s];
      const double r2 = x * x + y * y + z * z;
      double R = 0., dR = 0.;
      for (int k = 0; k < nprim && k < kCollPrims; ++k) {
        const double a = st.a[k][s], e = st.c[k][s] * exp(-a * r2);
        R += e;
        dR += -2. * a * e;
      }
      if (nprim > kCollPrims) {
        const DevShell* sh = shells + __builtin_amdgcn_readfirstlane(st.ish[s]);
        for (int k = kCollPrims; k < nprim; ++k) {
          const double e = sh->coeff[k] * exp(-sh->alpha[k] * r2);
          R += e;
          dR += -2. * sh->alpha[k] * e;
        }
      }
      body(l, pure, bf0, x, y, z, R, dR);
    }
  }
}

template <bool GRAD, typename T, int LMAX>
__global__ __launch_bounds__(256, coll_waves(LMAX, GRAD, sizeof(T) == 8)) void k_collocation(const WorkItem* __restrict__ blocks,
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
  const unsigned lane = threadIdx.x & 63;
  const int p = wi.a * kPtBlock + (int)lane;
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
  const int64_t m0 = (int64_t)wi.a * kPtBlock;
  // FP32 storage: bfloat16 residual plane of the values after the five
  // FP32 planes (B, dB/dx, dB/dy, dB/dz, Z).  Byte offsets within the
  // task's region are below 2^32 (checked when the chunks are planned); the
  // buffer unit adds them as unsigned 32-bit values.
  TaskStore ts;
  ts.rs = __builtin_amdgcn_make_buffer_rsrc(scratch + boff, 0, -1, 0x00020000);
  ts.vo4 = (int)(lane * 4u);
  ts.vo2 = (int)(lane * 2u);
  ts.lo_base = (unsigned)(5 * plane * (int64_t)sizeof(T));
  __shared__ ShellStage st;
  for_block_shells(t, shells, shell_list, lsh_first, gx, gy, gz, st,
                   [&](int l, bool pure, int bf0, double x, double y, double z, double R,
                       double dR) {
    const int64_t row0 = (int64_t)bf0 * ld + m0;
    if (l == 0) shell_eval<0, GRAD, T>(x, y, z, R, dR, pure, valid, ts, row0, ld, plane);
    else if (l == 1) shell_eval<1, GRAD, T>(x, y, z, R, dR, pure, valid, ts, row0, ld, plane);
    else if (LMAX == 2 || l == 2)
      shell_eval<2, GRAD, T>(x, y, z, R, dR, pure, valid, ts, row0, ld, plane);
    else if (LMAX == 3 || l == 3)
      shell_eval<(LMAX < 3 ? 2 : 3), GRAD, T>(x, y, z, R, dR, pure, valid, ts, row0, ld, plane);
    else
      shell_eval<(LMAX < 4 ? 2 : 4), GRAD, T>(x, y, z, R, dR, pure, valid, ts, row0, ld, plane);
  });
}

// Z = a B + c . grad B for one (task, 64-point block), stored as T (FP32
// for the mixed path, FP64 for the FP64 path), with B and grad B evaluated
// here in FP64 (the in-stage mode that leaves Z to this launch): the
// block's coefficients (a, cx, cy, cz per point) come from the density
// stage, zcoef[block][4][64] with the block numbered as the EXC/NEL
// partials.  Z goes to plane 4 of the task's scratch, one coalesced
// 64-point row per function.  Padded points have
// zero coefficients (Z = 0) and are evaluated at the task's last point.
// Also records the bound table's granule maxima of |B| and |Z|: per lane
// the bits of |x| under unsigned max (every NaN orders above infinity) per
// shell and per current granule; a wave reduces them (clamped to infinity)
// into LDS when it moves to another granule (its shells ascend), and a
// shell that spans two granules counts in both, which can only loosen the
// bounds.  Five waves per SIMD: with the gradient folded along the
// coefficient vector (shell_functions_z) up to f needs 114 registers; at
// five waves 8 spill and the launch is still faster (24 waters, FP32 Z
// 1.30 -> 1.20 ms, FP64 1.63 -> 1.49 ms).
template <bool GGA, int LMAX, typename T>
__global__ __launch_bounds__(256, 5) void k_z_in_stage(const WorkItem* __restrict__ blocks,
                                                      const DevTask* __restrict__ tasks,
                                                      const DevShell* __restrict__ shells,
                                                      const int32_t* __restrict__ shell_list,
                                                      const int32_t* __restrict__ lsh_first,
                                                      const double* __restrict__ px,
                                                      const double* __restrict__ py,
                                                      const double* __restrict__ pz,
                                                      const double* __restrict__ zcoef,
                                                      T* __restrict__ scratch,
                                                      unsigned* __restrict__ tmax_all) {
  const WorkItem wi = blocks[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x, lane = tid & 63;
  const int pl = min(wi.a * kPtBlock + lane, t.npts - 1);
  const double gx = px[t.pt_off + pl], gy = py[t.pt_off + pl], gz = pz[t.pt_off + pl];
  const double* zb = zcoef + (t.red_off + wi.a) * 256;
  const double za = zb[lane], zx = zb[64 + lane], zy = zb[128 + lane], zz = zb[192 + lane];
  const int64_t ld = t.npts_pad;
  const int64_t plane = (int64_t)t.nbe * ld;
  T* Zp = scratch + (sizeof(T) == 4 ? t.b_off_mx : t.b_off) + 4 * plane +
          (int64_t)wi.a * kPtBlock + lane;
  unsigned* tmax = tmax_all ? tmax_all + 2 * t.tmax_off : nullptr;
  __shared__ ShellStage st;
  __shared__ unsigned smax[kGranLds][2];
  if (tid < 2 * kGranLds) (&smax[0][0])[tid] = 0u;
  constexpr unsigned kInfBits = 0x7f800000u, kAbs = 0x7fffffffu;
  int cur = -1;               // granule of the running maxima (uniform)
  unsigned mb = 0u, mz = 0u;  // running maxima of granule cur
  auto flush = [&]() {
    unsigned b = mb, z = mz;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      b = max(b, (unsigned)__shfl_xor((int)b, off));
      z = max(z, (unsigned)__shfl_xor((int)z, off));
    }
    b = min(b, kInfBits);
    z = min(z, kInfBits);
    if (lane == 0) {
      if (cur < kGranLds) {
        atomicMax(&smax[cur][0], b);
        atomicMax(&smax[cur][1], z);
      } else {
        atomicMax(tmax + 2 * cur, b);
        atomicMax(tmax + 2 * cur + 1, z);
      }
    }
  };
  for_block_shells(t, shells, shell_list, lsh_first, gx, gy, gz, st,
                   [&](int l, bool pure, int bf0, double x, double y, double z, double R,
                       double dR) {
    unsigned sb = 0u, sz = 0u;  // this shell's maxima
    auto put = [&](int f, double v, double zv) {
      Zp[(int64_t)(bf0 + f) * ld] = (T)zv;
      sb = max(sb, __float_as_uint((float)v) & kAbs);
      sz = max(sz, __float_as_uint((float)zv) & kAbs);
    };
    if constexpr (GGA) {
      // Z_f = a B_f + c . grad B_f, the gradient folded along c.
      const double A = za * R + dR * (zx * x + zy * y + zz * z);
      shell_functions_z_l<LMAX>(l, x, y, z, R, A, zx, zy, zz, pure, put);
    } else {
      shell_functions_l<LMAX, false, true>(l, x, y, z, R, dR, pure,
                                           [&](int f, double v, double, double, double) {
        put(f, v, za * v);
      });
    }
    if (tmax) {
      const int size = pure ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
      const int g0 = bf0 / kGran, g1 = (bf0 + size - 1) / kGran;  // uniform
      if (g0 != cur) {
        if (cur >= 0) flush();
        cur = g0;
        mb = mz = 0u;
      }
      mb = max(mb, sb);
      mz = max(mz, sz);
      if (g1 != g0) {
        flush();
        cur = g1;
        mb = sb;
        mz = sz;
      }
    }
  });
  if (tmax) {
    if (cur >= 0) flush();
    __syncthreads();
    const int ng = min((t.nbe + kGran - 1) / kGran, kGranLds);
    if (tid < 2 * ng) atomicMax(tmax + tid, smax[tid >> 1][tid & 1]);
  }
}

int launch_z_in_stage(hipStream_t s, const WorkItem* items, int64_t n, const DevTask* tasks,
                      const DevShell* shells, const int32_t* shell_list,
                      const int32_t* lsh_first, const double* px, const double* py,
                      const double* pz, const double* zcoef, void* scratch, int precision,
                      unsigned* tmax, int gga, int lmax) {
  if (n == 0) return XCQ_SUCCESS;
  XCQ_REQUIRE(lmax >= 0 && lmax <= kInStageMaxL, XCQ_ERROR_INVALID_ARGUMENT,
              "in-stage Z: angular momentum out of range");
  auto go = [&](auto kern, auto* scr) {
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(256), 0, s, items, tasks, shells,
                       shell_list, lsh_first, px, py, pz, zcoef, scr, tmax);
  };
  auto pick = [&](auto* scr) {
    using T = std::remove_pointer_t<decltype(scr)>;
    if (gga) {
      if (lmax <= 2) go(k_z_in_stage<true, 2, T>, scr);
      else go(k_z_in_stage<true, 3, T>, scr);
    } else {
      if (lmax <= 2) go(k_z_in_stage<false, 2, T>, scr);
      else go(k_z_in_stage<false, 3, T>, scr);
    }
  };
  if (precision == XCQ_PREC_FP64) pick(static_cast<double*>(scratch));
  else pick(static_cast<float*>(scratch));
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// One block per (task, 64-point block) item; lmax is the highest angular
// momentum among the basis shells (the kernel evaluates l <= max(lmax, 2)).
template <typename T>
static int launch_collocation(hipStream_t s, const WorkItem* items, int64_t n,
                              const DevTask* tasks, const DevShell* shells,
                              const int32_t* shell_list, const int32_t* lsh_first,
                              const double* px, const double* py, const double* pz, T* scratch,
                              int deriv, int lmax) {
  if (n == 0) return XCQ_SUCCESS;
  XCQ_REQUIRE(lmax >= 0 && lmax <= kMaxL, XCQ_ERROR_INVALID_ARGUMENT,
              "collocation: angular momentum out of range");
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)n), dim3(256), 0, s, items, tasks, shells,
                       shell_list, lsh_first, px, py, pz, scratch);
  };
  if (deriv) {
    if (lmax <= 2) go(k_collocation<true, T, 2>);
    else if (lmax == 3) go(k_collocation<true, T, 3>);
    else go(k_collocation<true, T, 4>);
  } else {
    if (lmax <= 2) go(k_collocation<false, T, 2>);
    else if (lmax == 3) go(k_collocation<false, T, 3>);
    else go(k_collocation<false, T, 4>);
  }
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_collocation_fp64(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz,
                            double* scratch, int deriv, int lmax) {
  return launch_collocation<double>(s, items, n, tasks, shells, shell_list, lsh_first, px, py,
                                    pz, scratch, deriv, lmax);
}

// Mixed precision: values evaluated in FP64, stored as FP32 plus the value
// plane's bfloat16 residual.
int launch_collocation_fp32(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first,
                            const double* px, const double* py, const double* pz,
                            float* scratch, int deriv, int lmax) {
  return launch_collocation<float>(s, items, n, tasks, shells, shell_list, lsh_first, px, py,
                                   pz, scratch, deriv, lmax);
}

}  // namespace xcq
