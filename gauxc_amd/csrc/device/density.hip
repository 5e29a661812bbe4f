// Density stage of the XC path: X = fac B P_sub, rho / grad rho, the
// functional with weight folding, the EXC/NEL block sums and the Z matrix,
// in one launch.
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
// and gives the Z coefficients a = w vrho / 2, c = 2 w vsigma grad rho and
// the block's (w eps rho, w rho) sums; the tail then writes
//   Z = a B + c . grad B        (reference_local_host_work_driver.cxx:678-713)
// for the block's points.  X never reaches HBM.
//
// The K loop is software-pipelined: step s+1's global loads are issued into
// registers before step s's matrix instructions, and the LDS stage is double
// buffered so one barrier per step suffices.  Out-of-range rows/columns are
// read at a clamped (valid) index and zeroed by a select, so the loads are
// branch-free.  K indices are uniform across a wave, which makes the
// basis-map lookups scalar loads.
//
// Both kernels are templated on the function tile NT (32 or 64): a task's
// rows are covered by ceil(nbe / NT) tiles, so small tasks are not padded to
// 64.  A launch either runs one workgroup per (task, block) item or, in
// queued mode, one persistent workgroup per queue of items.
//
// Basis evaluation mode (template argument INSTAGE): 0 reads B and grad B
// from the planes the collocation stage stored; in the in-stage modes the
// fold evaluates them at the block's points from the shell parameters
// (fold_in_stage) and Z is formed either in the same kernel (mode 1,
// z_in_stage) or by a separate launch from the Z coefficients this kernel
// leaves (mode 3, k_z_in_stage in collocation.hip).
#include "internal.hpp"
#include "functional.hpp"
#include "shell_eval.hpp"

#include <climits>
#include <type_traits>

namespace xcq {

using v4d = double __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

namespace {

constexpr int kPad64 = 68;  // FP64 stage row (doubles)





__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }

__device__ __forceinline__ double bf16_val(unsigned short u) {
  return (double)__builtin_bit_cast(float, (unsigned)u << 16);
}

// Value plane of FP32 storage: FP32 value + bfloat16 residual (plane 5 of
// the task's scratch, 2 bytes per element).  nullptr for FP64 storage.
template <typename T>
__device__ __forceinline__ const unsigned short* lo_plane(const T* Bm, int plane) {
  if constexpr (sizeof(T) == 4)
    return reinterpret_cast<const unsigned short*>(Bm + 5 * (int64_t)plane);
  else
    return nullptr;
}

// Folds one X^T tile (rows = functions, columns = points) into the
// per-point partials in LDS.  frow(r) is the tile row (function) of
// accumulator register r of row block bi for this lane; the column (point)
// of lane l in column block bj is wn*32 + bj*16 + (l & 15).
template <typename T, int BI, bool GGA, typename FRow>
__device__ __forceinline__ void fold_tile(const double (&x)[BI][2][4], FRow frow, int m0, int n0,
                                          int wm, int wn, int lane, int nbe, int ld, int plane,
                                          const T* Bm, double (*dens)[64][4]) {
#pragma unroll
  for (int bj = 0; bj < 2; ++bj) {
    const int pl = wn * 32 + bj * 16 + (lane & 15);
    const int m = m0 + pl;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    const unsigned short* lo = lo_plane(Bm, plane);
#pragma unroll
    for (int bi = 0; bi < BI; ++bi) {
      T v0[4], v1[4], v2[4], v3[4];
      unsigned short l0[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int j = n0 + wm * 16 * BI + bi * 16 + frow(r);
        const int o = min(j, nbe - 1) * ld + m;
        v0[r] = Bm[o];
        if constexpr (sizeof(T) == 4) l0[r] = lo[o];
        if constexpr (GGA) {
          v1[r] = Bm[plane + o];
          v2[r] = Bm[2 * plane + o];
          v3[r] = Bm[3 * plane + o];
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        // Rows past nbe carry X = 0 (their P_sub column was zeroed), so the
        // clamped B values they read contribute nothing.
        const double xv = x[bi][bj][r];
        double b = (double)v0[r];
        // FP32 storage: X was formed from the FP32 part B32 of B alone, so
        // rho = B^T 2P B misses res^T X (res = B - B32, P symmetric), a
        // first-order term that does not average out (it biased EXC by
        // about -3e-10 relative): fold B32 + 2 res.
        if constexpr (sizeof(T) == 4) b += 2.0 * bf16_val(l0[r]);
        s0 += b * xv;
        if constexpr (GGA) {
          s1 += (double)v1[r] * xv;
          s2 += (double)v2[r] * xv;
          s3 += (double)v3[r] * xv;
        }
      }
    }
    // Sum over the four 16-lane groups (different functions, same point).
#pragma unroll
    for (int off = 16; off < 64; off <<= 1) {
      s0 += __shfl_xor(s0, off);
      if constexpr (GGA) {
        s1 += __shfl_xor(s1, off);
        s2 += __shfl_xor(s2, off);
        s3 += __shfl_xor(s3, off);
      }
    }
    if (lane < 16) {
      dens[wm][pl][0] += s0;
      if constexpr (GGA) {
        dens[wm][pl][1] += s1;
        dens[wm][pl][2] += s2;
        dens[wm][pl][3] += s3;
      }
    }
  }
}

// Fold through LDS (mixed kernel, tiles of up to 64 functions): the X^T
// tile (rows = functions, FP64) is written over the operand ring, then
// thread t takes points 4(t%16)..4(t%16)+3 of rows t/16, t/16 + 16, ... and
// reads B, its residual and grad B there with 16-byte vector loads (four
// points per load, against one in the accumulator layout).  Partials are
// summed over the wave's four row groups by shuffles; lanes 0..15 of wave w
// add them to dens[w].  The caller has passed a barrier after its last
// reads of the ring; this function ends after its last reads of it.
constexpr int kXLd = 66;  // LDS row stride of the X^T tile in doubles

template <int NT, int BI, bool GGA>
__device__ __forceinline__ void fold_tile_lds(const double (&x)[BI][2][4], double* xt, int m0,
                                              int n0, int wm, int wn, int lane, int tid, int nbe,
                                              int ld, int plane, const float* Bm,
                                              double (*dens)[64][4]) {
#pragma unroll
  for (int bi = 0; bi < BI; ++bi)
#pragma unroll
    for (int bj = 0; bj < 2; ++bj)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        xt[(wm * 16 * BI + bi * 16 + 4 * (lane >> 4) + r) * kXLd + wn * 32 + bj * 16 +
           (lane & 15)] = x[bi][bj][r];
  __syncthreads();
  const int q4 = 4 * (tid & 15);
  const int wave = tid >> 6;
  double s[4][4];
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int c = 0; c < 4; ++c) s[e][c] = 0.;
  const unsigned short* lo = reinterpret_cast<const unsigned short*>(Bm + 5 * (int64_t)plane);
#pragma unroll
  for (int k = 0; k < NT / 16; ++k) {
    const int f = (tid >> 4) + 16 * k;
    // Rows past nbe carry X = 0; their (clamped) B values contribute nothing.
    const int o = min(n0 + f, nbe - 1) * ld + m0 + q4;
    const float4 b = *reinterpret_cast<const float4*>(Bm + o);
    const uint2 l = *reinterpret_cast<const uint2*>(lo + o);
    float4 gxv, gyv, gzv;
    if constexpr (GGA) {
      gxv = *reinterpret_cast<const float4*>(Bm + plane + o);
      gyv = *reinterpret_cast<const float4*>(Bm + 2 * plane + o);
      gzv = *reinterpret_cast<const float4*>(Bm + 3 * plane + o);
    }
    const double* xr = xt + f * kXLd + q4;
    const float bb[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const unsigned w = e < 2 ? l.x : l.y;
      // B32 + 2 res: X lacks the residual (see fold_tile).
      const double bv = (double)bb[e] + 2.0 * (double)__builtin_bit_cast(
                                                  float, (e & 1) ? (w & 0xffff0000u) : (w << 16));
      const double xv = xr[e];
      s[e][0] += bv * xv;
      if constexpr (GGA) {
        const float gxa[4] = {gxv.x, gxv.y, gxv.z, gxv.w};
        const float gya[4] = {gyv.x, gyv.y, gyv.z, gyv.w};
        const float gza[4] = {gzv.x, gzv.y, gzv.z, gzv.w};
        s[e][1] += (double)gxa[e] * xv;
        s[e][2] += (double)gya[e] * xv;
        s[e][3] += (double)gza[e] * xv;
      }
    }
  }
  // Lanes l, l ^ 16, l ^ 32, l ^ 48 hold the same points (other rows).
#pragma unroll
  for (int off = 16; off < 64; off <<= 1)
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int c = 0; c < (GGA ? 4 : 1); ++c) s[e][c] += __shfl_xor(s[e][c], off);
  if ((tid & 63) < 16)
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int c = 0; c < (GGA ? 4 : 1); ++c) dens[wave][q4 + e][c] += s[e][c];
}

// Functional on the block's 64 points (threads 0..63 = wave 0).  Writes the
// Z coefficients (a, cx, cy, cz) of each point to zc and the block's
// (EXC, NEL) partial sums to red.
template <int ND = 2>
__device__ __forceinline__ void finish_points(const DevTask& t, int pb, int tid,
                                              double (*dens)[64][4], double (*zc)[64],
                                              const double* __restrict__ weights, int func,
                                              int gga, double* __restrict__ red) {
  if (tid >= 64) return;
  const int p = pb * kPtBlock + tid;
  double rho = 0., gx = 0., gy = 0., gz = 0.;
#pragma unroll
  for (int d = 0; d < ND; ++d) {
    rho += dens[d][tid][0];
    if (gga) {
      gx += dens[d][tid][1];
      gy += dens[d][tid][2];
      gz += dens[d][tid][3];
    }
  }
  gx *= 2.0;
  gy *= 2.0;
  gz *= 2.0;
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
  zc[0][tid] = za;
  zc[1][tid] = zx;
  zc[2][tid] = zy;
  zc[3][tid] = zz;
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * (t.red_off + pb)] = e;
    red[2 * (t.red_off + pb) + 1] = n;
  }
}

// Four consecutive elements of a plane (16-byte aligned for FP32, 32 for FP64).
template <typename T>
using Vec4 = T __attribute__((ext_vector_type(4)));

// Z = a B + c . grad B for the block's 64 points, every function row.
// Thread (q = tid & 15, g = tid >> 4) owns points m0 + 4q .. 4q + 3 of rows
// g, g + 16, ...  Each pass loads kZRows rows with 4-wide vector loads
// before any of them is used, so every thread keeps up to 5 kZRows loads in
// flight instead of waiting on each row in turn (4 rows of FP32, 2 of FP64,
// which fit the register budget of the GEMM phase).  Row offsets are clamped
// to a valid row; rows past nbe are not stored.
// Also records, per granule of 32 functions, max |B| and max |Z| over the
// block's points (B with its residual; float bits, so unsigned max orders
// them): LDS atomics into smax, then one device atomic per granule into the
// task's bound table (direct device atomics when the task has more granules
// than smax holds).  The B^T Z stage skips a tile pair when
// npts * (max|B_i| max|Z_j| + max|Z_i| max|B_j|) is below the drop tolerance.
template <typename T, bool GGA, int kZRows = sizeof(T) == 4 ? 4 : 2>
__device__ __forceinline__ void write_z(const DevTask& t, int m0, int tid, double (*zc)[64],
                                        T* Bm, unsigned (*smax)[2], unsigned* tmax) {
  const int q = tid & 15, g = tid >> 4;
  const int ld = t.npts_pad, nbe = t.nbe, plane = nbe * ld;
  double a[4], cx[4], cy[4], cz[4];
#pragma unroll
  for (int e = 0; e < 4; ++e) {
    a[e] = zc[0][4 * q + e];
    cx[e] = zc[1][4 * q + e];
    cy[e] = zc[2][4 * q + e];
    cz[e] = zc[3][4 * q + e];
  }
  T* B = Bm + m0 + 4 * q;
  const unsigned short* lo = lo_plane(Bm, plane);
  if constexpr (sizeof(T) == 4) lo += m0 + 4 * q;
  for (int j0 = g; j0 < nbe; j0 += 16 * kZRows) {
    Vec4<T> b[kZRows], x[kZRows], y[kZRows], z[kZRows];
    uint2 l[kZRows];
#pragma unroll
    for (int u = 0; u < kZRows; ++u) {
      const int o = min(j0 + 16 * u, nbe - 1) * ld;
      b[u] = *reinterpret_cast<const Vec4<T>*>(B + o);
      if constexpr (sizeof(T) == 4) l[u] = *reinterpret_cast<const uint2*>(lo + o);
      if constexpr (GGA) {
        x[u] = *reinterpret_cast<const Vec4<T>*>(B + plane + o);
        y[u] = *reinterpret_cast<const Vec4<T>*>(B + 2 * plane + o);
        z[u] = *reinterpret_cast<const Vec4<T>*>(B + 3 * plane + o);
      }
    }
#pragma unroll
    for (int u = 0; u < kZRows; ++u) {
      const int j = j0 + 16 * u;
      if (j >= nbe) break;
      Vec4<T> out;
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        double bv = (double)b[u][e];
        if constexpr (sizeof(T) == 4) {
          // bfloat16 residuals of points 2h, 2h + 1 are the low and high
          // halves of word h.
          const unsigned w = e < 2 ? l[u].x : l[u].y;
          bv += (double)__builtin_bit_cast(float, (e & 1) ? (w & 0xffff0000u) : (w << 16));
        }
        double zv = a[e] * bv;
        if constexpr (GGA) zv += cx[e] * (double)x[u][e] + cy[e] * (double)y[u][e] + cz[e] * (double)z[u][e];
        out[e] = (T)zv;
      }
      *reinterpret_cast<Vec4<T>*>(B + 4 * plane + j * ld) = out;
      if (tmax) {
        // A NaN counts as unbounded, so screening never hides it.
        float mb = 0.f, mz = 0.f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float fb = fabsf((float)b[u][e]), fz = fabsf((float)out[e]);
          mb = fmaxf(mb, fb != fb ? INFINITY : fb);
          mz = fmaxf(mz, fz != fz ? INFINITY : fz);
        }
        // The 16 lanes of this row (same g) hold its 64 points.
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) {
          mb = fmaxf(mb, __shfl_xor(mb, off));
          mz = fmaxf(mz, __shfl_xor(mz, off));
        }
        if (q == 0) {
          const int gr = j / kGran;
          if (gr < kGranLds) {
            atomicMax(&smax[gr][0], __float_as_uint(mb));
            atomicMax(&smax[gr][1], __float_as_uint(mz));
          } else {
            atomicMax(tmax + 2 * gr, __float_as_uint(mb));
            atomicMax(tmax + 2 * gr + 1, __float_as_uint(mz));
          }
        }
      }
    }
  }
  if (tmax) {
    __syncthreads();
    const int ng = min((nbe + kGran - 1) / kGran, kGranLds);
    if (tid < 2 * ng) atomicMax(tmax + tid, smax[tid >> 1][tid & 1]);
  }
}

// ---- basis functions evaluated inside the density stage -------------------
// With the in-stage evaluation the collocation stage stores only the value
// plane (and its residual), which the X product and the B^T Z stage read;
// the fold and Z evaluate B and grad B of the block's points from the shell
// parameters where they use them (FP64), so the three gradient planes are
// neither written nor read twice.  Wave w takes the shells s = w (mod 4)
// of a function range; lane = point of the 64-point block, so a shell's
// parameters are wave-uniform (scalar loads) and each function's values
// form one 64-point row.

// A shell as one wave evaluates it: its first local function, its index
// and the compact record's fields, all wave-uniform.
struct ShellU {
  int bf0, ish, info;
  double ox, oy, oz, a0, c0;
};

__device__ __forceinline__ double readlane_d(double v, int j) {
  const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = __builtin_amdgcn_readlane((unsigned)b, j);
  const unsigned hi = __builtin_amdgcn_readlane((unsigned)(b >> 32), j);
  return __builtin_bit_cast(double, (unsigned long long)hi << 32 | lo);
}

// Calls sink(f, v, gx, gy, gz) for the functions of shell u at the lane's
// point (rx, ry, rz): R = sum c e^{-a r2}, dR = -2 sum a c e^{-a r2}; the
// primitives past the first come from the full shell record.
// Displacement (x, y, z) of the lane's point from shell u's centre and the
// radial part R = sum c e^{-a r2}, dR = -2 sum a c e^{-a r2}; the
// primitives past the first come from the full shell record.
__device__ __forceinline__ void shell_radial(const ShellArgs& sa, const ShellU& u, double rx,
                                             double ry, double rz, double& x, double& y,
                                             double& z, double& R, double& dR) {
  const int nprim = u.info >> 16;
  x = rx - u.ox;
  y = ry - u.oy;
  z = rz - u.oz;
  const double r2 = x * x + y * y + z * z;
  const double e0 = u.c0 * exp(-u.a0 * r2);
  R = e0;
  dR = -2. * u.a0 * e0;
  if (nprim > 1) {
    const DevShell* sh = sa.shells + u.ish;
    for (int k = 1; k < nprim; ++k) {
      const double a = sh->alpha[k], e = sh->coeff[k] * exp(-a * r2);
      R += e;
      dR += -2. * a * e;
    }
  }
}

template <bool GRAD, typename Sink>
__device__ __forceinline__ void eval_shell(const ShellArgs& sa, const ShellU& u, double rx,
                                           double ry, double rz, Sink&& sink) {
  const int l = u.info & 0xff;
  const bool pure = (u.info >> 8) & 1;
  double x, y, z, R, dR;
  shell_radial(sa, u, rx, ry, rz, x, y, z, R, dR);
  shell_functions_l<kInStageMaxL, GRAD, true>(l, x, y, z, R, dR, pure, sink);
}

// Calls body(ShellU) for the task's shells that intersect local functions
// [.., nhi) from s_first (the shell containing the range's first
// function), this wave's share (every fourth); returns the shell
// containing nhi.  Lane i loads the map entries and the compact record of
// shell s_first + i (64 at a time, one vector load each), which the wave
// then broadcasts shell by shell.  The local first-function map is
// ascending, so the shells starting below nhi are a prefix counted by a
// ballot.
template <typename Body>
__device__ __forceinline__ int for_range_shells(const DevTask& t, const ShellArgs& sa,
                                                int s_first, int nhi, int wave, int lane,
                                                Body&& body) {
  int next = s_first;
  for (int sc = s_first; sc < t.nshells; sc += 64) {
    const int s = sc + lane;
    const bool ok = s < t.nshells;
    const int f0 = ok ? sa.lsh[t.lsh_off + s] : INT_MAX;
    const int ish = ok ? sa.shell_list[t.sh_off + s] : 0;
    const uint64_t in = __builtin_amdgcn_ballot_w64(f0 < nhi);
    const uint64_t le = __builtin_amdgcn_ballot_w64(f0 <= nhi);
    const int cnt = __popcll(in);
    if (le) next = sc + __popcll(le) - 1;
    if (wave < cnt) {
      const DevShellC c = sa.shc[ish];
      for (int j = wave; j < cnt; j += 4) {
        ShellU u;
        u.bf0 = __builtin_amdgcn_readlane(f0, j);
        u.ish = __builtin_amdgcn_readlane(ish, j);
        u.info = __builtin_amdgcn_readlane(c.info, j);
        u.ox = readlane_d(c.ox, j);
        u.oy = readlane_d(c.oy, j);
        u.oz = readlane_d(c.oz, j);
        u.a0 = readlane_d(c.a0, j);
        u.c0 = readlane_d(c.c0, j);
        body(u);
      }
    }
    if (cnt < 64) break;
  }
  return next;
}

// Folds rows [nlo, nhi) of the X^T tile in LDS (row n - nlo, column =
// point) into this lane's partials r: rho += B X, grad rho / 2 += grad B X.
// With B_f = R Y_f and grad B_f = R grad Y_f + dR r Y_f, a shell adds
// R sy to rho and R sg + dR r sy to grad rho / 2, where sy = sum Y_f X_f
// and sg = sum grad Y_f X_f over its functions in the range: the functions
// are formed from their angular factors alone.
// RES32: X was formed from the FP32-rounded value plane (mixed precision);
// rho then also takes sum (B - fl32(B)) X, the first-order term the X
// operand misses (see fold_tile).
template <bool GGA, bool RES32 = false>
__device__ __forceinline__ int fold_in_stage(const DevTask& t, const ShellArgs& sa, int s_first,
                                             int nlo, int nhi, int wave, int lane,
                                             const double* xt, double rx, double ry, double rz,
                                             double (&r)[4]) {
  return for_range_shells(t, sa, s_first, nhi, wave, lane, [&](const ShellU& u) {
    const int l = u.info & 0xff;
    const bool pure = (u.info >> 8) & 1;
    double x, y, z, R, dR;
    shell_radial(sa, u, rx, ry, rz, x, y, z, R, dR);
    double sy = 0., sgx = 0., sgy = 0., sgz = 0., sres = 0.;
    shell_functions_l<kInStageMaxL, GGA, true, true>(
        l, x, y, z, R, dR, pure, [&](int f, double v, double gx, double gy, double gz) {
          const int n = u.bf0 + f;  // uniform
          if (n >= nlo && n < nhi) {
            const double xv = xt[(n - nlo) * kXLd + lane];
            sy += v * xv;
            if constexpr (RES32) {
              const double b = R * v;
              sres += (b - (double)(float)b) * xv;
            }
            if constexpr (GGA) {
              sgx += gx * xv;
              sgy += gy * xv;
              sgz += gz * xv;
            }
          }
        });
    r[0] += R * sy;
    if constexpr (RES32) r[0] += sres;
    if constexpr (GGA) {
      const double d = dR * sy;
      r[1] += R * sgx + d * x;
      r[2] += R * sgy + d * y;
      r[3] += R * sgz + d * z;
    }
  });
}

// Wave-wide max of v (every lane gets it).
__device__ __forceinline__ unsigned wave_max(unsigned v) {
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) v = max(v, (unsigned)__shfl_xor((int)v, off));
  return v;
}

// Z = a B + c . grad B of the block's points for every function of the
// task, with B and grad B evaluated here; stored in FP32 (plane 4), one
// coalesced 64-point row per function.  Padded points have zero
// coefficients (so Z = 0) and are evaluated at the task's last point.
// Also records the granule maxima of |B| and |Z| as write_z does, per
// shell: each lane keeps the bits of |x| under unsigned max (which orders
// every NaN above infinity) for the shell and for the current granule; the
// wave reduces and clamps them to infinity when it moves to another granule
// (each wave meets the granules in ascending order).  A shell that spans
// two granules counts in both, which can only loosen the bounds.
template <bool GGA>
__device__ __forceinline__ void z_in_stage(const DevTask& t, const ShellArgs& sa, int m0,
                                           int wave, int lane, int tid, double rx, double ry,
                                           double rz, double (*zc)[64], float* Bm,
                                           unsigned (*smax)[2], unsigned* tmax) {
  const int ld = t.npts_pad, nbe = t.nbe;
  const int64_t plane = (int64_t)nbe * ld;
  const double za = zc[0][lane], zx = zc[1][lane], zy = zc[2][lane], zz = zc[3][lane];
  float* Zp = Bm + 4 * plane + m0 + lane;
  constexpr unsigned kInfBits = 0x7f800000u, kAbs = 0x7fffffffu;
  int cur = -1;               // granule of the running maxima (uniform)
  unsigned mb = 0u, mz = 0u;  // running maxima of granule cur
  auto flush = [&]() {
    const unsigned b = min(wave_max(mb), kInfBits), z = min(wave_max(mz), kInfBits);
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
  for_range_shells(t, sa, 0, nbe, wave, lane, [&](const ShellU& u) {
    unsigned sb = 0u, sz = 0u;  // this shell's maxima
    eval_shell<GGA>(sa, u, rx, ry, rz, [&](int f, double v, double gx, double gy, double gz) {
      double zv = za * v;
      if constexpr (GGA) zv += zx * gx + zy * gy + zz * gz;
      const float zf = (float)zv;
      Zp[(int64_t)(u.bf0 + f) * ld] = zf;
      sb = max(sb, __float_as_uint((float)v) & kAbs);
      sz = max(sz, __float_as_uint(zf) & kAbs);
    });
    if (tmax) {
      const int l = u.info & 0xff;
      const int size = (u.info >> 8) & 1 ? 2 * l + 1 : (l + 1) * (l + 2) / 2;
      const int g0 = u.bf0 / kGran, g1 = (u.bf0 + size - 1) / kGran;  // uniform
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
    const int ng = min((nbe + kGran - 1) / kGran, kGranLds);
    if (tid < 2 * ng) atomicMax(tmax + tid, smax[tid >> 1][tid & 1]);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// FP64 (v_mfma_f64_16x16x4_f64; accumulator register r of lane l holds row
// 4r + l/16).  Waves: wm = function half (NT/2 rows = BI blocks of 16),
// wn = point half of the 64-point block.  Loader roles: thread
// (c = tid & 63, q = tid >> 6) loads K rows 4q..4q+3 of column c of the B^T
// stage (point c) and, for c < NT, of the P_sub stage (function n0 + c).
// INSTAGE = 3: the fold evaluates B and grad B itself (fold_in_stage, from
// the X^T tile written over the stages) and the block's Z coefficients go
// to zbuf for the separate Z launch; 0: the fold and Z read the planes the
// collocation stage stored.  HIOCC (in-stage, 64-wide tiles): a register
// budget for three workgroups per CU, which spills a few registers of the
// fold and still hides more latency.
template <int NT, bool GGA, int INSTAGE, bool HIOCC = false>
__global__ __launch_bounds__(256, HIOCC ? 3 : 2) void k_density_f64(
    QueueArgs qa, const DevTask* __restrict__ tasks, const int32_t* __restrict__ bfmap,
    const double* __restrict__ Pall, double* __restrict__ scratch,
    const double* __restrict__ weights, int func, double fac, int want_z,
    double* __restrict__ red, unsigned* __restrict__ tmax, ShellArgs sa,
    double* __restrict__ zbuf, double* __restrict__ dens_out, int x_plane) {
  constexpr int BI = NT / 32;
  constexpr int kPtD = 2 * 16 * (NT + 4), kStageD = kPtD + 2 * 16 * kPad64;
  static_assert(INSTAGE == 0 || NT * kXLd <= kStageD, "X^T tile must fit the stages");
  __shared__ double stage[kStageD];
  // [buf][k][function]: A operand (P_sub); [buf][k][point]: B operand (B^T)
  double(*Pt)[16][NT + 4] = reinterpret_cast<double(*)[16][NT + 4]>(stage);
  double(*Bt)[16][kPad64] = reinterpret_cast<double(*)[16][kPad64]>(stage + kPtD);
  __shared__ double dens[4][64][4];
  __shared__ double zc[4][64];
  __shared__ unsigned smax[kGranLds][2];
  const int tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int lc = lane, lq = wave;  // loader column / K quad
  auto frow = [&](int r) { return 4 * r + (lane >> 4); };

  XCQ_FOR_EACH_ITEM(qa, it) {
    const int64_t t0 = qa.tstamp ? wall_clock64() : 0;
    const WorkItem wi = qa.items[it];
    const DevTask t = tasks[wi.task];
    const int m0 = wi.a * kPtBlock;
    const int ld = t.npts_pad;
    const int nbe = t.nbe;
    const int plane = nbe * ld;
    const int64_t ldp = t.ldp;
    const double* P = Pall + t.p_off;
    double* Bm = scratch + t.b_off;
    const int32_t* bf = bfmap + t.bf_off;
    for (int i = tid; i < 4 * 64 * 4; i += 256) (&dens[0][0][0])[i] = 0.;
    if (tid < 2 * kGranLds) (&smax[0][0])[tid] = 0u;
    __syncthreads();
    const double* bcol = Bm + m0 + lc;
    const int nks = (nbe + 15) / 16;
    // In-stage fold: this lane's point (padded points take the task's last
    // point: finite values, which meet zero X), the shell cursor and this
    // wave's density partials.
    double rx = 0., ry = 0., rz = 0.;
    if constexpr (INSTAGE != 0) {
      const int pl = min(m0 + lane, t.npts - 1);
      rx = sa.px[t.pt_off + pl];
      ry = sa.py[t.pt_off + pl];
      rz = sa.pz[t.pt_off + pl];
    }
    int s_cur = 0;
    double part[4] = {0., 0., 0., 0.};

    for (int n0 = 0; n0 < nbe; n0 += NT) {
      const int ncol = n0 + lc;
      const bool colok = lc < NT && ncol < nbe;
      const double* Pc = P + bf[min(ncol, nbe - 1)];  // column bf[n] of P (P symmetric)
      // Prefetch registers hold raw loads; the masks and the factor are
      // applied when they are staged, so no load is waited for before the
      // matrix instructions of the current step are issued.
      double rb[4], rp[4];
      auto load = [&](int k0) {
        int rowp[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) rowp[e] = bf[min(k0 + 4 * lq + e, nbe - 1)];  // scalar
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const int kc = min(k0 + 4 * lq + e, nbe - 1);
          rb[e] = bcol[kc * ld];
          rp[e] = Pc[(int64_t)rowp[e] * ldp];
        }
      };
      v4d acc[BI][2];
#pragma unroll
      for (int i = 0; i < BI; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = v4d{0., 0., 0., 0.};
      load(0);
      for (int s = 0; s < nks; ++s) {
        const int buf = s & 1;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const bool kok = 16 * s + 4 * lq + e < nbe;
          Bt[buf][4 * lq + e][lc] = kok ? rb[e] : 0.;
          if (lc < NT) Pt[buf][4 * lq + e][lc] = (kok && colok) ? fac * rp[e] : 0.;
        }
        __syncthreads();
        if (s + 1 < nks) load(16 * (s + 1));
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const int kr = ks * 4 + (lane >> 4);
          double a[BI];
#pragma unroll
          for (int bi = 0; bi < BI; ++bi) a[bi] = Pt[buf][kr][wm * 16 * BI + bi * 16 + (lane & 15)];
          const double b0 = Bt[buf][kr][wn * 32 + (lane & 15)];
          const double b1 = Bt[buf][kr][wn * 32 + 16 + (lane & 15)];
#pragma unroll
          for (int bi = 0; bi < BI; ++bi) {
            acc[bi][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b0, acc[bi][0], 0, 0, 0);
            acc[bi][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[bi], b1, acc[bi][1], 0, 0, 0);
          }
        }
      }
      if ((INSTAGE == 0 || !HIOCC) && x_plane >= 0) {
        // The X tile itself into plane x_plane (row = function, column =
        // point), for callers that contract it again (the EXC gradient; not
        // in the three-workgroups-per-CU configuration, whose register
        // budget it would overrun).
        double* xo = Bm + (int64_t)x_plane * plane + m0;
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int row = n0 + wm * 16 * BI + bi * 16 + frow(r);
            if (row < nbe) {
#pragma unroll
              for (int bj = 0; bj < 2; ++bj)
                xo[(int64_t)row * ld + wn * 32 + bj * 16 + (lane & 15)] = acc[bi][bj][r];
            }
          }
      }
      if constexpr (INSTAGE != 0) {
        double* xt = stage;
        __syncthreads();  // every wave is done with the last K step's stage
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              xt[(wm * 16 * BI + bi * 16 + frow(r)) * kXLd + wn * 32 + bj * 16 + (lane & 15)] =
                  acc[bi][bj][r];
        __syncthreads();
        s_cur = fold_in_stage<GGA>(t, sa, s_cur, n0, min(n0 + NT, nbe), wave, lane, xt, rx, ry,
                                   rz, part);
      } else {
        double x[BI][2][4];
#pragma unroll
        for (int i = 0; i < BI; ++i)
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) x[i][j][r] = acc[i][j][r];
        fold_tile<double, BI, GGA>(x, frow, m0, n0, wm, wn, lane, nbe, ld, plane, Bm, dens);
      }
      // The next tile's first stage write targets buffer 0, last read at step
      // nks-2 (or nks-1 when nks == 1), or the X^T tile; all waves pass this
      // barrier after it.
      __syncthreads();
    }
    if constexpr (INSTAGE != 0) {
#pragma unroll
      for (int c = 0; c < 4; ++c) dens[wave][lane][c] = part[c];
      __syncthreads();
    }
    if (dens_out) {
      // One spin component of a UKS evaluation: the block's raw sums
      // (sum B X, sum dB/dx X, ...) per point, [block][4][64]; the
      // polarised functional runs once both components are in.
      constexpr int ND = INSTAGE != 0 ? 4 : 2;
      if (tid < 64) {
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          double v = 0.;
#pragma unroll
          for (int d = 0; d < ND; ++d) v += dens[d][tid][c];
          dens_out[(t.red_off + wi.a) * 256 + c * 64 + tid] = v;
        }
      }
    } else {
      finish_points<INSTAGE != 0 ? 4 : 2>(t, wi.a, tid, dens, zc, weights, func, GGA, red);
      if (want_z) {
        __syncthreads();
        if constexpr (INSTAGE != 0)
          zbuf[(t.red_off + wi.a) * 256 + tid] = (&zc[0][0])[tid];
        else
          write_z<double, GGA>(t, m0, tid, zc, Bm, smax, tmax ? tmax + 2 * t.tmax_off : nullptr);
      }
    }
    stamp_item(qa, it, t0);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 operands on v_mfma_f32_16x16x4_f32 (lane l holds
// A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane l is row
// 4(l/16) + r), FP32 accumulation inside one 32-deep K block (eight chained
// MFMAs), FP64 across blocks.  The FP32 matrix cores round to nearest on
// every accumulation; the BF16 ones drop the low bits of the smaller
// products of a dot product, a bias towards zero that compounds into the
// integrated density (measured on gfx950: -0.6 ulp per 32-term BF16 dot
// product of positive terms spanning 20 binades, -0.004 ulp for the FP32
// MFMA on the same data).  The B^T operand is the FP32 part of the value
// plane; its bfloat16 residual enters the rho fold and Z.  P is gathered
// from an FP32 copy; the factor fac (a power of two) scales the
// accumulated X, which is exact.
//
// Operand staging goes straight from global memory to LDS
// (global_load_lds_dword: lane i of a wave writes dword i of a 256-byte LDS
// row), through a ring of kStages K blocks, so up to kStages - 1 blocks of
// loads are in flight while the matrix cores work on the current one, and
// no register holds staged data.  Thread (lane c, wave g) issues, per K block,
// the eight B^T rows 8g..8g+7 (lane = point) and the eight P_sub rows
// (lane = function c < NT).  P rows past nbe and functions past nbe are
// read from a zero row, B^T rows past nbe at a clamped valid row (finite,
// multiplied by those zero P rows).

// Two configurations of the kernel: the default one keeps three K blocks in
// the ring at two workgroups per CU; the high-occupancy one (HIOCC) keeps
// two blocks in a smaller ring and fits three workgroups per CU (LDS
// 41 KiB, at most 168 registers), which hides the fold's and Z's memory
// latency better on small tasks.
// 128-function tiles (large tasks) keep a two-block ring at two workgroups
// per CU (60 KiB of LDS; 76 KiB in the in-stage modes, whose buffer holds
// the whole X^T tile).
template <int NT, bool HIOCC>
constexpr int ring_stages() { return (HIOCC || NT == 128) ? 2 : 3; }
constexpr int density_waves(int nt, bool hiocc) { return (hiocc && nt != 128) ? 3 : 2; }

__device__ float g_zero_row[64];  // never written: zero-initialised device memory

// Waits until at most n of this wave's vector memory operations are
// outstanding, then joins the workgroup barrier.  The LDS rows written by
// global_load_lds are complete once the issuing wave's counter has passed
// them (vector memory operations complete in order), so after the barrier
// every wave's loads of the awaited block are in LDS.  One asm statement
// with a memory clobber keeps the compiler from moving LDS accesses across
// it.
__device__ __forceinline__ void wait_loads_barrier(int n) {
  switch (n) {
    case 0: asm volatile("s_waitcnt vmcnt(0)\n\ts_barrier" ::: "memory"); break;
    case 16: asm volatile("s_waitcnt vmcnt(16)\n\ts_barrier" ::: "memory"); break;
    case 24: asm volatile("s_waitcnt vmcnt(24)\n\ts_barrier" ::: "memory"); break;
    case 32: asm volatile("s_waitcnt vmcnt(32)\n\ts_barrier" ::: "memory"); break;
    default: asm volatile("s_waitcnt vmcnt(48)\n\ts_barrier" ::: "memory"); break;
  }
}
// Staging loads per thread per K block: eight B^T rows plus eight P_sub rows
// per 64 columns of the tile.
constexpr int loads_per_stage(int nt) { return 8 * (1 + (nt + 63) / 64); }
static_assert(loads_per_stage(64) == 16 && loads_per_stage(128) == 24, "ring wait counts");

// One dword per lane from global memory into the LDS row at lds_row (lane i
// writes lds_row[i]); lds_row must be the same for the whole wave.  Issued
// as inline asm so that the compiler's own wait insertion, which cannot tell
// the ring slots apart, adds no wait for these loads before LDS reads; the
// ring waits above cover them.  The low 32 bits of a generic pointer into
// LDS are the LDS byte offset that M0 takes.
__device__ __forceinline__ void load_lds(const float* src, float* lds_row) {
  const unsigned off = __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)lds_row);
  asm volatile("global_load_lds_dword %0, off" ::"v"(src), "{m0}"(off) : "memory");
}

// INSTAGE = 1: the fold and Z evaluate B and grad B themselves
// (fold_in_stage / z_in_stage); 3: the fold does, and the block's Z
// coefficients go to zbuf for a separate Z launch; 0: both read the planes
// the collocation stage stored.
template <int NT, bool GGA, bool HIOCC, int INSTAGE>
__global__ __launch_bounds__(256, density_waves(NT, HIOCC)) void k_density_mx(
    QueueArgs qa, const DevTask* __restrict__ tasks, const int32_t* __restrict__ bfmap,
    const float* __restrict__ Pfall, float* __restrict__ scratch,
    const double* __restrict__ weights, int func, double fac, int want_z,
    double* __restrict__ red, unsigned* __restrict__ tmax, ShellArgs sa,
    double* __restrict__ zbuf) {
  constexpr int BI = NT / 32;
  constexpr int kStages = ring_stages<NT, HIOCC>();
  constexpr int NPC = (NT + 63) / 64;  // 64-column chunks of a P_sub stage row
  constexpr int kLoadsPerStage = loads_per_stage(NT);
  // Operand ring: P_sub stages [slot][k][function] (A operand) then B^T
  // stages [slot][k][point] (B operand); after a tile's last K block the
  // X^T tile of the LDS fold (NT <= 64) is written over it.
  constexpr int kPtFloats = kStages * 32 * (NT + 4);
  constexpr int kStageFloats = kPtFloats + kStages * 32 * 68;
  // In-stage folds take the whole X^T tile at once (128-wide tiles: the
  // buffer grows to 66 KiB, still two workgroups per CU), so no wave holds
  // its accumulators while another row block is folded.
  constexpr bool kWholeXt = INSTAGE == 1 || INSTAGE == 3;
  constexpr int kRingFloats =
      kWholeXt && NT * kXLd * 2 > kStageFloats ? NT * kXLd * 2 : kStageFloats;
  static_assert(NT > 64 || NT * kXLd * 2 <= kRingFloats, "X^T tile must fit the ring");
  __shared__ __attribute__((aligned(16))) float ring[kRingFloats];
  float(*Pt)[32][NT + 4] = reinterpret_cast<float(*)[32][NT + 4]>(ring);
  float(*Bt)[32][68] = reinterpret_cast<float(*)[32][68]>(ring + kPtFloats);
  __shared__ double dens[4][64][4];
  __shared__ double zc[4][64];
  __shared__ unsigned smax[kGranLds][2];
  const int tid = threadIdx.x, lane = tid & 63, wave = uni(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int lc = lane, lg = wave;
  auto frow = [&](int r) { return 4 * (lane >> 4) + r; };

  XCQ_FOR_EACH_ITEM(qa, it) {
    const int64_t t0 = qa.tstamp ? wall_clock64() : 0;
    const WorkItem wi = qa.items[it];
    const DevTask t = tasks[wi.task];
    const int m0 = wi.a * kPtBlock;
    const int ld = t.npts_pad;
    const int nbe = t.nbe;
    const int plane = nbe * ld;
    const int64_t ldp = t.ldp;
    const float* Pf = Pfall + t.p_off;
    float* Bm = scratch + t.b_off_mx;
    const int32_t* bf = bfmap + t.bf_off;
    for (int i = tid; i < 4 * 64 * 4; i += 256) (&dens[0][0][0])[i] = 0.;
    if (tid < 2 * kGranLds) (&smax[0][0])[tid] = 0u;
    // Stores of the previous item may still count against vmcnt; the ring
    // waits below assume only this item's staging loads are outstanding.
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    const float* bcol = Bm + m0 + lc;
    const int nks = (nbe + 31) / 32;
    // In-stage basis evaluation: this lane's point (padded points take the
    // task's last point: finite values, which meet zero X and zero Z
    // coefficients), the shell cursor and this wave's density partials.
    const int pl = min(m0 + lane, t.npts - 1);
    double rx = 0., ry = 0., rz = 0.;
    if constexpr (INSTAGE != 0) {
      rx = sa.px[t.pt_off + pl];
      ry = sa.py[t.pt_off + pl];
      rz = sa.pz[t.pt_off + pl];
    }
    int s_cur = 0;
    double part[4] = {0., 0., 0., 0.};

    for (int n0 = 0; n0 < nbe; n0 += NT) {
      // P is symmetric: read P(k, n) as row bf[k], column bf[n] so that
      // neighbouring lanes (neighbouring n) touch neighbouring columns.
      const float* Pc[NPC];
      bool colok[NPC];
#pragma unroll
      for (int h = 0; h < NPC; ++h) {
        const int ncol = n0 + 64 * h + lc;
        colok[h] = lc < NT && ncol < nbe;
        Pc[h] = Pf + bf[min(ncol, nbe - 1)];
      }
      auto issue = [&](int s) {
        const int slot = s % kStages, k0 = 32 * s + 8 * lg;
        // Eight consecutive map entries (uniform: one scalar load); entries
        // past nbe lie in the map's zero padding and are not used.
        int rowp[8];
#pragma unroll
        for (int e = 0; e < 8; ++e) rowp[e] = bf[k0 + e];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int k = k0 + e;
          load_lds(bcol + min(k, nbe - 1) * ld, &Bt[slot][8 * lg + e][0]);
#pragma unroll
          for (int h = 0; h < NPC; ++h)
            if (lc < NT) {
              const float* src =
                  (k < nbe && colok[h]) ? Pc[h] + (int64_t)rowp[e] * ldp : g_zero_row + lc;
              load_lds(src, &Pt[slot][8 * lg + e][64 * h]);
            }
        }
      };
      double acc[BI][2][4];
#pragma unroll
      for (int a = 0; a < BI; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[a][b][r] = 0.;
      for (int s = 0; s < kStages - 1 && s < nks; ++s) issue(s);
      for (int s = 0; s < nks; ++s) {
        const int slot = s % kStages;
        // Blocks s+1 .. min(s + kStages - 2, nks - 1) may stay in flight.
        wait_loads_barrier(kLoadsPerStage * min(kStages - 2, nks - 1 - s));
        if (s + kStages - 1 < nks) issue(s + kStages - 1);
        v4f c[BI][2];
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj) c[bi][bj] = v4f{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int ks = 0; ks < 8; ++ks) {
          const int kr = ks * 4 + (lane >> 4);
          float a_[BI];
#pragma unroll
          for (int bi = 0; bi < BI; ++bi)
            a_[bi] = Pt[slot][kr][wm * 16 * BI + bi * 16 + (lane & 15)];
          const float b0 = Bt[slot][kr][wn * 32 + (lane & 15)];
          const float b1 = Bt[slot][kr][wn * 32 + 16 + (lane & 15)];
#pragma unroll
          for (int bi = 0; bi < BI; ++bi) {
            c[bi][0] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_[bi], b0, c[bi][0], 0, 0, 0);
            c[bi][1] = __builtin_amdgcn_mfma_f32_16x16x4f32(a_[bi], b1, c[bi][1], 0, 0, 0);
          }
        }
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[bi][bj][r] += (double)c[bi][bj][r];
      }
#pragma unroll
      for (int a = 0; a < BI; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
          for (int r = 0; r < 4; ++r) acc[a][b][r] *= fac;
      if constexpr (INSTAGE == 1 || INSTAGE == 3) {
        // The whole X^T tile goes to LDS (row = tile row), then is folded
        // in slices of up to 64 rows.
        double* xt = reinterpret_cast<double*>(ring);
        __syncthreads();  // every wave is done with the ring's last K block
#pragma unroll
        for (int bi = 0; bi < BI; ++bi)
#pragma unroll
          for (int bj = 0; bj < 2; ++bj)
#pragma unroll
            for (int r = 0; r < 4; ++r)
              xt[(wm * 16 * BI + bi * 16 + 4 * (lane >> 4) + r) * kXLd + wn * 32 + bj * 16 +
                 (lane & 15)] = acc[bi][bj][r];
        __syncthreads();
#pragma unroll
        for (int h = 0; h < (NT + 63) / 64; ++h) {
          const int nlo = n0 + 64 * h;
          if (nlo >= nbe) break;
          s_cur = fold_in_stage<GGA, true>(t, sa, s_cur, nlo, min(nlo + min(NT, 64), nbe), wave,
                                     lane, xt + 64 * h * kXLd, rx, ry, rz, part);
        }
      } else if constexpr (NT <= 64) {
        __syncthreads();  // every wave is done with the ring's last K block
        fold_tile_lds<NT, BI, GGA>(acc, reinterpret_cast<double*>(ring), m0, n0, wm, wn, lane,
                                   tid, nbe, ld, plane, Bm, dens);
      } else {
        fold_tile<float, BI, GGA>(acc, frow, m0, n0, wm, wn, lane, nbe, ld, plane, Bm, dens);
      }
      // The next tile's first loads overwrite the ring; every wave has
      // finished reading it (last K blocks, or the X^T tile) at this barrier.
      __syncthreads();
    }
    if constexpr (INSTAGE == 1 || INSTAGE == 3) {
#pragma unroll
      for (int c = 0; c < 4; ++c) dens[wave][lane][c] = part[c];
      __syncthreads();
    }
    finish_points<4>(t, wi.a, tid, dens, zc, weights, func, GGA, red);
    if (want_z) {
      __syncthreads();
      if constexpr (INSTAGE == 3)
        zbuf[(t.red_off + wi.a) * 256 + tid] = (&zc[0][0])[tid];
      else if constexpr (INSTAGE == 1)
        z_in_stage<GGA>(t, sa, m0, wave, lane, tid, rx, ry, rz, zc, Bm, smax,
                        tmax ? tmax + 2 * t.tmax_off : nullptr);
      else
        write_z<float, GGA, HIOCC ? 2 : 4>(t, m0, tid, zc, Bm, smax,
                                          tmax ? tmax + 2 * t.tmax_off : nullptr);
    }
    stamp_item(qa, it, t0);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// UKS tail (GauXC's generic host driver with spin_dim 2,
// reference_replicated_xc_host_integrator_exc_vxc.hpp:387-547, and the
// device UKS kernels it mirrors, cuda/kernels/zmat_vxc.cu:62-96,197-274).
// Inputs per point block: the raw sums of the P_s and P_z density passes
// (sum B X, sum dB X; X = B P with factor 1).  One thread per point:
//   rho_+- = (rho_s +- rho_z) / 2, grad n = 2 sum dB X_s, grad m = 2 sum dB X_z,
//   sigma_{aa, ab, bb} = (|gn|^2 + |gm|^2)/4 +- gn.gm/2, (|gn|^2 - |gm|^2)/4
//   (eval_uvvar_gga_uks, reference_local_host_work_driver.cxx:270-330);
// the polarised functional; weight folding; the block's (w n eps, w n) sums;
// and the Z coefficients of eval_zmat_{lda,gga}_vxc_uks (:607-632, 715-770):
//   Z_s = (v_a + v_b)/4 B + c_s . grad B,  Z_z = (v_a - v_b)/4 B + c_z . grad B,
//   c_s = g1 gn + g2 gm, c_z = g3 gm + g2 gn, with g1 = (v_aa + v_ab + v_bb)/2,
//   g2 = (v_aa - v_bb)/2, g3 = (v_aa - v_ab + v_bb)/2 (weights folded in).
__global__ __launch_bounds__(64) void k_uks_points(const WorkItem* __restrict__ items,
                                                   const DevTask* __restrict__ tasks,
                                                   const double* __restrict__ weights, int func,
                                                   int gga, const double* __restrict__ ds,
                                                   const double* __restrict__ dz,
                                                   double* __restrict__ red,
                                                   double* __restrict__ zs,
                                                   double* __restrict__ zz) {
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
    n = w * (ra + rb);
    if (func != 0) {
      const double saa = gga ? 0.25 * (nn + mm) + 0.5 * nm : 0.;
      const double sab = gga ? 0.25 * (nn - mm) : 0.;
      const double sbb = gga ? 0.25 * (nn + mm) - 0.5 * nm : 0.;
      const XCOutPol o = dev_eval_functional_pol(func, ra, rb, saa, sab, sbb);
      e = w * o.e;
      const double fp = 0.5 * w * o.vra, fm = 0.5 * w * o.vrb;
      as = 0.5 * (fp + fm);
      az = 0.5 * (fp - fm);
      if (gga) {
        const double vpp = w * o.vsaa, vpm = w * o.vsab, vmm = w * o.vsbb;
        const double g1 = 0.5 * (vpp + vpm + vmm), g2 = 0.5 * (vpp - vmm),
                     g3 = 0.5 * (vpp - vpm + vmm);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          cs[d] = g1 * gn[d] + g2 * gm[d];
          cz[d] = g3 * gm[d] + g2 * gn[d];
        }
      }
    }
  }
  zs[b * 256 + tid] = as;
  zz[b * 256 + tid] = az;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    zs[b * 256 + (1 + d) * 64 + tid] = cs[d];
    zz[b * 256 + (1 + d) * 64 + tid] = cz[d];
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * b] = e;
    red[2 * b + 1] = n;
  }
}

// ---------------------------------------------------------------------------
// FXC contraction tail, RKS (reference_replicated_xc_host_integrator_
// fxc_contraction.hpp:393-560 with spin_dim 1).  Inputs per point block: the
// raw sums of the P and trial-density passes with factor 2 (rho = sum B X,
// grad rho = 2 sum dB X; likewise trho, grad trho).  One thread per point:
//   tgamma = grad rho . grad trho,
//   A = v2rho2 trho + 2 v2rhogamma tgamma,
//   B = 2 (v2rhogamma trho + 2 v2gamma2 tgamma) grad rho + 2 vgamma grad trho
// (eval_tmat_{lda,gga}_vxc_rks, reference_local_host_work_driver.cxx:
// 1114-1143), weights folded in, and the Z coefficients of
// eval_zmat_gga_vxc_rks_ts (:1382-1408): a = A / 2, c = B.  The block's
// (w rho eps, w rho) sums of the ground-state density go to red.
__global__ __launch_bounds__(64) void k_fxc_points(const WorkItem* __restrict__ items,
                                                   const DevTask* __restrict__ tasks,
                                                   const double* __restrict__ weights, int func,
                                                   int gga, const double* __restrict__ dp,
                                                   const double* __restrict__ dt,
                                                   double* __restrict__ red,
                                                   double* __restrict__ zc) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x;
  const int64_t b = t.red_off + wi.a;
  const int p = wi.a * kPtBlock + tid;
  double sv[4], tv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    sv[c] = dp[b * 256 + c * 64 + tid];
    tv[c] = dt[b * 256 + c * 64 + tid];
  }
  double g[3], tg[3], sig = 0., tgam = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    g[d] = 2.0 * sv[1 + d];
    tg[d] = 2.0 * tv[1 + d];
    sig += g[d] * g[d];
    tgam += g[d] * tg[d];
  }
  double e = 0., n = 0., a = 0., cv[3] = {0., 0., 0.};
  if (p < t.npts) {
    const double w = weights[t.pt_off + p];
    n = w * sv[0];
    const XCOut o = dev_eval_functional(func, sv[0], sig);
    e = w * o.e;
    double vs, A, Bc;
    dev_fxc_point(func, sv[0], sig, tv[0], 2.0 * tgam, vs, A, Bc);
    a = 0.5 * w * A;
    if (gga) {
#pragma unroll
      for (int d = 0; d < 3; ++d) cv[d] = w * (2.0 * Bc * g[d] + 2.0 * vs * tg[d]);
    }
  }
  zc[b * 256 + tid] = a;
#pragma unroll
  for (int d = 0; d < 3; ++d) zc[b * 256 + (1 + d) * 64 + tid] = cv[d];
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * b] = e;
    red[2 * b + 1] = n;
  }
}

// UKS FXC contraction tail (reference_replicated_xc_host_integrator_
// fxc_contraction.hpp:94-600 with spin_dim 2; the trial kernels
// eval_tmat_{lda,gga}_vxc_uks and eval_zmat_*_uks_ts,
// reference_local_host_work_driver.cxx:1118-1124, 1149-1215, 1354-1367,
// 1441-1481).  FXC_s and FXC_z are the derivatives of VXC_s and VXC_z along
// the trial density, so this kernel writes the derivatives of k_uks_points'
// Z coefficients along (tP_s, tP_z): inputs are the raw sums of the four
// density passes (P_s, P_z, tP_s, tP_z; factor 1), the second derivatives
// come from dev_fxc_point_pol.  Padded points give zero coefficients.
__global__ __launch_bounds__(64) void k_fxc_points_uks(
    const WorkItem* __restrict__ items, const DevTask* __restrict__ tasks,
    const double* __restrict__ weights, int func, int gga, const double* __restrict__ ds,
    const double* __restrict__ dz, const double* __restrict__ dts,
    const double* __restrict__ dtz, double* __restrict__ red, double* __restrict__ zs,
    double* __restrict__ zz) {
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x;
  const int64_t b = t.red_off + wi.a;
  const int p = wi.a * kPtBlock + tid;
  double sv[4], zv[4], tsv[4], tzv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    sv[c] = ds[b * 256 + c * 64 + tid];
    zv[c] = dz[b * 256 + c * 64 + tid];
    tsv[c] = dts[b * 256 + c * 64 + tid];
    tzv[c] = dtz[b * 256 + c * 64 + tid];
  }
  double gn[3], gm[3], tgn[3], tgm[3];
  double nn = 0., mm = 0., nm = 0., dnn = 0., dmm = 0., dnm = 0.;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    gn[d] = 2.0 * sv[1 + d];
    gm[d] = 2.0 * zv[1 + d];
    tgn[d] = 2.0 * tsv[1 + d];
    tgm[d] = 2.0 * tzv[1 + d];
    nn += gn[d] * gn[d];
    mm += gm[d] * gm[d];
    nm += gn[d] * gm[d];
    dnn += 2.0 * gn[d] * tgn[d];
    dmm += 2.0 * gm[d] * tgm[d];
    dnm += tgn[d] * gm[d] + gn[d] * tgm[d];
  }
  double e = 0., n = 0., das = 0., daz = 0., dcs[3] = {0., 0., 0.}, dcz[3] = {0., 0., 0.};
  if (p < t.npts) {
    const double w = weights[t.pt_off + p];
    double x[5] = {fmax(0.5 * (sv[0] + zv[0]), 0.0), fmax(0.5 * (sv[0] - zv[0]), 0.0), 0., 0.,
                   0.};
    double dx[5] = {0.5 * (tsv[0] + tzv[0]), 0.5 * (tsv[0] - tzv[0]), 0., 0., 0.};
    if (gga) {
      const double stol2 = 2.1544346900318838e-43 * 2.1544346900318838e-43;
      x[2] = fmax(0.25 * (nn + mm) + 0.5 * nm, stol2);
      x[4] = fmax(0.25 * (nn + mm) - 0.5 * nm, stol2);
      const double hs = 0.5 * (x[2] + x[4]);
      x[3] = fmax(fmin(0.25 * (nn - mm), hs), -hs);
      dx[2] = 0.25 * (dnn + dmm) + 0.5 * dnm;
      dx[3] = 0.25 * (dnn - dmm);
      dx[4] = 0.25 * (dnn + dmm) - 0.5 * dnm;
    }
    n = w * sv[0];
    const XCOutPol o = dev_eval_functional_pol(func, x[0], x[1], x[2], x[3], x[4]);
    e = w * o.e;
    double out[10];
    dev_fxc_point_pol(func, x, dx, out);
    das = 0.25 * w * (out[5] + out[6]);
    daz = 0.25 * w * (out[5] - out[6]);
    if (gga) {
      const double g1 = 0.5 * w * (out[2] + out[3] + out[4]), g2 = 0.5 * w * (out[2] - out[4]),
                   g3 = 0.5 * w * (out[2] - out[3] + out[4]);
      const double dg1 = 0.5 * w * (out[7] + out[8] + out[9]), dg2 = 0.5 * w * (out[7] - out[9]),
                   dg3 = 0.5 * w * (out[7] - out[8] + out[9]);
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        dcs[d] = dg1 * gn[d] + g1 * tgn[d] + dg2 * gm[d] + g2 * tgm[d];
        dcz[d] = dg3 * gm[d] + g3 * tgm[d] + dg2 * gn[d] + g2 * tgn[d];
      }
    }
  }
  zs[b * 256 + tid] = das;
  zz[b * 256 + tid] = daz;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    zs[b * 256 + (1 + d) * 64 + tid] = dcs[d];
    zz[b * 256 + (1 + d) * 64 + tid] = dcz[d];
  }
  for (int off = 32; off > 0; off >>= 1) {
    e += __shfl_xor(e, off);
    n += __shfl_xor(n, off);
  }
  if (tid == 0) {
    red[2 * b] = e;
    red[2 * b + 1] = n;
  }
}

// Z = a B + c . grad B of one point block from stored coefficients
// ([block][4][64]) and the stored FP64 planes (no screening bounds).
template <bool GGA>
__global__ __launch_bounds__(256) void k_z_from_coeffs(const WorkItem* __restrict__ items,
                                                       const DevTask* __restrict__ tasks,
                                                       double* __restrict__ scratch,
                                                       const double* __restrict__ coef) {
  __shared__ double zc[4][64];
  __shared__ unsigned smax[kGranLds][2];
  const WorkItem wi = items[blockIdx.x];
  const DevTask t = tasks[wi.task];
  const int tid = threadIdx.x;
  (&zc[0][0])[tid] = coef[(t.red_off + wi.a) * 256 + tid];
  __syncthreads();
  write_z<double, GGA>(t, wi.a * kPtBlock, tid, zc, scratch + t.b_off, smax, nullptr);
}

int launch_uks_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* ds,
                      const double* dz, double* zs, double* zz) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_uks_points, dim3((unsigned)r.n), dim3(64), 0, v.stream, v.work + r.off,
                     v.tasks, v.w, func, gga, ds, dz, v.red, zs, zz);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_fxc_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* dp,
                      const double* dt, double* zc) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_fxc_points, dim3((unsigned)r.n), dim3(64), 0, v.stream, v.work + r.off,
                     v.tasks, v.w, func, gga, dp, dt, v.red, zc);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_fxc_points_uks(const CtxView& v, const WorkRange& r, int func, int gga,
                          const double* ds, const double* dz, const double* dts,
                          const double* dtz, double* zs, double* zz) {
  if (r.n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_fxc_points_uks, dim3((unsigned)r.n), dim3(64), 0, v.stream,
                     v.work + r.off, v.tasks, v.w, func, gga, ds, dz, dts, dtz, v.red, zs, zz);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_z_from_coeffs(const CtxView& v, const WorkRange& r, int gga, const double* coef) {
  if (r.n == 0) return XCQ_SUCCESS;
  double* scr = static_cast<double*>(v.scratch);
  if (gga)
    hipLaunchKernelGGL(k_z_from_coeffs<true>, dim3((unsigned)r.n), dim3(256), 0, v.stream,
                       v.work + r.off, v.tasks, scr, coef);
  else
    hipLaunchKernelGGL(k_z_from_coeffs<false>, dim3((unsigned)r.n), dim3(256), 0, v.stream,
                       v.work + r.off, v.tasks, scr, coef);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

int launch_density(const CtxView& v, const WorkRange& r, int nt, bool hiocc, int precision,
                   int func, int gga, double fac, int want_z, const double* P_override,
                   double* dens_out, int x_plane) {
  if (r.n == 0) return XCQ_SUCCESS;
  XCQ_REQUIRE(x_plane < 0 || (precision == XCQ_PREC_FP64 && !hiocc && x_plane < 5),
              XCQ_ERROR_INVALID_ARGUMENT,
              "the X plane is stored by the FP64 kernels in their two-workgroup configuration");
  const QueueArgs qa = queue_args(v, r);
  const dim3 grid(grid_of(r)), block(256);
  // The bound table is filled only when VXC is computed and contributions
  // may be dropped.
  unsigned* tm = (want_z && v.vxc_drop > 0.0) ? v.tmax : nullptr;
  const ShellArgs sa{v.shells, v.shc, v.shell_list, v.lsh, v.px, v.py, v.pz};
  if (precision == XCQ_PREC_FP64) {
    double* scr = static_cast<double*>(v.scratch);
    const double* P = P_override ? P_override : v.P;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, block, 0, v.stream, qa, v.tasks, v.bfmap, P, scr, v.w,
                         func, fac, want_z, v.red, tm, sa, v.zc, dens_out, x_plane);
    };
    // Any in-stage mode runs the FP64 kernel's one (mode 3).
    if (v.in_stage != 0) {
      if (nt == 32)
        gga ? go(k_density_f64<32, true, 3>) : go(k_density_f64<32, false, 3>);
      else if (hiocc)
        gga ? go(k_density_f64<64, true, 3, true>) : go(k_density_f64<64, false, 3, true>);
      else
        gga ? go(k_density_f64<64, true, 3>) : go(k_density_f64<64, false, 3>);
    } else {
      if (nt == 32)
        gga ? go(k_density_f64<32, true, 0>) : go(k_density_f64<32, false, 0>);
      else
        gga ? go(k_density_f64<64, true, 0>) : go(k_density_f64<64, false, 0>);
    }
  } else {
    float* scr = static_cast<float*>(v.scratch);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, grid, block, 0, v.stream, qa, v.tasks, v.bfmap, v.Pf32, scr, v.w,
                         func, fac, want_z, v.red, tm, sa, v.zc);
    };
    // Template arguments: tile width, GGA, high occupancy, in-stage mode.
    auto pick = [&](auto g, auto ins) {
      constexpr bool G = decltype(g)::value;
      constexpr int I = decltype(ins)::value;
      if (nt == 32) {
        go(k_density_mx<32, G, false, I>);
      } else if (nt == 128) {
        go(k_density_mx<128, G, false, I>);
      } else if constexpr (I != 1) {
        if (hiocc) go(k_density_mx<64, G, true, I>);
        else go(k_density_mx<64, G, false, I>);
      } else {
        // The in-stage fold with the in-kernel Z pass needs more than the
        // 168 registers of the three-workgroups-per-CU configuration; it
        // runs at two.
        go(k_density_mx<64, G, false, I>);
      }
    };
    using T_ = std::true_type;
    using F_ = std::false_type;
    using M0 = std::integral_constant<int, 0>;
    using M1 = std::integral_constant<int, 1>;
    using M3 = std::integral_constant<int, 3>;
    if (v.in_stage == 1)
      gga ? pick(T_{}, M1{}) : pick(F_{}, M1{});
    else if (v.in_stage == 3)
      gga ? pick(T_{}, M3{}) : pick(F_{}, M3{});
    else
      gga ? pick(T_{}, M0{}) : pick(F_{}, M0{});
  }
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
