// FP64 contractions of the XC path on gfx950 matrix cores
// (v_mfma_f64_16x16x4_f64).
//
//  (X = fac B P_sub, the density reduction and Z live in density.hip.)
//  k_syr2k    : V = B^T Z + Z^T B for lower-triangle 64x64 tiles with the
//               scatter-add into VXC fused into the epilogue (GauXC: DSYR2K
//               per task + sym_inc_by_submat, scheme1_base.cxx:1712-1766).
//
// Workgroup: 256 threads = 4 waves in a 2x2 arrangement.  K advances 16 per
// LDS stage.
// MFMA operand layout (16x16x4 f64, measured on gfx950): lane l holds
// A[l%16][l/16] and B[l/16][l%16]; accumulator register r of lane l is
// D[4*r + l/16][l%16] (rows interleave across the four 16-lane groups,
// unlike the FP32-output MFMAs whose register r holds row 4*(l/16) + r).
#include "internal.hpp"

namespace xcq {

using v4d = double __attribute__((ext_vector_type(4)));

__device__ __forceinline__ v4d mfma_f64(double a, double b, v4d c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// ---------------------------------------------------------------------------
// Templated on the function tile NT (32 or 64); each wave owns an
// (NT/2) x (NT/2) sub-tile = BI x BI MFMA blocks.  Loader: thread
// (p = tid & 15, f = tid >> 4) moves point k0 + p of rows f, f+16, ... of
// each panel; 16 lanes cover 16 consecutive points (128 B) of one row.  Rows
// past nbe are read at row nbe-1 and zeroed by a select.  The next K step is
// loaded into registers while the current one runs on the matrix cores.
template <int NT>
__global__ __launch_bounds__(256) void k_syr2k(QueueArgs qa, const DevTask* __restrict__ tasks,
                                               const int32_t* __restrict__ bfmap,
                                               const double* __restrict__ scratch,
                                               double* __restrict__ vxc_all, double drop,
    const unsigned* __restrict__ tmax) {
  constexpr int BI = NT / 32, NQ = NT / 16, LP = NT + 4;
  __shared__ double Bi[16][LP];
  __shared__ double Zi[16][LP];
  __shared__ double Bj[16][LP];
  __shared__ double Zj[16][LP];
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wave & 1, wn = wave >> 1;
  const int l_p = tid & 15, l_f = tid >> 4;

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
    const double* Bm = scratch + t.b_off;
    const double* Zm = Bm + 4 * plane;
    const int64_t ldv = t.ldp;
    double* vxc = vxc_all + t.p_off;

    int oi[NQ], oj[NQ];
    bool ki[NQ], kj[NQ];
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      const int fi = i0 + l_f + 16 * q, fj = j0 + l_f + 16 * q;
      ki[q] = fi < nbe;
      kj[q] = fj < nbe;
      oi[q] = min(fi, nbe - 1) * ld + l_p;
      oj[q] = min(fj, nbe - 1) * ld + l_p;
    }
    // Raw prefetch registers; the row masks are applied when staged, so the
    // loads stay in flight across the current step's matrix instructions.
    double rbi[NQ], rzi[NQ], rbj[NQ], rzj[NQ];
    auto load = [&](int k0) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        rbi[q] = Bm[oi[q] + k0];
        rzi[q] = Zm[oi[q] + k0];
      }
      if (!diag) {
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
          rbj[q] = Bm[oj[q] + k0];
          rzj[q] = Zm[oj[q] + k0];
        }
      }
    };

    v4d acc[BI][BI];
#pragma unroll
    for (int i = 0; i < BI; ++i)
#pragma unroll
      for (int j = 0; j < BI; ++j) acc[i][j] = v4d{0., 0., 0., 0.};

    load(0);
    for (int k0 = 0; k0 < ld; k0 += 16) {
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        Bi[l_p][l_f + 16 * q] = ki[q] ? rbi[q] : 0.;
        Zi[l_p][l_f + 16 * q] = ki[q] ? rzi[q] : 0.;
        if (!diag) {
          Bj[l_p][l_f + 16 * q] = kj[q] ? rbj[q] : 0.;
          Zj[l_p][l_f + 16 * q] = kj[q] ? rzj[q] : 0.;
        }
      }
      __syncthreads();
      if (k0 + 16 < ld) load(k0 + 16);
      const double(*BJ)[LP] = diag ? Bi : Bj;
      const double(*ZJ)[LP] = diag ? Zi : Zj;
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double bi_[BI], zi_[BI], bj_[BI], zj_[BI];
#pragma unroll
        for (int q = 0; q < BI; ++q) {
          const int ci = wm * 16 * BI + q * 16 + (lane & 15);
          const int cj = wn * 16 * BI + q * 16 + (lane & 15);
          bi_[q] = Bi[kr][ci];
          zi_[q] = Zi[kr][ci];
          bj_[q] = BJ[kr][cj];
          zj_[q] = ZJ[kr][cj];
        }
#pragma unroll
        for (int x = 0; x < BI; ++x)
#pragma unroll
          for (int y = 0; y < BI; ++y) {
            acc[x][y] = mfma_f64(bi_[x], zj_[y], acc[x][y]);
            acc[x][y] = mfma_f64(zi_[x], bj_[y], acc[x][y]);
          }
      }
      __syncthreads();
    }

    // Scatter-add into VXC.  Element (i, j) of the tile (i >= j locally) is
    // added at (bf[j], bf[i]) of the column-major VXC: the 16 lanes of a
    // lane group hold 16 consecutive j, so each atomic instruction touches a
    // few 128-byte column segments instead of 16 columns.  The
    // symmetrisation that follows sums the two triangles, so which triangle
    // an element lands in does not matter.
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
            const int i = i0 + wm * 16 * BI + bi * 16 + 4 * r + (lane >> 4);
            // Negligible contributions are not added (NaN still is).
            if (i < nbe && (!diag || i >= j) && !(fabs(acc[bi][bj][r]) < drop))
              atomicAdd(vxc + (int64_t)bf[i] * ldv + rj, acc[bi][bj][r]);
          }
      }
    }
    stamp_item(qa, it, t0);
  }
}

int launch_syr2k_mixed(const CtxView& v, const WorkRange& r, int nt);

int launch_syr2k(const CtxView& v, const WorkRange& r, int nt, int precision) {
  if (r.n == 0) return XCQ_SUCCESS;
  if (precision != XCQ_PREC_FP64) return launch_syr2k_mixed(v, r, nt);
  const QueueArgs qa = queue_args(v, r);
  const double* scr = static_cast<const double*>(v.scratch);
  if (nt == 32)
    hipLaunchKernelGGL(k_syr2k<32>, dim3(grid_of(r)), dim3(256), 0, v.stream, qa, v.tasks,
                       v.bfmap, scr, v.vxc, v.vxc_drop, v.vxc_drop > 0.0 ? v.tmax : nullptr);
  else
    hipLaunchKernelGGL(k_syr2k<64>, dim3(grid_of(r)), dim3(256), 0, v.stream, qa, v.tasks,
                       v.bfmap, scr, v.vxc, v.vxc_drop, v.vxc_drop > 0.0 ? v.tmax : nullptr);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
