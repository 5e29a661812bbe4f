// Point-wise kernels of the XC path on gfx950: SSF partition weights, the
// functional with weight folding and EXC/NEL block sums, the deterministic
// final reduction, and VXC symmetrisation.
#include "internal.hpp"

#include <algorithm>
#include "functional.hpp"


namespace xcq {

// ---------------------------------------------------------------------------
// SSF weights, one thread per point.
// GauXC semantics: host reference weights.cxx:117-237 (Frisch polynomial,
// a = 0.64, points closer to the parent than 0.5(1-a) R_nearest keep their
// weight).  The partition function of atom A is the product
// P_A = prod_{B != A} s(mu_AB), s(mu) = (1 - f(mu/a))/2 inside |mu| < a,
// 0 above, 1 below -- the same quantity the reference accumulates pairwise.
// Only atoms with d_A < (1+a)/(1-a) d_min can have P_A > 0 (their
// mu against the nearest atom is >= a otherwise), so the candidate set is
// filtered by that bound, slightly widened; exact zeros are still decided
// by the mu test itself.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ssf_s(double mu) {
  constexpr double a = 0.64;
  if (mu <= -a) return 1.0;
  if (mu >= a) return 0.0;
  const double s = mu / a, s2 = s * s, s3 = s * s2, s5 = s3 * s2, s7 = s5 * s2;
  const double f = (35. * (s - s3) + 21. * s5 - 5. * s7) / 16.;
  return 0.5 * (1.0 - f);
}

// With neighbour lists (nbr_idx / nbr_r: row A holds every atom in order of
// increasing R_AB, A itself first) each loop stops at a distance the
// triangle inequality fixes, instead of running over every atom:
//  * the nearest atom B* has d_B* <= d_P (P the parent), so R_PB* <= 2 d_P;
//  * a candidate A has d_A <= bound, so R_PA <= d_P + bound;
//  * a factor s(mu_AB) differs from 1 only if mu_AB > -a, i.e.
//    d_B < d_A + a R_AB, and d_B >= R_AB - d_A, so only R_AB < 2 d_A / (1 - a)
//    can contribute (the factors skipped are exactly 1).
// Each limit is widened by a relative 1e-10, far above rounding, so the
// result differs from the all-atom loops only in the order of the product's
// factors.  Without lists (nbr_idx null) the loops run over all atoms.
// The parent's product comes first: when it is exactly 0 the weight is
// w * 0 whatever the sum (the nearest atom's product is positive, so the
// sum is), and the point is done.  The nearest atom's product (usually the
// largest) comes next.  Every other candidate's product stops once it falls
// below 2^-70 times the running sum and is then left out, and a candidate
// whose factor against the nearest atom is already below that is not
// formed: factors lie in [0, 1], so each omitted term is below 2^-70 of the
// final sum and all of them together (at most 8,192) below 2^-57, under the
// rounding of the sum itself.
__global__ __launch_bounds__(256) void k_ssf(int32_t natoms, const double* __restrict__ xyz,
                                             const double* __restrict__ rab,
                                             const double* __restrict__ dnear, int64_t npts,
                                             const double* __restrict__ px,
                                             const double* __restrict__ py,
                                             const double* __restrict__ pz,
                                             int64_t stride,
                                             const int32_t* __restrict__ parent,
                                             double* __restrict__ w,
                                             int32_t* __restrict__ bad_parent,
                                             const int32_t* __restrict__ nbr_idx,
                                             const double* __restrict__ nbr_r) {
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= npts) return;
  constexpr double a = 0.64;
  const double x = px[ip * stride], y = py[ip * stride], z = pz[ip * stride];
  const int par = parent[ip];
  if (par < 0 || par >= natoms) {  // callers that cannot validate on the host
    if (bad_parent) *bad_parent = 1;
    return;
  }
  auto dist = [&](int A) {
    const double dx = x - xyz[3 * A], dy = y - xyz[3 * A + 1], dz = z - xyz[3 * A + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
  };
  const double dpar = dist(par);
  if (dpar < 0.5 * (1.0 - a) * dnear[par]) return;
  constexpr double kWiden = 1.0 + 1e-10;
  if (nbr_idx) {
    const int32_t* np = nbr_idx + (size_t)par * natoms;
    const double* nr = nbr_r + (size_t)par * natoms;
    double dmin = dpar;
    int inear = par;
    const double rnear = 2.0 * dpar * kWiden;
    for (int k = 1; k < natoms && nr[k] <= rnear; ++k) {
      const int B = np[k];
      const double d = dist(B);
      if (d < dmin) {
        dmin = d;
        inear = B;
      }
    }
    const double bound = (1.0 + a) / (1.0 - a) * dmin * kWiden + 1e-12;
    const double rcand = (dpar + bound) * kWiden;
    constexpr double kNegligible = 0x1p-70;
    // Product of candidate A, or 0 where it is 0 or below `floor_`.
    auto product = [&](int A, double dA, double floor_) {
      if (dA > bound) return 0.0;
      if (A != inear &&
          ssf_s((dA - dmin) / rab[(size_t)A + (size_t)inear * natoms]) <= floor_)
        return 0.0;
      const int32_t* ap = nbr_idx + (size_t)A * natoms;
      const double* ar = nbr_r + (size_t)A * natoms;
      const double rfac = 2.0 * dA / (1.0 - a) * kWiden;
      double PA = 1.0;
      for (int q = 1; q < natoms && ar[q] < rfac && PA > floor_; ++q)
        PA *= ssf_s((dA - dist(ap[q])) / ar[q]);
      return PA > floor_ ? PA : 0.0;
    };
    const double ppar = product(par, dpar, 0.0);
    if (ppar == 0.0) {
      w[ip] *= 0.0;
      return;
    }
    double sum = ppar;
    if (inear != par) sum += product(inear, dmin, kNegligible * sum);
    for (int k = 1; k < natoms && nr[k] <= rcand; ++k) {
      const int A = np[k];
      if (A == inear) continue;
      sum += product(A, dist(A), kNegligible * sum);
    }
    w[ip] *= ppar / sum;
    return;
  }
  double dmin = dpar;
  int inear = par;
  for (int A = 0; A < natoms; ++A) {
    const double d = dist(A);
    if (d < dmin) {
      dmin = d;
      inear = A;
    }
  }
  const double bound = (1.0 + a) / (1.0 - a) * dmin * (1.0 + 1e-10) + 1e-12;
  double sum = 0.0, ppar = 0.0;
  for (int A = 0; A < natoms; ++A) {
    const double dA = dist(A);
    if (dA > bound) continue;
    // The nearest atom's factor first: when it is exactly 0 the product
    // below is exactly 0 as well (every factor lies in [0, 1]).
    if (A != inear && ssf_s((dA - dmin) / rab[(size_t)A + (size_t)inear * natoms]) == 0.0)
      continue;
    double PA = 1.0;
    for (int B = 0; B < natoms && PA > 0.0; ++B) {
      if (B == A) continue;
      const double mu = (dA - dist(B)) / rab[(size_t)A + (size_t)B * natoms];
      PA *= ssf_s(mu);
    }
    sum += PA;
    if (A == par) ppar = PA;
  }
  w[ip] *= ppar / sum;
}

int launch_ssf(hipStream_t s, int32_t natoms, const double* xyz, const double* rab,
               const double* dnear, int64_t npts, const double* px, const double* py,
               const double* pz, int64_t stride, const int32_t* parent, double* w,
               int32_t* bad_parent, const int32_t* nbr_idx, const double* nbr_r) {
  if (npts == 0) return XCQ_SUCCESS;
  const int64_t nb = (npts + 255) / 256;
  hipLaunchKernelGGL(k_ssf, dim3((unsigned)nb), dim3(256), 0, s, natoms, xyz, rab, dnear, npts,
                     px, py, pz, stride, parent, w, bad_parent, nbr_idx, nbr_r);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// MolMeta on the device (GauXC src/molmeta.cxx:28-60): R_AB (column-major,
// leading dimension natoms) and each atom's nearest-neighbour distance, one
// thread per atom.  Products and sums are rounded separately (no fused
// multiply-add), as the host computes them.  Thread 0 also clears the
// bad-parent flag of the SSF launch that follows on the same stream.
__global__ __launch_bounds__(256) void k_molmeta(int32_t natoms, const double* __restrict__ xyz,
                                                 double* __restrict__ rab,
                                                 double* __restrict__ dnear,
                                                 int32_t* __restrict__ bad_parent) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i == 0 && bad_parent) *bad_parent = 0;
  if (i >= natoms) return;
  double dmin = INFINITY;
  for (int j = 0; j < natoms; ++j) {
    const double dx = xyz[3 * i] - xyz[3 * j];
    const double dy = xyz[3 * i + 1] - xyz[3 * j + 1];
    const double dz = xyz[3 * i + 2] - xyz[3 * j + 2];
    const double r = sqrt(__dadd_rn(__dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy)),
                                    __dmul_rn(dz, dz)));
    rab[(size_t)i + (size_t)j * natoms] = r;
    if (j != i) dmin = fmin(dmin, r);
  }
  dnear[i] = dmin;
}

int launch_molmeta(hipStream_t s, int32_t natoms, const double* xyz, double* rab, double* dnear,
                   int32_t* bad_parent) {
  hipLaunchKernelGGL(k_molmeta, dim3((unsigned)((natoms + 255) / 256)), dim3(256), 0, s, natoms,
                     xyz, rab, dnear, bad_parent);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// First-order correction of EXC for the FP32 rounding of P in the mixed
// path.  With rho = fac sum_ij P_ij phi_i phi_j, dEXC/dP_ij = fac VXC_ij
// (VXC symmetrised), so
//   EXC(P) = EXC(P32) + fac sum_ij (P - P32)_ij VXC_ij + O(|P - P32|^2),
// where P32 is the FP32 copy the density stage multiplies with.  Workgroup
// (b, m) sums elements b*256 + tid + k*kCorrBlocks*256 of molecule m's n x n
// matrices and writes its partial to part[m * kCorrBlocks + b];
// k_reduce_final adds the partials of each molecule in a fixed order.
__global__ __launch_bounds__(256) void k_p32_correction(const int32_t* __restrict__ mol_n,
                                                        const int64_t* __restrict__ mol_off,
                                                        const double* __restrict__ P,
                                                        const float* __restrict__ Pf,
                                                        const double* __restrict__ vxc,
                                                        double* __restrict__ part) {
  __shared__ double sh[256];
  const int m = blockIdx.y;
  const int64_t n2 = (int64_t)mol_n[m] * mol_n[m];
  const int64_t off = mol_off[m];
  double s = 0.;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n2; i += kCorrBlocks * 256)
    s += (P[off + i] - (double)Pf[off + i]) * vxc[off + i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[(int64_t)m * kCorrBlocks + blockIdx.x] = sh[0];
}

int launch_p32_correction(hipStream_t s, int32_t nmol, const int32_t* mol_n,
                          const int64_t* mol_off, const double* P, const float* Pf,
                          const double* vxc, double* part) {
  if (nmol == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_p32_correction, dim3(kCorrBlocks, (unsigned)nmol), dim3(256), 0, s, mol_n,
                     mol_off, P, Pf, vxc, part);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Per-molecule sums of the (e, n) block pairs in a fixed order:
// deterministic EXC and NEL.  First pass: workgroup (b, m) sums pairs
// blk_off[m] + b*256 + tid + k*kRedBlocks*256 into part[m][b]; second pass:
// one wave per molecule adds the kRedBlocks partials (and, when corr is
// given, cfac times the molecule's kCorrBlocks correction partials) in
// order.
__global__ __launch_bounds__(256) void k_reduce_partials(const int64_t* __restrict__ blk_off,
                                                         const double* __restrict__ red,
                                                         double* __restrict__ part) {
  __shared__ double se[256], sn[256];
  const int m = blockIdx.y;
  double e = 0., n = 0.;
  for (int64_t i = blk_off[m] + (int64_t)blockIdx.x * 256 + threadIdx.x; i < blk_off[m + 1];
       i += (int64_t)kRedBlocks * 256) {
    e += red[2 * i];
    n += red[2 * i + 1];
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = n;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      se[threadIdx.x] += se[threadIdx.x + st];
      sn[threadIdx.x] += sn[threadIdx.x + st];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    part[2 * ((int64_t)m * kRedBlocks + blockIdx.x)] = se[0];
    part[2 * ((int64_t)m * kRedBlocks + blockIdx.x) + 1] = sn[0];
  }
}

__global__ __launch_bounds__(64) void k_reduce_final(const double* __restrict__ part,
                                                     const double* __restrict__ corr, double cfac,
                                                     double* __restrict__ out) {
  const int m = blockIdx.x;
  if (threadIdx.x != 0) return;
  double e = 0., n = 0., c = 0.;
  for (int b = 0; b < kRedBlocks; ++b) {
    e += part[2 * ((int64_t)m * kRedBlocks + b)];
    n += part[2 * ((int64_t)m * kRedBlocks + b) + 1];
  }
  if (corr)
    for (int b = 0; b < kCorrBlocks; ++b) c += corr[(int64_t)m * kCorrBlocks + b];
  out[2 * m] = e + cfac * c;
  out[2 * m + 1] = n;
}

int launch_reduce_sums(hipStream_t s, int32_t nmol, const int64_t* blk_off, const double* red,
                       double* part, const double* corr, double cfac, double* out2) {
  if (nmol == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_reduce_partials, dim3(kRedBlocks, (unsigned)nmol), dim3(256), 0, s,
                     blk_off, red, part);
  hipLaunchKernelGGL(k_reduce_final, dim3((unsigned)nmol), dim3(64), 0, s, part, corr, cfac,
                     out2);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Zero fill (used instead of a memset node so the pipeline can be captured
// into a graph).
__global__ __launch_bounds__(256) void k_fill_zero(double* __restrict__ p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    p[i] = 0.;
}

int launch_fill(hipStream_t s, double* p, int64_t n) {
  if (n == 0) return XCQ_SUCCESS;
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Symmetrisation of the accumulated VXC: each off-diagonal pair (a, b) holds
// contributions in either triangle, and both entries become their sum
// (GauXC fills the upper triangle from the lower one,
// host_integrator_exc_vxc.hpp:577-599; here the scatter may land in either).
// Workgroup (bi, bj, m), bi >= bj, owns the 32x32 tile pair (bi, bj) /
// (bj, bi) of molecule m's nbf x nbf matrix at offset mol_off[m].
__global__ __launch_bounds__(256) void k_symmetrize(const int32_t* __restrict__ mol_n,
                                                    const int64_t* __restrict__ mol_off,
                                                    double* __restrict__ vxc) {
  __shared__ double lo[32][33];
  __shared__ double up[32][33];
  const int bi = blockIdx.x, bj = blockIdx.y;
  const int n = mol_n[blockIdx.z];
  if (bi < bj || bj * 32 >= n || bi * 32 >= n) return;
  const int64_t lda = n;
  double* a = vxc + mol_off[blockIdx.z];
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int k = ty; k < 32; k += 8) {
    const int r = bi * 32 + tx, c = bj * 32 + k;  // lower tile element (r, c)
    lo[k][tx] = (r < n && c < n) ? a[r + (int64_t)c * lda] : 0.;
    const int r2 = bj * 32 + tx, c2 = bi * 32 + k;  // upper tile element (r2, c2)
    up[k][tx] = (r2 < n && c2 < n) ? a[r2 + (int64_t)c2 * lda] : 0.;
  }
  __syncthreads();
  for (int k = ty; k < 32; k += 8) {
    // lower (r = bi*32+tx, c = bj*32+k) pairs with upper (c, r) = up[tx][k]
    const int r = bi * 32 + tx, c = bj * 32 + k;
    if (r < n && c < n && r > c) a[r + (int64_t)c * lda] = lo[k][tx] + up[tx][k];
    const int r2 = bj * 32 + tx, c2 = bi * 32 + k;
    if (r2 < n && c2 < n && r2 < c2) a[r2 + (int64_t)c2 * lda] = up[k][tx] + lo[tx][k];
  }
}

int launch_symmetrize(hipStream_t s, int32_t nmol, int32_t max_n, const int32_t* mol_n,
                      const int64_t* mol_off, double* vxc) {
  const int nt = (max_n + 31) / 32;
  if (nt == 0 || nmol == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_symmetrize, dim3(nt, nt, nmol), dim3(256), 0, s, mol_n, mol_off, vxc);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

// Functional on flat arrays (kernel-level parity entry point).
__global__ __launch_bounds__(256) void k_functional_points(int func, int64_t n,
                                                           const double* __restrict__ rho,
                                                           const double* __restrict__ sigma,
                                                           double* __restrict__ eps,
                                                           double* __restrict__ vrho,
                                                           double* __restrict__ vsigma) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double r = rho[i];
  const XCOut o = dev_eval_functional(func, r, sigma ? sigma[i] : 0.0);
  eps[i] = r > 0.0 ? o.e / r : 0.0;
  vrho[i] = o.vr;
  if (vsigma) vsigma[i] = o.vs;
}

int launch_functional_points(hipStream_t s, int func, int64_t n, const double* rho,
                             const double* sigma, double* eps, double* vrho, double* vsigma) {
  if (n == 0) return XCQ_SUCCESS;
  hipLaunchKernelGGL(k_functional_points, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     func, n, rho, sigma, eps, vrho, vsigma);
  XCQ_HIP(hipGetLastError());
  return XCQ_SUCCESS;
}

}  // namespace xcq
