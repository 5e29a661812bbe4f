// Batch screening of the load balancer on the device: which shells' cutoff
// spheres meet each grid batch's bounding box (GauXC's device load balancer,
// src/load_balancer/device/hip/replicated_hip_load_balancer.cxx:72-220, and
// its collision kernels hip_collision_detection.hip:75-220).  The host keeps
// the batching, builds the tasks from the returned bit rows and sorts and
// merges them as the host load balancer does, so both produce the same task
// list.
//
// One 256-thread workgroup per batch; a wave tests 64 consecutive shells at
// a time (lane = shell), collects the hits with a ballot and stores the two
// 32-bit words of that 64-shell group; hit counts and function counts are
// summed per wave and then across the four waves in LDS.  The test is the
// host's (include/gauxc/load_balancer.hpp geometry::cube_sphere_intersect)
// in the same operation order with FMA contraction off, so every decision
// is bit-identical to the host's.
#include "internal.hpp"

#include <algorithm>
#include <vector>

namespace xcq {

__device__ __forceinline__ bool box_sphere_meet(const double* lo, const double* up, double cx,
                                                double cy, double cz, double rad) {
#pragma clang fp contract(off)
  const double c[3] = {cx, cy, cz};
  double rem = rad * rad;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    double gap = 0.;
    if (c[d] < lo[d]) gap = lo[d] - c[d];
    else if (c[d] > up[d]) gap = c[d] - up[d];
    rem -= gap * gap;
    if (rem < 0.) return false;
  }
  return true;
}

__global__ __launch_bounds__(256) void k_collision(int64_t nbatches, const double* __restrict__ lo,
                                                   const double* __restrict__ up, int32_t nshells,
                                                   const double* __restrict__ centers,
                                                   const double* __restrict__ radii,
                                                   const int32_t* __restrict__ sizes,
                                                   int32_t ld32, uint32_t* __restrict__ bits,
                                                   int32_t* __restrict__ counts,
                                                   int32_t* __restrict__ nbe) {
  __shared__ int red[4][2];
  const int64_t b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const double blo[3] = {lo[3 * b], lo[3 * b + 1], lo[3 * b + 2]};
  const double bup[3] = {up[3 * b], up[3 * b + 1], up[3 * b + 2]};
  int cnt = 0, nf = 0;
  uint32_t* row = bits + b * ld32;
  for (int s0 = 64 * wave; s0 < nshells; s0 += 256) {
    const int s = s0 + lane;
    bool hit = false;
    int sz = 0;
    if (s < nshells) {
      hit = box_sphere_meet(blo, bup, centers[3 * s], centers[3 * s + 1], centers[3 * s + 2],
                            radii[s]);
      sz = hit ? sizes[s] : 0;
    }
    const unsigned long long m = __ballot(hit);
    const int w0 = s0 >> 5;
    if (lane == 0 && w0 < ld32) row[w0] = (uint32_t)m;
    if (lane == 32 && w0 + 1 < ld32) row[w0 + 1] = (uint32_t)(m >> 32);
    cnt += __popcll(m);
    for (int off = 32; off > 0; off >>= 1) sz += __shfl_xor(sz, off);
    nf += sz;
  }
  if (lane == 0) {
    red[wave][0] = cnt;
    red[wave][1] = nf;
  }
  __syncthreads();
  if (tid == 0) {
    counts[b] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
    nbe[b] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  }
}

}  // namespace xcq

using namespace xcq;

extern "C" int xcq_collision_detection(int device, int64_t nbatches, const double* lo,
                                       const double* up, int32_t nshells,
                                       const double* centers, const double* radii,
                                       const int32_t* shell_sizes, uint32_t* bits,
                                       int32_t* counts, int32_t* nbe) {
  XCQ_REQUIRE(nbatches >= 0 && nshells >= 0, XCQ_ERROR_INVALID_ARGUMENT, "negative sizes");
  if (nbatches == 0) return XCQ_SUCCESS;
  XCQ_REQUIRE(lo && up && bits && counts && nbe, XCQ_ERROR_INVALID_ARGUMENT,
              "null batch arrays");
  XCQ_REQUIRE(nshells == 0 || (centers && radii && shell_sizes), XCQ_ERROR_INVALID_ARGUMENT,
              "null shell arrays");
  const int32_t ld32 = (nshells + 31) / 32;
  if (nshells == 0) {
    std::fill(counts, counts + nbatches, 0);
    std::fill(nbe, nbe + nbatches, 0);
    return XCQ_SUCCESS;
  }
  int n = 0;
  if (xcq_device_count(&n) != XCQ_SUCCESS || device < 0 || device >= n) {
    set_error("no HIP device available");
    return XCQ_ERROR_NO_DEVICE;
  }
  XCQ_HIP(hipSetDevice(device));
  // Batches go through in slabs that keep the bit rows below 256 MiB.
  const int64_t slab = std::max<int64_t>(1, std::min<int64_t>(
                                                nbatches, (int64_t(256) << 20) / (4 * ld32)));
  double *d_lo = nullptr, *d_up = nullptr, *d_c = nullptr, *d_r = nullptr;
  int32_t *d_sz = nullptr, *d_cnt = nullptr, *d_nbe = nullptr;
  uint32_t* d_bits = nullptr;
  auto release = [&]() {
    for (void* p : {(void*)d_lo, (void*)d_up, (void*)d_c, (void*)d_r, (void*)d_sz,
                    (void*)d_cnt, (void*)d_nbe, (void*)d_bits})
      if (p) (void)hipFree(p);
  };
  auto fail = [&](hipError_t e, const char* what) {
    (void)hipGetLastError();
    release();
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return XCQ_ERROR_HIP;
  };
  hipError_t e;
  if ((e = hipMalloc(&d_lo, 24 * slab)) != hipSuccess ||
      (e = hipMalloc(&d_up, 24 * slab)) != hipSuccess ||
      (e = hipMalloc(&d_c, 24 * (size_t)nshells)) != hipSuccess ||
      (e = hipMalloc(&d_r, 8 * (size_t)nshells)) != hipSuccess ||
      (e = hipMalloc(&d_sz, 4 * (size_t)nshells)) != hipSuccess ||
      (e = hipMalloc(&d_cnt, 4 * slab)) != hipSuccess ||
      (e = hipMalloc(&d_nbe, 4 * slab)) != hipSuccess ||
      (e = hipMalloc(&d_bits, 4 * (size_t)ld32 * slab)) != hipSuccess)
    return fail(e, "hipMalloc (collision detection)");
  if ((e = hipMemcpy(d_c, centers, 24 * (size_t)nshells, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_r, radii, 8 * (size_t)nshells, hipMemcpyHostToDevice)) != hipSuccess ||
      (e = hipMemcpy(d_sz, shell_sizes, 4 * (size_t)nshells, hipMemcpyHostToDevice)) !=
          hipSuccess)
    return fail(e, "hipMemcpy (collision detection)");
  for (int64_t b0 = 0; b0 < nbatches; b0 += slab) {
    const int64_t nb = std::min(slab, nbatches - b0);
    if ((e = hipMemcpy(d_lo, lo + 3 * b0, 24 * nb, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(d_up, up + 3 * b0, 24 * nb, hipMemcpyHostToDevice)) != hipSuccess)
      return fail(e, "hipMemcpy (collision detection)");
    hipLaunchKernelGGL(k_collision, dim3((unsigned)nb), dim3(256), 0, 0, nb, d_lo, d_up, nshells,
                       d_c, d_r, d_sz, ld32, d_bits, d_cnt, d_nbe);
    if ((e = hipGetLastError()) != hipSuccess) return fail(e, "k_collision launch");
    if ((e = hipMemcpy(bits + b0 * ld32, d_bits, 4 * (size_t)ld32 * nb,
                       hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(counts + b0, d_cnt, 4 * nb, hipMemcpyDeviceToHost)) != hipSuccess ||
        (e = hipMemcpy(nbe + b0, d_nbe, 4 * nb, hipMemcpyDeviceToHost)) != hipSuccess)
      return fail(e, "hipMemcpy (collision detection)");
  }
  release();
  return XCQ_SUCCESS;
}
