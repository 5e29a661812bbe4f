// Internal declarations shared by the HIP translation units of
// libgauxc_amd_device.so.  Not part of the public C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include <gauxc_amd/device.h>

namespace xcq {

// ---- error handling -------------------------------------------------------
void set_error(const std::string& msg);

#define XCQ_HIP(call)                                                               \
  do {                                                                              \
    hipError_t e_ = (call);                                                         \
    if (e_ != hipSuccess) {                                                         \
      ::xcq::set_error(std::string(#call) + ": " + hipGetErrorString(e_));          \
      return XCQ_ERROR_HIP;                                                         \
    }                                                                               \
  } while (0)

#define XCQ_REQUIRE(cond, code, msg) \
  do {                               \
    if (!(cond)) {                   \
      ::xcq::set_error(msg);         \
      return code;                   \
    }                                \
  } while (0)

// ---- device data ------------------------------------------------------------
constexpr int kMaxL = 4;          // highest angular momentum the kernels evaluate
constexpr int kTile = 64;         // GEMM tile edge (points and basis functions)
constexpr int kPtBlock = 64;      // points per collocation / functional block

struct DevShell {
  double alpha[32];
  double coeff[32];
  double O[3];
  int32_t nprim, l, pure, size;
};

// One task as laid out on the device.  Offsets into resident arrays are
// absolute; offsets into the chunk scratch are relative to the chunk.
struct DevTask {
  int32_t npts, nbe, nshells, iparent;
  int32_t npts_pad;   // npts rounded up to kTile
  int32_t ntile;      // ceil(nbe / kTile)
  int64_t pt_off;     // points (SoA) / weights
  int64_t sh_off;     // shell list
  int64_t bf_off;     // local -> global basis function map (nbe entries)
  int64_t lsh_off;    // local shell -> first local basis function (nshells entries)
  int64_t b_off;      // chunk scratch: [5][nbe][npts_pad] (B, Bx, By, Bz, Z) in elements
  int64_t q_off;      // chunk per-point arrays (npts_pad entries)
  int64_t red_off;    // global index of the task's first point block (EXC/NEL partials)
};

struct WorkItem {  // (task, a, b): a/b are point-block or tile indices
  int32_t task, a, b, pad_;
};

struct Chunk {
  int32_t task_begin, task_end;
  int64_t scratch_elems;  // elements of T (FP64 or FP32) used by B-family arrays
  int64_t nq;             // per-point array length
  // work lists (offsets into the concatenated device lists)
  int64_t pblk_off, npblk;    // (task, point block): collocation, density, Z
  int64_t g2_off, ng2;        // SYR2K tiles (task, itile, jtile) with itile >= jtile
};

// View of a context's device buffers for the mixed-precision translation
// unit (defined in context.hip).
struct CtxView {
  hipStream_t stream;
  const DevTask* tasks;
  const DevShell* shells;
  const int32_t *shell_list, *bfmap, *lsh;
  const double *px, *py, *pz, *w;
  const WorkItem* work;
  double* P;
  double* vxc;
  int32_t nbf;
  double* red;
  void* scratch;
  size_t scratch_bytes;
  void** Psplit;  // nbf x nbf packed (p1, p2, p3, 0) bfloat16 slices of P
};
CtxView view_of(struct ::xcq_context_s* c);

// ---- kernel launchers (defined in the .hip files) ---------------------------

int upload_sph_table(hipStream_t s);
int launch_ssf(hipStream_t s, int32_t natoms, const double* atom_xyz, const double* rab,
               const double* dist_nearest, int64_t npts, const double* px, const double* py,
               const double* pz, const int32_t* parent, double* w);
int launch_collocation_fp64(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                            const DevTask* tasks, const DevShell* shells, const int32_t* shell_list,
                            const int32_t* lsh_first, const double* px, const double* py,
                            const double* pz, double* scratch, int deriv);
int launch_density_fp64(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                        const DevTask* tasks, const int32_t* bfmap, const double* P,
                        int64_t ldp, double* scratch, const double* weights, int func, int gga,
                        double fac, int want_z, double* red);
int launch_density_mixed(hipStream_t s, const Chunk& c, const WorkItem* pblk,
                         const DevTask* tasks, const int32_t* bfmap, const uint2* Ps,
                         int64_t ldp, float* scratch, const double* weights, int func, int gga,
                         double fac, int want_z, double* red);
int launch_syr2k_scatter_fp64(hipStream_t s, const Chunk& c, const WorkItem* g2,
                              const DevTask* tasks, const int32_t* bfmap, const double* scratch,
                              double* vxc, int64_t ldv);
int launch_symmetrize(hipStream_t s, int32_t n, double* a, int64_t lda);
int launch_reduce_sums(hipStream_t s, const double* red, int64_t nblocks, double* out2);
int launch_functional_points(hipStream_t s, int func, int64_t n, const double* rho,
                             const double* sigma, double* eps, double* vrho, double* vsigma);

}  // namespace xcq
