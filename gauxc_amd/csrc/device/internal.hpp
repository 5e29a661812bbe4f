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
constexpr int kPtBlock = 64;      // points per collocation / density block (npts_pad unit)
// Default |v| below which a task's B^T Z contribution is not scattered
// into VXC: far below the FP64 resolution of any VXC element above 1e-2,
// and its absolute effect (at most this times the number of tasks sharing a
// function pair) far below the parity bars.  Most scatter-adds of large,
// diffuse systems are of such negligible pairs (about three in four on a
// 150-water cluster), and each is an FP64 atomic.
constexpr double kDefaultVxcDrop = 1e-18;
constexpr int kBfPad = 32;        // zero entries after the last task's basis map: K blocks
                                  // of 32 rows read up to 31 entries past a task's nbe

struct DevShell {
  double alpha[32];
  double coeff[32];
  double O[3];
  int32_t nprim, l, pure, size;
};

// Compact copy of a shell's centre, first primitive and shape (48 bytes),
// which the density stage's basis evaluation loads 64 shells at a time.
struct DevShellC {
  double ox, oy, oz, a0, c0;
  int32_t info;  // l | pure << 8 | nprim << 16
  int32_t pad_;
};

// One task as laid out on the device.  Offsets into resident arrays are
// absolute; b_off is relative to the chunk scratch.
struct DevTask {
  int32_t npts, nbe, nshells, iparent;
  int32_t npts_pad;   // npts rounded up to kPtBlock
  int32_t mol;        // molecule of the task
  int32_t ldp;        // nbf of the molecule = leading dimension of its P and VXC
  int32_t nt_d;       // function tile of the density kernel (32, 64 or 128; the FP64
                      // kernel runs 128-wide tasks with 64-wide tiles)
  int32_t nt_s;       // function tile of the SYR2K kernel (32 or 64)
  int32_t dens_cls;   // density work list: 0 = 32-wide, 1 = 64-wide at three
                      // workgroups per CU, 2 = 64-wide at two, 3 = 128-wide
  int64_t pt_off;     // points (SoA) / weights
  int64_t sh_off;     // shell list
  int64_t bf_off;     // local -> molecule basis function map (nbe entries)
  int64_t lsh_off;    // local shell -> first local basis function (nshells entries)
  int64_t b_off;      // chunk scratch, FP64 storage: [5][nbe][npts_pad] doubles
                      // (B, Bx, By, Bz, Z), offset in doubles
  int64_t b_off_mx;   // chunk scratch, FP32 storage: the same five planes in
                      // floats + the value plane's bfloat16 residual
                      // [nbe][npts_pad], offset in floats (5.5 planes per task)
  int64_t red_off;    // global index of the task's first point block (EXC/NEL partials)
  int64_t tmax_off;   // first of the task's ceil(nbe / 32) function granules in the
                      // bound table (max |B| and max |Z| per 32 functions)
  int64_t p_off;      // offset of the molecule's P (and P slices, VXC) in elements
};

struct WorkItem {  // (task, a, b): a/b are point-block or tile indices
  int32_t task, a, b, pad_;
};

// A launchable work list: items [off, off + n) of the context's work array.
// nq > 0 runs it as a persistent launch of nq workgroups drawing items from
// dynamic queues: the list is split into nx sub-lists (nx = 8, one per XCD,
// when nq >= 8; else 1), sub-list x = items [qoff[x], qoff[x+1]) (queue
// table at qoff_off of the context's queue array, nx + 1 entries), and a
// workgroup b takes the next item of sub-list b % nx from that sub-list's
// ticket counter (counters at cnt_off of the context's counter array, zeroed
// at the start of every evaluation) until the sub-list is exhausted.
struct WorkRange {
  int64_t off = 0, n = 0;
  int64_t qoff_off = 0;
  int32_t nq = 0, nx = 0;
  int64_t cnt_off = 0;
};

// Launch arguments common to every queued stage kernel.
struct QueueArgs {
  const WorkItem* items;  // first item of the list
  const int32_t* qoff;    // sub-list table (nullptr: one workgroup per item)
  int64_t* tstamp;        // per-item (start, end) wall clock, or nullptr
  unsigned long long* counters;  // ticket counter per sub-list
  int32_t nx;             // sub-lists
};

struct Chunk {
  int32_t task_begin = 0, task_end = 0;
  int64_t scratch_elems = 0;  // elements of T (FP64 or FP32) used by B-family arrays
  WorkRange coll;             // all point blocks (collocation) = dens[0..3] back to back
  WorkRange dens[4];          // density blocks by class (DevTask::dens_cls)
  WorkRange syr2k[2];         // SYR2K tiles (it >= jt) with 32- / 64-function tiles
};

// View of a context's device buffers for the translation units that run
// pipeline stages (defined in context.hip).
struct CtxView {
  hipStream_t stream;
  const DevTask* tasks;
  const DevShell* shells;
  const int32_t *shell_list, *bfmap, *lsh;
  const double *px, *py, *pz, *w;
  const WorkItem* work;
  const int32_t* qoff;
  unsigned long long* counters;  // ticket counters of the dynamic queues
  int64_t* tstamp;  // nullptr unless task timing is on
  const double* P;
  const float* Pf32;  // FP32 copy of P (mixed-precision density stage)
  double* vxc;
  double* red;
  void* scratch;
  double vxc_drop;  // B^T Z contributions with |v| < vxc_drop are not added to VXC
  unsigned* tmax;   // per granule {max |B|, max |Z|} as float bits (the density stage
                    // fills it, the B^T Z stage skips tile pairs bounded below vxc_drop)
  int in_stage;     // density stage's basis evaluation: 0 = reads the stored B /
                    // grad B planes; 1 = evaluates them for its fold and Z (mixed; the
                    // FP64 path runs 3); 3 = for its fold, Z left to a separate launch
  const DevShellC* shc;
  double* zc;       // Z coefficients per point block [4][64] (in-stage mode 3)
};
// Basis data of the density stage's own evaluation of B and grad B.
struct ShellArgs {
  const DevShell* shells;       // primitives past the first
  const DevShellC* shc;
  const int32_t* shell_list;    // task shell lists (molecule shell indices)
  const int32_t* lsh;           // first local function of each task shell (ascending)
  const double *px, *py, *pz;   // point coordinates
};
constexpr int kInStageMaxL = 3;  // highest angular momentum it evaluates

constexpr int kGran = 32;      // functions per granule of the bound table
constexpr int kGranLds = 64;   // granules a density workgroup reduces in LDS (nbe <= 2048)

inline QueueArgs queue_args(const CtxView& v, const WorkRange& r) {
  return QueueArgs{v.work + r.off, r.nq ? v.qoff + r.qoff_off : nullptr,
                   v.tstamp ? v.tstamp + 2 * r.off : nullptr,
                   r.nq ? v.counters + r.cnt_off : nullptr, r.nx};
}
inline unsigned grid_of(const WorkRange& r) { return (unsigned)(r.nq ? r.nq : r.n); }

constexpr int64_t kNoItem = INT64_MAX;

// Next item of this workgroup's dynamic queue, or kNoItem.  Every thread of
// the workgroup calls it (uniform control flow): one lane takes a ticket
// with a vector atomic and the item index is broadcast through LDS.  The
// leading barrier keeps the previous item's readers of the slot ahead of
// the write.
__device__ __forceinline__ int64_t next_ticket(const QueueArgs& qa) {
  __shared__ int64_t slot;
  __syncthreads();
  if (threadIdx.x == 0) {
    const int x = (int)(blockIdx.x % (unsigned)qa.nx);
    const unsigned long long t = atomicAdd(qa.counters + x, 1ull);
    const int64_t n = qa.qoff[x + 1] - qa.qoff[x];
    slot = (int64_t)t < n ? qa.qoff[x] + (int64_t)t : kNoItem;
  }
  __syncthreads();
  return slot;
}

// Item loop of a queued stage kernel: one item per workgroup, or items from
// the workgroup's dynamic queue.  The body must end with all threads done
// with LDS.
#define XCQ_FOR_EACH_ITEM(qa, it_var)                                                  \
  for (int64_t it_var = (qa).qoff ? next_ticket(qa) : (int64_t)blockIdx.x,            \
               it_end_ = (qa).qoff ? kNoItem : it_var + 1;                             \
       it_var < it_end_; it_var = (qa).qoff ? next_ticket(qa) : it_var + 1)

// True when every element of the B^T Z + Z^T B tile pair (rows i0.., columns
// j0.., nt wide) of task t is bounded below `drop` by the granule maxima the
// density stage recorded: |sum_p B_pi Z_pj| <= npts max|B_i| max|Z_j|.
// The factor 1.01 covers the rounding of the stored maxima and of the
// products.
__device__ __forceinline__ bool tile_pair_negligible(const DevTask& t, const unsigned* tmax,
                                                     int i0, int j0, int nt, double drop) {
  if (!tmax) return false;
  const unsigned* tm = tmax + 2 * t.tmax_off;
  const int ng = (t.nbe + kGran - 1) / kGran;
  float bi = 0.f, zi = 0.f, bj = 0.f, zj = 0.f;
  for (int g = i0 / kGran; g < min((i0 + nt) / kGran, ng); ++g) {
    bi = fmaxf(bi, __uint_as_float(tm[2 * g]));
    zi = fmaxf(zi, __uint_as_float(tm[2 * g + 1]));
  }
  for (int g = j0 / kGran; g < min((j0 + nt) / kGran, ng); ++g) {
    bj = fmaxf(bj, __uint_as_float(tm[2 * g]));
    zj = fmaxf(zj, __uint_as_float(tm[2 * g + 1]));
  }
  const double bound = 1.01 * (double)t.npts * ((double)bi * zj + (double)zi * bj);
  return bound < drop;
}

__device__ __forceinline__ void stamp_item(const QueueArgs& qa, int64_t it, int64_t t0) {
  if (qa.tstamp && threadIdx.x == 0) {
    qa.tstamp[2 * it] = t0;
    qa.tstamp[2 * it + 1] = wall_clock64();
  }
}

// ---- kernel launchers (defined in the .hip files) ---------------------------
int upload_sph_table(hipStream_t s);
int launch_ssf(hipStream_t s, int32_t natoms, const double* atom_xyz, const double* rab,
               const double* dnear, int64_t npts, const double* px, const double* py,
               const double* pz, int64_t stride, const int32_t* parent, double* w,
               int32_t* bad_parent, const int32_t* nbr_idx = nullptr,
               const double* nbr_r = nullptr);
int launch_molmeta(hipStream_t s, int32_t natoms, const double* xyz, double* rab, double* dnear,
                   int32_t* bad_parent);
int launch_collocation_fp64(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first, const double* px,
                            const double* py, const double* pz, double* scratch, int deriv,
                            int lmax);
int launch_collocation_fp32(hipStream_t s, const WorkItem* items, int64_t n,
                            const DevTask* tasks, const DevShell* shells,
                            const int32_t* shell_list, const int32_t* lsh_first, const double* px,
                            const double* py, const double* pz, float* scratch, int deriv,
                            int lmax);
// Z = a B + c . grad B with B / grad B evaluated in FP64 (in-stage mode 3),
// stored in the precision's scratch layout, coefficients from zcoef, granule
// maxima into tmax (may be null).
int launch_z_in_stage(hipStream_t s, const WorkItem* items, int64_t n, const DevTask* tasks,
                      const DevShell* shells, const int32_t* shell_list,
                      const int32_t* lsh_first, const double* px, const double* py,
                      const double* pz, const double* zcoef, void* scratch, int precision,
                      unsigned* tmax, int gga, int lmax);
// Density stage (X, rho/grad rho, functional, EXC/NEL, Z) of one work range
// whose tasks use `nt`-function tiles (hiocc: the three-workgroups-per-CU
// configuration of the mixed kernel).  FP64 only, without hiocc:
// x_plane >= 0 also stores X = fac P_sub B into that scratch plane.
int launch_density(const CtxView& v, const WorkRange& r, int nt, bool hiocc, int precision,
                   int func, int gga, double fac, int want_z,
                   const double* P_override = nullptr, double* dens_out = nullptr,
                   int x_plane = -1);
// UKS: polarised functional and Z coefficients per point block from the two
// density passes' raw sums; Z planes from stored coefficients.
int launch_uks_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* ds,
                      const double* dz, double* zs, double* zz);
// FXC contraction (RKS): Z coefficients from the P and trial-density passes.
int launch_fxc_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* dp,
                      const double* dt, double* zc);
// UKS FXC contraction: derivatives of the UKS Z coefficients along the trial
// density from the four density passes.
int launch_fxc_points_uks(const CtxView& v, const WorkRange& r, int func, int gga,
                          const double* ds, const double* dz, const double* dts,
                          const double* dtz, double* zs, double* zz);
// EXC gradient: per-point coefficients, plane products out = fac P_sub in
// (+ out with accumulate; P defaults to the context's P), the
// basis-function part and the grid-weight part (gradient.hip).  RKS:
int launch_grad_points(const CtxView& v, const WorkRange& r, int func, int gga, const double* dp,
                       double* coef, double* wf);
int launch_plane_product(const CtxView& v, const WorkRange& r, int in_pl, int out_pl,
                         const double* P = nullptr, double fac = 2.0, int accumulate = 0);
int launch_grad_contract(const CtxView& v, const WorkRange& r, int gga,
                         const int32_t* shell_atom, const double* coef, int weight_derivs,
                         double* grad, int x_pl, int y_pl, int lmax);
// Za = a B + c . grad B from [block][4][64] coefficients into plane pl_s
// (and, with coef_z, a second one into pl_z; planes 3 and 4 only).  UKS:
// coefficients written over the P_s / P_z sums; contraction with Y
// (plane 1), X_N (2), X_Z (3).
int launch_grad_points_uks(const CtxView& v, const WorkRange& r, int func, int gga, double* ds,
                           double* dz, double* wf);
int launch_grad_zpair(const CtxView& v, const WorkRange& r, int gga, const double* coef_s,
                      const double* coef_z, int pl_s, int pl_z);
int launch_grad_contract_uks(const CtxView& v, const WorkRange& r, int gga,
                             const int32_t* shell_atom, const double* coef_s,
                             const double* coef_z, int weight_derivs, double* grad, int lmax);
int launch_ssf_grad(const CtxView& v, const WorkRange& r, int natoms, const double* xyz,
                    const double* dnear, const double* wf, double* grad,
                    const int32_t* nbr_idx = nullptr, const double* nbr_r = nullptr);
int launch_z_from_coeffs(const CtxView& v, const WorkRange& r, int gga, const double* coef);
// B^T Z + Z^T B with the VXC scatter-add, one work range of `nt`-wide tiles.
int launch_syr2k(const CtxView& v, const WorkRange& r, int nt, int precision);
int launch_P_to_f32(hipStream_t s, int64_t n, const double* P, float* out);
int launch_fill(hipStream_t s, double* p, int64_t n);
// Symmetrise the VXC of nmol molecules: mol_n[m] = nbf, mol_off[m] = offset.
int launch_symmetrize(hipStream_t s, int32_t nmol, int32_t max_n, const int32_t* mol_n,
                      const int64_t* mol_off, double* vxc);
// Partials per molecule of the FP32-rounding correction of EXC (four
// workgroups per CU, enough to stream a large molecule's P, P32 and VXC at
// HBM rate) and of the EXC/NEL block sums.
constexpr int kCorrBlocks = 1024;
constexpr int kRedBlocks = 64;
// part[m * kCorrBlocks + b] = partial sums of (P - P32) . VXC over molecule m.
int launch_p32_correction(hipStream_t s, int32_t nmol, const int32_t* mol_n,
                          const int64_t* mol_off, const double* P, const float* Pf,
                          const double* vxc, double* part);
// Per-molecule EXC/NEL: red pairs [blk_off[m], blk_off[m+1]) -> out2[2m], out2[2m+1],
// plus cfac * (the molecule's correction partials) on EXC when corr is not null;
// part holds 2 * kRedBlocks doubles per molecule.
int launch_reduce_sums(hipStream_t s, int32_t nmol, const int64_t* blk_off, const double* red,
                       double* part, const double* corr, double cfac, double* out2);
int launch_functional_points(hipStream_t s, int func, int64_t n, const double* rho,
                             const double* sigma, double* eps, double* vrho, double* vsigma);

}  // namespace xcq
