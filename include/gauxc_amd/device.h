/* gauxc_amd device layer: the C ABI between the GauXC-compatible C++ host
 * library (libgauxc_amd.so) and the gfx950 HIP kernels
 * (libgauxc_amd_device.so).  Plain-old-data arguments, int status codes, no
 * exceptions or C++ types cross it.  The host library is its only caller in
 * the product; tests call it directly for kernel-level parity.
 *
 * Each entry point states the GauXC step it replaces.  All GauXC citations
 * are relative to the GauXC source tree.
 *
 * Status codes: XCQ_SUCCESS (0) or a non-zero error; xcq_last_error()
 * returns the calling thread's last message.
 */
#ifndef GAUXC_AMD_DEVICE_H
#define GAUXC_AMD_DEVICE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XCQ_SUCCESS 0
#define XCQ_ERROR_INVALID_ARGUMENT 1
#define XCQ_ERROR_HIP 2
#define XCQ_ERROR_OUT_OF_MEMORY 3
#define XCQ_ERROR_STATE 4
#define XCQ_ERROR_NO_DEVICE 5

/* Functionals evaluated on the device.  XCQ_FUNC_PBE0 is the XC part of
 * PBE0 (0.75 PBE exchange + PBE correlation); the exact-exchange share is
 * not an XC-quadrature quantity. */
enum xcq_functional { XCQ_FUNC_SVWN5 = 1, XCQ_FUNC_PBE = 2, XCQ_FUNC_PBE0 = 3 };

/* Arithmetic of the per-task contractions.
 *   XCQ_PREC_FP64  : everything in FP64.
 *   XCQ_PREC_MIXED : basis functions evaluated in FP64 and consumed in FP32
 *                    (the value plane with a bfloat16 residual); X = B*P_sub
 *                    on FP32 MFMA (FP32 sums within 32-deep K blocks, FP64
 *                    across them); B^T*Z + Z^T*B on BF16 MFMA over exact
 *                    bfloat16 slices of the FP32 operands; SSF weights, the
 *                    functional at every point, EXC/NEL and the VXC
 *                    accumulation in FP64, with EXC corrected to first order
 *                    for the FP32 rounding of P. */
enum xcq_precision { XCQ_PREC_FP64 = 0, XCQ_PREC_MIXED = 1 };

/* Contracted shell with normalised coefficients (GauXC Shell<double>,
 * include/gauxc/shell.hpp:48-229). */
typedef struct xcq_shell {
  int32_t nprim, l, pure, pad_;
  double alpha[32];
  double coeff[32];
  double origin[3];
  double cutoff_radius;
} xcq_shell;

/* One quadrature task (GauXC XCTask, include/gauxc/xc_task.hpp:25-119).
 * Points are x,y,z interleaved; point_offset indexes points (x3) and
 * weights, shell_offset indexes the shell-list array.  Shell lists are
 * ascending. */
typedef struct xcq_task {
  int32_t iparent, npts, nshells, nbe;
  int64_t point_offset, shell_offset;
  double dist_nearest;
} xcq_task;

typedef struct xcq_context_s* xcq_context;

/* ---- runtime (GauXC DeviceRuntimeEnvironment / DeviceBackend,
 *      src/runtime_environment/device/hip/hip_backend.cxx:16-131) ---- */
int xcq_device_count(int* count);
const char* xcq_last_error(void);
/* arena_bytes = 0: size the working arena as fill_fraction of free memory. */
int xcq_context_create(int device, size_t arena_bytes, double fill_fraction, xcq_context* ctx);
/* Context whose chunk scratch arena (the per-task basis-function planes,
 * the bulk of the device memory an evaluation needs) is the caller's device
 * buffer `mem` of `mem_bytes` bytes, 256-byte aligned, owned by the caller
 * and never freed here (GauXC DeviceRuntimeEnvironment(mem, mem_sz),
 * include/gauxc/runtime_environment/decl.hpp:71-72).  Task tables, points,
 * P and VXC stay in the context's own allocations.  An upload whose largest
 * task does not fit the buffer fails with "Insufficient memory to even
 * start!" (xc_device_stack_data.cxx:720-721). */
int xcq_context_create_mem(int device, void* mem, size_t mem_bytes, xcq_context* ctx);
int xcq_context_destroy(xcq_context ctx);
int xcq_synchronize(xcq_context ctx);

/* ---- static data (XCDeviceStackData::send_static_data_density_basis,
 *      src/xc_integrator/xc_data/device/xc_device_stack_data.cxx:350-384) ---- */
int xcq_upload_basis(xcq_context ctx, int32_t nshells, const xcq_shell* shells);
/* Task table + points/weights, resident in HBM until replaced
 * (replaces the per-batch pack_and_send of generate_buffers,
 *  xc_device_stack_data.cxx:713-751). */
int xcq_upload_tasks(xcq_context ctx, int32_t ntasks, const xcq_task* tasks, int64_t npts_total,
                     const double* points, const double* weights, int64_t nshell_entries,
                     const int32_t* shell_lists);
/* Density matrix, nbf x nbf column-major (host pointer). */
int xcq_upload_density(xcq_context ctx, int32_t nbf, const double* P, int64_t ldp);

/* ---- batched evaluation of independent molecules in one context (no GauXC
 *      counterpart: GauXC runs one molecule per integrator; this batches the
 *      task lists of many molecules into the same launches) ----
 * Shells of nmol molecules, concatenated molecule by molecule; mol_nshells[m]
 * shells belong to molecule m.  xcq_upload_basis is the nmol = 1 case. */
int xcq_upload_molecules(xcq_context ctx, int32_t nmol, const int32_t* mol_nshells,
                         const xcq_shell* shells);
/* Tasks of several molecules: task_mol[i] is the molecule of task i (NULL:
 * every task belongs to molecule 0); shell lists hold global shell indices
 * of that molecule's shells.  xcq_upload_tasks is task_mol = NULL. */
int xcq_upload_tasks_batched(xcq_context ctx, int32_t ntasks, const xcq_task* tasks,
                             const int32_t* task_mol, int64_t npts_total, const double* points,
                             const double* weights, int64_t nshell_entries,
                             const int32_t* shell_lists);
int xcq_upload_density_mol(xcq_context ctx, int32_t mol, int32_t nbf, const double* P,
                           int64_t ldp);
int xcq_download_results_mol(xcq_context ctx, int32_t mol, double* exc, double* nel,
                             double* vxc, int64_t ldvxc);

/* ---- scheduling (no GauXC counterpart beyond the task sort of
 *      incore_replicated_xc_device_integrator_exc_vxc.hpp:253-257) ----
 * Tile policy: tasks with nbe <= dens32_max_nbe run the density stage with
 * 32-function tiles (else 64); likewise syr2k32_max_nbe for B^T Z. */
int xcq_set_tile_policy(xcq_context ctx, int32_t dens32_max_nbe, int32_t syr2k32_max_nbe);
/* Density tasks with 64-function tiles and nbe <= hiocc_max_nbe run in
 * the kernel configuration that fits three workgroups per CU (mixed: a
 * two-block operand ring; FP64: the in-stage kernel's register budget);
 * larger ones keep two workgroups per CU.  0 (default) disables it. */
int xcq_set_density_occupancy(xcq_context ctx, int32_t hiocc_max_nbe);
/* Mixed precision: density tasks with nbe >= wide_min_nbe (and above the
 * 32-wide threshold) use 128-function tiles (the FP64 path keeps 64).
 * Default: off (2^30). */
int xcq_set_density_wide(xcq_context ctx, int32_t wide_min_nbe);
/* Where the rho fold and Z get the basis functions' values and gradients
 * (GauXC stores B and grad B, collocation_device_masked_combined_kernel_deriv1,
 * and reads them back in its rho and Z kernels): 3 (default) = the density
 * stage evaluates them at its points for the fold and leaves Z to a
 * separate launch that evaluates them again; 1 = the mixed density stage
 * evaluates them for both (the FP64 path runs mode 3); 0 = both read the
 * planes the collocation stage stores.  In modes 1 and 3 the collocation
 * stage stores only the value plane.  Bases with angular momentum above 3
 * always use the stored planes.  Other values: error. */
int xcq_set_density_basis_eval(xcq_context ctx, int32_t in_stage);
/* The basis evaluation mode in effect (0 when the basis exceeds l = 3). */
int xcq_get_density_basis_eval(xcq_context ctx, int32_t* mode);
/* B^T Z contributions with |v| < tol are not scattered into VXC (each
 * scatter-add is an FP64 atomic; on large diffuse systems most are of
 * negligible function pairs).  Default 1e-18; 0 adds every contribution.
 * Applies to the RKS and UKS EXC/VXC and to both FXC contractions.
 * Allowed range [0, 1e-10].  No GauXC counterpart (GauXC adds all). */
int xcq_set_vxc_drop_tolerance(xcq_context ctx, double tol);
/* Whether the EXC gradient's grid-weight part bounds its loops with
 * R_AB-sorted neighbour lists (default 1; they are built up to 8,192
 * atoms) or loops over every atom (0).  Both give the same gradient to
 * rounding; the switch exists to exercise both.  No GauXC counterpart. */
int xcq_set_ssf_neighbour_lists(xcq_context ctx, int32_t enable);
/* Current settings: out5 = {dens32_max_nbe, syr2k32_max_nbe, hiocc_max_nbe,
 * nqueues, wide_min_nbe}. */
int xcq_get_tile_policy(xcq_context ctx, int32_t* out5);
/* nqueues = 0: one workgroup per work item (large-first upload order,
 * XCD-interleaved).  nqueues > 0: persistent launches of nqueues workgroups
 * per stage list drawing items from dynamic queues: the list is split into
 * one sub-list per XCD (tasks assigned largest predicted cost first to the
 * least-loaded sub-list, a task's items kept together in its sub-list, tasks
 * in decreasing cost order), and every workgroup takes the next item of its
 * XCD's sub-list from a ticket counter until the sub-list is empty.  Costs
 * are per-task predictions (dens_cost, syr2k_cost, ntasks entries each,
 * split evenly over a task's items), or the tile-flop heuristic when both
 * are NULL. */
int xcq_set_schedule(xcq_context ctx, int32_t nqueues, const double* dens_cost,
                     const double* syr2k_cost);
/* Per-item wall-clock stamps in the density and B^T Z kernels; after a run,
 * xcq_task_times returns per-task sums in microseconds (ntasks entries
 * each): the timings the cost model trains on. */
int xcq_enable_task_timing(xcq_context ctx, int enable);
int xcq_task_times(xcq_context ctx, double* dens_us, double* syr2k_us);
/* Capture the whole evaluation into a HIP graph on first use of a
 * (functional, precision, want_vxc) combination and replay it afterwards.
 * Stage times then report only the total (stage 6). */
int xcq_use_graph(xcq_context ctx, int enable);
/* Autotuner: times the uploaded workload on the device and keeps the
 * fastest settings.  Sweeps the 32/64 tile thresholds of the density and
 * B^T Z stages (nbe buckets {0, 32, 48, 64, 96, 128, 192, 256, all}); per
 * basis evaluation mode (3 when the basis allows it, then 0) the density
 * 32-wide threshold and, for mixed precision, the occupancy threshold over
 * the same buckets and the 128-wide threshold over {128, 192, 256, 384,
 * 512, off}, keeping the mode with the fastest collocation + density + Z
 * launch; then the persistent-queue count {0, 2, 4 per CU}.  Leaves the context on the best and reports the 32-wide
 * thresholds and queue count (any output pointer may be NULL;
 * xcq_get_tile_policy and xcq_get_density_basis_eval return the rest). */
int xcq_autotune(xcq_context ctx, int functional, int precision, int32_t* dens32_max_nbe,
                 int32_t* syr2k32_max_nbe, int32_t* nqueues, double* best_ms);

/* ---- SSF partition weights (GauXC: DeviceMolecularWeights::modify_weights,
 *      src/molecular_weights/device/device_molecular_weights.cxx:18, kernels
 *      cuda_ssf_1d.cu:24-131; host reference weights.cxx:117-237) ----
 * Host buffers in, host buffers out: points (3*npts), per-point parent atom,
 * weights modified in place. */
int xcq_ssf_weights(int device, int32_t natoms, const double* atom_xyz, int64_t npts,
                    const double* points, const int32_t* parent, double* weights);
/* Allocation-free variant of xcq_ssf_weights, for callers that cannot
 * allocate device memory (code under HIP graph capture is not a target: the
 * call synchronises).  No counterpart in GauXC.  points (x,y,z interleaved,
 * 3*npts), parent and weights are DEVICE buffers owned by the caller;
 * atom_xyz is a host array.  The inter-atomic distances live in a static
 * device array of XCQ_SSF_NOALLOC_MAX_ATOMS atoms (8.4 MB per device, held
 * until the process ends); larger molecules are rejected.  No hipMalloc or
 * hipFree is called.  Calls that overlap on one device race on that array.
 * A parent index out of range is reported as an error after the launch. */
#define XCQ_SSF_NOALLOC_MAX_ATOMS 1024
int xcq_ssf_weights_noalloc(int device, int32_t natoms, const double* atom_xyz, int64_t npts,
                            const double* points, const int32_t* parent, double* weights);
/* Same, on the weights of the uploaded task set (device resident). */
int xcq_ssf_weights_tasks(xcq_context ctx, int32_t natoms, const double* atom_xyz);
/* Copy the (possibly SSF-modified) resident weights back to the host. */
int xcq_download_weights(xcq_context ctx, double* weights);

/* ---- the EXC/VXC local work (IncoreReplicatedXCDeviceIntegrator::
 *      exc_vxc_local_work_, incore_replicated_xc_device_integrator_exc_vxc.hpp:221-386):
 *      collocation+gradient, X = 2 B P_sub, rho/grad rho/sigma, functional,
 *      weight folding, EXC/NEL, Z, V_sub = B^T Z + Z^T B, scatter into VXC,
 *      symmetrisation.  Runs asynchronously on the context's stream;
 *      results stay in HBM until xcq_download_results.  want_vxc = 0 gives
 *      the EXC-only path (eval_exc). ---- */
int xcq_exc_vxc(xcq_context ctx, int functional, int precision, int want_vxc);
/* ---- UKS EXC/VXC (GauXC eval_exc_vxc(Ps, Pz): the generic driver with
 *      spin_dim 2, reference_replicated_xc_host_integrator_exc_vxc.hpp:106-601;
 *      device kernels cuda/kernels/zmat_vxc.cu:62-96,197-274).  P_s = P_a + P_b
 *      goes through xcq_upload_density, P_z = P_a - P_b through
 *      xcq_upload_density_z (single-molecule contexts).  FP64, stored basis
 *      planes, every VXC contribution added.  Synchronous results through
 *      xcq_download_results_uks (vxc_s / vxc_z may be NULL). ---- */
int xcq_upload_density_z(xcq_context ctx, int32_t nbf, const double* Pz, int64_t ldp);
int xcq_exc_vxc_uks(xcq_context ctx, int functional, int want_vxc);
int xcq_download_results_uks(xcq_context ctx, double* exc, double* nel, double* vxc_s,
                             double* vxc_z, int64_t ldvxc);
/* ---- RKS FXC contraction (GauXC eval_fxc_contraction(P, tP):
 *      reference_replicated_xc_host_integrator_fxc_contraction.hpp:94-600,
 *      device zmat_fxc.cu; C API gauxc_integrator_eval_fxc_contraction_rks,
 *      include/gauxc/c/xc_integrator.h:351-362).  P through
 *      xcq_upload_density, the trial density tP through
 *      xcq_upload_trial_density (single-molecule contexts).  FXC =
 *      d VXC[P + e tP] / de is left in the context (FP64, with the VXC drop
 *      tolerance of the energy path) and read back with xcq_download_fxc, which
 *      also returns the integrated density of P. ---- */
int xcq_upload_trial_density(xcq_context ctx, int32_t nbf, const double* tP, int64_t ldp);
int xcq_fxc_contraction(xcq_context ctx, int functional);
int xcq_download_fxc(xcq_context ctx, double* nel, double* fxc, int64_t ldfxc);
/* UKS FXC contraction (GauXC eval_fxc_contraction(Ps, Pz, tPs, tPz), the same
 * driver with spin_dim 2; C API gauxc_integrator_eval_fxc_contraction_uks,
 * include/gauxc/c/xc_integrator.h:383-400): P_s / P_z through
 * xcq_upload_density / xcq_upload_density_z, tP_s / tP_z through
 * xcq_upload_trial_density / xcq_upload_trial_density_z.  FXC_s and FXC_z
 * (GauXC's (FXC_a + FXC_b) / 2 and (FXC_a - FXC_b) / 2) come back with
 * xcq_download_fxc_uks (either may be NULL). */
int xcq_upload_trial_density_z(xcq_context ctx, int32_t nbf, const double* tPz, int64_t ldp);
int xcq_fxc_contraction_uks(xcq_context ctx, int functional);
int xcq_download_fxc_uks(xcq_context ctx, double* nel, double* fxc_s, double* fxc_z,
                         int64_t ldfxc);
/* ---- RKS EXC nuclear gradient with the SSF grid-weight derivatives (GauXC
 *      eval_exc_grad(P) with include_weight_derivatives:
 *      reference_replicated_xc_host_integrator_exc_grad.hpp:106-600,
 *      weights.cxx:806-990; device increment_exc_grad.cu, cuda_ssf_1d.cu; C
 *      API gauxc_integrator_eval_exc_grad_rks, include/gauxc/c/xc_integrator.h:
 *      282-290).  P through xcq_upload_density (single-molecule contexts);
 *      xyz = the natoms atom positions (x, y, z), in the order of the tasks'
 *      parent indices; a shell belongs to the atom at exactly its centre.
 *      FP64, stored basis planes.  weight_derivs = 0 gives the Hellmann-
 *      Feynman form (IntegratorSettingsEXC_GRAD::include_weight_derivatives =
 *      false).  xcq_download_exc_grad returns EXC, NEL and grad[3 * natoms]
 *      (atom-major). ---- */
int xcq_exc_grad(xcq_context ctx, int functional, int32_t natoms, const double* xyz,
                 int weight_derivs);
/* UKS EXC nuclear gradient (GauXC eval_exc_grad(Ps, Pz): the same driver's
 * spin_dim 2 branches, reference_replicated_xc_host_integrator_exc_grad.hpp:
 * 64-100,106-600; C API gauxc_integrator_eval_exc_grad_uks, include/gauxc/c/
 * xc_integrator.h:304-313).  P_s through xcq_upload_density, P_z through
 * xcq_upload_density_z; otherwise as xcq_exc_grad; results through
 * xcq_download_exc_grad. */
int xcq_exc_grad_uks(xcq_context ctx, int functional, int32_t natoms, const double* xyz,
                     int weight_derivs);
int xcq_download_exc_grad(xcq_context ctx, double* exc, double* nel, double* grad);
/* Integrated density only (eval_integrate_den). */
int xcq_integrate_den(xcq_context ctx, double* nel);
int xcq_download_results(xcq_context ctx, double* exc, double* nel, double* vxc, int64_t ldvxc);

/* ---- load-balancer screening on the device (GauXC's device load balancer:
 *      src/load_balancer/device/hip/replicated_hip_load_balancer.cxx:72-220,
 *      hip_collision_detection.hip:75-220).  Host buffers in and out: for
 *      batch b with box [lo[3b..], up[3b..]] and shell s with cutoff sphere
 *      (centers[3s..], radii[s]), bit s % 32 of bits[b * ld32 + s / 32] is set
 *      when the sphere meets the box (ld32 = ceil(nshells / 32)); counts[b]
 *      is the number of such shells and nbe[b] the sum of their
 *      shell_sizes.  The test is geometry::cube_sphere_intersect's, operation
 *      for operation. ---- */
int xcq_collision_detection(int device, int64_t nbatches, const double* lo, const double* up,
                            int32_t nshells, const double* centers, const double* radii,
                            const int32_t* shell_sizes, uint32_t* bits, int32_t* counts,
                            int32_t* nbe);

/* ---- kernel-level entry points used by parity tests (host buffers) ---- */
/* Collocation (+gradient if deriv = 1) of uploaded task `task`, written as
 * [comp][nbe][npts] (comp = value, d/dx, d/dy, d/dz) in FP64.  precision
 * selects the FP64 or FP32 collocation kernel. */
int xcq_task_collocation(xcq_context ctx, int32_t task, int deriv, int precision, double* out);
/* Functional on n points (device evaluation of host arrays).  sigma/vsigma
 * may be NULL for SVWN5. */
int xcq_eval_functional(int device, int functional, int64_t n, const double* rho,
                        const double* sigma, double* eps, double* vrho, double* vsigma);

/* ---- instrumentation ----
 * Per-stage device times of the last xcq_exc_vxc call in milliseconds,
 * measured with HIP events on the context's stream.  Stage order:
 * 0 collocation, 1 X + density reduction, 2 functional + EXC/NEL,
 * 3 Z build, 4 B^T Z + scatter, 5 symmetrise, 6 total.  Returns the number
 * of stages written (at most n). */
int xcq_stage_times(xcq_context ctx, double* ms, int n);
/* Number of kernel launches and tiles of the last call, for roofline
 * accounting: out[0] = launches, out[1] = total GEMM tiles, out[2] = sum
 * npts, out[3] = sum over tasks of npts*nbe*nbe. */
int xcq_work_counters(xcq_context ctx, int64_t* out, int n);

/* ---- cost-model network of the task scheduler (no GauXC counterpart:
 *      GauXC orders tasks by the flop heuristic XCTask::cost_exc_vxc,
 *      include/gauxc/xc_task.hpp:112-114) ----
 * FP32 MLP  nin -> [Linear -> ReLU] x nlayers (width nhidden) -> Linear ->
 * nout, trained on the device by this library's own forward, backward and
 * Adam kernels.  Parameters are one vector W1, b1, W2, b2, ... with W_l
 * stored [out][in] row-major.  Rows of x / t are row-major [n][nin] /
 * [n][nout] host arrays.  Loss: mean squared error. */
typedef struct xcq_mlp_s* xcq_mlp;
int xcq_mlp_create(int device, int32_t nin, int32_t nhidden, int32_t nlayers, int32_t nout,
                   uint64_t seed, xcq_mlp* out);
int xcq_mlp_destroy(xcq_mlp m);
int xcq_mlp_num_params(xcq_mlp m, int64_t* n);
int xcq_mlp_get_params(xcq_mlp m, float* host);
/* Setting parameters also resets the optimiser state. */
int xcq_mlp_set_params(xcq_mlp m, const float* host);
int xcq_mlp_forward(xcq_mlp m, int64_t n, const float* x, float* y);
/* Loss and its gradient with respect to the parameter vector (no update). */
int xcq_mlp_gradients(xcq_mlp m, int64_t n, const float* x, const float* t, float* loss,
                      float* grad);
/* Adam (lr, beta 0.9 / 0.999) over shuffled minibatches of `batch` rows for
 * `epochs` passes; final_loss is the loss of the last minibatch. */
int xcq_mlp_train(xcq_mlp m, int64_t n, const float* x, const float* t, int32_t epochs,
                  int32_t batch, float lr, uint64_t seed, float* final_loss);

#ifdef __cplusplus
}
#endif
#endif
