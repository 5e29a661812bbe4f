/* GauXC-compatible C API for the XC-quadrature path.
 *
 * Function names, handle layout ({header, pointer} structs passed by value)
 * and the status convention follow GauXC's C API (the headers under include/gauxc/c,
 * implementation under src/c-api) so a C/Fortran/Python caller of GauXC's
 * eval_exc_vxc path links against this library unchanged.  Only the
 * functions of that path are provided.  Every function reports through
 * GauXCStatus: code 0 on success, 1 on error with a malloc'd message freed
 * by gauxc_status_delete (GauXC: include/gauxc/c/status.h:24-33,
 * src/c-api/c_xc_integrator.cxx:320-346).
 */
#ifndef GAUXC_AMD_C_GAUXC_H
#define GAUXC_AMD_C_GAUXC_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GauXCStatus {
  int code;
  char* message;
} GauXCStatus;
void gauxc_status_delete(GauXCStatus* status);

enum GauXC_Type {
  GauXC_Type_Molecule = 1,
  GauXC_Type_BasisSet = 2,
  GauXC_Type_MolGrid = 3,
  GauXC_Type_RuntimeEnvironment = 4,
  GauXC_Type_LoadBalancer = 5,
  GauXC_Type_LoadBalancerFactory = 6,
  GauXC_Type_MolecularWeights = 7,
  GauXC_Type_MolecularWeightsFactory = 8,
  GauXC_Type_Functional = 9,
  GauXC_Type_Integrator = 10
};
typedef struct GauXCHeader {
  enum GauXC_Type type;
} GauXCHeader;

/* Enumerations (GauXC include/gauxc/c/enums.h:24-84, same order). */
enum GauXC_RadialQuad {
  GauXC_RadialQuad_Becke,
  GauXC_RadialQuad_MuraKnowles,
  GauXC_RadialQuad_MurrayHandyLaming,
  GauXC_RadialQuad_TreutlerAhlrichs
};
enum GauXC_AtomicGridSizeDefault {
  GauXC_AtomicGridSizeDefault_FineGrid,
  GauXC_AtomicGridSizeDefault_UltraFineGrid,
  GauXC_AtomicGridSizeDefault_SuperFineGrid,
  GauXC_AtomicGridSizeDefault_GM3,
  GauXC_AtomicGridSizeDefault_GM5
};
enum GauXC_XCWeightAlg {
  GauXC_XCWeightAlg_NOTPARTITIONED,
  GauXC_XCWeightAlg_Becke,
  GauXC_XCWeightAlg_SSF,
  GauXC_XCWeightAlg_LKO
};
enum GauXC_ExecutionSpace { GauXC_ExecutionSpace_Host, GauXC_ExecutionSpace_Device };
enum GauXC_PruningScheme {
  GauXC_PruningScheme_Unpruned,
  GauXC_PruningScheme_Robust,
  GauXC_PruningScheme_Treutler
};

typedef struct GauXCAtom {
  int64_t Z;
  double x, y, z; /* bohr */
} GauXCAtom;

typedef struct GauXCShell {
  int32_t l;
  bool pure;
  int32_t nprim;
  double exponents[32];
  double coefficients[32];
  double origin[3];
  double shell_tolerance;
} GauXCShell;

#define GAUXC_HANDLE(NAME) \
  typedef struct NAME {    \
    GauXCHeader hdr;       \
    void* ptr;             \
  } NAME;
GAUXC_HANDLE(GauXCMolecule)
GAUXC_HANDLE(GauXCBasisSet)
GAUXC_HANDLE(GauXCMolGrid)
GAUXC_HANDLE(GauXCLoadBalancer)
GAUXC_HANDLE(GauXCLoadBalancerFactory)
GAUXC_HANDLE(GauXCMolecularWeights)
GAUXC_HANDLE(GauXCMolecularWeightsFactory)
GAUXC_HANDLE(GauXCFunctional)
GAUXC_HANDLE(GauXCIntegrator)
#undef GAUXC_HANDLE

/* Runtime environment handle of a device build of GauXC
 * (include/gauxc/c/runtime_environment.h:31-37 with GAUXC_HAS_DEVICE):
 * `ptr` holds a host RuntimeEnvironment, `device_ptr` a
 * DeviceRuntimeEnvironment; exactly one of them is set. */
typedef struct GauXCRuntimeEnvironment {
  GauXCHeader hdr;
  void* ptr;
  void* device_ptr;
} GauXCRuntimeEnvironment;

typedef struct GauXCMolecularWeightsSettings {
  enum GauXC_XCWeightAlg weight_alg;
  bool becke_size_adjustment;
} GauXCMolecularWeightsSettings;

/* Molecule (GauXC c/molecule.h) */
GauXCMolecule gauxc_molecule_new(GauXCStatus* status);
GauXCMolecule gauxc_molecule_new_from_atoms(GauXCStatus* status, const GauXCAtom* atoms,
                                            size_t natoms);
void gauxc_molecule_delete(GauXCStatus* status, GauXCMolecule* mol);
size_t gauxc_molecule_natoms(GauXCStatus* status, const GauXCMolecule mol);

/* Basis set (GauXC c/basisset.h) */
GauXCBasisSet gauxc_basisset_new(GauXCStatus* status);
GauXCBasisSet gauxc_basisset_new_from_shells(GauXCStatus* status, const GauXCShell* shells,
                                             size_t nshells, bool normalize);
void gauxc_basisset_delete(GauXCStatus* status, GauXCBasisSet* basis);
int32_t gauxc_basisset_nbf(GauXCStatus* status, const GauXCBasisSet basis);
/* Normalised shells and screening radii as the integrator sees them
 * (extension: GauXC exposes these only through C++). */
void gauxc_basisset_get_shells(GauXCStatus* status, const GauXCBasisSet basis,
                               GauXCShell* shells, double* cutoff_radii);

/* Molecular grid (GauXC c/molgrid.h) */
GauXCMolGrid gauxc_molgrid_new_default(GauXCStatus* status, const GauXCMolecule mol,
                                       enum GauXC_PruningScheme pruning_scheme,
                                       int64_t batchsize, enum GauXC_RadialQuad radial_quad,
                                       enum GauXC_AtomicGridSizeDefault grid_size);
void gauxc_molgrid_delete(GauXCStatus* status, GauXCMolGrid* molgrid);

/* Runtime environment (GauXC c/runtime_environment.h; no MPI communicator).
 * A host environment runs device work on device 0 with fill fraction 0.9.
 * The device environment's fill fraction sizes the integrator's working
 * arena from free device memory; _new_mem hands the integrator a caller-owned
 * device buffer (256-byte aligned) as that arena instead
 * (runtime_environment.h:89-108). */
GauXCRuntimeEnvironment gauxc_runtime_environment_new(GauXCStatus* status);
GauXCRuntimeEnvironment gauxc_device_runtime_environment_new(GauXCStatus* status,
                                                             double fill_fraction);
GauXCRuntimeEnvironment gauxc_device_runtime_environment_new_mem(GauXCStatus* status, void* mem,
                                                                 size_t mem_sz);
void gauxc_runtime_environment_delete(GauXCStatus* status, GauXCRuntimeEnvironment* env);
int gauxc_runtime_environment_comm_rank(GauXCStatus* status, const GauXCRuntimeEnvironment env);
int gauxc_runtime_environment_comm_size(GauXCStatus* status, const GauXCRuntimeEnvironment env);

/* Load balancer (GauXC c/load_balancer.h) */
GauXCLoadBalancerFactory gauxc_load_balancer_factory_new(GauXCStatus* status,
                                                         enum GauXC_ExecutionSpace ex,
                                                         const char* kernel_name);
void gauxc_load_balancer_factory_delete(GauXCStatus* status, GauXCLoadBalancerFactory* f);
GauXCLoadBalancer gauxc_load_balancer_factory_get_instance(GauXCStatus* status,
                                                           const GauXCLoadBalancerFactory f,
                                                           const GauXCRuntimeEnvironment env,
                                                           const GauXCMolecule mol,
                                                           const GauXCMolGrid mg,
                                                           const GauXCBasisSet basis);
void gauxc_load_balancer_delete(GauXCStatus* status, GauXCLoadBalancer* lb);
/* Task inspection (extension, mirrors LoadBalancer::get_tasks()):
 * counts[0] = ntasks, counts[1] = total points, counts[2] = total shell-list
 * entries.  export fills caller arrays sized from the counts: per task
 * iparent, npts, nshells, nbe, dist_nearest; points (3 per point),
 * weights, concatenated shell lists. */
void gauxc_load_balancer_task_counts(GauXCStatus* status, const GauXCLoadBalancer lb,
                                     int64_t* counts);
void gauxc_load_balancer_export_tasks(GauXCStatus* status, const GauXCLoadBalancer lb,
                                      int32_t* iparent, int32_t* npts, int32_t* nshells,
                                      int32_t* nbe, double* dist_nearest, double* points,
                                      double* weights, int32_t* shell_lists);
bool gauxc_load_balancer_weights_modified(GauXCStatus* status, const GauXCLoadBalancer lb);

/* Molecular weights (GauXC c/molecular_weights.h) */
GauXCMolecularWeightsFactory gauxc_molecular_weights_factory_new(
    GauXCStatus* status, enum GauXC_ExecutionSpace ex, const char* local_work_kernel_name,
    GauXCMolecularWeightsSettings settings);
void gauxc_molecular_weights_factory_delete(GauXCStatus* status,
                                            GauXCMolecularWeightsFactory* f);
GauXCMolecularWeights gauxc_molecular_weights_factory_get_instance(
    GauXCStatus* status, const GauXCMolecularWeightsFactory f);
void gauxc_molecular_weights_modify_weights(GauXCStatus* status, const GauXCMolecularWeights mw,
                                            const GauXCLoadBalancer lb);
void gauxc_molecular_weights_delete(GauXCStatus* status, GauXCMolecularWeights* mw);

/* Functional (GauXC c/functional.h; "SVWN5", "PBE", "PBE0") */
GauXCFunctional gauxc_functional_from_string(GauXCStatus* status, const char* spec,
                                             bool polarized);
void gauxc_functional_delete(GauXCStatus* status, GauXCFunctional* func);

/* Integrator (GauXC c/xc_integrator.h) */
GauXCIntegrator gauxc_integrator_new(GauXCStatus* status, const GauXCFunctional functional,
                                     const GauXCLoadBalancer lb,
                                     enum GauXC_ExecutionSpace execution_space,
                                     const char* integrator_input_type,
                                     const char* integrator_kernel_name,
                                     const char* local_work_kernel_name,
                                     const char* reduction_kernel_name);
void gauxc_integrator_delete(GauXCStatus* status, GauXCIntegrator* integrator);
void gauxc_integrator_integrate_den(GauXCStatus* status, const GauXCIntegrator integrator,
                                    const int64_t m, const int64_t n,
                                    const double* density_matrix, const int64_t ldp,
                                    double* den);
void gauxc_integrator_eval_exc_rks(GauXCStatus* status, const GauXCIntegrator integrator,
                                   const int64_t m, const int64_t n,
                                   const double* density_matrix, const int64_t ldp,
                                   double* exc);
void gauxc_integrator_eval_exc_vxc_rks(GauXCStatus* status, const GauXCIntegrator integrator,
                                       const int64_t m, const int64_t n,
                                       const double* density_matrix, const int64_t ldp,
                                       double* exc, double* vxc_matrix, const int64_t ldvxc);
/* UKS (GauXC include/gauxc/c/xc_integrator.h:113-133, 193-222): total density
 * P_s = P_a + P_b and spin density P_z = P_a - P_b; needs a polarized
 * functional (FP64 local work driver). */
void gauxc_integrator_eval_exc_uks(GauXCStatus* status, const GauXCIntegrator integrator,
                                   const int64_t m, const int64_t n,
                                   const double* density_matrix_s, const int64_t ldp_s,
                                   const double* density_matrix_z, const int64_t ldp_z,
                                   double* exc);
void gauxc_integrator_eval_exc_vxc_uks(GauXCStatus* status, const GauXCIntegrator integrator,
                                       const int64_t m, const int64_t n,
                                       const double* density_matrix_s, const int64_t ldp_s,
                                       const double* density_matrix_z, const int64_t ldp_z,
                                       double* exc, double* vxc_matrix_s, const int64_t vxc_ld_s,
                                       double* vxc_matrix_z, const int64_t vxc_ld_z);
/* RKS EXC nuclear gradient with the grid-weight derivatives (GauXC
 * include/gauxc/c/xc_integrator.h:282-290): exc_grad[3 * natoms], atom-major;
 * FP64 local work driver, unpolarized functional. */
void gauxc_integrator_eval_exc_grad_rks(GauXCStatus* status, const GauXCIntegrator integrator,
                                        const int64_t m, const int64_t n,
                                        const double* density_matrix, const int64_t ldp,
                                        double* exc_grad);
/* UKS EXC nuclear gradient (GauXC include/gauxc/c/xc_integrator.h:304-313):
 * P_s and P_z; FP64 local work driver, polarized functional. */
void gauxc_integrator_eval_exc_grad_uks(GauXCStatus* status, const GauXCIntegrator integrator,
                                        const int64_t m, const int64_t n,
                                        const double* density_matrix_s, const int64_t ldp_s,
                                        const double* density_matrix_z, const int64_t ldp_z,
                                        double* exc_grad);
/* Extension: the Hellmann-Feynman gradient (no grid-weight derivatives;
 * GauXC's IntegratorSettingsEXC_GRAD::include_weight_derivatives = false,
 * which its C API does not expose). */
void gauxc_integrator_eval_exc_grad_hf_rks(GauXCStatus* status,
                                           const GauXCIntegrator integrator, const int64_t m,
                                           const int64_t n, const double* density_matrix,
                                           const int64_t ldp, double* exc_grad);
void gauxc_integrator_eval_exc_grad_hf_uks(GauXCStatus* status,
                                           const GauXCIntegrator integrator, const int64_t m,
                                           const int64_t n, const double* density_matrix_s,
                                           const int64_t ldp_s, const double* density_matrix_z,
                                           const int64_t ldp_z, double* exc_grad);
/* RKS FXC contraction (GauXC include/gauxc/c/xc_integrator.h:351-362):
 * fxc = d VXC[P + e tP] / de for the trial density tP; FP64 local work
 * driver, unpolarized functional. */
void gauxc_integrator_eval_fxc_contraction_rks(GauXCStatus* status,
                                               const GauXCIntegrator integrator,
                                               const int64_t m, const int64_t n,
                                               const double* density_matrix, const int64_t ldp,
                                               const double* t_density_matrix,
                                               const int64_t ldtp, double* fxc,
                                               const int64_t ldfxc);
/* UKS FXC contraction (GauXC include/gauxc/c/xc_integrator.h:383-400): total
 * and spin densities and trial densities in, FXC_s and FXC_z out; FP64 local
 * work driver, polarized functional. */
void gauxc_integrator_eval_fxc_contraction_uks(
    GauXCStatus* status, const GauXCIntegrator integrator, const int64_t m, const int64_t n,
    const double* density_matrix_s, const int64_t ldp_s, const double* density_matrix_z,
    const int64_t ldp_z, const double* t_density_matrix_s, const int64_t ldtp_s,
    const double* t_density_matrix_z, const int64_t ldtp_z, double* fxc_s,
    const int64_t ldfxc_s, double* fxc_z, const int64_t ldfxc_z);
/* Extension: the integrator's device context (an xcq_context of
 * include/gauxc_amd/device.h) for stage timings; NULL before first use. */
void* gauxc_integrator_device_context(GauXCStatus* status, const GauXCIntegrator integrator);

/* Type-dispatched deletion (GauXC c/types.h:49-76): *handle points at any
 * handle struct above; its header selects the delete function, which
 * clears the handle's pointer.  NULL entries are skipped. */
void gauxc_object_delete(GauXCStatus* status, void** handle);
void gauxc_objects_delete(GauXCStatus* status, void** handles, size_t nhandles);

#ifdef __cplusplus
}
#endif
#endif
