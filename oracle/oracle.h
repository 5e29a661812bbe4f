/* CPU oracle for the XC-quadrature path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain C++/OpenMP restatement of GauXC's host reference integrator
 * (ReferenceReplicatedXCHostIntegrator::exc_vxc_local_work_,
 * src/xc_integrator/replicated/host/reference_replicated_xc_host_integrator_exc_vxc.hpp:106-601)
 * and of the pieces it calls.  It is the checker of the HIP path and the
 * CPU baseline that bench.py times; nothing in the product links it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.
 *
 * Layouts: basis values are [npts][nbe] (nbe fastest, as GauXC's host
 * driver), matrices are column-major with leading dimension ld.
 *
 * How each function is pinned (tests/test_oracle.py): GauXC's fixtures --
 * benzene SVWN5 / PBE0 EXC, VXC, FXC, EXC_GRAD and EXC_GRAD_HELLFEY on the
 * unpruned, Robust and Treutler ultra-fine grids; Li / SVWN5 UKS EXC, VXC
 * and FXC; doublet cytosine / SVWN5 UKS EXC, VXC and both gradients; the
 * SSF weights of 15 tasks; water collocation and GauXC's own gau2grid
 * compiled under oracle/_ref -- and mathematics: central differences of
 * every functional derivative, of VXC along a trial density (FXC) and of
 * EXC under nuclear displacement (gradients, RKS and UKS), closed-form
 * limits, the unpolarised limit of the polarised functionals, UKS with
 * equal spins against RKS, translational invariance, Gaussian quadrature.
 * Parity unpinned: per-point values of plain PBE (1.0 PBE_X) beyond the
 * PBE0 fixture's 0.75 PBE_X + PBE_C (exchange enters linearly, so only its
 * scale is left open, and the scale is the published 1.0), and polarised
 * PBE at spin polarisation != 0 against GauXC itself (pinned by the
 * mathematical checks above only).
 */
#ifndef GAUXC_ORACLE_H
#define GAUXC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_shell {
  int32_t nprim, l, pure, pad_;
  double alpha[32];
  double coeff[32]; /* normalised contraction coefficients */
  double origin[3];
  double cutoff_radius;
} oracle_shell;

typedef struct oracle_task {
  int32_t iparent, npts, nshells, nbe;
  int64_t point_offset; /* into the points (x,y,z interleaved) and weights arrays */
  int64_t shell_offset; /* into the shell-list array */
  double dist_nearest;
} oracle_task;

enum { ORACLE_SVWN5 = 1, ORACLE_PBE = 2, ORACLE_PBE0 = 3 };

/* SSF partition weights, in place (weights.cxx:117-237). */
void oracle_ssf_weights(int32_t natoms, const double* atom_xyz, int32_t ntasks,
                        const oracle_task* tasks, const double* points, double* weights,
                        int nthreads);

/* Collocation (+gradient when dx != NULL) of one shell list at npts points;
 * outputs [npts][nbe] (gau2grid_collocation.cxx:25-145 semantics). */
void oracle_collocation(const oracle_shell* shells, int32_t nlist, const int32_t* shell_list,
                        int32_t npts, const double* points, double* b, double* dx, double* dy,
                        double* dz);

/* Functional on n points: eps (per particle), vrho, vsigma (GGA only). */
int oracle_functional(int func, int64_t n, const double* rho, const double* sigma, double* eps,
                      double* vrho, double* vsigma);

/* Host BLAS for the two per-task products (dgemm_, dsyr2k_ with 32-bit
 * integers, as GauXC's host/blas.cxx:29,38 calls them), loaded with dlopen
 * from `path`; NULL or "" returns to the plain loops.  Returns 0 on success
 * (the loaded path is copied to `which`), 1 if the library does not load,
 * 2 if it lacks either symbol (the loops stay in use). */
int oracle_use_blas(const char* path, char* which, int nwhich);

/* Full EXC/VXC local work over the given tasks (weights already partitioned).
 * VXC is nbf x nbf column-major, symmetrised on return.  Returns 0 on success. */
int oracle_exc_vxc(int func, int32_t nbf, const oracle_shell* shells, const double* P,
                   int64_t ldp, int32_t ntasks, const oracle_task* tasks, const double* points,
                   const double* weights, const int32_t* shell_lists, double* exc, double* nel,
                   double* vxc, int64_t ldvxc, int nthreads);

/* Spin-polarised functional on n points: rho = (rho_a, rho_b) pairs, sigma =
 * (aa, ab, bb) triples (GGA only); eps per particle, vrho pairs, vsigma
 * triples. */
int oracle_functional_pol(int func, int64_t n, const double* rho, const double* sigma,
                          double* eps, double* vrho, double* vsigma);

/* UKS EXC/VXC local work: P_s = P_a + P_b and P_z = P_a - P_b in, VXC_s and
 * VXC_z out (symmetrised), GauXC's generic host driver with spin_dim 2. */
int oracle_exc_vxc_uks(int func, int32_t nbf, const oracle_shell* shells, const double* Ps,
                       int64_t ldps, const double* Pz, int64_t ldpz, int32_t ntasks,
                       const oracle_task* tasks, const double* points, const double* weights,
                       const int32_t* shell_lists, double* exc, double* nel, double* vxc_s,
                       double* vxc_z, int64_t ldvxc, int nthreads);

/* Unpolarised functional with directional second derivatives (FXC
 * contraction): vrho, vsigma (GGA) and
 *   a = v2rho2 drho + v2rhosigma dsigma,  b = v2rhosigma drho + v2sigma2 dsigma
 * per point (b, vsigma, sigma and dsigma GGA only). */
int oracle_functional_fxc(int func, int64_t n, const double* rho, const double* sigma,
                          const double* drho, const double* dsigma, double* vrho,
                          double* vsigma, double* a, double* b);

/* RKS FXC contraction: FXC = d VXC[P + e tP] / de at e = 0, GauXC's host
 * driver (nbf x nbf column-major, symmetrised); nel = integrated density. */
int oracle_fxc_contraction(int func, int32_t nbf, const oracle_shell* shells, const double* P,
                           int64_t ldp, const double* tP, int64_t ldtp, int32_t ntasks,
                           const oracle_task* tasks, const double* points,
                           const double* weights, const int32_t* shell_lists, double* nel,
                           double* fxc, int64_t ldfxc, int nthreads);

/* UKS EXC nuclear gradient (P_s, P_z; polarised functional), forms and
 * layout as oracle_exc_grad. */
int oracle_exc_grad_uks(int func, int32_t nbf, const oracle_shell* shells, int32_t natoms,
                        const double* xyz, const double* Ps, const double* Pz, int64_t ldp,
                        int32_t ntasks, const oracle_task* tasks, const double* points,
                        const double* weights, const int32_t* shell_lists, int weight_derivs,
                        double* grad, int nthreads);

/* UKS FXC contraction: FXC_s, FXC_z = d (VXC_s, VXC_z)[P + e tP] / de for
 * P = (P_s, P_z), tP = (tP_s, tP_z) (all with leading dimension ldp). */
int oracle_fxc_contraction_uks(int func, int32_t nbf, const oracle_shell* shells,
                               const double* Ps, const double* Pz, const double* tPs,
                               const double* tPz, int64_t ldp, int32_t ntasks,
                               const oracle_task* tasks, const double* points,
                               const double* weights, const int32_t* shell_lists,
                               double* fxc_s, double* fxc_z, int64_t ldfxc, int nthreads);

/* Collocation with gradient and Hessian of one shell list: out[0..9] =
 * B, dB/dx, dB/dy, dB/dz, d2B/dxx, dxy, dxz, dyy, dyz, dzz, each [npts][nbe]
 * (gau2grid's deriv2 semantics). */
void oracle_collocation_hessian(const oracle_shell* shells, int32_t nlist,
                                const int32_t* shell_list, int32_t npts, const double* points,
                                double* const* out);

/* RKS EXC nuclear gradient, with the SSF grid-weight derivatives
 * (weight_derivs = 1, GauXC's default) or the Hellmann-Feynman form without
 * them (IntegratorSettingsEXC_GRAD::include_weight_derivatives): grad[3 * natoms],
 * atom-major (x, y, z).  Shells are assigned to the atom at exactly their
 * origin. */
int oracle_exc_grad(int func, int32_t nbf, const oracle_shell* shells, int32_t natoms,
                    const double* xyz, const double* P, int64_t ldp, int32_t ntasks,
                    const oracle_task* tasks, const double* points, const double* weights,
                    const int32_t* shell_lists, int weight_derivs, double* grad,
                    int nthreads);

#ifdef __cplusplus
}
#endif
#endif
