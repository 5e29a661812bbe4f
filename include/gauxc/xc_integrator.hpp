// XC integrator and its factory.
//
// GauXC counterparts: include/gauxc/xc_integrator.hpp:29-94 (value API),
// include/gauxc/xc_integrator/integrator_factory.hpp:40-102 (string-keyed
// factory), src/xc_integrator/replicated/replicated_xc_integrator_impl.hpp:
// 57-59 (raw-pointer seam).  This library provides ExecutionSpace::Device:
// XCIntegrator<M>::eval_exc_vxc(P) runs the whole quadrature on one GPU
// through the device C ABI (include/gauxc_amd/device.h).
//
// Factory strings (upper-cased before matching, as GauXC does):
//   input type        "Replicated"
//   integrator kernel "Default" | "Incore"
//   local work driver "Default" | "Scheme1"   -> FP64 kernels
//                     "Scheme1-Mixed"         -> mixed precision kernels
//   reduction kernel  "Default" | "BasicMPI" | "NCCL" (one process: no-op)
// MatrixType needs value_type, rows(), cols(), data() (column-major) and a
// (rows, cols) constructor -- Eigen::MatrixXd or GauXC::HostMatrix.
#pragma once
#include <gauxc/exceptions.hpp>
#include <gauxc/load_balancer.hpp>
#include <gauxc/xc_functional.hpp>
#include <cstdint>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

namespace GauXC {

struct IntegratorSettingsXC {
  virtual ~IntegratorSettingsXC() noexcept = default;
};
struct IntegratorSettingsKS : public IntegratorSettingsXC {
  double gks_dtol = 1e-12;
};
/// GauXC include/gauxc/xc_integrator_settings.hpp:28-30: with the grid-weight
/// derivatives (and translational invariance for the parent atoms), or the
/// Hellmann-Feynman gradient alone.
struct IntegratorSettingsEXC_GRAD : public IntegratorSettingsKS {
  bool include_weight_derivatives = true;
};

struct LocalWorkSettings {};

enum class DevicePrecision { FP64 = 0, Mixed = 1 };

namespace util {
/// Named wall-clock timings in milliseconds (GauXC util::Timer).
class Timer {
  std::map<std::string, double> t_;
 public:
  void set(const std::string& k, double ms) { t_[k] = ms; }
  double get_duration(const std::string& k) const {
    auto it = t_.find(k);
    return it == t_.end() ? 0. : it->second;
  }
  const std::map<std::string, double>& all_timings() const { return t_; }
};
}  // namespace util

/// Minimal column-major matrix for callers without Eigen.
class HostMatrix {
  int64_t m_ = 0, n_ = 0;
  std::vector<double> d_;
 public:
  using value_type = double;
  HostMatrix() = default;
  HostMatrix(int64_t m, int64_t n) : m_(m), n_(n), d_(size_t(m) * size_t(n), 0.) {}
  int64_t rows() const { return m_; }
  int64_t cols() const { return n_; }
  double* data() { return d_.data(); }
  const double* data() const { return d_.data(); }
  double& operator()(int64_t i, int64_t j) { return d_[size_t(i) + size_t(j) * m_]; }
  double operator()(int64_t i, int64_t j) const { return d_[size_t(i) + size_t(j) * m_]; }
};

namespace detail {

/// Non-template device integrator over raw column-major pointers.
class DeviceXCIntegratorImpl {
  struct State;
  std::unique_ptr<State> s_;
  // The EXC gradient after the spin-specific checks (Pz null for RKS).
  void exc_grad_impl(int64_t m, const double* P, int64_t ldp, const double* Pz, int64_t ldpz,
                     double* grad, const IntegratorSettingsXC& s);

 public:
  /// The device, the fill fraction or the caller's device arena come from
  /// the load balancer's runtime environment (lb->runtime()).
  DeviceXCIntegratorImpl(std::shared_ptr<functional_type> func,
                         std::shared_ptr<LoadBalancer> lb, DevicePrecision prec);
  ~DeviceXCIntegratorImpl();
  void eval_exc_vxc(int64_t m, int64_t n, const double* P, int64_t ldp, double* VXC,
                    int64_t ldvxc, double* EXC, const IntegratorSettingsXC& s);
  void eval_exc(int64_t m, int64_t n, const double* P, int64_t ldp, double* EXC,
                const IntegratorSettingsXC& s);
  /// UKS (P_s = P_a + P_b, P_z = P_a - P_b), FP64; a polarized functional.
  void eval_exc_vxc_uks(int64_t m, int64_t n, const double* Ps, int64_t ldps, const double* Pz,
                        int64_t ldpz, double* VXCs, int64_t ldvxcs, double* VXCz,
                        int64_t ldvxcz, double* EXC, const IntegratorSettingsXC& s);
  void eval_exc_uks(int64_t m, int64_t n, const double* Ps, int64_t ldps, const double* Pz,
                    int64_t ldpz, double* EXC, const IntegratorSettingsXC& s);
  /// RKS EXC nuclear gradient, grad[3 * natoms] atom-major (FP64; unpolarized
  /// functional); settings of type IntegratorSettingsEXC_GRAD select the form.
  void eval_exc_grad(int64_t m, int64_t n, const double* P, int64_t ldp, double* grad,
                     const IntegratorSettingsXC& s);
  /// UKS EXC nuclear gradient (FP64; polarized functional).
  void eval_exc_grad_uks(int64_t m, int64_t n, const double* Ps, int64_t ldps, const double* Pz,
                         int64_t ldpz, double* grad, const IntegratorSettingsXC& s);
  /// RKS FXC contraction d VXC[P + e tP] / de (FP64; unpolarized functional).
  void eval_fxc_contraction(int64_t m, int64_t n, const double* P, int64_t ldp, const double* tP,
                            int64_t ldtp, double* FXC, int64_t ldfxc,
                            const IntegratorSettingsXC& s);
  /// UKS FXC contraction (FP64; polarized functional): FXC_s, FXC_z.
  void eval_fxc_contraction_uks(int64_t m, int64_t n, const double* Ps, int64_t ldps,
                                const double* Pz, int64_t ldpz, const double* tPs,
                                int64_t ldtps, const double* tPz, int64_t ldtpz, double* FXCs,
                                int64_t ldfxcs, double* FXCz, int64_t ldfxcz,
                                const IntegratorSettingsXC& s);
  void integrate_den(int64_t m, int64_t n, const double* P, int64_t ldp, double* N_EL);
  const util::Timer& get_timings() const;
  const LoadBalancer& load_balancer() const;
  LoadBalancer& load_balancer();
  DevicePrecision precision() const;
  double last_nel() const;
  /// Underlying device context (xcq_context of include/gauxc_amd/device.h).
  void* device_context();
};

}  // namespace detail

template <typename MatrixType>
class XCIntegrator {
  std::shared_ptr<detail::DeviceXCIntegratorImpl> pimpl_;

  detail::DeviceXCIntegratorImpl& impl() const {
    if (!pimpl_) GAUXC_PIMPL_NOT_INITIALIZED();
    return *pimpl_;
  }

 public:
  using matrix_type = MatrixType;
  using value_type = typename MatrixType::value_type;
  using exc_vxc_type_rks = std::tuple<value_type, matrix_type>;
  using exc_vxc_type_uks = std::tuple<value_type, matrix_type, matrix_type>;

  XCIntegrator() = default;
  explicit XCIntegrator(std::shared_ptr<detail::DeviceXCIntegratorImpl> p) : pimpl_(std::move(p)) {}

  value_type integrate_den(const MatrixType& P) {
    value_type n = 0.;
    impl().integrate_den(P.rows(), P.cols(), P.data(), P.rows(), &n);
    return n;
  }
  value_type eval_exc(const MatrixType& P, const IntegratorSettingsXC& s = IntegratorSettingsXC{}) {
    value_type e = 0.;
    impl().eval_exc(P.rows(), P.cols(), P.data(), P.rows(), &e, s);
    return e;
  }
  exc_vxc_type_rks eval_exc_vxc(const MatrixType& P,
                                const IntegratorSettingsXC& s = IntegratorSettingsXC{}) {
    matrix_type VXC(P.rows(), P.cols());
    value_type e = 0.;
    impl().eval_exc_vxc(P.rows(), P.cols(), P.data(), P.rows(), VXC.data(), VXC.rows(), &e, s);
    return std::make_tuple(e, std::move(VXC));
  }
  value_type eval_exc(const MatrixType& Ps, const MatrixType& Pz,
                      const IntegratorSettingsXC& s = IntegratorSettingsXC{}) {
    value_type e = 0.;
    impl().eval_exc_uks(Ps.rows(), Ps.cols(), Ps.data(), Ps.rows(), Pz.data(), Pz.rows(), &e, s);
    return e;
  }
  exc_vxc_type_uks eval_exc_vxc(const MatrixType& Ps, const MatrixType& Pz,
                                const IntegratorSettingsXC& s = IntegratorSettingsXC{}) {
    matrix_type VXCs(Ps.rows(), Ps.cols()), VXCz(Ps.rows(), Ps.cols());
    value_type e = 0.;
    impl().eval_exc_vxc_uks(Ps.rows(), Ps.cols(), Ps.data(), Ps.rows(), Pz.data(), Pz.rows(),
                            VXCs.data(), VXCs.rows(), VXCz.data(), VXCz.rows(), &e, s);
    return std::make_tuple(e, std::move(VXCs), std::move(VXCz));
  }
  using exc_grad_type = std::vector<value_type>;
  /// GauXC XCIntegrator::eval_exc_grad(P, settings)
  /// (include/gauxc/xc_integrator/impl.hpp:82-86), RKS.
  exc_grad_type eval_exc_grad(const MatrixType& P,
                              const IntegratorSettingsXC& s = IntegratorSettingsEXC_GRAD{}) {
    exc_grad_type g(3 * impl().load_balancer().molecule().size());
    impl().eval_exc_grad(P.rows(), P.cols(), P.data(), P.rows(), g.data(), s);
    return g;
  }
  /// GauXC XCIntegrator::eval_exc_grad(Ps, Pz, settings)
  /// (include/gauxc/xc_integrator/impl.hpp:89-93), UKS.
  exc_grad_type eval_exc_grad(const MatrixType& Ps, const MatrixType& Pz,
                              const IntegratorSettingsXC& s = IntegratorSettingsEXC_GRAD{}) {
    exc_grad_type g(3 * impl().load_balancer().molecule().size());
    impl().eval_exc_grad_uks(Ps.rows(), Ps.cols(), Ps.data(), Ps.rows(), Pz.data(), Pz.rows(),
                             g.data(), s);
    return g;
  }
  /// GauXC XCIntegrator::eval_fxc_contraction(P, tP)
  /// (include/gauxc/xc_integrator/impl.hpp:104-109), RKS.
  matrix_type eval_fxc_contraction(const MatrixType& P, const MatrixType& tP,
                                   const IntegratorSettingsXC& s = IntegratorSettingsXC{}) {
    matrix_type FXC(P.rows(), P.cols());
    impl().eval_fxc_contraction(P.rows(), P.cols(), P.data(), P.rows(), tP.data(), tP.rows(),
                                FXC.data(), FXC.rows(), s);
    return FXC;
  }
  using fxc_contraction_type_uks = std::tuple<matrix_type, matrix_type>;
  /// GauXC XCIntegrator::eval_fxc_contraction(Ps, Pz, tPs, tPz)
  /// (include/gauxc/xc_integrator/impl.hpp:112-117), UKS.
  fxc_contraction_type_uks eval_fxc_contraction(const MatrixType& Ps, const MatrixType& Pz,
                                                const MatrixType& tPs, const MatrixType& tPz,
                                                const IntegratorSettingsXC& s =
                                                    IntegratorSettingsXC{}) {
    matrix_type Fs(Ps.rows(), Ps.cols()), Fz(Ps.rows(), Ps.cols());
    impl().eval_fxc_contraction_uks(Ps.rows(), Ps.cols(), Ps.data(), Ps.rows(), Pz.data(),
                                    Pz.rows(), tPs.data(), tPs.rows(), tPz.data(), tPz.rows(),
                                    Fs.data(), Fs.rows(), Fz.data(), Fz.rows(), s);
    return std::make_tuple(std::move(Fs), std::move(Fz));
  }
  const util::Timer& get_timings() const { return impl().get_timings(); }
  const LoadBalancer& load_balancer() const { return impl().load_balancer(); }
  detail::DeviceXCIntegratorImpl& device_impl() { return impl(); }
};

/// Validates the factory strings and builds the device implementation.
std::shared_ptr<detail::DeviceXCIntegratorImpl> make_device_integrator_impl(
    ExecutionSpace ex, const std::string& input_type, const std::string& integrator_kernel,
    const std::string& lwd_kernel, const std::string& reduction_kernel,
    std::shared_ptr<functional_type> func, std::shared_ptr<LoadBalancer> lb);

template <typename MatrixType>
class XCIntegratorFactory {
  ExecutionSpace ex_;
  std::string input_type_, integrator_kernel_, lwd_kernel_, reduction_kernel_;

 public:
  XCIntegratorFactory(ExecutionSpace ex, std::string input_type,
                      std::string integrator_kernel = "Default",
                      std::string lwd_kernel = "Default",
                      std::string reduction_kernel = "Default",
                      LocalWorkSettings = LocalWorkSettings{})
      : ex_(ex), input_type_(std::move(input_type)),
        integrator_kernel_(std::move(integrator_kernel)), lwd_kernel_(std::move(lwd_kernel)),
        reduction_kernel_(std::move(reduction_kernel)) {}

  XCIntegrator<MatrixType> get_instance(std::shared_ptr<functional_type> func,
                                        std::shared_ptr<LoadBalancer> lb) {
    return XCIntegrator<MatrixType>(make_device_integrator_impl(
        ex_, input_type_, integrator_kernel_, lwd_kernel_, reduction_kernel_, func, lb));
  }
  XCIntegrator<MatrixType> get_instance(const functional_type& func, const LoadBalancer& lb) {
    // GauXC copies the functional and shares the caller's load balancer
    // state through a copy; weights/tasks travel with it.
    return get_instance(std::make_shared<functional_type>(func),
                        std::make_shared<LoadBalancer>(lb));
  }
  std::shared_ptr<XCIntegrator<MatrixType>> get_shared_instance(
      std::shared_ptr<functional_type> func, std::shared_ptr<LoadBalancer> lb) {
    return std::make_shared<XCIntegrator<MatrixType>>(get_instance(func, lb));
  }
};

}  // namespace GauXC
