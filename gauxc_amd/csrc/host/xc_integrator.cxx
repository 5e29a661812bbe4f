// Device XC integrator: validation, task upload and the calls into the
// device C ABI.  GauXC behaviour followed:
//  * argument checks of IncoreReplicatedXCDeviceIntegrator::eval_exc_vxc_
//    (incore_replicated_xc_device_integrator_exc_vxc.hpp:66-92) and the
//    "Weights Have Not Been Modified" check (:261-264);
//  * tasks processed in descending npts*nbe order (:254-257), without
//    reordering the caller's load balancer;
//  * fac = 2 for RKS: P is the alpha density (:315).
#include <gauxc/xc_integrator.hpp>
#include <gauxc/exceptions.hpp>
#include <gauxc_amd/device.h>

#include <algorithm>
#include <cctype>
#include <chrono>
#include <numeric>
#include <string>

namespace GauXC {

functional_type functional_from_string(const std::string& name, bool polarized) {
  std::string u = name;
  std::transform(u.begin(), u.end(), u.begin(), [](unsigned char c) { return std::toupper(c); });
  ExchCXX::Functional f;
  if (u == "SVWN5") f = ExchCXX::Functional::SVWN5;
  else if (u == "PBE") f = ExchCXX::Functional::PBE;
  else if (u == "PBE0") f = ExchCXX::Functional::PBE0;
  else GAUXC_GENERIC_EXCEPTION("Functional not provided by the device backend: " + name);
  return functional_type(ExchCXX::Backend::builtin, f,
                         polarized ? ExchCXX::Spin::Polarized : ExchCXX::Spin::Unpolarized);
}

namespace detail {

namespace {
void check(int rc, const char* what) {
  if (rc != XCQ_SUCCESS)
    GAUXC_GENERIC_EXCEPTION(std::string(what) + " failed: " + xcq_last_error());
}
int xcq_func_id(const functional_type& f) {
  switch (f.kind()) {
    case ExchCXX::Functional::SVWN5: return XCQ_FUNC_SVWN5;
    case ExchCXX::Functional::PBE: return XCQ_FUNC_PBE;
    case ExchCXX::Functional::PBE0: return XCQ_FUNC_PBE0;
  }
  GAUXC_GENERIC_EXCEPTION("unknown functional");
}
}  // namespace

struct DeviceXCIntegratorImpl::State {
  std::shared_ptr<functional_type> func;
  std::shared_ptr<LoadBalancer> lb;
  DevicePrecision prec;
  int device = 0;
  double fill_fraction = 0.9;
  void* user_mem = nullptr;  // caller-owned device arena (DeviceRuntimeEnvironment(mem, sz))
  size_t user_mem_size = 0;
  xcq_context ctx = nullptr;
  uint64_t uploaded_revision = ~uint64_t(0);
  bool basis_uploaded = false;
  util::Timer timer;
  double last_nel = 0.;

  ~State() {
    if (ctx) xcq_context_destroy(ctx);
  }

  void ensure_context() {
    if (ctx) return;
    if (user_mem)
      check(xcq_context_create_mem(device, user_mem, user_mem_size, &ctx),
            "xcq_context_create_mem");
    else
      check(xcq_context_create(device, 0, fill_fraction, &ctx), "xcq_context_create");
    // SCF calls repeat the same evaluation with a new P: capture it once as a
    // graph and replay it.
    check(xcq_use_graph(ctx, 1), "xcq_use_graph");
  }

  void ensure_static_data() {
    ensure_context();
    const auto& basis = lb->basis();
    if (!basis_uploaded) {
      std::vector<xcq_shell> sh(basis.size());
      for (size_t i = 0; i < basis.size(); ++i) {
        const auto& s = basis[i];
        auto& d = sh[i];
        d = xcq_shell{};
        d.nprim = s.nprim();
        d.l = s.l();
        d.pure = s.pure();
        for (int k = 0; k < s.nprim(); ++k) {
          d.alpha[k] = s.alpha()[k];
          d.coeff[k] = s.coeff()[k];
        }
        for (int k = 0; k < 3; ++k) d.origin[k] = s.O()[k];
        d.cutoff_radius = s.cutoff_radius();
      }
      check(xcq_upload_basis(ctx, static_cast<int32_t>(sh.size()), sh.data()),
            "xcq_upload_basis");
      basis_uploaded = true;
    }
    if (uploaded_revision == lb->revision()) return;
    const auto& tasks = lb->get_tasks();
    std::vector<size_t> order(tasks.size());
    std::iota(order.begin(), order.end(), size_t{0});
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      return tasks[a].points.size() * size_t(tasks[a].bfn_screening.nbe) >
             tasks[b].points.size() * size_t(tasks[b].bfn_screening.nbe);
    });
    std::vector<xcq_task> desc(tasks.size());
    size_t npts = 0, nsh = 0;
    for (const auto& t : tasks) {
      npts += t.points.size();
      nsh += t.bfn_screening.shell_list.size();
    }
    std::vector<double> pts(3 * npts), w(npts);
    std::vector<int32_t> sl(nsh);
    size_t po = 0, so = 0;
    for (size_t k = 0; k < order.size(); ++k) {
      const auto& t = tasks[order[k]];
      auto& d = desc[k];
      d.iparent = t.iParent;
      d.npts = static_cast<int32_t>(t.points.size());
      d.nshells = static_cast<int32_t>(t.bfn_screening.shell_list.size());
      d.nbe = t.bfn_screening.nbe;
      d.point_offset = static_cast<int64_t>(po);
      d.shell_offset = static_cast<int64_t>(so);
      d.dist_nearest = t.dist_nearest;
      for (size_t i = 0; i < t.points.size(); ++i, ++po) {
        pts[3 * po] = t.points[i][0];
        pts[3 * po + 1] = t.points[i][1];
        pts[3 * po + 2] = t.points[i][2];
        w[po] = t.weights[i];
      }
      for (int32_t s : t.bfn_screening.shell_list) sl[so++] = s;
    }
    check(xcq_upload_tasks(ctx, static_cast<int32_t>(desc.size()), desc.data(),
                           static_cast<int64_t>(npts), pts.data(), w.data(),
                           static_cast<int64_t>(nsh), sl.data()),
          "xcq_upload_tasks");
    uploaded_revision = lb->revision();
  }

  void validate(int64_t m, int64_t n, int64_t ldp, int64_t ldvxc, bool have_vxc) {
    const int64_t nbf = lb->basis().nbf();
    if (m != n) GAUXC_GENERIC_EXCEPTION("P Must Be Square");
    if (m != nbf) GAUXC_GENERIC_EXCEPTION("P Must Have Same Dimension as Basis");
    if (ldp < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDP");
    if (have_vxc && ldvxc < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDVXC");
    if (!lb->state().modified_weights_are_stored)
      GAUXC_GENERIC_EXCEPTION("Weights Have Not Been Modified");
  }
  void validate_rks(int64_t m, int64_t n, int64_t ldp, int64_t ldvxc, bool have_vxc) {
    validate(m, n, ldp, ldvxc, have_vxc);
    if (func->is_polarized())
      GAUXC_GENERIC_EXCEPTION("RKS evaluation needs an unpolarized functional");
  }
  void validate_uks(int64_t m, int64_t n, int64_t ldps, int64_t ldpz, int64_t ldvs, int64_t ldvz,
                    bool have_vxc) {
    validate(m, n, ldps, ldvs, have_vxc);
    if (ldpz < lb->basis().nbf()) GAUXC_GENERIC_EXCEPTION("Invalid LDPZ");
    if (have_vxc && ldvz < lb->basis().nbf()) GAUXC_GENERIC_EXCEPTION("Invalid LDVXCZ");
    if (!func->is_polarized())
      GAUXC_GENERIC_EXCEPTION("UKS evaluation needs a polarized functional");
    if (prec != DevicePrecision::FP64)
      GAUXC_GENERIC_EXCEPTION("UKS evaluation is provided in FP64 (local work driver Default)");
  }
};

DeviceXCIntegratorImpl::DeviceXCIntegratorImpl(std::shared_ptr<functional_type> func,
                                               std::shared_ptr<LoadBalancer> lb,
                                               DevicePrecision prec)
    : s_(std::make_unique<State>()) {
  if (!func || !lb) GAUXC_GENERIC_EXCEPTION("null functional or load balancer");
  s_->func = std::move(func);
  s_->lb = std::move(lb);
  s_->prec = prec;
  // GauXC sizes its device arena from the DeviceRuntimeEnvironment the load
  // balancer holds: fill_fraction of the free memory, or the caller's buffer
  // (device_runtime_environment_impl.hpp:42-50, xc_device_stack_data.cxx:713-751).
  if (const auto* drt = as_device_runtime_ptr(s_->lb->runtime())) {
    s_->device = drt->device_id();
    if (drt->owns_memory()) {
      s_->fill_fraction = drt->fill_fraction();
    } else {
      s_->user_mem = drt->device_memory();
      s_->user_mem_size = drt->device_memory_size();
    }
  }
}

DeviceXCIntegratorImpl::~DeviceXCIntegratorImpl() = default;

void DeviceXCIntegratorImpl::eval_exc_vxc(int64_t m, int64_t n, const double* P, int64_t ldp,
                                          double* VXC, int64_t ldvxc, double* EXC,
                                          const IntegratorSettingsXC&) {
  auto& s = *s_;
  s.validate_rks(m, n, ldp, ldvxc, true);
  if (!P || !VXC || !EXC) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  const auto t0 = std::chrono::steady_clock::now();
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), P, ldp), "xcq_upload_density");
  const auto t1 = std::chrono::steady_clock::now();
  check(xcq_exc_vxc(s.ctx, xcq_func_id(*s.func), static_cast<int>(s.prec), 1), "xcq_exc_vxc");
  check(xcq_synchronize(s.ctx), "xcq_synchronize");
  const auto t2 = std::chrono::steady_clock::now();
  check(xcq_download_results(s.ctx, EXC, &s.last_nel, VXC, ldvxc), "xcq_download_results");
  const auto t3 = std::chrono::steady_clock::now();
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  s.timer.set("XCIntegrator.HostToDevice_EXC_VXC", ms(t0, t1));
  s.timer.set("XCIntegrator.LocalWork_EXC_VXC", ms(t1, t2));
  s.timer.set("XCIntegrator.DeviceToHostCopy_EXC_VXC", ms(t2, t3));
}

void DeviceXCIntegratorImpl::eval_exc(int64_t m, int64_t n, const double* P, int64_t ldp,
                                      double* EXC, const IntegratorSettingsXC&) {
  auto& s = *s_;
  s.validate_rks(m, n, ldp, 0, false);
  if (!P || !EXC) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), P, ldp), "xcq_upload_density");
  check(xcq_exc_vxc(s.ctx, xcq_func_id(*s.func), static_cast<int>(s.prec), 0), "xcq_exc_vxc");
  check(xcq_download_results(s.ctx, EXC, &s.last_nel, nullptr, 0), "xcq_download_results");
}

void DeviceXCIntegratorImpl::eval_exc_vxc_uks(int64_t m, int64_t n, const double* Ps,
                                              int64_t ldps, const double* Pz, int64_t ldpz,
                                              double* VXCs, int64_t ldvxcs, double* VXCz,
                                              int64_t ldvxcz, double* EXC,
                                              const IntegratorSettingsXC&) {
  auto& s = *s_;
  s.validate_uks(m, n, ldps, ldpz, ldvxcs, ldvxcz, true);
  if (!Ps || !Pz || !VXCs || !VXCz || !EXC) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  if (ldvxcs != ldvxcz) {
    // The device download takes one leading dimension for both potentials.
    std::vector<double> vz(size_t(m) * size_t(m));
    eval_exc_vxc_uks(m, n, Ps, ldps, Pz, ldpz, VXCs, ldvxcs, vz.data(), m, EXC, {});
    for (int64_t j = 0; j < m; ++j)
      for (int64_t i = 0; i < m; ++i) VXCz[i + j * ldvxcz] = vz[size_t(i) + size_t(j) * m];
    return;
  }
  const auto t0 = std::chrono::steady_clock::now();
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), Ps, ldps), "xcq_upload_density");
  check(xcq_upload_density_z(s.ctx, static_cast<int32_t>(m), Pz, ldpz), "xcq_upload_density_z");
  const auto t1 = std::chrono::steady_clock::now();
  check(xcq_exc_vxc_uks(s.ctx, xcq_func_id(*s.func), 1), "xcq_exc_vxc_uks");
  check(xcq_synchronize(s.ctx), "xcq_synchronize");
  const auto t2 = std::chrono::steady_clock::now();
  check(xcq_download_results_uks(s.ctx, EXC, &s.last_nel, VXCs, VXCz, ldvxcs),
        "xcq_download_results_uks");
  const auto t3 = std::chrono::steady_clock::now();
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  s.timer.set("XCIntegrator.HostToDevice_EXC_VXC", ms(t0, t1));
  s.timer.set("XCIntegrator.LocalWork_EXC_VXC", ms(t1, t2));
  s.timer.set("XCIntegrator.DeviceToHostCopy_EXC_VXC", ms(t2, t3));
}

void DeviceXCIntegratorImpl::eval_exc_uks(int64_t m, int64_t n, const double* Ps, int64_t ldps,
                                          const double* Pz, int64_t ldpz, double* EXC,
                                          const IntegratorSettingsXC&) {
  auto& s = *s_;
  s.validate_uks(m, n, ldps, ldpz, 0, 0, false);
  if (!Ps || !Pz || !EXC) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), Ps, ldps), "xcq_upload_density");
  check(xcq_upload_density_z(s.ctx, static_cast<int32_t>(m), Pz, ldpz), "xcq_upload_density_z");
  check(xcq_exc_vxc_uks(s.ctx, xcq_func_id(*s.func), 0), "xcq_exc_vxc_uks");
  check(xcq_download_results_uks(s.ctx, EXC, &s.last_nel, nullptr, nullptr, m),
        "xcq_download_results_uks");
}

void DeviceXCIntegratorImpl::eval_exc_grad(int64_t m, int64_t n, const double* P, int64_t ldp,
                                           double* grad, const IntegratorSettingsXC& settings) {
  // Argument checks of reference_replicated_xc_host_integrator_exc_grad.hpp:29-37.
  auto& s = *s_;
  const int64_t nbf = s.lb->basis().nbf();
  if (m != n) GAUXC_GENERIC_EXCEPTION("P Must Be Square");
  if (m != nbf) GAUXC_GENERIC_EXCEPTION("P Must Have Same Dimension as Basis");
  if (ldp < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDP");
  if (s.func->is_polarized())
    GAUXC_GENERIC_EXCEPTION("RKS evaluation needs an unpolarized functional");
  exc_grad_impl(m, P, ldp, nullptr, 0, grad, settings);
}

void DeviceXCIntegratorImpl::eval_exc_grad_uks(int64_t m, int64_t n, const double* Ps,
                                               int64_t ldps, const double* Pz, int64_t ldpz,
                                               double* grad,
                                               const IntegratorSettingsXC& settings) {
  // Argument checks of reference_replicated_xc_host_integrator_exc_grad.hpp:70-79.
  auto& s = *s_;
  const int64_t nbf = s.lb->basis().nbf();
  if (m != n) GAUXC_GENERIC_EXCEPTION("P Must Be Square");
  if (m != nbf) GAUXC_GENERIC_EXCEPTION("P Must Have Same Dimension as Basis");
  if (ldps < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDPS");
  if (ldpz < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDPZ");
  if (!s.func->is_polarized())
    GAUXC_GENERIC_EXCEPTION("UKS evaluation needs a polarized functional");
  if (!Pz) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  exc_grad_impl(m, Ps, ldps, Pz, ldpz, grad, settings);
}

void DeviceXCIntegratorImpl::exc_grad_impl(int64_t m, const double* P, int64_t ldp,
                                           const double* Pz, int64_t ldpz, double* grad,
                                           const IntegratorSettingsXC& settings) {
  auto& s = *s_;
  if (!s.lb->state().modified_weights_are_stored)
    GAUXC_GENERIC_EXCEPTION("Weights Have Not Been Modified");
  if (s.prec != DevicePrecision::FP64)
    GAUXC_GENERIC_EXCEPTION("The EXC gradient is provided in FP64 (local work driver Default)");
  if (!P || !grad) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  bool weight_derivs = true;
  if (const auto* gs = dynamic_cast<const IntegratorSettingsEXC_GRAD*>(&settings))
    weight_derivs = gs->include_weight_derivatives;
  const Molecule& mol = s.lb->molecule();
  std::vector<double> xyz(3 * mol.size());
  for (size_t a = 0; a < mol.size(); ++a) {
    xyz[3 * a] = mol[a].x;
    xyz[3 * a + 1] = mol[a].y;
    xyz[3 * a + 2] = mol[a].z;
  }
  const auto t0 = std::chrono::steady_clock::now();
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), P, ldp), "xcq_upload_density");
  if (Pz)
    check(xcq_upload_density_z(s.ctx, static_cast<int32_t>(m), Pz, ldpz), "xcq_upload_density_z");
  const auto t1 = std::chrono::steady_clock::now();
  if (Pz)
    check(xcq_exc_grad_uks(s.ctx, xcq_func_id(*s.func), static_cast<int32_t>(mol.size()),
                           xyz.data(), weight_derivs ? 1 : 0),
          "xcq_exc_grad_uks");
  else
    check(xcq_exc_grad(s.ctx, xcq_func_id(*s.func), static_cast<int32_t>(mol.size()),
                       xyz.data(), weight_derivs ? 1 : 0),
          "xcq_exc_grad");
  check(xcq_synchronize(s.ctx), "xcq_synchronize");
  const auto t2 = std::chrono::steady_clock::now();
  double exc = 0.;
  check(xcq_download_exc_grad(s.ctx, &exc, &s.last_nel, grad), "xcq_download_exc_grad");
  const auto t3 = std::chrono::steady_clock::now();
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  s.timer.set("XCIntegrator.HostToDevice_EXC_GRAD", ms(t0, t1));
  s.timer.set("XCIntegrator.LocalWork_EXC_GRAD", ms(t1, t2));
  s.timer.set("XCIntegrator.DeviceToHostCopy_EXC_GRAD", ms(t2, t3));
}

void DeviceXCIntegratorImpl::eval_fxc_contraction(int64_t m, int64_t n, const double* P,
                                                  int64_t ldp, const double* tP, int64_t ldtp,
                                                  double* FXC, int64_t ldfxc,
                                                  const IntegratorSettingsXC&) {
  // Argument checks of reference_replicated_xc_host_integrator_fxc_contraction.hpp:39-57.
  auto& s = *s_;
  const int64_t nbf = s.lb->basis().nbf();
  if (m != n) GAUXC_GENERIC_EXCEPTION("P/FXC Must Be Square");
  if (m != nbf) GAUXC_GENERIC_EXCEPTION("P/FXC Must Have Same Dimension as Basis");
  if (ldp < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDPS");
  if (ldtp < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDTPS");
  if (ldfxc < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDFXCS");
  if (!s.lb->state().modified_weights_are_stored)
    GAUXC_GENERIC_EXCEPTION("Weights Have Not Been Modified");
  if (s.func->is_polarized())
    GAUXC_GENERIC_EXCEPTION("RKS evaluation needs an unpolarized functional");
  if (s.prec != DevicePrecision::FP64)
    GAUXC_GENERIC_EXCEPTION(
        "The FXC contraction is provided in FP64 (local work driver Default)");
  if (!P || !tP || !FXC) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  const auto t0 = std::chrono::steady_clock::now();
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), P, ldp), "xcq_upload_density");
  check(xcq_upload_trial_density(s.ctx, static_cast<int32_t>(m), tP, ldtp),
        "xcq_upload_trial_density");
  const auto t1 = std::chrono::steady_clock::now();
  check(xcq_fxc_contraction(s.ctx, xcq_func_id(*s.func)), "xcq_fxc_contraction");
  check(xcq_synchronize(s.ctx), "xcq_synchronize");
  const auto t2 = std::chrono::steady_clock::now();
  check(xcq_download_fxc(s.ctx, &s.last_nel, FXC, ldfxc), "xcq_download_fxc");
  const auto t3 = std::chrono::steady_clock::now();
  auto ms = [](auto a, auto b) { return std::chrono::duration<double, std::milli>(b - a).count(); };
  s.timer.set("XCIntegrator.HostToDevice_FXC", ms(t0, t1));
  s.timer.set("XCIntegrator.LocalWork_FXC", ms(t1, t2));
  s.timer.set("XCIntegrator.DeviceToHostCopy_FXC", ms(t2, t3));
}

void DeviceXCIntegratorImpl::eval_fxc_contraction_uks(
    int64_t m, int64_t n, const double* Ps, int64_t ldps, const double* Pz, int64_t ldpz,
    const double* tPs, int64_t ldtps, const double* tPz, int64_t ldtpz, double* FXCs,
    int64_t ldfxcs, double* FXCz, int64_t ldfxcz, const IntegratorSettingsXC&) {
  // Argument checks of reference_replicated_xc_host_integrator_fxc_contraction.hpp:39-57.
  auto& s = *s_;
  const int64_t nbf = s.lb->basis().nbf();
  if (m != n) GAUXC_GENERIC_EXCEPTION("P/FXC Must Be Square");
  if (m != nbf) GAUXC_GENERIC_EXCEPTION("P/FXC Must Have Same Dimension as Basis");
  if (ldps < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDPS");
  if (ldpz < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDPZ");
  if (ldtps < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDTPS");
  if (ldtpz < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDTZP");
  if (ldfxcs < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDFXCS");
  if (ldfxcz < nbf) GAUXC_GENERIC_EXCEPTION("Invalid LDFXCZ");
  if (!s.lb->state().modified_weights_are_stored)
    GAUXC_GENERIC_EXCEPTION("Weights Have Not Been Modified");
  if (!s.func->is_polarized())
    GAUXC_GENERIC_EXCEPTION("UKS evaluation needs a polarized functional");
  if (s.prec != DevicePrecision::FP64)
    GAUXC_GENERIC_EXCEPTION(
        "The FXC contraction is provided in FP64 (local work driver Default)");
  if (!Ps || !Pz || !tPs || !tPz || !FXCs || !FXCz)
    GAUXC_GENERIC_EXCEPTION("null pointer argument");
  s.ensure_static_data();
  const auto m32 = static_cast<int32_t>(m);
  check(xcq_upload_density(s.ctx, m32, Ps, ldps), "xcq_upload_density");
  check(xcq_upload_density_z(s.ctx, m32, Pz, ldpz), "xcq_upload_density_z");
  check(xcq_upload_trial_density(s.ctx, m32, tPs, ldtps), "xcq_upload_trial_density");
  check(xcq_upload_trial_density_z(s.ctx, m32, tPz, ldtpz), "xcq_upload_trial_density_z");
  check(xcq_fxc_contraction_uks(s.ctx, xcq_func_id(*s.func)), "xcq_fxc_contraction_uks");
  if (ldfxcs == ldfxcz) {
    check(xcq_download_fxc_uks(s.ctx, &s.last_nel, FXCs, FXCz, ldfxcs), "xcq_download_fxc_uks");
    return;
  }
  check(xcq_download_fxc_uks(s.ctx, &s.last_nel, FXCs, nullptr, ldfxcs), "xcq_download_fxc_uks");
  check(xcq_download_fxc_uks(s.ctx, nullptr, nullptr, FXCz, ldfxcz), "xcq_download_fxc_uks");
}

void DeviceXCIntegratorImpl::integrate_den(int64_t m, int64_t n, const double* P, int64_t ldp,
                                           double* N_EL) {
  auto& s = *s_;
  s.validate(m, n, ldp, 0, false);
  if (!P || !N_EL) GAUXC_GENERIC_EXCEPTION("null pointer argument");
  s.ensure_static_data();
  check(xcq_upload_density(s.ctx, static_cast<int32_t>(m), P, ldp), "xcq_upload_density");
  check(xcq_integrate_den(s.ctx, N_EL), "xcq_integrate_den");
  s.last_nel = *N_EL;
}

const util::Timer& DeviceXCIntegratorImpl::get_timings() const { return s_->timer; }
const LoadBalancer& DeviceXCIntegratorImpl::load_balancer() const { return *s_->lb; }
LoadBalancer& DeviceXCIntegratorImpl::load_balancer() { return *s_->lb; }
DevicePrecision DeviceXCIntegratorImpl::precision() const { return s_->prec; }
double DeviceXCIntegratorImpl::last_nel() const { return s_->last_nel; }
void* DeviceXCIntegratorImpl::device_context() {
  s_->ensure_static_data();
  return s_->ctx;
}

}  // namespace detail

std::shared_ptr<detail::DeviceXCIntegratorImpl> make_device_integrator_impl(
    ExecutionSpace ex, const std::string& input_type, const std::string& integrator_kernel,
    const std::string& lwd_kernel, const std::string& reduction_kernel,
    std::shared_ptr<functional_type> func, std::shared_ptr<LoadBalancer> lb) {
  auto up = [](std::string s) {
    std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::toupper(c); });
    return s;
  };
  if (ex != ExecutionSpace::Device)
    GAUXC_GENERIC_EXCEPTION("This library provides ExecutionSpace::Device only");
  if (up(input_type) != "REPLICATED")
    GAUXC_GENERIC_EXCEPTION("Integrator Input Type Not Recognized: " + input_type);
  const auto ik = up(integrator_kernel);
  if (ik != "DEFAULT" && ik != "INCORE")
    GAUXC_GENERIC_EXCEPTION("Integrator Kernel Not Recognized: " + integrator_kernel);
  const auto lk = up(lwd_kernel);
  DevicePrecision prec;
  if (lk == "DEFAULT" || lk == "SCHEME1") prec = DevicePrecision::FP64;
  else if (lk == "SCHEME1-MIXED") prec = DevicePrecision::Mixed;
  else GAUXC_GENERIC_EXCEPTION("Local Work Driver Not Recognized: " + lwd_kernel);
  const auto rk = up(reduction_kernel);
  if (rk != "DEFAULT" && rk != "BASICMPI" && rk != "NCCL")
    GAUXC_GENERIC_EXCEPTION("Reduction Kernel Not Recognized: " + reduction_kernel);
  if (!func || !func->initialized()) GAUXC_PIMPL_NOT_INITIALIZED();
  return std::make_shared<detail::DeviceXCIntegratorImpl>(std::move(func), std::move(lb), prec);
}

}  // namespace GauXC
