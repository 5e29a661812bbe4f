// GauXC-compatible C API over the C++ surface.  Exceptions are caught at
// the boundary and turned into GauXCStatus{1, message}
// (GauXC: src/c-api/c_xc_integrator.cxx:320-346).
#include <gauxc/c/gauxc.h>

#include <gauxc/basisset.hpp>
#include <gauxc/load_balancer.hpp>
#include <gauxc/molecular_weights.hpp>
#include <gauxc/molgrid.hpp>
#include <gauxc/runtime_environment.hpp>
#include <gauxc/xc_functional.hpp>
#include <gauxc/xc_integrator.hpp>

#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

using namespace GauXC;

namespace {

void set_status(GauXCStatus* st, int code, const char* msg) {
  if (!st) return;
  if (st->message) {
    std::free(st->message);
    st->message = nullptr;
  }
  st->code = code;
  if (msg) {
    st->message = static_cast<char*>(std::malloc(std::strlen(msg) + 1));
    if (st->message) std::strcpy(st->message, msg);
  }
}

template <typename F>
auto guarded(GauXCStatus* st, F&& f) -> decltype(f()) {
  try {
    auto r = f();
    set_status(st, 0, nullptr);
    return r;
  } catch (const std::exception& e) {
    set_status(st, 1, e.what());
  } catch (...) {
    set_status(st, 1, "unknown error");
  }
  return decltype(f()){};
}

template <typename F>
void guarded_void(GauXCStatus* st, F&& f) {
  try {
    f();
    set_status(st, 0, nullptr);
  } catch (const std::exception& e) {
    set_status(st, 1, e.what());
  } catch (...) {
    set_status(st, 1, "unknown error");
  }
}

template <typename H>
H make_handle(GauXC_Type t, void* p) {
  H h;
  h.hdr.type = t;
  h.ptr = p;
  return h;
}

template <typename T, typename H>
T& deref(const H& h, GauXC_Type t, const char* what) {
  if (h.hdr.type != t || !h.ptr) throw std::runtime_error(std::string("invalid ") + what + " handle");
  return *static_cast<T*>(h.ptr);
}

using IntegratorT = XCIntegrator<HostMatrix>;

struct LBHolder {
  std::shared_ptr<LoadBalancer> lb;
};

}  // namespace

extern "C" {

void gauxc_status_delete(GauXCStatus* st) {
  if (st && st->message) {
    std::free(st->message);
    st->message = nullptr;
  }
}

GauXCMolecule gauxc_molecule_new(GauXCStatus* st) {
  return guarded(st, [] { return make_handle<GauXCMolecule>(GauXC_Type_Molecule, new Molecule()); });
}

GauXCMolecule gauxc_molecule_new_from_atoms(GauXCStatus* st, const GauXCAtom* atoms, size_t n) {
  return guarded(st, [&] {
    if (!atoms && n) throw std::runtime_error("null atoms");
    auto* m = new Molecule();
    for (size_t i = 0; i < n; ++i)
      m->emplace_back(AtomicNumber(atoms[i].Z), atoms[i].x, atoms[i].y, atoms[i].z);
    return make_handle<GauXCMolecule>(GauXC_Type_Molecule, m);
  });
}

void gauxc_molecule_delete(GauXCStatus* st, GauXCMolecule* mol) {
  guarded_void(st, [&] {
    if (!mol) return;
    delete static_cast<Molecule*>(mol->ptr);
    mol->ptr = nullptr;
  });
}

size_t gauxc_molecule_natoms(GauXCStatus* st, const GauXCMolecule mol) {
  return guarded(st, [&] { return deref<Molecule>(mol, GauXC_Type_Molecule, "molecule").natoms(); });
}

GauXCBasisSet gauxc_basisset_new(GauXCStatus* st) {
  return guarded(st, [] {
    return make_handle<GauXCBasisSet>(GauXC_Type_BasisSet, new BasisSet<double>());
  });
}

GauXCBasisSet gauxc_basisset_new_from_shells(GauXCStatus* st, const GauXCShell* shells,
                                             size_t n, bool normalize) {
  return guarded(st, [&] {
    if (!shells && n) throw std::runtime_error("null shells");
    auto* b = new BasisSet<double>();
    for (size_t i = 0; i < n; ++i) {
      const auto& s = shells[i];
      if (s.nprim < 0 || s.nprim > 32) {
        delete b;
        throw std::runtime_error("nprim out of range");
      }
      Shell<double>::prim_array a{}, c{};
      for (int k = 0; k < s.nprim; ++k) {
        a[k] = s.exponents[k];
        c[k] = s.coefficients[k];
      }
      Shell<double> sh(PrimSize(s.nprim), AngularMomentum(s.l), SphericalType(s.pure ? 1 : 0), a,
                       c, {s.origin[0], s.origin[1], s.origin[2]}, normalize);
      if (s.shell_tolerance > 0. && s.shell_tolerance != sh.shell_tolerance())
        sh.set_shell_tolerance(s.shell_tolerance);
      b->push_back(sh);
    }
    return make_handle<GauXCBasisSet>(GauXC_Type_BasisSet, b);
  });
}

void gauxc_basisset_delete(GauXCStatus* st, GauXCBasisSet* b) {
  guarded_void(st, [&] {
    if (!b) return;
    delete static_cast<BasisSet<double>*>(b->ptr);
    b->ptr = nullptr;
  });
}

int32_t gauxc_basisset_nbf(GauXCStatus* st, const GauXCBasisSet b) {
  return guarded(st, [&] { return deref<BasisSet<double>>(b, GauXC_Type_BasisSet, "basis").nbf(); });
}

void gauxc_basisset_get_shells(GauXCStatus* st, const GauXCBasisSet b, GauXCShell* shells,
                               double* radii) {
  guarded_void(st, [&] {
    const auto& bs = deref<BasisSet<double>>(b, GauXC_Type_BasisSet, "basis");
    for (size_t i = 0; i < bs.size(); ++i) {
      const auto& s = bs[i];
      if (shells) {
        auto& d = shells[i];
        std::memset(&d, 0, sizeof(d));
        d.l = s.l();
        d.pure = s.pure() != 0;
        d.nprim = s.nprim();
        for (int k = 0; k < s.nprim(); ++k) {
          d.exponents[k] = s.alpha()[k];
          d.coefficients[k] = s.coeff()[k];
        }
        for (int k = 0; k < 3; ++k) d.origin[k] = s.O()[k];
        d.shell_tolerance = s.shell_tolerance();
      }
      if (radii) radii[i] = s.cutoff_radius();
    }
  });
}

GauXCMolGrid gauxc_molgrid_new_default(GauXCStatus* st, const GauXCMolecule mol,
                                       enum GauXC_PruningScheme ps, int64_t bsz,
                                       enum GauXC_RadialQuad rq,
                                       enum GauXC_AtomicGridSizeDefault gs) {
  return guarded(st, [&] {
    const auto& m = deref<Molecule>(mol, GauXC_Type_Molecule, "molecule");
    auto* g = new MolGrid(MolGridFactory::create_default_molgrid(
        m, static_cast<PruningScheme>(ps), BatchSize(bsz), static_cast<RadialQuad>(rq),
        static_cast<AtomicGridSizeDefault>(gs)));
    return make_handle<GauXCMolGrid>(GauXC_Type_MolGrid, g);
  });
}

void gauxc_molgrid_delete(GauXCStatus* st, GauXCMolGrid* g) {
  guarded_void(st, [&] {
    if (!g) return;
    delete static_cast<MolGrid*>(g->ptr);
    g->ptr = nullptr;
  });
}

GauXCRuntimeEnvironment gauxc_runtime_environment_new(GauXCStatus* st) {
  return guarded(st, [] {
    GauXCRuntimeEnvironment e{};
    e.hdr.type = GauXC_Type_RuntimeEnvironment;
    e.ptr = new RuntimeEnvironment();
    return e;
  });
}

GauXCRuntimeEnvironment gauxc_device_runtime_environment_new(GauXCStatus* st, double fill) {
  return guarded(st, [&] {
    if (!(fill > 0. && fill <= 1.)) throw std::runtime_error("fill_fraction must be in (0, 1]");
    GauXCRuntimeEnvironment e{};
    e.hdr.type = GauXC_Type_RuntimeEnvironment;
    e.device_ptr = new DeviceRuntimeEnvironment(fill);
    return e;
  });
}

GauXCRuntimeEnvironment gauxc_device_runtime_environment_new_mem(GauXCStatus* st, void* mem,
                                                                 size_t mem_sz) {
  return guarded(st, [&] {
    if (!mem && mem_sz) throw std::runtime_error("null device buffer with a non-zero size");
    GauXCRuntimeEnvironment e{};
    e.hdr.type = GauXC_Type_RuntimeEnvironment;
    e.device_ptr = new DeviceRuntimeEnvironment(mem, mem_sz);
    return e;
  });
}

void gauxc_runtime_environment_delete(GauXCStatus* st, GauXCRuntimeEnvironment* e) {
  guarded_void(st, [&] {
    if (!e) return;
    delete static_cast<RuntimeEnvironment*>(e->ptr);
    delete static_cast<DeviceRuntimeEnvironment*>(e->device_ptr);
    e->ptr = nullptr;
    e->device_ptr = nullptr;
  });
}

static const RuntimeEnvironment& runtime_of(const GauXCRuntimeEnvironment& e) {
  void* p = e.device_ptr ? e.device_ptr : e.ptr;
  if (e.hdr.type != GauXC_Type_RuntimeEnvironment || !p)
    throw std::runtime_error("Invalid RuntimeEnvironment handle");
  if (e.device_ptr) return *static_cast<const DeviceRuntimeEnvironment*>(e.device_ptr);
  return *static_cast<const RuntimeEnvironment*>(e.ptr);
}

int gauxc_runtime_environment_comm_rank(GauXCStatus* st, const GauXCRuntimeEnvironment e) {
  return guarded(st, [&] { return runtime_of(e).comm_rank(); });
}

int gauxc_runtime_environment_comm_size(GauXCStatus* st, const GauXCRuntimeEnvironment e) {
  return guarded(st, [&] { return runtime_of(e).comm_size(); });
}

GauXCLoadBalancerFactory gauxc_load_balancer_factory_new(GauXCStatus* st,
                                                         enum GauXC_ExecutionSpace ex,
                                                         const char* kernel) {
  return guarded(st, [&] {
    return make_handle<GauXCLoadBalancerFactory>(
        GauXC_Type_LoadBalancerFactory,
        new LoadBalancerFactory(static_cast<ExecutionSpace>(ex), kernel ? kernel : "Default"));
  });
}

void gauxc_load_balancer_factory_delete(GauXCStatus* st, GauXCLoadBalancerFactory* f) {
  guarded_void(st, [&] {
    if (!f) return;
    delete static_cast<LoadBalancerFactory*>(f->ptr);
    f->ptr = nullptr;
  });
}

GauXCLoadBalancer gauxc_load_balancer_factory_get_instance(GauXCStatus* st,
                                                           const GauXCLoadBalancerFactory f,
                                                           const GauXCRuntimeEnvironment env,
                                                           const GauXCMolecule mol,
                                                           const GauXCMolGrid mg,
                                                           const GauXCBasisSet basis) {
  return guarded(st, [&] {
    auto& fac = deref<LoadBalancerFactory>(f, GauXC_Type_LoadBalancerFactory, "factory");
    const auto& rt = runtime_of(env);
    const auto& m = deref<Molecule>(mol, GauXC_Type_Molecule, "molecule");
    const auto& g = deref<MolGrid>(mg, GauXC_Type_MolGrid, "molgrid");
    const auto& b = deref<BasisSet<double>>(basis, GauXC_Type_BasisSet, "basis");
    auto* h = new LBHolder{fac.get_shared_instance(rt, m, g, b)};
    return make_handle<GauXCLoadBalancer>(GauXC_Type_LoadBalancer, h);
  });
}

void gauxc_load_balancer_delete(GauXCStatus* st, GauXCLoadBalancer* lb) {
  guarded_void(st, [&] {
    if (!lb) return;
    delete static_cast<LBHolder*>(lb->ptr);
    lb->ptr = nullptr;
  });
}

void gauxc_load_balancer_task_counts(GauXCStatus* st, const GauXCLoadBalancer lb,
                                     int64_t* counts) {
  guarded_void(st, [&] {
    auto& h = deref<LBHolder>(lb, GauXC_Type_LoadBalancer, "load balancer");
    const auto& tasks = h.lb->get_tasks();
    int64_t np = 0, ns = 0;
    for (const auto& t : tasks) {
      np += static_cast<int64_t>(t.points.size());
      ns += static_cast<int64_t>(t.bfn_screening.shell_list.size());
    }
    counts[0] = static_cast<int64_t>(tasks.size());
    counts[1] = np;
    counts[2] = ns;
  });
}

void gauxc_load_balancer_export_tasks(GauXCStatus* st, const GauXCLoadBalancer lb,
                                      int32_t* iparent, int32_t* npts, int32_t* nshells,
                                      int32_t* nbe, double* dist_nearest, double* points,
                                      double* weights, int32_t* shell_lists) {
  guarded_void(st, [&] {
    auto& h = deref<LBHolder>(lb, GauXC_Type_LoadBalancer, "load balancer");
    const auto& tasks = h.lb->get_tasks();
    size_t po = 0, so = 0;
    for (size_t i = 0; i < tasks.size(); ++i) {
      const auto& t = tasks[i];
      if (iparent) iparent[i] = t.iParent;
      if (npts) npts[i] = static_cast<int32_t>(t.points.size());
      if (nshells) nshells[i] = static_cast<int32_t>(t.bfn_screening.shell_list.size());
      if (nbe) nbe[i] = t.bfn_screening.nbe;
      if (dist_nearest) dist_nearest[i] = t.dist_nearest;
      for (size_t k = 0; k < t.points.size(); ++k, ++po) {
        if (points)
          for (int d = 0; d < 3; ++d) points[3 * po + d] = t.points[k][d];
        if (weights) weights[po] = t.weights[k];
      }
      for (int32_t s : t.bfn_screening.shell_list) {
        if (shell_lists) shell_lists[so] = s;
        ++so;
      }
    }
  });
}

bool gauxc_load_balancer_weights_modified(GauXCStatus* st, const GauXCLoadBalancer lb) {
  return guarded(st, [&] {
    auto& h = deref<LBHolder>(lb, GauXC_Type_LoadBalancer, "load balancer");
    return h.lb->state().modified_weights_are_stored;
  });
}

GauXCMolecularWeightsFactory gauxc_molecular_weights_factory_new(
    GauXCStatus* st, enum GauXC_ExecutionSpace ex, const char* lwd,
    GauXCMolecularWeightsSettings s) {
  return guarded(st, [&] {
    MolecularWeightsSettings cs{static_cast<XCWeightAlg>(s.weight_alg), s.becke_size_adjustment};
    return make_handle<GauXCMolecularWeightsFactory>(
        GauXC_Type_MolecularWeightsFactory,
        new MolecularWeightsFactory(static_cast<ExecutionSpace>(ex), lwd ? lwd : "Default", cs));
  });
}

void gauxc_molecular_weights_factory_delete(GauXCStatus* st, GauXCMolecularWeightsFactory* f) {
  guarded_void(st, [&] {
    if (!f) return;
    delete static_cast<MolecularWeightsFactory*>(f->ptr);
    f->ptr = nullptr;
  });
}

GauXCMolecularWeights gauxc_molecular_weights_factory_get_instance(
    GauXCStatus* st, const GauXCMolecularWeightsFactory f) {
  return guarded(st, [&] {
    auto& fac = deref<MolecularWeightsFactory>(f, GauXC_Type_MolecularWeightsFactory, "factory");
    return make_handle<GauXCMolecularWeights>(GauXC_Type_MolecularWeights,
                                              new MolecularWeights(fac.get_instance()));
  });
}

void gauxc_molecular_weights_modify_weights(GauXCStatus* st, const GauXCMolecularWeights mw,
                                            const GauXCLoadBalancer lb) {
  guarded_void(st, [&] {
    auto& w = deref<MolecularWeights>(mw, GauXC_Type_MolecularWeights, "molecular weights");
    auto& h = deref<LBHolder>(lb, GauXC_Type_LoadBalancer, "load balancer");
    w.modify_weights(*h.lb);
  });
}

void gauxc_molecular_weights_delete(GauXCStatus* st, GauXCMolecularWeights* mw) {
  guarded_void(st, [&] {
    if (!mw) return;
    delete static_cast<MolecularWeights*>(mw->ptr);
    mw->ptr = nullptr;
  });
}

GauXCFunctional gauxc_functional_from_string(GauXCStatus* st, const char* spec, bool polarized) {
  return guarded(st, [&] {
    if (!spec) throw std::runtime_error("null functional spec");
    return make_handle<GauXCFunctional>(
        GauXC_Type_Functional, new functional_type(functional_from_string(spec, polarized)));
  });
}

void gauxc_functional_delete(GauXCStatus* st, GauXCFunctional* f) {
  guarded_void(st, [&] {
    if (!f) return;
    delete static_cast<functional_type*>(f->ptr);
    f->ptr = nullptr;
  });
}

GauXCIntegrator gauxc_integrator_new(GauXCStatus* st, const GauXCFunctional func,
                                     const GauXCLoadBalancer lb, enum GauXC_ExecutionSpace ex,
                                     const char* input_type, const char* ikern, const char* lwd,
                                     const char* red) {
  return guarded(st, [&] {
    const auto& f = deref<functional_type>(func, GauXC_Type_Functional, "functional");
    auto& h = deref<LBHolder>(lb, GauXC_Type_LoadBalancer, "load balancer");
    // The C API shares the caller's load balancer (weights and tasks).
    auto impl = make_device_integrator_impl(
        static_cast<ExecutionSpace>(ex), input_type ? input_type : "Replicated",
        ikern ? ikern : "Default", lwd ? lwd : "Default", red ? red : "Default",
        std::make_shared<functional_type>(f), h.lb);
    return make_handle<GauXCIntegrator>(GauXC_Type_Integrator, new IntegratorT(impl));
  });
}

void gauxc_integrator_delete(GauXCStatus* st, GauXCIntegrator* in) {
  guarded_void(st, [&] {
    if (!in) return;
    delete static_cast<IntegratorT*>(in->ptr);
    in->ptr = nullptr;
  });
}

void gauxc_integrator_integrate_den(GauXCStatus* st, const GauXCIntegrator in, const int64_t m,
                                    const int64_t n, const double* P, const int64_t ldp,
                                    double* den) {
  guarded_void(st, [&] {
    if (!P || !den) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().integrate_den(m, n, P, ldp, den);
  });
}

void gauxc_integrator_eval_exc_rks(GauXCStatus* st, const GauXCIntegrator in, const int64_t m,
                                   const int64_t n, const double* P, const int64_t ldp,
                                   double* exc) {
  guarded_void(st, [&] {
    if (!P || !exc) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc(m, n, P, ldp, exc, IntegratorSettingsXC{});
  });
}

void gauxc_integrator_eval_exc_vxc_rks(GauXCStatus* st, const GauXCIntegrator in,
                                       const int64_t m, const int64_t n, const double* P,
                                       const int64_t ldp, double* exc, double* vxc,
                                       const int64_t ldvxc) {
  guarded_void(st, [&] {
    if (!P || !exc || !vxc) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc_vxc(m, n, P, ldp, vxc, ldvxc, exc, IntegratorSettingsXC{});
  });
}

void gauxc_integrator_eval_exc_uks(GauXCStatus* st, const GauXCIntegrator in, const int64_t m,
                                   const int64_t n, const double* Ps, const int64_t ldps,
                                   const double* Pz, const int64_t ldpz, double* exc) {
  guarded_void(st, [&] {
    if (!Ps || !Pz || !exc) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc_uks(m, n, Ps, ldps, Pz, ldpz, exc, IntegratorSettingsXC{});
  });
}

void gauxc_integrator_eval_exc_vxc_uks(GauXCStatus* st, const GauXCIntegrator in,
                                       const int64_t m, const int64_t n, const double* Ps,
                                       const int64_t ldps, const double* Pz, const int64_t ldpz,
                                       double* exc, double* vxc_s, const int64_t ldvxc_s,
                                       double* vxc_z, const int64_t ldvxc_z) {
  guarded_void(st, [&] {
    if (!Ps || !Pz || !exc || !vxc_s || !vxc_z) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc_vxc_uks(m, n, Ps, ldps, Pz, ldpz, vxc_s, ldvxc_s, vxc_z, ldvxc_z,
                                     exc, IntegratorSettingsXC{});
  });
}

void gauxc_integrator_eval_exc_grad_rks(GauXCStatus* st, const GauXCIntegrator in,
                                        const int64_t m, const int64_t n, const double* P,
                                        const int64_t ldp, double* exc_grad) {
  guarded_void(st, [&] {
    if (!P || !exc_grad) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc_grad(m, n, P, ldp, exc_grad, IntegratorSettingsEXC_GRAD{});
  });
}

void gauxc_integrator_eval_exc_grad_hf_rks(GauXCStatus* st, const GauXCIntegrator in,
                                           const int64_t m, const int64_t n, const double* P,
                                           const int64_t ldp, double* exc_grad) {
  guarded_void(st, [&] {
    if (!P || !exc_grad) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    IntegratorSettingsEXC_GRAD gs;
    gs.include_weight_derivatives = false;
    I.device_impl().eval_exc_grad(m, n, P, ldp, exc_grad, gs);
  });
}

void gauxc_integrator_eval_exc_grad_uks(GauXCStatus* st, const GauXCIntegrator in,
                                        const int64_t m, const int64_t n, const double* Ps,
                                        const int64_t ldps, const double* Pz,
                                        const int64_t ldpz, double* exc_grad) {
  guarded_void(st, [&] {
    if (!Ps || !Pz || !exc_grad) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_exc_grad_uks(m, n, Ps, ldps, Pz, ldpz, exc_grad,
                                      IntegratorSettingsEXC_GRAD{});
  });
}

void gauxc_integrator_eval_exc_grad_hf_uks(GauXCStatus* st, const GauXCIntegrator in,
                                           const int64_t m, const int64_t n, const double* Ps,
                                           const int64_t ldps, const double* Pz,
                                           const int64_t ldpz, double* exc_grad) {
  guarded_void(st, [&] {
    if (!Ps || !Pz || !exc_grad) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    IntegratorSettingsEXC_GRAD gs;
    gs.include_weight_derivatives = false;
    I.device_impl().eval_exc_grad_uks(m, n, Ps, ldps, Pz, ldpz, exc_grad, gs);
  });
}

void gauxc_integrator_eval_fxc_contraction_rks(GauXCStatus* st, const GauXCIntegrator in,
                                               const int64_t m, const int64_t n,
                                               const double* P, const int64_t ldp,
                                               const double* tP, const int64_t ldtp,
                                               double* fxc, const int64_t ldfxc) {
  guarded_void(st, [&] {
    if (!P || !tP || !fxc) throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_fxc_contraction(m, n, P, ldp, tP, ldtp, fxc, ldfxc,
                                         IntegratorSettingsXC{});
  });
}

void gauxc_integrator_eval_fxc_contraction_uks(
    GauXCStatus* st, const GauXCIntegrator in, const int64_t m, const int64_t n,
    const double* Ps, const int64_t ldps, const double* Pz, const int64_t ldpz,
    const double* tPs, const int64_t ldtps, const double* tPz, const int64_t ldtpz,
    double* fxc_s, const int64_t ldfxc_s, double* fxc_z, const int64_t ldfxc_z) {
  guarded_void(st, [&] {
    if (!Ps || !Pz || !tPs || !tPz || !fxc_s || !fxc_z)
      throw std::runtime_error("null pointer argument");
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    I.device_impl().eval_fxc_contraction_uks(m, n, Ps, ldps, Pz, ldpz, tPs, ldtps, tPz, ldtpz,
                                             fxc_s, ldfxc_s, fxc_z, ldfxc_z,
                                             IntegratorSettingsXC{});
  });
}

void* gauxc_integrator_device_context(GauXCStatus* st, const GauXCIntegrator in) {
  return guarded(st, [&]() -> void* {
    auto& I = deref<IntegratorT>(in, GauXC_Type_Integrator, "integrator");
    return I.device_impl().device_context();
  });
}

// Type-dispatched deletion (GauXC src/c-api/c_types.cxx:29-109): the handle's
// header names its type.
void gauxc_object_delete(GauXCStatus* st, void** obj) {
  set_status(st, 0, nullptr);
  if (!obj || !*obj) return;
  switch (static_cast<const GauXCHeader*>(*obj)->type) {
    case GauXC_Type_Molecule: gauxc_molecule_delete(st, static_cast<GauXCMolecule*>(*obj)); break;
    case GauXC_Type_BasisSet: gauxc_basisset_delete(st, static_cast<GauXCBasisSet*>(*obj)); break;
    case GauXC_Type_MolGrid: gauxc_molgrid_delete(st, static_cast<GauXCMolGrid*>(*obj)); break;
    case GauXC_Type_RuntimeEnvironment:
      gauxc_runtime_environment_delete(st, static_cast<GauXCRuntimeEnvironment*>(*obj));
      break;
    case GauXC_Type_LoadBalancer:
      gauxc_load_balancer_delete(st, static_cast<GauXCLoadBalancer*>(*obj));
      break;
    case GauXC_Type_LoadBalancerFactory:
      gauxc_load_balancer_factory_delete(st, static_cast<GauXCLoadBalancerFactory*>(*obj));
      break;
    case GauXC_Type_MolecularWeights:
      gauxc_molecular_weights_delete(st, static_cast<GauXCMolecularWeights*>(*obj));
      break;
    case GauXC_Type_MolecularWeightsFactory:
      gauxc_molecular_weights_factory_delete(st,
                                             static_cast<GauXCMolecularWeightsFactory*>(*obj));
      break;
    case GauXC_Type_Functional:
      gauxc_functional_delete(st, static_cast<GauXCFunctional*>(*obj));
      break;
    case GauXC_Type_Integrator:
      gauxc_integrator_delete(st, static_cast<GauXCIntegrator*>(*obj));
      break;
    default: set_status(st, 1, "Unknown object type in gauxc_object_delete"); break;
  }
}

void gauxc_objects_delete(GauXCStatus* st, void** objs, size_t n) {
  set_status(st, 0, nullptr);
  for (size_t i = 0; objs && i < n; ++i) {
    if (!objs[i]) continue;
    gauxc_object_delete(st, &objs[i]);
    if (st && st->code != 0) return;
  }
}

}  // extern "C"
