// A C++ caller of the GauXC-compatible API, written as a GauXC user would
// (the call sequence of GauXC's tests/xc_integrator.cxx:163-212).  Built by
// tests/test_host.py (g++ against include/ and libgauxc_amd.so).
//
//   caller <input.bin> host    builds the tasks and prints their counts
//   caller <input.bin> device  also applies SSF weights on the device and
//                              prints EXC and a VXC checksum
//   caller <input.bin> device-mem <bytes>
//                              the same with a caller-owned device arena of
//                              <bytes> bytes (DeviceRuntimeEnvironment(mem, sz))
//
// input.bin: int64 natoms, natoms x (int64 Z, 3 x double xyz), int64 nshells,
// nshells x (int32 l, int32 pure, double alpha, double coeff, 3 x double O),
// int64 nbf, nbf*nbf doubles of P (column-major).
#include <gauxc/load_balancer.hpp>
#include <gauxc/molecular_weights.hpp>
#include <gauxc/molgrid.hpp>
#include <gauxc/runtime_environment.hpp>
#include <gauxc/xc_integrator.hpp>

#include <dlfcn.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

using namespace GauXC;

template <typename T>
T get(std::ifstream& f) {
  T v;
  f.read(reinterpret_cast<char*>(&v), sizeof(T));
  if (!f) throw std::runtime_error("short input file");
  return v;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: caller <input.bin> host|device|device-mem <bytes>\n");
    return 2;
  }
  try {
    std::ifstream f(argv[1], std::ios::binary);
    Molecule mol;
    const auto natoms = get<int64_t>(f);
    for (int64_t a = 0; a < natoms; ++a) {
      const auto Z = get<int64_t>(f);
      const double x = get<double>(f), y = get<double>(f), z = get<double>(f);
      mol.emplace_back(AtomicNumber(Z), x, y, z);
    }
    BasisSet<double> basis;
    const auto nsh = get<int64_t>(f);
    for (int64_t s = 0; s < nsh; ++s) {
      const auto l = get<int32_t>(f), pure = get<int32_t>(f);
      Shell<double>::prim_array alpha{}, coeff{};
      alpha[0] = get<double>(f);
      coeff[0] = get<double>(f);
      Shell<double>::cart_array O{get<double>(f), get<double>(f), get<double>(f)};
      basis.emplace_back(PrimSize(1), AngularMomentum(l), SphericalType(pure), alpha, coeff, O);
    }
    const auto nbf = get<int64_t>(f);
    HostMatrix P(nbf, nbf);
    for (int64_t k = 0; k < nbf * nbf; ++k) P.data()[k] = get<double>(f);

    const std::string mode = argv[2];
    void* mem = nullptr;
    size_t mem_sz = 0;
    void* hip = nullptr;
    using malloc_t = int (*)(void**, size_t);
    using free_t = int (*)(void*);
    if (mode == "device-mem") {
      if (argc < 4) throw std::runtime_error("device-mem needs a byte count");
      mem_sz = std::stoull(argv[3]);
      hip = dlopen("libamdhip64.so", RTLD_NOW | RTLD_GLOBAL);
      if (!hip) throw std::runtime_error("libamdhip64.so does not load");
      auto hip_malloc = reinterpret_cast<malloc_t>(dlsym(hip, "hipMalloc"));
      if (!hip_malloc || hip_malloc(&mem, mem_sz) != 0) throw std::runtime_error("hipMalloc");
    }
    std::unique_ptr<RuntimeEnvironment> rt;
    if (mem) rt = std::make_unique<DeviceRuntimeEnvironment>(mem, mem_sz);
    else rt = std::make_unique<DeviceRuntimeEnvironment>(0.9);

    auto mg = MolGridFactory::create_default_molgrid(mol, PruningScheme::Unpruned, BatchSize(512),
                                                     RadialQuad::MuraKnowles,
                                                     AtomicGridSizeDefault::FineGrid);
    LoadBalancerFactory lb_factory(ExecutionSpace::Host, "Default");
    auto lb = lb_factory.get_instance(*rt, mol, mg, basis);
    const auto& tasks = lb.get_tasks();
    size_t npts = 0, nsl = 0;
    for (const auto& t : tasks) {
      npts += t.points.size();
      nsl += t.bfn_screening.shell_list.size();
    }
    std::printf("tasks %zu points %zu shell_entries %zu nbf %d\n", tasks.size(), npts, nsl,
                (int)basis.nbf());
    if (mode == "host") return 0;

    MolecularWeightsFactory mw_factory(ExecutionSpace::Device, "Default",
                                       MolecularWeightsSettings{});
    mw_factory.get_instance().modify_weights(lb);
    functional_type func(ExchCXX::Backend::builtin, ExchCXX::Functional::PBE,
                         ExchCXX::Spin::Unpolarized);
    XCIntegratorFactory<HostMatrix> factory(ExecutionSpace::Device, "Replicated", "Default",
                                            "Default", "Default");
    auto integrator = factory.get_instance(func, lb);
    auto [EXC, VXC] = integrator.eval_exc_vxc(P);
    double sum = 0.;
    for (int64_t k = 0; k < nbf * nbf; ++k) sum += std::abs(VXC.data()[k]);
    std::printf("EXC %.15e VXC_abs_sum %.15e\n", EXC, sum);
    if (mem) {
      auto hip_free = reinterpret_cast<free_t>(dlsym(hip, "hipFree"));
      if (hip_free) hip_free(mem);
    }
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  return 0;
}
