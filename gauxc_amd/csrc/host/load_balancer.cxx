// Replicated task generation.  Behaviour follows
// src/load_balancer/host/replicated_host_load_balancer.cxx:22-193 and
// petite_replicated_load_balancer.cxx:31-65 (see load_balancer.hpp).
#include <gauxc/load_balancer.hpp>
#include <gauxc/exceptions.hpp>

#include <algorithm>
#include <cctype>
#include <numeric>

namespace GauXC {

namespace geometry {
bool cube_sphere_intersect(const double* lo, const double* up, const double* c, double rad) {
  // Squared distance from the centre to the box, compared with rad^2.
  double rem = rad * rad;
  for (int d = 0; d < 3; ++d) {
    double gap = 0.;
    if (c[d] < lo[d]) gap = lo[d] - c[d];
    else if (c[d] > up[d]) gap = c[d] - up[d];
    rem -= gap * gap;
    if (rem < 0.) return false;
  }
  return true;
}
}  // namespace geometry

LoadBalancer::LoadBalancer(const Molecule& mol, const MolGrid& mg,
                           const BasisSet<double>& basis, const RuntimeEnvironment& rt)
    : mol_(std::make_shared<Molecule>(mol)),
      mg_(std::make_shared<MolGrid>(mg)),
      basis_(std::make_shared<BasisSet<double>>(basis)),
      meta_(std::make_shared<MolMeta>(mol)),
      rt_(copy_runtime(rt)) {}

LoadBalancer::LoadBalancer(const LoadBalancer& o)
    : mol_(o.mol_), mg_(o.mg_), basis_(o.basis_), meta_(o.meta_), rt_(o.rt_), tasks_(o.tasks_),
      built_(o.built_), state_(o.state_), revision_(o.revision_) {}

void LoadBalancer::build_tasks() const {
  const auto& mol = *mol_;
  const auto& basis = *basis_;
  const size_t natoms = mol.natoms();
  std::vector<XCTask> local;
  for (size_t ia = 0; ia < natoms; ++ia) {
    const auto& atom = mol[ia];
    const std::array<double, 3> center = {atom.x, atom.y, atom.z};
    const auto& batcher = mg_->get_grid(atom.Z).batcher();
    const size_t nb = batcher.nbatches();
    std::vector<XCTask> atom_tasks(nb);
    std::vector<char> keep(nb, 0);
#pragma omp parallel for schedule(dynamic)
    for (size_t ib = 0; ib < nb; ++ib) {
      GridBatch b = batcher.at(ib, center);
      if (b.points.empty()) continue;
      XCTask t;
      int32_t nbe = 0;
      for (int32_t s = 0; s < basis.nshells(); ++s) {
        const auto& sh = basis[s];
        if (geometry::cube_sphere_intersect(b.lo.data(), b.up.data(), sh.O_data(),
                                            sh.cutoff_radius())) {
          t.bfn_screening.shell_list.push_back(s);
          nbe += sh.size();
        }
      }
      if (t.bfn_screening.shell_list.empty()) continue;
      t.iParent = static_cast<int32_t>(ia);
      t.npts = static_cast<int32_t>(b.points.size());
      t.points = std::move(b.points);
      t.weights = std::move(b.weights);
      t.bfn_screening.nbe = nbe;
      t.dist_nearest = meta_->dist_nearest()[ia];
      atom_tasks[ib] = std::move(t);
      keep[ib] = 1;
    }
    for (size_t ib = 0; ib < nb; ++ib)
      if (keep[ib]) local.push_back(std::move(atom_tasks[ib]));
  }
  std::stable_sort(local.begin(), local.end(), [](const XCTask& a, const XCTask& b) {
    if (a.iParent != b.iParent) return a.iParent < b.iParent;
    return a.bfn_screening.shell_list < b.bfn_screening.shell_list;
  });
  std::vector<XCTask> merged;
  for (auto& t : local) {
    if (!merged.empty() && merged.back().equiv_with(t)) merged.back().merge_with(t);
    else merged.push_back(std::move(t));
  }
  tasks_ = std::move(merged);
  built_ = true;
}

const std::vector<XCTask>& LoadBalancer::get_tasks() const {
  if (!built_) build_tasks();
  return tasks_;
}

std::vector<XCTask>& LoadBalancer::get_tasks() {
  if (!built_) build_tasks();
  return tasks_;
}

size_t LoadBalancer::max_npts() const {
  size_t m = 0;
  for (const auto& t : get_tasks()) m = std::max(m, t.points.size());
  return m;
}

size_t LoadBalancer::total_npts() const {
  size_t n = 0;
  for (const auto& t : get_tasks()) n += t.points.size();
  return n;
}

static std::string upper(std::string s) {
  std::transform(s.begin(), s.end(), s.begin(), [](unsigned char c) { return std::toupper(c); });
  return s;
}

LoadBalancerFactory::LoadBalancerFactory(ExecutionSpace ex, std::string kernel_name)
    : ex_(ex), kernel_(upper(std::move(kernel_name))) {
  if (kernel_ != "DEFAULT" && kernel_ != "REPLICATED" && kernel_ != "REPLICATED-PETITE")
    GAUXC_GENERIC_EXCEPTION("Load Balancer Kernel Not Recognized: " + kernel_);
}

LoadBalancer LoadBalancerFactory::get_instance(const RuntimeEnvironment& rt, const Molecule& mol,
                                               const MolGrid& mg,
                                               const BasisSet<double>& basis) {
  return LoadBalancer(mol, mg, basis, rt);
}

std::shared_ptr<LoadBalancer> LoadBalancerFactory::get_shared_instance(
    const RuntimeEnvironment& rt, const Molecule& mol, const MolGrid& mg,
    const BasisSet<double>& basis) {
  return std::make_shared<LoadBalancer>(get_instance(rt, mol, mg, basis));
}

}  // namespace GauXC
