// Task generation ("load balancing") for the replicated integrator.
//
// GauXC counterpart: include/gauxc/load_balancer.hpp:71-198 and
// src/load_balancer/host/replicated_host_load_balancer.cxx:22-193 with the
// "petite" screening of petite_replicated_load_balancer.cxx:31-65:
//  * for every atom and every micro-batch of its grid, keep each shell whose
//    screening sphere (centre, cutoff radius) intersects the batch's box;
//  * batches with no surviving shell are dropped;
//  * tasks are ordered by (parent atom, shell list) and consecutive tasks
//    with equal parent and shell list are merged (points concatenated).
// With a single process every task is local, so GauXC's greedy rank
// assignment reduces to keeping all of them.  ExecutionSpace::Device runs
// the screening on the GPU (GauXC's device load balancer,
// src/load_balancer/device/hip/replicated_hip_load_balancer.cxx:72-220);
// the batching, sort and merge stay on the host, so both spaces give the
// same task list.
#pragma once
#include <gauxc/basisset.hpp>
#include <gauxc/enums.hpp>
#include <gauxc/molecule.hpp>
#include <gauxc/molgrid.hpp>
#include <gauxc/runtime_environment.hpp>
#include <gauxc/xc_task.hpp>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace GauXC {

struct LoadBalancerState {
  bool modified_weights_are_stored = false;
};

class LoadBalancer {
  std::shared_ptr<const Molecule> mol_;
  std::shared_ptr<const MolGrid> mg_;
  std::shared_ptr<const BasisSet<double>> basis_;
  std::shared_ptr<const MolMeta> meta_;
  std::shared_ptr<const RuntimeEnvironment> rt_;
  ExecutionSpace ex_ = ExecutionSpace::Host;  // where the batch screening runs
  mutable std::vector<XCTask> tasks_;
  mutable bool built_ = false;
  LoadBalancerState state_;
  uint64_t revision_ = 0;  // bumped whenever tasks may have changed

  void build_tasks() const;
  /// Screening of every (batch, shell) pair: host loops, or the device
  /// kernel of xcq_collision_detection (GauXC's device load balancer).
  void screen_batches(const std::vector<GridBatch>& batches,
                      std::vector<std::vector<int32_t>>& lists,
                      std::vector<int32_t>& nbe) const;

 public:
  LoadBalancer(const Molecule& mol, const MolGrid& mg, const BasisSet<double>& basis,
               const RuntimeEnvironment& rt = RuntimeEnvironment(),
               ExecutionSpace ex = ExecutionSpace::Host);
  LoadBalancer(const LoadBalancer&);
  LoadBalancer(LoadBalancer&&) noexcept = default;
  ~LoadBalancer() = default;

  const std::vector<XCTask>& get_tasks() const;
  std::vector<XCTask>& get_tasks();
  const Molecule& molecule() const { return *mol_; }
  const MolMeta& molmeta() const { return *meta_; }
  const MolGrid& molgrid() const { return *mg_; }
  const BasisSet<double>& basis() const { return *basis_; }
  /// The runtime environment this load balancer was built with (GauXC
  /// include/gauxc/load_balancer.hpp:115-116).
  const RuntimeEnvironment& runtime() const { return *rt_; }
  LoadBalancerState& state() { return state_; }
  const LoadBalancerState& state() const { return state_; }
  size_t max_npts() const;
  size_t total_npts() const;
  /// Where the batch screening runs (the factory's execution space).
  ExecutionSpace execution_space() const { return ex_; }
  /// Monotone counter that changes whenever the task list or its weights change.
  uint64_t revision() const { return revision_; }
  void mark_modified() { ++revision_; }
};

class LoadBalancerFactory {
  ExecutionSpace ex_;
  std::string kernel_;

 public:
  LoadBalancerFactory(ExecutionSpace ex, std::string kernel_name);
  LoadBalancer get_instance(const RuntimeEnvironment& rt, const Molecule& mol, const MolGrid& mg,
                            const BasisSet<double>& basis);
  std::shared_ptr<LoadBalancer> get_shared_instance(const RuntimeEnvironment& rt,
                                                    const Molecule& mol, const MolGrid& mg,
                                                    const BasisSet<double>& basis);
};

namespace geometry {
/// True when the sphere (center, rad) intersects the axis-aligned box [lo, up].
bool cube_sphere_intersect(const double* lo, const double* up, const double* center, double rad);
}

}  // namespace GauXC
