// Partition (molecular) weights applied in place to a load balancer's tasks.
// GauXC counterpart: include/gauxc/molecular_weights.hpp:24-106 and
// src/molecular_weights/device/device_molecular_weights.cxx:18 (SSF only on
// the device).  The device path evaluates the Stratmann-Scuseria-Frisch
// weights with this build's HIP kernel.
#pragma once
#include <gauxc/enums.hpp>
#include <gauxc/load_balancer.hpp>
#include <memory>
#include <string>

namespace GauXC {

struct MolecularWeightsSettings {
  XCWeightAlg weight_alg = XCWeightAlg::SSF;
  bool becke_size_adjustment = false;
};

class MolecularWeights {
  ExecutionSpace ex_;
  MolecularWeightsSettings settings_;
  int device_;

 public:
  MolecularWeights(ExecutionSpace ex, MolecularWeightsSettings s, int device = 0)
      : ex_(ex), settings_(s), device_(device) {}
  /// Multiply every task weight by its SSF partition weight.  Throws if the
  /// load balancer's weights were already modified.
  void modify_weights(LoadBalancer& lb) const;
  const MolecularWeightsSettings& settings() const { return settings_; }
};

class MolecularWeightsFactory {
  ExecutionSpace ex_;
  std::string lwd_kernel_;
  MolecularWeightsSettings settings_;

 public:
  MolecularWeightsFactory(ExecutionSpace ex, std::string local_work_kernel_name,
                          MolecularWeightsSettings settings);
  MolecularWeights get_instance() const { return MolecularWeights(ex_, settings_); }
  std::shared_ptr<MolecularWeights> get_shared_instance() const {
    return std::make_shared<MolecularWeights>(ex_, settings_);
  }
};

}  // namespace GauXC
