// Runtime environments (GauXC: include/gauxc/runtime_environment/decl.hpp:26-86).
// This build runs one process per GPU and has no MPI layer: comm_rank() is 0
// and comm_size() is 1.  DeviceRuntimeEnvironment carries the device ordinal
// and either the fraction of free device memory the integrator may use as
// its working arena (GauXC's fill_fraction constructor, decl.hpp:73) or a
// caller-owned device buffer that becomes that arena (decl.hpp:71-72).  The
// load balancer keeps a copy of the environment it was built with, and the
// device integrator and molecular weights read the device, the fill fraction
// and the caller's buffer from it.
#pragma once
#include <cstddef>
#include <memory>

namespace GauXC {

class RuntimeEnvironment {
 public:
  RuntimeEnvironment() = default;
  virtual ~RuntimeEnvironment() = default;
  int comm_rank() const { return 0; }
  int comm_size() const { return 1; }
};

class DeviceRuntimeEnvironment : public RuntimeEnvironment {
  int device_ = 0;
  double fill_fraction_ = 0.9;
  void* user_mem_ = nullptr;
  size_t user_mem_size_ = 0;

 public:
  explicit DeviceRuntimeEnvironment(double fill_fraction = 0.9, int device = 0)
      : device_(device), fill_fraction_(fill_fraction) {}
  /// Caller-provided device arena (device pointer and size in bytes).
  DeviceRuntimeEnvironment(void* mem, size_t mem_sz, int device = 0)
      : device_(device), fill_fraction_(0.), user_mem_(mem), user_mem_size_(mem_sz) {}
  int device_id() const { return device_; }
  double fill_fraction() const { return fill_fraction_; }
  void* device_memory() const { return user_mem_; }
  size_t device_memory_size() const { return user_mem_size_; }
  bool owns_memory() const { return user_mem_ == nullptr; }
};

/// The device environment behind a runtime handle, or nullptr for a plain
/// (host) RuntimeEnvironment, whose device work then runs on device 0 with
/// the default fill fraction.
inline const DeviceRuntimeEnvironment* as_device_runtime_ptr(const RuntimeEnvironment& rt) {
  return dynamic_cast<const DeviceRuntimeEnvironment*>(&rt);
}

/// A copy of `rt` that keeps its dynamic type.
inline std::shared_ptr<const RuntimeEnvironment> copy_runtime(const RuntimeEnvironment& rt) {
  if (auto d = as_device_runtime_ptr(rt)) return std::make_shared<DeviceRuntimeEnvironment>(*d);
  return std::make_shared<RuntimeEnvironment>(rt);
}

}  // namespace GauXC
