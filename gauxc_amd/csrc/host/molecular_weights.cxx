// SSF molecular weights through the device C ABI.
// GauXC behaviour: src/molecular_weights/molecular_weights.cxx:27-31 and
// device/device_molecular_weights.cxx:18 (refuses a second application,
// SSF only on the device).
#include <gauxc/molecular_weights.hpp>
#include <gauxc/exceptions.hpp>
#include <gauxc_amd/device.h>

#include <algorithm>
#include <cctype>
#include <memory>
#include <string>
#include <vector>

namespace GauXC {

MolecularWeightsFactory::MolecularWeightsFactory(ExecutionSpace ex, std::string lwd,
                                                 MolecularWeightsSettings s)
    : ex_(ex), lwd_kernel_(std::move(lwd)), settings_(s) {
  std::transform(lwd_kernel_.begin(), lwd_kernel_.end(), lwd_kernel_.begin(),
                 [](unsigned char c) { return std::toupper(c); });
  if (lwd_kernel_ != "DEFAULT" && lwd_kernel_ != "SCHEME1" && lwd_kernel_ != "SCHEME1-MIXED")
    GAUXC_GENERIC_EXCEPTION("Local Work Driver Not Recognized: " + lwd_kernel_);
}

void MolecularWeights::modify_weights(LoadBalancer& lb) const {
  if (lb.state().modified_weights_are_stored)
    GAUXC_GENERIC_EXCEPTION("Attempting to Overwrite Modified Weights");
  if (settings_.weight_alg != XCWeightAlg::SSF)
    GAUXC_GENERIC_EXCEPTION("Only SSF partition weights are provided by the device backend");
  if (ex_ != ExecutionSpace::Device)
    GAUXC_GENERIC_EXCEPTION("MolecularWeights: only ExecutionSpace::Device is provided");

  auto& tasks = lb.get_tasks();
  const auto& mol = lb.molecule();
  const size_t ntasks = tasks.size();
  std::vector<size_t> off(ntasks + 1, 0);
  for (size_t it = 0; it < ntasks; ++it) off[it + 1] = off[it] + tasks[it].points.size();
  const size_t npts = off[ntasks];
  // Left uninitialised (every element is written below): the parallel
  // packing loop is then also the first touch of the pages, instead of a
  // serial zero fill of about 40 bytes per point.
  std::unique_ptr<double[]> pts(new double[3 * npts]), w(new double[npts]);
  std::unique_ptr<int32_t[]> parent(new int32_t[npts]);
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t it = 0; it < ntasks; ++it) {
    const auto& t = tasks[it];
    for (size_t i = 0, k = off[it]; i < t.points.size(); ++i, ++k) {
      pts[3 * k] = t.points[i][0];
      pts[3 * k + 1] = t.points[i][1];
      pts[3 * k + 2] = t.points[i][2];
      w[k] = t.weights[i];
      parent[k] = t.iParent;
    }
  }
  std::vector<double> xyz(3 * mol.natoms());
  for (size_t a = 0; a < mol.natoms(); ++a) {
    xyz[3 * a] = mol[a].x;
    xyz[3 * a + 1] = mol[a].y;
    xyz[3 * a + 2] = mol[a].z;
  }
  // The load balancer's device environment names the device (GauXC runs
  // the device weights on the runtime it was given).
  const auto* drt = as_device_runtime_ptr(lb.runtime());
  const int device = drt ? drt->device_id() : device_;
  if (xcq_ssf_weights(device, static_cast<int32_t>(mol.natoms()), xyz.data(),
                      static_cast<int64_t>(npts), pts.get(), parent.get(), w.get()) !=
      XCQ_SUCCESS)
    GAUXC_GENERIC_EXCEPTION(std::string("SSF weights failed: ") + xcq_last_error());
#pragma omp parallel for schedule(dynamic, 64)
  for (size_t it = 0; it < ntasks; ++it) {
    auto& t = tasks[it];
    for (size_t i = 0; i < t.weights.size(); ++i) t.weights[i] = w[off[it] + i];
  }
  lb.state().modified_weights_are_stored = true;
  lb.mark_modified();
}

}  // namespace GauXC
