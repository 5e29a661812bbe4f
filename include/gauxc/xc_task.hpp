// One unit of quadrature work: a batch of grid points of one parent atom and
// the shells that survive screening on it.
// GauXC counterpart: include/gauxc/xc_task.hpp:25-119 (same fields, merge
// rule "same parent and same shell list", and the cost heuristics
// cost() / cost_exc_vxc() that the load balancer and the task sort use).
#pragma once
#include <gauxc/exceptions.hpp>
#include <array>
#include <cstdint>
#include <limits>
#include <vector>

namespace GauXC {

struct XCTask {
  int32_t iParent = -1;
  std::vector<std::array<double, 3>> points;
  std::vector<double> weights;
  int32_t npts = 0;
  double dist_nearest = 0.;
  double max_weight = std::numeric_limits<double>::infinity();

  struct screening_data {
    std::vector<int32_t> shell_list;
    int32_t nbe = 0;
    bool equiv_with(const screening_data& o) const { return shell_list == o.shell_list; }
  };
  screening_data bfn_screening;

  bool equiv_with(const XCTask& o) const {
    return iParent == o.iParent && bfn_screening.equiv_with(o.bfn_screening);
  }

  void merge_with(const XCTask& o) {
    if (!equiv_with(o)) GAUXC_GENERIC_EXCEPTION("Cannot Perform Requested Merge: Incompatible Tasks");
    points.insert(points.end(), o.points.begin(), o.points.end());
    weights.insert(weights.end(), o.weights.begin(), o.weights.end());
    npts = static_cast<int32_t>(points.size());
  }

  /// Flop-count heuristic used for rank assignment (includes the SSF term).
  size_t cost(size_t n_deriv, size_t natoms) const {
    const size_t nbe = static_cast<size_t>(bfn_screening.nbe);
    return (nbe * (1 + nbe + n_deriv) + natoms * natoms) * static_cast<size_t>(npts);
  }
  /// Flop-count heuristic of the EXC/VXC work of this task.
  size_t cost_exc_vxc(size_t n_deriv) const {
    const size_t nbe = static_cast<size_t>(bfn_screening.nbe);
    return nbe * (1 + nbe + n_deriv) * static_cast<size_t>(npts);
  }
};

}  // namespace GauXC
