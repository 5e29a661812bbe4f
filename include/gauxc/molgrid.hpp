// Atom-centred quadratures, their micro-batching, and the per-element grid
// map of a molecule.
//
// GauXC counterparts: include/gauxc/grid.hpp, molgrid.hpp,
// molgrid/defaults.hpp and src/grid_factory.cxx:35-69,
// src/molgrid_defaults.cxx:52-71,144-202.  GauXC delegates the quadrature
// arithmetic and batching to IntegratorXX; this build implements them
// itself:
//  * Mura-Knowles radial rule: x_i = (i+1)/(n+1), r_i = -s ln(1 - x_i^3),
//    w_i = 3 s x_i^2 / ((1 - x_i^3)(n+1)) * r_i^2  (s = radial scale).
//  * Lebedev-Laikov angular rules with weights summing to 4 pi.
//  * Product grid: every radial node carries a full angular shell; a pruned
//    grid gives the radial nodes of each pruning region the region's
//    angular rule (src/grid_factory.cxx:83-148, the Robust and Treutler
//    schemes of :155-240).
//  * Batching: recursive octree bisection of the atomic grid's bounding
//    box until a cell holds at most BatchSize points; a batch's screening
//    box is the tight bounding box of its points.
#pragma once
#include <gauxc/enums.hpp>
#include <gauxc/molecule.hpp>
#include <gauxc/named_type.hpp>
#include <array>
#include <map>
#include <memory>
#include <tuple>
#include <variant>
#include <vector>

namespace GauXC {

/// Quadrature of one atom, centred at the origin.
struct AtomicQuadrature {
  std::vector<std::array<double, 3>> points;
  std::vector<double> weights;
  size_t npts() const { return weights.size(); }
};

/// Points of one batch (translated to the atom) and their bounding box.
struct GridBatch {
  std::array<double, 3> lo, up;
  std::vector<std::array<double, 3>> points;
  std::vector<double> weights;
};

class MicroBatcher {
  std::shared_ptr<const AtomicQuadrature> quad_;
  size_t max_batch_;
  std::vector<size_t> order_;    // permutation of quadrature points, grouped by batch
  std::vector<size_t> offsets_;  // batch b owns order_[offsets_[b] .. offsets_[b+1])
 public:
  MicroBatcher(size_t max_batch, std::shared_ptr<const AtomicQuadrature> q);
  size_t nbatches() const { return offsets_.size() - 1; }
  size_t max_batch_size() const { return max_batch_; }
  const AtomicQuadrature& quadrature() const { return *quad_; }
  /// Batch ibatch with points translated by center.
  GridBatch at(size_t ibatch, const std::array<double, 3>& center) const;
};

class Grid {
  std::shared_ptr<const AtomicQuadrature> quad_;
  std::shared_ptr<MicroBatcher> batcher_;
 public:
  Grid() = default;
  Grid(std::shared_ptr<const AtomicQuadrature> q, BatchSize bsz);
  const MicroBatcher& batcher() const { return *batcher_; }
  const AtomicQuadrature& quadrature() const { return *quad_; }
  size_t npts() const { return quad_->npts(); }
  size_t max_batch_size() const { return batcher_->max_batch_size(); }
};

struct UnprunedAtomicGridSpecification {
  RadialQuad radial_quad;
  RadialSize radial_size;
  RadialScale radial_scale;
  AngularSize angular_size;
};

/// Radial nodes [idx_st, idx_en) carry the angular rule of angular_size
/// (GauXC include/gauxc/grid_factory.hpp:32-45).
struct PruningRegion {
  size_t idx_st;
  size_t idx_en;
  AngularSize angular_size;
};

struct PrunedAtomicGridSpecification {
  RadialQuad radial_quad;
  RadialSize radial_size;
  RadialScale radial_scale;
  std::vector<PruningRegion> pruning_regions;
};

using atomic_grid_variant =
    std::variant<UnprunedAtomicGridSpecification, PrunedAtomicGridSpecification>;
using atomic_grid_spec_map = std::map<AtomicNumber, atomic_grid_variant>;

/// Psi4's "Robust" pruning of an unpruned Lebedev specification: the first
/// quarter of the radial nodes (n/4 + 1) on the order-7 rule, up to n/2 + 1
/// on the rule of order (base - 6) (rounded up to a tabulated order; the
/// base order when it is 6 or less), the rest on the base rule
/// (src/grid_factory.cxx:155-187).
PrunedAtomicGridSpecification robust_psi4_pruning_scheme(UnprunedAtomicGridSpecification unp);
/// Treutler's pruning: n/3 + 1 nodes on order 7, up to n/2 + 1 on order 11,
/// the rest on the base rule (src/grid_factory.cxx:191-220).
PrunedAtomicGridSpecification treutler_pruning_scheme(UnprunedAtomicGridSpecification unp);
/// The scheme's pruned specification; Unpruned gives a single region
/// (src/grid_factory.cxx:225-250).
PrunedAtomicGridSpecification create_pruned_spec(PruningScheme scheme,
                                                 UnprunedAtomicGridSpecification unp);

class MolGrid {
  std::map<AtomicNumber, Grid> grids_;
 public:
  MolGrid() = default;
  explicit MolGrid(const std::map<AtomicNumber, Grid>& g) : grids_(g) {}
  MolGrid(const atomic_grid_spec_map& specs, BatchSize bsz);
  size_t natoms_uniq() const { return grids_.size(); }
  const Grid& get_grid(AtomicNumber Z) const;
  size_t max_nbatches() const;
  int64_t max_batch_size() const;
};

RadialScale default_mk_radial_scaling_factor(AtomicNumber Z);
RadialScale default_radial_scaling_factor(RadialQuad rq, AtomicNumber Z);
std::tuple<RadialSize, AngularSize> default_grid_size(AtomicNumber Z, RadialQuad rq,
                                                      AtomicGridSizeDefault s);

struct AtomicGridFactory {
  static AtomicQuadrature mura_knowles(RadialSize nrad, RadialScale scale);
  static AtomicQuadrature lebedev_laikov(AngularSize nang);
  static std::shared_ptr<AtomicQuadrature> generate_unpruned_quadrature(RadialQuad rq,
                                                                        RadialSize nrad,
                                                                        AngularSize nang,
                                                                        RadialScale rscal);
  static Grid generate_unpruned_grid(RadialQuad rq, RadialSize nrad, AngularSize nang,
                                     RadialScale rscal, BatchSize bsz);
  static std::shared_ptr<AtomicQuadrature> generate_pruned_quadrature(
      RadialQuad rq, RadialSize nrad, const std::vector<PruningRegion>& regions,
      RadialScale rscal);
  static Grid generate_pruned_grid(RadialQuad rq, RadialSize nrad,
                                   const std::vector<PruningRegion>& regions, RadialScale rscal,
                                   BatchSize bsz);
  static Grid generate_grid(const atomic_grid_variant& spec, BatchSize bsz);
};

struct MolGridFactory {
  static UnprunedAtomicGridSpecification create_default_unpruned_grid_spec(
      AtomicNumber Z, RadialQuad rq, AtomicGridSizeDefault s);
  static atomic_grid_spec_map create_default_gridmap(const Molecule& mol, PruningScheme ps,
                                                     RadialQuad rq, AtomicGridSizeDefault s);
  static MolGrid create_default_molgrid(const Molecule& mol, PruningScheme ps, BatchSize bsz,
                                        RadialQuad rq, AtomicGridSizeDefault s);
};

}  // namespace GauXC
