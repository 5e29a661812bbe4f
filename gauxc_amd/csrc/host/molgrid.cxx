// Atomic quadratures, spherical micro-batching and molecular grid defaults.
// Reference behaviour: src/grid_factory.cxx:35-69 (unpruned product grids),
// :83-148 and :155-250 (pruned grids and the pruning schemes),
// src/molgrid_defaults.cxx:52-71 (Mura-Knowles scale 7.0 for alkali and
// alkaline-earth metals, else 5.0) and :144-202 (standard grid sizes).
#include <gauxc/molgrid.hpp>
#include <gauxc/exceptions.hpp>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <set>
#include <string>

namespace GauXC {

namespace {

struct LebedevOrbit {
  double a, b, c, w;
  int n;
};
struct LebedevRule {
  int npts, order, norbits;
  const LebedevOrbit* orbits;
};
#include "lebedev_orbits.inc"

// All distinct points reachable from (a,b,c) by coordinate permutation and
// sign change; the count is checked against the tabulated orbit size.
void expand_orbit(const LebedevOrbit& o, std::vector<std::array<double, 3>>& pts,
                  std::vector<double>& wts) {
  const double v[3] = {o.a, o.b, o.c};
  static const int perm[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
  std::set<std::array<double, 3>> seen;
  for (int s = 0; s < 8; ++s)
    for (const auto& p : perm) {
      std::array<double, 3> q;
      for (int d = 0; d < 3; ++d) {
        double x = v[p[d]];
        if ((s >> d) & 1) x = -x;
        q[d] = (x == 0.) ? 0. : x;  // fold -0 into +0
      }
      if (seen.insert(q).second) {
        pts.push_back(q);
        wts.push_back(o.w);
      }
    }
  if ((int)seen.size() != o.n) GAUXC_GENERIC_EXCEPTION("Lebedev orbit expansion mismatch");
}

// Tight bounding box of the points idx[begin, end).
void tight_box(const std::vector<std::array<double, 3>>& pts, const std::vector<size_t>& idx,
               size_t begin, size_t end, std::array<double, 3>& lo, std::array<double, 3>& up) {
  lo = up = pts[idx[begin]];
  for (size_t k = begin + 1; k < end; ++k)
    for (int d = 0; d < 3; ++d) {
      lo[d] = std::min(lo[d], pts[idx[k]][d]);
      up[d] = std::max(up[d], pts[idx[k]][d]);
    }
}

// Regroups idx[begin, end) by a cell label in [0, ncell) (stable within a
// cell) and returns the cell boundaries (ncell + 1 offsets).
template <class Label>
std::vector<size_t> group_by_cell(std::vector<size_t>& idx, size_t begin, size_t end, int ncell,
                                  Label label) {
  std::vector<size_t> count(ncell + 1, 0);
  std::vector<int> lab(end - begin);
  for (size_t k = begin; k < end; ++k) {
    lab[k - begin] = label(idx[k]);
    ++count[lab[k - begin] + 1];
  }
  for (int c = 0; c < ncell; ++c) count[c + 1] += count[c];
  std::vector<size_t> tmp(idx.begin() + begin, idx.begin() + end), pos(count.begin(), count.end() - 1);
  for (size_t k = 0; k < tmp.size(); ++k) idx[begin + pos[lab[k]]++] = tmp[k];
  for (auto& c : count) c += begin;
  return count;
}

// Octree over tight boxes: a cell with more than max_batch points is split at
// the midpoint of its points' bounding box; a coordinate equal to the midpoint
// goes to the lower half.
void octree_cells(const std::vector<std::array<double, 3>>& pts, std::vector<size_t>& idx,
                  size_t begin, size_t end, size_t max_batch, std::vector<size_t>& offs) {
  if (end == begin) return;
  if (end - begin <= max_batch) {
    offs.push_back(end);
    return;
  }
  std::array<double, 3> lo, up, mid;
  tight_box(pts, idx, begin, end, lo, up);
  for (int d = 0; d < 3; ++d) mid[d] = 0.5 * (lo[d] + up[d]);
  const auto cut = group_by_cell(idx, begin, end, 8, [&](size_t k) {
    return 4 * (pts[k][0] > mid[0]) + 2 * (pts[k][1] > mid[1]) + (pts[k][2] > mid[2]);
  });
  for (int c = 0; c < 8; ++c) {
    if (cut[c + 1] - cut[c] == end - begin) {  // no split possible (coincident points)
      offs.push_back(end);
      return;
    }
    octree_cells(pts, idx, cut[c], cut[c + 1], max_batch, offs);
  }
}

// Spherical partition of an atom-centred quadrature: a cell with more than
// max_batch points is cut into 3 x 3 x 3 sub-cells at the thirds of its
// points' bounding box; the central sub-cell (around the nucleus, where the
// radial grid is densest) is partitioned the same way, the 26 outer ones by
// the octree above.  This reproduces GauXC's batches (IntegratorXX's
// SphericalMicroBatcher, grid_impl.cxx:32-37) point set for point set: the
// 664 load-balancer tasks of benzene on the ultra-fine grid and the point
// sets of the fine-grid SSF fixture tasks.
void spherical_cells(const std::vector<std::array<double, 3>>& pts, std::vector<size_t>& idx,
                     size_t begin, size_t end, size_t max_batch, std::vector<size_t>& offs) {
  if (end == begin) return;
  if (end - begin <= max_batch) {
    offs.push_back(end);
    return;
  }
  std::array<double, 3> lo, up, e1, e2;
  tight_box(pts, idx, begin, end, lo, up);
  for (int d = 0; d < 3; ++d) {
    e1[d] = lo[d] + (up[d] - lo[d]) / 3.;
    e2[d] = lo[d] + 2. * (up[d] - lo[d]) / 3.;
  }
  auto third = [&](size_t k, int d) { return (pts[k][d] >= e1[d]) + (pts[k][d] >= e2[d]); };
  const auto cut = group_by_cell(idx, begin, end, 27, [&](size_t k) {
    return 9 * third(k, 0) + 3 * third(k, 1) + third(k, 2);
  });
  constexpr int kCentre = 13;  // (1, 1, 1)
  for (int c = 0; c < 27; ++c) {
    if (cut[c + 1] - cut[c] == end - begin) {
      offs.push_back(end);
      return;
    }
    if (c == kCentre) spherical_cells(pts, idx, cut[c], cut[c + 1], max_batch, offs);
    else octree_cells(pts, idx, cut[c], cut[c + 1], max_batch, offs);
  }
}

}  // namespace

MicroBatcher::MicroBatcher(size_t max_batch, std::shared_ptr<const AtomicQuadrature> q)
    : quad_(std::move(q)), max_batch_(max_batch) {
  if (max_batch_ == 0) GAUXC_GENERIC_EXCEPTION("BatchSize must be positive");
  const auto& pts = quad_->points;
  order_.resize(pts.size());
  std::iota(order_.begin(), order_.end(), size_t{0});
  offsets_.assign(1, 0);
  if (pts.empty()) return;
  spherical_cells(pts, order_, 0, pts.size(), max_batch_, offsets_);
}

GridBatch MicroBatcher::at(size_t ib, const std::array<double, 3>& c) const {
  if (ib >= nbatches()) GAUXC_GENERIC_EXCEPTION("batch index out of range");
  GridBatch b;
  const size_t beg = offsets_[ib], end = offsets_[ib + 1];
  b.points.reserve(end - beg);
  b.weights.reserve(end - beg);
  for (size_t k = beg; k < end; ++k) {
    const auto& p = quad_->points[order_[k]];
    b.points.push_back({p[0] + c[0], p[1] + c[1], p[2] + c[2]});
    b.weights.push_back(quad_->weights[order_[k]]);
  }
  b.lo = b.points.front();
  b.up = b.points.front();
  for (const auto& p : b.points)
    for (int d = 0; d < 3; ++d) {
      b.lo[d] = std::min(b.lo[d], p[d]);
      b.up[d] = std::max(b.up[d], p[d]);
    }
  return b;
}

Grid::Grid(std::shared_ptr<const AtomicQuadrature> q, BatchSize bsz)
    : quad_(std::move(q)),
      batcher_(std::make_shared<MicroBatcher>(static_cast<size_t>(bsz.get()), quad_)) {}

AtomicQuadrature AtomicGridFactory::mura_knowles(RadialSize nrad, RadialScale scale) {
  AtomicQuadrature q;
  const int64_t n = nrad.get();
  const double s = scale.get();
  q.points.resize(n);
  q.weights.resize(n);
  for (int64_t i = 0; i < n; ++i) {
    const double x = double(i + 1) / double(n + 1);
    const double x3 = x * x * x;
    const double r = -s * std::log(1. - x3);
    const double dr = 3. * s * x * x / ((1. - x3) * double(n + 1));
    q.points[i] = {r, 0., 0.};
    q.weights[i] = dr * r * r;
  }
  return q;
}

AtomicQuadrature AtomicGridFactory::lebedev_laikov(AngularSize nang) {
  for (const auto& rule : kLebedevRules) {
    if (rule.npts != nang.get()) continue;
    AtomicQuadrature q;
    for (int k = 0; k < rule.norbits; ++k) expand_orbit(rule.orbits[k], q.points, q.weights);
    return q;
  }
  GAUXC_GENERIC_EXCEPTION("Unsupported Lebedev grid size " + std::to_string(nang.get()));
}

std::shared_ptr<AtomicQuadrature> AtomicGridFactory::generate_unpruned_quadrature(
    RadialQuad rq, RadialSize nrad, AngularSize nang, RadialScale rscal) {
  if (rq != RadialQuad::MuraKnowles)
    GAUXC_GENERIC_EXCEPTION("Only the Mura-Knowles radial quadrature is provided");
  const auto rad = mura_knowles(nrad, rscal);
  const auto ang = lebedev_laikov(nang);
  auto q = std::make_shared<AtomicQuadrature>();
  q->points.reserve(rad.npts() * ang.npts());
  q->weights.reserve(rad.npts() * ang.npts());
  for (size_t ir = 0; ir < rad.npts(); ++ir) {
    const double r = rad.points[ir][0];
    for (size_t ia = 0; ia < ang.npts(); ++ia) {
      const auto& u = ang.points[ia];
      q->points.push_back({r * u[0], r * u[1], r * u[2]});
      q->weights.push_back(rad.weights[ir] * ang.weights[ia]);
    }
  }
  return q;
}

Grid AtomicGridFactory::generate_unpruned_grid(RadialQuad rq, RadialSize nrad, AngularSize nang,
                                               RadialScale rscal, BatchSize bsz) {
  return Grid(generate_unpruned_quadrature(rq, nrad, nang, rscal), bsz);
}

std::shared_ptr<AtomicQuadrature> AtomicGridFactory::generate_pruned_quadrature(
    RadialQuad rq, RadialSize nrad, const std::vector<PruningRegion>& regions,
    RadialScale rscal) {
  if (rq != RadialQuad::MuraKnowles)
    GAUXC_GENERIC_EXCEPTION("Only the Mura-Knowles radial quadrature is provided");
  const auto rad = mura_knowles(nrad, rscal);
  auto q = std::make_shared<AtomicQuadrature>();
  size_t next = 0;
  for (const auto& reg : regions) {
    if (reg.idx_st != next || reg.idx_en < reg.idx_st || reg.idx_en > rad.npts())
      GAUXC_GENERIC_EXCEPTION("Pruning regions must tile the radial nodes in order");
    const auto ang = lebedev_laikov(reg.angular_size);
    for (size_t ir = reg.idx_st; ir < reg.idx_en; ++ir) {
      const double r = rad.points[ir][0];
      for (size_t ia = 0; ia < ang.npts(); ++ia) {
        const auto& u = ang.points[ia];
        q->points.push_back({r * u[0], r * u[1], r * u[2]});
        q->weights.push_back(rad.weights[ir] * ang.weights[ia]);
      }
    }
    next = reg.idx_en;
  }
  if (next != rad.npts())
    GAUXC_GENERIC_EXCEPTION("Pruning regions must tile the radial nodes in order");
  return q;
}

Grid AtomicGridFactory::generate_pruned_grid(RadialQuad rq, RadialSize nrad,
                                             const std::vector<PruningRegion>& regions,
                                             RadialScale rscal, BatchSize bsz) {
  return Grid(generate_pruned_quadrature(rq, nrad, regions, rscal), bsz);
}

Grid AtomicGridFactory::generate_grid(const atomic_grid_variant& spec, BatchSize bsz) {
  if (const auto* u = std::get_if<UnprunedAtomicGridSpecification>(&spec))
    return generate_unpruned_grid(u->radial_quad, u->radial_size, u->angular_size,
                                  u->radial_scale, bsz);
  const auto& p = std::get<PrunedAtomicGridSpecification>(spec);
  return generate_pruned_grid(p.radial_quad, p.radial_size, p.pruning_regions, p.radial_scale,
                              bsz);
}

namespace {

// Algebraic order of a tabulated Lebedev rule of npts points (-1 if none),
// and the smallest tabulated order >= order.
int lebedev_order(int64_t npts) {
  for (const auto& rule : kLebedevRules)
    if (rule.npts == npts) return rule.order;
  return -1;
}
int lebedev_next_order(int order) {
  int best = -1;
  for (const auto& rule : kLebedevRules)
    if (rule.order >= order && (best < 0 || rule.order < best)) best = rule.order;
  if (best < 0) GAUXC_GENERIC_EXCEPTION("No Lebedev rule of order >= " + std::to_string(order));
  return best;
}
AngularSize lebedev_size(int order) {
  for (const auto& rule : kLebedevRules)
    if (rule.order == order) return AngularSize(rule.npts);
  GAUXC_GENERIC_EXCEPTION("No Lebedev rule of order " + std::to_string(order));
}

PrunedAtomicGridSpecification three_regions(const UnprunedAtomicGridSpecification& unp,
                                            size_t first_end, AngularSize first,
                                            AngularSize second) {
  const size_t n = static_cast<size_t>(unp.radial_size.get());
  const size_t mid_end = n / 2 + 1;
  return PrunedAtomicGridSpecification{
      unp.radial_quad, unp.radial_size, unp.radial_scale,
      {{0, first_end, first}, {first_end, mid_end, second}, {mid_end, n, unp.angular_size}}};
}

}  // namespace

PrunedAtomicGridSpecification robust_psi4_pruning_scheme(UnprunedAtomicGridSpecification unp) {
  const int base = lebedev_order(unp.angular_size.get());
  if (base < 0) GAUXC_GENERIC_EXCEPTION("Invalid Base Grid");
  const int med = lebedev_next_order(base > 6 ? base - 6 : base);
  const size_t n = static_cast<size_t>(unp.radial_size.get());
  return three_regions(unp, n / 4 + 1, lebedev_size(7), lebedev_size(med));
}

PrunedAtomicGridSpecification treutler_pruning_scheme(UnprunedAtomicGridSpecification unp) {
  const size_t n = static_cast<size_t>(unp.radial_size.get());
  return three_regions(unp, n / 3 + 1, lebedev_size(7), lebedev_size(11));
}

PrunedAtomicGridSpecification create_pruned_spec(PruningScheme scheme,
                                                 UnprunedAtomicGridSpecification unp) {
  switch (scheme) {
    case PruningScheme::Robust: return robust_psi4_pruning_scheme(unp);
    case PruningScheme::Treutler: return treutler_pruning_scheme(unp);
    default:
      return PrunedAtomicGridSpecification{
          unp.radial_quad, unp.radial_size, unp.radial_scale,
          {{0, static_cast<size_t>(unp.radial_size.get()), unp.angular_size}}};
  }
}

RadialScale default_mk_radial_scaling_factor(AtomicNumber Z) {
  switch (Z.get()) {
    case 3: case 4: case 11: case 12: case 19: case 20:
    case 37: case 38: case 55: case 56: case 87: case 88:
      return RadialScale(7.0);
    default:
      return RadialScale(5.0);
  }
}

RadialScale default_radial_scaling_factor(RadialQuad rq, AtomicNumber Z) {
  if (rq == RadialQuad::MuraKnowles) return default_mk_radial_scaling_factor(Z);
  GAUXC_GENERIC_EXCEPTION("Only the Mura-Knowles radial quadrature is provided");
}

std::tuple<RadialSize, AngularSize> default_grid_size(AtomicNumber Z, RadialQuad,
                                                      AtomicGridSizeDefault s) {
  switch (s) {
    case AtomicGridSizeDefault::GM3: return {RadialSize(35), AngularSize(110)};
    case AtomicGridSizeDefault::GM5: return {RadialSize(50), AngularSize(302)};
    case AtomicGridSizeDefault::FineGrid: return {RadialSize(75), AngularSize(302)};
    case AtomicGridSizeDefault::UltraFineGrid: return {RadialSize(99), AngularSize(590)};
    case AtomicGridSizeDefault::SuperFineGrid:
      return Z.get() <= 2 ? std::make_tuple(RadialSize(175), AngularSize(974))
                          : std::make_tuple(RadialSize(250), AngularSize(974));
  }
  GAUXC_GENERIC_EXCEPTION("Not A Recognized Standard Grid");
}

UnprunedAtomicGridSpecification MolGridFactory::create_default_unpruned_grid_spec(
    AtomicNumber Z, RadialQuad rq, AtomicGridSizeDefault s) {
  auto [rsz, asz] = default_grid_size(Z, rq, s);
  return UnprunedAtomicGridSpecification{rq, rsz, default_radial_scaling_factor(rq, Z), asz};
}

atomic_grid_spec_map MolGridFactory::create_default_gridmap(const Molecule& mol,
                                                            PruningScheme ps, RadialQuad rq,
                                                            AtomicGridSizeDefault s) {
  // GauXC builds every element's specification through create_pruned_spec
  // (include/gauxc/molgrid/defaults.hpp:43-64); Unpruned stays one region.
  atomic_grid_spec_map m;
  for (const auto& a : mol)
    if (!m.count(a.Z)) {
      auto unp = create_default_unpruned_grid_spec(a.Z, rq, s);
      if (ps == PruningScheme::Unpruned) m.emplace(a.Z, unp);
      else m.emplace(a.Z, create_pruned_spec(ps, unp));
    }
  return m;
}

MolGrid::MolGrid(const atomic_grid_spec_map& specs, BatchSize bsz) {
  for (const auto& [Z, sp] : specs) grids_.emplace(Z, AtomicGridFactory::generate_grid(sp, bsz));
}

MolGrid MolGridFactory::create_default_molgrid(const Molecule& mol, PruningScheme ps,
                                               BatchSize bsz, RadialQuad rq,
                                               AtomicGridSizeDefault s) {
  return MolGrid(create_default_gridmap(mol, ps, rq, s), bsz);
}

const Grid& MolGrid::get_grid(AtomicNumber Z) const {
  auto it = grids_.find(Z);
  if (it == grids_.end()) GAUXC_GENERIC_EXCEPTION("No grid for Z=" + std::to_string(Z.get()));
  return it->second;
}

size_t MolGrid::max_nbatches() const {
  size_t m = 0;
  for (const auto& [Z, g] : grids_) m = std::max(m, g.batcher().nbatches());
  return m;
}

int64_t MolGrid::max_batch_size() const {
  int64_t m = 0;
  for (const auto& [Z, g] : grids_) m = std::max<int64_t>(m, g.max_batch_size());
  return m;
}

}  // namespace GauXC
