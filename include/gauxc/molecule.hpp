// Atoms, molecules and per-molecule geometric metadata.
// GauXC counterparts: include/gauxc/atom.hpp, molecule.hpp, molmeta.hpp and
// src/molmeta.cxx:16-60 (inter-atomic distance matrix and nearest-neighbour
// distance per atom, both used by the SSF partition).
#pragma once
#include <gauxc/named_type.hpp>
#include <gauxc/exceptions.hpp>
#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

namespace GauXC {

struct Atom {
  AtomicNumber Z;
  double x = 0., y = 0., z = 0.;
  Atom() = default;
  Atom(AtomicNumber _Z, double _x, double _y, double _z) : Z(_Z), x(_x), y(_y), z(_z) {}
  bool operator==(const Atom& o) const { return Z == o.Z && x == o.x && y == o.y && z == o.z; }
};

class Molecule : public std::vector<Atom> {
 public:
  using std::vector<Atom>::vector;
  size_t natoms() const { return size(); }
  AtomicNumber maxZ() const {
    int64_t m = 0;
    for (const auto& a : *this) m = std::max(m, a.Z.get());
    return AtomicNumber(m);
  }
};

/// Geometry-derived data shared by the load balancer and the SSF weights.
/// rab() is natoms x natoms, symmetric, column-major (rab[i + j*natoms]).
class MolMeta {
  size_t natoms_ = 0;
  std::vector<double> rab_;
  std::vector<double> dist_nearest_;
  size_t sum_atomic_charges_ = 0;

 public:
  MolMeta() = default;
  explicit MolMeta(const Molecule& mol) : natoms_(mol.natoms()) {
    rab_.assign(natoms_ * natoms_, 0.);
    for (size_t j = 0; j < natoms_; ++j)
      for (size_t i = j + 1; i < natoms_; ++i) {
        const double dx = mol[i].x - mol[j].x, dy = mol[i].y - mol[j].y,
                     dz = mol[i].z - mol[j].z;
        const double d = std::sqrt(dx * dx + dy * dy + dz * dz);
        rab_[i + j * natoms_] = d;
        rab_[j + i * natoms_] = d;
      }
    dist_nearest_.assign(natoms_, std::numeric_limits<double>::infinity());
    for (size_t i = 0; i < natoms_; ++i)
      for (size_t j = 0; j < natoms_; ++j)
        if (i != j) dist_nearest_[i] = std::min(dist_nearest_[i], rab_[j + i * natoms_]);
    for (const auto& a : mol) sum_atomic_charges_ += static_cast<size_t>(a.Z.get());
  }
  size_t natoms() const { return natoms_; }
  const std::vector<double>& rab() const { return rab_; }
  const std::vector<double>& dist_nearest() const { return dist_nearest_; }
  size_t sum_atomic_charges() const { return sum_atomic_charges_; }
};

}  // namespace GauXC
