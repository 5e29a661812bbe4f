// Basis sets and the shell -> basis-function index map.
// GauXC counterparts: include/gauxc/basisset.hpp (nbf = sum of shell sizes,
// :83-88) and include/gauxc/basisset_map.hpp:45-80 (first AO of each shell,
// AO range, owning atom).
#pragma once
#include <gauxc/shell.hpp>
#include <gauxc/molecule.hpp>
#include <utility>
#include <vector>

namespace GauXC {

template <typename F>
struct BasisSet : public std::vector<Shell<F>> {
  using std::vector<Shell<F>>::vector;
  int32_t nshells() const { return static_cast<int32_t>(this->size()); }
  int32_t nbf() const {
    int32_t n = 0;
    for (const auto& s : *this) n += s.size();
    return n;
  }
};

class BasisSetMap {
  std::vector<int32_t> first_ao_;
  std::vector<int32_t> shell_size_;
  std::vector<int32_t> center_;
  int32_t nbf_ = 0;

 public:
  template <typename F>
  BasisSetMap(const BasisSet<F>& basis, const Molecule& mol) {
    first_ao_.reserve(basis.size());
    for (const auto& sh : basis) {
      first_ao_.push_back(nbf_);
      shell_size_.push_back(sh.size());
      nbf_ += sh.size();
      int32_t c = -1;
      for (size_t a = 0; a < mol.size(); ++a)
        if (mol[a].x == sh.O()[0] && mol[a].y == sh.O()[1] && mol[a].z == sh.O()[2]) {
          c = static_cast<int32_t>(a);
          break;
        }
      center_.push_back(c);
    }
  }
  int32_t nshells() const { return static_cast<int32_t>(first_ao_.size()); }
  int32_t nbf() const { return nbf_; }
  int32_t shell_to_first_ao(int32_t i) const { return first_ao_.at(i); }
  int32_t shell_size(int32_t i) const { return shell_size_.at(i); }
  std::pair<int32_t, int32_t> shell_to_ao_range(int32_t i) const {
    return {first_ao_.at(i), first_ao_.at(i) + shell_size_.at(i)};
  }
  int32_t shell_to_center(int32_t i) const { return center_.at(i); }
};

}  // namespace GauXC
