// Contracted Gaussian shell.
//
// GauXC counterpart: include/gauxc/shell.hpp:48-229 (same constructor,
// accessors and 32-primitive limit) and include/gauxc/util/gau_rad_eval.hpp
// (screening radius).  Two pieces of arithmetic must agree with GauXC for
// screening and normalisation to agree:
//  * normalize(): each primitive is scaled to a unit-norm x^l Gaussian,
//      N_p = sqrt( 2^l (2 a)^(l+3/2) / (pi^(3/2) (2l-1)!!) ),
//    then the contraction is rescaled so that its x^l component has unit
//    norm, using  <g_a|g_b> = (2l-1)!! pi^(3/2) / (2^l (a+b)^(l+3/2)).
//  * compute_shell_cutoff(): the radius where the |c|-weighted radial
//    function r^l sum_p |c_p| exp(-a_p r^2) drops to the tolerance, found
//    by GauXC's procedure (gau_rad_eval.hpp:31-71): start from the largest
//    single-primitive estimate sqrt((-ln tol + ln(a)/2)/a), then walk in
//    steps of 0.01 bohr to the crossing (right if still above tol,
//    otherwise left and back one step).  The walk is kept identical because
//    the radius decides which shells a grid batch keeps.
#pragma once
#include <gauxc/named_type.hpp>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdint>
#include <limits>
#include <ostream>

namespace GauXC {

namespace detail {
static constexpr size_t shell_nprim_max = 32;
static constexpr double default_shell_tolerance = 1e-10;

inline double double_factorial_odd(int n) {  // n!! for odd n, (-1)!! = 1
  double r = 1.;
  for (int k = n; k > 1; k -= 2) r *= k;
  return r;
}

template <typename T>
T radial_abs_eval(int32_t l, int32_t nprim, const T* alpha, const T* coeff, T r) {
  T s = 0.;
  const T r2 = r * r;
  for (int32_t i = 0; i < nprim; ++i) s += std::abs(coeff[i]) * std::exp(-alpha[i] * r2);
  return std::pow(r, l) * s;
}

template <typename T>
T shell_screening_radius(int32_t l, int32_t nprim, const T* alpha, const T* coeff, T tol) {
  if (tol <= 0.) return std::numeric_limits<T>::infinity();
  const double mlog_tol = -std::log(tol);
  double r = 0.;
  for (int32_t i = 0; i < nprim; ++i)
    r = std::max(r, std::sqrt((mlog_tol + 0.5 * std::log(alpha[i])) / alpha[i]));
  constexpr double step = 0.01;
  double v = radial_abs_eval<T>(l, nprim, alpha, coeff, r);
  if (v > tol) {
    while (v > tol) { r += step; v = radial_abs_eval<T>(l, nprim, alpha, coeff, r); }
  } else {
    while (v < tol) { r -= step; v = radial_abs_eval<T>(l, nprim, alpha, coeff, r); }
    r += step;
  }
  return r;
}
}  // namespace detail

template <typename F>
class alignas(256) Shell {
 public:
  using prim_array = std::array<F, detail::shell_nprim_max>;
  using cart_array = std::array<double, 3>;

 private:
  prim_array alpha_{};
  prim_array coeff_{};
  cart_array O_{};
  int32_t nprim_ = 0;
  int32_t l_ = 0;
  int32_t pure_ = 0;
  double cutoff_radius_ = 0.;
  double shell_tolerance_ = detail::default_shell_tolerance;

  void normalize() {
    const double pi32 = std::pow(M_PI, 1.5);
    const double dfac = detail::double_factorial_odd(2 * l_ - 1);
    const double two_l = std::ldexp(1.0, l_);
    for (int32_t i = 0; i < nprim_; ++i) {
      if (alpha_[i] == 0.) continue;
      const double ta = 2. * alpha_[i];
      coeff_[i] *= std::sqrt(two_l * std::pow(ta, l_ + 1.5) / (pi32 * dfac));
    }
    double ovl = 0.;
    for (int32_t i = 0; i < nprim_; ++i)
      for (int32_t j = 0; j < nprim_; ++j) {
        const double g = alpha_[i] + alpha_[j];
        ovl += coeff_[i] * coeff_[j] * dfac * pi32 / (two_l * std::pow(g, l_ + 1.5));
      }
    const double s = 1. / std::sqrt(ovl);
    for (int32_t i = 0; i < nprim_; ++i) coeff_[i] *= s;
  }

  void compute_shell_cutoff() {
    cutoff_radius_ = detail::shell_screening_radius<F>(l_, nprim_, alpha_.data(), coeff_.data(),
                                                      shell_tolerance_);
  }

 public:
  Shell() = default;
  Shell(PrimSize nprim, AngularMomentum l, SphericalType pure, prim_array alpha,
        prim_array coeff, cart_array O, bool _normalize = true)
      : alpha_(alpha), coeff_(coeff), O_(O), nprim_(nprim.get()), l_(l.get()),
        pure_(pure.get()) {
    assert(nprim_ >= 0 && nprim_ <= (int32_t)detail::shell_nprim_max);
    if (_normalize) normalize();
    compute_shell_cutoff();
  }

  void set_shell_tolerance(double tol) {
    if (tol != shell_tolerance_) {
      shell_tolerance_ = tol;
      compute_shell_cutoff();
    }
  }
  double shell_tolerance() const { return shell_tolerance_; }

  int32_t nprim() const { return nprim_; }
  int32_t l() const { return l_; }
  int32_t pure() const { return pure_; }
  void set_pure(bool p) { pure_ = p; }
  const F* alpha_data() const { return alpha_.data(); }
  const F* coeff_data() const { return coeff_.data(); }
  const double* O_data() const { return O_.data(); }
  F* alpha_data() { return alpha_.data(); }
  F* coeff_data() { return coeff_.data(); }
  double* O_data() { return O_.data(); }
  const prim_array& alpha() const { return alpha_; }
  const prim_array& coeff() const { return coeff_; }
  const cart_array& O() const { return O_; }
  prim_array& alpha() { return alpha_; }
  prim_array& coeff() { return coeff_; }
  cart_array& O() { return O_; }
  double cutoff_radius() const { return cutoff_radius_; }
  int32_t cart_size() const { return (l_ + 1) * (l_ + 2) / 2; }
  int32_t pure_size() const { return 2 * l_ + 1; }
  int32_t size() const { return pure_ ? pure_size() : cart_size(); }

  bool operator==(const Shell& o) const {
    if (nprim_ != o.nprim_ || l_ != o.l_ || pure_ != o.pure_ || O_ != o.O_) return false;
    for (int32_t i = 0; i < nprim_; ++i)
      if (alpha_[i] != o.alpha_[i] || coeff_[i] != o.coeff_[i]) return false;
    return true;
  }
};

template <typename F>
std::ostream& operator<<(std::ostream& os, const Shell<F>& sh) {
  os << "Shell(l=" << sh.l() << ", pure=" << sh.pure() << ", nprim=" << sh.nprim() << ", O={"
     << sh.O()[0] << "," << sh.O()[1] << "," << sh.O()[2] << "})";
  return os;
}

}  // namespace GauXC
