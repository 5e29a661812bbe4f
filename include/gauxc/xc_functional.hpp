// Functional handle of the GauXC-compatible surface.
//
// GauXC types its functional as ExchCXX::XCFunctional
// (include/gauxc/types.hpp:22) and evaluates it through ExchCXX's builtin
// kernels.  ExchCXX is a third-party library that this build does not use:
// the functionals on the quadrature path (Slater + VWN5 = SVWN5, PBE, and
// the XC part of PBE0) are evaluated by this build's own device kernels.
// This header keeps the construction syntax GauXC callers use,
//   functional_type(ExchCXX::Backend::builtin, ExchCXX::Functional::PBE0,
//                   ExchCXX::Spin::Unpolarized),
// and the queries the integrator needs (is_lda/is_gga/is_mgga/
// is_polarized/needs_laplacian).
#pragma once
#include <gauxc/exceptions.hpp>
#include <string>

namespace ExchCXX {

enum class Backend { builtin, libxc };
enum class Spin { Polarized, Unpolarized };
enum class Functional { SVWN5, PBE, PBE0 };

class XCFunctional {
  Functional f_ = Functional::SVWN5;
  Spin spin_ = Spin::Unpolarized;
  bool init_ = false;

 public:
  XCFunctional() = default;
  XCFunctional(Backend, Functional f, Spin s) : f_(f), spin_(s), init_(true) {}
  bool is_lda() const { return f_ == Functional::SVWN5; }
  bool is_gga() const { return f_ != Functional::SVWN5; }
  bool is_mgga() const { return false; }
  bool needs_laplacian() const { return false; }
  bool is_polarized() const { return spin_ == Spin::Polarized; }
  bool is_hyb() const { return f_ == Functional::PBE0; }
  double hyb_exx() const { return f_ == Functional::PBE0 ? 0.25 : 0.; }
  bool initialized() const { return init_; }
  Functional kind() const { return f_; }
  std::string name() const {
    switch (f_) {
      case Functional::SVWN5: return "SVWN5";
      case Functional::PBE: return "PBE";
      case Functional::PBE0: return "PBE0";
    }
    return "?";
  }
};

}  // namespace ExchCXX

namespace GauXC {
using functional_type = ExchCXX::XCFunctional;
/// Parse a functional name ("SVWN5", "PBE", "PBE0"; case-insensitive).
functional_type functional_from_string(const std::string& name, bool polarized = false);
}  // namespace GauXC
