// Strong scalar typedefs used by the GauXC-compatible C++ surface.
// Mirrors the role of include/gauxc/named_type.hpp in GauXC (a tagged
// wrapper with get()), so user code such as AtomicNumber(8).get() compiles
// unchanged against this library.
#pragma once
#include <cstdint>
#include <functional>

namespace GauXC {
namespace detail {

template <typename T, typename Tag>
class NamedType {
  T v_{};
 public:
  using value_type = T;
  constexpr NamedType() = default;
  constexpr explicit NamedType(T v) : v_(v) {}
  constexpr T get() const { return v_; }
  constexpr bool operator==(const NamedType& o) const { return v_ == o.v_; }
  constexpr bool operator!=(const NamedType& o) const { return v_ != o.v_; }
  constexpr bool operator<(const NamedType& o) const { return v_ < o.v_; }
};

}  // namespace detail

using AtomicNumber    = detail::NamedType<int64_t, struct AtomicNumberTag>;
using RadialSize      = detail::NamedType<int64_t, struct RadialSizeTag>;
using AngularSize     = detail::NamedType<int64_t, struct AngularSizeTag>;
using BatchSize       = detail::NamedType<int64_t, struct BatchSizeTag>;
using RadialScale     = detail::NamedType<double,  struct RadialScaleTag>;
using PrimSize        = detail::NamedType<int32_t, struct PrimSizeTag>;
using AngularMomentum = detail::NamedType<int32_t, struct AngularMomentumTag>;
using SphericalType   = detail::NamedType<int32_t, struct SphericalTypeTag>;

}  // namespace GauXC

namespace std {
template <typename T, typename Tag>
struct hash<GauXC::detail::NamedType<T, Tag>> {
  size_t operator()(const GauXC::detail::NamedType<T, Tag>& v) const noexcept {
    return std::hash<T>()(v.get());
  }
};
}  // namespace std
