// Exception type and throw macro of the GauXC-compatible surface
// (GauXC: include/gauxc/exceptions.hpp:40,79-87).  Messages keep GauXC's
// wording where callers may match on them ("Weights Have Not Been Modified",
// "PIMPL NOT INITIALIZED", ...).
#pragma once
#include <stdexcept>
#include <string>

namespace GauXC {

class generic_gauxc_exception : public std::runtime_error {
  std::string file_, function_;
  int line_;
 public:
  generic_gauxc_exception(const std::string& file, int line, const std::string& function,
                          const std::string& msg)
      : std::runtime_error(msg), file_(file), function_(function), line_(line) {}
  const std::string& file() const { return file_; }
  const std::string& function() const { return function_; }
  int line() const { return line_; }
};

}  // namespace GauXC

#define GAUXC_GENERIC_EXCEPTION(MSG) \
  throw ::GauXC::generic_gauxc_exception(__FILE__, __LINE__, __func__, MSG)

#define GAUXC_PIMPL_NOT_INITIALIZED() GAUXC_GENERIC_EXCEPTION("PIMPL NOT INITIALIZED")
