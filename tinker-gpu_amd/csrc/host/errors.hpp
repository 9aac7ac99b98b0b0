// Error plumbing for the C ABI: every entry point runs its body through
// guarded(), which turns C++ exceptions into status codes and records the
// message for t9_last_error().  The reference throws instead
// (include/tool/error.h:38-74); the host C++ above this ABI can map the codes
// back to exceptions.
#pragma once
#include <stdexcept>
#include <string>

namespace t9 {

struct ArgError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct HipError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct NotConverged : std::runtime_error {
  using std::runtime_error::runtime_error;
};
struct ListOverflow : std::runtime_error {
  using std::runtime_error::runtime_error;
};

void set_last_error(const std::string& msg);

template <class F>
int guarded(F&& f, int generic_code) {
  try {
    return f();
  } catch (const ArgError& e) {
    set_last_error(e.what());
    return 4;
  } catch (const HipError& e) {
    set_last_error(e.what());
    return 1;
  } catch (const NotConverged& e) {
    set_last_error(e.what());
    return 2;
  } catch (const ListOverflow& e) {
    set_last_error(e.what());
    return 3;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return generic_code;
  } catch (...) {
    set_last_error("unknown error");
    return generic_code;
  }
}

}  // namespace t9
