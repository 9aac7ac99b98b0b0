#include "errors.hpp"

#include "../../../include/t9hip.h"

namespace t9 {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }
}  // namespace t9

extern "C" const char* t9_last_error(void) { return t9::last_error_cstr(); }
