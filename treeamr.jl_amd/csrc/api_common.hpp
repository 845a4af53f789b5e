// Shared plumbing of the C ABI: the thread-local error message and the
// exception-to-status guard.
#pragma once

#include <exception>
#include <new>
#include <string>

#include "forest.hpp"
#include "tah.h"

namespace tah {

void set_last_error(const std::string& m);

template <class F>
tah_status guard(F&& f) {
  try {
    f();
    return TAH_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return tah_status(e.code);
  } catch (const std::bad_alloc&) {
    set_last_error("out of host memory");
    return TAH_ENOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return TAH_EINVAL;
  }
}

}  // namespace tah
