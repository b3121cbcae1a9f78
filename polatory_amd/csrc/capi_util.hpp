// Shared helpers of the extern "C" layer: exception-to-status guard and
// argument conversion.
#pragma once

#include <new>
#include <string>

#include "../../include/polatory_amd.h"
#include "common.hpp"

namespace pam {

std::string& last_error_ref();  // per thread

template <class F>
int guard(F&& f) {
  try {
    f();
    return PAM_OK;
  } catch (const Error& e) {
    last_error_ref() = e.what();
    return e.code;
  } catch (const std::bad_alloc&) {
    last_error_ref() = "host allocation failed";
    return PAM_ERR_OUT_OF_MEMORY;
  } catch (const std::exception& e) {
    last_error_ref() = e.what();
    return PAM_ERR_RUNTIME;
  } catch (...) {
    last_error_ref() = "unknown error";
    return PAM_ERR_RUNTIME;
  }
}

RbfDesc to_desc(const pam_rbf* r);

}  // namespace pam
