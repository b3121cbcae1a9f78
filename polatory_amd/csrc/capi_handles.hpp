// Opaque handle types shared by the extern "C" files, and the conversion of
// the FMM configuration struct.
#pragma once

#include <memory>

#include "capi_util.hpp"
#include "interpolant.hpp"

struct pam_model {
  pam::Model m;
};

struct pam_interpolant {
  std::unique_ptr<pam::Interpolant> impl;
};

namespace pam {

inline FmmConfig to_cfg(const pam_fmm_config& c) {
  FmmConfig r;
  r.precision = c.precision;
  r.order = c.order;
  r.leaf_level = c.leaf_level;
  r.mode = c.mode;
  r.svd_tol = c.svd_tolerance;
  return r;
}

inline pam_fmm_config from_cfg(const FmmConfig& c) {
  pam_fmm_config r;
  r.precision = c.precision;
  r.order = c.order;
  r.leaf_level = c.leaf_level;
  r.mode = c.mode;
  r.svd_tolerance = c.svd_tol;
  return r;
}

}  // namespace pam
