// Device-side RBF evaluation phi(r^2) with the constants folded per call site.
//
//   bh3  phi = -slope * sqrt(r^2 + c^2)        (rbf/polyharmonic_odd.hpp:32-43, kSign=-1)
//   th3  phi = +slope * (r^2 + c^2)^(3/2)      (rbf/polyharmonic_odd.hpp:120-136)
//   exp  phi = psill * exp(-3 r / range)       (rbf/cov_exponential.hpp:33-40)
//
// Call sites work in a local length unit h (r_phys = h * r_local); the folded
// constants (a, b) absorb h so that phi(r_local^2) = phi_phys(h r_local):
//   bh3: a = -slope*h,   b = (c/h)^2   -> a * sqrt(r2 + b)
//   th3: a =  slope*h^3, b = (c/h)^2   -> a * rho^3, rho = sqrt(r2 + b)
//   exp: a =  psill,     b = -3 h / range -> a * exp(b * sqrt(r2))
#pragma once

#include "common.hpp"

namespace pam {

struct PhiConsts {
  double a, b;
};

inline PhiConsts phi_consts(const RbfDesc& rbf, double h) {
  switch (rbf.kind) {
    case kBh3: return {-rbf.p0 * h, (rbf.p1 / h) * (rbf.p1 / h)};
    case kTh3: return {rbf.p0 * h * h * h, (rbf.p1 / h) * (rbf.p1 / h)};
    case kExp: return {rbf.p0, -3.0 * h / rbf.p1};
  }
  throw_invalid("unknown rbf kind");
}

template <int KIND>
__device__ __forceinline__ double phi_f64(double r2, double a, double b) {
  if constexpr (KIND == kBh3) {
    return a * sqrt(r2 + b);
  } else if constexpr (KIND == kTh3) {
    double rho = sqrt(r2 + b);
    return a * (rho * rho * rho);
  } else {
    return a * exp(b * sqrt(r2));
  }
}

// FP32 form: raw hardware square root (v_sqrt_f32, ~1 ulp) and exp2-based exp.
template <int KIND>
__device__ __forceinline__ float phi_f32(float r2, float a, float b) {
  if constexpr (KIND == kBh3) {
    return a * __builtin_amdgcn_sqrtf(r2 + b);
  } else if constexpr (KIND == kTh3) {
    float rho = __builtin_amdgcn_sqrtf(r2 + b);
    return a * (rho * rho * rho);
  } else {
    return a * __expf(b * __builtin_amdgcn_sqrtf(r2));
  }
}

// phi without its factor a, for loops that fold a into the weights:
// r2c = r^2 + b for bh3 / th3 (the caller seeds its sum of squares with b),
// r2c = r^2 for exp.
template <class T, int KIND>
__device__ __forceinline__ T phi_shape(T r2c, T b) {
  if constexpr (sizeof(T) == 8) {
    if constexpr (KIND == kBh3) {
      return sqrt(r2c);
    } else if constexpr (KIND == kTh3) {
      double rho = sqrt(r2c);
      return rho * rho * rho;
    } else {
      return exp(b * sqrt(r2c));
    }
  } else {
    if constexpr (KIND == kBh3) {
      return __builtin_amdgcn_sqrtf(r2c);
    } else if constexpr (KIND == kTh3) {
      float rho = __builtin_amdgcn_sqrtf(r2c);
      return rho * rho * rho;
    } else {
      return __expf(b * __builtin_amdgcn_sqrtf(r2c));
    }
  }
}

// Seed of the sum of squares that phi_shape expects.
template <class T, int KIND>
__device__ __forceinline__ T phi_seed(T b) {
  if constexpr (KIND == kExp) {
    return T(0);
  } else {
    return b;
  }
}

template <class T, int KIND>
__device__ __forceinline__ T phi_t(T r2, T a, T b) {
  if constexpr (sizeof(T) == 8) {
    return phi_f64<KIND>(r2, a, b);
  } else {
    return phi_f32<KIND>(r2, a, b);
  }
}

// Dispatch a templated functor on the runtime kind.
template <class F>
inline void dispatch_kind(int kind, F&& f) {
  switch (kind) {
    case kBh3: f(std::integral_constant<int, kBh3>{}); return;
    case kTh3: f(std::integral_constant<int, kTh3>{}); return;
    case kExp: f(std::integral_constant<int, kExp>{}); return;
  }
  throw_invalid("unknown rbf kind");
}

}  // namespace pam
