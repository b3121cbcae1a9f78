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

// Square root for the FP64 near field: the hardware reciprocal square root
// (v_rsq_f64, accurate to about 2^-23 relative) refined by one Goldschmidt
// step and a final Newton correction of the remainder (Markstein's method),
// within a few ulp of the correctly rounded sqrt in 8 instructions instead of
// the compiler's IEEE-rounded expansion. 0 for x = 0 (coincident points); x is
// a sum of squares, never negative.
__device__ __forceinline__ double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double s = x * y, h = 0.5 * y;
  const double r = fma(-s, h, 0.5);
  s = fma(s, r, s);
  h = fma(h, r, h);
  const double d = fma(-s, s, x);
  s = fma(d, h, s);
  return x > 0.0 ? s : 0.0;
}

// phi without its factor a, for loops that fold a into the weights:
// r2c = r^2 + b for bh3 / th3 (the caller seeds its sum of squares with b),
// r2c = r^2 for exp.
template <class T, int KIND>
__device__ __forceinline__ T phi_shape(T r2c, T b) {
  if constexpr (sizeof(T) == 8) {
    if constexpr (KIND == kBh3) {
      return sqrt_nr(r2c);
    } else if constexpr (KIND == kTh3) {
      double rho = sqrt_nr(r2c);
      return rho * rho * rho;
    } else {
      return exp(b * sqrt_nr(r2c));
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

// Gradient in physical units from differences d in a local unit h:
//   grad phi = ag * grad_shape(r2c) * d,  r2c = |d|^2 + phi_seed(b),
// with (ag, b) from grad_consts(rbf, h) (rbf/polyharmonic_odd.hpp:41-53,
// rbf/cov_exponential.hpp:42-50):
//   bh3: ag = -slope,          grad_shape = 1 / sqrt(r2c)  (0 where r2c = 0)
//   th3: ag = 3 slope h^2,     grad_shape = sqrt(r2c)
//   exp: ag = -3 psill / range, grad_shape = exp(b r) / r, r = sqrt(r2c)
// (at r = 0 the exponential's gradient is undefined, as in the reference).
inline PhiConsts grad_consts(const RbfDesc& rbf, double h) {
  const PhiConsts pc = phi_consts(rbf, h);
  switch (rbf.kind) {
    case kBh3: return {pc.a / h, pc.b};
    case kTh3: return {3.0 * pc.a / h, pc.b};
    case kExp: return {pc.a * pc.b / h, pc.b};
  }
  throw_invalid("unknown rbf kind");
}

template <class T, int KIND>
__device__ __forceinline__ T grad_shape(T r2c, T b) {
  if constexpr (sizeof(T) == 8) {
    if constexpr (KIND == kBh3) {
      return r2c > 0.0 ? 1.0 / sqrt(r2c) : 0.0;
    } else if constexpr (KIND == kTh3) {
      return sqrt(r2c);
    } else {
      const double r = sqrt(r2c);
      return exp(b * r) / r;
    }
  } else {
    if constexpr (KIND == kBh3) {
      return r2c > 0.0f ? __builtin_amdgcn_rsqf(r2c) : 0.0f;
    } else if constexpr (KIND == kTh3) {
      return __builtin_amdgcn_sqrtf(r2c);
    } else {
      const float rs = __builtin_amdgcn_rsqf(r2c);
      return __expf(b * (r2c * rs)) * rs;
    }
  }
}

// Hessian in physical units from local differences u (unit h):
//   H a = (ag / h) grad_shape(r2c) (a + hess_t(r2c) u (u . a))
// (rbf/polyharmonic_odd.hpp:56-72, rbf/cov_exponential.hpp:52-62):
//   bh3 / th3: hess_t = (K - 2) / r2c   (0 where r2c = 0, where the reference returns 0)
//   exp:       hess_t = -(1 - b r) / r^2, r = sqrt(r2c)
template <class T, int KIND>
__device__ __forceinline__ T hess_t(T r2c, T b) {
  if constexpr (KIND == kBh3) {
    return r2c > T(0) ? T(-1) / r2c : T(0);
  } else if constexpr (KIND == kTh3) {
    return r2c > T(0) ? T(1) / r2c : T(0);
  } else {
    const T r = sqrt(r2c);
    return -(T(1) - b * r) / r2c;
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

// Value, gradient and Hessian of the isotropic kernel at a difference u that
// is already transformed by the anisotropy (rbf/polyharmonic_odd.hpp:32-72,
// rbf/cov_exponential.hpp:33-62). Polyharmonic of degree K (bh3: K = 1, sign
// -1; th3: K = 3, sign +1), rho = sqrt(|u|^2 + c^2):
//   g = sign K slope rho^(K-2) u,  H = sign K slope rho^(K-2) (I + (K-2) u u^T / rho^2),
//   both zero at rho = 0.
// Exponential, r = |u|:
//   g = -3 psill e^(-3 r / range) / (range r) u,
//   H = (same factor) (I - (1/r^2 + 3 / (range r)) u u^T)   (no guard at r = 0).
template <int KIND>
__host__ __device__ inline double rbf_value_iso(double p0, double p1, const double u[3]) {
  const double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  if constexpr (KIND == kBh3) {
    return -p0 * sqrt(q + p1 * p1);
  } else if constexpr (KIND == kTh3) {
    const double rho = sqrt(q + p1 * p1);
    return p0 * (rho * rho * rho);
  } else {
    return p0 * exp(-3.0 * (sqrt(q) / p1));
  }
}

template <int KIND>
__host__ __device__ inline double rbf_grad_factor(double p0, double p1, const double u[3], double* t) {
  // Returns the factor f with g = f u; t receives the coefficient of u u^T in
  // H / f - I.
  const double q = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  if constexpr (KIND == kBh3 || KIND == kTh3) {
    constexpr int K = KIND == kBh3 ? 1 : 3;
    constexpr double sign = KIND == kBh3 ? -1.0 : 1.0;
    const double rho2 = q + p1 * p1;
    const double rho = sqrt(rho2);
    if (rho == 0.0) {
      *t = 0.0;
      return 0.0;
    }
    *t = (K - 2) / rho2;
    const double rk = K == 1 ? 1.0 / rho : rho;  // rho^(K-2)
    return sign * K * p0 * rk;
  } else {
    const double r = sqrt(q);
    *t = -(1.0 / (r * r) + 3.0 / (p1 * r));
    return -3.0 * p0 * exp(-3.0 * (r / p1)) / (p1 * r);
  }
}

template <int KIND>
__host__ __device__ inline void rbf_grad_iso(double p0, double p1, const double u[3], double g[3]) {
  double t;
  const double f = rbf_grad_factor<KIND>(p0, p1, u, &t);
  for (int k = 0; k < 3; k++) g[k] = f * u[k];
}

template <int KIND>
__host__ __device__ inline void rbf_hess_iso(double p0, double p1, const double u[3], double h[9]) {
  double t;
  const double f = rbf_grad_factor<KIND>(p0, p1, u, &t);
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) h[3 * i + j] = f * ((i == j ? 1.0 : 0.0) + t * u[i] * u[j]);
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
