// Switching-function evaluation shared by the host (construction: the stretch
// constants need s(0) and s(D_MAX)) and by the device kernels.
//
// Semantics follow src/tools/SwitchingFunction.cpp:
//   * baseSwitch::calculate (:133-184) with applystretch (:124-130):
//       r > D_MAX -> (0,0); rdist=(r-D_0)/R_0 <= 0 -> (stretch+shift, 0);
//       else f(rdist)*stretch+shift and dfunc = f'(rdist)*stretch/R_0/r
//   * squared fast paths: fixedRational<N>::calculateSqr (:203-215), fast
//     rational (:283-298), fastgaussian (:411-431)
//   * nativeqSwitch::calculate (:521-548)
// dfunc is (1/r) ds/dr, the quantity CoordinationBase multiplies by the
// distance vector.
#pragma once
#include <cmath>
#include <cfloat>
#include "plumed2_amd.h"

#if defined(__HIPCC__)
#define P2A_HD __host__ __device__ inline
#else
#define P2A_HD inline
#endif

namespace p2a {

// the reference's host build forms every product and sum on its own: no
// fused multiply-add contraction in this file, so that cut decisions and
// values round as the reference's do
#if defined(__clang__)
#pragma clang fp contract(off)
#endif

P2A_HD double ipow(double base, int exp) {
  // square-and-multiply, the evaluation order of Tools::fastpow (Tools.h:582-597)
  if (exp < 0) { exp = -exp; base = 1.0 / base; }
  double result = 1.0;
  while (exp) {
    if (exp & 1) result *= base;
    exp >>= 1;
    base *= base;
  }
  return result;
}

P2A_HD void rational_core(const p2a_switch& s, double rdist, double secDev, int N, int M, double dfunc0,
                          double result0, double& res, double& df) {
  const double moreThanOne = 1.0 + 5.0e10 * DBL_EPSILON;
  const double lessThanOne = 1.0 - 5.0e10 * DBL_EPSILON;
  if (s.simplified) {
    double rN = ipow(rdist, N - 1);
    double r = 1.0 / (1.0 + rN * rdist);
    res = r;
    df = -N * rN * r * r;
  } else if (!((rdist > lessThanOne) && (rdist < moreThanOne))) {
    double rN = ipow(rdist, N - 1);
    double rM = ipow(rdist, M - 1);
    double num = 1.0 - rN * rdist;
    double iden = 1.0 / (1.0 - rM * rdist);
    res = num * iden;
    df = ((M * res * rM) - (N * rN)) * iden;
  } else {
    // second-order expansion around rdist=1 (SwitchingFunction.cpp:271-276)
    double x = rdist - 1.0;
    res = result0 + x * (dfunc0 + 0.5 * x * secDev);
    df = dfunc0 + x * secDev;
  }
}

P2A_HD void base_function(const p2a_switch& s, double rdist, double& f, double& df) {
  switch (s.type) {
  case P2A_SW_RATIONAL_FIXED: {
    double rN = ipow(rdist, s.nn - 1);
    double r = 1.0 / (1.0 + rN * rdist);
    f = r;
    df = -s.nn * rN * r * r;
    return;
  }
  case P2A_SW_RATIONAL:
    rational_core(s, rdist, s.preSecDev, s.nn, s.mm, s.preDfunc, s.preRes, f, df);
    return;
  case P2A_SW_EXP: {
    double r = exp(-rdist);
    f = r;
    df = -r;
    return;
  }
  case P2A_SW_GAUSSIAN:
  case P2A_SW_FASTGAUSSIAN: {
    double r = exp(-0.5 * rdist * rdist);
    f = r;
    df = -rdist * r;
    return;
  }
  case P2A_SW_SMAP: {
    double sx = s.c * ipow(rdist, s.a);
    double r = pow(1.0 + sx, s.d);
    f = r;
    df = -s.b * sx / rdist * r / (1.0 + sx);
    return;
  }
  case P2A_SW_CUBIC: {
    double t1 = rdist - 1.0, t2 = 1.0 + 2.0 * rdist;
    f = t1 * t1 * t2;
    df = 2 * t1 * t2 + 2 * t1 * t1;
    return;
  }
  case P2A_SW_TANH: {
    double t = tanh(rdist);
    f = 1.0 - t;
    df = t * t - 1.0;
    return;
  }
  case P2A_SW_COSINUS: {
    f = 0.0;
    df = 0.0;
    if (rdist <= 1.0) {
      const double pi = 3.141592653589793238462643383279502884197169399375105820974944592307;
      double rp = rdist * pi;
      f = 0.5 * (cos(rp) + 1.0);
      df = -0.5 * pi * sin(rp);
    }
    return;
  }
  default:
    f = 0.0;
    df = 0.0;
    return;
  }
}

// s(r) and dfunc from the distance (non-squared entry point)
P2A_HD double switch_calc(const p2a_switch& s, double distance, double& dfunc) {
  if (s.type == P2A_SW_Q) {
    double res = 0.0, df = 0.0;
    if (distance <= s.dmax) {
      res = 1.0;
      if (distance > s.d0) {
        double rdist = s.beta * (distance - s.lambda * s.ref);
        double e = exp(rdist);
        res = 1.0 / (1.0 + e);
        df = -s.beta / (e + 2.0 + 1.0 / e) / distance;
        df *= s.stretch;
      }
      res = res * s.stretch + s.shift;
    }
    dfunc = df;
    return res;
  }
  double rdist = (distance - s.d0) * s.invr0;
  if (distance > s.dmax) {
    dfunc = 0.0;
    return 0.0;
  }
  if (!(rdist > 0.0)) {
    dfunc = 0.0;
    return s.stretch + s.shift;
  }
  double f, df;
  base_function(s, rdist, f, df);
  f = f * s.stretch + s.shift;
  df *= s.stretch;
  df *= s.invr0;
  df /= distance;
  dfunc = df;
  return f;
}

// s and dfunc from the squared distance, FP64 (every switch type)
P2A_HD double switch_calc_sqr(const p2a_switch& s, double d2, double& dfunc) {
  if (s.type == P2A_SW_RATIONAL_FIXED) {
    double result = 0.0, df = 0.0;
    if (d2 <= s.dmax_2) {
      double rdist = d2 * s.invr0_2;
      int P = s.nn / 2;
      double rN = ipow(rdist, P - 1);
      result = 1.0 / (1.0 + rN * rdist);
      df = -P * rN * result * result;
      df *= 2 * s.invr0_2;
      result = result * s.stretch + s.shift;
      df *= s.stretch;
    }
    dfunc = df;
    return result;
  }
  if (s.type == P2A_SW_RATIONAL && s.fast) {
    double result = 0.0, df = 0.0;
    if (d2 <= s.dmax_2) {
      double rdist = d2 * s.invr0_2;
      rational_core(s, rdist, s.preSecDevF, s.nnf, s.mmf, s.preDfuncF, s.preRes, result, df);
      df *= 2 * s.invr0_2;
      result = result * s.stretch + s.shift;
      df *= s.stretch;
    }
    dfunc = df;
    return result;
  }
  if (s.type == P2A_SW_FASTGAUSSIAN) {
    double result = 0.0, df = 0.0;
    if (d2 < s.dmax_2) {
      result = 1.0;
      if (d2 > 0.0) {
        result = exp(-0.5 * d2);
        df = -result;
        result = result * s.stretch + s.shift;
        df *= s.stretch;
      }
    }
    dfunc = df;
    return result;
  }
  return switch_calc(s, sqrt(d2), dfunc);
}

// whether a squared distance is inside the support of the switch
// (the inclusive/exclusive cut of each calculateSqr)
P2A_HD bool switch_inside(const p2a_switch& s, double d2) {
  if (s.type == P2A_SW_FASTGAUSSIAN) return d2 < s.dmax_2;
  if (s.type == P2A_SW_RATIONAL_FIXED || (s.type == P2A_SW_RATIONAL && s.fast)) return d2 <= s.dmax_2;
  return sqrt(d2) <= s.dmax;
}

#if defined(__clang__)
#pragma clang fp contract(fast)
#endif

}  // namespace p2a
