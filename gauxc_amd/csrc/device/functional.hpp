// Device evaluation of the unpolarised functionals on the quadrature path,
// in FP64.  Outputs follow the ExchCXX/libxc convention GauXC consumes
// (host_integrator_exc_vxc.hpp:445-450): energy per particle eps,
// vrho = d(rho eps)/d rho, vsigma = d(rho eps)/d sigma.
//
// Forms (written from the publications, constants as in libxc):
//  * Slater exchange: e = -(3/4)(3/pi)^(1/3) rho^(4/3).
//  * VWN correlation as GauXC's "SVWN5" evaluates it: the Vosko-Wilk-Nusair
//    (1980) fit to the RPA correlation energy, paramagnetic parameters
//    A=0.0310907, b=13.0720, c=42.7198, x0=-0.409286 (libxc LDA_C_VWN_RPA),
//    x = sqrt(rs).
//  * PBE exchange (Perdew, Burke, Ernzerhof 1996): kappa=0.804,
//    mu=0.2195149727645171, s^2 = sigma/(4 (3 pi^2)^(2/3) rho^(8/3)).
//  * PBE correlation: PW92 with libxc's modified parameters (a=0.0310907,
//    alpha1=0.21370, beta1..4 = 7.5957, 3.5876, 1.6382, 0.49294),
//    beta=0.06672455060314922, gamma=(1-ln 2)/pi^2,
//    t^2 = pi sigma / (16 (3 pi^2)^(1/3) rho^(7/3)).
// Thresholds: rho <= 1e-15 gives zero (1e-12 for PBE correlation); sigma is
// floored at (2.1544346900318838e-43)^2.
#pragma once
#include <hip/hip_runtime.h>

namespace xcq {

struct XCOut {
  double e;   // rho * eps
  double vr;  // vrho
  double vs;  // vsigma
};

__device__ __forceinline__ void dev_slater(double rho, double& e, double& v) {
  const double cx = -0.73855876638202240587;  // -(3/4)(3/pi)^(1/3)
  const double r13 = cbrt(rho);
  e = cx * rho * r13;
  v = (4.0 / 3.0) * cx * r13;
}

__device__ __forceinline__ void dev_vwn_rpa(double rho, double& e, double& v) {
  const double A = 0.0310907, b = 13.0720, c = 42.7198, x0 = -0.409286;
  const double pi = 3.14159265358979323846;
  const double rs = cbrt(3.0 / (4.0 * pi * rho));
  const double x = sqrt(rs);
  const double X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c;
  const double Q = sqrt(4.0 * c - b * b);
  const double tx = 2.0 * x + b;
  const double at = atan(Q / tx);
  const double bx0 = b * x0 / X0;
  const double eps = A * (log(x * x / X) + 2.0 * b / Q * at -
                          bx0 * (log((x - x0) * (x - x0) / X) + 2.0 * (b + 2.0 * x0) / Q * at));
  const double den = tx * tx + Q * Q;
  const double deps = A * (2.0 / x - tx / X - 4.0 * b / den -
                           bx0 * (2.0 / (x - x0) - tx / X - 4.0 * (b + 2.0 * x0) / den));
  e = rho * eps;
  v = eps - x / 6.0 * deps;
}

__device__ __forceinline__ void dev_pbe_x(double rho, double sigma, double& e, double& vr,
                                          double& vs) {
  const double kappa = 0.804, mu = 0.2195149727645171;
  const double ax = -0.73855876638202240587;
  const double C = 0.026121172985233605;  // 1 / (4 (3 pi^2)^(2/3))
  const double r13 = cbrt(rho);
  const double r43 = rho * r13;
  const double ir83 = 1.0 / (r43 * r43);
  const double s2 = C * sigma * ir83;
  const double den = 1.0 + mu * s2 / kappa;
  const double F = 1.0 + kappa - kappa / den;
  const double dF = mu / (den * den);
  e = ax * r43 * F;
  vr = (4.0 / 3.0) * ax * r13 * F - (8.0 / 3.0) * ax * r13 * dF * s2;
  vs = ax * r43 * dF * C * ir83;
}

__device__ __forceinline__ void dev_pbe_c(double rho, double sigma, double& e, double& vr,
                                          double& vs) {
  const double pi = 3.14159265358979323846;
  const double beta = 0.06672455060314922;
  const double gamma = 0.0310906908696549;  // (1 - ln 2) / pi^2
  const double y = beta / gamma;
  const double a = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const double rs = cbrt(3.0 / (4.0 * pi * rho));
  const double srs = sqrt(rs);
  const double Qd = 2.0 * a * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double dQd = 2.0 * a * (0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2.0 * b4 * rs);
  const double L = log1p(1.0 / Qd);
  const double ec = -2.0 * a * (1.0 + a1 * rs) * L;
  const double dec = -2.0 * a * a1 * L + 2.0 * a * (1.0 + a1 * rs) * dQd / (Qd * Qd + Qd);
  const double D = 0.0634682060977037;  // pi / (16 (3 pi^2)^(1/3))
  const double r73 = rho * rho * cbrt(rho);
  const double T = D * sigma / r73;
  const double E = exp(-ec / gamma);
  const double Aa = y / (E - 1.0);
  const double N = 1.0 + Aa * T, Dn = 1.0 + Aa * T + Aa * Aa * T * T;
  const double Q = T * N / Dn;
  const double H = gamma * log1p(y * Q);
  const double dH_dQ = gamma * y / (1.0 + y * Q);
  const double iDn2 = 1.0 / (Dn * Dn);
  const double dQ_dT = ((N + T * Aa) * Dn - T * N * (Aa + 2.0 * Aa * Aa * T)) * iDn2;
  const double dQ_dA = T * T * (Dn - N * (1.0 + 2.0 * Aa * T)) * iDn2;
  const double dA_de = y * E / (gamma * (E - 1.0) * (E - 1.0));
  e = rho * (ec + H);
  vr = ec + H - rs / 3.0 * dec * (1.0 + dH_dQ * dQ_dA * dA_de) - (7.0 / 3.0) * T * dH_dQ * dQ_dT;
  vs = rho * dH_dQ * dQ_dT * D / r73;
}

// func: XCQ_FUNC_SVWN5 / XCQ_FUNC_PBE / XCQ_FUNC_PBE0 (XC part).
__device__ __forceinline__ XCOut dev_eval_functional(int func, double rho, double sigma) {
  XCOut o{0.0, 0.0, 0.0};
  double e, vr, vs;
  if (func == XCQ_FUNC_SVWN5) {
    if (rho > 1e-15) {
      dev_slater(rho, e, vr);
      o.e += e;
      o.vr += vr;
      dev_vwn_rpa(rho, e, vr);
      o.e += e;
      o.vr += vr;
    }
    return o;
  }
  const double xs = (func == XCQ_FUNC_PBE0) ? 0.75 : 1.0;
  const double stol = 2.1544346900318838e-43;
  const double s = fmax(sigma, stol * stol);
  if (rho > 1e-15) {
    dev_pbe_x(rho, s, e, vr, vs);
    o.e += xs * e;
    o.vr += xs * vr;
    o.vs += xs * vs;
  }
  if (rho > 1e-12) {
    dev_pbe_c(rho, s, e, vr, vs);
    o.e += e;
    o.vr += vr;
    o.vs += vs;
  }
  return o;
}

}  // namespace xcq
