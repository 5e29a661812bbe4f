// Device evaluation of the functionals on the quadrature path (unpolarised,
// and spin-polarised for UKS below),
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

// ---- spin-polarised forms (UKS) -------------------------------------------
// Inputs rho_a, rho_b and sigma_aa, sigma_ab, sigma_bb; outputs e = n eps and
// the five derivatives (ExchCXX's polarised layout: vrho (a, b), vsigma
// (aa, ab, bb)).
//  * Exchange by spin scaling (Oliver-Perdew 1979):
//    e_x[ra, rb] = (e_x[2 ra, 4 saa] + e_x[2 rb, 4 sbb]) / 2, so
//    vrho_a = vrho(2 ra, 4 saa), vsigma_aa = 2 vsigma(2 ra, 4 saa), vsigma_ab = 0.
//  * Correlation in (n, zeta, sigma_tot = saa + 2 sab + sbb), differentiated
//    by forward-mode dual numbers in those three variables, then
//    d/d rho_a = d/dn + (1 - zeta)/n d/dzeta, d/d rho_b = d/dn - (1 + zeta)/n d/dzeta,
//    d/d sigma_aa = d/d sigma_bb = d/d sigma_tot, d/d sigma_ab = 2 d/d sigma_tot.
//    VWN(RPA): eps = eps_P + (eps_F - eps_P) f(zeta) with the paramagnetic and
//    ferromagnetic RPA fits (ferromagnetic A=0.01554535, b=20.1231,
//    c=101.578, x0=-0.743294), the interpolation of libxc's LDA_C_VWN_RPA.
//    PBE: PW92 with spin stiffness (modified parameters of libxc's GGA_C_PBE:
//    ferromagnetic a=0.01554535, alpha1=0.20548, beta=14.1189, 6.1977,
//    3.3662, 0.62517; -alpha_c a=0.0168869, alpha1=0.11125, beta=10.357,
//    3.6231, 0.88026, 0.49671; f''(0)=1.709920934161365617563962776245) and
//    phi(zeta) = ((1+zeta)^(2/3) + (1-zeta)^(2/3)) / 2 in H.
// Thresholds: no exchange from a channel with 2 rho_s <= 1e-15; no
// correlation for n <= 1e-15 (SVWN5) or 1e-12 (PBE).
struct XCOutPol {
  double e, vra, vrb, vsaa, vsab, vsbb;
};

struct D3 {
  double v, d0, d1, d2;
};
__device__ __forceinline__ D3 d3c(double v) { return D3{v, 0., 0., 0.}; }
__device__ __forceinline__ D3 d3ch(const D3& a, double f, double df) {
  return D3{f, df * a.d0, df * a.d1, df * a.d2};
}
__device__ __forceinline__ D3 operator+(const D3& a, const D3& b) {
  return D3{a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2};
}
__device__ __forceinline__ D3 operator-(const D3& a, const D3& b) {
  return D3{a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2};
}
__device__ __forceinline__ D3 operator*(const D3& a, const D3& b) {
  return D3{a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1,
            a.d2 * b.v + a.v * b.d2};
}
__device__ __forceinline__ D3 operator/(const D3& a, const D3& b) {
  const double i = 1.0 / b.v, q = a.v * i;
  return D3{q, (a.d0 - q * b.d0) * i, (a.d1 - q * b.d1) * i, (a.d2 - q * b.d2) * i};
}
__device__ __forceinline__ D3 operator+(const D3& a, double b) { return D3{a.v + b, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator+(double a, const D3& b) { return b + a; }
__device__ __forceinline__ D3 operator-(const D3& a, double b) { return D3{a.v - b, a.d0, a.d1, a.d2}; }
__device__ __forceinline__ D3 operator-(double a, const D3& b) {
  return D3{a - b.v, -b.d0, -b.d1, -b.d2};
}
__device__ __forceinline__ D3 operator*(const D3& a, double b) {
  return D3{a.v * b, a.d0 * b, a.d1 * b, a.d2 * b};
}
__device__ __forceinline__ D3 operator*(double a, const D3& b) { return b * a; }
__device__ __forceinline__ D3 operator/(const D3& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ D3 operator/(double a, const D3& b) { return d3c(a) / b; }
__device__ __forceinline__ D3 d3sqrt(const D3& a) {
  const double s = sqrt(a.v);
  return d3ch(a, s, 0.5 / s);
}
__device__ __forceinline__ D3 d3cbrt(const D3& a) {
  const double c = cbrt(a.v);
  return d3ch(a, c, 1.0 / (3.0 * c * c));
}
__device__ __forceinline__ D3 d3pow(const D3& a, double p) {
  const double x = pow(a.v, p);
  return d3ch(a, x, p * pow(a.v, p - 1.0));
}
__device__ __forceinline__ D3 d3log(const D3& a) { return d3ch(a, log(a.v), 1.0 / a.v); }
__device__ __forceinline__ D3 d3log1p(const D3& a) { return d3ch(a, log1p(a.v), 1.0 / (1.0 + a.v)); }
__device__ __forceinline__ D3 d3exp(const D3& a) {
  const double x = exp(a.v);
  return d3ch(a, x, x);
}
__device__ __forceinline__ D3 d3atan(const D3& a) { return d3ch(a, atan(a.v), 1.0 / (1.0 + a.v * a.v)); }

__device__ __forceinline__ D3 d3_vwn_fit(const D3& x, double A, double b, double c, double x0) {
  const D3 X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c, Q = sqrt(4.0 * c - b * b);
  const D3 at = d3atan(Q / (2.0 * x + b));
  return A * (d3log(x * x / X) + (2.0 * b / Q) * at -
              (b * x0 / X0) * (d3log((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}
__device__ __forceinline__ D3 d3_pw92(const D3& rs, double a, double a1, double b1, double b2,
                                      double b3, double b4) {
  const D3 sr = d3sqrt(rs);
  const D3 q = 2.0 * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs);
  return -2.0 * a * (1.0 + a1 * rs) * d3log1p(1.0 / q);
}
__device__ __forceinline__ D3 d3_fzeta(const D3& z) {
  return (d3pow(1.0 + z, 4.0 / 3.0) + d3pow(1.0 - z, 4.0 / 3.0) - 2.0) *
         (1.0 / 0.51984209978974632953);  // 2^(4/3) - 2
}

// Correlation energy density n eps_c(n, zeta, sigma_tot) (sigma_tot unused
// for SVWN5).
__device__ __forceinline__ D3 d3_corr(int func, const D3& n, const D3& z, const D3& st) {
  const double pi = 3.14159265358979323846;
  const D3 rs = d3cbrt(3.0 / (4.0 * pi * n));
  if (func == XCQ_FUNC_SVWN5) {
    const D3 x = d3sqrt(rs);
    const D3 ep = d3_vwn_fit(x, 0.0310907, 13.0720, 42.7198, -0.409286);
    const D3 ef = d3_vwn_fit(x, 0.01554535, 20.1231, 101.578, -0.743294);
    return n * (ep + (ef - ep) * d3_fzeta(z));
  }
  const D3 ep = d3_pw92(rs, 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const D3 ef = d3_pw92(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const D3 mac = d3_pw92(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const double fz20 = 1.709920934161365617563962776245;
  const D3 f = d3_fzeta(z), z2 = z * z, z4 = z2 * z2;
  const D3 ec = ep - mac * f * (1.0 / fz20) * (1.0 - z4) + (ef - ep) * f * z4;
  const double beta = 0.06672455060314922, gamma = 0.0310906908696549;
  const D3 phi = 0.5 * (d3pow(1.0 + z, 2.0 / 3.0) + d3pow(1.0 - z, 2.0 / 3.0));
  const D3 phi3 = phi * phi * phi;
  const double D = 0.0634682060977037;  // pi / (16 (3 pi^2)^(1/3))
  const D3 t2 = D * st / (phi * phi * n * n * d3cbrt(n));
  const D3 A = (beta / gamma) / (d3exp(-1.0 * ec / (gamma * phi3)) - 1.0);
  const D3 At2 = A * t2;
  const D3 H = gamma * phi3 * d3log1p((beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return n * (ec + H);
}

__device__ __forceinline__ XCOutPol dev_eval_functional_pol(int func, double ra, double rb,
                                                            double saa, double sab,
                                                            double sbb) {
  XCOutPol o{0., 0., 0., 0., 0., 0.};
  ra = fmax(ra, 0.0);
  rb = fmax(rb, 0.0);
  const bool gga = func != XCQ_FUNC_SVWN5;
  if (gga) {
    const double stol2 = 2.1544346900318838e-43 * 2.1544346900318838e-43;
    saa = fmax(saa, stol2);
    sbb = fmax(sbb, stol2);
    const double hs = 0.5 * (saa + sbb);
    sab = fmax(fmin(sab, hs), -hs);
  }
  const double xs = func == XCQ_FUNC_PBE0 ? 0.75 : 1.0;
  double e, vr, vs;
  if (2.0 * ra > 1e-15) {
    if (gga) {
      dev_pbe_x(2.0 * ra, 4.0 * saa, e, vr, vs);
      o.vsaa += 2.0 * xs * vs;
    } else {
      dev_slater(2.0 * ra, e, vr);
    }
    o.e += 0.5 * xs * e;
    o.vra += xs * vr;
  }
  if (2.0 * rb > 1e-15) {
    if (gga) {
      dev_pbe_x(2.0 * rb, 4.0 * sbb, e, vr, vs);
      o.vsbb += 2.0 * xs * vs;
    } else {
      dev_slater(2.0 * rb, e, vr);
    }
    o.e += 0.5 * xs * e;
    o.vrb += xs * vr;
  }
  const double nt = ra + rb;
  if (nt <= (gga ? 1e-12 : 1e-15)) return o;
  const double zt = (ra - rb) / nt;
  const D3 c = d3_corr(func, D3{nt, 1., 0., 0.}, D3{zt, 0., 1., 0.},
                       D3{saa + 2.0 * sab + sbb, 0., 0., 1.});
  o.e += c.v;
  o.vra += c.d0 + (1.0 - zt) / nt * c.d1;
  o.vrb += c.d0 - (1.0 + zt) / nt * c.d1;
  if (gga) {
    o.vsaa += c.d2;
    o.vsab += 2.0 * c.d2;
    o.vsbb += c.d2;
  }
  return o;
}

// ---- second derivatives (FXC contraction) --------------------------------
// The trial kernels of the FXC contraction need, at each point,
//   A  = v2rho2 drho + v2rhosigma dsigma,
//   Bc = v2rhosigma drho + v2sigma2 dsigma
// for one direction (drho, dsigma) (reference_local_host_work_driver.cxx:
// 1114-1143 with drho = trho, dsigma = 2 grad rho . grad trho).  Both are
// derivatives along that direction of vrho and vsigma, so the unpolarised
// energy density is evaluated once in a second-order forward-mode number:
// value, partials in rho (r) and sigma (s), the directional derivative (t)
// and its partials (rt = A, st = Bc).
struct HD {
  double v, r, s, t, rt, st;
};
__device__ __forceinline__ HD hdc(double v) { return HD{v, 0., 0., 0., 0., 0.}; }
// f(a) from f, f' and f'' at a.v.
__device__ __forceinline__ HD hd_chain(const HD& a, double f0, double f1, double f2) {
  return HD{f0, f1 * a.r, f1 * a.s, f1 * a.t, f2 * a.r * a.t + f1 * a.rt,
            f2 * a.s * a.t + f1 * a.st};
}
__device__ __forceinline__ HD operator+(const HD& a, const HD& b) {
  return HD{a.v + b.v, a.r + b.r, a.s + b.s, a.t + b.t, a.rt + b.rt, a.st + b.st};
}
__device__ __forceinline__ HD operator-(const HD& a, const HD& b) {
  return HD{a.v - b.v, a.r - b.r, a.s - b.s, a.t - b.t, a.rt - b.rt, a.st - b.st};
}
__device__ __forceinline__ HD operator*(const HD& a, const HD& b) {
  return HD{a.v * b.v,
            a.r * b.v + a.v * b.r,
            a.s * b.v + a.v * b.s,
            a.t * b.v + a.v * b.t,
            a.rt * b.v + a.r * b.t + a.t * b.r + a.v * b.rt,
            a.st * b.v + a.s * b.t + a.t * b.s + a.v * b.st};
}
__device__ __forceinline__ HD hd_inv(const HD& a) {
  const double i = 1.0 / a.v;
  return hd_chain(a, i, -i * i, 2.0 * i * i * i);
}
__device__ __forceinline__ HD operator/(const HD& a, const HD& b) { return a * hd_inv(b); }
__device__ __forceinline__ HD operator+(const HD& a, double b) {
  return HD{a.v + b, a.r, a.s, a.t, a.rt, a.st};
}
__device__ __forceinline__ HD operator+(double a, const HD& b) { return b + a; }
__device__ __forceinline__ HD operator-(const HD& a, double b) { return a + (-b); }
__device__ __forceinline__ HD operator-(double a, const HD& b) {
  return HD{a - b.v, -b.r, -b.s, -b.t, -b.rt, -b.st};
}
__device__ __forceinline__ HD operator*(const HD& a, double b) {
  return HD{a.v * b, a.r * b, a.s * b, a.t * b, a.rt * b, a.st * b};
}
__device__ __forceinline__ HD operator*(double a, const HD& b) { return b * a; }
__device__ __forceinline__ HD operator/(const HD& a, double b) { return a * (1.0 / b); }
__device__ __forceinline__ HD operator/(double a, const HD& b) { return a * hd_inv(b); }
__device__ __forceinline__ HD hd_sqrt(const HD& a) {
  const double q = sqrt(a.v);
  return hd_chain(a, q, 0.5 / q, -0.25 / (q * a.v));
}
__device__ __forceinline__ HD hd_cbrt(const HD& a) {
  const double c = cbrt(a.v), c2 = c * c;
  return hd_chain(a, c, 1.0 / (3.0 * c2), -2.0 / (9.0 * c2 * a.v));
}
__device__ __forceinline__ HD hd_log(const HD& a) {
  const double i = 1.0 / a.v;
  return hd_chain(a, log(a.v), i, -i * i);
}
__device__ __forceinline__ HD hd_log1p(const HD& a) {
  const double i = 1.0 / (1.0 + a.v);
  return hd_chain(a, log1p(a.v), i, -i * i);
}
__device__ __forceinline__ HD hd_exp(const HD& a) {
  const double e = exp(a.v);
  return hd_chain(a, e, e, e);
}
__device__ __forceinline__ HD hd_atan(const HD& a) {
  const double i = 1.0 / (1.0 + a.v * a.v);
  return hd_chain(a, atan(a.v), i, -2.0 * a.v * i * i);
}

// The unpolarised energy densities of the forms at the top of this file.
__device__ __forceinline__ HD hd_vwn_fit(const HD& x, double A, double b, double c, double x0) {
  const HD X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c, Q = sqrt(4.0 * c - b * b);
  const HD at = hd_atan(Q / (2.0 * x + b));
  return A * (hd_log(x * x / X) + (2.0 * b / Q) * at -
              (b * x0 / X0) * (hd_log((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
}
__device__ __forceinline__ HD hd_pw92(const HD& rs, double a, double a1, double b1, double b2,
                                      double b3, double b4) {
  const HD sr = hd_sqrt(rs);
  const HD q = 2.0 * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs);
  return -2.0 * a * (1.0 + a1 * rs) * hd_log1p(1.0 / q);
}
__device__ __forceinline__ HD hd_energy(int func, const HD& rho, const HD& sigma) {
  const double pi = 3.14159265358979323846;
  const double ax = -0.73855876638202240587;  // -(3/4)(3/pi)^(1/3)
  HD e = hdc(0.);
  if (func == XCQ_FUNC_SVWN5) {
    if (rho.v > 1e-15) {
      e = ax * rho * hd_cbrt(rho);
      const HD rs = hd_cbrt((3.0 / (4.0 * pi)) / rho);
      e = e + rho * hd_vwn_fit(hd_sqrt(rs), 0.0310907, 13.0720, 42.7198, -0.409286);
    }
    return e;
  }
  const double xs = (func == XCQ_FUNC_PBE0) ? 0.75 : 1.0;
  if (rho.v > 1e-15) {
    const double kappa = 0.804, mu = 0.2195149727645171;
    const double C = 0.026121172985233605;  // 1 / (4 (3 pi^2)^(2/3))
    const HD r43 = rho * hd_cbrt(rho);
    const HD s2 = C * sigma / (r43 * r43);
    const HD F = (1.0 + kappa) - kappa / (1.0 + (mu / kappa) * s2);
    e = (xs * ax) * r43 * F;
  }
  if (rho.v > 1e-12) {
    const double beta = 0.06672455060314922, gamma = 0.0310906908696549;
    const double y = beta / gamma;
    const double D = 0.0634682060977037;  // pi / (16 (3 pi^2)^(1/3))
    const HD rs = hd_cbrt((3.0 / (4.0 * pi)) / rho);
    const HD ec = hd_pw92(rs, 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
    const HD t2 = D * sigma / (rho * rho * hd_cbrt(rho));
    const HD Aa = y / (hd_exp((-1.0 / gamma) * ec) - 1.0);
    const HD At2 = Aa * t2;
    const HD H = gamma * hd_log1p(y * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
    e = e + rho * (ec + H);
  }
  return e;
}

// vsigma, A and Bc (above) at (rho, sigma) along (drho, dsigma); LDA leaves
// vs = Bc = 0.  Thresholds and the sigma floor as dev_eval_functional.
__device__ __forceinline__ void dev_fxc_point(int func, double rho, double sigma, double drho,
                                              double dsigma, double& vs, double& A, double& Bc) {
  const bool gga = func != XCQ_FUNC_SVWN5;
  const double stol = 2.1544346900318838e-43;
  const HD R{rho, 1., 0., drho, 0., 0.};
  const HD S{gga ? fmax(sigma, stol * stol) : 0., 0., 1., gga ? dsigma : 0., 0., 0.};
  const HD e = hd_energy(func, R, S);
  vs = gga ? e.s : 0.;
  A = e.rt;
  Bc = gga ? e.st : 0.;
}

// ---- generic forward-mode duals (UKS FXC contraction) -----------------------
// GD<S, N>: value and N derivative slots over the scalar S.  S = double gives
// first derivatives; GD<GD<double, 5>, 1> carries the five partials of the
// polarised variables (inner) and their derivatives along one direction
// (outer), i.e. the second derivatives the UKS trial kernels contract.
template <class S, int N>
struct GD {
  S v;
  S d[N];
};
template <class S>
struct GCst;
template <>
struct GCst<double> {
  __device__ __forceinline__ static double of(double x) { return x; }
};
template <class S, int N>
struct GCst<GD<S, N>> {
  __device__ __forceinline__ static GD<S, N> of(double x) {
    GD<S, N> r;
    r.v = GCst<S>::of(x);
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = GCst<S>::of(0.);
    return r;
  }
};
__device__ __forceinline__ double gval(double x) { return x; }
template <class S, int N>
__device__ __forceinline__ double gval(const GD<S, N>& a) { return gval(a.v); }
template <class S, int N>
__device__ __forceinline__ GD<S, N> gchain(const GD<S, N>& a, const S& f, const S& df) {
  GD<S, N> r;
  r.v = f;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = df * a.d[k];
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator+(const GD<S, N>& a, const GD<S, N>& b) {
  GD<S, N> r;
  r.v = a.v + b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator-(const GD<S, N>& a, const GD<S, N>& b) {
  GD<S, N> r;
  r.v = a.v - b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator*(const GD<S, N>& a, const GD<S, N>& b) {
  GD<S, N> r;
  r.v = a.v * b.v;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> ginv(const GD<S, N>& b) {
  const S i = 1.0 / b.v;
  return gchain(b, i, -1.0 * (i * i));
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator/(const GD<S, N>& a, const GD<S, N>& b) {
  return a * ginv(b);
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator+(const GD<S, N>& a, double b) {
  GD<S, N> r = a;
  r.v = a.v + b;
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator+(double a, const GD<S, N>& b) { return b + a; }
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator-(const GD<S, N>& a, double b) { return a + (-b); }
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator-(double a, const GD<S, N>& b) {
  return GCst<GD<S, N>>::of(a) - b;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator*(const GD<S, N>& a, double b) {
  GD<S, N> r;
  r.v = a.v * b;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
  return r;
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator*(double a, const GD<S, N>& b) { return b * a; }
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator/(const GD<S, N>& a, double b) { return a * (1.0 / b); }
template <class S, int N>
__device__ __forceinline__ GD<S, N> operator/(double a, const GD<S, N>& b) { return a * ginv(b); }

__device__ __forceinline__ double gsqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double gcbrt(double x) { return cbrt(x); }
__device__ __forceinline__ double gpow(double x, double e) { return pow(x, e); }
__device__ __forceinline__ double glog(double x) { return log(x); }
__device__ __forceinline__ double glog1p(double x) { return log1p(x); }
__device__ __forceinline__ double gexp(double x) { return exp(x); }
__device__ __forceinline__ double gatan(double x) { return atan(x); }
template <class S, int N>
__device__ __forceinline__ GD<S, N> gsqrt(const GD<S, N>& a) {
  const S q = gsqrt(a.v);
  return gchain(a, q, 0.5 / q);
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> gcbrt(const GD<S, N>& a) {
  const S c = gcbrt(a.v);
  return gchain(a, c, 1.0 / (3.0 * (c * c)));
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> gpow(const GD<S, N>& a, double e) {
  return gchain(a, gpow(a.v, e), e * gpow(a.v, e - 1.0));
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> glog(const GD<S, N>& a) { return gchain(a, glog(a.v), 1.0 / a.v); }
template <class S, int N>
__device__ __forceinline__ GD<S, N> glog1p(const GD<S, N>& a) {
  return gchain(a, glog1p(a.v), 1.0 / (1.0 + a.v));
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> gexp(const GD<S, N>& a) {
  const S e = gexp(a.v);
  return gchain(a, e, e);
}
template <class S, int N>
__device__ __forceinline__ GD<S, N> gatan(const GD<S, N>& a) {
  return gchain(a, gatan(a.v), 1.0 / (1.0 + a.v * a.v));
}

// The spin-polarised energy density n eps of the forms above (exchange by
// spin scaling, the spin-interpolated VWN(RPA) and PBE correlations), for
// any dual type T; thresholds as dev_eval_functional_pol.
template <class T>
__device__ __forceinline__ T gpol_energy(int func, const T& ra, const T& rb, const T& saa,
                                         const T& sab, const T& sbb) {
  const double pi = 3.14159265358979323846;
  const double ax = -0.73855876638202240587;  // -(3/4)(3/pi)^(1/3)
  const bool gga = func != XCQ_FUNC_SVWN5;
  const double xs = func == XCQ_FUNC_PBE0 ? 0.75 : 1.0;
  T e = GCst<T>::of(0.);
  auto ex = [&](const T& n, const T& s) {  // unpolarised exchange of (n, s)
    const T n43 = n * gcbrt(n);
    if (!gga) return ax * n43;
    const double kappa = 0.804, mu = 0.2195149727645171;
    const double C = 0.026121172985233605;  // 1 / (4 (3 pi^2)^(2/3))
    const T F = (1.0 + kappa) - kappa / (1.0 + (mu / kappa) * (C * s / (n43 * n43)));
    return ax * n43 * F;
  };
  if (2.0 * gval(ra) > 1e-15) e = e + (0.5 * xs) * ex(2.0 * ra, 4.0 * saa);
  if (2.0 * gval(rb) > 1e-15) e = e + (0.5 * xs) * ex(2.0 * rb, 4.0 * sbb);
  const T n = ra + rb;
  if (gval(n) <= (gga ? 1e-12 : 1e-15)) return e;
  const T z = (ra - rb) / n;
  const T rs = gcbrt((3.0 / (4.0 * pi)) / n);
  auto fz = [&](const T& zz) {
    return (gpow(1.0 + zz, 4.0 / 3.0) + gpow(1.0 - zz, 4.0 / 3.0) - 2.0) *
           (1.0 / 0.51984209978974632953);
  };
  auto vwn = [&](const T& x, double A, double b, double c, double x0) {
    const T X = x * x + b * x + c;
    const double X0 = x0 * x0 + b * x0 + c, Q = sqrt(4.0 * c - b * b);
    const T at = gatan(Q / (2.0 * x + b));
    return A * (glog(x * x / X) + (2.0 * b / Q) * at -
                (b * x0 / X0) * (glog((x - x0) * (x - x0) / X) + (2.0 * (b + 2.0 * x0) / Q) * at));
  };
  auto pw92 = [&](double a, double a1, double b1, double b2, double b3, double b4) {
    const T sr = gsqrt(rs);
    const T q = 2.0 * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs);
    return -2.0 * a * (1.0 + a1 * rs) * glog1p(1.0 / q);
  };
  if (!gga) {
    const T x = gsqrt(rs);
    const T ep = vwn(x, 0.0310907, 13.0720, 42.7198, -0.409286);
    const T ef = vwn(x, 0.01554535, 20.1231, 101.578, -0.743294);
    return e + n * (ep + (ef - ep) * fz(z));
  }
  const T ep = pw92(0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const T ef = pw92(0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const T mac = pw92(0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const double fz20 = 1.709920934161365617563962776245;
  const T f = fz(z), z2 = z * z, z4 = z2 * z2;
  const T ec = ep - mac * f * (1.0 / fz20) * (1.0 - z4) + (ef - ep) * f * z4;
  const double beta = 0.06672455060314922, gamma = 0.0310906908696549;
  const T phi = 0.5 * (gpow(1.0 + z, 2.0 / 3.0) + gpow(1.0 - z, 2.0 / 3.0));
  const T phi3 = phi * phi * phi;
  const double D = 0.0634682060977037;  // pi / (16 (3 pi^2)^(1/3))
  const T t2 = D * (saa + 2.0 * sab + sbb) / (phi * phi * n * n * gcbrt(n));
  const T A = (beta / gamma) / (gexp((-1.0 / gamma) * ec / phi3) - 1.0);
  const T At2 = A * t2;
  const T H = gamma * phi3 * glog1p((beta / gamma) * t2 * (1.0 + At2) / (1.0 + At2 + At2 * At2));
  return e + n * (ec + H);
}

// Polarised first derivatives at x = (ra, rb, saa, sab, sbb) and their
// derivatives along dx: out[0..4] = (vra, vrb, vsaa, vsab, vsbb),
// out[5..9] the same differentiated along dx.
__device__ __forceinline__ void dev_fxc_point_pol(int func, const double* x, const double* dx,
                                                  double* out) {
  using G1 = GD<double, 5>;
  using G2 = GD<G1, 1>;
  G2 v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    v[k].v = GCst<G1>::of(x[k]);
    v[k].v.d[k] = 1.0;
    v[k].d[0] = GCst<G1>::of(dx[k]);
  }
  const G2 e = gpol_energy(func, v[0], v[1], v[2], v[3], v[4]);
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    out[k] = e.v.d[k];
    out[5 + k] = e.d[0].d[k];
  }
}

}  // namespace xcq
