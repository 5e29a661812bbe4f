// CPU oracle of the XC-quadrature path -- test infrastructure, never shipped.
// Every function names the GauXC file:line it restates.  See oracle.h.
#include "oracle.h"

#include <dlfcn.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// ---------------------------------------------------------------------------
// Optional host BLAS (GauXC's host integrator calls dgemm_/dsyr2k_ through
// host/blas.cxx:29,38).  Loaded at run time by oracle_use_blas; when absent
// the plain loops below compute the same products.
// ---------------------------------------------------------------------------
using dgemm_t = void (*)(const char*, const char*, const int*, const int*, const int*,
                         const double*, const double*, const int*, const double*, const int*,
                         const double*, double*, const int*);
using dsyr2k_t = void (*)(const char*, const char*, const int*, const int*, const double*,
                          const double*, const int*, const double*, const int*, const double*,
                          double*, const int*);
dgemm_t g_dgemm = nullptr;
dsyr2k_t g_dsyr2k = nullptr;

// ---------------------------------------------------------------------------
// SSF weights: src/xc_integrator/local_work_driver/host/reference/weights.cxx:117-237,
// constants src/xc_integrator/integrator_util/integrator_constants.hpp:17-20.
// ---------------------------------------------------------------------------
constexpr double kSSFa = 0.64;
constexpr double kSSFTol = 1e-13;

inline double frisch_g(double mu) {
  const double s = mu / kSSFa, s2 = s * s, s3 = s * s2, s5 = s3 * s2, s7 = s5 * s2;
  return (35. * (s - s3) + 21. * s5 - 5. * s7) / 16.;
}

// ---------------------------------------------------------------------------
// Real solid harmonics in gau2grid's CCA order (m = -l..l), expressed on the
// Cartesian monomials x^a y^b z^c of the shell (Helgaker-Jorgensen-Olsen
// eq. 6.4.47-6.4.50; the normalisation N_lm reproduces gau2grid's
// GG_SPHERICAL_CCA coefficients, e.g. l=2: sqrt(3)xy, sqrt(3)yz,
// z^2-(x^2+y^2)/2, sqrt(3)xz, sqrt(3)/2 (x^2-y^2)).
// ---------------------------------------------------------------------------
struct SphTerm {
  int a, b, c;
  double coef;
};

double fact(int n) {
  double r = 1.;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
double binom(int n, int k) {
  if (k < 0 || k > n) return 0.;
  return fact(n) / (fact(k) * fact(n - k));
}

std::vector<std::vector<SphTerm>> solid_harmonics(int l) {
  std::vector<std::vector<SphTerm>> out;
  for (int m = -l; m <= l; ++m) {
    const int am = std::abs(m);
    const double N = std::sqrt(2. * fact(l + am) * fact(l - am) / (m == 0 ? 2. : 1.)) /
                     (std::ldexp(1.0, am) * fact(l));
    std::vector<SphTerm> terms;
    // v runs over integers (m >= 0) or half-integers (m < 0); use 2v.
    const int v2_start = (m < 0) ? 1 : 0;
    for (int t = 0; t <= (l - am) / 2; ++t)
      for (int u = 0; u <= t; ++u)
        for (int v2 = v2_start; v2 <= am; v2 += 2) {
          const int sgn_exp = t + (v2 - v2_start) / 2;
          const double C = ((sgn_exp % 2) ? -1. : 1.) * std::pow(0.25, t) * binom(l, t) *
                           binom(l - t, am + t) * binom(t, u) * binom(am, v2);
          if (C == 0.) continue;
          const int ax = 2 * t + am - 2 * u - v2;
          const int ay = 2 * u + v2;
          const int az = l - 2 * t - am;
          bool merged = false;
          for (auto& tm : terms)
            if (tm.a == ax && tm.b == ay && tm.c == az) {
              tm.coef += N * C;
              merged = true;
            }
          if (!merged) terms.push_back({ax, ay, az, N * C});
        }
    out.push_back(terms);
  }
  return out;
}

struct SphTables {
  std::vector<std::vector<SphTerm>> t[8];
  SphTables() {
    for (int l = 0; l < 8; ++l) t[l] = solid_harmonics(l);
  }
};
const SphTables& sph_tables() {
  static SphTables tb;
  return tb;
}

inline double ipow(double x, int n) {
  double r = 1.;
  for (int i = 0; i < n; ++i) r *= x;
  return r;
}

// Monomial x^a y^b z^c and its gradient times the radial part.
// d/dx [x^a y^b z^c R] = a x^(a-1) y^b z^c R + x^a y^b z^c dR/dx.
inline void monomial_eval(int a, int b, int c, double x, double y, double z, double R, double Rx,
                          double Ry, double Rz, double* v, double* gx, double* gy, double* gz) {
  const double xa = ipow(x, a), yb = ipow(y, b), zc = ipow(z, c);
  const double mono = xa * yb * zc;
  *v = mono * R;
  if (gx) {
    const double dxa = a ? a * ipow(x, a - 1) : 0.;
    const double dyb = b ? b * ipow(y, b - 1) : 0.;
    const double dzc = c ? c * ipow(z, c - 1) : 0.;
    *gx = dxa * yb * zc * R + mono * Rx;
    *gy = xa * dyb * zc * R + mono * Ry;
    *gz = xa * yb * dzc * R + mono * Rz;
  }
}

// ---------------------------------------------------------------------------
// Functionals (unpolarised).  ExchCXX's builtin kernels (libxc-derived) are
// the reference; they are third-party and absent, so these are written from
// the published forms:
//   Slater exchange; the VWN correlation GauXC's "SVWN5" evaluates, which is
//   the Vosko-Wilk-Nusair (1980) fit to the RPA correlation energy (libxc
//   LDA_C_VWN_RPA parameters, paramagnetic: A=0.0310907, b=13.0720,
//   c=42.7198, x0=-0.409286) -- pinned to 2e-15 in EXC by GauXC's
//   benzene_svwn5 fixture; PBE exchange and correlation (Perdew-Burke-
//   Ernzerhof 1996) with libxc's constants and the PW92 "modified"
//   correlation parameters libxc's GGA_C_PBE uses.
// Density thresholds follow the libxc defaults ExchCXX inherits: rho <= tol
// gives zero output (1e-15; 1e-12 for PBE correlation).
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;

inline void slater(double rho, double& e, double& v) {  // e = energy density
  const double cx = -0.75 * std::cbrt(3. / kPi);
  const double r13 = std::cbrt(rho);
  e = cx * rho * r13;
  v = 4. / 3. * cx * r13;
}

inline void vwn_rpa(double rho, double& e, double& v) {
  const double A = 0.0310907, b = 13.0720, c = 42.7198, x0 = -0.409286;
  const double rs = std::cbrt(3. / (4. * kPi * rho));
  const double x = std::sqrt(rs);
  const double X = x * x + b * x + c, X0 = x0 * x0 + b * x0 + c;
  const double Q = std::sqrt(4. * c - b * b);
  const double tx = 2. * x + b;
  const double at = std::atan(Q / tx);
  const double eps = A * (std::log(x * x / X) + 2. * b / Q * at -
                          b * x0 / X0 *
                              (std::log((x - x0) * (x - x0) / X) + 2. * (b + 2. * x0) / Q * at));
  const double den = tx * tx + Q * Q;
  const double deps_dx = A * (2. / x - tx / X - 4. * b / den -
                              b * x0 / X0 * (2. / (x - x0) - tx / X - 4. * (b + 2. * x0) / den));
  e = rho * eps;
  v = eps - x / 6. * deps_dx;
}

inline void pbe_x(double rho, double sigma, double& e, double& vr, double& vs) {
  const double kappa = 0.804, mu = 0.2195149727645171;
  const double ax = -0.75 * std::cbrt(3. / kPi);
  const double C = 1. / (4. * std::pow(3. * kPi * kPi, 2. / 3.));
  const double r13 = std::cbrt(rho);
  const double r43 = rho * r13;
  const double s2 = C * sigma / (r43 * r43);
  const double den = 1. + mu * s2 / kappa;
  const double F = 1. + kappa - kappa / den;
  const double dF = mu / (den * den);  // dF/d(s^2)
  e = ax * r43 * F;
  vr = 4. / 3. * ax * r13 * F - 8. / 3. * ax * r13 * dF * s2;
  vs = ax * r43 * dF * C / (r43 * r43);
}

inline void pw92_mod(double rs, double& eps, double& deps) {
  const double a = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const double srs = std::sqrt(rs);
  const double Qd = 2. * a * (b1 * srs + b2 * rs + b3 * rs * srs + b4 * rs * rs);
  const double dQd = 2. * a * (0.5 * b1 / srs + b2 + 1.5 * b3 * srs + 2. * b4 * rs);
  const double L = std::log1p(1. / Qd);
  eps = -2. * a * (1. + a1 * rs) * L;
  deps = -2. * a * a1 * L + 2. * a * (1. + a1 * rs) * dQd / (Qd * Qd + Qd);
}

inline void pbe_c(double rho, double sigma, double& e, double& vr, double& vs) {
  const double beta = 0.06672455060314922;
  const double gamma = (1. - std::log(2.)) / (kPi * kPi);
  const double y = beta / gamma;
  const double rs = std::cbrt(3. / (4. * kPi * rho));
  double ec, dec;
  pw92_mod(rs, ec, dec);
  const double D = kPi / (16. * std::cbrt(3. * kPi * kPi));
  const double T = D * sigma / std::pow(rho, 7. / 3.);
  const double E = std::exp(-ec / gamma);
  const double A = y / (E - 1.);
  const double N = 1. + A * T, Dn = 1. + A * T + A * A * T * T;
  const double Q = T * N / Dn;
  const double H = gamma * std::log1p(y * Q);
  const double dH_dQ = gamma * y / (1. + y * Q);
  const double dQ_dT = ((N + T * A) * Dn - T * N * (A + 2. * A * A * T)) / (Dn * Dn);
  const double dQ_dA = T * T * (Dn - N * (1. + 2. * A * T)) / (Dn * Dn);
  const double dA_de = y * E / (gamma * (E - 1.) * (E - 1.));
  e = rho * (ec + H);
  vr = ec + H - rs / 3. * dec * (1. + dH_dQ * dQ_dA * dA_de) - 7. / 3. * T * dH_dQ * dQ_dT;
  vs = rho * dH_dQ * dQ_dT * D / std::pow(rho, 7. / 3.);
}

struct FuncOut {
  double e, vr, vs;
};

inline FuncOut eval_func(int func, double rho, double sigma) {
  FuncOut o{0., 0., 0.};
  const double sig_tol = 2.1544346900318838e-43;
  if (func == ORACLE_SVWN5) {
    double e, v;
    if (rho > 1e-15) {
      slater(rho, e, v);
      o.e += e;
      o.vr += v;
      vwn_rpa(rho, e, v);
      o.e += e;
      o.vr += v;
    }
    return o;
  }
  const double xs = (func == ORACLE_PBE0) ? 0.75 : 1.0;
  const double s = std::max(sigma, sig_tol * sig_tol);
  double e, vr, vs;
  if (rho > 1e-15) {
    pbe_x(rho, s, e, vr, vs);
    o.e += xs * e;
    o.vr += xs * vr;
    o.vs += xs * vs;
  }
  if (rho > 1e-12) {
    pbe_c(rho, s, e, vr, vs);
    o.e += e;
    o.vr += vr;
    o.vs += vs;
  }
  return o;
}

// ---------------------------------------------------------------------------
// Spin-polarised functionals (UKS): energy density per volume as a function
// of (rho_a, rho_b, sigma_aa, sigma_ab, sigma_bb), differentiated exactly by
// forward-mode dual numbers (five directions), so vrho / vsigma need no
// hand-derived formulas.  Published forms:
//   exchange by spin scaling, E_x[ra, rb] = (E_x[2 ra, 4 saa] + E_x[2 rb, 4 sbb]) / 2
//   (Oliver-Perdew 1979), for Slater and PBE exchange;
//   VWN RPA correlation with the simple spin interpolation
//   eps = eps_P + (eps_F - eps_P) f(zeta) of libxc's LDA_C_VWN_RPA
//   (paramagnetic and ferromagnetic RPA fits of Vosko-Wilk-Nusair 1980);
//   PBE correlation (Perdew-Burke-Ernzerhof 1996, phi(zeta) spin scaling)
//   over PW92 with its spin stiffness, "modified" parameters as libxc's
//   GGA_C_PBE (f''(0) = 1.709920934161365617563962776245).
// Thresholds: a spin channel with 2 rho_s <= 1e-15 gives no exchange; no
// correlation when rho_a + rho_b <= 1e-15 (LDA) or 1e-12 (PBE).  GauXC's UKS
// path calls these through ExchCXX (reference_replicated_xc_host_integrator_
// exc_vxc.hpp:445-450 with spin_dim 2).
// ---------------------------------------------------------------------------
// Dual numbers over a scalar S with five derivative slots.  S = double
// gives first derivatives; S = DualT<double> nests them, so the outer
// slots carry derivatives of the inner ones (second derivatives, used by
// the FXC contraction).
template <class S>
struct DualT {
  S v;
  S d[5];
  DualT(double x = 0.) : v(x) {
    for (S& e : d) e = S(0.);
  }
  static DualT var(double x, int k) {
    DualT r(x);
    r.d[k] = S(1.);
    return r;
  }
};
using Dual = DualT<double>;

inline double val(double x) { return x; }
template <class S>
double val(const DualT<S>& a) { return val(a.v); }

template <class S>
DualT<S> chain(const DualT<S>& a, const S& f, const S& df) {
  DualT<S> r;
  r.v = f;
  for (int k = 0; k < 5; ++k) r.d[k] = df * a.d[k];
  return r;
}
template <class S>
DualT<S> operator+(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r;
  r.v = a.v + b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <class S>
DualT<S> operator-(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r;
  r.v = a.v - b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <class S>
DualT<S> operator-(const DualT<S>& a) { return DualT<S>(0.) - a; }
template <class S>
DualT<S> operator*(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r;
  r.v = a.v * b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <class S>
DualT<S> operator/(const DualT<S>& a, const DualT<S>& b) {
  DualT<S> r;
  r.v = a.v / b.v;
  for (int k = 0; k < 5; ++k) r.d[k] = (a.d[k] * b.v - a.v * b.d[k]) / (b.v * b.v);
  return r;
}
template <class S>
DualT<S> operator+(const DualT<S>& a, double b) { return a + DualT<S>(b); }
template <class S>
DualT<S> operator+(double a, const DualT<S>& b) { return DualT<S>(a) + b; }
template <class S>
DualT<S> operator-(const DualT<S>& a, double b) { return a - DualT<S>(b); }
template <class S>
DualT<S> operator-(double a, const DualT<S>& b) { return DualT<S>(a) - b; }
template <class S>
DualT<S> operator*(const DualT<S>& a, double b) { return chain(a, a.v * b, S(b)); }
template <class S>
DualT<S> operator*(double a, const DualT<S>& b) { return chain(b, a * b.v, S(a)); }
template <class S>
DualT<S> operator/(const DualT<S>& a, double b) { return chain(a, a.v / b, S(1. / b)); }
template <class S>
DualT<S> operator/(double a, const DualT<S>& b) { return DualT<S>(a) / b; }

// Elementary functions: the double overloads end the recursion.
inline double dsqrt(double x) { return std::sqrt(x); }
inline double dcbrt(double x) { return std::cbrt(x); }
inline double dpow(double x, double e) { return std::pow(x, e); }
inline double dlog(double x) { return std::log(x); }
inline double dlog1p(double x) { return std::log1p(x); }
inline double dexp(double x) { return std::exp(x); }
inline double datan(double x) { return std::atan(x); }
template <class S>
DualT<S> dsqrt(const DualT<S>& a) {
  const S s = dsqrt(a.v);
  return chain(a, s, 0.5 / s);
}
template <class S>
DualT<S> dcbrt(const DualT<S>& a) {
  const S c = dcbrt(a.v);
  return chain(a, c, 1. / (3. * c * c));
}
template <class S>
DualT<S> dpow(const DualT<S>& a, double e) {
  return chain(a, dpow(a.v, e), e * dpow(a.v, e - 1.));
}
template <class S>
DualT<S> dlog(const DualT<S>& a) { return chain(a, dlog(a.v), 1. / a.v); }
template <class S>
DualT<S> dlog1p(const DualT<S>& a) { return chain(a, dlog1p(a.v), 1. / (1. + a.v)); }
template <class S>
DualT<S> dexp(const DualT<S>& a) {
  const S e = dexp(a.v);
  return chain(a, e, e);
}
template <class S>
DualT<S> datan(const DualT<S>& a) { return chain(a, datan(a.v), 1. / (1. + a.v * a.v)); }

// Unpolarised exchange energy densities of total density n (gradient s2 = |grad n|^2).
template <class D>
D slater_e(const D& n) { return -0.75 * std::cbrt(3. / kPi) * n * dcbrt(n); }
template <class D>
D pbe_x_e(const D& n, const D& s2) {
  const double kappa = 0.804, mu = 0.2195149727645171;
  const double C = 1. / (4. * std::pow(3. * kPi * kPi, 2. / 3.));
  const D n43 = n * dcbrt(n);
  const D ss = C * s2 / (n43 * n43);
  const D F = 1. + kappa - kappa / (1. + mu * ss / kappa);
  return -0.75 * std::cbrt(3. / kPi) * n43 * F;
}
// VWN fit of one spin state (x = sqrt(rs)).
template <class D>
D vwn_fit(const D& x, double A, double b, double c, double x0) {
  const D X = x * x + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c, Q = std::sqrt(4. * c - b * b);
  const D at = datan(Q / (2. * x + b));
  return A * (dlog(x * x / X) + 2. * b / Q * at -
              b * x0 / X0 * (dlog((x - x0) * (x - x0) / X) + 2. * (b + 2. * x0) / Q * at));
}
// PW92 G(rs) of one parameter set.
template <class D>
D pw92_g(const D& rs, double a, double a1, double b1, double b2, double b3,
                   double b4) {
  const D sr = dsqrt(rs);
  const D q = 2. * a * (b1 * sr + b2 * rs + b3 * rs * sr + b4 * rs * rs);
  return -2. * a * (1. + a1 * rs) * dlog1p(1. / q);
}
template <class D>
D fzeta(const D& z) {
  return (dpow(1. + z, 4. / 3.) + dpow(1. - z, 4. / 3.) - 2.) / (std::pow(2., 4. / 3.) - 2.);
}

// Energy density of the polarised functional; all five inputs are duals.
template <class D>
D pol_energy(int func, const D& ra, const D& rb, const D& saa, const D& sab, const D& sbb) {
  const double dens_tol = 1e-15;
  D e(0.);
  const double xs = func == ORACLE_PBE0 ? 0.75 : 1.0;
  // Exchange by spin scaling.
  if (2. * val(ra) > dens_tol)
    e = e + 0.5 * xs * (func == ORACLE_SVWN5 ? slater_e(2. * ra) : pbe_x_e(2. * ra, 4. * saa));
  if (2. * val(rb) > dens_tol)
    e = e + 0.5 * xs * (func == ORACLE_SVWN5 ? slater_e(2. * rb) : pbe_x_e(2. * rb, 4. * sbb));
  const D n = ra + rb;
  if (val(n) <= (func == ORACLE_SVWN5 ? 1e-15 : 1e-12)) return e;
  const D zeta = (ra - rb) / n;
  const D rs = dcbrt(3. / (4. * kPi * n));
  if (func == ORACLE_SVWN5) {
    const D x = dsqrt(rs);
    const D ep = vwn_fit(x, 0.0310907, 13.0720, 42.7198, -0.409286);
    const D ef = vwn_fit(x, 0.01554535, 20.1231, 101.578, -0.743294);
    return e + n * (ep + (ef - ep) * fzeta(zeta));
  }
  const D ep = pw92_g(rs, 0.0310907, 0.21370, 7.5957, 3.5876, 1.6382, 0.49294);
  const D ef = pw92_g(rs, 0.01554535, 0.20548, 14.1189, 6.1977, 3.3662, 0.62517);
  const D mac = pw92_g(rs, 0.0168869, 0.11125, 10.357, 3.6231, 0.88026, 0.49671);
  const double fz20 = 1.709920934161365617563962776245;
  const D f = fzeta(zeta), z4 = zeta * zeta * zeta * zeta;
  const D ec = ep - mac * f / fz20 * (1. - z4) + (ef - ep) * f * z4;
  const double beta = 0.06672455060314922, gamma = (1. - std::log(2.)) / (kPi * kPi);
  const D phi = 0.5 * (dpow(1. + zeta, 2. / 3.) + dpow(1. - zeta, 2. / 3.));
  const D phi3 = phi * phi * phi;
  const double kD = kPi / (16. * std::cbrt(3. * kPi * kPi));
  const D stot = saa + 2. * sab + sbb;
  const D t2 = kD * stot / (phi * phi * dpow(n, 7. / 3.));
  const D A = (beta / gamma) / (dexp(-ec / (gamma * phi3)) - 1.);
  const D At2 = A * t2;
  const D H = gamma * phi3 *
                 dlog1p((beta / gamma) * t2 * (1. + At2) / (1. + At2 + At2 * At2));
  return e + n * (ec + H);
}

}  // namespace

extern "C" {

int oracle_functional_pol(int func, int64_t n, const double* rho, const double* sigma,
                          double* eps, double* vrho, double* vsigma) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  const double sig_tol2 = 2.1544346900318838e-43 * 2.1544346900318838e-43;
  for (int64_t i = 0; i < n; ++i) {
    const double ra = std::max(rho[2 * i], 0.), rb = std::max(rho[2 * i + 1], 0.);
    double saa = 0., sab = 0., sbb = 0.;
    if (gga) {
      saa = std::max(sigma[3 * i], sig_tol2);
      sbb = std::max(sigma[3 * i + 2], sig_tol2);
      // |sigma_ab| <= (sigma_aa + sigma_bb) / 2 keeps sigma_tot >= 0.
      sab = std::max(std::min(sigma[3 * i + 1], 0.5 * (saa + sbb)), -0.5 * (saa + sbb));
    }
    const Dual e = pol_energy(func, Dual::var(ra, 0), Dual::var(rb, 1), Dual::var(saa, 2),
                              Dual::var(sab, 3), Dual::var(sbb, 4));
    const double nt = ra + rb;
    eps[i] = nt > 0. ? e.v / nt : 0.;
    vrho[2 * i] = e.d[0];
    vrho[2 * i + 1] = e.d[1];
    if (gga) {
      vsigma[3 * i] = e.d[2];
      vsigma[3 * i + 1] = e.d[3];
      vsigma[3 * i + 2] = e.d[4];
    }
  }
  return 0;
}

}  // extern "C"

namespace {

}  // namespace

extern "C" {

void oracle_ssf_weights(int32_t natoms, const double* xyz, int32_t ntasks,
                        const oracle_task* tasks, const double* points, double* weights,
                        int nthreads) {
  // MolMeta::compute_rab / compute_dist_nearest (src/molmeta.cxx:28-60).
  std::vector<double> rab((size_t)natoms * natoms, 0.);
  for (int i = 0; i < natoms; ++i)
    for (int j = 0; j < natoms; ++j) {
      const double dx = xyz[3 * i] - xyz[3 * j], dy = xyz[3 * i + 1] - xyz[3 * j + 1],
                   dz = xyz[3 * i + 2] - xyz[3 * j + 2];
      rab[(size_t)i + (size_t)j * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> part(natoms), dist(natoms);
#pragma omp for schedule(dynamic)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const double cutoff = 0.5 * (1. - kSSFa) * t.dist_nearest;
      for (int32_t ip = 0; ip < t.npts; ++ip) {
        const double* p = points + 3 * (t.point_offset + ip);
        double& w = weights[t.point_offset + ip];
        for (int a = 0; a < natoms; ++a) {
          const double dx = p[0] - xyz[3 * a], dy = p[1] - xyz[3 * a + 1],
                       dz = p[2] - xyz[3 * a + 2];
          dist[a] = std::sqrt(dx * dx + dy * dy + dz * dz);
        }
        if (dist[t.iparent] < cutoff) continue;  // whole weight stays with the parent
        std::fill(part.begin(), part.end(), 1.);
        for (int i = 0; i < natoms; ++i)
          for (int j = 0; j < i; ++j) {
            if (!(part[i] > kSSFTol || part[j] > kSSFTol)) continue;
            const double mu = (dist[i] - dist[j]) / rab[(size_t)j + (size_t)i * natoms];
            if (mu <= -kSSFa) {
              part[j] = 0.;
            } else if (mu >= kSSFa) {
              part[i] = 0.;
            } else {
              const double g = 0.5 * (1. - frisch_g(mu));
              part[i] *= g;
              part[j] *= 1. - g;
            }
          }
        double sum = 0.;
        for (int a = 0; a < natoms; ++a) sum += part[a];
        w *= part[t.iparent] / sum;
      }
    }
  }
}

void oracle_collocation(const oracle_shell* shells, int32_t nlist, const int32_t* shell_list,
                        int32_t npts, const double* points, double* b, double* dx, double* dy,
                        double* dz) {
  int32_t nbe = 0;
  for (int32_t i = 0; i < nlist; ++i) {
    const auto& sh = shells[shell_list[i]];
    nbe += sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2;
  }
  const auto& sph = sph_tables();
  std::vector<double> cv, cgx, cgy, cgz;
  for (int32_t ip = 0; ip < npts; ++ip) {
    int32_t off = 0;
    for (int32_t i = 0; i < nlist; ++i) {
      const auto& sh = shells[shell_list[i]];
      const double x = points[3 * ip] - sh.origin[0];
      const double y = points[3 * ip + 1] - sh.origin[1];
      const double z = points[3 * ip + 2] - sh.origin[2];
      const double r2 = x * x + y * y + z * z;
      double R = 0., dR = 0.;  // dR = sum -2 a c exp(-a r^2)
      for (int k = 0; k < sh.nprim; ++k) {
        const double e = sh.coeff[k] * std::exp(-sh.alpha[k] * r2);
        R += e;
        dR += -2. * sh.alpha[k] * e;
      }
      const double Rx = dR * x, Ry = dR * y, Rz = dR * z;
      const int l = sh.l;
      const int ncart = (l + 1) * (l + 2) / 2;
      cv.assign(ncart, 0.);
      cgx.assign(ncart, 0.);
      cgy.assign(ncart, 0.);
      cgz.assign(ncart, 0.);
      // Cartesian CCA order: a from l down to 0, then b from l-a down to 0.
      int ic = 0;
      for (int a = l; a >= 0; --a)
        for (int bb = l - a; bb >= 0; --bb, ++ic)
          monomial_eval(a, bb, l - a - bb, x, y, z, R, Rx, Ry, Rz, &cv[ic],
                        dx ? &cgx[ic] : nullptr, &cgy[ic], &cgz[ic]);
      const size_t row = (size_t)ip * nbe;
      if (!sh.pure) {
        for (int c = 0; c < ncart; ++c) {
          b[row + off + c] = cv[c];
          if (dx) {
            dx[row + off + c] = cgx[c];
            dy[row + off + c] = cgy[c];
            dz[row + off + c] = cgz[c];
          }
        }
        off += ncart;
      } else {
        for (int m = 0; m < 2 * l + 1; ++m) {
          double v = 0., gx = 0., gy = 0., gz = 0.;
          for (const auto& tm : sph.t[l][m]) {
            double mv, mx, my, mz;
            monomial_eval(tm.a, tm.b, tm.c, x, y, z, R, Rx, Ry, Rz, &mv, &mx, &my, &mz);
            v += tm.coef * mv;
            gx += tm.coef * mx;
            gy += tm.coef * my;
            gz += tm.coef * mz;
          }
          b[row + off + m] = v;
          if (dx) {
            dx[row + off + m] = gx;
            dy[row + off + m] = gy;
            dz[row + off + m] = gz;
          }
        }
        off += 2 * l + 1;
      }
    }
  }
}

int oracle_functional(int func, int64_t n, const double* rho, const double* sigma, double* eps,
                      double* vrho, double* vsigma) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  for (int64_t i = 0; i < n; ++i) {
    const FuncOut o = eval_func(func, rho[i], sigma ? sigma[i] : 0.);
    eps[i] = rho[i] > 0. ? o.e / rho[i] : 0.;
    vrho[i] = o.vr;
    if (vsigma) vsigma[i] = o.vs;
  }
  return 0;
}

int oracle_use_blas(const char* path, char* which, int nwhich) {
  g_dgemm = nullptr;
  g_dsyr2k = nullptr;
  if (which && nwhich > 0) which[0] = 0;
  if (!path || !path[0]) return 0;
  void* h = dlopen(path, RTLD_NOW | RTLD_LOCAL);
  if (!h) return 1;
  // MKL: sequential BLAS inside the OpenMP task loop (as GauXC's host
  // integrator, which threads over tasks).
  using set_layer_t = int (*)(int);
  if (auto set_layer = (set_layer_t)dlsym(h, "MKL_Set_Threading_Layer")) set_layer(1);
  using set_threads_t = void (*)(int);
  if (auto set_threads = (set_threads_t)dlsym(h, "openblas_set_num_threads")) set_threads(1);
  auto gemm = (dgemm_t)dlsym(h, "dgemm_");
  auto syr2k = (dsyr2k_t)dlsym(h, "dsyr2k_");
  if (!gemm || !syr2k) {
    dlclose(h);
    return 2;
  }
  g_dgemm = gemm;
  g_dsyr2k = syr2k;
  if (which && nwhich > 0) std::snprintf(which, (size_t)nwhich, "%s", path);
  return 0;
}

int oracle_exc_vxc(int func, int32_t nbf, const oracle_shell* shells, const double* P,
                   int64_t ldp, int32_t ntasks, const oracle_task* tasks, const double* points,
                   const double* weights, const int32_t* shell_lists, double* exc, double* nel,
                   double* vxc, int64_t ldvxc, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  // Zero VXC (host_integrator_exc_vxc.hpp:173-195).
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = 0; i < nbf; ++i) vxc[i + (size_t)j * ldvxc] = 0.;
  // First AO of each shell (BasisSetMap).
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
  }
  double EXC = 0., NEL = 0.;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> B, Bx, By, Bz, X, Z, Psub, V, rho, sig, gx, gy, gz, eps, vr, vs;
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic) reduction(+ : EXC, NEL)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      // Local -> global basis-function index (the compressed submatrix map of
      // integrator_common.cxx:21-144 expressed per function).
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      B.assign((size_t)npts * nbe, 0.);
      if (gga) {
        Bx.assign(B.size(), 0.);
        By.assign(B.size(), 0.);
        Bz.assign(B.size(), 0.);
      }
      // eval_collocation(_gradient) (host_integrator_exc_vxc.hpp:378-380).
      oracle_collocation(shells, t.nshells, sl, npts, pts, B.data(), gga ? Bx.data() : nullptr,
                         gga ? By.data() : nullptr, gga ? Bz.data() : nullptr);
      // eval_xmat: X = 2 * B * P_sub (reference_local_host_work_driver.cxx:123-146).
      Psub.resize((size_t)nbe * nbe);
      for (int32_t j = 0; j < nbe; ++j)
        for (int32_t i = 0; i < nbe; ++i)
          Psub[(size_t)i + (size_t)j * nbe] = P[bf[i] + (size_t)bf[j] * ldp];
      X.assign((size_t)npts * nbe, 0.);
      if (g_dgemm) {
        // B is [npts][nbe] = column-major nbe x npts: X = 2 P_sub B.
        const double two = 2.0, zero = 0.0;
        g_dgemm("N", "N", &nbe, &npts, &nbe, &two, Psub.data(), &nbe, B.data(), &nbe, &zero,
                X.data(), &nbe);
      } else {
        for (int32_t p = 0; p < npts; ++p) {
          double* xr = X.data() + (size_t)p * nbe;
          const double* br = B.data() + (size_t)p * nbe;
          for (int32_t k = 0; k < nbe; ++k) {
            const double bk = 2.0 * br[k];
            const double* pc = Psub.data() + (size_t)k * nbe;  // P_sub symmetric: row k
            for (int32_t j = 0; j < nbe; ++j) xr[j] += bk * pc[j];
          }
        }
      }
      // eval_uvvar_{lda,gga}_rks (:150-163, :242-268).
      rho.assign(npts, 0.);
      sig.assign(npts, 0.);
      gx.assign(npts, 0.);
      gy.assign(npts, 0.);
      gz.assign(npts, 0.);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        double r = 0., ax = 0., ay = 0., az = 0.;
        for (int32_t k = 0; k < nbe; ++k) {
          r += B[o + k] * X[o + k];
          if (gga) {
            ax += Bx[o + k] * X[o + k];
            ay += By[o + k] * X[o + k];
            az += Bz[o + k] * X[o + k];
          }
        }
        rho[p] = r;
        if (gga) {
          gx[p] = 2. * ax;
          gy[p] = 2. * ay;
          gz[p] = 2. * az;
          sig[p] = gx[p] * gx[p] + gy[p] * gy[p] + gz[p] * gz[p];
        }
      }
      // Functional, weight folding, EXC/NEL (:445-502).
      eps.resize(npts);
      vr.resize(npts);
      vs.assign(npts, 0.);
      oracle_functional(func, npts, rho.data(), gga ? sig.data() : nullptr, eps.data(), vr.data(),
                        gga ? vs.data() : nullptr);
      for (int32_t p = 0; p < npts; ++p) {
        eps[p] *= w[p];
        vr[p] *= w[p];
        vs[p] *= w[p];
        NEL += w[p] * rho[p];
        EXC += eps[p] * rho[p];
      }
      // Z matrix (:586-604 LDA, :678-713 GGA).
      Z.resize((size_t)npts * nbe);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        const double fl = 0.5 * vr[p];
        const double fx = 2. * vs[p] * gx[p], fy = 2. * vs[p] * gy[p], fz = 2. * vs[p] * gz[p];
        for (int32_t k = 0; k < nbe; ++k) {
          double zv = fl * B[o + k];
          if (gga) zv += fx * Bx[o + k] + fy * By[o + k] + fz * Bz[o + k];
          Z[o + k] = zv;
        }
      }
      // inc_vxc: V = B^T Z + Z^T B (lower), scattered with atomics (:1678-1692,
      // host/util.hpp:130-168).
      V.assign((size_t)nbe * nbe, 0.);
      if (g_dsyr2k) {
        // Column-major nbe x npts operands: the upper triangle of B Z^T + Z B^T
        // is V[i * nbe + j] for j <= i, the layout of the loops below.
        const double one = 1.0, zero = 0.0;
        g_dsyr2k("U", "N", &nbe, &npts, &one, B.data(), &nbe, Z.data(), &nbe, &zero, V.data(),
                 &nbe);
      } else {
        for (int32_t p = 0; p < npts; ++p) {
          const size_t o = (size_t)p * nbe;
          for (int32_t i = 0; i < nbe; ++i) {
            const double bi = B[o + i], zi = Z[o + i];
            double* vrow = V.data() + (size_t)i * nbe;
            for (int32_t j = 0; j <= i; ++j) vrow[j] += bi * Z[o + j] + zi * B[o + j];
          }
        }
      }
      for (int32_t i = 0; i < nbe; ++i)
        for (int32_t j = 0; j <= i; ++j) {
#pragma omp atomic
          vxc[bf[i] + (size_t)bf[j] * ldvxc] += V[(size_t)i * nbe + j];
        }
    }
  }
  // Symmetrise lower -> upper (:577-599).
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = j + 1; i < nbf; ++i) vxc[j + (size_t)i * ldvxc] = vxc[i + (size_t)j * ldvxc];
  *exc = EXC;
  *nel = NEL;
  return 0;
}

}  // extern "C"

extern "C" {

// UKS EXC/VXC local work: GauXC's generic host driver with spin_dim 2
// (reference_replicated_xc_host_integrator_exc_vxc.hpp:106-601, is_uks):
// X_s = B P_s and X_z = B P_z (factor 1, :387-396); rho_+- = (rho_s +- rho_z)/2
// and the three sigmas of eval_uvvar_{lda,gga}_uks
// (reference_local_host_work_driver.cxx:166-187, 270-330); the polarised
// functional, weight folding (:452-470); Z_s, Z_z of eval_zmat_{lda,gga}_vxc_uks
// (:607-632, 715-770); VXC_s = B^T Z_s + Z_s^T B and VXC_z likewise (inc_vxc),
// symmetrised.
int oracle_exc_vxc_uks(int func, int32_t nbf, const oracle_shell* shells, const double* Ps,
                       int64_t ldps, const double* Pz, int64_t ldpz, int32_t ntasks,
                       const oracle_task* tasks, const double* points, const double* weights,
                       const int32_t* shell_lists, double* exc, double* nel, double* vxc_s,
                       double* vxc_z, int64_t ldvxc, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = 0; i < nbf; ++i) vxc_s[i + (size_t)j * ldvxc] = vxc_z[i + (size_t)j * ldvxc] = 0.;
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
  }
  double EXC = 0., NEL = 0.;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> B, Bx, By, Bz, X[2], Z[2], Psub[2], V, rho, sig, eps, vr, vs;
    std::vector<double> g[2][3];
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic) reduction(+ : EXC, NEL)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      B.assign((size_t)npts * nbe, 0.);
      if (gga) {
        Bx.assign(B.size(), 0.);
        By.assign(B.size(), 0.);
        Bz.assign(B.size(), 0.);
      }
      oracle_collocation(shells, t.nshells, sl, npts, pts, B.data(), gga ? Bx.data() : nullptr,
                         gga ? By.data() : nullptr, gga ? Bz.data() : nullptr);
      const double* Pm[2] = {Ps, Pz};
      const int64_t ldm[2] = {ldps, ldpz};
      for (int c = 0; c < 2; ++c) {
        Psub[c].resize((size_t)nbe * nbe);
        for (int32_t j = 0; j < nbe; ++j)
          for (int32_t i = 0; i < nbe; ++i)
            Psub[c][(size_t)i + (size_t)j * nbe] = Pm[c][bf[i] + (size_t)bf[j] * ldm[c]];
        X[c].assign((size_t)npts * nbe, 0.);
        for (int32_t p = 0; p < npts; ++p) {
          double* xr = X[c].data() + (size_t)p * nbe;
          const double* br = B.data() + (size_t)p * nbe;
          for (int32_t k = 0; k < nbe; ++k) {
            const double bk = br[k];
            const double* pc = Psub[c].data() + (size_t)k * nbe;
            for (int32_t j = 0; j < nbe; ++j) xr[j] += bk * pc[j];
          }
        }
      }
      rho.assign(2 * (size_t)npts, 0.);
      sig.assign(3 * (size_t)npts, 0.);
      for (int c = 0; c < 2; ++c)
        for (int d = 0; d < 3; ++d) g[c][d].assign(npts, 0.);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        double r[2] = {0., 0.}, gr[2][3] = {{0., 0., 0.}, {0., 0., 0.}};
        for (int c = 0; c < 2; ++c)
          for (int32_t k = 0; k < nbe; ++k) {
            const double x = X[c][o + k];
            r[c] += B[o + k] * x;
            if (gga) {
              gr[c][0] += Bx[o + k] * x;
              gr[c][1] += By[o + k] * x;
              gr[c][2] += Bz[o + k] * x;
            }
          }
        rho[2 * p] = 0.5 * (r[0] + r[1]);
        rho[2 * p + 1] = 0.5 * (r[0] - r[1]);
        if (gga) {
          double nn = 0., mm = 0., nm = 0.;
          for (int d = 0; d < 3; ++d) {
            g[0][d][p] = 2. * gr[0][d];  // grad n
            g[1][d][p] = 2. * gr[1][d];  // grad m_z
            nn += g[0][d][p] * g[0][d][p];
            mm += g[1][d][p] * g[1][d][p];
            nm += g[0][d][p] * g[1][d][p];
          }
          sig[3 * p] = 0.25 * (nn + mm) + 0.5 * nm;
          sig[3 * p + 1] = 0.25 * (nn - mm);
          sig[3 * p + 2] = 0.25 * (nn + mm) - 0.5 * nm;
        }
      }
      eps.resize(npts);
      vr.resize(2 * (size_t)npts);
      vs.assign(3 * (size_t)npts, 0.);
      oracle_functional_pol(func, npts, rho.data(), gga ? sig.data() : nullptr, eps.data(),
                            vr.data(), gga ? vs.data() : nullptr);
      for (int32_t p = 0; p < npts; ++p) {
        const double n = rho[2 * p] + rho[2 * p + 1];
        NEL += w[p] * n;
        EXC += w[p] * eps[p] * n;
      }
      for (int c = 0; c < 2; ++c) Z[c].resize((size_t)npts * nbe);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        const double fp = 0.5 * w[p] * vr[2 * p], fm = 0.5 * w[p] * vr[2 * p + 1];
        const double as = 0.5 * (fp + fm), az = 0.5 * (fp - fm);
        double cs[3] = {0., 0., 0.}, cz[3] = {0., 0., 0.};
        if (gga) {
          const double vpp = w[p] * vs[3 * p], vpm = w[p] * vs[3 * p + 1],
                       vmm = w[p] * vs[3 * p + 2];
          const double g1 = 0.5 * (vpp + vpm + vmm), g2 = 0.5 * (vpp - vmm),
                       g3 = 0.5 * (vpp - vpm + vmm);
          for (int d = 0; d < 3; ++d) {
            cs[d] = g1 * g[0][d][p] + g2 * g[1][d][p];
            cz[d] = g3 * g[1][d][p] + g2 * g[0][d][p];
          }
        }
        for (int32_t k = 0; k < nbe; ++k) {
          double zs = as * B[o + k], zz = az * B[o + k];
          if (gga) {
            zs += cs[0] * Bx[o + k] + cs[1] * By[o + k] + cs[2] * Bz[o + k];
            zz += cz[0] * Bx[o + k] + cz[1] * By[o + k] + cz[2] * Bz[o + k];
          }
          Z[0][o + k] = zs;
          Z[1][o + k] = zz;
        }
      }
      double* out[2] = {vxc_s, vxc_z};
      for (int c = 0; c < 2; ++c) {
        V.assign((size_t)nbe * nbe, 0.);
        for (int32_t p = 0; p < npts; ++p) {
          const size_t o = (size_t)p * nbe;
          for (int32_t i = 0; i < nbe; ++i) {
            const double bi = B[o + i], zi = Z[c][o + i];
            double* vrow = V.data() + (size_t)i * nbe;
            for (int32_t j = 0; j <= i; ++j) vrow[j] += bi * Z[c][o + j] + zi * B[o + j];
          }
        }
        for (int32_t i = 0; i < nbe; ++i)
          for (int32_t j = 0; j <= i; ++j) {
#pragma omp atomic
            out[c][bf[i] + (size_t)bf[j] * ldvxc] += V[(size_t)i * nbe + j];
          }
      }
    }
  }
  for (double* v : {vxc_s, vxc_z})
    for (int32_t j = 0; j < nbf; ++j)
      for (int32_t i = j + 1; i < nbf; ++i) v[j + (size_t)i * ldvxc] = v[i + (size_t)j * ldvxc];
  *exc = EXC;
  *nel = NEL;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// FXC contraction (RKS): GauXC's host driver
// reference_replicated_xc_host_integrator_fxc_contraction.hpp:94-600 with
// spin_dim 1.  Per task: X = 2 P_sub B and tX = 2 tP_sub B (:393-396,
// :441-444); rho, grad rho from X and trho, grad trho from tX
// (eval_uvvar_{lda,gga}_rks); the functional's first and second derivatives
// at rho (:433-439); the trial kernels of eval_tmat_{lda,gga}_vxc_rks
// (reference_local_host_work_driver.cxx:1114-1143):
//   tgamma = grad trho . grad rho,
//   A = v2rho2 trho + 2 v2rhogamma tgamma,
//   B = 2 (v2rhogamma trho + 2 v2gamma2 tgamma) grad rho + 2 vgamma grad trho,
// weight folding (:497-517); Z = A/2 B + B . grad B (eval_zmat_gga_vxc_rks_ts,
// :1382-1408); FXC = B^T Z + Z^T B (inc_vxc), symmetrised (:568-576).
// ---------------------------------------------------------------------------
namespace {

// vrho, vsigma and the directional second derivatives
//   A = v2rho2 drho + v2rhosigma dsig,  Bc = v2rhosigma drho + v2sigma2 dsig
// of the unpolarised functional at (rho, sigma), by nested duals over the
// spin-polarised energy at zeta = 0 (rho_a = rho_b = rho / 2, sigma_aa =
// sigma_ab = sigma_bb = sigma / 4): inner slots d/drho (0) and d/dsigma (1),
// outer slot 0 the direction (drho, dsig).
void fxc_point(int func, double rho, double sigma, double drho, double dsig, double& vr,
               double& vs, double& A, double& Bc) {
  using D2 = DualT<Dual>;
  const bool gga = func != ORACLE_SVWN5;
  const double sig_tol2 = 2.1544346900318838e-43 * 2.1544346900318838e-43;
  D2 R, S;
  R.v = Dual::var(std::max(rho, 0.), 0);
  R.d[0] = Dual(drho);
  S.v = Dual::var(gga ? std::max(sigma, sig_tol2) : 0., 1);
  S.d[0] = Dual(gga ? dsig : 0.);
  const D2 e = pol_energy(func, 0.5 * R, 0.5 * R, 0.25 * S, 0.25 * S, 0.25 * S);
  vr = e.v.d[0];
  vs = e.v.d[1];
  A = e.d[0].d[0];
  Bc = e.d[0].d[1];
}

// out[p][j] = fac * sum_k B[p][k] M_sub[k][j] (M_sub symmetric), the
// eval_xmat product of reference_local_host_work_driver.cxx:123-146.
void xmat(const std::vector<double>& B, const double* M, int64_t ldm,
          const std::vector<int32_t>& bf, int32_t npts, double fac, std::vector<double>& sub,
          std::vector<double>& out) {
  const int32_t nbe = (int32_t)bf.size();
  sub.resize((size_t)nbe * nbe);
  for (int32_t j = 0; j < nbe; ++j)
    for (int32_t i = 0; i < nbe; ++i) sub[(size_t)i + (size_t)j * nbe] = M[bf[i] + (size_t)bf[j] * ldm];
  out.assign((size_t)npts * nbe, 0.);
  if (g_dgemm) {
    const double zero = 0.0;
    g_dgemm("N", "N", &nbe, &npts, &nbe, &fac, sub.data(), &nbe, B.data(), &nbe, &zero,
            out.data(), &nbe);
    return;
  }
  for (int32_t p = 0; p < npts; ++p) {
    double* xr = out.data() + (size_t)p * nbe;
    const double* br = B.data() + (size_t)p * nbe;
    for (int32_t k = 0; k < nbe; ++k) {
      const double bk = fac * br[k];
      const double* pc = sub.data() + (size_t)k * nbe;
      for (int32_t j = 0; j < nbe; ++j) xr[j] += bk * pc[j];
    }
  }
}

}  // namespace

extern "C" {

int oracle_functional_fxc(int func, int64_t n, const double* rho, const double* sigma,
                          const double* drho, const double* dsigma, double* vrho,
                          double* vsigma, double* a, double* b) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int64_t i = 0; i < n; ++i) {
    double vr, vs, A, Bc;
    fxc_point(func, rho[i], gga ? sigma[i] : 0., drho[i], gga ? dsigma[i] : 0., vr, vs, A, Bc);
    vrho[i] = vr;
    a[i] = A;
    if (gga) {
      vsigma[i] = vs;
      b[i] = Bc;
    }
  }
  return 0;
}

// Spin-polarised point kernel of the UKS FXC contraction: the derivatives
// of the polarised vrho (a, b) and vsigma (aa, ab, bb) along the direction
// (dra, drb, dsaa, dsab, dsbb), by nested duals over pol_energy (inner slots
// the five variables, outer slot the direction).  out: vr[2], vs[3],
// dvr[2], dvs[3].
void fxc_point_pol(int func, const double* x, const double* dx, double* out) {
  using D2 = DualT<Dual>;
  D2 v[5];
  for (int k = 0; k < 5; ++k) {
    v[k].v = Dual::var(x[k], k);
    v[k].d[0] = Dual(dx[k]);
  }
  const D2 e = pol_energy(func, v[0], v[1], v[2], v[3], v[4]);
  for (int k = 0; k < 5; ++k) {
    out[k] = e.v.d[k];
    out[5 + k] = e.d[0].d[k];
  }
}

int oracle_fxc_contraction(int func, int32_t nbf, const oracle_shell* shells, const double* P,
                           int64_t ldp, const double* tP, int64_t ldtp, int32_t ntasks,
                           const oracle_task* tasks, const double* points,
                           const double* weights, const int32_t* shell_lists, double* nel,
                           double* fxc, int64_t ldfxc, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = 0; i < nbf; ++i) fxc[i + (size_t)j * ldfxc] = 0.;
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
  }
  double NEL = 0.;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> B, Bx, By, Bz, X, tX, Z, sub, V;
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic) reduction(+ : NEL)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      B.assign((size_t)npts * nbe, 0.);
      if (gga) {
        Bx.assign(B.size(), 0.);
        By.assign(B.size(), 0.);
        Bz.assign(B.size(), 0.);
      }
      oracle_collocation(shells, t.nshells, sl, npts, pts, B.data(), gga ? Bx.data() : nullptr,
                         gga ? By.data() : nullptr, gga ? Bz.data() : nullptr);
      xmat(B, P, ldp, bf, npts, 2.0, sub, X);
      xmat(B, tP, ldtp, bf, npts, 2.0, sub, tX);
      Z.resize((size_t)npts * nbe);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        double r = 0., tr = 0., g[3] = {0., 0., 0.}, tg[3] = {0., 0., 0.};
        for (int32_t k = 0; k < nbe; ++k) {
          r += B[o + k] * X[o + k];
          tr += B[o + k] * tX[o + k];
          if (gga) {
            g[0] += Bx[o + k] * X[o + k];
            g[1] += By[o + k] * X[o + k];
            g[2] += Bz[o + k] * X[o + k];
            tg[0] += Bx[o + k] * tX[o + k];
            tg[1] += By[o + k] * tX[o + k];
            tg[2] += Bz[o + k] * tX[o + k];
          }
        }
        double sig = 0., tgam = 0.;
        for (int d = 0; d < 3; ++d) {
          g[d] *= 2.;
          tg[d] *= 2.;
          sig += g[d] * g[d];
          tgam += g[d] * tg[d];
        }
        double vr, vs, A, Bc;
        fxc_point(func, r, sig, tr, 2. * tgam, vr, vs, A, Bc);
        NEL += w[p] * r;
        const double fa = 0.5 * w[p] * A;
        double c[3] = {0., 0., 0.};
        if (gga)
          for (int d = 0; d < 3; ++d) c[d] = w[p] * (2. * Bc * g[d] + 2. * vs * tg[d]);
        for (int32_t k = 0; k < nbe; ++k) {
          double zv = fa * B[o + k];
          if (gga) zv += c[0] * Bx[o + k] + c[1] * By[o + k] + c[2] * Bz[o + k];
          Z[o + k] = zv;
        }
      }
      V.assign((size_t)nbe * nbe, 0.);
      if (g_dsyr2k) {
        const double one = 1.0, zero = 0.0;
        g_dsyr2k("U", "N", &nbe, &npts, &one, B.data(), &nbe, Z.data(), &nbe, &zero, V.data(),
                 &nbe);
      } else {
        for (int32_t p = 0; p < npts; ++p) {
          const size_t o = (size_t)p * nbe;
          for (int32_t i = 0; i < nbe; ++i) {
            const double bi = B[o + i], zi = Z[o + i];
            double* vrow = V.data() + (size_t)i * nbe;
            for (int32_t j = 0; j <= i; ++j) vrow[j] += bi * Z[o + j] + zi * B[o + j];
          }
        }
      }
      for (int32_t i = 0; i < nbe; ++i)
        for (int32_t j = 0; j <= i; ++j) {
#pragma omp atomic
          fxc[bf[i] + (size_t)bf[j] * ldfxc] += V[(size_t)i * nbe + j];
        }
    }
  }
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = j + 1; i < nbf; ++i) fxc[j + (size_t)i * ldfxc] = fxc[i + (size_t)j * ldfxc];
  if (nel) *nel = NEL;
  return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// EXC nuclear gradient (RKS) with the grid-weight derivatives: GauXC's host
// driver reference_replicated_xc_host_integrator_exc_grad.hpp:106-600
// (include_weight_derivatives = true) and the SSF weight-derivative
// contraction of weights.cxx:806-990.
// ---------------------------------------------------------------------------
namespace {

// Monomial x^a y^b z^c times the radial part, with gradient and Hessian.
// With R(r^2), R1 = dR/d(r^2) * 2 = -2 sum a c e, R2 = 4 sum a^2 c e:
//   d_i f = d_i m R + m x_i R1,
//   d_i d_j f = d_i d_j m R + (d_i m x_j + d_j m x_i) R1 + m (delta_ij R1 + x_i x_j R2).
// out: v, gx, gy, gz, hxx, hxy, hxz, hyy, hyz, hzz.
void monomial_hess(int a, int b, int c, double x, double y, double z, double R, double R1,
                   double R2, double* out) {
  const int e[3] = {a, b, c};
  const double r[3] = {x, y, z};
  auto p = [](double t, int n) { return n < 0 ? 0. : ipow(t, n); };
  // d^k/dt^k t^n = n (n-1) .. t^(n-k)
  auto d1 = [&](int d) {
    double v = 1.;
    for (int k = 0; k < 3; ++k) v *= (k == d) ? (e[k] ? e[k] * p(r[k], e[k] - 1) : 0.) : p(r[k], e[k]);
    return v;
  };
  auto d2 = [&](int d, int f) {
    double v = 1.;
    for (int k = 0; k < 3; ++k) {
      const int n = e[k] - (k == d) - (k == f);
      double coef = 1.;
      for (int q = 0; q < (k == d) + (k == f); ++q) coef *= (e[k] - q);
      v *= n < 0 ? 0. : coef * p(r[k], n);
    }
    return v;
  };
  const double m = p(x, a) * p(y, b) * p(z, c);
  out[0] = m * R;
  double g[3];
  for (int d = 0; d < 3; ++d) {
    g[d] = d1(d);
    out[1 + d] = g[d] * R + m * r[d] * R1;
  }
  int k = 4;
  for (int d = 0; d < 3; ++d)
    for (int f = d; f < 3; ++f, ++k)
      out[k] = d2(d, f) * R + (g[d] * r[f] + g[f] * r[d]) * R1 +
               m * ((d == f ? R1 : 0.) + r[d] * r[f] * R2);
}

// SSF weight-derivative contraction of one task (weights.cxx:806-990):
// d/dR_B of sum_i w_i f_i with w_i = w0_i P_A(r_i) / sum_C P_C(r_i), points
// moving with their parent atom A (its row follows by translational
// invariance).  wf_i = w_i f_i.  rab = atom distance matrix.
void ssf_grad_contract(int32_t natoms, const double* xyz, const double* rab, const oracle_task& t,
                       const double* pts, const double* wf, double* grad) {
  const double bound = kSSFa - 1e-4;
  // d ln s(mu) / d mu for s(mu) = (1 - g(mu / a)) / 2, g the Frisch
  // polynomial: 1 - g(v) = (1 - v)^4 (5 v^3 + 20 v^2 + 29 v + 16) / 16 and
  // g'(v) = 35 (1 - v^2)^3 / 16, so t = 35 (1 + v)^3 / ((mu - a)(5 v^3 + 20 v^2 + 29 v + 16)).
  auto tfr = [](double mu) {
    const double v = mu / kSSFa;
    return 35. * (1. + v) * (1. + v) * (1. + v) /
           ((mu - kSSFa) * (((5. * v + 20.) * v + 29.) * v + 16.));
  };
  std::vector<double> part(natoms), dist(natoms);
  const int32_t A = t.iparent;
  auto X = [&](int at, int d) { return xyz[3 * at + d]; };
  for (int32_t i = 0; i < t.npts; ++i) {
    const double f = wf[i];
    if (std::fabs(f) < 1e-12) continue;
    const double* r = pts + 3 * i;
    for (int at = 0; at < natoms; ++at) {
      const double dx = r[0] - X(at, 0), dy = r[1] - X(at, 1), dz = r[2] - X(at, 2);
      dist[at] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (dist[A] < 0.5 * (1. - kSSFa) * t.dist_nearest) continue;
    std::fill(part.begin(), part.end(), 1.);
    for (int a = 0; a < natoms; ++a)
      for (int b = 0; b < a; ++b) {
        if (!(part[a] > kSSFTol || part[b] > kSSFTol)) continue;
        const double mu = (dist[a] - dist[b]) / rab[(size_t)b + (size_t)a * natoms];
        if (mu <= -kSSFa) {
          part[b] = 0.;
        } else if (mu >= kSSFa) {
          part[a] = 0.;
        } else {
          const double g = 0.5 * (1. - frisch_g(mu));
          part[a] *= g;
          part[b] *= 1. - g;
        }
      }
    double sum = 0.;
    for (int a = 0; a < natoms; ++a) sum += part[a];
    for (int B = 0; B < natoms; ++B) {
      if (B == A) continue;
      double gb[3] = {0., 0., 0.};
      const double rAB = rab[(size_t)B + (size_t)A * natoms];
      const double muAB = (dist[A] - dist[B]) / rAB;
      if (std::fabs(muAB) < bound) {
        const double c1 = tfr(muAB) / rAB * (part[A] - sum) / sum * f / dist[B];
        for (int d = 0; d < 3; ++d)
          gb[d] = c1 * ((X(B, d) - r[d]) + muAB * (X(B, d) - X(A, d)) / rAB * dist[B]);
      }
      if (part[B] > kSSFTol) {
        for (int C = 0; C < natoms; ++C) {
          if (C == B) continue;
          const double rBC = rab[(size_t)C + (size_t)B * natoms];
          const double muBC = (dist[B] - dist[C]) / rBC;
          if (std::fabs(muBC) >= bound) continue;
          const double c = part[B] * tfr(muBC) / rBC / sum * f;
          for (int d = 0; d < 3; ++d)
            gb[d] -= c * ((X(B, d) - r[d]) / dist[B] - muBC * (X(B, d) - X(C, d)) / rBC);
          if (C != A)
            for (int d = 0; d < 3; ++d) {
              const double gc = c * ((X(C, d) - r[d]) / dist[C] + muBC * (X(C, d) - X(B, d)) / rBC);
              grad[3 * C + d] += gc;
              grad[3 * A + d] -= gc;
            }
        }
      }
      for (int d = 0; d < 3; ++d) {
        grad[3 * B + d] += gb[d];
        grad[3 * A + d] -= gb[d];
      }
    }
  }
}

}  // namespace

extern "C" {

void oracle_collocation_hessian(const oracle_shell* shells, int32_t nlist,
                                const int32_t* shell_list, int32_t npts, const double* points,
                                double* const* out) {
  int32_t nbe = 0;
  for (int32_t i = 0; i < nlist; ++i) {
    const auto& sh = shells[shell_list[i]];
    nbe += sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2;
  }
  const auto& sph = sph_tables();
  double mv[10];
  for (int32_t ip = 0; ip < npts; ++ip) {
    int32_t off = 0;
    for (int32_t i = 0; i < nlist; ++i) {
      const auto& sh = shells[shell_list[i]];
      const double x = points[3 * ip] - sh.origin[0];
      const double y = points[3 * ip + 1] - sh.origin[1];
      const double z = points[3 * ip + 2] - sh.origin[2];
      const double r2 = x * x + y * y + z * z;
      double R = 0., R1 = 0., R2 = 0.;
      for (int k = 0; k < sh.nprim; ++k) {
        const double e = sh.coeff[k] * std::exp(-sh.alpha[k] * r2);
        R += e;
        R1 += -2. * sh.alpha[k] * e;
        R2 += 4. * sh.alpha[k] * sh.alpha[k] * e;
      }
      const int l = sh.l;
      const size_t row = (size_t)ip * nbe;
      if (!sh.pure) {
        int ic = 0;
        for (int a = l; a >= 0; --a)
          for (int bb = l - a; bb >= 0; --bb, ++ic) {
            monomial_hess(a, bb, l - a - bb, x, y, z, R, R1, R2, mv);
            for (int q = 0; q < 10; ++q) out[q][row + off + ic] = mv[q];
          }
        off += ic;
      } else {
        for (int m = 0; m < 2 * l + 1; ++m) {
          double acc[10] = {0., 0., 0., 0., 0., 0., 0., 0., 0., 0.};
          for (const auto& tm : sph.t[l][m]) {
            monomial_hess(tm.a, tm.b, tm.c, x, y, z, R, R1, R2, mv);
            for (int q = 0; q < 10; ++q) acc[q] += tm.coef * mv[q];
          }
          for (int q = 0; q < 10; ++q) out[q][row + off + m] = acc[q];
        }
        off += 2 * l + 1;
      }
    }
  }
}

int oracle_exc_grad(int func, int32_t nbf, const oracle_shell* shells, int32_t natoms,
                    const double* xyz, const double* P, int64_t ldp, int32_t ntasks,
                    const oracle_task* tasks, const double* points, const double* weights,
                    const int32_t* shell_lists, int weight_derivs, double* grad, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int k = 0; k < 3 * natoms; ++k) grad[k] = 0.;
  std::vector<double> rab((size_t)natoms * natoms);
  for (int a = 0; a < natoms; ++a)
    for (int b = 0; b < natoms; ++b) {
      const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
                   dz = xyz[3 * a + 2] - xyz[3 * b + 2];
      rab[(size_t)a + (size_t)b * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0), centre(nsh_total, -1);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
    // BasisSetMap: the atom at exactly the shell's origin (basisset_map.hpp:69-77).
    for (int a = 0; a < natoms && centre[s] < 0; ++a)
      if (xyz[3 * a] == sh.origin[0] && xyz[3 * a + 1] == sh.origin[1] &&
          xyz[3 * a + 2] == sh.origin[2])
        centre[s] = a;
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> pl[10], X[4], sub, wf, vr, vs, rho, sig, eps, gr[3];
    std::vector<double> gloc(3 * (size_t)natoms, 0.);
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      double* o[10];
      for (int q = 0; q < 10; ++q) {
        pl[q].assign((size_t)npts * nbe, 0.);
        o[q] = pl[q].data();
      }
      oracle_collocation_hessian(shells, t.nshells, sl, npts, pts, o);
      // X = 2 P_sub B and, for GGA, X_j = 2 P_sub dB/dj (:356-365).
      const int nx = gga ? 4 : 1;
      for (int q = 0; q < nx; ++q) xmat(pl[q], P, ldp, bf, npts, 2.0, sub, X[q]);
      rho.assign(npts, 0.);
      sig.assign(npts, 0.);
      for (int d = 0; d < 3; ++d) gr[d].assign(npts, 0.);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t off = (size_t)p * nbe;
        double r = 0., g[3] = {0., 0., 0.};
        for (int32_t k = 0; k < nbe; ++k) {
          r += pl[0][off + k] * X[0][off + k];
          for (int d = 0; d < 3; ++d) g[d] += pl[1 + d][off + k] * X[0][off + k];
        }
        rho[p] = r;
        for (int d = 0; d < 3; ++d) gr[d][p] = 2. * g[d];
        sig[p] = gr[0][p] * gr[0][p] + gr[1][p] * gr[1][p] + gr[2][p] * gr[2][p];
      }
      eps.resize(npts);
      vr.resize(npts);
      vs.assign(npts, 0.);
      oracle_functional(func, npts, rho.data(), gga ? sig.data() : nullptr, eps.data(), vr.data(),
                        gga ? vs.data() : nullptr);
      // Grid-weight part with w_i eps_i rho_i (:418-425).
      wf.resize(npts);
      if (weight_derivs) {
        for (int32_t p = 0; p < npts; ++p) wf[p] = eps[p] * rho[p] * w[p];
        ssf_grad_contract(natoms, xyz, rab.data(), t, pts, wf.data(), gloc.data());
      }
      // Basis-function part, shell by shell (:427-590); with the weight
      // derivatives, shells on the parent atom move with the grid and drop out.
      int32_t boff = 0;
      for (int32_t k = 0; k < t.nshells; ++k) {
        const int32_t s = sl[k];
        const int32_t n = first_ao[s + 1] - first_ao[s];
        const int32_t at = centre[s];
        if (at < 0 || (weight_derivs && at == t.iparent)) {
          boff += n;
          continue;
        }
        double g[3] = {0., 0., 0.};
        for (int32_t mu = boff; mu < boff + n; ++mu)
          for (int32_t p = 0; p < npts; ++p) {
            const size_t i = (size_t)p * nbe + mu;
            const double wv = w[p] * vr[p], xn = X[0][i];
            for (int d = 0; d < 3; ++d) g[d] += wv * xn * pl[1 + d][i];
            if (gga) {
              const double wg = w[p] * vs[p];
              const double dn[3] = {gr[0][p], gr[1][p], gr[2][p]};
              const double h[3][3] = {{pl[4][i], pl[5][i], pl[6][i]},
                                      {pl[5][i], pl[7][i], pl[8][i]},
                                      {pl[6][i], pl[8][i], pl[9][i]}};
              const double d11 = dn[0] * X[1][i] + dn[1] * X[2][i] + dn[2] * X[3][i];
              for (int d = 0; d < 3; ++d) {
                const double hd = h[d][0] * dn[0] + h[d][1] * dn[1] + h[d][2] * dn[2];
                g[d] += 2. * wg * (xn * hd + pl[1 + d][i] * d11);
              }
            }
          }
        for (int d = 0; d < 3; ++d) {
          gloc[3 * at + d] += -2. * g[d];
          if (weight_derivs) gloc[3 * t.iparent + d] -= -2. * g[d];
        }
        boff += n;
      }
    }
#pragma omp critical
    for (int k = 0; k < 3 * natoms; ++k) grad[k] += gloc[k];
  }
  return 0;
}

}  // extern "C"

extern "C" {

// UKS FXC contraction (GauXC's host driver with spin_dim 2,
// reference_replicated_xc_host_integrator_fxc_contraction.hpp:94-600):
// FXC_a,b = d VXC_a,b[P + e tP] / de through eval_tmat_{lda,gga}_vxc_uks and
// eval_zmat_{lda,gga}_vxc_uks_ts (reference_local_host_work_driver.cxx:
// 1118-1124, 1149-1215, 1354-1367, 1441-1481), returned as FXC_s = (FXC_a +
// FXC_b) / 2 and FXC_z = (FXC_a - FXC_b) / 2 (:577-600).  Written here as
// the derivative of the UKS Z coefficients of oracle_exc_vxc_uks along the
// trial direction, which is the same quantity.
int oracle_fxc_contraction_uks(int func, int32_t nbf, const oracle_shell* shells,
                               const double* Ps, const double* Pz, const double* tPs,
                               const double* tPz, int64_t ldp, int32_t ntasks,
                               const oracle_task* tasks, const double* points,
                               const double* weights, const int32_t* shell_lists,
                               double* fxc_s, double* fxc_z, int64_t ldfxc, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int32_t j = 0; j < nbf; ++j)
    for (int32_t i = 0; i < nbf; ++i) fxc_s[i + (size_t)j * ldfxc] = fxc_z[i + (size_t)j * ldfxc] = 0.;
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
  }
  const double sig_tol2 = 2.1544346900318838e-43 * 2.1544346900318838e-43;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> B, Bx, By, Bz, X[4], Z[2], sub, V;
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      B.assign((size_t)npts * nbe, 0.);
      if (gga) {
        Bx.assign(B.size(), 0.);
        By.assign(B.size(), 0.);
        Bz.assign(B.size(), 0.);
      }
      oracle_collocation(shells, t.nshells, sl, npts, pts, B.data(), gga ? Bx.data() : nullptr,
                         gga ? By.data() : nullptr, gga ? Bz.data() : nullptr);
      const double* Pm[4] = {Ps, Pz, tPs, tPz};
      for (int c = 0; c < 4; ++c) xmat(B, Pm[c], ldp, bf, npts, 1.0, sub, X[c]);
      for (int c = 0; c < 2; ++c) Z[c].resize((size_t)npts * nbe);
      for (int32_t p = 0; p < npts; ++p) {
        const size_t o = (size_t)p * nbe;
        double r[4] = {0., 0., 0., 0.}, g[4][3] = {};
        for (int c = 0; c < 4; ++c)
          for (int32_t k = 0; k < nbe; ++k) {
            const double xv = X[c][o + k];
            r[c] += B[o + k] * xv;
            if (gga) {
              g[c][0] += Bx[o + k] * xv;
              g[c][1] += By[o + k] * xv;
              g[c][2] += Bz[o + k] * xv;
            }
          }
        // n, m = P_s, P_z densities; tn, tm their trial counterparts.
        double gn[3], gm[3], tgn[3], tgm[3];
        for (int d = 0; d < 3; ++d) {
          gn[d] = 2. * g[0][d];
          gm[d] = 2. * g[1][d];
          tgn[d] = 2. * g[2][d];
          tgm[d] = 2. * g[3][d];
        }
        auto dot = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
        const double nn = dot(gn, gn), mm = dot(gm, gm), nm = dot(gn, gm);
        const double dnn = 2. * dot(gn, tgn), dmm = 2. * dot(gm, tgm),
                     dnm = dot(tgn, gm) + dot(gn, tgm);
        double x[5] = {std::max(0.5 * (r[0] + r[1]), 0.), std::max(0.5 * (r[0] - r[1]), 0.), 0., 0., 0.};
        double dx[5] = {0.5 * (r[2] + r[3]), 0.5 * (r[2] - r[3]), 0., 0., 0.};
        if (gga) {
          x[2] = std::max(0.25 * (nn + mm) + 0.5 * nm, sig_tol2);
          x[4] = std::max(0.25 * (nn + mm) - 0.5 * nm, sig_tol2);
          const double hs = 0.5 * (x[2] + x[4]);
          x[3] = std::max(std::min(0.25 * (nn - mm), hs), -hs);
          dx[2] = 0.25 * (dnn + dmm) + 0.5 * dnm;
          dx[3] = 0.25 * (dnn - dmm);
          dx[4] = 0.25 * (dnn + dmm) - 0.5 * dnm;
        }
        double out[10];
        fxc_point_pol(func, x, dx, out);
        // d/de of the UKS Z coefficients of oracle_exc_vxc_uks.
        const double wp = w[p];
        const double das = 0.25 * wp * (out[5] + out[6]), daz = 0.25 * wp * (out[5] - out[6]);
        double dcs[3] = {0., 0., 0.}, dcz[3] = {0., 0., 0.};
        if (gga) {
          const double g1 = 0.5 * wp * (out[2] + out[3] + out[4]), g2 = 0.5 * wp * (out[2] - out[4]),
                       g3 = 0.5 * wp * (out[2] - out[3] + out[4]);
          const double dg1 = 0.5 * wp * (out[7] + out[8] + out[9]),
                       dg2 = 0.5 * wp * (out[7] - out[9]),
                       dg3 = 0.5 * wp * (out[7] - out[8] + out[9]);
          for (int d = 0; d < 3; ++d) {
            dcs[d] = dg1 * gn[d] + g1 * tgn[d] + dg2 * gm[d] + g2 * tgm[d];
            dcz[d] = dg3 * gm[d] + g3 * tgm[d] + dg2 * gn[d] + g2 * tgn[d];
          }
        }
        for (int32_t k = 0; k < nbe; ++k) {
          double zs = das * B[o + k], zz = daz * B[o + k];
          if (gga) {
            zs += dcs[0] * Bx[o + k] + dcs[1] * By[o + k] + dcs[2] * Bz[o + k];
            zz += dcz[0] * Bx[o + k] + dcz[1] * By[o + k] + dcz[2] * Bz[o + k];
          }
          Z[0][o + k] = zs;
          Z[1][o + k] = zz;
        }
      }
      double* outm[2] = {fxc_s, fxc_z};
      for (int c = 0; c < 2; ++c) {
        V.assign((size_t)nbe * nbe, 0.);
        for (int32_t p = 0; p < npts; ++p) {
          const size_t o = (size_t)p * nbe;
          for (int32_t i = 0; i < nbe; ++i) {
            const double bi = B[o + i], zi = Z[c][o + i];
            double* vrow = V.data() + (size_t)i * nbe;
            for (int32_t j = 0; j <= i; ++j) vrow[j] += bi * Z[c][o + j] + zi * B[o + j];
          }
        }
        for (int32_t i = 0; i < nbe; ++i)
          for (int32_t j = 0; j <= i; ++j) {
#pragma omp atomic
            outm[c][bf[i] + (size_t)bf[j] * ldfxc] += V[(size_t)i * nbe + j];
          }
      }
    }
  }
  for (double* v : {fxc_s, fxc_z})
    for (int32_t j = 0; j < nbf; ++j)
      for (int32_t i = j + 1; i < nbf; ++i) v[j + (size_t)i * ldfxc] = v[i + (size_t)j * ldfxc];
  return 0;
}

}  // extern "C"

extern "C" {

// UKS EXC nuclear gradient: the same driver with spin_dim 2
// (reference_replicated_xc_host_integrator_exc_grad.hpp:106-600, is_uks
// branches): X_N = B P_s and X_Z = B P_z (factor 1) with their gradient
// products for GGAs, the polarised functional, the grid-weight part with
// w eps (rho_a + rho_b), and per shell
//   0.5 (vrho_n X_N + vrho_z X_Z) grad B
//   + GGA: 0.5 (v_pp + v_pm + v_mm)(H grad n X_N + grad B (grad n . Xg_N))
//        + 0.5 (v_pp - v_mm)(H grad m X_N + grad B (grad m . Xg_N))
//        + 0.5 (v_pp - v_mm)(H grad n X_Z + grad B (grad n . Xg_Z))
//        + 0.5 (v_pp - v_pm + v_mm)(H grad m X_Z + grad B (grad m . Xg_Z))
// (vrho_n,z = w (vrho_a +- vrho_b), v_pp/pm/mm = w vsigma_aa/ab/bb).
int oracle_exc_grad_uks(int func, int32_t nbf, const oracle_shell* shells, int32_t natoms,
                        const double* xyz, const double* Ps, const double* Pz, int64_t ldp,
                        int32_t ntasks, const oracle_task* tasks, const double* points,
                        const double* weights, const int32_t* shell_lists, int weight_derivs,
                        double* grad, int nthreads) {
  if (func != ORACLE_SVWN5 && func != ORACLE_PBE && func != ORACLE_PBE0) return 1;
  const bool gga = func != ORACLE_SVWN5;
  for (int k = 0; k < 3 * natoms; ++k) grad[k] = 0.;
  std::vector<double> rab((size_t)natoms * natoms);
  for (int a = 0; a < natoms; ++a)
    for (int b = 0; b < natoms; ++b) {
      const double dx = xyz[3 * a] - xyz[3 * b], dy = xyz[3 * a + 1] - xyz[3 * b + 1],
                   dz = xyz[3 * a + 2] - xyz[3 * b + 2];
      rab[(size_t)a + (size_t)b * natoms] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  int32_t nsh_total = 0;
  for (int32_t it = 0; it < ntasks; ++it)
    for (int32_t k = 0; k < tasks[it].nshells; ++k)
      nsh_total = std::max(nsh_total, shell_lists[tasks[it].shell_offset + k] + 1);
  std::vector<int32_t> first_ao(nsh_total + 1, 0), centre(nsh_total, -1);
  for (int32_t s = 0; s < nsh_total; ++s) {
    const auto& sh = shells[s];
    first_ao[s + 1] = first_ao[s] + (sh.pure ? 2 * sh.l + 1 : (sh.l + 1) * (sh.l + 2) / 2);
    for (int a = 0; a < natoms && centre[s] < 0; ++a)
      if (xyz[3 * a] == sh.origin[0] && xyz[3 * a + 1] == sh.origin[1] &&
          xyz[3 * a + 2] == sh.origin[2])
        centre[s] = a;
  }
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
#pragma omp parallel
  {
    std::vector<double> pl[10], XN[4], XZ[4], sub, wf, rho, sig, eps, vr, vs, gn[3], gm[3];
    std::vector<double> gloc(3 * (size_t)natoms, 0.);
    std::vector<int32_t> bf;
#pragma omp for schedule(dynamic)
    for (int32_t it = 0; it < ntasks; ++it) {
      const oracle_task& t = tasks[it];
      const int32_t npts = t.npts;
      const int32_t* sl = shell_lists + t.shell_offset;
      bf.clear();
      for (int32_t k = 0; k < t.nshells; ++k)
        for (int32_t f = first_ao[sl[k]]; f < first_ao[sl[k] + 1]; ++f) bf.push_back(f);
      const int32_t nbe = (int32_t)bf.size();
      const double* pts = points + 3 * t.point_offset;
      const double* w = weights + t.point_offset;
      double* o[10];
      for (int q = 0; q < 10; ++q) {
        pl[q].assign((size_t)npts * nbe, 0.);
        o[q] = pl[q].data();
      }
      oracle_collocation_hessian(shells, t.nshells, sl, npts, pts, o);
      const int nx = gga ? 4 : 1;
      for (int q = 0; q < nx; ++q) {
        xmat(pl[q], Ps, ldp, bf, npts, 1.0, sub, XN[q]);
        xmat(pl[q], Pz, ldp, bf, npts, 1.0, sub, XZ[q]);
      }
      rho.assign(2 * (size_t)npts, 0.);
      sig.assign(3 * (size_t)npts, 0.);
      for (int d = 0; d < 3; ++d) {
        gn[d].assign(npts, 0.);
        gm[d].assign(npts, 0.);
      }
      for (int32_t p = 0; p < npts; ++p) {
        const size_t off = (size_t)p * nbe;
        double rs = 0., rz = 0., an[3] = {0., 0., 0.}, am[3] = {0., 0., 0.};
        for (int32_t k = 0; k < nbe; ++k) {
          rs += pl[0][off + k] * XN[0][off + k];
          rz += pl[0][off + k] * XZ[0][off + k];
          for (int d = 0; d < 3; ++d) {
            an[d] += pl[1 + d][off + k] * XN[0][off + k];
            am[d] += pl[1 + d][off + k] * XZ[0][off + k];
          }
        }
        rho[2 * p] = 0.5 * (rs + rz);
        rho[2 * p + 1] = 0.5 * (rs - rz);
        double nn = 0., mm = 0., nm = 0.;
        for (int d = 0; d < 3; ++d) {
          gn[d][p] = 2. * an[d];
          gm[d][p] = 2. * am[d];
          nn += gn[d][p] * gn[d][p];
          mm += gm[d][p] * gm[d][p];
          nm += gn[d][p] * gm[d][p];
        }
        sig[3 * p] = 0.25 * (nn + mm) + 0.5 * nm;
        sig[3 * p + 1] = 0.25 * (nn - mm);
        sig[3 * p + 2] = 0.25 * (nn + mm) - 0.5 * nm;
      }
      eps.resize(npts);
      vr.resize(2 * (size_t)npts);
      vs.assign(3 * (size_t)npts, 0.);
      oracle_functional_pol(func, npts, rho.data(), gga ? sig.data() : nullptr, eps.data(),
                            vr.data(), gga ? vs.data() : nullptr);
      if (weight_derivs) {
        wf.resize(npts);
        for (int32_t p = 0; p < npts; ++p)
          wf[p] = eps[p] * (rho[2 * p] + rho[2 * p + 1]) * w[p];
        ssf_grad_contract(natoms, xyz, rab.data(), t, pts, wf.data(), gloc.data());
      }
      int32_t boff = 0;
      for (int32_t k = 0; k < t.nshells; ++k) {
        const int32_t s = sl[k];
        const int32_t n = first_ao[s + 1] - first_ao[s];
        const int32_t at = centre[s];
        if (at < 0 || (weight_derivs && at == t.iparent)) {
          boff += n;
          continue;
        }
        double g[3] = {0., 0., 0.};
        for (int32_t mu = boff; mu < boff + n; ++mu)
          for (int32_t p = 0; p < npts; ++p) {
            const size_t i = (size_t)p * nbe + mu;
            const double vn = w[p] * (vr[2 * p] + vr[2 * p + 1]),
                         vz = w[p] * (vr[2 * p] - vr[2 * p + 1]);
            const double xn = XN[0][i], xz = XZ[0][i];
            for (int d = 0; d < 3; ++d) g[d] += 0.5 * (vn * xn + vz * xz) * pl[1 + d][i];
            if (gga) {
              const double vpp = w[p] * vs[3 * p], vpm = w[p] * vs[3 * p + 1],
                           vmm = w[p] * vs[3 * p + 2];
              const double dn[3] = {gn[0][p], gn[1][p], gn[2][p]};
              const double dm[3] = {gm[0][p], gm[1][p], gm[2][p]};
              const double h[3][3] = {{pl[4][i], pl[5][i], pl[6][i]},
                                      {pl[5][i], pl[7][i], pl[8][i]},
                                      {pl[6][i], pl[8][i], pl[9][i]}};
              const double nN = dn[0] * XN[1][i] + dn[1] * XN[2][i] + dn[2] * XN[3][i];
              const double mN = dm[0] * XN[1][i] + dm[1] * XN[2][i] + dm[2] * XN[3][i];
              const double nZ = dn[0] * XZ[1][i] + dn[1] * XZ[2][i] + dn[2] * XZ[3][i];
              const double mZ = dm[0] * XZ[1][i] + dm[1] * XZ[2][i] + dm[2] * XZ[3][i];
              const double c1 = 0.5 * (vpp + vpm + vmm), c2 = 0.5 * (vpp - vmm),
                           c3 = 0.5 * (vpp - vpm + vmm);
              for (int d = 0; d < 3; ++d) {
                const double hn = h[d][0] * dn[0] + h[d][1] * dn[1] + h[d][2] * dn[2];
                const double hm = h[d][0] * dm[0] + h[d][1] * dm[1] + h[d][2] * dm[2];
                const double db = pl[1 + d][i];
                g[d] += c1 * (hn * xn + nN * db) + c2 * (hm * xn + mN * db) +
                        c2 * (hn * xz + nZ * db) + c3 * (hm * xz + mZ * db);
              }
            }
          }
        for (int d = 0; d < 3; ++d) {
          gloc[3 * at + d] += -2. * g[d];
          if (weight_derivs) gloc[3 * t.iparent + d] -= -2. * g[d];
        }
        boff += n;
      }
    }
#pragma omp critical
    for (int k = 0; k < 3 * natoms; ++k) grad[k] += gloc[k];
  }
  return 0;
}

}  // extern "C"
