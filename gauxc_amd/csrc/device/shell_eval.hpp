// Values and gradients of the functions of one basis shell at one point,
// shared by the collocation stage (which stores them) and the density stage
// (which evaluates them where it consumes them).
//
// Semantics: GauXC's collocation (gau2grid_collocation.cxx:25-145 on the
// host, collocation_device_masked_combined_kernel_deriv1 on the device):
//   B_mu(r) = R(|r-A|) Y_mu(r-A),  R(r) = sum_p c_p exp(-a_p r^2),
//   d/dx B = (dY/dx) R + Y x dR/dr / r = (dY/dx) R - 2 x Y sum_p a_p c_p e^{-a_p r^2},
// Cartesian functions in CCA order (x^l, x^(l-1)y, ..., z^l), spherical
// functions as real solid harmonics in CCA order m = -l..l, unnormalised as
// gau2grid's GG_SPHERICAL_CCA.  All arithmetic is FP64.
#pragma once
#include "internal.hpp"

namespace xcq {

// ---- real solid harmonics, tabulated at compile time --------------------
// Closed form of Helgaker, Jorgensen & Olsen, "Molecular Electronic-Structure
// Theory", eq. 6.4.47-6.4.50:
//   S_lm = N_lm sum_{t,u,v} C^{lm}_{tuv} x^{2t+|m|-2(u+v)} y^{2(u+v)} z^{l-2t-|m|},
//   C = (-1)^{t+v-v_m} (1/4)^t binom(l,t) binom(l-t,|m|+t) binom(t,u) binom(|m|,2v),
//   N_lm = sqrt(2 (l+|m|)! (l-|m|)! / 2^{delta_m0}) / (2^|m| l!),
// v_m = 0 (m >= 0) or 1/2 (m < 0), v running in steps of 1 from v_m.  The
// table is a constant expression, so the unrolled transforms below multiply
// by immediate coefficients and skip the zero ones.
constexpr int kNCart = (kMaxL + 1) * (kMaxL + 2) / 2;

constexpr double cx_fact(int n) {
  double r = 1.;
  for (int k = 2; k <= n; ++k) r *= k;
  return r;
}
constexpr double cx_binom(int n, int k) {
  return (k < 0 || k > n) ? 0. : cx_fact(n) / (cx_fact(k) * cx_fact(n - k));
}
// Newton's iteration from above; agrees with a correctly rounded square
// root to within one unit in the last place.
constexpr double cx_sqrt(double x) {
  double r = x > 1. ? x : 1.;
  for (int i = 0; i < 200; ++i) {
    const double n = 0.5 * (r + x / r);
    if (!(n < r)) break;
    r = n;
  }
  return r;
}
constexpr int cx_cart_index(int l, int a, int b) {  // CCA position of x^a y^b z^(l-a-b)
  int idx = 0;
  for (int aa = l; aa >= 0; --aa)
    for (int bb = l - aa; bb >= 0; --bb, ++idx)
      if (aa == a && bb == b) return idx;
  return -1;
}

struct SphTable {
  double c[kMaxL + 1][2 * kMaxL + 1][kNCart];
};

constexpr SphTable make_sph_table() {
  SphTable t{};
  for (int l = 0; l <= kMaxL; ++l)
    for (int m = -l; m <= l; ++m) {
      const int am = m < 0 ? -m : m;
      double pow2 = 1.;
      for (int k = 0; k < am; ++k) pow2 *= 2.;
      const double N =
          cx_sqrt(2. * cx_fact(l + am) * cx_fact(l - am) / (m == 0 ? 2. : 1.)) / (pow2 * cx_fact(l));
      const int v2s = m < 0 ? 1 : 0;  // 2v starts at 0 (m >= 0) or 1 (m < 0)
      for (int tt = 0; tt <= (l - am) / 2; ++tt) {
        double q = 1.;
        for (int k = 0; k < tt; ++k) q *= 0.25;
        for (int u = 0; u <= tt; ++u)
          for (int v2 = v2s; v2 <= am; v2 += 2) {
            const int se = tt + (v2 - v2s) / 2;
            const double C = ((se & 1) ? -1. : 1.) * q * cx_binom(l, tt) *
                             cx_binom(l - tt, am + tt) * cx_binom(tt, u) * cx_binom(am, v2);
            const int ax = 2 * tt + am - 2 * u - v2, ay = 2 * u + v2;
            t.c[l][m + l][cx_cart_index(l, ax, ay)] += N * C;
          }
      }
    }
  return t;
}

__device__ constexpr SphTable kSph = make_sph_table();

// Calls sink(f, v, gx, gy, gz) for each function f of an angular momentum L
// shell (in CCA order) at displacement (x, y, z) from its centre, with
// radial part R = sum c e^{-a r^2} and dR = -2 sum a c e^{-a r^2}.
// Without GRAD the gradient arguments are zero.  Spherical shells
// accumulate their 2L+1 outputs while the Cartesian monomials are
// generated, so no Cartesian array is kept; with LEAN they are formed one
// output at a time instead (fewer live registers, a few terms recomputed).
// With ANG only the angular factors are formed: v = Y_f and the gradient
// grad Y_f (R and dR are not read), for callers that apply the radial part
// once per shell.
template <int L, bool GRAD, bool LEAN = false, bool ANG = false, typename Sink>
__device__ __forceinline__ void shell_functions(double x, double y, double z, double R,
                                                double dR, bool pure, Sink&& sink) {
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.;
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  const double Rx = dR * x, Ry = dR * y, Rz = dR * z;
  // Cartesian term (a, b, c = L - a - b): value and gradient.
  auto term = [&](int a, int b, double& cv, double& gx, double& gy, double& gz) {
    const int c = L - a - b;
    const double mono = xp[a] * yp[b] * zp[c];
    cv = ANG ? mono : mono * R;
    if constexpr (GRAD) {
      const double dxm = a ? double(a) * xp[a > 0 ? a - 1 : 0] * yp[b] * zp[c] : 0.;
      const double dym = b ? double(b) * xp[a] * yp[b > 0 ? b - 1 : 0] * zp[c] : 0.;
      const double dzm = c ? double(c) * xp[a] * yp[b] * zp[c > 0 ? c - 1 : 0] : 0.;
      if constexpr (ANG) {
        gx = dxm;
        gy = dym;
        gz = dzm;
      } else {
        gx = dxm * R + mono * Rx;
        gy = dym * R + mono * Ry;
        gz = dzm * R + mono * Rz;
      }
    }
  };
  if (pure) {
    if constexpr (LEAN) {
      // One output at a time, each from the Cartesian terms with a nonzero
      // coefficient (about two per output): four accumulators live instead
      // of 4(2L+1).
#pragma unroll
      for (int m = 0; m < 2 * L + 1; ++m) {
        double v = 0., wx = 0., wy = 0., wz = 0.;
        bool first = true;  // the first nonzero term sets the sums (no "+ 0")
        int ic = 0;
#pragma unroll
        for (int a = L; a >= 0; --a) {
#pragma unroll
          for (int b = L - a; b >= 0; --b) {
            const double cf = kSph.c[L][m][ic];
            if (cf != 0.) {
              double cv, gx = 0., gy = 0., gz = 0.;
              term(a, b, cv, gx, gy, gz);
              v = first ? cf * cv : v + cf * cv;
              if constexpr (GRAD) {
                wx = first ? cf * gx : wx + cf * gx;
                wy = first ? cf * gy : wy + cf * gy;
                wz = first ? cf * gz : wz + cf * gz;
              }
              first = false;
            }
            ++ic;
          }
        }
        sink(m, v, wx, wy, wz);
      }
      return;
    }
    double v[2 * L + 1], wx[2 * L + 1], wy[2 * L + 1], wz[2 * L + 1];
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) v[m] = wx[m] = wy[m] = wz[m] = 0.;
    int ic = 0;
#pragma unroll
    for (int a = L; a >= 0; --a) {
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        double cv, gx = 0., gy = 0., gz = 0.;
        term(a, b, cv, gx, gy, gz);
#pragma unroll
        for (int m = 0; m < 2 * L + 1; ++m) {
          const double cf = kSph.c[L][m][ic];
          if (cf != 0.) {
            v[m] += cf * cv;
            if constexpr (GRAD) {
              wx[m] += cf * gx;
              wy[m] += cf * gy;
              wz[m] += cf * gz;
            }
          }
        }
        ++ic;
      }
    }
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) sink(m, v[m], wx[m], wy[m], wz[m]);
  } else {
    int ic = 0;
#pragma unroll
    for (int a = L; a >= 0; --a) {
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        double cv, gx = 0., gy = 0., gz = 0.;
        term(a, b, cv, gx, gy, gz);
        sink(ic, cv, gx, gy, gz);
        ++ic;
      }
    }
  }
}

// For the Z matrix: calls sink(f, v, zf) with v = B_f and
// zf = a B_f + u . grad B_f for each function f of the shell.  With
// B = R Y and grad R = dR r this is zf = A Y_f + R (u . grad Y_f),
// A = a R + dR (u . r), so each Cartesian term needs its monomial and the
// monomial's derivative along u, and a spherical output two sums (one
// output at a time, as the LEAN form above) instead of four.  The caller
// passes A.
template <int L, typename Sink>
__device__ __forceinline__ void shell_functions_z(double x, double y, double z, double R,
                                                  double A, double ux, double uy, double uz,
                                                  bool pure, Sink&& sink) {
  double xp[L + 1], yp[L + 1], zp[L + 1];
  xp[0] = yp[0] = zp[0] = 1.;
#pragma unroll
  for (int k = 1; k <= L; ++k) {
    xp[k] = xp[k - 1] * x;
    yp[k] = yp[k - 1] * y;
    zp[k] = zp[k - 1] * z;
  }
  // Cartesian term (a, b, c = L - a - b): monomial and its derivative along u.
  auto term = [&](int a, int b, double& mono, double& du) {
    const int c = L - a - b;
    mono = xp[a] * yp[b] * zp[c];
    du = 0.;
    if (a) du += double(a) * xp[a > 0 ? a - 1 : 0] * yp[b] * zp[c] * ux;
    if (b) du += double(b) * xp[a] * yp[b > 0 ? b - 1 : 0] * zp[c] * uy;
    if (c) du += double(c) * xp[a] * yp[b] * zp[c > 0 ? c - 1 : 0] * uz;
  };
  if (pure) {
#pragma unroll
    for (int m = 0; m < 2 * L + 1; ++m) {
      double y0 = 0., yu = 0.;  // sum cf * monomial, sum cf * derivative
      bool first = true;
      int ic = 0;
#pragma unroll
      for (int a = L; a >= 0; --a) {
#pragma unroll
        for (int b = L - a; b >= 0; --b) {
          const double cf = kSph.c[L][m][ic];
          if (cf != 0.) {
            double mono, du;
            term(a, b, mono, du);
            y0 = first ? cf * mono : y0 + cf * mono;
            yu = first ? cf * du : yu + cf * du;
            first = false;
          }
          ++ic;
        }
      }
      sink(m, R * y0, A * y0 + R * yu);
    }
  } else {
    int ic = 0;
#pragma unroll
    for (int a = L; a >= 0; --a) {
#pragma unroll
      for (int b = L - a; b >= 0; --b) {
        double mono, du;
        term(a, b, mono, du);
        sink(ic, R * mono, A * mono + R * du);
        ++ic;
      }
    }
  }
}

template <int LMAX, typename Sink>
__device__ __forceinline__ void shell_functions_z_l(int l, double x, double y, double z,
                                                    double R, double A, double ux, double uy,
                                                    double uz, bool pure, Sink&& sink) {
  if (l == 0) shell_functions_z<0>(x, y, z, R, A, ux, uy, uz, pure, sink);
  else if (l == 1) shell_functions_z<1>(x, y, z, R, A, ux, uy, uz, pure, sink);
  else if (LMAX <= 2 || l == 2) shell_functions_z<2>(x, y, z, R, A, ux, uy, uz, pure, sink);
  else if (LMAX == 3 || l == 3)
    shell_functions_z<(LMAX < 3 ? 2 : 3)>(x, y, z, R, A, ux, uy, uz, pure, sink);
  else shell_functions_z<(LMAX < 4 ? 2 : 4)>(x, y, z, R, A, ux, uy, uz, pure, sink);
}

// The same for a run-time l <= max(LMAX, 2) (l above that is not evaluated).
template <int LMAX, bool GRAD, bool LEAN = false, bool ANG = false, typename Sink>
__device__ __forceinline__ void shell_functions_l(int l, double x, double y, double z, double R,
                                                  double dR, bool pure, Sink&& sink) {
  if (l == 0) shell_functions<0, GRAD, LEAN, ANG>(x, y, z, R, dR, pure, sink);
  else if (l == 1) shell_functions<1, GRAD, LEAN, ANG>(x, y, z, R, dR, pure, sink);
  else if (LMAX <= 2 || l == 2) shell_functions<2, GRAD, LEAN, ANG>(x, y, z, R, dR, pure, sink);
  else if (LMAX == 3 || l == 3)
    shell_functions<(LMAX < 3 ? 2 : 3), GRAD, LEAN, ANG>(x, y, z, R, dR, pure, sink);
  else shell_functions<(LMAX < 4 ? 2 : 4), GRAD, LEAN, ANG>(x, y, z, R, dR, pure, sink);
}

}  // namespace xcq
