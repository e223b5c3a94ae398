// Point exchange-correlation functionals on the device (FP64).
//
// PBE (Perdew, Burke & Ernzerhof, PRL 77, 3865 (1996)) for a closed shell in
// the "Laplacian" potential form LIO uses: the local potential is
//   y2a = vx + vc,  evaluated from rho, |grad rho|, grad rho . grad|grad rho|
// and the Laplacian, with PW92 (Perdew & Wang, PRB 45, 13244 (1992)) for the
// uniform-gas correlation.  Behaviour follows g2g/pointxc/calc_ggaCS.h:35-160,
// pbeCS.h:29-185 and gcorc.h:21-33, including two quirks that change results
// at the 1e-8 level and are therefore kept:
//   * the reference writes every PBE and PW92 constant as a float literal, so
//     its FP64 build computes with the single-precision values of kappa, mu,
//     beta, the PW92 fit, 3 pi^2 and the quotients 4/3 and 1/3;
//   * density cut-offs: everything is zero below rho = 1e-18, exchange needs
//     rho^2 > 1e-18 and correlation rho >= 1e-12; |grad rho|^2 is floored at
//     DBL_MIN.
// LDA (iexch 1..3): Slater exchange with no correlation, Gunnarsson-Lundqvist
// or VWN correlation (g2g/pointxc/calc_ldaCS.h:32-102).
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>

namespace xcq {
namespace xc {

struct PbeConst {
  // single-precision literals widened to double, as the reference sees them
  static constexpr double kappa = (double)0.804f;
  static constexpr double mu = (double)0.2195149727645171f;
  static constexpr double mu_over_kappa = (double)0.273028573090195f;
  static constexpr double ax = (double)-0.738558766382022405884230032680836f;
  static constexpr double beta = (double)0.06672455060314922f;
  static constexpr double inv_gamma = (double)32.1639684429148f;
  static constexpr double beta_over_gamma = (double)2.14612633996736f;
  static constexpr double three_pi2 = (double)29.608813203268075856503472999628f;
  static constexpr double four_thirds = (double)(4.0f / 3.0f);
  static constexpr double one_third = (double)(1.0f / 3.0f);
  static constexpr double pw_a = (double)0.0310907f;
  static constexpr double pw_a1 = (double)0.21370f;
  static constexpr double pw_b1 = (double)7.5957f;
  static constexpr double pw_b2 = (double)3.5876f;
  static constexpr double pw_b3 = (double)1.6382f;
  static constexpr double pw_b4 = (double)0.49294f;
};

// PW92 unpolarised correlation per particle and d/d rs, from sqrt(rs).
__device__ __host__ inline void pw92_unpolarised(double srs, double& e, double& de_drs) {
  using C = PbeConst;
  const double poly = C::pw_b1 + srs * (C::pw_b2 + srs * (C::pw_b3 + C::pw_b4 * srs));
  const double den = 2.0 * C::pw_a * srs * poly;
  const double pre = -2.0 * C::pw_a * (1.0 + C::pw_a1 * srs * srs);
  const double lg = log(1.0 + 1.0 / den);
  e = pre * lg;
  const double dden = C::pw_a * (C::pw_b1 / srs + 2.0 * C::pw_b2 + srs * (3.0 * C::pw_b3 + 4.0 * C::pw_b4 * srs));
  de_drs = -2.0 * C::pw_a * C::pw_a1 * lg - pre * dden / (den * (1.0 + den));
}

// PBE energies per particle (ex, ec) and potential vx + vc.
__device__ __host__ inline void pbe_closed(double rho, double agrad, double delgrad, double lap,
                                           double& ex, double& ec, double& v) {
  using C = PbeConst;
  const double rho2 = rho * rho;
  const double rho13 = cbrt(rho);
  const double kf = cbrt(C::three_pi2) * rho13;
  double vx = 0.0;
  ex = 0.0;
  if (rho2 > 1e-18) {
    const double k2 = 2.0 * kf;
    const double s = agrad / (k2 * rho);
    const double s2 = s * s, s3 = s * s2;
    const double eunif = C::ax * rho13;
    const double p0 = 1.0 + C::mu_over_kappa * s2;
    const double fx = 1.0 + C::kappa - C::kappa / p0;
    ex = eunif * fx;
    const double vlap = lap / (k2 * k2 * rho);
    const double vdel = delgrad / (k2 * k2 * k2 * rho2);
    const double fs = 2.0 * C::mu / (p0 * p0);
    const double fss = -4.0 * C::mu_over_kappa * s * fs / p0;
    vx = eunif * (C::four_thirds * fx - (vdel - C::four_thirds * s3) * fss - vlap * fs);
  }
  ec = 0.0;
  double vc = 0.0;
  if (rho >= 1e-12) {
    const double rs = cbrt(3.0 / (4.0 * M_PI * rho));
    const double ks2 = 2.0 * sqrt(4.0 * kf / M_PI);
    const double t = agrad / (ks2 * rho);
    const double t2 = t * t, t4 = t2 * t2;
    const double uu = delgrad / (rho2 * ks2 * ks2 * ks2);
    const double vv = lap / (rho * ks2 * ks2);
    double el, el_rs;
    pw92_unpolarised(sqrt(rs), el, el_rs);
    const double vlsd = el - rs * C::one_third * el_rs;
    const double bg = C::beta_over_gamma;
    const double A = bg / (exp(-el * C::inv_gamma) - 1.0);
    const double At2 = A * t2;
    const double q4 = 1.0 + At2;
    const double q5 = 1.0 + At2 + A * A * t4;
    const double H = (C::beta / bg) * log(1.0 + bg * q4 * t2 / q5);
    ec = el + H;
    const double rs3 = rs / 3.0;
    const double bec = A * A * (bg / A + 1.0) / C::beta;
    const double q8 = (q5 + bg * q4 * t2) * q5;
    const double q9 = 1.0 + 2.0 * At2;
    const double hA = -C::beta * A * (t4 * t2 / q8) * (2.0 + At2);
    const double h_rs = -rs3 * hA * bec * el_rs;
    const double f0 = 2.0 * bg - 6.0 * A;
    const double f1 = q5 * q9 + q4 * q9 * q9;
    const double hAt = 2.0 * C::beta * t4 * ((q4 * q5 * f0 - bg * f1) / q8) / q8;
    const double h_rst = rs3 * t2 * hAt * bec * el_rs;
    const double ht = 2.0 * C::beta * (q9 / q8);
    const double f2 = q4 * q5 + At2 * (q4 * q9 + q5);
    const double f3 = 2.0 * A * q5 * q9 + bg * f2;
    const double htt = 4.0 * C::beta * t * (2.0 * A / q8 - ((q9 / q8) * f3) / q8);
    double comm = H + h_rs + h_rst + t2 * ht / 6.0 + 7.0 * t2 * t * htt / 6.0;
    comm = comm - uu * htt - vv * ht;
    vc = vlsd + comm;
  }
  v = vx + vc;
}

// GGA point: rho, grad rho (g), diagonal (h1 = xx,yy,zz) and mixed
// (h2 = xy,xz,yz) second derivatives of rho.
__device__ __host__ inline void gga_point(double rho, const double g[3], const double h1[3],
                                          const double h2[3], double& ex, double& ec,
                                          double& y2a) {
  if (rho < 1e-18) {
    ex = ec = y2a = 0.0;
    return;
  }
  double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  if (g2 < DBL_MIN) g2 = DBL_MIN;
  const double ag = sqrt(g2);
  const double del = (g[0] * g[0] * h1[0] + g[1] * g[1] * h1[1] + g[2] * g[2] * h1[2] +
                      2 * (g[0] * g[1] * h2[0] + g[0] * g[2] * h2[1] + g[1] * g[2] * h2[2])) / ag;
  const double lap = h1[0] + h1[1] + h1[2];
  pbe_closed(rho, ag, del, lap, ex, ec, y2a);
}

// LDA point functional, iexch in 1..3.
__device__ __host__ inline void lda_point(double rho, int iexch, double& ex, double& ec, double& y2a) {
  if (rho == 0) {
    ex = ec = y2a = 0;
    return;
  }
  const double y = cbrt(rho);
  const double v0 = -0.984745021842697 * y;  // (4/3) of the Slater exchange potential
  ex = -0.738558766382022447 * y;
  if (iexch == 1) {
    ec = 0;
    y2a = v0;
    return;
  }
  const double rs = 0.620350490899400087 / y;  // (3/(4 pi rho))^(1/3)
  if (iexch == 2) {  // Gunnarsson-Lundqvist
    const double x = rs / 11.4;
    double vcl;
    if (x > 1.0) {
      ec = -0.0333 * (0.5 * x - 0.33333333333333);
      vcl = 0.0111 * x * 0.5;
    } else {
      const double t1 = 1.0 + x * x * x, t2 = log(1.0 + 1.0 / x), t3 = x * x;
      ec = -0.0333 * (t1 * t2 - t3 + 0.5 * x - 0.33333333333333);
      vcl = 0.0111 * x * (3.0 * t3 * t2 - t1 / (x * (x + 1.0)) - 2.0 * x + 0.5);
    }
    y2a = v0 + ec + vcl;
    return;
  }
  // VWN paramagnetic fit, with LIO's tabulated derived constants
  const double A = 0.03109205, b = 3.72744, c = 12.9352, x0 = -0.10498, Q = 6.15199066246304849;
  const double T4 = -0.0311676086789438, C2 = 25.8704, B2Q = 1.21178337371132,
               BX0Q = 1.14352579286644, B4 = 14.90976, QSQ = 37.8469891110325,
               BX0 = 1.0329232240928, A6 = 0.005182008333;
  const double x = sqrt(rs);
  const double X = rs + b * x + c;
  const double t1 = 2.0 * x + b, t2 = log(X), t3 = atan(Q / t1);
  const double t5 = (b * x + C2) / x;
  ec = (2.0 * log(x) - t2 + B2Q * t3 - T4 * (2.0 * log(x - x0) - t2 + BX0Q * t3)) * A;
  const double vcl = ec - A6 * x * (t5 / X - B4 / (t1 * t1 + QSQ) * BX0 - T4 * (2.0 / (x - x0) - t1 / X));
  y2a = v0 + vcl;
}

}  // namespace xc
}  // namespace xcq
