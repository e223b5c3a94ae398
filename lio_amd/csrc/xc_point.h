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
// or VWN correlation (g2g/pointxc/calc_ldaCS.h:32-102).  The other GGAs of
// iexch 4-8 follow further below (gga_point_x).
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


// ---------------------------------------------------------------------------
// The other closed-shell GGAs of LIO (iexch 4-8, g2g/pointxc/calc_ggaCS.h:
// 68-238): exchange Perdew-Wang 86 (iexch 4, 8; PRB 33, 8800 (1986)) or
// Becke 88 (5, 6, 7; PRA 38, 3098 (1988)); correlation VWN (6), VWN plus
// Perdew's 1986 gradient correction (4, 5; PRB 33, 8822 (1986)) or
// Lee-Yang-Parr (7, 8; PRB 37, 785 (1988)).  Each part returns its energy per
// particle and its share of the Laplacian-form potential y2a, which are
// summed.  The expressions keep the reference's arithmetic where it changes
// results: its truncated decimal constants (1.33333333333 for 4/3 and the
// like), the literals it writes in single precision (the P86 and LYP
// coefficients, -1/3, 7/6, 8/3: widened here from their float values), the
// VWN fit evaluated from its four parameters (not the LDA path's tabulated
// quotients) with the reference's Q * Q2 in the potential, and pow where it
// calls pow.  Inputs: rho, y = rho^(1/3), |grad rho|^2 (floored at DBL_MIN),
// |grad rho|, the Laplacian and u0 = grad rho . (Hessian grad rho) / |grad rho|.
// ---------------------------------------------------------------------------
namespace gga2 {
constexpr double kAlpha = -0.738558766382022447;  // Slater exchange, -(3/4)(3/pi)^(1/3)
constexpr double kRsC = 0.620350490899400087;     // rs = kRsC / rho^(1/3)

// Perdew-Wang 86 exchange: F(s) = (1 + 1.296 s^2 + 14 s^4 + 0.2 s^6)^(1/15).
__device__ __host__ inline void x_pw86(double rho, double y, double ag, double lap, double u0,
                                       double& ex, double& vx) {
  const double kf = 3.0936677 * y;  // (3 pi^2 rho)^(1/3)
  const double s = ag / (2.0 * kf * rho);
  const double s2 = s * s, s3 = s2 * s;
  const double g0 = 1.0 + 1.296 * s2 + 14.0 * pow(s, 4) + 0.2 * pow(s, 6);
  const double fx = 1.0 / 15.0;
  const double F = pow(g0, fx);
  ex = kAlpha * F * y;
  const double t = lap / (rho * 4.0 * (kf * kf));
  const double u = u0 / (pow(2.0 * kf, 3) * (rho * rho));
  const double g2 = 2.592 * s + 56.0 * s3 + 1.2 * pow(s, 5);  // dg0/ds
  const double g3 = 2.592 + 56.0 * s2 + 1.2 * pow(s, 4);      // g2 / s
  const double g4 = 112.0 * s + 4.8 * s3;                     // dg3/ds
  const double dF = fx * F / g0 * g2;
  const double dsF = fx * F / g0 * (-14.0 * fx * g3 * g2 / g0 + g4);
  vx = kAlpha * y * (1.33333333333 * F - t / s * dF - (u - 1.3333333333 * s3) * dsF);
}

// Becke 88 exchange for the closed shell (each spin channel holds rho / 2);
// its potential needs the gradient of |grad rho| (from the Hessian).
__device__ __host__ inline void x_b88(double rho, double y, const double g[3], double ag,
                                      const double h1[3], const double h2[3], double lap,
                                      double& ex, double& vx) {
  const double beta = 0.0042;
  const double e0 = kAlpha * y;
  const double half = rho / 2.0;
  const double r13 = cbrt(half), r43 = cbrt(pow(half, 4));
  const double xs = ag / (2.0 * r43);  // reduced gradient of one spin channel
  const double ash = asinh(xs);
  const double dn = 1.0 + 6.0 * beta * xs * ash;
  ex = e0 + (-2.0 * beta * r43 * xs * xs / (dn * rho));
  const double v0 = 1.33333333333333 * e0;
  const double fb = 1.0 / dn;
  const double xa = xs / sqrt(1.0 + xs * xs);
  const double dn1 = 1.0 + fb * (1.0 - 6.0 * beta * xs * xa);
  const double dn2 = 1.0 / (1.0 + xs * xs) + 2.0 * fb * (2.0 - 6.0 * beta * xs * xa);
  const double dn3 = ash * (1.0 + 2.0 * fb) + xa * dn2;
  const double de1 = 1.33333333333333 / cbrt(pow(rho, 7));
  const double c2 = cbrt(2.0);
  // grad |grad rho| components, then the gradient of the channel's x
  const double dgx = (g[0] * h1[0] + g[1] * h2[0] + g[2] * h2[1]) / ag;
  const double dgy = (g[0] * h2[0] + g[1] * h1[1] + g[2] * h2[2]) / ag;
  const double dgz = (g[0] * h2[1] + g[1] * h2[2] + g[2] * h1[2]) / ag;
  const double inv = 1.0 / (rho * y);
  const double gxx = c2 * (inv * dgx - de1 * g[0] * ag);
  const double gxy = c2 * (inv * dgy - de1 * g[1] * ag);
  const double gxz = c2 * (inv * dgz - de1 * g[2] * ag);
  const double dn4 = 6.0 * beta * fb * (g[0] / 2.0 * gxx + g[1] / 2.0 * gxy + g[2] / 2.0 * gxz);
  const double dn5 = 1.33333333333333 * r43 * r13 * xs * xs;
  const double tot = dn5 - lap / 2.0 * dn1 + dn4 * dn3;
  vx = v0 + (-beta * fb / r43 * tot);
}

// VWN correlation as the GGA branch writes it (parameters A, b, c, x0, Q).
__device__ __host__ inline void c_vwn(double y, double& ec, double& vc) {
  const double A = 0.03109205, b = 3.72744, c = 12.9352, x0 = -0.10498,
               Q = 6.15199066246304849, A16 = 0.005182008333, Q2 = 4.7309269;
  const double rs = kRsC / y, x = sqrt(rs);
  const double X = rs + b * x + c;
  const double X0 = x0 * x0 + b * x0 + c;
  const double t1 = 2.0 * x + b, lX = log(X), at = atan(Q / t1);
  const double bx = b * x0 / X0;
  ec = A * (2.0 * log(x) - lX + 2.0 * b / Q * at -
            bx * (2.0 * log(x - x0) - lX + 2.0 * (b + 2.0 * x0) / Q * at));
  const double t5 = (b * x + 2.0 * c) / x;
  vc = ec - A16 * x * (t5 / X - 4.0 * b / (t1 * t1 + Q * Q2) * (1.0 - x0 / X0 * (b - 2.0 * x0)) -
                       bx * (2.0 / (x - x0) - t1 / X));
}

// Perdew 86 gradient correction on top of VWN: ec and vc are VWN's on entry.
__device__ __host__ inline void c_p86(double rho, double y, double g2, double ag, double lap,
                                      double u0, double& ec, double& vc) {
  const double alf = 0.023266, bet = 7.389, gam = 8.723, del = 0.472;
  const double rs = kRsC / y, rs2 = rs * rs;
  const double n1 = (double)0.002568f + alf * rs + bet * rs2;
  const double d1 = (double)1.0f + gam * rs + del * rs2 + 1.0e4 * bet * (rs * rs * rs);
  const double C = 0.001667 + n1 / d1;
  const double n1p = alf + (double)2.0f * bet * rs;
  const double d1p = gam + (double)2.0f * del * rs + 3.0e4 * bet * rs2;
  double dC = n1p / d1 - n1 / (d1 * d1) * d1p;
  dC = (double)-0.333333333333333f * dC * kRsC / (y * rho);
  const double phi = (double)0.0008129082f / C * ag / pow(rho, (double)(7.0f / 6.0f));
  const double ep = exp(-phi);
  const double e0 = ep * C;
  ec = ec + e0 * g2 / (y * (rho * rho));
  const double D1 = ((double)2.0f - phi) * lap / rho;
  const double phi2 = phi * phi;
  double D2 = (double)1.33333333333333333f - (double)3.666666666666666666f * phi +
              (double)1.166666666666666f * phi2;
  D2 = D2 * g2 / (rho * rho);
  const double D3 = phi * (phi - (double)3.0f) * u0 / (rho * ag);
  const double D4 = ep * g2 / (y * rho) * (phi2 - phi - (double)1.0f) * dC;
  vc = vc - 1.0 * (e0 / y * (D1 - D2 + D3) - D4);
}

// Lee-Yang-Parr correlation in the Laplacian form of the reference.
__device__ __host__ inline void c_lyp(double rho, double g2, double lap, double& ec, double& vc) {
  const double a = 0.04918, b = 0.132, c = 0.2533, c3 = 0.0844333333, d = 0.349,
               d3 = 0.116333333, cf = 2.87123400018819;
  const double rm13 = 1 / cbrt(rho);
  const double r53 = cbrt(pow(rho, 5));
  const double ecr = exp(-c * rm13);
  const double f1 = (double)1.0f / ((double)1.0f + d * rm13);
  const double tw = (double)(1.0f / 8.0f) * (g2 / rho - lap);  // Weizsaecker term / 8
  double term = (tw / (double)9.0f + lap / (double)18.0f) - (double)2.0f * tw + cf * r53;
  term = rho + b * (rm13 * rm13) * ecr * term;
  ec = -a * f1 * term / rho;
  const double h1 = ecr / r53;
  const double g1 = f1 * h1;
  const double tm1 = d3 * (rm13 / rho);
  const double fp1 = tm1 * (f1 * f1);
  const double tm2 = (double)-1.666666666f + c3 * rm13;
  const double hp1 = h1 * tm2 / rho;
  const double gp1 = fp1 * h1 + hp1 * f1;
  const double fp2 = tm1 * (double)2.0f * f1 * (fp1 - (double)0.6666666666f * f1 / rho);
  const double tm3 = (double)1.6666666666f - c3 * (double)1.3333333333f * rm13;
  const double hp2 = hp1 * tm2 / rho + h1 * tm3 / (rho * rho);
  const double gp2 = fp2 * h1 + (double)2.0f * fp1 * hp1 + hp2 * f1;
  const double term3 = -a * (fp1 * rho + f1) -
                       a * b * cf * (gp1 * rho + (double)(8.0f / 3.0f) * g1) * r53;
  const double term4 = (gp2 * rho * g2 + gp1 * ((double)3.0f * g2 + (double)2.0f * rho * lap) +
                        (double)4.0f * g1 * lap) * a * b / (double)4.0f;
  const double term5 = ((double)3.0f * gp2 * rho * g2 +
                        gp1 * ((double)5.0f * g2 + (double)6.0f * rho * lap) +
                        (double)4.0f * g1 * lap) * a * b / (double)72.0f;
  vc = term3 - term4 - term5;
}
}  // namespace gga2

// iexch 4-8 point (not inlined: the density kernels that evaluate the
// functional keep their register allocation for the PBE default).
__device__ __host__ __noinline__ void gga_point_other(double rho, const double g[3],
                                                      const double h1[3], const double h2[3],
                                                      int iexch, double& ex, double& ec,
                                                      double& y2a) {
  if (rho < 1e-18) {
    ex = ec = y2a = 0.0;
    return;
  }
  const double y = cbrt(rho);
  double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  if (g2 < DBL_MIN) g2 = DBL_MIN;
  const double ag = sqrt(g2);
  const double lap = h1[0] + h1[1] + h1[2];
  const double u0 = ((g[0] * g[0]) * h1[0] + 2.0 * g[0] * g[1] * h2[0] + 2.0 * g[1] * g[2] * h2[2] +
                     2.0 * g[0] * g[2] * h2[1] + (g[1] * g[1]) * h1[1] + (g[2] * g[2]) * h1[2]) /
                    ag;
  double vx, vc;
  if (iexch == 4 || iexch == 8)
    gga2::x_pw86(rho, y, ag, lap, u0, ex, vx);
  else
    gga2::x_b88(rho, y, g, ag, h1, h2, lap, ex, vx);
  if (iexch <= 6) {
    gga2::c_vwn(y, ec, vc);
    if (iexch != 6) gga2::c_p86(rho, y, g2, ag, lap, u0, ec, vc);
  } else {
    gga2::c_lyp(rho, g2, lap, ec, vc);
  }
  y2a = vx + vc;
}

// Closed-shell GGA point for LIO's iexch: 4-8 as above, PBE for the others
// (the reference's GGA branch sends iexch 0-3 and 9 to PBE).
__device__ __host__ inline void gga_point_x(double rho, const double g[3], const double h1[3],
                                            const double h2[3], int iexch, double& ex,
                                            double& ec, double& y2a) {
  if (iexch >= 4 && iexch <= 8)
    gga_point_other(rho, g, h1, h2, iexch, ex, ec, y2a);
  else
    gga_point(rho, g, h1, h2, ex, ec, y2a);
}

// ---------------------------------------------------------------------------
// Open shell (spin-polarised) PBE in the same Laplacian potential form
// (g2g/pointxc/calc_ggaOS.h:10-108, pbeOS_main.h:24-202, pbeOS_exch.h:36-72,
// pbeOS_corr.h:54-247): exchange by exact spin scaling
// Ex[a, b] = (Ex[2a] + Ex[2b]) / 2, correlation from the PW92 interpolation
// between the unpolarised, fully polarised and spin-stiffness fits with the
// PBE gradient term H(rs, zeta, t) and its potential (Perdew, Burke & Wang,
// appendix E).  Constants are the reference's float literals widened to
// double; its float quotients 4/3 and -1/6 are kept as such.
// ---------------------------------------------------------------------------
struct PwFit {
  double a, a1, b1, b2, b3, b4;
};

// G(rs) of Perdew & Wang (1992) eq. (10) and dG/drs, from sqrt(rs).
__device__ __host__ inline void pw92_fit(const PwFit& f, double srs, double& e, double& de_drs) {
  const double pre = -2.0 * f.a * (1.0 + f.a1 * srs * srs);
  const double den = 2.0 * f.a * srs * (f.b1 + srs * (f.b2 + srs * (f.b3 + f.b4 * srs)));
  const double lg = log(1.0 + 1.0 / den);
  e = pre * lg;
  const double dden = f.a * (f.b1 / srs + 2.0 * f.b2 + srs * (3.0 * f.b3 + 4.0 * f.b4 * srs));
  de_drs = -2.0 * f.a * f.a1 * lg - pre * dden / (den * (1.0 + den));
}

// Exchange per particle and potential of one spin channel, evaluated at the
// doubled density n = 2 rho_s with s, u, v of that doubled density.
__device__ __host__ inline void pbe_exchange_channel(double n, double s, double u, double v,
                                                     double& ex, double& vx) {
  using C = PbeConst;
  const double exunif = C::ax * cbrt(n);
  const double s2 = s * s, s3 = s * s2;
  const double p0 = 1.0 + C::mu_over_kappa * s2;
  const double fx = 1.0 + C::kappa - C::kappa / p0;
  ex = exunif * fx;
  const double fs = 2.0 * C::mu / (p0 * p0);
  const double fss = (-4.0 * C::mu_over_kappa * s * fs) / p0;
  vx = exunif * (C::four_thirds * fx - (u - C::four_thirds * s3) * fss - v * fs);
}

// Spin-resolved inputs: rho, |grad rho|, grad rho . grad |grad rho| and the
// Laplacian of each channel, |grad (a + b)| and its delgrad.  Outputs the
// exchange and correlation energies per particle and the two potentials.
__device__ __host__ inline void pbe_open(double ra, double ga, double da, double la, double rb,
                                         double gb, double db, double lb, double gt, double dt,
                                         double& ex, double& ec, double& va, double& vb) {
  using C = PbeConst;
  const double kmin = 1e-18;
  const double pi = (double)3.141592653589793238462643383279f;
  const double kf1 = cbrt(C::three_pi2);
  double exa = 0.0, vxa = 0.0, exb = 0.0, vxb = 0.0;
  ex = ec = va = vb = 0.0;
  const double chan[2][4] = {{ra, ga, da, la}, {rb, gb, db, lb}};
  for (int c = 0; c < 2; ++c) {
    const double n = 2.0 * chan[c][0];
    if (n > kmin) {
      const double tk = 2.0 * (kf1 * cbrt(n));
      const double s = (2.0 * chan[c][1]) / (n * tk);
      const double v = (2.0 * chan[c][3]) / (n * (tk * tk));
      const double u = (4.0 * chan[c][2]) / ((n * n) * (tk * tk * tk));
      if (c == 0)
        pbe_exchange_channel(n, s, u, v, exa, vxa);
      else
        pbe_exchange_channel(n, s, u, v, exb, vxb);
    }
  }
  const double rho = ra + rb;
  ex = exa * (ra / rho) + exb * (rb / rho);
  va = vxa;
  vb = vxb;
  if (rho < kmin * 1e5) return;

  const double zet = (ra - rb) / rho;
  const double kf = kf1 * cbrt(rho);
  const double rs = cbrt((9.0 * pi) / 4.0) / kf;
  const double phi = (cbrt((1.0 + zet) * (1.0 + zet)) + cbrt((1.0 - zet) * (1.0 - zet))) / 2.0;
  const double tks = 2.0 * sqrt(4.0 * kf / pi) * phi;
  const double t = gt / (tks * rho);
  const double uu = dt / (rho * rho * tks * tks * tks);
  const double vv = (la + lb) / (rho * tks * tks);
  const double ww = (ga * ga - gb * gb - zet * gt * gt) / ((rho * rho) * (tks * tks));

  // local spin density correlation (PW92 eqs. 8-9 and A1-A4)
  const PwFit fu{(double)0.0310907f, (double)0.21370f, (double)7.5957f, (double)3.5876f,
                 (double)1.6382f, (double)0.49294f};
  const PwFit fp{(double)0.01554535f, (double)0.20548f, (double)14.1189f, (double)6.1977f,
                 (double)3.3662f, (double)0.62517f};
  const PwFit fa{(double)0.0168869f, (double)0.11125f, (double)10.357f, (double)3.6231f,
                 (double)0.88026f, (double)0.49671f};
  const double srs = sqrt(rs);
  double eu, eurs, ep, eprs, alfm, alfrsm;
  pw92_fit(fu, srs, eu, eurs);
  pw92_fit(fp, srs, ep, eprs);
  pw92_fit(fa, srs, alfm, alfrsm);
  const double z4 = (zet * zet) * (zet * zet);
  const double gam = pow(2.0, 4.0 / 3.0) - 2.0;
  const double fzz = 8.0 / (9.0 * gam);
  const double f = (cbrt(pow(1.0 + zet, 4)) + cbrt(pow(1.0 - zet, 4)) - 2.0) / gam;
  const double elsd = eu * (1.0 - f * z4) + ep * f * z4 - alfm * f * (1.0 - z4) / fzz;
  const double ecrs = eurs * (1.0 - f * z4) + eprs * f * z4 - alfrsm * f * (1.0 - z4) / fzz;
  const double fz = C::four_thirds * (cbrt(1.0 + zet) - cbrt(1.0 - zet)) / gam;
  const double eczet = 4.0 * pow(zet, 3) * f * (ep - eu + alfm / fzz) +
                       fz * (z4 * ep - z4 * eu - (1.0 - z4) * alfm / fzz);
  const double lsd = elsd - rs * ecrs / 3.0 - zet * eczet;

  // gradient correction H and its potential
  const double gamma = (double)0.03109069086965489503494086371273f;
  const double delta = C::beta / gamma;
  const double g3 = pow(phi, 3);
  const double A = delta / (exp(-elsd / (g3 * gamma)) - 1.0);
  const double A2 = A * A;
  const double t2 = t * t, t4 = t2 * t2, t6 = t4 * t2;
  const double q4 = 1.0 + A * t2;
  const double q5 = 1.0 + A * t2 + A2 * t4;
  const double H = g3 * (C::beta / delta) * log(1.0 + delta * q4 * t2 / q5);
  const double g4 = g3 * phi;
  const double rs3 = rs / 3.0;
  const double eta = (double)1e-12f;
  const double sixth = (double)(-1.0f / 6.0f);
  const double gz = (pow(pow(1.0 + zet, 2) + eta, sixth) - pow(pow(1.0 - zet, 2) + eta, sixth)) / 3.0;
  const double fac = delta / A + 1.0;
  const double bg = -3.0 * A2 * elsd * fac / (C::beta * g4);
  const double bec = A2 * fac / (C::beta * g3);
  const double q8 = pow(q5, 2) + delta * q4 * q5 * t2;
  const double q9 = 1.0 + 2.0 * A * t2;
  const double hA = -C::beta * g3 * A * (t6 / q8) * (2.0 + A * t2);
  const double hrs = -rs3 * hA * bec * ecrs;
  const double f0 = 2.0 * delta - 6.0 * A;
  const double f1 = q5 * q9 + q4 * q9 * q9;
  const double hAt = 2.0 * C::beta * g3 * t4 * ((q4 * q5 * f0 - delta * f1) / q8) / q8;
  const double hrst = rs3 * t2 * hAt * bec * ecrs;
  const double hz = 3.0 * gz * H / phi + hA * (bg * gz + bec * eczet);
  const double ht = 2.0 * C::beta * g3 * q9 / q8;
  const double hzt = 3.0 * gz * ht / phi + hAt * (bg * gz + bec * eczet);
  const double f2 = q4 * q5 + A * t2 * (q4 * q9 + q5);
  const double f3 = 2.0 * A * q5 * q9 + delta * f2;
  const double htt = 4.0 * C::beta * g3 * t * (2.0 * A / q8 - ((q9 / q8) * f3) / q8);
  double comm = H + hrs + hrst + t2 * ht / 6.0 + 7.0 * t2 * t * htt / 6.0;
  const double pref = hz - gz * t2 * ht / phi;
  const double f5 = gz * (2.0 * ht + t * htt) / phi;
  comm = comm - pref * zet - uu * htt - vv * ht - ww * (hzt - f5);
  ec = elsd + H;
  va = vxa + ((lsd + eczet) + (comm + pref));
  vb = vxb + ((lsd - eczet) + (comm - pref));
}

// Open-shell GGA point from the two channels' rho, gradient and second
// derivatives (h1 = xx,yy,zz; h2 = xy,xz,yz).  Below 1e-18 in both channels
// everything is zero; a channel with a zero gradient divides by zero, as the
// reference does.
__device__ __host__ inline void gga_point_open(double ra, const double ga[3], const double h1a[3],
                                               const double h2a[3], double rb, const double gb[3],
                                               const double h1b[3], const double h2b[3],
                                               double& ex, double& ec, double& va, double& vb) {
  ex = ec = va = vb = 0.0;
  if (!(ra > 1e-18 || rb > 1e-18)) return;
  double g[3], h1[3], h2[3];
  for (int k = 0; k < 3; ++k) {
    g[k] = ga[k] + gb[k];
    h1[k] = h1a[k] + h1b[k];
    h2[k] = h2a[k] + h2b[k];
  }
  auto del_spin = [](const double* q, const double* p1, const double* p2, double n) {
    return (q[0] * q[0] * p1[0] + 2.0 * q[0] * q[1] * p2[0] + 2.0 * q[1] * q[2] * p2[2] +
            2.0 * q[0] * q[2] * p2[1] + q[1] * q[1] * p1[1] + q[2] * q[2] * p1[2]) / n;
  };
  const double na = sqrt(ga[0] * ga[0] + ga[1] * ga[1] + ga[2] * ga[2]);
  const double nb = sqrt(gb[0] * gb[0] + gb[1] * gb[1] + gb[2] * gb[2]);
  const double nt = sqrt(g[0] * g[0] + g[1] * g[1] + g[2] * g[2]);
  const double da = del_spin(ga, h1a, h2a, na);
  const double db = del_spin(gb, h1b, h2b, nb);
  const double dt = (g[0] * g[0] * h1[0] + g[1] * g[1] * h1[1] + g[2] * g[2] * h1[2] +
                     2.0 * g[0] * g[1] * h2[0] + 2.0 * g[1] * g[2] * h2[2] +
                     2.0 * g[0] * g[2] * h2[1]) / nt;
  pbe_open(ra, na, da, h1a[0] + h1a[1] + h1a[2], rb, nb, db, h1b[0] + h1b[1] + h1b[2], nt, dt,
           ex, ec, va, vb);
}

}  // namespace xc
}  // namespace xcq
