// CPU oracle: an FP64 restatement of LIO's CPU g2g exchange-correlation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP engine is held
// against; the shipped library never includes, links or calls it.  It follows
// the reference's CPU semantics step by step (file:line cited per function)
// so that its results equal those of LIO's own CPU path to rounding:
//
//   grid + partition ........ g2g/regenerate_partition.cpp:161-462
//   function screening ...... g2g/classify_functions.cpp:12-141,
//                             g2g/partition.cpp:227-254
//   group Becke weights ..... g2g/cpu/weight.cpp:13-112
//   basis values/derivs ..... g2g/cpu/functions.cpp:14-359 (restated as the
//                             general derivative of x^l y^m z^n g(r))
//   density + Fock .......... g2g/cpu/iteration.cpp:39-309 (closed shell)
//   XC forces ............... g2g/cpu/iteration.cpp:213-268
//   packed P / F I/O ........ g2g/partition.cpp:77-219
//   PBE ..................... g2g/pointxc/calc_ggaCS.h:35-160, pbeCS.h:29-185,
//                             gcorc.h:21-33 (PBE 1996 with the PW92 fit)
//   LDA ..................... g2g/pointxc/calc_ldaCS.h:32-102
//
// Parity status: pinned against LIO's CPU path built from its sources
// (tests/test_oracle_golden.py, tests/golden/*.npz), against closed
// forms for the functional and the basis, and against brute-force numpy
// contractions on small groups (tests/test_oracle_math.py).
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <set>
#include <stdexcept>
#include <vector>

#include "xc_input.h"

namespace {

constexpr int kMaxContr = 13;
constexpr long long kMinCost = 250000;  // LIO_MINCOST_OFFSET default

struct V3 {
  double x, y, z;
};
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline double norm(V3 a) { return std::sqrt(a.x * a.x + a.y * a.y + a.z * a.z); }
inline double norm2(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
inline double comp(const V3& v, int k) { return k == 0 ? v.x : (k == 1 ? v.y : v.z); }

struct Problem {
  int natom = 0, M = 0, iexch = 9, ns = 0, np = 0, nd = 0;  // np, nd: shells
  bool gga = true;
  double nf = 1.0;  // d-diagonal normalisation
  std::vector<V3> pos;
  std::vector<int> Z;
  std::vector<double> rm;
  std::vector<int> shells;
  std::vector<int> nuc, ncont;  // nuc 0-based
  std::vector<double> a, c;     // M x 13, row i = function i
  std::vector<int> lx, ly, lz;  // Cartesian powers of each function
  std::vector<double> rho;      // packed (alpha when open)
  bool open = false;
  std::vector<double> rho_b;    // packed beta density (open shell)
  int ngrid = 0;
  std::vector<V3> e;
  std::vector<double> wang;
  double lcs = 8.0, sr = 0.6;
  bool aaf = false, rzw = true;
  unsigned mppc = 1, mfe = 10;
};

struct Group {
  int kind = 0;
  std::vector<V3> pts;
  std::vector<double> w;
  std::vector<int> atom;
  std::vector<int> shell_first;  // selected shells: global index of 1st comp
  std::vector<int> shell_len;    // 1, 3 or 6
  int s = 0, p = 0, d = 0;
  std::vector<int> nucs;         // sorted atoms owning selected functions
  std::vector<int> fglobal;      // expanded local -> global function
  int m() const { return (int)fglobal.size(); }
  long long cost() const {
    long long n = (long long)pts.size(), g = m();
    return 10 * ((n * g * (1 + g)) / 2) + kMinCost;
  }
};

struct Handle {
  Problem pb;
  std::vector<Group> groups;  // cost-sorted: cubes first then spheres
  double grid_seconds = 0.0;
};

double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// ----------------------------------------------------------------------------
// problem capture (g2g/init.cpp:69-219, 238-306)
// ----------------------------------------------------------------------------
void load(Problem& pb, const XcInput* in) {
  pb.natom = in->natom;
  pb.M = in->M;
  pb.iexch = in->iexch;
  if (pb.iexch < 1 || pb.iexch > 9) throw std::runtime_error("iexch out of range");
  pb.gga = pb.iexch > 3;
  pb.nf = in->normalize ? 1.0 / std::sqrt(3.0) : 1.0;
  pb.ns = in->nshell[0];
  pb.np = in->nshell[1] / 3;
  pb.nd = in->nshell[2] / 6;
  pb.pos.resize(pb.natom);
  pb.Z.resize(pb.natom);
  pb.rm.resize(pb.natom);
  pb.shells.resize(pb.natom);
  const unsigned* nr = in->grid_type == 0 ? in->Nr : in->Nr2;
  for (int i = 0; i < pb.natom; ++i) {
    pb.pos[i] = {in->r[i], in->r[pb.natom + i], in->r[2 * pb.natom + i]};
    pb.Z[i] = (int)in->Iz[i];
    pb.rm[i] = in->Rm[in->Iz[i]];
    pb.shells[i] = (int)nr[in->Iz[i]];
  }
  const int M = pb.M;
  pb.nuc.resize(M);
  pb.ncont.resize(M);
  pb.a.assign((size_t)M * kMaxContr, 0.0);
  pb.c.assign((size_t)M * kMaxContr, 0.0);
  for (int i = 0; i < M; ++i) {
    pb.nuc[i] = (int)in->Nuc[i] - 1;
    pb.ncont[i] = (int)in->ncont[i];
    if (pb.ncont[i] > kMaxContr) throw std::runtime_error("too many contractions");
    for (int k = 0; k < pb.ncont[i]; ++k) {
      pb.a[(size_t)i * kMaxContr + k] = in->a[(size_t)k * M + i];
      pb.c[(size_t)i * kMaxContr + k] = in->c[(size_t)k * M + i];
    }
  }
  pb.lx.assign(M, 0);
  pb.ly.assign(M, 0);
  pb.lz.assign(M, 0);
  static const int P[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
  static const int D[6][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0},
                              {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};
  for (int i = pb.ns; i < pb.ns + 3 * pb.np; ++i) {
    int k = (i - pb.ns) % 3;
    pb.lx[i] = P[k][0];
    pb.ly[i] = P[k][1];
    pb.lz[i] = P[k][2];
  }
  for (int i = pb.ns + 3 * pb.np; i < M; ++i) {
    int k = (i - pb.ns - 3 * pb.np) % 6;
    pb.lx[i] = D[k][0];
    pb.ly[i] = D[k][1];
    pb.lz[i] = D[k][2];
  }
  const size_t packed = (size_t)M * (M + 1) / 2;
  pb.rho.assign(in->rho, in->rho + packed);
  pb.open = in->open != 0;
  if (pb.open) {
    if (!in->rho_b) throw std::runtime_error("open shell without a beta density");
    pb.rho_b.assign(in->rho_b, in->rho_b + packed);
  }
  static const int sizes[3] = {50, 116, 194};
  if (in->grid_type < 0 || in->grid_type > 2) throw std::runtime_error("grid type");
  pb.ngrid = sizes[in->grid_type];
  pb.e.resize(pb.ngrid);
  pb.wang.resize(pb.ngrid);
  const double* e = in->e[in->grid_type];
  for (int k = 0; k < pb.ngrid; ++k) {
    pb.e[k] = {e[k], e[pb.ngrid + k], e[2 * pb.ngrid + k]};
    pb.wang[k] = in->wang[in->grid_type][k];
  }
  pb.lcs = in->little_cube_size;
  pb.sr = in->sphere_radius;
  pb.aaf = in->assign_all_functions != 0;
  pb.rzw = in->remove_zero_weights != 0;
  pb.mppc = in->min_points_per_cube;
  pb.mfe = in->max_function_exponent;
}

// ----------------------------------------------------------------------------
// screening (classify_functions.cpp:12-65, partition.cpp:229-254)
// ----------------------------------------------------------------------------
bool significant(int len, double a, double d2, unsigned mfe) {
  if (len == 1) return a * d2 < mfe - std::log(std::pow(2.0 * a / M_PI, 3)) / 4;
  // p and d: radius where the most diffuse primitive times r^l falls to e^-mfe
  // relative to its normalisation, found by fixed-point iteration from r = 1.
  double f = std::pow(2.0 * a / M_PI, 3);
  f = std::sqrt(f * 4.0 * a);
  const double lognorm = std::log(len == 3 ? std::sqrt(f) : std::fabs(f));
  double r = 1.0, delta;
  do {
    double extra = (len == 3 ? std::log(r) : 2 * std::log(r));
    double r1 = std::sqrt((mfe - lognorm + extra) / a);
    delta = std::fabs(r - r1);
    r = r1;
  } while (delta > 0.1);
  return std::sqrt(d2) < r;
}

void select_functions(const Problem& pb, Group& g, const std::vector<double>& d2,
                      const std::vector<double>& minexp) {
  std::set<int> nucs;
  auto scan = [&](int begin, int end, int len, int& counter) {
    for (int f = begin; f < end; f += len) {
      int at = pb.nuc[f];
      if (pb.aaf || significant(len, minexp[f], d2[at], pb.mfe)) {
        g.shell_first.push_back(f);
        g.shell_len.push_back(len);
        ++counter;
        nucs.insert(at);
      }
    }
  };
  scan(0, pb.ns, 1, g.s);
  scan(pb.ns, pb.ns + 3 * pb.np, 3, g.p);
  scan(pb.ns + 3 * pb.np, pb.ns + 3 * pb.np + 6 * pb.nd, 6, g.d);
  g.nucs.assign(nucs.begin(), nucs.end());
  g.fglobal.clear();
  for (size_t k = 0; k < g.shell_first.size(); ++k)
    for (int c = 0; c < g.shell_len[k]; ++c) g.fglobal.push_back(g.shell_first[k] + c);
}

// ----------------------------------------------------------------------------
// Becke partition restricted to the group's atoms (cpu/weight.cpp:13-112)
// ----------------------------------------------------------------------------
double becke_s(const Problem& pb, const std::vector<double>& dist, V3 r, int j,
               int k) {
  double mu = (norm(sub(r, pb.pos[j])) - norm(sub(r, pb.pos[k]))) /
              dist[(size_t)j * pb.natom + k];
  double chi = pb.rm[j] / pb.rm[k];
  chi = (chi - 1.0) / (chi + 1.0);
  mu += (chi / (chi * chi - 1.0)) * (1.0 - mu * mu);
  for (int it = 0; it < 3; ++it) mu = 1.5 * mu - 0.5 * (mu * mu * mu);
  return 0.5 * (1.0 - mu);
}

void group_weights(const Problem& pb, const std::vector<double>& dist, Group& g) {
  const int n = (int)g.pts.size();
  for (int p = 0; p < n; ++p) {
    const int own = g.atom[p];
    const V3 r = g.pts[p];
    double total = 0.0, mine = 0.0;
    bool own_in = std::binary_search(g.nucs.begin(), g.nucs.end(), own);
    for (int j : g.nucs) {
      double prod = 1.0;
      for (int k : g.nucs) {
        if (k == j) continue;
        prod *= becke_s(pb, dist, r, j, k);
        if (prod == 0.0) break;
      }
      if (j == own) {
        mine = prod;
        if (mine == 0.0) break;  // the weight is zero whatever follows
      }
      total += prod;
    }
    if (!own_in) {
      mine = 1.0;
      for (int k : g.nucs) {
        mine *= becke_s(pb, dist, r, own, k);
        if (mine == 0.0) break;
      }
    }
    g.w[p] *= (total == 0.0 ? 0.0 : mine / total);
  }
  if (pb.rzw) {
    size_t k = 0;
    for (size_t p = 0; p < g.pts.size(); ++p) {
      if (g.w[p] != 0.0) {
        g.pts[k] = g.pts[p];
        g.w[k] = g.w[p];
        g.atom[k] = g.atom[p];
        ++k;
      }
    }
    g.pts.resize(k);
    g.w.resize(k);
    g.atom.resize(k);
  }
}

// ----------------------------------------------------------------------------
// grid and partition (regenerate_partition.cpp:161-466)
// ----------------------------------------------------------------------------
void build_partition(Handle& h) {
  Problem& pb = h.pb;
  const int M = pb.M, natom = pb.natom;
  // smallest exponent per element and per function
  std::vector<double> minexp_Z(130, std::numeric_limits<double>::max());
  std::vector<double> minexp_f(M, std::numeric_limits<double>::max());
  for (int i = 0; i < M; ++i) {
    int zt = pb.Z[pb.nuc[i]];
    for (int k = 0; k < pb.ncont[i]; ++k) {
      double a = pb.a[(size_t)i * kMaxContr + k];
      minexp_Z[zt] = std::min(minexp_Z[zt], a);
      if (a < minexp_f[i]) minexp_f[i] = a;
    }
  }
  // bounding prism of all atoms padded by each element's reach
  V3 x0{0, 0, 0}, x1{0, 0, 0};
  for (int at = 0; at < natom; ++at) {
    double rad = 2 * std::sqrt(pb.mfe / minexp_Z[pb.Z[at]]);
    V3 p = pb.pos[at];
    if (at == 0) {
      x0 = {p.x - rad, p.y - rad, p.z - rad};
      x1 = {p.x + rad, p.y + rad, p.z + rad};
    } else {
      x0 = {std::min(x0.x, p.x - rad), std::min(x0.y, p.y - rad), std::min(x0.z, p.z - rad)};
      x1 = {std::max(x1.x, p.x + rad), std::max(x1.y, p.y + rad), std::max(x1.z, p.z + rad)};
    }
  }
  const unsigned nx = (unsigned)std::ceil((x1.x - x0.x) / pb.lcs);
  const unsigned ny = (unsigned)std::ceil((x1.y - x0.y) / pb.lcs);
  const unsigned nz = (unsigned)std::ceil((x1.z - x0.z) / pb.lcs);

  // sphere radius per atom: radius of the outermost shell kept in the sphere
  std::vector<double> srad(natom, 0.0);
  for (int at = 0; at < natom; ++at) {
    const int n = pb.shells[at];
    const unsigned inc = (unsigned)std::ceil(pb.sr * n);
    if (pb.sr > 0 && inc > 0) {
      double x = std::cos((M_PI / (n + 1)) * (n - inc + 1));
      srad[at] = pb.rm[at] * (1.0 + x) / (1.0 - x);
    }
  }
  std::vector<double> dist((size_t)natom * natom);
  for (int i = 0; i < natom; ++i)
    for (int j = 0; j < natom; ++j) dist[(size_t)i * natom + j] = norm(sub(pb.pos[i], pb.pos[j]));

  // points: Gauss-Chebyshev (2nd kind) radial x Lebedev, clipped to the prism
  std::vector<Group> cube((size_t)nx * ny * nz), sphere(natom);
  for (int at = 0; at < natom; ++at) {
    const int n = pb.shells[at];
    const double t0 = M_PI / (n + 1), rm = pb.rm[at];
    const unsigned inc = (unsigned)std::ceil(pb.sr * n);
    const V3 A = pb.pos[at];
    for (int sh = 0; sh < n; ++sh) {
      const double t1 = t0 * (sh + 1);
      const double x = std::cos(t1);
      const double w = t0 * std::fabs(std::sin(t1));
      const double r1 = rm * (1.0 + x) / (1.0 - x);
      const double wrad = w * (r1 * r1) * rm * 2.0 / ((1.0 - x) * (1.0 - x));
      for (int k = 0; k < pb.ngrid; ++k) {
        V3 q{A.x + pb.e[k].x * r1, A.y + pb.e[k].y * r1, A.z + pb.e[k].z * r1};
        bool inside = x0.x <= q.x && q.x <= x1.x && x0.y <= q.y && q.y <= x1.y &&
                      x0.z <= q.z && q.z <= x1.z;
        if (!inside) continue;
        const double pw = wrad * pb.wang[k];
        Group* g;
        if (pb.sr > 0 && (unsigned)sh >= n - inc) {
          g = &sphere[at];
        } else {
          unsigned cx = (unsigned)std::floor((q.x - x0.x) / pb.lcs);
          unsigned cy = (unsigned)std::floor((q.y - x0.y) / pb.lcs);
          unsigned cz = (unsigned)std::floor((q.z - x0.z) / pb.lcs);
          if (cx >= nx || cy >= ny || cz >= nz) throw std::runtime_error("cube index");
          g = &cube[((size_t)cx * ny + cy) * nz + cz];
        }
        g->pts.push_back(q);
        g->w.push_back(pw);
        g->atom.push_back(at);
      }
    }
  }

  std::vector<Group> cubes, spheres;
  std::vector<double> d2(natom);
  for (unsigned i = 0; i < nx; ++i)
    for (unsigned j = 0; j < ny; ++j)
      for (unsigned k = 0; k < nz; ++k) {
        Group& g = cube[((size_t)i * ny + j) * nz + k];
        V3 c0{x0.x + i * pb.lcs, x0.y + j * pb.lcs, x0.z + k * pb.lcs};
        for (int at = 0; at < natom; ++at) {
          double dv[3];
          for (int ax = 0; ax < 3; ++ax) {
            double pa = comp(pb.pos[at], ax), lo = comp(c0, ax);
            if (pa < lo) dv[ax] = lo - pa;
            else if (pa > lo + pb.lcs) dv[ax] = pa - (lo + pb.lcs);
            else dv[ax] = 0;
          }
          d2[at] = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
        }
        g.kind = 0;
        select_functions(pb, g, d2, minexp_f);
        if (g.fglobal.empty() || g.pts.size() < pb.mppc) continue;
        group_weights(pb, dist, g);
        if (g.pts.size() < pb.mppc) continue;
        cubes.push_back(std::move(g));
      }
  if (pb.sr > 0) {
    for (int at = 0; at < natom; ++at) {
      Group& g = sphere[at];
      for (int i = 0; i < natom; ++i) {
        if (i == at) {
          d2[i] = 0;
        } else {
          double dd = norm(sub(pb.pos[i], pb.pos[at]));
          double s = srad[at] > dd ? 0 : dd - srad[at];
          d2[i] = s * s;
        }
      }
      g.kind = 1;
      select_functions(pb, g, d2, minexp_f);
      if (g.pts.size() < pb.mppc) continue;
      group_weights(pb, dist, g);
      if (g.pts.size() < pb.mppc) continue;
      spheres.push_back(std::move(g));
    }
  }
  auto by_cost = [](const Group& l, const Group& r) { return l.cost() < r.cost(); };
  std::stable_sort(cubes.begin(), cubes.end(), by_cost);
  std::stable_sort(spheres.begin(), spheres.end(), by_cost);
  h.groups.clear();
  for (auto& g : cubes) h.groups.push_back(std::move(g));
  for (auto& g : spheres) h.groups.push_back(std::move(g));
}

// ----------------------------------------------------------------------------
// basis functions: values, gradients and Hessians (cpu/functions.cpp:14-359)
// f = x^lx y^ly z^lz g(r), g = sum_k c_k exp(-a_k r^2); primitives with
// a r^2 > 70 are skipped.  d xx, yy, zz carry the NORM factor.
// out[10] = f, fx, fy, fz, fxx, fyy, fzz, fxy, fxz, fyz
// ----------------------------------------------------------------------------
inline double mono(const double v[3], const int L[3]) {
  double r = 1.0;
  for (int k = 0; k < 3; ++k)
    for (int q = 0; q < L[k]; ++q) r *= v[k];
  return r;
}

void eval_function(const Problem& pb, int f, V3 pt, double out[10]) {
  const V3 A = pb.pos[pb.nuc[f]];
  const double v[3] = {pt.x - A.x, pt.y - A.y, pt.z - A.z};
  const double r2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
  double t = 0, tg = 0, th = 0;
  for (int k = 0; k < pb.ncont[f]; ++k) {
    const double a = pb.a[(size_t)f * kMaxContr + k], c = pb.c[(size_t)f * kMaxContr + k];
    if (a * r2 > 70.0) continue;
    const double e = std::exp(-a * r2) * c;
    t += e;
    tg += e * a;
    th += e * (a * a);
  }
  const int L[3] = {pb.lx[f], pb.ly[f], pb.lz[f]};
  const int l = L[0] + L[1] + L[2];
  const double scale = (l == 2 && (L[0] == 2 || L[1] == 2 || L[2] == 2)) ? pb.nf : 1.0;
  // radial part derivatives: dg/da = -2 v_a tg; d2g/dadb = 4 v_a v_b th - 2 delta tg
  double gA[3], gAB[3][3];
  for (int i = 0; i < 3; ++i) {
    gA[i] = -2.0 * v[i] * tg;
    for (int j = 0; j < 3; ++j) gAB[i][j] = 4.0 * v[i] * v[j] * th - (i == j ? 2.0 * tg : 0.0);
  }
  // monomial and its derivatives
  const double P = mono(v, L);
  double PA[3], PAB[3][3];
  for (int i = 0; i < 3; ++i) {
    int Li[3] = {L[0], L[1], L[2]};
    if (Li[i] == 0) {
      PA[i] = 0;
    } else {
      double cf = Li[i];
      Li[i] -= 1;
      PA[i] = cf * mono(v, Li);
    }
    for (int j = 0; j < 3; ++j) {
      int Lij[3] = {L[0], L[1], L[2]};
      double cf = Lij[i];
      if (Lij[i] == 0) { PAB[i][j] = 0; continue; }
      Lij[i] -= 1;
      cf *= Lij[j];
      if (Lij[j] == 0) { PAB[i][j] = 0; continue; }
      Lij[j] -= 1;
      PAB[i][j] = cf * mono(v, Lij);
    }
  }
  out[0] = P * t;
  for (int i = 0; i < 3; ++i) out[1 + i] = PA[i] * t + P * gA[i];
  auto H = [&](int i, int j) {
    return PAB[i][j] * t + PA[i] * gA[j] + PA[j] * gA[i] + P * gAB[i][j];
  };
  out[4] = H(0, 0);
  out[5] = H(1, 1);
  out[6] = H(2, 2);
  out[7] = H(0, 1);
  out[8] = H(0, 2);
  out[9] = H(1, 2);
  for (int k = 0; k < 10; ++k) out[k] *= scale;
}

// ----------------------------------------------------------------------------
// functionals
// ----------------------------------------------------------------------------
// PW92 correlation energy per particle for the unpolarised gas and its rs
// derivative, as a function of sqrt(rs) (Perdew & Wang, PRB 45, 13244 (1992),
// eq. 10 with A=0.0310907, alpha1=0.21370, beta1..4 = 7.5957, 3.5876, 1.6382,
// 0.49294).  Restates gcorc.h:21-33.
void pw92(double srs, double& ec, double& dec_drs) {
  // The reference spells these fit parameters as float literals, so its FP64
  // build evaluates them rounded to single precision (gcorc.h:201-206).
  const double A = (double)0.0310907f, a1 = (double)0.21370f, b1 = (double)7.5957f,
               b2 = (double)3.5876f, b3 = (double)1.6382f, b4 = (double)0.49294f;
  const double q0 = -2.0 * A * (1.0 + a1 * srs * srs);
  const double q1 = 2.0 * A * srs * (b1 + srs * (b2 + srs * (b3 + b4 * srs)));
  const double lg = std::log(1.0 + 1.0 / q1);
  ec = q0 * lg;
  const double q1p = A * (b1 / srs + 2.0 * b2 + srs * (3.0 * b3 + 4.0 * b4 * srs));
  dec_drs = -2.0 * A * a1 * lg - q0 * q1p / (q1 * (1.0 + q1));
}

// PBE (Perdew, Burke & Ernzerhof, PRL 77, 3865 (1996)) for a closed shell, in
// the "Laplacian" potential form of Burke's reference routine that LIO uses
// (pbeCS.h:29-185): inputs rho, |grad rho|, grad rho . grad|grad rho| and
// laplacian; outputs energies per particle and the local potentials.
void pbe(double rho, double agrad, double delgrad, double lap, double& ex, double& vx,
         double& ec, double& vc) {
  // Constants as the reference's FP64 build sees them: every one is written
  // as a float literal in pbeCS.h:8-16, so it is the single-precision value
  // widened to double; 4/3 and 1/3 are float quotients (pbeCS.h:88,90,143).
  const double kappa = (double)0.804f, mu = (double)0.2195149727645171f;
  const double ul = (double)0.273028573090195f;             // mu / kappa
  const double ax = (double)-0.738558766382022405884230032680836f;  // -(3/4)(3/pi)^(1/3)
  const double beta = (double)0.06672455060314922f;
  const double gammainv = (double)32.1639684429148f;        // pi^2 / (1 - ln 2)
  const double bg = (double)2.14612633996736f;              // beta / gamma
  const double pi32 = (double)29.608813203268075856503472999628f;  // 3 pi^2
  const double four3 = (double)(4.0f / 3.0f), third = (double)(1.0f / 3.0f);
  const double kmin = 1e-18;
  const double rho2 = rho * rho;
  const double rho13 = std::cbrt(rho);
  const double kf = std::cbrt(pi32) * rho13;  // (3 pi^2 rho)^(1/3)
  if (rho2 > kmin) {
    const double twokf = 2.0 * kf;
    const double s = agrad / (twokf * rho);
    const double s2 = s * s, s3 = s * s2;
    const double exunif = ax * rho13;
    const double den = 1.0 + ul * s2;
    const double fx = 1.0 + kappa - kappa / den;
    ex = exunif * fx;
    const double v = lap / (twokf * twokf * rho);
    const double u = delgrad / (twokf * twokf * twokf * rho2);
    const double fs = 2.0 * mu / (den * den);   // (1/s) dFx/ds
    const double fss = -4.0 * ul * s * fs / den;  // d/ds of fs
    vx = exunif * (four3 * fx - (u - four3 * s3) * fss - v * fs);
  } else {
    ex = 0.0;
    vx = 0.0;
  }
  if (rho < kmin * 1e6) {
    ec = 0.0;
    vc = 0.0;
    return;
  }
  const double rs = std::cbrt(3.0 / (4.0 * M_PI * rho));
  const double ks = std::sqrt(4.0 * kf / M_PI);  // Thomas-Fermi screening
  const double twoks = 2.0 * ks;
  const double t = agrad / (twoks * rho);
  const double t2 = t * t, t4 = t2 * t2, t6 = t4 * t2;
  const double uu = delgrad / (rho2 * twoks * twoks * twoks);
  const double vv = lap / (rho * twoks * twoks);
  double eunif, eunif_rs;
  pw92(std::sqrt(rs), eunif, eunif_rs);
  const double vlsd = eunif - rs * third * eunif_rs;
  const double A = bg / (std::exp(-eunif * gammainv) - 1.0);
  const double A2 = A * A;
  const double q4 = 1.0 + A * t2;
  const double q5 = q4 + A2 * t4;
  const double H = (beta / bg) * std::log(1.0 + bg * q4 * t2 / q5);
  ec = eunif + H;
  // gradient-correction potential (the lgga branch of Burke's CORPBE)
  const double rs3 = rs / 3.0;
  const double bec = A2 * (bg / A + 1.0) / beta;
  const double q8 = (q5 + bg * q4 * t2) * q5;
  const double q9 = 1.0 + 2.0 * A * t2;
  const double hA = -beta * A * (t6 / q8) * (2.0 + A * t2);
  const double hrs = -rs3 * hA * bec * eunif_rs;
  const double f0 = 2.0 * bg - 6.0 * A;
  const double f1 = q5 * q9 + q4 * q9 * q9;
  const double hAt = 2.0 * beta * t4 * ((q4 * q5 * f0 - bg * f1) / q8) / q8;
  const double hrst = rs3 * t2 * hAt * bec * eunif_rs;
  const double ht = 2.0 * beta * (q9 / q8);
  const double f2 = q4 * q5 + A * t2 * (q4 * q9 + q5);
  const double f3 = 2.0 * A * q5 * q9 + bg * f2;
  const double htt = 4.0 * beta * t * (2.0 * A / q8 - ((q9 / q8) * f3) / q8);
  double comm = H + hrs + hrst + t2 * ht / 6.0 + 7.0 * t2 * t * htt / 6.0;
  comm = comm - uu * htt - vv * ht;
  vc = vlsd + comm;
}

// GGA point functional from rho, grad rho, (xx,yy,zz) and (xy,xz,yz) second
// derivatives (calc_ggaCS.h:35-160): the PBE branch (iexch 0-3 and 9).
void gga_point(double rho, const double g[3], const double h1[3], const double h2[3],
               double& ex, double& ec, double& y2a) {
  if (rho < 1e-18) {
    ex = ec = y2a = 0.0;
    return;
  }
  double grad2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  if (grad2 < std::numeric_limits<double>::min()) grad2 = std::numeric_limits<double>::min();
  const double agrad = std::sqrt(grad2);
  const double delgrad = (g[0] * g[0] * h1[0] + g[1] * g[1] * h1[1] + g[2] * g[2] * h1[2] +
                          2 * (g[0] * g[1] * h2[0] + g[0] * g[2] * h2[1] + g[1] * g[2] * h2[2])) /
                         agrad;
  const double lap = h1[0] + h1[1] + h1[2];
  double vx, vc;
  pbe(rho, agrad, delgrad, lap, ex, vx, ec, vc);
  y2a = vx + vc;
}

// Spin-polarised PBE (Burke's EASYPBE driver with PBE exchange by spin
// scaling and CORPBE correlation, as LIO ships them: calc_ggaOS.h:10-108,
// pbeOS_main.h:24-202, pbeOS_exch.h:36-72, pbeOS_corr.h:54-247, gcorc.h).
// Constants are the reference's float literals widened to double; 4/3 and
// -1/6 in the potentials are float quotients there and here.
namespace os {

void pw_fit(const double c[6], double srs, double& g, double& dg) {
  const double q0 = -2.0 * c[0] * (1.0 + c[1] * srs * srs);
  const double q1 = 2.0 * c[0] * srs * (c[2] + srs * (c[3] + srs * (c[4] + c[5] * srs)));
  const double q2 = std::log(1.0 + 1.0 / q1);
  g = q0 * q2;
  const double q3 = c[0] * (c[2] / srs + 2.0 * c[3] + srs * (3.0 * c[4] + 4.0 * c[5] * srs));
  dg = -2.0 * c[0] * c[1] * q2 - q0 * q3 / (q1 * (1.0 + q1));
}

// exchange of one channel at density n = 2 rho_sigma
void exch(double n, double s, double u, double v, double& ex, double& vx) {
  const double ax = (double)-0.738558766382022405884230032680836f;
  const double um = (double)0.2195149727645171f, uk = (double)0.804f;
  const double ul = (double)0.273028573090195f;
  const double f43 = (double)(4.0f / 3.0f);
  const double exunif = ax * std::cbrt(n);
  const double s2 = s * s, s3 = s * s2;
  const double p0 = 1.0 + ul * s2;
  const double fx = 1.0 + uk - uk / p0;
  ex = exunif * fx;
  const double Fs = 2.0 * um / (p0 * p0);
  const double Fss = -4.0 * ul * s * Fs / p0;
  const double t2 = f43 * fx;
  const double t3 = v * Fs;
  const double t4 = (u - f43 * s3) * Fss;
  vx = exunif * (t2 - t4 - t3);
}

void pbe(double ra, double ga, double da, double la, double rb, double gb, double db, double lb,
         double gt, double dt, double& ex, double& ec, double& va, double& vb) {
  const double tiny = 1e-18;
  const double pi = (double)3.141592653589793238462643383279f;
  const double pi32 = (double)29.608813203268075856503472999628f;
  double exs[2] = {0, 0}, vxs[2] = {0, 0};
  const double rs_[2] = {ra, rb}, gs_[2] = {ga, gb}, ds_[2] = {da, db}, ls_[2] = {la, lb};
  for (int c = 0; c < 2; ++c) {
    const double n = 2.0 * rs_[c];
    if (!(n > tiny)) continue;
    const double fk = std::cbrt(pi32) * std::cbrt(n);
    const double k2 = 2.0 * fk;
    const double s = (2.0 * gs_[c]) / (n * k2);
    const double v = (2.0 * ls_[c]) / (n * (k2 * k2));
    const double u = (4.0 * ds_[c]) / ((n * n) * (k2 * (k2 * k2)));
    exch(n, s, u, v, exs[c], vxs[c]);
  }
  const double rho = ra + rb;
  ex = exs[0] * (ra / rho) + exs[1] * (rb / rho);
  ec = 0.0;
  va = vxs[0];
  vb = vxs[1];
  if (rho < tiny * 1e5) return;
  const double zet = (ra - rb) / rho;
  const double fk = std::cbrt(pi32) * std::cbrt(rho);
  const double rs = std::cbrt((9.0 * pi) / 4.0) / fk;
  const double g = (std::cbrt((1.0 + zet) * (1.0 + zet)) + std::cbrt((1.0 - zet) * (1.0 - zet))) / 2.0;
  const double twoksg = 2.0 * std::sqrt(4.0 * fk / pi) * g;
  const double t = gt / (twoksg * rho);
  const double uu = dt / (rho * rho * twoksg * twoksg * twoksg);
  const double vv = (la + lb) / (rho * twoksg * twoksg);
  const double ww = (ga * ga - gb * gb - zet * gt * gt) / ((rho * rho) * (twoksg * twoksg));
  // LSD part (PW92)
  static const double cu[6] = {(double)0.0310907f, (double)0.21370f, (double)7.5957f,
                               (double)3.5876f, (double)1.6382f, (double)0.49294f};
  static const double cp[6] = {(double)0.01554535f, (double)0.20548f, (double)14.1189f,
                               (double)6.1977f, (double)3.3662f, (double)0.62517f};
  static const double ca[6] = {(double)0.0168869f, (double)0.11125f, (double)10.357f,
                               (double)3.6231f, (double)0.88026f, (double)0.49671f};
  const double rt = std::sqrt(rs);
  double EU, EURS, EP, EPRS, AM, AMRS;
  pw_fit(cu, rt, EU, EURS);
  pw_fit(cp, rt, EP, EPRS);
  pw_fit(ca, rt, AM, AMRS);
  const double Z4 = (zet * zet) * (zet * zet);
  const double GAM = std::pow(2.0, 4.0 / 3.0) - 2.0;
  const double F = (std::cbrt(std::pow(1.0 + zet, 4)) + std::cbrt(std::pow(1.0 - zet, 4)) - 2.0) / GAM;
  const double FZZ = 8.0 / (9.0 * GAM);
  const double e0 = EU * (1.0 - F * Z4) + EP * F * Z4 - AM * F * (1.0 - Z4) / FZZ;
  const double ECRS = EURS * (1.0 - F * Z4) + EPRS * F * Z4 - AMRS * F * (1.0 - Z4) / FZZ;
  const double FZ = (double)(4.0f / 3.0f) * (std::cbrt(1.0 + zet) - std::cbrt(1.0 - zet)) / GAM;
  const double ECZET = 4.0 * std::pow(zet, 3) * F * (EP - EU + AM / FZZ) +
                       FZ * (Z4 * EP - Z4 * EU - (1.0 - Z4) * AM / FZZ);
  const double COMM0 = e0 - rs * ECRS / 3.0 - zet * ECZET;
  // gradient part
  const double bet = (double)0.06672455060314922f;
  const double gama = (double)0.03109069086965489503494086371273f;
  const double delt = bet / gama;
  const double G = (std::cbrt(std::pow(1.0 + zet, 2)) + std::cbrt(std::pow(1.0 - zet, 2))) / 2.0;
  const double G3 = std::pow(G, 3);
  const double B = delt / (std::exp(-e0 / (G3 * gama)) - 1.0);
  const double B2 = B * B, T2 = t * t, T4 = T2 * T2, T6 = T4 * T2;
  const double Q4 = 1.0 + B * T2;
  const double Q5 = 1.0 + B * T2 + B2 * T4;
  const double H = G3 * (bet / delt) * std::log(1.0 + delt * Q4 * T2 / Q5);
  const double G4 = G3 * G;
  const double rsthrd = rs / 3.0;
  const double m16 = (double)(-1.0f / 6.0f);
  const double eta = (double)1e-12f;
  const double GZ = (std::pow(std::pow(1.0 + zet, 2) + eta, m16) -
                     std::pow(std::pow(1.0 - zet, 2) + eta, m16)) / 3.0;
  const double FAC = delt / B + 1.0;
  const double BG = -3.0 * B2 * e0 * FAC / (bet * G4);
  const double BEC = B2 * FAC / (bet * G3);
  const double Q8 = std::pow(Q5, 2) + delt * Q4 * Q5 * T2;
  const double Q9 = 1.0 + 2.0 * B * T2;
  const double hB = -bet * G3 * B * (T6 / Q8) * (2.0 + B * T2);
  const double hRS = -rsthrd * hB * BEC * ECRS;
  const double FACT0 = 2.0 * delt - 6.0 * B;
  const double FACT1 = Q5 * Q9 + Q4 * Q9 * Q9;
  const double hBT = 2.0 * bet * G3 * T4 * ((Q4 * Q5 * FACT0 - delt * FACT1) / Q8) / Q8;
  const double hRST = rsthrd * T2 * hBT * BEC * ECRS;
  const double hZ = 3.0 * GZ * H / G + hB * (BG * GZ + BEC * ECZET);
  const double hT = 2.0 * bet * G3 * Q9 / Q8;
  const double hZT = 3.0 * GZ * hT / G + hBT * (BG * GZ + BEC * ECZET);
  const double FACT2 = Q4 * Q5 + B * T2 * (Q4 * Q9 + Q5);
  const double FACT3 = 2.0 * B * Q5 * Q9 + delt * FACT2;
  const double hTT = 4.0 * bet * G3 * t * (2.0 * B / Q8 - ((Q9 / Q8) * FACT3) / Q8);
  double COMM = H + hRS + hRST + T2 * hT / 6.0 + 7.0 * T2 * t * hTT / 6.0;
  const double PREF = hZ - GZ * T2 * hT / G;
  const double FACT5 = GZ * (2.0 * hT + t * hTT) / G;
  COMM = COMM - PREF * zet - uu * hTT - vv * hT - ww * (hZT - FACT5);
  ec = e0 + H;
  va = vxs[0] + ((COMM0 + ECZET) + (COMM + PREF));
  vb = vxs[1] + ((COMM0 - ECZET) + (COMM - PREF));
}

// calc_ggaOS: the spin channels' rho, gradient and second derivatives
void point(double ra, const double ga[3], const double h1a[3], const double h2a[3], double rb,
           const double gb[3], const double h1b[3], const double h2b[3], double& ex, double& ec,
           double& va, double& vb) {
  ex = ec = va = vb = 0.0;
  if (!(ra > 1e-18) && !(rb > 1e-18)) return;
  const double g[3] = {ga[0] + gb[0], ga[1] + gb[1], ga[2] + gb[2]};
  const double h1[3] = {h1a[0] + h1b[0], h1a[1] + h1b[1], h1a[2] + h1b[2]};
  const double h2[3] = {h2a[0] + h2b[0], h2a[1] + h2b[1], h2a[2] + h2b[2]};
  auto mod = [](const double* q) { return std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]); };
  auto dspin = [](const double* q, const double* p1, const double* p2) {
    return q[0] * q[0] * p1[0] + 2.0 * q[0] * q[1] * p2[0] + 2.0 * q[1] * q[2] * p2[2] +
           2.0 * q[0] * q[2] * p2[1] + q[1] * q[1] * p1[1] + q[2] * q[2] * p1[2];
  };
  const double na = mod(ga), nb = mod(gb), nt = mod(g);
  const double dt = (g[0] * g[0] * h1[0] + g[1] * g[1] * h1[1] + g[2] * g[2] * h1[2] +
                     2.0 * g[0] * g[1] * h2[0] + 2.0 * g[1] * g[2] * h2[2] +
                     2.0 * g[0] * g[2] * h2[1]) / nt;
  pbe(ra, na, dspin(ga, h1a, h2a) / na, h1a[0] + h1a[1] + h1a[2], rb, nb, dspin(gb, h1b, h2b) / nb,
      h1b[0] + h1b[1] + h1b[2], nt, dt, ex, ec, va, vb);
}

}  // namespace os

// LDA: Slater exchange plus (1) nothing, (2) Gunnarsson-Lundqvist or (3) VWN
// correlation (calc_ldaCS.h:32-102).
// The GGA branches of iexch 4-8 (calc_ggaCS.h:43-236), step by step in the
// reference's order of operations, its constants (calc_ggaCS.h:10-33) and its
// single-precision literals widened to double:
//   exchange     4, 8: Perdew-Wang 86 (calc_ggaCS.h:67-90)
//                5-7:  Becke 88 (calc_ggaCS.h:91-135)
//   correlation  4-6:  VWN (calc_ggaCS.h:161-186), plus for 4 and 5 Perdew's
//                      1986 gradient term (calc_ggaCS.h:188-215)
//                7, 8: Lee-Yang-Parr (calc_ggaCS.h:217-253)
void gga_point_other(double rho, const double g[3], const double h1[3], const double h2[3],
                     int iexch, double& ex, double& ec, double& y2a) {
  if (rho < 1e-18) {  // calc_ggaCS.h:43-52 (FULL_DOUBLE)
    ex = ec = y2a = 0.0;
    return;
  }
  const double alpha = -0.738558766382022447, gl = 0.620350490899400087;
  const double y = std::cbrt(rho);
  double grad2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
  if (grad2 < std::numeric_limits<double>::min()) grad2 = std::numeric_limits<double>::min();
  const double dgrad = std::sqrt(grad2);
  const double d0 = h1[0] + h1[1] + h1[2];
  const double u0 = ((g[0] * g[0]) * h1[0] + 2.0 * g[0] * g[1] * h2[0] + 2.0 * g[1] * g[2] * h2[2] +
                     2.0 * g[0] * g[2] * h2[1] + (g[1] * g[1]) * h1[1] + (g[2] * g[2]) * h1[2]) /
                    dgrad;
  double vx;
  if (iexch == 4 || iexch == 8) {  // Perdew-Wang 86 exchange
    const double rho2 = rho * rho;
    const double ckf = 3.0936677 * y;
    const double sr = dgrad / (2.0 * ckf * rho);
    const double fx = 1.0 / 15.0;
    const double sr2 = sr * sr, sr3 = sr * sr * sr;
    const double g0 = 1.0 + 1.296 * sr2 + 14.0 * std::pow(sr, 4) + 0.2 * std::pow(sr, 6);
    const double F = std::pow(g0, fx);
    ex = alpha * F * y;
    const double t = d0 / (rho * 4.0 * (ckf * ckf));
    const double u = u0 / (std::pow(2.0 * ckf, 3) * rho2);
    const double g2 = 2.592 * sr + 56.0 * sr3 + 1.2 * std::pow(sr, 5);
    const double g3 = 2.592 + 56.0 * sr2 + 1.2 * std::pow(sr, 4);
    const double g4 = 112.0 * sr + 4.8 * sr3;
    const double dF = fx * F / g0 * g2;
    const double dsF = fx * F / g0 * (-14.0 * fx * g3 * g2 / g0 + g4);
    vx = alpha * y * (1.33333333333 * F - t / sr * dF - (u - 1.3333333333 * sr3) * dsF);
  } else {  // Becke 88 exchange
    const double beta = 0.0042;
    const double e0 = alpha * y;
    const double hr = rho / 2.0;
    const double r13 = std::cbrt(hr);
    const double r43 = std::cbrt(std::pow(hr, 4));
    const double xs = dgrad / (2.0 * r43);
    const double as = std::asinh(xs);
    const double DN = 1.0 + 6.0 * beta * xs * as;
    const double ect = -2.0 * beta * r43 * xs * xs / (DN * rho);
    ex = e0 + ect;
    const double v0 = 1.33333333333333 * e0;
    const double Fb = 1.0 / DN;
    const double XA1 = xs / std::sqrt(1.0 + xs * xs);
    const double DN1 = 1.0 + Fb * (1.0 - 6.0 * beta * xs * XA1);
    const double DN2 = 1.0 / (1.0 + xs * xs) + 2.0 * Fb * (2.0 - 6.0 * beta * xs * XA1);
    const double DN3 = as * (1.0 + 2.0 * Fb) + XA1 * DN2;
    const double D02 = d0 / 2.0;
    const double de1 = 1.33333333333333 / std::cbrt(std::pow(rho, 7));
    const double c2 = std::cbrt(2.0);
    const double dgx = (g[0] * h1[0] + g[1] * h2[0] + g[2] * h2[1]) / dgrad;
    const double dgy = (g[0] * h2[0] + g[1] * h1[1] + g[2] * h2[2]) / dgrad;
    const double dgz = (g[0] * h2[1] + g[1] * h2[2] + g[2] * h1[2]) / dgrad;
    const double gx = c2 * (1.0 / (rho * y) * dgx - de1 * g[0] * dgrad);
    const double gy = c2 * (1.0 / (rho * y) * dgy - de1 * g[1] * dgrad);
    const double gz = c2 * (1.0 / (rho * y) * dgz - de1 * g[2] * dgrad);
    const double DN4 = 6.0 * beta * Fb * (g[0] / 2.0 * gx + g[1] / 2.0 * gy + g[2] / 2.0 * gz);
    const double DN5 = 1.33333333333333 * r43 * r13 * xs * xs;
    const double TOT2 = DN5 - D02 * DN1 + DN4 * DN3;
    vx = v0 + (-beta * Fb / r43 * TOT2);
  }
  double vc;
  if (iexch <= 6) {  // VWN, then Perdew 86 for 4 and 5
    const double A1 = 0.03109205, B1 = 3.72744, C1 = 12.9352, X0 = -0.10498,
                 Q = 6.15199066246304849, A16 = 0.005182008333, Q2 = 4.7309269;
    const double rho2 = rho * rho;
    const double rs = gl / y;
    const double x1 = std::sqrt(rs);
    const double Xx = rs + B1 * x1 + C1;
    const double Xxo = (X0 * X0) + B1 * X0 + C1;
    const double t1 = 2.0 * x1 + B1;
    const double t2 = std::log(Xx);
    const double t3 = std::atan(Q / t1);
    const double t4 = B1 * X0 / Xxo;
    ec = A1 * (2.0 * std::log(x1) - t2 + 2.0 * B1 / Q * t3 -
               t4 * (2.0 * std::log(x1 - X0) - t2 + 2.0 * (B1 + 2.0 * X0) / Q * t3));
    const double t5 = (B1 * x1 + 2.0 * C1) / x1;
    const double t6 = X0 / Xxo;
    vc = ec - A16 * x1 *
                  (t5 / Xx - 4.0 * B1 / ((t1 * t1) + (Q * Q2)) * (1.0 - t6 * (B1 - 2.0 * X0)) -
                   t4 * (2.0 / (x1 - X0) - t1 / Xx));
    if (iexch != 6) {
      const double alf = 0.023266, bet = 7.389, gam = 8.723, del = 0.472;
      const double rs2 = rs * rs;
      const double Cx1 = double(0.002568f) + alf * rs + bet * rs2;
      const double Cx2 = double(1.0f) + gam * rs + del * rs2 + 1.0e4 * bet * (rs * rs * rs);
      const double C = 0.001667 + Cx1 / Cx2;
      const double Cx3 = alf + double(2.0f) * bet * rs;
      const double Cx4 = gam + double(2.0f) * del * rs + 3.0e4 * bet * rs2;
      double dC = Cx3 / Cx2 - Cx1 / (Cx2 * Cx2) * Cx4;
      dC = double(-0.333333333333333f) * dC * gl / (y * rho);
      const double phi = double(0.0008129082f) / C * dgrad / std::pow(rho, double(7.0f / 6.0f));
      const double expo = std::exp(-phi);
      const double ex0 = expo * C;
      ec = ec + ex0 * grad2 / (y * rho2);
      const double D1 = (double(2.0f) - phi) * d0 / rho;
      const double phi2 = phi * phi;
      double D2 = double(1.33333333333333333f) - double(3.666666666666666666f) * phi +
                  double(1.166666666666666f) * phi2;
      D2 = D2 * grad2 / rho2;
      const double D3 = phi * (phi - double(3.0f)) * u0 / (rho * dgrad);
      const double D4 = expo * grad2 / (y * rho) * (phi2 - phi - double(1.0f)) * dC;
      vc = vc - 1.0 * (ex0 / y * (D1 - D2 + D3) - D4);
    }
  } else {  // Lee-Yang-Parr
    const double AL = 0.04918, BL = 0.132, CL = 0.2533, CL3 = 0.0844333333, DL = 0.349,
                 DL3 = 0.116333333, CF = 2.87123400018819;
    const double rom13 = 1 / std::cbrt(rho);
    const double rom53 = std::cbrt(std::pow(rho, 5));
    const double ecro = std::exp(-CL * rom13);
    const double f1 = double(1.0f) / (double(1.0f) + DL * rom13);
    const double tw = double(1.0f / 8.0f) * (grad2 / rho - d0);
    double term = (tw / double(9.0f) + d0 / double(18.0f)) - double(2.0f) * tw + CF * rom53;
    term = rho + BL * (rom13 * rom13) * ecro * term;
    ec = -AL * f1 * term / rho;
    const double hh1 = ecro / rom53;
    const double gg1 = f1 * hh1;
    const double tm1 = DL3 * (rom13 / rho);
    const double fp1 = tm1 * (f1 * f1);
    const double tm2 = double(-1.666666666f) + CL3 * rom13;
    const double hp1 = hh1 * tm2 / rho;
    const double gp1 = fp1 * hh1 + hp1 * f1;
    const double fp2 = tm1 * double(2.0f) * f1 * (fp1 - double(0.6666666666f) * f1 / rho);
    const double tm3 = double(1.6666666666f) - CL3 * double(1.3333333333f) * rom13;
    const double hp2 = hp1 * tm2 / rho + hh1 * tm3 / (rho * rho);
    const double gp2 = fp2 * hh1 + double(2.0f) * fp1 * hp1 + hp2 * f1;
    const double term3 = -AL * (fp1 * rho + f1) -
                         AL * BL * CF * (gp1 * rho + double(8.0f / 3.0f) * gg1) * rom53;
    const double term4 = (gp2 * rho * grad2 + gp1 * (double(3.0f) * grad2 + double(2.0f) * rho * d0) +
                          double(4.0f) * gg1 * d0) * AL * BL / double(4.0f);
    const double term5 = (double(3.0f) * gp2 * rho * grad2 +
                          gp1 * (double(5.0f) * grad2 + double(6.0f) * rho * d0) +
                          double(4.0f) * gg1 * d0) * AL * BL / double(72.0f);
    vc = term3 - term4 - term5;
  }
  y2a = vx + vc;
}

// The GGA point of LIO's closed-shell path for any iexch of the GGA branch.
void gga_point_lio(double rho, const double g[3], const double h1[3], const double h2[3],
                   int iexch, double& ex, double& ec, double& y2a) {
  if (iexch >= 4 && iexch <= 8)
    gga_point_other(rho, g, h1, h2, iexch, ex, ec, y2a);
  else
    gga_point(rho, g, h1, h2, ex, ec, y2a);
}

void lda_point(double rho, int iexch, double& ex, double& ec, double& y2a) {
  if (rho == 0) {
    ex = ec = y2a = 0;
    return;
  }
  const double y = std::cbrt(rho);
  const double alpha = -0.738558766382022447;  // -(3/4)(3/pi)^(1/3)
  const double gl = 0.620350490899400087;      // (3/(4 pi))^(1/3)
  const double v0 = -0.984745021842697 * y;    // (4/3) alpha rho^(1/3)
  ex = alpha * y;
  if (iexch == 1) {
    ec = 0;
    y2a = v0;
  } else if (iexch == 2) {
    const double rs = gl / y;
    const double x = rs / 11.4;
    double vcl;
    if (x > 1.0) {
      ec = -0.0333 * (0.5 * x - 0.33333333333333);
      vcl = 0.0111 * x * 0.5;
    } else {
      const double t1 = 1.0 + x * x * x, t2 = std::log(1.0 + 1.0 / x), t3 = x * x;
      ec = -0.0333 * (t1 * t2 - t3 + 0.5 * x - 0.33333333333333);
      vcl = 0.0111 * x * (3.0 * t3 * t2 - t1 / (x * (x + 1.0)) - 2.0 * x + 0.5);
    }
    y2a = v0 + ec + vcl;
  } else {
    // VWN (Vosko, Wilk & Nusair, Can. J. Phys. 58, 1200 (1980)), paramagnetic
    const double A = 0.03109205, b = 3.72744, cc = 12.9352, x0 = -0.10498;
    const double Q = 6.15199066246304849;
    const double rs = gl / y;
    const double x = std::sqrt(rs);
    const double X = rs + b * x + cc;
    const double X0 = x0 * x0 + b * x0 + cc;
    const double t1 = 2.0 * x + b, t2 = std::log(X), t3 = std::atan(Q / t1);
    const double t4 = b * x0 / X0;
    ec = A * (2.0 * std::log(x) - t2 + 2.0 * b / Q * t3 -
              t4 * (2.0 * std::log(x - x0) - t2 + 2.0 * (b + 2.0 * x0) / Q * t3));
    const double t5 = (b * x + 2.0 * cc) / x;
    const double t6 = x0 / X0;
    const double vcl = ec - A / 6.0 * x *
                                (t5 / X - 4.0 * b / (t1 * t1 + Q * Q) * (1.0 - t6 * (b - 2.0 * x0)) -
                                 t4 * (2.0 / (x - x0) - t1 / X));
    y2a = v0 + vcl;
  }
}

// ----------------------------------------------------------------------------
// per-group solve (cpu/iteration.cpp:39-309, closed shell, Fock + energy)
// ----------------------------------------------------------------------------
inline size_t packed_index(int i, int j, int M) {
  if (i > j) std::swap(i, j);
  return (size_t)i * M - ((size_t)i * (i - 1)) / 2 + (j - i);
}

void solve_group(const Problem& pb, const Group& g, const double* rho, double& energy,
                 double& electrons, double* F, double* forces) {
  const int m = g.m(), n = (int)g.pts.size(), M = pb.M;
  // the gradient is also needed for LDA forces (cpu/functions.cpp:25-32)
  const int nv = pb.gga ? 10 : (forces ? 4 : 1);
  // values: [point][value][function]
  std::vector<double> fv((size_t)n * nv * m);
  for (int p = 0; p < n; ++p)
    for (int i = 0; i < m; ++i) {
      double o[10];
      eval_function(pb, g.fglobal[i], g.pts[p], o);
      for (int v = 0; v < nv; ++v) fv[((size_t)p * nv + v) * m + i] = o[v];
    }
  // P' block: packed values (off-diagonals already doubled), symmetric fill
  std::vector<double> Pl((size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) Pl[(size_t)i * m + j] = rho[packed_index(g.fglobal[i], g.fglobal[j], M)];
  std::vector<double> factor(n);
  for (int p = 0; p < n; ++p) {
    const double* f = &fv[(size_t)p * nv * m];
    double ex = 0, ec = 0, y2a = 0, dens = 0;
    if (!pb.gga) {
      for (int i = 0; i < m; ++i) {
        double w = 0;
        for (int j = i; j < m; ++j) w += Pl[(size_t)j * m + i] * f[j];
        dens += f[i] * w;
      }
      lda_point(dens, pb.iexch, ex, ec, y2a);
    } else {
      const double *gx = f + m, *gy = f + 2 * m, *gz = f + 3 * m;
      const double *hxx = f + 4 * m, *hyy = f + 5 * m, *hzz = f + 6 * m;
      const double *hxy = f + 7 * m, *hxz = f + 8 * m, *hyz = f + 9 * m;
      double dx = 0, dy = 0, dz = 0, d1x = 0, d1y = 0, d1z = 0, d2x = 0, d2y = 0, d2z = 0;
      for (int i = 0; i < m; ++i) {
        double w = 0, wx = 0, wy = 0, wz = 0, w1x = 0, w1y = 0, w1z = 0, w2x = 0, w2y = 0,
               w2z = 0;
        const double* Pi = &Pl[(size_t)i * m];
        for (int j = 0; j <= i; ++j) {
          const double pij = Pi[j];
          w += f[j] * pij;
          wx += gx[j] * pij;
          wy += gy[j] * pij;
          wz += gz[j] * pij;
          w1x += hxx[j] * pij;
          w1y += hyy[j] * pij;
          w1z += hzz[j] * pij;
          w2x += hxy[j] * pij;
          w2y += hxz[j] * pij;
          w2z += hyz[j] * pij;
        }
        const double fi = f[i];
        dens += fi * w;
        dx += gx[i] * w + wx * fi;
        dy += gy[i] * w + wy * fi;
        dz += gz[i] * w + wz * fi;
        d1x += gx[i] * wx * 2 + hxx[i] * w + w1x * fi;
        d1y += gy[i] * wy * 2 + hyy[i] * w + w1y * fi;
        d1z += gz[i] * wz * 2 + hzz[i] * w + w1z * fi;
        d2x += gx[i] * wy + gy[i] * wx + hxy[i] * w + w2x * fi;
        d2y += gx[i] * wz + gz[i] * wx + hxz[i] * w + w2y * fi;
        d2z += gy[i] * wz + gz[i] * wy + hyz[i] * w + w2z * fi;
      }
      const double gr[3] = {dx, dy, dz}, h1[3] = {d1x, d1y, d1z}, h2[3] = {d2x, d2y, d2z};
      gga_point_lio(dens, gr, h1, h2, pb.iexch, ex, ec, y2a);
    }
    const double wp = g.w[p];
    energy += (dens * wp) * (ex + ec);
    electrons += dens * wp;
    factor[p] = wp * y2a;
  }
  // XC forces (cpu/iteration.cpp:213-268): per point and nucleus,
  // dd = -sum_{i on the nucleus} w_i grad phi_i with w_i = sum_j P'_ji phi_j
  // (x2 on the diagonal), scaled by the point's factor; summed over points in
  // order, then added to the nucleus' row of the (natom x 3) force array
  if (forces) {
    std::vector<double> fsum((size_t)g.nucs.size() * 3, 0.0);
    std::vector<double> dd(fsum.size());
    for (int p = 0; p < n; ++p) {
      const double* f = &fv[(size_t)p * nv * m];
      std::fill(dd.begin(), dd.end(), 0.0);
      for (int i = 0; i < m; ++i) {
        double w = 0.0;
        for (int j = 0; j < m; ++j) w += Pl[(size_t)j * m + i] * f[j] * (i == j ? 2 : 1);
        const int a = (int)(std::lower_bound(g.nucs.begin(), g.nucs.end(),
                                             pb.nuc[g.fglobal[i]]) - g.nucs.begin());
        dd[3 * a + 0] -= w * f[m + i];
        dd[3 * a + 1] -= w * f[2 * m + i];
        dd[3 * a + 2] -= w * f[3 * m + i];
      }
      for (size_t k = 0; k < dd.size(); ++k) fsum[k] += dd[k] * factor[p];
    }
    for (size_t a = 0; a < g.nucs.size(); ++a)
      for (int c = 0; c < 3; ++c) forces[3 * (size_t)g.nucs[a] + c] += fsum[3 * a + c];
  }
  if (!F) return;
  // Fock: F_big += sum_p phi_i phi_j factor_p over local pairs i <= j
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) {
      double acc = 0.0;
      for (int p = 0; p < n; ++p) {
        const double* f = &fv[(size_t)p * nv * m];
        acc += f[i] * f[j] * factor[p];
      }
      F[packed_index(g.fglobal[i], g.fglobal[j], M)] += acc;
    }
}

// per-group open-shell solve (cpu/iteration.cpp:311-622).  The reference
// computes nothing for an LDA in open shell (its LDA branch is empty): no
// energy, no Fock, no forces; for PBE it builds both spin densities, calls the
// spin-polarised functional and adds sum_p phi_i phi_j w_p v_sigma(p) to each
// spin's Fock.  Forces add both spins' contributions per nucleus.
void solve_group_open(const Problem& pb, const Group& g, const double* rho_a, const double* rho_b,
                      double& energy, double& electrons, double* Fa, double* Fb, double* forces) {
  if (!pb.gga) return;
  const int m = g.m(), n = (int)g.pts.size(), M = pb.M;
  const int nv = 10;
  std::vector<double> fv((size_t)n * nv * m);
  for (int p = 0; p < n; ++p)
    for (int i = 0; i < m; ++i) {
      double o[10];
      eval_function(pb, g.fglobal[i], g.pts[p], o);
      for (int v = 0; v < nv; ++v) fv[((size_t)p * nv + v) * m + i] = o[v];
    }
  std::vector<double> Pa((size_t)m * m), Pb((size_t)m * m);
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) {
      const size_t k = packed_index(g.fglobal[i], g.fglobal[j], M);
      Pa[(size_t)i * m + j] = rho_a[k];
      Pb[(size_t)i * m + j] = rho_b[k];
    }
  // rho, gradient, (xx,yy,zz), (xy,xz,yz) of one spin at point p
  auto channel = [&](const std::vector<double>& Pl, const double* f, double q[10]) {
    const double *gx = f + m, *gy = f + 2 * m, *gz = f + 3 * m;
    const double *hxx = f + 4 * m, *hyy = f + 5 * m, *hzz = f + 6 * m;
    const double *hxy = f + 7 * m, *hxz = f + 8 * m, *hyz = f + 9 * m;
    for (int k = 0; k < 10; ++k) q[k] = 0.0;
    for (int i = 0; i < m; ++i) {
      double w = 0, wx = 0, wy = 0, wz = 0, w1x = 0, w1y = 0, w1z = 0, w2x = 0, w2y = 0, w2z = 0;
      const double* Pi = &Pl[(size_t)i * m];
      for (int j = 0; j <= i; ++j) {
        const double pij = Pi[j];
        w += f[j] * pij;
        wx += gx[j] * pij;
        wy += gy[j] * pij;
        wz += gz[j] * pij;
        w1x += hxx[j] * pij;
        w1y += hyy[j] * pij;
        w1z += hzz[j] * pij;
        w2x += hxy[j] * pij;
        w2y += hxz[j] * pij;
        w2z += hyz[j] * pij;
      }
      const double fi = f[i];
      q[0] += fi * w;
      q[1] += gx[i] * w + wx * fi;
      q[4] += gx[i] * wx * 2 + hxx[i] * w + w1x * fi;
      q[7] += gx[i] * wy + gy[i] * wx + hxy[i] * w + w2x * fi;
      q[2] += gy[i] * w + wy * fi;
      q[5] += gy[i] * wy * 2 + hyy[i] * w + w1y * fi;
      q[8] += gx[i] * wz + gz[i] * wx + hxz[i] * w + w2y * fi;
      q[3] += gz[i] * w + wz * fi;
      q[6] += gz[i] * wz * 2 + hzz[i] * w + w1z * fi;
      q[9] += gy[i] * wz + gz[i] * wy + hyz[i] * w + w2z * fi;
    }
  };
  std::vector<double> fa(n), fb(n);
  for (int p = 0; p < n; ++p) {
    const double* f = &fv[(size_t)p * nv * m];
    double qa[10], qb[10];
    channel(Pa, f, qa);
    channel(Pb, f, qb);
    double ex, ec, va, vb;
    os::point(qa[0], qa + 1, qa + 4, qa + 7, qb[0], qb + 1, qb + 4, qb + 7, ex, ec, va, vb);
    const double wp = g.w[p];
    energy += ((qa[0] + qb[0]) * wp) * (ex + ec);
    electrons += (qa[0] + qb[0]) * wp;
    fa[p] = wp * va;
    fb[p] = wp * vb;
  }
  if (forces) {
    const size_t nn = g.nucs.size();
    std::vector<double> sa(3 * nn, 0.0), sb(3 * nn, 0.0), da(3 * nn), db(3 * nn);
    for (int p = 0; p < n; ++p) {
      const double* f = &fv[(size_t)p * nv * m];
      std::fill(da.begin(), da.end(), 0.0);
      std::fill(db.begin(), db.end(), 0.0);
      for (int i = 0; i < m; ++i) {
        double wa = 0.0, wb = 0.0;
        for (int j = 0; j < m; ++j) {
          const double two = (i == j ? 2 : 1);
          wa += Pa[(size_t)j * m + i] * f[j] * two;
          wb += Pb[(size_t)j * m + i] * f[j] * two;
        }
        const size_t a = std::lower_bound(g.nucs.begin(), g.nucs.end(), pb.nuc[g.fglobal[i]]) -
                         g.nucs.begin();
        for (int c = 0; c < 3; ++c) {
          da[3 * a + c] -= wa * f[(1 + c) * m + i];
          db[3 * a + c] -= wb * f[(1 + c) * m + i];
        }
      }
      for (size_t k = 0; k < 3 * nn; ++k) {
        sa[k] += da[k] * fa[p];
        sb[k] += db[k] * fb[p];
      }
    }
    for (size_t a = 0; a < nn; ++a)
      for (int c = 0; c < 3; ++c) forces[3 * (size_t)g.nucs[a] + c] += sa[3 * a + c] + sb[3 * a + c];
  }
  if (!Fa) return;
  for (int i = 0; i < m; ++i)
    for (int j = i; j < m; ++j) {
      double acc_a = 0.0, acc_b = 0.0;
      for (int p = 0; p < n; ++p) {
        const double* f = &fv[(size_t)p * nv * m];
        acc_a += f[i] * f[j] * fa[p];
        acc_b += f[i] * f[j] * fb[p];
      }
      const size_t k = packed_index(g.fglobal[i], g.fglobal[j], M);
      Fa[k] += acc_a;
      Fb[k] += acc_b;
    }
}

double solve_range(Handle& h, const double* rho_in, int first, int count, double* energy,
                   double* F, double* electrons, double* points, double* forces = nullptr,
                   const double* rho_b_in = nullptr, double* Fb = nullptr) {
  const Problem& pb = h.pb;
  const double* rho = rho_in ? rho_in : pb.rho.data();
  const double* rho_b = pb.open ? (rho_b_in ? rho_b_in : pb.rho_b.data()) : nullptr;
  const size_t packed = (size_t)pb.M * (pb.M + 1) / 2;
  const size_t nforce = forces ? 3 * (size_t)pb.natom : 0;
  const int last = std::min<int>((int)h.groups.size(), first + count);
  double e = 0.0, ne = 0.0, pts = 0.0;
  const double t0 = now();
#pragma omp parallel reduction(+ : e, ne, pts)
  {
    std::vector<double> Fl(F ? packed : 0, 0.0);
    std::vector<double> Flb(F && pb.open ? packed : 0, 0.0);
    std::vector<double> fl(nforce, 0.0);
#pragma omp for schedule(dynamic, 1)
    for (int gi = first; gi < last; ++gi) {
      double ge = 0.0, gn = 0.0;
      if (pb.open)
        solve_group_open(pb, h.groups[gi], rho, rho_b, ge, gn, F ? Fl.data() : nullptr,
                         F ? Flb.data() : nullptr, forces ? fl.data() : nullptr);
      else
        solve_group(pb, h.groups[gi], rho, ge, gn, F ? Fl.data() : nullptr,
                    forces ? fl.data() : nullptr);
      e += ge;
      ne += gn;
      pts += (double)h.groups[gi].pts.size();
    }
    if (F) {
#pragma omp critical
      for (size_t k = 0; k < packed; ++k) F[k] += Fl[k];
    }
    if (F && pb.open && Fb) {
#pragma omp critical
      for (size_t k = 0; k < packed; ++k) Fb[k] += Flb[k];
    }
    if (forces) {
#pragma omp critical
      for (size_t k = 0; k < nforce; ++k) forces[k] += fl[k];
    }
  }
  const double t1 = now();
  if (energy) *energy = e;
  if (electrons) *electrons = ne;
  if (points) *points = pts;
  return t1 - t0;
}

}  // namespace

extern "C" {

void* oracle_setup(const XcInput* in) {
  try {
    auto* h = new Handle;
    load(h->pb, in);
    const double t0 = now();
    build_partition(*h);
    h->grid_seconds = now() - t0;
    return h;
  } catch (...) {
    return nullptr;
  }
}

double oracle_grid_seconds(void* hp) { return static_cast<Handle*>(hp)->grid_seconds; }

// Full XC build: energy assigned, fock accumulated (+=), electrons assigned.
double oracle_solve(void* hp, const double* rho, double* energy, double* fock,
                    double* electrons) {
  auto* h = static_cast<Handle*>(hp);
  return solve_range(*h, rho, 0, (int)h->groups.size(), energy, fock, electrons, nullptr);
}

// Energy and XC forces (computation type 2, no Fock): energy assigned,
// forces (natom x 3, atom-major) accumulated (+=), electrons assigned.
double oracle_solve_forces(void* hp, const double* rho, double* energy, double* forces,
                           double* electrons) {
  auto* h = static_cast<Handle*>(hp);
  return solve_range(*h, rho, 0, (int)h->groups.size(), energy, nullptr, electrons, nullptr,
                     forces);
}

// Open shell: energy assigned, the two spins' Fock matrices accumulated (+=)
// when fock_a is given, forces (natom x 3, atom-major) accumulated when
// given, electrons assigned.
double oracle_solve_open(void* hp, const double* rho_a, const double* rho_b, double* energy,
                         double* fock_a, double* fock_b, double* forces, double* electrons) {
  auto* h = static_cast<Handle*>(hp);
  if (!h->pb.open) return -1.0;
  return solve_range(*h, rho_a, 0, (int)h->groups.size(), energy, fock_a, electrons, nullptr,
                     forces, rho_b, fock_b);
}

// Bounded sample for CPU timing: groups [first, first+count) of the cost-sorted
// list (cubes then spheres), Fock accumulation included but discarded.
double oracle_solve_sample(void* hp, const double* rho, int first, int count, double* energy,
                           double* points) {
  auto* h = static_cast<Handle*>(hp);
  const size_t packed = (size_t)h->pb.M * (h->pb.M + 1) / 2;
  std::vector<double> F(packed, 0.0);
  return solve_range(*h, rho, first, count, energy, F.data(), nullptr, points);
}

int oracle_groups(void* hp, XcGroupInfo* out, int max) {
  auto* h = static_cast<Handle*>(hp);
  int k = 0;
  for (const Group& g : h->groups) {
    if (out && k < max) {
      XcGroupInfo& o = out[k];
      o.kind = g.kind;
      o.npoints = (int)g.pts.size();
      o.nfuncs = g.m();
      o.s = g.s;
      o.p = g.p;
      o.d = g.d;
      o.nnuc = (int)g.nucs.size();
      o.wsum = 0;
      o.xsum = 0;
      for (size_t p = 0; p < g.pts.size(); ++p) {
        o.wsum += g.w[p];
        o.xsum += g.pts[p].x;
      }
      std::vector<unsigned> fl(g.shell_first.begin(), g.shell_first.end());
      unsigned long long hh = 1469598103934665603ull;
      for (unsigned x : fl)
        for (int b = 0; b < 4; ++b) {
          hh ^= (x >> (8 * b)) & 0xffu;
          hh *= 1099511628211ull;
        }
      o.fhash = hh;
    }
    ++k;
  }
  return k;
}

// (kind, npoints, nfuncs, 10 n m^2) per group, cost-sorted order.
int oracle_group_work(void* hp, long long* out) {
  auto* h = static_cast<Handle*>(hp);
  int k = 0;
  for (const Group& g : h->groups) {
    if (out) {
      long long n = (long long)g.pts.size(), m = g.m();
      out[4 * k + 0] = g.kind;
      out[4 * k + 1] = n;
      out[4 * k + 2] = m;
      out[4 * k + 3] = 10 * n * m * m;
    }
    ++k;
  }
  return k;
}

void oracle_free(void* hp) { delete static_cast<Handle*>(hp); }

void oracle_gga_point(double rho, const double* g, const double* h1, const double* h2,
                      int iexch, double* ex, double* ec, double* y2a) {
  gga_point_lio(rho, g, h1, h2, iexch, *ex, *ec, *y2a);
}

void oracle_lda_point(double rho, int iexch, double* ex, double* ec, double* y2a) {
  lda_point(rho, iexch, *ex, *ec, *y2a);
}

// Basis values at arbitrary points: out[(p*M + f)*10 + v].
int oracle_basis_values(const XcInput* in, int npts, const double* pts, int M, double* out) {
  try {
    Problem pb;
    load(pb, in);
    if (M != pb.M) return 1;
    for (int p = 0; p < npts; ++p) {
      V3 q{pts[3 * p], pts[3 * p + 1], pts[3 * p + 2]};
      for (int f = 0; f < M; ++f) eval_function(pb, f, q, &out[((size_t)p * M + f) * 10]);
    }
    return 0;
  } catch (...) {
    return 2;
  }
}

// Vectorised point functionals for tests: arrays of n points; grad, h1 (xx,
// yy, zz) and h2 (xy, xz, yz) are n x 3 row-major.
void oracle_gga_points(int n, const double* rho, const double* grad, const double* h1,
                       const double* h2, double* ex, double* ec, double* y2a) {
  for (int i = 0; i < n; ++i) gga_point(rho[i], grad + 3 * i, h1 + 3 * i, h2 + 3 * i, ex[i], ec[i], y2a[i]);
}

// The same for any iexch of the GGA branch (4-8, PBE otherwise).
void oracle_gga_points_x(int n, const double* rho, const double* grad, const double* h1,
                         const double* h2, int iexch, double* ex, double* ec, double* y2a) {
  for (int i = 0; i < n; ++i)
    gga_point_lio(rho[i], grad + 3 * i, h1 + 3 * i, h2 + 3 * i, iexch, ex[i], ec[i], y2a[i]);
}

// Spin-polarised PBE points: per point the two channels' rho, gradient (3),
// (xx,yy,zz) and (xy,xz,yz); outputs ex, ec per particle and both potentials.
void oracle_gga_points_open(int n, const double* ra, const double* ga, const double* h1a,
                            const double* h2a, const double* rb, const double* gb,
                            const double* h1b, const double* h2b, double* ex, double* ec,
                            double* va, double* vb) {
  for (int i = 0; i < n; ++i)
    os::point(ra[i], ga + 3 * i, h1a + 3 * i, h2a + 3 * i, rb[i], gb + 3 * i, h1b + 3 * i,
              h2b + 3 * i, ex[i], ec[i], va[i], vb[i]);
}

void oracle_lda_points(int n, const double* rho, int iexch, double* ex, double* ec, double* y2a) {
  for (int i = 0; i < n; ++i) lda_point(rho[i], iexch, ex[i], ec[i], y2a[i]);
}

void oracle_pw92(double srs, double* e, double* de_drs) { pw92(srs, *e, *de_drs); }

// Grid export for tests: total point count, then points (x, y, z, w) and the
// group index of each point, groups in cost-sorted order.
long long oracle_grid_points(void* hp, double* xyzw, int* group) {
  auto* h = static_cast<Handle*>(hp);
  long long k = 0;
  for (size_t g = 0; g < h->groups.size(); ++g) {
    const Group& G = h->groups[g];
    for (size_t p = 0; p < G.pts.size(); ++p, ++k) {
      if (xyzw) {
        xyzw[4 * k + 0] = G.pts[p].x;
        xyzw[4 * k + 1] = G.pts[p].y;
        xyzw[4 * k + 2] = G.pts[p].z;
        xyzw[4 * k + 3] = G.w[p];
      }
      if (group) group[k] = (int)g;
    }
  }
  return k;
}

// Expanded function list of group g (global indices); returns its length.
int oracle_group_functions(void* hp, int g, int* out) {
  auto* h = static_cast<Handle*>(hp);
  const Group& G = h->groups[g];
  if (out)
    for (int i = 0; i < G.m(); ++i) out[i] = G.fglobal[i];
  return G.m();
}

}  // extern "C"
