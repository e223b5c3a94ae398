// Host grid and point-group partition.
//
// Produces exactly the groups LIO's g2g defines (the discrete decisions --
// prism clipping, sphere/cube assignment, significant functions and
// zero-weight removal -- decide which basis products enter the quadrature,
// so they are reproduced expression for expression):
//   points and groups ......... g2g/regenerate_partition.cpp:161-462
//   significant functions ..... g2g/classify_functions.cpp:12-141,
//                               g2g/partition.cpp:229-254
//   group-local Becke weights . g2g/cpu/weight.cpp:13-112
// The result is flattened (structure of arrays) for a single upload.
#include <omp.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <limits>
#include <stdexcept>

#include "xcq.h"

namespace xcq {

namespace {

double wall() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// Cartesian powers of the components of a p (x,y,z) and d (xx,xy,yy,xz,yz,zz)
// shell in g2g's order (g2g/cpu/functions.cpp:105-170).
const int kP[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
const int kD[6][3] = {{2, 0, 0}, {1, 1, 0}, {0, 2, 0}, {1, 0, 1}, {0, 1, 1}, {0, 0, 2}};

// Significance test of one shell (g2g/partition.cpp:229-254): an s shell is
// kept when a d^2 < mfe - ln((2a/pi)^3)/4; p and d shells when d is inside
// the radius found by fixed-point iteration on r = sqrt((mfe - ln N + l ln r)/a).
// shell_radius gives the right-hand side (s) or the radius (p, d) once per
// shell; shell_keeps applies it to a squared distance.
double shell_radius(int len, double a, unsigned mfe) {
  if (len == 1) return mfe - std::log(std::pow((2. * a / M_PI), 3)) / 4;
  double nrm = std::pow((2.0 * a / M_PI), 3);
  nrm = std::sqrt(nrm * 4.0 * a);
  const double lnn = std::log(len == 3 ? std::sqrt(nrm) : std::fabs(nrm));
  double r = 1.0, step;
  do {
    const double lr = (len == 3 ? std::log(r) : 2 * std::log(r));
    const double r_next = std::sqrt((mfe - lnn + lr) / a);
    step = std::fabs(r - r_next);
    r = r_next;
  } while (step > 0.1);
  return r;
}

inline bool shell_keeps(int len, double a, double thr, double d2) {
  return len == 1 ? a * d2 < thr : std::sqrt(d2) < thr;
}

struct Staging {
  std::vector<double> x, y, z, w;
  std::vector<int> atom;
};

struct Selected {
  std::vector<int> first;  // first component of each selected shell
  std::vector<int> len;
  std::vector<int> atoms;  // sorted owners
  int ns = 0, np = 0, nd = 0;
};

void select(const Molecule& mol, const Options& opt, const GridFrame& fr,
            const std::vector<double>& d2, Selected& s) {
  std::vector<char> owner(mol.natom, 0);
  for (size_t k = 0; k < fr.sh_first.size(); ++k) {
    const int at = fr.sh_atom[k], len = fr.sh_len[k];
    if (opt.assign_all_functions || shell_keeps(len, fr.sh_a[k], fr.sh_thr[k], d2[at])) {
      s.first.push_back(fr.sh_first[k]);
      s.len.push_back(len);
      owner[at] = 1;
      (len == 1 ? s.ns : len == 3 ? s.np : s.nd)++;
    }
  }
  for (int a = 0; a < mol.natom; ++a)
    if (owner[a]) s.atoms.push_back(a);
}

// The group's atoms nearest-first from (x, y, z) (its first staged point):
// list positions sorted by distance, ties in list order.  The device forms
// the same order with the same operations (k_becke_order).
std::vector<int> becke_order(const Molecule& mol, const std::vector<int>& atoms, double x,
                             double y, double z) {
  const int k = (int)atoms.size();
  std::vector<double> d(k);
  for (int q = 0; q < k; ++q) {
    const int a = atoms[q];
    const double dx = x - mol.x[a], dy = y - mol.y[a], dz = z - mol.z[a];
    d[q] = std::sqrt(dx * dx + dy * dy + dz * dz);
  }
  std::vector<int> ord(k);
  for (int q = 0; q < k; ++q) ord[q] = q;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return d[a] < d[b]; });
  return ord;
}

// Becke cell function s(mu_jk) with the size adjustment of the atom pair.
inline double cell(double dj, double dk, double rjk, double rmj, double rmk) {
  double u = (dj - dk) / rjk;
  double chi = rmj / rmk;
  chi = (chi - 1.0) / (chi + 1.0);
  u += (chi / (chi * chi - 1.0)) * (1.0 - u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  return 0.5 * (1.0 - u);
}

// Multiplies the staged weights by the Becke partition over the group's own
// atoms and drops points whose weight became exactly zero:
// w *= P_own / sum_j P_j, P_j = prod_{q != j} s(mu_jq), every factor in [0, 1].
// P_own is formed exactly as the reference forms it (a zero test against the
// point's nearest atom n, then the atoms in list order), so a weight is zero
// exactly when the reference's is, and a zero P_own ends the point; so is P_n
// (no zero test), a term of the sum and in general its largest.  Every other
// P_j runs over the group's atoms nearest-first (becke_order: the same order
// for all points of the group) and is dropped once its running product -- an
// upper bound of P_j -- falls below 2^-60 of the larger of P_own and P_n:
// the sum holds that term, so a dropped one lies far below half an ulp of it.  The sum runs over
// kBeckeLanes interleaved classes of j (j mod kBeckeLanes, each in list
// order from 0.0), which are then added in class order: the device gives each
// class to its own wave.  The sum differs from the reference's by a few ulps
// at most.  A point whose own atom is not among the group's does the same
// with P_own formed over all of the group's atoms, left out of the sum and
// out of the bound.
void weigh(const Molecule& mol, const Options& opt, const std::vector<double>& rij,
           const std::vector<int>& atoms, Staging& st) {
  const int n = (int)st.x.size(), k = (int)atoms.size(), na = mol.natom;
  const std::vector<int> ord = becke_order(mol, atoms, n > 0 ? st.x[0] : 0.0,
                                           n > 0 ? st.y[0] : 0.0, n > 0 ? st.z[0] : 0.0);
#pragma omp parallel for schedule(dynamic, 64) if (n > 256)
  for (int p = 0; p < n; ++p) {
    const int own = st.atom[p];
    double dist[512];
    std::vector<double> big;
    double* d = dist;
    if (k + 1 > 512) {
      big.resize(k + 1);
      d = big.data();
    }
    int qn = -1, qo = -1;
    for (int q = 0; q < k; ++q) {
      const int a = atoms[q];
      const double dx = st.x[p] - mol.x[a], dy = st.y[p] - mol.y[a], dz = st.z[p] - mol.z[a];
      d[q] = std::sqrt(dx * dx + dy * dy + dz * dz);
      if (a == own) qo = q;
      if (qn < 0 || d[q] < d[qn]) qn = q;
    }
    if (k == 0) {
      st.w[p] = 0.0;  // no atoms: the sum is empty
      continue;
    }
    const int an = atoms[qn];
    // the reference's P_j of the atom at distance dj (zero test against the
    // nearest atom unless j is the nearest, then the atoms in list order)
    auto exact = [&](int aj, double dj) {
      if (aj != an && cell(dj, d[qn], rij[(size_t)aj * na + an], mol.rm[aj], mol.rm[an]) == 0.0)
        return 0.0;
      double prod = 1.0;
      for (int q = 0; q < k; ++q) {
        const int ak = atoms[q];
        if (ak == aj) continue;
        prod *= cell(dj, d[q], rij[(size_t)aj * na + ak], mol.rm[aj], mol.rm[ak]);
        if (prod == 0.0) break;
      }
      return prod;
    };
    double down = 0.0;
    if (qo >= 0) {
      down = d[qo];
    } else {
      const double dx = st.x[p] - mol.x[own], dy = st.y[p] - mol.y[own], dz = st.z[p] - mol.z[own];
      down = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
    const double mine = exact(own, down);
    if (mine == 0.0) {
      st.w[p] = 0.0;
      continue;
    }
    const double pn = qn == qo ? mine : exact(an, d[qn]);
    // a bound below the sum: the largest of its exact terms
    const double thr = std::ldexp(qo >= 0 ? std::max(mine, pn) : pn, -60);
    double total = 0.0;
    for (int c = 0; c < kBeckeLanes; ++c) {
      double sum = 0.0;
      for (int j = c; j < k; j += kBeckeLanes) {
        if (j == qo) {
          sum += mine;
          continue;
        }
        if (j == qn) {
          sum += pn;
          continue;
        }
        const int aj = atoms[j];
        if (cell(d[j], d[qn], rij[(size_t)aj * na + an], mol.rm[aj], mol.rm[an]) == 0.0) continue;
        double prod = 1.0;
        for (int t = 0; t < k; ++t) {
          const int q = ord[t], ak = atoms[q];
          if (ak == aj) continue;
          prod *= cell(d[j], d[q], rij[(size_t)aj * na + ak], mol.rm[aj], mol.rm[ak]);
          if (prod < thr) {
            prod = 0.0;
            break;
          }
        }
        sum += prod;
      }
      total += sum;
    }
    st.w[p] *= (total == 0.0 ? 0.0 : mine / total);
  }
  if (!opt.remove_zero_weights) return;
  int keep = 0;
  for (int p = 0; p < n; ++p) {
    if (st.w[p] == 0.0) continue;
    st.x[keep] = st.x[p];
    st.y[keep] = st.y[p];
    st.z[keep] = st.z[p];
    st.w[keep] = st.w[p];
    st.atom[keep] = st.atom[p];
    ++keep;
  }
  st.x.resize(keep);
  st.y.resize(keep);
  st.z.resize(keep);
  st.w.resize(keep);
  st.atom.resize(keep);
}

uint64_t fnv1a(const std::vector<int>& v) {
  uint64_t h = 1469598103934665603ull;
  for (int xi : v) {
    const unsigned x = (unsigned)xi;
    for (int b = 0; b < 4; ++b) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

}  // namespace

Molecule capture(const ProblemView& pv, unsigned grid_type) {
  if (grid_type > 2) throw std::runtime_error("g2g: unknown grid type");
  if (pv.iexch < 1 || pv.iexch > 9) throw std::runtime_error("g2g: Iexch must be in 1..9");
  Molecule mol;
  mol.natom = (int)pv.natom;
  mol.M = (int)pv.M;
  mol.iexch = (int)pv.iexch;
  mol.gga = pv.iexch > 3;
  mol.open = pv.open;
  mol.ns = (int)pv.nshell[0];
  mol.np = (int)pv.nshell[1] / 3;
  mol.nd = (int)pv.nshell[2] / 6;
  if (mol.ns + 3 * mol.np + 6 * mol.nd != mol.M)
    throw std::runtime_error("g2g: nshell does not add up to M");
  const unsigned ld = pv.max_atoms;
  mol.x.resize(mol.natom);
  mol.y.resize(mol.natom);
  mol.z.resize(mol.natom);
  mol.Z.resize(mol.natom);
  mol.rm.resize(mol.natom);
  mol.shells.resize(mol.natom);
  const unsigned* nr = grid_type == 0 ? pv.Nr : pv.Nr2;
  for (int i = 0; i < mol.natom; ++i) {
    mol.x[i] = pv.r[i];
    mol.y[i] = pv.r[ld + i];
    mol.z[i] = pv.r[2 * ld + i];
    mol.Z[i] = (int)pv.Iz[i];
    mol.rm[i] = pv.Rm[pv.Iz[i]];
    mol.shells[i] = (int)nr[pv.Iz[i]];
  }
  const double nf = pv.normalize ? 1.0 / std::sqrt(3.0) : 1.0;
  mol.fn.resize(mol.M);
  const unsigned lda = pv.ngaussians;
  for (int f = 0; f < mol.M; ++f) {
    BasisFn& b = mol.fn[f];
    b.atom = (int)pv.Nuc[f] - 1;
    if (b.atom < 0 || b.atom >= mol.natom) throw std::runtime_error("g2g: Nuc out of range");
    b.ncont = (int)pv.ncont[f];
    if (b.ncont < 1 || b.ncont > kMaxContractions)
      throw std::runtime_error("Maximum functions per contraction reached!");
    int comp = 0;
    if (f < mol.ns) {
      b.shell_len = 1;
      b.lx = b.ly = b.lz = 0;
    } else if (f < mol.ns + 3 * mol.np) {
      b.shell_len = 3;
      comp = (f - mol.ns) % 3;
      b.lx = kP[comp][0];
      b.ly = kP[comp][1];
      b.lz = kP[comp][2];
    } else {
      b.shell_len = 6;
      comp = (f - mol.ns - 3 * mol.np) % 6;
      b.lx = kD[comp][0];
      b.ly = kD[comp][1];
      b.lz = kD[comp][2];
    }
    const double scale = (b.shell_len == 6 && (b.lx == 2 || b.ly == 2 || b.lz == 2)) ? nf : 1.0;
    for (int k = 0; k < kMaxContractions; ++k) {
      b.a[k] = k < b.ncont ? pv.a[(size_t)k * lda + f] : 0.0;
      b.c[k] = k < b.ncont ? pv.c[(size_t)k * lda + f] * scale : 0.0;
    }
  }
  static const int kSizes[3] = {50, 116, 194};
  mol.ngrid = kSizes[grid_type];
  const double* e = pv.e[grid_type];
  const double* w = pv.wang[grid_type];
  mol.ex.assign(e, e + mol.ngrid);
  mol.ey.assign(e + mol.ngrid, e + 2 * mol.ngrid);
  mol.ez.assign(e + 2 * mol.ngrid, e + 3 * mol.ngrid);
  mol.ew.assign(w, w + mol.ngrid);
  return mol;
}

GridFrame grid_frame(const Molecule& mol, const Options& opt) {
  const int na = mol.natom, M = mol.M;
  if (na == 0) throw std::runtime_error("g2g: no atoms");
  GridFrame fr;
  // most diffuse exponent per element and per function (regenerate:165-186)
  std::vector<double> amin_z(130, std::numeric_limits<double>::max());
  std::vector<double> amin_f(M, std::numeric_limits<double>::max());
  for (int f = 0; f < M; ++f) {
    const BasisFn& b = mol.fn[f];
    const int z = mol.Z[b.atom];
    for (int k = 0; k < b.ncont; ++k) {
      amin_z[z] = std::min(amin_z[z], b.a[k]);
      if (b.a[k] < amin_f[f]) amin_f[f] = b.a[k];
    }
  }
  // prism around all atoms (regenerate:189-219)
  for (int a = 0; a < na; ++a) {
    const double reach = 2 * std::sqrt(opt.max_function_exponent / amin_z[mol.Z[a]]);
    const double p[3] = {mol.x[a], mol.y[a], mol.z[a]};
    for (int k = 0; k < 3; ++k) {
      if (a == 0) {
        fr.lo[k] = p[k] - reach;
        fr.hi[k] = p[k] + reach;
      } else {
        fr.lo[k] = std::min(fr.lo[k], p[k] - reach);
        fr.hi[k] = std::max(fr.hi[k], p[k] + reach);
      }
    }
  }
  const double lcs = opt.little_cube_size;
  for (int k = 0; k < 3; ++k) fr.dim[k] = (unsigned)std::ceil((fr.hi[k] - fr.lo[k]) / lcs);
  fr.ncubes = (size_t)fr.dim[0] * fr.dim[1] * fr.dim[2];

  fr.srad.assign(na, 0.0);
  fr.sphere_from.assign(na, 0);
  for (int a = 0; a < na; ++a) {
    const unsigned n = (unsigned)mol.shells[a];
    const unsigned inc = (unsigned)std::ceil(opt.sphere_radius * n);
    fr.sphere_from[a] = opt.sphere_radius > 0 ? n - inc : n;
    if (opt.sphere_radius > 0 && inc > 0) {
      const double x = std::cos((M_PI / (n + 1)) * (n - inc + 1));
      fr.srad[a] = mol.rm[a] * (1.0 + x) / (1.0 - x);
    }
  }
  fr.rij.resize((size_t)na * na);
  for (int i = 0; i < na; ++i)
    for (int j = 0; j < na; ++j) {
      const double dx = mol.x[i] - mol.x[j], dy = mol.y[i] - mol.y[j], dz = mol.z[i] - mol.z[j];
      fr.rij[(size_t)i * na + j] = std::sqrt(dx * dx + dy * dy + dz * dz);
    }
  // radial Gauss-Chebyshev (second kind) nodes of every atom (regenerate:274-300)
  fr.rad_off.resize(na + 1);
  for (int a = 0; a < na; ++a) {
    fr.rad_off[a] = (int)fr.r1.size();
    const int n = mol.shells[a];
    const double t0 = M_PI / (n + 1), rm = mol.rm[a];
    for (int sh = 0; sh < n; ++sh) {
      const double t1 = t0 * (sh + 1);
      const double x = std::cos(t1);
      const double wr = t0 * std::fabs(std::sin(t1));
      const double r1 = rm * (1.0 + x) / (1.0 - x);
      fr.r1.push_back(r1);
      fr.wrad.push_back(wr * (r1 * r1) * rm * 2.0 / ((1.0 - x) * (1.0 - x)));
    }
  }
  fr.rad_off[na] = (int)fr.r1.size();
  // significance test of every basis shell (partition.cpp:229-254)
  const int pend = mol.ns + 3 * mol.np, dend = pend + 6 * mol.nd;
  auto shells = [&](int b, int e, int len) {
    for (int f = b; f < e; f += len) {
      const double a = amin_f[f];
      fr.sh_first.push_back(f);
      fr.sh_len.push_back(len);
      fr.sh_atom.push_back(mol.fn[f].atom);
      fr.sh_a.push_back(a);
      fr.sh_thr.push_back(shell_radius(len, a, opt.max_function_exponent));
    }
  };
  shells(0, mol.ns, 1);
  shells(mol.ns, pend, 3);
  shells(pend, dend, 6);
  return fr;
}

StagedPartition stage_partition(const Molecule& mol, const Options& opt) {
  const double t_start = wall();
  const int na = mol.natom;
  const GridFrame fr = grid_frame(mol, opt);
  const double* lo = fr.lo;
  const double* hi = fr.hi;
  const unsigned* dim = fr.dim;
  const double lcs = opt.little_cube_size;
  const std::vector<double>& rij = fr.rij;
  const std::vector<double>& srad = fr.srad;

  // points of every atom shell and angular node (regenerate:274-322)
  std::vector<Staging> cubes(fr.ncubes), spheres(na);
  for (int a = 0; a < na; ++a) {
    const int n = mol.shells[a];
    for (int sh = 0; sh < n; ++sh) {
      const double r1 = fr.r1[fr.rad_off[a] + sh], wrad = fr.wrad[fr.rad_off[a] + sh];
      const bool in_sphere = (unsigned)sh >= fr.sphere_from[a];
      for (int k = 0; k < mol.ngrid; ++k) {
        const double q[3] = {mol.x[a] + mol.ex[k] * r1, mol.y[a] + mol.ey[k] * r1,
                             mol.z[a] + mol.ez[k] * r1};
        if (!(lo[0] <= q[0] && q[0] <= hi[0] && lo[1] <= q[1] && q[1] <= hi[1] &&
              lo[2] <= q[2] && q[2] <= hi[2]))
          continue;
        Staging* st;
        if (in_sphere) {
          st = &spheres[a];
        } else {
          unsigned c[3];
          for (int d = 0; d < 3; ++d) c[d] = (unsigned)std::floor((q[d] - lo[d]) / lcs);
          if (c[0] >= dim[0] || c[1] >= dim[1] || c[2] >= dim[2])
            throw std::runtime_error("grid point outside the cube lattice of the bounding prism");
          st = &cubes[((size_t)c[0] * dim[1] + c[1]) * dim[2] + c[2]];
        }
        st->x.push_back(q[0]);
        st->y.push_back(q[1]);
        st->z.push_back(q[2]);
        st->w.push_back(wrad * mol.ew[k]);
        st->atom.push_back(a);
      }
    }
  }

  StagedPartition out;
  out.rij = rij;
  std::vector<double> d2(na);
  auto emit = [&](int kind, Staging& st, const Selected& s) {
    GridPartition& o = out.groups;
    o.ngroups++;
    o.kind.push_back(kind);
    o.npts.push_back((int)st.x.size());
    o.pt_off.push_back((int)o.x.size());
    o.x.insert(o.x.end(), st.x.begin(), st.x.end());
    o.y.insert(o.y.end(), st.y.begin(), st.y.end());
    o.z.insert(o.z.end(), st.z.begin(), st.z.end());
    o.w.insert(o.w.end(), st.w.begin(), st.w.end());
    out.owner.insert(out.owner.end(), st.atom.begin(), st.atom.end());
    o.fn_off.push_back((int)o.fn.size());
    int m = 0;
    for (size_t k = 0; k < s.first.size(); ++k)
      for (int c = 0; c < s.len[k]; ++c, ++m) o.fn.push_back(s.first[k] + c);
    o.m.push_back(m);
    o.nshells.push_back((int)s.first.size());
    o.ns.push_back(s.ns);
    o.np.push_back(s.np);
    o.nd.push_back(s.nd);
    o.nnuc.push_back((int)s.atoms.size());
    o.fhash.push_back(fnv1a(s.first));
    out.atom_off.push_back((int)out.atoms.size());
    out.atoms.insert(out.atoms.end(), s.atoms.begin(), s.atoms.end());
  };
  for (unsigned i = 0; i < dim[0]; ++i)
    for (unsigned j = 0; j < dim[1]; ++j)
      for (unsigned k = 0; k < dim[2]; ++k) {
        Staging& st = cubes[((size_t)i * dim[1] + j) * dim[2] + k];
        const double corner[3] = {lo[0] + i * lcs, lo[1] + j * lcs, lo[2] + k * lcs};
        for (int a = 0; a < na; ++a) {
          const double p[3] = {mol.x[a], mol.y[a], mol.z[a]};
          double v[3];
          for (int d = 0; d < 3; ++d) {
            if (p[d] < corner[d]) v[d] = corner[d] - p[d];
            else if (p[d] > (corner[d] + lcs)) v[d] = p[d] - (corner[d] + lcs);
            else v[d] = 0;
          }
          d2[a] = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
        }
        Selected s;
        select(mol, opt, fr, d2, s);
        if (s.first.empty() || st.x.size() < opt.min_points_per_cube) continue;
        emit(0, st, s);
      }
  if (opt.sphere_radius > 0) {
    for (int a = 0; a < na; ++a) {
      Staging& st = spheres[a];
      for (int b = 0; b < na; ++b) {
        if (b == a) {
          d2[b] = 0;
        } else {
          const double dab = rij[(size_t)b * na + a];
          const double gap = (srad[a] > dab ? 0 : dab - srad[a]);
          d2[b] = gap * gap;
        }
      }
      Selected s;
      select(mol, opt, fr, d2, s);
      if (st.x.size() < opt.min_points_per_cube) continue;
      emit(1, st, s);
    }
  }
  out.atom_off.push_back((int)out.atoms.size());
  out.groups.seconds = wall() - t_start;
  return out;
}

GridPartition finish_partition(const Molecule& mol, const Options& opt,
                               const StagedPartition& sp) {
  const double t_start = wall();
  const GridPartition& in = sp.groups;
  GridPartition out;
  for (int g = 0; g < in.ngroups; ++g) {
    Staging st;
    const int b = in.pt_off[g], n = in.npts[g];
    st.x.assign(in.x.begin() + b, in.x.begin() + b + n);
    st.y.assign(in.y.begin() + b, in.y.begin() + b + n);
    st.z.assign(in.z.begin() + b, in.z.begin() + b + n);
    st.w.assign(in.w.begin() + b, in.w.begin() + b + n);
    st.atom.assign(sp.owner.begin() + b, sp.owner.begin() + b + n);
    const std::vector<int> atoms(sp.atoms.begin() + sp.atom_off[g],
                                 sp.atoms.begin() + sp.atom_off[g + 1]);
    weigh(mol, opt, sp.rij, atoms, st);
    if (st.x.size() < opt.min_points_per_cube) continue;
    out.ngroups++;
    out.kind.push_back(in.kind[g]);
    out.npts.push_back((int)st.x.size());
    out.pt_off.push_back((int)out.x.size());
    out.x.insert(out.x.end(), st.x.begin(), st.x.end());
    out.y.insert(out.y.end(), st.y.begin(), st.y.end());
    out.z.insert(out.z.end(), st.z.begin(), st.z.end());
    out.w.insert(out.w.end(), st.w.begin(), st.w.end());
    out.fn_off.push_back((int)out.fn.size());
    out.fn.insert(out.fn.end(), in.fn.begin() + in.fn_off[g],
                  in.fn.begin() + in.fn_off[g] + in.m[g]);
    out.m.push_back(in.m[g]);
    out.nshells.push_back(in.nshells[g]);
    out.ns.push_back(in.ns[g]);
    out.np.push_back(in.np[g]);
    out.nd.push_back(in.nd[g]);
    out.nnuc.push_back(in.nnuc[g]);
    out.fhash.push_back(in.fhash[g]);
  }
  out.seconds = in.seconds + (wall() - t_start);
  return out;
}

GridPartition build_partition(const Molecule& mol, const Options& opt) {
  return finish_partition(mol, opt, stage_partition(mol, opt));
}

}  // namespace xcq
