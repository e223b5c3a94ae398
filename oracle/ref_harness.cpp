// Harness around LIO's own CPU g2g path (FP64), built from the reference
// sources where they lie (see oracle/Makefile) into oracle/_ref/.
//
// Test infrastructure only: it produces golden vectors and pins the CPU
// oracle.  It drives the reference through its extern "C" ABI exactly as
// lioamber does (g2g_set_options_, g2g_init_, g2g_parameter_init_,
// g2g_reload_atom_positions_, g2g_solve_groups_; g2g/init.cpp:38-433) and
// reads the reference's global partition (init.cpp:27) for the electron count
// and the group inventory, which the ABI does not return.
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "common.h"
#include "init.h"
#include "matrix.h"
#include "partition.h"

#include "xc_input.h"

extern G2G::Partition partition;  // g2g/init.cpp:27

extern "C" {
void g2g_init_(void);
void g2g_deinit_(void);
void g2g_set_options_(double*, double*, double*, bool*, bool*, bool*, unsigned&,
                      unsigned&, unsigned&, unsigned&);
void g2g_parameter_init_(const unsigned&, const unsigned&, const unsigned&,
                         const unsigned&, double*, double*, const unsigned*,
                         const unsigned*, const unsigned*, unsigned*,
                         const unsigned&, unsigned*, const unsigned*, double*,
                         double*, double*, double*, double*, double*, double*,
                         const unsigned&, bool&, const unsigned&, const unsigned&,
                         const unsigned&, double*, double*, double*, double*,
                         double*, double*, bool&, const unsigned&, const unsigned&);
void g2g_reload_atom_positions_(const unsigned&);
void g2g_solve_groups_(const unsigned&, double*, double*);
}

namespace {

struct State {
  // Copies that stay alive while the reference holds raw pointers to them.
  std::vector<double> r, Rm, a, c, rho, fock, fockb, rho_a, rho_b, e1, e2, e3, w1, w2, w3;
  std::vector<unsigned> Iz, Nr, Nr2, Nuc, ncont, nshell;
  unsigned natom = 0, M = 0;
  bool open = false, use_libxc = false;
  bool live = false;
};
State S;

template <class T>
std::vector<T> copy_of(const T* p, size_t n) {
  return p ? std::vector<T>(p, p + n) : std::vector<T>(n, T(0));
}

double now() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

unsigned long long fnv(const std::vector<unsigned>& v) {
  unsigned long long h = 1469598103934665603ull;
  for (unsigned x : v) {
    for (int b = 0; b < 4; ++b) {
      h ^= (x >> (8 * b)) & 0xffu;
      h *= 1099511628211ull;
    }
  }
  return h;
}

}  // namespace

extern "C" {

// Builds the grid/partition; returns seconds spent in the reload.
double ref_setup(const XcInput* in) {
  if (S.live) {
    g2g_deinit_();
    S.live = false;
  }
  S.natom = in->natom;
  S.M = in->M;
  const size_t M = in->M, packed = M * (M + 1) / 2;
  S.r = copy_of(in->r, 3 * S.natom);
  S.Rm = copy_of(in->Rm, 55);
  S.Iz = copy_of(in->Iz, S.natom);
  S.Nr = copy_of(in->Nr, 55);
  S.Nr2 = copy_of(in->Nr2, 55);
  S.Nuc = copy_of(in->Nuc, M);
  S.ncont = copy_of(in->ncont, M);
  S.nshell = copy_of(in->nshell, 3);
  S.a = copy_of(in->a, M * MAX_CONTRACTIONS);
  S.c = copy_of(in->c, M * MAX_CONTRACTIONS);
  S.open = in->open != 0;
  if (S.open) {
    // lioamber passes the total density as rho_vec and the two spins apart
    S.rho_a = copy_of(in->rho, packed);
    S.rho_b = copy_of(in->rho_b, packed);
    S.rho.resize(packed);
    for (size_t k = 0; k < packed; ++k) S.rho[k] = S.rho_a[k] + S.rho_b[k];
  } else {
    S.rho = copy_of(in->rho, packed);
    S.rho_a.clear();
    S.rho_b.clear();
  }
  S.fock.assign(packed, 0.0);
  S.fockb.assign(packed, 0.0);
  S.e1 = copy_of(in->e[0], 50 * 3);
  S.e2 = copy_of(in->e[1], 116 * 3);
  S.e3 = copy_of(in->e[2], 194 * 3);
  S.w1 = copy_of(in->wang[0], 50);
  S.w2 = copy_of(in->wang[1], 116);
  S.w3 = copy_of(in->wang[2], 194);

  double fgm = 0.0, lcs = in->little_cube_size, sr = in->sphere_radius;
  bool aaf = in->assign_all_functions != 0, eai = true,
       rzw = in->remove_zero_weights != 0;
  unsigned mppc = in->min_points_per_cube, mfe = in->max_function_exponent,
           tim = 0, verb = 0;
  g2g_set_options_(&fgm, &lcs, &sr, &aaf, &eai, &rzw, mppc, mfe, tim, verb);
  g2g_init_();
  const unsigned norm = in->normalize ? 1u : 0u, natom = S.natom, Mu = S.M;
  const unsigned nco = 1, nunp = 0, nopt = 0, iexch = in->iexch, exid = 0,
                 ecid = 0;
  g2g_parameter_init_(norm, natom, natom, Mu, S.r.data(), S.Rm.data(),
                      S.Iz.data(), S.Nr.data(), S.Nr2.data(), S.Nuc.data(), Mu,
                      S.ncont.data(), S.nshell.data(), S.c.data(), S.a.data(),
                      S.rho.data(), S.fock.data(), S.fockb.data(),
                      S.open ? S.rho_a.data() : nullptr, S.open ? S.rho_b.data() : nullptr,
                      nco, S.open, nunp, nopt, iexch, S.e1.data(),
                      S.e2.data(), S.e3.data(), S.w1.data(), S.w2.data(),
                      S.w3.data(), S.use_libxc, exid, ecid);
  const unsigned gt = in->grid_type;
  double t0 = now();
  g2g_reload_atom_positions_(gt);
  double t1 = now();
  S.live = true;
  return t1 - t0;
}

// One XC Fock build (computation_type 0 with energy): energy is assigned,
// fock_out receives the packed Fock accumulated onto zeros. Returns seconds.
double ref_solve(const double* rho, double* energy, double* fock_out) {
  if (rho) std::memcpy(S.rho.data(), rho, S.rho.size() * sizeof(double));
  std::fill(S.fock.begin(), S.fock.end(), 0.0);
  const unsigned ct = 0;
  double t0 = now();
  g2g_solve_groups_(ct, energy, nullptr);
  double t1 = now();
  if (fock_out) std::memcpy(fock_out, S.fock.data(), S.fock.size() * sizeof(double));
  return t1 - t0;
}

// computation_type ct (0 Fock, 1 energy, 2 energy + forces, 3 forces) with
// energy_all_iterations set: energy assigned, fock_out receives the packed
// Fock accumulated onto zeros, forces_out the (natom x 3, column-major,
// leading dimension natom) forces accumulated onto zeros.  Returns seconds.
double ref_solve_ct(unsigned ct, const double* rho, double* energy, double* fock_out,
                    double* forces_out) {
  if (rho) std::memcpy(S.rho.data(), rho, S.rho.size() * sizeof(double));
  std::fill(S.fock.begin(), S.fock.end(), 0.0);
  std::vector<double> forces(3 * (size_t)S.natom, 0.0);
  double t0 = now();
  g2g_solve_groups_(ct, energy, forces.data());
  double t1 = now();
  if (fock_out) std::memcpy(fock_out, S.fock.data(), S.fock.size() * sizeof(double));
  if (forces_out) std::memcpy(forces_out, forces.data(), forces.size() * sizeof(double));
  return t1 - t0;
}

// Open shell, computation type ct: the spin densities are copied into the
// arrays the reference reads, fock_a / fock_b receive the two spins' packed
// Fock matrices accumulated onto zeros, forces_out the forces (natom x 3,
// column-major).  Returns seconds, or -1 when the problem is closed shell.
double ref_solve_open(unsigned ct, const double* rho_a, const double* rho_b, double* energy,
                      double* fock_a, double* fock_b, double* forces_out) {
  if (!S.open) return -1.0;
  if (rho_a) std::memcpy(S.rho_a.data(), rho_a, S.rho_a.size() * sizeof(double));
  if (rho_b) std::memcpy(S.rho_b.data(), rho_b, S.rho_b.size() * sizeof(double));
  std::fill(S.fock.begin(), S.fock.end(), 0.0);
  std::fill(S.fockb.begin(), S.fockb.end(), 0.0);
  std::vector<double> forces(3 * (size_t)S.natom, 0.0);
  double t0 = now();
  g2g_solve_groups_(ct, energy, forces.data());
  double t1 = now();
  if (fock_a) std::memcpy(fock_a, S.fock.data(), S.fock.size() * sizeof(double));
  if (fock_b) std::memcpy(fock_b, S.fockb.data(), S.fockb.size() * sizeof(double));
  if (forces_out) std::memcpy(forces_out, forces.data(), forces.size() * sizeof(double));
  return t1 - t0;
}

// Integrated electron count sum_p w_p rho(r_p) over the reference's groups,
// using the function values the CPU path keeps after a solve.
double ref_electrons(void) {
  double n = 0.0;
  auto one = [&](G2G::PointGroup<double>* g0) {
    auto* g = static_cast<G2G::PointGroupCPU<double>*>(g0);
    const unsigned m = g->total_functions();
    g->compute_functions(false, false);
    G2G::HostMatrix<double> P(m, m);
    g->get_rmm_input(P);
    for (unsigned p = 0; p < g->number_of_points; ++p) {
      double rho = 0.0;
      for (unsigned i = 0; i < m; ++i) {
        double w = 0.0;
        for (unsigned j = 0; j <= i; ++j) w += P(i, j) * g->function_values(j, p);
        rho += g->function_values(i, p) * w;
      }
      n += g->points[p].weight * rho;
    }
  };
  for (auto* g : partition.cubes) one(g);
  for (auto* g : partition.spheres) one(g);
  return n;
}

// Group inventory; returns the number of groups (fills at most max entries).
int ref_groups(XcGroupInfo* out, int max) {
  int k = 0;
  auto one = [&](G2G::PointGroup<double>* g, int kind) {
    if (k < max && out) {
      XcGroupInfo& o = out[k];
      o.kind = kind;
      o.npoints = g->number_of_points;
      o.nfuncs = g->total_functions();
      o.s = g->s_functions;
      o.p = g->p_functions;
      o.d = g->d_functions;
      o.nnuc = g->total_nucleii();
      o.wsum = 0.0;
      o.xsum = 0.0;
      for (const auto& pt : g->points) {
        o.wsum += pt.weight;
        o.xsum += pt.position.x;
      }
      o.fhash = fnv(g->local2global_func);
    }
    ++k;
  };
  for (auto* g : partition.cubes) one(g, 0);
  for (auto* g : partition.spheres) one(g, 1);
  return k;
}

void ref_teardown(void) {
  if (S.live) g2g_deinit_();
  S.live = false;
}

}  // extern "C"
