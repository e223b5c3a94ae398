// C entry points: the handle API (xcq_*) and the g2g Fortran ABI on top of a
// default context.  The g2g semantics follow g2g/init.cpp:38-433 and
// g2g/partition.cpp:426-591 (energy assigned, Fock accumulated, NaN energy
// terminates).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <memory>
#include <stdexcept>
#include <string>

#include "../../include/g2g.h"
#include "../../include/xcq.h"
#include "xcq.h"

namespace xcq {
// Named Fortran timers (timers.cpp) need the options' timer level.
void timers_configure(unsigned level);
}

struct xcq_context {
  std::unique_ptr<xcq::Engine> engine;
  xcq::ProblemView pv;
  xcq::Options opt;
  bool have_problem = false;
  int grid_type = -1;
  xcq::Molecule mol;
  xcq::GridPartition part;
  double upload_seconds = 0.0;
  long long npts_pad = 0;
  double flops = 0.0;
  int density_tiles = 0, fock_items = 0;
  xcq::Precision prec = xcq::Precision::FP64;
  int grid_mode = XCQ_GRID_HOST;  // where the grid and partition are built
  bool points_on_host = true;
};

namespace {

thread_local std::string g_error;

double wall() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const std::exception& e) {
    g_error = e.what();
    return -1;
  } catch (...) {
    g_error = "unknown error";
    return -2;
  }
}

void view_from_problem(const xcq_problem* p, xcq::ProblemView& pv, xcq::Options& opt) {
  pv.natom = p->natom;
  pv.open = p->open != 0;
  pv.max_atoms = p->max_atoms ? p->max_atoms : p->natom;
  pv.M = p->M;
  pv.ngaussians = p->ngaussians ? p->ngaussians : p->M;
  pv.normalize = p->normalize != 0;
  pv.iexch = p->iexch;
  pv.r = p->r;
  pv.Rm = p->Rm;
  pv.Iz = p->Iz;
  pv.Nr = p->Nr;
  pv.Nr2 = p->Nr2;
  pv.Nuc = p->Nuc;
  pv.ncont = p->ncont;
  pv.nshell = p->nshell;
  pv.c = p->c;
  pv.a = p->a;
  for (int k = 0; k < 3; ++k) {
    pv.e[k] = p->e[k];
    pv.wang[k] = p->wang[k];
  }
  opt.little_cube_size = p->little_cube_size;
  opt.sphere_radius = p->sphere_radius;
  opt.assign_all_functions = p->assign_all_functions != 0;
  opt.remove_zero_weights = p->remove_zero_weights != 0;
  opt.min_points_per_cube = p->min_points_per_cube;
  opt.max_function_exponent = p->max_function_exponent;
}

void fill_groups(const xcq::GridPartition& part, xcq_group_info* out, int max) {
  for (int g = 0; g < part.ngroups && g < max; ++g) {
    xcq_group_info& o = out[g];
    o.kind = part.kind[g];
    o.npoints = part.npts[g];
    o.nfuncs = part.m[g];
    o.s = part.ns[g];
    o.p = part.np[g];
    o.d = part.nd[g];
    o.nnuc = part.nnuc[g];
    o.wsum = 0.0;
    o.xsum = 0.0;
    for (int k = 0; k < part.npts[g]; ++k) {
      o.wsum += part.w[part.pt_off[g] + k];
      o.xsum += part.x[part.pt_off[g] + k];
    }
    o.fhash = part.fhash[g];
  }
}

void check_problem(const xcq::ProblemView& pv) {
  if (!pv.r || !pv.Rm || !pv.Iz || !pv.Nuc || !pv.ncont || !pv.nshell || !pv.a || !pv.c)
    throw std::runtime_error("g2g: problem arrays missing");
  if (pv.natom == 0 || pv.M == 0) throw std::runtime_error("g2g: empty problem");
  if (pv.open && pv.iexch <= 3)
    throw std::runtime_error("g2g: open-shell LDA is not provided (LIO's CPU path evaluates no "
                             "functional there and reads unset factors)");
}

void build(xcq_context* ctx, unsigned grid_type) {
  if (!ctx->have_problem) throw std::runtime_error("g2g: grid requested before parameter init");
  check_problem(ctx->pv);
  ctx->mol = xcq::capture(ctx->pv, grid_type);
  if (ctx->grid_mode == XCQ_GRID_DEVICE) {
    const double t0 = wall();
    ctx->part = ctx->engine->load_generated({&ctx->mol}, ctx->opt, ctx->prec)[0];
    ctx->upload_seconds = wall() - t0 - ctx->part.seconds;
    ctx->points_on_host = false;
  } else if (ctx->grid_mode == XCQ_GRID_STAGED) {
    const xcq::StagedPartition st = xcq::stage_partition(ctx->mol, ctx->opt);
    const double t0 = wall();
    ctx->part = ctx->engine->load_staged({&ctx->mol}, {&st}, ctx->opt, ctx->prec)[0];
    ctx->upload_seconds = wall() - t0;
    ctx->points_on_host = false;
  } else {
    ctx->part = xcq::build_partition(ctx->mol, ctx->opt);
    const double t0 = wall();
    ctx->engine->load(ctx->mol, ctx->part, ctx->prec);
    ctx->upload_seconds = wall() - t0;
    ctx->points_on_host = true;
  }
  ctx->grid_type = (int)grid_type;
}

// Device-built grids keep their points on the device: copy them into the
// host partition (unpadded) when a caller asks for per-group sums.
void ensure_points(xcq_context* ctx) {
  if (ctx->points_on_host) return;
  xcq::GridPartition& part = ctx->part;
  const long long np = ctx->engine->padded_points();
  std::vector<double> x(np), y(np), z(np), w(np);
  ctx->engine->download_points(x.data(), y.data(), z.data(), w.data());
  part.x.clear();
  part.y.clear();
  part.z.clear();
  part.w.clear();
  long long off = 0;
  for (int g = 0; g < part.ngroups; ++g) {
    const int n = part.npts[g];
    part.x.insert(part.x.end(), x.begin() + off, x.begin() + off + n);
    part.y.insert(part.y.end(), y.begin() + off, y.begin() + off + n);
    part.z.insert(part.z.end(), z.begin() + off, z.begin() + off + n);
    part.w.insert(part.w.end(), w.begin() + off, w.begin() + off + n);
    off += (n + 15) / 16 * 16;
  }
  ctx->points_on_host = true;
}

}  // namespace

extern "C" {

const char* xcq_last_error(void) { return g_error.c_str(); }

int xcq_create(int device, xcq_context** out) {
  return guarded([&] {
    auto* ctx = new xcq_context;
    try {
      ctx->engine.reset(new xcq::Engine(device));
    } catch (...) {
      delete ctx;
      throw;
    }
    *out = ctx;
  });
}

int xcq_destroy(xcq_context* ctx) {
  return guarded([&] { delete ctx; });
}

int xcq_set_problem(xcq_context* ctx, const xcq_problem* p) {
  return guarded([&] {
    view_from_problem(p, ctx->pv, ctx->opt);
    ctx->have_problem = true;
    ctx->grid_type = -1;
  });
}

int xcq_set_precision(xcq_context* ctx, int precision) {
  return guarded([&] {
    if (precision != XCQ_FP64 && precision != XCQ_FP32 && precision != XCQ_BF16X6)
      throw std::runtime_error("g2g: unknown precision mode");
    ctx->prec = static_cast<xcq::Precision>(precision);
    if (ctx->grid_type >= 0) {
      if (!ctx->points_on_host) {  // device-built grid: rebuild it in the new mode
        build(ctx, (unsigned)ctx->grid_type);
        return;
      }
      const double t0 = wall();
      ctx->engine->load(ctx->mol, ctx->part, ctx->prec);
      ctx->upload_seconds = wall() - t0;
    }
  });
}

int xcq_build_grid(xcq_context* ctx, unsigned grid_type) {
  return guarded([&] { build(ctx, grid_type); });
}

int xcq_release_device(xcq_context* ctx) {
  return guarded([&] {
    ctx->engine->unload();
    ctx->grid_type = -1;
  });
}

int xcq_set_grid_mode(xcq_context* ctx, int mode) {
  return guarded([&] {
    if (mode != XCQ_GRID_HOST && mode != XCQ_GRID_DEVICE && mode != XCQ_GRID_STAGED)
      throw std::runtime_error("g2g: unknown grid mode");
    ctx->grid_mode = mode;
  });
}

int xcq_query_grid(xcq_context* ctx, xcq_grid_info* o) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: no grid built");
    const xcq::GridPartition& part = ctx->part;
    o->ngroups = part.ngroups;
    o->npoints = part.total_points();
    long long pad = 0;
    double fl = 0.0, by = 0.0;
    for (int g = 0; g < part.ngroups; ++g) {
      pad += (part.npts[g] + 15) / 16 * 16;
      fl += 10.0 * part.npts[g] * (double)part.m[g] * part.m[g];
      by += 2.0 * 10.0 * part.npts[g] * (double)part.m[g] * 8.0;
    }
    o->npoints_padded = pad;
    o->flops = fl;
    o->bytes_values = by;
    o->grid_seconds = part.seconds;
    o->upload_seconds = ctx->upload_seconds;
    o->eval_ms = ctx->engine->eval_ms();
    o->arena_bytes = ctx->engine->arena_bytes();
    o->density_tiles = (int)(pad / 16);
    o->fock_items = 0;
    ctx->engine->load_phases(o->load_ms);
  });
}

int xcq_groups(xcq_context* ctx, xcq_group_info* out, int max, int* n) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: no grid built");
    if (out) ensure_points(ctx);
    if (out) fill_groups(ctx->part, out, max);
    if (n) *n = ctx->part.ngroups;
  });
}

int xcq_solve(xcq_context* ctx, const double* rho, double* fock_accum, xcq_solve_info* info) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: solve before the grid was built");
    xcq::SolveResult r = ctx->engine->solve(rho, fock_accum, true);
    if (info) {
      info->energy = r.energy;
      info->electrons = r.electrons;
      info->seconds = r.seconds;
      for (int k = 0; k < 4; ++k) info->kernel_ms[k] = r.kernel_ms[k];
    }
  });
}

int xcq_solve_forces(xcq_context* ctx, const double* rho, double* fock_accum,
                     double* forces_accum, int ld, xcq_solve_info* info) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: solve before the grid was built");
    if (!forces_accum) throw std::runtime_error("g2g: no force array");
    xcq::SolveResult r = ctx->engine->solve(rho, fock_accum, true, forces_accum, ld);
    if (info) {
      info->energy = r.energy;
      info->electrons = r.electrons;
      info->seconds = r.seconds;
      for (int k = 0; k < 4; ++k) info->kernel_ms[k] = r.kernel_ms[k];
    }
  });
}

int xcq_solve_open(xcq_context* ctx, const double* rho_a, const double* rho_b, double* fock_a,
                   double* fock_b, double* forces_accum, int ld, xcq_solve_info* info) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: solve before the grid was built");
    if (!ctx->engine->open_shell()) throw std::runtime_error("g2g: the problem is closed shell");
    xcq::SolveResult r =
        ctx->engine->solve_open(rho_a, rho_b, fock_a, fock_b, true, forces_accum, ld);
    if (info) {
      info->energy = r.energy;
      info->electrons = r.electrons;
      info->seconds = r.seconds;
      for (int k = 0; k < 4; ++k) info->kernel_ms[k] = r.kernel_ms[k];
    }
  });
}

int xcq_reserve_slab(int device, unsigned long long bytes) {
  return guarded([&] {
    xcq::Engine::reserve_slab(device, (size_t)bytes);
  });
}

int xcq_solve_device(xcq_context* ctx, const double* d_rho, double* d_fock) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: solve before the grid was built");
    ctx->engine->solve_device(d_rho, d_fock, true);
  });
}

int xcq_solve_async(xcq_context* ctx, const double* rho) {
  return guarded([&] {
    if (ctx->grid_type < 0) throw std::runtime_error("g2g: solve before the grid was built");
    ctx->engine->solve_async(rho);
  });
}

int xcq_query(xcq_context* ctx) {
  int r = -1;
  const int rc = guarded([&] { r = ctx->engine->idle() ? 1 : 0; });
  return rc != 0 ? rc : r;
}

int xcq_collect(xcq_context* ctx, double* fock_accum, xcq_solve_info* info) {
  return guarded([&] {
    xcq::SolveResult r = ctx->engine->collect(fock_accum);
    if (info) {
      info->energy = r.energy;
      info->electrons = r.electrons;
      info->seconds = 0.0;
      for (int k = 0; k < 4; ++k) info->kernel_ms[k] = 0.0;
    }
  });
}

int xcq_sync_energy(xcq_context* ctx, double* energy, double* electrons) {
  return guarded([&] { ctx->engine->read_energy(energy, electrons); });
}

int xcq_kernel_times(xcq_context* ctx, double ms[4]) {
  return guarded([&] { ctx->engine->kernel_times(ms); });
}

void* xcq_stream(xcq_context* ctx) { return ctx ? ctx->engine->stream() : nullptr; }

int xcq_host_partition(const xcq_problem* p, unsigned grid_type, xcq_group_info* out, int max,
                       int* n) {
  return guarded([&] {
    xcq::ProblemView pv;
    xcq::Options opt;
    view_from_problem(p, pv, opt);
    check_problem(pv);
    xcq::Molecule mol = xcq::capture(pv, grid_type);
    xcq::GridPartition part = xcq::build_partition(mol, opt);
    if (out) fill_groups(part, out, max);
    if (n) *n = part.ngroups;
  });
}

}  // extern "C"

// ---------------------------------------------------------------------------
// batches
// ---------------------------------------------------------------------------

struct xcq_batch {
  std::unique_ptr<xcq::Engine> engine;
  std::vector<xcq::Molecule> mols;
  std::vector<xcq::GridPartition> parts;
  std::vector<xcq::StagedPartition> staged;  // staged-grid mode
  xcq::Options opt;                           // device and staged grid modes
  int grid_mode = XCQ_GRID_HOST;
  xcq::Precision prec = xcq::Precision::FP64;
};

namespace {

void load_batch(xcq_batch* b) {
  std::vector<const xcq::Molecule*> m;
  for (size_t k = 0; k < b->mols.size(); ++k) m.push_back(&b->mols[k]);
  if (b->grid_mode == XCQ_GRID_DEVICE) {
    b->parts = b->engine->load_generated(m, b->opt, b->prec);
    return;
  }
  if (b->grid_mode == XCQ_GRID_STAGED) {
    std::vector<const xcq::StagedPartition*> st;
    for (size_t k = 0; k < b->staged.size(); ++k) st.push_back(&b->staged[k]);
    b->parts = b->engine->load_staged(m, st, b->opt, b->prec);
    return;
  }
  std::vector<const xcq::GridPartition*> p;
  for (size_t k = 0; k < b->parts.size(); ++k) p.push_back(&b->parts[k]);
  b->engine->load_batch(m, p, b->prec);
}

}  // namespace

extern "C" {

int xcq_batch_create(int device, int n, const xcq_problem* problems, const unsigned* grid_types,
                     int precision, xcq_batch** out) {
  return xcq_batch_create_ex(device, n, problems, grid_types, precision, XCQ_GRID_HOST, out);
}

int xcq_batch_create_ex(int device, int n, const xcq_problem* problems,
                        const unsigned* grid_types, int precision, int grid_mode,
                        xcq_batch** out) {
  return guarded([&] {
    if (grid_mode != XCQ_GRID_HOST && grid_mode != XCQ_GRID_DEVICE &&
        grid_mode != XCQ_GRID_STAGED)
      throw std::runtime_error("g2g: unknown grid mode");
    if (n <= 0 || !problems || !grid_types) throw std::runtime_error("g2g: empty batch");
    if (precision != XCQ_FP64 && precision != XCQ_FP32 && precision != XCQ_BF16X6)
      throw std::runtime_error("g2g: unknown precision mode");
    std::unique_ptr<xcq_batch> b(new xcq_batch);
    b->prec = static_cast<xcq::Precision>(precision);
    b->engine.reset(new xcq::Engine(device));
    b->mols.resize(n);
    b->parts.resize(n);
    b->grid_mode = grid_mode;
    std::vector<xcq::StagedPartition>& staged = b->staged;
    staged.resize(grid_mode == XCQ_GRID_STAGED ? n : 0);
    xcq::Options& opt0 = b->opt;
    for (int k = 0; k < n; ++k) {
      xcq::ProblemView pv;
      xcq::Options opt;
      view_from_problem(&problems[k], pv, opt);
      check_problem(pv);
      if (k == 0) opt0 = opt;
      b->mols[k] = xcq::capture(pv, grid_types[k]);
      if (grid_mode == XCQ_GRID_STAGED) {
        if (opt.remove_zero_weights != opt0.remove_zero_weights ||
            opt.min_points_per_cube != opt0.min_points_per_cube)
          throw std::runtime_error("g2g: a staged-grid batch needs one weight-removal setting");
        staged[k] = xcq::stage_partition(b->mols[k], opt);
      } else if (grid_mode == XCQ_GRID_DEVICE) {
        if (opt.remove_zero_weights != opt0.remove_zero_weights ||
            opt.min_points_per_cube != opt0.min_points_per_cube ||
            opt.little_cube_size != opt0.little_cube_size ||
            opt.sphere_radius != opt0.sphere_radius ||
            opt.assign_all_functions != opt0.assign_all_functions ||
            opt.max_function_exponent != opt0.max_function_exponent)
          throw std::runtime_error("g2g: a device-grid batch needs one set of grid options");
      } else {
        b->parts[k] = xcq::build_partition(b->mols[k], opt);
      }
    }
    load_batch(b.get());
    *out = b.release();
  });
}

int xcq_batch_destroy(xcq_batch* b) {
  return guarded([&] { delete b; });
}

int xcq_batch_offsets(xcq_batch* b, long long* offsets) {
  return guarded([&] {
    const int n = b->engine->molecules();
    for (int k = 0; k <= n; ++k) offsets[k] = b->engine->packed_offset(k);
  });
}

int xcq_batch_solve_device(xcq_batch* b, const double* d_rho, double* d_fock) {
  return guarded([&] { b->engine->solve_device(d_rho, d_fock, true); });
}

int xcq_batch_solve_forces(xcq_batch* b, const double* d_rho, double* d_fock, double* d_forces) {
  return guarded([&] {
    if (!d_forces) throw std::runtime_error("g2g: no force array");
    b->engine->solve_device(d_rho, d_fock, true, d_forces);
  });
}

int xcq_batch_solve_forces_static(xcq_batch* b, const double* d_rho, double* d_fock,
                                  double* d_forces) {
  return guarded([&] {
    if (!d_forces) throw std::runtime_error("g2g: no force array");
    b->engine->solve_device_static(d_rho, d_fock, d_forces);
  });
}

long long xcq_static_scratch_bytes(void) { return (long long)xcq::Engine::static_scratch_bytes(); }

int xcq_batch_atoms(xcq_batch* b, int* first_atom) {
  return guarded([&] {
    const int n = b->engine->molecules();
    for (int k = 0; k <= n; ++k) first_atom[k] = b->engine->atoms(k);
  });
}

int xcq_batch_capture(xcq_batch* b, const double* d_rho, double* d_fock) {
  return guarded([&] { b->engine->capture_graph(d_rho, d_fock); });
}

int xcq_batch_replay(xcq_batch* b) {
  return guarded([&] { b->engine->replay_graph(); });
}

int xcq_batch_energies(xcq_batch* b, double* energy, double* electrons) {
  return guarded([&] {
    const int n = b->engine->molecules();
    std::vector<double> en(2 * n);
    b->engine->read_energies(en.data());
    for (int k = 0; k < n; ++k) {
      if (energy) energy[k] = en[2 * k];
      if (electrons) electrons[k] = en[2 * k + 1];
    }
  });
}

int xcq_batch_kernel_times(xcq_batch* b, double ms[4]) {
  return guarded([&] { b->engine->kernel_times(ms); });
}

void* xcq_batch_stream(xcq_batch* b) { return b ? b->engine->stream() : nullptr; }

long long xcq_batch_arena_bytes(xcq_batch* b) {
  return b ? (long long)b->engine->arena_bytes() : -1;
}

int xcq_batch_load_ms(xcq_batch* b, double ms[4]) {
  return guarded([&] { b->engine->load_phases(ms); });
}

static xcq::Tuning schedule(xcq::Tuning t, int fock_slice, int density_variant) {
  if (fock_slice > 0) t.fock_slice = fock_slice;
  if (density_variant >= 0) t.density_variant = density_variant;
  return t;
}

int xcq_set_schedule(xcq_context* ctx, int fock_slice, int density_variant) {
  return guarded([&] {
    ctx->engine->set_tuning(schedule(ctx->engine->tuning(), fock_slice, density_variant));
  });
}

int xcq_batch_set_schedule(xcq_batch* b, int fock_slice, int density_variant) {
  return guarded([&] {
    b->engine->set_tuning(schedule(b->engine->tuning(), fock_slice, density_variant));
    load_batch(b);
  });
}

int xcq_batch_set_bucket_slices(xcq_batch* b, int nb, const int* m_hi, const int* slices) {
  return guarded([&] {
    if (nb < 0 || (nb > 0 && (!m_hi || !slices))) throw std::runtime_error("g2g: bad buckets");
    xcq::Tuning t = b->engine->tuning();
    t.bucket_mhi.assign(m_hi, m_hi + nb);
    t.bucket_slice.assign(slices, slices + nb);
    b->engine->set_tuning(t);
    load_batch(b);
  });
}

int xcq_batch_group_sizes(xcq_batch* b, int* m_pad, long long* npts_pad, int max, int* ngroups) {
  return guarded([&] {
    int n = 0;
    for (const xcq::GridPartition& p : b->parts) {
      for (int g = 0; g < p.ngroups; ++g, ++n) {
        if (n >= max) continue;
        const int align = b->prec == xcq::Precision::BF16X6 ? 32 : 16;
        if (m_pad) m_pad[n] = (p.m[g] + align - 1) / align * align;
        if (npts_pad) npts_pad[n] = (p.npts[g] + 15) / 16 * 16;
      }
    }
    if (ngroups) *ngroups = n;
  });
}

int xcq_set_persistent(xcq_context* ctx, int on) {
  return guarded([&] {
    xcq::Tuning t = ctx->engine->tuning();
    t.persistent = on ? 1 : 0;
    ctx->engine->set_tuning(t);
  });
}

int xcq_batch_set_persistent(xcq_batch* b, int on) {
  return guarded([&] {
    xcq::Tuning t = b->engine->tuning();
    t.persistent = on ? 1 : 0;
    b->engine->set_tuning(t);
    load_batch(b);
  });
}

int xcq_batch_get_schedule(xcq_batch* b, int* fock_slice, int* density_variant, int* persistent) {
  return guarded([&] {
    const xcq::Tuning& t = b->engine->tuning();
    if (fock_slice) *fock_slice = t.fock_slice;
    if (density_variant) *density_variant = t.density_variant;
    if (persistent) *persistent = t.persistent;
  });
}

int xcq_batch_tune(xcq_batch* b, const double* d_rho, double* d_fock, int reps, const int* slices,
                   const int* variants, int ncand, double* ms_out, int* best) {
  return guarded([&] {
    if (ncand <= 0 || reps <= 0) throw std::runtime_error("g2g: nothing to tune");
    int ibest = -1;
    double tbest = 0.0;
    for (int c = 0; c < ncand; ++c) {
      xcq::Tuning t = b->engine->tuning();
      t.fock_slice = slices[c];
      if (variants) t.density_variant = variants[c];
      b->engine->set_tuning(t);
      load_batch(b);
      b->engine->solve_device(d_rho, d_fock, true);  // warm-up
      double e[2];
      b->engine->read_energy(&e[0], &e[1]);
      const double t0 = wall();
      for (int r = 0; r < reps; ++r) b->engine->solve_device(d_rho, d_fock, true);
      b->engine->read_energy(&e[0], &e[1]);
      const double ms = 1e3 * (wall() - t0) / reps;
      if (ms_out) ms_out[c] = ms;
      if (ibest < 0 || ms < tbest) {
        ibest = c;
        tbest = ms;
      }
    }
    xcq::Tuning t = b->engine->tuning();
    t.fock_slice = slices[ibest];
    if (variants) t.density_variant = variants[ibest];
    b->engine->set_tuning(t);
    load_batch(b);
    if (best) *best = ibest;
  });
}

// ---------------------------------------------------------------------------
// g2g Fortran ABI on the default context
// ---------------------------------------------------------------------------

}  // extern "C"

namespace {

xcq_context* g_ctx = nullptr;
xcq::Options g_opt;  // set by g2g_set_options_, possibly before g2g_init_

[[noreturn]] void die(const char* where, const std::string& msg) {
  std::fprintf(stderr, "g2g: %s failed: %s\n", where, msg.c_str());
  std::fflush(stderr);
  std::exit(1);
}

template <class F>
void abi(const char* where, F&& f) {
  try {
    f();
  } catch (const std::exception& e) {
    die(where, e.what());
  } catch (...) {
    die(where, "unknown error");
  }
}

// LIO_XC_PRECISION = fp64 (default) | fp32 | bf16x6 selects the contraction
// arithmetic of the Fortran-facing singleton.
xcq::Precision precision_from_env() {
  const char* v = std::getenv("LIO_XC_PRECISION");
  if (!v || !*v) return xcq::Precision::FP64;
  const std::string s(v);
  if (s == "fp64") return xcq::Precision::FP64;
  if (s == "fp32") return xcq::Precision::FP32;
  if (s == "bf16x6") return xcq::Precision::BF16X6;
  throw std::runtime_error("LIO_XC_PRECISION must be fp64, fp32 or bf16x6, got '" + s + "'");
}

xcq_context* ctx_or_die(const char* where) {
  if (!g_ctx) die(where, "g2g_init_ was not called");
  return g_ctx;
}

}  // namespace

extern "C" {

void g2g_init_(void) {
  abi("g2g_init_", [] {
    if (!g_ctx) {
      g_ctx = new xcq_context;
      g_ctx->engine.reset(new xcq::Engine(-1));
    }
    g_ctx->prec = precision_from_env();
    if (const char* v = std::getenv("LIO_XC_DEVICE_GRID"))
      g_ctx->grid_mode = v[0] == '1' ? XCQ_GRID_DEVICE : v[0] == '2' ? XCQ_GRID_STAGED
                                                                       : XCQ_GRID_HOST;
    g_ctx->opt = g_opt;
    if (g_opt.verbose > 2) std::printf("  Using HIP device %d for the XC quadrature.\n", g_ctx->engine->device());
  });
}

void g2g_set_options_(double* fgm, double* lcs, double* sr, bool* aaf, bool* eai, bool* rzw,
                      unsigned* mppc, unsigned* mfe, unsigned* tl, unsigned* verbose) {
  g_opt.free_global_memory = *fgm;
  g_opt.little_cube_size = *lcs;
  g_opt.sphere_radius = *sr;
  g_opt.assign_all_functions = *aaf;
  g_opt.energy_all_iterations = *eai;
  g_opt.remove_zero_weights = *rzw;
  g_opt.min_points_per_cube = *mppc;
  g_opt.max_function_exponent = *mfe;
  g_opt.timer_level = *tl;
  g_opt.verbose = *verbose;
  xcq::timers_configure(*tl);
  if (g_ctx) g_ctx->opt = g_opt;
}

void g2g_parameter_init_(const unsigned* norm, const unsigned* natom, const unsigned* max_atoms,
                         const unsigned* ngaussians, double* r, double* Rm, const unsigned* Iz,
                         const unsigned* Nr, const unsigned* Nr2, unsigned* Nuc, const unsigned* M,
                         unsigned* ncont, const unsigned* nshell, double* c, double* a,
                         double* rho_vec, double* fock_vec, double* fockb_vec, double* rhoalpha,
                         double* rhobeta, const unsigned* nco, bool* OPEN, const unsigned* nunp,
                         const unsigned* nopt, const unsigned* Iexch, double* e, double* e2,
                         double* e3, double* wang, double* wang2, double* wang3, bool* use_libxc,
                         const unsigned* ex_functional_id, const unsigned* ec_functional_id) {
  (void)ex_functional_id;
  (void)ec_functional_id;
  abi("g2g_parameter_init_", [&] {
    xcq_context* ctx = ctx_or_die("g2g_parameter_init_");
    if (*Iexch < 1 || *Iexch > 9) throw std::runtime_error("Iexch must be in 1..9");
    if (*use_libxc) throw std::runtime_error("libxc functionals are not provided by this build");
    xcq::ProblemView& pv = ctx->pv;
    pv = xcq::ProblemView{};
    pv.natom = *natom;
    pv.max_atoms = *max_atoms;
    pv.M = *M;
    pv.ngaussians = *ngaussians;
    pv.normalize = *norm != 0;
    pv.iexch = *Iexch;
    pv.open = *OPEN;
    pv.nco = *nco;
    pv.nunp = *nunp;
    pv.nopt = *nopt;
    pv.r = r;
    pv.Rm = Rm;
    pv.Iz = Iz;
    pv.Nr = Nr;
    pv.Nr2 = Nr2;
    pv.Nuc = Nuc;
    pv.ncont = ncont;
    pv.nshell = nshell;
    pv.c = c;
    pv.a = a;
    // open shell reads the two spin densities and writes both Fock matrices
    // (g2g/init.cpp:161-178); rho_vec (the total) is not used by the solve
    pv.rho = *OPEN ? rhoalpha : rho_vec;
    pv.fock = fock_vec;
    pv.rho_b = *OPEN ? rhobeta : nullptr;
    pv.fock_b = *OPEN ? fockb_vec : nullptr;
    if (*OPEN && (!rhoalpha || !rhobeta || !fockb_vec))
      throw std::runtime_error("open shell without alpha/beta densities or beta Fock");
    pv.e[0] = e;
    pv.e[1] = e2;
    pv.e[2] = e3;
    pv.wang[0] = wang;
    pv.wang[1] = wang2;
    pv.wang[2] = wang3;
    ctx->opt = g_opt;
    ctx->have_problem = true;
    ctx->grid_type = -1;
  });
}

void g2g_reload_atom_positions_(const unsigned* grid_type) {
  abi("g2g_reload_atom_positions_", [&] { build(ctx_or_die("g2g_reload_atom_positions_"), *grid_type); });
}

void g2g_new_grid_(const unsigned* grid_type) {
  abi("g2g_new_grid_", [&] {
    xcq_context* ctx = ctx_or_die("g2g_new_grid_");
    if ((int)*grid_type != ctx->grid_type) build(ctx, *grid_type);
  });
}

void g2g_solve_groups_(const unsigned* computation_type, double* energy, double* forces) {
  abi("g2g_solve_groups_", [&] {
    xcq_context* ctx = ctx_or_die("g2g_solve_groups_");
    const unsigned ct = *computation_type;
    if (ct > 3) throw std::runtime_error("unknown computation type");
    if (ctx->grid_type < 0) throw std::runtime_error("solve before g2g_reload_atom_positions_");
    // 0 Fock, 1 energy, 2 energy + forces, 3 forces (init.cpp:340-351)
    const bool compute_rmm = (ct == 0);
    const bool compute_forces = (ct == 2 || ct == 3);
    const bool compute_energy = (ct == 1 || ct == 2) || g_opt.energy_all_iterations;
    if (compute_forces && !forces) throw std::runtime_error("forces requested without a force array");
    xcq::SolveResult r =
        ctx->pv.open
            ? ctx->engine->solve_open(ctx->pv.rho, ctx->pv.rho_b,
                                      compute_rmm ? ctx->pv.fock : nullptr,
                                      compute_rmm ? ctx->pv.fock_b : nullptr, compute_energy,
                                      compute_forces ? forces : nullptr, (int)ctx->pv.max_atoms)
            : ctx->engine->solve(ctx->pv.rho, compute_rmm ? ctx->pv.fock : nullptr,
                                 compute_energy, compute_forces ? forces : nullptr,
                                 (int)ctx->pv.max_atoms);
    const double e = compute_energy ? r.energy : 0.0;
    *energy = e;
    if (std::isnan(e)) die("g2g_solve_groups_", "XC energy is NaN");
  });
}

void g2g_deinit_(void) {
  abi("g2g_deinit_", [] {
    delete g_ctx;
    g_ctx = nullptr;
  });
}

}  // extern "C"
