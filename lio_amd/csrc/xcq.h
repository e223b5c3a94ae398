// Internal C++ interface of the MI355X XC quadrature engine.
//
// Layers:
//   g2g_abi.cpp  : the Fortran-facing g2g C ABI (drop-in for libg2g.so)
//   grid.cpp     : host grid + point-group partition (cubes and spheres of
//                  significant basis functions, group-local Becke weights)
//   engine.hip   : device arena, uploads, fused HIP kernels, solve
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace xcq {

constexpr int kMaxContractions = 13;
// Classes of j (j mod kBeckeLanes) the Becke sum sum_j P_j is split into:
// each class is summed in list order, then the classes in order, on the host
// and on the device (one wave per class) alike.
constexpr int kBeckeLanes = 8;

// g2g_set_options_ knobs and their defaults (g2g/init.cpp:394-406).
struct Options {
  double free_global_memory = 0.0;
  double little_cube_size = 8.0;
  double sphere_radius = 0.6;
  bool assign_all_functions = false;
  bool energy_all_iterations = false;
  bool remove_zero_weights = true;
  unsigned min_points_per_cube = 1;
  unsigned max_function_exponent = 10;
  unsigned timer_level = 0;
  unsigned verbose = 0;
};

// The caller's arrays (Fortran layouts), referenced, not owned, exactly as
// g2g_parameter_init_ receives them (g2g/init.cpp:69-219).
struct ProblemView {
  unsigned natom = 0, max_atoms = 0, M = 0, ngaussians = 0;
  bool normalize = true;
  unsigned iexch = 9;
  bool open = false;
  unsigned nco = 0, nunp = 0, nopt = 0;
  const double* r = nullptr;        // natom x 3, ld = max_atoms
  const double* Rm = nullptr;       // by Z
  const unsigned* Iz = nullptr;     // natom
  const unsigned* Nr = nullptr;     // shells by Z (grid type 0)
  const unsigned* Nr2 = nullptr;    // shells by Z (grid types 1, 2)
  const unsigned* Nuc = nullptr;    // M, 1-based
  const unsigned* ncont = nullptr;  // M
  const unsigned* nshell = nullptr; // 3
  const double* c = nullptr;        // M x 13, ld = ngaussians
  const double* a = nullptr;        // M x 13, ld = ngaussians
  const double* rho = nullptr;      // packed density (alpha when open)
  double* fock = nullptr;           // packed Fock, accumulated (+=) (alpha when open)
  const double* rho_b = nullptr;    // open shell: packed beta density
  double* fock_b = nullptr;         // open shell: packed beta Fock (+=)
  const double* e[3] = {nullptr, nullptr, nullptr};
  const double* wang[3] = {nullptr, nullptr, nullptr};
};

// Basis in the engine's form: one record per expanded function.
struct BasisFn {
  double a[kMaxContractions];
  double c[kMaxContractions];  // includes the NORM factor of d xx/yy/zz
  int atom;                    // 0-based
  int ncont;
  int lx, ly, lz;
  int shell_len;               // 1 (s), 3 (p), 6 (d)
};

struct Molecule {
  int natom = 0, M = 0, ns = 0, np = 0, nd = 0;  // np/nd count shells
  int iexch = 9;
  bool gga = true;
  bool open = false;            // open shell (two spin densities)
  std::vector<double> x, y, z;  // atom positions (bohr)
  std::vector<int> Z;
  std::vector<double> rm;       // Becke size parameter per atom
  std::vector<int> shells;      // radial shells per atom
  std::vector<BasisFn> fn;      // M records
  int ngrid = 0;
  std::vector<double> ex, ey, ez, ew;  // angular rule
};

// Builds a Molecule from the caller's arrays for grid type gt (0, 1, 2).
// Throws std::runtime_error on inconsistent input.
Molecule capture(const ProblemView& pv, unsigned grid_type);

// The point-group partition, flattened for upload.  Groups are cubes (in
// prism order) followed by spheres (in atom order); points keep generation
// order inside each group.
struct GridPartition {
  int ngroups = 0;
  std::vector<int> kind;     // 0 cube, 1 sphere
  std::vector<int> npts;     // points after zero-weight removal
  std::vector<int> pt_off;   // offset into x/y/z/w
  std::vector<int> m;        // expanded functions
  std::vector<int> fn_off;   // offset into fn
  std::vector<int> nshells;  // selected shells (s + p + d)
  std::vector<int> ns, np, nd;
  std::vector<int> nnuc;
  std::vector<uint64_t> fhash;  // FNV-1a of the first-component indices
  std::vector<double> x, y, z, w;
  std::vector<int> fn;       // global expanded function indices per group
  double seconds = 0.0;
  long long total_points() const { return (long long)x.size(); }
};

GridPartition build_partition(const Molecule& mol, const Options& opt);

// The partition before the Becke weights: the candidate groups (same order)
// with their significant functions, staged points (radial x angular weight)
// and, per group, its nuclei (atoms[atom_off[g] .. atom_off[g+1]), ascending)
// and per point the atom that generated it.  finish_partition applies the
// group-local Becke weights and zero-weight removal on the host
// (build_partition = finish_partition(stage_partition)); the engine can do
// the same on the device (Engine::load_staged).
struct StagedPartition {
  GridPartition groups;     // points unweighted by Becke
  std::vector<int> owner;   // per staged point
  std::vector<int> atom_off, atoms;
  std::vector<double> rij;  // natom x natom distances
};
StagedPartition stage_partition(const Molecule& mol, const Options& opt);
GridPartition finish_partition(const Molecule& mol, const Options& opt,
                               const StagedPartition& staged);

// Everything about a molecule's grid that the partition rules need before
// any point exists (g2g/regenerate_partition.cpp:161-322,
// g2g/classify_functions.cpp:12-141): the bounding prism and its cube
// lattice, the sphere shells and radii of every atom, the radial nodes
// (radius, radial weight) of every atom shell, the atom distances and, per
// basis shell, the significance test of g2g/partition.cpp:229-254 reduced to
// a threshold.  The host staging and the device grid build both use it, so
// their points, groups and function sets are the same.
struct GridFrame {
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  unsigned dim[3] = {0, 0, 0};
  size_t ncubes = 0;
  std::vector<double> srad;          // per atom: sphere radius (0: no sphere)
  std::vector<unsigned> sphere_from; // per atom: first radial shell inside the sphere
  std::vector<int> rad_off;          // per atom: first entry of r1/wrad (atom-major)
  std::vector<double> r1, wrad;      // per atom shell: radius and radial weight
  std::vector<double> rij;           // natom x natom distances
  // basis shells in function order (s, then p, then d): first function,
  // length (1/3/6), atom, and the test: len 1: keep when a d^2 < thr;
  // len 3/6: keep when sqrt(d^2) < thr
  std::vector<int> sh_first, sh_len, sh_atom;
  std::vector<double> sh_a, sh_thr;
};
GridFrame grid_frame(const Molecule& mol, const Options& opt);

// Stable device radix sort of (key, value) pairs on the low `bits` key bits
// (gridsort.hip); temp == nullptr asks for the scratch size.  HIP status.
int radix_sort_pairs(void* temp, size_t* temp_bytes, const unsigned* keys_in,
                     unsigned* keys_out, const unsigned* vals_in, unsigned* vals_out, int n,
                     int bits, void* stream);

// ---------------------------------------------------------------------------
// Device engine
// ---------------------------------------------------------------------------
// Arithmetic of the two contractions (density and Fock GEMMs).  The
// functional, the row-dot epilogue, the Exc/N reductions and the Fock
// accumulation across slices and groups are FP64 in every mode.
//   FP64: FP64 values, FP64 MFMA.
//   FP32: FP32 values (half the bytes), FP32 MFMA, Fock products summed in
//         the matrix unit over at most 32 points before FP64 accumulation.
//   BF16X6: FP32 values stored point-major; every operand is split exactly
//         into three BF16 pieces (8 significant bits each) and a tile
//         product is the six piece products of weight >= 2^-16 relative
//         (hi*hi, hi*mid, mid*hi, hi*lo, mid*mid, lo*hi) on BF16 MFMA with
//         FP32 accumulation.
enum class Precision : int { FP64 = 0, FP32 = 1, BF16X6 = 2 };

// Launch schedule knobs, picked per grid by the autotuner (api.cpp).
//   fock_slice: points per Fock split-K slice (multiple of 128): longer
//               slices leave fewer partial tiles to add, shorter ones more
//               blocks for small groups.
//   density_variant (split-BF16 mode): 0 = 64-function row chunks with each
//               wave reading its P_g pieces; 1 = 128-function chunks and
//               2 = 64-function chunks with the pieces staged once per block
//               in LDS; 3 = 2 for groups of up to 64 functions and, above,
//               16-point blocks whose waves split the rows (6 row tiles a
//               wave per pass) and share the B operand through LDS; 4 = 3
//               with 4 row tiles a wave per pass; 5 = 2 in the radial-Hessian
//               layout (phi and its gradient as planes, the second
//               derivatives rebuilt from per-shell radial sums: 4 value
//               planes instead of 10 in HBM, more arithmetic per point);
//               6 and 7 = the basis evaluated inside the density kernel
//               (128- and 64-function chunks; no value planes, FP64 values);
//               8 = 2 with one more plane holding phi's FP32
//               remainder, Y_0 = P_g phi from four-piece operands (ten piece
//               products) and phi in FP64 in the epilogue; 9 = 2 with the
//               remainder added to first order in the epilogue only; 10
//               (default) = 8 with one block per (64-point tile, 64-function
//               row chunk): a tile's blocks run side by side on one XCD and
//               share its points' values through that L2; a second kernel
//               adds the chunks' row-dot sums in chunk order and evaluates
//               the functional; 11 = 10 with phi's four pieces split once
//               per grid (one more BF16 plane set, 8 bytes per point and
//               function) instead of in every build.
//               Variants 0-5 and 9 leave |dN| up to about 1e-6 at a few
//               hundred electrons; 6-8, 10 and 11 keep it near 1e-8 (measured errors
//               against FP64 are recorded by the tests).
//   persistent: 1 = the whole closed-shell split-BF16 build of density
//               variant 8 as one persistent launch over a queue of every
//               phase's items (k_build_persist; dependencies by per-group
//               counters), 0 = one launch per phase.  Bitwise the results of
//               variant 8's phase launches; other variants ignore it.
//   bucket_mhi / bucket_slice: a Fock slice per group size bucket: groups
//               of at most bucket_mhi[i] padded functions (ascending; the
//               first bucket that holds the group) take bucket_slice[i];
//               larger groups, and every group when the lists are empty,
//               take fock_slice.
struct Tuning {
  int fock_slice = 1024;
  int density_variant = 10;
  int persistent = 0;
  std::vector<int> bucket_mhi, bucket_slice;
  int slice_for(int m_pad) const {
    for (size_t i = 0; i < bucket_mhi.size(); ++i)
      if (m_pad <= bucket_mhi[i]) return bucket_slice[i];
    return fock_slice;
  }
};

struct SolveResult {
  double energy = 0.0;
  double electrons = 0.0;
  double seconds = 0.0;      // wall time of the device solve incl. copies
  double kernel_ms[4] = {0, 0, 0, 0};  // expand, density, fock, pack (events)
};

class Engine;

// One molecule resident on the device: its grid, work lists and buffers.
struct DeviceGrid;

class Engine {
 public:
  explicit Engine(int device = -1);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  // Uploads basis, grid and work lists; replaces any previous grid.
  void load(const Molecule& mol, const GridPartition& part, Precision prec = Precision::FP64);
  // Several molecules resident together: every solve phase is one launch over
  // all their groups; their packed rho/F are concatenated in molecule order
  // (packed_offset).  Results are bitwise those of each molecule alone.
  void load_batch(const std::vector<const Molecule*>& mols,
                  const std::vector<const GridPartition*>& parts, Precision prec);
  // Device grid rebuild: uploads the staged partitions, applies the
  // group-local Becke weights and zero-weight removal on the device, keeps
  // the groups with at least min_points_per_cube points and loads the batch
  // from there.  Returns the final partitions (point arrays left empty: the
  // points live on the device, see download_points).
  std::vector<GridPartition> load_staged(const std::vector<const Molecule*>& mols,
                                         const std::vector<const StagedPartition*>& staged,
                                         const Options& opt, Precision prec);
  // The whole rebuild on the device: points, cube/sphere groups and their
  // significant functions (the host computes only each molecule's grid
  // frame), then the Becke weights, zero-weight removal and the load as in
  // load_staged.  Returns the final partitions (points on the device).
  std::vector<GridPartition> load_generated(const std::vector<const Molecule*>& mols,
                                            const Options& opt, Precision prec);
  // Frees the device grid (arena, pinned buffers, graph); a load restores it.
  void unload();
  // Padded point arrays of the loaded grid (padded_points() entries each).
  void download_points(double* x, double* y, double* z, double* w);
  long long padded_points() const;
  // One XC Fock build from packed rho (host memory).  fock_accum (host) gets
  // += F_xc; pass nullptr to skip the Fock matrix.  forces_accum (host,
  // Fortran layout: component c of atom a at c * forces_ld + a) gets += the
  // XC forces of g2g/cpu/iteration.cpp:213-268; nullptr skips them.
  SolveResult solve(const double* rho_packed, double* fock_accum, bool want_energy,
                    double* forces_accum = nullptr, int forces_ld = 0);
  // Open shell (g2g/cpu/iteration.cpp:311-622): the alpha and beta densities
  // in, both spins' Fock matrices accumulated (both or neither), forces as
  // in solve.  Also accepts a closed-shell grid with rho_b = fock_b = null.
  SolveResult solve_open(const double* rho_a, const double* rho_b, double* fock_a, double* fock_b,
                         bool want_energy, double* forces_accum = nullptr, int forces_ld = 0);
  bool open_shell() const;
  // Reserves `bytes` of device memory once for the process; from then on the
  // grid arenas of every engine on that device are carved from it on the
  // host (no device allocation per grid).  Kept until the process ends.
  static void reserve_slab(int device, size_t bytes);  // device -1: the current one
  // Queues a build from a packed host density (copied to pinned memory at
  // once) with its copy-back; idle() tells when it has finished and
  // collect() waits for it and adds F onto fock_accum (may be null).  One
  // build may be queued at a time: a second solve_async, a synchronous solve
  // or a grid (re)load before collect() throws, since they reuse the pinned
  // buffers the queued build still reads and writes.
  void solve_async(const double* rho_packed);
  bool idle() const;
  SolveResult collect(double* fock_accum);
  // Same with device-resident input/output (no PCIe copies).  d_fock (may be
  // null) receives F_xc; d_forces (may be null) the forces of every batch
  // atom, atom-major (3 a + c), molecule k's atoms from atoms(k).
  SolveResult solve_device(const double* d_rho_packed, double* d_fock, bool want_energy,
                           double* d_forces = nullptr);
  // solve_device without any device allocation: where the split-BF16 force
  // pass needs its Y buffer (allocated by solve_device at the first forces
  // build), it takes it from a static device array of static_scratch_bytes()
  // instead, which every call on the device shares (calls that overlap on
  // one device race; each call writes Y in full before reading it).  Throws
  // when the grid needs more than that.
  SolveResult solve_device_static(const double* d_rho_packed, double* d_fock, double* d_forces);
  static size_t static_scratch_bytes();
  // Waits for the stream and reads the last solve's energy and electron count
  // (of the first molecule); read_energies: (Exc, N) of every molecule.
  void read_energy(double* energy, double* electrons);
  void read_energies(double* e_n);
  // Records one full solve on these device buffers as a hipGraph; replay
  // launches it (same buffers) on the engine stream.  A new load drops it.
  void capture_graph(const double* d_rho_packed, double* d_fock);
  void replay_graph();
  bool has_graph() const { return graph_ != nullptr; }
  int molecules() const;
  int atoms(int k) const;  // first batch atom of molecule k; k = molecules() gives the total
  long long packed_offset(int k) const;  // k = molecules() gives the total
  // Waits for the stream; HIP-event durations (ms) of the last solve's phases.
  void kernel_times(double ms[4]);

  // Takes effect at the next load / load_batch.
  void set_tuning(const Tuning& t) { tuning_ = t; }
  const Tuning& tuning() const { return tuning_; }
  int device() const { return device_; }
  size_t arena_bytes() const;
  // Duration (ms, HIP events) of the basis evaluation at the last load.
  double eval_ms() const;
  // Host wall time (ms) of the last load's phases: point weights, planning,
  // uploads, evaluation.
  void load_phases(double ms[4]) const {
    for (int k = 0; k < 4; ++k) ms[k] = load_ms_[k];
  }
  void* stream() const { return stream_; }
  const DeviceGrid* grid() const { return grid_; }

 private:
  int device_ = 0;
  void* stream_ = nullptr;
  DeviceGrid* grid_ = nullptr;
  void* events_[8] = {};
  void* graph_ = nullptr;
  void* stage_ = nullptr;     // device grid rebuild staging (kept, grow-only)
  size_t stage_cap_ = 0;
  void* becke_ = nullptr;     // Becke tables in near-first order (kept, grow-only)
  size_t becke_cap_ = 0;
  double load_ms_[4] = {};
  bool pending_ = false;      // an asynchronous build is queued and not yet collected
  void check_not_pending(const char* what) const;
  Tuning tuning_;
  void drop_graph();
  char* stage_buffer(size_t bytes);
  char* becke_buffer(size_t bytes);
  std::vector<GridPartition> weigh_and_load(const std::vector<const Molecule*>& mols,
                                            const struct StageHost& h, const struct StageDev& d,
                                            const Options& opt, Precision prec, double t_start);
  void load_batch_impl(const std::vector<const Molecule*>& mols,
                       const std::vector<const GridPartition*>& parts, Precision prec,
                       const struct DevicePoints* dpts);
};

}  // namespace xcq
