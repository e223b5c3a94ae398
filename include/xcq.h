/* Handle-based C API of the XC quadrature engine (libg2g.so).
 *
 * The g2g Fortran ABI (g2g.h) is a process-global singleton, like the
 * reference.  This API exposes the same engine per context so that callers
 * can hold several molecules at once and receive status codes instead of a
 * terminated process.  It has no counterpart in the reference; the g2g_*
 * entry points are implemented on top of one default context.
 *
 * All functions return 0 on success and a negative code on failure; the
 * message of the last failure on the calling thread is xcq_last_error().
 */
#ifndef LIO_AMD_XCQ_H
#define LIO_AMD_XCQ_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct xcq_context xcq_context;

/* The arrays of g2g_parameter_init_ in their Fortran layouts plus the
 * g2g_set_options_ knobs.  Referenced, not copied: they must outlive the
 * context or the next xcq_set_problem. */
typedef struct {
  unsigned natom, max_atoms, M, ngaussians;
  int normalize;            /* NORM: d xx/yy/zz carry 1/sqrt(3) */
  unsigned iexch;           /* 1..3 LDA, 9 PBE */
  const double* r;          /* max_atoms x 3 column-major, bohr */
  const double* Rm;         /* Becke size parameter by Z (0..54) */
  const unsigned* Iz;       /* natom */
  const unsigned* Nr;       /* radial shells by Z, grid type 0 */
  const unsigned* Nr2;      /* radial shells by Z, grid types 1, 2 */
  const unsigned* Nuc;      /* M, 1-based atom of each function */
  const unsigned* ncont;    /* M */
  const unsigned* nshell;   /* 3: #s, #p functions, #d functions */
  const double* c;          /* M x 13 column-major (ld = ngaussians) */
  const double* a;          /* M x 13 column-major (ld = ngaussians) */
  const double* e[3];       /* angular rules 50/116/194: n x 3 column-major */
  const double* wang[3];    /* their weights (sum 4 pi) */
  double little_cube_size;  /* default 8.0 */
  double sphere_radius;     /* default 0.6 */
  int assign_all_functions; /* default 0 */
  int remove_zero_weights;  /* default 1 */
  unsigned min_points_per_cube;   /* default 1 */
  unsigned max_function_exponent; /* default 10 */
  int open;                 /* 1: open shell (PBE only); solve with xcq_solve_open */
} xcq_problem;

/* Partition inventory of one point group. */
typedef struct {
  int kind;      /* 0 cube, 1 sphere */
  int npoints;   /* after zero-weight removal */
  int nfuncs;    /* expanded functions */
  int s, p, d;   /* selected shells per kind */
  int nnuc;      /* atoms owning selected functions */
  double wsum;   /* sum of point weights */
  double xsum;   /* sum of point x coordinates */
  unsigned long long fhash; /* FNV-1a of the selected shells' first indices */
} xcq_group_info;

/* Grid summary after xcq_build_grid. */
typedef struct {
  int ngroups;
  long long npoints;          /* real points */
  long long npoints_padded;
  double flops;               /* algorithmic 10 n m^2 summed over groups */
  double bytes_values;        /* algorithmic 2*10*n*m*8 evaluation bytes */
  double grid_seconds;        /* host partition build */
  double upload_seconds;
  double eval_ms;             /* basis evaluation kernel (HIP events) */
  unsigned long long arena_bytes;
  int density_tiles, fock_items;
  /* host wall time (ms) of the load's phases: point weights (staging copies,
   * Becke weights, group counts; device-built grids only), planning (work
   * lists and arena), uploads, basis evaluation until the stream is idle */
  double load_ms[4];
} xcq_grid_info;

typedef struct {
  double energy, electrons;
  double seconds;             /* host wall time of the call */
  double kernel_ms[4];        /* expand, density, fock, pack (HIP events) */
} xcq_solve_info;

const char* xcq_last_error(void);

int xcq_create(int device, xcq_context** out);
int xcq_destroy(xcq_context* ctx);
int xcq_set_problem(xcq_context* ctx, const xcq_problem* problem);
/* Arithmetic of the density and Fock contractions (the functional, the
 * energy/electron reductions and the Fock accumulation stay FP64):
 *   XCQ_FP64 (default) FP64 values and FP64 MFMA;
 *   XCQ_FP32           FP32 values and FP32 MFMA;
 *   XCQ_BF16X6         FP32 values, each split exactly into three BF16
 *                      pieces; six BF16 MFMA products per tile (the pieces
 *                      of combined weight above 2^-24), FP32 accumulation.
 * Takes effect at the next xcq_build_grid, or at once (re-uploading and
 * re-evaluating the basis) when a grid is already built.  No reference
 * counterpart: LIO's GPU path picks precision at compile time. */
enum { XCQ_FP64 = 0, XCQ_FP32 = 1, XCQ_BF16X6 = 2 };
int xcq_set_precision(xcq_context* ctx, int precision);

/* Builds grid and partition for grid_type (0/1/2) and uploads them. */
int xcq_build_grid(xcq_context* ctx, unsigned grid_type);
/* Where the next xcq_build_grid builds the grid and partition:
 * XCQ_GRID_HOST (default; OpenMP, as g2g/regenerate_partition.cpp and
 * g2g/cpu/weight.cpp), XCQ_GRID_DEVICE (all on the GPU: points, cube and
 * sphere groups, significant functions, group-local Becke weights, zero-weight
 * removal and compaction; the host computes only the bounding prism, the
 * radial nodes and the shell thresholds) or XCQ_GRID_STAGED (points, groups
 * and functions on the host, weights and compaction on the GPU).  The groups,
 * points and weights are the same in every mode.  No reference counterpart
 * (LIO's GPU build generates and groups points on the host and weighs them on
 * the GPU). */
enum { XCQ_GRID_HOST = 0, XCQ_GRID_DEVICE = 1, XCQ_GRID_STAGED = 2 };
int xcq_set_grid_mode(xcq_context* ctx, int mode);
/* Frees the context's device memory (grid, values, work buffers); the
 * problem stays set and the next xcq_build_grid restores it.  No reference
 * counterpart: lets a server hold more molecules than fit in HBM at once. */
int xcq_release_device(xcq_context* ctx);
int xcq_query_grid(xcq_context* ctx, xcq_grid_info* out);
/* Group inventory of the built grid; returns the count through *ngroups. */
int xcq_groups(xcq_context* ctx, xcq_group_info* out, int max, int* ngroups);
/* One XC Fock build from a packed host density; fock_accum (+=, may be NULL). */
int xcq_solve(xcq_context* ctx, const double* rho_packed, double* fock_accum,
              xcq_solve_info* info);
/* Same with device pointers (rho and F packed, in device memory); no PCIe
 * copies and no synchronisation: the work is queued on the context stream
 * (xcq_stream); energy and electrons are read back by xcq_sync_energy. */
int xcq_solve_device(xcq_context* ctx, const double* d_rho_packed, double* d_fock);
/* Reserves `bytes` of device memory once per process (device -1: the current
 * one); the grid arenas of every context and batch on that device are then
 * carved from it on the host, with no device allocation per grid, and a grid
 * that does not fit fails with "out of memory".  Kept until the process ends.
 * No reference counterpart: for serving loops that rebuild and evict grids of
 * different sizes at every request. */
int xcq_reserve_slab(int device, unsigned long long bytes);
/* Energy and XC forces (g2g_solve_groups_ computation types 2/3,
 * g2g/cpu/iteration.cpp:213-268 for the force, g2g/partition.cpp:516-526 for
 * the accumulation): forces_accum is the caller's natom x 3 column-major
 * array with leading dimension ld >= natom, accumulated (+=); fock_accum
 * (may be NULL) additionally receives += F_xc. */
int xcq_solve_forces(xcq_context* ctx, const double* rho_packed, double* fock_accum,
                     double* forces_accum, int ld, xcq_solve_info* info);
/* Open shell (g2g/cpu/iteration.cpp:311-622, spin-polarised PBE of
 * g2g/pointxc/calc_ggaOS.h): the packed alpha and beta densities in (same
 * packing as rho), both spins' Fock matrices accumulated (+=) -- both or
 * neither -- and the forces as in xcq_solve_forces (forces_accum may be
 * NULL).  Batches of open-shell problems use the xcq_batch_* entry points
 * with d_rho and d_fock holding the alpha then the beta packed blocks
 * (2 x offsets[n] entries each). */
int xcq_solve_open(xcq_context* ctx, const double* rho_a, const double* rho_b, double* fock_a,
                   double* fock_b, double* forces_accum, int ld, xcq_solve_info* info);
int xcq_sync_energy(xcq_context* ctx, double* energy, double* electrons);
/* Asynchronous build from a packed host density (copied at once): queued on
 * the context stream with the copy-back of F, Exc and N.  xcq_query returns
 * 1 when it has finished, 0 while it runs (negative on error); xcq_collect
 * waits for it, adds F onto fock_accum (may be NULL) and fills info.  No
 * reference counterpart: lets one host thread keep many molecules' builds in
 * flight on their own streams. */
int xcq_solve_async(xcq_context* ctx, const double* rho_packed);
int xcq_query(xcq_context* ctx);
int xcq_collect(xcq_context* ctx, double* fock_accum, xcq_solve_info* info);
/* Waits for the context stream and returns the HIP-event durations (ms) of
 * the last solve's phases: [0] density expansion, [1] density + functional,
 * [2] Fock tiles + slice sums, [3] Fock packing + energy reduction. */
int xcq_kernel_times(xcq_context* ctx, double ms[4]);
/* HIP stream (hipStream_t) the context launches on. */
void* xcq_stream(xcq_context* ctx);

/* Host-only partition build (no device needed): inventory of the groups the
 * engine would build for this problem and grid type. */
int xcq_host_partition(const xcq_problem* problem, unsigned grid_type,
                       xcq_group_info* out, int max, int* ngroups);

/* ---- Batches: several molecules resident on one device --------------------
 * No reference counterpart (LIO solves one molecule per process); for callers
 * that build many independent XC Fock matrices at once.  Each solve phase is
 * one launch over the groups of all molecules (heavy groups first), so
 * uneven molecules share the machine; the per-molecule results are bitwise
 * those of the molecule solved alone.  The molecules' packed rho and F are
 * concatenated in molecule order; offsets[k] is the first packed entry of
 * molecule k and offsets[n] the total (n + 1 entries).  All molecules of a
 * batch use the same functional. */
typedef struct xcq_batch xcq_batch;
int xcq_batch_create(int device, int n, const xcq_problem* problems,
                     const unsigned* grid_types, int precision, xcq_batch** out);
/* Same with the grid weighting on the host or the device (XCQ_GRID_*). */
int xcq_batch_create_ex(int device, int n, const xcq_problem* problems,
                        const unsigned* grid_types, int precision, int grid_mode,
                        xcq_batch** out);
int xcq_batch_destroy(xcq_batch* b);
int xcq_batch_offsets(xcq_batch* b, long long* offsets);
/* Queues one build of every molecule on the batch stream (device buffers). */
int xcq_batch_solve_device(xcq_batch* b, const double* d_rho, double* d_fock);
/* Records one build on these buffers as a hipGraph (replayed by
 * xcq_batch_replay, one launch call per build). */
/* Full build plus the XC forces of every batch atom into d_forces (device,
 * atom-major: component c of batch atom a at 3 a + c, overwritten); d_fock
 * may be NULL.  xcq_batch_atoms gives the first batch atom of each molecule
 * (n + 1 entries, the last the total). */
int xcq_batch_solve_forces(xcq_batch* b, const double* d_rho, double* d_fock, double* d_forces);
/* xcq_batch_solve_forces without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory during
 * the call (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  xcq_batch_solve_forces allocates, at its first call on a
 * split-BF16 grid, a buffer for Y = P_g phi of the force pass; this entry
 * point takes Y from a static device array of the library instead
 * (xcq_static_scratch_bytes(), 64 MiB, held by each device the library runs
 * on until the process ends) and fails with an error when the batch needs
 * more.  Every call on one device shares that array: calls that overlap on
 * one device (several batches' streams at once) race.  Same results as
 * xcq_batch_solve_forces. */
int xcq_batch_solve_forces_static(xcq_batch* b, const double* d_rho, double* d_fock,
                                  double* d_forces);
long long xcq_static_scratch_bytes(void);
int xcq_batch_atoms(xcq_batch* b, int* first_atom);
int xcq_batch_capture(xcq_batch* b, const double* d_rho, double* d_fock);
int xcq_batch_replay(xcq_batch* b);
/* Waits; Exc and N of every molecule (n entries each). */
int xcq_batch_energies(xcq_batch* b, double* energy, double* electrons);
int xcq_batch_kernel_times(xcq_batch* b, double ms[4]);
void* xcq_batch_stream(xcq_batch* b);
long long xcq_batch_arena_bytes(xcq_batch* b);
/* Host wall time (ms) of the last load's phases, as xcq_grid_info.load_ms. */
int xcq_batch_load_ms(xcq_batch* b, double ms[4]);
/* Autotuner: times `reps` builds for every candidate schedule c < ncand --
 * Fock slice length slices[c] (points, multiple of 128) and, when variants is
 * not NULL, split-BF16 density kernel variants[c] (0: 64-function chunks, 1 and 2:
 * 128- and 64-function chunks with LDS-staged density pieces, 3 and 4:
 * row-split 16-point blocks for groups above 64 functions, 5: 2 in the
 * radial-Hessian layout, 4 value planes instead of 10, 6 and 7: the basis
 * evaluated in the density kernel, 8: 2 with phi's FP32 remainder plane and
 * four-piece Y_0 products, 9: 2 with the remainder added in the epilogue,
 * 10 (default): 8 with one block per (64-point tile, 64-function row chunk),
 * the blocks of a tile consecutive on one XCD so that its points' values
 * are fetched into that XCD's L2 once for all chunks, 11: 10 with phi's
 * pieces split once per grid; 0-5 and 9 are the faster, less accurate
 * forms: |dN| near 1e-6 at a few hundred electrons) -- on these buffers
 * (rebuilding the work lists for each), keeps the fastest and returns the
 * per-candidate ms in ms_out; *best is its index.  A captured graph is
 * dropped. */
int xcq_batch_tune(xcq_batch* b, const double* d_rho, double* d_fock, int reps,
                   const int* slices, const int* variants, int ncand, double* ms_out,
                   int* best);
/* Sets the schedule directly (a zero slice or a negative variant keeps that
 * knob): for a context it takes effect at the next xcq_build_grid, a batch
 * reloads at once.  No reference counterpart. */
int xcq_set_schedule(xcq_context* ctx, int fock_slice, int density_variant);
int xcq_batch_set_schedule(xcq_batch* b, int fock_slice, int density_variant);
/* The batch's schedule knobs (any pointer may be NULL).  No reference
 * counterpart. */
int xcq_batch_get_schedule(xcq_batch* b, int* fock_slice, int* density_variant, int* persistent);
/* on = 1: a closed-shell split-BF16 build with density variant 8 runs as one
 * persistent launch over an ordered queue of every phase's items (expansion,
 * density and functional, Fock tile slices, tile sums, packing, energies),
 * dependencies tracked by per-group counters; on = 0 (default): one launch
 * per phase.  The results are bitwise those of variant 8's phase launches.
 * Takes effect at the next xcq_build_grid (a batch reloads at once).  Other
 * modes (FP64, FP32, open shell, other density variants, the default 10
 * included) keep the phase launches. */
int xcq_set_persistent(xcq_context* ctx, int on);
/* Fock slice per group size bucket (the autotuner's schedule): groups of at
 * most m_hi[i] padded functions (m_hi ascending; the first bucket that holds
 * the group) take slices[i] points per Fock slice (multiples of 128), larger
 * groups the batch's fock_slice; nb = 0 clears the buckets.  Reloads. */
int xcq_batch_set_bucket_slices(xcq_batch* b, int nb, const int* m_hi, const int* slices);
/* Padded function and point counts of every group of the batch (molecule
 * order); *ngroups gets the count, at most max entries are written. */
int xcq_batch_group_sizes(xcq_batch* b, int* m_pad, long long* npts_pad, int max, int* ngroups);
int xcq_batch_set_persistent(xcq_batch* b, int on);

#ifdef __cplusplus
}
#endif
#endif
