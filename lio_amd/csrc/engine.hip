// Device engine for the XC Fock build on MI355X (gfx950), FP64 path.
//
// Grid build (once per geometry, like the reference's cached GPU groups,
// g2g/cuda/iteration.cu:792-881):
//   k_eval_basis       phi, grad phi and the Hessian of every significant basis
//                      function at every point of its group, written once to
//                      HBM as planes [value][function][point] (points
//                      contiguous, coalesced).  HBM-bound.
// Per XC Fock build (one solve):
//   k_gather_density   packed rho -> dense symmetric block P_g per group
//   k_density          Y_t^T = P_g Phi_t^T for t = phi, d/dx, d/dy, d/dz on
//                      FP64 MFMA (16x16x4), 64 points per block; row-dot
//                      epilogue -> rho, grad rho, Hessian of rho -> PBE/LDA
//                      -> factor_p = w_p v_p and per-block (Exc, N)
//   k_fock             F_g tile (64 x 64) over a 512-point slice,
//                      phi^T diag(factor) phi on FP64 MFMA, one partial/slice
//   k_fock_group       sums the slices of each tile in fixed order
//   k_fock_pack        packed F_IJ = sum over groups holding I and J, in
//                      ascending group order (no float atomics)
//   k_energy           fixed-order reduction of the per-block (Exc, N)
// Every reduction has a fixed order, so repeated solves are bitwise equal.
//
// The maths follows g2g/cpu/iteration.cpp:116-291 (closed shell): with P the
// true symmetric density (packed off-diagonals halved) and Y = Phi P,
//   rho = sum_i phi_i Y_i,  d_a rho = 2 sum_i d_a phi_i Y_i,
//   d_a d_b rho = 2 sum_i (d_ab phi_i Y_i + d_a phi_i (Y_b)_i),
//   F_ij += sum_p phi_i phi_j w_p y2a_p,  Exc = sum_p rho_p w_p (ex + ec).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <map>
#include <atomic>
#include <mutex>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include "xc_point.h"
#include "xcq.h"

#define XCQ_HIP(call)                                                                    \
  do {                                                                                   \
    hipError_t e_ = (call);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e_) + " (" + \
                               #call + ") at " + __FILE__ + ":" + std::to_string(__LINE__)); \
  } while (0)

namespace xcq {

namespace {

constexpr int kPtAlign = 16;     // point padding (one MFMA column tile)
constexpr int kFnAlign = 16;     // function padding (one MFMA row tile)
constexpr int kDensPts = 64;     // points per k_density block (16 per wave)
constexpr int kRowSplitMin = 64; // split-BF16 row-split density above this m_pad
constexpr int kFnChunk = 64;     // functions per density pass / Fock tile side
constexpr int kXcds = 8;         // MI355X: XCDs, each with its own L2 (block b runs on b % 8)
constexpr int kSliceDefault = 1024;  // points per Fock slice (split-K over blocks)
constexpr int kPackMapMax = 16384;  // groups per molecule the packing kernel maps in LDS
constexpr int kEvalPts = 256;    // threads per evaluation block
constexpr int kEvalSpan = 4096;  // points (or elements) one evaluation block covers

typedef double d4 __attribute__((ext_vector_type(4)));

struct DevFn {
  double a[kMaxContractions];
  double c[kMaxContractions];
  double cx, cy, cz;
  int ncont, lx, ly, lz;
};

struct DevGroup {
  int pt_off;          // first padded point
  int npts;            // real points
  int npts_pad;        // multiple of kPtAlign
  int fn_off;          // first entry of the padded function list
  int m;               // real functions
  int m_pad;           // multiple of kFnAlign
  int ntile;           // kFnChunk tiles per side
  int nslice;          // Fock slices
  long long val_off;   // value planes: nval x m_pad x npts_pad
  long long pg_off;    // P_g / F_g block: m_pad x m_pad
  int fock_base;       // first Fock item
  int slen;            // points per Fock slice (multiple of 128; per size bucket)
  long long phi_off;   // function-major phi plane (split-BF16 mode): m_pad x npts_pad
  int mol;             // molecule of the batch
  int pad2_;
  long long force_off; // force partials: one 3 x m_pad block per 64-point block
  int shell_off;       // radial-Hessian layout: first shell in the shell table
  int nsh_pad;         // its shells, padded to a multiple of 4
  long long rad_off;   // radial planes R0, R1, R2: 3 x npts_pad x nsh_pad
};

// Per expanded function of a group in the radial-Hessian and in-kernel
// evaluation layouts: its centre, the factor relating its contraction to its
// shell's first function (the d xx/yy/zz normalisation) and (shell, lx, ly,
// lz) packed as shell | lx << 16 | ly << 20 | lz << 24, shell counted from
// the group's first shell.  Padding functions carry scale 0 and the last
// real function's shell.
struct FnMeta {
  double cx, cy, cz;
  double scale;
  int code;
  int pad_;
};

// One molecule of a batch: its functions are [fn_base, fn_base + M) of the
// batch numbering, its packed rho/F start at rho_off of the concatenated
// buffers, and its density blocks are eidx[eb .. ee) (ascending).
struct DevMol {
  int M, fn_base;
  long long rho_off;
  int eb, ee;
  int g0, ng;  // its groups [g0, g0 + ng) in the batch
};

struct EvalItem {
  int group, k, p0, pad_;
};
struct DensityItem {
  int group, p0;
};
struct FockItem {
  int group, ti, tj, slice;
};
struct GatherItem {
  int group, start;  // first element of the m_pad^2 block handled
};

struct Axis {
  double m, d, dd;
};

__device__ inline Axis axis(double v, int l) {
  if (l == 0) return {1.0, 0.0, 0.0};
  if (l == 1) return {v, 1.0, 0.0};
  return {v * v, 2.0 * v, 2.0};
}

// Contracted Cartesian Gaussian x^lx y^ly z^lz sum_k c_k exp(-a_k r^2) and
// its derivatives (g2g/cpu/functions.cpp:56-354): NV = 1 value, 4 adds the
// gradient (LDA with forces), 10 also the Hessian (xx, yy, zz, xy, xz, yz).  Primitives with
// a r^2 > 70 are skipped as in the reference.
// Radial sums of a contraction at squared distance r2: t = sum c e, tg =
// sum c a e, th = sum c a^2 e with e = exp(-a r2), primitives with a r2 > 70
// skipped (g2g/cpu/functions.cpp:56-82).  NV > 1 adds tg, NV > 4 th.
template <int NV>
__device__ inline void radial_sums(const DevFn& b, double r2, double& t, double& tg, double& th) {
  t = tg = th = 0.0;
  const int nc = b.ncont;
  for (int k = 0; k < nc; ++k) {
    const double a = b.a[k];
    if (a * r2 > 70.0) continue;
    const double e = exp(-a * r2) * b.c[k];
    t += e;
    if (NV > 1) tg += e * a;
    if (NV > 4) th += e * (a * a);
  }
}

// Values of x^lx y^ly z^lz g(r) from the radial sums of g at offset v from
// the centre: NV = 1 the value, 4 adds the gradient, 10 the Hessian (xx, yy,
// zz, xy, xz, yz); the operations of g2g/cpu/functions.cpp:83-354.
template <int NV>
__device__ inline void basis_from_radial(int lx, int ly, int lz, double vx, double vy, double vz,
                                         double t, double tg, double th, double* o) {
  const Axis X = axis(vx, lx), Y = axis(vy, ly), Z = axis(vz, lz);
  const double P = X.m * Y.m * Z.m;
  o[0] = P * t;
  if (NV == 1) return;
  const double Px = X.d * Y.m * Z.m, Py = X.m * Y.d * Z.m, Pz = X.m * Y.m * Z.d;
  const double gx = -2.0 * vx * tg, gy = -2.0 * vy * tg, gz = -2.0 * vz * tg;
  o[1] = Px * t + P * gx;
  o[2] = Py * t + P * gy;
  o[3] = Pz * t + P * gz;
  if (NV == 4) return;
  const double Pxx = X.dd * Y.m * Z.m, Pyy = X.m * Y.dd * Z.m, Pzz = X.m * Y.m * Z.dd;
  const double Pxy = X.d * Y.d * Z.m, Pxz = X.d * Y.m * Z.d, Pyz = X.m * Y.d * Z.d;
  o[4] = Pxx * t + 2.0 * Px * gx + P * (4.0 * vx * vx * th - 2.0 * tg);
  o[5] = Pyy * t + 2.0 * Py * gy + P * (4.0 * vy * vy * th - 2.0 * tg);
  o[6] = Pzz * t + 2.0 * Pz * gz + P * (4.0 * vz * vz * th - 2.0 * tg);
  o[7] = Pxy * t + Px * gy + Py * gx + P * (4.0 * vx * vy * th);
  o[8] = Pxz * t + Px * gz + Pz * gx + P * (4.0 * vx * vz * th);
  o[9] = Pyz * t + Py * gz + Pz * gy + P * (4.0 * vy * vz * th);
}

template <int NV>
__device__ inline void eval_basis(const DevFn& b, double px, double py, double pz, double* o) {
  const double vx = px - b.cx, vy = py - b.cy, vz = pz - b.cz;
  const double r2 = vx * vx + vy * vy + vz * vz;
  double t, tg, th;
  radial_sums<NV>(b, r2, t, tg, th);
  basis_from_radial<NV>(b.lx, b.ly, b.lz, vx, vy, vz, t, tg, th, o);
}

// 16x16x4 MFMA in FP64 or FP32.  Operands: lane l supplies A[l%16][l/16] and
// B[l/16][l%16].  Results (measured on the MI355X): the FP64 form leaves
// D[r][c] in lane (r%4)*16 + c, item r/4 (lane l, item i holds row
// l/16 + 4 i); the FP32 form leaves it in lane (r/4)*16 + c, item r%4 (row
// 4 (l/16) + i).
typedef float f4 __attribute__((ext_vector_type(4)));

template <class T>
struct Mma;

template <>
struct Mma<double> {
  typedef d4 acc;
  __device__ static inline void op(double a, double b, d4& c) {
    c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int lk, int i) { return lk + 4 * i; }
};

template <>
struct Mma<float> {
  typedef f4 acc;
  __device__ static inline void op(float a, float b, f4& c) {
    c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static constexpr int row(int lk, int i) { return 4 * lk + i; }
};

// Packed index of (i, j), i <= j, column-major lower triangle (partition.cpp:132-134).
__device__ __host__ inline long long packed_index(long long i, long long j, long long M) {
  return i * M - (i * (i - 1)) / 2 + (j - i);
}

// ---------------------------------------------------------------------------
// basis evaluation: one function x 256 points per block
// ---------------------------------------------------------------------------
template <int NV, class T>
__global__ __launch_bounds__(256) void k_eval_basis(const EvalItem* __restrict__ items,
                                                    const DevGroup* __restrict__ groups,
                                                    const int* __restrict__ fnl,
                                                    const DevFn* __restrict__ bas,
                                                    const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    T* __restrict__ vals, int to_phi) {
  const EvalItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int f = fnl[g.fn_off + it.k];
  const long long plane = (long long)g.m_pad * g.npts_pad;
  const int pe = min(it.p0 + kEvalSpan, g.npts_pad);
  for (int p = it.p0 + threadIdx.x; p < pe; p += kEvalPts) {
    double v[NV];
    if (f >= 0) {
      const int q = g.pt_off + p;
      eval_basis<NV>(bas[f], px[q], py[q], pz[q], v);
    } else {
      for (int k = 0; k < NV; ++k) v[k] = 0.0;
    }
    // to_phi: the function-major phi plane of the split-BF16 path
    T* out = vals + (to_phi ? g.phi_off : g.val_off) + (long long)it.k * g.npts_pad + p;
    for (int k = 0; k < NV; ++k) out[k * plane] = (T)v[k];
  }
}

// ---------------------------------------------------------------------------
// density blocks P_g[k][j] = P(f_k, f_j) with the true symmetric P
// ---------------------------------------------------------------------------
// Entry (fk, fj) of the true symmetric density (packed off-diagonals
// halved); fk, fj are batch function indices of molecule mol, or -1.
__device__ inline double density_entry(const DevMol& mol, const double* __restrict__ rho, int fk,
                                       int fj) {
  if (fk < 0 || fj < 0) return 0.0;
  const int lo = min(fk, fj) - mol.fn_base, hi = max(fk, fj) - mol.fn_base;
  const double v = rho[mol.rho_off + packed_index(lo, hi, mol.M)];
  return lo != hi ? 0.5 * v : v;
}

template <class T>
__global__ __launch_bounds__(256) void k_gather_density(const GatherItem* __restrict__ items,
                                                        const DevGroup* __restrict__ groups,
                                                        const DevMol* __restrict__ mols,
                                                        const int* __restrict__ fnl,
                                                        const double* __restrict__ rho,
                                                        T* __restrict__ pg) {
  const GatherItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int e = it.start + threadIdx.x;
  if (e >= g.m_pad * g.m_pad) return;
  const int k = e / g.m_pad, j = e % g.m_pad;
  const double v = density_entry(mols[g.mol], rho, fnl[g.fn_off + k], fnl[g.fn_off + j]);
  pg[g.pg_off + e] = (T)v;
}

// Functional stage of a density launch for point q from its row-dot sums s
// (rho and the halves of grad rho and of its second derivatives).  spin 0:
// closed shell.  Open shell runs the density launch once per spin: spin 1
// (alpha) stores its ten sums in `sums` (plane k at k * stride + q) and
// contributes no energy; spin 2 (beta) reads them back and evaluates the
// spin-polarised functional (g2g/cpu/iteration.cpp:311-488), writing both
// spins' factors w v_sigma.  Open shell is GGA only.
template <int NT>
__device__ inline void point_stage(const double* s, int q, int iexch, int spin,
                                   const double* __restrict__ pw, double* __restrict__ factor,
                                   double* __restrict__ factor_b, double* __restrict__ sums,
                                   long long stride, double& e_w, double& n_w) {
  if constexpr (NT == 4) {
    if (spin == 1) {
      for (int k = 0; k < 10; ++k) sums[k * stride + q] = s[k];
      e_w = n_w = 0.0;
      return;
    }
    if (spin == 2) {
      double a[10];
      for (int k = 0; k < 10; ++k) a[k] = sums[k * stride + q];
      const double ga[3] = {2.0 * a[1], 2.0 * a[2], 2.0 * a[3]};
      const double h1a[3] = {2.0 * a[4], 2.0 * a[5], 2.0 * a[6]};
      const double h2a[3] = {2.0 * a[7], 2.0 * a[8], 2.0 * a[9]};
      const double gb[3] = {2.0 * s[1], 2.0 * s[2], 2.0 * s[3]};
      const double h1b[3] = {2.0 * s[4], 2.0 * s[5], 2.0 * s[6]};
      const double h2b[3] = {2.0 * s[7], 2.0 * s[8], 2.0 * s[9]};
      double ex, ec, va, vb;
      xc::gga_point_open(a[0], ga, h1a, h2a, s[0], gb, h1b, h2b, ex, ec, va, vb);
      const double w = pw[q];
      factor[q] = w * va;
      factor_b[q] = w * vb;
      const double dens = a[0] + s[0];
      e_w = (dens * w) * (ex + ec);
      n_w = dens * w;
      return;
    }
  }
  const double dens = s[0];
  double ex, ec, y2a;
  if constexpr (NT == 4) {
    const double gr[3] = {2.0 * s[1], 2.0 * s[2], 2.0 * s[3]};
    const double h1[3] = {2.0 * s[4], 2.0 * s[5], 2.0 * s[6]};
    const double h2[3] = {2.0 * s[7], 2.0 * s[8], 2.0 * s[9]};
    xc::gga_point_x(dens, gr, h1, h2, iexch, ex, ec, y2a);
  } else {
    xc::lda_point(dens, iexch, ex, ec, y2a);
  }
  const double w = pw[q];
  factor[q] = w * y2a;
  e_w = (dens * w) * (ex + ec);
  n_w = dens * w;
}

// ---------------------------------------------------------------------------
// density + functional: 64 points per block, 16 per wave
// ---------------------------------------------------------------------------
// Wave w owns points p0 = 64 b + 16 w (the 16 columns of its MFMA tiles) and
// walks the functions in chunks of 64 (4 row tiles).  For each row tile rt
// and derivative t it accumulates (Y_t)^T[j][p] = sum_k P[j][k] Phi_t[k][p];
// lane l then holds, for point p0 + l%16, the four functions
// j = 16 rt + l/16 + 4 i (i = item) of all NT derivatives, and contracts them
// with the values of those functions at that point (10 row-dot partials).
template <int NT, class T>
__global__ __launch_bounds__(256, 2) void k_density(const DensityItem* __restrict__ items,
                                                 const DevGroup* __restrict__ groups,
                                                 const T* __restrict__ pg,
                                                 const T* __restrict__ vals,
                                                 const double* __restrict__ pw,
                                                 double* __restrict__ factor,
                                                 double* __restrict__ epart, int iexch,
                                                 int spin, double* __restrict__ factor_b,
                                                 double* __restrict__ sums, long long sstride) {
  __shared__ double sE[4][2];
  constexpr int NS = (NT == 4) ? 10 : 1;
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  double e_w = 0.0, n_w = 0.0;
  if (p0 < g.npts_pad) {
    const int np = g.npts_pad, mp = g.m_pad;
    const long long plane = (long long)mp * np;
    const T* V = vals + g.val_off;
    const T* Pm = pg + g.pg_off;
    double s[NS];
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
    for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
      const int nrt = min(4, (mp - c0) / 16);
      typename Mma<T>::acc acc[4][NT];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[rt][t] = typename Mma<T>::acc{0, 0, 0, 0};
      const T* Bk = V + (long long)lk * np + p0 + lrow;   // Phi_t[k][p]
      const T* Ak = Pm + (long long)lk * mp + c0 + lrow;  // P[k][j] = P[j][k]
      // operands of step k0 in (a, b); the next step's are loaded before the
      // MFMAs of this one so the loads overlap the matrix pipe
      T a[4], b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = Bk[t * plane];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) a[rt] = (rt < nrt) ? Ak[16 * rt] : T(0);
      for (int k0 = 0; k0 < mp; k0 += 4) {
        T an[4], bn[NT];
        const bool more = k0 + 4 < mp;
        const long long kb = (long long)(more ? k0 + 4 : k0);
#pragma unroll
        for (int t = 0; t < NT; ++t) bn[t] = Bk[t * plane + kb * np];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) an[rt] = (rt < nrt) ? Ak[kb * mp + 16 * rt] : T(0);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          if (rt < nrt) {
#pragma unroll
            for (int t = 0; t < NT; ++t) Mma<T>::op(a[rt], b[t], acc[rt][t]);
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) b[t] = bn[t];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) a[rt] = an[rt];
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        if (rt >= nrt) continue;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int j = c0 + 16 * rt + Mma<T>::row(lk, i);
          const T* vj = V + (long long)j * np + p0 + lrow;
          if constexpr (NT == 4) {
            double v[10];
#pragma unroll
            for (int q = 0; q < 10; ++q) v[q] = (double)vj[q * plane];
            const double y0 = acc[rt][0][i], yx = acc[rt][1][i], yy = acc[rt][2][i],
                         yz = acc[rt][3][i];
            s[0] += v[0] * y0;
            s[1] += v[1] * y0;
            s[2] += v[2] * y0;
            s[3] += v[3] * y0;
            s[4] += v[4] * y0 + v[1] * yx;
            s[5] += v[5] * y0 + v[2] * yy;
            s[6] += v[6] * y0 + v[3] * yz;
            s[7] += v[7] * y0 + v[1] * yy;
            s[8] += v[8] * y0 + v[1] * yz;
            s[9] += v[9] * y0 + v[2] * yz;
          } else {
            s[0] += (double)vj[0] * (double)acc[rt][0][i];
          }
        }
      }
    }
    for (int k = 0; k < NS; ++k) {  // the 4 lane groups holding the same point
      s[k] += __shfl_xor(s[k], 16);
      s[k] += __shfl_xor(s[k], 32);
    }
    if (lk == 0)
      point_stage<NT>(s, g.pt_off + p0 + lrow, iexch, spin, pw, factor, factor_b, sums, sstride,
                      e_w, n_w);
    for (int o = 8; o >= 1; o >>= 1) {
      e_w += __shfl_xor(e_w, o);
      n_w += __shfl_xor(n_w, o);
    }
  }
  if (lane == 0) {
    sE[wave][0] = e_w;
    sE[wave][1] = n_w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    epart[2 * blockIdx.x] = ((sE[0][0] + sE[1][0]) + sE[2][0]) + sE[3][0];
    epart[2 * blockIdx.x + 1] = ((sE[0][1] + sE[1][1]) + sE[2][1]) + sE[3][1];
  }
}

// Item i of an accumulator with a runtime index (selects, no scratch).
template <class A>
__device__ inline double pick(const A& v, int i) {
  return (double)(i == 0 ? v[0] : (i == 1 ? v[1] : (i == 2 ? v[2] : v[3])));
}

// ---------------------------------------------------------------------------
// Fock: one block (4 waves) per (group, tile pair, 512-point slice)
// ---------------------------------------------------------------------------
// out[i][j] = sum_p phi_i(p) factor_p phi_j(p) for i in tile ti, j in tile tj.
// Wave w takes the points pb + [128 w, 128 w + 128) of the slice; the four
// partial tiles are added through LDS in wave order (fixed, reproducible).
// The K order is permuted so that the four K slots a lane feeds to four
// consecutive MFMAs are four consecutive points: one 32-byte load per lane
// and operand row instead of four scattered 8-byte loads.
template <class T>
__global__ __launch_bounds__(256, 2) void k_fock(const FockItem* __restrict__ items,
                                                 const DevGroup* __restrict__ groups,
                                                 const T* __restrict__ vals,
                                                 const double* __restrict__ factor,
                                                 double* __restrict__ fpart) {
  typedef typename Mma<T>::acc A4;
  typedef T V4 __attribute__((ext_vector_type(4)));
  // FP32 products are summed over at most 32 points in the matrix unit and
  // then added into FP64; FP64 accumulates in place.
  constexpr bool kFlush = sizeof(T) == 4;
  __shared__ double sF[kFnChunk * kFnChunk];
  const FockItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lrow = lane & 15, lk = lane >> 4;
  const int np = g.npts_pad, mp = g.m_pad;
  const int i0 = kFnChunk * it.ti, j0 = kFnChunk * it.tj;
  const int nri = min(4, (mp - i0) / 16), nrj = min(4, (mp - j0) / 16);
  const bool diag = it.ti == it.tj;
  const int slen = g.slen;
  const int pb = it.slice * slen + wave * (slen / 4);
  const int pe = min(pb + slen / 4, np);
  const T* V = vals + g.val_off;
  const double* fac = factor + g.pt_off;
  A4 acc[4][4];
  d4 dacc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      acc[a][b] = A4{0, 0, 0, 0};
      dacc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    }
  int step = 0;
  for (int p = pb + 4 * lk; p < pe; p += 16, ++step) {
    const d4 fd = *reinterpret_cast<const d4*>(fac + p);
    const V4 f = V4{(T)fd[0], (T)fd[1], (T)fd[2], (T)fd[3]};
    V4 A[4], B[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      A[r] = (r < nri) ? *reinterpret_cast<const V4*>(V + (long long)(i0 + 16 * r + lrow) * np + p)
                       : V4{0, 0, 0, 0};
      if (diag)  // the same function rows on both sides: one load
        B[r] = A[r] * f;
      else
        B[r] = (r < nrj)
                   ? *reinterpret_cast<const V4*>(V + (long long)(j0 + 16 * r + lrow) * np + p) * f
                   : V4{0, 0, 0, 0};
    }
#pragma unroll
    for (int s = 0; s < 4; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (r < nri) {
#pragma unroll
          for (int c = 0; c < 4; ++c)
            if (c < nrj) Mma<T>::op(A[r][s], B[c][s], acc[r][c]);
        }
      }
    if constexpr (kFlush) {
      if ((step & 1) == 1 || p + 16 >= pe) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              dacc[r][c][i] += (double)acc[r][c][i];
              acc[r][c][i] = 0;
            }
      }
    }
  }
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int e = (16 * r + Mma<T>::row(lk, i)) * kFnChunk + 16 * c + lrow;
            const double x = kFlush ? dacc[r][c][i] : (double)acc[r][c][i];
            sF[e] = (w == 0) ? x : sF[e] + x;
          }
    }
    __syncthreads();
  }
  double* out = fpart + (long long)blockIdx.x * kFnChunk * kFnChunk;
  for (int e = threadIdx.x; e < kFnChunk * kFnChunk; e += 256) out[e] = sF[e];
}

// ---------------------------------------------------------------------------
// split-BF16 (BF16X6) contractions
// ---------------------------------------------------------------------------
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f8v __attribute__((ext_vector_type(8)));

// Eight BF16 pieces from their FP32 values (four v_cvt_pk_bf16_f32).  Every
// piece of the splits below has at most 8 significant bits, so the
// round-to-nearest conversion is exact: the upper half of the FP32 bits.
__device__ inline bf8 pack_bf8(const float (&v)[8]) {
  f8v x;
#pragma unroll
  for (int q = 0; q < 8; ++q) x[q] = v[q];
  return __builtin_convertvector(x, bf8);
}
typedef __bf16 bf4 __attribute__((ext_vector_type(4)));
typedef unsigned int u4 __attribute__((ext_vector_type(4)));
typedef float fl4 __attribute__((ext_vector_type(4)));

// Block fixed-point split of the 32 values of one K step (8 per lane, in the
// 4 lanes l, l^16, l^32, l^48 that hold the same row or point): they share
// one exponent e (2^(e-1) <= max |x| < 2^e) and each value becomes integers
// a0 + a1 2^-8 + a2 2^-16 (|a0| <= 256, |a1|, |a2| <= 128, round to nearest)
// in units of 2^(e-8), kept as the BF16 pieces a_s 2^(e-8-8s) (exact: an
// integer of magnitude at most 256 has at most 8 significant bits).  The error is below 2^-24 of the block maximum.  With
// both operands split this way every product of one MFMA output is an
// integer multiple of one unit and their sum (below 2^21 units) is exact in
// the matrix unit, whose adder otherwise truncates products that lie far
// below the largest one (measured) -- a bias on quadratic forms such as rho.
// All 64 lanes must call it (cross-lane maximum).
__device__ inline void split3(const float (&x)[8], bf8& h, bf8& m, bf8& l) {
  float mx = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) mx = fmaxf(mx, fabsf(x[q]));
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  int e;
  (void)frexpf(mx, &e);
  // Adding and removing 1.5 * 2^(e+15) rounds a value of magnitude below 2^e
  // to the nearest multiple of 2^(e-8) (ties to even): the sum lies in
  // [2^(e+15), 2^(e+16)), whose spacing is 2^(e-8).  The same with
  // 1.5 * 2^(e+7) and 1.5 * 2^(e-1) takes the remainders to multiples of
  // 2^(e-16) and 2^(e-24).  Every piece then has at most 8 significant bits,
  // so its BF16 form is the upper half of its FP32 bits.
  const float c0 = ldexpf(1.5f, e + 15), c1 = ldexpf(1.5f, e + 7), c2 = ldexpf(1.5f, e - 1);
  float vh[8], vm[8], vl[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    vh[q] = (x[q] + c0) - c0;
    const float r = x[q] - vh[q];  // exact
    vm[q] = (r + c1) - c1;
    vl[q] = ((r - vm[q]) + c2) - c2;
  }
  h = pack_bf8(vh);
  m = pack_bf8(vm);
  l = pack_bf8(vl);
}

// 16x16x32 BF16 MFMA (measured on the MI355X): lane l supplies
// A[l%16][8 (l/16) + q] and B[8 (l/16) + q][l%16], q = 0..7; the result
// D[4 (l/16) + i][l%16] is item i of lane l.  The unit sums the 32 products
// exactly for block fixed-point pieces, and rounds C + sum to nearest when C
// is the larger term; a C much smaller than the products is aligned to them
// and loses its low bits.  So the six piece products of weight >= 2^-16
// relative run largest first (hi*hi from C = 0), each later one adding
// smaller products onto a larger C.  Returns the step's tile, which callers
// add into their running sum on the VALU.
__device__ inline f4 mma6(const bf8& ah, const bf8& am, const bf8& al, const bf8& bh,
                          const bf8& bm, const bf8& bl) {
  f4 c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, f4{0, 0, 0, 0}, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bm, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bh, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(am, bm, c, 0, 0, 0);
  c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, c, 0, 0, 0);
  return c;
}

__device__ inline void load8(const float* p, float (&x)[8]) {
  const fl4 a = *reinterpret_cast<const fl4*>(p);
  const fl4 b = *reinterpret_cast<const fl4*>(p + 4);
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[q] = a[q];
    x[4 + q] = b[q];
  }
}

// Point-major FP32 basis planes [value][point][function] (functions
// contiguous) for the split-BF16 path.  A block of eight waves covers 64
// points and 32 functions of one group: each wave evaluates four functions
// (one at a time, the function record wave-uniform) for its 64 points and
// keeps their values in registers; plane by plane the values are transposed
// through a small LDS tile (two buffers, one barrier a plane) and each
// thread stores 4 consecutive functions of one point in one 16-byte store,
// so a point's row piece of a plane is a whole 128-byte line (16-function
// pieces, half lines, wrote at 3.1 TB/s against the function-major FP64
// planes' 5.7).  The tile's phi values also go to the function-major phi
// plane the Fock kernel reads (rows of 64 points).  Items of group -1 are
// padding of the XCD-interleaved item list.  The split-BF16 function padding
// is a multiple of 32, so a block never passes its group's functions.
constexpr int kPmPts = 64, kPmFns = 32;
// LO: lo16 (point-major, the planes' layout, group blocks at phi_off) gets
// phi's remainder phi - (double)(float)phi rounded to BF16 (nearest even).
__device__ inline unsigned short float_to_bf16(float f) {
  unsigned u = __builtin_bit_cast(unsigned, f);
  u += 0x7fffu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

template <int NV, bool LO = false>
__global__ __launch_bounds__(512) void k_eval_pm(const EvalItem* __restrict__ items,
                                                 const DevGroup* __restrict__ groups,
                                                 const int* __restrict__ fnl,
                                                 const DevFn* __restrict__ bas,
                                                 const double* __restrict__ px,
                                                 const double* __restrict__ py,
                                                 const double* __restrict__ pz,
                                                 float* __restrict__ vals,
                                                 float* __restrict__ phiT,
                                                 unsigned short* __restrict__ lo16) {
  constexpr int NP = NV + (LO ? 1 : 0), FW = kPmFns / 8;  // functions per wave
  __shared__ float sT[2][kPmFns][kPmPts + 1];
  const EvalItem it = items[blockIdx.x];
  if (it.group < 0) return;
  const DevGroup g = groups[it.group];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int p = it.p0 + lane;
  const int q = g.pt_off + (p < g.npts_pad ? p : it.p0);
  const double x = px[q], y = py[q], z = pz[q];
  float o[FW][NP];
#pragma unroll
  for (int j = 0; j < FW; ++j) {
    const int kk = wave + 8 * j;
    const int f = fnl[g.fn_off + it.k + kk];
    double v[NV];
    if (f >= 0) {
      eval_basis<NV>(bas[f], x, y, z, v);
    } else {
#pragma unroll
      for (int c = 0; c < NV; ++c) v[c] = 0.0;
    }
#pragma unroll
    for (int c = 0; c < NV; ++c) o[j][c] = (float)v[c];
    if (LO) o[j][NP - 1] = (float)(v[0] - (double)(float)v[0]);
    if (p < g.npts_pad) phiT[g.phi_off + (long long)(it.k + kk) * g.npts_pad + p] = o[j][0];
  }
  const int wp = threadIdx.x >> 3, wf = 4 * (threadIdx.x & 7);
  const int pp = it.p0 + wp;
  const long long plane = (long long)g.m_pad * g.npts_pad;
#pragma unroll
  for (int c = 0; c < NP; ++c) {
    const int b = c & 1;  // the buffer of plane c - 2 was read before plane c - 1's barrier
#pragma unroll
    for (int j = 0; j < FW; ++j) sT[b][wave + 8 * j][lane] = o[j][c];
    __syncthreads();
    if (pp < g.npts_pad) {
      if (!LO || c < NV) {
        float* out = vals + g.val_off + c * plane + (long long)pp * g.m_pad + it.k + wf;
        *reinterpret_cast<fl4*>(out) =
            fl4{sT[b][wf][wp], sT[b][wf + 1][wp], sT[b][wf + 2][wp], sT[b][wf + 3][wp]};
      } else {
        unsigned short* o16 = lo16 + g.phi_off + (long long)pp * g.m_pad + it.k + wf;
        const unsigned a = float_to_bf16(sT[b][wf][wp]) |
                           ((unsigned)float_to_bf16(sT[b][wf + 1][wp]) << 16);
        const unsigned d = float_to_bf16(sT[b][wf + 2][wp]) |
                           ((unsigned)float_to_bf16(sT[b][wf + 3][wp]) << 16);
        *reinterpret_cast<uint2*>(o16) = uint2{a, d};
      }
    }
  }
}

// Radial sums of every shell of a group at every point (radial-Hessian
// layout): R0 = sum_k c_k exp(-a_k r^2), R1 = -2 sum_k c_k a_k exp(..),
// R2 = 4 sum_k c_k a_k^2 exp(..), with the contraction of the shell's first
// function and primitives with a r^2 > 70 skipped (g2g/cpu/functions.cpp:
// 56-82).  One (point, shell) per thread, shells contiguous per point.
__global__ __launch_bounds__(256) void k_eval_radial(const GatherItem* __restrict__ items,
                                                     const DevGroup* __restrict__ groups,
                                                     const int* __restrict__ shells,
                                                     const DevFn* __restrict__ bas,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     float* __restrict__ rad) {
  const GatherItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const long long nplane = (long long)g.npts_pad * g.nsh_pad;
  const int ee = (int)min((long long)it.start + kEvalSpan, nplane);
  for (int e = it.start + threadIdx.x; e < ee; e += kEvalPts) {
    const int p = e / g.nsh_pad, sh = e % g.nsh_pad;
    const int f = shells[g.shell_off + sh];
    double t = 0.0, tg = 0.0, th = 0.0;
    if (f >= 0) {
      const DevFn& b = bas[f];
      const int q = g.pt_off + p;
      const double vx = px[q] - b.cx, vy = py[q] - b.cy, vz = pz[q] - b.cz;
      const double r2 = vx * vx + vy * vy + vz * vz;
      for (int k = 0; k < b.ncont; ++k) {
        const double a = b.a[k];
        if (a * r2 > 70.0) continue;
        const double ex = exp(-a * r2) * b.c[k];
        t += ex;
        tg += ex * a;
        th += ex * (a * a);
      }
    }
    float* out = rad + g.rad_off + e;
    out[0] = (float)t;
    out[nplane] = (float)(-2.0 * tg);
    out[2 * nplane] = (float)(4.0 * th);
  }
}

// The six second derivatives (xx, yy, zz, xy, xz, yz) of function meta at a
// point from its shell's radial sums and the offset v from its centre:
// x^lx y^ly z^lz g(r) with g' terms in R1 and g'' terms in R2.
__device__ inline void radial_hessian(const FnMeta& mt, float vx, float vy, float vz, float r0,
                                      float r1, float r2, float (&h)[6]) {
  const int lx = (mt.code >> 16) & 3, ly = (mt.code >> 20) & 3, lz = (mt.code >> 24) & 3;
  auto ax = [](float v, int l, float& m, float& d, float& dd) {
    m = l == 0 ? 1.f : (l == 1 ? v : v * v);
    d = l == 0 ? 0.f : (l == 1 ? 1.f : 2.f * v);
    dd = l == 2 ? 2.f : 0.f;
  };
  float Xm, Xd, Xdd, Ym, Yd, Ydd, Zm, Zd, Zdd;
  ax(vx, lx, Xm, Xd, Xdd);
  ax(vy, ly, Ym, Yd, Ydd);
  ax(vz, lz, Zm, Zd, Zdd);
  const float P = Xm * Ym * Zm;
  const float Px = Xd * Ym * Zm, Py = Xm * Yd * Zm, Pz = Xm * Ym * Zd;
  const float s = mt.scale;
  h[0] = s * (Xdd * Ym * Zm * r0 + 2.f * Px * vx * r1 + P * (vx * vx * r2 + r1));
  h[1] = s * (Xm * Ydd * Zm * r0 + 2.f * Py * vy * r1 + P * (vy * vy * r2 + r1));
  h[2] = s * (Xm * Ym * Zdd * r0 + 2.f * Pz * vz * r1 + P * (vz * vz * r2 + r1));
  h[3] = s * (Xd * Yd * Zm * r0 + (Px * vy + Py * vx) * r1 + P * vx * vy * r2);
  h[4] = s * (Xd * Ym * Zd * r0 + (Px * vz + Pz * vx) * r1 + P * vx * vz * r2);
  h[5] = s * (Xm * Yd * Zd * r0 + (Py * vz + Pz * vy) * r1 + P * vy * vz * r2);
}

// Density block P_g as NPC BF16 piece planes [piece][j][k] (true symmetric
// P), split in blocks of 32 consecutive k of one row exactly as split3 does
// (the 32 threads of a block are 32 consecutive lanes of one wave); a fourth
// piece carries the next 8 bits (units of 2^(e-32)) for the products that
// need P to 2^-32 of the block maximum.
template <int NPC>
__device__ __forceinline__ void gather_bf_item(const int bi, const GatherItem* __restrict__ items,
                                                           const DevGroup* __restrict__ groups,
                                                           const DevMol* __restrict__ mols,
                                                           const int* __restrict__ fnl,
                                                           const double* __restrict__ rho,
                                                           unsigned short* __restrict__ pgs) {
  const GatherItem it = items[bi];
  const DevGroup g = groups[it.group];
  const int e = it.start + threadIdx.x;
  const long long psz = (long long)g.m_pad * g.m_pad;
  const bool in = e < psz;  // whole 32-lane blocks are in or out (m_pad % 32 == 0)
  double v = 0.0;
  if (in) {
    const int k = e / g.m_pad, j = e % g.m_pad;
    v = density_entry(mols[g.mol], rho, fnl[g.fn_off + k], fnl[g.fn_off + j]);
  }
  double mx = fabs(v);
  for (int o = 1; o < 32; o <<= 1) mx = fmax(mx, __shfl_xor(mx, o));
  int ex;
  (void)frexp(mx, &ex);
  const double xs = ldexp(v, 8 - ex);
  const double a0 = rint(xs);
  const double r1 = (xs - a0) * 256.0;
  const double a1 = rint(r1);
  const double r2 = (r1 - a1) * 256.0;
  const double a2 = rint(r2);
  if (!in) return;
  __bf16* out = reinterpret_cast<__bf16*>(pgs) + NPC * g.pg_off + e;
  out[0] = (__bf16)(float)ldexp(a0, ex - 8);
  out[psz] = (__bf16)(float)ldexp(a1, ex - 16);
  out[2 * psz] = (__bf16)(float)ldexp(a2, ex - 24);
  if (NPC == 4) out[3 * psz] = (__bf16)(float)ldexp(rint((r2 - a2) * 256.0), ex - 32);
}

template <int NPC>
__global__ __launch_bounds__(256) void k_gather_density_bf(const GatherItem* __restrict__ items,
                                                           const DevGroup* __restrict__ groups,
                                                           const DevMol* __restrict__ mols,
                                                           const int* __restrict__ fnl,
                                                           const double* __restrict__ rho,
                                                           unsigned short* __restrict__ pgs) {
  gather_bf_item<NPC>(blockIdx.x, items, groups, mols, fnl, rho, pgs);
}

// Density + functional on split-BF16 MFMA: same blocking as k_density (64
// points per block, 16 per wave, 64-function row chunks) with K steps of 32
// functions.  The B operand of a lane is 8 consecutive functions of its point
// (32 contiguous bytes of the point-major planes), split in registers; the A
// operand is 8 consecutive columns of a P_g row, read as pre-split pieces.
// The epilogue reads the four functions a lane holds per row tile as one
// 16-byte load per plane and runs in FP64.
template <int NT>
__global__ __launch_bounds__(256, 2) void k_density_bf(const DensityItem* __restrict__ items,
                                                    const DevGroup* __restrict__ groups,
                                                    const unsigned short* __restrict__ pg3,
                                                    const float* __restrict__ vals,
                                                    const double* __restrict__ pw,
                                                    double* __restrict__ factor,
                                                    double* __restrict__ epart, int iexch) {
  __shared__ double sE[4][2];
  constexpr int NS = (NT == 4) ? 10 : 1;
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  double e_w = 0.0, n_w = 0.0;
  if (p0 < g.npts_pad) {
    const int np = g.npts_pad, mp = g.m_pad;
    const long long plane = (long long)mp * np;
    const long long psz = (long long)mp * mp;
    const float* Vp = vals + g.val_off + (long long)(p0 + lrow) * mp;  // this lane's point
    const unsigned short* Pp = pg3 + 3 * g.pg_off;
    double s[NS];
    for (int k = 0; k < NS; ++k) s[k] = 0.0;
    for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
      const int nrt = min(4, (mp - c0) / 16);
      f4 acc[4][NT];
#pragma unroll
      for (int rt = 0; rt < 4; ++rt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[rt][t] = f4{0, 0, 0, 0};
      for (int k0 = 0; k0 < mp; k0 += 32) {
        const int kk = k0 + 8 * lk;
        bf8 bh[NT], bm[NT], bl[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          float x[8];
          load8(Vp + t * plane + kk, x);
          split3(x, bh[t], bm[t], bl[t]);
        }
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          if (rt < nrt) {
            const unsigned short* pa = Pp + (long long)(c0 + 16 * rt + lrow) * mp + kk;
            const bf8 ah = *reinterpret_cast<const bf8*>(pa);
            const bf8 am = *reinterpret_cast<const bf8*>(pa + psz);
            const bf8 al = *reinterpret_cast<const bf8*>(pa + 2 * psz);
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[rt][t] += mma6(ah, am, al, bh[t], bm[t], bl[t]);
          }
        }
      }
#pragma unroll
      for (int rt = 0; rt < 4; ++rt) {
        if (rt >= nrt) continue;
        const float* vj = Vp + c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
        if constexpr (NT == 4) {
          fl4 v[10];
#pragma unroll
          for (int q = 0; q < 10; ++q) v[q] = *reinterpret_cast<const fl4*>(vj + q * plane);
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const double y0 = acc[rt][0][i], yx = acc[rt][1][i], yy = acc[rt][2][i],
                         yz = acc[rt][3][i];
            const double v1 = v[1][i], v2 = v[2][i], v3 = v[3][i];
            s[0] += (double)v[0][i] * y0;
            s[1] += v1 * y0;
            s[2] += v2 * y0;
            s[3] += v3 * y0;
            s[4] += (double)v[4][i] * y0 + v1 * yx;
            s[5] += (double)v[5][i] * y0 + v2 * yy;
            s[6] += (double)v[6][i] * y0 + v3 * yz;
            s[7] += (double)v[7][i] * y0 + v1 * yy;
            s[8] += (double)v[8][i] * y0 + v1 * yz;
            s[9] += (double)v[9][i] * y0 + v2 * yz;
          }
        } else {
          const fl4 v0 = *reinterpret_cast<const fl4*>(vj);
#pragma unroll
          for (int i = 0; i < 4; ++i) s[0] += (double)v0[i] * (double)acc[rt][0][i];
        }
      }
    }
    for (int k = 0; k < NS; ++k) {  // the 4 lane groups holding the same point
      s[k] += __shfl_xor(s[k], 16);
      s[k] += __shfl_xor(s[k], 32);
    }
    if (lk == 0) {
      const int q = g.pt_off + p0 + lrow;
      const double dens = s[0];
      double ex, ec, y2a;
      if constexpr (NT == 4) {
        const double gr[3] = {2.0 * s[1], 2.0 * s[2], 2.0 * s[3]};
        const double h1[3] = {2.0 * s[4], 2.0 * s[5], 2.0 * s[6]};
        const double h2[3] = {2.0 * s[7], 2.0 * s[8], 2.0 * s[9]};
        xc::gga_point_x(dens, gr, h1, h2, iexch, ex, ec, y2a);
      } else {
        xc::lda_point(dens, iexch, ex, ec, y2a);
      }
      const double w = pw[q];
      factor[q] = w * y2a;
      e_w = (dens * w) * (ex + ec);
      n_w = dens * w;
    }
    for (int o = 8; o >= 1; o >>= 1) {
      e_w += __shfl_xor(e_w, o);
      n_w += __shfl_xor(n_w, o);
    }
  }
  if (lane == 0) {
    sE[wave][0] = e_w;
    sE[wave][1] = n_w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    epart[2 * blockIdx.x] = ((sE[0][0] + sE[1][0]) + sE[2][0]) + sE[3][0];
    epart[2 * blockIdx.x + 1] = ((sE[0][1] + sE[1][1]) + sE[2][1]) + sE[3][1];
  }
}

// Density + functional on split-BF16 MFMA, wide form: 64 points per block
// (16 per wave) and 128-function row chunks (8 row tiles per wave), with the
// P_g pieces of a chunk and K step (ROWS rows x 32 columns x 3 pieces)
// staged once per block in LDS (double-buffered: the next step's tile is
// copied from global memory straight into LDS while this step's MFMAs run)
// instead of being read by each of the four waves; the wider chunk halves the
// re-reads of the B operand.  The B operand of the next K step (FP32 values
// of the wave's points) is loaded while this step's products run and split
// in registers; the ten row-dot sums of each point live in LDS between row
// chunks, which keeps the kernel at two waves per SIMD without spills.
// Epilogue and reductions as in k_density_bf.
// RAD: radial-Hessian layout -- phi and its gradient are planes, the six
// second derivatives are rebuilt from the shells' radial sums (FP32) and the
// functions' angular factors in the epilogue, which then reads 4 planes
// instead of 10.
// BF16 bit patterns to FP32 (exact: the upper half of the FP32 bits).
__device__ inline void bf16x8_to_float(const u4& p, float (&x)[8]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    x[2 * q] = __builtin_bit_cast(float, p[q] << 16);
    x[2 * q + 1] = __builtin_bit_cast(float, p[q] & 0xffff0000u);
  }
}
__device__ inline fl4 bf16x4_to_float(const unsigned short* p) {
  const uint2 u = *reinterpret_cast<const uint2*>(p);
  return fl4{__builtin_bit_cast(float, u.x << 16), __builtin_bit_cast(float, u.x & 0xffff0000u),
             __builtin_bit_cast(float, u.y << 16), __builtin_bit_cast(float, u.y & 0xffff0000u)};
}

__device__ inline void split4f(const float (&hi)[8], const float (&lo)[8], bf8 (&pc)[4]);
__device__ inline void mma10(const bf8 (&a)[4], const bf8 (&b)[4], f4& hi, f4& lo);

// LO: phi's FP32 rounding, which does not average out over the grid (every
// point of a radial shell sees the same value of its atom's s functions),
// is what bounds the electron count of the three-piece form; lo16 holds its
// remainder l = phi - h (h = the FP32 plane 0) as BF16 (point-major, the
// layout of the planes; eight bits of l carry phi to about 2^-33, and the
// four-piece split below needs l to 2^-8 of itself).
// LO == 1 (density variant 8): Y_0 = P_g phi from the pair (h, l) split into
// four BF16 pieces against four-piece P_g (the ten piece products of weight
// >= 2^-24, mma10), and the epilogue's phi is h + l in FP64.
// LO == 2 (density variant 9): the products stay those of h (Y = P_g h, the
// six-product chain) and the epilogue adds l to first order: with P_g
// symmetric, rho = sum_ij P_ij (h + l)_i (h + l)_j = sum_i (h_i + 2 l_i) Y_i
// + O(l^2), and the gradient sums sum_i d_a phi_i (P_g phi)_i gain
// sum_j l_j (P_g d_a h)_j = l . Y_a, all from values the lane already holds.
// Y_x, Y_y, Y_z keep the six-product chain in both forms.
// LDS of one density block (the persistent build kernel overlays it with the
// other phases' LDS).
template <int NT, int ROWS, int LO>
struct DensBf2Smem {
  u4 sA[2][(LO == 1 ? 4 : 3) * ROWS * 4];
  double sE[4][2];
  double sS[(NT == 4) ? 10 : 1][256];
};

template <int NT, int ROWS, bool RAD, int LO>
__device__ __forceinline__ void density_bf2_item(const int bi, const DensityItem* __restrict__ items,
                                                     const DevGroup* __restrict__ groups,
                                                     const unsigned short* __restrict__ pg3,
                                                     const float* __restrict__ vals,
                                                     const double* __restrict__ pw,
                                                     double* __restrict__ factor,
                                                     double* __restrict__ epart, int iexch,
                                                     int spin, double* __restrict__ factor_b,
                                                     double* __restrict__ sums, long long sstride,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const FnMeta* __restrict__ fmeta,
                                                     const float* __restrict__ rad,
                                                     float* __restrict__ ybuf,
                                                     const unsigned short* __restrict__ lo16,
    DensBf2Smem<NT, ROWS, LO>& sm) {
  constexpr int NS = (NT == 4) ? 10 : 1;
  constexpr int RT = ROWS / 16;
  constexpr int NPC = LO == 1 ? 4 : 3;       // P_g pieces
  constexpr int UNITS = NPC * ROWS * 4;      // 16-byte units of one staged tile
  auto& sA = sm.sA;                          // [buf][piece][row][k octet]
  auto& sE = sm.sE;
  auto& sS = sm.sS;
  const DensityItem it = items[bi];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;
  const int np = g.npts_pad, mp = g.m_pad;
  const long long plane = (long long)mp * np;
  const long long psz = (long long)mp * mp;
  const int pt = active ? p0 + lrow : it.p0;  // inactive waves still stage tiles
  const float* Vp = vals + g.val_off + (long long)pt * mp;  // this lane's point
  const unsigned short* Vl = lo16 + g.phi_off + (long long)pt * mp;  // phi's remainder (LO)
  const unsigned short* Pp = pg3 + NPC * g.pg_off;
  double ppx = 0.0, ppy = 0.0, ppz = 0.0;
  if constexpr (RAD) {
    ppx = px[g.pt_off + pt];
    ppy = py[g.pt_off + pt];
    ppz = pz[g.pt_off + pt];
  }
  for (int k = 0; k < NS; ++k) sS[k][threadIdx.x] = 0.0;
  // staging map: 16-byte unit c = threadIdx.x + 256 i of a tile is piece
  // c / (4 ROWS), row (c / 4) % ROWS, k octet c % 4, copied from global
  // memory straight into LDS (no staging registers): lane l of wave w writes
  // unit 256 i + 64 w + l.  Rows past the group's functions are never read
  // (their address is clamped to the last row).
  auto dma = [&](int c0, int k0, int b) {
#pragma unroll
    for (int i = 0; i < UNITS / 256; ++i) {
      const int c = threadIdx.x + 256 * i;
      const int sp = c / (4 * ROWS), row = (c >> 2) & (ROWS - 1), oc = c & 3;
      const int j = min(c0 + row, mp - 1);
      __builtin_amdgcn_global_load_lds(
          (const void*)(Pp + sp * psz + (long long)j * mp + k0 + 8 * oc),
          (__attribute__((address_space(3))) void*)&sA[b][256 * i + 64 * wave], 16, 0, 0);
    }
  };
  for (int c0 = 0; c0 < mp; c0 += ROWS) {
    const int nrt = min(RT, (mp - c0) / 16);
    f4 acc[RT][NT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[rt][t] = f4{0, 0, 0, 0};
    float xb[NT][8];
#pragma unroll
    for (int t = 0; t < NT; ++t) load8(Vp + t * plane + 8 * lk, xb[t]);
    u4 xlp = u4{0, 0, 0, 0};  // phi's remainder of the step, 8 BF16 (LO == 1), loaded ahead
    if constexpr (LO == 1) xlp = *reinterpret_cast<const u4*>(Vl + 8 * lk);
    dma(c0, 0, 0);
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < mp; k0 += 32, buf ^= 1) {
      const bool more = k0 + 32 < mp;
      if (more) dma(c0, k0 + 32, buf ^ 1);
      const int kk = k0 + 8 * lk;
      bf8 bh[NT], bm[NT], bl[NT];
      bf8 b4[LO == 1 ? 4 : 1];
      if constexpr (LO == 1) {
        float xl[8];
        bf16x8_to_float(xlp, xl);
        split4f(xb[0], xl, b4);
      }
#pragma unroll
      for (int t = LO == 1 ? 1 : 0; t < NT; ++t) split3(xb[t], bh[t], bm[t], bl[t]);
      if (more) {
#pragma unroll
        for (int t = 0; t < NT; ++t) load8(Vp + t * plane + kk + 32, xb[t]);
        if constexpr (LO == 1) xlp = *reinterpret_cast<const u4*>(Vl + kk + 32);
      }
      if (active) {
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          if (rt < nrt) {
            const int u = (16 * rt + lrow) * 4 + lk;
            const bf8 ah = __builtin_bit_cast(bf8, sA[buf][u]);
            const bf8 am = __builtin_bit_cast(bf8, sA[buf][ROWS * 4 + u]);
            const bf8 al = __builtin_bit_cast(bf8, sA[buf][2 * ROWS * 4 + u]);
            if constexpr (LO == 1) {
              const bf8 a4[4] = {ah, am, al, __builtin_bit_cast(bf8, sA[buf][3 * ROWS * 4 + u])};
              f4 hi, lo;
              mma10(a4, b4, hi, lo);
              acc[rt][0] += hi + lo;
            }
#pragma unroll
            for (int t = LO == 1 ? 1 : 0; t < NT; ++t)
              acc[rt][t] += mma6(ah, am, al, bh[t], bm[t], bl[t]);
          }
        }
      }
      __syncthreads();
    }
    if (!active) continue;
    if (ybuf) {  // Y = P_g phi of this point for the force pass (exact FP32 sums)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
        if (rt < nrt)
          *reinterpret_cast<fl4*>(ybuf + g.phi_off + (long long)pt * mp + c0 + 16 * rt + 4 * lk) =
              fl4{acc[rt][0][0], acc[rt][0][1], acc[rt][0][2], acc[rt][0][3]};
    }
    double s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = sS[k][threadIdx.x];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      if (rt >= nrt) continue;
      const float* vj = Vp + c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
      if constexpr (NT == 4) {
        fl4 v[10];
        if constexpr (RAD) {
#pragma unroll
          for (int q = 0; q < 4; ++q) v[q] = *reinterpret_cast<const fl4*>(vj + q * plane);
          const long long rplane = (long long)np * g.nsh_pad;
          const float* R = rad + g.rad_off + (long long)pt * g.nsh_pad;
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const FnMeta mt = fmeta[g.fn_off + c0 + 16 * rt + 4 * lk + i];
            const int sh = mt.code & 0xffff;
            float h[6];
            radial_hessian(mt, (float)(ppx - mt.cx), (float)(ppy - mt.cy), (float)(ppz - mt.cz),
                           R[sh], R[rplane + sh], R[2 * rplane + sh], h);
#pragma unroll
            for (int q = 0; q < 6; ++q) v[4 + q][i] = h[q];
          }
        } else {
#pragma unroll
          for (int q = 0; q < 10; ++q) v[q] = *reinterpret_cast<const fl4*>(vj + q * plane);
        }
        fl4 vlo = fl4{0, 0, 0, 0};
        if constexpr (LO != 0) vlo = bf16x4_to_float(Vl + c0 + 16 * rt + 4 * lk);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double y0 = acc[rt][0][i], yx = acc[rt][1][i], yy = acc[rt][2][i],
                       yz = acc[rt][3][i];
          const double v1 = v[1][i], v2 = v[2][i], v3 = v[3][i];
          if constexpr (LO == 2) {
            const double l = vlo[i];
            s[0] += ((double)v[0][i] + 2.0 * l) * y0;
            s[1] += v1 * y0 + l * yx;
            s[2] += v2 * y0 + l * yy;
            s[3] += v3 * y0 + l * yz;
          } else {
            s[0] += ((double)v[0][i] + (double)vlo[i]) * y0;
            s[1] += v1 * y0;
            s[2] += v2 * y0;
            s[3] += v3 * y0;
          }
          s[4] += (double)v[4][i] * y0 + v1 * yx;
          s[5] += (double)v[5][i] * y0 + v2 * yy;
          s[6] += (double)v[6][i] * y0 + v3 * yz;
          s[7] += (double)v[7][i] * y0 + v1 * yy;
          s[8] += (double)v[8][i] * y0 + v1 * yz;
          s[9] += (double)v[9][i] * y0 + v2 * yz;
        }
      } else {
        const fl4 v0 = *reinterpret_cast<const fl4*>(vj);
        fl4 vlo = fl4{0, 0, 0, 0};
        if constexpr (LO != 0) vlo = bf16x4_to_float(Vl + c0 + 16 * rt + 4 * lk);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          s[0] += ((double)v0[i] + (LO == 2 ? 2.0 : 1.0) * (double)vlo[i]) *
                  (double)acc[rt][0][i];
      }
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) sS[k][threadIdx.x] = s[k];
  }
  double e_w = 0.0, n_w = 0.0;
  if (active) {
    double s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) s[k] = sS[k][threadIdx.x];
    for (int k = 0; k < NS; ++k) {  // the 4 lane groups holding the same point
      s[k] += __shfl_xor(s[k], 16);
      s[k] += __shfl_xor(s[k], 32);
    }
    if (lk == 0)
      point_stage<NT>(s, g.pt_off + p0 + lrow, iexch, spin, pw, factor, factor_b, sums, sstride,
                      e_w, n_w);
    for (int o = 8; o >= 1; o >>= 1) {
      e_w += __shfl_xor(e_w, o);
      n_w += __shfl_xor(n_w, o);
    }
  }
  if (lane == 0) {
    sE[wave][0] = e_w;
    sE[wave][1] = n_w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    epart[2 * bi] = ((sE[0][0] + sE[1][0]) + sE[2][0]) + sE[3][0];
    epart[2 * bi + 1] = ((sE[0][1] + sE[1][1]) + sE[2][1]) + sE[3][1];
  }
}


template <int NT, int ROWS, bool RAD = false, int LO = 0>
__global__ __launch_bounds__(256, ROWS == 128 ? 1 : 2) void k_density_bf2(const DensityItem* __restrict__ items,
                                                     const DevGroup* __restrict__ groups,
                                                     const unsigned short* __restrict__ pg3,
                                                     const float* __restrict__ vals,
                                                     const double* __restrict__ pw,
                                                     double* __restrict__ factor,
                                                     double* __restrict__ epart, int iexch,
                                                     int spin, double* __restrict__ factor_b,
                                                     double* __restrict__ sums, long long sstride,
                                                     const double* __restrict__ px,
                                                     const double* __restrict__ py,
                                                     const double* __restrict__ pz,
                                                     const FnMeta* __restrict__ fmeta,
                                                     const float* __restrict__ rad,
                                                     float* __restrict__ ybuf,
                                                     const unsigned short* __restrict__ lo16) {
  __shared__ DensBf2Smem<NT, ROWS, LO> sm;
  density_bf2_item<NT, ROWS, RAD, LO>(blockIdx.x, items, groups, pg3, vals, pw, factor, epart, iexch, spin, factor_b, sums, sstride, px, py, pz, fmeta, rad, ybuf, lo16, sm);
}

// phi's BF16 pieces of density variant 11, once per grid: for every point
// and 32-function K block, phi's pair (FP32 plane, BF16 remainder) in four
// pieces (split4f), in the lanes of the density kernels (a wave holds 16
// points, lane group lk the functions 8 lk .. 8 lk + 7 of a K block), so the
// pieces are those the density kernel would make, bit for bit.  Layout per
// group: four piece planes of m_pad x npts_pad BF16 (point-major, functions
// contiguous) at 4 x phi_off.
__global__ __launch_bounds__(256) void k_split_pm(const DensityItem* __restrict__ items,
                                                  const DevGroup* __restrict__ groups,
                                                  const float* __restrict__ vals,
                                                  const unsigned short* __restrict__ lo16,
                                                  unsigned short* __restrict__ pcs) {
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  if (p0 >= g.npts_pad) return;  // whole waves (the splits exchange across lanes)
  const int np = g.npts_pad, mp = g.m_pad, pt = p0 + lrow;
  const long long plane = (long long)mp * np;
  const float* Vp = vals + g.val_off + (long long)pt * mp;
  const unsigned short* Vl = lo16 + g.phi_off + (long long)pt * mp;
  unsigned short* Pc = pcs + 4 * g.phi_off + (long long)pt * mp;
  for (int k0 = 0; k0 < mp; k0 += 32) {
    const int kk = k0 + 8 * lk;
    float x[8], xl[8];
    load8(Vp + kk, x);
    bf16x8_to_float(*reinterpret_cast<const u4*>(Vl + kk), xl);
    bf8 b4[4];
    split4f(x, xl, b4);
#pragma unroll
    for (int j = 0; j < 4; ++j)
      *reinterpret_cast<u4*>(Pc + j * plane + kk) = __builtin_bit_cast(u4, b4[j]);
  }
}

// Density row-dot sums on split-BF16 MFMA, one (point tile, row chunk) per
// block (density variant 10): the products and epilogue of variant 8 for the
// 64 functions c0 .. c0 + 63 of a 64-point tile.  The blocks of one tile are
// consecutive in their XCD's dispatch order (the list is built so), so they
// run side by side and the B operand of the tile's points (read for every
// chunk) comes from HBM once and from the XCD's L2 for the other chunks.
// Each block leaves its ten row-dot sums per point in cpart (one slot per
// chunk) and k_density_cs_finish adds them in chunk order, then evaluates
// the functional with one lane per point (the density blocks hold each point
// on four lanes, a quarter of the wave).  (A last-block-finishes form, with
// an agent-scope counter per tile, needs agent-scope fences in every block;
// on this GPU they write back and invalidate the XCD's L2, which undid the
// sharing: 9.0 ms against 1.4 ms.)
struct ChunkItem {
  int group, p0, c0;
  int tile;   // the point tile's density item: its energy slot
  int pbase;  // first cpart slot of the tile (one per chunk)
  int nch;    // chunks of the group
};

// PS (density variant 11): phi's four BF16 pieces come split once per grid
// (k_split_pm, the same pieces bit for bit), so the K loop loads them (four
// 16-byte loads a step, one step ahead) instead of the FP32 plane and the
// remainder and splits only the gradient planes.  (All thirteen B pieces
// split once per grid leave the loop with almost no VALU work but read 26
// bytes per point and function for every row chunk instead of 18, and the
// launch turned bandwidth-bound: 1.62 ms against 1.38 ms at (H2O)32.)
template <int NT, bool PS>
__global__ __launch_bounds__(256, 2) void k_density_cs(const ChunkItem* __restrict__ items,
                                                       const DevGroup* __restrict__ groups,
                                                       const unsigned short* __restrict__ pg4,
                                                       const float* __restrict__ vals,
                                                       float* __restrict__ ybuf,
                                                       const unsigned short* __restrict__ lo16,
                                                       double* __restrict__ cpart,
                                                       const unsigned short* __restrict__ pcs) {
  constexpr int NS = (NT == 4) ? 10 : 1;
  constexpr int ROWS = 64, RT = 4, NPC = 4;
  constexpr int UNITS = NPC * ROWS * 4;  // 16-byte units of one staged P_g tile
  __shared__ u4 sA[2][UNITS];            // [buf][piece][row][k octet]
  const ChunkItem it = items[blockIdx.x];
  if (it.group < 0) return;  // padding of a short XCD list
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;
  const int np = g.npts_pad, mp = g.m_pad, c0 = it.c0;
  const long long plane = (long long)mp * np;
  const long long psz = (long long)mp * mp;
  const int pt = active ? p0 + lrow : it.p0;  // inactive waves still stage tiles
  const float* Vp = vals + g.val_off + (long long)pt * mp;
  const unsigned short* Vl = lo16 + g.phi_off + (long long)pt * mp;
  const unsigned short* Pp = pg4 + NPC * g.pg_off;
  // P_g tile of a K step: 16-byte unit c = threadIdx.x + 256 i is piece
  // c / (4 ROWS), row (c / 4) % ROWS, k octet c % 4, loaded into registers
  // before the step's products and written to LDS after them.  (Copied
  // straight into LDS instead, the copy's completion has to be awaited
  // before the next LDS instruction -- the split's cross-lane exchange --
  // which left its latency exposed at every step.)  Rows past the group's
  // functions are never read (their address is clamped to the last row).
  u4 pa[UNITS / 256];
  auto gload = [&](int k0) {
#pragma unroll
    for (int i = 0; i < UNITS / 256; ++i) {
      const int c = threadIdx.x + 256 * i;
      const int sp = c / (4 * ROWS), row = (c >> 2) & (ROWS - 1), oc = c & 3;
      const int j = min(c0 + row, mp - 1);
      pa[i] = *reinterpret_cast<const u4*>(Pp + sp * psz + (long long)j * mp + k0 + 8 * oc);
    }
  };
  auto lstore = [&](int b) {
#pragma unroll
    for (int i = 0; i < UNITS / 256; ++i) sA[b][threadIdx.x + 256 * i] = pa[i];
  };
  const int nrt = min(RT, (mp - c0) / 16);
  f4 acc[RT][NT];
#pragma unroll
  for (int rt = 0; rt < RT; ++rt)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[rt][t] = f4{0, 0, 0, 0};
  const unsigned short* Pc = pcs + 4 * g.phi_off + (long long)pt * mp;  // PS
  float xb[NT][8];      // the next step's FP32 values (plane 0 unused with PS)
  u4 xlp = u4{0, 0, 0, 0};
  u4 nb[PS ? 4 : 1];    // the next step's phi pieces (PS)
  if constexpr (PS) {
#pragma unroll
    for (int j = 0; j < 4; ++j) nb[j] = *reinterpret_cast<const u4*>(Pc + j * plane + 8 * lk);
  } else {
    load8(Vp + 8 * lk, xb[0]);
    xlp = *reinterpret_cast<const u4*>(Vl + 8 * lk);
  }
#pragma unroll
  for (int t = 1; t < NT; ++t) load8(Vp + t * plane + 8 * lk, xb[t]);
  gload(0);
  lstore(0);
  __syncthreads();
  for (int k0 = 0, buf = 0; k0 < mp; k0 += 32, buf ^= 1) {
    const bool more = k0 + 32 < mp;
    const int kk = k0 + 8 * lk;
    bf8 bh[NT], bm[NT], bl[NT];
    bf8 b4[4];
    if constexpr (PS) {
#pragma unroll
      for (int j = 0; j < 4; ++j) b4[j] = __builtin_bit_cast(bf8, nb[j]);
    } else {
      float xl[8];
      bf16x8_to_float(xlp, xl);
      split4f(xb[0], xl, b4);
    }
#pragma unroll
    for (int t = 1; t < NT; ++t) split3(xb[t], bh[t], bm[t], bl[t]);
    if (more) {
      gload(k0 + 32);
      if constexpr (PS) {
#pragma unroll
        for (int j = 0; j < 4; ++j) nb[j] = *reinterpret_cast<const u4*>(Pc + j * plane + kk + 32);
      } else {
        load8(Vp + kk + 32, xb[0]);
        xlp = *reinterpret_cast<const u4*>(Vl + kk + 32);
      }
#pragma unroll
      for (int t = 1; t < NT; ++t) load8(Vp + t * plane + kk + 32, xb[t]);
    }
    if (active) {
#pragma unroll
      for (int rt = 0; rt < RT; ++rt) {
        if (rt < nrt) {
          const int u = (16 * rt + lrow) * 4 + lk;
          const bf8 ah = __builtin_bit_cast(bf8, sA[buf][u]);
          const bf8 am = __builtin_bit_cast(bf8, sA[buf][ROWS * 4 + u]);
          const bf8 al = __builtin_bit_cast(bf8, sA[buf][2 * ROWS * 4 + u]);
          const bf8 a4[4] = {ah, am, al, __builtin_bit_cast(bf8, sA[buf][3 * ROWS * 4 + u])};
          f4 hi, lo;
          mma10(a4, b4, hi, lo);
          acc[rt][0] += hi + lo;
#pragma unroll
          for (int t = 1; t < NT; ++t) acc[rt][t] += mma6(ah, am, al, bh[t], bm[t], bl[t]);
        }
      }
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
  }
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = 0.0;
  if (active) {
    if (ybuf) {  // Y = P_g phi of this point for the force pass (exact FP32 sums)
#pragma unroll
      for (int rt = 0; rt < RT; ++rt)
        if (rt < nrt)
          *reinterpret_cast<fl4*>(ybuf + g.phi_off + (long long)pt * mp + c0 + 16 * rt + 4 * lk) =
              fl4{acc[rt][0][0], acc[rt][0][1], acc[rt][0][2], acc[rt][0][3]};
    }
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
      if (rt >= nrt) continue;
      const float* vj = Vp + c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
      const fl4 vlo = bf16x4_to_float(Vl + c0 + 16 * rt + 4 * lk);
      if constexpr (NT == 4) {
        fl4 v[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) v[q] = *reinterpret_cast<const fl4*>(vj + q * plane);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const double y0 = acc[rt][0][i], yx = acc[rt][1][i], yy = acc[rt][2][i],
                       yz = acc[rt][3][i];
          const double v1 = v[1][i], v2 = v[2][i], v3 = v[3][i];
          s[0] += ((double)v[0][i] + (double)vlo[i]) * y0;
          s[1] += v1 * y0;
          s[2] += v2 * y0;
          s[3] += v3 * y0;
          s[4] += (double)v[4][i] * y0 + v1 * yx;
          s[5] += (double)v[5][i] * y0 + v2 * yy;
          s[6] += (double)v[6][i] * y0 + v3 * yz;
          s[7] += (double)v[7][i] * y0 + v1 * yy;
          s[8] += (double)v[8][i] * y0 + v1 * yz;
          s[9] += (double)v[9][i] * y0 + v2 * yz;
        }
      } else {
        const fl4 v0 = *reinterpret_cast<const fl4*>(vj);
#pragma unroll
        for (int i = 0; i < 4; ++i) s[0] += ((double)v0[i] + (double)vlo[i]) * (double)acc[rt][0][i];
      }
    }
    for (int k = 0; k < NS; ++k) {  // the 4 lane groups holding the same point
      s[k] += __shfl_xor(s[k], 16);
      s[k] += __shfl_xor(s[k], 32);
    }
  }
  if (active && lk == 0) {
    double* my = cpart + (long long)(it.pbase + c0 / ROWS) * NS * kDensPts + 16 * wave + lrow;
#pragma unroll
    for (int k = 0; k < NS; ++k) my[k * kDensPts] = s[k];
  }
}

// Density variant 10, after k_density_cs: one wave per point tile, lane l
// its point l; the chunks' row-dot sums added in chunk order, then the
// functional and the tile's energy partials in the order of the one-block
// form (16-point groups by a butterfly, then the groups in order).
template <int NT>
__global__ __launch_bounds__(256) void k_density_cs_finish(const ChunkItem* __restrict__ fin, int nfin,
                                                           const DevGroup* __restrict__ groups,
                                                           const double* __restrict__ pw,
                                                           double* __restrict__ factor,
                                                           double* __restrict__ epart, int iexch,
                                                           int spin, double* __restrict__ factor_b,
                                                           double* __restrict__ sums,
                                                           long long sstride,
                                                           const double* __restrict__ cpart) {
  constexpr int NS = (NT == 4) ? 10 : 1;
  const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  if (w >= nfin) return;  // whole waves
  const ChunkItem it = fin[w];
  const DevGroup g = groups[it.group];
  const bool active = it.p0 + lane < g.npts_pad;
  double e_w = 0.0, n_w = 0.0;
  if (active) {
    const double* ps = cpart + (long long)it.pbase * NS * kDensPts + lane;
    double s[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
      double t = ps[k * kDensPts];
      for (int c = 1; c < it.nch; ++c) t += ps[((long long)c * NS + k) * kDensPts];
      s[k] = t;
    }
    point_stage<NT>(s, g.pt_off + it.p0 + lane, iexch, spin, pw, factor, factor_b, sums, sstride,
                    e_w, n_w);
  }
  for (int o = 8; o >= 1; o >>= 1) {
    e_w += __shfl_xor(e_w, o);
    n_w += __shfl_xor(n_w, o);
  }
  const double e1 = __shfl(e_w, 16), e2 = __shfl(e_w, 32), e3 = __shfl(e_w, 48);
  const double n1 = __shfl(n_w, 16), n2 = __shfl(n_w, 32), n3 = __shfl(n_w, 48);
  if (lane == 0) {
    epart[2 * it.tile] = ((e_w + e1) + e2) + e3;
    epart[2 * it.tile + 1] = ((n_w + n1) + n2) + n3;
  }
}

// Density + functional on split-BF16 MFMA, row-split form for groups of
// more than 64 functions: a block owns 16 points and its four waves split the
// function row tiles (wave w takes tiles w, w + 4, ..., RTW of them a pass).
// The B operand of the block's points is read from HBM and split into BF16
// pieces once per K step (wave t stages derivative t) and shared by the four
// waves through LDS (double-buffered); each wave reads the P_g pieces of its
// own rows, which stay L2-resident.  The row-dot partials of the waves are
// added in wave order before the functional.
template <int NT, int RTW>
__global__ __launch_bounds__(256, 2) void k_density_bf3(const DensityItem* __restrict__ items,
                                                     const DevGroup* __restrict__ groups,
                                                     const unsigned short* __restrict__ pg3,
                                                     const float* __restrict__ vals,
                                                     const double* __restrict__ pw,
                                                     double* __restrict__ factor,
                                                     double* __restrict__ epart, int iexch) {
  constexpr int NS = (NT == 4) ? 10 : 1;
  constexpr int UNITS = NT * 3 * 64;  // 16-byte units of one staged B step
  __shared__ u4 sB[2][UNITS];         // [buf][t][piece][point][k octet]
  __shared__ double sPart[4][NS][16];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0;
  const int np = g.npts_pad, mp = g.m_pad, nrt = mp / 16;
  const long long plane = (long long)mp * np;
  const long long psz = (long long)mp * mp;
  const float* Vp = vals + g.val_off + (long long)(p0 + lrow) * mp;  // lane's point
  const unsigned short* Pp = pg3 + 3 * g.pg_off;
  const bool stager = wave < NT;  // wave t splits derivative t
  double s[NS];
#pragma unroll
  for (int k = 0; k < NS; ++k) s[k] = 0.0;
  auto stage = [&](int buf, const float (&x)[8]) {
    bf8 h, m, l;
    split3(x, h, m, l);
    const int u = (3 * wave) * 64 + lrow * 4 + lk;
    sB[buf][u] = __builtin_bit_cast(u4, h);
    sB[buf][u + 64] = __builtin_bit_cast(u4, m);
    sB[buf][u + 128] = __builtin_bit_cast(u4, l);
  };
  const int npass = (nrt + 4 * RTW - 1) / (4 * RTW);
  for (int pass = 0; pass < npass; ++pass) {
    const int rbase = pass * 4 * RTW;
    f4 acc[RTW][NT];
#pragma unroll
    for (int i = 0; i < RTW; ++i)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[i][t] = f4{0, 0, 0, 0};
    float x[8];
    if (stager) {
      load8(Vp + wave * plane + 8 * lk, x);
      stage(0, x);
    }
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < mp; k0 += 32, buf ^= 1) {
      const bool more = k0 + 32 < mp;
      if (stager && more) load8(Vp + wave * plane + k0 + 32 + 8 * lk, x);
      bf8 bh[NT], bm[NT], bl[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int u = (3 * t) * 64 + lrow * 4 + lk;
        bh[t] = __builtin_bit_cast(bf8, sB[buf][u]);
        bm[t] = __builtin_bit_cast(bf8, sB[buf][u + 64]);
        bl[t] = __builtin_bit_cast(bf8, sB[buf][u + 128]);
      }
#pragma unroll
      for (int i = 0; i < RTW; ++i) {
        const int r = rbase + wave + 4 * i;
        if (r < nrt) {
          const unsigned short* pa = Pp + (long long)(16 * r + lrow) * mp + k0 + 8 * lk;
          const bf8 ah = *reinterpret_cast<const bf8*>(pa);
          const bf8 am = *reinterpret_cast<const bf8*>(pa + psz);
          const bf8 al = *reinterpret_cast<const bf8*>(pa + 2 * psz);
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[i][t] += mma6(ah, am, al, bh[t], bm[t], bl[t]);
        }
      }
      if (stager && more) stage(buf ^ 1, x);
      __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < RTW; ++i) {
      const int r = rbase + wave + 4 * i;
      if (r >= nrt) continue;
      const float* vj = Vp + 16 * r + 4 * lk;  // functions j .. j+3 of this point
      if constexpr (NT == 4) {
        fl4 v[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) v[q] = *reinterpret_cast<const fl4*>(vj + q * plane);
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) {
          const double y0 = acc[i][0][ii], yx = acc[i][1][ii], yy = acc[i][2][ii],
                       yz = acc[i][3][ii];
          const double v1 = v[1][ii], v2 = v[2][ii], v3 = v[3][ii];
          s[0] += (double)v[0][ii] * y0;
          s[1] += v1 * y0;
          s[2] += v2 * y0;
          s[3] += v3 * y0;
          s[4] += (double)v[4][ii] * y0 + v1 * yx;
          s[5] += (double)v[5][ii] * y0 + v2 * yy;
          s[6] += (double)v[6][ii] * y0 + v3 * yz;
          s[7] += (double)v[7][ii] * y0 + v1 * yy;
          s[8] += (double)v[8][ii] * y0 + v1 * yz;
          s[9] += (double)v[9][ii] * y0 + v2 * yz;
        }
      } else {
        const fl4 v0 = *reinterpret_cast<const fl4*>(vj);
#pragma unroll
        for (int ii = 0; ii < 4; ++ii) s[0] += (double)v0[ii] * (double)acc[i][0][ii];
      }
    }
  }
#pragma unroll
  for (int k = 0; k < NS; ++k) {  // the 4 lane groups holding the same point
    s[k] += __shfl_xor(s[k], 16);
    s[k] += __shfl_xor(s[k], 32);
  }
  if (lk == 0) {
#pragma unroll
    for (int k = 0; k < NS; ++k) sPart[wave][k][lrow] = s[k];
  }
  __syncthreads();
  if (wave == 0) {
    double e_w = 0.0, n_w = 0.0;
    if (lk == 0) {
      double t[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k)
        t[k] = ((sPart[0][k][lrow] + sPart[1][k][lrow]) + sPart[2][k][lrow]) + sPart[3][k][lrow];
      const double dens = t[0];
      double ex, ec, y2a;
      if constexpr (NT == 4) {
        const double gr[3] = {2.0 * t[1], 2.0 * t[2], 2.0 * t[3]};
        const double h1[3] = {2.0 * t[4], 2.0 * t[5], 2.0 * t[6]};
        const double h2[3] = {2.0 * t[7], 2.0 * t[8], 2.0 * t[9]};
        xc::gga_point_x(dens, gr, h1, h2, iexch, ex, ec, y2a);
      } else {
        xc::lda_point(dens, iexch, ex, ec, y2a);
      }
      const int q = g.pt_off + p0 + lrow;
      const double w = pw[q];
      factor[q] = w * y2a;
      e_w = (dens * w) * (ex + ec);
      n_w = dens * w;
    }
    for (int o = 8; o >= 1; o >>= 1) {
      e_w += __shfl_xor(e_w, o);
      n_w += __shfl_xor(n_w, o);
    }
    if (lane == 0) {
      epart[2 * blockIdx.x] = e_w;
      epart[2 * blockIdx.x + 1] = n_w;
    }
  }
}

// Fock on split-BF16 MFMA, same blocking as k_fock (4 waves x 128 points of a
// 512-point slice, 64 x 64 tile).  K steps of 32 points: a lane's operand is
// 8 consecutive points of one function row of the function-major phi plane,
// times the factor for the B side (rounded to FP32 from the FP64 product),
// split in registers.  The next K step's rows and factors are loaded while
// this step's products run: the B side is split first, then each A row just
// before its products, so the raw values of both steps fit in registers
// (253 VGPRs, two waves per SIMD).  Each wave sums its 128 points in FP32 in the matrix
// unit; the four partial tiles are added in FP64 in wave order.
__device__ __forceinline__ void fock_bf_item(const int bi, const FockItem* __restrict__ items,
                                                  const DevGroup* __restrict__ groups,
                                                  const float* __restrict__ phiT,
                                                  const double* __restrict__ factor,
                                                  double* __restrict__ fpart,
                                             double (&sm)[kFnChunk * kFnChunk]) {
  double* sF = sm;
  const FockItem it = items[bi];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lrow = lane & 15, lk = lane >> 4;
  const int np = g.npts_pad, mp = g.m_pad;
  const int i0 = kFnChunk * it.ti, j0 = kFnChunk * it.tj;
  const int nri = min(4, (mp - i0) / 16), nrj = min(4, (mp - j0) / 16);
  const bool diag = it.ti == it.tj;
  const int slen = g.slen;
  const int pb = it.slice * slen + wave * (slen / 4);
  const int pe = min(pb + slen / 4, np);
  const float* Phi = phiT + g.phi_off;
  const double* fac = factor + g.pt_off;
  f4 acc[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = f4{0, 0, 0, 0};
  // raw values of the current K step, loaded one step ahead
  float xa[4][8], xb[4][8];
  double f[8];
  auto load_fb = [&](int p) {
    const bool ok = p < pe;
    if (ok) {
      const d4 f0 = *reinterpret_cast<const d4*>(fac + p);
      const d4 f1 = *reinterpret_cast<const d4*>(fac + p + 4);
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        f[q] = f0[q];
        f[4 + q] = f1[q];
      }
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) f[q] = 0.0;
    }
    if (!diag) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if (ok && r < nrj) {
          load8(Phi + (long long)(j0 + 16 * r + lrow) * np + p, xb[r]);
        } else {
#pragma unroll
          for (int q = 0; q < 8; ++q) xb[r][q] = 0.f;
        }
      }
    }
  };
  auto load_a = [&](int p, int r) {
    if (p < pe && r < nri) {
      load8(Phi + (long long)(i0 + 16 * r + lrow) * np + p, xa[r]);
    } else {
#pragma unroll
      for (int q = 0; q < 8; ++q) xa[r][q] = 0.f;
    }
  };
  load_fb(pb + 8 * lk);
#pragma unroll
  for (int r = 0; r < 4; ++r) load_a(pb + 8 * lk, r);
  for (int p = pb + 8 * lk; p - 8 * lk < pe; p += 32) {
    const bool ok = p < pe;
    const bool next = p + 32 - 8 * lk < pe;
    bf8 bh[4], bm[4], bl[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float x[8];
      if (ok && r < nrj) {
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = (float)((double)(diag ? xa[r][q] : xb[r][q]) * f[q]);
      } else {
#pragma unroll
        for (int q = 0; q < 8; ++q) x[q] = 0.f;
      }
      split3(x, bh[r], bm[r], bl[r]);
    }
    if (next) load_fb(p + 32);
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      bf8 ah, am, al;
      split3(xa[r], ah, am, al);
      if (next) load_a(p + 32, r);
      if (r < nri) {
#pragma unroll
        for (int c = 0; c < 4; ++c)
          if (c < nrj) acc[r][c] += mma6(ah, am, al, bh[c], bm[c], bl[c]);
      }
    }
  }
  for (int w = 0; w < 4; ++w) {
    if (wave == w) {
#pragma unroll
      for (int r = 0; r < 4; ++r)
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int e = (16 * r + 4 * lk + i) * kFnChunk + 16 * c + lrow;
            const double x = (double)acc[r][c][i];
            sF[e] = (w == 0) ? x : sF[e] + x;
          }
    }
    __syncthreads();
  }
  double* out = fpart + (long long)bi * kFnChunk * kFnChunk;
  for (int e = threadIdx.x; e < kFnChunk * kFnChunk; e += 256) out[e] = sF[e];
}

// Block b takes item order[b] (its partial keeps the item's slot): the tile
// pairs of one group and slice are consecutive in one XCD's dispatch order,
// so the function rows they share come from that XCD's L2; -1 pads a short
// XCD list.
__global__ __launch_bounds__(256, 2) void k_fock_bf(const FockItem* __restrict__ items,
                                                  const DevGroup* __restrict__ groups,
                                                  const float* __restrict__ phiT,
                                                  const double* __restrict__ factor,
                                                  double* __restrict__ fpart,
                                                  const int* __restrict__ order) {
  __shared__ double sF[kFnChunk * kFnChunk];
  const int it = order[blockIdx.x];
  if (it < 0) return;
  fock_bf_item(it, items, groups, phiT, factor, fpart, sF);
}

// ---------------------------------------------------------------------------
// split-BF16 density with the basis evaluated in the kernel (density variant 6)
// ---------------------------------------------------------------------------
// No value planes are read from HBM.  A block of 64 points (16 per wave)
// evaluates phi, its gradient and its Hessian from per-shell radial sums
// (sum c e, sum c a e, sum c a^2 e with e = exp(-a r^2)) that each wave
// computes in FP64 for its own points and keeps in LDS: once per K step for
// the K operand and once per row tile for the row-dot epilogue.  The group
// metadata says which shell every function belongs to (FnMeta) and which
// shells a K step or row tile touches (kbsh, tsh).
//
// Precision.  Y_0 = P_g phi gives rho and its gradient, hence Exc and N.
// Its K operand is split from FP64 values into four BF16 pieces (block
// fixed point, units down to 2^-32 of the block maximum), P_g has four
// pieces too, and the tile product is the ten piece products of weight
// >= 2^-24 relative; the epilogue multiplies by FP64 values.  Single-precision
// operands instead leave a rounding error that does not average out: on a
// Becke/Lebedev grid all points of a radial shell see the same value of the
// atom's own s functions, so their rounding repeats over every point of the
// shell (and over every atom of the same element) and moved N by 1e-6 at
// (H2O)32.  Y_a = P_g d_a phi (a = x, y, z) only enter the second-derivative
// terms of the potential; they keep three-piece operands (the FP64 gradient
// rounded to FP32) and the six-product chain.
//
// Schedule per K step: a single LDS buffer holds the step's P_g tile (four
// pieces, copied straight from global memory); the tile and the radial sums
// of the next step are produced while the block waits at the barrier that
// ends this step, so each step runs [radial sums of K shells, DMA in flight]
// -> barrier -> [K operand, MFMAs] -> barrier.  Two blocks share a CU.


__device__ inline void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block fixed-point split of an FP32 pair (hi + lo, |lo| <= ulp(hi) / 2: a
// double rounded to FP32 and its remainder; a lane's 8 K values of one
// point, the same 4-lane block as split3) into four BF16 pieces in units
// 2^(e-8), 2^(e-16), 2^(e-24), 2^(e-32), 2^(e-1) <= max |hi| < 2^e: the
// first three pieces come from hi as in split3, the fourth from hi's
// remainder plus lo (their FP32 sum is off by at most 2^-48 of the block
// maximum, far below the piece's unit 2^(e-32)), so the pieces are those of
// the double hi + lo up to that rounding.  All 64 lanes must call it.
__device__ inline void split4f(const float (&hi)[8], const float (&lo)[8], bf8 (&pc)[4]) {
  float mx = 0.f;
#pragma unroll
  for (int q = 0; q < 8; ++q) mx = fmaxf(mx, fabsf(hi[q]));
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  int e;
  (void)frexpf(mx, &e);
  const float c0 = ldexpf(1.5f, e + 15), c1 = ldexpf(1.5f, e + 7), c2 = ldexpf(1.5f, e - 1),
              c3 = ldexpf(1.5f, e - 9);
  float v[4][8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[0][q] = (hi[q] + c0) - c0;
    float r = hi[q] - v[0][q];
    v[1][q] = (r + c1) - c1;
    r -= v[1][q];
    v[2][q] = (r + c2) - c2;
    r = (r - v[2][q]) + lo[q];
    v[3][q] = (r + c3) - c3;
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) pc[k] = pack_bf8(v[k]);
}

// The ten piece products of weight >= 2^-24 of four-piece operands, as two
// chains from zero, largest first: hi = the products of weight 1 and 2^-8,
// lo = those of weight 2^-16 and 2^-24 (each chain keeps 24 bits of its own
// scale, so the small products are not rounded away against the large ones).
__device__ inline void mma10(const bf8 (&a)[4], const bf8 (&b)[4], f4& hi, f4& lo) {
  hi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], f4{0, 0, 0, 0}, 0, 0, 0);
  hi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], hi, 0, 0, 0);
  hi = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], hi, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], f4{0, 0, 0, 0}, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[3], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[2], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[1], lo, 0, 0, 0);
  lo = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3], b[0], lo, 0, 0, 0);
}

// Radial sums of shells [r.x, r.x + r.y) of group g (at most ST / 16) at
// the wave's 16 points into the wave's LDS table rw: t at [s][p], tg at
// ST + [s][p], th (NV = 10) at 2 ST + [s][p] (16 doubles per shell row).
// Lane (lrow, lk) takes shells lk, lk + 4, ... at point lrow.
template <int NV, int ST>
__device__ inline void radial_table(const DevGroup& g, int2 r, const int* __restrict__ shells,
                                    const DevFn* __restrict__ bas, double ppx, double ppy,
                                    double ppz, int lrow, int lk, double* rw) {
  for (int sh = lk; sh < r.y; sh += 4) {
    const DevFn& b = bas[shells[g.shell_off + r.x + sh]];
    const double vx = ppx - b.cx, vy = ppy - b.cy, vz = ppz - b.cz;
    double t, tg, th;
    radial_sums<NV>(b, vx * vx + vy * vy + vz * vz, t, tg, th);
    rw[sh * 16 + lrow] = t;
    if (NV > 1) rw[ST + sh * 16 + lrow] = tg;
    if (NV > 4) rw[2 * ST + sh * 16 + lrow] = th;
  }
}

// Values of local function jl of group g at the lane's point from the radial
// table whose first shell is sh0 (zeros for padding functions).
template <int NV, int ST>
__device__ inline void basis_from_table(const DevGroup& g, const FnMeta* __restrict__ fmeta,
                                        int jl, int sh0, double ppx, double ppy, double ppz,
                                        int lrow, const double* rw, double (&o)[NV]) {
  if (jl >= g.m) {
#pragma unroll
    for (int k = 0; k < NV; ++k) o[k] = 0.0;
    return;
  }
  const FnMeta mt = fmeta[g.fn_off + jl];
  const int sl = ((mt.code & 0xffff) - sh0) * 16 + lrow;
  const double t = rw[sl];
  const double tg = NV > 1 ? rw[ST + sl] : 0.0;
  const double th = NV > 4 ? rw[2 * ST + sl] : 0.0;
  basis_from_radial<NV>((mt.code >> 16) & 3, (mt.code >> 20) & 3, (mt.code >> 24) & 3,
                        ppx - mt.cx, ppy - mt.cy, ppz - mt.cz, t, tg, th, o);
#pragma unroll
  for (int k = 0; k < NV; ++k) o[k] *= mt.scale;
}

// LDS layout of the K operand values of one K step (32 functions at the
// block's 64 points), written by the waves that evaluate them and read by the
// MFMA lanes, [plane][point][function] in FP32 with rows of 36 floats (the 16
// lanes of a lane group -- 16 consecutive points, 8 consecutive functions
// each -- then hit distinct banks): planes 0 and 1 are phi as an FP32 pair
// (the FP64 value rounded, and its remainder), planes 2-4 the gradient.
constexpr int kEvRow = 36, kEvPlane = 64 * kEvRow;

template <int NT, int ROWS>
__global__ __launch_bounds__(256, 2) void k_density_ev(const DensityItem* __restrict__ items,
                                                    const DevGroup* __restrict__ groups,
                                                    const unsigned short* __restrict__ pg4,
                                                    const FnMeta* __restrict__ fmeta,
                                                    const int* __restrict__ shells,
                                                    const int2* __restrict__ sfn,
                                                    const int2* __restrict__ kbsh,
                                                    const int2* __restrict__ tsh,
                                                    const DevFn* __restrict__ bas,
                                                    const double* __restrict__ px,
                                                    const double* __restrict__ py,
                                                    const double* __restrict__ pz,
                                                    const double* __restrict__ pw,
                                                    double* __restrict__ factor,
                                                    double* __restrict__ epart, int iexch,
                                                    int spin, double* __restrict__ factor_b,
                                                    double* __restrict__ sums, long long sstride,
                                                    float* __restrict__ ybuf) {
  constexpr int NS = (NT == 4) ? 10 : 1;
  constexpr int RT = ROWS / 16;
  constexpr int UNITS = 4 * ROWS * 4;  // 16-byte units of one staged P_g tile
  constexpr int VBYTES = (NT == 4 ? 5 : 2) * kEvPlane * 4;
  constexpr int RWS = (NT == 4) ? 768 : 256;  // doubles of one wave's epilogue radial table
  static_assert(VBYTES >= 4 * RWS * 8, "the epilogue's radial tables reuse the value area");
  __shared__ u4 sA[UNITS];                      // [piece][row][k octet]
  __shared__ __attribute__((aligned(16))) char sV[VBYTES];  // K values, or radial tables
  __shared__ double sE[4][2];
  __shared__ double sS[NS][64];  // row-dot sums of the block's points between row chunks
  float* sK = reinterpret_cast<float*>(sV);  // [plane][point][kEvRow]
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;  // uniform over the wave
  const int mp = g.m_pad, nkb = mp >> 5, nchunk = (mp + ROWS - 1) / ROWS;
  const long long psz = (long long)mp * mp;
  // this lane's MFMA point and its evaluation point (block point `lane`);
  // their coordinates are read where they are needed
  const int pt = active ? p0 + lrow : it.p0;
  const int q = g.pt_off + pt;
  const int qe = g.pt_off + min(it.p0 + lane, g.npts_pad - 1);
  const unsigned short* Pp = pg4 + 4 * g.pg_off;
  double* rw = reinterpret_cast<double*>(sV) + RWS * wave;  // epilogue radial table
  // unit c = threadIdx.x + 256 i of a tile is piece c / (4 ROWS), row
  // (c / 4) % ROWS, k octet c % 4 (rows past the functions are clamped and
  // never read)
  auto dma = [&](int c0, int k0) {
#pragma unroll
    for (int i = 0; i < UNITS / 256; ++i) {
      const int c = threadIdx.x + 256 * i;
      const int sp = c / (4 * ROWS), row = (c >> 2) & (ROWS - 1), oc = c & 3;
      const int j = min(c0 + row, mp - 1);
      __builtin_amdgcn_global_load_lds(
          (const void*)(Pp + sp * psz + (long long)j * mp + k0 + 8 * oc),
          (__attribute__((address_space(3))) void*)&sA[256 * i + 64 * wave], 16, 0, 0);
    }
  };
  // K values of K block kb at the block's 64 points: wave w evaluates shells
  // w, w + 4, ... of the block (one point per lane, so the shell, its
  // primitives and its functions are the same over the wave); phi in FP64
  // (kept as an FP32 pair), the gradient in FP32 from the FP64 radial sums.
  // Padding functions are zero.
  auto eval_k = [&](int kb) {
    const int2 kr = kbsh[(g.fn_off >> 5) + kb];
    const int k0 = 32 * kb;
    float* row = sK + lane * kEvRow - k0;
    const double epx = px[qe], epy = py[qe], epz = pz[qe];
    for (int i = wave; i < kr.y; i += 4) {
      const int sh = g.shell_off + kr.x + i;
      const DevFn& b = bas[shells[sh]];
      const double vx = epx - b.cx, vy = epy - b.cy, vz = epz - b.cz;
      double t, tg, th;
      radial_sums<(NT == 4 ? 4 : 1)>(b, vx * vx + vy * vy + vz * vz, t, tg, th);
      const float fx = (float)vx, fy = (float)vy, fz = (float)vz;
      const float ft = (float)t, fg = (float)(-2.0 * tg);
      const int2 fr = sfn[sh];
      const int jb = max(fr.x, k0), je = min(fr.x + fr.y, k0 + 32);
      for (int jl = jb; jl < je; ++jl) {
        const FnMeta mt = fmeta[g.fn_off + jl];
        const int lx = (mt.code >> 16) & 3, ly = (mt.code >> 20) & 3, lz = (mt.code >> 24) & 3;
        const double Pd = (lx == 0 ? 1.0 : lx == 1 ? vx : vx * vx) *
                          (ly == 0 ? 1.0 : ly == 1 ? vy : vy * vy) *
                          (lz == 0 ? 1.0 : lz == 1 ? vz : vz * vz);
        const double ph = mt.scale * (Pd * t);
        const float hi = (float)ph;
        row[jl] = hi;
        row[kEvPlane + jl] = (float)(ph - (double)hi);
        if constexpr (NT == 4) {
          const float Xm = lx == 0 ? 1.f : lx == 1 ? fx : fx * fx;
          const float Ym = ly == 0 ? 1.f : ly == 1 ? fy : fy * fy;
          const float Zm = lz == 0 ? 1.f : lz == 1 ? fz : fz * fz;
          const float Xd = lx == 0 ? 0.f : lx == 1 ? 1.f : 2.f * fx;
          const float Yd = ly == 0 ? 0.f : ly == 1 ? 1.f : 2.f * fy;
          const float Zd = lz == 0 ? 0.f : lz == 1 ? 1.f : 2.f * fz;
          const float fs = (float)mt.scale, Pf = Xm * Ym * Zm * fg;
          row[2 * kEvPlane + jl] = fs * (Xd * Ym * Zm * ft + Pf * fx);
          row[3 * kEvPlane + jl] = fs * (Xm * Yd * Zm * ft + Pf * fy);
          row[4 * kEvPlane + jl] = fs * (Xm * Ym * Zd * ft + Pf * fz);
        }
      }
    }
    // padding functions of the block (past the group's last function)
    if (wave == 0)
      for (int jl = max(g.m, k0); jl < k0 + 32; ++jl)
#pragma unroll
        for (int pl = 0; pl < (NT == 4 ? 5 : 2); ++pl) row[pl * kEvPlane + jl] = 0.f;
  };
  if (threadIdx.x < 64)
#pragma unroll
    for (int k = 0; k < NS; ++k) sS[k][threadIdx.x] = 0.0;
  dma(0, 0);
  eval_k(0);
  for (int c = 0; c < nchunk; ++c) {
    const int c0 = c * ROWS, nrt = min(RT, (mp - c0) >> 4);
    f4 acc[RT][NT];  // plane 0 unused: Y_0 is summed in FP64 (y0s)
    d4 y0s[RT];
#pragma unroll
    for (int rt = 0; rt < RT; ++rt) {
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[rt][t] = f4{0, 0, 0, 0};
      y0s[rt] = d4{0, 0, 0, 0};
    }
    for (int kb = 0; kb < nkb; ++kb) {
      __syncthreads();  // the step's P_g tile and K values are in LDS
      if (active) {
        // K operand: functions 32 kb + 8 lk + q of this lane's point
        const float* kv = sK + (16 * wave + lrow) * kEvRow + 8 * lk;
        float xh[8], xl[8];
        float xg[3][8];
        load8(kv, xh);
        load8(kv + kEvPlane, xl);
        if constexpr (NT == 4) {
#pragma unroll
          for (int a = 0; a < 3; ++a) load8(kv + (2 + a) * kEvPlane, xg[a]);
        }
        bf8 B0[4];
        split4f(xh, xl, B0);
        bf8 Bg[3][3];
        if constexpr (NT == 4) {
#pragma unroll
          for (int a = 0; a < 3; ++a) split3(xg[a], Bg[a][0], Bg[a][1], Bg[a][2]);
        }
#pragma unroll
        for (int rt = 0; rt < RT; ++rt) {
          if (rt < nrt) {
            const int u = (16 * rt + lrow) * 4 + lk;
            bf8 A[4];
#pragma unroll
            for (int pc = 0; pc < 4; ++pc) A[pc] = __builtin_bit_cast(bf8, sA[pc * ROWS * 4 + u]);
            f4 hi, lo;
            mma10(A, B0, hi, lo);
#pragma unroll
            for (int i = 0; i < 4; ++i) y0s[rt][i] += (double)hi[i] + (double)lo[i];
            if constexpr (NT == 4) {
#pragma unroll
              for (int a = 0; a < 3; ++a)
                acc[rt][1 + a] += mma6(A[0], A[1], A[2], Bg[a][0], Bg[a][1], Bg[a][2]);
            }
          }
        }
      }
      __syncthreads();  // every wave is done with the tile and the K values
      if (kb == nkb - 1) {
        // row-dot epilogue of this row chunk with FP64 values (its radial
        // tables take the K value area, one per wave)
        if (active) {
          const double ppx = px[q], ppy = py[q], ppz = pz[q];
          double s[NS];
#pragma unroll
          for (int k = 0; k < NS; ++k) s[k] = 0.0;
          if (ybuf) {  // Y = P_g phi of this point for the force pass
#pragma unroll
            for (int rt = 0; rt < RT; ++rt)
              if (rt < nrt)
                *reinterpret_cast<fl4*>(ybuf + g.phi_off + (long long)pt * mp + c0 + 16 * rt +
                                        4 * lk) = fl4{(float)y0s[rt][0], (float)y0s[rt][1],
                                                      (float)y0s[rt][2], (float)y0s[rt][3]};
          }
#pragma unroll
          for (int rt = 0; rt < RT; ++rt) {
            if (rt >= nrt) continue;
            const int2 tr = tsh[(g.fn_off >> 4) + (c0 >> 4) + rt];
            wave_lds_sync();  // the previous readers of the table are done
            radial_table<NS, 256>(g, tr, shells, bas, ppx, ppy, ppz, lrow, lk, rw);
            wave_lds_sync();
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              double v[NS];
              basis_from_table<NS, 256>(g, fmeta, c0 + 16 * rt + 4 * lk + i, tr.x, ppx, ppy,
                                        ppz, lrow, rw, v);
              const double y0 = y0s[rt][i];
              s[0] += v[0] * y0;
              if constexpr (NT == 4) {
                const double yx = acc[rt][1][i], yy = acc[rt][2][i], yz = acc[rt][3][i];
                s[1] += v[1] * y0;
                s[2] += v[2] * y0;
                s[3] += v[3] * y0;
                s[4] += v[4] * y0 + v[1] * yx;
                s[5] += v[5] * y0 + v[2] * yy;
                s[6] += v[6] * y0 + v[3] * yz;
                s[7] += v[7] * y0 + v[1] * yy;
                s[8] += v[8] * y0 + v[1] * yz;
                s[9] += v[9] * y0 + v[2] * yz;
              }
            }
          }
          // this chunk's sums of the lane's point (its 4 lane groups, in a
          // fixed order) onto the point's running sums
#pragma unroll
          for (int k = 0; k < NS; ++k) {
            s[k] += __shfl_xor(s[k], 16);
            s[k] += __shfl_xor(s[k], 32);
          }
          if (lk == 0)
#pragma unroll
            for (int k = 0; k < NS; ++k) sS[k][16 * wave + lrow] += s[k];
        }
        __syncthreads();  // the radial tables are done before the next K values
      }
      const int nc = kb + 1 < nkb ? c : c + 1, nk = kb + 1 < nkb ? kb + 1 : 0;
      if (nc < nchunk) {
        dma(nc * ROWS, 32 * nk);
        eval_k(nk);
      }
    }
  }
  double e_w = 0.0, n_w = 0.0;
  if (active) {
    if (lk == 0) {
      double s[NS];
#pragma unroll
      for (int k = 0; k < NS; ++k) s[k] = sS[k][16 * wave + lrow];
      point_stage<NT>(s, g.pt_off + p0 + lrow, iexch, spin, pw, factor, factor_b, sums, sstride,
                      e_w, n_w);
    }
    for (int o = 8; o >= 1; o >>= 1) {
      e_w += __shfl_xor(e_w, o);
      n_w += __shfl_xor(n_w, o);
    }
  }
  if (lane == 0) {
    sE[wave][0] = e_w;
    sE[wave][1] = n_w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    epart[2 * blockIdx.x] = ((sE[0][0] + sE[1][0]) + sE[2][0]) + sE[3][0];
    epart[2 * blockIdx.x + 1] = ((sE[0][1] + sE[1][1]) + sE[2][1]) + sE[3][1];
  }
}

// ---------------------------------------------------------------------------
// XC forces (computation types 2 and 3)
// ---------------------------------------------------------------------------
// LIO's XC force on nucleus A (g2g/cpu/iteration.cpp:213-268) is
//   f_A = sum_p factor_p dd_p(A),  dd_p(A) = -2 sum_{j on A} Y_j(p) grad phi_j(p)
// with Y = P phi (true symmetric P) and factor_p = w_p y2a_p from the density
// launch.  Per 64-point block the force kernels recompute Y on the matrix unit
// (one product instead of the density's four) and write, for every function
// j of the group, gpart[j][a] = sum over the block's points of
// factor_p Y_j(p) d_a phi_j(p): lanes of one function are added by a fixed
// butterfly over the wave's 16 points, the four waves in wave order.
// k_force_group sums a group's blocks in order and k_force_atom the
// functions of each nucleus (ascending) over the groups holding them
// (ascending), so the forces are as reproducible as the energy.

// Adds c over the 16 lanes of a lane group (one function, 16 points) and
// stores the sum of wave w at sG[w][jj] from the lane group's first lane.
__device__ inline void force_lanes(double (&c)[3], int lrow, double* slot) {
#pragma unroll
  for (int o = 1; o < 16; o <<= 1) {
#pragma unroll
    for (int a = 0; a < 3; ++a) c[a] += __shfl_xor(c[a], o);
  }
  if (lrow == 0) {
    slot[0] = c[0];
    slot[1] = c[1];
    slot[2] = c[2];
  }
}

// Sums w[3 i + a] (four functions i, three components a) over the 16 lanes
// of a row (the 16 points of a wave's tile) in a fixed order: two halving
// exchanges (xor 8, then 4) leave each lane three partial sums of one
// function, a butterfly (xor 2, 1) completes them, and the lanes with
// lrow % 4 == 0 store function i = lrow / 4 into slot jj[i].
__device__ inline void force_rows(const double (&w)[12], int lrow, double (*slot)[3],
                                  const int (&jj)[4]) {
  const bool h8 = (lrow & 8) != 0, h4 = (lrow & 4) != 0;
  double u[6], v[3];
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    const double keep = h8 ? w[6 + t] : w[t], send = h8 ? w[t] : w[6 + t];
    u[t] = keep + __shfl_xor(send, 8);
  }
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const double keep = h4 ? u[3 + t] : u[t], send = h4 ? u[t] : u[3 + t];
    v[t] = keep + __shfl_xor(send, 4);
  }
#pragma unroll
  for (int o = 2; o >= 1; o >>= 1) {
#pragma unroll
    for (int t = 0; t < 3; ++t) v[t] += __shfl_xor(v[t], o);
  }
  if ((lrow & 3) == 0) {
    double* dst = slot[jj[(h8 ? 2 : 0) + (h4 ? 1 : 0)]];
    dst[0] = v[0];
    dst[1] = v[1];
    dst[2] = v[2];
  }
}

// FP64 / FP32 values (planes [value][function][point]).
template <class T>
__global__ __launch_bounds__(256, 2) void k_force(const DensityItem* __restrict__ items,
                                                  const DevGroup* __restrict__ groups,
                                                  const T* __restrict__ pg,
                                                  const T* __restrict__ vals,
                                                  const double* __restrict__ factor,
                                                  double* __restrict__ gpart) {
  __shared__ double sG[4][kFnChunk][3];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;  // uniform over the wave
  const int np = g.npts_pad, mp = g.m_pad;
  const long long plane = (long long)mp * np;
  const T* V = vals + g.val_off;
  const T* Pm = pg + g.pg_off;
  const double f = active ? factor[g.pt_off + p0 + lrow] : 0.0;
  double* out = gpart + g.force_off + (long long)(it.p0 / kDensPts) * mp * 3;
  for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
    const int nrt = min(4, (mp - c0) / 16);
    typename Mma<T>::acc acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) acc[rt] = typename Mma<T>::acc{0, 0, 0, 0};
    if (active) {
      const T* Bk = V + (long long)lk * np + p0 + lrow;   // phi[k][p]
      const T* Ak = Pm + (long long)lk * mp + c0 + lrow;  // P[k][j] = P[j][k]
      for (int k0 = 0; k0 < mp; k0 += 4) {
        const T b = Bk[(long long)k0 * np];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
          if (rt < nrt) Mma<T>::op(Ak[(long long)k0 * mp + 16 * rt], b, acc[rt]);
      }
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int jj = 16 * rt + Mma<T>::row(lk, i);
        double c[3] = {0.0, 0.0, 0.0};
        if (active && rt < nrt) {
          const T* vj = V + (long long)(c0 + jj) * np + p0 + lrow;
          const double y = f * (double)acc[rt][i];
#pragma unroll
          for (int a = 0; a < 3; ++a) c[a] = y * (double)vj[(a + 1) * plane];
        }
        force_lanes(c, lrow, sG[wave][jj]);
      }
    }
    __syncthreads();
    if (threadIdx.x < 3 * kFnChunk) {
      const int jj = threadIdx.x / 3, a = threadIdx.x % 3;
      if (c0 + jj < mp)
        out[(long long)(c0 + jj) * 3 + a] =
            ((sG[0][jj][a] + sG[1][jj][a]) + sG[2][jj][a]) + sG[3][jj][a];
    }
    __syncthreads();
  }
}

// Split-BF16 values (point-major FP32 planes, P_g as three BF16 pieces): Y on
// the six-product BF16 MFMA chain as in the density kernels, the P_g pieces of
// each chunk and K step copied once per block straight into LDS
// (double-buffered), the next K step's B operand loaded during this step.
__global__ __launch_bounds__(256, 2) void k_force_bf(const DensityItem* __restrict__ items,
                                                     const DevGroup* __restrict__ groups,
                                                     const unsigned short* __restrict__ pg3,
                                                     const float* __restrict__ vals,
                                                     const double* __restrict__ factor,
                                                     double* __restrict__ gpart) {
  constexpr int ROWS = kFnChunk;
  constexpr int UNITS = 3 * ROWS * 4;  // 16-byte units of one staged tile
  __shared__ u4 sA[2][UNITS];          // [buf][piece][row][k octet]
  __shared__ double sG[4][kFnChunk][3];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;  // uniform over the wave
  const int np = g.npts_pad, mp = g.m_pad;
  const long long plane = (long long)mp * np;
  const long long psz = (long long)mp * mp;
  const float* Vp = vals + g.val_off + (long long)(active ? p0 + lrow : 0) * mp;
  const unsigned short* Pp = pg3 + 3 * g.pg_off;
  const double f = active ? factor[g.pt_off + p0 + lrow] : 0.0;
  double* out = gpart + g.force_off + (long long)(it.p0 / kDensPts) * mp * 3;
  // staging map as in the density kernel: unit c = threadIdx.x + 256 i is
  // piece c / (4 ROWS), row (c / 4) % ROWS, k octet c % 4, copied from global
  // memory straight into LDS (lane l of wave w writes unit 256 i + 64 w + l;
  // rows past the group's functions are clamped and never read)
  auto dma = [&](int c0, int k0, int b) {
#pragma unroll
    for (int i = 0; i < UNITS / 256; ++i) {
      const int c = threadIdx.x + 256 * i;
      const int sp = c / (4 * ROWS), row = (c >> 2) & (ROWS - 1), oc = c & 3;
      const int j = min(c0 + row, mp - 1);
      __builtin_amdgcn_global_load_lds(
          (const void*)(Pp + sp * psz + (long long)j * mp + k0 + 8 * oc),
          (__attribute__((address_space(3))) void*)&sA[b][256 * i + 64 * wave], 16, 0, 0);
    }
  };
  for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
    const int nrt = min(4, (mp - c0) / 16);
    f4 acc[4];
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) acc[rt] = f4{0, 0, 0, 0};
    float xb[8];  // the B operand of the current K step, loaded one step ahead
    load8(Vp + 8 * lk, xb);
    dma(c0, 0, 0);
    __syncthreads();
    for (int k0 = 0, buf = 0; k0 < mp; k0 += 32, buf ^= 1) {
      const bool more = k0 + 32 < mp;
      if (more) dma(c0, k0 + 32, buf ^ 1);
      if (active) {
        bf8 bh, bm, bl;
        split3(xb, bh, bm, bl);
        if (more) load8(Vp + k0 + 32 + 8 * lk, xb);
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) {
          if (rt < nrt) {
            const int u = (16 * rt + lrow) * 4 + lk;
            const bf8 ah = __builtin_bit_cast(bf8, sA[buf][u]);
            const bf8 am = __builtin_bit_cast(bf8, sA[buf][ROWS * 4 + u]);
            const bf8 al = __builtin_bit_cast(bf8, sA[buf][2 * ROWS * 4 + u]);
            acc[rt] += mma6(ah, am, al, bh, bm, bl);
          }
        }
      }
      __syncthreads();
    }
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      fl4 v[3] = {fl4{0, 0, 0, 0}, fl4{0, 0, 0, 0}, fl4{0, 0, 0, 0}};
      if (active && rt < nrt) {
        const float* vj = Vp + c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = *reinterpret_cast<const fl4*>(vj + (a + 1) * plane);
      }
      double w[12];
      int jj[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        jj[i] = 16 * rt + 4 * lk + i;
        const double y = f * (double)acc[rt][i];
        w[3 * i] = y * (double)v[0][i];
        w[3 * i + 1] = y * (double)v[1][i];
        w[3 * i + 2] = y * (double)v[2][i];
      }
      force_rows(w, lrow, sG[wave], jj);
    }
    __syncthreads();
    if (threadIdx.x < 3 * kFnChunk) {
      const int jj = threadIdx.x / 3, a = threadIdx.x % 3;
      if (c0 + jj < mp)
        out[(long long)(c0 + jj) * 3 + a] =
            ((sG[0][jj][a] + sG[1][jj][a]) + sG[2][jj][a]) + sG[3][jj][a];
    }
    __syncthreads();
  }
}

// Split-BF16 forces from the Y = P_g phi the density kernel left in ybuf
// (point-major, exact FP32 sums): the same per-function sums as k_force_bf,
// in the same lane and wave order, without recomputing the product.
__global__ __launch_bounds__(256) void k_force_y(const DensityItem* __restrict__ items,
                                                 const DevGroup* __restrict__ groups,
                                                 const float* __restrict__ ybuf,
                                                 const float* __restrict__ vals,
                                                 const double* __restrict__ factor,
                                                 double* __restrict__ gpart) {
  __shared__ double sG[4][kFnChunk][3];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;  // uniform over the wave
  const int np = g.npts_pad, mp = g.m_pad;
  const long long plane = (long long)mp * np;
  const long long prow = (long long)(active ? p0 + lrow : 0) * mp;
  const float* Vp = vals + g.val_off + prow;
  const float* Yp = ybuf + g.phi_off + prow;
  const double f = active ? factor[g.pt_off + p0 + lrow] : 0.0;
  double* out = gpart + g.force_off + (long long)(it.p0 / kDensPts) * mp * 3;
  for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
    const int nrt = min(4, (mp - c0) / 16);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      fl4 y = fl4{0, 0, 0, 0};
      fl4 v[3] = {fl4{0, 0, 0, 0}, fl4{0, 0, 0, 0}, fl4{0, 0, 0, 0}};
      if (active && rt < nrt) {
        const int j = c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
        y = *reinterpret_cast<const fl4*>(Yp + j);
#pragma unroll
        for (int a = 0; a < 3; ++a) v[a] = *reinterpret_cast<const fl4*>(Vp + j + (a + 1) * plane);
      }
      double w[12];
      int jj[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        jj[i] = 16 * rt + 4 * lk + i;
        const double yy = f * (double)y[i];
        w[3 * i] = yy * (double)v[0][i];
        w[3 * i + 1] = yy * (double)v[1][i];
        w[3 * i + 2] = yy * (double)v[2][i];
      }
      force_rows(w, lrow, sG[wave], jj);
    }
    __syncthreads();
    if (threadIdx.x < 3 * kFnChunk) {
      const int jj = threadIdx.x / 3, a = threadIdx.x % 3;
      if (c0 + jj < mp)
        out[(long long)(c0 + jj) * 3 + a] =
            ((sG[0][jj][a] + sG[1][jj][a]) + sG[2][jj][a]) + sG[3][jj][a];
    }
    __syncthreads();
  }
}

// Forces of density variant 6: Y = P_g phi from the density kernel (ybuf,
// point-major, the FP32 sums of the ten-product chain) and the gradient of
// every function evaluated here in FP64 from the per-shell radial sums of
// each row tile; the same per-function sums, lane and wave order as
// k_force_y.
__global__ __launch_bounds__(256) void k_force_ev(const DensityItem* __restrict__ items,
                                                  const DevGroup* __restrict__ groups,
                                                  const float* __restrict__ ybuf,
                                                  const FnMeta* __restrict__ fmeta,
                                                  const int* __restrict__ shells,
                                                  const int2* __restrict__ tsh,
                                                  const DevFn* __restrict__ bas,
                                                  const double* __restrict__ px,
                                                  const double* __restrict__ py,
                                                  const double* __restrict__ pz,
                                                  const double* __restrict__ factor,
                                                  double* __restrict__ gpart) {
  __shared__ double sG[4][kFnChunk][3];
  __shared__ double sR[4][512];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int p0 = it.p0 + 16 * wave;
  const bool active = p0 < g.npts_pad;  // uniform over the wave
  const int mp = g.m_pad;
  const int pt = active ? p0 + lrow : it.p0;
  const int q = g.pt_off + pt;
  const double ppx = px[q], ppy = py[q], ppz = pz[q];
  const float* Yp = ybuf + g.phi_off + (long long)pt * mp;
  const double f = active ? factor[q] : 0.0;
  double* rw = sR[wave];
  double* out = gpart + g.force_off + (long long)(it.p0 / kDensPts) * mp * 3;
  for (int c0 = 0; c0 < mp; c0 += kFnChunk) {
    const int nrt = min(4, (mp - c0) / 16);
#pragma unroll
    for (int rt = 0; rt < 4; ++rt) {
      fl4 y = fl4{0, 0, 0, 0};
      double v[4][4];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int a = 0; a < 4; ++a) v[i][a] = 0.0;
      if (active && rt < nrt) {
        const int2 tr = tsh[(g.fn_off >> 4) + (c0 >> 4) + rt];
        wave_lds_sync();
        radial_table<4, 256>(g, tr, shells, bas, ppx, ppy, ppz, lrow, lk, rw);
        wave_lds_sync();
        const int j = c0 + 16 * rt + 4 * lk;  // functions j .. j+3 of this point
        y = *reinterpret_cast<const fl4*>(Yp + j);
#pragma unroll
        for (int i = 0; i < 4; ++i)
          basis_from_table<4, 256>(g, fmeta, j + i, tr.x, ppx, ppy, ppz, lrow, rw, v[i]);
      }
      double w[12];
      int jj[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        jj[i] = 16 * rt + 4 * lk + i;
        const double yy = f * (double)y[i];
        w[3 * i] = yy * v[i][1];
        w[3 * i + 1] = yy * v[i][2];
        w[3 * i + 2] = yy * v[i][3];
      }
      force_rows(w, lrow, sG[wave], jj);
    }
    __syncthreads();
    if (threadIdx.x < 3 * kFnChunk) {
      const int jj = threadIdx.x / 3, a = threadIdx.x % 3;
      if (c0 + jj < mp)
        out[(long long)(c0 + jj) * 3 + a] =
            ((sG[0][jj][a] + sG[1][jj][a]) + sG[2][jj][a]) + sG[3][jj][a];
    }
    __syncthreads();
  }
}

// gsum[fn_off + j][a] = sum over the group's blocks (in order) of gpart; one
// block per group.
// accumulate: add onto gsum (the second spin of an open shell).
// One thread per (group, element), blockIdx.x the group; the loads of
// several blocks are in flight together, the sum stays in block order.
__global__ __launch_bounds__(256) void k_force_group(const DevGroup* __restrict__ groups,
                                                     const double* __restrict__ gpart,
                                                     double* __restrict__ gsum, int accumulate) {
  const DevGroup g = groups[blockIdx.x];
  const int e = blockIdx.y * 256 + threadIdx.x;
  if (e >= 3 * g.m_pad) return;
  const int nblk = (g.npts_pad + kDensPts - 1) / kDensPts;
  const long long stride = 3LL * g.m_pad;
  const double* src = gpart + g.force_off + e;
  double s = 0.0;
#pragma unroll 8
  for (int b = 0; b < nblk; ++b) s += src[b * stride];
  double* dst = gsum + 3LL * g.fn_off + e;
  *dst = accumulate ? *dst + s : s;
}

// Per batch function I, the sum of its group sums over the groups holding it
// (ascending group order); one thread per function.
__global__ __launch_bounds__(256) void k_force_fn(const int* __restrict__ memb_off,
                                                  const int2* __restrict__ memb,
                                                  const DevGroup* __restrict__ groups,
                                                  const double* __restrict__ gsum, int M,
                                                  double* __restrict__ fsum) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= M) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int k = memb_off[I]; k < memb_off[I + 1]; ++k) {
    const int2 u = memb[k];
    const double* gs = gsum + 3LL * (groups[u.x].fn_off + u.y);
    s[0] += gs[0];
    s[1] += gs[1];
    s[2] += gs[2];
  }
  fsum[3LL * I] = s[0];
  fsum[3LL * I + 1] = s[1];
  fsum[3LL * I + 2] = s[2];
}

// Force on batch atom A (atom-major out[3 A + a]) = -2 sum over A's functions
// (ascending) of their function sums.
__global__ __launch_bounds__(64) void k_force_atom(const int* __restrict__ afn_off,
                                                   const int* __restrict__ afn,
                                                   const double* __restrict__ fsum, int natom,
                                                   double* __restrict__ out) {
  const int A = blockIdx.x * blockDim.x + threadIdx.x;
  if (A >= natom) return;
  double s[3] = {0.0, 0.0, 0.0};
  for (int q = afn_off[A]; q < afn_off[A + 1]; ++q) {
    const double* fs = fsum + 3LL * afn[q];
    s[0] += fs[0];
    s[1] += fs[1];
    s[2] += fs[2];
  }
  out[3 * A + 0] = -2.0 * s[0];
  out[3 * A + 1] = -2.0 * s[1];
  out[3 * A + 2] = -2.0 * s[2];
}

// F_g tile (ti, tj) = sum over slices, in slice order.  Block b of a tile
// pair handles elements [256 (b % 16), +256) of the 64 x 64 tile.
__device__ __forceinline__ void fock_group_item(const int bi, const DevGroup* __restrict__ groups,
                                                    const int2* __restrict__ tiles,
                                                    const double* __restrict__ fpart,
                                                    double* __restrict__ fg) {
  const int2 gt = tiles[bi >> 4];  // x: group, y: tile-pair index
  const DevGroup g = groups[gt.x];
  int ti = 0, rem = gt.y;
  while (rem >= g.ntile - ti) {
    rem -= g.ntile - ti;
    ++ti;
  }
  const int tj = ti + rem;
  const int e = ((bi & 15) << 8) + threadIdx.x;
  const int r = e / kFnChunk, c = e % kFnChunk;
  const int i = kFnChunk * ti + r, j = kFnChunk * tj + c;
  if (i >= g.m_pad || j >= g.m_pad) return;
  const double* src = fpart + (long long)(g.fock_base + gt.y * g.nslice) * kFnChunk * kFnChunk + e;
  double s = 0.0;
#pragma unroll 8
  for (int sl = 0; sl < g.nslice; ++sl) s += src[(long long)sl * kFnChunk * kFnChunk];
  fg[g.pg_off + (long long)i * g.m_pad + j] = s;
}

__global__ __launch_bounds__(256) void k_fock_group(const DevGroup* __restrict__ groups,
                                                    const int2* __restrict__ tiles,
                                                    const double* __restrict__ fpart,
                                                    double* __restrict__ fg) {
  fock_group_item(blockIdx.x, groups, tiles, fpart, fg);
}

// Packed F: entry (I <= J) sums F_g(li, lj) over the groups holding both
// functions, in ascending group order.  memb holds, per function, its
// (group, local index) pairs sorted by group.  I (blockIdx.y) and J are
// batch function indices; J runs over the rest of I's molecule.  The block
// first maps each group of I's molecule to I's local index in it (-1 where
// I is absent) in LDS; each thread then walks J's list alone, in ascending
// group order, so the sum is the same as a merge of the two lists.
__global__ void k_fock_pack(const int* __restrict__ memb_off, const int2* __restrict__ memb,
                            const DevGroup* __restrict__ groups, const DevMol* __restrict__ mols,
                            const int* __restrict__ fn_mol, const double* __restrict__ fg,
                            double* __restrict__ F) {
  extern __shared__ int smap[];
  const int I = blockIdx.y;
  const DevMol mol = mols[fn_mol[I]];
  for (int q = threadIdx.x; q < mol.ng; q += blockDim.x) smap[q] = -1;
  __syncthreads();
  for (int a = memb_off[I] + threadIdx.x; a < memb_off[I + 1]; a += blockDim.x) {
    const int2 u = memb[a];
    smap[u.x - mol.g0] = u.y;
  }
  __syncthreads();
  const int J = mol.fn_base + blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= mol.fn_base + mol.M || J < I) return;
  double s = 0.0;
  for (int b = memb_off[J]; b < memb_off[J + 1]; ++b) {
    const int2 v = memb[b];
    const int ui = smap[v.x - mol.g0];
    if (ui >= 0) {
      const DevGroup g = groups[v.x];
      const int li = min(ui, v.y), lj = max(ui, v.y);
      s += fg[g.pg_off + (long long)li * g.m_pad + lj];
    }
  }
  F[mol.rho_off + packed_index(I - mol.fn_base, J - mol.fn_base, mol.M)] = s;
}

// The same by merging the two sorted lists, for molecules with more groups
// than the LDS map holds.
__global__ void k_fock_pack_merge(const int* __restrict__ memb_off, const int2* __restrict__ memb,
                            const DevGroup* __restrict__ groups, const DevMol* __restrict__ mols,
                            const int* __restrict__ fn_mol, const double* __restrict__ fg,
                            double* __restrict__ F) {
  const int I = blockIdx.y;
  const DevMol mol = mols[fn_mol[I]];
  const int J = mol.fn_base + blockIdx.x * blockDim.x + threadIdx.x;
  if (J >= mol.fn_base + mol.M || J < I) return;
  int a = memb_off[I], ae = memb_off[I + 1];
  int b = memb_off[J], be = memb_off[J + 1];
  double s = 0.0;
  while (a < ae && b < be) {
    const int2 u = memb[a], v = memb[b];
    if (u.x < v.x) {
      ++a;
    } else if (v.x < u.x) {
      ++b;
    } else {
      const DevGroup g = groups[u.x];
      const int li = min(u.y, v.y), lj = max(u.y, v.y);
      s += fg[g.pg_off + (long long)li * g.m_pad + lj];
      ++a;
      ++b;
    }
  }
  F[mol.rho_off + packed_index(I - mol.fn_base, J - mol.fn_base, mol.M)] = s;
}

// Fixed-order reduction of the per-block (Exc, N) pairs of each molecule
// (block b: molecule b, its density blocks eidx[eb .. ee)).
__global__ __launch_bounds__(1024) void k_energy(const double* __restrict__ epart,
                                                 const DevMol* __restrict__ mols,
                                                 const int* __restrict__ eidx,
                                                 double* __restrict__ out) {
  __shared__ double se[1024], sn[1024];
  const DevMol mol = mols[blockIdx.x];
  double e = 0.0, c = 0.0;
  for (int i = mol.eb + threadIdx.x; i < mol.ee; i += 1024) {
    e += epart[2 * eidx[i]];
    c += epart[2 * eidx[i] + 1];
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = c;
  __syncthreads();
  for (int s = 512; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sn[threadIdx.x] += sn[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[2 * blockIdx.x] = se[0];
    out[2 * blockIdx.x + 1] = sn[0];
  }
}

// ---------------------------------------------------------------------------
// One persistent launch per build (split-BF16 density variant 8, closed shell)
// ---------------------------------------------------------------------------
// Every phase of a build -- density expansion, density and functional, Fock
// tiles, tile sums over slices, packing, energies -- as items of one queue,
// taken in order by a fixed set of resident blocks (atomic head counter).
// An item waits only for items earlier in the queue: a density block for
// its group's expansion blocks, a Fock tile slice for its group's density
// blocks, a tile sum for its slices, the packing for every tile sum, an
// energy for every density block.  An item is taken only after every item
// before it was taken by a running block that does not wait for anything
// later, so every wait ends (no co-residency of the whole grid is needed,
// and builds on other streams may share the device).  Producers publish
// with an agent-scope release after the block's stores, consumers poll with
// relaxed loads and take an agent-scope acquire before reading.  A wait
// that has not ended after about 2^26 polls sets the error word and the
// energies come out NaN (a safety net; the order above rules it out).
// Every item computes exactly what its block computes in the phase
// launches, with the same item index, so the results are bitwise the same.

struct PersistArgs {
  int n0, n1, n2, n3, n4, n5;  // queue: [0,n0) expansion, [n0,n1) density, [n1,n2) Fock,
                               // [n2,n3) tile sums (16 per tile pair), [n3,n4) packing,
                               // [n4,n5) energies
  int* sync;                   // head, error, tile sums done, density done, then per group
                               // expansion done, density done, then per tile pair slices done
  int ngroups, ntiles;
  const int* gneed;            // expansion blocks per group
  const int* dneed;            // density blocks per group
  const int* tbase;            // first tile pair (gtiles index) of each group
  int fgneed, dallneed;
  const GatherItem* gitems;
  const DevGroup* groups;
  const DevMol* mols;
  const int* fnl;
  unsigned short* pg;
  const DensityItem* ditems;
  const float* vals;
  const double* pw;
  double* factor;
  double* epart;
  int iexch;
  const double *px, *py, *pz;
  const FnMeta* fmeta;
  const float* rad;
  const unsigned short* lo16;
  const FockItem* fitems;
  const int* fperm;            // queue order of the Fock items (heavy groups first)
  const float* phiT;
  double* fpart;
  const int2* gtiles;
  const int* tperm;            // queue order of the tile-sum items
  double* fg;
  const int2* pitems;          // packing: (batch function I, 256-wide block of J)
  const int* memb_off;
  const int2* memb;
  const int* fn_mol;
  const int* eidx;
  double* eout;
};

// Once a wait has timed out every other wait gives up at once, so a failed
// build drains quickly (its energies are NaN).
__device__ inline void persist_wait(int* cnt, int need, int* err) {
  if (threadIdx.x == 0) {
    long long spins = 0;
    while (__hip_atomic_load(cnt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
      __builtin_amdgcn_s_sleep(2);
      if ((++spins & 1023) == 0 &&
          __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
        break;
      if (spins > (1LL << 26)) {
        __hip_atomic_fetch_or(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

__device__ inline void persist_signal(int* cnt) {
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(cnt, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}

// Each phase of the persistent build is its own (not inlined) function, so
// each is register-allocated alone, as in its phase kernel: inlined into one
// loop, values the phases share got hoisted and stayed live through the
// density's 256-register body (100+ registers spilled).
template <int NT>
__device__ __attribute__((noinline)) void persist_density(const PersistArgs* pa, int d,
                                                          float* ybuf,
                                                          DensBf2Smem<NT, 64, 1>* sm) {
  const PersistArgs& a = *pa;
  density_bf2_item<NT, 64, false, 1>(d, a.ditems, a.groups, a.pg, a.vals, a.pw, a.factor, a.epart,
                                     a.iexch, 0, nullptr, nullptr, 0, a.px, a.py, a.pz, a.fmeta,
                                     a.rad, ybuf, a.lo16, *sm);
}

__device__ __attribute__((noinline)) void persist_fock(const PersistArgs* pa, int f,
                                                       double (*sm)[kFnChunk * kFnChunk]) {
  const PersistArgs& a = *pa;
  fock_bf_item(f, a.fitems, a.groups, a.phiT, a.factor, a.fpart, *sm);
}

union PersistSmem {
  DensBf2Smem<4, 64, 1> d4;
  DensBf2Smem<1, 64, 1> d1;
  double f[kFnChunk * kFnChunk];
  double e[2][1024];
};

// The grid's tables live in device memory (written once per grid load) and
// are read again for every item; the buffers of the call (packed rho and F,
// the force pass's Y) are kernel arguments, so a captured graph records them.
template <int NT>
__global__ __launch_bounds__(256, 2) void k_build_persist(PersistArgs* pargs, int* sync,
                                                         const double* rho, double* F,
                                                         float* ybuf) {
  __shared__ PersistSmem sm;
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = __hip_atomic_fetch_add(sync, 1, __ATOMIC_RELAXED,
                                                          __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    const PersistArgs& a = *pargs;
    int* err = a.sync + 1;
    int* fgdone = a.sync + 2;
    int* dall = a.sync + 3;
    int* gdone = a.sync + 4;
    int* ddone = gdone + a.ngroups;
    int* tdone = ddone + a.ngroups;
    if (it >= a.n5) break;
    if (it < a.n0) {
      gather_bf_item<4>(it, a.gitems, a.groups, a.mols, a.fnl, rho, a.pg);
      persist_signal(&gdone[a.gitems[it].group]);
    } else if (it < a.n1) {
      const int d = it - a.n0, g = a.ditems[d].group;
      persist_wait(&gdone[g], a.gneed[g], err);
      if constexpr (NT == 4)
        persist_density<4>(pargs, d, ybuf, &sm.d4);
      else
        persist_density<1>(pargs, d, ybuf, &sm.d1);
      __syncthreads();
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(&ddone[g], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
      if (threadIdx.x == 0)
        __hip_atomic_fetch_add(dall, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    } else if (it < a.n2) {
      const int f = a.fperm[it - a.n1];
      const FockItem fi = a.fitems[f];
      persist_wait(&ddone[fi.group], a.dneed[fi.group], err);
      persist_fock(pargs, f, &sm.f);
      const DevGroup& G = a.groups[fi.group];
      persist_signal(&tdone[a.tbase[fi.group] + (f - G.fock_base) / G.nslice]);
    } else if (it < a.n3) {
      const int q = a.tperm[it - a.n2], t = q >> 4;
      persist_wait(&tdone[t], a.groups[a.gtiles[t].x].nslice, err);
      fock_group_item(q, a.groups, a.gtiles, a.fpart, a.fg);
      persist_signal(fgdone);
    } else if (it < a.n4) {
      persist_wait(fgdone, a.fgneed, err);
      const int2 pi = a.pitems[it - a.n3];
      const int I = pi.x;
      const DevMol mol = a.mols[a.fn_mol[I]];
      const int J = mol.fn_base + pi.y * 256 + (int)threadIdx.x;
      if (J < mol.fn_base + mol.M && J >= I) {
        int u0 = a.memb_off[I], ue = a.memb_off[I + 1];
        int v0 = a.memb_off[J], ve = a.memb_off[J + 1];
        double sum = 0.0;
        while (u0 < ue && v0 < ve) {
          const int2 u = a.memb[u0], v = a.memb[v0];
          if (u.x < v.x) {
            ++u0;
          } else if (v.x < u.x) {
            ++v0;
          } else {
            const DevGroup gg = a.groups[u.x];
            const int li = min(u.y, v.y), lj = max(u.y, v.y);
            sum += a.fg[gg.pg_off + (long long)li * gg.m_pad + lj];
            ++u0;
            ++v0;
          }
        }
        F[mol.rho_off + packed_index(I - mol.fn_base, J - mol.fn_base, mol.M)] = sum;
      }
    } else {
      // energies: the 1024-lane fixed-order reduction of the phase launch,
      // four lanes per thread
      persist_wait(dall, a.dallneed, err);
      const int m = it - a.n4;
      const DevMol mol = a.mols[m];
      for (int v = threadIdx.x; v < 1024; v += 256) {
        double e = 0.0, c = 0.0;
        for (int i = mol.eb + v; i < mol.ee; i += 1024) {
          e += a.epart[2 * a.eidx[i]];
          c += a.epart[2 * a.eidx[i] + 1];
        }
        sm.e[0][v] = e;
        sm.e[1][v] = c;
      }
      __syncthreads();
      for (int st = 512; st > 0; st >>= 1) {
        for (int v = threadIdx.x; v < st; v += 256) {
          sm.e[0][v] += sm.e[0][v + st];
          sm.e[1][v] += sm.e[1][v + st];
        }
        __syncthreads();
      }
      if (threadIdx.x == 0) {
        const bool bad = __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        a.eout[2 * m] = bad ? __builtin_nan("") : sm.e[0][0];
        a.eout[2 * m + 1] = bad ? __builtin_nan("") : sm.e[1][0];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// device grid rebuild: group-local Becke weights and zero-weight removal
// ---------------------------------------------------------------------------
// A staged group (before weights) and its molecule's atom tables.
struct StGroup {
  int pt_off, npts;    // staged points
  int atom_off, natoms;  // nuclei (molecule-local, ascending) in gatoms
  int mol;
  int sq_off;  // first entry of the group's natoms x natoms tables in near-first order
};
struct StMol {
  int atom_base, natom;  // atoms in the batch-wide atom arrays
  long long rij_off;     // natom x natom distances and size terms
};

// Size term of the Becke cell for atoms j, k: chi = rm_j / rm_k,
// chi <- (chi - 1) / (chi + 1), aw = chi / (chi^2 - 1) (g2g/cpu/weight.cpp:
// 13-112, the same operations in the same order as the host partition).
static double becke_size_term(double rmj, double rmk) {
#pragma clang fp contract(off)
  double chi = rmj / rmk;
  chi = (chi - 1.0) / (chi + 1.0);
  return chi / (chi * chi - 1.0);
}

__device__ inline double dist3(double x, double y, double z, double ax, double ay, double az) {
#pragma clang fp contract(off)
  const double dx = x - ax, dy = y - ay, dz = z - az;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

// Becke cell function with the pair's size term aw = chi / (chi^2 - 1)
// taken from a table (the host computes it with the same operations), so a
// factor costs one division instead of three.
__device__ inline double becke_cell_w(double dj, double dk, double rjk, double aw) {
#pragma clang fp contract(off)
  double u = (dj - dk) / rjk;
  u += aw * (1.0 - u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  u = 1.5 * u - 0.5 * (u * u * u);
  return 0.5 * (1.0 - u);
}

// Staged points per Becke block; the block has kBeckeLanes waves, one per
// class of j, over the same points.  The distances of the block's points to
// the first kBeckeNear nuclei of the near-first order are formed once per
// block, in LDS.
constexpr int kBeckePts = 64, kBeckeNear = 64;

// The near-first order of every staged group's nuclei: the nuclei by
// distance from the group's first staged point, ties in list order (the
// rank of q counts the nuclei strictly nearer, or as near and earlier), the
// order the host's becke_order forms with the same operations.  gord holds
// the nucleus indices in that order, gpos the rank of each list position, and
// the group's tables in that order follow: the nuclei's coordinates (ox, oy,
// oz) and, for every list position j and rank t, the pair distance and size
// term of nuclei j and gord[t] (td, tw, rows of natoms from sq_off), so that
// the Becke products load by rank, with no index load in between.  One block
// per group.
__global__ __launch_bounds__(256) void k_becke_order(
    const StGroup* __restrict__ groups, const StMol* __restrict__ mols,
    const int* __restrict__ gatoms, const double* __restrict__ ax,
    const double* __restrict__ ay, const double* __restrict__ az,
    const double* __restrict__ aw, const double* __restrict__ rij,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, int* __restrict__ gord, int* __restrict__ gpos,
    double* __restrict__ ox, double* __restrict__ oy, double* __restrict__ oz,
    double* __restrict__ td, double* __restrict__ tw) {
  const StGroup G = groups[blockIdx.x];
  if (G.npts == 0) return;
  const StMol m = mols[G.mol];
  const int* at = gatoms + G.atom_off;
  const int k = G.natoms, na = m.natom;
  const double* A = ax + m.atom_base;
  const double* B = ay + m.atom_base;
  const double* C = az + m.atom_base;
  const double x = sx[G.pt_off], y = sy[G.pt_off], z = sz[G.pt_off];
  for (int q = threadIdx.x; q < k; q += blockDim.x) {
    const double dq = dist3(x, y, z, A[at[q]], B[at[q]], C[at[q]]);
    int rank = 0;
    for (int r = 0; r < k; ++r) {
      const double dr = dist3(x, y, z, A[at[r]], B[at[r]], C[at[r]]);
      rank += (dr < dq || (dr == dq && r < q)) ? 1 : 0;
    }
    gord[G.atom_off + rank] = at[q];
    gpos[G.atom_off + q] = rank;
  }
  __syncthreads();  // the order, written by the block, is read back below
  const int* ord = gord + G.atom_off;
  for (int t = threadIdx.x; t < k; t += blockDim.x) {
    const int ak = ord[t];
    ox[G.atom_off + t] = A[ak];
    oy[G.atom_off + t] = B[ak];
    oz[G.atom_off + t] = C[ak];
  }
  const double* D = rij + m.rij_off;
  const double* W = aw + m.rij_off;
  for (int e = threadIdx.x; e < k * k; e += blockDim.x) {
    const int j = e / k, t = e % k;
    const long long src = (long long)at[j] * na + ord[t];
    td[(long long)G.sq_off + e] = D[src];
    tw[(long long)G.sq_off + e] = W[src];
  }
}

// The exact terms of the Becke weights, one staged point per lane: the
// point's nearest nucleus n (first in list order at the least distance) and
// own nucleus (list positions, -1 when it is not among the group's), P_own
// and P_n formed as the reference forms them (a zero test against n unless
// j = n, then the nuclei in list order; g2g/cpu/weight.cpp:13-112), so a
// weight is zero exactly when the reference's is.  A point whose own nucleus
// is not among the group's forms P_own over all of them.  The group is
// uniform in the block, so its nuclei and their coordinates are wave-uniform
// loads; the two products run in one pass over the nuclei.
__global__ __launch_bounds__(kBeckePts) void k_becke_exact(
    const StGroup* __restrict__ groups, const int2* __restrict__ items,
    const StMol* __restrict__ mols, const int* __restrict__ gatoms,
    const int* __restrict__ owner, const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, const double* __restrict__ aw, const double* __restrict__ rij,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, double2* __restrict__ bterm, int2* __restrict__ bnuc) {
#pragma clang fp contract(off)
  const int2 it = items[blockIdx.x];
  const int gi = __builtin_amdgcn_readfirstlane(it.x);
  const int p = __builtin_amdgcn_readfirstlane(it.y) + (int)threadIdx.x;
  const StGroup G = groups[gi];
  if (p >= G.pt_off + G.npts) return;
  const StMol m = mols[G.mol];
  const int* at = gatoms + G.atom_off;
  const int k = G.natoms, own = owner[p], na = m.natom;
  const double* A = ax + m.atom_base;
  const double* B = ay + m.atom_base;
  const double* C = az + m.atom_base;
  const double* W = aw + m.rij_off;
  const double* D = rij + m.rij_off;
  const double x = sx[p], y = sy[p], z = sz[p];
  int an = -1, qo = -1, qn = -1;
  double dn = 0.0;
  for (int q = 0; q < k; ++q) {
    const int aq = at[q];
    const double dq = dist3(x, y, z, A[aq], B[aq], C[aq]);
    if (aq == own) qo = q;
    if (an < 0 || dq < dn) {
      dn = dq;
      an = aq;
      qn = q;
    }
  }
  // P_own and P_n in one pass over the nuclei (each distance formed once),
  // each with the reference's factor order and zero exit
  double mine = 0.0, pn = 0.0;
  if (k > 0) {
    const double down = dist3(x, y, z, A[own], B[own], C[own]);
    const double* Do = D + (long long)own * na;
    const double* Wo = W + (long long)own * na;
    const double* Dn = D + (long long)an * na;
    const double* Wn = W + (long long)an * na;
    double po = 1.0, pq = 1.0;
    bool so = true, sn = qn != qo;  // products still running (P_n is not
                                    // needed once P_own is zero)
    if (own != an && becke_cell_w(down, dn, Do[an], Wo[an]) == 0.0) {
      po = 0.0;
      so = sn = false;
    }
    for (int q = 0; q < k && (so || sn); ++q) {
      const int ak = at[q];
      const double dq = dist3(x, y, z, A[ak], B[ak], C[ak]);
      if (so && ak != own) {
        po *= becke_cell_w(down, dq, Do[ak], Wo[ak]);
        so = po != 0.0;
        sn = sn && so;
      }
      if (sn && ak != an) {
        pq *= becke_cell_w(dn, dq, Dn[ak], Wn[ak]);
        sn = pq != 0.0;
      }
    }
    mine = po;
    pn = (mine == 0.0 || qn == qo) ? mine : pq;
  }
  bterm[p] = make_double2(mine, pn);
  bnuc[p] = make_int2(qo, qn);
}

// Becke weights of kBeckePts staged points of one group: weight *=
// P_own / sum_j P_j over the group's nuclei, P_j = prod_{q != j} s(mu_jq),
// every factor in [0, 1], with P_own and P_n from k_becke_exact.  Lane l of
// every wave takes point l; wave c sums the j of class c (j mod kBeckeLanes)
// in list order, and the first wave adds the classes' sums in order.  Every
// P_j other than P_own and P_n runs over the nuclei nearest-first (the order
// from the group's first point, with the coordinates and pair tables by rank
// from k_becke_order) and is dropped once its running product, an
// upper bound of P_j, falls below 2^-60 of the larger exact term of the sum
// (P_own, P_n): a dropped term lies far below half an ulp of the sum.  The
// distances to the first kBeckeNear nuclei of that order come from LDS
// (formed once per block by the same function), the rest are recomputed.  A
// point whose own nucleus is not among the group's leaves P_own out of the
// sum and the bound; a zero P_own ends the point.  The group is uniform in
// the block, so the loops over j and the nuclei's loads are wave-uniform;
// distances are recomputed instead of stored per point, with the host's
// values.  The host's weigh (grid.cpp) does the same arithmetic, so the
// weights are bitwise its weights.
__global__ __launch_bounds__(kBeckePts * kBeckeLanes) void k_becke(
    const StGroup* __restrict__ groups, const int2* __restrict__ items,
    const StMol* __restrict__ mols, const int* __restrict__ gatoms,
    const int* __restrict__ gpos, const double* __restrict__ ox, const double* __restrict__ oy,
    const double* __restrict__ oz, const double* __restrict__ td, const double* __restrict__ tw,
    const double* __restrict__ ax, const double* __restrict__ ay, const double* __restrict__ az,
    const double* __restrict__ aw, const double* __restrict__ rij,
    const double* __restrict__ sx, const double* __restrict__ sy,
    const double* __restrict__ sz, const double2* __restrict__ bterm,
    const int2* __restrict__ bnuc, double* __restrict__ sw) {
#pragma clang fp contract(off)
  __shared__ double part[kBeckeLanes][kBeckePts];
  __shared__ double dnear[kBeckeNear][kBeckePts];
  const int2 it = items[blockIdx.x];
  const int gi = __builtin_amdgcn_readfirstlane(it.x);
  const int lane = (int)threadIdx.x % kBeckePts;
  const int cls = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kBeckePts);
  const int p = __builtin_amdgcn_readfirstlane(it.y) + lane;
  const StGroup G = groups[gi];
  const bool live = p < G.pt_off + G.npts;
  const int k = G.natoms, kn = min(k, kBeckeNear);
  const StMol m = mols[G.mol];
  const int* at = gatoms + G.atom_off;
  const double* A = ax + m.atom_base;
  const double* B = ay + m.atom_base;
  const double* C = az + m.atom_base;
  const double* Xo = ox + G.atom_off;  // the nuclei's coordinates by rank
  const double* Yo = oy + G.atom_off;
  const double* Zo = oz + G.atom_off;
  double x = 0.0, y = 0.0, z = 0.0;
  if (live) {
    x = sx[p];
    y = sy[p];
    z = sz[p];
    for (int t = cls; t < kn; t += kBeckeLanes) dnear[t][lane] = dist3(x, y, z, Xo[t], Yo[t], Zo[t]);
  }
  __syncthreads();
  double mine = 0.0, sum = 0.0;
  if (live && k > 0) {
    const int na = m.natom;
    const double* W = aw + m.rij_off;
    const double* D = rij + m.rij_off;
    const double2 tm = bterm[p];
    const int2 nq = bnuc[p];
    const int qo = nq.x, qn = nq.y, an = at[qn];
    const double dn = dist3(x, y, z, A[an], B[an], C[an]);
    const double pn = tm.y;
    mine = tm.x;
    if (mine != 0.0) {
      const double thr = ldexp(qo >= 0 ? fmax(mine, pn) : pn, -60);
      for (int j = cls; j < k; j += kBeckeLanes) {
        if (j == qo) {
          sum += mine;
          continue;
        }
        if (j == qn) {
          sum += pn;
          continue;
        }
        const int aj = at[j];
        const double dj = dist3(x, y, z, A[aj], B[aj], C[aj]);
        const double* Dj = D + (long long)aj * na;
        const double* Wj = W + (long long)aj * na;
        if (becke_cell_w(dj, dn, Dj[an], Wj[an]) == 0.0) continue;
        const int tj = gpos[G.atom_off + j];  // j's own rank, skipped
        const double* Dr = td + G.sq_off + (long long)j * k;
        const double* Wr = tw + G.sq_off + (long long)j * k;
        double prod = 1.0;
        for (int t = 0; t < k; ++t) {
          if (t == tj) continue;
          const double dk = t < kBeckeNear ? dnear[t][lane] : dist3(x, y, z, Xo[t], Yo[t], Zo[t]);
          prod *= becke_cell_w(dj, dk, Dr[t], Wr[t]);
          if (prod < thr) {
            prod = 0.0;
            break;
          }
        }
        sum += prod;
      }
    }
  }
  part[cls][lane] = sum;
  __syncthreads();
  if (cls != 0 || !live) return;
  double total = 0.0;
#pragma unroll
  for (int c = 0; c < kBeckeLanes; ++c) total += part[c][lane];
  if (k == 0 || mine == 0.0)
    sw[p] = 0.0;
  else
    sw[p] *= (total == 0.0 ? 0.0 : mine / total);
}

// Points a staged group keeps (all, or those of non-zero weight).
__global__ __launch_bounds__(256) void k_group_count(const StGroup* __restrict__ groups,
                                                     const double* __restrict__ sw, int drop_zero,
                                                     int* __restrict__ count) {
  __shared__ int sc[4];
  const StGroup G = groups[blockIdx.x];
  int c = 0;
  for (int i = threadIdx.x; i < G.npts; i += 256) c += (!drop_zero || sw[G.pt_off + i] != 0.0);
  for (int o = 32; o >= 1; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) sc[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) count[blockIdx.x] = sc[0] + sc[1] + sc[2] + sc[3];
}

// Final group f takes the kept points of staged group src[f] in staged
// order into [pt_off, pt_off + n), then pads to npts_pad with copies of its
// first point of zero weight (the host layout).
struct CompactItem {
  int src, pt_off, n, npts_pad;
};
__global__ __launch_bounds__(256) void k_compact(const CompactItem* __restrict__ items,
                                                 const StGroup* __restrict__ groups,
                                                 const double* __restrict__ sx,
                                                 const double* __restrict__ sy,
                                                 const double* __restrict__ sz,
                                                 const double* __restrict__ sw, int drop_zero,
                                                 double* __restrict__ px, double* __restrict__ py,
                                                 double* __restrict__ pz, double* __restrict__ pw) {
  __shared__ int swave[4];
  __shared__ int sfirst;
  const CompactItem it = items[blockIdx.x];
  const StGroup G = groups[it.src];
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (threadIdx.x == 0) sfirst = -1;
  int base = 0;
  for (int c0 = 0; c0 < G.npts; c0 += 256) {
    const int i = c0 + threadIdx.x;
    const bool keep = i < G.npts && (!drop_zero || sw[G.pt_off + i] != 0.0);
    const unsigned long long mask = __ballot(keep);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) swave[wave] = __popcll(mask);
    __syncthreads();
    int before = base;
    for (int w = 0; w < wave; ++w) before += swave[w];
    if (keep) {
      const int q = it.pt_off + before + rank;
      px[q] = sx[G.pt_off + i];
      py[q] = sy[G.pt_off + i];
      pz[q] = sz[G.pt_off + i];
      pw[q] = sw[G.pt_off + i];
      if (before + rank == 0) sfirst = i;
    }
    base += swave[0] + swave[1] + swave[2] + swave[3];
    __syncthreads();
  }
  const int f = sfirst;
  for (int q = it.n + threadIdx.x; q < it.npts_pad; q += 256) {
    px[it.pt_off + q] = sx[G.pt_off + f];
    py[it.pt_off + q] = sy[G.pt_off + f];
    pz[it.pt_off + q] = sz[G.pt_off + f];
    pw[it.pt_off + q] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// Device grid build: points, cube/sphere groups and significant functions
// ---------------------------------------------------------------------------
// The same points, groups and function sets as the host staging
// (g2g/regenerate_partition.cpp:274-462, g2g/classify_functions.cpp:12-141,
// g2g/partition.cpp:229-254), from the molecule's grid frame (prism, cube
// lattice, sphere shells, radial nodes, shell thresholds; grid_frame on the
// host, O(atoms^2 + shells)).  Every raw point (atom, radial shell, angular
// node) gets the key of its group -- a cube of the prism's lattice, the
// atom's sphere, or none when it lies outside the prism -- and a stable sort
// by key puts each group's points together in generation order, the order
// in which the host appends them.  The arithmetic that decides a point's
// position, weight, cube and a shell's significance is the host's, operation
// for operation (no contraction into fused multiply-adds), so every decision
// is the same.

constexpr int kGenMaxAtoms = 8192;  // atoms of one molecule (LDS flags of the screening)

struct GenAtom {
  double x, y, z;
  long long raw_off;  // first raw point of the atom (batch-wide)
  int mol, local;     // molecule, molecule-local index
  int rad_off;        // first radial node (r1, wrad)
  int sphere_from;    // first radial shell inside the atom's sphere
};
struct GenMol {
  double lo[3], hi[3], lcs;
  unsigned dim[3];
  int ncubes, natom;
  int bin_base;       // first group key of the molecule (cubes, then spheres)
  int atom_base;      // first GenAtom / batch atom
  int ang_off, ngrid; // angular rule
  int sh_off, nsh;    // basis shells (shell tables)
  long long slot_off; // shell slots: key b at slot_off + (b - bin_base) * nsh
  long long aslot_off;  // atom slots: key b at aslot_off + (b - bin_base) * natom
  long long rij_off;    // natom x natom distances
};
struct BinInfo {
  int nsel, m, ns, np, nd, nnuc;
  unsigned long long fhash;
};

__global__ __launch_bounds__(256) void k_fill_int(int* __restrict__ p, int n, int v) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}

// Largest a with atoms[a].raw_off <= r.
__device__ inline int gen_atom_of(const GenAtom* __restrict__ atoms, int n, long long r) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (atoms[mid].raw_off <= r)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Position and weight of raw point r (regenerate:274-300): the atom's centre
// plus the angular node times the radial node, weight wrad * wang.
__device__ inline void gen_point(const GenAtom& A, const GenMol& M, long long r,
                                 const double* __restrict__ r1, const double* __restrict__ wrad,
                                 const double* __restrict__ ex, const double* __restrict__ ey,
                                 const double* __restrict__ ez, const double* __restrict__ ew,
                                 double (&q)[3], double& w, int& sh) {
#pragma clang fp contract(off)
  const long long local = r - A.raw_off;
  sh = (int)(local / M.ngrid);
  const int k = (int)(local - (long long)sh * M.ngrid);
  const double rr = r1[A.rad_off + sh];
  q[0] = A.x + ex[M.ang_off + k] * rr;
  q[1] = A.y + ey[M.ang_off + k] * rr;
  q[2] = A.z + ez[M.ang_off + k] * rr;
  w = wrad[A.rad_off + sh] * ew[M.ang_off + k];
}

// Group key of every raw point (no_group, past the last key, for a point
// outside the prism); value = the raw index.  err gets 1 when a point inside
// the prism falls outside its cube lattice (the host throws).
__global__ __launch_bounds__(256) void k_gen_keys(const GenAtom* __restrict__ atoms, int natoms,
                                                  const GenMol* __restrict__ mols,
                                                  const double* __restrict__ r1,
                                                  const double* __restrict__ wrad,
                                                  const double* __restrict__ ex,
                                                  const double* __restrict__ ey,
                                                  const double* __restrict__ ez,
                                                  const double* __restrict__ ew, long long nraw,
                                                  unsigned no_group, unsigned* __restrict__ keys,
                                                  unsigned* __restrict__ vals,
                                                  int* __restrict__ err) {
#pragma clang fp contract(off)
  const long long r = (long long)blockIdx.x * 256 + threadIdx.x;
  if (r >= nraw) return;
  const GenAtom A = atoms[gen_atom_of(atoms, natoms, r)];
  const GenMol& M = mols[A.mol];
  double q[3], w;
  int sh;
  gen_point(A, M, r, r1, wrad, ex, ey, ez, ew, q, w, sh);
  unsigned key = no_group;
  if (M.lo[0] <= q[0] && q[0] <= M.hi[0] && M.lo[1] <= q[1] && q[1] <= M.hi[1] &&
      M.lo[2] <= q[2] && q[2] <= M.hi[2]) {
    if (sh >= A.sphere_from) {
      key = (unsigned)(M.bin_base + M.ncubes + A.local);
    } else {
      unsigned c[3];
      for (int d = 0; d < 3; ++d) c[d] = (unsigned)floor((q[d] - M.lo[d]) / M.lcs);
      if (c[0] >= M.dim[0] || c[1] >= M.dim[1] || c[2] >= M.dim[2])
        atomicOr(err, 1);
      else
        key = (unsigned)M.bin_base + (c[0] * M.dim[1] + c[1]) * M.dim[2] + c[2];
    }
  }
  keys[r] = key;
  vals[r] = (unsigned)r;
}

// The run of every group key in the sorted keys: bounds[k] its first
// position, bounds[nkeys + k] one past its last (both zero for a key with no
// points), so its point count is their difference.  A thread per sorted
// key writes the bounds where the key differs from a neighbour: no atomics.
__global__ __launch_bounds__(256) void k_gen_bounds(const unsigned* __restrict__ keys,
                                                    long long n, unsigned no_group,
                                                    int* __restrict__ bounds) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const unsigned k = keys[i];
  if (k >= no_group) return;
  if (i == 0 || keys[i - 1] != k) bounds[k] = (int)i;
  if (i == n - 1 || keys[i + 1] != k) bounds[no_group + k] = (int)(i + 1);
}

// Staged point arrays in sorted (group, generation) order: coordinates,
// weight wrad * wang, owning atom (molecule-local).
__global__ __launch_bounds__(256) void k_gen_gather(const GenAtom* __restrict__ atoms, int natoms,
                                                    const GenMol* __restrict__ mols,
                                                    const double* __restrict__ r1,
                                                    const double* __restrict__ wrad,
                                                    const double* __restrict__ ex,
                                                    const double* __restrict__ ey,
                                                    const double* __restrict__ ez,
                                                    const double* __restrict__ ew,
                                                    const unsigned* __restrict__ vals, long long n,
                                                    double* __restrict__ sx, double* __restrict__ sy,
                                                    double* __restrict__ sz, double* __restrict__ sw,
                                                    int* __restrict__ owner) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const long long r = vals[i];
  const GenAtom A = atoms[gen_atom_of(atoms, natoms, r)];
  double q[3], w;
  int sh;
  gen_point(A, mols[A.mol], r, r1, wrad, ex, ey, ez, ew, q, w, sh);
  sx[i] = q[0];
  sy[i] = q[1];
  sz[i] = q[2];
  sw[i] = w;
  owner[i] = A.local;
}

// One block per group key: the significant shells of the group (in shell
// order: s, p, d) into its shell slots, the atoms owning them (ascending)
// into its atom slots, and the counts and FNV-1a hash of the first function
// indices.  The squared distance of an atom to the group is the host's: to
// the cube's box, or to the sphere's surface (zero for the sphere's own
// atom).
__global__ __launch_bounds__(256) void k_gen_screen(const GenMol* __restrict__ mols,
                                                    const int* __restrict__ bin_mol,
                                                    const GenAtom* __restrict__ atoms,
                                                    const double* __restrict__ srad,
                                                    const double* __restrict__ rij,
                                                    const int* __restrict__ sh_first,
                                                    const int* __restrict__ sh_len,
                                                    const int* __restrict__ sh_atom,
                                                    const double* __restrict__ sh_a,
                                                    const double* __restrict__ sh_thr,
                                                    int assign_all, int* __restrict__ slots,
                                                    int* __restrict__ aslots,
                                                    BinInfo* __restrict__ info) {
#pragma clang fp contract(off)
  __shared__ unsigned char sown[kGenMaxAtoms];
  __shared__ int swave[4];
  const int key = blockIdx.x;
  const GenMol M = mols[bin_mol[key]];
  const int lb = key - M.bin_base;
  const bool cube = lb < M.ncubes;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int a = threadIdx.x; a < M.natom; a += 256) sown[a] = 0;
  double corner[3] = {0, 0, 0};
  int sa = -1;
  if (cube) {
    const unsigned ci = (unsigned)lb / (M.dim[1] * M.dim[2]);
    const unsigned cj = ((unsigned)lb / M.dim[2]) % M.dim[1];
    const unsigned ck = (unsigned)lb % M.dim[2];
    corner[0] = M.lo[0] + ci * M.lcs;
    corner[1] = M.lo[1] + cj * M.lcs;
    corner[2] = M.lo[2] + ck * M.lcs;
  } else {
    sa = lb - M.ncubes;
  }
  __syncthreads();
  int* slot = slots + M.slot_off + (long long)lb * M.nsh;
  int base = 0;
  for (int c0 = 0; c0 < M.nsh; c0 += 256) {
    const int k = c0 + threadIdx.x;
    bool keep = false;
    int len = 0;
    if (k < M.nsh) {
      const int s = M.sh_off + k;
      const int at = sh_atom[s];
      len = sh_len[s];
      double d2;
      if (cube) {
        const GenAtom& A = atoms[M.atom_base + at];
        const double p[3] = {A.x, A.y, A.z};
        double v[3];
        for (int d = 0; d < 3; ++d) {
          if (p[d] < corner[d])
            v[d] = corner[d] - p[d];
          else if (p[d] > (corner[d] + M.lcs))
            v[d] = p[d] - (corner[d] + M.lcs);
          else
            v[d] = 0;
        }
        d2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
      } else if (at == sa) {
        d2 = 0;
      } else {
        const double dab = rij[M.rij_off + (long long)at * M.natom + sa];
        const double sr = srad[M.atom_base + sa];
        const double gap = (sr > dab ? 0 : dab - sr);
        d2 = gap * gap;
      }
      keep = assign_all || (len == 1 ? sh_a[s] * d2 < sh_thr[s] : sqrt(d2) < sh_thr[s]);
      if (keep) sown[at] = 1;
    }
    const unsigned long long mask = __ballot(keep);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) swave[wave] = __popcll(mask);
    __syncthreads();
    int before = base;
    for (int w = 0; w < wave; ++w) before += swave[w];
    if (keep) slot[before + rank] = k;
    base += swave[0] + swave[1] + swave[2] + swave[3];
    __syncthreads();
  }
  // atoms owning a kept shell, ascending
  int* aslot = aslots + M.aslot_off + (long long)lb * M.natom;
  int abase = 0;
  for (int a0 = 0; a0 < M.natom; a0 += 256) {
    const int a = a0 + threadIdx.x;
    const bool own = a < M.natom && sown[a];
    const unsigned long long mask = __ballot(own);
    const int rank = __popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0) swave[wave] = __popcll(mask);
    __syncthreads();
    int before = abase;
    for (int w = 0; w < wave; ++w) before += swave[w];
    if (own) aslot[before + rank] = a;
    abase += swave[0] + swave[1] + swave[2] + swave[3];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    BinInfo bi;
    bi.nsel = base;
    bi.nnuc = abase;
    bi.ns = bi.np = bi.nd = bi.m = 0;
    unsigned long long h = 1469598103934665603ull;
    for (int i = 0; i < base; ++i) {
      const int s = M.sh_off + slot[i];
      const int len = sh_len[s];
      bi.m += len;
      if (len == 1)
        ++bi.ns;
      else if (len == 3)
        ++bi.np;
      else
        ++bi.nd;
      const unsigned x = (unsigned)sh_first[s];
      for (int b = 0; b < 4; ++b) {
        h ^= (x >> (8 * b)) & 0xffu;
        h *= 1099511628211ull;
      }
    }
    bi.fhash = h;
    info[key] = bi;
  }
}

// Kept groups' shell and atom lists packed densely for the copy to the host.
struct PackItem {
  int key, nsel, nnuc, soff, aoff, pad_;
};
__global__ __launch_bounds__(256) void k_gen_pack(const PackItem* __restrict__ items,
                                                  const GenMol* __restrict__ mols,
                                                  const int* __restrict__ bin_mol,
                                                  const int* __restrict__ slots,
                                                  const int* __restrict__ aslots,
                                                  int* __restrict__ pk, int* __restrict__ pa) {
  const PackItem it = items[blockIdx.x];
  const GenMol M = mols[bin_mol[it.key]];
  const int lb = it.key - M.bin_base;
  const int* slot = slots + M.slot_off + (long long)lb * M.nsh;
  const int* aslot = aslots + M.aslot_off + (long long)lb * M.natom;
  for (int i = threadIdx.x; i < it.nsel; i += 256) pk[it.soff + i] = slot[i];
  for (int i = threadIdx.x; i < it.nnuc; i += 256) pa[it.aoff + i] = aslot[i];
}

double wall() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

}  // namespace

// ---------------------------------------------------------------------------
// Device-resident molecule: every buffer lives in one arena allocation.
// ---------------------------------------------------------------------------
struct DeviceGrid {
  int M = 0, natom = 0, iexch = 9, nval = 10;  // M, natom: whole batch
  int nmol = 1, max_M = 0, max_ng = 0, max_mpad = 0;
  int fock_slice = kSliceDefault;
  int density_variant = 10;
  int nrowsplit = 0;  // leading density items on the row-split kernel
  long long packed = 0;                 // packed entries of the whole batch
  std::vector<long long> mol_packed;    // per molecule offset, then the total
  DevMol* mols = nullptr;
  int* fn_mol = nullptr;
  int* eidx = nullptr;
  bool gga = true;
  int ngroups = 0, neval = 0, ngather = 0, ndensity = 0, nfock = 0, ngroup_tiles = 0;
  int nforce = 0;                        // force blocks (64 points each)
  long long npts_pad = 0, val_elems = 0, pg_elems = 0;
  char* arena = nullptr;
  size_t arena_bytes = 0;   // used by this grid
  size_t arena_cap = 0;     // allocated (reused by the next grid when it fits)
  size_t pinned_packed = 0; // capacity of h_rho / h_F (entries) and h_e (pairs)
  int pinned_mols = 0;
  double eval_ms = 0.0;
  DevFn* bas = nullptr;
  DevGroup* groups = nullptr;
  EvalItem* eitems = nullptr;
  GatherItem* gitems = nullptr;
  EvalItem* pmitems = nullptr;  // point-major evaluation blocks (split-BF16)
  int npm = 0;
  float* phiT = nullptr;          // function-major phi (split-BF16 Fock operand)
  float* ybuf = nullptr;          // Y = P_g phi per point (split-BF16 forces), point-major;
                                  // allocated at the first forces build (not in the arena)
  size_t ybuf_cap = 0;            // its bytes
  DensityItem* ditems = nullptr;
  FockItem* fitems = nullptr;
  int2* gtiles = nullptr;
  int* fnl = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr, *pw = nullptr;
  double* rho = nullptr;
  double* F = nullptr;
  void* vals = nullptr;   // value planes, FP64 or FP32 by precision
  void* pg = nullptr;     // P_g blocks (per solve), same type
  Precision prec = Precision::FP64;
  bool open = false;      // open shell: rho/F hold alpha then beta packed blocks
  void* pg2 = nullptr;    // beta P_g blocks (open shell)
  double* factor_b = nullptr;  // beta factors (open shell)
  double* sums = nullptr;      // alpha row-dot sums, 10 x npts_pad (open shell)
  double* fg = nullptr;  // F_g blocks (per solve)
  double* factor = nullptr;
  double* epart = nullptr;
  double* fpart = nullptr;
  int* memb_off = nullptr;
  int2* memb = nullptr;
  double* eout = nullptr;
  DensityItem* citems = nullptr;  // force blocks
  double* gpart = nullptr;        // per block and function (3)
  double* gsum = nullptr;         // per group and function (3)
  double* fsum = nullptr;         // per batch function (3): its groups' sums
  int* afn_off = nullptr;         // functions of each batch atom
  int* afn = nullptr;
  double* fout = nullptr;         // forces, natom x 3 (atom-major)
  bool radial = false;            // radial-Hessian layout (split-BF16 variant 5)
  bool inkernel = false;          // basis evaluated in the density kernel (variants 6, 7)
  bool lo = false;                // phi's remainder (lo16), four-piece Y_0 (variants 8-10)
  bool chunked = false;           // one density block per (point tile, row chunk) (variants 10, 11)
  bool presplit = false;          // phi's pieces split once per grid (variant 11)
  unsigned short* pcs = nullptr;  // those pieces: 4 planes of phi_elems BF16
  ChunkItem* chitems = nullptr;   // its blocks in dispatch order (XCD-interleaved)
  int nchunk = 0;
  double* cpart = nullptr;        // per (point tile, chunk): 10 x 64 row-dot sums
  ChunkItem* finitems = nullptr;  // per point tile (k_density_cs_finish)
  int* forder = nullptr;          // split-BF16 Fock blocks: item per block (XCD-interleaved)
  int nfock_order = 0;
  int nfin = 0;
  unsigned short* lo16 = nullptr; // phi's remainder in BF16, point-major (phi_elems)
  bool persistent = false;        // one persistent launch per build (k_build_persist)
  PersistArgs pargs{};            // its queue and tables (device pointers)
  PersistArgs* dpargs = nullptr;  // device copy of pargs, read by the kernel
  int nsync = 0;                  // counters the launch clears first
  int persist_blocks = 0;         // resident blocks of the persistent launch (2 per CU)
  int2* kbsh = nullptr;           // per 32-function K block: first shell, shell count
  int2* sfn = nullptr;            // per group shell: first local function, function count
  int2* tsh = nullptr;            // per 16-function row tile: first shell, shell count
  long long phi_elems = 0;        // elements of the function-major phi plane
  int* shells = nullptr;          // first (batch) function of every group shell
  FnMeta* fmeta = nullptr;        // per padded function of every group
  float* rad = nullptr;           // radial planes
  GatherItem* ritems = nullptr;   // radial evaluation blocks
  int nrad = 0;
  std::vector<int> mol_atom;      // first batch atom of each molecule, then natom
  double* h_rho = nullptr;
  double* h_F = nullptr;
  double* h_e = nullptr;
};

namespace {

struct Carver {
  char* base;
  size_t used = 0;
  template <class T>
  T* take(size_t n) {
    used = (used + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + used);
    used += n * sizeof(T);
    return p;
  }
};

}  // namespace

Engine::Engine(int device) {
  int n = 0;
  XCQ_HIP(hipGetDeviceCount(&n));
  if (n <= 0) throw std::runtime_error("g2g: no HIP device available");
  if (device < 0) XCQ_HIP(hipGetDevice(&device));
  device_ = device;
  XCQ_HIP(hipSetDevice(device_));
  hipStream_t s;
  XCQ_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  stream_ = s;
  hipMemPool_t pool;
  XCQ_HIP(hipDeviceGetDefaultMemPool(&pool, device_));
  uint64_t keep = ~0ull;  // freed blocks stay in the pool until the process ends
  XCQ_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
  for (auto& e : events_) {
    hipEvent_t ev;
    XCQ_HIP(hipEventCreate(&ev));
    e = ev;
  }
}

// Device memory comes from the device's stream-ordered pool (hipMallocAsync /
// hipFreeAsync on the engine stream, release threshold at its maximum): a
// freed arena stays in the pool for the next grid of any engine on the
// device, and freeing does not wait for the whole device, which would stall
// the builds of other molecules in flight on other streams.
// Optional device slab: one allocation per device made up front
// (Engine::reserve_slab), from which every engine on that device carves its
// grid arenas with a first-fit free list on the host.  A serving loop that
// rebuilds and evicts grids of different sizes at every request then makes
// no device allocation at all.  A block is returned only when its engine's
// stream is idle (after a synchronisation), so another engine may reuse it
// at once.
struct Slab {
  std::mutex mu;
  // set once (size first, then base with release order) and never changed,
  // so engines on other threads may test it without the lock
  std::atomic<char*> base{nullptr};
  size_t size = 0;
  std::map<size_t, size_t> free_;  // offset -> length, coalesced
  bool active() const { return base.load(std::memory_order_acquire) != nullptr; }
  bool owns(const void* p) const {
    const char* b = base.load(std::memory_order_acquire);
    return b && p >= b && static_cast<const char*>(p) < b + size;
  }
  char* take(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    std::lock_guard<std::mutex> l(mu);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      const size_t off = it->first, len = it->second;
      free_.erase(it);
      if (len > bytes) free_[off + bytes] = len - bytes;
      return base.load(std::memory_order_relaxed) + off;
    }
    return nullptr;
  }
  void give(char* p, size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    std::lock_guard<std::mutex> l(mu);
    size_t off = p - base.load(std::memory_order_relaxed), len = bytes;
    auto next = free_.lower_bound(off);
    if (next != free_.end() && off + len == next->first) {
      len += next->second;
      next = free_.erase(next);
    }
    if (next != free_.begin()) {
      auto prev = std::prev(next);
      if (prev->first + prev->second == off) {
        off = prev->first;
        len += prev->second;
        free_.erase(prev);
      }
    }
    free_[off] = len;
  }
};
static Slab g_slab[64];

// Allocates `want` bytes (at least `need`) from the pool on stream s.  When
// the pool's cached blocks do not fit the request it returns them to the
// device and tries again, then falls back to `need`.  Returns the size
// obtained.
static size_t pool_alloc(void** p, size_t want, size_t need, hipStream_t s, int device) {
  if (hipMallocAsync(p, want, s) == hipSuccess) return want;
  (void)hipGetLastError();
  hipMemPool_t pool;
  XCQ_HIP(hipDeviceGetDefaultMemPool(&pool, device));
  XCQ_HIP(hipStreamSynchronize(s));
  XCQ_HIP(hipMemPoolTrimTo(pool, 0));
  if (hipMallocAsync(p, want, s) == hipSuccess) return want;
  (void)hipGetLastError();
  XCQ_HIP(hipMallocAsync(p, need, s));
  return need;
}

// Returns an arena to the slab it came from (after the stream's work on it
// has finished) or to the pool (stream-ordered).  Never throws: destructors
// call it.
static void free_arena(char* a, size_t cap, int device, hipStream_t s) {
  if (!a) return;
  Slab& sl = g_slab[device];
  if (sl.owns(a)) {  // another engine may take the block at once
    (void)hipStreamSynchronize(s);
    sl.give(a, cap);
  } else
    (void)hipFreeAsync(a, s);
}

static void release(DeviceGrid* g, hipStream_t s, int device) {
  if (!g) return;
  free_arena(g->arena, g->arena_cap, device, s);
  free_arena(reinterpret_cast<char*>(g->ybuf), g->ybuf_cap, device, s);
  if (g->h_rho) (void)hipHostFree(g->h_rho);
  if (g->h_F) (void)hipHostFree(g->h_F);
  if (g->h_e) (void)hipHostFree(g->h_e);
  delete g;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
  drop_graph();
  release(grid_, (hipStream_t)stream_, device_);
  if (stage_) (void)hipFreeAsync(stage_, (hipStream_t)stream_);
  if (becke_) (void)hipFreeAsync(becke_, (hipStream_t)stream_);
  if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
  for (auto& e : events_)
    if (e) (void)hipEventDestroy((hipEvent_t)e);
  if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
}

size_t Engine::arena_bytes() const { return grid_ ? grid_->arena_bytes : 0; }
double Engine::eval_ms() const { return grid_ ? grid_->eval_ms : 0.0; }

void Engine::load(const Molecule& mol, const GridPartition& part, Precision prec) {
  load_batch({&mol}, {&part}, prec);
}

void Engine::load_batch(const std::vector<const Molecule*>& mols,
                        const std::vector<const GridPartition*>& parts, Precision prec) {
  load_batch_impl(mols, parts, prec, nullptr);
}

// Staged points on the device after the Becke weights (load_staged).
struct DevicePoints {
  const StGroup* groups;  // staged groups
  const double *x, *y, *z, *w;
  int drop_zero;
  std::vector<int> src;   // staged group of each final group (batch order)
};

void Engine::load_batch_impl(const std::vector<const Molecule*>& mols,
                             const std::vector<const GridPartition*>& parts, Precision prec,
                             const DevicePoints* dpts) {
  XCQ_HIP(hipSetDevice(device_));
  check_not_pending("a grid load");
  XCQ_HIP(hipStreamSynchronize((hipStream_t)stream_));
  const double t_plan = wall();
  load_ms_[0] = load_ms_[1] = load_ms_[2] = load_ms_[3] = 0.0;
  drop_graph();
  // the previous grid's arena and pinned buffers are reused when large enough
  // (a rebuilt grid of a moved molecule is about the same size): freeing
  // device memory waits for the whole device, which would stall the builds
  // of other molecules in flight on other streams
  struct Old {
    DeviceGrid* g;
    hipStream_t s;
    int device;
    ~Old() { release(g, s, device); }
  } old{grid_, (hipStream_t)stream_, device_};
  grid_ = nullptr;
  const int nmol = (int)mols.size();
  if (nmol == 0 || parts.size() != mols.size())
    throw std::runtime_error("g2g: a batch needs one partition per molecule");
  for (const Molecule* m : mols)
    if (m->gga != mols[0]->gga || m->iexch != mols[0]->iexch)
      throw std::runtime_error("g2g: the molecules of a batch must share the functional");
  if (tuning_.fock_slice < 128 || tuning_.fock_slice % 128 != 0)
    throw std::runtime_error("g2g: the Fock slice must be a positive multiple of 128 points");
  if (tuning_.bucket_mhi.size() != tuning_.bucket_slice.size())
    throw std::runtime_error("g2g: one Fock slice per size bucket");
  for (size_t i = 0; i < tuning_.bucket_slice.size(); ++i)
    if (tuning_.bucket_slice[i] < 128 || tuning_.bucket_slice[i] % 128 != 0 ||
        (i > 0 && tuning_.bucket_mhi[i] <= tuning_.bucket_mhi[i - 1]))
      throw std::runtime_error("g2g: bucket slices must be positive multiples of 128 points and "
                               "the bucket bounds ascending");
  auto* dg = new DeviceGrid;
  // until the grid is committed below, every exit frees what it holds (its
  // arena -- possibly the previous grid's, handed over -- and pinned buffers)
  struct Guard {
    DeviceGrid* g;
    hipStream_t s;
    int device;
    ~Guard() { release(g, s, device); }
  } guard{dg, (hipStream_t)stream_, device_};
  dg->nmol = nmol;
  dg->fock_slice = tuning_.fock_slice;
  dg->density_variant = tuning_.density_variant;

  if (dg->density_variant < 0 || dg->density_variant > 11)
    throw std::runtime_error("g2g: unknown density kernel variant");
  dg->iexch = mols[0]->iexch;
  dg->gga = mols[0]->gga;
  dg->open = mols[0]->open;
  for (const Molecule* m : mols)
    if (m->open != dg->open)
      throw std::runtime_error("g2g: the molecules of a batch must all be open or all closed shell");
  if (dg->open && !dg->gga) {
    throw std::runtime_error("g2g: open-shell LDA is not provided (LIO's CPU path evaluates no "
                             "functional there and reads unset factors)");
  }
  if (dg->open && prec == Precision::BF16X6 && dg->density_variant != 1 &&
      dg->density_variant != 2 && dg->density_variant < 5) {
    throw std::runtime_error("g2g: open shell in split-BF16 needs density variant 1, 2 or 5-11");
  }
  const int nspin = dg->open ? 2 : 1;
  // LDA keeps the gradient too: the XC forces need it.  The radial-Hessian
  // layout keeps phi and its gradient as planes and rebuilds the second
  // derivatives from per-shell radial sums.
  dg->radial = prec == Precision::BF16X6 && tuning_.density_variant == 5;
  dg->inkernel = prec == Precision::BF16X6 &&
                 (tuning_.density_variant == 6 || tuning_.density_variant == 7);
  dg->lo = prec == Precision::BF16X6 && tuning_.density_variant >= 8;
  // variant 10: variant 8's products, one block per (point tile, row chunk);
  // 11: 10 with phi's four pieces split once per grid
  dg->chunked = prec == Precision::BF16X6 && tuning_.density_variant >= 10;
  dg->presplit = prec == Precision::BF16X6 && tuning_.density_variant == 11;
  // in-kernel evaluation keeps no value planes (phi alone, function-major,
  // for the Fock kernel); variant 8 one more plane (phi's remainder)
  dg->nval = dg->inkernel ? 0 : (dg->gga && !dg->radial) ? 10 : 4;
  dg->prec = prec;
  const size_t vbytes = prec == Precision::FP64 ? sizeof(double) : sizeof(float);
  const bool bf = prec == Precision::BF16X6;
  // split-BF16 K steps are 32 functions; its P_g holds three 2-byte pieces
  const int fn_align = bf ? 32 : kFnAlign;
  const size_t pbytes =
      bf ? (dg->inkernel || dg->density_variant == 8 || dg->chunked ? 4 : 3) *
               sizeof(unsigned short)
         : vbytes;

  // batch numbering: molecule k owns functions [fn_base, fn_base + M) and
  // packed entries [rho_off, rho_off + M (M + 1) / 2)
  std::vector<DevMol> hmol(nmol);
  dg->mol_atom.assign(nmol + 1, 0);
  for (int k = 0; k < nmol; ++k) dg->mol_atom[k + 1] = dg->mol_atom[k] + mols[k]->natom;
  int M = 0, G = 0;
  long long packed = 0;
  for (int k = 0; k < nmol; ++k) {
    hmol[k].M = mols[k]->M;
    hmol[k].fn_base = M;
    hmol[k].rho_off = packed;
    hmol[k].g0 = G;
    hmol[k].ng = parts[k]->ngroups;
    dg->max_ng = std::max(dg->max_ng, parts[k]->ngroups);
    M += mols[k]->M;
    G += parts[k]->ngroups;
    packed += (long long)mols[k]->M * (mols[k]->M + 1) / 2;
    dg->max_M = std::max(dg->max_M, mols[k]->M);
    dg->natom += mols[k]->natom;
  }
  dg->M = M;
  dg->packed = packed;
  dg->ngroups = G;
  dg->mol_packed.resize(nmol + 1);
  for (int k = 0; k < nmol; ++k) dg->mol_packed[k] = hmol[k].rho_off;
  dg->mol_packed[nmol] = packed;

  std::vector<DevFn> bas(M);
  std::vector<int> fn_mol(M);
  // functions of each batch atom, ascending
  std::vector<int> afn_off(dg->natom + 1, 0), afn(M);
  for (int k = 0; k < nmol; ++k)
    for (int f = 0; f < mols[k]->M; ++f) ++afn_off[dg->mol_atom[k] + mols[k]->fn[f].atom + 1];
  for (int a = 0; a < dg->natom; ++a) afn_off[a + 1] += afn_off[a];
  {
    std::vector<int> fill(afn_off.begin(), afn_off.end() - 1);
    for (int k = 0; k < nmol; ++k)
      for (int f = 0; f < mols[k]->M; ++f)
        afn[fill[dg->mol_atom[k] + mols[k]->fn[f].atom]++] = hmol[k].fn_base + f;
  }
  for (int k = 0; k < nmol; ++k) {
    const Molecule& mol = *mols[k];
    for (int f = 0; f < mol.M; ++f) {
      const BasisFn& b = mol.fn[f];
      DevFn& d = bas[hmol[k].fn_base + f];
      std::memcpy(d.a, b.a, sizeof d.a);
      std::memcpy(d.c, b.c, sizeof d.c);
      d.cx = mol.x[b.atom];
      d.cy = mol.y[b.atom];
      d.cz = mol.z[b.atom];
      d.ncont = b.ncont;
      d.lx = b.lx;
      d.ly = b.ly;
      d.lz = b.lz;
      fn_mol[hmol[k].fn_base + f] = k;
    }
  }
  std::vector<DevGroup> groups(G);
  std::vector<int> fnl;
  std::vector<double> hx, hy, hz, hw;
  long long pt = 0, vo = 0, po = 0, fo = 0, go = 0, ro = 0;
  std::vector<int> shells;
  std::vector<FnMeta> fmeta;
  std::vector<int2> tsh, kbsh, sfn;
  int fock_base = 0;
  std::vector<std::vector<int2>> membership(M);
  for (int g = 0, k = 0, gk = 0; g < G; ++g, ++gk) {
    while (gk == parts[k]->ngroups) {
      ++k;
      gk = 0;
    }
    const GridPartition& part = *parts[k];
    DevGroup& d = groups[g];
    d.mol = k;
    d.pad2_ = 0;
    const int n = part.npts[gk], m = part.m[gk];
    d.npts = n;
    d.npts_pad = (n + kPtAlign - 1) / kPtAlign * kPtAlign;
    d.m = m;
    d.m_pad = (m + fn_align - 1) / fn_align * fn_align;
    dg->max_mpad = std::max(dg->max_mpad, d.m_pad);
    d.pt_off = (int)pt;
    d.fn_off = (int)fnl.size();
    d.ntile = (d.m_pad + kFnChunk - 1) / kFnChunk;
    d.nslice = (d.npts_pad + tuning_.slice_for(d.m_pad) - 1) / tuning_.slice_for(d.m_pad);
    d.val_off = vo;
    d.pg_off = po;
    d.fock_base = fock_base;
    d.slen = tuning_.slice_for(d.m_pad);
    d.phi_off = fo;
    d.force_off = go;
    go += (long long)(d.npts_pad + kDensPts - 1) / kDensPts * d.m_pad * 3;
    fock_base += d.ntile * (d.ntile + 1) / 2 * d.nslice;
    for (int i = 0; i < d.m_pad; ++i) {
      const int f = i < m ? hmol[k].fn_base + part.fn[part.fn_off[gk] + i] : -1;
      fnl.push_back(f);
      if (f >= 0) membership[f].push_back(make_int2(g, i));
    }
    d.shell_off = (int)shells.size();
    d.nsh_pad = 0;
    d.rad_off = ro;
    if (dg->radial || dg->inkernel) {
      const Molecule& mol = *mols[k];
      int first = -1;
      const size_t meta0 = fmeta.size();
      for (int i = 0; i < d.m_pad; ++i) {
        FnMeta mt{0.0, 0.0, 0.0, 0.0, 0, 0};
        if (i >= m && i > 0) {  // padding: scale 0, the last real function's shell
          mt = fmeta[meta0 + m - 1];
          mt.scale = 0.0;
        }
        if (i < m) {
          const int fm = part.fn[part.fn_off[gk] + i], f = hmol[k].fn_base + fm;
          const int comp = fm < mol.ns ? 0
                           : fm < mol.ns + 3 * mol.np ? (fm - mol.ns) % 3
                                                      : (fm - mol.ns - 3 * mol.np) % 6;
          if (comp == 0) {
            first = f;
            shells.push_back(f);
          }
          if (first < 0) throw std::runtime_error("g2g: a group starts inside a shell");
          const DevFn& b = bas[f];
          const DevFn& b0 = bas[first];
          const double ratio = b.c[0] / b0.c[0];
          bool same = b.ncont == b0.ncont && b.cx == b0.cx && b.cy == b0.cy && b.cz == b0.cz;
          for (int q = 0; same && q < b.ncont; ++q)
            same = b.a[q] == b0.a[q] &&
                   std::fabs(b.c[q] - ratio * b0.c[q]) <= 1e-12 * std::fabs(b.c[q]);
          if (!same)
            throw std::runtime_error("g2g: density variants 5 and 6 need the functions of a "
                                     "shell to share centre, exponents and contraction");
          mt.cx = b.cx;
          mt.cy = b.cy;
          mt.cz = b.cz;
          mt.scale = ratio;
          mt.code = ((int)shells.size() - 1 - d.shell_off) | (b.lx << 16) | (b.ly << 20) |
                    (b.lz << 24);
        }
        fmeta.push_back(mt);
      }
      const int nsh = (int)shells.size() - d.shell_off;
      d.nsh_pad = (nsh + 3) / 4 * 4;
      for (int q = nsh; q < d.nsh_pad; ++q) shells.push_back(-1);
      // first local function and function count of every shell
      for (int q = 0; q < d.nsh_pad; ++q) sfn.push_back(make_int2(0, 0));
      for (int i = 0; i < m; ++i) {
        int2& e = sfn[d.shell_off + (fmeta[meta0 + i].code & 0xffff)];
        if (e.y == 0) e.x = i;
        ++e.y;
      }
      if (dg->radial) ro += 3LL * d.npts_pad * d.nsh_pad;
      // shells each 16-function row tile and each 32-function K block touch
      auto span = [&](int lo, int len) {
        const int hi = std::min(lo + len, m) - 1;
        const int last = (fmeta[meta0 + std::max(m - 1, 0)].code & 0xffff);
        if (lo > hi) return make_int2(last, 0);
        const int a = fmeta[meta0 + lo].code & 0xffff, b = fmeta[meta0 + hi].code & 0xffff;
        return make_int2(a, b - a + 1);
      };
      for (int t0 = 0; t0 < d.m_pad; t0 += 16) tsh.push_back(span(t0, 16));
      for (int k0 = 0; k0 < d.m_pad; k0 += 32) kbsh.push_back(span(k0, 32));
    }
    for (int p = 0; !dpts && p < d.npts_pad; ++p) {
      const int src = part.pt_off[gk] + (p < n ? p : 0);  // padding repeats point 0 with w = 0
      hx.push_back(part.x[src]);
      hy.push_back(part.y[src]);
      hz.push_back(part.z[src]);
      hw.push_back(p < n ? part.w[src] : 0.0);
    }
    pt += d.npts_pad;
    vo += (long long)dg->nval * d.m_pad * d.npts_pad;
    po += (long long)d.m_pad * d.m_pad;
    if (bf) fo += (long long)d.m_pad * d.npts_pad;
  }
  dg->npts_pad = pt;
  dg->phi_elems = fo;
  dg->val_elems = vo;
  dg->pg_elems = po;
  dg->nfock = fock_base;

  // evaluation, gather and density work: heavy groups first
  std::vector<int> order(G);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return groups[a].m_pad > groups[b].m_pad; });
  std::vector<EvalItem> eitems;
  std::vector<GatherItem> gitems, ritems;
  std::vector<EvalItem> pmitems;
  size_t pm_len = 0;
  // split-BF16 variant 3: groups of more than 64 functions run the row-split
  // kernel on 16-point blocks (listed first), the rest 64-point blocks
  const bool rowsplit = bf && (tuning_.density_variant == 3 || tuning_.density_variant == 4);
  std::vector<DensityItem> ditems, dsmall, citems;
  // the function-major evaluation runs unless the point-major one does
  const bool fm_eval = !bf || dg->inkernel, pm_eval = bf && !dg->inkernel;
  {  // the lists' lengths first: millions of items for large molecules
    size_t ne = 0, ng = 0, npm = 0, nt = 0;
    for (int g = 0; g < G; ++g) {
      const DevGroup& d = groups[g];
      const size_t tiles = (d.npts_pad + kDensPts - 1) / kDensPts;
      nt += tiles;
      if (fm_eval) ne += (size_t)d.m_pad * ((d.npts_pad + kEvalSpan - 1) / kEvalSpan);
      ng += ((size_t)d.m_pad * d.m_pad + 255) / 256;
      if (pm_eval)
        npm += (size_t)((d.npts_pad + kPmPts - 1) / kPmPts) * ((d.m_pad + kPmFns - 1) / kPmFns);
    }
    citems.reserve(nt);
    dsmall.reserve(nt);
    eitems.reserve(ne);
    gitems.reserve(ng);
    // the point-major items XCD-interleaved: block b runs on XCD b % 8 and
    // takes item b / 8 of the list's (b % 8)-th eighth, so the function
    // chunks of one point tile, the rows of the same points, are written
    // through one L2 (items of group -1 pad the eighths)
    pm_len = (npm + kXcds - 1) / kXcds;
    pmitems.assign(kXcds * pm_len, EvalItem{-1, 0, 0, 0});
  }
  size_t pm_next = 0;
  for (int g : order) {
    const DevGroup& d = groups[g];
    for (int p0 = 0; p0 < d.npts_pad; p0 += kDensPts) citems.push_back(DensityItem{g, p0});
    for (int k = 0; fm_eval && k < d.m_pad; ++k)
      for (int p0 = 0; p0 < d.npts_pad; p0 += kEvalSpan) eitems.push_back(EvalItem{g, k, p0, 0});
    for (int e0 = 0; e0 < d.m_pad * d.m_pad; e0 += 256) gitems.push_back(GatherItem{g, e0});
    if (pm_eval)
      for (int p0 = 0; p0 < d.npts_pad; p0 += kPmPts)
        for (int k0 = 0; k0 < d.m_pad; k0 += kPmFns, ++pm_next)
          pmitems[kXcds * (pm_next % pm_len) + pm_next / pm_len] = EvalItem{g, k0, p0, 0};
    if (dg->radial)
      for (int e0 = 0; e0 < d.nsh_pad * d.npts_pad; e0 += kEvalSpan)
        ritems.push_back(GatherItem{g, e0});
    if (rowsplit && d.m_pad > kRowSplitMin)
      for (int p0 = 0; p0 < d.npts_pad; p0 += 16) ditems.push_back(DensityItem{g, p0});
    else
      for (int p0 = 0; p0 < d.npts_pad; p0 += kDensPts) dsmall.push_back(DensityItem{g, p0});
  }
  dg->nrowsplit = rowsplit ? (int)ditems.size() : 0;
  ditems.insert(ditems.end(), dsmall.begin(), dsmall.end());
  // per molecule, its density blocks in ascending block order (the energy
  // reduction order, the same whether the molecule is alone or batched)
  std::vector<int> eidx;
  for (int k = 0; k < nmol; ++k) {
    hmol[k].eb = (int)eidx.size();
    for (int i = 0; i < (int)ditems.size(); ++i)
      if (groups[ditems[i].group].mol == k) eidx.push_back(i);
    hmol[k].ee = (int)eidx.size();
  }
  // variants 10, 11: every point tile's row chunks.  Block b runs on XCD
  // b % kXcds, so the list is the XCDs' lists interleaved: each tile (heavy
  // groups first) goes whole to the XCD with the least work so far, its
  // chunks consecutive there, and the short lists are padded with empty
  // items (a few).  Contiguous runs of tiles per XCD (a group's P_g in one
  // L2) needed thousands of empty items for the uneven item counts and ran
  // slower.
  std::vector<ChunkItem> chitems, finitems;
  long long npart = 0;
  if (dg->chunked) {
    std::vector<std::vector<ChunkItem>> lane(kXcds);
    std::vector<long long> load(kXcds, 0);
    for (int i = 0; i < (int)ditems.size(); ++i) {
      const int mp = groups[ditems[i].group].m_pad, nch = (mp + kFnChunk - 1) / kFnChunk;
      const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
      load[x] += (long long)nch * (mp + 48);  // K steps and the epilogue
      finitems.push_back(ChunkItem{ditems[i].group, ditems[i].p0, 0, i, (int)npart, nch});
      for (int c = 0; c < nch; ++c)
        lane[x].push_back(ChunkItem{ditems[i].group, ditems[i].p0, kFnChunk * c, i, (int)npart, nch});
      npart += nch;
    }
    size_t len = 0;
    for (const auto& l : lane) len = std::max(len, l.size());
    chitems.assign(len * kXcds, ChunkItem{-1, 0, 0, 0, 0, 0});
    for (int x = 0; x < kXcds; ++x)
      for (size_t j = 0; j < lane[x].size(); ++j) chitems[j * kXcds + x] = lane[x][j];
  }
  dg->nchunk = (int)chitems.size();
  // split-BF16 Fock blocks: the tile pairs of each (group, slice), heavy
  // groups first, go whole to the XCD with the fewest blocks so far
  std::vector<int> forder;
  if (bf) {
    std::vector<std::vector<int>> lane(kXcds);
    for (int g : order) {
      const DevGroup& d = groups[g];
      const int pairs = d.ntile * (d.ntile + 1) / 2;
      for (int sl = 0; sl < d.nslice; ++sl) {
        size_t x = 0;
        for (int q = 1; q < kXcds; ++q)
          if (lane[q].size() < lane[x].size()) x = q;
        for (int pr = 0; pr < pairs; ++pr) lane[x].push_back(d.fock_base + pr * d.nslice + sl);
      }
    }
    size_t len = 0;
    for (const auto& l : lane) len = std::max(len, l.size());
    forder.assign(len * kXcds, -1);
    for (int x = 0; x < kXcds; ++x)
      for (size_t j = 0; j < lane[x].size(); ++j) forder[j * kXcds + x] = lane[x][j];
  }
  dg->nfock_order = (int)forder.size();
  dg->nfin = (int)finitems.size();
  // Fock partial k belongs to (group, tile pair, slice) in slot order
  std::vector<FockItem> fitems(fock_base);
  std::vector<int2> gtiles;
  for (int g = 0; g < G; ++g) {
    const DevGroup& d = groups[g];
    int pair = 0;
    for (int ti = 0; ti < d.ntile; ++ti)
      for (int tj = ti; tj < d.ntile; ++tj, ++pair) {
        gtiles.push_back(make_int2(g, pair));
        for (int s = 0; s < d.nslice; ++s)
          fitems[d.fock_base + pair * d.nslice + s] = FockItem{g, ti, tj, s};
      }
  }
  // persistent build: per-group item counts, each group's first tile pair,
  // the packing items and the queue orders of the Fock and tile-sum items
  // (groups in the density order, heavy first)
  dg->persistent = tuning_.persistent != 0 && bf && dg->density_variant == 8 && !dg->open;
  std::vector<int> gneed, dneed, tbase, fperm, tperm;
  std::vector<int2> pitems;
  if (dg->persistent) {
    gneed.assign(G, 0);
    dneed.assign(G, 0);
    tbase.assign(G, 0);
    for (const GatherItem& gi : gitems) gneed[gi.group]++;
    for (const DensityItem& di : ditems) dneed[di.group]++;
    for (int g = 0, t = 0; g < G; ++g) {
      tbase[g] = t;
      t += groups[g].ntile * (groups[g].ntile + 1) / 2;
    }
    for (int g : order) {
      const DevGroup& d = groups[g];
      const int pairs = d.ntile * (d.ntile + 1) / 2;
      for (int q = 0; q < pairs * d.nslice; ++q) fperm.push_back(d.fock_base + q);
      for (int q = 0; q < pairs * 16; ++q) tperm.push_back(16 * tbase[g] + q);
    }
    for (int k = 0; k < nmol; ++k)
      for (int I = 0; I < hmol[k].M; ++I)
        for (int xb = I / 256; xb * 256 < hmol[k].M; ++xb)
          pitems.push_back(make_int2(hmol[k].fn_base + I, xb));
  }
  dg->nsync = dg->persistent ? 4 + 2 * G + (int)gtiles.size() : 0;
  if (dg->persistent) {
    int cus = 0;
    XCQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device_));
    dg->persist_blocks = 2 * std::max(1, cus);
  }
  dg->neval = (int)eitems.size();
  dg->ngather = (int)gitems.size();
  dg->npm = (int)pmitems.size();
  dg->nrad = (int)ritems.size();
  dg->ndensity = (int)ditems.size();
  dg->nforce = (int)citems.size();
  dg->ngroup_tiles = (int)gtiles.size();
  std::vector<int> moff(M + 1, 0);
  std::vector<int2> memb;
  for (int f = 0; f < M; ++f) {
    moff[f] = (int)memb.size();
    memb.insert(memb.end(), membership[f].begin(), membership[f].end());
  }
  moff[M] = (int)memb.size();

  auto plan = [&](Carver& c) {
    dg->bas = c.take<DevFn>(M);
    dg->groups = c.take<DevGroup>(G);
    dg->eitems = c.take<EvalItem>(eitems.size());
    dg->gitems = c.take<GatherItem>(gitems.size());
    dg->pmitems = c.take<EvalItem>(pmitems.size());
    dg->phiT = c.take<float>(fo);
    dg->ditems = c.take<DensityItem>(ditems.size());
    dg->fitems = c.take<FockItem>(fitems.size());
    dg->gtiles = c.take<int2>(gtiles.size());
    dg->fnl = c.take<int>(fnl.size());
    dg->px = c.take<double>(pt);
    dg->py = c.take<double>(pt);
    dg->pz = c.take<double>(pt);
    dg->pw = c.take<double>(pt);
    dg->rho = c.take<double>(nspin * packed);
    dg->F = c.take<double>(nspin * packed);
    dg->vals = c.take<char>(vo * vbytes);
    dg->pg = c.take<char>(po * pbytes);
    dg->fg = c.take<double>(po);
    dg->factor = c.take<double>(pt);
    dg->epart = c.take<double>(2 * ditems.size());
    dg->fpart = c.take<double>((size_t)fitems.size() * kFnChunk * kFnChunk);
    dg->memb_off = c.take<int>(M + 1);
    dg->memb = c.take<int2>(memb.size() + 1);
    dg->mols = c.take<DevMol>(nmol);
    dg->fn_mol = c.take<int>(M);
    dg->eidx = c.take<int>(eidx.size() + 1);
    dg->eout = c.take<double>(2 * nmol);
    dg->citems = c.take<DensityItem>(citems.size());
    dg->gpart = c.take<double>(go);
    dg->gsum = c.take<double>(3 * fnl.size());
    dg->fsum = c.take<double>(3 * (size_t)M + 3);
    dg->afn_off = c.take<int>(afn_off.size());
    dg->afn = c.take<int>(afn.size());
    dg->fout = c.take<double>(3 * (size_t)dg->natom);
    dg->pg2 = c.take<char>(dg->open ? po * pbytes : 0);
    dg->factor_b = c.take<double>(dg->open ? pt : 0);
    dg->sums = c.take<double>(dg->open ? 10 * pt : 0);
    dg->shells = c.take<int>(shells.size() + 1);
    dg->fmeta = c.take<FnMeta>(fmeta.size() + 1);
    dg->rad = c.take<float>(ro);
    dg->ritems = c.take<GatherItem>(ritems.size());
    dg->tsh = c.take<int2>(tsh.size() + 1);
    dg->kbsh = c.take<int2>(kbsh.size() + 1);
    dg->sfn = c.take<int2>(sfn.size() + 1);
    dg->lo16 = c.take<unsigned short>(dg->lo ? fo : 0);
    dg->chitems = c.take<ChunkItem>(chitems.size());
    dg->pcs = c.take<unsigned short>(dg->presplit ? 4 * fo : 0);
    dg->cpart = c.take<double>((size_t)npart * 10 * kDensPts);
    dg->finitems = c.take<ChunkItem>(finitems.size());
    dg->forder = c.take<int>(forder.size());
    PersistArgs& pa = dg->pargs;
    dg->dpargs = c.take<PersistArgs>(1);
    pa.sync = c.take<int>(dg->nsync + 1);
    pa.gneed = c.take<int>(gneed.size() + 1);
    pa.dneed = c.take<int>(dneed.size() + 1);
    pa.tbase = c.take<int>(tbase.size() + 1);
    pa.fperm = c.take<int>(fperm.size() + 1);
    pa.tperm = c.take<int>(tperm.size() + 1);
    pa.pitems = c.take<int2>(pitems.size() + 1);
  };
  Carver sizing{nullptr};
  plan(sizing);
  dg->arena_bytes = sizing.used + 256;
  if (old.g && old.g->arena && old.g->arena_cap >= dg->arena_bytes) {
    dg->arena = old.g->arena;
    dg->arena_cap = old.g->arena_cap;
    old.g->arena = nullptr;
  } else {
    const size_t reclaim = old.g && old.g->arena ? old.g->arena_cap : 0;
    // headroom (12.5 %, 25 % for a grid that outgrew its arena) so that the
    // grids of the next steps of a moving molecule fit the same arena
    const size_t want = dg->arena_bytes + (reclaim ? dg->arena_bytes / 4 : dg->arena_bytes / 8);
    size_t free_b = 0, total_b = 0;
    const bool slab = g_slab[device_].active();
    XCQ_HIP(hipMemGetInfo(&free_b, &total_b));
    {  // memory freed into the pool is available again
      hipMemPool_t pool;
      uint64_t reserved = 0, used = 0;
      XCQ_HIP(hipDeviceGetDefaultMemPool(&pool, device_));
      XCQ_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrReservedMemCurrent, &reserved));
      XCQ_HIP(hipMemPoolGetAttribute(pool, hipMemPoolAttrUsedMemCurrent, &used));
      if (reserved > used) free_b += reserved - used;
    }
    if (!slab && dg->arena_bytes > free_b + reclaim) {
      const size_t need = dg->arena_bytes;
      throw std::runtime_error("g2g: grid needs " + std::to_string(need >> 20) +
                               " MiB of device memory, " + std::to_string(free_b >> 20) +
                               " MiB free");
    }
    if (reclaim) {
      free_arena(old.g->arena, old.g->arena_cap, device_, (hipStream_t)stream_);
      old.g->arena = nullptr;
    }
    Slab& sl = g_slab[device_];
    if (sl.active()) {  // carve from the slab: exact size, no device call
      dg->arena = sl.take(dg->arena_bytes);
      if (!dg->arena) {
        const size_t need = dg->arena_bytes;
        throw std::runtime_error("g2g: device slab full: grid needs " +
                                 std::to_string(need >> 20) + " MiB (out of memory)");
      }
      dg->arena_cap = dg->arena_bytes;
    } else {
      const size_t alloc = want <= free_b + reclaim ? want : dg->arena_bytes;
      void* a = nullptr;
      dg->arena_cap = pool_alloc(&a, alloc, dg->arena_bytes, (hipStream_t)stream_, device_);
      dg->arena = static_cast<char*>(a);
    }
  }
  Carver carve{dg->arena};
  plan(carve);
  // the previous grid's Y buffer serves this one when large enough
  if (old.g && old.g->ybuf && old.g->ybuf_cap >= (size_t)(dg->open ? 2 : 1) * fo * sizeof(float)) {
    dg->ybuf = old.g->ybuf;
    dg->ybuf_cap = old.g->ybuf_cap;
    old.g->ybuf = nullptr;
  }
  const double t_up = wall();
  if (old.g && old.g->h_rho && old.g->pinned_packed >= (size_t)(nspin * packed) &&
      old.g->pinned_mols >= nmol) {
    dg->h_rho = old.g->h_rho;
    dg->h_F = old.g->h_F;
    dg->h_e = old.g->h_e;
    dg->pinned_packed = old.g->pinned_packed;
    dg->pinned_mols = old.g->pinned_mols;
    old.g->h_rho = old.g->h_F = old.g->h_e = nullptr;
  } else {
    dg->pinned_packed = (size_t)(nspin * packed);
    dg->pinned_mols = nmol;
    XCQ_HIP(hipHostMalloc(&dg->h_rho, nspin * packed * sizeof(double), hipHostMallocDefault));
    XCQ_HIP(hipHostMalloc(&dg->h_F, nspin * packed * sizeof(double), hipHostMallocDefault));
    XCQ_HIP(hipHostMalloc(&dg->h_e, 2 * nmol * sizeof(double), hipHostMallocDefault));
  }

  hipStream_t s = (hipStream_t)stream_;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) XCQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up(dg->bas, bas.data(), bas.size() * sizeof(DevFn));
  up(dg->groups, groups.data(), groups.size() * sizeof(DevGroup));
  up(dg->eitems, eitems.data(), eitems.size() * sizeof(EvalItem));
  up(dg->gitems, gitems.data(), gitems.size() * sizeof(GatherItem));
  up(dg->pmitems, pmitems.data(), pmitems.size() * sizeof(EvalItem));
  up(dg->ditems, ditems.data(), ditems.size() * sizeof(DensityItem));
  up(dg->fitems, fitems.data(), fitems.size() * sizeof(FockItem));
  up(dg->gtiles, gtiles.data(), gtiles.size() * sizeof(int2));
  up(dg->fnl, fnl.data(), fnl.size() * sizeof(int));
  up(dg->px, hx.data(), hx.size() * sizeof(double));
  up(dg->py, hy.data(), hy.size() * sizeof(double));
  up(dg->pz, hz.data(), hz.size() * sizeof(double));
  up(dg->pw, hw.data(), hw.size() * sizeof(double));
  up(dg->memb_off, moff.data(), moff.size() * sizeof(int));
  up(dg->memb, memb.data(), memb.size() * sizeof(int2));
  up(dg->mols, hmol.data(), hmol.size() * sizeof(DevMol));
  up(dg->fn_mol, fn_mol.data(), fn_mol.size() * sizeof(int));
  up(dg->eidx, eidx.data(), eidx.size() * sizeof(int));
  up(dg->citems, citems.data(), citems.size() * sizeof(DensityItem));
  up(dg->shells, shells.data(), shells.size() * sizeof(int));
  up(dg->fmeta, fmeta.data(), fmeta.size() * sizeof(FnMeta));
  up(dg->ritems, ritems.data(), ritems.size() * sizeof(GatherItem));
  up(dg->tsh, tsh.data(), tsh.size() * sizeof(int2));
  up(dg->kbsh, kbsh.data(), kbsh.size() * sizeof(int2));
  up(dg->sfn, sfn.data(), sfn.size() * sizeof(int2));
  up(dg->afn_off, afn_off.data(), afn_off.size() * sizeof(int));
  up(dg->afn, afn.data(), afn.size() * sizeof(int));
  up(dg->chitems, chitems.data(), chitems.size() * sizeof(ChunkItem));
  up(dg->finitems, finitems.data(), finitems.size() * sizeof(ChunkItem));
  up(dg->forder, forder.data(), forder.size() * sizeof(int));
  if (dg->persistent) {
    PersistArgs& pa = dg->pargs;
    up(const_cast<int*>(pa.gneed), gneed.data(), gneed.size() * sizeof(int));
    up(const_cast<int*>(pa.dneed), dneed.data(), dneed.size() * sizeof(int));
    up(const_cast<int*>(pa.tbase), tbase.data(), tbase.size() * sizeof(int));
    up(const_cast<int*>(pa.fperm), fperm.data(), fperm.size() * sizeof(int));
    up(const_cast<int*>(pa.tperm), tperm.data(), tperm.size() * sizeof(int));
    up(const_cast<int2*>(pa.pitems), pitems.data(), pitems.size() * sizeof(int2));
    pa.n0 = (int)gitems.size();
    pa.n1 = pa.n0 + (int)ditems.size();
    pa.n2 = pa.n1 + (int)fitems.size();
    pa.n3 = pa.n2 + 16 * (int)gtiles.size();
    pa.n4 = pa.n3 + (int)pitems.size();
    pa.n5 = pa.n4 + nmol;
    pa.ngroups = G;
    pa.ntiles = (int)gtiles.size();
    pa.fgneed = 16 * (int)gtiles.size();
    pa.dallneed = (int)ditems.size();
    pa.gitems = dg->gitems;
    pa.groups = dg->groups;
    pa.mols = dg->mols;
    pa.fnl = dg->fnl;
    pa.pg = static_cast<unsigned short*>(dg->pg);
    pa.ditems = dg->ditems;
    pa.vals = static_cast<const float*>(dg->vals);
    pa.pw = dg->pw;
    pa.factor = dg->factor;
    pa.epart = dg->epart;
    pa.iexch = dg->iexch;
    pa.px = dg->px;
    pa.py = dg->py;
    pa.pz = dg->pz;
    pa.fmeta = dg->fmeta;
    pa.rad = dg->rad;
    pa.lo16 = dg->lo16;
    pa.fitems = dg->fitems;
    pa.phiT = dg->phiT;
    pa.fpart = dg->fpart;
    pa.gtiles = dg->gtiles;
    pa.fg = dg->fg;
    pa.memb_off = dg->memb_off;
    pa.memb = dg->memb;
    pa.fn_mol = dg->fn_mol;
    pa.eidx = dg->eidx;
    pa.eout = dg->eout;
    up(dg->dpargs, &pa, sizeof(PersistArgs));
  }
  if (dpts) {  // points come from the device rebuild: compact them into place
    if ((int)dpts->src.size() != G) throw std::runtime_error("g2g: staged group map mismatch");
    std::vector<CompactItem> ci(G);
    for (int g = 0; g < G; ++g)
      ci[g] = CompactItem{dpts->src[g], groups[g].pt_off, groups[g].npts, groups[g].npts_pad};
    CompactItem* dci = nullptr;
    XCQ_HIP(hipMallocAsync((void**)&dci, std::max<size_t>(1, G) * sizeof(CompactItem), s));
    up(dci, ci.data(), ci.size() * sizeof(CompactItem));
    if (G > 0)
      hipLaunchKernelGGL(k_compact, dim3(G), dim3(256), 0, s, dci, dpts->groups, dpts->x, dpts->y,
                         dpts->z, dpts->w, dpts->drop_zero, dg->px, dg->py, dg->pz, dg->pw);
    XCQ_HIP(hipGetLastError());
    XCQ_HIP(hipFreeAsync(dci, s));
  }
  // basis values of every group, kept for all solves on this grid
  hipEvent_t* ev = reinterpret_cast<hipEvent_t*>(events_);
  const double t_eval = wall();
  XCQ_HIP(hipEventRecord(ev[6], s));
  if (dg->neval > 0 || dg->npm > 0) {
    const dim3 grid(dg->neval), block(kEvalPts);
    if (prec == Precision::FP64) {
      double* v = static_cast<double*>(dg->vals);
      if (dg->gga)
        hipLaunchKernelGGL((k_eval_basis<10, double>), grid, block, 0, s, dg->eitems, dg->groups,
                           dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, 0);
      else
        hipLaunchKernelGGL((k_eval_basis<4, double>), grid, block, 0, s, dg->eitems, dg->groups,
                           dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, 0);
    } else if (prec == Precision::FP32) {
      float* v = static_cast<float*>(dg->vals);
      if (dg->gga)
        hipLaunchKernelGGL((k_eval_basis<10, float>), grid, block, 0, s, dg->eitems, dg->groups,
                           dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, 0);
      else
        hipLaunchKernelGGL((k_eval_basis<4, float>), grid, block, 0, s, dg->eitems, dg->groups,
                           dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, 0);
    } else if (dg->inkernel) {  // phi alone, function-major, for the Fock kernel
      hipLaunchKernelGGL((k_eval_basis<1, float>), grid, block, 0, s, dg->eitems, dg->groups,
                         dg->fnl, dg->bas, dg->px, dg->py, dg->pz, dg->phiT, 1);
    } else {
      float* v = static_cast<float*>(dg->vals);
      if (dg->npm > 0) {
        if (dg->lo && dg->gga)
          hipLaunchKernelGGL((k_eval_pm<10, true>), dim3(dg->npm), dim3(512), 0, s,
                             dg->pmitems, dg->groups, dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v,
                             dg->phiT, dg->lo16);
        else if (dg->lo)
          hipLaunchKernelGGL((k_eval_pm<4, true>), dim3(dg->npm), dim3(512), 0, s, dg->pmitems,
                             dg->groups, dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, dg->phiT, dg->lo16);
        else if (dg->gga && !dg->radial)
          hipLaunchKernelGGL((k_eval_pm<10>), dim3(dg->npm), dim3(512), 0, s,
                             dg->pmitems, dg->groups, dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v,
                             dg->phiT, dg->lo16);
        else
          hipLaunchKernelGGL(k_eval_pm<4>, dim3(dg->npm), dim3(512), 0, s, dg->pmitems, dg->groups,
                             dg->fnl, dg->bas, dg->px, dg->py, dg->pz, v, dg->phiT, dg->lo16);
      }
      if (dg->presplit && dg->ndensity > 0)
        hipLaunchKernelGGL(k_split_pm, dim3(dg->ndensity), dim3(256), 0, s, dg->ditems,
                           dg->groups, v, dg->lo16, dg->pcs);
      if (dg->nrad > 0)
        hipLaunchKernelGGL(k_eval_radial, dim3(dg->nrad), block, 0, s, dg->ritems, dg->groups,
                           dg->shells, dg->bas, dg->px, dg->py, dg->pz, dg->rad);
    }
    XCQ_HIP(hipGetLastError());
  }
  XCQ_HIP(hipEventRecord(ev[7], s));
  XCQ_HIP(hipStreamSynchronize(s));
  float ms = 0.f;
  XCQ_HIP(hipEventElapsedTime(&ms, ev[6], ev[7]));
  dg->eval_ms = ms;
  grid_ = dg;
  guard.g = nullptr;
  const double t_end = wall();
  load_ms_[1] = 1e3 * (t_up - t_plan);
  load_ms_[2] = 1e3 * (t_eval - t_up);
  load_ms_[3] = 1e3 * (t_end - t_eval);
}

// Host side of a staged batch: the staged groups (points on the device), the
// nuclei of each group, the Becke blocks and the atom tables, and per
// molecule the staged groups' metadata (kind, functions, counts, hash).
struct StageHost {
  std::vector<StGroup> sg;
  std::vector<StMol> sm;
  std::vector<int> gatoms;
  std::vector<int2> bitems;  // (staged group, first point) per Becke block
  std::vector<double> ax, ay, az, aw, rij;
  std::vector<GridPartition> meta;  // per molecule (point arrays unused)
};
// Device side: the staged points and the tables of StageHost.
struct StageDev {
  StGroup* groups;
  StMol* mols;
  int2* bitems;
  int *owner, *gatoms, *gord, *count;
  double *ax, *ay, *az, *aw, *rij, *x, *y, *z, *w;
  double2* bterm;  // P_own, P_n of every staged point
  int2* bnuc;      // own and nearest nucleus (list positions) of every staged point
};

namespace {

// Atom tables of one molecule for the Becke weights: positions, the size
// terms of every atom pair and the distances.
void stage_atoms(const Molecule& mol, const std::vector<double>& rij, StageHost& h) {
  StMol m;
  m.atom_base = (int)h.ax.size();
  m.natom = mol.natom;
  m.rij_off = (long long)h.rij.size();
  h.sm.push_back(m);
  h.ax.insert(h.ax.end(), mol.x.begin(), mol.x.end());
  h.ay.insert(h.ay.end(), mol.y.begin(), mol.y.end());
  h.az.insert(h.az.end(), mol.z.begin(), mol.z.end());
  h.rij.insert(h.rij.end(), rij.begin(), rij.end());
  for (int a = 0; a < mol.natom; ++a)
    for (int b = 0; b < mol.natom; ++b) h.aw.push_back(becke_size_term(mol.rm[a], mol.rm[b]));
}

// Carves the tables of h and n_points staged points.
void carve_stage(Carver& c, const StageHost& h, long long n_points, StageDev& d) {
  d.groups = c.take<StGroup>(h.sg.size() + 1);
  d.mols = c.take<StMol>(h.sm.size() + 1);
  d.bitems = c.take<int2>(h.bitems.size() + 1);
  d.owner = c.take<int>(n_points + 1);
  d.gatoms = c.take<int>(h.gatoms.size() + 1);
  d.gord = c.take<int>(h.gatoms.size() + 1);
  d.count = c.take<int>(h.sg.size() + 1);
  d.ax = c.take<double>(h.ax.size() + 1);
  d.ay = c.take<double>(h.ay.size() + 1);
  d.az = c.take<double>(h.az.size() + 1);
  d.aw = c.take<double>(h.aw.size() + 1);
  d.rij = c.take<double>(h.rij.size() + 1);
  d.x = c.take<double>(n_points + 1);
  d.y = c.take<double>(n_points + 1);
  d.z = c.take<double>(n_points + 1);
  d.w = c.take<double>(n_points + 1);
  d.bterm = c.take<double2>(n_points + 1);
  d.bnuc = c.take<int2>(n_points + 1);
}

}  // namespace

// The grow-only staging buffer of the device grid rebuilds, at least `bytes`.
char* Engine::stage_buffer(size_t bytes) {
  if (stage_cap_ < bytes + 256) {
    if (stage_) XCQ_HIP(hipFreeAsync(stage_, (hipStream_t)stream_));
    stage_ = nullptr;
    stage_cap_ = 0;
    const size_t want = bytes + 256 + (bytes + 256) / 4;
    stage_cap_ = pool_alloc(&stage_, want, bytes + 256, (hipStream_t)stream_, device_);
  }
  return static_cast<char*>(stage_);
}

// The grow-only buffer of the Becke tables in near-first order, at least
// `bytes`.
char* Engine::becke_buffer(size_t bytes) {
  if (becke_cap_ < bytes + 256) {
    if (becke_) XCQ_HIP(hipFreeAsync(becke_, (hipStream_t)stream_));
    becke_ = nullptr;
    becke_cap_ = 0;
    const size_t want = bytes + 256 + (bytes + 256) / 4;
    becke_cap_ = pool_alloc(&becke_, want, bytes + 256, (hipStream_t)stream_, device_);
  }
  return static_cast<char*>(becke_);
}

// Becke weights, zero-weight removal and the group counts on the device,
// then the final groups (at least min_points_per_cube points) are compacted
// into the arena by the batch load.
std::vector<GridPartition> Engine::weigh_and_load(const std::vector<const Molecule*>& mols,
                                                  const StageHost& h, const StageDev& d,
                                                  const Options& opt, Precision prec,
                                                  double t_start) {
  const int nmol = (int)mols.size();
  const int nst = (int)h.sg.size();
  hipStream_t s = (hipStream_t)stream_;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) XCQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  // each group's natoms x natoms Becke tables in near-first order, in their
  // own grow-only buffer (sized now that the groups are known)
  std::vector<StGroup> sg = h.sg;
  long long nsq = 0;
  for (StGroup& G : sg) {
    G.sq_off = (int)nsq;
    nsq += (long long)G.natoms * G.natoms;
  }
  if (nsq > (1LL << 31) - 256) throw std::runtime_error("g2g: Becke tables too large");
  const size_t nat = h.gatoms.size() + 1;
  double *td, *tw, *ox, *oy, *oz;
  int* gpos;
  auto carve_becke = [&](Carver& c) {
    td = c.take<double>((size_t)nsq + 1);
    tw = c.take<double>((size_t)nsq + 1);
    ox = c.take<double>(nat);
    oy = c.take<double>(nat);
    oz = c.take<double>(nat);
    gpos = c.take<int>(nat);
  };
  Carver bsize{nullptr};
  carve_becke(bsize);
  Carver bc{becke_buffer(bsize.used)};
  carve_becke(bc);
  up(d.groups, sg.data(), sg.size() * sizeof(StGroup));
  up(d.mols, h.sm.data(), h.sm.size() * sizeof(StMol));
  // the Becke blocks heaviest first (cost ~ natoms^2), so that the longest
  // ones do not start last
  std::vector<int2> bitems = h.bitems;
  std::stable_sort(bitems.begin(), bitems.end(), [&](const int2& a, const int2& b) {
    return h.sg[a.x].natoms > h.sg[b.x].natoms;
  });
  up(d.bitems, bitems.data(), bitems.size() * sizeof(int2));
  up(d.gatoms, h.gatoms.data(), h.gatoms.size() * sizeof(int));
  up(d.ax, h.ax.data(), h.ax.size() * sizeof(double));
  up(d.ay, h.ay.data(), h.ay.size() * sizeof(double));
  up(d.az, h.az.data(), h.az.size() * sizeof(double));
  up(d.aw, h.aw.data(), h.aw.size() * sizeof(double));
  up(d.rij, h.rij.data(), h.rij.size() * sizeof(double));
  if (nst > 0)
    hipLaunchKernelGGL(k_becke_order, dim3(nst), dim3(256), 0, s, d.groups, d.mols, d.gatoms,
                       d.ax, d.ay, d.az, d.aw, d.rij, d.x, d.y, d.z, d.gord, gpos, ox, oy, oz,
                       td, tw);
  if (!h.bitems.empty()) {
    hipLaunchKernelGGL(k_becke_exact, dim3((unsigned)h.bitems.size()), dim3(kBeckePts), 0, s,
                       d.groups, d.bitems, d.mols, d.gatoms, d.owner, d.ax, d.ay, d.az, d.aw,
                       d.rij, d.x, d.y, d.z, d.bterm, d.bnuc);
    hipLaunchKernelGGL(k_becke, dim3((unsigned)h.bitems.size()), dim3(kBeckePts * kBeckeLanes), 0,
                       s, d.groups, d.bitems, d.mols, d.gatoms, gpos, ox, oy, oz, td, tw, d.ax,
                       d.ay, d.az, d.aw, d.rij, d.x, d.y, d.z, d.bterm, d.bnuc, d.w);
  }
  const int drop = opt.remove_zero_weights ? 1 : 0;
  if (nst > 0)
    hipLaunchKernelGGL(k_group_count, dim3(nst), dim3(256), 0, s, d.groups, d.w, drop, d.count);
  XCQ_HIP(hipGetLastError());
  std::vector<int> count(nst);
  if (nst > 0) XCQ_HIP(hipMemcpyAsync(count.data(), d.count, nst * sizeof(int),
                                      hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  const double weights_ms = 1e3 * (wall() - t_start);
  // the final partitions: groups keeping at least min_points_per_cube points
  std::vector<GridPartition> parts(nmol);
  DevicePoints dp;
  dp.groups = d.groups;
  dp.x = d.x;
  dp.y = d.y;
  dp.z = d.z;
  dp.w = d.w;
  dp.drop_zero = drop;
  for (int k = 0, gs = 0; k < nmol; ++k) {
    const GridPartition& in = h.meta[k];
    GridPartition& out = parts[k];
    int pts = 0;
    for (int g = 0; g < in.ngroups; ++g, ++gs) {
      if ((unsigned)count[gs] < opt.min_points_per_cube) continue;
      out.ngroups++;
      out.kind.push_back(in.kind[g]);
      out.npts.push_back(count[gs]);
      out.pt_off.push_back(pts);
      pts += count[gs];
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
      dp.src.push_back(gs);
    }
    out.seconds = in.seconds;
  }
  std::vector<const GridPartition*> pp;
  for (auto& p : parts) pp.push_back(&p);
  load_batch_impl(mols, pp, prec, &dp);
  load_ms_[0] = weights_ms;
  return parts;
}

std::vector<GridPartition> Engine::load_staged(const std::vector<const Molecule*>& mols,
                                              const std::vector<const StagedPartition*>& staged,
                                              const Options& opt, Precision prec) {
  XCQ_HIP(hipSetDevice(device_));
  check_not_pending("a grid load");
  const double t_start = wall();
  const int nmol = (int)mols.size();
  if (nmol == 0 || staged.size() != mols.size())
    throw std::runtime_error("g2g: a staged batch needs one staging per molecule");
  StageHost h;
  std::vector<int> owner;
  std::vector<double> sx, sy, sz, sw;
  for (int k = 0; k < nmol; ++k) {
    const Molecule& mol = *mols[k];
    const StagedPartition& st = *staged[k];
    stage_atoms(mol, st.rij, h);
    const GridPartition& gp = st.groups;
    h.meta.push_back(gp);
    for (int g = 0; g < gp.ngroups; ++g) {
      StGroup G;
      G.pt_off = (int)sx.size();
      G.npts = gp.npts[g];
      G.atom_off = (int)h.gatoms.size();
      G.natoms = st.atom_off[g + 1] - st.atom_off[g];
      G.mol = k;
      G.sq_off = 0;
      h.gatoms.insert(h.gatoms.end(), st.atoms.begin() + st.atom_off[g],
                      st.atoms.begin() + st.atom_off[g + 1]);
      const int b = gp.pt_off[g];
      sx.insert(sx.end(), gp.x.begin() + b, gp.x.begin() + b + G.npts);
      sy.insert(sy.end(), gp.y.begin() + b, gp.y.begin() + b + G.npts);
      sz.insert(sz.end(), gp.z.begin() + b, gp.z.begin() + b + G.npts);
      sw.insert(sw.end(), gp.w.begin() + b, gp.w.begin() + b + G.npts);
      owner.insert(owner.end(), st.owner.begin() + b, st.owner.begin() + b + G.npts);
      for (int o = 0; o < G.npts; o += kBeckePts)
        h.bitems.push_back(make_int2((int)h.sg.size(), G.pt_off + o));
      h.sg.push_back(G);
    }
  }
  const long long npts = (long long)sx.size();
  if (npts > (1LL << 31) - 256) throw std::runtime_error("g2g: too many staged points");
  StageDev d{};
  Carver plan{nullptr};
  carve_stage(plan, h, npts, d);
  Carver c{stage_buffer(plan.used)};
  carve_stage(c, h, npts, d);
  hipStream_t s = (hipStream_t)stream_;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) XCQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up(d.owner, owner.data(), owner.size() * sizeof(int));
  up(d.x, sx.data(), npts * sizeof(double));
  up(d.y, sy.data(), npts * sizeof(double));
  up(d.z, sz.data(), npts * sizeof(double));
  up(d.w, sw.data(), npts * sizeof(double));
  return weigh_and_load(mols, h, d, opt, prec, t_start);
}

// The whole grid rebuild on the device: points, groups and significant
// functions (k_gen_*), then the weights and the load as for a staged batch.
std::vector<GridPartition> Engine::load_generated(const std::vector<const Molecule*>& mols,
                                                  const Options& opt, Precision prec) {
  XCQ_HIP(hipSetDevice(device_));
  check_not_pending("a grid load");
  const double t_start = wall();
  const int nmol = (int)mols.size();
  if (nmol == 0) throw std::runtime_error("g2g: empty batch");
  // host: the grid frames (O(atoms^2 + shells) per molecule) and the tables
  std::vector<GenAtom> gatom;
  std::vector<GenMol> gmol(nmol);
  std::vector<double> r1, wrad, ex, ey, ez, ew, srad, sh_a, sh_thr;
  std::vector<int> sh_first, sh_len, sh_atom, bin_mol;
  StageHost h;
  long long nraw = 0, nslot = 0, naslot = 0;
  std::vector<GridFrame> frames(nmol);
  for (int k = 0; k < nmol; ++k) {
    const Molecule& mol = *mols[k];
    if (mol.natom > kGenMaxAtoms)
      throw std::runtime_error("g2g: the device grid build takes at most " +
                               std::to_string(kGenMaxAtoms) + " atoms per molecule");
    GridFrame& fr = frames[k];
    fr = grid_frame(mol, opt);
    stage_atoms(mol, fr.rij, h);
    GenMol& M = gmol[k];
    for (int d = 0; d < 3; ++d) {
      M.lo[d] = fr.lo[d];
      M.hi[d] = fr.hi[d];
      M.dim[d] = fr.dim[d];
    }
    M.lcs = opt.little_cube_size;
    if (fr.ncubes + (size_t)mol.natom + bin_mol.size() >= (size_t)1 << 30)
      throw std::runtime_error("g2g: too many cubes in the bounding prism");
    M.ncubes = (int)fr.ncubes;
    M.natom = mol.natom;
    M.bin_base = (int)bin_mol.size();
    M.atom_base = (int)gatom.size();
    M.ang_off = (int)ex.size();
    M.ngrid = mol.ngrid;
    M.sh_off = (int)sh_first.size();
    M.nsh = (int)fr.sh_first.size();
    M.slot_off = nslot;
    M.aslot_off = naslot;
    M.rij_off = h.sm[k].rij_off;
    const long long nbins = (long long)fr.ncubes + mol.natom;
    nslot += nbins * M.nsh;
    naslot += nbins * mol.natom;
    bin_mol.insert(bin_mol.end(), (size_t)nbins, k);
    ex.insert(ex.end(), mol.ex.begin(), mol.ex.end());
    ey.insert(ey.end(), mol.ey.begin(), mol.ey.end());
    ez.insert(ez.end(), mol.ez.begin(), mol.ez.end());
    ew.insert(ew.end(), mol.ew.begin(), mol.ew.end());
    for (int a = 0; a < mol.natom; ++a) {
      GenAtom A;
      A.x = mol.x[a];
      A.y = mol.y[a];
      A.z = mol.z[a];
      A.raw_off = nraw;
      A.mol = k;
      A.local = a;
      A.rad_off = (int)r1.size() + fr.rad_off[a];
      A.sphere_from = (int)fr.sphere_from[a];
      gatom.push_back(A);
      nraw += (long long)mol.shells[a] * mol.ngrid;
    }
    r1.insert(r1.end(), fr.r1.begin(), fr.r1.end());
    wrad.insert(wrad.end(), fr.wrad.begin(), fr.wrad.end());
    srad.insert(srad.end(), fr.srad.begin(), fr.srad.end());
    sh_first.insert(sh_first.end(), fr.sh_first.begin(), fr.sh_first.end());
    sh_len.insert(sh_len.end(), fr.sh_len.begin(), fr.sh_len.end());
    sh_atom.insert(sh_atom.end(), fr.sh_atom.begin(), fr.sh_atom.end());
    sh_a.insert(sh_a.end(), fr.sh_a.begin(), fr.sh_a.end());
    sh_thr.insert(sh_thr.end(), fr.sh_thr.begin(), fr.sh_thr.end());
  }
  if (nraw >= (1LL << 31) - 256) throw std::runtime_error("g2g: too many grid points");
  const int nbins = (int)bin_mol.size();
  const double frame_s = wall() - t_start;
  int bits = 1;  // keys 0 .. nbins (nbins: no group)
  while ((1u << bits) <= (unsigned)nbins) ++bits;
  // one staging plan: generation scratch, then the Becke tables at their
  // upper bounds (groups <= keys, nuclei lists <= keys x atoms)
  size_t sort_bytes = 0;
  XCQ_HIP((hipError_t)radix_sort_pairs(nullptr, &sort_bytes, nullptr, nullptr, nullptr, nullptr,
                                       (int)nraw, bits, stream_));
  struct Gen {
    GenAtom* atoms;
    GenMol* mols;
    double *r1, *wrad, *ex, *ey, *ez, *ew, *srad, *sh_a, *sh_thr;
    int *sh_first, *sh_len, *sh_atom, *bin_mol, *count, *err, *slots, *aslots, *pk, *pa;
    unsigned *keys, *vals, *skeys, *svals;
    BinInfo* info;
    PackItem* pitems;
    void* sort;
  } g{};
  StageHost bound;  // sizes only
  bound.sg.resize(nbins);
  bound.sm = h.sm;
  bound.gatoms.resize((size_t)naslot);
  bound.bitems.resize((size_t)(nraw / kBeckePts + nbins));
  bound.ax = h.ax;
  bound.ay = h.ay;
  bound.az = h.az;
  bound.aw = h.aw;
  bound.rij = h.rij;
  StageDev d{};
  auto carve = [&](Carver& c) {
    g.atoms = c.take<GenAtom>(gatom.size());
    g.mols = c.take<GenMol>(nmol);
    g.r1 = c.take<double>(r1.size());
    g.wrad = c.take<double>(wrad.size());
    g.ex = c.take<double>(ex.size());
    g.ey = c.take<double>(ey.size());
    g.ez = c.take<double>(ez.size());
    g.ew = c.take<double>(ew.size());
    g.srad = c.take<double>(srad.size());
    g.sh_a = c.take<double>(sh_a.size() + 1);
    g.sh_thr = c.take<double>(sh_thr.size() + 1);
    g.sh_first = c.take<int>(sh_first.size() + 1);
    g.sh_len = c.take<int>(sh_len.size() + 1);
    g.sh_atom = c.take<int>(sh_atom.size() + 1);
    g.bin_mol = c.take<int>(nbins);
    g.count = c.take<int>(2 * nbins);
    g.err = c.take<int>(1);
    g.slots = c.take<int>(nslot + 1);
    g.aslots = c.take<int>(naslot + 1);
    g.pk = c.take<int>(nslot + 1);
    g.pa = c.take<int>(naslot + 1);
    g.keys = c.take<unsigned>(nraw);
    g.vals = c.take<unsigned>(nraw);
    g.skeys = c.take<unsigned>(nraw);
    g.svals = c.take<unsigned>(nraw);
    g.info = c.take<BinInfo>(nbins);
    g.pitems = c.take<PackItem>(nbins);
    g.sort = c.take<char>(sort_bytes);
    carve_stage(c, bound, nraw, d);
  };
  Carver plan{nullptr};
  carve(plan);
  Carver c{stage_buffer(plan.used)};
  carve(c);
  hipStream_t s = (hipStream_t)stream_;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) XCQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up(g.atoms, gatom.data(), gatom.size() * sizeof(GenAtom));
  up(g.mols, gmol.data(), gmol.size() * sizeof(GenMol));
  up(g.r1, r1.data(), r1.size() * sizeof(double));
  up(g.wrad, wrad.data(), wrad.size() * sizeof(double));
  up(g.ex, ex.data(), ex.size() * sizeof(double));
  up(g.ey, ey.data(), ey.size() * sizeof(double));
  up(g.ez, ez.data(), ez.size() * sizeof(double));
  up(g.ew, ew.data(), ew.size() * sizeof(double));
  up(g.srad, srad.data(), srad.size() * sizeof(double));
  up(g.sh_a, sh_a.data(), sh_a.size() * sizeof(double));
  up(g.sh_thr, sh_thr.data(), sh_thr.size() * sizeof(double));
  up(g.sh_first, sh_first.data(), sh_first.size() * sizeof(int));
  up(g.sh_len, sh_len.data(), sh_len.size() * sizeof(int));
  up(g.sh_atom, sh_atom.data(), sh_atom.size() * sizeof(int));
  up(g.bin_mol, bin_mol.data(), bin_mol.size() * sizeof(int));
  // run bounds and the error flag start at zero (a fill kernel, no memset)
  hipLaunchKernelGGL(k_fill_int, dim3((2 * nbins + 255) / 256), dim3(256), 0, s, g.count,
                     2 * nbins, 0);
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(256), 0, s, g.err, 1, 0);
  const int natoms = (int)gatom.size();
  if (nraw > 0)
    hipLaunchKernelGGL(k_gen_keys, dim3((unsigned)((nraw + 255) / 256)), dim3(256), 0, s,
                       g.atoms, natoms, g.mols, g.r1, g.wrad, g.ex, g.ey, g.ez, g.ew, nraw,
                       (unsigned)nbins, g.keys, g.vals, g.err);
  XCQ_HIP(hipGetLastError());
  XCQ_HIP((hipError_t)radix_sort_pairs(g.sort, &sort_bytes, g.keys, g.skeys, g.vals, g.svals,
                                       (int)nraw, bits, stream_));
  if (nraw > 0)
    hipLaunchKernelGGL(k_gen_bounds, dim3((unsigned)((nraw + 255) / 256)), dim3(256), 0, s,
                       g.skeys, nraw, (unsigned)nbins, g.count);
  up(d.rij, h.rij.data(), h.rij.size() * sizeof(double));
  hipLaunchKernelGGL(k_gen_screen, dim3(nbins), dim3(256), 0, s, g.mols, g.bin_mol, g.atoms,
                     g.srad, d.rij, g.sh_first, g.sh_len, g.sh_atom, g.sh_a, g.sh_thr,
                     opt.assign_all_functions ? 1 : 0, g.slots, g.aslots, g.info);
  XCQ_HIP(hipGetLastError());
  std::vector<int> count(2 * nbins);
  std::vector<BinInfo> info(nbins);
  int err = 0;
  XCQ_HIP(hipMemcpyAsync(count.data(), g.count, 2 * nbins * sizeof(int), hipMemcpyDeviceToHost,
                         s));
  XCQ_HIP(hipMemcpyAsync(info.data(), g.info, nbins * sizeof(BinInfo), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipMemcpyAsync(&err, g.err, sizeof(int), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  if (err) throw std::runtime_error("grid point outside the cube lattice of the bounding prism");
  for (int b = 0; b < nbins; ++b) count[b] = count[nbins + b] - count[b];
  // the staged groups, as the host keeps them: cubes with significant
  // functions and enough points, in lattice order, then the spheres
  std::vector<long long> start(nbins + 1, 0);
  for (int b = 0; b < nbins; ++b) start[b + 1] = start[b] + count[b];
  std::vector<PackItem> items;
  std::vector<int> kept;
  int soff = 0, aoff = 0;
  for (int b = 0; b < nbins; ++b) {
    const GenMol& M = gmol[bin_mol[b]];
    const bool cube = b - M.bin_base < M.ncubes;
    const BinInfo& bi = info[b];
    if (cube && (bi.nsel == 0 || (unsigned)count[b] < opt.min_points_per_cube)) continue;
    if (!cube && (opt.sphere_radius <= 0 || (unsigned)count[b] < opt.min_points_per_cube))
      continue;
    items.push_back(PackItem{b, bi.nsel, bi.nnuc, soff, aoff, 0});
    kept.push_back(b);
    soff += bi.nsel;
    aoff += bi.nnuc;
  }
  std::vector<int> pk(soff), pa(aoff);
  if (!items.empty()) {
    up(g.pitems, items.data(), items.size() * sizeof(PackItem));
    hipLaunchKernelGGL(k_gen_pack, dim3((unsigned)items.size()), dim3(256), 0, s, g.pitems,
                       g.mols, g.bin_mol, g.slots, g.aslots, g.pk, g.pa);
    XCQ_HIP(hipGetLastError());
    if (soff) XCQ_HIP(hipMemcpyAsync(pk.data(), g.pk, soff * sizeof(int), hipMemcpyDeviceToHost, s));
    if (aoff) XCQ_HIP(hipMemcpyAsync(pa.data(), g.pa, aoff * sizeof(int), hipMemcpyDeviceToHost, s));
  }
  // staged points in sorted order (the points of every key, kept or not)
  const long long nvalid = start[nbins];
  if (nvalid > 0)
    hipLaunchKernelGGL(k_gen_gather, dim3((unsigned)((nvalid + 255) / 256)), dim3(256), 0, s,
                       g.atoms, natoms, g.mols, g.r1, g.wrad, g.ex, g.ey, g.ez, g.ew, g.svals,
                       nvalid, d.x, d.y, d.z, d.w, d.owner);
  XCQ_HIP(hipGetLastError());
  XCQ_HIP(hipStreamSynchronize(s));
  h.meta.assign(nmol, GridPartition());
  for (size_t i = 0; i < kept.size(); ++i) {
    const int b = kept[i], k = bin_mol[b];
    const GenMol& M = gmol[k];
    const BinInfo& bi = info[b];
    const PackItem& it = items[i];
    GridPartition& o = h.meta[k];
    o.ngroups++;
    o.kind.push_back(b - M.bin_base < M.ncubes ? 0 : 1);
    o.npts.push_back(count[b]);
    o.pt_off.push_back(0);
    o.fn_off.push_back((int)o.fn.size());
    for (int q = 0; q < bi.nsel; ++q) {
      const int sh = pk[it.soff + q];
      const GridFrame& fr = frames[k];
      for (int cpt = 0; cpt < fr.sh_len[sh]; ++cpt) o.fn.push_back(fr.sh_first[sh] + cpt);
    }
    o.m.push_back(bi.m);
    o.nshells.push_back(bi.nsel);
    o.ns.push_back(bi.ns);
    o.np.push_back(bi.np);
    o.nd.push_back(bi.nd);
    o.nnuc.push_back(bi.nnuc);
    o.fhash.push_back(bi.fhash);
    StGroup G;
    G.pt_off = (int)start[b];
    G.npts = count[b];
    G.atom_off = (int)h.gatoms.size();
    G.natoms = bi.nnuc;
    G.mol = k;
    G.sq_off = 0;
    h.gatoms.insert(h.gatoms.end(), pa.begin() + it.aoff, pa.begin() + it.aoff + bi.nnuc);
    for (int o2 = 0; o2 < G.npts; o2 += kBeckePts)
      h.bitems.push_back(make_int2((int)h.sg.size(), G.pt_off + o2));
    h.sg.push_back(G);
  }
  for (int k = 0; k < nmol; ++k) h.meta[k].seconds = frame_s;
  return weigh_and_load(mols, h, d, opt, prec, t_start);
}

void Engine::download_points(double* x, double* y, double* z, double* w) {
  if (!grid_) throw std::runtime_error("g2g: no grid");
  const DeviceGrid& g = *grid_;
  hipStream_t s = (hipStream_t)stream_;
  const size_t b = (size_t)g.npts_pad * sizeof(double);
  if (x) XCQ_HIP(hipMemcpyAsync(x, g.px, b, hipMemcpyDeviceToHost, s));
  if (y) XCQ_HIP(hipMemcpyAsync(y, g.py, b, hipMemcpyDeviceToHost, s));
  if (z) XCQ_HIP(hipMemcpyAsync(z, g.pz, b, hipMemcpyDeviceToHost, s));
  if (w) XCQ_HIP(hipMemcpyAsync(w, g.pw, b, hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
}

long long Engine::padded_points() const { return grid_ ? grid_->npts_pad : 0; }

// Phase boundaries are marked with ev[k] when ev is given (not while a graph
// is being captured).
static void mark(hipEvent_t* ev, int k, hipStream_t s) {
  if (ev) XCQ_HIP(hipEventRecord(ev[k], s));
}

// Gather and density + functional launches of one spin.  spin 0 closed
// shell, 1 the alpha pass (P_g from the first packed block into pg), 2 the
// beta pass (the second block into pg2).
template <class T>
static void launch_density(DeviceGrid& g, hipStream_t s, const double* d_rho, int spin,
                           hipEvent_t* ev) {
  T* pg = static_cast<T*>(spin == 2 ? g.pg2 : g.pg);
  const double* rho = spin == 2 ? d_rho + g.packed : d_rho;
  const T* vals = static_cast<const T*>(g.vals);
  if (g.ngather > 0)
    hipLaunchKernelGGL(k_gather_density<T>, dim3(g.ngather), dim3(256), 0, s, g.gitems, g.groups,
                       g.mols, g.fnl, rho, pg);
  if (spin < 2) mark(ev, 1, s);
  if (g.ndensity > 0) {
    if (g.gga)
      hipLaunchKernelGGL((k_density<4, T>), dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                         pg, vals, g.pw, g.factor, g.epart, g.iexch, spin, g.factor_b, g.sums,
                         g.npts_pad);
    else
      hipLaunchKernelGGL((k_density<1, T>), dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                         pg, vals, g.pw, g.factor, g.epart, g.iexch, spin, g.factor_b, g.sums,
                         g.npts_pad);
  }
  XCQ_HIP(hipGetLastError());
}

// yb: where the density kernel leaves Y = P_g phi for the force pass (or null).
static void launch_density_bf(DeviceGrid& g, hipStream_t s, const double* d_rho, int spin,
                              hipEvent_t* ev, float* yb = nullptr) {
  unsigned short* pg3 = static_cast<unsigned short*>(spin == 2 ? g.pg2 : g.pg);
  const double* rho = spin == 2 ? d_rho + g.packed : d_rho;
  const float* vals = static_cast<const float*>(g.vals);
  if (g.inkernel) {  // variants 6 (128-function row chunks) and 7 (64)
    if (g.ngather > 0)
      hipLaunchKernelGGL(k_gather_density_bf<4>, dim3(g.ngather), dim3(256), 0, s, g.gitems,
                         g.groups, g.mols, g.fnl, rho, pg3);
    if (spin < 2) mark(ev, 1, s);
    float* y = yb ? yb + (spin == 2 ? g.phi_elems : 0) : nullptr;
    if (g.ndensity > 0) {
      const dim3 grid(g.ndensity), block(256);
#define XCQ_EV(NT, R)                                                                         \
  hipLaunchKernelGGL((k_density_ev<NT, R>), grid, block, 0, s, g.ditems, g.groups, pg3, g.fmeta, \
                     g.shells, g.sfn, g.kbsh, g.tsh, g.bas, g.px, g.py, g.pz, g.pw, g.factor, g.epart,  \
                     g.iexch, spin, g.factor_b, g.sums, g.npts_pad, y)
      if (g.gga && g.density_variant == 6)
        XCQ_EV(4, 128);
      else if (g.gga)
        XCQ_EV(4, 64);
      else if (g.density_variant == 6)
        XCQ_EV(1, 128);
      else
        XCQ_EV(1, 64);
#undef XCQ_EV
    }
    XCQ_HIP(hipGetLastError());
    return;
  }
  if (g.lo) {  // variants 8 and 9: phi's remainder plane (8: four-piece P_g)
    const bool four = g.density_variant == 8 || g.chunked;
    if (g.ngather > 0) {
      if (four)
        hipLaunchKernelGGL(k_gather_density_bf<4>, dim3(g.ngather), dim3(256), 0, s, g.gitems,
                           g.groups, g.mols, g.fnl, rho, pg3);
      else
        hipLaunchKernelGGL(k_gather_density_bf<3>, dim3(g.ngather), dim3(256), 0, s, g.gitems,
                           g.groups, g.mols, g.fnl, rho, pg3);
    }
    if (spin < 2) mark(ev, 1, s);
    float* y = yb ? yb + (spin == 2 ? g.phi_elems : 0) : nullptr;

    if (g.chunked && g.nchunk > 0) {
      const dim3 fgrid((g.nfin + 3) / 4);
      if (g.gga) {
        if (g.presplit)
          hipLaunchKernelGGL((k_density_cs<4, true>), dim3(g.nchunk), dim3(256), 0, s, g.chitems,
                             g.groups, pg3, vals, y, g.lo16, g.cpart, g.pcs);
        else
          hipLaunchKernelGGL((k_density_cs<4, false>), dim3(g.nchunk), dim3(256), 0, s, g.chitems,
                             g.groups, pg3, vals, y, g.lo16, g.cpart, g.pcs);
        if (g.nfin > 0)
          hipLaunchKernelGGL(k_density_cs_finish<4>, fgrid, dim3(256), 0, s, g.finitems, g.nfin,
                             g.groups, g.pw, g.factor, g.epart, g.iexch, spin, g.factor_b, g.sums,
                             g.npts_pad, g.cpart);
      } else {
        if (g.presplit)
          hipLaunchKernelGGL((k_density_cs<1, true>), dim3(g.nchunk), dim3(256), 0, s, g.chitems,
                             g.groups, pg3, vals, y, g.lo16, g.cpart, g.pcs);
        else
          hipLaunchKernelGGL((k_density_cs<1, false>), dim3(g.nchunk), dim3(256), 0, s, g.chitems,
                             g.groups, pg3, vals, y, g.lo16, g.cpart, g.pcs);
        if (g.nfin > 0)
          hipLaunchKernelGGL(k_density_cs_finish<1>, fgrid, dim3(256), 0, s, g.finitems, g.nfin,
                             g.groups, g.pw, g.factor, g.epart, g.iexch, spin, g.factor_b, g.sums,
                             g.npts_pad, g.cpart);
      }
    } else if (g.ndensity > 0) {
#define XCQ_LO(NT, L)                                                                           \
  hipLaunchKernelGGL((k_density_bf2<NT, 64, false, L>), dim3(g.ndensity), dim3(256), 0, s,        \
                     g.ditems, g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,        \
                     g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, y, g.lo16)
      if (g.gga && four)
        XCQ_LO(4, 1);
      else if (g.gga)
        XCQ_LO(4, 2);
      else if (four)
        XCQ_LO(1, 1);
      else
        XCQ_LO(1, 2);
#undef XCQ_LO
    }
    XCQ_HIP(hipGetLastError());
    return;
  }
  if (g.ngather > 0)
    hipLaunchKernelGGL(k_gather_density_bf<3>, dim3(g.ngather), dim3(256), 0, s, g.gitems,
                       g.groups, g.mols, g.fnl, rho, pg3);
  if (spin < 2) mark(ev, 1, s);
  if (g.ndensity > 0) {
    if (g.density_variant == 1) {
      if (g.gga)
        hipLaunchKernelGGL((k_density_bf2<4, 128>), dim3(g.ndensity), dim3(256), 0, s, g.ditems,
                           g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
      else
        hipLaunchKernelGGL((k_density_bf2<1, 128>), dim3(g.ndensity), dim3(256), 0, s, g.ditems,
                           g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
    } else if (g.density_variant == 2) {
      if (g.gga)
        hipLaunchKernelGGL((k_density_bf2<4, 64>), dim3(g.ndensity), dim3(256), 0, s, g.ditems,
                           g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
      else
        hipLaunchKernelGGL((k_density_bf2<1, 64>), dim3(g.ndensity), dim3(256), 0, s, g.ditems,
                           g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
    } else if (g.density_variant == 5) {
      if (g.gga)
        hipLaunchKernelGGL((k_density_bf2<4, 64, true>), dim3(g.ndensity), dim3(256), 0, s,
                           g.ditems, g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
      else
        hipLaunchKernelGGL((k_density_bf2<1, 64, true>), dim3(g.ndensity), dim3(256), 0, s,
                           g.ditems, g.groups, pg3, vals, g.pw, g.factor, g.epart, g.iexch, spin,
                           g.factor_b, g.sums, g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, yb, nullptr);
    } else if (g.density_variant == 3 || g.density_variant == 4) {
      const int nr = g.nrowsplit, ns = g.ndensity - nr;
      const DensityItem* si = g.ditems + nr;
      double* se = g.epart + 2 * (size_t)nr;
      if (g.gga) {
        if (nr > 0 && g.density_variant == 3)
          hipLaunchKernelGGL((k_density_bf3<4, 6>), dim3(nr), dim3(256), 0, s, g.ditems, g.groups,
                             pg3, vals, g.pw, g.factor, g.epart, g.iexch);
        if (nr > 0 && g.density_variant == 4)
          hipLaunchKernelGGL((k_density_bf3<4, 4>), dim3(nr), dim3(256), 0, s, g.ditems, g.groups,
                             pg3, vals, g.pw, g.factor, g.epart, g.iexch);
        if (ns > 0)
          hipLaunchKernelGGL((k_density_bf2<4, 64>), dim3(ns), dim3(256), 0, s, si, g.groups, pg3,
                             vals, g.pw, g.factor, se, g.iexch, 0, g.factor_b, g.sums,
                             g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, (float*)nullptr, nullptr);
      } else {
        if (nr > 0)
          hipLaunchKernelGGL((k_density_bf3<1, 6>), dim3(nr), dim3(256), 0, s, g.ditems, g.groups,
                             pg3, vals, g.pw, g.factor, g.epart, g.iexch);
        if (ns > 0)
          hipLaunchKernelGGL((k_density_bf2<1, 64>), dim3(ns), dim3(256), 0, s, si, g.groups, pg3,
                             vals, g.pw, g.factor, se, g.iexch, 0, g.factor_b, g.sums,
                             g.npts_pad, g.px, g.py, g.pz, g.fmeta, g.rad, (float*)nullptr, nullptr);
      }
    } else {
      if (g.gga)
        hipLaunchKernelGGL(k_density_bf<4>, dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                           pg3, vals, g.pw, g.factor, g.epart, g.iexch);
      else
        hipLaunchKernelGGL(k_density_bf<1>, dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                           pg3, vals, g.pw, g.factor, g.epart, g.iexch);
    }
  }
  XCQ_HIP(hipGetLastError());
}

// F_g blocks of one spin from its factors w v (slices, then slice sums).
static void launch_fock(DeviceGrid& g, hipStream_t s, const double* factor) {
  if (g.nfock <= 0) return;
  if (g.prec == Precision::FP64)
    hipLaunchKernelGGL(k_fock<double>, dim3(g.nfock), dim3(256), 0, s, g.fitems, g.groups,
                       static_cast<const double*>(g.vals), factor, g.fpart);
  else if (g.prec == Precision::FP32)
    hipLaunchKernelGGL(k_fock<float>, dim3(g.nfock), dim3(256), 0, s, g.fitems, g.groups,
                       static_cast<const float*>(g.vals), factor, g.fpart);
  else
    hipLaunchKernelGGL(k_fock_bf, dim3(g.nfock_order), dim3(256), 0, s, g.fitems, g.groups,
                       g.phiT, factor, g.fpart, g.forder);
  hipLaunchKernelGGL(k_fock_group, dim3(16 * g.ngroup_tiles), dim3(256), 0, s, g.groups, g.gtiles,
                     g.fpart, g.fg);
  XCQ_HIP(hipGetLastError());
}

// XC forces after the density launches (P_g blocks and factors in place):
// per-block partials and group sums per spin, then per-atom sums into
// d_forces (natom x 3, atom-major).
// Closed-shell split-BF16 builds whose density kernel is k_density_bf2
// (variants 1, 2, 5) leave Y for the forces in ybuf.
static bool ystore_forces(const DeviceGrid& g) {
  if (g.inkernel || g.lo) return g.ybuf != nullptr;  // both spins when open
  return g.prec == Precision::BF16X6 && !g.open && g.ybuf &&
         (g.density_variant == 1 || g.density_variant == 2 || g.density_variant == 5);
}

static void launch_forces(DeviceGrid& g, hipStream_t s, double* d_forces) {
  const int nspin = g.open ? 2 : 1;
  for (int sp = 0; sp < nspin; ++sp) {
    const void* pg = sp == 1 ? g.pg2 : g.pg;
    const double* factor = sp == 1 ? g.factor_b : g.factor;
    if (g.nforce > 0) {
      if (g.prec == Precision::FP64)
        hipLaunchKernelGGL(k_force<double>, dim3(g.nforce), dim3(256), 0, s, g.citems, g.groups,
                           static_cast<const double*>(pg), static_cast<const double*>(g.vals),
                           factor, g.gpart);
      else if (g.prec == Precision::FP32)
        hipLaunchKernelGGL(k_force<float>, dim3(g.nforce), dim3(256), 0, s, g.citems, g.groups,
                           static_cast<const float*>(pg), static_cast<const float*>(g.vals),
                           factor, g.gpart);
      else if (g.inkernel)
        hipLaunchKernelGGL(k_force_ev, dim3(g.nforce), dim3(256), 0, s, g.citems, g.groups,
                           g.ybuf + (sp == 1 ? g.phi_elems : 0), g.fmeta, g.shells, g.tsh, g.bas,
                           g.px, g.py, g.pz, factor, g.gpart);
      else if (ystore_forces(g))
        hipLaunchKernelGGL(k_force_y, dim3(g.nforce), dim3(256), 0, s, g.citems, g.groups,
                           g.ybuf + (sp == 1 ? g.phi_elems : 0),
                           static_cast<const float*>(g.vals), factor, g.gpart);
      else
        hipLaunchKernelGGL(k_force_bf, dim3(g.nforce), dim3(256), 0, s, g.citems, g.groups,
                           static_cast<const unsigned short*>(pg),
                           static_cast<const float*>(g.vals), factor, g.gpart);
    }
    if (g.ngroups > 0)
      hipLaunchKernelGGL(k_force_group, dim3(g.ngroups, (3 * g.max_mpad + 255) / 256), dim3(256),
                         0, s, g.groups, g.gpart, g.gsum, sp);
  }
  if (g.M > 0)
    hipLaunchKernelGGL(k_force_fn, dim3((g.M + 255) / 256), dim3(256), 0, s, g.memb_off, g.memb,
                       g.groups, g.gsum, g.M, g.fsum);
  if (g.natom > 0)
    hipLaunchKernelGGL(k_force_atom, dim3((g.natom + 63) / 64), dim3(64), 0, s, g.afn_off, g.afn,
                       g.fsum, g.natom, d_forces);
  XCQ_HIP(hipGetLastError());
}

// The whole build: gather, density + functional, Fock, packed reduction,
// per-molecule energies, and the XC forces when d_forces is given.  Every
// launch covers all molecules of the batch.  Open shell: d_rho and d_fock
// hold the alpha then the beta packed blocks; the density runs once per spin
// and the Fock matrix is built and packed per spin.
static void launch_all(DeviceGrid& g, hipStream_t s, const double* d_rho, double* d_fock,
                       hipEvent_t* ev, double* d_forces = nullptr) {
  mark(ev, 0, s);
  if (g.persistent && d_fock) {
    // one launch for the whole build (the counters cleared by a fill kernel,
    // not a memset, so the pair can be captured in a graph); the HIP-event
    // phases then read: expand = clearing, density = the whole build
    const PersistArgs& pa = g.pargs;
    float* yb = d_forces && ystore_forces(g) ? g.ybuf : nullptr;
    hipLaunchKernelGGL(k_fill_int, dim3((g.nsync + 255) / 256), dim3(256), 0, s, pa.sync, g.nsync,
                       0);
    mark(ev, 1, s);
    const int nblocks = std::max(1, std::min(g.persist_blocks, pa.n5));
    if (g.gga)
      hipLaunchKernelGGL(k_build_persist<4>, dim3(nblocks), dim3(256), 0, s, g.dpargs, pa.sync,
                         d_rho, d_fock, yb);
    else
      hipLaunchKernelGGL(k_build_persist<1>, dim3(nblocks), dim3(256), 0, s, g.dpargs, pa.sync,
                         d_rho, d_fock, yb);
    XCQ_HIP(hipGetLastError());
    mark(ev, 2, s);
    mark(ev, 3, s);
    if (d_forces) launch_forces(g, s, d_forces);
    mark(ev, 4, s);
    return;
  }
  const int nspin = g.open ? 2 : 1;
  // closed-shell split-BF16 forces take Y = P_g phi from the density pass
  const bool ystore = d_forces && ystore_forces(g);
  for (int sp = 0; sp < nspin; ++sp) {
    const int spin = g.open ? sp + 1 : 0;
    if (g.prec == Precision::FP64)
      launch_density<double>(g, s, d_rho, spin, ev);
    else if (g.prec == Precision::FP32)
      launch_density<float>(g, s, d_rho, spin, ev);
    else
      launch_density_bf(g, s, d_rho, spin, ev, ystore ? g.ybuf : nullptr);
  }
  mark(ev, 2, s);
  for (int sp = 0; d_fock && sp < nspin; ++sp) {
    launch_fock(g, s, sp == 1 ? g.factor_b : g.factor);
    if (sp == 0) mark(ev, 3, s);
    if (g.max_ng <= kPackMapMax)
      hipLaunchKernelGGL(k_fock_pack, dim3((g.max_M + 127) / 128, g.M), dim3(128),
                         (size_t)g.max_ng * sizeof(int), s, g.memb_off, g.memb, g.groups, g.mols,
                         g.fn_mol, g.fg, d_fock + sp * g.packed);
    else
      hipLaunchKernelGGL(k_fock_pack_merge, dim3((g.max_M + 127) / 128, g.M), dim3(128), 0, s,
                         g.memb_off, g.memb, g.groups, g.mols, g.fn_mol, g.fg,
                         d_fock + sp * g.packed);
  }
  if (!d_fock) mark(ev, 3, s);
  hipLaunchKernelGGL(k_energy, dim3(g.nmol), dim3(1024), 0, s, g.epart, g.mols, g.eidx, g.eout);
  XCQ_HIP(hipGetLastError());
  if (d_forces) launch_forces(g, s, d_forces);
  mark(ev, 4, s);
}

// The split-BF16 force passes read Y = P_g phi, which the density kernel
// leaves in a buffer of its own (one FP32 value per padded point and function,
// per spin): allocated at the first forces build of a grid, from the device
// slab when one is reserved, else from the stream-ordered pool.
static void ensure_ybuf(DeviceGrid& g, hipStream_t s, int device) {
  if (g.prec != Precision::BF16X6 || g.ybuf) return;
  const size_t bytes = (size_t)(g.open ? 2 : 1) * g.phi_elems * sizeof(float) + 256;
  Slab& sl = g_slab[device];
  void* p = nullptr;
  if (sl.active()) {
    p = sl.take(bytes);
    if (!p)
      throw std::runtime_error("g2g: device slab full: the force pass needs " +
                               std::to_string(bytes >> 20) + " MiB (out of memory)");
    g.ybuf_cap = bytes;
  } else {
    g.ybuf_cap = pool_alloc(&p, bytes, bytes, s, device);
  }
  g.ybuf = static_cast<float*>(p);
}

SolveResult Engine::solve_device(const double* d_rho, double* d_fock, bool want_energy,
                                 double* d_forces) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  if (d_forces) ensure_ybuf(*grid_, (hipStream_t)stream_, device_);
  launch_all(*grid_, (hipStream_t)stream_, d_rho, d_fock, reinterpret_cast<hipEvent_t*>(events_),
             d_forces);
  (void)want_energy;
  return SolveResult{};
}

// Scratch of solve_device_static: the split-BF16 force pass's Y = P_g phi
// (FP32, one value per padded point and function, per spin) in a file-scope
// device array instead of a buffer allocated at the first forces build.
// Each device holds its own copy from the module's load until the process
// ends; every call on a device shares it and leaves its contents behind.
constexpr size_t kStaticScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// Device address of this device's scratch (looked up once per device, with
// that device current; the host-side name of the array is not an address).
static void* scratch_address(int device) {
  if (device < 0 || device >= 64) throw std::runtime_error("g2g: device index out of range");
  std::lock_guard<std::mutex> l(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    void* p = nullptr;
    XCQ_HIP(hipSetDevice(device));
    XCQ_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[device] = p;
  }
  return g_scratch_ptr[device];
}

size_t Engine::static_scratch_bytes() { return kStaticScratchBytes; }

SolveResult Engine::solve_device_static(const double* d_rho, double* d_fock, double* d_forces) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  DeviceGrid& g = *grid_;
  const bool ystore = g.inkernel || g.lo || (!g.open && (g.density_variant == 1 ||
                                                  g.density_variant == 2 ||
                                                  g.density_variant == 5));
  if (!d_forces || g.prec != Precision::BF16X6 || !ystore) {
    // no Y buffer on these paths: no scratch beyond the grid's arena
    launch_all(g, (hipStream_t)stream_, d_rho, d_fock, reinterpret_cast<hipEvent_t*>(events_),
               d_forces);
    return SolveResult{};
  }
  const size_t need = (size_t)(g.open ? 2 : 1) * (size_t)g.phi_elems * sizeof(float);
  if (need > kStaticScratchBytes)
    throw std::runtime_error("g2g: this grid's force pass needs " + std::to_string(need >> 20) +
                             " MiB of scratch; the static scratch holds " +
                             std::to_string(kStaticScratchBytes >> 20) +
                             " MiB (use the allocating entry point)");
  float* y = static_cast<float*>(scratch_address(device_));
  XCQ_HIP(hipSetDevice(device_));
  float* own = g.ybuf;  // the grid's own buffer (if any) is left untouched
  g.ybuf = y;
  try {
    launch_all(g, (hipStream_t)stream_, d_rho, d_fock, reinterpret_cast<hipEvent_t*>(events_),
               d_forces);
  } catch (...) {
    g.ybuf = own;
    throw;
  }
  g.ybuf = own;
  return SolveResult{};
}

void Engine::capture_graph(const double* d_rho, double* d_fock) {
  if (!grid_) throw std::runtime_error("g2g: graph capture before the grid was built");
  XCQ_HIP(hipSetDevice(device_));
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipStreamSynchronize(s));
  drop_graph();
  hipGraph_t graph = nullptr;
  XCQ_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  try {
    launch_all(*grid_, s, d_rho, d_fock, nullptr);
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  XCQ_HIP(hipStreamEndCapture(s, &graph));
  hipGraphExec_t exec = nullptr;
  const hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  XCQ_HIP(e);
  graph_ = exec;
}

void Engine::unload() {
  XCQ_HIP(hipSetDevice(device_));
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipStreamSynchronize(s));
  pending_ = false;  // a queued build's result is dropped with the grid
  drop_graph();
  release(grid_, s, device_);
  grid_ = nullptr;
  if (stage_) (void)hipFreeAsync(stage_, s);
  stage_ = nullptr;
  stage_cap_ = 0;
  if (becke_) (void)hipFreeAsync(becke_, s);
  becke_ = nullptr;
  becke_cap_ = 0;
  // the frees have run: the memory is back in the pool for any stream
  XCQ_HIP(hipStreamSynchronize(s));
}

void Engine::replay_graph() {
  if (!graph_) throw std::runtime_error("g2g: no captured graph");
  XCQ_HIP(hipGraphLaunch((hipGraphExec_t)graph_, (hipStream_t)stream_));
}

void Engine::drop_graph() {
  if (graph_) (void)hipGraphExecDestroy((hipGraphExec_t)graph_);
  graph_ = nullptr;
}

void Engine::read_energy(double* energy, double* electrons) {
  std::vector<double> e(2 * (size_t)std::max(1, molecules()));
  read_energies(e.data());
  if (energy) *energy = e[0];
  if (electrons) *electrons = e[1];
}

void Engine::read_energies(double* out) {
  if (!grid_) throw std::runtime_error("g2g: no grid");
  hipStream_t s = (hipStream_t)stream_;
  const size_t bytes = 2 * grid_->nmol * sizeof(double);
  XCQ_HIP(hipMemcpyAsync(grid_->h_e, grid_->eout, bytes, hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  std::memcpy(out, grid_->h_e, bytes);
}

int Engine::molecules() const { return grid_ ? grid_->nmol : 0; }

int Engine::atoms(int k) const {
  if (!grid_ || k < 0 || k > grid_->nmol) throw std::runtime_error("g2g: molecule out of range");
  return grid_->mol_atom[k];
}

long long Engine::packed_offset(int k) const {
  if (!grid_ || k < 0 || k > grid_->nmol) throw std::runtime_error("g2g: molecule out of range");
  return grid_->mol_packed[k];
}

void Engine::kernel_times(double ms[4]) {
  XCQ_HIP(hipStreamSynchronize((hipStream_t)stream_));
  hipEvent_t* ev = reinterpret_cast<hipEvent_t*>(events_);
  for (int k = 0; k < 4; ++k) {
    float t = 0.f;
    XCQ_HIP(hipEventElapsedTime(&t, ev[k], ev[k + 1]));
    ms[k] = t;
  }
}

void Engine::check_not_pending(const char* what) const {
  if (pending_)
    throw std::runtime_error(std::string("g2g: ") + what +
                             " while an asynchronous build is queued: collect it first");
}

void Engine::solve_async(const double* rho_packed) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  check_not_pending("a second asynchronous build");
  if (grid_->open) throw std::runtime_error("g2g: asynchronous builds are closed shell only");
  DeviceGrid& g = *grid_;
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipSetDevice(device_));
  const size_t packed = (size_t)g.packed;
  std::memcpy(g.h_rho, rho_packed, packed * sizeof(double));
  XCQ_HIP(hipMemcpyAsync(g.rho, g.h_rho, packed * sizeof(double), hipMemcpyHostToDevice, s));
  solve_device(g.rho, g.F, true);
  XCQ_HIP(hipMemcpyAsync(g.h_F, g.F, packed * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipMemcpyAsync(g.h_e, g.eout, 2 * g.nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  pending_ = true;
}

bool Engine::idle() const {
  const hipError_t e = hipStreamQuery((hipStream_t)stream_);
  if (e == hipSuccess) return true;
  if (e == hipErrorNotReady) return false;
  throw std::runtime_error(std::string("HIP error ") + hipGetErrorString(e) + " (stream query)");
}

SolveResult Engine::collect(double* fock_accum) {
  if (!grid_) throw std::runtime_error("g2g: no grid");
  if (!pending_) throw std::runtime_error("g2g: collect without a queued asynchronous build");
  DeviceGrid& g = *grid_;
  XCQ_HIP(hipStreamSynchronize((hipStream_t)stream_));
  pending_ = false;
  if (fock_accum)
    for (long long k = 0; k < g.packed; ++k) fock_accum[k] += g.h_F[k];
  SolveResult res;
  res.energy = g.h_e[0];
  res.electrons = g.h_e[1];
  return res;
}

SolveResult Engine::solve(const double* rho_packed, double* fock_accum, bool want_energy,
                          double* forces_accum, int forces_ld) {
  if (grid_ && grid_->open)
    throw std::runtime_error("g2g: this grid is open shell: solve it with both spin densities");
  return solve_open(rho_packed, nullptr, fock_accum, nullptr, want_energy, forces_accum, forces_ld);
}

SolveResult Engine::solve_open(const double* rho_a, const double* rho_b, double* fock_a,
                               double* fock_b, bool want_energy, double* forces_accum,
                               int forces_ld) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  check_not_pending("a synchronous build");
  const double t0 = wall();
  DeviceGrid& g = *grid_;
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipSetDevice(device_));
  if (forces_accum && forces_ld < g.natom)
    throw std::runtime_error("g2g: the force array's leading dimension is below the atom count");
  if (g.open && (!rho_b || (fock_a != nullptr) != (fock_b != nullptr)))
    throw std::runtime_error("g2g: an open-shell solve needs both spin densities (and both Fock "
                             "matrices or neither)");
  const size_t packed = (size_t)g.packed;
  const int nspin = g.open ? 2 : 1;
  std::memcpy(g.h_rho, rho_a, packed * sizeof(double));
  if (g.open) std::memcpy(g.h_rho + packed, rho_b, packed * sizeof(double));
  XCQ_HIP(hipMemcpyAsync(g.rho, g.h_rho, nspin * packed * sizeof(double), hipMemcpyHostToDevice, s));
  SolveResult res = solve_device(g.rho, fock_a ? g.F : nullptr, want_energy,
                                 forces_accum ? g.fout : nullptr);
  if (fock_a)
    XCQ_HIP(hipMemcpyAsync(g.h_F, g.F, nspin * packed * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipMemcpyAsync(g.h_e, g.eout, 2 * g.nmol * sizeof(double), hipMemcpyDeviceToHost, s));
  std::vector<double> f(forces_accum ? 3 * (size_t)g.natom : 0);
  if (forces_accum)
    XCQ_HIP(hipMemcpyAsync(f.data(), g.fout, f.size() * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  if (fock_a)
    for (size_t k = 0; k < packed; ++k) fock_a[k] += g.h_F[k];
  if (fock_b)
    for (size_t k = 0; k < packed; ++k) fock_b[k] += g.h_F[packed + k];
  // Fortran layout of the caller's forces: column c of atom a at c * ld + a
  if (forces_accum)
    for (int a = 0; a < g.natom; ++a)
      for (int c = 0; c < 3; ++c) forces_accum[(size_t)c * forces_ld + a] += f[3 * (size_t)a + c];
  res.energy = g.h_e[0];
  res.electrons = g.h_e[1];
  hipEvent_t* ev = reinterpret_cast<hipEvent_t*>(events_);
  for (int k = 0; k < 4; ++k) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, ev[k], ev[k + 1]) == hipSuccess) res.kernel_ms[k] = ms;
  }
  res.seconds = wall() - t0;
  return res;
}

bool Engine::open_shell() const { return grid_ && grid_->open; }

void Engine::reserve_slab(int device, size_t bytes) {
  if (device < 0) XCQ_HIP(hipGetDevice(&device));
  if (device >= 64) throw std::runtime_error("g2g: device out of range");
  Slab& sl = g_slab[device];
  std::lock_guard<std::mutex> l(sl.mu);
  if (sl.active()) {
    if (sl.size >= bytes) return;  // a second server in the same process
    throw std::runtime_error("g2g: the device slab is already reserved with " +
                             std::to_string(sl.size >> 20) + " MiB");
  }
  int cur = 0;
  XCQ_HIP(hipGetDevice(&cur));
  XCQ_HIP(hipSetDevice(device));
  {  // blocks earlier grids returned to the stream-ordered pool go back first
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      (void)hipDeviceSynchronize();
      (void)hipMemPoolTrimTo(pool, 0);
    }
  }
  void* p = nullptr;
  const hipError_t e = hipMalloc(&p, bytes);
  (void)hipSetDevice(cur);
  XCQ_HIP(e);
  sl.free_.clear();
  sl.free_[0] = bytes;
  sl.size = bytes;
  sl.base.store(static_cast<char*>(p), std::memory_order_release);
}

}  // namespace xcq
