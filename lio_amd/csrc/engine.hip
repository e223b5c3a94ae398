// Device engine for the XC Fock build on MI355X (gfx950), FP64 path.
//
// One solve = five launches on one stream, all reading one device arena:
//   k_expand_density   packed rho (off-diagonals x2) -> dense symmetric P
//   k_density          per 16-point tile: Y = [phi; dphi/dx; dphi/dy; dphi/dz] P
//                      on FP64 MFMA (16x16x4), basis values computed in-kernel,
//                      row-dot epilogue -> rho, grad rho, Hessian of rho ->
//                      PBE/LDA -> factor_p = w_p v_p, per-tile Exc and N.
//                      Also stores phi(p, j) for the Fock pass.
//   k_fock             F_g tile += phi^T diag(factor) phi over a 256-point
//                      slice of the group (FP64 MFMA), one partial per slice
//   k_fock_group       sums the slices of each tile in fixed order
//   k_fock_pack        packed F_IJ = sum over groups holding I and J, in fixed
//                      group order (no float atomics: bitwise reproducible)
// plus a one-block reduction of the per-tile energies.
//
// The maths follows g2g/cpu/iteration.cpp:116-291 (closed shell): with P the
// true symmetric density (packed off-diagonals halved),
//   rho = sum_i phi_i Y_i,  d_a rho = 2 sum_i d_a phi_i Y_i,
//   d_a d_b rho = 2 sum_i (d_ab phi_i Y_i + d_a phi_i (Y_b)_i),
//   F_ij += sum_p phi_i phi_j w_p y2a_p,  Exc = sum_p rho_p w_p (ex + ec).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
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

constexpr int kTile = 16;        // points per density tile (x4 derivative rows)
constexpr int kBlock = 64;       // columns per density column block / Fock tile
constexpr int kChunk = 16;       // K chunk staged through LDS
constexpr int kSlice = 256;      // points per Fock slice (split-K)

typedef double d4 __attribute__((ext_vector_type(4)));

struct DevFn {
  double a[kMaxContractions];
  double c[kMaxContractions];
  double cx, cy, cz;
  int ncont, lx, ly, lz;
};

struct DevGroup {
  int pt_off;      // first padded point
  int npts;        // real points
  int npts_pad;    // multiple of kTile
  int fn_off;      // first entry of the padded function list
  int m;           // real functions
  int m_pad;       // multiple of 16
  int ntile;       // 64-column tiles per side
  int nslice;      // Fock slices
  long long phi_off;   // phi block: npts_pad x m_pad
  long long fg_off;    // F_g block: m_pad x m_pad
  int fock_base;   // first Fock item
  int pad_;
};

struct DensityItem {
  int group;
  int p0;  // first point (group-local)
};

struct FockItem {
  int group;
  int ti, tj;
  int slice;
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
// gradient, 10 adds the Hessian (xx, yy, zz, xy, xz, yz).  Primitives with
// a r^2 > 70 are skipped as in the reference.
template <int NV>
__device__ inline void eval_basis(const DevFn& b, double px, double py, double pz, double* o) {
  const double vx = px - b.cx, vy = py - b.cy, vz = pz - b.cz;
  const double r2 = vx * vx + vy * vy + vz * vz;
  double t = 0.0, tg = 0.0, th = 0.0;
  const int nc = b.ncont;
  for (int k = 0; k < nc; ++k) {
    const double a = b.a[k];
    if (a * r2 > 70.0) continue;
    const double e = exp(-a * r2) * b.c[k];
    t += e;
    if (NV > 1) tg += e * a;
    if (NV > 4) th += e * (a * a);
  }
  const Axis X = axis(vx, b.lx), Y = axis(vy, b.ly), Z = axis(vz, b.lz);
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

__device__ inline double mfma64(double a, double b, d4& c) {
  c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  return 0.0;
}

// Packed index of (i, j), i <= j, column-major lower triangle (partition.cpp:132-134).
__device__ __host__ inline long long packed_index(long long i, long long j, long long M) {
  return i * M - (i * (i - 1)) / 2 + (j - i);
}

__global__ void k_expand_density(const double* __restrict__ rho, double* __restrict__ P, int M) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)M * M) return;
  const int i = (int)(idx / M), j = (int)(idx % M);
  const int lo = i < j ? i : j, hi = i < j ? j : i;
  const double v = rho[packed_index(lo, hi, M)];
  P[idx] = (i == j) ? v : 0.5 * v;
}

// One 16-point tile of one group.  256 threads = 4 waves.  The FP64 16x16x4
// MFMA leaves D[r][c] in lane (r % 4) * 16 + c, item r / 4 (lane l holds
// rows l/16 + 4j); A-operand row r of row-tile rt is therefore the
// derivative t = r / 4 at point 4 rt + r % 4, which puts the 4 derivative
// rows of one point in the 4 accumulator items of one lane.
__global__ __launch_bounds__(256) void k_density_gga(
    const DensityItem* __restrict__ items, const DevGroup* __restrict__ groups,
    const int* __restrict__ fnl, const DevFn* __restrict__ bas, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ pw,
    const double* __restrict__ P, int M, double* __restrict__ phi, double* __restrict__ factor,
    double* __restrict__ epart) {
  __shared__ double sA[kTile][4][kChunk + 1];
  __shared__ double sB[kChunk][kBlock + 1];
  __shared__ double sPart[4][kTile][10];
  __shared__ int sCol[kBlock];

  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int* fl = fnl + g.fn_off;
  const int pbase = g.pt_off + it.p0;

  for (int k = tid; k < 4 * kTile * 10; k += 256) (&sPart[0][0][0])[k] = 0.0;

  for (int c0 = 0; c0 < g.m_pad; c0 += kBlock) {
    __syncthreads();
    if (tid < kBlock) sCol[tid] = (c0 + tid < g.m_pad) ? fl[c0 + tid] : -1;
    const bool active = (c0 + 16 * wave) < g.m_pad;
    d4 acc[4];
    for (int rt = 0; rt < 4; ++rt) acc[rt] = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.m_pad; k0 += kChunk) {
      __syncthreads();
      {  // A: phi and gradient of function k0+kk at point pp
        const int pp = tid >> 4, kk = tid & 15;
        const int f = fl[k0 + kk];
        double v[4] = {0.0, 0.0, 0.0, 0.0};
        if (f >= 0) {
          const int q = pbase + pp;
          eval_basis<4>(bas[f], px[q], py[q], pz[q], v);
        }
        for (int t = 0; t < 4; ++t) sA[pp][t][kk] = v[t];
      }
      for (int e = tid; e < kChunk * kBlock; e += 256) {  // B: P rows k, columns of the block
        const int kk = e / kBlock, cc = e % kBlock;
        const int fr = fl[k0 + kk], fc = sCol[cc];
        sB[kk][cc] = (fr >= 0 && fc >= 0) ? P[(long long)fr * M + fc] : 0.0;
      }
      __syncthreads();
      if (active) {
        for (int ks = 0; ks < kChunk / 4; ++ks) {
          const double b = sB[4 * ks + lk][16 * wave + lrow];
          for (int rt = 0; rt < 4; ++rt) {
            const double a = sA[4 * rt + (lrow & 3)][lrow >> 2][4 * ks + lk];
            mfma64(a, b, acc[rt]);
          }
        }
      }
    }
    if (active) {
      const int cl = 16 * wave + lrow;
      const int col = c0 + cl;
      const int f = sCol[cl];
      for (int rt = 0; rt < 4; ++rt) {
        const int pp = 4 * rt + lk;
        const int q = pbase + pp;
        double v[10];
        if (f >= 0) {
          eval_basis<10>(bas[f], px[q], py[q], pz[q], v);
        } else {
          for (int k = 0; k < 10; ++k) v[k] = 0.0;
        }
        phi[g.phi_off + (long long)(it.p0 + pp) * g.m_pad + col] = v[0];
        const double y0 = acc[rt][0], yx = acc[rt][1], yy = acc[rt][2], yz = acc[rt][3];
        double s[10];
        s[0] = v[0] * y0;
        s[1] = v[1] * y0;
        s[2] = v[2] * y0;
        s[3] = v[3] * y0;
        s[4] = v[4] * y0 + v[1] * yx;
        s[5] = v[5] * y0 + v[2] * yy;
        s[6] = v[6] * y0 + v[3] * yz;
        s[7] = v[7] * y0 + v[1] * yy;
        s[8] = v[8] * y0 + v[1] * yz;
        s[9] = v[9] * y0 + v[2] * yz;
        for (int k = 0; k < 10; ++k) {
          double x = s[k];
          x += __shfl_xor(x, 1);
          x += __shfl_xor(x, 2);
          x += __shfl_xor(x, 4);
          x += __shfl_xor(x, 8);
          s[k] = x;
        }
        if (lrow == 0)
          for (int k = 0; k < 10; ++k) sPart[wave][pp][k] += s[k];
      }
    }
  }
  __syncthreads();
  if (tid < 64) {
    double e = 0.0, n = 0.0;
    if (tid < kTile) {
      const int pp = tid;
      double t[10];
      for (int k = 0; k < 10; ++k)
        t[k] = sPart[0][pp][k] + sPart[1][pp][k] + sPart[2][pp][k] + sPart[3][pp][k];
      const double dens = t[0];
      const double gr[3] = {2.0 * t[1], 2.0 * t[2], 2.0 * t[3]};
      const double h1[3] = {2.0 * t[4], 2.0 * t[5], 2.0 * t[6]};
      const double h2[3] = {2.0 * t[7], 2.0 * t[8], 2.0 * t[9]};
      double ex, ec, y2a;
      xc::gga_point(dens, gr, h1, h2, ex, ec, y2a);
      const int q = pbase + pp;
      const double w = pw[q];
      factor[q] = w * y2a;
      e = (dens * w) * (ex + ec);
      n = dens * w;
    }
    for (int o = 8; o >= 1; o >>= 1) {
      e += __shfl_xor(e, o);
      n += __shfl_xor(n, o);
    }
    if (tid == 0) {
      epart[2 * blockIdx.x] = e;
      epart[2 * blockIdx.x + 1] = n;
    }
  }
}

// LDA variant: only phi enters; rows of the A operand are 16 points.
__global__ __launch_bounds__(256) void k_density_lda(
    const DensityItem* __restrict__ items, const DevGroup* __restrict__ groups,
    const int* __restrict__ fnl, const DevFn* __restrict__ bas, const double* __restrict__ px,
    const double* __restrict__ py, const double* __restrict__ pz, const double* __restrict__ pw,
    const double* __restrict__ P, int M, double* __restrict__ phi, double* __restrict__ factor,
    double* __restrict__ epart, int iexch) {
  __shared__ double sA[kTile][kChunk + 1];
  __shared__ double sB[kChunk][kBlock + 1];
  __shared__ double sPart[4][kTile];
  __shared__ int sCol[kBlock];
  const DensityItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int* fl = fnl + g.fn_off;
  const int pbase = g.pt_off + it.p0;
  if (tid < 4 * kTile) (&sPart[0][0])[tid] = 0.0;
  for (int c0 = 0; c0 < g.m_pad; c0 += kBlock) {
    __syncthreads();
    if (tid < kBlock) sCol[tid] = (c0 + tid < g.m_pad) ? fl[c0 + tid] : -1;
    const bool active = (c0 + 16 * wave) < g.m_pad;
    d4 acc = d4{0.0, 0.0, 0.0, 0.0};
    for (int k0 = 0; k0 < g.m_pad; k0 += kChunk) {
      __syncthreads();
      {
        const int pp = tid >> 4, kk = tid & 15;
        const int f = fl[k0 + kk];
        double v = 0.0;
        if (f >= 0) {
          const int q = pbase + pp;
          eval_basis<1>(bas[f], px[q], py[q], pz[q], &v);
        }
        sA[pp][kk] = v;
      }
      for (int e = tid; e < kChunk * kBlock; e += 256) {
        const int kk = e / kBlock, cc = e % kBlock;
        const int fr = fl[k0 + kk], fc = sCol[cc];
        sB[kk][cc] = (fr >= 0 && fc >= 0) ? P[(long long)fr * M + fc] : 0.0;
      }
      __syncthreads();
      if (active) {
        for (int ks = 0; ks < kChunk / 4; ++ks) {
          const double b = sB[4 * ks + lk][16 * wave + lrow];
          const double a = sA[lrow][4 * ks + lk];
          mfma64(a, b, acc);
        }
      }
    }
    if (active) {
      const int cl = 16 * wave + lrow;
      const int col = c0 + cl;
      const int f = sCol[cl];
      for (int j = 0; j < 4; ++j) {
        const int pp = lk + 4 * j;  // D row lk + 4 j
        const int q = pbase + pp;
        double v = 0.0;
        if (f >= 0) eval_basis<1>(bas[f], px[q], py[q], pz[q], &v);
        phi[g.phi_off + (long long)(it.p0 + pp) * g.m_pad + col] = v;
        double x = v * acc[j];
        x += __shfl_xor(x, 1);
        x += __shfl_xor(x, 2);
        x += __shfl_xor(x, 4);
        x += __shfl_xor(x, 8);
        if (lrow == 0) sPart[wave][pp] += x;
      }
    }
  }
  __syncthreads();
  if (tid < 64) {
    double e = 0.0, n = 0.0;
    if (tid < kTile) {
      const double dens = sPart[0][tid] + sPart[1][tid] + sPart[2][tid] + sPart[3][tid];
      double ex, ec, y2a;
      xc::lda_point(dens, iexch, ex, ec, y2a);
      const int q = pbase + tid;
      const double w = pw[q];
      factor[q] = w * y2a;
      e = (dens * w) * (ex + ec);
      n = dens * w;
    }
    for (int o = 8; o >= 1; o >>= 1) {
      e += __shfl_xor(e, o);
      n += __shfl_xor(n, o);
    }
    if (tid == 0) {
      epart[2 * blockIdx.x] = e;
      epart[2 * blockIdx.x + 1] = n;
    }
  }
}

// Fock tile (ti, tj) of one group over one slice of points:
// out[i][j] = sum_p phi(p, i) factor_p phi(p, j), 64 x 64, FP64 MFMA.
__global__ __launch_bounds__(256) void k_fock(const FockItem* __restrict__ items,
                                              const DevGroup* __restrict__ groups,
                                              const double* __restrict__ phi,
                                              const double* __restrict__ factor,
                                              double* __restrict__ fpart) {
  __shared__ double sI[kTile][kBlock + 1];
  __shared__ double sJ[kTile][kBlock + 1];
  const FockItem it = items[blockIdx.x];
  const DevGroup g = groups[it.group];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int lrow = lane & 15, lk = lane >> 4;
  const int i0 = kBlock * it.ti, j0 = kBlock * it.tj;
  const int pb = it.slice * kSlice;
  const int pe = min(pb + kSlice, g.npts_pad);
  const double* ph = phi + g.phi_off;
  d4 acc[4];
  for (int ct = 0; ct < 4; ++ct) acc[ct] = d4{0.0, 0.0, 0.0, 0.0};
  for (int p0 = pb; p0 < pe; p0 += kTile) {
    __syncthreads();
    for (int e = tid; e < kTile * kBlock; e += 256) {
      const int pp = e / kBlock, cc = e % kBlock;
      const long long row = (long long)(p0 + pp) * g.m_pad;
      const double fq = factor[g.pt_off + p0 + pp];
      sI[pp][cc] = (i0 + cc < g.m_pad) ? ph[row + i0 + cc] : 0.0;
      sJ[pp][cc] = (j0 + cc < g.m_pad) ? fq * ph[row + j0 + cc] : 0.0;
    }
    __syncthreads();
    for (int ks = 0; ks < kTile / 4; ++ks) {
      const double a = sI[4 * ks + lk][16 * wave + lrow];
      for (int ct = 0; ct < 4; ++ct) {
        const double b = sJ[4 * ks + lk][16 * ct + lrow];
        mfma64(a, b, acc[ct]);
      }
    }
  }
  double* out = fpart + (long long)blockIdx.x * kBlock * kBlock;
  for (int ct = 0; ct < 4; ++ct)
    for (int j = 0; j < 4; ++j) out[(16 * wave + lk + 4 * j) * kBlock + 16 * ct + lrow] = acc[ct][j];
}

// F_g tile (ti, tj) = sum over slices, in slice order.
__global__ __launch_bounds__(256) void k_fock_group(const DevGroup* __restrict__ groups,
                                                    const int2* __restrict__ tiles,
                                                    const double* __restrict__ fpart,
                                                    double* __restrict__ fg) {
  const int2 gt = tiles[blockIdx.x];  // x: group, y: tile-pair index
  const DevGroup g = groups[gt.x];
  // tile pair index -> (ti, tj), tj >= ti, row-major over the upper triangle
  int ti = 0, rem = gt.y;
  while (rem >= g.ntile - ti) {
    rem -= g.ntile - ti;
    ++ti;
  }
  const int tj = ti + rem;
  const int base = g.fock_base + gt.y * g.nslice;
  for (int e = threadIdx.x; e < kBlock * kBlock; e += 256) {
    const int r = e / kBlock, c = e % kBlock;
    const int i = kBlock * ti + r, j = kBlock * tj + c;
    if (i >= g.m_pad || j >= g.m_pad) continue;
    double s = 0.0;
    for (int sl = 0; sl < g.nslice; ++sl) s += fpart[(long long)(base + sl) * kBlock * kBlock + e];
    fg[g.fg_off + (long long)i * g.m_pad + j] = s;
  }
}

// Packed F: entry (I <= J) sums F_g(li, lj) over the groups holding both
// functions, in ascending group order.  memb holds, per function, its
// (group, local index) pairs sorted by group.
__global__ void k_fock_pack(const int* __restrict__ memb_off, const int2* __restrict__ memb,
                            const DevGroup* __restrict__ groups, const double* __restrict__ fg,
                            double* __restrict__ F, int M) {
  const int J = blockIdx.x * blockDim.x + threadIdx.x;
  const int I = blockIdx.y;
  if (J >= M || J < I) return;
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
      s += fg[g.fg_off + (long long)li * g.m_pad + lj];
      ++a;
      ++b;
    }
  }
  F[packed_index(I, J, M)] = s;
}

// Fixed-order reduction of the per-tile (Exc, N) pairs.
__global__ __launch_bounds__(256) void k_energy(const double* __restrict__ epart, int n,
                                                double* __restrict__ out) {
  __shared__ double se[256], sn[256];
  double e = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    e += epart[2 * i];
    c += epart[2 * i + 1];
  }
  se[threadIdx.x] = e;
  sn[threadIdx.x] = c;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      se[threadIdx.x] += se[threadIdx.x + s];
      sn[threadIdx.x] += sn[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = se[0];
    out[1] = sn[0];
  }
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
  int M = 0, natom = 0, iexch = 9;
  bool gga = true;
  int ngroups = 0, ndensity = 0, nfock = 0, ngroup_tiles = 0;
  long long npts_pad = 0, phi_elems = 0, fg_elems = 0;
  long long npts_real = 0;
  double flops = 0.0;  // 10 n m^2 over groups (algorithmic)
  char* arena = nullptr;
  size_t arena_bytes = 0;
  // device pointers into the arena
  DevFn* bas = nullptr;
  DevGroup* groups = nullptr;
  DensityItem* ditems = nullptr;
  FockItem* fitems = nullptr;
  int2* gtiles = nullptr;
  int* fnl = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr, *pw = nullptr;
  double* P = nullptr;
  double* rho = nullptr;
  double* F = nullptr;
  double* phi = nullptr;
  double* factor = nullptr;
  double* epart = nullptr;
  double* fpart = nullptr;
  double* fg = nullptr;
  int* memb_off = nullptr;
  int2* memb = nullptr;
  double* eout = nullptr;
  // pinned host staging
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
  for (auto& e : events_) {
    hipEvent_t ev;
    XCQ_HIP(hipEventCreate(&ev));
    e = ev;
  }
}

static void release(DeviceGrid* g) {
  if (!g) return;
  if (g->arena) (void)hipFree(g->arena);
  if (g->h_rho) (void)hipHostFree(g->h_rho);
  if (g->h_F) (void)hipHostFree(g->h_F);
  if (g->h_e) (void)hipHostFree(g->h_e);
  delete g;
}

Engine::~Engine() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize((hipStream_t)stream_);
  release(grid_);
  for (auto& e : events_)
    if (e) (void)hipEventDestroy((hipEvent_t)e);
  if (stream_) (void)hipStreamDestroy((hipStream_t)stream_);
}

size_t Engine::arena_bytes() const { return grid_ ? grid_->arena_bytes : 0; }

void Engine::load(const Molecule& mol, const GridPartition& part) {
  XCQ_HIP(hipSetDevice(device_));
  XCQ_HIP(hipStreamSynchronize((hipStream_t)stream_));
  release(grid_);
  grid_ = nullptr;
  auto* dg = new DeviceGrid;
  dg->M = mol.M;
  dg->natom = mol.natom;
  dg->iexch = mol.iexch;
  dg->gga = mol.gga;
  const int G = part.ngroups, M = mol.M;
  dg->ngroups = G;

  // host images ------------------------------------------------------------
  std::vector<DevFn> bas(M);
  for (int f = 0; f < M; ++f) {
    const BasisFn& b = mol.fn[f];
    DevFn& d = bas[f];
    std::memcpy(d.a, b.a, sizeof d.a);
    std::memcpy(d.c, b.c, sizeof d.c);
    d.cx = mol.x[b.atom];
    d.cy = mol.y[b.atom];
    d.cz = mol.z[b.atom];
    d.ncont = b.ncont;
    d.lx = b.lx;
    d.ly = b.ly;
    d.lz = b.lz;
  }
  std::vector<DevGroup> groups(G);
  std::vector<int> fnl;
  std::vector<double> hx, hy, hz, hw;
  long long pt = 0, phi = 0, fgo = 0;
  int fock_base = 0;
  std::vector<std::vector<int2>> membership(M);
  for (int g = 0; g < G; ++g) {
    DevGroup& d = groups[g];
    const int n = part.npts[g], m = part.m[g];
    d.npts = n;
    d.npts_pad = (n + kTile - 1) / kTile * kTile;
    d.m = m;
    d.m_pad = (m + 15) / 16 * 16;
    d.pt_off = (int)pt;
    d.fn_off = (int)fnl.size();
    d.ntile = (d.m_pad + kBlock - 1) / kBlock;
    d.nslice = (d.npts_pad + kSlice - 1) / kSlice;
    d.phi_off = phi;
    d.fg_off = fgo;
    d.fock_base = fock_base;
    d.pad_ = 0;
    fock_base += d.ntile * (d.ntile + 1) / 2 * d.nslice;
    for (int i = 0; i < d.m_pad; ++i) {
      const int f = i < m ? part.fn[part.fn_off[g] + i] : -1;
      fnl.push_back(f);
      if (f >= 0) membership[f].push_back(make_int2(g, i));
    }
    for (int p = 0; p < d.npts_pad; ++p) {
      const int src = part.pt_off[g] + (p < n ? p : 0);  // padding repeats point 0 with w = 0
      hx.push_back(part.x[src]);
      hy.push_back(part.y[src]);
      hz.push_back(part.z[src]);
      hw.push_back(p < n ? part.w[src] : 0.0);
    }
    pt += d.npts_pad;
    phi += (long long)d.npts_pad * d.m_pad;
    fgo += (long long)d.m_pad * d.m_pad;
    dg->flops += 10.0 * n * (double)m * m;
    dg->npts_real += n;
  }
  dg->npts_pad = pt;
  dg->phi_elems = phi;
  dg->fg_elems = fgo;
  dg->nfock = fock_base;

  // work lists: heavy groups first (longest-processing-time order)
  std::vector<int> order(G);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
    return (long long)groups[a].m_pad * groups[a].m_pad > (long long)groups[b].m_pad * groups[b].m_pad;
  });
  std::vector<DensityItem> ditems;
  for (int g : order)
    for (int p0 = 0; p0 < groups[g].npts_pad; p0 += kTile) ditems.push_back(DensityItem{g, p0});
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
  dg->ndensity = (int)ditems.size();
  dg->ngroup_tiles = (int)gtiles.size();
  std::vector<int> moff(M + 1, 0);
  std::vector<int2> memb;
  for (int f = 0; f < M; ++f) {
    moff[f] = (int)memb.size();
    memb.insert(memb.end(), membership[f].begin(), membership[f].end());
  }
  moff[M] = (int)memb.size();

  // arena --------------------------------------------------------------------
  const size_t packed = (size_t)M * (M + 1) / 2;
  auto plan = [&](Carver& c) {
    dg->bas = c.take<DevFn>(M);
    dg->groups = c.take<DevGroup>(G);
    dg->ditems = c.take<DensityItem>(ditems.size());
    dg->fitems = c.take<FockItem>(fitems.size());
    dg->gtiles = c.take<int2>(gtiles.size());
    dg->fnl = c.take<int>(fnl.size());
    dg->px = c.take<double>(pt);
    dg->py = c.take<double>(pt);
    dg->pz = c.take<double>(pt);
    dg->pw = c.take<double>(pt);
    dg->P = c.take<double>((size_t)M * M);
    dg->rho = c.take<double>(packed);
    dg->F = c.take<double>(packed);
    dg->phi = c.take<double>(phi);
    dg->factor = c.take<double>(pt);
    dg->epart = c.take<double>(2 * ditems.size());
    dg->fpart = c.take<double>((size_t)fitems.size() * kBlock * kBlock);
    dg->fg = c.take<double>(fgo);
    dg->memb_off = c.take<int>(M + 1);
    dg->memb = c.take<int2>(memb.size() + 1);
    dg->eout = c.take<double>(2);
  };
  Carver sizing{nullptr};
  plan(sizing);
  dg->arena_bytes = sizing.used + 256;
  XCQ_HIP(hipMalloc(&dg->arena, dg->arena_bytes));
  Carver carve{dg->arena};
  plan(carve);
  XCQ_HIP(hipHostMalloc(&dg->h_rho, packed * sizeof(double), hipHostMallocDefault));
  XCQ_HIP(hipHostMalloc(&dg->h_F, packed * sizeof(double), hipHostMallocDefault));
  XCQ_HIP(hipHostMalloc(&dg->h_e, 2 * sizeof(double), hipHostMallocDefault));

  hipStream_t s = (hipStream_t)stream_;
  auto up = [&](void* dst, const void* src, size_t bytes) {
    if (bytes) XCQ_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  };
  up(dg->bas, bas.data(), bas.size() * sizeof(DevFn));
  up(dg->groups, groups.data(), groups.size() * sizeof(DevGroup));
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
  XCQ_HIP(hipStreamSynchronize(s));
  grid_ = dg;
}

SolveResult Engine::solve_device(const double* d_rho, double* d_fock, bool want_energy) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  DeviceGrid& g = *grid_;
  hipStream_t s = (hipStream_t)stream_;
  const int M = g.M;
  SolveResult res;
  hipEvent_t* ev = reinterpret_cast<hipEvent_t*>(events_);
  XCQ_HIP(hipEventRecord(ev[0], s));
  {
    const long long n = (long long)M * M;
    hipLaunchKernelGGL(k_expand_density, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, d_rho,
                       g.P, M);
  }
  XCQ_HIP(hipEventRecord(ev[1], s));
  if (g.ndensity > 0) {
    if (g.gga)
      hipLaunchKernelGGL(k_density_gga, dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                         g.fnl, g.bas, g.px, g.py, g.pz, g.pw, g.P, M, g.phi, g.factor, g.epart);
    else
      hipLaunchKernelGGL(k_density_lda, dim3(g.ndensity), dim3(256), 0, s, g.ditems, g.groups,
                         g.fnl, g.bas, g.px, g.py, g.pz, g.pw, g.P, M, g.phi, g.factor, g.epart,
                         g.iexch);
  }
  XCQ_HIP(hipGetLastError());
  XCQ_HIP(hipEventRecord(ev[2], s));
  if (g.nfock > 0) {
    hipLaunchKernelGGL(k_fock, dim3(g.nfock), dim3(256), 0, s, g.fitems, g.groups, g.phi,
                       g.factor, g.fpart);
    hipLaunchKernelGGL(k_fock_group, dim3(g.ngroup_tiles), dim3(256), 0, s, g.groups, g.gtiles,
                       g.fpart, g.fg);
  }
  XCQ_HIP(hipGetLastError());
  XCQ_HIP(hipEventRecord(ev[3], s));
  if (d_fock)
    hipLaunchKernelGGL(k_fock_pack, dim3((M + 127) / 128, M), dim3(128), 0, s, g.memb_off, g.memb,
                       g.groups, g.fg, d_fock, M);
  hipLaunchKernelGGL(k_energy, dim3(1), dim3(256), 0, s, g.epart, g.ndensity, g.eout);
  XCQ_HIP(hipGetLastError());
  XCQ_HIP(hipEventRecord(ev[4], s));
  (void)want_energy;
  return res;
}

void Engine::read_energy(double* energy, double* electrons) {
  if (!grid_) throw std::runtime_error("g2g: no grid");
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipMemcpyAsync(grid_->h_e, grid_->eout, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  if (energy) *energy = grid_->h_e[0];
  if (electrons) *electrons = grid_->h_e[1];
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

SolveResult Engine::solve(const double* rho_packed, double* fock_accum, bool want_energy) {
  if (!grid_) throw std::runtime_error("g2g: solve before the grid was built");
  const double t0 = wall();
  DeviceGrid& g = *grid_;
  hipStream_t s = (hipStream_t)stream_;
  XCQ_HIP(hipSetDevice(device_));
  const size_t packed = (size_t)g.M * (g.M + 1) / 2;
  std::memcpy(g.h_rho, rho_packed, packed * sizeof(double));
  XCQ_HIP(hipMemcpyAsync(g.rho, g.h_rho, packed * sizeof(double), hipMemcpyHostToDevice, s));
  SolveResult res = solve_device(g.rho, fock_accum ? g.F : nullptr, want_energy);
  if (fock_accum)
    XCQ_HIP(hipMemcpyAsync(g.h_F, g.F, packed * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipMemcpyAsync(g.h_e, g.eout, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  XCQ_HIP(hipStreamSynchronize(s));
  if (fock_accum)
    for (size_t k = 0; k < packed; ++k) fock_accum[k] += g.h_F[k];
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

}  // namespace xcq
