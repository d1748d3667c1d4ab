// Variable-density pressure Poisson of the standard projection on gfx950.
//
// Reference: sim_solve_standard_projection's Poisson block
// (src/module_simulation.f90:2477-2588), poisson_matvec (:2668-2707),
// poisson_ssor (:2709-2758), the Jacobi preconditioner
// (src/module_linear_solvers.f90:88-135), BiCGSTAB(2) (:361-579).
//
// Operator storage ("face form"): the five reference arrays a_W..a_N, a_P
// (56 B/cell) are stored as kx((nx+1) x ny) -- kx(i,j) = a_W(i,j) and
// kx(nx+1,j) = a_E(nx,j) --, ky(nx x (ny+1)) likewise, and diag = a_P.  The
// operator is symmetric (a_E(i,j) = a_W(i+1,j), a_N(i,j) = a_S(i,j+1)), so
// this holds every coefficient once and the matvec evaluates the reference's
// expression bit-for-bit: a_P x - a_W x_W - a_E x_E - a_S x_S - a_N x_N in that
// order, neighbours dropped at non-periodic edges, wrapped when periodic.
//
// Krylov iterations are device-driven: each fused kernel ends with a
// deterministic two-level reduction (fixed per-block partials summed in fixed
// order by the last block to finish), and that last block also computes the
// scalar recurrences (alpha, beta, stop flags), so no host round trip is
// needed inside an iteration.  The host polls the stop flag every few dozen
// iterations.
#include "covof_internal.h"
#include "seq_sweep.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <initializer_list>
#include <utility>

namespace covof {

// ---------------------------------------------------------------- scalars
enum Slot {
  S_BNORM = 0, S_RZ = 1, /* and 2 */ S_STEP = 3, S_RR = 4, S_PEND = 5, S_RELRES = 6,
  S_FLAG = 7, S_ITERS = 8, S_TOL = 9, S_FIRST = 10, S_MAXIT = 11, S_DOT0 = 12,
  S_RMEAN = 13,  // constant-mode handling of the mixed mode
  S_BEST = 14, S_STALL = 15,  // stall guard of the mixed mode
  // inner (FP32) solve of the mixed mode
  I_RZ = 16, /* and 17 */ I_RELRES = 18, I_RR0 = 19, I_FLAG = 20, I_ITERS = 21, I_STEP = 22,
  I_RR = 23, I_ALPHA = 24, /* and 25 */ I_FIRST = 26, I_MAXIT = 27, I_PEND = 28,
  // BiCGSTAB(2)
  B_RHO0 = 32, B_RHO1, B_ALPHA, B_OMEGA, B_BETA, B_GAMMA, B_R0N, B_G1, B_G2, B_RN,
  B_S11, B_S12, B_S22, B_RHS1, B_RHS2,
  // 48-55: the history row of diagnostics.hip
  S_ZERO = 63,  // never written: the stop slot of launches outside a solve
  NSLOT = 64
};
enum Flag { F_RUN = 0, F_CONV = 1, F_BREAK = 2, F_NONFINITE = 3, F_MAXIT = 5, F_STALL = 6 };

struct DevOp {
  int nx, ny;
  int perW, perE, perS, perN;
  double eps;  // the 1e-12 * dvol diagonal shift of the assembly
};

template <typename T>
struct OpT {
  const T *kx, *ky, *diag;
};

// kx rows are padded to a multiple of 64 entries so every row starts on a
// 256-byte boundary (the (nx+1)-long rows would otherwise shift by one word).
__host__ __device__ __forceinline__ long kx_pitch(int nx) { return ((long)(nx + 1) + 63) & ~63L; }
__device__ __forceinline__ long kxi(int nx, int i, int j) { return (long)(j - 1) * kx_pitch(nx) + (i - 1); }
__device__ __forceinline__ long kyi(int nx, int i, int j) { return (long)(j - 1) * nx + (i - 1); }

// y(i,j) of poisson_matvec for a vector given by an accessor xv(k).
template <typename T, typename XV>
__device__ __forceinline__ T apply_op(const DevOp &o, const OpT<T> &A, XV xv, int i, int j, long k) {
  const int nx = o.nx, ny = o.ny;
  T val = A.diag[k] * xv(k);
  if (i > 1) val = val - A.kx[kxi(nx, i, j)] * xv(k - 1);
  else if (o.perW) val = val - A.kx[kxi(nx, 1, j)] * xv(k + (nx - 1));
  if (i < nx) val = val - A.kx[kxi(nx, i + 1, j)] * xv(k + 1);
  else if (o.perE) val = val - A.kx[kxi(nx, nx + 1, j)] * xv(k - (nx - 1));
  if (j > 1) val = val - A.ky[kyi(nx, i, j)] * xv(k - nx);
  else if (o.perS) val = val - A.ky[kyi(nx, i, 1)] * xv(k + (long)(ny - 1) * nx);
  if (j < ny) val = val - A.ky[kyi(nx, i, j + 1)] * xv(k + nx);
  else if (o.perN) val = val - A.ky[kyi(nx, i, ny + 1)] * xv(k - (long)(ny - 1) * nx);
  return val;
}

// The same product with the diagonal formed from the four face values,
// a_P = a_W + a_E + a_S + a_N + 1e-12 dvol in the assembly's order (bit-equal to
// the stored a_P in FP64), so the diagonal array is not streamed.
template <typename T, typename XV>
__device__ __forceinline__ T apply_op_fd(const DevOp &o, const OpT<T> &A, XV xv, int i, int j, long k) {
  const int nx = o.nx, ny = o.ny;
  const T aW = A.kx[kxi(nx, i, j)], aE = A.kx[kxi(nx, i + 1, j)];
  const T aS = A.ky[kyi(nx, i, j)], aN = A.ky[kyi(nx, i, j + 1)];
  const T aP = aW + aE + aS + aN + (T)o.eps;
  T val = aP * xv(k);
  if (i > 1) val = val - aW * xv(k - 1);
  else if (o.perW) val = val - aW * xv(k + (nx - 1));
  if (i < nx) val = val - aE * xv(k + 1);
  else if (o.perE) val = val - aE * xv(k - (nx - 1));
  if (j > 1) val = val - aS * xv(k - nx);
  else if (o.perS) val = val - aS * xv(k + (long)(ny - 1) * nx);
  if (j < ny) val = val - aN * xv(k + nx);
  else if (o.perN) val = val - aN * xv(k - (long)(ny - 1) * nx);
  return val;
}

// Jacobi z = r / d if |d| > 1e-15 else r (src/module_linear_solvers.f90:94-101).
template <typename T>
__device__ __forceinline__ T jacobi(T r, T d) {
  return (fabs((double)d) > 1.0e-15) ? r / d : r;
}

// ---------------------------------------------------- reductions (deterministic)
//
// No atomics and no device-scope fences (a release fence on a multi-XCD part
// writes the whole L2 back): a kernel leaves one partial per block and slot
// with plain stores, slot-major (slot q of block b at part[q * nblk + b], so a
// slot is one contiguous run), and the NEXT kernel on the stream -- the kernel
// boundary makes the stores visible -- has every block sum the slots it needs
// in the same fixed order.  Every block thus derives bit-identical scalars; persistent
// scalars are written by block 0 only, into slots nobody else reads during
// that launch (parity-indexed where a value is both read and updated).
constexpr int BLK = 256;  // threads per block: tile 64 (x) by 4 (y)
constexpr int NP = 4;     // partial slots per block (at most)

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

__device__ __forceinline__ int block_id() { return blockIdx.y * gridDim.x + blockIdx.x; }
__device__ __forceinline__ bool lead() { return threadIdx.x == 0 && blockIdx.x == 0 && blockIdx.y == 0; }

// Block sums of v[0..NV) (wave sums in T, across waves in double) into slot
// first + q of this block.
template <int NV, typename T>
__device__ void block_partials(const T (&v)[NV], double *part, int first = 0) {
  __shared__ double sh[NV][BLK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const T w = wave_sum<T>(v[q]);
    if (lane == 0) sh[q][wid] = (double)w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double t = 0.0;
      for (int w = 0; w < BLK / 64; ++w) t += sh[q][w];
      part[(long)(first + q) * (gridDim.x * gridDim.y) + block_id()] = t;
    }
  }
}

// Grid totals tot[q] of slots q0 <= q < q1 (bounds uniform over the grid;
// the other entries are 0) of the partials left by the previous kernel (nblk
// blocks, even): the same value in every thread of every block.  Each thread
// reads pairs of partials with 16-byte loads.
template <int NV>
__device__ void grid_totals(const double *part, int nblk, double (&tot)[NV], int q0 = 0, int q1 = NV) {
  __shared__ double sh[NV][BLK / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // every slot's loads of a round are issued before any is added (a thread's
  // pairs of a slot are still added in index order)
  constexpr int U = 2;  // pairs per thread and slot per round: 1024 partials in one round
  double t[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) t[q] = 0.0;
  for (int b0 = threadIdx.x; 2 * b0 < nblk; b0 += U * BLK) {
    double2 v[NV][U];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      const double2 *p2 = reinterpret_cast<const double2 *>(part + (long)q * nblk);
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int b = b0 + u * BLK;
        v[q][u] = (q >= q0 && q < q1 && 2 * b < nblk) ? p2[b] : make_double2(0.0, 0.0);
      }
    }
#pragma unroll
    for (int q = 0; q < NV; ++q)
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (2 * (b0 + u * BLK) < nblk) {
          t[q] += v[q][u].x;
          t[q] += v[q][u].y;
        }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if (q < q0 || q >= q1) continue;
    const double w = wave_sum<double>(t[q]);
    if (lane == 0) sh[q][wid] = w;
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double t = 0.0;
    if (q >= q0 && q < q1)
      for (int w = 0; w < BLK / 64; ++w) t += sh[q][w];
    tot[q] = t;
  }
}

// ------------------------------------ reference-order sums (serial, exact)
//
// The reference sums an inner product as sum(x*y) (or dot_product) over a
// column-major array: one accumulator, elements in index order (i fastest),
// each product rounded before it is added.  That chain is sequential, so the
// reference-order mode gives each sum one wave: all 64 lanes form the
// products of the next 512-element chunk and stage them in LDS while lane 0
// adds the current chunk's products, in order, into its accumulator.  Several
// sums run side by side in the waves of one workgroup.
constexpr int SER_CH = 512;  // products per chunk and sum

__device__ __forceinline__ double ser_term(const SerialSet &s, int q, long t) {
  const double x = s.a[q][t];
  if (s.b[q]) return x * s.b[q][t];
  if (s.dg[q]) return x * jacobi<double>(x, s.dg[q][t]);
  return x;
}

__global__ void __launch_bounds__(64 * SER_MAX) k_serial_sums(long n, SerialSet s, const double *scal, int f0,
                                                              int f1) {
  if (scal[f0] != 0.0 || scal[f1] != 0.0) return;
  constexpr int U = SER_CH / 64;
  __shared__ double buf[SER_MAX][2][SER_CH];
  const int q = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const long nch = (n + SER_CH - 1) / SER_CH;
  double v[U];
#pragma unroll
  for (int u = 0; u < U; ++u) {
    const long t = u * 64 + lane;
    v[u] = t < n ? ser_term(s, q, t) : 0.0;
  }
#pragma unroll
  for (int u = 0; u < U; ++u) buf[q][0][u * 64 + lane] = v[u];
  __syncthreads();
  double acc = 0.0;
  for (long c = 0; c < nch; ++c) {
    const bool more = c + 1 < nch;  // uniform over the workgroup
    if (more) {
      const long base = (c + 1) * SER_CH;
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long t = base + u * 64 + lane;
        v[u] = t < n ? ser_term(s, q, t) : 0.0;
      }
    }
    if (lane == 0) {
      const double *bb = buf[q][c & 1];
      const long left = n - c * SER_CH;
      if (left >= SER_CH) {
#pragma unroll 32
        for (int m = 0; m < SER_CH; ++m) acc = acc + bb[m];
      } else {
        for (int m = 0; m < (int)left; ++m) acc = acc + bb[m];
      }
    }
    if (more) {
#pragma unroll
      for (int u = 0; u < U; ++u) buf[q][(c + 1) & 1][u * 64 + lane] = v[u];
    }
    __syncthreads();
  }
  if (lane == 0) s.dst[q][0] = acc;
  for (int k = lane + 1; k < s.nzero; k += 64) s.dst[q][k] = 0.0;
}

void serial_sums(Ctx &c, long n, const SerialSet &s, int f0, int f1) {
  hipLaunchKernelGGL(k_serial_sums, dim3(1), dim3(64 * s.nq), 0, c.stream, n, s, (const double *)c.ps.scal, f0,
                     f1);
}

// Cell loop over a 1-D grid of 8 * gx * chunks blocks (gx = 64-column strips).
// The dispatcher deals workgroup ids round-robin over the 8 XCDs, so block id
// goes to XCD id % 8; that XCD owns a horizontal band of ny/8 rows, and its
// blocks split the band into 64-column strips and row chunks walked 4 rows at
// a time.  Stencil neighbours left/right and above/below then come from the
// same XCD's L2 (only the band edges cross), and the rows a block reads as
// "above" it touched a moment before.
// CELL_LOOP's partition without the loop: i, the row range [jb_ + 1, je_] and
// active_ (the block has a piece of the grid).
#define CELL_BOUNDS(o)                                                                         \
  const int gx_ = ((o).nx + 63) / 64;                                                          \
  const int xcd_ = blockIdx.x & 7, local_ = blockIdx.x >> 3;                                   \
  const int chunks_ = (gridDim.x >> 3) / gx_;                                                  \
  const int band_ = ((((o).ny + 7) / 8) + 3) & ~3;                                             \
  const int rows_ = (((band_ + chunks_ - 1) / chunks_) + 3) & ~3;                              \
  const int i = (local_ % gx_) * 64 + (threadIdx.x & 63) + 1;                                  \
  const int jb_ = xcd_ * band_ + (local_ / gx_) * rows_;                                        \
  const int je_ = min(min(jb_ + rows_, xcd_ * band_ + band_), (o).ny);                          \
  const bool active_ = local_ < gx_ * chunks_

#define CELL_LOOP(o)                                                                           \
  const int gx_ = ((o).nx + 63) / 64;                                                          \
  const int xcd_ = blockIdx.x & 7, local_ = blockIdx.x >> 3;                                   \
  const int chunks_ = (gridDim.x >> 3) / gx_;                                                  \
  const int band_ = ((((o).ny + 7) / 8) + 3) & ~3;                                             \
  const int rows_ = (((band_ + chunks_ - 1) / chunks_) + 3) & ~3;                              \
  const int i = (local_ % gx_) * 64 + (threadIdx.x & 63) + 1;                                  \
  const int jb_ = xcd_ * band_ + (local_ / gx_) * rows_;                                        \
  const int je_ = min(min(jb_ + rows_, xcd_ * band_ + band_), (o).ny);                          \
  if (local_ < gx_ * chunks_)                                                                  \
    for (int j = jb_ + (threadIdx.x >> 6) + 1; j <= je_; j += 4)                               \
      if (i <= (o).nx)

// ---------------------------------------------------------------- assembly

// A solid obstacle as the Poisson block sees it: on = obs_enabled and
// obs_perm <= 1e-12, then a face is a wall when its point lies in the box or
// disc (is_in_obstacle, src/module_mesh.f90:203-222, at (x_face, y_center)
// for u faces and (x_center, y_face) for v faces,
// src/module_navier_stokes.f90:82-131).
struct DevObs {
  int on, type;
  double x0, y0, x1, y1, xc, yc, r;
};
__device__ __forceinline__ bool in_obs(const DevObs &o, double x, double y) {
  if (o.type == 1) return x >= o.x0 - 1.0e-9 && x <= o.x1 + 1.0e-9 && y >= o.y0 - 1.0e-9 && y <= o.y1 + 1.0e-9;
  if (o.type == 2) return (x - o.xc) * (x - o.xc) + (y - o.yc) * (y - o.yc) <= o.r * o.r;
  return false;
}
__device__ __forceinline__ bool solid_u(const DevObs &o, int i, int j, double dx, double dy) {
  return o.on && in_obs(o, (double)(i - 1) * dx, ((double)(j - 1) + 0.5) * dy);
}
__device__ __forceinline__ bool solid_v(const DevObs &o, int i, int j, double dx, double dy) {
  return o.on && in_obs(o, ((double)(i - 1) + 0.5) * dx, (double)(j - 1) * dy);
}

// src/module_simulation.f90:2483-2533 on faces: one thread per cell writes
// b, kx(i,j) (and kx(nx+1,j) on the last column), ky(i,j) (and ky(i,ny+1)),
// then diag from the four face values in the reference's summation order.
__global__ void k_assemble(const double *__restrict__ u, const double *__restrict__ v,
                           const double *__restrict__ rho_u, const double *__restrict__ rho_v,
                           int nx, int ny, double dx, double dy, double dt, int wallW, int wallE,
                           int wallS, int wallN, DevObs ob, double *kx, double *ky, double *diag, double *b,
                           float *kx32, float *ky32, float *diag32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > nx || j > ny) return;
  auto H = [&](int a, int c) { return (long)c * (nx + 2) + a; };
  auto cx = [&](int a) {  // face a of row j
    double r = rho_u[H(a, j)];
    if (r < 1.0e-12) r = 1.0;
    return (dt * dy) / (r * dx);
  };
  auto cy = [&](int c) {
    double r = rho_v[H(i, c)];
    if (r < 1.0e-12) r = 1.0;
    return (dt * dx) / (r * dy);
  };
  const double aW = ((i == 1 && wallW) || solid_u(ob, i, j, dx, dy)) ? 0.0 : cx(i);
  const double aE = ((i == nx && wallE) || solid_u(ob, i + 1, j, dx, dy)) ? 0.0 : cx(i + 1);
  const double aS = ((j == 1 && wallS) || solid_v(ob, i, j, dx, dy)) ? 0.0 : cy(j);
  const double aN = ((j == ny && wallN) || solid_v(ob, i, j + 1, dx, dy)) ? 0.0 : cy(j + 1);
  const long k = (long)(j - 1) * nx + (i - 1);
  kx[kxi(nx, i, j)] = aW;
  kx32[kxi(nx, i, j)] = (float)aW;
  if (i == nx) {
    kx[kxi(nx, nx + 1, j)] = aE;
    kx32[kxi(nx, nx + 1, j)] = (float)aE;
  }
  ky[kyi(nx, i, j)] = aS;
  ky32[kyi(nx, i, j)] = (float)aS;
  if (j == ny) {
    ky[kyi(nx, i, ny + 1)] = aN;
    ky32[kyi(nx, i, ny + 1)] = (float)aN;
  }
  const double aP = aW + aE + aS + aN + 1.0e-12 * (dx * dy);
  diag[k] = aP;
  // the FP32 operator's diagonal is summed from its own FP32 faces (what the
  // FP32 PCG kernels form on the fly), not rounded from the FP64 one
  diag32[k] = (float)aW + (float)aE + (float)aS + (float)aN + (float)(1.0e-12 * (dx * dy));
  const double div = (u[H(i + 1, j)] - u[H(i, j)]) * dy + (v[H(i, j + 1)] - v[H(i, j)]) * dx;
  if (b) b[k] = -div;
}

// ----------------------------------------------------------- basic kernels

template <typename T>
__global__ void __launch_bounds__(BLK) k_matvec(DevOp o, OpT<T> A, const T *__restrict__ x, T *__restrict__ y) {
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    y[k] = apply_op<T>(o, A, [&](long q) { return x[q]; }, i, j, k);
  }
}

__global__ void k_precond_simple(int precond, long n, const double *__restrict__ d,
                                 const double *__restrict__ r, double *__restrict__ z) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  z[t] = (precond == 1) ? jacobi<double>(r[t], d[t]) : r[t];
}

template <typename T>
__global__ void k_fill(long n, T *__restrict__ x, T v) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) x[t] = v;
}

__global__ void k_set_scalars(double *scal, int first, int n, double v) {
  const int t = threadIdx.x;
  if (t < n) scal[first + t] = v;
}

// The reference's symmetric Gauss-Seidel (poisson_ssor, :2709-2758): a
// forward sweep (D + L) z = r, then a backward sweep (D + U) z = D z, both in
// lexicographic order, with the periodic wrap terms reading what that order
// reads (still 0 forward, the forward value backward).  It runs on the strip
// sweeps of seq_sweep.h in their Poisson mode: per cell a record of a_P and
// the four face coefficients, zero where the reference adds no term (a wall
// face), so every cell evaluates the reference's expression on the same
// operands, bit for bit.  A flag slot (the solve's convergence flag, or null)
// skips the application once the solve has stopped.
__global__ void __launch_bounds__(BLK) k_sgs_records(DevOp o, OpT<double> A, int S, double *__restrict__ rec) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  const int nx = o.nx, ny = o.ny;
  if (k >= (long)nx * ny) return;
  const int j = (int)(k / nx) + 1, i = (int)(k - (long)(j - 1) * nx) + 1;
  double *r = rec + sk_of(S, i, j) * 8;
  r[SR_P] = A.diag[k];
  r[SR_W] = (i > 1 || o.perW) ? A.kx[kxi(nx, i, j)] : 0.0;
  r[SR_E] = (i < nx || o.perE) ? A.kx[kxi(nx, i + 1, j)] : 0.0;
  r[SR_S] = (j > 1 || o.perS) ? A.ky[kyi(nx, i, j)] : 0.0;
  r[SR_N] = (j < ny || o.perN) ? A.ky[kyi(nx, i, j + 1)] : 0.0;
}

__global__ void __launch_bounds__(BLK) k_sgs_in(DevOp o, int S, const double *__restrict__ r, double *rec,
                                                const double *flag) {
  if (flag && *flag != 0.0) return;
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)o.nx * o.ny) return;
  const int j = (int)(k / o.nx) + 1, i = (int)(k - (long)(j - 1) * o.nx) + 1;
  rec[sk_of(S, i, j) * 8 + SR_R] = r[k];
}

__global__ void __launch_bounds__(64 * SEQ_MAXW) k_sgs_sweep(DevOp o, SeqBlk b, const double *flag) {
  if (flag && *flag != 0.0) return;
  seq_sweeps<SEQ_POISSON_SGS>(o.perW != 0, o.perS != 0, b, 1, 1.0);
}

__global__ void __launch_bounds__(BLK) k_sgs_out(DevOp o, int S, const double *__restrict__ x,
                                                 double *__restrict__ z, const double *flag) {
  if (flag && *flag != 0.0) return;
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)o.nx * o.ny) return;
  const int j = (int)(k / o.nx) + 1, i = (int)(k - (long)(j - 1) * o.nx) + 1;
  z[k] = x[sk_of(S, i, j)];
}

// ------------------------------------------------------------ PCG (fused)
//
// Iteration k of the SP PCG loop (src/module_simulation.f90:2553-2575), two
// kernels, each opening with the grid totals of the previous kernel:
//   A_k: totals of B_{k-1} (or of the initial residual when k = 1): r.r, r.z
//        (and b.b at k = 1).  ||r||/||b|| <= tol -> converged (the exit of
//        iteration k-1); |rz_old| < 1e-35 -> breakdown; beta = rz / rz_old.
//        Body: x += alpha_{k-1} d_prev (the solution update of iteration
//        k-1, one kernel late), d = z + beta d_prev (d = z at k = 1,
//        d_prev = 0), Ad = A d, partial d.Ad.  z is formed on the fly from r
//        (Jacobi: z = r / a_P with a_P summed from the four face
//        coefficients, bit-equal to the assembled diagonal; none: z = r) or
//        read (explicit preconditioners), so A_k reads r or z, d_prev, x, kx,
//        ky and writes d, Ad, x.
//   B_k: totals of A_k: d.Ad; |dAd| < 1e-35 -> breakdown; alpha = rz / dAd.
//        Body: r -= alpha Ad, partials r.r and r.z (z = r / a_P from the
//        stored diagonal).  B_k reads r, Ad, diag and writes r.
// Each x / r value is formed by the reference's expression on the same
// operands; only the x update moves into the next A.  The update of the last
// iteration of a solve (A stopped, or the FP32 recurrence reached a reliable
// update) is left pending: B records alpha and which d buffer it used, A
// clears the record when it applies the update, and k_flush_x (or the
// reliable update) applies a record left standing.
// d is ping-ponged (A_k reads neighbours' d_prev while writing d); rz is kept
// in two parity slots (A_k reads rz_{k-1} and stores rz_k).
struct PcgSlots {
  int rz;        // rz at parity 0, rz at parity 1 in rz + 1
  int rr, relres, flag, iters, ref, tol, first, maxit;
  int is_b_norm;  // reference norm: ||b|| (< 1e-30 -> 1) from the first totals;
                  // 0: the norm already in slot ref (the outer ||b||)
  int inner;      // FP32 recurrence of the mixed mode: no convergence test of
                  // its own (the FP64 residual of the reliable updates decides)
  int step;       // alpha of the last B (as the value of type T)
  int pend;       // 0: no pending x update; 1 / 2: alpha * d0 / d1 is pending
};

// Rows per load batch of k_pcg_a (FP64 / FP32) and the target block count
// of the reduction grid (build-time tunables).
#ifndef PCG_ROWS_F64
#define PCG_ROWS_F64 2
#endif
#ifndef PCG_ROWS_F32
#define PCG_ROWS_F32 4
#endif
#ifndef RED_BLOCKS
#define RED_BLOCKS 1024
#endif

template <typename T, int PC>
__global__ void __launch_bounds__(BLK) k_pcg_a(DevOp o, OpT<T> A, const T *__restrict__ zr,
                                               const T *__restrict__ dprev, T *__restrict__ d,
                                               T *__restrict__ Ad, T *__restrict__ x, double *scal,
                                               PcgSlots sl, const double *pin, double *pout, int k) {
  // The block's four waves walk the same run of rows over four adjacent
  // 64-column strips, each keeping d, a_P and the face coefficients of the
  // rows below / at / above in registers: every value is loaded once,
  // x-neighbours come from the adjacent lanes, and at the strip edges inside
  // the block from the adjacent wave through LDS (one exchange of the west
  // faces and one of the edge lanes' d per batch of rows); only the block's
  // outer edges and the periodic wraps load extra values.  The row above is
  // loaded one step ahead (its z needs its own a_P).  Rows go in batches of G
  // whose loads are all issued before the batch is evaluated, so each wave
  // keeps several rows of loads in flight.  The matvec evaluates
  // poisson_matvec's expression in its order on the same operands.
  constexpr int G = sizeof(T) == 8 ? PCG_ROWS_F64 : PCG_ROWS_F32;
  __shared__ T s_kw[4][G], s_c0[4][G], s_c63[4][G], s_ke0[4];
  const int nx = o.nx, ny = o.ny;
  const int gx_ = (nx + 63) / 64;
  const int gq_ = (gx_ + 3) / 4;  // groups of four strips
  const int xcd_ = blockIdx.x & 7, local_ = blockIdx.x >> 3;
  const int chunks_ = (gridDim.x >> 3) / gq_;
  const int band_ = (((ny + 7) / 8) + 3) & ~3;
  const int rows_ = (band_ + chunks_ - 1) / chunks_;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int i = ((local_ % gq_) * 4 + w) * 64 + lane + 1;
  const int jb = xcd_ * band_ + (local_ / gq_) * rows_;
  const int je = min(min(jb + rows_, xcd_ * band_ + band_), ny);
  const int jw0 = jb + 1, jw1 = je;
  const bool vi = i <= nx;
  const bool act = local_ < gq_ * chunks_ && jw0 <= jw1;  // uniform over the block
  // the right neighbour: the next lane, the next wave's lane 0 (through LDS),
  // or a load (the block's last strip, the mesh's last column)
  const bool right_lds = lane == 63 && w < 3 && i < nx;
  const bool right_load = i == nx || (lane == 63 && w == 3);
  const bool left_lds = lane == 0 && w > 0;
  const bool left_load = lane == 0 && w == 0;
  auto at = [&](int a, int b) { return (long)(b - 1) * nx + (a - 1); };
  auto zval = [&](T rv, T ap) -> T { return PC == 1 ? jacobi<T>(rv, ap) : rv; };
  // The first row's operands and the row below do not depend on beta: they
  // are loaded before the stop flags and the grid totals are read, so the
  // latencies overlap.
  // Row chunks alternate their walking direction: an odd chunk walks down
  // from its last row.  Two chunks that share a boundary then read the rows
  // on either side of it at the same moment (both at their start, or both at
  // their end), so the second reader finds them in L2; walking the same way,
  // one of them would read the row at its start and the other at its end.
  // "prev" is the row before the first one in walking order (the row below,
  // or the row above when walking down).
  const bool rev = ((local_ / gq_) & 1) != 0;
  const int jf = rev ? jw1 : jw0;  // first row of the walk
  T r0 = T(0), p0 = T(0), rd = T(0), pd = T(0), dgd = T(0), aS = T(0), kw = T(0), kn = T(0), ke = T(0);
  if (act && vi) {
    r0 = zr[at(i, jf)];
    p0 = dprev[at(i, jf)];
    const int jd = rev ? (jf < ny ? jf + 1 : (o.perN ? 1 : 0)) : (jf > 1 ? jf - 1 : (o.perS ? ny : 0));
    if (jd > 0) {
      rd = zr[at(i, jd)];
      pd = dprev[at(i, jd)];
      if (PC == 1) dgd = A.diag[at(i, jd)];
    }
    aS = A.ky[kyi(nx, i, jf)];
    kw = A.kx[kxi(nx, i, jf)];
    kn = A.ky[kyi(nx, i, jf + 1)];
    if (i == nx) ke = A.kx[kxi(nx, nx + 1, jf)];
    else if (right_load) ke = A.kx[kxi(nx, i + 1, jf)];
  }
  const double flag_own = scal[sl.flag], flag_outer = scal[S_FLAG];
  const bool first = scal[sl.first] != 0.0;
  const T alpha_prev = (T)scal[sl.step];
  // slots: 0 r.r, 1 r.z, 2 b.b (first outer iteration only); the inner
  // recurrence of the mixed mode needs r.z only -- and, at the first
  // iteration after a reliable update, slots 2 and 3: z.(A d_prev) and
  // d_prev.(A d_prev) (see k_refresh_conj)
  const bool post_refresh = sl.inner && k == 1 && !first;
  double tot[4];
  grid_totals<4>(pin, gridDim.x, tot, sl.inner ? 1 : 0, post_refresh ? 4 : ((first && sl.is_b_norm) ? 3 : 2));
  if (flag_own != 0.0 || flag_outer != 0.0) return;
  const double rz_prev = scal[sl.rz + ((k - 1) & 1)];
  int stop = F_RUN;
  double ref = 1.0, rn = 0.0;
  if (sl.inner) {
    if (!(tot[1] == tot[1]) || fabs(tot[1]) > 1.0e300) stop = F_NONFINITE;
    else if (!first && !post_refresh && fabs(rz_prev) < 1.0e-35) stop = F_BREAK;
    else if (scal[sl.iters] >= scal[sl.maxit]) stop = F_MAXIT;
  } else {
    if (first && sl.is_b_norm) {
      ref = sqrt(tot[2]);
      if (ref < 1.0e-30) ref = 1.0;
    } else {
      ref = scal[sl.ref];
    }
    rn = sqrt(tot[0]);
    if (!(rn == rn) || rn > 1.0e300) stop = F_NONFINITE;
    else if (rn / ref <= scal[sl.tol]) stop = F_CONV;
    else if (!first && fabs(rz_prev) < 1.0e-35) stop = F_BREAK;
    else if (scal[sl.iters] >= scal[sl.maxit]) stop = F_MAXIT;
  }
  if (lead()) {
    if (!sl.inner) {
      if (first && sl.is_b_norm) scal[sl.ref] = ref;
      scal[sl.rr] = tot[0];
      scal[sl.relres] = rn / ref;
    }
    if (stop != F_RUN) {
      scal[sl.flag] = (double)stop;
    } else {
      scal[sl.iters] += 1.0;
      scal[sl.rz + (k & 1)] = tot[1];
      scal[sl.pend] = 0.0;  // this kernel applies the pending update
    }
  }
  if (stop != F_RUN) return;  // the pending update stays recorded (uniform: every block returns)
  // the update of iteration k-1 is pending unless this is the first
  // iteration (after a reliable update it was folded into the correction)
  const bool pending = !first && !(sl.inner && k == 1);
  // After a reliable update the residual was replaced, and beta = rz / rz_prev
  // no longer makes d conjugate to d_prev (once the FP32 recurrence's
  // residual has fallen below the true one, the ratio is large and the old
  // direction swamps the new one); the conjugate choice
  // beta = -z.(A d_prev) / d_prev.(A d_prev) keeps the step sound.
  double bd = 0.0;
  if (post_refresh) bd = (tot[3] > 0.0) ? -tot[2] / tot[3] : 0.0;
  else if (!first) bd = tot[1] / rz_prev;
  const T beta = (T)bd;
  T acc[1] = {T(0)};
  // the first row's east face at the strip edges inside the block
  if (lane == 0) s_ke0[w] = kw;
  lds_barrier();
  if (right_lds) ke = s_ke0[w + 1];
  T aE = __shfl_down(kw, 1, 64);
  if (right_lds || right_load) aE = ke;
  T aP = kw + aE + aS + kn + (T)o.eps;
  T pcur = p0;
  T cur = zval(r0, aP) + beta * p0;
  T down = zval(rd, dgd) + beta * pd;  // d of the walk's previous row
  if (act) {
    const int nrow = jw1 - jw0 + 1;
    for (int t0 = 0; t0 < nrow; t0 += G) {
      const int j0 = rev ? jw1 - t0 : jw0 + t0;  // batch row g: j0 + g (j0 - g walking down)
      // loads of the batch: the next row in walking order (r or z, d_prev,
      // kx(i), the face away from the row being formed -- ky(i, +1) walking
      // up, ky(i) walking down -- and kx(i+1) at the block's right edge), x,
      // and at the block's edges the x-neighbours (r or z, d_prev, diagonal)
      T ru[G], pu[G], kwu[G], fnu[G], fwu[G], keu[G], xo[G], rl[G], pl[G], dl[G], rr[G], pr[G], dr[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int j = rev ? j0 - g : j0 + g;
        ru[g] = pu[g] = kwu[g] = fnu[g] = fwu[g] = keu[g] = xo[g] = T(0);
        rl[g] = pl[g] = dl[g] = rr[g] = pr[g] = dr[g] = T(0);
        if (t0 + g < nrow && vi) {
          const int ju = rev ? (j > 1 ? j - 1 : (o.perS ? ny : 0)) : (j < ny ? j + 1 : (o.perN ? 1 : 0));
          if (ju > 0) {
            ru[g] = zr[at(i, ju)];
            pu[g] = dprev[at(i, ju)];
            kwu[g] = A.kx[kxi(nx, i, ju)];
            fnu[g] = A.ky[kyi(nx, i, rev ? ju : ju + 1)];
            // periodic wrap: the row's own face toward the row being formed
            if (ju != (rev ? j - 1 : j + 1)) fwu[g] = A.ky[kyi(nx, i, rev ? ju + 1 : ju)];
            if (i == nx) keu[g] = A.kx[kxi(nx, nx + 1, ju)];
            else if (right_load) keu[g] = A.kx[kxi(nx, i + 1, ju)];
          }
          if (pending) xo[g] = x[at(i, j)];
          if (left_load) {
            const int il = i > 1 ? i - 1 : (o.perW ? nx : 0);
            if (il > 0) {
              rl[g] = zr[at(il, j)];
              pl[g] = dprev[at(il, j)];
              if (PC == 1) dl[g] = A.diag[at(il, j)];
            }
          }
          if (i == nx) {
            if (o.perE) {
              rr[g] = zr[at(1, j)];
              pr[g] = dprev[at(1, j)];
              if (PC == 1) dr[g] = A.diag[at(1, j)];
            }
          } else if (right_load) {
            rr[g] = zr[at(i + 1, j)];
            pr[g] = dprev[at(i + 1, j)];
            if (PC == 1) dr[g] = A.diag[at(i + 1, j)];
          }
        }
      }
      // the next wave's west faces of the next rows (this wave's east faces at lane 63)
      if (lane == 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) s_kw[w][g] = kwu[g];
      }
      lds_barrier();
      if (right_lds) {
#pragma unroll
        for (int g = 0; g < G; ++g) keu[g] = s_kw[w + 1][g];
      }
      // the next rows: a_P from their four faces in west, east, south, north
      // order (the face toward the row being formed is that row's face away
      // from them unless the row wraps), z, d; curs[g] = d of batch row g
      T ups[G], aEus[G], aPus[G], aSus[G], aNus[G], curs[G];
      {
        T fsh = rev ? aS : kn;  // face of the row being formed toward the next row
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const int j = rev ? j0 - g : j0 + g;
          T aEu = __shfl_down(kwu[g], 1, 64);
          if (right_lds || right_load) aEu = keu[g];
          const bool wraps = rev ? j == 1 : j == ny;
          const T ftow = wraps ? fwu[g] : fsh;
          const T aSu = rev ? fnu[g] : ftow, aNu = rev ? ftow : fnu[g];
          const T aPu = kwu[g] + aEu + aSu + aNu + (T)o.eps;
          ups[g] = zval(ru[g], aPu) + beta * pu[g];
          aEus[g] = aEu;
          aPus[g] = aPu;
          aSus[g] = aSu;
          aNus[g] = aNu;
          curs[g] = g == 0 ? cur : ups[g - 1];
          fsh = fnu[g];
        }
      }
      // the edge lanes' d of the batch's rows for the adjacent waves
      if (lane == 0) {
#pragma unroll
        for (int g = 0; g < G; ++g) s_c0[w][g] = curs[g];
      }
      if (lane == 63) {
#pragma unroll
        for (int g = 0; g < G; ++g) s_c63[w][g] = curs[g];
      }
      lds_barrier();
#pragma unroll
      for (int g = 0; g < G; ++g) {
        if (t0 + g >= nrow) break;  // uniform over the block
        const int j = rev ? j0 - g : j0 + g;
        T left = __shfl_up(cur, 1, 64);
        T right = __shfl_down(cur, 1, 64);
        if (left_lds) left = s_c63[w - 1][g];
        if (left_load) left = zval(rl[g], dl[g]) + beta * pl[g];
        if (right_lds) right = s_c0[w + 1][g];
        if (right_load) right = zval(rr[g], dr[g]) + beta * pr[g];
        if (vi) {
          const T south = rev ? ups[g] : down, north = rev ? down : ups[g];
          T val = aP * cur;
          if (i > 1 || o.perW) val = val - kw * left;
          if (i < nx || o.perE) val = val - aE * right;
          if (j > 1 || o.perS) val = val - aS * south;
          if (j < ny || o.perN) val = val - kn * north;
          const long q = at(i, j);
          d[q] = cur;
          Ad[q] = val;
          if (pending) x[q] = xo[g] + alpha_prev * pcur;
          acc[0] += cur * val;
        }
        down = cur;
        cur = ups[g];
        pcur = pu[g];
        aS = aSus[g];
        kn = aNus[g];
        kw = kwu[g];
        aE = aEus[g];
        aP = aPus[g];
      }
      // no barrier here: the next batch's writes of s_kw follow this batch's
      // second barrier (every read of s_kw precedes it) and its writes of
      // s_c0 / s_c63 its own first barrier (every read of them precedes it)
    }
  }
  block_partials<1, T>(acc, pout);
}

template <typename T, int PC>
__global__ void __launch_bounds__(BLK) k_pcg_b(DevOp o, const T *__restrict__ diag, T *__restrict__ r,
                                               const T *__restrict__ Ad, double *scal, PcgSlots sl,
                                               const double *pin, double *pout, int k) {
  const double flag_own = scal[sl.flag], flag_outer = scal[S_FLAG];
  const double rz = scal[sl.rz + (k & 1)];
  double tot[1];
  grid_totals<1>(pin, gridDim.x * gridDim.y, tot);
  if (flag_own != 0.0 || flag_outer != 0.0) return;
  const double dAd = tot[0];
  if (!(fabs(dAd) >= 1.0e-35)) {
    if (lead()) scal[sl.flag] = (dAd == dAd) ? (double)F_BREAK : (double)F_NONFINITE;
    return;
  }
  const T alpha = (T)(rz / dAd);
  if (lead()) {
    scal[sl.first] = 0.0;
    scal[sl.step] = (double)alpha;
    scal[sl.pend] = (k & 1) ? 2.0 : 1.0;  // d of iteration k: d1 when k is odd
  }
  T acc[3] = {T(0), T(0), T(0)};
  // CELL_LOOP's rows, U of a wave's rows per batch with the batch's loads
  // issued first (the same rows in the same order: the partials are those of
  // the one-row loop); FP32 loads move half the bytes per instruction, so
  // FP32 batches four rows
  constexpr int U = sizeof(T) == 4 ? 4 : 1;
  CELL_BOUNDS(o);
  if (active_ && i <= o.nx) {
    for (int j0 = jb_ + (threadIdx.x >> 6) + 1; j0 <= je_; j0 += 4 * U) {
      T rv[U], av[U], dv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + 4 * u;
        rv[u] = av[u] = T(0);
        dv[u] = T(1);
        if (j <= je_) {
          const long q = (long)(j - 1) * o.nx + (i - 1);
          rv[u] = r[q];
          av[u] = Ad[q];
          if (PC == 1) dv[u] = diag[q];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int j = j0 + 4 * u;
        if (j > je_) break;
        const long q = (long)(j - 1) * o.nx + (i - 1);
        const T rk = rv[u] - alpha * av[u];
        r[q] = rk;
        acc[0] += rk * rk;
        if (PC == 0) acc[1] += rk * rk;
        if (PC == 1) acc[1] += rk * jacobi<T>(rk, dv[u]);
      }
    }
  }
  block_partials<3, T>(acc, pout);
}

// The solution update left pending by the last B of a loop: x += alpha d
// with the d buffer B recorded (d0 or d1), unless the guard slot is set.
template <typename T>
__global__ void __launch_bounds__(BLK) k_flush_x(DevOp o, T *__restrict__ x, const T *__restrict__ d0,
                                                 const T *__restrict__ d1, const double *scal, int pend, int step,
                                                 int guard) {
  const double pv = scal[pend];
  if (pv == 0.0 || scal[guard] != 0.0) return;
  const T *dl = (pv == 1.0) ? d0 : d1;
  const T alpha = (T)scal[step];
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    x[q] = x[q] + alpha * dl[q];
  }
}

// ---------------------------------------------- single-reduction CG (FP32)
//
// The FP32 recurrence of the mixed mode with Jacobi as one kernel and one
// grid reduction per iteration: the Chronopoulos-Gear form of PCG
// (A. T. Chronopoulos, C. W. Gear, J. Comput. Appl. Math. 25 (1989) 153-168),
// mathematically the same iterates as the two-kernel loop:
//   u = M r, w = A u, gamma = (r, u), delta = (w, u)
//   beta = gamma / gamma_prev, alpha = gamma / (delta - beta gamma / alpha_prev)
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, then the next
//   u = M r, w = A u and the partials of (r, u), (w, u).
// A reads r, w, s, p, x, kx, ky and writes p, s, x, r, w: 12 words per cell
// against the 14 of the two passes.  The next w needs u = M r at the four
// neighbours, formed on the fly from their r, w, s (the same register window
// over a 64-column strip as k_pcg_a).  After a reliable update the residual
// is replaced and the recurrences' orthogonality assumptions no longer hold,
// so the first iteration of a cycle makes the new direction A-conjugate to
// the kept one explicitly: beta = -(u, A p_prev) / (p_prev, A p_prev) and
// (p, A p) = delta + 2 beta (u, A p_prev) + beta^2 (p_prev, A p_prev), from
// the partials of k_cg2_init (which recomputes s = A p_prev); with
// beta = gamma / gamma_prev there instead, the FP32 loop stalls near 1e-6 on
// pressure-Dirichlet (NEUMANN-velocity) boundaries.  r, w, p, s ping-pong
// between two buffers with the parity of k.
struct Cg2Row {
  float s, rn, uo, un, ue;  // s, new r, old / new M r, new M r of the strip-edge neighbour
  float kW, kS, kN, kE, d;
};

__global__ void __launch_bounds__(BLK) k_cg2(DevOp o, OpT<float> A, const float *__restrict__ r,
                                             const float *__restrict__ w, const float *__restrict__ so,
                                             const float *__restrict__ po, float *__restrict__ rn,
                                             float *__restrict__ wn, float *__restrict__ pn, float *__restrict__ sn,
                                             float *__restrict__ x, double *scal, const double *pin, double *pout,
                                             int k) {
  constexpr int G = 4;
  const int nx = o.nx, ny = o.ny;
  const int gx_ = (nx + 63) / 64;
  const int xcd_ = blockIdx.x & 7, local_ = blockIdx.x >> 3;
  const int chunks_ = (gridDim.x >> 3) / gx_;
  const int band_ = (((ny + 7) / 8) + 3) & ~3;
  const int rows_ = (((band_ + chunks_ - 1) / chunks_) + 3) & ~3;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int i = (local_ % gx_) * 64 + lane + 1;
  const int jb = xcd_ * band_ + (local_ / gx_) * rows_;
  const int je = min(min(jb + rows_, xcd_ * band_ + band_), ny);
  const int rw = (je - jb + 3) / 4;
  const int jw0 = jb + 1 + wv * rw, jw1 = min(jw0 + rw - 1, je);
  const bool vi = i <= nx;
  const bool act = local_ < gx_ * chunks_ && jw0 <= jw1;  // uniform over the wave
  const double flag_own = scal[I_FLAG], flag_outer = scal[S_FLAG];
  const bool first = scal[I_FIRST] != 0.0;
  double tot[4];
  grid_totals<4>(pin, gridDim.x, tot, 0, k == 1 ? 4 : 2);
  if (flag_own != 0.0 || flag_outer != 0.0) return;
  const double gam = tot[0], del = tot[1];
  const double gam_prev = scal[I_RZ + ((k - 1) & 1)], alpha_prev = scal[I_ALPHA + ((k - 1) & 1)];
  double beta = 0.0, den = del;
  int stop = F_RUN;
  if (!(gam == gam) || fabs(gam) > 1.0e300) stop = F_NONFINITE;
  else if (!first && fabs(gam_prev) < 1.0e-35) stop = F_BREAK;
  else if (scal[I_ITERS] >= scal[I_MAXIT]) stop = F_MAXIT;
  if (stop == F_RUN && !first) {
    if (k == 1) {
      beta = (tot[3] > 0.0) ? -tot[2] / tot[3] : 0.0;
      den = del + 2.0 * beta * tot[2] + beta * beta * tot[3];
    } else {
      beta = gam / gam_prev;
      den = del - beta * gam / alpha_prev;
    }
  }
  if (stop == F_RUN && !(fabs(den) >= 1.0e-35)) stop = (den == den) ? F_BREAK : F_NONFINITE;
  const double alpha = (stop == F_RUN) ? gam / den : 0.0;
  if (lead()) {
    if (stop != F_RUN) {
      scal[I_FLAG] = (double)stop;
    } else {
      scal[I_ITERS] += 1.0;
      scal[I_RZ + (k & 1)] = gam;
      scal[I_ALPHA + (k & 1)] = alpha;
      if (k == 2) scal[I_FIRST] = 0.0;
    }
  }
  if (stop != F_RUN) return;
  const float fb = (float)beta, fa = (float)alpha, eps = (float)o.eps;
  auto at = [&](int a, int b) { return (long)(b - 1) * nx + (a - 1); };
  // M r_new of cell (a, b) off the strip: its own faces and r, w, s
  auto u_off = [&](int a, int b) {
    const long q = at(a, b);
    const float d = A.kx[kxi(nx, a, b)] + A.kx[kxi(nx, a + 1, b)] + A.ky[kyi(nx, a, b)] + A.ky[kyi(nx, a, b + 1)] + eps;
    const float sv = w[q] + fb * so[q];
    return jacobi<float>(r[q] - fa * sv, d);
  };
  // row R entering the window: everything but the strip-edge neighbour
  auto enter = [&](int R, Cg2Row &E) {
    E.s = E.rn = E.uo = E.un = E.ue = 0.f;
    E.kW = E.kS = E.kN = E.kE = 0.f;
    E.d = 1.f;
    if (!vi) return;
    const long q = at(i, R);
    E.kW = A.kx[kxi(nx, i, R)];
    E.kS = A.ky[kyi(nx, i, R)];
    E.kN = A.ky[kyi(nx, i, R + 1)];
    const float rv = r[q], wv2 = w[q], sv = so[q];
    E.s = wv2 + fb * sv;
    E.rn = rv - fa * E.s;
    if (i == nx || lane == 63) E.kE = A.kx[kxi(nx, i + 1, R)];
    E.uo = rv;  // M applied once d is known (needs kE from the neighbour lane)
    E.un = E.rn;
    if (lane == 0) {
      const int il = i > 1 ? i - 1 : (o.perW ? nx : 0);
      if (il > 0) E.ue = u_off(il, R);
    }
    if (i == nx) {
      if (o.perE) E.ue = u_off(1, R);
    } else if (lane == 63) {
      E.ue = u_off(i + 1, R);
    }
  };
  // finish a row: kE from the right lane, d, M r (all lanes take part in the shuffle)
  auto finish = [&](Cg2Row &E) {
    const float kEs = __shfl_down(E.kW, 1, 64);
    if (!(i == nx || lane == 63)) E.kE = kEs;
    E.d = E.kW + E.kE + E.kS + E.kN + eps;
    E.uo = jacobi<float>(E.uo, E.d);
    E.un = jacobi<float>(E.un, E.d);
  };
  double acc[2] = {0.0, 0.0};
  float accf0 = 0.f, accf1 = 0.f;
  if (act) {
    Cg2Row dn, cu;
    const int jd = jw0 > 1 ? jw0 - 1 : (o.perS ? ny : 0);
    if (jd > 0) enter(jd, dn);
    else {
      dn.s = dn.rn = dn.uo = dn.un = dn.ue = 0.f;
      dn.kW = dn.kS = dn.kN = dn.kE = 0.f;
      dn.d = 1.f;
    }
    enter(jw0, cu);
    finish(dn);
    finish(cu);
    for (int j0 = jw0; j0 <= jw1; j0 += G) {
      Cg2Row up[G];
      float pov[G], xv[G];
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int j = j0 + g;
        pov[g] = xv[g] = 0.f;
        up[g].s = up[g].rn = up[g].uo = up[g].un = up[g].ue = 0.f;
        up[g].kW = up[g].kS = up[g].kN = up[g].kE = 0.f;
        up[g].d = 1.f;
        if (j <= jw1) {
          const int ju = j < ny ? j + 1 : (o.perN ? 1 : 0);
          if (ju > 0) enter(ju, up[g]);
          if (vi) {
            pov[g] = po[at(i, j)];
            xv[g] = x[at(i, j)];
          }
        }
      }
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const int j = j0 + g;
        if (j > jw1) break;  // uniform over the wave
        finish(up[g]);
        float left = __shfl_up(cu.un, 1, 64);
        float right = __shfl_down(cu.un, 1, 64);
        if (lane == 0) left = cu.ue;
        if (i == nx || lane == 63) right = cu.ue;
        if (vi) {
          const long q = at(i, j);
          const float pv = cu.uo + fb * pov[g];
          const float xnew = xv[g] + fa * pv;
          float val = cu.d * cu.un;
          if (i > 1 || o.perW) val = val - cu.kW * left;
          if (i < nx || o.perE) val = val - cu.kE * right;
          if (j > 1 || o.perS) val = val - cu.kS * dn.un;
          if (j < ny || o.perN) val = val - cu.kN * up[g].un;
          pn[q] = pv;
          sn[q] = cu.s;
          x[q] = xnew;
          rn[q] = cu.rn;
          wn[q] = val;
          accf0 += cu.rn * cu.un;
          accf1 += val * cu.un;
        }
        dn = cu;
        cu = up[g];
      }
    }
  }
  acc[0] = accf0;
  acc[1] = accf1;
  block_partials<2, double>(acc, pout);
}

// Start of a cycle: u = M r32 (r32 = the refreshed residual), w = A u, and
// s = A p recomputed from the kept direction p (the recurrence s = w + beta s
// drifts from A p in FP32 over many iterations; the refresh resets it with
// the residual), then the partials (r, u), (w, u), (u, s), (p, s).
__global__ void __launch_bounds__(BLK) k_cg2_init(DevOp o, OpT<float> A, const float *__restrict__ r,
                                                  float *__restrict__ w, float *__restrict__ sp,
                                                  const float *__restrict__ pp, double *pout, const double *scal) {
  if (scal[S_FLAG] != 0.0) return;
  float acc[4] = {0.f, 0.f, 0.f, 0.f};
  const float eps = (float)o.eps;
  auto u_at = [&](long q, int a, int b) {
    const float d = A.kx[kxi(o.nx, a, b)] + A.kx[kxi(o.nx, a + 1, b)] + A.ky[kyi(o.nx, a, b)] +
                    A.ky[kyi(o.nx, a, b + 1)] + eps;
    return jacobi<float>(r[q], d);
  };
  CELL_LOOP(o) {
    const int nx = o.nx, ny = o.ny;
    const long q = (long)(j - 1) * nx + (i - 1);
    const float u = u_at(q, i, j);
    const float aW = A.kx[kxi(nx, i, j)], aE = A.kx[kxi(nx, i + 1, j)];
    const float aS = A.ky[kyi(nx, i, j)], aN = A.ky[kyi(nx, i, j + 1)];
    float val = (aW + aE + aS + aN + eps) * u;
    if (i > 1) val = val - aW * u_at(q - 1, i - 1, j);
    else if (o.perW) val = val - aW * u_at(q + (nx - 1), nx, j);
    if (i < nx) val = val - aE * u_at(q + 1, i + 1, j);
    else if (o.perE) val = val - aE * u_at(q - (nx - 1), 1, j);
    if (j > 1) val = val - aS * u_at(q - nx, i, j - 1);
    else if (o.perS) val = val - aS * u_at(q + (long)(ny - 1) * nx, i, ny);
    if (j < ny) val = val - aN * u_at(q + nx, i, j + 1);
    else if (o.perN) val = val - aN * u_at(q - (long)(ny - 1) * nx, i, 1);
    w[q] = val;
    const float pc = pp[q];
    float ap = (aW + aE + aS + aN + eps) * pc;
    if (i > 1) ap = ap - aW * pp[q - 1];
    else if (o.perW) ap = ap - aW * pp[q + (nx - 1)];
    if (i < nx) ap = ap - aE * pp[q + 1];
    else if (o.perE) ap = ap - aE * pp[q - (nx - 1)];
    if (j > 1) ap = ap - aS * pp[q - nx];
    else if (o.perS) ap = ap - aS * pp[q + (long)(ny - 1) * nx];
    if (j < ny) ap = ap - aN * pp[q + nx];
    else if (o.perN) ap = ap - aN * pp[q - (long)(ny - 1) * nx];
    sp[q] = ap;
    acc[0] += r[q] * u;
    acc[1] += val * u;
    acc[2] += u * ap;
    acc[3] += pc * ap;
  }
  block_partials<4, float>(acc, pout);
}

// r.z with an explicit z (SGS, RB-SGS, multigrid) into slot 1 of B's partials.
template <typename T>
__global__ void __launch_bounds__(BLK) k_pcg_rz(DevOp o, const T *__restrict__ r, const T *__restrict__ z,
                                                const double *scal, PcgSlots sl, double *pout) {
  if (scal[sl.flag] != 0.0 || scal[S_FLAG] != 0.0) return;
  T acc[1] = {T(0)};
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    acc[0] += r[q] * z[q];
  }
  block_partials<1, T>(acc, pout, 1);
}


// Initial residual of the SP loop (:2538-2549): r = b - A x, z = r / a_P
// (Jacobi), partials r.r, r.z, b.b and sum(r).
template <int PC>
__global__ void __launch_bounds__(BLK) k_residual(DevOp o, OpT<double> A, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *__restrict__ r,
                                                  double *__restrict__ z, double *pout, const double *scal) {
  if (scal[S_FLAG] != 0.0) return;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    const double ax = apply_op<double>(o, A, [&](long m) { return x[m]; }, i, j, q);
    const double rk = b[q] - ax;
    r[q] = rk;
    acc[0] += rk * rk;
    if (PC == 0) acc[1] += rk * rk;
    if (PC == 1) {
      const double zk = jacobi<double>(rk, A.diag[q]);
      z[q] = zk;
      acc[1] += rk * zk;
    }
    acc[2] += b[q] * b[q];
    acc[3] += rk;
  }
  block_partials<4, double>(acc, pout);
}

// FP64 residual check of the mixed mode's reliable updates (one block):
// ||b|| (first time), ||r||, converged when ||r||/||b|| <= tol; the FP32
// recurrence's iteration cap ends the solve too.
//
// Constant mode: when no face is a NEUMANN (pressure-Dirichlet) edge, the
// operator is the singular face Laplacian L plus the 1e-12*dvol shift, which
// vanishes from the FP32 copy, so the FP32 operator is singular.  Its
// recurrence must then see a consistent right-hand side and must not carry
// the null vector into x: the inner solve gets r - mean(r) (S_RMEAN) and the
// update adds the correction without its mean (k_add_correction).  x keeps
// the constant of the initial guess; for a right-hand side whose mean is a
// rounding residue (the projection's b = -div u* telescopes to the boundary
// flux) the FP64 residual then converges like the FP64 solve's.
__global__ void __launch_bounds__(BLK) k_residual_check(const double *pin, int nblk, double *scal, int first,
                                                        int singular, double inv_n, int stall_max) {
  if (scal[S_FLAG] != 0.0) return;
  double tot[4];
  grid_totals<4>(pin, nblk, tot);
  if (threadIdx.x != 0) return;
  scal[S_RMEAN] = singular ? tot[3] * inv_n : 0.0;
  scal[I_PEND] = 0.0;  // folded into x by the flush before the correction
  if (first) {
    double bn = sqrt(tot[2]);
    if (bn < 1.0e-30) bn = 1.0;
    scal[S_BNORM] = bn;
  }
  // the residual of the x the solve returns, whatever ends it
  const double rn = sqrt(tot[0]);
  const double rel = rn / scal[S_BNORM];
  scal[S_RR] = tot[0];
  scal[S_RELRES] = rel;
  if (!(rn == rn) || rn > 1.0e300) {
    scal[S_FLAG] = (double)F_NONFINITE;
  } else if (rel <= scal[S_TOL]) {
    scal[S_FLAG] = (double)F_CONV;
  } else if (scal[I_FLAG] == (double)F_MAXIT || scal[I_ITERS] >= scal[I_MAXIT]) {
    scal[S_FLAG] = (double)F_MAXIT;
  } else if (first || rel < 0.9 * scal[S_BEST]) {
    scal[S_BEST] = rel;
    scal[S_STALL] = 0.0;
  } else if ((scal[S_STALL] += 1.0) >= (double)stall_max) {
    // stall_max reliable updates without a 10 % gain over the best residual:
    // the FP32 recurrence has hit the attainable accuracy of this system (the
    // FP32 operator's rounding times its condition number), and the host
    // finishes the solve in FP64 from this x
    scal[S_FLAG] = (double)F_STALL;
  }
}

// ||b - A x|| and ||b|| partials with no stop flag (the true residual of a
// finished solve), then ||b - A x|| / ||b|| (||b|| < 1e-30 counts as 1) into
// *out.
__global__ void __launch_bounds__(BLK) k_true_res(DevOp o, OpT<double> A, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *pout) {
  double acc[2] = {0.0, 0.0};
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    const double rk = b[q] - apply_op<double>(o, A, [&](long m) { return x[m]; }, i, j, q);
    acc[0] += rk * rk;
    acc[1] += b[q] * b[q];
  }
  block_partials<2, double>(acc, pout);
}

__global__ void __launch_bounds__(BLK) k_true_res_finish(const double *pin, int nblk, double *out) {
  double tot[2];
  grid_totals<2>(pin, nblk, tot);
  if (threadIdx.x != 0) return;
  double bn = sqrt(tot[1]);
  if (bn < 1.0e-30) bn = 1.0;
  *out = sqrt(tot[0]) / bn;
}

// sum(dx32) into slot 0 of this kernel's partials (singular operator only).
__global__ void __launch_bounds__(BLK) k_sum32(DevOp o, const float *__restrict__ x, double *pout,
                                               const double *scal) {
  if (scal[S_FLAG] != 0.0) return;
  double acc[1] = {0.0};
  CELL_LOOP(o) { acc[0] += (double)x[(long)(j - 1) * o.nx + (i - 1)]; }
  block_partials<1, double>(acc, pout);
}

// Reliable update x += (double) dx32 (without dx32's mean when the operator
// is singular), dx32 = 0.  pin: the partials of k_sum32 (nblk blocks), read
// when singular.
__global__ void __launch_bounds__(BLK) k_add_correction(long n, double *__restrict__ x, float *__restrict__ dx32,
                                                        const double *scal, const double *pin, int nblk,
                                                        int singular, double inv_n) {
  if (scal[S_FLAG] != 0.0) return;
  double shift = 0.0;
  if (singular) {
    double tot[1];
    grid_totals<1>(pin, nblk, tot);
    shift = -(tot[0] * inv_n);
  }
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    x[t] = x[t] + ((double)dx32[t] + shift);
    dx32[t] = 0.0f;
  }
}

// FP32 inner-solve setup: r32 = r - S_RMEAN rounded to FP32, z32 = r32 / a_P
// (Jacobi) or r32, partials r.r, r.z, r.r.
template <int PC>
__global__ void __launch_bounds__(BLK) k_inner_init(DevOp o, const double *__restrict__ r,
                                                    float *__restrict__ r32, const float *__restrict__ d32,
                                                    float *__restrict__ z32, double *pout, double *scal) {
  const double flag = scal[S_FLAG], rmean = scal[S_RMEAN];
  if (flag != 0.0) return;
  if (lead()) {  // a cycle the FP32 recurrence left early (breakdown, non-finite) restarts it
    if (scal[I_FLAG] != 0.0) scal[I_FIRST] = 1.0;
    scal[I_FLAG] = 0.0;
  }
  float acc[3] = {0.f, 0.f, 0.f};
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    const float rk = (float)(r[q] - rmean);
    r32[q] = rk;
    acc[0] += rk * rk;
    if (PC == 0) acc[1] += rk * rk;
    if (PC == 1) {
      const float zk = jacobi<float>(rk, d32[q]);
      z32[q] = zk;
      acc[1] += rk * zk;
    }
  }
  acc[2] = acc[0];
  block_partials<3, float>(acc, pout);
}

// Slots 2, 3 of a reliable update's partials: z.(A d_prev), d_prev.(A d_prev)
// for the conjugate restart of the FP32 recurrence (d_prev, A d_prev: the
// last iteration's direction and product).
__global__ void __launch_bounds__(BLK) k_refresh_conj(DevOp o, const float *__restrict__ z, const float *__restrict__ dp,
                                                      const float *__restrict__ adp, double *pout,
                                                      const double *scal) {
  if (scal[S_FLAG] != 0.0) return;
  float acc[2] = {0.f, 0.f};
  CELL_LOOP(o) {
    const long q = (long)(j - 1) * o.nx + (i - 1);
    const float a = adp[q];
    acc[0] += z[q] * a;
    acc[1] += dp[q] * a;
  }
  block_partials<2, float>(acc, pout, 2);
}

// p = x - mean(x) (src/module_simulation.f90:2578-2585) into p(1:nx,1:ny).
__global__ void __launch_bounds__(BLK) k_sum(DevOp o, const double *__restrict__ x, double *pout) {
  double acc[1] = {0.0};
  CELL_LOOP(o) { acc[0] += x[(long)(j - 1) * o.nx + (i - 1)]; }
  block_partials<1, double>(acc, pout);
}

__global__ void __launch_bounds__(BLK) k_sum_finish(const double *pin, int nblk, double *scal, int slot) {
  double tot[1];
  grid_totals<1>(pin, nblk, tot);
  if (threadIdx.x == 0) scal[slot] = tot[0];
}

__global__ void k_write_pressure(DevOp o, const double *__restrict__ x, const double *scal, int slot,
                                 double *p) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > o.nx || j > o.ny) return;
  const double mean = scal[slot] / (double)((long)o.nx * o.ny);
  p[(long)j * (o.nx + 2) + i] = x[(long)(j - 1) * o.nx + (i - 1)] - mean;
}

// sync_cell_scalar on one rank: periodic wrap of the halo (src/module_mpi_comm.f90:46-56).
__global__ void k_wrap_halo(int nx, int ny, int perx, int pery, double *p) {
  const int t = threadIdx.x;
  auto H = [&](int a, int c) { return (long)c * (nx + 2) + a; };
  if (perx) {
    for (int j = t; j <= ny + 1; j += blockDim.x) p[H(0, j)] = p[H(nx, j)];
    __syncthreads();
    for (int j = t; j <= ny + 1; j += blockDim.x) p[H(nx + 1, j)] = p[H(1, j)];
    __syncthreads();
  }
  if (pery) {
    for (int i = t; i <= nx + 1; i += blockDim.x) p[H(i, 0)] = p[H(i, ny)];
    __syncthreads();
    for (int i = t; i <= nx + 1; i += blockDim.x) p[H(i, ny + 1)] = p[H(i, 1)];
  }
}

__global__ void k_gather_interior(int nx, int ny, const double *__restrict__ p, double *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > nx || j > ny) return;
  x[(long)(j - 1) * nx + (i - 1)] = p[(long)j * (nx + 2) + i];
}

// Velocity correction (src/module_simulation.f90:2590-2654); solid obstacle
// faces are skipped like walls.
__global__ void k_correct(int nx, int ny, double dx, double dy, double dt, int bl, int br, int bb,
                          int bt, int wallL, int wallR, int wallB, int wallT, int perL, int perR,
                          int perB, int perT, DevObs ob, const double *__restrict__ p,
                          const double *__restrict__ rho_u, const double *__restrict__ rho_v, double *u,
                          double *v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  auto H = [&](int a, int c) { return (long)c * (nx + 2) + a; };
  const int u_nx = br ? nx + 1 : nx, v_ny = bt ? ny + 1 : ny;
  if (i <= u_nx && j <= ny) {
    const bool wall = (i == 1 && bl && wallL) || (i == u_nx && br && wallR) || solid_u(ob, i, j, dx, dy);
    if (!wall) {
      double r = rho_u[H(i, j)];
      if (r < 1.0e-12) r = 1.0;
      const double pw = (i > 1) ? p[H(i - 1, j)] : (perL ? p[H(nx, j)] : p[H(1, j)]);
      const double pe = (i <= nx) ? p[H(i, j)] : (perR ? p[H(1, j)] : p[H(nx, j)]);
      const double g = (pe - pw) / dx;
      u[H(i, j)] = u[H(i, j)] - (dt / r) * g;
    }
  }
  if (i <= nx && j <= v_ny) {
    const bool wall = (j == 1 && bb && wallB) || (j == v_ny && bt && wallT) || solid_v(ob, i, j, dx, dy);
    if (!wall) {
      double r = rho_v[H(i, j)];
      if (r < 1.0e-12) r = 1.0;
      const double ps = (j > 1) ? p[H(i, j - 1)] : (perB ? p[H(i, ny)] : p[H(i, 1)]);
      const double pn = (j <= ny) ? p[H(i, j)] : (perT ? p[H(i, 1)] : p[H(i, ny)]);
      const double g = (pn - ps) / dy;
      v[H(i, j)] = v[H(i, j)] - (dt / r) * g;
    }
  }
}

// ================================================================= host side

static DevOp dev_op(const Mesh &m) {
  DevOp o;
  o.nx = m.nx;
  o.ny = m.ny;
  o.perW = m.bc[0] == BC_PERIODIC;
  o.perE = m.bc[1] == BC_PERIODIC;
  o.perS = m.bc[2] == BC_PERIODIC;
  o.perN = m.bc[3] == BC_PERIODIC;
  o.eps = 1.0e-12 * (m.dx * m.dy);
  return o;
}

// 1-D grid of 8 XCD bands x gx strips x chunks (about 1024 blocks); see CELL_LOOP.
static dim3 red_grid(const Mesh &m) {
  const int gx = (m.nx + 63) / 64;
  const int band = (((m.ny + 7) / 8) + 3) & ~3;
  int chunks = RED_BLOCKS / (8 * gx);
  if (chunks < 1) chunks = 1;
  const int max_chunks = (band + 3) / 4;  // at least 4 rows per chunk
  if (chunks > max_chunks) chunks = max_chunks;
  return dim3(8 * gx * chunks);
}

int poisson_alloc(Ctx &c) {
  Poisson &P = c.ps;
  const long n = (long)c.m.nx * c.m.ny;
  const long nkx = kx_pitch(c.m.nx) * c.m.ny, nky = (long)c.m.nx * (c.m.ny + 1);
  COVOF_CHECK(hipMalloc(&P.kx, nkx * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.ky, nky * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.diag, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.kx32, nkx * sizeof(float)));
  COVOF_CHECK(hipMalloc(&P.ky32, nky * sizeof(float)));
  COVOF_CHECK(hipMalloc(&P.diag32, n * sizeof(float)));
  double **d64[] = {&P.b, &P.x, &P.r, &P.p, &P.q, &P.z, &P.t0, &P.t1, &P.t2, &P.t3, &P.t4, &P.t5};
  for (double **pp : d64) COVOF_CHECK(hipMalloc(pp, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.fx, nkx * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.fy, nky * sizeof(double)));
  float **d32[] = {&P.r32, &P.q32, &P.x32, &P.z32};
  for (float **pp : d32) COVOF_CHECK(hipMalloc(pp, n * sizeof(float)));
  // p32 holds the two ping-pong direction buffers (2n)
  COVOF_CHECK(hipMalloc(&P.p32, 2 * n * sizeof(float)));
  const dim3 g = red_grid(c.m);
  P.nblocks = g.x * g.y;
  // partials: up to 4 values per block of the fused kernels, 5 per block of
  // the 512-block inner-product kernel of BiCGSTAB(2)
  const long npart = std::max<long>((long)P.nblocks * NP * 3, 512L * 9);
  COVOF_CHECK(hipMalloc(&P.part, npart * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.scal, NSLOT * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.ticket, sizeof(unsigned)));
  COVOF_CHECK(hipMemset(P.ticket, 0, sizeof(unsigned)));
  COVOF_CHECK(hipMemset(P.scal, 0, NSLOT * sizeof(double)));
  COVOF_CHECK(hipHostMalloc(&P.hscal, NSLOT * sizeof(double), hipHostMallocDefault));
  COVOF_CHECK(hipEventCreateWithFlags(&P.cycle_ev[0], hipEventDisableTiming));
  COVOF_CHECK(hipEventCreateWithFlags(&P.cycle_ev[1], hipEventDisableTiming));
  return COVOF_OK;
}

void poisson_free(Ctx &c) {
  Poisson &P = c.ps;
  void *ptrs[] = {P.kx, P.ky, P.diag, P.kx32, P.ky32, P.diag32, P.b, P.x, P.r, P.p, P.q, P.z,
                  P.t0, P.t1, P.t2, P.t3, P.t4, P.t5, P.r32, P.p32, P.q32, P.x32, P.z32, P.part, P.scal, P.ticket,
                  P.fx, P.fy, P.cg2, P.sgs_rec, P.sgs_xa, P.sgs_xb, P.sgs_h};
  for (void *q : ptrs) hipFree(q);
  if (P.hscal) hipHostFree(P.hscal);
  for (PGraph &G : P.graph)
    if (G.exec) hipGraphExecDestroy(G.exec);
  mg_free(c);
  if (P.cycle_ev[0]) hipEventDestroy(P.cycle_ev[0]);
  if (P.cycle_ev[1]) hipEventDestroy(P.cycle_ev[1]);
}

static int is_wall_type(int t) { return t == BC_DIRICHLET || t == BC_FREESLIP; }

static DevObs dev_obs(const Ctx &c) {
  const Obstacle &o = c.obs;
  DevObs d{o.enabled && o.perm <= 1.0e-12 ? 1 : 0, o.type, o.x0, o.y0, o.x1, o.y1, o.xc, o.yc, o.r};
  return d;
}

int poisson_assemble(Ctx &c, const double *u, const double *v, const double *rho_u,
                     const double *rho_v, double dt, double *b) {
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 63) / 64, (m.ny + 3) / 4);
  // is_wall_* of :2490-2526: wall faces need is_boundary_* (non-periodic) and a wall BC type
  const int wW = !m.perx && is_wall_type(m.bc[0]);
  const int wE = !m.perx && is_wall_type(m.bc[1]);
  const int wS = !m.pery && is_wall_type(m.bc[2]);
  const int wN = !m.pery && is_wall_type(m.bc[3]);
  hipLaunchKernelGGL(k_assemble, grd, blk, 0, c.stream, u, v, rho_u, rho_v, m.nx, m.ny, m.dx, m.dy,
                     dt, wW, wE, wS, wN, dev_obs(c), c.ps.kx, c.ps.ky, c.ps.diag, b, c.ps.kx32, c.ps.ky32,
                     c.ps.diag32);
  COVOF_CHECK_LAUNCH();
  c.ps.assembled = 1;
  ++c.ps.assembly_id;
  return COVOF_OK;
}

int poisson_matvec(Ctx &c, const double *x, double *y) {
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  const DevOp o = dev_op(c.m);
  OpT<double> A{c.ps.kx, c.ps.ky, c.ps.diag};
  hipLaunchKernelGGL((k_matvec<double>), red_grid(c.m), dim3(BLK), 0, c.stream, o, A, x, y);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// Preconditioner codes of the fused PCG: 0 none, 1 Jacobi, 2 exact SGS
// (lexicographic, sequential), 3 red-black SGS, 4 multigrid V-cycle.
static int pc_code(int precond) {
  switch (precond) {
    case COVOF_PRECOND_NONE: return 0;
    case COVOF_PRECOND_JACOBI: return 1;
    case COVOF_PRECOND_SSOR: return 2;
    case COVOF_PRECOND_SGS_RB: return 3;
    case COVOF_PRECOND_MG: return 4;
    default: return -1;
  }
}
static int mg_kind(int pc) { return pc == 3 ? MG_KIND_SGS_RB : MG_KIND_VCYCLE; }

// z = M^-1 r for the explicit preconditioners of precision T.

// z = SGS(r) on the strip sweeps (records rebuilt after each assembly);
// flag: a device scalar that skips the application when nonzero, or null.
int sgs_reserve(Ctx &c) {
  Poisson &P = c.ps;
  if (P.sgs_rec) return COVOF_OK;
  const long len = seq_len(c.m.nx, c.m.ny);
  COVOF_CHECK(hipMalloc(&P.sgs_rec, 8 * len * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.sgs_xa, len * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.sgs_xb, len * sizeof(double)));
  const long nh = c.m.nx > c.m.ny ? c.m.nx : c.m.ny;
  COVOF_CHECK(hipMalloc(&P.sgs_h, 4 * nh * sizeof(double)));
  return COVOF_OK;
}

static int sgs_apply(Ctx &c, const double *r, double *z, const double *flag) {
  Poisson &P = c.ps;
  const DevOp o = dev_op(c.m);
  const int S = o.nx + 63;
  const long n = (long)o.nx * o.ny;
  const dim3 g((unsigned)((n + BLK - 1) / BLK));
  const int rs = sgs_reserve(c);
  if (rs) return rs;
  if (P.sgs_id != P.assembly_id) {
    OpT<double> A{P.kx, P.ky, P.diag};
    hipLaunchKernelGGL(k_sgs_records, g, dim3(BLK), 0, c.stream, o, A, S, P.sgs_rec);
    P.sgs_id = P.assembly_id;
  }
  const long nh = o.nx > o.ny ? o.nx : o.ny;
  SeqBlk b{o.nx, o.ny, (o.ny + 63) / 64, S, P.sgs_rec, P.sgs_xa, P.sgs_xb,
           P.sgs_h, P.sgs_h + nh, P.sgs_h + 2 * nh, P.sgs_h + 3 * nh};
  hipLaunchKernelGGL(k_sgs_in, g, dim3(BLK), 0, c.stream, o, S, r, P.sgs_rec, flag);
  hipLaunchKernelGGL(k_sgs_sweep, dim3(1), dim3(64 * seq_waves(o.ny)), 0, c.stream, o, b, flag);
  hipLaunchKernelGGL(k_sgs_out, g, dim3(BLK), 0, c.stream, o, S, (const double *)P.sgs_xa, z, flag);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}
template <typename T>
static bool explicit_precond(Ctx &c, int pc, const T *r, T *z, int fslot, double *rz = nullptr);
// z = M^-1 r; rz: the reduction grid's r.z slot, filled by the multigrid
// cycle's finest pass when it can (the return value says whether it did)
template <>
bool explicit_precond<double>(Ctx &c, int pc, const double *r, double *z, int fslot, double *rz) {
  if (pc == 2) {
    sgs_apply(c, r, z, (const double *)(c.ps.scal + fslot));
    return false;
  }
  if (pc == 4 && c.ps.mg_fp32_cycle) return mg_apply_fp32_cycle(c, r, z, S_FLAG, fslot, rz);
  return mg_apply64(c, mg_kind(pc), r, z, S_FLAG, fslot, rz);
}
template <>
bool explicit_precond<float>(Ctx &c, int pc, const float *r, float *z, int fslot, double *) {
  mg_apply32(c, mg_kind(pc), r, z, S_FLAG, fslot);
  return false;
}

int poisson_precond(Ctx &c, int precond, const double *r, double *z) {
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  const long n = (long)c.m.nx * c.m.ny;
  if (precond == COVOF_PRECOND_SSOR) {
    const int st = sgs_apply(c, r, z, nullptr);
    if (st) return st;
  } else if (precond == COVOF_PRECOND_JACOBI || precond == COVOF_PRECOND_NONE) {
    hipLaunchKernelGGL(k_precond_simple, dim3((n + 255) / 256), dim3(256), 0, c.stream, precond, n,
                       c.ps.diag, r, z);
  } else if (precond == COVOF_PRECOND_SGS_RB || precond == COVOF_PRECOND_MG) {
    const int pc = pc_code(precond);
    if (mg_check(c, mg_kind(pc))) return COVOF_BAD_ARGUMENT;
    int st;
    if ((st = mg_prepare(c, false))) return st;
    mg_apply64(c, mg_kind(pc), r, z, S_ZERO, S_ZERO);
  } else {
    return COVOF_BAD_ARGUMENT;
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// Reads back the scalar slots [0, n) (synchronises the stream).
static int read_scalars(Ctx &c, int n) {
  COVOF_CHECK(hipMemcpyAsync(c.ps.hscal, c.ps.scal, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  COVOF_CHECK(hipStreamSynchronize(c.stream));
  prof_harvest(c);
  return COVOF_OK;
}

static int set_scalar(Ctx &c, int slot, double v) {
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, c.ps.scal, slot, 1, v);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

template <typename T>
static int fill(Ctx &c, T *x, long n, T v) {
  hipLaunchKernelGGL((k_fill<T>), dim3(1024), dim3(256), 0, c.stream, n, x, v);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static const PcgSlots OUTER = {S_RZ,  S_RR,    S_RELRES, S_FLAG, S_ITERS, S_BNORM, S_TOL,
                               S_FIRST, S_MAXIT, 1, 0, S_STEP, S_PEND};
static const PcgSlots INNER = {I_RZ,    I_RR,    I_RELRES, I_FLAG, I_ITERS, S_BNORM, S_TOL,
                               I_FIRST, I_MAXIT, 0, 1, I_STEP, I_PEND};
constexpr int NREAD = 32;  // scalar slots read back by the host

// Partial buffers: [0] written by A kernels, [1] by B / residual / rz kernels,
// [2] by the sum kernel.
static double *part_buf(Ctx &c, int w) { return c.ps.part + (long)w * c.ps.nblocks * NP; }
static_assert(NP % 2 == 0, "partial buffers start on 16-byte boundaries");

static int status_of_flag(double f) {
  if (f == F_CONV) return COVOF_OK;
  if (f == F_BREAK) return COVOF_BREAKDOWN;
  if (f == F_NONFINITE) return COVOF_NONFINITE;
  return COVOF_NOT_CONVERGED;
}

// The serial sums that replace a B kernel's partials in reference-order mode:
// r.r into slot 0 and r.z into slot 1 of pb (z = r without a preconditioner,
// Jacobi's r / a_P formed on the fly, else the stored z), and b.b into slot 2
// for the initial residual.
static SerialSet rz_terms(Ctx &c, int pc, const double *r, const double *z, const double *diag, double *pb,
                          bool with_b) {
  SerialSet s{};
  s.nzero = c.ps.nblocks;
  s.a[0] = r;
  s.b[0] = r;
  s.dst[0] = pb;
  s.a[1] = r;
  if (pc == 0) s.b[1] = r;
  else if (pc == 1) s.dg[1] = diag;
  else s.b[1] = z;
  s.dst[1] = pb + c.ps.nblocks;
  s.nq = 2;
  if (with_b) {
    s.a[2] = c.ps.b;
    s.b[2] = c.ps.b;
    s.dst[2] = pb + 2L * c.ps.nblocks;
    s.nq = 3;
  }
  return s;
}

// One PCG iteration k (A_k, B_k) on (x, r, z); A reads r (PC 0/1: z is
// formed from r) or z (explicit preconditioners); d_prev / d alternate between
// d0 and d1 with the parity of k (the first iteration, k odd, reads d0, which
// must hold zeros).  When `timed`, the two launches are bracketed by events.
template <typename T, int PC>
static void pcg_iteration(Ctx &c, OpT<T> A, T *x, T *r, T *z, T *d0, T *d1, T *Ad, const PcgSlots &sl,
                          int k, bool timed) {
  const DevOp o = dev_op(c.m);
  const dim3 g = red_grid(c.m);
  double *pa = part_buf(c, 0), *pb = part_buf(c, 1);
  const T *zr = (PC <= 1) ? r : z;
  T *dprev = (k & 1) ? d0 : d1;
  T *dnew = (k & 1) ? d1 : d0;
  const int tag = timed ? (sizeof(T) == 8 ? COVOF_KERNEL_PCG_A_FP64 : COVOF_KERNEL_PCG_A_FP32) : -1;
  // reference-order mode (FP64): each grid reduction is replaced by the
  // serial sum, written as the slot's only nonzero partial
  const bool ser = sizeof(T) == 8 && c.ref_order;
  const long n = (long)o.nx * o.ny;
  launch_tagged(c, tag, k_pcg_a<T, PC>, g, dim3(BLK), o, A, zr, (const T *)dprev, dnew, Ad, x, c.ps.scal, sl,
                (const double *)pb, pa, k);
  if (ser) {
    SerialSet s{};
    s.nq = 1;
    s.nzero = c.ps.nblocks;
    s.a[0] = (const double *)dnew;
    s.b[0] = (const double *)Ad;
    s.dst[0] = pa;
    serial_sums(c, n, s, sl.flag, S_FLAG);
  }
  launch_tagged(c, timed ? tag + 1 : -1, k_pcg_b<T, PC>, g, dim3(BLK), o, (const T *)A.diag, r, (const T *)Ad,
                c.ps.scal, sl, (const double *)pa, pb, k);
  // (FP64 with multigrid: the cycle's finest pass forms the r.z partials)
  bool rz_done = false;
  if (PC >= 2) rz_done = explicit_precond<T>(c, PC, (const T *)r, z, sl.flag, ser ? nullptr : pb + c.ps.nblocks);
  if (ser) {
    serial_sums(c, n, rz_terms(c, PC, (const double *)r, (const double *)z, (const double *)A.diag, pb, false),
                sl.flag, S_FLAG);
  } else if (PC >= 2 && !rz_done) {
    hipLaunchKernelGGL(k_pcg_rz<T>, g, dim3(BLK), 0, c.stream, o, (const T *)r, (const T *)z, c.ps.scal, sl, pb);
  }
}

// Applies the solution update a loop left pending (see k_pcg_a).
template <typename T>
static void flush_x(Ctx &c, T *x, const T *d0, const T *d1, const PcgSlots &sl, int guard) {
  hipLaunchKernelGGL(k_flush_x<T>, red_grid(c.m), dim3(BLK), 0, c.stream, dev_op(c.m), x, d0, d1,
                     (const double *)c.ps.scal, sl.pend, sl.step, guard);
}

// The single-reduction loop's buffers: r, w, p, s at parity 0 and 1.
struct Cg2Bufs {
  float *r[2], *w[2], *p[2], *s[2];
};
static Cg2Bufs cg2_bufs(Ctx &c) {
  const long n = (long)c.m.nx * c.m.ny;
  float *b = c.ps.cg2;
  return Cg2Bufs{{b, b + n}, {b + 2 * n, b + 3 * n}, {b + 4 * n, b + 5 * n}, {b + 6 * n, b + 7 * n}};
}

// Iteration k of the single-reduction loop: reads parity (k+1)&1, writes k&1.
static void cg2_iteration(Ctx &c, int k, bool timed) {
  Poisson &P = c.ps;
  const Cg2Bufs B = cg2_bufs(c);
  const int ri = (k + 1) & 1, wo = k & 1;
  OpT<float> A32{P.kx32, P.ky32, P.diag32};
  launch_tagged(c, timed ? COVOF_KERNEL_PCG_A_FP32 : -1, k_cg2, red_grid(c.m), dim3(BLK), dev_op(c.m), A32,
                (const float *)B.r[ri], (const float *)B.w[ri], (const float *)B.s[ri], (const float *)B.p[ri],
                B.r[wo], B.w[wo], B.p[wo], B.s[wo], P.x32, P.scal, (const double *)part_buf(c, ri),
                part_buf(c, wo), k);
}

// Whether the mixed Jacobi loop runs as the single-reduction kernel (the
// context's mixed_loop option) or as the two-kernel PCG (the default: at
// 2048^2 it is as fast per iteration and needs 4 % fewer iterations).
static bool mixed_cg2(const Ctx &c, int pc) { return pc == 1 && c.opts.mixed_loop == 1; }

// The reliable update of the mixed mode: x += x32, x32 = 0, r = b - A x in
// FP64 (+ its FP32 copy), the convergence test, and the FP32 recurrence's
// restart values (z32, r.r, r.z) from the true residual.
static void mixed_refresh(Ctx &c, int pc, const double *b, double *x, int first) {
  Poisson &P = c.ps;
  const long n = (long)c.m.nx * c.m.ny;
  const DevOp o = dev_op(c.m);
  const dim3 g = red_grid(c.m);
  OpT<double> A{P.kx, P.ky, P.diag};
  double *pb = part_buf(c, 1), *ps = part_buf(c, 2);
  const Mesh &m = c.m;
  const int singular = (m.bc[0] != BC_NEUMANN && m.bc[1] != BC_NEUMANN && m.bc[2] != BC_NEUMANN &&
                        m.bc[3] != BC_NEUMANN) ? 1 : 0;
  const double inv_n = 1.0 / (double)n;
  if (!first) {
    // the FP32 loop's last solution update, then the correction
    flush_x<float>(c, P.x32, P.p32, P.p32 + n, INNER, S_FLAG);
    if (singular)
      hipLaunchKernelGGL(k_sum32, g, dim3(BLK), 0, c.stream, o, (const float *)P.x32, ps, (const double *)P.scal);
    hipLaunchKernelGGL(k_add_correction, dim3(1024), dim3(BLK), 0, c.stream, n, x, P.x32, (const double *)P.scal,
                       (const double *)ps, P.nblocks, singular, inv_n);
  }
  hipLaunchKernelGGL((k_residual<0>), g, dim3(BLK), 0, c.stream, o, A, b, (const double *)x, P.r, P.z, pb,
                     (const double *)P.scal);
  // stall window: 48 updates of 128 Jacobi iterations (CG's residual norm
  // plateaus for about a thousand Jacobi iterations at a time on the
  // density-jump systems; this is far past that), 16 multigrid iterations
  // (a converging MG-PCG gains far more than 10 % per 8 iterations; fitted
  // on the 4096 x 2048 bubble step); the context's mixed_stall option
  // overrides it
  const int stall_max = c.opts.mixed_stall > 0 ? c.opts.mixed_stall : (pc == 4 ? 2 : 48);
  hipLaunchKernelGGL(k_residual_check, dim3(1), dim3(BLK), 0, c.stream, (const double *)pb, P.nblocks, P.scal,
                     first, singular, inv_n, stall_max);
  const bool cg2 = mixed_cg2(c, pc);
  // the single-reduction loop restarts from r32 in its parity-0 r buffer
  float *r32 = cg2 ? cg2_bufs(c).r[0] : P.r32;
  if (pc == 1 && !cg2)
    hipLaunchKernelGGL((k_inner_init<1>), g, dim3(BLK), 0, c.stream, o, (const double *)P.r, r32,
                       (const float *)P.diag32, P.z32, pb, P.scal);
  else
    hipLaunchKernelGGL((k_inner_init<0>), g, dim3(BLK), 0, c.stream, o, (const double *)P.r, r32,
                       (const float *)P.diag32, P.z32, pb, P.scal);
  if (pc >= 3) {  // explicit z32 = M^-1 r32, r.z into slot 1
    explicit_precond<float>(c, pc, (const float *)P.r32, P.z32, INNER.flag);
    hipLaunchKernelGGL(k_pcg_rz<float>, g, dim3(BLK), 0, c.stream, o, (const float *)P.r32, (const float *)P.z32,
                       P.scal, INNER, pb);
  }
  if (!cg2) {  // conjugate restart of the two-kernel loop (d_prev = d0, A d_prev = q32)
    const float *z32 = (pc == 0) ? (const float *)P.r32 : (const float *)P.z32;
    hipLaunchKernelGGL(k_refresh_conj, g, dim3(BLK), 0, c.stream, o, z32, (const float *)P.p32,
                       (const float *)P.q32, pb, (const double *)P.scal);
  }
  if (cg2) {
    const Cg2Bufs B = cg2_bufs(c);
    OpT<float> A32{P.kx32, P.ky32, P.diag32};
    hipLaunchKernelGGL(k_cg2_init, g, dim3(BLK), 0, c.stream, o, A32, (const float *)B.r[0], B.w[0], B.s[0],
                       (const float *)B.p[0], part_buf(c, 0), (const double *)P.scal);
  }
}

// One solver cycle: ITERS PCG iterations (FP64 solve) or ITERS FP32
// iterations + one reliable update (mixed solve), then a copy of the scalar
// slots to pinned host memory.  All pointers are context buffers.  With
// profiling, every prof.on-th iteration's two kernels are bracketed by events.
static void enqueue_cycle(Ctx &c, int pc, bool mixed, int iters) {
  Poisson &P = c.ps;
  const long n = (long)c.m.nx * c.m.ny;
  OpT<double> A{P.kx, P.ky, P.diag};
  OpT<float> A32{P.kx32, P.ky32, P.diag32};
  for (int k = 1; k <= iters; ++k) {
    const bool timed = prof_sampled(c, k);
    if (mixed) {
      float *d0 = P.p32, *d1 = P.p32 + n;
      if (mixed_cg2(c, pc)) cg2_iteration(c, k, timed);
      else if (pc == 1) pcg_iteration<float, 1>(c, A32, P.x32, P.r32, P.z32, d0, d1, P.q32, INNER, k, timed);
      else if (pc == 3) pcg_iteration<float, 3>(c, A32, P.x32, P.r32, P.z32, d0, d1, P.q32, INNER, k, timed);
      else if (pc == 4) pcg_iteration<float, 4>(c, A32, P.x32, P.r32, P.z32, d0, d1, P.q32, INNER, k, timed);
      else pcg_iteration<float, 0>(c, A32, P.x32, P.r32, P.z32, d0, d1, P.q32, INNER, k, timed);
    } else {
      if (pc == 1) pcg_iteration<double, 1>(c, A, P.x, P.r, P.z, P.p, P.t0, P.q, OUTER, k, timed);
      else if (pc == 3) pcg_iteration<double, 3>(c, A, P.x, P.r, P.z, P.p, P.t0, P.q, OUTER, k, timed);
      else if (pc == 4) pcg_iteration<double, 4>(c, A, P.x, P.r, P.z, P.p, P.t0, P.q, OUTER, k, timed);
      else pcg_iteration<double, 0>(c, A, P.x, P.r, P.z, P.p, P.t0, P.q, OUTER, k, timed);
    }
  }
  if (mixed) mixed_refresh(c, pc, P.b, P.x, 0);
  hipMemcpyAsync(P.hscal, P.scal, NREAD * sizeof(double), hipMemcpyDeviceToHost, c.stream);
}

// The cycle captured once into a HIP graph (one instantiation serves every
// solve of a context with the same selectors: only context buffers appear).
static int capture_cycle(Ctx &c, PGraph &G, int key, int pc, bool mixed, int iters) {
  if (G.exec && G.key == key) return COVOF_OK;
  if (G.exec) {
    hipGraphExecDestroy(G.exec);
    G.exec = nullptr;
  }
  COVOF_CHECK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeRelaxed));
  enqueue_cycle(c, pc, mixed, iters);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(c.stream, &graph);
  if (ec != hipSuccess || !graph) return COVOF_HIP_ERROR;
  const hipError_t ei = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ei != hipSuccess) return COVOF_HIP_ERROR;
  G.key = key;
  return COVOF_OK;
}

// Runs cycles until the stop flag is set, keeping one cycle in flight while
// the previous one's scalars are checked.  Graph replays normally; with
// profiling on, the same kernels are launched directly so events can bracket
// the sampled iterations.
static int run_cycles(Ctx &c, PGraph &G, int key, int pc, bool mixed, int iters, int max_cycles) {
  Poisson &P = c.ps;
  const bool direct = c.prof.on > 0;
  if (!direct) {
    const int st = capture_cycle(c, G, key, pc, mixed, iters);
    if (st) return st;
  }
  auto launch = [&]() -> int {
    if (direct) {
      enqueue_cycle(c, pc, mixed, iters);
      COVOF_CHECK_LAUNCH();
    } else {
      COVOF_CHECK(hipGraphLaunch(G.exec, c.stream));
    }
    return COVOF_OK;
  };
  int launched = 0, st;
  if ((st = launch())) return st;
  COVOF_CHECK(hipEventRecord(P.cycle_ev[0], c.stream));
  ++launched;
  for (int cyc = 0;; ++cyc) {
    if (launched < max_cycles) {
      if ((st = launch())) return st;
      COVOF_CHECK(hipEventRecord(P.cycle_ev[launched & 1], c.stream));
      ++launched;
    }
    COVOF_CHECK(hipEventSynchronize(P.cycle_ev[cyc & 1]));
    if (P.hscal[S_FLAG] != 0.0 || cyc + 1 >= launched) break;
  }
  COVOF_CHECK(hipStreamSynchronize(c.stream));
  prof_harvest(c);
  return COVOF_OK;
}

static int set_scalars(Ctx &c, std::initializer_list<std::pair<int, double>> kv) {
  for (const auto &p : kv) {
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, c.ps.scal, p.first, 1, p.second);
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// PCG iterations per captured FP64 cycle (a multigrid iteration costs many
// launches and the solve takes few of them: shorter cycles)
static int cycle64(int pc) { return pc == 4 ? 8 : (pc == 3 ? 32 : 64); }

// FP64 PCG of the SP loop on the context vectors P.b -> P.x.  NONE / JACOBI
// run as captured cycles of fused kernels; SSOR (the sequential SGS
// wavefront) runs launch by launch with the same kernels.
static int solve_pcg64(Ctx &c, int precond, double tol, int maxit, int *iters, double *relres,
                       bool fp32_cycle = false) {
  Poisson &P = c.ps;
  // the multigrid cycle in FP32 (its launches read the flag while this solve
  // enqueues or captures them)
  struct CycleFlag {
    int &f;
    CycleFlag(int &f_, bool on) : f(f_) { f = on ? 1 : 0; }
    ~CycleFlag() { f = 0; }
  } cycle_flag(P.mg_fp32_cycle, fp32_cycle && pc_code(precond) == 4);
  const Mesh &m = c.m;
  const long n = (long)m.nx * m.ny;
  const DevOp o = dev_op(m);
  const dim3 g = red_grid(m);
  OpT<double> A{P.kx, P.ky, P.diag};
  int st;
  if ((st = fill<double>(c, P.p, n, 0.0))) return st;
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, P.scal, 0, NREAD, 0.0);
  if ((st = set_scalars(c, {{S_TOL, tol}, {S_FIRST, 1.0}, {S_MAXIT, (double)maxit}}))) return st;
  const int pc = pc_code(precond);
  if (pc >= 3 && (st = mg_prepare(c, P.mg_fp32_cycle != 0))) return st;
  double *pb = part_buf(c, 1);
  if (pc == 1)
    hipLaunchKernelGGL((k_residual<1>), g, dim3(BLK), 0, c.stream, o, A, (const double *)P.b, (const double *)P.x,
                       P.r, P.z, pb, (const double *)P.scal);
  else
    hipLaunchKernelGGL((k_residual<0>), g, dim3(BLK), 0, c.stream, o, A, (const double *)P.b, (const double *)P.x,
                       P.r, P.z, pb, (const double *)P.scal);
  COVOF_CHECK_LAUNCH();
  // z = M^-1 r for the explicit preconditioners, then r.z into slot 1 of the
  // residual partials (reference-order mode: r.r, r.z and b.b as serial sums)
  if (pc == 2) {
    if ((st = sgs_apply(c, (const double *)P.r, P.z, nullptr))) return st;
  } else if (pc >= 3) {
    explicit_precond<double>(c, pc, (const double *)P.r, P.z, OUTER.flag);
  }
  if (c.ref_order) {
    serial_sums(c, n, rz_terms(c, pc, P.r, P.z, P.diag, pb, true), S_FLAG, S_FLAG);
  } else if (pc >= 2) {
    hipLaunchKernelGGL(k_pcg_rz<double>, g, dim3(BLK), 0, c.stream, o, (const double *)P.r, (const double *)P.z,
                       (const double *)P.scal, OUTER, pb);
  }
  COVOF_CHECK_LAUNCH();
  if (pc == 2) {  // the sequential SGS: launch by launch
    for (int k = 1;; ++k) {
      pcg_iteration<double, 2>(c, A, P.x, P.r, P.z, P.p, P.t0, P.q, OUTER, k, prof_sampled(c, k));
      COVOF_CHECK_LAUNCH();
      if (k % 16 == 0 || k > maxit) {
        if ((st = read_scalars(c, NREAD))) return st;
        if (P.hscal[S_FLAG] != 0.0) break;
      }
    }
  } else {
    const int cyc = cycle64(pc);
    const int key = 1 + pc + (c.ref_order ? 64 : 0) + (P.mg_fp32_cycle ? 128 : 0);
    if ((st = run_cycles(c, P.graph[0], key, pc, false, cyc, maxit / cyc + 2))) return st;
  }
  flush_x<double>(c, P.x, P.p, P.t0, OUTER, S_ZERO);
  COVOF_CHECK_LAUNCH();
  if ((st = read_scalars(c, NREAD))) return st;
  const int it = (int)P.hscal[S_ITERS];
  if (iters) *iters = it;
  if (relres) *relres = P.hscal[S_RELRES];
  return status_of_flag(P.hscal[S_FLAG]);
}

// Mixed precision on P.b -> P.x: Jacobi (or unpreconditioned) CG whose Krylov
// recurrence -- the correction x32, r32, z32, d, Ad and the scalars -- runs on
// the FP32 operator copy with FP32 wave sums (FP64 across waves and blocks),
// with a reliable update every `cycle` iterations: the FP32 correction is
// folded into the FP64 x, the true residual r = b - A x is recomputed in FP64,
// the FP32 residual restarts from it while the search direction is kept (beta
// is formed with the refreshed residual), and the FP64 ||r||/||b|| <= tol test
// ends the solve.  Method: reliable updates (Sleijpen & van der Vorst, 1996) as
// used for mixed-precision CG by Clark et al., Comput. Phys. Commun. 181
// (2010) 1517.  When the FP64 residual stops improving (no 10 % gain over
// a window of reliable updates) before tol, FP64 PCG continues from the
// refined x; the iteration count is then the sum of both.
int cg2_reserve(Ctx &c);

static int true_relres(Ctx &c, double *relres);

static int solve_mixed(Ctx &c, int precond, double tol, int maxit, int *iters, double *relres) {
  Poisson &P = c.ps;
  const long n = (long)c.m.nx * c.m.ny;
  const int pc = pc_code(precond);
  if (pc == 4 && c.opts.mixed_mg == 0) {
    // FP64 CG whose preconditioner, the V-cycle, runs in FP32: the FP64
    // recurrence keeps FP64 MG-PCG's iteration count and attainable accuracy,
    // the cycle's passes move half the bytes
    P.last_it32 = 0;
    P.last_it64 = 0;
    int it = 0;
    const int st = solve_pcg64(c, precond, tol, maxit, &it, relres, true);
    P.last_it64 = it;
    if (iters) *iters = it;
    if (st == COVOF_HIP_ERROR || st == COVOF_BAD_ARGUMENT) return st;
    const int st2 = true_relres(c, relres);
    return st2 ? st2 : st;
  }
  int cycle = c.opts.mixed_cycle > 0 ? c.opts.mixed_cycle : (pc == 4 ? 8 : (pc == 3 ? 32 : 128));
  if (cycle < 2) cycle = 2;
  cycle &= ~1;  // even: the parity pattern of k repeats from one replay to the next
  int st;
  if (pc >= 3 && (st = mg_prepare(c, true))) return st;
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, P.scal, 0, NREAD, 0.0);
  if ((st = set_scalars(c, {{S_TOL, tol}, {I_FIRST, 1.0}, {I_MAXIT, (double)maxit}, {S_MAXIT, (double)maxit}})))
    return st;
  if ((st = fill<float>(c, P.x32, n, 0.0f))) return st;
  if ((st = fill<float>(c, P.p32, n, 0.0f))) return st;
  if ((st = fill<float>(c, P.q32, n, 0.0f))) return st;
  if (mixed_cg2(c, pc)) {
    if ((st = cg2_reserve(c))) return st;
    const Cg2Bufs B = cg2_bufs(c);
    if ((st = fill<float>(c, B.p[0], n, 0.0f))) return st;
    if ((st = fill<float>(c, B.s[0], n, 0.0f))) return st;
  }
  P.last_it32 = 0;
  P.last_it64 = 0;
  mixed_refresh(c, pc, P.b, P.x, 1);
  COVOF_CHECK_LAUNCH();
  if ((st = read_scalars(c, NREAD))) return st;
  if (P.hscal[S_FLAG] == 0.0) {
    const int key = 1000 + cycle * 16 + pc * 2 + (mixed_cg2(c, pc) ? 1 : 0);
    if ((st = run_cycles(c, P.graph[1], key, pc, true, cycle, maxit / cycle + 2))) return st;
  }
  if ((st = read_scalars(c, NREAD))) return st;
  const int it32 = (int)P.hscal[I_ITERS];
  P.last_it32 = it32;
  if (iters) *iters = it32;
  if (relres) *relres = P.hscal[S_RELRES];
  const double f = P.hscal[S_FLAG];
  if (f == F_STALL && it32 < maxit) {
    // the FP32 recurrence stagnated above tol: FP64 PCG from the refined x
    // (which holds every correction so far) for the remaining iterations
    int it64 = 0;
    double rec = 0.0;
    st = solve_pcg64(c, precond, tol, maxit - it32, &it64, &rec);
    P.last_it64 = it64;
    if (iters) *iters = it32 + it64;
    if (st == COVOF_HIP_ERROR || st == COVOF_BAD_ARGUMENT) return st;
    // the mixed mode reports the true FP64 residual of the x it returns; the
    // status is the continuation's, which stops on its recurrence residual
    // like every FP64 PCG solve (where the system's attainable accuracy is
    // above tol -- 4096 x 2048 bubbles at density ratio 1000: FP64 MG-PCG
    // alone ends at a true 1.7e-9 when its recurrence reads 9e-11 -- the
    // true residual can exceed tol)
    const int st2 = true_relres(c, relres);
    return st2 ? st2 : st;
  }
  if (f == F_CONV) return COVOF_OK;
  if (f == F_NONFINITE) return COVOF_NONFINITE;
  return COVOF_NOT_CONVERGED;
}

// ||b - A x|| / ||b|| of the context vectors P.b, P.x in FP64, recomputed.
static int true_relres(Ctx &c, double *relres) {
  Poisson &P = c.ps;
  OpT<double> A{P.kx, P.ky, P.diag};
  double *pb = part_buf(c, 1);
  hipLaunchKernelGGL(k_true_res, red_grid(c.m), dim3(BLK), 0, c.stream, dev_op(c.m), A, (const double *)P.b,
                     (const double *)P.x, pb);
  hipLaunchKernelGGL(k_true_res_finish, dim3(1), dim3(BLK), 0, c.stream, (const double *)pb, P.nblocks,
                     P.scal + 58);
  COVOF_CHECK_LAUNCH();
  COVOF_CHECK(hipMemcpyAsync(P.hscal + 58, P.scal + 58, sizeof(double), hipMemcpyDeviceToHost, c.stream));
  COVOF_CHECK(hipStreamSynchronize(c.stream));
  if (relres) *relres = P.hscal[58];
  return COVOF_OK;
}

int poisson_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol,
                  int maxit, int precision, int *iters, double *relres) {
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  const int pc = pc_code(precond);
  if (pc < 0) return COVOF_BAD_ARGUMENT;
  if (pc >= 3 && mg_check(c, mg_kind(pc))) return COVOF_BAD_ARGUMENT;
  if (maxit < 0 || !(tol >= 0.0)) return COVOF_BAD_ARGUMENT;
  if (lin_solver == COVOF_LINEAR_CG) {
    if (precision == COVOF_PRECISION_MIXED && precond == COVOF_PRECOND_SSOR) return COVOF_BAD_ARGUMENT;
    if (precision != COVOF_PRECISION_MIXED && precision != COVOF_PRECISION_FP64) return COVOF_BAD_ARGUMENT;
    // the captured cycles work on the context vectors P.b, P.x
    const long n = (long)c.m.nx * c.m.ny;
    Poisson &P = c.ps;
    if (b != P.b) COVOF_CHECK(hipMemcpyAsync(P.b, b, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    if (x != P.x) COVOF_CHECK(hipMemcpyAsync(P.x, x, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    const int st = precision == COVOF_PRECISION_MIXED ? solve_mixed(c, precond, tol, maxit, iters, relres)
                                                      : solve_pcg64(c, precond, tol, maxit, iters, relres);
    if (st == COVOF_HIP_ERROR || st == COVOF_BAD_ARGUMENT) return st;
    if (x != P.x) COVOF_CHECK(hipMemcpyAsync(x, P.x, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return st;
  }
  if (lin_solver == COVOF_LINEAR_BICGSTAB2) {
    extern int solve_bicgstab2(Ctx &, const double *, double *, int, double, int, int *, double *);
    if (precision != COVOF_PRECISION_FP64) return COVOF_BAD_ARGUMENT;
    return solve_bicgstab2(c, b, x, precond, tol, maxit, iters, relres);
  }
  return COVOF_BAD_ARGUMENT;
}

int cg2_reserve(Ctx &c) {
  if (!c.ps.cg2) COVOF_CHECK(hipMalloc(&c.ps.cg2, 8L * c.m.nx * c.m.ny * sizeof(float)));
  return COVOF_OK;
}

void solver_invalidate(Ctx &c) {
  for (PGraph &G : c.ps.graph) {
    if (G.exec) hipGraphExecDestroy(G.exec);
    G.exec = nullptr;
    G.key = -1;
  }
  c.ps.mg64.built = -1;
  c.ps.mg32.built = -1;
}

int sp_poisson(Ctx &c, const double *u, const double *v, const double *rho_u, const double *rho_v,
               double *p, double dt, int lin_solver, int precond, double tol, int maxit,
               int precision, int *iters, double *relres) {
  Poisson &P = c.ps;
  const Mesh &m = c.m;
  int st = poisson_assemble(c, u, v, rho_u, rho_v, dt, P.b);
  if (st) return st;
  const dim3 blk(64, 4), grd((m.nx + 63) / 64, (m.ny + 3) / 4);
  hipLaunchKernelGGL(k_gather_interior, grd, blk, 0, c.stream, m.nx, m.ny, p, P.x);
  COVOF_CHECK_LAUNCH();
  const int solve_st = poisson_solve(c, P.b, P.x, lin_solver, precond, tol, maxit, precision, iters, relres);
  if (solve_st == COVOF_BAD_ARGUMENT || solve_st == COVOF_HIP_ERROR) return solve_st;
  const int pst = sp_pressure(c, P.x, p);
  if (pst) return pst;
  return solve_st;
}

// p(1:nx,1:ny) = x - mean(x) and the periodic halo wrap (the end of the
// Poisson block, src/module_simulation.f90:2577-2588).  x: (nx, ny).
int sp_pressure(Ctx &c, const double *x, double *p) {
  Poisson &P = c.ps;
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 63) / 64, (m.ny + 3) / 4);
  const DevOp o = dev_op(m);
  if (c.ref_order) {  // sum(p_sol) in the reference's order
    SerialSet s{};
    s.nq = 1;
    s.nzero = 1;
    s.a[0] = x;
    s.dst[0] = P.scal + S_DOT0;
    serial_sums(c, (long)m.nx * m.ny, s, S_ZERO, S_ZERO);
  } else {
    hipLaunchKernelGGL(k_sum, red_grid(m), dim3(BLK), 0, c.stream, o, x, part_buf(c, 2));
    hipLaunchKernelGGL(k_sum_finish, dim3(1), dim3(BLK), 0, c.stream, (const double *)part_buf(c, 2), P.nblocks,
                       P.scal, (int)S_DOT0);
  }
  hipLaunchKernelGGL(k_write_pressure, grd, blk, 0, c.stream, o, x, P.scal, S_DOT0, p);
  hipLaunchKernelGGL(k_wrap_halo, dim3(1), dim3(1024), 0, c.stream, m.nx, m.ny, m.perx, m.pery, p);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int sp_correct(Ctx &c, const double *p, const double *rho_u, const double *rho_v, double dt,
               double *u, double *v) {
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 1 + 63) / 64, (m.ny + 1 + 3) / 4);
  const int bl = !m.perx, br = !m.perx, bb = !m.pery, bt = !m.pery;
  hipLaunchKernelGGL(k_correct, grd, blk, 0, c.stream, m.nx, m.ny, m.dx, m.dy, dt, bl, br, bb, bt,
                     is_wall_type(m.bc[0]), is_wall_type(m.bc[1]), is_wall_type(m.bc[2]),
                     is_wall_type(m.bc[3]), m.bc[0] == BC_PERIODIC, m.bc[1] == BC_PERIODIC,
                     m.bc[2] == BC_PERIODIC, m.bc[3] == BC_PERIODIC, dev_obs(c), p, rho_u, rho_v, u, v);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
