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

#include <algorithm>
#include <cmath>

namespace covof {

// ---------------------------------------------------------------- scalars
enum Slot {
  S_BNORM = 0, S_RZ = 1, S_ALPHA = 2, S_BETA = 3, S_RR = 4, S_DAD = 5, S_RELRES = 6,
  S_FLAG = 7, S_ITERS = 8, S_TOL = 9, S_DOT0 = 10,  // 10..13 generic dot outputs
  // inner (FP32) solve of the mixed mode
  I_RZ = 16, I_ALPHA = 17, I_BETA = 18, I_RR0 = 19, I_FLAG = 20, I_ITERS = 21, I_ETA = 22,
  I_RR = 23, I_DAD = 24,
  // BiCGSTAB(2)
  B_RHO0 = 32, B_RHO1, B_ALPHA, B_OMEGA, B_BETA, B_GAMMA, B_R0N, B_G1, B_G2, B_RN,
  B_S11, B_S12, B_S22, B_RHS1, B_RHS2,
  NSLOT = 64
};
enum Flag { F_RUN = 0, F_CONV = 1, F_BREAK = 2, F_NONFINITE = 3 };

struct DevOp {
  int nx, ny;
  int perW, perE, perS, perN;
};

template <typename T>
struct OpT {
  const T *kx, *ky, *diag;
};

__device__ __forceinline__ long kxi(int nx, int i, int j) { return (long)(j - 1) * (nx + 1) + (i - 1); }
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

// Jacobi z = r / d if |d| > 1e-15 else r (src/module_linear_solvers.f90:94-101).
template <typename T>
__device__ __forceinline__ T jacobi(T r, T d) {
  return (fabs((double)d) > 1.0e-15) ? r / d : r;
}

// ---------------------------------------------------- reductions (deterministic)
constexpr int BLK = 256;  // threads per block: tile 64 (x) by 4 (y)

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Sums NV per-thread values over the block (waves in T, across waves in
// double) and the grid; returns true in thread 0 of the last block to
// finish, which then finds the grid totals in tot[].
template <int NV, typename T>
__device__ bool grid_reduce(const T (&v)[NV], double *part, unsigned *ticket, double (&tot)[NV]) {
  __shared__ double sh[NV][BLK / 64];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int nblk = gridDim.x * gridDim.y;
  const int bid = blockIdx.y * gridDim.x + blockIdx.x;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const T w = wave_sum<T>(v[q]);
    if (lane == 0) sh[q][wid] = (double)w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[q][w];
      part[(long)bid * NV + q] = s;
    }
    __threadfence();
    const unsigned t = atomicAdd(ticket, 1u);
    last = (t == (unsigned)(nblk - 1));
  }
  __syncthreads();
  if (!last) return false;
  __threadfence();
  // last block: fixed-order sum of the partials
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblk; b += BLK) s += __builtin_nontemporal_load(&part[(long)b * NV + q]);
    s = wave_sum<double>(s);
    if (lane == 0) sh[q][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[q][w];
      tot[q] = s;
    }
    *ticket = 0u;
  }
  return threadIdx.x == 0;
}

// Cell loop over the block's tile column and its rows (grid-stride in y).
#define CELL_LOOP(o)                                                                   \
  const int i = blockIdx.x * 64 + (threadIdx.x & 63) + 1;                              \
  for (int j = blockIdx.y * 4 + (threadIdx.x >> 6) + 1; j <= (o).ny; j += gridDim.y * 4) \
    if (i <= (o).nx)

// ---------------------------------------------------------------- assembly

// src/module_simulation.f90:2483-2533 on faces: one thread per cell writes
// b, kx(i,j) (and kx(nx+1,j) on the last column), ky(i,j) (and ky(i,ny+1)),
// then diag from the four face values in the reference's summation order.
__global__ void k_assemble(const double *__restrict__ u, const double *__restrict__ v,
                           const double *__restrict__ rho_u, const double *__restrict__ rho_v,
                           int nx, int ny, double dx, double dy, double dt, int wallW, int wallE,
                           int wallS, int wallN, double *kx, double *ky, double *diag, double *b,
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
  const double aW = (i == 1 && wallW) ? 0.0 : cx(i);
  const double aE = (i == nx && wallE) ? 0.0 : cx(i + 1);
  const double aS = (j == 1 && wallS) ? 0.0 : cy(j);
  const double aN = (j == ny && wallN) ? 0.0 : cy(j + 1);
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
  diag32[k] = (float)aP;
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

// Symmetric Gauss-Seidel of poisson_ssor (src/module_simulation.f90:2709-2758)
// in one workgroup by anti-diagonals i + j = s: the forward sweep reads only
// west/south values of earlier diagonals, the backward sweep east/north values
// of later ones, and the periodic wraps read exactly the states the
// lexicographic order sees (still 0 forward, the forward value backward), so
// the result equals the sequential recurrence bit-for-bit.
__global__ void __launch_bounds__(1024) k_sgs_wavefront(DevOp o, OpT<double> A, const double *__restrict__ r,
                                                        double *z) {
  const int nx = o.nx, ny = o.ny;
  const long n = (long)nx * ny;
  for (long t = threadIdx.x; t < n; t += blockDim.x) z[t] = 0.0;
  __syncthreads();
  for (int s = 2; s <= nx + ny; ++s) {
    const int ilo = s - ny > 1 ? s - ny : 1, ihi = s - 1 < nx ? s - 1 : nx;
    for (int i = ilo + threadIdx.x; i <= ihi; i += blockDim.x) {
      const int j = s - i;
      const long k = (long)(j - 1) * nx + (i - 1);
      double val = r[k];
      if (i > 1) val = val + A.kx[kxi(nx, i, j)] * z[k - 1];
      else if (o.perW) val = val + A.kx[kxi(nx, 1, j)] * z[k + (nx - 1)];
      if (j > 1) val = val + A.ky[kyi(nx, i, j)] * z[k - nx];
      else if (o.perS) val = val + A.ky[kyi(nx, i, 1)] * z[k + (long)(ny - 1) * nx];
      z[k] = val / A.diag[k];
    }
    __syncthreads();
  }
  for (int s = nx + ny; s >= 2; --s) {
    const int ilo = s - ny > 1 ? s - ny : 1, ihi = s - 1 < nx ? s - 1 : nx;
    for (int i = ihi - threadIdx.x; i >= ilo; i -= blockDim.x) {
      const int j = s - i;
      const long k = (long)(j - 1) * nx + (i - 1);
      double val = A.diag[k] * z[k];
      if (i < nx) val = val + A.kx[kxi(nx, i + 1, j)] * z[k + 1];
      else if (o.perE) val = val + A.kx[kxi(nx, nx + 1, j)] * z[k - (nx - 1)];
      if (j < ny) val = val + A.ky[kyi(nx, i, j + 1)] * z[k + nx];
      else if (o.perN) val = val + A.ky[kyi(nx, i, ny + 1)] * z[k - (long)(ny - 1) * nx];
      z[k] = val / A.diag[k];
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ PCG (fused)
//
// Iteration of the SP PCG loop (src/module_simulation.f90:2553-2575):
//   kernel A: d = z + beta d_prev (d = z on the first iteration, d_prev = 0),
//             Ad = A d, partial d.Ad;  last block: iters += 1, |dAd| < 1e-35
//             -> breakdown, alpha = rz / dAd.
//   kernel B: x += alpha d, r -= alpha Ad, partial r.r and r.z;  last block:
//             ||r||/||b|| <= tol -> converged, |rz_old| < 1e-35 -> breakdown,
//             beta = rz_new / rz_old.
// PC: 0 none (z = r), 1 Jacobi (z = r/diag on the fly), 2 explicit z array.
// d is ping-ponged (kernel A reads neighbours' d_prev while writing d).
struct PcgSlots {
  int rz, alpha, beta, dad, rr, flag, iters, bnorm_or_rr0, tol;
  int relative_to_rr0;  // inner solve: stop on ||r|| <= tol * ||r0||
};

template <typename T, int PC>
__device__ __forceinline__ T zval(const T *r, const T *diag, const T *z, long k) {
  if (PC == 0) return r[k];
  if (PC == 1) return jacobi<T>(r[k], diag[k]);
  return z[k];
}

template <typename T, int PC>
__global__ void __launch_bounds__(BLK) k_pcg_a(DevOp o, OpT<T> A, const T *__restrict__ r,
                                               const T *__restrict__ z, const T *__restrict__ dprev,
                                               T *__restrict__ d, T *__restrict__ Ad, double *scal,
                                               PcgSlots sl, double *part, unsigned *ticket) {
  if (scal[sl.flag] != 0.0) return;
  const T beta = (T)scal[sl.beta];
  T acc[1] = {T(0)};
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    auto dv = [&](long q) { return zval<T, PC>(r, A.diag, z, q) + beta * dprev[q]; };
    const T dk = dv(k);
    const T q = apply_op<T>(o, A, dv, i, j, k);
    d[k] = dk;
    Ad[k] = q;
    acc[0] += dk * q;
  }
  double tot[1];
  if (grid_reduce<1, T>(acc, part, ticket, tot)) {
    scal[sl.iters] += 1.0;
    const double dAd = tot[0];
    scal[sl.dad] = dAd;
    if (!(fabs(dAd) >= 1.0e-35)) {
      scal[sl.flag] = (dAd == dAd) ? (double)F_BREAK : (double)F_NONFINITE;
    } else {
      scal[sl.alpha] = scal[sl.rz] / dAd;
    }
  }
}

template <typename T, int PC>
__global__ void __launch_bounds__(BLK) k_pcg_b(DevOp o, const T *__restrict__ diag, T *__restrict__ x,
                                               T *__restrict__ r, const T *__restrict__ d,
                                               const T *__restrict__ Ad, double *scal, PcgSlots sl,
                                               double *part, unsigned *ticket) {
  if (scal[sl.flag] != 0.0) return;
  const T alpha = (T)scal[sl.alpha];
  T acc[2] = {T(0), T(0)};
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    x[k] = x[k] + alpha * d[k];
    const T rk = r[k] - alpha * Ad[k];
    r[k] = rk;
    acc[0] += rk * rk;
    if (PC == 0) acc[1] += rk * rk;
    if (PC == 1) acc[1] += rk * jacobi<T>(rk, diag[k]);
  }
  double tot[2];
  if (grid_reduce<2, T>(acc, part, ticket, tot)) {
    const double rn = sqrt(tot[0]);
    scal[sl.rr] = tot[0];
    const double ref = sl.relative_to_rr0 ? sqrt(scal[sl.bnorm_or_rr0]) : scal[sl.bnorm_or_rr0];
    if (!(rn == rn) || rn > 1.0e300) {
      scal[sl.flag] = (double)F_NONFINITE;
    } else if (rn / ref <= scal[sl.tol]) {
      scal[sl.flag] = (double)F_CONV;
    } else if (PC != 2) {
      if (fabs(scal[sl.rz]) < 1.0e-35) {
        scal[sl.flag] = (double)F_BREAK;
      } else {
        scal[sl.beta] = tot[1] / scal[sl.rz];
        scal[sl.rz] = tot[1];
      }
    }
  }
}

// r.z with an explicit z (SGS preconditioner) and the beta update of kernel B.
__global__ void __launch_bounds__(BLK) k_pcg_rz(DevOp o, const double *__restrict__ r, const double *__restrict__ z,
                                                double *scal, PcgSlots sl, double *part, unsigned *ticket) {
  if (scal[sl.flag] != 0.0) return;
  double acc[1] = {0.0};
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    acc[0] += r[k] * z[k];
  }
  double tot[1];
  if (grid_reduce<1, double>(acc, part, ticket, tot)) {
    if (fabs(scal[sl.rz]) < 1.0e-35) {
      scal[sl.flag] = (double)F_BREAK;
    } else {
      scal[sl.beta] = tot[0] / scal[sl.rz];
      scal[sl.rz] = tot[0];
    }
  }
}

// Initial residual of the SP loop (:2538-2549): r = b - A x, ||b||, ||r||,
// converged if ||r||/||b|| <= tol; r.z for PC 0/1.  r32 (optional) receives
// the FP32 copy of the residual for the inner solve of the mixed mode.
template <int PC>
__global__ void __launch_bounds__(BLK) k_residual(DevOp o, OpT<double> A, const double *__restrict__ b,
                                                  const double *__restrict__ x, double *__restrict__ r,
                                                  float *__restrict__ r32, double *scal, PcgSlots sl,
                                                  int want_bnorm, double *part, unsigned *ticket) {
  double acc[3] = {0.0, 0.0, 0.0};
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    const double ax = apply_op<double>(o, A, [&](long q) { return x[q]; }, i, j, k);
    const double rk = b[k] - ax;
    r[k] = rk;
    if (r32) r32[k] = (float)rk;
    acc[0] += b[k] * b[k];
    acc[1] += rk * rk;
    if (PC == 0) acc[2] += rk * rk;
    if (PC == 1) acc[2] += rk * jacobi<double>(rk, A.diag[k]);
  }
  double tot[3];
  if (grid_reduce<3, double>(acc, part, ticket, tot)) {
    if (want_bnorm) {
      double bn = sqrt(tot[0]);
      if (bn < 1.0e-30) bn = 1.0;
      scal[S_BNORM] = bn;
    }
    const double rn = sqrt(tot[1]);
    scal[sl.rr] = tot[1];
    scal[S_RELRES] = rn / scal[S_BNORM];
    if (!(rn == rn) || rn > 1.0e300) scal[sl.flag] = (double)F_NONFINITE;
    else if (rn / scal[S_BNORM] <= scal[sl.tol]) scal[sl.flag] = (double)F_CONV;
    if (PC != 2) scal[sl.rz] = tot[2];
    scal[sl.beta] = 0.0;
  }
}

// Refinement update x += (double) dx32.
__global__ void k_add_correction(long n, double *__restrict__ x, const float *__restrict__ dx32) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    x[t] = x[t] + (double)dx32[t];
}

// FP32 inner-solve setup: r.z and r.r of r32 (z = Jacobi or identity), beta = 0.
template <int PC>
__global__ void __launch_bounds__(BLK) k_inner_init(DevOp o, const float *__restrict__ r32, const float *__restrict__ d32,
                                                    double *scal, PcgSlots sl, double *part, unsigned *ticket) {
  if (scal[S_FLAG] != 0.0) return;
  float acc[2] = {0.f, 0.f};
  CELL_LOOP(o) {
    const long k = (long)(j - 1) * o.nx + (i - 1);
    const float rk = r32[k];
    acc[0] += rk * rk;
    acc[1] += rk * zval<float, PC>(r32, d32, (const float *)nullptr, k);
  }
  double tot[2];
  if (grid_reduce<2, float>(acc, part, ticket, tot)) {
    scal[sl.bnorm_or_rr0] = tot[0];
    scal[sl.rz] = tot[1];
    scal[sl.beta] = 0.0;
    scal[sl.flag] = 0.0;
    if (!(tot[0] > 0.0)) scal[sl.flag] = (double)F_CONV;
  }
}

// p = x - mean(x) (src/module_simulation.f90:2578-2585) into p(1:nx,1:ny).
__global__ void __launch_bounds__(BLK) k_sum(DevOp o, const double *__restrict__ x, double *scal, int slot,
                                             double *part, unsigned *ticket) {
  double acc[1] = {0.0};
  CELL_LOOP(o) { acc[0] += x[(long)(j - 1) * o.nx + (i - 1)]; }
  double tot[1];
  if (grid_reduce<1, double>(acc, part, ticket, tot)) scal[slot] = tot[0];
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

// Velocity correction (src/module_simulation.f90:2590-2654), no obstacles.
__global__ void k_correct(int nx, int ny, double dx, double dy, double dt, int bl, int br, int bb,
                          int bt, int wallL, int wallR, int wallB, int wallT, int perL, int perR,
                          int perB, int perT, const double *__restrict__ p, const double *__restrict__ rho_u,
                          const double *__restrict__ rho_v, double *u, double *v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  auto H = [&](int a, int c) { return (long)c * (nx + 2) + a; };
  const int u_nx = br ? nx + 1 : nx, v_ny = bt ? ny + 1 : ny;
  if (i <= u_nx && j <= ny) {
    const bool wall = (i == 1 && bl && wallL) || (i == u_nx && br && wallR);
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
    const bool wall = (j == 1 && bb && wallB) || (j == v_ny && bt && wallT);
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
  return o;
}

static dim3 red_grid(const Mesh &m) {
  const int gx = (m.nx + 63) / 64;
  int gy = (m.ny + 3) / 4;
  const int want = 2048;
  int cap = want / gx;
  if (cap < 1) cap = 1;
  if (gy > cap) gy = cap;
  return dim3(gx, gy);
}

int poisson_alloc(Ctx &c) {
  Poisson &P = c.ps;
  const long n = (long)c.m.nx * c.m.ny;
  const long nkx = (long)(c.m.nx + 1) * c.m.ny, nky = (long)c.m.nx * (c.m.ny + 1);
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
  float **d32[] = {&P.r32, &P.p32, &P.q32, &P.x32};
  for (float **pp : d32) COVOF_CHECK(hipMalloc(pp, n * sizeof(float)));
  // z32 doubles as the second ping-pong buffer of p32: allocate 2n for p32
  hipFree(P.p32);
  COVOF_CHECK(hipMalloc(&P.p32, 2 * n * sizeof(float)));
  const dim3 g = red_grid(c.m);
  P.nblocks = g.x * g.y;
  // partials: up to 4 values per block of the fused kernels, 5 per block of
  // the 512-block inner-product kernel of BiCGSTAB(2)
  const long npart = std::max<long>((long)P.nblocks * 4, 512L * 5);
  COVOF_CHECK(hipMalloc(&P.part, npart * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.scal, NSLOT * sizeof(double)));
  COVOF_CHECK(hipMalloc(&P.ticket, sizeof(unsigned)));
  COVOF_CHECK(hipMemset(P.ticket, 0, sizeof(unsigned)));
  COVOF_CHECK(hipMemset(P.scal, 0, NSLOT * sizeof(double)));
  COVOF_CHECK(hipHostMalloc(&P.hscal, NSLOT * sizeof(double), hipHostMallocDefault));
  return COVOF_OK;
}

void poisson_free(Ctx &c) {
  Poisson &P = c.ps;
  void *ptrs[] = {P.kx, P.ky, P.diag, P.kx32, P.ky32, P.diag32, P.b, P.x, P.r, P.p, P.q, P.z,
                  P.t0, P.t1, P.t2, P.t3, P.t4, P.t5, P.r32, P.p32, P.q32, P.x32, P.part, P.scal, P.ticket,
                  P.fx, P.fy};
  for (void *q : ptrs) hipFree(q);
  if (P.hscal) hipHostFree(P.hscal);
}

static int is_wall_type(int t) { return t == BC_DIRICHLET || t == BC_FREESLIP; }

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
                     dt, wW, wE, wS, wN, c.ps.kx, c.ps.ky, c.ps.diag, b, c.ps.kx32, c.ps.ky32,
                     c.ps.diag32);
  COVOF_CHECK_LAUNCH();
  c.ps.assembled = 1;
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

int poisson_precond(Ctx &c, int precond, const double *r, double *z) {
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  const long n = (long)c.m.nx * c.m.ny;
  if (precond == COVOF_PRECOND_SSOR) {
    const DevOp o = dev_op(c.m);
    OpT<double> A{c.ps.kx, c.ps.ky, c.ps.diag};
    hipLaunchKernelGGL(k_sgs_wavefront, dim3(1), dim3(1024), 0, c.stream, o, A, r, z);
  } else if (precond == COVOF_PRECOND_JACOBI || precond == COVOF_PRECOND_NONE) {
    hipLaunchKernelGGL(k_precond_simple, dim3((n + 255) / 256), dim3(256), 0, c.stream, precond, n,
                       c.ps.diag, r, z);
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

static const PcgSlots OUTER = {S_RZ, S_ALPHA, S_BETA, S_DAD, S_RR, S_FLAG, S_ITERS, S_BNORM, S_TOL, 0};
static const PcgSlots INNER = {I_RZ, I_ALPHA, I_BETA, I_DAD, I_RR, I_FLAG, I_ITERS, I_RR0, I_ETA, 1};

// Launches up to `count` fused PCG iterations (T, PC) on (x, r) with the
// ping-pong direction buffers d0/d1; *parity tracks which holds d_prev.
template <typename T, int PC>
static int pcg_iterations(Ctx &c, OpT<T> A, T *x, T *r, const T *z, T *d0, T *d1, T *Ad,
                          const PcgSlots &sl, int count, int *parity) {
  const DevOp o = dev_op(c.m);
  const dim3 g = red_grid(c.m);
  for (int it = 0; it < count; ++it) {
    T *dprev = (*parity == 0) ? d0 : d1;
    T *dnew = (*parity == 0) ? d1 : d0;
    hipLaunchKernelGGL((k_pcg_a<T, PC>), g, dim3(BLK), 0, c.stream, o, A, r, z, dprev, dnew, Ad,
                       c.ps.scal, sl, c.ps.part, c.ps.ticket);
    hipLaunchKernelGGL((k_pcg_b<T, PC>), g, dim3(BLK), 0, c.stream, o, A.diag, x, r, dnew, Ad,
                       c.ps.scal, sl, c.ps.part, c.ps.ticket);
    *parity ^= 1;
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static int status_of_flag(double f, int iters, int maxit) {
  if (f == F_CONV) return COVOF_OK;
  if (f == F_BREAK) return COVOF_BREAKDOWN;
  if (f == F_NONFINITE) return COVOF_NONFINITE;
  return iters >= maxit ? COVOF_NOT_CONVERGED : COVOF_OK;
}

// FP64 PCG of the SP loop with NONE / JACOBI (fused) or SSOR (SGS wavefront).
static int solve_pcg64(Ctx &c, const double *b, double *x, int precond, double tol, int maxit,
                       int *iters, double *relres) {
  Poisson &P = c.ps;
  const Mesh &m = c.m;
  const long n = (long)m.nx * m.ny;
  const DevOp o = dev_op(m);
  const dim3 g = red_grid(m);
  OpT<double> A{P.kx, P.ky, P.diag};
  int st;
  if ((st = fill<double>(c, P.p, n, 0.0))) return st;
  if ((st = fill<double>(c, P.q, n, 0.0))) return st;
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, P.scal, 0, 16, 0.0);
  if ((st = set_scalar(c, S_TOL, tol))) return st;
  const int pc = precond == COVOF_PRECOND_SSOR ? 2 : (precond == COVOF_PRECOND_JACOBI ? 1 : 0);
  if (pc == 2)
    hipLaunchKernelGGL((k_residual<2>), g, dim3(BLK), 0, c.stream, o, A, b, x, P.r, (float *)nullptr,
                       P.scal, OUTER, 1, P.part, P.ticket);
  else if (pc == 1)
    hipLaunchKernelGGL((k_residual<1>), g, dim3(BLK), 0, c.stream, o, A, b, x, P.r, (float *)nullptr,
                       P.scal, OUTER, 1, P.part, P.ticket);
  else
    hipLaunchKernelGGL((k_residual<0>), g, dim3(BLK), 0, c.stream, o, A, b, x, P.r, (float *)nullptr,
                       P.scal, OUTER, 1, P.part, P.ticket);
  COVOF_CHECK_LAUNCH();
  int parity = 0, done = 0;
  if (pc == 2) {
    // z = SGS(r), rz = r.z before the loop; every iteration: A, B, SGS, rz
    if ((st = poisson_precond(c, COVOF_PRECOND_SSOR, P.r, P.z))) return st;
    // rz = r.z via k_pcg_rz (seeded with rz_old = 1 so its breakdown test passes)
    if ((st = set_scalar(c, S_RZ, 1.0))) return st;
    hipLaunchKernelGGL(k_pcg_rz, g, dim3(BLK), 0, c.stream, o, P.r, P.z, P.scal, OUTER, P.part, P.ticket);
    if ((st = set_scalar(c, S_BETA, 0.0))) return st;
    while (done < maxit) {
      const int chunk = std::min(maxit - done, done < 64 ? 16 : 64);
      for (int it = 0; it < chunk; ++it) {
        double *dprev = parity == 0 ? P.p : P.t0;
        double *dnew = parity == 0 ? P.t0 : P.p;
        hipLaunchKernelGGL((k_pcg_a<double, 2>), g, dim3(BLK), 0, c.stream, o, A, P.r, P.z, dprev,
                           dnew, P.q, P.scal, OUTER, P.part, P.ticket);
        hipLaunchKernelGGL((k_pcg_b<double, 2>), g, dim3(BLK), 0, c.stream, o, P.diag, x, P.r, dnew,
                           P.q, P.scal, OUTER, P.part, P.ticket);
        if ((st = poisson_precond(c, COVOF_PRECOND_SSOR, P.r, P.z))) return st;
        hipLaunchKernelGGL(k_pcg_rz, g, dim3(BLK), 0, c.stream, o, P.r, P.z, P.scal, OUTER, P.part, P.ticket);
        parity ^= 1;
      }
      COVOF_CHECK_LAUNCH();
      done += chunk;
      if ((st = read_scalars(c, 16))) return st;
      if (P.hscal[S_FLAG] != 0.0) break;
    }
  } else {
    // d_prev must be 0 on the first iteration (d = z + 0 * d_prev)
    while (done < maxit) {
      const int chunk = std::min(maxit - done, done < 64 ? 16 : (done < 1024 ? 64 : 256));
      if (pc == 1) st = pcg_iterations<double, 1>(c, A, x, P.r, nullptr, P.p, P.t0, P.q, OUTER, chunk, &parity);
      else st = pcg_iterations<double, 0>(c, A, x, P.r, nullptr, P.p, P.t0, P.q, OUTER, chunk, &parity);
      if (st) return st;
      done += chunk;
      if ((st = read_scalars(c, 16))) return st;
      if (P.hscal[S_FLAG] != 0.0) break;
    }
  }
  if ((st = read_scalars(c, 16))) return st;
  const int it = (int)P.hscal[S_ITERS];
  if (iters) *iters = it;
  if (relres) *relres = (P.hscal[S_ITERS] > 0.0) ? sqrt(P.hscal[S_RR]) / P.hscal[S_BNORM] : P.hscal[S_RELRES];
  return status_of_flag(P.hscal[S_FLAG], it, maxit);
}

// Mixed precision: FP64 iterative refinement around an FP32 PCG (NONE/JACOBI).
// Outer: r = b - A x in FP64 (with the previous correction folded into x in
// the same pass); stop on ||r||/||b|| <= tol.  Inner: PCG on A32 dx = r32 from
// dx = 0 to ||r_in|| <= eta ||r_in,0|| (FP32 vectors and operator, FP32 wave
// sums, FP64 across waves and blocks).
static int solve_mixed(Ctx &c, const double *b, double *x, int precond, double tol, int maxit,
                       int *iters, double *relres) {
  Poisson &P = c.ps;
  const Mesh &m = c.m;
  const long n = (long)m.nx * m.ny;
  const DevOp o = dev_op(m);
  const dim3 g = red_grid(m);
  OpT<double> A{P.kx, P.ky, P.diag};
  OpT<float> A32{P.kx32, P.ky32, P.diag32};
  const int pc = precond == COVOF_PRECOND_JACOBI ? 1 : 0;
  const double eta = 1.0e-4;
  const int max_outer = 60;
  int st;
  hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, P.scal, 0, 32, 0.0);
  if ((st = set_scalar(c, S_TOL, tol))) return st;
  if ((st = set_scalar(c, I_ETA, eta))) return st;
  if ((st = fill<float>(c, P.x32, n, 0.0f))) return st;
  float *d0 = P.p32, *d1 = P.p32 + n;
  int total_inner = 0;
  for (int outer = 0; outer < max_outer; ++outer) {
    // x += dx (dx = 0 before the first inner solve), then the FP64 residual
    if (outer > 0)
      hipLaunchKernelGGL(k_add_correction, dim3(1024), dim3(256), 0, c.stream, n, x, P.x32);
    if (pc == 1)
      hipLaunchKernelGGL((k_residual<1>), g, dim3(BLK), 0, c.stream, o, A, b, x, P.r, P.r32, P.scal,
                         OUTER, outer == 0, P.part, P.ticket);
    else
      hipLaunchKernelGGL((k_residual<0>), g, dim3(BLK), 0, c.stream, o, A, b, x, P.r, P.r32, P.scal,
                         OUTER, outer == 0, P.part, P.ticket);
    COVOF_CHECK_LAUNCH();
    if ((st = read_scalars(c, 32))) return st;
    if (P.hscal[S_FLAG] != 0.0 || total_inner >= maxit) break;
    // inner FP32 PCG from dx = 0
    if ((st = fill<float>(c, P.x32, n, 0.0f))) return st;
    if ((st = fill<float>(c, d0, n, 0.0f))) return st;
    hipLaunchKernelGGL(k_set_scalars, dim3(1), dim3(64), 0, c.stream, P.scal, I_ITERS, 1, 0.0);
    if (pc == 1)
      hipLaunchKernelGGL((k_inner_init<1>), g, dim3(BLK), 0, c.stream, o, P.r32, P.diag32, P.scal, INNER, P.part, P.ticket);
    else
      hipLaunchKernelGGL((k_inner_init<0>), g, dim3(BLK), 0, c.stream, o, P.r32, P.diag32, P.scal, INNER, P.part, P.ticket);
    COVOF_CHECK_LAUNCH();
    int parity = 0, done = 0;
    const int inner_cap = maxit - total_inner;
    while (done < inner_cap) {
      const int chunk = std::min(inner_cap - done, done < 64 ? 32 : (done < 1024 ? 128 : 256));
      if (pc == 1) st = pcg_iterations<float, 1>(c, A32, P.x32, P.r32, nullptr, d0, d1, P.q32, INNER, chunk, &parity);
      else st = pcg_iterations<float, 0>(c, A32, P.x32, P.r32, nullptr, d0, d1, P.q32, INNER, chunk, &parity);
      if (st) return st;
      done += chunk;
      if ((st = read_scalars(c, 32))) return st;
      if (P.hscal[I_FLAG] != 0.0) break;
    }
    total_inner += (int)P.hscal[I_ITERS];
    if (P.hscal[I_FLAG] == F_NONFINITE) break;
  }
  if ((st = read_scalars(c, 32))) return st;
  if (iters) *iters = total_inner;
  if (relres) *relres = P.hscal[S_RELRES];
  const double f = P.hscal[S_FLAG];
  if (f == F_CONV) return COVOF_OK;
  if (f == F_NONFINITE || P.hscal[I_FLAG] == F_NONFINITE) return COVOF_NONFINITE;
  return COVOF_NOT_CONVERGED;
}

int poisson_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol,
                  int maxit, int precision, int *iters, double *relres) {
  if (!c.ps.assembled) return COVOF_BAD_ARGUMENT;
  if (precond != COVOF_PRECOND_NONE && precond != COVOF_PRECOND_JACOBI && precond != COVOF_PRECOND_SSOR)
    return COVOF_BAD_ARGUMENT;
  if (maxit < 0 || !(tol >= 0.0)) return COVOF_BAD_ARGUMENT;
  if (lin_solver == COVOF_LINEAR_CG) {
    if (precision == COVOF_PRECISION_MIXED) {
      if (precond == COVOF_PRECOND_SSOR) return COVOF_BAD_ARGUMENT;
      return solve_mixed(c, b, x, precond, tol, maxit, iters, relres);
    }
    if (precision != COVOF_PRECISION_FP64) return COVOF_BAD_ARGUMENT;
    return solve_pcg64(c, b, x, precond, tol, maxit, iters, relres);
  }
  if (lin_solver == COVOF_LINEAR_BICGSTAB2) {
    extern int solve_bicgstab2(Ctx &, const double *, double *, int, double, int, int *, double *);
    if (precision != COVOF_PRECISION_FP64) return COVOF_BAD_ARGUMENT;
    return solve_bicgstab2(c, b, x, precond, tol, maxit, iters, relres);
  }
  return COVOF_BAD_ARGUMENT;
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
  const DevOp o = dev_op(m);
  hipLaunchKernelGGL(k_sum, red_grid(m), dim3(BLK), 0, c.stream, o, P.x, P.scal, S_DOT0, P.part, P.ticket);
  hipLaunchKernelGGL(k_write_pressure, grd, blk, 0, c.stream, o, P.x, P.scal, S_DOT0, p);
  hipLaunchKernelGGL(k_wrap_halo, dim3(1), dim3(1024), 0, c.stream, m.nx, m.ny, m.perx, m.pery, p);
  COVOF_CHECK_LAUNCH();
  return solve_st;
}

int sp_correct(Ctx &c, const double *p, const double *rho_u, const double *rho_v, double dt,
               double *u, double *v) {
  const Mesh &m = c.m;
  const dim3 blk(64, 4), grd((m.nx + 1 + 63) / 64, (m.ny + 1 + 3) / 4);
  const int bl = !m.perx, br = !m.perx, bb = !m.pery, bt = !m.pery;
  hipLaunchKernelGGL(k_correct, grd, blk, 0, c.stream, m.nx, m.ny, m.dx, m.dy, dt, bl, br, bb, bt,
                     is_wall_type(m.bc[0]), is_wall_type(m.bc[1]), is_wall_type(m.bc[2]),
                     is_wall_type(m.bc[3]), m.bc[0] == BC_PERIODIC, m.bc[1] == BC_PERIODIC,
                     m.bc[2] == BC_PERIODIC, m.bc[3] == BC_PERIODIC, p, rho_u, rho_v, u, v);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
