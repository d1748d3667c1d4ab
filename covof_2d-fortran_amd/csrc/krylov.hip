// BiCGSTAB(2) (src/module_linear_solvers.f90:361-579) and CG (:735-838) on
// any operator: the pressure operator (BiCGSTAB(2) re-targeted from the
// reference's velocity system) and the SP momentum operator (momentum.hip).
//
// The reference's left-preconditioned BiCGSTAB(2): apply_PA = precond(A x)
// (:323-356).  Vector updates are elementwise kernels that evaluate the
// reference's expressions in the same order; inner products are deterministic
// grid reductions; the scalar recurrences (with every breakdown test of the
// reference) run on the host, which reads the inner products back once per
// dependent scalar.
#include "covof_internal.h"

#include <cmath>

namespace covof {

namespace {

constexpr int BLK = 256;

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// Up to five inner products over at most three distinct vectors: each
// element of every distinct vector is loaded once and all their pairwise
// products are accumulated (BiCGSTAB(2)'s five-product step reads three
// vectors instead of ten).
struct Pairs {
  const double *v[3];  // the distinct vectors
  int nq;              // products wanted
  int ia[5], ib[5];    // operands of product q (indices into v, ia <= ib)
};

// per-block partials of the NV(NV+1)/2 pairwise products, then the last block
// sums the partials in block order and writes out[q]
template <int NV>
__global__ void __launch_bounds__(BLK) k_dots(long n, Pairs P, double *part, unsigned *ticket, double *out) {
  constexpr int NP = NV * (NV + 1) / 2;
  __shared__ double sh[NP][BLK / 64];
  __shared__ double tot[NP];
  __shared__ bool last;
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; ++p) acc[p] = 0.0;
  // the grid-stride loop four strides at a time, the four elements' loads
  // issued together (the same elements in the same order as one stride at a
  // time)
  const long stride = (long)gridDim.x * BLK;
  for (long t0 = (long)blockIdx.x * BLK + threadIdx.x; t0 < n; t0 += 4 * stride) {
    double x[4][NV];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long t = t0 + u * stride;
#pragma unroll
      for (int k = 0; k < NV; ++k) x[u][k] = t < n ? P.v[k][t] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (t0 + u * stride >= n) break;
      int p = 0;
#pragma unroll
      for (int a = 0; a < NV; ++a)
#pragma unroll
        for (int b = a; b < NV; ++b) acc[p++] += x[u][a] * x[u][b];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NP; ++p) {
    const double w = wsum(acc[p]);
    if (lane == 0) sh[p][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int p = 0; p < NP; ++p) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[p][w];
      part[(long)blockIdx.x * 6 + p] = s;
    }
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int p = 0; p < NP; ++p) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BLK) s += __builtin_nontemporal_load(&part[(long)b * 6 + p]);
    s = wsum(s);
    if (lane == 0) sh[p][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int p = 0; p < NP; ++p) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s += sh[p][w];
      tot[p] = s;
    }
    for (int q = 0; q < P.nq; ++q) {
      const int a = P.ia[q], b = P.ib[q];
      out[q] = tot[a * NV - a * (a - 1) / 2 + (b - a)];
    }
    *ticket = 0u;
  }
}

// out = x + a*y
__global__ void k_axpy(long n, double *out, const double *x, double a, const double *y) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    out[t] = x[t] + a * y[t];
}
// out = x - y
__global__ void k_sub(long n, double *out, const double *x, const double *y) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    out[t] = x[t] - y[t];
}
// out = x + a*y + b*z   ((x + a y) + b z)
__global__ void k_axpbz(long n, double *out, const double *x, double a, const double *y, double b,
                        const double *z) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    out[t] = x[t] + a * y[t] + b * z[t];
}
// out = x - (a*y + b*z)
__global__ void k_xmin(long n, double *out, const double *x, double a, const double *y, double b,
                       const double *z) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    out[t] = x[t] - (a * y[t] + b * z[t]);
}
__global__ void k_zero(long n, double *x) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    x[t] = 0.0;
}

}  // namespace

struct Bicg {
  Ctx &c;
  LinOp &op;
  long n;
  double *tmp;
  dim3 g{1024}, b{256};
  int err = 0;

  // apply_PA (src/module_linear_solvers.f90:323-356): dst = M^-1 (A src)
  void PA(const double *src, double *dst) {
    if (!err) err = op.apply_pa(src, dst, tmp);
  }
  // inner products -> host (synchronises); k <= 5 products of at most three
  // distinct vectors
  void dots(int k, const double *const *a, const double *const *bb, double *res) {
    if (err) return;
    Pairs P;
    P.nq = k;
    int nv = 0;
    auto slot = [&](const double *p) {
      for (int i = 0; i < nv; ++i)
        if (P.v[i] == p) return i;
      if (nv == 3) return -1;
      P.v[nv] = p;
      return nv++;
    };
    for (int q = 0; q < k; ++q) {
      int ia = slot(a[q]), ib = slot(bb[q]);
      if (ia < 0 || ib < 0) { err = COVOF_BAD_ARGUMENT; return; }
      P.ia[q] = ia < ib ? ia : ib;
      P.ib[q] = ia < ib ? ib : ia;
    }
    if (c.ref_order) {
      // the reference's dot_product: one serial sum per product, in index order
      SerialSet S{};
      S.nq = k;
      S.nzero = 1;
      for (int q = 0; q < k; ++q) {
        S.a[q] = a[q];
        S.b[q] = bb[q];
        S.dst[q] = c.ps.scal + 40 + q;
      }
      serial_sums(c, n, S, 63, 63);
    } else if (nv == 1)
      hipLaunchKernelGGL(k_dots<1>, dim3(512), dim3(BLK), 0, c.stream, n, P, c.ps.part, c.ps.ticket, c.ps.scal + 40);
    else if (nv == 2)
      hipLaunchKernelGGL(k_dots<2>, dim3(512), dim3(BLK), 0, c.stream, n, P, c.ps.part, c.ps.ticket, c.ps.scal + 40);
    else
      hipLaunchKernelGGL(k_dots<3>, dim3(512), dim3(BLK), 0, c.stream, n, P, c.ps.part, c.ps.ticket, c.ps.scal + 40);
    if (hipGetLastError() != hipSuccess) { err = COVOF_HIP_ERROR; return; }
    if (hipMemcpyAsync(c.ps.hscal, c.ps.scal + 40, k * sizeof(double), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess) { err = COVOF_HIP_ERROR; return; }
    for (int q = 0; q < k; ++q) res[q] = c.ps.hscal[q];
  }
  double dot(const double *x, const double *y) {
    double r = 0.0;
    const double *a[1] = {x}, *bb[1] = {y};
    dots(1, a, bb, &r);
    return r;
  }
  void axpy(double *out, const double *x, double a, const double *y) {
    hipLaunchKernelGGL(k_axpy, g, b, 0, c.stream, n, out, x, a, y);
  }
};

// src/module_linear_solvers.f90:361-579 (iteration counter += 2 per sweep;
// stop on ||r||/||r0|| < tol or ||r|| < 1e-12 of the preconditioned residual).
// w: nine work vectors of op.n entries, none aliasing b or x.
int krylov_bicgstab2(Ctx &c, LinOp &op, const double *bvec, double *x, double *const w[9], double tol,
                     int maxit, int *iters, double *relres) {
  const long n = op.n;
  double *bh = w[0], *rt = w[1], *r0 = w[2], *r1 = w[3], *r2 = w[4], *u0 = w[5], *u1 = w[6], *u2 = w[7];
  Bicg K{c, op, n, w[8]};
  for (int k = 0; k < 9; ++k)
    if (w[k] == x || w[k] == bvec) return COVOF_BAD_ARGUMENT;
  const dim3 g(1024), bl(256);
  for (double *v : {u0, u1, u2, r1, r2}) hipLaunchKernelGGL(k_zero, g, bl, 0, c.stream, n, v);
  K.err = op.precond(bvec, bh);
  K.PA(x, r0);
  hipLaunchKernelGGL(k_sub, g, bl, 0, c.stream, n, r0, bh, r0);
  const double r0n = sqrt(K.dot(r0, r0));
  if (K.err) return K.err;
  int iter = 0;
  double rn = r0n;
  int status = COVOF_NOT_CONVERGED;
  if (r0n < 1.0e-12) {
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    return COVOF_OK;
  }
  hipMemcpyAsync(rt, r0, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
  double rho0 = 1.0, rho1, alpha = 0.0, omega = 1.0, beta, gamma;
  while (iter < maxit) {
    iter += 2;
    rho0 = -omega * rho0;
    rho1 = K.dot(rt, r0);
    if (fabs(rho0) < 1.0e-300) { status = COVOF_BREAKDOWN; break; }
    beta = alpha * (rho1 / rho0);
    rho0 = rho1;
    K.axpy(u0, r0, -beta, u0);
    K.PA(u0, u1);
    gamma = K.dot(rt, u1);
    if (fabs(gamma) < 1.0e-300) { status = COVOF_BREAKDOWN; break; }
    alpha = rho0 / gamma;
    K.axpy(r0, r0, -alpha, u1);
    K.axpy(x, x, alpha, u0);
    K.PA(r0, r1);
    rho1 = K.dot(rt, r1);
    if (fabs(rho0) < 1.0e-300) { status = COVOF_BREAKDOWN; break; }
    beta = alpha * (rho1 / rho0);
    rho0 = rho1;
    K.axpy(u0, r0, -beta, u0);
    K.axpy(u1, r1, -beta, u1);
    K.PA(u1, u2);
    gamma = K.dot(rt, u2);
    if (fabs(gamma) < 1.0e-300) { status = COVOF_BREAKDOWN; break; }
    alpha = rho0 / gamma;
    K.axpy(x, x, alpha, u0);
    K.axpy(r0, r0, -alpha, u1);
    K.axpy(r1, r1, -alpha, u2);
    K.PA(r1, r2);
    double s[5];
    const double *A5[5] = {r1, r2, r2, r0, r0}, *B5[5] = {r1, r1, r2, r1, r2};
    K.dots(5, A5, B5, s);
    const double s11 = s[0], s12 = s[1], s22 = s[2], rhs1 = s[3], rhs2 = s[4];
    const double rn1 = sqrt(s11), rn2 = sqrt(s22);
    double g1, g2;
    if (rn1 < 1.0e-300) {
      status = COVOF_BREAKDOWN;
      break;
    } else if (rn2 < 1.0e-300) {
      g2 = 0.0;
      g1 = rhs1 / s11;
    } else {
      const double i1 = 1.0 / rn1, i2 = 1.0 / rn2;
      const double c12 = s12 * i1 * i2, b1 = rhs1 * i1, b2 = rhs2 * i2;
      const double det = 1.0 - c12 * c12;
      if (fabs(det) < 1.0e-14) {
        g2 = 0.0;
        g1 = b1 * i1;
      } else {
        g1 = (b1 - c12 * b2) / (det * rn1);
        g2 = (b2 - c12 * b1) / (det * rn2);
      }
    }
    hipLaunchKernelGGL(k_axpbz, g, bl, 0, c.stream, n, x, x, g1, r0, g2, r1);
    hipLaunchKernelGGL(k_xmin, g, bl, 0, c.stream, n, r0, r0, g1, r1, g2, r2);
    hipLaunchKernelGGL(k_xmin, g, bl, 0, c.stream, n, u0, u0, g1, u1, g2, u2);
    if (g2 != 0.0) omega = g2;
    else if (fabs(g1) > 1.0e-300) omega = g1;
    else omega = 1.0;
    rn = sqrt(K.dot(r0, r0));
    if (K.err) return K.err;
    if (rn != rn || rn > 1.0e30) { status = COVOF_NONFINITE; break; }
    if (rn / r0n < tol || rn < 1.0e-12) { status = COVOF_OK; break; }
  }
  if (K.err) return K.err;
  if (hipStreamSynchronize(c.stream) != hipSuccess) return COVOF_HIP_ERROR;
  if (iters) *iters = iter;
  if (relres) *relres = rn / r0n;
  return status;
}

// solve_cg (src/module_linear_solvers.f90:735-838): split PCG from x0, stop
// on ||r||/||r0|| < tol or ||r|| < 1e-13; |p.Ap| < 1e-30 ends the loop
// (COVOF_BREAKDOWN).  iters follows the reference: the stopping iteration,
// else maxit.  z is formed before the stop test (its dot rides with r.r), so
// each iteration synchronises twice.  w: four work vectors (r, z, p, Ap).
int krylov_cg(Ctx &c, LinOp &op, const double *bvec, double *x, double *const w[4], double tol, int maxit,
              int *iters, double *relres) {
  const long n = op.n;
  double *r = w[0], *z = w[1], *p = w[2], *Ap = w[3];
  for (int k = 0; k < 4; ++k)
    if (w[k] == x || w[k] == bvec) return COVOF_BAD_ARGUMENT;
  Bicg K{c, op, n, nullptr};
  const dim3 g(1024), bl(256);
  K.err = op.apply(x, Ap);
  hipLaunchKernelGGL(k_sub, g, bl, 0, c.stream, n, r, bvec, Ap);
  const double r0n = sqrt(K.dot(r, r));
  if (K.err) return K.err;
  if (iters) *iters = maxit;
  if (relres) *relres = 1.0;
  if (r0n < 1.0e-12) {
    if (iters) *iters = 0;
    if (relres) *relres = 0.0;
    return COVOF_OK;
  }
  if (!K.err) K.err = op.precond(r, z);
  hipMemcpyAsync(p, z, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream);
  double rz_old = K.dot(r, z);
  int status = COVOF_NOT_CONVERGED;
  for (int it = 1; it <= maxit && !K.err; ++it) {
    K.err = op.apply(p, Ap);
    const double pAp = K.dot(p, Ap);
    if (K.err) break;
    if (fabs(pAp) < 1.0e-30) { status = COVOF_BREAKDOWN; break; }
    const double alpha = rz_old / pAp;
    K.axpy(x, x, alpha, p);
    K.axpy(r, r, -alpha, Ap);
    if (!K.err) K.err = op.precond(r, z);
    double s[2];
    const double *A2[2] = {r, r}, *B2[2] = {r, z};
    K.dots(2, A2, B2, s);
    if (K.err) break;
    const double rn = sqrt(s[0]);
    if (relres) *relres = rn / r0n;
    if (rn != rn) { status = COVOF_NONFINITE; break; }
    if (rn / r0n < tol || rn < 1.0e-13) {
      if (iters) *iters = it;
      status = COVOF_OK;
      break;
    }
    const double beta = s[1] / rz_old;
    rz_old = s[1];
    K.axpy(p, z, beta, p);
  }
  if (K.err) return K.err;
  if (hipStreamSynchronize(c.stream) != hipSuccess) return COVOF_HIP_ERROR;
  return status;
}

namespace {
// The Poisson operator and its preconditioners as a LinOp.
struct PoissonOp : LinOp {
  Ctx &c;
  int pc;
  PoissonOp(Ctx &cc, int precond) : c(cc), pc(precond) { n = (long)cc.m.nx * cc.m.ny; }
  int apply(const double *x, double *y) override { return poisson_matvec(c, x, y); }
  int precond(const double *r, double *z) override { return poisson_precond(c, pc, r, z); }
};
}  // namespace

// BiCGSTAB(2) re-targeted to the pressure operator.
int solve_bicgstab2(Ctx &c, const double *bvec, double *x, int precond, double tol, int maxit,
                    int *iters, double *relres) {
  Poisson &P = c.ps;
  PoissonOp op(c, precond);
  double *bh = P.b == bvec ? P.z : P.b;  // b_hat must not alias b
  double *const w[9] = {bh, P.r, P.t4, P.q, P.t0, P.t1, P.t2, P.t3, P.t5};
  return krylov_bicgstab2(c, op, bvec, x, w, tol, maxit, iters, relres);
}

}  // namespace covof
