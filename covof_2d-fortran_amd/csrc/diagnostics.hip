// Output-side kernels: the per-step history row of vtk_write_diagnostics
// (src/module_vtk_writer.f90:399-531) and the derived cell fields of
// vtk_write_step (:90-394: divergence and cell-centred velocity).
//
// The history row is one pass over the interior cells: seven sums (bubble
// volume and its y and v moments, gas volume, gas mass and pressure
// integrals, kinetic energy) and two maxima (speed, |div u|), each cell's
// terms evaluated with the reference's expressions; per-block partials in
// fixed order, folded by the last block (the sums differ from the
// reference's serial ones only by summation order; the maxima are exact).
#include "covof_internal.h"

#include <cmath>

namespace covof {

namespace {

constexpr int BLK = 256;
constexpr int NS = 7;  // sums
constexpr int NM = 2;  // maxima

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}
__device__ __forceinline__ double wmax(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    const double o = __shfl_down(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}

#define HX(i, j) ((long)(j) * (m.nx + 2) + (i))

// out: yc, rise velocity, bubble volume, kinetic energy, max speed, max
// |div|, mean gas density, mean gas pressure
__global__ void __launch_bounds__(BLK) k_history(Mesh m, const double *__restrict__ C, const double *__restrict__ u,
                                                 const double *__restrict__ v, const double *__restrict__ p,
                                                 const double *__restrict__ rho, const double *__restrict__ T,
                                                 double *part, unsigned *ticket, double *out) {
  __shared__ double sh[NS + NM][BLK / 64];
  __shared__ bool last;
  const long ncell = (long)m.nx * m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  double acc[NS] = {0, 0, 0, 0, 0, 0, 0};
  double mx[NM] = {0, 0};
  for (long t = (long)blockIdx.x * BLK + threadIdx.x; t < ncell; t += (long)gridDim.x * BLK) {
    const int j = (int)(t / m.nx) + 1, i = (int)(t - (long)(j - 1) * m.nx) + 1;
    const double y_pos = ((double)(j - 1) + 0.5) * dy;  // y_center (src/module_mesh.f90:192-196)
    const double c_val = C[HX(i, j)];
    const double ind = 1.0 - c_val;
    const double pc = p[HX(i, j)];
    const double Tc = T ? T[HX(i, j)] : 300.0;
    acc[0] += ind * dvol;                                  // bubble volume
    acc[1] += y_pos * ind * dvol;                          // y moment
    const double rho_gas_c = pc / (287.05 * (Tc > 1.0 ? Tc : 1.0));
    acc[2] += ind * rho_gas_c * dvol;                      // gas mass
    acc[3] += ind * pc * dvol;                             // gas pressure integral
    acc[4] += ind * dvol;                                  // gas volume
    const double uw = u[HX(i, j)], ue = u[HX(i + 1, j)], vs = v[HX(i, j)], vn = v[HX(i, j + 1)];
    const double uc = 0.5 * (uw + ue), vc = 0.5 * (vs + vn);
    acc[5] += vc * ind * dvol;                             // v moment
    const double speed2 = uc * uc + vc * vc;
    const double u2p = 0.5 * (uw * uw + ue * ue), v2p = 0.5 * (vs * vs + vn * vn);
    acc[6] += 0.5 * rho[HX(i, j)] * (u2p + v2p) * dvol;    // kinetic energy
    const double speed = sqrt(speed2);
    if (speed > mx[0]) mx[0] = speed;
    const double dv = fabs((ue - uw) / dx + (vn - vs) / dy);
    if (dv > mx[1]) mx[1] = dv;
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NS; ++q) {
    const double w = wsum(acc[q]);
    if (lane == 0) sh[q][wid] = w;
  }
#pragma unroll
  for (int q = 0; q < NM; ++q) {
    const double w = wmax(mx[q]);
    if (lane == 0) sh[NS + q][wid] = w;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 0; q < NS + NM; ++q) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s = q < NS ? s + sh[q][w] : (sh[q][w] > s ? sh[q][w] : s);
      part[(long)blockIdx.x * (NS + NM) + q] = s;
    }
    __threadfence();
    last = atomicAdd(ticket, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  for (int q = 0; q < NS + NM; ++q) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += BLK) {
      const double x = __builtin_nontemporal_load(&part[(long)b * (NS + NM) + q]);
      s = q < NS ? s + x : (x > s ? x : s);
    }
    s = q < NS ? wsum(s) : wmax(s);
    if (lane == 0) sh[q][wid] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot[NS + NM];
    for (int q = 0; q < NS + NM; ++q) {
      double s = 0.0;
      for (int w = 0; w < BLK / 64; ++w) s = q < NS ? s + sh[q][w] : (sh[q][w] > s ? sh[q][w] : s);
      tot[q] = s;
    }
    const double vol = tot[0], gvol = tot[4];
    out[0] = vol > 1.0e-12 ? tot[1] / vol : 0.0;
    out[1] = vol > 1.0e-12 ? tot[5] / vol : 0.0;
    out[2] = vol;
    out[3] = tot[6];
    out[4] = tot[NS];
    out[5] = tot[NS + 1];
    out[6] = gvol > 1.0e-12 ? tot[2] / gvol : 0.0;
    out[7] = gvol > 1.0e-12 ? tot[3] / gvol : 0.0;
    *ticket = 0u;
  }
}

// divergence and cell-centred velocity of vtk_write_step, (nx, ny) arrays
__global__ void __launch_bounds__(BLK) k_cell_output(Mesh m, const double *__restrict__ u,
                                                     const double *__restrict__ v, double *__restrict__ div,
                                                     double *__restrict__ uc, double *__restrict__ vc) {
  const long t = (long)blockIdx.x * BLK + threadIdx.x;
  if (t >= (long)m.nx * m.ny) return;
  const int j = (int)(t / m.nx) + 1, i = (int)(t - (long)(j - 1) * m.nx) + 1;
  const double uw = u[HX(i, j)], ue = u[HX(i + 1, j)], vs = v[HX(i, j)], vn = v[HX(i, j + 1)];
  div[t] = (ue - uw) / m.dx + (vn - vs) / m.dy;
  uc[t] = 0.5 * (uw + ue);
  vc[t] = 0.5 * (vs + vn);
}

#undef HX

}  // namespace

// One history row into out[8] (host memory; synchronises the stream).
int history_row(Ctx &c, const double *C, const double *u, const double *v, const double *p, const double *rho,
                const double *T, double *out) {
  Poisson &P = c.ps;
  hipLaunchKernelGGL(k_history, dim3(512), dim3(BLK), 0, c.stream, c.m, C, u, v, p, rho, T, P.part, P.ticket,
                     P.scal + 48);
  COVOF_CHECK_LAUNCH();
  COVOF_CHECK(hipMemcpyAsync(P.hscal, P.scal + 48, 8 * sizeof(double), hipMemcpyDeviceToHost, c.stream));
  COVOF_CHECK(hipStreamSynchronize(c.stream));
  for (int q = 0; q < 8; ++q) out[q] = P.hscal[q];
  return COVOF_OK;
}

int cell_output(Ctx &c, const double *u, const double *v, double *div, double *uc, double *vc) {
  const long n = (long)c.m.nx * c.m.ny;
  hipLaunchKernelGGL(k_cell_output, dim3((unsigned)((n + BLK - 1) / BLK)), dim3(BLK), 0, c.stream, c.m, u, v, div,
                     uc, vc);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
