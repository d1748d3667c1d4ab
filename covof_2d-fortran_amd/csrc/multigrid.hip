// Geometric multigrid and red-black symmetric Gauss-Seidel preconditioners
// for the pressure-Poisson CG (no reference counterpart: they plug in where
// the reference's solve_cg takes a preconditioner, the role of PRECOND_CUSTOM
// in src/module_linear_solvers.f90:22-41, :306-312).
//
// Both are fixed symmetric positive (semi-)definite linear operators, as CG
// requires:
//  * RB-SGS: one symmetric red-black Gauss-Seidel sweep from z = 0
//    (red, black, red): the parallel counterpart of the lexicographic SGS of
//    poisson_ssor, with different iterates.
//  * MG: one V-cycle.  Level l+1 aggregates 2 x 2 cells of level l (ceil at
//    odd sizes).  Coarse operator A_{l+1} = R A_l P with P piecewise-constant
//    injection and R = P^T / 2, which keeps the face form: a coarse face
//    coefficient is half the sum of the fine face coefficients it covers and
//    the coarse diagonal shift is half the children's.  Smoothing is red-black
//    Gauss-Seidel, red then black before the coarse correction and black then
//    red after it (the adjoint order), so the cycle is symmetric; the coarsest
//    level (at most 4096 cells) is relaxed by one workgroup with a fixed
//    palindromic sequence of sweeps.
// Every kernel returns at once when the solve's stop flag is set, so the
// launches of a captured cycle after convergence cost only their dispatch.
#include <algorithm>
#include <cstdlib>

#include "covof_internal.h"

namespace covof {

namespace {

template <typename T>
struct Lev {
  int nx, ny;
  long kxp;  // kx row pitch
  const T *kx, *ky, *diag;
  T *x;
  const T *b;
  int perx, pery;
};

__host__ __device__ __forceinline__ long pitch_of(int nx) { return ((long)(nx + 1) + 63) & ~63L; }

// The solve's stop flags (two scalar slots; a launch outside a solve names a
// slot that is never set).
__device__ __forceinline__ bool stopped(const double *scal, int f0, int f1) {
  return scal[f0] != 0.0 || scal[f1] != 0.0;
}

// Gauss-Seidel value of cell (i, j) from the current x of its neighbours.
template <typename T>
__device__ __forceinline__ T gs_value(const Lev<T> &L, int i, int j) {
  const int nx = L.nx, ny = L.ny;
  const long k = (long)(j - 1) * nx + (i - 1);
  const T *kxr = L.kx + (long)(j - 1) * L.kxp;
  T s = L.b[k];
  if (i > 1) s = s + kxr[i - 1] * L.x[k - 1];
  else if (L.perx) s = s + kxr[0] * L.x[k + (nx - 1)];
  if (i < nx) s = s + kxr[i] * L.x[k + 1];
  else if (L.perx) s = s + kxr[nx] * L.x[k - (nx - 1)];
  if (j > 1) s = s + L.ky[k] * L.x[k - nx];
  else if (L.pery) s = s + L.ky[k] * L.x[k + (long)(ny - 1) * nx];
  if (j < ny) s = s + L.ky[k + nx] * L.x[k + nx];
  else if (L.pery) s = s + L.ky[k + nx] * L.x[k - (long)(ny - 1) * nx];
  return s / L.diag[k];
}

// One half-sweep over the cells of `color` ((i + j) & 1 == color).  Thread =
// the column pair (2p+1, 2p+2) of row j.  INIT: x was zero, so the colour's
// cells take b / diag and the other colour's cells are set to 0.
template <typename T, bool INIT>
__global__ void __launch_bounds__(256) k_mg_smooth(Lev<T> L, int color, const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int i0 = 2 * p + 1;
  if (j > L.ny || i0 > L.nx) return;
  const int ia = (((i0 + j) & 1) == color) ? i0 : i0 + 1;  // the pair's cell of this colour
  const int io = (ia == i0) ? i0 + 1 : i0;
  const long row = (long)(j - 1) * L.nx;
  if (INIT) {
    if (ia <= L.nx) L.x[row + ia - 1] = L.b[row + ia - 1] / L.diag[row + ia - 1];
    if (io <= L.nx) L.x[row + io - 1] = T(0);
  } else if (ia <= L.nx) {
    L.x[row + ia - 1] = gs_value(L, ia, j);
  }
}

// Coarse operator from the fine one (once per assembly).  One thread per
// coarse face.  scheme 0: half the sum of the covered fine faces (R A P with
// R = P^T / 2).  scheme 1 (series-parallel): along each fine row (column) the
// path between the two coarse centres -- the internal faces of the two
// aggregates, where the centres sit -- crosses half of the left aggregate's
// internal face, the fine face and half of the right one's, in series:
// G = 1 / (1/(2 k_l) + 1/k_f + 1/(2 k_r)); the rows add in parallel.  For a
// constant coefficient both give k; across a density jump scheme 1 keeps the
// coarse conductance of the thin low-conductance layers that scheme 0 averages
// away.  Faces next to a one-column aggregate (odd sizes) and domain-edge
// faces use scheme 0.
template <typename T>
__device__ __forceinline__ T series3(T kl, T kf, T kr) {
  if (kf == T(0) || kl == T(0) || kr == T(0)) return T(0);
  return T(1) / (T(0.5) / kl + T(1) / kf + T(0.5) / kr);
}

template <typename T>
__global__ void k_mg_coarsen_faces(int nxf, int nyf, long kxpf, const T *__restrict__ kxf,
                                   const T *__restrict__ kyf, int nxc, int nyc, long kxpc, T *kxc, T *kyc,
                                   int perx, int pery, int scheme) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  auto KX = [&](int i, int j) { return kxf[(long)(j - 1) * kxpf + (i - 1)]; };
  auto KY = [&](int i, int j) { return kyf[(long)(j - 1) * nxf + (i - 1)]; };
  // x face I of coarse row J: fine face 2I-1 (nxf+1 for the east edge) of
  // fine rows 2J-1, 2J
  if (I <= nxc + 1 && J <= nyc) {
    const int fi = (I <= nxc) ? 2 * I - 1 : nxf + 1;
    const int j0 = 2 * J - 1, j1 = 2 * J;
    // internal faces of the aggregates on either side (0: none usable)
    int il = 0, ir = 0;
    if (scheme == 1 && nxf % 2 == 0) {
      if (I > 1 && I <= nxc) {
        il = 2 * I - 2;
        ir = 2 * I;
      } else if (perx) {
        il = nxf;
        ir = 2;
      }
    }
    T s;
    if (il > 0) {
      s = series3(KX(il, j0), KX(fi, j0), KX(ir, j0));
      if (j1 <= nyf) s = s + series3(KX(il, j1), KX(fi, j1), KX(ir, j1));
    } else {
      s = KX(fi, j0);
      if (j1 <= nyf) s = s + KX(fi, j1);
      s = T(0.5) * s;
    }
    kxc[(long)(J - 1) * kxpc + (I - 1)] = s;
  }
  if (I <= nxc && J <= nyc + 1) {
    const int fj = (J <= nyc) ? 2 * J - 1 : nyf + 1;
    const int i0 = 2 * I - 1, i1 = 2 * I;
    int jl = 0, jr = 0;
    if (scheme == 1 && nyf % 2 == 0) {
      if (J > 1 && J <= nyc) {
        jl = 2 * J - 2;
        jr = 2 * J;
      } else if (pery) {
        jl = nyf;
        jr = 2;
      }
    }
    T s;
    if (jl > 0) {
      s = series3(KY(i0, jl), KY(i0, fj), KY(i0, jr));
      if (i1 <= nxf) s = s + series3(KY(i1, jl), KY(i1, fj), KY(i1, jr));
    } else {
      s = KY(i0, fj);
      if (i1 <= nxf) s = s + KY(i1, fj);
      s = T(0.5) * s;
    }
    kyc[(long)(J - 1) * nxc + (I - 1)] = s;
  }
}

// diag_c = its four faces + its diagonal shift: with R = P^T / 2 at every
// level, the shift of a level-l cell is eps0 * (level-0 cells it covers) / 2^l
// (eps0 = the assembly's 1e-12 dvol).
template <typename T>
__global__ void k_mg_coarsen_diag(int nx0, int ny0, int l, double eps0, int nxc, int nyc, long kxpc,
                                  const T *__restrict__ kxc, const T *__restrict__ kyc, T *diagc) {
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (I > nxc || J > nyc) return;
  const long span = 1L << l;
  const long cx = std::min<long>((long)I * span, nx0) - (long)(I - 1) * span;
  const long cy = std::min<long>((long)J * span, ny0) - (long)(J - 1) * span;
  const double shift = eps0 * (double)(cx * cy) / (double)span;
  const long kc = (long)(J - 1) * nxc + (I - 1);
  diagc[kc] = kxc[(long)(J - 1) * kxpc + (I - 1)] + kxc[(long)(J - 1) * kxpc + I] + kyc[kc] + kyc[kc + nxc] + (T)shift;
}

// Residual of the fine level restricted to the coarse right-hand side:
// b_c(I,J) = 1/2 * sum over the children of (b - A x).
template <typename T>
__global__ void k_mg_restrict(Lev<T> F, int nxc, int nyc, T *bc, T scale, const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  const int I = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int J = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (I > nxc || J > nyc) return;
  const int nx = F.nx, ny = F.ny;
  T s = T(0);
  for (int dj = 0; dj < 2; ++dj)
    for (int di = 0; di < 2; ++di) {
      const int i = 2 * I - 1 + di, j = 2 * J - 1 + dj;
      if (i > nx || j > ny) continue;
      const long k = (long)(j - 1) * nx + (i - 1);
      const T *kxr = F.kx + (long)(j - 1) * F.kxp;
      T ax = F.diag[k] * F.x[k];
      if (i > 1) ax = ax - kxr[i - 1] * F.x[k - 1];
      else if (F.perx) ax = ax - kxr[0] * F.x[k + (nx - 1)];
      if (i < nx) ax = ax - kxr[i] * F.x[k + 1];
      else if (F.perx) ax = ax - kxr[nx] * F.x[k - (nx - 1)];
      if (j > 1) ax = ax - F.ky[k] * F.x[k - nx];
      else if (F.pery) ax = ax - F.ky[k] * F.x[k + (long)(ny - 1) * nx];
      if (j < ny) ax = ax - F.ky[k + nx] * F.x[k + nx];
      else if (F.pery) ax = ax - F.ky[k + nx] * F.x[k - (long)(ny - 1) * nx];
      s = s + (F.b[k] - ax);
    }
  bc[(long)(J - 1) * nxc + (I - 1)] = scale * s;
}

// Coarse correction x += P x_c fused with the first post-smoothing half-sweep
// (black, colour 1).  A black cell's Gauss-Seidel value is formed from its
// red neighbours' corrected values x + P x_c; the red cells are not written:
// the red half-sweep that follows overwrites them with values that do not
// depend on their own previous ones, so the result equals correction, black
// sweep, red sweep -- and no thread reads a cell another thread writes.
template <typename T>
__global__ void __launch_bounds__(256) k_mg_prolong_smooth(Lev<T> F, int nxc, const T *__restrict__ xc,
                                                           const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const int i0 = 2 * p + 1;
  const int nx = F.nx, ny = F.ny;
  if (j > ny || i0 > nx) return;
  const int ib = (((i0 + j) & 1) == 1) ? i0 : i0 + 1;  // black cell of the pair
  if (ib > nx) return;
  auto xv = [&](int i, int jj) {
    return F.x[(long)(jj - 1) * nx + (i - 1)] + xc[(long)((jj + 1) / 2 - 1) * nxc + ((i + 1) / 2 - 1)];
  };
  const T *kxr = F.kx + (long)(j - 1) * F.kxp;
  const long k = (long)(j - 1) * nx + (ib - 1);
  T s = F.b[k];
  if (ib > 1) s = s + kxr[ib - 1] * xv(ib - 1, j);
  else if (F.perx) s = s + kxr[0] * xv(nx, j);
  if (ib < nx) s = s + kxr[ib] * xv(ib + 1, j);
  else if (F.perx) s = s + kxr[nx] * xv(1, j);
  if (j > 1) s = s + F.ky[k] * xv(ib, j - 1);
  else if (F.pery) s = s + F.ky[k] * xv(ib, ny);
  if (j < ny) s = s + F.ky[k + nx] * xv(ib, j + 1);
  else if (F.pery) s = s + F.ky[k + nx] * xv(ib, 1);
  F.x[k] = s / F.diag[k];
}

// The bottom of the V-cycle in one workgroup: every level from the first one
// with at most MG_BOTTOM_CELLS cells down to the coarsest, with x and the
// right-hand sides in LDS (the face coefficients are read from global memory,
// where they stay cached).  Same cycle as the multi-block levels: NU
// red-black sweeps from x = 0, restriction, ..., K red-black sweeps and K in
// the adjoint order on the coarsest level, then prolongation and NU sweeps in
// the adjoint order.  Within a half-sweep every update reads the state before
// the half-sweep (values are staged in registers across a barrier), so odd
// periodic sizes (same-colour wrap neighbours) stay deterministic.
constexpr int MG_BOTTOM_CELLS = 4096;
constexpr int MG_BOTTOM_LDS = 6144;  // cells of all bottom levels together

struct BotLev {
  int nx, ny;
  long kxp;
  const void *kx, *ky, *diag;
  int off;  // offset of the level's cells in the LDS arrays
};
struct Bottom {
  int nlev;  // levels handled here
  BotLev lev[MG_MAX_LEVELS];
  int perx, pery, K, NU;
  double scale;  // restriction weight (R = scale * P^T)
};

template <typename T>
__global__ void __launch_bounds__(1024) k_mg_bottom(Bottom B, const T *__restrict__ b_in, T *__restrict__ x_out,
                                                    const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  __shared__ T xs[MG_BOTTOM_LDS];
  __shared__ T bs[MG_BOTTOM_LDS];
  const int tid = threadIdx.x;
  {
    const int n0 = B.lev[0].nx * B.lev[0].ny;
    for (int t = tid; t < n0; t += 1024) bs[t] = b_in[t];
  }
  // coefficients of this thread's cells (at most 4 per level) in registers,
  // loaded once per visit of a level: a half-sweep then reads only LDS
  T cw[4], ce[4], cs[4], cn[4], cd[4];
  auto load_coef = [&](int l) {
    const BotLev &L = B.lev[l];
    const int nx = L.nx, n = nx * L.ny;
    const T *kx = (const T *)L.kx, *ky = (const T *)L.ky, *dg = (const T *)L.diag;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = tid + q * 1024;
      cw[q] = ce[q] = cs[q] = cn[q] = T(0);
      cd[q] = T(1);
      if (t >= n) continue;
      const int i = t % nx + 1, j = t / nx + 1;
      const T *kxr = kx + (long)(j - 1) * L.kxp;
      cw[q] = kxr[i - 1];
      ce[q] = kxr[i];
      cs[q] = ky[t];
      cn[q] = ky[t + nx];
      cd[q] = dg[t];
    }
  };
  auto half = [&](int l, int color) {
    const BotLev &L = B.lev[l];
    const int nx = L.nx, ny = L.ny, n = nx * ny;
    T *x = xs + L.off;
    const T *b = bs + L.off;
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = tid + q * 1024;
      v[q] = T(0);
      if (t >= n) continue;
      const int i = t % nx + 1, j = t / nx + 1;
      if (((i + j) & 1) != color) continue;
      T s = b[t];
      if (i > 1) s = s + cw[q] * x[t - 1];
      else if (B.perx) s = s + cw[q] * x[t + (nx - 1)];
      if (i < nx) s = s + ce[q] * x[t + 1];
      else if (B.perx) s = s + ce[q] * x[t - (nx - 1)];
      if (j > 1) s = s + cs[q] * x[t - nx];
      else if (B.pery) s = s + cs[q] * x[t + (ny - 1) * nx];
      if (j < ny) s = s + cn[q] * x[t + nx];
      else if (B.pery) s = s + cn[q] * x[t - (ny - 1) * nx];
      v[q] = s / cd[q];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = tid + q * 1024;
      if (t >= n) continue;
      const int i = t % nx + 1, j = t / nx + 1;
      if (((i + j) & 1) == color) x[t] = v[q];
    }
    __syncthreads();
  };
  auto zero = [&](int l) {
    const int n = B.lev[l].nx * B.lev[l].ny;
    for (int t = tid; t < n; t += 1024) xs[B.lev[l].off + t] = T(0);
    __syncthreads();
  };
  const int last = B.nlev - 1;
  for (int l = 0; l < last; ++l) {
    zero(l);
    load_coef(l);
    for (int s = 0; s < B.NU; ++s) {
      half(l, 0);
      half(l, 1);
    }
    // restriction of the residual into the next level's right-hand side
    const BotLev &F = B.lev[l], &C = B.lev[l + 1];
    const T *kx = (const T *)F.kx, *ky = (const T *)F.ky, *dg = (const T *)F.diag;
    const T *x = xs + F.off, *b = bs + F.off;
    const int nx = F.nx, ny = F.ny, nc = C.nx * C.ny;
    for (int t = tid; t < nc; t += 1024) {
      const int I = t % C.nx + 1, J = t / C.nx + 1;
      T sum = T(0);
      for (int dj = 0; dj < 2; ++dj)
        for (int di = 0; di < 2; ++di) {
          const int i = 2 * I - 1 + di, j = 2 * J - 1 + dj;
          if (i > nx || j > ny) continue;
          const int k = (j - 1) * nx + (i - 1);
          const T *kxr = kx + (long)(j - 1) * F.kxp;
          T ax = dg[k] * x[k];
          if (i > 1) ax = ax - kxr[i - 1] * x[k - 1];
          else if (B.perx) ax = ax - kxr[0] * x[k + (nx - 1)];
          if (i < nx) ax = ax - kxr[i] * x[k + 1];
          else if (B.perx) ax = ax - kxr[nx] * x[k - (nx - 1)];
          if (j > 1) ax = ax - ky[k] * x[k - nx];
          else if (B.pery) ax = ax - ky[k] * x[k + (ny - 1) * nx];
          if (j < ny) ax = ax - ky[k + nx] * x[k + nx];
          else if (B.pery) ax = ax - ky[k + nx] * x[k - (ny - 1) * nx];
          sum = sum + (b[k] - ax);
        }
      bs[C.off + t] = (T)B.scale * sum;
    }
    __syncthreads();
  }
  zero(last);
  load_coef(last);
  for (int s = 0; s < B.K; ++s) {
    half(last, 0);
    half(last, 1);
  }
  for (int s = 0; s < B.K; ++s) {
    half(last, 1);
    half(last, 0);
  }
  for (int l = last - 1; l >= 0; --l) {
    const BotLev &F = B.lev[l], &C = B.lev[l + 1];
    const int n = F.nx * F.ny;
    for (int t = tid; t < n; t += 1024) {
      const int i = t % F.nx + 1, j = t / F.nx + 1;
      xs[F.off + t] += xs[C.off + ((j + 1) / 2 - 1) * C.nx + ((i + 1) / 2 - 1)];
    }
    __syncthreads();
    load_coef(l);
    for (int s = 0; s < B.NU; ++s) {
      half(l, 1);
      half(l, 0);
    }
  }
  {
    const int n0 = B.lev[0].nx * B.lev[0].ny;
    for (int t = tid; t < n0; t += 1024) x_out[t] = xs[t];
  }
}

template <typename T>
struct Hier;
template <>
struct Hier<double> {
  static Mg &get(Ctx &c) { return c.ps.mg64; }
  static const double *kx(Ctx &c) { return c.ps.kx; }
  static const double *ky(Ctx &c) { return c.ps.ky; }
  static const double *diag(Ctx &c) { return c.ps.diag; }
};
template <>
struct Hier<float> {
  static Mg &get(Ctx &c) { return c.ps.mg32; }
  static const float *kx(Ctx &c) { return c.ps.kx32; }
  static const float *ky(Ctx &c) { return c.ps.ky32; }
  static const float *diag(Ctx &c) { return c.ps.diag32; }
};

template <typename T>
Lev<T> level(Ctx &c, int l, const T *b, T *x) {
  Mg &G = Hier<T>::get(c);
  Lev<T> L;
  L.nx = G.lev[l].nx;
  L.ny = G.lev[l].ny;
  L.kxp = pitch_of(L.nx);
  L.perx = c.m.perx;
  L.pery = c.m.pery;
  if (l == 0) {
    L.kx = Hier<T>::kx(c);
    L.ky = Hier<T>::ky(c);
    L.diag = Hier<T>::diag(c);
  } else {
    L.kx = (const T *)G.lev[l].kx;
    L.ky = (const T *)G.lev[l].ky;
    L.diag = (const T *)G.lev[l].diag;
  }
  L.x = (l == 0) ? x : (T *)G.lev[l].x;
  L.b = (l == 0) ? b : (const T *)G.lev[l].b;
  return L;
}

// Coarse-face scheme (see k_mg_coarsen_faces): averaged faces by default
// (fewer iterations on the density-jump workloads measured), series faces
// with the context's mg_coarse option.
inline int coarse_scheme(const Ctx &c) { return c.opts.mg_coarse ? 1 : 0; }

// Red-black sweeps before and after each coarse correction (1..4, 1 by
// default).
inline int smoothing_sweeps(const Ctx &c) { return c.opts.mg_sweeps; }

// Weight of the coarse-grid correction (scales R), in (0, 2], 1.8 by
// default.  With the piecewise-constant transfer the Galerkin correction
// (weight 1) is too weak -- the coarse operators of aggregation are known to
// over-estimate the energy of smooth modes -- and an over-correction in
// (1, 2) restores the convergence (D. Braess, Computing 55 (1995) 379-393);
// a weight <= 2 keeps the cycle symmetric positive definite.  At 2048^2 with
// density ratio 1000 it takes the CG from 231 to 34 iterations.
inline double correction_weight(const Ctx &c) { return c.opts.mg_omega; }

// Coarsen while the level has more than 16 cells and both sides of the next
// level keep at least 2 cells (a side of 1 in a walled box would leave a cell
// whose only coupling is the 1e-12 dvol shift).
inline bool coarsen_more(int nx, int ny) { return (long)nx * ny > 16 && nx >= 4 && ny >= 4; }

inline dim3 pair_grid(int nx, int ny) { return dim3(((nx + 1) / 2 + 63) / 64, (ny + 3) / 4); }
inline dim3 cell_grid(int nx, int ny) { return dim3((nx + 63) / 64, (ny + 3) / 4); }

// The level buffers of the hierarchy (mesh-dependent only).
template <typename T>
int reserve(Ctx &c) {
  Mg &G = Hier<T>::get(c);
  const Mesh &m = c.m;
  if (!G.nlev) {
    // levels: halve (ceil) down to a few cells (coarsen_more); a periodic
    // direction is halved only while even
    int nx = m.nx, ny = m.ny;
    G.lev[0].nx = nx;
    G.lev[0].ny = ny;
    int nl = 1;
    while (coarsen_more(nx, ny) && nl < MG_MAX_LEVELS) {
      const bool cx = nx > 1 && (!m.perx || nx % 2 == 0);
      const bool cy = ny > 1 && (!m.pery || ny % 2 == 0);
      if (!cx || !cy) break;
      nx = (nx + 1) / 2;
      ny = (ny + 1) / 2;
      MgLevelBuf &B = G.lev[nl];
      B.nx = nx;
      B.ny = ny;
      const long n = (long)nx * ny;
      COVOF_CHECK(hipMalloc(&B.kx, pitch_of(nx) * ny * sizeof(T)));
      COVOF_CHECK(hipMalloc(&B.ky, (long)nx * (ny + 1) * sizeof(T)));
      COVOF_CHECK(hipMalloc(&B.diag, n * sizeof(T)));
      COVOF_CHECK(hipMalloc(&B.x, n * sizeof(T)));
      COVOF_CHECK(hipMalloc(&B.b, n * sizeof(T)));
      ++nl;
    }
    G.nlev = nl;
  }
  return COVOF_OK;
}

template <typename T>
int prepare(Ctx &c) {
  Mg &G = Hier<T>::get(c);
  const Mesh &m = c.m;
  const int st = reserve<T>(c);
  if (st) return st;
  if (G.built != c.ps.assembly_id) {
    for (int l = 1; l < G.nlev; ++l) {
      const Lev<T> F = level<T>(c, l - 1, nullptr, nullptr);
      MgLevelBuf &B = G.lev[l];
      const int nxc = B.nx, nyc = B.ny;
      hipLaunchKernelGGL(k_mg_coarsen_faces<T>, cell_grid(nxc + 1, nyc + 1), dim3(64, 4), 0, c.stream, F.nx, F.ny,
                         F.kxp, F.kx, F.ky, nxc, nyc, pitch_of(nxc), (T *)B.kx, (T *)B.ky, m.perx, m.pery,
                         coarse_scheme(c));
      hipLaunchKernelGGL(k_mg_coarsen_diag<T>, cell_grid(nxc, nyc), dim3(64, 4), 0, c.stream, m.nx, m.ny, l,
                         1.0e-12 * (m.dx * m.dy), nxc, nyc, pitch_of(nxc), (const T *)B.kx, (const T *)B.ky,
                         (T *)B.diag);
    }
    COVOF_CHECK_LAUNCH();
    G.built = c.ps.assembly_id;
  }
  return COVOF_OK;
}

template <typename T>
void smooth(Ctx &c, const Lev<T> &L, int color, bool init, int f0, int f1) {
  if (init)
    hipLaunchKernelGGL((k_mg_smooth<T, true>), pair_grid(L.nx, L.ny), dim3(64, 4), 0, c.stream, L, color,
                       (const double *)c.ps.scal, f0, f1);
  else
    hipLaunchKernelGGL((k_mg_smooth<T, false>), pair_grid(L.nx, L.ny), dim3(64, 4), 0, c.stream, L, color,
                       (const double *)c.ps.scal, f0, f1);
}

template <typename T>
void apply(Ctx &c, int kind, const T *r, T *z, int f0, int f1) {
  Mg &G = Hier<T>::get(c);
  const int NU = smoothing_sweeps(c);  // red-black sweeps before and after the coarse correction
  if (kind == MG_KIND_SGS_RB) {
    const Lev<T> L = level<T>(c, 0, r, z);
    smooth(c, L, 0, true, f0, f1);
    smooth(c, L, 1, false, f0, f1);
    smooth(c, L, 0, false, f0, f1);
    return;
  }
  const int nl = G.nlev;
  // first level handled by the single-workgroup bottom cycle: as many of the
  // coarsest levels as fit its LDS arrays (each at most MG_BOTTOM_CELLS)
  int lb = nl - 1;
  long held = (long)G.lev[lb].nx * G.lev[lb].ny;
  while (lb > 0) {
    const long cells = (long)G.lev[lb - 1].nx * G.lev[lb - 1].ny;
    if (cells > MG_BOTTOM_CELLS || held + cells > MG_BOTTOM_LDS) break;
    held += cells;
    --lb;
  }
  const T scale = (T)(0.5 * correction_weight(c));
  for (int l = 0; l < lb; ++l) {
    const Lev<T> L = level<T>(c, l, r, z);
    for (int s = 0; s < NU; ++s) {
      smooth(c, L, 0, s == 0, f0, f1);
      smooth(c, L, 1, false, f0, f1);
    }
    const MgLevelBuf &B = G.lev[l + 1];
    hipLaunchKernelGGL(k_mg_restrict<T>, cell_grid(B.nx, B.ny), dim3(64, 4), 0, c.stream, L, B.nx, B.ny, (T *)B.b,
                       scale, (const double *)c.ps.scal, f0, f1);
  }
  {
    Bottom Bt{};
    Bt.nlev = nl - lb;
    int off = 0;
    for (int l = lb; l < nl; ++l) {
      const Lev<T> L = level<T>(c, l, r, z);
      BotLev &D = Bt.lev[l - lb];
      D.nx = L.nx;
      D.ny = L.ny;
      D.kxp = L.kxp;
      D.kx = L.kx;
      D.ky = L.ky;
      D.diag = L.diag;
      D.off = off;
      off += L.nx * L.ny;
    }
    Bt.perx = c.m.perx;
    Bt.pery = c.m.pery;
    const Lev<T> Lc = level<T>(c, nl - 1, r, z);
    Bt.K = std::min(64, 2 * std::max(Lc.nx, Lc.ny));
    Bt.NU = NU;
    Bt.scale = 0.5 * correction_weight(c);
    const Lev<T> Lb = level<T>(c, lb, r, z);
    hipLaunchKernelGGL(k_mg_bottom<T>, dim3(1), dim3(1024), 0, c.stream, Bt, Lb.b, Lb.x, (const double *)c.ps.scal,
                       f0, f1);
  }
  for (int l = lb - 1; l >= 0; --l) {
    const Lev<T> L = level<T>(c, l, r, z);
    const MgLevelBuf &B = G.lev[l + 1];
    hipLaunchKernelGGL(k_mg_prolong_smooth<T>, pair_grid(L.nx, L.ny), dim3(64, 4), 0, c.stream, L, B.nx,
                       (const T *)B.x, (const double *)c.ps.scal, f0, f1);
    smooth(c, L, 0, false, f0, f1);
    for (int s = 1; s < NU; ++s) {
      smooth(c, L, 1, false, f0, f1);
      smooth(c, L, 0, false, f0, f1);
    }
  }
}

}  // namespace

int mg_check(const Ctx &c, int kind) {
  const Mesh &m = c.m;
  // the multi-block red-black half-sweeps need even periodic sizes
  if ((m.perx && m.nx % 2) || (m.pery && m.ny % 2)) return COVOF_BAD_ARGUMENT;
  if (kind == MG_KIND_VCYCLE) {
    // the coarsest level must fit one workgroup
    int nx = m.nx, ny = m.ny, nl = 1;
    while (coarsen_more(nx, ny) && nl < MG_MAX_LEVELS) {
      const bool cx = nx > 1 && (!m.perx || nx % 2 == 0);
      const bool cy = ny > 1 && (!m.pery || ny % 2 == 0);
      if (!cx || !cy) break;
      nx = (nx + 1) / 2;
      ny = (ny + 1) / 2;
      ++nl;
    }
    if ((long)nx * ny > MG_BOTTOM_CELLS) return COVOF_BAD_ARGUMENT;
  }
  return COVOF_OK;
}

int mg_prepare(Ctx &c, bool fp32) { return fp32 ? prepare<float>(c) : prepare<double>(c); }
int mg_reserve(Ctx &c) {
  const int st = reserve<double>(c);
  return st ? st : reserve<float>(c);
}

void mg_apply64(Ctx &c, int kind, const double *r, double *z, int f0, int f1) {
  apply<double>(c, kind, r, z, f0, f1);
}
void mg_apply32(Ctx &c, int kind, const float *r, float *z, int f0, int f1) {
  apply<float>(c, kind, r, z, f0, f1);
}

void mg_free(Ctx &c) {
  for (Mg *G : {&c.ps.mg64, &c.ps.mg32}) {
    for (int l = 1; l < G->nlev; ++l) {
      MgLevelBuf &B = G->lev[l];
      for (void *p : {B.kx, B.ky, B.diag, B.x, B.b})
        if (p) (void)hipFree(p);
      B = MgLevelBuf();
    }
    G->nlev = 0;
    G->built = 0;
  }
}

}  // namespace covof
