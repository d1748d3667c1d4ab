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
#include <type_traits>

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
  const double a = scal[f0], b = scal[f1];  // both loads issued together
  return (a != 0.0) | (b != 0.0);
}

// Keeps the loads issued above it in place (the compiler would otherwise sink
// them past a following early return, behind the stop flags' load).
// (p, q: the restrict-qualified arrays among them, which the barrier must see)
__device__ __forceinline__ void loads_stay_above(const void *p, const void *q) {
  asm volatile("" ::"s"(p), "s"(q) : "memory");
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

// ---------------------------------------------------------------- fused level
//
// With one red-black sweep before and after the coarse correction (the
// default), a level's five launches above fold into two, each one pass over
// the level's arrays:
//  * k_mg_pre: the pre-smoothing from x = 0 leaves red = b / diag and black
//    = the Gauss-Seidel value from those reds; the kernel forms both on a
//    tile with a two-cell halo in LDS, then the residual of the tile's cells
//    and its restriction.  It writes only the coarse right-hand side: the
//    post-smoothing needs no pre-smoothed value but the reds, which it forms
//    again from b / diag.
//  * k_mg_post: red + P x_c from b / diag, the black sweep from those, the
//    red sweep from the new blacks; writes the level's x.
// Every value is the expression of the kernels above on the same operands
// (same terms, same order), so the cycle is bitwise the five-launch form.
// Tile: 64 x 16 fine cells (32 x 8 coarse), 256 threads; region = tile + 2
// (the blacks next to the tile need the reds two cells out); a periodic side
// wraps the region's cells (even sizes: the colours agree across the wrap).
constexpr int FT_X = 64, FT_Y = 16, FT_H = 2;
constexpr int FR_X = FT_X + 2 * FT_H, FR_Y = FT_Y + 2 * FT_H;
constexpr int FR_N = FR_X * FR_Y;
constexpr int FT_BLK = 256;
static_assert((FT_X / 2) * (FT_Y / 2) == FT_BLK, "one thread per coarse cell of the tile");

__device__ __forceinline__ int wrap1(int g, int n) {
  int q = (g - 1) % n;
  return (q < 0 ? q + n : q) + 1;
}

// Tile of block b: consecutive blocks go to the eight XCDs in turn, so XCD x
// takes the contiguous run of tiles [start_x, start_x + count_x): neighbour
// tiles, whose cells the halos re-read, mostly share that XCD's L2.
__device__ __forceinline__ int mg_xcd_tile(int b, int nt) {
  const int q = nt >> 3, rem = nt & 7, x = b & 7;
  return x * q + min(x, rem) + (b >> 3);
}

// The region's cells are taken in horizontal pairs, region columns (2p, 2p+1)
// of a row c.  Region column 0 is global column i0 = 64 tx - 1 (odd) and a
// periodic side is even, so a pair is an (odd, even) global column pair: the
// two fine columns of one coarse column, with one red cell, the first iff c is
// even (j0 = 16 ty - 1 is odd).  A thread keeps its pairs' b, diag and face
// coefficients in registers from the loads to the last phase; only x goes
// through LDS (a cell needs its neighbours' x and nothing else of theirs).
constexpr int FP_X = FR_X / 2;                      // pairs per region row
constexpr int FP_N = FP_X * FR_Y;                   // pairs per region
constexpr int FP_S = (FP_N + FT_BLK - 1) / FT_BLK;  // pair slots per thread

template <typename T>
struct Two;
template <>
struct Two<float> {
  using type = float2;
};
template <>
struct Two<double> {
  using type = double2;
};

// One pair: cell 0 = region column a0, cell 1 = a0 + 1, region row c.
// w, e, s, n: the cell's west, east, south and north face coefficients.
template <typename T, typename TB>
struct Pair {
  TB b0, b1;
  T d0, d1, w0, e0, w1, e1, s0, s1, n0, n1;
  T xc;          // the coarse cell's x (post pass)
  int c, a0;     // region row and first column
  int gi0, gj;   // global column of cell 0 and global row (wrapped)
  bool ok0, ok1; // inside the domain
};

// The loads of a thread's pairs.  INNER: the tile's whole region lies inside
// the domain and every array is aligned for two-value loads (even nx), so a
// pair's values come in two-value loads with no wrap or validity to settle.
// Otherwise every load is a single value and unconditional: a cell outside
// the domain reads cell (1, 1), whose value its users never take.
template <bool INNER, bool POST, typename T, typename TB>
__device__ __forceinline__ void pair_loads(const Lev<T> &F, const TB *__restrict__ bin, int i0, int j0, int nxc,
                                           const T *__restrict__ xcv, Pair<T, TB> (&p)[FP_S]) {
  const int nx = F.nx, ny = F.ny, kxp = (int)F.kxp;
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) P = FP_N - 1;  // a slot past the region reads the last pair again (never used)
    const int c = P / FP_X, a0 = 2 * (P - c * FP_X);
    Pair<T, TB> &q = p[s];
    q.c = c;
    q.a0 = a0;
    int gi0 = i0 + a0, gj = j0 + c;
    if (INNER) {
      using T2 = typename Two<T>::type;
      using TB2 = typename Two<TB>::type;
      const int k = (gj - 1) * nx + (gi0 - 1), kk = (gj - 1) * kxp + (gi0 - 1);  // both even
      const TB2 bv = *reinterpret_cast<const TB2 *>(bin + k);
      const T2 dv = *reinterpret_cast<const T2 *>(F.diag + k);
      const T2 sv = *reinterpret_cast<const T2 *>(F.ky + k);
      const T2 nv = *reinterpret_cast<const T2 *>(F.ky + k + nx);
      const T2 xa = *reinterpret_cast<const T2 *>(F.kx + kk);
      const T2 xb = *reinterpret_cast<const T2 *>(F.kx + kk + 2);  // second value: row padding at most
      q.b0 = bv.x;
      q.b1 = bv.y;
      q.d0 = dv.x;
      q.d1 = dv.y;
      q.s0 = sv.x;
      q.s1 = sv.y;
      q.n0 = nv.x;
      q.n1 = nv.y;
      q.w0 = xa.x;
      q.e0 = xa.y;
      q.w1 = xa.y;
      q.e1 = xb.x;
      q.ok0 = q.ok1 = true;
      if (POST) q.xc = xcv[((gj + 1) / 2 - 1) * nxc + ((gi0 + 1) / 2 - 1)];
    } else {
      bool okx = true, oky = true;
      if (gi0 < 1 || gi0 > nx) {
        if (F.perx) gi0 = wrap1(gi0, nx);
        else okx = false;
      }
      if (gj < 1 || gj > ny) {
        if (F.pery) gj = wrap1(gj, ny);
        else oky = false;
      }
      // gi0 is odd: outside on the west it is -1 (so is cell 1, column 0);
      // wrapped it is at most nx - 1 (nx even), so cell 1 needs no wrap
      q.ok0 = okx && oky;
      q.ok1 = q.ok0 && gi0 + 1 <= nx;
      const int k0 = q.ok0 ? (gj - 1) * nx + (gi0 - 1) : 0, k1 = q.ok1 ? (gj - 1) * nx + gi0 : 0;
      const int kk0 = q.ok0 ? (gj - 1) * kxp + (gi0 - 1) : 0, kk1 = q.ok1 ? (gj - 1) * kxp + gi0 : 0;
      q.b0 = bin[k0];
      q.b1 = bin[k1];
      q.d0 = F.diag[k0];
      q.d1 = F.diag[k1];
      q.s0 = F.ky[k0];
      q.s1 = F.ky[k1];
      q.n0 = F.ky[k0 + nx];
      q.n1 = F.ky[k1 + nx];
      q.w0 = F.kx[kk0];
      q.e0 = F.kx[kk0 + 1];
      q.w1 = F.kx[kk1];
      q.e1 = F.kx[kk1 + 1];
      if (POST) {
        // the red cell's coarse cell (both cells of a pair share it)
        const bool okr = (c & 1) ? q.ok1 : q.ok0;
        q.xc = xcv[okr ? ((gj + 1) / 2 - 1) * nxc + ((gi0 + 1) / 2 - 1) : 0];
      }
    }
    q.gi0 = gi0;
    q.gj = gj;
  }
}

// x of the pairs' reds from x = 0 (b / diag; POST: + the coarse correction),
// 0 at the blacks.
template <bool POST, typename T, typename TB>
__device__ __forceinline__ void pair_reds(const Pair<T, TB> (&p)[FP_S], T *xs) {
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    const int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) break;
    const Pair<T, TB> &q = p[s];
    const int rc = q.c & 1;  // which cell is red
    const T br = rc ? (T)q.b1 : (T)q.b0, dr = rc ? q.d1 : q.d0;
    const bool okr = rc ? q.ok1 : q.ok0;
    T v = br / dr;
    if (POST) v = v + q.xc;
    xs[2 * P + rc] = okr ? v : T(0);
    xs[2 * P + 1 - rc] = T(0);
  }
}

// Gauss-Seidel value of region cell e (global column gi, row gj) from the x
// of its neighbours in LDS (gs_value's terms and order).
template <bool INNER, typename T>
__device__ __forceinline__ T pair_gs(const Lev<T> &F, int gi, int gj, T bv, T dv, T w, T e, T s, T n,
                                     const T *xs, int k) {
  T a = bv;
  if (INNER || gi > 1 || F.perx) a = a + w * xs[k - 1];
  if (INNER || gi < F.nx || F.perx) a = a + e * xs[k + 1];
  if (INNER || gj > 1 || F.pery) a = a + s * xs[k - FR_X];
  if (INNER || gj < F.ny || F.pery) a = a + n * xs[k + FR_X];
  return a / dv;
}

// The black half-sweep over the region minus its outer ring, from the reds.
template <bool INNER, typename T, typename TB>
__device__ __forceinline__ void pair_blacks(const Lev<T> &F, const Pair<T, TB> (&p)[FP_S], T *xs) {
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    const int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) break;
    const Pair<T, TB> &q = p[s];
    const int bk = 1 - (q.c & 1);  // which cell is black
    const int ab = q.a0 + bk;
    const bool okb = bk ? q.ok1 : q.ok0;
    if (q.c < 1 || q.c > FR_Y - 2 || ab < 1 || ab > FR_X - 2 || !okb) continue;
    const T v = bk ? pair_gs<INNER>(F, q.gi0 + 1, q.gj, (T)q.b1, q.d1, q.w1, q.e1, q.s1, q.n1, xs, 2 * P + 1)
                   : pair_gs<INNER>(F, q.gi0, q.gj, (T)q.b0, q.d0, q.w0, q.e0, q.s0, q.n0, xs, 2 * P);
    xs[2 * P + bk] = v;
  }
}

// Whether pair slot q holds tile cells (region columns 2..65, rows 2..17),
// and for each cell whether it is inside the domain (a tile cell is never
// west or south of it, and is not wrapped).
template <typename T, typename TB>
__device__ __forceinline__ bool pair_in_tile(const Pair<T, TB> &q, int i0, int j0, int nx, int ny, bool &in0,
                                             bool &in1) {
  if (q.c < FT_H || q.c >= FT_H + FT_Y || q.a0 < FT_H || q.a0 >= FT_H + FT_X) return false;
  const bool row = j0 + q.c <= ny;
  in0 = row && i0 + q.a0 <= nx;
  in1 = row && i0 + q.a0 + 1 <= nx;
  return true;
}

template <bool INNER, typename T, typename TB>
__device__ __forceinline__ void pre_pass(const Lev<T> &F, const TB *__restrict__ bin, int tx, int ty, int i0, int j0,
                                         int nxc, int nyc, T *__restrict__ bc, T scale, const double *scal, int f0,
                                         int f1, T *xs, T *rs) {
  const int nx = F.nx, ny = F.ny;
  Pair<T, TB> p[FP_S];
  pair_loads<INNER, false>(F, bin, i0, j0, nxc, (const T *)nullptr, p);
  loads_stay_above(bin, bin);
  if (stopped(scal, f0, f1)) return;
  pair_reds<false>(p, xs);
  __syncthreads();
  pair_blacks<INNER>(F, p, xs);
  __syncthreads();
  // residual b - A x of the tile's cells (k_mg_restrict's terms and order)
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    const int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) break;
    const Pair<T, TB> &q = p[s];
    bool in0 = true, in1 = true;
    if (!pair_in_tile(q, i0, j0, nx, ny, in0, in1)) continue;
    const int gi = i0 + q.a0, gj = j0 + q.c;
    const int r = (q.c - FT_H) * FT_X + (q.a0 - FT_H);
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      if (!(h ? in1 : in0)) continue;
      const int k = 2 * P + h;
      const T w = h ? q.w1 : q.w0, e = h ? q.e1 : q.e0, so = h ? q.s1 : q.s0, no = h ? q.n1 : q.n0;
      T ax = (h ? q.d1 : q.d0) * xs[k];
      if (INNER || gi + h > 1 || F.perx) ax = ax - w * xs[k - 1];
      if (INNER || gi + h < nx || F.perx) ax = ax - e * xs[k + 1];
      if (INNER || gj > 1 || F.pery) ax = ax - so * xs[k - FR_X];
      if (INNER || gj < ny || F.pery) ax = ax - no * xs[k + FR_X];
      rs[r + h] = (T)(h ? q.b1 : q.b0) - ax;
    }
  }
  __syncthreads();
  // restriction: b_c(I, J) = scale * (sum of the children's residuals)
  const int ci = threadIdx.x % (FT_X / 2), cj = threadIdx.x / (FT_X / 2);
  const int I = tx * (FT_X / 2) + ci + 1, J = ty * (FT_Y / 2) + cj + 1;
  if (I > nxc || J > nyc) return;
  T sum = T(0);
  for (int dj = 0; dj < 2; ++dj)
    for (int di = 0; di < 2; ++di) {
      if (!INNER && (2 * I - 1 + di > nx || 2 * J - 1 + dj > ny)) continue;
      sum = sum + rs[(2 * cj + dj) * FT_X + 2 * ci + di];
    }
  bc[(long)(J - 1) * nxc + (I - 1)] = scale * sum;
}

// Lane 0's value: the sum of the wave's 64 values (a fixed tree).
__device__ __forceinline__ double wave_total(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// rzt: when not null, the tile's sum of r z (r = the right-hand side as
// loaded, z = the value stored) goes to rzt[tile] (wsum: four doubles of LDS).
template <bool INNER, typename T, typename TB, typename TX>
__device__ __forceinline__ void post_pass(const Lev<T> &F, const TB *__restrict__ bin, TX *__restrict__ xout,
                                          int i0, int j0, int nxc, const T *__restrict__ xc, const double *scal,
                                          int f0, int f1, T *xs, double *rzt, int tile, double *wsum) {
  const int nx = F.nx, ny = F.ny;
  Pair<T, TB> p[FP_S];
  pair_loads<INNER, true>(F, bin, i0, j0, nxc, xc, p);
  loads_stay_above(bin, xc);
  if (stopped(scal, f0, f1)) return;
  // reds corrected by the coarse level: the pre-smoothed red b / diag + x_c
  pair_reds<true>(p, xs);
  __syncthreads();
  // black half-sweep from the corrected reds (k_mg_prolong_smooth)
  pair_blacks<INNER>(F, p, xs);
  __syncthreads();
  // red half-sweep from the new blacks; the tile's x
  double rz = 0.0;
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    const int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) break;
    const Pair<T, TB> &q = p[s];
    bool in0 = true, in1 = true;
    if (!pair_in_tile(q, i0, j0, nx, ny, in0, in1)) continue;
    const int gi = i0 + q.a0, gj = j0 + q.c;
    const int rc = q.c & 1;
    const T v0 = rc ? xs[2 * P] : pair_gs<INNER>(F, gi, gj, (T)q.b0, q.d0, q.w0, q.e0, q.s0, q.n0, xs, 2 * P);
    const T v1 = rc ? pair_gs<INNER>(F, gi + 1, gj, (T)q.b1, q.d1, q.w1, q.e1, q.s1, q.n1, xs, 2 * P + 1)
                    : xs[2 * P + 1];
    const int k = (gj - 1) * nx + (gi - 1);
    const TX z0 = (TX)v0, z1 = (TX)v1;
    if (INNER) {
      using TX2 = typename Two<TX>::type;
      TX2 o;
      o.x = z0;
      o.y = z1;
      *reinterpret_cast<TX2 *>(xout + k) = o;
    } else {
      if (in0) xout[k] = z0;
      if (in1) xout[k + 1] = z1;
    }
    if (rzt) {
      if (in0) rz += (double)q.b0 * (double)z0;
      if (in1) rz += (double)q.b1 * (double)z1;
    }
  }
  if (rzt) {
    rz = wave_total(rz);
    if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = rz;
    __syncthreads();
    if (threadIdx.x == 0) rzt[tile] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
  }
}

// The tile partials of r.z as the finest fused post pass forms them (same
// tiles, same pairs per thread, same order), for a cycle that ran without the
// fused passes: the solve then takes the same r.z either way.
__global__ void __launch_bounds__(FT_BLK) k_rz_tiles(int nx, int ny, const double *__restrict__ r,
                                                     const double *__restrict__ z, double *__restrict__ rzt,
                                                     const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  __shared__ double wsum[FT_BLK / 64];
  const int ntx = (nx + FT_X - 1) / FT_X;
  const int t = mg_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = t / ntx;
  const int i0 = tx * FT_X + 1 - FT_H, j0 = ty * FT_Y + 1 - FT_H;
  double rz = 0.0;
#pragma unroll
  for (int s = 0; s < FP_S; ++s) {
    const int P = threadIdx.x + s * FT_BLK;
    if (P >= FP_N) break;
    const int c = P / FP_X, a0 = 2 * (P - c * FP_X);
    if (c < FT_H || c >= FT_H + FT_Y || a0 < FT_H || a0 >= FT_H + FT_X) continue;
    const bool row = j0 + c <= ny, in0 = row && i0 + a0 <= nx, in1 = row && i0 + a0 + 1 <= nx;
    const int k = (j0 + c - 1) * nx + (i0 + a0 - 1);
    if (in0) rz += r[k] * z[k];
    if (in1) rz += r[k + 1] * z[k + 1];
  }
  rz = wave_total(rz);
  if ((threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = rz;
  __syncthreads();
  if (threadIdx.x == 0) rzt[t] = ((wsum[0] + wsum[1]) + wsum[2]) + wsum[3];
}

// The reduction grid's r.z partials from the finest post pass's tile
// partials: entry b = the sum of tiles [b per, (b + 1) per) in tile order.
__global__ void __launch_bounds__(256) k_rz_fold(const double *__restrict__ tp, int nt, int per,
                                                 double *__restrict__ pout, int nblk, const double *scal, int f0,
                                                 int f1) {
  if (stopped(scal, f0, f1)) return;
  const int b = blockIdx.x * 256 + threadIdx.x;
  if (b >= nblk) return;
  const int t0 = b * per, t1 = min(nt, t0 + per);
  double s = 0.0;
  for (int t = t0; t < t1; t += 8) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = tp[t + u < t1 ? t + u : t0];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t + u < t1) s += v[u];
  }
  pout[b] = s;
}

// The region of tile (tx, ty) lies inside the domain and the arrays allow
// two-value loads (vec).
__device__ __forceinline__ bool inner_tile(int vec, int nx, int ny, int tx, int ty, int i0, int j0) {
  return vec && tx >= 1 && ty >= 1 && i0 + FR_X - 1 <= nx && j0 + FR_Y - 1 <= ny;
}

template <typename T, typename TB>
__global__ void __launch_bounds__(FT_BLK) k_mg_pre(Lev<T> F, const TB *__restrict__ bin, int nxc, int nyc,
                                                   T *__restrict__ bc, T scale, int vec, const double *scal, int f0,
                                                   int f1) {
  __shared__ T xs[FR_N];
  __shared__ T rs[FT_X * FT_Y];
  const int ntx = (F.nx + FT_X - 1) / FT_X;
  const int t = mg_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = t / ntx;
  const int i0 = tx * FT_X + 1 - FT_H, j0 = ty * FT_Y + 1 - FT_H;  // global cell of region cell (0, 0)
  if (inner_tile(vec, F.nx, F.ny, tx, ty, i0, j0))
    pre_pass<true>(F, bin, tx, ty, i0, j0, nxc, nyc, bc, scale, scal, f0, f1, xs, rs);
  else
    pre_pass<false>(F, bin, tx, ty, i0, j0, nxc, nyc, bc, scale, scal, f0, f1, xs, rs);
}

template <typename T, typename TB, typename TX>
__global__ void __launch_bounds__(FT_BLK) k_mg_post(Lev<T> F, const TB *__restrict__ bin, TX *__restrict__ xout,
                                                    int nxc, const T *__restrict__ xc, int vec, const double *scal,
                                                    int f0, int f1, double *rzt) {
  __shared__ T xs[FR_N];
  __shared__ double wsum[FT_BLK / 64];
  const int ntx = (F.nx + FT_X - 1) / FT_X;
  const int t = mg_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % ntx, ty = t / ntx;
  const int i0 = tx * FT_X + 1 - FT_H, j0 = ty * FT_Y + 1 - FT_H;
  if (inner_tile(vec, F.nx, F.ny, tx, ty, i0, j0))
    post_pass<true>(F, bin, xout, i0, j0, nxc, xc, scal, f0, f1, xs, rzt, t, wsum);
  else
    post_pass<false>(F, bin, xout, i0, j0, nxc, xc, scal, f0, f1, xs, rzt, t, wsum);
}

// The fused passes' two-value loads and stores: an even level width and
// 16-byte aligned arrays (the face pitch is a multiple of 64).
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15) == 0; }
template <typename T>
inline int pair_vec(const Lev<T> &L, const void *b, const void *x) {
  return L.nx % 2 == 0 && L.kxp % 2 == 0 && aligned16(L.kx) && aligned16(L.ky) && aligned16(L.diag) &&
         aligned16(b) && (x == nullptr || aligned16(x));
}

inline unsigned fused_tiles(int nx, int ny) {
  return (unsigned)(((nx + FT_X - 1) / FT_X) * ((ny + FT_Y - 1) / FT_Y));
}

// rz (the reduction grid's r.z slot) from the tile partials rzt.
inline void rz_fold(Ctx &c, const double *rzt, double *rz, int f0, int f1) {
  const int nt = (int)fused_tiles(c.m.nx, c.m.ny), nblk = c.ps.nblocks, per = (nt + nblk - 1) / nblk;
  hipLaunchKernelGGL(k_rz_fold, dim3((nblk + 255) / 256), dim3(256), 0, c.stream, rzt, nt, per, rz, nblk,
                     (const double *)c.ps.scal, f0, f1);
}
// rz from r and z after a cycle without the fused finest pass.
inline void rz_unfused(Ctx &c, double *rzt, const double *r, const double *z, double *rz, int f0, int f1) {
  hipLaunchKernelGGL(k_rz_tiles, dim3(fused_tiles(c.m.nx, c.m.ny)), dim3(FT_BLK), 0, c.stream, c.m.nx, c.m.ny, r,
                     z, rzt, (const double *)c.ps.scal, f0, f1);
  rz_fold(c, rzt, rz, f0, f1);
}

// The bottom of the V-cycle in one workgroup: every level from the first one
// with at most MG_BOTTOM_CELLS cells down to the coarsest, with x and the
// right-hand sides in LDS.  Same cycle as the multi-block levels: NU
// red-black sweeps from x = 0, restriction, ..., K red-black sweeps and K in
// the adjoint order on the coarsest level, then prolongation and NU sweeps in
// the adjoint order.  Within a half-sweep every update reads the state before
// the half-sweep (values are staged in registers across a barrier), so odd
// periodic sizes (same-colour wrap neighbours) stay deterministic.
// The levels from the first one of at most MG_WAVE_CELLS cells down are
// cycled by the first wave alone, with their face coefficients in LDS (loaded
// by the whole workgroup at the start) and wave-level ordering in place of
// the workgroup barriers: the coarsest level's 4K half-sweeps then cost no
// barrier.  Same operations in the same order either way.
constexpr int MG_BOTTOM_CELLS = 4096;
constexpr int MG_BOTTOM_LDS = 6144;  // cells of all bottom levels together
// largest level above the coarsest that the bottom takes (build-time tunable)
#ifndef MG_BOTTOM_TAKE
#define MG_BOTTOM_TAKE 4096
#endif
constexpr int MG_WAVE_CELLS = 256;   // a wave level: at most 4 cells per lane
constexpr int MG_WAVE_LDS = 512;     // cells of all wave levels together

struct BotLev {
  int nx, ny;
  long kxp;
  const void *kx, *ky, *diag;
  int off;   // offset of the level's cells in the LDS arrays
  int woff;  // wave levels: offset of its coefficients in the LDS coefficient arrays
};
struct Bottom {
  int nlev;  // levels handled here
  int lw;    // first wave level (nlev: none)
  BotLev lev[MG_MAX_LEVELS];
  int perx, pery, K, NU;
  int project;   // singular operator: the coarsest solve is P S P (see k_mg_bottom)
  double scale;  // restriction weight (R = scale * P^T)
};

// Ordering of the first wave's LDS accesses across its lanes (LDS operations
// of one wave complete in issue order; the fences keep the compiler's order).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <typename T>
__global__ void __launch_bounds__(1024) k_mg_bottom(Bottom B, const T *__restrict__ b_in, T *__restrict__ x_out,
                                                    const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  __shared__ T xs[MG_BOTTOM_LDS];
  __shared__ T bs[MG_BOTTOM_LDS];
  __shared__ T wc[5][MG_WAVE_LDS];  // west, east, south, north, diagonal of the wave levels
  __shared__ T rsd[MG_BOTTOM_CELLS];  // residuals of a workgroup level, for its restriction
  __shared__ T s_mean;
  const int tid = threadIdx.x;
  const int last = B.nlev - 1, lw = B.lw;
  {
    // the first level's right-hand side (at most 4096 cells) and the wave
    // levels' coefficients (one cell of them per thread): all loads issued
    // before any is stored
    const int n0 = B.lev[0].nx * B.lev[0].ny;
    T bl[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = tid + q * 1024;
      bl[q] = b_in[t < n0 ? t : 0];
    }
    // the thread's wave level (a uniform walk over the levels, so each level's
    // record is read with scalar loads) and its cell t there
    bool wave = false;
    int t = 0, nx = B.lev[last].nx;
    long kxp = B.lev[last].kxp;
    const T *pkx = (const T *)B.lev[last].kx, *pky = (const T *)B.lev[last].ky, *pdg = (const T *)B.lev[last].diag;
    for (int l = lw; l <= last; ++l) {
      const BotLev &L = B.lev[l];
      if (tid >= L.woff && tid < L.woff + L.nx * L.ny) {
        wave = true;
        t = tid - L.woff;
        nx = L.nx;
        kxp = L.kxp;
        pkx = (const T *)L.kx;
        pky = (const T *)L.ky;
        pdg = (const T *)L.diag;
      }
    }
    T cf[5];
    {
      const int i = t % nx + 1, j = t / nx + 1;
      const T *kxr = pkx + (long)(j - 1) * kxp;
      cf[0] = kxr[i - 1];
      cf[1] = kxr[i];
      cf[2] = pky[t];
      cf[3] = pky[t + nx];
      cf[4] = pdg[t];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = tid + q * 1024;
      if (t < n0) bs[t] = bl[q];
    }
    if (wave)
#pragma unroll
      for (int q = 0; q < 5; ++q) wc[q][tid] = cf[q];
  }
  // coefficients of this thread's cells (at most 4 per level) and their
  // (i, j) in registers, loaded once per visit of a level: a half-sweep then
  // reads only LDS
  T cw[4], ce[4], cs[4], cn[4], cd[4];
  int ci[4], cj[4];
  auto load_coef = [&](int l) {
    const BotLev &L = B.lev[l];
    const int nx = L.nx, n = nx * L.ny;
    const T *kx = (const T *)L.kx, *ky = (const T *)L.ky, *dg = (const T *)L.diag;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t0 = tid + q * 1024, t = t0 < n ? t0 : 0;  // past the level: cell (1, 1), never stored
      const int i = t % nx + 1, j = t / nx + 1;
      ci[q] = i;
      cj[q] = j;
      const T *kxr = kx + (long)(j - 1) * L.kxp;
      cw[q] = kxr[i - 1];
      ce[q] = kxr[i];
      cs[q] = ky[t];
      cn[q] = ky[t + nx];
      cd[q] = dg[t];
    }
  };
  // The x of the four neighbours of cell t = (i, j) of a level: every read is
  // issued (a neighbour across a non-periodic side reads the wrapped cell,
  // which is in range, and its term is then left out by a select, not a
  // branch, so no LDS read waits inside a branch).
  struct Nb {
    T w, e, s, n;
    bool hw, he, hs, hn;
  };
  auto nbrs = [&](const BotLev &L, const T *x, int t, int i, int j) {
    const int nx = L.nx, ny = L.ny;
    Nb r;
    r.w = x[i > 1 ? t - 1 : t + (nx - 1)];
    r.e = x[i < nx ? t + 1 : t - (nx - 1)];
    r.s = x[j > 1 ? t - nx : t + (ny - 1) * nx];
    r.n = x[j < ny ? t + nx : t - (ny - 1) * nx];
    r.hw = i > 1 || B.perx;
    r.he = i < nx || B.perx;
    r.hs = j > 1 || B.pery;
    r.hn = j < ny || B.pery;
    return r;
  };
  // Gauss-Seidel value of cell t = (i, j) of a level from x (LDS)
  auto gs = [&](const BotLev &L, const T *x, T bv, T w, T e, T so, T no, T d, int t, int i, int j) {
    const Nb v = nbrs(L, x, t, i, j);
    T s = bv;
    const T s1 = s + w * v.w;
    s = v.hw ? s1 : s;
    const T s2 = s + e * v.e;
    s = v.he ? s2 : s;
    const T s3 = s + so * v.s;
    s = v.hs ? s3 : s;
    const T s4 = s + no * v.n;
    s = v.hn ? s4 : s;
    return s / d;
  };
  // residual b - A x of cell t = (i, j) (k_mg_restrict's terms and order)
  auto resid = [&](const BotLev &L, const T *x, T bv, T w, T e, T so, T no, T d, int t, int i, int j) {
    const Nb v = nbrs(L, x, t, i, j);
    T ax = d * x[t];
    const T a1 = ax - w * v.w;
    ax = v.hw ? a1 : ax;
    const T a2 = ax - e * v.e;
    ax = v.he ? a2 : ax;
    const T a3 = ax - so * v.s;
    ax = v.hs ? a3 : ax;
    const T a4 = ax - no * v.n;
    ax = v.hn ? a4 : ax;
    return bv - ax;
  };
  // coarse right-hand side of cell t of level l + 1: scale * sum of the
  // children's residuals; coef(k, q) gives face q of child k
  auto restrict_cell = [&](int l, int t, auto coef) {
    const BotLev &F = B.lev[l], &C = B.lev[l + 1];
    const T *x = xs + F.off, *b = bs + F.off;
    const int nx = F.nx, ny = F.ny;
    const int I = t % C.nx + 1, J = t / C.nx + 1;
    T sum = T(0);
#pragma unroll 1
    for (int c4 = 0; c4 < 4; ++c4) {
      const int i = 2 * I - 1 + (c4 & 1), j = 2 * J - 1 + (c4 >> 1);
      const bool has = i <= nx && j <= ny;
      const int ii = has ? i : 2 * I - 1, jj = has ? j : 2 * J - 1;  // a missing child reads its sibling
      const int k = (jj - 1) * nx + (ii - 1);
      const T r = resid(F, x, b[k], coef(k, 0), coef(k, 1), coef(k, 2), coef(k, 3), coef(k, 4), k, ii, jj);
      const T s1 = sum + r;
      sum = has ? s1 : sum;
    }
    bs[C.off + t] = (T)B.scale * sum;
  };
  // ---- the workgroup's levels
  auto half = [&](int l, int color) {
    const BotLev &L = B.lev[l];
    const int n = L.nx * L.ny;
    T *x = xs + L.off;
    const T *b = bs + L.off;
    T v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q * 1024 >= n) break;
      const int t0 = tid + q * 1024, t = t0 < n ? t0 : 0;
      v[q] = gs(L, x, b[t], cw[q], ce[q], cs[q], cn[q], cd[q], t, ci[q], cj[q]);
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      if (q * 1024 >= n) break;
      const int t = tid + q * 1024;
      if (t < n && ((ci[q] + cj[q]) & 1) == color) x[t] = v[q];
    }
    __syncthreads();
  };
  auto zero = [&](int l) {
    const int n = B.lev[l].nx * B.lev[l].ny;
    for (int t = tid; t < n; t += 1024) xs[B.lev[l].off + t] = T(0);
    __syncthreads();
  };
  const int lblock = lw < last ? lw : last;  // levels [0, lblock) go down and up with the workgroup
  for (int l = 0; l < lblock; ++l) {
    load_coef(l);
    zero(l);
    for (int s = 0; s < B.NU; ++s) {
      half(l, 0);
      half(l, 1);
    }
    // restriction: the residuals of the level's cells from the coefficients
    // in registers (restrict_cell's terms and order), then the children's sums
    const BotLev &F = B.lev[l], &C = B.lev[l + 1];
    {
      const int nx = F.nx, ny = F.ny, n = nx * ny;
      const T *x = xs + F.off, *b = bs + F.off;
      T r[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q * 1024 >= n) break;
        const int t0 = tid + q * 1024, t = t0 < n ? t0 : 0;
        r[q] = resid(F, x, b[t], cw[q], ce[q], cs[q], cn[q], cd[q], t, ci[q], cj[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q * 1024 >= n) break;
        const int t = tid + q * 1024;
        if (t < n) rsd[t] = r[q];
      }
      __syncthreads();
      const int nc = C.nx * C.ny;
      for (int t = tid; t < nc; t += 1024) {
        const int I = t % C.nx + 1, J = t / C.nx + 1;
        T c[4];
        bool has[4];
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
          const int i = 2 * I - 1 + (c4 & 1), j = 2 * J - 1 + (c4 >> 1);
          has[c4] = i <= nx && j <= ny;
          c[c4] = rsd[has[c4] ? (j - 1) * nx + (i - 1) : 0];
        }
        T sum = T(0);
#pragma unroll
        for (int c4 = 0; c4 < 4; ++c4) {
          const T s1 = sum + c[c4];
          sum = has[c4] ? s1 : sum;
        }
        bs[C.off + t] = (T)B.scale * sum;
      }
      __syncthreads();
    }
  }
  // With a singular operator (no Dirichlet pressure side: the constant is in
  // its null space up to the assembly's 1e-12 dvol shift) the coarsest solve
  // is P S P, S its sweeps and P the removal of the mean (a serial sum in
  // cell order): the sweeps would otherwise grow the constant mode of the
  // coarsest right-hand side's mean without bound, and the prolonged
  // constant, invisible to A, would cost the fine levels' stencils their
  // digits by cancellation.  P S P keeps the cycle symmetric.
  auto mean_free = [&](T *a, int n, bool wave) {
    if ((wave ? threadIdx.x : tid) == 0) {
      T sum = T(0);
      for (int t = 0; t < n; ++t) sum = sum + a[t];
      s_mean = sum / (T)n;
    }
    if (wave) wave_sync();
    else __syncthreads();
    const T mu = s_mean;
    for (int t = wave ? threadIdx.x : tid; t < n; t += wave ? 64 : 1024) a[t] = a[t] - mu;
    if (wave) wave_sync();
    else __syncthreads();
  };
  if (lw > last) {
    // no wave level: the coarsest level with the workgroup
    const BotLev &L = B.lev[last];
    if (B.project) mean_free(bs + L.off, L.nx * L.ny, false);
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
    if (B.project) mean_free(xs + L.off, L.nx * L.ny, false);
  } else if (tid < 64) {
    // ---- the wave levels, first wave only
    const int lane = tid;
    // (i, j) of the lane's cells of the wave level being visited
    int wi[4], wj[4];
    auto wave_cells = [&](int l) {
      const int nx = B.lev[l].nx, n = nx * B.lev[l].ny;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int t0 = lane + q * 64, t = t0 < n ? t0 : 0;
        wi[q] = t % nx + 1;
        wj[q] = t / nx + 1;
      }
    };
    auto half_w = [&](int l, int color) {
      const BotLev &L = B.lev[l];
      const int n = L.nx * L.ny, o = L.woff;
      T *x = xs + L.off;
      const T *b = bs + L.off;
      T v[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q * 64 >= n) break;
        const int t0 = lane + q * 64, t = t0 < n ? t0 : 0;
        v[q] = gs(L, x, b[t], wc[0][o + t], wc[1][o + t], wc[2][o + t], wc[3][o + t], wc[4][o + t], t, wi[q],
                  wj[q]);
      }
      wave_sync();
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        if (q * 64 >= n) break;
        const int t = lane + q * 64;
        if (t < n && ((wi[q] + wj[q]) & 1) == color) x[t] = v[q];
      }
      wave_sync();
    };
    auto zero_w = [&](int l) {
      const int n = B.lev[l].nx * B.lev[l].ny;
      for (int t = lane; t < n; t += 64) xs[B.lev[l].off + t] = T(0);
      wave_sync();
    };
    for (int l = lw; l < last; ++l) {
      wave_cells(l);
      zero_w(l);
      for (int s = 0; s < B.NU; ++s) {
        half_w(l, 0);
        half_w(l, 1);
      }
      const int o = B.lev[l].woff;
      auto lcoef = [&](int k, int q) -> T { return wc[q][o + k]; };
      const int nc = B.lev[l + 1].nx * B.lev[l + 1].ny;
      for (int t = lane; t < nc; t += 64) restrict_cell(l, t, lcoef);
      wave_sync();
    }
    if (B.lev[last].nx * B.lev[last].ny <= 64) {
      // the coarsest level, one cell per lane: its operands, neighbour
      // indices (wrapped, in range) and which terms it has in registers for
      // the 4K half-sweeps
      const BotLev &L = B.lev[last];
      const int nx = L.nx, ny = L.ny, n = nx * ny, o = L.woff;
      T *x = xs + L.off;
      if (B.project) mean_free(bs + L.off, n, true);
      const bool live = lane < n;
      const int t = live ? lane : 0;
      const int i = t % nx + 1, j = t / nx + 1;
      const int col = (i + j) & 1;
      const int iw = i > 1 ? t - 1 : t + (nx - 1), ie = i < nx ? t + 1 : t - (nx - 1);
      const int is = j > 1 ? t - nx : t + (ny - 1) * nx, in = j < ny ? t + nx : t - (ny - 1) * nx;
      const bool hw = i > 1 || B.perx, he = i < nx || B.perx, hs = j > 1 || B.pery, hn = j < ny || B.pery;
      const T bv = bs[L.off + t], w = wc[0][o + t], e = wc[1][o + t], so = wc[2][o + t], no = wc[3][o + t],
              d = wc[4][o + t];
      if (live) x[t] = T(0);
      wave_sync();
      auto half_c = [&](int color) {
        const T xw = x[iw], xe = x[ie], xs_ = x[is], xn = x[in];
        T s = bv;
        const T s1 = s + w * xw;
        s = hw ? s1 : s;
        const T s2 = s + e * xe;
        s = he ? s2 : s;
        const T s3 = s + so * xs_;
        s = hs ? s3 : s;
        const T s4 = s + no * xn;
        s = hn ? s4 : s;
        const T v = s / d;
        wave_sync();
        if (live && col == color) x[t] = v;
        wave_sync();
      };
      for (int s = 0; s < B.K; ++s) {
        half_c(0);
        half_c(1);
      }
      for (int s = 0; s < B.K; ++s) {
        half_c(1);
        half_c(0);
      }
      if (B.project) mean_free(x, n, true);
    } else {
      const BotLev &L = B.lev[last];
      if (B.project) mean_free(bs + L.off, L.nx * L.ny, true);
      wave_cells(last);
      zero_w(last);
      for (int s = 0; s < B.K; ++s) {
        half_w(last, 0);
        half_w(last, 1);
      }
      for (int s = 0; s < B.K; ++s) {
        half_w(last, 1);
        half_w(last, 0);
      }
      if (B.project) mean_free(xs + L.off, L.nx * L.ny, true);
    }
    for (int l = last - 1; l >= lw; --l) {
      const BotLev &F = B.lev[l], &C = B.lev[l + 1];
      const int n = F.nx * F.ny;
      wave_cells(l);
#pragma unroll
      for (int q = 0; q < 4; ++q) {  // prolongation: x += x of the coarse cell
        const int t = lane + q * 64;
        if (t < n) xs[F.off + t] += xs[C.off + ((wj[q] + 1) / 2 - 1) * C.nx + ((wi[q] + 1) / 2 - 1)];
      }
      wave_sync();
      for (int s = 0; s < B.NU; ++s) {
        half_w(l, 1);
        half_w(l, 0);
      }
    }
  }
  __syncthreads();
  for (int l = lblock - 1; l >= 0; --l) {
    // (FP32: the last level the workgroup went down through still has its
    // coefficients in registers when the wave levels took the coarsest; FP64
    // loads them again rather than keep them across the wave levels)
    if (sizeof(T) == 8 || l < lblock - 1 || lw > last) load_coef(l);
    const BotLev &F = B.lev[l], &C = B.lev[l + 1];
    const int n = F.nx * F.ny;
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // prolongation: x += x of the coarse cell
      const int t = tid + q * 1024;
      if (t < n) xs[F.off + t] += xs[C.off + ((cj[q] + 1) / 2 - 1) * C.nx + ((ci[q] + 1) / 2 - 1)];
    }
    __syncthreads();
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

// No side with a Dirichlet pressure (BC_NEUMANN): the operator is singular up
// to the assembly's 1e-12 dvol shift, its null space the constant.
inline bool singular_operator(const Mesh &m) {
  return m.bc[0] != BC_NEUMANN && m.bc[1] != BC_NEUMANN && m.bc[2] != BC_NEUMANN && m.bc[3] != BC_NEUMANN;
}

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
    // r.z partials of the finest level's fused post pass, one per tile
    COVOF_CHECK(hipMalloc(&G.rzt, (size_t)fused_tiles(m.nx, m.ny) * sizeof(double)));
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

// First level of the single-workgroup bottom cycle: as many of the coarsest
// levels as fit its LDS arrays (each above the coarsest at most
// MG_BOTTOM_TAKE cells).
inline int bottom_level(const Mg &G) {
  const int nl = G.nlev;
  int lb = nl - 1;
  long held = (long)G.lev[lb].nx * G.lev[lb].ny;
  while (lb > 0) {
    const long cells = (long)G.lev[lb - 1].nx * G.lev[lb - 1].ny;
    if (cells > MG_BOTTOM_TAKE || held + cells > MG_BOTTOM_LDS) break;
    held += cells;
    --lb;
  }
  return lb;
}

// Whether the V-cycle runs its multi-workgroup levels as the fused passes.
template <typename T>
inline bool fused_cycle(Ctx &c) {
  return smoothing_sweeps(c) == 1 && c.opts.mg_fused && bottom_level(Hier<T>::get(c)) >= 1;
}

// z = M^-1 r with the cycle in precision T; r and z may be of another type
// (TB, TZ: FP64 around an FP32 cycle) only on the fused path, where the first
// level's passes round r on load and widen z on store.
template <typename T, typename TB = T, typename TZ = T>
bool apply(Ctx &c, int kind, const TB *r, TZ *z, int f0, int f1, double *rz = nullptr) {
  constexpr bool same = std::is_same<TB, T>::value && std::is_same<TZ, T>::value;
  const T *r0 = same ? (const T *)r : nullptr;
  T *z0 = same ? (T *)z : nullptr;
  Mg &G = Hier<T>::get(c);
  const int NU = smoothing_sweeps(c);  // red-black sweeps before and after the coarse correction
  if (kind == MG_KIND_SGS_RB) {
    const Lev<T> L = level<T>(c, 0, r0, z0);
    smooth(c, L, 0, true, f0, f1);
    smooth(c, L, 1, false, f0, f1);
    smooth(c, L, 0, false, f0, f1);
    return false;
  }
  const int nl = G.nlev;
  const int lb = bottom_level(G);
  const T scale = (T)(0.5 * correction_weight(c));
  const bool fused = fused_cycle<T>(c);
  bool done_rz = false;
  for (int l = 0; l < lb; ++l) {
    const Lev<T> L = level<T>(c, l, r0, z0);
    const MgLevelBuf &B = G.lev[l + 1];
    if (fused) {
      const dim3 grid(fused_tiles(L.nx, L.ny));
      if (l == 0)
        hipLaunchKernelGGL((k_mg_pre<T, TB>), grid, dim3(FT_BLK), 0, c.stream, L, r, B.nx, B.ny, (T *)B.b, scale,
                           pair_vec(L, r, nullptr), (const double *)c.ps.scal, f0, f1);
      else
        hipLaunchKernelGGL((k_mg_pre<T, T>), grid, dim3(FT_BLK), 0, c.stream, L, L.b, B.nx, B.ny, (T *)B.b, scale,
                           pair_vec(L, L.b, nullptr), (const double *)c.ps.scal, f0, f1);
      continue;
    }
    for (int s = 0; s < NU; ++s) {
      smooth(c, L, 0, s == 0, f0, f1);
      smooth(c, L, 1, false, f0, f1);
    }
    hipLaunchKernelGGL(k_mg_restrict<T>, cell_grid(B.nx, B.ny), dim3(64, 4), 0, c.stream, L, B.nx, B.ny, (T *)B.b,
                       scale, (const double *)c.ps.scal, f0, f1);
  }
  {
    Bottom Bt{};
    Bt.nlev = nl - lb;
    int off = 0;
    for (int l = lb; l < nl; ++l) {
      const Lev<T> L = level<T>(c, l, r0, z0);
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
    // wave levels: the coarsest ones of at most MG_WAVE_CELLS cells each whose
    // coefficients fit MG_WAVE_LDS together
    Bt.lw = Bt.nlev;
    int woff = 0;
    while (Bt.lw > 0) {
      const BotLev &D = Bt.lev[Bt.lw - 1];
      const int cells = D.nx * D.ny;
      if (cells > MG_WAVE_CELLS || woff + cells > MG_WAVE_LDS) break;
      woff += cells;
      --Bt.lw;
    }
    for (int l = Bt.lw, o = 0; l < Bt.nlev; ++l) {
      Bt.lev[l].woff = o;
      o += Bt.lev[l].nx * Bt.lev[l].ny;
    }
    Bt.perx = c.m.perx;
    Bt.pery = c.m.pery;
    const Lev<T> Lc = level<T>(c, nl - 1, r0, z0);
    Bt.K = std::min(64, 2 * std::max(Lc.nx, Lc.ny));
    Bt.NU = NU;
    Bt.project = c.opts.mg_project && singular_operator(c.m);
    Bt.scale = 0.5 * correction_weight(c);
    const Lev<T> Lb = level<T>(c, lb, r0, z0);
    hipLaunchKernelGGL(k_mg_bottom<T>, dim3(1), dim3(1024), 0, c.stream, Bt, Lb.b, Lb.x, (const double *)c.ps.scal,
                       f0, f1);
  }
  for (int l = lb - 1; l >= 0; --l) {
    const Lev<T> L = level<T>(c, l, r0, z0);
    const MgLevelBuf &B = G.lev[l + 1];
    if (fused) {
      const dim3 grid(fused_tiles(L.nx, L.ny));
      // the finest pass also forms r.z when asked (an FP64 z)
      double *rzt = (l == 0 && rz && std::is_same<TZ, double>::value && G.rzt) ? G.rzt : nullptr;
      if (l == 0)
        hipLaunchKernelGGL((k_mg_post<T, TB, TZ>), grid, dim3(FT_BLK), 0, c.stream, L, r, z, B.nx, (const T *)B.x,
                           pair_vec(L, r, z), (const double *)c.ps.scal, f0, f1, rzt);
      else
        hipLaunchKernelGGL((k_mg_post<T, T, T>), grid, dim3(FT_BLK), 0, c.stream, L, L.b, L.x, B.nx,
                           (const T *)B.x, pair_vec(L, L.b, L.x), (const double *)c.ps.scal, f0, f1,
                           (double *)nullptr);
      if (rzt) {
        rz_fold(c, rzt, rz, f0, f1);
        done_rz = true;
      }
      continue;
    }
    hipLaunchKernelGGL(k_mg_prolong_smooth<T>, pair_grid(L.nx, L.ny), dim3(64, 4), 0, c.stream, L, B.nx,
                       (const T *)B.x, (const double *)c.ps.scal, f0, f1);
    smooth(c, L, 0, false, f0, f1);
    for (int s = 1; s < NU; ++s) {
      smooth(c, L, 1, false, f0, f1);
      smooth(c, L, 0, false, f0, f1);
    }
  }
  if (!fused && rz && std::is_same<TZ, double>::value && std::is_same<TB, double>::value && G.rzt) {
    rz_unfused(c, G.rzt, (const double *)r, (const double *)z, rz, f0, f1);
    done_rz = true;
  }
  return done_rz;
}

// Rounding / widening copy (the FP32 cycle of an FP64 solve off the fused path).
template <typename TI, typename TO>
__global__ void __launch_bounds__(256) k_mg_convert(long n, const TI *__restrict__ in, TO *__restrict__ out,
                                                    const double *scal, int f0, int f1) {
  if (stopped(scal, f0, f1)) return;
  for (long k = blockIdx.x * 256L + threadIdx.x; k < n; k += (long)gridDim.x * 256) out[k] = (TO)in[k];
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

bool mg_apply64(Ctx &c, int kind, const double *r, double *z, int f0, int f1, double *rz) {
  return apply<double>(c, kind, r, z, f0, f1, rz);
}
void mg_apply32(Ctx &c, int kind, const float *r, float *z, int f0, int f1) {
  apply<float>(c, kind, r, z, f0, f1);
}
bool mg_apply_fp32_cycle(Ctx &c, const double *r, double *z, int f0, int f1, double *rz) {
  if (fused_cycle<float>(c)) return apply<float, double, double>(c, MG_KIND_VCYCLE, r, z, f0, f1, rz);
  const long n = (long)c.m.nx * c.m.ny;
  const double *scal = (const double *)c.ps.scal;
  hipLaunchKernelGGL((k_mg_convert<double, float>), dim3(1024), dim3(256), 0, c.stream, n, r, c.ps.r32, scal, f0, f1);
  apply<float>(c, MG_KIND_VCYCLE, (const float *)c.ps.r32, c.ps.z32, f0, f1);
  hipLaunchKernelGGL((k_mg_convert<float, double>), dim3(1024), dim3(256), 0, c.stream, n, (const float *)c.ps.z32,
                     z, scal, f0, f1);
  if (!rz || !c.ps.mg32.rzt) return false;
  rz_unfused(c, c.ps.mg32.rzt, r, z, rz, f0, f1);
  return true;
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
    if (G->rzt) (void)hipFree(G->rzt);
    G->rzt = nullptr;
  }
}

}  // namespace covof
