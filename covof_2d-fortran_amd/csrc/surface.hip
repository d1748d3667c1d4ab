// The steps advect_vof runs after the transport, on gfx950: SVOF Helmholtz
// smoothing, diffusion smoothing, interface curvature and the CSF surface
// tension forces (single rank, no markers).
//
// Reference: smooth_svof_helmholtz (src/module_interface_vof.f90:387-476),
// vof_smooth_diffusion (:1416-1461), compute_curvature_and_csf (:627-789),
// exchange_halos_scalar on one rank (src/module_mpi_comm.f90:46-56).
//
// Every value is the reference's expression on the same operands (the library
// is built with -ffp-contract=off), so the results are bit-identical to the
// reference's.  Halo fields use the reference's (0:nx+1, 0:ny+1) layout.
//
//   curvature + CSF   three stencil passes (normals, curvature, face forces);
//                     the halos of the normals and of the curvature, which the
//                     reference zero-initialises and then exchanges, are
//                     written in closed form by the threads of the edge cells.
//   diffusion         one pass per explicit step, ping-ponging between the
//                     output and a work array; a step reads its input's halo
//                     in closed form (periodic image or zero-gradient copy),
//                     the final pass materialises the output's halo.
//   SVOF              the 25 symmetric Gauss-Seidel iterations are a
//                     sequential recurrence: they run on the strip sweeps of
//                     seq_sweep.h (one wave per 64-row strip, each lane
//                     walking a row), so every cell sees exactly the
//                     neighbour states the lexicographic sweep sees.
#include "covof_internal.h"
#include "seq_sweep.h"

namespace covof {

namespace {

struct SMesh {
  int nx, ny;
  double dx, dy;
  int perx, pery;
};

SMesh smesh(const Mesh &m) { return SMesh{m.nx, m.ny, m.dx, m.dy, m.perx, m.pery}; }

__device__ __forceinline__ long HS(const SMesh &m, int i, int j) { return (long)j * (m.nx + 2) + i; }

constexpr int SBX = 64, SBY = 4;

// Halo images of a field the reference zero-initialises and then exchanges
// (exchange_halos_scalar on one rank: the x wrap over all rows, then the y
// wrap over all columns): the thread of interior cell (i, j) with value V
// writes the halo cells that copy it, and the zeros of the physical sides.
__device__ __forceinline__ void exchanged_halo(double *F, const SMesh &m, int i, int j, double V) {
  const int nx = m.nx, ny = m.ny;
  if (m.perx) {
    if (i == nx) F[HS(m, 0, j)] = V;
    if (i == 1) F[HS(m, nx + 1, j)] = V;
  } else {
    if (i == 1) F[HS(m, 0, j)] = 0.0;
    if (i == nx) F[HS(m, nx + 1, j)] = 0.0;
  }
  if (m.pery) {
    if (j == ny) F[HS(m, i, 0)] = V;
    if (j == 1) F[HS(m, i, ny + 1)] = V;
  } else {
    if (j == 1) F[HS(m, i, 0)] = 0.0;
    if (j == ny) F[HS(m, i, ny + 1)] = 0.0;
  }
  if (m.perx && m.pery) {
    if (i == nx && j == ny) F[HS(m, 0, 0)] = V;
    if (i == 1 && j == ny) F[HS(m, nx + 1, 0)] = V;
    if (i == nx && j == 1) F[HS(m, 0, ny + 1)] = V;
    if (i == 1 && j == 1) F[HS(m, nx + 1, ny + 1)] = V;
  } else if (i == 1 && j == 1) {
    F[HS(m, 0, 0)] = 0.0;
    F[HS(m, nx + 1, 0)] = 0.0;
    F[HS(m, 0, ny + 1)] = 0.0;
    F[HS(m, nx + 1, ny + 1)] = 0.0;
  }
}

// 1. Unit normal of grad G by central differences where |grad G| > 1e-3 / dx
//    (:675-696), else 0.
__global__ void __launch_bounds__(SBX *SBY) k_normals(const double *__restrict__ G, SMesh m, double *nxs,
                                                     double *nys) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  const double gx = (G[HS(m, i + 1, j)] - G[HS(m, i - 1, j)]) / (2.0 * m.dx);
  const double gy = (G[HS(m, i, j + 1)] - G[HS(m, i, j - 1)]) / (2.0 * m.dy);
  const double gm = sqrt(gx * gx + gy * gy);
  double ux = 0.0, uy = 0.0;
  if (gm > 1.0e-3 / m.dx) {
    ux = gx / gm;
    uy = gy / gm;
  }
  nxs[HS(m, i, j)] = ux;
  nys[HS(m, i, j)] = uy;
  exchanged_halo(nxs, m, i, j, ux);
  exchanged_halo(nys, m, i, j, uy);
}

// Height-function curvature (no reference counterpart; Cummins, Francois &
// Kothe, Computers & Structures 83 (2005) 425-434): C at (ii, jj) for the
// 3 x 7 stencils, the periodic image or else the nearest interior cell.
__device__ __forceinline__ double hf_c(const double *C, const SMesh &m, int ii, int jj) {
  if (ii < 1) ii = m.perx ? ii + m.nx : 1;
  else if (ii > m.nx) ii = m.perx ? ii - m.nx : m.nx;
  if (jj < 1) jj = m.pery ? jj + m.ny : 1;
  else if (jj > m.ny) jj = m.pery ? jj - m.ny : m.ny;
  return C[HS(m, ii, jj)];
}

// Curvature at (i, j) from three 7-cell column sums of C along the dominant
// direction of the unit normal (nxv, nyv) (towards increasing C):
// kappa = -h'' / (1 + h'^2)^(3/2), the reference's sign (+1/R for a liquid
// drop).  False when a column does not run from empty (C < 1e-6) to full
// (C > 1 - 1e-6) along the normal.
__device__ bool hf_kappa(const double *C, const SMesh &m, int i, int j, double nxv, double nyv, double &kap) {
  const bool vert = fabs(nyv) >= fabs(nxv);
  const double sgn = vert ? nyv : nxv;
  double h[3];
#pragma unroll
  for (int q = -1; q <= 1; ++q) {
    double sum = 0.0;
#pragma unroll
    for (int r = -3; r <= 3; ++r) sum = sum + (vert ? hf_c(C, m, i + q, j + r) : hf_c(C, m, i + r, j + q));
    const double lo = vert ? hf_c(C, m, i + q, j - 3) : hf_c(C, m, i - 3, j + q);
    const double hi = vert ? hf_c(C, m, i + q, j + 3) : hf_c(C, m, i + 3, j + q);
    if (sgn > 0.0 ? (lo > 1.0e-6 || hi < 1.0 - 1.0e-6) : (lo < 1.0 - 1.0e-6 || hi > 1.0e-6)) return false;
    h[q + 1] = sum;
  }
  const double a = vert ? m.dy : m.dx, b = vert ? m.dx : m.dy;  // along the heights / across
  const double hp = (h[2] - h[0]) * a / (2.0 * b);
  const double hpp = (h[2] - 2.0 * h[1] + h[0]) * a / (b * b);
  const double den = 1.0 + hp * hp;
  kap = -hpp / (den * sqrt(den));
  return true;
}

// 2. kappa = -div(n) where n is non-zero (:703-714), written to the curvature
//    output with its exchanged halo (:716, :728).  With hf (C given), the
//    height-function curvature where the heights bracket the interface.
__global__ void __launch_bounds__(SBX *SBY) k_curvature(const double *__restrict__ nxs,
                                                       const double *__restrict__ nys, SMesh m, double *kap,
                                                       const double *__restrict__ hfC, double *__restrict__ hfok) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  double k = 0.0, ok = 0.0;
  const double nxv = nxs[HS(m, i, j)], nyv = nys[HS(m, i, j)];
  if (nxv != 0.0 || nyv != 0.0) {
    k = -((nxs[HS(m, i + 1, j)] - nxs[HS(m, i - 1, j)]) / (2.0 * m.dx) +
          (nys[HS(m, i, j + 1)] - nys[HS(m, i, j - 1)]) / (2.0 * m.dy));
    double khf;
    if (hfC && hf_kappa(hfC, m, i, j, nxv, nyv, khf)) {
      k = khf;
      ok = 1.0;
    }
  }
  kap[HS(m, i, j)] = k;
  if (hfok) hfok[HS(m, i, j)] = ok;
  exchanged_halo(kap, m, i, j, k);
}

// 2b. Height-function cells whose columns do not bracket the interface take
//     the mean of their valid neighbours' height-function curvatures (rows
//     j-1..j+1, i fastest; periodic images, no cells beyond a physical side),
//     else keep -div(n).  Only such cells are written and only valid cells
//     are read, so the pass runs in place.
__global__ void __launch_bounds__(SBX *SBY) k_hf_fill(const double *__restrict__ nxs,
                                                     const double *__restrict__ nys, SMesh m,
                                                     const double *__restrict__ hfok, double *kap) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  if (hfok[HS(m, i, j)] != 0.0 || (nxs[HS(m, i, j)] == 0.0 && nys[HS(m, i, j)] == 0.0)) return;
  double sum = 0.0;
  int cnt = 0;
  for (int b = -1; b <= 1; ++b)
    for (int a = -1; a <= 1; ++a) {
      int ii = i + a, jj = j + b;
      if (ii < 1 || ii > m.nx) {
        if (!m.perx) continue;
        ii = ii < 1 ? ii + m.nx : ii - m.nx;
      }
      if (jj < 1 || jj > m.ny) {
        if (!m.pery) continue;
        jj = jj < 1 ? jj + m.ny : jj - m.ny;
      }
      if (hfok[HS(m, ii, jj)] != 0.0) {
        sum = sum + kap[HS(m, ii, jj)];
        ++cnt;
      }
    }
  if (cnt == 0) return;
  const double k = sum / (double)cnt;
  kap[HS(m, i, j)] = k;
  exchanged_halo(kap, m, i, j, k);
}

// 3. Balanced-force CSF on the faces (:731-787): fx on x-faces (1..nx+1) x
//    (1..ny), fy on y-faces (1..nx) x (1..ny+1); other entries untouched.
__global__ void __launch_bounds__(SBX *SBY) k_csf(const double *__restrict__ kap, const double *__restrict__ C,
                                                 SMesh m, double sigma, double *fx, double *fy) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx + 1 || j > m.ny + 1) return;
  const long k = HS(m, i, j);
  if (j <= m.ny) {
    const double ka = 0.5 * (kap[HS(m, i - 1, j)] + kap[k]);
    const double gc = (C[k] - C[HS(m, i - 1, j)]) / m.dx;
    fx[k] = sigma * ka * gc;
  }
  if (i <= m.nx) {
    const double ka = 0.5 * (kap[HS(m, i, j - 1)] + kap[k]);
    const double gc = (C[k] - C[HS(m, i, j - 1)]) / m.dy;
    fy[k] = sigma * ka * gc;
  }
}

__global__ void k_zero3(long n, double *a, double *b, double *c) {
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x)
    a[t] = b[t] = c[t] = 0.0;
}

// Value of the diffusion state's halo cell next to interior cell (i, j) as
// vof_smooth_diffusion's fill leaves it (exchange, then zero-gradient copies
// on the physical sides): a periodic image or the adjacent interior value.
__device__ __forceinline__ double dif_x(const double *F, const SMesh &m, int i, int j) {
  if (i == 0) return F[HS(m, m.perx ? m.nx : 1, j)];
  if (i == m.nx + 1) return F[HS(m, m.perx ? 1 : m.nx, j)];
  return F[HS(m, i, j)];
}
__device__ __forceinline__ double dif_y(const double *F, const SMesh &m, int i, int j) {
  if (j == 0) return F[HS(m, i, m.pery ? m.ny : 1)];
  if (j == m.ny + 1) return F[HS(m, i, m.pery ? 1 : m.ny)];
  return F[HS(m, i, j)];
}

// One explicit step (:1441-1448): dst interior from src (halo in closed form).
__global__ void __launch_bounds__(SBX *SBY) k_diffusion_step(const double *__restrict__ src, SMesh m, double dtau,
                                                            double *__restrict__ dst) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  const double c = src[HS(m, i, j)];
  const double lx = (dif_x(src, m, i + 1, j) - 2.0 * c + dif_x(src, m, i - 1, j)) / (m.dx * m.dx);
  const double ly = (dif_y(src, m, i, j + 1) - 2.0 * c + dif_y(src, m, i, j - 1)) / (m.dy * m.dy);
  dst[HS(m, i, j)] = c + dtau * (lx + ly);
}

// The output's final halo (:1452-1456) over the (nx+2) x (ny+2) grid.  prev:
// the state before the last step, whose filled halo the corners inherit when
// only y is periodic (the y wrap copies halo column 0 before the x copies
// refresh it).
__global__ void k_diffusion_halo(double *F, const double *__restrict__ prev, SMesh m) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = m.nx, ny = m.ny;
  if (i > nx + 1 || j > ny + 1) return;
  const bool hx = i == 0 || i == nx + 1, hy = j == 0 || j == ny + 1;
  if (!hx && !hy) return;
  double v;
  if (hx && !hy) {
    v = dif_x(F, m, i, j);
  } else if (hy && !hx) {
    v = dif_y(F, m, i, j);
  } else {  // corner
    const int ii = (i == 0) ? (m.perx ? nx : 1) : (m.perx ? 1 : nx);
    const int jj = (j == 0) ? (m.pery ? ny : 1) : (m.pery ? 1 : ny);
    if (!m.perx && m.pery) {
      // y wrap of halo column i (prev's fill: zero-gradient copy of column 1 / nx)
      v = prev[HS(m, i == 0 ? 1 : nx, jj)];
    } else {
      v = F[HS(m, ii, jj)];
    }
  }
  F[HS(m, i, j)] = v;
}

// smooth_svof_helmholtz (:387-476): 25 symmetric Gauss-Seidel iterations of
// (I - D lap) C_s = C from C_s = C, the halo filled (zero-gradient copies on
// the physical sides, then the exchange) before each iteration.  The
// recurrence runs on the strip sweeps of seq_sweep.h (their SVOF mode): per
// cell a record of a_P, a_W, a_E, a_S, a_N and C, the state starting from C.
__global__ void __launch_bounds__(SBX *SBY) k_svof_records(const double *__restrict__ C, SMesh m, double wx,
                                                          double wy, int S, double *__restrict__ rec,
                                                          double *__restrict__ x) {
  const int i = blockIdx.x * SBX + threadIdx.x + 1, j = blockIdx.y * SBY + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  const bool bl = !m.perx, br = !m.perx, bb = !m.pery, bt = !m.pery;
  const double aE = (i == m.nx && br) ? 0.0 : wx;
  const double aW = (i == 1 && bl) ? 0.0 : wx;
  const double aN = (j == m.ny && bt) ? 0.0 : wy;
  const double aS = (j == 1 && bb) ? 0.0 : wy;
  const long q = sk_of(S, i, j);
  double *r = rec + q * 8;
  r[SR_P] = 1.0 + aE + aW + aN + aS;
  r[SR_W] = aW;
  r[SR_E] = aE;
  r[SR_S] = aS;
  r[SR_N] = aN;
  r[SR_R] = C[HS(m, i, j)];
  x[q] = C[HS(m, i, j)];
}

__global__ void __launch_bounds__(64 * SEQ_MAXW) k_svof_sweeps(SMesh m, SeqBlk b) {
  seq_sweeps<SEQ_SVOF>(m.perx != 0, m.pery != 0, b, 25, 1.0);
}

// The end of smooth_svof_helmholtz from the sweeps' state x (skewed), one
// workgroup.
__global__ void __launch_bounds__(1024) k_svof_finish(const double *__restrict__ x, int S, SMesh m, double *Cs) {
  const int nx = m.nx, ny = m.ny;
  const bool bl = !m.perx, br = !m.perx, bb = !m.pery, bt = !m.pery;
  for (long t = threadIdx.x; t < (long)nx * ny; t += blockDim.x) {
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    Cs[HS(m, i, j)] = x[sk_of(S, i, j)];
  }
  __syncthreads();
  // clamp to [0, 1] (:453-458), Neumann copies, the four corners, exchange (:460-474)
  for (long t = threadIdx.x; t < (long)nx * ny; t += blockDim.x) {
    const int i = (int)(t % nx) + 1, j = (int)(t / nx) + 1;
    const double c = Cs[HS(m, i, j)];
    const double tc = 1.0 < c ? 1.0 : c;
    Cs[HS(m, i, j)] = 0.0 > tc ? 0.0 : tc;
  }
  __syncthreads();
  for (int j = 1 + threadIdx.x; j <= ny; j += blockDim.x) {
    if (bl) Cs[HS(m, 0, j)] = Cs[HS(m, 1, j)];
    if (br) Cs[HS(m, nx + 1, j)] = Cs[HS(m, nx, j)];
  }
  for (int i = 1 + threadIdx.x; i <= nx; i += blockDim.x) {
    if (bb) Cs[HS(m, i, 0)] = Cs[HS(m, i, 1)];
    if (bt) Cs[HS(m, i, ny + 1)] = Cs[HS(m, i, ny)];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    Cs[HS(m, 0, 0)] = Cs[HS(m, 1, 1)];
    Cs[HS(m, nx + 1, 0)] = Cs[HS(m, nx, 1)];
    Cs[HS(m, 0, ny + 1)] = Cs[HS(m, 1, ny)];
    Cs[HS(m, nx + 1, ny + 1)] = Cs[HS(m, nx, ny)];
  }
  __syncthreads();
  if (m.perx)
    for (int j = threadIdx.x; j <= ny + 1; j += blockDim.x) {
      Cs[HS(m, 0, j)] = Cs[HS(m, nx, j)];
      Cs[HS(m, nx + 1, j)] = Cs[HS(m, 1, j)];
    }
  __syncthreads();
  if (m.pery)
    for (int i = threadIdx.x; i <= nx + 1; i += blockDim.x) {
      Cs[HS(m, i, 0)] = Cs[HS(m, i, ny)];
      Cs[HS(m, i, ny + 1)] = Cs[HS(m, i, 1)];
    }
}

}  // namespace

// ================================================================= host side

static long halo_size(const Mesh &m) { return (long)(m.nx + 2) * (m.ny + 2); }

// Three halo-size work arrays (the normals and the height-function flags;
// the diffusion ping-pong), made on first use and kept with the context.
int surface_reserve(Ctx &c) {
  if (!c.surf.work) COVOF_CHECK(hipMalloc(&c.surf.work, 3 * halo_size(c.m) * sizeof(double)));
  Surface &F = c.surf;
  if (!F.sv_rec) {
    const long len = seq_len(c.m.nx, c.m.ny);
    COVOF_CHECK(hipMalloc(&F.sv_rec, 8 * len * sizeof(double)));
    COVOF_CHECK(hipMalloc(&F.sv_xa, len * sizeof(double)));
    COVOF_CHECK(hipMalloc(&F.sv_xb, len * sizeof(double)));
    COVOF_CHECK(hipMalloc(&F.sv_h, 4 * (long)(c.m.nx > c.m.ny ? c.m.nx : c.m.ny) * sizeof(double)));
  }
  return COVOF_OK;
}

static int surface_work(Ctx &c, double **w0, double **w1) {
  // the third array: the height-function validity flags of curvature_csf
  if (!c.surf.work) COVOF_CHECK(hipMalloc(&c.surf.work, 3 * halo_size(c.m) * sizeof(double)));
  *w0 = c.surf.work;
  *w1 = c.surf.work + halo_size(c.m);
  return COVOF_OK;
}

void surface_free(Ctx &c) {
  for (double *p : {c.surf.work, c.surf.sv_rec, c.surf.sv_xa, c.surf.sv_xb, c.surf.sv_h}) hipFree(p);
  c.surf = Surface();
}

int curvature_csf(Ctx &c, const double *G, const double *C, int enable, double sigma, double *curv,
                  double *fx, double *fy, int method) {
  const SMesh m = smesh(c.m);
  if (!enable || !(sigma > 0.0)) {
    hipLaunchKernelGGL(k_zero3, dim3(512), dim3(256), 0, c.stream, halo_size(c.m), curv, fx, fy);
    COVOF_CHECK_LAUNCH();
    return COVOF_OK;
  }
  double *nxs, *nys;
  int st = surface_work(c, &nxs, &nys);
  if (st) return st;
  const dim3 blk(SBX, SBY);
  const dim3 g_in((m.nx + SBX - 1) / SBX, (m.ny + SBY - 1) / SBY);
  const dim3 g_face((m.nx + 1 + SBX - 1) / SBX, (m.ny + 1 + SBY - 1) / SBY);
  hipLaunchKernelGGL(k_normals, g_in, blk, 0, c.stream, G, m, nxs, nys);
  const bool hf = method == COVOF_CURVATURE_HEIGHT_FUNCTION;
  double *hfok = hf ? c.surf.work + 2 * halo_size(c.m) : nullptr;
  hipLaunchKernelGGL(k_curvature, g_in, blk, 0, c.stream, (const double *)nxs, (const double *)nys, m, curv,
                     hf ? C : (const double *)nullptr, hfok);
  if (hf)
    hipLaunchKernelGGL(k_hf_fill, g_in, blk, 0, c.stream, (const double *)nxs, (const double *)nys, m,
                       (const double *)hfok, curv);
  hipLaunchKernelGGL(k_csf, g_face, blk, 0, c.stream, (const double *)curv, C, m, sigma, fx, fy);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int smooth_diffusion(Ctx &c, const double *C, int steps, double *Ct) {
  const SMesh m = smesh(c.m);
  if (steps < 1) steps = 3;
  double *w0, *w1;
  int st = surface_work(c, &w0, &w1);
  if (st) return st;
  const double dx = m.dx, dy = m.dy;
  const double dtau = 0.2 * (dx * dx < dy * dy ? dx * dx : dy * dy);
  const dim3 blk(SBX, SBY);
  const dim3 g_in((m.nx + SBX - 1) / SBX, (m.ny + SBY - 1) / SBY);
  // step s reads the previous state and writes the next; the last one lands in Ct
  const double *src = C;
  const double *prev = C;
  for (int s = 0; s < steps; ++s) {
    double *dst = ((steps - 1 - s) % 2 == 0) ? Ct : w1;
    hipLaunchKernelGGL(k_diffusion_step, g_in, blk, 0, c.stream, src, m, dtau, dst);
    prev = src;
    src = dst;
  }
  hipLaunchKernelGGL(k_diffusion_halo, dim3((m.nx + 2 + 63) / 64, (m.ny + 2 + 3) / 4), dim3(64, 4), 0, c.stream,
                     Ct, prev, m);
  COVOF_CHECK_LAUNCH();
  (void)w0;
  return COVOF_OK;
}

int svof_smooth(Ctx &c, const double *C, double width, double *Cs) {
  const SMesh m = smesh(c.m);
  double k = width;
  if (k <= 0.0) k = 1.0;
  const double D = k * m.dx * m.dy;
  const double wx = D / (m.dx * m.dx), wy = D / (m.dy * m.dy);
  Surface &F = c.surf;
  const int S = m.nx + 63;
  if (!F.sv_rec) {
    const int rs = surface_reserve(c);
    if (rs) return rs;
  }
  const long nh = m.nx > m.ny ? m.nx : m.ny;
  const SeqBlk b{m.nx, m.ny, (m.ny + 63) / 64, S, F.sv_rec, F.sv_xa, F.sv_xb,
                 F.sv_h, F.sv_h + nh, F.sv_h + 2 * nh, F.sv_h + 3 * nh};
  const dim3 blk(SBX, SBY), g_in((m.nx + SBX - 1) / SBX, (m.ny + SBY - 1) / SBY);
  hipLaunchKernelGGL(k_svof_records, g_in, blk, 0, c.stream, C, m, wx, wy, S, F.sv_rec, F.sv_xa);
  hipLaunchKernelGGL(k_svof_sweeps, dim3(1), dim3(64 * seq_waves(m.ny)), 0, c.stream, m, b);
  hipLaunchKernelGGL(k_svof_finish, dim3(1), dim3(1024), 0, c.stream, (const double *)F.sv_xa, S, m, Cs);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
