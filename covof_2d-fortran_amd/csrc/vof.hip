// Weymouth-Yue VOF advection with PLIC / ELVIRA reconstruction on gfx950.
//
// Reference: advect_vof (src/module_interface_vof.f90:792-825),
// advect_vof_weymouth_yue (:1162-1377), reconstruct_interface and the PLIC /
// ELVIRA reconstructions (src/module_interface_reconstruction.f90:283-457,
// :580-683), exchange_halos_scalar / apply_vof_bc on one rank
// (src/module_mpi_comm.f90:46-56, :636-683), update_physical_properties
// (src/module_interface_vof.f90:479-616).
//
// Pipeline of one directional pass (src -> dst, both halo layout):
//   k_recon_list   reconstruction of the MIXED interior cells only, taken from
//                  a compacted list (full/empty cells have the fixed segment
//                  n = (1,0), alpha = +-1e9 and are never stored).  ELVIRA
//                  spreads the 7 candidates of a cell over 8 lanes.
//   k_ghost        one workgroup: ghost-normal exchange / outward copy and
//                  re-inversion of alpha in mixed ghost cells.
//   k_flux_update  per interior cell: the two face fluxes (cut_box_area of the
//                  donor segment over the swept box) and the conservative
//                  update with the frozen dilation term; writes dst, the
//                  1-byte dilation indicator (first pass) and appends mixed
//                  cells of dst to the list of the next pass.
//   k_halo         one workgroup: dst's halo ring = src's ring, then the
//                  exchange / VOF-BC sequence, then (two-phase, after the
//                  second pass) the physical-corner fix.
// Segments of mixed cells live in full-size arrays that are written and read
// only at mixed cells, so HBM traffic stays with the streaming C/u/v fields.
#include "covof_internal.h"
#include "geometry.h"

namespace covof {

struct DevMesh {
  int nx, ny;
  double dx, dy;
  int bcl, bcb;        // left / bottom BC types (periodic flux fix, apply_vof_bc)
  int perx, pery;
};

static DevMesh dev_mesh(const Mesh &m) {
  DevMesh d;
  d.nx = m.nx;
  d.ny = m.ny;
  d.dx = m.dx;
  d.dy = m.dy;
  d.bcl = m.bc[0];
  d.bcb = m.bc[2];
  d.perx = m.perx;
  d.pery = m.pery;
  return d;
}

__device__ __forceinline__ long H(const DevMesh &m, int i, int j) { return (long)j * (m.nx + 2) + i; }

// ------------------------------------------------------------ reconstruction

// Youngs normal of an interior cell (src/module_interface_reconstruction.f90:313-320).
__device__ __forceinline__ void youngs(const double *C, const DevMesh &m, int i, int j,
                                       double &mx, double &my) {
  const double c_mm = C[H(m, i - 1, j - 1)], c_0m = C[H(m, i, j - 1)], c_pm = C[H(m, i + 1, j - 1)];
  const double c_m0 = C[H(m, i - 1, j)], c_p0 = C[H(m, i + 1, j)];
  const double c_mp = C[H(m, i - 1, j + 1)], c_0p = C[H(m, i, j + 1)], c_pp = C[H(m, i + 1, j + 1)];
  mx = -(((c_pp + 2.0 * c_p0 + c_pm) - (c_mp + 2.0 * c_m0 + c_mm)) / (8.0 * m.dx));
  my = -(((c_pp + 2.0 * c_0p + c_mp) - (c_pm + 2.0 * c_0m + c_mm)) / (8.0 * m.dy));
}

// PLIC segment of a mixed interior cell (:283-343).
__device__ void plic_mixed(const double *C, const DevMesh &m, int i, int j, double c, double &onx,
                           double &ony, double &oal) {
  double mx, my;
  youngs(C, m, i, j, mx, my);
  const double nrm = hypot_ref(mx, my);
  double pnx = 1.0, pny = 0.0;
  if (nrm > 1.0e-12) {
    pnx = mx / nrm;
    pny = my / nrm;
  }
  onx = pnx;
  ony = pny;
  oal = compute_alpha(face_coord(i, m.dx), face_coord(i + 1, m.dx), face_coord(j, m.dy),
                      face_coord(j + 1, m.dy), pnx, pny, c);
}

// One ELVIRA candidate k (0..6) of a mixed interior cell (:365-440): unit
// normal, alpha and the squared 3x3 volume error.  valid = false when the
// candidate's norm is below 1e-12 (skipped by the reference).
__device__ void elvira_candidate(const double *C, const DevMesh &m, int i, int j, double c, int k,
                                 double &ux, double &uy, double &al, double &err, bool &valid) {
  const double dx = m.dx, dy = m.dy;
#define CC(a, b) C[H(m, i + (a), j + (b))]
  double cnx, cny;
  if (k < 3) {
    const double sgn_y = (CC(0, 1) < CC(0, -1)) ? 1.0 : -1.0;
    const double H_m1 = (CC(-1, -1) + CC(-1, 0) + CC(-1, 1)) * dy;
    const double H_0 = (CC(0, -1) + CC(0, 0) + CC(0, 1)) * dy;
    const double H_p1 = (CC(1, -1) + CC(1, 0) + CC(1, 1)) * dy;
    if (k == 0) cnx = -(sgn_y * (H_p1 - H_m1) / (2.0 * dx));
    else if (k == 1) cnx = -(sgn_y * (H_p1 - H_0) / dx);
    else cnx = -(sgn_y * (H_0 - H_m1) / dx);
    cny = sgn_y * 1.0;
  } else if (k < 6) {
    const double sgn_x = (CC(1, 0) < CC(-1, 0)) ? 1.0 : -1.0;
    const double W_m1 = (CC(-1, -1) + CC(0, -1) + CC(1, -1)) * dx;
    const double W_0 = (CC(-1, 0) + CC(0, 0) + CC(1, 0)) * dx;
    const double W_p1 = (CC(-1, 1) + CC(0, 1) + CC(1, 1)) * dx;
    cnx = sgn_x * 1.0;
    if (k == 3) cny = -(sgn_x * (W_p1 - W_m1) / (2.0 * dy));
    else if (k == 4) cny = -(sgn_x * (W_p1 - W_0) / dy);
    else cny = -(sgn_x * (W_0 - W_m1) / dy);
  } else {
    cnx = -(((CC(1, 1) + 2.0 * CC(1, 0) + CC(1, -1)) - (CC(-1, 1) + 2.0 * CC(-1, 0) + CC(-1, -1))) / (8.0 * dx));
    cny = -(((CC(1, 1) + 2.0 * CC(0, 1) + CC(-1, 1)) - (CC(1, -1) + 2.0 * CC(0, -1) + CC(-1, -1))) / (8.0 * dy));
  }
  const double cn = hypot_ref(cnx, cny);
  valid = !(cn < 1.0e-12);
  ux = 1.0;
  uy = 0.0;
  al = 0.0;
  err = 0.0;
  if (!valid) return;
  ux = cnx / cn;
  uy = cny / cn;
  const double xmin = face_coord(i, dx), xmax = face_coord(i + 1, dx);
  const double ymin = face_coord(j, dy), ymax = face_coord(j + 1, dy);
  al = compute_alpha(xmin, xmax, ymin, ymax, ux, uy, c);
  double e = 0.0;
  for (int dj = -1; dj <= 1; ++dj)
    for (int di = -1; di <= 1; ++di) {
      if (di == 0 && dj == 0) continue;
      const double vt = CC(di, dj) * dx * dy;
      const double vr = cut_box_area(face_coord(i + di, dx), face_coord(i + di + 1, dx),
                                     face_coord(j + dj, dy), face_coord(j + dj + 1, dy), ux, uy, al);
      const double df = vr - vt;
      e = e + df * df;
    }
  err = e;
#undef CC
}

// PLIC over the list: one lane per mixed cell.
__global__ void k_recon_plic(const double *__restrict__ C, DevMesh m, const int *__restrict__ list,
                             const int *__restrict__ count, int slot, int *__restrict__ count_clear,
                             double *snx, double *sny, double *sal) {
  const int n = count[slot];
  if (blockIdx.x == 0 && threadIdx.x == 0) count_clear[0] = 0;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
    const int k = list[t];
    const int j = k / (m.nx + 2), i = k - j * (m.nx + 2);
    double a, b, c;
    plic_mixed(C, m, i, j, C[k], a, b, c);
    snx[k] = a;
    sny[k] = b;
    sal[k] = c;
  }
}

// ELVIRA over the list: 8 lanes per mixed cell, lane l < 7 evaluates
// candidate l; the winner is the first candidate reaching the smallest error
// below 1e30 (the reference's strict '<' scan from best_err = 1e30).
__global__ void k_recon_elvira(const double *__restrict__ C, DevMesh m, const int *__restrict__ list,
                               const int *__restrict__ count, int slot, int *__restrict__ count_clear,
                               double *snx, double *sny, double *sal) {
  const int n = count[slot];
  if (blockIdx.x == 0 && threadIdx.x == 0) count_clear[0] = 0;
  const int lane8 = threadIdx.x & 7;
  const int groups_per_grid = (gridDim.x * blockDim.x) >> 3;
  for (int g = (blockIdx.x * blockDim.x + threadIdx.x) >> 3; g - ((threadIdx.x & 63) >> 3) < n;
       g += groups_per_grid) {
    // every lane of the wave stays in the loop while any group of the wave has work
    const bool active = g < n;
    int k = 0, i = 0, j = 0;
    double ux = 1.0, uy = 0.0, al = 0.0, err = 0.0;
    bool valid = false;
    if (active) {
      k = list[g];
      j = k / (m.nx + 2);
      i = k - j * (m.nx + 2);
      if (lane8 < 7) elvira_candidate(C, m, i, j, C[k], lane8, ux, uy, al, err, valid);
    }
    // key: error (or +inf if not a contender), ties by candidate index
    double key = (valid && err < 1.0e30) ? err : __builtin_inf();
    int idx = lane8;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {
      const double ok = __shfl_xor(key, off, 8);
      const int oi = __shfl_xor(idx, off, 8);
      if (ok < key || (ok == key && oi < idx)) {
        key = ok;
        idx = oi;
      }
    }
    const double bx = __shfl(ux, idx, 8), by = __shfl(uy, idx, 8), ba = __shfl(al, idx, 8);
    if (active && lane8 == 0) {
      if (key == __builtin_inf()) {
        snx[k] = 1.0;
        sny[k] = 0.0;
        sal[k] = 0.0;
      } else {
        snx[k] = bx;
        sny[k] = by;
        sal[k] = ba;
      }
    }
  }
}

// ----------------------------------------------------------- ghost segments

// Halo ring of the (0:nx+1, 0:ny+1) grid: rows j = 0 and j = ny+1 (nx+2 each),
// then columns i = 0 and i = nx+1 for j = 1..ny.
__device__ __forceinline__ int ring_len(const DevMesh &m) { return 2 * (m.nx + 2) + 2 * m.ny; }
__device__ __forceinline__ int ring_of(const DevMesh &m, int i, int j) {
  if (j == 0) return i;
  if (j == m.ny + 1) return (m.nx + 2) + i;
  if (i == 0) return 2 * (m.nx + 2) + (j - 1);
  return 2 * (m.nx + 2) + m.ny + (j - 1);
}
__device__ __forceinline__ void ring_cell(const DevMesh &m, int r, int &i, int &j) {
  const int W = m.nx + 2;
  if (r < W) { i = r; j = 0; }
  else if (r < 2 * W) { i = r - W; j = m.ny + 1; }
  else if (r < 2 * W + m.ny) { i = 0; j = r - 2 * W + 1; }
  else { i = m.nx + 1; j = r - 2 * W - m.ny + 1; }
}
__device__ __forceinline__ bool on_ring(const DevMesh &m, int i, int j) {
  return i == 0 || j == 0 || i == m.nx + 1 || j == m.ny + 1;
}

// Normal of an interior cell as stored in the reference's segments array.
__device__ __forceinline__ void interior_normal(const double *C, const DevMesh &m, const double *snx,
                                                const double *sny, int i, int j, double &gx,
                                                double &gy) {
  const long k = H(m, i, j);
  if (is_mixed(C[k])) {
    gx = snx[k];
    gy = sny[k];
  } else {
    gx = 1.0;
    gy = 0.0;
  }
}

// reconstruct_interface ghost handling (:602-681) in one workgroup.  gx/gy
// hold seg_nx/seg_ny on the ring (interior values are read from the segment
// arrays); the phases follow the reference's statement order.
__global__ void __launch_bounds__(1024) k_ghost(const double *__restrict__ C, DevMesh m,
                                                double *snx, double *sny, double *sal,
                                                double *gx, double *gy) {
  const int L = ring_len(m), nx = m.nx, ny = m.ny;
  const int tid = threadIdx.x, nt = blockDim.x;
  auto get = [&](int i, int j, double &a, double &b) {
    if (on_ring(m, i, j)) {
      const int r = ring_of(m, i, j);
      a = gx[r];
      b = gy[r];
    } else {
      interior_normal(C, m, snx, sny, i, j, a, b);
    }
  };
  auto put = [&](int i, int j, double a, double b) {
    const int r = ring_of(m, i, j);
    gx[r] = a;
    gy[r] = b;
  };
  for (int r = tid; r < L; r += nt) { gx[r] = 0.0; gy[r] = 0.0; }
  __syncthreads();
  if (m.perx) {  // exchange_halos_scalar, x
    for (int j = tid; j <= ny + 1; j += nt) { double a, b; get(nx, j, a, b); put(0, j, a, b); }
    __syncthreads();
    for (int j = tid; j <= ny + 1; j += nt) { double a, b; get(1, j, a, b); put(nx + 1, j, a, b); }
    __syncthreads();
  }
  if (m.pery) {  // exchange_halos_scalar, y
    for (int i = tid; i <= nx + 1; i += nt) { double a, b; get(i, ny, a, b); put(i, 0, a, b); }
    __syncthreads();
    for (int i = tid; i <= nx + 1; i += nt) { double a, b; get(i, 1, a, b); put(i, ny + 1, a, b); }
    __syncthreads();
  }
  if (!m.perx) {  // is_boundary_left / right: copy outward
    for (int j = 1 + tid; j <= ny; j += nt) { double a, b; get(1, j, a, b); put(0, j, a, b); }
    __syncthreads();
    for (int j = 1 + tid; j <= ny; j += nt) { double a, b; get(nx, j, a, b); put(nx + 1, j, a, b); }
    __syncthreads();
  }
  if (!m.pery) {  // is_boundary_bottom / top
    for (int i = tid; i <= nx + 1; i += nt) { double a, b; get(i, 1, a, b); put(i, 0, a, b); }
    __syncthreads();
    for (int i = tid; i <= nx + 1; i += nt) { double a, b; get(i, ny, a, b); put(i, ny + 1, a, b); }
    __syncthreads();
  }
  // segments of ghost cells: full / empty keep (1,0,+-1e9) (never stored);
  // mixed ones take the renormalised ring normal and a re-inverted alpha, or
  // the reconstruction's own ghost value n = (1,0) when the norm is tiny.
  for (int r = tid; r < L; r += nt) {
    int i, j;
    ring_cell(m, r, i, j);
    const long k = H(m, i, j);
    const double c = C[k];
    if (!is_mixed(c)) continue;
    const double xmin = face_coord(i, m.dx), xmax = face_coord(i + 1, m.dx);
    const double ymin = face_coord(j, m.dy), ymax = face_coord(j + 1, m.dy);
    double sx = gx[r], sy = gy[r];
    const double nr = hypot_ref(sx, sy);
    if (nr > 1.0e-12) {
      sx = sx / nr;
      sy = sy / nr;
    } else {
      sx = 1.0;
      sy = 0.0;
    }
    snx[k] = sx;
    sny[k] = sy;
    sal[k] = compute_alpha(xmin, xmax, ymin, ymax, sx, sy, c);
  }
}

// ------------------------------------------------------------ flux + update

// Volume through face (i,j) of the DIR pass (:1204-1222 for x, :1249-1267 for y).
template <int DIR>
__device__ __forceinline__ double face_flux(const double *__restrict__ C, const double *__restrict__ vel,
                                            const DevMesh &m, const double *snx, const double *sny,
                                            const double *sal, int i, int j, double dt) {
  const double uf = vel[H(m, i, j)];
  const double del = fabs(uf) * dt;
  if (del <= 1.0e-14) return 0.0;
  int di, dj;
  double xmin, xmax, ymin, ymax;
  if (DIR == 0) {
    const double xp = face_coord(i, m.dx);
    ymin = face_coord(j, m.dy);
    ymax = face_coord(j + 1, m.dy);
    if (uf > 0.0) { di = i - 1; dj = j; xmin = xp - del; xmax = xp; }
    else { di = i; dj = j; xmin = xp; xmax = xp + del; }
  } else {
    const double yp = face_coord(j, m.dy);
    xmin = face_coord(i, m.dx);
    xmax = face_coord(i + 1, m.dx);
    if (uf > 0.0) { di = i; dj = j - 1; ymin = yp - del; ymax = yp; }
    else { di = i; dj = j; ymin = yp; ymax = yp + del; }
  }
  const long kd = H(m, di, dj);
  const double cdn = C[kd];
  double gnx = 1.0, gny = 0.0, gal;
  if (is_mixed(cdn)) {
    gnx = snx[kd];
    gny = sny[kd];
    gal = sal[kd];
  } else {
    gal = (cdn <= 1.0e-6) ? 1.0e9 : -1.0e9;
  }
  const double a = cut_box_area(xmin, xmax, ymin, ymax, gnx, gny, gal);
  return uf > 0.0 ? a : -a;
}

// Wave-aggregated append of cell k to a list.
__device__ __forceinline__ void list_append(bool pred, int k, int *list, int *count) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(count, __popcll(mask));
  base = __shfl(base, leader);
  if (pred) list[base + __popcll(mask & ((1ull << lane) - 1ull))] = k;
}

// DIR 0: x-pass (Eq. 15), DIR 1: y-pass (Eq. 16).  FIRST: first pass of the
// step (dilation indicator from src = C^n, written to cd); later pass reads cd
// and clamps to [0,1] (:1234-1239 / :1276-1283).
template <int DIR, bool FIRST>
__global__ void __launch_bounds__(256) k_flux_update(const double *__restrict__ src, double *__restrict__ dst,
                                                     const double *__restrict__ vel, DevMesh m, double dt,
                                                     const double *__restrict__ snx, const double *__restrict__ sny,
                                                     const double *__restrict__ sal, uint8_t *__restrict__ cd,
                                                     int *__restrict__ list_next, int *__restrict__ count_next) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const bool inside = i <= m.nx && j <= m.ny;
  long k = 0;
  double cn = 0.0;
  if (inside) {
    k = H(m, i, j);
    const double c0 = src[k];
    double flo, fhi, du, h;
    if (DIR == 0) {
      if (m.bcl == BC_PERIODIC && (i == 1 || i == m.nx)) {
        // :1223-1231: the boundary face keeps the flux computed with the real donor
        const double fb = (vel[H(m, 1, j)] >= 0.0) ? face_flux<0>(src, vel, m, snx, sny, sal, m.nx + 1, j, dt)
                                                    : face_flux<0>(src, vel, m, snx, sny, sal, 1, j, dt);
        flo = (i == 1) ? fb : face_flux<0>(src, vel, m, snx, sny, sal, i, j, dt);
        fhi = (i == m.nx) ? fb : face_flux<0>(src, vel, m, snx, sny, sal, i + 1, j, dt);
      } else {
        flo = face_flux<0>(src, vel, m, snx, sny, sal, i, j, dt);
        fhi = face_flux<0>(src, vel, m, snx, sny, sal, i + 1, j, dt);
      }
      du = vel[H(m, i + 1, j)] - vel[k];
      h = m.dx;
    } else {
      if (m.bcb == BC_PERIODIC && (j == 1 || j == m.ny)) {
        const double fb = (vel[H(m, i, 1)] >= 0.0) ? face_flux<1>(src, vel, m, snx, sny, sal, i, m.ny + 1, dt)
                                                    : face_flux<1>(src, vel, m, snx, sny, sal, i, 1, dt);
        flo = (j == 1) ? fb : face_flux<1>(src, vel, m, snx, sny, sal, i, j, dt);
        fhi = (j == m.ny) ? fb : face_flux<1>(src, vel, m, snx, sny, sal, i, j + 1, dt);
      } else {
        flo = face_flux<1>(src, vel, m, snx, sny, sal, i, j, dt);
        fhi = face_flux<1>(src, vel, m, snx, sny, sal, i, j + 1, dt);
      }
      du = vel[H(m, i, j + 1)] - vel[k];
      h = m.dy;
    }
    double cdv;
    if (FIRST) {
      const bool b = c0 >= 0.5;
      cd[k] = b ? 1 : 0;
      cdv = b ? 1.0 : 0.0;
    } else {
      cdv = cd[k] ? 1.0 : 0.0;
    }
    const double area = m.dx * m.dy;
    cn = c0 - (fhi - flo) / area + cdv * (dt / h) * du;
    if (!FIRST) {
      const double t = 1.0 < cn ? 1.0 : cn;  // min(1, c)
      cn = 0.0 > t ? 0.0 : t;                // max(0, .)
    }
    dst[k] = cn;
  }
  list_append(inside && is_mixed(cn), (int)k, list_next, count_next);
}

// Mixed interior cells of C -> list (first pass of a call).
__global__ void k_classify(const double *__restrict__ C, DevMesh m, int *list, int *count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  const bool inside = i <= m.nx && j <= m.ny;
  const long k = inside ? H(m, i, j) : 0;
  list_append(inside && is_mixed(C[k]), (int)k, list, count);
}

__global__ void k_zero_int(int *p, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) p[t] = 0;
}

// ------------------------------------------------------------------ halo ring

// exchange_halos_scalar (src/module_mpi_comm.f90:46-56) then apply_vof_bc
// (:636-683) on one rank, as consecutive phases of one workgroup.
__device__ void halo_sequence(double *F, const DevMesh &m) {
  const int nx = m.nx, ny = m.ny, tid = threadIdx.x, nt = blockDim.x;
  if (m.perx) {
    for (int j = tid; j <= ny + 1; j += nt) F[H(m, 0, j)] = F[H(m, nx, j)];
    __syncthreads();
    for (int j = tid; j <= ny + 1; j += nt) F[H(m, nx + 1, j)] = F[H(m, 1, j)];
    __syncthreads();
  }
  if (m.pery) {
    for (int i = tid; i <= nx + 1; i += nt) F[H(m, i, 0)] = F[H(m, i, ny)];
    __syncthreads();
    for (int i = tid; i <= nx + 1; i += nt) F[H(m, i, ny + 1)] = F[H(m, i, 1)];
    __syncthreads();
  }
  if (m.bcl == BC_PERIODIC) {
    for (int j = 1 + tid; j <= ny; j += nt) F[H(m, 0, j)] = F[H(m, nx, j)];
    __syncthreads();
    for (int j = 1 + tid; j <= ny; j += nt) F[H(m, nx + 1, j)] = F[H(m, 1, j)];
    __syncthreads();
  } else if (!m.perx) {
    for (int j = 1 + tid; j <= ny; j += nt) F[H(m, 0, j)] = F[H(m, 1, j)];
    __syncthreads();
    for (int j = 1 + tid; j <= ny; j += nt) F[H(m, nx + 1, j)] = F[H(m, nx, j)];
    __syncthreads();
  }
  if (m.bcb == BC_PERIODIC) {
    for (int i = 1 + tid; i <= nx; i += nt) F[H(m, i, 0)] = F[H(m, i, ny)];
    __syncthreads();
    for (int i = 1 + tid; i <= nx; i += nt) F[H(m, i, ny + 1)] = F[H(m, i, 1)];
    __syncthreads();
  } else if (!m.pery) {
    for (int i = 1 + tid; i <= nx; i += nt) F[H(m, i, 0)] = F[H(m, i, 1)];
    __syncthreads();
  }
  if (!m.pery) {
    for (int i = 1 + tid; i <= nx; i += nt) F[H(m, i, ny + 1)] = F[H(m, i, ny)];
    __syncthreads();
  }
}

// dst's ring starts as src's (c_star = C, :1233 / c_final = c_star, :1275),
// then one exchange + VOF BC: inside the first pass (:1240-1243) and, after the
// second, in advect_vof (:809-810).  CORNERS: two-phase physical-corner fix of
// update_physical_properties (:532-535).
__global__ void __launch_bounds__(1024) k_halo(const double *__restrict__ src, double *dst, DevMesh m,
                                               int corners) {
  const int L = ring_len(m);
  for (int r = threadIdx.x; r < L; r += blockDim.x) {
    int i, j;
    ring_cell(m, r, i, j);
    dst[H(m, i, j)] = src[H(m, i, j)];
  }
  __syncthreads();
  halo_sequence(dst, m);
  if (corners && threadIdx.x == 0) {
    const int nx = m.nx, ny = m.ny;
    const bool bx = !m.perx, by = !m.pery;
    if (bx && by) {
      dst[H(m, 0, 0)] = dst[H(m, 1, 1)];
      dst[H(m, nx + 1, 0)] = dst[H(m, nx, 1)];
      dst[H(m, 0, ny + 1)] = dst[H(m, 1, ny)];
      dst[H(m, nx + 1, ny + 1)] = dst[H(m, nx, ny)];
    }
  }
}

// ------------------------------------------------------------ properties

// Physical-corner fix of C (src/module_interface_vof.f90:532-535).
__global__ void k_corner_fix(double *C, DevMesh m) {
  if (threadIdx.x != 0) return;
  const int nx = m.nx, ny = m.ny;
  if (!m.perx && !m.pery) {
    C[H(m, 0, 0)] = C[H(m, 1, 1)];
    C[H(m, nx + 1, 0)] = C[H(m, nx, 1)];
    C[H(m, 0, ny + 1)] = C[H(m, 1, ny)];
    C[H(m, nx + 1, ny + 1)] = C[H(m, nx, ny)];
  }
}

__device__ __forceinline__ double clamp01(double c) {
  const double t = 1.0 < c ? 1.0 : c;
  return 0.0 > t ? 0.0 : t;
}

// update_physical_properties, two-phase incompressible (:538-613), C already
// corner-fixed.  One thread per node / cell of the (0:nx+1, 0:ny+1) grid.
__global__ void k_properties(const double *__restrict__ C, DevMesh m, double rho1, double rho2,
                             double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                             double *mu, double *mu_node) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = m.nx, ny = m.ny;
  if (i > nx + 1 || j > ny + 1) return;
  auto rho_at = [&](int a, int b) {
    const double c = clamp01(C[H(m, a, b)]);
    return c * rho1 + (1.0 - c) * rho2;
  };
  auto mu_raw = [&](int a, int b) {
    const double c = clamp01(C[H(m, a, b)]);
    return c * mu1 + (1.0 - c) * mu2;
  };
  // mu corners are copied from the diagonal interior cell (:562-565)
  auto mu_at = [&](int a, int b) {
    if (a == 0 && b == 0) return mu_raw(1, 1);
    if (a == nx + 1 && b == 0) return mu_raw(nx, 1);
    if (a == 0 && b == ny + 1) return mu_raw(1, ny);
    if (a == nx + 1 && b == ny + 1) return mu_raw(nx, ny);
    return mu_raw(a, b);
  };
  const long k = H(m, i, j);
  rho[k] = rho_at(i, j);
  mu[k] = mu_at(i, j);
  if (j >= 1 && j <= ny && i >= 1) rho_u[k] = 0.5 * (rho_at(i - 1, j) + rho_at(i, j));
  if (i >= 1 && i <= nx && j >= 1) rho_v[k] = 0.5 * (rho_at(i, j - 1) + rho_at(i, j));
  if (i >= 1 && j >= 1) {
    const double sw = mu_at(i - 1, j - 1), se = mu_at(i, j - 1), nw = mu_at(i - 1, j), ne = mu_at(i, j);
    mu_node[k] = 4.0 / (1.0 / sw + 1.0 / se + 1.0 / nw + 1.0 / ne);
  }
}

__global__ void k_geom_batch(int n, const double *xmin, const double *xmax, const double *ymin,
                             const double *ymax, const double *gnx, const double *gny,
                             const double *alpha, const double *c, double *area, double *aout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  area[t] = cut_box_area(xmin[t], xmax[t], ymin[t], ymax[t], gnx[t], gny[t], alpha[t]);
  aout[t] = compute_alpha(xmin[t], xmax[t], ymin[t], ymax[t], gnx[t], gny[t], c[t]);
}

// Ghost-cell segments of the halo grid for the full reconstruct_interface
// output (test entry point): every halo-grid cell gets its segment, full and
// empty ones the fixed (1,0,+-1e9).
__global__ void k_expand_segments(const double *__restrict__ C, DevMesh m, const double *snx,
                                  const double *sny, const double *sal, double *onx, double *ony,
                                  double *oal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.nx + 1 || j > m.ny + 1) return;
  const long k = H(m, i, j);
  const double c = C[k];
  if (is_mixed(c)) {
    onx[k] = snx[k];
    ony[k] = sny[k];
    oal[k] = sal[k];
  } else {
    onx[k] = 1.0;
    ony[k] = 0.0;
    oal[k] = (c <= 1.0e-6) ? 1.0e9 : -1.0e9;
  }
}

// ================================================================= host side

int vof_alloc(Ctx &c) {
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const long L = 2L * (c.m.nx + 2) + 2L * c.m.ny;
  COVOF_CHECK(hipMalloc(&c.vof.Cb, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.snx, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.sny, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.sal, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.cd, nh));
  COVOF_CHECK(hipMalloc(&c.vof.list[0], nh * sizeof(int)));
  COVOF_CHECK(hipMalloc(&c.vof.list[1], nh * sizeof(int)));
  COVOF_CHECK(hipMalloc(&c.vof.count, 4 * sizeof(int)));
  COVOF_CHECK(hipMalloc(&c.vof.ring, 2 * L * sizeof(double)));
  COVOF_CHECK(hipMemset(c.vof.count, 0, 4 * sizeof(int)));
  COVOF_CHECK(hipMemset(c.vof.cd, 0, nh));
  return COVOF_OK;
}

void vof_free(Ctx &c) {
  hipFree(c.vof.Cb);
  hipFree(c.vof.snx);
  hipFree(c.vof.sny);
  hipFree(c.vof.sal);
  hipFree(c.vof.cd);
  hipFree(c.vof.list[0]);
  hipFree(c.vof.list[1]);
  hipFree(c.vof.count);
  hipFree(c.vof.ring);
}

static int recon_grid_blocks() { return 2048; }

// Reconstruction of src for one pass: list[slot] holds its mixed interior
// cells; the kernel clears count[1-slot] for the pass's update to append to.
static int launch_recon(Ctx &c, int method, const double *src, int slot) {
  const DevMesh m = dev_mesh(c.m);
  const long L = 2L * (c.m.nx + 2) + 2L * c.m.ny;
  if (method == 1)
    hipLaunchKernelGGL(k_recon_elvira, dim3(recon_grid_blocks()), dim3(256), 0, c.stream, src, m,
                       c.vof.list[slot], c.vof.count, slot, c.vof.count + (1 - slot), c.vof.snx,
                       c.vof.sny, c.vof.sal);
  else
    hipLaunchKernelGGL(k_recon_plic, dim3(recon_grid_blocks()), dim3(256), 0, c.stream, src, m,
                       c.vof.list[slot], c.vof.count, slot, c.vof.count + (1 - slot), c.vof.snx,
                       c.vof.sny, c.vof.sal);
  COVOF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ghost, dim3(1), dim3(1024), 0, c.stream, src, m, c.vof.snx, c.vof.sny,
                     c.vof.sal, c.vof.ring, c.vof.ring + L);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static dim3 cell_grid(const Mesh &m, dim3 blk) {
  return dim3((m.nx + blk.x - 1) / blk.x, (m.ny + blk.y - 1) / blk.y);
}

static int launch_classify(Ctx &c, const double *C, int slot) {
  const DevMesh m = dev_mesh(c.m);
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, c.stream, c.vof.count + slot, 1);
  COVOF_CHECK_LAUNCH();
  const dim3 blk(64, 4);
  hipLaunchKernelGGL(k_classify, cell_grid(c.m, blk), blk, 0, c.stream, C, m, c.vof.list[slot],
                     c.vof.count + slot);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// One directional pass src -> dst.  dir 0 = x, 1 = y.
static int launch_pass(Ctx &c, int method, int dir, bool first, const double *src, double *dst,
                       const double *u, const double *v, double dt, int slot, bool corners) {
  int st = launch_recon(c, method, src, slot);
  if (st) return st;
  const DevMesh m = dev_mesh(c.m);
  const dim3 blk(64, 4);
  const dim3 grd = cell_grid(c.m, blk);
  int *ln = c.vof.list[1 - slot], *cn = c.vof.count + (1 - slot);
  if (dir == 0) {
    if (first)
      hipLaunchKernelGGL((k_flux_update<0, true>), grd, blk, 0, c.stream, src, dst, u, m, dt,
                         c.vof.snx, c.vof.sny, c.vof.sal, c.vof.cd, ln, cn);
    else
      hipLaunchKernelGGL((k_flux_update<0, false>), grd, blk, 0, c.stream, src, dst, u, m, dt,
                         c.vof.snx, c.vof.sny, c.vof.sal, c.vof.cd, ln, cn);
  } else {
    if (first)
      hipLaunchKernelGGL((k_flux_update<1, true>), grd, blk, 0, c.stream, src, dst, v, m, dt,
                         c.vof.snx, c.vof.sny, c.vof.sal, c.vof.cd, ln, cn);
    else
      hipLaunchKernelGGL((k_flux_update<1, false>), grd, blk, 0, c.stream, src, dst, v, m, dt,
                         c.vof.snx, c.vof.sny, c.vof.sal, c.vof.cd, ln, cn);
  }
  COVOF_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_halo, dim3(1), dim3(1024), 0, c.stream, src, dst, m, corners ? 1 : 0);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int vof_corner_fix(Ctx &c, double *C) {
  hipLaunchKernelGGL(k_corner_fix, dim3(1), dim3(64), 0, c.stream, C, dev_mesh(c.m));
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int vof_properties(Ctx &c, const double *C, double rho1, double rho2, double mu1, double mu2,
                   double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node) {
  const DevMesh m = dev_mesh(c.m);
  const dim3 blk(64, 4);
  const dim3 grd((c.m.nx + 2 + 63) / 64, (c.m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_properties, grd, blk, 0, c.stream, C, m, rho1, rho2, mu1, mu2, rho, rho_u,
                     rho_v, mu, mu_node);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// nsteps of advect_vof(WEYMOUTH_YUE) on C in place (device pointers).
int vof_advect(Ctx &c, double *C, const double *u, const double *v, double dt, int method,
               int *step_count, int nsteps, int two_phase, double rho1, double rho2, double mu1,
               double mu2, double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node) {
  int st = launch_classify(c, C, 0);
  if (st) return st;
  int slot = 0;
  for (int s = 0; s < nsteps; ++s) {
    const bool x_first = ((*step_count) % 2) == 0;
    const int d1 = x_first ? 0 : 1, d2 = x_first ? 1 : 0;
    st = launch_pass(c, method, d1, true, C, c.vof.Cb, u, v, dt, slot, false);
    if (st) return st;
    slot = 1 - slot;
    st = launch_pass(c, method, d2, false, c.vof.Cb, C, u, v, dt, slot, two_phase != 0);
    if (st) return st;
    slot = 1 - slot;
    *step_count += 1;
    if (two_phase) {
      st = vof_properties(c, C, rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node);
      if (st) return st;
    }
  }
  return COVOF_OK;
}

int vof_reconstruct_interface(Ctx &c, int method, const double *C, double *onx, double *ony,
                              double *oal) {
  int st = launch_classify(c, C, 0);
  if (st) return st;
  st = launch_recon(c, method, C, 0);
  if (st) return st;
  const DevMesh m = dev_mesh(c.m);
  const dim3 blk(64, 4);
  const dim3 grd((c.m.nx + 2 + 63) / 64, (c.m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_expand_segments, grd, blk, 0, c.stream, C, m, c.vof.snx, c.vof.sny,
                     c.vof.sal, onx, ony, oal);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int vof_geom_batch(Ctx &c, int n, const double *xmin, const double *xmax, const double *ymin,
                   const double *ymax, const double *gnx, const double *gny, const double *alpha,
                   const double *cc, double *area, double *aout) {
  if (n <= 0) return COVOF_OK;
  hipLaunchKernelGGL(k_geom_batch, dim3((n + 255) / 256), dim3(256), 0, c.stream, n, xmin, xmax,
                     ymin, ymax, gnx, gny, alpha, cc, area, aout);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace covof
