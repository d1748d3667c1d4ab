// Weymouth-Yue VOF advection with PLIC / ELVIRA reconstruction on gfx950.
//
// Reference: advect_vof (src/module_interface_vof.f90:792-825),
// advect_vof_weymouth_yue (:1162-1377), reconstruct_interface and the PLIC /
// ELVIRA reconstructions (src/module_interface_reconstruction.f90:283-457,
// :580-683), exchange_halos_scalar / apply_vof_bc on one rank
// (src/module_mpi_comm.f90:46-56, :636-683), update_physical_properties
// (src/module_interface_vof.f90:479-616).
//
// Pipeline of one directional pass (src -> dst, both halo layout), two
// launches:
//   k_recon        one block per 64x16 tile reconstructs the tile's MIXED
//                  cells from the tile's list (full/empty cells have the fixed
//                  segment n = (1,0), alpha = +-1e9 and are never stored);
//                  extra blocks reconstruct the ghost segments of the halo
//                  ring (ghost-normal exchange / outward copy of
//                  reconstruct_interface in closed form, alpha re-inverted in
//                  the ghost box).  ELVIRA spreads the 7 candidates of a cell
//                  over 8 lanes.
//   k_flux_update  per interior cell: the two face fluxes (cut_box_area of the
//                  donor segment over the swept box) and the conservative
//                  update with the frozen dilation term; writes dst, its halo
//                  (exchange + VOF BC in closed form), the 1-byte dilation
//                  indicator (first pass) and the tile's list of mixed cells
//                  of dst, compacted in LDS, for the next reconstruction.
// Segments of mixed cells live in full-size arrays that are written and read
// only at mixed cells, so HBM traffic stays with the streaming C/u/v fields.
#include "covof_internal.h"
#include "geometry.h"

#include <algorithm>
#include <mutex>

namespace covof {

struct DevMesh {
  int nx, ny;
  double dx, dy;
  int bcl, bcb;        // left / bottom BC types (periodic flux fix, apply_vof_bc)
  int perx, pery;
  int jet;             // JET_INLET top profile (non-periodic y only)
  double jet_xc, jet_hw;  // its centre and half width
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
  d.jet = 0;
  d.jet_xc = 0.0;
  d.jet_hw = 0.0;
  return d;
}

// The mesh with the context's top profile (apply_vof_bc, src/module_mpi_comm.f90:
// 664-676: xc_top = jet_center if > 0 else lx/2, half width jet_width/2 if
// jet_width > 0 else 0.04).
static DevMesh dev_mesh(const Ctx &c) {
  DevMesh d = dev_mesh(c.m);
  if (c.jet.on && !c.m.pery) {
    d.jet = 1;
    d.jet_xc = c.jet.center > 0.0 ? c.jet.center : 0.5 * c.m.lx;
    d.jet_hw = c.jet.width > 0.0 ? 0.5 * c.jet.width : 0.04;
  }
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
  oal = compute_alpha_fast(face_coord(i, m.dx), face_coord(i + 1, m.dx), face_coord(j, m.dy),
                      face_coord(j + 1, m.dy), pnx, pny, c);
}

// Neighbour q = 0..7 of a 3x3 block in row-major order without the centre.
__device__ __forceinline__ int nb_di(int q) { return q < 3 ? q - 1 : q == 3 ? -1 : q == 4 ? 1 : q - 6; }
__device__ __forceinline__ int nb_dj(int q) { return q < 3 ? -1 : q < 5 ? 0 : 1; }

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
  al = compute_alpha_fast(xmin, xmax, ymin, ymax, ux, uy, c);
  // cut areas of the 8 neighbours (row-major, centre skipped): the line
  // leaves every corner of most of them on one side -- area 0, or the box's
  // own shoelace sum, which is the clip's with all four corners kept -- and
  // each lane clips only the neighbours it cuts, so a wave runs as many clips
  // as its busiest lane rather than eight.
  double vr[8];
  unsigned cut = 0u;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int di = nb_di(q), dj = nb_dj(q);
    const double x0 = face_coord(i + di, dx), x1 = face_coord(i + di + 1, dx);
    const double y0 = face_coord(j + dj, dy), y1 = face_coord(j + dj + 1, dy);
    const int kin = (int)(ux * x0 + uy * y0 + al <= 0.0) + (int)(ux * x1 + uy * y0 + al <= 0.0) +
                    (int)(ux * x1 + uy * y1 + al <= 0.0) + (int)(ux * x0 + uy * y1 + al <= 0.0);
    vr[q] = kin == 4 ? full_box_area(x0, x1, y0, y1) : 0.0;
    if (kin > 0 && kin < 4) cut |= 1u << q;
  }
  while (cut) {
    const int q = __builtin_ctz(cut);
    cut &= cut - 1u;
    const int di = nb_di(q), dj = nb_dj(q);
    const double a = cut_box_area(face_coord(i + di, dx), face_coord(i + di + 1, dx), face_coord(j + dj, dy),
                                  face_coord(j + dj + 1, dy), ux, uy, al);
#pragma unroll
    for (int r = 0; r < 8; ++r) vr[r] = (q == r) ? a : vr[r];
  }
  double e = 0.0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double vt = CC(nb_di(q), nb_dj(q)) * dx * dy;
    const double df = vr[q] - vt;
    e = e + df * df;
  }
  err = e;
#undef CC
}

// ------------------------------------------------------------------ tiles
//
// Cells are processed in tiles of TX x TY = 64 x 16 (1024 cells, 256 threads,
// 4 rows per thread).  The update kernel of a pass leaves, per tile, the list
// of its mixed cells (block-local compaction through LDS: no global atomics)
// for the next reconstruction, which runs one block per tile.
constexpr int TX = 64, TY = 16, TCELLS = TX * TY, TBLK = 256;
constexpr int TRG = TBLK / TX, TROWS = TY / TRG;  // a thread's rows: its row group g, g + TRG, ...
static_assert(TBLK % TX == 0 && TY % TRG == 0, "tile shape");
// ELVIRA reconstruction: 9 cells of 7 candidate lanes per wave
constexpr int ELV_PER_WAVE = 9;

__device__ __forceinline__ int tiles_x(const DevMesh &m) { return (m.nx + TX - 1) / TX; }
__device__ __forceinline__ int tiles_y(const DevMesh &m) { return (m.ny + TY - 1) / TY; }

// Appends cell k (if pred) to the block's LDS list; returns nothing.
__device__ __forceinline__ void lds_append(bool pred, int k, int *s_list, int *s_count) {
  const unsigned long long mask = __ballot(pred);
  if (mask == 0ull) return;
  const int lane = threadIdx.x & 63;
  const int leader = __ffsll((long long)mask) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(s_count, __popcll(mask));
  base = __shfl(base, leader);
  if (pred) s_list[base + __popcll(mask & ((1ull << lane) - 1ull))] = k;
}

// ------------------------------------------------------------ ghost sources
//
// Ghost-cell normals of reconstruct_interface (:602-681): the halo exchange
// of seg_nx / seg_ny followed by the outward copies at physical edges gives
// every ring cell (i,j) the normal of one interior cell, in closed form:
// periodic directions wrap (0 -> n, n+1 -> 1), physical ones clamp (0 -> 1,
// n+1 -> n).  Exception: periodic in y only, the four corners keep the zero
// the halo was initialised with (no copy reaches them).
__device__ __forceinline__ int wrap_or_clamp(int i, int n, int per) {
  if (i < 1) return per ? n : 1;
  if (i > n) return per ? 1 : n;
  return i;
}
__device__ __forceinline__ bool ghost_zero(const DevMesh &m, int i, int j) {
  return !m.perx && m.pery && (i == 0 || i == m.nx + 1) && (j == 0 || j == m.ny + 1);
}

__device__ __forceinline__ int ring_len(const DevMesh &m) { return 2 * (m.nx + 2) + 2 * m.ny; }
static inline int ring_cells(const Mesh &m) { return 2 * (m.nx + 2) + 2 * m.ny; }
__device__ __forceinline__ void ring_cell(const DevMesh &m, int r, int &i, int &j) {
  const int W = m.nx + 2;
  if (r < W) { i = r; j = 0; }
  else if (r < 2 * W) { i = r - W; j = m.ny + 1; }
  else if (r < 2 * W + m.ny) { i = 0; j = r - 2 * W + 1; }
  else { i = m.nx + 1; j = r - 2 * W - m.ny + 1; }
}

// Segment of a mixed ghost cell from the (unnormalised) source normal.
__device__ void ghost_segment(const DevMesh &m, int i, int j, double c, double gx, double gy,
                              double *snx, double *sny, double *sal) {
  const double nr = hypot_ref(gx, gy);
  double sx = 1.0, sy = 0.0;  // the reconstruction's own ghost value when |n| <= 1e-12
  if (nr > 1.0e-12) {
    sx = gx / nr;
    sy = gy / nr;
  }
  const long k = H(m, i, j);
  snx[k] = sx;
  sny[k] = sy;
  sal[k] = compute_alpha_fast(face_coord(i, m.dx), face_coord(i + 1, m.dx), face_coord(j, m.dy),
                         face_coord(j + 1, m.dy), sx, sy, c);
}

// --------------------------------------------------------- reconstruction

// Exclusive prefix sum of the per-tile counts (one workgroup of 1024
// threads): offs[t] = first global index of tile t's mixed cells,
// offs[ntiles] = their total.  Each of the 16 waves owns a contiguous run of
// tiles (a multiple of 64) and reads and writes it 64 consecutive tiles at a
// time: first the run's total, then -- after the 16 totals are scanned
// through LDS -- the run's offsets, each chunk scanned by lane shifts and
// carried to the next.  Integer sums, so any order gives the same offsets.
__global__ void __launch_bounds__(1024) k_tile_scan(const int *__restrict__ count, int ntiles, int *offs) {
  __shared__ int wsum[16];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int per = ((ntiles + 15) / 16 + 63) & ~63;  // tiles per wave
  const int t0 = w * per, t1 = min(t0 + per, ntiles);
  int s = 0;
  for (int base = t0; base < t1; base += 256) {
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = base + q * 64 + lane;
      v[q] = (t < t1) ? count[t] : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) s += v[q];
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) wsum[w] = s;
  __syncthreads();
  if (w == 0) {
    int x = lane < 16 ? wsum[lane] : 0;
#pragma unroll
    for (int off = 1; off < 16; off <<= 1) {
      const int u = __shfl_up(x, off, 64);
      if (lane >= off) x += u;
    }
    if (lane < 16) wsum[lane] = x;  // inclusive totals of the waves
  }
  __syncthreads();
  int carry = w > 0 ? wsum[w - 1] : 0;
  for (int base = t0; base < t1; base += 256) {
    int v[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {  // four chunks' loads first
      const int t = base + q * 64 + lane;
      v[q] = (t < t1) ? count[t] : 0;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int t = base + q * 64 + lane;
      int inc = v[q];
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const int u = __shfl_up(inc, off, 64);
        if (lane >= off) inc += u;
      }
      if (t < t1) offs[t] = carry + inc - v[q];
      carry += __shfl(inc, 63, 64);
    }
  }
  if (threadIdx.x == 1023) offs[ntiles] = wsum[15];
}

// METHOD 0 = PLIC, 1 = ELVIRA.  The work is a list of units, each one wave's
// batch of 64 cells (PLIC, a lane per cell) or 9 (ELVIRA, the 7 candidates of
// a cell on 7 lanes, lane 63 idle): first the halo ring, then the mixed
// interior cells of all tile lists (an interior index maps to its tile by
// binary search in offs).  Wave w of the grid takes units w, w + (waves in
// the grid), ...; the grid has a wave per unit of a typical interface, so the
// hardware dispatcher balances the workgroups over the CUs, and the ring
// comes first because most of its cells are not mixed: its waves end at once
// and free their slots for interior waves instead of queueing behind them.
// (A counter handing out units to a resident grid was slower: its atomics
// are serialised at one address.)  For every mixed ring cell the normal of
// its source cell is reconstructed again (the same arithmetic),
// renormalised and alpha re-inverted in the ghost box.  The kernel is held to
// 128 VGPRs, four waves to a SIMD.  ELVIRA's winner is the first candidate
// reaching the smallest error below 1e30 (the reference's strict '<' scan
// from best_err = 1e30).
template <int METHOD>
__global__ void __launch_bounds__(TBLK, 4) k_recon(const double *__restrict__ C, DevMesh m, const int *__restrict__ list,
                                                   const int *__restrict__ offs, double *snx, double *sny, double *sal) {
  const int ntiles = tiles_x(m) * tiles_y(m);
  const int nring = ring_len(m);
  const int nmix = offs[ntiles];
  constexpr int PER_UNIT = (METHOD == 1) ? ELV_PER_WAVE : 64;
  const int nru = (nring + PER_UNIT - 1) / PER_UNIT;
  const int nunits = nru + (nmix + PER_UNIT - 1) / PER_UNIT;
  const int lane = threadIdx.x & 63;
  const int cand = lane % 7, gbase = lane - cand;  // ELVIRA: candidate and its cell's first lane
  const int slot = (METHOD == 1) ? lane / 7 : lane;
  const bool lane_used = METHOD == 0 || lane < 63;
  const int nwaves = gridDim.x * (TBLK / 64);
  int u = blockIdx.x * (TBLK / 64) + (threadIdx.x >> 6);
  while (u < nunits) {
    const bool ring = u < nru;
    int i = 0, j = 0, si = 0, sj = 0;
    bool work = false, zero = false;
    double c = 0.0;
    if (ring) {
      const int r = PER_UNIT * u + slot;
      if (lane_used && r < nring) {
        ring_cell(m, r, i, j);
        c = C[H(m, i, j)];
        if (is_mixed(c)) {
          work = true;
          zero = ghost_zero(m, i, j);
          si = wrap_or_clamp(i, m.nx, m.perx);
          sj = wrap_or_clamp(j, m.ny, m.pery);
        }
      }
    } else {
      // the tile of the unit's first entry e0: the last tile t with
      // offs[t] <= e0, found by the wave in rounds of 64 probes (two rounds
      // for 4096 tiles) instead of a per-lane bisection's chain of dependent
      // loads; each lane then steps on to its own entry's tile
      const int e0 = PER_UNIT * (u - nru), e = e0 + slot;
      int lo = 0;
      for (int stride = (ntiles + 63) / 64;; stride = (stride + 63) / 64) {
        const int t = min(lo + lane * stride, ntiles);
        const unsigned long long le = __ballot(offs[t] <= e0 && lo + lane * stride <= ntiles);
        lo += (63 - __builtin_clzll(le)) * stride;  // lane 0 always qualifies (offs[lo] <= e0)
        if (stride == 1) break;
      }
      if (lane_used && e < nmix) {
        while (offs[lo + 1] <= e) ++lo;  // offs[lo] <= e < offs[lo + 1]
        const int k = list[(long)lo * TCELLS + (e - offs[lo])];
        j = k / (m.nx + 2);
        i = k - j * (m.nx + 2);
        si = i;
        sj = j;
        c = C[k];
        work = true;
      }
    }
    // normal (and alpha for interior cells) of the source cell (si, sj)
    const double cs = (work && !zero) ? C[H(m, si, sj)] : 0.0;
    const bool src_mixed = work && !zero && is_mixed(cs);
    double bx = 1.0, by = 0.0, ba = 0.0;
    if (METHOD == 0) {
      if (src_mixed) plic_mixed(C, m, si, sj, cs, bx, by, ba);
    } else {
      double ux = 1.0, uy = 0.0, al = 0.0, err = 0.0;
      bool valid = false;
      if (src_mixed) elvira_candidate(C, m, si, sj, cs, cand, ux, uy, al, err, valid);
      const double key = (valid && err < 1.0e30) ? err : __builtin_inf();
      // strict '<' over the candidates in order: the first smallest wins
      double best = __builtin_inf();
      int win = gbase;
#pragma unroll
      for (int q = 0; q < 7; ++q) {
        const double kq = __shfl(key, gbase + q);
        if (kq < best) {
          best = kq;
          win = gbase + q;
        }
      }
      bx = __shfl(ux, win);
      by = __shfl(uy, win);
      ba = __shfl(al, win);
      if (best == __builtin_inf()) {
        bx = 1.0;
        by = 0.0;
        ba = 0.0;
      }
      if (cand != 0) work = false;
    }
    if (work) {
      if (!ring) {
        const long k = H(m, i, j);
        snx[k] = bx;
        sny[k] = by;
        sal[k] = ba;
      } else {
        double gx = bx, gy = by;
        if (zero) {
          gx = 0.0;
          gy = 0.0;
        } else if (!src_mixed) {
          gx = 1.0;
          gy = 0.0;
        }
        ghost_segment(m, i, j, c, gx, gy, snx, sny, sal);
      }
    }
    u += nwaves;
  }
}

// Mixed interior cells of C -> per-tile lists (first pass of a call).
__global__ void __launch_bounds__(TBLK) k_classify(const double *__restrict__ C, DevMesh m, int *list, int *count) {
  __shared__ int s_list[TCELLS];
  __shared__ int s_count;
  if (threadIdx.x == 0) s_count = 0;
  __syncthreads();
  const int tx = blockIdx.x % tiles_x(m), ty = blockIdx.x / tiles_x(m);
  const int i = tx * TX + (threadIdx.x % TX) + 1;
  for (int r = 0; r < TROWS; ++r) {
    const int j = ty * TY + (threadIdx.x / TX) + TRG * r + 1;
    const bool inside = i <= m.nx && j <= m.ny;
    const long k = inside ? H(m, i, j) : 0;
    lds_append(inside && is_mixed(C[k]), (int)k, s_list, &s_count);
  }
  __syncthreads();
  const int n = s_count;
  for (int e = threadIdx.x; e < n; e += TBLK) list[(long)blockIdx.x * TCELLS + e] = s_list[e];
  if (threadIdx.x == 0) count[blockIdx.x] = n;
}

// ------------------------------------------------------------ flux + update

// Swept box and donor cell of face (i,j) of the DIR pass (:1204-1222 for x,
// :1249-1267 for y); zero: |u| dt <= 1e-14 (no flux through the face).
struct FaceBox {
  double xmin, xmax, ymin, ymax;
  long kd;
  bool pos, zero;
};

template <int DIR>
__device__ __forceinline__ FaceBox face_box(double uf, const DevMesh &m, int i, int j, double dt) {
  FaceBox b;
  const double del = fabs(uf) * dt;
  b.zero = del <= 1.0e-14;
  b.pos = uf > 0.0;
  int di, dj;
  if (DIR == 0) {
    const double xp = face_coord(i, m.dx);
    b.ymin = face_coord(j, m.dy);
    b.ymax = face_coord(j + 1, m.dy);
    if (b.pos) { di = i - 1; dj = j; b.xmin = xp - del; b.xmax = xp; }
    else { di = i; dj = j; b.xmin = xp; b.xmax = xp + del; }
  } else {
    const double yp = face_coord(j, m.dy);
    b.xmin = face_coord(i, m.dx);
    b.xmax = face_coord(i + 1, m.dx);
    if (b.pos) { di = i; dj = j - 1; b.ymin = yp - del; b.ymax = yp; }
    else { di = i; dj = j; b.ymin = yp; b.ymax = yp + del; }
  }
  b.kd = H(m, di, dj);
  return b;
}

// Face (i, j) as the update uses it: on a periodic axis the faces 1 and n+1
// both carry the flux computed with the real donor (face n+1 when the
// velocity at face 1 is >= 0, face 1 otherwise; :1223-1231 for x,
// :1268-1275 for y).
template <int DIR>
__device__ __forceinline__ void pass_face(const double *__restrict__ vel, const DevMesh &m, int &i, int &j) {
  if (DIR == 0) {
    if (m.bcl == BC_PERIODIC && (i == 1 || i == m.nx + 1)) i = (vel[H(m, 1, j)] >= 0.0) ? m.nx + 1 : 1;
  } else {
    if (m.bcb == BC_PERIODIC && (j == 1 || j == m.ny + 1)) j = (vel[H(m, i, 1)] >= 0.0) ? m.ny + 1 : 1;
  }
}

// Halo of dst after one exchange_halos_scalar + apply_vof_bc on one rank
// (src/module_mpi_comm.f90:46-56, :636-683), in closed form: the edge halo
// cells copy one interior cell (wrap for periodic directions, zero-gradient
// copy otherwise), which the thread computing that interior value writes.
// The corners depend on the statement order: fully periodic they take the
// diagonally opposite interior cell; periodic in one direction only they take
// a halo value of src (the copy made before the boundary conditions ran);
// with no periodic direction they keep src's value, or with the two-phase
// physical-corner fix of update_physical_properties (:532-535) the diagonal
// interior cell.
__device__ __forceinline__ void write_halo(double *dst, const double *src, const DevMesh &m, int i, int j,
                                           double V, bool corner_fix) {
  const int nx = m.nx, ny = m.ny;
  if (i == (m.perx ? nx : 1)) dst[H(m, 0, j)] = V;
  if (i == (m.perx ? 1 : nx)) dst[H(m, nx + 1, j)] = V;
  if (j == (m.pery ? ny : 1)) dst[H(m, i, 0)] = V;
  if (j == (m.pery ? 1 : ny)) {
    if (m.jet) {  // JET_INLET: liquid (C = 0) inside the injector, C = 1 outside
      const double x_c = ((double)(i - 1) + 0.5) * m.dx;
      dst[H(m, i, ny + 1)] = (fabs(x_c - m.jet_xc) <= m.jet_hw) ? 0.0 : 1.0;
    } else {
      dst[H(m, i, ny + 1)] = V;
    }
  }
  if (m.perx && m.pery) {
    if (i == nx && j == ny) dst[H(m, 0, 0)] = V;
    if (i == 1 && j == ny) dst[H(m, nx + 1, 0)] = V;
    if (i == nx && j == 1) dst[H(m, 0, ny + 1)] = V;
    if (i == 1 && j == 1) dst[H(m, nx + 1, ny + 1)] = V;
  } else if (!m.perx && !m.pery && corner_fix) {
    if (i == 1 && j == 1) dst[H(m, 0, 0)] = V;
    if (i == nx && j == 1) dst[H(m, nx + 1, 0)] = V;
    if (i == 1 && j == ny) dst[H(m, 0, ny + 1)] = V;
    if (i == nx && j == ny) dst[H(m, nx + 1, ny + 1)] = V;
  } else if (i == 1 && j == 1) {
    double c00, c10, c01, c11;  // (0,0), (nx+1,0), (0,ny+1), (nx+1,ny+1)
    if (m.perx && !m.pery) {
      c00 = src[H(m, nx, 0)];
      c10 = src[H(m, 1, 0)];
      c01 = src[H(m, nx, ny + 1)];
      c11 = src[H(m, 1, ny + 1)];
    } else if (!m.perx && m.pery) {
      c00 = src[H(m, 0, ny)];
      c10 = src[H(m, nx + 1, ny)];
      c01 = src[H(m, 0, 1)];
      c11 = src[H(m, nx + 1, 1)];
    } else {
      c00 = src[H(m, 0, 0)];
      c10 = src[H(m, nx + 1, 0)];
      c01 = src[H(m, 0, ny + 1)];
      c11 = src[H(m, nx + 1, ny + 1)];
    }
    dst[H(m, 0, 0)] = c00;
    dst[H(m, nx + 1, 0)] = c10;
    dst[H(m, 0, ny + 1)] = c01;
    dst[H(m, nx + 1, ny + 1)] = c11;
  }
}

// Tile of block b: consecutive workgroups go to the eight XCDs in turn, so
// XCD x is given the contiguous run of tiles [start_x, start_x + count_x)
// (row-major over tiles): the tiles above and below a tile, whose edge rows
// the y-pass reads, mostly sit in the same XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nt) {
  const int q = nt >> 3, rem = nt & 7, x = b & 7;
  return x * q + min(x, rem) + (b >> 3);
}

// One directional pass on a tile.  DIR 0: x-pass (Eq. 15), DIR 1: y-pass
// (Eq. 16).  FIRST: first pass of the step (dilation indicator from src = C^n,
// written to cd); the second pass reads cd and clamps to [0,1]
// (:1234-1239 / :1276-1283).  Writes dst's interior and halo, and the tile's
// list of mixed cells of dst for the next reconstruction.
// Phase 1 evaluates every face of the tile once (65 x 16 x-faces or 64 x 17
// y-faces) into LDS: a face whose donor is empty or
// full has the closed-form value of the clip (0, or the swept box's
// shoelace area), a face whose donor is mixed joins the tile's list, and
// phase 2 clips those with every lane busy; phase 3 updates the cells from
// the two face values each, with the cell's C and face velocities taken from
// the copies phase 1 left in LDS (face (i, j) is cell (i, j)), so the tile is
// read from memory once.
template <int DIR, bool FIRST, bool P2>
__global__ void __launch_bounds__(TBLK) k_flux_update(const double *__restrict__ src, double *__restrict__ dst,
                                                      const double *__restrict__ vel, DevMesh m, double dt,
                                                      const double *__restrict__ snx, const double *__restrict__ sny,
                                                      const double *__restrict__ sal, uint8_t *__restrict__ cd,
                                                      int *__restrict__ list, int *__restrict__ count,
                                                      int corner_fix) {
  constexpr int FW = DIR == 0 ? TX + 1 : TX;  // faces per LDS row
  constexpr int NF = DIR == 0 ? (TX + 1) * TY : TX * (TY + 1);
  // the list of mixed-donor faces is dead once phase 2 has run: the list of
  // mixed cells of dst reuses its memory (29 KiB of LDS per block: five
  // blocks per CU)
  static_assert(NF >= TCELLS, "the face list's memory holds the cell list");
  __shared__ double s_flux[NF];
  __shared__ double s_c[NF], s_u[NF];  // src and vel at each face's own index
  __shared__ int s_mix[NF];
  int *s_list = s_mix;
  __shared__ int s_count, s_nmix;
  if (threadIdx.x == 0) {
    s_count = 0;
    s_nmix = 0;
  }
  __syncthreads();
  const int t = xcd_tile(blockIdx.x, gridDim.x);
  const int tx = t % tiles_x(m), ty = t / tiles_x(m);
  const int i0 = tx * TX + 1, j0 = ty * TY + 1;
  const int imax = DIR == 0 ? m.nx + 1 : m.nx, jmax = DIR == 0 ? m.ny : m.ny + 1;
  const int col = threadIdx.x % TX;
  const int i = i0 + col;
  constexpr int NRC = TROWS;  // rows per thread
  bool cdo[NRC];              // second pass: the dilation indicator of the step
#pragma unroll
  for (int r = 0; r < NRC; ++r) {
    const int j = j0 + (threadIdx.x / TX) + TRG * r;
    cdo[r] = !FIRST && i <= m.nx && j <= m.ny && cd[H(m, i, j)] != 0;
  }
  // all loads of a thread's faces first -- the face velocity and both
  // candidate donors (the velocity's sign picks one) -- so they are in flight
  // together; then the closed forms and the list of mixed-donor faces
  constexpr int NR = (NF + TBLK - 1) / TBLK;
  double uf[NR], cpos[NR], cneg[NR];
  int fis[NR], fjs[NR];
  bool ok[NR];
  // (loads are unconditional, at the face's own index -- a face outside the
  // tile's range reads face (1, 1) and is never used -- so no load waits on
  // another; a periodic edge face, whose flux comes from the real donor's
  // face, is loaded again at that face below)
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int f = r * TBLK + threadIdx.x;
    const int fr = f / FW;
    const int fi = i0 + (f - fr * FW), fj = j0 + fr;
    ok[r] = f < NF && fi <= imax && fj <= jmax;
    const int li = ok[r] ? fi : 1, lj = ok[r] ? fj : 1;
    uf[r] = vel[H(m, li, lj)];
    cpos[r] = DIR == 0 ? src[H(m, li - 1, lj)] : src[H(m, li, lj - 1)];
    cneg[r] = src[H(m, li, lj)];
    fis[r] = fi;
    fjs[r] = fj;
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    int fi = fis[r], fj = fjs[r];
    if (ok[r]) {  // the cell update reads them here: face (i, j) is cell (i, j)
      const int f = r * TBLK + threadIdx.x;
      s_c[f] = cneg[r];
      s_u[f] = uf[r];
    }
    const bool edge = ok[r] && (DIR == 0 ? (m.bcl == BC_PERIODIC && (fi == 1 || fi == m.nx + 1))
                                         : (m.bcb == BC_PERIODIC && (fj == 1 || fj == m.ny + 1)));
    if (edge) {
      pass_face<DIR>(vel, m, fi, fj);
      uf[r] = vel[H(m, fi, fj)];
      cpos[r] = DIR == 0 ? src[H(m, fi - 1, fj)] : src[H(m, fi, fj - 1)];
      cneg[r] = src[H(m, fi, fj)];
      fis[r] = fi;
      fjs[r] = fj;
    }
  }
#pragma unroll
  for (int r = 0; r < NR; ++r) {
    const int f = r * TBLK + threadIdx.x;
    bool mixed = false;
    if (ok[r]) {
      const FaceBox b = face_box<DIR>(uf[r], m, fis[r], fjs[r], dt);
      double fl = 0.0;
      if (!b.zero) {
        const double cdn = b.pos ? cpos[r] : cneg[r];
        if (is_mixed(cdn)) {
          mixed = true;
        } else if (cdn <= 1.0e-6) {
          fl = b.pos ? 0.0 : -0.0;  // empty donor: no vertex inside
        } else {
          const double a = full_box_area(b.xmin, b.xmax, b.ymin, b.ymax);
          fl = b.pos ? a : -a;
        }
      }
      s_flux[f] = fl;
    }
    lds_append(mixed, f, s_mix, &s_nmix);  // every lane of the wave takes part
  }
  __syncthreads();
  const int nmix = s_nmix;
  for (int e = threadIdx.x; e < nmix; e += TBLK) {
    const int f = s_mix[e];
    const int fr = f / FW;
    const int fi0 = i0 + (f - fr * FW), fj0 = j0 + fr;
    int fi = fi0, fj = fj0;
    pass_face<DIR>(vel, m, fi, fj);
    // the face's own velocity is in LDS; only a periodic edge face moved to
    // the real donor's side loads it again
    double u;
    if (fi == fi0 && fj == fj0)
      u = s_u[f];
    else
      u = vel[H(m, fi, fj)];
    const FaceBox b = face_box<DIR>(u, m, fi, fj, dt);
    const double a = cut_box_area(b.xmin, b.xmax, b.ymin, b.ymax, snx[b.kd], sny[b.kd], sal[b.kd]);
    s_flux[f] = b.pos ? a : -a;
  }
  __syncthreads();
  // P2: power-of-two spacings, whose reciprocals are exact: a quotient by
  // them is the product with the reciprocal, the same value
  const double area = m.dx * m.dy;
  const double rarea = P2 ? 1.0 / area : 0.0;
  const double dth = P2 ? dt * (1.0 / (DIR == 0 ? m.dx : m.dy)) : dt / (DIR == 0 ? m.dx : m.dy);
#pragma unroll
  for (int r = 0; r < NRC; ++r) {
    const int row = (threadIdx.x / TX) + TRG * r;
    const int j = j0 + row;
    const bool inside = i <= m.nx && j <= m.ny;
    long k = 0;
    double cn = 0.0;
    if (inside) {
      k = H(m, i, j);
      const int lo = row * FW + col, hi = DIR == 0 ? lo + 1 : lo + FW;
      const double flo = s_flux[lo], fhi = s_flux[hi];
      const double c0 = s_c[lo];
      const double du = s_u[hi] - s_u[lo];
      double cdv;
      if (FIRST) {
        const bool bb = c0 >= 0.5;
        cd[k] = bb ? 1 : 0;
        cdv = bb ? 1.0 : 0.0;
      } else {
        cdv = cdo[r] ? 1.0 : 0.0;
      }
      cn = c0 - (P2 ? (fhi - flo) * rarea : (fhi - flo) / area) + cdv * dth * du;
      if (!FIRST) {
        const double tc = 1.0 < cn ? 1.0 : cn;  // min(1, c)
        cn = 0.0 > tc ? 0.0 : tc;               // max(0, .)
      }
      dst[k] = cn;
      if (i == 1 || j == 1 || i == m.nx || j == m.ny) write_halo(dst, src, m, i, j, cn, corner_fix != 0);
    }
    lds_append(inside && is_mixed(cn), (int)k, s_list, &s_count);
  }
  __syncthreads();
  const int n = s_count;
  for (int e = threadIdx.x; e < n; e += TBLK) list[(long)t * TCELLS + e] = s_list[e];
  if (threadIdx.x == 0) count[t] = n;
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
// Cs non-null: the SVOF branch (:540-545, :596-600), rho and mu from the
// clamped C_s, mu_node harmonically from the clamped node mean of C_s.
__global__ void k_properties(const double *__restrict__ C, const double *__restrict__ Cs, DevMesh m, double rho1,
                             double rho2, double mu1, double mu2, double *rho, double *rho_u, double *rho_v,
                             double *mu, double *mu_node) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = m.nx, ny = m.ny;
  if (i > nx + 1 || j > ny + 1) return;
  const double *F = Cs ? Cs : C;
  auto rho_at = [&](int a, int b) {
    const double c = clamp01(F[H(m, a, b)]);
    return c * rho1 + (1.0 - c) * rho2;
  };
  auto mu_raw = [&](int a, int b) {
    const double c = clamp01(F[H(m, a, b)]);
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
    if (Cs) {
      const double cn = clamp01(0.25 * (Cs[H(m, i - 1, j - 1)] + Cs[H(m, i, j - 1)] + Cs[H(m, i - 1, j)] +
                                        Cs[H(m, i, j)]));
      const double den = mu2 * cn + mu1 * (1.0 - cn);
      mu_node[k] = (mu1 * mu2) / (1.0e-30 > den ? 1.0e-30 : den);
    } else {
      const double sw = mu_at(i - 1, j - 1), se = mu_at(i, j - 1), nw = mu_at(i - 1, j), ne = mu_at(i, j);
      mu_node[k] = 4.0 / (1.0 / sw + 1.0 / se + 1.0 / nw + 1.0 / ne);
    }
  }
}

__global__ void k_geom_batch(int n, const double *xmin, const double *xmax, const double *ymin,
                             const double *ymax, const double *gnx, const double *gny,
                             const double *alpha, const double *c, double *area, double *aout) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  area[t] = cut_box_area(xmin[t], xmax[t], ymin[t], ymax[t], gnx[t], gny[t], alpha[t]);
  aout[t] = compute_alpha_fast(xmin[t], xmax[t], ymin[t], ymax[t], gnx[t], gny[t], c[t]);
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

static int n_tiles(const Mesh &m) { return ((m.nx + TX - 1) / TX) * ((m.ny + TY - 1) / TY); }

int vof_alloc(Ctx &c) {
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const long nt = n_tiles(c.m);
  COVOF_CHECK(hipMalloc(&c.vof.Cb, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.snx, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.sny, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.sal, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&c.vof.cd, nh));
  COVOF_CHECK(hipMalloc(&c.vof.list[0], nt * TCELLS * sizeof(int)));
  COVOF_CHECK(hipMalloc(&c.vof.count, nt * sizeof(int)));
  COVOF_CHECK(hipMalloc(&c.vof.offs, (nt + 1) * sizeof(int)));
  COVOF_CHECK(hipMemset(c.vof.count, 0, nt * sizeof(int)));
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
  hipFree(c.vof.count);
  hipFree(c.vof.offs);
  hipFree(c.vof.adj);
}

// Workgroups of k_recon: a wave per unit of an interface with up to one
// mixed cell in 28 (ELVIRA) or 4 (PLIC) plus the ring's units; waves past the
// units return at once, and larger interfaces take further units in stride.
static dim3 recon_grid(const Mesh &mesh, int method) {
  const int per_unit = method == 1 ? ELV_PER_WAVE : 64;
  const int L = 2 * (mesh.nx + 2) + 2 * mesh.ny;
  const int waves = n_tiles(mesh) * (TBLK / 64) + (L + per_unit - 1) / per_unit;
  return dim3((waves + TBLK / 64 - 1) / (TBLK / 64));
}

// Reconstruction of src: tile lists (left by the previous update or by
// k_classify) plus the halo ring.
static int launch_recon(Ctx &c, int method, const double *src) {
  const DevMesh m = dev_mesh(c.m);
  const int nt = n_tiles(c.m);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, c.stream, (const int *)c.vof.count, nt, c.vof.offs);
  launch_tagged(c, COVOF_KERNEL_VOF_RECON, method == 1 ? k_recon<1> : k_recon<0>, recon_grid(c.m, method),
                dim3(TBLK), src, m, (const int *)c.vof.list[0], (const int *)c.vof.offs, c.vof.snx, c.vof.sny,
                c.vof.sal);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static int launch_classify(Ctx &c, const double *C) {
  hipLaunchKernelGGL(k_classify, dim3(n_tiles(c.m)), dim3(TBLK), 0, c.stream, C, dev_mesh(c.m), c.vof.list[0],
                     c.vof.count);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// One directional pass src -> dst.  dir 0 = x, 1 = y.
static int launch_pass(Ctx &c, int method, int dir, bool first, const double *src, double *dst,
                       const double *u, const double *v, double dt, bool corner_fix) {
  int st = launch_recon(c, method, src);
  if (st) return st;
  const DevMesh m = dev_mesh(c);  // the halo written after the pass sees the top profile
  const dim3 grd(n_tiles(c.m)), blk(TBLK);
  int *ln = c.vof.list[0], *cn = c.vof.count;
  const int cf = corner_fix ? 1 : 0;
  const bool p2 = is_pow2(c.m.dx) && is_pow2(c.m.dy);
  auto kern = dir == 0 ? (first ? (p2 ? k_flux_update<0, true, true> : k_flux_update<0, true, false>)
                                : (p2 ? k_flux_update<0, false, true> : k_flux_update<0, false, false>))
                       : (first ? (p2 ? k_flux_update<1, true, true> : k_flux_update<1, true, false>)
                                : (p2 ? k_flux_update<1, false, true> : k_flux_update<1, false, false>));
  launch_tagged(c, COVOF_KERNEL_VOF_UPDATE, kern, grd, blk, src, dst, dir == 0 ? u : v, m, dt,
                (const double *)c.vof.snx, (const double *)c.vof.sny, (const double *)c.vof.sal, c.vof.cd, ln, cn, cf);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int vof_corner_fix(Ctx &c, double *C) {
  hipLaunchKernelGGL(k_corner_fix, dim3(1), dim3(64), 0, c.stream, C, dev_mesh(c.m));
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int vof_properties(Ctx &c, const double *C, double rho1, double rho2, double mu1, double mu2,
                   double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node, const double *Cs) {
  const DevMesh m = dev_mesh(c.m);
  const dim3 blk(64, 4);
  const dim3 grd((c.m.nx + 2 + 63) / 64, (c.m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_properties, grd, blk, 0, c.stream, C, Cs, m, rho1, rho2, mu1, mu2, rho, rho_u,
                     rho_v, mu, mu_node);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// nsteps of advect_vof(WEYMOUTH_YUE) on C in place (device pointers).
int vof_advect(Ctx &c, double *C, const double *u, const double *v, double dt, int method,
               int *step_count, int nsteps, int two_phase, double rho1, double rho2, double mu1,
               double mu2, double *rho, double *rho_u, double *rho_v, double *mu, double *mu_node) {
  if (nsteps <= 0) return COVOF_OK;
  int st = launch_classify(c, C);
  if (st) return st;
  for (int s = 0; s < nsteps; ++s) {
    const bool x_first = ((*step_count) % 2) == 0;
    const int d1 = x_first ? 0 : 1, d2 = x_first ? 1 : 0;
    st = launch_pass(c, method, d1, true, C, c.vof.Cb, u, v, dt, false);
    if (st) return st;
    st = launch_pass(c, method, d2, false, c.vof.Cb, C, u, v, dt, two_phase != 0);
    if (st) return st;
    *step_count += 1;
    if (two_phase) {
      st = vof_properties(c, C, rho1, rho2, mu1, mu2, rho, rho_u, rho_v, mu, mu_node, nullptr);
      if (st) return st;
    }
  }
  return COVOF_OK;
}

int vof_reconstruct_interface(Ctx &c, int method, const double *C, double *onx, double *ony,
                              double *oal) {
  int st = launch_classify(c, C);
  if (st) return st;
  st = launch_recon(c, method, C);
  if (st) return st;
  const DevMesh m = dev_mesh(c.m);
  const dim3 blk(64, 4);
  const dim3 grd((c.m.nx + 2 + 63) / 64, (c.m.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_expand_segments, grd, blk, 0, c.stream, C, m, c.vof.snx, c.vof.sny,
                     c.vof.sal, onx, ony, oal);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// ------------------------------------------------ allocation-free variant
//
// covof_reconstruct_interface_noalloc_d needs no context: its only scratch
// (the per-tile lists of mixed cells, their sizes and offsets) lives in the
// file-scope array below, one copy per GPU, held until the process ends; the
// segments are written straight into the caller's output arrays.  Nothing in
// it must start at zero (k_classify writes every count and list entry the
// later kernels read, k_tile_scan every offset).  Calls that overlap on one
// GPU share the array and race.
constexpr int NOALLOC_MAX_TILES = 1024;  // e.g. 1024 x 1024, 2048 x 512 cells
constexpr long NOALLOC_INTS = (long)NOALLOC_MAX_TILES * TCELLS + 2L * NOALLOC_MAX_TILES + 1;
static __device__ __align__(256) int scratch[NOALLOC_INTS];  // 4.2 MB

constexpr int NOALLOC_MAX_DEVICES = 64;
static int *g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

// The device address of this GPU's copy of `scratch` (the current device),
// looked up once per device; the name itself is a host-side stand-in.
int noalloc_scratch(int **out) {
  int dev = 0;
  COVOF_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return COVOF_BAD_ARGUMENT;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int *p = nullptr;
    COVOF_CHECK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
    if (!p) return COVOF_HIP_ERROR;
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return COVOF_OK;
}

int noalloc_max_tiles() { return NOALLOC_MAX_TILES; }
int noalloc_tiles(const Mesh &m) { return n_tiles(m); }

// reconstruct_interface on stream s with the tile lists in the scratch array:
// classify, scan, reconstruct into (onx, ony, oal), then fill the non-mixed
// cells in place.
int vof_reconstruct_noalloc(const Mesh &mesh, int method, const double *C, double *onx, double *ony, double *oal,
                            hipStream_t s) {
  const int nt = n_tiles(mesh);
  if (nt > NOALLOC_MAX_TILES) return COVOF_BAD_ARGUMENT;
  int *base = nullptr;
  const int st = noalloc_scratch(&base);
  if (st) return st;
  int *list = base, *count = base + (long)NOALLOC_MAX_TILES * TCELLS, *offs = count + NOALLOC_MAX_TILES;
  const DevMesh m = dev_mesh(mesh);
  hipLaunchKernelGGL(k_classify, dim3(nt), dim3(TBLK), 0, s, C, m, list, count);
  hipLaunchKernelGGL(k_tile_scan, dim3(1), dim3(1024), 0, s, (const int *)count, nt, offs);
  hipLaunchKernelGGL(method == 1 ? k_recon<1> : k_recon<0>, recon_grid(mesh, method), dim3(TBLK), 0, s, C, m,
                     (const int *)list, (const int *)offs, onx, ony, oal);
  const dim3 blk(64, 4), grd((mesh.nx + 2 + 63) / 64, (mesh.ny + 2 + 3) / 4);
  hipLaunchKernelGGL(k_expand_segments, grd, blk, 0, s, C, m, (const double *)onx, (const double *)ony,
                     (const double *)oal, onx, ony, oal);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// ======================================================= adjoint of one step
//
// Reverse mode of one advect_vof(WEYMOUTH_YUE) step C^n -> C^{n+1} on a mesh
// with walls on all four sides or periodic in both directions (no property
// update, no top profile).  The discrete map is differentiated with its
// branch pattern frozen: the dilation indicator c_d, the donor of every face,
// the empty / mixed / full class of every cell, ELVIRA's chosen candidate
// and its sign factors, and the clamp of the second pass (derivative 1 where
// the clamp left the value alone).  compute_alpha's 45-step bisection is
// differentiated as the exact root it approximates (implicit function:
// area(alpha, n) = c |cell|).  For a unit normal n and a box, with L the
// length of the chord n.x + alpha = 0 inside the box and M its midpoint:
//   d area / d alpha = -L,   d area / d n = -L M,
//   d area / d xmin = -h(xmin), d area / d xmax = +h(xmax)
// (h: the covered length of that box edge), so d alpha / d c = -|cell| / L and
// d alpha / d n = -M in a cell.  The normal is m / |m| with m linear in the
// 3x3 stencil (Youngs, or the chosen ELVIRA candidate), a ghost cell's normal
// the renormalised normal of its source cell.

// Chord of n.x + alpha = 0 in the box: length and midpoint (0 if it misses).
__device__ void chord(double xmin, double xmax, double ymin, double ymax, double nx, double ny, double alpha,
                      double &L, double &mx, double &my) {
  const double px[4] = {xmin, xmax, xmax, xmin}, py[4] = {ymin, ymin, ymax, ymax};
  double d[4];
  for (int k = 0; k < 4; ++k) d[k] = nx * px[k] + ny * py[k] + alpha;
  double qx[2] = {0.0, 0.0}, qy[2] = {0.0, 0.0};
  int n = 0;
  for (int k = 0; k < 4 && n < 2; ++k) {
    const int k1 = (k + 1) & 3;
    if (d[k] == 0.0) {
      if (n == 0 || qx[0] != px[k] || qy[0] != py[k]) {
        qx[n] = px[k];
        qy[n] = py[k];
        ++n;
      }
    } else if ((d[k] < 0.0) != (d[k1] < 0.0) && d[k1] != 0.0) {
      const double t = d[k] / (d[k] - d[k1]);
      qx[n] = px[k] + t * (px[k1] - px[k]);
      qy[n] = py[k] + t * (py[k1] - py[k]);
      ++n;
    }
  }
  if (n < 2) {
    L = 0.0;
    mx = my = 0.0;
    return;
  }
  L = sqrt((qx[1] - qx[0]) * (qx[1] - qx[0]) + (qy[1] - qy[0]) * (qy[1] - qy[0]));
  mx = 0.5 * (qx[0] + qx[1]);
  my = 0.5 * (qy[0] + qy[1]);
}

// Covered length of the segment [a0, a1] of the line x = X (vert) or y = Y
// by {nx x + ny y + alpha <= 0}: t ranges over the free coordinate.
__device__ double covered(double fixed, double a0, double a1, double nfix, double nfree, double alpha) {
  const double r = -(nfix * fixed + alpha);  // nfree * t <= r
  if (nfree == 0.0) return r >= 0.0 ? a1 - a0 : 0.0;
  const double t = r / nfree;
  double lo = a0, hi = a1;
  if (nfree > 0.0) hi = t < hi ? t : hi;
  else lo = t > lo ? t : lo;
  return hi > lo ? hi - lo : 0.0;
}

// Flux of face (i, j) of pass DIR as k_flux_update computes it there (the
// same closed forms for empty and full donors).
template <int DIR>
__device__ double adj_face_flux(const double *src, const double *vel, const double *sx, const double *sy,
                                const double *sa, const DevMesh &m, int i, int j, double dt) {
  const FaceBox b = face_box<DIR>(vel[H(m, i, j)], m, i, j, dt);
  if (b.zero) return 0.0;
  const double cdn = src[b.kd];
  double a;
  if (is_mixed(cdn)) a = cut_box_area(b.xmin, b.xmax, b.ymin, b.ymax, sx[b.kd], sy[b.kd], sa[b.kd]);
  else if (cdn <= 1.0e-6) return b.pos ? 0.0 : -0.0;
  else a = full_box_area(b.xmin, b.xmax, b.ymin, b.ymax);
  return b.pos ? a : -a;
}

// g = dJ/dC_out * (clamp derivative) per interior cell, written to G and as
// the direct part of dJ/dC_in.  The halo of dJ/dC_in starts at zero, except
// with corners_through its four corners: dJ/dC_out's (the pass copied src's
// corners into dst).
template <int DIR>
__global__ void k_adj_g(DevMesh m, const double *__restrict__ src, const double *__restrict__ Cn,
                        const double *__restrict__ vel, const double *__restrict__ sx, const double *__restrict__ sy,
                        const double *__restrict__ sa, const double *__restrict__ cbar_out, int clamp,
                        int corners_through, double dt, double *__restrict__ G, double *__restrict__ cbar_in) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.nx + 1 || j > m.ny + 1) return;
  const long k = H(m, i, j);
  if (i < 1 || j < 1 || i > m.nx || j > m.ny) {
    const bool corner = (i == 0 || i == m.nx + 1) && (j == 0 || j == m.ny + 1);
    G[k] = 0.0;
    cbar_in[k] = (corner && corners_through) ? cbar_out[k] : 0.0;
    return;
  }
  double g = cbar_out[k];
  if (clamp) {
    int il = i, jl = j, ih = DIR == 0 ? i + 1 : i, jh = DIR == 0 ? j : j + 1;
    pass_face<DIR>(vel, m, il, jl);
    pass_face<DIR>(vel, m, ih, jh);
    const double flo = adj_face_flux<DIR>(src, vel, sx, sy, sa, m, il, jl, dt);
    const double fhi = adj_face_flux<DIR>(src, vel, sx, sy, sa, m, ih, jh, dt);
    const double area = m.dx * m.dy;
    const double dth = dt / (DIR == 0 ? m.dx : m.dy);
    const double du = (DIR == 0 ? vel[H(m, i + 1, j)] : vel[H(m, i, j + 1)]) - vel[k];
    const double cdv = Cn[k] >= 0.5 ? 1.0 : 0.0;
    const double cpre = src[k] - (fhi - flo) / area + cdv * dth * du;  // k_flux_update's expression
    // a clamp that removes no more than rounding (|overshoot| <= 1e-12, as
    // full and empty cells carry) passes the derivative through: the
    // one-sided derivative from inside [0, 1], the same whatever the
    // rounding of the forward
    if (cpre < -1.0e-12 || cpre > 1.0 + 1.0e-12) g = 0.0;
  }
  G[k] = g;
  cbar_in[k] = g;
}

// Per face of the pass: the flux adjoint (update and periodic copy), the
// velocity adjoint (dilation term and swept-box width, accumulated into vbar)
// and the donor segment's adjoint (fs: d/dnx, d/dny, d/dalpha per face).
template <int DIR>
__global__ void k_adj_face(DevMesh m, const double *__restrict__ src, const double *__restrict__ Cn,
                           const double *__restrict__ vel, const double *__restrict__ sx,
                           const double *__restrict__ sy, const double *__restrict__ sa,
                           const double *__restrict__ G, double dt, double *__restrict__ vbar,
                           double *__restrict__ fs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = m.nx, ny = m.ny;
  const long k = H(m, i, j);
  const long nh = (long)(nx + 2) * (ny + 2);
  const bool face = DIR == 0 ? (i >= 1 && i <= nx + 1 && j >= 1 && j <= ny)
                             : (i >= 1 && i <= nx && j >= 1 && j <= ny + 1);
  if (i > nx + 1 || j > ny + 1) return;
  fs[k] = fs[nh + k] = fs[2 * nh + k] = 0.0;
  if (!face) return;
  const double area = m.dx * m.dy, dh = dt / (DIR == 0 ? m.dx : m.dy);
  // g and c_d of the cells on the high (this face is their low face) and low side
  auto gcell = [&](int a, int b) { return (a >= 1 && a <= nx && b >= 1 && b <= ny) ? G[H(m, a, b)] : 0.0; };
  auto cdcell = [&](int a, int b) { return (a >= 1 && a <= nx && b >= 1 && b <= ny && Cn[H(m, a, b)] >= 0.5) ? 1.0 : 0.0; };
  const int ia = DIR == 0 ? i - 1 : i, ja = DIR == 0 ? j : j - 1;  // the cell below / left
  // dJ/dF as the update used it: + g(high) / |cell|, - g(low) / |cell|
  auto fbar_used = [&](int fi, int fj) {
    const int la = DIR == 0 ? fi - 1 : fi, lb = DIR == 0 ? fj : fj - 1;
    return (gcell(fi, fj) - gcell(la, lb)) / area;
  };
  double fb = fbar_used(i, j);
  // periodic copy: faces 1 and n+1 both use the flux computed at the face whose donor is real
  const bool per = DIR == 0 ? (m.bcl == BC_PERIODIC) : (m.bcb == BC_PERIODIC);
  const int last = DIR == 0 ? nx + 1 : ny + 1;
  const int pos = DIR == 0 ? i : j;
  if (per && (pos == 1 || pos == last)) {
    const double v1 = DIR == 0 ? vel[H(m, 1, j)] : vel[H(m, i, 1)];
    const int keep = v1 >= 0.0 ? last : 1;
    const double both = DIR == 0 ? fbar_used(1, j) + fbar_used(last, j) : fbar_used(i, 1) + fbar_used(i, last);
    fb = (pos == keep) ? both : 0.0;
  }
  // dilation term of the two cells: + cd(high) dh u_f ... - : d/du_f
  double ub = dh * (cdcell(ia, ja) * gcell(ia, ja) - cdcell(i, j) * gcell(i, j));
  const FaceBox b = face_box<DIR>(vel[k], m, i, j, dt);
  if (!b.zero && fb != 0.0) {
    const double sgn = b.pos ? 1.0 : -1.0;
    const double c = src[b.kd];
    const double nxd = sx[b.kd], nyd = sy[b.kd], ald = sa[b.kd];
    // the swept box's far edge moves by -dt * sign(u) per unit u: dF/du = dt * h(far edge)
    double hfar;
    if (DIR == 0) {
      const double X = b.pos ? b.xmin : b.xmax;
      hfar = covered(X, b.ymin, b.ymax, nxd, nyd, ald);
    } else {
      const double Y = b.pos ? b.ymin : b.ymax;
      hfar = covered(Y, b.xmin, b.xmax, nyd, nxd, ald);
    }
    if (c >= 1.0 - 1.0e-6) hfar = DIR == 0 ? b.ymax - b.ymin : b.xmax - b.xmin;
    if (c <= 1.0e-6) hfar = 0.0;
    ub += fb * dt * hfar;
    if (is_mixed(c)) {
      double L, mx, my;
      chord(b.xmin, b.xmax, b.ymin, b.ymax, nxd, nyd, ald, L, mx, my);
      const double w = fb * sgn;
      fs[k] = -w * L * mx;
      fs[nh + k] = -w * L * my;
      fs[2 * nh + k] = -w * L;
    }
  }
  vbar[k] += ub;
}

// Segment adjoint of every cell: the sum over the (at most two) faces of the
// pass it is the donor of.
template <int DIR>
__global__ void k_adj_segbar(DevMesh m, const double *__restrict__ vel, const double *__restrict__ fs, double dt,
                             double *__restrict__ sb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int nx = m.nx, ny = m.ny;
  if (i > nx + 1 || j > ny + 1) return;
  const long nh = (long)(nx + 2) * (ny + 2);
  const long k = H(m, i, j);
  double a = 0.0, bq = 0.0, c = 0.0;
  auto take = [&](int fi, int fj, bool want_pos) {
    const bool face = DIR == 0 ? (fi >= 1 && fi <= nx + 1 && fj >= 1 && fj <= ny)
                               : (fi >= 1 && fi <= nx && fj >= 1 && fj <= ny + 1);
    if (!face) return;
    const long f = H(m, fi, fj);
    const double uf = vel[f];
    if (fabs(uf) * dt <= 1.0e-14 || (uf > 0.0) != want_pos) return;
    a += fs[f];
    bq += fs[nh + f];
    c += fs[2 * nh + f];
  };
  if (DIR == 0) {
    take(i + 1, j, true);   // the face above-right carries this cell's volume when u > 0
    take(i, j, false);      // the low face when u < 0
  } else {
    take(i, j + 1, true);
    take(i, j, false);
  }
  sb[k] = a;
  sb[nh + k] = bq;
  sb[2 * nh + k] = c;
}

// Source cell of a ghost (ring) cell's normal (reconstruct_interface's
// exchange + outward copies): wrap on periodic axes, clamp on physical ones.
__device__ __forceinline__ long ghost_src(const DevMesh &m, int i, int j) {
  return H(m, wrap_or_clamp(i, m.nx, m.perx), wrap_or_clamp(j, m.ny, m.pery));
}

// Ring cells: the ghost segment's alpha and renormalised normal -> dJ/dC of
// the ring cell, and per ring entry the source normal's adjoint (gathered in
// ring order by k_adj_ghost_acc).
__global__ void k_adj_ghost(DevMesh m, const double *__restrict__ src, const double *__restrict__ sx,
                            const double *__restrict__ sy, const double *__restrict__ sa,
                            const double *__restrict__ sb, double *__restrict__ cbar, double *__restrict__ ring) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ring_len(m)) return;
  int i, j;
  ring_cell(m, r, i, j);
  const long nh = (long)(m.nx + 2) * (m.ny + 2);
  const long k = H(m, i, j);
  const long ks = ghost_src(m, i, j);
  ring[3 * r] = ring[3 * r + 1] = 0.0;
  ring[3 * r + 2] = (double)ks;
  const double c = src[k];
  double nbx = sb[k], nby = sb[nh + k], ab = sb[2 * nh + k];
  if (!is_mixed(c) || (nbx == 0.0 && nby == 0.0 && ab == 0.0)) return;
  const double xmin = face_coord(i, m.dx), xmax = face_coord(i + 1, m.dx);
  const double ymin = face_coord(j, m.dy), ymax = face_coord(j + 1, m.dy);
  const double nx = sx[k], ny = sy[k];
  double L, mx, my;
  chord(xmin, xmax, ymin, ymax, nx, ny, sa[k], L, mx, my);
  if (L > 0.0) {
    cbar[k] += ab * (-(m.dx * m.dy) / L);
    nbx += ab * (-mx);
    nby += ab * (-my);
  }
  // n = s / |s| with s the source cell's unit normal (a non-mixed source
  // gives the constant (1, 0))
  if (!is_mixed(src[ks])) return;
  const double s = hypot_ref(sx[ks], sy[ks]);
  if (!(s > 1.0e-12)) return;
  const double dot = nbx * nx + nby * ny;
  ring[3 * r] = (nbx - dot * nx) / s;
  ring[3 * r + 1] = (nby - dot * ny) / s;
}

// Whether interior cell (i, j) of a ring entry's wrap / clamp target is one
// of the four corner cells, the only targets more than one ring cell maps to
// (on meshes of at least 2 x 2 cells): the others take one contribution each,
// in any order; the corner targets take theirs in ring order, as one serial
// loop would.
__device__ __forceinline__ bool corner_target(const DevMesh &m, int ti, int tj) {
  return (ti == 1 || ti == m.nx) && (tj == 1 || tj == m.ny);
}
__device__ __forceinline__ bool ring_serial_only(const DevMesh &m) { return m.nx < 2 || m.ny < 2; }

__global__ void k_adj_ghost_acc(DevMesh m, const double *__restrict__ ring, double *__restrict__ sb) {
  const long nh = (long)(m.nx + 2) * (m.ny + 2);
  const int L = ring_len(m);
  const bool serial = ring_serial_only(m);
  auto target_of = [&](int r, int &ti, int &tj) {
    int i, j;
    ring_cell(m, r, i, j);
    ti = wrap_or_clamp(i, m.nx, m.perx);
    tj = wrap_or_clamp(j, m.ny, m.pery);
  };
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (!serial && r < L) {
    int ti, tj;
    target_of(r, ti, tj);
    if (!corner_target(m, ti, tj)) {
      const long ks = (long)ring[3 * r + 2];
      sb[ks] += ring[3 * r];
      sb[nh + ks] += ring[3 * r + 1];
    }
  }
  if (r != 0) return;
  for (int q = 0; q < L; ++q) {  // the corner targets (every target on a degenerate mesh), in ring order
    int ti, tj;
    target_of(q, ti, tj);
    if (!serial && !corner_target(m, ti, tj)) continue;
    const long ks = (long)ring[3 * q + 2];
    sb[ks] += ring[3 * q];
    sb[nh + ks] += ring[3 * q + 1];
  }
}

// d(cnx, cny)/dC of ELVIRA candidate k (0..6, 6 = Youngs) over the 3x3
// stencil, index (di + 1) + 3 (dj + 1), with the sign factors frozen.
__device__ void cand_weights(int k, double sgn_x, double sgn_y, double dx, double dy, double *wx, double *wy) {
  for (int q = 0; q < 9; ++q) wx[q] = wy[q] = 0.0;
  auto at = [](int di, int dj) { return (di + 1) + 3 * (dj + 1); };
  const double w3[3] = {1.0, 2.0, 1.0};
  for (int d = -1; d <= 1; ++d) {
    if (k == 0) { wx[at(1, d)] += -sgn_y * dy / (2.0 * dx); wx[at(-1, d)] += sgn_y * dy / (2.0 * dx); }
    if (k == 1) { wx[at(1, d)] += -sgn_y * dy / dx; wx[at(0, d)] += sgn_y * dy / dx; }
    if (k == 2) { wx[at(0, d)] += -sgn_y * dy / dx; wx[at(-1, d)] += sgn_y * dy / dx; }
    if (k == 3) { wy[at(d, 1)] += -sgn_x * dx / (2.0 * dy); wy[at(d, -1)] += sgn_x * dx / (2.0 * dy); }
    if (k == 4) { wy[at(d, 1)] += -sgn_x * dx / dy; wy[at(d, 0)] += sgn_x * dx / dy; }
    if (k == 5) { wy[at(d, 0)] += -sgn_x * dx / dy; wy[at(d, -1)] += sgn_x * dx / dy; }
    if (k == 6) {
      wx[at(1, d)] += -w3[d + 1] / (8.0 * dx);
      wx[at(-1, d)] += w3[d + 1] / (8.0 * dx);
      wy[at(d, 1)] += -w3[d + 1] / (8.0 * dy);
      wy[at(d, -1)] += w3[d + 1] / (8.0 * dy);
    }
  }
}

// Interior mixed cells: alpha's dependence on the cell's own C and on the
// normal, then the normal's on the 3x3 stencil (W: 9 weights per cell).
template <int METHOD>
__global__ void k_adj_interior(DevMesh m, const double *__restrict__ src, const double *__restrict__ sx,
                               const double *__restrict__ sy, const double *__restrict__ sa,
                               const double *__restrict__ sb, double *__restrict__ cbar, double *__restrict__ W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
  const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
  if (i > m.nx || j > m.ny) return;
  const long nh = (long)(m.nx + 2) * (m.ny + 2);
  const long k = H(m, i, j);
  for (int q = 0; q < 9; ++q) W[q * nh + k] = 0.0;
  const double c = src[k];
  double nbx = sb[k], nby = sb[nh + k], ab = sb[2 * nh + k];
  if (!is_mixed(c) || (nbx == 0.0 && nby == 0.0 && ab == 0.0)) return;
  const double dx = m.dx, dy = m.dy;
  // the candidate the reconstruction chose (PLIC: Youngs).  ELVIRA: the
  // derivative is taken along the lowest-index candidate that ties with the
  // winner to rounding -- error within 1e-10 relative and the same unit
  // normal to 1e-9 (Youngs' normal and the central height-function one
  // coincide wherever the outer rows or columns of the block are uniform,
  // and rounding alone then decides the winner)
  int kc = 6;
  if (METHOD == 1) {
    double cux[7], cuy[7], cer[7];
    bool cok[7];
    double best = 1.0e30;
    int win = -1;
    for (int q = 0; q < 7; ++q) {
      double al;
      elvira_candidate(src, m, i, j, c, q, cux[q], cuy[q], al, cer[q], cok[q]);
      if (cok[q] && cer[q] < best) {
        best = cer[q];
        win = q;
      }
    }
    if (win < 0) return;  // no candidate: the constant segment (1, 0, 0)
    const double tol = 1.0e-10 * best + 1.0e-20 * (dx * dy) * (dx * dy);
    kc = win;
    for (int q = 0; q < win; ++q)
      if (cok[q] && cer[q] <= best + tol && fabs(cux[q] - cux[win]) + fabs(cuy[q] - cuy[win]) <= 1.0e-9) {
        kc = q;
        break;
      }
  }
  const double nx = sx[k], ny = sy[k];
  double L, mx, my;
  chord(face_coord(i, dx), face_coord(i + 1, dx), face_coord(j, dy), face_coord(j + 1, dy), nx, ny, sa[k], L, mx,
        my);
  if (L > 0.0) {
    cbar[k] += ab * (-(dx * dy) / L);
    nbx += ab * (-mx);
    nby += ab * (-my);
  }
#define CC(a, b) src[H(m, i + (a), j + (b))]
  const double sgn_y = (CC(0, 1) < CC(0, -1)) ? 1.0 : -1.0;
  const double sgn_x = (CC(1, 0) < CC(-1, 0)) ? 1.0 : -1.0;
  double cx, cy;  // the unnormalised candidate
  if (kc == 6) {
    youngs(src, m, i, j, cx, cy);
  } else if (kc < 3) {
    const double H_m1 = (CC(-1, -1) + CC(-1, 0) + CC(-1, 1)) * dy;
    const double H_0 = (CC(0, -1) + CC(0, 0) + CC(0, 1)) * dy;
    const double H_p1 = (CC(1, -1) + CC(1, 0) + CC(1, 1)) * dy;
    cx = kc == 0 ? -(sgn_y * (H_p1 - H_m1) / (2.0 * dx)) : kc == 1 ? -(sgn_y * (H_p1 - H_0) / dx)
                                                                   : -(sgn_y * (H_0 - H_m1) / dx);
    cy = sgn_y;
  } else {
    const double W_m1 = (CC(-1, -1) + CC(0, -1) + CC(1, -1)) * dx;
    const double W_0 = (CC(-1, 0) + CC(0, 0) + CC(1, 0)) * dx;
    const double W_p1 = (CC(-1, 1) + CC(0, 1) + CC(1, 1)) * dx;
    cx = sgn_x;
    cy = kc == 3 ? -(sgn_x * (W_p1 - W_m1) / (2.0 * dy)) : kc == 4 ? -(sgn_x * (W_p1 - W_0) / dy)
                                                                   : -(sgn_x * (W_0 - W_m1) / dy);
  }
#undef CC
  const double cn = hypot_ref(cx, cy);
  if (!(cn > 1.0e-12)) return;
  const double dot = nbx * nx + nby * ny;
  const double mbx = (nbx - dot * nx) / cn, mby = (nby - dot * ny) / cn;
  double wx[9], wy[9];
  cand_weights(kc, sgn_x, sgn_y, dx, dy, wx, wy);
  for (int q = 0; q < 9; ++q) W[q * nh + k] = mbx * wx[q] + mby * wy[q];
}

// dJ/dC of every cell from the stencil weights of the interior cells around it.
__global__ void k_adj_gather(DevMesh m, const double *__restrict__ W, double *__restrict__ cbar) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  if (i > m.nx + 1 || j > m.ny + 1) return;
  const long nh = (long)(m.nx + 2) * (m.ny + 2);
  double s = 0.0;
  for (int dj = -1; dj <= 1; ++dj)
    for (int di = -1; di <= 1; ++di) {
      const int a = i - di, b = j - dj;  // the cell whose stencil holds (i, j) at offset (di, dj)
      if (a < 1 || a > m.nx || b < 1 || b > m.ny) continue;
      s += W[(long)((di + 1) + 3 * (dj + 1)) * nh + H(m, a, b)];
    }
  cbar[H(m, i, j)] += s;
}

// Halo of a field copied from its interior (zero-gradient copy or wrap, the
// corners from the diagonal cell): moves the halo entries of cbar onto their
// interior sources in ring order (one thread, a fixed summation order).
// corners == 0 leaves the four corner entries in place (their values came
// from another field).
__global__ void k_adj_fold(DevMesh m, double *__restrict__ cbar, int corners) {
  const int nx = m.nx, ny = m.ny;
  const int L = ring_len(m);
  const bool serial = ring_serial_only(m);
  // ring cell q moved onto its target; false: a skipped corner, or (with
  // `corner_pass`) a target another pass handles
  auto fold = [&](int q, bool corner_pass) {
    int i, j;
    ring_cell(m, q, i, j);
    const bool corner = (i == 0 || i == nx + 1) && (j == 0 || j == ny + 1);
    if (corner && !corners) return;
    const int ti = wrap_or_clamp(i, nx, m.perx), tj = wrap_or_clamp(j, ny, m.pery);
    if (!serial && corner_target(m, ti, tj) != corner_pass) return;
    const long k = H(m, i, j);
    const double v = cbar[k];
    cbar[k] = 0.0;
    cbar[H(m, ti, tj)] += v;
  };
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (!serial && r < L) fold(r, false);
  if (r != 0) return;
  for (int q = 0; q < L; ++q) fold(q, true);  // the corner targets, in ring order
}

// Halo of C as the transport leaves it with the two-phase corner fix, from
// the interior alone: zero-gradient copies on physical sides, the wrap on
// periodic ones, the corners from the diagonal cell (one thread per ring cell).
__global__ void k_fill_halo(DevMesh m, double *C) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= ring_len(m)) return;
  int i, j;
  ring_cell(m, r, i, j);
  C[H(m, i, j)] = C[H(m, wrap_or_clamp(i, m.nx, m.perx), wrap_or_clamp(j, m.ny, m.pery))];
}

int vof_fill_halo(Ctx &c, double *C) {
  if (c.m.perx != c.m.pery) return COVOF_BAD_ARGUMENT;
  const int L = 2 * (c.m.nx + 2) + 2 * c.m.ny;
  hipLaunchKernelGGL(k_fill_halo, dim3((L + 255) / 256), dim3(256), 0, c.stream, dev_mesh(c.m), C);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// Work arrays of the step adjoint (made on the first call): the first
// pass's segments (3), g, the per-face segment adjoints (3), the per-cell
// segment adjoints (3), the stencil weights (9), dJ/dC* and a copy of
// dJ/dC^{n+1} (halo-size each), and the ring records (3 per ring cell).
static int adj_alloc(Ctx &c) {
  if (c.vof.adj) return COVOF_OK;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const long L = 2L * (c.m.nx + 2) + 2L * c.m.ny;
  COVOF_CHECK(hipMalloc(&c.vof.adj, (21 * nh + 3 * L) * sizeof(double)));
  return COVOF_OK;
}

// Reverse mode of one pass (src -> dst with the segments sx, sy, sa of src):
// dJ/dC_out -> dJ/dC_in, and dJ/d(face velocity) added to velbar.
template <int DIR>
static int adj_pass(Ctx &c, int method, const double *src, const double *Cn, const double *vel, const double *sx,
                    const double *sy, const double *sa, const double *cbar_out, int clamp, int corners_through,
                    double dt, double *cbar_in, double *velbar) {
  const DevMesh m = dev_mesh(c.m);
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  double *G = c.vof.adj + 3 * nh, *fs = G + nh, *sb = fs + 3 * nh, *W = sb + 3 * nh, *ring = W + 9 * nh + 2 * nh;
  const dim3 blk(64, 4), grd((c.m.nx + 2 + 63) / 64, (c.m.ny + 2 + 3) / 4);
  const dim3 grd_in((c.m.nx + 63) / 64, (c.m.ny + 3) / 4);
  const int L = 2 * (c.m.nx + 2) + 2 * c.m.ny;
  hipLaunchKernelGGL(k_adj_g<DIR>, grd, blk, 0, c.stream, m, src, Cn, vel, sx, sy, sa, cbar_out, clamp,
                     corners_through, dt, G, cbar_in);
  hipLaunchKernelGGL(k_adj_face<DIR>, grd, blk, 0, c.stream, m, src, Cn, vel, sx, sy, sa, (const double *)G, dt,
                     velbar, fs);
  hipLaunchKernelGGL(k_adj_segbar<DIR>, grd, blk, 0, c.stream, m, vel, (const double *)fs, dt, sb);
  hipLaunchKernelGGL(k_adj_ghost, dim3((L + 255) / 256), dim3(256), 0, c.stream, m, src, sx, sy, sa,
                     (const double *)sb, cbar_in, ring);
  hipLaunchKernelGGL(k_adj_ghost_acc, dim3((ring_cells(c.m) + 255) / 256), dim3(256), 0, c.stream, m,
                     (const double *)ring, sb);
  hipLaunchKernelGGL(method == 1 ? k_adj_interior<1> : k_adj_interior<0>, grd_in, blk, 0, c.stream, m, src, sx, sy,
                     sa, (const double *)sb, cbar_in, W);
  hipLaunchKernelGGL(k_adj_gather, grd, blk, 0, c.stream, m, (const double *)W, cbar_in);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// Reverse mode of one vof_advect step (WEYMOUTH_YUE, pass order from
// step_count, the two-phase corner fix with two_phase): recomputes the
// step's intermediate field C* and both passes' segments, then runs the
// passes backwards.  cbar_new = dJ/dC^{n+1} (halo layout, halo entries
// included); cbar = dJ/dC^n (halo layout; with fold_in the halo entries are
// moved onto the interior cells they copy, else they are the derivatives
// with respect to C^n's halo values as given); ubar, vbar += dJ/du, dJ/dv.
int vof_adjoint_reserve(Ctx &c) { return adj_alloc(c); }

int vof_advect_adjoint(Ctx &c, int method, const double *C, const double *u, const double *v, double dt,
                       int step_count, int two_phase, int fold_in, const double *cbar_new, double *cbar,
                       double *ubar, double *vbar) {
  if (c.m.perx != c.m.pery || (c.jet.on && !c.m.pery)) return COVOF_BAD_ARGUMENT;
  int st = adj_alloc(c);
  if (st) return st;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const DevMesh m = dev_mesh(c.m);
  const int per = c.m.perx;
  const bool x_first = (step_count % 2) == 0;
  double *seg1 = c.vof.adj, *cmid = c.vof.adj + 19 * nh, *cout = cmid + nh;
  // the forward step's first pass (C -> C*) and the segments of C and of C*
  st = launch_classify(c, C);
  if (!st) st = launch_pass(c, method, x_first ? 0 : 1, true, C, c.vof.Cb, u, v, dt, false);
  if (st) return st;
  COVOF_CHECK(hipMemcpyAsync(seg1, c.vof.snx, nh * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  COVOF_CHECK(hipMemcpyAsync(seg1 + nh, c.vof.sny, nh * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  COVOF_CHECK(hipMemcpyAsync(seg1 + 2 * nh, c.vof.sal, nh * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  st = launch_recon(c, method, c.vof.Cb);
  if (st) return st;
  // second pass backwards: dJ/dC^{n+1} (its halo onto the interior it copies)
  COVOF_CHECK(hipMemcpyAsync(cout, cbar_new, nh * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  hipLaunchKernelGGL(k_adj_fold, dim3((ring_cells(c.m) + 255) / 256), dim3(256), 0, c.stream, m, cout,
                     (per || two_phase) ? 1 : 0);
  COVOF_CHECK_LAUNCH();
  const double *Cs = c.vof.Cb;
  if (x_first)
    st = adj_pass<1>(c, method, Cs, C, v, c.vof.snx, c.vof.sny, c.vof.sal, cout, 1, (!per && !two_phase) ? 1 : 0, dt,
                     cmid, vbar);
  else
    st = adj_pass<0>(c, method, Cs, C, u, c.vof.snx, c.vof.sny, c.vof.sal, cout, 1, (!per && !two_phase) ? 1 : 0, dt,
                     cmid, ubar);
  if (st) return st;
  // C*'s halo: edges copied from its interior, corners from C's (walls)
  hipLaunchKernelGGL(k_adj_fold, dim3((ring_cells(c.m) + 255) / 256), dim3(256), 0, c.stream, m, cmid, per);
  COVOF_CHECK_LAUNCH();
  if (x_first)
    st = adj_pass<0>(c, method, C, C, u, seg1, seg1 + nh, seg1 + 2 * nh, cmid, 0, per ? 0 : 1, dt, cbar, ubar);
  else
    st = adj_pass<1>(c, method, C, C, v, seg1, seg1 + nh, seg1 + 2 * nh, cmid, 0, per ? 0 : 1, dt, cbar, vbar);
  if (st) return st;
  if (fold_in) {
    hipLaunchKernelGGL(k_adj_fold, dim3((ring_cells(c.m) + 255) / 256), dim3(256), 0, c.stream, m, cbar, 1);
    COVOF_CHECK_LAUNCH();
  }
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
