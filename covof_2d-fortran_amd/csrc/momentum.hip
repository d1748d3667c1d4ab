// The momentum block of the standard projection (the SP velocity predictor):
// src/module_simulation.f90:2454-2469 —
//   assemble_preconditioner_blocks (the Jacobi diagonal)
//       src/module_navier_stokes.f90:1268-1595
//   assemble_rhs                        src/module_navier_stokes.f90:785-1101
//   pack_vector / unpack_vector         src/module_navier_stokes.f90:231-375
//   apply_system_matrix                 src/module_navier_stokes.f90:380-780
//   solve_linear_system (BiCGSTAB(2), CG; NONE / JACOBI)
//       src/module_linear_solvers.f90:88-135, :361-579, :735-838, :1167-1208
//   sync_u / sync_v + apply_velocity_bc src/module_mpi_comm.f90:46-56, :290-537
//
// Supported subset (everything else returns COVOF_BAD_ARGUMENT at bind time):
// the standard projection (no augmented-Lagrangian or monolithic terms), no
// obstacles, UNIFORM boundary profiles, no energy equation, no SSP-RK3
// momentum, and no NEUMANN top side (the reference's v-row at j = ny+1 then
// reads past its array).
//
// Layout: the unknown vector is u(1:u_nx, 1:ny) then v(1:nx, 1:v_ny), i
// fastest (u_nx = nx+1 and v_ny = ny+1 on physical sides), device FP64.
// Every kernel evaluates the reference's expressions in its operand order
// (-ffp-contract=off), so diagonal, right-hand side and products are bitwise
// the reference's; the Krylov loops differ from it only in the summation
// order of their inner products.
//
// The operator keeps the reference's persistent work fields (its SAVEd `tf`):
// each product unpacks x into them, exchanges the periodic halos and applies
// the homogeneous velocity BC in the reference's order, and ghost cells that
// sequence does not write keep what the previous product left (zero at first
// use), exactly as the reference's do.
#include "covof_internal.h"

#include <cmath>

namespace covof {

namespace {

constexpr int BLK = 256;

// dof -> (i, j, is_v)
struct Dof {
  int i, j, isv;
};
__device__ __forceinline__ Dof dof_ij(long k, long nu, int u_nx, int nx) {
  Dof d;
  if (k < nu) {
    d.isv = 0;
    d.j = (int)(k / u_nx) + 1;
    d.i = (int)(k - (long)(d.j - 1) * u_nx) + 1;
  } else {
    const long kv = k - nu;
    d.isv = 1;
    d.j = (int)(kv / nx) + 1;
    d.i = (int)(kv - (long)(d.j - 1) * nx) + 1;
  }
  return d;
}

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }
__device__ __forceinline__ bool wallbc(int t) { return t == BC_DIRICHLET || t == BC_FREESLIP; }

#define HX(i, j) ((long)(j) * (m.nx + 2) + (i))

// conv_face_flux (src/module_navier_stokes.f90:2388-2416): centered, upwind,
// or the superbee-limited TVD flux.
__device__ double conv_face_flux(int scheme, double flux, double phi_L, double phi_R, double q_LL, double q_L,
                                 double q_R, double q_RR, bool has_LL, bool has_RR) {
  if (scheme == 2) return 0.5 * flux * (phi_L + phi_R);
  if (scheme == 1) return flux >= 0.0 ? flux * phi_L : flux * phi_R;
  const double dq = q_R - q_L;
  double psi = 0.0;
  if (flux >= 0.0) {
    if (has_LL && fabs(dq) > 1.0e-14) {
      const double r = (q_L - q_LL) / dq;
      psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
    }
    return flux * (phi_L + 0.5 * psi * (phi_R - phi_L));
  }
  if (has_RR && fabs(dq) > 1.0e-14) {
    const double r = (q_RR - q_R) / dq;
    psi = dmax(dmax(0.0, dmin(2.0 * r, 1.0)), dmin(r, 2.0));
  }
  return flux * (phi_R + 0.5 * psi * (phi_L - phi_R));
}

// compute_u_conv_fluxes (src/module_navier_stokes.f90:2418-2457)
__device__ void u_fluxes(const Mesh &m, const MomCfg &c, const MomFields &F, int i, int j, double &FE,
                         double &FW, double &FN, double &FS) {
  const double *uu = c.lin == 2 ? F.uo : F.u, *vv = c.lin == 2 ? F.vo : F.v;
  double u_n = uu[HX(i, j)], u_n_E = uu[HX(i + 1, j)], u_n_W = uu[HX(i - 1, j)];
  double v_n_N = 0.5 * (vv[HX(i - 1, j + 1)] + vv[HX(i, j + 1)]);
  double v_n_S = 0.5 * (vv[HX(i - 1, j)] + vv[HX(i, j)]);
  if (c.ab_nm1) {
    u_n = 2.0 * u_n - F.un[HX(i, j)];
    u_n_E = 2.0 * u_n_E - F.un[HX(i + 1, j)];
    u_n_W = 2.0 * u_n_W - F.un[HX(i - 1, j)];
    const double v_nm1_N = 0.5 * (F.vn[HX(i - 1, j + 1)] + F.vn[HX(i, j + 1)]);
    const double v_nm1_S = 0.5 * (F.vn[HX(i - 1, j)] + F.vn[HX(i, j)]);
    v_n_N = 2.0 * v_n_N - v_nm1_N;
    v_n_S = 2.0 * v_n_S - v_nm1_S;
  }
  FE = 0.5 * (u_n + u_n_E) * m.dy;
  FW = 0.5 * (u_n_W + u_n) * m.dy;
  FN = v_n_N * m.dx;
  FS = v_n_S * m.dx;
}

// compute_v_conv_fluxes (src/module_navier_stokes.f90:2459-2498)
__device__ void v_fluxes(const Mesh &m, const MomCfg &c, const MomFields &F, int i, int j, double &FE,
                         double &FW, double &FN, double &FS) {
  const double *uu = c.lin == 2 ? F.uo : F.u, *vv = c.lin == 2 ? F.vo : F.v;
  double v_n = vv[HX(i, j)], v_n_N = vv[HX(i, j + 1)], v_n_S = vv[HX(i, j - 1)];
  double u_n_E = 0.5 * (uu[HX(i + 1, j - 1)] + uu[HX(i + 1, j)]);
  double u_n_W = 0.5 * (uu[HX(i, j - 1)] + uu[HX(i, j)]);
  if (c.ab_nm1) {
    v_n = 2.0 * v_n - F.vn[HX(i, j)];
    v_n_N = 2.0 * v_n_N - F.vn[HX(i, j + 1)];
    v_n_S = 2.0 * v_n_S - F.vn[HX(i, j - 1)];
    const double u_nm1_E = 0.5 * (F.un[HX(i + 1, j - 1)] + F.un[HX(i + 1, j)]);
    const double u_nm1_W = 0.5 * (F.un[HX(i, j - 1)] + F.un[HX(i, j)]);
    u_n_E = 2.0 * u_n_E - u_nm1_E;
    u_n_W = 2.0 * u_n_W - u_nm1_W;
  }
  FE = u_n_E * m.dy;
  FW = u_n_W * m.dy;
  FN = 0.5 * (v_n + v_n_N) * m.dx;
  FS = 0.5 * (v_n_S + v_n) * m.dx;
}

// x -> interior unknowns of the halo fields u, v (and p beyond nuv: the
// monolithic system) -- ns_unpack_vector's copy loops
__global__ void __launch_bounds__(BLK) k_unpack(Mesh m, long nu, long nuv, long n, int u_nx,
                                                const double *__restrict__ x, double *__restrict__ u,
                                                double *__restrict__ v, double *__restrict__ p) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1;
    p[HX((int)(kp - (long)(j - 1) * m.nx) + 1, j)] = x[k];
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  (d.isv ? v : u)[HX(d.i, d.j)] = x[k];
}

// ns_pack_vector
__global__ void __launch_bounds__(BLK) k_pack(Mesh m, long nu, long nuv, long n, int u_nx,
                                              const double *__restrict__ u, const double *__restrict__ v,
                                              const double *__restrict__ p, double *__restrict__ x) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1;
    x[k] = p[HX((int)(kp - (long)(j - 1) * m.nx) + 1, j)];
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  x[k] = (d.isv ? v : u)[HX(d.i, d.j)];
}

// sync_u, sync_v (exchange_halos_scalar on one rank: periodic columns over
// all rows, then periodic rows over all columns) and apply_velocity_bc, in
// the reference's statement order: one workgroup, a barrier between every
// step whose reads depend on an earlier step's writes.
// The single-rank exchange_halos_scalar of one halo field inside a
// workgroup (all threads of the block call it).
__device__ void wg_exchange(const Mesh &m, double *f) {
  const int nx = m.nx, ny = m.ny, t = threadIdx.x, T = blockDim.x;
  if (m.perx) {
    for (int j = t; j <= ny + 1; j += T) f[HX(0, j)] = f[HX(nx, j)];
    __syncthreads();
    for (int j = t; j <= ny + 1; j += T) f[HX(nx + 1, j)] = f[HX(1, j)];
    __syncthreads();
  }
  if (m.pery) {
    for (int i = t; i <= nx + 1; i += T) f[HX(i, 0)] = f[HX(i, ny)];
    __syncthreads();
    for (int i = t; i <= nx + 1; i += T) f[HX(i, ny + 1)] = f[HX(i, 1)];
    __syncthreads();
  }
}

__global__ void __launch_bounds__(1024) k_vel_halo(Mesh m, MomCfg c, double *__restrict__ u,
                                                   double *__restrict__ v, int homogeneous, int exchange,
                                                   double *__restrict__ p) {
  const int nx = m.nx, ny = m.ny, t = threadIdx.x, T = blockDim.x;
  const int *bc = m.bc;
  const bool bl = !m.perx, bb = !m.pery;  // physical sides (single rank)
  if (exchange) {
    if (p) wg_exchange(m, p);  // sync_cell_scalar of the monolithic pressure
    wg_exchange(m, u);
    wg_exchange(m, v);
  }
  // apply_velocity_bc (src/module_mpi_comm.f90:290-537), UNIFORM profiles
  const double u_l = homogeneous ? 0.0 : c.bcu[0], v_l = homogeneous ? 0.0 : c.bcv[0];
  const double u_r = homogeneous ? 0.0 : c.bcu[1], v_r = homogeneous ? 0.0 : c.bcv[1];
  const double u_b = homogeneous ? 0.0 : c.bcu[2], v_b = homogeneous ? 0.0 : c.bcv[2];
  const double u_t = homogeneous ? 0.0 : c.bcu[3], v_t = homogeneous ? 0.0 : c.bcv[3];
  if (bc[0] == BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      u[HX(0, j)] = u[HX(nx, j)];
      u[HX(nx + 1, j)] = u[HX(1, j)];
      v[HX(0, j)] = v[HX(nx, j)];
      v[HX(nx + 1, j)] = v[HX(1, j)];
    }
    __syncthreads();
  }
  if (bc[2] == BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      v[HX(i, 0)] = v[HX(i, ny)];
      v[HX(i, ny + 1)] = v[HX(i, 1)];
    }
    __syncthreads();
    for (int i = t; i <= nx + 1; i += T) {
      u[HX(i, 0)] = u[HX(i, ny)];
      u[HX(i, ny + 1)] = u[HX(i, 1)];
    }
    __syncthreads();
  }
  if (bl && bc[0] != BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      if (bc[0] == BC_DIRICHLET) {
        if (!homogeneous) u[HX(1, j)] = u_l;
        u[HX(0, j)] = u_l;
        v[HX(0, j)] = 2.0 * v_l - v[HX(1, j)];
      } else if (bc[0] == BC_FREESLIP) {
        if (!homogeneous) u[HX(1, j)] = 0.0;
        u[HX(0, j)] = 0.0;
        v[HX(0, j)] = v[HX(1, j)];
      } else if (bc[0] == BC_NEUMANN) {
        u[HX(0, j)] = u[HX(1, j)];
        v[HX(0, j)] = v[HX(1, j)];
      }
    }
    __syncthreads();
  }
  if (bl && bc[1] != BC_PERIODIC) {
    for (int j = 1 + t; j <= ny; j += T) {
      if (bc[1] == BC_DIRICHLET) {
        if (!homogeneous) u[HX(nx + 1, j)] = u_r;
        v[HX(nx + 1, j)] = 2.0 * v_r - v[HX(nx, j)];
      } else if (bc[1] == BC_FREESLIP) {
        if (!homogeneous) u[HX(nx + 1, j)] = 0.0;
        v[HX(nx + 1, j)] = v[HX(nx, j)];
      } else if (bc[1] == BC_NEUMANN) {
        if (!homogeneous) u[HX(nx + 1, j)] = u[HX(nx, j)];
        v[HX(nx + 1, j)] = v[HX(nx, j)];
      }
    }
    __syncthreads();
  }
  if (bb && bc[2] != BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      if (bc[2] == BC_DIRICHLET) {
        if (!homogeneous) v[HX(i, 1)] = v_b;
        v[HX(i, 0)] = v_b;
        if (ny >= 2) u[HX(i, 0)] = (8.0 * u_b - 6.0 * u[HX(i, 1)] + u[HX(i, 2)]) / 3.0;
        else u[HX(i, 0)] = 2.0 * u_b - u[HX(i, 1)];
      } else if (bc[2] == BC_FREESLIP) {
        if (!homogeneous) v[HX(i, 1)] = 0.0;
        v[HX(i, 0)] = 0.0;
        u[HX(i, 0)] = u[HX(i, 1)];
      } else if (bc[2] == BC_NEUMANN) {
        u[HX(i, 0)] = u[HX(i, 1)];
        v[HX(i, 0)] = v[HX(i, 1)];
      }
    }
    __syncthreads();
  }
  if (bb && bc[3] != BC_PERIODIC) {
    for (int i = 1 + t; i <= nx; i += T) {
      if (bc[3] == BC_DIRICHLET) {
        if (!homogeneous) v[HX(i, ny + 1)] = v_t;
        if (ny >= 2) u[HX(i, ny + 1)] = (8.0 * u_t - 6.0 * u[HX(i, ny)] + u[HX(i, ny - 1)]) / 3.0;
        else u[HX(i, ny + 1)] = 2.0 * u_t - u[HX(i, ny)];
      } else if (bc[3] == BC_FREESLIP) {
        if (!homogeneous) v[HX(i, ny + 1)] = 0.0;
        u[HX(i, ny + 1)] = u[HX(i, ny)];
      } else if (bc[3] == BC_NEUMANN) {
        u[HX(i, ny + 1)] = u[HX(i, ny)];
        if (!homogeneous) v[HX(i, ny + 1)] = v[HX(i, ny)];
      }
    }
    __syncthreads();
  }
  if (!homogeneous && t == 0) {
    if (bl && bc[1] != BC_PERIODIC) {
      if (bb && bc[2] != BC_PERIODIC) {
        u[HX(nx + 1, 0)] = u[HX(nx + 1, 1)];
        v[HX(nx + 1, 0)] = v[HX(nx, 0)];
      }
      if (bb && bc[3] != BC_PERIODIC) {
        u[HX(nx + 1, ny + 1)] = u[HX(nx + 1, ny)];
        v[HX(nx + 1, ny + 1)] = v[HX(nx, ny + 1)];
      }
    }
    if (bl && bc[0] != BC_PERIODIC && bb && bc[3] != BC_PERIODIC) v[HX(0, ny + 1)] = v[HX(1, ny + 1)];
  }
}

// ns_apply_system_matrix, SP branch (src/module_navier_stokes.f90:380-780):
// row k of A x from the unpacked work fields tu, tv.
__device__ __forceinline__ double mom_row(const Mesh &m, const MomCfg &c, const MomFields &F, double dt, long k,
                                          long nu, int u_nx, int v_ny, const double *__restrict__ tu,
                                          const double *__restrict__ tv, const double *__restrict__ tp) {
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  const int sch = c.scheme;
  const bool has_ee = i + 2 <= nx + 1, has_ww = i - 2 >= 0, has_nn = j + 2 <= ny + 1, has_ss = j - 2 >= 0;
#define TU(a, b) tu[HX(a, b)]
#define TV(a, b) tv[HX(a, b)]
  if (!d.isv) {
    const double *u = F.u;
    const double rho_u = F.ru[HX(i, j)];
    const double diag_bc = (at * rho_u / dt) * dvol;
    if (i == 1 && bl) {
      return m.bc[0] == BC_NEUMANN ? diag_bc * (TU(1, j) - TU(2, j)) : diag_bc * TU(1, j);
    }
    if (i == u_nx && bl) {
      return m.bc[1] == BC_NEUMANN ? diag_bc * (TU(u_nx, j) - TU(u_nx - 1, j)) : diag_bc * TU(u_nx, j);
    }
    double val = (at * rho_u / dt) * TU(i, j) * dvol;
    if (c.conv) {
      double cE, cW, cN, cS;
      u_fluxes(m, c, F, i, j, cE, cW, cN, cS);
      const double q_ee = has_ee ? u[HX(i + 2, j)] : 0.0, q_ww = has_ww ? u[HX(i - 2, j)] : 0.0;
      const double q_nn = has_nn ? u[HX(i, j + 2)] : 0.0, q_ss = has_ss ? u[HX(i, j - 2)] : 0.0;
      const double uW = u[HX(i - 1, j)], uP = u[HX(i, j)], uE = u[HX(i + 1, j)];
      const double uS = u[HX(i, j - 1)], uN = u[HX(i, j + 1)];
      const double pE = conv_face_flux(sch, cE, TU(i, j), TU(i + 1, j), uW, uP, uE, q_ee, true, has_ee);
      const double pW = conv_face_flux(sch, cW, TU(i - 1, j), TU(i, j), q_ww, uW, uP, uE, has_ww, true);
      const double pN = conv_face_flux(sch, cN, TU(i, j), TU(i, j + 1), uS, uP, uN, q_nn, true, has_nn);
      const double pS = conv_face_flux(sch, cS, TU(i, j - 1), TU(i, j), q_ss, uS, uP, uN, has_ss, true);
      val = val + rho_u * (pE - pW + pN - pS);
      if (c.lin == 3) {
        const double du_dx = (uE - uW) / (2.0 * dx);
        const double du_dy = (uN - uS) / (2.0 * dy);
        const double tvf = 0.25 * (TV(i - 1, j) + TV(i, j) + TV(i - 1, j + 1) + TV(i, j + 1));
        val = val + rho_u * (TU(i, j) * du_dx + tvf * du_dy) * dvol;
      }
    }
    const double mu_E = F.mu[HX(i, j)], mu_W = F.mu[HX(i - 1, j)];
    const double deE = 2.0 * mu_E * (TU(i + 1, j) - TU(i, j)) / dx * dy;
    const double deW = 2.0 * mu_W * (TU(i, j) - TU(i - 1, j)) / dx * dy;
    val = val - (deE - deW);
    const double mnN = F.mn[HX(i, j + 1)], mnS = F.mn[HX(i, j)];
    const double shN = mnN * ((TU(i, j + 1) - TU(i, j)) / dy + (TV(i, j + 1) - TV(i - 1, j + 1)) / dx) * dx;
    const double shS = mnS * ((TU(i, j) - TU(i, j - 1)) / dy + (TV(i, j) - TV(i - 1, j)) / dx) * dx;
    val = val - (shN - shS);
    if (c.mono) val = val + (tp[HX(i, j)] - tp[HX(i - 1, j)]) * dy;  // B_u^T p
    return val;
  } else {
    const double *v = F.v;
    const double rho_v = F.rv[HX(i, j)];
    const double diag_bc = (at * rho_v / dt) * dvol;
    if ((j == 1 && bb && wallbc(m.bc[2])) || (j == v_ny && bb && wallbc(m.bc[3]))) {
      return diag_bc * TV(i, j);
    }
    double val = (at * rho_v / dt) * TV(i, j) * dvol;
    if (c.conv) {
      double cE, cW, cN, cS;
      v_fluxes(m, c, F, i, j, cE, cW, cN, cS);
      const double q_ee = has_ee ? v[HX(i + 2, j)] : 0.0, q_ww = has_ww ? v[HX(i - 2, j)] : 0.0;
      const double q_nn = has_nn ? v[HX(i, j + 2)] : 0.0, q_ss = has_ss ? v[HX(i, j - 2)] : 0.0;
      const double vW = v[HX(i - 1, j)], vP = v[HX(i, j)], vE = v[HX(i + 1, j)];
      const double vS = v[HX(i, j - 1)], vN = v[HX(i, j + 1)];
      const double pE = conv_face_flux(sch, cE, TV(i, j), TV(i + 1, j), vW, vP, vE, q_ee, true, has_ee);
      const double pW = conv_face_flux(sch, cW, TV(i - 1, j), TV(i, j), q_ww, vW, vP, vE, has_ww, true);
      const double pN = conv_face_flux(sch, cN, TV(i, j), TV(i, j + 1), vS, vP, vN, q_nn, true, has_nn);
      const double pS = conv_face_flux(sch, cS, TV(i, j - 1), TV(i, j), q_ss, vS, vP, vN, has_ss, true);
      val = val + rho_v * (pE - pW + pN - pS);
      if (c.lin == 3) {
        const double dv_dx = (vE - vW) / (2.0 * dx);
        const double dv_dy = (vN - vS) / (2.0 * dy);
        const double tuf = 0.25 * (TU(i, j - 1) + TU(i + 1, j - 1) + TU(i, j) + TU(i + 1, j));
        val = val + rho_v * (tuf * dv_dx + TV(i, j) * dv_dy) * dvol;
      }
    }
    const double mu_N = F.mu[HX(i, j)], mu_S = F.mu[HX(i, j - 1)];
    const double deN = 2.0 * mu_N * (TV(i, j + 1) - TV(i, j)) / dy * dx;
    const double deS = 2.0 * mu_S * (TV(i, j) - TV(i, j - 1)) / dy * dx;
    val = val - (deN - deS);
    const double mnE = F.mn[HX(i + 1, j)], mnW = F.mn[HX(i, j)];
    const double shE = mnE * ((TV(i + 1, j) - TV(i, j)) / dx + (TU(i + 1, j) - TU(i + 1, j - 1)) / dy) * dy;
    const double shW = mnW * ((TV(i, j) - TV(i - 1, j)) / dx + (TU(i, j) - TU(i, j - 1)) / dy) * dy;
    val = val - (shE - shW);
    if (c.mono) val = val + (tp[HX(i, j)] - tp[HX(i, j - 1)]) * dx;  // B_v^T p
    return val;
  }
#undef TU
#undef TV
}

// y = A x, one unknown per thread; with JAC, y = M^-1 A x for the Jacobi
// preconditioner (apply_PA, src/module_linear_solvers.f90:323-356): the same
// row value divided as apply_preconditioner divides it.
template <bool JAC>
__global__ void __launch_bounds__(BLK) k_mom_matvec(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                    long n, int u_nx, int v_ny, const double *__restrict__ tu,
                                                    const double *__restrict__ tv, const double *__restrict__ tp,
                                                    const double *__restrict__ diag, double *__restrict__ y) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  double val;
  if (k >= nuv) {
    // continuity row of the monolithic system: div(u) dvol, and the
    // regularisation 1e-10 p(1,1) dvol at cell (1,1)
    const long kp = k - nuv;
    const int j = (int)(kp / m.nx) + 1, i = (int)(kp - (long)(j - 1) * m.nx) + 1;
    const double dvol = m.dx * m.dy;
    const double div = (tu[HX(i + 1, j)] - tu[HX(i, j)]) / m.dx + (tv[HX(i, j + 1)] - tv[HX(i, j)]) / m.dy;
    val = (i == 1 && j == 1) ? div * dvol + 1.0e-10 * tp[HX(1, 1)] * dvol : div * dvol;
  } else {
    val = mom_row(m, c, F, dt, k, nu, u_nx, v_ny, tu, tv, tp);
  }
  if (JAC) {
    const double dk = diag[k];
    y[k] = fabs(dk) > 1.0e-15 ? val / dk : val;
  } else {
    y[k] = val;
  }
}

// Outputs of the blocks kernel; any pointer may be null.  u*: (u_nx, ny),
// v*: (nx, v_ny) block coefficients; uaP2, vaP2: u_aP, v_aP on the halo
// grid, the Schur coefficients' source.
struct BlkOut {
  double *diag;
  double *uW, *uE, *uS, *uN, *uP;
  double *vW, *vE, *vS, *vN, *vP;
  double *uaP2, *vaP2;
};
#define PUT(arr, val)                    \
  do {                                   \
    if (o.arr) o.arr[dof] = (val);       \
  } while (0)

// ns_assemble_preconditioner_blocks (src/module_navier_stokes.f90:1268-1595;
// no AL, Brinkman or obstacle terms) over the u and v unknowns: diag_u,
// diag_v, the SSOR coefficients u_a*, v_a*, and u_aP / v_aP on the halo grid.
__global__ void __launch_bounds__(BLK) k_mom_blocks(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                    int u_nx, int v_ny, BlkOut o) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= nuv) return;
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double dt_safe = dt > 1.0e-30 ? dt : 1.0;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  if (!d.isv) {
    const long dof = k;
    const double rho_u = F.ru[HX(i, j)];
    double Du = (at * rho_u / dt_safe) * dvol;
    if (Du <= 1.0e-30) Du = 1.0;
    if ((i == 1 && bl) || (i == u_nx && bl)) {
      const bool west = i == 1;
      if (o.diag) o.diag[k] = Du;
      if (o.uaP2) o.uaP2[HX(i, j)] = Du;
      const double nb = (m.bc[west ? 0 : 1] == BC_NEUMANN && u_nx >= 2) ? Du : 0.0;
      PUT(uW, west ? 0.0 : nb);
      PUT(uE, west ? nb : 0.0);
      PUT(uS, 0.0);
      PUT(uN, 0.0);
      PUT(uP, Du);
      return;
    }
    const double mu_e = F.mu[HX(i, j)], mu_w = F.mu[HX(i - 1, j)];
    const double cE = 2.0 * mu_e * (dy / dx), cW = 2.0 * mu_w * (dy / dx);
    const double cN_raw = F.mn[HX(i, j + 1)] * (dx / dy), cS_raw = F.mn[HX(i, j)] * (dx / dy);
    double cN_visc = cN_raw, cS_visc = cS_raw, dwy = 0.0, cN_extra = 0.0, cS_extra = 0.0;
    if (j == 1 && bb) {
      if (m.bc[2] == BC_DIRICHLET) {
        if (ny >= 2) {
          dwy = dwy + 3.0 * cS_raw;
          cN_extra = cN_extra + cS_raw / 3.0;
        } else {
          dwy = dwy + 2.0 * cS_raw;
        }
      }
      cS_visc = 0.0;
    }
    if (j == ny && bb) {
      if (m.bc[3] == BC_DIRICHLET) {
        if (ny >= 2) {
          dwy = dwy + 3.0 * cN_raw;
          cS_extra = cS_extra + cN_raw / 3.0;
        } else {
          dwy = dwy + 2.0 * cN_raw;
        }
      }
      cN_visc = 0.0;
    }
    double conv_diag = 0.0, conv_W = 0.0, conv_E = 0.0, conv_S = 0.0, conv_N = 0.0, newt = 0.0;
    if (c.conv) {
      double FE, FW, FN, FS;
      u_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_u * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      conv_E = rho_u * dmax(-FE, 0.0);
      conv_W = rho_u * dmax(FW, 0.0);
      conv_N = (j == ny && bb) ? 0.0 : rho_u * dmax(-FN, 0.0);
      conv_S = (j == 1 && bb) ? 0.0 : rho_u * dmax(FS, 0.0);
      if (c.lin == 3) {
        const double du_dx = (F.u[HX(i + 1, j)] - F.u[HX(i - 1, j)]) / (2.0 * dx);
        newt = rho_u * dmax(0.0, du_dx) * dvol;
      }
    }
    PUT(uW, cW + conv_W + 0.0);
    PUT(uE, cE + conv_E + 0.0);
    PUT(uS, cS_visc + conv_S + cS_extra);
    PUT(uN, cN_visc + conv_N + cN_extra);
    double aP = Du + dwy + conv_diag + 0.0 + newt + cW + cE + cS_visc + cN_visc + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    if (o.diag) o.diag[k] = aP;
    PUT(uP, aP);
    if (o.uaP2) o.uaP2[HX(i, j)] = aP;
  } else {
    const long dof = k - nu;
    const double rho_v = F.rv[HX(i, j)];
    double Dv = (at * rho_v / dt_safe) * dvol;
    if (Dv <= 1.0e-30) Dv = 1.0;
    const bool is_wall = (j == 1 && bb && wallbc(m.bc[2])) || (j == v_ny && bb && wallbc(m.bc[3]));
    const bool nm_s = j == 1 && bb && m.bc[2] == BC_NEUMANN && v_ny >= 2;
    const bool nm_n = j == v_ny && bb && m.bc[3] == BC_NEUMANN && v_ny >= 2;
    if (is_wall || nm_s || nm_n) {
      if (o.diag) o.diag[k] = Dv;
      if (o.vaP2) o.vaP2[HX(i, j)] = Dv;
      PUT(vE, 0.0);
      PUT(vW, 0.0);
      PUT(vN, (!is_wall && nm_s) ? Dv : 0.0);
      PUT(vS, (!is_wall && !nm_s && nm_n) ? Dv : 0.0);
      PUT(vP, Dv);
      return;
    }
    const double mu_n = F.mu[HX(i, j)], mu_s = F.mu[HX(i, j - 1)];
    const double cN = 2.0 * mu_n * (dx / dy), cS = 2.0 * mu_s * (dx / dy);
    const double cE_raw = F.mn[HX(i + 1, j)] * (dy / dx), cW_raw = F.mn[HX(i, j)] * (dy / dx);
    double cE_visc = cE_raw, cW_visc = cW_raw, dwx = 0.0;
    if (i == 1 && bl) {
      if (m.bc[0] == BC_DIRICHLET) dwx = dwx + 2.0 * cW_raw;
      cW_visc = 0.0;
    }
    if (i == nx && bl) {
      if (m.bc[1] == BC_DIRICHLET) dwx = dwx + 2.0 * cE_raw;
      cE_visc = 0.0;
    }
    double conv_diag = 0.0, conv_W = 0.0, conv_E = 0.0, conv_S = 0.0, conv_N = 0.0, newt = 0.0;
    if (c.conv) {
      double FE, FW, FN, FS;
      v_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_v * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      conv_E = (i == nx && bl) ? 0.0 : rho_v * dmax(-FE, 0.0);
      conv_W = (i == 1 && bl) ? 0.0 : rho_v * dmax(FW, 0.0);
      conv_N = rho_v * dmax(-FN, 0.0);
      conv_S = rho_v * dmax(FS, 0.0);
      if (c.lin == 3) {
        const double dv_dy = (F.v[HX(i, j + 1)] - F.v[HX(i, j - 1)]) / (2.0 * dy);
        newt = rho_v * dmax(0.0, dv_dy) * dvol;
      }
    }
    PUT(vW, cW_visc + conv_W);
    PUT(vE, cE_visc + conv_E);
    PUT(vS, cS + conv_S + 0.0);
    PUT(vN, cN + conv_N + 0.0);
    double aP = Dv + dwx + conv_diag + 0.0 + newt + cW_visc + cE_visc + cS + cN + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    if (o.diag) o.diag[k] = aP;
    PUT(vP, aP);
    if (o.vaP2) o.vaP2[HX(i, j)] = aP;
  }
}
#undef PUT

// The Schur block of the monolithic system (same routine, :1540-1590): the
// 5-point coefficients of B diag(A)^-1 B^T from u_aP, v_aP (halo, exchanged),
// no coupling through wall sides, and the regularisation (1e-10 dvol at cell
// (1,1), 1e-12 dvol elsewhere).  diag: the pressure rows of diag_p.
__global__ void __launch_bounds__(BLK) k_schur(Mesh m, const double *__restrict__ uaP2,
                                               const double *__restrict__ vaP2, double *__restrict__ sW,
                                               double *__restrict__ sE, double *__restrict__ sS,
                                               double *__restrict__ sN, double *__restrict__ sP,
                                               double *__restrict__ diag) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= (long)m.nx * m.ny) return;
  const int j = (int)(k / m.nx) + 1, i = (int)(k - (long)(j - 1) * m.nx) + 1;
  const bool bl = !m.perx, bb = !m.pery;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  double aW = 0.0, aE = 0.0, aS = 0.0, aN = 0.0;
  if (!(i == 1 && bl && wallbc(m.bc[0]))) aW = (dy * dy) / dmax(1.0e-30, uaP2[HX(i, j)]);
  if (!(i == m.nx && bl && wallbc(m.bc[1]))) aE = (dy * dy) / dmax(1.0e-30, uaP2[HX(i + 1, j)]);
  if (!(j == 1 && bb && wallbc(m.bc[2]))) aS = (dx * dx) / dmax(1.0e-30, vaP2[HX(i, j)]);
  if (!(j == m.ny && bb && wallbc(m.bc[3]))) aN = (dx * dx) / dmax(1.0e-30, vaP2[HX(i, j + 1)]);
  const double reg = (i == 1 && j == 1) ? 1.0e-10 * dvol : 1.0e-12 * dvol;
  const double aP = aW + aE + aS + aN + 0.0 + reg;
  if (sW) {
    sW[k] = aW;
    sE[k] = aE;
    sS[k] = aS;
    sN[k] = aN;
    sP[k] = aP;
  }
  if (diag) diag[k] = aP;
}

__global__ void __launch_bounds__(BLK) k_fill(long n, double val, double *__restrict__ a) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k < n) a[k] = val;
}

// exchange_halos_scalar of up to two halo fields, one workgroup
__global__ void __launch_bounds__(1024) k_exchange2(Mesh m, double *a, double *b) {
  if (a) wg_exchange(m, a);
  if (b) wg_exchange(m, b);
}

// ----- the sequential preconditioners of apply_preconditioner
// (src/module_linear_solvers.f90:137-316) with ssor_solve_u / _v / _p
// (:1278-1403).  Lexicographic Gauss-Seidel is order dependent; it runs here
// in one workgroup per block along anti-diagonals i + j = const: a cell
// depends only on its W/S neighbours of the previous diagonal (already
// updated) and its E/N neighbours of the next one (not yet updated), so the
// wavefront reproduces the lexicographic sweep exactly, value for value.
struct Blk5 {
  const double *aW, *aE, *aS, *aN, *aP;
};

// The right-hand side of a block: the residual, or (kind 1: u, 2: v) the
// residual less the pressure gradient of the block-triangular step.
__device__ __forceinline__ double blk_rhs(const Mesh &m, const double *src, const double *xp, int kind, int i,
                                          int j, long dof, int ni, int nj) {
  const bool bl = !m.perx, bb = !m.pery;
  if (kind == 1) {
    const bool is_wall = (i == 1 && bl) || (i == ni && bl);
    return is_wall ? src[dof] : src[dof] - (xp[HX(i, j)] - xp[HX(i - 1, j)]) * m.dy;
  }
  if (kind == 2) {
    const bool is_wall = (j == 1 && bb && wallbc(m.bc[2])) || (j == nj && bb && wallbc(m.bc[3]));
    return is_wall ? src[dof] : src[dof] - (xp[HX(i, j)] - xp[HX(i, j - 1)]) * m.dx;
  }
  return src[dof];
}

// n_sweeps symmetric sweeps with relaxation omega, from zero, on the (ni, nj)
// block held in the halo field x; sync: the single-rank exchange of x before
// every sweep but the first and between the halves of a sweep.
__device__ void wg_ssor(const Mesh &m, int ni, int nj, Blk5 A, const double *src, const double *xp, int kind,
                        double *x, int n_sweeps, double omega, bool sync) {
  const int t = threadIdx.x, T = blockDim.x;
  const long nh = (long)(m.nx + 2) * (m.ny + 2);
  for (long k = t; k < nh; k += T) x[k] = 0.0;
  __syncthreads();
  auto cell = [&](int i, int j) {
    const long dof = (long)(j - 1) * ni + (i - 1);
    double d = A.aP[dof];
    if (fabs(d) < 1.0e-30) d = 1.0;
    double val = blk_rhs(m, src, xp, kind, i, j, dof, ni, nj);
    const double aw = A.aW[dof], ae = A.aE[dof], as = A.aS[dof], an = A.aN[dof];
    if (aw != 0.0) val = val + aw * x[HX(i - 1, j)];
    if (ae != 0.0) val = val + ae * x[HX(i + 1, j)];
    if (as != 0.0) val = val + as * x[HX(i, j - 1)];
    if (an != 0.0) val = val + an * x[HX(i, j + 1)];
    x[HX(i, j)] = (1.0 - omega) * x[HX(i, j)] + omega * (val / d);
  };
  for (int sweep = 1; sweep <= n_sweeps; ++sweep) {
    if (sync && sweep > 1) wg_exchange(m, x);
    for (int dg = 2; dg <= ni + nj; ++dg) {
      const int lo = dg - nj > 1 ? dg - nj : 1, hi = dg - 1 < ni ? dg - 1 : ni;
      for (int i = lo + t; i <= hi; i += T) cell(i, dg - i);
      __syncthreads();
    }
    if (sync) wg_exchange(m, x);
    for (int dg = ni + nj; dg >= 2; --dg) {
      const int lo = dg - nj > 1 ? dg - nj : 1, hi = dg - 1 < ni ? dg - 1 : ni;
      for (int i = hi - t; i >= lo; i -= T) cell(i, dg - i);
      __syncthreads();
    }
  }
}

// interior of the halo field x -> the block's part of dst
__device__ void wg_store(const Mesh &m, int ni, int nj, const double *x, double *dst) {
  const long nb = (long)ni * nj;
  for (long k = threadIdx.x; k < nb; k += blockDim.x) {
    const int j = (int)(k / ni) + 1, i = (int)(k - (long)(j - 1) * ni) + 1;
    dst[k] = x[HX(i, j)];
  }
}

struct SeqArgs {
  Blk5 blk[3];     // u, v, Schur
  int ni[3], nj[3];
  long off[3];     // offsets of the blocks in the vectors
  double *x[3];    // halo work fields
};

// Blocks 0, 1 (and 2, the pressure) of the SSOR preconditioner, one
// workgroup each (kind 0, 2 sweeps, omega 1), or the u and v solves of the
// block-triangular one (kind 1 / 2 with the pressure correction xp, 4 sweeps).
__global__ void __launch_bounds__(1024) k_pc_blocks(Mesh m, SeqArgs a, const double *__restrict__ src,
                                                    double *__restrict__ dst, int bt) {
  const int b = blockIdx.x;
  const bool sync = m.perx || m.pery;
  const int kind = bt ? b + 1 : 0;
  wg_ssor(m, a.ni[b], a.nj[b], a.blk[b], src + a.off[b], a.x[2], kind, a.x[b], bt ? 4 : 2, 1.0, sync);
  wg_store(m, a.ni[b], a.nj[b], a.x[b], dst + a.off[b]);
}

// The pressure step of the block-triangular preconditioner: 12 SSOR sweeps
// with omega 1.25 on the Schur block, the mean removed (serial sum in the
// reference's order), the halo exchanged.
__global__ void __launch_bounds__(1024) k_pc_bt_pressure(Mesh m, SeqArgs a, const double *__restrict__ src,
                                                         double *__restrict__ dst) {
  __shared__ double mean;
  const bool sync = m.perx || m.pery;
  double *xp = a.x[2];
  wg_ssor(m, m.nx, m.ny, a.blk[2], src + a.off[2], nullptr, 0, xp, 12, 1.25, sync);
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int j = 1; j <= m.ny; ++j)
      for (int i = 1; i <= m.nx; ++i) sum = sum + xp[HX(i, j)];
    mean = sum / (double)((long)m.nx * m.ny);
  }
  __syncthreads();
  const long np = (long)m.nx * m.ny;
  for (long k = threadIdx.x; k < np; k += blockDim.x) {
    const int j = (int)(k / m.nx) + 1, i = (int)(k - (long)(j - 1) * m.nx) + 1;
    xp[HX(i, j)] = xp[HX(i, j)] - mean;
  }
  __syncthreads();
  if (sync) wg_exchange(m, xp);
  wg_store(m, m.nx, m.ny, xp, dst + a.off[2]);
}

// ns_assemble_rhs, SP branch (src/module_navier_stokes.f90:785-1101)
// (the continuity rows of the monolithic system, k >= nuv, are zero)
__global__ void __launch_bounds__(BLK) k_mom_rhs(Mesh m, MomCfg c, MomFields F, double dt, long nu, long nuv,
                                                 long n, int u_nx, int v_ny, double *__restrict__ rhs) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  if (k >= nuv) {
    rhs[k] = 0.0;
    return;
  }
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  const double *u = F.u, *v = F.v;
  if (!d.isv) {
    const double rho_u = F.ru[HX(i, j)];
    const double diag_bc = (at * rho_u / dt) * dvol;
    if (i == 1 && bl) {
      rhs[k] = diag_bc * (m.bc[0] == BC_DIRICHLET ? c.bcu[0] : 0.0);
      return;
    }
    if (i == u_nx && bl) {
      rhs[k] = diag_bc * (m.bc[1] == BC_DIRICHLET ? c.bcu[1] : 0.0);
      return;
    }
    double val;
    if (c.gear) {
      const double ui = rho_u * (2.0 * F.uo[HX(i, j)] - 0.5 * F.un[HX(i, j)]);
      val = (ui / dt) * dvol;
    } else {
      val = (rho_u / dt) * F.uo[HX(i, j)] * dvol;
    }
    val = val + rho_u * c.gx * dvol;
    val = val + F.fx[HX(i, j)] * dvol;
    if (c.lin == 3) {
      const double du_dx = (u[HX(i + 1, j)] - u[HX(i - 1, j)]) / (2.0 * dx);
      const double du_dy = (u[HX(i, j + 1)] - u[HX(i, j - 1)]) / (2.0 * dy);
      const double vf = 0.25 * (v[HX(i - 1, j)] + v[HX(i, j)] + v[HX(i - 1, j + 1)] + v[HX(i, j + 1)]);
      val = val + rho_u * (u[HX(i, j)] * du_dx + vf * du_dy) * dvol;
    }
    if (j == ny && bb && m.bc[3] == BC_DIRICHLET) {
      const double mu_N = F.mn[HX(i, j + 1)];
      val = ny >= 2 ? val + mu_N * ((8.0 / 3.0) * c.bcu[3] / dy) * dx : val + mu_N * (2.0 * c.bcu[3] / dy) * dx;
    }
    if (j == 1 && bb && m.bc[2] == BC_DIRICHLET) {
      const double mu_S = F.mn[HX(i, j)];
      val = ny >= 2 ? val + mu_S * ((8.0 / 3.0) * c.bcu[2] / dy) * dx : val + mu_S * (2.0 * c.bcu[2] / dy) * dx;
    }
    rhs[k] = val;
  } else {
    const double rho_v = F.rv[HX(i, j)];
    const double diag_bc = (at * rho_v / dt) * dvol;
    if (j == 1 && bb && wallbc(m.bc[2])) {
      rhs[k] = diag_bc * (m.bc[2] == BC_DIRICHLET ? c.bcv[2] : 0.0);
      return;
    }
    if (j == v_ny && bb && wallbc(m.bc[3])) {
      rhs[k] = diag_bc * (m.bc[3] == BC_DIRICHLET ? c.bcv[3] : 0.0);
      return;
    }
    double val;
    if (c.gear) {
      const double vi = rho_v * (2.0 * F.vo[HX(i, j)] - 0.5 * F.vn[HX(i, j)]);
      val = (vi / dt) * dvol;
    } else {
      val = (rho_v / dt) * F.vo[HX(i, j)] * dvol;
    }
    val = val + rho_v * c.gy * dvol;
    val = val + F.fy[HX(i, j)] * dvol;
    if (c.lin == 3) {
      const double dv_dx = (v[HX(i + 1, j)] - v[HX(i - 1, j)]) / (2.0 * dx);
      const double dv_dy = (v[HX(i, j + 1)] - v[HX(i, j - 1)]) / (2.0 * dy);
      const double uf = 0.25 * (u[HX(i, j - 1)] + u[HX(i + 1, j - 1)] + u[HX(i, j)] + u[HX(i + 1, j)]);
      val = val + rho_v * (uf * dv_dx + v[HX(i, j)] * dv_dy) * dvol;
    }
    if (i == 1 && bl && m.bc[0] == BC_DIRICHLET) {
      const double mu_W = F.mn[HX(i, j)];
      val = val + mu_W * (2.0 * c.bcv[0] / dx) * dy;
    }
    if (i == nx && bl && m.bc[1] == BC_DIRICHLET) {
      const double mu_E = F.mn[HX(i + 1, j)];
      val = val + mu_E * (2.0 * c.bcv[1] / dx) * dy;
    }
    rhs[k] = val;
  }
}

// apply_preconditioner JACOBI (src/module_linear_solvers.f90:88-135)
__global__ void __launch_bounds__(BLK) k_mom_jacobi(long n, const double *__restrict__ d,
                                                    const double *__restrict__ r, double *__restrict__ z) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const double dk = d[k];
  z[k] = fabs(dk) > 1.0e-15 ? r[k] / dk : r[k];
}

#undef HX

inline dim3 grid_of(long n) { return dim3((unsigned)((n + BLK - 1) / BLK)); }

int mom_alloc(Ctx &c) {
  Mom &M = c.mom;
  if (M.tu) return COVOF_OK;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  // at most: u, v and p unknowns
  const long n = (long)(c.m.nx + 1) * c.m.ny + (long)c.m.nx * (c.m.ny + 1) + (long)c.m.nx * c.m.ny;
  for (double **f : {&M.tu, &M.tv, &M.tp}) {
    COVOF_CHECK(hipMalloc(f, nh * sizeof(double)));
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 0.0, *f);
  }
  COVOF_CHECK_LAUNCH();
  COVOF_CHECK(hipMalloc(&M.diag, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.rhs, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.x, n * sizeof(double)));
  for (double *&w : M.w) COVOF_CHECK(hipMalloc(&w, n * sizeof(double)));
  return COVOF_OK;
}

// the SSOR / Schur coefficients and their work fields, made on first use
int mom_alloc_blocks(Ctx &c) {
  Mom &M = c.mom;
  if (M.bk[0]) return COVOF_OK;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  const long nb = (long)(c.m.nx + 1) * (c.m.ny + 1);  // covers every block
  for (double *&b : M.bk) COVOF_CHECK(hipMalloc(&b, nb * sizeof(double)));
  for (double *&b : M.aP2) COVOF_CHECK(hipMalloc(&b, nh * sizeof(double)));
  for (double *&b : M.sx) COVOF_CHECK(hipMalloc(&b, nh * sizeof(double)));
  return COVOF_OK;
}

int mom_product(Ctx &c, const double *x, double *y, bool jacobi);

// The momentum (or monolithic) operator with its preconditioner as a LinOp;
// the Jacobi-preconditioned product is one fused kernel.
struct MomOp : LinOp {
  Ctx &c;
  int pc;
  MomOp(Ctx &cc, int precond) : c(cc), pc(precond) { n = cc.mom.n; }
  int apply(const double *x, double *y) override { return mom_matvec(c, x, y); }
  int apply_pa(const double *x, double *y, double *tmp) override {
    if (pc == COVOF_PRECOND_JACOBI) return mom_product(c, x, y, true);
    return LinOp::apply_pa(x, y, tmp);
  }
  int precond(const double *r, double *z) override { return mom_precond(c, pc, r, z); }
};

// assemble_preconditioner_blocks into M.diag and (full) the SSOR / Schur
// coefficients of M.bk
int mom_blocks(Ctx &c, bool full) {
  Mom &M = c.mom;
  const long nh = (long)(c.m.nx + 2) * (c.m.ny + 2);
  if (full || M.cfg.mono) {
    const int st = mom_alloc_blocks(c);
    if (st) return st;
  }
  BlkOut o{};
  o.diag = M.diag;
  if (full) {
    o.uW = M.bk[0], o.uE = M.bk[1], o.uS = M.bk[2], o.uN = M.bk[3], o.uP = M.bk[4];
    o.vW = M.bk[5], o.vE = M.bk[6], o.vS = M.bk[7], o.vN = M.bk[8], o.vP = M.bk[9];
  }
  if (M.cfg.mono) {
    o.uaP2 = M.aP2[0], o.vaP2 = M.aP2[1];
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 1.0, M.aP2[0]);
    hipLaunchKernelGGL(k_fill, grid_of(nh), dim3(BLK), 0, c.stream, nh, 1.0, M.aP2[1]);
  }
  hipLaunchKernelGGL(k_mom_blocks, grid_of(M.nuv), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.nuv,
                     M.u_nx, M.v_ny, o);
  if (M.cfg.mono) {
    hipLaunchKernelGGL(k_exchange2, dim3(1), dim3(1024), 0, c.stream, c.m, M.aP2[0], M.aP2[1]);
    double *s[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    if (full)
      for (int q = 0; q < 5; ++q) s[q] = M.bk[10 + q];
    hipLaunchKernelGGL(k_schur, grid_of((long)c.m.nx * c.m.ny), dim3(BLK), 0, c.stream, c.m,
                       (const double *)M.aP2[0], (const double *)M.aP2[1], s[0], s[1], s[2], s[3], s[4],
                       M.diag + M.nuv);
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

}  // namespace

void mom_free(Ctx &c) {
  Mom &M = c.mom;
  for (double *p : {M.tu, M.tv, M.tp, M.diag, M.rhs, M.x}) hipFree(p);
  for (double *p : M.w) hipFree(p);
  for (double *p : M.bk) hipFree(p);
  for (double *p : M.aP2) hipFree(p);
  for (double *p : M.sx) hipFree(p);
  M = Mom();
}

int mom_bind(Ctx &c, const MomCfg &cfg, const MomFields &f, double dt) {
  const int *bc = c.m.bc;
  if (!(dt > 0.0) || !f.u || !f.v || !f.uo || !f.vo || !f.ru || !f.rv || !f.mu || !f.mn || !f.fx || !f.fy)
    return COVOF_BAD_ARGUMENT;
  if ((cfg.gear || cfg.ab_nm1) && (!f.un || !f.vn)) return COVOF_BAD_ARGUMENT;
  if (cfg.lin < 1 || cfg.lin > 3 || cfg.scheme < 1 || cfg.scheme > 3) return COVOF_BAD_ARGUMENT;
  // periodic sides come in pairs; the reference's v-row at j = ny+1 reads past
  // its array with a NEUMANN top, so that side is not supported
  if ((bc[0] == BC_PERIODIC) != (bc[1] == BC_PERIODIC) || (bc[2] == BC_PERIODIC) != (bc[3] == BC_PERIODIC) ||
      bc[3] == BC_NEUMANN)
    return COVOF_BAD_ARGUMENT;
  const int st = mom_alloc(c);
  if (st) return st;
  Mom &M = c.mom;
  M.cfg = cfg;
  M.f = f;
  M.dt = dt;
  M.u_nx = c.m.nx + (c.m.perx ? 0 : 1);
  M.v_ny = c.m.ny + (c.m.pery ? 0 : 1);
  M.nu = (long)M.u_nx * c.m.ny;
  M.nuv = M.nu + (long)c.m.nx * M.v_ny;
  M.n = M.nuv + (cfg.mono ? (long)c.m.nx * c.m.ny : 0);
  M.bound = 1;
  return COVOF_OK;
}

int mom_diag(Ctx &c, double *diag) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  int st = mom_blocks(c, false);
  if (st) return st;
  if (diag != M.diag)
    COVOF_CHECK(hipMemcpyAsync(diag, M.diag, M.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
  return COVOF_OK;
}

int mom_rhs(Ctx &c, double *rhs) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_mom_rhs, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.nuv, M.n,
                     M.u_nx, M.v_ny, rhs);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int mom_pack(Ctx &c, const double *u, const double *v, const double *p, double *x) {
  Mom &M = c.mom;
  if (!M.bound || (M.cfg.mono && !p)) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_pack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx, u, v, p, x);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

namespace {
// unpack_vector(homogeneous) into the work fields, then the operator (with
// the Jacobi division when jacobi: M.diag must hold the diagonal).
int mom_product(Ctx &c, const double *x, double *y, bool jacobi) {
  Mom &M = c.mom;
  double *tp = M.cfg.mono ? M.tp : nullptr;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx, x, M.tu,
                     M.tv, tp);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, M.tu, M.tv, 1, 1, tp);
  launch_tagged(c, COVOF_KERNEL_MOM_MATVEC, jacobi ? k_mom_matvec<true> : k_mom_matvec<false>, grid_of(M.n),
                dim3(BLK), c.m, M.cfg, M.f, M.dt, M.nu, M.nuv, M.n, M.u_nx, M.v_ny, (const double *)M.tu,
                (const double *)M.tv, (const double *)M.tp, (const double *)M.diag, y);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// the blocks of the sequential preconditioners, as the kernels take them
SeqArgs seq_args(Ctx &c) {
  Mom &M = c.mom;
  SeqArgs a{};
  for (int b = 0; b < 3; ++b)
    a.blk[b] = Blk5{M.bk[5 * b], M.bk[5 * b + 1], M.bk[5 * b + 2], M.bk[5 * b + 3], M.bk[5 * b + 4]};
  a.ni[0] = M.u_nx, a.nj[0] = c.m.ny, a.off[0] = 0;
  a.ni[1] = c.m.nx, a.nj[1] = M.v_ny, a.off[1] = M.nu;
  a.ni[2] = c.m.nx, a.nj[2] = c.m.ny, a.off[2] = M.nuv;
  for (int b = 0; b < 3; ++b) a.x[b] = M.sx[b];
  return a;
}
}  // namespace

int mom_matvec(Ctx &c, const double *x, double *y) {
  if (!c.mom.bound) return COVOF_BAD_ARGUMENT;
  return mom_product(c, x, y, false);
}

// apply_preconditioner (src/module_linear_solvers.f90:48-316) with the blocks
// last assembled: NONE copies, JACOBI divides by the diagonal, SSOR runs two
// symmetric sweeps per block, BLOCK_TRIANGULAR (monolithic) the pressure
// Schur solve then the velocity solves against its gradient (SSOR otherwise).
int mom_precond(Ctx &c, int precond, const double *r, double *z) {
  Mom &M = c.mom;
  if (precond == COVOF_PRECOND_JACOBI) {
    hipLaunchKernelGGL(k_mom_jacobi, grid_of(M.n), dim3(BLK), 0, c.stream, M.n, (const double *)M.diag, r, z);
  } else if (precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR) {
    const SeqArgs a = seq_args(c);
    if (precond == COVOF_PRECOND_BLOCK_TRIANGULAR && M.cfg.mono) {
      hipLaunchKernelGGL(k_pc_bt_pressure, dim3(1), dim3(1024), 0, c.stream, c.m, a, r, z);
      hipLaunchKernelGGL(k_pc_blocks, dim3(2), dim3(1024), 0, c.stream, c.m, a, r, z, 1);
    } else {
      hipLaunchKernelGGL(k_pc_blocks, dim3(M.cfg.mono ? 3 : 2), dim3(1024), 0, c.stream, c.m, a, r, z, 0);
    }
  } else {
    if (z != r) COVOF_CHECK(hipMemcpyAsync(z, r, M.n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return COVOF_OK;
  }
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

static bool precond_ok(int precond) {
  return precond == COVOF_PRECOND_NONE || precond == COVOF_PRECOND_JACOBI || precond == COVOF_PRECOND_SSOR ||
         precond == COVOF_PRECOND_BLOCK_TRIANGULAR;
}

// assemble_preconditioner_blocks of the bound fields, then apply_preconditioner
int mom_apply_precond(Ctx &c, int precond, const double *r, double *z) {
  if (!c.mom.bound || !precond_ok(precond)) return COVOF_BAD_ARGUMENT;
  const int st = mom_blocks(c, precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR);
  return st ? st : mom_precond(c, precond, r, z);
}

// solve_linear_system on the momentum (or monolithic) operator, the
// preconditioner from the blocks of the bound fields.  b, x: n unknowns, not
// a work vector.
int mom_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol, int maxit, int *iters,
              double *relres) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  if ((lin_solver != COVOF_LINEAR_BICGSTAB2 && lin_solver != COVOF_LINEAR_CG) || !precond_ok(precond) ||
      maxit < 0)
    return COVOF_BAD_ARGUMENT;
  int st = mom_blocks(c, precond == COVOF_PRECOND_SSOR || precond == COVOF_PRECOND_BLOCK_TRIANGULAR);
  if (st) return st;
  MomOp op(c, precond);
  if (lin_solver == COVOF_LINEAR_CG) return krylov_cg(c, op, b, x, M.w, tol, maxit, iters, relres);
  return krylov_bicgstab2(c, op, b, x, M.w, tol, maxit, iters, relres);
}

// sync_u, sync_v and apply_boundary_conditions of one rank
// (src/module_mpi_comm.f90:733-775, :290-537) on halo fields u, v.
int velocity_bc(Ctx &c, const MomCfg &cfg, double *u, double *v, int homogeneous) {
  const int *bc = c.m.bc;
  if ((bc[0] == BC_PERIODIC) != (bc[1] == BC_PERIODIC) || (bc[2] == BC_PERIODIC) != (bc[3] == BC_PERIODIC))
    return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, cfg, u, v, homogeneous ? 1 : 0, 1,
                     (double *)nullptr);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

// The momentum block of sim_solve_standard_projection (:2454-2469): the
// current u, v are the initial guess and the operator's fields_old; the
// solution is unpacked into them with the inhomogeneous BC, then synced and
// the BC applied once more, as the reference does.
int mom_predictor(Ctx &c, double *u, double *v, int lin_solver, int precond, double tol, int maxit, int *iters,
                  double *relres) {
  Mom &M = c.mom;
  if (!M.bound || M.cfg.mono) return COVOF_BAD_ARGUMENT;
  int st = mom_rhs(c, M.rhs);
  if (!st) st = mom_pack(c, u, v, nullptr, M.x);
  if (st) return st;
  const int sst = mom_solve(c, M.rhs, M.x, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx,
                     (const double *)M.x, u, v, (double *)nullptr);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, (double *)nullptr);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, (double *)nullptr);
  COVOF_CHECK_LAUNCH();
  return sst;
}

// The monolithic branch of the time step (src/module_simulation.f90:553-583,
// one Uzawa pass): blocks, right-hand side with zero continuity rows, pack
// (u, v, p: the initial guess), solve, unpack_vector(.false.) into u, v, p
// (pressure exchanged, velocity exchanged and its BC applied).
int mono_step(Ctx &c, double *u, double *v, double *p, int lin_solver, int precond, double tol, int maxit,
              int *iters, double *relres) {
  Mom &M = c.mom;
  if (!M.bound || !M.cfg.mono || !p) return COVOF_BAD_ARGUMENT;
  int st = mom_rhs(c, M.rhs);
  if (!st) st = mom_pack(c, u, v, p, M.x);
  if (st) return st;
  const int sst = mom_solve(c, M.rhs, M.x, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.nuv, M.n, M.u_nx,
                     (const double *)M.x, u, v, p);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1, p);
  COVOF_CHECK_LAUNCH();
  return sst;
}

}  // namespace covof
