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

// x -> interior unknowns of the halo fields u, v (ns_unpack_vector's copy loops)
__global__ void __launch_bounds__(BLK) k_unpack(Mesh m, long nu, long n, int u_nx, const double *__restrict__ x,
                                                double *__restrict__ u, double *__restrict__ v) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  (d.isv ? v : u)[HX(d.i, d.j)] = x[k];
}

// ns_pack_vector
__global__ void __launch_bounds__(BLK) k_pack(Mesh m, long nu, long n, int u_nx, const double *__restrict__ u,
                                              const double *__restrict__ v, double *__restrict__ x) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  x[k] = (d.isv ? v : u)[HX(d.i, d.j)];
}

// sync_u, sync_v (exchange_halos_scalar on one rank: periodic columns over
// all rows, then periodic rows over all columns) and apply_velocity_bc, in
// the reference's statement order: one workgroup, a barrier between every
// step whose reads depend on an earlier step's writes.
__global__ void __launch_bounds__(1024) k_vel_halo(Mesh m, MomCfg c, double *__restrict__ u,
                                                   double *__restrict__ v, int homogeneous, int exchange) {
  const int nx = m.nx, ny = m.ny, t = threadIdx.x, T = blockDim.x;
  const int *bc = m.bc;
  const bool bl = !m.perx, bb = !m.pery;  // physical sides (single rank)
  if (exchange && m.perx) {
    for (int j = t; j <= ny + 1; j += T) {
      u[HX(0, j)] = u[HX(nx, j)];
      v[HX(0, j)] = v[HX(nx, j)];
    }
    __syncthreads();
    for (int j = t; j <= ny + 1; j += T) {
      u[HX(nx + 1, j)] = u[HX(1, j)];
      v[HX(nx + 1, j)] = v[HX(1, j)];
    }
    __syncthreads();
  }
  if (exchange && m.pery) {
    for (int i = t; i <= nx + 1; i += T) {
      u[HX(i, 0)] = u[HX(i, ny)];
      v[HX(i, 0)] = v[HX(i, ny)];
    }
    __syncthreads();
    for (int i = t; i <= nx + 1; i += T) {
      u[HX(i, ny + 1)] = u[HX(i, 1)];
      v[HX(i, ny + 1)] = v[HX(i, 1)];
    }
    __syncthreads();
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
                                          const double *__restrict__ tv) {
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
    return val;
  }
#undef TU
#undef TV
}

// y = A x, one unknown per thread; with JAC, y = M^-1 A x for the Jacobi
// preconditioner (apply_PA, src/module_linear_solvers.f90:323-356): the same
// row value divided as apply_preconditioner divides it.
template <bool JAC>
__global__ void __launch_bounds__(BLK) k_mom_matvec(Mesh m, MomCfg c, MomFields F, double dt, long nu, long n,
                                                    int u_nx, int v_ny, const double *__restrict__ tu,
                                                    const double *__restrict__ tv, const double *__restrict__ diag,
                                                    double *__restrict__ y) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const double val = mom_row(m, c, F, dt, k, nu, u_nx, v_ny, tu, tv);
  if (JAC) {
    const double dk = diag[k];
    y[k] = fabs(dk) > 1.0e-15 ? val / dk : val;
  } else {
    y[k] = val;
  }
}

// ns_assemble_preconditioner_blocks: diag_u, diag_v
// (src/module_navier_stokes.f90:1268-1595; no AL, Brinkman or obstacle terms)
__global__ void __launch_bounds__(BLK) k_mom_diag(Mesh m, MomCfg c, MomFields F, double dt, long nu, long n,
                                                  int u_nx, int v_ny, double *__restrict__ diag) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
  const Dof d = dof_ij(k, nu, u_nx, m.nx);
  const int i = d.i, j = d.j, nx = m.nx, ny = m.ny;
  const double dx = m.dx, dy = m.dy, dvol = dx * dy;
  const double dt_safe = dt > 1.0e-30 ? dt : 1.0;
  const double at = c.gear ? 1.5 : 1.0;
  const bool bl = !m.perx, bb = !m.pery;
  if (!d.isv) {
    const double rho_u = F.ru[HX(i, j)];
    double Du = (at * rho_u / dt_safe) * dvol;
    if (Du <= 1.0e-30) Du = 1.0;
    if ((i == 1 && bl) || (i == u_nx && bl)) {
      diag[k] = Du;
      return;
    }
    const double mu_e = F.mu[HX(i, j)], mu_w = F.mu[HX(i - 1, j)];
    const double cE = 2.0 * mu_e * (dy / dx), cW = 2.0 * mu_w * (dy / dx);
    const double cN_raw = F.mn[HX(i, j + 1)] * (dx / dy), cS_raw = F.mn[HX(i, j)] * (dx / dy);
    double cN_visc = cN_raw, cS_visc = cS_raw, dwy = 0.0;
    if (j == 1 && bb) {
      if (m.bc[2] == BC_DIRICHLET) dwy = ny >= 2 ? dwy + 3.0 * cS_raw : dwy + 2.0 * cS_raw;
      cS_visc = 0.0;
    }
    if (j == ny && bb) {
      if (m.bc[3] == BC_DIRICHLET) dwy = ny >= 2 ? dwy + 3.0 * cN_raw : dwy + 2.0 * cN_raw;
      cN_visc = 0.0;
    }
    double conv_diag = 0.0, newt = 0.0;
    if (c.conv) {
      double FE, FW, FN, FS;
      u_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_u * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      if (c.lin == 3) {
        const double du_dx = (F.u[HX(i + 1, j)] - F.u[HX(i - 1, j)]) / (2.0 * dx);
        newt = rho_u * dmax(0.0, du_dx) * dvol;
      }
    }
    double aP = Du + dwy + conv_diag + 0.0 + newt + cW + cE + cS_visc + cN_visc + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    diag[k] = aP;
  } else {
    const double rho_v = F.rv[HX(i, j)];
    double Dv = (at * rho_v / dt_safe) * dvol;
    if (Dv <= 1.0e-30) Dv = 1.0;
    const bool is_wall = (j == 1 && bb && wallbc(m.bc[2])) || (j == v_ny && bb && wallbc(m.bc[3]));
    if (is_wall || (j == 1 && bb && m.bc[2] == BC_NEUMANN && v_ny >= 2) ||
        (j == v_ny && bb && m.bc[3] == BC_NEUMANN && v_ny >= 2)) {
      diag[k] = Dv;
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
    double conv_diag = 0.0, newt = 0.0;
    if (c.conv) {
      double FE, FW, FN, FS;
      v_fluxes(m, c, F, i, j, FE, FW, FN, FS);
      conv_diag = rho_v * (dmax(FE, 0.0) + dmax(-FW, 0.0) + dmax(FN, 0.0) + dmax(-FS, 0.0));
      if (c.lin == 3) {
        const double dv_dy = (F.v[HX(i, j + 1)] - F.v[HX(i, j - 1)]) / (2.0 * dy);
        newt = rho_v * dmax(0.0, dv_dy) * dvol;
      }
    }
    double aP = Dv + dwx + conv_diag + 0.0 + newt + cW_visc + cE_visc + cS + cN + 0.0 + 0.0;
    if (aP <= 1.0e-15) aP = 1.0;
    diag[k] = aP;
  }
}

// ns_assemble_rhs, SP branch (src/module_navier_stokes.f90:785-1101)
__global__ void __launch_bounds__(BLK) k_mom_rhs(Mesh m, MomCfg c, MomFields F, double dt, long nu, long n,
                                                 int u_nx, int v_ny, double *__restrict__ rhs) {
  const long k = (long)blockIdx.x * BLK + threadIdx.x;
  if (k >= n) return;
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
  const long n = (long)(c.m.nx + 1) * c.m.ny + (long)c.m.nx * (c.m.ny + 1);  // at most
  COVOF_CHECK(hipMalloc(&M.tu, nh * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.tv, nh * sizeof(double)));
  COVOF_CHECK(hipMemsetAsync(M.tu, 0, nh * sizeof(double), c.stream));
  COVOF_CHECK(hipMemsetAsync(M.tv, 0, nh * sizeof(double), c.stream));
  COVOF_CHECK(hipMalloc(&M.diag, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.rhs, n * sizeof(double)));
  COVOF_CHECK(hipMalloc(&M.x, n * sizeof(double)));
  for (double *&w : M.w) COVOF_CHECK(hipMalloc(&w, n * sizeof(double)));
  return COVOF_OK;
}

int mom_product(Ctx &c, const double *x, double *y, bool jacobi);

// The momentum operator with its NONE / JACOBI preconditioner as a LinOp;
// the preconditioned product is one fused kernel.
struct MomOp : LinOp {
  Ctx &c;
  int pc;
  MomOp(Ctx &cc, int precond) : c(cc), pc(precond) { n = cc.mom.n; }
  int apply(const double *x, double *y) override { return mom_matvec(c, x, y); }
  int apply_pa(const double *x, double *y, double *tmp) override {
    if (pc == COVOF_PRECOND_JACOBI) return mom_product(c, x, y, true);
    return LinOp::apply_pa(x, y, tmp);
  }
  int precond(const double *r, double *z) override {
    if (pc == COVOF_PRECOND_JACOBI) {
      hipLaunchKernelGGL(k_mom_jacobi, grid_of(n), dim3(BLK), 0, c.stream, n, c.mom.diag, r, z);
      return hipGetLastError() == hipSuccess ? COVOF_OK : COVOF_HIP_ERROR;
    }
    if (z != r) COVOF_CHECK(hipMemcpyAsync(z, r, n * sizeof(double), hipMemcpyDeviceToDevice, c.stream));
    return COVOF_OK;
  }
};

}  // namespace

void mom_free(Ctx &c) {
  Mom &M = c.mom;
  for (double *p : {M.tu, M.tv, M.diag, M.rhs, M.x}) hipFree(p);
  for (double *p : M.w) hipFree(p);
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
  M.n = M.nu + (long)c.m.nx * M.v_ny;
  M.bound = 1;
  return COVOF_OK;
}

int mom_diag(Ctx &c, double *diag) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_mom_diag, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.n, M.u_nx,
                     M.v_ny, diag);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int mom_rhs(Ctx &c, double *rhs) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_mom_rhs, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.cfg, M.f, M.dt, M.nu, M.n, M.u_nx,
                     M.v_ny, rhs);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

int mom_pack(Ctx &c, const double *u, const double *v, double *x) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  hipLaunchKernelGGL(k_pack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.n, M.u_nx, u, v, x);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}

namespace {
// unpack_vector(homogeneous) into the work fields, then the operator (with
// the Jacobi division when jacobi: M.diag must hold the diagonal).
int mom_product(Ctx &c, const double *x, double *y, bool jacobi) {
  Mom &M = c.mom;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.n, M.u_nx, x, M.tu, M.tv);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, M.tu, M.tv, 1, 1);
  launch_tagged(c, COVOF_KERNEL_MOM_MATVEC, jacobi ? k_mom_matvec<true> : k_mom_matvec<false>, grid_of(M.n),
                dim3(BLK), c.m, M.cfg, M.f, M.dt, M.nu, M.n, M.u_nx, M.v_ny, (const double *)M.tu,
                (const double *)M.tv, (const double *)M.diag, y);
  COVOF_CHECK_LAUNCH();
  return COVOF_OK;
}
}  // namespace

int mom_matvec(Ctx &c, const double *x, double *y) {
  if (!c.mom.bound) return COVOF_BAD_ARGUMENT;
  return mom_product(c, x, y, false);
}

// solve_linear_system on the momentum operator, the diagonal of the bound
// fields as the Jacobi preconditioner.  b, x: n unknowns, not a work vector.
int mom_solve(Ctx &c, const double *b, double *x, int lin_solver, int precond, double tol, int maxit, int *iters,
              double *relres) {
  Mom &M = c.mom;
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  if ((lin_solver != COVOF_LINEAR_BICGSTAB2 && lin_solver != COVOF_LINEAR_CG) ||
      (precond != COVOF_PRECOND_NONE && precond != COVOF_PRECOND_JACOBI) || maxit < 0)
    return COVOF_BAD_ARGUMENT;
  int st = mom_diag(c, M.diag);
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
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, cfg, u, v, homogeneous ? 1 : 0, 1);
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
  if (!M.bound) return COVOF_BAD_ARGUMENT;
  int st = mom_rhs(c, M.rhs);
  if (!st) st = mom_pack(c, u, v, M.x);
  if (st) return st;
  const int sst = mom_solve(c, M.rhs, M.x, lin_solver, precond, tol, maxit, iters, relres);
  if (sst == COVOF_BAD_ARGUMENT || sst == COVOF_HIP_ERROR) return sst;
  hipLaunchKernelGGL(k_unpack, grid_of(M.n), dim3(BLK), 0, c.stream, c.m, M.nu, M.n, M.u_nx, (const double *)M.x,
                     u, v);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1);
  hipLaunchKernelGGL(k_vel_halo, dim3(1), dim3(1024), 0, c.stream, c.m, M.cfg, u, v, 0, 1);
  COVOF_CHECK_LAUNCH();
  return sst;
}

}  // namespace covof
