"""Dense float64 torch restatement of the two-phase step reduced to
transport, density update, inertial predictor and projection, differentiated
by torch autograd: the reference gradient of the library's multi-step
kinetic-energy adjoint.  Test infrastructure only (CPU or GPU tensors).

* Weymouth-Yue transport (src/module_interface_vof.f90:1162-1377): two split
  passes, the dilation term with c_d = [C^n >= 0.5], the second pass clamped
  to [0, 1]; the periodic face copy (:1223-1231, :1268-1275); the halo of
  apply_vof_bc (src/module_mpi_comm.f90:636-683) with the two-phase corner
  fix (src/module_interface_vof.f90:532-535).
* Reconstruction (src/module_interface_reconstruction.f90): Youngs normals
  (:313-320), ELVIRA's seven candidates and first-minimum choice (:365-440),
  ghost normals as the renormalised normal of the cell the halo copies
  (:602-681).
* Cut areas (:33-101) by Green's theorem over the clipped box (the same
  polygon as the reference's Sutherland-Hodgman clip, summed edge by edge);
  the line constant (:105-147) as the exact root of the volume equation:
  bisection to machine precision without gradient, then one Newton step,
  which carries the implicit-function derivative.
* The density of update_physical_properties (:552-553) and its face means;
  the SP Poisson block (src/module_simulation.f90:2477-2588) as a dense
  solve; the correction (:2590-2654); the history row's kinetic energy
  (src/module_vtk_writer.f90:399-531).

Every branch (donor, class of a cell, ELVIRA's choice and sign factors,
clamps, c_d) is taken from this model's own forward values, so autograd
differentiates the same piecewise-smooth map as the library's adjoint.
"""
from __future__ import annotations

import torch

MIX = 1.0e-6
F64 = torch.float64


def is_mixed(c):
    return ~(c <= MIX) & ~(c >= 1.0 - MIX)


class Grid:
    """nx x ny cells on [0, lx] x [0, ly], walls on every side or periodic
    in both directions; halo layout (ny + 2, nx + 2), index [j, i]."""

    def __init__(self, nx, ny, lx=1.0, ly=1.0, periodic=False, device="cpu"):
        self.nx, self.ny = nx, ny
        self.dx, self.dy = lx / nx, ly / ny
        self.per = periodic
        self.dev = device
        ii = torch.arange(nx + 2, dtype=F64, device=device)
        jj = torch.arange(ny + 2, dtype=F64, device=device)
        self.X0 = ((ii - 1.0) * self.dx)[None, :].expand(ny + 2, nx + 2)
        self.X1 = (ii * self.dx)[None, :].expand(ny + 2, nx + 2)
        self.Y0 = ((jj - 1.0) * self.dy)[:, None].expand(ny + 2, nx + 2)
        self.Y1 = (jj * self.dy)[:, None].expand(ny + 2, nx + 2)

        def src(n):  # wrap_or_clamp of every halo index
            s = torch.arange(n + 2, device=device)
            s[0] = n if periodic else 1
            s[n + 1] = 1 if periodic else n
            return s
        self.si = src(nx)[None, :].expand(ny + 2, nx + 2)
        self.sj = src(ny)[:, None].expand(ny + 2, nx + 2)
        ring = torch.ones(ny + 2, nx + 2, dtype=torch.bool, device=device)
        ring[1:ny + 1, 1:nx + 1] = False
        self.ring = ring
        corner = torch.zeros_like(ring)
        for a in (0, ny + 1):
            for b in (0, nx + 1):
                corner[a, b] = True
        self.corner = corner

    def copy_halo(self, F):
        """Halo entries copied from the interior (wrap or zero-gradient,
        corners from the diagonal cell)."""
        return F[self.sj, self.si]


# ---------------------------------------------------------------- geometry

def cut_area(x0, x1, y0, y1, nx, ny, al):
    """Area of {nx x + ny y + al <= 0} in the box and the length of the
    chord of the line inside it (both broadcast over the inputs)."""
    px = (x0, x1, x1, x0)
    py = (y0, y0, y1, y1)
    d = [nx * px[k] + ny * py[k] + al for k in range(4)]
    ins = [dk <= 0.0 for dk in d]
    zero = torch.zeros_like(d[0])
    S = zero
    ex = ey = enx = eny = zero
    has = torch.zeros_like(ins[0])
    for k in range(4):
        k1 = (k + 1) % 4
        cross = ins[k] != ins[k1]
        den = torch.where(cross, d[k] - d[k1], torch.ones_like(d[k]))
        t = d[k] / den
        X = px[k] + t * (px[k1] - px[k])
        Y = py[k] + t * (py[k1] - py[k])
        both = ins[k] & ins[k1]
        S = S + torch.where(both, px[k] * py[k1] - px[k1] * py[k], zero)
        io = ins[k] & ~ins[k1]
        S = S + torch.where(io, px[k] * Y - X * py[k], zero)
        ex = torch.where(io, X, ex)
        ey = torch.where(io, Y, ey)
        oi = ~ins[k] & ins[k1]
        S = S + torch.where(oi, X * py[k1] - px[k1] * Y, zero)
        enx = torch.where(oi, X, enx)
        eny = torch.where(oi, Y, eny)
        has = has | io
    S = S + torch.where(has, ex * eny - enx * ey, zero)
    L2 = torch.where(has, (ex - enx) ** 2 + (ey - eny) ** 2, torch.ones_like(S))
    L = torch.where(has, torch.sqrt(L2), zero)
    return 0.5 * S, L


def box_area(x0, x1, y0, y1):
    """The whole box's area as the reference's clip sums it for a full
    donor: the shoelace terms of its corners in absolute coordinates."""
    s = (x1 * y1 - x1 * y0) + (x1 * y1 - x0 * y1)
    s = s + (x0 * y0 - x0 * y1) + (x0 * y0 - x1 * y0)
    return 0.5 * s


def alpha_of(x0, x1, y0, y1, nx, ny, c):
    """Line constant with cut area c |box| for the normal (nx, ny)."""
    box = (x1 - x0) * (y1 - y0)
    with torch.no_grad():
        w = torch.stack([nx * x0 + ny * y0, nx * x1 + ny * y0, nx * x1 + ny * y1, nx * x0 + ny * y1])
        lo, hi = -w.max(0).values, -w.min(0).values  # full at lo, empty at hi
        target = c * box
        for _ in range(110):
            mid = 0.5 * (lo + hi)
            A, _ = cut_area(x0, x1, y0, y1, nx, ny, mid)
            big = A > target
            lo = torch.where(big, mid, lo)
            hi = torch.where(big, hi, mid)
        a0 = 0.5 * (lo + hi)
        _, L = cut_area(x0, x1, y0, y1, nx.detach(), ny.detach(), a0)
        L = torch.where(L > 0.0, L, torch.ones_like(L))
    A, _ = cut_area(x0, x1, y0, y1, nx, ny, a0)
    return a0 + (A - c * box) / L


def unit(mx, my):
    """(mx, my) / |(mx, my)| where the norm exceeds 1e-12, else (1, 0); and
    that mask."""
    n2 = mx * mx + my * my
    ok = n2 > 1.0e-24
    nr = torch.sqrt(torch.where(ok, n2, torch.ones_like(n2)))
    return torch.where(ok, mx / nr, torch.ones_like(mx)), torch.where(ok, my / nr, torch.zeros_like(my)), ok


# ---------------------------------------------------------- reconstruction

def _stencil(C, g):
    ny, nx = g.ny, g.nx
    return lambda di, dj: C[1 + dj:ny + 1 + dj, 1 + di:nx + 1 + di]


def youngs(C, g):
    c = _stencil(C, g)
    mx = -(((c(1, 1) + 2.0 * c(1, 0) + c(1, -1)) - (c(-1, 1) + 2.0 * c(-1, 0) + c(-1, -1))) / (8.0 * g.dx))
    my = -(((c(1, 1) + 2.0 * c(0, 1) + c(-1, 1)) - (c(1, -1) + 2.0 * c(0, -1) + c(-1, -1))) / (8.0 * g.dy))
    return mx, my


def elvira(C, g):
    """Unit normal of the chosen candidate per interior cell and the mask of
    cells with a valid candidate.  The value is the first-minimum winner's;
    the derivative is that of the lowest-index candidate tying with it to
    rounding (error within 1e-10 relative, unit normal within 1e-9)."""
    c = _stencil(C, g)
    dx, dy = g.dx, g.dy
    with torch.no_grad():
        sgn_y = torch.where(c(0, 1) < c(0, -1), 1.0, -1.0).to(F64)
        sgn_x = torch.where(c(1, 0) < c(-1, 0), 1.0, -1.0).to(F64)
    Hm = (c(-1, -1) + c(-1, 0) + c(-1, 1)) * dy
    H0 = (c(0, -1) + c(0, 0) + c(0, 1)) * dy
    Hp = (c(1, -1) + c(1, 0) + c(1, 1)) * dy
    Wm = (c(-1, -1) + c(0, -1) + c(1, -1)) * dx
    W0 = (c(-1, 0) + c(0, 0) + c(1, 0)) * dx
    Wp = (c(-1, 1) + c(0, 1) + c(1, 1)) * dx
    ym, ymy = youngs(C, g)
    cands = [(-(sgn_y * (Hp - Hm) / (2.0 * dx)), sgn_y * 1.0), (-(sgn_y * (Hp - H0) / dx), sgn_y * 1.0),
             (-(sgn_y * (H0 - Hm) / dx), sgn_y * 1.0), (sgn_x * 1.0, -(sgn_x * (Wp - Wm) / (2.0 * dy))),
             (sgn_x * 1.0, -(sgn_x * (Wp - W0) / dy)), (sgn_x * 1.0, -(sgn_x * (W0 - Wm) / dy)), (ym, ymy)]
    I = (slice(1, g.ny + 1), slice(1, g.nx + 1))
    x0, x1, y0, y1 = g.X0[I], g.X1[I], g.Y0[I], g.Y1[I]
    cc = C[I].detach()
    cs = torch.where(is_mixed(cc), cc, torch.full_like(cc, 0.5))
    errs = []
    with torch.no_grad():
        for cx, cy in cands:
            ux, uy, ok = unit(cx.detach(), cy.detach())
            al = alpha_of(x0, x1, y0, y1, ux, uy, cs)
            e = torch.zeros_like(cc)
            for dj in (-1, 0, 1):
                for di in (-1, 0, 1):
                    if di == 0 and dj == 0:
                        continue
                    J = (slice(1 + dj, g.ny + 1 + dj), slice(1 + di, g.nx + 1 + di))
                    vr, _ = cut_area(g.X0[J], g.X1[J], g.Y0[J], g.Y1[J], ux, uy, al)
                    df = vr - C[J].detach() * dx * dy
                    e = e + df * df
            errs.append(torch.where(ok & (e < 1.0e30), e, torch.full_like(e, float("inf"))))
        E = torch.stack(errs)
        best = E.argmin(0)
        Emin = E.min(0).values
        valid = torch.isfinite(Emin)
    units = [unit(cx, cy) for cx, cy in cands]
    UX = torch.stack([a for a, _, _ in units])
    UY = torch.stack([b for _, b, _ in units])
    with torch.no_grad():
        bx = UX.gather(0, best[None])
        by = UY.gather(0, best[None])
        tol = 1.0e-10 * Emin + 1.0e-20 * (dx * dy) ** 2
        tie = torch.isfinite(E) & (E <= Emin + tol) & ((UX - bx).abs() + (UY - by).abs() <= 1.0e-9)
        adj = torch.where(valid, tie.to(torch.int64).argmax(0), best)  # first tying index
    ax = UX.gather(0, adj[None]).squeeze(0)
    ay = UY.gather(0, adj[None]).squeeze(0)
    ux = bx.squeeze(0) + (ax - ax.detach())
    uy = by.squeeze(0) + (ay - ay.detach())
    return ux, uy, valid


def reconstruct(C, g, method):
    """Segments (nx, ny, alpha) of every halo-grid cell (meaningful for the
    mixed ones), as reconstruct_interface leaves them for the transport."""
    if method == "PLIC":
        mx, my = youngs(C, g)
        ok_cand = torch.ones_like(mx, dtype=torch.bool)
        ux, uy, _ = unit(mx, my)
    else:
        ux, uy, ok_cand = elvira(C, g)
    # interior normals on the halo grid, gathered for ring cells from their source
    pad = torch.nn.functional.pad
    UX = pad(ux, (1, 1, 1, 1))[g.sj, g.si]
    UY = pad(uy, (1, 1, 1, 1))[g.sj, g.si]
    OK = pad(ok_cand, (1, 1, 1, 1))[g.sj, g.si]
    Csrc = C[g.sj, g.si]
    src_mixed = is_mixed(Csrc)
    UX = torch.where(src_mixed & OK, UX, torch.ones_like(UX))
    UY = torch.where(src_mixed & OK, UY, torch.zeros_like(UY))
    rx, ry, _ = unit(UX, UY)  # ghost normals are renormalised
    NX = torch.where(g.ring, rx, UX)
    NY = torch.where(g.ring, ry, UY)
    mixed = is_mixed(C)
    cs = torch.where(mixed, C, torch.full_like(C, 0.5))
    AL = alpha_of(g.X0, g.X1, g.Y0, g.Y1, NX, NY, cs)
    # ELVIRA without a valid candidate: the constant segment (1, 0, 0)
    AL = torch.where(~g.ring & ~OK, torch.zeros_like(AL), AL)
    return NX, NY, AL


# ---------------------------------------------------------------- transport

def wy_pass(src, Cn, vel, g, dt, d, first, corner_fix, method):
    """One directional pass src -> dst (halo layout)."""
    nx, ny, dx, dy = g.nx, g.ny, g.dx, g.dy
    NX, NY, AL = reconstruct(src, g, method)
    if d == 0:
        F = (slice(1, ny + 1), slice(1, nx + 2))     # x faces
        lo_d = (slice(1, ny + 1), slice(0, nx + 1))  # donor when u > 0
        hi_d = F                                     # donor when u < 0
        uf = vel[F]
        xp = g.X0[F]
        b0, b1 = g.Y0[F], g.Y1[F]
    else:
        F = (slice(1, ny + 2), slice(1, nx + 1))
        lo_d = (slice(0, ny + 1), slice(1, nx + 1))
        hi_d = F
        uf = vel[F]
        xp = g.Y0[F]
        b0, b1 = g.X0[F], g.X1[F]
    dl = torch.abs(uf) * dt
    zero = dl.detach() <= 1.0e-14
    pos = uf.detach() > 0.0
    a0 = torch.where(pos, xp - dl, xp)
    a1 = torch.where(pos, xp, xp + dl)
    pick = lambda A: torch.where(pos, A[lo_d], A[hi_d])
    cd, nxd, nyd, ald = pick(src), pick(NX), pick(NY), pick(AL)
    if d == 0:
        Am, _ = cut_area(a0, a1, b0, b1, nxd, nyd, ald)
        full = box_area(a0, a1, b0, b1)
    else:
        Am, _ = cut_area(b0, b1, a0, a1, nxd, nyd, ald)
        full = box_area(b0, b1, a0, a1)
    cdd = cd.detach()
    A = torch.where(is_mixed(cdd), Am, torch.where(cdd <= MIX, torch.zeros_like(Am), full))
    Fx = torch.where(zero, torch.zeros_like(A), torch.where(pos, A, -A))
    if g.per:  # faces 1 and n+1 carry the flux computed with the real donor
        if d == 0:
            keep_last = (vel[1:ny + 1, 1] >= 0.0)[:, None]
            fk = torch.where(keep_last, Fx[:, -1:], Fx[:, :1])
            Fx = torch.cat([fk, Fx[:, 1:-1], fk], dim=1)
        else:
            keep_last = (vel[1, 1:nx + 1] >= 0.0)[None, :]
            fk = torch.where(keep_last, Fx[-1:, :], Fx[:1, :])
            Fx = torch.cat([fk, Fx[1:-1, :], fk], dim=0)
    I = (slice(1, ny + 1), slice(1, nx + 1))
    cdv = (Cn[I].detach() >= 0.5).to(F64)
    if d == 0:
        flo, fhi = Fx[:, :nx], Fx[:, 1:]
        du = vel[1:ny + 1, 2:nx + 2] - vel[1:ny + 1, 1:nx + 1]
        dth = dt / dx
    else:
        flo, fhi = Fx[:ny, :], Fx[1:, :]
        du = vel[2:ny + 2, 1:nx + 1] - vel[1:ny + 1, 1:nx + 1]
        dth = dt / dy
    cn = src[I] - (fhi - flo) / (dx * dy) + cdv * dth * du
    if not first:
        # the clamp; an overshoot of at most 1e-12 (rounding) keeps derivative 1
        cd_ = cn.detach()
        cn = torch.where(cn > 1.0, torch.where(cd_ <= 1.0 + 1.0e-12, cn - cd_ + 1.0, torch.ones_like(cn)), cn)
        cd_ = cn.detach()
        cn = torch.where(cn < 0.0, torch.where(cd_ >= -1.0e-12, cn - cd_, torch.zeros_like(cn)), cn)
    out = torch.nn.functional.pad(cn, (1, 1, 1, 1))[g.sj, g.si]
    if not g.per and not corner_fix:
        out = torch.where(g.corner, src, out)
    return out


def wy_step(C, u, v, g, dt, step_count, method="PLIC", two_phase=True):
    """advect_vof(WEYMOUTH_YUE) once: C^n -> C^{n+1} (halo layout)."""
    x_first = step_count % 2 == 0
    d1, d2 = (0, 1) if x_first else (1, 0)
    vel = (u, v)
    Cs = wy_pass(C, C, vel[d1], g, dt, d1, True, False, method)
    return wy_pass(Cs, C, vel[d2], g, dt, d2, False, two_phase, method)


# ------------------------------------------------------- the coupled model

def clamp01(c):
    t = torch.where(c > 1.0, torch.ones_like(c), c)
    return torch.where(t < 0.0, torch.zeros_like(t), t)


def density(C, rho1, rho2):
    cc = clamp01(C)
    return cc * rho1 + (1.0 - cc) * rho2


def kinetic_energy(C, u, v, g, rho1, rho2):
    nx, ny = g.nx, g.ny
    I = (slice(1, ny + 1), slice(1, nx + 1))
    uw, ue = u[1:ny + 1, 1:nx + 1], u[1:ny + 1, 2:nx + 2]
    vs, vn = v[1:ny + 1, 1:nx + 1], v[2:ny + 2, 1:nx + 1]
    e = 0.5 * density(C[I], rho1, rho2) * (0.5 * (uw * uw + ue * ue) + 0.5 * (vs * vs + vn * vn)) * (g.dx * g.dy)
    return e.sum()


def predictor(u, v, u1, v1, g, dt, gx, gy, first):
    """BDF2 inertia with gravity (explicit Euler when first), zero on the
    wall faces; walls on every side."""
    nx, ny = g.nx, g.ny
    us = u + dt * gx if first else (4.0 * u - u1 + 2.0 * dt * gx) / 3.0
    vs = v + dt * gy if first else (4.0 * v - v1 + 2.0 * dt * gy) / 3.0
    mu = torch.zeros_like(u, dtype=torch.bool)
    mu[1:ny + 1, 2:nx + 1] = True
    mv = torch.zeros_like(v, dtype=torch.bool)
    mv[2:ny + 1, 1:nx + 1] = True
    return torch.where(mu, us, torch.zeros_like(us)), torch.where(mv, vs, torch.zeros_like(vs))


def projection(us, vs, C, g, dt, rho1, rho2):
    """SP Poisson block as a dense solve, mean removal, correction (walls)."""
    nx, ny, dx, dy = g.nx, g.ny, g.dx, g.dy
    rho = density(C, rho1, rho2)
    rho_u = 0.5 * (rho[1:ny + 1, 0:nx + 1] + rho[1:ny + 1, 1:nx + 2])  # faces i = 1..nx+1
    rho_v = 0.5 * (rho[0:ny + 1, 1:nx + 1] + rho[1:ny + 2, 1:nx + 1])  # faces j = 1..ny+1
    kx = (dt * dy) / (rho_u * dx)
    ky = (dt * dx) / (rho_v * dy)
    kx = torch.cat([torch.zeros_like(kx[:, :1]), kx[:, 1:nx], torch.zeros_like(kx[:, :1])], dim=1)
    ky = torch.cat([torch.zeros_like(ky[:1, :]), ky[1:ny, :], torch.zeros_like(ky[:1, :])], dim=0)
    aW, aE, aS, aN = kx[:, :nx], kx[:, 1:], ky[:ny, :], ky[1:, :]
    aP = aW + aE + aS + aN + 1.0e-12 * (dx * dy)
    n = nx * ny
    idx = torch.arange(n, device=g.dev).reshape(ny, nx)
    rows = [idx.reshape(-1)]
    cols = [idx.reshape(-1)]
    vals = [aP.reshape(-1)]
    for a, sh in ((aW, (0, -1)), (aE, (0, 1)), (aS, (-1, 0)), (aN, (1, 0))):
        r0 = idx[max(0, -sh[0]):ny - max(0, sh[0]), max(0, -sh[1]):nx - max(0, sh[1])]
        c0 = idx[max(0, sh[0]):ny + min(0, sh[0]) or None, max(0, sh[1]):nx + min(0, sh[1]) or None]
        av = a[max(0, -sh[0]):ny - max(0, sh[0]), max(0, -sh[1]):nx - max(0, sh[1])]
        rows.append(r0.reshape(-1))
        cols.append(c0.reshape(-1))
        vals.append(-av.reshape(-1))
    A = torch.zeros(n, n, dtype=F64, device=g.dev).index_put(
        (torch.cat(rows), torch.cat(cols)), torch.cat(vals), accumulate=True)
    div = (us[1:ny + 1, 2:nx + 2] - us[1:ny + 1, 1:nx + 1]) * dy + (vs[2:ny + 2, 1:nx + 1] - vs[1:ny + 1, 1:nx + 1]) * dx
    x = torch.linalg.solve(A, (-div).reshape(-1)).reshape(ny, nx)
    p = x - x.mean()
    gu = (p[:, 1:] - p[:, :-1]) / dx        # faces i = 2..nx
    gv = (p[1:, :] - p[:-1, :]) / dy        # faces j = 2..ny
    du = torch.zeros_like(us)
    dv = torch.zeros_like(vs)
    du = du.index_put((torch.arange(1, ny + 1, device=g.dev)[:, None], torch.arange(2, nx + 1, device=g.dev)[None, :]),
                      (dt / rho_u[:, 1:nx]) * gu)
    dv = dv.index_put((torch.arange(2, ny + 1, device=g.dev)[:, None], torch.arange(1, nx + 1, device=g.dev)[None, :]),
                      (dt / rho_v[1:ny, :]) * gv)
    return us - du, vs - dv


def coupled_ke(C0, u0, v0, g, nsteps, dt, rho1, rho2, gx=0.0, gy=-0.98, method="PLIC"):
    """J = KE after nsteps steps of the coupled model from (C0, u0, v0); C0's
    halo is refilled from its interior."""
    C = g.copy_halo(C0)
    u, v = u0, v0
    u1 = v1 = None
    for n in range(nsteps):
        C = wy_step(C, u, v, g, dt, n, method, two_phase=True)
        us, vs = predictor(u, v, u1, v1, g, dt, gx, gy, n == 0)
        u1, v1 = u, v
        u, v = projection(us, vs, C, g, dt, rho1, rho2)
    return kinetic_energy(C, u, v, g, rho1, rho2)
