"""Adjoint of the whole pressure projection (covof_projection_adjoint_d and
the torch.autograd entry point projection): b = -div(u*), A(rho) x = b,
p = x - mean(x), u = u* - (dt/rho_u) dp/dx, v = v* - (dt/rho_v) dp/dy
(src/module_simulation.f90:2477-2654).

Checks, with J = sum(w_u u) + sum(w_v v) + sum(w_p p):
  * against an independent FP64 reference: the same projection written with
    dense torch ops on the CPU (the reference's per-cell assembly, a dense
    solve, the mean removal and halo wrap, the correction with its index
    rules) and differentiated by torch autograd, on walls, periodic and
    pressure-Dirichlet (NEUMANN) sides;
  * against central finite differences along 16 seeded directions of log rho
    (h = 1e-6) and along seeded directions of u*, v*.
"""
import numpy as np
import pytest
import torch

from conftest import to_gpu

pytestmark = pytest.mark.gpu


def _halo(t, perx, pery):
    """Halo fill of an interior (ny, nx) tensor: periodic wrap or a
    zero-gradient copy (the cell-density halo the face means read)."""
    left = t[:, -1:] if perx else t[:, :1]
    right = t[:, :1] if perx else t[:, -1:]
    tx = torch.cat([left, t, right], dim=1)
    down = tx[-1:, :] if pery else tx[:1, :]
    up = tx[:1, :] if pery else tx[-1:, :]
    return torch.cat([down, tx, up], dim=0)


def dense_projection(log_rho, us, vs, w_u, w_v, w_p, bc, dt):
    """J and dJ/dlog_rho, dJ/du*, dJ/dv* by CPU autograd through the
    projection written out with dense torch ops."""
    ny, nx = log_rho.shape
    lr = torch.tensor(log_rho, dtype=torch.float64, requires_grad=True)
    u = torch.tensor(us, dtype=torch.float64, requires_grad=True)
    v = torch.tensor(vs, dtype=torch.float64, requires_grad=True)
    perx, pery = bc[0] == 3, bc[2] == 3
    wall = [b in (1, 4) for b in bc]
    dx, dy = 1.0 / nx, 1.0 / ny
    rh = _halo(torch.exp(lr), perx, pery)
    ru = 0.5 * (rh[:, :-1] + rh[:, 1:])          # ru[j, f - 1]: x face f, between halo cells f-1 and f
    rv = 0.5 * (rh[:-1, :] + rh[1:, :])          # rv[f - 1, i]: y face f, between halo rows f-1 and f
    N = nx * ny
    idx = lambda i, j: (j - 1) * nx + (i - 1)
    rows, cols, vals = [], [], []
    b = []
    for j in range(1, ny + 1):
        for i in range(1, nx + 1):
            k = idx(i, j)
            aW = 0.0 if (i == 1 and not perx and wall[0]) else dt * dy / (ru[j, i - 1] * dx)
            aE = 0.0 if (i == nx and not perx and wall[1]) else dt * dy / (ru[j, i] * dx)
            aS = 0.0 if (j == 1 and not pery and wall[2]) else dt * dx / (rv[j - 1, i] * dy)
            aN = 0.0 if (j == ny and not pery and wall[3]) else dt * dx / (rv[j, i] * dy)
            aP = aW + aE + aS + aN + 1e-12 * dx * dy
            rows.append(k), cols.append(k), vals.append(aP)
            for a, c in ((aW, idx(i - 1, j) if i > 1 else (idx(nx, j) if perx else None)),
                         (aE, idx(i + 1, j) if i < nx else (idx(1, j) if perx else None)),
                         (aS, idx(i, j - 1) if j > 1 else (idx(i, ny) if pery else None)),
                         (aN, idx(i, j + 1) if j < ny else (idx(i, 1) if pery else None))):
                if c is not None and not (isinstance(a, float) and a == 0.0):
                    rows.append(k), cols.append(c), vals.append(-a)
            b.append(-((u[j, i + 1] - u[j, i]) * dy + (v[j + 1, i] - v[j, i]) * dx))
    vals = [x if torch.is_tensor(x) else torch.tensor(x, dtype=torch.float64) for x in vals]
    A = torch.zeros((N, N), dtype=torch.float64).index_put(
        (torch.tensor(rows), torch.tensor(cols)), torch.stack(vals), accumulate=True)
    x = torch.linalg.solve(A, torch.stack(b))
    P = (x - x.mean()).reshape(ny, nx)
    # pressure halo: interior, then the periodic wrap (x first, then y)
    p = torch.zeros((ny + 2, nx + 2), dtype=torch.float64)
    p = p.index_put((torch.arange(1, ny + 1)[:, None], torch.arange(1, nx + 1)[None, :]), P)
    if perx:
        p = torch.cat([p[:, nx:nx + 1], p[:, 1:nx + 1], p[:, 1:2]], dim=1)
    if pery:
        p = torch.cat([p[ny:ny + 1, :], p[1:ny + 1, :], p[1:2, :]], dim=0)
    # correction (k_correct's index rules; walls skipped)
    u_nx, v_ny = (nx if perx else nx + 1), (ny if pery else ny + 1)
    fu = torch.zeros_like(u)
    for f in range(1, u_nx + 1):
        if (f == 1 and not perx and wall[0]) or (f == u_nx and not perx and wall[1]):
            continue
        cw = f - 1 if f > 1 else (nx if perx else 1)
        ce = f if f <= nx else (1 if perx else nx)
        g = (P[:, ce - 1] - P[:, cw - 1]) / dx
        col = -(dt / ru[1:ny + 1, f - 1]) * g
        fu = fu + torch.nn.functional.pad(col[:, None], (f, nx + 1 - f, 1, 1))
    fv = torch.zeros_like(v)
    for f in range(1, v_ny + 1):
        if (f == 1 and not pery and wall[2]) or (f == v_ny and not pery and wall[3]):
            continue
        cs = f - 1 if f > 1 else (ny if pery else 1)
        cn = f if f <= ny else (1 if pery else ny)
        g = (P[cn - 1, :] - P[cs - 1, :]) / dy
        row = -(dt / rv[f - 1, 1:nx + 1]) * g
        fv = fv + torch.nn.functional.pad(row[None, :], (1, 1, f, ny + 1 - f))
    uo, vo = u + fu, v + fv
    J = (torch.tensor(w_u) * uo).sum() + (torch.tensor(w_v) * vo).sum() + (torch.tensor(w_p) * p).sum()
    J.backward()
    return J.item(), lr.grad.numpy(), u.grad.numpy(), v.grad.numpy()


def _case(nx, ny, seed, bc):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid((np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx, indexing="ij")
    C = ((xx - 0.45) ** 2 + (yy - 0.55) ** 2 < 0.28 ** 2).astype(float)
    C = 0.3 * C + 0.7 * np.clip(C + 0.3 * rng.random(C.shape), 0, 1) * C
    log_rho = np.log(1.0 + 999.0 * C)
    us = 0.1 * rng.normal(size=(ny + 2, nx + 2))
    vs = 0.1 * rng.normal(size=(ny + 2, nx + 2))
    # periodic twins are equal and wall faces carry the no-penetration value,
    # as after the exchange and the velocity BC
    if bc[0] == 3:
        us[:, nx + 1] = us[:, 1]
    if bc[2] == 3:
        vs[ny + 1, :] = vs[1, :]
    if bc[0] in (1, 4):
        us[:, 1] = 0.0
    if bc[1] in (1, 4):
        us[:, nx + 1] = 0.0
    if bc[2] in (1, 4):
        vs[1, :] = 0.0
    if bc[3] in (1, 4):
        vs[ny + 1, :] = 0.0
    w_u, w_v, w_p = (rng.normal(size=(ny + 2, nx + 2)) for _ in range(3))
    return log_rho, us, vs, w_u, w_v, w_p


BCS = {"walls": (4, 1, 4, 1), "periodic": (3, 3, 3, 3), "periodic-x-neumann-top": (3, 3, 1, 2),
       "neumann-right": (1, 2, 4, 1)}


def _singular(bc):
    return 2 not in bc


def _face_mask(nx, ny, bc, axis):
    """The face entries whose derivative is defined independently of the
    operator's near-null constant mode: with no pressure-Dirichlet side the
    boundary-normal faces (and periodic twins) change the right-hand side's
    mean, which the 1e-12 dvol shift turns into an arbitrary constant."""
    m = np.ones((ny + 2, nx + 2), bool)
    if not _singular(bc):
        return m
    if axis == 0:
        m[:, [0, 1, nx + 1]] = False
    else:
        m[[0, 1, ny + 1], :] = False
    return m


@pytest.mark.parametrize("name", list(BCS))
def test_projection_adjoint_matches_dense_autograd(cv, name):
    from covof_amd.autograd import projection
    bc = BCS[name]
    nx, ny = 24, 20
    dt = 0.05
    log_rho, us, vs, w_u, w_v, w_p = _case(nx, ny, 31, bc)
    J_ref, g_lr, g_u, g_v = dense_projection(log_rho, us, vs, w_u, w_v, w_p, bc, dt)
    ctx = cv.Context(nx, ny, bc=bc)
    lr = to_gpu(log_rho).requires_grad_(True)
    ug = to_gpu(us).requires_grad_(True)
    vg = to_gpu(vs).requires_grad_(True)
    u, v, p = projection(ctx, lr, ug, vg, dt, tol=1e-14, max_iter=100000)
    J = (to_gpu(w_u) * u).sum() + (to_gpu(w_v) * v).sum() + (to_gpu(w_p) * p).sum()
    J.backward()
    print(f"[{name}] J {J.item():.12e} (dense {J_ref:.12e})")
    assert abs(J.item() - J_ref) <= 1e-8 * abs(J_ref)
    e_lr = np.abs(lr.grad.cpu().numpy() - g_lr).max() / np.abs(g_lr).max()
    mu, mv = _face_mask(nx, ny, bc, 0), _face_mask(nx, ny, bc, 1)
    e_u = np.abs(ug.grad.cpu().numpy() - g_u)[mu].max() / np.abs(g_u[mu]).max()
    e_v = np.abs(vg.grad.cpu().numpy() - g_v)[mv].max() / np.abs(g_v[mv]).max()
    print(f"[{name}] rel err dJ/dlogrho {e_lr:.2e}, dJ/du* {e_u:.2e}, dJ/dv* {e_v:.2e}")
    assert e_lr < 1e-8 and e_u < 1e-8 and e_v < 1e-8


@pytest.mark.parametrize("name", ["walls", "periodic-x-neumann-top"])
def test_projection_adjoint_central_differences(cv, name):
    """16 seeded directions of log rho (central differences, h = 1e-6) and
    4 seeded directions of (u*, v*) (J is linear in them), on 96 x 80."""
    from covof_amd.autograd import projection
    bc = BCS[name]
    nx, ny, dt = 96, 80, 0.02
    log_rho, us, vs, w_u, w_v, w_p = _case(nx, ny, 41, bc)
    ctx = cv.Context(nx, ny, bc=bc)
    W = [to_gpu(a) for a in (w_u, w_v, w_p)]
    opts = dict(tol=1e-14, max_iter=400000)

    def J(lr, a, b):
        with torch.no_grad():
            out = projection(ctx, to_gpu(lr), to_gpu(a), to_gpu(b), dt, **opts)
            return sum(float((w * o).sum()) for w, o in zip(W, out))

    lr = to_gpu(log_rho).requires_grad_(True)
    ug = to_gpu(us).requires_grad_(True)
    vg = to_gpu(vs).requires_grad_(True)
    out = projection(ctx, lr, ug, vg, dt, **opts)
    sum((w * o).sum() for w, o in zip(W, out)).backward()
    g_lr, g_u, g_v = lr.grad.cpu().numpy(), ug.grad.cpu().numpy(), vg.grad.cpu().numpy()
    rng = np.random.default_rng(17)
    h = 1e-6
    worst = 0.0
    for _ in range(16):
        d = rng.normal(size=log_rho.shape)
        fd = (J(log_rho + h * d, us, vs) - J(log_rho - h * d, us, vs)) / (2 * h)
        ad = float((g_lr * d).sum())
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"[{name}] log rho: worst rel err over 16 directions {worst:.2e}")
    assert worst < 1e-6
    mu, mv = _face_mask(nx, ny, bc, 0), _face_mask(nx, ny, bc, 1)
    base = J(log_rho, us, vs)
    for _ in range(4):
        du = rng.normal(size=us.shape) * mu
        dv = rng.normal(size=vs.shape) * mv
        fd = J(log_rho, us + du, vs + dv) - base
        ad = float((g_u * du).sum() + (g_v * dv).sum())
        assert abs(fd - ad) <= 1e-8 * abs(ad), (fd, ad)
