"""Adjoint of the variable-density pressure solve (covof_poisson_adjoint_d and
the torch.autograd entry point pressure_solve).

Checks, with J = sum(w * p), p = A(exp(log_rho))^-1 rhs:
  * against an independent FP64 reference gradient: the operator assembled
    from log_rho with dense torch ops on the CPU (arithmetic face means,
    k = dt*dy/(rho_f*dx), walls dropped, +1e-12*dvol on the diagonal), solved
    with torch.linalg.solve and differentiated by torch autograd;
  * against central finite differences along seeded random directions;
  * at 1024^2 with the mixed-precision solves (the adjoint workload: density
    of 64-circle type, w seeded Gaussian).
"""
import numpy as np
import pytest
import torch

from conftest import to_gpu

pytestmark = pytest.mark.gpu


def _dense_reference(log_rho, rhs, w, bc, dt=1.0):
    """J and its gradients by CPU autograd through a dense FP64 solve."""
    ny, nx = log_rho.shape
    lr = torch.tensor(log_rho, dtype=torch.float64, requires_grad=True)
    b = torch.tensor(rhs, dtype=torch.float64, requires_grad=True)
    perx, pery = bc[0] == 3, bc[2] == 3
    rho = torch.exp(lr)
    # halo fill as the reference's (periodic wrap or zero-gradient copy)
    left = rho[:, -1:] if perx else rho[:, :1]
    right = rho[:, :1] if perx else rho[:, -1:]
    rx = torch.cat([left, rho, right], dim=1)           # (ny, nx+2)
    down = rx[-1:, :] if pery else rx[:1, :]
    up = rx[:1, :] if pery else rx[-1:, :]
    rh = torch.cat([down, rx, up], dim=0)               # (ny+2, nx+2)
    dx, dy = 1.0 / nx, 1.0 / ny
    rho_u = 0.5 * (rh[1:ny + 1, 0:nx + 1] + rh[1:ny + 1, 1:nx + 2])   # faces i = 1..nx+1
    rho_v = 0.5 * (rh[0:ny + 1, 1:nx + 1] + rh[1:ny + 2, 1:nx + 1])   # faces j = 1..ny+1
    kx = dt * dy / (rho_u * dx)
    ky = dt * dx / (rho_v * dy)
    N = nx * ny
    A = torch.zeros((N, N), dtype=torch.float64)
    idx = lambda i, j: j * nx + i  # 0-based interior
    diag = torch.zeros(N, dtype=torch.float64) + 1e-12 * dx * dy
    rows, cols, vals = [], [], []
    for j in range(ny):
        for i in range(nx + 1):   # x faces; face i sits between cells i-1 and i
            k = kx[j, i]
            L, R = i - 1, i
            if 0 < i < nx:
                a, c = idx(L, j), idx(R, j)
            elif perx:
                a, c = idx(nx - 1, j), idx(0, j)
                if i == nx:
                    continue  # the periodic face is face 0's twin
            elif bc[0 if i == 0 else 1] in (1, 4):
                continue  # wall face: coefficient 0
            else:
                c = idx(0 if i == 0 else nx - 1, j)
                diag = diag.index_add(0, torch.tensor([c]), k.reshape(1))
                continue
            diag = diag.index_add(0, torch.tensor([a, c]), torch.stack([k, k]))
            rows += [a, c]
            cols += [c, a]
            vals += [-k, -k]
    for j in range(ny + 1):
        for i in range(nx):
            k = ky[j, i]
            if 0 < j < ny:
                a, c = idx(i, j - 1), idx(i, j)
            elif pery:
                a, c = idx(i, ny - 1), idx(i, 0)
                if j == ny:
                    continue
            elif bc[2 if j == 0 else 3] in (1, 4):
                continue
            else:
                c = idx(i, 0 if j == 0 else ny - 1)
                diag = diag.index_add(0, torch.tensor([c]), k.reshape(1))
                continue
            diag = diag.index_add(0, torch.tensor([a, c]), torch.stack([k, k]))
            rows += [a, c]
            cols += [c, a]
            vals += [-k, -k]
    A = A.index_put((torch.tensor(rows), torch.tensor(cols)), torch.stack(vals), accumulate=True)
    A = A + torch.diag(diag)
    p = torch.linalg.solve(A, b.reshape(N))
    J = (torch.tensor(w, dtype=torch.float64).reshape(N) * p).sum()
    J.backward()
    return J.item(), lr.grad.numpy(), b.grad.numpy()


def _case(nx, ny, seed, periodic):
    rng = np.random.default_rng(seed)
    yy, xx = np.meshgrid((np.arange(ny) + 0.5) / ny, (np.arange(nx) + 0.5) / nx, indexing="ij")
    C = ((xx - 0.45) ** 2 + (yy - 0.55) ** 2 < 0.25 ** 2).astype(float)
    C = 0.25 * C + 0.75 * np.clip(C + 0.3 * rng.random(C.shape), 0, 1) * C  # a ragged heavy blob
    log_rho = np.log(1.0 + 999.0 * C)
    rhs = rng.normal(size=(ny, nx))
    rhs -= rhs.mean()
    w = rng.normal(size=(ny, nx))
    w -= w.mean()
    bc = (3, 3, 3, 3) if periodic else (4, 4, 4, 4)
    return log_rho, rhs, w, bc


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_matches_dense_autograd(cv, periodic):
    from covof_amd.autograd import pressure_solve
    nx, ny = 24, 20
    log_rho, rhs, w, bc = _case(nx, ny, 11 + periodic, periodic)
    J_ref, g_lr_ref, g_b_ref = _dense_reference(log_rho, rhs, w, bc)
    ctx = cv.Context(nx, ny, bc=bc)
    lr = to_gpu(log_rho).requires_grad_(True)
    b = to_gpu(rhs).requires_grad_(True)
    p = pressure_solve(ctx, lr, b, 1.0, tol=1e-14, max_iter=100000)
    J = (to_gpu(w) * p).sum()
    J.backward()
    assert abs(J.item() - J_ref) <= 1e-9 * abs(J_ref)
    scale = np.abs(g_lr_ref).max()
    err_lr = np.abs(lr.grad.cpu().numpy() - g_lr_ref).max() / scale
    # dJ/drhs = lam = A^-1 w: its constant component is set by the 1e-12*dvol
    # shift and the rounding of mean(w) (A is otherwise singular), so compare
    # the mean-free parts (all that acts on a zero-mean rhs)
    gb = b.grad.cpu().numpy()
    gb, gbr = gb - gb.mean(), g_b_ref - g_b_ref.mean()
    err_b = np.abs(gb - gbr).max() / np.abs(gbr).max()
    print(f"dJ/dlogrho rel err {err_lr:.2e}, dJ/drhs rel err {err_b:.2e}")
    assert err_lr < 1e-8 and err_b < 1e-8


@pytest.mark.parametrize("periodic", [False, True])
def test_adjoint_central_differences(cv, periodic):
    """<dJ/dlogrho, d> against (J(lr + h d) - J(lr - h d)) / 2h, h = 1e-6."""
    from covof_amd.autograd import pressure_solve
    nx, ny = 48, 40
    log_rho, rhs, w, bc = _case(nx, ny, 21 + periodic, periodic)
    ctx = cv.Context(nx, ny, bc=bc)
    wg, bg = to_gpu(w), to_gpu(rhs)

    def J(lr):
        with torch.no_grad():
            return (wg * pressure_solve(ctx, to_gpu(lr), bg, 1.0, tol=1e-14, max_iter=200000)).sum().item()

    lr = to_gpu(log_rho).requires_grad_(True)
    b = bg.clone().requires_grad_(True)
    (wg * pressure_solve(ctx, lr, b, 1.0, tol=1e-14, max_iter=200000)).sum().backward()
    g = lr.grad.cpu().numpy()
    rng = np.random.default_rng(5)
    h = 1e-6
    for _ in range(4):
        d = rng.normal(size=log_rho.shape)
        fd = (J(log_rho + h * d) - J(log_rho - h * d)) / (2 * h)
        ad = float((g * d).sum())
        assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-300), (fd, ad)
    # J is linear in rhs: dJ/drhs = lam exactly, checked along a zero-mean
    # direction (a mean would excite the near-null constant mode)
    d = rng.normal(size=rhs.shape)
    d -= d.mean()
    lin = J(log_rho)
    with torch.no_grad():
        Jd = (wg * pressure_solve(ctx, to_gpu(log_rho), to_gpu(rhs + d), 1.0, tol=1e-14,
                                  max_iter=200000)).sum().item()
    assert abs((Jd - lin) - float((b.grad.cpu().numpy() * d).sum())) <= 1e-8 * abs(Jd - lin)


def test_adjoint_1024_mixed_precision_fd(cv):
    """The adjoint workload at 1024^2: density of seeded circles (rho 1 vs 1000,
    one-cell smoothing), J = sum(w p) with w seeded Gaussian, forward and
    transpose solves in the mixed mode; central FD along seeded directions."""
    from covof_amd import workloads as wl
    from covof_amd.autograd import pressure_solve
    n = 1024
    case = wl.poisson_case(n, seed=0, n_circles=64, rmin=8, rmax=64)
    C = case["C"][1:n + 1, 1:n + 1]
    log_rho = np.log(1000.0 * C + (1.0 - C))
    rng = np.random.default_rng(7)
    w = rng.normal(size=(n, n))
    w -= w.mean()
    ctx = cv.Context(n, n, bc=case["bc"])
    wg, bg = to_gpu(w), to_gpu(case["rhs"])
    opts = dict(tol=1e-12, max_iter=400000, precision=cv.PRECISION_MIXED)

    def J(lr):
        with torch.no_grad():
            return (wg * pressure_solve(ctx, to_gpu(lr), bg, 1.0, **opts)).sum().item()

    lr = to_gpu(log_rho).requires_grad_(True)
    (wg * pressure_solve(ctx, lr, bg, 1.0, **opts)).sum().backward()
    g = lr.grad.cpu().numpy()
    # fourth-order central differences (Richardson on h and 2h) with h = 1e-3:
    # truncation O(h^4) ~ 1e-12, and the solves' 1e-12 noise amplified by 1/h
    # only; the 1e-6 of a second-order difference would leave the noise of a
    # mixed-precision solve (attainable ~1e-13) amplified a million-fold
    h = 1e-3
    worst = 0.0
    for _ in range(3):
        d = rng.normal(size=log_rho.shape)
        d1 = J(log_rho + h * d) - J(log_rho - h * d)
        d2 = J(log_rho + 2 * h * d) - J(log_rho - 2 * h * d)
        fd = (8.0 * d1 - d2) / (12.0 * h)
        ad = float((g * d).sum())
        worst = max(worst, abs(fd - ad) / abs(ad))
    print(f"1024^2 mixed-precision adjoint vs FD: worst rel err {worst:.2e}")
    assert worst < 1e-6
