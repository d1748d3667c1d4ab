"""A numpy restatement of the multigrid V-cycle preconditioner (PRECOND_MG,
FP64, one red-black sweep before and after each coarse correction) with the
same operations in the same order as the GPU kernels, so M r agrees bit for
bit.  Test infrastructure: the V-cycle has no reference counterpart, and this
file is the independent statement the GPU cycle is checked against.

The cycle: level l+1 aggregates 2 x 2 cells of level l (ceil at odd sizes, a
periodic side halved only while even) while the level has more than 16 cells
and both sides at least 4.  Coarse faces are half the sum of the covered fine
faces; a coarse diagonal is its four faces plus eps0 * (level-0 cells
covered) / 2^l.  Levels with more than 4096 cells (or beyond 6144 cells in
all) start from x = 0 with the reds set to b / diag; the others start from
x = 0 with a full Gauss-Seidel red half-sweep.  The coarsest level takes K
red-black sweeps and K in the adjoint order, K = min(64, 2 max(nx, ny)); with
a singular operator its right-hand side and its result are made mean-free
(a serial sum in cell order).
"""
import numpy as np

BOTTOM_CELLS, BOTTOM_LDS = 4096, 6144


def faces_from_assembly(co):
    """kx (ny, nx+1) = west faces + the last east face; ky (ny+1, nx) = south
    faces + the last north face; from the oracle's aW, aE, aS, aN."""
    ny, nx = co["aP"].shape
    kx = np.zeros((ny, nx + 1))
    kx[:, :nx] = co["aW"]
    kx[:, nx] = co["aE"][:, nx - 1]
    ky = np.zeros((ny + 1, nx))
    ky[:ny, :] = co["aS"]
    ky[ny, :] = co["aN"][ny - 1, :]
    return kx, ky


def _coarsen_more(nx, ny):
    return nx * ny > 16 and nx >= 4 and ny >= 4


def hierarchy(kx, ky, diag, perx, pery, eps0):
    levels = [(kx, ky, diag)]
    ny0, nx0 = diag.shape
    nx, ny = nx0, ny0
    while _coarsen_more(nx, ny) and len(levels) < 16:
        if not (nx > 1 and (not perx or nx % 2 == 0)) or not (ny > 1 and (not pery or ny % 2 == 0)):
            break
        kxf, kyf, _ = levels[-1]
        nxc, nyc = (nx + 1) // 2, (ny + 1) // 2
        kxc = np.zeros((nyc, nxc + 1))
        for I in range(1, nxc + 2):
            fi = 2 * I - 1 if I <= nxc else nx + 1
            s = kxf[0::2, fi - 1].copy()                      # fine rows 2J-1
            if ny % 2 == 0:
                s = s + kxf[1::2, fi - 1]
            else:
                s[:-1] = s[:-1] + kxf[1::2, fi - 1]
            kxc[:, I - 1] = 0.5 * s
        kyc = np.zeros((nyc + 1, nxc))
        for J in range(1, nyc + 2):
            fj = 2 * J - 1 if J <= nyc else ny + 1
            s = kyf[fj - 1, 0::2].copy()
            if nx % 2 == 0:
                s = s + kyf[fj - 1, 1::2]
            else:
                s[:-1] = s[:-1] + kyf[fj - 1, 1::2]
            kyc[J - 1, :] = 0.5 * s
        lvl = len(levels)
        span = 1 << lvl
        I = np.arange(1, nxc + 1)
        J = np.arange(1, nyc + 1)
        cx = np.minimum(I * span, nx0) - (I - 1) * span
        cy = np.minimum(J * span, ny0) - (J - 1) * span
        shift = eps0 * (cy[:, None] * cx[None, :]).astype(np.float64) / float(span)
        dg = kxc[:, :-1] + kxc[:, 1:] + kyc[:-1, :] + kyc[1:, :] + shift
        levels.append((kxc, kyc, dg))
        nx, ny = nxc, nyc
    return levels


def _nbrs(x, perx, pery):
    """x of the west, east, south, north neighbours (wrapped; zero outside)."""
    xw, xe, xs, xn = (np.zeros_like(x) for _ in range(4))
    xw[:, 1:], xe[:, :-1], xs[1:, :], xn[:-1, :] = x[:, :-1], x[:, 1:], x[:-1, :], x[1:, :]
    if perx:
        xw[:, 0], xe[:, -1] = x[:, -1], x[:, 0]
    if pery:
        xs[0, :], xn[-1, :] = x[-1, :], x[0, :]
    return xw, xe, xs, xn


def _has(shape, perx, pery):
    ny, nx = shape
    i = np.arange(1, nx + 1)[None, :].repeat(ny, 0)
    j = np.arange(1, ny + 1)[:, None].repeat(nx, 1)
    return (i > 1) | bool(perx), (i < nx) | bool(perx), (j > 1) | bool(pery), (j < ny) | bool(pery), (i + j) & 1


def _gs(lev, x, b, color, perx, pery):
    kx, ky, dg = lev
    hw, he, hs, hn, col = _has(dg.shape, perx, pery)
    xw, xe, xs, xn = _nbrs(x, perx, pery)
    s = b.copy()
    s = np.where(hw, s + kx[:, :-1] * xw, s)
    s = np.where(he, s + kx[:, 1:] * xe, s)
    s = np.where(hs, s + ky[:-1, :] * xs, s)
    s = np.where(hn, s + ky[1:, :] * xn, s)
    return np.where(col == color, s / dg, x)


def _restrict(lev, x, b, scale, perx, pery, shape_c):
    kx, ky, dg = lev
    hw, he, hs, hn, _ = _has(dg.shape, perx, pery)
    xw, xe, xs, xn = _nbrs(x, perx, pery)
    ax = dg * x
    ax = np.where(hw, ax - kx[:, :-1] * xw, ax)
    ax = np.where(he, ax - kx[:, 1:] * xe, ax)
    ax = np.where(hs, ax - ky[:-1, :] * xs, ax)
    ax = np.where(hn, ax - ky[1:, :] * xn, ax)
    r = b - ax
    nyc, nxc = shape_c
    s = np.zeros(shape_c)
    for dj in (0, 1):
        for di in (0, 1):
            sub = r[dj::2, di::2]
            s[:sub.shape[0], :sub.shape[1]] = s[:sub.shape[0], :sub.shape[1]] + sub
    return scale * s


def _prolong(xc, shape):
    ny, nx = shape
    return np.repeat(np.repeat(xc, 2, axis=0), 2, axis=1)[:ny, :nx]


def _serial_mean(a):
    s = 0.0
    for v in a.ravel():  # cell order, one accumulator
        s = s + float(v)
    return s / a.size


def vcycle(levels, r, perx, pery, omega=1.8, project=False):
    """M r; project: the operator is singular (no Dirichlet pressure side) and
    the coarsest solve removes the mean of its right-hand side and of its
    result (P S P)."""
    nl = len(levels)
    cells = [lev[2].size for lev in levels]
    lb = nl - 1
    held = cells[lb]
    while lb > 0:
        if cells[lb - 1] > BOTTOM_CELLS or held + cells[lb - 1] > BOTTOM_LDS:
            break
        held += cells[lb - 1]
        lb -= 1
    nyc, nxc = levels[-1][2].shape
    K = min(64, 2 * max(nxc, nyc))
    scale = 0.5 * omega

    def cycle(l, b):
        lev = levels[l]
        _, _, dg = lev
        col = _has(dg.shape, perx, pery)[4]
        if l == nl - 1 and project:
            b = b - _serial_mean(b)
        if l < lb:
            x = np.where(col == 0, b / dg, 0.0)
        else:
            x = _gs(lev, np.zeros_like(b), b, 0, perx, pery)
        if l == nl - 1:
            x = _gs(lev, x, b, 1, perx, pery)
            for s in range(1, K):
                x = _gs(lev, _gs(lev, x, b, 0, perx, pery), b, 1, perx, pery)
            for s in range(K):
                x = _gs(lev, _gs(lev, x, b, 1, perx, pery), b, 0, perx, pery)
            return x - _serial_mean(x) if project else x
        x = _gs(lev, x, b, 1, perx, pery)
        bc = _restrict(lev, x, b, scale, perx, pery, levels[l + 1][2].shape)
        x = x + _prolong(cycle(l + 1, bc), dg.shape)
        x = _gs(lev, x, b, 1, perx, pery)
        return _gs(lev, x, b, 0, perx, pery)

    return cycle(0, r)
