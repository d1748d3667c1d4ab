"""The allocation-free entry point covof_reconstruct_interface_noalloc_d
(static device scratch, no context) against covof_reconstruct_interface_d
through the C-ABI: bit for bit on two calls in a row, the second on a
smaller mesh (the scratch keeps the first call's lists), and the size limit
refused on the host."""
import ctypes as ct

import numpy as np
import pytest

from conftest import to_gpu

pytestmark = pytest.mark.gpu


def _case(nx, ny, seed, periodic):
    from covof_amd import workloads as wl
    C = wl.ellipses(nx, ny, 12, 3.0, 0.2 * min(nx, ny), seed=seed)
    (wl.fill_halo_periodic if periodic else wl.fill_halo_zero_gradient)(C, nx, ny)
    return C


def _noalloc(cv, nx, ny, bc, method, C):
    import torch
    lib = cv._lib.lib()
    out = [torch.full_like(C, float("nan")) for _ in range(3)]
    st = lib.covof_reconstruct_interface_noalloc_d(nx, ny, 1.0, 1.0, (ct.c_int * 4)(*bc), method.encode(),
                                                   C.data_ptr(), *[o.data_ptr() for o in out],
                                                   ct.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    return st, out


def test_noalloc_reconstruction_matches_context_entry(cv):
    cases = [(320, 208, (1, 1, 1, 1), "ELVIRA", 11, False),   # 5 x 13 tiles
             (96, 80, (3, 3, 3, 3), "PLIC", 12, True),        # smaller, after the first
             (130, 66, (3, 3, 1, 1), "ELVIRA", 13, False)]    # partial tiles
    for nx, ny, bc, method, seed, per in cases:
        print(f"[noalloc] {nx}x{ny} bc {bc} {method}", flush=True)
        C = to_gpu(_case(nx, ny, seed, per))
        st, got = _noalloc(cv, nx, ny, bc, method, C)
        assert st == cv._lib.OK, cv._lib.lib().covof_last_error()
        ctx = cv.Context(nx, ny, 1.0, 1.0, bc=bc)
        ref = ctx.reconstruct_interface(method, C)
        ctx.synchronize()
        for a, b, name in zip(got, ref, ("nx", "ny", "alpha")):
            assert np.array_equal(a.cpu().numpy(), b.cpu().numpy()), name
        ctx.close()


def test_noalloc_rejects_meshes_beyond_its_scratch(cv):
    import torch
    T = [torch.zeros((4, 4), dtype=torch.float64, device="cuda") for _ in range(4)]
    lib = cv._lib.lib()
    bc = (ct.c_int * 4)(1, 1, 1, 1)
    # 2048 x 2048 needs 32 x 128 = 4096 tiles: refused before any launch
    st = lib.covof_reconstruct_interface_noalloc_d(2048, 2048, 1.0, 1.0, bc, b"PLIC", *[t.data_ptr() for t in T],
                                                   None)
    assert st == cv._lib.BAD_ARGUMENT
    assert b"tiles" in lib.covof_last_error()
