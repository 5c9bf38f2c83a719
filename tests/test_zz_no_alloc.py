"""The allocation-free site absorption (ipeps_absorb_noalloc) through the C
ABI, on device buffers, against the allocating entry point ipeps_absorb
(same fused kernels, so the results must be bit-identical): two calls in a
row, the second smaller, with the context allocator's live and peak bytes
unchanged by the call."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# (lx, la, lad): left-move enlarged corner and a permuted-leg variant
CASES = [
    (6, 23, 9, ([-1, -4, 11, 12, 21, 22], [21, 22, -2, -5, 3], [11, 12, -3, -6, 3])),
    (4, 11, 5, ([-3, 11, -1, 21, 12, 22], [21, 22, -2, -5, 3], [11, 12, -4, -6, 3])),
]


def _shape_x(lx, chi1, chi2, D):
    opens = [chi1, chi2]
    return [opens.pop(0) if l < 0 else D for l in lx]


def _out_shape(lx, sx, la, sa, lad, sad):
    sizes = {}
    for ind, sh in ((lx, sx), (la, sa), (lad, sad)):
        for l, d in zip(ind, sh):
            if l < 0:
                sizes[-l - 1] = d
    return tuple(sizes[k] for k in range(len(sizes)))


def _dev(x):
    import torch
    return torch.from_numpy(np.asarray(x, dtype=np.float32).ravel(order="F").copy()).cuda()


def _host(t, shape):
    return t.cpu().numpy().reshape(shape, order="F")


def test_absorb_noalloc_matches_absorb(pkg):
    import torch
    from ipeps_amd import _lib
    lib = _lib.lib()
    ctx = pkg.Context()
    ints = lambda l: (C.c_int * len(l))(*[int(v) for v in l])
    i64 = lambda l: (C.c_int64 * len(l))(*[int(v) for v in l])
    ptr = lambda t: C.c_void_p(t.data_ptr())
    try:
        for D, chi1, chi2, (lx, la, lad) in CASES:
            print(f"noalloc case D={D} chi=({chi1},{chi2}) lx={lx}", flush=True)
            rng = np.random.default_rng(100 + D)
            sx = _shape_x(lx, chi1, chi2, D)
            sa = (D, D, D, D, 2)
            X = rng.standard_normal(sx).astype(np.float32)
            A = rng.standard_normal(sa).astype(np.float32)
            Ad = rng.standard_normal(sa).astype(np.float32)
            so = _out_shape(lx, sx, la, sa, lad, sa)
            G = rng.standard_normal(so).astype(np.float32)
            dX, dA, dAd, dG = _dev(X), _dev(A), _dev(Ad), _dev(G)
            dY = torch.empty(int(np.prod(so)), dtype=torch.float32, device="cuda")
            dgX = torch.empty_like(dX)
            dgA = torch.empty_like(dA)
            dgAd = torch.empty_like(dAd)
            torch.cuda.synchronize()
            live0, peak0 = ctx.mem()
            _lib.check(lib.ipeps_absorb_noalloc(ctx.h, i64(sx), ptr(dX), ints(lx), i64(sa), ptr(dA), ints(la),
                                                ptr(dAd), ints(lad), ptr(dG), ptr(dY), ptr(dgX), ptr(dgA),
                                                ptr(dgAd)))
            ctx.sync()
            assert ctx.mem() == (live0, peak0)
            Y0, gX0, gA0, gAd0 = pkg.absorb(X, lx, A, la, Ad, lad, G=G, fused=True, ctx=ctx)
            for got, ref, name in ((_host(dY, so), Y0, "Y"), (_host(dgX, sx), gX0, "gX"),
                                   (_host(dgA, sa), gA0, "gA"), (_host(dgAd, sa), gAd0, "gAd")):
                assert got.shape == ref.shape, name
                np.testing.assert_array_equal(got, ref, err_msg=name)
    finally:
        ctx.close()


def test_absorb_noalloc_rejects_unfused_shapes(pkg):
    import torch
    from ipeps_amd import _lib
    lib = _lib.lib()
    ctx = pkg.Context()
    ints = lambda l: (C.c_int * len(l))(*[int(v) for v in l])
    i64 = lambda l: (C.c_int64 * len(l))(*[int(v) for v in l])
    try:
        D = 3  # below the fused kernel's range: rejected before any launch
        lx, la, lad = CASES[0][3]
        sx, sa = _shape_x(lx, 7, 5, D), (D, D, D, D, 2)
        buf = torch.zeros(4096, dtype=torch.float32, device="cuda")
        p = C.c_void_p(buf.data_ptr())
        st = lib.ipeps_absorb_noalloc(ctx.h, i64(sx), p, ints(lx), i64(sa), p, ints(la), p, ints(lad), None, p,
                                      None, None, None)
        assert st != 0
        assert b"fused FP32 kernel" in lib.ipeps_last_error()
    finally:
        ctx.close()
