"""qlg_linear_combine_noalloc (the entry point that allocates nothing) against
qlg_linear_combine, through the C ABI: real and complex data, two calls in a
row with the second one smaller (fewer elements and fewer basis vectors), a
nonzero beta; results bit-identical (same kernel, same order of operations);
the context's device block cache sees no allocation or reuse during the
no-allocation calls; k beyond the staging limit is refused."""
import ctypes as C

import pytest
import torch

pytestmark = pytest.mark.gpu

if torch.cuda.is_available():
    from ultradmrg_amd import _lib as L
    from ultradmrg_amd.tensor import default_context


def _stats(ctx):
    s = (C.c_int64 * 4)()
    ctx.check(ctx.lib.qlg_ctx_memory_stats(ctx.h, s))
    return list(s)


def _combine(ctx, fn, dtype, n, bases, coefs, beta, y):
    k = len(bases)
    ptrs = (C.c_void_p * max(k, 1))(*[b.data_ptr() for b in bases])
    cf = (C.c_double * max(len(coefs), 1))(*coefs)
    bt = (C.c_double * 2)(*beta)
    torch.cuda.synchronize()
    return fn(ctx.h, dtype, n, k, cf, ptrs, bt, C.c_void_p(y.data_ptr()))


@pytest.mark.parametrize("cplx", [False, True])
def test_noalloc_combine_matches_allocating_entry_point(cplx):
    ctx = default_context(0)
    lib = ctx.lib
    dt = torch.complex128 if cplx else torch.float64
    dtype = L.QLG_C128 if cplx else L.QLG_F64
    g = torch.Generator(device="cuda").manual_seed(7)
    for n, k in ((100_003, 7), (5_001, 3)):      # the second call is smaller
        print(f"[case] noalloc combine cplx={cplx} n={n} k={k}", flush=True)
        bases = [torch.randn(n, dtype=dt, device="cuda", generator=g) for _ in range(k)]
        y0 = torch.randn(n, dtype=dt, device="cuda", generator=g)
        coefs = torch.randn(k * (2 if cplx else 1), dtype=torch.float64, generator=torch.Generator().manual_seed(n)).tolist()
        beta = [0.75, -0.5 if cplx else 0.0]
        ya, yb = y0.clone(), y0.clone()
        ctx.check(_combine(ctx, lib.qlg_linear_combine, dtype, n, bases, coefs, beta, ya))
        before = _stats(ctx)
        ctx.check(_combine(ctx, lib.qlg_linear_combine_noalloc, dtype, n, bases, coefs, beta, yb))
        after = _stats(ctx)
        assert after[1:] == before[1:], "the no-allocation entry point touched the block cache"
        assert torch.equal(ya, yb)
        # sanity against torch on the host
        ref = 0.75 * y0.cpu() if not cplx else (0.75 - 0.5j) * y0.cpu()
        for i, b in enumerate(bases):
            c = complex(coefs[2 * i], coefs[2 * i + 1]) if cplx else coefs[i]
            ref = ref + c * b.cpu()
        assert float((ya.cpu() - ref).abs().max()) <= 1e-12 * float(ref.abs().max())


def test_noalloc_combine_refuses_beyond_limit():
    ctx = default_context(0)
    n, k = 16, 1025
    base = torch.zeros(n, dtype=torch.float64, device="cuda")
    y = torch.zeros(n, dtype=torch.float64, device="cuda")
    rc = _combine(ctx, ctx.lib.qlg_linear_combine_noalloc, L.QLG_F64, n, [base] * k, [1.0] * k, [0.0, 0.0], y)
    assert rc == 1   # QLG_ERR_INVALID
    assert b"staging limit" in ctx.lib.qlg_last_error(ctx.h)
