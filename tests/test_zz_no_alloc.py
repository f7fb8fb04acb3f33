"""GPU test of phub_cf_eval_noalloc (the entry point for callers that cannot
allocate): on two calls in a row, the second smaller, it returns exactly what
phub_cf_eval returns for the same inputs (the same kernel on the same staged
values), makes no allocation (phub_alloc_count unchanged), and rejects inputs
beyond its fixed scratch with PHUB_ERR_UNSUPPORTED before touching the GPU."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chains(rng, nch, m, nw):
    a = np.ascontiguousarray(rng.uniform(-6, 6, (nch, m)))
    b = np.ascontiguousarray(rng.uniform(0.2, 3, (nch, m)))
    b[:, -1] = 0.0
    a[: nch // 3, m // 2:] = 0.0            # some chains end early (trailing zero levels)
    b[: nch // 3, m // 2 - 1:] = 0.0
    n2 = rng.uniform(0.1, 1.0, nch)
    sh = rng.normal(0, 3, nch)
    sg = rng.choice([-1.0, 1.0], nch)
    w = np.linspace(-15, 15, nw)
    return a, b, n2, sh, sg, w


def _call(fn, ctx, a, b, n2, sh, sg, w, eta):
    nch, m = a.shape
    out = np.zeros(2 * nch * len(w))
    rc = fn(ctx.h, nch, m, _dp(a), _dp(b), _dp(n2), _dp(sh), _dp(sg), len(w), _dp(w), eta, _dp(out))
    return rc, out


def test_cf_eval_noalloc_matches_cf_eval(ctx):
    lib = ctx.lib
    rng = np.random.default_rng(17)
    big = _chains(rng, 300, 120, 2048)        # 1.30e6 doubles of the 4.19e6 scratch
    small = _chains(rng, 37, 50, 500)
    ref_big = _call(lib.phub_cf_eval, ctx, *big, 0.05)
    ref_small = _call(lib.phub_cf_eval, ctx, *small, 0.3)
    assert ref_big[0] == 0 and ref_small[0] == 0
    count0 = lib.phub_alloc_count()
    rc1, o1 = _call(lib.phub_cf_eval_noalloc, ctx, *big, 0.05)
    rc2, o2 = _call(lib.phub_cf_eval_noalloc, ctx, *small, 0.3)
    assert rc1 == 0 and rc2 == 0, lib.phub_last_error(ctx.h)
    assert lib.phub_alloc_count() == count0
    assert np.array_equal(o1, ref_big[1])
    assert np.array_equal(o2, ref_small[1])
    assert np.all(np.isfinite(o1)) and np.abs(o1).max() > 0


def test_cf_eval_noalloc_rejects_oversize(ctx):
    from pyhub_amd import _lib as L
    lib = ctx.lib
    rng = np.random.default_rng(18)
    huge = _chains(rng, 1100, 20, 2048)       # 4.6e6 doubles > 2^22
    rc, _ = _call(lib.phub_cf_eval_noalloc, ctx, *huge, 0.05)
    assert rc == L.ERR_UNSUPPORTED
    assert b"exceeds the fixed scratch" in lib.phub_last_error(ctx.h)
