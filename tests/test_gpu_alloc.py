"""GPU test of the allocation contract of include/phub.h: after a first call (which
may build tables kept with the operator and grow the context's reduction buffer and
scratch arena), every hot entry point called again on the same operator and
context, with caller workspaces where it takes one, makes no device or pinned-host
allocation (phub_alloc_count unchanged); the repeat's results equal the first's to
1e-10 (the two-body sums accumulate with atomics, so their last bits may differ)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _hubbard_calls(ctx):
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp, cf_eval
    n, nu, nd = 8, 4, 3
    T = O.ring_T(n)
    T[0, 4] = T[4, 0] = -0.3
    op = HubbardOp(ctx, n, nu, nd, T, 4.0)
    tgt_p = HubbardOp(ctx, n, nu + 1, nd, T, 4.0)
    v = torch.as_tensor(np.random.default_rng(3).normal(size=op.dim), device="cuda")
    ws = {m: op._workspace(m, 200, 0, 2, True) for m in (L.LCZ_PLAIN64, L.LCZ_FULL64, L.LCZ_MIXED)}
    cws = tgt_p.cf_chain_workspace(64)
    mws = op.minres_workspace()
    bws = op.band_workspace(2)
    xout = op.vec(torch.float64)

    def run():
        out = {}
        out["hv64"] = op.hv(v).cpu().numpy()
        out["hv32"] = op.hv(v.float()).double().cpu().numpy()
        for m in (L.LCZ_PLAIN64, L.LCZ_FULL64, L.LCZ_MIXED):
            r = op.lanczos(mode=m, max_steps=200, tol=1e-12, refine_steps=2, seed=5,
                           workspace=ws[m])
            out[f"e0_{m}"] = r.e0
            psi = r.psi
        out["rdm"] = op.rdm(psi)
        out["rdm_spin"] = op.one_rdm_spin(psi)
        out["hf"] = op.hf_grad(psi)
        out["grad"] = op.grad(v, psi)
        out["dbl"] = op.two_rdm(psi)
        _, x = op.excite(psi, 2, 0, True, target=tgt_p)
        out["chain"] = tgt_p.cf_chain(x, 40, 64, workspace=cws)
        out["cf"] = cf_eval(ctx, out["chain"][0], out["chain"][1], out["chain"][2], out["e0_0"],
                            1.0, np.linspace(-5, 5, 64), 0.1)
        b = op.hv(psi)
        b -= torch.dot(psi, b) * psi
        out["minres"] = op.minres(psi, out["e0_0"], b, max_iter=400, tol=1e-10, workspace=mws,
                                  out=xout)[0].cpu().numpy()
        V2 = torch.stack([op.hv(psi), psi])
        out["band"] = op.band_lanczos(V2, max_steps=60, tol=1e-8, workspace=bws)
        return out

    return run


def _heisenberg_calls(ctx):
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HeisenbergOp
    Jstd = O.j1j2_square(4, 3, 1.0, 0.5)
    op = HeisenbergOp(ctx, 12, (Jstd + Jstd.T) / 2, sign_convention=1)
    ws = op._workspace(L.LCZ_FULL64, 200, 0, 0, True)

    def run():
        r = op.lanczos(mode=L.LCZ_FULL64, max_steps=200, tol=1e-12, workspace=ws)
        return {"e0": r.e0, "hv": op.hv(r.psi).cpu().numpy(), "sisj": op.sisj(r.psi),
                "hf": op.hf_grad(r.psi)}

    return run


def _same(a, b):
    if isinstance(a, dict):
        return all(_same(a[k], b[k]) for k in a)
    if isinstance(a, (tuple, list)):
        return all(_same(x, y) for x, y in zip(a, b))
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.allclose(a, b, rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("model", ["hubbard", "heisenberg"])
def test_repeated_calls_allocate_nothing(ctx, model):
    lib = ctx.lib
    run = (_hubbard_calls if model == "hubbard" else _heisenberg_calls)(ctx)
    first = run()
    torch.cuda.synchronize()
    before = lib.phub_alloc_count()
    again = run()
    torch.cuda.synchronize()
    after = lib.phub_alloc_count()
    print(f"{model}: library allocations in the repeated calls: {after - before}", flush=True)
    assert after == before
    assert _same(first, again)
