"""GPU tests of the graph-captured Lanczos steps: the three-term recurrence replayed
from hipGraphs (scalars in device memory, the host waiting only at the convergence
checks) gives the same bits as the host-driven loop -- alpha, beta, the step count,
E0 and psi -- in FP64 and FP32 vectors, with the vectors kept (psi in one pass) and
without (three rotating vectors), for a fixed step count and through a breakdown
(a sector smaller than the step budget)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _run(op, monkeypatch, graph, **kw):
    monkeypatch.setenv("PHUB_LANCZOS_GRAPH", "1" if graph else "0")
    r = op.lanczos(**kw)
    torch.cuda.synchronize()
    psi = r.psi.cpu().numpy() if r.psi is not None else None
    return r, psi


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("want_psi", [True, False])
@pytest.mark.parametrize("case", ["converge", "fixed", "breakdown"])
def test_graph_lanczos_same_bits_as_host_driven(ctx, monkeypatch, prec, want_psi, case):
    from pyhub_amd import _lib as L
    from pyhub_amd.engine import HubbardOp
    rng = np.random.default_rng(60)
    if case == "breakdown":
        n, nu, nd = 4, 2, 1                   # 24 states, 60-step budget
    else:
        n, nu, nd = 10, 5, 4                  # 52,920 states
    T = O.ring_T(n)
    T[0, n // 2] = T[n // 2, 0] = -0.4
    T[np.arange(n), np.arange(n)] = rng.normal(0, 0.3, n)
    op = HubbardOp(ctx, n, nu, nd, T, rng.uniform(2, 6, n))
    v0 = torch.as_tensor(rng.uniform(-0.5, 0.5, op.dim), device="cuda")
    mode = L.LCZ_PLAIN64 if prec == 64 else L.LCZ_PLAIN32
    kw = dict(mode=mode, max_steps=60 if case == "breakdown" else 200, tol=1e-11, v0=v0,
              want_psi=want_psi)
    if case == "fixed":
        kw["fixed_steps"] = 37
    r0, p0 = _run(op, monkeypatch, False, **kw)
    r1, p1 = _run(op, monkeypatch, True, **kw)
    print(f"{case} fp{prec} psi={want_psi}: {r0.steps} steps, E0 {r0.e0:.12f}", flush=True)
    assert r1.steps == r0.steps and r1.converged == r0.converged
    assert np.array_equal(r1.alpha, r0.alpha) and np.array_equal(r1.beta, r0.beta)
    assert r1.e0 == r0.e0
    if want_psi:
        assert np.array_equal(p1, p0)
    if case == "breakdown":
        assert r0.steps <= op.dim
