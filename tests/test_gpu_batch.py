"""GPU tests of the batched persistent launch (many clusters, ground states and
diagonal Green's functions) against the single-cluster FP64 path."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_batched_clusters_vs_single_path(ctx):
    from pyhub_amd import _lib as L
    from pyhub_amd.batch import gii_omega, random_cluster, run_batch
    from pyhub_amd.engine import HubbardOp, cf_eval
    import time
    clusters = [random_cluster(c) for c in range(12)]
    t0 = time.perf_counter()
    e0, ga, gb, gn = run_batch(clusters, gs_steps=256, gf_steps=80, seed=3, gs_tol=1e-9)
    print(f"batched launch: {len(clusters)} clusters in {time.perf_counter() - t0:.2f} s", flush=True)
    omega = np.linspace(-15, 15, 256)
    eta = 1.0
    G = gii_omega(e0, ga, gb, gn, clusters, omega, eta=eta)
    errs, cerrs = [], []
    for ci, cl in enumerate(clusters):
        op = HubbardOp(ctx, cl.n, cl.n_up, cl.n_dn, cl.T, cl.U)
        r = op.lanczos(mode=L.LCZ_FULL64, max_steps=min(300, op.dim), tol=1e-12)
        print(f"cluster {ci}: n={cl.n} ({cl.n_up},{cl.n_dn}) dim={op.dim} e0 {e0[ci]:.12f} ref {r.e0:.12f}",
              flush=True)
        # ground state in FP64 vectors: E0 to the FP64 solver's accuracy
        assert abs(e0[ci] - r.e0) <= 1e-8 * max(1.0, abs(r.e0)), (ci, e0[ci], r.e0)
        for site in (1, cl.n):
            for spin in (0, 1):
                # sum rule <c c^dag> + <c^dag c> = 1
                assert abs(gn[ci, site - 1, spin].sum() - 1.0) < 1e-6
                Gref = np.zeros(len(omega), dtype=complex)
                for kind, create, sign in ((0, True, 1.0), (1, False, -1.0)):
                    nu = cl.n_up + ((1 if create else -1) if spin == 0 else 0)
                    nd = cl.n_dn + ((1 if create else -1) if spin == 1 else 0)
                    if not (0 <= nu <= cl.n and 0 <= nd <= cl.n):
                        continue
                    tgt, x = op.excite(r.psi, site, spin, create)
                    a, b, n2 = tgt.cf_chain(x, min(80, tgt.dim), precision=64)
                    if n2 > 0:
                        Gref += cf_eval(ctx, a, b, n2, r.e0, sign, omega, eta)[0]
                        # the leading recurrence coefficients fix the low moments of
                        # the spectral function: FP32 chain vectors against FP64
                        # compare up to the first beta that (nearly) terminates the
                        # chain: past it the coefficients are rounding noise in both
                        # (alpha_0, beta_0, alpha_1: the first four moments; later ones
                        # amplify the FP32 rounding by the chain's own conditioning)
                        sc = max(1.0, np.max(np.abs(a[:3])))
                        k = min(2, len(a))
                        small = np.nonzero(b[:k] < 1e-3 * sc)[0]
                        if len(small):
                            k = int(small[0]) + 1
                        ca = np.max(np.abs(ga[ci, site - 1, spin, kind, :k] - a[:k])) / sc
                        cb = (np.max(np.abs(gb[ci, site - 1, spin, kind, :k - 1] - b[:k - 1])) / sc
                              if k > 1 else 0.0)
                        cerrs.append((max(ca, cb), ci, site, spin, kind))
                        # norm of the FP32-rounded excitation against FP64
                        assert abs(gn[ci, site - 1, spin, kind] - n2) < 1e-6 * n2
                err = np.max(np.abs(G[ci, site - 1, spin] - Gref)) / max(1e-12, np.max(np.abs(Gref)))
                errs.append((err, ci, site, spin))
    errs.sort(reverse=True)
    cerrs.sort(reverse=True)
    print("largest G_ii errors (rel. to max|G|, eta=1):", [(f"{e:.2e}", c, i, s) for e, c, i, s in errs[:5]],
          flush=True)
    print("largest leading-coefficient errors:", [(f"{e:.2e}",) + tuple(t) for e, *t in cerrs[:5]],
          flush=True)
    assert cerrs[0][0] < 1e-5, cerrs[0]
    assert errs[0][0] < 5e-3, errs[0]
