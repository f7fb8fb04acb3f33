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


def _lehmann(model, psi, e0, site, spin, omega, eta, cache):
    """Exact G_ii,s(omega) of a small cluster from the oracle's dense spectra of the
    N +- 1 sectors: particle poles at E_n - E0, hole poles at E0 - E_n, weights
    |<n|c^dag|0>|^2 and |<n|c|0>|^2; also the two weights' sums."""
    from oracle import oracle as O
    z = omega + 1j * eta
    G = np.zeros(len(omega), dtype=complex)
    w2 = [0.0, 0.0]
    for kind, create in ((0, True), (1, False)):
        d = 1 if create else -1
        nu = model.nup + (d if spin == 0 else 0)
        nd = model.ndn + (d if spin == 1 else 0)
        if not (0 <= nu <= model.n and 0 <= nd <= model.n):
            continue
        tgt, x = model.excite(psi, site, spin, create)
        key = (nu, nd)
        if key not in cache:
            cache[key] = np.linalg.eigh(tgt.dense())
        E, V = cache[key]
        w = (V.T @ x) ** 2
        w2[kind] = float(x @ x)
        if create:
            G += (w[None, :] / (z[:, None] - (E - e0)[None, :])).sum(1)
        else:
            G += (w[None, :] / (z[:, None] + (E - e0)[None, :])).sum(1)
    return G, w2


def test_batched_edge_sectors_and_site_dependent_u(ctx):
    """Clusters the random batch does not draw, against the oracle's dense spectra
    (E0 and the Lehmann sum of G_ii) and closed forms: one site (E0 = 2 eps + U), the
    Hubbard dimer (U/2 - sqrt(U^2/4 + 4 t^2)), no up electrons (free down fermions),
    a full up species (each down electron sees eps_i + U_i: per-site U read by the
    right site), site-dependent U on 5 and 7 sites (the per-site branch of the
    diagonal), and 16 sites, the batched path's maximum.  Chains of 120 steps exhaust
    the sectors up to 120 states; eta = 0.5."""
    from oracle import oracle as O
    from pyhub_amd.batch import Cluster, gii_omega, run_batch
    rng = np.random.default_rng(404)

    def rand_T(n, scale=0.8):
        T = rng.normal(size=(n, n)) * scale
        T = (T + T.T) / 2
        return T

    T1 = np.array([[-0.3]])
    T2 = np.array([[0.0, -1.0], [-1.0, 0.0]])
    T6, T7, T5, T7b = rand_T(6), rand_T(7), rand_T(5), rand_T(7)
    clusters = [
        Cluster(1, 1, 1, T1, np.array([5.0])),
        Cluster(2, 1, 1, T2, np.full(2, 4.0)),
        Cluster(6, 0, 3, T6, rng.uniform(2, 8, 6)),
        Cluster(7, 7, 2, T7, rng.uniform(1, 9, 7)),
        Cluster(5, 2, 3, T5, rng.uniform(1, 9, 5)),
        Cluster(7, 3, 2, T7b, rng.uniform(1, 9, 7)),
    ]
    closed = [2 * -0.3 + 5.0, O.dimer_e0(1.0, 4.0), O.free_fermion_e0(T6, 0, 3),
              float(np.trace(T7)) + O.free_fermion_e0(T7 + np.diag(clusters[3].U), 0, 2), None, None]
    # the largest cluster the batched path takes (16 sites), one electron of each spin
    clusters.append(Cluster(16, 1, 1, rand_T(16, 0.5), rng.uniform(1, 9, 16)))
    closed.append(None)
    e0, ga, gb, gn = run_batch(clusters, gs_steps=256, gf_steps=120, seed=5, gs_tol=1e-10)
    omega = np.linspace(-20, 20, 401)
    eta = 0.5
    G = gii_omega(e0, ga, gb, gn, clusters, omega, eta=eta)
    worst = []
    for ci, cl in enumerate(clusters):
        model = O.HubbardModel(cl.n, cl.n_up, cl.n_dn, cl.T, cl.U)
        E, V = np.linalg.eigh(model.dense())
        psi = V[:, 0]
        print(f"cluster {ci}: n={cl.n} ({cl.n_up},{cl.n_dn}) dim={model.dim} e0 {e0[ci]:.12f} "
              f"dense {E[0]:.12f} closed {closed[ci]}", flush=True)
        assert abs(e0[ci] - E[0]) <= 1e-9 * max(1.0, abs(E[0])), (ci, e0[ci], E[0])
        if closed[ci] is not None:
            assert abs(E[0] - closed[ci]) <= 1e-10 * max(1.0, abs(E[0])), (ci, E[0], closed[ci])
        cache = {}
        for site in range(1, cl.n + 1):
            for spin in (0, 1):
                Gref, w2 = _lehmann(model, psi, E[0], site, spin, omega, eta, cache)
                # weights: FP32-rounded excitations of the FP64 ground state
                for kind in (0, 1):
                    assert abs(gn[ci, site - 1, spin, kind] - w2[kind]) <= 1e-6 * max(w2[kind], 1e-3), \
                        (ci, site, spin, kind, gn[ci, site - 1, spin, kind], w2[kind])
                assert abs(gn[ci, site - 1, spin].sum() - 1.0) < 1e-6
                err = np.max(np.abs(G[ci, site - 1, spin] - Gref)) / np.max(np.abs(Gref))
                worst.append((err, ci, site, spin))
        # the slots of absent sites stay zero
        assert not np.any(G[ci, cl.n:])
    worst.sort(reverse=True)
    print("largest G_ii errors (rel. to max|G|, eta=0.5):",
          [(f"{e:.2e}", c, i, s) for e, c, i, s in worst[:5]], flush=True)
    assert worst[0][0] < 1e-4, worst[0]   # 3.8e-6 measured


def test_batched_argument_errors(ctx):
    """The batched entry point rejects, before any launch, what it cannot take, with the
    library's error codes: more than 16 sites (unsupported), particle numbers out
    of [0, n], a non-symmetric or non-finite t matrix, non-finite U, and step counts
    outside [2, 256] (argument errors).  A good batch on the same context still runs."""
    from pyhub_amd import _lib as L
    from pyhub_amd.batch import Cluster, run_batch
    from oracle import oracle as O
    T6 = O.ring_T(6)
    good = Cluster(6, 3, 3, T6, np.full(6, 4.0))
    with pytest.raises(L.PhubError) as ei:
        run_batch([good, Cluster(17, 1, 1, O.ring_T(17), np.full(17, 4.0))], ctx=ctx)
    assert ei.value.code == L.ERR_UNSUPPORTED
    for bad in (Cluster(6, 7, 3, T6, np.full(6, 4.0)), Cluster(6, 3, -1, T6, np.full(6, 4.0))):
        with pytest.raises(ValueError):
            run_batch([good, bad], ctx=ctx)
    Tn = T6.copy()
    Tn[0, 1] += 0.1
    with pytest.raises(ValueError, match="symmetric"):
        run_batch([Cluster(6, 3, 3, Tn, np.full(6, 4.0))], ctx=ctx)
    Tf = T6.copy()
    Tf[2, 2] = np.nan
    with pytest.raises(ValueError, match="non-finite"):
        run_batch([Cluster(6, 3, 3, Tf, np.full(6, 4.0))], ctx=ctx)
    with pytest.raises(ValueError, match="non-finite"):
        run_batch([Cluster(6, 3, 3, T6, np.array([4.0, 4.0, np.inf, 4.0, 4.0, 4.0]))], ctx=ctx)
    for gs, gf in ((1, 120), (256, 0), (300, 120), (256, 257)):
        with pytest.raises(ValueError, match="steps"):
            run_batch([good], gs_steps=gs, gf_steps=gf, ctx=ctx)
    e0, ga, gb, gn = run_batch([good], gs_steps=200, gf_steps=40, ctx=ctx)
    m = O.HubbardModel(6, 3, 3, T6, 4.0)
    assert abs(e0[0] - np.linalg.eigvalsh(m.dense())[0]) < 1e-9
